// mask.hpp -- the bit mask of a host-side leaf view (GPU-backed Bonxai drop-in).
//
// Public surface of the reference's Bonxai::Mask (bonxai_core/include/bonxai/mask.hpp:19-140):
// bit n lives in word n >> 6 at bit n & 63 (mask.hpp:329-365), the word count is
// max(bits / 64, 1) (mask.hpp:256-257), and an on-iterator walks the set bits upward.
#pragma once

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <vector>

namespace Bonxai {

class Mask {
 public:
  /// 2^(3 log2dim) bits, all off (or all on).
  explicit Mask(size_t log2dim, bool on = false)
      : size_(1u << (3 * log2dim)),
        words_(std::max<uint32_t>(size_ >> 6, 1u), on ? ~uint64_t(0) : uint64_t(0)) {}

  [[nodiscard]] size_t memUsage() const { return sizeof(Mask) + words_.size() * sizeof(uint64_t); }
  [[nodiscard]] uint32_t bitCount() const { return size_; }
  [[nodiscard]] uint32_t wordCount() const { return static_cast<uint32_t>(words_.size()); }
  [[nodiscard]] uint64_t getWord(size_t n) const { return words_[n]; }
  void setWord(size_t n, uint64_t v) { words_[n] = v; }
  [[nodiscard]] uint64_t* data() { return words_.data(); }
  [[nodiscard]] const uint64_t* data() const { return words_.data(); }

  [[nodiscard]] uint32_t countOn() const {
    uint32_t s = 0;
    for (uint64_t w : words_) s += static_cast<uint32_t>(__builtin_popcountll(w));
    return s;
  }
  [[nodiscard]] bool isOn(uint32_t n) const { return (words_[n >> 6] >> (n & 63)) & 1u; }
  [[nodiscard]] bool isOff(uint32_t n) const { return !isOn(n); }
  /// all bits on
  [[nodiscard]] bool isOn() const {
    for (uint64_t w : words_)
      if (w != ~uint64_t(0)) return false;
    return true;
  }
  /// all bits off
  [[nodiscard]] bool isOff() const {
    for (uint64_t w : words_)
      if (w) return false;
    return true;
  }
  /// set bit n; returns its previous state
  bool setOn(uint32_t n) {
    uint64_t& w = words_[n >> 6];
    const uint64_t b = uint64_t(1) << (n & 63);
    const bool was = (w & b) != 0;
    w |= b;
    return was;
  }
  /// clear bit n; returns its previous state
  bool setOff(uint32_t n) {
    uint64_t& w = words_[n >> 6];
    const uint64_t b = uint64_t(1) << (n & 63);
    const bool was = (w & b) != 0;
    w &= ~b;
    return was;
  }
  void set(uint32_t n, bool on) { on ? (void)setOn(n) : (void)setOff(n); }
  void setOn() { std::fill(words_.begin(), words_.end(), ~uint64_t(0)); }
  void setOff() { std::fill(words_.begin(), words_.end(), uint64_t(0)); }
  void set(bool on) { on ? setOn() : setOff(); }

  bool operator==(const Mask& o) const { return size_ == o.size_ && words_ == o.words_; }
  bool operator!=(const Mask& o) const { return !(*this == o); }

  /// first set bit at or after `start`, or bitCount() if none
  [[nodiscard]] uint32_t findNextOn(uint32_t start) const {
    uint32_t n = start >> 6;
    if (n >= words_.size()) return size_;
    uint64_t b = words_[n] & (~uint64_t(0) << (start & 63));
    while (!b) {
      if (++n >= words_.size()) return size_;
      b = words_[n];
    }
    const uint32_t pos = (n << 6) + static_cast<uint32_t>(__builtin_ctzll(b));
    return pos < size_ ? pos : size_;
  }
  [[nodiscard]] uint32_t findFirstOn() const { return findNextOn(0); }

  /// iterator over the set bits, upward (contextually true while valid)
  class Iterator {
   public:
    Iterator(const Mask* m, uint32_t pos) : mask_(m), pos_(pos) {}
    uint32_t operator*() const { return pos_; }
    uint32_t pos() const { return pos_; }
    explicit operator bool() const { return pos_ < mask_->size_; }
    bool isValid() const { return pos_ < mask_->size_; }
    Iterator& operator++() {
      pos_ = mask_->findNextOn(pos_ + 1);
      return *this;
    }
    Iterator operator++(int) {
      Iterator t = *this;
      ++*this;
      return t;
    }

   private:
    const Mask* mask_;
    uint32_t pos_;
  };
  [[nodiscard]] Iterator beginOn() const { return Iterator(this, findFirstOn()); }

 private:
  uint32_t size_;
  std::vector<uint64_t> words_;
};

}  // namespace Bonxai
