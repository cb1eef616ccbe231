// gpu_runtime.hpp -- the device context behind the GPU-backed Bonxai drop-in:
// one bxg_ctx per GPU (a stream, scratch and a memory budget), shared by every
// grid created on that GPU, and the translation of the C ABI's status codes
// into the exceptions the reference's C++ API raises.
//
// Environment (read once, at first use of the default context):
//   BONXAI_GPU_DEVICE     device ordinal of the default context (default 0)
//   BONXAI_GPU_BUDGET_MB  device-memory budget of the default context (0 = none)
#pragma once

#include <cstdint>
#include <cstdlib>
#include <memory>
#include <new>
#include <stdexcept>
#include <string>

#include "bonxai_gpu.h"

namespace Bonxai {
namespace gpu {

/// A failed call into the device library; status() is the BXG_* code.
class Error : public std::runtime_error {
 public:
  Error(int status, const std::string& what) : std::runtime_error(what), status_(status) {}
  int status() const { return status_; }

 private:
  int status_;
};

/// Device memory or node-pool exhaustion (BXG_ERR_OUT_OF_POOL, BXG_ERR_BUDGET).
class OutOfMemory : public Error {
 public:
  using Error::Error;
};

inline void check(int status, bxg_ctx* ctx = nullptr) {
  if (status == BXG_OK) return;
  std::string msg = bxg_status_string(status);
  if (ctx) {
    const char* detail = bxg_ctx_last_error(ctx);
    if (detail && *detail) msg += std::string(": ") + detail;
  }
  if (status == BXG_ERR_OUT_OF_POOL || status == BXG_ERR_BUDGET) throw OutOfMemory(status, msg);
  throw Error(status, msg);
}

/// Owner of one bxg_ctx.
class Context {
 public:
  explicit Context(int device = 0, uint64_t budget_bytes = 0) {
    bxg_ctx* c = nullptr;
    check(bxg_ctx_create(device, budget_bytes, &c));
    ctx_ = c;
  }
  ~Context() {
    if (ctx_) bxg_ctx_destroy(ctx_);
  }
  Context(const Context&) = delete;
  Context& operator=(const Context&) = delete;

  bxg_ctx* get() const { return ctx_; }
  void sync() const { check(bxg_ctx_sync(ctx_), ctx_); }

  /// The process-wide context, created on first use.  It is never destroyed:
  /// grids with static storage may outlive any other object, and the device
  /// memory goes back to the driver when the process ends.
  static Context& Default() {
    static Context* inst = [] {
      const char* d = std::getenv("BONXAI_GPU_DEVICE");
      const char* b = std::getenv("BONXAI_GPU_BUDGET_MB");
      const int dev = d ? std::atoi(d) : 0;
      const uint64_t mb = b ? std::strtoull(b, nullptr, 10) : 0;
      return new Context(dev, mb << 20);
    }();
    return *inst;
  }

 private:
  bxg_ctx* ctx_ = nullptr;
};

}  // namespace gpu
}  // namespace Bonxai
