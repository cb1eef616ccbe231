/*
 * bonxai_gpu_probes.h -- measurement instruments beside the map library.
 *
 * Not part of the map API (include/bonxai_gpu.h) and not in its library:
 * these entry points live in libbonxai_probes.so, which links against
 * libbonxai_gpu.so only to run on a context's stream (bxg_ctx_stream).  They
 * have no counterpart in the reference; they exist to calibrate and bound the
 * map kernels' measurements:
 *   - bxg_copy_probe: a kernel of known traffic, the scale for HBM byte
 *     counters of a profile;
 *   - bxg_gather_probe: the memory system's rate for scattered cell reads,
 *     the floor of a lookup batch in caller order;
 *   - bxg_fp64_probe: the FP64 operations of the scan front end, built with
 *     the map library's flags, for the correctly-rounded checks.
 * Each call enqueues its kernel on the context's stream and returns; the
 * context's calls that wait for its stream (bxg_ctx_sync, bxg_memcpy_d2h)
 * wait for it, and bxg_event_record brackets it for timing.
 */
#ifndef BONXAI_GPU_PROBES_H
#define BONXAI_GPU_PROBES_H

#include <stdint.h>

#include "bonxai_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

/* a plain device-to-device copy kernel (16 bytes a lane), bytes % 16 == 0:
 * the known-size reference that scales the HBM byte counters of a profile */
int bxg_copy_probe(bxg_ctx* ctx, const void* src_dev, void* dst_dev, uint64_t bytes);

/* n scattered 4-byte reads from a device table of table_bytes (a power of two,
 * at least 4), written in order to out_dev (n words): the rate the memory
 * system gives to randomly placed cell reads, the part of a lookup batch in
 * caller order that no kernel can coalesce */
int bxg_gather_probe(bxg_ctx* ctx, const void* table_dev, uint64_t table_bytes, uint64_t n,
                     void* out_dev);

/* element-wise FP64 operations of the scan front end on device arrays:
 * out[5i..5i+4] = sqrt(a), a/b, c + (a/b)*c, floor(a*b), (a*a + b*b) + c*c */
int bxg_fp64_probe(bxg_ctx* ctx, const double* a_dev, const double* b_dev, const double* c_dev,
                   uint64_t n, double* out_dev);

#ifdef __cplusplus
}
#endif

#endif /* BONXAI_GPU_PROBES_H */
