/* Batched device decode of HDF5-style Waltz chunks into one training tensor.
 * Same struct, entry point and wire format as the reference's
 * include/waltz_gpu_batch.h:7-17 (implemented in src/batch_decode.cuh:74-186),
 * bound by python/pywaltz/gpu_dataloader.py:34-36.
 *
 * `host` and `packed` hold identical bytes (host copy + device copy): 16-byte
 * aligned canonical archives followed by `jobs` little-endian uint64
 * (offset, size) pairs at the end, in sample / side / tile-ZYX order (x
 * fastest); size 0 marks a hole, filled with `fill`. One job is one tile of
 * x*y*z samples at origin (tx*x, ty*y, tz*z) of an fx*fy*fz frame; tiles that
 * cross the frame edge are clipped. Output: contiguous (side, sample, z, y, x)
 * of `width`-byte floats, output_bytes = fx*fy*fz*samples*sides*width.
 * Scratch: device memory of at least waltz_gpu_batch_scratch_bytes() (the
 * reference's 3*stride*jobs*width + 8*256*bz*4 + 4 also suffices).
 * Returns NULL on success or a thread-local diagnostic; synchronous on
 * `stream` (a hipStream_t).
 */
#ifndef WALTZ_GPU_BATCH_H
#define WALTZ_GPU_BATCH_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif
typedef struct {
    uint32_t x, y, z, fx, fy, fz, samples, sides, width;
} waltz_batch_shape;

const char* waltz_gpu_decode_batch(const waltz_batch_shape* shape, const void* host, const void* packed, size_t bytes,
                                   void* scratch, size_t scratch_bytes, void* output, size_t output_bytes,
                                   double fill, void* stream);
/* Scratch bytes waltz_gpu_decode_batch needs for this shape (no reference
 * counterpart; the reference documents the formula in its header comment). */
size_t waltz_gpu_batch_scratch_bytes(const waltz_batch_shape* shape);
#ifdef __cplusplus
}
#endif
#endif
