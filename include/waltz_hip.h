/* Waltz HIP C ABI for MI355X (gfx950).
 *
 * Replaces the reference's HIP/DCU session API (platform/hip/waltz_hip.h:7-18),
 * which its Python GPU loader binds in serial mode
 * (python/pywaltz/gpu_dataloader.py:41-47), and the CUDA chunk decoder
 * (include/waltz_cuda.h:11-13). All payload pointers are device pointers; the
 * calls are synchronous on the session's stream (the default stream) or the
 * given stream. Functions returning int give 0 on success and -1 on failure,
 * with the message in waltz_hip_error(). One caller per session.
 */
#ifndef WALTZ_HIP_H
#define WALTZ_HIP_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct waltz_hip_session waltz_hip_session;
/* x fastest; width 4 (float) or 8 (double); bound = absolute error bound. */
typedef struct { uint32_t x, y, z; int width; double bound; } waltz_hip_config;

/* platform/hip/waltz_hip.h:9-12 */
waltz_hip_session* waltz_hip_create(int device);
void waltz_hip_destroy(waltz_hip_session*);
const char* waltz_hip_error(const waltz_hip_session*);
size_t waltz_hip_max_bytes(const waltz_hip_config*);
/* platform/hip/waltz_hip.h:17-18: canonical archive bytes, byte-identical to
 * the reference CPU codec; decode checks the archive against `bytes`. */
int waltz_hip_encode(waltz_hip_session*, const waltz_hip_config*, const void* input, void* archive, size_t capacity,
                     size_t* written);
int waltz_hip_decode(waltz_hip_session*, const waltz_hip_config*, const void* archive, size_t bytes, void* output,
                     size_t output_bytes);

/* include/waltz_cuda.h:11-13: decode one canonical chunk on `stream`
 * (a hipStream_t). Returns NULL on success or a thread-local diagnostic. */
const char* waltz_cuda_decode_chunk(uint32_t nx, uint32_t ny, uint32_t nz, int width, const void* archive,
                                    size_t archive_bytes, void* output, size_t output_bytes, void* stream);

/* No counterpart in the reference. For callers that cannot allocate device
 * memory (code under stream capture, serving loops, harnesses that forbid
 * hipMalloc): waltz_cuda_decode_chunk with the same arguments, rules and
 * results, whose device scratch is one fixed array inside the library (about
 * 26.4 MiB per device, held until the process ends) instead of a workspace
 * grown with hipMalloc. Chunks of at most WALTZ_NOALLOC_MAX_SAMPLES samples;
 * larger ones are rejected with a diagnostic. Calls share the array: they are
 * serialised within a process, and each process has its own copy. */
#define WALTZ_NOALLOC_MAX_SAMPLES 1048576u
const char* waltz_cuda_decode_chunk_noalloc(uint32_t nx, uint32_t ny, uint32_t nz, int width, const void* archive,
                                            size_t archive_bytes, void* output, size_t output_bytes, void* stream);

#ifdef __cplusplus
}
#endif
#endif
