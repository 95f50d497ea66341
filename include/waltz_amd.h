/* Stage-level C ABI of the gfx950 Waltz library (no reference counterpart:
 * the reference links its stages inside one C++ library, src/waltz.cu; these
 * entry points expose the same stages individually so that callers and parity
 * tests can drive one stage on device buffers). Device pointers, hipStream_t
 * passed as void*, synchronous; 0 on success, -1 on failure with the message
 * in waltz_amd_last_error() (thread-local).
 */
#ifndef WALTZ_AMD_H
#define WALTZ_AMD_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

const char* waltz_amd_last_error(void);
/* Library build identity: "gfx950 ..." */
const char* waltz_amd_version(void);

/* Whole codec on `stream` (cpu/include/waltz_cpu.h:24-29 semantics with device
 * buffers). width = 4 or 8. */
size_t waltz_amd_max_compressed_bytes(uint32_t nx, uint32_t ny, uint32_t nz, int width);
int waltz_amd_compress(int width, const void* input, uint32_t nx, uint32_t ny, uint32_t nz, double abs_bound,
                       void* archive, size_t capacity, size_t* written, void* stream);
int waltz_amd_decompress(int width, const void* archive, size_t bytes, uint32_t nx, uint32_t ny, uint32_t nz,
                         void* output, void* stream);

/* CDF 9/7 lifting of an nx*ny*nz volume (cpu/src/transform.cpp:409-456):
 * forward reads `input`, writes coefficients to `output`; inverse reads
 * coefficients from `input`, writes samples to `output`. input may equal
 * output. Needs two scratch volumes tmp0/tmp1 (may be NULL: allocated). */
int waltz_amd_transform(int width, int inverse, const void* input, void* output, uint32_t nx, uint32_t ny,
                        uint32_t nz, unsigned levels_xy, unsigned levels_z, int dyadic, void* stream);

/* Transpose of waltz_amd_transform with the same geometry (no reference
 * counterpart; the gradient of the transform for autograd): of_inverse = 0
 * maps a coefficient-space gradient to sample space (DWT^T), of_inverse = 1
 * maps a sample-space gradient to coefficient space (IDWT^T). input may
 * equal output. */
int waltz_amd_transform_adjoint(int width, int of_inverse, const void* input, void* output, uint32_t nx, uint32_t ny,
                                uint32_t nz, unsigned levels_xy, unsigned levels_z, int dyadic, void* stream);

/* Asynchronous forms of the two calls above (no reference counterpart), for
 * autograd and stream capture: the caller passes the two scratch volumes
 * tmp0/tmp1 (nx*ny*nz elements each, distinct from input and output) and the
 * call only enqueues work on `stream`: no synchronisation, allocation or lock. */
int waltz_amd_transform_async(int width, int inverse, const void* input, void* output, void* tmp0, void* tmp1,
                              uint32_t nx, uint32_t ny, uint32_t nz, unsigned levels_xy, unsigned levels_z, int dyadic,
                              void* stream);
int waltz_amd_transform_adjoint_async(int width, int of_inverse, const void* input, void* output, void* tmp0,
                                      void* tmp1, uint32_t nx, uint32_t ny, uint32_t nz, unsigned levels_xy,
                                      unsigned levels_z, int dyadic, void* stream);

/* Bit-neutral launch tuning (no reference counterpart): autotune times the
 * fused level kernels of both directions at each candidate z segment for this
 * geometry and type on this device, then the per-axis passes' line tiles and
 * LDS pair limit, and keeps the fastest (process-wide, per geometry and
 * direction); set_tuning fixes the number of
 * fused levels and their five z segments (positive multiples of 16) for one
 * direction; clear_tuning restores the built-in heuristic; tuning_report
 * writes the table as JSON (returns its full length). Results never depend on
 * these choices. */
int waltz_amd_autotune(int width, uint32_t nx, uint32_t ny, uint32_t nz, void* stream);
int waltz_amd_set_tuning(int width, uint32_t nx, uint32_t ny, uint32_t nz, int inverse, unsigned fused_levels,
                         const uint32_t* z_segments);
/* The per-axis passes' share of the tuning (the plane schedule, and dyadic
 * levels that are not fused): line tiles of the X, Y and Z passes (8, 16 or
 * 32; an inverse direction uses only its X tile, its Y and Z tiles being part
 * of the format) and the largest plane, in samples, whose X and Y passes run
 * as one LDS-staged launch (0, 1024 or 4096). Results never depend on them. */
int waltz_amd_set_pass_tuning(int width, uint32_t nx, uint32_t ny, uint32_t nz, int inverse, const uint32_t* tiles,
                              uint32_t lds_pair_max);
void waltz_amd_clear_tuning(void);
size_t waltz_amd_tuning_report(char* buf, size_t cap);

/* Quantization + Morton packing (cpu/src/codec.cpp:198-234): coeff is
 * replaced by dequantized values; mag[n], sign[ceil(n/32)] packed order;
 * q-outliers sorted by index into qidx/qval (capacity n), count in *nq. */
int waltz_amd_quantize(int width, void* coeff, uint32_t nx, uint32_t ny, uint32_t nz, uint32_t bz, double quantum,
                       void* mag, void* sign, void* qidx, void* qval, uint32_t* nq, void* stream);

/* WZP stream (cpu/src/wzp_cpu.cpp:780-949): encode mag/sign of n elements
 * into out (capacity bytes), *written = stream bytes; decode back. */
int waltz_amd_wzp_encode(const void* mag, const void* sign, uint32_t n, void* out, size_t capacity, size_t* written,
                         void* stream);
int waltz_amd_wzp_decode(const void* in, size_t bytes, uint32_t n, void* mag, void* sign, void* stream);

/* Batched decode of many archives whose tile shapes may differ (no reference
 * counterpart; the reference's waltz_gpu_decode_batch takes one tile shape).
 * create: dims = jobs x {nx, ny, nz} (x fastest), width 4 or 8; NULL on error
 * (message in waltz_amd_batch_error()). Tiles are written to the output
 * concatenated in job order (output_elements() samples in total). decode:
 * `packed` holds the archives (device, packed_bytes long; every span must lie
 * inside it or the job is reported corrupt), `spans` is a device array of jobs
 * little-endian uint64 (offset, size) pairs (offsets 8-byte aligned; size 0
 * = tile of zeros); it only enqueues work on `stream` (no synchronisation,
 * allocation or memset), so it can be captured in a hipGraph and replayed.
 * check: synchronises `stream` and reports the device status of the last
 * decode (0, or -1 for corrupt archives / nonfinite output). */
typedef struct waltz_amd_batch waltz_amd_batch;
waltz_amd_batch* waltz_amd_batch_create(int width, uint32_t jobs, const uint32_t* dims);
const char* waltz_amd_batch_error(void);
void waltz_amd_batch_destroy(waltz_amd_batch* batch);
size_t waltz_amd_batch_workspace_bytes(const waltz_amd_batch* batch);
size_t waltz_amd_batch_output_elements(const waltz_amd_batch* batch);
int waltz_amd_batch_decode(waltz_amd_batch* batch, const void* packed, size_t packed_bytes, const void* spans,
                           void* output, void* workspace, size_t workspace_bytes, void* stream);
int waltz_amd_batch_check(waltz_amd_batch* batch, void* stream);

/* Batched compress with the same plan (no reference counterpart: the
 * reference compresses one field per call, src/waltz.cu:327-609). `input`
 * holds every job's field concatenated in job order (output_elements()
 * samples, the layout batch_decode writes), compressed with one absolute
 * bound; archive j is written at a 16-byte aligned slot of `archives`
 * (archive_capacity() bytes) and `spans` (device, jobs x little-endian uint64
 * {offset, size}) receives each archive's place and length (size 0 for a job
 * that failed), i.e. exactly what batch_decode takes. Every archive is the
 * canonical one the single-field compress writes. Work: one launch per stage
 * for all jobs; the call only enqueues work on `stream` (no synchronisation,
 * allocation or memset), so it can be captured and replayed. compress_check
 * synchronises `stream` and reports the first failed job ("job j: ...").
 * The plan must not run a decode and a compress at the same time. */
size_t waltz_amd_batch_compress_workspace_bytes(const waltz_amd_batch* batch);
size_t waltz_amd_batch_archive_capacity(const waltz_amd_batch* batch);
int waltz_amd_batch_compress(waltz_amd_batch* batch, const void* input, double abs_bound, void* archives,
                             size_t archives_bytes, void* spans, void* workspace, size_t workspace_bytes,
                             void* stream);
int waltz_amd_batch_compress_check(waltz_amd_batch* batch, void* stream);

/* Per-stage timing with HIP events on the pipeline's stream (off by default).
 * enable returns the previous state; report writes a JSON object
 * {"stage": [total_ms, launches], ...} accumulated since the last report and
 * resets it; it returns the full length (write is truncated to cap-1). */
int waltz_amd_profile_enable(int on);
size_t waltz_amd_profile_report(char* buf, size_t cap);

#ifdef __cplusplus
}
#endif
#endif
