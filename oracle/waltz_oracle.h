/*
 * ORACLE — test infrastructure only.
 *
 * A plain-C restatement of the reference Waltz CPU scalar codec (the normative
 * byte/bit profile of the archive format). Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it, and there only to check against;
 * the product library never links, loads or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the reference repository root). Pinning (what checks each function against
 * something other than itself) is listed in tests/test_oracle_*.py:
 *   - whole-codec archives and decoded bits: golden digests written by the
 *     reference CPU codec (tests/golden/), and oracle/_ref when it is built;
 *   - transform: the reference's FNV-1a known answers for four shapes and both
 *     types, oracle/_ref bitwise, and perfect-reconstruction/linearity checks;
 *   - WZP: oracle/_ref bytes, size identities, brute-force BIT_2 permutation;
 *   - Morton rank: brute-force octree recursion on tiny bricks;
 *   - quantization and repairs: closed forms (llrint, TwoDiff) on crafted values.
 *
 * All functions return 0 on success and -1 on failure (wo_last_error() holds a
 * thread-local message), except where documented.
 */
#ifndef WALTZ_ORACLE_H
#define WALTZ_ORACLE_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

const char* wo_last_error(void);

/* Geometry (include/utils/Blob.hpp:22-29, cpu/src/codec.cpp:89-107). */
unsigned wo_dwt_levels(uint32_t n);
uint32_t wo_zband(uint32_t nx, uint32_t ny, uint32_t nz, unsigned lz, int dyadic);
uint32_t wo_morton_rank(uint32_t x, uint32_t y, uint32_t z, uint32_t nx, uint32_t ny, uint32_t nz);

/* In-place CDF 9/7 lifting over an nx*ny*nz row-major volume (x fastest),
 * cpu/src/transform.cpp:409-456. Sets round-to-nearest, FTZ/DAZ off. */
int wo_transform_f32(float* data, uint32_t nx, uint32_t ny, uint32_t nz, unsigned lxy, unsigned lz,
                     int dyadic, int inverse);
int wo_transform_f64(double* data, uint32_t nx, uint32_t ny, uint32_t nz, unsigned lxy, unsigned lz,
                     int dyadic, int inverse);

/* Quantization + Morton brick packing, cpu/src/codec.cpp:198-234.
 * coeff is replaced by the dequantized coefficients; mag[n] (packed order),
 * sign[ceil(n/32)] (packed order, cleared here), q-outliers sorted by index. */
int wo_quantize_f32(float* coeff, uint32_t nx, uint32_t ny, uint32_t nz, uint32_t bz, double quantum,
                    uint16_t* mag, uint32_t* sign, uint32_t* qidx, uint32_t* qval, uint32_t* nq);
int wo_quantize_f64(double* coeff, uint32_t nx, uint32_t ny, uint32_t nz, uint32_t bz, double quantum,
                    uint16_t* mag, uint32_t* sign, uint32_t* qidx, uint32_t* qval, uint32_t* nq);

/* Packed magnitudes/signs -> dequantized row-major coefficients,
 * cpu/src/codec.cpp:139-150 (no q-outlier patching). */
void wo_dequantize_f32(const uint16_t* mag, const uint32_t* sign, uint32_t nx, uint32_t ny, uint32_t nz,
                       uint32_t bz, double quantum, float* out);
void wo_dequantize_f64(const uint16_t* mag, const uint32_t* sign, uint32_t nx, uint32_t ny, uint32_t nz,
                       uint32_t bz, double quantum, double* out);

/* Native residual repairs, cpu/src/repair_avx2.cpp:48-74 (ascending indices). */
int wo_repairs_f32(const float* original, const float* reconstructed, uint32_t n, double bound,
                   uint32_t* idx, float* err, uint32_t* count);
int wo_repairs_f64(const double* original, const double* reconstructed, uint32_t n, double bound,
                   uint32_t* idx, double* err, uint32_t* count);
/* Exact |value-original| <= bound via TwoDiff, cpu/src/repair_avx2.cpp:48-56. */
int wo_within_bound(double value, double original, double bound);

/* WZP stream, cpu/src/wzp_cpu.cpp:780-949. encode returns the byte count or -1. */
size_t wo_wzp_max_bytes(uint32_t n);
int64_t wo_wzp_encode(const uint16_t* mag, const uint32_t* sign, uint32_t n, uint8_t* out, size_t capacity);
int wo_wzp_decode(const uint8_t* in, size_t bytes, uint32_t n, uint16_t* mag, uint32_t* sign);

/* Whole codec, cpu/src/codec.cpp:177-336,344-355. */
size_t wo_max_compressed_bytes(uint32_t nx, uint32_t ny, uint32_t nz, int width);
int wo_compress_f32(const float* in, uint32_t nx, uint32_t ny, uint32_t nz, double abs_bound,
                    uint8_t* out, size_t capacity, size_t* written);
int wo_compress_f64(const double* in, uint32_t nx, uint32_t ny, uint32_t nz, double abs_bound,
                    uint8_t* out, size_t capacity, size_t* written);
int wo_decompress_f32(const uint8_t* in, size_t bytes, uint32_t nx, uint32_t ny, uint32_t nz, float* out);
int wo_decompress_f64(const uint8_t* in, size_t bytes, uint32_t nx, uint32_t ny, uint32_t nz, double* out);

/* FNV-1a-64 over bytes (the digest the reference's transform bench prints,
 * cpu/tests/test_transform.cpp:246-247). */
uint64_t wo_fnv1a64(const void* data, size_t bytes);

#ifdef __cplusplus
}
#endif
#endif
