// Quantization with Morton brick packing, and its inverse (dequantization of
// packed magnitudes into row-major coefficients).
//
// Bricks are 16x16xBZ boxes of the full coefficient volume, enumerated z-slab,
// then y, then x; brick (x0,y0,z0) with extents (bx,by,bzz) starts at packed
// offset z0*nx*ny + y0*nx*bzz + x0*by*bzz (the running sum of
// cpu/src/codec.cpp:117-125 in closed form). Inside a brick, the packed order is
// the recursive ceil-half octant rank (cpu/src/codec.cpp:96-107), tabulated
// once per geometry for the eight edge shapes.
//
// One workgroup owns one brick: it reads the brick's coefficients (rows of
// <=16 contiguous values), writes the dequantized value back in place, scatters
// u16 magnitudes inside the brick's own packed span, and assembles the brick's
// sign bits in LDS before writing whole 32-bit words (only the two words that a
// brick shares with its neighbours use atomics).
#include "common.hpp"
#include "launch.hpp"

namespace wz {

BrickGeometry brick_geometry(Dims d, uint32_t bz) {
    BrickGeometry g;
    g.d = d;
    g.bz = bz;
    g.nbx = (d.nx + 15) / 16;
    g.nby = (d.ny + 15) / 16;
    g.nbz = (d.nz + bz - 1) / bz;
    return g;
}

__global__ void __launch_bounds__(256) rank_tables(Dims d, uint32_t bz, uint16_t* rank, uint16_t* inv) {
    const unsigned s = blockIdx.x;
    const uint32_t nx = (s & 1) ? d.nx % 16 : 16, ny = (s & 2) ? d.ny % 16 : 16, nz = (s & 4) ? d.nz % bz : bz;
    const uint32_t vol = nx * ny * nz;
    for (uint32_t i = threadIdx.x; i < vol; i += 256) {
        const uint32_t x = i % nx, y = (i / nx) % ny, z = i / (nx * ny);
        const uint32_t r = morton_rank(x, y, z, nx, ny, nz);
        rank[s * 8192 + i] = uint16_t(r);
        inv[s * 8192 + r] = uint16_t(i);
    }
}

void build_rank_tables(Dims d, uint32_t bz, uint16_t* rank, uint16_t* inv, hipStream_t s) {
    hipLaunchKernelGGL(rank_tables, dim3(8), dim3(256), 0, s, d, bz, rank, inv);
}

struct Brick {
    uint32_t x0, y0, z0, bx, by, bz, shape;
    uint64_t base;
};

WZ_D Brick brick_of(const BrickGeometry& g, uint32_t b) {
    Brick k;
    const uint32_t bxi = b % g.nbx, r = b / g.nbx, byi = r % g.nby, bzi = r / g.nby;
    k.x0 = bxi * 16;
    k.y0 = byi * 16;
    k.z0 = bzi * g.bz;
    k.bx = min(16u, g.d.nx - k.x0);
    k.by = min(16u, g.d.ny - k.y0);
    k.bz = min(g.bz, g.d.nz - k.z0);
    k.shape = unsigned(k.bx != 16) | (unsigned(k.by != 16) << 1) | (unsigned(k.bz != g.bz) << 2);
    k.base = uint64_t(k.z0) * g.d.nx * g.d.ny + uint64_t(k.y0) * g.d.nx * k.bz + uint64_t(k.x0) * k.by * k.bz;
    return k;
}

template <class T>
__global__ void __launch_bounds__(256) quantize_kernel(T* __restrict__ coeff, BrickGeometry g, T q, T inv,
                                                       const uint16_t* __restrict__ rank, uint16_t* __restrict__ mag,
                                                       uint32_t* __restrict__ sign, SideList qo,
                                                       uint32_t* __restrict__ status) {
    __shared__ uint32_t bits[257];
    const Brick k = brick_of(g, blockIdx.x);
    const uint32_t vol = k.bx * k.by * k.bz;
    const uint64_t plane = uint64_t(g.d.nx) * g.d.ny;
    for (uint32_t w = threadIdx.x; w < 257; w += 256) bits[w] = 0;
    __syncthreads();
    const uint32_t shift = uint32_t(k.base & 31u);
    bool bad = false;
    for (uint32_t i = threadIdx.x; i < vol; i += 256) {
        const uint32_t x = i % k.bx, y = (i / k.bx) % k.by, z = i / (k.bx * k.by);
        const uint64_t at = (uint64_t(k.z0 + z) * g.d.ny + (k.y0 + y)) * g.d.nx + k.x0 + x;
        const uint32_t r = rank[k.shape * 8192 + i];
        const T scaled = coeff[at] * inv;
        // cpu/src/quantize_avx2.cpp:45-62: finite, |scaled| < 2^63, llrint, INT32 saturation.
        if (!__builtin_isfinite(scaled) || !(abs_t(scaled) < T(9223372036854775808.0))) bad = true;
        const T rounded = rint_t(scaled);
        const bool negative = rounded < T(0);
        const T am = abs_t(rounded);
        const int32_t m = am >= T(2147483648.0) ? INT32_MAX : int32_t(am);
        coeff[at] = m ? T(negative ? -m : m) * q : T(0);
        mag[k.base + r] = uint16_t(m > 65535 ? 65535 : m);
        if (negative && m) atomicOr(&bits[(r + shift) >> 5], 1u << ((r + shift) & 31u));
        if (m > 65535) {
            atomicOr(&qo.bitmap[at >> 5], 1u << (at & 31u));
            const uint32_t slot = atomicAdd(qo.count, 1u);
            if (slot < qo.capacity) {
                qo.index[slot] = uint32_t(at);
                static_cast<uint32_t*>(qo.payload)[slot] = uint32_t(m);
            }
        }
    }
    if (bad) atomicOr(status, uint32_t(kQuantDomain));
    __syncthreads();
    // Brick bits occupy global bits [base, base+vol); LDS word w maps to global
    // word base/32 + w. The first and last words may be shared with neighbours.
    const uint32_t words = (shift + vol + 31) >> 5;
    const uint64_t w0 = k.base >> 5;
    for (uint32_t w = threadIdx.x; w < words; w += 256) {
        const uint32_t v = bits[w];
        if (w == 0 || w == words - 1) {
            if (v) atomicOr(&sign[w0 + w], v);
        } else {
            sign[w0 + w] = v;
        }
    }
}

template <class T>
void quantize_bricks(T* coeff, const BrickGeometry& g, double quantum, const uint16_t* rank, uint16_t* mag,
                     uint32_t* sign, SideList q, uint32_t* status, hipStream_t s) {
    const uint64_t bricks = uint64_t(g.nbx) * g.nby * g.nbz;
    check_launch_size(bricks * 256);
    hipLaunchKernelGGL((quantize_kernel<T>), dim3(uint32_t(bricks)), dim3(256), 0, s, coeff, g, T(quantum),
                       T(1.0 / quantum), rank, mag, sign, q, status);
}

template <class T>
__global__ void __launch_bounds__(256) dequantize_kernel(const uint16_t* __restrict__ mag,
                                                         const uint32_t* __restrict__ sign, BrickGeometry g, T q,
                                                         const uint16_t* __restrict__ rank, T* __restrict__ out) {
    const Brick k = brick_of(g, blockIdx.x);
    const uint32_t vol = k.bx * k.by * k.bz;
    for (uint32_t i = threadIdx.x; i < vol; i += 256) {
        const uint32_t x = i % k.bx, y = (i / k.bx) % k.by, z = i / (k.bx * k.by);
        const uint64_t at = (uint64_t(k.z0 + z) * g.d.ny + (k.y0 + y)) * g.d.nx + k.x0 + x;
        const uint64_t p = k.base + rank[k.shape * 8192 + i];
        const T v = T(mag[p]) * q;  // cpu/src/codec.cpp:146-147
        out[at] = ((sign[p >> 5] >> (p & 31u)) & 1u) ? -v : v;
    }
}

template <class T>
void dequantize_bricks(const uint16_t* mag, const uint32_t* sign, const BrickGeometry& g, double quantum,
                       const uint16_t* rank, T* out, hipStream_t s) {
    const uint64_t bricks = uint64_t(g.nbx) * g.nby * g.nbz;
    check_launch_size(bricks * 256);
    hipLaunchKernelGGL((dequantize_kernel<T>), dim3(uint32_t(bricks)), dim3(256), 0, s, mag, sign, g, T(quantum),
                       rank, out);
}

template void quantize_bricks<float>(float*, const BrickGeometry&, double, const uint16_t*, uint16_t*, uint32_t*,
                                     SideList, uint32_t*, hipStream_t);
template void quantize_bricks<double>(double*, const BrickGeometry&, double, const uint16_t*, uint16_t*, uint32_t*,
                                      SideList, uint32_t*, hipStream_t);
template void dequantize_bricks<float>(const uint16_t*, const uint32_t*, const BrickGeometry&, double,
                                       const uint16_t*, float*, hipStream_t);
template void dequantize_bricks<double>(const uint16_t*, const uint32_t*, const BrickGeometry&, double,
                                        const uint16_t*, double*, hipStream_t);

}  // namespace wz
