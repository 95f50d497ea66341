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
#include "bricks.hpp"

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

// Packed rank from the geometry's tabulated edge shapes.
struct TableRank {
    const uint16_t* rank;
    WZ_D uint32_t operator()(const Brick& k, uint32_t i, uint32_t, uint32_t, uint32_t) const {
        return rank[k.shape * 8192 + i];
    }
};

template <class T>
__global__ void __launch_bounds__(256) quantize_kernel(T* __restrict__ coeff, BrickGeometry g, T q, T inv,
                                                       const uint16_t* __restrict__ rank, uint16_t* __restrict__ mag,
                                                       uint32_t* __restrict__ sign, SideList qo,
                                                       uint32_t* __restrict__ status) {
    quantize_brick(coeff, g, q, inv, TableRank{rank}, mag, sign, qo, status, blockIdx.x);
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
                                                         const DecodeRecord* __restrict__ plan,
                                                         const uint16_t* __restrict__ rank, T* __restrict__ out) {
    if (plan) {  // a planned decode: the quantum of the header checked on the device
        if (!plan->ok) return;
        q = T(plan->quantum);
    }
    const Brick k = brick_of(g, blockIdx.x);
    const uint32_t vol = k.bx * k.by * k.bz;
    constexpr uint32_t kBurst = 8;  // all loads of a burst issued before its stores
    for (uint32_t i0 = 0; i0 < vol; i0 += kBurst * 256u) {
        uint32_t mv[kBurst], sv[kBurst], pv[kBurst];
#pragma unroll
        for (uint32_t u = 0; u < kBurst; ++u) {
            const uint32_t i = i0 + u * 256u + threadIdx.x;
            const uint64_t p = k.base + rank[k.shape * 8192 + (i < vol ? i : 0u)];
            pv[u] = uint32_t(p & 31u);
            mv[u] = mag[p];
            sv[u] = sign[p >> 5];
        }
#pragma unroll
        for (uint32_t u = 0; u < kBurst; ++u) {
            const uint32_t i = i0 + u * 256u + threadIdx.x;
            if (i >= vol) continue;
            uint32_t x, y, z;
            brick_xyz(k, i, x, y, z);
            const uint64_t at = (uint64_t(k.z0 + z) * g.d.ny + (k.y0 + y)) * g.d.nx + k.x0 + x;
            const T v = T(mv[u]) * q;  // cpu/src/codec.cpp:146-147
            out[at] = ((sv[u] >> pv[u]) & 1u) ? -v : v;
        }
    }
}

template <class T>
void dequantize_bricks(const uint16_t* mag, const uint32_t* sign, const BrickGeometry& g, double quantum,
                       const uint16_t* rank, T* out, hipStream_t s, const DecodeRecord* plan) {
    const uint64_t bricks = uint64_t(g.nbx) * g.nby * g.nbz;
    check_launch_size(bricks * 256);
    hipLaunchKernelGGL((dequantize_kernel<T>), dim3(uint32_t(bricks)), dim3(256), 0, s, mag, sign, g, T(quantum),
                       plan, rank, out);
}

template void quantize_bricks<float>(float*, const BrickGeometry&, double, const uint16_t*, uint16_t*, uint32_t*,
                                     SideList, uint32_t*, hipStream_t);
template void quantize_bricks<double>(double*, const BrickGeometry&, double, const uint16_t*, uint16_t*, uint32_t*,
                                      SideList, uint32_t*, hipStream_t);
template void dequantize_bricks<float>(const uint16_t*, const uint32_t*, const BrickGeometry&, double,
                                       const uint16_t*, float*, hipStream_t, const DecodeRecord*);
template void dequantize_bricks<double>(const uint16_t*, const uint32_t*, const BrickGeometry&, double,
                                        const uint16_t*, double*, hipStream_t, const DecodeRecord*);

}  // namespace wz
