// Fused one-launch-per-level CDF 9/7 lifting for the dyadic schedule.
//
// Forward level (order Y, X, Z; cpu/src/transform.cpp:425-430): a workgroup of
// 576 threads (9 wave64s) owns a 64 x 32 output tile in XY and one z segment.
// For every pair of input planes of the segment (plus a 4-sample reflected halo
// on each side) it loads the two 72 x 40 reflected XY footprints into LDS,
// lifts Y (tile 8) and X (tile 8) from LDS, and feeds the result into a
// per-position Z lifting pipeline held in registers (5 values per position,
// the steps of cpu/src/transform.cpp:136-145 in streaming order). Forward
// lifting is bit-neutral to tiling, so every coefficient equals the
// reference's. Outputs of each z pair are staged in LDS and written with
// 16-byte stores: the LLL band to a compact next-level box, the seven other
// bands to the coefficient volume at their split-band positions.
//
// Inverse level (order Z, X, Y; cpu/src/transform.cpp:420-424) streams Z over
// the 72 x 40 footprint of split coefficients in canonical tiles of 16 (f32) or
// 8 (f64) samples: the normal lifting chain runs per pair, and the three values
// that a tile's endpoint fold changes (its first two and its last sample,
// cpu/src/transform.cpp:118-129) are computed by side chains at fixed phases
// of the tile. Each pair of reconstructed z planes is then lifted in X (tile 8,
// unfolded) and Y (tile 8, folded) from LDS and emitted with coalesced rows.
#include "common.hpp"
#include "fused.hpp"
#include "launch.hpp"

#include <type_traits>
#include <utility>

namespace wz {
namespace {

constexpr int TX = 64, TY = 32;            // output tile
constexpr int FW = TX + 8, FH = TY + 8;    // reflected footprint
constexpr int PA = FW + 1;                 // LDS pitch (odd: spreads banks)
constexpr int FT = 576;                    // threads (9 wave64s) per workgroup
constexpr int FSLOTS = FW * FH / FT;       // 5 footprint samples per thread and plane
static_assert(FW * FH == FSLOTS * FT, "footprint must tile the workgroup");
// The inverse level runs 8 waves (two per SIMD): six footprint slots per
// thread, the last slot of threads past the footprint repeating their fifth
// (the same column computed twice and written to the same place).
constexpr int IT = 512;
constexpr int ISLOTS = (FW * FH + IT - 1) / IT;
static_assert(ISLOTS == 6 && FW * FH > (ISLOTS - 1) * IT, "inverse footprint slots");

// Workgroup -> tile, XCD-aware. The dispatcher deals consecutive workgroups
// round-robin to the 8 XCDs, each with its own L2; XCD k is given the k-th
// contiguous slice of the x-fastest tile order instead, so tiles that share
// halo rows and columns run on the same XCD at about the same time and the
// halo lines come from its L2. Grid = 8 * ceil(tiles / 8); spare workgroups
// return at once.
struct Tile {
    uint32_t bx, by, bz, id;
};
WZ_D bool tile_of(Dims grid, Tile& t) {
    const uint32_t total = grid.nx * grid.ny * grid.nz, per = (total + 7) / 8;
    t.id = (blockIdx.x & 7u) * per + (blockIdx.x >> 3);
    if (t.id >= total) return false;
    t.bx = t.id % grid.nx;
    t.by = (t.id / grid.nx) % grid.ny;
    t.bz = t.id / (grid.nx * grid.ny);
    return true;
}
WZ_HD Dims tile_grid(Dims n, uint32_t sz) {
    return Dims{(n.nx + TX - 1) / TX, (n.ny + TY - 1) / TY, (n.nz + sz - 1) / sz};
}
WZ_HD uint32_t xcd_blocks(Dims g) { return 8 * ((g.nx * g.ny * g.nz + 7) / 8); }

template <class T, int P>
WZ_D void fwd_lift(T (&e)[P], T (&o)[P]) {
    using C = Lift<T>;
#pragma unroll
    for (int j = 0; j < P - 1; ++j) o[j] = fmadd(C::a, e[j] + e[j + 1], o[j]);
#pragma unroll
    for (int j = 1; j < P - 1; ++j) e[j] = fmadd(C::b, o[j - 1] + o[j], e[j]);
#pragma unroll
    for (int j = 1; j < P - 2; ++j) o[j] = fmadd(C::g, e[j] + e[j + 1], o[j]);
#pragma unroll
    for (int j = 2; j < P - 2; ++j) e[j] = fmadd(C::d, o[j - 1] + o[j], e[j]);
#pragma unroll
    for (int j = 2; j < P - 2; ++j) {
        e[j] = e[j] * C::k;
        o[j] = o[j] * (-C::ik);
    }
}

// Raw buffer access with a wave-uniform base (a 4 GiB window: stride 0, no
// range limit, dword data format) and a 32-bit per-lane byte offset: no
// per-lane 64-bit address arithmetic.
WZ_D __amdgpu_buffer_rsrc_t window(const void* base) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, -1, 0x00020000);
}
// The same with a byte range: loads beyond it read 0, stores beyond it are
// dropped (how the kernels mask rows and lanes without branching, which keeps
// every step's memory operations and so the memory-counter waits the same).
WZ_D __amdgpu_buffer_rsrc_t window_n(const void* base, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, int(bytes), 0x00020000);
}
constexpr uint32_t kOutside = 0xFFFFFFF0u;  // a byte offset beyond every window_n range
template <class T>
WZ_D T buf_load(__amdgpu_buffer_rsrc_t r, uint32_t off);
template <>
WZ_D float buf_load<float>(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, int(off), 0, 0));
}
template <>
WZ_D double buf_load<double>(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, int(off), 0, 0));
}
// Streaming (non-temporal) load of data read exactly once (the repair
// emitter's originals), so it does not evict the coefficient lines that
// neighbouring tiles' halos read again.
WZ_D float buf_load_nt(float, __amdgpu_buffer_rsrc_t r, uint32_t off) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, int(off), 0, 2));
}
WZ_D double buf_load_nt(double, __amdgpu_buffer_rsrc_t r, uint32_t off) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, int(off), 0, 2));
}
// Streaming (non-temporal) stores of a level's output: they would otherwise
// evict the input lines that neighbouring tiles' halos read again from L2.
WZ_D void buf_store_nt(float v, __amdgpu_buffer_rsrc_t r, uint32_t off) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, int(off), 0, 2);
}
WZ_D void buf_store_nt(double v, __amdgpu_buffer_rsrc_t r, uint32_t off) {
    using U2 = unsigned __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(U2, v), r, int(off), 0, 2);
}
WZ_D void buf_store(float v, __amdgpu_buffer_rsrc_t r, uint32_t off) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, int(off), 0, 0);
}
WZ_D void buf_store(double v, __amdgpu_buffer_rsrc_t r, uint32_t off) {
    using U2 = unsigned __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(U2, v), r, int(off), 0, 0);
}

// inv_lift streamed left to right: ld(k, even, odd) supplies pair k and
// out(j, even, odd) receives finished pair j (2 <= j <= P-3) as soon as it is
// final, so only four lifting values stay live instead of 2P. Every output is
// the same sequence of operations on the same operands as in inv_lift, hence
// the same bits.
template <class T, int P, bool Folded, class Ld, class Out>
WZ_D void inv_lift_stream(Ld&& ld, Out&& out) {
    using C = Lift<T>;
    T first_odd = T(0), po0 = T(0), pe1 = T(0), po2 = T(0), pe3 = T(0);
#pragma unroll
    for (int k = 0; k < P; ++k) {
        T E, O;
        ld(k, E, O);
        if (k == 0) {  // e[0] is never read; o[0] only through the sum (or the fold)
            if (Folded) first_odd = O;
            else po0 = O * (-C::k);
            continue;
        }
        const T o0 = O * (-C::k), e0 = E * C::ik;
        T s;
        if (Folded && k == 1) s = fmadd(first_odd, -C::k, o0);
        else if (Folded && k == P - 1) s = fmadd(O, -C::k, po0);
        else s = po0 + o0;
        const T e1 = fmadd(-C::d, s, e0);                 // e1[k]
        if (k >= 2) {
            const T o2 = fmadd(-C::g, pe1 + e1, po0);     // o2[k-1]
            if (k >= 3) {
                const T e3 = fmadd(-C::b, po2 + o2, pe1); // e3[k-1]
                if (k >= 4) out(k - 2, pe3, fmadd(-C::a, pe3 + e3, po2));  // o4[k-2]
                pe3 = e3;
            }
            po2 = o2;
        }
        po0 = o0;
        pe1 = e1;
    }
}

// Lane L of v set to the scalar s (v_writelane: one vector instruction, no
// lane-index compare; the lane is an immediate).
template <int L>
WZ_D uint32_t write_lane(uint32_t v, uint32_t s) {
    static_assert(L >= 0 && L < 64, "lane");
    asm("v_writelane_b32 %0, %1, %2" : "+v"(v) : "s"(s), "n"(L));
    return v;
}
// Lane 2r + h of the result: word h of the 64-bit wave mask m[r].
template <int... R>
WZ_D uint32_t bitmap_words(const unsigned long long (&m)[8], std::integer_sequence<int, R...>) {
    uint32_t v = 0;
    ((v = write_lane<2 * R>(v, uint32_t(m[R])), v = write_lane<2 * R + 1>(v, uint32_t(m[R] >> 32))), ...);
    return v;
}

// Four consecutive staged values to a 16-byte aligned destination.
// (non-temporal, see buf_store_nt)
WZ_D void store_vec4(float* dst, const float* v) {
    using F4 = float __attribute__((ext_vector_type(4)));
    __builtin_nontemporal_store(*reinterpret_cast<const F4*>(v), reinterpret_cast<F4*>(dst));
}
WZ_D void store_vec4(double* dst, const double* v) {
    using D2 = double __attribute__((ext_vector_type(2)));
    __builtin_nontemporal_store(reinterpret_cast<const D2*>(v)[0], reinterpret_cast<D2*>(dst));
    __builtin_nontemporal_store(reinterpret_cast<const D2*>(v)[1], reinterpret_cast<D2*>(dst) + 1);
}

}  // namespace

// Forward level, 576 threads (9 wave64s) per 64 x 32 x sz tile; one step per
// z pair (two planes):
//   1. every thread stores its 5 footprint samples of both planes (loaded one
//      step ahead) into A;
//   2. Y: one (plane, column, 16-row window) per thread, into B;
//   3. X: one (plane, row, 16-column window) per thread (512 of them), into X
//      (aliasing A), low band at columns 0..31 and high band at 32..63;
//   4. Z: 512 threads each stream 4 columns of a row through the Z lifting
//      pipeline (five registers per column) and stage the two z bands of the
//      finished pair in C (aliasing B);
//   5. the staged rows are written with 16-byte stores: the LLL band to the
//      next level's box, the seven other bands to the coefficient volume.
constexpr int CP = TX + 4;             // row pitch of the X and C stages (16-byte rows)

template <class T>
__global__ void __launch_bounds__(FT) fwd_level_kernel(FwdLevel<T> a) {
    constexpr int ASZ = 2 * FH * PA, BSZ = 2 * TY * PA;
    static_assert(2 * TY * CP <= ASZ && 2 * TY * CP <= BSZ, "stage aliasing");
    __shared__ __attribute__((aligned(16))) T A[ASZ];  // two input planes; then X results [plane][row][CP]
    __shared__ __attribute__((aligned(16))) T B[BSZ];  // after Y [plane][row][PA]; then C [zband][row][CP]
    T* const Xs = A;
    T* const Cs = B;
    const uint32_t t = threadIdx.x;
    Tile tl;
    if (!tile_of(tile_grid(a.n, a.sz), tl)) return;
    const uint32_t x0 = tl.bx * TX, y0 = tl.by * TY, z0 = tl.bz * a.sz;
    const Dims n = a.n;
    const uint32_t hx = half_len(n.nx), hy = half_len(n.ny), hz = half_len(n.nz);

    uint32_t loff[FSLOTS], aoff[FSLOTS];
#pragma unroll
    for (int s = 0; s < FSLOTS; ++s) {
        const uint32_t idx = t + FT * s, r = idx / FW, c = idx % FW;
        loff[s] = (reflect_pos(int32_t(y0) - 4 + int32_t(r), n.ny) * uint32_t(a.spy) +
                   reflect_pos(int32_t(x0) - 4 + int32_t(c), n.nx)) *
                  uint32_t(sizeof(T));  // byte offset in the plane
        aoff[s] = r * PA + c;
    }
    // Z ownership (threads < 512): row zr, columns 4 zq .. 4 zq + 3 of the X stage
    const uint32_t zr = t & 31u, zq = (t >> 5) & 15u;
    const bool zown = t < 512;
    T E[4], O[4], O1[4], E2[4], O3[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) E[k] = O[k] = O1[k] = E2[k] = O3[k] = T(0);

    // 16-byte stores need 4-aligned rows and band offsets in both destinations
    const bool vec_hi = (a.hpy & 3u) == 0 && (a.hpz & 3u) == 0 && (hx & 3u) == 0;
    const bool vec_lo = (a.lpy & 3u) == 0 && (a.lpz & 3u) == 0;

    const int npairs = int(a.sz / 2) + 4;
    // the two planes of the next step are in flight while this step runs
    T pe[FSLOTS], po[FSLOTS];
    {
        const auto w0 = window(a.src + uint64_t(reflect_pos(int32_t(z0) - 4, n.nz)) * a.spz);
        const auto w1 = window(a.src + uint64_t(reflect_pos(int32_t(z0) - 3, n.nz)) * a.spz);
#pragma unroll
        for (int s = 0; s < FSLOTS; ++s) {
            pe[s] = buf_load<T>(w0, loff[s]);
            po[s] = buf_load<T>(w1, loff[s]);
        }
    }
    for (int i = 0; i < npairs; ++i) {
        // 1. footprints of planes 2i, 2i+1 into A; next step's loads issued
        bool ok = true;
#pragma unroll
        for (int s = 0; s < FSLOTS; ++s) {
            ok = ok && __builtin_isfinite(pe[s]) && __builtin_isfinite(po[s]);
            A[aoff[s]] = pe[s];
            A[FH * PA + aoff[s]] = po[s];
        }
        if (i < npairs - 1) {
            const int32_t pn = int32_t(z0) - 4 + 2 * (i + 1);
            const auto w0 = window(a.src + uint64_t(reflect_pos(pn, n.nz)) * a.spz);
            const auto w1 = window(a.src + uint64_t(reflect_pos(pn + 1, n.nz)) * a.spz);
#pragma unroll
            for (int s = 0; s < FSLOTS; ++s) {
                pe[s] = buf_load<T>(w0, loff[s]);
                po[s] = buf_load<T>(w1, loff[s]);
            }
        }
        if (a.status && !ok) atomicOr(a.status, uint32_t(kNonFiniteInput));
        __syncthreads();
        {  // 2. Y: plane h, column c, rows 8s..8s+15 -> 4 low + 4 high pairs
            const uint32_t h = t / (4 * FW), rem = t % (4 * FW), c = rem % FW, sg = rem / FW;
            const T* src = A + h * FH * PA + c;
            T e[8], o[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                e[j] = src[(8 * sg + 2 * j) * PA];
                o[j] = src[(8 * sg + 2 * j + 1) * PA];
            }
            fwd_lift<T, 8>(e, o);
            T* dst = B + h * TY * PA + c;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                dst[(4 * sg + j) * PA] = e[j + 2];
                dst[(16 + 4 * sg + j) * PA] = o[j + 2];
            }
        }
        __syncthreads();
        if (t < 512) {  // 3. X: plane h, row r, columns 8u..8u+15 -> 4 low + 4 high pairs
            const uint32_t h = t >> 8, r = t & 31u, u = (t >> 5) & 7u;
            const T* src = B + h * TY * PA + r * PA + 8 * u;
            T e[8], o[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                e[j] = src[2 * j];
                o[j] = src[2 * j + 1];
            }
            fwd_lift<T, 8>(e, o);
            T* dst = Xs + h * TY * CP + r * CP + 4 * u;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                dst[q] = e[q + 2];
                dst[32 + q] = o[q + 2];
            }
        }
        __syncthreads();
        // 4. Z pipeline: pair i of the segment (cpu/src/transform.cpp:136-145, streamed)
        if (zown) {
            using C = Lift<T>;
            T ev[4], ov[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                ev[k] = Xs[zr * CP + 4 * zq + k];
                ov[k] = Xs[TY * CP + zr * CP + 4 * zq + k];
            }
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const T o1 = fmadd(C::a, E[k] + ev[k], O[k]);   // o1[i-1]
                const T e2 = fmadd(C::b, O1[k] + o1, E[k]);     // e2[i-1]
                const T o3 = fmadd(C::g, E2[k] + e2, O1[k]);    // o3[i-2]
                const T e4 = fmadd(C::d, O3[k] + o3, E2[k]);    // e4[i-2]
                if (i >= 4) {
                    Cs[zr * CP + 4 * zq + k] = e4 * C::k;
                    Cs[TY * CP + zr * CP + 4 * zq + k] = o3 * (-C::ik);
                }
                E[k] = ev[k];
                O[k] = ov[k];
                O1[k] = o1;
                E2[k] = e2;
                O3[k] = o3;
            }
        }
        __syncthreads();  // X (in A) consumed; C (in B) complete
        if (i < 4) continue;
        // 5. rows of the finished pair: 64 rows (z band, y row) x 16 groups of 4
        const uint32_t iz = z0 / 2 + uint32_t(i - 4);
        for (uint32_t w = t; w < 2 * TY * 16; w += FT) {
            const uint32_t row = w >> 4, q = w & 15u;
            const uint32_t zb = row / TY, rr = row % TY, yb = rr >> 4, iy = y0 / 2 + (rr & 15u);
            const uint32_t xb = q >> 3, ix = x0 / 2 + 4 * (q & 7u);
            if (!(zb ? 2 * iz + 1 < n.nz : 2 * iz < n.nz) || !(yb ? 2 * iy + 1 < n.ny : 2 * iy < n.ny)) continue;
            const T* v = Cs + row * CP + 4 * q;
            const bool low = !(xb | yb | zb);
            T* dst;
            if (low) {
                dst = a.lo + uint64_t(iz) * a.lpz + uint64_t(iy) * a.lpy + ix;
            } else {
                const uint64_t X = xb ? hx + ix : ix, Y = yb ? hy + iy : iy, Z = zb ? hz + iz : iz;
                dst = a.hi + Z * a.hpz + Y * a.hpy + X;
            }
            // samples 2*ix+xb .. of the level box exist up to nx
            const bool full = 2 * (ix + 3) + xb < n.nx;
            if (full && (low ? vec_lo : vec_hi)) {
                store_vec4(dst, v);
            } else {
#pragma unroll
                for (int k = 0; k < 4; ++k)
                    if (2 * (ix + k) + xb < n.nx) __builtin_nontemporal_store(v[k], dst + k);
            }
        }
    }
}

// Inverse level, 512 threads (8 wave64s, two per SIMD) per 64 x 32 x sz tile;
// one step per z pair: every thread streams 6 footprint columns (the last one
// a repeat for the threads past the 2880 columns) through the Z synthesis (with the
// tile-end fold side chain), refilling each column's registers with the next
// pair as soon as they are consumed, and stages the two reconstructed planes
// in Zp; X then runs on 320 (plane, row, 16-sample) items and Y on 512
// (plane, column, 8-sample) items, one per thread, whose lanes are 64
// consecutive x for coalesced emission; both lift their line streamed. Loads
// and stores go through buffer windows at wave-uniform plane or row bases.
// For the repair emitter each Y thread's 8 original samples are loaded as
// the step begins.
template <class T, int Emit, bool Requant = false, bool RowWords = true, bool Patch = false>
__global__ void __launch_bounds__(IT) inv_level_kernel(InvLevel<T> a) {
    constexpr int H = sizeof(T) == 4 ? 8 : 4;   // canonical Z tile / 2 (pairs per tile)
    __shared__ T Zp[2 * FH * PA];                // two reconstructed z planes (split XY)
    __shared__ T Xp[2 * FH * (TX + 1)];          // after X: rows still split in Y
    __shared__ uint32_t seg_fill;                // repair entries of this workgroup
    // Patch: each Y wave's slots of its step's 8 rows x 64 lanes (the lane's
    // repair, NaN for none)
    __shared__ T pslot[Patch ? IT * 8 : 1];
    using C = Lift<T>;
    const uint32_t t = threadIdx.x;
    Tile tl;
    if (!tile_of(tile_grid(a.n, a.sz), tl)) return;
    const uint32_t x0 = tl.bx * TX, y0 = tl.by * TY, zs = tl.bz * a.sz;
    const uint32_t wg = tl.id;
    const Dims n = a.n;
    const uint32_t hx = half_len(n.nx), hy = half_len(n.ny), hz = half_len(n.nz);
    // RowWords (repair emitter): nx % 32 == 0, so bitmap words never straddle rows
    if (t == 0) seg_fill = 0;

    // Footprint positions (row r, column c of the 72 x 40 window) in slot
    // order: first the 36 x 20 positions with even r and c (which read the
    // coarse LLL box in even z planes), then the rest row pair by row pair.
    // Waves then read from one base pointer (wave-uniform) with 32-bit byte
    // offsets; only the one wave that straddles the two groups selects per lane.
    constexpr uint32_t kLow = (FW / 2) * (FH / 2);
    // A position reads the LLL box exactly when its row and column are even
    // (a reflection keeps the parity, and even positions split into the low
    // band): slot order puts them first, so slot 0 is all low, slot 1 below
    // thread kLow - IT, the later slots never. Compile-time per slot, so no
    // per-slot lane masks stay live through the steps.
    static_assert(kLow >= uint32_t(IT) && kLow < 2u * uint32_t(IT), "slot classes");
    auto slot_low = [&](int s) -> bool { return s == 0 || (s == 1 && t < kLow - uint32_t(IT)); };
    uint32_t eoff[ISLOTS], ooff[ISLOTS];   // even-plane and odd-plane byte offsets
    uint32_t zpk[(ISLOTS + 1) / 2] = {};     // Zp offsets, two 16-bit halves per word
#pragma unroll
    for (int s = 0; s < ISLOTS; ++s) {
        const uint32_t idx = t + IT * s < uint32_t(FW * FH) ? t + IT * s : t + IT * (s - 1);
        uint32_t r, c;
        if (idx < kLow) {
            r = 2 * (idx / (FW / 2));
            c = 2 * (idx % (FW / 2));
        } else {
            const uint32_t j = idx - kLow, k = j % (FW / 2 + FW);
            r = 2 * (j / (FW / 2 + FW)) + (k < FW / 2 ? 0u : 1u);
            c = k < FW / 2 ? 2 * k + 1 : k - FW / 2;
        }
        const uint32_t cy = split_pos(int32_t(y0) - 4 + int32_t(r), n.ny);
        const uint32_t cx = split_pos(int32_t(x0) - 4 + int32_t(c), n.nx);
        const bool low = idx < kLow;  // = cx < hx && cy < hy
        const uint32_t ohi = (cy * uint32_t(a.hpy) + cx) * uint32_t(sizeof(T));
        eoff[s] = low ? (cy * uint32_t(a.lpy) + cx) * uint32_t(sizeof(T)) : ohi;
        ooff[s] = ohi;
        zpk[s / 2] |= (r * PA + c) << (16 * (s & 1));
    }
    auto at = [](const T* base, uint32_t off) {
        return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + off);
    };
    // Sources of pair g: the coefficient volume's planes of the even and odd
    // z sample and the LLL box's even plane (an even z sample always lies in
    // the low half, an odd one never)
    struct PairSrc {
        const T *he, *ho, *le;
    };
    auto pair_src = [&](int32_t g) {
        const uint32_t ze = split_pos(2 * g, n.nz), zo = split_pos(2 * g + 1, n.nz);
        return PairSrc{a.hi + uint64_t(ze) * a.hpz, a.hi + uint64_t(zo) * a.hpz, a.lo + uint64_t(ze) * a.lpz};
    };
    auto load_slot = [&](const PairSrc& q, int s, T& ev, T& ov) {
        ev = at(slot_low(s) ? q.le : q.he, eoff[s]);  // one load on every path
        ov = buf_load<T>(window(q.ho), ooff[s]);
    };
    // Values of slot s as the Z synthesis takes them (at use, so the loads
    // stay in flight): high-band values of the repair emitter on unquantized
    // bands are quantized and dequantized; LLL-box values never are.
    auto band = [&](T v) {
        if constexpr (Requant) return requantize<T>(v, a.q, a.inv);
        else return v;
    };
    auto even_band = [&](int s, T v) {
        if constexpr (!Requant) return v;
        else return slot_low(s) ? v : band(v);
    };
    // Y item of this thread (t < 512): plane yp, column yx, y tile yt; a wave
    // is one (plane, tile) and its lanes are 64 consecutive x
    const uint32_t wave = __builtin_amdgcn_readfirstlane(t >> 6);
    const bool yown = wave < 2 * (TY / 8);
    const uint32_t yp = wave / (TY / 8), yt = wave % (TY / 8), yx = t & 63u;
    const uint32_t X = x0 + yx;
    const bool xin = X < n.nx;
    // row Y of plane Z of a full-size field (wave-uniform) and this lane's byte offset in it
    auto row_of = [&](auto* base, uint32_t Z, uint32_t Y) { return base + uint64_t(Z) * a.opz + uint64_t(Y) * a.opy; };
    // status of all steps, reported once at the end
    bool all_finite = true, bound_ok = true;
    // original samples of the Y item for step g (repair emitter)
    // one window over plane Z (wave-uniform; none past the field) and per-lane
    // offsets, beyond the window for samples outside the field
    const uint32_t plane_bytes = n.ny * a.opy * uint32_t(sizeof(T));
    auto plane_window = [&](const T* base, uint32_t Z) {
        const bool in = Z < n.nz;
        return window_n(base + uint64_t(in ? Z : 0u) * a.opz, in ? plane_bytes : 0u);
    };
    // Patch (the decoder's repairs applied as samples are stored): in a Y
    // wave's step, lanes 8r .. 8r+7 hold list entries first_r - 1 .. first_r
    // + 6 of row r, first_r being the key table's entry for the row's
    // 64-sample tile. Lane 8r proves the entry current (the entry before it
    // lies below the tile: a stale entry fails this or the next test), lanes
    // 8r+1 .. 8r+7 hold the tile's first repairs. The table entries of the
    // next step load one step ahead (lane r holds row r's, spread to its
    // lanes by a cross-lane permute; measured faster than scalar loads, whose
    // waits share a counter with the LDS traffic), and the window loads as its
    // step begins, ahead of the coefficient loads, so waiting for it never
    // waits for those.
    const uint32_t pl = t & 63u;
    const uint32_t txn = (n.nx + TX - 1) / TX, tw = min(uint32_t(TX), n.nx - x0);
    uint32_t np = a.np;  // the list (planned decode: from the device record)
    const uint32_t* plist = a.pidx;
    const T* perrs = a.perr;
    if constexpr (Patch) {
        if (a.plan) {
            np = a.plan->ok ? a.plan->np : 0u;
            plist = reinterpret_cast<const uint32_t*>(a.archive + a.plan->pi);
            perrs = reinterpret_cast<const T*>(a.archive + a.plan->pv);
        }
    }
    const auto pidx = window_n(plist, Patch ? 4u * np : 0u);
    const auto perr = window_n(perrs, Patch ? uint32_t(sizeof(T)) * np : 0u);
    const auto ptab = window_n(a.ptable, Patch ? 4u * a.keys : 0u);
    auto tab_off = [&](int32_t g) {
        const uint32_t Z = 2 * uint32_t(g - 2) + yp, Y = y0 + 8 * yt + pl;
        return pl < 8 && g >= 2 && Z < n.nz && Y < n.ny ? 4u * ((Z * n.ny + Y) * txn + tl.bx) : kOutside;
    };
    uint32_t ptv = 0;           // lanes 0..7: row r's table entry of the next step
    uint32_t pfv = 0, pci = 0;  // first_r of the lane's row, the lane's entry index
    T pce = T(0);               // and its repair
    auto lane_off = [&](uint32_t Y) { return xin && Y < n.ny ? (Y * a.opy + X) * uint32_t(sizeof(T)) : kOutside; };
    // repair emitter: this workgroup's record segment (cap entries)
    const auto seg_index = window_n(a.rep.index + uint64_t(wg) * a.rep.cap, 4 * a.rep.cap);
    const auto seg_payload =
        window_n(static_cast<T*>(a.rep.payload) + uint64_t(wg) * a.rep.cap, uint32_t(sizeof(T)) * a.rep.cap);
    auto fetch_orig = [&](int32_t g, T (&ref)[8]) {
        const uint32_t Z = 2 * uint32_t(g - 2) + yp;
        const auto w = plane_window(a.orig, yown ? Z : n.nz);
#pragma unroll
        for (int j = 0; j < 8; ++j) ref[j] = buf_load_nt(T(0), w, lane_off(y0 + 8 * yt + j));
    };
    // streaming state per slot: OS = os[g-1], EA = eA[g-1], OB = oB[g-2],
    // EC = eC[g-2]; SF rotates through o_raw[B-2], eA'[B-1], oB'[B-1], eC'[B]
    T OS[ISLOTS], EA[ISLOTS], OB[ISLOTS], EC[ISLOTS], SF[ISLOTS];
#pragma unroll
    for (int s = 0; s < ISLOTS; ++s) OS[s] = EA[s] = OB[s] = EC[s] = SF[s] = T(0);

    const int32_t g0 = int32_t(zs / 2) - 2, g1 = int32_t((zs + a.sz) / 2) + 1;
    // float repair emitter: a step's originals are loaded as the step begins,
    // ahead of the next pair's coefficients, so they arrive during the Z and X
    // passes and the in-order memory counter never makes a column wait for
    // them (double: at use, for registers)
    constexpr bool early = Emit == kEmitRepair && sizeof(T) == 4;
    T ref[8];
    // ve/vo hold pair g's coefficients; each slot's registers are refilled
    // with pair g+1 as soon as the slot has consumed them (in flight during
    // the rest of the step)
    T ve[ISLOTS], vo[ISLOTS];
    {
        const PairSrc q = pair_src(g0);
#pragma unroll
        for (int s = 0; s < ISLOTS; ++s) load_slot(q, s, ve[s], vo[s]);
    }
    // One step per z pair. The warm-up steps (Z only) and the emitting steps
    // run as separate loops (and a wave without Y items would run its own),
    // so each loop issues the same memory operations every iteration and the
    // memory-counter waits stay exact (a merged loop makes a wave wait for
    // its own previous stores before it starts the Z pass).
    auto step = [&](int32_t g, auto emit_c, auto y_c) {
        const bool emit = emit_c, yw = y_c;  // compile-time for float, run-time for double
        if constexpr (Patch) {
            if (emit) {
                pfv = uint32_t(__builtin_amdgcn_ds_bpermute(int((pl >> 3) << 2), int(ptv)));
                const uint32_t pos = pfv - 1u + (pl & 7u);
                const bool held = pos < np;
                pci = __builtin_amdgcn_raw_buffer_load_b32(pidx, int(held ? 4u * pos : kOutside), 0, 0);
                pce = buf_load<T>(perr, held ? uint32_t(sizeof(T)) * pos : kOutside);
            }
            ptv = __builtin_amdgcn_raw_buffer_load_b32(ptab, int(tab_off(g + 1)), 0, 0);
        }
        if constexpr (early) {
            if (emit) fetch_orig(g, ref);
        }
        const PairSrc nq = pair_src(g < g1 ? g + 1 : g);
        const int phase = ((g % H) + H) % H;
#pragma unroll
        for (int s = 0; s < ISLOTS; ++s) {
            const T oraw = band(vo[s]);
            const T es = even_band(s, ve[s]) * C::ik, os = oraw * (-C::k);
            const T eA = fmadd(-C::d, OS[s] + os, es);   // eA[g]
            const T oB = fmadd(-C::g, EA[s] + eA, OS[s]); // oB[g-1]
            const T eC = fmadd(-C::b, OB[s] + oB, EA[s]); // eC[g-1]
            const T oD = fmadd(-C::a, EC[s] + eC, OB[s]); // oD[g-2]
            T even = EC[s], odd = oD;                     // samples 2(g-2), 2(g-2)+1
            if (phase == H - 1) {  // g = B-1: fold of tile B's first pair, eA'[B-1]
                SF[s] = fmadd(-C::d, fmadd(SF[s], -C::k, os), es);
            } else if (phase == 0) {  // g = B: oB'[B-1]
                SF[s] = fmadd(-C::g, SF[s] + eA, OS[s]);
            } else if (phase == 1) {  // g = B+1: eC'[B]; tile B-H's last sample oD''[B-1]
                SF[s] = fmadd(-C::b, SF[s] + oB, EA[s]);
                const T eAe = fmadd(-C::d, fmadd(oraw, -C::k, OS[s]), es);   // eA''[B+1]
                const T oBe = fmadd(-C::g, EA[s] + eAe, OS[s]);             // oB''[B]
                const T eCe = fmadd(-C::b, OB[s] + oBe, EA[s]);             // eC''[B]
                odd = fmadd(-C::a, EC[s] + eCe, OB[s]);
            } else if (phase == 2) {  // g = B+2: tile B's first two samples
                even = SF[s];
                odd = fmadd(-C::a, SF[s] + eC, OB[s]);
            }
            if (phase == H - 2) SF[s] = oraw;  // o_raw[B-2] for the next tile's fold
            if (emit) {
                const uint32_t zo = (zpk[s / 2] >> (16 * (s & 1))) & 0xFFFFu;
                Zp[zo] = even;
                Zp[FH * PA + zo] = odd;
            }
            OS[s] = os;
            EA[s] = eA;
            OB[s] = oB;
            EC[s] = eC;
            // the slot's registers take the next pair once nothing reads them
            // (the last step reloads its own pair, unused)
            load_slot(nq, s, ve[s], vo[s]);
        }
        if (!emit) return;
        __syncthreads();
        // X (unfolded, so any tile width gives the same bits; 16 samples per
        // item keeps the 320 items in one round): plane p, row rr, tile xt
        for (uint32_t w = t; w < 2 * FH * (TX / 16); w += IT) {
            const uint32_t p = w / (FH * (TX / 16)), rem = w % (FH * (TX / 16));
            const uint32_t rr = rem % FH, xt = rem / FH;
            const T* src = Zp + p * FH * PA + rr * PA + 16 * xt;
            T* dst = Xp + p * FH * (TX + 1) + rr * (TX + 1) + 16 * xt;
            inv_lift_stream<T, 12, false>(
                [&](int k, T& e, T& o) {
                    e = src[2 * k];
                    o = src[2 * k + 1];
                },
                [&](int j, T e, T o) {
                    dst[2 * (j - 2)] = e;
                    dst[2 * (j - 2) + 1] = o;
                });
        }
        __syncthreads();
        if (!yw) return;
        // Y (folded, tile 8) of this thread's column -> 8 samples
        const uint32_t Z = 2 * uint32_t(g - 2) + yp;
        const bool xz = xin && Z < n.nz;
        const T* src = Xp + yp * FH * (TX + 1) + 8 * yt * (TX + 1) + yx;
        auto ld = [&](int k, T& e, T& o) {
            e = src[(2 * k) * (TX + 1)];
            o = src[(2 * k + 1) * (TX + 1)];
        };
        if constexpr (Emit != kEmitRepair) {
            bool fin = true;
            const auto wout = plane_window(a.out, Z);
            T pv[8];  // Patch: this lane's repair in each row, NaN for none
            if constexpr (Patch) {
                const uint32_t r = pl >> 3, k = pl & 7u, Yr = y0 + 8 * yt + r;
                const bool rin = Z < n.nz && Yr < n.ny;
                const uint32_t lo = (Z * n.ny + Yr) * n.nx + x0;  // index of the row tile's first sample
                const bool held = rin && pfv - 1u + k < np;
                const bool in_tile = held && k && pci - lo < tw;
                const unsigned long long cur = __ballot(k == 0 && rin && (pfv == 0 || (held && pci < lo)));
                const unsigned long long opens = __ballot(k == 1 && in_tile);
                const bool apply = in_tile && ((cur >> (8 * r)) & 1ull) && ((opens >> (8 * r + 1)) & 1ull);
                T* slot = pslot + wave * 512u;
#pragma unroll
                for (int j = 0; j < 8; ++j) slot[64 * j + pl] = T(__builtin_nan(""));
                // lanes of one wave exchange through LDS in issue order
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                if (apply) slot[64 * r + (pci - lo)] = pce;
                // a window still in its tile at its last lane: the rest of the
                // tile's repairs (at most 64 per tile), 64 entries at a time
                unsigned long long more = __ballot(apply && k == 7);
                while (more) {
                    const uint32_t L = uint32_t(__builtin_ctzll(more));
                    more &= more - 1;
                    const uint32_t lo2 = __builtin_amdgcn_readlane(lo, L);
                    uint32_t base = __builtin_amdgcn_readlane(pfv, L) + 7u;
                    for (;;) {
                        const uint32_t pos = base + pl;
                        const bool in = pos < np;
                        const uint32_t ci = __builtin_amdgcn_raw_buffer_load_b32(pidx, int(in ? 4u * pos : kOutside), 0, 0);
                        const T ce = buf_load<T>(perr, in ? uint32_t(sizeof(T)) * pos : kOutside);
                        const bool hit = in && ci - lo2 < tw;
                        if (hit) slot[64 * (L >> 3) + (ci - lo2)] = ce;
                        if (!((__ballot(hit) >> 63) & 1ull)) break;
                        base += 64;
                    }
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
                for (int j = 0; j < 8; ++j) pv[j] = slot[64 * j + pl];
            }
            auto put = [&](uint32_t j, T v) {
                const uint32_t Y = y0 + 8 * yt + j;
                if constexpr (Patch) {
                    const T e = pv[j];
                    if (e == e) v = v - e;  // cpu/src/codec.cpp:330
                }
                buf_store_nt(v, wout, lane_off(Y));
                fin = fin && (!(xz && Y < n.ny) || __builtin_isfinite(v));
            };
            inv_lift_stream<T, 8, true>(ld, [&](int j, T e, T o) {
                put(2 * (j - 2), e);
                put(2 * (j - 2) + 1, o);
            });
            all_finite = all_finite && fin;
        } else {
            if constexpr (!early) fetch_orig(g, ref);
            // decide each sample as the lifting finishes it; then the step's
            // eight row masks go out as one 16-lane store of bitmap words and
            // its records take one LDS reservation
            const uint32_t lane = t & 63u;
            const T nb = T(a.bound);
            // Wave masks: ballots of single compares combined in scalar
            // registers (a ballot of a combined predicate costs a vector
            // select and compare); lane predicates come back from them.
            const unsigned long long xz_m = __ballot(xz);  // lanes inside the field
            // rows of the Y item inside the field; the empty asm keeps the
            // eight row tests in the step (hoisted, they were loop-invariant
            // scalars spilled to vector lanes and restored every step)
            uint32_t rows_in = __builtin_amdgcn_readfirstlane(n.ny > y0 + 8 * yt ? n.ny - (y0 + 8 * yt) : 0u);
            asm volatile("" : "+s"(rows_in));
            bool ok = true;
            unsigned long long m8[8];  // wave-uniform patch masks of the eight rows
            T err8[8];                 // read only where the sample is patched
            auto decide = [&](uint32_t j, T v, T orig) {
                const unsigned long long in_m = j < rows_in ? xz_m : 0ull;
                if constexpr (sizeof(T) == 4) {
                    // the same float operations as repair_decision; its exact
                    // bound test is settled by the rounded difference d of
                    // the final value and the original: |d| < the largest
                    // float <= bound implies |exact difference| < bound
                    // (round to nearest is monotonic), which also implies
                    // every operand was finite. Other lanes (ties at the
                    // bound, non-finite samples) take the full test.
                    // The full test gives the fast lanes the same result, so a
                    // wave with any slow lane runs it on all lanes (one
                    // wave-uniform branch, no per-lane one).
                    const T e = v - orig;
                    const bool p = abs_t(e) > nb;
                    const T d = p ? (v - e) - orig : e;
                    const bool fast = abs_t(d) < T(a.bound_lo);
                    unsigned long long pm = __ballot(p) & in_m;
                    T err = e;
                    if (in_m & ~__ballot(fast)) {
                        const bool inb = __builtin_amdgcn_inverse_ballot_w64(in_m);
                        bool okj = true;
                        T ej = T(0);
                        const bool pj = repair_decision<T>(v, orig, nb, a.bound, &ej, &okj);
                        pm = __ballot(pj) & in_m;
                        err = ej;
                        ok = ok && (!inb || okj);
                    }
                    m8[j] = pm;
                    err8[j] = err;
                } else {
                    const bool inb = __builtin_amdgcn_inverse_ballot_w64(in_m);
                    bool patch = false;
                    T err = T(0);
                    if (inb) patch = repair_decision<T>(v, orig, nb, a.bound, &err, &ok);
                    m8[j] = __ballot(patch);
                    err8[j] = err;
                }
            };
            inv_lift_stream<T, 8, true>(ld, [&](int j, T e, T o) {
                decide(2 * (j - 2), e, ref[2 * (j - 2)]);
                decide(2 * (j - 2) + 1, o, ref[2 * (j - 2) + 1]);
            });
            // sample index of this lane in row y0 + 8 yt of plane Z (indices fit 32 bits)
            const uint32_t row0 = uint32_t((uint64_t(Z) * n.ny + y0 + 8 * yt) * n.nx) + X;
            if constexpr (!RowWords) {  // bitmap words straddle rows: one atomic per patched sample
#pragma unroll
                for (int r = 0; r < 8; ++r)
                    if (__builtin_amdgcn_inverse_ballot_w64(m8[r])) {
                        const uint32_t at = row0 + uint32_t(r) * n.nx;
                        atomicOr(&a.rep.bitmap[at >> 5], 1u << (at & 31u));
                    }
            } else {  // 64 consecutive x at a 32-aligned row start: each row owns two words
                // lane 2r + h holds word h of row r, written from the masks
                const uint32_t wv = bitmap_words(m8, std::make_integer_sequence<int, 8>{});
                // lanes 2r + h that store: rows inside the field, and the
                // second word only where the row tile reaches x0 + 32
                const uint64_t rows_m = rows_in >= 8 ? 0xFFFFull : (1ull << (2 * rows_in)) - 1ull;
                const uint64_t put_m = Z < n.nz ? rows_m & (x0 + 32 < n.nx ? 0xFFFFull : 0x5555ull) : 0ull;
                const bool put = __builtin_amdgcn_inverse_ballot_w64(put_m);
                const uint32_t words = uint32_t((uint64_t(n.nx) * n.ny * n.nz + 31) / 32);
                __builtin_amdgcn_raw_buffer_store_b32(
                    wv, window_n(a.rep.bitmap, 4 * words),
                    int(put ? 4 * (((row0 - yx + (lane >> 1) * n.nx) >> 5) + (lane & 1u)) : kOutside), 0, 0);
            }
            uint32_t total = 0;
#pragma unroll
            for (int r = 0; r < 8; ++r) total += uint32_t(__popcll(m8[r]));
            uint32_t base = 0;
            if (total && lane == 0) base = atomicAdd(&seg_fill, total);
            base = __builtin_amdgcn_readfirstlane(base);
            // this step's entries start at a wave-uniform slot of the segment;
            // lanes without an entry store beyond the windows (dropped)
            uint32_t before = base;  // entries of the earlier steps and rows
#pragma unroll
            for (int r = 0; r < 8; ++r) {
                const unsigned long long m = m8[r];
                // entries of lower lanes: mbcnt counts the mask bits below this lane
                const uint32_t k = before + __builtin_amdgcn_mbcnt_hi(uint32_t(m >> 32),
                                                                      __builtin_amdgcn_mbcnt_lo(uint32_t(m), 0u));
                const bool mine = __builtin_amdgcn_inverse_ballot_w64(m);
                __builtin_amdgcn_raw_buffer_store_b32(row0 + uint32_t(r) * n.nx, seg_index, int(mine ? 4 * k : kOutside), 0, 0);
                buf_store(err8[r], seg_payload, mine ? uint32_t(sizeof(T)) * k : kOutside);
                before += uint32_t(__popcll(m));
            }
            bound_ok = bound_ok && ok;
        }
    };
    const int32_t gE = int32_t(zs / 2) + 2;  // first step that emits
    int32_t g = g0;
    if constexpr (sizeof(T) == 4) {
        if (yown) {
            for (; g < gE; ++g) step(g, std::false_type{}, std::true_type{});
            // as many dropped stores as an emitting step issues, so the
            // emitting loop is entered with the memory-counter state it keeps
            // (the counter tracks stores too) and its waits stay exact; each
            // at its own offset past the empty window, or the compiler keeps
            // only the last of identical stores
            constexpr int kStores = Emit == kEmitRepair ? 17 : 8;
            const auto none = window_n(a.out ? static_cast<const void*>(a.out) : a.hi, 0u);
#pragma unroll
            for (int k = 0; k < kStores; ++k)
                __builtin_amdgcn_raw_buffer_store_b32(0u, none, int(kOutside - 16u * uint32_t(k)), 0, 0);
            for (; g <= g1; ++g) step(g, std::true_type{}, std::true_type{});
        } else {
            for (; g < gE; ++g) step(g, std::false_type{}, std::false_type{});
            for (; g <= g1; ++g) step(g, std::true_type{}, std::false_type{});
        }
    } else {  // double: one loop (three copies would not fit the registers)
        for (; g <= g1; ++g) step(g, g >= gE, yown);
    }
    {
        const bool lead = (t & 63u) == 0;
        if (Emit == kEmitStoreScreen && __ballot(!all_finite) && lead) atomicOr(a.status, uint32_t(kNonFiniteOutput));
        if (Emit == kEmitRepair && __ballot(!bound_ok) && lead) atomicOr(a.status, uint32_t(kRepairBound));
    }
    if constexpr (Emit == kEmitRepair) {
        __syncthreads();
        if (t == 0) a.rep.count[wg] = seg_fill;
    }
}

// ------------------------------------------------------------------ launchers
template <class T>
void launch_fwd_level(const FwdLevel<T>& a, hipStream_t s) {
    // the kernel addresses an input plane with 32-bit byte offsets
    if (!fused_plane_fits(a.spy, a.n.ny, sizeof(T))) throw std::logic_error("fused forward level: plane exceeds 4 GiB");
    const dim3 grid(xcd_blocks(tile_grid(a.n, a.sz)));
    hipLaunchKernelGGL((fwd_level_kernel<T>), grid, dim3(FT), 0, s, a);
}
template <class T>
void launch_inv_level(const InvLevel<T>& a, hipStream_t s) {
    // the kernel addresses a plane with 32-bit byte offsets
    if (!fused_plane_fits(a.hpy, a.n.ny, sizeof(T)) || !fused_plane_fits(a.lpy, half_len(a.n.ny), sizeof(T)) ||
        !fused_plane_fits(a.opy, a.n.ny, sizeof(T)))
        throw std::logic_error("fused inverse level: plane exceeds 4 GiB");
    const dim3 grid(xcd_blocks(tile_grid(a.n, a.sz)));
    const bool words = (a.n.nx & 31u) == 0;
    if (a.emit == kEmitRepair && a.q > T(0) && words)
        hipLaunchKernelGGL((inv_level_kernel<T, kEmitRepair, true>), grid, dim3(IT), 0, s, a);
    else if (a.emit == kEmitRepair && a.q > T(0))
        hipLaunchKernelGGL((inv_level_kernel<T, kEmitRepair, true, false>), grid, dim3(IT), 0, s, a);
    else if (a.emit == kEmitRepair && words)
        hipLaunchKernelGGL((inv_level_kernel<T, kEmitRepair>), grid, dim3(IT), 0, s, a);
    else if (a.emit == kEmitRepair)
        hipLaunchKernelGGL((inv_level_kernel<T, kEmitRepair, false, false>), grid, dim3(IT), 0, s, a);
    else if (a.emit == kEmitStoreScreen && a.np)
        hipLaunchKernelGGL((inv_level_kernel<T, kEmitStoreScreen, false, true, true>), grid, dim3(IT), 0, s, a);
    else if (a.emit == kEmitStoreScreen)
        hipLaunchKernelGGL((inv_level_kernel<T, kEmitStoreScreen>), grid, dim3(IT), 0, s, a);
    else if (a.q > T(0))  // a coarser level of the split compress inverse
        hipLaunchKernelGGL((inv_level_kernel<T, kEmitStore, true>), grid, dim3(IT), 0, s, a);
    else
        hipLaunchKernelGGL((inv_level_kernel<T, kEmitStore>), grid, dim3(IT), 0, s, a);
}

template <class T>
uint32_t repair_z_segment(Dims n, uint32_t tuned) {
    // workgroups of the repair emitter one CU holds at once (registers, LDS)
    static const uint32_t per_cu = [] {
        int b = 0;
        int b2 = 0;  // both repair variants: the smaller residency
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, inv_level_kernel<T, kEmitRepair>, IT, 0) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&b2, inv_level_kernel<T, kEmitRepair, true>, IT, 0) !=
                hipSuccess)
            b = b2 = 1;
        b = b < b2 ? b : b2;
        if (b < 1) b = 1;
        return uint32_t(b);
    }();
    int dev = 0, cus = 0;
    hip_check(hipGetDevice(&dev), "device");
    hip_check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "CU count");
    const uint64_t resident = uint64_t(cus) * per_cu;
    const uint32_t top = (n.nz + 15) & ~15u;
    uint32_t sz = tuned;
    auto tiles = [&](uint32_t z) {
        const Dims g = tile_grid(n, z);
        return uint64_t(g.nx) * g.ny * g.nz;
    };
    while (sz < top && tiles(sz) > resident) sz = sz * 2 < top ? sz * 2 : top;
    return sz;
}

template uint32_t repair_z_segment<float>(Dims, uint32_t);
template uint32_t repair_z_segment<double>(Dims, uint32_t);
template void launch_fwd_level<float>(const FwdLevel<float>&, hipStream_t);
template void launch_fwd_level<double>(const FwdLevel<double>&, hipStream_t);
template void launch_inv_level<float>(const InvLevel<float>&, hipStream_t);
template void launch_inv_level<double>(const InvLevel<double>&, hipStream_t);

}  // namespace wz
