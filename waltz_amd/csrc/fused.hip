// Fused one-launch-per-level CDF 9/7 lifting for the dyadic schedule.
//
// Forward level (order Y, X, Z; cpu/src/transform.cpp:425-430): a workgroup of
// 320 threads (5 wave64s) owns a 64 x 32 output tile in XY and one z segment.
// For every input plane of the segment (plus a 4-sample reflected halo on each
// side) it loads the 72 x 40 reflected XY footprint into LDS, lifts Y (tile 8)
// and X (tile 8) from LDS, and feeds the result into a per-position Z lifting
// pipeline held in registers (5 values per position, the steps of
// cpu/src/transform.cpp:136-145 in streaming order). Forward lifting is
// bit-neutral to tiling, so every coefficient equals the reference's. Outputs of
// each z pair are staged in LDS and written as 128-byte row segments: the LLL
// band to a compact next-level box, the seven other bands to the coefficient
// volume at their split-band positions.
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

namespace wz {
namespace {

constexpr int TX = 64, TY = 32;            // output tile
constexpr int FW = TX + 8, FH = TY + 8;    // reflected footprint
constexpr int PA = FW + 1;                 // LDS pitch (odd: spreads banks)
constexpr int THREADS = 320;
constexpr int SLOTS = FW * FH / THREADS;   // 9 footprint samples per thread
static_assert(FW * FH == SLOTS * THREADS, "footprint must tile the workgroup");

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

template <class T, int P, bool Folded>
WZ_D void inv_lift(T (&e)[P], T (&o)[P]) {
    using C = Lift<T>;
    const T first_odd = o[0], last_odd = o[P - 1];
#pragma unroll
    for (int j = 0; j < P; ++j) {
        e[j] = e[j] * C::ik;
        o[j] = o[j] * (-C::k);
    }
#pragma unroll
    for (int j = 1; j < P; ++j) {
        T s;
        if (Folded && j == 1) s = fmadd(first_odd, -C::k, o[1]);
        else if (Folded && j == P - 1) s = fmadd(last_odd, -C::k, o[P - 2]);
        else s = o[j - 1] + o[j];
        e[j] = fmadd(-C::d, s, e[j]);
    }
#pragma unroll
    for (int j = 1; j < P - 1; ++j) o[j] = fmadd(-C::g, e[j] + e[j + 1], o[j]);
#pragma unroll
    for (int j = 2; j < P - 1; ++j) e[j] = fmadd(-C::b, o[j - 1] + o[j], e[j]);
#pragma unroll
    for (int j = 2; j < P - 2; ++j) o[j] = fmadd(-C::a, e[j] + e[j + 1], o[j]);
}

}  // namespace

template <class T>
__global__ void __launch_bounds__(THREADS) fwd_level_kernel(FwdLevel<T> a) {
    __shared__ T A[FH * PA];            // input plane footprint
    __shared__ T B[TY * PA];            // after Y: rows 0..15 low pairs, 16..31 high pairs
    __shared__ T Cst[2 * TY * (TX + 1)]; // outputs of one z pair: [zband][row][x-low 32 | x-high 32]
    const uint32_t t = threadIdx.x;
    Tile tl;
    if (!tile_of(tile_grid(a.n, a.sz), tl)) return;
    const uint32_t x0 = tl.bx * TX, y0 = tl.by * TY, z0 = tl.bz * a.sz;
    const Dims n = a.n;
    const uint32_t hx = half_len(n.nx), hy = half_len(n.ny), hz = half_len(n.nz);

    uint32_t loff[SLOTS];
#pragma unroll
    for (int s = 0; s < SLOTS; ++s) {
        const uint32_t idx = t + THREADS * s, r = idx / FW, c = idx % FW;
        loff[s] = reflect_pos(int32_t(y0) - 4 + int32_t(r), n.ny) * uint32_t(a.spy) +
                  reflect_pos(int32_t(x0) - 4 + int32_t(c), n.nx);
    }
    // X-pass / Z-pipeline ownership (threads < 256): row r, x tile u (4 pairs)
    const uint32_t r = t & 31u, u = t >> 5;
    T E[8], O[8], O1[8], E2[8], O3[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) E[k] = O[k] = O1[k] = E2[k] = O3[k] = T(0);

    const int npairs = int(a.sz / 2) + 4;
    // software pipeline, two planes deep: while plane p is lifted, the
    // footprints of planes p+1 and p+2 are in flight (even planes in `pe`,
    // odd planes in `po`)
    T pe[SLOTS], po[SLOTS];
    {
        const T* plane0 = a.src + uint64_t(reflect_pos(int32_t(z0) - 4, n.nz)) * a.spz;
        const T* plane1 = a.src + uint64_t(reflect_pos(int32_t(z0) - 3, n.nz)) * a.spz;
#pragma unroll
        for (int s = 0; s < SLOTS; ++s) {
            pe[s] = plane0[loff[s]];
            po[s] = plane1[loff[s]];
        }
    }
    for (int i = 0; i < npairs; ++i) {
        T ev[8], ov[8];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            T cur[SLOTS];
            T (&buf)[SLOTS] = h ? po : pe;
#pragma unroll
            for (int s = 0; s < SLOTS; ++s) cur[s] = buf[s];
            if (i < npairs - 1) {  // plane 2(i+1)+h of the segment
                const int32_t pn = int32_t(z0) - 4 + 2 * (i + 1) + h;
                const T* plane = a.src + uint64_t(reflect_pos(pn, n.nz)) * a.spz;
#pragma unroll
                for (int s = 0; s < SLOTS; ++s) buf[s] = plane[loff[s]];
            }
            bool ok = true;
#pragma unroll
            for (int s = 0; s < SLOTS; ++s) {
                const uint32_t idx = t + THREADS * s;
                ok = ok && __builtin_isfinite(cur[s]);
                A[(idx / FW) * PA + idx % FW] = cur[s];
            }
            if (a.status && !ok) atomicOr(a.status, uint32_t(kNonFiniteInput));
            __syncthreads();
            if (t < 4 * FW) {  // Y: column c, rows 8s..8s+15 -> 4 low + 4 high pairs
                const uint32_t c = t % FW, s = t / FW;
                T e[8], o[8];
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    e[j] = A[(8 * s + 2 * j) * PA + c];
                    o[j] = A[(8 * s + 2 * j + 1) * PA + c];
                }
                fwd_lift<T, 8>(e, o);
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    B[(4 * s + j) * PA + c] = e[j + 2];
                    B[(16 + 4 * s + j) * PA + c] = o[j + 2];
                }
            }
            __syncthreads();
            if (t < 256) {  // X: row r, columns 8u..8u+15 -> 4 low + 4 high pairs
                T e[8], o[8];
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    e[j] = B[r * PA + 8 * u + 2 * j];
                    o[j] = B[r * PA + 8 * u + 2 * j + 1];
                }
                fwd_lift<T, 8>(e, o);
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    (h ? ov : ev)[q] = e[q + 2];
                    (h ? ov : ev)[4 + q] = o[q + 2];
                }
            }
        }
        // Z pipeline: pair i of the segment (cpu/src/transform.cpp:136-145, streamed)
        if (t < 256) {
            using C = Lift<T>;
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const T o1 = fmadd(C::a, E[k] + ev[k], O[k]);   // o1[i-1]
                const T e2 = fmadd(C::b, O1[k] + o1, E[k]);     // e2[i-1]
                const T o3 = fmadd(C::g, E2[k] + e2, O1[k]);    // o3[i-2]
                const T e4 = fmadd(C::d, O3[k] + o3, E2[k]);    // e4[i-2]
                if (i >= 4) {
                    const uint32_t col = (k < 4 ? 0u : 32u) + 4 * u + (k & 3);
                    Cst[r * (TX + 1) + col] = e4 * C::k;
                    Cst[(TY + r) * (TX + 1) + col] = o3 * (-C::ik);
                }
                E[k] = ev[k];
                O[k] = ov[k];
                O1[k] = o1;
                E2[k] = e2;
                O3[k] = o3;
            }
        }
        if (i >= 4) {
            __syncthreads();
            const uint32_t iz = z0 / 2 + uint32_t(i - 4);
            const uint32_t lane = t & 63u, xb = lane >> 5, ix = x0 / 2 + (lane & 31u);
            const bool xok = xb ? 2 * ix + 1 < n.nx : 2 * ix < n.nx;
            for (uint32_t job = t >> 6; job < 2 * TY; job += THREADS / 64) {
                const uint32_t zb = job / TY, rr = job % TY, yb = rr >> 4, iy = y0 / 2 + (rr & 15u);
                const bool ok = xok && (yb ? 2 * iy + 1 < n.ny : 2 * iy < n.ny) && (zb ? 2 * iz + 1 < n.nz : 2 * iz < n.nz);
                if (!ok) continue;
                const T v = Cst[job * (TX + 1) + lane];
                if (!(xb | yb | zb)) {
                    a.lo[uint64_t(iz) * a.lpz + uint64_t(iy) * a.lpy + ix] = v;
                } else {
                    const uint64_t X = xb ? hx + ix : ix, Y = yb ? hy + iy : iy, Z = zb ? hz + iz : iz;
                    a.hi[Z * a.hpz + Y * a.hpy + X] = v;
                }
            }
        }
    }
}

template <class T, int Emit>
__global__ void __launch_bounds__(THREADS) inv_level_kernel(InvLevel<T> a) {
    constexpr int H = sizeof(T) == 4 ? 8 : 4;   // canonical Z tile / 2 (pairs per tile)
    __shared__ T Zp[2 * FH * PA];                // two reconstructed z planes (split XY)
    __shared__ T Xp[2 * FH * (TX + 1)];          // after X: rows still split in Y
    __shared__ uint32_t seg_fill;                // repair entries of this workgroup
    using C = Lift<T>;
    const uint32_t t = threadIdx.x;
    Tile tl;
    if (!tile_of(tile_grid(a.n, a.sz), tl)) return;
    const uint32_t x0 = tl.bx * TX, y0 = tl.by * TY, zs = tl.bz * a.sz;
    const uint32_t wg = tl.id;
    const Dims n = a.n;
    const uint32_t hx = half_len(n.nx), hy = half_len(n.ny), hz = half_len(n.nz);
    const bool rowwords = (n.nx & 31u) == 0;  // bitmap words never straddle rows
    if (t == 0) seg_fill = 0;

    // per footprint slot: offsets of its (cx, cy) in the coefficient volume and
    // in the coarse LLL box (used when cx < hx, cy < hy and cz < hz)
    uint32_t ohi[SLOTS], olo[SLOTS];
    uint32_t lowmask = 0;
#pragma unroll
    for (int s = 0; s < SLOTS; ++s) {
        const uint32_t idx = t + THREADS * s, r = idx / FW, c = idx % FW;
        const uint32_t cy = split_pos(int32_t(y0) - 4 + int32_t(r), n.ny);
        const uint32_t cx = split_pos(int32_t(x0) - 4 + int32_t(c), n.nx);
        ohi[s] = cy * uint32_t(a.hpy) + cx;
        olo[s] = cy * uint32_t(a.lpy) + cx;
        if (cx < hx && cy < hy) lowmask |= 1u << s;
    }
    // Issue the 2 x SLOTS coefficient loads of pair g (even and odd z sample).
    auto issue = [&](int32_t g, T (&ev)[SLOTS], T (&ov)[SLOTS]) {
        const uint32_t ze = split_pos(2 * g, n.nz), zo = split_pos(2 * g + 1, n.nz);
        const T* he = a.hi + uint64_t(ze) * a.hpz;
        const T* ho = a.hi + uint64_t(zo) * a.hpz;
        const T* le = a.lo + uint64_t(ze) * a.lpz;
        const T* lo = a.lo + uint64_t(zo) * a.lpz;
        const uint32_t le_ok = ze < hz ? lowmask : 0u, lo_ok = zo < hz ? lowmask : 0u;
#pragma unroll
        for (int s = 0; s < SLOTS; ++s) {
            const T* pe = ((le_ok >> s) & 1u) ? le + olo[s] : he + ohi[s];
            const T* po = ((lo_ok >> s) & 1u) ? lo + olo[s] : ho + ohi[s];
            ev[s] = *pe;
            ov[s] = *po;
        }
    };
    // streaming state per slot: OS = os[g-1], EA = eA[g-1], OB = oB[g-2],
    // EC = eC[g-2]; SF rotates through o_raw[B-2], eA'[B-1], oB'[B-1], eC'[B]
    T OS[SLOTS], EA[SLOTS], OB[SLOTS], EC[SLOTS], SF[SLOTS];
#pragma unroll
    for (int s = 0; s < SLOTS; ++s) OS[s] = EA[s] = OB[s] = EC[s] = SF[s] = T(0);

    const int32_t g0 = int32_t(zs / 2) - 2, g1 = int32_t((zs + a.sz) / 2) + 1;
    T nev[SLOTS], nov[SLOTS];
    issue(g0, nev, nov);
    for (int32_t g = g0; g <= g1; ++g) {
        T cev[SLOTS], cov[SLOTS];
#pragma unroll
        for (int s = 0; s < SLOTS; ++s) {
            cev[s] = nev[s];
            cov[s] = nov[s];
        }
        if (g < g1) issue(g + 1, nev, nov);  // in flight during this step
        const int phase = ((g % H) + H) % H;
        const bool emit = g - 2 >= int32_t(zs / 2);
#pragma unroll
        for (int s = 0; s < SLOTS; ++s) {
            const T oraw = cov[s];
            const T es = cev[s] * C::ik, os = oraw * (-C::k);
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
                const uint32_t idx = t + THREADS * s, row = idx / FW, col = idx % FW;
                Zp[row * PA + col] = even;
                Zp[FH * PA + row * PA + col] = odd;
            }
            OS[s] = os;
            EA[s] = eA;
            OB[s] = oB;
            EC[s] = eC;
        }
        if (!emit) continue;
        __syncthreads();
        // X (unfolded, tile 8): plane p, row rr, x tile xt -> 8 samples
        for (uint32_t w = t; w < 2 * FH * (TX / 8); w += THREADS) {
            const uint32_t p = w / (FH * (TX / 8)), rem = w % (FH * (TX / 8));
            const uint32_t rr = rem % FH, xt = rem / FH;
            const T* src = Zp + p * FH * PA + rr * PA + 8 * xt;
            T e[8], o[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                e[j] = src[2 * j];
                o[j] = src[2 * j + 1];
            }
            inv_lift<T, 8, false>(e, o);
            T* dst = Xp + p * FH * (TX + 1) + rr * (TX + 1) + 8 * xt;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                dst[2 * j] = e[j + 2];
                dst[2 * j + 1] = o[j + 2];
            }
        }
        __syncthreads();
        // Y (folded, tile 8): plane p, column x, y tile yt -> 8 samples. A wave
        // is one (p, yt) and its lanes are 64 consecutive x (64 | 256, 64 | 320).
        const uint32_t zbase = 2 * uint32_t(g - 2);
        for (uint32_t w = t; w < 2 * TX * (TY / 8); w += THREADS) {
            const uint32_t p = w / (TX * (TY / 8)), rem = w % (TX * (TY / 8));
            const uint32_t x = rem % TX, yt = rem / TX;
            const T* src = Xp + p * FH * (TX + 1) + 8 * yt * (TX + 1) + x;
            const uint32_t X = x0 + x, Z = zbase + p;
            const bool xz = X < n.nx && Z < n.nz;
            // the original samples of this column (repair emit): all eight
            // loads issued before the lifting, none waits behind a store
            // (float; for double the registers are not there and each load
            // is issued where it is used)
            constexpr bool hoist = Emit == kEmitRepair && sizeof(T) == 4;
            T ref[8];
            if constexpr (hoist) {
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const uint32_t Y = y0 + 8 * yt + j;
                    ref[j] = xz && Y < n.ny ? a.orig[uint64_t(Z) * a.opz + uint64_t(Y) * a.opy + X] : T(0);
                }
            }
            T e[8], o[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                e[j] = src[(2 * j) * (TX + 1)];
                o[j] = src[(2 * j + 1) * (TX + 1)];
            }
            inv_lift<T, 8, true>(e, o);
            if constexpr (Emit != kEmitRepair) {
                bool fin = true;
                if (xz) {
                    T* dst = a.out + uint64_t(Z) * a.opz + X;
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const uint32_t Y = y0 + 8 * yt + 2 * j;
                        if (Y < n.ny) {
                            dst[uint64_t(Y) * a.opy] = e[j + 2];
                            fin = fin && __builtin_isfinite(e[j + 2]);
                        }
                        if (Y + 1 < n.ny) {
                            dst[uint64_t(Y + 1) * a.opy] = o[j + 2];
                            fin = fin && __builtin_isfinite(o[j + 2]);
                        }
                    }
                }
                if (Emit == kEmitStoreScreen && __ballot(!fin) && (t & 63u) == 0)
                    atomicOr(a.status, uint32_t(kNonFiniteOutput));
            } else {
                // decide the column's 8 samples, then reserve all of the
                // wave's records with one LDS atomic and write them
                const uint32_t lane = t & 63u;
                const unsigned long long below = (1ull << lane) - 1ull;
                const T nb = T(a.bound);
                bool ok = true;
                T err[8];
                uint32_t pbits = 0, before[8], total = 0;
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const uint32_t Y = y0 + 8 * yt + j;
                    const T v = (j & 1) ? o[j / 2 + 2] : e[j / 2 + 2];
                    bool patch = false;
                    err[j] = T(0);
                    if (xz && Y < n.ny) {
                        const T r = hoist ? ref[j] : a.orig[uint64_t(Z) * a.opz + uint64_t(Y) * a.opy + X];
                        patch = repair_decision<T>(v, r, nb, a.bound, &err[j], &ok);
                    }
                    const unsigned long long m = __ballot(patch);
                    pbits |= uint32_t(patch) << j;
                    before[j] = total + uint32_t(__popcll(m & below));
                    total += uint32_t(__popcll(m));
                    if (rowwords) {  // 64 consecutive x at a 32-aligned row start: own two words
                        if (lane == 0 && Z < n.nz && Y < n.ny) {
                            const uint64_t w0 = ((uint64_t(Z) * n.ny + Y) * n.nx + x0) >> 5;
                            a.rep.bitmap[w0] = uint32_t(m);
                            if (x0 + 32 < n.nx) a.rep.bitmap[w0 + 1] = uint32_t(m >> 32);
                        }
                    } else if (patch) {
                        const uint64_t at = (uint64_t(Z) * n.ny + Y) * n.nx + X;
                        atomicOr(&a.rep.bitmap[at >> 5], 1u << (at & 31u));
                    }
                }
                if (total) {
                    uint32_t base = 0;
                    if (lane == 0) base = atomicAdd(&seg_fill, total);
                    base = __shfl(base, 0, 64);
                    const uint64_t seg0 = uint64_t(wg) * a.rep.cap + base;
#pragma unroll
                    for (int j = 0; j < 8; ++j) {
                        if ((pbits >> j) & 1u) {
                            const uint32_t Y = y0 + 8 * yt + j;
                            a.rep.index[seg0 + before[j]] = uint32_t((uint64_t(Z) * n.ny + Y) * n.nx + X);
                            static_cast<T*>(a.rep.payload)[seg0 + before[j]] = err[j];
                        }
                    }
                }
                if (__ballot(!ok) && lane == 0) atomicOr(a.status, uint32_t(kRepairBound));
            }
        }
    }
    if constexpr (Emit == kEmitRepair) {
        __syncthreads();
        if (t == 0) a.rep.count[wg] = seg_fill;
    }
}

// ------------------------------------------------------------------ launchers
template <class T>
void launch_fwd_level(const FwdLevel<T>& a, hipStream_t s) {
    const dim3 grid(xcd_blocks(tile_grid(a.n, a.sz)));
    hipLaunchKernelGGL((fwd_level_kernel<T>), grid, dim3(THREADS), 0, s, a);
}
template <class T>
void launch_inv_level(const InvLevel<T>& a, hipStream_t s) {
    const dim3 grid(xcd_blocks(tile_grid(a.n, a.sz)));
    if (a.emit == kEmitRepair)
        hipLaunchKernelGGL((inv_level_kernel<T, kEmitRepair>), grid, dim3(THREADS), 0, s, a);
    else if (a.emit == kEmitStoreScreen)
        hipLaunchKernelGGL((inv_level_kernel<T, kEmitStoreScreen>), grid, dim3(THREADS), 0, s, a);
    else
        hipLaunchKernelGGL((inv_level_kernel<T, kEmitStore>), grid, dim3(THREADS), 0, s, a);
}

template void launch_fwd_level<float>(const FwdLevel<float>&, hipStream_t);
template void launch_fwd_level<double>(const FwdLevel<double>&, hipStream_t);
template void launch_inv_level<float>(const InvLevel<float>&, hipStream_t);
template void launch_inv_level<double>(const InvLevel<double>&, hipStream_t);

}  // namespace wz
