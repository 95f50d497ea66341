// Batched decode: many canonical archives, of one tile shape or of a mix of
// shapes, decoded with one launch per pipeline stage for the whole batch.
//
// A plan is built on the host from the tile shapes alone: per job the level
// schedule, the WZP chunk and Morton brick counts, offsets into three work
// banks (coefficients, and two scratch banks that also hold the decoded
// magnitudes and sign words), and the output placement (offset, pitches and
// clipped extents). Each stage then runs as one launch whose blocks are dealt
// to jobs through block prefix tables:
//   plan     one workgroup parses every 32-byte header on the device, checks
//            it against the job's shape and span, and scans the side-channel
//            counts over jobs;
//   wzp      one workgroup per WZP chunk of every job (the single-field
//            chunk decoder, writing into the job's magnitude/sign banks);
//   dequant  one workgroup per Morton brick (rank computed on the fly);
//   qpatch   q-outliers of all jobs, grid-stride;
//   passes   the inverse lifting schedule, one launch per schedule step, each
//            job running its own axis, box and tile size at that step; the
//            job's last pass writes straight into the output with clipping
//            (holes are filled there);
//   rpatch   native repairs of all jobs into the output.
// No host synchronisation, allocation or memset happens between the stages,
// so a batch can be captured in a hipGraph and replayed. Every value equals
// the single-field decoder's (same kernels' arithmetic, same tile sizes).
//
// Entry points: the reference's waltz_gpu_decode_batch (one tile shape,
// frames of tiles; src/batch_decode.cuh:29-186) and waltz_amd_batch_* (any
// mix of shapes, tiles concatenated in job order).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/waltz_amd.h"
#include "../../include/waltz_gpu_batch.h"
#include "common.hpp"
#include "launch.hpp"
#include "bricks.hpp"
#include "lifting.hpp"
#include "sections.hpp"
#include "wzp_encode.hpp"

namespace wz {
namespace {

enum PassKind : uint8_t { kPassX = 0, kPassY = 1, kPassZ = 2, kPassCopy = 3, kPassNone = 255 };
enum BankId : uint8_t { kHome = 0, kTmp0 = 1, kTmp1 = 2, kOutput = 3 };
enum JobState : uint32_t { kJobOk = 0, kJobHole = 1, kJobBad = 2 };
// (De)quantization of a job: separate brick passes around the WZP coders, or
// inside them for full 16^3 brick pairs (pair_bricks_fit) and for the other
// fields of full bricks (rank_bricks_fit).
enum JobCoder : uint32_t { kCoderSeparate = 0, kCoderPairs = 1, kCoderRanks = 2 };

struct JobShape {
    Dims d;
    uint32_t lx, lz, bz, dyadic;
    uint32_t n, chunks, groups, nbx, nby, nbz;
    uint32_t rank;      // the job's Morton rank table set (one per distinct geometry)
    uint32_t coder;     // JobCoder: where (de)quantization runs
    uint64_t bank;      // element offset in each work bank
    uint64_t mag;       // u16 offset of the magnitudes (in the tmp0 bank)
    uint64_t sign;      // word offset of the sign words (in the tmp1 bank)
    uint64_t out;       // element offset of the tile origin in the output
    uint64_t opy, opz;  // output pitches (elements)
    Dims clip;          // tile extents that lie inside the output
};

struct JobRun {
    uint64_t arc;       // byte offset of the archive in the packed buffer
    uint64_t wzp;       // WZP stream bytes
    double quantum;
    uint32_t np, nq;
    uint64_t pi, pv, qi, qv;  // section offsets from the archive start
    uint32_t state, pad;
};

struct Pass {
    uint8_t kind, src, dst, pad;
    Dims act;           // box of this pass
    uint32_t ntiles;    // tiles per line
    uint32_t work;      // work items (threads)
};

struct Banks {
    void* home;
    void* tmp0;
    void* tmp1;
    void* out;
};

// Largest j with pfx[j] <= b (pfx[jobs] > b): the job owning item b.
WZ_D uint32_t job_of(const uint32_t* __restrict__ pfx, uint32_t jobs, uint32_t b) {
    uint32_t lo = 0, hi = jobs;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (pfx[mid] <= b) lo = mid;
        else hi = mid;
    }
    return lo;
}

// The same for a block index the whole workgroup shares (every lane of each
// wave calls it): two rounds of 64 parallel reads and a ballot instead of a
// chain of dependent reads, for up to 4096 jobs.
WZ_D uint32_t job_of_wave(const uint32_t* __restrict__ pfx, uint32_t jobs, uint32_t b) {
    if (jobs <= 1) return 0;
    if (jobs > 4096) return job_of(pfx, jobs, b);
    const uint32_t lane = __lane_id();
    const uint32_t step = (jobs + 63) / 64;
    const uint32_t p1 = lane * step;  // the segment [p1, p1 + step) of lane l
    const unsigned long long m1 = __ballot(p1 < jobs && pfx[p1 < jobs ? p1 : 0u] <= b);
    const uint32_t seg = 63u - uint32_t(__builtin_clzll(m1));  // lane 0 always holds (pfx[0] = 0)
    const uint32_t p2 = seg * step + lane;
    const unsigned long long m2 = __ballot(lane < step && p2 < jobs && pfx[p2 < jobs ? p2 : 0u] <= b);
    return seg * step + (63u - uint32_t(__builtin_clzll(m2)));
}

WZ_D uint32_t rd32(const uint8_t* p) {
    return uint32_t(p[0]) | (uint32_t(p[1]) << 8) | (uint32_t(p[2]) << 16) | (uint32_t(p[3]) << 24);
}

// Header checks of archive_size_from_header plus the job's own geometry
// (include/utils/Blob.hpp:71-90, cpu/src/codec.cpp:276-296).
WZ_D bool parse_header(const uint8_t* h, uint64_t size, const JobShape& sh, uint32_t width, JobRun& r) {
    if (size < 32) return false;
    const uint32_t f = h[0];
    bool ok = !(f & 0xE0u) && (f & 0x18u) && (((f & 4u) ? 8u : 4u) == width) && ((f >> 3) & 3u) == 3u && !(f & 2u);
    ok = ok && h[1] == sh.lx && h[2] == sh.lz && h[3] == sh.bz && (f & 1u) == sh.dyadic;
    const uint64_t qb = uint64_t(rd32(h + 8)) | (uint64_t(rd32(h + 12)) << 32);
    const double q = __longlong_as_double((long long)qb);
    ok = ok && q > 0 && __builtin_isfinite(q);
    if (width == 4) {
        const float qf = float(q);
        ok = ok && qf > 0.0f && __builtin_isfinite(qf);
    }
    const uint32_t pb = rd32(h + 4), np = rd32(h + 24), nq = rd32(h + 28);
    const uint64_t wzp = uint64_t(rd32(h + 16)) | (uint64_t(rd32(h + 20)) << 32);
    const WzpShape ws = wzp_shape(sh.n);
    ok = ok && np <= sh.n && nq <= sh.n && uint64_t(pb) == uint64_t(np) * width;
    ok = ok && wzp >= ws.header && wzp <= ws.header + uint64_t(sh.n) * 2 + (uint64_t(sh.n) + 7) / 8;
    if (!ok) return false;
    const Layout L = archive_layout(wzp, np, uint64_t(np) * width, nq);
    if (L.end != size) return false;
    r.wzp = wzp;
    r.quantum = q;
    r.np = np;
    r.nq = nq;
    r.pi = L.pi;
    r.pv = L.pv;
    r.qi = L.qi;
    r.qv = L.qv;
    return true;
}

// One workgroup of 1024 threads: per-job headers, then exclusive scans of
// the repair and q-outlier counts over decodable jobs.
__global__ void __launch_bounds__(1024) batch_plan_kernel(const uint8_t* __restrict__ packed, uint64_t packed_bytes,
                                                          const uint64_t* __restrict__ spans,
                                                          const JobShape* __restrict__ shape, JobRun* run,
                                                          uint32_t jobs, uint32_t width, uint32_t* np_pfx,
                                                          uint32_t* nq_pfx, uint32_t* status) {
    __shared__ uint32_t wsum[2][16];
    __shared__ uint32_t carry[2];
    const uint32_t t = threadIdx.x;
    for (uint32_t j = t; j < jobs; j += 1024) {
        JobRun r{};
        r.arc = spans[2 * j];
        const uint64_t size = spans[2 * j + 1];
        if (size == 0) {
            r.state = kJobHole;
        } else if ((r.arc & 7u) || r.arc > packed_bytes || size > packed_bytes - r.arc ||
                   !parse_header(packed + r.arc, size, shape[j], width, r)) {
            r.state = kJobBad;
            atomicOr(status, uint32_t(kCorrupt));
        }
        run[j] = r;
    }
    if (t < 2) carry[t] = 0;
    __syncthreads();
    const uint32_t lane = t & 63u, wv = t >> 6;
    for (uint32_t base = 0; base <= jobs; base += 1024) {
        const uint32_t j = base + t;
        uint32_t vp = 0, vq = 0;
        if (j < jobs && run[j].state == kJobOk) {
            vp = run[j].np;
            vq = run[j].nq;
        }
        uint32_t ip = vp, iq = vq;
#pragma unroll
        for (uint32_t d = 1; d < 64; d <<= 1) {
            const uint32_t a = __shfl_up(ip, d, 64), b = __shfl_up(iq, d, 64);
            if (lane >= d) {
                ip += a;
                iq += b;
            }
        }
        if (lane == 63) {
            wsum[0][wv] = ip;
            wsum[1][wv] = iq;
        }
        __syncthreads();
        uint32_t bp = carry[0], bq = carry[1], tp = 0, tq = 0;
        for (uint32_t k = 0; k < 16; ++k) {
            bp += k < wv ? wsum[0][k] : 0u;
            bq += k < wv ? wsum[1][k] : 0u;
            tp += wsum[0][k];
            tq += wsum[1][k];
        }
        if (j <= jobs) {
            np_pfx[j] = bp + ip - vp;
            nq_pfx[j] = bq + iq - vq;
        }
        __syncthreads();
        if (t == 0) {
            carry[0] += tp;
            carry[1] += tq;
        }
        __syncthreads();
    }
}

// Chunks of the jobs that are not brick pairs (mag_pfx counts only theirs):
// magnitudes and signs, dequantized by batch_dequant_kernel.
__global__ void __launch_bounds__(256) batch_wzp_kernel(const uint8_t* __restrict__ packed,
                                                        const JobShape* __restrict__ shape,
                                                        const JobRun* __restrict__ run,
                                                        const uint32_t* __restrict__ mag_pfx, uint32_t jobs,
                                                        uint16_t* mag, uint32_t* sign, uint32_t* status) {
    const uint32_t j = job_of_wave(mag_pfx, jobs, blockIdx.x);
    const JobRun& r = run[j];
    if (r.state != kJobOk) return;
    const JobShape& sh = shape[j];
    decode_chunk(packed + r.arc + 32, r.wzp, sh.n, sh.chunks, sh.groups, blockIdx.x - mag_pfx[j],
                 MagSink{mag + sh.mag, sign + sh.sign}, status);
}

// Chunks of the brick-pair jobs (pair_pfx): decoded straight into dequantized
// coefficients, as the single-field decoder's BrickSink does.
template <class T>
__global__ void __launch_bounds__(256) batch_wzp_pairs_kernel(const uint8_t* __restrict__ packed,
                                                              const JobShape* __restrict__ shape,
                                                              const JobRun* __restrict__ run,
                                                              const uint32_t* __restrict__ pair_pfx, uint32_t jobs,
                                                              T* __restrict__ home, uint32_t* status) {
    const uint32_t j = job_of_wave(pair_pfx, jobs, blockIdx.x);
    const JobRun& r = run[j];
    if (r.state != kJobOk) return;
    const JobShape& sh = shape[j];
    decode_chunk(packed + r.arc + 32, r.wzp, sh.n, sh.chunks, sh.groups, blockIdx.x - pair_pfx[j],
                 BrickSink<T>{home + sh.bank, sh.d.nx, sh.d.ny, T(r.quantum)}, status);
}

// Rank table sets of a plan: per distinct (geometry, z band) the packed rank
// of every local index of the eight brick edge shapes (quantize.hip's tables;
// the full brick's first).
constexpr uint64_t kRankSet = 2ull * 8 * 8192;  // u16 entries per set (rank + inverse)

// Chunks of the other full-brick jobs (rank_pfx): decoded straight into
// dequantized coefficients, as the single-field decoder's RankSink does.
template <class T>
__global__ void __launch_bounds__(256) batch_wzp_ranks_kernel(const uint8_t* __restrict__ packed,
                                                              const JobShape* __restrict__ shape,
                                                              const JobRun* __restrict__ run,
                                                              const uint32_t* __restrict__ rank_pfx, uint32_t jobs,
                                                              const uint16_t* __restrict__ ranks,
                                                              T* __restrict__ home, uint32_t* status) {
    const uint32_t j = job_of_wave(rank_pfx, jobs, blockIdx.x);
    const JobRun& r = run[j];
    if (r.state != kJobOk) return;
    const JobShape& sh = shape[j];
    const RankBricks g{ranks + uint64_t(sh.rank) * kRankSet, sh.d.nx, sh.d.ny, sh.bz};
    decode_chunk(packed + r.arc + 32, r.wzp, sh.n, sh.chunks, sh.groups, blockIdx.x - rank_pfx[j],
                 RankSink<T>{home + sh.bank, g, T(r.quantum)}, status);
}
// Bricks per workgroup of the separate (de)quantization passes: the bricks
// of small z bands hold a few hundred samples, and a workgroup per brick
// spent most of its life finding its job and brick
constexpr uint32_t kBricksPerBlock = 8;

struct SetRank {
    const uint16_t* set;
    WZ_D uint32_t operator()(const Brick& k, uint32_t i, uint32_t, uint32_t, uint32_t) const {
        return set[k.shape * 8192 + i];
    }
};

// Dequantization of one brick (quantize.hip's brick order, packed base and ranks).
// Dequantization of up to kBricksPerBlock consecutive bricks of one job
// (quantize.hip's brick order, packed base and ranks), their samples taken
// as one range so that every burst is full.
template <class T>
__global__ void __launch_bounds__(256) batch_dequant_kernel(const JobShape* __restrict__ shape,
                                                            const JobRun* __restrict__ run,
                                                            const uint32_t* __restrict__ brick_pfx, uint32_t jobs,
                                                            const uint16_t* __restrict__ ranks,
                                                            const uint16_t* __restrict__ mag,
                                                            const uint32_t* __restrict__ sign, T* __restrict__ home) {
    __shared__ Brick kb[kBricksPerBlock];
    __shared__ uint32_t kstart[kBricksPerBlock + 1];  // sample offsets of the bricks in the range
    const uint32_t j = job_of_wave(brick_pfx, jobs, blockIdx.x);
    const JobRun& r = run[j];
    if (r.state != kJobOk) return;
    const JobShape& sh = shape[j];
    const BrickGeometry g{sh.d, sh.bz, sh.nbx, sh.nby, sh.nbz};
    const uint32_t b0 = (blockIdx.x - brick_pfx[j]) * kBricksPerBlock;
    const uint32_t cnt = min(kBricksPerBlock, sh.nbx * sh.nby * sh.nbz - b0);
    if (threadIdx.x < cnt) kb[threadIdx.x] = brick_of(g, b0 + threadIdx.x);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t at = 0;
        for (uint32_t b = 0; b < cnt; ++b) {
            kstart[b] = at;
            at += kb[b].bx * kb[b].by * kb[b].bz;
        }
        for (uint32_t b = cnt; b <= kBricksPerBlock; ++b) kstart[b] = at;
    }
    __syncthreads();
    uint32_t ks[kBricksPerBlock + 1];
#pragma unroll
    for (uint32_t b = 0; b <= kBricksPerBlock; ++b) ks[b] = kstart[b];
    const uint32_t total = ks[kBricksPerBlock];
    const T q = T(r.quantum);
    const uint16_t* m = mag + sh.mag;
    const uint32_t* sg = sign + sh.sign;
    const uint16_t* rset = ranks + uint64_t(sh.rank) * kRankSet;
    T* c = home + sh.bank;
    constexpr uint32_t kBurst = 8;  // all loads of a burst issued before its stores
    for (uint32_t i0 = 0; i0 < total; i0 += kBurst * 256u) {
        uint32_t mv[kBurst], sv[kBurst], pv[kBurst], bi[kBurst], li[kBurst];
#pragma unroll
        for (uint32_t u = 0; u < kBurst; ++u) {
            const uint32_t s = min(i0 + u * 256u + threadIdx.x, total - 1);
            uint32_t b = 0;  // the brick holding sample s of the range
#pragma unroll
            for (uint32_t e = 1; e < kBricksPerBlock; ++e) b += s >= ks[e] ? 1u : 0u;
            bi[u] = b;
            li[u] = s - ks[b];
            const Brick& k = kb[b];
            const uint64_t p = k.base + rset[k.shape * 8192 + li[u]];
            pv[u] = uint32_t(p & 31u);
            mv[u] = m[p];
            sv[u] = sg[p >> 5];
        }
#pragma unroll
        for (uint32_t u = 0; u < kBurst; ++u) {
            if (i0 + u * 256u + threadIdx.x >= total) continue;
            const Brick& k = kb[bi[u]];
            const uint32_t i = li[u];
            uint32_t x, y, z;
            brick_xyz(k, i, x, y, z);
            const T v = T(mv[u]) * q;  // cpu/src/codec.cpp:146-147
            c[(uint64_t(k.z0 + z) * sh.d.ny + (k.y0 + y)) * sh.d.nx + k.x0 + x] = ((sv[u] >> pv[u]) & 1u) ? -v : v;
        }
    }
}

template <class T>
__global__ void __launch_bounds__(256) batch_qpatch_kernel(const uint8_t* __restrict__ packed,
                                                           const JobShape* __restrict__ shape,
                                                           const JobRun* __restrict__ run,
                                                           const uint32_t* __restrict__ nq_pfx, uint32_t jobs,
                                                           T* __restrict__ home, uint32_t* status) {
    const uint32_t total = nq_pfx[jobs];
    for (uint32_t k = blockIdx.x * 256u + threadIdx.x; k < total; k += gridDim.x * 256u) {
        const uint32_t j = job_of(nq_pfx, jobs, k), e = k - nq_pfx[j];
        const JobRun& r = run[j];
        const JobShape& sh = shape[j];
        const uint8_t* a = packed + r.arc;
        const uint32_t i = reinterpret_cast<const uint32_t*>(a + r.qi)[e];
        const uint32_t v = reinterpret_cast<const uint32_t*>(a + r.qv)[e];
        if (i >= sh.n || v <= 65535u || v > 0x7FFFFFFFu) {
            atomicOr(status, uint32_t(kCorrupt));
            continue;
        }
        T* c = home + sh.bank;
        const T x = T(v) * T(r.quantum);  // cpu/src/codec.cpp:299-302
        c[i] = c[i] < T(0) ? -x : x;
    }
}

// Stores one reconstructed sample: into a work bank, or (last pass) into the
// output with clipping, the hole fill and the non-finite screen.
template <class T>
struct Put {
    T* bank;  // null: output
    T* out;
    const JobShape* sh;
    bool hole;
    T fill;
    WZ_D bool operator()(uint64_t at, uint32_t x, uint32_t y, uint32_t z, T v) const {
        if (bank) {
            bank[at] = v;
            return true;
        }
        if (x >= sh->clip.nx || y >= sh->clip.ny || z >= sh->clip.nz) return true;
        out[sh->out + uint64_t(z) * sh->opz + uint64_t(y) * sh->opy + x] = hole ? fill : v;
        return hole || __builtin_isfinite(v);
    }
};

// One tile of one line of an inverse axis pass (axis_pass of transform.hip
// with the job's tile extents as the volume).
template <class T, int Axis, int Tile>
WZ_D bool inverse_line(const T* __restrict__ src, const Put<T>& put, Dims full, Dims act, uint32_t ntiles,
                       uint32_t i) {
    constexpr int P = Tile / 2 + 4;
    uint32_t x = 0, y = 0, z = 0, t;
    if constexpr (Axis == 0) {
        t = i % ntiles;
        const uint32_t r = i / ntiles;
        y = r % act.ny;
        z = r / act.ny;
    } else if constexpr (Axis == 1) {
        x = i % act.nx;
        const uint32_t r = i / act.nx;
        t = r % ntiles;
        z = r / ntiles;
    } else {
        x = i % act.nx;
        const uint32_t r = i / act.nx;
        y = r % act.ny;
        t = r / act.ny;
    }
    const uint64_t plane = uint64_t(full.nx) * full.ny;
    const uint32_t n = Axis == 0 ? act.nx : Axis == 1 ? act.ny : act.nz;
    const uint64_t stride = Axis == 0 ? 1u : Axis == 1 ? uint64_t(full.nx) : plane;
    const uint64_t base = uint64_t(z) * plane + uint64_t(y) * full.nx + x;
    const int32_t start = int32_t(t) * Tile;
    T e[P], o[P];
#pragma unroll
    for (int j = 0; j < P; ++j) {
        const int32_t p = start - 4 + 2 * j;
        e[j] = src[base + uint64_t(split_pos(p, n)) * stride];
        o[j] = src[base + uint64_t(split_pos(p + 1, n)) * stride];
    }
    lift_inverse<T, P, Axis != 0>(e, o);
    bool fin = true;
#pragma unroll
    for (int j = 2; j < P - 2; ++j) {
        const uint32_t p = uint32_t(start) + 2u * uint32_t(j - 2);
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const uint32_t q = p + h;
            if (q < n) {
                const uint32_t X = Axis == 0 ? q : x, Y = Axis == 1 ? q : y, Z = Axis == 2 ? q : z;
                fin = put(base + uint64_t(q) * stride, X, Y, Z, h ? o[j] : e[j]) && fin;
            }
        }
    }
    return fin;
}

template <class T>
__global__ void __launch_bounds__(256) batch_pass_kernel(const JobShape* __restrict__ shape,
                                                         const JobRun* __restrict__ run,
                                                         const Pass* __restrict__ pass,
                                                         const uint32_t* __restrict__ pfx, uint32_t jobs, Banks bk,
                                                         T fill, uint32_t* status) {
    constexpr int ZTile = sizeof(T) == 4 ? 16 : 8;  // canonical inverse Z tile
    const uint32_t j = job_of_wave(pfx, jobs, blockIdx.x);
    const Pass ps = pass[j];
    const uint32_t i = (blockIdx.x - pfx[j]) * 256u + threadIdx.x;
    const uint32_t state = run[j].state;
    if (ps.kind == kPassNone || state == kJobBad || (state == kJobHole && ps.dst != kOutput)) return;
    const JobShape& sh = shape[j];
    auto bank = [&](uint8_t id) -> T* {
        T* b = static_cast<T*>(id == kHome ? bk.home : id == kTmp0 ? bk.tmp0 : bk.tmp1);
        return b + sh.bank;
    };
    const T* src = bank(ps.src);
    const Put<T> put{ps.dst == kOutput ? nullptr : bank(ps.dst), static_cast<T*>(bk.out), &sh, state == kJobHole,
                     fill};
    bool fin = true;
    if (i < ps.work) {
        switch (ps.kind) {
            case kPassX: fin = inverse_line<T, 0, 8>(src, put, sh.d, ps.act, ps.ntiles, i); break;
            case kPassY: fin = inverse_line<T, 1, 8>(src, put, sh.d, ps.act, ps.ntiles, i); break;
            case kPassZ: fin = inverse_line<T, 2, ZTile>(src, put, sh.d, ps.act, ps.ntiles, i); break;
            default: {
                const uint32_t x = i % ps.act.nx, r = i / ps.act.nx, y = r % ps.act.ny, z = r / ps.act.ny;
                const uint64_t at = (uint64_t(z) * sh.d.ny + y) * sh.d.nx + x;
                fin = put(at, x, y, z, src[at]);
            }
        }
    }
    if (__ballot(!fin) && (threadIdx.x & 63u) == 0) atomicOr(status, uint32_t(kNonFiniteOutput));
}

constexpr uint32_t kRecords = 4;         // consecutive repair records per thread and round
constexpr uint32_t kRpatchBlocks = 2048;  // fixed grid (the record count is known on the device only)

template <class T>
__global__ void __launch_bounds__(256) batch_rpatch_kernel(const uint8_t* __restrict__ packed,
                                                           const JobShape* __restrict__ shape,
                                                           const JobRun* __restrict__ run,
                                                           const uint32_t* __restrict__ np_pfx, uint32_t jobs,
                                                           T* __restrict__ out, uint32_t* status) {
    // each thread takes kRecords consecutive records (of one job or a few)
    // per round: one job search, then a walk to the next job
    const uint32_t total = np_pfx[jobs];
    for (uint32_t k0 = (blockIdx.x * 256u + threadIdx.x) * kRecords; k0 < total;
         k0 += gridDim.x * 256u * kRecords) {
        uint32_t j = job_of(np_pfx, jobs, k0);
        for (uint32_t k = k0; k < min(k0 + kRecords, total); ++k) {
            while (k >= np_pfx[j + 1]) ++j;  // jobs without repairs are skipped
            const uint32_t e = k - np_pfx[j];
            const JobRun& r = run[j];
            const JobShape& sh = shape[j];
            const uint8_t* a = packed + r.arc;
            const uint32_t i = reinterpret_cast<const uint32_t*>(a + r.pi)[e];
            const T err = reinterpret_cast<const T*>(a + r.pv)[e];
            if (i >= sh.n || !__builtin_isfinite(err)) {
                atomicOr(status, uint32_t(kCorrupt));
                continue;
            }
            const uint32_t x = i % sh.d.nx, y = (i / sh.d.nx) % sh.d.ny, z = i / (sh.d.nx * sh.d.ny);
            if (x >= sh.clip.nx || y >= sh.clip.ny || z >= sh.clip.nz) continue;
            T* o = out + sh.out + uint64_t(z) * sh.opz + uint64_t(y) * sh.opy + x;
            const T v = *o - err;  // cpu/src/codec.cpp:330
            *o = v;
            if (!__builtin_isfinite(v)) atomicOr(status, uint32_t(kNonFiniteOutput));
        }
    }
}

// ------------------------------------------------------------------ batched compress
// The single-field compress pipeline (codec.hip) with one launch per stage for
// every job of the batch: forward passes, brick quantization, WZP chunks, WZP
// groups, inverse passes into a reconstruction bank, repair tiles, rank
// prefixes, layout + header per job, placement of both side lists. Per-job
// offsets into the work banks and into the caller's archive buffer come from
// the plan; failures are recorded per job (EncodeResult::status).
struct JobComp {
    uint64_t arc, cap;      // the job's archive slot in the output buffer (bytes)
    uint64_t staging;       // byte offset of its WZP staging slots
    uint64_t qlist;         // entry offset of its q-outlier list
    uint64_t rlist;         // entry offset of its repair segments
    uint64_t words;         // word offset of its flag bitmaps (16-byte aligned)
    uint64_t prefix;        // word offset of its rank scratch
    uint32_t qcap, tiles;   // q-outlier capacity, repair tiles of kRepairTile
    uint32_t chunk0, group0, tile0, pblocks;  // global first chunk / group / tile; prefix blocks
    uint32_t qblocks, pad;  // q-outlier placement blocks
};

// Device pointers of one batched compress call.
struct CompBufs {
    const void* input;      // jobs' fields, concatenated (JobShape::out offsets)
    void* home;             // coefficients
    void* tmp0;             // lifting scratch; then u16 magnitudes
    void* tmp1;             // lifting scratch; then sign words
    void* recon;            // reconstruction (the input's layout)
    uint8_t* staging;
    uint32_t *qbm, *rbm;    // flag bitmaps
    uint32_t *qidx, *qval;  // q-outlier lists
    uint32_t* ridx;         // repair segments: indices
    void* rval;             // repair segments: corrections
    uint32_t *rcount, *qcount, *prefix, *sizes, *ticket;
    unsigned long long* desc;
    EncodeResult* res;
    uint8_t* out;           // archive slots
    uint64_t* spans;        // jobs x {offset, size}
};

// One tile of one line of a forward axis pass (transform.hip's lift_tile,
// forward: reflected halo, band-split outputs).
template <class T, int Axis, int Tile>
WZ_D bool forward_line(const T* __restrict__ src, T* __restrict__ dst, Dims full, Dims act, uint32_t ntiles,
                       uint32_t i, bool screen) {
    constexpr int P = Tile / 2 + 4;
    uint32_t x = 0, y = 0, z = 0, t;
    if constexpr (Axis == 0) {
        t = i % ntiles;
        const uint32_t r = i / ntiles;
        y = r % act.ny;
        z = r / act.ny;
    } else if constexpr (Axis == 1) {
        x = i % act.nx;
        const uint32_t r = i / act.nx;
        t = r % ntiles;
        z = r / ntiles;
    } else {
        x = i % act.nx;
        const uint32_t r = i / act.nx;
        y = r % act.ny;
        t = r / act.ny;
    }
    const uint64_t plane = uint64_t(full.nx) * full.ny;
    const uint32_t n = Axis == 0 ? act.nx : Axis == 1 ? act.ny : act.nz;
    const uint64_t stride = Axis == 0 ? 1u : Axis == 1 ? uint64_t(full.nx) : plane;
    const uint64_t base = uint64_t(z) * plane + uint64_t(y) * full.nx + x;
    const int32_t start = int32_t(t) * Tile;
    T e[P], o[P];
#pragma unroll
    for (int j = 0; j < P; ++j) {
        const int32_t p = start - 4 + 2 * j;
        e[j] = src[base + uint64_t(reflect_pos(p, n)) * stride];
        o[j] = src[base + uint64_t(reflect_pos(p + 1, n)) * stride];
    }
    bool ok = true;
    if (screen) {  // the input screen of the first pass (cpu/src/codec.cpp:183)
#pragma unroll
        for (int j = 0; j < P; ++j) ok = ok && __builtin_isfinite(e[j]) && __builtin_isfinite(o[j]);
    }
    lift_forward<T, P>(e, o);
    const uint32_t h = half_len(n);
#pragma unroll
    for (int j = 2; j < P - 2; ++j) {
        const uint32_t p = uint32_t(start) + 2u * uint32_t(j - 2);
        if (p < n) dst[base + uint64_t(p >> 1) * stride] = e[j];
        if (p + 1 < n) dst[base + uint64_t(h + (p >> 1)) * stride] = o[j];
    }
    return ok;
}

enum : uint8_t { kInput = 4, kRecon = 5 };

template <class T>
__global__ void __launch_bounds__(256) batch_fwd_pass_kernel(const JobShape* __restrict__ shape,
                                                             const Pass* __restrict__ pass,
                                                             const uint32_t* __restrict__ pfx, uint32_t jobs,
                                                             CompBufs b) {
    constexpr int ZTile = sizeof(T) == 4 ? 16 : 8;
    const uint32_t j = job_of_wave(pfx, jobs, blockIdx.x);
    const Pass ps = pass[j];
    const uint32_t i = (blockIdx.x - pfx[j]) * 256u + threadIdx.x;
    if (ps.kind == kPassNone) return;
    const JobShape& sh = shape[j];
    auto bank = [&](uint8_t id) -> T* {
        if (id == kInput) return const_cast<T*>(static_cast<const T*>(b.input)) + sh.out;
        T* p = static_cast<T*>(id == kHome ? b.home : id == kTmp0 ? b.tmp0 : b.tmp1);
        return p + sh.bank;
    };
    const T* src = bank(ps.src);
    T* dst = bank(ps.dst);
    const bool screen = ps.src == kInput;
    bool ok = true;
    if (i < ps.work) {
        switch (ps.kind) {
            case kPassX: ok = forward_line<T, 0, 8>(src, dst, sh.d, ps.act, ps.ntiles, i, screen); break;
            case kPassY: ok = forward_line<T, 1, 8>(src, dst, sh.d, ps.act, ps.ntiles, i, screen); break;
            // the Z tile make_pass sized the pass for (forward tiles change no bit)
            case kPassZ: ok = forward_line<T, 2, ZTile>(src, dst, sh.d, ps.act, ps.ntiles, i, screen); break;
            default: {
                const uint32_t x = i % ps.act.nx, r = i / ps.act.nx, y = r % ps.act.ny, z = r / ps.act.ny;
                const uint64_t at = (uint64_t(z) * sh.d.ny + y) * sh.d.nx + x;
                dst[at] = src[at];
            }
        }
    }
    if (__ballot(!ok) && (threadIdx.x & 63u) == 0) atomicOr(&b.res[j].status, uint32_t(kNonFiniteInput));
}

template <class T>
__global__ void __launch_bounds__(256) batch_quantize_kernel(const JobShape* __restrict__ shape,
                                                             const JobComp* __restrict__ comp,
                                                             const uint32_t* __restrict__ brick_pfx, uint32_t jobs,
                                                             const uint16_t* __restrict__ ranks, double quantum,
                                                             CompBufs b) {
    const uint32_t j = job_of_wave(brick_pfx, jobs, blockIdx.x);
    const JobShape& sh = shape[j];
    const JobComp& jc = comp[j];
    const BrickGeometry g{sh.d, sh.bz, sh.nbx, sh.nby, sh.nbz};
    const SideList qo{b.qbm + jc.words, b.qidx + jc.qlist, b.qval + jc.qlist, b.qcount + j, jc.qcap};
    // up to kBricksPerBlock consecutive bricks of the job, one after another
    const uint32_t b0 = (blockIdx.x - brick_pfx[j]) * kBricksPerBlock;
    const uint32_t b1 = min(b0 + kBricksPerBlock, sh.nbx * sh.nby * sh.nbz);
    for (uint32_t k = b0; k < b1; ++k) {
        if (k != b0) __syncthreads();  // the previous brick's sign words are out of LDS
        quantize_brick(static_cast<T*>(b.home) + sh.bank, g, T(quantum), T(1.0 / quantum),
                       SetRank{ranks + uint64_t(sh.rank) * kRankSet},
                       static_cast<uint16_t*>(b.tmp0) + sh.mag, static_cast<uint32_t*>(b.tmp1) + sh.sign, qo,
                       &b.res[j].status, k);
    }
}

WZ_D WzpScratch job_wzp(const CompBufs& b, const JobComp& jc) {
    return WzpScratch{b.desc + jc.group0, b.staging + jc.staging, b.sizes + jc.chunk0, nullptr};
}

// Chunks of the jobs that are not brick pairs (mag_pfx counts only theirs):
// the magnitudes and signs of batch_quantize_kernel.
__global__ void __launch_bounds__(256) batch_encode_kernel(const JobShape* __restrict__ shape,
                                                           const JobComp* __restrict__ comp,
                                                           const uint32_t* __restrict__ mag_pfx, uint32_t jobs,
                                                           CompBufs b) {
    const uint32_t j = job_of_wave(mag_pfx, jobs, blockIdx.x);
    const JobShape& sh = shape[j];
    const JobComp& jc = comp[j];
    const MagSource src{static_cast<const uint16_t*>(b.tmp0) + sh.mag, static_cast<const uint32_t*>(b.tmp1) + sh.sign};
    encode_chunk(src, blockIdx.x - mag_pfx[j], sh.n, sh.groups, b.out + jc.arc + 32, jc.cap - 32, job_wzp(b, jc),
                 &b.res[j].status);
}

// Chunks of the brick-pair jobs (pair_pfx): quantized inside the encoder, as
// the single-field BrickSource does, with the dequantized coefficients
// written back over the whole job for the inverse passes.
template <class T>
__global__ void __launch_bounds__(256) batch_encode_pairs_kernel(const JobShape* __restrict__ shape,
                                                                 const JobComp* __restrict__ comp,
                                                                 const uint32_t* __restrict__ pair_pfx,
                                                                 uint32_t jobs, double quantum, CompBufs b) {
    const uint32_t j = job_of_wave(pair_pfx, jobs, blockIdx.x);
    const JobShape& sh = shape[j];
    const JobComp& jc = comp[j];
    const SideList qo{b.qbm + jc.words, b.qidx + jc.qlist, b.qval + jc.qlist, b.qcount + j, jc.qcap};
    const BrickSource<T> src{static_cast<T*>(b.home) + sh.bank, sh.d.nx, sh.d.ny, T(quantum), T(1.0 / quantum), qo,
                             sh.d.nx, sh.d.ny, sh.d.nz};
    encode_chunk(src, blockIdx.x - pair_pfx[j], sh.n, sh.groups, b.out + jc.arc + 32, jc.cap - 32, job_wzp(b, jc),
                 &b.res[j].status);
}

// Chunks of the other full-brick jobs (rank_pfx): quantized inside the
// encoder, as the single-field RankSource does, with the dequantized
// coefficients written back over the whole job for the inverse passes.
template <class T>
__global__ void __launch_bounds__(256) batch_encode_ranks_kernel(const JobShape* __restrict__ shape,
                                                                 const JobComp* __restrict__ comp,
                                                                 const uint32_t* __restrict__ rank_pfx,
                                                                 uint32_t jobs, const uint16_t* __restrict__ ranks,
                                                                 double quantum, CompBufs b) {
    const uint32_t j = job_of_wave(rank_pfx, jobs, blockIdx.x);
    const JobShape& sh = shape[j];
    const JobComp& jc = comp[j];
    const SideList qo{b.qbm + jc.words, b.qidx + jc.qlist, b.qval + jc.qlist, b.qcount + j, jc.qcap};
    const RankBricks g{ranks + uint64_t(sh.rank) * kRankSet, sh.d.nx, sh.d.ny, sh.bz};
    const RankSource<T> src{static_cast<T*>(b.home) + sh.bank, g, T(quantum), T(1.0 / quantum), qo,
                            sh.d.nx, sh.d.ny, sh.d.nz};
    encode_chunk(src, blockIdx.x - rank_pfx[j], sh.n, sh.groups, b.out + jc.arc + 32, jc.cap - 32, job_wzp(b, jc),
                 &b.res[j].status);
}

// Groups of all jobs in one ticket order (job by job, group by group), so
// every group's predecessors in its own stream are held by running workgroups.
__global__ void __launch_bounds__(kPackThreads) batch_pack_kernel(const JobShape* __restrict__ shape,
                                                                  const JobComp* __restrict__ comp,
                                                                  const uint32_t* __restrict__ group_pfx,
                                                                  uint32_t jobs, CompBufs b) {
    __shared__ uint32_t g_sh;
    if (threadIdx.x == 0) g_sh = atomicAdd(b.ticket, 1u);
    __syncthreads();
    const uint32_t j = job_of_wave(group_pfx, jobs, g_sh);
    const JobShape& sh = shape[j];
    const JobComp& jc = comp[j];
    pack_group(g_sh - group_pfx[j], b.out + jc.arc + 32, jc.cap - 32, sh.chunks, sh.groups, job_wzp(b, jc),
               &b.res[j].wzp_bytes, &b.res[j].status);
}

template <class T>
WZ_D SegList job_repairs(const CompBufs& b, const JobComp& jc) {
    return SegList{b.rbm + jc.words, b.ridx + jc.rlist, static_cast<T*>(b.rval) + jc.rlist, b.rcount + jc.tile0,
                   jc.tiles, kRepairTile};
}

template <class T>
__global__ void __launch_bounds__(256) batch_repair_kernel(const JobShape* __restrict__ shape,
                                                           const JobComp* __restrict__ comp,
                                                           const uint32_t* __restrict__ tile_pfx, uint32_t jobs,
                                                           double bound, CompBufs b) {
    const uint32_t j = job_of_wave(tile_pfx, jobs, blockIdx.x);
    const JobShape& sh = shape[j];
    const JobComp& jc = comp[j];
    repair_tile(static_cast<const T*>(b.input) + sh.out, static_cast<const T*>(b.recon) + sh.out, sh.n, bound,
                job_repairs<T>(b, jc), blockIdx.x - tile_pfx[j], &b.res[j].status);
}

// Rank scratch of a job at jc.prefix: [repair wpre | repair bsum | q wpre | q bsum]
// (finish_scratch_words(n) words, each part a multiple of 64 words).
struct JobRanks {
    Ranks r, q;
    uint32_t *rbsum, *qbsum;
    uint32_t words, blocks;
};
WZ_D JobRanks job_ranks(const CompBufs& b, const JobShape& sh, const JobComp& jc) {
    const uint32_t words = (sh.n + 31) / 32, blocks = (words + kPrefixWords - 1) / kPrefixWords;
    const uint64_t pw = (uint64_t(words) + 63) & ~uint64_t(63), pb = (uint64_t(blocks) + 63) & ~uint64_t(63);
    uint32_t* base = b.prefix + jc.prefix;
    JobRanks k;
    k.rbsum = base + pw;
    k.qbsum = base + pw + pb + pw;
    k.r = Ranks{b.rbm + jc.words, base, k.rbsum};
    k.q = Ranks{b.qbm + jc.words, base + pw + pb, k.qbsum};
    k.words = words;
    k.blocks = blocks;
    return k;
}

// Word prefixes of both flag bitmaps of every job: block k of the launch is
// prefix block (k / 2 - pblock_pfx[j]) of job j, bitmap k % 2.
__global__ void __launch_bounds__(256) batch_prefix_kernel(const JobShape* __restrict__ shape,
                                                           const JobComp* __restrict__ comp,
                                                           const uint32_t* __restrict__ pblock_pfx, uint32_t jobs,
                                                           CompBufs b) {
    const uint32_t k = blockIdx.x >> 1, which = blockIdx.x & 1u;
    const uint32_t j = job_of_wave(pblock_pfx, jobs, k);
    const JobShape& sh = shape[j];
    const JobRanks r = job_ranks(b, sh, comp[j]);
    const uint32_t blk = k - pblock_pfx[j];
    if (which == 0) word_prefix_block(r.r.bm, r.words, const_cast<uint32_t*>(r.r.wpre), r.rbsum, blk);
    else if (b.qcount[j] != 0) word_prefix_block(r.q.bm, r.words, const_cast<uint32_t*>(r.q.wpre), r.qbsum, blk);
}

// One 1024-thread workgroup per job: block sums, layout, header, padding.
template <class T>
__global__ void __launch_bounds__(1024) batch_layout_kernel(const JobShape* __restrict__ shape,
                                                            const JobComp* __restrict__ comp, double quantum,
                                                            CompBufs b) {
    const uint32_t j = blockIdx.x;
    const JobShape& sh = shape[j];
    const JobComp& jc = comp[j];
    const JobRanks r = job_ranks(b, sh, jc);
    const SideList qo{b.qbm + jc.words, b.qidx + jc.qlist, b.qval + jc.qlist, b.qcount + j, jc.qcap};
    const uint8_t flags = uint8_t((sh.dyadic ? 1u : 0u) | (sizeof(T) == 8 ? 4u : 0u) | (3u << 3));
    scan_layout<T>(b.out + jc.arc, jc.cap, sh.n, r.rbsum, r.qbsum, r.blocks, &b.res[j].pad, qo, flags,
                   uint8_t(sh.lx), uint8_t(sh.lz), uint8_t(sh.bz), quantum, &b.res[j]);
    if (threadIdx.x == 0) {
        b.spans[2 * j] = jc.arc;
        b.spans[2 * j + 1] = b.res[j].status ? 0ull : b.res[j].end;
    }
}

// Both side lists of every job to their index-order slots: per job, one
// workgroup per repair tile (at most 4096 entries, about 2% of them in use at
// typical bounds), then qblocks for the q-outliers. The grid is fixed before
// the counts are known, so it stays small: with eight workgroups per tile
// and up to 256 q-outlier blocks per job, most of them only returned and the
// launch took 262 µs for 256 jobs.
template <class T>
__global__ void __launch_bounds__(256) batch_place_kernel(const JobShape* __restrict__ shape,
                                                          const JobComp* __restrict__ comp,
                                                          const uint32_t* __restrict__ place_pfx, uint32_t jobs,
                                                          CompBufs b) {
    const uint32_t j = job_of_wave(place_pfx, jobs, blockIdx.x);
    const JobShape& sh = shape[j];
    const JobComp& jc = comp[j];
    const EncodeResult* res = &b.res[j];
    if (res->status) return;
    const JobRanks r = job_ranks(b, sh, jc);
    const uint32_t g = blockIdx.x - place_pfx[j], rblocks = jc.tiles;
    uint8_t* archive = b.out + jc.arc;
    if (g < rblocks) {
        place_repairs<T>(archive, job_repairs<T>(b, jc), r.r, res, g, 0, 1);
        return;
    }
    const uint32_t np = res->np, nq = res->nq;
    const Layout L = archive_layout(res->wzp_bytes, np, uint64_t(np) * sizeof(T), nq);
    const uint32_t* qi = b.qidx + jc.qlist;
    const uint32_t* qv = b.qval + jc.qlist;
    for (uint32_t e = (g - rblocks) * 256u + threadIdx.x; e < nq; e += jc.qblocks * 256u) {
        const uint32_t idx = qi[e];
        const uint32_t pos = r.q.of(idx);
        reinterpret_cast<uint32_t*>(archive + L.qi)[pos] = idx;
        reinterpret_cast<uint32_t*>(archive + L.qv)[pos] = qv[e];
    }
}

// ------------------------------------------------------------------ LDS-staged passes
// Two pass kinds that stage a whole piece of the box in LDS and run several
// lifting passes there, reading and writing the box once:
//   kPassXYPlane  one workgroup per z plane of the box: the X and Y passes of
//                 `levels` XY levels (Pass::pad) in canonical order (inverse:
//                 coarse to fine, X then Y; forward: fine to coarse, Y then X);
//   kPassZCol     one workgroup per row of 64 columns: the Z passes of
//                 `levels` Z levels (the plane schedule's whole Z phase).
// Tiles and arithmetic are those of the per-axis passes (lift_tile), so every
// value is the same. In place is allowed (src may equal dst).
enum : uint8_t { kPassXYPlane = 6, kPassZCol = 7 };
constexpr uint32_t kColumns = 64;          // columns of a kPassZCol workgroup
// Samples of the one LDS buffer a workgroup stages (odd plane pitches: rows
// of up to 128 + 1 samples for the large variant).
template <class T, bool Large>
struct PlaneCap {
    static constexpr uint32_t value = (Large ? 66560u : 16640u) / sizeof(T);
};
template <class T>
constexpr uint32_t zcol_cap() { return PlaneCap<T, true>::value / kColumns; }  // z extent of a column row

// Banks of the LDS passes: input (compress forward), the three work banks, the
// output (decode) or reconstruction (compress inverse), and where failures go.
struct StageBanks {
    const void* input;
    void* home;
    void* tmp0;
    void* tmp1;
    void* out;
    EncodeResult* res;  // compress: per-job status; decode: null (status word)
};

template <class T>
WZ_D T* stage_bank(const StageBanks& b, const JobShape& sh, uint8_t id) {
    if (id == kInput) return const_cast<T*>(static_cast<const T*>(b.input)) + sh.out;
    return static_cast<T*>(id == kHome ? b.home : id == kTmp0 ? b.tmp0 : b.tmp1) + sh.bank;
}

template <class T, bool Large, bool Inverse>
__global__ void __launch_bounds__(kLdsThreads) batch_lds_kernel(const JobShape* __restrict__ shape,
                                                                const JobRun* __restrict__ run,
                                                                const Pass* __restrict__ pass,
                                                                const uint32_t* __restrict__ pfx, uint32_t jobs,
                                                                StageBanks b, T fill, uint32_t* status) {
    constexpr uint32_t Cap = PlaneCap<T, Large>::value;
    constexpr int ZTile = Inverse ? (sizeof(T) == 4 ? 16 : 8) : 8;
    __shared__ T A[Cap];
    const uint32_t j = job_of_wave(pfx, jobs, blockIdx.x);
    const Pass ps = pass[j];
    if (ps.kind != kPassXYPlane && ps.kind != kPassZCol) return;
    const uint32_t state = run ? run[j].state : uint32_t(kJobOk);
    if (state == kJobBad || (state == kJobHole && ps.dst != kOutput)) return;
    const JobShape& sh = shape[j];
    const uint32_t t = threadIdx.x, item = blockIdx.x - pfx[j];
    const Dims full = sh.d, act = ps.act;
    const uint64_t plane = uint64_t(full.nx) * full.ny;
    const T* src = stage_bank<T>(b, sh, ps.src);
    const bool screen = ps.src == kInput;
    const bool xy = ps.kind == kPassXYPlane;
    // the staged piece: plane z (rows of act.nx samples at an odd pitch), or
    // columns x0.. of row y (rows of up to 64 z-ordered columns)
    const uint32_t z = item, y = xy ? 0u : item % act.ny, x0 = xy ? 0u : (item / act.ny) * kColumns;
    const uint32_t cols = xy ? act.nx : min(kColumns, act.nx - x0);
    const uint32_t rows = xy ? act.ny : act.nz;
    const uint32_t pitch = xy ? (act.nx | 1u) : kColumns;
    const uint32_t cells = cols * rows;
    const FastDiv by_cols(cols);  // cells fit LDS: exact
    auto gat = [&](uint32_t c, uint32_t r) -> uint64_t {  // global index of staged cell (c, r)
        return xy ? uint64_t(z) * plane + uint64_t(r) * full.nx + c
                  : uint64_t(r) * plane + uint64_t(y) * full.nx + x0 + c;
    };
    // staging: every thread issues kBurst loads before it writes any of them to LDS
    constexpr uint32_t kBurst = 8;
    bool ok = true;
    for (uint32_t i0 = 0; i0 < cells; i0 += kBurst * kLdsThreads) {
        T v[kBurst];
#pragma unroll
        for (uint32_t k = 0; k < kBurst; ++k) {
            const uint32_t i = i0 + k * kLdsThreads + t;
            const uint32_t r = by_cols.div(i);
            v[k] = i < cells ? src[gat(i - r * cols, r)] : T(0);
        }
#pragma unroll
        for (uint32_t k = 0; k < kBurst; ++k) {
            const uint32_t i = i0 + k * kLdsThreads + t;
            if (i < cells) {
                ok = ok && (!screen || __builtin_isfinite(v[k]));
                const uint32_t r = by_cols.div(i);
                A[r * pitch + (i - r * cols)] = v[k];
            }
        }
    }
    __syncthreads();
    const uint32_t levels = ps.pad;
    uint32_t ex[5], ey[5];  // extents of each level inside the piece
    ex[0] = xy ? act.nx : act.nz;
    ey[0] = act.ny;
    for (uint32_t l = 1; l < 5; ++l) {
        ex[l] = half_len(ex[l - 1]);
        ey[l] = half_len(ey[l - 1]);
    }
    for (uint32_t k = 0; k < levels; ++k) {
        const uint32_t l = Inverse ? levels - 1 - k : k;
        if (xy) {
            if constexpr (Inverse) {
                lds_pass<T, true, 0, 8>(A, pitch, ex[l], ey[l]);  // X (unfolded), rows of the level box
                lds_pass<T, true, 1, 8>(A, pitch, ey[l], ex[l]);  // Y (folded), its columns
            } else {
                lds_pass<T, false, 1, 8>(A, pitch, ey[l], ex[l]);  // Y
                lds_pass<T, false, 0, 8>(A, pitch, ex[l], ey[l]);  // X
            }
        } else {
            lds_pass<T, Inverse, 2, ZTile>(A, pitch, ex[l], cols);  // Z (folded) of every column
        }
    }
    bool fin = true;
    if (ps.dst == kOutput) {
        const Put<T> put{nullptr, static_cast<T*>(b.out), &sh, state == kJobHole, fill};
        for (uint32_t i = t; i < cells; i += kLdsThreads) {
            const uint32_t r = by_cols.div(i), c = i - r * cols;
            const uint32_t X = xy ? c : x0 + c, Y = xy ? r : y, Zc = xy ? z : r;
            fin = put(gat(c, r), X, Y, Zc, A[r * pitch + c]) && fin;
        }
        if (__ballot(!fin) && (t & 63u) == 0 && status) atomicOr(status, uint32_t(kNonFiniteOutput));
    } else {
        T* dst = stage_bank<T>(b, sh, ps.dst);
        for (uint32_t i = t; i < cells; i += kLdsThreads) {
            const uint32_t r = by_cols.div(i), c = i - r * cols;
            dst[gat(c, r)] = A[r * pitch + c];
        }
    }
    if (screen && __ballot(!ok) && (t & 63u) == 0 && b.res) atomicOr(&b.res[j].status, uint32_t(kNonFiniteInput));
}

// ------------------------------------------------------------------ host plan
Dims halve(Dims d, bool xy, bool z) {
    if (xy) {
        d.nx = half_len(d.nx);
        d.ny = half_len(d.ny);
    }
    if (z) d.nz = half_len(d.nz);
    return d;
}

Pass make_pass(uint8_t kind, uint8_t src, uint8_t dst, Dims act, unsigned width) {
    Pass p{};
    p.kind = kind;
    p.src = src;
    p.dst = dst;
    p.act = act;
    const uint64_t vol = uint64_t(act.nx) * act.ny * act.nz;
    if (kind == kPassCopy) {
        p.ntiles = 1;
        p.work = uint32_t(vol);
        return p;
    }
    const uint32_t n = kind == kPassX ? act.nx : kind == kPassY ? act.ny : act.nz;
    const uint32_t tile = kind == kPassZ ? (width == 4 ? 16u : 8u) : 8u;
    p.ntiles = (n + tile - 1) / tile;
    p.work = uint32_t(vol / n * p.ntiles);
    return p;
}

// Whether a box's XY planes / Z columns fit the LDS-staged passes.
bool plane_fits(Dims act, unsigned width) {
    // rows at an odd pitch; every tile column of a line fits one lds_pass group
    const uint64_t cap = width == 4 ? PlaneCap<float, true>::value : PlaneCap<double, true>::value;
    const uint64_t k = width == 4 ? 4 : 2;
    return uint64_t(act.nx | 1u) * act.ny <= cap && (act.nx + 7) / 8 <= k * kLdsThreads &&
           (act.ny + 7) / 8 <= k * kLdsThreads;
}
bool zcol_fits(Dims act, unsigned width) { return act.nz <= (width == 4 ? zcol_cap<float>() : zcol_cap<double>()); }

Pass lds_pass(uint8_t kind, uint8_t src, uint8_t dst, Dims act, uint32_t levels) {
    Pass p{};
    p.kind = kind;
    p.src = src;
    p.dst = dst;
    p.pad = uint8_t(levels);
    p.act = act;
    p.ntiles = 1;
    p.work = kind == kPassXYPlane ? act.nz : act.ny * ((act.nx + kColumns - 1) / kColumns);  // workgroups
    return p;
}

// The inverse schedule of transform_inverse (cpu/src/transform.cpp:409-447):
// every level ends in the home bank, the very last pass writes the output.
// A dyadic level is its Z pass and one XY-plane pass (X then Y) when the
// level's planes fit LDS; the plane schedule is one XY-plane pass for all XY
// levels and one Z-column pass for all Z levels when they fit. Otherwise per
// axis passes through the scratch banks (copied back to home).
std::vector<Pass> inverse_schedule(const JobShape& sh, unsigned width) {
    std::vector<Pass> out;
    if (sh.lx == 0 && sh.lz == 0) {
        out.push_back(make_pass(kPassCopy, kHome, kOutput, sh.d, width));
        return out;
    }
    struct Op {
        uint8_t kind;
        Dims act;
        uint32_t levels;
    };
    std::vector<std::vector<Op>> levels;  // ops of each level (one level ends in home)
    auto axes = [&](Dims act, std::initializer_list<uint8_t> ks) {
        std::vector<Op> v;
        for (uint8_t k : ks) v.push_back({k, act, 1});
        return v;
    };
    if (sh.dyadic) {
        Dims sizes[5], act = sh.d;
        for (unsigned l = 0; l < sh.lx; ++l) {
            sizes[l] = act;
            act = halve(act, true, true);
        }
        for (int l = int(sh.lx) - 1; l >= 0; --l) {
            if (plane_fits(sizes[l], width)) levels.push_back({{kPassZ, sizes[l], 1}, {kPassXYPlane, sizes[l], 1}});
            else levels.push_back(axes(sizes[l], {kPassZ, kPassX, kPassY}));
        }
    } else {
        Dims xy[5], zs[5], act = sh.d;
        for (unsigned l = 0; l < sh.lx; ++l) {
            xy[l] = act;
            act = halve(act, true, false);
        }
        act = sh.d;
        for (unsigned l = 0; l < sh.lz; ++l) {
            zs[l] = act;
            act = halve(act, false, true);
        }
        if (sh.lx && plane_fits(sh.d, width)) levels.push_back({{kPassXYPlane, sh.d, sh.lx}});
        else
            for (int l = int(sh.lx) - 1; l >= 0; --l) levels.push_back(axes(xy[l], {kPassX, kPassY}));
        if (sh.lz && zcol_fits(sh.d, width)) levels.push_back({{kPassZCol, sh.d, sh.lz}});
        else
            for (int l = int(sh.lz) - 1; l >= 0; --l) levels.push_back(axes(zs[l], {kPassZ}));
    }
    for (size_t k = 0; k < levels.size(); ++k) {
        const bool last = k + 1 == levels.size();
        const uint8_t fin = last ? uint8_t(kOutput) : uint8_t(kHome);
        uint8_t src = kHome;
        const std::vector<Op>& L = levels[k];
        for (size_t a = 0; a < L.size(); ++a) {
            const Op& op = L[a];
            const bool lds = op.kind == kPassXYPlane || op.kind == kPassZCol;
            uint8_t dst;
            if (a + 1 == L.size()) dst = fin;
            else dst = src == kTmp0 ? uint8_t(kTmp1) : uint8_t(kTmp0);
            if (dst == src && !lds) dst = src == kTmp0 ? uint8_t(kTmp1) : uint8_t(kTmp0);
            out.push_back(lds ? lds_pass(op.kind, src, dst, op.act, op.levels) : make_pass(op.kind, src, dst, op.act, width));
            src = dst;
        }
        if (src != fin) out.push_back(make_pass(kPassCopy, src, fin, L.back().act, width));
    }
    return out;
}

// The forward schedule of transform_forward (cpu/src/transform.cpp:425-456):
// the first pass reads the job's input, every level ends in the home bank.
// A dyadic level is one XY-plane pass (Y then X) and its Z pass; the plane
// schedule is one Z-column pass and one XY-plane pass, when they fit LDS.
// Forward tiles do not change a bit.
std::vector<Pass> forward_schedule(const JobShape& sh, unsigned width) {
    std::vector<Pass> out;
    if (sh.lx == 0 && sh.lz == 0) {
        out.push_back(make_pass(kPassCopy, kInput, kHome, sh.d, width));
        return out;
    }
    struct Op {
        uint8_t kind;
        Dims act;
        uint32_t levels;
    };
    std::vector<std::vector<Op>> levels;
    if (sh.dyadic) {
        Dims act = sh.d;
        for (unsigned l = 0; l < sh.lx; ++l) {
            if (plane_fits(act, width)) levels.push_back({{kPassXYPlane, act, 1}, {kPassZ, act, 1}});
            else levels.push_back({{kPassY, act, 1}, {kPassX, act, 1}, {kPassZ, act, 1}});
            act = halve(act, true, true);
        }
    } else {
        if (sh.lz && zcol_fits(sh.d, width)) {
            levels.push_back({{kPassZCol, sh.d, sh.lz}});
        } else {
            Dims act = sh.d;
            for (unsigned l = 0; l < sh.lz; ++l) {
                levels.push_back({{kPassZ, act, 1}});
                act = halve(act, false, true);
            }
        }
        if (sh.lx && plane_fits(sh.d, width)) {
            levels.push_back({{kPassXYPlane, sh.d, sh.lx}});
        } else {
            Dims act = sh.d;
            for (unsigned l = 0; l < sh.lx; ++l) {
                levels.push_back({{kPassY, act, 1}, {kPassX, act, 1}});
                act = halve(act, true, false);
            }
        }
    }
    uint8_t src = kInput;
    for (const std::vector<Op>& L : levels) {
        for (size_t a = 0; a < L.size(); ++a) {
            const Op& op = L[a];
            const bool lds = op.kind == kPassXYPlane || op.kind == kPassZCol;
            uint8_t dst;
            if (a + 1 == L.size()) dst = kHome;
            else dst = src == kTmp0 ? uint8_t(kTmp1) : uint8_t(kTmp0);
            if (dst == src && !lds) dst = src == kTmp0 ? uint8_t(kTmp1) : uint8_t(kTmp0);
            out.push_back(lds ? lds_pass(op.kind, src, dst, op.act, op.levels) : make_pass(op.kind, src, dst, op.act, width));
            src = dst;
        }
        if (src != kHome) {
            out.push_back(make_pass(kPassCopy, src, kHome, L.back().act, width));
            src = kHome;
        }
    }
    return out;
}

// Per-step launch tables: every job's pass at step st, the block prefixes of
// the per-axis kernel (256-thread blocks of work items) and of the LDS kernel
// (one workgroup per plane / column row), and whether that step's LDS pieces
// need the large LDS variant.
void build_steps(const std::vector<std::vector<Pass>>& sched, uint32_t J, unsigned width,
                 std::vector<std::vector<Pass>>& step, std::vector<std::vector<uint32_t>>& pfx,
                 std::vector<std::vector<uint32_t>>& lpfx, std::vector<uint8_t>& large) {
    size_t steps = 0;
    for (const auto& v : sched) steps = std::max(steps, v.size());
    step.assign(steps, std::vector<Pass>(J));
    pfx.assign(steps, std::vector<uint32_t>(J + 1, 0));
    lpfx.assign(steps, std::vector<uint32_t>(J + 1, 0));
    large.assign(steps, 0);
    const uint64_t small = width == 4 ? PlaneCap<float, false>::value : PlaneCap<double, false>::value;
    for (size_t st = 0; st < steps; ++st) {
        uint64_t blocks = 0, lblocks = 0;
        for (uint32_t j = 0; j < J; ++j) {
            Pass ps{};
            ps.kind = kPassNone;
            if (st < sched[j].size()) ps = sched[j][st];
            step[st][j] = ps;
            pfx[st][j] = uint32_t(blocks);
            lpfx[st][j] = uint32_t(lblocks);
            if (ps.kind == kPassXYPlane || ps.kind == kPassZCol) {
                lblocks += ps.work;
                const uint64_t piece = ps.kind == kPassXYPlane ? uint64_t(ps.act.nx | 1u) * ps.act.ny
                                                               : uint64_t(ps.act.nz) * kColumns;
                if (piece > small) large[st] = 1;
            } else if (ps.kind != kPassNone) {
                blocks += (uint64_t(ps.work) + 255) / 256;
            }
        }
        if (blocks > 0x7FFFFFFFull || lblocks > 0x7FFFFFFFull) throw std::invalid_argument("batch launch limit");
        pfx[st][J] = uint32_t(blocks);
        lpfx[st][J] = uint32_t(lblocks);
    }
}

// Byte offsets of one batched compress call's workspace (all 256-byte aligned).
struct CompressLayout {
    uint64_t home, tmp0, tmp1, recon, staging, qbm, rbm, qidx, qval, ridx, rval, rcount, prefix, sizes, desc;
    uint64_t zero0, qcount, ticket, res, zero_end, total;
};

struct HostPlan {
    unsigned width = 4;
    uint32_t jobs = 0;
    std::vector<JobShape> shape;
    std::vector<uint32_t> chunk_pfx, brick_pfx;
    // chunks of the brick-pair jobs and of the others (WZP coder launches)
    std::vector<uint32_t> pair_pfx, rank_pfx, mag_pfx;
    std::vector<std::vector<Pass>> step;      // step[s][job]
    std::vector<std::vector<uint32_t>> pfx;   // block prefixes per step (per-axis kernel)
    std::vector<std::vector<uint32_t>> lpfx;  // block prefixes per step (LDS kernel)
    std::vector<uint8_t> large;               // step needs the large LDS variant
    uint64_t bank_elems = 0, mag_elems = 0, sign_words = 0, out_elems = 0;
    // compress
    std::vector<std::vector<Pass>> fstep;     // forward passes, fstep[s][job]
    std::vector<std::vector<uint32_t>> fpfx, flpfx;
    std::vector<uint8_t> flarge;
    std::vector<JobComp> comp;
    std::vector<uint32_t> group_pfx, tile_pfx, pblock_pfx, place_pfx;
    uint64_t archive_bytes = 0;               // archive slots of all jobs
    CompressLayout cl{};
    std::vector<std::pair<Dims, uint32_t>> rank_geo;  // distinct (geometry, z band) of the rank table sets
};

// Geometry checks of the single-field codec (codec.hip geometry()).
JobShape job_shape(Dims d, unsigned width) {
    if (!d.nx || !d.ny || !d.nz) throw std::invalid_argument("nonzero dimensions required");
    const uint64_t n = uint64_t(d.nx) * d.ny * d.nz;
    if (n > 0xFFFFFFFFull / width || n > 0xFFFFFFFFull / 2) throw std::invalid_argument("tile exceeds format limits");
    JobShape s{};
    s.d = d;
    s.n = uint32_t(n);
    s.lx = dwt_levels(d.nx < d.ny ? d.nx : d.ny);
    s.lz = dwt_levels(d.nz);
    s.dyadic = s.lx == s.lz;
    s.bz = zband(d, s.lz, s.dyadic);
    const WzpShape ws = wzp_shape(s.n);
    s.chunks = ws.chunks;
    s.groups = ws.groups;
    s.nbx = (d.nx + 15) / 16;
    s.nby = (d.ny + 15) / 16;
    s.nbz = (d.nz + s.bz - 1) / s.bz;
    s.clip = d;
    return s;
}

// Stage tables and workspace layout of the batched compress (for the same
// job shapes; inputs use the JobShape::out offsets).
void finish_compress_plan(HostPlan& p) {
    const uint32_t J = p.jobs;
    const unsigned w = p.width;
    std::vector<std::vector<Pass>> sched(J);
    size_t steps = 0;
    p.comp.assign(J, JobComp{});
    p.group_pfx.assign(J + 1, 0);
    p.tile_pfx.assign(J + 1, 0);
    p.pblock_pfx.assign(J + 1, 0);
    p.place_pfx.assign(J + 1, 0);
    uint64_t arc = 0, staging = 0, qlist = 0, words = 0, prefix = 0, groups = 0, tiles = 0, pblocks = 0, places = 0;
    for (uint32_t j = 0; j < J; ++j) {
        const JobShape& s = p.shape[j];
        JobComp& c = p.comp[j];
        c.arc = arc;
        c.cap = max_compressed_bytes(s.d, w);
        arc += (c.cap + 15) & ~uint64_t(15);
        c.staging = staging;
        staging += uint64_t(s.chunks) * kChunkSlot;
        c.qcap = s.n <= (1u << 20) ? s.n : (1u << 20);
        c.qlist = qlist;
        qlist += c.qcap;
        c.tiles = (s.n + kRepairTile - 1) / kRepairTile;
        c.words = words;
        const uint64_t wds = (uint64_t(s.n) + 31) / 32;
        words += (wds + 3) & ~uint64_t(3);
        c.prefix = prefix;
        prefix += finish_scratch_words(s.n);
        c.chunk0 = p.chunk_pfx[j];
        c.group0 = uint32_t(groups);
        c.tile0 = uint32_t(tiles);
        c.pblocks = uint32_t((wds + kPrefixWords - 1) / kPrefixWords);
        c.qblocks = std::max(1u, std::min(16u, (c.qcap + 4095) / 4096));
        c.rlist = uint64_t(c.tile0) * kRepairTile;
        p.group_pfx[j] = c.group0;
        p.tile_pfx[j] = c.tile0;
        p.pblock_pfx[j] = uint32_t(pblocks);
        p.place_pfx[j] = uint32_t(places);
        groups += s.groups;
        tiles += c.tiles;
        pblocks += c.pblocks;
        places += uint64_t(c.tiles) + c.qblocks;
        sched[j] = forward_schedule(s, w);
        steps = std::max(steps, sched[j].size());
    }
    if (groups > 0x7FFFFFFFull || tiles > 0x7FFFFFFFull || 2 * pblocks > 0x7FFFFFFFull || places > 0x7FFFFFFFull)
        throw std::invalid_argument("batch launch limit");
    p.group_pfx[J] = uint32_t(groups);
    p.tile_pfx[J] = uint32_t(tiles);
    p.pblock_pfx[J] = uint32_t(pblocks);
    p.place_pfx[J] = uint32_t(places);
    p.archive_bytes = arc;
    (void)steps;
    build_steps(sched, J, w, p.fstep, p.fpfx, p.flpfx, p.flarge);
    CompressLayout& L = p.cl;
    uint64_t at = 0;
    auto take = [&](uint64_t bytes) {
        const uint64_t o = at;
        at += (bytes + 255) & ~uint64_t(255);
        return o;
    };
    const uint64_t bank = p.bank_elems * w;
    L.home = take(bank);
    L.tmp0 = take(std::max(bank, p.mag_elems * 2));
    L.tmp1 = take(std::max(bank, p.sign_words * 4));
    L.recon = take(p.out_elems * w);
    L.staging = take(staging + 64);
    L.rbm = take(words * 4);
    L.qidx = take(qlist * 4);
    L.qval = take(qlist * 4);
    L.ridx = take(tiles * kRepairTile * 4);
    L.rval = take(tiles * kRepairTile * w);
    L.rcount = take(tiles * 4);
    L.prefix = take(prefix * 4);
    L.sizes = take(uint64_t(p.chunk_pfx[J]) * 4);
    L.zero0 = at;  // zero-initialised from here on
    L.qbm = take(words * 4);
    L.desc = take(groups * 8);
    L.qcount = take(uint64_t(J) * 4);
    L.ticket = take(4);
    L.res = take(uint64_t(J) * sizeof(EncodeResult));
    L.zero_end = at;
    L.total = at;
}

// Finishes a plan whose shapes (and output placement) are set: bank offsets,
// block prefixes and the per-step pass tables.
void finish_plan(HostPlan& p) {
    const uint32_t J = p.jobs;
    p.chunk_pfx.assign(J + 1, 0);
    p.brick_pfx.assign(J + 1, 0);
    p.pair_pfx.assign(J + 1, 0);
    p.rank_pfx.assign(J + 1, 0);
    p.mag_pfx.assign(J + 1, 0);
    std::vector<std::vector<Pass>> sched(J);
    size_t steps = 0;
    uint64_t bank = 0, mag = 0, sign = 0, chunks = 0, bricks = 0, pchunks = 0, rchunks = 0, mchunks = 0;
    p.rank_geo.clear();
    for (uint32_t j = 0; j < J; ++j) {
        JobShape& s = p.shape[j];
        uint32_t set = 0;
        while (set < p.rank_geo.size() &&
               !(p.rank_geo[set].first.nx == s.d.nx && p.rank_geo[set].first.ny == s.d.ny &&
                 p.rank_geo[set].first.nz == s.d.nz && p.rank_geo[set].second == s.bz))
            ++set;
        if (set == p.rank_geo.size()) p.rank_geo.push_back({s.d, s.bz});
        s.rank = set;
        s.bank = bank;
        s.mag = mag;
        s.sign = sign;
        bank += (uint64_t(s.n) + 3) & ~uint64_t(3);
        mag += (uint64_t(s.n) + 7) & ~uint64_t(7);
        sign += (uint64_t(s.n) + 31) / 32;
        p.chunk_pfx[j] = uint32_t(chunks);
        p.brick_pfx[j] = uint32_t(bricks);
        p.pair_pfx[j] = uint32_t(pchunks);
        p.rank_pfx[j] = uint32_t(rchunks);
        p.mag_pfx[j] = uint32_t(mchunks);
        chunks += s.chunks;
        // full-brick jobs (de)quantize inside the WZP coder: no bricks in the
        // separate passes
        s.coder = pair_bricks_fit(s.d.nx, s.d.ny, s.d.nz, s.bz)   ? kCoderPairs
                  : rank_bricks_fit(s.d.nx, s.d.ny, s.d.nz, s.bz) ? kCoderRanks
                                                                  : kCoderSeparate;
        if (s.coder == kCoderPairs) {
            pchunks += s.chunks;
        } else if (s.coder == kCoderRanks) {
            rchunks += s.chunks;
        } else {
            mchunks += s.chunks;
            bricks += (uint64_t(s.nbx) * s.nby * s.nbz + kBricksPerBlock - 1) / kBricksPerBlock;  // brick blocks
        }
        sched[j] = inverse_schedule(s, p.width);
        steps = std::max(steps, sched[j].size());
    }
    if (chunks > 0x7FFFFFFFull || bricks > 0x7FFFFFFFull) throw std::invalid_argument("batch launch limit");
    p.chunk_pfx[J] = uint32_t(chunks);
    p.brick_pfx[J] = uint32_t(bricks);
    p.pair_pfx[J] = uint32_t(pchunks);
    p.rank_pfx[J] = uint32_t(rchunks);
    p.mag_pfx[J] = uint32_t(mchunks);
    p.bank_elems = bank;
    p.mag_elems = mag;
    p.sign_words = sign;
    (void)steps;
    build_steps(sched, J, p.width, p.step, p.pfx, p.lpfx, p.large);
    finish_compress_plan(p);
}

// Device copy of a plan: tables in one allocation owned by the plan.
struct DevicePlan {
    HostPlan host;
    uint8_t* mem = nullptr;
    JobShape* shape = nullptr;
    JobRun* run = nullptr;
    uint32_t *chunk_pfx = nullptr, *brick_pfx = nullptr, *np_pfx = nullptr, *nq_pfx = nullptr, *status = nullptr;
    uint32_t *pair_pfx = nullptr, *rank_pfx = nullptr, *mag_pfx = nullptr;
    Pass* pass = nullptr;        // steps x jobs
    uint32_t* pass_pfx = nullptr;  // steps x (jobs + 1)
    uint32_t* host_status = nullptr;  // pinned
    // compress tables
    JobComp* comp = nullptr;
    uint32_t *group_pfx = nullptr, *tile_pfx = nullptr, *pblock_pfx = nullptr, *place_pfx = nullptr;
    Pass* fpass = nullptr;          // forward steps x jobs
    uint32_t* fpass_pfx = nullptr;  // forward steps x (jobs + 1)
    uint32_t* lds_pfx = nullptr;    // inverse steps x (jobs + 1), LDS kernel
    uint32_t* flds_pfx = nullptr;   // forward steps x (jobs + 1), LDS kernel
    EncodeResult* last_res = nullptr;  // results of the last compress (device, in its workspace)
    uint16_t* ranks = nullptr;         // kRankSet entries per distinct geometry
    // compress: the WZP group pass runs on `side` (default priority) beside
    // the inverse passes, between the two events
    hipStream_t side = nullptr;
    hipEvent_t pack_ready = nullptr, pack_done = nullptr;
    std::string error;

    ~DevicePlan() {
        if (side) (void)hipStreamSynchronize(side);
        if (pack_ready) (void)hipEventDestroy(pack_ready);
        if (pack_done) (void)hipEventDestroy(pack_done);
        if (side) (void)hipStreamDestroy(side);
        if (mem) (void)hipFree(mem);
        if (ranks) (void)hipFree(ranks);
        if (host_status) (void)hipHostFree(host_status);
    }

    void upload(hipStream_t s) {
        const HostPlan& p = host;
        const uint32_t J = p.jobs;
        const size_t S = p.step.size();
        size_t used = 0;
        auto carve = [&](size_t bytes) {
            const size_t at = used;
            used += (bytes + 255) & ~size_t(255);
            return at;
        };
        const size_t o_shape = carve(sizeof(JobShape) * J), o_run = carve(sizeof(JobRun) * J);
        const size_t o_pair = carve(4 * (J + 1)), o_rankp = carve(4 * (J + 1)), o_magp = carve(4 * (J + 1));
        const size_t o_cp = carve(4 * (J + 1)), o_bp = carve(4 * (J + 1)), o_np = carve(4 * (J + 1)),
                     o_nq = carve(4 * (J + 1)), o_st = carve(4), o_pass = carve(sizeof(Pass) * J * S),
                     o_pp = carve(4 * (J + 1) * S);
        const size_t FS = p.fstep.size();
        const size_t o_comp = carve(sizeof(JobComp) * J), o_gp = carve(4 * (J + 1)), o_tp = carve(4 * (J + 1)),
                     o_pbp = carve(4 * (J + 1)), o_plp = carve(4 * (J + 1)), o_fpass = carve(sizeof(Pass) * J * FS),
                     o_fpp = carve(4 * (J + 1) * FS), o_lp = carve(4 * (J + 1) * S), o_flp = carve(4 * (J + 1) * FS);
        hip_check(hipMalloc(reinterpret_cast<void**>(&mem), used), "batch plan allocation");
        hip_check(hipHostMalloc(reinterpret_cast<void**>(&host_status), 4), "batch status");
        hip_check(hipStreamCreateWithFlags(&side, hipStreamNonBlocking), "batch side stream");
        hip_check(hipEventCreateWithFlags(&pack_ready, hipEventDisableTiming), "event");
        hip_check(hipEventCreateWithFlags(&pack_done, hipEventDisableTiming), "event");
        shape = reinterpret_cast<JobShape*>(mem + o_shape);
        run = reinterpret_cast<JobRun*>(mem + o_run);
        chunk_pfx = reinterpret_cast<uint32_t*>(mem + o_cp);
        brick_pfx = reinterpret_cast<uint32_t*>(mem + o_bp);
        pair_pfx = reinterpret_cast<uint32_t*>(mem + o_pair);
        rank_pfx = reinterpret_cast<uint32_t*>(mem + o_rankp);
        mag_pfx = reinterpret_cast<uint32_t*>(mem + o_magp);
        np_pfx = reinterpret_cast<uint32_t*>(mem + o_np);
        nq_pfx = reinterpret_cast<uint32_t*>(mem + o_nq);
        status = reinterpret_cast<uint32_t*>(mem + o_st);
        pass = reinterpret_cast<Pass*>(mem + o_pass);
        pass_pfx = reinterpret_cast<uint32_t*>(mem + o_pp);
        comp = reinterpret_cast<JobComp*>(mem + o_comp);
        group_pfx = reinterpret_cast<uint32_t*>(mem + o_gp);
        tile_pfx = reinterpret_cast<uint32_t*>(mem + o_tp);
        pblock_pfx = reinterpret_cast<uint32_t*>(mem + o_pbp);
        place_pfx = reinterpret_cast<uint32_t*>(mem + o_plp);
        fpass = reinterpret_cast<Pass*>(mem + o_fpass);
        fpass_pfx = reinterpret_cast<uint32_t*>(mem + o_fpp);
        lds_pfx = reinterpret_cast<uint32_t*>(mem + o_lp);
        flds_pfx = reinterpret_cast<uint32_t*>(mem + o_flp);
        std::vector<Pass> flat;
        std::vector<uint32_t> flat_pfx, flat_lpfx;
        for (size_t st = 0; st < S; ++st) {
            flat.insert(flat.end(), p.step[st].begin(), p.step[st].end());
            flat_pfx.insert(flat_pfx.end(), p.pfx[st].begin(), p.pfx[st].end());
            flat_lpfx.insert(flat_lpfx.end(), p.lpfx[st].begin(), p.lpfx[st].end());
        }
        hip_check(hipMemcpyAsync(shape, p.shape.data(), sizeof(JobShape) * J, hipMemcpyHostToDevice, s), "plan");
        hip_check(hipMemcpyAsync(chunk_pfx, p.chunk_pfx.data(), 4 * (J + 1), hipMemcpyHostToDevice, s), "plan");
        hip_check(hipMemcpyAsync(brick_pfx, p.brick_pfx.data(), 4 * (J + 1), hipMemcpyHostToDevice, s), "plan");
        hip_check(hipMemcpyAsync(pair_pfx, p.pair_pfx.data(), 4 * (J + 1), hipMemcpyHostToDevice, s), "plan");
        hip_check(hipMemcpyAsync(rank_pfx, p.rank_pfx.data(), 4 * (J + 1), hipMemcpyHostToDevice, s), "plan");
        hip_check(hipMemcpyAsync(mag_pfx, p.mag_pfx.data(), 4 * (J + 1), hipMemcpyHostToDevice, s), "plan");
        if (S) {
            hip_check(hipMemcpyAsync(pass, flat.data(), sizeof(Pass) * flat.size(), hipMemcpyHostToDevice, s), "plan");
            hip_check(hipMemcpyAsync(pass_pfx, flat_pfx.data(), 4 * flat_pfx.size(), hipMemcpyHostToDevice, s),
                      "plan");
            hip_check(hipMemcpyAsync(lds_pfx, flat_lpfx.data(), 4 * flat_lpfx.size(), hipMemcpyHostToDevice, s),
                      "plan");
        }
        std::vector<Pass> fflat;
        std::vector<uint32_t> fflat_pfx, fflat_lpfx;
        for (size_t st = 0; st < FS; ++st) {
            fflat.insert(fflat.end(), p.fstep[st].begin(), p.fstep[st].end());
            fflat_pfx.insert(fflat_pfx.end(), p.fpfx[st].begin(), p.fpfx[st].end());
            fflat_lpfx.insert(fflat_lpfx.end(), p.flpfx[st].begin(), p.flpfx[st].end());
        }
        hip_check(hipMemcpyAsync(comp, p.comp.data(), sizeof(JobComp) * J, hipMemcpyHostToDevice, s), "plan");
        hip_check(hipMemcpyAsync(group_pfx, p.group_pfx.data(), 4 * (J + 1), hipMemcpyHostToDevice, s), "plan");
        hip_check(hipMemcpyAsync(tile_pfx, p.tile_pfx.data(), 4 * (J + 1), hipMemcpyHostToDevice, s), "plan");
        hip_check(hipMemcpyAsync(pblock_pfx, p.pblock_pfx.data(), 4 * (J + 1), hipMemcpyHostToDevice, s), "plan");
        hip_check(hipMemcpyAsync(place_pfx, p.place_pfx.data(), 4 * (J + 1), hipMemcpyHostToDevice, s), "plan");
        if (FS) {
            hip_check(hipMemcpyAsync(fpass, fflat.data(), sizeof(Pass) * fflat.size(), hipMemcpyHostToDevice, s),
                      "plan");
            hip_check(hipMemcpyAsync(fpass_pfx, fflat_pfx.data(), 4 * fflat_pfx.size(), hipMemcpyHostToDevice, s),
                      "plan");
            hip_check(hipMemcpyAsync(flds_pfx, fflat_lpfx.data(), 4 * fflat_lpfx.size(), hipMemcpyHostToDevice, s),
                      "plan");
        }
        hip_check(hipMalloc(reinterpret_cast<void**>(&ranks), std::max<size_t>(1, p.rank_geo.size()) * kRankSet * 2),
                  "batch rank tables");
        for (size_t k = 0; k < p.rank_geo.size(); ++k)
            build_rank_tables(p.rank_geo[k].first, p.rank_geo[k].second, ranks + k * kRankSet,
                              ranks + k * kRankSet + 8 * 8192, s);
        hip_check(hipStreamSynchronize(s), "plan upload");
    }

    // The LDS-staged passes of one step (blocks workgroups), small or large variant.
    template <class T, bool Inverse>
    void launch_lds(uint32_t blocks, bool large_lds, const JobRun* r, const Pass* ps, const uint32_t* pf,
                    const StageBanks& sb, T fill, uint32_t* st, hipStream_t s) const {
        if (!blocks) return;
        if (large_lds)
            hipLaunchKernelGGL((batch_lds_kernel<T, true, Inverse>), dim3(blocks), dim3(kLdsThreads), 0, s, shape, r,
                               ps, pf, host.jobs, sb, fill, st);
        else
            hipLaunchKernelGGL((batch_lds_kernel<T, false, Inverse>), dim3(blocks), dim3(kLdsThreads), 0, s, shape,
                               r, ps, pf, host.jobs, sb, fill, st);
    }

    // Enqueues the batched compress on `s`: no synchronisation, allocation or
    // memset (graph-capturable). `ws` holds host.cl.total bytes.
    template <class T>
    void enqueue_compress(const T* input, double bound, uint8_t* out, uint64_t* spans, uint8_t* ws,
                          hipStream_t s) {
        const HostPlan& p = host;
        const uint32_t J = p.jobs;
        const CompressLayout& L = p.cl;
        const double quantum = 1.5 * bound;
        CompBufs b{};
        b.input = input;
        b.home = ws + L.home;
        b.tmp0 = ws + L.tmp0;
        b.tmp1 = ws + L.tmp1;
        b.recon = ws + L.recon;
        b.staging = ws + L.staging;
        b.qbm = reinterpret_cast<uint32_t*>(ws + L.qbm);
        b.rbm = reinterpret_cast<uint32_t*>(ws + L.rbm);
        b.qidx = reinterpret_cast<uint32_t*>(ws + L.qidx);
        b.qval = reinterpret_cast<uint32_t*>(ws + L.qval);
        b.ridx = reinterpret_cast<uint32_t*>(ws + L.ridx);
        b.rval = ws + L.rval;
        b.rcount = reinterpret_cast<uint32_t*>(ws + L.rcount);
        b.qcount = reinterpret_cast<uint32_t*>(ws + L.qcount);
        b.prefix = reinterpret_cast<uint32_t*>(ws + L.prefix);
        b.sizes = reinterpret_cast<uint32_t*>(ws + L.sizes);
        b.ticket = reinterpret_cast<uint32_t*>(ws + L.ticket);
        b.desc = reinterpret_cast<unsigned long long*>(ws + L.desc);
        b.res = reinterpret_cast<EncodeResult*>(ws + L.res);
        b.out = out;
        b.spans = spans;
        last_res = b.res;
        zero_words(ws + L.zero0, (L.zero_end - L.zero0) / 4, s);
        zero_words(run, uint64_t(J) * sizeof(JobRun) / 4, s);  // every job decodable for the inverse passes
        const StageBanks fsb{input, b.home, b.tmp0, b.tmp1, nullptr, b.res};
        for (size_t st = 0; st < p.fstep.size(); ++st) {
            const uint32_t blocks = p.fpfx[st][J];
            if (blocks)
                hipLaunchKernelGGL((batch_fwd_pass_kernel<T>), dim3(blocks), dim3(256), 0, s, shape,
                                   fpass + st * J, fpass_pfx + st * (J + 1), J, b);
            launch_lds<T, false>(p.flpfx[st][J], p.flarge[st], nullptr, fpass + st * J, flds_pfx + st * (J + 1), fsb,
                                 T(0), nullptr, s);
        }
        zero_words(b.tmp1, p.sign_words, s);  // sign words: bricks OR their edge words
        if (p.brick_pfx[J])
            hipLaunchKernelGGL((batch_quantize_kernel<T>), dim3(p.brick_pfx[J]), dim3(256), 0, s, shape, comp,
                               brick_pfx, J, ranks, quantum, b);
        if (p.mag_pfx[J])
            hipLaunchKernelGGL(batch_encode_kernel, dim3(p.mag_pfx[J]), dim3(256), 0, s, shape, comp, mag_pfx, J, b);
        if (p.pair_pfx[J])
            hipLaunchKernelGGL((batch_encode_pairs_kernel<T>), dim3(p.pair_pfx[J]), dim3(256), 0, s, shape, comp,
                               pair_pfx, J, quantum, b);
        if (p.rank_pfx[J])
            hipLaunchKernelGGL((batch_encode_ranks_kernel<T>), dim3(p.rank_pfx[J]), dim3(256), 0, s, shape, comp,
                               rank_pfx, J, ranks, quantum, b);
        // the group pass beside the inverse passes and the repair pass (none of
        // them reads or writes its staging slots, descriptors or archive
        // payload); the layout waits for it
        const bool pack = p.group_pfx[J] != 0;
        if (pack) {
            hip_check(hipEventRecord(pack_ready, s), "event record");
            hip_check(hipStreamWaitEvent(side, pack_ready, 0), "stream wait");
            hipLaunchKernelGGL(batch_pack_kernel, dim3(p.group_pfx[J]), dim3(kPackThreads), 0, side, shape, comp,
                               group_pfx, J, b);
            hip_check(hipEventRecord(pack_done, side), "event record");
        }
        const Banks bk{b.home, b.tmp0, b.tmp1, b.recon};
        const StageBanks isb{nullptr, b.home, b.tmp0, b.tmp1, b.recon, nullptr};
        for (size_t st = 0; st < p.step.size(); ++st) {
            const uint32_t blocks = p.pfx[st][J];
            if (blocks)
                hipLaunchKernelGGL((batch_pass_kernel<T>), dim3(blocks), dim3(256), 0, s, shape, run, pass + st * J,
                                   pass_pfx + st * (J + 1), J, bk, T(0), status);
            launch_lds<T, true>(p.lpfx[st][J], p.large[st], run, pass + st * J, lds_pfx + st * (J + 1), isb, T(0),
                                status, s);
        }
        if (p.tile_pfx[J])
            hipLaunchKernelGGL((batch_repair_kernel<T>), dim3(p.tile_pfx[J]), dim3(256), 0, s, shape, comp, tile_pfx,
                               J, bound, b);
        if (p.pblock_pfx[J])
            hipLaunchKernelGGL(batch_prefix_kernel, dim3(2 * p.pblock_pfx[J]), dim3(256), 0, s, shape, comp,
                               pblock_pfx, J, b);
        if (pack) hip_check(hipStreamWaitEvent(s, pack_done, 0), "stream wait");
        hipLaunchKernelGGL((batch_layout_kernel<T>), dim3(J), dim3(1024), 0, s, shape, comp, quantum, b);
        hipLaunchKernelGGL((batch_place_kernel<T>), dim3(p.place_pfx[J]), dim3(256), 0, s, shape, comp, place_pfx,
                           J, b);
    }

    // Per-job results of the last compress (synchronises `s`).
    std::vector<EncodeResult> compress_results(hipStream_t s) const {
        hip_check(hipGetLastError(), "batch launch");
        std::vector<EncodeResult> r(host.jobs);
        if (!last_res) throw std::runtime_error("no batched compress enqueued");
        hip_check(hipMemcpyAsync(r.data(), last_res, sizeof(EncodeResult) * host.jobs, hipMemcpyDeviceToHost, s),
                  "batch results");
        hip_check(hipStreamSynchronize(s), "batch compress");
        return r;
    }

    // Enqueues the whole decode on `s`: no synchronisation, allocation or memset.
    template <class T>
    void enqueue(const uint8_t* packed, uint64_t packed_bytes, const uint64_t* spans, Banks bk, double fill,
                 hipStream_t s) const {
        const HostPlan& p = host;
        const uint32_t J = p.jobs;
        zero_words(status, 1, s);
        hipLaunchKernelGGL(batch_plan_kernel, dim3(1), dim3(1024), 0, s, packed, packed_bytes, spans, shape, run, J,
                           p.width, np_pfx, nq_pfx, status);
        if (p.mag_pfx[J])
            hipLaunchKernelGGL(batch_wzp_kernel, dim3(p.mag_pfx[J]), dim3(256), 0, s, packed, shape, run, mag_pfx, J,
                               static_cast<uint16_t*>(bk.tmp0), static_cast<uint32_t*>(bk.tmp1), status);
        if (p.pair_pfx[J])
            hipLaunchKernelGGL((batch_wzp_pairs_kernel<T>), dim3(p.pair_pfx[J]), dim3(256), 0, s, packed, shape, run,
                               pair_pfx, J, static_cast<T*>(bk.home), status);
        if (p.rank_pfx[J])
            hipLaunchKernelGGL((batch_wzp_ranks_kernel<T>), dim3(p.rank_pfx[J]), dim3(256), 0, s, packed, shape, run,
                               rank_pfx, J, ranks, static_cast<T*>(bk.home), status);
        if (p.brick_pfx[J])
            hipLaunchKernelGGL((batch_dequant_kernel<T>), dim3(p.brick_pfx[J]), dim3(256), 0, s, shape, run,
                               brick_pfx, J, ranks, static_cast<const uint16_t*>(bk.tmp0),
                               static_cast<const uint32_t*>(bk.tmp1), static_cast<T*>(bk.home));
        hipLaunchKernelGGL((batch_qpatch_kernel<T>), dim3(256), dim3(256), 0, s, packed, shape, run, nq_pfx, J,
                           static_cast<T*>(bk.home), status);
        const StageBanks sb{nullptr, bk.home, bk.tmp0, bk.tmp1, bk.out, nullptr};
        for (size_t st = 0; st < p.step.size(); ++st) {
            const uint32_t blocks = p.pfx[st][J];
            if (blocks)
                hipLaunchKernelGGL((batch_pass_kernel<T>), dim3(blocks), dim3(256), 0, s, shape, run,
                                   pass + st * J, pass_pfx + st * (J + 1), J, bk, T(fill), status);
            launch_lds<T, true>(p.lpfx[st][J], p.large[st], run, pass + st * J, lds_pfx + st * (J + 1), sb, T(fill),
                                status, s);
        }
        hipLaunchKernelGGL((batch_rpatch_kernel<T>), dim3(kRpatchBlocks), dim3(256), 0, s, packed, shape, run, np_pfx, J,
                           static_cast<T*>(bk.out), status);
    }

    void check(hipStream_t s) {
        hip_check(hipGetLastError(), "batch launch");
        hip_check(hipMemcpyAsync(host_status, status, 4, hipMemcpyDeviceToHost, s), "batch status");
        hip_check(hipStreamSynchronize(s), "batch decode");
        const uint32_t st = *host_status;
        if (st & kCorrupt) throw std::runtime_error("corrupt Waltz archive in batch");
        if (st & kNonFiniteOutput) throw std::runtime_error("nonfinite reconstruction");
        if (st) throw std::runtime_error("Waltz device status error");
    }

    uint64_t bank_bytes() const { return host.bank_elems * host.width; }
};

thread_local std::string g_batch_error;

void need(bool ok, const char* what) {
    if (!ok) throw std::runtime_error(what);
}

// ------------------------------------------------------------------ reference ABI
struct FrameGeometry {
    uint32_t gx, gy, gz, jobs;
    uint64_t n, frame;
};

FrameGeometry frame_geometry(const waltz_batch_shape& d) {
    need(d.x && d.y && d.z && d.fx && d.fy && d.fz && d.samples && (d.sides == 1 || d.sides == 2),
         "invalid batch geometry");
    need(d.width == 4 || d.width == 8, "width must be 4 or 8");
    FrameGeometry g;
    g.gx = (d.fx + d.x - 1) / d.x;
    g.gy = (d.fy + d.y - 1) / d.y;
    g.gz = (d.fz + d.z - 1) / d.z;
    const uint64_t jobs = uint64_t(g.gx) * g.gy * g.gz * d.samples * d.sides;
    g.n = uint64_t(d.x) * d.y * d.z;
    need(g.n * d.width <= 0x7FFFFFFFull && jobs <= 65535, "batch launch limit");
    g.jobs = uint32_t(jobs);
    g.frame = uint64_t(d.fx) * d.fy * d.fz;
    return g;
}

// Tiles of one shape in (sample, side, tile z, tile y, tile x) job order,
// placed in (side, sample, z, y, x) frames with clipping.
HostPlan frame_plan(const waltz_batch_shape& d, const FrameGeometry& g) {
    HostPlan p;
    p.width = d.width;
    p.jobs = g.jobs;
    const JobShape proto = job_shape(Dims{d.x, d.y, d.z}, d.width);
    p.shape.assign(g.jobs, proto);
    for (uint32_t j = 0; j < g.jobs; ++j) {
        uint32_t r = j;
        const uint32_t ox = (r % g.gx) * d.x;
        r /= g.gx;
        const uint32_t oy = (r % g.gy) * d.y;
        r /= g.gy;
        const uint32_t oz = (r % g.gz) * d.z;
        r /= g.gz;
        const uint64_t base = (uint64_t(r % d.sides) * d.samples + r / d.sides) * g.frame;
        JobShape& s = p.shape[j];
        s.out = base + (uint64_t(oz) * d.fy + oy) * d.fx + ox;
        s.opy = d.fx;
        s.opz = uint64_t(d.fx) * d.fy;
        s.clip = Dims{std::min(d.x, d.fx - ox), std::min(d.y, d.fy - oy), std::min(d.z, d.fz - oz)};
    }
    p.out_elems = g.frame * d.samples * d.sides;
    finish_plan(p);
    return p;
}

template <class T>
void run_frames(const waltz_batch_shape& d, const uint8_t* host, const uint8_t* packed, size_t bytes, void* scratch,
                size_t scratch_bytes, T* out, size_t out_bytes, double fill, hipStream_t s) {
    const FrameGeometry g = frame_geometry(d);
    need(out_bytes == g.frame * d.samples * d.sides * sizeof(T), "incorrect output size");
    need(bytes >= 16ull * g.jobs, "short batch buffers");
    const size_t table = bytes - 16ull * g.jobs;
    need(table % 16 == 0, "unaligned batch spans");
    // host-side validation of every span and header (src/batch_decode.cuh:90-120)
    for (uint32_t j = 0; j < g.jobs; ++j) {
        uint64_t pair[2];
        std::memcpy(pair, host + table + 16ull * j, 16);
        if (!pair[1]) continue;
        need(pair[0] % 16 == 0 && pair[0] <= table && pair[1] >= 32 && pair[1] <= table - pair[0],
             "invalid batch span");
        need(archive_size_from_header(host + pair[0], Dims{d.x, d.y, d.z}, sizeof(T)) == pair[1],
             "invalid canonical batch archive");
    }
    DevicePlan plan;
    plan.host = frame_plan(d, g);
    const uint64_t bank = plan.bank_bytes();
    need(scratch_bytes >= 3 * bank, "short batch buffers");
    plan.upload(s);
    uint8_t* sc = static_cast<uint8_t*>(scratch);
    const Banks bk{sc, sc + bank, sc + 2 * bank, out};
    plan.enqueue<T>(packed, table, reinterpret_cast<const uint64_t*>(packed + table), bk, fill, s);
    plan.check(s);
}
}  // namespace
}  // namespace wz

struct waltz_amd_batch {
    wz::DevicePlan plan;
};

extern "C" {

size_t waltz_gpu_batch_scratch_bytes(const waltz_batch_shape* shape) {
    try {
        const wz::FrameGeometry g = wz::frame_geometry(*shape);
        return size_t(3 * g.jobs * ((g.n + 3) & ~uint64_t(3)) * shape->width);
    } catch (...) {
        return 0;
    }
}

const char* waltz_gpu_decode_batch(const waltz_batch_shape* shape, const void* host, const void* packed, size_t bytes,
                                   void* scratch, size_t scratch_bytes, void* output, size_t output_bytes,
                                   double fill, void* stream) {
    try {
        if (!shape || !host || !packed || !scratch || !output) throw std::invalid_argument("null argument");
        const auto* h = static_cast<const uint8_t*>(host);
        const auto* p = static_cast<const uint8_t*>(packed);
        hipStream_t s = static_cast<hipStream_t>(stream);
        if (shape->width == 4)
            wz::run_frames<float>(*shape, h, p, bytes, scratch, scratch_bytes, static_cast<float*>(output),
                                  output_bytes, fill, s);
        else
            wz::run_frames<double>(*shape, h, p, bytes, scratch, scratch_bytes, static_cast<double*>(output),
                                   output_bytes, fill, s);
        return nullptr;
    } catch (const std::exception& e) {
        wz::g_batch_error = e.what();
    } catch (...) {
        wz::g_batch_error = "unknown batch error";
    }
    return wz::g_batch_error.c_str();
}

waltz_amd_batch* waltz_amd_batch_create(int width, uint32_t jobs, const uint32_t* dims) {
    try {
        if (width != 4 && width != 8) throw std::invalid_argument("width must be 4 or 8");
        if (!jobs || !dims) throw std::invalid_argument("empty batch");
        if (jobs > 1000000) throw std::invalid_argument("batch launch limit");
        auto b = std::make_unique<waltz_amd_batch>();
        wz::HostPlan& p = b->plan.host;
        p.width = unsigned(width);
        p.jobs = jobs;
        p.shape.reserve(jobs);
        uint64_t out = 0;
        for (uint32_t j = 0; j < jobs; ++j) {
            wz::JobShape s = wz::job_shape(wz::Dims{dims[3 * j], dims[3 * j + 1], dims[3 * j + 2]}, unsigned(width));
            s.out = out;
            s.opy = s.d.nx;
            s.opz = uint64_t(s.d.nx) * s.d.ny;
            out += s.n;
            p.shape.push_back(s);
        }
        p.out_elems = out;
        wz::finish_plan(p);
        b->plan.upload(nullptr);
        return b.release();
    } catch (const std::exception& e) {
        wz::g_batch_error = e.what();
    } catch (...) {
        wz::g_batch_error = "unknown batch error";
    }
    return nullptr;
}

const char* waltz_amd_batch_error(void) { return wz::g_batch_error.c_str(); }

void waltz_amd_batch_destroy(waltz_amd_batch* b) { delete b; }

size_t waltz_amd_batch_workspace_bytes(const waltz_amd_batch* b) { return b ? size_t(3 * b->plan.bank_bytes()) : 0; }

size_t waltz_amd_batch_compress_workspace_bytes(const waltz_amd_batch* b) {
    return b ? size_t(b->plan.host.cl.total) : 0;
}

size_t waltz_amd_batch_archive_capacity(const waltz_amd_batch* b) { return b ? size_t(b->plan.host.archive_bytes) : 0; }

int waltz_amd_batch_compress(waltz_amd_batch* b, const void* input, double abs_bound, void* archives,
                             size_t archives_bytes, void* spans, void* workspace, size_t workspace_bytes,
                             void* stream) {
    try {
        if (!b || !input || !archives || !spans || !workspace) throw std::invalid_argument("null argument");
        const wz::HostPlan& p = b->plan.host;
        if (!(abs_bound > 0) || !std::isfinite(abs_bound))
            throw std::invalid_argument("positive finite absolute error required");
        const double quantum = 1.5 * abs_bound;
        const bool f32 = p.width == 4;
        const bool q_ok = f32 ? (float(quantum) > 0 && std::isfinite(float(quantum)) && std::isfinite(float(1.0 / quantum)))
                              : (quantum > 0 && std::isfinite(quantum) && std::isfinite(1.0 / quantum));
        if (!q_ok) throw std::invalid_argument("quantum not representable in native type");
        if (archives_bytes < p.archive_bytes) throw std::invalid_argument("archive capacity too small");
        if (workspace_bytes < p.cl.total) throw std::invalid_argument("short batch workspace");
        if (reinterpret_cast<uintptr_t>(archives) % 16 || reinterpret_cast<uintptr_t>(workspace) % 256 ||
            reinterpret_cast<uintptr_t>(input) % p.width || reinterpret_cast<uintptr_t>(spans) % 8)
            throw std::invalid_argument("misaligned batch buffers");
        hipStream_t s = static_cast<hipStream_t>(stream);
        uint8_t* out = static_cast<uint8_t*>(archives);
        uint8_t* ws = static_cast<uint8_t*>(workspace);
        if (f32)
            b->plan.enqueue_compress<float>(static_cast<const float*>(input), abs_bound, out,
                                            static_cast<uint64_t*>(spans), ws, s);
        else
            b->plan.enqueue_compress<double>(static_cast<const double*>(input), abs_bound, out,
                                             static_cast<uint64_t*>(spans), ws, s);
        return 0;
    } catch (const std::exception& e) {
        wz::g_batch_error = e.what();
    } catch (...) {
        wz::g_batch_error = "unknown batch error";
    }
    return -1;
}

int waltz_amd_batch_compress_check(waltz_amd_batch* b, void* stream) {
    try {
        if (!b) throw std::invalid_argument("null batch");
        const std::vector<wz::EncodeResult> r = b->plan.compress_results(static_cast<hipStream_t>(stream));
        for (size_t j = 0; j < r.size(); ++j) {
            const uint32_t st = r[j].status;
            if (!st) continue;
            const char* what = (st & wz::kNonFiniteInput)   ? "nonfinite input is unsupported"
                               : (st & wz::kQuantDomain)    ? "coefficient exceeds portable integer conversion domain"
                               : (st & wz::kRepairBound)    ? "native repair cannot meet strict bound or nonfinite reconstruction"
                               : (st & wz::kSideCapacity)   ? "portable GPU side-channel capacity reached"
                               : (st & wz::kWzpCapacity)    ? "archive capacity too small"
                               : (st & wz::kLookbackTimeout) ? "WZP lookback did not complete"
                                                            : "Waltz device status error";
            throw std::runtime_error("job " + std::to_string(j) + ": " + what);
        }
        return 0;
    } catch (const std::exception& e) {
        wz::g_batch_error = e.what();
    } catch (...) {
        wz::g_batch_error = "unknown batch error";
    }
    return -1;
}

size_t waltz_amd_batch_output_elements(const waltz_amd_batch* b) { return b ? size_t(b->plan.host.out_elems) : 0; }

int waltz_amd_batch_decode(waltz_amd_batch* b, const void* packed, size_t packed_bytes, const void* spans,
                           void* output, void* workspace, size_t workspace_bytes, void* stream) {
    try {
        if (!b || !packed || !spans || !output || !workspace) throw std::invalid_argument("null argument");
        const uint64_t bank = b->plan.bank_bytes();
        if (workspace_bytes < 3 * bank) throw std::invalid_argument("short batch workspace");
        uint8_t* w = static_cast<uint8_t*>(workspace);
        const wz::Banks bk{w, w + bank, w + 2 * bank, output};
        hipStream_t s = static_cast<hipStream_t>(stream);
        if (b->plan.host.width == 4)
            b->plan.enqueue<float>(static_cast<const uint8_t*>(packed), packed_bytes,
                                   static_cast<const uint64_t*>(spans), bk, 0.0, s);
        else
            b->plan.enqueue<double>(static_cast<const uint8_t*>(packed), packed_bytes,
                                    static_cast<const uint64_t*>(spans), bk, 0.0, s);
        return 0;
    } catch (const std::exception& e) {
        wz::g_batch_error = e.what();
    } catch (...) {
        wz::g_batch_error = "unknown batch error";
    }
    return -1;
}

int waltz_amd_batch_check(waltz_amd_batch* b, void* stream) {
    try {
        if (!b) throw std::invalid_argument("null batch");
        b->plan.check(static_cast<hipStream_t>(stream));
        return 0;
    } catch (const std::exception& e) {
        wz::g_batch_error = e.what();
    } catch (...) {
        wz::g_batch_error = "unknown batch error";
    }
    return -1;
}

}  // extern "C"
