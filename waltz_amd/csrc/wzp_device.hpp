// Device pieces of the WZP stream shared by the single-field and batched
// kernels: block scans, byte loads, the decoder's sample sinks and the
// per-chunk decoder (cpu/src/wzp_cpu.cpp:595-949).
#pragma once

#include "bits.hpp"
#include "common.hpp"
#include "launch.hpp"
#include "pairbrick.hpp"

namespace wz {

namespace {
constexpr uint32_t kThreads = 256;

WZ_D uint32_t wave_inclusive(uint32_t v) {
    const uint32_t lane = threadIdx.x & 63u;
#pragma unroll
    for (uint32_t d = 1; d < 64; d <<= 1) {
        const uint32_t t = __shfl_up(v, d, 64);
        if (lane >= d) v += t;
    }
    return v;
}

// Block-wide exclusive scan for 256 threads; `sh` holds >= 4 words.
WZ_D uint32_t block_exclusive(uint32_t v, uint32_t* sh, uint32_t& total) {
    const uint32_t inc = wave_inclusive(v);
    const uint32_t w = threadIdx.x >> 6;
    if ((threadIdx.x & 63u) == 63u) sh[w] = inc;
    __syncthreads();
    uint32_t before = 0, sum = 0;
#pragma unroll
    for (uint32_t i = 0; i < kThreads / 64; ++i) {
        const uint32_t s = sh[i];
        before += i < w ? s : 0u;
        sum += s;
    }
    __syncthreads();
    total = sum;
    return before + inc - v;
}

WZ_D uint32_t load24(const uint8_t* p) { return uint32_t(p[0]) | (uint32_t(p[1]) << 8) | (uint32_t(p[2]) << 16); }
WZ_D uint32_t load32(const uint8_t* p) { return load24(p) | (uint32_t(p[3]) << 24); }

constexpr unsigned long long kFlagAggregate = 1ull << 62;
constexpr unsigned long long kFlagInclusive = 2ull << 62;
constexpr unsigned long long kValueMask = (1ull << 62) - 1;
constexpr uint32_t kSpinLimit = 1u << 26;

WZ_D uint32_t chunk_elements(uint32_t n, uint32_t c) {
    const uint32_t rest = n - c * kChunk;
    return rest < kChunk ? rest : kChunk;
}
}  // namespace

// Four consecutive samples of one row (16-byte aligned in the paired-brick
// geometry: rows start at multiples of 4 samples).
WZ_D void load4(const float* p, float (&v)[4]) {
    const float4 x = *reinterpret_cast<const float4*>(p);
    v[0] = x.x;
    v[1] = x.y;
    v[2] = x.z;
    v[3] = x.w;
}
WZ_D void load4(const double* p, double (&v)[4]) {
    const double2 a = reinterpret_cast<const double2*>(p)[0], b = reinterpret_cast<const double2*>(p)[1];
    v[0] = a.x;
    v[1] = a.y;
    v[2] = b.x;
    v[3] = b.y;
}
WZ_D void store4(float* p, const float (&v)[4]) { *reinterpret_cast<float4*>(p) = make_float4(v[0], v[1], v[2], v[3]); }
WZ_D void store4(double* p, const double (&v)[4]) {
    reinterpret_cast<double2*>(p)[0] = make_double2(v[0], v[1]);
    reinterpret_cast<double2*>(p)[1] = make_double2(v[2], v[3]);
}

// Decoder outputs. store() takes thread t's magnitudes [32t, 32t + 32) of
// chunk c and their sign word; a collective sink (kCollective) is called by
// every thread of the workgroup at once and may use the decoder's staging
// buffer `lds` (kChunk / 2 + kThreads words), whose contents are dead by then.
struct MagSink {  // packed magnitude / sign arrays
    static constexpr bool kCollective = false;
    uint16_t* mag;
    uint32_t* sign;
    WZ_D void store(uint32_t c, uint32_t t, uint32_t E, const uint32_t (&m)[32], uint32_t sw, uint32_t*) const {
        const uint32_t first = t * 32;
        uint16_t* p = mag + uint64_t(c) * kChunk + first;
        if (first < E) sign[uint64_t(c) * (kChunk / 32) + t] = sw;
        if (first + 32 <= E) {
            uint4* v = reinterpret_cast<uint4*>(p);
#pragma unroll
            for (int q = 0; q < 4; ++q)
                v[q] = make_uint4(m[8 * q] | (m[8 * q + 1] << 16), m[8 * q + 2] | (m[8 * q + 3] << 16),
                                  m[8 * q + 4] | (m[8 * q + 5] << 16), m[8 * q + 6] | (m[8 * q + 7] << 16));
        } else {
#pragma unroll
            for (int i = 0; i < 32; ++i)
                if (first + i < E) p[i] = uint16_t(m[i]);
        }
    }
};

// Dequantizes straight into the coefficient volume (paired-brick geometry),
// as dequantize_kernel does per sample (cpu/src/codec.cpp:146-147).
template <class T>
struct BrickSink {
    static constexpr bool kCollective = false;
    T* coeff;
    uint32_t nx, ny;
    T q;
    WZ_D void store(uint32_t c, uint32_t t, uint32_t, const uint32_t (&m)[32], uint32_t sw, uint32_t*) const {
        uint32_t X, Y, Z;
        pair_subblock(nx, ny, c, t, X, Y, Z);
#pragma unroll
        for (uint32_t r = 0; r < 8; ++r) {
            const uint32_t dy = r & 3u, dz = r >> 2;
            T d[4];
#pragma unroll
            for (uint32_t dx = 0; dx < 4; ++dx) {
                const uint32_t i = subblock_rank(dx, dy, dz);
                const T v = T(m[i]) * q;
                d[dx] = ((sw >> i) & 1u) ? -v : v;
            }
            store4(coeff + (uint64_t(Z + dz) * ny + (Y + dy)) * nx + X, d);
        }
    }
};

// The bricks of chunk c in a field of full bricks (rank_bricks_fit): the
// chunk's first packed position, its first brick and their count; thread i <
// count writes the row-major index of brick first + i's origin to kb[i] and,
// when ko is given, the origin's x, y, z to ko[3i .. 3i + 2].
struct RankBricks {
    const uint16_t* rank;  // local row-major index -> rank of a full 16 x 16 x bz brick
    uint32_t nx, ny, bz;
};
struct RankChunk {
    uint32_t p0, first, count;
};
WZ_D RankChunk rank_chunk(const RankBricks& g, uint32_t c, uint32_t E, uint32_t* kb, uint32_t* ko = nullptr) {
    const uint32_t p0 = c * kChunk;
    const uint32_t first = (p0 >> 8) / g.bz, last = ((p0 + E - 1u) >> 8) / g.bz;
    const uint32_t count = last - first + 1u;
    if (threadIdx.x < count) {
        const uint32_t nbx = g.nx >> 4, nby = g.ny >> 4, b = first + threadIdx.x;
        const uint32_t r = b / nbx, bx = b - r * nbx, bzi = r / nby, by = r - bzi * nby;
        kb[threadIdx.x] = (bzi * g.bz * g.ny + 16u * by) * g.nx + 16u * bx;
        if (ko) {
            ko[3 * threadIdx.x] = 16u * bx;
            ko[3 * threadIdx.x + 1] = 16u * by;
            ko[3 * threadIdx.x + 2] = bzi * g.bz;
        }
    }
    return RankChunk{p0, first, count};
}

// Quad f (4 x-consecutive samples) of a chunk's bricks: the row-major index
// of its first sample, its brick (of the chunk's) and index inside the brick,
// and its four packed positions less the chunk's first (those outside
// [0, E) belong to a neighbouring chunk).
struct RankQuad {
    uint32_t at, lb, li;
    uint32_t j[4];
};
WZ_D RankQuad rank_quad(const RankBricks& g, const RankChunk& rc, const FastDiv& by_q, const uint32_t* kb,
                        uint32_t f) {
    const uint32_t q = 64u * g.bz;  // quads per brick
    const uint32_t lb = by_q.div(f), li = 4u * (f - lb * q);
    const uint2 rr = *reinterpret_cast<const uint2*>(g.rank + li);  // ranks of li .. li + 3
    const uint32_t jb = (rc.first + lb) * (256u * g.bz) - rc.p0;    // (wraps below the chunk)
    RankQuad out;
    out.at = kb[lb] + ((li >> 8) * g.ny + ((li >> 4) & 15u)) * g.nx + (li & 15u);
    out.lb = lb;
    out.li = li;
    out.j[0] = jb + (rr.x & 0xFFFFu);
    out.j[1] = jb + (rr.x >> 16);
    out.j[2] = jb + (rr.y & 0xFFFFu);
    out.j[3] = jb + (rr.y >> 16);
    return out;
}

// Dequantizes straight into the coefficient volume of a field of full
// bricks (rank_bricks_fit), as dequantize_kernel does per sample
// (cpu/src/codec.cpp:146-147): the chunk's magnitudes and sign words go to
// LDS in packed order, then the workgroup writes the chunk's bricks row by
// row, each sample read at its rank.
template <class T>
struct RankSink {
    static constexpr bool kCollective = true;
    T* coeff;
    RankBricks g;
    T q;
    WZ_D void store(uint32_t c, uint32_t t, uint32_t E, const uint32_t (&m)[32], uint32_t sw, uint32_t* lds) const {
        __shared__ uint32_t kb[kRankBricksMax];
        const uint16_t* mg = reinterpret_cast<const uint16_t*>(lds);
        uint32_t* const sg = lds + kChunk / 2;
        __syncthreads();  // every thread is done with the staging buffer
        uint4* v = reinterpret_cast<uint4*>(lds) + 4u * t;
#pragma unroll
        for (int k = 0; k < 4; ++k)
            v[k] = make_uint4(m[8 * k] | (m[8 * k + 1] << 16), m[8 * k + 2] | (m[8 * k + 3] << 16),
                              m[8 * k + 4] | (m[8 * k + 5] << 16), m[8 * k + 6] | (m[8 * k + 7] << 16));
        sg[t] = sw;
        const RankChunk rc = rank_chunk(g, c, E, kb);
        __syncthreads();
        const FastDiv by_q(64u * g.bz);  // f * q < 2^32: exact
        const uint32_t nq = rc.count * 64u * g.bz;
        for (uint32_t f = t; f < nq; f += kThreads) {
            const RankQuad rq = rank_quad(g, rc, by_q, kb, f);
            T d[4];
            bool all = true;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const bool in = rq.j[k] < E;
                all = all && in;
                const uint32_t j = in ? rq.j[k] : 0u;
                const T x = T(mg[j]) * q;
                d[k] = ((sg[j >> 5] >> (j & 31u)) & 1u) ? -x : x;
            }
            if (all) {
                store4(coeff + rq.at, d);
            } else {
#pragma unroll
                for (int k = 0; k < 4; ++k)
                    if (rq.j[k] < E) coeff[rq.at + k] = d[k];
            }
        }
    }
};

// v_perm_b32 selector that spreads the low bytes of a word to the byte
// positions whose bits are set in the 4-bit mask m (selector 4 = a zero byte
// of the first operand elsewhere).
WZ_HD uint32_t expansion_selector(uint32_t m) {
    uint32_t sel = 0, k = 0;
    for (uint32_t b = 0; b < 4; ++b) {
        const uint32_t v = ((m >> b) & 1u) ? k++ : 4u;
        sel |= v << (8 * b);
    }
    return sel;
}

// Decodes chunk c of a WZP stream of `bytes` bytes holding n magnitudes;
// called by all 256 threads of a workgroup.
// The 32 magnitudes [32t, 32t + 32) of a chunk of E from its staged BIT_2
// planes (plane p of block b at u16 index p * blocks + b, the tail past the
// last whole block verbatim); zero past E.
WZ_D void bit2_unstage(const uint32_t* tr, uint32_t E, uint32_t t, uint32_t (&m)[32]) {
    const uint16_t* tr16 = reinterpret_cast<const uint16_t*>(tr);
    const uint32_t blocks = E / 16, first = 32 * t;
    // every thread runs the two-block transform (indices stay inside the
    // 8192-entry buffer even for threads past the last whole block) ...
    uint32_t pp[16];
    if (!(blocks & 1u)) {
#pragma unroll
        for (int pl = 0; pl < 16; ++pl) pp[pl] = tr[pl * (blocks >> 1) + t];
    } else {
#pragma unroll
        for (int pl = 0; pl < 16; ++pl)
            pp[pl] = uint32_t(tr16[pl * blocks + 2 * t]) | (uint32_t(tr16[pl * blocks + 2 * t + 1]) << 16);
    }
    bit2_inverse_pair(pp, m);
    // ... and threads whose second block is not whole take the samples past
    // the last whole block verbatim (selects only: m stays in registers)
    if (2 * t + 1 >= blocks) {
        const bool whole0 = 2 * t < blocks;
#pragma unroll
        for (int j = 0; j < 32; ++j) {
            const uint32_t i = first + j;
            const uint32_t raw = i < E ? uint32_t(tr16[i]) : 0u;
            m[j] = (j < 16 && whole0) ? m[j] : raw;
        }
    }
}

template <class Sink>
WZ_D void decode_chunk(const uint8_t* __restrict__ in, uint64_t bytes, uint32_t n, uint32_t chunks, uint32_t groups,
                       uint32_t c, const Sink& sink, uint32_t* __restrict__ status) {
    // the staged chunk payload and, once every lane holds its bytes in
    // registers, the BIT_2 planes: one LDS buffer; the sign bytes are copied
    // aside before it is reused
    __shared__ __attribute__((aligned(16))) uint32_t ib32[(kChunkBytes + kSignSlot) / 4 + 4];
    uint32_t* const tr = ib32;
    __shared__ uint32_t sg32[kSignSlot / 4];  // sign bytes
    __shared__ uint32_t esel[16];           // byte-expansion selectors
    __shared__ uint32_t b1w[kThreads / 4];  // B1 as 64 words (256 bytes)
    __shared__ uint32_t b2w[8];             // B2 as 8 words (32 bytes)
    __shared__ uint32_t sh[8];
    __shared__ uint32_t info[4];            // offset, ok

    const uint32_t t = threadIdx.x;
    const uint32_t E = chunk_elements(n, c);
    const uint32_t first = t * 32;
    const uint64_t header = uint64_t(groups) * 4 + uint64_t(chunks) * 3;
    const uint8_t* meta = in + uint64_t(groups) * 4;
    const uint32_t mc = load24(meta + uint64_t(c) * 3);
    const uint32_t code = mc & kRawCode, sc = mc >> 14;
    const uint32_t raw_bytes = 2 * E;
    const bool raw = code == kRawCode;
    const uint32_t mbytes = code == 0 ? 0u : raw ? raw_bytes : code;
    const uint32_t sbytes = code == 0 ? 0u : sc ? sc : kSignSlot;
    if (t < 16) esel[t] = expansion_selector(t);
    if (t < 64) {
        // the chunk's offset: its group's offset plus the sizes of the
        // group's earlier chunks, one meta per lane of the first wave (all
        // loads in flight at once), summed across the 16 lanes
        const uint32_t g = c / kGroup, j = g * kGroup + t;
        const uint32_t goff = t == 0 ? load32(in + uint64_t(g) * 4) : 0u;
        uint32_t size = 0;
        if (t < kGroup && j < c) {
            const uint32_t mj = load24(meta + uint64_t(j) * 3);
            const uint32_t cj = mj & kRawCode, sj = mj >> 14;
            if (cj) size = (cj == kRawCode ? 2 * chunk_elements(n, j) : cj) + (sj ? sj : kSignSlot);
        }
#pragma unroll
        for (int o = 1; o < int(kGroup); o <<= 1) size += __shfl_xor(size, o);
        if (t == 0) {
            const uint64_t off = uint64_t(goff) + size;
            uint32_t ok = 1;
            if (code == 0 && sc != 0) ok = 0;
            if (code && !raw && (code < 4 || code + 2 >= raw_bytes)) ok = 0;
            if (code && sbytes > (E + 7) / 8) ok = 0;
            if (header + off + mbytes + sbytes > bytes) ok = 0;
            info[0] = uint32_t(off);
            info[1] = ok;
            if (!ok) atomicOr(status, uint32_t(kCorrupt));
        }
    }
    __syncthreads();
    if (!info[1]) return;
    if (code == 0) {
        uint32_t zero[32];
#pragma unroll
        for (int i = 0; i < 32; ++i) zero[i] = 0;
        sink.store(c, t, E, zero, 0u, ib32);
        return;
    }
    // stage the chunk payload in LDS with aligned word loads that never leave
    // the stream [in, in + bytes); six words per thread are in flight per
    // round (a typical chunk takes one or two rounds)
    const uint8_t* const p0 = in + header + info[0];
    const uint32_t mis = uint32_t(reinterpret_cast<uintptr_t>(p0) & 3u);
    const uint32_t* const w0 = reinterpret_cast<const uint32_t*>(p0 - mis);
    const uint32_t nwords = (mis + mbytes + sbytes + 3) / 4;
    constexpr uint32_t kRound = 6;
    if (reinterpret_cast<const uint8_t*>(w0) >= in &&
        reinterpret_cast<const uint8_t*>(w0 + nwords) <= in + bytes) {  // the usual case
        for (uint32_t base = t; base < nwords; base += kRound * kThreads) {
            uint32_t v[kRound];
#pragma unroll
            for (uint32_t k = 0; k < kRound; ++k) {
                const uint32_t i = base + k * kThreads;
                v[k] = i < nwords ? w0[i] : 0u;
            }
#pragma unroll
            for (uint32_t k = 0; k < kRound; ++k) {
                const uint32_t i = base + k * kThreads;
                if (i < nwords) ib32[i] = v[k];
            }
        }
    } else {  // a word at an end of the stream: bytes inside it only
        for (uint32_t i = t; i < nwords; i += kThreads) {
            const uint8_t* a = reinterpret_cast<const uint8_t*>(w0 + i);
            uint32_t v = 0;
            for (uint32_t k = 0; k < 4; ++k)
                if (a + k >= in && a + k < in + bytes) v |= uint32_t(a[k]) << (8 * k);
            ib32[i] = v;
        }
    }
    __syncthreads();
    const uint8_t* src = reinterpret_cast<const uint8_t*>(ib32) + mis;
    uint8_t* const ps = reinterpret_cast<uint8_t*>(sg32);
    for (uint32_t i = t; i < sbytes; i += kThreads) ps[i] = src[mbytes + i];
    uint32_t m[32];
    bool bad = false;
    if (raw) {
#pragma unroll
        for (int i = 0; i < 32; ++i)
            m[i] = first + i < E ? uint32_t(src[2 * (first + i)]) | (uint32_t(src[2 * (first + i) + 1]) << 8) : 0u;
    } else {
        // top-down: B3 -> B2 -> B1 -> B0 -> bytes, each level's changed bytes
        // located by popcount ranks (cpu/src/wzp_cpu.cpp:595-660)
        const uint32_t b3 = load32(src + code - 4);
        const uint32_t K3 = popc32(b3);
        const int32_t pos3 = int32_t(code) - 4 - int32_t(K3);
        if (pos3 < 0) bad = true;
        if (!bad && t < 32) {
            const uint32_t r = popc32(t == 31 ? b3 : b3 & ((2u << t) - 1u));
            const uint32_t v = r ? uint32_t(src[pos3 + int32_t(r) - 1]) : 0u;
            reinterpret_cast<uint8_t*>(b2w)[t] = uint8_t(v);
        }
        __syncthreads();
        uint32_t K2 = 0;
#pragma unroll
        for (int i = 0; i < 8; ++i) K2 += popc32(b2w[i]);
        const int32_t pos2 = pos3 - int32_t(K2);
        if (pos2 < 0) bad = true;
        if (!bad) {
            const uint32_t wd = t >> 5, bit = t & 31u;
            uint32_t r = popc32(bit == 31 ? b2w[wd] : b2w[wd] & ((2u << bit) - 1u));
            for (uint32_t i = 0; i < wd; ++i) r += popc32(b2w[i]);
            const uint32_t v = r ? uint32_t(src[pos2 + int32_t(r) - 1]) : 0u;
            reinterpret_cast<uint8_t*>(b1w)[t] = uint8_t(v);
        }
        __syncthreads();
        const uint32_t b1t = reinterpret_cast<const uint8_t*>(b1w)[t];
        uint32_t K1;
        const uint32_t e1 = block_exclusive(popc32(b1t), sh, K1);
        const int32_t pos1 = pos2 - int32_t(K1);
        if (pos1 < 0) bad = true;
        unsigned long long b0t = 0;
        if (!bad) {
            uint32_t r = e1;
            uint32_t cur = r ? uint32_t(src[pos1 + int32_t(r) - 1]) : 0u;
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                if ((b1t >> k) & 1u) cur = src[pos1 + int32_t(r++)];
                b0t |= (unsigned long long)(cur & 0xFFu) << (8 * k);
            }
        }
        {
            const int32_t valid = int32_t(raw_bytes) - int32_t(64 * t);
            const unsigned long long keep = valid <= 0 ? 0ull : valid >= 64 ? ~0ull : (1ull << valid) - 1ull;
            if (b0t & ~keep) bad = true;
        }
        uint32_t K0;
        const uint32_t e0 = block_exclusive(popc64(b0t), sh, K0);
        if (pos1 - int32_t(K0) != 0) bad = true;
        uint32_t tw[16];
        if (!bad) {
            // this lane's nonzero bytes src[e0, e0 + popcount(b0t)), spread
            // word by word with byte permutes from an 8-byte window of
            // aligned loads (never past the staging buffer)
            const uint32_t byte0 = uint32_t(src - reinterpret_cast<const uint8_t*>(ib32)) + e0;
            uint32_t wi = byte0 >> 2;
            const uint32_t sh0 = byte0 & 3u;
            unsigned long long acc = (((unsigned long long)ib32[wi + 1] << 32) | ib32[wi]) >> (8 * sh0);
            uint32_t navail = 8 - sh0;
            wi += 2;
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                if (navail < 4) {
                    acc |= (unsigned long long)ib32[wi++] << (8 * navail);
                    navail += 4;
                }
                const uint32_t m = uint32_t(b0t >> (4 * q)) & 15u;
                tw[q] = __builtin_amdgcn_perm(0u, uint32_t(acc), esel[m]);
                const uint32_t c = popc32(m);
                acc >>= 8 * c;
                navail -= c;
            }
        }
        __syncthreads();  // every lane's bytes are in registers: the buffer takes the planes
        if (!bad) {
#pragma unroll
            for (int q = 0; q < 4; ++q)
                reinterpret_cast<uint4*>(tr)[4 * t + q] = make_uint4(tw[4 * q], tw[4 * q + 1], tw[4 * q + 2],
                                                                      tw[4 * q + 3]);
        }
        __syncthreads();
        bit2_unstage(tr, E, t, m);
    }
    uint32_t nzmask = 0;
#pragma unroll
    for (int i = 0; i < 32; ++i) nzmask |= uint32_t(m[i] != 0) << i;
    uint32_t NNZ;
    const uint32_t es = block_exclusive(popc32(nzmask), sh, NNZ);
    if (NNZ == 0 || (NNZ + 7) / 8 != sbytes) bad = true;
    // (a lane's zero bitmap past the chunk end is its own check: one verdict
    // for the workgroup before a collective sink's barriers)
    if constexpr (Sink::kCollective) bad = __syncthreads_or(bad) != 0;
    if (bad) {
        if (t == 0 || (threadIdx.x & 63u) == 0) atomicOr(status, uint32_t(kCorrupt));
        return;
    }
    const uint32_t cnt = popc32(nzmask);
    uint32_t bits = 0;
    if (cnt) {
        const uint32_t b = es >> 3, sh8 = es & 7u;
        const uint32_t need = (sh8 + cnt + 7) >> 3;  // <= 5 bytes
        unsigned long long v = 0;
        for (uint32_t i = 0; i < need; ++i) v |= (unsigned long long)ps[b + i] << (8 * i);
        bits = uint32_t(v >> sh8) & (cnt == 32 ? 0xFFFFFFFFu : ((1u << cnt) - 1u));
    }
    if (t == 0 && (NNZ & 7u) && (ps[sbytes - 1] >> (NNZ & 7u))) atomicOr(status, uint32_t(kCorrupt));
    sink.store(c, t, E, m, expand_bits(bits, nzmask), ib32);
}

// Groups of the largest stream (n < 2^31 samples, 8192 per chunk, 16 chunks per group).
constexpr uint32_t kMaxGroups = 16384;
constexpr uint32_t kDirThreads = 1024;

// Exclusive scan over a workgroup of kDirThreads threads; `sh` holds 16 words.
WZ_D uint32_t dir_exclusive(uint32_t v, uint32_t* sh, uint32_t& total) {
    const uint32_t inc = wave_inclusive(v);
    const uint32_t w = threadIdx.x >> 6;
    if ((threadIdx.x & 63u) == 63u) sh[w] = inc;
    __syncthreads();
    uint32_t before = 0, sum = 0;
#pragma unroll
    for (uint32_t i = 0; i < kDirThreads / 64; ++i) {
        const uint32_t s = sh[i];
        before += i < w ? s : 0u;
        sum += s;
    }
    __syncthreads();
    total = sum;
    return before + inc - v;
}

// Directory of one WZP stream of `bytes` >= header bytes, checked by one
// workgroup of kDirThreads threads (cpu/src/wzp_cpu.cpp:713-761): each group's
// span [offset, offset + length) lies inside the payload, an empty group has
// offset 0, and the non-empty spans tile the payload exactly (no overlap, gap,
// trailing or missing bytes). Lengths follow from the chunk metas. The encoder
// writes the spans in group order (offset = sum of the earlier lengths), which
// one scan confirms; any other order is checked by ranks: the starts must be
// distinct and each span must start at the summed length of the spans that
// start before it. `len` is kMaxGroups words of LDS.
WZ_D void check_directory(const uint8_t* __restrict__ in, uint64_t bytes, uint32_t n, uint32_t chunks,
                          uint32_t groups, uint32_t* len, uint32_t* sh, uint32_t* __restrict__ status) {
    const uint32_t t = threadIdx.x;
    const uint64_t header = uint64_t(groups) * 4 + uint64_t(chunks) * 3;
    const uint64_t payload = bytes - header;
    const uint8_t* meta = in + uint64_t(groups) * 4;
    if (payload > 0xFFFFFFFFull || groups > kMaxGroups) {  // uniform: no workgroup barrier follows
        if (t == 0) atomicOr(status, uint32_t(kCorrupt));
        return;
    }
    int bad = 0, chain = 1;
    uint64_t carry = 0;
    for (uint32_t base = 0; base < groups; base += kDirThreads) {
        const uint32_t g = base + t;
        uint32_t L = 0, O = 0;
        if (g < groups) {
            auto add = [&](uint32_t c, uint32_t mc) {
                const uint32_t code = mc & kRawCode, sc = mc >> 14;
                if (code) L += (code == kRawCode ? 2 * chunk_elements(n, c) : code) + (sc ? sc : kSignSlot);
            };
            if ((g + 1) * kGroup <= chunks && !(reinterpret_cast<uintptr_t>(meta) & 3u)) {
                // a whole group: its 48 meta bytes as 12 aligned words
                const uint32_t* mw = reinterpret_cast<const uint32_t*>(meta) + uint64_t(g) * 12;
                uint32_t w[12];
#pragma unroll
                for (int k = 0; k < 12; ++k) w[k] = mw[k];
#pragma unroll
                for (uint32_t j = 0; j < kGroup; ++j) {
                    const uint32_t k = 3 * j / 4, sh8 = 8 * (3 * j % 4);
                    const unsigned long long pair = (k + 1 < 12 ? (unsigned long long)w[k + 1] << 32 : 0ull) | w[k];
                    add(g * kGroup + j, uint32_t(pair >> sh8) & 0xFFFFFFu);
                }
            } else {
                for (uint32_t c = g * kGroup; c < chunks; ++c) add(c, load24(meta + uint64_t(c) * 3));
            }
            O = load32(in + uint64_t(g) * 4);
            len[g] = L;
            if (O > payload || L > payload - O || (L == 0 && O != 0)) bad = 1;
        }
        uint32_t round_total;
        const uint32_t ex = dir_exclusive(L, sh, round_total);
        if (L && uint64_t(O) != carry + ex) chain = 0;
        carry += round_total;
    }
    if (carry != payload) bad = 1;
    bad = __syncthreads_or(bad);
    chain = __syncthreads_and(chain);
    if (!bad && !chain) {
        for (uint32_t g = t; g < groups; g += kDirThreads) {
            const uint32_t L = len[g];
            if (!L) continue;
            const uint32_t O = load32(in + uint64_t(g) * 4);
            uint64_t before = 0;
            for (uint32_t h = 0; h < groups; ++h) {
                const uint32_t Lh = len[h];
                if (!Lh || h == g) continue;
                const uint32_t Oh = load32(in + uint64_t(h) * 4);
                if (Oh == O) bad = 1;
                before += Oh < O ? Lh : 0u;
            }
            if (before != O) bad = 1;
        }
        bad = __syncthreads_or(bad);
    }
    if (bad && t == 0) atomicOr(status, uint32_t(kCorrupt));
}

}  // namespace wz
