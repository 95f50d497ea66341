// WZP encoder pieces shared by the single-field and batched pipelines:
// encoder sources (packed magnitudes, or quantization of paired bricks on
// the fly), the per-chunk encoder and the per-group pass (payload placement
// by decoupled lookback). See wzp.hip for the stream format.
#pragma once

#include "wzp_device.hpp"

namespace wz {

// v_perm_b32 selector that packs the bytes of a word whose bits are set in
// the 4-bit mask m, in order, into the low bytes (selector 4 = a zero byte of
// the first operand for the rest).
WZ_HD uint32_t compaction_selector(uint32_t m) {
    uint32_t sel = 0x04040404u, k = 0;
    for (uint32_t b = 0; b < 4; ++b)
        if ((m >> b) & 1u) {
            sel = (sel & ~(0xFFu << (8 * k))) | (b << (8 * k));
            ++k;
        }
    return sel;
}

// Encoder inputs. load() gives thread t of chunk c the chunk's magnitudes
// [32t, 32t + 32) (zero past the chunk's E elements) and returns their sign
// word; a collective source (kCollective) is called by every thread of the
// workgroup at once and may use the encoder's staging buffer `lds` (kChunk / 2
// + kThreads words), free at that point, which it leaves free again.
struct MagSource {  // packed magnitude / sign arrays
    static constexpr bool kCollective = false;
    const uint16_t* mag;
    const uint32_t* sign;
    WZ_D uint32_t load(uint32_t c, uint32_t t, uint32_t E, uint32_t (&m)[32], uint32_t*, uint32_t*) const {
        const uint32_t first = t * 32;
        const uint16_t* p = mag + uint64_t(c) * kChunk + first;
        if (first + 32 <= E) {
            const uint4* v = reinterpret_cast<const uint4*>(p);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const uint4 x = v[q];
                const uint32_t wds[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
                for (int h = 0; h < 4; ++h) {
                    m[8 * q + 2 * h] = wds[h] & 0xFFFFu;
                    m[8 * q + 2 * h + 1] = wds[h] >> 16;
                }
            }
        } else {
#pragma unroll
            for (int i = 0; i < 32; ++i) m[i] = first + i < E ? uint32_t(p[i]) : 0u;
        }
        return first < E ? sign[uint64_t(c) * (kChunk / 32) + t] : 0u;
    }
};

// Quantizes the coefficient volume on the fly (paired-brick geometry): reads
// the thread's 4x4x2 sub-block, writes the dequantized coefficients back in
// place and appends q-outliers, exactly as quantize_kernel does per sample.
template <class T>
struct BrickSource {
    static constexpr bool kCollective = false;
    T* coeff;
    uint32_t nx, ny;
    T q, inv;
    SideList qo;
    // dequantized coefficients are written back inside [0,hx) x [0,hy) x
    // [0,hz) only: the whole volume, or the coarser levels' box when the
    // finest inverse level requantizes its high bands on load
    uint32_t hx, hy, hz;
    WZ_D uint32_t load(uint32_t c, uint32_t t, uint32_t, uint32_t (&m)[32], uint32_t* status, uint32_t*) const {
        uint32_t X, Y, Z;
        pair_subblock(nx, ny, c, t, X, Y, Z);
        uint32_t sw = 0;
        bool bad = false;
        // all eight rows in flight at once (the write-backs and q-outlier
        // atomics below would otherwise hold each load until the previous
        // row is done)
        T v[8][4];
#pragma unroll
        for (uint32_t r = 0; r < 8; ++r)
            load4(coeff + (uint64_t(Z + (r >> 2)) * ny + (Y + (r & 3u))) * nx + X, v[r]);
#pragma unroll
        for (uint32_t r = 0; r < 8; ++r) {
            const uint32_t dy = r & 3u, dz = r >> 2;
            const uint64_t at = (uint64_t(Z + dz) * ny + (Y + dy)) * nx + X;
            uint32_t mv[4], over = 0;
#pragma unroll
            for (uint32_t dx = 0; dx < 4; ++dx) {
                // cpu/src/quantize_avx2.cpp:45-62: finite and |scaled| < 2^63
                // (one compare: false for NaN and infinities), llrint, INT32
                // saturation
                const T scaled = v[r][dx] * inv;
                bad = bad || !(abs_t(scaled) < T(9223372036854775808.0));
                const T rounded = rint_t(scaled);
                const T am = abs_t(rounded);
                mv[dx] = am >= T(2147483648.0) ? uint32_t(INT32_MAX) : uint32_t(int32_t(am));
                const uint32_t i = subblock_rank(dx, dy, dz);
                m[i] = mv[dx] > 65535u ? 65535u : mv[dx];
                sw |= uint32_t(rounded < T(0)) << i;
                over |= uint32_t(mv[dx] > 65535u) << dx;
            }
            if (__ballot(over != 0)) {  // q-outliers (rare): one wave-uniform branch per row
#pragma unroll
                for (uint32_t dx = 0; dx < 4; ++dx)
                    append_qoutlier(qo, (over >> dx) & 1u, uint32_t(at + dx), mv[dx]);
            }
            if (X < hx && Y + dy < hy && Z + dz < hz) {  // dequantized write-back inside the box
                T d[4];
#pragma unroll
                for (uint32_t dx = 0; dx < 4; ++dx) d[dx] = requantize<T>(v[r][dx], q, inv);
                store4(coeff + at, d);
            }
        }
        if (bad) atomicOr(status, uint32_t(kQuantDomain));
        return sw;
    }
};

// Quantizes the coefficient volume of a field of full bricks on the fly
// (rank_bricks_fit): the workgroup reads the chunk's bricks row by row (four
// x-consecutive samples per thread and step), quantizes them as
// quantize_kernel does (write-back inside [0,hx) x [0,hy) x [0,hz) as in
// BrickSource, q-outliers appended) and places magnitudes and sign bits at
// their ranks in LDS; each thread then takes its 32 magnitudes from there.
template <class T>
struct RankSource {
    static constexpr bool kCollective = true;
    T* coeff;
    RankBricks g;
    T q, inv;
    SideList qo;
    uint32_t hx, hy, hz;
    WZ_D uint32_t load(uint32_t c, uint32_t t, uint32_t E, uint32_t (&m)[32], uint32_t* status, uint32_t* lds) const {
        __shared__ uint32_t kb[kRankBricksMax], ko[3 * kRankBricksMax];
        uint16_t* const mg = reinterpret_cast<uint16_t*>(lds);
        uint32_t* const sg = lds + kChunk / 2;
        sg[t] = 0;
        const RankChunk rc = rank_chunk(g, c, E, kb, ko);
        __syncthreads();
        const FastDiv by_q(64u * g.bz);  // f * q < 2^32: exact
        const uint32_t nq = rc.count * 64u * g.bz;
        bool bad = false;
        // a burst of quads per thread: their rank and coefficient loads are
        // all issued before the first LDS and write-back stores
        constexpr uint32_t kBurst = 4;
        for (uint32_t f0 = 0; f0 < nq; f0 += kBurst * kThreads) {
            RankQuad rq[kBurst];
            T v[kBurst][4];
#pragma unroll
            for (uint32_t u = 0; u < kBurst; ++u) {
                const uint32_t f = f0 + u * kThreads + t;
                rq[u] = rank_quad(g, rc, by_q, kb, f < nq ? f : 0u);
                load4(coeff + rq[u].at, v[u]);
            }
#pragma unroll
            for (uint32_t u = 0; u < kBurst; ++u) {
                if (f0 + u * kThreads + t >= nq) continue;
                const RankQuad& r = rq[u];
                T d[4];
                bool all = true;
                uint32_t mvs[4], over = 0;
#pragma unroll
                for (uint32_t k = 0; k < 4; ++k) {
                    const bool in = r.j[k] < E;
                    all = all && in;
                    // cpu/src/quantize_avx2.cpp:45-62: finite and |scaled| < 2^63
                    // (one compare: false for NaN and infinities), llrint,
                    // INT32 saturation
                    const T scaled = v[u][k] * inv;
                    const T rounded = rint_t(scaled);
                    const T am = abs_t(rounded);
                    const uint32_t mv = am >= T(2147483648.0) ? uint32_t(INT32_MAX) : uint32_t(int32_t(am));
                    mvs[k] = mv;
                    d[k] = requantize<T>(v[u][k], q, inv);
                    if (!in) continue;
                    over |= uint32_t(mv > 65535u) << k;
                    bad = bad || !(abs_t(scaled) < T(9223372036854775808.0));
                    mg[r.j[k]] = uint16_t(mv > 65535u ? 65535u : mv);
                    if (rounded < T(0)) atomicOr(&sg[r.j[k] >> 5], 1u << (r.j[k] & 31u));
                }
                if (__ballot(over != 0)) {  // q-outliers (rare): one wave-uniform branch per quad
#pragma unroll
                    for (uint32_t k = 0; k < 4; ++k) append_qoutlier(qo, (over >> k) & 1u, r.at + k, mvs[k]);
                }
                // dequantized write-back inside the box (a quad never
                // straddles hx: hx is 0, nx or nx / 2, multiples of 8)
                const uint32_t X = ko[3 * r.lb] + (r.li & 15u), Y = ko[3 * r.lb + 1] + ((r.li >> 4) & 15u),
                               Z = ko[3 * r.lb + 2] + (r.li >> 8);
                if (X < hx && Y < hy && Z < hz) {
                    if (all) {
                        store4(coeff + r.at, d);
                    } else {
#pragma unroll
                        for (uint32_t k = 0; k < 4; ++k)
                            if (r.j[k] < E) coeff[r.at + k] = d[k];
                    }
                }
            }
        }
        __syncthreads();
        const uint32_t first = 32u * t;
        const uint4* w = reinterpret_cast<const uint4*>(lds) + 4u * t;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const uint4 x = w[k];
            const uint32_t wds[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
            for (int h = 0; h < 4; ++h) {
                m[8 * k + 2 * h] = first + 8 * k + 2 * h < E ? wds[h] & 0xFFFFu : 0u;
                m[8 * k + 2 * h + 1] = first + 8 * k + 2 * h + 1 < E ? wds[h] >> 16 : 0u;
            }
        }
        const uint32_t sw = sg[t];
        __syncthreads();  // the buffer is free again for the caller
        if (bad) atomicOr(status, uint32_t(kQuantDomain));
        return sw;
    }
};

// Encodes chunk c of an n-sample stream with `groups` groups (one workgroup of
// kThreads threads): meta at out + 4 groups + 3 c (when inside `capacity`),
// payload bytes to the chunk's staging slot, its length to w.size[c].
template <class Source>
WZ_D void encode_chunk(const Source& src, uint32_t c, uint32_t n, uint32_t groups, uint8_t* __restrict__ out,
                       uint64_t capacity, const WzpScratch& w, uint32_t* __restrict__ status) {
    // BIT_2 planes (16 KiB) and, once every thread holds its share of them in
    // registers, the assembled payload: one LDS buffer
    __shared__ __attribute__((aligned(16))) uint32_t stage[(kChunkBytes + kSignSlot) / 4 + 4];
    uint32_t* const tr = stage;
    uint32_t* const ob32 = stage;
    __shared__ uint32_t csel[16];                      // byte-compaction selectors
    __shared__ unsigned long long b0s[kThreads];       // B0, 2048 bytes
    __shared__ __attribute__((aligned(8))) uint8_t b1s[kThreads];
    __shared__ __attribute__((aligned(8))) uint8_t b2s[32];
    __shared__ uint32_t signbuf[kThreads + 1];
    __shared__ uint32_t sh[8];
    __shared__ uint32_t n3s[4], b3s[4];

    const uint32_t t = threadIdx.x;
    signbuf[t] = 0;
    if (t < 16) csel[t] = compaction_selector(t);
    if (t == 0) signbuf[kThreads] = 0;
    __syncthreads();
    const uint32_t E = chunk_elements(n, c);
    const uint32_t raw_bytes = 2 * E;
    const uint32_t first = t * 32;

    // 1. this thread's 32 magnitudes and their sign word
    uint32_t m[32];
    const uint32_t sw = src.load(c, t, E, m, status, stage);
    uint32_t nzmask = 0;
#pragma unroll
    for (int i = 0; i < 32; ++i) nzmask |= uint32_t(m[i] != 0) << i;
    const uint32_t cnt = popc32(nzmask);

    // 2. BIT_2 planes into LDS (plane p of block b at u16 index p*blocks+b)
    uint16_t* tr16 = reinterpret_cast<uint16_t*>(tr);
    const uint32_t blocks = E / 16;
    if (2 * t + 1 < blocks) {  // both blocks whole: one two-block transform
        uint32_t pp[16];
        bit2_forward_pair(m, pp);
        if (!(blocks & 1u)) {  // the pair's u16 slots share one aligned word
#pragma unroll
            for (int pl = 0; pl < 16; ++pl) tr[pl * (blocks >> 1) + t] = pp[pl];
        } else {
#pragma unroll
            for (int pl = 0; pl < 16; ++pl) {
                tr16[pl * blocks + 2 * t] = uint16_t(pp[pl]);
                tr16[pl * blocks + 2 * t + 1] = uint16_t(pp[pl] >> 16);
            }
        }
    } else {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const uint32_t b = 2 * t + h;
            if (b < blocks) {
                uint32_t p[16];
#pragma unroll
                for (int j = 0; j < 16; ++j) p[j] = m[16 * h + j];
                bit2_forward(p);
#pragma unroll
                for (int pl = 0; pl < 16; ++pl) tr16[pl * blocks + b] = uint16_t(p[pl]);
            } else {
#pragma unroll
                for (int j = 0; j < 16; ++j) {
                    const uint32_t i = first + 16 * h + j;
                    if (i < E) tr16[i] = uint16_t(m[16 * h + j]);
                }
            }
        }
    }
    __syncthreads();

    // 3. RZE bitmaps: B0 (nonzero bytes), B1/B2/B3 (changed bytes of the level below)
    // this thread's 64 transformed bytes stay in registers for the payload
    uint32_t tw[16];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const uint4 v = reinterpret_cast<const uint4*>(tr)[4 * t + q];
        tw[4 * q] = v.x;
        tw[4 * q + 1] = v.y;
        tw[4 * q + 2] = v.z;
        tw[4 * q + 3] = v.w;
    }
    unsigned long long b0t = 0;
#pragma unroll
    for (int q = 0; q < 16; ++q) b0t |= (unsigned long long)nonzero_bytes4(tw[q]) << (4 * q);
    {
        const int32_t valid = int32_t(raw_bytes) - int32_t(64 * t);
        if (valid <= 0) b0t = 0;
        else if (valid < 64) b0t &= (1ull << valid) - 1ull;
    }
    b0s[t] = b0t;
    __syncthreads();
    const uint32_t b1t = change_flags8(b0t, t ? uint32_t(b0s[t - 1] >> 56) : 0u);
    b1s[t] = uint8_t(b1t);
    __syncthreads();
    uint32_t b2t = 0;
    if (t < 32) {
        const unsigned long long w8 = *reinterpret_cast<const unsigned long long*>(&b1s[8 * t]);
        b2t = change_flags8(w8, t ? uint32_t(b1s[8 * t - 1]) : 0u);
        b2s[t] = uint8_t(b2t);
    }
    __syncthreads();
    if (t < 4) {
        const unsigned long long w8 = *reinterpret_cast<const unsigned long long*>(&b2s[8 * t]);
        const uint32_t b3t = change_flags8(w8, t ? uint32_t(b2s[8 * t - 1]) : 0u);
        b3s[t] = b3t;
        n3s[t] = popc32(b3t);
    }
    const uint32_t nzb = popc64(b0t), n1 = popc32(b1t), n2 = popc32(b2t);
    uint32_t totA, totB;
    const uint32_t exA = block_exclusive(nzb | (n1 << 16), sh, totA);   // also orders b3s/n3s
    const uint32_t exB = block_exclusive(cnt | (n2 << 16), sh, totB);
    const uint32_t e0 = exA & 0xFFFFu, e1 = exA >> 16, es = exB & 0xFFFFu, e2 = exB >> 16;
    const uint32_t K0 = totA & 0xFFFFu, K1 = totA >> 16, NNZ = totB & 0xFFFFu, K2 = totB >> 16;
    const uint32_t K3 = n3s[0] + n3s[1] + n3s[2] + n3s[3];
    const uint32_t rze = K0 + K1 + K2 + K3 + 4;
    const bool raw = rze + 2 >= raw_bytes;
    const uint32_t stored = raw ? raw_bytes : rze;
    const uint32_t sign_bytes = (NNZ + 7) / 8;
    const uint32_t chunk_bytes = NNZ ? stored + sign_bytes : 0u;
    // the planes are in registers (tw) since the B0 barrier: clear the buffer
    // for the OR-assembled payload (a raw chunk re-reads its planes instead;
    // the barrier after step 5 orders both before the payload writes)
    if (!raw)
        for (uint32_t i = t; i < (kChunkBytes + kSignSlot) / 16 + 1; i += kThreads)
            reinterpret_cast<uint4*>(ob32)[i] = make_uint4(0, 0, 0, 0);

    // 4. compacted signs of the nonzero magnitudes, LSB first, into LDS
    {
        const uint32_t comp = compress_bits(sw, nzmask);
        if (cnt) {
            const uint32_t off = es & 31u, wd = es >> 5;
            atomicOr(&signbuf[wd], comp << off);
            if (off && off + cnt > 32) atomicOr(&signbuf[wd + 1], comp >> (32 - off));
        }
    }

    // 5. size and meta of the chunk (its payload offset is found by the group pass)
    if (t == 0) {
        w.size[c] = chunk_bytes;
        if (uint64_t(groups) * 4 + uint64_t(c) * 3 + 3 <= capacity) {
            const uint32_t sign_code = sign_bytes == kSignSlot ? 0u : sign_bytes;
            const uint32_t meta = NNZ ? ((raw ? kRawCode : stored) | (sign_code << 14)) : 0u;
            uint8_t* mp = out + uint64_t(groups) * 4 + uint64_t(c) * 3;
            mp[0] = uint8_t(meta);
            mp[1] = uint8_t(meta >> 8);
            mp[2] = uint8_t(meta >> 16);
        }
    }
    __syncthreads();  // the payload buffer is cleared (above) before any thread ORs bytes into it
    if (!NNZ) return;
    // 7. assemble the chunk payload in LDS
    uint8_t* ob = reinterpret_cast<uint8_t*>(ob32);
    if (raw) {  // LE u16 magnitudes in place: whole blocks from their staged
                // planes; the samples past the last whole block are staged
                // verbatim at their own positions already
        const uint16_t* tr16c = reinterpret_cast<const uint16_t*>(tr);
        uint32_t pp[16], mr[32];
        if (!(blocks & 1u)) {  // (indices stay inside the buffer for every thread)
#pragma unroll
            for (int pl = 0; pl < 16; ++pl) pp[pl] = tr[pl * (blocks >> 1) + t];
        } else {
#pragma unroll
            for (int pl = 0; pl < 16; ++pl)
                pp[pl] = uint32_t(tr16c[pl * blocks + 2 * t]) | (uint32_t(tr16c[pl * blocks + 2 * t + 1]) << 16);
        }
        bit2_inverse_pair(pp, mr);
        const uint32_t whole = 2 * t + 1 < blocks ? 32u : (2 * t < blocks ? 16u : 0u);
        __syncthreads();  // all planes read before the buffer is overwritten
        uint16_t* ob16 = reinterpret_cast<uint16_t*>(ob32);
#pragma unroll
        for (int i = 0; i < 32; ++i)
            if (uint32_t(i) < whole) ob16[first + i] = uint16_t(mr[i]);
    } else {
        {  // nonzero bytes in order: each word's nonzero bytes packed by one
           // byte permute, gathered in an 8-byte accumulator and OR-ed into
           // the zeroed payload a word at a time (a lane's two edge words are
           // shared with its neighbours; the bytes are disjoint)
            uint32_t wi = e0 >> 2, nacc = e0 & 3u;
            unsigned long long acc = 0;
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const uint32_t m = uint32_t(b0t >> (4 * q)) & 15u;
                const uint32_t packed = __builtin_amdgcn_perm(0u, tw[q], csel[m]);
                acc |= (unsigned long long)packed << (8 * nacc);
                nacc += popc32(m);
                if (nacc >= 4) {
                    atomicOr(&ob32[wi], uint32_t(acc));
                    acc >>= 32;
                    nacc -= 4;
                    ++wi;
                }
            }
            if (nacc) atomicOr(&ob32[wi], uint32_t(acc));
        }
        {  // changed B0 bytes
            uint32_t k = K0 + e1, bits = b1t;
            while (bits) {
                const uint32_t j = uint32_t(__builtin_ctz(bits));
                ob[k++] = uint8_t(b0t >> (8 * j));
                bits &= bits - 1;
            }
        }
        if (t < 32) {  // changed B1 bytes
            uint32_t k = K0 + K1 + e2, bits = b2t;
            while (bits) {
                const uint32_t j = uint32_t(__builtin_ctz(bits));
                ob[k++] = b1s[8 * t + j];
                bits &= bits - 1;
            }
        }
        if (t < 4) {  // changed B2 bytes, then B3 verbatim
            uint32_t k = K0 + K1 + K2;
            for (uint32_t i = 0; i < t; ++i) k += n3s[i];
            uint32_t bits = b3s[t];
            while (bits) {
                const uint32_t j = uint32_t(__builtin_ctz(bits));
                ob[k++] = b2s[8 * t + j];
                bits &= bits - 1;
            }
            ob[rze - 4 + t] = uint8_t(b3s[t]);
        }
    }
    for (uint32_t i = t; i < sign_bytes; i += kThreads) ob[stored + i] = uint8_t(signbuf[i >> 2] >> (8 * (i & 3u)));
    __syncthreads();

    // 8. the payload into the chunk's staging slot (16-byte aligned)
    uint4* const slot = reinterpret_cast<uint4*>(w.staging + uint64_t(c) * kChunkSlot);
    for (uint32_t q = t; q < (chunk_bytes + 15) / 16; q += kThreads) slot[q] = reinterpret_cast<const uint4*>(ob32)[q];
}

// Group pass (cpu/src/wzp_cpu.cpp:789-830 for the directory): workgroup =
// one 16-chunk group, in ticket order. Wave 0 scans the group's chunk sizes
// and publishes the group length; the whole workgroup looks back over the
// earlier groups' descriptors (one per thread per round trip, nearest
// inclusive prefix by ballots, block reduction of the aggregates up to it).
// Descriptors are the only data passed between workgroups, so relaxed
// agent-scope atomics suffice (the slots were written by the previous
// kernel). Then the group's directory entry (offset of its first payload
// byte, 0 for an empty group) and its payloads, moved
// from the slots with 16-byte stores over funnel-shifted 16-byte loads.
constexpr uint32_t kPackThreads = 1024, kPackUnroll = 4;
// Group g (taken in ticket order by the caller: every earlier group of the
// stream is held by a running workgroup) by a workgroup of kPackThreads threads.
WZ_D void pack_group(uint32_t g, uint8_t* __restrict__ out, uint64_t capacity, uint32_t chunks, uint32_t groups,
                     const WzpScratch& w, unsigned long long* __restrict__ total, uint32_t* __restrict__ status) {
    __shared__ uint32_t off_sh[kGroup + 1];  // chunk offsets inside the group; [kGroup] = group length
    __shared__ uint32_t qpre[kGroup + 1];    // 16-byte destination groups before each chunk
    __shared__ uint32_t ok_sh, wait_sh;
    __shared__ uint32_t near_w[kPackThreads / 64];        // per wave: nearest inclusive lane (or 64)
    __shared__ unsigned long long part_w[kPackThreads / 64];
    __shared__ unsigned long long excl_sh;
    const uint32_t t = threadIdx.x, lane = t & 63u, wv = t >> 6;
    const uint32_t first = g * kGroup, count = min(kGroup, chunks - first);
    const uint64_t header = uint64_t(groups) * 4 + uint64_t(chunks) * 3;
    // the group's chunk offsets and length (wave 0), published at once
    if (t < 64) {
        const uint32_t size = lane < count ? w.size[first + lane] : 0u;
        const uint32_t incl = wave_inclusive(size);
        if (lane <= kGroup) off_sh[lane] = incl - size;  // lane kGroup: the group length
        if (lane == 63)
            __hip_atomic_store(&w.desc[g], (g == 0 ? kFlagInclusive : kFlagAggregate) | incl, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    const uint32_t glen = off_sh[kGroup];
    // decoupled lookback with the whole workgroup: thread t reads the
    // descriptor of group j - t, so one round trip covers kPackThreads groups;
    // stop at the nearest inclusive prefix and sum the aggregates up to it
    unsigned long long acc = 0;
    if (g > 0) {
        int64_t j = int64_t(g) - 1;
        uint32_t spins = 0;
        for (;;) {
            const int64_t idx = j - int64_t(t);
            const unsigned long long d =
                idx >= 0 ? __hip_atomic_load(&w.desc[idx], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                         : kFlagInclusive;  // nothing before group 0
            const unsigned long long flag = d & ~kValueMask;
            const unsigned long long im = __ballot(flag == kFlagInclusive);
            if (lane == 0) near_w[wv] = im ? uint32_t(__ffsll((long long)im) - 1) : 64u;
            if (t == 0) wait_sh = 0;
            __syncthreads();
            uint32_t stop = kPackThreads;  // block position of the nearest inclusive descriptor
            for (uint32_t k = 0; k < kPackThreads / 64; ++k)
                if (near_w[k] < 64) {
                    stop = 64 * k + near_w[k];
                    break;
                }
            if (t <= stop && flag == 0) wait_sh = 1;  // a descriptor we need is not published yet
            unsigned long long v = t <= stop ? (d & kValueMask) : 0ull;
#pragma unroll
            for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
            if (lane == 0) part_w[wv] = v;
            __syncthreads();
            const bool wait = wait_sh != 0;
            unsigned long long sum = 0;
            for (uint32_t k = 0; k < kPackThreads / 64; ++k) sum += part_w[k];
            __syncthreads();  // near_w / part_w / wait_sh are rewritten by the next round
            if (wait) {
                if (++spins > kSpinLimit) {
                    if (t == 0) atomicOr(status, uint32_t(kLookbackTimeout));
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
                continue;
            }
            acc += sum;
            if (stop < kPackThreads) break;
            j -= kPackThreads;
        }
        if (t == 0)
            __hip_atomic_store(&w.desc[g], kFlagInclusive | (acc + glen), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (t == 0) {
        excl_sh = acc;
        const bool fits = header + acc + glen <= capacity;
        ok_sh = fits;
        if (!fits) atomicOr(status, uint32_t(kWzpCapacity));
        if (g == groups - 1) *total = header + acc + glen;
        if (uint64_t(g) * 4 + 4 <= capacity) {
            const uint32_t off = glen ? uint32_t(acc) : 0u;
            uint8_t* p = out + uint64_t(g) * 4;
            p[0] = uint8_t(off);
            p[1] = uint8_t(off >> 8);
            p[2] = uint8_t(off >> 16);
            p[3] = uint8_t(off >> 24);
        }
    }
    if (t < 64) {  // destination 16-byte groups per chunk, then their prefix
        const uint32_t size = lane < kGroup ? off_sh[lane + 1] - off_sh[lane] : 0u;
        const uint64_t dst0 = reinterpret_cast<uintptr_t>(out) + header + acc + (lane < kGroup ? off_sh[lane] : 0u);
        const uint32_t lead = uint32_t(dst0 & 15u);
        const uint32_t nq = size ? (lead + size + 15) / 16 : 0u;
        const uint32_t qi = wave_inclusive(nq);
        if (lane <= kGroup) qpre[lane] = qi - nq;
    }
    __syncthreads();
    if (!ok_sh) return;
    // every destination group of every chunk of the group, kPackUnroll per
    // thread per round with all their loads issued before any store
    const uint32_t all = qpre[kGroup];
    struct Unit {
        uint8_t* dst;        // 16-byte aligned destination group
        const uint8_t* src;  // the chunk's slot
        int32_t o;           // payload byte at dst
        uint32_t size;
    };
    auto unit = [&](uint32_t u) {
        uint32_t i = 0;
#pragma unroll
        for (uint32_t k = 1; k < kGroup; ++k) i += u >= qpre[k] ? 1u : 0u;
        uint8_t* const d = out + header + excl_sh + off_sh[i];
        const uint32_t lead = uint32_t(reinterpret_cast<uintptr_t>(d) & 15u);
        const uint32_t q = u - qpre[i];
        return Unit{d - lead + 16 * q, w.staging + uint64_t(first + i) * kChunkSlot, int32_t(16 * q) - int32_t(lead),
                    off_sh[i + 1] - off_sh[i]};
    };
    // the archive is not read again here: non-temporal stores (staging keeps
    // ordinary loads, which neighbouring destination groups share)
    using U4 = uint32_t __attribute__((ext_vector_type(4)));
    for (uint32_t u0 = t; u0 < all; u0 += kPackThreads * kPackUnroll) {
        U4 lo[kPackUnroll], hi[kPackUnroll];
#pragma unroll
        for (uint32_t r = 0; r < kPackUnroll; ++r) {
            const uint32_t u = u0 + r * kPackThreads;
            if (u >= all) break;
            const Unit x = unit(u);
            if (x.o >= 0 && uint32_t(x.o) + 16 <= x.size) {  // the 16 bytes lie in two aligned 16-byte loads
                lo[r] = *reinterpret_cast<const U4*>(x.src + (uint32_t(x.o) & ~15u));
                hi[r] = *reinterpret_cast<const U4*>(x.src + (uint32_t(x.o) & ~15u) + 16);
            }
        }
#pragma unroll
        for (uint32_t r = 0; r < kPackUnroll; ++r) {
            const uint32_t u = u0 + r * kPackThreads;
            if (u >= all) break;
            const Unit x = unit(u);
            if (x.o >= 0 && uint32_t(x.o) + 16 <= x.size) {
                // starting at word k of lo, byte sb
                const uint32_t k = (uint32_t(x.o) >> 2) & 3u, sb = uint32_t(x.o) & 3u;
                const uint32_t wv[8] = {lo[r].x, lo[r].y, lo[r].z, lo[r].w, hi[r].x, hi[r].y, hi[r].z, hi[r].w};
                uint32_t v[5];
#pragma unroll
                for (int j = 0; j < 5; ++j)
                    v[j] = k == 0 ? wv[j] : k == 1 ? wv[j + 1] : k == 2 ? wv[j + 2] : wv[j + 3];
                const U4 o = {__builtin_amdgcn_alignbyte(v[1], v[0], sb), __builtin_amdgcn_alignbyte(v[2], v[1], sb),
                              __builtin_amdgcn_alignbyte(v[3], v[2], sb), __builtin_amdgcn_alignbyte(v[4], v[3], sb)};
                __builtin_nontemporal_store(o, reinterpret_cast<U4*>(x.dst));
            } else {  // an edge group: the chunk's bytes only
                for (int j = 0; j < 16; ++j) {
                    const int32_t p = x.o + j;
                    if (p >= 0 && uint32_t(p) < x.size) x.dst[j] = x.src[p];
                }
            }
        }
    }
}


}  // namespace wz
