/*
 * ORACLE — test infrastructure only (see waltz_oracle.h). A plain-C restatement
 * of the reference Waltz CPU scalar codec; the product never links it.
 *
 * The WZP bit/byte format restated here (BIT_2 bit transpose, RZE zero and
 * repetition elimination) descends from the LC framework (BSD-3-Clause,
 * https://github.com/burtscher/LC-framework) through the reference's
 * cpu/src/wzp_cpu.cpp; this file is an independent restatement of that
 * format, not a copy of either implementation.
 */
#define _GNU_SOURCE
#include "waltz_oracle.h"

#include <fenv.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#if defined(__x86_64__) || defined(__i386__)
#include <xmmintrin.h>
#endif

/* ---------------------------------------------------------------- errors */
static _Thread_local char g_error[256];
const char* wo_last_error(void) { return g_error; }
static int fail(const char* message) {
    snprintf(g_error, sizeof g_error, "%s", message);
    return -1;
}

/* Round-to-nearest, FTZ/DAZ cleared for the duration of a call:
 * cpu/src/codec.cpp:52-72. */
typedef struct {
    fenv_t env;
    unsigned csr;
} fpguard;
static void guard_enter(fpguard* g) {
    fegetenv(&g->env);
    fesetround(FE_TONEAREST);
#if defined(__x86_64__) || defined(__i386__)
    g->csr = _mm_getcsr();
    _mm_setcsr(g->csr & ~((1u << 15) | (1u << 6)));
#endif
}
static void guard_leave(fpguard* g) {
    fesetenv(&g->env);
#if defined(__x86_64__) || defined(__i386__)
    _mm_setcsr(g->csr);
#endif
}

/* ---------------------------------------------------------------- geometry */
typedef struct { uint32_t nx, ny, nz; } oshape;
static uint32_t min_u(uint32_t a, uint32_t b) { return a < b ? a : b; }
static uint32_t half_len(uint32_t n) { return n / 2 + n % 2; }
static oshape shrink(oshape s, int xy, int z) {
    if (xy) { s.nx = half_len(s.nx); s.ny = half_len(s.ny); }
    if (z) s.nz = half_len(s.nz);
    return s;
}
/* Whole-sample mirror with period 2n-2: cpu/src/transform.cpp:60-66. */
static uint32_t reflect_pos(int64_t p, uint32_t n) {
    if (n <= 1) return 0;
    const int64_t period = 2 * (int64_t)n - 2;
    int64_t r = p % period;
    if (r < 0) r += period;
    return (uint32_t)(r < (int64_t)n ? r : period - r);
}
/* Split-band position of a (reflected) sample: cpu/src/transform.cpp:68-71. */
static uint32_t split_pos(int64_t p, uint32_t n) {
    const uint32_t r = reflect_pos(p, n);
    return r / 2 + (r % 2 ? half_len(n) : 0);
}

/* include/utils/Blob.hpp:22-29 */
unsigned wo_dwt_levels(uint32_t n) {
    unsigned l = 0;
    while (n > 8 && l < 5) { n -= n / 2; ++l; }
    return l;
}
/* cpu/src/codec.cpp:89-95 */
uint32_t wo_zband(uint32_t nx, uint32_t ny, uint32_t nz, unsigned lz, int dyadic) {
    if (dyadic) return min_u(16u, (nz + ((1u << lz) - 1u)) >> lz);
    if (nz < 4) return nz;
    const uint64_t blocks = (uint64_t)((nx + 15u) / 16) * ((ny + 15u) / 16);
    const uint64_t full = (uint64_t)(nx / 16) * (ny / 16);
    return (blocks - full) * 32 >= blocks ? 4u : 2u;
}
/* Recursive ceil-half octant rank (z half, then y, then x): cpu/src/codec.cpp:96-107. */
uint32_t wo_morton_rank(uint32_t x, uint32_t y, uint32_t z, uint32_t nx, uint32_t ny, uint32_t nz) {
    uint32_t rank = 0;
    while (nx > 1 || ny > 1 || nz > 1) {
        const uint32_t lx = nx - nx / 2, ly = ny - ny / 2, lz = nz - nz / 2;
        const int cx = x >= lx, cy = y >= ly, cz = z >= lz;
        const uint32_t zn = cz ? nz / 2 : lz, yn = cy ? ny / 2 : ly;
        if (cz) rank += lz * nx * ny;
        if (cy) rank += zn * ly * nx;
        if (cx) rank += zn * yn * lx;
        if (cx) x -= lx;
        if (cy) y -= ly;
        if (cz) z -= lz;
        nx = cx ? nx / 2 : lx;
        ny = yn;
        nz = zn;
    }
    return rank;
}
/* Ranks for the eight brick shapes (full / x-edge / y-edge / z-edge combos),
 * cpu/src/codec.cpp:108-116. Table s*8192 + (z*ny+y)*nx+x. */
static uint16_t* brick_ranks(oshape d, uint32_t bz) {
    uint16_t* rank = (uint16_t*)calloc(8 * 8192, 2);
    if (!rank) return NULL;
    for (unsigned s = 0; s < 8; ++s) {
        const uint32_t nx = (s & 1) ? d.nx % 16 : 16, ny = (s & 2) ? d.ny % 16 : 16, nz = (s & 4) ? d.nz % bz : bz;
        for (uint32_t z = 0; z < nz; ++z)
            for (uint32_t y = 0; y < ny; ++y)
                for (uint32_t x = 0; x < nx; ++x)
                    rank[s * 8192 + (z * ny + y) * nx + x] = (uint16_t)wo_morton_rank(x, y, z, nx, ny, nz);
    }
    return rank;
}

/* Keys are unique element indices, so any correct sort yields the reference's
 * order (cpu/src/codec.cpp:167-176,225-234). */
typedef struct { uint32_t key, value; } kv;
static int kv_cmp(const void* a, const void* b) {
    const uint32_t x = ((const kv*)a)->key, y = ((const kv*)b)->key;
    return x < y ? -1 : x > y;
}
static void sort_index_pairs(uint32_t* index, uint32_t* value, uint32_t count) {
    if (count < 2) return;
    kv* t = (kv*)malloc(sizeof(kv) * count);
    if (!t) abort();
    for (uint32_t i = 0; i < count; ++i) { t[i].key = index[i]; t[i].value = value[i]; }
    qsort(t, count, sizeof(kv), kv_cmp);
    for (uint32_t i = 0; i < count; ++i) { index[i] = t[i].key; value[i] = t[i].value; }
    free(t);
}

/* ---------------------------------------------------------------- bytes */
static void put16(uint8_t* p, uint16_t v) { p[0] = (uint8_t)v; p[1] = (uint8_t)(v >> 8); }
static void put24(uint8_t* p, uint32_t v) { p[0] = (uint8_t)v; p[1] = (uint8_t)(v >> 8); p[2] = (uint8_t)(v >> 16); }
static void put32(uint8_t* p, uint32_t v) { put24(p, v); p[3] = (uint8_t)(v >> 24); }
static void put64(uint8_t* p, uint64_t v) { put32(p, (uint32_t)v); put32(p + 4, (uint32_t)(v >> 32)); }
static uint16_t get16(const uint8_t* p) { return (uint16_t)(p[0] | (p[1] << 8)); }
static uint32_t get24(const uint8_t* p) { return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16); }
static uint32_t get32(const uint8_t* p) { return get24(p) | ((uint32_t)p[3] << 24); }
static uint64_t get64(const uint8_t* p) { return get32(p) | ((uint64_t)get32(p + 4) << 32); }

/* Section layout after the 32-byte header, include/utils/Blob.hpp:71-79. */
typedef struct { uint64_t pi, pv, qi, qv, end; } blob_layout;
static uint64_t pad8(uint64_t n) { return (n + 7) & ~(uint64_t)7; }
static blob_layout layout(uint64_t wzp, uint32_t np, uint32_t repair_bytes, uint32_t nq) {
    blob_layout L;
    L.pi = 32 + pad8(wzp);
    L.pv = L.pi + pad8((uint64_t)np * 4);
    L.qi = L.pv + pad8(repair_bytes);
    L.qv = L.qi + pad8((uint64_t)nq * 4);
    L.end = L.qv + pad8((uint64_t)nq * 4);
    return L;
}
/* Indices < n and pairwise distinct, cpu/src/codec.cpp:260-275. */
static int check_indices(const uint8_t* p, uint32_t count, uint32_t n) {
    int ordered = 1;
    uint32_t prev = 0;
    for (uint32_t i = 0; i < count; ++i) {
        const uint32_t v = get32(p + 4 * (size_t)i);
        if (v >= n) return fail("outlier index exceeds volume");
        if (i && v <= prev) ordered = 0;
        prev = v;
    }
    if (ordered) return 0;
    uint8_t* seen = (uint8_t*)calloc(n, 1);
    if (!seen) return fail("out of memory");
    for (uint32_t i = 0; i < count; ++i) {
        const uint32_t v = get32(p + 4 * (size_t)i);
        if (seen[v]) { free(seen); return fail("duplicate outlier index"); }
        seen[v] = 1;
    }
    free(seen);
    return 0;
}

/* Exact strict bound: cpu/src/repair_avx2.cpp:48-56. */
int wo_within_bound(double value, double original, double bound) {
    const double high = value - original, magnitude = fabs(high);
    if (!isfinite(high) || magnitude > bound) return 0;
    if (magnitude < bound) return 1;
    const double bv = value - high, av = high + bv;
    const double br = bv - original, ar = value - av, low = ar + br;
    return !((high > 0 && low > 0) || (high < 0 && low < 0));
}

/* ---------------------------------------------------------------- WZP
 * Stream: [groups x u32 offset][chunks x 24-bit meta][payload]
 * chunk = 8192 magnitudes, group = 16 chunks (cpu/src/wzp_cpu.cpp:57-62,86-99). */
enum { CHUNK = 8192, GROUP = 16, SIGN_SLOT = 1024, RAW_CODE = 0x3fff };

typedef struct { uint32_t chunks, groups; size_t header, max_bytes; } wzp_shape;
static int wzp_geometry(uint32_t n, wzp_shape* s) {
    if (n > UINT32_MAX / 2u) return fail("WZP element count exceeds the GPU format limit");
    s->chunks = n == 0 ? 0 : 1u + (n - 1u) / CHUNK;
    s->groups = (s->chunks + GROUP - 1u) / GROUP;
    s->header = (size_t)s->groups * 4u + (size_t)s->chunks * 3u;
    s->max_bytes = s->header + (size_t)n * 2u + ((size_t)n + 7u) / 8u;
    return 0;
}
size_t wo_wzp_max_bytes(uint32_t n) {
    wzp_shape s;
    return wzp_geometry(n, &s) ? 0 : s.max_bytes;
}
static uint32_t chunk_len(uint32_t n, uint32_t chunk) {
    const uint32_t rest = n - chunk * CHUNK;
    return rest < CHUNK ? rest : CHUNK;
}

/* BIT_2: within each block of 16 magnitudes, bit k of element j becomes bit
 * 15-j of plane 15-k; plane p of block b is the u16 at p*blocks+b; trailing
 * elements%16 magnitudes stay raw (cpu/src/wzp_cpu.cpp:176-193). */
static void bit2_encode(const uint16_t* mag, uint32_t elements, uint8_t* out) {
    const uint32_t blocks = elements / 16u;
    for (uint32_t b = 0; b < blocks; ++b)
        for (unsigned p = 0; p < 16; ++p) {
            uint16_t word = 0;
            for (unsigned j = 0; j < 16; ++j)
                word |= (uint16_t)(((mag[b * 16 + j] >> (15 - p)) & 1u) << (15 - j));
            put16(out + 2 * (p * blocks + b), word);
        }
    for (uint32_t i = blocks * 16u; i < elements; ++i) put16(out + 2 * i, mag[i]);
}
static void bit2_decode(const uint8_t* in, uint32_t elements, uint16_t* mag) {
    const uint32_t blocks = elements / 16u;
    for (uint32_t b = 0; b < blocks; ++b)
        for (unsigned j = 0; j < 16; ++j) {
            uint16_t v = 0;
            for (unsigned p = 0; p < 16; ++p)
                v |= (uint16_t)(((get16(in + 2 * (p * blocks + b)) >> (15 - j)) & 1u) << (15 - p));
            mag[b * 16 + j] = v;
        }
    for (uint32_t i = blocks * 16u; i < elements; ++i) mag[i] = get16(in + 2 * i);
}

/* RZE over one chunk's bytes: B0 (2048 B) flags nonzero bytes; each level R in
 * 2048, 256, 32 keeps the bytes of B that differ from their predecessor (byte
 * -1 is 0) and flags them in the next bitmap; B3's 4 bytes close the stream.
 * Output: [nonzero bytes][changed B0 bytes][changed B1 bytes][changed B2 bytes][B3].
 * cpu/src/wzp_cpu.cpp:540-586. Returns the encoded size (0 if all zero). */
static uint32_t rze_encode(const uint8_t* in, uint32_t bytes, uint8_t* out /* may be NULL */) {
    uint8_t bm[2048 + 256 + 32 + 4];
    memset(bm, 0, sizeof bm);
    uint32_t pos = 0;
    for (uint32_t i = 0; i < bytes; ++i)
        if (in[i]) {
            bm[i >> 3] |= (uint8_t)(1u << (i & 7));
            if (out) out[pos] = in[i];
            ++pos;
        }
    if (pos == 0) return 0;
    uint32_t base = 0;
    for (uint32_t range = 2048; range != 4; range /= 8) {
        uint8_t prev = 0;
        for (uint32_t i = 0; i < range; ++i) {
            const uint8_t v = bm[base + i];
            if (v != prev) {
                bm[base + range + (i >> 3)] |= (uint8_t)(1u << (i & 7));
                if (out) out[pos] = v;
                ++pos;
            }
            prev = v;
        }
        base += range;
    }
    if (out) memcpy(out + pos, bm + base, 4);
    return pos + 4;
}
/* Inverse of rze_encode with the reference's consistency checks,
 * cpu/src/wzp_cpu.cpp:595-660. */
static int rze_decode(const uint8_t* in, uint32_t stored, uint32_t raw_bytes, uint8_t* out) {
    uint8_t bm[2048 + 256 + 32 + 4];
    if (stored < 4) return fail("WZP truncated RZE bitmap");
    uint32_t pos = stored - 4;
    memcpy(bm + 2336, in + pos, 4);
    uint32_t base = 2336;
    for (uint32_t range = 32; range <= 2048; range *= 8) {
        const uint8_t* ctl = bm + base;
        uint32_t changes = 0;
        for (uint32_t i = 0; i < range / 8; ++i) changes += (uint32_t)__builtin_popcount(ctl[i]);
        if (changes > pos) return fail("WZP RZE transition stream underflows");
        pos -= changes;
        uint32_t rd = pos;
        base -= range;
        uint8_t v = 0;
        for (uint32_t i = 0; i < range; ++i) {
            if ((ctl[i >> 3] >> (i & 7)) & 1u) {
                const uint8_t next = in[rd++];
                if (next == v) return fail("WZP redundant RZE transition");
                v = next;
            }
            bm[base + i] = v;
        }
    }
    uint32_t nonzero = 0;
    for (uint32_t i = 0; i < 2048; ++i) nonzero += (uint32_t)__builtin_popcount(bm[i]);
    if (nonzero == 0 || nonzero != pos) return fail("WZP RZE value stream length mismatch");
    const uint32_t full = raw_bytes / 8, part = raw_bytes & 7;
    if (part && (bm[full] >> part)) return fail("WZP nonzero partial-chunk RZE padding");
    for (uint32_t i = full + (part != 0); i < 2048; ++i)
        if (bm[i]) return fail("WZP nonzero partial-chunk RZE padding");
    uint32_t rd = 0;
    for (uint32_t i = 0; i < raw_bytes; ++i) {
        uint8_t v = 0;
        if ((bm[i >> 3] >> (i & 7)) & 1u) {
            v = in[rd++];
            if (!v) return fail("WZP zero value in sparse RZE stream");
        }
        out[i] = v;
    }
    return 0;
}

/* Encoder: cpu/src/wzp_cpu.cpp:769-867. Meta = code (bits 0-13: 0 zero chunk,
 * 0x3fff raw, else RZE bytes; raw iff RZE+2 >= 2*elements) | sign code << 14
 * (0 means a full 1024-byte slot). Signs of nonzero magnitudes, LSB first. */
int64_t wo_wzp_encode(const uint16_t* mag, const uint32_t* sign, uint32_t n, uint8_t* out, size_t capacity) {
    wzp_shape s;
    if (wzp_geometry(n, &s)) return -1;
    if (n == 0) return 0;
    if (!mag || !sign || !out) return fail("WZP null encode buffer");
    if (capacity < s.header) return fail("WZP output capacity is too small");
    uint8_t transformed[2 * CHUNK];
    uint8_t* meta = out + (size_t)s.groups * 4u;
    size_t cursor = s.header;
    for (uint32_t g = 0; g < s.groups; ++g) {
        put32(out + 4 * (size_t)g, 0);
        int seen = 0;
        const uint32_t first = g * GROUP, last = min_u(s.chunks, first + GROUP);
        for (uint32_t c = first; c < last; ++c) {
            const uint32_t elements = chunk_len(n, c), base = c * CHUNK;
            const uint16_t* src = mag + base;
            uint32_t nnz = 0;
            for (uint32_t i = 0; i < elements; ++i) nnz += src[i] != 0;
            if (!nnz) { put24(meta + 3 * (size_t)c, 0); continue; }
            const uint32_t raw_bytes = 2 * elements;
            bit2_encode(src, elements, transformed);
            const uint32_t rze = rze_encode(transformed, raw_bytes, NULL);
            const int raw = rze + 2 >= raw_bytes;
            const uint32_t stored = raw ? raw_bytes : rze;
            const uint32_t sign_bytes = (nnz + 7) / 8;
            const uint32_t chunk_bytes = stored + sign_bytes;
            if (chunk_bytes > capacity - cursor) return fail("WZP output capacity is too small");
            if (cursor - s.header > (size_t)UINT32_MAX - chunk_bytes) return fail("WZP payload exceeds its 32-bit directory");
            if (!seen) { put32(out + 4 * (size_t)g, (uint32_t)(cursor - s.header)); seen = 1; }
            if (raw) for (uint32_t i = 0; i < elements; ++i) put16(out + cursor + 2 * i, src[i]);
            else rze_encode(transformed, raw_bytes, out + cursor);
            uint8_t* ps = out + cursor + stored;
            memset(ps, 0, sign_bytes);
            uint32_t r = 0;
            for (uint32_t i = 0; i < elements; ++i) {
                if (!src[i]) continue;
                const uint32_t gi = base + i;
                ps[r >> 3] |= (uint8_t)(((sign[gi >> 5] >> (gi & 31)) & 1u) << (r & 7));
                ++r;
            }
            const uint32_t sign_code = sign_bytes == SIGN_SLOT ? 0 : sign_bytes;
            put24(meta + 3 * (size_t)c, (raw ? RAW_CODE : stored) | (sign_code << 14));
            cursor += chunk_bytes;
        }
    }
    return (int64_t)cursor;
}

typedef struct { uint32_t magnitude, signs; int raw; } chunk_size;
static int inspect_meta(uint32_t m, uint32_t elements, chunk_size* cs) {
    const uint32_t code = m & RAW_CODE, sign_code = m >> 14;
    if (code == 0) {
        if (sign_code) return fail("WZP zero chunk carries sign metadata");
        cs->magnitude = cs->signs = 0;
        cs->raw = 0;
        return 0;
    }
    const uint32_t raw_bytes = 2 * elements;
    cs->raw = code == RAW_CODE;
    if (!cs->raw && (code < 4 || code + 2 >= raw_bytes)) return fail("WZP invalid compressed magnitude length");
    cs->signs = sign_code ? sign_code : SIGN_SLOT;
    if (cs->signs > (elements + 7) / 8) return fail("WZP sign stream exceeds the chunk size");
    cs->magnitude = cs->raw ? raw_bytes : code;
    return 0;
}

/* Decoder: cpu/src/wzp_cpu.cpp:720-761,868-949. Groups may appear in any
 * payload order as long as they tile the payload exactly. */
int wo_wzp_decode(const uint8_t* in, size_t bytes, uint32_t n, uint16_t* mag, uint32_t* sign) {
    wzp_shape s;
    if (wzp_geometry(n, &s)) return -1;
    if (n == 0) return bytes ? fail("WZP nonempty archive for zero elements") : 0;
    if (bytes < s.header || bytes > s.max_bytes) return fail("WZP archive size is outside its valid bounds");
    const uint8_t* meta = in + (size_t)s.groups * 4u;
    const uint8_t* payload = in + s.header;
    const size_t payload_bytes = bytes - s.header;
    if (payload_bytes > UINT32_MAX) return fail("WZP payload exceeds its 32-bit directory");
    /* directory: group spans must tile the payload */
    uint64_t* span = (uint64_t*)malloc(sizeof(uint64_t) * s.groups);
    if (!span) return fail("out of memory");
    for (uint32_t g = 0; g < s.groups; ++g) {
        uint32_t len = 0;
        for (uint32_t c = g * GROUP; c < min_u(s.chunks, g * GROUP + GROUP); ++c) {
            chunk_size cs;
            if (inspect_meta(get24(meta + 3 * (size_t)c), chunk_len(n, c), &cs)) { free(span); return -1; }
            len += cs.magnitude + cs.signs;
        }
        const uint32_t off = get32(in + 4 * (size_t)g);
        if (off > payload_bytes || len > payload_bytes - off || (len == 0 && off != 0)) {
            free(span);
            return fail("WZP directory range is outside the payload");
        }
        span[g] = ((uint64_t)off << 32) | len;
    }
    if (s.groups == 1 && (span[0] >> 32) != 0) { free(span); return fail("WZP directory does not cover the payload"); }
    {   /* sort spans by offset (insertion order does not matter: offsets of
         * nonempty groups are distinct if they tile) */
        uint64_t cursor = 0;
        uint64_t* sorted = (uint64_t*)malloc(sizeof(uint64_t) * s.groups);
        if (!sorted) { free(span); return fail("out of memory"); }
        memcpy(sorted, span, sizeof(uint64_t) * s.groups);
        for (uint32_t i = 1; i < s.groups; ++i) {
            const uint64_t v = sorted[i];
            uint32_t j = i;
            while (j > 0 && (sorted[j - 1] >> 32) > (v >> 32)) { sorted[j] = sorted[j - 1]; --j; }
            sorted[j] = v;
        }
        for (uint32_t i = 0; i < s.groups; ++i) {
            const uint32_t len = (uint32_t)sorted[i];
            if (!len) continue;
            if ((sorted[i] >> 32) != cursor) { free(sorted); free(span); return fail("WZP payload groups overlap or leave a gap"); }
            cursor += len;
        }
        free(sorted);
        if (cursor != payload_bytes) { free(span); return fail("WZP trailing or missing payload bytes"); }
    }
    free(span);
    memset(sign, 0, ((size_t)n + 31) / 32 * 4);
    uint8_t expanded[2 * CHUNK];
    for (uint32_t g = 0; g < s.groups; ++g) {
        size_t cursor = get32(in + 4 * (size_t)g);
        for (uint32_t c = g * GROUP; c < min_u(s.chunks, g * GROUP + GROUP); ++c) {
            const uint32_t elements = chunk_len(n, c), base = c * CHUNK;
            chunk_size cs;
            if (inspect_meta(get24(meta + 3 * (size_t)c), elements, &cs)) return -1;
            uint16_t* dst = mag + base;
            if (!cs.magnitude) { memset(dst, 0, 2 * (size_t)elements); continue; }
            const uint8_t* src = payload + cursor;
            if (cs.raw) for (uint32_t i = 0; i < elements; ++i) dst[i] = get16(src + 2 * i);
            else {
                if (rze_decode(src, cs.magnitude, 2 * elements, expanded)) return -1;
                bit2_decode(expanded, elements, dst);
            }
            uint32_t nnz = 0;
            for (uint32_t i = 0; i < elements; ++i) nnz += dst[i] != 0;
            if (!nnz || (nnz + 7) / 8 != cs.signs) return fail("WZP compact sign length does not match nonzero magnitudes");
            const uint8_t* ps = src + cs.magnitude;
            if ((nnz & 7) && (ps[cs.signs - 1] >> (nnz & 7))) return fail("WZP nonzero compact sign padding");
            uint32_t r = 0;
            for (uint32_t i = 0; i < elements; ++i) {
                if (!dst[i]) continue;
                const uint32_t gi = base + i;
                sign[gi >> 5] |= (uint32_t)((ps[r >> 3] >> (r & 7)) & 1u) << (gi & 31);
                ++r;
            }
            cursor += cs.magnitude + cs.signs;
        }
    }
    return 0;
}

/* Largest archive: include/utils/Blob.hpp:80-86 with every side channel full,
 * cpu/src/codec.cpp:344. */
size_t wo_max_compressed_bytes(uint32_t nx, uint32_t ny, uint32_t nz, int width) {
    const uint64_t n = (uint64_t)nx * ny * nz;
    if (!n || (width != 4 && width != 8) || n > UINT32_MAX / (unsigned)width) return 0;
    const size_t w = wo_wzp_max_bytes((uint32_t)n);
    if (!w) return 0;
    return (size_t)layout(w, (uint32_t)n, (uint32_t)(n * (unsigned)width), (uint32_t)n).end;
}

uint64_t wo_fnv1a64(const void* data, size_t bytes) {
    const uint8_t* p = (const uint8_t*)data;
    uint64_t h = 1469598103934665603ull;
    for (size_t i = 0; i < bytes; ++i) { h ^= p[i]; h *= 1099511628211ull; }
    return h;
}

/* ---------------------------------------------------------------- typed halves */
#define T float
#define SFX f32
#define WIDTH 4
#define FMA fmaf
#define FABS fabsf
#define LLRINT llrintf
#include "oracle_impl.h"
#undef T
#undef SFX
#undef WIDTH
#undef FMA
#undef FABS
#undef LLRINT

#define T double
#define SFX f64
#define WIDTH 8
#define FMA fma
#define FABS fabs
#define LLRINT llrint
#include "oracle_impl.h"
#undef T
#undef SFX
#undef WIDTH
#undef FMA
#undef FABS
#undef LLRINT
