// Shared device/host definitions for the gfx950 Waltz codec.
//
// Arithmetic profile: every lifting step is an explicit correctly rounded fma
// (v_fma_f32 / v_fma_f64) and every other + and * is a separately rounded
// IEEE operation; the library is compiled with -ffp-contract=off, no fast-math
// and IEEE denormals, so device results are bitwise those of the reference CPU
// codec (cpu/src/transform.cpp:26-36,41-48).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#define WZ_HD __host__ __device__ __forceinline__
#define WZ_D __device__ __forceinline__

namespace wz {

// ------------------------------------------------------------------ geometry
struct Dims {
    uint32_t nx, ny, nz;
};

WZ_HD uint32_t half_len(uint32_t n) { return n / 2 + n % 2; }

// Whole-sample mirror with period 2n-2 (cpu/src/transform.cpp:60-66). Inputs
// lie within a few tiles of the axis, so the modulo path is cold.
WZ_HD uint32_t reflect_pos(int32_t p, uint32_t n) {
    if (p >= 0 && uint32_t(p) < n) return uint32_t(p);
    if (n <= 1) return 0;
    const int32_t period = 2 * int32_t(n) - 2;
    int32_t r = p % period;
    if (r < 0) r += period;
    return uint32_t(r < int32_t(n) ? r : period - r);
}
// Split-band position of a reflected sample (cpu/src/transform.cpp:68-71).
WZ_HD uint32_t split_pos(int32_t p, uint32_t n) {
    const uint32_t r = reflect_pos(p, n);
    return (r >> 1) + ((r & 1u) ? half_len(n) : 0u);
}

// Level count per axis group (include/utils/Blob.hpp:22-29).
WZ_HD unsigned dwt_levels(uint32_t n) {
    unsigned l = 0;
    while (n > 8 && l < 5) { n -= n / 2; ++l; }
    return l;
}

// Morton z band (cpu/src/codec.cpp:89-95).
WZ_HD uint32_t zband(Dims d, unsigned lz, bool dyadic) {
    if (dyadic) {
        const uint32_t v = (d.nz + ((1u << lz) - 1u)) >> lz;
        return v < 16u ? v : 16u;
    }
    if (d.nz < 4) return d.nz;
    const uint64_t blocks = uint64_t((d.nx + 15u) / 16) * ((d.ny + 15u) / 16);
    const uint64_t full = uint64_t(d.nx / 16) * (d.ny / 16);
    return (blocks - full) * 32 >= blocks ? 4u : 2u;
}

// Recursive ceil-half octant rank inside one brick: z half, then y, then x
// (cpu/src/codec.cpp:96-107).
WZ_HD uint32_t morton_rank(uint32_t x, uint32_t y, uint32_t z, uint32_t nx, uint32_t ny, uint32_t nz) {
    uint32_t rank = 0;
    while (nx > 1 || ny > 1 || nz > 1) {
        const uint32_t lx = nx - nx / 2, ly = ny - ny / 2, lz = nz - nz / 2;
        const bool cx = x >= lx, cy = y >= ly, cz = z >= lz;
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

// ------------------------------------------------------------------ lifting constants
// CDF 9/7 lifting steps and scale, rounded to T (cpu/src/transform.cpp:41-48).
template <class T> struct Lift {
    static constexpr T a = T(-1.586134342059924);
    static constexpr T b = T(-0.052980118572961);
    static constexpr T g = T(0.882911075530934);
    static constexpr T d = T(0.443506852043971);
    static constexpr T k = T(1.1496043988602418);
    static constexpr T ik = T(1) / T(1.1496043988602418);
};

// Per-type |x|, round-to-nearest-even and copysign (the plain builtins take
// and return double, which would widen float work to f64 instructions).
WZ_HD float abs_t(float x) { return __builtin_fabsf(x); }
WZ_HD double abs_t(double x) { return __builtin_fabs(x); }
WZ_HD float rint_t(float x) { return __builtin_rintf(x); }
WZ_HD double rint_t(double x) { return __builtin_rint(x); }
WZ_HD float copysign_t(float m, float s) { return __builtin_copysignf(m, s); }
WZ_HD double copysign_t(double m, double s) { return __builtin_copysign(m, s); }

WZ_HD float fmadd(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
WZ_HD double fmadd(double a, double b, double c) { return __builtin_fma(a, b, c); }

// Exact |value - original| <= bound with a TwoDiff tie check
// (cpu/src/repair_avx2.cpp:48-56).
// Branch-free: the tie check's TwoDiff is computed either way.
WZ_D bool within_bound_exact(double value, double original, double bound) {
    const double high = value - original, magnitude = __builtin_fabs(high);
    const double bv = value - high, av = high + bv;
    const double br = bv - original, ar = value - av, low = ar + br;
    const bool tie_ok = !((high > 0 && low > 0) || (high < 0 && low < 0));
    return __builtin_isfinite(high) && !(magnitude > bound) && (magnitude < bound || tie_ok);
}

// One native repair decision (cpu/src/repair_avx2.cpp:58-74): returns whether
// `rec` needs a record (error in *err); clears *ok on a strict-bound failure.
template <class T>
WZ_D bool repair_decision(T rec, T orig, T native_bound, double bound, T* err, bool* ok) {
    T value = rec;
    const T e = value - orig;
    bool patch = false;
    if (!__builtin_isfinite(value) || !__builtin_isfinite(e)) *ok = false;
    if (abs_t(e) > native_bound) {
        patch = true;
        value = value - e;
    }
    if (!__builtin_isfinite(value) || !within_bound_exact(double(value), double(orig), bound)) *ok = false;
    *err = e;
    return patch;
}

// Quantize one coefficient and dequantize it again (cpu/src/quantize_avx2.cpp:45-62,
// cpu/src/codec.cpp:299-302): round to nearest even of v / quantum, magnitude
// saturated at INT32_MAX, zero stays +0. Callers that get here have screened
// the domain (finite, |v / quantum| < 2^63).
// The integer round trip is the identity on the rounded value except for zero
// (always +0) and magnitudes of 2^31 and more (T(INT32_MAX) with its sign), so
// it is done with two selects.
template <class T>
WZ_D T requantize(T v, T q, T inv) {
    const T rounded = rint_t(v * inv);
    const T sat = T(INT32_MAX);  // INT32_MAX in T: 2^31 for float, exact for double
    const T m = abs_t(rounded) >= T(2147483648.0) ? copysign_t(sat, rounded) : rounded;
    return rounded == T(0) ? T(0) : m * q;
}

// Segmented side list: each producer workgroup appends to its own segment of
// `cap` entries and records the count; a complete row-major flag bitmap lets
// the archive assembly find every entry's index-order position.
struct SegList {
    uint32_t* bitmap;   // ceil(n/32) words
    uint32_t* index;    // segments * cap
    void* payload;      // segments * cap of the payload type
    uint32_t* count;    // per segment
    uint32_t segments, cap;
};

// ------------------------------------------------------------------ WZP format
constexpr uint32_t kChunk = 8192;        // magnitudes per chunk
constexpr uint32_t kChunkBytes = 16384;  // raw magnitude bytes per chunk
constexpr uint32_t kGroup = 16;          // chunks per group
constexpr uint32_t kSignSlot = 1024;     // sign code 0 = full 1024-byte slot
constexpr uint32_t kChunkSlot = kChunkBytes + kSignSlot;  // largest chunk payload (a multiple of 16)
constexpr uint32_t kRawCode = 0x3fff;

struct WzpShape {
    uint32_t chunks, groups;
    uint64_t header;
};
WZ_HD WzpShape wzp_shape(uint32_t n) {
    WzpShape s;
    s.chunks = n == 0 ? 0u : 1u + (n - 1u) / kChunk;
    s.groups = (s.chunks + kGroup - 1u) / kGroup;
    s.header = uint64_t(s.groups) * 4u + uint64_t(s.chunks) * 3u;
    return s;
}

// Division by a divisor that stays fixed over a loop, by one multiply:
// floor(n / d) = (n * ceil(2^32 / d)) >> 32, exact while n * d < 2^32.
// ceil(2^32 / d) = floor((2^32 - 1) / d) + 1 for every d >= 1, a 32-bit
// division (a 64-bit one is a long software routine on the GPU).
struct FastDiv {
    uint32_t d;
    uint64_t m;
    WZ_HD explicit FastDiv(uint32_t divisor) : d(divisor), m(uint64_t(0xFFFFFFFFu / divisor) + 1u) {}
    WZ_HD uint32_t div(uint32_t n) const { return uint32_t((uint64_t(n) * m) >> 32); }
    WZ_HD uint32_t mod(uint32_t n) const { return n - div(n) * d; }
};

// ------------------------------------------------------------------ archive layout
// 32-byte little-endian header followed by 8-byte aligned sections
// (include/utils/Blob.hpp:38-79).
WZ_HD uint64_t pad8(uint64_t n) { return (n + 7) & ~uint64_t(7); }
struct Layout {
    uint64_t pi, pv, qi, qv, end;
};
WZ_HD Layout archive_layout(uint64_t wzp, uint32_t np, uint64_t repair_bytes, uint32_t nq) {
    Layout L;
    L.pi = 32 + pad8(wzp);
    L.pv = L.pi + pad8(uint64_t(np) * 4);
    L.qi = L.pv + pad8(repair_bytes);
    L.qv = L.qi + pad8(uint64_t(nq) * 4);
    L.end = L.qv + pad8(uint64_t(nq) * 4);
    return L;
}

// ------------------------------------------------------------------ decode plan
// A decode launched before its header reaches the host: what the launch
// assumes of the archive (the geometry's canonical flags, levels and z band)
// and the record the device fills from the header (cpu/src/codec.cpp:
// 279-300, include/utils/Blob.hpp:38-90): the header bytes, the sections'
// places, and `ok` when the header is valid, canonical for this launch and
// its sections end exactly at the archive's size. Every planned kernel does
// nothing unless `ok`; the host then checks the header it reads back.
struct PlanExpect {
    uint32_t n, width;
    uint8_t flags, lx, lz, bz;
};
struct DecodeRecord {
    uint32_t status, ok, np, nq;
    unsigned long long wzp, pi, pv, qi, qv;
    double quantum;
    uint8_t header[32];
};
constexpr uint64_t kRecordBytes = 256;  // device / pinned record size (either result record)
static_assert(sizeof(DecodeRecord) <= kRecordBytes, "decode record");

WZ_D void plan_decode(const uint8_t* a, uint64_t bytes, const PlanExpect& e, DecodeRecord* out) {
    DecodeRecord r{};
    if (bytes >= 32)
        for (int k = 0; k < 32; ++k) r.header[k] = a[k];
    const uint8_t* h = r.header;
    auto rd32 = [&](int o) {
        return uint32_t(h[o]) | (uint32_t(h[o + 1]) << 8) | (uint32_t(h[o + 2]) << 16) | (uint32_t(h[o + 3]) << 24);
    };
    const uint64_t wzp = uint64_t(rd32(16)) | (uint64_t(rd32(20)) << 32);
    const uint32_t pb = rd32(4), np = rd32(24), nq = rd32(28);
    const double quantum = __longlong_as_double(
        static_cast<long long>(uint64_t(rd32(8)) | (uint64_t(rd32(12)) << 32)));
    const double qn = e.width == 4 ? double(float(quantum)) : quantum;
    const uint64_t header = wzp_shape(e.n).header;
    bool ok = bytes >= 32 && h[0] == e.flags && h[1] == e.lx && h[2] == e.lz && h[3] == e.bz;
    ok = ok && quantum > 0 && __builtin_isfinite(quantum) && qn > 0 && __builtin_isfinite(qn);
    ok = ok && np <= e.n && nq <= e.n && uint64_t(pb) == uint64_t(np) * e.width;
    ok = ok && wzp >= header && wzp <= header + 2ull * e.n + (e.n + 7ull) / 8 && wzp <= bytes - 32;
    if (ok) {
        const Layout L = archive_layout(wzp, np, pb, nq);
        ok = L.end == bytes;
        r.np = np;
        r.nq = nq;
        r.wzp = wzp;
        r.pi = L.pi;
        r.pv = L.pv;
        r.qi = L.qi;
        r.qv = L.qv;
        r.quantum = quantum;
    }
    if (!ok) r.np = r.nq = 0;
    r.ok = ok;
    *out = r;
}

// Device status bits shared by all kernels of one call.
enum Status : uint32_t {
    kNonFiniteInput = 1u << 0,
    kQuantDomain = 1u << 1,
    kRepairBound = 1u << 2,
    kWzpCapacity = 1u << 3,
    kLookbackTimeout = 1u << 4,
    kCorrupt = 1u << 5,
    kNonFiniteOutput = 1u << 6,
    kSideCapacity = 1u << 7,
    kUnordered = 1u << 8,  // a side list is not strictly increasing: decode again on the general path
};

}  // namespace wz
