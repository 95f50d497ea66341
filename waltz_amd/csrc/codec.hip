// Whole-field compress / decompress pipelines on one device and stream.
//
// Compress (cpu/src/codec.cpp:177-258 order):
//   forward lifting (input screened for non-finite samples on its first pass)
//   -> brick quantization (magnitudes, signs, dequantized coefficients, q-outliers)
//   -> WZP stream written straight into the archive at offset 32
//   -> inverse lifting of the dequantized coefficients
//   -> native repair detection against the input
//   -> side channels ordered by index into their sections, header, padding.
// One 48-byte status/result read back at the end decides success.
//
// Decompress (cpu/src/codec.cpp:276-336): 32-byte header read and validated on
// the host, WZP decode, dequantization into row-major coefficients, q-outlier
// patches, inverse lifting into the caller's output, repair patches, finite
// screen.
//
// Scratch lives in one cached, grow-only device allocation per device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <vector>

#include "common.hpp"
#include "fused.hpp"
#include "launch.hpp"
#include "pairbrick.hpp"
#include "profile.hpp"

namespace wz {

__global__ void zero_kernel(uint32_t* __restrict__ p, uint64_t words) {
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < words; i += stride) p[i] = 0;
}

void zero_words(void* p, uint64_t words, hipStream_t s) {
    if (!words) return;
    const uint64_t blocks = (words + 255) / 256;
    hipLaunchKernelGGL(zero_kernel, dim3(uint32_t(blocks < 4096 ? blocks : 4096)), dim3(256), 0, s,
                       static_cast<uint32_t*>(p), words);
}

namespace {

struct DeviceWorkspace {
    std::mutex lock;
    void* base = nullptr;
    size_t capacity = 0;
    uint16_t* rank = nullptr;  // 8*8192 forward + 8*8192 inverse
    Dims rank_dims{0, 0, 0};
    uint32_t rank_bz = 0;
    EncodeResult* result = nullptr;      // device
    EncodeResult* host_result = nullptr;  // pinned
    SideStream side{};                    // the WZP group pass beside the inverse levels (default priority)

    ~DeviceWorkspace() = default;  // process-lifetime cache

    void* reserve(size_t bytes) {
        if (bytes <= capacity) return base;
        if (base) hip_check(hipFree(base), "hipFree");
        base = nullptr;
        capacity = 0;
        hip_check(hipMalloc(&base, bytes), "Waltz workspace allocation");
        capacity = bytes;
        return base;
    }
    void init() {
        if (!rank) {
            hip_check(hipMalloc(reinterpret_cast<void**>(&rank), 2 * 8 * 8192 * sizeof(uint16_t)), "rank tables");
            // one record for compress (EncodeResult) or decode (DecodeRecord)
            hip_check(hipMalloc(reinterpret_cast<void**>(&result), kRecordBytes), "result");
            hip_check(hipHostMalloc(reinterpret_cast<void**>(&host_result), kRecordBytes), "pinned result");
            hip_check(hipStreamCreateWithFlags(&side.stream, hipStreamNonBlocking), "side stream");
            hip_check(hipEventCreateWithFlags(&side.ready, hipEventDisableTiming), "event");
            hip_check(hipEventCreateWithFlags(&side.done, hipEventDisableTiming), "event");
            hip_check(hipEventCreateWithFlags(&side.mid, hipEventDisableTiming), "event");
        }
    }
    const uint16_t* ranks(Dims d, uint32_t bz, hipStream_t s) {
        if (!(rank_dims.nx == d.nx && rank_dims.ny == d.ny && rank_dims.nz == d.nz && rank_bz == bz)) {
            build_rank_tables(d, bz, rank, rank + 8 * 8192, s);
            rank_dims = d;
            rank_bz = bz;
        }
        return rank;
    }
};

DeviceWorkspace& workspace() {
    static std::mutex m;
    static std::vector<std::unique_ptr<DeviceWorkspace>> all;
    int dev = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    std::lock_guard<std::mutex> g(m);
    if (all.size() <= size_t(dev)) all.resize(size_t(dev) + 1);
    if (!all[size_t(dev)]) all[size_t(dev)] = std::make_unique<DeviceWorkspace>();
    return *all[size_t(dev)];
}

// Bump allocator over the workspace with 256-byte alignment.
struct Carve {
    uint8_t* base;
    size_t used = 0;
    template <class U>
    U* take(uint64_t count) {
        U* p = reinterpret_cast<U*>(base + used);
        used += (count * sizeof(U) + 255) & ~size_t(255);
        return p;
    }
};

struct Geometry {
    Dims d;
    uint32_t n;
    unsigned lx, lz;
    bool dyadic;
    uint32_t bz;
};

Geometry geometry(Dims d, unsigned width) {
    if (!d.nx || !d.ny || !d.nz) throw std::invalid_argument("nonzero dimensions required");
    const uint64_t n = uint64_t(d.nx) * d.ny * d.nz;
    if (n > 0xFFFFFFFFull / width) throw std::invalid_argument("volume exceeds WLT2 section limits");
    if (n > 0xFFFFFFFFull / 2) throw std::invalid_argument("WZP element count exceeds the GPU format limit");
    Geometry g;
    g.d = d;
    g.n = uint32_t(n);
    g.lx = dwt_levels(d.nx < d.ny ? d.nx : d.ny);
    g.lz = dwt_levels(d.nz);
    g.dyadic = g.lx == g.lz;
    g.bz = zband(d, g.lz, g.dyadic);
    return g;
}

[[noreturn]] void raise_status(uint32_t st) {
    if (st & kNonFiniteInput) throw std::runtime_error("nonfinite input is unsupported");
    if (st & kQuantDomain) throw std::runtime_error("coefficient exceeds portable integer conversion domain");
    if (st & kRepairBound) throw std::runtime_error("native repair cannot meet strict bound or nonfinite reconstruction");
    if (st & kSideCapacity) throw std::runtime_error("portable GPU side-channel capacity reached");
    if (st & kWzpCapacity) throw std::runtime_error("archive capacity too small");
    if (st & kLookbackTimeout) throw std::runtime_error("WZP lookback did not complete");
    if (st & kCorrupt) throw std::runtime_error("corrupt Waltz archive");
    if (st & kNonFiniteOutput) throw std::runtime_error("nonfinite reconstruction");
    throw std::runtime_error("Waltz device status error");
}

uint32_t qcap(uint32_t n) { return n <= (1u << 20) ? n : (1u << 20); }

// An early exit (exception) must not leave side-stream work of this call
// running into the next one.
struct SideJoin {
    const SideStream* sd;
    ~SideJoin() {
        if (sd) (void)hipStreamSynchronize(sd->stream);
    }
};

// Caller buffers must not overlap (cpu/src/codec.cpp:82-86).
void check_disjoint(const void* a, uint64_t as, const void* b, uint64_t bs) {
    const uintptr_t x = reinterpret_cast<uintptr_t>(a), y = reinterpret_cast<uintptr_t>(b);
    if (as > UINTPTR_MAX - x || bs > UINTPTR_MAX - y) throw std::invalid_argument("pointer range overflow");
    if (x < y + bs && y < x + as) throw std::invalid_argument("input and output must not overlap");
}

}  // namespace

uint64_t max_compressed_bytes(Dims d, unsigned width) {
    const Geometry g = geometry(d, width);
    const WzpShape ws = wzp_shape(g.n);
    const uint64_t wzp = ws.header + uint64_t(g.n) * 2 + (uint64_t(g.n) + 7) / 8;
    return archive_layout(wzp, g.n, uint64_t(g.n) * width, g.n).end;
}

template <class T>
uint64_t compress(const T* input, Dims d, double abs_bound, uint8_t* archive, uint64_t capacity, hipStream_t s) {
    if (!(abs_bound > 0) || !std::isfinite(abs_bound))
        throw std::invalid_argument("positive finite absolute error required");
    if (reinterpret_cast<uintptr_t>(archive) % 8 || reinterpret_cast<uintptr_t>(input) % sizeof(T))
        throw std::invalid_argument("archive must be 8-byte aligned and input naturally aligned");
    const Geometry g = geometry(d, sizeof(T));
    const double quantum = 1.5 * abs_bound;
    const T q = T(quantum), inv = T(1.0 / quantum);
    if (!(q > T(0)) || !std::isfinite(q) || !std::isfinite(inv))
        throw std::invalid_argument("quantum not representable in native type");
    if (capacity < 32) throw std::runtime_error("archive capacity too small");
    check_disjoint(input, uint64_t(g.n) * sizeof(T), archive, capacity);
    const uint32_t n = g.n;
    const uint64_t words = (uint64_t(n) + 31) / 32;
    const WzpShape ws = wzp_shape(n);
    const uint32_t nq = qcap(n);
    const uint32_t rtiles = (n + kRepairTile - 1) / kRepairTile;
    uint32_t fsegs = 0, fcap = 0;  // repair segments of a fused finest inverse level
    level0_repair_segments(g.d, g.lx, g.dyadic, sizeof(T), &fsegs, &fcap);
    const uint64_t rslots = std::max<uint64_t>(uint64_t(rtiles) * kRepairTile, uint64_t(fsegs) * fcap);
    const uint32_t rsegs = std::max(rtiles, fsegs);

    DeviceWorkspace& w = workspace();
    std::lock_guard<std::mutex> guard(w.lock);
    w.init();
    // layout of the per-call scratch
    // brick-pair fields quantize inside the encoder: no sign words to clear
    const bool pairs = pair_bricks_fit(g.d.nx, g.d.ny, g.d.nz, g.bz);
    // other fields of full bricks quantize inside the encoder too
    const bool ranks = rank_bricks_fit(g.d.nx, g.d.ny, g.d.nz, g.bz);
    const bool fusedq = pairs || ranks;
    auto plan = [&](Carve& c) {
        struct P {
            T *coeff, *t0, *t1, *rec;
            uint8_t* stage;
            uint16_t* mag;
            uint32_t *sign, *qbm, *rbm, *qidx, *qval, *ridx, *rcount, *seg, *zeroed;
            T* rerr;
            unsigned long long* desc;
            uint32_t *size, *counters;
        } p;
        p.coeff = c.take<T>(n);
        p.t0 = c.take<T>(n);
        p.t1 = c.take<T>(n);
        // the reconstruction of an unfused finest level and the encoder's
        // staging slots: one region with a fused finest level (which never
        // writes the reconstruction), else two, so that the WZP group pass
        // can run beside the inverse
        if (fsegs) {
            p.rec = reinterpret_cast<T*>(
                c.take<uint8_t>(std::max<uint64_t>(uint64_t(n) * sizeof(T), wzp_staging_bytes(ws.chunks))));
            p.stage = reinterpret_cast<uint8_t*>(p.rec);
        } else {
            p.rec = c.take<T>(n);
            p.stage = c.take<uint8_t>(wzp_staging_bytes(ws.chunks));
        }
        p.mag = c.take<uint16_t>(n);
        p.ridx = c.take<uint32_t>(rslots);
        p.rerr = c.take<T>(rslots);
        p.rcount = c.take<uint32_t>(rsegs);
        p.rbm = c.take<uint32_t>(words);
        p.qidx = c.take<uint32_t>(nq);
        p.qval = c.take<uint32_t>(nq);
        p.seg = c.take<uint32_t>(finish_scratch_words(n));
        p.size = c.take<uint32_t>(ws.chunks);
        // zero-initialised region: sign, q bitmap, lookback descriptors, counters
        const size_t z0 = c.used;
        p.sign = c.take<uint32_t>(fusedq ? 0 : words);
        p.qbm = c.take<uint32_t>(words);
        p.desc = c.take<unsigned long long>(ws.chunks);
        p.counters = c.take<uint32_t>(4);  // ticket, q count, repair total
        p.zeroed = reinterpret_cast<uint32_t*>(c.base + z0);
        return std::make_pair(p, c.used - z0);
    };
    Carve probe{nullptr};
    plan(probe);
    Carve carve{static_cast<uint8_t*>(w.reserve(probe.used))};
    auto [p, zero_bytes] = plan(carve);

    EncodeResult* res = w.result;
    uint32_t* status = &res->status;
    zero_words(res, sizeof(EncodeResult) / 4, s);
    zero_words(p.zeroed, zero_bytes / 4, s);

    transform_forward<T>(input, p.coeff, p.t0, p.t1, g.d, g.lx, g.lz, g.dyadic, status, s);
    SideList qo{p.qbm, p.qidx, p.qval, p.counters + 1, nq};
    WzpScratch wsc{p.desc, p.stage, p.size, p.counters};
    // with the fused repair emitter and a quantizing encoder the finest
    // level's high bands stay unquantized in coeff; that kernel requantizes
    // them on load
    const bool requant = fsegs && fusedq;
    // The group pass runs on the side stream beside the inverse levels (the
    // staging slots it reads are never written by them); the archive sections
    // wait for it.
    const SideStream* side = &w.side;
    const SideJoin join{side};
    // When the coarser inverse levels requantize what they load as well (the
    // two leading levels fused), the encoder writes nothing back. They run
    // after the encoder, beside its group pass: beside the encoder they slowed
    // it as much as they hid, and the group pass then collided with the
    // finest level, which needs every CU's two workgroup slots.
    const bool split = requant && coarse_split_fits(g.d, g.lx, sizeof(T));
    // a plane-schedule field whose inverse runs as fused XY levels quantizes
    // and dequantizes the coefficients as they load: nothing written back
    const bool plane_requant = !g.dyadic && fusedq && plane_fused_fits(g.d, g.lx, g.lz, sizeof(T));
    const WriteBack wb = split || plane_requant ? kWriteNone : requant ? kWriteCoarseBox : kWriteAll;
    if (pairs) {
        StageClock c("quantize_encode", s);
        quantize_encode_pairs<T>(p.coeff, g.d, quantum, qo, archive + 32, capacity - 32, wsc, &res->wzp_bytes,
                                 status, s, wb, side);
    } else if (ranks) {
        const uint16_t* rank = w.ranks(g.d, g.bz, s);
        StageClock c("quantize_encode", s);
        quantize_encode_ranks<T>(p.coeff, g.d, g.bz, rank, quantum, qo, archive + 32, capacity - 32, wsc,
                                 &res->wzp_bytes, status, s, wb, side);
    } else {
        const BrickGeometry bg = brick_geometry(g.d, g.bz);
        const uint16_t* rank = w.ranks(g.d, g.bz, s);
        {
            StageClock c("quantize", s);
            quantize_bricks<T>(p.coeff, bg, quantum, rank, p.mag, p.sign, qo, status, s);
        }
        StageClock c("wzp_encode", s);
        wzp_encode(p.mag, p.sign, n, archive + 32, capacity - 32, wsc, &res->wzp_bytes, status, s, side);
    }
    // finest inverse level with native repair detection fused in when that
    // level runs as a fused kernel; otherwise reconstruct, then a repair pass
    Level0Emit<T> emit{kEmitRepair, status, input, abs_bound, SegList{p.rbm, p.ridx, p.rerr, p.rcount, fsegs, fcap},
                       requant ? quantum : 0.0};
    SegList rp = emit.rep;
    if (split) {
        inverse_coarse_requant<T>(p.coeff, p.t0, p.t1, g.d, g.lx, quantum, s);
        inverse_finest_requant<T>(p.coeff, p.t0, g.d, g.lx, emit, s);
    } else if (!transform_inverse<T>(p.coeff, p.rec, p.t0, p.t1, g.d, g.lx, g.lz, g.dyadic, s,
                                     fsegs ? &emit : nullptr, plane_requant ? quantum : 0.0)) {
        rp = SegList{p.rbm, p.ridx, p.rerr, p.rcount, rtiles, kRepairTile};
        StageClock c("repairs", s);
        detect_repairs<T>(input, p.rec, n, abs_bound, rp, status, s);
    }
    if (side) hip_check(hipStreamWaitEvent(s, side->done, 0), "stream wait");
    const uint8_t flags = uint8_t((g.dyadic ? 1u : 0u) | (sizeof(T) == 8 ? 4u : 0u) | (3u << 3));
    {
        StageClock c("sections", s);
        finish_archive<T>(archive, capacity, n, rp, qo, p.seg, flags, uint8_t(g.lx), uint8_t(g.lz), uint8_t(g.bz),
                          quantum, res, s);
    }
    hip_check(hipMemcpyAsync(w.host_result, res, sizeof(EncodeResult), hipMemcpyDeviceToHost, s), "result copy");
    hip_check(hipStreamSynchronize(s), "compress");
    hip_check(hipGetLastError(), "compress launch");
    profile_collect();
    const EncodeResult r = *w.host_result;
    if (r.status) raise_status(r.status);
    return r.end;
}

uint64_t archive_size_from_header(const uint8_t* h, Dims d, unsigned width) {
    const Geometry g = geometry(d, width);
    const uint8_t f = h[0];
    if ((f & 0xE0u) || !(f & 0x18u) || ((f & 4u) ? 8u : 4u) != width)
        throw std::runtime_error("invalid Waltz archive header");
    if (((f >> 3) & 3u) != 3u) throw std::runtime_error("unsupported Waltz synthesis mode");
    if (f & 2u) throw std::runtime_error("compact repairs are not part of the canonical profile");
    const unsigned lx = h[1], lz = h[2], bz = h[3];
    const bool dyadic = f & 1u;
    if (lx > 5 || lz > 5 || bz < 1 || bz > 32 || (dyadic && lx != lz))
        throw std::runtime_error("invalid levels or z band");
    if (lx != g.lx || lz != g.lz) throw std::runtime_error("dimensions/levels mismatch");
    auto rd32 = [&](int o) {
        return uint32_t(h[o]) | (uint32_t(h[o + 1]) << 8) | (uint32_t(h[o + 2]) << 16) | (uint32_t(h[o + 3]) << 24);
    };
    const uint64_t wzp = uint64_t(rd32(16)) | (uint64_t(rd32(20)) << 32);
    const uint32_t pb = rd32(4), np = rd32(24), nq = rd32(28);
    double quantum;
    const uint64_t qb = uint64_t(rd32(8)) | (uint64_t(rd32(12)) << 32);
    std::memcpy(&quantum, &qb, 8);
    const double qn = width == 4 ? double(float(quantum)) : quantum;
    if (!(quantum > 0) || !std::isfinite(quantum) || !(qn > 0) || !std::isfinite(qn))
        throw std::runtime_error("invalid quantum");
    if (np > g.n || nq > g.n) throw std::runtime_error("invalid section counts");
    if (uint64_t(pb) != uint64_t(np) * width) throw std::runtime_error("invalid native correction section");
    if (wzp > (1ull << 40)) throw std::runtime_error("invalid section counts");
    return archive_layout(wzp, np, pb, nq).end;
}

// Scratch of one decode: coefficient volume, two lifting scratch volumes,
// magnitudes and sign words, each 256-byte aligned.
uint64_t decode_scratch_bytes(uint32_t n, unsigned width) {
    Carve probe{nullptr};
    probe.take<uint8_t>(uint64_t(n) * width);
    probe.take<uint8_t>(uint64_t(n) * width);
    probe.take<uint8_t>(uint64_t(n) * width);
    probe.take<uint16_t>(n);
    probe.take<uint32_t>((uint64_t(n) + 31) / 32);
    return probe.used;
}

namespace {

// Where one decode's memory comes from: the device scratch (decode_scratch_bytes),
// the rank tables for geometry d and band bz, the device record and a host
// copy of it.
struct DecodeSource {
    uint8_t* scratch;
    DecodeRecord* result;
    DecodeRecord* host;
    DeviceWorkspace* cache;  // cached rank tables and the side stream, or nullptr: ranks rebuilt into `rank`
    uint16_t* rank;
    const uint16_t* ranks(Dims d, uint32_t bz, hipStream_t s) const {
        if (cache) return cache->ranks(d, bz, s);
        build_rank_tables(d, bz, rank, rank + 8 * 8192, s);
        return rank;
    }
};

// One decode of a validated archive into `output`; returns the device status.
// canonical: both side lists are taken as strictly increasing (the archives
// compress writes), and a list found out of order comes back as kUnordered;
// otherwise both lists are first checked for duplicate indices.
template <class T>
uint32_t decode_pass(const uint8_t* archive, const uint8_t* hdr, const Geometry& g, T* output, bool canonical,
                     const DecodeSource& m, hipStream_t s) {
    const uint32_t n = g.n;
    const uint32_t bz = hdr[3];
    // the archive's own schedule: a plane-schedule archive with equal level
    // counts is valid (cpu/src/codec.cpp:279-283)
    const bool dyadic = (hdr[0] & 1u) != 0;
    uint64_t wzp = 0;
    for (int k = 7; k >= 0; --k) wzp = (wzp << 8) | hdr[16 + k];
    auto rd32 = [&](int o) {
        return uint32_t(hdr[o]) | (uint32_t(hdr[o + 1]) << 8) | (uint32_t(hdr[o + 2]) << 16) |
               (uint32_t(hdr[o + 3]) << 24);
    };
    const uint32_t np = rd32(24), nq = rd32(28);
    double quantum;
    const uint64_t qb = uint64_t(rd32(8)) | (uint64_t(rd32(12)) << 32);
    std::memcpy(&quantum, &qb, 8);
    const Layout L = archive_layout(wzp, np, uint64_t(np) * sizeof(T), nq);
    const uint64_t words = (uint64_t(n) + 31) / 32;
    Carve c{m.scratch};
    T* coeff = c.take<T>(n);
    T* t0 = c.take<T>(n);
    T* t1 = c.take<T>(n);
    uint16_t* mag = c.take<uint16_t>(n);
    uint32_t* sign = c.take<uint32_t>(words);
    uint32_t* status = &m.result->status;
    zero_words(m.result, sizeof(DecodeRecord) / 4, s);
    if (!canonical) {  // duplicate indices fail (cpu/src/codec.cpp:260-275); `sign` is free until decode
        check_unique(archive + L.qi, nq, n, sign, status, s);
        check_unique(archive + L.pi, np, n, sign, status, s);
    }
    Level0Emit<T> screen{kEmitStoreScreen, status, nullptr, 0.0, SegList{}, 0.0};
    // a canonical repair list is applied as the finest fused level stores its
    // samples (its key table in the magnitude scratch, free once the
    // coefficients are dequantized); a list out of order comes back as
    // kUnordered, and the general pass patches separately
    const uint64_t keys = canonical && dyadic ? patch_table_words(g.d, np, sizeof(T)) : 0;
    if (keys) {
        screen.pidx = reinterpret_cast<const uint32_t*>(archive + L.pi);
        screen.perr = reinterpret_cast<const T*>(archive + L.pv);
        screen.np = np;
        screen.ptable = reinterpret_cast<uint32_t*>(mag);
    }
    // the WZP directory check runs on the side stream beside the chunk decoders
    const SideStream* side = m.cache ? &m.cache->side : nullptr;
    const SideJoin join{side};
    const bool pairs = pair_bricks_fit(g.d.nx, g.d.ny, g.d.nz, bz);
    const uint16_t* rank = pairs ? nullptr : m.ranks(g.d, bz, s);
    const bool ranks = rank_bricks_fit(g.d.nx, g.d.ny, g.d.nz, bz) && reinterpret_cast<uintptr_t>(rank) % 8 == 0;
    if (pairs || ranks) {
        StageClock ck("decode_dequantize", s);
        if (pairs) decode_dequantize_pairs<T>(archive + 32, wzp, g.d, quantum, coeff, status, s, side);
        else decode_dequantize_ranks<T>(archive + 32, wzp, g.d, bz, rank, quantum, coeff, status, s, side);
        if (keys && side) {
            // the key table reads only the archive, and the dequantizing
            // decoders leave the magnitude scratch unused: built on the side
            // stream after the directory check, beside the chunk decoders
            repair_keys<T>(screen.pidx, screen.perr, np, g.d, screen.ptable, status, side->stream);
            hip_check(hipEventRecord(side->mid, side->stream), "event record");
            screen.ptable_built = true;
        }
    } else {
        {
            StageClock ck("wzp_decode", s);
            wzp_decode(archive + 32, wzp, n, mag, sign, status, s, side);
        }
        const BrickGeometry bg = brick_geometry(g.d, bz);
        StageClock ck("dequantize", s);
        dequantize_bricks<T>(mag, sign, bg, quantum, rank, coeff, s);
    }
    {
        StageClock ck("qoutliers", s);
        apply_qoutliers<T>(coeff, archive + L.qi, archive + L.qv, nq, n, quantum, status, s);
    }
    if (screen.ptable_built) hip_check(hipStreamWaitEvent(s, side->mid, 0), "stream wait");
    const bool screened = transform_inverse<T>(coeff, output, t0, t1, g.d, g.lx, g.lz, dyadic, s, &screen);
    {
        StageClock ck("patches", s);
        if (!screened || !keys)  // screens patched samples
            apply_repairs<T>(output, archive + L.pi, archive + L.pv, np, n, status, s);
        if (!screened) check_finite<T>(output, n, status, s);
    }
    if (side) hip_check(hipStreamWaitEvent(s, side->done, 0), "stream wait");
    hip_check(hipMemcpyAsync(m.host, m.result, sizeof(DecodeRecord), hipMemcpyDeviceToHost, s), "status copy");
    hip_check(hipStreamSynchronize(s), "decompress");
    hip_check(hipGetLastError(), "decompress launch");
    return m.host->status;
}

// A decode planned on the device: a kernel checks the header into the
// device record and every stage reads the stream size, quantum and side
// lists from it, so the call synchronises once, at its end, instead of
// first waiting for the header. It assumes the canonical archive of the
// geometry (the one compress writes); any other header leaves every planned
// kernel idle (record not ok) and the caller decodes on the host-header path.
// Applies to dyadic fields whose finest inverse level runs fused (it applies
// the repairs, so every count the launch needs comes from the record), and to
// plane-schedule fields whose inverse runs as fused XY levels and one Z phase
// (which screens the output; the repairs follow in a pass of their own).
bool planned_decode_fits(const Geometry& g, uint64_t archive_bytes, unsigned width) {
    if (archive_bytes == ~uint64_t(0) || archive_bytes < 32) return false;
    if (!g.dyadic) return plane_fused_fits(g.d, g.lx, g.lz, width);
    return g.lx >= 1 && tuning_for(g.d, g.lx, width, true).fused >= 1 && patch_table_words(g.d, g.n, width) != 0;
}

template <class T>
uint32_t decode_planned(const uint8_t* archive, uint64_t archive_bytes, const Geometry& g, T* output,
                        const DecodeSource& m, hipStream_t s) {
    const uint32_t n = g.n;
    Carve c{m.scratch};
    T* coeff = c.take<T>(n);
    T* t0 = c.take<T>(n);
    T* t1 = c.take<T>(n);
    uint16_t* mag = c.take<uint16_t>(n);
    uint32_t* sign = c.take<uint32_t>((uint64_t(n) + 31) / 32);
    uint32_t* table = reinterpret_cast<uint32_t*>(mag);  // repair key table, once the magnitudes are dead
    DecodeRecord* rec = m.result;
    uint32_t* status = &rec->status;
    const uint8_t flags = uint8_t((g.dyadic ? 1u : 0u) | (sizeof(T) == 8 ? 4u : 0u) | (3u << 3));
    {
        StageClock ck("decode_plan", s);
        decode_plan(archive, archive_bytes,
                    PlanExpect{n, uint32_t(sizeof(T)), flags, uint8_t(g.lx), uint8_t(g.lz), uint8_t(g.bz)}, rec, s);
    }
    // Beside the chunk decoders, the side stream builds the repair key table
    // (the dequantizing decoders leave the magnitude scratch unused; it reads
    // only the archive), then checks the WZP directory; the caller's stream
    // waits for the table only before the finest level, after the coarse
    // levels.
    const SideStream* side = m.cache ? &m.cache->side : nullptr;
    const SideJoin join{side};
    const bool pairs = pair_bricks_fit(g.d.nx, g.d.ny, g.d.nz, g.bz);
    const uint16_t* rank = pairs ? nullptr : m.ranks(g.d, g.bz, s);
    const bool ranks = rank_bricks_fit(g.d.nx, g.d.ny, g.d.nz, g.bz) && reinterpret_cast<uintptr_t>(rank) % 8 == 0;
    const bool fold = g.dyadic;  // repairs folded into the finest fused level (else a pass of their own)
    const bool keys_on_side = fold && (pairs || ranks) && side;
    if (side) hip_check(hipEventRecord(side->ready, s), "event record");
    if (pairs || ranks) {
        StageClock ck("decode_dequantize", s);
        if (pairs) decode_dequantize_pairs_planned<T>(archive + 32, rec, g.d, coeff, status, s);
        else decode_dequantize_ranks_planned<T>(archive + 32, rec, g.d, g.bz, rank, coeff, status, s);
    } else {
        {
            StageClock ck("wzp_decode", s);
            wzp_decode_planned(archive + 32, rec, n, mag, sign, status, s);
        }
        StageClock ck("dequantize", s);
        dequantize_bricks<T>(mag, sign, brick_geometry(g.d, g.bz), 0.0, rank, coeff, s, rec);
    }
    if (side) {
        hipStream_t ss = side->stream;
        hip_check(hipStreamWaitEvent(ss, side->ready, 0), "stream wait");
        if (keys_on_side) {
            StageClock ck("patch_keys", ss);
            repair_keys_planned<T>(archive, rec, g.d, table, status, ss);
        }
        hip_check(hipEventRecord(side->mid, ss), "event record");
        wzp_directory_planned(archive + 32, rec, n, status, ss);
        hip_check(hipEventRecord(side->done, ss), "event record");
    } else {
        wzp_directory_planned(archive + 32, rec, n, status, s);
    }
    if (fold && !keys_on_side) {
        StageClock ck("patch_keys", s);
        repair_keys_planned<T>(archive, rec, g.d, table, status, s);
    }
    {
        StageClock ck("qoutliers", s);
        apply_qoutliers_planned<T>(coeff, archive, rec, n, status, s);
    }
    if (!fold) {
        Level0Emit<T> screen{kEmitStoreScreen, status, nullptr, 0.0, SegList{}, 0.0};
        const bool screened = transform_inverse<T>(coeff, output, t0, t1, g.d, g.lx, g.lz, false, s, &screen);
        {
            StageClock ck("patches", s);
            apply_repairs_planned<T>(output, archive, rec, n, status, s);
        }
        if (!screened) check_finite<T>(output, n, status, s);
        if (side) hip_check(hipStreamWaitEvent(s, side->done, 0), "stream wait");
        hip_check(hipMemcpyAsync(m.host, rec, sizeof(DecodeRecord), hipMemcpyDeviceToHost, s), "status copy");
        hip_check(hipStreamSynchronize(s), "decompress");
        hip_check(hipGetLastError(), "decompress launch");
        return m.host->status;
    }
    Level0Emit<T> screen{kEmitStoreScreen, status, nullptr, 0.0, SegList{}, 0.0};
    screen.np = n;  // an upper bound: the count comes from the record
    screen.ptable = table;
    screen.ptable_built = true;
    screen.plan = rec;
    screen.archive = archive;
    const LowBox<T> low = inverse_coarse_levels<T>(coeff, t0, t1, g.d, g.lx, s);
    if (keys_on_side) hip_check(hipStreamWaitEvent(s, side->mid, 0), "stream wait");
    inverse_finest_level<T>(coeff, low, output, g.d, g.lx, screen, s);
    if (side) hip_check(hipStreamWaitEvent(s, side->done, 0), "stream wait");
    hip_check(hipMemcpyAsync(m.host, rec, sizeof(DecodeRecord), hipMemcpyDeviceToHost, s), "status copy");
    hip_check(hipStreamSynchronize(s), "decompress");
    hip_check(hipGetLastError(), "decompress launch");
    return m.host->status;
}

// Header read and checks, then the decode (with the duplicate-checking pass
// when a side list is out of order). The scratch of `m` must hold
// decode_scratch_bytes(n, sizeof(T)).
template <class T>
void decompress_with(const uint8_t* archive, uint64_t archive_bytes, Dims d, T* output, hipStream_t s,
                     bool canonical_chunk, const DecodeSource& m) {
    const Geometry g = geometry(d, sizeof(T));
    uint8_t hdr[32];
    uint32_t st = 0;
    bool decoded = false;  // by the planned decode (its header was valid and canonical)
    if (planned_decode_fits(g, archive_bytes, sizeof(T))) {
        check_disjoint(archive, archive_bytes, output, uint64_t(g.n) * sizeof(T));
        st = decode_planned<T>(archive, archive_bytes, g, output, m, s);
        std::memcpy(hdr, m.host->header, 32);
        decoded = m.host->ok != 0;
    } else {
        uint8_t* h = m.host->header;
        hip_check(hipMemcpyAsync(h, archive, 32, hipMemcpyDeviceToHost, s), "header copy");
        hip_check(hipStreamSynchronize(s), "header copy");
        std::memcpy(hdr, h, 32);
    }
    const uint64_t wzp = uint64_t(hdr[16]) | (uint64_t(hdr[17]) << 8) | (uint64_t(hdr[18]) << 16) |
                         (uint64_t(hdr[19]) << 24) | (uint64_t(hdr[20]) << 32) | (uint64_t(hdr[21]) << 40) |
                         (uint64_t(hdr[22]) << 48) | (uint64_t(hdr[23]) << 56);
    const WzpShape ws = wzp_shape(g.n);
    if (canonical_chunk) {  // the chunk decoder takes canonical archives only (src/waltz.cu:631-646)
        const uint8_t flags = uint8_t((g.dyadic ? 1u : 0u) | (sizeof(T) == 8 ? 4u : 0u) | (3u << 3));
        if (hdr[0] != flags || hdr[1] != g.lx || hdr[2] != g.lz || hdr[3] != g.bz || wzp < ws.header ||
            wzp > archive_bytes - 32)
            throw std::runtime_error("invalid canonical Waltz chunk shape, header or length");
    }
    const uint64_t end = archive_size_from_header(hdr, d, sizeof(T));
    if (archive_bytes != ~uint64_t(0) && end != archive_bytes)
        throw std::runtime_error(canonical_chunk ? "invalid canonical Waltz chunk shape, header or length"
                                                 : "truncated archive or trailing data");
    check_disjoint(archive, end, output, uint64_t(g.n) * sizeof(T));
    if (wzp < ws.header || wzp > ws.header + uint64_t(g.n) * 2 + (uint64_t(g.n) + 7) / 8)
        throw std::runtime_error("WZP archive size is outside its valid bounds");
    if (!decoded) st = decode_pass<T>(archive, hdr, g, output, true, m, s);
    // a side list out of order (not written by compress, but valid when its
    // indices are distinct): decode again with the duplicate checks
    if (st & kUnordered) st = decode_pass<T>(archive, hdr, g, output, false, m, s) & ~uint32_t(kUnordered);
    profile_collect();
    if (st) raise_status(st);
}

}  // namespace

template <class T>
void decompress(const uint8_t* archive, uint64_t archive_bytes, Dims d, T* output, hipStream_t s,
                bool canonical_chunk) {
    if (reinterpret_cast<uintptr_t>(archive) % 8 || reinterpret_cast<uintptr_t>(output) % sizeof(T))
        throw std::invalid_argument("archive must be 8-byte aligned and output naturally aligned");
    const Geometry g = geometry(d, sizeof(T));
    DeviceWorkspace& w = workspace();
    std::lock_guard<std::mutex> guard(w.lock);
    w.init();
    const DecodeSource m{static_cast<uint8_t*>(w.reserve(decode_scratch_bytes(g.n, sizeof(T)))),
                         reinterpret_cast<DecodeRecord*>(w.result), reinterpret_cast<DecodeRecord*>(w.host_result), &w,
                         nullptr};
    decompress_with<T>(archive, archive_bytes, d, output, s, canonical_chunk, m);
}

template <class T>
void decompress_into(const uint8_t* archive, uint64_t archive_bytes, Dims d, T* output, hipStream_t s,
                     bool canonical_chunk, const DecodeMemory& mem) {
    if (reinterpret_cast<uintptr_t>(archive) % 8 || reinterpret_cast<uintptr_t>(output) % sizeof(T))
        throw std::invalid_argument("archive must be 8-byte aligned and output naturally aligned");
    const Geometry g = geometry(d, sizeof(T));
    if (decode_scratch_bytes(g.n, sizeof(T)) > mem.scratch_bytes)
        throw std::invalid_argument("field exceeds the decoder's fixed scratch");
    DecodeRecord host{};
    const DecodeSource m{mem.scratch, mem.result, &host, nullptr, mem.rank};
    decompress_with<T>(archive, archive_bytes, d, output, s, canonical_chunk, m);
}

template uint64_t compress<float>(const float*, Dims, double, uint8_t*, uint64_t, hipStream_t);
template uint64_t compress<double>(const double*, Dims, double, uint8_t*, uint64_t, hipStream_t);
template void decompress<float>(const uint8_t*, uint64_t, Dims, float*, hipStream_t, bool);
template void decompress<double>(const uint8_t*, uint64_t, Dims, double*, hipStream_t, bool);
template void decompress_into<float>(const uint8_t*, uint64_t, Dims, float*, hipStream_t, bool, const DecodeMemory&);
template void decompress_into<double>(const uint8_t*, uint64_t, Dims, double*, hipStream_t, bool,
                                      const DecodeMemory&);

}  // namespace wz
