// Autotuner of the bit-neutral launch choices (no reference counterpart). For
// one geometry and type it times, on this device and a synthetic finite
// field: for the dyadic schedule every level's fused kernel of both
// directions at each candidate z segment (the fastest per level), then the
// number of leading levels to run fused (the rest as axis passes) from whole
// transforms; for every schedule the per-axis passes' line tiles (X, Y, Z;
// the inverse only X) and the largest plane whose X and Y passes run as one
// LDS-staged launch. Every candidate produces the same bits (a z segment only
// partitions the z range among workgroups, a line tile only the line among
// threads, fused and pass forms are bitwise equal), so the choice changes
// speed only.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "common.hpp"
#include "fused.hpp"
#include "launch.hpp"

namespace wz {
namespace {

template <class T>
__global__ void __launch_bounds__(256) tune_fill(T* __restrict__ p, uint64_t n) {
    const uint64_t i = uint64_t(blockIdx.x) * 256u + threadIdx.x;
    if (i >= n) return;
    uint32_t h = uint32_t(i) * 2654435761u ^ uint32_t(i >> 32);
    h ^= h >> 15;
    h *= 2246822519u;
    h ^= h >> 13;
    p[i] = T(int32_t(h)) * T(1.0 / 2147483648.0);
}

struct DeviceBuffer {
    void* p = nullptr;
    explicit DeviceBuffer(size_t bytes) { hip_check(hipMalloc(&p, bytes), "autotune buffer"); }
    ~DeviceBuffer() {
        if (p) hipFree(p);
    }
};

struct Timer {
    hipEvent_t a, b;
    Timer() {
        hip_check(hipEventCreate(&a), "event");
        hip_check(hipEventCreate(&b), "event");
    }
    ~Timer() {
        hipEventDestroy(a);
        hipEventDestroy(b);
    }
    // best of `reps` timed launches after one warm-up
    template <class F>
    float best_ms(F&& launch, hipStream_t s, int reps = 3) {
        launch();
        float best = 1e30f;
        for (int r = 0; r < reps; ++r) {
            hip_check(hipEventRecord(a, s), "event");
            launch();
            hip_check(hipEventRecord(b, s), "event");
            hip_check(hipEventSynchronize(b), "event");
            float ms = 0;
            hip_check(hipEventElapsedTime(&ms, a, b), "event");
            best = std::min(best, ms);
        }
        return best;
    }
};

std::vector<uint32_t> candidates(uint32_t nz) {
    std::vector<uint32_t> c;
    const uint32_t top = (nz + 15) & ~15u;
    for (uint32_t sz = 16; sz <= 256; sz *= 2)
        if (sz <= top) c.push_back(sz);
    if (c.empty()) c.push_back(16);
    return c;
}
}  // namespace

// The per-axis passes' line tiles and LDS pair limit for one direction, by
// coordinate descent over whole transforms (each parameter in turn over its
// candidates, the others fixed; a change has to win by 1.5%).
template <class T>
void tune_passes(Dims full, unsigned lxy, unsigned lz, bool dyadic, bool inverse, T* x, T* c, T* t0, T* t1,
                 Timer& timer, hipStream_t s) {
    Tuning t = tuning_for(full, lxy, sizeof(T), inverse);
    auto measure = [&](const Tuning& cand) {
        set_tuning(full, sizeof(T), inverse, cand);
        if (inverse) return timer.best_ms([&] { transform_inverse<T>(c, x, t0, t1, full, lxy, lz, dyadic, s); }, s, 5);
        return timer.best_ms([&] { transform_forward<T>(x, c, t0, t1, full, lxy, lz, dyadic, nullptr, s); }, s, 5);
    };
    float best = measure(t);
    const uint32_t tiles[3] = {8, 16, 32}, pairs[3] = {0, 1024, 4096};
    for (int param = 0; param < 4; ++param) {
        if (inverse && (param == 1 || param == 2)) continue;  // canonical inverse Y / Z tiles
        for (int k = 0; k < 3; ++k) {
            Tuning cand = t;
            if (param < 3) cand.tile[param] = tiles[k];
            else cand.pair_max = pairs[k];
            if ((param < 3 && cand.tile[param] == t.tile[param]) || (param == 3 && cand.pair_max == t.pair_max))
                continue;
            const float ms = measure(cand);
            if (ms < best * 0.985f) {
                best = ms;
                t = cand;
            }
        }
    }
    set_tuning(full, sizeof(T), inverse, t);
}

template <class T>
void autotune(Dims full, hipStream_t s) {
    const unsigned lxy = dwt_levels(full.nx < full.ny ? full.nx : full.ny), lz = dwt_levels(full.nz);
    if (lxy == 0 && lz == 0) return;
    const bool dyadic = lxy == lz;
    const uint64_t n = uint64_t(full.nx) * full.ny * full.nz;
    if (!dyadic || !fused_plane_fits(full.nx, full.ny, sizeof(T))) {  // per-axis passes only
        DeviceBuffer a(n * sizeof(T)), b(n * sizeof(T)), c(n * sizeof(T)), d(n * sizeof(T));
        hipLaunchKernelGGL((tune_fill<T>), dim3(uint32_t((n + 255) / 256)), dim3(256), 0, s, static_cast<T*>(a.p), n);
        Timer timer;
        T *x = static_cast<T*>(a.p), *co = static_cast<T*>(b.p), *t0 = static_cast<T*>(c.p), *t1 = static_cast<T*>(d.p);
        tune_passes<T>(full, lxy, lz, dyadic, false, x, co, t0, t1, timer, s);
        transform_forward<T>(x, co, t0, t1, full, lxy, lz, dyadic, nullptr, s);
        tune_passes<T>(full, lxy, lz, dyadic, true, x, co, t0, t1, timer, s);
        hip_check(hipGetLastError(), "autotune launch");
        hip_check(hipStreamSynchronize(s), "autotune");
        return;
    }
    const Tuning base = heuristic_tuning(full, lxy);
    DeviceBuffer in(n * sizeof(T)), coeff(n * sizeof(T)), low(n * sizeof(T));
    hipLaunchKernelGGL((tune_fill<T>), dim3(uint32_t((n + 255) / 256)), dim3(256), 0, s, static_cast<T*>(in.p), n);
    hipLaunchKernelGGL((tune_fill<T>), dim3(uint32_t((n + 255) / 256)), dim3(256), 0, s, static_cast<T*>(coeff.p), n);
    hipLaunchKernelGGL((tune_fill<T>), dim3(uint32_t((n + 255) / 256)), dim3(256), 0, s, static_cast<T*>(low.p), n);
    Timer timer;
    Tuning fw = base, iv = base;
    Dims box = full;
    for (unsigned l = 0; l < lxy; ++l) {
        const Dims h{half_len(box.nx), half_len(box.ny), half_len(box.nz)};
        float best_f = 1e30f, best_i = 1e30f;
        for (const uint32_t sz : candidates(box.nz)) {
            FwdLevel<T> f{};
            f.src = static_cast<const T*>(in.p);
            f.spy = full.nx;
            f.spz = uint64_t(full.nx) * full.ny;
            f.n = box;
            f.hi = static_cast<T*>(coeff.p);
            f.hpy = full.nx;
            f.hpz = uint64_t(full.nx) * full.ny;
            f.lo = static_cast<T*>(low.p);
            f.lpy = h.nx;
            f.lpz = uint64_t(h.nx) * h.ny;
            f.sz = sz;
            const float tf = timer.best_ms([&] { launch_fwd_level<T>(f, s); }, s);
            if (tf < best_f) {
                best_f = tf;
                fw.sz[l] = sz;
            }
            InvLevel<T> v{};
            v.hi = static_cast<const T*>(coeff.p);
            v.hpy = full.nx;
            v.hpz = uint64_t(full.nx) * full.ny;
            v.lo = static_cast<const T*>(low.p);
            v.lpy = h.nx;
            v.lpz = uint64_t(h.nx) * h.ny;
            v.n = box;
            v.out = static_cast<T*>(in.p);
            v.opy = full.nx;
            v.opz = uint64_t(full.nx) * full.ny;
            v.sz = sz;
            v.emit = kEmitStore;
            const float ti = timer.best_ms([&] { launch_inv_level<T>(v, s); }, s);
            if (ti < best_i) {
                best_i = ti;
                iv.sz[l] = sz;
            }
        }
        box = h;
    }
    // then the number of fused levels: whole transforms with 0 .. lxy leading
    // levels fused (coarser boxes underfill the GPU but save launches)
    DeviceBuffer spare(n * sizeof(T));
    T* const x = static_cast<T*>(in.p);
    T* const c = static_cast<T*>(coeff.p);
    T* const t0 = static_cast<T*>(low.p);
    T* const t1 = static_cast<T*>(spare.p);
    float best_f = 1e30f, best_i = 1e30f;
    unsigned fuse_f = base.fused, fuse_i = base.fused;
    for (unsigned f = 0; f <= lxy; ++f) {
        Tuning tf = fw, ti = iv;
        tf.fused = ti.fused = f;
        set_tuning(full, sizeof(T), false, tf);
        set_tuning(full, sizeof(T), true, ti);
        const float a =
            timer.best_ms([&] { transform_forward<T>(x, c, t0, t1, full, lxy, lz, true, nullptr, s); }, s, 5);
        const float b = timer.best_ms([&] { transform_inverse<T>(c, x, t0, t1, full, lxy, lz, true, s); }, s, 5);
        // one more fused level has to win by 1.5%: the coarsest boxes give the
        // fused kernel a handful of workgroups, and timing noise would
        // otherwise flip between near-equal choices from run to run
        if (a < best_f * 0.985f) {
            best_f = a;
            fuse_f = f;
        }
        if (b < best_i * 0.985f) {
            best_i = b;
            fuse_i = f;
        }
    }
    fw.fused = fuse_f;
    iv.fused = fuse_i;
    set_tuning(full, sizeof(T), false, fw);
    set_tuning(full, sizeof(T), true, iv);
    // levels left to the per-axis passes: their tiles and LDS pair limit
    if (fuse_f < lxy) tune_passes<T>(full, lxy, lz, true, false, x, c, t0, t1, timer, s);
    if (fuse_i < lxy) {
        transform_forward<T>(x, c, t0, t1, full, lxy, lz, true, nullptr, s);
        tune_passes<T>(full, lxy, lz, true, true, x, c, t0, t1, timer, s);
    }
    hip_check(hipGetLastError(), "autotune launch");
    hip_check(hipStreamSynchronize(s), "autotune");
}

template void autotune<float>(Dims, hipStream_t);
template void autotune<double>(Dims, hipStream_t);

}  // namespace wz
