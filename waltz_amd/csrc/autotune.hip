// Autotuner of the bit-neutral launch choices of the dyadic schedule (no
// reference counterpart). For one geometry and type it times, on this device
// and a synthetic finite field, every level's fused kernel of both directions
// at each candidate z segment, keeps the fastest per level, and then picks the
// number of leading levels to run fused (the rest as axis passes) from whole
// transforms. Every candidate produces the same bits (a z segment only
// partitions the z range among workgroups; fused and pass forms are bitwise
// equal), so the choice changes speed only.
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

template <class T>
void autotune(Dims full, hipStream_t s) {
    const unsigned lxy = dwt_levels(full.nx < full.ny ? full.nx : full.ny), lz = dwt_levels(full.nz);
    if (lxy != lz || lxy == 0) return;  // the plane schedule runs axis passes only
    if (!fused_plane_fits(full.nx, full.ny, sizeof(T))) return;  // so do planes of 4 GiB and more
    const Tuning base = heuristic_tuning(full, lxy);
    const uint64_t n = uint64_t(full.nx) * full.ny * full.nz;
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
    hip_check(hipGetLastError(), "autotune launch");
    hip_check(hipStreamSynchronize(s), "autotune");
    set_tuning(full, sizeof(T), false, fw);
    set_tuning(full, sizeof(T), true, iv);
}

template void autotune<float>(Dims, hipStream_t);
template void autotune<double>(Dims, hipStream_t);

}  // namespace wz
