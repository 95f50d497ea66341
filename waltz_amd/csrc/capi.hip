// Exported entry points: the C++ API of include/waltz.hpp, the HIP C ABI of
// include/waltz_hip.h and the stage ABI of include/waltz_amd.h. Thin wrappers
// over the pipelines in codec.hip; exceptions become -1 / diagnostics here.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <type_traits>
#include <vector>
#include <mutex>

#include "../../include/waltz.hpp"
#include "../../include/waltz_amd.h"
#include "../../include/waltz_hip.h"
#include "common.hpp"
#include "launch.hpp"

namespace {
thread_local std::string g_error;

template <class F>
int guarded(F&& f) {
    try {
        g_error.clear();
        f();
        return 0;
    } catch (const std::exception& e) {
        g_error = e.what();
    } catch (...) {
        g_error = "unknown Waltz error";
    }
    return -1;
}

wz::Dims dims_of(uint32_t nx, uint32_t ny, uint32_t nz) { return wz::Dims{nx, ny, nz}; }

wz::Dims dims_of(const WALTZ::Config& c) {
    if (c.dimx > 0xFFFFFFFFull || c.dimy > 0xFFFFFFFFull || c.dimz > 0xFFFFFFFFull)
        throw std::invalid_argument("dimension exceeds 32 bits");
    return wz::Dims{uint32_t(c.dimx), uint32_t(c.dimy), uint32_t(c.dimz)};
}

hipStream_t as_stream(void* s) { return static_cast<hipStream_t>(s); }

// Two full-volume scratch arrays for the stage entry points, kept per device
// and grown on demand (calls are synchronous; the lock serialises them).
struct StageScratch {
    std::mutex mu;
    void* p = nullptr;
    size_t cap = 0;
    void* get(size_t bytes) {
        if (bytes > cap) {
            if (p) hipFree(p);
            p = nullptr;
            cap = 0;
            wz::hip_check(hipMalloc(&p, bytes), "stage scratch");
            cap = bytes;
        }
        return p;
    }
};

StageScratch& stage_scratch() {
    static std::mutex m;
    static std::vector<std::unique_ptr<StageScratch>> all;
    int dev = 0;
    wz::hip_check(hipGetDevice(&dev), "device");
    std::lock_guard<std::mutex> g(m);
    if (all.size() <= size_t(dev)) all.resize(size_t(dev) + 1);
    if (!all[size_t(dev)]) all[size_t(dev)] = std::make_unique<StageScratch>();
    return *all[size_t(dev)];
}

// Order-preserving integer images of floating-point values for atomic min/max.
__device__ __forceinline__ uint32_t ordered(float v) {
    const uint32_t b = __float_as_uint(v);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ unsigned long long ordered(double v) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    return (b >> 63) ? ~b : (b | (1ull << 63));
}
inline double from_ordered(uint32_t u) {
    const uint32_t b = (u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u;
    float f;
    std::memcpy(&f, &b, 4);
    return double(f);
}
inline double from_ordered(unsigned long long u) {
    const unsigned long long b = (u >> 63) ? (u & ~(1ull << 63)) : ~u;
    double d;
    std::memcpy(&d, &b, 8);
    return d;
}

template <class T, class U>
__global__ void __launch_bounds__(256) extrema_kernel(const T* __restrict__ d, uint64_t n, U* __restrict__ out) {
    U lo = ~U(0), hi = U(0);
    const uint64_t stride = uint64_t(gridDim.x) * 256u;
    for (uint64_t i = uint64_t(blockIdx.x) * 256u + threadIdx.x; i < n; i += stride) {
        const U v = ordered(d[i]);
        lo = v < lo ? v : lo;
        hi = v > hi ? v : hi;
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        const U a = __shfl_xor(lo, o, 64), b = __shfl_xor(hi, o, 64);
        lo = a < lo ? a : lo;
        hi = b > hi ? b : hi;
    }
    if ((threadIdx.x & 63u) == 0) {
        atomicMin(&out[0], lo);
        atomicMax(&out[1], hi);
    }
}

template <class T>
void extrema(const T* d, uint64_t n, double& mn, double& mx, hipStream_t s) {
    using U = typename std::conditional<sizeof(T) == 4, uint32_t, unsigned long long>::type;
    if (!n) throw std::invalid_argument("empty array");
    U init[2] = {~U(0), U(0)};
    U* dev = nullptr;
    wz::hip_check(hipMalloc(&dev, sizeof(init)), "extrema scratch");
    wz::hip_check(hipMemcpyAsync(dev, init, sizeof(init), hipMemcpyHostToDevice, s), "extrema init");
    const uint64_t blocks = (n + 255) / 256;
    hipLaunchKernelGGL((extrema_kernel<T, U>), dim3(uint32_t(blocks < 2048 ? blocks : 2048)), dim3(256), 0, s, d, n,
                       dev);
    wz::hip_check(hipMemcpyAsync(init, dev, sizeof(init), hipMemcpyDeviceToHost, s), "extrema result");
    wz::hip_check(hipStreamSynchronize(s), "extrema");
    hipFree(dev);
    mn = from_ordered(init[0]);
    mx = from_ordered(init[1]);
}
}  // namespace

// ------------------------------------------------------------------ C++ API
namespace WALTZ {

// Shape and buffer checks of src/waltz.cu:335-349: nonzero dimensions whose
// product fits the WZP element limit and equals config.num, non-null buffers,
// and an input that does not overlap the archive's capacity.
template <typename T>
wz::Dims checked_shape(const Config& c, const void* input, const void* archive) {
    if (!c.dimx || !c.dimy || !c.dimz || c.dimx > 0xFFFFFFFFull / 2 || c.dimy > (0xFFFFFFFFull / 2) / c.dimx ||
        c.dimz > (0xFFFFFFFFull / 2) / (c.dimx * c.dimy) || c.num != c.dimx * c.dimy * c.dimz || !input || !archive)
        throw std::invalid_argument("Waltz: invalid buffers or shape for WZP");
    return dims_of(c);
}

template <typename T, int NDIMS>
size_t d_WALTZ_compress(T* d_oridata, char* d_cmpData, const Config& config, const char*, void* stream) {
    static_assert(NDIMS == 3, "3-D only");
    const wz::Dims d = checked_shape<T>(config, d_oridata, d_cmpData);
    const uint64_t n = uint64_t(d.nx) * d.ny * d.nz;
    const size_t capacity = compressed_buffer_capacity<T>(n);
    const uintptr_t src = reinterpret_cast<uintptr_t>(d_oridata), dst = reinterpret_cast<uintptr_t>(d_cmpData);
    if (src <= dst ? dst - src < n * sizeof(T) : src - dst < capacity)
        throw std::invalid_argument("Waltz: input and output buffers overlap");
    return size_t(wz::compress<T>(d_oridata, d, config.absErrorBound, reinterpret_cast<uint8_t*>(d_cmpData),
                                  capacity, as_stream(stream)));
}

template <typename T, int NDIMS>
size_t d_WALTZ_decompress(char* d_cmpData, T* d_decData, const Config& config, const char*, void* stream) {
    static_assert(NDIMS == 3, "3-D only");
    const wz::Dims d = checked_shape<T>(config, d_decData, d_cmpData);
    wz::decompress<T>(reinterpret_cast<const uint8_t*>(d_cmpData), ~uint64_t(0), d, d_decData, as_stream(stream));
    return size_t(uint64_t(d.nx) * d.ny * d.nz * sizeof(T));
}

template <typename T>
void d_extrema(T* d_data, uint64_t n, double& out_min, double& out_max, void* stream) {
    extrema<T>(d_data, n, out_min, out_max, as_stream(stream));
}

template size_t d_WALTZ_compress<float, 3>(float*, char*, const Config&, const char*, void*);
template size_t d_WALTZ_compress<double, 3>(double*, char*, const Config&, const char*, void*);
template size_t d_WALTZ_decompress<float, 3>(char*, float*, const Config&, const char*, void*);
template size_t d_WALTZ_decompress<double, 3>(char*, double*, const Config&, const char*, void*);
template void d_extrema<float>(float*, uint64_t, double&, double&, void*);
template void d_extrema<double>(double*, uint64_t, double&, double&, void*);

}  // namespace WALTZ

// ------------------------------------------------------------------ C ABIs
struct waltz_hip_session {
    int device;
    std::string error;
};

extern "C" {

const char* waltz_amd_last_error(void) { return g_error.c_str(); }
const char* waltz_amd_version(void) { return "waltz-amd gfx950 (wave64, lifting + WZP + native repairs)"; }

size_t waltz_amd_max_compressed_bytes(uint32_t nx, uint32_t ny, uint32_t nz, int width) {
    try {
        return size_t(wz::max_compressed_bytes(dims_of(nx, ny, nz), unsigned(width)));
    } catch (...) {
        return 0;
    }
}

int waltz_amd_compress(int width, const void* input, uint32_t nx, uint32_t ny, uint32_t nz, double abs_bound,
                       void* archive, size_t capacity, size_t* written, void* stream) {
    if (written) *written = 0;
    return guarded([&] {
        if (!input || !archive || !written) throw std::invalid_argument("null argument");
        const wz::Dims d = dims_of(nx, ny, nz);
        if (width == 4)
            *written = wz::compress<float>(static_cast<const float*>(input), d, abs_bound,
                                           static_cast<uint8_t*>(archive), capacity, as_stream(stream));
        else if (width == 8)
            *written = wz::compress<double>(static_cast<const double*>(input), d, abs_bound,
                                            static_cast<uint8_t*>(archive), capacity, as_stream(stream));
        else
            throw std::invalid_argument("type must be FP32 or FP64");
    });
}

int waltz_amd_decompress(int width, const void* archive, size_t bytes, uint32_t nx, uint32_t ny, uint32_t nz,
                         void* output, void* stream) {
    return guarded([&] {
        if (!archive || !output) throw std::invalid_argument("null argument");
        const wz::Dims d = dims_of(nx, ny, nz);
        if (bytes < 32) throw std::runtime_error("invalid Waltz archive header");
        if (width == 4)
            wz::decompress<float>(static_cast<const uint8_t*>(archive), bytes, d, static_cast<float*>(output),
                                  as_stream(stream));
        else if (width == 8)
            wz::decompress<double>(static_cast<const uint8_t*>(archive), bytes, d, static_cast<double*>(output),
                                   as_stream(stream));
        else
            throw std::invalid_argument("type must be FP32 or FP64");
    });
}

int waltz_amd_transform(int width, int inverse, const void* input, void* output, uint32_t nx, uint32_t ny,
                        uint32_t nz, unsigned lxy, unsigned lz, int dyadic, void* stream) {
    return guarded([&] {
        if (lxy > 5 || lz > 5 || (dyadic && lxy != lz)) throw std::invalid_argument("invalid transform levels");
        if (width != 4 && width != 8) throw std::invalid_argument("type must be FP32 or FP64");
        const wz::Dims d = dims_of(nx, ny, nz);
        const uint64_t n = uint64_t(nx) * ny * nz;
        hipStream_t s = as_stream(stream);
        StageScratch& sc = stage_scratch();
        std::lock_guard<std::mutex> hold(sc.mu);
        void* tmp = sc.get(2 * ((n * size_t(width) + 255) & ~size_t(255)));
        uint32_t* status = nullptr;  // the stage transform does not screen its input
        char* t0 = static_cast<char*>(tmp);
        char* t1 = t0 + ((n * size_t(width) + 255) & ~size_t(255));
        if (width == 4) {
            if (inverse) {
                if (input != output)
                    wz::hip_check(hipMemcpyAsync(output, input, n * 4, hipMemcpyDeviceToDevice, s), "copy");
                wz::transform_inverse<float>(static_cast<float*>(output), static_cast<float*>(output),
                                             reinterpret_cast<float*>(t0), reinterpret_cast<float*>(t1), d, lxy, lz,
                                             dyadic != 0, s);
            } else {
                wz::transform_forward<float>(static_cast<const float*>(input), static_cast<float*>(output),
                                             reinterpret_cast<float*>(t0), reinterpret_cast<float*>(t1), d, lxy, lz,
                                             dyadic != 0, status, s);
            }
        } else {
            if (inverse) {
                if (input != output)
                    wz::hip_check(hipMemcpyAsync(output, input, n * 8, hipMemcpyDeviceToDevice, s), "copy");
                wz::transform_inverse<double>(static_cast<double*>(output), static_cast<double*>(output),
                                              reinterpret_cast<double*>(t0), reinterpret_cast<double*>(t1), d, lxy,
                                              lz, dyadic != 0, s);
            } else {
                wz::transform_forward<double>(static_cast<const double*>(input), static_cast<double*>(output),
                                              reinterpret_cast<double*>(t0), reinterpret_cast<double*>(t1), d, lxy,
                                              lz, dyadic != 0, status, s);
            }
        }
        wz::hip_check(hipGetLastError(), "transform launch");
        wz::hip_check(hipStreamSynchronize(s), "transform");
    });
}

int waltz_amd_transform_adjoint(int width, int of_inverse, const void* input, void* output, uint32_t nx, uint32_t ny,
                                uint32_t nz, unsigned lxy, unsigned lz, int dyadic, void* stream) {
    return guarded([&] {
        if (lxy > 5 || lz > 5 || (dyadic && lxy != lz)) throw std::invalid_argument("invalid transform levels");
        if (width != 4 && width != 8) throw std::invalid_argument("type must be FP32 or FP64");
        const wz::Dims d = dims_of(nx, ny, nz);
        const uint64_t n = uint64_t(nx) * ny * nz;
        hipStream_t s = as_stream(stream);
        if (input != output)
            wz::hip_check(hipMemcpyAsync(output, input, n * size_t(width), hipMemcpyDeviceToDevice, s), "copy");
        StageScratch& sc = stage_scratch();
        std::lock_guard<std::mutex> hold(sc.mu);
        void* tmp = sc.get(2 * ((n * size_t(width) + 255) & ~size_t(255)));
        char* t0 = static_cast<char*>(tmp);
        char* t1 = t0 + ((n * size_t(width) + 255) & ~size_t(255));
        if (width == 4)
            wz::transform_adjoint<float>(static_cast<float*>(output), reinterpret_cast<float*>(t0),
                                         reinterpret_cast<float*>(t1), d, lxy, lz, dyadic != 0, of_inverse != 0, s);
        else
            wz::transform_adjoint<double>(static_cast<double*>(output), reinterpret_cast<double*>(t0),
                                          reinterpret_cast<double*>(t1), d, lxy, lz, dyadic != 0, of_inverse != 0, s);
        wz::hip_check(hipGetLastError(), "adjoint launch");
        wz::hip_check(hipStreamSynchronize(s), "transform adjoint");
    });
}

int waltz_amd_transform_async(int width, int inverse, const void* input, void* output, void* tmp0, void* tmp1,
                              uint32_t nx, uint32_t ny, uint32_t nz, unsigned lxy, unsigned lz, int dyadic,
                              void* stream) {
    return guarded([&] {
        if (lxy > 5 || lz > 5 || (dyadic && lxy != lz)) throw std::invalid_argument("invalid transform levels");
        if (width != 4 && width != 8) throw std::invalid_argument("type must be FP32 or FP64");
        if (!input || !output || !tmp0 || !tmp1) throw std::invalid_argument("null argument");
        const wz::Dims d = dims_of(nx, ny, nz);
        const uint64_t n = uint64_t(nx) * ny * nz;
        hipStream_t s = as_stream(stream);
        auto run = [&](auto* out) {
            using T = std::remove_pointer_t<decltype(out)>;
            T* t0 = static_cast<T*>(tmp0);
            T* t1 = static_cast<T*>(tmp1);
            if (inverse) {
                // a fused finest level reads the input's high bands in place
                // and writes the output once
                if (dyadic && input != output &&
                    wz::transform_inverse_preserving<T>(static_cast<const T*>(input), out, t0, t1, d, lxy, s))
                    return;
                if (input != output)
                    wz::hip_check(hipMemcpyAsync(out, input, n * sizeof(T), hipMemcpyDeviceToDevice, s), "copy");
                wz::transform_inverse<T>(out, out, t0, t1, d, lxy, lz, dyadic != 0, s);
            } else {
                wz::transform_forward<T>(static_cast<const T*>(input), out, t0, t1, d, lxy, lz, dyadic != 0, nullptr,
                                         s);
            }
        };
        if (width == 4) run(static_cast<float*>(output));
        else run(static_cast<double*>(output));
        wz::hip_check(hipGetLastError(), "transform launch");
    });
}

int waltz_amd_transform_adjoint_async(int width, int of_inverse, const void* input, void* output, void* tmp0,
                                      void* tmp1, uint32_t nx, uint32_t ny, uint32_t nz, unsigned lxy, unsigned lz,
                                      int dyadic, void* stream) {
    return guarded([&] {
        if (lxy > 5 || lz > 5 || (dyadic && lxy != lz)) throw std::invalid_argument("invalid transform levels");
        if (width != 4 && width != 8) throw std::invalid_argument("type must be FP32 or FP64");
        if (!input || !output || !tmp0 || !tmp1) throw std::invalid_argument("null argument");
        const wz::Dims d = dims_of(nx, ny, nz);
        hipStream_t s = as_stream(stream);
        if (width == 4)
            wz::transform_adjoint<float>(static_cast<float*>(output), static_cast<float*>(tmp0),
                                         static_cast<float*>(tmp1), d, lxy, lz, dyadic != 0, of_inverse != 0, s,
                                         static_cast<const float*>(input));
        else
            wz::transform_adjoint<double>(static_cast<double*>(output), static_cast<double*>(tmp0),
                                          static_cast<double*>(tmp1), d, lxy, lz, dyadic != 0, of_inverse != 0, s,
                                          static_cast<const double*>(input));
        wz::hip_check(hipGetLastError(), "adjoint launch");
    });
}

int waltz_amd_autotune(int width, uint32_t nx, uint32_t ny, uint32_t nz, void* stream) {
    return guarded([&] {
        if (width != 4 && width != 8) throw std::invalid_argument("type must be FP32 or FP64");
        const wz::Dims d = dims_of(nx, ny, nz);
        if (width == 4) wz::autotune<float>(d, as_stream(stream));
        else wz::autotune<double>(d, as_stream(stream));
    });
}

int waltz_amd_set_tuning(int width, uint32_t nx, uint32_t ny, uint32_t nz, int inverse, unsigned fused_levels,
                         const uint32_t* z_segments) {
    return guarded([&] {
        if (width != 4 && width != 8) throw std::invalid_argument("type must be FP32 or FP64");
        if (!z_segments) throw std::invalid_argument("null z segments");
        const wz::Dims d = dims_of(nx, ny, nz);
        wz::Tuning t = wz::tuning_for(d, wz::dwt_levels(nx < ny ? nx : ny), unsigned(width), inverse != 0);
        t.fused = fused_levels;
        for (int l = 0; l < 5; ++l) t.sz[l] = z_segments[l];
        wz::set_tuning(d, unsigned(width), inverse != 0, t);
    });
}

int waltz_amd_set_pass_tuning(int width, uint32_t nx, uint32_t ny, uint32_t nz, int inverse, const uint32_t* tiles,
                              uint32_t lds_pair_max) {
    return guarded([&] {
        if (width != 4 && width != 8) throw std::invalid_argument("type must be FP32 or FP64");
        if (!tiles) throw std::invalid_argument("null tiles");
        const wz::Dims d = dims_of(nx, ny, nz);
        wz::Tuning t = wz::tuning_for(d, wz::dwt_levels(nx < ny ? nx : ny), unsigned(width), inverse != 0);
        for (int a = 0; a < 3; ++a) t.tile[a] = tiles[a];
        t.pair_max = lds_pair_max;
        wz::set_tuning(d, unsigned(width), inverse != 0, t);
    });
}

void waltz_amd_clear_tuning(void) { wz::clear_tuning(); }

size_t waltz_amd_tuning_report(char* buf, size_t cap) {
    const std::string r = wz::tuning_report();
    if (buf && cap) {
        const size_t k = std::min(cap - 1, r.size());
        std::memcpy(buf, r.data(), k);
        buf[k] = 0;
    }
    return r.size();
}

int waltz_amd_quantize(int width, void* coeff, uint32_t nx, uint32_t ny, uint32_t nz, uint32_t bz, double quantum,
                       void* mag, void* sign, void* qidx, void* qval, uint32_t* nq, void* stream) {
    return guarded([&] {
        const wz::Dims d = dims_of(nx, ny, nz);
        const uint64_t n = uint64_t(nx) * ny * nz;
        const uint64_t words = (n + 31) / 32;
        hipStream_t s = as_stream(stream);
        uint32_t* scratch = nullptr;  // q bitmap, counter, status
        wz::hip_check(hipMalloc(reinterpret_cast<void**>(&scratch), (words + 2) * 4), "quantize scratch");
        wz::zero_words(scratch, words + 2, s);
        wz::zero_words(sign, words, s);
        uint16_t* rank = nullptr;
        wz::hip_check(hipMalloc(reinterpret_cast<void**>(&rank), 2 * 8 * 8192 * 2), "rank tables");
        wz::build_rank_tables(d, bz, rank, rank + 8 * 8192, s);
        const wz::BrickGeometry g = wz::brick_geometry(d, bz);
        wz::SideList q{scratch, static_cast<uint32_t*>(qidx), qval, scratch + words, uint32_t(n)};
        if (width == 4)
            wz::quantize_bricks<float>(static_cast<float*>(coeff), g, quantum, rank, static_cast<uint16_t*>(mag),
                                       static_cast<uint32_t*>(sign), q, scratch + words + 1, s);
        else
            wz::quantize_bricks<double>(static_cast<double*>(coeff), g, quantum, rank, static_cast<uint16_t*>(mag),
                                        static_cast<uint32_t*>(sign), q, scratch + words + 1, s);
        uint32_t host[2];
        wz::hip_check(hipMemcpyAsync(host, scratch + words, 8, hipMemcpyDeviceToHost, s), "count");
        wz::hip_check(hipStreamSynchronize(s), "quantize");
        hipFree(scratch);
        hipFree(rank);
        *nq = host[0];
        if (host[1]) throw std::runtime_error("coefficient exceeds portable integer conversion domain");
    });
}

int waltz_amd_wzp_encode(const void* mag, const void* sign, uint32_t n, void* out, size_t capacity, size_t* written,
                         void* stream) {
    return guarded([&] {
        const wz::WzpShape ws = wz::wzp_shape(n);
        hipStream_t s = as_stream(stream);
        const uint64_t bytes = ws.chunks * 16ull + 64;
        char* scratch = nullptr;
        uint8_t* staging = nullptr;
        wz::hip_check(hipMalloc(reinterpret_cast<void**>(&scratch), bytes + 512), "wzp scratch");
        if (hipMalloc(reinterpret_cast<void**>(&staging), wz::wzp_staging_bytes(ws.chunks)) != hipSuccess) {
            hipFree(scratch);
            throw std::runtime_error("wzp staging allocation failed");
        }
        wz::zero_words(scratch, (bytes + 512) / 4, s);
        unsigned long long* desc = reinterpret_cast<unsigned long long*>(scratch);
        uint32_t* size = reinterpret_cast<uint32_t*>(scratch + ws.chunks * 8ull) + ws.chunks;
        uint32_t* tail = size + ws.chunks;  // ticket, status, pad, total(u64)
        unsigned long long* total = reinterpret_cast<unsigned long long*>(scratch + ((ws.chunks * 16ull + 16 + 7) & ~7ull));
        wz::WzpScratch w{desc, staging, size, tail};
        wz::wzp_encode(static_cast<const uint16_t*>(mag), static_cast<const uint32_t*>(sign), n,
                       static_cast<uint8_t*>(out), capacity, w, total, tail + 1, s);
        uint32_t st = 0;
        unsigned long long tot = 0;
        wz::hip_check(hipMemcpyAsync(&st, tail + 1, 4, hipMemcpyDeviceToHost, s), "status");
        wz::hip_check(hipMemcpyAsync(&tot, total, 8, hipMemcpyDeviceToHost, s), "total");
        wz::hip_check(hipStreamSynchronize(s), "wzp encode");
        hipFree(scratch);
        hipFree(staging);
        if (st) throw std::runtime_error("WZP encode failed (capacity or lookback)");
        *written = n ? size_t(tot) : 0;
    });
}

int waltz_amd_wzp_decode(const void* in, size_t bytes, uint32_t n, void* mag, void* sign, void* stream) {
    return guarded([&] {
        const wz::WzpShape ws = wz::wzp_shape(n);
        if (n > 0xFFFFFFFFu / 2) throw std::invalid_argument("WZP element count exceeds the GPU format limit");
        if (bytes < ws.header || bytes > ws.header + uint64_t(n) * 2 + (uint64_t(n) + 7) / 8)
            throw std::runtime_error("WZP archive size is outside its valid bounds");
        hipStream_t s = as_stream(stream);
        uint32_t* status = nullptr;
        wz::hip_check(hipMalloc(reinterpret_cast<void**>(&status), 4), "status");
        wz::zero_words(status, 1, s);
        wz::wzp_decode(static_cast<const uint8_t*>(in), bytes, n, static_cast<uint16_t*>(mag),
                       static_cast<uint32_t*>(sign), status, s);
        uint32_t st = 0;
        wz::hip_check(hipMemcpyAsync(&st, status, 4, hipMemcpyDeviceToHost, s), "status");
        wz::hip_check(hipStreamSynchronize(s), "wzp decode");
        hipFree(status);
        if (st) throw std::runtime_error("corrupt WZP stream");
    });
}

// ------------------------------------------------------------------ HIP session ABI
waltz_hip_session* waltz_hip_create(int device) {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) return nullptr;
    auto* s = new (std::nothrow) waltz_hip_session{device, {}};
    return s;
}
void waltz_hip_destroy(waltz_hip_session* s) { delete s; }
const char* waltz_hip_error(const waltz_hip_session* s) { return s ? s->error.c_str() : "null session"; }
size_t waltz_hip_max_bytes(const waltz_hip_config* c) {
    if (!c) return 0;
    return waltz_amd_max_compressed_bytes(c->x, c->y, c->z, c->width);
}
int waltz_hip_encode(waltz_hip_session* s, const waltz_hip_config* c, const void* input, void* archive,
                     size_t capacity, size_t* written) {
    if (!s || !c) return -1;
    int rc = -1;
    if (hipSetDevice(s->device) == hipSuccess)
        rc = waltz_amd_compress(c->width, input, c->x, c->y, c->z, c->bound, archive, capacity, written, nullptr);
    else
        g_error = "cannot select device";
    s->error = rc ? g_error : std::string();
    return rc;
}
int waltz_hip_decode(waltz_hip_session* s, const waltz_hip_config* c, const void* archive, size_t bytes,
                     void* output, size_t output_bytes) {
    if (!s || !c) return -1;
    int rc = -1;
    if (uint64_t(c->x) * c->y * c->z * uint64_t(c->width) > output_bytes)
        g_error = "output capacity too small";
    else if (hipSetDevice(s->device) == hipSuccess)
        rc = waltz_amd_decompress(c->width, archive, bytes, c->x, c->y, c->z, output, nullptr);
    else
        g_error = "cannot select device";
    s->error = rc ? g_error : std::string();
    return rc;
}

// src/waltz.cu:836-865: argument rule (width 4 or 8, nonzero dimensions, a
// volume of at most 2^32 - 1 bytes, out_bytes exactly the volume's bytes, an
// archive of at least the header) and a canonical archive of exactly
// archive_bytes. Calls are serialised; the stream is synchronised before
// returning, also on failure.
const char* waltz_cuda_decode_chunk(uint32_t nx, uint32_t ny, uint32_t nz, int width, const void* archive,
                                    size_t archive_bytes, void* output, size_t output_bytes, void* stream) {
    static std::mutex chunk_lock;
    const int rc = guarded([&] {
        std::lock_guard<std::mutex> hold(chunk_lock);
        if ((width != 4 && width != 8) || !nx || !ny || !nz || !archive || !output || archive_bytes < 32 ||
            uint64_t(nx) * ny > 0xFFFFFFFFull / unsigned(width) / nz ||
            output_bytes != uint64_t(nx) * ny * nz * unsigned(width))
            throw std::invalid_argument("invalid chunk dimensions or buffer sizes");
        const wz::Dims d = dims_of(nx, ny, nz);
        if (width == 4)
            wz::decompress<float>(static_cast<const uint8_t*>(archive), archive_bytes, d, static_cast<float*>(output),
                                  as_stream(stream), true);
        else
            wz::decompress<double>(static_cast<const uint8_t*>(archive), archive_bytes, d,
                                   static_cast<double*>(output), as_stream(stream), true);
    });
    if (!rc) return nullptr;
    (void)hipStreamSynchronize(as_stream(stream));
    return g_error.c_str();
}

}  // extern "C"
