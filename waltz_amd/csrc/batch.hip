// Batched decode ABI (include/waltz_gpu_batch.h; reference semantics
// src/batch_decode.cuh:29-186): spans and headers are validated from the host
// copy, every archive is decoded on the device, and each tile is placed (with
// clipping at the frame edge) into the (side, sample, z, y, x) output; holes
// are filled.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>

#include "../../include/waltz_gpu_batch.h"
#include "common.hpp"
#include "launch.hpp"

namespace wz {
namespace {
thread_local std::string g_batch_error;

struct BatchGeometry {
    uint32_t gx, gy, gz, jobs;
    uint64_t n, frame;
};

void need(bool ok, const char* what) {
    if (!ok) throw std::runtime_error(what);
}

BatchGeometry batch_geometry(const waltz_batch_shape& d) {
    need(d.x && d.y && d.z && d.fx && d.fy && d.fz && d.samples && (d.sides == 1 || d.sides == 2),
         "invalid batch geometry");
    need(d.width == 4 || d.width == 8, "width must be 4 or 8");
    BatchGeometry g;
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

struct Placement {
    uint32_t tx, ty, tz, fx, fy, fz, ox, oy, oz;
    uint64_t base;  // element offset of this job's frame in the output
};

Placement place(const waltz_batch_shape& d, const BatchGeometry& g, uint32_t j) {
    Placement p{d.x, d.y, d.z, d.fx, d.fy, d.fz, 0, 0, 0, 0};
    p.ox = (j % g.gx) * d.x;
    j /= g.gx;
    p.oy = (j % g.gy) * d.y;
    j /= g.gy;
    p.oz = (j % g.gz) * d.z;
    j /= g.gz;
    p.base = (uint64_t(j % d.sides) * d.samples + j / d.sides) * g.frame;
    return p;
}

template <class T>
__global__ void __launch_bounds__(256) place_tile(const T* __restrict__ tile, T* __restrict__ out, Placement p,
                                                  uint32_t n, bool fill_only, T fill) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    const uint32_t x = p.ox + i % p.tx, y = p.oy + (i / p.tx) % p.ty, z = p.oz + i / (p.tx * p.ty);
    if (x >= p.fx || y >= p.fy || z >= p.fz) return;
    out[p.base + (uint64_t(z) * p.fy + y) * p.fx + x] = fill_only ? fill : tile[i];
}

template <class T>
void run(const waltz_batch_shape& d, const uint8_t* host, const uint8_t* packed, size_t bytes, void* scratch,
         size_t scratch_bytes, T* out, size_t out_bytes, double fill, hipStream_t s) {
    const BatchGeometry g = batch_geometry(d);
    need(out_bytes == g.frame * d.samples * d.sides * sizeof(T), "incorrect output size");
    need(scratch_bytes >= g.n * sizeof(T), "short batch buffers");
    need(bytes >= 16ull * g.jobs, "short batch buffers");
    const size_t table = bytes - 16ull * g.jobs;
    need(table % 16 == 0, "unaligned batch spans");
    const Dims tile{d.x, d.y, d.z};
    T* buf = static_cast<T*>(scratch);
    for (uint32_t j = 0; j < g.jobs; ++j) {
        uint64_t pair[2];
        std::memcpy(pair, host + table + 16ull * j, 16);
        const Placement p = place(d, g, j);
        const uint32_t blocks = uint32_t((g.n + 255) / 256);
        if (!pair[1]) {
            hipLaunchKernelGGL((place_tile<T>), dim3(blocks), dim3(256), 0, s, buf, out, p, uint32_t(g.n), true,
                               T(fill));
            continue;
        }
        const uint64_t at = pair[0], sz = pair[1];
        need(at % 16 == 0 && at <= table && sz >= 32 && sz <= table - at, "invalid batch span");
        need(archive_size_from_header(host + at, tile, sizeof(T)) == sz, "invalid canonical batch archive");
        decompress<T>(packed + at, sz, tile, buf, s);
        hipLaunchKernelGGL((place_tile<T>), dim3(blocks), dim3(256), 0, s, buf, out, p, uint32_t(g.n), false,
                           T(fill));
    }
    hip_check(hipStreamSynchronize(s), "batch decode");
}
}  // namespace
}  // namespace wz

extern "C" {

size_t waltz_gpu_batch_scratch_bytes(const waltz_batch_shape* shape) {
    try {
        const wz::BatchGeometry g = wz::batch_geometry(*shape);
        return size_t((g.n * shape->width + 255) & ~uint64_t(255));
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
            wz::run<float>(*shape, h, p, bytes, scratch, scratch_bytes, static_cast<float*>(output), output_bytes,
                           fill, s);
        else
            wz::run<double>(*shape, h, p, bytes, scratch, scratch_bytes, static_cast<double*>(output), output_bytes,
                            fill, s);
        return nullptr;
    } catch (const std::exception& e) {
        wz::g_batch_error = e.what();
    } catch (...) {
        wz::g_batch_error = "unknown batch error";
    }
    return wz::g_batch_error.c_str();
}

}  // extern "C"
