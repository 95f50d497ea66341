// Chunk decoder that allocates nothing: waltz_cuda_decode_chunk's rules and
// results, with every piece of device scratch taken from one fixed array of
// this library instead of the cached workspace (which grows with hipMalloc).
// For callers that must not allocate device memory: code under stream
// capture, serving loops, harnesses that forbid hipMalloc.
//
// The array is a file-scope __device__ variable, so every device the library
// runs on holds its own copy (kScratchBytes, about 26.4 MiB) from module load
// until the process ends. Host code reaches a device's copy only through the
// address hipGetSymbolAddress returns on that device (looked up once per
// device, on the entry point's first call there); the variable's own name is a
// host-side stand-in, not a device address. Calls share the array and it
// keeps the last call's contents: each call clears what must start at zero
// (the result record, the duplicate-index bitmaps) and rebuilds the rank
// tables it needs. Calls are serialised inside this process; calls from two
// processes cannot overlap on a copy, since each process has its own.
#include <hip/hip_runtime.h>

#include <mutex>
#include <stdexcept>
#include <string>

#include "../../include/waltz_hip.h"
#include "launch.hpp"

namespace {

constexpr uint32_t kMaxSamples = WALTZ_NOALLOC_MAX_SAMPLES;
// decode_scratch_bytes(kMaxSamples, 8): three f64 volumes, u16 magnitudes,
// sign words (each a multiple of 256 bytes at this size)
constexpr uint64_t kDecodeBytes = 3ull * 8 * kMaxSamples + 2ull * kMaxSamples + kMaxSamples / 8;
constexpr uint64_t kRankBytes = 2ull * 8 * 8192 * sizeof(uint16_t);
constexpr uint64_t kResultBytes = 256;
constexpr uint64_t kScratchBytes = kDecodeBytes + kRankBytes + kResultBytes;
static_assert(kDecodeBytes % 256 == 0 && wz::kRecordBytes <= kResultBytes, "scratch layout");

constexpr int kMaxDevices = 64;

}  // namespace

static __device__ __align__(256) uint8_t scratch[kScratchBytes];

namespace {

std::mutex g_lock;
uint8_t* g_scratch[kMaxDevices] = {};  // per device: the address of its copy of `scratch`

// Device address of the current device's `scratch`; called with g_lock held.
uint8_t* device_scratch() {
    int dev = 0;
    wz::hip_check(hipGetDevice(&dev), "hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("device index out of range");
    if (!g_scratch[dev]) {
        void* p = nullptr;
        wz::hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "fixed decode scratch address");
        if (!p) throw std::runtime_error("fixed decode scratch address");
        g_scratch[dev] = static_cast<uint8_t*>(p);
    }
    return g_scratch[dev];
}

thread_local std::string g_error;

}  // namespace

extern "C" const char* waltz_cuda_decode_chunk_noalloc(uint32_t nx, uint32_t ny, uint32_t nz, int width,
                                                       const void* archive, size_t archive_bytes, void* output,
                                                       size_t output_bytes, void* stream) {
    hipStream_t s = static_cast<hipStream_t>(stream);
    try {
        std::lock_guard<std::mutex> hold(g_lock);
        if ((width != 4 && width != 8) || !nx || !ny || !nz || !archive || !output || archive_bytes < 32 ||
            uint64_t(nx) * ny > 0xFFFFFFFFull / unsigned(width) / nz ||
            output_bytes != uint64_t(nx) * ny * nz * unsigned(width))
            throw std::invalid_argument("invalid chunk dimensions or buffer sizes");
        if (uint64_t(nx) * ny * nz > kMaxSamples)
            throw std::invalid_argument("chunk exceeds the fixed decode scratch (1048576 samples)");
        uint8_t* base = device_scratch();
        const wz::DecodeMemory mem{base, kDecodeBytes, reinterpret_cast<uint16_t*>(base + kDecodeBytes),
                                   reinterpret_cast<wz::DecodeRecord*>(base + kDecodeBytes + kRankBytes)};
        const wz::Dims d{nx, ny, nz};
        if (width == 4)
            wz::decompress_into<float>(static_cast<const uint8_t*>(archive), archive_bytes, d,
                                       static_cast<float*>(output), s, true, mem);
        else
            wz::decompress_into<double>(static_cast<const uint8_t*>(archive), archive_bytes, d,
                                        static_cast<double*>(output), s, true, mem);
        return nullptr;
    } catch (const std::exception& e) {
        g_error = e.what();
    } catch (...) {
        g_error = "unknown Waltz decode error";
    }
    (void)hipStreamSynchronize(s);
    return g_error.c_str();
}
