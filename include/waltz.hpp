// Waltz C++ API (drop-in for the reference's include/waltz.hpp:17-81).
// All buffers are caller-owned device pointers; `stream` is a hipStream_t
// passed as void* (nullptr = default stream). Both calls synchronize the
// stream before returning. Errors are C++ exceptions (std::invalid_argument
// for arguments, std::runtime_error for device, capacity or strict-bound
// failures). Config.absErrorBound must already be absolute (resolve REL with
// d_extrema first). The library is built for gfx950 (MI355X) and exports the
// explicit instantiations for T in {float, double}, NDIMS = 3.
#pragma once

#include <cstddef>
#include <cstdint>
#include <limits>
#include <stdexcept>

#include "utils/Blob.hpp"
#include "utils/Config.hpp"

namespace WALTZ {

// Output capacity: one input volume of magnitude/header storage plus the
// fixed side-channel caps (2^22 repairs, 2^20 q-outliers) and 1 MiB of slack
// (include/waltz.hpp:21-33 of the reference).
template <typename T>
size_t compressed_buffer_capacity(uint64_t num) {
    constexpr uint64_t pwe_cap = uint64_t{1} << 22;
    constexpr uint64_t qout_cap = uint64_t{1} << 20;
    constexpr uint64_t slack = uint64_t{1} << 20;
    const uint64_t pwe = num < pwe_cap ? num : pwe_cap;
    const uint64_t qout = num < qout_cap ? num : qout_cap;
    const uint64_t side = pwe * (sizeof(uint32_t) + sizeof(T)) + qout * (sizeof(uint32_t) + sizeof(int32_t)) + slack;
    const uint64_t max_bytes = std::numeric_limits<size_t>::max();
    if (side > max_bytes || num > (max_bytes - side) / sizeof(T))
        throw std::overflow_error("Waltz compressed buffer capacity exceeds size_t");
    return static_cast<size_t>(num * sizeof(T) + side);
}

template <typename T, int NDIMS>
size_t d_WALTZ_compress(T* d_oridata, char* d_cmpData, const WALTZ::Config& config, const char* cmpPath,
                        void* stream);
template <typename T, int NDIMS>
size_t d_WALTZ_decompress(char* d_cmpData, T* d_decData, const WALTZ::Config& config, const char* decPath,
                          void* stream);
// Exact minimum and maximum of a device array (used to resolve REL bounds).
template <typename T>
void d_extrema(T* d_data, uint64_t n, double& out_min, double& out_max, void* stream);

template <typename T>
size_t d_compress(const WALTZ::Config& config, T* d_oridata, char* d_cmpData, const char* cmpPath, void* stream) {
    Config conf(config);
    if (conf.ndims != 3) throw std::invalid_argument("compression currently only supports 3D data.");
    return d_WALTZ_compress<T, 3>(d_oridata, d_cmpData, conf, cmpPath, stream);
}

// Decompress an archive produced by d_compress (device -> device). Config must
// carry the original dimensions; the archive does not store them.
template <typename T>
size_t d_decompress(const WALTZ::Config& config, char* d_cmpData, T* d_decData, const char* decPath, void* stream) {
    Config conf(config);
    if (conf.ndims != 3) throw std::invalid_argument("decompression currently only supports 3D data.");
    return d_WALTZ_decompress<T, 3>(d_cmpData, d_decData, conf, decPath, stream);
}

}  // namespace WALTZ
