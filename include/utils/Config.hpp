// WALTZ::Config — field geometry and error-bound settings passed to the C++ API.
// Same members, enum values and dimension rule as the reference's
// include/utils/Config.hpp:35,135-202: the variadic constructor keeps only
// dimensions > 1 (in order) and counts them in ndims; chunk callers that need
// singleton axes set ndims/dimx/dimy/dimz directly.
#pragma once

#include <cstdint>
#include <initializer_list>
#include <limits>
#include <stdexcept>

namespace WALTZ {

enum EB { EB_ABS, EB_REL, EB_PSNR, EB_L2NORM, EB_ABS_AND_REL, EB_ABS_OR_REL };

class Config {
  public:
    Config() : Config(1) {}

    template <class... Dims>
    Config(Dims... args) {
        const std::initializer_list<uint64_t> dims = {static_cast<uint64_t>(args)...};
        setDims(dims.begin(), dims.end());
    }

    template <class Iter>
    uint64_t setDims(Iter begin, Iter end) {
        dimx = dimy = dimz = dimw = 1;
        ndims = 0;
        num = 1;
        for (Iter it = begin; it != end; ++it) {
            const uint64_t d = static_cast<uint64_t>(*it);
            if (d <= 1) continue;
            if (ndims >= 4) throw std::invalid_argument("Config supports at most 4 dimensions");
            setDim(ndims, d);
            ++ndims;
            if (d > std::numeric_limits<uint64_t>::max() / num)
                throw std::overflow_error("Config element count exceeds uint64 capacity");
            num *= d;
        }
        if (ndims == 0) {
            ndims = 1;
            dimx = 1;
            num = 1;
        }
        return num;
    }

    void setDim(uint8_t i, uint64_t v) {
        switch (i) {
        case 0: dimx = v; break;
        case 1: dimy = v; break;
        case 2: dimz = v; break;
        case 3: dimw = v; break;
        default: throw std::invalid_argument("dimension index out of range");
        }
    }

    uint8_t ndims;
    uint64_t dimx, dimy, dimz, dimw;
    uint64_t chunkx{256}, chunky{256}, chunkz{256}, chunkw{256};
    uint64_t num;
    uint64_t blockSize = 0;
    uint8_t errorBoundMode = EB_ABS;
    double absErrorBound = 0;
    double relErrorBound = 0;
    double psnrErrorBound = 0;
    double l2normErrorBound = 0;
};

}  // namespace WALTZ
