// C-linkage entry points into the reference CPU codec's internal stages, so the
// oracle tests can pin stage outputs (transform, WZP) and not only whole
// archives. Built only together with the reference sources into oracle/_ref/
// by oracle/Makefile; never part of the product library.
//
// The declarations restate the reference's internal interfaces:
//   cpu/src/transform.hpp:9-38  (Shape, workspace_elements, forward/inverse_f32/f64)
//   cpu/src/wzp_cpu.hpp:23-35   (wzp_max_bytes, wzp_encode, wzp_decode)
#include <cstddef>
#include <cstdint>
#include <exception>

namespace waltz_cpu {
struct Shape { std::uint32_t nx, ny, nz; };
std::size_t workspace_elements(Shape shape);
void forward_f32(float*, float*, Shape, std::uint8_t, std::uint8_t, bool, std::uint8_t, bool);
void forward_f64(double*, double*, Shape, std::uint8_t, std::uint8_t, bool, std::uint8_t, bool);
void inverse_f32(float*, float*, Shape, std::uint8_t, std::uint8_t, bool, std::uint8_t, bool);
void inverse_f64(double*, double*, Shape, std::uint8_t, std::uint8_t, bool, std::uint8_t, bool);
}  // namespace waltz_cpu

namespace waltz_cpu_detail {
std::size_t wzp_max_bytes(std::uint32_t n);
std::size_t wzp_encode(const std::uint16_t*, const std::uint32_t*, std::uint32_t, std::uint8_t*,
                       std::size_t, bool);
void wzp_decode(const std::uint8_t*, std::size_t, std::uint32_t, std::uint16_t*, std::uint32_t*, bool);
}  // namespace waltz_cpu_detail

extern "C" {

std::size_t ref_workspace_elements(std::uint32_t nx, std::uint32_t ny, std::uint32_t nz) {
    try { return waltz_cpu::workspace_elements({nx, ny, nz}); } catch (...) { return 0; }
}

// width 4/8; data and scratch hold ref_workspace_elements() samples. Returns 0 on success.
int ref_transform(int width, int inverse, void* data, void* scratch, std::uint32_t nx, std::uint32_t ny,
                  std::uint32_t nz, unsigned lxy, unsigned lz, int dyadic, int avx2) {
    try {
        const waltz_cpu::Shape s{nx, ny, nz};
        if (width == 4) {
            auto* d = static_cast<float*>(data);
            auto* t = static_cast<float*>(scratch);
            if (inverse) waltz_cpu::inverse_f32(d, t, s, lxy, lz, dyadic != 0, 2, avx2 != 0);
            else waltz_cpu::forward_f32(d, t, s, lxy, lz, dyadic != 0, 2, avx2 != 0);
        } else {
            auto* d = static_cast<double*>(data);
            auto* t = static_cast<double*>(scratch);
            if (inverse) waltz_cpu::inverse_f64(d, t, s, lxy, lz, dyadic != 0, 2, avx2 != 0);
            else waltz_cpu::forward_f64(d, t, s, lxy, lz, dyadic != 0, 2, avx2 != 0);
        }
        return 0;
    } catch (...) {
        return -1;
    }
}

std::size_t ref_wzp_max_bytes(std::uint32_t n) {
    try { return waltz_cpu_detail::wzp_max_bytes(n); } catch (...) { return 0; }
}

// Returns the encoded byte count, or (size_t)-1 on failure.
std::size_t ref_wzp_encode(const std::uint16_t* mag, const std::uint32_t* sign, std::uint32_t n,
                           std::uint8_t* out, std::size_t capacity, int avx2) {
    try { return waltz_cpu_detail::wzp_encode(mag, sign, n, out, capacity, avx2 != 0); }
    catch (...) { return ~std::size_t(0); }
}

int ref_wzp_decode(const std::uint8_t* in, std::size_t bytes, std::uint32_t n, std::uint16_t* mag,
                   std::uint32_t* sign, int avx2) {
    try { waltz_cpu_detail::wzp_decode(in, bytes, n, mag, sign, avx2 != 0); return 0; }
    catch (...) { return -1; }
}

}  // extern "C"
