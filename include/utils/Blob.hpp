// The 32-byte Waltz archive header and its section layout (little endian, no
// magic or version). Field names, offsets and flag bits are the wire format
// of the reference's include/utils/Blob.hpp:15-90; the code is this
// project's own.
//
//  off  field          meaning
//   0   flags          bit0 dyadic, bit1 compact repairs (never set by the
//                      canonical encoder), bit2 FP64, bits3-4 synthesis+1
//                      (canonical all-level lifting = 2 -> 3), bits5-7 zero
//   1   levels_xy      2   levels_z      3   reorder_bz
//   4   pwe_bytes (u32)  8 q (f64)  16 mag_comp_bytes (u64)  24 n_pwe  28 n_qo
//
// Sections after the header, each zero-padded to 8 bytes: WZP stream, repair
// indices (u32, increasing), repair errors (T), q-outlier indices (u32,
// increasing), q-outlier magnitudes (u32).
#pragma once

#include <cstddef>
#include <cstdint>
#include <stdexcept>

namespace WALTZ {

enum class PweEncoding : uint8_t { Native = 0, Small = 1 };
enum class Synthesis : uint8_t { Convolution = 0, HaloLifting3D = 1, HaloAllLevels = 2 };

// Levels per axis group: halve (rounding up) while longer than 8, at most 5.
inline constexpr uint8_t dwt_levels(uint32_t length) {
    uint8_t levels = 0;
    while (length > 8 && levels < 5) {
        length -= length / 2;
        ++levels;
    }
    return levels;
}

struct WaltzBlobHeader {
    static constexpr uint8_t CudaFusedPlaneZ = 0x20;
    uint8_t flags, levels_xy, levels_z, reorder_bz;
    uint32_t pwe_bytes;
    double q;
    uint64_t mag_comp_bytes;
    uint32_t n_pwe, n_qo;

    static constexpr uint8_t pack_flags(bool dyadic, unsigned width, PweEncoding pwe, Synthesis synth) {
        return uint8_t(unsigned(dyadic) | (unsigned(pwe) << 1) | (width == 8 ? 4u : 0u) |
                       ((unsigned(synth) + 1u) << 3));
    }
    static constexpr bool valid_flags(uint8_t f) { return !(f & 0xE0u) && (f & 0x18u); }
    static constexpr unsigned width(uint8_t f) { return (f & 4u) ? 8u : 4u; }
    static constexpr Synthesis synthesis_of(uint8_t f) { return Synthesis(int((f >> 3) & 3u) - 1); }
    static constexpr PweEncoding repair_of(uint8_t f) { return PweEncoding((f >> 1) & 1u); }
};
inline constexpr size_t BlobSize = 32;
inline constexpr size_t BlobPweBytes = 4;
inline constexpr size_t BlobQuantum = 8;
inline constexpr size_t BlobWzpBytes = 16;
inline constexpr size_t BlobPweCount = 24;
inline constexpr size_t BlobQoutCount = 28;
static_assert(sizeof(WaltzBlobHeader) == BlobSize, "header is 32 bytes");

struct BlobLayout {
    uint64_t pi, pv, qi, qv, end;
    static constexpr uint64_t pad(uint64_t n) { return (n + 7) & ~uint64_t(7); }
    static constexpr BlobLayout from_sizes(uint64_t wzp, uint32_t np, uint32_t repairs, uint32_t nq) {
        const uint64_t pi = BlobSize + pad(wzp), pv = pi + pad(uint64_t(np) * 4);
        const uint64_t qi = pv + pad(repairs), qb = pad(uint64_t(nq) * 4);
        return {pi, pv, qi, qi + qb, qi + 2 * qb};
    }
    static BlobLayout checked(uint64_t wzp, uint32_t np, uint32_t repairs, uint32_t nq) {
        const uint64_t fixed = from_sizes(0, np, repairs, nq).end;
        const uint64_t limit = uint64_t(SIZE_MAX) & ~uint64_t(7);
        if (fixed > limit || wzp > limit - fixed) throw std::overflow_error("Waltz archive size overflow");
        return from_sizes(wzp, np, repairs, nq);
    }
};

}  // namespace WALTZ
