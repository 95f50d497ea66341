// Host-side declarations of the stage launchers (one per pipeline stage).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

#include "common.hpp"

namespace wz {

inline void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}
inline void check_launch_size(uint64_t total) {
    if (total > 0xFFFFFF00ull) throw std::invalid_argument("Waltz launch exceeds the 32-bit grid range");
}

// transform.hip
template <class T>
void transform_forward(const T* input, T* coeff, T* tmp0, T* tmp1, Dims full, unsigned lxy, unsigned lz,
                       bool dyadic, uint32_t* status, hipStream_t s);
// What the last (finest) inverse level does with its reconstruction when it
// runs as a fused kernel: store it (optionally screening for non-finite
// samples) or, without storing, derive native repairs against `orig`.
template <class T>
struct Level0Emit {
    int mode;          // InvEmit (fused.hpp)
    uint32_t* status;
    const T* orig;
    double bound;
    SegList rep;       // segments/cap per level0_repair_segments()
    double quantum;    // kEmitRepair: > 0 when the finest level's high bands hold the
                       // unquantized coefficients (quantized and dequantized on load)
    // kEmitStoreScreen: a repair list (strictly increasing indices, np > 0
    // entries; np = 0: none) applied to the samples as they are stored, and
    // scratch for its key table (patch_table_words(full) words)
    const uint32_t* pidx = nullptr;
    const T* perr = nullptr;
    uint32_t np = 0;
    uint32_t* ptable = nullptr;
    bool ptable_built = false;  // the caller already ran repair_keys into ptable
    // a planned decode: the list's count and places come from this device
    // record (np above is then only an upper bound, for the size checks)
    const DecodeRecord* plan = nullptr;
    const uint8_t* archive = nullptr;
};
// Words of the folded repair patches' key table (one per row and 64-sample x
// tile), or 0 when a list of np entries of `width`-byte values cannot be
// folded into the finest level's stores for this geometry.
uint64_t patch_table_words(Dims full, uint32_t np, unsigned width);
// Returns true when `emit` was carried out inside the fused finest level;
// false means the caller must run the equivalent separate pass on `output`.
// requant > 0 (plane-schedule fields with plane_fused_fits only): coeff holds
// the unquantized coefficients, quantized with this quantum and dequantized
// as the levels load them.
template <class T>
bool transform_inverse(T* coeff, T* output, T* tmp0, T* tmp1, Dims full, unsigned lxy, unsigned lz, bool dyadic,
                       hipStream_t s, const Level0Emit<T>* emit = nullptr, double requant = 0.0);
// transform_inverse of a dyadic field with a fused finest level, in two
// parts, so that a caller can order other work between them: the levels
// above the finest one (they read and write coeff inside the level-1 box
// [0, ceil(nx/2)) x [0, ceil(ny/2)) x [0, ceil(nz/2)) only, and tmp0/tmp1),
// which return where their reconstruction of that box lies; then the finest
// level into `output` (not coeff) with `emit`.
template <class T>
struct LowBox {
    const T* p;
    uint32_t py, pz;  // row and plane pitch (elements)
};
template <class T>
LowBox<T> inverse_coarse_levels(T* coeff, T* tmp0, T* tmp1, Dims full, unsigned lxy, hipStream_t s);
// transform_inverse of a dyadic field whose finest level runs fused, leaving
// `coeff` intact: the level-1 box is copied into `output` and the coarse
// levels run there, the finest level reads the high bands from coeff. False
// (nothing done) when the geometry has no fused finest level.
template <class T>
bool transform_inverse_preserving(const T* coeff, T* output, T* tmp0, T* tmp1, Dims full, unsigned lxy,
                                  hipStream_t s);
template <class T>
void inverse_finest_level(const T* coeff, const LowBox<T>& low, T* output, Dims full, unsigned lxy,
                          const Level0Emit<T>& emit, hipStream_t s);
// Transpose of transform_forward (of_inverse = false) or transform_inverse
// (true) with the same geometry, in place on `data`, or from `src` (left
// intact) into `data`; tmp0/tmp1 as above.
template <class T>
void transform_adjoint(T* data, T* tmp0, T* tmp1, Dims full, unsigned lxy, unsigned lz, bool dyadic,
                       bool of_inverse, hipStream_t s, const T* src = nullptr);
// The compress inverse of a brick-pair field whose coefficients stay
// unquantized (the encoder writes nothing back), in two halves: levels >= 2
// on a quantized and dequantized compact copy of their box (in t1), level 1
// as a fused kernel that requantizes the high bands it loads, its
// reconstruction (compact, the level-1 box) at the start of t0; then the
// finest level fused with the same requantization and `emit`. Applies when
// coarse_split_fits (dyadic, >= 2 levels, both leading levels fused).
bool coarse_split_fits(Dims full, unsigned lxy, unsigned width);
// Whether the plane schedule's inverse of this geometry runs as fused XY
// levels and one LDS-staged Z phase; that inverse screens its output for
// non-finite samples when asked (a kEmitStoreScreen emit without repairs).
bool plane_fused_fits(Dims full, unsigned lxy, unsigned lz, unsigned width);
template <class T>
void inverse_coarse_requant(const T* coeff, T* t0, T* t1, Dims full, unsigned lxy, double quantum, hipStream_t s);
template <class T>
void inverse_finest_requant(const T* coeff, const T* t0, Dims full, unsigned lxy, const Level0Emit<T>& emit,
                            hipStream_t s);
// Segments and per-segment capacity the fused finest inverse level uses for
// repairs (0 segments when that level is not fused for this geometry).
void level0_repair_segments(Dims full, unsigned lxy, bool dyadic, unsigned width, uint32_t* segments,
                            uint32_t* cap);


// Bit-neutral launch choices of the dyadic schedule for one geometry, type
// and direction: leading levels run as fused kernels, and their z segments
// (positive multiples of 16).
struct Tuning {
    unsigned fused;
    uint32_t sz[5];
    // per-axis passes (the plane schedule, and the dyadic levels not fused):
    // line tile of the X, Y, Z passes (8, 16 or 32; forward passes take any,
    // the inverse only its X tile: its Y and Z tiles are part of the format),
    // and the largest plane (samples) whose X and Y passes run as one
    // LDS-staged launch (0, 1024 or 4096)
    uint32_t tile[3];
    uint32_t pair_max;
};
Tuning tuning_for(Dims full, unsigned lxy, unsigned width, bool inverse);
Tuning heuristic_tuning(Dims full, unsigned lxy);
void set_tuning(Dims full, unsigned width, bool inverse, const Tuning& t);
void clear_tuning();
std::string tuning_report();
// Measures the candidates on this device and keeps the fastest (autotune.hip).
template <class T>
void autotune(Dims full, hipStream_t s);

// quantize.hip
struct BrickGeometry {
    Dims d;
    uint32_t bz, nbx, nby, nbz;
};
BrickGeometry brick_geometry(Dims d, uint32_t bz);
// rank: 8*8192 u16 (local row-major -> Morton rank); inv: 8*8192 u16 (rank -> local index).
void build_rank_tables(Dims d, uint32_t bz, uint16_t* rank, uint16_t* inv, hipStream_t s);

struct SideList {          // unsorted (index, payload) list + row-major flag bitmap
    uint32_t* bitmap;      // ceil(n/32) words, zeroed
    uint32_t* index;       // capacity entries
    void* payload;         // capacity entries of payload type
    uint32_t* count;       // device counter (zeroed)
    uint32_t capacity;
};

// Appends this lane's q-outlier (row-major index g, magnitude m) when `over`,
// with one counter atomic per wave for all of the wave's entries (thousands
// of entries on one counter otherwise serialise); the list's order is free,
// the archive orders it by the flag bitmap. Called by every active lane.
WZ_D void append_qoutlier(const SideList& qo, bool over, uint32_t g, uint32_t m) {
    const unsigned long long mask = __ballot(over);
    if (!mask) return;
    const uint32_t lane = __lane_id(), leader = uint32_t(__ffsll((long long)mask)) - 1u;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(qo.count, uint32_t(__popcll(mask)));
    base = uint32_t(__builtin_amdgcn_readlane(int(base), int(leader)));  // leader: wave-uniform
    if (!over) return;
    atomicOr(&qo.bitmap[g >> 5], 1u << (g & 31u));
    const uint32_t slot = base + uint32_t(__popcll(mask & ((1ull << lane) - 1ull)));
    if (slot < qo.capacity) {
        qo.index[slot] = g;
        static_cast<uint32_t*>(qo.payload)[slot] = m;
    }
}

template <class T>
void quantize_bricks(T* coeff, const BrickGeometry& g, double quantum, const uint16_t* rank, uint16_t* mag,
                     uint32_t* sign, SideList q, uint32_t* status, hipStream_t s);
// plan: a planned decode's device record (quantum from it; nothing unless ok)
template <class T>
void dequantize_bricks(const uint16_t* mag, const uint32_t* sign, const BrickGeometry& g, double quantum,
                       const uint16_t* rank, T* out, hipStream_t s, const DecodeRecord* plan = nullptr);

// wzp.hip
struct WzpScratch {
    unsigned long long* desc;  // groups lookback descriptors (zeroed)
    uint8_t* staging;          // wzp_staging_bytes(chunks): one payload slot per chunk
    uint32_t* size;            // chunks payload sizes
    uint32_t* ticket;          // 1 counter (zeroed)
};
// A second stream for the WZP group pass, so that it runs beside the work
// queued after the encoder on the caller's stream: the pass waits for `ready`
// (recorded after the encoder) and records `done`, which the caller's stream
// must wait for before it reads the stream's bytes or total.
struct SideStream {
    hipStream_t stream;
    hipEvent_t ready, done;
    hipEvent_t mid;  // a join point inside the side stream's work (decode: the repair key table built)
};
// Staging of the encoder: kChunkSlot bytes per chunk plus read-ahead padding.
inline uint64_t wzp_staging_bytes(uint32_t chunks) { return uint64_t(chunks) * kChunkSlot + 64; }
// Encodes mag/sign into out (WZP stream at out[0..]); payload capacity bound
// `capacity` (bytes available from out). Total bytes -> *total (device u64).
void wzp_encode(const uint16_t* mag, const uint32_t* sign, uint32_t n, uint8_t* out, uint64_t capacity,
                WzpScratch w, unsigned long long* total, uint32_t* status, hipStream_t s,
                const SideStream* pack = nullptr);
// Decodes a WZP stream of `bytes` bytes into packed mag/sign arrays. The
// decoders and the decode_* calls below also check the group directory
// (spans inside the payload and tiling it exactly, else kCorrupt), on `side`
// when given, beside the chunk decoders (both only read the stream); the
// caller's stream must wait for side->done before it reads the status.
void wzp_decode(const uint8_t* in, uint64_t bytes, uint32_t n, uint16_t* mag, uint32_t* sign, uint32_t* status,
                hipStream_t s, const SideStream* side = nullptr);
// Paired-brick geometry only (pair_bricks_fit): quantization fused into the
// encoder (q-outliers appended to `qo`; the dequantized coefficients written
// back in place over the whole volume, over the coarser levels' box only, or
// nowhere), and dequantization fused into the decoder (writes the
// coefficient volume).
enum WriteBack { kWriteAll, kWriteCoarseBox, kWriteNone };
template <class T>
void quantize_encode_pairs(T* coeff, Dims d, double quantum, SideList qo, uint8_t* out, uint64_t capacity,
                           WzpScratch w, unsigned long long* total, uint32_t* status, hipStream_t s,
                           WriteBack wb = kWriteAll, const SideStream* pack = nullptr);
template <class T>
void decode_dequantize_pairs(const uint8_t* in, uint64_t bytes, Dims d, double quantum, T* coeff, uint32_t* status,
                             hipStream_t s, const SideStream* side = nullptr);
// The same two, planned on the device: stream size and quantum from `rec`
// (nothing happens unless rec->ok); the directory check is a launch of its
// own (wzp_directory_planned), on whichever stream the caller picks.
template <class T>
void decode_dequantize_pairs_planned(const uint8_t* in, const DecodeRecord* rec, Dims d, T* coeff, uint32_t* status,
                                     hipStream_t s);
void wzp_decode_planned(const uint8_t* in, const DecodeRecord* rec, uint32_t n, uint16_t* mag, uint32_t* sign,
                        uint32_t* status, hipStream_t s);
void wzp_directory_planned(const uint8_t* in, const DecodeRecord* rec, uint32_t n, uint32_t* status, hipStream_t s);
// The same fusions for the other fields of full bricks (rank_bricks_fit, z
// band bz): `rank` is the full brick's rank table (build_rank_tables, shape
// 0), 8-byte aligned.
template <class T>
void quantize_encode_ranks(T* coeff, Dims d, uint32_t bz, const uint16_t* rank, double quantum, SideList qo,
                           uint8_t* out, uint64_t capacity, WzpScratch w, unsigned long long* total,
                           uint32_t* status, hipStream_t s, WriteBack wb = kWriteAll,
                           const SideStream* pack = nullptr);
template <class T>
void decode_dequantize_ranks(const uint8_t* in, uint64_t bytes, Dims d, uint32_t bz, const uint16_t* rank,
                             double quantum, T* coeff, uint32_t* status, hipStream_t s,
                             const SideStream* side = nullptr);
template <class T>
void decode_dequantize_ranks_planned(const uint8_t* in, const DecodeRecord* rec, Dims d, uint32_t bz,
                                     const uint16_t* rank, T* coeff, uint32_t* status, hipStream_t s);

// sections.hip
constexpr uint32_t kRepairTile = 4096;  // row-major indices per standalone repair segment
// Standalone repair pass: segment k = indices [4096k, 4096k+4096), cap 4096.
template <class T>
void detect_repairs(const T* original, const T* reconstructed, uint32_t n, double bound, SegList r,
                    uint32_t* status, hipStream_t s);
struct EncodeResult {
    unsigned long long wzp_bytes;  // written by wzp_encode
    unsigned long long end;        // archive bytes
    uint32_t np, nq, status, pad;
};
// Orders both side lists by index into the archive sections and writes the
// 32-byte header and all padding. `scratch` needs finish_scratch_words(n) u32.
uint64_t finish_scratch_words(uint32_t n);
template <class T>
void finish_archive(uint8_t* archive, uint64_t capacity, uint32_t n, SegList repairs, SideList qout,
                    uint32_t* scratch, uint8_t flags, uint8_t lx, uint8_t lz, uint8_t bz, double quantum,
                    EncodeResult* result, hipStream_t s);
template <class T>
void apply_qoutliers(T* coeff, const uint8_t* idx, const uint8_t* val, uint32_t count, uint32_t n, double quantum,
                     uint32_t* status, hipStream_t s);
template <class T>
void apply_repairs(T* out, const uint8_t* idx, const uint8_t* err, uint32_t count, uint32_t n, uint32_t* status,
                   hipStream_t s);
// The key table of a repair list folded into the finest inverse level
// (entry of row r and x tile t at r * ceil(nx/64) + t: the list position of
// the tile's first repair; tiles without repairs keep stale entries, which
// the reader rejects against the list), with rpatch's record checks.
template <class T>
void repair_keys(const uint32_t* idx, const T* err, uint32_t count, Dims d, uint32_t* table, uint32_t* status,
                 hipStream_t s);
// Planned decode (DecodeRecord): the header check into `rec` (its status
// cleared), and the list kernels reading counts and places from it.
void decode_plan(const uint8_t* archive, uint64_t archive_bytes, const PlanExpect& e, DecodeRecord* rec,
                 hipStream_t s);
template <class T>
void apply_qoutliers_planned(T* coeff, const uint8_t* archive, const DecodeRecord* rec, uint32_t n, uint32_t* status,
                             hipStream_t s);
template <class T>
void repair_keys_planned(const uint8_t* archive, const DecodeRecord* rec, Dims d, uint32_t* table, uint32_t* status,
                         hipStream_t s);
template <class T>
void apply_repairs_planned(T* out, const uint8_t* archive, const DecodeRecord* rec, uint32_t n, uint32_t* status,
                           hipStream_t s);
// Flags kCorrupt when an index of the list repeats (or is >= n); `bitmap`:
// ceil(n/32) words, cleared here.
void check_unique(const uint8_t* idx, uint32_t count, uint32_t n, uint32_t* bitmap, uint32_t* status, hipStream_t s);
// Fills `words` 32-bit words with zero (a kernel, so calls stay graph-capturable).
void zero_words(void* p, uint64_t words, hipStream_t s);

// codec.hip: whole-field pipelines. Device pointers; synchronous on `s`.
// Throw std::invalid_argument / std::runtime_error with the reference's messages.
template <class T>
uint64_t compress(const T* input, Dims d, double abs_bound, uint8_t* archive, uint64_t capacity, hipStream_t s);
// archive_bytes = ~0: size taken from the header. canonical_chunk: the header
// must be exactly the one compress writes for this geometry and type, and the
// archive exactly archive_bytes long (the chunk decoder's rule).
template <class T>
void decompress(const uint8_t* archive, uint64_t archive_bytes, Dims d, T* output, hipStream_t s,
                bool canonical_chunk = false);
// Decode with memory the caller supplies (no device allocation inside):
// `scratch` holds scratch_bytes >= decode_scratch_bytes(n, width) bytes,
// 256-byte aligned; `rank` 2 * 8 * 8192 u16 (rebuilt by every call that
// needs them); `result` kRecordBytes of device memory. Same checks and
// results as decompress.
struct DecodeMemory {
    uint8_t* scratch;
    uint64_t scratch_bytes;
    uint16_t* rank;
    DecodeRecord* result;
};
uint64_t decode_scratch_bytes(uint32_t n, unsigned width);
template <class T>
void decompress_into(const uint8_t* archive, uint64_t archive_bytes, Dims d, T* output, hipStream_t s,
                     bool canonical_chunk, const DecodeMemory& mem);
// Archive bytes implied by a 32-byte header (validated), or throws.
uint64_t archive_size_from_header(const uint8_t* header, Dims d, unsigned width);
uint64_t max_compressed_bytes(Dims d, unsigned width);
template <class T>
void check_finite(const T* out, uint32_t n, uint32_t* status, hipStream_t s);

}  // namespace wz
