// WZP magnitude/sign stream (cpu/src/wzp_cpu.cpp:769-949) on gfx950.
//
// Stream: [groups x u32 group offset][chunks x 24-bit meta][payload], one chunk
// = 8192 u16 magnitudes, one group = 16 chunks; payload in logical chunk order.
// Chunk payload = magnitude bytes (RZE over BIT_2 planes, or raw LE u16 when
// RZE+2 >= raw size) then the signs of the nonzero magnitudes, LSB first.
//
// Encoder: one 256-thread workgroup (4 wave64s) per chunk. Each thread owns
// 32 magnitudes = one sign word = 64 transformed bytes = 8 bytes of the zero
// bitmap B0 = one byte of B1. All counts are packed two per 32-bit word into
// block-wide scans (wave shuffles + LDS). Each chunk writes its meta, its size
// and its payload into a 16-byte aligned staging slot and ends; no workgroup
// waits for another. A second kernel, one workgroup per 16-chunk group taken
// in ticket order, finds the group's payload offset by a decoupled lookback
// over group descriptors (aggregate / inclusive flags, relaxed agent-scope
// atomics), writes the group directory entry and moves the group's payloads
// from their slots to their final place in logical order.
//
// Decoder: one workgroup per chunk; the chunk offset is the group offset plus
// the sizes of the earlier chunks of its group (from their metas). RZE is
// undone top-down with popcount prefix ranks, then BIT_2, then signs are
// expanded under the nonzero mask. Every read is bounds-checked against the
// stream; inconsistencies raise kCorrupt.
#include "profile.hpp"
#include "wzp_encode.hpp"

namespace wz {

template <class Source>
__global__ void __launch_bounds__(256) wzp_encode_kernel(Source src, uint32_t n, uint32_t groups,
                                                         uint8_t* __restrict__ out, uint64_t capacity, WzpScratch w,
                                                         uint32_t* __restrict__ status) {
    encode_chunk(src, blockIdx.x, n, groups, out, capacity, w, status);
}

__global__ void __launch_bounds__(kPackThreads) wzp_pack_kernel(uint8_t* __restrict__ out, uint64_t capacity,
                                                                uint32_t chunks, uint32_t groups, WzpScratch w,
                                                                unsigned long long* __restrict__ total,
                                                                uint32_t* __restrict__ status) {
    __shared__ uint32_t g_sh;
    if (threadIdx.x == 0) g_sh = atomicAdd(w.ticket, 1u);
    __syncthreads();
    pack_group(g_sh, out, capacity, chunks, groups, w, total, status);
}

// The group pass after the encoder, on `s` or on the side stream.
void launch_pack(uint8_t* out, uint64_t capacity, const WzpShape& ws, const WzpScratch& w, unsigned long long* total,
                 uint32_t* status, hipStream_t s, const SideStream* side) {
    hipStream_t ps = s;
    if (side) {
        hip_check(hipEventRecord(side->ready, s), "event record");
        hip_check(hipStreamWaitEvent(side->stream, side->ready, 0), "stream wait");
        ps = side->stream;
    }
    {
        StageClock c("wzp_pack", ps);
        hipLaunchKernelGGL(wzp_pack_kernel, dim3(ws.groups), dim3(kPackThreads), 0, ps, out, capacity, ws.chunks,
                           ws.groups, w, total, status);
    }
    if (side) hip_check(hipEventRecord(side->done, ps), "event record");
}

void wzp_encode(const uint16_t* mag, const uint32_t* sign, uint32_t n, uint8_t* out, uint64_t capacity, WzpScratch w,
                unsigned long long* total, uint32_t* status, hipStream_t s, const SideStream* pack) {
    const WzpShape ws = wzp_shape(n);
    if (ws.chunks == 0) {
        if (pack) hip_check(hipEventRecord(pack->done, s), "event record");
        return;
    }
    hipLaunchKernelGGL(wzp_encode_kernel<MagSource>, dim3(ws.chunks), dim3(kThreads), 0, s, MagSource{mag, sign}, n,
                       ws.groups, out, capacity, w, status);
    launch_pack(out, capacity, ws, w, total, status, s, pack);
}

template <class T>
void quantize_encode_pairs(T* coeff, Dims d, double quantum, SideList qo, uint8_t* out, uint64_t capacity,
                           WzpScratch w, unsigned long long* total, uint32_t* status, hipStream_t s,
                           WriteBack wb, const SideStream* pack) {
    if (!pair_bricks_fit(d.nx, d.ny, d.nz, 16)) throw std::invalid_argument("field does not tile into brick pairs");
    const uint32_t n = d.nx * d.ny * d.nz;
    const WzpShape ws = wzp_shape(n);
    const BrickSource<T> src{coeff,
                             d.nx,
                             d.ny,
                             T(quantum),
                             T(1.0 / quantum),
                             qo,
                             wb == kWriteNone ? 0u : wb == kWriteCoarseBox ? half_len(d.nx) : d.nx,
                             wb == kWriteNone ? 0u : wb == kWriteCoarseBox ? half_len(d.ny) : d.ny,
                             wb == kWriteNone ? 0u : wb == kWriteCoarseBox ? half_len(d.nz) : d.nz};
    hipLaunchKernelGGL(wzp_encode_kernel<BrickSource<T>>, dim3(ws.chunks), dim3(kThreads), 0, s, src, n, ws.groups,
                       out, capacity, w, status);
    launch_pack(out, capacity, ws, w, total, status, s, pack);
}

// The write-back box of a fused quantizing encoder.
Dims writeback_box(Dims d, WriteBack wb) {
    if (wb == kWriteNone) return Dims{0, 0, 0};
    if (wb == kWriteCoarseBox) return Dims{half_len(d.nx), half_len(d.ny), half_len(d.nz)};
    return d;
}

void check_rank_bricks(Dims d, uint32_t bz, const uint16_t* rank) {
    if (!rank_bricks_fit(d.nx, d.ny, d.nz, bz) || reinterpret_cast<uintptr_t>(rank) % 8)
        throw std::invalid_argument("field does not tile into full bricks");
}

template <class T>
void quantize_encode_ranks(T* coeff, Dims d, uint32_t bz, const uint16_t* rank, double quantum, SideList qo,
                           uint8_t* out, uint64_t capacity, WzpScratch w, unsigned long long* total,
                           uint32_t* status, hipStream_t s, WriteBack wb, const SideStream* pack) {
    check_rank_bricks(d, bz, rank);
    const uint32_t n = d.nx * d.ny * d.nz;
    const WzpShape ws = wzp_shape(n);
    const Dims h = writeback_box(d, wb);
    const RankSource<T> src{coeff, RankBricks{rank, d.nx, d.ny, bz}, T(quantum), T(1.0 / quantum), qo, h.nx, h.ny,
                            h.nz};
    hipLaunchKernelGGL(wzp_encode_kernel<RankSource<T>>, dim3(ws.chunks), dim3(kThreads), 0, s, src, n, ws.groups,
                       out, capacity, w, status);
    launch_pack(out, capacity, ws, w, total, status, s, pack);
}

template <class Sink>
__global__ void __launch_bounds__(256) wzp_decode_kernel(const uint8_t* __restrict__ in, uint64_t bytes, uint32_t n,
                                                         uint32_t chunks, uint32_t groups, Sink sink,
                                                         uint32_t* __restrict__ status) {
    decode_chunk(in, bytes, n, chunks, groups, blockIdx.x, sink, status);
}

__global__ void __launch_bounds__(kDirThreads) wzp_directory_kernel(const uint8_t* __restrict__ in, uint64_t bytes,
                                                                   uint32_t n, uint32_t chunks, uint32_t groups,
                                                                   uint32_t* __restrict__ status) {
    __shared__ uint32_t len[kMaxGroups];
    __shared__ uint32_t sh[kDirThreads / 64];
    check_directory(in, bytes, n, chunks, groups, len, sh, status);
}

// Planned decode: the stream size and quantum come from the device record.
__global__ void __launch_bounds__(kDirThreads) wzp_directory_planned_kernel(const uint8_t* __restrict__ in,
                                                                           const DecodeRecord* __restrict__ rec,
                                                                           uint32_t n, uint32_t chunks,
                                                                           uint32_t groups,
                                                                           uint32_t* __restrict__ status) {
    __shared__ uint32_t len[kMaxGroups];
    __shared__ uint32_t sh[kDirThreads / 64];
    if (!rec->ok) return;
    check_directory(in, rec->wzp, n, chunks, groups, len, sh, status);
}

template <class T>
__global__ void __launch_bounds__(256) wzp_decode_pairs_planned_kernel(const uint8_t* __restrict__ in,
                                                                      const DecodeRecord* __restrict__ rec,
                                                                      uint32_t n, uint32_t chunks, uint32_t groups,
                                                                      T* __restrict__ coeff, uint32_t nx, uint32_t ny,
                                                                      uint32_t* __restrict__ status) {
    if (!rec->ok) return;
    decode_chunk(in, rec->wzp, n, chunks, groups, blockIdx.x, BrickSink<T>{coeff, nx, ny, T(rec->quantum)}, status);
}

template <class T>
__global__ void __launch_bounds__(256) wzp_decode_ranks_planned_kernel(const uint8_t* __restrict__ in,
                                                                      const DecodeRecord* __restrict__ rec,
                                                                      uint32_t n, uint32_t chunks, uint32_t groups,
                                                                      T* __restrict__ coeff, RankBricks g,
                                                                      uint32_t* __restrict__ status) {
    if (!rec->ok) return;
    decode_chunk(in, rec->wzp, n, chunks, groups, blockIdx.x, RankSink<T>{coeff, g, T(rec->quantum)}, status);
}

__global__ void __launch_bounds__(256) wzp_decode_mag_planned_kernel(const uint8_t* __restrict__ in,
                                                                    const DecodeRecord* __restrict__ rec, uint32_t n,
                                                                    uint32_t chunks, uint32_t groups, MagSink sink,
                                                                    uint32_t* __restrict__ status) {
    if (!rec->ok) return;
    decode_chunk(in, rec->wzp, n, chunks, groups, blockIdx.x, sink, status);
}

// A stream's chunk decoders, enqueued on s first (the device starts them as
// soon as the work before them ends), and its directory check, on the side
// stream when given (both only read the stream and report into the status).
template <class Decoders, class Directory>
void stream_decode(uint32_t n, hipStream_t s, const SideStream* side, Decoders decoders, Directory directory) {
    const WzpShape ws = wzp_shape(n);
    if (side) hip_check(hipEventRecord(side->ready, s), "event record");
    if (ws.chunks) decoders(ws);
    hipStream_t ds = s;
    if (side) {
        hip_check(hipStreamWaitEvent(side->stream, side->ready, 0), "stream wait");
        ds = side->stream;
    }
    if (ws.chunks) directory(ws, ds);
    if (side) hip_check(hipEventRecord(side->done, ds), "event record");
}

template <class T>
void decode_dequantize_pairs_planned(const uint8_t* in, const DecodeRecord* rec, Dims d, T* coeff, uint32_t* status,
                                     hipStream_t s) {
    if (!pair_bricks_fit(d.nx, d.ny, d.nz, 16)) throw std::invalid_argument("field does not tile into brick pairs");
    const uint32_t n = d.nx * d.ny * d.nz;
    const WzpShape ws = wzp_shape(n);
    if (!ws.chunks) return;
    hipLaunchKernelGGL(wzp_decode_pairs_planned_kernel<T>, dim3(ws.chunks), dim3(kThreads), 0, s, in, rec, n,
                       ws.chunks, ws.groups, coeff, d.nx, d.ny, status);
}

template <class T>
void decode_dequantize_ranks_planned(const uint8_t* in, const DecodeRecord* rec, Dims d, uint32_t bz,
                                     const uint16_t* rank, T* coeff, uint32_t* status, hipStream_t s) {
    check_rank_bricks(d, bz, rank);
    const uint32_t n = d.nx * d.ny * d.nz;
    const WzpShape ws = wzp_shape(n);
    if (!ws.chunks) return;
    hipLaunchKernelGGL(wzp_decode_ranks_planned_kernel<T>, dim3(ws.chunks), dim3(kThreads), 0, s, in, rec, n,
                       ws.chunks, ws.groups, coeff, RankBricks{rank, d.nx, d.ny, bz}, status);
}

void wzp_decode_planned(const uint8_t* in, const DecodeRecord* rec, uint32_t n, uint16_t* mag, uint32_t* sign,
                        uint32_t* status, hipStream_t s) {
    const WzpShape ws = wzp_shape(n);
    if (!ws.chunks) return;
    hipLaunchKernelGGL(wzp_decode_mag_planned_kernel, dim3(ws.chunks), dim3(kThreads), 0, s, in, rec, n, ws.chunks,
                       ws.groups, MagSink{mag, sign}, status);
}

void wzp_directory_planned(const uint8_t* in, const DecodeRecord* rec, uint32_t n, uint32_t* status, hipStream_t s) {
    const WzpShape ws = wzp_shape(n);
    if (!ws.chunks) return;
    hipLaunchKernelGGL(wzp_directory_planned_kernel, dim3(1), dim3(kDirThreads), 0, s, in, rec, n, ws.chunks,
                       ws.groups, status);
}

void wzp_decode(const uint8_t* in, uint64_t bytes, uint32_t n, uint16_t* mag, uint32_t* sign, uint32_t* status,
                hipStream_t s, const SideStream* side) {
    if (n && bytes < wzp_shape(n).header) throw std::runtime_error("WZP archive size is outside its valid bounds");
    stream_decode(
        n, s, side,
        [&](const WzpShape& ws) {
            hipLaunchKernelGGL(wzp_decode_kernel<MagSink>, dim3(ws.chunks), dim3(kThreads), 0, s, in, bytes, n,
                               ws.chunks, ws.groups, MagSink{mag, sign}, status);
        },
        [&](const WzpShape& ws, hipStream_t ds) {
            hipLaunchKernelGGL(wzp_directory_kernel, dim3(1), dim3(kDirThreads), 0, ds, in, bytes, n, ws.chunks,
                               ws.groups, status);
        });
}

template <class T>
void decode_dequantize_pairs(const uint8_t* in, uint64_t bytes, Dims d, double quantum, T* coeff, uint32_t* status,
                             hipStream_t s, const SideStream* side) {
    if (!pair_bricks_fit(d.nx, d.ny, d.nz, 16)) throw std::invalid_argument("field does not tile into brick pairs");
    const uint32_t n = d.nx * d.ny * d.nz;
    if (bytes < wzp_shape(n).header) throw std::runtime_error("WZP archive size is outside its valid bounds");
    stream_decode(
        n, s, side,
        [&](const WzpShape& ws) {
            hipLaunchKernelGGL(wzp_decode_kernel<BrickSink<T>>, dim3(ws.chunks), dim3(kThreads), 0, s, in, bytes, n,
                               ws.chunks, ws.groups, BrickSink<T>{coeff, d.nx, d.ny, T(quantum)}, status);
        },
        [&](const WzpShape& ws, hipStream_t ds) {
            hipLaunchKernelGGL(wzp_directory_kernel, dim3(1), dim3(kDirThreads), 0, ds, in, bytes, n, ws.chunks,
                               ws.groups, status);
        });
}

template <class T>
void decode_dequantize_ranks(const uint8_t* in, uint64_t bytes, Dims d, uint32_t bz, const uint16_t* rank,
                             double quantum, T* coeff, uint32_t* status, hipStream_t s, const SideStream* side) {
    check_rank_bricks(d, bz, rank);
    const uint32_t n = d.nx * d.ny * d.nz;
    if (bytes < wzp_shape(n).header) throw std::runtime_error("WZP archive size is outside its valid bounds");
    const RankSink<T> sink{coeff, RankBricks{rank, d.nx, d.ny, bz}, T(quantum)};
    stream_decode(
        n, s, side,
        [&](const WzpShape& ws) {
            hipLaunchKernelGGL(wzp_decode_kernel<RankSink<T>>, dim3(ws.chunks), dim3(kThreads), 0, s, in, bytes, n,
                               ws.chunks, ws.groups, sink, status);
        },
        [&](const WzpShape& ws, hipStream_t ds) {
            hipLaunchKernelGGL(wzp_directory_kernel, dim3(1), dim3(kDirThreads), 0, ds, in, bytes, n, ws.chunks,
                               ws.groups, status);
        });
}

#define WZ_RANKS(T)                                                                                               \
    template void quantize_encode_ranks<T>(T*, Dims, uint32_t, const uint16_t*, double, SideList, uint8_t*,      \
                                           uint64_t, WzpScratch, unsigned long long*, uint32_t*, hipStream_t,    \
                                           WriteBack, const SideStream*);                                          \
    template void decode_dequantize_ranks<T>(const uint8_t*, uint64_t, Dims, uint32_t, const uint16_t*, double,  \
                                             T*, uint32_t*, hipStream_t, const SideStream*);                       \
    template void decode_dequantize_ranks_planned<T>(const uint8_t*, const DecodeRecord*, Dims, uint32_t,        \
                                                     const uint16_t*, T*, uint32_t*, hipStream_t);
WZ_RANKS(float)
WZ_RANKS(double)
#undef WZ_RANKS

template void quantize_encode_pairs<float>(float*, Dims, double, SideList, uint8_t*, uint64_t, WzpScratch,
                                           unsigned long long*, uint32_t*, hipStream_t, WriteBack, const SideStream*);
template void quantize_encode_pairs<double>(double*, Dims, double, SideList, uint8_t*, uint64_t, WzpScratch,
                                            unsigned long long*, uint32_t*, hipStream_t, WriteBack,
                                            const SideStream*);
template void decode_dequantize_pairs<float>(const uint8_t*, uint64_t, Dims, double, float*, uint32_t*,
                                             hipStream_t, const SideStream*);
template void decode_dequantize_pairs<double>(const uint8_t*, uint64_t, Dims, double, double*, uint32_t*,
                                              hipStream_t, const SideStream*);
template void decode_dequantize_pairs_planned<float>(const uint8_t*, const DecodeRecord*, Dims, float*, uint32_t*,
                                                     hipStream_t);
template void decode_dequantize_pairs_planned<double>(const uint8_t*, const DecodeRecord*, Dims, double*, uint32_t*,
                                                      hipStream_t);

}  // namespace wz
