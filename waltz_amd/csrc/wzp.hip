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
                           bool coarse_box_only, const SideStream* pack) {
    if (!pair_bricks_fit(d.nx, d.ny, d.nz, 16)) throw std::invalid_argument("field does not tile into brick pairs");
    const uint32_t n = d.nx * d.ny * d.nz;
    const WzpShape ws = wzp_shape(n);
    const BrickSource<T> src{coeff,
                             d.nx,
                             d.ny,
                             T(quantum),
                             T(1.0 / quantum),
                             qo,
                             coarse_box_only ? half_len(d.nx) : d.nx,
                             coarse_box_only ? half_len(d.ny) : d.ny,
                             coarse_box_only ? half_len(d.nz) : d.nz};
    hipLaunchKernelGGL(wzp_encode_kernel<BrickSource<T>>, dim3(ws.chunks), dim3(kThreads), 0, s, src, n, ws.groups,
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

// The stream's directory check runs beside its chunk decoders (both only
// read the stream and report into `status`). bytes >= the stream header.
void wzp_check_directory(const uint8_t* in, uint64_t bytes, uint32_t n, uint32_t* status, hipStream_t s,
                         const SideStream* side) {
    const WzpShape ws = wzp_shape(n);
    if (ws.chunks == 0) {
        if (side) hip_check(hipEventRecord(side->done, s), "event record");
        return;
    }
    if (bytes < ws.header) throw std::runtime_error("WZP archive size is outside its valid bounds");
    hipStream_t ds = s;
    if (side) {
        hip_check(hipEventRecord(side->ready, s), "event record");
        hip_check(hipStreamWaitEvent(side->stream, side->ready, 0), "stream wait");
        ds = side->stream;
    }
    hipLaunchKernelGGL(wzp_directory_kernel, dim3(1), dim3(kDirThreads), 0, ds, in, bytes, n, ws.chunks, ws.groups,
                       status);
    if (side) hip_check(hipEventRecord(side->done, ds), "event record");
}

void wzp_decode(const uint8_t* in, uint64_t bytes, uint32_t n, uint16_t* mag, uint32_t* sign, uint32_t* status,
                hipStream_t s, const SideStream* side) {
    const WzpShape ws = wzp_shape(n);
    wzp_check_directory(in, bytes, n, status, s, side);
    if (ws.chunks == 0) return;
    hipLaunchKernelGGL(wzp_decode_kernel<MagSink>, dim3(ws.chunks), dim3(kThreads), 0, s, in, bytes, n, ws.chunks,
                       ws.groups, MagSink{mag, sign}, status);
}

template <class T>
void decode_dequantize_pairs(const uint8_t* in, uint64_t bytes, Dims d, double quantum, T* coeff, uint32_t* status,
                             hipStream_t s, const SideStream* side) {
    if (!pair_bricks_fit(d.nx, d.ny, d.nz, 16)) throw std::invalid_argument("field does not tile into brick pairs");
    const uint32_t n = d.nx * d.ny * d.nz;
    const WzpShape ws = wzp_shape(n);
    wzp_check_directory(in, bytes, n, status, s, side);
    hipLaunchKernelGGL(wzp_decode_kernel<BrickSink<T>>, dim3(ws.chunks), dim3(kThreads), 0, s, in, bytes, n,
                       ws.chunks, ws.groups, BrickSink<T>{coeff, d.nx, d.ny, T(quantum)}, status);
}

template void quantize_encode_pairs<float>(float*, Dims, double, SideList, uint8_t*, uint64_t, WzpScratch,
                                           unsigned long long*, uint32_t*, hipStream_t, bool, const SideStream*);
template void quantize_encode_pairs<double>(double*, Dims, double, SideList, uint8_t*, uint64_t, WzpScratch,
                                            unsigned long long*, uint32_t*, hipStream_t, bool, const SideStream*);
template void decode_dequantize_pairs<float>(const uint8_t*, uint64_t, Dims, double, float*, uint32_t*,
                                             hipStream_t, const SideStream*);
template void decode_dequantize_pairs<double>(const uint8_t*, uint64_t, Dims, double, double*, uint32_t*,
                                              hipStream_t, const SideStream*);

}  // namespace wz
