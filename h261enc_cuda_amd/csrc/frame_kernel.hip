/*
 * frame_kernel.hip -- per-stream frame control launches: the fused frame
 * tail (denoise update, frame size, rate-control update, advance: one wave
 * per stream) of the engine's default mode, and the separate kernels of its
 * unfused mode and of the legacy path's denoise update.
 */
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include "frame_control.h"
#include "h261gpu_internal.h"

namespace h261g {

namespace {

__global__ void rc_begin_kernel(RcState *rc_all, const StreamDesc *streams, int n, int32_t mb_lo, int32_t mb_hi)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const StreamDesc &d = streams[s];
    if (!(d.flags & kStreamActive) || d.mb_base < mb_lo || d.mb_base >= mb_hi) return;
    RcState &rc = rc_all[s];
    fc::rc_begin_one(rc, d, rc.cost_accum);
    rc.cost_accum = 0;
}

__global__ __launch_bounds__(64 * fc::kTailWaves) void frame_bits_kernel(RcState *rc_all, const StreamDesc *streams, int n,
                                                         const MbSide *side, int32_t *bytes_out, const SeqState *seq,
                                                         MbHdr *hdr)
{
    const int s = blockIdx.x;
    const StreamDesc &d = streams[s];
    if (!(d.flags & kStreamActive)) return;
    RcState &rc = rc_all[s];
    const bool write = hdr && seq && seq[s].bits;
    const int bits = fc::frame_bits_group(d, side, rc.pic_quant, threadIdx.x >> 6, threadIdx.x & 63,
                                          write ? hdr : nullptr, write ? seq[s].bits : nullptr,
                                          write ? seq[s].bits_words : 0);
    if (threadIdx.x == 0) {
        const int bytes = (bits + 7) >> 3;
        fc::rc_update_one(rc, d.frame_type, bytes);
        rc.frame_bytes = bytes;
        rc.frame_bits = bits;
        if (bytes_out) bytes_out[s] = bytes;
    }
}

__global__ __launch_bounds__(64) void denoise_update_kernel(DenoiseState *dn_all, const StreamDesc *streams)
{
    const int s = blockIdx.x;
    if (!(streams[s].flags & kStreamActive)) return;
    DenoiseState *dn = dn_all + s;
    const int ci = dn->inter_count, ca = dn->intra_count;
    fc::denoise_update_lane(dn, threadIdx.x, ci, ca);
    __syncthreads();
    if (threadIdx.x == 0) {
        dn->inter_count = 0;
        dn->intra_count = 0;
    }
}

__global__ __launch_bounds__(64 * fc::kTailWaves) void frame_tail_kernel(StreamDesc *streams, SeqState *seq, DenoiseState *dn_all,
                                                         RcState *rc_all, const MbSide *side, MbHdr *hdr,
                                                         int32_t *bytes_out, int32_t mb_lo, int32_t mb_hi)
{
    const int s = blockIdx.x;
    StreamDesc &d = streams[s];
    if (!(d.flags & kStreamActive) || d.mb_base < mb_lo || d.mb_base >= mb_hi) return;
    fc::frame_tail_group(d, seq[s], dn_all + s, rc_all[s], side, hdr, bytes_out ? bytes_out + s : nullptr,
                         threadIdx.x >> 6, threadIdx.x & 63);
}

__global__ void advance_kernel(StreamDesc *streams, SeqState *seq, int n)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    StreamDesc &d = streams[s];
    if (!(d.flags & kStreamActive)) return;
    fc::advance_one(d, seq[s]);
}

}  // namespace

int launch_frame_tail(StreamDesc *streams, SeqState *seq, DenoiseState *dn, RcState *rc, const MbSide *side,
                      MbHdr *hdr, int n_streams, void *stream, void *ev_start, void *ev_stop, int32_t mb_lo,
                      int32_t mb_hi)
{
    if (n_streams <= 0) {
        if (ev_start) { /* nothing to launch: an empty timed interval */
            (void)hipEventRecord((hipEvent_t)ev_start, (hipStream_t)stream);
            (void)hipEventRecord((hipEvent_t)ev_stop, (hipStream_t)stream);
        }
        return 0;
    }
    int32_t *no_bytes = nullptr;
    if (ev_start)
        hipExtLaunchKernelGGL(frame_tail_kernel, dim3(n_streams), dim3(64 * fc::kTailWaves), 0, (hipStream_t)stream,
                              (hipEvent_t)ev_start, (hipEvent_t)ev_stop, 0, streams, seq, dn, rc, side, hdr, no_bytes, mb_lo,
                              mb_hi);
    else
        hipLaunchKernelGGL(frame_tail_kernel, dim3(n_streams), dim3(64 * fc::kTailWaves), 0, (hipStream_t)stream, streams, seq, dn, rc,
                           side, hdr, no_bytes, mb_lo, mb_hi);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_advance(StreamDesc *streams, SeqState *seq, int n_streams, void *stream)
{
    if (n_streams <= 0) return 0;
    hipLaunchKernelGGL(advance_kernel, dim3((n_streams + 63) / 64), dim3(64), 0, (hipStream_t)stream, streams, seq,
                       n_streams);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_rc_begin(RcState *rc, const StreamDesc *streams, int n_streams, void *stream, int32_t mb_lo, int32_t mb_hi)
{
    if (n_streams <= 0) return 0;
    hipLaunchKernelGGL(rc_begin_kernel, dim3((n_streams + 63) / 64), dim3(64), 0, (hipStream_t)stream, rc, streams,
                       n_streams, mb_lo, mb_hi);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_frame_bits(RcState *rc, const StreamDesc *streams, int n_streams, const MbSide *side, int32_t *bytes_out,
                      void *stream, const SeqState *seq, MbHdr *hdr)
{
    if (n_streams <= 0) return 0;
    hipLaunchKernelGGL(frame_bits_kernel, dim3(n_streams), dim3(64 * fc::kTailWaves), 0, (hipStream_t)stream, rc, streams, n_streams,
                       side, bytes_out, seq, hdr);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_denoise_update(DenoiseState *dn, const StreamDesc *streams, int n_streams, void *stream)
{
    if (n_streams <= 0) return 0;
    hipLaunchKernelGGL(denoise_update_kernel, dim3(n_streams), dim3(64), 0, (hipStream_t)stream, dn, streams);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace h261g
