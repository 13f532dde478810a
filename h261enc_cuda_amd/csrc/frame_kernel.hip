/*
 * frame_kernel.hip -- per-stream frame control on the device: rate control
 * (picture init and update, src/h261_ratectrl.c:107-245), the exact size
 * of the coded frame, and the advance of every stream to its next frame.
 *
 * The frame size needs the syntax state the reference writer carries along
 * each GOB (src/h261_gob.c:58-170, src/h261_macroblock.c:94-245): the MBA
 * increment since the last coded MB, the MVD predictor (reset at MBs 1, 12,
 * 23 and after every non-MC MB) and the MQUANT state.  One wave per stream
 * scans its GOBs, one lane per GOB, adding those bits to the per-MB
 * CBP + TCOEFF bits the macroblock coder counted.
 */
#include <hip/hip_runtime.h>

#include "h261_vlc.h"
#include "h261gpu_internal.h"

namespace h261g {

namespace {

__constant__ int32_t c_qscale[63] = {
    0,    1,    1,    1,    1,    1,    2,    2,    2,    3,    3,    4,    4,    5,    6,    6,
    8,    9,    10,   12,   13,   16,   18,   21,   24,   27,   32,   36,   42,   48,   55,   64,
    73,   84,   97,   111,  128,  147,  168,  194,  222,  256,  294,  337,  388,  445,  512,  588,
    675,  776,  891,  1024, 1176, 1351, 1552, 1782, 2048, 2352, 2702, 3104, 3565, 4096, 4705};

__constant__ Vlc c_mba[35] = H261G_MBA_TABLE;
__constant__ Vlc c_mtype[10] = H261G_MTYPE_TABLE;
__constant__ Vlc c_mvd[32] = H261G_MVD_TABLE;

constexpr int kRcBits = 6;
constexpr int kRcOne = 1 << kRcBits;
constexpr int kRcHalf = 1 << (kRcBits - 1);
constexpr int kRcTimeScale = 1918; /* 29.97 frames/s, Q6 (ratectrl.c:7-8) */

__device__ __forceinline__ int imin(int a, int b) { return a < b ? a : b; }
__device__ __forceinline__ int imax(int a, int b) { return a > b ? a : b; }
__device__ __forceinline__ int rc_mul(int a, int b) { return (int)(((int64_t)a * (int64_t)b) >> kRcBits); }
__device__ __forceinline__ int rc_div(int a, int b) { return (int)((((int64_t)a) << kRcBits) / b); }
__device__ __forceinline__ int rc_qscale_for_quant(int q) { return c_qscale[imin(imax(q, 0), 31) + 31]; }

/* ratectrl.c:63-104 */
__device__ int rc_quant_for_qscale(int qs)
{
    int pos = 31, step = 16, lo, hi;
    while (step) {
        if (c_qscale[pos] < qs)
            pos += step;
        else
            pos -= step;
        step >>= 1;
    }
    if (c_qscale[pos] < qs) {
        lo = pos;
        hi = imin(pos + 1, 62);
    } else {
        lo = imax(0, pos - 1);
        hi = pos;
    }
    pos = (qs - c_qscale[lo] < c_qscale[hi] - qs) ? lo : hi;
    return pos - 31;
}

/* h261_ratectrl_collect_picture_stats + h261_ratectrl_init_picture +
 * the clamp of h261_init_picture (ratectrl.c:130-195, 225-245,
 * picture.c:25-33) */
__global__ void rc_begin_kernel(RcState *rc_all, const StreamDesc *streams, int n)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const StreamDesc &d = streams[s];
    if (!(d.flags & kStreamActive)) return;
    RcState &rc = rc_all[s];
    const int num_mb = d.mb_w * d.mb_h;
    rc.picture_cost = imax(rc.cost_accum, num_mb * 100);
    rc.cost_accum = 0;
    const int ft = d.frame_type;
    rc.vbv_fill += rc_div(rc.bitrate, kRcTimeScale);
    rc.vbv_fill = imin(rc.vbv_fill, rc.vbv_size);
    const int cost_scale = rc_div(rc.picture_cost, rc.prediction_cost[ft]);
    int pred = rc_mul(cost_scale, rc.prediction_size[ft]);
    pred = rc_div(pred, rc_qscale_for_quant(rc.current_quant));
    int target = rc_div(rc.bitrate, kRcTimeScale);
    if (ft == H261_FRAME_TYPE_INTRA) target *= 20;
    const int up = rc_div(rc.vbv_fill, rc.vbv_size);
    const int lower = target - rc_mul(target, kRcOne - up);
    target = imax(lower, 10); /* the reference keeps only this bound (ratectrl.c:177-178) */
    int qss = rc_div(pred, target);
    if (qss < kRcOne) qss = rc_mul(kRcOne - up, kRcOne) + rc_mul(up, qss);
    rc.current_quant += rc_quant_for_qscale(qss);
    rc.current_quant = imin(imax(rc.current_quant, 1), 31);
    int q = rc.current_quant + (((kRcOne - up) * 4) >> kRcBits);
    if (rc.fixed_quant > 0) q = rc.fixed_quant;
    rc.pic_quant = imin(31, imax(1, q));
}

/* frame size (GOB scan) and h261_ratectrl_update_picture
 * (ratectrl.c:198-222) */
__global__ __launch_bounds__(64) void frame_bits_kernel(RcState *rc_all, const StreamDesc *streams, int n,
                                                         const MbSide *side, int32_t *bytes_out)
{
    const int s = blockIdx.x;
    const StreamDesc &d = streams[s];
    if (!(d.flags & kStreamActive)) return;
    RcState &rc = rc_all[s];
    const int lane = threadIdx.x;
    const bool cif = d.mb_w == 22;
    const int n_gob = cif ? 12 : (d.mb_h / 3) * (d.mb_w / 11);
    const int gob_cols = d.mb_w / 11;
    const int picq = rc.pic_quant;
    int bits = 0;
    if (lane < n_gob) {
        /* GOB origin in MBs: CIF GOBs 1..12 in two columns, QCIF 1, 3, 5 */
        const int gx = (lane % gob_cols) * 11, gy = (lane / gob_cols) * 3;
        bits = kGobHeaderBits;
        int cq = picq, mv_valid = 0, lmx = 0, lmy = 0, addr = 0;
        for (int i = 0; i < 33; i++) {
            addr++;
            const int mx = gx + i % 11, my = gy + i / 11;
            if (i == 0 || i == 11 || i == 22) mv_valid = 0;
            const MbSide sd = side[d.mb_base + my * d.mb_w + mx];
            const int flags = sd.flags;
            if (!flags) {
                mv_valid = 0;
                continue;
            }
            bits += c_mba[addr].len;
            addr = 0;
            if (sd.type == H261_MB_TYPE_INTRA) {
                bits += sd.quant != cq ? c_mtype[1].len + 5 : c_mtype[0].len;
            } else {
                const bool coded_q = sd.quant != cq && (flags & H261_MB_COEFF);
                /* MTYPE by [quant coded][flags] (macroblock.c:131-150) */
                const int t0[8] = {-1, 4, -1, 7, 2, 5, -1, 8}, t1[8] = {-1, -1, -1, -1, 3, 6, -1, 9};
                const int t = coded_q ? t1[flags & 7] : t0[flags & 7];
                bits += (t >= 0 ? c_mtype[t].len : 0) + (coded_q ? 5 : 0);
                if (flags & H261_MB_MC) {
                    int dx = sd.mvx, dy = sd.mvy;
                    if (mv_valid) {
                        dx -= lmx;
                        dy -= lmy;
                    }
                    bits += c_mvd[(dx + 48) & 31].len + c_mvd[(dy + 48) & 31].len;
                }
            }
            bits += sd.bits;
            if (flags & H261_MB_COEFF) cq = sd.quant;
            if (flags & H261_MB_MC) {
                lmx = sd.mvx;
                lmy = sd.mvy;
                mv_valid = 1;
            } else
                mv_valid = 0;
        }
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) bits += __shfl_xor(bits, off, 64);
    if (lane == 0) {
        const int total = bits + kPictureHeaderBits;
        const int bytes = (total + 7) >> 3;
        const int ft = d.frame_type;
        rc.prediction_cost[ft] = rc_mul(rc.prediction_cost[ft], kRcHalf);
        rc.prediction_size[ft] = rc_mul(rc.prediction_size[ft], kRcHalf);
        rc.prediction_cost[ft] += rc.picture_cost;
        rc.prediction_size[ft] += rc_mul(bytes, rc_qscale_for_quant(picq));
        rc.vbv_fill -= bytes;
        rc.frame_bytes = bytes;
        rc.frame_bits = total;
        if (bytes_out) bytes_out[s] = bytes;
    }
}

}  // namespace

int launch_rc_begin(RcState *rc, const StreamDesc *streams, int n_streams, int, void *stream)
{
    if (n_streams <= 0) return 0;
    hipLaunchKernelGGL(rc_begin_kernel, dim3((n_streams + 63) / 64), dim3(64), 0, (hipStream_t)stream, rc, streams,
                       n_streams);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_frame_bits(RcState *rc, const StreamDesc *streams, int n_streams, const MbSide *side, int32_t *bytes_out,
                      void *stream)
{
    if (n_streams <= 0) return 0;
    hipLaunchKernelGGL(frame_bits_kernel, dim3(n_streams), dim3(64), 0, (hipStream_t)stream, rc, streams, n_streams,
                       side, bytes_out);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace h261g
