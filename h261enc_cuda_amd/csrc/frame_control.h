/*
 * frame_control.h -- per-stream frame control as device functions, shared by
 * the separate launches and the fused forms (frame_kernel.hip, the tail of
 * the motion-estimation kernel):
 *
 *   rc_begin_one        picture cost, h261_ratectrl_init_picture, the
 *                       fixed-quant override, h261_init_picture's clamp
 *                       (src/h261_ratectrl.c:130-195, 225-245,
 *                       src/h261_picture.c:25-33)
 *   frame_bits_wave     exact size of the coded frame: picture and GOB
 *                       headers plus, per coded MB, the MBA / MTYPE / MQUANT
 *                       / MVD bits the writer's GOB state produces
 *                       (src/h261_gob.c:58-199, src/h261_macroblock.c:94-245)
 *                       added to the MB coder's CBP + TCOEFF count.  One
 *                       wave: lane i = MB i of the GOB; the previous coded MB
 *                       (MBA increment), the last MB with coefficients
 *                       (MQUANT state) and the MV predictor (MB i-1, reset at
 *                       MBs 1, 12, 23) come from ballots and lane shuffles.
 *   rc_update_one       h261_ratectrl_update_picture (ratectrl.c:198-222)
 *   denoise_update_lane macroblocks_update_dct_denoise (mbgpu:478-503)
 *   advance_one         the reconstruction becomes the next reference, the
 *                       next input frame is selected
 *   frame_tail_wave     the four above that end a frame, in one wave
 *
 * Fusing stays fence-free: an agent-scope fence on this device writes back
 * and invalidates the XCD's L2, which costs far more than a launch when
 * every wave issues one.  The only fused piece inside a data-parallel
 * kernel is rate-control begin, whose single input (the frame's cost sum)
 * travels in the same 64-bit atomic as the completion count.
 */
#ifndef H261G_FRAME_CONTROL_H
#define H261G_FRAME_CONTROL_H

#include <hip/hip_runtime.h>

#include "h261_vlc.h"
#include "h261gpu_internal.h"

namespace h261g {
namespace fc {

__constant__ int32_t c_qscale[63] = {
    0,    1,    1,    1,    1,    1,    2,    2,    2,    3,    3,    4,    4,    5,    6,    6,
    8,    9,    10,   12,   13,   16,   18,   21,   24,   27,   32,   36,   42,   48,   55,   64,
    73,   84,   97,   111,  128,  147,  168,  194,  222,  256,  294,  337,  388,  445,  512,  588,
    675,  776,  891,  1024, 1176, 1351, 1552, 1782, 2048, 2352, 2702, 3104, 3565, 4096, 4705};
__constant__ Vlc c_mba[35] = H261G_MBA_TABLE;
__constant__ Vlc c_mtype[10] = H261G_MTYPE_TABLE;
__constant__ Vlc c_mvd[32] = H261G_MVD_TABLE;
/* MTYPE index by [MQUANT coded][flags MC|FILTER|COEFF] (macroblock.c:131-150) */
__constant__ int8_t c_mtype_of_flags[2][8] = {{-1, 4, -1, 7, 2, 5, -1, 8}, {-1, -1, -1, -1, 3, 6, -1, 9}};

constexpr int kRcBits = 6;
constexpr int kRcOne = 1 << kRcBits;
constexpr int kRcHalf = 1 << (kRcBits - 1);
constexpr int kRcTimeScale = 1918; /* 29.97 frames/s in Q6 (ratectrl.c:7-8) */

__device__ __forceinline__ int imin(int a, int b) { return a < b ? a : b; }
__device__ __forceinline__ int imax(int a, int b) { return a > b ? a : b; }
__device__ __forceinline__ int rc_mul(int a, int b) { return (int)(((int64_t)a * (int64_t)b) >> kRcBits); }
__device__ __forceinline__ int rc_div(int a, int b) { return (int)((((int64_t)a) << kRcBits) / b); }
__device__ __forceinline__ int rc_qscale_for_quant(int q) { return c_qscale[imin(imax(q, 0), 31) + 31]; }

/* ratectrl.c:63-104 */
__device__ __forceinline__ int rc_quant_for_qscale(int qs)
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

/* cost_sum: the MD costs of the frame's MBs summed */
__device__ __forceinline__ void rc_begin_one(RcState &rc, const StreamDesc &d, int cost_sum)
{
    const int num_mb = d.mb_w * d.mb_h;
    rc.picture_cost = imax(cost_sum, num_mb * 100);
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

__device__ __forceinline__ void rc_update_one(RcState &rc, int ft, int bytes)
{
    const int picq = rc.pic_quant;
    rc.prediction_cost[ft] = rc_mul(rc.prediction_cost[ft], kRcHalf);
    rc.prediction_size[ft] = rc_mul(rc.prediction_size[ft], kRcHalf);
    rc.prediction_cost[ft] += rc.picture_cost;
    rc.prediction_size[ft] += rc_mul(bytes, rc_qscale_for_quant(picq));
    rc.vbv_fill -= bytes;
}

/* all 64 lanes call; returns the frame's total bits on every lane.  The side
 * information of every GOB is loaded before any is scanned, so the loads
 * overlap instead of costing one memory latency per GOB. */
__device__ __forceinline__ int frame_bits_wave(const StreamDesc &d, const MbSide *side, int picq, int lane)
{
    constexpr int kMaxGob = 12;
    const bool cif = d.mb_w == 22;
    const int n_gob = cif ? 12 : 3;
    const int gob_cols = cif ? 2 : 1;
    const uint64_t below = lane ? ((~0ull) >> (64 - lane)) : 0ull;
    const bool valid = lane < 33;
    /* unconditional loads from clamped in-frame indexes (lanes 33-63 and
     * GOBs past the last read a valid entry and discard it), so all loads
     * issue back to back */
    const int ml = valid ? lane : 32;
    uint64_t raw[kMaxGob];
#pragma unroll
    for (int g = 0; g < kMaxGob; g++) {
        const int gg = g < n_gob ? g : n_gob - 1;
        const int gx = (gg % gob_cols) * 11, gy = (gg / gob_cols) * 3;
        raw[g] = *(const uint64_t *)(side + d.mb_base + (gy + ml / 11) * d.mb_w + gx + ml % 11);
    }
    MbSide sds[kMaxGob];
#pragma unroll
    for (int g = 0; g < kMaxGob; g++) {
        const uint64_t v = (valid && g < n_gob) ? raw[g] : 0ull;
        __builtin_memcpy(&sds[g], &v, sizeof(v));
    }
    int bits = 0;
#pragma unroll
    for (int g = 0; g < kMaxGob; g++) {
        if (g >= n_gob) break;
        const MbSide sd = sds[g];
        const int flags = sd.flags;
        const bool coded = flags != 0;
        const uint64_t cm = __ballot(coded);
        const uint64_t qm = __ballot(coded && (flags & H261_MB_COEFF));
        /* MBA increment since the previous coded MB of the GOB */
        const uint64_t cb = cm & below;
        const int prev = cb ? 63 - __builtin_clzll(cb) : -1;
        const int addr = lane - prev;
        /* quantiser state: the last MB with coefficients before this one */
        const uint64_t qb = qm & below;
        const int qsrc = qb ? 63 - __builtin_clzll(qb) : lane;
        const int qv = __shfl((int)sd.quant, qsrc, 64);
        const int cq = qb ? qv : picq;
        /* MV predictor: MB i-1 coded with MC, not at the start of a GOB row */
        const int src = lane ? lane - 1 : 0;
        const int pflags = __shfl(flags, src, 64);
        const int pmvx = __shfl((int)sd.mvx, src, 64), pmvy = __shfl((int)sd.mvy, src, 64);
        const bool mv_valid = lane % 11 != 0 && (pflags & H261_MB_MC);
        int b = 0;
        if (coded) {
            b += c_mba[addr].len;
            if (sd.type == H261_MB_TYPE_INTRA) {
                b += sd.quant != cq ? c_mtype[1].len + 5 : c_mtype[0].len;
            } else {
                const bool coded_q = sd.quant != cq && (flags & H261_MB_COEFF);
                const int t = c_mtype_of_flags[coded_q ? 1 : 0][flags & 7];
                b += (t >= 0 ? c_mtype[t].len : 0) + (coded_q ? 5 : 0);
                if (flags & H261_MB_MC) {
                    int dx = sd.mvx, dy = sd.mvy;
                    if (mv_valid) {
                        dx -= pmvx;
                        dy -= pmvy;
                    }
                    b += c_mvd[(dx + 48) & 31].len + c_mvd[(dy + 48) & 31].len;
                }
            }
            b += sd.bits;
        }
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) b += __shfl_xor(b, off, 64);
        bits += b + kGobHeaderBits;
    }
    return bits + kPictureHeaderBits;
}

/* 64 lanes of one wave; the caller provides a workgroup barrier between the
 * statistic reads and the count reset */
__device__ __forceinline__ void denoise_update_lane(DenoiseState *dn, int i, int ci, int ca)
{
    const int ni = dn->inter_noise[i], na = dn->intra_noise[i];
    dn->inter_offset[i] = (ci * dn->denoise + ni / 2) / (ni + 1);
    dn->intra_offset[i] = (ca * dn->denoise + na / 2) / (na + 1);
    dn->inter_noise[i] = 0;
    dn->intra_noise[i] = 0;
}

__device__ __forceinline__ void advance_one(StreamDesc &d, SeqState &q)
{
    q.pp ^= 1;
    const size_t ysz = (size_t)d.width * d.height, csz = ysz / 4;
    uint8_t *ref = q.recon[q.pp], *rec = q.recon[q.pp ^ 1];
    d.ref[0] = ref;
    d.ref[1] = ref + ysz;
    d.ref[2] = ref + ysz + csz;
    d.rec[0] = rec;
    d.rec[1] = rec + ysz;
    d.rec[2] = rec + ysz + csz;
    d.frame_num++;
    d.frame_type = H261_FRAME_TYPE_INTER;
    if (d.flags & kStreamSeqInput) {
        q.remaining--;
        if (q.remaining <= 0) {
            d.flags &= ~kStreamActive;
        } else {
            const uint8_t *cur = q.base + (size_t)q.frame_size * d.frame_num;
            d.cur[0] = cur;
            d.cur[1] = cur + ysz;
            d.cur[2] = cur + ysz + csz;
        }
    } else {
        d.flags &= ~kStreamActive; /* a host-fed stream waits for its next frame */
    }
}

/* the end of one stream's frame, all 64 lanes of one wave: denoise offsets
 * for the next frame, exact frame size, rate-control update, advance.  Reads
 * only what earlier launches wrote (statistics, side information). */
__device__ __forceinline__ void frame_tail_wave(StreamDesc &d, SeqState &q, DenoiseState *dn, RcState &rc,
                                                const MbSide *side, int32_t *bytes_out, int lane)
{
    const int ci = dn->inter_count, ca = dn->intra_count;
    denoise_update_lane(dn, lane, ci, ca);
    const int bits = frame_bits_wave(d, side, rc.pic_quant, lane);
    __syncthreads();
    if (lane == 0) {
        dn->inter_count = 0;
        dn->intra_count = 0;
        const int bytes = (bits + 7) >> 3;
        rc.frame_bits = bits;
        rc.frame_bytes = bytes;
        rc_update_one(rc, d.frame_type, bytes);
        if (bytes_out) *bytes_out = bytes;
        advance_one(d, q);
    }
}

}  // namespace fc
}  // namespace h261g

#endif
