/*
 * frame_control.h -- per-stream frame control as device functions, shared by
 * the separate launches and the fused forms (frame_kernel.hip, the tail of
 * the motion-estimation kernel):
 *
 *   rc_begin_one        picture cost, h261_ratectrl_init_picture, the
 *                       fixed-quant override, h261_init_picture's clamp
 *                       (src/h261_ratectrl.c:130-195, 225-245,
 *                       src/h261_picture.c:25-33)
 *   frame_bits_group    exact size of the coded frame: picture and GOB
 *                       headers plus, per coded MB, the MBA / MTYPE / MQUANT
 *                       / MVD bits the writer's GOB state produces
 *                       (src/h261_gob.c:58-199, src/h261_macroblock.c:94-245)
 *                       added to the MB coder's CBP + TCOEFF count.  The
 *                       waves of a workgroup share out the GOBs; lane i = MB
 *                       i of a GOB; the previous coded MB
 *                       (MBA increment), the last MB with coefficients
 *                       (MQUANT state) and the MV predictor (MB i-1, reset at
 *                       MBs 1, 12, 23) come from ballots and lane shuffles.
 *   rc_update_one       h261_ratectrl_update_picture (ratectrl.c:198-222)
 *   denoise_update_lane macroblocks_update_dct_denoise (mbgpu:478-503)
 *   advance_one         the reconstruction becomes the next reference, the
 *                       next input frame is selected
 *   frame_tail_group    the four above that end a frame, in one workgroup
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
__constant__ Vlc c_cbp[64] = H261G_CBP_TABLE;
/* MTYPE index by [MQUANT coded][flags MC|FILTER|COEFF] (macroblock.c:131-150),
 * -1 where no MTYPE exists; packed as value + 1 in 4-bit fields so the lookup
 * is a shift of a constant */
constexpr uint64_t kMtypeOfFlags = [] {
    constexpr int8_t t[2][8] = {{-1, 4, -1, 7, 2, 5, -1, 8}, {-1, -1, -1, -1, 3, 6, -1, 9}};
    uint64_t v = 0;
    for (int q = 0; q < 2; q++)
        for (int f = 0; f < 8; f++) v |= (uint64_t)(t[q][f] + 1) << (4 * (8 * q + f));
    return v;
}();
__device__ __forceinline__ int mtype_of_flags(bool coded_q, int flags)
{
    return (int)((kMtypeOfFlags >> (4 * (8 * (coded_q ? 1 : 0) + flags))) & 15u) - 1;
}

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

/* inclusive prefix sum over the wave's lanes: a scan within each row of 16
 * lanes by DPP row shifts (lanes shifted in from outside the row read 0),
 * then the totals of the rows below added from three readlanes.  VALU
 * latency only: no LDS round trip as a lane shuffle would take. */
__device__ __forceinline__ int wave_inclusive_sum(int x, int lane)
{
    x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, true); /* row_shr:1 */
    x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, true); /* row_shr:2 */
    x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, true); /* row_shr:4 */
    x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, true); /* row_shr:8 */
    const int r0 = __builtin_amdgcn_readlane(x, 15), r1 = __builtin_amdgcn_readlane(x, 31),
              r2 = __builtin_amdgcn_readlane(x, 47);
    const int row = lane >> 4;
    return x + (row > 0 ? r0 : 0) + (row > 1 ? r1 : 0) + (row > 2 ? r2 : 0);
}

/* exclusive prefix sum over the wave's lanes */
__device__ __forceinline__ int wave_exclusive_sum(int v, int lane) { return wave_inclusive_sum(v, lane) - v; }

/* ORs the low `len` (<= 32) bits of code into a bitstream of 32-bit words at
 * bit position pos.  Words hold stream bytes in order (the first bit of the
 * stream is the most significant bit of byte 0), so the big-endian bit
 * layout is byte-swapped before the OR; neighbours share boundary words. */
__device__ __forceinline__ void or_bits(uint32_t *words, int pos, uint32_t code, int len)
{
    if (len <= 0) return;
    const int i = pos >> 5, off = pos & 31;
    if (off + len <= 32) {
        atomicOr(words + i, __builtin_bswap32(code << (32 - off - len)));
    } else {
        const int spill = off + len - 32;
        atomicOr(words + i, __builtin_bswap32(code >> spill));
        atomicOr(words + i + 1, __builtin_bswap32(code << (32 - spill)));
    }
}

/* the waves of a frame-tail workgroup share out the GOBs (wave w takes GOBs
 * w, w + kTailWaves, ...) and each scans its own: lane i = MB i of the GOB.
 * Every thread of the workgroup calls; returns the frame's total bits on
 * every thread.  The side information of a wave's GOBs is loaded before any
 * is scanned, and the code tables are copied to LDS first, so no step of a
 * scan waits on memory.
 *
 * With hdr and bits (the stream's bitstream words, at least the frame's
 * size): each coded MB's bit position, header codes (MBA, MTYPE, MQUANT,
 * MVD, CBP) and total bits go to hdr for the bitstream writer, the frame's
 * words are zeroed and the picture and GOB headers written
 * (src/h261_picture.c:38-71, src/h261_gob.c:58-199,
 * src/h261_macroblock.c:94-245). */
constexpr int kTailWaves = 4;
__device__ __forceinline__ int frame_bits_group(const StreamDesc &d, const MbSide *side, int picq, int wave, int lane,
                                                MbHdr *hdr = nullptr, uint32_t *bits = nullptr, int bits_words = 0)
{
    constexpr int kMaxGob = 12;
    constexpr int kPer = (kMaxGob + kTailWaves - 1) / kTailWaves; /* GOBs per wave */
    const bool cif = d.mb_w == 22;
    const int n_gob = cif ? 12 : 3;
    const int gob_cols = cif ? 2 : 1;
    const uint64_t below = lane ? ((~0ull) >> (64 - lane)) : 0ull;
    const bool valid = lane < 33;
    /* unconditional loads from clamped in-frame indexes (lanes 33-63 and
     * GOBs past the last read a valid entry and discard it), so all loads
     * issue back to back */
    const int ml = valid ? lane : 32;
    uint64_t raw[kPer];
#pragma unroll
    for (int k = 0; k < kPer; k++) {
        const int g = wave + k * kTailWaves;
        const int gg = g < n_gob ? g : n_gob - 1;
        const int gx = (gg % gob_cols) * 11, gy = (gg / gob_cols) * 3;
        raw[k] = *(const uint64_t *)(side + d.mb_base + (gy + ml / 11) * d.mb_w + gx + ml % 11);
    }
    /* MBA, MTYPE, MVD and CBP codes: code | len << 16; GOB totals */
    __shared__ uint32_t t_mba[35], t_mtype[10], t_mvd[32], t_cbp[64];
    __shared__ int gob_bits[kMaxGob];
    if (wave == 0) {
        const uint32_t *g_mba = (const uint32_t *)c_mba, *g_mtype = (const uint32_t *)c_mtype,
                       *g_mvd = (const uint32_t *)c_mvd, *g_cbp = (const uint32_t *)c_cbp;
        const uint32_t a = g_mba[lane < 35 ? lane : 0], b = g_mtype[lane < 10 ? lane : 0],
                       c = g_mvd[lane < 32 ? lane : 0], e = g_cbp[lane];
        if (lane < 35) t_mba[lane] = a;
        if (lane < 10) t_mtype[lane] = b;
        if (lane < 32) t_mvd[lane] = c;
        t_cbp[lane] = e;
    }
    __syncthreads();
    auto vcode = [](uint32_t e) { return e & 0xffffu; };
    auto vlen = [](uint32_t e) { return (int)((e >> 16) & 0xffu); };
    /* per own GOB and lane: header codes, their length, the MB's bits, and
     * the inclusive prefix of the bits over the GOB's MBs */
    uint64_t hcs[kPer];
    int hls[kPer], bs[kPer], incls[kPer];
#pragma unroll
    for (int k = 0; k < kPer; k++) {
        const int g = wave + k * kTailWaves;
        uint64_t v = (valid && g < n_gob) ? raw[k] : 0ull;
        MbSide sd;
        __builtin_memcpy(&sd, &v, sizeof(v));
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
        uint64_t hc = 0; /* header codes of the MB, first bit most significant */
        int hl = 0;
        auto app = [&](uint32_t c, int l) {
            hc = (hc << l) | c;
            hl += l;
        };
        if (coded) {
            const uint32_t ea = t_mba[addr];
            app(vcode(ea), vlen(ea));
            if (sd.type == H261_MB_TYPE_INTRA) {
                const uint32_t et = t_mtype[sd.quant != cq ? 1 : 0];
                app(vcode(et), vlen(et));
                if (sd.quant != cq) app((uint32_t)sd.quant, 5);
                b = hl + sd.bits;
            } else {
                const bool coded_q = sd.quant != cq && (flags & H261_MB_COEFF);
                const int t = mtype_of_flags(coded_q, flags & 7);
                if (t >= 0) app(vcode(t_mtype[t]), vlen(t_mtype[t]));
                if (coded_q) app((uint32_t)sd.quant, 5);
                if (flags & H261_MB_MC) {
                    int dx = sd.mvx, dy = sd.mvy;
                    if (mv_valid) {
                        dx -= pmvx;
                        dy -= pmvy;
                    }
                    const uint32_t vx = t_mvd[(dx + 48) & 31], vy = t_mvd[(dy + 48) & 31];
                    app(vcode(vx), vlen(vx));
                    app(vcode(vy), vlen(vy));
                }
                b = hl + sd.bits; /* sd.bits: CBP + TCOEFF */
                if (flags & H261_MB_COEFF) {
                    const uint32_t ec = t_cbp[(uint8_t)sd.cbp & 63];
                    app(vcode(ec), vlen(ec));
                }
            }
        }
        const int incl = wave_inclusive_sum(b, lane);
        hcs[k] = hc;
        hls[k] = hl;
        bs[k] = b;
        incls[k] = incl;
        if (g < n_gob && lane == 0) gob_bits[g] = __builtin_amdgcn_readlane(incl, 63);
    }
    __syncthreads();
    /* where each GOB starts: after the picture header and every earlier
     * GOB (its header and its MBs) */
    int start[kMaxGob];
    int pos = kPictureHeaderBits;
#pragma unroll
    for (int g = 0; g < kMaxGob; g++) {
        start[g] = pos;
        if (g < n_gob) pos += kGobHeaderBits + gob_bits[g];
    }
    if (hdr) {
#pragma unroll
        for (int k = 0; k < kPer; k++) {
            const int g = wave + k * kTailWaves;
            const int gs = g < n_gob ? start[g] : 0;
            if (g < n_gob && valid && bs[k] > 0) {
                const int gx = (g % gob_cols) * 11, gy = (g / gob_cols) * 3;
                MbHdr h;
                h.code = hcs[k];
                h.bitpos = gs + kGobHeaderBits + incls[k] - bs[k];
                h.len = (uint8_t)hls[k];
                h.valid = 1;
                h.bits = (uint16_t)bs[k];
                hdr[d.mb_base + (gy + lane / 11) * d.mb_w + gx + lane % 11] = h;
            }
        }
    }
    if (bits && ((pos + 31) >> 5) <= bits_words) {
        /* the frame's words start at zero; then the headers (the MB bits are
         * ORed in by the writer launch) */
        const int words = (pos + 31) >> 5;
        for (int i = wave * 64 + lane; i < words; i += kTailWaves * 64) bits[i] = 0u;
        __syncthreads();
        if (wave == 0 && lane == 0) {
            /* PSC, TR, PTYPE (split, document camera, freeze release, source
             * format, hi-res and spare bits as the reference sets them), PEI */
            const uint32_t ph = (16u << 12) | ((uint32_t)(d.frame_num & 31) << 7) | (1u << 4) | ((cif ? 1u : 0u) << 3) |
                                (1u << 2) | (1u << 1);
            bits[0] = __builtin_bswap32(ph);
        }
        if (lane < kPer) {
            /* GBSC, GN, GQUANT, GEI of this wave's GOBs, a lane each */
            const int g = wave + lane * kTailWaves;
            if (g < n_gob) {
                const uint32_t gn = cif ? g + 1 : 2 * g + 1;
                int gs = 0;
#pragma unroll
                for (int k = 0; k < kMaxGob; k++) gs = k == g ? start[k] : gs;
                or_bits(bits, gs, (1u << 10) | (gn << 6) | ((uint32_t)picq << 1), kGobHeaderBits);
            }
        }
    }
    return pos;
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
__device__ __forceinline__ void frame_tail_group(StreamDesc &d, SeqState &q, DenoiseState *dn, RcState &rc,
                                                 const MbSide *side, MbHdr *hdr, int32_t *bytes_out, int wave,
                                                 int lane)
{
    const int ci = dn->inter_count, ca = dn->intra_count;
    if (wave == 0) denoise_update_lane(dn, lane, ci, ca);
    const bool write = hdr && q.bits;
    const int bits = frame_bits_group(d, side, rc.pic_quant, wave, lane, write ? hdr : nullptr,
                                      write ? q.bits : nullptr, q.bits_words);
    __syncthreads();
    if (wave == 0 && lane == 0) {
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
