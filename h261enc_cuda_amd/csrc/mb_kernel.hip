/*
 * mb_kernel.hip -- the fused macroblock coder for gfx950.
 *
 * The reference runs 12 kernels per frame with a thread per 8x8 block and a
 * global-memory round trip through an interleaved Int32 coefficient buffer
 * between each (src/gpu_device.cu:592-815).  Here one wave64 codes a whole
 * macroblock (6 blocks, 384 samples) in registers and LDS:
 *
 *   lanes 0-47 = (block b, row r):   MC / loop filter, residual, DCT rows
 *   LDS transpose
 *   lanes 0-47 = (block b, col c):   DCT columns, denoise, quantiser, CBP,
 *                                    dequantiser, IDCT columns
 *   LDS transpose
 *   lanes 0-47 = (block b, row r):   IDCT rows, reconstruction
 *
 * Arithmetic follows the reference op for op (compiled with
 * -ffp-contract=off and '#pragma clang fp contract(off)'): products with an
 * unsuffixed literal are double multiplies rounded to float; the forward
 * DCT rounds with floorf((float)(x + 0.49999)), the IDCT with
 * floor((double)x + 0.49999) (src/mbgpu_kernels.cuh:337-442, 827-931).
 * The denoise uses the previous frame's offsets and accumulates this
 * frame's |coefficient| statistics (LDS atomics, one global flush per
 * workgroup); the CBP is the OR of this frame's coded blocks, the intended
 * result of the reference's racy reset/atomicOr (mbgpu:676-721).
 */
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include "h261_vlc.h"
#include "h261gpu_internal.h"

#pragma clang fp contract(off)

namespace h261g {

namespace {

/* float tables initialised from double literals (mbgpu:335, 823) */
#define H261G_FWD_POST {1.0, 0.720960, 0.765367, 0.850430, 1.0, 1.272759, 1.847759, 3.624510}
#define H261G_INV_PRE {1.0, 1.387040, 1.306563, 1.175876, 1.0, 0.785695, 0.541196, 0.275899}
/* indexed by the lane's column (memory) and by unrolled constants (folded) */
__constant__ float c_fwd_post[8] = H261G_FWD_POST;
__constant__ float c_inv_pre[8] = H261G_INV_PRE;
constexpr float kFwdPost[8] = H261G_FWD_POST;
constexpr float kInvPre[8] = H261G_INV_PRE;
#undef H261G_FWD_POST
#undef H261G_INV_PRE

/* f_inv_quant = 1.0 / (float)(2q) as a double quotient rounded to float
 * (mbgpu:685); folded exactly at compile time */
#define H261G_INVQ(q) ((float)(1.0 / (double)(float)(2 * (q))))
__constant__ float c_inv_quant[32] = {
    0.0f,           H261G_INVQ(1),  H261G_INVQ(2),  H261G_INVQ(3),  H261G_INVQ(4),  H261G_INVQ(5),
    H261G_INVQ(6),  H261G_INVQ(7),  H261G_INVQ(8),  H261G_INVQ(9),  H261G_INVQ(10), H261G_INVQ(11),
    H261G_INVQ(12), H261G_INVQ(13), H261G_INVQ(14), H261G_INVQ(15), H261G_INVQ(16), H261G_INVQ(17),
    H261G_INVQ(18), H261G_INVQ(19), H261G_INVQ(20), H261G_INVQ(21), H261G_INVQ(22), H261G_INVQ(23),
    H261G_INVQ(24), H261G_INVQ(25), H261G_INVQ(26), H261G_INVQ(27), H261G_INVQ(28), H261G_INVQ(29),
    H261G_INVQ(30), H261G_INVQ(31)};
#undef H261G_INVQ

__constant__ uint8_t c_zigzag[64] = H261G_ZIGZAG;

/* bits of each TCOEFF (run, |level|) event: code length plus the sign bit,
 * kEscapeBits where the table has no code; four entries per dword */
constexpr Vlc kTcoeff[kTcoeffRuns][kTcoeffLevels] = H261G_TCOEFF_TABLE;
constexpr int kTlenWords = kTcoeffRuns * kTcoeffLevels / 4; /* 108 */
struct TlenTable {
    uint32_t w[kTlenWords];
};
constexpr TlenTable make_tlen()
{
    TlenTable t{};
    for (int i = 0; i < kTcoeffRuns * kTcoeffLevels; i++) {
        const int len = kTcoeff[i / kTcoeffLevels][i % kTcoeffLevels].len;
        const uint32_t b = (uint32_t)(len ? len + 1 : kEscapeBits);
        t.w[i / 4] |= b << (8 * (i % 4));
    }
    return t;
}
__constant__ TlenTable c_tlen = make_tlen();
__constant__ Vlc c_cbp[64] = H261G_CBP_TABLE;

constexpr int kCbpBit[6] = {32, 16, 8, 4, 2, 1};
constexpr int kBlkPitch = 72; /* floats per block in the transpose buffer */

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* the 1-D AAN forward transform shared by the row and column passes
 * (mbgpu:350-387 and 399-437); o[k] gets output k before the postscale */
struct Fwd8 {
    float o[8];
};

__device__ __forceinline__ void fdct_core(float tmp0, float tmp1, float tmp2, float tmp3, float tmp4,
                                          float tmp5, float tmp6, float tmp7, float *o)
{
    float tmp10 = tmp0 + tmp3, tmp13 = tmp0 - tmp3, tmp11 = tmp1 + tmp2, tmp12 = tmp1 - tmp2;
    o[0] = tmp10 + tmp11;
    o[4] = tmp10 - tmp11;
    float z1 = tmp12 + tmp13;
    z1 = (float)((double)z1 * 0.707107);
    o[2] = tmp13 + z1;
    o[6] = tmp13 - z1;
    tmp4 += tmp5;
    tmp5 += tmp6;
    tmp6 += tmp7;
    float z5 = (float)((double)(tmp4 - tmp6) * 0.382683);
    float z2 = (float)((double)tmp4 * 0.541196 + (double)z5);
    float z4 = (float)((double)tmp6 * 1.306563 + (double)z5);
    float z3 = (float)((double)tmp5 * 0.707107);
    float z11 = tmp7 + z3, z13 = tmp7 - z3;
    o[5] = z13 + z2;
    o[3] = z13 - z2;
    o[1] = z11 + z4;
    o[7] = z11 - z4;
}

/* the 1-D AAN inverse transform of the column and row passes
 * (mbgpu:851-887 and 892-919); returns outputs 0..7 before rounding */
__device__ __forceinline__ void idct_core(const float *t, float *o)
{
    float tmp13 = t[2] + t[6];
    float tmp12 = t[2] - t[6];
    tmp12 = (float)((double)tmp12 * 1.414214);
    tmp12 -= tmp13;
    float tmp10 = t[0] + t[4], tmp11 = t[0] - t[4];
    float tmp0 = tmp10 + tmp13, tmp1 = tmp11 + tmp12, tmp2 = tmp11 - tmp12, tmp3 = tmp10 - tmp13;
    float z11 = t[1] + t[7], z12 = t[1] - t[7], z13 = t[5] + t[3], z10 = t[5] - t[3];
    tmp13 = z11 + z13;
    tmp11 = (float)((double)(z11 - z13) * 1.414214);
    float z5 = (float)((double)(z12 + z10) * 1.84776);
    tmp10 = (float)((double)z12 * 1.082392 - (double)z5);
    tmp12 = (float)((double)z10 * -2.613126 + (double)z5);
    tmp12 -= tmp13;
    tmp11 -= tmp12;
    tmp10 += tmp11;
    o[0] = tmp0 + tmp13;
    o[1] = tmp1 + tmp12;
    o[2] = tmp2 + tmp11;
    o[3] = tmp3 - tmp10;
    o[4] = tmp3 + tmp10;
    o[5] = tmp2 - tmp11;
    o[6] = tmp1 - tmp12;
    o[7] = tmp0 - tmp13;
}

/* bits of one TCOEFF event: the code and its sign bit, or an escape; tlen
 * is the wave's LDS copy of the code lengths plus one (kEscapeBits where the
 * table has no code) */
__device__ __forceinline__ int tcoeff_bits(const uint8_t *tlen, int run, int level)
{
    const int a = level < 0 ? -level : level;
    return (run < kTcoeffRuns && a < kTcoeffLevels) ? (int)tlen[run * kTcoeffLevels + a] : kEscapeBits;
}

/* flushes this wave's denoise statistics of stream s (LDS -> global
 * atomics); the frame tail reads them in a later launch */
__device__ __forceinline__ void flush_segment(const MbLaunch &p, int s, int32_t (*stat)[64], int32_t *scount, int lane)
{
    __syncthreads();
    DenoiseState *dn = p.denoise + s;
    if (stat[0][lane]) atomicAdd(&dn->intra_noise[lane], stat[0][lane]);
    if (stat[1][lane]) atomicAdd(&dn->inter_noise[lane], stat[1][lane]);
    if (lane == 0 && scount[0]) atomicAdd(&dn->intra_count, scount[0]);
    if (lane == 1 && scount[1]) atomicAdd(&dn->inter_count, scount[1]);
    __syncthreads();
    stat[0][lane] = 0;
    stat[1][lane] = 0;
    if (lane < 2) scount[lane] = 0;
    __syncthreads();
}

__global__ __launch_bounds__(64, 5) void mb_kernel(MbLaunch p)
{
    __shared__ float tf[6 * kBlkPitch];
    __shared__ __align__(16) int32_t lv[6 * 64];
    __shared__ int32_t stat[2][64];
    __shared__ int32_t scount[2];
    __shared__ int32_t sblk[8];
    __shared__ __align__(16) uint32_t tlen32[kTlenWords];
    const uint8_t *tlen = (const uint8_t *)tlen32;

    const int lane = threadIdx.x;
    const int b = lane >> 3; /* block 0-5 (6, 7 idle) */
    const int rc = lane & 7; /* row in phases 1/3, column in phase 2 */
    const bool act = b < 6;

    stat[0][lane] = 0;
    stat[1][lane] = 0;
    if (lane < 2) scount[lane] = 0;
    if (lane < kTlenWords / 2) ((uint2 *)tlen32)[lane] = ((const uint2 *)c_tlen.w)[lane];
    /* per-lane constants of the loop: zig-zag position of lane, column
     * scale factors of lane (b, c) */
    const int zz = c_zigzag[lane];
    const float post_c = (float)((double)c_fwd_post[rc] * 0.125);
    const float pre_c = c_inv_pre[rc];

    const int n_wg = (p.n_mb + p.mbs_per_wave - 1) / p.mbs_per_wave;
    const int wg = launch_index(blockIdx.x, n_wg, p.xcd_remap);
    if (wg >= n_wg) return;
    for (int it = 0; it < p.mbs_per_wave; it++) {
        const int g = wg * p.mbs_per_wave + it;
        if (g >= p.n_mb) break;
        const int s = p.mb_stream[g];
        if (s < 0) continue; /* unmapped range of a closed stream */
        const StreamDesc &d = p.streams[s];
        if (!(d.flags & kStreamActive)) continue;
        const int m = g - d.mb_base;
        const MbMode md = p.mode[g];
        const int W = d.width;
        const int bx = (m % d.mb_w) * 16, by = (m / d.mb_w) * 16;
        const int picq = (d.flags & kStreamDeviceQuant) ? p.rc[s].pic_quant : d.pic_quant;
        const int quant = clampi(picq + md.quant_adjust, 1, 31); /* gpu_device.cu:625-627 */
        const bool intra = md.type == H261_MB_TYPE_INTRA;
        const bool filter = (md.flags & H261_MB_FILTER) != 0;
        const int t = intra ? 0 : 1;
        (void)t;

        /* ---- phase 1: lane (b, r) ---- */
        const int plane = b < 4 ? 0 : b - 3;
        int px, py, stride, mvx, mvy;
        if (b < 4) {
            px = bx + (b & 1) * 8;
            py = by + ((b >> 1) & 1) * 8;
            stride = W;
            mvx = md.mvx;
            mvy = md.mvy;
        } else { /* chroma vector = luma vector / 2, truncated (mbgpu:266-269) */
            px = bx >> 1;
            py = by >> 1;
            stride = W >> 1;
            mvx = md.mvx / 2;
            mvy = md.mvy / 2;
        }
        /* issue every global load of the MB before using any: the current
         * row, three reference rows (rc-1, rc, rc+1 clamped to the block --
         * the loop filter is 1-D at the block edges; an intra MB has a zero
         * vector, so its addresses are in the frame as well) as aligned
         * dwords, and this lane's eight denoise offsets */
        const int t_off = intra ? 0 : 1;
        uint32_t offw[8];
        uint32_t offp[4]; /* the offsets as UInt16 (mbgpu:525-528), two per register */
        int pred[8], res[8];
        uint2 cw = make_uint2(0, 0);
        uint32_t rw[3][3];
        int o = 0;
        if (act) {
            const int32_t *offs = t_off == 0 ? p.denoise[s].intra_offset : p.denoise[s].inter_offset;
#pragma unroll
            for (int r = 0; r < 8; r++) offw[r] = (uint32_t)offs[rc + 8 * r];
#pragma unroll
            for (int r = 0; r < 4; r++) offp[r] = (offw[2 * r] & 0xffffu) | (offw[2 * r + 1] << 16);
            cw = *(const uint2 *)(d.cur[plane] + (size_t)(py + rc) * stride + px);
            const int rx = px + mvx;
            const int a = rx & ~3;
            o = rx & 3;
            /* the third dword is needed only when o > 0, and then it lies in the row */
            const int a2 = a + 8 <= stride - 4 ? a + 8 : a;
            const uint8_t *rbase = d.ref[plane] + (size_t)(py + mvy) * stride;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const int rr = clampi(rc + k - 1, 0, 7);
                const uint8_t *q = rbase + (size_t)rr * stride;
                rw[k][0] = *(const uint32_t *)(q + a);
                rw[k][1] = *(const uint32_t *)(q + a + 4);
                rw[k][2] = *(const uint32_t *)(q + a2);
            }
        }
        if (act) {
            int cpel[8];
#pragma unroll
            for (int x = 0; x < 4; x++) {
                cpel[x] = (cw.x >> (8 * x)) & 0xff;
                cpel[x + 4] = (cw.y >> (8 * x)) & 0xff;
            }
            /* the 8 reference pels of rows rc-1, rc, rc+1 */
            int a8[3][8];
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const uint32_t lo = __builtin_amdgcn_alignbyte(rw[k][1], rw[k][0], o);
                const uint32_t hi = __builtin_amdgcn_alignbyte(rw[k][2], rw[k][1], o);
#pragma unroll
                for (int x = 0; x < 4; x++) {
                    a8[k][x] = (lo >> (8 * x)) & 0xff;
                    a8[k][x + 4] = (hi >> (8 * x)) & 0xff;
                }
            }
            if (intra) {
#pragma unroll
                for (int x = 0; x < 8; x++) pred[x] = 0;
            } else if (!filter) {
#pragma unroll
                for (int x = 0; x < 8; x++) pred[x] = a8[1][x];
            } else {
                /* H.261 loop filter, separable [1 2 1]: 1-D at the block
                 * edges, corners unfiltered (mbgpu:79-150) */
                int h[3][8];
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    h[k][0] = 4 * a8[k][0];
#pragma unroll
                    for (int x = 1; x < 7; x++) h[k][x] = a8[k][x - 1] + 2 * a8[k][x] + a8[k][x + 1];
                    h[k][7] = 4 * a8[k][7];
                }
                if (rc == 0 || rc == 7) {
#pragma unroll
                    for (int x = 0; x < 8; x++) pred[x] = (h[1][x] + 2) >> 2;
                } else {
#pragma unroll
                    for (int x = 0; x < 8; x++) pred[x] = (h[0][x] + 2 * h[1][x] + h[2][x] + 8) >> 4;
                }
            }
#pragma unroll
            for (int x = 0; x < 8; x++) res[x] = cpel[x] - pred[x];

            /* forward DCT, row pass */
            float o[8];
            fdct_core((float)(res[0] + res[7]), (float)(res[1] + res[6]), (float)(res[2] + res[5]),
                      (float)(res[3] + res[4]), (float)(res[3] - res[4]), (float)(res[2] - res[5]),
                      (float)(res[1] - res[6]), (float)(res[0] - res[7]), o);
#pragma unroll
            for (int k = 0; k < 8; k++) tf[b * kBlkPitch + rc * 8 + k] = o[k];
        }
        __syncthreads();

        /* ---- phase 2: lane (b, c) ---- */
        int coef[8];
        int lvl[8];
        int blk_accum = 0;
        if (act) {
            const int c = rc;
            float v[8];
#pragma unroll
            for (int r = 0; r < 8; r++) v[r] = tf[b * kBlkPitch + r * 8 + c];
            const float cs = post_c;
            float o[8];
            fdct_core(v[0] + v[7], v[1] + v[6], v[2] + v[5], v[3] + v[4], v[3] - v[4], v[2] - v[5], v[1] - v[6],
                      v[0] - v[7], o);
#pragma unroll
            for (int k = 0; k < 8; k++)
                coef[k] = (int)floorf((float)((double)(cs * kFwdPost[k] * o[k]) + 0.49999));

            /* denoise with the previous frame's offsets, staged as UInt16
             * (mbgpu:525-528, 586-604) */
#pragma unroll
            for (int r = 0; r < 8; r++) {
                int i = c + 8 * r;
                const int off = (int)((offp[r >> 1] >> (16 * (r & 1))) & 0xffffu);
                int cc = coef[r];
                int a = cc < 0 ? -cc : cc;
                atomicAdd(&stat[t][i], a);
                a = a - off;
                a = a > 0 ? a : 0;
                coef[r] = cc < 0 ? -a : a;
            }

            /* forward quantiser (mbgpu:685-716) */
            const float inv = c_inv_quant[quant];
#pragma unroll
            for (int r = 0; r < 8; r++) {
                int i = c + 8 * r;
                int l;
                if (intra && i == 0) {
                    l = coef[r] / 8;
                    l = clampi(l, 1, 254);
                    if (l == 128) l = 255;
                    blk_accum += 100;
                } else {
                    float f = (float)coef[r] * inv;
                    l = clampi((int)f, -127, 127);
                    int al = l < 0 ? -l : l;
                    blk_accum += al > 1 ? 100 : al;
                }
                lvl[r] = l;
                lv[b * 64 + i] = l;
            }
        }
        /* sum the 8 columns of the block */
        blk_accum += __shfl_xor(blk_accum, 1, 64);
        blk_accum += __shfl_xor(blk_accum, 2, 64);
        blk_accum += __shfl_xor(blk_accum, 4, 64);
        const bool coded = act && blk_accum > 2;
        /* cbp of the MB: ballot of the lanes c == 0 of coded blocks */
        const uint64_t coded_mask = __ballot(coded && rc == 0);
        int cbp = 0;
#pragma unroll
        for (int k = 0; k < 6; k++)
            if ((coded_mask >> (8 * k)) & 1) cbp |= kCbpBit[k];

        if (act) {
            const int c = rc;
            /* dequantiser (mbgpu:763-814); an uncoded block is all zero */
            const int qadd = coded ? ((quant - 1) | 1) : 0, qmul = coded ? quant * 2 : 0;
            int deq[8];
#pragma unroll
            for (int r = 0; r < 8; r++) {
                int i = c + 8 * r;
                int l = lvl[r];
                if (intra && i == 0) {
                    deq[r] = l == 255 ? 1024 : l * 8;
                } else {
                    int x = l * qmul;
                    if (x > 0)
                        x += qadd;
                    else if (x < 0)
                        x -= qadd;
                    deq[r] = clampi(x, -2048, 2047);
                }
            }
            /* IDCT prescale and column pass (mbgpu:836-888) */
            float tv[8], o[8];
#pragma unroll
            for (int r = 0; r < 8; r++) {
                const float cs = (float)((double)kInvPre[r] * 0.125);
                tv[r] = ((float)deq[r] * pre_c) * cs;
            }
            idct_core(tv, o);
            __builtin_amdgcn_s_barrier(); /* all phase-2 reads of tf are done */
#pragma unroll
            for (int r = 0; r < 8; r++) tf[b * kBlkPitch + r * 8 + c] = o[r];
        } else {
            __builtin_amdgcn_s_barrier();
        }
        __syncthreads();

        /* ---- phase 3: lane (b, r) ---- */
        if (act) {
            float v[8], o[8];
#pragma unroll
            for (int k = 0; k < 8; k++) v[k] = tf[b * kBlkPitch + rc * 8 + k];
            idct_core(v, o);
            uint32_t w0 = 0, w1 = 0;
#pragma unroll
            for (int x = 0; x < 8; x++) {
                int rres = clampi((int)floor((double)o[x] + 0.49999), -256, 255);
                int val = clampi(pred[x] + rres, 0, 255);
                if (x < 4)
                    w0 |= (uint32_t)val << (8 * x);
                else
                    w1 |= (uint32_t)val << (8 * (x - 4));
            }
            *(uint2 *)(d.rec[plane] + (size_t)(py + rc) * stride + px) = make_uint2(w0, w1);
        }

        /* ---- bit count of the MB's CBP + TCOEFF layer ---- */
        int bits = 0;
        const bool written_any = intra || cbp;
        if (p.side && written_any) {
            /* lane k = transmission position k of each block in turn */
            for (int k = 0; k < 6; k++) {
                const bool blk_written = intra || (cbp & kCbpBit[k]);
                if (!blk_written) continue;
                int l = lv[k * 64 + zz];
                uint64_t nz = __ballot(l != 0);
                if (l != 0) {
                    uint64_t below = lane ? (nz & ((~0ull) >> (64 - lane))) : 0ull;
                    int prev = below ? 63 - __builtin_clzll(below) : -1;
                    int run = lane - prev - 1;
                    if (lane == 0) {
                        if (intra)
                            bits += 8; /* DC: 8-bit fixed length */
                        else if (l == 1 || l == -1)
                            bits += 2; /* first coefficient '1s' */
                        else
                            bits += tcoeff_bits(tlen, 0, l);
                    } else {
                        bits += tcoeff_bits(tlen, run, l);
                    }
                }
                if (lane == 0) bits += 2; /* EOB */
            }
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) bits += __shfl_xor(bits, off, 64);
            if (!intra) bits += c_cbp[cbp].len;
        }
        __syncthreads();

        /* ---- outputs ---- */
        const int flags = md.flags | (cbp ? H261_MB_COEFF : 0);
        if (p.levels) {
            /* 384 levels as 96 dwords of 4 bytes: AC levels are within
             * [-127, 127], an intra DC within [1, 255] (read back as uint8) */
            uint32_t *out = (uint32_t *)(p.levels + (size_t)g * 384);
#pragma unroll
            for (int j = 0; j < 2; j++) {
                const int w = lane + 64 * j;
                if (w < 96) {
                    const int4 v = *(const int4 *)&lv[4 * w];
                    out[w] = (uint32_t)(v.x & 0xff) | ((uint32_t)(v.y & 0xff) << 8) | ((uint32_t)(v.z & 0xff) << 16) |
                             ((uint32_t)(v.w & 0xff) << 24);
                }
            }
        }
        if (p.mb_struct) {
            h261_macroblock_t *o = p.mb_struct + g;
#pragma unroll
            for (int k = 0; k < 6; k++) o->rgi_tcoeff[k][lane] = lv[k * 64 + lane];
            if (lane == 0) {
                o->i_macroblock_address = 0;
                o->i_macroblock_type = md.type;
                o->i_macroblock_type_flags = flags;
                o->i_macroblock_quant = quant;
                o->rgi_mv[0] = md.mvx;
                o->rgi_mv[1] = md.mvy;
                o->i_coded_block_pattern = cbp;
                o->i_mb_x = bx;
                o->i_mb_y = by;
            }
        }
        if (p.side && lane == 0) {
            MbSide sd;
            sd.type = (int8_t)md.type;
            sd.flags = (int8_t)flags;
            sd.quant = (int8_t)quant;
            sd.cbp = (int8_t)cbp;
            sd.mvx = (int8_t)md.mvx;
            sd.mvy = (int8_t)md.mvy;
            sd.bits = (int16_t)bits;
            p.side[g] = sd;
        }
        if (lane == 0) atomicAdd(&scount[t], 6);
        __syncthreads();
        /* the segment of stream s in this wave ends here when the wave's next
         * MB belongs to another stream (or there is none) */
        const int gn = g + 1;
        const bool last = it + 1 == p.mbs_per_wave || gn >= p.n_mb || p.mb_stream[gn] != s;
        if (last) flush_segment(p, s, stat, scount, lane);
    }
    (void)sblk;
}

}  // namespace

int launch_mb(const MbLaunch &p, void *stream, void *ev_start, void *ev_stop)
{
    if (p.n_mb <= 0) {
        if (ev_start) { /* nothing to launch: an empty timed interval */
            (void)hipEventRecord((hipEvent_t)ev_start, (hipStream_t)stream);
            (void)hipEventRecord((hipEvent_t)ev_stop, (hipStream_t)stream);
        }
        return 0;
    }
    const int per = p.mbs_per_wave > 0 ? p.mbs_per_wave : 1;
    MbLaunch q = p;
    q.mbs_per_wave = per;
    const int grid = launch_grid((p.n_mb + per - 1) / per, p.xcd_remap);
    if (ev_start)
        hipExtLaunchKernelGGL(mb_kernel, dim3(grid), dim3(64), 0, (hipStream_t)stream, (hipEvent_t)ev_start,
                              (hipEvent_t)ev_stop, 0, q);
    else
        hipLaunchKernelGGL(mb_kernel, dim3(grid), dim3(64), 0, (hipStream_t)stream, q);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace h261g
