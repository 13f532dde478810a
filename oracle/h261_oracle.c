/*
 * h261_oracle.c -- CPU restatement of the reference H.261 encoder core.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product): see
 * h261_oracle.h.  Built with -O2 -ffp-contract=off by oracle/Makefile.
 *
 * Each function follows one part of the reference and cites it.  Where the
 * reference has data races (SATD reductions src/megpu_kernels.cuh:681-696,
 * CBP reset vs atomicOr src/mbgpu_kernels.cuh:676-720) this file states
 * the intended result: the plain sum of the 16 sub-block SATDs and the OR
 * of this frame's coded blocks.
 */
#include "h261_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define MB 16

static int imin(int a, int b) { return a < b ? a : b; }
static int imax(int a, int b) { return a > b ? a : b; }

/* ------------------------------------------------------------------ */
/* motion estimation                                                   */
/* ------------------------------------------------------------------ */

/* src/gpu_device.cu:226-265 */
void ora_mv_limits(int mb_w, int mb_h, int8_t *limits)
{
    int gmax_x = (mb_w - 1) * 16, gmax_y = (mb_h - 1) * 16;
    for (int y = 0; y < mb_h; y++)
        for (int x = 0; x < mb_w; x++) {
            int bx = x * 16, by = y * 16;
            int8_t *l = limits + 4 * (x + y * mb_w);
            l[0] = (int8_t)(imax(bx - 15, 0) - bx);
            l[1] = (int8_t)(imin(bx + 15, gmax_x) - bx);
            l[2] = (int8_t)(imax(by - 15, 0) - by);
            l[3] = (int8_t)(imin(by + 15, gmax_y) - by);
        }
}

/* src/megpu_kernels.cuh:3-21 */
int ora_sad16(const uint8_t *cur, const uint8_t *ref, int stride, int cx, int cy, int rx, int ry)
{
    int sad = 0;
    for (int y = 0; y < 16; y++)
        for (int x = 0; x < 16; x++)
            sad += abs((int)cur[(cy + y) * stride + cx + x] - (int)ref[(ry + y) * stride + rx + x]);
    return sad;
}

/* src/megpu_kernels.cuh:95-154 (coarse) and 375-434 (refine): every thread
 * of a dim x dim block compares itself with its +x, +y and +xy neighbours
 * at distance stride (strict <, in that order) and the block synchronises
 * between the read and the write phase, so each stride is a function of
 * the previous stride's array.  Returns the winning candidate index. */
int ora_select_quadtree(const uint16_t *cost, int dim)
{
    uint32_t c[121], c2[121];
    int idx[121], idx2[121];
    int n = dim * dim;
    for (int t = 0; t < n; t++) {
        c[t] = cost[t];
        idx[t] = t;
    }
    for (int s = 1; s < dim; s *= 2) {
        for (int ty = 0; ty < dim; ty++)
            for (int tx = 0; tx < dim; tx++) {
                int t = tx + ty * dim;
                uint32_t best = c[t];
                int bi = idx[t];
                const int ox[3] = {s, 0, s}, oy[3] = {0, s, s};
                for (int k = 0; k < 3; k++) {
                    int x = tx + ox[k], y = ty + oy[k];
                    if (x < dim && y < dim) {
                        int u = x + y * dim;
                        if (c[u] < best) {
                            best = c[u];
                            bi = idx[u];
                        }
                    }
                }
                c2[t] = best;
                idx2[t] = bi;
            }
        memcpy(c, c2, sizeof(uint32_t) * n);
        memcpy(idx, idx2, sizeof(int) * n);
    }
    return idx[0];
}

/* src/megpu_kernels.cuh:467-500 */
int ora_satd4x4(const int16_t d[16])
{
    int t[4][4], satd = 0;
    for (int i = 0; i < 4; i++) {
        const int16_t *p = d + 4 * i;
        int a1 = p[0] + p[1], a2 = p[0] - p[1], a3 = p[2] + p[3], a4 = p[2] - p[3];
        t[i][0] = a1 + a3;
        t[i][1] = a2 + a4;
        t[i][2] = a1 - a3;
        t[i][3] = a2 - a4;
    }
    for (int i = 0; i < 4; i++) {
        int a1 = t[0][i] + t[1][i], a2 = t[0][i] - t[1][i];
        int a3 = t[2][i] + t[3][i], a4 = t[2][i] - t[3][i];
        satd += abs(a1 + a3) + abs(a2 + a4) + abs(a1 - a3) + abs(a2 - a4);
    }
    return satd;
}

static const int8_t k_coarse[11] = {-14, -11, -8, -5, -2, 0, 2, 5, 8, 11, 14}; /* megpu:29-32 */
static const int8_t k_refine[7] = {-3, -2, -1, 0, 1, 2, 3};                    /* megpu:309-312 */

static long g_last_sad_candidates;

static int mv_in_limits(int x, int y, const int8_t *l)
{
    /* megpu:67-83 */
    return !(x < l[0] || x > l[1] || y < l[2] || y > l[3]);
}

/* cost of one sub-block for the three SATD kernels.
 * kind 0: inter (megpu:617-721), 1: inter+filter (megpu:724-905, live
 * branch 833-899), 2: intra (megpu:907-987). */
static int sub_block_satd(const uint8_t *cur, const uint8_t *ref, int stride, int sbx, int sby,
                          int sub_x, int sub_y, int mvx, int mvy, int kind)
{
    int16_t d[16];
    int rx = sbx + mvx, ry = sby + mvy;
    for (int y = 0; y < 4; y++)
        for (int x = 0; x < 4; x++) {
            int c = cur[(sby + y) * stride + sbx + x];
            if (kind == 2)
                d[x + 4 * y] = (int16_t)c;
            else
                d[x + 4 * y] = (int16_t)(c - ref[(ry + y) * stride + rx + x]);
        }
    if (kind == 1) {
        /* interior 6x6 of each 8x8 block: odd sub-block index covers
         * pels 0..2, even covers 1..3 (megpu:833-853) */
        int fx0 = (sub_x & 1) ? 0 : 1, fx1 = (sub_x & 1) ? 3 : 4;
        int fy0 = (sub_y & 1) ? 0 : 1, fy1 = (sub_y & 1) ? 3 : 4;
        for (int y = fy0; y < fy1; y++)
            for (int x = fx0; x < fx1; x++) {
                const uint8_t *r = ref + (ry + y) * stride + rx + x;
                int a1 = r[-stride - 1], a2 = r[-stride], a3 = r[-stride + 1];
                int a4 = r[-1], a5 = r[0], a6 = r[1];
                int a7 = r[stride - 1], a8 = r[stride], a9 = r[stride + 1];
                int p = (a1 + a2 * 2 + a3 + a4 * 2 + a5 * 4 + a6 * 2 + a7 + a8 * 2 + a9 + 8) / 16;
                d[x + 4 * y] = (int16_t)(cur[(sby + y) * stride + sbx + x] - p);
            }
    }
    return ora_satd4x4(d);
}

/* src/gpu_device.cu:356-432: coarse (megpu:27-166), refine (megpu:307-446),
 * then the inter / inter+filter / intra SATD costs. */
void ora_me_frame(const uint8_t *cur, const uint8_t *ref, int width, int height, me_gpu_mb_t *out)
{
    int mb_w = width / 16, mb_h = height / 16;
    int8_t *lim = (int8_t *)malloc(4 * mb_w * mb_h);
    long ncand = 0;
    ora_mv_limits(mb_w, mb_h, lim);
    for (int my = 0; my < mb_h; my++)
        for (int mx = 0; mx < mb_w; mx++) {
            int m = mx + my * mb_w, bx = mx * 16, by = my * 16;
            const int8_t *l = lim + 4 * m;
            uint16_t cost[121];
            for (int ty = 0; ty < 11; ty++)
                for (int tx = 0; tx < 11; tx++) {
                    int vx = k_coarse[tx], vy = k_coarse[ty];
                    if (mv_in_limits(vx, vy, l)) {
                        cost[tx + ty * 11] = (uint16_t)ora_sad16(cur, ref, width, bx, by, bx + vx, by + vy);
                        ncand++;
                    } else
                        cost[tx + ty * 11] = 65535;
                }
            int w = ora_select_quadtree(cost, 11);
            int sx = k_coarse[w % 11], sy = k_coarse[w / 11];
            uint16_t rcost[49];
            for (int ty = 0; ty < 7; ty++)
                for (int tx = 0; tx < 7; tx++) {
                    int vx = sx + k_refine[tx], vy = sy + k_refine[ty];
                    if (mv_in_limits(vx, vy, l)) {
                        rcost[tx + ty * 7] = (uint16_t)ora_sad16(cur, ref, width, bx, by, bx + vx, by + vy);
                        ncand++;
                    } else
                        rcost[tx + ty * 7] = 65535;
                }
            w = ora_select_quadtree(rcost, 7);
            int mvx = sx + k_refine[w % 7], mvy = sy + k_refine[w / 7];

            int cost_inter = 0, cost_filter = 0, cost_intra = 0;
            for (int sub_y = 0; sub_y < 4; sub_y++)
                for (int sub_x = 0; sub_x < 4; sub_x++) {
                    int sbx = bx + 4 * sub_x, sby = by + 4 * sub_y;
                    cost_inter += sub_block_satd(cur, ref, width, sbx, sby, sub_x, sub_y, mvx, mvy, 0);
                    cost_filter += sub_block_satd(cur, ref, width, sbx, sby, sub_x, sub_y, mvx, mvy, 1);
                    cost_intra += sub_block_satd(cur, ref, width, sbx, sby, sub_x, sub_y, 0, 0, 2);
                }
            me_gpu_mb_t *o = out + m;
            o->i_intra_cost = cost_intra;
            o->i_zero_cost = 0; /* never written by the reference (h261_types.h:53) */
            o->i_16x16_cost = cost_inter;
            o->i_16x16_filter_cost = cost_filter;
            o->s_16x16_mv.i8_mx = (Int8)mvx;
            o->s_16x16_mv.i8_my = (Int8)mvy;
        }
    g_last_sad_candidates = ncand;
    free(lim);
}

/* ------------------------------------------------------------------ */
/* DCT / IDCT / quantiser                                              */
/* ------------------------------------------------------------------ */

/* float tables initialised from double literals (mbgpu:335, 823): each
 * element is the double nearest the decimal, rounded to float */
static const float k_fwd_post[8] = {1.0, 0.720960, 0.765367, 0.850430, 1.0, 1.272759, 1.847759, 3.624510};
static const float k_inv_pre[8] = {1.0, 1.387040, 1.306563, 1.175876, 1.0, 0.785695, 0.541196, 0.275899};

/* src/mbgpu_kernels.cuh:337-442.  Every product with an unsuffixed literal
 * is evaluated in double and rounded to float on assignment; the final
 * rounding is floorf() of (float)(prod + 0.49999).  Parity unpinned (op
 * typing; see h261_oracle.h). */
void ora_fdct8x8(const int32_t in[64], int32_t out[64])
{
    float t[64];
    for (int r = 0; r < 8; r++) {
        const int32_t *p = in + 8 * r;
        float tmp0 = (float)(p[0] + p[7]), tmp1 = (float)(p[1] + p[6]);
        float tmp2 = (float)(p[2] + p[5]), tmp3 = (float)(p[3] + p[4]);
        float tmp4 = (float)(p[3] - p[4]), tmp5 = (float)(p[2] - p[5]);
        float tmp6 = (float)(p[1] - p[6]), tmp7 = (float)(p[0] - p[7]);
        float tmp10 = tmp0 + tmp3, tmp13 = tmp0 - tmp3, tmp11 = tmp1 + tmp2, tmp12 = tmp1 - tmp2;
        float *o = t + 8 * r;
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
    for (int c = 0; c < 8; c++) {
        const float *p = t + c;
        float cs = k_fwd_post[c];
        cs = (float)((double)cs * 0.125);
        float tmp0 = p[0] + p[56], tmp1 = p[8] + p[48], tmp2 = p[16] + p[40], tmp3 = p[24] + p[32];
        float tmp4 = p[24] - p[32], tmp5 = p[16] - p[40], tmp6 = p[8] - p[48], tmp7 = p[0] - p[56];
        float tmp10 = tmp0 + tmp3, tmp13 = tmp0 - tmp3, tmp11 = tmp1 + tmp2, tmp12 = tmp1 - tmp2;
#define FQ(k, v) out[8 * (k) + c] = (int32_t)floorf((float)((double)(cs * k_fwd_post[k] * (v)) + 0.49999))
        FQ(0, tmp10 + tmp11);
        FQ(4, tmp10 - tmp11);
        float z1 = tmp12 + tmp13;
        z1 = (float)((double)z1 * 0.707107);
        FQ(2, tmp13 + z1);
        FQ(6, tmp13 - z1);
        tmp4 += tmp5;
        tmp5 += tmp6;
        tmp6 += tmp7;
        float z5 = (float)((double)(tmp4 - tmp6) * 0.382683);
        float z2 = (float)((double)tmp4 * 0.541196 + (double)z5);
        float z4 = (float)((double)tmp6 * 1.306563 + (double)z5);
        float z3 = (float)((double)tmp5 * 0.707107);
        float z11 = tmp7 + z3, z13 = tmp7 - z3;
        FQ(5, z13 + z2);
        FQ(3, z13 - z2);
        FQ(1, z11 + z4);
        FQ(7, z11 - z4);
#undef FQ
    }
}

/* src/mbgpu_kernels.cuh:827-931: prescale, column pass, row pass,
 * (int)floor(double) rounding, clamp to [-256, 255].  Parity unpinned (op
 * typing; see h261_oracle.h). */
void ora_idct8x8(const int32_t in[64], int32_t out[64])
{
    float t[64];
    for (int r = 0; r < 8; r++) {
        float cs = (float)((double)k_inv_pre[r] * 0.125);
        for (int k = 0; k < 8; k++)
            t[8 * r + k] = ((float)in[8 * r + k]) * k_inv_pre[k] * cs;
    }
    for (int i = 0; i < 8; i++) {
        float tmp12 = t[16 + i] - t[48 + i];
        float tmp13 = t[16 + i] + t[48 + i];
        tmp12 = (float)((double)tmp12 * 1.414214);
        tmp12 -= tmp13;
        float tmp10 = t[0 + i] + t[32 + i];
        float tmp11 = t[0 + i] - t[32 + i];
        float tmp0 = tmp10 + tmp13, tmp1 = tmp11 + tmp12, tmp2 = tmp11 - tmp12, tmp3 = tmp10 - tmp13;
        float z11 = t[8 + i] + t[56 + i], z12 = t[8 + i] - t[56 + i];
        float z13 = t[40 + i] + t[24 + i], z10 = t[40 + i] - t[24 + i];
        tmp13 = z11 + z13;
        tmp11 = (float)((double)(z11 - z13) * 1.414214);
        float z5 = (float)((double)(z12 + z10) * 1.84776);
        tmp10 = (float)((double)z12 * 1.082392 - (double)z5);
        tmp12 = (float)((double)z10 * -2.613126 + (double)z5);
        tmp12 -= tmp13;
        tmp11 -= tmp12;
        tmp10 += tmp11;
        t[0 + i] = tmp0 + tmp13;
        t[8 + i] = tmp1 + tmp12;
        t[16 + i] = tmp2 + tmp11;
        t[24 + i] = tmp3 - tmp10;
        t[32 + i] = tmp3 + tmp10;
        t[40 + i] = tmp2 - tmp11;
        t[48 + i] = tmp1 - tmp12;
        t[56 + i] = tmp0 - tmp13;
    }
    for (int r = 0; r < 8; r++) {
        const float *p = t + 8 * r;
        float tmp13 = p[2] + p[6];
        float tmp12 = p[2] - p[6];
        tmp12 = (float)((double)tmp12 * 1.414214);
        tmp12 -= tmp13;
        float tmp10 = p[0] + p[4], tmp11 = p[0] - p[4];
        float tmp0 = tmp10 + tmp13, tmp1 = tmp11 + tmp12, tmp2 = tmp11 - tmp12, tmp3 = tmp10 - tmp13;
        float z11 = p[1] + p[7], z12 = p[1] - p[7], z13 = p[5] + p[3], z10 = p[5] - p[3];
        tmp13 = z11 + z13;
        tmp11 = (float)((double)(z11 - z13) * 1.414214);
        float z5 = (float)((double)(z12 + z10) * 1.84776);
        tmp10 = (float)((double)z12 * 1.082392 - (double)z5);
        tmp12 = (float)((double)z10 * -2.613126 + (double)z5);
        tmp12 -= tmp13;
        tmp11 -= tmp12;
        tmp10 += tmp11;
#define IQ(k, v) out[8 * r + (k)] = imin(255, imax(-256, (int32_t)floor((double)(v) + 0.49999)))
        IQ(0, tmp0 + tmp13);
        IQ(1, tmp1 + tmp12);
        IQ(2, tmp2 + tmp11);
        IQ(3, tmp3 - tmp10);
        IQ(4, tmp3 + tmp10);
        IQ(5, tmp2 - tmp11);
        IQ(6, tmp1 - tmp12);
        IQ(7, tmp0 - tmp13);
#undef IQ
    }
}

/* src/mbgpu_kernels.cuh:685-722 (one block).  The reciprocal is the double
 * quotient 1.0/(float)(2q) rounded to float; the level is the float product
 * truncated toward zero and clamped to [-127, 127]. */
int ora_quant_block(int32_t c[64], int quant, int intra)
{
    float inv = (float)(1.0 / (double)(float)(quant * 2));
    int accum = 0, start = 0;
    if (intra) {
        int dc = c[0] / 8;
        dc = imax(dc, 1);
        dc = imin(dc, 254);
        if (dc == 128) dc = 255;
        c[0] = dc;
        accum = 100;
        start = 1;
    }
    for (int i = start; i < 64; i++) {
        float f = (float)c[i] * inv;
        int l = imin(127, imax(-127, (int32_t)f));
        c[i] = l;
        l = abs(l);
        accum += l > 1 ? 100 : l;
    }
    return accum > 2;
}

/* src/mbgpu_kernels.cuh:763-814 */
void ora_dequant_block(int32_t c[64], int quant, int intra, int coded)
{
    int add = (quant - 1) | 1, mul = quant * 2, start = 0;
    if (!coded) add = mul = 0;
    if (intra) {
        c[0] = c[0] == 255 ? 1024 : c[0] * 8;
        start = 1;
    }
    for (int i = start; i < 64; i++) {
        int l = c[i] * mul;
        if (l > 0)
            l += add;
        else if (l < 0)
            l -= add;
        c[i] = imin(2047, imax(-2048, l));
    }
}

/* src/mbgpu_kernels.cuh:75-166 (setup) and 1006-1096 (reconstruction):
 * plain copy, or the H.261 loop filter -- separable [1 2 1], interior
 * (sum+8)>>4, block edges 1-D (sum+2)>>2, corners unfiltered.  Both of the
 * reference's pass orders give the same integers. */
void ora_predict8x8(const uint8_t *ref, int stride, int px, int py, int filter, int32_t pred[64])
{
    if (!filter) {
        for (int y = 0; y < 8; y++)
            for (int x = 0; x < 8; x++) pred[x + 8 * y] = ref[(py + y) * stride + px + x];
        return;
    }
    int h[64];
    for (int y = 0; y < 8; y++) {
        const uint8_t *r = ref + (py + y) * stride + px;
        h[8 * y + 0] = r[0] * 4;
        for (int x = 1; x < 7; x++) h[8 * y + x] = r[x - 1] + 2 * r[x] + r[x + 1];
        h[8 * y + 7] = r[7] * 4;
    }
    for (int x = 0; x < 8; x++) {
        pred[x] = (h[x] + 2) >> 2;
        for (int y = 1; y < 7; y++) pred[x + 8 * y] = (h[x + 8 * (y - 1)] + 2 * h[x + 8 * y] + h[x + 8 * (y + 1)] + 8) >> 4;
        pred[x + 56] = (h[x + 56] + 2) >> 2;
    }
}

/* ------------------------------------------------------------------ */
/* MB coder: encode + decode                                           */
/* ------------------------------------------------------------------ */

static const int k_cbp_bit[6] = {32, 16, 8, 4, 2, 1}; /* h261_types.h:33-38 */

/* src/gpu_device.cu:592-729 (encode) and 732-815 (decode), kernels
 * src/mbgpu_kernels.cuh:172-1277. */
void ora_encode_macroblocks(const uint8_t *cur[3], const uint8_t *ref[3], int width, int height,
                            const h261_mb_mode_decision_t *md, int pic_quant,
                            h261_dct_denoise_t *dn, h261_macroblock_t *mbs, uint8_t *recon[3])
{
    int mb_w = width / 16, mb_h = height / 16, n_mb = mb_w * mb_h;
    int32_t (*coef)[6][64] = (int32_t (*)[6][64])malloc(sizeof(int32_t) * 384 * n_mb);
    int32_t (*pred)[6][64] = (int32_t (*)[6][64])malloc(sizeof(int32_t) * 384 * n_mb);
    uint16_t off[2][64];
    long noise[2][64];
    long count[2] = {0, 0};
    /* offsets staged as UInt16 (mbgpu:513, 525-528); index 0 intra, 1 inter */
    for (int i = 0; i < 64; i++) {
        off[0][i] = (uint16_t)dn->rgi_intra_signal_offset[i];
        off[1][i] = (uint16_t)dn->rgi_inter_signal_offset[i];
        noise[0][i] = noise[1][i] = 0;
    }
    for (int m = 0; m < n_mb; m++) {
        const h261_mb_mode_decision_t *d = md + m;
        h261_macroblock_t *o = mbs + m;
        int quant = imin(31, imax(1, pic_quant + d->i_quantiser_adjust)); /* gpu_device.cu:625-627 */
        int mbx = (m % mb_w) * 16, mby = (m / mb_w) * 16;
        int intra = d->i_mb_type == H261_MB_TYPE_INTRA;
        /* setup_luma copies MB fields (mbgpu:205-212) */
        o->i_macroblock_address = 0;
        o->i_macroblock_type = d->i_mb_type;
        o->i_macroblock_type_flags = d->i_mb_flags;
        o->i_macroblock_quant = quant;
        o->rgi_mv[0] = d->i_mv_x;
        o->rgi_mv[1] = d->i_mv_y;
        o->i_coded_block_pattern = 0;
        o->i_mb_x = mbx;
        o->i_mb_y = mby;
        int filter = (d->i_mb_flags & H261_MB_FILTER) != 0;
        for (int b = 0; b < 6; b++) {
            int plane = b < 4 ? 0 : (b == 4 ? 1 : 2);
            int stride = plane ? width / 2 : width;
            int px, py, mvx, mvy;
            if (b < 4) {
                px = mbx + (b & 1) * 8;
                py = mby + ((b >> 1) & 1) * 8;
                mvx = d->i_mv_x;
                mvy = d->i_mv_y;
            } else { /* chroma: mv / 2 truncated toward zero (mbgpu:266-269) */
                px = mbx / 2;
                py = mby / 2;
                mvx = d->i_mv_x / 2;
                mvy = d->i_mv_y / 2;
            }
            int32_t blk[64];
            if (intra) {
                for (int i = 0; i < 64; i++) {
                    blk[i] = cur[plane][(py + i / 8) * stride + px + i % 8];
                    pred[m][b][i] = 0;
                }
            } else {
                ora_predict8x8(ref[plane], stride, px + mvx, py + mvy, filter, pred[m][b]);
                for (int i = 0; i < 64; i++)
                    blk[i] = cur[plane][(py + i / 8) * stride + px + i % 8] - pred[m][b][i];
            }
            ora_fdct8x8(blk, coef[m][b]);
            /* denoise (mbgpu:506-636): stats use |c| before the subtraction */
            int t = intra ? 0 : 1;
            count[t]++;
            for (int i = 0; i < 64; i++) {
                int c = coef[m][b][i], a = abs(c);
                noise[t][i] += a;
                a = imax(a - (int)off[t][i], 0);
                coef[m][b][i] = c < 0 ? -a : a;
            }
        }
    }
    /* macroblocks_update_dct_denoise (mbgpu:478-503) runs after the frame's
     * denoise pass: the new offsets apply to the next frame */
    for (int i = 0; i < 64; i++) {
        int ni = (int)noise[1][i], na = (int)noise[0][i];
        dn->rgi_inter_signal_offset[i] = ((int)count[1] * dn->i_denoise + ni / 2) / (ni + 1);
        dn->rgi_intra_signal_offset[i] = ((int)count[0] * dn->i_denoise + na / 2) / (na + 1);
        dn->rgi_inter_noise[i] = 0;
        dn->rgi_intra_noise[i] = 0;
    }
    dn->i_inter_count = 0;
    dn->i_intra_count = 0;

    for (int m = 0; m < n_mb; m++) {
        h261_macroblock_t *o = mbs + m;
        int intra = o->i_macroblock_type == H261_MB_TYPE_INTRA;
        int quant = o->i_macroblock_quant;
        for (int b = 0; b < 6; b++) {
            int32_t *c = coef[m][b];
            int coded = ora_quant_block(c, quant, intra);
            memcpy(o->rgi_tcoeff[b], c, sizeof(int32_t) * 64);
            if (coded) {
                o->i_coded_block_pattern |= k_cbp_bit[b];
                o->i_macroblock_type_flags |= H261_MB_COEFF;
            }
        }
    }
    /* decode: dequant, IDCT of coded blocks, reconstruction */
    for (int m = 0; m < n_mb; m++) {
        const h261_macroblock_t *o = mbs + m;
        int intra = o->i_macroblock_type == H261_MB_TYPE_INTRA;
        int mbx = o->i_mb_x, mby = o->i_mb_y;
        for (int b = 0; b < 6; b++) {
            int32_t c[64], res[64];
            int coded = (o->i_coded_block_pattern & k_cbp_bit[b]) != 0;
            memcpy(c, o->rgi_tcoeff[b], sizeof(c));
            ora_dequant_block(c, o->i_macroblock_quant, intra, coded);
            if (coded)
                ora_idct8x8(c, res);
            else
                memcpy(res, c, sizeof(res)); /* all zero after dequant */
            int plane = b < 4 ? 0 : (b == 4 ? 1 : 2);
            int stride = plane ? width / 2 : width;
            int px = b < 4 ? mbx + (b & 1) * 8 : mbx / 2;
            int py = b < 4 ? mby + ((b >> 1) & 1) * 8 : mby / 2;
            for (int i = 0; i < 64; i++) {
                int v = intra ? res[i] : pred[m][b][i] + res[i];
                recon[plane][(py + i / 8) * stride + px + i % 8] = (uint8_t)imin(255, imax(0, v));
            }
        }
    }
    (void)mb_h;
    free(coef);
    free(pred);
}

/* ------------------------------------------------------------------ */
/* mode decision and rate control                                      */
/* ------------------------------------------------------------------ */

/* src/h261_moded.c:31-100 (quant adjust stays 0: its code is commented out) */
void ora_mode_decision(const me_gpu_mb_t *me, int num_mb, int frame_type, h261_mb_mode_decision_t *md)
{
    for (int i = 0; i < num_mb; i++) {
        const me_gpu_mb_t *g = me + i;
        h261_mb_mode_decision_t *d = md + i;
        int intra = frame_type == H261_FRAME_TYPE_INTRA ||
                    (g->i_intra_cost <= g->i_16x16_cost && g->i_intra_cost <= g->i_16x16_filter_cost);
        d->i_quantiser_adjust = 0;
        if (intra) {
            d->i_mb_type = H261_MB_TYPE_INTRA;
            d->i_cost = g->i_intra_cost;
            d->i_mb_flags = 0;
            d->i_mv_x = d->i_mv_y = 0;
        } else {
            d->i_mb_type = H261_MB_TYPE_INTER;
            if (g->i_16x16_filter_cost < g->i_16x16_cost) {
                d->i_mb_flags = H261_MB_MC | H261_MB_FILTER;
                d->i_cost = g->i_16x16_filter_cost;
            } else {
                d->i_mb_flags = 0;
                d->i_cost = g->i_16x16_cost;
            }
            d->i_mv_x = g->s_16x16_mv.i8_mx;
            d->i_mv_y = g->s_16x16_mv.i8_my;
            if (d->i_mv_x != 0 || d->i_mv_y != 0) d->i_mb_flags |= H261_MB_MC;
        }
    }
}

#define RC_BITS 6
#define RC_ONE (1 << RC_BITS)
#define RC_HALF (1 << (RC_BITS - 1))
#define RC_TIME_SCALE 1918 /* 29.97 in 6-bit fixed point (ratectrl.c:7-8) */

/* pow(2, f/5) in 6-bit fixed point, f = -31..31 (ratectrl.c:11-19) */
static const int k_qscale[63] = {
    0,    1,    1,    1,    1,    1,    2,    2,    2,    3,    3,    4,    4,    5,    6,    6,
    8,    9,    10,   12,   13,   16,   18,   21,   24,   27,   32,   36,   42,   48,   55,   64,
    73,   84,   97,   111,  128,  147,  168,  194,  222,  256,  294,  337,  388,  445,  512,  588,
    675,  776,  891,  1024, 1176, 1351, 1552, 1782, 2048, 2352, 2702, 3104, 3565, 4096, 4705};

static int rc_mul(int a, int b) { return (int)(((int64_t)a * (int64_t)b) >> RC_BITS); }
static int rc_div(int a, int b) { return (int)((((int64_t)a) << RC_BITS) / b); }
static int rc_qscale_for_quant(int q) { return k_qscale[imin(imax(q, 0), 31) + 31]; }

/* ratectrl.c:63-104 */
static int rc_quant_for_qscale(int qs)
{
    int pos = 31, step = 16, lo, hi;
    while (step) {
        if (k_qscale[pos] < qs)
            pos += step;
        else
            pos -= step;
        step >>= 1;
    }
    if (k_qscale[pos] < qs) {
        lo = pos;
        hi = imin(pos + 1, 62);
    } else {
        lo = imax(0, pos - 1);
        hi = pos;
    }
    pos = (qs - k_qscale[lo] < k_qscale[hi] - qs) ? lo : hi;
    return pos - 31;
}

/* ratectrl.c:107-127 */
void ora_rc_init(h261_ratectrl_t *rc, int quant, int bitrate, int vbv_size, int mode)
{
    memset(rc, 0, sizeof(*rc));
    rc->i_ratectrl_mode = mode;
    rc->i_current_quant = quant;
    for (int i = 0; i < 2; i++) {
        rc->i_prediction_cost[i] = 100;
        rc->i_prediction_size[i] = 1;
    }
    rc->i_bitrate = bitrate / 8;
    rc->i_vbv_size = vbv_size / 8;
    rc->i_vbv_fill = rc->i_vbv_size;
}

/* ratectrl.c:225-245 */
int ora_rc_picture_cost(const h261_mb_mode_decision_t *md, int num_mb)
{
    int c = 0;
    for (int i = 0; i < num_mb; i++) c += md[i].i_cost;
    return imax(c, num_mb * 100);
}

/* ratectrl.c:130-195 */
int ora_rc_init_picture(h261_ratectrl_t *rc, int ft)
{
    rc->i_vbv_fill += rc_div(rc->i_bitrate, RC_TIME_SCALE);
    rc->i_vbv_fill = imin(rc->i_vbv_fill, rc->i_vbv_size);
    int cost_scale = rc_div(rc->i_current_picture_cost, rc->i_prediction_cost[ft]);
    int pred = rc_mul(cost_scale, rc->i_prediction_size[ft]);
    pred = rc_div(pred, rc_qscale_for_quant(rc->i_current_quant));
    int target = rc_div(rc->i_bitrate, RC_TIME_SCALE);
    if (ft == H261_FRAME_TYPE_INTRA) target *= 20;
    int up = rc_div(rc->i_vbv_fill, rc->i_vbv_size);
    int upper = target + rc_mul(target, up);
    int lower = target - rc_mul(target, RC_ONE - up);
    (void)upper; /* the reference overwrites the clamped target (ratectrl.c:177-178) */
    target = imax(lower, 10);
    int qss = rc_div(pred, target);
    if (qss < RC_ONE) qss = rc_mul(RC_ONE - up, RC_ONE) + rc_mul(up, qss);
    rc->i_current_quant += rc_quant_for_qscale(qss);
    rc->i_current_quant = imin(imax(rc->i_current_quant, 1), 31);
    int penalty = ((RC_ONE - up) * 4) >> RC_BITS;
    return rc->i_current_quant + penalty;
}

/* ratectrl.c:198-222 */
void ora_rc_update_picture(h261_ratectrl_t *rc, int ft, int quant, int bytes)
{
    rc->i_prediction_cost[ft] = rc_mul(rc->i_prediction_cost[ft], RC_HALF);
    rc->i_prediction_size[ft] = rc_mul(rc->i_prediction_size[ft], RC_HALF);
    rc->i_prediction_cost[ft] += rc->i_current_picture_cost;
    rc->i_prediction_size[ft] += rc_mul(bytes, rc_qscale_for_quant(quant));
    rc->i_vbv_fill -= bytes;
}

/* ------------------------------------------------------------------ */
/* bitstream writer                                                    */
/* ------------------------------------------------------------------ */

typedef struct {
    uint8_t *buf;
    int cap;
    long bitpos; /* bits written */
    int overflow;
} bw_t;

/* MSB-first, equivalent to src/h261_bitstream.c:28-61 on a zeroed buffer */
static void bw_put(bw_t *w, uint32_t code, int len)
{
    for (int i = len - 1; i >= 0; i--) {
        long byte = w->bitpos >> 3;
        if (byte >= w->cap) {
            w->overflow = 1;
            return;
        }
        if ((code >> i) & 1) w->buf[byte] |= (uint8_t)(0x80 >> (w->bitpos & 7));
        w->bitpos++;
    }
}

/* VLC tables of ITU-T H.261 as laid out in src/h261_tables.c:4-325 */
static const uint8_t k_mba[35][2] = {
    {1, 16}, {1, 1},   {3, 3},   {2, 3},   {3, 4},   {2, 4},   {3, 5},   {2, 5},   {7, 7},
    {6, 7},  {11, 8},  {10, 8},  {9, 8},   {8, 8},   {7, 8},   {6, 8},   {23, 10}, {22, 10},
    {21, 10}, {20, 10}, {19, 10}, {18, 10}, {35, 11}, {34, 11}, {33, 11}, {32, 11}, {31, 11},
    {30, 11}, {29, 11}, {28, 11}, {27, 11}, {26, 11}, {25, 11}, {24, 11}, {15, 11}};
static const uint8_t k_mtype[10][2] = {{1, 4}, {1, 7}, {1, 1}, {1, 5}, {1, 9},
                                       {1, 8}, {1, 10}, {1, 3}, {1, 2}, {1, 6}};
static const uint8_t k_cbp[64][2] = {
    {0, 0},  {11, 5}, {9, 5},  {13, 6}, {13, 4}, {23, 7}, {19, 7}, {31, 8}, {12, 4}, {22, 7},
    {18, 7}, {30, 8}, {19, 5}, {27, 8}, {23, 8}, {19, 8}, {11, 4}, {21, 7}, {17, 7}, {29, 8},
    {17, 5}, {25, 8}, {21, 8}, {17, 8}, {15, 6}, {15, 8}, {13, 8}, {3, 9},  {15, 5}, {11, 8},
    {7, 8},  {7, 9},  {10, 4}, {20, 7}, {16, 7}, {28, 8}, {14, 6}, {14, 8}, {12, 8}, {2, 9},
    {16, 5}, {24, 8}, {20, 8}, {16, 8}, {14, 5}, {10, 8}, {6, 8},  {6, 9},  {18, 5}, {26, 8},
    {22, 8}, {18, 8}, {13, 5}, {9, 8},  {5, 8},  {5, 9},  {12, 5}, {8, 8},  {4, 8},  {4, 9},
    {7, 3},  {10, 5}, {8, 5},  {12, 6}};
static const uint8_t k_mvd[32][2] = {
    {25, 11}, {27, 11}, {29, 11}, {31, 11}, {33, 11}, {35, 11}, {19, 10}, {21, 10},
    {23, 10}, {7, 8},   {9, 8},   {11, 8},  {7, 7},   {3, 5},   {3, 4},   {3, 3},
    {1, 1},   {2, 3},   {2, 4},   {2, 5},   {6, 7},   {10, 8},  {8, 8},   {6, 8},
    {22, 10}, {20, 10}, {18, 10}, {34, 11}, {32, 11}, {30, 11}, {28, 11}, {26, 11}};
/* (run, level, code, length); h261_tables.c:230-313 */
static const uint8_t k_tcoeff[][4] = {
    {0, 1, 3, 2},    {0, 2, 4, 4},    {0, 3, 5, 5},    {0, 4, 6, 7},    {0, 5, 38, 8},
    {0, 6, 33, 8},   {0, 7, 10, 10},  {0, 8, 29, 12},  {0, 9, 24, 12},  {0, 10, 19, 12},
    {0, 11, 16, 12}, {0, 12, 26, 13}, {0, 13, 25, 13}, {0, 14, 24, 13}, {0, 15, 23, 13},
    {1, 1, 3, 3},    {1, 2, 6, 6},    {1, 3, 37, 8},   {1, 4, 12, 10},  {1, 5, 27, 12},
    {1, 6, 22, 13},  {1, 7, 21, 13},  {2, 1, 5, 4},    {2, 2, 4, 7},    {2, 3, 11, 10},
    {2, 4, 20, 12},  {2, 5, 20, 13},  {3, 1, 7, 5},    {3, 2, 36, 8},   {3, 3, 28, 12},
    {3, 4, 19, 13},  {4, 1, 6, 5},    {4, 2, 15, 10},  {4, 3, 18, 12},  {5, 1, 7, 6},
    {5, 2, 9, 10},   {5, 3, 18, 13},  {6, 1, 5, 6},    {6, 2, 30, 12},  {7, 1, 4, 6},
    {7, 2, 21, 12},  {8, 1, 7, 7},    {8, 2, 17, 12},  {9, 1, 5, 7},    {9, 2, 17, 13},
    {10, 1, 39, 8},  {10, 2, 16, 13}, {11, 1, 35, 8},  {12, 1, 34, 8},  {13, 1, 32, 8},
    {14, 1, 14, 10}, {15, 1, 13, 10}, {16, 1, 8, 10},  {17, 1, 31, 12}, {18, 1, 26, 12},
    {19, 1, 25, 12}, {20, 1, 23, 12}, {21, 1, 22, 12}, {22, 1, 31, 13}, {23, 1, 30, 13},
    {24, 1, 29, 13}, {25, 1, 28, 13}, {26, 1, 27, 13}};
/* transmission (zig-zag) order, h261_tables.c:316-325 */
static const uint8_t k_zigzag[64] = {0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,
                                     12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13, 6,  7,  14, 21, 28,
                                     35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51,
                                     58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};

/* src/h261_macroblock.c:5-92 */
static void write_tcoeff(bw_t *w, const int32_t *c, int intra)
{
    int run, start;
    if (intra) {
        bw_put(w, (uint32_t)c[0], 8);
        run = 0;
        start = 1;
    } else if (c[0]) {
        int l = c[0];
        if (l == 1 || l == -1) {
            bw_put(w, 1, 1);
            bw_put(w, l < 0, 1);
            start = 1;
        } else
            start = 0;
        run = 0;
    } else {
        run = 1;
        start = 1;
    }
    for (int i = start; i < 64; i++) {
        int l = c[k_zigzag[i]];
        if (!l) {
            run++;
            continue;
        }
        int ul = l < 0 ? -l : l;
        int found = -1;
        for (unsigned k = 0; k < sizeof(k_tcoeff) / sizeof(k_tcoeff[0]); k++)
            if (k_tcoeff[k][0] == run && k_tcoeff[k][1] == ul) {
                found = (int)k;
                break;
            }
        if (found < 0) {
            bw_put(w, 1, 6);
            bw_put(w, (uint32_t)run, 6);
            bw_put(w, (uint32_t)l, 8);
        } else {
            bw_put(w, k_tcoeff[found][2], k_tcoeff[found][3]);
            bw_put(w, l < 0, 1);
        }
        run = 0;
    }
    bw_put(w, 2, 2); /* EOB */
}

typedef struct {
    int quant, mv_valid, mv_x, mv_y;
} coding_vars_t;

/* src/h261_macroblock.c:94-245 */
static void write_macroblock(bw_t *w, const h261_macroblock_t *mb, int addr, coding_vars_t *cv)
{
    /* h261_write_macroblock (macroblock.c:250-262) writes nothing at all for
     * a type other than INTRA/INTER; MBA is written by the two writers */
    if (mb->i_macroblock_type != H261_MB_TYPE_INTRA && mb->i_macroblock_type != H261_MB_TYPE_INTER) return;
    bw_put(w, k_mba[addr][0], k_mba[addr][1]);
    if (mb->i_macroblock_type == H261_MB_TYPE_INTRA) {
        if (mb->i_macroblock_quant != cv->quant) {
            bw_put(w, k_mtype[1][0], k_mtype[1][1]);
            bw_put(w, (uint32_t)mb->i_macroblock_quant, 5);
        } else
            bw_put(w, k_mtype[0][0], k_mtype[0][1]);
        for (int b = 0; b < 6; b++) write_tcoeff(w, mb->rgi_tcoeff[b], 1);
        return;
    }
    /* MTYPE index by [quant coded][flags] (macroblock.c:131-150) */
    static const int type_tab[2][8] = {{-1, 4, -1, 7, 2, 5, -1, 8}, {-1, -1, -1, -1, 3, 6, -1, 9}};
    int flags = mb->i_macroblock_type_flags;
    int cq = mb->i_macroblock_quant != cv->quant && (flags & H261_MB_COEFF);
    int t = (flags >= 0 && flags < 8) ? type_tab[cq][flags] : -1;
    if (t < 0) {
        w->overflow = 2; /* the reference crashes on purpose here (macroblock.c:186-192) */
        return;
    }
    bw_put(w, k_mtype[t][0], k_mtype[t][1]);
    if (cq) bw_put(w, (uint32_t)mb->i_macroblock_quant, 5);
    if (flags & H261_MB_MC) {
        int dx = mb->rgi_mv[0], dy = mb->rgi_mv[1];
        if (cv->mv_valid) {
            dx -= cv->mv_x;
            dy -= cv->mv_y;
        }
        /* rgi_mvd_translation_table: d in [-30,30] -> index (d+16) mod 32 */
        if (dx < -30 || dx > 30 || dy < -30 || dy > 30) {
            w->overflow = 2;
            return;
        }
        int ix = (dx + 16 + 32) & 31, iy = (dy + 16 + 32) & 31;
        bw_put(w, k_mvd[ix][0], k_mvd[ix][1]);
        bw_put(w, k_mvd[iy][0], k_mvd[iy][1]);
    }
    if (flags & H261_MB_COEFF) {
        int cbp = mb->i_coded_block_pattern;
        if (!cbp) {
            w->overflow = 2;
            return;
        }
        bw_put(w, k_cbp[cbp][0], k_cbp[cbp][1]);
        for (int b = 0; b < 6; b++)
            if (cbp & k_cbp_bit[b]) write_tcoeff(w, mb->rgi_tcoeff[b], 0);
    }
}

/* src/h261_picture.c:38-71, src/h261_gob.c:6-199 */
int ora_write_picture(const h261_macroblock_t *mbs, int mb_w, int source_format, int temporal_reference,
                      int pic_quant, uint8_t *out, int cap)
{
    bw_t w = {out, cap, 0, 0};
    memset(out, 0, cap);
    bw_put(&w, 16, 20);                              /* PSC */
    bw_put(&w, (uint32_t)(temporal_reference & 31), 5); /* TR */
    bw_put(&w, 0, 1);                                /* split screen */
    bw_put(&w, 0, 1);                                /* document camera */
    bw_put(&w, 1, 1);                                /* freeze picture release */
    bw_put(&w, (uint32_t)source_format, 1);
    bw_put(&w, 1, 1); /* hi-res bit as the reference sets it (picture.c:12) */
    bw_put(&w, 1, 1); /* spare */
    bw_put(&w, 0, 1); /* PEI */
    int n_gob = source_format == H261_SOURCE_FORMAT_CIF ? 12 : 3;
    for (int g = 0; g < n_gob; g++) {
        int gn, ox, oy;
        if (source_format == H261_SOURCE_FORMAT_CIF) {
            gn = g + 1;
            ox = (g & 1) * 176;
            oy = (g >> 1) * 48;
        } else {
            gn = 2 * g + 1;
            ox = 0;
            oy = g * 48;
        }
        bw_put(&w, 1, 16);
        bw_put(&w, (uint32_t)gn, 4);
        bw_put(&w, (uint32_t)pic_quant, 5);
        bw_put(&w, 0, 1); /* GEI */
        coding_vars_t cv = {pic_quant, 0, 0, 0};
        int addr = 0;
        for (int i = 1; i < 34; i++) {
            addr++;
            int x = ((i - 1) % 11) * 16 + ox, y = ((i - 1) / 11) * 16 + oy;
            if (i == 1 || i == 12 || i == 23) cv.mv_valid = 0;
            const h261_macroblock_t *mb = mbs + (y >> 4) * mb_w + (x >> 4);
            int flags = mb->i_macroblock_type_flags;
            if (flags) {
                write_macroblock(&w, mb, addr, &cv);
                addr = 0;
                if (flags & H261_MB_COEFF) cv.quant = mb->i_macroblock_quant;
                if (flags & H261_MB_MC) {
                    cv.mv_x = mb->rgi_mv[0];
                    cv.mv_y = mb->rgi_mv[1];
                    cv.mv_valid = 1;
                } else
                    cv.mv_valid = 0;
            } else
                cv.mv_valid = 0;
        }
    }
    if (w.overflow) return -1;
    return (int)((w.bitpos + 7) >> 3);
}

/* ------------------------------------------------------------------ */
/* stream: the frame loop of src/h261_main.c:152-249                   */
/* ------------------------------------------------------------------ */

struct ora_stream {
    int width, height, mb_w, mb_h, n_mb, source_format, fixed_quant;
    int frame;
    uint8_t *ref[3];
    h261_dct_denoise_t dn;
    h261_ratectrl_t rc;
    me_gpu_mb_t *me;
    h261_mb_mode_decision_t *md;
    h261_macroblock_t *mbs;
    long sad_candidates;
};

ora_stream_t *ora_stream_create(int width, int height, int denoise, int init_quant, int bitrate,
                                int vbv_size, int fixed_quant)
{
    if (width % 16 || height % 16 || width <= 0 || height <= 0) return NULL;
    ora_stream_t *s = (ora_stream_t *)calloc(1, sizeof(*s));
    s->width = width;
    s->height = height;
    s->mb_w = width / 16;
    s->mb_h = height / 16;
    s->n_mb = s->mb_w * s->mb_h;
    s->source_format = width == 352 ? H261_SOURCE_FORMAT_CIF : H261_SOURCE_FORMAT_QCIF;
    s->fixed_quant = fixed_quant;
    s->ref[0] = (uint8_t *)calloc(width * height, 1); /* h261_init_frame zeroes (main.c:26-28) */
    s->ref[1] = (uint8_t *)calloc(width * height / 4, 1);
    s->ref[2] = (uint8_t *)calloc(width * height / 4, 1);
    s->dn.i_denoise = denoise;
    s->me = (me_gpu_mb_t *)calloc(s->n_mb, sizeof(me_gpu_mb_t));
    s->md = (h261_mb_mode_decision_t *)calloc(s->n_mb, sizeof(h261_mb_mode_decision_t));
    s->mbs = (h261_macroblock_t *)calloc(s->n_mb, sizeof(h261_macroblock_t));
    ora_rc_init(&s->rc, init_quant, bitrate, vbv_size, H261_RATECTRL_MODE_CBR);
    return s;
}

void ora_stream_destroy(ora_stream_t *s)
{
    if (!s) return;
    for (int i = 0; i < 3; i++) free(s->ref[i]);
    free(s->me);
    free(s->md);
    free(s->mbs);
    free(s);
}

long ora_stream_last_sad_candidates(const ora_stream_t *s) { return s->sad_candidates; }

/* the last frame's picture cost as rate control used it
 * (s_ratectrl.i_current_picture_cost, printed by h261_main.c:257-264) */
int ora_stream_last_picture_cost(const ora_stream_t *s) { return s->rc.i_current_picture_cost; }

int ora_stream_encode_frame(ora_stream_t *s, const uint8_t *y, const uint8_t *cb, const uint8_t *cr,
                            uint8_t *bits, int cap, uint8_t *recon_y, uint8_t *recon_cb, uint8_t *recon_cr,
                            me_gpu_mb_t *me_out, h261_macroblock_t *mb_out, int *quant_out)
{
    int ft = s->frame == 0 ? H261_FRAME_TYPE_INTRA : H261_FRAME_TYPE_INTER; /* main.c:206-213 */
    ora_me_frame(y, s->ref[0], s->width, s->height, s->me);
    s->sad_candidates = g_last_sad_candidates;
    ora_mode_decision(s->me, s->n_mb, ft, s->md);
    s->rc.i_current_picture_cost = ora_rc_picture_cost(s->md, s->n_mb);
    int q = ora_rc_init_picture(&s->rc, ft);
    if (s->fixed_quant > 0) q = s->fixed_quant;
    q = imin(31, imax(1, q)); /* h261_init_picture (picture.c:25-33) */
    const uint8_t *cur[3] = {y, cb, cr};
    const uint8_t *ref[3] = {s->ref[0], s->ref[1], s->ref[2]};
    uint8_t *rec[3];
    int ysz = s->width * s->height;
    rec[0] = (uint8_t *)malloc(ysz);
    rec[1] = (uint8_t *)malloc(ysz / 4);
    rec[2] = (uint8_t *)malloc(ysz / 4);
    ora_encode_macroblocks(cur, ref, s->width, s->height, s->md, q, &s->dn, s->mbs, rec);
    int n = ora_write_picture(s->mbs, s->mb_w, s->source_format, s->frame & 31, q, bits, cap);
    if (n >= 0) ora_rc_update_picture(&s->rc, ft, q, n);
    /* the reconstruction becomes the next reference (main.c:247-249) */
    for (int i = 0; i < 3; i++) {
        free(s->ref[i]);
        s->ref[i] = rec[i];
    }
    if (recon_y) memcpy(recon_y, rec[0], ysz);
    if (recon_cb) memcpy(recon_cb, rec[1], ysz / 4);
    if (recon_cr) memcpy(recon_cr, rec[2], ysz / 4);
    if (me_out) memcpy(me_out, s->me, sizeof(me_gpu_mb_t) * s->n_mb);
    if (mb_out) memcpy(mb_out, s->mbs, sizeof(h261_macroblock_t) * s->n_mb);
    if (quant_out) *quant_out = q;
    s->frame++;
    return n;
}

/* sizes and offsets of the ABI structs as include/h261gpu_abi.h lays them
 * out, in the order of refdrv_layout (oracle/ref_driver.c) */
#include <stddef.h>
int ora_layout(long *out, int cap)
{
    long v[] = {
        (long)sizeof(me_gpu_mv_t),
        (long)sizeof(me_gpu_mb_t),
        (long)offsetof(me_gpu_mb_t, i_16x16_cost),
        (long)offsetof(me_gpu_mb_t, s_16x16_mv),
        (long)sizeof(me_gpu_t),
        (long)offsetof(me_gpu_t, prgs_macroblocks_result),
        (long)offsetof(me_gpu_t, p_cuda),
        (long)sizeof(h261_mb_mode_decision_t),
        (long)sizeof(h261_mode_decision_t),
        (long)sizeof(h261_frame_t),
        (long)sizeof(h261_bitstream_t),
        (long)sizeof(h261_dct_denoise_t),
        (long)offsetof(h261_dct_denoise_t, rgi_inter_signal_offset),
        (long)sizeof(h261_macroblock_t),
        (long)offsetof(h261_macroblock_t, rgi_tcoeff),
        (long)offsetof(h261_macroblock_t, i_mb_x),
        (long)sizeof(h261_macroblocks_t),
        (long)offsetof(h261_macroblocks_t, ps_macroblocks),
        (long)offsetof(h261_macroblocks_t, pui8_reconstructed_Y_host),
        (long)offsetof(h261_macroblocks_t, pui8_reconstructed_Cb_host),
        (long)offsetof(h261_macroblocks_t, pui8_reconstructed_Cr_host),
        (long)offsetof(h261_macroblocks_t, i_num_blocks),
        (long)offsetof(h261_macroblocks_t, p_cuda),
        (long)sizeof(h261_gob_t),
        (long)sizeof(h261_picture_t),
        (long)offsetof(h261_picture_t, i_quantiser),
        (long)sizeof(h261_ratectrl_t),
        (long)sizeof(h261_context_t),
        (long)offsetof(h261_context_t, s_picture),
        (long)offsetof(h261_context_t, ps_mode_decision),
        (long)offsetof(h261_context_t, ps_macroblocks),
        (long)sizeof(h261_picture_parameters_t),
    };
    int n = (int)(sizeof(v) / sizeof(v[0]));
    for (int i = 0; i < n && i < cap; i++) out[i] = v[i];
    return n;
}
