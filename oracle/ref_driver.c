/*
 * ref_driver.c -- thin C entry points over the reference's own host code.
 *
 * TEST INFRASTRUCTURE ONLY.  oracle/Makefile compiles this file together
 * with the reference's unmodified host sources (h261_bitstream.c,
 * h261_gob.c, h261_macroblock.c, h261_picture.c, h261_ratectrl.c,
 * h261_tables.c, h261_moded.c) read in place from the reference tree, with
 * the reference's own headers, into oracle/_ref/libh261ref.so.  Tests use
 * it to pin the oracle's writer, mode decision and rate control, and to
 * measure struct layouts with the reference's headers.
 */
#include <stddef.h>
#include <stdint.h>
#include <string.h>

#include "h261_decl.h" /* the reference's header (-I<reference>/src) */

/* sizes and offsets of every ABI struct, as the reference headers lay them
 * out; the same list is produced from include/h261gpu_abi.h by the product */
int refdrv_layout(long *out, int cap)
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

/* one picture through the reference writer (h261_picture.c, h261_gob.c,
 * h261_macroblock.c, h261_bitstream.c) */
int refdrv_write_picture(const h261_macroblock_t *mbs, int mb_w, int mb_h, int source_format,
                         int frame_num, int quant, int frame_type, unsigned char *out, int cap)
{
    h261_context_t ctx;
    h261_macroblocks_t m;
    h261_picture_parameters_t pp;
    UInt8 *p;
    UInt32 len;
    memset(&ctx, 0, sizeof(ctx));
    memset(&m, 0, sizeof(m));
    ctx.i_source_format = source_format;
    pp.i_frame_num = frame_num;
    pp.i_source_format = source_format;
    pp.i_quantiser = quant;
    pp.i_frame_type = frame_type;
    h261_init_picture(&ctx, &pp);
    m.i_mb_width = mb_w;
    m.i_mb_height = mb_h;
    m.i_num_mb = mb_w * mb_h;
    m.ps_macroblocks = (h261_macroblock_t *)mbs;
    ctx.ps_macroblocks = &m;
    h261_bitstream_init(&ctx.s_bitstream, 0x100000);
    h261_bitstream_reset(&ctx.s_bitstream);
    h261_write_picture_header(&ctx);
    h261_write_gobs(&ctx);
    h261_bitstream_get(&ctx.s_bitstream, &p, &len);
    int n = (int)len;
    if (n > cap) n = -1;
    else memcpy(out, p, len);
    h261_bitstream_deinit(&ctx.s_bitstream);
    return n;
}

/* h261_mode_decision (h261_moded.c:31-123) over given ME results */
void refdrv_mode_decision(const me_gpu_mb_t *me, int mb_w, int mb_h, int frame_type,
                          h261_mb_mode_decision_t *out)
{
    me_gpu_t g;
    h261_mode_decision_t *md;
    h261_picture_parameters_t pp;
    memset(&g, 0, sizeof(g));
    g.i_mb_width = mb_w;
    g.i_mb_height = mb_h;
    g.i_num_mb = mb_w * mb_h;
    g.prgs_macroblocks_result = (me_gpu_mb_t *)me;
    memset(&pp, 0, sizeof(pp));
    pp.i_frame_type = frame_type;
    h261_init_mode_decision(&md, mb_w, mb_h);
    h261_mode_decision(md, &g, &pp);
    memcpy(out, md->prgs_mb, sizeof(h261_mb_mode_decision_t) * mb_w * mb_h);
    h261_deinit_mode_decision(md);
}

/* rate control as driven by h261_main.c:165, 219-224, 243 */
typedef struct {
    h261_context_t ctx;
    h261_mode_decision_t md;
    h261_picture_parameters_t pp;
} refdrv_rc_t;

static refdrv_rc_t g_rc[16];

int refdrv_rc_init(int slot, int quant, int bitrate, int vbv)
{
    if (slot < 0 || slot >= 16) return -1;
    memset(&g_rc[slot], 0, sizeof(g_rc[slot]));
    h261_init_ratectrl(&g_rc[slot].ctx, quant, bitrate, vbv, H261_RATECTRL_MODE_CBR);
    return 0;
}

/* returns the picture quantiser after h261_init_picture's clamp; a
 * fixed_quant > 0 replaces the rate-control quantiser before the clamp */
int refdrv_rc_begin(int slot, h261_mb_mode_decision_t *mds, int mb_w, int mb_h, int frame_num,
                    int frame_type, int fixed_quant)
{
    refdrv_rc_t *r = &g_rc[slot];
    r->md.i_mb_width = mb_w;
    r->md.i_mb_height = mb_h;
    r->md.prgs_mb = mds;
    r->ctx.ps_mode_decision = &r->md;
    r->pp.i_frame_num = frame_num;
    r->pp.i_frame_type = frame_type;
    h261_ratectrl_collect_picture_stats(&r->ctx, &r->pp);
    h261_ratectrl_init_picture(&r->ctx, &r->pp);
    if (fixed_quant > 0) r->pp.i_quantiser = fixed_quant;
    h261_init_picture(&r->ctx, &r->pp);
    return r->ctx.s_picture.i_quantiser;
}

void refdrv_rc_end(int slot, int bytes) { h261_ratectrl_update_picture(&g_rc[slot].ctx, bytes); }

void refdrv_rc_state(int slot, int *out /* 9 ints */)
{
    h261_ratectrl_t *rc = &g_rc[slot].ctx.s_ratectrl;
    out[0] = rc->i_ratectrl_mode;
    out[1] = rc->i_current_quant;
    out[2] = rc->i_current_picture_cost;
    out[3] = rc->i_prediction_cost[0];
    out[4] = rc->i_prediction_cost[1];
    out[5] = rc->i_prediction_size[0];
    out[6] = rc->i_prediction_size[1];
    out[7] = rc->i_bitrate;
    out[8] = rc->i_vbv_fill;
}
