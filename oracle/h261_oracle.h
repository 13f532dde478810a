/*
 * h261_oracle.h -- CPU restatement of the reference H.261 encoder core.
 *
 * TEST INFRASTRUCTURE ONLY.  This code is the checker that the HIP device
 * layer is compared against; it is never linked into, loaded by or called
 * from the product library.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may use it.
 *
 * Every function restates one piece of the reference and cites the
 * reference file:line it follows.  Build flags: -O2 -ffp-contract=off (the
 * forward/inverse DCT and the forward quantiser depend on op-by-op float /
 * double rounding with no fused multiply-add).
 *
 * Parity unpinned: the float/double op typing of ora_fdct8x8 and
 * ora_idct8x8 (their rare rounding cases).  The reference's CUDA kernels
 * cannot be built here, so these two are checked only against an exact
 * float64 transform (within +-1) and round trips; every other function is
 * pinned by the reference's host sources built under oracle/_ref, by brute
 * force, closed forms or planted-mistake checks (tests/test_oracle_*.py).
 */
#ifndef H261_ORACLE_H
#define H261_ORACLE_H

#include <stdint.h>
#include "../include/h261gpu_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- unit-level functions (one reference helper each) ---- */

/* per-MB [dx_min, dx_max, dy_min, dy_max]; src/gpu_device.cu:226-265 */
void ora_mv_limits(int mb_w, int mb_h, int8_t *limits /* 4*mb_w*mb_h */);

/* 16x16 SAD; src/megpu_kernels.cuh:3-21 */
int ora_sad16(const uint8_t *cur, const uint8_t *ref, int stride, int cx, int cy, int rx, int ry);

/* unnormalised 4x4 Hadamard SATD; src/megpu_kernels.cuh:467-500 */
int ora_satd4x4(const int16_t d[16]);

/* quad-tree minimum over a dim x dim candidate grid; src/megpu_kernels.cuh:95-154 */
int ora_select_quadtree(const uint16_t *cost, int dim);

/* whole-frame motion estimation: coarse grid, refine, 3 SATD costs.
 * src/gpu_device.cu:356-432 launch order, src/megpu_kernels.cuh:27-987 */
void ora_me_frame(const uint8_t *cur_y, const uint8_t *ref_y, int width, int height,
                  me_gpu_mb_t *out /* mb_w*mb_h */);

/* forward AAN DCT of an 8x8 block (raster in/out); src/mbgpu_kernels.cuh:337-442 */
void ora_fdct8x8(const int32_t in[64], int32_t out[64]);
/* inverse AAN DCT (raster in/out, output clamped to [-256,255]);
 * src/mbgpu_kernels.cuh:827-931 */
void ora_idct8x8(const int32_t in[64], int32_t out[64]);

/* forward quantiser of one block in place; returns 1 when the block is coded.
 * src/mbgpu_kernels.cuh:644-724 */
int ora_quant_block(int32_t c[64], int quant, int intra);
/* H.261 dequantiser of one block in place; src/mbgpu_kernels.cuh:731-815 */
void ora_dequant_block(int32_t c[64], int quant, int intra, int coded);

/* motion-compensated prediction of one 8x8 block, with the H.261 loop
 * filter when filter != 0; src/mbgpu_kernels.cuh:75-166 and 1006-1096 */
void ora_predict8x8(const uint8_t *ref, int stride, int px, int py, int filter, int32_t pred[64]);

/* ---- frame-level encoder (restatement of the whole reference loop) ---- */

typedef struct ora_stream ora_stream_t;

/* width/height: 352x288 (CIF) or 176x144 (QCIF).  denoise: DCT denoise
 * strength (src/h261_main.c:162 uses 5).  init_quant/bitrate/vbv: rate
 * control init (src/h261_main.c:165 uses 10, 512000, 512000).
 * fixed_quant > 0 overrides the picture quantiser after rate control (the
 * reference has no fixed-quant mode; rate control state still evolves). */
ora_stream_t *ora_stream_create(int width, int height, int denoise, int init_quant,
                                int bitrate, int vbv_size, int fixed_quant);
void ora_stream_destroy(ora_stream_t *s);

/* Encodes one frame exactly as src/h261_main.c:186-249 does.  Writes the
 * frame's bytes to bits (cap bytes) and returns their count, or -1 on
 * overflow.  Optional outputs (may be NULL): the reconstruction (also kept
 * as the next reference), the ME results, the macroblock structs, the
 * picture quantiser used. */
int ora_stream_encode_frame(ora_stream_t *s, const uint8_t *y, const uint8_t *cb, const uint8_t *cr,
                            uint8_t *bits, int cap, uint8_t *recon_y, uint8_t *recon_cb,
                            uint8_t *recon_cr, me_gpu_mb_t *me_out, h261_macroblock_t *mb_out,
                            int *quant_out);

/* Number of SAD candidates the two-stage search evaluated in the last frame
 * (valid coarse + valid refine); used to price the SAD roofline. */
long ora_stream_last_sad_candidates(const ora_stream_t *s);
int ora_stream_last_picture_cost(const ora_stream_t *s);

/* ---- pieces of the frame loop, exposed for tests ---- */

/* src/h261_moded.c:31-100 */
void ora_mode_decision(const me_gpu_mb_t *me, int num_mb, int frame_type,
                       h261_mb_mode_decision_t *md);

/* src/h261_ratectrl.c:107-245 */
void ora_rc_init(h261_ratectrl_t *rc, int quant, int bitrate, int vbv_size, int mode);
int ora_rc_picture_cost(const h261_mb_mode_decision_t *md, int num_mb);
int ora_rc_init_picture(h261_ratectrl_t *rc, int frame_type); /* returns unclamped quantiser */
void ora_rc_update_picture(h261_ratectrl_t *rc, int frame_type, int quant, int bytes);

/* MB coder: encode (setup, DCT, denoise, denoise update, quant) followed by
 * decode (dequant, IDCT, reconstruction).  src/gpu_device.cu:592-815 */
void ora_encode_macroblocks(const uint8_t *cur[3], const uint8_t *ref[3], int width, int height,
                            const h261_mb_mode_decision_t *md, int pic_quant,
                            h261_dct_denoise_t *denoise, h261_macroblock_t *mbs,
                            uint8_t *recon[3]);

/* bitstream of one picture: src/h261_picture.c:38-71, src/h261_gob.c:6-199,
 * src/h261_macroblock.c:5-262.  Returns byte count or -1 on overflow. */
int ora_write_picture(const h261_macroblock_t *mbs, int mb_w, int source_format,
                      int temporal_reference, int pic_quant, uint8_t *out, int cap);

#ifdef __cplusplus
}
#endif

#endif
