/*
 * ref_dropin.c -- the reference's frame loop over the device C ABI.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the loop of the reference's
 * src/h261_main.c:140-296 (without file I/O, printing and timing, which need
 * the Win32 timer the image lacks) compiled against the reference's own
 * headers and host sources, and linked against the product library
 * h261enc_cuda_amd/libh261gpu.so, which resolves the 16 entry points of the
 * reference's src/gpu_device.h.  It demonstrates the drop-in: the
 * reference's mode decision, rate control and bitstream writer run
 * unchanged on top of the HIP device layer.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "h261_decl.h" /* the reference's header (-I<reference>/src) */

static void frame_alloc(h261_frame_t *f, int w, int h)
{
    /* h261_init_frame (h261_main.c:10-29) */
    f->i_width = w;
    f->i_height = h;
    f->i_stride_y = w;
    f->i_stride_c = w / 2;
    f->pui8_Y = (UInt8 *)calloc(w * h, 1);
    f->pui8_Cb = (UInt8 *)calloc(w * h / 4, 1);
    f->pui8_Cr = (UInt8 *)calloc(w * h / 4, 1);
}

static void frame_free(h261_frame_t *f)
{
    free(f->pui8_Y);
    free(f->pui8_Cb);
    free(f->pui8_Cr);
}

/* frames: n_frames I420 pictures back to back.  Returns the total number of
 * bitstream bytes, or -1 if bits_cap is too small. */
long refdrv_dropin_encode(int width, int height, int source_format, int n_frames,
                          const unsigned char *frames, int denoise, int init_quant, int bitrate,
                          int vbv, int fixed_quant, unsigned char *bits_out, long bits_cap,
                          int *frame_bytes, unsigned char *recon_out, int *quants_out,
                          h261_macroblock_t *last_mbs_out)
{
    int fsize = width * height, csize = fsize / 4;
    long total = 0;
    h261_context_t *ctx = (h261_context_t *)calloc(1, sizeof(h261_context_t));
    me_gpu_t *me = 0;
    h261_mode_decision_t *md = 0;
    h261_macroblocks_t *mbs = 0;

    ctx->i_source_format = source_format;
    frame_alloc(&ctx->s_current_frame, width, height);
    frame_alloc(&ctx->s_reference_frame, width, height);
    h261_bitstream_init(&ctx->s_bitstream, 0x100000);

    h261_gpu_init();
    h261_gpu_device_init_textures(width, height);
    h261_init_macroblocks(&mbs, width, height, denoise);
    h261_gpu_device_init_me(&me, width, height);
    h261_init_mode_decision(&md, me->i_mb_width, me->i_mb_height);
    h261_init_ratectrl(ctx, init_quant, bitrate, vbv, H261_RATECTRL_MODE_CBR);

    for (int f = 0; f < n_frames; f++) {
        h261_frame_t *src = &ctx->s_current_frame, *ref = &ctx->s_reference_frame;
        h261_picture_parameters_t pp;
        UInt8 *p;
        UInt32 len;
        const unsigned char *in = frames + (long)f * (fsize + 2 * csize);
        memcpy(src->pui8_Y, in, fsize);
        memcpy(src->pui8_Cb, in + fsize, csize);
        memcpy(src->pui8_Cr, in + fsize + csize, csize);

        h261_gpu_device_set_current(src->pui8_Y, width, height);
        h261_gpu_device_set_current_chroma_cb(src->pui8_Cb, width / 2, height / 2);
        h261_gpu_device_set_current_chroma_cr(src->pui8_Cr, width / 2, height / 2);
        h261_gpu_device_set_reference(ref->pui8_Y, width, height);
        h261_gpu_device_set_reference_chroma_cb(ref->pui8_Cb, width / 2, height / 2);
        h261_gpu_device_set_reference_chroma_cr(ref->pui8_Cr, width / 2, height / 2);

        h261_gpu_device_me(me);

        memset(&pp, 0, sizeof(pp));
        pp.i_frame_num = f;
        pp.i_source_format = ctx->i_source_format;
        pp.i_frame_type = f == 0 ? H261_FRAME_TYPE_INTRA : H261_FRAME_TYPE_INTER;

        h261_mode_decision(md, me, &pp);
        ctx->ps_mode_decision = md;
        h261_ratectrl_collect_picture_stats(ctx, &pp);
        h261_ratectrl_init_picture(ctx, &pp);
        if (fixed_quant > 0) pp.i_quantiser = fixed_quant;
        h261_init_picture(ctx, &pp);
        ctx->ps_macroblocks = mbs;

        h261_gpu_device_encode_macroblocks(ctx);

        h261_bitstream_reset(&ctx->s_bitstream);
        h261_write_picture_header(ctx);
        h261_write_gobs(ctx);

        h261_gpu_device_decode_macroblocks(ctx);

        h261_bitstream_get(&ctx->s_bitstream, &p, &len);
        if (total + (long)len > bits_cap) {
            total = -1;
            break;
        }
        memcpy(bits_out + total, p, len);
        total += len;
        if (frame_bytes) frame_bytes[f] = (int)len;
        if (quants_out) quants_out[f] = ctx->s_picture.i_quantiser;

        h261_ratectrl_update_picture(ctx, (Int32)len);

        memcpy(ref->pui8_Y, mbs->pui8_reconstructed_Y_host, fsize);
        memcpy(ref->pui8_Cb, mbs->pui8_reconstructed_Cb_host, csize);
        memcpy(ref->pui8_Cr, mbs->pui8_reconstructed_Cr_host, csize);
        if (recon_out) {
            unsigned char *o = recon_out + (long)f * (fsize + 2 * csize);
            memcpy(o, ref->pui8_Y, fsize);
            memcpy(o + fsize, ref->pui8_Cb, csize);
            memcpy(o + fsize + csize, ref->pui8_Cr, csize);
        }
        if (last_mbs_out && f == n_frames - 1)
            memcpy(last_mbs_out, mbs->ps_macroblocks, sizeof(h261_macroblock_t) * mbs->i_num_mb);
    }

    h261_gpu_device_deinit_textures();
    h261_gpu_device_deinit_me(&me);
    h261_deinit_mode_decision(md);
    h261_deinit_macroblocks(mbs);
    h261_bitstream_deinit(&ctx->s_bitstream);
    frame_free(&ctx->s_current_frame);
    frame_free(&ctx->s_reference_frame);
    free(ctx);
    return total;
}
