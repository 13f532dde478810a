/*
 * legacy_abi.hip -- the 16 entry points of the reference's src/gpu_device.h
 * on the gfx950 kernels.
 *
 * Behaviour follows src/gpu_device.cu: one process-wide encoder, current
 * and reference planes set by synchronous copies, motion estimation and
 * macroblock coding as blocking calls, output in the caller-visible pinned
 * structs.  The reference's 5 ME launches become one (me_kernel), its 7
 * encode + 5 decode launches become one fused launch plus the denoise
 * update; decode therefore only copies the reconstruction the encode call
 * already made (the device state the reference's decode reads is exactly
 * what its encode left behind, so the results are the same).
 */
#include <hip/hip_runtime.h>

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/h261gpu.h"
#include "h261gpu_internal.h"

using namespace h261g;

namespace {

int g_last_error = 0;

bool check(hipError_t e, const char *what)
{
    if (e == hipSuccess) return true;
    fprintf(stderr, "h261gpu: %s failed: %s\n", what, hipGetErrorString(e));
    g_last_error = H261G_ERR_DEVICE;
    return false;
}

bool verbose() { return getenv("H261GPU_VERBOSE") != nullptr; }

/* the six planes the reference keeps in textures (gpu_device.cu:26-42) */
struct Planes {
    int width = 0, height = 0;
    uint8_t *cur[3] = {nullptr, nullptr, nullptr};
    uint8_t *ref[3] = {nullptr, nullptr, nullptr};
};
Planes g_planes;

/* backend state behind me_gpu_t::p_cuda and h261_macroblocks_t::p_cuda */
struct Backend {
    hipStream_t stream = nullptr;
    StreamDesc *desc_host = nullptr; /* pinned */
    StreamDesc *desc_dev = nullptr;
    int32_t *mb_stream_dev = nullptr;
    MeResult *me_dev = nullptr;
    MbMode *mode_host = nullptr; /* pinned */
    MbMode *mode_dev = nullptr;
};

bool backend_init(Backend *b, int n_mb)
{
    if (!check(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking), "hipStreamCreate")) return false;
    if (!check(hipHostMalloc((void **)&b->desc_host, sizeof(StreamDesc), hipHostMallocDefault), "hipHostMalloc"))
        return false;
    memset(b->desc_host, 0, sizeof(StreamDesc));
    if (!check(hipMalloc((void **)&b->desc_dev, sizeof(StreamDesc)), "hipMalloc")) return false;
    if (!check(hipMalloc((void **)&b->mb_stream_dev, sizeof(int32_t) * n_mb), "hipMalloc")) return false;
    if (!check(hipMemset(b->mb_stream_dev, 0, sizeof(int32_t) * n_mb), "hipMemset")) return false;
    return true;
}

void backend_free(Backend *b)
{
    if (!b) return;
    if (b->stream) hipStreamDestroy(b->stream);
    if (b->desc_host) hipHostFree(b->desc_host);
    if (b->desc_dev) hipFree(b->desc_dev);
    if (b->mb_stream_dev) hipFree(b->mb_stream_dev);
    if (b->me_dev) hipFree(b->me_dev);
    if (b->mode_host) hipHostFree(b->mode_host);
    if (b->mode_dev) hipFree(b->mode_dev);
    delete b;
}

void set_plane(uint8_t *dst, const UInt8 *src, Int32 w, Int32 h, const char *what)
{
    if (!dst || !src || w <= 0 || h <= 0) {
        fprintf(stderr, "h261gpu: %s: planes not initialised or invalid arguments\n", what);
        g_last_error = H261G_ERR_ARG;
        return;
    }
    check(hipMemcpy(dst, src, (size_t)w * h, hipMemcpyHostToDevice), what);
}

template <class T>
T *host_alloc(size_t n)
{
    T *p = nullptr;
    if (!check(hipHostMalloc((void **)&p, sizeof(T) * n, hipHostMallocDefault), "hipHostMalloc")) return nullptr;
    memset(p, 0, sizeof(T) * n);
    return p;
}

template <class T>
T *dev_alloc(size_t n)
{
    T *p = nullptr;
    if (!check(hipMalloc((void **)&p, sizeof(T) * n), "hipMalloc")) return nullptr;
    return p;
}

}  // namespace

extern "C" {

int h261gpu_last_error(void)
{
    int e = g_last_error;
    g_last_error = 0;
    return e;
}

const char *h261gpu_error_string(int err)
{
    switch (err) {
    case H261G_OK: return "ok";
    case H261G_ERR_ARG: return "invalid argument";
    case H261G_ERR_DEVICE: return "HIP runtime failure";
    case H261G_ERR_MEMORY: return "allocation failed";
    case H261G_ERR_FULL: return "no free stream slot or HBM budget exhausted";
    case H261G_ERR_STATE: return "call out of order";
    default: return "unknown error";
    }
}

/* src/gpu_device.cu:58-87 */
Void h261_gpu_init(void)
{
    int n = 0;
    if (!check(hipGetDeviceCount(&n), "hipGetDeviceCount") || n <= 0) {
        fprintf(stderr, "h261gpu: no HIP device\n");
        g_last_error = H261G_ERR_DEVICE;
        return;
    }
    if (verbose()) {
        for (int i = 0; i < n; i++) {
            hipDeviceProp_t pr;
            if (hipGetDeviceProperties(&pr, i) == hipSuccess)
                printf("device: %s (%d) %s, %d CUs, %zu bytes\n", pr.name, i, pr.gcnArchName,
                       pr.multiProcessorCount, (size_t)pr.totalGlobalMem);
        }
    }
    check(hipSetDevice(0), "hipSetDevice");
}

/* src/gpu_device.cu:90-145 */
Void h261_gpu_device_init_textures(Int32 i_width, Int32 i_height)
{
    if (i_width <= 0 || i_height <= 0 || i_width % 16 || i_height % 16) {
        fprintf(stderr, "h261gpu: init_textures: %dx%d is not a multiple of 16\n", i_width, i_height);
        g_last_error = H261G_ERR_ARG;
        return;
    }
    h261_gpu_device_deinit_textures();
    g_planes.width = i_width;
    g_planes.height = i_height;
    size_t sz[3] = {(size_t)i_width * i_height, (size_t)i_width * i_height / 4, (size_t)i_width * i_height / 4};
    for (int i = 0; i < 3; i++) {
        g_planes.cur[i] = dev_alloc<uint8_t>(sz[i]);
        g_planes.ref[i] = dev_alloc<uint8_t>(sz[i]);
        if (g_planes.cur[i]) hipMemset(g_planes.cur[i], 0, sz[i]);
        if (g_planes.ref[i]) hipMemset(g_planes.ref[i], 0, sz[i]);
    }
}

/* src/gpu_device.cu:148-156 */
Void h261_gpu_device_deinit_textures(void)
{
    for (int i = 0; i < 3; i++) {
        if (g_planes.cur[i]) hipFree(g_planes.cur[i]);
        if (g_planes.ref[i]) hipFree(g_planes.ref[i]);
        g_planes.cur[i] = g_planes.ref[i] = nullptr;
    }
    g_planes.width = g_planes.height = 0;
}

/* src/gpu_device.cu:158-215 */
Void h261_gpu_device_set_current(UInt8 *p, Int32 w, Int32 h) { set_plane(g_planes.cur[0], p, w, h, "set_current"); }
Void h261_gpu_device_set_current_chroma_cb(UInt8 *p, Int32 w, Int32 h)
{
    set_plane(g_planes.cur[1], p, w, h, "set_current_chroma_cb");
}
Void h261_gpu_device_set_current_chroma_cr(UInt8 *p, Int32 w, Int32 h)
{
    set_plane(g_planes.cur[2], p, w, h, "set_current_chroma_cr");
}
Void h261_gpu_device_set_reference(UInt8 *p, Int32 w, Int32 h) { set_plane(g_planes.ref[0], p, w, h, "set_reference"); }
Void h261_gpu_device_set_reference_chroma_cb(UInt8 *p, Int32 w, Int32 h)
{
    set_plane(g_planes.ref[1], p, w, h, "set_reference_chroma_cb");
}
Void h261_gpu_device_set_reference_chroma_cr(UInt8 *p, Int32 w, Int32 h)
{
    set_plane(g_planes.ref[2], p, w, h, "set_reference_chroma_cr");
}

/* src/gpu_device.cu:286-322 (limits 226-265, zero candidates 267-283) */
void h261_gpu_device_init_me(me_gpu_t **pps, Int32 i_width, Int32 i_height)
{
    if (!pps || i_width <= 0 || i_height <= 0 || i_width % 16 || i_height % 16) {
        fprintf(stderr, "h261gpu: init_me: invalid arguments\n");
        g_last_error = H261G_ERR_ARG;
        return;
    }
    me_gpu_t *me = host_alloc<me_gpu_t>(1);
    if (!me) return;
    me->i_mb_width = i_width / 16;
    me->i_mb_height = i_height / 16;
    me->i_num_mb = me->i_mb_width * me->i_mb_height;
    const int n = me->i_num_mb;
    Backend *b = new Backend();
    if (!backend_init(b, n)) {
        backend_free(b);
        hipHostFree(me);
        return;
    }
    b->me_dev = dev_alloc<MeResult>(n);
    me->prgs_macroblocks_device = (me_gpu_mb_t *)b->me_dev;
    me->prgs_macroblocks_result = host_alloc<me_gpu_mb_t>(n);
    me->ps_candidate_vector_device = dev_alloc<me_gpu_mv_t>(n);
    me->ps_candidate_vector_host = host_alloc<me_gpu_mv_t>(n);
    me->ps_starting_vector_device = dev_alloc<me_gpu_mv_t>(n);
    me->ps_starting_vector_host = host_alloc<me_gpu_mv_t>(n);
    me->pi8_motion_vector_limits_device = dev_alloc<Int8>(4 * n);
    me->pi8_motion_vector_limits_host = host_alloc<Int8>(4 * n);
    me->p_cuda = b;
    /* motion vector limits, kept for callers that read them; the kernel
     * derives the same limits from the MB position */
    const int gmx = (me->i_mb_width - 1) * 16, gmy = (me->i_mb_height - 1) * 16;
    for (int y = 0; y < me->i_mb_height; y++)
        for (int x = 0; x < me->i_mb_width; x++) {
            Int8 *l = me->pi8_motion_vector_limits_host + 4 * (x + y * me->i_mb_width);
            int bx = x * 16, by = y * 16;
            l[0] = (Int8)((bx - 15 > 0 ? bx - 15 : 0) - bx);
            l[1] = (Int8)((bx + 15 < gmx ? bx + 15 : gmx) - bx);
            l[2] = (Int8)((by - 15 > 0 ? by - 15 : 0) - by);
            l[3] = (Int8)((by + 15 < gmy ? by + 15 : gmy) - by);
        }
    if (me->pi8_motion_vector_limits_device)
        check(hipMemcpy(me->pi8_motion_vector_limits_device, me->pi8_motion_vector_limits_host, 4 * n,
                        hipMemcpyHostToDevice),
              "hipMemcpy");
    if (me->ps_candidate_vector_device)
        check(hipMemcpy(me->ps_candidate_vector_device, me->ps_candidate_vector_host, sizeof(me_gpu_mv_t) * n,
                        hipMemcpyHostToDevice),
              "hipMemcpy");
    *pps = me;
}

/* src/gpu_device.cu:324-353 */
void h261_gpu_device_deinit_me(me_gpu_t **pps)
{
    if (!pps || !*pps) return;
    me_gpu_t *me = *pps;
    Backend *b = (Backend *)me->p_cuda;
    b->me_dev = nullptr; /* freed below as prgs_macroblocks_device */
    if (me->prgs_macroblocks_device) hipFree(me->prgs_macroblocks_device);
    if (me->prgs_macroblocks_result) hipHostFree(me->prgs_macroblocks_result);
    if (me->ps_candidate_vector_device) hipFree(me->ps_candidate_vector_device);
    if (me->ps_candidate_vector_host) hipHostFree(me->ps_candidate_vector_host);
    if (me->ps_starting_vector_device) hipFree(me->ps_starting_vector_device);
    if (me->ps_starting_vector_host) hipHostFree(me->ps_starting_vector_host);
    if (me->pi8_motion_vector_limits_device) hipFree(me->pi8_motion_vector_limits_device);
    if (me->pi8_motion_vector_limits_host) hipHostFree(me->pi8_motion_vector_limits_host);
    backend_free(b);
    hipHostFree(me);
    *pps = nullptr;
}

/* src/gpu_device.cu:356-432 */
void h261_gpu_device_me(me_gpu_t *me)
{
    if (!me || !me->p_cuda || !g_planes.cur[0] || g_planes.width != me->i_mb_width * 16 ||
        g_planes.height != me->i_mb_height * 16) {
        fprintf(stderr, "h261gpu: me: encoder not initialised for this size\n");
        g_last_error = H261G_ERR_STATE;
        return;
    }
    Backend *b = (Backend *)me->p_cuda;
    const int n = me->i_num_mb;
    StreamDesc *d = b->desc_host;
    memset(d, 0, sizeof(*d));
    for (int i = 0; i < 3; i++) {
        d->cur[i] = g_planes.cur[i];
        d->ref[i] = g_planes.ref[i];
    }
    d->width = g_planes.width;
    d->height = g_planes.height;
    d->mb_w = me->i_mb_width;
    d->mb_h = me->i_mb_height;
    d->frame_type = H261_FRAME_TYPE_INTER; /* the reference searches on every frame */
    d->flags = kStreamActive;
    if (!check(hipMemcpyAsync(b->desc_dev, d, sizeof(StreamDesc), hipMemcpyHostToDevice, b->stream), "hipMemcpyAsync"))
        return;
    MeLaunch L{};
    L.streams = b->desc_dev;
    L.mb_stream = b->mb_stream_dev;
    L.n_mb = n;
    L.result = b->me_dev;
    L.mv_start = (int8_t *)me->ps_starting_vector_device;
    L.skip_search_on_intra = 0;
    /* H261GPU_ME_VARIANT=0|1 selects the search kernel (same results) */
    const char *v = getenv("H261GPU_ME_VARIANT");
    L.variant = v && v[0] == '1' ? kMeSurface : kMeTwoStage;
    if (launch_me(L, b->stream) != 0) {
        check(hipGetLastError(), "me_kernel launch");
        return;
    }
    check(hipMemcpyAsync(me->ps_starting_vector_host, me->ps_starting_vector_device, sizeof(me_gpu_mv_t) * n,
                         hipMemcpyDeviceToHost, b->stream),
          "hipMemcpyAsync");
    check(hipMemcpyAsync(me->prgs_macroblocks_result, me->prgs_macroblocks_device, sizeof(me_gpu_mb_t) * n,
                         hipMemcpyDeviceToHost, b->stream),
          "hipMemcpyAsync");
    check(hipMemcpyAsync(me->ps_candidate_vector_device, me->ps_starting_vector_device, sizeof(me_gpu_mv_t) * n,
                         hipMemcpyDeviceToDevice, b->stream),
          "hipMemcpyAsync");
    if (!check(hipStreamSynchronize(b->stream), "hipStreamSynchronize")) return;
    /* merge the starting vectors into the results (gpu_device.cu:418-422) */
    for (int i = 0; i < n; i++) me->prgs_macroblocks_result[i].s_16x16_mv = me->ps_starting_vector_host[i];
}

/* src/gpu_device.cu:442-539 */
Void h261_init_macroblocks(h261_macroblocks_t **pps, Int32 i_width, Int32 i_height, Int32 i_denoise)
{
    if (!pps || i_width <= 0 || i_height <= 0 || i_width % 16 || i_height % 16) {
        fprintf(stderr, "h261gpu: init_macroblocks: invalid arguments\n");
        g_last_error = H261G_ERR_ARG;
        return;
    }
    const int mw = i_width >> 4, mh = i_height >> 4, n = mw * mh;
    h261_macroblocks_t *m = host_alloc<h261_macroblocks_t>(1);
    if (!m) return;
    m->i_mb_width = mw;
    m->i_mb_height = mh;
    m->i_num_mb = n;
    m->ps_macroblocks_device = dev_alloc<h261_macroblock_t>(n);
    m->ps_macroblocks = host_alloc<h261_macroblock_t>(n);
    for (int y = 0; y < mh; y++)
        for (int x = 0; x < mw; x++) {
            m->ps_macroblocks[y * mw + x].i_mb_x = x * 16;
            m->ps_macroblocks[y * mw + x].i_mb_y = y * 16;
        }
    if (m->ps_macroblocks_device)
        check(hipMemcpy(m->ps_macroblocks_device, m->ps_macroblocks, sizeof(h261_macroblock_t) * n,
                        hipMemcpyHostToDevice),
              "hipMemcpy");
    m->pi_mb_types_host = host_alloc<Int32>(n);
    m->pi_mb_flags_host = host_alloc<Int32>(n);
    m->pi_mb_quant_host = host_alloc<Int32>(n);
    m->pi_mb_mv_x_host = host_alloc<Int32>(n);
    m->pi_mb_mv_y_host = host_alloc<Int32>(n);
    /* the per-field device arrays of the reference are replaced by one
     * packed MbMode array (Backend::mode_dev) */
    m->pi_mb_types_device = m->pi_mb_flags_device = m->pi_mb_quant_device = nullptr;
    m->pi_mb_mv_x_device = m->pi_mb_mv_y_device = nullptr;

    m->ps_denoise_host = host_alloc<h261_dct_denoise_t>(1);
    m->ps_denoise_device = dev_alloc<h261_dct_denoise_t>(1);
    if (m->ps_denoise_host) m->ps_denoise_host->i_denoise = i_denoise;
    if (m->ps_denoise_device && m->ps_denoise_host)
        check(hipMemcpy(m->ps_denoise_device, m->ps_denoise_host, sizeof(h261_dct_denoise_t), hipMemcpyHostToDevice),
              "hipMemcpy");

    /* reconstruction: the reference sizes all three planes n*256 bytes */
    m->pui8_reconstructed_Y_device = dev_alloc<UInt8>((size_t)n * 256);
    m->pui8_reconstructed_Y_host = host_alloc<UInt8>((size_t)n * 256);
    m->pui8_reconstructed_Cb_device = dev_alloc<UInt8>((size_t)n * 256);
    m->pui8_reconstructed_Cb_host = host_alloc<UInt8>((size_t)n * 256);
    m->pui8_reconstructed_Cr_device = dev_alloc<UInt8>((size_t)n * 256);
    m->pui8_reconstructed_Cr_host = host_alloc<UInt8>((size_t)n * 256);

    /* the interleaved coefficient buffer of the reference does not exist:
     * coefficients never leave the kernel's registers and LDS */
    m->i_num_blocks = (n * 6 + 15) / 16;
    m->i_num_coeff_blocks = m->i_num_blocks * 16;
    m->pi_coeff_blocks_device = nullptr;
    m->pi_coeff_blocks_host = nullptr;

    Backend *b = new Backend();
    if (!backend_init(b, n)) {
        backend_free(b);
        return;
    }
    b->mode_host = host_alloc<MbMode>(n);
    b->mode_dev = dev_alloc<MbMode>(n);
    m->p_cuda = b;
    *pps = m;
}

/* src/gpu_device.cu:542-590 */
Void h261_deinit_macroblocks(h261_macroblocks_t *m)
{
    if (!m) return;
    hipFree(m->ps_macroblocks_device);
    hipHostFree(m->ps_macroblocks);
    hipHostFree(m->pi_mb_types_host);
    hipHostFree(m->pi_mb_flags_host);
    hipHostFree(m->pi_mb_quant_host);
    hipHostFree(m->pi_mb_mv_x_host);
    hipHostFree(m->pi_mb_mv_y_host);
    hipHostFree(m->ps_denoise_host);
    hipFree(m->ps_denoise_device);
    hipFree(m->pui8_reconstructed_Y_device);
    hipHostFree(m->pui8_reconstructed_Y_host);
    hipFree(m->pui8_reconstructed_Cb_device);
    hipHostFree(m->pui8_reconstructed_Cb_host);
    hipFree(m->pui8_reconstructed_Cr_device);
    hipHostFree(m->pui8_reconstructed_Cr_host);
    backend_free((Backend *)m->p_cuda);
    hipHostFree(m);
}

/* src/gpu_device.cu:592-729: pack loop, setup/DCT/denoise/quant -- here
 * fused with dequant/IDCT/reconstruction -- and the denoise update */
Void h261_gpu_device_encode_macroblocks(h261_context_t *ctx)
{
    if (!ctx || !ctx->ps_macroblocks || !ctx->ps_mode_decision || !ctx->ps_macroblocks->p_cuda || !g_planes.cur[0]) {
        fprintf(stderr, "h261gpu: encode_macroblocks: encoder not initialised\n");
        g_last_error = H261G_ERR_STATE;
        return;
    }
    h261_macroblocks_t *m = ctx->ps_macroblocks;
    h261_mode_decision_t *md = ctx->ps_mode_decision;
    Backend *b = (Backend *)m->p_cuda;
    if (g_planes.width != m->i_mb_width * 16 || g_planes.height != m->i_mb_height * 16) {
        fprintf(stderr, "h261gpu: encode_macroblocks: size differs from init_textures\n");
        g_last_error = H261G_ERR_STATE;
        return;
    }
    const int n = m->i_num_mb;
    const int picq = ctx->s_picture.i_quantiser;
    for (int y = 0; y < m->i_mb_height; y++)
        for (int x = 0; x < m->i_mb_width; x++) {
            const h261_mb_mode_decision_t *d = &md->prgs_mb[y * md->i_mb_width + x];
            const int i = y * m->i_mb_width + x;
            int q = picq + d->i_quantiser_adjust;
            q = q < 1 ? 1 : (q > 31 ? 31 : q);
            m->pi_mb_types_host[i] = d->i_mb_type;
            m->pi_mb_flags_host[i] = d->i_mb_flags;
            m->pi_mb_mv_x_host[i] = d->i_mv_x;
            m->pi_mb_mv_y_host[i] = d->i_mv_y;
            m->pi_mb_quant_host[i] = q;
            /* the reference fetches through wrap-mode textures; here a
             * vector must keep the 16x16 block inside the frame, as every
             * vector its motion estimation returns does */
            if (d->i_mb_type != H261_MB_TYPE_INTRA &&
                (x * 16 + d->i_mv_x < 0 || y * 16 + d->i_mv_y < 0 || x * 16 + d->i_mv_x + 16 > g_planes.width ||
                 y * 16 + d->i_mv_y + 16 > g_planes.height)) {
                fprintf(stderr, "h261gpu: encode_macroblocks: MB %d vector (%d,%d) leaves the frame\n", i,
                        d->i_mv_x, d->i_mv_y);
                g_last_error = H261G_ERR_ARG;
                return;
            }
            MbMode &o = b->mode_host[i];
            o.type = d->i_mb_type;
            o.flags = d->i_mb_flags;
            o.mvx = d->i_mv_x;
            o.mvy = d->i_mv_y;
            o.quant_adjust = d->i_quantiser_adjust;
            o.cost = d->i_cost;
        }
    StreamDesc *sd = b->desc_host;
    memset(sd, 0, sizeof(*sd));
    for (int i = 0; i < 3; i++) {
        sd->cur[i] = g_planes.cur[i];
        sd->ref[i] = g_planes.ref[i];
    }
    sd->rec[0] = m->pui8_reconstructed_Y_device;
    sd->rec[1] = m->pui8_reconstructed_Cb_device;
    sd->rec[2] = m->pui8_reconstructed_Cr_device;
    sd->width = g_planes.width;
    sd->height = g_planes.height;
    sd->mb_w = m->i_mb_width;
    sd->mb_h = m->i_mb_height;
    sd->pic_quant = picq;
    sd->flags = kStreamActive;
    if (!check(hipMemcpyAsync(b->desc_dev, sd, sizeof(StreamDesc), hipMemcpyHostToDevice, b->stream), "hipMemcpyAsync"))
        return;
    if (!check(hipMemcpyAsync(b->mode_dev, b->mode_host, sizeof(MbMode) * n, hipMemcpyHostToDevice, b->stream),
               "hipMemcpyAsync"))
        return;
    MbLaunch L{};
    L.streams = b->desc_dev;
    L.mb_stream = b->mb_stream_dev;
    L.n_mb = n;
    L.mode = b->mode_dev;
    L.denoise = (DenoiseState *)m->ps_denoise_device;
    L.mb_struct = m->ps_macroblocks_device;
    L.mbs_per_wave = 1;
    if (launch_mb(L, b->stream) != 0 || launch_denoise_update((DenoiseState *)m->ps_denoise_device, b->desc_dev, 1, b->stream) != 0) {
        check(hipGetLastError(), "mb_kernel launch");
        return;
    }
    check(hipMemcpyAsync(m->ps_macroblocks, m->ps_macroblocks_device, sizeof(h261_macroblock_t) * n,
                         hipMemcpyDeviceToHost, b->stream),
          "hipMemcpyAsync");
    check(hipStreamSynchronize(b->stream), "hipStreamSynchronize");
}

/* src/gpu_device.cu:732-815: the reconstruction was made by the fused
 * encode launch; copy it out */
Void h261_gpu_device_decode_macroblocks(h261_context_t *ctx)
{
    if (!ctx || !ctx->ps_macroblocks || !ctx->ps_macroblocks->p_cuda) {
        fprintf(stderr, "h261gpu: decode_macroblocks: encoder not initialised\n");
        g_last_error = H261G_ERR_STATE;
        return;
    }
    h261_macroblocks_t *m = ctx->ps_macroblocks;
    Backend *b = (Backend *)m->p_cuda;
    const size_t n = (size_t)m->i_num_mb;
    check(hipMemcpyAsync(m->pui8_reconstructed_Y_host, m->pui8_reconstructed_Y_device, n * 256,
                         hipMemcpyDeviceToHost, b->stream),
          "hipMemcpyAsync");
    check(hipMemcpyAsync(m->pui8_reconstructed_Cb_host, m->pui8_reconstructed_Cb_device, n * 64,
                         hipMemcpyDeviceToHost, b->stream),
          "hipMemcpyAsync");
    check(hipMemcpyAsync(m->pui8_reconstructed_Cr_host, m->pui8_reconstructed_Cr_device, n * 64,
                         hipMemcpyDeviceToHost, b->stream),
          "hipMemcpyAsync");
    check(hipStreamSynchronize(b->stream), "hipStreamSynchronize");
}

}  // extern "C"
