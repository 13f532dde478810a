/*
 * content.hip -- synthetic input frames (content.h) on the device.  The
 * kernels write whole frame sequences straight into HBM so that encoding
 * throughput is measured with inputs already resident: per scene, the
 * texture once into a 512x512 table, then the scene's frames from it.  The
 * same frames on the host: content_host.cpp.
 */
#include <hip/hip_runtime.h>

#include <stdlib.h>
#include <string.h>

#include "../../include/h261gpu.h"
#include "content.h"
#include "h261gpu_internal.h"

namespace h261g {

using content::FrameCtx;

namespace {

/* the texture of one seed over the 512x512 torus, a byte per pel
 * (content::texture, 16..235) */
__global__ __launch_bounds__(256) void texture_table_kernel(uint64_t seed, uint8_t *tab)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < content::kTextureTableBytes) tab[i] = (uint8_t)content::texture(seed, i & 511, i >> 9);
}

/* blockIdx.y = frame of the job (frames of one scene: their texture is
 * tab).  A thread takes a 4x2 block of luma pels and the two chroma pels
 * under it.  The block lies inside one macroblock, so its displacement is
 * computed once, and a chroma pel's texture sum is the sum of the
 * noise-free luma textures of its 2x2 pels: 8 table reads per thread serve
 * both planes, and only the per-pel noise is hashed. */
__global__ __launch_bounds__(256) void content_kernel(const FrameCtx *ctx, const uint8_t *tab, uint8_t *out, int width,
                                                      int height, size_t frame_stride)
{
    const FrameCtx f = ctx[blockIdx.y];
    uint8_t *y = out + frame_stride * blockIdx.y;
    uint8_t *cb = y + (size_t)width * height;
    uint8_t *cr = cb + (size_t)width * height / 4;
    const int bw = width / 4, i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= bw * (height / 2)) return;
    const int x0 = 4 * (i % bw), y0 = 2 * (i / bw);
    int dx, dy;
    content::displacement(f, x0, y0, &dx, &dy);
    int t[2][4];
#pragma unroll
    for (int r = 0; r < 2; r++)
#pragma unroll
        for (int k = 0; k < 4; k++) t[r][k] = tab[((y0 + r + dy) & 511) * 512 + ((x0 + k + dx) & 511)];
#pragma unroll
    for (int r = 0; r < 2; r++) {
        uint32_t w = 0;
#pragma unroll
        for (int k = 0; k < 4; k++)
            w |= (uint32_t)content::clamp255(t[r][k] + content::noise(f.noise_seed, f.frame, x0 + k, y0 + r, f.sigma))
                 << (8 * k);
        *(uint32_t *)(y + (size_t)(y0 + r) * width + x0) = w;
    }
    uint32_t wb = 0, wr = 0;
#pragma unroll
    for (int c = 0; c < 2; c++) {
        const int sum = t[0][2 * c] + t[0][2 * c + 1] + t[1][2 * c] + t[1][2 * c + 1];
        wb |= (uint32_t)content::clamp255(128 + (sum - 504) / 4) << (8 * c);
        wr |= (uint32_t)content::clamp255(128 - (sum - 504) / 8) << (8 * c);
    }
    const size_t co = (size_t)(y0 / 2) * (width / 2) + x0 / 2;
    *(uint16_t *)(cb + co) = (uint16_t)wb;
    *(uint16_t *)(cr + co) = (uint16_t)wr;
}

/* FrameCtx of frames first .. first + n - 1, one workgroup: the same
 * values as content_frame_ctx.  Global motion accumulates the per-frame
 * steps of frames 1..t: a running workgroup scan over the frames in chunks
 * of kCtxThreads, from frame 1. */
constexpr int kCtxThreads = 256;
__global__ __launch_bounds__(kCtxThreads) void content_ctx_kernel(h261g_content_params_t c, int width, int height,
                                                                  int first, int n, FrameCtx *ctx)
{
    __shared__ int sx[kCtxThreads], sy[kCtxThreads];
    const int tid = threadIdx.x;
    int carry_x = 0, carry_y = 0; /* displacement through the previous chunk */
    for (int t0 = 0; t0 < first + n; t0 += kCtxThreads) {
        const int t = t0 + tid;
        int dx = 0, dy = 0;
        if (!c.local && t >= 1 && t < first + n) {
            dx = content::global_step(c.seed, t, 0, c.motion);
            dy = content::global_step(c.seed, t, 1, c.motion);
        }
        /* inclusive scan over the chunk */
        sx[tid] = dx;
        sy[tid] = dy;
        __syncthreads();
        for (int off = 1; off < kCtxThreads; off <<= 1) {
            const int ax = tid >= off ? sx[tid - off] : 0, ay = tid >= off ? sy[tid - off] : 0;
            __syncthreads();
            sx[tid] += ax;
            sy[tid] += ay;
            __syncthreads();
        }
        if (t >= first && t < first + n) {
            FrameCtx f;
            f.noise_seed = c.seed;
            f.frame = t;
            f.sigma = c.sigma;
            f.local = c.local;
            f.width = width;
            f.height = height;
            f.ox = f.oy = 0;
            f.age = 0;
            if (c.local) {
                const int scene = c.scene_cut > 0 ? t / c.scene_cut : 0;
                f.tex_seed = content::scene_seed(c.seed, scene);
                f.age = c.scene_cut > 0 ? t - scene * c.scene_cut : t;
            } else {
                f.tex_seed = c.seed;
                f.ox = carry_x + sx[tid];
                f.oy = carry_y + sy[tid];
            }
            ctx[t - first] = f;
        }
        carry_x += sx[kCtxThreads - 1];
        carry_y += sy[kCtxThreads - 1];
        __syncthreads();
    }
}

}  // namespace

/* frames [first, first + n) of a stream into out (I420, frame after frame);
 * tab: content::kTextureTableBytes of device memory, used on this stream only */
int generate_frames_device(const h261g_content_params_t &c, int width, int height, int first, int n, uint8_t *out,
                           FrameCtx *ctx, uint8_t *tab, void *stream)
{
    if (n <= 0) return 0;
    /* per-frame parameters on the device, then scene by scene the texture
     * table and the scene's frames: nothing for the host to wait for */
    hipLaunchKernelGGL(content_ctx_kernel, dim3(1), dim3(kCtxThreads), 0, (hipStream_t)stream, c, width, height, first,
                       n, ctx);
    if (hipGetLastError() != hipSuccess) return H261G_ERR_DEVICE;
    const int threads = (width / 4) * (height / 2);
    const size_t fsz = (size_t)width * height * 3 / 2;
    for (int a = first; a < first + n;) {
        /* the texture seed of frame a's scene, as content_ctx_kernel sets it */
        int b = first + n;
        uint64_t seed = c.seed;
        if (c.local) {
            const int scene = c.scene_cut > 0 ? a / c.scene_cut : 0;
            if (c.scene_cut > 0) b = min(b, (scene + 1) * c.scene_cut);
            seed = content::scene_seed(c.seed, scene);
        }
        hipLaunchKernelGGL(texture_table_kernel, dim3(content::kTextureTableBytes / 256), dim3(256), 0,
                           (hipStream_t)stream, seed, tab);
        hipLaunchKernelGGL(content_kernel, dim3((threads + 255) / 256, b - a), dim3(256), 0, (hipStream_t)stream,
                           ctx + (a - first), tab, out + (size_t)(a - first) * fsz, width, height, fsz);
        if (hipGetLastError() != hipSuccess) return H261G_ERR_DEVICE;
        a = b;
    }
    return 0;
}

}  // namespace h261g
