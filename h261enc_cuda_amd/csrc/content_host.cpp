/*
 * content_host.cpp -- synthetic input frames (content.h) on the host: the
 * same pels the device generator (content.hip) writes into HBM.
 *
 * Plain C++ with no HIP dependency.  It is linked into libh261gpu.so
 * (h261g_content_frame_host of include/h261gpu.h) and also built alone as
 * libh261content.so, so that processes which only need input frames (CPU
 * checkers, file writers) do not load the HIP runtime.
 */
#include "../../include/h261gpu.h"
#include "content.h"

namespace h261g {

/* frame context of frame t of a stream: displacement sums for the global
 * mode, scene seed and age for the local mode */
content::FrameCtx content_frame_ctx(const h261g_content_params_t &c, int width, int height, int t)
{
    content::FrameCtx f;
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
        for (int k = 1; k <= t; k++) {
            f.ox += content::global_step(c.seed, k, 0, c.motion);
            f.oy += content::global_step(c.seed, k, 1, c.motion);
        }
    }
    return f;
}

}  // namespace h261g

extern "C" int h261g_content_frame_host(const h261g_content_params_t *c, int32_t width, int32_t height, int32_t frame,
                                        uint8_t *y, uint8_t *cb, uint8_t *cr)
{
    if (!c || !y || !cb || !cr || width <= 0 || height <= 0 || width % 16 || height % 16 || frame < 0)
        return H261G_ERR_ARG;
    const h261g::content::FrameCtx f = h261g::content_frame_ctx(*c, width, height, frame);
    for (int py = 0; py < height; py++)
        for (int px = 0; px < width; px++) y[py * width + px] = h261g::content::luma(f, px, py);
    for (int py = 0; py < height / 2; py++)
        for (int px = 0; px < width / 2; px++)
            h261g::content::chroma(f, px, py, &cb[py * (width / 2) + px], &cr[py * (width / 2) + px]);
    return H261G_OK;
}
