/*
 * content.h -- deterministic synthetic video, integer arithmetic only, the
 * same function on the host and on the device.
 *
 * Texture: value noise on a 512x512 torus, 4 octaves with lattice cells of
 * 32/16/8/4 pels and amplitudes 8/4/2/1, smoothstep (3w^2 - 2w^3, Q16)
 * bilinear interpolation of 8-bit lattice values, scaled to 16..235.
 * Luma pel (x, y) of frame t = texture(x + O_t.x, y + O_t.y) + noise,
 * clamped to 0..255, where O_t is the frame's displacement:
 *   global mode: O_t = sum of per-frame vectors drawn from [-M, M]^2;
 *   local mode:  O_t(mb) = (t - t_scene) * v(mb), v an affine field
 *                (rotation <= 2 deg, scale 0.98..1.02, translation <= 12)
 *                rounded per macroblock, 10% of macroblocks replaced by a
 *                random vector in [-15, 15]^2; a scene cut every
 *                scene_cut frames draws a new texture and field.
 * Noise: sigma * (sum of 12 uniform 16-bit values - 6*65536) / 65536,
 * rounded half away from zero (Irwin-Hall approximation of N(0, sigma^2)).
 * Chroma (x, y) from the noise-free luma texture sum S over the 2x2 pels
 * (2x, 2y) .. (2x+1, 2y+1): Cb = 128 + (S - 504) / 4, Cr = 128 - (S - 504) / 8,
 * clamped to 0..255.
 */
#ifndef H261G_CONTENT_H
#define H261G_CONTENT_H

#include <stdint.h>

#ifdef __HIPCC__
#define H261G_CHD __host__ __device__ __forceinline__
#else
#define H261G_CHD inline
#endif

namespace h261g {
namespace content {

H261G_CHD uint64_t mix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

H261G_CHD uint32_t lattice(uint64_t seed, int oct, int i, int j)
{
    uint64_t k = seed * 0x100000001B3ull + ((uint64_t)oct << 48) + ((uint64_t)(uint16_t)i << 24) + (uint16_t)j;
    return (uint32_t)(mix64(k) & 255u);
}

H261G_CHD int64_t smooth_q16(int f, int shift)
{
    int64_t w = (int64_t)f << (16 - shift);
    return ((w * w) >> 16) * (3 * 65536 - 2 * w) >> 16;
}

/* noise-free texture value 16..235 at torus position (x, y) */
H261G_CHD int texture(uint64_t seed, int x, int y)
{
    x &= 511;
    y &= 511;
    int64_t sum = 0;
    for (int o = 0; o < 4; o++) {
        const int shift = 5 - o; /* cell 32 >> o */
        const int L = 1 << shift, n = 512 >> shift;
        const int ci = x >> shift, cj = y >> shift, fx = x & (L - 1), fy = y & (L - 1);
        const int ci1 = (ci + 1) & (n - 1), cj1 = (cj + 1) & (n - 1);
        const int64_t sx = smooth_q16(fx, shift), sy = smooth_q16(fy, shift);
        const int64_t v00 = lattice(seed, o, ci, cj), v10 = lattice(seed, o, ci1, cj);
        const int64_t v01 = lattice(seed, o, ci, cj1), v11 = lattice(seed, o, ci1, cj1);
        const int64_t top = v00 * (65536 - sx) + v10 * sx;
        const int64_t bot = v01 * (65536 - sx) + v11 * sx;
        const int64_t val = (top * (65536 - sy) + bot * sy) >> 16;
        sum += val << (3 - o); /* amplitudes 8, 4, 2, 1 */
    }
    return 16 + (int)((sum * 219) / (15ll * 255 * 65536));
}

H261G_CHD int noise(uint64_t seed, int frame, int x, int y, int sigma)
{
    if (sigma <= 0) return 0;
    uint64_t k = mix64(seed ^ 0x6E6F697365ull) + (uint64_t)frame * 0x9E3779B1ull + ((uint64_t)y << 20) + (uint64_t)x;
    int64_t s = 0;
    for (int i = 0; i < 3; i++) {
        uint64_t h = mix64(k * 4 + i);
        s += (int64_t)(h & 0xffff) + (int64_t)((h >> 16) & 0xffff) + (int64_t)((h >> 32) & 0xffff) +
             (int64_t)((h >> 48) & 0xffff);
    }
    int64_t n = (s - 6 * 65536) * sigma;
    return (int)(n >= 0 ? (n + 32768) / 65536 : -((-n + 32768) / 65536));
}

H261G_CHD int clamp255(int v) { return v < 0 ? 0 : (v > 255 ? 255 : v); }

/* per-frame global vector component in [-m, m] */
H261G_CHD int global_step(uint64_t seed, int frame, int axis, int m)
{
    if (m <= 0) return 0;
    return (int)(mix64(seed ^ 0x676C6F62ull ^ ((uint64_t)frame << 8) ^ (uint64_t)axis) % (uint64_t)(2 * m + 1)) - m;
}

/* local mode: the scene's texture seed */
H261G_CHD uint64_t scene_seed(uint64_t seed, int scene) { return mix64(seed ^ 0x7363656E65ull ^ ((uint64_t)scene << 32)); }

/* local mode: vector of macroblock (i, j) in a scene, Q16 affine field
 * around the frame centre, 10% of macroblocks random */
H261G_CHD void local_vector(uint64_t sseed, int i, int j, int width, int height, int *vx, int *vy)
{
    const uint64_t h0 = mix64(sseed ^ 0x616666ull);
    /* rotation in [-2, 2] degrees as Q16 sin/cos from a small-angle series
     * (|angle| <= 0.0349 rad): sin ~ a - a^3/6, cos ~ 1 - a^2/2 */
    const int64_t a = ((int64_t)(h0 & 0xffff) - 32768) * 2287 / 32768; /* Q16 rad, 2287 = 0.0349*65536 */
    const int64_t sn = a - (a * a / 65536) * a / (6 * 65536);
    const int64_t cs = 65536 - (a * a) / (2 * 65536);
    const int64_t sc = 65536 + (((int64_t)((h0 >> 16) & 0xffff) - 32768) * 1311) / 32768; /* 0.98..1.02 */
    const int tx = (int)((h0 >> 32) % 25) - 12, ty = (int)((h0 >> 40) % 25) - 12;
    const uint64_t hm = mix64(sseed ^ ((uint64_t)i << 16) ^ (uint64_t)j ^ 0x6D62ull);
    if ((hm % 10) == 0) {
        *vx = (int)((hm >> 8) % 31) - 15;
        *vy = (int)((hm >> 16) % 31) - 15;
        return;
    }
    const int64_t px = (int64_t)(i * 16 + 8 - width / 2), py = (int64_t)(j * 16 + 8 - height / 2);
    const int64_t ax = (sc * ((cs * px - sn * py) / 65536)) / 65536;
    const int64_t ay = (sc * ((sn * px + cs * py) / 65536)) / 65536;
    int dx = (int)(ax - px) + tx, dy = (int)(ay - py) + ty;
    *vx = dx < -15 ? -15 : (dx > 15 ? 15 : dx);
    *vy = dy < -15 ? -15 : (dy > 15 ? 15 : dy);
}

/* the noise-free texture of one seed over the whole 512x512 torus, one byte
 * per pel (values 16..235): the device generator fills it once per scene
 * and every frame of the scene reads its pels from it */
constexpr int kTextureTableBytes = 512 * 512;

struct FrameCtx {
    uint64_t tex_seed; /* texture of this frame (scene seed in local mode) */
    uint64_t noise_seed;
    int frame;
    int sigma;
    int local;
    int ox, oy;   /* global displacement */
    int age;      /* frames since the scene start (local mode) */
    int width, height;
};

/* luma texture displacement of pel (x, y) */
H261G_CHD void displacement(const FrameCtx &f, int x, int y, int *dx, int *dy)
{
    if (!f.local) {
        *dx = f.ox;
        *dy = f.oy;
        return;
    }
    int vx, vy;
    local_vector(f.tex_seed, x >> 4, y >> 4, f.width, f.height, &vx, &vy);
    *dx = vx * f.age;
    *dy = vy * f.age;
}

H261G_CHD uint8_t luma(const FrameCtx &f, int x, int y)
{
    int dx, dy;
    displacement(f, x, y, &dx, &dy);
    return (uint8_t)clamp255(texture(f.tex_seed, x + dx, y + dy) + noise(f.noise_seed, f.frame, x, y, f.sigma));
}

H261G_CHD void chroma(const FrameCtx &f, int x, int y, uint8_t *cb, uint8_t *cr)
{
    int dx, dy;
    displacement(f, 2 * x, 2 * y, &dx, &dy);
    int X = 2 * x + dx, Y = 2 * y + dy;
    int s = texture(f.tex_seed, X, Y) + texture(f.tex_seed, X + 1, Y) + texture(f.tex_seed, X, Y + 1) +
            texture(f.tex_seed, X + 1, Y + 1);
    *cb = (uint8_t)clamp255(128 + (s - 504) / 4);
    *cr = (uint8_t)clamp255(128 - (s - 504) / 8);
}

}  // namespace content
}  // namespace h261g

#endif
