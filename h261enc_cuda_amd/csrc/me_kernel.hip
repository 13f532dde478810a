/*
 * me_kernel.hip -- motion estimation for gfx950.
 *
 * One wave64 per 16x16 macroblock:
 *   1. the reference search window (rows by-15..by+30, cols bx-16..bx+31)
 *      is staged in LDS with aligned dword loads, the current MB next to it;
 *   2. stage 1 evaluates the 11x11 coarse grid
 *      {-14,-11,-8,-5,-2,0,2,5,8,11,14}^2 (src/megpu_kernels.cuh:29-32),
 *      stage 2 the 7x7 refine around the winner (megpu:309-312, 338-339),
 *      each candidate one lane, 16 rows x 4 packed v_sad_u8 (alignbyte
 *      realigns the window bytes to the candidate's column);
 *   3. the winner is the wave minimum of (SAD << 8 | morton(tx, ty)):
 *      the reference's quad-tree reduction with strict '<' in the order
 *      self, +x, +y, +xy (megpu:95-154, 375-434) selects exactly the
 *      lexicographic minimum of (cost, Z-order index), so a butterfly
 *      min over that key reproduces its tie-break;
 *   4. 48 lanes compute the 4x4 Hadamard SATDs of the inter, inter+filter
 *      and intra costs (megpu:617-987), summed per MB (the reference's
 *      racy shared-memory sum is meant as this plain sum).
 * Out-of-limit candidates cost 65535 (megpu:67-92); limits keep every
 * fetch inside the frame (gpu_device.cu:226-265).
 */
#include <hip/hip_runtime.h>

#include "frame_control.h"
#include "h261gpu_internal.h"

namespace h261g {

namespace {

constexpr int kWinRows = 46;  /* 16 + 2*15 */
constexpr int kWinDw = 12;    /* 48 bytes per row: frame cols bx-16 .. bx+31 */

__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        uint32_t o = (uint32_t)__shfl_xor((int)v, off, 64);
        v = o < v ? o : v;
    }
    return v;
}

__device__ __forceinline__ int group16_sum(int v)
{
#pragma unroll
    for (int off = 8; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

/* Z-order index with x in the even bits: the reference reduction order */
__device__ __forceinline__ uint32_t morton(uint32_t x, uint32_t y)
{
    uint32_t k = 0;
#pragma unroll
    for (int b = 0; b < 4; b++) k |= (((x >> b) & 1u) << (2 * b)) | (((y >> b) & 1u) << (2 * b + 1));
    return k;
}

__device__ __forceinline__ void unmorton(uint32_t k, int &x, int &y)
{
    x = 0;
    y = 0;
#pragma unroll
    for (int b = 0; b < 4; b++) {
        x |= (int)((k >> (2 * b)) & 1u) << b;
        y |= (int)((k >> (2 * b + 1)) & 1u) << b;
    }
}

__device__ __forceinline__ int coarse_vec(int i)
{
    /* {-14,-11,-8,-5,-2,0,2,5,8,11,14} */
    return i < 5 ? 3 * i - 14 : (i == 5 ? 0 : 3 * i - 16);
}

/* 16x16 SAD of the candidate (dx, dy) against the staged window */
__device__ __forceinline__ uint32_t sad16(const uint32_t *win, const uint32_t *cur, int dx, int dy)
{
    int c0 = dx + 16;
    int q = c0 >> 2, sh = c0 & 3;
    const uint32_t *row = win + (dy + 15) * kWinDw + q;
    uint32_t acc = 0;
#pragma unroll 4
    for (int k = 0; k < 16; k++) {
        uint32_t w0 = row[0], w1 = row[1], w2 = row[2], w3 = row[3], w4 = row[4];
        uint32_t r0 = __builtin_amdgcn_alignbyte(w1, w0, sh);
        uint32_t r1 = __builtin_amdgcn_alignbyte(w2, w1, sh);
        uint32_t r2 = __builtin_amdgcn_alignbyte(w3, w2, sh);
        uint32_t r3 = __builtin_amdgcn_alignbyte(w4, w3, sh);
        acc = __builtin_amdgcn_sad_u8(r0, cur[4 * k + 0], acc);
        acc = __builtin_amdgcn_sad_u8(r1, cur[4 * k + 1], acc);
        acc = __builtin_amdgcn_sad_u8(r2, cur[4 * k + 2], acc);
        acc = __builtin_amdgcn_sad_u8(r3, cur[4 * k + 3], acc);
        row += kWinDw;
    }
    return acc;
}

/* megpu_satd_4x4_kernel (src/megpu_kernels.cuh:467-500) */
__device__ __forceinline__ int satd4x4(const int d[16])
{
    int t[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        int a1 = d[4 * i] + d[4 * i + 1], a2 = d[4 * i] - d[4 * i + 1];
        int a3 = d[4 * i + 2] + d[4 * i + 3], a4 = d[4 * i + 2] - d[4 * i + 3];
        t[i][0] = a1 + a3;
        t[i][1] = a2 + a4;
        t[i][2] = a1 - a3;
        t[i][3] = a2 - a4;
    }
    int s = 0;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        int a1 = t[0][i] + t[1][i], a2 = t[0][i] - t[1][i];
        int a3 = t[2][i] + t[3][i], a4 = t[2][i] - t[3][i];
        s += abs(a1 + a3) + abs(a2 + a4) + abs(a1 - a3) + abs(a2 - a4);
    }
    return s;
}

/* the full SAD surface of one MB for dy, dx in [-15, 15] (dx = -16 is a
 * padding column), surf[(dy + 15) * 32 + dx + 16].  Lane (dy, half) covers
 * 16 horizontal offsets: window dwords g0..g0+7 of each row (g0 = 4 half,
 * 16-byte aligned loads) against the 4 dwords of the current row, one
 * v_qsad_pk_u16_u8 per 4-byte column and group of 4 offsets. */
__device__ __forceinline__ void sad_surface(const uint32_t *win, const uint32_t *curw, uint16_t *surf, int lane)
{
    if (lane >= 62) return;
    const int dyi = lane >> 1, half = lane & 1;
    const uint32_t *row = win + dyi * kWinDw + 4 * half;
    uint64_t acc0 = 0, acc1 = 0, acc2 = 0, acc3 = 0;
#pragma unroll 2
    for (int k = 0; k < 16; k++) {
        const uint4 a = *(const uint4 *)(row + k * kWinDw);
        const uint4 b = *(const uint4 *)(row + k * kWinDw + 4);
        const uint4 c = *(const uint4 *)(curw + 4 * k);
        const uint32_t w[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
        const uint32_t cq[4] = {c.x, c.y, c.z, c.w};
#pragma unroll
        for (int q = 0; q < 4; q++) {
            acc0 = __builtin_amdgcn_qsad_pk_u16_u8(((uint64_t)w[q + 1] << 32) | w[q], cq[q], acc0);
            acc1 = __builtin_amdgcn_qsad_pk_u16_u8(((uint64_t)w[q + 2] << 32) | w[q + 1], cq[q], acc1);
            acc2 = __builtin_amdgcn_qsad_pk_u16_u8(((uint64_t)w[q + 3] << 32) | w[q + 2], cq[q], acc2);
            acc3 = __builtin_amdgcn_qsad_pk_u16_u8(((uint64_t)w[q + 4] << 32) | w[q + 3], cq[q], acc3);
        }
    }
    uint4 *out = (uint4 *)(surf + dyi * 32 + 16 * half);
    out[0] = make_uint4((uint32_t)acc0, (uint32_t)(acc0 >> 32), (uint32_t)acc1, (uint32_t)(acc1 >> 32));
    out[1] = make_uint4((uint32_t)acc2, (uint32_t)(acc2 >> 32), (uint32_t)acc3, (uint32_t)(acc3 >> 32));
}

template <int kVariant>
__global__ __launch_bounds__(64) void me_kernel(MeLaunch p)
{
    __shared__ __align__(16) uint32_t win[kWinRows * kWinDw];
    __shared__ __align__(16) uint32_t curw[64];
    __shared__ __align__(16) uint16_t surf[kVariant == kMeSurface ? 31 * 32 : 8];

    const int g = launch_index(blockIdx.x, p.n_mb, p.xcd_remap);
    if (g >= p.n_mb) return;
    const int lane = threadIdx.x;
    const int s = p.mb_stream[g];
    if (s < 0) return; /* unmapped range of a closed stream */
    const StreamDesc &d = p.streams[s];
    if (!(d.flags & kStreamActive)) return;
    const int m = g - d.mb_base;
    const int W = d.width, H = d.height;
    const int bx = (m % d.mb_w) * 16, by = (m / d.mb_w) * 16;
    const uint8_t *cur = d.cur[0];
    const uint8_t *ref = d.ref[0];

    /* motion vector limits (gpu_device.cu:226-265) */
    const int lx0 = max(bx - 15, 0) - bx, lx1 = min(bx + 15, (d.mb_w - 1) * 16) - bx;
    const int ly0 = max(by - 15, 0) - by, ly1 = min(by + 15, (d.mb_h - 1) * 16) - by;

    /* stage the window: all loads issued before any wait.  Window dwords
     * outside the frame are loaded from a clamped in-frame address; no
     * in-limit candidate or filter tap ever reads them. */
    constexpr int kWinWords = kWinRows * kWinDw; /* 552 */
    constexpr int kPerLane = (kWinWords + 63) / 64;
    uint32_t staged[kPerLane];
#pragma unroll
    for (int j = 0; j < kPerLane; j++) {
        const int i = lane + 64 * j;
        const int r = i / kWinDw, c4 = i - r * kWinDw;
        const int fy = min(max(by - 15 + r, 0), H - 1);
        const int fx = min(max(bx - 16 + 4 * c4, 0), W - 4);
        staged[j] = *(const uint32_t *)(ref + (size_t)fy * W + fx);
    }
    const uint32_t cw = *(const uint32_t *)(cur + (size_t)(by + (lane >> 2)) * W + bx + 4 * (lane & 3));
#pragma unroll
    for (int j = 0; j < kPerLane; j++) {
        const int i = lane + 64 * j;
        if (i < kWinWords) win[i] = staged[j];
    }
    curw[lane] = cw;
    __syncthreads();

    const bool search = !(p.skip_search_on_intra && d.frame_type == H261_FRAME_TYPE_INTRA);
    int mvx = 0, mvy = 0;
    if (search) {
        if (kVariant == kMeSurface) {
            sad_surface(win, curw, surf, lane);
            __syncthreads();
        }
        auto cost_of = [&](int vx, int vy) -> uint32_t {
            if (!(vx >= lx0 && vx <= lx1 && vy >= ly0 && vy <= ly1)) return 65535u;
            return kVariant == kMeSurface ? (uint32_t)surf[(vy + 15) * 32 + vx + 16] : sad16(win, curw, vx, vy);
        };
        /* stage 1: coarse grid */
        uint32_t best = 0xffffffffu;
        for (int ci = lane; ci < 121; ci += 64) {
            int tx = ci % 11, ty = ci / 11;
            int vx = coarse_vec(tx), vy = coarse_vec(ty);
            uint32_t cost = cost_of(vx, vy);
            uint32_t key = (cost << 8) | morton((uint32_t)tx, (uint32_t)ty);
            best = key < best ? key : best;
        }
        best = wave_min_u32(best);
        int wtx, wty;
        unmorton(best & 0xffu, wtx, wty);
        const int sx = coarse_vec(wtx), sy = coarse_vec(wty);

        /* stage 2: 7x7 refine around the coarse winner */
        best = 0xffffffffu;
        if (lane < 49) {
            int tx = lane % 7, ty = lane / 7;
            int vx = sx + tx - 3, vy = sy + ty - 3;
            uint32_t cost = cost_of(vx, vy);
            best = (cost << 8) | morton((uint32_t)tx, (uint32_t)ty);
        }
        best = wave_min_u32(best);
        unmorton(best & 0xffu, wtx, wty);
        mvx = sx + wtx - 3;
        mvy = sy + wty - 3;
    }

    /* SATD costs: lanes 0-15 inter, 16-31 inter+filter, 32-47 intra */
    int satd = 0;
    const int kind = lane >> 4;
    if (lane < 48 && (search || kind == 2)) {
        const int sub = lane & 15, sub_x = sub & 3, sub_y = sub >> 2;
        const uint8_t *cb = (const uint8_t *)curw;
        const uint8_t *wb = (const uint8_t *)win;
        int dd[16];
        /* window byte of MB pel (x, y) displaced by the mv */
        const int wr0 = mvy + 15 + 4 * sub_y, wc0 = mvx + 16 + 4 * sub_x;
#pragma unroll
        for (int y = 0; y < 4; y++)
#pragma unroll
            for (int x = 0; x < 4; x++) {
                int c = cb[(4 * sub_y + y) * 16 + 4 * sub_x + x];
                dd[4 * y + x] = kind == 2 ? c : c - (int)wb[(wr0 + y) * 48 + wc0 + x];
            }
        if (kind == 1) {
            /* 3x3 [1 2 1] filter on the 6x6 interior of each 8x8 block
             * (megpu:833-879); border pels stay unfiltered */
            const int fx0 = (sub_x & 1) ? 0 : 1, fx1 = (sub_x & 1) ? 3 : 4;
            const int fy0 = (sub_y & 1) ? 0 : 1, fy1 = (sub_y & 1) ? 3 : 4;
            for (int y = fy0; y < fy1; y++)
                for (int x = fx0; x < fx1; x++) {
                    const uint8_t *r = wb + (wr0 + y) * 48 + wc0 + x;
                    int f = r[-49] + 2 * r[-48] + r[-47] + 2 * r[-1] + 4 * r[0] + 2 * r[1] + r[47] +
                            2 * r[48] + r[49];
                    f = (f + 8) >> 4;
                    dd[4 * y + x] = (int)cb[(4 * sub_y + y) * 16 + 4 * sub_x + x] - f;
                }
        }
        satd = satd4x4(dd);
    }
    satd = group16_sum(satd);
    const int c_inter = __shfl(satd, 0, 64), c_filter = __shfl(satd, 16, 64), c_intra = __shfl(satd, 32, 64);

    if (lane == 0) {
        MeResult r;
        r.intra_cost = c_intra;
        r.zero_cost = 0;
        r.inter_cost = c_inter;
        r.filter_cost = c_filter;
        r.mvx = (int8_t)mvx;
        r.mvy = (int8_t)mvy;
        r.pad = 0;
        p.result[g] = r;
        if (p.mv_start) {
            p.mv_start[2 * g] = (int8_t)mvx;
            p.mv_start[2 * g + 1] = (int8_t)mvy;
        }
        if (p.mode) {
            /* h261_mode_decision (src/h261_moded.c:31-100) */
            MbMode md;
            md.quant_adjust = 0;
            bool intra = d.frame_type == H261_FRAME_TYPE_INTRA ||
                         (c_intra <= c_inter && c_intra <= c_filter);
            if (intra) {
                md.type = H261_MB_TYPE_INTRA;
                md.cost = c_intra;
                md.flags = 0;
                md.mvx = md.mvy = 0;
            } else {
                md.type = H261_MB_TYPE_INTER;
                if (c_filter < c_inter) {
                    md.flags = H261_MB_MC | H261_MB_FILTER;
                    md.cost = c_filter;
                } else {
                    md.flags = 0;
                    md.cost = c_inter;
                }
                md.mvx = mvx;
                md.mvy = mvy;
                if (mvx != 0 || mvy != 0) md.flags |= H261_MB_MC;
            }
            p.mode[g] = md;
            if (p.rc) {
                if (p.fuse_tail) {
                    /* count and cost sum in one atomic: the MB that completes
                     * the count holds the frame's cost sum and starts the
                     * picture (rate control, quantiser) */
                    RcState &rc = p.rc[s];
                    const unsigned long long old =
                        atomicAdd((unsigned long long *)&rc.me_acc, (1ull << 32) | (uint32_t)md.cost);
                    if ((int)(old >> 32) == d.mb_w * d.mb_h - 1) {
                        fc::rc_begin_one(rc, d, (int)(uint32_t)old + md.cost);
                        atomicExch((unsigned long long *)&rc.me_acc, 0ull);
                    }
                } else {
                    atomicAdd(&p.rc[s].cost_accum, md.cost);
                }
            }
        }
    }
}

}  // namespace

int launch_me(const MeLaunch &p, void *stream)
{
    if (p.n_mb <= 0) return 0;
    if (p.variant == kMeSurface)
        hipLaunchKernelGGL(me_kernel<kMeSurface>, dim3(launch_grid(p.n_mb, p.xcd_remap)), dim3(64), 0, (hipStream_t)stream, p);
    else
        hipLaunchKernelGGL(me_kernel<kMeTwoStage>, dim3(launch_grid(p.n_mb, p.xcd_remap)), dim3(64), 0, (hipStream_t)stream, p);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace h261g
