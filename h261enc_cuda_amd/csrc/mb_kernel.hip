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
 *   lanes 0-47 = (block b, col c):   DCT columns, denoise, quantiser, CBP
 *   lanes 0-47 = (block b, row r):   every block's prediction goes to the
 *                                    reconstruction plane: an uncoded
 *                                    block's reconstruction is its
 *                                    prediction (all-zero levels give a
 *                                    zero IDCT, mbgpu:966-976)
 * and then, for the coded blocks only, packed 8 to a pass:
 *   lane (slot, col c):              dequantiser, IDCT columns
 *   LDS transpose
 *   lane (slot, row r):              IDCT rows, the prediction row read
 *                                    back, reconstruction
 *
 * Arithmetic follows the reference op for op (compiled with
 * -ffp-contract=off and '#pragma clang fp contract(off)'): products with an
 * unsuffixed literal are double multiplies rounded to float; the forward
 * DCT rounds with floorf((float)(x + 0.49999)) (computed as the FP32 sum
 * x + 0.49999f, the same for every float x in range), the IDCT with
 * floor((double)x + 0.49999) (src/mbgpu_kernels.cuh:337-442, 827-931).
 * A wave codes groups of consecutive MBs; per group, each MB's stream,
 * position, quantiser, mode and plane pointers, and the denoise offset
 * tables of the group's streams, are read once into LDS.  The denoise uses
 * the previous frame's offsets and accumulates this frame's |coefficient|
 * statistics (LDS atomics, one global flush per group); the CBP is the OR
 * of this frame's coded blocks, the intended result of the reference's racy
 * reset/atomicOr (mbgpu:676-721).
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

/* The plane pointers come from stream descriptors in memory, so the compiler
 * sees generic pointers and would issue flat loads and stores.  A flat
 * access also counts against the LDS wait counter: every wait for an LDS
 * result would then wait for the outstanding plane accesses as well.  The
 * plane accesses name the global address space instead. */
template <class T>
__device__ __forceinline__ T load_global(const uint8_t *p)
{
    return *(const __attribute__((address_space(1))) T *)(p);
}
/* two words through a plain vector type: uint2's member operators take
 * their operands as generic pointers and would bring the flat access back */
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void load_global2(const uint8_t *p, uint32_t &w0, uint32_t &w1)
{
    const u32x2 v = *(const __attribute__((address_space(1))) u32x2 *)(p);
    w0 = v.x;
    w1 = v.y;
}
__device__ __forceinline__ void store_global2(uint8_t *p, uint32_t w0, uint32_t w1)
{
    u32x2 v;
    v.x = w0;
    v.y = w1;
    *(__attribute__((address_space(1))) u32x2 *)(p) = v;
}

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

/* sum over the wave without LDS round trips: quads, half rows and rows by
 * DPP, the four rows by readlane */
__device__ __forceinline__ int wave_sum_i32(int x)
{
    /* "old" = 0, the sum's identity: each move folds into the add */
    x += __builtin_amdgcn_update_dpp(0, x, 0xB1, 0xf, 0xf, false);  /* quad_perm [1,0,3,2] */
    x += __builtin_amdgcn_update_dpp(0, x, 0x4E, 0xf, 0xf, false);  /* quad_perm [2,3,0,1] */
    x += __builtin_amdgcn_update_dpp(0, x, 0x141, 0xf, 0xf, false); /* row_half_mirror */
    x += __builtin_amdgcn_update_dpp(0, x, 0x140, 0xf, 0xf, false); /* row_mirror */
    return __builtin_amdgcn_readlane(x, 0) + __builtin_amdgcn_readlane(x, 16) + __builtin_amdgcn_readlane(x, 32) +
           __builtin_amdgcn_readlane(x, 48);
}

/* bits of one TCOEFF event: the code and its sign bit, or an escape; the
 * code lengths plus one (kEscapeBits where the table has no code) are read
 * from constant memory (only MBs with coded blocks count bits, and an LDS
 * copy would cost the coder a wave per SIMD) */
__device__ __forceinline__ int tcoeff_bits(int run, int level)
{
    const int a = level < 0 ? -level : level;
    const uint8_t *tlen = (const uint8_t *)c_tlen.w;
    return (run < kTcoeffRuns && a < kTcoeffLevels) ? (int)tlen[run * kTcoeffLevels + a] : kEscapeBits;
}

/* One MB of a group as the lanes of its blocks need it: lane mi computes it
 * once per group (the stream's descriptor, the mode, the picture quantiser)
 * and keeps it in LDS, where every pass reads it back. */
struct MbInfo {
    const uint8_t *cur[3];
    const uint8_t *ref[3];
    uint8_t *rec[3];
    int32_t s;             /* stream; -1: no active stream */
    int32_t bx, by, width; /* luma position and width */
    int32_t quant, intra, filter, mvx, mvy;
    int32_t type, flags;   /* the mode's MB type and flags */
    int32_t pad[3];
};
static_assert(sizeof(MbInfo) == 128, "four 32-byte rows");

__device__ __forceinline__ MbInfo mb_info(const MbLaunch &p, int g)
{
    /* loads in two rounds: the MB's stream and mode, then the stream's
     * descriptor and picture quantiser (indices clamped, checked after) */
    MbInfo o{};
    const int gc = g < p.n_mb ? g : p.n_mb - 1;
    const int s = p.mb_stream[gc];
    const MbMode md = p.mode[gc];
    const int sc = s > 0 ? s : 0;
    const StreamDesc &d = p.streams[sc];
    const int rcq = p.rc[sc].pic_quant;
    const uint32_t flags = d.flags;
    o.s = (g < p.n_mb && s >= 0 && (flags & kStreamActive)) ? s : -1;
    if (o.s < 0) return o;
    const int m = g - d.mb_base;
    o.bx = (m % d.mb_w) * 16;
    o.by = (m / d.mb_w) * 16;
    o.width = d.width;
    const int picq = (flags & kStreamDeviceQuant) ? rcq : d.pic_quant;
    o.quant = clampi(picq + md.quant_adjust, 1, 31); /* gpu_device.cu:625-627 */
    o.intra = md.type == H261_MB_TYPE_INTRA;
    o.filter = (md.flags & H261_MB_FILTER) != 0;
    o.mvx = md.mvx;
    o.mvy = md.mvy;
    o.type = md.type;
    o.flags = md.flags;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        o.cur[k] = d.cur[k];
        o.ref[k] = d.ref[k];
        o.rec[k] = d.rec[k];
    }
    return o;
}

/* Per-lane view of the block a lane works on in one pass: lane (k, r) of a
 * pass takes block 8 * pass + k of its group of G macroblocks (6 blocks
 * each), row or column r. */
struct LaneBlock {
    bool act;      /* the block exists and its stream is active */
    int s;         /* stream */
    int b;         /* block in the MB: 0-3 luma, 4 Cb, 5 Cr */
    int slot;      /* statistics slot: 0 = the group's first stream, 1 = another */
    bool intra, filter;
    int quant, t;  /* t: 0 intra, 1 inter (statistics table) */
    int px, py, stride, mvx, mvy, plane;
    int mi;        /* MB of the group (its MbInfo holds the plane pointers) */
};

__device__ __forceinline__ LaneBlock lane_block(const MbInfo *info, int bi, int n_blocks, int s0)
{
    LaneBlock L{};
    const int mi = bi / 6, b = bi - 6 * mi;
    const MbInfo &o = info[bi < n_blocks ? mi : 0];
    L.b = b;
    L.mi = mi;
    L.s = bi < n_blocks ? o.s : -1;
    L.act = L.s >= 0;
    if (!L.act) return L;
    L.slot = L.s == s0 ? 0 : 1;
    L.quant = o.quant;
    L.intra = o.intra;
    L.filter = o.filter;
    L.t = L.intra ? 0 : 1;
    L.plane = b < 4 ? 0 : b - 3;
    if (b < 4) {
        L.px = o.bx + (b & 1) * 8;
        L.py = o.by + ((b >> 1) & 1) * 8;
        L.stride = o.width;
        L.mvx = o.mvx;
        L.mvy = o.mvy;
    } else { /* chroma vector = luma vector / 2, truncated (mbgpu:266-269) */
        L.px = o.bx >> 1;
        L.py = o.by >> 1;
        L.stride = o.width >> 1;
        L.mvx = o.mvx / 2;
        L.mvy = o.mvy / 2;
    }
    return L;
}

/* flushes the wave's denoise statistics of one stream (LDS -> global
 * atomics) and clears them; the frame tail reads them in a later launch */
__device__ __forceinline__ void flush_stats(const MbLaunch &p, int s, int32_t (*st)[64], int32_t *cnt, int lane)
{
    DenoiseState *dn = p.denoise + s;
    if (st[0][lane]) atomicAdd(&dn->intra_noise[lane], st[0][lane]);
    if (st[1][lane]) atomicAdd(&dn->inter_noise[lane], st[1][lane]);
    if (lane == 0 && cnt[0]) atomicAdd(&dn->intra_count, cnt[0]);
    if (lane == 1 && cnt[1]) atomicAdd(&dn->inter_count, cnt[1]);
}

/* One wave codes groups of G consecutive MBs (G = 1: one pass of 6 blocks,
 * lanes 48-63 idle; G = 4: 24 blocks in three full passes of 8).  Per pass:
 * phase 1 lane (block, row), phase 2 lane (block, column), then lane (block,
 * row) writes an uncoded block's reconstruction or keeps a coded block's
 * prediction in LDS.  The group's coded blocks then go through the dequantiser,
 * IDCT and reconstruction 8 at a time (at the batched bitrates about 9% of
 * inter blocks are coded, so mostly one pass or none).  Last, per MB of the
 * group, the CBP, bit count and outputs. */
template <int G>
/* 6 waves per SIMD: 80 VGPRs and 5.9 KB of LDS per wave (at 7.9 KB, with the
 * prediction rows and the TCOEFF lengths in LDS, the LDS held it at 5).  The
 * two-chain step gains more than the coder alone (0.331 -> 0.318 ms per full
 * step, 133 -> 131 us per launch): the other half's search waves fit beside
 * the coder's */
__global__ __launch_bounds__(64, 6) void mb_kernel(MbLaunch p)
{
    constexpr int kBlocks = 6 * G;
    constexpr int kPasses = (kBlocks + 7) / 8;
    __shared__ float tf[8 * kBlkPitch];
    __shared__ __align__(16) int8_t lv[kBlocks * 64]; /* levels; an intra DC (1..255) as uint8 */
    __shared__ int32_t stat[2][2][64]; /* [stream slot][intra, inter][coefficient] */
    __shared__ int32_t scount[2][2];
    __shared__ MbInfo info[G];
    /* the previous frame's denoise offsets staged as UInt16 (mbgpu:525-528):
     * [stream slot][intra, inter][coefficient] */
    __shared__ uint16_t dn_off[2][2][64];

    const int lane = threadIdx.x;
    const int k8 = lane >> 3; /* block of the pass */
    const int rc = lane & 7;  /* row in phases 1/3, column in phase 2 */

    stat[0][0][lane] = stat[0][1][lane] = stat[1][0][lane] = stat[1][1][lane] = 0;
    if (lane < 4) scount[lane >> 1][lane & 1] = 0;
    /* per-lane constants: zig-zag position of lane, column scale factors */
    const int zz = c_zigzag[lane];
    const float post_c = (float)((double)c_fwd_post[rc] * 0.125);
    const float pre_c = c_inv_pre[rc];

    const int n_wg = (p.n_mb - p.mb_first + p.mbs_per_wave - 1) / p.mbs_per_wave;
    const int wg = launch_index(blockIdx.x, n_wg, p.xcd_remap);
    if (wg >= n_wg) return;
    for (int grp = 0; grp * G < p.mbs_per_wave; grp++) {
        const int g0 = p.mb_first + wg * p.mbs_per_wave + grp * G;
        if (g0 >= p.n_mb) break;
        /* lane mi < G describes MB g0 + mi; the group's first stream with an
         * active MB comes from a ballot over those lanes */
        int my_s = -1;
        if (lane < G) {
            const MbInfo o = mb_info(p, g0 + lane);
            info[lane] = o;
            my_s = o.s;
        }
        const uint64_t has_s = __ballot(my_s >= 0);
        if (!has_s) continue;
        const int s0 = __builtin_amdgcn_readlane(my_s, (int)__builtin_ctzll(has_s));
        /* the group's other stream, if any (a stream has at least 99 MBs, so
         * a group of 4 MBs meets at most two) */
        const uint64_t has_o = __ballot(my_s >= 0 && my_s != s0);
        const int s1 = has_o ? __builtin_amdgcn_readlane(my_s, (int)__builtin_ctzll(has_o)) : -1;
        /* lane i loads entry i of the four offset tables now; they go to LDS
         * after the first pass's phase 1, which hides the loads */
        uint32_t dnv[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int st = k < 2 ? s0 : (s1 >= 0 ? s1 : s0);
            const int32_t *tab = (k & 1) ? p.denoise[st].inter_offset : p.denoise[st].intra_offset;
            dnv[k] = (uint32_t)tab[lane];
        }
        uint32_t coded_bits = 0;
        __syncthreads();

#pragma unroll 1
        for (int pass = 0; pass < kPasses; pass++) {
            const int bi = pass * 8 + k8; /* block of the group */
            const LaneBlock L = lane_block(info, bi, kBlocks, s0);

            /* ---- phase 1: loads first (current row, reference rows rc-1..rc+1
             * clamped to the block -- the loop filter is 1-D at the block
             * edges; an intra MB has a zero vector -- as aligned dwords, and
             * this lane's eight denoise offsets), then MC / loop filter,
             * residual, DCT rows ---- */
            uint32_t predw[2] = {0, 0};
            if (L.act) {
                const MbInfo &mbi = info[L.mi];
                uint32_t cw0, cw1;
                load_global2(mbi.cur[L.plane] + (size_t)(L.py + rc) * L.stride + L.px, cw0, cw1);
                const int rx = L.px + L.mvx;
                const int a = rx & ~3, o = rx & 3;
                /* the third dword is needed only when o > 0, and then it lies in the row */
                const int a2 = a + 8 <= L.stride - 4 ? a + 8 : a;
                const uint8_t *rbase = mbi.ref[L.plane] + (size_t)(L.py + L.mvy) * L.stride;
                uint32_t rw[3][3];
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    const int rr = clampi(rc + k - 1, 0, 7);
                    const uint8_t *q = rbase + (size_t)rr * L.stride;
                    rw[k][0] = load_global<uint32_t>(q + a);
                    rw[k][1] = load_global<uint32_t>(q + a + 4);
                    rw[k][2] = load_global<uint32_t>(q + a2);
                }
                /* MC, loop filter and residual on exact FP32 integers (pels,
                 * filter sums up to 4 x 1020 + 8, residuals): FP32 issues at
                 * twice the integer rate, the rounding shifts become products
                 * by powers of two and floor, and v_cvt_f32_ubyte / v_cvt_pk_u8
                 * move the pels between bytes and floats.  The 8 reference pels
                 * of rows rc-1, rc, rc+1: */
                float a8[3][8];
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    const uint32_t lo = __builtin_amdgcn_alignbyte(rw[k][1], rw[k][0], o);
                    const uint32_t hi = __builtin_amdgcn_alignbyte(rw[k][2], rw[k][1], o);
#pragma unroll
                    for (int x = 0; x < 4; x++) {
                        a8[k][x] = (float)((lo >> (8 * x)) & 0xffu);
                        a8[k][x + 4] = (float)((hi >> (8 * x)) & 0xffu);
                    }
                }
                float pred[8];
                if (L.intra) {
#pragma unroll
                    for (int x = 0; x < 8; x++) pred[x] = 0.0f;
                } else if (!L.filter) {
#pragma unroll
                    for (int x = 0; x < 8; x++) pred[x] = a8[1][x];
                } else {
                    /* H.261 loop filter, separable [1 2 1]: 1-D at the block
                     * edges, corners unfiltered (mbgpu:79-150).  Row weights
                     * (1, 2, 1), or (0, 4, 0) on the block's first and last
                     * rows: (4 h + 8) >> 4 == (h + 2) >> 2 */
                    float h[3][8];
#pragma unroll
                    for (int k = 0; k < 3; k++) {
                        h[k][0] = 4.0f * a8[k][0];
#pragma unroll
                        for (int x = 1; x < 7; x++) h[k][x] = __builtin_fmaf(2.0f, a8[k][x], a8[k][x - 1]) + a8[k][x + 1];
                        h[k][7] = 4.0f * a8[k][7];
                    }
                    const bool edge_row = rc == 0 || rc == 7;
                    const float w02 = edge_row ? 0.0f : 1.0f, w1 = edge_row ? 4.0f : 2.0f;
#pragma unroll
                    for (int x = 0; x < 8; x++) {
                        const float sum = __builtin_fmaf(w02, h[0][x] + h[2][x], __builtin_fmaf(w1, h[1][x], 8.0f));
                        pred[x] = floorf(sum * 0.0625f);
                    }
                }
                float res[8];
#pragma unroll
                for (int x = 0; x < 4; x++) {
                    res[x] = (float)((cw0 >> (8 * x)) & 0xffu) - pred[x];
                    res[x + 4] = (float)((cw1 >> (8 * x)) & 0xffu) - pred[x + 4];
                    predw[0] = __builtin_amdgcn_cvt_pk_u8_f32(pred[x], (uint32_t)x, predw[0]);
                    predw[1] = __builtin_amdgcn_cvt_pk_u8_f32(pred[x + 4], (uint32_t)x, predw[1]);
                }
                /* forward DCT, row pass */
                float o8[8];
                fdct_core(res[0] + res[7], res[1] + res[6], res[2] + res[5], res[3] + res[4], res[3] - res[4],
                          res[2] - res[5], res[1] - res[6], res[0] - res[7], o8);
#pragma unroll
                for (int k = 0; k < 8; k++) tf[k8 * kBlkPitch + rc * 8 + k] = o8[k];
            }
            if (pass == 0) {
#pragma unroll
                for (int k = 0; k < 4; k++) dn_off[k >> 1][k & 1][lane] = (uint16_t)dnv[k];
            }
            __syncthreads();

            /* ---- phase 2: lane (block, column c).  From the postscale
             * rounding on, the coefficients stay floats: every value below
             * is an integer of magnitude under 2^24, exact in FP32, so the
             * denoise and the quantiser give the reference's integer results
             * with FP32 instructions (twice the integer issue rate on this
             * GPU): |c| - offset and the clamp at zero are exact, a float
             * product's sign is the operands' ((float)(-a) * inv ==
             * -(a * inv)), and truncation and the +-127 clamp commute with
             * the sign ---- */
            float blk_accum = 0.0f;
            if (L.act) {
                const int c = rc;
                float v[8];
#pragma unroll
                for (int r = 0; r < 8; r++) v[r] = tf[k8 * kBlkPitch + r * 8 + c];
                const float cs = post_c;
                float o8[8];
                fdct_core(v[0] + v[7], v[1] + v[6], v[2] + v[5], v[3] + v[4], v[3] - v[4], v[2] - v[5],
                          v[1] - v[6], v[0] - v[7], o8);
                /* the reference's floorf(x + 0.49999): a double sum rounded
                 * to float, then floored.  For every float x with |x| <=
                 * 1e7 (coefficients stay below 2^13) that equals floorf of
                 * the FP32 sum x + 0.49999f (checked over every float bit
                 * pattern), so no conversions or FP64 add are needed */
                float cf[8];
#pragma unroll
                for (int k = 0; k < 8; k++) cf[k] = floorf(cs * kFwdPost[k] * o8[k] + 0.49999f);

                /* denoise with the previous frame's offsets, staged as UInt16
                 * (mbgpu:525-528, 586-604); the statistics stay integer LDS
                 * atomics (float ones are far slower when the 8 blocks of a
                 * pass add to one address) */
#pragma unroll
                for (int r = 0; r < 8; r++) {
                    const int i = c + 8 * r;
                    const float off = (float)dn_off[L.slot][L.t][i];
                    const float a = fabsf(cf[r]);
                    atomicAdd(&stat[L.slot][L.t][i], (int)a);
                    cf[r] = copysignf(fmaxf(a - off, 0.0f), cf[r]);
                }

                /* forward quantiser (mbgpu:685-716): level = clamp(trunc(c *
                 * inv), -127, 127), taken as sign(c) * min(trunc(|c| * inv), 127) */
                const float inv = c_inv_quant[L.quant];
#pragma unroll
                for (int r = 0; r < 8; r++) {
                    const int i = c + 8 * r;
                    int l;
                    if (L.intra && i == 0) {
                        l = (int)cf[r] / 8;
                        l = clampi(l, 1, 254);
                        if (l == 128) l = 255;
                        blk_accum += 100.0f;
                    } else {
                        const float t = fminf(truncf(fabsf(cf[r]) * inv), 127.0f);
                        l = (int)copysignf(t, cf[r]);
                        blk_accum += t > 1.0f ? 100.0f : t;
                    }
                    lv[bi * 64 + i] = (int8_t)l;
                }
            }
            /* sum the 8 columns of the block: quads, then the two quads of
             * the 8 lanes (row_half_mirror pairs lane j with 7 - j) */
            blk_accum += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(blk_accum), 0xB1, 0xf, 0xf, false));
            blk_accum += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(blk_accum), 0x4E, 0xf, 0xf, false));
            blk_accum += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(blk_accum), 0x141, 0xf, 0xf, false));
            const bool coded = L.act && blk_accum > 2.0f;
            /* coded blocks of the pass: ballot of the lanes c == 0 */
            const uint64_t coded_mask = __ballot(coded && rc == 0);
#pragma unroll
            for (int k = 0; k < 8; k++)
                if ((coded_mask >> (8 * k)) & 1) coded_bits |= 1u << (pass * 8 + k);

            /* lane (block, row rc) holds row rc of the block's prediction: the
             * reconstruction of an uncoded block, and for a coded block the
             * prediction its residual is added to in the coded-block passes
             * below, which read it back (the barriers between order the
             * wave's global stores and loads).  Kept in LDS instead, the
             * coder would hold a wave per SIMD less */
            if (L.act) {
                uint8_t *dst = info[L.mi].rec[L.plane] + (size_t)L.py * L.stride + L.px;
                store_global2(dst + (size_t)rc * L.stride, predw[0], predw[1]);
            }
            __syncthreads(); /* tf is reused by the next pass */
        }

        /* ---- the group's coded blocks, 8 per pass: slot k8 takes the
         * (8 cp + k8)-th set bit of coded_bits ---- */
        const int n_coded = __builtin_popcount(coded_bits);
        uint32_t rest = coded_bits;
#pragma unroll 1
        for (int cp = 0; cp * 8 < n_coded; cp++) {
            int bi = -1;
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const int b = rest ? __builtin_ctz(rest) : -1;
                bi = k8 == j ? b : bi;
                rest &= rest - 1;
            }
            const bool act = bi >= 0;
            /* the block's MB descriptor gives its quantiser, type and place */
            const LaneBlock C = lane_block(info, act ? bi : 0, kBlocks, s0);
            const bool intra = C.intra;
            const int qmul = C.quant * 2, qadd = (C.quant - 1) | 1;
            /* the prediction row this lane adds to in the row pass, loaded now
             * so the load overlaps the column pass */
            uint32_t predw[2] = {0, 0};
            uint8_t *rdst = info[C.mi].rec[C.plane] + (size_t)(C.py + rc) * C.stride + C.px;
            if (act) load_global2(rdst, predw[0], predw[1]);
            if (act) {
                const int c = rc;
                /* dequantiser (mbgpu:763-814) */
                float tv[8], o8[8];
#pragma unroll
                for (int r = 0; r < 8; r++) {
                    const int i = c + 8 * r;
                    const int raw = lv[bi * 64 + i];
                    int dq;
                    if (intra && i == 0) {
                        const int l = raw & 0xff;
                        dq = l == 255 ? 1024 : l * 8;
                    } else {
                        int x = raw * qmul;
                        if (x > 0)
                            x += qadd;
                        else if (x < 0)
                            x -= qadd;
                        dq = clampi(x, -2048, 2047);
                    }
                    /* IDCT prescale (mbgpu:836-849) */
                    const float cs = (float)((double)kInvPre[r] * 0.125);
                    tv[r] = ((float)dq * pre_c) * cs;
                }
                /* IDCT column pass (mbgpu:851-888) */
                idct_core(tv, o8);
#pragma unroll
                for (int r = 0; r < 8; r++) tf[k8 * kBlkPitch + r * 8 + c] = o8[r];
            }
            __syncthreads();

            /* lane (slot, row rc): IDCT rows, reconstruction */
            if (act) {
                float v[8], o8[8];
#pragma unroll
                for (int k = 0; k < 8; k++) v[k] = tf[k8 * kBlkPitch + rc * 8 + k];
                idct_core(v, o8);
                uint32_t w[2] = {0, 0};
#pragma unroll
                for (int x = 0; x < 8; x++) {
                    const int rres = clampi((int)floor((double)o8[x] + 0.49999), -256, 255);
                    const int pv = (int)((predw[x >> 2] >> (8 * (x & 3))) & 0xff);
                    const int val = clampi(pv + rres, 0, 255);
                    w[x >> 2] |= (uint32_t)val << (8 * (x & 3));
                }
                store_global2(rdst, w[0], w[1]);
            }
            __syncthreads(); /* tf and slots are reused by the next pass */
        }

        /* ---- per MB of the group: CBP, bit count of the CBP + TCOEFF
         * layer, outputs ---- */
#pragma unroll 1
        for (int mi = 0; mi < G; mi++) {
            const int g = g0 + mi;
            /* the MB as the passes saw it (wave-uniform values from LDS) */
            const int s = __builtin_amdgcn_readfirstlane(info[mi].s);
            if (s < 0) continue; /* past the launch, unmapped, or its stream inactive */
            const int quant = __builtin_amdgcn_readfirstlane(info[mi].quant);
            const int mtype = __builtin_amdgcn_readfirstlane(info[mi].type);
            const int mflags = __builtin_amdgcn_readfirstlane(info[mi].flags);
            const int mvx = __builtin_amdgcn_readfirstlane(info[mi].mvx);
            const int mvy = __builtin_amdgcn_readfirstlane(info[mi].mvy);
            const bool intra = mtype == H261_MB_TYPE_INTRA;
            int cbp = 0;
#pragma unroll
            for (int k = 0; k < 6; k++)
                if ((coded_bits >> (6 * mi + k)) & 1) cbp |= kCbpBit[k];
            const int8_t *mlv = lv + 6 * mi * 64;

            int bits = 0;
            if (p.side && (intra || cbp)) {
                /* lane k = transmission position k of each block in turn; the
                 * written blocks' levels are read up front */
                int lk[6];
#pragma unroll
                for (int k = 0; k < 6; k++)
                    lk[k] = (intra || (cbp & kCbpBit[k])) ? (int)mlv[k * 64 + zz] : 0;
#pragma unroll
                for (int k = 0; k < 6; k++) {
                    const bool blk_written = intra || (cbp & kCbpBit[k]);
                    if (!blk_written) continue;
                    const int l = lk[k]; /* an intra DC reads as non-zero either way */
                    const uint64_t nz = __ballot(l != 0);
                    if (l != 0) {
                        const uint64_t below = lane ? (nz & ((~0ull) >> (64 - lane))) : 0ull;
                        const int prev = below ? 63 - __builtin_clzll(below) : -1;
                        const int run = lane - prev - 1;
                        if (lane == 0) {
                            if (intra)
                                bits += 8; /* DC: 8-bit fixed length */
                            else if (l == 1 || l == -1)
                                bits += 2; /* first coefficient '1s' */
                            else
                                bits += tcoeff_bits(0, l);
                        } else {
                            bits += tcoeff_bits(run, l);
                        }
                    }
                    if (lane == 0) bits += 2; /* EOB */
                }
                bits = wave_sum_i32(bits);
                if (!intra) bits += c_cbp[cbp].len;
            }

            const int flags = mflags | (cbp ? H261_MB_COEFF : 0);
            if (p.levels) {
                /* 384 levels as 96 dwords of 4 bytes: AC levels are within
                 * [-127, 127], an intra DC within [1, 255] (read back as uint8) */
                uint32_t *out = (uint32_t *)(p.levels + (size_t)g * 384);
#pragma unroll
                for (int j = 0; j < 2; j++) {
                    const int w = lane + 64 * j;
                    if (w < 96) out[w] = ((const uint32_t *)mlv)[w];
                }
            }
            if (p.mb_struct) {
                h261_macroblock_t *o = p.mb_struct + g;
#pragma unroll
                for (int k = 0; k < 6; k++)
                    o->rgi_tcoeff[k][lane] = intra && lane == 0 ? (int)(uint8_t)mlv[k * 64] : (int)mlv[k * 64 + lane];
                if (lane == 0) {
                    o->i_macroblock_address = 0;
                    o->i_macroblock_type = mtype;
                    o->i_macroblock_type_flags = flags;
                    o->i_macroblock_quant = quant;
                    o->rgi_mv[0] = mvx;
                    o->rgi_mv[1] = mvy;
                    o->i_coded_block_pattern = cbp;
                    o->i_mb_x = info[mi].bx;
                    o->i_mb_y = info[mi].by;
                }
            }
            if (p.side && lane == 0) {
                MbSide sd;
                sd.type = (int8_t)mtype;
                sd.flags = (int8_t)flags;
                sd.quant = (int8_t)quant;
                sd.cbp = (int8_t)cbp;
                sd.mvx = (int8_t)mvx;
                sd.mvy = (int8_t)mvy;
                sd.bits = (int16_t)bits;
                p.side[g] = sd;
            }
            if (lane == 0) atomicAdd(&scount[s == s0 ? 0 : 1][intra ? 0 : 1], 6);
        }
        __syncthreads();

        /* statistics of the group's streams go out at the group's end.  A
         * stream has at least 99 MBs, so a group of 4 MBs meets at most two */
        flush_stats(p, s0, stat[0], scount[0], lane);
        const uint64_t has1 = __ballot(s1 >= 0);
        if (has1) flush_stats(p, __builtin_amdgcn_readlane(s1, (int)__builtin_ctzll(has1)), stat[1], scount[1], lane);
        __syncthreads();
        stat[0][0][lane] = stat[0][1][lane] = stat[1][0][lane] = stat[1][1][lane] = 0;
        if (lane < 4) scount[lane >> 1][lane & 1] = 0;
    }
}

}  // namespace

int launch_mb(const MbLaunch &p, void *stream, void *ev_start, void *ev_stop)
{
    if (p.n_mb - p.mb_first <= 0) {
        if (ev_start) { /* nothing to launch: an empty timed interval */
            (void)hipEventRecord((hipEvent_t)ev_start, (hipStream_t)stream);
            (void)hipEventRecord((hipEvent_t)ev_stop, (hipStream_t)stream);
        }
        return 0;
    }
    const int per = p.mbs_per_wave > 0 ? p.mbs_per_wave : 1;
    MbLaunch q = p;
    q.mbs_per_wave = per;
    const int grid = launch_grid((p.n_mb - p.mb_first + per - 1) / per, p.xcd_remap);
    /* groups of 4 MBs fill all 64 lanes; otherwise one MB per pass */
    auto k = per % 4 == 0 ? mb_kernel<4> : mb_kernel<1>;
    if (ev_start)
        hipExtLaunchKernelGGL(k, dim3(grid), dim3(64), 0, (hipStream_t)stream, (hipEvent_t)ev_start,
                              (hipEvent_t)ev_stop, 0, q);
    else
        hipLaunchKernelGGL(k, dim3(grid), dim3(64), 0, (hipStream_t)stream, q);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace h261g
