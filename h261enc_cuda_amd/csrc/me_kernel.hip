/*
 * me_kernel.hip -- motion estimation for gfx950.
 *
 * One wave64 per 16x16 macroblock, a run of consecutive MBs per wave:
 *   1. the reference search window (rows by-15..by+30, cols bx-16..bx+31)
 *      is staged in LDS with aligned dword loads.  MBs of a run that lie
 *      side by side in one MB row share one staged tile of kTile windows
 *      (rows by-15..by+30, cols bx-16..bx+16 kTile+15): each reference pel
 *      is fetched once per tile instead of up to three times;
 *   2. stage 1 evaluates the 11x11 coarse grid
 *      {-14,-11,-8,-5,-2,0,2,5,8,11,14}^2 (src/megpu_kernels.cuh:29-32),
 *      stage 2 the 7x7 refine around the winner (megpu:309-312, 338-339),
 *      each candidate one lane, 16 rows x 4 packed v_sad_u8 (alignbyte
 *      realigns the window bytes to the candidate's column); which lane
 *      takes which candidate comes from constant tables laid out so that a
 *      half-wave's LDS reads never meet in a bank, and validity is a lane
 *      mask built with scalar instructions;
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
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include "frame_control.h"
#include "h261gpu_internal.h"

namespace h261g {

namespace {

constexpr int kWinRows = 46;  /* 16 + 2*15 */
/* word 3 of a gfx9 buffer resource: 32-bit data format, raw addressing */
constexpr int kBufferDword3 = 0x00020000;
constexpr int kWinDw = 12;    /* 48 bytes per row: frame cols bx-16 .. bx+31 */
/* dwords per tile row: kTile windows side by side, 16 pels apart */
constexpr int tile_dw(int tile)
{
    return 4 * tile + 8;
}
/* LDS dwords per tile row.  Two-stage search: an odd pitch whose residue
 * mod 32 lets the lane tables below read without bank conflicts (15 for one
 * window; a tile needs at least tile_dw dwords: 17 for two, 29 for four).
 * Surface variant: 12, its 16-byte row loads need aligned rows. */
template <int kVariant, int kTile>
constexpr int win_pitch()
{
    return kVariant == 1 ? kWinDw : kTile == 1 ? 15 : kTile == 2 ? 17 : 29;
}
/* The two-stage tile has two spare rows above and below and one spare
 * dword in front: a refine candidate outside the limits (up to 17 px from
 * the MB) then still reads LDS inside the array; its cost is replaced by
 * 65535. win = winmem + win_base. */
template <int kVariant, int kTile>
constexpr int win_base()
{
    return kVariant == 1 ? 0 : 2 * win_pitch<kVariant, kTile>() + 1;
}
template <int kVariant, int kTile>
constexpr int win_alloc()
{
    return kVariant == 1 ? kWinRows * kWinDw
                         : (win_base<kVariant, kTile>() + (kWinRows + 2) * win_pitch<kVariant, kTile>() + 3) & ~3;
}
/* a refine read of the tile's last MB reaches dword tile_dw (for candidates
 * outside the limits, whose cost is replaced): still inside the row */
static_assert(tile_dw(1) < 15 && tile_dw(2) < 17 && tile_dw(4) < 29, "tile rows fit their pitch");

/* minimum over the wave: quads, half rows and rows by DPP, rows by readlane.
 * Every lane reads a lane of its own row, so the DPP "old" operand is never
 * used; giving it min's identity lets the compiler fold each move into the
 * v_min_u32 as a DPP source (one instruction per stage instead of a move,
 * a wait state and the min) */
__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v)
{
    uint32_t o;
    o = (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)v, 0xB1, 0xf, 0xf, false); /* quad_perm [1,0,3,2] */
    v = o < v ? o : v;
    o = (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)v, 0x4E, 0xf, 0xf, false); /* quad_perm [2,3,0,1] */
    v = o < v ? o : v;
    o = (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)v, 0x141, 0xf, 0xf, false); /* row_half_mirror */
    v = o < v ? o : v;
    o = (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)v, 0x140, 0xf, 0xf, false); /* row_mirror */
    v = o < v ? o : v;
    const uint32_t r0 = (uint32_t)__builtin_amdgcn_readlane((int)v, 0), r1 = (uint32_t)__builtin_amdgcn_readlane((int)v, 16);
    const uint32_t r2 = (uint32_t)__builtin_amdgcn_readlane((int)v, 32), r3 = (uint32_t)__builtin_amdgcn_readlane((int)v, 48);
    const uint32_t a = r0 < r1 ? r0 : r1, b = r2 < r3 ? r2 : r3;
    return a < b ? a : b;
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

__host__ __device__ constexpr int coarse_vec(int i)
{
    /* {-14,-11,-8,-5,-2,0,2,5,8,11,14} */
    return i < 5 ? 3 * i - 14 : (i == 5 ? 0 : 3 * i - 16);
}

constexpr uint32_t morton_c(uint32_t x, uint32_t y)
{
    uint32_t k = 0;
    for (int b = 0; b < 4; b++) k |= (((x >> b) & 1u) << (2 * b)) | (((y >> b) & 1u) << (2 * b + 1));
    return k;
}

/* Per-lane candidate tables of the two-stage search, one set per tile pitch
 * P (15, 17, 29).  A candidate (vx, vy) starts at LDS byte
 * 4 P (vy + 15) + vx + 16 of its MB's window: its dword is that >> 2, its
 * byte shift that & 3.
 *
 * Lane layout.  LDS serves a wave's dword reads a half-wave (32 lanes) at a
 * time over 32 banks; lanes reading different dwords of one bank wait for
 * each other.  All lanes of a read add the same row and dword to their
 * candidate's start (and all lanes the same window offset within the tile),
 * so a half-wave reads without conflict when its candidates' distinct start
 * dwords lie in distinct banks.
 *   coarse: kCoarseOrder[P][2 i + h][j] is the candidate (tx + 11 ty) of
 *           lane 32 h + j in slot i, an assignment of the 121 candidates
 *           found by a search for that property (88 distinct start dwords
 *           in 4 x 32 bank slots; checked below).  Spare lanes repeat the
 *           first candidate of their half-wave (the same dword, so a
 *           broadcast) and are masked off;
 *   refine: the 7 x 7 rows ty go to lanes 7 j + tx of half-wave 0 in the
 *           order 0, 1, 4, 5 and of half-wave 1 in the order 2, 3, 6: for
 *           each pitch, rows P dwords apart then never meet in a bank,
 *           whatever the winner's byte alignment.  Spare lanes repeat the
 *           first lane of their half-wave and are masked off.
 * Entries: byte address (coarse) or 4 P ty + tx added to the winner's byte
 * address (refine) | Z-order key << 16. */
constexpr int kTwoStagePitches[3] = {15, 17, 29};
constexpr int pitch_set(int pitch)
{
    return pitch == 15 ? 0 : pitch == 17 ? 1 : 2;
}
constexpr int8_t kCoarseOrder[3][4][32] = {
    {/* pitch 15 */
     {0, 5, 6, 7, 13, 14, 16, 17, 24, 25, 26, 27, 28, 29, 32, 49, 50, 52, 53, 57, 58, 73, 74, 75, 76, 88, 89, 104, 105, -1, -1, -1},
     {8, 9, 10, 11, 34, 35, 36, 43, 54, 55, 56, 60, 61, 62, 66, 67, 71, 72, 78, 85, 86, 100, 101, 102, 103, 107, 108, 110, 118, 119, 120, -1},
     {4, 12, 18, 19, 20, 21, 23, 30, 31, 37, 48, 59, 63, 64, 65, 77, 82, 83, 84, 87, 90, 91, 92, 96, 97, 112, 113, 114, 117, -1, -1, -1},
     {1, 2, 3, 15, 22, 33, 38, 39, 40, 41, 42, 44, 45, 46, 47, 51, 68, 69, 70, 79, 80, 81, 93, 94, 95, 98, 99, 106, 109, 111, 115, 116}},
    {/* pitch 17 */
     {4, 7, 10, 12, 16, 17, 24, 25, 26, 27, 28, 29, 41, 42, 56, 60, 61, 65, 66, 70, 74, 75, 92, 95, 106, 111, 112, 113, 115, 116, -1, -1},
     {1, 5, 6, 13, 14, 15, 30, 31, 34, 37, 38, 39, 45, 46, 47, 48, 51, 62, 76, 77, 79, 80, 81, 82, 83, 90, 91, 99, 100, 109, 114, -1},
     {0, 2, 3, 11, 19, 20, 22, 23, 32, 40, 44, 49, 50, 52, 53, 59, 63, 64, 84, 85, 86, 87, 88, 89, 93, 94, 96, 97, 101, 102, -1, -1},
     {8, 9, 18, 21, 33, 35, 36, 43, 54, 55, 57, 58, 67, 68, 69, 71, 72, 73, 78, 98, 103, 104, 105, 107, 108, 110, 117, 118, 119, 120, -1, -1}},
    {/* pitch 29 */
     {0, 1, 4, 8, 9, 16, 17, 19, 20, 30, 31, 32, 37, 43, 52, 53, 56, 57, 58, 59, 67, 85, 86, 88, 89, 98, 109, 112, 113, 120, -1, -1},
     {2, 3, 10, 11, 18, 22, 23, 24, 25, 27, 28, 38, 39, 41, 42, 45, 46, 47, 48, 51, 54, 60, 61, 70, 74, 75, 78, 81, 104, 105, -1, -1},
     {5, 6, 7, 21, 34, 44, 62, 68, 69, 71, 72, 73, 77, 82, 83, 87, 90, 91, 93, 94, 100, 101, 102, 103, 106, 110, 114, 115, 116, 117, -1, -1},
     {12, 13, 14, 15, 26, 29, 33, 35, 36, 40, 49, 50, 55, 63, 64, 65, 66, 76, 79, 80, 84, 92, 95, 96, 97, 99, 107, 108, 111, 118, 119, -1}},
};
constexpr int kRefineRowBase[7] = {0, 7, 32, 39, 14, 21, 46}; /* first lane of refine row ty */

constexpr int coarse_start(int pitch, int ci)
{
    return 4 * pitch * (coarse_vec(ci / 11) + 15) + coarse_vec(ci % 11) + 16;
}
constexpr int coarse_at(int ps, int i, int l) /* candidate of lane l in slot i, spare lanes resolved */
{
    const int c = kCoarseOrder[ps][2 * i + l / 32][l % 32];
    return c >= 0 ? c : kCoarseOrder[ps][2 * i + l / 32][0];
}
struct CoarseLanes {
    uint32_t e[2][64];
};
constexpr CoarseLanes make_coarse_lanes(int ps)
{
    CoarseLanes t{};
    for (int i = 0; i < 2; i++)
        for (int l = 0; l < 64; l++) {
            const int ci = coarse_at(ps, i, l), tx = ci % 11, ty = ci / 11;
            t.e[i][l] = (uint32_t)coarse_start(kTwoStagePitches[ps], ci) | (morton_c((uint32_t)tx, (uint32_t)ty) << 16);
        }
    return t;
}
/* every candidate in exactly one lane; distinct start dwords of a half-wave
 * in distinct banks */
constexpr bool coarse_order_ok(int ps)
{
    const int pitch = kTwoStagePitches[ps];
    int seen[121] = {};
    for (int k = 0; k < 4; k++)
        for (int j = 0; j < 32; j++) {
            const int c = kCoarseOrder[ps][k][j];
            if (c < -1 || c > 120) return false;
            if (c >= 0) seen[c]++;
            for (int j2 = 0; j2 < j; j2++) {
                const int c2 = kCoarseOrder[ps][k][j2];
                if (c < 0 || c2 < 0) continue;
                const int a = coarse_start(pitch, c) >> 2, a2 = coarse_start(pitch, c2) >> 2;
                if (a != a2 && a % 32 == a2 % 32) return false;
            }
        }
    for (int c = 0; c < 121; c++)
        if (seen[c] != 1) return false;
    return true;
}
static_assert(coarse_order_ok(0) && coarse_order_ok(1) && coarse_order_ok(2), "coarse lane layouts");

struct RefineLanes {
    uint32_t e[64];
};
constexpr RefineLanes make_refine_lanes(int ps)
{
    RefineLanes t{};
    for (int l = 0; l < 64; l++) {
        int tx = 0, ty = l < 32 ? 0 : 2; /* spare lanes: the first lane of the half-wave */
        for (int r = 0; r < 7; r++)
            if (l >= kRefineRowBase[r] && l < kRefineRowBase[r] + 7) {
                tx = l - kRefineRowBase[r];
                ty = r;
            }
        t.e[l] = (uint32_t)(4 * kTwoStagePitches[ps] * ty + tx) | (morton_c((uint32_t)tx, (uint32_t)ty) << 16);
    }
    return t;
}
/* rows of a half-wave: start dwords P ty + (a + tx) / 4, a = 0..3, in
 * distinct banks for distinct dwords */
constexpr bool refine_rows_ok(int pitch)
{
    for (int a = 0; a < 4; a++)
        for (int r = 0; r < 7; r++)
            for (int r2 = 0; r2 < 7; r2++)
                for (int x = 0; x < 7; x++)
                    for (int x2 = 0; x2 < 7; x2++) {
                        if ((kRefineRowBase[r] < 32) != (kRefineRowBase[r2] < 32)) continue;
                        const int d = pitch * r + (a + x) / 4, d2 = pitch * r2 + (a + x2) / 4;
                        if (d != d2 && d % 32 == d2 % 32) return false;
                    }
    return kRefineRowBase[3] + 7 <= 64;
}
static_assert(refine_rows_ok(15) && refine_rows_ok(17) && refine_rows_ok(29), "refine lane layout");

/* Lane masks of the in-limit coarse candidates by frame edge.  An MB's
 * limits are +-15 clipped at the frame border (gpu_device.cu:226-265) and MB
 * positions are multiples of 16, so each limit has one of two values: -15
 * or 0 on the left and top, 15 or 0 on the right and bottom.  Entry
 * (left | right << 1 | top << 2 | bottom << 3) holds the masks of the two
 * coarse slots; spare lanes are never valid. */
struct CoarseEdgeMasks {
    uint64_t m[16][2];
};
constexpr CoarseEdgeMasks make_coarse_edge_masks(int ps)
{
    CoarseEdgeMasks t{};
    for (int e = 0; e < 16; e++) {
        const int lx0 = (e & 1) ? 0 : -15, lx1 = (e & 2) ? 0 : 15, ly0 = (e & 4) ? 0 : -15, ly1 = (e & 8) ? 0 : 15;
        for (int i = 0; i < 2; i++)
            for (int l = 0; l < 64; l++) {
                const int c = kCoarseOrder[ps][2 * i + l / 32][l % 32];
                if (c < 0) continue;
                const int vx = coarse_vec(c % 11), vy = coarse_vec(c / 11);
                if (vx >= lx0 && vx <= lx1 && vy >= ly0 && vy <= ly1) t.m[e][i] |= 1ull << l;
            }
    }
    return t;
}

struct TwoStageTables {
    CoarseLanes coarse;
    RefineLanes refine;
    CoarseEdgeMasks edge;
};
constexpr TwoStageTables make_two_stage_tables(int ps)
{
    return TwoStageTables{make_coarse_lanes(ps), make_refine_lanes(ps), make_coarse_edge_masks(ps)};
}
__constant__ TwoStageTables c_two_stage[3] = {make_two_stage_tables(0), make_two_stage_tables(1),
                                              make_two_stage_tables(2)};

__device__ __forceinline__ uint64_t refine_mask(int sx, int sy, int lx0, int lx1, int ly0, int ly1)
{
    const int x0 = max(lx0 - sx + 3, 0), x1 = min(lx1 - sx + 3, 6);
    const int y0 = max(ly0 - sy + 3, 0), y1 = min(ly1 - sy + 3, 6);
    const uint32_t xm = x1 >= x0 ? (2u << x1) - (1u << x0) : 0u;
    uint64_t m = 0;
#pragma unroll
    for (int ty = 0; ty < 7; ty++)
        if (ty >= y0 && ty <= y1) m |= (uint64_t)xm << kRefineRowBase[ty];
    return m;
}

/* one row of a candidate: window dwords o..o+4 realigned by sh bytes against
 * the current row's four dwords (scalar operands) */
__device__ __forceinline__ uint32_t sad_row(const uint32_t *row, int sh, uint32_t c0, uint32_t c1, uint32_t c2,
                                            uint32_t c3, uint32_t acc)
{
    const uint32_t w0 = row[0], w1 = row[1], w2 = row[2], w3 = row[3], w4 = row[4];
    acc = __builtin_amdgcn_sad_u8(__builtin_amdgcn_alignbyte(w1, w0, sh), c0, acc);
    acc = __builtin_amdgcn_sad_u8(__builtin_amdgcn_alignbyte(w2, w1, sh), c1, acc);
    acc = __builtin_amdgcn_sad_u8(__builtin_amdgcn_alignbyte(w3, w2, sh), c2, acc);
    acc = __builtin_amdgcn_sad_u8(__builtin_amdgcn_alignbyte(w4, w3, sh), c3, acc);
    return acc;
}

/* 16x16 SADs of NS candidates per lane against the staged window.  The
 * current MB's rows are wave-uniform: they are read with scalar loads (the
 * constant address space) and enter v_sad_u8 as scalar operands. */
typedef const __attribute__((address_space(4))) uint32_t *scalar_u32_ptr;

template <int NS, int kPitch>
__device__ __forceinline__ void sad16(const uint32_t *win, const uint8_t *cur_mb, int pitch, const int (&off)[NS],
                                      const int (&sh)[NS], uint32_t (&acc)[NS])
{
#pragma unroll
    for (int i = 0; i < NS; i++) acc[i] = 0;
    /* each search stage loads the current rows anew, next to their use: kept
     * from one stage to the next they would hold 64 SGPRs and spill */
    const uint64_t cm = (uint64_t)(uintptr_t)cur_mb;
    uint64_t base = ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(cm >> 32)) << 32) |
                    (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)cm);
    asm volatile("" : "+s"(base));
    /* the rows come in batches of 4 (16 SGPRs), each waited for once: a
     * scalar load shares the LDS wait counter and returns out of order, so
     * while one is in flight every wait for a window read waits for all
     * outstanding LDS reads too.  Inside a batch the window reads of
     * successive rows then overlap (ME -2.5% in a same-box A/B) */
#pragma unroll
    for (int kb = 0; kb < 16; kb += 4) {
        uint32_t cr[4][4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const scalar_u32_ptr c = (scalar_u32_ptr)(uintptr_t)(base + (uint64_t)(kb + k) * pitch);
            cr[k][0] = c[0];
            cr[k][1] = c[1];
            cr[k][2] = c[2];
            cr[k][3] = c[3];
        }
        asm volatile("" ::"s"(cr[0][0]), "s"(cr[0][1]), "s"(cr[0][2]), "s"(cr[0][3]), "s"(cr[1][0]), "s"(cr[1][1]),
                     "s"(cr[1][2]), "s"(cr[1][3]));
        asm volatile("" ::"s"(cr[2][0]), "s"(cr[2][1]), "s"(cr[2][2]), "s"(cr[2][3]), "s"(cr[3][0]), "s"(cr[3][1]),
                     "s"(cr[3][2]), "s"(cr[3][3]));
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int k = 0; k < 4; k++)
#pragma unroll
            for (int i = 0; i < NS; i++)
                acc[i] = sad_row(win + off[i] + (kb + k) * kPitch, sh[i], cr[k][0], cr[k][1], cr[k][2], cr[k][3], acc[i]);
        __builtin_amdgcn_sched_barrier(0);
    }
}

typedef short short2v __attribute__((ext_vector_type(2)));

/* a plane pointer from a stream descriptor is a generic pointer to the
 * compiler; the load names the global address space, so it is not issued as
 * a flat load (which would also count against the LDS wait counter) */
__device__ __forceinline__ uint32_t load_global_u32(const uint8_t *p)
{
    return *(const __attribute__((address_space(1))) uint32_t *)(p);
}

__device__ __forceinline__ short2v as_s2(uint32_t v)
{
    short2v r;
    __builtin_memcpy(&r, &v, 4);
    return r;
}

/* partner value across the lanes of a quad (DPP quad_perm) */
template <int kCtrl>
__device__ __forceinline__ short2v quad_swap(short2v v)
{
    int x;
    __builtin_memcpy(&x, &v, 4);
    x = __builtin_amdgcn_mov_dpp(x, kCtrl, 0xf, 0xf, false);
    short2v r;
    __builtin_memcpy(&r, &x, 4);
    return r;
}

/* 4x4 Hadamard of a sub-block held by the 4 lanes of a quad, lane y = row y,
 * as pairs (d0, d2), (d1, d3) of packed int16; returns this lane's sum of
 * |coefficient| over its 4 coefficients (the sub-block's SATD summed over
 * the quad is megpu_satd_4x4_kernel, src/megpu_kernels.cuh:467-500) */
__device__ __forceinline__ int satd_quad(short2v e, short2v o, short2v s1, short2v s2)
{
    const short2v sum = e + o, dif = e - o;   /* (a1, a3), (a2, a4) */
    const short2v u = {sum.x, dif.x}, v = {sum.y, dif.y};
    short2v p = u + v, m = u - v;             /* (t0, t1), (t2, t3) */
    /* columns: butterflies across the quad; an odd partner negates itself */
    p = p * s1 + quad_swap<0xB1>(p);
    m = m * s1 + quad_swap<0xB1>(m);
    p = p * s2 + quad_swap<0x4E>(p);
    m = m * s2 + quad_swap<0x4E>(m);
    const short2v ap = __builtin_elementwise_max(p, -p), am = __builtin_elementwise_max(m, -m);
    const short2v t = ap + am;
    return (int)(uint16_t)t.x + (int)(uint16_t)t.y;
}

/* sum over the wave: quads, half rows and rows by DPP, rows by readlane */
__device__ __forceinline__ int wave_sum(int x)
{
    /* "old" = 0, the sum's identity: each move folds into the add */
    x += __builtin_amdgcn_update_dpp(0, x, 0xB1, 0xf, 0xf, false);  /* quad_perm [1,0,3,2] */
    x += __builtin_amdgcn_update_dpp(0, x, 0x4E, 0xf, 0xf, false);  /* quad_perm [2,3,0,1] */
    x += __builtin_amdgcn_update_dpp(0, x, 0x141, 0xf, 0xf, false); /* row_half_mirror */
    x += __builtin_amdgcn_update_dpp(0, x, 0x140, 0xf, 0xf, false); /* row_mirror */
    return __builtin_amdgcn_readlane(x, 0) + __builtin_amdgcn_readlane(x, 16) + __builtin_amdgcn_readlane(x, 32) +
           __builtin_amdgcn_readlane(x, 48);
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


/* The staged tile of a wave: global MBs g .. end-1, side by side in one MB
 * row of one stream (end = 0: nothing staged yet) */
struct Tile {
    int g, end;
};

/* Stages the tile whose first MB is at (bx, by): rows by-15..by+30, dwords
 * c = 0..D-1 of frame columns bx-16+4c (D = 4 kTile + 8; 12 for the surface
 * variant).  All loads are issued before any wait: lane (c, rr) of the
 * first RPI x D lanes loads dword c of rows rr, rr + RPI, ... (RPI = 64 / D
 * rows per pass).  Tile dwords outside the frame come from a clamped
 * in-frame address (no in-limit candidate or filter tap reads them).  When
 * all 46 rows lie inside the frame, each staged row has a scalar base and
 * the lane a 32-bit offset. */
template <int kVariant, int W, int kTile, int kFirst, int kCount>
__device__ __forceinline__ void stage_passes(const __amdgpu_buffer_rsrc_t rsrc, bool inside, int by, int fx, int c,
                                             int rr, uint32_t *win)
{
    constexpr int H = W == 352 ? 288 : 144;
    constexpr int D = kVariant == kMeSurface ? kWinDw : tile_dw(kTile);
    constexpr int kPitch = win_pitch<kVariant, kTile>();
    constexpr int RPI = 64 / D;
    constexpr int kFull = kWinRows / RPI; /* passes with every row inside the tile */
    uint32_t staged[kCount];
    if (inside) {
        const uint32_t lo = (uint32_t)(rr * W + fx);
#pragma unroll
        for (int i = 0; i < kCount; i++) {
            const int t = kFirst + i;
            if (t < kFull) {
                staged[i] = __builtin_amdgcn_raw_buffer_load_b32(rsrc, lo, (by - 15 + RPI * t) * W, 0);
            } else {
                const int fy = by - 15 + min(RPI * t + rr, kWinRows - 1);
                staged[i] = __builtin_amdgcn_raw_buffer_load_b32(rsrc, (uint32_t)(fy * W + fx), 0, 0);
            }
        }
    } else {
#pragma unroll
        for (int i = 0; i < kCount; i++) {
            const int t = kFirst + i;
            const int fy = min(max(by - 15 + min(RPI * t + rr, kWinRows - 1), 0), H - 1);
            staged[i] = __builtin_amdgcn_raw_buffer_load_b32(rsrc, (uint32_t)(fy * W + fx), 0, 0);
        }
    }
    if (threadIdx.x < RPI * D) {
#pragma unroll
        for (int i = 0; i < kCount; i++) {
            const int t = kFirst + i;
            if (t < kFull || RPI * t + rr < kWinRows) win[(RPI * t + rr) * kPitch + c] = staged[i];
        }
    }
}

/* Stages the tile whose first MB is at (bx, by): rows by-15..by+30, dwords
 * c = 0..D-1 of frame columns bx-16+4c (D = 4 kTile + 8; 12 for the surface
 * variant).  Lane (c, rr) of the first RPI x D lanes loads dword c of rows
 * rr, rr + RPI, ... (RPI = 64 / D rows per pass); the loads of up to 12
 * passes are issued before any wait (a four-MB tile takes 23 passes, in
 * batches of 8, so that the staged values do not spill).  Tile dwords
 * outside the frame come from a clamped in-frame address (no in-limit
 * candidate or filter tap reads them).  When all 46 rows lie inside the
 * frame, each staged row has a scalar base and the lane a 32-bit offset. */
template <int kVariant, int W, int kTile>
__device__ __forceinline__ void stage_tile(const uint8_t *ref, int bx, int by, uint32_t *win)
{
    constexpr int H = W == 352 ? 288 : 144;
    constexpr int D = kVariant == kMeSurface ? kWinDw : tile_dw(kTile);
    constexpr int RPI = 64 / D;                        /* 5, 4 or 2 */
    constexpr int kIters = (kWinRows + RPI - 1) / RPI; /* 10, 12 or 23 */
    constexpr int kBatch = kIters <= 12 ? kIters : 8;
    const int lane = threadIdx.x;
    const int c = lane % D, rr = min(lane / D, RPI - 1);
    const int fx = min(max(bx - 16 + 4 * c, 0), W - 4);
    /* buffer loads over the luma plane: a 32-bit lane offset, and in the
     * interior case a scalar row offset (addresses stay in the plane, so its
     * range check never applies).  The plane's address is wave-uniform: said
     * so explicitly, else the resource is built per lane and each load loops
     * over lanes */
    const uint64_t ref_u = ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)((uintptr_t)ref >> 32)) << 32) |
                           (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(uintptr_t)ref);
    const __amdgpu_buffer_rsrc_t rsrc =
        __builtin_amdgcn_make_buffer_rsrc((void *)(uintptr_t)ref_u, 0, W * H, kBufferDword3);
    const bool inside = by >= 16 && by + 32 <= H;
    stage_passes<kVariant, W, kTile, 0, kBatch>(rsrc, inside, by, fx, c, rr, win);
    if constexpr (kIters > kBatch)
        stage_passes<kVariant, W, kTile, kBatch, (kIters - kBatch < kBatch ? kIters - kBatch : kBatch)>(rsrc, inside, by, fx, c, rr, win);
    if constexpr (kIters > 2 * kBatch)
        stage_passes<kVariant, W, kTile, 2 * kBatch, kIters - 2 * kBatch>(rsrc, inside, by, fx, c, rr, win);
    static_assert(kIters <= 3 * kBatch, "three batches at most");
}

/* SATD costs (inter, inter + loop filter, intra), the reference's mode
 * decision and the MB's result stores, once the vector (mvx, mvy) is found.
 * Lane (4x4 sub-block, row y) holds its 4 current pels in crow and, for the
 * prediction rows R - 1, R, R + 1 (R = the lane's MB row), pels -1..2 of its
 * 4-pel group in lo[i] and pels 3, 4 in the low bytes of hi[i]. */
__device__ __forceinline__ int me_finish(const MeLaunch &p, int g, const StreamDesc &d, uint32_t crow,
                                         const uint32_t (&lo)[3], const uint32_t (&hi)[3], int R, int sub_x, int yq,
                                         int mvx, int mvy, bool search)
{
    const int lane = threadIdx.x & 63;
    /* SATD costs (inter, inter + loop filter, intra); 4 pels per lane as
     * packed int16 pairs (p0, p2), (p1, p3) */
    const short2v ce = as_s2(__builtin_amdgcn_perm(crow, crow, 0x0c020c00u));
    const short2v co = as_s2(__builtin_amdgcn_perm(crow, crow, 0x0c030c01u));
    const int yl = R & 7;
    const bool row_border = yl == 0 || yl == 7;
    /* [1 2 1] along rows R - 1, R, R + 1 of the prediction as packed pairs
     * (h0, h2) and (h1, h3).  On a block's border rows the filtered values
     * are discarded below, so every lane reads the same three rows and no
     * lane branches */
    uint32_t h02[3], h13[3], pm02 = 0, pm13 = 0;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const uint32_t A = __builtin_amdgcn_perm(hi[i], lo[i], 0x0c020c00u); /* (p-1, p1) */
        const uint32_t B = __builtin_amdgcn_perm(hi[i], lo[i], 0x0c030c01u); /* (p0, p2) */
        const uint32_t C = __builtin_amdgcn_perm(hi[i], lo[i], 0x0c040c02u); /* (p1, p3) */
        const uint32_t D = __builtin_amdgcn_perm(hi[i], lo[i], 0x0c050c03u); /* (p2, p4) */
        h02[i] = A + 2 * B + C;
        h13[i] = B + 2 * C + D;
        if (i == 1) {
            pm02 = B;
            pm13 = C;
        }
    }
    /* 3x3 [1 2 1] filter, rounded, on the 6x6 interior of each 8x8 block
     * (megpu:833-879); pels on the block border keep the raw prediction */
    const uint32_t f02 = (((h02[0] + 2 * h02[1] + h02[2]) + 0x00080008u) >> 4) & 0x00ff00ffu;
    const uint32_t f13 = (((h13[0] + 2 * h13[1] + h13[2]) + 0x00080008u) >> 4) & 0x00ff00ffu;
    const uint32_t m02 = row_border ? 0xffffffffu : ((sub_x & 1) ? 0u : 0x0000ffffu);
    const uint32_t m13 = row_border ? 0xffffffffu : ((sub_x & 1) ? 0xffff0000u : 0u);
    const uint32_t fp02 = (pm02 & m02) | (f02 & ~m02), fp13 = (pm13 & m13) | (f13 & ~m13);
    const short2v s1 = (yq & 1) ? short2v{-1, -1} : short2v{1, 1};
    const short2v s2 = (yq & 2) ? short2v{-1, -1} : short2v{1, 1};
    const int t_inter = satd_quad(ce - as_s2(pm02), co - as_s2(pm13), s1, s2);
    const int t_filter = satd_quad(ce - as_s2(fp02), co - as_s2(fp13), s1, s2);
    const int t_intra = satd_quad(ce, co, s1, s2);
    const int c_inter = search ? wave_sum(t_inter) : 0;
    const int c_filter = search ? wave_sum(t_filter) : 0;
    const int c_intra = wave_sum(t_intra);

    /* h261_mode_decision (src/h261_moded.c:31-100), on wave-uniform values */
    const bool md_intra = d.frame_type == H261_FRAME_TYPE_INTRA || (c_intra <= c_inter && c_intra <= c_filter);
    const bool md_filter = !md_intra && c_filter < c_inter;
    int md_flags = md_filter ? (H261_MB_MC | H261_MB_FILTER) : 0;
    if (!md_intra && (mvx != 0 || mvy != 0)) md_flags |= H261_MB_MC;
    const int md_cost = md_intra ? c_intra : (md_filter ? c_filter : c_inter);

    /* results: MeResult (5 dwords) from lanes 0-4, MbMode (6 dwords) from
     * lanes 8-13 */
    if (lane < 5) {
        const uint32_t mv = (uint32_t)(uint8_t)(int8_t)mvx | ((uint32_t)(uint8_t)(int8_t)mvy << 8);
        const uint32_t w = lane == 0 ? (uint32_t)c_intra : lane == 1 ? 0u : lane == 2 ? (uint32_t)c_inter
                         : lane == 3 ? (uint32_t)c_filter : mv;
        ((uint32_t *)(p.result + g))[lane] = w;
    }
    if (p.mode && lane >= 8 && lane < 14) {
        /* MbMode {type, flags, mvx, mvy, quant_adjust, cost} */
        const int k = lane - 8;
        const int w = k == 0 ? (md_intra ? H261_MB_TYPE_INTRA : H261_MB_TYPE_INTER) : k == 1 ? md_flags
                    : k == 2 ? (md_intra ? 0 : mvx) : k == 3 ? (md_intra ? 0 : mvy) : k == 4 ? 0 : md_cost;
        ((int32_t *)(p.mode + g))[k] = w;
    }
    if (lane == 0 && p.mv_start) {
        p.mv_start[2 * g] = (int8_t)mvx;
        p.mv_start[2 * g + 1] = (int8_t)mvy;
    }
    return md_cost;
}

/* One macroblock of a W-pel-wide frame (352 CIF, 176 QCIF: every size and
 * divisor below is a compile-time constant, so MB positions, plane offsets
 * and the current MB's scalar row loads need no run-time arithmetic).  The
 * MB's window is the tile's columns from 16 (mx - tile.mx) pels on; a new
 * tile is staged when the MB lies outside the current one. */
template <int kVariant, int W, int kTile>
__device__ __forceinline__ int me_mb(const MeLaunch &p, int g, const StreamDesc &d, uint32_t *tile_win,
                                     uint32_t *curw, uint16_t *surf, Tile &tile)
{
    constexpr int H = W == 352 ? 288 : 144, MBW = W / 16, MBH = H / 16;
    constexpr int kPitch = win_pitch<kVariant, kTile>();
    const int lane = threadIdx.x;
    const int m = g - d.mb_base;
    const int mx = m % MBW, my = m / MBW;
    const int bx = 16 * mx, by = 16 * my;
    const uint8_t *cur = d.cur[0];
    const uint8_t *cur_mb = cur + by * W + bx;
    const int edges = (mx == 0 ? 1 : 0) | (mx == MBW - 1 ? 2 : 0) | (my == 0 ? 4 : 0) | (my == MBH - 1 ? 8 : 0);

    /* motion vector limits (gpu_device.cu:226-265) */
    const int lx0 = (edges & 1) ? 0 : -15, lx1 = (edges & 2) ? 0 : 15;
    const int ly0 = (edges & 4) ? 0 : -15, ly1 = (edges & 8) ? 0 : 15;

    /* SATD lanes: (4x4 sub-block, row y); this lane's 4 current pels, loaded
     * first so the load overlaps the staging */
    const int sub = lane >> 2, yq = lane & 3, sub_x = sub & 3, sub_y = sub >> 2;
    const int R = 4 * sub_y + yq; /* MB row */
    const uint32_t crow = load_global_u32(cur_mb + R * W + 4 * sub_x);
    const uint32_t cw = kVariant == kMeSurface ? load_global_u32(cur_mb + (lane >> 2) * W + 4 * (lane & 3)) : 0u;

    if (g >= tile.end) {
        if (tile.end) __syncthreads(); /* the previous MBs' reads of the tile are done */
        stage_tile<kVariant, W, kTile>(d.ref[0], bx, by, tile_win);
        tile = Tile{g, g + min(kTile, MBW - mx)};
        if (kVariant == kMeSurface) curw[lane] = cw;
        __syncthreads();
    }
    uint32_t *const win = tile_win + 4 * (g - tile.g);

    const bool search = !(p.skip_search_on_intra && d.frame_type == H261_FRAME_TYPE_INTRA);
    int mvx = 0, mvy = 0;
    if (search) {
        if (kVariant == kMeSurface) {
            sad_surface(win, curw, surf, lane);
            __syncthreads();
        }
        /* surface variant: the cost of an in-limit candidate is a lookup */
        auto cost_of = [&](int vx, int vy) -> uint32_t {
            if (!(vx >= lx0 && vx <= lx1 && vy >= ly0 && vy <= ly1)) return 65535u;
            return (uint32_t)surf[(vy + 15) * 32 + vx + 16];
        };
        /* stage 1: coarse grid, candidates lane and lane + 64 */
        uint32_t best = 0xffffffffu;
        if (kVariant == kMeSurface) {
            for (int ci = lane; ci < 121; ci += 64) {
                const int tx = ci % 11, ty = ci / 11;
                const uint32_t cost = cost_of(coarse_vec(tx), coarse_vec(ty));
                const uint32_t key = (cost << 8) | morton((uint32_t)tx, (uint32_t)ty);
                best = key < best ? key : best;
            }
        } else {
            /* every coarse candidate lies inside the window, so each is
             * evaluated where it is; the mask bit of an out-of-limit one (or
             * of a spare lane) replaces its cost by 65535 */
            const TwoStageTables &tab = c_two_stage[pitch_set(kPitch)];
            int off[2], sh[2];
            uint32_t e[2];
#pragma unroll
            for (int i = 0; i < 2; i++) {
                e[i] = tab.coarse.e[i][lane];
                off[i] = (int)(e[i] & 0xffffu) >> 2;
                sh[i] = (int)(e[i] & 3u);
            }
            uint32_t acc[2];
            sad16<2, kPitch>(win, cur_mb, W, off, sh, acc);
#pragma unroll
            for (int i = 0; i < 2; i++) {
                const uint32_t cost = __builtin_amdgcn_inverse_ballot_w64(tab.edge.m[edges][i]) ? acc[i] : 65535u;
                const uint32_t key = (cost << 8) | (e[i] >> 16);
                best = key < best ? key : best;
            }
        }
        best = wave_min_u32(best);
        int wtx, wty;
        unmorton(best & 0xffu, wtx, wty);
        const int sx = coarse_vec(wtx), sy = coarse_vec(wty);

        /* stage 2: 7x7 refine around the coarse winner (always in limits,
         * since (0, 0) is and no SAD reaches 65535) */
        if (kVariant == kMeSurface) {
            const int tx = (lane < 49 ? lane : 0) % 7, ty = (lane < 49 ? lane : 0) / 7;
            const uint32_t cost = cost_of(sx + tx - 3, sy + ty - 3);
            best = lane < 49 ? (cost << 8) | morton((uint32_t)tx, (uint32_t)ty) : 0xffffffffu;
        } else {
            /* lane (tx, ty) starts at the byte address of (sx - 3, sy - 3)
             * plus 60 ty + tx; out-of-limit candidates (at most 17 px out)
             * read the window's spare rows and are masked off, as are the
             * spare lanes */
            const uint64_t m2 = refine_mask(sx, sy, lx0, lx1, ly0, ly1);
            const uint32_t e2 = c_two_stage[pitch_set(kPitch)].refine.e[lane];
            const int a = 4 * kPitch * (sy + 12) + (sx + 13) + (int)(e2 & 0xffffu);
            int off[1] = {a >> 2}, sh[1] = {a & 3};
            uint32_t acc[1];
            sad16<1, kPitch>(win, cur_mb, W, off, sh, acc);
            const uint32_t cost = __builtin_amdgcn_inverse_ballot_w64(m2) ? acc[0] : 65535u;
            best = (cost << 8) | (e2 >> 16);
        }
        best = wave_min_u32(best);
        unmorton(best & 0xffu, wtx, wty);
        mvx = sx + wtx - 3;
        mvy = sy + wty - 3;
    }

    /* SATD rows: window row of the predicted MB row, byte of pel -1 of
     * this lane; rows -1 and 46 are spare window rows */
    const int wr = mvy + 15 + R;
    const int b0 = mvx + 16 + 4 * sub_x - 1;
    const int q = b0 >> 2, bsh = b0 & 3, q2 = min(q + 2, kWinDw - 1);
    uint32_t lo[3], hi[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const uint32_t *w = win + (wr - 1 + i) * kPitch;
        const uint32_t w0 = w[q], w1 = w[q + 1], w2 = w[q2];
        lo[i] = __builtin_amdgcn_alignbyte(w1, w0, bsh); /* pels -1, 0, 1, 2 */
        hi[i] = __builtin_amdgcn_alignbyte(w2, w1, bsh); /* pels 3, 4 */
    }
    return me_finish(p, g, d, crow, lo, hi, R, sub_x, yq, mvx, mvy, search);
}

/* The MB costs a wave found for one stream enter its rate control with one
 * atomic.  Fused tail: count and cost sum travel in one 64-bit atomic, and
 * the wave whose MBs complete the count holds the frame's cost sum and
 * starts the picture (rate control, quantiser); else the costs only
 * accumulate for the rate-control begin launch. */
__device__ __forceinline__ void add_costs(const MeLaunch &p, int s, const StreamDesc &d, int count, uint32_t cost)
{
    if ((threadIdx.x & 63) != 0 || count == 0 || !p.mode || !p.rc) return;
    RcState &rc = p.rc[s];
    if (p.fuse_tail) {
        const unsigned long long old = atomicAdd((unsigned long long *)&rc.me_acc, ((unsigned long long)count << 32) | cost);
        if ((int)(old >> 32) + count == d.mb_w * d.mb_h) {
            fc::rc_begin_one(rc, d, (int)((uint32_t)old + cost));
            atomicExch((unsigned long long *)&rc.me_acc, 0ull);
        }
    } else {
        atomicAdd(&rc.cost_accum, (int)cost);
    }
}

/* One wave searches kMbs consecutive MBs of the global MB space in turn (the
 * launch's MB range in runs of kMbs): the wave's start-up and the tail of
 * each MB's stores and atomics overlap the next MB's work, a run's MBs of
 * one stream add their costs with one atomic, and up to kTile MBs of a run
 * that lie side by side in one MB row search one staged tile. */
template <int kVariant, int kMbs, int kTile>
/* 8 waves per SIMD (64 VGPRs; 11-19 kernel arguments and descriptor fields
 * then live in VGPR lanes, about 15 v_readlane per MB): 2% faster than 7 */
__global__ __launch_bounds__(64, 8) void me_kernel(MeLaunch p)
{
    static_assert(kTile == 1 || (kVariant == kMeTwoStage && kTile <= kMbs), "tiles: two-stage search, within a run");
    __shared__ __align__(16) uint32_t winmem[win_alloc<kVariant, kTile>()];
    __shared__ __align__(16) uint32_t curw[kVariant == kMeSurface ? 64 : 4];
    __shared__ __align__(16) uint16_t surf[kVariant == kMeSurface ? 31 * 32 : 8];
    uint32_t *const win = winmem + win_base<kVariant, kTile>();

    const int runs = (p.n_mb - p.mb_first + kMbs - 1) / kMbs;
    const int r = launch_index(blockIdx.x, runs, p.xcd_remap);
    if (r >= runs) return;
    const int g0 = p.mb_first + r * kMbs;
    int seg = -1, count = 0;
    uint32_t cost = 0;
    const StreamDesc *seg_d = nullptr;
    Tile tile{0, 0};
#pragma unroll 1
    for (int i = 0; i < kMbs; i++) {
        const int g = g0 + i;
        if (g >= p.n_mb) break;
        const int s = p.mb_stream[g]; /* < 0: unmapped range of a closed stream */
        if (s != seg) {
            if (seg >= 0) add_costs(p, seg, *seg_d, count, cost);
            seg = s;
            count = 0;
            cost = 0;
            seg_d = s >= 0 ? &p.streams[s] : nullptr;
        }
        if (s < 0 || !(seg_d->flags & kStreamActive)) continue;
        cost += (uint32_t)(seg_d->width == 352 ? me_mb<kVariant, 352, kTile>(p, g, *seg_d, win, curw, surf, tile)
                                               : me_mb<kVariant, 176, kTile>(p, g, *seg_d, win, curw, surf, tile));
        count++;
    }
    if (seg >= 0) add_costs(p, seg, *seg_d, count, cost);
}

}  // namespace

int launch_me(const MeLaunch &p, void *stream, void *ev_start, void *ev_stop)
{
    if (p.n_mb - p.mb_first <= 0) {
        if (ev_start) { /* nothing to launch: an empty timed interval */
            (void)hipEventRecord((hipEvent_t)ev_start, (hipStream_t)stream);
            (void)hipEventRecord((hipEvent_t)ev_stop, (hipStream_t)stream);
        }
        return 0;
    }
    /* the surface variant searches one MB per wave (its LDS surface and
     * registers leave no room for a run); a tile is at most a run wide */
    const int mbs = p.variant == kMeSurface ? 1
                    : p.mbs_per_wave == 8 ? 8 : p.mbs_per_wave == 4 ? 4 : p.mbs_per_wave == 2 ? 2 : 1;
    const int tile = min(p.tile == 4 ? 4 : p.tile == 2 ? 2 : 1, mbs);
    const dim3 grid(launch_grid((p.n_mb - p.mb_first + mbs - 1) / mbs, p.xcd_remap));
    /* [log2 run][log2 tile] */
    static void (*const two_stage[4][3])(MeLaunch) = {
        {me_kernel<kMeTwoStage, 1, 1>, nullptr, nullptr},
        {me_kernel<kMeTwoStage, 2, 1>, me_kernel<kMeTwoStage, 2, 2>, nullptr},
        {me_kernel<kMeTwoStage, 4, 1>, me_kernel<kMeTwoStage, 4, 2>, me_kernel<kMeTwoStage, 4, 4>},
        {me_kernel<kMeTwoStage, 8, 1>, me_kernel<kMeTwoStage, 8, 2>, me_kernel<kMeTwoStage, 8, 4>},
    };
    auto k = p.variant == kMeSurface ? me_kernel<kMeSurface, 1, 1>
                                     : two_stage[mbs == 8 ? 3 : mbs == 4 ? 2 : mbs == 2 ? 1 : 0][tile == 4 ? 2 : tile - 1];
    if (ev_start)
        hipExtLaunchKernelGGL(k, grid, dim3(64), 0, (hipStream_t)stream, (hipEvent_t)ev_start, (hipEvent_t)ev_stop, 0,
                              p);
    else
        hipLaunchKernelGGL(k, grid, dim3(64), 0, (hipStream_t)stream, p);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace h261g
