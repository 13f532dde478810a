/*
 * bitstream.hip -- the H.261 bitstream writer on the device.
 *
 * After the frame tail has placed every coded MB of a frame (bit position
 * and the MBA / MTYPE / MQUANT / MVD / CBP codes in MbHdr, picture and GOB
 * headers already in the stream's words), a wave takes 8 consecutive MBs of
 * the global MB space, loads their coded blocks' levels into LDS in one
 * round, and writes each coded MB's bits: its header codes and the TCOEFF
 * layer of its coded blocks
 * (src/h261_macroblock.c:5-92): the intra DC as 8 bits, an inter block's
 * first coefficient +-1 as '1s', every other non-zero level as the run/level
 * VLC and its sign bit or as the 20-bit escape, EOB '10' after each block.
 *
 * Lane k of a block is transmission position k: a ballot of the non-zero
 * positions gives every lane its run, an exclusive prefix sum of the code
 * lengths its bit offset.  The MB's bits are assembled in LDS at the word
 * phase of its position, then stored: whole words inside the MB with plain
 * stores, the first and last words (shared with the neighbouring MBs or
 * headers) with atomic ORs.  Words hold stream bytes in order.
 */
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include "frame_control.h"
#include "h261_vlc.h"
#include "h261gpu_internal.h"

namespace h261g {

namespace {

__constant__ uint8_t c_zz[64] = H261G_ZIGZAG;
__constant__ Vlc c_tc[kTcoeffRuns][kTcoeffLevels] = H261G_TCOEFF_TABLE;

/* LDS words for one MB: its header (<= 64 bits), 6 blocks of at most
 * 8 + 63 * 20 + 2 bits, and the leading partial word */
constexpr int kMbWords = 256;
static_assert(kMbWords * 32 >= 31 + 64 + 6 * (8 + 63 * 20 + 2), "MB bit buffer");
/* consecutive global MBs per wave: their levels arrive in one round of
 * loads (16-byte pieces of coded blocks only) */
constexpr int kMbsPerWave = 8;
constexpr int kPieces = kMbsPerWave * 384 / 16;
constexpr int kPieceRounds = (kPieces + 63) / 64;
/* independent waves per workgroup: a compute unit holds a limited number of
 * workgroups, so one-wave workgroups would cap the waves resident per SIMD */
constexpr int kWavesPerGroup = 4;
constexpr int kTc = kTcoeffRuns * kTcoeffLevels;
constexpr int kTcRounds = (kTc + 64 * kWavesPerGroup - 1) / (64 * kWavesPerGroup);

/* ORs the low len (<= 32) bits of code into LDS words at bit position pos,
 * first bit most significant */
__device__ __forceinline__ void lds_or(uint32_t *buf, int pos, uint32_t code, int len)
{
    if (len <= 0) return;
    const int i = pos >> 5, off = pos & 31;
    if (off + len <= 32) {
        atomicOr(buf + i, code << (32 - off - len));
    } else {
        const int spill = off + len - 32;
        atomicOr(buf + i, code >> spill);
        atomicOr(buf + i + 1, code << (32 - spill));
    }
}

/* each wave codes into its own LDS buffer, and a wave's LDS operations
 * complete in issue order: ordering the accesses of different lanes needs
 * only a compiler barrier, not __syncthreads, whose fence would also wait
 * for every global store and atomic of the wave to complete */
__device__ __forceinline__ void wave_lds_order() { __builtin_amdgcn_wave_barrier(); }

typedef __attribute__((address_space(1))) uint32_t global_u32;

/* ORs code (len <= 22 bits, len 0 allowed) into LDS words at bit position
 * pos with two unconditional ds_or: a part that does not exist goes to the
 * lane's own spare word, so lanes without a code neither branch nor contend
 * for one address */
__device__ __forceinline__ void or_code(uint32_t *buf, int pos, uint32_t code, int len, int lane)
{
    const int i = pos >> 5, end = (pos & 31) + len;
    const bool split = end > 32;
    const uint32_t hi = split ? (code >> ((end - 32) & 31)) : (code << ((32 - end) & 31));
    const uint32_t lo = code << ((64 - end) & 31);
    atomicOr(buf + (len ? i : kMbWords + lane), hi);
    atomicOr(buf + (split ? i + 1 : kMbWords + lane), split ? lo : 0u);
}

__device__ __forceinline__ int rl(int v, int lane) { return __builtin_amdgcn_readlane(v, lane); }

/* coded blocks of an MB as bits 32 >> k for block k */
__device__ __forceinline__ int coded_blocks(const MbSide &sd)
{
    if (sd.type == H261_MB_TYPE_INTRA) return 63;
    return (sd.flags & H261_MB_COEFF) ? ((int)(uint8_t)sd.cbp & 63) : 0;
}

__global__ __launch_bounds__(64 * kWavesPerGroup) void bitstream_kernel(BitLaunch p)
{
    __shared__ uint32_t buf_all[kWavesPerGroup][kMbWords + 64]; /* + a spare word per lane */
    __shared__ uint4 lvs_all[kWavesPerGroup][kPieces]; /* each wave's MBs' levels, [mb][6][64] */
    __shared__ uint32_t tc[kTc];                        /* run/level table: code | len << 16 */
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int n_waves = (p.n_mb + kMbsPerWave - 1) / kMbsPerWave;
    const int n_groups = (n_waves + kWavesPerGroup - 1) / kWavesPerGroup;
    const int grp = launch_index(blockIdx.x, n_groups, p.xcd_remap);
    if (grp >= n_groups) return; /* the whole workgroup */
    const int g0 = (grp * kWavesPerGroup + wid) * kMbsPerWave;
    uint32_t *const buf = buf_all[wid];
    uint4 *const lvs = lvs_all[wid];

    /* every load is issued before its value is needed: two rounds of
     * memory latency per wave (places and side information, then levels
     * and bitstream words), none inside the coding loop */
    static_assert(sizeof(Vlc) == 4, "one word per table entry");
    const uint32_t *tcg = (const uint32_t *)&c_tc[0][0];
    uint32_t tcv[kTcRounds];
#pragma unroll
    for (int i = 0; i < kTcRounds; i++) {
        const int k = threadIdx.x + 64 * kWavesPerGroup * i;
        tcv[i] = tcg[k < kTc ? k : 0];
    }
    const int zz = c_zz[lane];
    /* lanes 0..7: the MBs' places, side information, streams (other lanes,
     * and waves past the last MB, read a clamped index and discard it) */
    const int gl = g0 + (lane < kMbsPerWave ? lane : 0);
    const bool mine = lane < kMbsPerWave && gl < p.n_mb;
    const int gc = gl < p.n_mb ? gl : p.n_mb - 1;
    const int s = p.mb_stream[gc];
    const MbHdr h = p.hdr[gc];
    const MbSide sd = p.side[gc];
    const int valid = (mine && s >= 0 && h.valid) ? 1 : 0;
    const int cb = valid ? coded_blocks(sd) : 0;
    const int hlo = (int)(uint32_t)h.code, hhi = (int)(uint32_t)(h.code >> 32);
    const int bitpos = h.bitpos, hlen = h.len;
    const int intra = sd.type == H261_MB_TYPE_INTRA;
    uint64_t vmask = __ballot(valid);
    /* the coded blocks' levels and each MB's bitstream words */
    const SeqState &q = p.seq[valid ? s : 0];
    uint32_t *const words = q.bits;
    const int cap = q.bits_words;
    if (valid) p.hdr[gc].valid = 0; /* placed again by the next frame tail */
    const uint4 *src = (const uint4 *)(p.levels + (size_t)g0 * 384);
    uint4 v[kPieceRounds];
#pragma unroll
    for (int j = 0; j < kPieceRounds; j++) {
        const int i = lane + 64 * j;
        const int m = i < kPieces ? i / 24 : 0, k = (i % 24) / 4;
        v[j] = make_uint4(0, 0, 0, 0);
        if (i < kPieces && (__shfl(cb, m, 64) & (32 >> k))) v[j] = src[i];
    }
    /* the run/level table in LDS: a lookup per coefficient stays off the
     * vector memory path */
#pragma unroll
    for (int i = 0; i < kTcRounds; i++) {
        const int k = threadIdx.x + 64 * kWavesPerGroup * i;
        if (k < kTc) tc[k] = tcv[i];
    }
#pragma unroll
    for (int i = 0; i < kMbWords / 64; i++) buf[lane + 64 * i] = 0u;
#pragma unroll
    for (int j = 0; j < kPieceRounds; j++)
        if (lane + 64 * j < kPieces) lvs[lane + 64 * j] = v[j];
    __syncthreads(); /* the table from every wave */

    const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
    const int8_t *lvb = (const int8_t *)lvs;
    while (vmask) {
        const int m = __builtin_ctzll(vmask);
        vmask &= vmask - 1;
        const int pos0 = rl(bitpos, m), len0 = rl(hlen, m), cbm = rl(cb, m);
        const bool intra_mb = rl(intra, m) != 0;
        const int p0 = pos0 & 31;
        if (lane == 0) {
            const uint32_t lo = (uint32_t)rl(hlo, m), hi = (uint32_t)rl(hhi, m);
            if (len0 > 32) {
                lds_or(buf, p0, hi, len0 - 32);
                lds_or(buf, p0 + len0 - 32, lo, 32);
            } else {
                lds_or(buf, p0, lo, len0);
            }
        }
        int pos = p0 + len0;
        int raws[6];
#pragma unroll
        for (int k = 0; k < 6; k++) raws[k] = lvb[m * 384 + k * 64 + zz];
        const bool dc = intra_mb && lane == 0;
#pragma unroll
        for (int k = 0; k < 6; k++) {
            if (!(cbm & (32 >> k))) continue;
            /* branch-free: every lane computes every candidate code and
             * selects (lanes of a wave diverge here on every block) */
            const int raw = raws[k];
            const int l = dc ? (raw & 0xff) : raw; /* the intra DC is an unsigned 8-bit code */
            const bool nzl = l != 0 || dc;
            const uint64_t bl = __ballot(nzl) & below;
            /* run: zero positions since the previous non-zero one (an intra
             * DC counts as non-zero; an inter block starts before position 0) */
            const int run = lane - (bl ? 63 - __builtin_clzll(bl) : -1) - 1;
            const int a = l < 0 ? -l : l;
            const uint32_t sign = l < 0 ? 1u : 0u;
            const bool in_range = run < kTcoeffRuns && a < kTcoeffLevels;
            const uint32_t e = tc[in_range ? run * kTcoeffLevels + a : 0]; /* entry 0 has length 0 */
            const int tlen = (int)((e >> 16) & 0xffu);
            /* escape (000001, 6-bit run, 8-bit level), else the table code
             * and its sign, else an inter block's first +-1 as '1s', else
             * the intra DC as 8 bits */
            uint32_t code = tlen ? (((e & 0xffffu) << 1) | sign) : ((1u << 14) | ((uint32_t)run << 8) | ((uint32_t)l & 0xffu));
            int len = tlen ? tlen + 1 : 20;
            const bool first = lane == 0 && a == 1;
            code = first ? (2u | sign) : code;
            len = first ? 2 : len;
            code = dc ? (uint32_t)l : code;
            len = dc ? 8 : len;
            code = nzl ? code : 0u;
            len = nzl ? len : 0;
            /* EOB '10' closes the block: appended to lane 63's code */
            code = lane == 63 ? ((code << 2) | 2u) : code;
            len = lane == 63 ? len + 2 : len;
            const int incl = fc::wave_inclusive_sum(len, lane);
            or_code(buf, pos + incl - len, code, len, lane);
            pos += __builtin_amdgcn_readlane(incl, 63);
        }
        wave_lds_order();
        /* words inside the MB are its own: plain stores; the first and last
         * are shared with neighbours: atomic ORs.  The buffer is left zero. */
        const int n = (pos + 31) >> 5;
        const int w0 = pos0 >> 5;
        /* a global-address-space pointer: flat stores would also count
         * against the LDS wait counter and make every LDS wait in the loop
         * wait for them */
        global_u32 *dst = (global_u32 *)(((uint64_t)(uint32_t)rl((int)(uint32_t)(uint64_t)words, m)) |
                                         ((uint64_t)(uint32_t)rl((int)(uint32_t)((uint64_t)words >> 32), m) << 32));
        const bool fits = w0 + n <= rl(cap, m);
        for (int i = lane; i < n; i += 64) {
            const uint32_t v = __builtin_bswap32(buf[i]);
            buf[i] = 0u;
            if (!fits) continue;
            if (i == 0 || i == n - 1) {
                if (v) __hip_atomic_fetch_or(dst + w0 + i, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            } else {
                dst[w0 + i] = v;
            }
        }
        wave_lds_order();
    }
}

}  // namespace

int launch_bitstream(const BitLaunch &p, void *stream, void *ev_start, void *ev_stop)
{
    if (p.n_mb <= 0) {
        if (ev_start) (void)hipEventRecord((hipEvent_t)ev_start, (hipStream_t)stream);
        if (ev_stop) (void)hipEventRecord((hipEvent_t)ev_stop, (hipStream_t)stream);
        return 0;
    }
    const int n_waves = (p.n_mb + kMbsPerWave - 1) / kMbsPerWave;
    const int n_groups = (n_waves + kWavesPerGroup - 1) / kWavesPerGroup;
    const dim3 grid(launch_grid(n_groups, p.xcd_remap));
    if (ev_start || ev_stop)
        hipExtLaunchKernelGGL(bitstream_kernel, grid, dim3(64 * kWavesPerGroup), 0, (hipStream_t)stream, (hipEvent_t)ev_start,
                              (hipEvent_t)ev_stop, 0, p);
    else
        hipLaunchKernelGGL(bitstream_kernel, grid, dim3(64 * kWavesPerGroup), 0, (hipStream_t)stream, p);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace h261g
