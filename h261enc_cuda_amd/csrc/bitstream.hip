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

#include <stdio.h>

#include <mutex>

#include "frame_control.h"
#include "h261_vlc.h"
#include "h261gpu_internal.h"

namespace h261g {

namespace {

__constant__ uint8_t c_zz[64] = H261G_ZIGZAG;
/* The run/level table compacted: word 0 has length 0 (no code); then, run
 * by run, the codes of levels 1..max as code | len << 16; then, per run,
 * its first code's word | its largest level << 8.  About 90 words instead
 * of the 432 of the full table, so less LDS per wave. */
constexpr Vlc kTcFull[kTcoeffRuns][kTcoeffLevels] = H261G_TCOEFF_TABLE;
constexpr int tc_codes()
{
    int n = 0;
    for (int r = 0; r < kTcoeffRuns; r++)
        for (int l = 1; l < kTcoeffLevels; l++) n += kTcFull[r][l].len ? 1 : 0;
    return n;
}
constexpr int kTcCodes = tc_codes();
constexpr int kTcRunBase = 1 + kTcCodes;
constexpr int kTc = kTcRunBase + kTcoeffRuns;
struct TcCompact {
    uint32_t w[kTc];
};
/* a run's codes must be levels 1..max without gaps for the compact form */
constexpr bool tc_contiguous()
{
    for (int r = 0; r < kTcoeffRuns; r++) {
        bool gap = false;
        for (int l = 1; l < kTcoeffLevels; l++) {
            if (!kTcFull[r][l].len) gap = true;
            else if (gap) return false;
        }
    }
    return kTcFull[0][0].len == 0;
}
static_assert(tc_contiguous(), "TCOEFF codes of a run cover levels 1..max");
constexpr TcCompact make_tc()
{
    TcCompact t{};
    int k = 1;
    for (int r = 0; r < kTcoeffRuns; r++) {
        const int first = k;
        int maxl = 0;
        for (int l = 1; l < kTcoeffLevels; l++)
            if (kTcFull[r][l].len) {
                t.w[k++] = (uint32_t)kTcFull[r][l].code | ((uint32_t)kTcFull[r][l].len << 16);
                maxl = l;
            }
        t.w[kTcRunBase + r] = (uint32_t)first | ((uint32_t)maxl << 8);
    }
    return t;
}
__constant__ TcCompact c_tc = make_tc();

/* LDS words of one MB's bits: its header (<= 64 bits), 6 blocks of at most
 * 8 + 63 * 20 + 2 bits, and the leading partial word */
constexpr int kMbWords = 248;
static_assert(kMbWords * 32 >= 31 + 64 + 6 * (8 + 63 * 20 + 2) + 32, "MB bit buffer");
/* consecutive global MBs per wave: their levels arrive in one round of
 * loads (16-byte pieces of coded blocks only) */
constexpr int kMbsPerWave = 4;
constexpr int kPieces = kMbsPerWave * 384 / 16;
constexpr int kPieceRounds = (kPieces + 63) / 64;
/* independent waves per workgroup (each with its own LDS buffers; measured
 * at 512 streams: 1 wave 38 us, 2 or 4 waves 40 us, 8 waves 54 us per
 * launch at full load) */
constexpr int kWavesPerGroup = 1;
constexpr int kTcRounds = (kTc + 64 * kWavesPerGroup - 1) / (64 * kWavesPerGroup);
/* non-zero coefficients buffered before they are coded (a block is never
 * split between two flushes) */
constexpr int kEntryCap = 128;

/* a buffered coefficient: transmission position, level byte and flags */
enum : uint32_t {
    kEntFirst = 1u << 6,   /* first of its block */
    kEntLast = 1u << 7,    /* last of its block: EOB follows */
    kEntDc = 1u << 8,      /* intra DC: 8-bit fixed length code */
    kEntEmpty = 1u << 9,   /* coded block without coefficients: EOB only */
    kEntMbFirst = 1u << 12 /* first of its MB */
};
__device__ __forceinline__ uint32_t ent_mb(uint32_t e) { return (e >> 10) & 3u; }

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

/* each wave codes into its own LDS buffers, and a wave's LDS operations
 * complete in issue order: ordering the accesses of different lanes needs
 * only a compiler barrier, not __syncthreads, whose fence would also wait
 * for every global store and atomic of the wave to complete */
__device__ __forceinline__ void wave_lds_order() { __builtin_amdgcn_wave_barrier(); }

typedef __attribute__((address_space(1))) uint32_t global_u32;

__device__ __forceinline__ int rl(int v, int lane) { return __builtin_amdgcn_readlane(v, lane); }

/* coded blocks of an MB as bits 32 >> k for block k */
__device__ __forceinline__ int coded_blocks(const MbSide &sd)
{
    if (sd.type == H261_MB_TYPE_INTRA) return 63;
    return (sd.flags & H261_MB_COEFF) ? ((int)(uint8_t)sd.cbp & 63) : 0;
}

/* ORs code (len <= 22 bits, len 0 allowed) into LDS words at word w, bit
 * offset off, with two unconditional ds_or: a part that does not exist goes
 * to the lane's own spare word, so lanes without a code neither branch nor
 * contend for one address */
__device__ __forceinline__ void or_code(uint32_t *lds, int w, int off, uint32_t code, int len, int spare)
{
    const int end = off + len;
    const bool split = end > 32;
    const uint32_t hi = split ? (code >> ((end - 32) & 31)) : (code << ((32 - end) & 31));
    const uint32_t lo = code << ((64 - end) & 31);
    atomicOr(lds + (len ? w : spare), hi);
    atomicOr(lds + (split ? w + 1 : spare), split ? lo : 0u);
}

struct WaveLds {
    uint32_t region[kMbsPerWave][kMbWords]; /* each MB's bits from its word phase */
    uint32_t spare[64];
    uint32_t ent[kEntryCap + 64];           /* buffered coefficients (+ a discard slot per lane) */
    int32_t mb_off[kMbsPerWave];            /* bit position in the MB's region minus the running code sum */
    uint4 lvs[kPieces];                     /* the MBs' levels, [mb][6][64] */
};

/* codes the n buffered coefficients: VLC and sign or escape, EOB after a
 * block's last, prefix sum of the lengths, OR into the MB regions.  *sum is
 * the running sum of code lengths over everything flushed before. */
__device__ __forceinline__ void flush_entries(WaveLds &w, const uint32_t *tc, int n, int &sum, int lane)
{
    for (int base = 0; base < n; base += 64) {
        const int j = base + lane;
        const bool has = j < n;
        const uint32_t e = has ? w.ent[j] : 0u;
        const uint32_t ep = j > 0 ? w.ent[j - 1] : 0u;
        const int pos = (int)(e & 63u);
        const int m = (int)ent_mb(e);
        const int prev = (e & kEntFirst) ? -1 : (int)(ep & 63u);
        const int run = pos - prev - 1;
        const bool dc = (e & kEntDc) != 0;
        const int l = dc ? (int)((e >> 16) & 0xffu) : (int)(int8_t)((e >> 16) & 0xffu);
        const int a = l < 0 ? -l : l;
        const uint32_t sign = l < 0 ? 1u : 0u;
        const bool run_ok = run >= 0 && run < kTcoeffRuns;
        const uint32_t ri = tc[kTcRunBase + (run_ok ? run : 0)];
        const bool in_range = run_ok && a >= 1 && a <= (int)(ri >> 8);
        const uint32_t te = tc[in_range ? (int)(ri & 0xffu) + a - 1 : 0]; /* word 0 has length 0 */
        const int tlen = (int)((te >> 16) & 0xffu);
        /* escape (000001, 6-bit run, 8-bit level), else the table code and
         * its sign, else an inter block's first coefficient +-1 at position
         * 0 as '1s', else the intra DC as 8 bits */
        uint32_t code = tlen ? (((te & 0xffffu) << 1) | sign)
                             : ((1u << 14) | ((uint32_t)(run & 63) << 8) | ((uint32_t)l & 0xffu));
        int len = tlen ? tlen + 1 : 20;
        const bool one = pos == 0 && !dc && a == 1;
        code = one ? (2u | sign) : code;
        len = one ? 2 : len;
        code = dc ? (uint32_t)l : code;
        len = dc ? 8 : len;
        const bool none = !has || (e & kEntEmpty);
        code = none ? 0u : code;
        len = none ? 0 : len;
        /* EOB '10' after a block's last coefficient */
        const bool last = has && (e & kEntLast);
        code = last ? ((code << 2) | 2u) : code;
        len = last ? len + 2 : len;
        const int incl = fc::wave_inclusive_sum(len, lane);
        const int at_sum = sum + incl - len; /* running sum before this code */
        /* the MB's first coefficient fixes its offset between the running
         * sum and its region */
        if (has && (e & kEntMbFirst)) w.mb_off[m] += -at_sum;
        wave_lds_order();
        const int at = has ? w.mb_off[m] + at_sum : 0;
        or_code(&w.region[0][0], m * kMbWords + (at >> 5), at & 31, code, len,
                (int)((&w.spare[lane]) - &w.region[0][0]));
        sum += __builtin_amdgcn_readlane(incl, 63);
        wave_lds_order();
    }
}

__global__ __launch_bounds__(64 * kWavesPerGroup) void bitstream_kernel(BitLaunch p)
{
    __shared__ WaveLds lds_all[kWavesPerGroup];
    __shared__ uint32_t tc[kTc]; /* the compacted run/level table */
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int n_waves = (p.n_mb - p.mb_first + kMbsPerWave - 1) / kMbsPerWave;
    const int n_groups = (n_waves + kWavesPerGroup - 1) / kWavesPerGroup;
    const int grp = launch_index(blockIdx.x, n_groups, p.xcd_remap);
    if (grp >= n_groups) return; /* the whole workgroup */
    const int g0 = p.mb_first + (grp * kWavesPerGroup + wid) * kMbsPerWave;
    WaveLds &w = lds_all[wid];

    /* every load is issued before its value is needed: two rounds of
     * memory latency per wave (places and side information, then levels
     * and bitstream words), none inside the coding loop */
    const uint32_t *tcg = c_tc.w;
    uint32_t tcv[kTcRounds];
#pragma unroll
    for (int i = 0; i < kTcRounds; i++) {
        const int k = threadIdx.x + 64 * kWavesPerGroup * i;
        tcv[i] = tcg[k < kTc ? k : 0];
    }
    const int zz = c_zz[lane];
    /* lanes 0..3: the MBs' places, side information, streams (other lanes,
     * and waves past the last MB, read a clamped index and discard it) */
    const int gl = g0 + (lane < kMbsPerWave ? lane : 0);
    const bool mine = lane < kMbsPerWave && gl < p.n_mb;
    const int gc = gl < p.n_mb ? gl : p.n_mb - 1;
    const int s = p.mb_stream[gc];
    const MbHdr h = p.hdr[gc];
    const MbSide sd = p.side[gc];
    const int valid = (mine && s >= 0 && h.valid) ? 1 : 0;
    const int cb = valid ? coded_blocks(sd) : 0;
    const int intra = sd.type == H261_MB_TYPE_INTRA;
    const uint64_t vmask = __ballot(valid);
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
    for (int i = lane; i < kMbsPerWave * kMbWords; i += 64) (&w.region[0][0])[i] = 0u;
    if (lane < kMbsPerWave) w.mb_off[lane] = (h.bitpos & 31) + h.len; /* where the TCOEFF layer starts */
#pragma unroll
    for (int j = 0; j < kPieceRounds; j++)
        if (lane + 64 * j < kPieces) w.lvs[lane + 64 * j] = v[j];
    __syncthreads(); /* the table from every wave */

    /* the MB headers, lane m for MB m */
    if (valid) {
        const int p0 = h.bitpos & 31, len0 = h.len;
        const uint32_t lo = (uint32_t)h.code, hi = (uint32_t)(h.code >> 32);
        uint32_t *reg = w.region[lane];
        if (len0 > 32) {
            lds_or(reg, p0, hi, len0 - 32);
            lds_or(reg, p0 + len0 - 32, lo, 32);
        } else {
            lds_or(reg, p0, lo, len0);
        }
    }

    /* the non-zero coefficients of every coded block in bitstream order,
     * compacted: a lane per coefficient when coded */
    const int8_t *lvb = (const int8_t *)w.lvs;
    /* every block's level at this lane's transmission position, read once */
    int raw_all[kMbsPerWave][6];
#pragma unroll
    for (int m = 0; m < kMbsPerWave; m++)
#pragma unroll
        for (int k = 0; k < 6; k++) raw_all[m][k] = lvb[m * 384 + k * 64 + zz];
    int n_ent = 0, sum = 0;
#pragma unroll
    for (int m = 0; m < kMbsPerWave; m++) {
        if (!((vmask >> m) & 1)) continue;
        const int cbm = rl(cb, m);
        const bool intra_mb = rl(intra, m) != 0;
        uint32_t mb_first = kEntMbFirst;
        const int(&raws)[6] = raw_all[m];
        const bool dc = intra_mb && lane == 0;
#pragma unroll
        for (int k = 0; k < 6; k++) {
            if (!(cbm & (32 >> k))) continue;
            if (n_ent + 64 > kEntryCap) {
                wave_lds_order();
                flush_entries(w, tc, n_ent, sum, lane);
                n_ent = 0;
            }
            const int raw = raws[k];
            const int l = dc ? (raw & 0xff) : raw; /* the intra DC is an unsigned 8-bit code */
            const bool nzl = l != 0 || dc;
            const uint64_t nz = __ballot(nzl);
            const int cnt = __builtin_popcountll(nz);
            const uint32_t tag = ((uint32_t)m << 10) | mb_first;
            if (cnt == 0) {
                if (lane == 0) w.ent[n_ent] = tag | kEntFirst | kEntLast | kEntEmpty;
                n_ent += 1;
            } else {
                const int rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(nz >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)nz, 0u));
                const uint32_t e = (uint32_t)lane | tag | (rank == 0 ? kEntFirst : 0u) |
                                   (rank == cnt - 1 ? kEntLast : 0u) | (dc ? kEntDc : 0u) |
                                   (((uint32_t)l & 0xffu) << 16);
                /* the first entry alone carries the MB-first flag */
                w.ent[nzl ? n_ent + rank : kEntryCap + lane] = rank == 0 ? e : (e & ~kEntMbFirst);
                n_ent += cnt;
            }
            mb_first = 0;
        }
    }
    if (n_ent) {
        wave_lds_order();
        flush_entries(w, tc, n_ent, sum, lane);
    }
    wave_lds_order();

    /* the MBs' words, all MBs at once: lane i stores word i of the MBs'
     * words laid end to end.  Inside an MB a word is its own (plain store);
     * its first and last are shared with neighbours (atomic OR).  An MB whose
     * frame outgrew the buffer stores nothing (written on the host). */
    int nw = 0;
    if (valid) {
        const int n = ((h.bitpos & 31) + h.bits + 31) >> 5;
        nw = (h.bitpos >> 5) + n <= cap ? n : 0;
    }
    const uint64_t dst_mb = (uint64_t)(words + (h.bitpos >> 5));
    int n_of[kMbsPerWave], o_of[kMbsPerWave];
    uint64_t d_of[kMbsPerWave];
    int total = 0;
#pragma unroll
    for (int m = 0; m < kMbsPerWave; m++) {
        n_of[m] = rl(nw, m);
        o_of[m] = total;
        total += n_of[m];
        d_of[m] = (uint64_t)(uint32_t)rl((int)(uint32_t)dst_mb, m) |
                  ((uint64_t)(uint32_t)rl((int)(uint32_t)(dst_mb >> 32), m) << 32);
    }
    for (int i = lane; i < total; i += 64) {
        int m = 0;
#pragma unroll
        for (int k = 1; k < kMbsPerWave; k++) m += i >= o_of[k] ? 1 : 0;
        int n = n_of[0], o = o_of[0];
        uint64_t d = d_of[0];
#pragma unroll
        for (int k = 1; k < kMbsPerWave; k++) {
            n = m == k ? n_of[k] : n;
            o = m == k ? o_of[k] : o;
            d = m == k ? d_of[k] : d;
        }
        const int k = i - o;
        /* a global-address-space pointer: flat stores would also count
         * against the LDS wait counter */
        global_u32 *dst = (global_u32 *)d;
        const uint32_t val = __builtin_bswap32(w.region[m][k]);
        if (k == 0 || k == n - 1) {
            if (val) __hip_atomic_fetch_or(dst + k, val, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            dst[k] = val;
        }
    }
}

/* sizes and places of fetched bitstreams, one workgroup: stream ids[i]'s
 * last frame is on the device when it fit the stream's buffer.  Places are
 * 4-byte aligned, a running scan over chunks of the list; a frame that is
 * not on the device, or whose place ends past the host buffer's capacity,
 * gets size -1 (written on the host instead).  meta (host memory): sizes[n],
 * then offsets[n + 1], offsets[n] being the bytes all frames would take. */
/* one wave: it needs only one wave slot, so it is dispatched at once even
 * while other work fills the device */
__global__ __launch_bounds__(64) void fetch_plan_kernel(const int32_t *ids, int n, const StreamDesc *desc,
                                                        int n_slots, const SeqState *seq, const RcState *rc,
                                                        int32_t *dev_size, int32_t *dev_off, int32_t *meta,
                                                        int64_t data_cap)
{
    const int lane = threadIdx.x;
    int carry = 0;
    for (int base = 0; base < n; base += 64) {
        const int i = base + lane;
        int size = -1;
        /* an id outside the slots or of a slot with no open stream (its
         * descriptor is zero) has nothing to fetch */
        const int s = i < n ? ids[i] : -1;
        if (s >= 0 && s < n_slots && desc[s].mb_w > 0) {
            const SeqState &q = seq[s];
            const RcState &r = rc[s];
            if (q.bits && r.frame_bits > 0 && (r.frame_bits + 31) / 32 <= q.bits_words) size = r.frame_bytes;
        }
        const int span = size > 0 ? (size + 3) & ~3 : 0;
        const int incl = fc::wave_inclusive_sum(span, lane);
        const int off = carry + incl - span;
        if ((int64_t)off + span > data_cap) size = -1;
        if (i < n) {
            dev_size[i] = size;
            dev_off[i] = off;
            meta[i] = size;
            meta[n + i] = off;
        }
        carry += __builtin_amdgcn_readlane(incl, 63);
    }
    if (lane == 0) meta[2 * n] = carry;
}

/* one workgroup per listed stream: its frame's words into host memory */
__global__ __launch_bounds__(256) void fetch_copy_kernel(const int32_t *ids, const SeqState *seq,
                                                         const int32_t *dev_size, const int32_t *dev_off,
                                                         uint32_t *host_data)
{
    const int i = blockIdx.x;
    const int size = dev_size[i];
    if (size <= 0) return;
    const uint32_t *src = seq[ids[i]].bits;
    uint32_t *dst = host_data + dev_off[i] / 4;
    const int words = (size + 3) / 4;
    for (int k = threadIdx.x; k < words; k += blockDim.x) dst[k] = src[k];
}

}  // namespace

int launch_bitstream(const BitLaunch &p, void *stream, void *ev_start, void *ev_stop)
{
    if (p.n_mb - p.mb_first <= 0) {
        if (ev_start) (void)hipEventRecord((hipEvent_t)ev_start, (hipStream_t)stream);
        if (ev_stop) (void)hipEventRecord((hipEvent_t)ev_stop, (hipStream_t)stream);
        return 0;
    }
    const int n_waves = (p.n_mb - p.mb_first + kMbsPerWave - 1) / kMbsPerWave;
    const int n_groups = (n_waves + kWavesPerGroup - 1) / kWavesPerGroup;
    const dim3 grid(launch_grid(n_groups, p.xcd_remap));
    if (ev_start || ev_stop)
        hipExtLaunchKernelGGL(bitstream_kernel, grid, dim3(64 * kWavesPerGroup), 0, (hipStream_t)stream, (hipEvent_t)ev_start,
                              (hipEvent_t)ev_stop, 0, p);
    else
        hipLaunchKernelGGL(bitstream_kernel, grid, dim3(64 * kWavesPerGroup), 0, (hipStream_t)stream, p);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

/* Scratch of the allocation-free fetch: per listed stream its frame size
 * (first kFetchIntoMaxStreams entries) and its place (next
 * kFetchIntoMaxStreams).  32 KiB, held by every device the library runs on
 * until the process ends.  The plan kernel writes every entry the copy kernel
 * reads, so nothing needs clearing between calls; calls that overlap on one
 * device share it and race.  Host code never passes the array's name to a
 * kernel (on the host it is only a stand-in): it passes the device address
 * from fetch_into_scratch. */
static __device__ __align__(256) int32_t scratch[2 * kFetchIntoMaxStreams];

int32_t *fetch_into_scratch()
{
    constexpr int kMaxDevices = 64;
    static std::mutex mu;
    static int32_t *addr[kMaxDevices] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[dev]) {
        void *p = nullptr;
        const hipError_t err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (err != hipSuccess) {
            fprintf(stderr, "h261gpu: hipGetSymbolAddress(scratch): %s\n", hipGetErrorString(err));
            return nullptr;
        }
        addr[dev] = (int32_t *)p;
    }
    return addr[dev];
}

int launch_fetch(const int32_t *ids, int n, const StreamDesc *desc, int n_slots, const SeqState *seq, const RcState *rc,
                 int32_t *dev_size, int32_t *dev_off, int32_t *meta, uint8_t *host_data, int64_t data_cap, void *stream)
{
    if (n <= 0) return 0;
    hipLaunchKernelGGL(fetch_plan_kernel, dim3(1), dim3(64), 0, (hipStream_t)stream, ids, n, desc, n_slots, seq, rc,
                       dev_size, dev_off, meta, data_cap);
    if (hipGetLastError() != hipSuccess) return -1;
    hipLaunchKernelGGL(fetch_copy_kernel, dim3(n), dim3(256), 0, (hipStream_t)stream, ids, seq, dev_size, dev_off,
                       (uint32_t *)host_data);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace h261g
