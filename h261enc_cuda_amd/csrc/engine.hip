/*
 * engine.hip -- the multi-stream encoder (C ABI part 2 of h261gpu.h).
 *
 * Every open stream owns, in HBM: its input (one host-fed frame buffer or a
 * device-generated frame sequence), two reconstruction buffers used as
 * reference / reconstruction in turn, and its DenoiseState and RcState.
 * All streams' macroblocks sit in one global MB index space (mb_base per
 * stream, a map from global MB to stream), so one frame step is three
 * launches over all streams at once:
 *
 *   me_kernel          search + SATD costs + mode decision; the MB that
 *                      completes its stream runs rate-control begin
 *                      (picture cost, rate control, picture quantiser)
 *   mb_kernel          the fused macroblock coder, bit counts, statistics
 *   frame_tail_kernel  one wave per stream: denoise offsets for the next
 *                      frame, exact frame size (GOB scan), rate-control
 *                      update, swap reference/reconstruction, next input
 *
 * (H261G_OPT_FUSED_TAILS 0: the same work as six launches: me_kernel,
 * rc_begin, mb_kernel, denoise_update, frame_bits, advance.)
 *
 * Nothing in a step depends on the host, so h261g_engine_run captures the
 * step once in a HIP graph and replays it.
 */
#include <hip/hip_runtime.h>

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <vector>

#include "../../include/h261gpu.h"
#include "content.h"
#include "frame_control.h"
#include "h261gpu_internal.h"

namespace h261g {

int generate_frames_device(const h261g_content_params_t &c, int width, int height, int first, int n, uint8_t *out,
                           content::FrameCtx *ctx, uint8_t *tab, void *stream);

namespace {

bool ok(hipError_t e, const char *what)
{
    if (e == hipSuccess) return true;
    fprintf(stderr, "h261gpu engine: %s: %s\n", what, hipGetErrorString(e));
    return false;
}

struct Slot {
    bool open = false;
    int width = 0, height = 0, mb_w = 0, mb_h = 0, n_mb = 0;
    int mb_base = -1;
    uint8_t *input = nullptr;    /* host-fed frame buffer (I420) */
    uint8_t *recon = nullptr;    /* two I420 buffers */
    uint8_t *seq = nullptr;      /* device frame sequence */
    uint8_t *bits = nullptr;     /* the last coded frame's bitstream (device writer) */
    size_t input_cap = 0, recon_cap = 0, seq_cap = 0, bits_cap = 0; /* allocation sizes (device memory cache) */
    int bits_words = 0;          /* usable capacity of bits in 32-bit words */
    int seq_frames = 0;
    int64_t footprint = 0;
    bool pending = false;        /* host frame queued for the next step */
    bool fresh = false;          /* no frame step since open: desc0/seq0 mirror the device */
    StreamDesc desc0{};
    SeqState seq0{};
    /* asynchronous input: generation on the input stream, activation by the
     * first step after gen_ev completes */
    bool gen_pending = false;
    hipEvent_t gen_ev = nullptr;
    StreamDesc gen_d{};
    SeqState gen_q{};
    h261g_stream_params_t params{};
};

/* writes one stream's device state at open (kernel arguments carry it, so
 * nothing waits for the host): descriptor, sequence state, denoise and rate
 * control state, its MB range in the map, and a zeroed reconstruction chain
 * (frame 0's reference is all zero, h261_main.c:26-28) */
__global__ void stream_init_kernel(StreamDesc *desc, SeqState *seq, DenoiseState *dn, RcState *rc,
                                   int32_t *mb_stream, int id, StreamDesc d, SeqState q, DenoiseState dnv,
                                   RcState rcv, int base, int n_mb, uint32_t *recon, int recon_words)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, n = gridDim.x * blockDim.x;
    for (int k = i; k < recon_words; k += n) recon[k] = 0;
    for (int k = i; k < n_mb; k += n) mb_stream[base + k] = id;
    if (i == 0) {
        desc[id] = d;
        seq[id] = q;
        dn[id] = dnv;
        rc[id] = rcv;
    }
}

/* descriptor and sequence state of one stream (input selection) */
__global__ void stream_set_kernel(StreamDesc *desc, SeqState *seq, int id, StreamDesc d, SeqState q)
{
    desc[id] = d;
    seq[id] = q;
}

/* re-packing of the MB index space: the new map, and each open stream's new
 * MB base (bases[id] < 0: slot not open) */
__global__ void remap_kernel(StreamDesc *desc, int32_t *mb_stream, const int32_t *bases)
{
    /* one workgroup per stream slot: an open stream's new range of the map
     * and its descriptor's base */
    const int s = blockIdx.x, base = bases[s];
    if (base < 0) return;
    const int n = desc[s].mb_w * desc[s].mb_h;
    for (int k = threadIdx.x; k < n; k += blockDim.x) mb_stream[base + k] = s;
    if (threadIdx.x == 0) desc[s].mb_base = base;
}

/* host-fed input of up to kPutIds streams, packed in one staging buffer:
 * frame i (I420, planes back to back) at stage + off[i] goes to its
 * stream's input buffer (desc.cur[0], planes back to back as well), and the
 * descriptor becomes active with the device quantiser (what
 * h261g_stream_put_frame writes with a copy).  Offsets are multiples of 16
 * and frame sizes are too (CIF and QCIF), so the copy moves 16-byte words.
 * Blocks (x, i): frame i, a grid-stride share of its words. */
constexpr int kPutIds = 256;
struct PutFrames {
    const uint8_t *stage;
    int32_t n;
    int32_t id[kPutIds];
    int64_t off[kPutIds];
};
static_assert(sizeof(PutFrames) <= 4096 - 64, "kernel argument size");
__global__ void put_frames_kernel(StreamDesc *desc, PutFrames a)
{
    const int i = blockIdx.y;
    if (i >= a.n) return;
    StreamDesc &d = desc[a.id[i]];
    const int words = d.width * d.height * 3 / 32; /* 16-byte words of an I420 frame */
    const uint4 *src = (const uint4 *)(a.stage + a.off[i]);
    uint4 *dst = (uint4 *)d.cur[0];
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < words; k += gridDim.x * blockDim.x) dst[k] = src[k];
    if (blockIdx.x == 0 && threadIdx.x == 0) d.flags = kStreamActive | kStreamDeviceQuant;
}

/* a closed stream: inactive descriptor, MB range unmapped */
__global__ void stream_close_kernel(StreamDesc *desc, int32_t *mb_stream, int id, int base, int n_mb)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n_mb) mb_stream[base + i] = -1;
    if (i == 0) desc[id] = StreamDesc{};
}

}  // namespace

}  // namespace h261g

using namespace h261g;

struct h261g_engine {
    int max_streams = 0;
    int64_t budget = 0, used = 0;
    /* device memory of stream buffers (inputs, reconstructions, frame
     * sequences, bitstreams): chunks from hipMalloc carved first-fit, ranges
     * coalesced when released.  Open, close and input generation neither
     * allocate nor free nor wait while a chunk has room.  All frame work is
     * on one stream, so a released range is reused only by work queued after
     * the work that used it (input generation on the second stream first
     * waits for the work queued so far). */
    struct DevChunk {
        uint8_t *base;
        size_t size;
        std::map<size_t, size_t> free; /* offset -> length */
    };
    std::vector<DevChunk> chunks;
    /* re-packing: pinned host staging (bases then map), its device copy, and
     * the event after which the host staging may be rewritten */
    int32_t *remap_host = nullptr, *remap_dev = nullptr;
    size_t remap_cap = 0;
    hipEvent_t remap_ev = nullptr;
    bool remap_pending = false;
    /* input generation stream (default priority, like the compute stream) and
     * the event marking the compute work a generation must follow */
    hipStream_t in_stream = nullptr;
    hipEvent_t compute_ev = nullptr;
    /* the bitstream writer of a step runs on its own stream beside the next
     * step's motion estimation (which does not touch its inputs); the next
     * step's coder, which overwrites them, waits for it.  writer_pending:
     * the last step's writer has not been launched yet */
    hipStream_t wr_stream = nullptr;
    hipEvent_t wr_fork = nullptr, wr_join = nullptr;
    int overlap_writer = 2; /* 0 off, 1 on, 2 when a launch fills the device */
    bool writer_pending = false;
    /* two chains: the second half of the MB space is stepped on chain_stream,
     * forked from and joined back into the engine stream around a run;
     * chain_me_a / chain_me_b[step & 1]: each chain's motion estimation
     * done, the rungs that keep the chains together */
    hipStream_t chain_stream = nullptr;
    hipEvent_t chain_fork = nullptr, chain_join = nullptr;
    hipEvent_t chain_me_a = nullptr, chain_me_b[2] = {nullptr, nullptr};
    /* h261g_engine_put_frames: two device staging buffers of the packed host
     * frames, used in turn; the copies run on put_stream (default priority)
     * beside the engine stream's steps.  put_copied[b]: buffer b's copy done
     * (the engine stream waits for it before placing the frames);
     * put_placed[b]: the kernel that read buffer b done (the next copy into
     * b waits for it) */
    uint8_t *put_stage[2] = {nullptr, nullptr};
    size_t put_stage_cap[2] = {0, 0};
    int put_calls = 0;
    hipStream_t put_stream = nullptr;
    hipEvent_t put_copied[2] = {nullptr, nullptr}, put_placed[2] = {nullptr, nullptr};
    int chains = 0; /* 0 by size, 1 one chain, 2 two chains */
    int ladder = 0; /* two chains: 0 by size, 1 strict, 2 one step of slack */
    int async_input = 0;
    int mbs_per_wave = 0; /* 0: chosen per launch from the MB count */
    int me_mbs_per_wave = 0; /* motion estimation: 0 chosen per launch, else 1, 2, 4 or 8 */
    int me_tile = 0;         /* MBs per staged search-window tile: 0 chosen with the run, else 1, 2 or 4 */
    int me_variant = kMeTwoStage;
    /* 1: rate-control begin in the ME tail; 2: its own launch; 3: its own
     * launch when a chain fills the device (a short step is latency bound),
     * else 1; 0: no fused tails */
    int fused_tails = 3;
    /* bit 0: motion estimation, bit 1: macroblock coder, bit 2: bitstream
     * writer.  Measured at 512 streams: the remap cuts the coder's HBM reads
     * ~9x and its time ~4-14% (at 4 MBs per wave); it cuts motion
     * estimation's HBM reads ~7.6x (509 -> 67 MB per launch) at the same
     * time within noise (that kernel is VALU/LDS bound); the writer is ~20%
     * faster without it.  Default: motion estimation and coder. */
    int xcd_remap = 3;
    /* 1: bitstreams written on the device into a per-stream buffer of
     * kBitsBytes*; 2: buffers of the worst-case frame size; 0: written on
     * the host from the levels when read */
    int device_bits = 1;
    hipStream_t stream = nullptr;
    std::vector<Slot> slots;
    /* device arrays indexed by stream slot */
    StreamDesc *desc = nullptr;
    SeqState *seq = nullptr;
    DenoiseState *dn = nullptr;
    RcState *rc = nullptr;
    /* global MB space */
    int mb_cap = 0, mb_used = 0;
    int32_t *mb_stream = nullptr;
    MeResult *me = nullptr;
    MbMode *mode = nullptr;
    MbSide *side = nullptr;
    int8_t *levels = nullptr; /* [mb][6][64]; intra DC levels (1..255) as uint8 */
    MbHdr *hdr = nullptr;     /* [mb]: bitstream places, frame tail -> writer (all invalid between steps) */
    std::vector<int> free_cif, free_qcif; /* MB ranges of closed streams */
    h261_macroblock_t *mbstruct = nullptr; /* scratch for read_macroblocks */
    /* bulk bitstream fetch: ids, sizes/offsets and bytes in pinned host
     * memory (the gather kernels write the host memory directly), device
     * scratch, and the event of the fetch in flight */
    int32_t *fetch_ids = nullptr, *fetch_meta = nullptr, *fetch_dev = nullptr;
    uint8_t *fetch_data = nullptr;
    int fetch_cap = 0;
    size_t fetch_data_cap = 0;
    int fetch_n = 0;
    bool fetch_pending = false;
    hipEvent_t fetch_ev = nullptr;
    /* device address of the allocation-free fetch's scratch on the engine's
     * device (looked up at creation) */
    int32_t *fetch_scratch = nullptr;
    int mbstruct_cap = 0;
    /* graph of one step */
    /* graphs of one step and of kGraphSteps steps: kernels inside a graph
     * follow each other without a gap, a graph launch costs a few us */
    static constexpr int kGraphSteps = 10;
    hipGraph_t graph = nullptr, graph_k = nullptr;
    hipGraphExec_t graph_exec = nullptr, graph_exec_k = nullptr;
    bool graph_valid = false;
    /* the captured step leaves its bitstream writer to the next step (one
     * overlapped chain): every replay ends with that writer pending */
    bool graph_leaves_writer = false;
    /* profiling */
    bool profile = false;
    /* profiling: a ring of event sets, one per step, read back lazily so the
     * host keeps enqueueing ahead of the device and no host gap lands
     * between an event and the kernel it brackets */
    static constexpr int kEvRing = 64;
    hipEvent_t ev[kEvRing][8] = {}; /* start/stop of ME, MB coder, frame tail, bitstream writer */
    int64_t ev_recorded = 0, ev_read = 0;
    double me_ms = 0, mb_ms = 0, other_ms = 0, bits_ms = 0;
    int64_t launches = 0;
};

namespace {

/* bitstream buffer of one stream: four times the largest coded picture
 * H.261 allows (256 kbit CIF, 64 kbit QCIF); a larger frame is written on
 * the host from its levels when read.  mode 2: the worst case of any
 * frame, so the device always writes it. */
constexpr int64_t kBitsBytesCif = 128 * 1024, kBitsBytesQcif = 32 * 1024;
int64_t bits_bytes(int w, int h, int mode)
{
    if (mode == 0) return 0;
    const int n_mb = (w / 16) * (h / 16);
    if (mode == 2) return ((frame_bits_bound(n_mb) + 31) / 32) * 4;
    return n_mb == 396 ? kBitsBytesCif : kBitsBytesQcif;
}

int64_t stream_footprint(int w, int h, int bits_mode)
{
    const int64_t fsz = (int64_t)w * h * 3 / 2, n_mb = (int64_t)(w / 16) * (h / 16);
    /* input + two reconstructions + bitstream + per-MB arrays */
    return 3 * fsz + bits_bytes(w, h, bits_mode) +
           n_mb * ((int64_t)sizeof(int32_t) + sizeof(MeResult) + sizeof(MbMode) + sizeof(MbSide) + sizeof(MbHdr) +
                   768);
}

bool valid_size(int w, int h) { return (w == 352 && h == 288) || (w == 176 && h == 144); }

/* grow the global MB arrays to hold n MBs (contents are per step, so no copy) */
int ensure_mb_capacity(h261g_engine *e, int n)
{
    if (n <= e->mb_cap) return H261G_OK;
    int cap = e->mb_cap ? e->mb_cap : 1024;
    while (cap < n) cap *= 2;
    if (!ok(hipStreamSynchronize(e->stream), "sync")) return H261G_ERR_DEVICE;
    int32_t *ms = nullptr;
    MeResult *me = nullptr;
    MbMode *mo = nullptr;
    MbSide *sd = nullptr;
    int8_t *lv = nullptr;
    MbHdr *hd = nullptr;
    if (!ok(hipMalloc((void **)&ms, sizeof(int32_t) * cap), "hipMalloc") ||
        !ok(hipMalloc((void **)&hd, sizeof(MbHdr) * cap), "hipMalloc") ||
        !ok(hipMalloc((void **)&me, sizeof(MeResult) * cap), "hipMalloc") ||
        !ok(hipMalloc((void **)&mo, sizeof(MbMode) * cap), "hipMalloc") ||
        !ok(hipMalloc((void **)&sd, sizeof(MbSide) * cap), "hipMalloc") ||
        !ok(hipMalloc((void **)&lv, (size_t)384 * cap), "hipMalloc")) {
        hipFree(ms);
        hipFree(me);
        hipFree(mo);
        hipFree(sd);
        hipFree(lv);
        hipFree(hd);
        return H261G_ERR_MEMORY;
    }
    /* no MB has a pending bitstream place between steps */
    if (!ok(hipMemset(hd, 0, sizeof(MbHdr) * cap), "hipMemset")) {
        hipFree(ms);
        hipFree(me);
        hipFree(mo);
        hipFree(sd);
        hipFree(lv);
        hipFree(hd);
        return H261G_ERR_DEVICE;
    }
    /* the map and the last step's results of the MBs in use survive growth */
    if (e->mb_stream && e->mb_used) {
        const size_t n = (size_t)e->mb_used;
        (void)hipMemcpy(ms, e->mb_stream, sizeof(int32_t) * n, hipMemcpyDeviceToDevice);
        (void)hipMemcpy(me, e->me, sizeof(MeResult) * n, hipMemcpyDeviceToDevice);
        (void)hipMemcpy(mo, e->mode, sizeof(MbMode) * n, hipMemcpyDeviceToDevice);
        (void)hipMemcpy(sd, e->side, sizeof(MbSide) * n, hipMemcpyDeviceToDevice);
        (void)hipMemcpy(lv, e->levels, (size_t)384 * n, hipMemcpyDeviceToDevice);
    }
    hipFree(e->mb_stream);
    hipFree(e->me);
    hipFree(e->mode);
    hipFree(e->side);
    hipFree(e->levels);
    hipFree(e->hdr);
    e->mb_stream = ms;
    e->hdr = hd;
    e->me = me;
    e->mode = mo;
    e->side = sd;
    e->levels = lv;
    e->mb_cap = cap;
    e->graph_valid = false;
    return H261G_OK;
}

void invalidate_graph(h261g_engine *e)
{
    e->graph_valid = false;
}

constexpr size_t kChunkBytes = (size_t)256 << 20;
constexpr size_t kRangeAlign = 256;

/* chunks with no range in use go back to the device (after the queued work) */
void release_free_chunks(h261g_engine *e)
{
    bool synced = false;
    for (size_t i = 0; i < e->chunks.size();) {
        auto &c = e->chunks[i];
        if (c.free.size() == 1 && c.free.begin()->first == 0 && c.free.begin()->second == c.size) {
            if (!synced) {
                (void)hipStreamSynchronize(e->stream);
                (void)hipStreamSynchronize(e->in_stream);
                synced = true;
            }
            hipFree(c.base);
            e->chunks.erase(e->chunks.begin() + (long)i);
        } else {
            i++;
        }
    }
}

/* a device range of at least `bytes`; *cap gets its (rounded) size */
uint8_t *cache_get(h261g_engine *e, size_t bytes, size_t *cap)
{
    const size_t b = (bytes + kRangeAlign - 1) / kRangeAlign * kRangeAlign;
    for (int attempt = 0; attempt < 2; attempt++) {
        for (auto &c : e->chunks)
            for (auto it = c.free.begin(); it != c.free.end(); ++it)
                if (it->second >= b) {
                    const size_t off = it->first, rest = it->second - b;
                    c.free.erase(it);
                    if (rest) c.free.emplace(off + b, rest);
                    *cap = b;
                    return c.base + off;
                }
        if (attempt) break;
        /* a new chunk: the standard size, or exactly this range when that
         * is larger or the device is nearly full */
        uint8_t *base = nullptr;
        size_t size = b > kChunkBytes ? b : kChunkBytes;
        if (hipMalloc((void **)&base, size) != hipSuccess) {
            (void)hipGetLastError();
            release_free_chunks(e);
            size = b;
            if (!ok(hipMalloc((void **)&base, size), "hipMalloc")) return nullptr;
        }
        h261g_engine::DevChunk c{base, size, {}};
        c.free.emplace(0, size);
        e->chunks.push_back(std::move(c));
    }
    return nullptr;
}

/* returns a range to its chunk, merged with free neighbours */
void cache_put(h261g_engine *e, uint8_t *ptr, size_t cap)
{
    if (!ptr) return;
    for (auto &c : e->chunks) {
        if (ptr < c.base || ptr >= c.base + c.size) continue;
        size_t off = (size_t)(ptr - c.base), len = cap;
        auto next = c.free.lower_bound(off);
        if (next != c.free.end() && next->first == off + len) {
            len += next->second;
            next = c.free.erase(next);
        }
        if (next != c.free.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == off) {
                off = prev->first;
                len += prev->second;
                c.free.erase(prev);
            }
        }
        c.free.emplace(off, len);
        return;
    }
}

int stream_count(const h261g_engine *e) { return e->max_streams; }

int read_oldest_times(h261g_engine *e);

/* waves of the macroblock coder that keep every SIMD of the device busy
 * (256 CUs x 4 SIMDs x 5 waves) */
constexpr int kFillWaves = 256 * 4 * 5;
/* motion-estimation waves resident on the device (256 CUs x 4 SIMDs x 7) and
 * the MBs per wave of a launch that holds several times as many MBs */
constexpr int kMeResidentWaves = 256 * 4 * 8;
constexpr int kMeRunDefault = 2;
constexpr int kMeTileDefault = 2;

/* one frame step: three launches (default), or the six standalone launches */
/* the bitstream writer over the whole MB space (MBs not placed by a frame
 * tail since the last writer are skipped) */
BitLaunch writer_launch(const h261g_engine *e)
{
    BitLaunch wl{};
    wl.mb_stream = e->mb_stream;
    wl.n_mb = e->mb_used;
    wl.seq = e->seq;
    wl.hdr = e->hdr;
    wl.side = e->side;
    wl.levels = e->levels;
    wl.xcd_remap = (e->xcd_remap >> 2) & 1;
    return wl;
}

/* the last step's bitstreams, now, on the engine stream: before anything
 * reads them or changes the MB space they are indexed by */
int flush_writer(h261g_engine *e)
{
    if (!e->writer_pending) return H261G_OK;
    e->writer_pending = false;
    return launch_bitstream(writer_launch(e), e->stream) ? H261G_ERR_DEVICE : H261G_OK;
}

/* The MB split of a two-chain step: the end of the open stream whose range
 * ends nearest the middle of the MB space (ranges are contiguous, so both
 * halves hold whole streams), or 0 for one chain. */
int chain_split(const h261g_engine *e, bool timed)
{
    const int n = e->mb_used;
    if (timed || e->chains == 1 || (e->chains == 0 && n < 2 * kFillWaves)) return 0;
    int best = 0;
    int64_t best_d = INT64_MAX;
    for (const auto &s : e->slots) {
        if (!s.open) continue;
        const int end = s.mb_base + s.n_mb;
        const int64_t d = end > n / 2 ? end - n / 2 : n / 2 - end;
        if (end < n && d < best_d) {
            best = end;
            best_d = d;
        }
    }
    return best;
}

/* MBs one motion-estimation wave searches in turn: the option, else by
 * launch size -- one while a launch is short (latency), more once the
 * launch holds several times the waves the device keeps resident */
int me_mbs_per_wave(const h261g_engine *e, int n_mb)
{
    if (e->me_mbs_per_wave) return e->me_mbs_per_wave;
    return n_mb >= 4 * kMeResidentWaves ? kMeRunDefault : 1;
}

/* one frame step of the streams whose MBs lie in [lo, hi), on stream st */
int enqueue_chain(h261g_engine *e, hipStream_t st, int lo, int hi, bool whole, bool timed, hipEvent_t *ev,
                  hipEvent_t me_done = nullptr)
{
    const int ns = stream_count(e);
    /* rate-control begin in the tail of motion estimation makes each wave
     * wait for a returning atomic (about 6% of a full launch); as its own
     * launch it costs a few us, hidden by the other chain */
    const int fused = e->fused_tails == 3 ? (hi - lo >= kFillWaves ? 2 : 1) : e->fused_tails;
    MeLaunch ml{};
    ml.streams = e->desc;
    ml.mb_stream = e->mb_stream;
    ml.n_mb = hi;
    ml.mb_first = lo;
    ml.result = e->me;
    ml.mode = e->mode;
    ml.rc = e->rc;
    ml.skip_search_on_intra = 1;
    ml.variant = e->me_variant;
    ml.fuse_tail = fused == 1;
    ml.xcd_remap = e->xcd_remap & 1;
    ml.mbs_per_wave = me_mbs_per_wave(e, hi - lo);
    ml.tile = e->me_tile ? e->me_tile : min(ml.mbs_per_wave, kMeTileDefault);
    /* one chain: the previous step's bitstreams beside this step's motion
     * estimation (with nothing pending every MB is skipped) */
    const bool overlap =
        whole && e->device_bits && (e->overlap_writer == 1 || (e->overlap_writer == 2 && hi - lo >= kFillWaves));
    if (overlap) {
        if (!ok(hipEventRecord(e->wr_fork, st), "event") || !ok(hipStreamWaitEvent(e->wr_stream, e->wr_fork, 0), "wait") ||
            launch_bitstream(writer_launch(e), e->wr_stream, timed ? ev[6] : nullptr, timed ? ev[7] : nullptr) ||
            !ok(hipEventRecord(e->wr_join, e->wr_stream), "event"))
            return H261G_ERR_DEVICE;
    }
    if (launch_me(ml, st, timed ? ev[0] : nullptr, timed ? ev[1] : nullptr)) return H261G_ERR_DEVICE;
    if (me_done && !ok(hipEventRecord(me_done, st), "event")) return H261G_ERR_DEVICE;
    if (overlap && !ok(hipStreamWaitEvent(st, e->wr_join, 0), "wait")) return H261G_ERR_DEVICE;
    if (fused != 1 && launch_rc_begin(e->rc, e->desc, ns, st, whole ? 0 : lo, whole ? INT32_MAX : hi))
        return H261G_ERR_DEVICE;
    MbLaunch bl{};
    bl.streams = e->desc;
    bl.mb_stream = e->mb_stream;
    bl.n_mb = hi;
    bl.mb_first = lo;
    bl.mode = e->mode;
    bl.denoise = e->dn;
    bl.rc = e->rc;
    bl.side = e->side;
    bl.levels = e->levels;
    /* 0 = by batch size: one MB per wave while the launch is too small to
     * fill the device, up to four when there are MBs to spare (fewer waves,
     * fewer statistic flushes) */
    const int n_mb = hi - lo;
    bl.mbs_per_wave = e->mbs_per_wave ? e->mbs_per_wave : (n_mb >= 4 * kFillWaves ? 4 : (n_mb >= 2 * kFillWaves ? 2 : 1));
    bl.xcd_remap = (e->xcd_remap >> 1) & 1;
    if (launch_mb(bl, st, timed ? ev[2] : nullptr, timed ? ev[3] : nullptr)) return H261G_ERR_DEVICE;
    /* with device bitstreams the frame tail also places every coded MB in
     * its frame's bitstream and writes the picture and GOB headers */
    MbHdr *hdr = e->device_bits ? e->hdr : nullptr;
    if (fused) {
        if (launch_frame_tail(e->desc, e->seq, e->dn, e->rc, e->side, hdr, ns, st, timed ? ev[4] : nullptr,
                              timed ? ev[5] : nullptr, whole ? 0 : lo, whole ? INT32_MAX : hi))
            return H261G_ERR_DEVICE;
    } else {
        /* separate kernels (one chain only): "other" spans the three
         * end-of-frame launches (rate-control begin, between the two big
         * kernels, is not timed) */
        if (timed) (void)hipEventRecord(ev[4], st);
        if (launch_denoise_update(e->dn, e->desc, ns, st)) return H261G_ERR_DEVICE;
        if (launch_frame_bits(e->rc, e->desc, ns, e->side, nullptr, st, e->seq, hdr)) return H261G_ERR_DEVICE;
        if (launch_advance(e->desc, e->seq, ns, st)) return H261G_ERR_DEVICE;
        if (timed) (void)hipEventRecord(ev[5], st);
    }
    if (hdr && !overlap) {
        BitLaunch wl = writer_launch(e);
        wl.mb_first = lo;
        wl.n_mb = hi;
        if (launch_bitstream(wl, st, timed ? ev[6] : nullptr, timed ? ev[7] : nullptr)) return H261G_ERR_DEVICE;
    } else if (!hdr && timed) {
        (void)hipEventRecord(ev[6], st);
        (void)hipEventRecord(ev[7], st);
    }
    if (whole) e->writer_pending = overlap;
    return H261G_OK;
}

/* k frame steps on the engine stream: one chain, or two chains forked onto
 * chain_stream and joined back after the k steps (the chains share no data,
 * so they need not meet between steps) */
int enqueue_steps(h261g_engine *e, int k, bool timed)
{
    const int split = e->fused_tails ? chain_split(e, timed) : 0;
    hipStream_t st = e->stream;
    if (!split) {
        for (int i = 0; i < k; i++) {
            if (timed && e->ev_recorded - e->ev_read == h261g_engine::kEvRing) {
                int rcode = read_oldest_times(e);
                if (rcode) return rcode;
            }
            hipEvent_t *ev = e->ev[e->ev_recorded % h261g_engine::kEvRing];
            int rcode = enqueue_chain(e, st, 0, e->mb_used, true, timed, ev);
            if (rcode) return rcode;
            if (timed) e->ev_recorded++;
        }
        return H261G_OK;
    }
    /* A ladder of motion-estimation events ties the chains: chain B's
     * search of step i starts after chain A's.  Without the rungs nothing
     * ties them, and a captured run's launches reach the device one chain
     * after the other: a chain would run most of its steps alone.
     *   strict: chain A's search of step i+1 starts after chain B's of step
     *   i.  The searches alternate, so each half's coder, frame tail and
     *   writer run beside the other half's search, and the latency-bound
     *   tails never meet on an otherwise idle device.
     *   one step of slack: chain A's search of step i+1 starts after chain
     *   B's of step i-1, so A's next search may overlap B's current one.
     *   Once the halves run at the same pace, B trails A by a whole step and
     *   both halves' tails coincide, but when finished streams leave the
     *   halves small and uneven, a half's search alone does not fill the
     *   device and the strict form loses (6% over the benchmark's job at 4
     *   MBs per search wave).
     * By size: strict while each half holds four times the search waves the
     * device keeps resident. */
    const int half = split < e->mb_used - split ? split : e->mb_used - split;
    const bool strict = e->ladder == 1 || (e->ladder == 0 && half >= 4 * kMeResidentWaves);
    if (!ok(hipEventRecord(e->chain_fork, st), "event") || !ok(hipStreamWaitEvent(e->chain_stream, e->chain_fork, 0), "wait"))
        return H261G_ERR_DEVICE;
    for (int i = 0; i < k; i++) {
        const int rung = strict ? i - 1 : i - 2; /* B's search that A's of step i waits for */
        if (rung >= 0 && !ok(hipStreamWaitEvent(st, e->chain_me_b[rung & 1], 0), "wait")) return H261G_ERR_DEVICE;
        int rcode = enqueue_chain(e, st, 0, split, false, false, nullptr, e->chain_me_a);
        if (!rcode && !ok(hipStreamWaitEvent(e->chain_stream, e->chain_me_a, 0), "wait")) rcode = H261G_ERR_DEVICE;
        if (!rcode)
            rcode = enqueue_chain(e, e->chain_stream, split, e->mb_used, false, false, nullptr, e->chain_me_b[i & 1]);
        if (rcode) return rcode;
    }
    if (!ok(hipEventRecord(e->chain_join, e->chain_stream), "event") || !ok(hipStreamWaitEvent(st, e->chain_join, 0), "wait"))
        return H261G_ERR_DEVICE;
    return H261G_OK;
}

/* reads back the oldest recorded event set */
int read_oldest_times(h261g_engine *e)
{
    hipEvent_t *ev = e->ev[e->ev_read % h261g_engine::kEvRing];
    if (!ok(hipEventSynchronize(ev[7]), "event")) return H261G_ERR_DEVICE;
    float a = 0, c = 0, f = 0, b = 0;
    (void)hipEventElapsedTime(&a, ev[0], ev[1]);
    (void)hipEventElapsedTime(&c, ev[2], ev[3]);
    (void)hipEventElapsedTime(&f, ev[4], ev[5]);
    (void)hipEventElapsedTime(&b, ev[6], ev[7]);
    e->me_ms += a;
    e->other_ms += f;
    e->mb_ms += c;
    e->bits_ms += b;
    e->launches++;
    e->ev_read++;
    return H261G_OK;
}

int read_all_times(h261g_engine *e)
{
    while (e->ev_read < e->ev_recorded) {
        int rcode = read_oldest_times(e);
        if (rcode) return rcode;
    }
    return H261G_OK;
}

Slot *get_slot(h261g_engine *e, int id)
{
    if (!e || id < 0 || id >= e->max_streams || !e->slots[id].open) return nullptr;
    return &e->slots[id];
}

}  // namespace

extern "C" {

int h261g_engine_config(h261g_engine_t *e, int32_t option, int32_t value)
{
    if (!e) return H261G_ERR_ARG;
    if (flush_writer(e)) return H261G_ERR_DEVICE;
    switch (option) {
    case H261G_OPT_ME_VARIANT:
        if (value != kMeTwoStage && value != kMeSurface) return H261G_ERR_ARG;
        e->me_variant = value;
        break;
    case H261G_OPT_MBS_PER_WAVE:
        if (value < 0 || value > 64) return H261G_ERR_ARG;
        e->mbs_per_wave = value;
        break;
    case H261G_OPT_FUSED_TAILS:
        if (value < 0 || value > 3) return H261G_ERR_ARG;
        e->fused_tails = value;
        break;
    case H261G_OPT_XCD_REMAP:
        if (value < 0 || value > 7) return H261G_ERR_ARG;
        e->xcd_remap = value;
        break;
    case H261G_OPT_ASYNC_INPUT:
        if (value != 0 && value != 1) return H261G_ERR_ARG;
        e->async_input = value;
        return H261G_OK;
    case H261G_OPT_OVERLAP_WRITER:
        if (value < 0 || value > 2) return H261G_ERR_ARG;
        e->overlap_writer = value;
        break;
    case H261G_OPT_CHAINS:
        if (value < 0 || value > 2) return H261G_ERR_ARG;
        e->chains = value;
        break;
    case H261G_OPT_CHAIN_LADDER:
        if (value < 0 || value > 2) return H261G_ERR_ARG;
        e->ladder = value;
        break;
    case H261G_OPT_ME_MBS_PER_WAVE:
        if (value != 0 && value != 1 && value != 2 && value != 4 && value != 8) return H261G_ERR_ARG;
        e->me_mbs_per_wave = value;
        break;
    case H261G_OPT_ME_TILE:
        if (value != 0 && value != 1 && value != 2 && value != 4) return H261G_ERR_ARG;
        e->me_tile = value;
        break;
    case H261G_OPT_DEVICE_BITSTREAM:
        if (value < 0 || value > 2) return H261G_ERR_ARG;
        /* bitstream buffers are sized at open */
        for (const auto &sl : e->slots)
            if (sl.open) return H261G_ERR_STATE;
        e->device_bits = value;
        break;
    default:
        return H261G_ERR_ARG;
    }
    e->graph_valid = false;
    return H261G_OK;
}

int h261g_device_info(h261g_device_info_t *out)
{
    if (!out) return H261G_ERR_ARG;
    int dev = 0;
    hipDeviceProp_t prop;
    if (!ok(hipGetDevice(&dev), "hipGetDevice") || !ok(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties"))
        return H261G_ERR_DEVICE;
    *out = h261g_device_info_t{};
    out->compute_units = prop.multiProcessorCount;
    out->clock_khz = prop.clockRate;
    out->l2_bytes = prop.l2CacheSize;
    out->memory_bytes = (int64_t)prop.totalGlobalMem;
    snprintf(out->name, sizeof(out->name), "%s", prop.name);
    snprintf(out->arch, sizeof(out->arch), "%s", prop.gcnArchName);
    return H261G_OK;
}

int64_t h261g_stream_footprint(int32_t width, int32_t height)
{
    if (!valid_size(width, height)) return H261G_ERR_ARG;
    return stream_footprint(width, height, 1);
}

int h261g_engine_create(h261g_engine_t **out, int32_t max_streams, int64_t hbm_budget, int32_t mbs_per_wave)
{
    if (!out || max_streams <= 0 || max_streams > kMaxStreams || hbm_budget < 0 || mbs_per_wave < 0)
        return H261G_ERR_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return H261G_ERR_DEVICE;
    h261g_engine *e = new h261g_engine();
    e->max_streams = max_streams;
    e->budget = hbm_budget;
    e->mbs_per_wave = mbs_per_wave;
    e->slots.resize(max_streams);
    bool good = ok(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking), "stream") &&
                ok(hipStreamCreateWithFlags(&e->in_stream, hipStreamNonBlocking), "stream") &&
                ok(hipStreamCreateWithFlags(&e->wr_stream, hipStreamNonBlocking), "stream") &&
                ok(hipEventCreateWithFlags(&e->wr_fork, hipEventDisableTiming), "event") &&
                ok(hipEventCreateWithFlags(&e->wr_join, hipEventDisableTiming), "event") &&
                ok(hipStreamCreateWithFlags(&e->chain_stream, hipStreamNonBlocking), "stream") &&
                ok(hipEventCreateWithFlags(&e->chain_fork, hipEventDisableTiming), "event") &&
                ok(hipEventCreateWithFlags(&e->chain_join, hipEventDisableTiming), "event") &&
                ok(hipEventCreateWithFlags(&e->chain_me_a, hipEventDisableTiming), "event") &&
                ok(hipEventCreateWithFlags(&e->chain_me_b[0], hipEventDisableTiming), "event") &&
                ok(hipEventCreateWithFlags(&e->chain_me_b[1], hipEventDisableTiming), "event") &&
                ok(hipEventCreateWithFlags(&e->compute_ev, hipEventDisableTiming), "event") &&
                ok(hipMalloc((void **)&e->desc, sizeof(StreamDesc) * max_streams), "hipMalloc") &&
                ok(hipMalloc((void **)&e->seq, sizeof(SeqState) * max_streams), "hipMalloc") &&
                ok(hipMalloc((void **)&e->dn, sizeof(DenoiseState) * max_streams), "hipMalloc") &&
                ok(hipMalloc((void **)&e->rc, sizeof(RcState) * max_streams), "hipMalloc") &&
                ok(hipMemset(e->desc, 0, sizeof(StreamDesc) * max_streams), "hipMemset") &&
                ok(hipMemset(e->seq, 0, sizeof(SeqState) * max_streams), "hipMemset") &&
                ok(hipHostMalloc((void **)&e->remap_host, sizeof(int32_t) * max_streams, hipHostMallocDefault),
                   "hipHostMalloc") &&
                ok(hipMalloc((void **)&e->remap_dev, sizeof(int32_t) * max_streams), "hipMalloc") &&
                ok(hipEventCreateWithFlags(&e->remap_ev, hipEventDisableTiming), "event") &&
                (e->fetch_scratch = fetch_into_scratch()) != nullptr;
    e->remap_cap = good ? (size_t)max_streams : 0;
    for (int r = 0; good && r < h261g_engine::kEvRing; r++)
        for (int i = 0; good && i < 8; i++) good = ok(hipEventCreate(&e->ev[r][i]), "event");
    if (!good) {
        h261g_engine_destroy(e);
        return H261G_ERR_DEVICE;
    }
    *out = e;
    return H261G_OK;
}

int h261g_engine_destroy(h261g_engine_t *e)
{
    if (!e) return H261G_ERR_ARG;
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    if (e->in_stream) (void)hipStreamSynchronize(e->in_stream);
    if (e->wr_stream) (void)hipStreamSynchronize(e->wr_stream);
    if (e->chain_stream) (void)hipStreamSynchronize(e->chain_stream);
    for (auto &s : e->slots)
        if (s.gen_ev) (void)hipEventDestroy(s.gen_ev);
    for (auto &c : e->chunks) hipFree(c.base);
    e->chunks.clear();
    hipHostFree(e->remap_host);
    hipFree(e->remap_dev);
    if (e->remap_ev) (void)hipEventDestroy(e->remap_ev);
    if (e->graph_exec) (void)hipGraphExecDestroy(e->graph_exec);
    if (e->graph) (void)hipGraphDestroy(e->graph);
    if (e->graph_exec_k) (void)hipGraphExecDestroy(e->graph_exec_k);
    if (e->graph_k) (void)hipGraphDestroy(e->graph_k);
    hipFree(e->desc);
    hipFree(e->seq);
    hipFree(e->dn);
    hipFree(e->rc);
    hipFree(e->mb_stream);
    hipFree(e->me);
    hipFree(e->mode);
    hipFree(e->side);
    hipFree(e->levels);
    hipFree(e->hdr);
    hipHostFree(e->mbstruct);
    hipHostFree(e->fetch_ids);
    hipHostFree(e->fetch_meta);
    hipHostFree(e->fetch_data);
    hipFree(e->fetch_dev);
    if (e->put_stream) (void)hipStreamSynchronize(e->put_stream);
    for (int b = 0; b < 2; b++) {
        hipFree(e->put_stage[b]);
        if (e->put_copied[b]) (void)hipEventDestroy(e->put_copied[b]);
        if (e->put_placed[b]) (void)hipEventDestroy(e->put_placed[b]);
    }
    if (e->put_stream) (void)hipStreamDestroy(e->put_stream);
    if (e->fetch_ev) (void)hipEventDestroy(e->fetch_ev);
    for (int r = 0; r < h261g_engine::kEvRing; r++)
        for (int i = 0; i < 8; i++)
            if (e->ev[r][i]) (void)hipEventDestroy(e->ev[r][i]);
    if (e->compute_ev) (void)hipEventDestroy(e->compute_ev);
    if (e->in_stream) (void)hipStreamDestroy(e->in_stream);
    if (e->wr_stream) (void)hipStreamDestroy(e->wr_stream);
    if (e->wr_fork) (void)hipEventDestroy(e->wr_fork);
    if (e->wr_join) (void)hipEventDestroy(e->wr_join);
    if (e->chain_stream) (void)hipStreamDestroy(e->chain_stream);
    if (e->chain_fork) (void)hipEventDestroy(e->chain_fork);
    if (e->chain_join) (void)hipEventDestroy(e->chain_join);
    if (e->chain_me_a) (void)hipEventDestroy(e->chain_me_a);
    for (hipEvent_t ev : e->chain_me_b)
        if (ev) (void)hipEventDestroy(ev);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
    return H261G_OK;
}

int h261g_stream_open(h261g_engine_t *e, const h261g_stream_params_t *p)
{
    if (!e || !p || !valid_size(p->width, p->height) || p->denoise < 0 || p->init_quant < 1 || p->init_quant > 31 ||
        p->bitrate < 8 || p->vbv_size < 8 || p->fixed_quant < 0 || p->fixed_quant > 31)
        return H261G_ERR_ARG;
    int id = -1;
    for (int i = 0; i < e->max_streams; i++)
        if (!e->slots[i].open) {
            id = i;
            break;
        }
    if (id < 0) return H261G_ERR_FULL;
    if (flush_writer(e)) return H261G_ERR_DEVICE;
    const int64_t fp = stream_footprint(p->width, p->height, e->device_bits);
    if (e->budget && e->used + fp > e->budget) return H261G_ERR_FULL;
    Slot &s = e->slots[id];
    const int64_t fsz = (int64_t)p->width * p->height * 3 / 2;
    hipEvent_t gen_ev = s.gen_ev; /* a slot keeps its event object across streams */
    s = Slot();
    s.gen_ev = gen_ev;
    s.width = p->width;
    s.height = p->height;
    s.mb_w = p->width / 16;
    s.mb_h = p->height / 16;
    s.n_mb = s.mb_w * s.mb_h;
    s.params = *p;
    s.input = cache_get(e, fsz, &s.input_cap);
    s.recon = s.input ? cache_get(e, 2 * fsz, &s.recon_cap) : nullptr;
    const int64_t bbytes = bits_bytes(p->width, p->height, e->device_bits);
    s.bits = (s.recon && bbytes) ? cache_get(e, (size_t)bbytes, &s.bits_cap) : nullptr;
    if (!s.input || !s.recon || (bbytes && !s.bits)) {
        cache_put(e, s.input, s.input_cap);
        cache_put(e, s.recon, s.recon_cap);
        s = Slot();
        s.gen_ev = gen_ev;
        return H261G_ERR_MEMORY;
    }
    /* MB range: a free range of a closed stream of the same size, else append */
    int base = -1;
    auto &fl = s.n_mb == 396 ? e->free_cif : e->free_qcif;
    if (!fl.empty()) {
        base = fl.back();
        fl.pop_back();
    } else {
        base = e->mb_used;
        int rcode = ensure_mb_capacity(e, e->mb_used + s.n_mb);
        if (rcode) {
            cache_put(e, s.input, s.input_cap);
            cache_put(e, s.recon, s.recon_cap);
            cache_put(e, s.bits, s.bits_cap);
            s = Slot();
            s.gen_ev = gen_ev;
            return rcode;
        }
        e->mb_used += s.n_mb;
    }
    s.mb_base = base;
    s.open = true;
    s.footprint = fp;
    e->used += fp;
    /* device state: descriptor, reconstruction chain (frame 0 reference is
     * all zero, h261_main.c:26-28), denoise, rate control */
    std::vector<int32_t> map(s.n_mb, id);
    StreamDesc d{};
    const size_t ysz = (size_t)s.width * s.height, csz = ysz / 4;
    d.cur[0] = s.input;
    d.cur[1] = s.input + ysz;
    d.cur[2] = s.input + ysz + csz;
    d.ref[0] = s.recon;
    d.ref[1] = s.recon + ysz;
    d.ref[2] = s.recon + ysz + csz;
    d.rec[0] = s.recon + fsz;
    d.rec[1] = s.recon + fsz + ysz;
    d.rec[2] = s.recon + fsz + ysz + csz;
    d.width = s.width;
    d.height = s.height;
    d.mb_w = s.mb_w;
    d.mb_h = s.mb_h;
    d.mb_base = base;
    d.frame_type = H261_FRAME_TYPE_INTRA;
    d.pic_quant = 1;
    d.flags = kStreamDeviceQuant; /* inactive until it has input */
    d.frame_num = 0;
    d.source_format = s.width == 352 ? H261_SOURCE_FORMAT_CIF : H261_SOURCE_FORMAT_QCIF;
    SeqState q{};
    q.frame_size = fsz;
    q.pp = 0;
    q.recon[0] = s.recon;
    q.recon[1] = s.recon + fsz;
    s.bits_words = (int)(bbytes / 4);
    q.bits = (uint32_t *)s.bits;
    q.bits_words = s.bits_words;
    DenoiseState dn{};
    dn.denoise = p->denoise;
    RcState rc{};
    rc.mode = H261_RATECTRL_MODE_CBR;
    rc.current_quant = p->init_quant; /* h261_init_ratectrl (ratectrl.c:107-127) */
    rc.prediction_cost[0] = rc.prediction_cost[1] = 100;
    rc.prediction_size[0] = rc.prediction_size[1] = 1;
    rc.bitrate = p->bitrate / 8;
    rc.vbv_size = p->vbv_size / 8;
    rc.vbv_fill = rc.vbv_size;
    rc.fixed_quant = p->fixed_quant;
    hipLaunchKernelGGL(stream_init_kernel, dim3(64), dim3(256), 0, e->stream, e->desc, e->seq, e->dn, e->rc,
                       e->mb_stream, id, d, q, dn, rc, base, s.n_mb, (uint32_t *)s.recon, (int)(2 * fsz / 4));
    if (!ok(hipGetLastError(), "stream init")) return H261G_ERR_DEVICE;
    s.desc0 = d;
    s.seq0 = q;
    s.fresh = true;
    invalidate_graph(e);
    return id;
}

int h261g_stream_close(h261g_engine_t *e, int32_t id)
{
    Slot *s = get_slot(e, id);
    if (!s) return H261G_ERR_ARG;
    if (flush_writer(e)) return H261G_ERR_DEVICE;
    /* an input still being generated finishes before its buffer is released */
    if (s->gen_pending && !ok(hipStreamWaitEvent(e->stream, s->gen_ev, 0), "wait")) return H261G_ERR_DEVICE;
    hipEvent_t gen_ev = s->gen_ev;
    /* ordered after every step that used the stream; the buffers go to the
     * cache (a later open reuses them after this point on the same stream) */
    hipLaunchKernelGGL(stream_close_kernel, dim3((s->n_mb + 255) / 256), dim3(256), 0, e->stream, e->desc, e->mb_stream,
                       id, s->mb_base, s->n_mb);
    if (!ok(hipGetLastError(), "stream close")) return H261G_ERR_DEVICE;
    cache_put(e, s->input, s->input_cap);
    cache_put(e, s->recon, s->recon_cap);
    cache_put(e, s->seq, s->seq_cap);
    cache_put(e, s->bits, s->bits_cap);
    e->used -= s->footprint;
    (s->n_mb == 396 ? e->free_cif : e->free_qcif).push_back(s->mb_base);
    *s = Slot();
    s->gen_ev = gen_ev; /* the slot keeps its event for the next stream */
    invalidate_graph(e);
    return H261G_OK;
}

int h261g_stream_put_frame(h261g_engine_t *e, int32_t id, const uint8_t *y, const uint8_t *cb, const uint8_t *cr)
{
    Slot *s = get_slot(e, id);
    if (!s || !y || !cb || !cr || s->seq) return H261G_ERR_ARG;
    if (flush_writer(e)) return H261G_ERR_DEVICE;
    const size_t ysz = (size_t)s->width * s->height, csz = ysz / 4;
    int32_t flags = kStreamActive | kStreamDeviceQuant;
    bool good = ok(hipMemcpyAsync(s->input, y, ysz, hipMemcpyHostToDevice, e->stream), "hipMemcpy") &&
                ok(hipMemcpyAsync(s->input + ysz, cb, csz, hipMemcpyHostToDevice, e->stream), "hipMemcpy") &&
                ok(hipMemcpyAsync(s->input + ysz + csz, cr, csz, hipMemcpyHostToDevice, e->stream), "hipMemcpy") &&
                ok(hipMemcpyAsync(&e->desc[id].flags, &flags, sizeof(flags), hipMemcpyHostToDevice, e->stream),
                   "hipMemcpy") &&
                ok(hipStreamSynchronize(e->stream), "sync");
    s->pending = true;
    return good ? H261G_OK : H261G_ERR_DEVICE;
}

int h261g_engine_put_frames(h261g_engine_t *e, int32_t n, const int32_t *ids, const uint8_t *frames,
                            int64_t frames_bytes, const int64_t *offsets)
{
    if (!e || n < 0 || (n > 0 && (!ids || !frames || !offsets))) return H261G_ERR_ARG;
    /* everything is checked before anything is queued: open host-fed
     * streams, each at most once, 16-byte aligned offsets, frames inside
     * the buffer */
    std::vector<uint8_t> seen(e->slots.size(), 0);
    int64_t lo = INT64_MAX, hi = 0;
    for (int i = 0; i < n; i++) {
        Slot *s = get_slot(e, ids[i]);
        if (!s || s->seq || seen[ids[i]] || offsets[i] < 0 || (offsets[i] & 15)) return H261G_ERR_ARG;
        seen[ids[i]] = 1;
        lo = offsets[i] < lo ? offsets[i] : lo;
        const int64_t end = offsets[i] + (int64_t)s->width * s->height * 3 / 2;
        if (end > frames_bytes) return H261G_ERR_ARG;
        hi = end > hi ? end : hi;
    }
    if (n == 0) return H261G_OK;
    if (flush_writer(e)) return H261G_ERR_DEVICE;
    if (!e->put_stream) {
        /* the copy stream and its events are kept only when all exist */
        hipStream_t st = nullptr;
        hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
        bool good = ok(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "stream");
        for (int k = 0; good && k < 4; k++) good = ok(hipEventCreateWithFlags(&ev[k], hipEventDisableTiming), "event");
        if (!good) {
            for (hipEvent_t x : ev)
                if (x) (void)hipEventDestroy(x);
            if (st) (void)hipStreamDestroy(st);
            return H261G_ERR_DEVICE;
        }
        e->put_stream = st;
        e->put_copied[0] = ev[0];
        e->put_copied[1] = ev[1];
        e->put_placed[0] = ev[2];
        e->put_placed[1] = ev[3];
    }
    const int b = e->put_calls & 1;
    const size_t span = (size_t)(hi - lo);
    if (span > e->put_stage_cap[b]) {
        /* the previous copy into buffer b and the kernel that read it may
         * still run */
        if (!ok(hipStreamSynchronize(e->put_stream), "sync") || !ok(hipStreamSynchronize(e->stream), "sync"))
            return H261G_ERR_DEVICE;
        if (e->put_stage[b]) (void)hipFree(e->put_stage[b]);
        e->put_stage[b] = nullptr;
        e->put_stage_cap[b] = 0;
        const size_t cap = span + span / 4;
        if (!ok(hipMalloc((void **)&e->put_stage[b], cap), "hipMalloc")) return H261G_ERR_MEMORY;
        e->put_stage_cap[b] = cap;
    }
    /* one copy of the whole span (a per-plane copy costs microseconds of
     * runtime overhead each, far more than its bytes take on PCIe), on the
     * copy stream once the frames last staged in buffer b are placed, so it
     * runs beside the steps queued so far */
    if (e->put_calls >= 2 && !ok(hipStreamWaitEvent(e->put_stream, e->put_placed[b], 0), "wait"))
        return H261G_ERR_DEVICE;
    if (!ok(hipMemcpyAsync(e->put_stage[b], frames + lo, span, hipMemcpyHostToDevice, e->put_stream), "hipMemcpy") ||
        !ok(hipEventRecord(e->put_copied[b], e->put_stream), "event") ||
        !ok(hipStreamWaitEvent(e->stream, e->put_copied[b], 0), "wait"))
        return H261G_ERR_DEVICE;
    for (int i0 = 0; i0 < n; i0 += kPutIds) {
        PutFrames a{};
        a.stage = e->put_stage[b];
        a.n = min(kPutIds, n - i0);
        for (int k = 0; k < a.n; k++) {
            a.id[k] = ids[i0 + k];
            a.off[k] = offsets[i0 + k] - lo;
        }
        hipLaunchKernelGGL(put_frames_kernel, dim3(8, a.n), dim3(256), 0, e->stream, e->desc, a);
        if (!ok(hipGetLastError(), "put_frames")) return H261G_ERR_DEVICE;
    }
    if (!ok(hipEventRecord(e->put_placed[b], e->stream), "event")) return H261G_ERR_DEVICE;
    e->put_calls++;
    for (int i = 0; i < n; i++) e->slots[ids[i]].pending = true;
    return H261G_OK;
}

int h261g_stream_input_pending(h261g_engine_t *e, int32_t id)
{
    Slot *s = get_slot(e, id);
    if (!s) return H261G_ERR_ARG;
    return s->gen_pending ? 1 : 0;
}

int h261g_stream_frames_done(h261g_engine_t *e, int32_t id)
{
    Slot *s = get_slot(e, id);
    if (!s) return H261G_ERR_ARG;
    if (flush_writer(e)) return H261G_ERR_DEVICE;
    StreamDesc d;
    if (!ok(hipStreamSynchronize(e->stream), "sync") ||
        !ok(hipMemcpy(&d, e->desc + id, sizeof(d), hipMemcpyDeviceToHost), "hipMemcpy"))
        return H261G_ERR_DEVICE;
    return d.frame_num;
}

/* per-MB arrays only carry data within a step, so re-packing needs only the
 * map and each descriptor's mb_base rewritten.  The host knows both, so the
 * update is queued (pinned staging copy + a kernel) and nothing waits */
int h261g_engine_compact(h261g_engine_t *e)
{
    if (!e) return H261G_ERR_ARG;
    if (flush_writer(e)) return H261G_ERR_DEVICE;
    /* only the new bases travel (a few KB): the remap kernel rebuilds the
     * map from them; entries past the packed range are never read (launches
     * cover mb_used MBs) */
    const size_t words = (size_t)e->max_streams;
    if (e->remap_pending && !ok(hipEventSynchronize(e->remap_ev), "event")) return H261G_ERR_DEVICE;
    e->remap_pending = false;
    if (words > e->remap_cap) {
        hipHostFree(e->remap_host);
        hipFree(e->remap_dev);
        e->remap_host = nullptr;
        e->remap_dev = nullptr;
        e->remap_cap = 0;
        if (!ok(hipHostMalloc((void **)&e->remap_host, words * sizeof(int32_t), hipHostMallocDefault), "hipHostMalloc") ||
            !ok(hipMalloc((void **)&e->remap_dev, words * sizeof(int32_t)), "hipMalloc"))
            return H261G_ERR_MEMORY;
        e->remap_cap = words;
    }
    if (!e->remap_ev && !ok(hipEventCreateWithFlags(&e->remap_ev, hipEventDisableTiming), "event"))
        return H261G_ERR_DEVICE;
    std::vector<int32_t> bases(words, -1);
    int used = 0;
    for (int i = 0; i < e->max_streams; i++) {
        Slot &s = e->slots[i];
        if (!s.open) continue;
        s.mb_base = used;
        s.desc0.mb_base = used;
        /* an input still generating writes gen_d whole when it activates:
         * it must carry the new base, which remap_kernel gives the device */
        s.gen_d.mb_base = used;
        bases[i] = used;
        used += s.n_mb;
    }
    memcpy(e->remap_host, bases.data(), words * sizeof(int32_t));
    e->mb_used = used;
    e->free_cif.clear();
    e->free_qcif.clear();
    if (!ok(hipMemcpyAsync(e->remap_dev, e->remap_host, words * sizeof(int32_t), hipMemcpyHostToDevice, e->stream),
            "hipMemcpy") ||
        !ok(hipEventRecord(e->remap_ev, e->stream), "event"))
        return H261G_ERR_DEVICE;
    e->remap_pending = true;
    hipLaunchKernelGGL(remap_kernel, dim3(e->max_streams), dim3(256), 0, e->stream, e->desc, e->mb_stream, e->remap_dev);
    if (!ok(hipGetLastError(), "remap")) return H261G_ERR_DEVICE;
    invalidate_graph(e);
    return H261G_OK;
}

int h261g_stream_generate_input(h261g_engine_t *e, int32_t id, const h261g_content_params_t *c, int32_t n_frames)
{
    Slot *s = get_slot(e, id);
    if (!s || !c || n_frames <= 0 || s->seq) return H261G_ERR_ARG;
    if (flush_writer(e)) return H261G_ERR_DEVICE;
    const int64_t fsz = (int64_t)s->width * s->height * 3 / 2;
    const int64_t bytes = fsz * n_frames;
    if (e->budget && e->used + bytes > e->budget) return H261G_ERR_FULL;
    /* the frames, then the generator's per-frame parameters and its texture
     * table (one per stream, so generations on different HIP streams never
     * share one) */
    const size_t ctx_off = ((size_t)bytes + 255) & ~(size_t)255;
    const size_t tab_off = (ctx_off + sizeof(content::FrameCtx) * n_frames + 255) & ~(size_t)255;
    s->seq = cache_get(e, tab_off + content::kTextureTableBytes, &s->seq_cap);
    if (!s->seq) return H261G_ERR_MEMORY;
    /* asynchronous input (a fresh stream only): generated on the input
     * stream after the compute work queued so far (a cached buffer may still
     * be read by it); the stream joins the first step after it is ready */
    const bool async = e->async_input && s->fresh;
    if (async && !s->gen_ev && !ok(hipEventCreateWithFlags(&s->gen_ev, hipEventDisableTiming), "event"))
        return H261G_ERR_DEVICE;
    if (async && (!ok(hipEventRecord(e->compute_ev, e->stream), "event") ||
                  !ok(hipStreamWaitEvent(e->in_stream, e->compute_ev, 0), "wait")))
        return H261G_ERR_DEVICE;
    int rcode = generate_frames_device(*c, s->width, s->height, 0, n_frames, s->seq,
                                       (content::FrameCtx *)(s->seq + ctx_off), s->seq + tab_off,
                                       async ? e->in_stream : e->stream);
    if (rcode) {
        cache_put(e, s->seq, s->seq_cap);
        s->seq = nullptr;
        return rcode;
    }
    e->used += bytes;
    s->footprint += bytes;
    s->seq_frames = n_frames;
    /* point the descriptor at the next frame to code and activate: from the
     * host mirror while no step has run since open, else read back */
    StreamDesc d;
    SeqState q;
    if (s->fresh) {
        d = s->desc0;
        q = s->seq0;
    } else if (!ok(hipStreamSynchronize(e->stream), "sync") ||
               !ok(hipMemcpy(&d, e->desc + id, sizeof(d), hipMemcpyDeviceToHost), "hipMemcpy") ||
               !ok(hipMemcpy(&q, e->seq + id, sizeof(q), hipMemcpyDeviceToHost), "hipMemcpy")) {
        return H261G_ERR_DEVICE;
    }
    const size_t ysz = (size_t)s->width * s->height, csz = ysz / 4;
    const uint8_t *cur = s->seq + (size_t)fsz * d.frame_num;
    d.cur[0] = cur;
    d.cur[1] = cur + ysz;
    d.cur[2] = cur + ysz + csz;
    d.flags = kStreamActive | kStreamDeviceQuant | kStreamSeqInput;
    q.base = s->seq;
    q.remaining = n_frames - d.frame_num;
    if (async) {
        if (!ok(hipEventRecord(s->gen_ev, e->in_stream), "event")) return H261G_ERR_DEVICE;
        s->gen_pending = true;
        s->gen_d = d;
        s->gen_q = q;
        s->desc0 = d;
        s->seq0 = q;
        return H261G_OK;
    }
    hipLaunchKernelGGL(stream_set_kernel, dim3(1), dim3(1), 0, e->stream, e->desc, e->seq, id, d, q);
    if (!ok(hipGetLastError(), "stream set")) return H261G_ERR_DEVICE;
    if (s->fresh) {
        s->desc0 = d;
        s->seq0 = q;
    }
    return H261G_OK;
}

/* streams whose asynchronous input is ready join the next step */
int activate_ready_inputs(h261g_engine *e)
{
    for (int id = 0; id < e->max_streams; id++) {
        Slot &s = e->slots[id];
        if (!s.open || !s.gen_pending) continue;
        const hipError_t q = hipEventQuery(s.gen_ev);
        if (q == hipErrorNotReady) continue;
        if (!ok(q, "event")) return H261G_ERR_DEVICE;
        hipLaunchKernelGGL(stream_set_kernel, dim3(1), dim3(1), 0, e->stream, e->desc, e->seq, id, s.gen_d, s.gen_q);
        if (!ok(hipGetLastError(), "stream set")) return H261G_ERR_DEVICE;
        s.gen_pending = false;
    }
    return H261G_OK;
}

int h261g_engine_step(h261g_engine_t *e)
{
    if (!e) return H261G_ERR_ARG;
    if (e->async_input) {
        int rc0 = activate_ready_inputs(e);
        if (rc0) return rc0;
    }
    int n = 0;
    for (auto &s : e->slots)
        if (s.open && (s.pending || s.seq)) n++;
    /* a two-chain step runs its writers in-chain: a writer left pending by
     * one-chain steps goes first */
    if (chain_split(e, e->profile) && flush_writer(e)) return H261G_ERR_DEVICE;
    int rcode = enqueue_steps(e, 1, e->profile);
    if (rcode) return rcode;
    for (auto &s : e->slots) s.fresh = false;

    for (auto &s : e->slots) s.pending = false;
    return n;
}

int h261g_engine_run(h261g_engine_t *e, int32_t n_steps)
{
    if (!e || n_steps < 0) return H261G_ERR_ARG;
    if (e->async_input) {
        int rc0 = activate_ready_inputs(e);
        if (rc0) return rc0;
    }
    if (e->profile) {
        for (int i = 0; i < n_steps; i++) {
            int rcode = h261g_engine_step(e);
            if (rcode < 0) return rcode;
        }
        return H261G_OK;
    }
    if (chain_split(e, false) && flush_writer(e)) return H261G_ERR_DEVICE;
    if (!e->graph_valid) {
        /* capture queues nothing, so a writer left by earlier steps is
         * launched now; the captured step's own writer flag is recorded and
         * applied after each replay below */
        if (flush_writer(e)) return H261G_ERR_DEVICE;
        if (e->graph_exec) (void)hipGraphExecDestroy(e->graph_exec);
        if (e->graph) (void)hipGraphDestroy(e->graph);
        if (e->graph_exec_k) (void)hipGraphExecDestroy(e->graph_exec_k);
        if (e->graph_k) (void)hipGraphDestroy(e->graph_k);
        e->graph_exec = e->graph_exec_k = nullptr;
        e->graph = e->graph_k = nullptr;
        for (int k : {1, h261g_engine::kGraphSteps}) {
            hipGraph_t &g = k == 1 ? e->graph : e->graph_k;
            hipGraphExec_t &x = k == 1 ? e->graph_exec : e->graph_exec_k;
            if (!ok(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal), "capture"))
                return H261G_ERR_DEVICE;
            int rcode = enqueue_steps(e, k, false);
            hipError_t ce = hipStreamEndCapture(e->stream, &g);
            if (rcode || !ok(ce, "end capture")) return H261G_ERR_DEVICE;
            if (!ok(hipGraphInstantiate(&x, g, nullptr, nullptr, 0), "instantiate")) return H261G_ERR_DEVICE;
        }
        /* enqueue_steps set the flag as if the captured steps had run */
        e->graph_leaves_writer = e->writer_pending;
        e->writer_pending = false;
        e->graph_valid = true;
    }
    /* a replayed step launches a pending writer first only when it was
     * captured with one (the same overlapped shape) */
    if (n_steps > 0 && !e->graph_leaves_writer && flush_writer(e)) return H261G_ERR_DEVICE;
    int left = n_steps;
    for (; left >= h261g_engine::kGraphSteps; left -= h261g_engine::kGraphSteps)
        if (!ok(hipGraphLaunch(e->graph_exec_k, e->stream), "graph launch")) return H261G_ERR_DEVICE;
    for (; left > 0; left--)
        if (!ok(hipGraphLaunch(e->graph_exec, e->stream), "graph launch")) return H261G_ERR_DEVICE;
    if (n_steps > 0) {
        for (auto &s : e->slots) s.fresh = false;
        e->writer_pending = e->graph_leaves_writer;
    }
    for (auto &s : e->slots) s.pending = false;
    return H261G_OK;
}

int h261g_engine_sync(h261g_engine_t *e)
{
    if (!e) return H261G_ERR_ARG;
    if (flush_writer(e)) return H261G_ERR_DEVICE;
    return ok(hipStreamSynchronize(e->stream), "sync") ? H261G_OK : H261G_ERR_DEVICE;
}

int h261g_stream_info(h261g_engine_t *e, int32_t id, h261g_frame_info_t *info)
{
    Slot *s = get_slot(e, id);
    if (!s || !info) return H261G_ERR_ARG;
    StreamDesc d;
    RcState rc;
    if (!ok(hipMemcpy(&d, e->desc + id, sizeof(d), hipMemcpyDeviceToHost), "hipMemcpy") ||
        !ok(hipMemcpy(&rc, e->rc + id, sizeof(rc), hipMemcpyDeviceToHost), "hipMemcpy"))
        return H261G_ERR_DEVICE;
    /* the descriptor has already advanced to the next frame */
    info->frame_num = d.frame_num - 1;
    info->frame_type = d.frame_num - 1 == 0 ? H261_FRAME_TYPE_INTRA : H261_FRAME_TYPE_INTER;
    info->quant = rc.pic_quant;
    info->bytes = rc.frame_bytes;
    info->cost = rc.picture_cost;
    return H261G_OK;
}

int h261g_stream_read_macroblocks(h261g_engine_t *e, int32_t id, h261_macroblock_t *mbs)
{
    Slot *s = get_slot(e, id);
    if (!s || !mbs) return H261G_ERR_ARG;
    std::vector<MbSide> sd(s->n_mb);
    std::vector<int8_t> lv((size_t)s->n_mb * 384);
    if (!ok(hipMemcpy(sd.data(), e->side + s->mb_base, sizeof(MbSide) * s->n_mb, hipMemcpyDeviceToHost), "hipMemcpy") ||
        !ok(hipMemcpy(lv.data(), e->levels + (size_t)s->mb_base * 384, (size_t)384 * s->n_mb,
                      hipMemcpyDeviceToHost),
            "hipMemcpy"))
        return H261G_ERR_DEVICE;
    for (int m = 0; m < s->n_mb; m++) {
        h261_macroblock_t &o = mbs[m];
        const MbSide &q = sd[m];
        o.i_macroblock_address = 0;
        o.i_macroblock_type = q.type;
        o.i_macroblock_type_flags = q.flags;
        o.i_macroblock_quant = q.quant;
        o.rgi_mv[0] = q.mvx;
        o.rgi_mv[1] = q.mvy;
        o.i_coded_block_pattern = q.cbp;
        const int8_t *l = lv.data() + (size_t)m * 384;
        const bool intra = q.type == H261_MB_TYPE_INTRA;
        for (int k = 0; k < 384; k++)
            o.rgi_tcoeff[k / 64][k % 64] = (intra && k % 64 == 0) ? (int)(uint8_t)l[k] : (int)l[k];
        o.i_mb_x = (m % s->mb_w) * 16;
        o.i_mb_y = (m / s->mb_w) * 16;
    }
    return H261G_OK;
}

int h261g_stream_read_recon(h261g_engine_t *e, int32_t id, uint8_t *y, uint8_t *cb, uint8_t *cr)
{
    Slot *s = get_slot(e, id);
    if (!s || !y || !cb || !cr) return H261G_ERR_ARG;
    StreamDesc d;
    if (!ok(hipMemcpy(&d, e->desc + id, sizeof(d), hipMemcpyDeviceToHost), "hipMemcpy")) return H261G_ERR_DEVICE;
    /* after the step's advance the last reconstruction is the reference */
    const size_t ysz = (size_t)s->width * s->height, csz = ysz / 4;
    bool good = ok(hipMemcpy(y, d.ref[0], ysz, hipMemcpyDeviceToHost), "hipMemcpy") &&
                ok(hipMemcpy(cb, d.ref[1], csz, hipMemcpyDeviceToHost), "hipMemcpy") &&
                ok(hipMemcpy(cr, d.ref[2], csz, hipMemcpyDeviceToHost), "hipMemcpy");
    return good ? H261G_OK : H261G_ERR_DEVICE;
}

int h261g_stream_read_me(h261g_engine_t *e, int32_t id, me_gpu_mb_t *me)
{
    Slot *s = get_slot(e, id);
    if (!s || !me) return H261G_ERR_ARG;
    return ok(hipMemcpy(me, e->me + s->mb_base, sizeof(MeResult) * s->n_mb, hipMemcpyDeviceToHost), "hipMemcpy")
               ? H261G_OK
               : H261G_ERR_DEVICE;
}

int h261g_stream_read_input(h261g_engine_t *e, int32_t id, uint8_t *y, uint8_t *cb, uint8_t *cr)
{
    Slot *s = get_slot(e, id);
    if (!s || !y || !cb || !cr) return H261G_ERR_ARG;
    /* the input of the next step: generation and selection queued so far
     * complete first */
    if (!ok(hipStreamSynchronize(e->in_stream), "sync") || !ok(hipStreamSynchronize(e->stream), "sync"))
        return H261G_ERR_DEVICE;
    StreamDesc d;
    if (!ok(hipMemcpy(&d, e->desc + id, sizeof(d), hipMemcpyDeviceToHost), "hipMemcpy")) return H261G_ERR_DEVICE;
    const size_t ysz = (size_t)s->width * s->height, csz = ysz / 4;
    bool good = ok(hipMemcpy(y, d.cur[0], ysz, hipMemcpyDeviceToHost), "hipMemcpy") &&
                ok(hipMemcpy(cb, d.cur[1], csz, hipMemcpyDeviceToHost), "hipMemcpy") &&
                ok(hipMemcpy(cr, d.cur[2], csz, hipMemcpyDeviceToHost), "hipMemcpy");
    return good ? H261G_OK : H261G_ERR_DEVICE;
}

int h261g_stream_read_bitstream(h261g_engine_t *e, int32_t id, uint8_t *out, int32_t cap)
{
    Slot *s = get_slot(e, id);
    if (!s || !out || cap < 0) return H261G_ERR_ARG;
    if (flush_writer(e) || (s->bits && !ok(hipStreamSynchronize(e->stream), "sync"))) return H261G_ERR_DEVICE;
    if (s->bits) {
        /* written by the device when the frame fit the stream's buffer */
        RcState rc;
        if (!ok(hipMemcpy(&rc, e->rc + id, sizeof(rc), hipMemcpyDeviceToHost), "hipMemcpy")) return H261G_ERR_DEVICE;
        if (rc.frame_bits > 0 && (rc.frame_bits + 31) / 32 <= s->bits_words) {
            if (rc.frame_bytes > cap) return H261G_ERR_FULL;
            if (!ok(hipMemcpy(out, s->bits, (size_t)rc.frame_bytes, hipMemcpyDeviceToHost), "hipMemcpy"))
                return H261G_ERR_DEVICE;
            return rc.frame_bytes;
        }
    }
    if (e->mbstruct_cap < s->n_mb) {
        hipHostFree(e->mbstruct);
        e->mbstruct = nullptr;
        e->mbstruct_cap = 0;
        if (!ok(hipHostMalloc((void **)&e->mbstruct, sizeof(h261_macroblock_t) * s->n_mb, hipHostMallocDefault),
                "hipHostMalloc"))
            return H261G_ERR_MEMORY;
        e->mbstruct_cap = s->n_mb;
    }
    int rcode = h261g_stream_read_macroblocks(e, id, e->mbstruct);
    if (rcode) return rcode;
    h261g_frame_info_t info;
    if ((rcode = h261g_stream_info(e, id, &info))) return rcode;
    return h261g_write_picture(e->mbstruct, s->width, s->height, info.frame_num, info.quant, out, cap);
}

int h261g_engine_fetch_bitstreams(h261g_engine_t *e, const int32_t *ids, int32_t n)
{
    if (!e || n < 0 || (n && !ids)) return H261G_ERR_ARG;
    for (int i = 0; i < n; i++)
        if (!get_slot(e, ids[i])) return H261G_ERR_ARG;
    if (flush_writer(e)) return H261G_ERR_DEVICE;
    /* the previous fetch's buffers are rewritten below */
    if (e->fetch_pending && !ok(hipEventSynchronize(e->fetch_ev), "event")) return H261G_ERR_DEVICE;
    e->fetch_pending = false;
    /* the host buffer grows to twice what the previous fetch needed; frames
     * that do not fit are written on the host when collected */
    size_t data = e->fetch_data_cap ? e->fetch_data_cap : ((size_t)4 << 20);
    if (e->fetch_n && e->fetch_meta && (size_t)e->fetch_meta[2 * e->fetch_n] > e->fetch_data_cap)
        data = 2 * (size_t)e->fetch_meta[2 * e->fetch_n];
    if (!e->fetch_ev && !ok(hipEventCreateWithFlags(&e->fetch_ev, hipEventDisableTiming), "event"))
        return H261G_ERR_DEVICE;
    if (n > e->fetch_cap) {
        if (!ok(hipStreamSynchronize(e->stream), "sync")) return H261G_ERR_DEVICE;
        hipHostFree(e->fetch_ids);
        hipHostFree(e->fetch_meta);
        hipFree(e->fetch_dev);
        e->fetch_ids = e->fetch_meta = e->fetch_dev = nullptr;
        e->fetch_cap = 0;
        e->fetch_n = 0;
        const int cap = n > e->max_streams ? n + n / 2 : e->max_streams;
        if (!ok(hipHostMalloc((void **)&e->fetch_ids, sizeof(int32_t) * cap, hipHostMallocDefault), "hipHostMalloc") ||
            !ok(hipHostMalloc((void **)&e->fetch_meta, sizeof(int32_t) * (2 * cap + 1), hipHostMallocNonCoherent),
                "hipHostMalloc") ||
            !ok(hipMalloc((void **)&e->fetch_dev, sizeof(int32_t) * 3 * cap), "hipMalloc"))
            return H261G_ERR_MEMORY;
        e->fetch_cap = cap;
    }
    if (data > e->fetch_data_cap) {
        /* no fetch is in flight (waited above), so the old buffer is free */
        hipHostFree(e->fetch_data);
        e->fetch_data = nullptr;
        e->fetch_data_cap = 0;
        /* written by the gather kernels only, read by the host after the
         * fetch's event: non-coherent, so the device writes it through its
         * caches in whole lines rather than word by word across the bus */
        if (!ok(hipHostMalloc((void **)&e->fetch_data, data, hipHostMallocNonCoherent), "hipHostMalloc"))
            return H261G_ERR_MEMORY;
        e->fetch_data_cap = data;
    }
    memcpy(e->fetch_ids, ids, sizeof(int32_t) * n);
    e->fetch_n = n;
    /* the ids to device memory by DMA; scratch: ids, sizes, places */
    int32_t *dev_ids = e->fetch_dev + 2 * e->fetch_cap;
    if (n && (!ok(hipMemcpyAsync(dev_ids, e->fetch_ids, sizeof(int32_t) * n, hipMemcpyHostToDevice, e->stream),
                  "hipMemcpy") ||
              launch_fetch(dev_ids, n, e->desc, e->max_streams, e->seq, e->rc, e->fetch_dev, e->fetch_dev + e->fetch_cap,
                           e->fetch_meta, e->fetch_data, (int64_t)e->fetch_data_cap, e->stream)))
        return H261G_ERR_DEVICE;
    if (!ok(hipEventRecord(e->fetch_ev, e->stream), "event")) return H261G_ERR_DEVICE;
    e->fetch_pending = true;
    return H261G_OK;
}

/* memory the device can access in place: device memory, or host memory
 * allocated pinned (hipHostMalloc) */
static bool device_accessible(const void *p)
{
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeHost || a.type == hipMemoryTypeManaged;
}

int h261g_engine_fetch_bitstreams_into(h261g_engine_t *e, const int32_t *ids, int32_t n, uint8_t *out, int64_t cap,
                                       int32_t *meta)
{
    if (!e || n < 0 || cap < 0) return H261G_ERR_ARG;
    if (n > kFetchIntoMaxStreams) {
        fprintf(stderr, "h261gpu: h261g_engine_fetch_bitstreams_into: %d streams; at most %d per call\n", n,
                kFetchIntoMaxStreams);
        return H261G_ERR_ARG;
    }
    if (n == 0) return H261G_OK;
    if (!ids || !out || !meta || ((uintptr_t)out & 3) || ((uintptr_t)meta & 3) || ((uintptr_t)ids & 3)) {
        fprintf(stderr, "h261gpu: h261g_engine_fetch_bitstreams_into: null or unaligned buffer\n");
        return H261G_ERR_ARG;
    }
    if (!device_accessible(ids) || !device_accessible(out) || !device_accessible(meta)) {
        fprintf(stderr, "h261gpu: h261g_engine_fetch_bitstreams_into: ids, out and meta must be device memory or "
                        "pinned host memory\n");
        return H261G_ERR_ARG;
    }
    if (!e->fetch_scratch) return H261G_ERR_STATE;
    if (flush_writer(e)) return H261G_ERR_DEVICE;
    /* sizes and places in the scratch's two halves */
    return launch_fetch(ids, n, e->desc, e->max_streams, e->seq, e->rc, e->fetch_scratch,
                        e->fetch_scratch + kFetchIntoMaxStreams, meta, out, cap, e->stream)
               ? H261G_ERR_DEVICE
               : H261G_OK;
}

int h261g_engine_fetched(h261g_engine_t *e, const uint8_t **data, const int32_t **sizes, const int32_t **offsets)
{
    if (!e || !data || !sizes || !offsets) return H261G_ERR_ARG;
    if (!e->fetch_ev || (!e->fetch_pending && !e->fetch_n)) return H261G_ERR_STATE;
    if (e->fetch_pending && !ok(hipEventSynchronize(e->fetch_ev), "event")) return H261G_ERR_DEVICE;
    e->fetch_pending = false;
    *data = e->fetch_data;
    *sizes = e->fetch_meta;
    *offsets = e->fetch_meta + e->fetch_n;
    return e->fetch_n;
}

int h261g_engine_profile(h261g_engine_t *e, int32_t enable)
{
    if (!e) return H261G_ERR_ARG;
    if (e->ev_read < e->ev_recorded && !ok(hipStreamSynchronize(e->stream), "sync")) return H261G_ERR_DEVICE;
    e->ev_read = e->ev_recorded;
    e->profile = enable != 0;
    e->me_ms = e->mb_ms = e->other_ms = e->bits_ms = 0;
    e->launches = 0;
    return H261G_OK;
}

int h261g_engine_kernel_times(h261g_engine_t *e, double *me_ms, double *mb_ms, double *other_ms, double *bits_ms,
                              int64_t *launches)
{
    if (!e) return H261G_ERR_ARG;
    int rcode = read_all_times(e);
    if (rcode) return rcode;
    if (me_ms) *me_ms = e->me_ms;
    if (mb_ms) *mb_ms = e->mb_ms;
    if (other_ms) *other_ms = e->other_ms;
    if (bits_ms) *bits_ms = e->bits_ms;
    if (launches) *launches = e->launches;
    return H261G_OK;
}

}  // extern "C"
