/*
 * h261gpu_internal.h -- definitions shared by the HIP kernels and the host
 * driver of libh261gpu.so (not part of the public C ABI).
 *
 * Device data layout (one stream, plain path; the batched engine keeps
 * many of these side by side):
 *   planes      Y w*h, Cb/Cr (w/2)*(h/2) bytes, stride = width, for the
 *               current frame, the reference and the reconstruction
 *   per MB      MeResult (costs + mv), MbMode (type/flags/mv/quant), side
 *               info, int16 levels[6][64] (raster)
 *   per stream  DenoiseState (offsets + this frame's statistics), RcState
 */
#ifndef H261GPU_INTERNAL_H
#define H261GPU_INTERNAL_H

#include <stdint.h>

#include "../../include/h261gpu_abi.h"

#ifdef __HIPCC__
#define H261G_HD __host__ __device__
#else
#define H261G_HD
#endif

namespace h261g {

constexpr int kMaxStreams = 65536;

/* per-stream descriptor read by every kernel of a frame step */
struct StreamDesc {
    const uint8_t *cur[3]; /* Y, Cb, Cr of the frame being coded */
    const uint8_t *ref[3]; /* previous reconstruction */
    uint8_t *rec[3];       /* this frame's reconstruction */
    int32_t width, height, mb_w, mb_h;
    int32_t mb_base;    /* first global MB index of this stream */
    int32_t frame_type; /* H261_FRAME_TYPE_* */
    int32_t pic_quant;  /* picture quantiser (1..31) when set by the host */
    int32_t flags;      /* kStream* bits */
    int32_t frame_num;
    int32_t source_format;
    int32_t pad[2];
};

enum : int32_t {
    kStreamActive = 1,       /* encode a frame for this stream in this step */
    kStreamDeviceQuant = 2,  /* picture quantiser comes from RcState.pic_quant */
    kStreamSeqInput = 4,     /* input is a device frame sequence */
};

/* device-side per-stream input cursor and reconstruction ping-pong */
struct SeqState {
    uint8_t *base;      /* frame 0 of a device frame sequence */
    int64_t frame_size; /* bytes of one I420 frame */
    int32_t remaining;  /* frames left including the current one */
    int32_t pp;         /* which reconstruction buffer is the reference */
    uint8_t *recon[2];
    uint32_t *bits;     /* optional: the coded frame's bitstream (bytes in stream order) */
    int32_t bits_words; /* capacity of bits in 32-bit words */
    int32_t pad;
};

/* one coded MB's place in its frame's bitstream, from the frame tail to the
 * bitstream writer */
struct MbHdr {
    uint64_t code;  /* MBA, MTYPE, MQUANT, MVD and CBP codes, first bit most significant */
    int32_t bitpos; /* bit offset of the MBA code in the frame */
    uint8_t len;    /* bits in code */
    uint8_t valid;  /* 1: written by the next writer launch, which clears it */
    uint16_t bits;  /* all bits of the MB: header codes and TCOEFF layer */
};
static_assert(sizeof(MbHdr) == 16, "MbHdr layout");

/* worst-case bits of one frame: per MB an escape-coded level at every
 * coefficient of 6 blocks plus headers, plus picture and GOB headers */
constexpr int64_t frame_bits_bound(int n_mb) { return (int64_t)n_mb * (6 * (8 + 63 * 20 + 2) + 64) + 32 + 12 * 26; }

/* motion estimation result of one MB; same fields as me_gpu_mb_t */
struct MeResult {
    int32_t intra_cost, zero_cost, inter_cost, filter_cost;
    int8_t mvx, mvy;
    int16_t pad;
};
static_assert(sizeof(MeResult) == 20, "MeResult mirrors me_gpu_mb_t");

/* coding mode of one MB (h261_mb_mode_decision_t without cost) */
struct MbMode {
    int32_t type, flags, mvx, mvy, quant_adjust, cost;
};

/* side information of one coded MB, written by the MB coder */
struct MbSide {
    int8_t type;  /* 0 intra, 2 inter */
    int8_t flags; /* MC|FILTER|COEFF */
    int8_t quant;
    int8_t cbp;
    int8_t mvx, mvy;
    int16_t bits; /* MTYPE(no MQUANT)+CBP+TCOEFF bits, for the frame-size count */
};
static_assert(sizeof(MbSide) == 8, "MbSide is loaded as one 64-bit word");

/* DCT denoise state, device resident; same layout as h261_dct_denoise_t */
struct DenoiseState {
    int32_t denoise;
    int32_t intra_count;
    int32_t intra_noise[64];
    int32_t intra_offset[64];
    int32_t inter_count;
    int32_t inter_noise[64];
    int32_t inter_offset[64];
};
static_assert(sizeof(DenoiseState) == sizeof(h261_dct_denoise_t), "layout");

/* rate control state, device resident; first 9 fields as h261_ratectrl_t */
struct RcState {
    int32_t mode, current_quant, picture_cost;
    int32_t prediction_cost[2], prediction_size[2];
    int32_t bitrate, vbv_size, vbv_fill;
    int32_t fixed_quant; /* > 0: picture quantiser override */
    int32_t pic_quant;   /* quantiser chosen for the current frame */
    int32_t cost_accum;  /* sum of MD costs of the current frame (separate rate-control launch) */
    int32_t frame_bytes; /* bytes of the last coded frame */
    int32_t frame_bits;
    int32_t pad;
    /* fused rate-control begin: MBs through motion estimation (high 32 bits)
     * and the sum of their MD costs (low 32 bits), one 64-bit atomic per MB,
     * so the MB that completes the count also holds the frame's total cost
     * and no memory fence is needed */
    uint64_t me_acc;
};
static_assert(sizeof(RcState) == 72, "RcState layout");

/* Workgroups are dispatched to the 8 XCDs of the device in turn, each XCD
 * with its own L2.  Neighbouring macroblocks share reference and search
 * window lines, so logical workgroups are dealt to the XCDs in chunks of
 * kXcdChunk consecutive ones: workgroup i (XCD i % 8, its k = i / 8-th
 * there) is given logical index ((k / C) * 8 + i % 8) * C + k % C.  The grid
 * is padded to a multiple of 8 C; indexes >= n exit.  Chunks rather than one
 * contiguous run per XCD keep the XCDs' loads equal when a range of
 * macroblocks is idle (a finished stream awaiting re-packing). */
constexpr int kXcds = 8;
#ifndef H261G_XCD_CHUNK
#define H261G_XCD_CHUNK 64
#endif
constexpr int kXcdChunk = H261G_XCD_CHUNK;
H261G_HD inline int xcd_grid(int n) { return ((n + kXcds * kXcdChunk - 1) / (kXcds * kXcdChunk)) * kXcds * kXcdChunk; }
H261G_HD inline int xcd_logical(int i, int n)
{
    (void)n;
    const int k = i / kXcds;
    return ((k / kXcdChunk) * kXcds + i % kXcds) * kXcdChunk + k % kXcdChunk;
}
H261G_HD inline int launch_grid(int n, bool remap) { return remap ? xcd_grid(n) : n; }
H261G_HD inline int launch_index(int i, int n, bool remap) { return remap ? xcd_logical(i, n) : i; }

/* ---- launchers (implemented in the .hip files) ---- */

/* one wave per MB: two-stage SAD search, the three SATD costs, and (when
 * mode != nullptr) the reference mode decision for the MB.
 * Launch covers global MBs [mb_first, n_mb); mb_stream maps a global MB to
 * its stream.  Per-MB arrays are indexed by the global MB. */
struct MeLaunch {
    const StreamDesc *streams;
    const int32_t *mb_stream;
    int32_t n_mb;       /* end of the launch's MB range */
    int32_t mb_first;   /* start of the launch's MB range */
    MeResult *result;   /* [n_mb] */
    int8_t *mv_start;   /* optional [2*n_mb]: refined vector (me_gpu_t starting vector) */
    MbMode *mode;       /* optional [n_mb]: device mode decision */
    RcState *rc;        /* optional per-stream: cost accumulation */
    int32_t skip_search_on_intra; /* 1: intra frames only compute the intra cost */
    int32_t variant;              /* kMeTwoStage or kMeSurface */
    int32_t fuse_tail;            /* 1: the MB that completes a stream runs rate-control begin;
                                     0: costs only accumulate (rate-control begin is launched) */
    int32_t xcd_remap;            /* 1: contiguous macroblock runs per XCD (xcd_logical) */
    int32_t mbs_per_wave;         /* consecutive MBs one wave searches in turn (1, 2, 4 or 8; 0 = 1) */
    int32_t tile;                 /* MBs of a run that share one staged window tile (1, 2 or 4; at most
                                     mbs_per_wave; 0 = 1) */
};
enum : int32_t {
    kMeTwoStage = 0, /* the reference's candidates only, one v_sad_u8 lane per candidate */
    kMeSurface = 1,  /* full +-15 SAD surface by v_qsad_pk_u16_u8, then the same selection */
};
/* ev_start / ev_stop (hipEvent_t, optional): timestamps of the dispatch
 * itself (hipExtLaunchKernelGGL), for profiling; never under capture */
int launch_me(const MeLaunch &p, void *stream, void *ev_start = nullptr, void *ev_stop = nullptr);

/* one wave per MB: MC/loop filter, DCT, denoise, quant, dequant, IDCT,
 * reconstruction; statistics for the denoise update; bit counts. */
struct MbLaunch {
    const StreamDesc *streams;
    const int32_t *mb_stream;
    int32_t n_mb;                 /* launch covers global MBs [mb_first, n_mb) */
    int32_t mb_first;
    const MbMode *mode;           /* [n_mb] */
    DenoiseState *denoise;        /* per stream */
    RcState *rc;                  /* per stream: kStreamDeviceQuant, fused tail */
    MbSide *side;                 /* optional [n_mb] */
    int8_t *levels;               /* optional [n_mb][6][64]; intra DC (1..255) as uint8 */
    h261_macroblock_t *mb_struct; /* optional [n_mb] legacy layout */
    int32_t mbs_per_wave;         /* MBs coded by one wave in sequence */
    int32_t xcd_remap;            /* 1: contiguous macroblock runs per XCD (xcd_logical) */
};
int launch_mb(const MbLaunch &p, void *stream, void *ev_start = nullptr, void *ev_stop = nullptr);

/* per stream, 64 lanes: new denoise offsets from this frame's statistics,
 * statistics reset (src/mbgpu_kernels.cuh:478-503) */
int launch_denoise_update(DenoiseState *dn, const StreamDesc *streams, int n_streams, void *stream);

/* per stream, one thread: picture cost, rate-control picture init (streams
 * whose MB range starts in [mb_lo, mb_hi)) */
int launch_rc_begin(RcState *rc, const StreamDesc *streams, int n_streams, void *stream, int32_t mb_lo = 0,
                    int32_t mb_hi = 0x7fffffff);
/* per stream, one wave: GOB scan of the MBA / MVD / MQUANT bits, frame
 * size, rate-control update */
int launch_frame_bits(RcState *rc, const StreamDesc *streams, int n_streams, const MbSide *side,
                      int32_t *frame_bytes_out, void *stream, const SeqState *seq = nullptr, MbHdr *hdr = nullptr);
/* per stream: the reconstruction becomes the reference, next input frame */
int launch_advance(StreamDesc *streams, SeqState *seq, int n_streams, void *stream);
/* per stream, one wave: denoise update, frame size, rate-control update and
 * advance in one launch; with hdr, also each coded MB's bitstream position
 * and header codes, and the picture and GOB headers in SeqState.bits.
 * Streams whose MB range starts outside [mb_lo, mb_hi) are left alone. */
int launch_frame_tail(StreamDesc *streams, SeqState *seq, DenoiseState *dn, RcState *rc, const MbSide *side,
                      MbHdr *hdr, int n_streams, void *stream, void *ev_start = nullptr, void *ev_stop = nullptr,
                      int32_t mb_lo = 0, int32_t mb_hi = 0x7fffffff);

/* one wave per MB: the coded MBs' bits (headers from the frame tail, then
 * the TCOEFF codes of their levels) into their streams' bitstreams */
struct BitLaunch {
    const int32_t *mb_stream;
    int32_t n_mb;     /* launch covers global MBs [mb_first, n_mb) */
    int32_t mb_first;
    const SeqState *seq;
    MbHdr *hdr;
    const MbSide *side;
    const int8_t *levels;
    int32_t xcd_remap;
};
int launch_bitstream(const BitLaunch &p, void *stream, void *ev_start = nullptr, void *ev_stop = nullptr);

/* the last frame's bitstream of streams ids[0..n) (ids readable by the
 * device) into host_data (host or device memory the device can write):
 * sizes and 4-byte-aligned places into meta (sizes[n], offsets[n + 1]; size
 * -1: not on the device, past data_cap bytes, or an id of no open stream
 * among the n_slots descriptors), the bytes into host_data; dev_size /
 * dev_off: device scratch of n entries each */
int launch_fetch(const int32_t *ids, int n, const StreamDesc *desc, int n_slots, const SeqState *seq, const RcState *rc,
                 int32_t *dev_size, int32_t *dev_off, int32_t *meta, uint8_t *host_data, int64_t data_cap, void *stream);

/* h261g_engine_fetch_bitstreams_into: at most this many streams per call;
 * its scratch (sizes and places, kFetchIntoMaxStreams entries each) is a
 * file-scope device array of bitstream.hip.  fetch_into_scratch gives that
 * array's address on the current device (looked up once per device), or
 * nullptr when the lookup fails. */
constexpr int kFetchIntoMaxStreams = 4096;
int32_t *fetch_into_scratch();

}  // namespace h261g

#endif
