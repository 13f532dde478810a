/*
 * h261gpu.h -- C ABI of libh261gpu.so, the MI355X (gfx950) H.261 encoder
 * core.
 *
 * Part 1 is the drop-in: the 16 entry points of the reference's
 * src/gpu_device.h:7-32, with identical prototypes and the struct layouts
 * of h261gpu_abi.h.  The reference host (src/h261_main.c,
 * src/h261_moded.c, src/h261_ratectrl.c, src/h261_picture.c,
 * src/h261_gob.c, src/h261_macroblock.c, src/h261_bitstream.c) links
 * against this library unchanged in place of src/gpu_device.cu.  As in the
 * reference these calls are synchronous, operate on one process-wide
 * encoder and report errors by printing; h261gpu_last_error() additionally
 * returns the last failure.
 *
 * Part 2 is a handle-based engine with no reference counterpart: many
 * independent streams (QCIF/CIF) coded one frame step at a time with the
 * frame chain kept on the device (reference frames, denoise and
 * rate-control state), the picture quantiser chosen on the device and the
 * exact frame size counted on the device.
 */
#ifndef H261GPU_H
#define H261GPU_H

#include <stdint.h>

#include "h261gpu_abi.h"

/* exported from libh261gpu.so (built with -fvisibility=hidden) */
#if defined(__GNUC__)
#define H261G_API __attribute__((visibility("default")))
#else
#define H261G_API
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ */
/* Part 1: drop-in for src/gpu_device.h                                */
/* ------------------------------------------------------------------ */

/* device selection; src/gpu_device.cu:58-87 */
H261G_API Void h261_gpu_init(void);

/* current / reference planes; src/gpu_device.cu:90-215.  The reference
 * holds them in six cudaArray textures; here they are linear device planes
 * with stride = width.  set_* copy synchronously from the caller's buffer. */
H261G_API Void h261_gpu_device_init_textures(Int32 i_width, Int32 i_height);
H261G_API Void h261_gpu_device_deinit_textures(void);
H261G_API Void h261_gpu_device_set_current(UInt8 *pui8_frame_data, Int32 i_width, Int32 i_height);
H261G_API Void h261_gpu_device_set_current_chroma_cb(UInt8 *pui8_frame_data, Int32 i_width, Int32 i_height);
H261G_API Void h261_gpu_device_set_current_chroma_cr(UInt8 *pui8_frame_data, Int32 i_width, Int32 i_height);
H261G_API Void h261_gpu_device_set_reference(UInt8 *pui8_frame_data, Int32 i_width, Int32 i_height);
H261G_API Void h261_gpu_device_set_reference_chroma_cb(UInt8 *pui8_frame_data, Int32 i_width, Int32 i_height);
H261G_API Void h261_gpu_device_set_reference_chroma_cr(UInt8 *pui8_frame_data, Int32 i_width, Int32 i_height);

/* motion estimation; src/gpu_device.cu:286-432.  h261_gpu_device_me fills
 * ps_me_gpu->prgs_macroblocks_result[] (costs + vector per MB). */
H261G_API void h261_gpu_device_init_me(me_gpu_t **pps_me_gpu, Int32 i_width, Int32 i_height);
H261G_API void h261_gpu_device_deinit_me(me_gpu_t **pps_me_gpu);
H261G_API void h261_gpu_device_me(me_gpu_t *ps_me_gpu);

/* macroblock coder; src/gpu_device.cu:442-815.  encode fills
 * ps_ctx->ps_macroblocks->ps_macroblocks[] (type, flags, quant, mv, cbp,
 * levels); decode fills pui8_reconstructed_{Y,Cb,Cr}_host. */
H261G_API Void h261_init_macroblocks(h261_macroblocks_t **pps_mbs, Int32 i_width, Int32 i_height, Int32 i_denoise);
H261G_API Void h261_deinit_macroblocks(h261_macroblocks_t *ps_mbs);
H261G_API Void h261_gpu_device_encode_macroblocks(h261_context_t *ps_ctx);
H261G_API Void h261_gpu_device_decode_macroblocks(h261_context_t *ps_ctx);

/* last error of any call of this library (0 = none); clears it.  The
 * reference only prints cudaGetLastError (gpu_device.cu:430, 727, 813). */
H261G_API int h261gpu_last_error(void);
/* human-readable text of an error code */
H261G_API const char *h261gpu_error_string(int err);

/* ------------------------------------------------------------------ */
/* Part 2: multi-stream engine                                         */
/* ------------------------------------------------------------------ */

#define H261G_OK 0
#define H261G_ERR_ARG (-1)     /* invalid argument */
#define H261G_ERR_DEVICE (-2)  /* HIP runtime failure */
#define H261G_ERR_MEMORY (-3)  /* device or host allocation failed */
#define H261G_ERR_FULL (-4)    /* no free stream slot / HBM budget exhausted */
#define H261G_ERR_STATE (-5)   /* call out of order */

typedef struct h261g_engine h261g_engine_t;

typedef struct {
    int32_t width;       /* 352 (CIF) or 176 (QCIF) */
    int32_t height;      /* 288 or 144 */
    int32_t denoise;     /* DCT denoise strength (the reference app uses 5) */
    int32_t init_quant;  /* rate control: initial quantiser (10) */
    int32_t bitrate;     /* rate control: bit/s (512000) */
    int32_t vbv_size;    /* rate control: VBV bits (512000) */
    int32_t fixed_quant; /* > 0: picture quantiser override after rate control */
    int32_t reserved;
} h261g_stream_params_t;

/* synthetic content generated on the device (value-noise texture) */
typedef struct {
    uint64_t seed;     /* texture seed */
    int32_t sigma;     /* additive noise standard deviation (pels) */
    int32_t motion;    /* per-frame global motion drawn from [-motion, motion] */
    int32_t local;     /* 1: per-macroblock local motion field + scene cuts */
    int32_t scene_cut; /* frames between scene cuts (local mode), 0 = none */
} h261g_content_params_t;

/* max_streams: stream slots; hbm_budget: bytes of device memory the engine
 * may use for stream state (0 = no limit).  mbs_per_wave: MBs one wave of
 * the macroblock coder codes in sequence (0 = chosen per launch from the
 * batch size). */
H261G_API int h261g_engine_create(h261g_engine_t **out, int32_t max_streams, int64_t hbm_budget, int32_t mbs_per_wave);
H261G_API int h261g_engine_destroy(h261g_engine_t *e);

/* engine options (h261g_engine_config) */
#define H261G_OPT_ME_VARIANT 1   /* 0: the reference's candidates, v_sad_u8 per candidate;
                                    1: full +-15 SAD surface by v_qsad_pk_u16_u8 (same selection,
                                    same results) */
#define H261G_OPT_MBS_PER_WAVE 2 /* macroblocks one wave of the macroblock coder codes
                                    (0: chosen per launch from the batch size) */
#define H261G_OPT_FUSED_TAILS 3  /* 1: per-stream frame control runs in the tails of the two
                                    frame-step kernels; 2: rate-control begin as its own small
                                    launch between them, the rest fused; 3 (default): 2 when a
                                    step (or half of a two-chain step) has enough macroblocks to
                                    fill the device, else 1; 0: as four separate launches */
#define H261G_OPT_XCD_REMAP 4    /* bit 0: motion estimation, bit 1: macroblock coder, bit 2:
                                    bitstream writer; a set bit
                                    gives each XCD a contiguous run of macroblocks (neighbours
                                    share its L2), a clear bit has workgroup i code macroblock i.
                                    Default 3. */
#define H261G_OPT_ASYNC_INPUT 5  /* 1: h261g_stream_generate_input on a new stream generates on a
                                    second HIP stream, overlapped with frame steps; the stream
                                    joins the first step after its input is ready.  0 (default):
                                    the input is ready for the next step */
#define H261G_OPT_DEVICE_BITSTREAM 6 /* 1 (default): each step also writes every coded frame's H.261
                                        bitstream on the device, into a buffer per stream of four
                                        times the largest picture H.261 allows; 2: buffers of the
                                        worst-case frame size; 0: no device bitstream (written on the
                                        host when read).  Set while no stream is open
                                        (H261G_ERR_STATE otherwise). */
#define H261G_OPT_OVERLAP_WRITER 7 /* 1: a step's bitstream writer runs beside the next step's motion
                                      estimation, on a second HIP stream (h261g_engine_sync, reads and
                                      stream changes launch a pending writer first); 0: at the end of its
                                      own step; 2 (default): 1 when a step has enough macroblocks to
                                      fill the device, else 0 (a short step is latency bound) */
#define H261G_OPT_CHAINS 8 /* 2: the streams are split at a stream boundary near the middle of the
                              macroblock space into two halves, each stepped by its own sequence
                              of launches on its own HIP stream (the halves share nothing, so one
                              half's kernel tails overlap the other half's work; the bitstream
                              writer then runs at the end of each half's step); 1: one sequence;
                              0 (default): 2 when a step has at least twice the macroblocks that
                              fill the device, else 1.  Profiled steps
                              (h261g_engine_profile) always run as one sequence. */
#define H261G_OPT_ME_MBS_PER_WAVE 9 /* consecutive macroblocks one motion-estimation wave searches in
                                       turn: 1, 2, 4 or 8; 0 (default): chosen per launch from its
                                       size (1 while a launch is short) */
#define H261G_OPT_ME_TILE 10 /* search-window tiling: consecutive macroblocks of a motion-estimation
                                wave's run that lie side by side in one macroblock row share one
                                staged reference tile of up to this many search windows: 1
                                (every macroblock stages its own window), 2 or 4 (at most the
                                run); 0 (default): as wide as the run, at most 2 */
#define H261G_OPT_CHAIN_LADDER 11 /* two chains (H261G_OPT_CHAINS): 1: a half's motion estimation of step
                                     i+1 starts after the other half's of step i, so the halves stay half
                                     a step apart (one half's coder, frame tail and writer beside the
                                     other's search); 2: one step of slack (a half's search may overlap
                                     the other's); 0 (default): 1 while each half holds enough
                                     macroblocks to fill the device with its search alone, else 2 */
H261G_API int h261g_engine_config(h261g_engine_t *e, int32_t option, int32_t value);

/* the device the library runs on (the current HIP device) */
typedef struct {
    int32_t compute_units;
    int32_t clock_khz;
    int64_t l2_bytes;
    int64_t memory_bytes;
    char name[64];
    char arch[64];
} h261g_device_info_t;
H261G_API int h261g_device_info(h261g_device_info_t *out);

/* device bytes one stream of the given size occupies */
H261G_API int64_t h261g_stream_footprint(int32_t width, int32_t height);

/* opens a stream; returns its id (>= 0) or an H261G_ERR_* code */
H261G_API int h261g_stream_open(h261g_engine_t *e, const h261g_stream_params_t *p);
H261G_API int h261g_stream_close(h261g_engine_t *e, int32_t id);

/* input of the next step: copy host planes (Y, Cb, Cr with stride = width
 * resp. width/2) */
H261G_API int h261g_stream_put_frame(h261g_engine_t *e, int32_t id, const uint8_t *y, const uint8_t *cb, const uint8_t *cr);
/* input of the next step for n host-fed streams at once, from one host
 * buffer of frames_bytes bytes: stream ids[i]'s frame is the I420 frame at
 * frames + offsets[i] (Y, Cb, Cr planes back to back, strides width and
 * width/2; offsets multiples of 16, every frame inside the buffer).  One copy of the span from the lowest offset to the end
 * of the highest frame crosses PCIe into a device staging buffer of the
 * engine (grown on demand, kept until h261g_engine_destroy), and one kernel
 * moves each frame into its stream's input buffer and activates the stream;
 * nothing waits.  From pageable memory the runtime stages the span before
 * the call returns; from pinned memory (hipHostMalloc) it is read
 * asynchronously and must stay unchanged until the next step has finished
 * (h261g_engine_sync, a read, or a completed fetch).  Every id must name an
 * open stream without device-resident input, at most once; otherwise
 * H261G_ERR_ARG and nothing is queued.  Frames may share bytes (they are
 * only read). */
H261G_API int h261g_engine_put_frames(h261g_engine_t *e, int32_t n, const int32_t *ids, const uint8_t *frames,
                                      int64_t frames_bytes, const int64_t *offsets);
/* device-resident input: generates frames 0..n_frames-1 of synthetic
 * content into HBM; the stream then codes one per step until the sequence
 * ends (the stream stays open, inactive, afterwards) */
H261G_API int h261g_stream_generate_input(h261g_engine_t *e, int32_t id, const h261g_content_params_t *c, int32_t n_frames);
/* 1 while a stream's asynchronous input (H261G_OPT_ASYNC_INPUT) is still
 * being generated, 0 once it takes part in frame steps (from the step call
 * that activated it); host state only, no device access */
H261G_API int h261g_stream_input_pending(h261g_engine_t *e, int32_t id);
/* number of frames a stream has coded (synchronises) */
H261G_API int h261g_stream_frames_done(h261g_engine_t *e, int32_t id);
/* re-packs the macroblock index space after streams closed: the new stream
 * bases are queued behind the work queued so far and the call returns at
 * once (it waits only for the previous re-packing's staging copy) */
H261G_API int h261g_engine_compact(h261g_engine_t *e);

/* encodes one frame for every stream that has an input; asynchronous.
 * Returns the number of streams coded. */
H261G_API int h261g_engine_step(h261g_engine_t *e);
/* steps repeatedly: n_steps frame steps replayed from a captured HIP graph
 * for streams with device-generated content */
H261G_API int h261g_engine_run(h261g_engine_t *e, int32_t n_steps);
H261G_API int h261g_engine_sync(h261g_engine_t *e);

/* results of the last step of a stream (synchronise first) */
typedef struct {
    int32_t frame_num;  /* index of the coded frame */
    int32_t frame_type; /* H261_FRAME_TYPE_* */
    int32_t quant;      /* picture quantiser */
    int32_t bytes;      /* exact H.261 size of the frame in bytes */
    int32_t cost;       /* picture cost rate control used: the sum of the MBs' mode-decision
                           costs, at least 100 per MB (h261_ratectrl.c:225-245) */
} h261g_frame_info_t;
H261G_API int h261g_stream_info(h261g_engine_t *e, int32_t id, h261g_frame_info_t *info);
/* macroblocks of the last frame in the reference's h261_macroblock_t
 * layout (n_mb entries) */
H261G_API int h261g_stream_read_macroblocks(h261g_engine_t *e, int32_t id, h261_macroblock_t *mbs);
/* reconstruction of the last frame */
H261G_API int h261g_stream_read_recon(h261g_engine_t *e, int32_t id, uint8_t *y, uint8_t *cb, uint8_t *cr);
/* motion estimation results of the last frame (n_mb entries) */
H261G_API int h261g_stream_read_me(h261g_engine_t *e, int32_t id, me_gpu_mb_t *me);
/* the last frame's H.261 bitstream; returns its byte count (equal to
 * h261g_frame_info_t.bytes) or an error.  With H261G_OPT_DEVICE_BITSTREAM
 * (default) the device wrote it during the step and only its bytes are
 * copied; otherwise, or when the frame outgrew the stream's bitstream
 * buffer, it is written here on the host from the device's levels and side
 * information (the same bytes) */
H261G_API int h261g_stream_read_bitstream(h261g_engine_t *e, int32_t id, uint8_t *out, int32_t cap);
/* bulk bitstream delivery: queues, after the work queued so far, a copy of
 * the last frame's bitstream of each listed stream into pinned host memory
 * owned by the engine, and returns at once.  A later fetch first waits for
 * this one. */
H261G_API int h261g_engine_fetch_bitstreams(h261g_engine_t *e, const int32_t *ids, int32_t n);
/* waits for the last fetch; returns its stream count n.  Stream ids[i]'s
 * bytes are data[offsets[i] .. offsets[i] + sizes[i]); sizes[i] < 0: that
 * frame is not in the host buffer -- it outgrew its device buffer, device
 * bitstreams are off, or the host buffer was full (it grows for the next
 * fetch) -- read it with h261g_stream_read_bitstream.  Valid until the next
 * fetch. */
H261G_API int h261g_engine_fetched(h261g_engine_t *e, const uint8_t **data, const int32_t **sizes,
                                   const int32_t **offsets);
/* Allocation-free bulk bitstream delivery.  No counterpart in the
 * reference: an entry point for callers that cannot allocate device or host
 * memory (code under HIP graph capture, serving loops that must not call
 * hipMalloc / hipHostMalloc).  The same as h261g_engine_fetch_bitstreams,
 * but into buffers the caller owns, with the plan's scratch (per stream its
 * size and place) in a fixed device array of the library rather than in
 * memory the engine allocates; it allocates and frees nothing.
 *   ids   n stream ids, n <= H261G_FETCH_INTO_MAX_STREAMS
 *   out   cap bytes for the frames, 4-byte aligned
 *   meta  2n + 1 int32: sizes[n], then offsets[n + 1] (offsets[n]: the bytes
 *         all frames would take), as h261g_engine_fetched gives them;
 *         sizes[i] < 0: frame not on the device, past cap, or ids[i] names no
 *         open stream
 * ids, out and meta must be device memory or pinned host memory
 * (H261G_ERR_ARG otherwise, and for n above the maximum).  Queued behind the
 * work queued so far on the engine stream; returns at once, and
 * h261g_engine_sync waits for it.  Every call on one device shares the one
 * scratch array: calls that overlap on a device (two engines, or two threads)
 * race. */
#define H261G_FETCH_INTO_MAX_STREAMS 4096
H261G_API int h261g_engine_fetch_bitstreams_into(h261g_engine_t *e, const int32_t *ids, int32_t n, uint8_t *out,
                                                 int64_t cap, int32_t *meta);
/* the current input planes, for checking device-generated content
 * (synchronises the queued input generation and frame steps) */
H261G_API int h261g_stream_read_input(h261g_engine_t *e, int32_t id, uint8_t *y, uint8_t *cb, uint8_t *cr);

/* per-kernel device time accumulated since the last reset, in ms
 * (HIP events around each launch; enabled by h261g_engine_profile(e, 1)) */
H261G_API int h261g_engine_profile(h261g_engine_t *e, int32_t enable);
H261G_API int h261g_engine_kernel_times(h261g_engine_t *e, double *me_ms, double *mb_ms, double *other_ms,
                                        double *bits_ms, int64_t *launches);

/* synthetic content on the host (the same pels the device generator makes)
 * for frame `frame` of a stream */
H261G_API int h261g_content_frame_host(const h261g_content_params_t *c, int32_t width, int32_t height, int32_t frame,
                             uint8_t *y, uint8_t *cb, uint8_t *cr);

/* the H.261 writer over a frame's macroblocks, on the host: picture header,
 * GOBs, MBs, TCOEFF, in the syntax of src/h261_picture.c, src/h261_gob.c,
 * src/h261_macroblock.c.  Returns the byte count or an error. */
H261G_API int h261g_write_picture(const h261_macroblock_t *mbs, int32_t width, int32_t height, int32_t frame_num,
                        int32_t quant, uint8_t *out, int32_t cap);

#ifdef __cplusplus
}
#endif

#endif /* H261GPU_H */
