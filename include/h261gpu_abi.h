/*
 * h261gpu_abi.h -- data layouts shared between the encoder host (mode
 * decision, rate control, bitstream writer) and the device layer.
 *
 * These structs are the drop-in contract: they must be byte-for-byte
 * identical to the ones in the reference's src/h261_types.h:46-192 and
 * src/common.h:1-21, because the reference host code (src/h261_moded.c,
 * src/h261_gob.c, src/h261_macroblock.c, src/h261_main.c) reads and writes
 * them directly and passes pointers to them across the device C ABI.
 * Field order, field types and therefore offsets are fixed by that
 * contract; tests/test_oracle_reference.py compares every size and offset
 * against the reference headers compiled by gcc.
 *
 * Sizes on x86-64 (gcc, LP64):
 *   me_gpu_mv_t 2, me_gpu_mb_t 20, me_gpu_t 88, h261_mb_mode_decision_t 24,
 *   h261_macroblock_t 1572, h261_macroblocks_t 208, h261_dct_denoise_t 1036.
 */
#ifndef H261GPU_ABI_H
#define H261GPU_ABI_H

#ifdef __cplusplus
extern "C" {
#endif

/* scalar names used by the reference's prototypes (src/common.h:2-15) */
typedef void Void;
typedef char Int8;
typedef unsigned char UInt8;
typedef short Int16;
typedef unsigned short UInt16;
typedef int Int32;
typedef unsigned int UInt32;
typedef long long Int64;
typedef unsigned long long UInt64;

/* picture / macroblock constants (src/h261_types.h:2-42) */
#define H261_FRAME_TYPE_INTRA 0
#define H261_FRAME_TYPE_INTER 1
#define H261_SOURCE_FORMAT_QCIF 0
#define H261_SOURCE_FORMAT_CIF 1
#define H261_MB_TYPE_INTRA 0
#define H261_MB_TYPE_INTER 2
#define H261_MB_MC 1
#define H261_MB_FILTER 2
#define H261_MB_COEFF 4
#define H261_RATECTRL_MODE_QUANT 1
#define H261_RATECTRL_MODE_CBR 2

/* motion estimation output (src/h261_types.h:46-79) */
typedef struct {
    Int8 i8_mx;
    Int8 i8_my;
} me_gpu_mv_t;

typedef struct {
    Int32 i_intra_cost;
    Int32 i_zero_cost; /* never written by the reference device layer */
    Int32 i_16x16_cost;
    Int32 i_16x16_filter_cost;
    me_gpu_mv_t s_16x16_mv;
} me_gpu_mb_t;

typedef struct me_gpu_s {
    Int32 i_mb_width;
    Int32 i_mb_height;
    Int32 i_num_mb;
    me_gpu_mb_t *prgs_macroblocks_device;
    me_gpu_mb_t *prgs_macroblocks_result; /* host-visible results */
    me_gpu_mv_t *ps_candidate_vector_device;
    me_gpu_mv_t *ps_candidate_vector_host;
    me_gpu_mv_t *ps_starting_vector_device;
    me_gpu_mv_t *ps_starting_vector_host;
    Int8 *pi8_motion_vector_limits_device;
    Int8 *pi8_motion_vector_limits_host;
    Void *p_cuda; /* opaque backend handle */
} me_gpu_t;

/* mode decision (src/h261_types.h:84-98) */
typedef struct {
    Int32 i_cost;
    Int32 i_quantiser_adjust;
    Int32 i_mb_type;
    Int32 i_mb_flags;
    Int32 i_mv_x, i_mv_y;
} h261_mb_mode_decision_t;

typedef struct {
    Int32 i_mb_width;
    Int32 i_mb_height;
    h261_mb_mode_decision_t *prgs_mb;
} h261_mode_decision_t;

/* frames, bitstream, denoise (src/h261_types.h:102-132) */
typedef struct {
    Int32 i_width;
    Int32 i_height;
    Int32 i_stride_y;
    Int32 i_stride_c;
    UInt8 *pui8_Y;
    UInt8 *pui8_Cr;
    UInt8 *pui8_Cb;
} h261_frame_t;

typedef struct {
    UInt8 *pui8_bitstream;
    Int32 i_length;
    Int32 i_byte_count;
    Int32 i_next_bit;
} h261_bitstream_t;

typedef struct {
    Int32 i_denoise;
    Int32 i_intra_count;
    Int32 rgi_intra_noise[64];
    Int32 rgi_intra_signal_offset[64];
    Int32 i_inter_count;
    Int32 rgi_inter_noise[64];
    Int32 rgi_inter_signal_offset[64];
} h261_dct_denoise_t;

/* one coded macroblock as consumed by the writer (src/h261_types.h:137-151) */
typedef struct h261_macroblock_s {
    Int32 i_macroblock_address;
    Int32 i_macroblock_type;
    Int32 i_macroblock_type_flags;
    Int32 i_macroblock_quant;
    Int32 rgi_mv[2];
    Int32 i_coded_block_pattern;
    Int32 rgi_tcoeff[6][64]; /* quantised levels, raster order */
    Int32 i_mb_x;            /* pel coordinates */
    Int32 i_mb_y;
} h261_macroblock_t;

/* macroblock coder state (src/h261_types.h:154-192) */
typedef struct {
    Int32 i_mb_width;
    Int32 i_mb_height;
    Int32 i_num_mb;
    h261_macroblock_t *ps_macroblocks;
    h261_macroblock_t *ps_macroblocks_device;
    h261_dct_denoise_t *ps_denoise_host;
    h261_dct_denoise_t *ps_denoise_device;
    Int32 *pi_mb_types_host;
    Int32 *pi_mb_types_device;
    Int32 *pi_mb_flags_host;
    Int32 *pi_mb_flags_device;
    Int32 *pi_mb_quant_host;
    Int32 *pi_mb_quant_device;
    Int32 *pi_mb_mv_x_host;
    Int32 *pi_mb_mv_x_device;
    Int32 *pi_mb_mv_y_host;
    Int32 *pi_mb_mv_y_device;
    UInt8 *pui8_reconstructed_Y_device;
    UInt8 *pui8_reconstructed_Y_host;
    UInt8 *pui8_reconstructed_Cr_device;
    UInt8 *pui8_reconstructed_Cr_host;
    UInt8 *pui8_reconstructed_Cb_device;
    UInt8 *pui8_reconstructed_Cb_host;
    Int32 i_num_blocks;
    Int32 i_num_coeff_blocks;
    Int32 *pi_coeff_blocks_device;
    Int32 *pi_coeff_blocks_host;
    Void *p_cuda;
} h261_macroblocks_t;

/* picture-level syntax state (src/h261_types.h:195-242) */
typedef struct h261_mb_coding_cache_t {
    Int32 i_quantiser;
    Int32 i_last_mv_valid;
    Int32 i_last_mv_x;
    Int32 i_last_mv_y;
} h261_mb_coding_vars_t;

typedef struct {
    Int32 i_group_of_blocks_start_code;
    Int32 i_group_number;
    Int32 i_quantiser;
    Int32 i_num_extra_insertion_information;
    h261_mb_coding_vars_t s_coding_vars;
} h261_gob_t;

typedef struct {
    Int32 i_picture_start_code;
    Int32 i_temporal_reference;
    Int32 i_split_screen_indicator;
    Int32 i_document_camera_indicator;
    Int32 i_freeze_picture_release;
    Int32 i_source_format;
    Int32 i_highres_mode;
    Int32 i_spare_bit;
    Int32 i_num_extra_insertion_information;
    UInt8 rgui8_extra_insertion_information[1];
    h261_gob_t s_groups_of_blocks;
    Int32 i_quantiser;
    Int32 i_frame_type;
} h261_picture_t;

/* rate control (src/h261_types.h:252-264) */
typedef struct {
    Int32 i_ratectrl_mode;
    Int32 i_current_quant;
    Int32 i_current_picture_cost;
    Int32 i_prediction_cost[2];
    Int32 i_prediction_size[2];
    Int32 i_bitrate;
    Int32 i_vbv_size;
    Int32 i_vbv_fill;
} h261_ratectrl_t;

/* encoder context passed to encode/decode (src/h261_types.h:267-284) */
typedef struct {
    Int32 i_source_format;
    Int32 i_frame_counter;
    h261_frame_t s_current_frame;
    h261_frame_t s_reference_frame;
    h261_picture_t s_picture;
    h261_bitstream_t s_bitstream;
    h261_mode_decision_t *ps_mode_decision;
    h261_ratectrl_t s_ratectrl;
    h261_macroblocks_t *ps_macroblocks;
} h261_context_t;

typedef struct {
    Int32 i_frame_num;
    Int32 i_source_format;
    Int32 i_quantiser;
    Int32 i_frame_type;
} h261_picture_parameters_t;

#ifdef __cplusplus
}
#endif

#endif /* H261GPU_ABI_H */
