/*
 * Parameter blocks shared by every entry point of the toolkit.
 *
 * Drop-in ABI for cuda_toolkit/src/public/cuda_beamformer_parameters.h:15-178.
 * Type names, enum values, field order and therefore every byte offset are
 * those of the reference: a caller built against the reference header can pass
 * its structs to this library unchanged.  Offsets are pinned by the
 * static_asserts at the end of this file (checked by every C++ translation unit
 * that includes it, and by tests/test_abi.py through ctypes).
 */
#ifndef CUT_BEAMFORMER_PARAMETERS_H
#define CUT_BEAMFORMER_PARAMETERS_H

#ifdef __cplusplus
extern "C" {
#else
#include <stdbool.h> /* the structs hold bool fields; C clients include this header too */
#endif

#define MAX_ACQ_COUNT 256
#define MAX_CHANNEL_COUNT 256
#define MAX_FILTER_LENGTH 1024

typedef unsigned int uint;

/* Row/column array orientation of the transmit or receive aperture. */
typedef enum RCAOrientation {
    ORIENT_NONE = 0,
    ORIENT_ROWS = 1,
    ORIENT_COLS = 2,
} RCAOrientation;

/* NORMAL = real RF samples; BANDPASS = interleaved IQ pairs (not on this path). */
typedef enum SampleType {
    SAMPLE_NORMAL = 0,
    SAMPLE_BANDPASS = 1,
} SampleType;

/* TX orientation in the high nibble, RX orientation in the low nibble. */
typedef enum {
    INVALID = 0,
    NO_TX_ROW_RX = 0x01,
    NO_TX_COL_RX = 0x02,
    ROW_TX_ROW_RX = 0x11,
    ROW_TX_COL_RX = 0x12,
    COL_TX_ROW_RX = 0x21,
    COL_TX_COL_RX = 0x22,
} PackedTROrientation;

typedef enum BeamformPlane {
    PLANE_XZ = 0,
    PLANE_YZ = 1,
    PLANE_XY = 2,
    PLANE_ARBITRARY = 3,
} BeamformPlane;

typedef enum SequenceId {
    FORCES = 0,
    UFORCES = 1,
    HERCULES = 2,
    RCA_VLS = 3,
    RCA_TPW = 4,
    UHURCULES = 5,
    RACES = 6,
    EPIC_FORCES = 7,
    EPIC_UFORCES = 8,
    EPIC_UHERCULES = 9,
    FLASH = 10,
    HERO_PA = 11,
    MIXES_S = 100,
} SequenceId;

typedef enum InputDataTypes {
    INVALID_TYPE = 0,
    TYPE_I16 = 1,
    TYPE_F32 = 2,
    TYPE_F32C = 3,
    TYPE_U8 = 4,
} InputDataTypes;

/* Transmit encoding: row order of the decode matrix. */
typedef enum EncodingMatrix {
    NONE = 0,
    HADAMARD = 1,
    WALSH = 2,
} EncodingMatrix;

typedef enum ApoType {
    RX_HANN = 0,
    RX_TO_SIN = 1,
    TX_TO_SIN = 2,
    BOTH_TO_SIN = 3,
} ApoType;

/* Block-matching settings.  Two-element arrays are indexed [0] = x (image
 * width / columns), [1] = y (rows), which is how the reference reads them
 * (cuda_toolkit/src/image_processing/image_processor.cpp:136-177). */
typedef struct NccMotionParameters {
    uint patch_size;            /* square template edge, pixels */
    uint motion_grid_spacing;   /* pixels between template centres */
    uint motion_grid_dims[2];   /* template grid size */
    uint scale_dims[2];         /* not read by the reference */
    uint search_margins[2];     /* symmetric search margin around a template */
    float abs_cor_threshold;    /* minimum accepted peak correlation */
    float rel_cor_threshold;    /* peak must beat the zero-shift value by this fraction */
    float min_patch_variance;   /* minimum paraboloid curvature at the peak */
    bool use_subpixel;          /* add the paraboloid vertex offset */
    uint reference_frame;       /* frame every other frame is compared with */
    bool neighbour_compare;     /* compare with the neighbour toward the reference */
    uint frame_count;
    uint image_dims[2];         /* [width, height] */
    InputDataTypes data_type;   /* not read by the reference */
} NCCMotionParameters;

typedef struct CudaBeamformerParameters {
    /* transducer and sequence */
    float xdc_transform[16];      /* 4x4 column-major; [12],[13] = half aperture */
    float xdc_element_pitch[2];   /* [m] {x, y} */
    uint rf_raw_dim[4];           /* [samples*transmits + pad, raw channels] */
    uint dec_data_dim[4];         /* [samples, channels, transmits] */
    EncodingMatrix decode;
    SequenceId das_shader_id;
    float time_offset;            /* [s] */

    /* output grid */
    uint output_points[4];        /* [X, Y, Z, frames] */
    float output_min_coordinate[4];
    float output_max_coordinate[4];

    float sampling_frequency;     /* [Hz] */
    float center_frequency;       /* [Hz] */
    float speed_of_sound;         /* [m/s] */

    float off_axis_pos;           /* not read */
    BeamformPlane beamform_plane; /* not read */

    float fn_tx;
    float fn_rx;
    bool interpolate;             /* not read by the reference (always cubic) */
    float coherency_weighting;

    short channel_mapping[MAX_CHANNEL_COUNT];
    short sparse_elements[MAX_CHANNEL_COUNT];
    float foci[MAX_ACQ_COUNT * 3];

    RCAOrientation tx_orientation;
    RCAOrientation rx_orientation;

    uint readi_group_count;
    uint readi_group_id;

    uint mixes_count;
    uint mixes_offset;
    uint mixes_rows[128];

    uint filter_length;
    float rf_filter[MAX_FILTER_LENGTH];

    InputDataTypes data_type;
    ApoType apo_type;
    int to_power;

    SampleType sample_type;
} CudaBeamformerParameters;

typedef struct CorrImagesParameters {
    unsigned int template_dims[2];  /* [width, height] */
    unsigned int source_dims[2];
} CorrImagesParameters;

#ifdef __cplusplus
} /* extern "C" */

#include <cstddef>
/* Byte layout of the reference structs (g++ x86-64 / hipcc host). */
static_assert(sizeof(CudaBeamformerParameters) == 8948, "BP size");
static_assert(alignof(CudaBeamformerParameters) == 4, "BP align");
static_assert(offsetof(CudaBeamformerParameters, xdc_element_pitch) == 64, "");
static_assert(offsetof(CudaBeamformerParameters, rf_raw_dim) == 72, "");
static_assert(offsetof(CudaBeamformerParameters, dec_data_dim) == 88, "");
static_assert(offsetof(CudaBeamformerParameters, decode) == 104, "");
static_assert(offsetof(CudaBeamformerParameters, das_shader_id) == 108, "");
static_assert(offsetof(CudaBeamformerParameters, time_offset) == 112, "");
static_assert(offsetof(CudaBeamformerParameters, output_points) == 116, "");
static_assert(offsetof(CudaBeamformerParameters, output_min_coordinate) == 132, "");
static_assert(offsetof(CudaBeamformerParameters, output_max_coordinate) == 148, "");
static_assert(offsetof(CudaBeamformerParameters, sampling_frequency) == 164, "");
static_assert(offsetof(CudaBeamformerParameters, center_frequency) == 168, "");
static_assert(offsetof(CudaBeamformerParameters, speed_of_sound) == 172, "");
static_assert(offsetof(CudaBeamformerParameters, off_axis_pos) == 176, "");
static_assert(offsetof(CudaBeamformerParameters, beamform_plane) == 180, "");
static_assert(offsetof(CudaBeamformerParameters, fn_tx) == 184, "");
static_assert(offsetof(CudaBeamformerParameters, fn_rx) == 188, "");
static_assert(offsetof(CudaBeamformerParameters, interpolate) == 192, "");
static_assert(offsetof(CudaBeamformerParameters, coherency_weighting) == 196, "");
static_assert(offsetof(CudaBeamformerParameters, channel_mapping) == 200, "");
static_assert(offsetof(CudaBeamformerParameters, sparse_elements) == 712, "");
static_assert(offsetof(CudaBeamformerParameters, foci) == 1224, "");
static_assert(offsetof(CudaBeamformerParameters, tx_orientation) == 4296, "");
static_assert(offsetof(CudaBeamformerParameters, rx_orientation) == 4300, "");
static_assert(offsetof(CudaBeamformerParameters, readi_group_count) == 4304, "");
static_assert(offsetof(CudaBeamformerParameters, readi_group_id) == 4308, "");
static_assert(offsetof(CudaBeamformerParameters, mixes_count) == 4312, "");
static_assert(offsetof(CudaBeamformerParameters, mixes_offset) == 4316, "");
static_assert(offsetof(CudaBeamformerParameters, mixes_rows) == 4320, "");
static_assert(offsetof(CudaBeamformerParameters, filter_length) == 4832, "");
static_assert(offsetof(CudaBeamformerParameters, rf_filter) == 4836, "");
static_assert(offsetof(CudaBeamformerParameters, data_type) == 8932, "");
static_assert(offsetof(CudaBeamformerParameters, apo_type) == 8936, "");
static_assert(offsetof(CudaBeamformerParameters, to_power) == 8940, "");
static_assert(offsetof(CudaBeamformerParameters, sample_type) == 8944, "");
static_assert(sizeof(NCCMotionParameters) == 72, "NCC params size");
static_assert(offsetof(NCCMotionParameters, search_margins) == 24, "");
static_assert(offsetof(NCCMotionParameters, use_subpixel) == 44, "");
static_assert(offsetof(NCCMotionParameters, reference_frame) == 48, "");
static_assert(offsetof(NCCMotionParameters, neighbour_compare) == 52, "");
static_assert(offsetof(NCCMotionParameters, frame_count) == 56, "");
static_assert(offsetof(NCCMotionParameters, image_dims) == 60, "");
static_assert(offsetof(NCCMotionParameters, data_type) == 68, "");
static_assert(sizeof(CorrImagesParameters) == 16, "corr params size");
#endif /* __cplusplus */

#endif /* CUT_BEAMFORMER_PARAMETERS_H */
