/*
 * Out-of-process transfer protocol: the Linux form of the reference harness's
 * wire format (test_app/src/communication_params.h:1-65) and of its client
 * API (test_app/client_lib/cuda_transfer.h:16-28).  It lets a process that
 * does not link the GPU library (MATLAB, Python, a different runtime) drive
 * a server process that does.
 *
 * Transport.  The reference uses a Windows named pipe for commands and two
 * named file mappings for data and parameters.  Here:
 *   command channel  a UNIX-domain SOCK_SEQPACKET socket (message boundaries
 *                    kept, bidirectional like the named pipe), path
 *                    /tmp/<name>_command.sock
 *   data segment     POSIX shared memory /<name>_data (CUT_DATA_SMEM_SIZE
 *                    bytes by default; the client reads the size with fstat)
 *   parameters       POSIX shared memory /<name>_parameters
 *                    (one SharedMemoryParams)
 * <name> is CUT_TRANSFER_DEFAULT_NAME unless the environment variable
 * CUT_TRANSFER_NAME sets it (client) or --name does (server).
 *
 * Exchange, as in the reference: the client copies its parameters into the
 * parameters segment and its input into the data segment, sends a
 * CommandPipeMessage {opcode, data_size}; the server answers ACK, computes,
 * writes the output to the start of the data segment and answers SUCCESS
 * with the output size in data_size, or ERR.  The server is the program
 * cut_transfer_server (built next to libcuda_toolkit.so); the client is
 * libcut_transfer.so (plain C, no GPU dependency).
 *
 * Differences from the reference, all on error paths: SVD_FILTER (which the
 * reference server logs and leaves unanswered, so its client waits out the
 * 20-minute timeout) and unknown opcodes are answered with ERR; a failed
 * motion detection is answered with ERR (the reference reports SUCCESS with
 * a zero map); a handler error is answered with ERR and the server keeps
 * serving (the reference server throws and exits); the client checks the
 * output size the server reports against the size the call expects.
 */
#ifndef CUT_TRANSFER_H
#define CUT_TRANSFER_H

#include "cuda_beamformer_parameters.h"

#ifdef __cplusplus
extern "C" {
#endif

#define CUT_TRANSFER_DEFAULT_NAME "cut_cuda"

#define MEGABYTE (1024U * 1024U)
#define GIGABYTE (1024U * 1024U * 1024U)

#define COMMAND_PIPE_SIZE MEGABYTE
#define DATA_SMEM_SIZE 0x80000000UL /* 2 GiB: the server's default data segment */

/* What the client asks for (communication_params.h:21-30). */
typedef enum CudaCommand {
    ERR = 0,
    ACK = 1,
    SUCCESS = 2,
    BEAMFORM_VOLUME = 3,
    SVD_FILTER = 4,
    NCC_MOTION_DETECT = 5,
    CORR_IMAGES = 6,
} CudaCommand;

typedef struct CommandPipeMessage {
    CudaCommand opcode;
    unsigned long long data_size; /* input bytes (request) or output bytes (SUCCESS) */
    int frame_count;
} CommandPipeMessage;

typedef struct SVDParameters {
    int frame_count;
    int filter_indicies[256]; /* singular values to remove (name as in the reference) */
} SVDParameters;

typedef struct SharedMemoryParams {
    CudaBeamformerParameters beamformerParameters;
    SVDParameters svdParameters;
    NCCMotionParameters nccMotionParameters;
    CorrImagesParameters corrImagesParameters;
} SharedMemoryParams;

/* ---- client API (libcut_transfer.so), cuda_transfer.h:16-28 ------------- */
/* Raw RF int16 / f32 [rf_raw_dim[1]][rf_raw_dim[0]] -> complex64 volume
 * (output_points[0..2] x 2 floats). */
void beamform_i16(const short* data, CudaBeamformerParameters bp, float* output);
void beamform_f32(const float* data, CudaBeamformerParameters bp, float* output);
/* f32 frames [frame_count][image_dims[1]][image_dims[0]] -> float4 maps
 * [frame_count][motion_grid_dims[1]][motion_grid_dims[0]]. */
void motion_detect_f32(const float* images, NCCMotionParameters params, float* motion_maps);
/* Valid-region ZNCC map ((source - template + 1) per axis, floats). */
void corr_images_f32(const float* template_image, const float* source_image, CorrImagesParameters params,
                     float* corr_map);
/* Unmap the segments and close the command channel. */
void deinit(void);

/* Extension (no counterpart in the reference, whose calls return void and
 * only print errors): 1 if the last call above completed, 0 if it failed. */
int cut_transfer_last_ok(void);

#ifdef __cplusplus
} /* extern "C" */
#endif

#ifdef __cplusplus
static_assert(sizeof(CommandPipeMessage) == 24, "CommandPipeMessage layout");
static_assert(sizeof(SVDParameters) == 1028, "SVDParameters layout");
static_assert(sizeof(SharedMemoryParams) == 8948 + 1028 + 72 + 16, "SharedMemoryParams layout");
#else
_Static_assert(sizeof(CommandPipeMessage) == 24, "CommandPipeMessage layout");
_Static_assert(sizeof(SVDParameters) == 1028, "SVDParameters layout");
_Static_assert(sizeof(SharedMemoryParams) == 8948 + 1028 + 72 + 16, "SharedMemoryParams layout");
#endif

#endif /* CUT_TRANSFER_H */
