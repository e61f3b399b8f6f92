/*
 * C API over OpenGL shader-storage buffers.  Drop-in for
 * cuda_toolkit/src/public/cuda_toolkit_ogl.h:37-75 (implementation replaced:
 * cuda_toolkit/src/api/cuda_toolkit_ogl.cpp:105-273).  GL buffers are shared
 * with the GPU through HIP-GL interop (hipGraphicsGLRegisterBuffer).
 *
 * Call order (as in the reference): init_cuda_configuration and
 * register_cuda_buffers in any order; cuda_set_channel_mapping before the
 * first cuda_decode; cuda_set_match_filter after init.  All calls are
 * synchronous and return false with a message on std::cerr on failure.
 */
#ifndef CUT_CUDA_TOOLKIT_OGL_H
#define CUT_CUDA_TOOLKIT_OGL_H

#include <stddef.h>
#include <stdint.h>
#ifndef __cplusplus
#include <stdbool.h>
#endif

#ifndef MAX_CHANNEL_COUNT
#define MAX_CHANNEL_COUNT 256
#endif

#define CUT_OGL_EXPORT __attribute__((visibility("default")))

#ifdef __cplusplus
extern "C" {
#endif
typedef unsigned int uint;
typedef uint16_t u16;
typedef int16_t i16;

/* input_dims = {samples*transmits + pad, raw channels};
 * decoded_dims = {samples, channels, transmits}.  Replaces
 * cuda_toolkit_ogl.cpp:105-117 (Hadamard decode order). */
CUT_OGL_EXPORT bool init_cuda_configuration(const uint* input_dims, const uint* decoded_dims);

/* Registers up to 16 complex64 RF buffers and one raw int16 buffer.
 * Replaces cuda_toolkit_ogl.cpp:127-160. */
CUT_OGL_EXPORT bool register_cuda_buffers(const uint* rf_data_ssbos, uint rf_buffer_count, uint raw_data_ssbo);

/* 256-entry map from decoded channel to raw channel.  Replaces cuda_toolkit_ogl.cpp:162-175. */
CUT_OGL_EXPORT bool cuda_set_channel_mapping(const i16 channel_mapping[MAX_CHANNEL_COUNT]);

/* Match filter taps; length 0 disables filtering.  The spectrum stored is the
 * DFT of the zero-padded taps (circular convolution, no time reversal), as in
 * cuda_toolkit/src/rf_processing/rf_ffts/hilbert_handler.cu:55-78. */
CUT_OGL_EXPORT bool cuda_set_match_filter(const float* match_filter, uint length);

/* Hilbert transform of the real parts of RF buffer in -> complex RF buffer out.
 * Replaces cuda_toolkit_ogl.cpp:237-273. */
CUT_OGL_EXPORT bool cuda_hilbert(uint input_buffer_idx, uint output_buffer_idx);

/* Convert + channel-remap + Hadamard-decode raw int16 at a byte offset into
 * the real parts of RF buffer out.  Replaces cuda_toolkit_ogl.cpp:192-235. */
CUT_OGL_EXPORT bool cuda_decode(size_t input_offset, uint output_buffer_idx);

/* Frees decode state and unregisters every buffer.  Replaces cuda_toolkit_ogl.cpp:119-125. */
CUT_OGL_EXPORT void deinit_cuda_configuration(void);

#ifdef __cplusplus
}
#endif
#endif /* CUT_CUDA_TOOLKIT_OGL_H */
