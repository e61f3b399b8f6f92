/*
 * By-value C shims for MATLAB's loadlibrary/calllib.  Drop-in for
 * cuda_toolkit/src/public/cuda_toolkit_matlab.h:16-23 (implementation
 * replaced: cuda_toolkit/src/api/cuda_toolkit_matlab.cpp:9-75).  Each shim
 * sizes the spans from the parameter block and forwards to the C++ API;
 * failures are reported on std::cerr only, as in the reference.
 */
#ifndef CUT_CUDA_TOOLKIT_MATLAB_H
#define CUT_CUDA_TOOLKIT_MATLAB_H

#include "cuda_beamformer_parameters.h"

#define CUT_MATLAB_EXPORT __attribute__((visibility("default")))

#ifdef __cplusplus
extern "C" {
#endif

CUT_MATLAB_EXPORT void beamform_i16_matlab(const short* data, CudaBeamformerParameters bp, float* output);
CUT_MATLAB_EXPORT void beamform_f32_matlab(const float* data, CudaBeamformerParameters bp, float* output);
CUT_MATLAB_EXPORT void motion_detect_f32_matlab(const float* images, NCCMotionParameters params, float* motion_maps);
CUT_MATLAB_EXPORT void corr_images_f32_matlab(const float* template_image, const float* source_image,
                                              CorrImagesParameters params, float* corr_map);

#ifdef __cplusplus
}
#endif
#endif /* CUT_CUDA_TOOLKIT_MATLAB_H */
