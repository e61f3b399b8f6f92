/*
 * C++ span API.  Drop-in for cuda_toolkit/src/public/cuda_toolkit.hpp:17-29
 * (implementation replaced: cuda_toolkit/src/api/cuda_toolkit.cpp:109-330).
 *
 * Exported symbols keep the reference's Itanium mangling, e.g.
 *   _ZN12cuda_toolkit8beamformESt4spanIKhLm18446744073709551615EES0_IhLm18446744073709551615EERK24CudaBeamformerParameters
 * Every call is synchronous: host spans in, host spans out.  Errors return
 * false and print one line on std::cerr; no exception crosses the boundary.
 */
#ifndef CUT_CUDA_TOOLKIT_HPP
#define CUT_CUDA_TOOLKIT_HPP

#ifdef __cplusplus
#include <cstdint>
#include <span>
#include "cuda_beamformer_parameters.h"

#define EXPORT_FN __attribute__((visibility("default")))

namespace cuda_toolkit {
/* raw RF (i16 or f32, [raw_ch][rf_raw_dim[0]]) -> complex64 volume [Nz][Ny][Nx]. */
EXPORT_FN bool beamform(std::span<const uint8_t> input_data,
                        std::span<uint8_t> output_data,
                        const CudaBeamformerParameters& bp);

/* f32 frames [frame][height][width] -> float4 motion maps [frame][grid_y][grid_x]. */
EXPORT_FN bool motion_detection(std::span<const uint8_t> images,
                                std::span<uint8_t> motion_maps,
                                const NccMotionParameters& params);

/* Full-image zero-mean normalised cross correlation, valid region only. */
EXPORT_FN bool corr_images(std::span<const float> template_image,
                           std::span<const float> source_image,
                           std::span<float> corr_map,
                           const uint template_dims[2],
                           const uint source_dims[2]);
}  // namespace cuda_toolkit

#endif /* __cplusplus */
#endif /* CUT_CUDA_TOOLKIT_HPP */
