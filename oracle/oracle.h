/*
 * FP64 CPU oracle of the RF -> volume path and of the block-matching path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker that the HIP path is
 * compared against; it is never linked into, called by or shipped with the
 * library.  Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of
 * bench.py load it (through oracle/__init__.py).
 *
 * Every function restates the algorithm of a reference file (cited per
 * function) in double precision.  Parameters that the reference derives in
 * single precision on the host (samples per metre, voxel pitch, element
 * positions) are derived the same way, in float, so the oracle evaluates the
 * reference's geometry exactly and only the per-tap arithmetic is promoted.
 */
#ifndef CUT_ORACLE_H
#define CUT_ORACLE_H

#include <stddef.h>
#include <stdint.h>
#include <stdbool.h>
#include "../include/cuda_beamformer_parameters.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- Hadamard (cuda_toolkit/src/rf_processing/hadamard/hadamard_decoder.cpp) ---- */
int or_hadamard(int n, int ordering, double* H);
void or_decode(const double* conv, const double* H, int T, size_t cols, double* out);

/* ---- convert (data_conversion_kernels.cuh:25-41) ---- */
void or_convert(const void* raw, int type, const int16_t* map, size_t R0, int T, int C, int S, double* out);
/* ---- bandpass convert + demodulation + centred FIR (data_conversion_kernels.cuh:14-102) ---- */
void or_convert_demod(const void* raw, int type, const int16_t* map, size_t R0, int T, int C, int S, double fd,
                      double fs, const double* h, int L, double* out);

/* ---- FFT / Hilbert (hilbert_handler.cu:55-155, hilbert_kernels.cuh:17-40) ---- */
void or_dft(const double* re, const double* im, int n, int sign, double* ore, double* oim);
void or_fft(double* re, double* im, int n, int sign);
void or_filter_spectrum(const double* h, int L, int S, double* spec /* 2*S interleaved */);
void or_hilbert(const double* x, size_t rows, int S, const double* h, int L, double* out);

/* ---- DAS (beamformer.cu:17-98, templated_beamformer.cuh:16-324, beamformer_utils.cuh:99-213) ---- */
typedef struct or_das_consts {
    int S, C, T;
    int Nx, Ny, Nz;
    float min_x, min_y, min_z;
    float res_x, res_y, res_z;
    float xdc_min_x, xdc_min_y, xdc_max_x, xdc_max_y;
    float pitch_x, pitch_y;
    float spm;
    int delay;
    float focal_z;
    int focus_type; /* 0 none, 1 XZ_PLANE, 2 YZ_PLANE, 3 XZ_FOCUS, 4 YZ_FOCUS */
    float fn_tx, fn_rx, cw;
    int readi_groups, readi_id, encoding;
    int apo_type, to_power;
    int sequence;
    /* per-transmit geometry of the sequences of beamformer_kernels.cu */
    float fs, fc;                       /* sampling and centre frequency [Hz] */
    float tx_angle[MAX_ACQ_COUNT];      /* RCA_TPW: steering angle of transmit t = foci[3t+2] */
    int tx_element[MAX_ACQ_COUNT];      /* UFORCES: sparse_elements[t] */
} or_das_consts;

int or_das_constants(const CudaBeamformerParameters* bp, or_das_consts* k);
/* interp: 0 = Catmull-Rom cubic (reference), 1 = linear.  cf_mode: 0 = reference
 * behaviour (FORCES, UFORCES and RCA_TPW apply the coherence factor, HERCULES
 * does not), 1 = apply it for HERCULES as well.  Returns the number of taps
 * with apo above the sequence's threshold through *active (may be NULL).  out:
 * interleaved complex [Nz][Ny][Nx]. */
int or_das(const double* iq, const or_das_consts* k, int interp, int cf_mode, double* out, long long* active);
/* The same on the depth planes planes[0..n_planes) of the grid only (NULL with
 * n_planes = Nz: every plane); out: interleaved complex [n_planes][Ny][Nx]. */
int or_das_planes(const double* iq, const or_das_consts* k, int interp, int cf_mode, const int* planes, int n_planes,
                  double* out, long long* active);
int or_readi_row_bits(const or_das_consts* k, uint64_t* bits);

/* ---- 2-D block matching (image_processor.cpp:106-325, block_match_kernels.cu) ---- */
void or_zncc_valid(const double* src, int sw, int sh, int sstride,
                   const double* tpl, int tw, int th, int tstride, double* out);
int or_motion_detection(const float* images, const NCCMotionParameters* p, float* maps);

/* ---- 3-D full-search ZNCC block matching (extension of the 2-D reference to
 * volumes: blocks of `block`^3 at `stride`, shifts |d| <= margin per axis,
 * shifts that leave the volume excluded; volumes [nz][ny][nx]) ---- */
double or_zncc3d_at(const float* tpl_vol, const float* src_vol, int nx, int ny, int nz,
                    int bx, int by, int bz, int block, int dx, int dy, int dz);
/* disp: int [n_blocks][3] (dx, dy, dz), peak: float [n_blocks]; blocks in
 * (bz, by, bx) raster order; the first maximum in (dz, dy, dx) raster order wins. */
void or_ncc3d_full(const float* tpl_vol, const float* src_vol, int nx, int ny, int nz,
                   int block, int stride, int margin, int* disp, float* peak);
/* Same search restricted to |d - centre| <= radius for each block. */
void or_ncc3d_window(const float* tpl_vol, const float* src_vol, int nx, int ny, int nz,
                     int block, int stride, int margin, const int* centre, int radius, int* disp, float* peak);

/* ---- displacement-proposal CNN (no reference counterpart) ----
 * Input [B][2][16][16][16] (template, source blocks, each zero-mean unit-std);
 * L1 conv 2->16 k3 s2 p1 -> 8^3, L2 16->32 s2 -> 4^3, L3 32->32 s1 -> 4^3,
 * L4 32->32 s2 -> 2^3 (ReLU after each), FC 256 -> 3.  Parameters packed as
 * w1 b1 w2 b2 w3 b3 w4 b4 wf bf, conv weights [co][ci][kz][ky][kx].
 * or_pnet_forward: out [B][3].  or_pnet_loss_grad: loss = (1/B) sum 0.5|y-t|^2,
 * grad of the loss w.r.t. every parameter (same packing). */
int or_pnet_param_count(void);
void or_pnet_forward(const double* params, const double* in, int B, double* out);
double or_pnet_loss_grad(const double* params, const double* in, const double* target, int B, double* grad);

#ifdef __cplusplus
}
#endif
#endif
