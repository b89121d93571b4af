// Unnormalised 2-D complex-to-real inverse FFT of the half-plane spectrum.
//
// The reference scatters F[y][x0] into an Nx x (Ny/2+1) Hermitian array
// (arrange_samples, meshless_vis.cu:318-332: the x0 == cx column is dropped,
// x0 > cx goes to row x0 - 2cx + Nx, rows/columns not written stay zero) and
// runs cufftExecC2R / fftwf_execute on it (meshless_vis.cu:146,349,
// meshless_vis_cpu.cpp:62-64,261): a complex inverse DFT along x, then a
// Hermitian (c2r) inverse DFT along y.  Here, for any Nx, Ny the check accepts:
//
//   pass 1 (x): spectrum row y < cy (contiguous in x0) is scattered into a
//     length-Nx line in LDS (the reference's arrange; zero elsewhere), inverse
//     transformed, and written as row y of work[cy][Nx] (contiguous writes).
//   pass 2 (y): rows r and r + 1 of the arranged array are completed to
//     Hermitian length-Ny lines X1, X2 (imaginary part of the y = 0 term
//     ignored, y >= cy zero) and transformed together as X1 + i X2: the
//     transform of a Hermitian line is real, so the real part of the result is
//     image row r and the imaginary part image row r + 1.  The reads
//     work[y][r..r+1] of neighbouring workgroups share cache lines, so
//     workgroups are remapped to keep neighbours on one XCD (one L2).
//
// Each line is transformed by Stockham stages of radix 4 (one radix-8 or -2
// stage for an odd power of two), 3, 5, 7, 11, 13: the first stage gathers its
// inputs from global memory, the last scatters its outputs to global memory,
// the stages between run in place in LDS (a stage's butterflies are read into
// registers, then written back after a barrier).  FP32 twiddles
// exp(+2 pi i k/L), computed in FP64 on the host, come from two 128-entry
// tables in LDS.  A length with a prime factor above 13 goes
// through Bluestein's chirp-z form: X_k = c_k sum_n (x_n c_n) conj(c_{k-n}),
// c_n = exp(+pi i n^2/N), the convolution by power-of-two transforms of length
// M >= 2N - 1 against the precomputed transform of the wrapped conj(c).
// Lines up to 16384 values (Bluestein: N <= 8192) fit one workgroup's LDS.
#pragma once
#include "internal.h"

namespace mv {

constexpr int kFftEmax = 16;           // values of a line a thread holds per stage
constexpr int kFftThreads = 1024;      // largest block (lines of more than 8192 values)
constexpr int kFftThreadsSmall = 512;  // largest block otherwise (twice the registers per thread)
constexpr int kFftMaxLds = 16384;      // complex values of LDS per workgroup
constexpr int kFftTargetElems = 4096;  // LDS values per workgroup when lines are plentiful
constexpr int kXcds = 8;
constexpr int kTwiddleWords = 256;
// LDS row pitch of a line of L values (one line per row, no padding).
__host__ __device__ constexpr int lds_pitch(int L) { return L; }

struct C2rArgs {
  const float2* F;
  float2* work;
  float* image;
  int cx, cy, Nx, Ny;
  int lines;  // pass 1: spectrum rows per workgroup; pass 2: row pairs per workgroup
  int blocks;  // logical blocks per render
  // Pass 1 finalizes the spectrum itself when acc is set: F = scale * (re, -im) of the FP64 sums,
  // written to F_out (the render's spectrum) as it is gathered, instead of a finalize launch.
  const double2* acc;
  float2* F_out;
  double scale;
};

// The two passes (fft_kernels.hip; every variant compiled in a translation unit of its own):
// ODD - the line length has a radix 3, 5, 7, 11 or 13 stage; MAXT - the largest block.
template <bool ODD, int MAXT>
__global__ __launch_bounds__(MAXT) void c2r_x_kernel(C2rArgs a, FftAxis ax, const RenderOut* __restrict__ batch);
template <bool ODD, int MAXT>
__global__ __launch_bounds__(MAXT) void c2r_y_kernel(C2rArgs a, FftAxis ay, const RenderOut* __restrict__ batch);

}  // namespace mv
