// Unnormalised 2-D complex-to-real inverse FFT of the half-plane spectrum.
//
// The reference scatters F[y][x0] into an Nx x (Ny/2+1) Hermitian array
// (arrange_samples, meshless_vis.cu:318-332: the x0 == cx column is dropped,
// x0 > cx goes to row x0 - 2cx + Nx, rows/columns not written stay zero) and
// runs cufftExecC2R / fftwf_execute on it (meshless_vis.cu:349,
// meshless_vis_cpu.cpp:261).  Here the scatter is fused into the first pass:
//
//   pass 1 (cols): for each y < cy, the F row y is contiguous in x0; it is
//     scattered into a length-Nx line in LDS (zero elsewhere) and inverse-
//     transformed along x; the result goes to work[r][y] (the arranged layout).
//   pass 2 (rows): for each r, work[r][0..cy) is completed to a Hermitian
//     length-Ny line (imaginary part of the y = 0 term ignored, columns
//     cy..Ny/2 zero) and inverse-transformed; the real part is image[r][*].
//
// Both passes run radix-2 Stockham autosort transforms in LDS with an FP32
// twiddle table exp(+2*pi*i*k/N) computed in FP64 on the host; lengths that
// are not powers of two fall back to a direct O(N^2) DFT from the same table.
#include "internal.h"

namespace mv {

constexpr int kFftThreads = 256;
constexpr int kFftLdsComplex = 4096;  // complex values per buffer (two buffers)

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// Inverse (+i) transform of `lines` lines of length N held in src; returns the
// buffer that holds the result.  Every thread of the block must call it.
__device__ float2* lds_inverse_fft(float2* src, float2* dst, int lines, int N, int log2N, const float2* __restrict__ tw) {
  const int tid = threadIdx.x;
  if (log2N >= 0) {
    const int half = N >> 1;
    const int log2half = log2N - 1;
    int s = 1, log2s = 0;
    for (int n = N; n > 1; n >>= 1) {
      const int m = n >> 1;
      for (int i = tid; i < lines * half; i += kFftThreads) {
        const int l = i >> log2half;
        const int b = i & (half - 1);
        const int p = b >> log2s;
        const int q = b & (s - 1);
        const float2* in = src + l * N;
        float2* out = dst + l * N;
        const float2 a = in[q + s * p];
        const float2 c = in[q + s * (p + m)];
        out[q + s * 2 * p] = cadd(a, c);
        out[q + s * (2 * p + 1)] = cmul(csub(a, c), tw[p * s]);
      }
      __syncthreads();
      float2* t = src;
      src = dst;
      dst = t;
      s <<= 1;
      ++log2s;
    }
    return src;
  }
  // Direct DFT for lengths that are not powers of two.
  for (int i = tid; i < lines * N; i += kFftThreads) {
    const int l = i / N;
    const int k = i - l * N;
    const float2* in = src + l * N;
    float2 acc = make_float2(0.0f, 0.0f);
    int idx = 0;
    for (int r = 0; r < N; ++r) {
      acc = cadd(acc, cmul(in[r], tw[idx]));
      idx += k;
      if (idx >= N) idx -= N;
    }
    dst[l * N + k] = acc;
  }
  __syncthreads();
  return dst;
}

// Lines y0.. of the spectrum through the first pass (x) into work[r][y].
__device__ __forceinline__ void c2r_cols_body(const float2* __restrict__ F, int cx, int cy, int Nx, int log2Nx, int Ny,
                                              int lines, int y0, const float2* __restrict__ tw,
                                              float2* __restrict__ work) {
  __shared__ float2 buf[2 * kFftLdsComplex];
  float2* b0 = buf;
  float2* b1 = buf + kFftLdsComplex;
  const int nl = min(lines, cy - y0);
  const int two_cx = 2 * cx;
  const int tid = threadIdx.x;
  for (int i = tid; i < lines * Nx; i += kFftThreads) b0[i] = make_float2(0.0f, 0.0f);
  __syncthreads();
  for (int i = tid; i < nl * two_cx; i += kFftThreads) {
    const int l = i / two_cx;
    const int x0 = i - l * two_cx;
    if (x0 == cx) continue;
    const int r = (x0 < cx) ? x0 : x0 - two_cx + Nx;
    b0[l * Nx + r] = F[static_cast<size_t>(y0 + l) * two_cx + x0];
  }
  __syncthreads();
  const float2* res = lds_inverse_fft(b0, b1, lines, Nx, log2Nx, tw);
  const int ld = Ny / 2 + 1;
  for (int i = tid; i < Nx * nl; i += kFftThreads) {
    const int r = i / nl;
    const int l = i - r * nl;
    work[static_cast<size_t>(r) * ld + y0 + l] = res[l * Nx + r];
  }
}

// Rows r0.. of work through the second pass (y, Hermitian completion) into the image.
__device__ __forceinline__ void c2r_rows_body(const float2* __restrict__ work, int cy, int Nx, int Ny, int log2Ny,
                                              int rows, int r0, const float2* __restrict__ tw,
                                              float* __restrict__ image) {
  __shared__ float2 buf[2 * kFftLdsComplex];
  float2* b0 = buf;
  float2* b1 = buf + kFftLdsComplex;
  const int nr = min(rows, Nx - r0);
  const int ld = Ny / 2 + 1;
  const int tid = threadIdx.x;
  for (int i = tid; i < rows * Ny; i += kFftThreads) b0[i] = make_float2(0.0f, 0.0f);
  __syncthreads();
  for (int i = tid; i < nr * cy; i += kFftThreads) {
    const int l = i / cy;
    const int y = i - l * cy;
    const float2 v = work[static_cast<size_t>(r0 + l) * ld + y];
    if (y == 0) {
      b0[l * Ny] = make_float2(v.x, 0.0f);
    } else {
      b0[l * Ny + y] = v;
      b0[l * Ny + Ny - y] = make_float2(v.x, -v.y);
    }
  }
  __syncthreads();
  const float2* res = lds_inverse_fft(b0, b1, rows, Ny, log2Ny, tw);
  for (int i = tid; i < nr * Ny; i += kFftThreads) {
    image[static_cast<size_t>(r0) * Ny + i] = res[i].x;
  }
}

__global__ __launch_bounds__(kFftThreads) void c2r_cols_kernel(const float2* __restrict__ F, int cx, int cy, int Nx,
                                                              int log2Nx, int Ny, int lines,
                                                              const float2* __restrict__ tw,
                                                              float2* __restrict__ work) {
  c2r_cols_body(F, cx, cy, Nx, log2Nx, Ny, lines, blockIdx.x * lines, tw, work);
}

__global__ __launch_bounds__(kFftThreads) void c2r_rows_kernel(const float2* __restrict__ work, int cy, int Nx, int Ny,
                                                              int log2Ny, int rows, const float2* __restrict__ tw,
                                                              float* __restrict__ image) {
  c2r_rows_body(work, cy, Nx, Ny, log2Ny, rows, blockIdx.x * rows, tw, image);
}

// Batched: block row blockIdx.y = render.
__global__ __launch_bounds__(kFftThreads) void c2r_cols_batch_kernel(const RenderOut* __restrict__ r, int cx, int cy,
                                                                    int Nx, int log2Nx, int Ny, int lines,
                                                                    const float2* __restrict__ tw) {
  const RenderOut ro = r[blockIdx.y];
  c2r_cols_body(ro.F, cx, cy, Nx, log2Nx, Ny, lines, blockIdx.x * lines, tw, ro.work);
}

__global__ __launch_bounds__(kFftThreads) void c2r_rows_batch_kernel(const RenderOut* __restrict__ r, int cy, int Nx,
                                                                    int Ny, int log2Ny, int rows,
                                                                    const float2* __restrict__ tw) {
  const RenderOut ro = r[blockIdx.y];
  c2r_rows_body(ro.work, cy, Nx, Ny, log2Ny, rows, blockIdx.x * rows, tw, ro.image);
}

static int ilog2_pow2(int n) {
  if (n <= 0 || (n & (n - 1))) return -1;
  int l = 0;
  while ((1 << l) < n) ++l;
  return l;
}

hipError_t launch_c2r(const float2* F, int cx, int cy, int Nx, int Ny, const float2* tw_x, const float2* tw_y,
                      float2* work, float* image, hipStream_t s) {
  if (Nx > kFftLdsComplex || Ny > kFftLdsComplex) return hipErrorInvalidValue;
  // A single render's passes have only cy and Nx lines: as many lines per workgroup as the LDS
  // holds would leave most CUs idle (256x256: 8 and 16 workgroups), so a workgroup takes just
  // enough lines for about kC2rBlocks workgroups per pass.
  constexpr int kC2rBlocks = 512;
  if (cy > 0) {
    const int lines = max(1, min(min(cy, kFftLdsComplex / Nx), (cy + kC2rBlocks - 1) / kC2rBlocks));
    hipLaunchKernelGGL(c2r_cols_kernel, dim3((cy + lines - 1) / lines), dim3(kFftThreads), 0, s, F, cx, cy, Nx,
                       ilog2_pow2(Nx), Ny, lines, tw_x, work);
  }
  const int rows = max(1, min(min(Nx, kFftLdsComplex / Ny), (Nx + kC2rBlocks - 1) / kC2rBlocks));
  hipLaunchKernelGGL(c2r_rows_kernel, dim3((Nx + rows - 1) / rows), dim3(kFftThreads), 0, s, work, cy, Nx, Ny,
                     ilog2_pow2(Ny), rows, tw_y, image);
  return hipGetLastError();
}

hipError_t launch_c2r_batch(const RenderOut* r, int n_renders, int cx, int cy, int Nx, int Ny, const float2* tw_x,
                            const float2* tw_y, hipStream_t s) {
  if (Nx > kFftLdsComplex || Ny > kFftLdsComplex) return hipErrorInvalidValue;
  if (n_renders <= 0) return hipSuccess;
  if (cy > 0) {
    const int lines = max(1, min(cy, kFftLdsComplex / Nx));
    hipLaunchKernelGGL(c2r_cols_batch_kernel, dim3((cy + lines - 1) / lines, n_renders), dim3(kFftThreads), 0, s, r, cx,
                       cy, Nx, ilog2_pow2(Nx), Ny, lines, tw_x);
  }
  const int rows = max(1, min(Nx, kFftLdsComplex / Ny));
  hipLaunchKernelGGL(c2r_rows_batch_kernel, dim3((Nx + rows - 1) / rows, n_renders), dim3(kFftThreads), 0, s, r, cy,
                     Nx, Ny, ilog2_pow2(Ny), rows, tw_y);
  return hipGetLastError();
}

}  // namespace mv
