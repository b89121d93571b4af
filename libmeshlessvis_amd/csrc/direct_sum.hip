// Plain FP32 direct sum over the Hermitian half of the slice plane.
//
// Computes, for every frequency (x, y), x in [-cx+1, cx], y in [0, cy),
//     S(x, y) = sum_i w_i * Phi(r(x,y) * h_i) * exp(+2*pi*i*(x*a_i + y*b_i))
// with r = sqrt((step.x*x)^2 + (step.y*y)^2), the sum that
// sample_fourier_transform_over_grid evaluates (fourier_transform_cpu.cpp:73-134,
// fourier_transform.cu:67-141); the conjugate and the step-size scale are
// applied when the spectrum is finalised.
//
// Work decomposition (gfx950, wave64):
//  * a workgroup is 4 waves; a wave owns 64 consecutive x (one per lane) and
//    F consecutive y, so a workgroup owns a 64 x (4F) frequency tile;
//  * blockIdx.y splits the particles (split-K); each split writes its FP64
//    partial sums, reduced in order by reduce_partials (deterministic);
//  * particles stream through LDS in tiles of 256 16-byte packets (one
//    coalesced 4 KiB load per tile); every lane reads the same packet
//    (LDS broadcast);
//  * phase: the packet holds the particle's phase steps as 32-bit fixed-point
//    revolutions, so a term's phase is one wrapping integer add, converted
//    to FP32 revolutions in [-1/2, 1/2) for v_sin_f32 / v_cos_f32;
//  * accumulation: FP32 within a 256-particle tile, FP64 across tiles.
#include "internal.h"
#include "phi.h"

namespace mv {

constexpr int kTile = 256;   // packets per LDS tile
constexpr int kWaves = 4;    // waves per workgroup
constexpr int kF = 8;        // y-frequencies per lane
constexpr float kTwoPowM32 = 2.3283064365386962890625e-10f;  // 2^-32

template <bool UNIFORM, int BASIS, int MODE>
__global__ __launch_bounds__(256) void direct_sum_kernel(const Packet* __restrict__ packets, int ntiles,
                                                         int tiles_per_split, ViewParams vp, float radius,
                                                         int tiles_x, double2* __restrict__ partial) {
  __shared__ Packet tile[kTile];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int tx = blockIdx.x % tiles_x;
  const int ty = blockIdx.x / tiles_x;
  const int two_cx = 2 * vp.cx;
  const int j = tx * 64 + lane;
  const int x = j - vp.cx + 1;
  const int y0 = ty * (kWaves * kF) + wave * kF;
  const int nfreq = two_cx * vp.cy;

  // Per-frequency constants.  r is formed exactly as the reference forms it.
  const float fu = vp.sx * static_cast<float>(x);
  float fconst[kF];  // UNIFORM: Phi(r*radius); Gaussian: +-pi*log2(e)*r^2; else r
#pragma unroll
  for (int f = 0; f < kF; ++f) {
    const float fv = vp.sy * static_cast<float>(y0 + f);
    const float r = sqrtf(fu * fu + fv * fv);
    if (UNIFORM) {
      fconst[f] = phi<BASIS, MODE>(r * radius);
    } else if (BASIS == GAUSSIAN) {
      const float sgn = (MODE == VIS_PHI_REFERENCE) ? 1.0f : -1.0f;
      fconst[f] = sgn * kPi * 1.44269504088896340736f * (r * r);
    } else {
      fconst[f] = r;
    }
  }

  double dre[kF], dim[kF];
#pragma unroll
  for (int f = 0; f < kF; ++f) dre[f] = dim[f] = 0.0;

  const uint32_t ux = static_cast<uint32_t>(x);
  const uint32_t uy = static_cast<uint32_t>(y0);
  const int t_begin = blockIdx.y * tiles_per_split;
  const int t_end = min(ntiles, t_begin + tiles_per_split);
  for (int t = t_begin; t < t_end; ++t) {
    __syncthreads();
    tile[threadIdx.x] = packets[static_cast<size_t>(t) * kTile + threadIdx.x];
    __syncthreads();

    float are[kF], aim[kF];
#pragma unroll
    for (int f = 0; f < kF; ++f) are[f] = aim[f] = 0.0f;

#pragma unroll 2
    for (int p = 0; p < kTile; ++p) {
      const Packet q = tile[p];
      uint32_t P = ux * q.pu + uy * q.pv;
      const float hh = q.h * q.h;
#pragma unroll
      for (int f = 0; f < kF; ++f) {
        const float ph = static_cast<float>(static_cast<int32_t>(P)) * kTwoPowM32;
        const float s = sin_rev(ph);
        const float c = cos_rev(ph);
        float wt;
        if (UNIFORM) {
          wt = q.w;
        } else if (BASIS == GAUSSIAN) {
          wt = q.w * __builtin_amdgcn_exp2f(fconst[f] * hh);
        } else {
          wt = q.w * phi<BASIS, MODE>(fconst[f] * q.h);
        }
        are[f] = fmaf(wt, c, are[f]);
        aim[f] = fmaf(wt, s, aim[f]);
        P += q.pv;
      }
    }
#pragma unroll
    for (int f = 0; f < kF; ++f) {
      dre[f] += static_cast<double>(are[f]);
      dim[f] += static_cast<double>(aim[f]);
    }
  }

  if (j < two_cx) {
    const int x0 = (x >= 0) ? x : x + two_cx;
    double2* out = partial + static_cast<size_t>(blockIdx.y) * nfreq;
#pragma unroll
    for (int f = 0; f < kF; ++f) {
      const int y = y0 + f;
      if (y < vp.cy) {
        double re = dre[f], im = dim[f];
        if (UNIFORM) {
          re *= static_cast<double>(fconst[f]);
          im *= static_cast<double>(fconst[f]);
        }
        out[y * two_cx + x0] = make_double2(re, im);
      }
    }
  }
}

template <bool UNIFORM, int BASIS, int MODE>
static void launch_variant(dim3 grid, hipStream_t s, const Packet* packets, int ntiles, int tps, const ViewParams& vp,
                           float radius, int tiles_x, double2* partial) {
  hipLaunchKernelGGL((direct_sum_kernel<UNIFORM, BASIS, MODE>), grid, dim3(256), 0, s, packets, ntiles, tps, vp,
                     radius, tiles_x, partial);
}

template <bool UNIFORM, int BASIS>
static void launch_mode(dim3 grid, hipStream_t s, const Packet* packets, int ntiles, int tps, const ViewParams& vp,
                        float radius, int tiles_x, double2* partial) {
  if (vp.phi_mode == VIS_PHI_ANALYTIC)
    launch_variant<UNIFORM, BASIS, VIS_PHI_ANALYTIC>(grid, s, packets, ntiles, tps, vp, radius, tiles_x, partial);
  else
    launch_variant<UNIFORM, BASIS, VIS_PHI_REFERENCE>(grid, s, packets, ntiles, tps, vp, radius, tiles_x, partial);
}

template <bool UNIFORM>
static void launch_basis(dim3 grid, hipStream_t s, const Packet* packets, int ntiles, int tps, const ViewParams& vp,
                         float radius, int tiles_x, double2* partial) {
  if (vp.basis == SPH)
    launch_mode<UNIFORM, SPH>(grid, s, packets, ntiles, tps, vp, radius, tiles_x, partial);
  else if (vp.basis == GAUSSIAN)
    launch_mode<UNIFORM, GAUSSIAN>(grid, s, packets, ntiles, tps, vp, radius, tiles_x, partial);
  else
    launch_mode<UNIFORM, WENDLAND_D3_C2>(grid, s, packets, ntiles, tps, vp, radius, tiles_x, partial);
}

hipError_t launch_direct_sum(const Packet* packets, int n_padded, int ksplit, int uniform, float radius,
                             const ViewParams& vp, double2* partial, hipStream_t s) {
  const int ntiles = n_padded / kTile;
  const int tiles_x = (2 * vp.cx + 63) / 64;
  const int tiles_y = (vp.cy + kWaves * kF - 1) / (kWaves * kF);
  const int tps = (ntiles + ksplit - 1) / ksplit;
  const dim3 grid(tiles_x * tiles_y, ksplit);
  if (uniform)
    launch_basis<true>(grid, s, packets, ntiles, tps, vp, radius, tiles_x, partial);
  else
    launch_basis<false>(grid, s, packets, ntiles, tps, vp, radius, tiles_x, partial);
  return hipGetLastError();
}

// ------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void reduce_partials_kernel(const double2* __restrict__ partial, int ksplit,
                                                              int nfreq, int first, double2* __restrict__ acc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nfreq) return;
  double re = 0.0, im = 0.0;
  for (int k = 0; k < ksplit; ++k) {
    const double2 v = partial[static_cast<size_t>(k) * nfreq + i];
    re += v.x;
    im += v.y;
  }
  if (!first) {
    const double2 a = acc[i];
    re += a.x;
    im += a.y;
  }
  acc[i] = make_double2(re, im);
}

hipError_t launch_reduce_partials(const double2* partial, int ksplit, int nfreq, int first, double2* acc,
                                  hipStream_t s) {
  if (nfreq <= 0) return hipSuccess;
  hipLaunchKernelGGL(reduce_partials_kernel, dim3((nfreq + 255) / 256), dim3(256), 0, s, partial, ksplit, nfreq,
                     first, acc);
  return hipGetLastError();
}

// The reference stores (scale*sum_cos, -scale*sum_sin) (fourier_transform_cpu.cpp:305-312).
__global__ __launch_bounds__(256) void finalize_spectrum_kernel(const double2* __restrict__ acc, int nfreq,
                                                                float scale, float2* __restrict__ F) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nfreq) return;
  const double2 a = acc[i];
  const double sc = static_cast<double>(scale);
  F[i] = make_float2(static_cast<float>(a.x * sc), static_cast<float>(-a.y * sc));
}

// Batched reduction + finalize: block row r = render, 64 frequencies per block (four waves share
// each frequency's chunks, sum_chunks_block); partial index f = j * cy + y (j = x + cx - 1) is
// stored at F[y * 2cx + x0], x0 the reference column of x.
__global__ __launch_bounds__(256) void batch_finish_kernel(const RenderOut* __restrict__ r,
                                                            const JobPart* __restrict__ jp, int cx, int cy) {
  __shared__ double2 lds[256];
  const int f = blockIdx.x * 64 + (threadIdx.x & 63);
  const int two_cx = 2 * cx;
  const int nfreq = two_cx * cy;
  const bool valid = f < nfreq;
  const RenderOut ro = r[blockIdx.y];
  double re = 0.0, im = 0.0;
  for (int j = ro.job_begin; j < ro.job_end; ++j) {  // block-uniform: every thread reaches the barriers
    const JobPart p = jp[j];
    const double2 sum = sum_chunks_block(p.partial + (valid ? f : 0), p.nchunks, static_cast<size_t>(nfreq), valid, lds);
    re += sum.x;
    im += sum.y;
  }
  if (threadIdx.x >= 64 || !valid) return;
  const int j = f / cy;
  const int y = f - j * cy;
  const int x = j - cx + 1;
  const int x0 = (x >= 0) ? x : x + two_cx;
  const double sc = static_cast<double>(ro.scale);
  ro.F[y * two_cx + x0] = make_float2(static_cast<float>(re * sc), static_cast<float>(-im * sc));
}

hipError_t launch_batch_finish(const RenderOut* r, const JobPart* jp, int n_renders, int cx, int cy, hipStream_t s) {
  const int nfreq = 2 * cx * cy;
  if (nfreq <= 0 || n_renders <= 0) return hipSuccess;
  hipLaunchKernelGGL(batch_finish_kernel, dim3((nfreq + 63) / 64, n_renders), dim3(256), 0, s, r, jp, cx, cy);
  return hipGetLastError();
}

hipError_t launch_finalize_spectrum(const double2* acc, int nfreq, float scale, float2* F, hipStream_t s) {
  if (nfreq <= 0) return hipSuccess;
  hipLaunchKernelGGL(finalize_spectrum_kernel, dim3((nfreq + 255) / 256), dim3(256), 0, s, acc, nfreq, scale, F);
  return hipGetLastError();
}

}  // namespace mv
