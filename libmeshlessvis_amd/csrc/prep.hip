// Per-view particle preparation: one pass over the registered AoS particles
// that writes the 16-byte packets the direct-sum kernels stream.
//
// The phase of particle i at frequency (x, y) of the slice plane is, in
// revolutions, x*a_i + y*b_i with a_i = step.x*dot(u_axis, p_i) and
// b_i = step.y*dot(v_axis, p_i) (the reference forms 2*pi*dot(k, p) in FP32,
// fourier_transform_cpu.cpp:300).  Only the fractional parts of a_i and b_i
// matter because x and y are integers, so they are reduced here in FP64 and
// stored as unsigned 32-bit fixed point; the kernels then form x*a + y*b
// modulo one revolution exactly, by wrapping integer arithmetic.  The
// quantisation error is at most (|x|+|y|) * 2^-33 revolutions.
#include "internal.h"

#include <algorithm>

namespace mv {

__device__ __forceinline__ uint32_t to_fixed_rev(double a) {
  const double f = a - floor(a);                       // [0, 1)
  const double scaled = f * 4294967296.0 + 0.5;        // round to nearest
  return static_cast<uint32_t>(static_cast<uint64_t>(scaled));  // 2^32 wraps to 0
}

// The packet of particle i (a zero packet for i < 0).  cull > 0: weight 0 when the
// footprint square of half-side cull * r0 around (du, dv) misses the period window
// |du| <= 1/(2 sx), |dv| <= 1/(2 sy) (the image is periodic in u and v with period
// 1/step, centred on the origin).
__device__ __forceinline__ Packet make_packet(const Constraint* __restrict__ c, const float* __restrict__ radii, int i,
                                              double sx, double sy, double3 u, double3 v, float cull) {
  Packet p;
  if (i >= 0) {
    const float4 q = reinterpret_cast<const float4*>(c)[i];
    const double px = q.x, py = q.y, pz = q.z;
    const double du = fma(u.x, px, fma(u.y, py, u.z * pz));
    const double dv = fma(v.x, px, fma(v.y, py, v.z * pz));
    p.pu = to_fixed_rev(sx * du);
    p.pv = to_fixed_rev(sy * dv);
    p.w = q.w;
    p.h = radii ? radii[i] : 1.0f;
    if (cull > 0.0f) {
      const double r = static_cast<double>(cull) * p.h, hu = 0.5 / fabs(sx), hv = 0.5 / fabs(sy);
      if (du + r < -hu || du - r > hu || dv + r < -hv || dv - r > hv) p.w = 0.0f;
    }
  } else {
    p.pu = 0u;
    p.pv = 0u;
    p.w = 0.0f;
    p.h = 0.0f;
  }
  return p;
}

// Block 0 also clears `nzero` (< 256) ints at `zero`: the work-queue counters of the MFMA launches
// that follow, cleared here instead of by a launch of their own.
__device__ __forceinline__ void clear_counters(int* zero, int nzero) {
  if (blockIdx.x == 0 && static_cast<int>(threadIdx.x) < nzero) zero[threadIdx.x] = 0;
}

__global__ __launch_bounds__(256) void prep_packets_kernel(const Constraint* __restrict__ c,
                                                           const float* __restrict__ radii, int n, int n_padded,
                                                           double sx, double sy, double3 u, double3 v, float cull,
                                                           Packet* __restrict__ out, int* zero, int nzero) {
  clear_counters(zero, nzero);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_padded) return;
  out[i] = make_packet(c, radii, i < n ? i : -1, sx, sy, u, v, cull);
}

// Every job's packets in one launch: block b prepares slots [(b - begin[j]) * kPrepBlock, ...)
// of job j = block_job[b] (a table lookup, uniform over the block).
__global__ __launch_bounds__(kPrepBlock) void prep_jobs_kernel(const PrepJob* __restrict__ jobs,
                                                               const long long* __restrict__ begin,
                                                               const int* __restrict__ block_job, long long n_blocks,
                                                               int* zero, int nzero) {
  clear_counters(zero, nzero);
  const long long b = blockIdx.x;
  if (b >= n_blocks) return;
  const int j = block_job[b];
  const PrepJob& J = jobs[j];
  const int k = static_cast<int>((b - begin[j]) * kPrepBlock) + static_cast<int>(threadIdx.x);
  if (k >= J.n_slots) return;
  const int i = k < J.n ? k : -1;
  J.out[k] = make_packet(J.c, J.radii, i, J.sx, J.sy, J.u, J.v, J.cull);
}

// Grid of a prep launch: a block per 256 slots, at least one when it clears counters.
static unsigned prep_grid(long long items, int nzero) {
  return static_cast<unsigned>(std::max<long long>(nzero > 0 ? 1 : 0, (items + 255) / 256));
}

hipError_t launch_prep_jobs(const PrepJob* jobs, const long long* begin, const int* block_job, long long n_blocks,
                            hipStream_t s, int* zero, int nzero) {
  if (nzero > 256) return hipErrorInvalidValue;
  n_blocks = std::max(0LL, n_blocks);
  const long long grid = std::max<long long>(nzero > 0 ? 1 : 0, n_blocks);
  if (grid == 0) return hipSuccess;
  if (grid > 0x7fffffffLL) return hipErrorInvalidValue;
  hipLaunchKernelGGL(prep_jobs_kernel, dim3(static_cast<unsigned>(grid)), dim3(kPrepBlock), 0, s, jobs, begin,
                     block_job, n_blocks, zero, nzero);
  return hipGetLastError();
}

hipError_t launch_prep_packets(const Constraint* d_constraints, const float* d_radii, int n, int n_padded,
                               double sx, double sy, double3 u, double3 v, float cull, Packet* out, hipStream_t s,
                               int* zero, int nzero) {
  if (nzero > 256) return hipErrorInvalidValue;
  const unsigned grid = prep_grid(std::max(0, n_padded), nzero);
  if (grid == 0) return hipSuccess;
  hipLaunchKernelGGL(prep_packets_kernel, dim3(grid), dim3(256), 0, s, d_constraints, d_radii, n, n_padded, sx, sy,
                     u, v, cull, out, zero, nzero);
  return hipGetLastError();
}

float cull_factor(const VisConfig* c, int basis) {
  if (!c || !c->cull_fully_aliased_terms) return 0.0f;
  return basis == SPH ? 2.0f : basis == GAUSSIAN ? 3.0f : 1.0f;
}

}  // namespace mv
