// Internal declarations shared by the host runtime and the HIP kernel files of
// libmeshless_vis.so.  Nothing here is part of the C ABI.
#pragma once

#include "../../include/meshless_vis.h"
#include "errors.h"

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

namespace mv {

// ---------------------------------------------------------------- errors ---
bool hip_ok(hipError_t e, const char* what);  // records VIS_ERROR_HIP / OOM on failure
// hipMalloc through the library's allocation-failure hook; false (error recorded) on failure.
bool device_malloc(void** p, size_t bytes, const char* what);

// ------------------------------------------------------- particle packets ---
// Per-view, per-particle record written by the prep kernel and streamed by the
// direct-sum kernels: the two fixed-point phase steps (revolutions * 2^32,
// reduced mod 1 in FP64), the weight and the radius.
struct alignas(16) Packet {
  uint32_t pu;  // frac(step.x * dot(u_axis, position)) * 2^32
  uint32_t pv;  // frac(step.y * dot(v_axis, position)) * 2^32
  float w;      // weight
  float h;      // radius r0 (1 when the group has no radii)
};

// Registration facts about a group, keyed by its device constraint pointer.
struct GroupInfo {
  int n = 0;                 // real particle count (number_of_terms)
  int padded = 0;            // d_number_of_terms
  bool has_radii = false;
  bool uniform_radius = false;  // every radius equal (or no radii)
  float radius = 1.0f;          // the common radius when uniform_radius
  // Range of |r0| over the nonzero radii (Phi is even in r0, so a radius bins by its
  // magnitude); has_zero_h: some radius is 0 (those particles get a bin of their own).
  float h_min = 1.0f, h_max = 1.0f;
  bool has_zero_h = false;
  float w_absmax = 0.0f;
  // Unique per registration (set by registry_put): device state derived from a
  // group (bins, captured graphs) is keyed by it, because a later registration can
  // reuse the same device addresses.
  unsigned long long serial = 0;
};
// The facts above from a group's host (or host-copied) arrays; radii may be nullptr.
void group_stats(const Constraint* constraints, const float* radii, int n, GroupInfo* info);
void registry_put(const void* key, const GroupInfo& info);
bool registry_get(const void* key, GroupInfo* info);
void registry_erase(const void* key);

// ------------------------------------------------------------ device buffers
// Grow-only device scratch owned by a plan.
struct DevBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
  bool reserve(size_t need);  // false (error recorded) on failure
  void release();
  template <class T> T* as() const { return static_cast<T*>(ptr); }
};

// Start/stop events of one timed kernel launch.
using EventPair = std::pair<hipEvent_t, hipEvent_t>;

// ----------------------------------------------------- view description ---
// Everything a kernel needs to know about one view, passed by value.
struct ViewParams {
  int cx, cy;          // cutoff frequency: x in [-cx+1, cx], y in [0, cy)
  float sx, sy;        // step_size
  int basis;           // BasisFunctionId
  int phi_mode;        // VisPhiMode
};

// sum_k p[k * stride] over k < n in FP64, in a fixed order (deterministic run to run): four
// running sums s_q over the chunks k = q mod 4, each in increasing k, then (s0 + s1) + (s2 + s3).
// The four waves of a 256-thread block share the work, one frequency per lane: wave q keeps
// s_q and the block adds the four through LDS, so four waves' loads are in flight per frequency.
// Every thread of the block calls it (it holds two barriers); `valid` lanes read p, the others
// add nothing.  The sum is returned in wave 0 (threads 0..63); lds holds 256 double2.
__device__ __forceinline__ double2 sum_chunks_block(const float2* __restrict__ p, int n, size_t stride, bool valid,
                                                    double2* lds) {
  const int lane = threadIdx.x & 63, q = threadIdx.x >> 6;
  double r = 0.0, s = 0.0;
  if (valid) {
#pragma unroll 4
    for (int k = q; k < n; k += 4) {
      const float2 v = p[static_cast<size_t>(k) * stride];
      r += static_cast<double>(v.x);
      s += static_cast<double>(v.y);
    }
  }
  lds[q * 64 + lane] = make_double2(r, s);
  __syncthreads();
  double2 out = make_double2(0.0, 0.0);
  if (q == 0) {
    const double2 a = lds[lane], b = lds[64 + lane], c = lds[128 + lane], d = lds[192 + lane];
    out = make_double2((a.x + b.x) + (c.x + d.x), (a.y + b.y) + (c.y + d.y));
  }
  __syncthreads();  // lds free for the next call
  return out;
}

// ------------------------------------------------------ kernel launchers ---
// prep.hip
// cull > 0: a particle whose projected footprint, the square of half-side cull * r0
// around (p.u, p.v), lies wholly outside the image's period window
// [-1/(2 sx), 1/(2 sx)] x [-1/(2 sy), 1/(2 sy)] (a fully aliased term) gets weight 0.
// zero, nzero (at most 256): ints the launch also clears (the work-queue counters of the MFMA
// launches that follow); nullptr, 0 for none.
hipError_t launch_prep_packets(const Constraint* d_constraints, const float* d_radii, int n, int n_padded,
                               double sx, double sy, double3 u, double3 v, float cull, Packet* out, hipStream_t s,
                               int* zero = nullptr, int nzero = 0);

// The cull factor of a group under a config: 0 unless cull_fully_aliased_terms is set;
// else the basis's support in units of r0 (SPH 2, Wendland 1, Gaussian 3: exp(-9 pi)).
float cull_factor(const VisConfig* c, int basis);

// One job of a batched packet preparation: the packets of a (view, group) pair.
struct PrepJob {
  const Constraint* c;  // slot k holds particle k (k < n), zeros past n
  const float* radii;   // nullptr: no radii
  Packet* out;
  int n, n_slots;
  double sx, sy;
  double3 u, v;
  float cull;  // as launch_prep_packets
  int pad;
};
// Every job's packets in one launch of kPrepBlock-thread blocks, each within one job:
// block_job[b] = the job of block b, begin[j] = the first block of job j, n_blocks = all
// blocks (the sum over jobs of ceil(n_slots / kPrepBlock)).
constexpr int kPrepBlock = 256;
hipError_t launch_prep_jobs(const PrepJob* jobs, const long long* begin, const int* block_job, long long n_blocks,
                            hipStream_t s, int* zero = nullptr, int nzero = 0);

// One render of a batch after the MFMA launches: its jobs' split-K partials
// (JobPart[job_begin, job_end), summed in FP64 in a fixed order) give F = scale * (re, -im);
// then the C2R through work into image.
struct JobPart {
  const float2* partial;  // [nchunks][2cx][cy]
  int nchunks;
  int pad;
};
struct RenderOut {
  float2* F;
  float2* work;
  float* image;
  float scale;
  int job_begin, job_end;
  int pad;
};
hipError_t launch_batch_finish(const RenderOut* r, const JobPart* jp, int n_renders, int cx, int cy, hipStream_t s);
// ------------------------------------------------------------ C2R (fft.hip) ---
// One axis of the 2-D C2R: the inverse DFT of length N as Stockham stages over LDS lines of
// length L (L = N, or the power-of-two Bluestein length when N has a prime factor above 13).
constexpr int kFftMaxStages = 16;
struct FftAxis {
  int N, L, nst, pad;
  int radix[kFftMaxStages];
  const float2* tw;     // W^j, then W^(128 j), j < 128 (W = exp(+2 pi i / L))
  const float2* chirp;  // Bluestein: exp(+pi i n^2 / N), n < N; nullptr for a direct transform
  const float2* gb;     // Bluestein: transform of the wrapped conj(chirp), divided by L
};
struct FftPlan {
  FftAxis x{}, y{};
  DevBuf mem;  // the tables
};
// Whether a line length is supported (1..16384; at most 8192 with a prime factor above 13).
bool fft_length_supported(int n, std::string* why);
// Tables of an Nx x Ny C2R (synchronous upload); false (error recorded) on failure.
bool fft_plan_build(int nx, int ny, FftPlan* plan);
// Batched C2R of n_renders spectra of one grid (cx, cy, Nx, Ny).
hipError_t launch_c2r_batch(const RenderOut* r, int n_renders, int cx, int cy, int Nx, int Ny, const FftPlan& plan,
                            hipStream_t s);

// direct_sum.hip: plain FP32 direct sum over [first, first+count) packets.
//   phi_kind: 0 = uniform radius (Phi hoisted per frequency), 1 = per-particle radius.
//   partial: [ksplit][2cx*cy] double2, slot k holds the sum of split k (unscaled, +i sin).
hipError_t launch_direct_sum(const Packet* packets, int n_padded, int ksplit, int uniform, float radius,
                             const ViewParams& vp, double2* partial, hipStream_t s);
// Sum the ksplit partials (in order) into acc (assign when first), FP64.
hipError_t launch_reduce_partials(const double2* partial, int ksplit, int nfreq, int first, double2* acc,
                                  hipStream_t s);
// F[i] = (scale*re, -scale*im) as float2 (slot 0 of _d_freq_image).
hipError_t launch_finalize_spectrum(const double2* acc, int nfreq, float scale, float2* F, hipStream_t s);

// Zero `words` 32-bit words at p with a kernel (no memset node when a caller captures the stream).
hipError_t launch_zero_words(void* p, size_t words, hipStream_t s);

// fft.hip: 2-D C2R of the half-plane spectrum F[cy][2cx] into image[Nx][Ny]; work is the
// Nx*(Ny/2+1) complex buffer (_d_freq_image_arranged), used as the intermediate [cy][Nx].
hipError_t launch_c2r(const float2* F, int cx, int cy, int Nx, int Ny, const FftPlan& plan, float2* work,
                      float* image, hipStream_t s);
// The finalize and the C2R of a single render in the two C2R launches: the first pass writes the
// spectrum F = scale * (re, -im) of the FP64 sums acc[cy][2cx] as it gathers it (the same values as
// launch_finalize_spectrum) and transforms it.
hipError_t launch_finalize_c2r(const double2* acc, float scale, float2* F, int cx, int cy, int Nx, int Ny,
                               const FftPlan& plan, float2* work, float* image, hipStream_t s);

}  // namespace mv
