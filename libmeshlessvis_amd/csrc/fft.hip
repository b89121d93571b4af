// Host side of the 2-D C2R: line factorisation, FP64 twiddle / Bluestein tables, launch shapes
// and the launches of the two pass kernels (fft_kernels.hip).
#include "fft.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <mutex>
#include <vector>

namespace mv {

static unsigned xcd_grid(int n) { return static_cast<unsigned>((n + kXcds - 1) / kXcds * kXcds); }

// ------------------------------------------------------------------ host ---

static int largest_prime_factor(int n) {
  int lpf = 1;
  for (int f = 2; static_cast<long long>(f) * f <= n; ++f)
    while (n % f == 0) {
      lpf = f;
      n /= f;
    }
  return n > 1 ? std::max(lpf, n) : lpf;
}
static int pow2_at_least(int n) {
  int m = 1;
  while (m < n) m <<= 1;
  return m;
}

bool fft_length_supported(int n, std::string* why) {
  if (n < 1 || n > kFftMaxLds) {
    if (why) *why = "number_of_samples must be 1.." + std::to_string(kFftMaxLds);
    return false;
  }
  if (largest_prime_factor(n) > 13 && pow2_at_least(2 * n - 1) > kFftMaxLds) {
    if (why) *why = "number_of_samples above 8192 with a prime factor above 13";
    return false;
  }
  return true;
}

// Radices of a line of length L: the power of two in radix-4 stages (one radix-8 or -2
// stage for an odd power), then 3, 5, 7, 11, 13.  Measured on 256- and 1024-value lines
// (single renders at 256^2 and 1024^2): radix 4 12.6 and 20.7 us per C2R, radix 8 13.9 and
// 21.4, radix 16 17.6 and 24.1 (the larger radices leave most threads idle per stage and
// write LDS with larger strides).
static bool factor(int L, FftAxis* ax) {
  ax->nst = 0;
  int n = L, p2 = 0;
  while (n % 2 == 0) {
    n /= 2;
    ++p2;
  }
  auto push = [&](int r) {
    if (ax->nst == kFftMaxStages) return false;
    ax->radix[ax->nst++] = r;
    return true;
  };
  if (p2 & 1) {
    if (!push(p2 >= 3 ? 8 : 2)) return false;
    p2 -= p2 >= 3 ? 3 : 1;
  }
  for (; p2 >= 2; p2 -= 2)
    if (!push(4)) return false;
  for (int r : {3, 5, 7, 11, 13})
    while (n % r == 0) {
      if (!push(r)) return false;
      n /= r;
    }
  return n == 1;
}

// exp(+2 pi i k / n) in FP64 (k reduced mod n first).
static std::complex<double> root(long long k, long long n) {
  const double a = 2.0 * M_PI * static_cast<double>(k % n) / static_cast<double>(n);
  return {std::cos(a), std::sin(a)};
}

// In-place FP64 transform sum_j x_j exp(+2 pi i jk / n), n a power of two (host, plan time).
static void host_fft_pow2(std::vector<std::complex<double>>& x) {
  const size_t n = x.size();
  for (size_t i = 1, j = 0; i < n; ++i) {
    size_t bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) std::swap(x[i], x[j]);
  }
  for (size_t len = 2; len <= n; len <<= 1) {
    for (size_t i = 0; i < n; i += len)
      for (size_t k = 0; k < len / 2; ++k) {
        const std::complex<double> w = root(static_cast<long long>(k), static_cast<long long>(len));
        const std::complex<double> u = x[i + k], v = x[i + k + len / 2] * w;
        x[i + k] = u + v;
        x[i + k + len / 2] = u - v;
      }
  }
}

static float2 f2(std::complex<double> z) { return make_float2(static_cast<float>(z.real()), static_cast<float>(z.imag())); }

// Host tables of one axis; device pointers are filled in by fft_plan_build.
static bool axis_tables(int N, FftAxis* ax, std::vector<float2>* tables, size_t* tw_off, size_t* chirp_off,
                        size_t* gb_off) {
  std::string why;
  if (!fft_length_supported(N, &why)) {
    set_error(VIS_ERROR_INVALID_CONFIG, why);
    return false;
  }
  ax->N = N;
  const bool blue = largest_prime_factor(N) > 13;
  ax->L = blue ? pow2_at_least(2 * N - 1) : N;
  if (!factor(ax->L, ax)) {
    set_error(VIS_ERROR_INVALID_CONFIG, "unsupported FFT length " + std::to_string(N));
    return false;
  }
  auto pad = [&]() { tables->resize((tables->size() + 31) / 32 * 32); };  // 256-byte aligned sections
  pad();
  *tw_off = tables->size();
  // W^j and W^(128 j), j < 128 (W = exp(+2 pi i / L), L <= 16384 = 128 * 128).
  for (int k = 0; k < 128; ++k) tables->push_back(f2(root(k, ax->L)));
  for (int k = 0; k < 128; ++k) tables->push_back(f2(root(128LL * k, ax->L)));
  *chirp_off = *gb_off = static_cast<size_t>(-1);
  if (blue) {
    const int M = ax->L;
    // c_n = exp(+pi i n^2 / N) = exp(+2 pi i (n^2 mod 2N) / 2N).
    std::vector<std::complex<double>> c(N), b(M, 0.0);
    for (int n = 0; n < N; ++n) c[n] = root(static_cast<long long>(n) * n % (2LL * N), 2LL * N);
    for (int m = 0; m < N; ++m) {
      b[m] = std::conj(c[m]);
      if (m) b[M - m] = std::conj(c[m]);
    }
    host_fft_pow2(b);
    pad();
    *chirp_off = tables->size();
    for (int n = 0; n < N; ++n) tables->push_back(f2(c[n]));
    pad();
    *gb_off = tables->size();
    for (int k = 0; k < M; ++k) tables->push_back(f2(b[k] / static_cast<double>(M)));
  }
  return true;
}

bool fft_plan_build(int nx, int ny, FftPlan* p) {
  std::vector<float2> t;
  size_t ox[3], oy[3];
  FftAxis x{}, y{};
  if (!axis_tables(nx, &x, &t, &ox[0], &ox[1], &ox[2]) || !axis_tables(ny, &y, &t, &oy[0], &oy[1], &oy[2]))
    return false;
  if (!p->mem.reserve(sizeof(float2) * t.size())) return false;
  if (!hip_ok(hipMemcpy(p->mem.ptr, t.data(), sizeof(float2) * t.size(), hipMemcpyHostToDevice), "upload FFT tables"))
    return false;
  const float2* base = p->mem.as<float2>();
  auto at = [&](size_t o) { return o == static_cast<size_t>(-1) ? nullptr : base + o; };
  x.tw = at(ox[0]);
  x.chirp = at(ox[1]);
  x.gb = at(ox[2]);
  y.tw = at(oy[0]);
  y.chirp = at(oy[1]);
  y.gb = at(oy[2]);
  p->x = x;
  p->y = y;
  return true;
}

// Launch shape of a pass: `lines` lines of length L per block out of `total` per render.
struct PassShape {
  int lines, blocks, threads;
  size_t lds;
};
static PassShape pass_shape(int total, int L, int renders) {
  PassShape sh{};
  // About 512 blocks over the launch when lines are few; otherwise kFftTargetElems values of
  // LDS per block, so that several blocks share a CU.
  const long long all = static_cast<long long>(total) * renders;
  int lines = static_cast<int>(std::max<long long>(1, (all + 511) / 512));
  lines = std::min(lines, std::max(1, kFftTargetElems / L));
  lines = std::max(1, std::min(lines, total));
  sh.lines = lines;
  sh.blocks = (total + lines - 1) / lines;
  const int elems = lines * L;  // <= kFftMaxLds: L <= kFftMaxLds and lines * L <= max(L, kFftTargetElems)
  // A thread per four values, within the largest block; at least kFftEmax values a thread.
  int th = (elems / 4 + 63) / 64 * 64;
  th = std::min(th, elems > kFftEmax * kFftThreadsSmall ? kFftThreads : kFftThreadsSmall);
  th = std::max(th, (elems + kFftEmax - 1) / kFftEmax);
  sh.threads = std::max(64, (th + 63) / 64 * 64);
  sh.lds = sizeof(float2) * static_cast<size_t>(kTwiddleWords + lines * lds_pitch(L));
  return sh;
}

static bool odd_axis(const FftAxis& ax) {
  for (int i = 0; i < ax.nst; ++i)
    if (ax.radix[i] & 1) return true;
  return false;
}

typedef void (*C2rKernel)(C2rArgs, FftAxis, const RenderOut*);
// The kernel variant of a pass: odd radices or not, block limit 512 or 1024 threads.
static C2rKernel pass_kernel(bool y_pass, bool odd, int threads) {
  const bool big = threads > kFftThreadsSmall;
  if (y_pass)
    return odd ? (big ? c2r_y_kernel<true, kFftThreads> : c2r_y_kernel<true, kFftThreadsSmall>)
               : (big ? c2r_y_kernel<false, kFftThreads> : c2r_y_kernel<false, kFftThreadsSmall>);
  return odd ? (big ? c2r_x_kernel<true, kFftThreads> : c2r_x_kernel<true, kFftThreadsSmall>)
             : (big ? c2r_x_kernel<false, kFftThreads> : c2r_x_kernel<false, kFftThreadsSmall>);
}

// Dynamic LDS above the default limit for every kernel variant (once per process).
static hipError_t set_lds_limits() {
  static std::once_flag once;
  static hipError_t err = hipSuccess;
  std::call_once(once, [] {
    const int bytes = static_cast<int>(sizeof(float2) * (kTwiddleWords + lds_pitch(kFftMaxLds)));
    for (int v = 0; v < 8; ++v) {
      const C2rKernel k = pass_kernel(v & 1, v & 2, (v & 4) ? kFftThreads : 64);
      const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(k),
                                               hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
      if (e != hipSuccess && err == hipSuccess) err = e;
    }
  });
  return err;
}

static hipError_t launch_passes(const float2* F, float2* work, float* image, const RenderOut* batch, int renders,
                                int cx, int cy, int Nx, int Ny, const FftPlan& plan, hipStream_t s,
                                const double2* acc = nullptr, float scale = 0.0f, float2* F_out = nullptr) {
  if (plan.x.N != Nx || plan.y.N != Ny) return hipErrorInvalidValue;
  if (renders <= 0 || Nx <= 0 || Ny <= 0) return hipSuccess;
  hipError_t e = set_lds_limits();
  if (e != hipSuccess) return e;
  C2rArgs a{F, work, image, cx, cy, Nx, Ny, 0, 0, acc, F_out, static_cast<double>(scale)};
  if (cy > 0) {
    const PassShape p1 = pass_shape(cy, plan.x.L, renders);
    a.lines = p1.lines;
    a.blocks = p1.blocks;
    hipLaunchKernelGGL(pass_kernel(false, odd_axis(plan.x), p1.threads), dim3(xcd_grid(p1.blocks), renders),
                       dim3(p1.threads), p1.lds, s, a, plan.x, batch);
  }
  const PassShape p2 = pass_shape((Nx + 1) / 2, plan.y.L, renders);
  a.lines = p2.lines;
  a.blocks = p2.blocks;
  hipLaunchKernelGGL(pass_kernel(true, odd_axis(plan.y), p2.threads), dim3(xcd_grid(p2.blocks), renders),
                     dim3(p2.threads), p2.lds, s, a, plan.y, batch);
  return hipGetLastError();
}

hipError_t launch_c2r(const float2* F, int cx, int cy, int Nx, int Ny, const FftPlan& plan, float2* work,
                      float* image, hipStream_t s) {
  return launch_passes(F, work, image, nullptr, 1, cx, cy, Nx, Ny, plan, s);
}

hipError_t launch_finalize_c2r(const double2* acc, float scale, float2* F, int cx, int cy, int Nx, int Ny,
                               const FftPlan& plan, float2* work, float* image, hipStream_t s) {
  return launch_passes(F, work, image, nullptr, 1, cx, cy, Nx, Ny, plan, s, acc, scale, F);
}

hipError_t launch_c2r_batch(const RenderOut* r, int n_renders, int cx, int cy, int Nx, int Ny, const FftPlan& plan,
                            hipStream_t s) {
  return launch_passes(nullptr, nullptr, nullptr, r, n_renders, cx, cy, Nx, Ny, plan, s);
}

}  // namespace mv
