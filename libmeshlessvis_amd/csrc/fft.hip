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
#include "internal.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <mutex>
#include <vector>

namespace mv {

constexpr int kFftEmax = 16;           // values of a line a thread holds per stage
constexpr int kFftThreads = 1024;      // largest block (lines of more than 8192 values)
constexpr int kFftThreadsSmall = 512;  // largest block otherwise (twice the registers per thread)
constexpr int kFftMaxLds = 16384;      // complex values of LDS per workgroup
constexpr int kFftTargetElems = 4096;  // LDS values per workgroup when lines are plentiful
constexpr int kXcds = 8;

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 conjf2(float2 a) { return make_float2(a.x, -a.y); }

// Twiddle factors W^e = exp(+2 pi i e / L), e < L, from two tables held in LDS:
// tws[e mod 128] * tws[128 + e / 128] (the tables hold W^j and W^(128 j), j < 128, each
// computed in FP64 on the host and rounded to FP32).
constexpr int kTwiddleWords = 256;
__device__ __forceinline__ float2 twiddle(const float2* tws, int e) {
  return cmul(tws[e & 127], tws[128 + (e >> 7)]);
}

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 times_i(float2 a) { return make_float2(-a.y, a.x); }

// In-register inverse DFTs y_k = sum_j a_j exp(+2 pi i jk / R).
__device__ __forceinline__ void dft2(float2* a) {
  const float2 t = a[0];
  a[0] = cadd(t, a[1]);
  a[1] = csub(t, a[1]);
}
__device__ __forceinline__ void dft4(float2& a0, float2& a1, float2& a2, float2& a3) {
  const float2 t0 = cadd(a0, a2), t1 = csub(a0, a2), t2 = cadd(a1, a3), t3 = times_i(csub(a1, a3));
  a0 = cadd(t0, t2);
  a2 = csub(t0, t2);
  a1 = cadd(t1, t3);
  a3 = csub(t1, t3);
}
// Radix 2 x (R/2): the even and odd halves' DFTs joined by the roots exp(+2 pi i k / R).
__device__ __forceinline__ void dft8(float2* a) {
  float2 e[4] = {a[0], a[2], a[4], a[6]}, o[4] = {a[1], a[3], a[5], a[7]};
  dft4(e[0], e[1], e[2], e[3]);
  dft4(o[0], o[1], o[2], o[3]);
  constexpr float c = 0.70710678118654752440f;
  const float2 w[4] = {o[0], make_float2((o[1].x - o[1].y) * c, (o[1].x + o[1].y) * c), times_i(o[2]),
                       make_float2(-(o[3].x + o[3].y) * c, (o[3].x - o[3].y) * c)};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    a[k] = cadd(e[k], w[k]);
    a[k + 4] = csub(e[k], w[k]);
  }
}
__device__ __forceinline__ void dft16(float2* a) {
  float2 e[8], o[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    e[j] = a[2 * j];
    o[j] = a[2 * j + 1];
  }
  dft8(e);
  dft8(o);
  constexpr float c1 = 0.92387953251128675613f, s1 = 0.38268343236508977173f, c2 = 0.70710678118654752440f;
  const float2 r[8] = {make_float2(1.0f, 0.0f), make_float2(c1, s1),  make_float2(c2, c2),  make_float2(s1, c1),
                       make_float2(0.0f, 1.0f), make_float2(-s1, c1), make_float2(-c2, c2), make_float2(-c1, s1)};
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const float2 w = k ? cmul(o[k], r[k]) : o[k];
    a[k] = cadd(e[k], w);
    a[k + 8] = csub(e[k], w);
  }
}
// Odd R: output k of the direct sum, roots from the twiddle table (w_R^t = W^{t L / R}).
template <int R>
__device__ __forceinline__ float2 dft_odd_out(const float2* a, const float2* tws, int per_line, int k) {
  float2 y = a[0];
#pragma unroll
  for (int i = 1; i < R; ++i) {
    const float2 c = twiddle(tws, ((i * k) % R) * per_line);
    y.x = fmaf(a[i].x, c.x, fmaf(-a[i].y, c.y, y.x));
    y.y = fmaf(a[i].x, c.y, fmaf(a[i].y, c.x, y.y));
  }
  return y;
}

// The twiddle tables (kTwiddleWords values; blocks have at least 64 threads) are fetched
// from global memory into registers at the start of a kernel and stored to LDS by the first
// stage after it has issued its own loads, so both wait out one memory latency together.
struct Twiddles {
  float2 v[kTwiddleWords / 64];
};
__device__ __forceinline__ Twiddles fetch_twiddles(const FftAxis& ax) {
  Twiddles t;
#pragma unroll
  for (int k = 0; k < kTwiddleWords / 64; ++k) {
    const int j = threadIdx.x + k * blockDim.x;
    t.v[k] = j < kTwiddleWords ? ax.tw[j] : make_float2(0.0f, 0.0f);
  }
  return t;
}
__device__ __forceinline__ void store_twiddles(float2* tws, const Twiddles& t) {
#pragma unroll
  for (int k = 0; k < kTwiddleWords / 64; ++k) {
    const int j = threadIdx.x + k * blockDim.x;
    if (j < kTwiddleWords) tws[j] = t.v[k];
  }
}

// Lines held in LDS (line l at lds + l * L): the reads and writes of the inner stages.
// (One pad word per 16 values against bank conflicts of the strided writes measured 5%
// slower on 256- and 1024-value lines: the stages are not LDS-bandwidth bound.)
__host__ __device__ constexpr int lds_pitch(int L) { return L; }
struct LdsLines {
  static constexpr bool kLds = true;
  float2* lds;
  int pitch;
  __device__ __forceinline__ float2 load(int l, int n) const { return lds[l * pitch + n]; }
  __device__ __forceinline__ void store(int l, int n, float2 v) const { lds[l * pitch + n] = v; }
};

// One radix-R Stockham stage over `lines` lines of length L, s = product of the radices
// before it:
//   a_j = in[q + s (p + j m)],  y_k = sum_j a_j w_R^{jk},  out[q + s (R p + k)] = y_k W^{k p s}
// with m = L / (s R), p < m, q < s, w_R = exp(+2 pi i / R), W = exp(+2 pi i / L).  The
// inputs come from `in` and the outputs go to `out`: LDS lines, or for the first and last
// stages of a transform, the kernel's global gather and scatter.  In place in LDS: a thread's
// butterflies are read into registers before the barrier, written after it.  twp: the twiddle
// tables (global in the first stage, while their LDS copy tv is being stored, LDS after).
// Every thread of the block calls it.
template <int R, bool POW2, class In, class Out>
__device__ __forceinline__ void fft_stage(const float2* twp, int lines, int L, int s, const In& in, const Out& out,
                                          const Twiddles* tv, float2* tws) {
  // POW2: L is a power of two (so are L / R and s): index splits by shifts, not divisions.
  const int lpl = POW2 ? __ffs(L / R) - 1 : 0, ls = POW2 ? __ffs(s) - 1 : 0;
  constexpr int J = (kFftEmax + R - 1) / R;
  const int nthr = blockDim.x, tid = threadIdx.x;
  const int per_line = L / R;
  const int m = per_line / s;
  const int T = lines * per_line;
  float2 v[J][R];
  int lj[J], oj[J], pj[J];
#pragma unroll
  for (int j = 0; j < J; ++j) {
    const int b = tid + j * nthr;
    lj[j] = -1;
    oj[j] = pj[j] = 0;
    if (b < T) {
      const int l = POW2 ? b >> lpl : b / per_line, bb = b - l * per_line;
      const int p = POW2 ? bb >> ls : bb / s, q = bb - p * s;
#pragma unroll
      for (int i = 0; i < R; ++i) v[j][i] = in.load(l, q + s * (p + i * m));
      lj[j] = l;
      oj[j] = q + s * R * p;  // output of k = 0
      pj[j] = p * s;          // twiddle of output k: W^{k p s}
    }
  }
  if (tv) store_twiddles(tws, *tv);
  if (In::kLds && Out::kLds) __syncthreads();
#pragma unroll
  for (int j = 0; j < J; ++j) {
    if (lj[j] < 0) break;  // butterflies b = tid + j nthr: the first one past T ends the list
    if (R & 1) {
#pragma unroll
      for (int k = 0; k < R; ++k) {
        const float2 y = dft_odd_out<R>(v[j], twp, per_line, k);
        out.store(lj[j], oj[j] + s * k, (k && pj[j]) ? cmul(y, twiddle(twp, k * pj[j])) : y);
      }
    } else {
      if (R == 2) dft2(v[j]);
      else if (R == 4) dft4(v[j][0], v[j][1], v[j][2], v[j][3]);
      else if (R == 8) dft8(v[j]);
      else dft16(v[j]);
      out.store(lj[j], oj[j], v[j][0]);
#pragma unroll
      for (int k = 1; k < R; ++k)
        out.store(lj[j], oj[j] + s * k, pj[j] ? cmul(v[j][k], twiddle(twp, k * pj[j])) : v[j][k]);
    }
  }
  if (Out::kLds) __syncthreads();
}

template <bool ODD, class In, class Out>
__device__ __forceinline__ void stage_r(int R, const float2* twp, int lines, int L, int s, const In& in,
                                        const Out& out, const Twiddles* tv, float2* tws) {
  switch (R) {
    case 16: fft_stage<16, !ODD>(twp, lines, L, s, in, out, tv, tws); break;
    case 8: fft_stage<8, !ODD>(twp, lines, L, s, in, out, tv, tws); break;
    case 4: fft_stage<4, !ODD>(twp, lines, L, s, in, out, tv, tws); break;
    case 2: fft_stage<2, !ODD>(twp, lines, L, s, in, out, tv, tws); break;
    default:
      if (ODD) {
        switch (R) {
          case 3: fft_stage<3, false>(twp, lines, L, s, in, out, tv, tws); break;
          case 5: fft_stage<5, false>(twp, lines, L, s, in, out, tv, tws); break;
          case 7: fft_stage<7, false>(twp, lines, L, s, in, out, tv, tws); break;
          case 11: fft_stage<11, false>(twp, lines, L, s, in, out, tv, tws); break;
          default: fft_stage<13, false>(twp, lines, L, s, in, out, tv, tws); break;
        }
      }
  }
}

// The inverse (+i) DFT of length ax.L of `lines` lines, read through `in` by the first stage
// and written through `out` by the last, the stages between in LDS.  With tv, the first stage
// stores the twiddle tables' LDS copy (and reads the global ones).  ODD: the line length has
// a factor 3, 5, 7, 11 or 13 (the variant without those stages needs fewer registers).
template <bool ODD, class In, class Out>
__device__ __forceinline__ void transform(const FftAxis& ax, int lines, float2* lds, float2* tws, const In& in, const Out& out,
                          const Twiddles* tv) {
  const LdsLines mid{lds, lds_pitch(ax.L)};
  const int nst = ax.nst;
  const float2* tw0 = tv ? ax.tw : tws;
  if (nst == 0) {  // L = 1
    if (tv) store_twiddles(tws, *tv);
    for (int l = threadIdx.x; l < lines; l += blockDim.x) {
      const float2 x = in.load(l, 0);
      if (In::kLds && Out::kLds) __syncthreads();
      out.store(l, 0, x);
    }
    __syncthreads();
    return;
  }
  if (nst == 1) {
    stage_r<ODD>(ax.radix[0], tw0, lines, ax.L, 1, in, out, tv, tws);
    return;
  }
  stage_r<ODD>(ax.radix[0], tw0, lines, ax.L, 1, in, mid, tv, tws);
  int s = ax.radix[0];
  for (int st = 1; st + 1 < nst; ++st) {
    stage_r<ODD>(ax.radix[st], tws, lines, ax.L, s, mid, mid, nullptr, tws);
    s *= ax.radix[st];
  }
  stage_r<ODD>(ax.radix[nst - 1], tws, lines, ax.L, s, mid, out, nullptr, tws);
}

// The inverse DFT of length ax.N of every line: a direct transform, or Bluestein's: the
// gather multiplies x_n by chirp[n] (zero for n >= N), the first transform ends in LDS, the
// pointwise product by gb and a conjugation, the second transform, and the scatter takes
// chirp[n] * conj(z_n) (finish_value).
template <bool ODD, class In, class Out>
__device__ __forceinline__ void line_transform(const FftAxis& ax, int lines, float2* lds, float2* tws, const In& in, const Out& out,
                               const Twiddles& tv) {
  if (!ax.chirp) {
    transform<ODD>(ax, lines, lds, tws, in, out, &tv);
    return;
  }
  const LdsLines mid{lds, lds_pitch(ax.L)};
  transform<ODD>(ax, lines, lds, tws, in, mid, &tv);
  for (int i = threadIdx.x; i < lines * ax.L; i += blockDim.x) {
    const int l = i / ax.L, k = i - l * ax.L;
    mid.store(l, k, conjf2(cmul(mid.load(l, k), ax.gb[k])));
  }
  __syncthreads();
  transform<ODD>(ax, lines, lds, tws, mid, out, nullptr);
}
__device__ __forceinline__ float2 load_value(const float2* chirp, int n, float2 x) {
  return chirp ? cmul(x, chirp[n]) : x;
}
__device__ __forceinline__ float2 finish_value(const float2* chirp, int n, float2 z) {
  return chirp ? cmul(chirp[n], conjf2(z)) : z;
}

// Blocks of a launch in the order the hardware spreads them over the XCDs (block b on XCD
// b mod 8): logical block = (b mod 8) * per + b / 8, so consecutive logical blocks share an
// XCD and its L2.  The grid is rounded up to a multiple of 8; logical blocks past n exit.
__device__ __forceinline__ int xcd_block(int n) {
  const int per = (n + kXcds - 1) / kXcds;
  return static_cast<int>(blockIdx.x % kXcds) * per + static_cast<int>(blockIdx.x / kXcds);
}
static unsigned xcd_grid(int n) { return static_cast<unsigned>((n + kXcds - 1) / kXcds * kXcds); }

struct C2rArgs {
  const float2* F;
  float2* work;
  float* image;
  int cx, cy, Nx, Ny;
  int lines;  // pass 1: spectrum rows per workgroup; pass 2: row pairs per workgroup
  int blocks;  // logical blocks per render
};

// Pass 1: spectrum rows y0 .. y0 + lines - 1 -> work rows.  Value n of line l is the
// arranged spectrum's row n: F[y][n] (n < cx), F[y][n - Nx + 2cx] (n > Nx - cx), else 0.
struct XGather {
  static constexpr bool kLds = false;
  const float2* F;
  const float2* chirp;
  int y0, nl, cx, Nx;
  __device__ __forceinline__ float2 load(int l, int n) const {
    if (l >= nl || n >= Nx) return make_float2(0.0f, 0.0f);
    const float2* row = F + static_cast<size_t>(y0 + l) * (2 * cx);
    if (n < cx) return load_value(chirp, n, row[n]);
    if (n > Nx - cx) return load_value(chirp, n, row[n - Nx + 2 * cx]);
    return make_float2(0.0f, 0.0f);
  }
};
struct XScatter {
  static constexpr bool kLds = false;
  float2* work;
  const float2* chirp;
  int y0, nl, Nx;
  __device__ __forceinline__ void store(int l, int n, float2 v) const {
    if (l < nl && n < Nx) work[static_cast<size_t>(y0 + l) * Nx + n] = finish_value(chirp, n, v);
  }
};

template <bool ODD, int MAXT>
__global__ __launch_bounds__(MAXT) void c2r_x_kernel(C2rArgs a, FftAxis ax, const RenderOut* __restrict__ batch) {
  extern __shared__ float2 smem[];
  float2* tws = smem;
  float2* lds = smem + kTwiddleWords;
  const int blk = xcd_block(a.blocks);
  if (blk >= a.blocks) return;
  const float2* F = a.F;
  float2* work = a.work;
  if (batch) {
    const RenderOut ro = batch[blockIdx.y];
    F = ro.F;
    work = ro.work;
  }
  const int y0 = blk * a.lines;
  const int nl = min(a.lines, a.cy - y0);
  const Twiddles tv = fetch_twiddles(ax);
  const XGather in{F, ax.chirp, y0, nl, a.cx, a.Nx};
  const XScatter out{work, ax.chirp, y0, nl, a.Nx};
  line_transform<ODD>(ax, a.lines, lds, tws, in, out, tv);
}

// Pass 2: image rows r0 .. r0 + 2 lines - 1 from work, two rows per transform.  Value n of
// pair p is the Hermitian completion of rows r = r0 + 2p, r + 1 (v1, v2 = work[y][r], [r + 1]):
//   Z[n] = X1[n] + i X2[n] = (v1.x - v2.y, v1.y + v2.x)             (0 < n < cy, y = n),
//   Z[n] = conj X1[y] + i conj X2[y] = (v1.x + v2.y, v2.x - v1.y)  (Ny - cy < n < Ny, y = Ny - n),
//   Z[0] = (Re v1, Re v2) (the imaginary parts of the y = 0 terms are ignored), else 0.
// The transform of a Hermitian line is real: Re z -> image row r, Im z -> row r + 1.
struct YGather {
  static constexpr bool kLds = false;
  const float2* work;
  const float2* chirp;
  int r0, nr, np, cy, Nx, Ny;
  __device__ __forceinline__ float2 load(int p, int n) const {
    if (p >= np || n >= Ny) return make_float2(0.0f, 0.0f);
    const bool lo = n < cy, hi = n > Ny - cy && n > 0;
    if (!lo && !hi) return make_float2(0.0f, 0.0f);
    const int y = lo ? n : Ny - n;
    const float2* src = work + static_cast<size_t>(y) * Nx + r0 + 2 * p;
    const float2 v1 = src[0];
    const float2 v2 = (2 * p + 1 < nr) ? src[1] : make_float2(0.0f, 0.0f);
    float2 z;
    if (n == 0)
      z = make_float2(v1.x, v2.x);
    else if (lo)
      z = make_float2(v1.x - v2.y, v1.y + v2.x);
    else
      z = make_float2(v1.x + v2.y, v2.x - v1.y);
    return load_value(chirp, n, z);
  }
};
struct YScatter {
  static constexpr bool kLds = false;
  float* image;
  const float2* chirp;
  int r0, nr, np, Ny;
  __device__ __forceinline__ void store(int p, int n, float2 v) const {
    if (p >= np || n >= Ny) return;
    const float2 z = finish_value(chirp, n, v);
    const int r = r0 + 2 * p;
    image[static_cast<size_t>(r) * Ny + n] = z.x;
    if (2 * p + 1 < nr) image[static_cast<size_t>(r + 1) * Ny + n] = z.y;
  }
};

template <bool ODD, int MAXT>
__global__ __launch_bounds__(MAXT) void c2r_y_kernel(C2rArgs a, FftAxis ay, const RenderOut* __restrict__ batch) {
  extern __shared__ float2 smem[];
  float2* tws = smem;
  float2* lds = smem + kTwiddleWords;
  const int blk = xcd_block(a.blocks);
  if (blk >= a.blocks) return;
  const float2* work = a.work;
  float* image = a.image;
  if (batch) {
    const RenderOut ro = batch[blockIdx.y];
    work = ro.work;
    image = ro.image;
  }
  const int r0 = 2 * a.lines * blk;
  const int nr = min(2 * a.lines, a.Nx - r0);  // image rows of this block
  const int np = (nr + 1) / 2;                 // row pairs with data
  const Twiddles tv = fetch_twiddles(ay);
  const YGather in{work, ay.chirp, r0, nr, np, a.cy, a.Nx, a.Ny};
  const YScatter out{image, ay.chirp, r0, nr, np, a.Ny};
  line_transform<ODD>(ay, a.lines, lds, tws, in, out, tv);
}

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
                                int cx, int cy, int Nx, int Ny, const FftPlan& plan, hipStream_t s) {
  if (plan.x.N != Nx || plan.y.N != Ny) return hipErrorInvalidValue;
  if (renders <= 0 || Nx <= 0 || Ny <= 0) return hipSuccess;
  hipError_t e = set_lds_limits();
  if (e != hipSuccess) return e;
  C2rArgs a{F, work, image, cx, cy, Nx, Ny, 0, 0};
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

hipError_t launch_c2r_batch(const RenderOut* r, int n_renders, int cx, int cy, int Nx, int Ny, const FftPlan& plan,
                            hipStream_t s) {
  return launch_passes(nullptr, nullptr, nullptr, r, n_renders, cx, cy, Nx, Ny, plan, s);
}

}  // namespace mv
