// Device code of the 2-D C2R (the design is described in fft.h): line transforms by Stockham
// stages in LDS, the pass kernels and their gathers and scatters.  Compiled once per kernel
// variant: MV_FFT_PASS (0: x pass, 1: y pass), MV_FFT_ODD (0/1) and MV_FFT_MAXT (512/1024) name
// the one instantiated here, so the variants build in parallel.
#include "fft.h"

namespace mv {

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 conjf2(float2 a) { return make_float2(a.x, -a.y); }

// Twiddle factors W^e = exp(+2 pi i e / L), e < L, from two tables held in LDS:
// tws[e mod 128] * tws[128 + e / 128] (the tables hold W^j and W^(128 j), j < 128, each
// computed in FP64 on the host and rounded to FP32).
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
      else dft8(v[j]);
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


// Pass 1: spectrum rows y0 .. y0 + lines - 1 -> work rows.  Value n of line l is the
// arranged spectrum's row n: F[y][n] (n < cx), F[y][n - Nx + 2cx] (n > Nx - cx), else 0.
// With acc (fused finalize), F[y][x0] = scale * (re, -im) of acc[y][x0] in the same FP64
// arithmetic as finalize_spectrum_kernel, stored to F_out; the first stage gathers every value
// exactly once, and the load of n = cx also stores the dropped column x0 = cx (2cx <= Nx, so n = cx
// is never a gathered column).
struct XGather {
  static constexpr bool kLds = false;
  const float2* F;
  const float2* chirp;
  int y0, nl, cx, Nx;
  const double2* acc;
  float2* F_out;
  double scale;
  __device__ __forceinline__ float2 final_value(size_t i) const {
    const double2 a = acc[i];
    const float2 v = make_float2(static_cast<float>(a.x * scale), static_cast<float>(-a.y * scale));
    F_out[i] = v;
    return v;
  }
  __device__ __forceinline__ float2 load(int l, int n) const {
    if (l >= nl || n >= Nx) return make_float2(0.0f, 0.0f);
    const size_t row = static_cast<size_t>(y0 + l) * (2 * cx);
    if (acc) {
      if (n < cx) return load_value(chirp, n, final_value(row + n));
      if (n > Nx - cx) return load_value(chirp, n, final_value(row + n - Nx + 2 * cx));
      if (n == cx && cx > 0) (void)final_value(row + cx);
      return make_float2(0.0f, 0.0f);
    }
    if (n < cx) return load_value(chirp, n, F[row + n]);
    if (n > Nx - cx) return load_value(chirp, n, F[row + n - Nx + 2 * cx]);
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
  const XGather in{F, ax.chirp, y0, nl, a.cx, a.Nx, batch ? nullptr : a.acc, a.F_out, a.scale};
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

#if defined(MV_FFT_PASS) && defined(MV_FFT_ODD) && defined(MV_FFT_MAXT)
#if MV_FFT_PASS == 0
template __global__ void c2r_x_kernel<MV_FFT_ODD != 0, MV_FFT_MAXT>(C2rArgs, FftAxis, const RenderOut*);
#else
template __global__ void c2r_y_kernel<MV_FFT_ODD != 0, MV_FFT_MAXT>(C2rArgs, FftAxis, const RenderOut*);
#endif
#endif

}  // namespace mv
