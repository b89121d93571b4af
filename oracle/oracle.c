/*
 * oracle.c -- CPU oracle for the Fourier-projection hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load this (through oracle/__init__.py), and
 * only as the checker or the timed CPU baseline; the product library never
 * links or calls it.
 *
 * Contents (each a clean-room restatement; reference lines cited):
 *   mvo_phi64          Phi in FP64, reference formulas or exact transforms
 *                      (fourier_transform_cpu.cpp:37-66)
 *   mvo_phi32_port     Phi in FP32 exactly as fourier_transform_cpu.cpp:37-66
 *   mvo_spectrum64     FP64 direct sum, all groups, stored as the reference
 *                      stores it (fourier_transform_cpu.cpp:73-134,136-171)
 *   mvo_port_spectrum32  FP32/OpenMP restatement of the same loop nest, same
 *                      loop order, schedule(dynamic, block_length), serial FP32
 *                      accumulation; the timed CPU baseline
 *   mvo_arrange64      arrange_samples (meshless_vis_cpu.cpp:209-234)
 *   mvo_c2r64          unnormalised 2-D C2R (FFTW/cuFFT plan semantics,
 *                      meshless_vis_cpu.cpp:62-64,261): inverse complex DFT
 *                      along axis 0, then Hermitian c2r along axis 1
 *                      (imaginary parts of the DC and Nyquist terms ignored)
 *
 * Pins (tests/test_oracle_*.py): Phi against numerical quadrature of the
 * basis functions' 3-D Fourier integrals and against closed-form special
 * values; the spectrum against single-particle closed forms, translation,
 * linearity and Hermitian symmetry; arrange against its closed-form index
 * map; the C2R against numpy.fft.irfft2(norm="forward") and direct sums.
 * Built with -ffp-contract=off so the FP32 port rounds like the reference's
 * x86-64 build (no FMA contraction).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

#define MVO_SPH 0
#define MVO_GAUSSIAN 1
#define MVO_WENDLAND 2
#define MVO_PHI_REFERENCE 0
#define MVO_PHI_ANALYTIC 1

static const double PI = 3.14159265358979323846264338327950288;

/* ---------------------------------------------------------------- Phi ---- */

/* Exact SPH transform 3*pi*(sin m/m)^3*(sin m - m cos m)/m^3, m = pi*r. */
static double phi_sph_exact(double r) {
  const double m = PI * r;
  const double m2 = m * m;
  double sinc, g;
  if (m < 0.5) {
    /* sin m / m and (sin m - m cos m)/m^3 as Maclaurin series */
    double t = 1.0, s = 0.0;
    for (int n = 0; n < 12; ++n) {
      s += t;
      t *= -m2 / ((2.0 * n + 2.0) * (2.0 * n + 3.0));
    }
    sinc = s;
    /* g = sum_{n>=1} (-1)^(n+1) 2n m^(2n-2) / (2n+1)! */
    double fact = 6.0; /* (2n+1)! at n=1 */
    double pw = 1.0;   /* m^(2n-2) */
    g = 0.0;
    for (int n = 1; n < 14; ++n) {
      g += ((n & 1) ? 1.0 : -1.0) * 2.0 * n * pw / fact;
      pw *= m2;
      fact *= (2.0 * n + 2.0) * (2.0 * n + 3.0);
    }
  } else {
    sinc = sin(m) / m;
    g = (sin(m) - m * cos(m)) / (m2 * m);
  }
  return 3.0 * PI * sinc * sinc * sinc * g;
}

/* Exact Wendland C2 transform 7.5*pi*(4m^2-6+(6-m^2)cos 2m+4.5 m sin 2m)/m^8.
   Below m = 2 it is summed as 7.5*pi*sum_{k>=4} c_k m^(2k-8), where c_k is the
   m^(2k) coefficient of the numerator's Maclaurin series:
   c_k = 6(-1)^k 4^k/(2k)! + (-1)^k 4^(k-1)/(2k-2)! + 4.5(-1)^(k-1) 2^(2k-1)/(2k-1)!  */
#define MVO_WEND_TERMS 24
static double wend_coef[MVO_WEND_TERMS];
static int wend_ready = 0;
static void init_tables(void) {
  if (wend_ready) return;
  for (int j = 0; j < MVO_WEND_TERMS; ++j) {
    const int k = j + 4;
    const double sg = (k & 1) ? -1.0 : 1.0;
    const double f2k = tgamma(2.0 * k + 1.0), f2km2 = tgamma(2.0 * k - 1.0), f2km1 = tgamma(2.0 * k);
    wend_coef[j] = 7.5 * PI * (6.0 * sg * pow(4.0, k) / f2k + sg * pow(4.0, k - 1) / f2km2 - 4.5 * sg * pow(2.0, 2 * k - 1) / f2km1);
  }
  wend_ready = 1;
}

static double phi_wendland_exact(double r) {
  const double m = PI * r;
  const double m2 = m * m;
  if (m < 2.0) {
    double s = 0.0;
    for (int j = MVO_WEND_TERMS - 1; j >= 0; --j) s = s * m2 + wend_coef[j];
    return s;
  }
  const double c = cos(2.0 * m), sn = sin(2.0 * m);
  return 7.5 * PI * (4.0 * m2 - 6.0 + (6.0 - m2) * c + 4.5 * m * sn) / (m2 * m2 * m2 * m2);
}

/* The reference's formulas (fourier_transform_cpu.cpp:37-66) evaluated in FP64. */
static double phi_reference64(int basis, double r) {
  if (basis == MVO_SPH) {
    if (r >= 0.06) {
      const double m = PI * r;
      const double c = cos(2.0 * m), s = sin(2.0 * m);
      return (2.3561944901923448 / (m * m * m * m * m * m)) * (c - 1.0) * (c + m * s - 1.0);
    }
    return PI + r * (-0.007968913156311 + r * (-18.293608272337678));
  }
  if (basis == MVO_GAUSSIAN) return exp(PI * r * r);
  if (r >= 0.24) {
    const double m = PI * r;
    const double c = cos(2.0 * m), s = sin(2.0 * m);
    const double m2 = m * m, m4 = m2 * m2;
    return ((PI * 7.5) / (m4 * m4)) * (4.0 * m2 - 6.0 + (6.0 - m2) * c + 4.5 * m * s);
  }
  return 0.299199300341890 + r * (-0.002379178586124 + r * (-0.370530218163545));
}

double mvo_phi64(int basis, int mode, double r) {
  if (mode == MVO_PHI_REFERENCE) return phi_reference64(basis, r);
  if (basis == MVO_SPH) return phi_sph_exact(r);
  if (basis == MVO_GAUSSIAN) return exp(-PI * r * r);
  return phi_wendland_exact(r);
}

void mvo_phi64_array(int basis, int mode, const double* r, double* out, int n) {
  init_tables();
  for (int i = 0; i < n; ++i) out[i] = mvo_phi64(basis, mode, r[i]);
}

/* FP32 exactly as fourier_transform_cpu.cpp:37-66 (std::cos/sin/exp of float). */
static const float PI_F = 3.141592653589793238462643383279f;
static const float TWO_PI_F = 6.283185307179586476925286766559f;

float mvo_phi32_port(int basis, float r) {
  if (basis == MVO_SPH) {
    if (r >= 0.06f) {
      float m = PI_F * r;
      float m2x = 2.0f * m;
      float c = cosf(m2x), s = sinf(m2x);
      return ((2.3561944901923448f) / (m * m * m * m * m * m)) * (c - 1.0f) * (c + m * s - 1.0f);
    }
    return (PI_F + r * (-0.007968913156311f + r * (-18.293608272337678f)));
  }
  if (basis == MVO_GAUSSIAN) return expf(PI_F * r * r);
  if (r >= 0.24f) {
    float m = PI_F * r;
    float m2x = 2.0f * m;
    float c = cosf(m2x), s = sinf(m2x);
    float m_2 = m * m;
    float m_4 = m_2 * m_2;
    return ((PI_F * 7.5f) / (m_4 * m_4)) * (4.0f * m_2 - 6.0f + (6.0f - m_2) * c + 4.5f * m * s);
  }
  return 0.299199300341890f + r * (-0.002379178586124f + r * (-0.370530218163545f));
}

void mvo_phi32_port_array(int basis, const float* r, float* out, int n) {
  for (int i = 0; i < n; ++i) out[i] = mvo_phi32_port(basis, r[i]);
}

/* ------------------------------------------------------------ views ---- */

/* Index -> (x, y) of the half-plane grid (fourier_transform_cpu.cpp:269-271). */
static void index_to_xy(int index, int cx, int cy, int* x, int* y) {
  int xx = index % (2 * cx);
  if (xx > cx) xx -= 2 * cx;
  *x = xx;
  *y = (index % (2 * cx * cy)) / (2 * cx);
}

/*
 * FP64 oracle of the stored spectrum: out[2*j], out[2*j+1] =
 *   s * sum_i w_i Phi(r h_i) cos(theta_i), -s * sum_i w_i Phi(r h_i) sin(theta_i)
 * at the frequency index idx[j] (all 2*cx*cy indices when idx == NULL),
 * theta_i = 2*pi*dot(k, p_i), k = fu*u + fv*v, fu = sx*x, fv = sy*y,
 * r = sqrt(fu^2 + fv^2), s = sx*sy; everything from the FP32 inputs in FP64.
 * Each frequency's sum runs in blocks of 1024 terms added pairwise.
 */
void mvo_spectrum64(int ngroups, const int* counts, const int* bases, const float* const* cons,
                    const float* const* radii, float sx, float sy, const float* u, const float* v, int cx, int cy,
                    int mode, const int* idx, int nidx, double* out, int nthreads) {
  const int total = 2 * cx * cy;
  const int n_out = idx ? nidx : total;
  const double s = (double)sx * (double)sy;
  init_tables();
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 16)
  for (int j = 0; j < n_out; ++j) {
    const int index = idx ? idx[j] : j;
    int x, y;
    index_to_xy(index, cx, cy, &x, &y);
    const double fu = (double)sx * x, fv = (double)sy * y;
    const double kx = fu * u[0] + fv * v[0], ky = fu * u[1] + fv * v[1], kz = fu * u[2] + fv * v[2];
    const double r = sqrt(fu * fu + fv * fv);
    double re = 0.0, im = 0.0;
    for (int g = 0; g < ngroups; ++g) {
      const float* c = cons[g];
      const float* h = radii[g];
      const int basis = bases[g];
      const double phi_const = h ? 0.0 : mvo_phi64(basis, mode, r);
      for (int b0 = 0; b0 < counts[g]; b0 += 1024) {
        const int b1 = counts[g] < b0 + 1024 ? counts[g] : b0 + 1024;
        double bre = 0.0, bim = 0.0;
        for (int i = b0; i < b1; ++i) {
          const double t = kx * c[4 * i] + ky * c[4 * i + 1] + kz * c[4 * i + 2];
          const double frac = t - floor(t);
          const double th = 2.0 * PI * frac;
          const double phi = h ? mvo_phi64(basis, mode, r * (double)h[i]) : phi_const;
          const double a = (double)c[4 * i + 3] * phi;
          bre += a * cos(th);
          bim += a * sin(th);
        }
        re += bre;
        im += bim;
      }
    }
    out[2 * j] = s * re;
    out[2 * j + 1] = -s * im;
  }
}

/* Phi of the port: the reference's FP32 formulas (mode MVO_PHI_REFERENCE), or the exact
   transform evaluated in FP64 and rounded to FP32 (MVO_PHI_ANALYTIC: the reference's loop,
   phase and FP32 sums with the Phi of configurations whose reference formulas overflow). */
static float port_phi(int basis, int mode, float r) {
  return mode == MVO_PHI_REFERENCE ? mvo_phi32_port(basis, r) : (float)mvo_phi64(basis, MVO_PHI_ANALYTIC, (double)r);
}

/*
 * FP32 restatement of fourier_transform_cpu.cpp:73-134 for all groups
 * (first group assigns, later groups accumulate, :305-312).  Only rows
 * y % row_stride == 0 are computed when row_stride > 1 (timing samples); the
 * other rows are left untouched.  With idx (nidx entries), only the flat
 * frequency indices idx[j] are computed, each into out[2j], out[2j+1] (the same
 * per-frequency loop, for parity checks at sizes where whole rows cost too much).
 * `padded` particles per group are summed (the reference loops over
 * d_number_of_terms, zero padding included); entries beyond counts[g] are
 * treated as zero particles.
 */
void mvo_port_spectrum32(int ngroups, const int* counts, const int* padded, const int* bases,
                         const float* const* cons, const float* const* radii, float sx, float sy, const float* u,
                         const float* v, int cx, int cy, float* out, int block_length, int row_stride,
                         int nthreads, int mode, const int* idx, int nidx) {
  const int image_size = 2 * cx * cy;
  const int n_out = idx ? nidx : image_size;
  const float scale = sx * sy;
  if (nthreads > 0) omp_set_num_threads(nthreads);
  if (block_length < 1) block_length = 1;
  if (row_stride < 1) row_stride = 1;
  init_tables();
  for (int g = 0; g < ngroups; ++g) {
    const float* c = cons[g];
    const float* h = radii[g];
    const int basis = bases[g];
    const int n = counts[g];
    const int np = padded[g] > n ? padded[g] : n;
#pragma omp parallel for schedule(dynamic, block_length)
    for (int j = 0; j < n_out; j++) {
      const int index = idx ? idx[j] : j;
      int x = index % (2 * cx);
      if (x > cx) x = x - (2 * cx);
      const int y = (index % image_size) / (2 * cx);
      if (!idx && y % row_stride) continue;
      const float fu = sx * x, fv = sy * y;
      const float fx = fu * u[0] + fv * v[0], fy = fu * u[1] + fv * v[1], fz = fu * u[2] + fv * v[2];
      const float r = sqrtf((float)(fu * fu + fv * fv));
      float sre = 0.0f, sim = 0.0f;
      for (int k = 0; k < np; k++) {
        float px = 0.0f, py = 0.0f, pz = 0.0f, w = 0.0f, r0 = 0.0f;
        if (k < n) {
          px = c[4 * k];
          py = c[4 * k + 1];
          pz = c[4 * k + 2];
          w = c[4 * k + 3];
          r0 = h ? h[k] : 0.0f;
        }
        const float term = h ? port_phi(basis, mode, r * r0) : port_phi(basis, mode, r);
        const float vv = TWO_PI_F * (fx * px + fy * py + fz * pz);
        sre += w * term * cosf(vv);
        sim += w * term * sinf(vv);
      }
      if (g == 0) {
        out[2 * j] = sre * scale;
        out[2 * j + 1] = -sim * scale;
      } else {
        out[2 * j] += sre * scale;
        out[2 * j + 1] += -sim * scale;
      }
    }
  }
}

/* arrange_samples (meshless_vis_cpu.cpp:209-234) on an FP64 spectrum; out is zeroed first. */
void mvo_arrange64(const double* F, int cx, int cy, int Nx, int Ny, double* out) {
  const int ld = Ny / 2 + 1;
  memset(out, 0, sizeof(double) * 2 * (size_t)Nx * ld);
  const int size = 2 * cx * cy;
  for (int index = 0; index < size; ++index) {
    int x = index % (2 * cx);
    const int y = index / (2 * cx);
    int index_x = x;
    if (x > cx) {
      x = x - 2 * cx;
      index_x = x + Nx;
    }
    if (x != cx) {
      out[2 * ((size_t)index_x * ld + y)] = F[2 * index];
      out[2 * ((size_t)index_x * ld + y) + 1] = F[2 * index + 1];
    }
  }
}

/* Unnormalised 2-D C2R of an Nx x (Ny/2+1) Hermitian half array, by direct DFTs. */
void mvo_c2r64(const double* A, int Nx, int Ny, double* img, int nthreads) {
  const int ld = Ny / 2 + 1;
  double* T = (double*)malloc(sizeof(double) * 2 * (size_t)Nx * ld);
  double* cx_tab = (double*)malloc(sizeof(double) * 2 * (size_t)Nx);
  double* cy_tab = (double*)malloc(sizeof(double) * 2 * (size_t)Ny);
  for (int k = 0; k < Nx; ++k) {
    cx_tab[2 * k] = cos(2.0 * PI * k / Nx);
    cx_tab[2 * k + 1] = sin(2.0 * PI * k / Nx);
  }
  for (int k = 0; k < Ny; ++k) {
    cy_tab[2 * k] = cos(2.0 * PI * k / Ny);
    cy_tab[2 * k + 1] = sin(2.0 * PI * k / Ny);
  }
  if (nthreads > 0) omp_set_num_threads(nthreads);
  /* axis 0: T[r][c] = sum_r' A[r'][c] e^{+2 pi i r r'/Nx} */
#pragma omp parallel for schedule(static)
  for (int r = 0; r < Nx; ++r) {
    for (int c = 0; c < ld; ++c) {
      double re = 0.0, im = 0.0;
      for (int q = 0; q < Nx; ++q) {
        const int t = (int)(((long long)r * q) % Nx);
        const double ar = A[2 * ((size_t)q * ld + c)], ai = A[2 * ((size_t)q * ld + c) + 1];
        re += ar * cx_tab[2 * t] - ai * cx_tab[2 * t + 1];
        im += ar * cx_tab[2 * t + 1] + ai * cx_tab[2 * t];
      }
      T[2 * ((size_t)r * ld + c)] = re;
      T[2 * ((size_t)r * ld + c) + 1] = im;
    }
  }
  /* axis 1: img[r][n] = Re T[r][0] + 2 Re sum_{0<c<Ny/2} T[r][c] e^{2 pi i c n/Ny} + Re T[r][Ny/2] (-1)^n (Ny even) */
#pragma omp parallel for schedule(static)
  for (int r = 0; r < Nx; ++r) {
    for (int n = 0; n < Ny; ++n) {
      double acc = T[2 * ((size_t)r * ld)];
      const int top = (Ny % 2 == 0) ? Ny / 2 : ld;
      for (int c = 1; c < top; ++c) {
        const int t = (int)(((long long)c * n) % Ny);
        acc += 2.0 * (T[2 * ((size_t)r * ld + c)] * cy_tab[2 * t] - T[2 * ((size_t)r * ld + c) + 1] * cy_tab[2 * t + 1]);
      }
      if (Ny % 2 == 0) acc += T[2 * ((size_t)r * ld + Ny / 2)] * ((n & 1) ? -1.0 : 1.0);
      img[(size_t)r * Ny + n] = acc;
    }
  }
  free(T);
  free(cx_tab);
  free(cy_tab);
}
