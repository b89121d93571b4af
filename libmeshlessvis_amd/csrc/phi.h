// Device-side Fourier transforms Phi(r) of the three bases, in FP32.
//
// Two families:
//  * reference: the formulas of fourier_transform_cpu.cpp:37-66 (and the
//    identical fourier_transform.cu:31-60): closed forms above a switch radius
//    (0.06 SPH, 0.24 Wendland), a quadratic fit below it, and exp(+pi r^2) for
//    the Gaussian.  The three fit coefficients of each quadratic are data that
//    must equal the reference's to reproduce its Phi.
//  * analytic: the exact transforms.  With m = pi*r:
//      SPH (M4 spline, support 2):  3*pi * (sin m / m)^3 * (sin m - m cos m) / m^3
//      Wendland C2 (unit ball):     7.5*pi * (4m^2 - 6 + (6 - m^2) cos 2m + 4.5 m sin 2m) / m^8
//      Gaussian exp(-pi|x|^2):      exp(-pi r^2)
//    evaluated through their Maclaurin series near 0 (where the closed forms
//    cancel catastrophically in FP32) and the closed forms beyond.
//
// Trigonometric arguments are formed in revolutions and reduced to [-1/2, 1/2]
// before the gfx950 v_sin_f32 / v_cos_f32 instructions, which take revolutions.
#pragma once
#include <hip/hip_runtime.h>

namespace mv {

constexpr float kPi = 3.14159265358979323846f;

// sin(2*pi*t), cos(2*pi*t); t must already be reduced to [-1/2, 1/2].
__device__ __forceinline__ float sin_rev(float t) { return __builtin_amdgcn_sinf(t); }
__device__ __forceinline__ float cos_rev(float t) { return __builtin_amdgcn_cosf(t); }
__device__ __forceinline__ float reduce_rev(float t) { return t - __builtin_rintf(t); }

__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }

// exp(x) through the hardware 2^x.
__device__ __forceinline__ float exp_hw(float x) { return __builtin_amdgcn_exp2f(x * 1.44269504088896340736f); }

// ---- analytic transforms ----------------------------------------------------
// SPH: g(m) = (sin m - m cos m)/m^3 = sum_{n>=1} (-1)^(n+1) 2n m^(2n-2) / (2n+1)!
__device__ __forceinline__ float phi_sph_analytic(float r) {
  const float m = kPi * r;
  const float t = reduce_rev(0.5f * r);  // m in revolutions
  const float sm = sin_rev(t), cm = cos_rev(t);
  const float m2 = m * m;
  const float sinc = (m > 1e-4f) ? sm * rcp(m) : 1.0f - m2 * (1.0f / 6.0f);
  float g;
  if (m < 1.0f) {
    g = -4.4983316069528332e-14f;
    g = g * m2 + 1.0706029224547743e-11f;
    g = g * m2 - 1.9270852604185938e-09f;
    g = g * m2 + 2.5052108385441719e-07f;
    g = g * m2 - 2.2045855379188713e-05f;
    g = g * m2 + 1.1904761904761905e-03f;
    g = g * m2 - 3.3333333333333333e-02f;
    g = g * m2 + 3.3333333333333333e-01f;
  } else {
    g = (sm - m * cm) * rcp(m2 * m);
  }
  return (3.0f * kPi) * (sinc * sinc * sinc) * g;
}

__device__ __forceinline__ float phi_gaussian_analytic(float r) { return exp_hw(-kPi * r * r); }

// Wendland: 7.5*pi * sum_{k>=4} c_k m^(2k-8); coefficients below are 7.5*pi*c_k.
__device__ __forceinline__ float phi_wendland_analytic(float r) {
  const float m = kPi * r;
  const float m2 = m * m;
  if (m < 2.0f) {
    float p = -4.3128532568967781e-16f;
    p = p * m2 + 5.7341344438286709e-14f;
    p = p * m2 - 6.3304844259868527e-12f;
    p = p * m2 + 5.6868851760115227e-10f;
    p = p * m2 - 4.0519056879082099e-08f;
    p = p * m2 + 2.2140770366069861e-06f;
    p = p * m2 - 8.8563081464279445e-05f;
    p = p * m2 + 2.4177721239748289e-03f;
    p = p * m2 - 3.9893240045584676e-02f;
    p = p * m2 + 2.9919930034188507e-01f;
    return p;
  }
  const float t = reduce_rev(r);
  const float c = cos_rev(t), s = sin_rev(t);
  const float m4 = m2 * m2;
  return ((kPi * 7.5f) * rcp(m4 * m4)) * (4.0f * m2 - 6.0f + (6.0f - m2) * c + 4.5f * m * s);
}

// ---- reference formulas (fourier_transform_cpu.cpp:37-66) -------------------
// Above the switch radius the reference's closed forms are algebraically the
// exact transforms ((2.356.../m^6)(cos2m-1)(cos2m+m sin2m-1) equals
// 3 pi sin^3 m (sin m - m cos m)/m^6, with 2.356... = 3 pi/4), so they are
// evaluated through the cancellation-free analytic routines; below it the
// reference's quadratic fits are used as they are.
__device__ __forceinline__ float phi_sph_reference(float r) {
  if (r >= 0.06f) return phi_sph_analytic(r);
  return 3.14159265358979f + r * (-0.007968913156311f + r * (-18.293608272337678f));
}

__device__ __forceinline__ float phi_gaussian_reference(float r) { return exp_hw(kPi * r * r); }

__device__ __forceinline__ float phi_wendland_reference(float r) {
  if (r >= 0.24f) return phi_wendland_analytic(r);
  return 0.299199300341890f + r * (-0.002379178586124f + r * (-0.370530218163545f));
}

template <int BASIS, int MODE>
__device__ __forceinline__ float phi(float r) {
  if (MODE == VIS_PHI_REFERENCE) {
    if (BASIS == SPH) return phi_sph_reference(r);
    if (BASIS == GAUSSIAN) return phi_gaussian_reference(r);
    return phi_wendland_reference(r);
  } else {
    if (BASIS == SPH) return phi_sph_analytic(r);
    if (BASIS == GAUSSIAN) return phi_gaussian_analytic(r);
    return phi_wendland_analytic(r);
  }
}

}  // namespace mv
