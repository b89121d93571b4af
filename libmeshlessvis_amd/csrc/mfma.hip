// Split-precision MFMA complex-GEMM form of the direct sum.
//
// On the slice plane the phase separates, x*a_i + y*b_i, so for a group whose
// Phi(r h_i) factors into an x part and a y part the half-plane sum is a
// complex matrix product over the particles i:
//
//     S(x, y) = E(x, y) * sum_i A(x, i) B(i, y)
//     A(x, i) = alpha_i(x) exp(+2 pi i x a_i),  B(i, y) = w_i' beta_i(y) exp(+2 pi i y b_i)
//
//  * uniform radius h (or no radii, h = 1), any basis:
//      alpha = beta = 1, E = Phi(r h) * wmax
//  * Gaussian with per-particle radii (exact, no binning: exp(s pi h^2 r^2)
//    = exp(s pi h^2 fu^2) exp(s pi h^2 fv^2), s = -1 analytic, +1 reference):
//      alpha_i(x) = exp(s pi (h_i^2 - h_ref^2) fu^2), beta_i(y) likewise with fv,
//      E = exp(s pi h_ref^2 r^2) * wmax, h_ref the group's min (s=-1) or max (s=+1)
//      radius so that |alpha|, |beta| <= 1;
//  w_i' = w_i / wmax keeps |B| <= 1 (FP16 range).
//
// Kernel (gfx950, wave64, v_mfma_f32_32x32x16_{f16,bf16}):
//  * a workgroup is 8 waves and owns a 128 (x) x 128 (y) complex output tile;
//    a wave owns 64 x 32 (two 32x32 blocks, real and imaginary accumulators);
//  * the complex product runs as real MFMAs with K = (re, im) pairs of 8
//    particles: acc_re += [Ar Ai].[Br -Bi], acc_im += [Ar Ai].[Bi Br]; the B
//    operand forms are a sign flip and a 16-bit rotation of the packed (Br, Bi);
//  * operands are generated on the fly per 16-particle step into double-
//    buffered LDS as packed half pairs: phase from the particle's fixed-point
//    phase steps (exact modular arithmetic), v_sin/v_cos in revolutions, the
//    Gaussian factor, then hi = rnd(v), lo = rnd(v - hi) (split forms);
//    one barrier per step, generation of step s+1 overlaps the MFMAs of step s;
//  * split products: hi*hi + hi*lo + lo*hi into the same FP32 accumulators;
//  * accumulation: MFMA FP32 accumulators are folded into a second FP32 set
//    every 512 particles; split-K partials (FP64) are reduced in order.
#include "mfma_common.h"

namespace mv {

template <int OP, bool SPLIT, bool GAUSS, int BASIS, int MODE>
__global__ __launch_bounds__(kThreads, 1) void mfma_sum_kernel(const Packet* __restrict__ pk, MfmaParams p,
                                                                 double2* __restrict__ partial) {
  // [buffer][A_hi, A_lo, B_hi, B_lo][row * kPitch + particle], one packed (re, im) pair per word
  __shared__ uint32_t lds[2][4][kTileX * kPitch];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wx = wave & 1;
  const int wy = wave >> 1;
  const int j0 = (blockIdx.x % p.tiles_x) * kTileX;
  const int y0 = (blockIdx.x / p.tiles_x) * kTileY;
  const int ks = blockIdx.y;

  // Generation role: particle gp of the step, rows/columns gg + 32m.
  const int gp = tid & (kStep - 1);
  const int gg = tid >> 4;
  uint32_t ux[4], uy[4];
  float ca[4], cb[4];
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    const int x = j0 + gg + 32 * m - p.cx + 1;
    const int y = y0 + gg + 32 * m;
    ux[m] = static_cast<uint32_t>(x);
    uy[m] = static_cast<uint32_t>(y);
    const float fu = p.sx * static_cast<float>(x);
    const float fv = p.sy * static_cast<float>(y);
    ca[m] = p.sigma_pi_log2e * fu * fu;
    cb[m] = p.sigma_pi_log2e * fv * fv;
  }

  auto generate = [&](int step, int buf) {
    const Packet q = pk[static_cast<size_t>(step) * kStep + gp];
    const float4 bi = p.step_bin ? p.bininfo[p.step_bin[step]] : p.bin0;
    const float dh = q.h * q.h - (p.ref_hi ? bi.y : bi.x);
    const float wq = q.w * bi.w;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      // A(x, i)
      {
        const uint32_t P = ux[m] * q.pu;
        const float ph = static_cast<float>(static_cast<int32_t>(P)) * kTwoPowM32;
        float amp = 1.0f;
        if (GAUSS) amp = __builtin_amdgcn_exp2f(fminf(ca[m] * dh, 0.0f));
        const float re = amp * cos_rev(ph), im = amp * sin_rev(ph);
        const uint32_t hi = pack2<OP>(re, im);
        const int o = (gg + 32 * m) * kPitch + gp;
        lds[buf][0][o] = hi;
        if (SPLIT) {
          float hr, hm;
          unpack2<OP>(hi, &hr, &hm);
          lds[buf][1][o] = pack2<OP>(re - hr, im - hm);
        }
      }
      // B(i, y)
      {
        const uint32_t P = uy[m] * q.pv;
        const float ph = static_cast<float>(static_cast<int32_t>(P)) * kTwoPowM32;
        float amp = wq;
        if (GAUSS) amp *= __builtin_amdgcn_exp2f(fminf(cb[m] * dh, 0.0f));
        const float re = amp * cos_rev(ph), im = amp * sin_rev(ph);
        const uint32_t hi = pack2<OP>(re, im);
        const int o = (gg + 32 * m) * kPitch + gp;
        lds[buf][2][o] = hi;
        if (SPLIT) {
          float hr, hm;
          unpack2<OP>(hi, &hr, &hm);
          lds[buf][3][o] = pack2<OP>(re - hr, im - hm);
        }
      }
    }
  };

  f32x16_t acc[2][2], sum[2][2];  // [row block][re, im]
#pragma unroll
  for (int rb = 0; rb < 2; ++rb)
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      acc[rb][c] = (f32x16_t){};
      sum[rb][c] = (f32x16_t){};
    }

  auto consume = [&](int buf) {
#pragma unroll
    for (int kk = 0; kk < kStep / 8; ++kk) {
      const int pofs = kk * 8 + (lane >> 5) * 4;
      const int col = wy * 32 + (lane & 31);
      const u32x4_t bhi = *reinterpret_cast<const u32x4_t*>(&lds[buf][2][col * kPitch + pofs]);
      const u32x4_t bre_hi = negate_imag(bhi), bim_hi = swap_halves(bhi);
      u32x4_t bre_lo, bim_lo;
      if (SPLIT) {
        const u32x4_t blo = *reinterpret_cast<const u32x4_t*>(&lds[buf][3][col * kPitch + pofs]);
        bre_lo = negate_imag(blo);
        bim_lo = swap_halves(blo);
      }
#pragma unroll
      for (int rb = 0; rb < 2; ++rb) {
        const int row = wx * 64 + rb * 32 + (lane & 31);
        const u32x4_t ahi = *reinterpret_cast<const u32x4_t*>(&lds[buf][0][row * kPitch + pofs]);
        acc[rb][0] = mfma<OP>(ahi, bre_hi, acc[rb][0]);
        acc[rb][1] = mfma<OP>(ahi, bim_hi, acc[rb][1]);
        if (SPLIT) {
          const u32x4_t alo = *reinterpret_cast<const u32x4_t*>(&lds[buf][1][row * kPitch + pofs]);
          acc[rb][0] = mfma<OP>(ahi, bre_lo, acc[rb][0]);
          acc[rb][1] = mfma<OP>(ahi, bim_lo, acc[rb][1]);
          acc[rb][0] = mfma<OP>(alo, bre_hi, acc[rb][0]);
          acc[rb][1] = mfma<OP>(alo, bim_hi, acc[rb][1]);
        }
      }
    }
  };

  // sum += E_b(x, y) * acc; acc = 0.  E_b = exp(s pi h_ref^2 r^2) * W_b (Gaussian), Phi(r h) * W_b (uniform).
  const int y_out = y0 + wy * 32 + (lane & 31);
  const float fv = p.sy * static_cast<float>(y_out);
  auto fold = [&](int b) {
    const float4 bi = p.step_bin ? p.bininfo[b] : p.bin0;
    const float hh_ref = p.ref_hi ? bi.y : bi.x;
#pragma unroll
    for (int rb = 0; rb < 2; ++rb) {
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        const int j = j0 + wx * 64 + rb * 32 + (i >> 2) * 8 + (lane >> 5) * 4 + (i & 3);
        const float fu = p.sx * static_cast<float>(j - p.cx + 1);
        const float r2 = fu * fu + fv * fv;
        float e;
        if (GAUSS)
          e = __builtin_amdgcn_exp2f(p.sigma_pi_log2e * hh_ref * r2) * bi.z;
        else
          e = phi<BASIS, MODE>(sqrtf(r2) * p.radius) * bi.z;
        sum[rb][0][i] = fmaf(e, acc[rb][0][i], sum[rb][0][i]);
        sum[rb][1][i] = fmaf(e, acc[rb][1][i], sum[rb][1][i]);
        __builtin_amdgcn_sched_barrier(0);
      }
      acc[rb][0] = acc[rb][1] = (f32x16_t){};
    }
  };

  const int s_begin = ks * p.steps_per_split;
  const int s_end = min(p.steps_total, s_begin + p.steps_per_split);
  int cur = (p.step_bin && s_begin < s_end) ? p.step_bin[s_begin] : 0;
  int since = 0;
  if (s_begin < s_end) generate(s_begin, 0);
  __syncthreads();
  for (int s = s_begin; s < s_end; ++s) {
    const int buf = (s - s_begin) & 1;
    if (s + 1 < s_end) generate(s + 1, buf ^ 1);
    if (p.step_bin) {
      const int b = p.step_bin[s];
      if (b != cur) {
        fold(cur);
        cur = b;
        since = 0;
      }
    }
    consume(buf);
    if (++since == kFlushSteps) {
      fold(cur);
      since = 0;
    }
    __syncthreads();
  }
  if (since) fold(cur);

  // partial[ks][j][y]
  const int two_cx = 2 * p.cx;
  if (y_out >= p.cy) return;
#pragma unroll
  for (int rb = 0; rb < 2; ++rb) {
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const int j = j0 + wx * 64 + rb * 32 + (i >> 2) * 8 + (lane >> 5) * 4 + (i & 3);
      if (j >= two_cx) continue;
      partial[(static_cast<size_t>(ks) * two_cx + j) * p.cy + y_out] =
          make_double2(static_cast<double>(sum[rb][0][i]), static_cast<double>(sum[rb][1][i]));
    }
  }
}

// acc[y*2cx + x0] (+)= sum_k partial[k][j][y], x0 the reference column of x = j - cx + 1.
__global__ __launch_bounds__(256) void mfma_reduce_kernel(const double2* __restrict__ partial, int ksplit, int cx,
                                                          int cy, int first, double2* __restrict__ acc) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  const int two_cx = 2 * cx;
  const int nfreq = two_cx * cy;
  if (f >= nfreq) return;
  double re = 0.0, im = 0.0;
  for (int k = 0; k < ksplit; ++k) {
    const double2 v = partial[static_cast<size_t>(k) * nfreq + f];
    re += v.x;
    im += v.y;
  }
  const int j = f / cy;
  const int y = f - j * cy;
  const int x = j - cx + 1;
  const int x0 = (x >= 0) ? x : x + two_cx;
  const int o = y * two_cx + x0;
  if (!first) {
    re += acc[o].x;
    im += acc[o].y;
  }
  acc[o] = make_double2(re, im);
}

// ------------------------------------------------------------------ host ---
template <int OP, bool SPLIT, bool GAUSS, int BASIS, int MODE>
static void launch_k(dim3 grid, hipStream_t s, const Packet* pk, const MfmaParams& p, double2* partial) {
  hipLaunchKernelGGL((mfma_sum_kernel<OP, SPLIT, GAUSS, BASIS, MODE>), grid, dim3(kThreads), 0, s, pk, p, partial);
}

template <int OP, bool SPLIT>
static void launch_kind(dim3 grid, hipStream_t s, const Packet* pk, const MfmaParams& p, double2* partial,
                        bool gauss) {
  if (gauss) {
    launch_k<OP, SPLIT, true, GAUSSIAN, VIS_PHI_ANALYTIC>(grid, s, pk, p, partial);
    return;
  }
  const bool an = p.phi_mode == VIS_PHI_ANALYTIC;
  switch (p.basis) {
    case SPH:
      an ? launch_k<OP, SPLIT, false, SPH, VIS_PHI_ANALYTIC>(grid, s, pk, p, partial)
         : launch_k<OP, SPLIT, false, SPH, VIS_PHI_REFERENCE>(grid, s, pk, p, partial);
      break;
    case GAUSSIAN:
      an ? launch_k<OP, SPLIT, false, GAUSSIAN, VIS_PHI_ANALYTIC>(grid, s, pk, p, partial)
         : launch_k<OP, SPLIT, false, GAUSSIAN, VIS_PHI_REFERENCE>(grid, s, pk, p, partial);
      break;
    default:
      an ? launch_k<OP, SPLIT, false, WENDLAND_D3_C2, VIS_PHI_ANALYTIC>(grid, s, pk, p, partial)
         : launch_k<OP, SPLIT, false, WENDLAND_D3_C2, VIS_PHI_REFERENCE>(grid, s, pk, p, partial);
  }
}

bool mfma_group_spectrum(const Group& g, const GroupInfo& info, int n, const ViewParams& vp, double3 u, double3 v,
                         const VisConfig* c, bool first, hipStream_t s, MfmaWorkspace* ws, double2* acc,
                         EventPair* ev) {
  const bool uniform = info.uniform_radius;
  const bool gauss = !uniform && vp.basis == GAUSSIAN;
  if (!uniform && !gauss) return mfma_binned_group_spectrum(g, info, n, vp, u, v, c, first, s, ws, acc, ev);
  const int nfreq = 2 * vp.cx * vp.cy;

  MfmaParams p;
  p.cx = vp.cx;
  p.cy = vp.cy;
  p.sx = vp.sx;
  p.sy = vp.sy;
  p.basis = vp.basis;
  p.phi_mode = vp.phi_mode;
  const float sigma = (vp.phi_mode == VIS_PHI_REFERENCE) ? 1.0f : -1.0f;
  p.sigma_pi_log2e = gauss ? sigma * kPi * kLog2e : 0.0f;
  p.ref_hi = sigma > 0.0f ? 1 : 0;
  p.radius = info.has_radii ? info.radius : 1.0f;
  p.tiles_x = (2 * vp.cx + kTileX - 1) / kTileX;
  const int tiles_y = (vp.cy + kTileY - 1) / kTileY;
  if (gauss) {
    // Bins of h ratio 1.25: per-bin weight scale and reference radius.
    const int nb = c->number_of_h_bins > 0 ? c->number_of_h_bins : auto_bins(info.h_min, info.h_max, 1.25);
    if (nb > 65535) {
      set_error(VIS_ERROR_INVALID_CONFIG, "number_of_h_bins above 65535");
      return false;
    }
    const BinCache* bc = bin_cache(g, info, n, nb, 1, s);
    if (!bc) return false;
    if (!ws->packets.reserve(sizeof(Packet) * bc->n_slots)) return false;
    if (!hip_ok(launch_prep_packets_mapped(g.d_constraints, g.d_radii, bc->slot_map.as<int>(), bc->n_slots,
                                           c->step_size.x, c->step_size.y, u, v, ws->packets.as<Packet>(), s),
                "prep_packets_mapped launch"))
      return false;
    p.steps_total = bc->n_slots / kStep;
    p.step_bin = bc->step_bin.as<uint16_t>();
    p.bininfo = bc->bininfo.as<float4>();
    p.bin0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  } else {
    const int n_padded = ((n + 255) / 256) * 256;
    if (!ws->packets.reserve(sizeof(Packet) * std::max(n_padded, 256))) return false;
    if (!hip_ok(launch_prep_packets(g.d_constraints, g.d_radii, n, n_padded, c->step_size.x, c->step_size.y, u, v,
                                    ws->packets.as<Packet>(), s),
                "prep_packets launch"))
      return false;
    p.steps_total = n_padded / kStep;
    p.step_bin = nullptr;
    p.bininfo = nullptr;
    p.bin0 = make_float4(0.0f, 0.0f, info.w_absmax, info.w_absmax > 0.0f ? 1.0f / info.w_absmax : 0.0f);
  }
  const int tiles = p.tiles_x * tiles_y;
  // Enough workgroups to fill 256 CUs a few times, at least 64 steps (1024 particles) per split.
  int ksplit = std::max(1, (1024 + tiles - 1) / tiles);
  ksplit = std::min(ksplit, std::max(1, p.steps_total / 64));
  p.steps_per_split = (p.steps_total + ksplit - 1) / ksplit;
  ksplit = std::max(1, (p.steps_total + p.steps_per_split - 1) / p.steps_per_split);
  if (!ws->partial.reserve(sizeof(double2) * static_cast<size_t>(ksplit) * nfreq)) return false;

  const dim3 grid(tiles, ksplit);
  if (ev && !hip_ok(hipEventRecord(ev->first, s), "hipEventRecord")) return false;
  const int form = c->form;
  if (form == VIS_FORM_MFMA_F16)
    launch_kind<OP_F16, false>(grid, s, ws->packets.as<Packet>(), p, ws->partial.as<double2>(), gauss);
  else if (form == VIS_FORM_MFMA_SPLIT_BF16)
    launch_kind<OP_BF16, true>(grid, s, ws->packets.as<Packet>(), p, ws->partial.as<double2>(), gauss);
  else
    launch_kind<OP_F16, true>(grid, s, ws->packets.as<Packet>(), p, ws->partial.as<double2>(), gauss);
  if (!hip_ok(hipGetLastError(), "mfma_sum launch")) return false;
  if (ev && !hip_ok(hipEventRecord(ev->second, s), "hipEventRecord")) return false;
  return launch_mfma_reduce(ws->partial.as<double2>(), ksplit, vp.cx, vp.cy, first, acc, s);
}

bool launch_mfma_reduce(const double2* partial, int ksplit, int cx, int cy, bool first, double2* acc, hipStream_t s) {
  const int nfreq = 2 * cx * cy;
  hipLaunchKernelGGL(mfma_reduce_kernel, dim3((nfreq + 255) / 256), dim3(256), 0, s, partial, ksplit, cx, cy,
                     first ? 1 : 0, acc);
  return hip_ok(hipGetLastError(), "mfma_reduce launch");
}


}  // namespace mv
