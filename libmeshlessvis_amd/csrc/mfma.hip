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
                                                                 const int4* __restrict__ chunks,
                                                                 float2* __restrict__ partial) {
  // [buffer][A_hi, A_lo, B_hi, B_lo][row * kPitch + particle], one packed (re, im) pair per word
  __shared__ uint32_t lds[2][4][kTileX * kPitch];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int tiles = p.tiles_x * p.tiles_y;
  const int tile = blockIdx.x % tiles;
  const int chunk = blockIdx.x / tiles;
  const int j0 = (tile % p.tiles_x) * kTileX;
  const int y0 = (tile / p.tiles_x) * kTileY;
  const int4 ch = chunks[chunk];  // (bin, first step, end step, -)
  const float4 bi = p.bininfo ? p.bininfo[ch.x] : p.bin0;
  const float hh_ref = p.ref_hi ? bi.y : bi.x;

  // Waves w and w+4 of a workgroup share a SIMD on gfx950, so every SIMD runs one producer
  // and one consumer wave and the MFMAs of one overlap the VALU work of the other.
  const int role_idx = wave & (kProducerWaves - 1);
  if (wave < kProducerWaves) {
    // ---- producer: operands of step s+1 while the consumers multiply step s ----
    const int ptid = role_idx * 64 + lane;
    const int gp = ptid & (kStep - 1);  // particle of the step
    const int gg = ptid >> 4;           // 0..15: rows / columns gg + 16m
    uint32_t ux[8], uy[8];
    float ca[8], cb[8];
#pragma unroll
    for (int m = 0; m < 8; ++m) {
      const int x = j0 + gg + 16 * m - p.cx + 1;
      const int y = y0 + gg + 16 * m;
      ux[m] = static_cast<uint32_t>(x);
      uy[m] = static_cast<uint32_t>(y);
      const float fu = p.sx * static_cast<float>(x);
      const float fv = p.sy * static_cast<float>(y);
      ca[m] = p.sigma_pi_log2e * fu * fu;
      cb[m] = p.sigma_pi_log2e * fv * fv;
    }
    auto generate = [&](int step, int buf) {
      const Packet q = pk[static_cast<size_t>(step) * kStep + gp];
      const float dh = q.h * q.h - hh_ref;
      const float wq = q.w * bi.w;
#pragma unroll
      for (int m = 0; m < 8; ++m) {
        const int o = (gg + 16 * m) * kPitch + gp;
        {  // A(x, i)
          const uint32_t P = ux[m] * q.pu;
          const float ph = static_cast<float>(static_cast<int32_t>(P)) * kTwoPowM32;
          float amp = 1.0f;
          if (GAUSS) amp = __builtin_amdgcn_exp2f(fminf(ca[m] * dh, 0.0f));
          const float re = amp * cos_rev(ph), im = amp * sin_rev(ph);
          const uint32_t hi = pack2<OP>(re, im);
          lds[buf][0][o] = hi;
          if (SPLIT) {
            float hr, hm;
            unpack2<OP>(hi, &hr, &hm);
            lds[buf][1][o] = pack2<OP>(re - hr, im - hm);
          }
        }
        {  // B(i, y)
          const uint32_t P = uy[m] * q.pv;
          const float ph = static_cast<float>(static_cast<int32_t>(P)) * kTwoPowM32;
          float amp = wq;
          if (GAUSS) amp *= __builtin_amdgcn_exp2f(fminf(cb[m] * dh, 0.0f));
          const float re = amp * cos_rev(ph), im = amp * sin_rev(ph);
          const uint32_t hi = pack2<OP>(re, im);
          lds[buf][2][o] = hi;
          if (SPLIT) {
            float hr, hm;
            unpack2<OP>(hi, &hr, &hm);
            lds[buf][3][o] = pack2<OP>(re - hr, im - hm);
          }
        }
      }
    };
    if (ch.y < ch.z) generate(ch.y, 0);
    __syncthreads();
    for (int s = ch.y; s < ch.z; ++s) {
      if (s + 1 < ch.z) generate(s + 1, (s - ch.y + 1) & 1);
      __syncthreads();
    }
    return;
  }

  // ---- consumer: a 64 (x) x 64 (y) quarter of the tile, four 32x32 blocks ----
  const int cw = role_idx;
  const int wx = cw & 1;
  const int wy = cw >> 1;
  f32x16_t acc[2][2][2];  // [row block][column block][re, im]
#pragma unroll
  for (int rb = 0; rb < 2; ++rb)
#pragma unroll
    for (int cb = 0; cb < 2; ++cb) acc[rb][cb][0] = acc[rb][cb][1] = (f32x16_t){};

  __syncthreads();
  for (int s = ch.y; s < ch.z; ++s) {
    const int buf = (s - ch.y) & 1;
#pragma unroll
    for (int kk = 0; kk < kStep / 8; ++kk) {
      const int pofs = kk * 8 + (lane >> 5) * 4;
      u32x4_t bre_hi[2], bim_hi[2], bre_lo[2], bim_lo[2];
#pragma unroll
      for (int cb = 0; cb < 2; ++cb) {
        const int col = wy * 64 + cb * 32 + (lane & 31);
        const u32x4_t bhi = *reinterpret_cast<const u32x4_t*>(&lds[buf][2][col * kPitch + pofs]);
        bre_hi[cb] = negate_imag(bhi);
        bim_hi[cb] = swap_halves(bhi);
        if (SPLIT) {
          const u32x4_t blo = *reinterpret_cast<const u32x4_t*>(&lds[buf][3][col * kPitch + pofs]);
          bre_lo[cb] = negate_imag(blo);
          bim_lo[cb] = swap_halves(blo);
        }
      }
#pragma unroll
      for (int rb = 0; rb < 2; ++rb) {
        const int row = wx * 64 + rb * 32 + (lane & 31);
        const u32x4_t ahi = *reinterpret_cast<const u32x4_t*>(&lds[buf][0][row * kPitch + pofs]);
        u32x4_t alo;
        if (SPLIT) alo = *reinterpret_cast<const u32x4_t*>(&lds[buf][1][row * kPitch + pofs]);
#pragma unroll
        for (int cb = 0; cb < 2; ++cb) {
          acc[rb][cb][0] = mfma<OP>(ahi, bre_hi[cb], acc[rb][cb][0]);
          acc[rb][cb][1] = mfma<OP>(ahi, bim_hi[cb], acc[rb][cb][1]);
          if (SPLIT) {
            acc[rb][cb][0] = mfma<OP>(ahi, bre_lo[cb], acc[rb][cb][0]);
            acc[rb][cb][1] = mfma<OP>(ahi, bim_lo[cb], acc[rb][cb][1]);
            acc[rb][cb][0] = mfma<OP>(alo, bre_hi[cb], acc[rb][cb][0]);
            acc[rb][cb][1] = mfma<OP>(alo, bim_hi[cb], acc[rb][cb][1]);
          }
        }
      }
    }
    __syncthreads();
  }

  // Epilogue: partial[chunk][j][y] = E_b(x, y) * acc,
  // E_b = exp(s pi h_ref^2 r^2) W_b (Gaussian) or Phi(r h) W_b (uniform radius).
  const int two_cx = 2 * p.cx;
#pragma unroll
  for (int cb = 0; cb < 2; ++cb) {
    const int y = y0 + wy * 64 + cb * 32 + (lane & 31);
    if (y >= p.cy) continue;
    const float fv = p.sy * static_cast<float>(y);
#pragma unroll
    for (int rb = 0; rb < 2; ++rb) {
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        const int j = j0 + wx * 64 + rb * 32 + (i >> 2) * 8 + (lane >> 5) * 4 + (i & 3);
        if (j >= two_cx) continue;
        const float fu = p.sx * static_cast<float>(j - p.cx + 1);
        const float r2 = fu * fu + fv * fv;
        float e;
        if (GAUSS)
          e = __builtin_amdgcn_exp2f(p.sigma_pi_log2e * hh_ref * r2) * bi.z;
        else
          e = phi<BASIS, MODE>(sqrtf(r2) * p.radius) * bi.z;
        partial[(static_cast<size_t>(chunk) * two_cx + j) * p.cy + y] =
            make_float2(e * acc[rb][cb][0][i], e * acc[rb][cb][1][i]);
      }
    }
  }
}

// acc[y*2cx + x0] (+)= sum_k partial[k][j][y] (in order, FP64), x0 the reference column of x = j - cx + 1.
template <class PT>
__global__ __launch_bounds__(256) void mfma_reduce_kernel(const PT* __restrict__ partial, int ksplit, int cx, int cy,
                                                          int first, double2* __restrict__ acc) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  const int two_cx = 2 * cx;
  const int nfreq = two_cx * cy;
  if (f >= nfreq) return;
  double re = 0.0, im = 0.0;
  for (int k = 0; k < ksplit; ++k) {
    const PT v = partial[static_cast<size_t>(k) * nfreq + f];
    re += static_cast<double>(v.x);
    im += static_cast<double>(v.y);
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
static void launch_k(dim3 grid, hipStream_t s, const Packet* pk, const MfmaParams& p, const int4* chunks,
                     float2* partial) {
  hipLaunchKernelGGL((mfma_sum_kernel<OP, SPLIT, GAUSS, BASIS, MODE>), grid, dim3(kThreads), 0, s, pk, p, chunks,
                     partial);
}

template <int OP, bool SPLIT>
static void launch_kind(dim3 grid, hipStream_t s, const Packet* pk, const MfmaParams& p, const int4* chunks,
                        float2* partial, bool gauss) {
  if (gauss) {
    launch_k<OP, SPLIT, true, GAUSSIAN, VIS_PHI_ANALYTIC>(grid, s, pk, p, chunks, partial);
    return;
  }
  const bool an = p.phi_mode == VIS_PHI_ANALYTIC;
  switch (p.basis) {
    case SPH:
      an ? launch_k<OP, SPLIT, false, SPH, VIS_PHI_ANALYTIC>(grid, s, pk, p, chunks, partial)
         : launch_k<OP, SPLIT, false, SPH, VIS_PHI_REFERENCE>(grid, s, pk, p, chunks, partial);
      break;
    case GAUSSIAN:
      an ? launch_k<OP, SPLIT, false, GAUSSIAN, VIS_PHI_ANALYTIC>(grid, s, pk, p, chunks, partial)
         : launch_k<OP, SPLIT, false, GAUSSIAN, VIS_PHI_REFERENCE>(grid, s, pk, p, chunks, partial);
      break;
    default:
      an ? launch_k<OP, SPLIT, false, WENDLAND_D3_C2, VIS_PHI_ANALYTIC>(grid, s, pk, p, chunks, partial)
         : launch_k<OP, SPLIT, false, WENDLAND_D3_C2, VIS_PHI_REFERENCE>(grid, s, pk, p, chunks, partial);
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
  p.tiles_y = (vp.cy + kTileY - 1) / kTileY;
  std::vector<int> bin_step;  // first step of every bin
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
    p.bininfo = bc->bininfo.as<float4>();
    p.bin0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    bin_step = bc->bin_step;
  } else {
    const int n_padded = ((n + 255) / 256) * 256;
    if (!ws->packets.reserve(sizeof(Packet) * std::max(n_padded, 256))) return false;
    if (!hip_ok(launch_prep_packets(g.d_constraints, g.d_radii, n, n_padded, c->step_size.x, c->step_size.y, u, v,
                                    ws->packets.as<Packet>(), s),
                "prep_packets launch"))
      return false;
    p.bininfo = nullptr;
    p.bin0 = make_float4(0.0f, 0.0f, info.w_absmax, info.w_absmax > 0.0f ? 1.0f / info.w_absmax : 0.0f);
    bin_step = {0, n_padded / kStep};
  }
  // Work items: (bin, step range) chunks of at most kMaxChunkSteps steps, sized so that
  // tiles x chunks fills the chip several times over.
  const int tiles = p.tiles_x * p.tiles_y;
  const long long total_steps = bin_step.back();
  const int chunk_steps = static_cast<int>(
      std::min<long long>(kMaxChunkSteps, std::max<long long>(32, total_steps * tiles / 2048)));
  std::vector<int4> chunks;
  for (size_t b = 0; b + 1 < bin_step.size(); ++b)
    for (int s0 = bin_step[b]; s0 < bin_step[b + 1]; s0 += chunk_steps)
      chunks.push_back(make_int4(static_cast<int>(b), s0, std::min(bin_step[b + 1], s0 + chunk_steps), 0));
  if (chunks.empty()) chunks.push_back(make_int4(0, 0, 0, 0));
  const int nchunks = static_cast<int>(chunks.size());
  const bool same = ws->chunks_host.size() == chunks.size() &&
                    std::equal(chunks.begin(), chunks.end(), ws->chunks_host.begin(), [](const int4& a, const int4& b) {
                      return a.x == b.x && a.y == b.y && a.z == b.z;
                    });
  if (!same) {
    if (!ws->chunks.reserve(sizeof(int4) * nchunks)) return false;
    if (!hip_ok(hipStreamSynchronize(s), "synchronize") ||
        !hip_ok(hipMemcpy(ws->chunks.ptr, chunks.data(), sizeof(int4) * nchunks, hipMemcpyHostToDevice), "upload"))
      return false;
    ws->chunks_host = chunks;
  }
  if (!ws->partial.reserve(sizeof(float2) * static_cast<size_t>(nchunks) * nfreq)) return false;

  const dim3 grid(tiles * nchunks);
  if (ev && !hip_ok(hipEventRecord(ev->first, s), "hipEventRecord")) return false;
  const int form = c->form;
  const int4* ck = ws->chunks.as<int4>();
  float2* part = ws->partial.as<float2>();
  if (form == VIS_FORM_MFMA_F16)
    launch_kind<OP_F16, false>(grid, s, ws->packets.as<Packet>(), p, ck, part, gauss);
  else if (form == VIS_FORM_MFMA_SPLIT_BF16)
    launch_kind<OP_BF16, true>(grid, s, ws->packets.as<Packet>(), p, ck, part, gauss);
  else
    launch_kind<OP_F16, true>(grid, s, ws->packets.as<Packet>(), p, ck, part, gauss);
  if (!hip_ok(hipGetLastError(), "mfma_sum launch")) return false;
  if (ev && !hip_ok(hipEventRecord(ev->second, s), "hipEventRecord")) return false;
  return launch_mfma_reduce(part, nchunks, vp.cx, vp.cy, first, acc, s);
}

bool launch_mfma_reduce(const double2* partial, int ksplit, int cx, int cy, bool first, double2* acc, hipStream_t s) {
  const int nfreq = 2 * cx * cy;
  hipLaunchKernelGGL(mfma_reduce_kernel<double2>, dim3((nfreq + 255) / 256), dim3(256), 0, s, partial, ksplit, cx, cy,
                     first ? 1 : 0, acc);
  return hip_ok(hipGetLastError(), "mfma_reduce launch");
}

bool launch_mfma_reduce(const float2* partial, int ksplit, int cx, int cy, bool first, double2* acc, hipStream_t s) {
  const int nfreq = 2 * cx * cy;
  hipLaunchKernelGGL(mfma_reduce_kernel<float2>, dim3((nfreq + 255) / 256), dim3(256), 0, s, partial, ksplit, cx, cy,
                     first ? 1 : 0, acc);
  return hip_ok(hipGetLastError(), "mfma_reduce launch");
}

}  // namespace mv
