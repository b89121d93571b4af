// Split-precision MFMA complex-GEMM form of the direct sum.
//
// On the slice plane the phase separates, x*a_i + y*b_i, so for a group whose
// Phi(r h_i) factors into an x part and a y part the half-plane sum is a
// complex matrix product over the particles i:
//
//     S(x, y) = E_b(x, y) * sum_{i in b} A(x, i) B(i, y)        (b: h-bin of i)
//     A(x, i) = alpha_i(x) exp(+2 pi i x a_i),  B(i, y) = (w_i / W_b) beta_i(y) exp(+2 pi i y b_i)
//
//  * uniform radius h (or no radii, h = 1), any basis, one bin:
//      alpha = beta = 1, E = Phi(r h) * W
//  * Gaussian with per-particle radii (exact, no interpolation: exp(s pi h^2 r^2)
//    = exp(s pi h^2 fu^2) exp(s pi h^2 fv^2), s = -1 analytic, +1 reference),
//    log-spaced h-bins of ratio 2 (default):
//      alpha_i(x) = exp(s pi (h_i^2 - h_b^2) fu^2), beta_i(y) likewise with fv,
//      E_b = exp(s pi h_b^2 r^2) * W_b, h_b the bin's lower (s=-1) or upper (s=+1)
//      edge so that |alpha|, |beta| <= 1;
//  W_b = the bin's largest |w| keeps |B| <= 1 and FP16 operands out of the
//  subnormal range (weights correlate with h).
//
// Kernel (gfx950, wave64, v_mfma_f32_32x32x16_{f16,bf16}), persistent:
//  * the workgroups pull work items (job, 128x128 output tile, range of
//    16-particle steps inside one bin) from an atomic counter, items longest
//    first; one launch covers every view and group of a render batch;
//  * 8 waves: waves 0-3 generate operands, waves 4-7 run the MFMAs.  Waves w
//    and w+4 share a SIMD on gfx950, so each SIMD overlaps one wave's VALU
//    work with the other's matrix work;
//  * operands are generated per 16-particle step into double-buffered LDS as
//    packed half pairs: phase from the particle's fixed-point phase steps
//    (exact modular arithmetic), v_sin/v_cos in revolutions, the Gaussian
//    factor, then hi = rnd(v), lo = rnd(v - hi) (split forms); one barrier
//    per step;
//  * the complex product runs as real MFMAs with K = (re, im) pairs of 8
//    particles: acc_re += [Ar Ai].[Br -Bi], acc_im += [Ar Ai].[Bi Br]; the B
//    forms are a sign flip and a 16-bit rotation of the packed (Br, Bi);
//    split products hi*hi + hi*lo + lo*hi share the FP32 accumulators;
//  * an item spans at most 16384 particles of FP32 accumulation; its partial
//    tile (FP32) is reduced with the job's other items in FP64, in a fixed order.
#include "mfma_common.h"

#include <atomic>
#include <cstring>
#include <type_traits>
#include <tuple>

namespace mv {

// One stage = 32 particles; a word of LDS holds one component of two consecutive
// particles, so an LDS row is 16 words (+4 of padding against bank conflicts).
constexpr int kStage = 32;
constexpr int kRowPitch = kStage / 2 + 4;

// Threads of the sum kernel with a TX x TY tile: 4 producer waves and one consumer wave per
// 32 |x| x 64 y block.
constexpr int sum_threads(int TX, int TY) { return (kProducerWaves + (TX / 32) * (TY / 64)) * 64; }

template <int OP, bool SPLIT, bool GAUSS, int BASIS, int MODE, int TX, int TY>
__global__ __launch_bounds__(sum_threads(TX, TY), 1) void mfma_sum_kernel(const MfmaJob* __restrict__ jobs,
                                                                            const MfmaItem* __restrict__ items,
                                                                            MfmaShared sh, int* __restrict__ counter) {
  static_assert(TX % 32 == 0 && TY % 64 == 0 && TX <= kTileX && TY <= kTileY, "tile");
  // [buffer][re_hi, re_lo, im_hi, im_lo][row * kRowPitch + particle pair]; A rows are |x|, B rows are y.
  // With the default 128 x 128 tile together exactly 160 KiB: the work-item index lives in a
  // padding word (words 16..19 of every row are never read or written by the operand code).
  __shared__ uint32_t ldsA[2][4][TX * kRowPitch];
  __shared__ uint32_t ldsB[2][4][TY * kRowPitch];
  int* next_item = reinterpret_cast<int*>(&ldsA[0][0][kStage / 2]);

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  // Waves w, w+4 and w+8 of the workgroup share SIMD w (mod 4): every SIMD runs one
  // producer and two consumer waves, and the MFMAs of the consumers overlap the VALU
  // work of the producer.
  const int two_cx = 2 * sh.cx;

  for (;;) {
    if (tid == 0) *next_item = atomicAdd(counter, 1);
    __syncthreads();
    const int it = *next_item;
    __syncthreads();
    if (it >= sh.n_items) break;
    const MfmaItem item = items[it];
    const MfmaJob job = jobs[item.job];
    const int tile_x = item.tile % job.tiles_x;
    const int xa0 = 1 + tile_x * TX;  // first |x| of the tile
    const int y0 = (item.tile / job.tiles_x) * TY;
    const float4 bi = job.bininfo ? job.bininfo[item.bin] : job.bin0;
    const float hh_ref = sh.ref_hi ? bi.y : bi.x;
    // Particles [p0, p1) of the item, in stages of 32 (the last one may be partly empty:
    // its missing particles are zero packets, whose B operands are 0).
    const int p0 = item.s0 * kStep, p1 = item.s1 * kStep;
    const int n_stages = (p1 - p0 + kStage - 1) / kStage;

    if (wave < kProducerWaves) {
      // ---- producer: operands of stage k+1 while the consumers multiply stage k ----
      const int pp = lane & 15;  // particles 2pp, 2pp+1 of the stage
      // 0..15: A rows gg + 16m (m < TX / 16) and B rows gg + 16m (m < TY / 16).  A wave's four
      // row groups are 4 apart, so
      // its stores (rows 20 words apart) hit each of the 32 LDS banks exactly twice; row
      // groups 1 apart would collide 3-way.
      const int gg = (lane >> 4) * 4 + wave;
      // Row gg + 16m: phase = P0 + m * 16 * (phase step), exact in 32-bit modular arithmetic.
      const uint32_t xa_first = static_cast<uint32_t>(xa0 + gg);
      const uint32_t y_first = static_cast<uint32_t>(y0 + gg);
      constexpr int MA = TX / 16, MB = TY / 16;
      float ca[MA], cb[MB];
#pragma unroll
      for (int m = 0; m < MA; ++m) {
        const float fu = sh.sx * static_cast<float>(static_cast<int>(xa_first) + 16 * m);
        ca[m] = sh.sigma_pi_log2e * fu * fu;
      }
#pragma unroll
      for (int m = 0; m < MB; ++m) {
        const float fv = sh.sy * static_cast<float>(static_cast<int>(y_first) + 16 * m);
        cb[m] = sh.sigma_pi_log2e * fv * fv;
      }
      // x = 0: A(0, i) = 1, so S(0, y) = E(0, y) * sum_i B(i, y), summed here in FP32 by the
      // tiles of the first |x| range (cx = 0: an empty half plane).
      const bool zero_col = tile_x == 0 && sh.cx > 0;
      float zr[MB], zi[MB];
#pragma unroll
      for (int m = 0; m < MB; ++m) zr[m] = zi[m] = 0.0f;
      // Loads without a branch, clamped into the item: a conditional load would make the
      // compiler wait for it at the branch's join, right after issuing it.  Slots past p1
      // are zeroed through their weight in generate.
      auto fetch = [&](int k, Packet* q) {
        const int p = p0 + k * kStage + 2 * pp;
        q[0] = GAUSS ? load_packet(job.packets + min(p, p1 - 1)) : load_packet_puw(job.packets + min(p, p1 - 1));
        q[1] = GAUSS ? load_packet(job.packets + min(p + 1, p1 - 1))
                     : load_packet_puw(job.packets + min(p + 1, p1 - 1));
      };
      auto generate = [&](const Packet* q, int k, auto zero) {
        const int buf = k & 1;
        const int left = p1 - (p0 + k * kStage + 2 * pp);  // particles of the item from this lane's first
        float dh[2], wq[2];
        uint32_t PA[2], PB[2], dA[2], dB[2];
#pragma unroll
        for (int u = 0; u < 2; ++u) {
          // h_ref is the bin edge that keeps the exponents <= 0 (|alpha|, |beta| <= 1).
          const float d = q[u].h * q[u].h - hh_ref;
          dh[u] = sh.ref_hi ? fminf(d, 0.0f) : fmaxf(d, 0.0f);
          wq[u] = u < left ? q[u].w * bi.w : 0.0f;
          PA[u] = xa_first * q[u].pu + kPhaseRound;
          PB[u] = y_first * q[u].pv + kPhaseRound;
          dA[u] = q[u].pu << 4;
          dB[u] = q[u].pv << 4;
        }
#pragma unroll
        for (int m = 0; m < MA; ++m) {  // A(|x|, i)
          float re[2], im[2];
#pragma unroll
          for (int u = 0; u < 2; ++u) {
            const float ph = phase_rev(PA[u]);
            PA[u] += dA[u];
            re[u] = cos_rev(ph);
            im[u] = sin_rev(ph);
            if (GAUSS) {
              const float amp = __builtin_amdgcn_exp2f(ca[m] * dh[u]);
              re[u] *= amp;
              im[u] *= amp;
            }
          }
          store_pair<OP, SPLIT, TX * kRowPitch>(ldsA[buf], (gg + 16 * m) * kRowPitch + pp, re, im);
        }
#pragma unroll
        for (int m = 0; m < MB; ++m) {  // B(i, y)
          float re[2], im[2];
#pragma unroll
          for (int u = 0; u < 2; ++u) {
            const float ph = phase_rev(PB[u]);
            PB[u] += dB[u];
            float amp = wq[u];
            if (GAUSS) amp *= __builtin_amdgcn_exp2f(cb[m] * dh[u]);
            re[u] = amp * cos_rev(ph);
            im[u] = amp * sin_rev(ph);
          }
          if (decltype(zero)::value) {
            zr[m] += re[0] + re[1];
            zi[m] += im[0] + im[1];
          }
          store_pair<OP, SPLIT, TY * kRowPitch>(ldsB[buf], (gg + 16 * m) * kRowPitch + pp, re, im);
        }
      };
      // Packets are loaded one stage ahead of their use, so the load's latency
      // hides behind a whole stage of operand generation.  n_stages + 1 barriers, as
      // the consumers.
      auto produce = [&](auto zero) {
        if (n_stages > 0) {
          Packet q0[2], qn[2];
          fetch(0, q0);
          fetch(min(1, n_stages - 1), qn);
          generate(q0, 0, zero);
          __syncthreads();
          for (int k = 1; k < n_stages; ++k) {  // stage k while the consumers multiply k - 1
            const Packet q[2] = {qn[0], qn[1]};
            fetch(min(k + 1, n_stages - 1), qn);
            generate(q, k, zero);
            __syncthreads();
          }
        }
        __syncthreads();
      };
      if (zero_col)
        produce(std::true_type{});
      else
        produce(std::false_type{});
      if (zero_col) {
        // Sum over the 16 lanes of a row group (pp = 0..15), then one lane per row writes.
#pragma unroll
        for (int m = 0; m < MB; ++m) {
#pragma unroll
          for (int d = 1; d < 16; d <<= 1) {
            zr[m] += __shfl_xor(zr[m], d);
            zi[m] += __shfl_xor(zi[m], d);
          }
          const int y = static_cast<int>(y_first) + 16 * m;
          if (pp == 0 && y < sh.cy) {
            const float fv = sh.sy * static_cast<float>(y);
            float e;
            if (GAUSS)
              e = __builtin_amdgcn_exp2f(sh.sigma_pi_log2e * hh_ref * fv * fv) * bi.z;
            else
              e = phi<BASIS, MODE>(fabsf(fv) * job.radius) * bi.z;
            job.partial[(static_cast<size_t>(item.chunk) * two_cx + (sh.cx - 1)) * sh.cy + y] =
                make_float2(e * zr[m], e * zi[m]);
          }
        }
      }
      continue;
    }

    // ---- consumer: 32 |x| rows x 64 y columns; real products rr, ii, ri, ir ----
    const int c = wave - kProducerWaves;  // 0 .. (TX / 32) (TY / 64) - 1
    const int wx = c % (TX / 32);
    const int wy = c / (TX / 32);
    f32x16_t acc[4][2];  // [Ar.Br, Ai.Bi, Ar.Bi, Ai.Br][column block]
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[q][0] = acc[q][1] = (f32x16_t){};

    __syncthreads();
    for (int k = 0; k < n_stages; ++k) {
      const int buf = k & 1;
#pragma unroll
      for (int kk = 0; kk < kStage / 16; ++kk) {
        const int pofs = kk * 8 + (lane >> 5) * 4;
        const int row = wx * 32 + (lane & 31);
        const u32x4_t ar = *reinterpret_cast<const u32x4_t*>(&ldsA[buf][0][row * kRowPitch + pofs]);
        const u32x4_t ai = *reinterpret_cast<const u32x4_t*>(&ldsA[buf][2][row * kRowPitch + pofs]);
        u32x4_t arl, ail;
        if (SPLIT) {
          arl = *reinterpret_cast<const u32x4_t*>(&ldsA[buf][1][row * kRowPitch + pofs]);
          ail = *reinterpret_cast<const u32x4_t*>(&ldsA[buf][3][row * kRowPitch + pofs]);
        }
#pragma unroll
        for (int cb = 0; cb < 2; ++cb) {
          const int col = wy * 64 + cb * 32 + (lane & 31);
          const u32x4_t br = *reinterpret_cast<const u32x4_t*>(&ldsB[buf][0][col * kRowPitch + pofs]);
          const u32x4_t bim = *reinterpret_cast<const u32x4_t*>(&ldsB[buf][2][col * kRowPitch + pofs]);
          acc[0][cb] = mfma<OP>(ar, br, acc[0][cb]);
          acc[1][cb] = mfma<OP>(ai, bim, acc[1][cb]);
          acc[2][cb] = mfma<OP>(ar, bim, acc[2][cb]);
          acc[3][cb] = mfma<OP>(ai, br, acc[3][cb]);
          if (SPLIT) {
            const u32x4_t brl = *reinterpret_cast<const u32x4_t*>(&ldsB[buf][1][col * kRowPitch + pofs]);
            const u32x4_t bil = *reinterpret_cast<const u32x4_t*>(&ldsB[buf][3][col * kRowPitch + pofs]);
            acc[0][cb] = mfma<OP>(ar, brl, acc[0][cb]);
            acc[1][cb] = mfma<OP>(ai, bil, acc[1][cb]);
            acc[2][cb] = mfma<OP>(ar, bil, acc[2][cb]);
            acc[3][cb] = mfma<OP>(ai, brl, acc[3][cb]);
            acc[0][cb] = mfma<OP>(arl, br, acc[0][cb]);
            acc[1][cb] = mfma<OP>(ail, bim, acc[1][cb]);
            acc[2][cb] = mfma<OP>(arl, bim, acc[2][cb]);
            acc[3][cb] = mfma<OP>(ail, br, acc[3][cb]);
          }
        }
      }
      __syncthreads();
    }

    // Epilogue: A(-x, i) = conj(A(x, i)), so with P = sum A B split into its real products
    //   S(+x, y) = E (Prr - Pii, Pri + Pir),  S(-x, y) = E (Prr + Pii, Pri - Pir),
    // E = exp(s pi h_b^2 r^2) W_b (Gaussian) or Phi(r h) W (uniform radius); partial[chunk][j][y]
    // with j = x + cx - 1.  x = -cx is outside the half plane.
#pragma unroll
    for (int cb = 0; cb < 2; ++cb) {
      const int y = y0 + wy * 64 + cb * 32 + (lane & 31);
      if (y >= sh.cy) continue;
      const float fv = sh.sy * static_cast<float>(y);
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        const int xa = xa0 + wx * 32 + (i >> 2) * 8 + (lane >> 5) * 4 + (i & 3);
        if (xa > sh.cx) continue;
        const float fu = sh.sx * static_cast<float>(xa);
        const float r2 = fu * fu + fv * fv;
        float e;
        if (GAUSS)
          e = __builtin_amdgcn_exp2f(sh.sigma_pi_log2e * hh_ref * r2) * bi.z;
        else
          e = phi<BASIS, MODE>(sqrtf(r2) * job.radius) * bi.z;
        const float prr = acc[0][cb][i], pii = acc[1][cb][i], pri = acc[2][cb][i], pir = acc[3][cb][i];
        float2* col = job.partial + static_cast<size_t>(item.chunk) * two_cx * sh.cy + y;
        col[static_cast<size_t>(xa + sh.cx - 1) * sh.cy] = make_float2(e * (prr - pii), e * (pri + pir));
        if (xa < sh.cx) col[static_cast<size_t>(sh.cx - 1 - xa) * sh.cy] = make_float2(e * (prr + pii), e * (pri - pir));
      }
    }
  }
}

// acc[y*2cx + x0] (+)= sum_k partial[k][j][y] (FP64, fixed order), x0 the reference column of x = j - cx + 1.
// 64 frequencies per block, four waves sharing each frequency's chunks (sum_chunks_block).
__global__ __launch_bounds__(256) void mfma_reduce_kernel(const float2* __restrict__ partial, int nchunks, int cx,
                                                          int cy, int first, double2* __restrict__ acc) {
  __shared__ double2 lds[256];
  const int f = blockIdx.x * 64 + (threadIdx.x & 63);
  const int two_cx = 2 * cx;
  const int nfreq = two_cx * cy;
  const bool valid = f < nfreq;
  const double2 sum = sum_chunks_block(partial + (valid ? f : 0), nchunks, static_cast<size_t>(nfreq), valid, lds);
  if (threadIdx.x >= 64 || !valid) return;
  double re = sum.x, im = sum.y;
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

// Zero fill by a kernel (grid-stride), so that code a caller captures into a HIP graph records
// no memset node.
__global__ void zero_ints_kernel(int* p, size_t n) {
  for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<size_t>(gridDim.x) * blockDim.x)
    p[i] = 0;
}

hipError_t launch_zero_words(void* p, size_t words, hipStream_t s) {
  if (words == 0) return hipSuccess;
  const size_t blocks = std::min<size_t>(1024, (words + 255) / 256);
  hipLaunchKernelGGL(zero_ints_kernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s, static_cast<int*>(p),
                     words);
  return hipGetLastError();
}

// ------------------------------------------------------------------ host ---
template <int OP, bool SPLIT, bool GAUSS, int BASIS, int MODE, int TX, int TY>
static void launch_sum_tile(int grid, hipStream_t s, const MfmaJob* jobs, const MfmaItem* items,
                            const MfmaShared& sh, int* counter) {
  hipLaunchKernelGGL((mfma_sum_kernel<OP, SPLIT, GAUSS, BASIS, MODE, TX, TY>), dim3(grid),
                     dim3(sum_threads(TX, TY)), 0, s, jobs, items, sh, counter);
}

// The tile variants (sum_tile) are built for the split forms; the single-FP16 form keeps the
// default tile.
template <int OP, bool SPLIT, bool GAUSS, int BASIS, int MODE>
static void launch_sum(int grid, hipStream_t s, const MfmaJob* jobs, const MfmaItem* items, const MfmaShared& sh,
                       int* counter) {
  switch (SPLIT ? sh.tile : VIS_TILE_DEFAULT) {
    case VIS_TILE_HALF_X:
      launch_sum_tile<OP, SPLIT, GAUSS, BASIS, MODE, kTileX / 2, kTileY>(grid, s, jobs, items, sh, counter);
      break;
    case VIS_TILE_HALF_Y:
      launch_sum_tile<OP, SPLIT, GAUSS, BASIS, MODE, kTileX, kTileY / 2>(grid, s, jobs, items, sh, counter);
      break;
    case VIS_TILE_QUARTER:
      launch_sum_tile<OP, SPLIT, GAUSS, BASIS, MODE, kTileX / 2, kTileY / 2>(grid, s, jobs, items, sh, counter);
      break;
    default:
      launch_sum_tile<OP, SPLIT, GAUSS, BASIS, MODE, kTileX, kTileY>(grid, s, jobs, items, sh, counter);
  }
}

template <int OP, bool SPLIT>
static void launch_sum_kind(int grid, hipStream_t s, const MfmaJob* jobs, const MfmaItem* items, const MfmaShared& sh,
                            int* counter, int kind, int basis, int mode) {
  if (kind == KIND_GAUSS) {
    launch_sum<OP, SPLIT, true, GAUSSIAN, VIS_PHI_ANALYTIC>(grid, s, jobs, items, sh, counter);
    return;
  }
  const bool an = mode == VIS_PHI_ANALYTIC;
  switch (basis) {
    case SPH:
      an ? launch_sum<OP, SPLIT, false, SPH, VIS_PHI_ANALYTIC>(grid, s, jobs, items, sh, counter)
         : launch_sum<OP, SPLIT, false, SPH, VIS_PHI_REFERENCE>(grid, s, jobs, items, sh, counter);
      break;
    case GAUSSIAN:
      an ? launch_sum<OP, SPLIT, false, GAUSSIAN, VIS_PHI_ANALYTIC>(grid, s, jobs, items, sh, counter)
         : launch_sum<OP, SPLIT, false, GAUSSIAN, VIS_PHI_REFERENCE>(grid, s, jobs, items, sh, counter);
      break;
    default:
      an ? launch_sum<OP, SPLIT, false, WENDLAND_D3_C2, VIS_PHI_ANALYTIC>(grid, s, jobs, items, sh, counter)
         : launch_sum<OP, SPLIT, false, WENDLAND_D3_C2, VIS_PHI_REFERENCE>(grid, s, jobs, items, sh, counter);
  }
}

// Extra time per stage of a two-node tile that also sums the x = 0 column, relative to the
// other tiles: measured, one render of the 262k clustered view in one round of items, their
// items took 196 against 178 us for the same steps (their producers' stage loop issues 256
// instead of 181 VALU instructions, but the producers are not all of an item's time).
constexpr double kZeroColExtra = 0.10;

// CUs of the device current at the first call (thread-safe static initialisation).
static int device_cus() {
  static const int cus = [] {
    int dev = 0, n = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0)
      n = 256;
    return n;
  }();
  return cus;
}

int mfma_kind(const GroupInfo& info, int basis) {
  if (info.uniform_radius) return KIND_UNIFORM;
  return basis == GAUSSIAN ? KIND_GAUSS : KIND_BINNED;
}

// The h-bins (count nb) and interpolation nodes (T) a binned kind takes under a config; false
// (error recorded) on an invalid setting.
static bool bin_setting(const GroupInfo& info, int kind, const VisConfig* c, int* nb_out, int* T_out) {
  int nb, T = 1;
  if (kind == KIND_GAUSS) {
    // Default ratio 2: the bins only keep alpha, beta <= 1 in FP16 range, and a factor small
    // enough to leave it belongs to a negligible term; fewer bins mean fewer bin-aligned work
    // items (the 4M-particle log-normal view: 19.8 ms at 1.25, 19.3 ms at 2, same accuracy).
    const double ratio = c->h_bin_ratio > 1.0f ? c->h_bin_ratio : 2.0;
    nb = c->number_of_h_bins > 0 ? c->number_of_h_bins : auto_bins(info.h_min, info.h_max, ratio);
  } else {
    T = c->number_of_interpolation_nodes > 0 ? c->number_of_interpolation_nodes : 2;
    if (T < 2 || T > kMaxNodes) {
      set_error(VIS_ERROR_INVALID_CONFIG, "number_of_interpolation_nodes must be 0 (auto) or 2..5");
      return false;
    }
    // Defaults: 2 Chebyshev nodes per bin of ratio 1.01 (the two-node kernel; interpolation
    // error < 1.2e-5 of Phi(0) pointwise), 3 nodes: 1.04 (< 3e-6), more: 1.1.  A group of fewer
    // than kFewParticles particles takes two-node bins of ratio 1.0025 (16 times smaller error):
    // with a few particles per bin the error does not average out (2 to 70 particles, every one
    // at a bin edge: 0.9-1.5e-5 against FP64 at ratio 1.01, 0.7-1.1e-6 at 1.0025), and the
    // narrower bins cost such a group nothing measurable.
    constexpr int kFewParticles = 1024;
    const double t2 = info.n < kFewParticles ? 1.0025 : 1.01;
    const double ratio = c->h_bin_ratio > 1.0f ? c->h_bin_ratio : (T == 2 ? t2 : T == 3 ? 1.04 : 1.1);
    nb = c->number_of_h_bins > 0 ? c->number_of_h_bins : auto_bins(info.h_min, info.h_max, ratio);
  }
  // Radii of 0 take a bin of their own (bin 0, width 0) ahead of the log-spaced bins of the
  // nonzero radii.
  if (info.has_zero_h) ++nb;
  if (nb > 65535) {
    set_error(VIS_ERROR_INVALID_CONFIG, "number_of_h_bins above 65535");
    return false;
  }
  *nb_out = nb;
  *T_out = T;
  return true;
}

bool mfma_prebin(const Group& g, const GroupInfo& info, const VisConfig* c, hipStream_t s) {
  if (c->form == VIS_FORM_PLAIN || !g.d_constraints || !g.h_radii || !g.h_constraints || info.n <= 0) return true;
  const int kind = mfma_kind(info, g.basis_function_id);
  if (kind == KIND_UNIFORM) return true;
  int nb, T;
  if (!bin_setting(info, kind, c, &nb, &T)) return false;
  return bin_cache_from_host(g, info, info.n, nb, T, s) != nullptr;
}

// Bins, packet slots and work items of one job (host side; builds the group's
// bins on first use).
bool mfma_plan(const MfmaJobSpec& spec, hipStream_t s, long long target_steps, MfmaPlanned* out) {
  const VisConfig* c = spec.c;
  const GroupInfo& info = spec.info;
  out->kind = mfma_kind(info, spec.vp.basis);
  out->basis = spec.vp.basis;
  out->T = 1;
  out->bc = nullptr;
  out->items.clear();
  MfmaJob& job = out->job;
  job = MfmaJob{};
  job.radius = info.has_radii ? info.radius : 1.0f;
  std::vector<int> bin_step;
  if (out->kind == KIND_UNIFORM) {
    out->n_slots = std::max(kStep, ((spec.n + kStep - 1) / kStep) * kStep);
    job.bin0 = make_float4(0.0f, 0.0f, info.w_absmax, info.w_absmax > 0.0f ? 1.0f / info.w_absmax : 0.0f);
    bin_step = {0, out->n_slots / kStep};
  } else {
    int nb, T;
    if (!bin_setting(info, out->kind, c, &nb, &T)) return false;
    const BinCache* bc = bin_cache(*spec.g, info, spec.n, nb, T, s);
    if (!bc) return false;
    out->bc = bc;
    out->T = T;
    out->n_slots = bc->n_slots;
    job.bininfo = bc->bininfo.as<float4>();
    job.step_bin = bc->step_bin.as<uint16_t>();
    job.nodes = bc->nodes.as<float>();
    job.lag = bc->lag.as<float4>();
    bin_step = bc->bin_step;
  }
  // Work items.  Sum kernel: bin-aligned step ranges of at most kMaxChunkSteps; binned
  // kernel: ranges may cross bins (it folds at bin changes).  Sized so that a single
  // render still fills the chip several times over.
  const bool binned = out->kind == KIND_BINNED;
  // Sum kernel tiles cover |x| in [1, cx] (each row gives x and -x; x = 0 comes with the
  // first range); binned kernel tiles (T >= 3) cover j = x + cx - 1 in [0, 2cx).
  // Two-node binned kernel: |x| rows as the sum kernel.
  const bool paired = binned && out->T == 2;
  // The tile variant: split forms only (the single-FP16 form and the T >= 3 kernel keep theirs).
  const bool split = c->form != VIS_FORM_MFMA_F16;
  out->tile = split && (!binned || paired) ? c->tile : VIS_TILE_DEFAULT;
  int tx, ty;
  if (!binned)
    tile_dims(out->tile, kTileX, kTileY, &tx, &ty);
  else if (paired)
    tile_dims(out->tile, kPairTileX, kPairTileY, &tx, &ty);
  else
    tx = kBinnedTileX, ty = kBinnedTileY;
  job.tiles_x = !binned || paired ? std::max(1, (spec.vp.cx + tx - 1) / tx) : (2 * spec.vp.cx + tx - 1) / tx;
  const int tiles = job.tiles_x * ((spec.vp.cy + ty - 1) / ty);
  const long long total_steps = bin_step.back();
  const int cap = binned ? 4096 : kMaxChunkSteps;
  const int floor_steps = binned ? 128 : 32;
  out->tile_steps = total_steps * tiles;
  // One render of the two-node kernel: the tiles of the first |x| range, whose producers also
  // sum the x = 0 column (kZeroColExtra more work per stage), get proportionally more and
  // shorter step ranges than the other tiles, so that one round of items ends together.
  const int zero_tiles = tiles / job.tiles_x;  // tiles t with t % tiles_x == 0
  const bool split_zero = target_steps <= 0 && paired && job.tiles_x > 1 && spec.vp.cx > 0;
  long long item_budget = 0;  // one render: the items of its rounds (rounds x CUs)
  if (target_steps <= 0) {
    // One render: items of equal size whose count is a whole number (1 or 2) of rounds over
    // the CUs - the atomic queue hands equal items out in rounds, and a last round of a few
    // items leaves the other CUs idle (a plan of 2.03 rounds takes 3).  Few rounds: every
    // item writes a partial spectrum tile that the reduction reads back (the 1M-particle
    // uniform view: 8 rounds 1.72 ms per view, 1.41 of them in the MFMA kernel; 2 rounds
    // 1.48 ms, 1.38 in the kernel).  Binned kernels take one round: an item also opens and
    // folds its bins, and their FP32 sums are folded per bin, so longer items cost no accuracy
    // (the 262k clustered view: 2.81e13 -> 2.97e13 terms/s); the sum kernel keeps two, its FP32
    // accumulations lasting a whole item (one round: the 1M view 1% faster, 2.2e-6 -> 3.2e-6).
    const int cus = device_cus();
    const double w = static_cast<double>(total_steps) * tiles;
    const long long max_rounds = binned ? 1 : 2;
    const long long rounds =
        std::max(1LL, std::min(max_rounds, std::llround(w / (static_cast<double>(floor_steps) * cus))));
    item_budget = rounds * cus;
    const long long chunks =
        split_zero ? std::max(1LL, std::min<long long>(std::llround(rounds * cus / (tiles + zero_tiles * kZeroColExtra)),
                                                      rounds * cus / tiles))
                   : std::max(1LL, (rounds * cus + tiles - 1) / tiles);
    target_steps = (total_steps + chunks - 1) / chunks;
  }
  int chunk_steps = static_cast<int>(std::min<long long>(cap, std::max<long long>(floor_steps, target_steps)));
  chunk_steps += chunk_steps & 1;  // even: items of the two-node kernel are whole 32-particle stages
  int chunk = 0;
  auto add_item = [&](int t, int bin, int s0, int s1, int cost, int ck) {
    MfmaItem it{};
    it.tile = t;
    it.s0 = s0;
    it.s1 = s1;
    it.bin = bin;
    it.chunk = ck;
    it.cost = cost;
    out->items.push_back(it);
  };
  auto add_range = [&](int bin, int s0, int s1, int cost) {
    for (int t = 0; t < tiles; ++t) {
      MfmaItem it{};
      it.tile = t;
      it.s0 = s0;
      it.s1 = s1;
      it.bin = bin;
      it.chunk = chunk;
      // Tiles of the first |x| range also sum the x = 0 column (binned kernels: about a
      // quarter more generation work per stage): queued first, so that no workgroup gets
      // two of them in a row while another gets none.
      it.cost = t % job.tiles_x != 0 ? cost : binned ? cost + cost / 4 : cost + cost / 16;
      out->items.push_back(it);
    }
    ++chunk;
  };
  if (binned) {
    // Ranges of equal cost rather than equal length: a step costs 1 and every bin start
    // kFoldCost more (the fold of the accumulators and of the x = 0 column at a bin change,
    // and the stage it holds up), so the ranges over the sparse tails of the h distribution,
    // one short bin after another, are not the stragglers of the launch.  The value is
    // measured: a least-squares fit of the item durations of one render of the 262k-particle
    // clustered view gives 1.11 us per 32-particle stage and 2.74 us per fold, about 5 steps.
    constexpr double kFoldCost = 5.0;
    std::vector<double> cum(static_cast<size_t>(total_steps) + 1, 0.0);
    std::vector<char> starts(static_cast<size_t>(total_steps) + 1, 0);
    for (size_t b = 0; b + 1 < bin_step.size(); ++b)
      if (bin_step[b] < bin_step[b + 1]) starts[bin_step[b]] = 1;
    for (long long k = 0; k < total_steps; ++k) cum[k + 1] = cum[k] + 1.0 + (starts[k] ? kFoldCost : 0.0);
    // n ranges [s0, s1) of about equal cost, in whole 32-particle stages.
    auto partition = [&](long long n) {
      std::vector<std::pair<int, int>> r;
      long long s0 = 0;
      for (long long k = 1; k <= n && s0 < total_steps; ++k) {
        long long s1 = total_steps;
        if (k < n) {
          const double goal = cum[total_steps] * static_cast<double>(k) / static_cast<double>(n);
          s1 = std::lower_bound(cum.begin(), cum.end(), goal) - cum.begin();
          s1 += s1 & 1;
          s1 = std::min(total_steps, std::max(s1, s0 + 2));
        }
        r.emplace_back(static_cast<int>(s0), static_cast<int>(s1));
        s0 = s1;
      }
      return r;
    };
    const long long n_ranges = std::max<long long>(1, (total_steps + chunk_steps - 1) / chunk_steps);
    auto range_cost = [&](const std::pair<int, int>& r) { return static_cast<int>(cum[r.second] - cum[r.first] + 0.5); };
    if (split_zero) {
      // Chunk k: range r0[k] of the x = 0 tiles, r1[k] of the others (an empty item, which
      // writes zeros, where a tile has no range k: the reduction sums every chunk slot).
      // As many items as the render's rounds hold: x = 0 tiles x n0 + others x n1 <= budget.
      const long long others = tiles - zero_tiles;
      long long n0 = std::llround(n_ranges * (1.0 + kZeroColExtra));
      n0 = std::max<long long>(n_ranges, std::min<long long>(n0, (item_budget - others * n_ranges) / zero_tiles));
      const auto r1 = partition(n_ranges);
      const auto r0 = partition(n0);
      const size_t n = std::max(r0.size(), r1.size());
      for (size_t k = 0; k < n; ++k) {
        for (int t = 0; t < tiles; ++t) {
          const auto& r = t % job.tiles_x == 0 ? r0 : r1;
          const double f = t % job.tiles_x == 0 ? 1.0 + kZeroColExtra : 1.0;
          if (k < r.size())
            add_item(t, 0, r[k].first, r[k].second, static_cast<int>(range_cost(r[k]) * f + 0.5), chunk);
          else
            add_item(t, 0, 0, 0, 0, chunk);
        }
        ++chunk;
      }
    } else {
      for (const auto& r : partition(n_ranges)) add_range(0, r.first, r.second, range_cost(r));
    }
  } else {
    for (size_t b = 0; b + 1 < bin_step.size(); ++b)
      for (int s0 = bin_step[b]; s0 < bin_step[b + 1]; s0 += chunk_steps)
        add_range(static_cast<int>(b), s0, std::min(bin_step[b + 1], s0 + chunk_steps),
                  std::min(bin_step[b + 1], s0 + chunk_steps) - s0);
  }
  if (chunk == 0) add_range(0, 0, 0, 0);
  out->item_chunks = chunk;
  out->slots_per_chunk = paired ? 2 : 1;  // the two-node kernel writes one slot per node
  out->nchunks = out->slots_per_chunk * chunk;
  out->pass_chunks = chunk;
  return true;
}

static std::atomic<long long> g_partial_budget{kDefaultPartialBudget};
void mfma_set_partial_budget(long long bytes) { g_partial_budget = bytes > 0 ? bytes : kDefaultPartialBudget; }

// Host half of a batch: plans, buffer sizes, work tables uploaded (synchronous).
static std::vector<unsigned long long> plan_key(const std::vector<MfmaJobSpec>& specs) {
  std::vector<unsigned long long> k;
  k.push_back(specs.size());
  k.push_back(static_cast<unsigned long long>(g_partial_budget.load()));
  for (const MfmaJobSpec& sp : specs) {
    uint32_t ratio;
    std::memcpy(&ratio, &sp.c->h_bin_ratio, sizeof(ratio));
    for (unsigned long long v :
         {sp.info.serial, static_cast<unsigned long long>(reinterpret_cast<uintptr_t>(sp.g->d_constraints)),
          static_cast<unsigned long long>(sp.n), static_cast<unsigned long long>(sp.vp.cx),
          static_cast<unsigned long long>(sp.vp.cy), static_cast<unsigned long long>(sp.vp.basis),
          static_cast<unsigned long long>(sp.c->number_of_h_bins),
          static_cast<unsigned long long>(sp.c->number_of_interpolation_nodes), static_cast<unsigned long long>(ratio),
          static_cast<unsigned long long>(sp.c->tile), static_cast<unsigned long long>(sp.c->form)})
      k.push_back(v);
  }
  return k;
}

bool mfma_prepare(const std::vector<MfmaJobSpec>& specs, hipStream_t s, MfmaWorkspace* ws, MfmaBatch* bt) {
  std::vector<unsigned long long> key = plan_key(specs);
  if (bt == &ws->batch && !ws->plan_key.empty() && key == ws->plan_key) return true;  // tables already uploaded
  ws->plan_key.clear();
  bt->plans.assign(specs.size(), MfmaPlanned{});
  bt->keys.clear();
  bt->job_begin.clear();
  bt->item_begin.clear();
  bt->passes = 1;
  if (specs.empty()) return true;
  // Items per batch: about 16 per CU over the whole batch, so a batch of many small
  // renders is not split more finely than it needs to be.
  // A plan does not depend on the view: jobs of one group under equal settings (the views of a
  // timestep) share it, so each distinct one is planned once.
  auto same_plan = [](const MfmaJobSpec& a, const MfmaJobSpec& b) {
    return a.g->d_constraints == b.g->d_constraints && a.info.serial == b.info.serial && a.n == b.n &&
           a.vp.cx == b.vp.cx && a.vp.cy == b.vp.cy && a.vp.basis == b.vp.basis &&
           a.c->number_of_h_bins == b.c->number_of_h_bins &&
           a.c->number_of_interpolation_nodes == b.c->number_of_interpolation_nodes &&
           a.c->h_bin_ratio == b.c->h_bin_ratio && a.c->tile == b.c->tile && a.c->form == b.c->form;
  };
  std::vector<size_t> first(specs.size());
  for (size_t k = 0; k < specs.size(); ++k) {
    first[k] = k;
    for (size_t j = 0; j < k; ++j)
      if (first[j] == j && same_plan(specs[j], specs[k])) {
        first[k] = j;
        break;
      }
  }
  long long work = 0;
  for (size_t k = 0; k < specs.size(); ++k) {
    if (first[k] != k)
      bt->plans[k] = bt->plans[first[k]];
    else if (!mfma_plan(specs[k], s, 0, &bt->plans[k]))
      return false;
    work += bt->plans[k].tile_steps;
  }
  const long long target = std::max<long long>(1, work / (16LL * device_cus()));
  size_t slots = 0, full = 0;
  for (size_t k = 0; k < specs.size(); ++k) {
    if (specs.size() > 1) {
      if (first[k] != k)
        bt->plans[k] = bt->plans[first[k]];
      else if (!mfma_plan(specs[k], s, target, &bt->plans[k]))
        return false;
    }
    bt->plans[k].pk_off = slots;
    slots += bt->plans[k].n_slots;
    full += static_cast<size_t>(bt->plans[k].nchunks) * 2 * specs[k].vp.cx * specs[k].vp.cy;
  }
  // The partial slots grow with the particle count (one spectrum per 16,384 particles of the sum
  // kernel): past the budget the batch runs in passes, each over a 1/passes share of every job's
  // chunks (at least one), its slots reduced into the FP64 accumulators before the next pass.
  const size_t budget = std::max<size_t>(1, static_cast<size_t>(g_partial_budget.load()) / sizeof(float2));
  int passes = 1;
  if (full > budget) {
    const size_t want = (full + budget - 1) / budget;
    int most = 1;
    for (const MfmaPlanned& pl : bt->plans) most = std::max(most, pl.item_chunks);
    passes = static_cast<int>(std::min<size_t>(want, static_cast<size_t>(most)));
  }
  size_t parts = 0;
  int used = 1;
  for (size_t k = 0; k < specs.size(); ++k) {
    MfmaPlanned& pl = bt->plans[k];
    pl.pass_chunks = (pl.item_chunks + passes - 1) / passes;
    used = std::max(used, (pl.item_chunks + pl.pass_chunks - 1) / pl.pass_chunks);
    pl.part_off = parts;
    parts += static_cast<size_t>(pl.pass_chunks) * pl.slots_per_chunk * 2 * specs[k].vp.cx * specs[k].vp.cy;
  }
  bt->passes = used;
  if (!ws->packets.reserve(sizeof(Packet) * slots) || !ws->partial.reserve(sizeof(float2) * std::max<size_t>(parts, 1)))
    return false;
  for (auto& pl : bt->plans) {
    pl.job.packets = ws->packets.as<Packet>() + pl.pk_off;
    pl.job.partial = ws->partial.as<float2>() + pl.part_off;
  }
  // One persistent launch per kernel variant present (kind, basis, nodes, tile).
  auto key_of = [](const MfmaPlanned& pl) {
    return std::make_tuple(pl.kind, pl.kind == KIND_GAUSS ? static_cast<int>(GAUSSIAN) : pl.basis, pl.T, pl.tile);
  };
  for (auto& pl : bt->plans)
    if (std::find(bt->keys.begin(), bt->keys.end(), key_of(pl)) == bt->keys.end()) bt->keys.push_back(key_of(pl));
  std::vector<MfmaJob> jobs;
  std::vector<int> jid(bt->plans.size(), 0);  // a plan's job index within its key's jobs
  for (auto& key : bt->keys) {
    bt->job_begin.push_back(static_cast<int>(jobs.size()));
    for (size_t j = 0; j < bt->plans.size(); ++j) {
      if (key_of(bt->plans[j]) != key) continue;
      jid[j] = static_cast<int>(jobs.size()) - bt->job_begin.back();
      jobs.push_back(bt->plans[j].job);
    }
  }
  bt->job_begin.push_back(static_cast<int>(jobs.size()));
  std::vector<MfmaItem> items;
  for (int p = 0; p < bt->passes; ++p) {
    for (auto& key : bt->keys) {
      bt->item_begin.push_back(static_cast<int>(items.size()));
      std::vector<MfmaItem> its;
      for (size_t j = 0; j < bt->plans.size(); ++j) {
        const MfmaPlanned& pl = bt->plans[j];
        if (key_of(pl) != key) continue;
        const int c0 = p * pl.pass_chunks;
        for (MfmaItem it : pl.items) {
          if (it.chunk < c0 || it.chunk >= c0 + pl.pass_chunks) continue;
          it.job = jid[j];
          it.chunk -= c0;  // the pass's own slots
          its.push_back(it);
        }
      }
      // Costliest items first: the atomic queue then balances the tail.
      std::stable_sort(its.begin(), its.end(), [](const MfmaItem& a, const MfmaItem& b) { return a.cost > b.cost; });
      items.insert(items.end(), its.begin(), its.end());
    }
  }
  bt->item_begin.push_back(static_cast<int>(items.size()));
  const size_t jbytes = sizeof(MfmaJob) * jobs.size(), ibytes = sizeof(MfmaItem) * items.size();
  // The previous upload (and the launches reading the previous tables) must be done before
  // the host copy and the device tables are reused.
  if (!hip_ok(hipStreamSynchronize(s), "synchronize")) return false;
  if (!ws->tables.reserve(jbytes + ibytes) || !ws->counters.reserve(sizeof(int) * bt->passes * bt->keys.size()))
    return false;
  ws->tables_host.resize(jbytes + ibytes);
  std::memcpy(ws->tables_host.data(), jobs.data(), jbytes);
  std::memcpy(ws->tables_host.data() + jbytes, items.data(), ibytes);
  // Stream-ordered upload: earlier launches on the stream have read the old tables by the time
  // it runs.  The host copy stays untouched until the next re-plan, which first waits for it.
  if (!hip_ok(hipMemcpyAsync(ws->tables.ptr, ws->tables_host.data(), jbytes + ibytes, hipMemcpyHostToDevice, s),
              "upload work tables"))
    return false;
  bt->jobs = reinterpret_cast<const MfmaJob*>(ws->tables.ptr);
  bt->items = reinterpret_cast<const MfmaItem*>(static_cast<char*>(ws->tables.ptr) + jbytes);
  bt->counters = ws->counters.as<int>();
  if (bt == &ws->batch) ws->plan_key = std::move(key);
  return true;
}

const Constraint* mfma_ordered(const MfmaPlanned& pl) { return pl.bc ? pl.bc->ordered.as<Constraint>() : nullptr; }
const float* mfma_ordered_radii(const MfmaPlanned& pl) { return pl.bc ? pl.bc->ordered_r.as<float>() : nullptr; }

// Device half: kernel launches only (safe to capture into a HIP graph).
bool mfma_enqueue(const std::vector<MfmaJobSpec>& specs, const MfmaBatch& bt, hipStream_t s, EventPair* ev) {
  if (specs.empty()) return true;
  // The first packet prep also clears the work-queue counters (a prep launch that clears
  // counters runs at least one block), one launch fewer per render.
  const int ncounters = bt.passes * static_cast<int>(bt.keys.size());
  const bool fused = ncounters <= 256;  // one block clears them
  for (size_t k = 0; k < specs.size(); ++k) {
    const MfmaJobSpec& sp = specs[k];
    const MfmaPlanned& pl = bt.plans[k];
    int* zero = k == 0 && fused ? bt.counters : nullptr;
    const int nzero = k == 0 && fused ? ncounters : 0;
    hipError_t e;
    if (pl.bc)  // the binned particles in slot order: every slot read in place
      e = launch_prep_packets(pl.bc->ordered.as<Constraint>(), pl.bc->ordered_r.as<float>(), pl.n_slots, pl.n_slots,
                              sp.c->step_size.x, sp.c->step_size.y, sp.u, sp.v,
                              cull_factor(sp.c, sp.g->basis_function_id), const_cast<Packet*>(pl.job.packets), s,
                              zero, nzero);
    else
      e = launch_prep_packets(sp.g->d_constraints, sp.g->d_radii, sp.n, pl.n_slots, sp.c->step_size.x,
                              sp.c->step_size.y, sp.u, sp.v, cull_factor(sp.c, sp.g->basis_function_id),
                              const_cast<Packet*>(pl.job.packets), s, zero, nzero);
    if (!hip_ok(e, "prep_packets launch")) return false;
  }
  // In passes, the timed span (ev) runs from the first pass's kernels to the last pass's and
  // includes the reductions between them.
  const bool one = bt.passes == 1;
  if (ev && !one && !hip_ok(hipEventRecord(ev->first, s), "hipEventRecord")) return false;
  for (int p = 0; p < bt.passes; ++p) {
    if (!mfma_launch_sums(specs, bt, s, one ? ev : nullptr, fused, p)) return false;
    if (ev && !one && p + 1 == bt.passes && !hip_ok(hipEventRecord(ev->second, s), "hipEventRecord")) return false;
    for (size_t k = 0; k < specs.size(); ++k) {
      const MfmaPlanned& pl = bt.plans[k];
      const int chunks = std::min(pl.pass_chunks, pl.item_chunks - p * pl.pass_chunks);
      if (chunks <= 0) continue;  // this job's chunks all ran in earlier passes
      const int nfreq = 2 * specs[k].vp.cx * specs[k].vp.cy;
      const bool assign = p == 0 && specs[k].first;
      hipLaunchKernelGGL(mfma_reduce_kernel, dim3((nfreq + 63) / 64), dim3(256), 0, s, pl.job.partial,
                         chunks * pl.slots_per_chunk, specs[k].vp.cx, specs[k].vp.cy, assign ? 1 : 0, specs[k].acc);
      if (!hip_ok(hipGetLastError(), "mfma_reduce launch")) return false;
    }
  }
  return true;
}

bool mfma_launch_sums(const std::vector<MfmaJobSpec>& specs, const MfmaBatch& bt, hipStream_t s, EventPair* ev,
                      bool counters_cleared, int pass) {
  if (specs.empty()) return true;
  if (pass < 0 || pass >= bt.passes) return hip_ok(hipErrorInvalidValue, "mfma launch: pass out of range");
  const VisConfig* c0 = specs[0].c;
  const size_t nkeys = bt.keys.size();
  int* counters = bt.counters + static_cast<size_t>(pass) * nkeys;
  if (!counters_cleared && !hip_ok(launch_zero_words(counters, nkeys, s), "zero counters")) return false;

  MfmaShared sh{};
  sh.cx = specs[0].vp.cx;
  sh.cy = specs[0].vp.cy;
  sh.sx = specs[0].vp.sx;
  sh.sy = specs[0].vp.sy;
  const int mode = specs[0].vp.phi_mode;
  const float sigma = (mode == VIS_PHI_REFERENCE) ? 1.0f : -1.0f;
  sh.ref_hi = sigma > 0.0f ? 1 : 0;
  if (ev && !hip_ok(hipEventRecord(ev->first, s), "hipEventRecord")) return false;
  for (size_t k = 0; k < nkeys; ++k) {
    const int kind = std::get<0>(bt.keys[k]), basis = std::get<1>(bt.keys[k]), T = std::get<2>(bt.keys[k]);
    const size_t ib = static_cast<size_t>(pass) * nkeys + k;
    sh.tile = std::get<3>(bt.keys[k]);
    sh.n_items = bt.item_begin[ib + 1] - bt.item_begin[ib];
    if (sh.n_items == 0) continue;  // (a later pass, every job of this variant done)
    sh.sigma_pi_log2e = kind == KIND_GAUSS ? sigma * kPi * kLog2e : 0.0f;
    const int grid = std::max(1, std::min(sh.n_items, device_cus()));
    const MfmaJob* kj = bt.jobs + bt.job_begin[k];
    const MfmaItem* ki = bt.items + bt.item_begin[ib];
    int* ctr = counters + k;
    if (kind == KIND_BINNED) {
      if (!launch_binned(grid, s, kj, ki, sh, ctr, c0->form, T, basis, mode)) return false;
    } else if (c0->form == VIS_FORM_MFMA_F16) {
      launch_sum_kind<OP_F16, false>(grid, s, kj, ki, sh, ctr, kind, basis, mode);
    } else if (c0->form == VIS_FORM_MFMA_SPLIT_BF16) {
      launch_sum_kind<OP_BF16, true>(grid, s, kj, ki, sh, ctr, kind, basis, mode);
    } else {
      launch_sum_kind<OP_F16, true>(grid, s, kj, ki, sh, ctr, kind, basis, mode);
    }
    if (!hip_ok(hipGetLastError(), "mfma launch")) return false;
  }
  if (ev && !hip_ok(hipEventRecord(ev->second, s), "hipEventRecord")) return false;
  return true;
}

bool mfma_execute(const std::vector<MfmaJobSpec>& specs, hipStream_t s, MfmaWorkspace* ws, EventPair* ev) {
  return mfma_prepare(specs, s, ws, &ws->batch) && mfma_enqueue(specs, ws->batch, s, ev);
}

}  // namespace mv
