// MFMA form for SPH / Wendland groups with per-particle radii.
//
// Phi(r h_i) does not factor into x and y parts for these bases, so the
// particles are sorted into log-spaced bins of h (once per registration) and,
// inside bin b, Phi is interpolated in h through T Chebyshev nodes h_bt:
//
//     Phi(r h_i) ~= sum_t l_bt(h_i) Phi(r h_bt)          (l_bt: Lagrange basis)
//     S(x, y)    ~= sum_b sum_t Phi(r h_bt) * sum_{i in b} A(x, i) B_t(i, y)
//     A(x, i) = exp(+2 pi i x a_i),  B_t(i, y) = (w_i / W_b) l_bt(h_i) exp(+2 pi i y b_i)
//
// (l_bt(h_i) / W_b is computed once per binning, in FP64 on the host, and kept
// per particle slot)
// with W_b the largest |w| of bin b, folded back in with the node factors
// (weights correlate with h, so per-bin scaling keeps the FP16 operands away
// from the subnormal range);
// i.e. one complex GEMM per (bin, node) sharing the A operand; the node factors
// Phi(r h_bt) are applied when the accumulators are folded (every 1024
// particles and at every bin change).  The interpolation error per term is
// bounded by max|Phi^(T)| (r dh)^T / (T! 2^(2T-1)); with bins of ratio 1.1 and
// T = 4 it is below 2e-6 of Phi(0).
//
// Kernel: 8 waves, 128 (x) x 64 (y) complex tile per workgroup, 32 x 32 per
// wave, T accumulator pairs per wave; operand generation and split products
// as in mfma.hip.
#include "mfma_common.h"

#include <algorithm>
#include <cmath>
#include <map>
#include <tuple>
#include <vector>

namespace mv {

constexpr int kBFlushSteps = 64;  // fold every 1024 particles

template <int OP, bool SPLIT, int T, int BASIS, int MODE>
__global__ __launch_bounds__(kThreads, 1) void mfma_binned_kernel(const MfmaJob* __restrict__ jobs,
                                                                    const MfmaItem* __restrict__ items,
                                                                    MfmaShared sh, int* __restrict__ counter) {
  __shared__ uint32_t ldsA[2][2][kBinnedTileX * kPitch];     // [buf][hi, lo]
  __shared__ uint32_t ldsB[2][T][2][kBinnedTileY * kPitch];  // [buf][node][hi, lo]
  __shared__ int next_item;

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wx = wave & 3;   // 32-row block
  const int wy = wave >> 2;  // 32-column block
  const int gp = tid & (kStep - 1);
  const int gg = tid >> 4;  // 0..31
  const int two_cx = 2 * sh.cx;

  for (;;) {
    if (tid == 0) next_item = atomicAdd(counter, 1);
    __syncthreads();
    const int it = next_item;
    __syncthreads();
    if (it >= sh.n_items) break;
    const MfmaItem item = items[it];
    const MfmaJob job = jobs[item.job];
    const int j0 = (item.tile % job.tiles_x) * kBinnedTileX;
    const int y0 = (item.tile / job.tiles_x) * kBinnedTileY;

    // Rows gg + 32m, columns gg + 32m: phase = P0 + m * 32 * (phase step), exact
    // in 32-bit modular arithmetic.
    const uint32_t x_first = static_cast<uint32_t>(j0 + gg - sh.cx + 1);
    const uint32_t y_first = static_cast<uint32_t>(y0 + gg);
    constexpr int LW = lag_words(T);
    struct Ahead {
      Packet q;
      float4 l[LW];
    };
    auto fetch = [&](int step, Ahead* a) {
      const size_t slot = static_cast<size_t>(step) * kStep + gp;
      a->q = load_packet(job.packets + slot);
#pragma unroll
      for (int k = 0; k < LW; ++k) a->l[k] = load_float4(job.lag + slot * LW + k);
    };
    auto step_bin = [&](int step) -> int {
      return *(const __attribute__((address_space(1))) uint16_t*)(job.step_bin + step);
    };

    auto generate = [&](const Ahead& a, int buf) {
      const Packet& q = a.q;
      float l[T];
#pragma unroll
      for (int t = 0; t < T; ++t) {
        const float4 v = a.l[t >> 2];
        l[t] = (t & 3) == 0 ? v.x : (t & 3) == 1 ? v.y : (t & 3) == 2 ? v.z : v.w;
      }
      uint32_t PA = x_first * q.pu + kPhaseRound;
      const uint32_t dA = q.pu << 5;
#pragma unroll
      for (int m = 0; m < 4; ++m) {
        const float ph = phase_rev(PA);
        PA += dA;
        const float re = cos_rev(ph), im = sin_rev(ph);
        const uint32_t hi = pack2<OP>(re, im);
        const int o = (gg + 32 * m) * kPitch + gp;
        ldsA[buf][0][o] = hi;
        if (SPLIT) {
          float hr, hm;
          unpack2<OP>(hi, &hr, &hm);
          ldsA[buf][1][o] = pack2<OP>(re - hr, im - hm);
        }
      }
      uint32_t PB = y_first * q.pv + kPhaseRound;
      const uint32_t dB = q.pv << 5;
#pragma unroll
      for (int m = 0; m < 2; ++m) {
        const float ph = phase_rev(PB);
        PB += dB;
        const float re0 = q.w * cos_rev(ph), im0 = q.w * sin_rev(ph);
        const int o = (gg + 32 * m) * kPitch + gp;
#pragma unroll
        for (int t = 0; t < T; ++t) {
          const float re = l[t] * re0, im = l[t] * im0;
          const uint32_t hi = pack2<OP>(re, im);
          ldsB[buf][t][0][o] = hi;
          if (SPLIT) {
            float hr, hm;
            unpack2<OP>(hi, &hr, &hm);
            ldsB[buf][t][1][o] = pack2<OP>(re - hr, im - hm);
          }
        }
      }
    };

    f32x16_t acc[T][2], sum[2];
#pragma unroll
    for (int t = 0; t < T; ++t) acc[t][0] = acc[t][1] = (f32x16_t){};
    sum[0] = sum[1] = (f32x16_t){};

    auto consume = [&](int buf) {
#pragma unroll
      for (int kk = 0; kk < kStep / 8; ++kk) {
        const int pofs = kk * 8 + (lane >> 5) * 4;
        const int row = wx * 32 + (lane & 31);
        const int col = wy * 32 + (lane & 31);
        const u32x4_t ahi = *reinterpret_cast<const u32x4_t*>(&ldsA[buf][0][row * kPitch + pofs]);
        u32x4_t alo;
        if (SPLIT) alo = *reinterpret_cast<const u32x4_t*>(&ldsA[buf][1][row * kPitch + pofs]);
#pragma unroll
        for (int t = 0; t < T; ++t) {
          const u32x4_t bhi = *reinterpret_cast<const u32x4_t*>(&ldsB[buf][t][0][col * kPitch + pofs]);
          const u32x4_t bre_hi = negate_imag(bhi), bim_hi = swap_halves(bhi);
          acc[t][0] = mfma<OP>(ahi, bre_hi, acc[t][0]);
          acc[t][1] = mfma<OP>(ahi, bim_hi, acc[t][1]);
          if (SPLIT) {
            const u32x4_t blo = *reinterpret_cast<const u32x4_t*>(&ldsB[buf][t][1][col * kPitch + pofs]);
            acc[t][0] = mfma<OP>(ahi, negate_imag(blo), acc[t][0]);
            acc[t][1] = mfma<OP>(ahi, swap_halves(blo), acc[t][1]);
            acc[t][0] = mfma<OP>(alo, bre_hi, acc[t][0]);
            acc[t][1] = mfma<OP>(alo, bim_hi, acc[t][1]);
          }
        }
      }
    };

    const int y = y0 + wy * 32 + (lane & 31);
    const float fv = sh.sy * static_cast<float>(y);
    float rr[16];  // |k| of the lane's 16 outputs
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const int j = j0 + wx * 32 + (i >> 2) * 8 + (lane >> 5) * 4 + (i & 3);
      const float fu = sh.sx * static_cast<float>(j - sh.cx + 1);
      rr[i] = sqrtf(fu * fu + fv * fv);
    }
    // sum += W_b sum_t Phi(r h_bt) acc_t; acc = 0.
    auto fold = [&](int b) {
      const float* nd = job.nodes + b * kMaxNodes;
      const float ws = job.bininfo[b].z;
#pragma unroll
      for (int t = 0; t < T; ++t) {
        const float hn = nd[t];
#pragma unroll
        for (int i = 0; i < 16; ++i) {
          const float f = phi<BASIS, MODE>(rr[i] * hn) * ws;
          sum[0][i] = fmaf(f, acc[t][0][i], sum[0][i]);
          sum[1][i] = fmaf(f, acc[t][1][i], sum[1][i]);
          __builtin_amdgcn_sched_barrier(0);  // one output at a time: keeps the fold's registers few
        }
        acc[t][0] = acc[t][1] = (f32x16_t){};
      }
    };

    // Packets, Lagrange weights and step bins are loaded one step ahead of their
    // use, so their latency hides behind a whole step.
    Ahead an{};
    int cur = 0, bn = 0;  // bin of the first step; bin of the step after the current one
    if (item.s0 < item.s1) {
      Ahead a0;
      fetch(item.s0, &a0);
      cur = step_bin(item.s0);
      if (item.s0 + 1 < item.s1) {
        fetch(item.s0 + 1, &an);
        bn = step_bin(item.s0 + 1);
      }
      generate(a0, 0);
    }
    __syncthreads();
    int since = 0;
    int b = cur;  // bin of step s
    for (int s = item.s0; s < item.s1; ++s) {
      const int buf = (s - item.s0) & 1;
      const int b_next = bn;
      if (s + 1 < item.s1) {
        const Ahead a = an;
        if (s + 2 < item.s1) {
          fetch(s + 2, &an);
          bn = step_bin(s + 2);
        }
        generate(a, buf ^ 1);
      }
      if (b != cur) {
        fold(cur);
        cur = b;
        since = 0;
      }
      consume(buf);
      if (++since == kBFlushSteps) {
        fold(cur);
        since = 0;
      }
      b = b_next;
      __syncthreads();
    }
    if (since) fold(cur);

    if (y < sh.cy) {
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        const int j = j0 + wx * 32 + (i >> 2) * 8 + (lane >> 5) * 4 + (i & 3);
        if (j >= two_cx) continue;
        job.partial[(static_cast<size_t>(item.chunk) * two_cx + j) * sh.cy + y] = make_float2(sum[0][i], sum[1][i]);
      }
    }
  }
}

// ---- Two nodes per bin: x and -x from one |x| row -------------------------
//
// A(-x, i) = conj A(x, i), so the four real products of A = (Ar, Ai) and
// B_t = (Br, Bi) over a bin, Prr = Ar.Br, Pii = Ai.Bi, Pri = Ar.Bi, Pir = Ai.Br,
// give both S(+x) = (Prr - Pii, Pri + Pir) and S(-x) = (Prr + Pii, Pri - Pir): one
// |x| row costs the MFMAs of one x row of the kernel above and yields two.  The
// real and imaginary parts are separate LDS arrays (two particles per word), so
// the MFMA's K index is the particle.  x = 0 (A = 1) is the column sum of B_t,
// kept by the lanes that generate B.
//
// Tile: 64 |x| rows x 64 columns.  Waves 0-3 multiply node 0, waves 4-7 node 1,
// each a 32 x 32 block (4 products of accumulators, folded S(+-x) sums); at the end
// of an item the node-1 waves hand their sums to the node-0 waves through LDS.
constexpr int kPairPitch = kStep / 2 + 4;  // 8 words of 16 particles + 4 against bank conflicts

template <int OP, bool SPLIT, int BASIS, int MODE>
__global__ __launch_bounds__(kThreads, 1) void mfma_binned_pair_kernel(const MfmaJob* __restrict__ jobs,
                                                                         const MfmaItem* __restrict__ items,
                                                                         MfmaShared sh, int* __restrict__ counter) {
  constexpr int T = 2;
  // [buf][re_hi, re_lo, im_hi, im_lo] rows of A (|x|) and, per node, of B (y).
  __shared__ uint32_t ldsA[2][4][kPairTileX * kPairPitch];
  __shared__ uint32_t ldsB[2][T][4][kPairTileY * kPairPitch];
  // End of an item: the node-1 waves' folded sums, [wave & 3][component][i][lane].
  __shared__ float xchg[4 * 4 * 16 * 64];
  __shared__ int next_item;

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int node = wave >> 2;       // MFMA node of this wave
  const int wx = wave & 1;          // 32-row |x| block
  const int wy = (wave >> 1) & 1;   // 32-column block
  const int pp = tid & 7;           // generation: particles 2pp, 2pp+1 of the step
  const int gg = tid >> 3;          // generation: A row gg, B row gg (0..63)
  const int two_cx = 2 * sh.cx;

  for (;;) {
    if (tid == 0) next_item = atomicAdd(counter, 1);
    __syncthreads();
    const int it = next_item;
    __syncthreads();
    if (it >= sh.n_items) break;
    const MfmaItem item = items[it];
    const MfmaJob job = jobs[item.job];
    const int tile_x = item.tile % job.tiles_x;
    const int xa0 = 1 + tile_x * kPairTileX;
    const int y0 = (item.tile / job.tiles_x) * kPairTileY;
    const bool zero_col = tile_x == 0 && sh.cx > 0;  // (cx = 0: an empty half plane)
    const uint32_t xa_first = static_cast<uint32_t>(xa0 + gg);
    const uint32_t y_first = static_cast<uint32_t>(y0 + gg);

    struct Ahead {
      Packet q[2];
      float2 l[2];
    };
    auto fetch = [&](int step, Ahead* a) {
      const size_t slot = static_cast<size_t>(step) * kStep + 2 * pp;
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        a->q[u] = load_packet(job.packets + slot + u);
        const float4 l = load_float4(job.lag + slot + u);
        a->l[u] = make_float2(l.x, l.y);
      }
    };
    auto step_bin = [&](int step) -> int {
      return *(const __attribute__((address_space(1))) uint16_t*)(job.step_bin + step);
    };
    float zr[T], zi[T], z0r = 0.0f, z0i = 0.0f;  // x = 0: this lane's column sums of B_t; folded
#pragma unroll
    for (int t = 0; t < T; ++t) zr[t] = zi[t] = 0.0f;

    // Also returns this step's x = 0 contributions (its two particles' B_t) in gr, gi.
    auto generate = [&](const Ahead& a, int buf, float* gr, float* gi) {
      float re[2], im[2];
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        const float ph = phase_rev(xa_first * a.q[u].pu + kPhaseRound);
        re[u] = cos_rev(ph);
        im[u] = sin_rev(ph);
      }
      store_pair<OP, SPLIT, kPairTileX * kPairPitch>(ldsA[buf], gg * kPairPitch + pp, re, im);
      float re0[2], im0[2];
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        const float ph = phase_rev(y_first * a.q[u].pv + kPhaseRound);
        re0[u] = a.q[u].w * cos_rev(ph);
        im0[u] = a.q[u].w * sin_rev(ph);
      }
#pragma unroll
      for (int t = 0; t < T; ++t) {
#pragma unroll
        for (int u = 0; u < 2; ++u) {
          const float l = t == 0 ? a.l[u].x : a.l[u].y;
          re[u] = l * re0[u];
          im[u] = l * im0[u];
        }
        gr[t] = re[0] + re[1];
        gi[t] = im[0] + im[1];
        store_pair<OP, SPLIT, kPairTileY * kPairPitch>(ldsB[buf][t], gg * kPairPitch + pp, re, im);
      }
    };

    f32x16_t acc[4], sp[2], sm[2];  // [Prr, Pii, Pri, Pir] of this wave's node; folded S(+x), S(-x)
    acc[0] = acc[1] = acc[2] = acc[3] = (f32x16_t){};
    sp[0] = sp[1] = sm[0] = sm[1] = (f32x16_t){};

    auto consume = [&](int buf) {
      const int pofs = (lane >> 5) * 4;
      const int row = wx * 32 + (lane & 31);
      const int col = wy * 32 + (lane & 31);
      const uint32_t* bt = &ldsB[buf][node][0][0];
      const u32x4_t ar = *reinterpret_cast<const u32x4_t*>(&ldsA[buf][0][row * kPairPitch + pofs]);
      const u32x4_t ai = *reinterpret_cast<const u32x4_t*>(&ldsA[buf][2][row * kPairPitch + pofs]);
      const u32x4_t br = *reinterpret_cast<const u32x4_t*>(bt + 0 * kPairTileY * kPairPitch + col * kPairPitch + pofs);
      const u32x4_t bi = *reinterpret_cast<const u32x4_t*>(bt + 2 * kPairTileY * kPairPitch + col * kPairPitch + pofs);
      acc[0] = mfma<OP>(ar, br, acc[0]);
      acc[1] = mfma<OP>(ai, bi, acc[1]);
      acc[2] = mfma<OP>(ar, bi, acc[2]);
      acc[3] = mfma<OP>(ai, br, acc[3]);
      if (SPLIT) {
        const u32x4_t arl = *reinterpret_cast<const u32x4_t*>(&ldsA[buf][1][row * kPairPitch + pofs]);
        const u32x4_t ail = *reinterpret_cast<const u32x4_t*>(&ldsA[buf][3][row * kPairPitch + pofs]);
        const u32x4_t brl = *reinterpret_cast<const u32x4_t*>(bt + 1 * kPairTileY * kPairPitch + col * kPairPitch + pofs);
        const u32x4_t bil = *reinterpret_cast<const u32x4_t*>(bt + 3 * kPairTileY * kPairPitch + col * kPairPitch + pofs);
        acc[0] = mfma<OP>(ar, brl, acc[0]);
        acc[1] = mfma<OP>(ai, bil, acc[1]);
        acc[2] = mfma<OP>(ar, bil, acc[2]);
        acc[3] = mfma<OP>(ai, brl, acc[3]);
        acc[0] = mfma<OP>(arl, br, acc[0]);
        acc[1] = mfma<OP>(ail, bi, acc[1]);
        acc[2] = mfma<OP>(arl, bi, acc[2]);
        acc[3] = mfma<OP>(ail, br, acc[3]);
      }
    };

    const int y = y0 + wy * 32 + (lane & 31);
    const float fv = sh.sy * static_cast<float>(y);
    const float fv0 = sh.sy * static_cast<float>(static_cast<int>(y_first));
    // S(+-x) += W_b Phi(r h_b,node) S_node(+-x); acc = 0.  Every lane also folds the x = 0
    // column sums of both nodes (its generation duty).
    auto fold = [&](int b) {
      const float* nd = job.nodes + b * kMaxNodes;
      const float ws = job.bininfo[b].z;
      const float hn = nd[node];
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        const int xa = xa0 + wx * 32 + (i >> 2) * 8 + (lane >> 5) * 4 + (i & 3);
        const float fu = sh.sx * static_cast<float>(xa);
        const float f = phi<BASIS, MODE>(sqrtf(fu * fu + fv * fv) * hn) * ws;
        const float prr = acc[0][i], pii = acc[1][i], pri = acc[2][i], pir = acc[3][i];
        sp[0][i] = fmaf(f, prr - pii, sp[0][i]);
        sp[1][i] = fmaf(f, pri + pir, sp[1][i]);
        sm[0][i] = fmaf(f, prr + pii, sm[0][i]);
        sm[1][i] = fmaf(f, pri - pir, sm[1][i]);
        __builtin_amdgcn_sched_barrier(0);  // one output at a time: keeps the fold's registers few
      }
      acc[0] = acc[1] = acc[2] = acc[3] = (f32x16_t){};
      if (zero_col) {
#pragma unroll
        for (int t = 0; t < T; ++t) {
          const float f0 = phi<BASIS, MODE>(fabsf(fv0) * nd[t]) * ws;
          z0r = fmaf(f0, zr[t], z0r);
          z0i = fmaf(f0, zi[t], z0i);
          zr[t] = zi[t] = 0.0f;
        }
      }
    };

    // Packets, Lagrange weights and step bins are loaded one step ahead of their
    // use, so their latency hides behind a whole step.  The x = 0 contributions of a
    // step are generated with it and added to the column sums when the step is
    // multiplied, i.e. after the fold of the previous bin.
    Ahead an{};
    float pr[T] = {}, pi[T] = {};  // x = 0 contributions of the step generated last
    int cur = 0, bn = 0;
    if (item.s0 < item.s1) {
      Ahead a0;
      fetch(item.s0, &a0);
      cur = step_bin(item.s0);
      if (item.s0 + 1 < item.s1) {
        fetch(item.s0 + 1, &an);
        bn = step_bin(item.s0 + 1);
      }
      generate(a0, 0, pr, pi);
    }
    __syncthreads();
    int since = 0;
    int b = cur;
    for (int s = item.s0; s < item.s1; ++s) {
      const int buf = (s - item.s0) & 1;
      const int b_next = bn;
      float cr[T], ci[T];  // step s
#pragma unroll
      for (int t = 0; t < T; ++t) {
        cr[t] = pr[t];
        ci[t] = pi[t];
      }
      if (s + 1 < item.s1) {
        const Ahead a = an;
        if (s + 2 < item.s1) {
          fetch(s + 2, &an);
          bn = step_bin(s + 2);
        }
        generate(a, buf ^ 1, pr, pi);
      }
      if (b != cur) {
        fold(cur);
        cur = b;
        since = 0;
      }
      if (zero_col) {
#pragma unroll
        for (int t = 0; t < T; ++t) {
          zr[t] += cr[t];
          zi[t] += ci[t];
        }
      }
      consume(buf);
      if (++since == kBFlushSteps) {
        fold(cur);
        since = 0;
      }
      b = b_next;
      __syncthreads();
    }
    if (since) fold(cur);

    // Node 1 hands its sums to node 0.
    if (node == 1) {
      float* xw = xchg + (wave & 3) * (4 * 16 * 64);
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        xw[(0 * 16 + i) * 64 + lane] = sp[0][i];
        xw[(1 * 16 + i) * 64 + lane] = sp[1][i];
        xw[(2 * 16 + i) * 64 + lane] = sm[0][i];
        xw[(3 * 16 + i) * 64 + lane] = sm[1][i];
      }
    }
    __syncthreads();
    if (node == 0 && y < sh.cy) {
      const float* xw = xchg + (wave & 3) * (4 * 16 * 64);
      float2* out = job.partial + static_cast<size_t>(item.chunk) * two_cx * sh.cy + y;
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        const int xa = xa0 + wx * 32 + (i >> 2) * 8 + (lane >> 5) * 4 + (i & 3);
        if (xa > sh.cx) continue;
        out[static_cast<size_t>(xa + sh.cx - 1) * sh.cy] =
            make_float2(sp[0][i] + xw[(0 * 16 + i) * 64 + lane], sp[1][i] + xw[(1 * 16 + i) * 64 + lane]);
        if (xa < sh.cx)
          out[static_cast<size_t>(sh.cx - 1 - xa) * sh.cy] =
              make_float2(sm[0][i] + xw[(2 * 16 + i) * 64 + lane], sm[1][i] + xw[(3 * 16 + i) * 64 + lane]);
      }
    }
    if (zero_col) {
#pragma unroll
      for (int d = 1; d < 8; d <<= 1) {
        z0r += __shfl_xor(z0r, d);
        z0i += __shfl_xor(z0i, d);
      }
      if (pp == 0 && static_cast<int>(y_first) < sh.cy)
        job.partial[(static_cast<size_t>(item.chunk) * two_cx + (sh.cx - 1)) * sh.cy + y_first] = make_float2(z0r, z0i);
    }
  }
}

// ------------------------------------------------------------------ host ---

// Per-group binning, built once per registration and kept until unregister
// (one entry per (bins, nodes) setting in use).
static std::mutex g_bin_mu;
static std::map<std::tuple<const void*, int, int>, BinCache>& bin_caches() {
  static std::map<std::tuple<const void*, int, int>, BinCache> m;
  return m;
}

void mfma_forget_group(const void* key) {
  std::lock_guard<std::mutex> lk(g_bin_mu);
  auto& m = bin_caches();
  for (auto it = m.begin(); it != m.end();) {
    if (std::get<0>(it->first) == key) {
      it->second.release();
      it = m.erase(it);
    } else {
      ++it;
    }
  }
}

int auto_bins(float h_min, float h_max, double ratio) {
  if (!(h_min > 0.0f) || !(h_max > h_min)) return 1;
  const double nb = std::ceil(std::log(static_cast<double>(h_max) / h_min) / std::log(ratio));
  return static_cast<int>(std::min(1024.0, std::max(1.0, nb)));
}

// Sort the group's particles into nb bins of h (log-spaced when h_min > 0),
// each padded to whole 16-particle steps, with T Chebyshev nodes per bin and
// the bin's largest |w|.
static bool build_bins(const Group& g, const GroupInfo& info, int n, int nb, int T, hipStream_t s, BinCache* bc) {
  std::vector<float> h(n);
  std::vector<Constraint> cons(n);
  if (n > 0) {
    if (!hip_ok(hipMemcpyAsync(h.data(), g.d_radii, sizeof(float) * n, hipMemcpyDeviceToHost, s), "hipMemcpy(radii)") ||
        !hip_ok(hipMemcpyAsync(cons.data(), g.d_constraints, sizeof(Constraint) * n, hipMemcpyDeviceToHost, s),
                "hipMemcpy(constraints)"))
      return false;
    if (!hip_ok(hipStreamSynchronize(s), "synchronize")) return false;
  }
  const double lo = info.h_min, hi = info.h_max;
  const bool logb = lo > 0.0 && hi > lo;
  std::vector<double> edge(nb + 1);
  for (int b = 0; b <= nb; ++b)
    edge[b] = logb ? lo * std::pow(hi / lo, static_cast<double>(b) / nb) : lo + (hi - lo) * b / nb;
  auto bin_of = [&](float x) {
    if (!(hi > lo)) return 0;
    const double f = logb ? std::log(x / lo) / std::log(hi / lo) : (x - lo) / (hi - lo);
    const int b = static_cast<int>(f * nb);
    return std::min(nb - 1, std::max(0, b));
  };
  std::vector<int> count(nb, 0);
  std::vector<float4> binfo(nb);
  for (int b = 0; b < nb; ++b)
    binfo[b] = make_float4(static_cast<float>(edge[b] * edge[b]), static_cast<float>(edge[b + 1] * edge[b + 1]), 0.0f,
                           0.0f);
  for (int i = 0; i < n; ++i) {
    const int b = bin_of(h[i]);
    ++count[b];
    binfo[b].z = std::max(binfo[b].z, std::fabs(cons[i].weight));
  }
  for (int b = 0; b < nb; ++b) binfo[b].w = binfo[b].z > 0.0f ? 1.0f / binfo[b].z : 0.0f;
  std::vector<int> start(nb + 1, 0);  // in slots, each bin padded to kStep
  for (int b = 0; b < nb; ++b) start[b + 1] = start[b] + ((count[b] + kStep - 1) / kStep) * kStep;
  const int n_slots = std::max(start[nb], kStep);
  std::vector<int> map(n_slots, -1);
  std::vector<int> fill(start.begin(), start.end() - 1);
  for (int i = 0; i < n; ++i) map[fill[bin_of(h[i])]++] = i;
  const int steps = n_slots / kStep;
  std::vector<uint16_t> sbin(steps, 0);
  for (int b = 0; b < nb; ++b)
    for (int k = start[b] / kStep; k < start[b + 1] / kStep; ++k) sbin[k] = static_cast<uint16_t>(b);
  std::vector<float> nodes(static_cast<size_t>(nb) * kMaxNodes, 0.0f);
  std::vector<double> ndd(static_cast<size_t>(nb) * kMaxNodes, 0.0), inv(static_cast<size_t>(nb) * kMaxNodes, 0.0);
  for (int b = 0; b < nb; ++b) {
    const double c = 0.5 * (edge[b] + edge[b + 1]), d = 0.5 * (edge[b + 1] - edge[b]);
    double* nd = &ndd[static_cast<size_t>(b) * kMaxNodes];
    for (int t = 0; t < T; ++t) nd[t] = c + d * std::cos((2.0 * t + 1.0) * M_PI / (2.0 * T));
    for (int t = 0; t < T; ++t) {
      double den = 1.0;
      for (int q = 0; q < T; ++q)
        if (q != t) den *= (nd[t] - nd[q]);
      nodes[b * kMaxNodes + t] = static_cast<float>(nd[t]);
      inv[b * kMaxNodes + t] = 1.0 / den;
    }
  }
  // Lagrange weights of every particle at its bin's nodes, scaled by 1 / max |w|
  // of the bin (view-independent, so computed once here).
  const int lw = 4 * lag_words(T);
  std::vector<float> lag(static_cast<size_t>(n_slots) * lw, 0.0f);
  for (int k = 0; k < n_slots; ++k) {
    const int i = map[k];
    if (i < 0) continue;
    const int b = bin_of(h[i]);
    const double* nd = &ndd[static_cast<size_t>(b) * kMaxNodes];
    for (int t = 0; t < T; ++t) {
      double l = inv[static_cast<size_t>(b) * kMaxNodes + t] * binfo[b].w;
      for (int q = 0; q < T; ++q)
        if (q != t) l *= (static_cast<double>(h[i]) - nd[q]);
      lag[static_cast<size_t>(k) * lw + t] = static_cast<float>(l);
    }
  }
  bc->n = n;
  bc->nb = nb;
  bc->T = T;
  bc->n_slots = n_slots;
  bc->bin_step.resize(nb + 1);
  for (int b = 0; b <= nb; ++b) bc->bin_step[b] = start[b] / kStep;
  if (!bc->slot_map.reserve(sizeof(int) * n_slots) || !bc->step_bin.reserve(sizeof(uint16_t) * steps) ||
      !bc->nodes.reserve(sizeof(float) * nodes.size()) || !bc->lag.reserve(sizeof(float) * lag.size()) ||
      !bc->bininfo.reserve(sizeof(float4) * nb))
    return false;
  return hip_ok(hipMemcpy(bc->slot_map.ptr, map.data(), sizeof(int) * n_slots, hipMemcpyHostToDevice), "upload") &&
         hip_ok(hipMemcpy(bc->step_bin.ptr, sbin.data(), sizeof(uint16_t) * steps, hipMemcpyHostToDevice), "upload") &&
         hip_ok(hipMemcpy(bc->nodes.ptr, nodes.data(), sizeof(float) * nodes.size(), hipMemcpyHostToDevice), "upload") &&
         hip_ok(hipMemcpy(bc->lag.ptr, lag.data(), sizeof(float) * lag.size(), hipMemcpyHostToDevice), "upload") &&
         hip_ok(hipMemcpy(bc->bininfo.ptr, binfo.data(), sizeof(float4) * nb, hipMemcpyHostToDevice), "upload");
}

const BinCache* bin_cache(const Group& g, const GroupInfo& info, int n, int nb, int T, hipStream_t s) {
  std::lock_guard<std::mutex> lk(g_bin_mu);
  auto key = std::make_tuple(static_cast<const void*>(g.d_constraints), nb, T);
  BinCache& e = bin_caches()[key];
  if (e.n != n || !e.slot_map.ptr) {
    e.release();
    if (!build_bins(g, info, n, nb, T, s, &e)) {
      e.release();
      bin_caches().erase(key);
      return nullptr;
    }
  }
  return &e;
}

template <int OP, bool SPLIT, int T, int BASIS, int MODE>
static void launch_b(int grid, hipStream_t s, const MfmaJob* jobs, const MfmaItem* items, const MfmaShared& sh,
                     int* counter) {
  hipLaunchKernelGGL((mfma_binned_kernel<OP, SPLIT, T, BASIS, MODE>), dim3(grid), dim3(kThreads), 0, s, jobs, items,
                     sh, counter);
}

template <int OP, bool SPLIT, int T>
static void launch_bt(int grid, hipStream_t s, const MfmaJob* jobs, const MfmaItem* items, const MfmaShared& sh,
                      int* counter, int basis, int mode) {
  const bool an = mode == VIS_PHI_ANALYTIC;
  if (basis == SPH)
    an ? launch_b<OP, SPLIT, T, SPH, VIS_PHI_ANALYTIC>(grid, s, jobs, items, sh, counter)
       : launch_b<OP, SPLIT, T, SPH, VIS_PHI_REFERENCE>(grid, s, jobs, items, sh, counter);
  else
    an ? launch_b<OP, SPLIT, T, WENDLAND_D3_C2, VIS_PHI_ANALYTIC>(grid, s, jobs, items, sh, counter)
       : launch_b<OP, SPLIT, T, WENDLAND_D3_C2, VIS_PHI_REFERENCE>(grid, s, jobs, items, sh, counter);
}

template <int OP, bool SPLIT, int BASIS, int MODE>
static void launch_pair(int grid, hipStream_t s, const MfmaJob* jobs, const MfmaItem* items, const MfmaShared& sh,
                        int* counter) {
  hipLaunchKernelGGL((mfma_binned_pair_kernel<OP, SPLIT, BASIS, MODE>), dim3(grid), dim3(kThreads), 0, s, jobs, items,
                     sh, counter);
}

template <int OP, bool SPLIT>
static void launch_bf(int grid, hipStream_t s, const MfmaJob* jobs, const MfmaItem* items, const MfmaShared& sh,
                      int* counter, int T, int basis, int mode) {
  if (T == 2) {
    const bool an = mode == VIS_PHI_ANALYTIC;
    if (basis == SPH)
      an ? launch_pair<OP, SPLIT, SPH, VIS_PHI_ANALYTIC>(grid, s, jobs, items, sh, counter)
         : launch_pair<OP, SPLIT, SPH, VIS_PHI_REFERENCE>(grid, s, jobs, items, sh, counter);
    else
      an ? launch_pair<OP, SPLIT, WENDLAND_D3_C2, VIS_PHI_ANALYTIC>(grid, s, jobs, items, sh, counter)
         : launch_pair<OP, SPLIT, WENDLAND_D3_C2, VIS_PHI_REFERENCE>(grid, s, jobs, items, sh, counter);
    return;
  }
  switch (T) {
    case 3: launch_bt<OP, SPLIT, 3>(grid, s, jobs, items, sh, counter, basis, mode); break;
    case 4: launch_bt<OP, SPLIT, 4>(grid, s, jobs, items, sh, counter, basis, mode); break;
    default: launch_bt<OP, SPLIT, 5>(grid, s, jobs, items, sh, counter, basis, mode); break;
  }
}

bool launch_binned(int grid, hipStream_t s, const MfmaJob* jobs, const MfmaItem* items, const MfmaShared& sh,
                   int* counter, int form, int T, int basis, int mode) {
  if (form == VIS_FORM_MFMA_F16)
    launch_bf<OP_F16, false>(grid, s, jobs, items, sh, counter, T, basis, mode);
  else if (form == VIS_FORM_MFMA_SPLIT_BF16)
    launch_bf<OP_BF16, true>(grid, s, jobs, items, sh, counter, T, basis, mode);
  else
    launch_bf<OP_F16, true>(grid, s, jobs, items, sh, counter, T, basis, mode);
  return hip_ok(hipGetLastError(), "mfma_binned launch");
}

}  // namespace mv
