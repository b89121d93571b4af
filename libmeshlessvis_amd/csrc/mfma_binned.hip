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
// Phi(r h_bt) are applied when the accumulators are folded (every 2048
// particles and at every bin change).  The interpolation error per term is
// bounded by max|Phi^(T)| (r dh)^T / (T! 2^(2T-1)); with bins of ratio 1.1 and
// T = 4 it is below 2e-6 of Phi(0).
//
// mfma_binned_kernel (T >= 3): 8 waves, 128 (x) x 64 (y) complex tile per
// workgroup, 32 x 32 per wave, T accumulator pairs per wave; operand generation
// and split products as in mfma.hip.  T = 2 (the default) runs
// mfma_binned_pair_kernel below, with its own layout.
#include "mfma_common.h"

#include <algorithm>
#include <cmath>
#include <map>
#include <tuple>
#include <type_traits>
#include <vector>

namespace mv {

// Fold the FP32 accumulators at least every 2048 particles (also at every bin change).  Twice
// the former 1024: the 262k clustered view 0.217 -> 0.213 ms, the drifting-cluster batch 46.3 ->
// 45.5 ms, rel-L2 against FP64 unchanged (1.78e-6, 2.42e-6).
constexpr int kBFlushSteps = 128;

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
      float l[T];
    };
    // Exactly the words generate reads (see load_packet_puw).
    auto fetch = [&](int step, Ahead* a) {
      const size_t slot = static_cast<size_t>(step) * kStep + gp;
      a->q = load_packet_puw(job.packets + slot);
      load_floats<T>(reinterpret_cast<const float*>(job.lag + slot * LW), a->l);
    };
    auto step_bin = [&](int step) -> int {
      return *(const __attribute__((address_space(1))) uint16_t*)(job.step_bin + step);
    };
    // Bins of 64 steps at a time, one per lane (bin_at(k): bin of step s0 + k, k ascending).
    const int n_steps = item.s1 - item.s0;
    int bin_blk = -1, bin_v = 0;
    auto bin_at = [&](int k) -> int {
      if ((k >> 6) != bin_blk) {
        bin_blk = k >> 6;
        const int j = (k & ~63) + lane;
        bin_v = j < n_steps ? step_bin(item.s0 + j) : 0;
      }
      return __builtin_amdgcn_readlane(bin_v, k & 63);
    };

    auto generate = [&](const Ahead& a, int buf) {
      const Packet& q = a.q;
      const float* l = a.l;
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
        if (SPLIT) ldsA[buf][1][o] = split_lo<OP>(re, im, hi);
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
          if (SPLIT) ldsB[buf][t][1][o] = split_lo<OP>(re, im, hi);
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

    // Packets and Lagrange weights are loaded one step ahead of their use, so their
    // latency hides behind a whole step; the loads take no branch (clamped into the item),
    // since a conditional load makes the compiler wait for it at the branch's join.
    Ahead an{};
    int cur = 0;
    if (n_steps > 0) {
      Ahead a0;
      fetch(item.s0, &a0);
      fetch(item.s0 + min(1, n_steps - 1), &an);
      cur = bin_at(0);
      generate(a0, 0);
    }
    __syncthreads();
    int since = 0;
    for (int k = 0; k < n_steps; ++k) {
      const int buf = k & 1;
      const int b = bin_at(k);
      if (k + 1 < n_steps) {
        const Ahead a = an;
        fetch(item.s0 + min(k + 2, n_steps - 1), &an);
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
// kept by the waves that generate B.
//
// Tile: TX |x| rows x TY columns (default 64 x 64: 12 waves).  Waves 0-3 generate the
// operands of the next 32-particle stage (bins are padded to 32 particles, so a stage lies
// in one bin); the other waves multiply the current one: consumer c = wave - 4 holds node
// t = c / NB of the 32 x 32 block c % NB (NB = TX TY / 1024 blocks) - 4 products of
// accumulators and the folded S(+-x) sums.  With the default tile, waves w, w + 4 and w + 8
// share a SIMD: one producer and both nodes' consumers of block w.  An item owns chunk
// slots 2 * chunk and 2 * chunk + 1 (the reduction adds them): node t's consumers write its
// folded sums into slot 2 * chunk + t and its producers the x = 0 column.
//
// Split FP16: the two correction products of each node (A_lo B_t,hi + A_hi B_t,lo) are
// summed for both nodes under one node's factor.  With E = w e^(2 pi i y b) / W_b (node
// free: l_0 + l_1 = 1), B_1,hi = rnd(E - B_0,hi) and D = E - B_0,hi - B_1,hi (the nodes'
// low parts together),
//     node 0 accumulates A_hi B_0,hi + A_lo E_hi,  node 1  A_hi B_1,hi + A_hi rnd(D),
// E_hi = rnd(B_0,hi + B_1,hi): with equal factors this is A.E to the split's precision, and
// each consumer issues 8 MFMAs per 16 particles instead of 12.  A_lo B_1 goes with node 0's
// factor and A_hi B_0,lo with node 1's: an error of (Phi(r h_1) - Phi(r h_0)) / Phi(0)
// (<= ~1% for ratio-1.01 bins) times a 2^-11-relative term, about 1e-6 relative.  Split
// BF16 (low parts 2^-8) keeps the three products of each node.
constexpr int kPairStage = 2 * kStep;             // particles per stage
constexpr int kPairPitch = kPairStage / 2 + 4;    // 16 words of 32 particles + 4 against bank conflicts
// 4 producer waves and two consumer waves (nodes 0, 1) per 32 x 32 block of the tile.
constexpr int pair_threads(int TX, int TY) { return (kProducerWaves + 2 * (TX / 32) * (TY / 32)) * 64; }

template <int OP, bool SPLIT, int BASIS, int MODE, int TX, int TY>
__global__ __launch_bounds__(pair_threads(TX, TY), 1) void mfma_binned_pair_kernel(
    const MfmaJob* __restrict__ jobs, const MfmaItem* __restrict__ items, MfmaShared sh, int* __restrict__ counter) {
  static_assert(TX % 32 == 0 && TY % 32 == 0, "tile");
  constexpr int T = 2;
  constexpr bool kShared = SPLIT && OP == OP_F16;  // shared correction products (above)
  constexpr int NB = (TX / 32) * (TY / 32);  // 32 x 32 blocks of the tile
  constexpr int kA = TX * kPairPitch;  // words of one A part
  constexpr int kB = TY * kPairPitch;  // words of one B part
  constexpr int kPairAWords = 2 * 4 * kA;      // [buf][re_hi, re_lo, im_hi, im_lo][row]
  constexpr int kPairBWords = 2 * 2 * 4 * kB;  // [buf][node][...][row]
  __shared__ uint32_t lds[kPairAWords + kPairBWords];
  // The work-item index lives in a padding word (words 16..19 of a row are never read or
  // written by the operand code).
  int* next_item = reinterpret_cast<int*>(&lds[kPairStage / 2]);
  auto ldsA = [&](int buf, int part) { return lds + (buf * 4 + part) * kA; };
  auto ldsB = [&](int buf, int node, int part) { return lds + kPairAWords + ((buf * T + node) * 4 + part) * kB; };

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform (scalar register)
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
    const int xa0 = 1 + tile_x * TX;
    const int y0 = (item.tile / job.tiles_x) * TY;
    // Steps [s0, s1) in stages of two steps; s0 and s1 are even (bins and work items are).
    const int n_stages = (item.s1 - item.s0) / 2;
    auto stage_bin = [&](int k) -> int {
      return *(const __attribute__((address_space(1))) uint16_t*)(job.step_bin + item.s0 + 2 * k);
    };
    // Bins of 64 stages at a time, one per lane (a wave-uniform global load per stage would
    // stall every stage for its latency); bin_at(k) needs k ascending.
    int bin_blk = -1, bin_v = 0;
    auto bin_at = [&](int k) -> int {
      if ((k >> 6) != bin_blk) {
        bin_blk = k >> 6;
        const int j = (k & ~63) + lane;
        bin_v = j < n_stages ? stage_bin(j) : 0;
      }
      return __builtin_amdgcn_readlane(bin_v, k & 63);
    };

    if (wave < kProducerWaves) {
      // ---- producer: operands of stage k+1 while the consumers multiply stage k ----
      const int pp = lane & 15;  // particles 2pp, 2pp+1 of the stage
      // 0..15: A rows gg + 16m (m < TX / 16) and B rows gg + 16m (m < TY / 16).  A wave's four
      // row groups are 4 apart, so
      // its stores (rows 20 words apart) hit each of the 32 LDS banks exactly twice; row
      // groups 1 apart would collide 3-way.
      const int gg = (lane >> 4) * 4 + wave;
      const uint32_t xa_first = static_cast<uint32_t>(xa0 + gg);
      const uint32_t y_first = static_cast<uint32_t>(y0 + gg);
      const bool zero_col = tile_x == 0 && sh.cx > 0;  // (cx = 0: an empty half plane)
      struct Ahead {
        Packet q[2];
        float2 l[2];
      };
      auto fetch = [&](int k, Ahead* a) {
        const size_t slot = static_cast<size_t>(item.s0) * kStep + k * kPairStage + 2 * pp;
#pragma unroll
        for (int u = 0; u < 2; ++u) {
          a->q[u] = load_packet_puw(job.packets + slot + u);
          a->l[u] = load_float2(job.lag + slot + u);
        }
      };
      // x = 0: this lane's column sums of B_t per column m over the current bin (zr, zi), and
      // their node-weighted sum over the item's bins (Zr, Zi).  Each column's 16 lanes
      // (pp = 0..15) are summed once, at the end of the item, by shuffles in a fixed order:
      // every column has one owner row group, so there is nothing to share through LDS.
      constexpr int MA = TX / 16, MB = TY / 16;
      float zr[T][MB], zi[T][MB], Zr[T][MB], Zi[T][MB];
#pragma unroll
      for (int m = 0; m < MB; ++m)
#pragma unroll
        for (int t = 0; t < T; ++t) zr[t][m] = zi[t][m] = Zr[t][m] = Zi[t][m] = 0.0f;
      auto zfold = [&](int b) {
        const float* nd = job.nodes + b * kMaxNodes;
        const float ws = job.bininfo[b].z;
#pragma unroll
        for (int m = 0; m < MB; ++m) {
          const float fv = sh.sy * static_cast<float>(static_cast<int>(y_first) + 16 * m);
#pragma unroll
          for (int t = 0; t < T; ++t) {
            const float f = phi<BASIS, MODE>(fabsf(fv) * nd[t]) * ws;
            Zr[t][m] = fmaf(f, zr[t][m], Zr[t][m]);
            Zi[t][m] = fmaf(f, zi[t][m], Zi[t][m]);
            zr[t][m] = zi[t][m] = 0.0f;
          }
        }
      };
      auto generate = [&](const Ahead& a, int buf, auto zero) {
        uint32_t PA[2], PB[2], dA[2], dB[2];
#pragma unroll
        for (int u = 0; u < 2; ++u) {
          PA[u] = xa_first * a.q[u].pu + kPhaseRound;
          PB[u] = y_first * a.q[u].pv + kPhaseRound;
          dA[u] = a.q[u].pu << 4;
          dB[u] = a.q[u].pv << 4;
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
          }
          store_pair<OP, SPLIT, kA>(reinterpret_cast<uint32_t(*)[kA]>(ldsA(buf, 0)), (gg + 16 * m) * kPairPitch + pp, re,
                                    im);
        }
        // Node weights w * l_t once per particle, then one multiply per value and node.
        float wl[2][T], ws[2];
#pragma unroll
        for (int u = 0; u < 2; ++u) {
          wl[u][0] = a.q[u].w * a.l[u].x;
          wl[u][1] = a.q[u].w * a.l[u].y;
          ws[u] = wl[u][0] + wl[u][1];  // w / W_b: the weights of E
        }
#pragma unroll
        for (int m = 0; m < MB; ++m) {  // B_t(i, y)
          float c0[2], s0[2];
#pragma unroll
          for (int u = 0; u < 2; ++u) {
            const float ph = phase_rev(PB[u]);
            PB[u] += dB[u];
            c0[u] = cos_rev(ph);
            s0[u] = sin_rev(ph);
          }
          const int o = (gg + 16 * m) * kPairPitch + pp;
          if (kShared) {
            // Parts 0, 2 of node t: B_t,hi; parts 1, 3: E_hi for node 0, rnd(D) for node 1.
            // Node 0: B_0 in FP32 and its high part.  Node 1 from v = E - B_0,hi (one FMA per
            // value, from the FP32 E = (w / W_b) e^(...)): B_1,hi = rnd(v), D = v - B_1,hi
            // (exact), so B_0,hi + B_1,hi + D = E; v is B_1 plus B_0's low part, which node 1
            // thus carries - under its own factor, like the shared corrections.
            float r0[2], i0[2];
#pragma unroll
            for (int u = 0; u < 2; ++u) {
              r0[u] = wl[u][0] * c0[u];
              i0[u] = wl[u][0] * s0[u];
            }
            const uint32_t h0r = pack2<OP>(r0[0], r0[1]), h0i = pack2<OP>(i0[0], i0[1]);
            const float vr0 = fma_minus_half<0>(ws[0], c0[0], h0r), vr1 = fma_minus_half<1>(ws[1], c0[1], h0r);
            const float vi0 = fma_minus_half<0>(ws[0], s0[0], h0i), vi1 = fma_minus_half<1>(ws[1], s0[1], h0i);
            const uint32_t h1r = pack2<OP>(vr0, vr1), h1i = pack2<OP>(vi0, vi1);
            if (decltype(zero)::value) {  // x = 0: the exact FP32 values of each node
              zr[0][m] += r0[0] + r0[1];
              zi[0][m] += i0[0] + i0[1];
              zr[1][m] += wl[0][1] * c0[0] + wl[1][1] * c0[1];
              zi[1][m] += wl[0][1] * s0[0] + wl[1][1] * s0[1];
            }
            ldsB(buf, 0, 0)[o] = h0r;
            ldsB(buf, 0, 2)[o] = h0i;
            ldsB(buf, 1, 0)[o] = h1r;
            ldsB(buf, 1, 2)[o] = h1i;
            // rnd(B_0,hi + B_1,hi) for rnd(E): within 2^-11 of it, an error of 2^-22 in A_lo E_hi.
            ldsB(buf, 0, 1)[o] = add_f16x2(h0r, h1r);
            ldsB(buf, 0, 3)[o] = add_f16x2(h0i, h1i);
            ldsB(buf, 1, 1)[o] = pack2<OP>(minus_half<0>(vr0, h1r), minus_half<1>(vr1, h1r));
            ldsB(buf, 1, 3)[o] = pack2<OP>(minus_half<0>(vi0, h1i), minus_half<1>(vi1, h1i));
            continue;
          }
#pragma unroll
          for (int t = 0; t < T; ++t) {
            float re[2], im[2];
#pragma unroll
            for (int u = 0; u < 2; ++u) {
              re[u] = wl[u][t] * c0[u];
              im[u] = wl[u][t] * s0[u];
            }
            if (decltype(zero)::value) {
              zr[t][m] += re[0] + re[1];
              zi[t][m] += im[0] + im[1];
            }
            store_pair<OP, SPLIT, kB>(reinterpret_cast<uint32_t(*)[kB]>(ldsB(buf, t, 0)), o, re, im);
          }
        }
      };
      // Packets and Lagrange weights are loaded two stages ahead of their use, without a
      // branch (a conditional load would make the compiler wait for it at the branch's
      // join, right after issuing it).  The x = 0 sums are folded when the stage about to
      // be generated opens a new bin, and every kBFlushSteps steps, as the consumers fold
      // theirs.  n_stages + 1 barriers, as the consumers.
      auto produce = [&](auto zero) {
        int zb = 0, since = 0;
        if (n_stages > 0) {
          // Three fixed register sets, stage k's packets in set k mod 3: the loads of stage
          // k + 2 are issued as stage k is generated, two stages ahead of their use.  (With a
          // set copied per stage the compiler sinks the loads into the middle of the stage, to
          // reuse their registers, leaving half a stage to hide their latency.)
          Ahead s0, s1, s2;
          if (decltype(zero)::value) zb = bin_at(0);
          fetch(0, &s0);
          fetch(min(1, n_stages - 1), &s1);
          fetch(min(2, n_stages - 1), &s2);
          generate(s0, 0, zero);
          since = 2;
          __syncthreads();
          // Stage k (k >= 1) while the consumers multiply k - 1; refill the set of stage k - 1.
          auto stage = [&](int k, const Ahead& use, Ahead* refill) {
            if (decltype(zero)::value) {
              const int nb = bin_at(k);
              if (nb != zb || since >= kBFlushSteps) {
                zfold(zb);
                zb = nb;
                since = 0;
              }
            }
            fetch(min(k + 2, n_stages - 1), refill);
            __builtin_amdgcn_sched_barrier(0);  // the loads stay ahead of the generation
            generate(use, k & 1, zero);
            since += 2;
            __syncthreads();
          };
          for (int k = 1; k < n_stages; k += 3) {
            stage(k, s1, &s0);
            if (k + 1 >= n_stages) break;
            stage(k + 1, s2, &s1);
            if (k + 2 >= n_stages) break;
            stage(k + 2, s0, &s2);
          }
        }
        __syncthreads();
        if (decltype(zero)::value && n_stages > 0) zfold(zb);
      };
      if (zero_col)
        produce(std::true_type{});
      else
        produce(std::false_type{});
      __syncthreads();  // as the consumers' end-of-item barrier
      if (zero_col) {
        // Sum over the 16 lanes of a column (pp = 0..15), then one lane per column writes node
        // t's x = 0 value into chunk slot 2 * chunk + t.
#pragma unroll
        for (int t = 0; t < T; ++t)
#pragma unroll
          for (int m = 0; m < MB; ++m) {
#pragma unroll
            for (int d = 1; d < 16; d <<= 1) {
              Zr[t][m] += __shfl_xor(Zr[t][m], d);
              Zi[t][m] += __shfl_xor(Zi[t][m], d);
            }
            const int y = y0 + gg + 16 * m;
            if (pp == 0 && y < sh.cy)
              job.partial[(static_cast<size_t>(2 * item.chunk + t) * two_cx + (sh.cx - 1)) * sh.cy + y] =
                  make_float2(Zr[t][m], Zi[t][m]);
          }
      }
      continue;
    }

    // ---- consumer: node t of a 32 |x| x 32 y block ----
    // Waves w, w + 4 and w + 8 share SIMD w: the producer w and the two consumers of block w
    // (nodes 0 and 1), so every SIMD interleaves one wave's operand generation with two
    // waves' MFMAs and folds.
    const int c = wave - kProducerWaves;  // 0 .. 2 NB - 1
    const int blk = c % NB, t = c / NB;
    const int xp = t == 0 ? 1 : 0;  // shared corrections: A part of the second product (lo, hi)
    const int wx = blk % (TX / 32);
    const int wy = blk / (TX / 32);
    f32x16_t acc[4], sp[2], sm[2];  // [Prr, Pii, Pri, Pir] of node t; folded S(+x), S(-x)
    acc[0] = acc[1] = acc[2] = acc[3] = (f32x16_t){};
    sp[0] = sp[1] = sm[0] = sm[1] = (f32x16_t){};

    const int y = y0 + wy * 32 + (lane & 31);
    const float fv = sh.sy * static_cast<float>(y);
    const int xl = xa0 + wx * 32 + (lane >> 5) * 4;  // |x| of output i is xl + (i >> 2) * 8 + (i & 3)
    // S(+-x) += W_b Phi(r h_bt) S_t(+-x) of the accumulators; acc = 0.
    // The bin's node and weight scale through the scalar data cache (b and t are wave-uniform):
    // no vector load to wait for at every fold.  (Recomputing each output's |f| here instead of
    // keeping the 16 radii, which the compiler spills across the stage loop and reloads one by
    // one, removed every spill but measured 3-7% slower: the reloads cost less.)
    auto fold = [&](int b) {
      const float hn = uniform_load(job.nodes + b * kMaxNodes + t);
      const float ws = uniform_load(&job.bininfo[b].z);
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        const float fu = sh.sx * static_cast<float>(xl + (i >> 2) * 8 + (i & 3));
        const float f = phi<BASIS, MODE>(sqrtf(fu * fu + fv * fv) * hn) * ws;
        const float prr = acc[0][i], pii = acc[1][i], pri = acc[2][i], pir = acc[3][i];
        sp[0][i] = fmaf(f, prr - pii, sp[0][i]);
        sp[1][i] = fmaf(f, pri + pir, sp[1][i]);
        sm[0][i] = fmaf(f, prr + pii, sm[0][i]);
        sm[1][i] = fmaf(f, pri - pir, sm[1][i]);
        __builtin_amdgcn_sched_barrier(0);  // one output at a time: keeps the fold's registers few
      }
      acc[0] = acc[1] = acc[2] = acc[3] = (f32x16_t){};
    };

    __syncthreads();
    int cur = n_stages > 0 ? bin_at(0) : 0, since = 0;
    for (int k = 0; k < n_stages; ++k) {
      const int b = bin_at(k);
      if (b != cur || since >= kBFlushSteps) {
        fold(cur);
        cur = b;
        since = 0;
      }
      const int buf = k & 1;
      const uint32_t* A = ldsA(buf, 0);
      const uint32_t* B = ldsB(buf, t, 0);
      // The lane's LDS offsets are recomputed every stage from the lane id (a few VALU
      // instructions) rather than kept live across the fold, where they would spill; the
      // lane id comes from a volatile v_mbcnt pair, which the compiler cannot hoist (a hoisted
      // lane id is spilled, and its reload stalls every stage).
      int lid;
      asm volatile("v_mbcnt_lo_u32_b32 %0, -1, 0\n\tv_mbcnt_hi_u32_b32 %0, -1, %0" : "=v"(lid));
      const int rowa = (wx * 32 + (lid & 31)) * kPairPitch + (lid >> 5) * 4;
      const int rowb = (wy * 32 + (lid & 31)) * kPairPitch + (lid >> 5) * 4;
      auto lda = [&](int kk, int part) { return *reinterpret_cast<const u32x4_t*>(A + part * kA + rowa + kk * 8); };
      auto ldb = [&](int kk, int part) { return *reinterpret_cast<const u32x4_t*>(B + part * kB + rowb + kk * 8); };
      // acc += (x_re, x_im) x (y_re, y_im): the four real products.
      auto quart = [&](const u32x4_t& xr, const u32x4_t& xi, const u32x4_t& yr, const u32x4_t& yi) {
        acc[0] = mfma<OP>(xr, yr, acc[0]);
        acc[1] = mfma<OP>(xi, yi, acc[1]);
        acc[2] = mfma<OP>(xr, yi, acc[2]);
        acc[3] = mfma<OP>(xi, yr, acc[3]);
      };
#pragma unroll
      for (int kk = 0; kk < kPairStage / 16; ++kk) {
        const u32x4_t ahr = lda(kk, 0), ahi = lda(kk, 2), bhr = ldb(kk, 0), bhi = ldb(kk, 2);
        quart(ahr, ahi, bhr, bhi);  // hi.hi
        if (kShared) {  // node 0: A_lo E_hi; node 1: A_hi rnd(D)
          const u32x4_t xr = lda(kk, xp), xi = lda(kk, xp + 2), yr = ldb(kk, 1), yi = ldb(kk, 3);
          quart(xr, xi, yr, yi);
        } else if (SPLIT) {
          const u32x4_t alr = lda(kk, 1), ali = lda(kk, 3);
          quart(alr, ali, bhr, bhi);  // lo.hi
          const u32x4_t blr = ldb(kk, 1), bli = ldb(kk, 3);
          quart(ahr, ahi, blr, bli);  // hi.lo
        }
      }
      since += 2;
      __syncthreads();
    }
    if (n_stages > 0) fold(cur);
    // Node t's sums into chunk slot 2 * chunk + t (every item writes its whole tiles, also
    // without particles).
    if (y < sh.cy) {
      float2* out = job.partial + static_cast<size_t>(2 * item.chunk + t) * two_cx * sh.cy + y;
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        const int xa = xl + (i >> 2) * 8 + (i & 3);
        if (xa > sh.cx) continue;
        out[static_cast<size_t>(xa + sh.cx - 1) * sh.cy] = make_float2(sp[0][i], sp[1][i]);
        if (xa < sh.cx) out[static_cast<size_t>(sh.cx - 1 - xa) * sh.cy] = make_float2(sm[0][i], sm[1][i]);
      }
    }
    __syncthreads();  // end of the item: the producers wait here too
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

// Sort the group's particles into nb bins of |h| (Phi is even in h), each padded to whole
// 16-particle steps, with T Chebyshev nodes per bin and the bin's largest |w|.  The nonzero
// radii get log-spaced bins over [h_min, h_max] (linear if that range is empty); radii of 0,
// when the group has any, get bin 0 of width 0, whose first node is 0 with Lagrange weight 1.
// From host arrays h[n] (radii) and cons[n]; the tables are uploaded on stream s from host
// copies kept in bc.
static bool build_bins_host(const float* h, const Constraint* cons, const GroupInfo& info, int n, int nb, int T,
                            hipStream_t s, BinCache* bc) {
  const double lo = info.h_min, hi = info.h_max;
  const int z = info.has_zero_h ? 1 : 0;  // bins ahead of the nonzero radii' bins
  const int nbp = nb - z;                 // bins of the nonzero radii
  if (nbp < 1) {
    set_error(VIS_ERROR_INVALID_CONFIG, "h-binning: no bin left for the nonzero radii");
    return false;
  }
  const bool logb = lo > 0.0 && hi > lo;
  // Bin b spans [e_lo[b], e_hi[b]]; the zero bin is the point 0.
  std::vector<double> e_lo(nb, 0.0), e_hi(nb, 0.0);
  for (int b = 0; b < nbp; ++b) {
    auto edge = [&](int k) { return logb ? lo * std::pow(hi / lo, static_cast<double>(k) / nbp) : lo + (hi - lo) * k / nbp; };
    e_lo[z + b] = edge(b);
    e_hi[z + b] = edge(b + 1);
  }
  const double log_ratio = logb ? std::log(hi / lo) : 1.0;
  auto bin_of = [&](float x) {
    const double a = std::fabs(static_cast<double>(x));
    if (z && !(a > 0.0)) return 0;
    if (!(hi > lo)) return z;
    const double f = logb ? std::log(a / lo) / log_ratio : (a - lo) / (hi - lo);
    const int b = static_cast<int>(f * nbp);
    return z + std::min(nbp - 1, std::max(0, b));
  };
  // Each particle's bin, once (one logarithm per particle).
  std::vector<uint16_t> bin(n);
  std::vector<int> count(nb, 0);
  std::vector<float4>& binfo = bc->h_binfo;
  binfo.assign(nb, make_float4(0.0f, 0.0f, 0.0f, 0.0f));
  for (int b = 0; b < nb; ++b)
    binfo[b] = make_float4(static_cast<float>(e_lo[b] * e_lo[b]), static_cast<float>(e_hi[b] * e_hi[b]), 0.0f,
                           0.0f);
  for (int i = 0; i < n; ++i) {
    const int b = bin_of(h[i]);
    bin[i] = static_cast<uint16_t>(b);
    ++count[b];
    binfo[b].z = std::max(binfo[b].z, std::fabs(cons[i].weight));
  }
  for (int b = 0; b < nb; ++b) binfo[b].w = binfo[b].z > 0.0f ? 1.0f / binfo[b].z : 0.0f;
  // In slots, each bin padded to two steps (32 particles): the two-node kernel's stages
  // never straddle a bin.
  constexpr int kBinPad = 2 * kStep;
  std::vector<int> start(nb + 1, 0);
  for (int b = 0; b < nb; ++b) start[b + 1] = start[b] + ((count[b] + kBinPad - 1) / kBinPad) * kBinPad;
  const int n_slots = std::max(start[nb], kBinPad);
  std::vector<int> map(n_slots, -1);
  std::vector<int> fill(start.begin(), start.end() - 1);
  for (int i = 0; i < n; ++i) map[fill[bin[i]]++] = i;
  const int steps = n_slots / kStep;
  std::vector<uint16_t>& sbin = bc->h_sbin;
  sbin.assign(steps, 0);
  for (int b = 0; b < nb; ++b)
    for (int k = start[b] / kStep; k < start[b + 1] / kStep; ++k) sbin[k] = static_cast<uint16_t>(b);
  std::vector<float>& nodes = bc->h_nodes;
  nodes.assign(static_cast<size_t>(nb) * kMaxNodes, 0.0f);
  std::vector<double> ndd(static_cast<size_t>(nb) * kMaxNodes, 0.0), inv(static_cast<size_t>(nb) * kMaxNodes, 0.0);
  std::vector<char> point(nb, 0);  // bins of width 0: every particle sits on the first node
  for (int b = 0; b < nb; ++b) {
    const double c = 0.5 * (e_lo[b] + e_hi[b]), d = 0.5 * (e_hi[b] - e_lo[b]);
    double* nd = &ndd[static_cast<size_t>(b) * kMaxNodes];
    point[b] = !(d > 0.0);
    for (int t = 0; t < T; ++t) nd[t] = point[b] ? c : c + d * std::cos((2.0 * t + 1.0) * M_PI / (2.0 * T));
    for (int t = 0; t < T; ++t) {
      double den = 1.0;
      for (int q = 0; q < T; ++q)
        if (q != t) den *= (nd[t] - nd[q]);
      nodes[b * kMaxNodes + t] = static_cast<float>(nd[t]);
      inv[b * kMaxNodes + t] = point[b] ? 0.0 : 1.0 / den;
    }
  }
  // Lagrange weights of every particle at its bin's nodes, scaled by 1 / max |w|
  // of the bin (view-independent, so computed once here).
  const int lw = 4 * lag_words(T);
  std::vector<float>& lag = bc->h_lag;
  lag.assign(static_cast<size_t>(n_slots) * lw, 0.0f);
  for (int k = 0; k < n_slots; ++k) {
    const int i = map[k];
    if (i < 0) continue;
    const int b = bin[i];
    const double* nd = &ndd[static_cast<size_t>(b) * kMaxNodes];
    const double a = std::fabs(static_cast<double>(h[i]));
    for (int t = 0; t < T; ++t) {
      double l;
      if (point[b]) {
        l = t == 0 ? binfo[b].w : 0.0;
      } else {
        l = inv[static_cast<size_t>(b) * kMaxNodes + t] * binfo[b].w;
        for (int q = 0; q < T; ++q)
          if (q != t) l *= (a - nd[q]);
      }
      lag[static_cast<size_t>(k) * lw + t] = static_cast<float>(l);
    }
  }
  bc->n = n;
  bc->nb = nb;
  bc->T = T;
  bc->n_slots = n_slots;
  bc->bin_step.resize(nb + 1);
  for (int b = 0; b <= nb; ++b) bc->bin_step[b] = start[b] / kStep;
  // The particles in slot order (zero on padding slots): the packet prep reads them contiguously.
  std::vector<Constraint>& ord = bc->h_ordered;
  std::vector<float>& ord_r = bc->h_ordered_r;
  ord.assign(n_slots, Constraint{});
  ord_r.assign(n_slots, 0.0f);
  for (int k = 0; k < n_slots; ++k)
    if (map[k] >= 0) {
      ord[k] = cons[map[k]];
      ord_r[k] = h[map[k]];
    }
  if (!bc->ordered.reserve(sizeof(Constraint) * n_slots) || !bc->ordered_r.reserve(sizeof(float) * n_slots) ||
      !bc->step_bin.reserve(sizeof(uint16_t) * steps) ||
      !bc->nodes.reserve(sizeof(float) * nodes.size()) || !bc->lag.reserve(sizeof(float) * lag.size()) ||
      !bc->bininfo.reserve(sizeof(float4) * nb))
    return false;
  auto up = [&](void* dst, const void* src, size_t bytes) {
    return hip_ok(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s), "upload bins");
  };
  return up(bc->ordered.ptr, ord.data(), sizeof(Constraint) * n_slots) &&
         up(bc->ordered_r.ptr, ord_r.data(), sizeof(float) * n_slots) &&
         up(bc->step_bin.ptr, sbin.data(), sizeof(uint16_t) * steps) &&
         up(bc->nodes.ptr, nodes.data(), sizeof(float) * nodes.size()) &&
         up(bc->lag.ptr, lag.data(), sizeof(float) * lag.size()) &&
         up(bc->bininfo.ptr, binfo.data(), sizeof(float4) * nb);
}

// From a copy of the group's device arrays (the render's stream s orders it after the
// registration's copies).  Returns with the tables on the device: a cached binning may be used
// next from another config's stream.
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
  return build_bins_host(h.data(), cons.data(), info, n, nb, T, s, bc) &&
         hip_ok(hipStreamSynchronize(s), "synchronize bins");
}

static const BinCache* bin_cache_impl(const Group& g, const GroupInfo& info, int n, int nb, int T, hipStream_t s,
                                      bool from_host) {
  std::lock_guard<std::mutex> lk(g_bin_mu);
  auto key = std::make_tuple(static_cast<const void*>(g.d_constraints), nb, T);
  BinCache& e = bin_caches()[key];
  if (e.n != n || !e.ordered.ptr) {
    e.release();
    const bool ok = from_host ? build_bins_host(g.h_radii, g.h_constraints, info, n, nb, T, s, &e)
                              : build_bins(g, info, n, nb, T, s, &e);
    if (!ok) {
      e.release();
      bin_caches().erase(key);
      return nullptr;
    }
  }
  return &e;
}

const BinCache* bin_cache(const Group& g, const GroupInfo& info, int n, int nb, int T, hipStream_t s) {
  return bin_cache_impl(g, info, n, nb, T, s, false);
}

const BinCache* bin_cache_from_host(const Group& g, const GroupInfo& info, int n, int nb, int T, hipStream_t s) {
  return bin_cache_impl(g, info, n, nb, T, s, true);
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
  // Tile variants (tile_dims) for the split forms; the single-FP16 form keeps the default.
  constexpr int X = kPairTileX, Y = kPairTileY;
  switch (SPLIT ? sh.tile : VIS_TILE_DEFAULT) {
    case VIS_TILE_HALF_X:
      hipLaunchKernelGGL((mfma_binned_pair_kernel<OP, SPLIT, BASIS, MODE, X / 2, Y>), dim3(grid),
                         dim3(pair_threads(X / 2, Y)), 0, s, jobs, items, sh, counter);
      break;
    case VIS_TILE_HALF_Y:
      hipLaunchKernelGGL((mfma_binned_pair_kernel<OP, SPLIT, BASIS, MODE, X, Y / 2>), dim3(grid),
                         dim3(pair_threads(X, Y / 2)), 0, s, jobs, items, sh, counter);
      break;
    case VIS_TILE_QUARTER:
      hipLaunchKernelGGL((mfma_binned_pair_kernel<OP, SPLIT, BASIS, MODE, X / 2, Y / 2>), dim3(grid),
                         dim3(pair_threads(X / 2, Y / 2)), 0, s, jobs, items, sh, counter);
      break;
    default:
      hipLaunchKernelGGL((mfma_binned_pair_kernel<OP, SPLIT, BASIS, MODE, X, Y>), dim3(grid), dim3(pair_threads(X, Y)),
                         0, s, jobs, items, sh, counter);
  }
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
