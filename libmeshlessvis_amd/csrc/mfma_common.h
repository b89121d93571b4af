// Helpers shared by the MFMA direct-sum kernels (gfx950, wave64).
#pragma once
#include "mfma.h"
#include "phi.h"

#include <algorithm>
#include <vector>

namespace mv {

typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));
typedef uint32_t u32x4_t __attribute__((ext_vector_type(4)));

constexpr int kStep = 16;          // particles per LDS stage
constexpr int kPitch = kStep + 4;  // words per LDS row (padding against bank conflicts)
constexpr int kTileX = 128;        // sum kernel: |x| rows per workgroup tile (x and -x each)
constexpr int kTileY = 128;        // sum kernel: y columns per workgroup tile
constexpr int kSumThreads = 768;   // sum kernel: 12 waves, 4 producers and 8 consumers
constexpr int kThreads = 512;      // 8 waves
constexpr int kProducerWaves = 4;  // waves 0-3 generate operands, waves 4-7 run the MFMAs
constexpr int kMaxChunkSteps = 1024;  // at most 16384 particles per FP32 accumulation
constexpr int kMaxNodes = 5;
constexpr float kTwoPowM32 = 2.3283064365386962890625e-10f;
constexpr float kLog2e = 1.44269504088896340736f;

enum OperandType { OP_F16 = 0, OP_BF16 = 1 };

// Constants shared by every job of a launch.
struct MfmaShared {
  int cx, cy;
  float sx, sy;
  float sigma_pi_log2e;  // Gaussian kind: +-pi*log2(e); 0 otherwise
  int ref_hi;            // Gaussian h_ref: 0 the bin's lower edge, 1 its upper edge
  int n_items;
  int tile;  // enum VisTile of the launch's jobs
};

// Tile (|x| rows, y columns per workgroup) of the sum kernel and of the two-node binned kernel
// for a VisTile value: the default, half the |x| rows, half the y columns, or both halved.
// The split forms build every variant; the single-FP16 form and the T >= 3 binned kernel keep
// their default tile.
inline void tile_dims(int tile, int dx, int dy, int* tx, int* ty) {
  *tx = (tile == VIS_TILE_HALF_X || tile == VIS_TILE_QUARTER) ? dx / 2 : dx;
  *ty = (tile == VIS_TILE_HALF_Y || tile == VIS_TILE_QUARTER) ? dy / 2 : dy;
}


// Two values rounded to nearest even into one word (v_cvt_pk_f16_f32 / v_cvt_pk_bf16_f32).
template <int OP>
__device__ __forceinline__ uint32_t pack2(float a, float b) {
  if (OP == OP_F16) {
    const _Float16 __attribute__((ext_vector_type(2))) h = {static_cast<_Float16>(a), static_cast<_Float16>(b)};
    return __builtin_bit_cast(uint32_t, h);
  } else {
    const __bf16 ha = static_cast<__bf16>(a), hb = static_cast<__bf16>(b);
    return static_cast<uint32_t>(__builtin_bit_cast(uint16_t, ha)) |
           (static_cast<uint32_t>(__builtin_bit_cast(uint16_t, hb)) << 16);
  }
}

template <int OP>
__device__ __forceinline__ void unpack2(uint32_t u, float* a, float* b) {
  if (OP == OP_F16) {
    const auto h = __builtin_bit_cast(_Float16 __attribute__((ext_vector_type(2))), u);
    *a = static_cast<float>(h[0]);
    *b = static_cast<float>(h[1]);
  } else {
    *a = __builtin_bit_cast(float, u << 16);
    *b = __builtin_bit_cast(float, u & 0xffff0000u);
  }
}

// Packet load through a global-address-space pointer (a global load, which the
// wait counters track apart from LDS traffic, instead of a flat load).
typedef float f32x4_t __attribute__((ext_vector_type(4)));
__device__ __forceinline__ Packet load_packet(const Packet* p) {
  const u32x4_t v = *(const __attribute__((address_space(1))) u32x4_t*)p;
  Packet q;
  q.pu = v[0];
  q.pv = v[1];
  // Elements are copied to scalars before the bit cast: clang's
  // __builtin_bit_cast of a vector-element lvalue reads element 0.
  const uint32_t wb = v[2], hb = v[3];
  q.w = __builtin_bit_cast(float, wb);
  q.h = __builtin_bit_cast(float, hb);
  return q;
}
__device__ __forceinline__ float4 load_float4(const float4* p) {
  const f32x4_t v = *(const __attribute__((address_space(1))) f32x4_t*)p;
  return make_float4(v[0], v[1], v[2], v[3]);
}
// Only the fields a kernel reads.  A loaded register the kernel never reads is free to
// the register allocator while the load is in flight, and reusing it makes the compiler
// wait for the load right there - in a prefetch, at once.
typedef uint32_t u32x3_t __attribute__((ext_vector_type(3)));
typedef float f32x2_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ Packet load_packet_puw(const Packet* p) {  // pu, pv, w (h = 0)
  const u32x3_t v = *(const __attribute__((address_space(1))) u32x3_t*)p;
  Packet q;
  q.pu = v[0];
  q.pv = v[1];
  const uint32_t wb = v[2];
  q.w = __builtin_bit_cast(float, wb);
  q.h = 0.0f;
  return q;
}
__device__ __forceinline__ float2 load_float2(const float4* p) {  // x, y
  const f32x2_t v = *(const __attribute__((address_space(1))) f32x2_t*)p;
  return make_float2(v[0], v[1]);
}
// N consecutive floats from a 16-byte aligned address: loads of 4 and one of the rest.
template <int N>
__device__ __forceinline__ void load_floats(const float* p, float* out) {
  typedef const __attribute__((address_space(1))) float* gptr;
#pragma unroll
  for (int i = 0; i + 4 <= N; i += 4) {
    const f32x4_t v = *(const __attribute__((address_space(1))) f32x4_t*)(p + i);
    out[i] = v[0], out[i + 1] = v[1], out[i + 2] = v[2], out[i + 3] = v[3];
  }
  constexpr int r = N & 3, b = N & ~3;
  if (r == 3) {
    typedef float f3 __attribute__((ext_vector_type(3)));
    const f3 v = *(const __attribute__((address_space(1))) f3*)(p + b);
    out[b] = v[0], out[b + 1] = v[1], out[b + 2] = v[2];
  } else if (r == 2) {
    const f32x2_t v = *(const __attribute__((address_space(1))) f32x2_t*)(p + b);
    out[b] = v[0], out[b + 1] = v[1];
  } else if (r == 1) {
    out[b] = *gptr(p + b);
  }
}

// Half an LSB of the 23 phase bits phase_rev keeps; added once to a phase so
// that phase_rev rounds instead of truncating.
constexpr uint32_t kPhaseRound = 1u << 8;

// Fixed-point phase P (revolutions * 2^32) -> 1 + P / 2^32 rounded down to 23
// bits, a float in [1, 2): its fraction is the phase, and v_sin/v_cos take
// revolutions, so sin_rev/cos_rev of it are the sine/cosine of the phase.  One
// v_alignbit_b32: the exponent bits of 1.0 (0x7f << 23) above P's top 23 bits.
__device__ __forceinline__ float phase_rev(uint32_t P) {
  return __builtin_bit_cast(float, __builtin_amdgcn_alignbit(0x7fu, P, 9));
}

// The FP16 residuals rnd(a - hi.x), rnd(b - hi.y) of a packed pair hi: v_fma_mix
// takes the FP16 halves of hi as FP32 operands, forms a - hi exactly (one FMA) and
// rounds it to FP16 into the low / high half of the result - two instructions in
// place of two conversions, two subtractions and a pack.
__device__ __forceinline__ uint32_t split_lo_f16(float a, float b, uint32_t hi) {
  uint32_t d;
  asm("v_fma_mixlo_f16 %0, %1, 1.0, -%2 op_sel_hi:[0,0,1]" : "=v"(d) : "v"(a), "v"(hi));
  asm("v_fma_mixhi_f16 %0, %1, 1.0, -%2 op_sel:[0,0,1] op_sel_hi:[0,0,1]" : "+v"(d) : "v"(b), "v"(hi));
  return d;
}

// a * b - h rounded once (one v_fma_mix_f32), h the FP16 half U of the word hi taken as
// FP32.
template <int U>
__device__ __forceinline__ float fma_minus_half(float a, float b, uint32_t hi) {
  float d;
  if (U == 0)
    asm("v_fma_mix_f32 %0, %1, %2, -%3 op_sel_hi:[0,0,1]" : "=v"(d) : "v"(a), "v"(b), "v"(hi));
  else
    asm("v_fma_mix_f32 %0, %1, %2, -%3 op_sel:[0,0,1] op_sel_hi:[0,0,1]" : "=v"(d) : "v"(a), "v"(b), "v"(hi));
  return d;
}

// a - h, h the FP16 half U of hi (one v_fma_mix_f32).
template <int U>
__device__ __forceinline__ float minus_half(float a, uint32_t hi) {
  float d;
  if (U == 0)
    asm("v_fma_mix_f32 %0, %1, 1.0, -%2 op_sel_hi:[0,0,1]" : "=v"(d) : "v"(a), "v"(hi));
  else
    asm("v_fma_mix_f32 %0, %1, 1.0, -%2 op_sel:[0,0,1] op_sel_hi:[0,0,1]" : "=v"(d) : "v"(a), "v"(hi));
  return d;
}

// A wave-uniform value read through the scalar data cache (s_load: no vector registers, no
// wait on the vector memory counter).  A read only; p must be the same in every lane.
__device__ __forceinline__ float uniform_load(const float* p) {
  return *(const __attribute__((address_space(4))) float*)p;
}

// The FP16 pair sum x + y of two packed words (v_pk_add_f16, round to nearest even).
__device__ __forceinline__ uint32_t add_f16x2(uint32_t x, uint32_t y) {
  using h2 = _Float16 __attribute__((ext_vector_type(2)));
  return __builtin_bit_cast(uint32_t, __builtin_bit_cast(h2, x) + __builtin_bit_cast(h2, y));
}

// rnd(a - hi.x), rnd(b - hi.y): the low words of the split a = hi + lo.
template <int OP>
__device__ __forceinline__ uint32_t split_lo(float a, float b, uint32_t hi) {
  if (OP == OP_F16) return split_lo_f16(a, b, hi);
  float ha, hb;
  unpack2<OP>(hi, &ha, &hb);
  return pack2<OP>(a - ha, b - hb);
}

// hi = rnd(v), lo = rnd(v - hi) of two consecutive particles' real and imaginary
// parts, stored as words [re_hi, re_lo, im_hi, im_lo][o] of an LDS array of N words
// per part (one 16-bit value of each particle per word).
template <int OP, bool SPLIT, int N>
__device__ __forceinline__ void store_pair(uint32_t (*arr)[N], int o, const float* re, const float* im) {
  const uint32_t hr = pack2<OP>(re[0], re[1]), hm = pack2<OP>(im[0], im[1]);
  arr[0][o] = hr;
  arr[2][o] = hm;
  if (SPLIT) {
    arr[1][o] = split_lo<OP>(re[0], re[1], hr);
    arr[3][o] = split_lo<OP>(im[0], im[1], hm);
  }
}

template <int OP>
__device__ __forceinline__ f32x16_t mfma(u32x4_t a, u32x4_t b, f32x16_t c) {
  if (OP == OP_F16)
    return __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(half8_t, a), __builtin_bit_cast(half8_t, b), c,
                                                  0, 0, 0);
  else
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8_t, a), __builtin_bit_cast(bf16x8_t, b),
                                                   c, 0, 0, 0);
}

__device__ __forceinline__ u32x4_t negate_imag(u32x4_t v) {  // (r, i) -> (r, -i) per word
  return v ^ (u32x4_t){0x80000000u, 0x80000000u, 0x80000000u, 0x80000000u};
}
__device__ __forceinline__ u32x4_t swap_halves(u32x4_t v) {  // (r, i) -> (i, r) per word
  return (v >> 16) | (v << 16);
}

// h-binning of a group (mfma_binned.hip): slot k of the per-view packet array holds one
// particle of the group (the particles sorted by bin, each bin padded with zero slots to whole
// 16-particle steps); step_bin gives each step's bin.  The particles are kept in that slot
// order (ordered, ordered_r), so the per-view packet prep reads them contiguously.
// float4 words per slot holding the T Lagrange weights of a particle.
constexpr int lag_words(int T) { return (T + 3) / 4; }

struct BinCache {
  int n = 0, nb = 0, T = 0;
  int n_slots = 0;
  DevBuf ordered;    // Constraint[n_slots]: the particle of slot k, zero on padding
  DevBuf ordered_r;  // float[n_slots]: its radius, zero on padding
  DevBuf step_bin;   // uint16_t[n_slots / 16]
  DevBuf nodes;     // float[nb][kMaxNodes] Chebyshev nodes in h
  DevBuf lag;       // float4[n_slots][lag_words(T)]: Lagrange weights l_t(h_i) / max |w| of the bin, 0 on padding
  DevBuf bininfo;   // float4[nb] (h_lo^2, h_hi^2, max |w|, 1 / max |w|)
  std::vector<int> bin_step;  // host: first step of every bin, nb + 1 entries
  // Host copies of the tables: the sources of their stream-ordered uploads, kept until release.
  std::vector<Constraint> h_ordered;
  std::vector<float> h_ordered_r;
  std::vector<uint16_t> h_sbin;
  std::vector<float> h_nodes, h_lag;
  std::vector<float4> h_binfo;
  void release() {
    ordered.release();  // hipFree waits for the device, so the uploads above are done
    ordered_r.release();
    step_bin.release();
    nodes.release();
    lag.release();
    bininfo.release();
    h_ordered.clear();
    h_ordered_r.clear();
    h_sbin.clear();
    h_nodes.clear();
    h_lag.clear();
    h_binfo.clear();
  }
};
// The group's binning for (nb, T), built on first use (from a copy of the device arrays);
// nullptr (error recorded) on failure.
const BinCache* bin_cache(const Group& g, const GroupInfo& info, int n, int nb, int T, hipStream_t s);
// The same, built from the group's host arrays (which must hold what was registered).
const BinCache* bin_cache_from_host(const Group& g, const GroupInfo& info, int n, int nb, int T, hipStream_t s);
// Bins of ratio `ratio` over [h_min, h_max], at least 1, at most 1024.
int auto_bins(float h_min, float h_max, double ratio);

// Tile of the binned kernel (mfma_binned.hip).
constexpr int kBinnedTileX = 128;
constexpr int kBinnedTileY = 64;
// Tile of the two-node binned kernel: |x| rows (x and -x each) x y columns.
constexpr int kPairTileX = 64;
constexpr int kPairTileY = 64;

// Persistent binned kernel over a job/item table (mfma_binned.hip); false (error recorded) on failure.
bool launch_binned(int grid, hipStream_t s, const MfmaJob* jobs, const MfmaItem* items, const MfmaShared& sh,
                   int* counter, int form, int T, int basis, int mode);

}  // namespace mv
