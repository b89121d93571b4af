// Host plan of one MFMA job (separate so that mfma.h stays free of kernel types).
#pragma once
#include "mfma.h"

namespace mv {

struct MfmaJob {
  const Packet* packets;     // the job's packets in slot order
  const float4* bininfo;     // [bin] (h_lo^2, h_hi^2, max |w|, 1 / max |w|); nullptr: bin0
  const uint16_t* step_bin;  // binned kernel: bin of every 16-particle step
  const float* nodes;        // binned kernel: [bin][kMaxNodes] Chebyshev nodes in h
  const float4* lag;         // binned kernel: [slot][lag_words(T)] l_t(h) / max |w| of the slot's bin
  float2* partial;           // [chunk][2cx][cy] partial spectra of the job
  float4 bin0;               // the single bin when bininfo is nullptr
  float radius;              // uniform-radius kind: Phi(r * radius)
  int tiles_x;
  int pad0, pad1;
};

// A unit of work: one output tile of one job over a range of 16-particle steps.
struct MfmaItem {
  int job, tile, s0, s1;  // steps [s0, s1)
  int bin, chunk;         // the range's bin (sum kernel) and its partial slot
  int cost;               // host-side estimate in steps (queue order); not read by the kernels
  int pad;
};

struct MfmaPlanned {
  int kind = 0, basis = 0, T = 1;
  int tile = 0;  // enum VisTile of the kernel that runs the job
  const BinCache* bc = nullptr;
  int n_slots = 0, nchunks = 0;  // nchunks: partial slots over all passes (two per chunk, two nodes)
  int item_chunks = 0;           // distinct MfmaItem::chunk values
  int slots_per_chunk = 1;       // 2 for the two-node kernel (one slot per node)
  int pass_chunks = 0;           // chunks per pass (item_chunks when the batch runs in one pass)
  long long tile_steps = 0;  // steps x tiles: the job's work
  size_t pk_off = 0, part_off = 0;
  MfmaJob job{};
  std::vector<MfmaItem> items;
};

}  // namespace mv
