// MFMA (matrix-core) forms of the direct sum: host interface.
#pragma once
#include "internal.h"

#include <tuple>
#include <vector>

namespace mv {

enum MfmaKind { KIND_UNIFORM = 0, KIND_GAUSS = 1, KIND_BINNED = 2 };

struct MfmaJob;
struct MfmaItem;
struct BinCache;

// One (view, group) pair to sum, and where its unscaled spectrum goes:
// acc[2cx*cy] (+)= sum_i w_i Phi(r h_i) exp(+2 pi i (x a_i + y b_i)).
struct MfmaJobSpec {
  const Group* g = nullptr;
  GroupInfo info;
  int n = 0;
  ViewParams vp{};
  double3 u{}, v{};
  const VisConfig* c = nullptr;
  double2* acc = nullptr;
  bool first = true;
};

// Host plan of one job.
struct MfmaPlanned;

// Work tables of a prepared batch (device pointers into the workspace).  A batch whose chunk
// slots would exceed the partial budget runs in passes: pass p launches every kernel variant over
// the items of its share of each job's chunks (slots rebased to 0) and reduces them into the FP64
// accumulators before the next pass reuses the slots.
struct MfmaBatch {
  std::vector<MfmaPlanned> plans;
  std::vector<std::tuple<int, int, int, int>> keys;  // (kind, basis, nodes, tile) of every launch
  std::vector<int> job_begin;   // jobs of key k: [job_begin[k], job_begin[k + 1])
  std::vector<int> item_begin;  // items of pass p, key k: [item_begin[p K + k], item_begin[p K + k + 1])
  int passes = 1;
  const MfmaJob* jobs = nullptr;
  const MfmaItem* items = nullptr;
  int* counters = nullptr;  // one per (pass, key), p K + k
};

// Device buffers the MFMA forms keep between calls.
struct MfmaWorkspace {
  DevBuf packets;   // per-view packets of every job, bin-ordered
  DevBuf partial;   // float2 partial spectra of every work chunk
  DevBuf tables;    // MfmaJob[] then MfmaItem[]
  DevBuf counters;  // one work-queue counter per launch
  std::vector<char> tables_host;
  MfmaBatch batch;
  // What the batch was planned for (groups by registration serial, grid, bin settings): a
  // render with the same plan reuses the uploaded tables (the plan does not depend on the view).
  std::vector<unsigned long long> plan_key;
  void release() {
    plan_key.clear();
    packets.release();
    partial.release();
    tables.release();
    counters.release();
  }
};

// Plans (bins, packet slots, work items) and uploads the work tables of a batch.
bool mfma_prepare(const std::vector<MfmaJobSpec>& specs, hipStream_t s, MfmaWorkspace* ws, MfmaBatch* bt);
// Launches a prepared batch: packet prep, persistent MFMA kernels, reductions (graph-capturable).
bool mfma_enqueue(const std::vector<MfmaJobSpec>& specs, const MfmaBatch& bt, hipStream_t s, EventPair* ev);
// Only the work-queue reset (unless a prep launch cleared the counters) and the persistent MFMA
// kernels (packets prepared, partials reduced by the caller).
bool mfma_launch_sums(const std::vector<MfmaJobSpec>& specs, const MfmaBatch& bt, hipStream_t s, EventPair* ev,
                      bool counters_cleared = false, int pass = 0);
// prepare + enqueue.
bool mfma_execute(const std::vector<MfmaJobSpec>& specs, hipStream_t s, MfmaWorkspace* ws, EventPair* ev);
// Bytes of FP32 partial spectra a batch may hold at once before it runs in passes (<= 0: the
// default, kDefaultPartialBudget).  Takes effect at the next plan.
constexpr long long kDefaultPartialBudget = 8LL << 30;
void mfma_set_partial_budget(long long bytes);

// A planned job's particles and radii in slot order (nullptr: slot i holds particle i of the
// group, zeros past n).
const Constraint* mfma_ordered(const MfmaPlanned& pl);
const float* mfma_ordered_radii(const MfmaPlanned& pl);

// Which kernel variant a group takes (MfmaKind).
int mfma_kind(const GroupInfo& info, int basis);

// Drops what the MFMA forms cached about a group (called on unregister).
void mfma_forget_group(const void* key);

// At registration (host arrays at hand): builds the group's h-bins for the config's MFMA setting
// from the host arrays, with the table uploads on stream s, so that the first render finds them
// (no copy back from the device).  Nothing for the plain form or a uniform radius; false (error
// recorded) on failure.
bool mfma_prebin(const Group& g, const GroupInfo& info, const VisConfig* c, hipStream_t s);

}  // namespace mv

#include "mfma_plan.h"
