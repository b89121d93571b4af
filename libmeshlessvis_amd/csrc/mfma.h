// Split-precision MFMA form of the direct sum (see mfma.hip).
#pragma once
#include "internal.h"

namespace mv {

// Device buffers the MFMA form keeps between renders of one config.
struct MfmaWorkspace {
  DevBuf packets;   // per-view operand records, h-sorted
  DevBuf partial;   // split-K partial spectra
  DevBuf chunks;    // int4 (bin, first step, end step, -) work items of the MFMA kernel
  std::vector<int4> chunks_host;
  void release() {
    packets.release();
    partial.release();
    chunks.release();
    chunks_host.clear();
  }
};

// Adds (assigns when first) the group's unscaled sum
// sum_i w_i Phi(r h_i) exp(+2 pi i (x a_i + y b_i)) into acc[2cx*cy].
bool mfma_group_spectrum(const Group& g, const GroupInfo& info, int n, const ViewParams& vp, double3 u, double3 v,
                         const VisConfig* c, bool first, hipStream_t s, MfmaWorkspace* ws, double2* acc,
                         EventPair* ev);

// Drops what the MFMA form cached about a group (called on unregister).
void mfma_forget_group(const void* key);

}  // namespace mv
