// Host runtime of libmeshless_vis.so: VisConfig lifecycle, dataset
// registration, the per-view render pipeline and the additive entry points.
//
// Reference counterparts: meshless_vis/meshless_vis.cu:129-355 (GPU runtime)
// and meshless_vis/fourier_transform.cu:161-181 (dispatch).  The reference
// aborts on any CUDA error; here every entry point records an error code and
// message instead (vis_get_last_error_code / vis_get_last_error).
#include "internal.h"
#include "mfma.h"

#include <hip/hip_gl_interop.h>
#include <rocprofiler-sdk-roctx/roctx.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <atomic>
#include <cstring>
#include <limits>
#include <unordered_map>

namespace mv {

// ---------------------------------------------------------------- errors ---
bool hip_ok(hipError_t e, const char* what) {
  if (e == hipSuccess) return true;
  const int code = (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? VIS_ERROR_OUT_OF_MEMORY : VIS_ERROR_HIP;
  set_error(code, std::string(what) + ": " + hipGetErrorString(e));
  (void)hipGetLastError();  // clear the sticky per-thread error of a failed launch
  return false;
}

// ------------------------------------------------------------ roctx ranges ---
// Every public entry point opens a roctx range named after it, so a marker trace
// (rocprofv3 --marker-trace) shows the library's calls beside their kernels.  Without a tool
// attached the calls do nothing.
namespace {
struct Range {
  explicit Range(const char* name) { roctxRangePushA(name); }
  ~Range() { roctxRangePop(); }
  Range(const Range&) = delete;
  Range& operator=(const Range&) = delete;
};
}  // namespace

// ---------------------------------------------------- device allocation ---
// Every device allocation of the library goes through device_malloc, which also serves the
// allocation-failure hook (vis_debug_fail_device_allocations): with the hook set to k >= 0, k
// more allocations succeed and every one after them fails as out of memory, without calling
// hipMalloc, until the hook is set to a negative value.
static std::atomic<long long> g_fail_after{-1};

bool device_malloc(void** p, size_t bytes, const char* what) {
  *p = nullptr;
  long long left = g_fail_after.load();
  while (left >= 0) {
    if (left == 0) {
      set_error(VIS_ERROR_OUT_OF_MEMORY, std::string(what) + ": allocation failure (injected)");
      return false;
    }
    if (g_fail_after.compare_exchange_weak(left, left - 1)) break;
  }
  return hip_ok(hipMalloc(p, bytes), what);
}

// -------------------------------------------------------------- registry ---
static std::mutex g_reg_mu;
static std::unordered_map<const void*, GroupInfo>& reg() {
  static std::unordered_map<const void*, GroupInfo> m;
  return m;
}
void group_stats(const Constraint* constraints, const float* radii, int n, GroupInfo* info) {
  float wmax = 0.0f;
  for (int i = 0; constraints && i < n; ++i) wmax = std::max(wmax, std::fabs(constraints[i].weight));
  info->w_absmax = wmax;
  info->has_zero_h = false;
  if (radii && n > 0) {
    info->has_radii = true;
    float lo = radii[0], hi = radii[0];
    float amin = std::numeric_limits<float>::infinity(), amax = 0.0f;
    for (int i = 0; i < n; ++i) {
      lo = std::min(lo, radii[i]);
      hi = std::max(hi, radii[i]);
      const float a = std::fabs(radii[i]);
      if (a > 0.0f)
        amin = std::min(amin, a);
      else
        info->has_zero_h = true;
      amax = std::max(amax, a);
    }
    info->uniform_radius = (lo == hi);
    info->radius = lo;
    info->h_min = std::isfinite(amin) ? amin : 0.0f;
    info->h_max = amax;
  } else {
    info->has_radii = false;
    info->uniform_radius = true;
    info->radius = 1.0f;
    info->h_min = info->h_max = 1.0f;
  }
}

void registry_put(const void* key, const GroupInfo& info) {
  static unsigned long long next_serial = 0;
  std::lock_guard<std::mutex> lk(g_reg_mu);
  GroupInfo& e = reg()[key];
  e = info;
  e.serial = ++next_serial;
}
bool registry_get(const void* key, GroupInfo* info) {
  std::lock_guard<std::mutex> lk(g_reg_mu);
  auto it = reg().find(key);
  if (it == reg().end()) return false;
  *info = it->second;
  return true;
}
void registry_erase(const void* key) {
  std::lock_guard<std::mutex> lk(g_reg_mu);
  reg().erase(key);
}

// -------------------------------------------------------------- DevBuf ---
bool DevBuf::reserve(size_t need) {
  if (need <= bytes && ptr) return true;
  release();
  if (need == 0) return true;
  void* p = nullptr;
  if (!device_malloc(&p, need, "hipMalloc(workspace)")) return false;
  ptr = p;
  bytes = need;
  return true;
}
void DevBuf::release() {
  if (ptr) (void)hipFree(ptr);
  ptr = nullptr;
  bytes = 0;
}

}  // namespace mv

// ------------------------------------------------------------------ plan ---
// A captured batch render (vis_fourier_volume_rendering_batched), replayed while
// its signature (configs, datasets, registrations, views, buffers) is unchanged.
struct BatchGraph {
  std::vector<uint64_t> key;   // the captured batch
  std::vector<uint64_t> seen;  // the last batch run without a graph
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipStream_t capture = nullptr;  // default priority
  mv::MfmaWorkspace ws;
  mv::DevBuf io;  // batched prep / finish / C2R tables
  int captures = 0;        // graphs captured (over the config's life)
  long long launches = 0;  // graph launches (the capturing call's included)
  void drop() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    key.clear();
  }
  void release() {
    drop();
    ws.release();
    io.release();
    if (capture) (void)hipStreamDestroy(capture);
    capture = nullptr;
  }
};

struct VisPlan {
  int device = 0;
  hipStream_t stream = nullptr;
  int nx = 0, ny = 0;
  mv::FftPlan fft;  // tables of the C2R of the config's number_of_samples
  mv::DevBuf acc, partial, packets;
  mv::MfmaWorkspace mfma;
  BatchGraph batch;
  hipStream_t copy_stream = nullptr;  // asynchronous registration (created on first use)
  bool timing = false;
  std::vector<mv::EventPair> events;
  size_t events_used = 0;
};

namespace mv {

// The C2R tables of a size this build supports; for another size the config is still created
// (as the reference's vis_config_create, which validates nothing) and vis_config_check says why
// it cannot render.
static bool plan_set_samples(VisPlan* p, int nx, int ny) {
  p->nx = nx;
  p->ny = ny;
  p->fft.x.N = p->fft.y.N = 0;
  if (!fft_length_supported(nx, nullptr) || !fft_length_supported(ny, nullptr)) return true;
  return fft_plan_build(nx, ny, &p->fft);
}

static void plan_destroy(VisPlan* p) {
  if (!p) return;
  p->fft.mem.release();
  p->acc.release();
  p->partial.release();
  p->packets.release();
  p->mfma.release();
  p->batch.release();
  if (p->copy_stream) {
    (void)hipStreamSynchronize(p->copy_stream);
    (void)hipStreamDestroy(p->copy_stream);
  }
  for (auto& e : p->events) {
    (void)hipEventDestroy(e.first);
    (void)hipEventDestroy(e.second);
  }
  delete p;
}

static size_t spectrum_count(const VisConfig* c) {
  return static_cast<size_t>(2) * c->_cutoff_frequency.x * c->_cutoff_frequency.y;
}

static bool alloc_freq_image(VisConfig* c) {
  if (c->_d_freq_image) (void)hipFree(c->_d_freq_image);
  c->_d_freq_image = nullptr;
  const size_t n = spectrum_count(c) * c->_number_of_partial_sums;
  if (n == 0) return true;
  return device_malloc(reinterpret_cast<void**>(&c->_d_freq_image), sizeof(float2) * n, "hipMalloc(_d_freq_image)");
}

static bool alloc_samples(VisConfig* c) {
  if (c->_d_freq_image_arranged) (void)hipFree(c->_d_freq_image_arranged);
  c->_d_freq_image_arranged = nullptr;
  const size_t na = static_cast<size_t>(c->_number_of_samples.x) * (c->_number_of_samples.y / 2 + 1);
  if (!device_malloc(reinterpret_cast<void**>(&c->_d_freq_image_arranged), sizeof(float2) * std::max<size_t>(na, 1),
                     "hipMalloc(_d_freq_image_arranged)"))
    return false;
  if (c->_automatic_d_image) {
    if (c->_d_image) (void)hipFree(c->_d_image);
    c->_d_image = nullptr;
    const size_t ni = static_cast<size_t>(c->_number_of_samples.x) * c->_number_of_samples.y;
    if (!device_malloc(reinterpret_cast<void**>(&c->_d_image), sizeof(float) * std::max<size_t>(ni, 1),
                       "hipMalloc(_d_image)"))
      return false;
  }
  return plan_set_samples(c->_plan, c->_number_of_samples.x, c->_number_of_samples.y);
}

struct DeviceLimits {
  int warp = 64, max_threads = 1024, max_dim0 = 1024;
};
static bool device_limits(DeviceLimits* out) {
  static std::mutex mu;
  static std::unordered_map<int, DeviceLimits> cache;
  int dev = 0;
  if (!hip_ok(hipGetDevice(&dev), "hipGetDevice")) return false;
  std::lock_guard<std::mutex> lk(mu);
  auto it = cache.find(dev);
  if (it == cache.end()) {
    hipDeviceProp_t prop;
    if (!hip_ok(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties")) return false;
    DeviceLimits l;
    l.warp = prop.warpSize;
    l.max_threads = prop.maxThreadsPerBlock;
    l.max_dim0 = prop.maxThreadsDim[0];
    it = cache.emplace(dev, l).first;
  }
  *out = it->second;
  return true;
}

// The vis_config_check rules (meshless_vis.cu:191-210) plus this build's limits.
static bool config_valid(const VisConfig* c, std::string* why) {
  DeviceLimits lim;
  if (!device_limits(&lim)) {
    *why = "cannot query device limits";
    return false;
  }
  const int bl = c->block_length;
  if (bl < 64) { *why = "block_length < 64"; return false; }
  if (bl < lim.warp) { *why = "block_length < wavefront size"; return false; }
  if (bl > lim.max_threads || bl > lim.max_dim0) { *why = "block_length > max threads per block"; return false; }
  if ((2 * c->_cutoff_frequency.x * c->_cutoff_frequency.y) % bl) { *why = "2*cutoff.x*cutoff.y not a multiple of block_length"; return false; }
  if (2 * c->_cutoff_frequency.x > c->_number_of_samples.x) { *why = "2*cutoff.x > number_of_samples.x"; return false; }
  if (2 * c->_cutoff_frequency.y > c->_number_of_samples.y) { *why = "2*cutoff.y > number_of_samples.y"; return false; }
  if (c->_number_of_partial_sums < 1) { *why = "number_of_partial_sums < 1"; return false; }
  // Limits of this build (not in the reference).
  if (c->_cutoff_frequency.x < 0 || c->_cutoff_frequency.y < 0) { *why = "negative cutoff"; return false; }
  // Buffers a failed allocation (vis_config_change_*) left missing.
  if (spectrum_count(c) > 0 && !c->_d_freq_image) { *why = "no spectrum buffer (its allocation failed)"; return false; }
  if (!c->_d_freq_image_arranged) { *why = "no FFT work buffer (its allocation failed)"; return false; }
  if (!fft_length_supported(c->_number_of_samples.x, why) || !fft_length_supported(c->_number_of_samples.y, why))
    return false;
  if (c->_plan->fft.x.N != c->_number_of_samples.x || c->_plan->fft.y.N != c->_number_of_samples.y) {
    *why = "number_of_samples changed without vis_config_change_number_of_samples";
    return false;
  }
  if (c->form < VIS_FORM_PLAIN || c->form > VIS_FORM_MFMA_SPLIT_F16) { *why = "unknown form"; return false; }
  if (c->phi_mode != VIS_PHI_REFERENCE && c->phi_mode != VIS_PHI_ANALYTIC) { *why = "unknown phi_mode"; return false; }
  if (!(c->h_bin_ratio == 0.0f || c->h_bin_ratio > 1.0f)) { *why = "h_bin_ratio must be 0 (auto) or above 1"; return false; }
  if (c->tile < VIS_TILE_DEFAULT || c->tile > VIS_TILE_QUARTER) { *why = "unknown tile"; return false; }
  return true;
}

static EventPair* next_events(VisPlan* p) {
  if (p->events_used == p->events.size()) {
    hipEvent_t a, b;
    if (!hip_ok(hipEventCreate(&a), "hipEventCreate")) return nullptr;
    if (!hip_ok(hipEventCreate(&b), "hipEventCreate")) {
      (void)hipEventDestroy(a);
      return nullptr;
    }
    p->events.emplace_back(a, b);
  }
  return &p->events[p->events_used++];
}

// Number of particle splits for the plain kernel: enough workgroups to fill
// the chip several times over, at least one 256-particle tile per split.
static int choose_ksplit(int ntiles, int freq_tiles) {
  const int target = 2048;
  int k = (target + freq_tiles - 1) / std::max(1, freq_tiles);
  k = std::min(k, std::max(1, ntiles));
  return std::max(1, k);
}

// A group whose device arrays the caller filled itself (not through
// vis_register_meshless_dataset, as the reference allows: the kernels read only
// d_constraints, d_radii and the counts) has no registry entry.  Its facts (weight and
// radius ranges, which the MFMA forms' scaling and bins need) come from a copy of those
// arrays taken at every render, since the caller may rewrite them in place between
// renders; each such copy gets a serial of its own, and the MFMA forms' cached bins of the
// arrays are dropped so that they are rebuilt from the same data.
static std::atomic<unsigned long long> g_transient_serial{1ull << 62};
static bool group_info_for(const Group& g, hipStream_t s, GroupInfo* info) {
  if (registry_get(g.d_constraints, info)) return true;
  *info = GroupInfo{};
  int n = std::max(0, g.number_of_terms);
  if (g.d_number_of_terms > 0) n = std::min(n, g.d_number_of_terms);
  info->n = n;
  info->padded = g.d_number_of_terms;
  std::vector<Constraint> hc(n);
  std::vector<float> hr(g.d_radii ? n : 0);
  if (n > 0 && g.d_constraints) {
    if (!hip_ok(hipMemcpyAsync(hc.data(), g.d_constraints, sizeof(Constraint) * n, hipMemcpyDeviceToHost, s),
                "hipMemcpy(d_constraints)") ||
        (g.d_radii && !hip_ok(hipMemcpyAsync(hr.data(), g.d_radii, sizeof(float) * n, hipMemcpyDeviceToHost, s),
                              "hipMemcpy(d_radii)")) ||
        !hip_ok(hipStreamSynchronize(s), "synchronize"))
      return false;
  }
  group_stats(hc.data(), g.d_radii ? hr.data() : nullptr, n, info);
  info->serial = ++g_transient_serial;
  // Renders on other streams may still read the bins about to be dropped.
  if (g.d_constraints && !(hip_ok(hipDeviceSynchronize(), "synchronize"))) return false;
  if (g.d_constraints) mfma_forget_group(g.d_constraints);
  return true;
}

// Validated description of every group of a render (one MFMA job spec per group).
static bool collect_groups(MeshlessDataset* ds, VisConfig* c, double2* acc, std::vector<MfmaJobSpec>* specs) {
  specs->clear();
  ViewParams vp;
  vp.cx = c->_cutoff_frequency.x;
  vp.cy = c->_cutoff_frequency.y;
  vp.sx = c->step_size.x;
  vp.sy = c->step_size.y;
  vp.phi_mode = c->phi_mode;
  const double3 u = make_double3(c->u_axis.x, c->u_axis.y, c->u_axis.z);
  const double3 v = make_double3(c->v_axis.x, c->v_axis.y, c->v_axis.z);
  for (int gi = 0; gi < ds->number_of_groups; ++gi) {
    const Group& g = ds->groups[gi];
    if (g.number_of_terms > 0 && g.d_constraints == nullptr) {
      set_error(VIS_ERROR_NOT_REGISTERED, "group " + std::to_string(gi) + " is not registered");
      return false;
    }
    MfmaJobSpec sp;
    sp.g = &g;
    if (!group_info_for(g, c->_plan->stream, &sp.info)) return false;
    sp.n = std::max(0, std::min(sp.info.n, g.d_number_of_terms > 0 ? g.d_number_of_terms : sp.info.n));
    sp.vp = vp;
    sp.vp.basis = g.basis_function_id;
    if (sp.vp.basis < SPH || sp.vp.basis > WENDLAND_D3_C2) {
      set_error(VIS_ERROR_INVALID_VALUE, "unknown basis_function_id");
      return false;
    }
    sp.u = u;
    sp.v = v;
    sp.c = c;
    sp.acc = acc;
    sp.first = gi == 0;
    specs->push_back(sp);
  }
  return true;
}

// Plain FP32 direct sum of one group, added into acc.
static bool plain_group(VisPlan* p, const MfmaJobSpec& sp, int nfreq) {
  const Group& g = *sp.g;
  const int n = sp.n;
  const int n_padded = ((n + 255) / 256) * 256;
  if (!p->packets.reserve(sizeof(Packet) * std::max(n_padded, 256))) return false;
  if (!hip_ok(launch_prep_packets(g.d_constraints, g.d_radii, n, n_padded, sp.c->step_size.x, sp.c->step_size.y, sp.u,
                                  sp.v, cull_factor(sp.c, g.basis_function_id), p->packets.as<Packet>(), p->stream),
              "prep_packets launch"))
    return false;
  const int ntiles = n_padded / 256;
  const int freq_tiles = ((2 * sp.vp.cx + 63) / 64) * ((sp.vp.cy + 31) / 32);
  const int ksplit = choose_ksplit(ntiles, freq_tiles);
  if (!p->partial.reserve(sizeof(double2) * static_cast<size_t>(ksplit) * nfreq)) return false;
  EventPair* ev = p->timing ? next_events(p) : nullptr;
  if (ev && !hip_ok(hipEventRecord(ev->first, p->stream), "hipEventRecord")) return false;
  const bool uniform = sp.info.uniform_radius;
  const float radius = sp.info.has_radii ? sp.info.radius : 1.0f;
  if (!hip_ok(launch_direct_sum(p->packets.as<Packet>(), n_padded, ksplit, uniform ? 1 : 0, radius, sp.vp,
                                p->partial.as<double2>(), p->stream),
              "direct_sum launch"))
    return false;
  if (ev && !hip_ok(hipEventRecord(ev->second, p->stream), "hipEventRecord")) return false;
  return hip_ok(launch_reduce_partials(p->partial.as<double2>(), ksplit, nfreq, sp.first ? 1 : 0, sp.acc, p->stream),
                "reduce_partials launch");
}

// The spectrum of all groups: the FP64 sums into the plan's accumulator and, with finalize,
// scaled into _d_freq_image slot 0 (without: the caller's first C2R pass does that).
static bool spectrum(MeshlessDataset* ds, VisConfig* c, bool finalize = true) {
  VisPlan* p = c->_plan;
  const int nfreq = static_cast<int>(spectrum_count(c));
  if (nfreq == 0) return true;
  if (!p->acc.reserve(sizeof(double2) * nfreq)) return false;
  std::vector<MfmaJobSpec> specs;
  if (!collect_groups(ds, c, p->acc.as<double2>(), &specs)) return false;
  if (c->form != VIS_FORM_PLAIN) {
    if (!mfma_execute(specs, p->stream, &p->mfma, p->timing ? next_events(p) : nullptr)) return false;
  } else {
    for (const MfmaJobSpec& sp : specs)
      if (!plain_group(p, sp, nfreq)) return false;
  }
  if (!finalize) return true;
  return hip_ok(launch_finalize_spectrum(p->acc.as<double2>(), nfreq, c->_scale, c->_d_freq_image, p->stream),
                "finalize launch");
}


// ------------------------------------------------------------ batched render ---
// All renders of a batch go through one captured HIP graph when they share a
// stream, an MFMA form, the Phi mode and the grid (cutoff, samples, steps):
// per-job packet prep, one persistent MFMA launch per kernel variant, the
// reductions, and every render's finalize + C2R.
static bool batch_fusable(MeshlessDataset** ds, VisConfig** cf, int n) {
  const VisConfig* c0 = cf[0];
  if (c0->form == VIS_FORM_PLAIN) return false;
  for (int i = 0; i < n; ++i) {
    const VisConfig* c = cf[i];
    if (c->_plan->stream != c0->_plan->stream || c->form != c0->form || c->phi_mode != c0->phi_mode ||
        c->_cutoff_frequency.x != c0->_cutoff_frequency.x || c->_cutoff_frequency.y != c0->_cutoff_frequency.y ||
        c->step_size.x != c0->step_size.x || c->step_size.y != c0->step_size.y ||
        c->number_of_h_bins != c0->number_of_h_bins ||
        c->number_of_interpolation_nodes != c0->number_of_interpolation_nodes ||
        c->h_bin_ratio != c0->h_bin_ratio || c->tile != c0->tile || ds[i]->number_of_groups < 1)
      return false;
    // A group filled by the caller (no registry entry) is re-read at every render; a
    // captured graph could not follow it, so such a batch renders view by view.
    for (int g = 0; g < ds[i]->number_of_groups; ++g) {
      GroupInfo info;
      const Group& gr = ds[i]->groups[g];
      if (gr.d_constraints && !registry_get(gr.d_constraints, &info)) return false;
    }
    for (int j = 0; j < i; ++j)
      if (cf[j] == c) return false;  // one render per config (each owns its output)
  }
  return true;
}

static void key_put(std::vector<uint64_t>* k, const void* p) { k->push_back(reinterpret_cast<uintptr_t>(p)); }
static void key_put(std::vector<uint64_t>* k, float f) {
  uint32_t u;
  std::memcpy(&u, &f, sizeof(u));
  k->push_back(u);
}
static void key_put(std::vector<uint64_t>* k, int i) { k->push_back(static_cast<uint32_t>(i)); }

static std::vector<uint64_t> batch_key(MeshlessDataset** ds, VisConfig** cf, int n) {
  std::vector<uint64_t> k;
  key_put(&k, n);
  for (int i = 0; i < n; ++i) {
    const VisConfig* c = cf[i];
    key_put(&k, c);
    key_put(&k, ds[i]);
    for (float f : {c->step_size.x, c->step_size.y, c->u_axis.x, c->u_axis.y, c->u_axis.z, c->v_axis.x, c->v_axis.y,
                    c->v_axis.z, c->h_bin_ratio})
      key_put(&k, f);
    for (int v : {c->_cutoff_frequency.x, c->_cutoff_frequency.y, c->_number_of_samples.x, c->_number_of_samples.y,
                  c->form, c->phi_mode, c->number_of_h_bins, c->number_of_interpolation_nodes, c->tile,
                  static_cast<int>(c->cull_fully_aliased_terms)})
      key_put(&k, v);
    for (const void* q : {static_cast<const void*>(c->_d_freq_image), static_cast<const void*>(c->_d_image),
                          static_cast<const void*>(c->_d_freq_image_arranged), static_cast<const void*>(c->_plan->fft.mem.ptr), static_cast<const void*>(c->_plan->acc.ptr),
                          static_cast<const void*>(c->_plan->stream)})
      key_put(&k, q);
    key_put(&k, ds[i]->number_of_groups);
    for (int g = 0; g < ds[i]->number_of_groups; ++g) {
      const Group& gr = ds[i]->groups[g];
      // The registration serial: a group unregistered and registered again (possibly at
      // the same device addresses) has new bins, so its old graph must not be replayed.
      GroupInfo info;
      const unsigned long long serial = registry_get(gr.d_constraints, &info) ? info.serial : 0ull;
      k.push_back(serial);
      key_put(&k, gr.d_constraints);
      key_put(&k, gr.d_radii);
      key_put(&k, gr.number_of_terms);
      key_put(&k, gr.d_number_of_terms);
      key_put(&k, static_cast<int>(gr.basis_function_id));
    }
  }
  return k;
}

// Device tables of the batched prep / finish / C2R launches (uploaded before capture).
static bool upload_io_tables(const std::vector<MfmaJobSpec>& specs, const MfmaBatch& bt, VisConfig** cf, int n,
                             const std::vector<int>& render_of, BatchGraph* bg, const PrepJob** jobs,
                             const long long** begin, const int** block_job, const RenderOut** outs,
                             const JobPart** parts, long long* total) {
  const size_t nj = specs.size();
  std::vector<PrepJob> pj(nj);
  std::vector<long long> bg_begin(nj);  // first prep block of each job
  std::vector<int> bj;                  // the job of each prep block
  std::vector<JobPart> jp(nj);
  std::vector<RenderOut> ro(n);
  long long acc = 0;
  for (size_t k = 0; k < nj; ++k) {
    const MfmaJobSpec& sp = specs[k];
    const MfmaPlanned& pl = bt.plans[k];
    PrepJob& J = pj[k];
    // Binned jobs read their particles in slot order (every slot in place, zero on padding).
    const Constraint* ord = mfma_ordered(pl);
    J.c = ord ? ord : sp.g->d_constraints;
    J.radii = ord ? mfma_ordered_radii(pl) : sp.g->d_radii;
    J.out = const_cast<Packet*>(pl.job.packets);
    J.n = ord ? pl.n_slots : sp.n;
    J.n_slots = pl.n_slots;
    J.sx = sp.c->step_size.x;
    J.sy = sp.c->step_size.y;
    J.u = sp.u;
    J.v = sp.v;
    J.cull = cull_factor(sp.c, sp.g->basis_function_id);
    bg_begin[k] = acc;
    const long long blocks = (pl.n_slots + kPrepBlock - 1) / kPrepBlock;
    acc += blocks;
    bj.insert(bj.end(), static_cast<size_t>(blocks), static_cast<int>(k));
    jp[k].partial = pl.job.partial;
    jp[k].nchunks = pl.nchunks;
  }
  for (int i = 0; i < n; ++i) {
    ro[i] = RenderOut{cf[i]->_d_freq_image, cf[i]->_d_freq_image_arranged, cf[i]->_d_image, cf[i]->_scale, 0, 0, 0};
    ro[i].job_begin = static_cast<int>(std::find(render_of.begin(), render_of.end(), i) - render_of.begin());
    ro[i].job_end = ro[i].job_begin;
    while (ro[i].job_end < static_cast<int>(nj) && render_of[ro[i].job_end] == i) ++ro[i].job_end;
  }
  const size_t b_pj = sizeof(PrepJob) * nj, b_bg = sizeof(long long) * nj, b_ro = sizeof(RenderOut) * n,
               b_jp = sizeof(JobPart) * nj, b_bj = sizeof(int) * bj.size();
  auto up = [](size_t b) { return (b + 255) / 256 * 256; };
  const size_t o_bg = up(b_pj), o_ro = o_bg + up(b_bg), o_jp = o_ro + up(b_ro), o_bj = o_jp + up(b_jp),
               all = o_bj + up(b_bj);
  // The tables of the previous batch (an eager run or a replay of its graph, queued on the
  // configs' stream, possibly a non-blocking one) may still be read: wait for them before the
  // buffer is reused.
  if (!hip_ok(hipStreamSynchronize(cf[0]->_plan->stream), "synchronize")) return false;
  if (!bg->io.reserve(all)) return false;
  std::vector<char> h(all, 0);
  std::memcpy(h.data(), pj.data(), b_pj);
  std::memcpy(h.data() + o_bg, bg_begin.data(), b_bg);
  std::memcpy(h.data() + o_ro, ro.data(), b_ro);
  std::memcpy(h.data() + o_jp, jp.data(), b_jp);
  if (b_bj) std::memcpy(h.data() + o_bj, bj.data(), b_bj);
  if (!hip_ok(hipMemcpy(bg->io.ptr, h.data(), all, hipMemcpyHostToDevice), "upload batch tables")) return false;
  char* base = static_cast<char*>(bg->io.ptr);
  *jobs = reinterpret_cast<const PrepJob*>(base);
  *begin = reinterpret_cast<const long long*>(base + o_bg);
  *outs = reinterpret_cast<const RenderOut*>(base + o_ro);
  *parts = reinterpret_cast<const JobPart*>(base + o_jp);
  *block_job = reinterpret_cast<const int*>(base + o_bj);
  *total = acc;
  return true;
}

// Plans the batch and enqueues its launches: captured into bg's graph (capture = true), or run
// at once on the configs' stream.  Either way the same launches in the same order.
static bool enqueue_batch(MeshlessDataset** ds, VisConfig** cf, int n, BatchGraph* bg, bool capture) {
  bg->drop();
  hipStream_t s = cf[0]->_plan->stream;
  std::vector<MfmaJobSpec> specs, one;
  std::vector<int> render_of;  // render of every job
  bool same_grid = true;       // every render has the samples of render 0: batched C2R
  for (int i = 0; i < n; ++i) {
    VisConfig* c = cf[i];
    std::string why;
    if (!config_valid(c, &why)) {
      set_error(VIS_ERROR_INVALID_CONFIG, "invalid VisConfig: " + why);
      return false;
    }
    if (!c->_d_image) {
      set_error(VIS_ERROR_INVALID_VALUE, "vis_fourier_volume_rendering_batched: no _d_image (manual image not set)");
      return false;
    }
    vis_config_compute_scale(c);
    if (!c->_plan->acc.reserve(sizeof(double2) * std::max<size_t>(spectrum_count(c), 1))) return false;
    if (!collect_groups(ds[i], c, c->_plan->acc.as<double2>(), &one)) return false;
    specs.insert(specs.end(), one.begin(), one.end());
    render_of.insert(render_of.end(), one.size(), i);
    same_grid = same_grid && c->_number_of_samples.x == cf[0]->_number_of_samples.x &&
                c->_number_of_samples.y == cf[0]->_number_of_samples.y;
  }
  same_grid = same_grid && n <= 65535;  // render = grid row of the batched launches
  if (!mfma_prepare(specs, s, &bg->ws, &bg->ws.batch)) return false;
  // A batch too large for its partial slots at once runs in passes, each reduced by
  // mfma_enqueue; the batched finish reduces a single pass, so such a batch takes the
  // per-render finish.
  if (bg->ws.batch.passes > 1) same_grid = false;
  const PrepJob* pjobs = nullptr;
  const long long* pbegin = nullptr;
  const int* pblock_job = nullptr;
  const RenderOut* outs = nullptr;
  const JobPart* parts = nullptr;
  long long total = 0;
  if (same_grid &&
      !upload_io_tables(specs, bg->ws.batch, cf, n, render_of, bg, &pjobs, &pbegin, &pblock_job, &outs, &parts,
                        &total))
    return false;
  if (capture) {
    if (!bg->capture && !hip_ok(hipStreamCreateWithFlags(&bg->capture, hipStreamNonBlocking), "hipStreamCreate"))
      return false;
    if (!hip_ok(hipStreamBeginCapture(bg->capture, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture"))
      return false;
  }
  const hipStream_t q = capture ? bg->capture : s;
  bool ok;
  if (same_grid) {
    // One launch each: every job's packets, the persistent MFMA kernels, every render's
    // reduction + finalize, and the two C2R passes of every render.
    const VisConfig* c0 = cf[0];
    const int nkeys = static_cast<int>(bg->ws.batch.keys.size());
    const bool fused = nkeys <= 256;  // the prep launch clears the work-queue counters
    ok = hip_ok(launch_prep_jobs(pjobs, pbegin, pblock_job, total, q, fused ? bg->ws.batch.counters : nullptr,
                                 fused ? nkeys : 0),
                "prep launch") &&
         mfma_launch_sums(specs, bg->ws.batch, q, nullptr, fused) &&
         hip_ok(launch_batch_finish(outs, parts, n, c0->_cutoff_frequency.x, c0->_cutoff_frequency.y, q),
                "finish launch") &&
         hip_ok(launch_c2r_batch(outs, n, c0->_cutoff_frequency.x, c0->_cutoff_frequency.y, c0->_number_of_samples.x,
                                 c0->_number_of_samples.y, c0->_plan->fft, q),
                "c2r launch");
  } else {
    ok = mfma_enqueue(specs, bg->ws.batch, q, nullptr);
    for (int i = 0; ok && i < n; ++i) {
      VisConfig* c = cf[i];
      const int nfreq = static_cast<int>(spectrum_count(c));
      ok = nfreq == 0 || hip_ok(launch_finalize_c2r(c->_plan->acc.as<double2>(), c->_scale, c->_d_freq_image,
                                                    c->_cutoff_frequency.x, c->_cutoff_frequency.y,
                                                    c->_number_of_samples.x, c->_number_of_samples.y, c->_plan->fft,
                                                    c->_d_freq_image_arranged, c->_d_image, q),
                                "c2r launch");
      if (ok && nfreq == 0)
        ok = hip_ok(launch_c2r(c->_d_freq_image, c->_cutoff_frequency.x, c->_cutoff_frequency.y,
                               c->_number_of_samples.x, c->_number_of_samples.y, c->_plan->fft,
                               c->_d_freq_image_arranged, c->_d_image, q),
                    "c2r launch");
    }
  }
  if (!capture) return ok;
  hipGraph_t graph = nullptr;
  const hipError_t end = hipStreamEndCapture(bg->capture, &graph);
  if (!ok) {
    if (graph) (void)hipGraphDestroy(graph);
    return false;
  }
  if (!hip_ok(end, "hipStreamEndCapture")) return false;
  bg->graph = graph;
  if (!hip_ok(hipGraphInstantiate(&bg->exec, graph, nullptr, nullptr, 0), "hipGraphInstantiate")) {
    bg->drop();
    return false;
  }
  return true;
}

// A batch seen for the first time runs at once, without a graph (a one-off batch, such as one
// timestep of an animation, would not repay the capture and instantiation); the same batch again
// is captured, and replayed while it stays the same.
static void render_batch_graph(MeshlessDataset** ds, VisConfig** cf, int n) {
  VisPlan* p0 = cf[0]->_plan;
  BatchGraph& bg = p0->batch;
  std::vector<uint64_t> key = batch_key(ds, cf, n);
  EventPair* ev = p0->timing ? next_events(p0) : nullptr;
  if (!bg.exec || key != bg.key) {
    if (key != bg.seen) {
      if (ev && !hip_ok(hipEventRecord(ev->first, p0->stream), "hipEventRecord")) return;
      const bool ok = enqueue_batch(ds, cf, n, &bg, false);
      // The key as the same batch will present it next time: the first render of a config
      // allocates buffers the key names (the FP64 accumulator, the C2R workspace).
      bg.seen = ok ? batch_key(ds, cf, n) : std::vector<uint64_t>();
      if (ok && ev) hip_ok(hipEventRecord(ev->second, p0->stream), "hipEventRecord");
      return;
    }
    if (!enqueue_batch(ds, cf, n, &bg, true)) {
      bg.drop();
      return;
    }
    bg.key = batch_key(ds, cf, n);
    ++bg.captures;
  }
  if (ev && !hip_ok(hipEventRecord(ev->first, p0->stream), "hipEventRecord")) return;
  if (!hip_ok(hipGraphLaunch(bg.exec, p0->stream), "hipGraphLaunch")) return;
  ++bg.launches;
  if (ev) hip_ok(hipEventRecord(ev->second, p0->stream), "hipEventRecord");
}


// ------------------------------------------------------- allocation-free render ---
// vis_fourier_volume_rendering_noalloc: the plain-form render with every scratch buffer
// (per-view packets, split-K FP64 partials, FP64 spectrum accumulator) in one static device
// array per GPU, so the call itself allocates no device memory.  The array is reached through
// its address from hipGetSymbolAddress, taken once per device with that device current (in
// host code the array's name is not a device address).
constexpr size_t kNoallocScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

constexpr int kMaxScratchDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxScratchDevices] = {};

static unsigned char* scratch_for_current_device() {
  int dev = 0;
  if (!hip_ok(hipGetDevice(&dev), "hipGetDevice")) return nullptr;
  if (dev < 0 || dev >= kMaxScratchDevices) {
    set_error(VIS_ERROR_UNSUPPORTED, "vis_fourier_volume_rendering_noalloc: device index out of range");
    return nullptr;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    if (!hip_ok(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)")) return nullptr;
    g_scratch_ptr[dev] = ptr;
  }
  return static_cast<unsigned char*>(g_scratch_ptr[dev]);
}

static size_t align256(size_t b) { return (b + 255) / 256 * 256; }

static bool render_noalloc(MeshlessDataset* ds, VisConfig* c) {
  VisPlan* p = c->_plan;
  const int nfreq = static_cast<int>(spectrum_count(c));
  const int ngroups = ds->number_of_groups;
  if (ngroups < 1 || nfreq == 0) {
    if (!hip_ok(launch_zero_words(c->_d_freq_image, 2 * static_cast<size_t>(nfreq), p->stream), "zero spectrum"))
      return false;
  } else {
    // Sizes first: the largest group's packets, the largest split-K partials.
    const int freq_tiles = ((2 * c->_cutoff_frequency.x + 63) / 64) * ((c->_cutoff_frequency.y + 31) / 32);
    size_t max_padded = 0, max_ksplit = 1;
    for (int gi = 0; gi < ngroups; ++gi) {
      const Group& g = ds->groups[gi];
      if (g.basis_function_id < SPH || g.basis_function_id > WENDLAND_D3_C2) {
        set_error(VIS_ERROR_INVALID_VALUE, "unknown basis_function_id");
        return false;
      }
      GroupInfo info;
      if (g.number_of_terms > 0 && (!g.d_constraints || !registry_get(g.d_constraints, &info))) {
        set_error(VIS_ERROR_NOT_REGISTERED,
                  "vis_fourier_volume_rendering_noalloc: every group must be registered (vis_register_meshless_dataset)");
        return false;
      }
      const size_t n = g.number_of_terms > 0 ? static_cast<size_t>(info.n) : 0;
      const size_t padded = (n + 255) / 256 * 256;
      max_padded = std::max(max_padded, padded);
      max_ksplit = std::max<size_t>(max_ksplit, choose_ksplit(static_cast<int>(padded / 256), freq_tiles));
    }
    const size_t acc_b = align256(sizeof(double2) * nfreq);
    const size_t pk_b = align256(sizeof(Packet) * std::max<size_t>(max_padded, 256));
    const size_t part_one = sizeof(double2) * static_cast<size_t>(nfreq);
    if (acc_b + pk_b + part_one > kNoallocScratchBytes) {
      set_error(VIS_ERROR_INVALID_VALUE,
                "vis_fourier_volume_rendering_noalloc: particles and spectrum need more than its 64 MiB scratch "
                "(16 B per particle of the largest group + 32 B per frequency at least)");
      return false;
    }
    // As many split-K slots as the render would take, fewer when the scratch holds fewer.
    const size_t ksplit_cap = (kNoallocScratchBytes - acc_b - pk_b) / part_one;
    unsigned char* base = scratch_for_current_device();
    if (!base) return false;
    double2* acc = reinterpret_cast<double2*>(base);
    Packet* packets = reinterpret_cast<Packet*>(base + acc_b);
    double2* partial = reinterpret_cast<double2*>(base + acc_b + pk_b);
    ViewParams vp;
    vp.cx = c->_cutoff_frequency.x;
    vp.cy = c->_cutoff_frequency.y;
    vp.sx = c->step_size.x;
    vp.sy = c->step_size.y;
    vp.phi_mode = c->phi_mode;
    const double3 u = make_double3(c->u_axis.x, c->u_axis.y, c->u_axis.z);
    const double3 v = make_double3(c->v_axis.x, c->v_axis.y, c->v_axis.z);
    for (int gi = 0; gi < ngroups; ++gi) {
      const Group& g = ds->groups[gi];
      GroupInfo info;
      if (g.number_of_terms > 0) registry_get(g.d_constraints, &info);
      const int n = g.number_of_terms > 0 ? info.n : 0;
      const int n_padded = (n + 255) / 256 * 256;
      const int ksplit = static_cast<int>(
          std::min<size_t>(ksplit_cap, choose_ksplit(n_padded / 256, freq_tiles)));
      vp.basis = g.basis_function_id;
      const float radius = info.has_radii ? info.radius : 1.0f;
      if (!hip_ok(launch_prep_packets(g.d_constraints, g.d_radii, n, n_padded, c->step_size.x, c->step_size.y, u, v,
                                      cull_factor(c, g.basis_function_id), packets, p->stream),
                  "prep_packets launch") ||
          !hip_ok(launch_direct_sum(packets, n_padded, ksplit, info.uniform_radius ? 1 : 0, radius, vp, partial,
                                    p->stream),
                  "direct_sum launch") ||
          !hip_ok(launch_reduce_partials(partial, ksplit, nfreq, gi == 0 ? 1 : 0, acc, p->stream),
                  "reduce_partials launch"))
        return false;
    }
    // The first C2R pass finalizes the spectrum from acc as it gathers it.
    return hip_ok(launch_finalize_c2r(acc, c->_scale, c->_d_freq_image, c->_cutoff_frequency.x,
                                      c->_cutoff_frequency.y, c->_number_of_samples.x, c->_number_of_samples.y, p->fft,
                                      c->_d_freq_image_arranged, c->_d_image, p->stream),
                  "c2r launch");
  }
  return hip_ok(launch_c2r(c->_d_freq_image, c->_cutoff_frequency.x, c->_cutoff_frequency.y, c->_number_of_samples.x,
                           c->_number_of_samples.y, p->fft, c->_d_freq_image_arranged, c->_d_image,
                           p->stream),
                "c2r launch");
}

}  // namespace mv

using namespace mv;

extern "C" {

VisConfig* vis_config_create(bool automatic_d_image, float2 step_size, int2 cutoff_frequency, float3 u_axis,
                             float3 v_axis, int2 number_of_samples, int block_length, int number_of_partial_sums) {
  clear_error();
  if (cutoff_frequency.x < 0 || cutoff_frequency.y < 0 || number_of_samples.x <= 0 || number_of_samples.y <= 0 ||
      number_of_partial_sums < 1) {
    set_error(VIS_ERROR_INVALID_VALUE, "vis_config_create: negative cutoff, non-positive samples or partial sums < 1");
    return nullptr;
  }
  VisConfig* c = static_cast<VisConfig*>(std::calloc(1, sizeof(VisConfig)));
  if (!c) {
    set_error(VIS_ERROR_OUT_OF_MEMORY, "vis_config_create: host allocation failed");
    return nullptr;
  }
  c->step_size = step_size;
  c->u_axis = u_axis;
  c->v_axis = v_axis;
  c->cull_fully_aliased_terms = false;
  c->_number_of_samples = number_of_samples;
  c->_cutoff_frequency = cutoff_frequency;
  c->block_length = block_length;
  c->_number_of_partial_sums = number_of_partial_sums;
  c->_automatic_d_image = automatic_d_image;
  c->form = VIS_FORM_PLAIN;
  c->phi_mode = VIS_PHI_REFERENCE;
  c->_plan = new VisPlan();
  (void)hipGetDevice(&c->_plan->device);
  if (!alloc_freq_image(c) || !alloc_samples(c)) {
    const int code = vis_get_last_error_code();
    const std::string msg = vis_get_last_error();
    vis_config_destroy(c);
    set_error(code, msg);
    return nullptr;
  }
  vis_config_compute_scale(c);
  return c;
}

VisConfig* vis_config_get_default() {
  return vis_config_create(true, make_float2(0.1f, 0.1f), make_int2(128, 128), make_float3(1.0f, 0.0f, 0.0f),
                           make_float3(0.0f, 1.0f, 0.0f), make_int2(512, 512), 256, 1);
}

bool vis_config_check(VisConfig* vis_config) {
  if (!vis_config) return false;
  std::string why;
  return config_valid(vis_config, &why);
}

void vis_config_manual_d_image(VisConfig* vis_config, float* d_image) {
  if (!vis_config) return;
  vis_config->_d_image = d_image;
}

void vis_config_change_number_of_samples(VisConfig* c, int2 number_of_samples) {
  clear_error();
  if (!c) return;
  if (number_of_samples.x <= 0 || number_of_samples.y <= 0) {
    set_error(VIS_ERROR_INVALID_VALUE, "vis_config_change_number_of_samples: non-positive size");
    return;
  }
  (void)hipStreamSynchronize(c->_plan->stream);
  c->_number_of_samples = number_of_samples;
  alloc_samples(c);
}

void vis_config_change_cutoff_frequency(VisConfig* c, int2 cutoff_frequency) {
  clear_error();
  if (!c) return;
  if (cutoff_frequency.x < 0 || cutoff_frequency.y < 0) {
    set_error(VIS_ERROR_INVALID_VALUE, "vis_config_change_cutoff_frequency: negative cutoff");
    return;
  }
  (void)hipStreamSynchronize(c->_plan->stream);
  c->_cutoff_frequency = cutoff_frequency;
  alloc_freq_image(c);
}

void vis_config_change_number_of_partial_sums(VisConfig* c, int number_of_partial_sums) {
  clear_error();
  if (!c) return;
  if (number_of_partial_sums < 1) {
    set_error(VIS_ERROR_INVALID_VALUE, "vis_config_change_number_of_partial_sums: value < 1");
    return;
  }
  (void)hipStreamSynchronize(c->_plan->stream);
  c->_number_of_partial_sums = number_of_partial_sums;
  alloc_freq_image(c);
}

void vis_config_destroy(VisConfig* c) {
  if (!c) return;
  if (c->_plan) (void)hipStreamSynchronize(c->_plan->stream);
  if (c->_automatic_d_image && c->_d_image) (void)hipFree(c->_d_image);
  if (c->_d_freq_image_arranged) (void)hipFree(c->_d_freq_image_arranged);
  if (c->_d_freq_image) (void)hipFree(c->_d_freq_image);
  plan_destroy(c->_plan);
  std::free(c);
}

void vis_config_compute_scale(VisConfig* c) {
  if (!c) return;
  c->_scale = c->step_size.x * c->step_size.y;
}

// Registration: zero-padded AoS copies, as load_into_device (meshless_vis.cu:264-290).
// Synchronous: copies on the config's stream, then waits for it.  Asynchronous:
// copies on the plan's copy stream; the config's stream waits for them through an
// event, the host does not.
static void register_groups(VisConfig* c, MeshlessDataset* ds, bool async, const char* what) {
  clear_error();
  if (!c || !ds) {
    set_error(VIS_ERROR_INVALID_VALUE, std::string(what) + ": null argument");
    return;
  }
  const long long multiple = static_cast<long long>(c->_number_of_partial_sums) * c->block_length;
  if (multiple <= 0) {
    set_error(VIS_ERROR_INVALID_CONFIG, std::string(what) + ": number_of_partial_sums*block_length <= 0");
    return;
  }
  VisPlan* p = c->_plan;
  if (async && !p->copy_stream &&
      !hip_ok(hipStreamCreateWithFlags(&p->copy_stream, hipStreamNonBlocking), "hipStreamCreate(copy)"))
    return;
  hipStream_t s = async ? p->copy_stream : p->stream;
  std::vector<std::pair<const void*, GroupInfo>> infos;
  int processed = 0;  // groups whose device pointers this call has set
  // On any failure the registration is undone: the copies already queued are waited for, every
  // device array this call allocated is freed (and any bins built for it dropped), and those
  // groups' device pointers are cleared (the first error stays the one reported).
  auto rollback = [&]() {
    const int code = vis_get_last_error_code();
    const std::string msg = vis_get_last_error();
    (void)hipStreamSynchronize(s);
    for (int j = 0; j < processed; ++j) {
      Group& g = ds->groups[j];
      if (g.d_constraints) {
        mfma_forget_group(g.d_constraints);
        (void)hipFree(g.d_constraints);
      }
      if (g.d_radii) (void)hipFree(g.d_radii);
      g.d_constraints = nullptr;
      g.d_radii = nullptr;
    }
    set_error(code, msg);
  };
  for (int j = 0; j < ds->number_of_groups; ++j) {
    Group& g = ds->groups[j];
    g.d_constraints = nullptr;
    g.d_radii = nullptr;
    processed = j + 1;
    const int n = g.number_of_terms;
    if (n < 0) {
      set_error(VIS_ERROR_INVALID_VALUE, "negative number_of_terms");
      return rollback();
    }
    const long long rounded = ((n + multiple - 1) / multiple) * multiple;
    g.d_number_of_terms = static_cast<int>(rounded);
    GroupInfo info;
    info.n = n;
    info.padded = static_cast<int>(rounded);
    group_stats(g.h_constraints, g.h_radii, n, &info);
    if (g.h_constraints && rounded > 0) {
      if (!device_malloc(reinterpret_cast<void**>(&g.d_constraints), sizeof(Constraint) * rounded,
                         "hipMalloc(d_constraints)"))
        return rollback();
      if (!hip_ok(hipMemsetAsync(g.d_constraints, 0, sizeof(Constraint) * rounded, s), "hipMemset") ||
          !hip_ok(hipMemcpyAsync(g.d_constraints, g.h_constraints, sizeof(Constraint) * n, hipMemcpyHostToDevice, s),
                  "hipMemcpy(d_constraints)"))
        return rollback();
    }
    if (g.h_radii && rounded > 0) {
      if (!device_malloc(reinterpret_cast<void**>(&g.d_radii), sizeof(float) * rounded, "hipMalloc(d_radii)"))
        return rollback();
      if (!hip_ok(hipMemsetAsync(g.d_radii, 0, sizeof(float) * rounded, s), "hipMemset") ||
          !hip_ok(hipMemcpyAsync(g.d_radii, g.h_radii, sizeof(float) * n, hipMemcpyHostToDevice, s), "hipMemcpy(d_radii)"))
        return rollback();
    }
    if (g.d_constraints) {
      infos.emplace_back(g.d_constraints, info);
      // The config's MFMA form bins the group from the host arrays now, on the same stream as
      // the copies: the first render needs no copy back from the device, and with asynchronous
      // registration this host work overlaps the work already queued.
      if (!mfma_prebin(g, info, c, s)) return rollback();
    }
  }
  if (async) {
    hipEvent_t ev = nullptr;
    if (!hip_ok(hipEventCreateWithFlags(&ev, hipEventDisableTiming), "hipEventCreate")) return rollback();
    const bool ok = hip_ok(hipEventRecord(ev, s), "hipEventRecord") &&
                    hip_ok(hipStreamWaitEvent(p->stream, ev, 0), "hipStreamWaitEvent");
    (void)hipEventDestroy(ev);
    if (!ok) return rollback();
  } else if (!hip_ok(hipStreamSynchronize(s), "register synchronize")) {
    return rollback();
  }
  for (const auto& e : infos) registry_put(e.first, e.second);
}

void vis_register_meshless_dataset(VisConfig* c, MeshlessDataset* ds) {
  const Range range("vis_register_meshless_dataset");
  register_groups(c, ds, false, "vis_register_meshless_dataset");
}

void vis_register_meshless_dataset_async(VisConfig* c, MeshlessDataset* ds) {
  const Range range("vis_register_meshless_dataset_async");
  register_groups(c, ds, true, "vis_register_meshless_dataset_async");
}

void vis_unregister_meshless_dataset(VisConfig* c, MeshlessDataset* ds) {
  const Range range("vis_unregister_meshless_dataset");
  clear_error();
  if (!ds) return;
  if (c && c->_plan) (void)hipStreamSynchronize(c->_plan->stream);
  for (int j = 0; j < ds->number_of_groups; ++j) {
    Group& g = ds->groups[j];
    if (g.d_constraints) {
      registry_erase(g.d_constraints);
      mfma_forget_group(g.d_constraints);
      (void)hipFree(g.d_constraints);
    }
    if (g.d_radii) (void)hipFree(g.d_radii);
    g.d_constraints = nullptr;
    g.d_radii = nullptr;
  }
}

void fourier_transform(MeshlessDataset* ds, VisConfig* c) {
  const Range range("fourier_transform");
  clear_error();
  if (!ds || !c) {
    set_error(VIS_ERROR_INVALID_VALUE, "fourier_transform: null argument");
    return;
  }
  if (ds->number_of_groups < 1) return;  // as fourier_transform.cu:163
  vis_config_compute_scale(c);
  std::string why;
  if (!config_valid(c, &why)) {
    set_error(VIS_ERROR_INVALID_CONFIG, "invalid VisConfig: " + why);
    return;
  }
  spectrum(ds, c);
}

void vis_fourier_volume_rendering(MeshlessDataset* ds, VisConfig* c) {
  const Range range("vis_fourier_volume_rendering");
  clear_error();
  if (!ds || !c) {
    set_error(VIS_ERROR_INVALID_VALUE, "vis_fourier_volume_rendering: null argument");
    return;
  }
  std::string why;
  if (!config_valid(c, &why)) {
    set_error(VIS_ERROR_INVALID_CONFIG, "invalid VisConfig: " + why);
    return;
  }
  if (!c->_d_image) {
    set_error(VIS_ERROR_INVALID_VALUE, "vis_fourier_volume_rendering: no _d_image (manual image not set)");
    return;
  }
  VisPlan* p = c->_plan;
  const size_t nfreq = spectrum_count(c);
  if (ds->number_of_groups < 1 || nfreq == 0) {
    // No groups: the reference leaves the spectrum as it was; an empty dataset renders zero here.
    if (!hip_ok(launch_zero_words(c->_d_freq_image, 2 * nfreq, p->stream), "zero spectrum")) return;
    hip_ok(launch_c2r(c->_d_freq_image, c->_cutoff_frequency.x, c->_cutoff_frequency.y, c->_number_of_samples.x,
                      c->_number_of_samples.y, p->fft, c->_d_freq_image_arranged, c->_d_image, p->stream),
           "c2r launch");
    return;
  }
  vis_config_compute_scale(c);
  if (!spectrum(ds, c, false)) return;
  // The first C2R pass finalizes the spectrum (scale, sign) from the FP64 sums as it gathers it.
  hip_ok(launch_finalize_c2r(p->acc.as<double2>(), c->_scale, c->_d_freq_image, c->_cutoff_frequency.x,
                             c->_cutoff_frequency.y, c->_number_of_samples.x, c->_number_of_samples.y, p->fft,
                             c->_d_freq_image_arranged, c->_d_image, p->stream),
         "c2r launch");
}

void vis_fourier_volume_rendering_noalloc(MeshlessDataset* ds, VisConfig* c) {
  const Range range("vis_fourier_volume_rendering_noalloc");
  clear_error();
  if (!ds || !c) {
    set_error(VIS_ERROR_INVALID_VALUE, "vis_fourier_volume_rendering_noalloc: null argument");
    return;
  }
  std::string why;
  if (!config_valid(c, &why)) {
    set_error(VIS_ERROR_INVALID_CONFIG, "invalid VisConfig: " + why);
    return;
  }
  if (c->form != VIS_FORM_PLAIN) {
    set_error(VIS_ERROR_INVALID_CONFIG, "vis_fourier_volume_rendering_noalloc: renders the plain form only");
    return;
  }
  if (!c->_d_image) {
    set_error(VIS_ERROR_INVALID_VALUE, "vis_fourier_volume_rendering_noalloc: no _d_image (manual image not set)");
    return;
  }
  vis_config_compute_scale(c);
  render_noalloc(ds, c);
}

// The reference (meshless_vis.cu:357-364) registers the GL buffer object with CUDA, maps it,
// points _d_image at it, renders, unmaps and unregisters.  Here through HIP's GL interop, on
// the config's stream: the buffer is mapped on the stream (GL work on it is ordered before the
// render), the render writes the image into it, and the unmap orders the render before later GL
// work.  Deliberate difference: the config's own _d_image is restored afterwards (the
// reference left it pointing into the unmapped buffer).  The buffer must hold at least
// Nx * Ny floats.  A failure of the interop calls (no current GL context, a name that is not a
// buffer object) is reported as VIS_ERROR_HIP and renders nothing.
void vis_opengl_fourier_volume_rendering(MeshlessDataset* ds, VisConfig* c, unsigned int buffer_object) {
  const Range range("vis_opengl_fourier_volume_rendering");
  clear_error();
  if (!ds || !c) {
    set_error(VIS_ERROR_INVALID_VALUE, "vis_opengl_fourier_volume_rendering: null argument");
    return;
  }
  std::string why;
  if (!config_valid(c, &why)) {
    set_error(VIS_ERROR_INVALID_CONFIG, "invalid VisConfig: " + why);
    return;
  }
  hipStream_t s = c->_plan->stream;
  hipGraphicsResource* res = nullptr;
  if (!hip_ok(hipGraphicsGLRegisterBuffer(&res, buffer_object, hipGraphicsRegisterFlagsWriteDiscard),
              "hipGraphicsGLRegisterBuffer"))
    return;
  auto unregister = [&] { (void)hipGraphicsUnregisterResource(res); };
  if (!hip_ok(hipGraphicsMapResources(1, &res, s), "hipGraphicsMapResources")) {
    unregister();
    return;
  }
  void* ptr = nullptr;
  size_t bytes = 0;
  const size_t need = sizeof(float) * static_cast<size_t>(c->_number_of_samples.x) * c->_number_of_samples.y;
  bool ok = hip_ok(hipGraphicsResourceGetMappedPointer(&ptr, &bytes, res), "hipGraphicsResourceGetMappedPointer");
  if (ok && bytes < need) {
    set_error(VIS_ERROR_INVALID_VALUE, "vis_opengl_fourier_volume_rendering: buffer object smaller than the image");
    ok = false;
  }
  if (ok) {
    float* own = c->_d_image;
    c->_d_image = static_cast<float*>(ptr);
    vis_fourier_volume_rendering(ds, c);
    c->_d_image = own;
    ok = vis_get_last_error_code() == VIS_SUCCESS;
  }
  // Unmap and unregister whatever happened; the first error stays the one reported.
  const hipError_t e1 = hipGraphicsUnmapResources(1, &res, s);
  const hipError_t e2 = hipGraphicsUnregisterResource(res);
  if (ok) hip_ok(e1 != hipSuccess ? e1 : e2, "hipGraphicsUnmapResources / hipGraphicsUnregisterResource");
}

void vis_copy_to_host(VisConfig* c, float* h_image) {
  const Range range("vis_copy_to_host");
  clear_error();
  if (!c || !h_image || !c->_d_image) {
    set_error(VIS_ERROR_INVALID_VALUE, "vis_copy_to_host: null argument or no image");
    return;
  }
  const size_t n = static_cast<size_t>(c->_number_of_samples.x) * c->_number_of_samples.y;
  if (!hip_ok(hipMemcpyAsync(h_image, c->_d_image, sizeof(float) * n, hipMemcpyDeviceToHost, c->_plan->stream),
              "hipMemcpy(image)"))
    return;
  hip_ok(hipStreamSynchronize(c->_plan->stream), "synchronize");
}

void vis_copy_spectrum_to_host(VisConfig* c, float2* h_spectrum) {
  const Range range("vis_copy_spectrum_to_host");
  clear_error();
  if (!c || !h_spectrum) {
    set_error(VIS_ERROR_INVALID_VALUE, "vis_copy_spectrum_to_host: null argument");
    return;
  }
  const size_t n = spectrum_count(c);
  if (n == 0) return;
  if (!hip_ok(hipMemcpyAsync(h_spectrum, c->_d_freq_image, sizeof(float2) * n, hipMemcpyDeviceToHost, c->_plan->stream),
              "hipMemcpy(spectrum)"))
    return;
  hip_ok(hipStreamSynchronize(c->_plan->stream), "synchronize");
}

void vis_set_stream(VisConfig* c, hipStream_t stream) {
  if (!c) return;
  c->_plan->stream = stream;
}

hipStream_t vis_get_stream(VisConfig* c) { return c ? c->_plan->stream : nullptr; }

void vis_synchronize(VisConfig* c) {
  clear_error();
  if (!c) return;
  hip_ok(hipStreamSynchronize(c->_plan->stream), "synchronize");
}

void vis_fourier_volume_rendering_batched(MeshlessDataset** datasets, VisConfig** configs, int n) {
  const Range range("vis_fourier_volume_rendering_batched");
  clear_error();
  if (n < 0 || (n > 0 && (!datasets || !configs))) {
    set_error(VIS_ERROR_INVALID_VALUE, "vis_fourier_volume_rendering_batched: bad arguments");
    return;
  }
  if (n == 0) return;
  for (int i = 0; i < n; ++i) {
    if (!datasets[i] || !configs[i]) {
      set_error(VIS_ERROR_INVALID_VALUE, "vis_fourier_volume_rendering_batched: null dataset or config");
      return;
    }
  }
  if (!batch_fusable(datasets, configs, n)) {
    for (int i = 0; i < n; ++i) {
      vis_fourier_volume_rendering(datasets[i], configs[i]);
      if (vis_get_last_error_code() != VIS_SUCCESS) return;
    }
    return;
  }
  render_batch_graph(datasets, configs, n);
}

double vis_autotune(MeshlessDataset* ds, VisConfig* c, double tolerance) {
  const Range range("vis_autotune");
  clear_error();
  if (!ds || !c || !(tolerance > 0.0)) {
    set_error(VIS_ERROR_INVALID_VALUE, "vis_autotune: null argument or non-positive tolerance");
    return -1.0;
  }
  if (c->form == VIS_FORM_PLAIN) {
    set_error(VIS_ERROR_INVALID_CONFIG, "vis_autotune: the config's form is not an MFMA form");
    return -1.0;
  }
  std::string why;
  if (!config_valid(c, &why)) {
    set_error(VIS_ERROR_INVALID_CONFIG, "invalid VisConfig: " + why);
    return -1.0;
  }
  VisPlan* p = c->_plan;
  const size_t nfreq = spectrum_count(c);
  if (!p->acc.reserve(sizeof(double2) * std::max<size_t>(nfreq, 1))) return -1.0;
  std::vector<MfmaJobSpec> specs;
  if (!collect_groups(ds, c, p->acc.as<double2>(), &specs)) return -1.0;
  bool binned = false, gauss = false;
  for (const MfmaJobSpec& sp : specs) {
    const int k = mfma_kind(sp.info, sp.vp.basis);
    binned |= k == KIND_BINNED;
    gauss |= k == KIND_GAUSS;
  }
  vis_config_compute_scale(c);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (!hip_ok(hipEventCreate(&e0), "hipEventCreate") || !hip_ok(hipEventCreate(&e1), "hipEventCreate")) return -1.0;
  auto timed = [&](double* ms) {
    if (!spectrum(ds, c)) return false;  // warm: bins, buffers
    if (!hip_ok(hipEventRecord(e0, p->stream), "hipEventRecord")) return false;
    for (int r = 0; r < 2; ++r)
      if (!spectrum(ds, c)) return false;
    if (!hip_ok(hipEventRecord(e1, p->stream), "hipEventRecord") || !hip_ok(hipEventSynchronize(e1), "sync"))
      return false;
    float t = 0.0f;
    if (!hip_ok(hipEventElapsedTime(&t, e0, e1), "hipEventElapsedTime")) return false;
    *ms = t / 2.0;
    return true;
  };
  auto fetch = [&](std::vector<float2>* out) {
    out->resize(nfreq);
    return nfreq == 0 || (hip_ok(hipMemcpyAsync(out->data(), c->_d_freq_image, sizeof(float2) * nfreq,
                                                hipMemcpyDeviceToHost, p->stream),
                                 "hipMemcpy(spectrum)") &&
                          hip_ok(hipStreamSynchronize(p->stream), "synchronize"));
  };
  const int form = c->form, nodes0 = c->number_of_interpolation_nodes, bins0 = c->number_of_h_bins,
            tile0 = c->tile;
  const float ratio0 = c->h_bin_ratio;
  double best_ms = -1.0, best_err = 0.0;
  int best_nodes = nodes0, best_tile = tile0;
  float best_ratio = ratio0;
  bool best_ok = false;
  std::vector<float2> ref, got;
  bool ok = true;
  c->form = VIS_FORM_PLAIN;  // the reference spectrum: the plain FP32 kernel
  ok = spectrum(ds, c) && fetch(&ref);
  c->form = form;
  double ref_norm = 0.0;
  for (const float2& z : ref) ref_norm += static_cast<double>(z.x) * z.x + static_cast<double>(z.y) * z.y;
  // Times one candidate and keeps it if it is the fastest within the tolerance (or, while none
  // is, the most accurate).
  auto consider = [&](int nodes, float ratio, int tile) {
    c->number_of_interpolation_nodes = nodes;
    c->h_bin_ratio = ratio;
    c->tile = tile;
    double ms = 0.0;
    if (!(ok = timed(&ms) && fetch(&got))) return;
    double d = 0.0;
    for (size_t i = 0; i < nfreq; ++i) {
      const double dx = static_cast<double>(got[i].x) - ref[i].x, dy = static_cast<double>(got[i].y) - ref[i].y;
      d += dx * dx + dy * dy;
    }
    const double err = ref_norm > 0.0 ? std::sqrt(d / ref_norm) : std::sqrt(d);
    const bool meets = err <= tolerance;
    if (best_ms < 0.0 || (meets && (!best_ok || ms < best_ms)) || (!meets && !best_ok && err < best_err)) {
      best_ms = ms;
      best_err = err;
      best_ok = meets;
      best_nodes = nodes;
      best_ratio = ratio;
      best_tile = tile;
    }
  };
  // First the interpolation nodes and bin ratio (with the config's tile), then the tile variants
  // of the kernel that setting runs (split forms: the Gaussian / uniform-radius kernel and the
  // two-node kernel build them).
  std::vector<std::pair<int, float>> cand;
  if (binned)
    cand = {{2, 1.005f}, {2, 1.01f}, {2, 1.02f}, {3, 1.03f}, {3, 1.04f}, {3, 1.06f}, {4, 1.1f}, {4, 1.15f},
            {5, 1.2f}, {5, 1.3f}};
  else if (gauss)
    cand = {{nodes0, 1.25f}, {nodes0, 1.5f}, {nodes0, 2.0f}, {nodes0, 4.0f}};
  else
    cand = {{nodes0, ratio0}};
  c->number_of_h_bins = 0;
  for (size_t k = 0; ok && k < cand.size(); ++k) consider(cand[k].first, cand[k].second, tile0);
  const bool tiles = form != VIS_FORM_MFMA_F16 && (!binned || best_nodes == 2);
  for (int t = VIS_TILE_DEFAULT; ok && tiles && t <= VIS_TILE_QUARTER; ++t)
    if (t != tile0) consider(best_nodes, best_ratio, t);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  c->number_of_h_bins = bins0;
  c->number_of_interpolation_nodes = nodes0;
  c->h_bin_ratio = ratio0;
  c->tile = tile0;
  if (!ok) return -1.0;
  c->number_of_h_bins = 0;
  c->number_of_interpolation_nodes = best_nodes;
  c->h_bin_ratio = best_ratio;
  c->tile = best_tile;
  return best_ms;
}

void vis_debug_fail_device_allocations(int after) { g_fail_after.store(after < 0 ? -1 : after); }

void vis_set_partial_budget(long long bytes) {
  const Range range("vis_set_partial_budget");
  clear_error();
  mfma_set_partial_budget(bytes);
}

void vis_debug_batch_graph_counts(const VisConfig* c, int* captures, long long* launches) {
  clear_error();
  if (!c || !c->_plan || !captures || !launches) {
    set_error(VIS_ERROR_INVALID_VALUE, "vis_debug_batch_graph_counts: null argument");
    return;
  }
  *captures = c->_plan->batch.captures;
  *launches = c->_plan->batch.launches;
}

void vis_set_kernel_timing(VisConfig* c, int enable) {
  if (!c) return;
  c->_plan->timing = enable != 0;
}

int vis_get_kernel_timing(VisConfig* c, double* total_ms) {
  clear_error();
  if (!c) return 0;
  VisPlan* p = c->_plan;
  if (!hip_ok(hipStreamSynchronize(p->stream), "synchronize")) return 0;
  double sum = 0.0;
  for (size_t i = 0; i < p->events_used; ++i) {
    float ms = 0.0f;
    if (!hip_ok(hipEventElapsedTime(&ms, p->events[i].first, p->events[i].second), "hipEventElapsedTime")) return 0;
    sum += ms;
  }
  if (total_ms) *total_ms = sum;
  return static_cast<int>(p->events_used);
}

void vis_reset_kernel_timing(VisConfig* c) {
  if (!c) return;
  (void)hipStreamSynchronize(c->_plan->stream);
  c->_plan->events_used = 0;
}

}  // extern "C"
