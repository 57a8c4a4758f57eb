// extern "C" boundary of the MI355X winding-number engine. Declarations and
// the reference interface each entry point replaces: include/wn.h.
//
// Two families:
//  * host-buffer calls (mli_winding_gpu, wn_winding_gpu_ex): the reference's
//    contract -- the caller owns every host buffer, the library never retains
//    a pointer (src/capi.mojo:1-5). Unlike the reference, which creates a
//    DeviceContext and four device buffers per call (src/capi.mojo:366-398),
//    one stream and grow-only device buffers persist per process, guarded by a
//    mutex so concurrent callers are safe.
//  * device-buffer calls (wn_forward, wn_pack_soa, wn_batch_*, wn_backward):
//    inputs already in HBM, caller-provided workspace, caller's stream,
//    asynchronous. No allocation inside.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <set>
#include <string>
#include <vector>

#include "wn.h"
#include "wn_internal.h"

namespace {

thread_local std::string g_error;

void set_error(const std::string& msg) { g_error = msg; }

int64_t fail(int64_t code, const std::string& msg) {
  set_error(msg);
  return code;
}

constexpr int64_t kOk = 0;
constexpr int64_t kErrArg = -1;
constexpr int64_t kErrNoDevice = -2;
constexpr int64_t kErrHip = -3;
constexpr int64_t kErrWorkspace = -4;
constexpr int64_t kErrInjected = -5;

// Test hook: WN_FAULT=alloc|launch|sync makes the host-buffer path fail at that
// step, from the host, without any device-side fault.
bool injected(const char* step) {
  const char* v = std::getenv("WN_FAULT");
  return v && std::strcmp(v, step) == 0;
}

int device_count() {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return n;
}

// The strip mesh kept by the host-buffer strip path (wn_winding_gpu_mesh):
// its handle and the host bytes of the V and F it was built from.
struct HostStripCache {
  int64_t mesh = 0;
  int fdtype = 0;
  int64_t nv = 0;
  int64_t nf = 0;
  std::vector<char> F;
  std::vector<double> V;
};

struct HostPathContext {
  std::mutex mu;
  HostStripCache strips;
  bool ready = false;
  int device = -1;
  hipStream_t stream = nullptr;
  void* buf = nullptr;  // tri | q | dst | workspace, grow-only
  size_t cap = 0;
};

// The host-buffer entries (mli_winding_gpu, wn_winding_gpu_ex / _vf /
// _mesh) run on one of kHostContexts contexts, each with its own stream,
// grow-only device buffer and kept strip mesh, so that callers on several
// threads overlap their copies and kernels instead of queueing on one lock
// (as many as the device's default hardware queues).
constexpr int kHostContexts = 4;

HostPathContext* host_contexts() {
  static HostPathContext ctx[kHostContexts];
  return ctx;
}

// A context for the calling thread, held for the call: the one the thread
// used last (its kept strip mesh and buffers) when free, else any free one,
// else the thread's own once it is released. A thread's first context is
// dealt round robin, so concurrent first calls spread out.
class HostLease {
 public:
  HostLease() {
    static std::atomic<int> deal{0};
    thread_local int last = -1;
    if (last < 0) last = deal.fetch_add(1) % kHostContexts;
    HostPathContext* all = host_contexts();
    int pick = last;
    std::unique_lock<std::mutex> l(all[pick].mu, std::try_to_lock);
    for (int i = 1; i < kHostContexts && !l.owns_lock(); ++i) {
      const int k = (last + i) % kHostContexts;
      std::unique_lock<std::mutex> t(all[k].mu, std::try_to_lock);
      if (t.owns_lock()) {
        l = std::move(t);
        pick = k;
      }
    }
    if (!l.owns_lock()) {
      l = std::unique_lock<std::mutex>(all[last].mu);
      pick = last;
    }
    last = pick;
    ctx_ = &all[pick];
    lock_ = std::move(l);
  }
  HostPathContext& ctx() { return *ctx_; }

 private:
  HostPathContext* ctx_ = nullptr;
  std::unique_lock<std::mutex> lock_;
};

std::string hip_msg(const char* what, hipError_t e) {
  return std::string(what) + ": " + hipGetErrorString(e);
}

// Resident blocks per (precision, qpl) on this device, cached; 0 = unknown.
int resident_blocks(int precision, int qpl) {
  static std::mutex mu;
  static int cache[3][5] = {};
  static bool asked[3][5] = {};
  if (precision < 0 || precision > 2 || qpl < 1 || qpl > 4) return 0;
  std::lock_guard<std::mutex> lock(mu);
  if (!asked[precision][qpl]) {
    asked[precision][qpl] = true;
    cache[precision][qpl] = device_count() > 0 ? wn::fwd_resident_blocks(precision, qpl) : 0;
  }
  return cache[precision][qpl];
}

// Face splits for a launch: enough (query block x split) blocks for about 24
// rounds of the device's resident blocks, so that the uneven end of the last
// round is short, capped at one 256-face tile (or strip tile) per split.
// Measured on the 28,320-face torus (strip and face kernels): 1M queries ->
// 12 splits (+4 % over 5), 65k queries -> >= 64 splits, 4k queries -> one
// tile per split; more rounds than ~24 start to cost block set-up, and so do
// one-tile splits while two-tile splits still give >= 4 rounds (65k queries:
// 56 splits beat 112).
int64_t best_splits(int64_t tiles, int64_t qblocks, int64_t resident) {
  resident = std::max<int64_t>(resident, 1);
  int64_t max_s = std::max<int64_t>(1, std::min<int64_t>(4096, tiles));
  if (tiles >= 2 && qblocks * (tiles / 2) >= 4 * resident) max_s = std::min<int64_t>(max_s, tiles / 2);
  const int64_t want = (24 * std::max<int64_t>(resident, 1) + qblocks - 1) / std::max<int64_t>(qblocks, 1);
  return std::max<int64_t>(1, std::min<int64_t>(max_s, want));
}

wn::FwdConfig make_config(int64_t nf, int64_t nq, int64_t precision, int64_t qpl,
                          int64_t splits, int64_t sort, double kappa) {
  wn::FwdConfig c;
  c.precision = (int)precision;
  c.qpl = (int)qpl;
  c.splits = (int)splits;
  c.sort = (int)sort;
  c.kappa = (float)kappa;
  const bool known =
      precision == wn::kFP64 || precision == wn::kFP32C || precision == wn::kFP64Plain;
  auto qblocks_of = [nq](int lanes) { return std::max<int64_t>(1, (nq + 256 * lanes - 1) / (256 * lanes)); };
  auto resident_of = [&](int lanes) {
    const int r = known ? resident_blocks((int)precision, lanes) : 0;
    return (int64_t)(r > 0 ? r : 2048);
  };
  // One query per lane: measured fastest for both precisions on gfx950
  // (full 8-wave occupancy beats the LDS-read amortisation of 2 or 4).
  if (c.qpl <= 0) c.qpl = 1;
  if (c.splits <= 0 && c.qpl >= 1 && c.qpl <= 4)
    c.splits = (int)best_splits((nf + wn::kFaceTile - 1) / wn::kFaceTile, qblocks_of(c.qpl),
                                resident_of(c.qpl));
  if (c.sort < 0) c.sort = (c.precision == wn::kFP32C && nq >= 4096) ? 1 : 0;
  if (kappa < 0.0) c.kappa = 1.0f;
  return c;
}

// The mesh of a host-buffer call: either the face SoA (tri, 9 x nf float64,
// packed by the caller as python/igl/__init__.py:248-250 does) or V (nv x 3
// float64) and F (nf x 3; fdtype 0 = int64, 1 = int32), uploaded as they are
// and gathered into the SoA on the device (half the bytes of the SoA for a
// closed mesh, and no host gather).
struct HostMesh {
  bool gather = false;  // false: tri; true: V and F
  const double* tri = nullptr;
  const double* V = nullptr;
  const void* F = nullptr;
  int64_t nv = 0;
  int fdtype = 0;
};

// First face with a corner outside [0, nv), or -1.
template <typename FT>
int64_t first_bad_face(const FT* F, int64_t nv, int64_t nf) {
  for (int64_t f = 0; f < nf; ++f)
    for (int c = 0; c < 3; ++c) {
      const int64_t vi = (int64_t)F[3 * f + c];
      if (vi < 0 || vi >= nv) return f;
    }
  return -1;
}

int64_t host_path(const HostMesh& mesh, const double* q, double* dst, int64_t nf, int64_t nq,
                  const wn::FwdConfig& cfg) {
  if (nq < 0 || nf < 0 || mesh.nv < 0) return fail(kErrArg, "negative size");
  if (nq == 0) return kOk;
  const bool gather = mesh.gather;
  if (nf > 0 && !gather && !mesh.tri) return fail(kErrArg, "null tri");
  if (nf > 0 && gather) {
    if (!mesh.V || !mesh.F) return fail(kErrArg, "null V or F");
    if (mesh.fdtype != 0 && mesh.fdtype != 1) return fail(kErrArg, "fdtype must be 0 (int64) or 1 (int32)");
    const int64_t bad = mesh.fdtype == 0
                            ? first_bad_face((const int64_t*)mesh.F, mesh.nv, nf)
                            : first_bad_face((const int32_t*)mesh.F, mesh.nv, nf);
    if (bad >= 0) return fail(kErrArg, "face " + std::to_string(bad) + " has a corner outside V");
  }
  if (!q || !dst) return fail(kErrArg, "null query or output buffer");
  if (const char* why = wn::fwd_check(nf, nq, cfg)) return fail(kErrArg, why);
  HostLease lease;
  HostPathContext& ctx = lease.ctx();
  hipError_t e;
  if (!ctx.ready) {
    if (device_count() < 1) return fail(kErrNoDevice, "no HIP device available");
    e = hipGetDevice(&ctx.device);
    if (e != hipSuccess) return fail(kErrHip, hip_msg("hipGetDevice", e));
    e = hipStreamCreateWithFlags(&ctx.stream, hipStreamNonBlocking);
    if (e != hipSuccess) return fail(kErrHip, hip_msg("hipStreamCreate", e));
    ctx.ready = true;
  }
  e = hipSetDevice(ctx.device);
  if (e != hipSuccess) return fail(kErrHip, hip_msg("hipSetDevice", e));

  const size_t tri_b = wn::align_up((size_t)nf * 9 * sizeof(double));
  const size_t q_b = wn::align_up((size_t)nq * 3 * sizeof(double));
  const size_t dst_b = wn::align_up((size_t)nq * sizeof(double));
  const size_t ws_b = wn::fwd_workspace_bytes(nf, nq, cfg);
  const size_t v_b = gather ? wn::align_up((size_t)mesh.nv * 3 * sizeof(double)) : 0;
  const size_t f_b = gather ? wn::align_up((size_t)nf * 3 * (mesh.fdtype == 0 ? 8 : 4)) : 0;
  const size_t need = tri_b + q_b + dst_b + ws_b + v_b + f_b;
  if (injected("alloc")) return fail(kErrInjected, "injected allocation failure (WN_FAULT=alloc)");
  if (need > ctx.cap) {
    if (ctx.buf) {
      (void)hipStreamSynchronize(ctx.stream);
      (void)hipFree(ctx.buf);
      ctx.buf = nullptr;
      ctx.cap = 0;
    }
    e = hipMalloc(&ctx.buf, need);
    if (e != hipSuccess) {
      ctx.buf = nullptr;
      return fail(kErrHip, hip_msg("hipMalloc", e));
    }
    ctx.cap = need;
  }
  // failures after the first copy is queued wait for the stream, so that no
  // copy still reads or writes the caller's host buffers after the return
  auto fail_queued = [&ctx](int64_t code, const std::string& msg) {
    (void)hipStreamSynchronize(ctx.stream);
    return fail(code, msg);
  };
  char* base = (char*)ctx.buf;
  double* d_tri = (double*)base;
  double* d_q = (double*)(base + tri_b);
  double* d_dst = (double*)(base + tri_b + q_b);
  void* d_ws = base + tri_b + q_b + dst_b;

  if (nf > 0 && !gather) {
    e = hipMemcpyAsync(d_tri, mesh.tri, (size_t)nf * 9 * sizeof(double), hipMemcpyHostToDevice,
                       ctx.stream);
    if (e != hipSuccess) return fail_queued(kErrHip, hip_msg("hipMemcpy tri", e));
  } else if (nf > 0) {
    char* d_v = base + tri_b + q_b + dst_b + ws_b;
    char* d_f = d_v + v_b;
    e = hipMemcpyAsync(d_v, mesh.V, (size_t)mesh.nv * 3 * sizeof(double), hipMemcpyHostToDevice,
                       ctx.stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(d_f, mesh.F, (size_t)nf * 3 * (mesh.fdtype == 0 ? 8 : 4),
                         hipMemcpyHostToDevice, ctx.stream);
    if (e != hipSuccess) return fail_queued(kErrHip, hip_msg("hipMemcpy V/F", e));
    // indices were checked above: the gather reads only inside V
    e = wn::pack_soa_launch(d_v, d_f, mesh.nv, nf, 0, mesh.fdtype, d_tri, nullptr, ctx.stream);
    if (e != hipSuccess) return fail_queued(kErrHip, hip_msg("pack launch", e));
  }
  e = hipMemcpyAsync(d_q, q, (size_t)nq * 3 * sizeof(double), hipMemcpyHostToDevice, ctx.stream);
  if (e != hipSuccess) return fail_queued(kErrHip, hip_msg("hipMemcpy q", e));
  if (injected("launch"))
    return fail_queued(kErrInjected, "injected launch failure (WN_FAULT=launch)");
  e = wn::fwd_launch(d_tri, d_q, d_dst, nf, nq, cfg, d_ws, ws_b, ctx.stream);
  if (e != hipSuccess) return fail_queued(kErrHip, hip_msg("kernel launch", e));
  e = hipMemcpyAsync(dst, d_dst, (size_t)nq * sizeof(double), hipMemcpyDeviceToHost, ctx.stream);
  if (e != hipSuccess) return fail_queued(kErrHip, hip_msg("hipMemcpy dst", e));
  e = hipStreamSynchronize(ctx.stream);
  if (e != hipSuccess) return fail(kErrHip, hip_msg("hipStreamSynchronize", e));
  if (injected("sync")) return fail(kErrInjected, "injected synchronisation failure (WN_FAULT=sync)");
  return kOk;
}

// Live batch plans; handles are checked against this set before use.
std::mutex& batch_registry_mu() {
  static std::mutex mu;
  return mu;
}

std::set<wn::BatchPlan*>& batch_registry() {
  static std::set<wn::BatchPlan*> live;
  return live;
}

wn::BatchPlan* lookup_batch(int64_t handle) {
  std::lock_guard<std::mutex> lock(batch_registry_mu());
  auto it = batch_registry().find((wn::BatchPlan*)(intptr_t)handle);
  return it == batch_registry().end() ? nullptr : *it;
}

// A device-resident mesh in strip-record form (wn_mesh_*).
struct MeshHandle {
  int64_t nf = 0;
  int64_t nv = 0;
  int64_t nrec = 0;
  void* d_rec = nullptr;     // one allocation: records, FP64 sweep records, vertex ids, flags
  wn::StripRec64* d_rec64 = nullptr;
  int32_t* d_idx = nullptr;
  int8_t* d_flag = nullptr;
  // boundary edges (net directed, from F) as record index pairs + weights:
  // the FP32-compensated forward's boundary correction
  int64_t nbe = 0;
  int32_t* d_bpairs = nullptr;
  double* d_bw = nullptr;

  wn::RibbonSrc boundary() const {
    wn::RibbonSrc b;
    b.pts = (const double*)d_rec;
    b.ids = d_bpairs;
    b.w = d_bw;
    b.n = nbe;
    return b;
  }
};

// Device copy of a mesh's strip topology (and its records when V is given on
// the host): one allocation, records first.
int64_t mesh_upload(MeshHandle* m, const double* V, const void* F, int fdtype) {
  std::vector<int32_t> idx;
  std::vector<int8_t> flag;
  m->nrec = wn::strip_topology(F, m->nf, fdtype, idx, flag);
  if (m->nrec == 0) return kOk;
  std::vector<int32_t> bpairs;
  std::vector<double> bw;
  m->nbe = wn::strip_boundary(F, m->nf, fdtype, m->nv, idx.data(), m->nrec, bpairs, bw);
  const size_t rec_b = wn::align_up((size_t)m->nrec * 4 * sizeof(double));
  const size_t rec64_b = wn::align_up((size_t)m->nrec * sizeof(wn::StripRec64));
  const size_t idx_b = wn::align_up((size_t)m->nrec * sizeof(int32_t));
  const size_t flag_b = wn::align_up((size_t)m->nrec);
  const size_t bp_b = wn::align_up((size_t)m->nbe * 2 * sizeof(int32_t));
  const size_t bw_b = wn::align_up((size_t)m->nbe * sizeof(double));
  hipError_t e = hipMalloc(&m->d_rec, rec_b + rec64_b + idx_b + flag_b + bp_b + bw_b);
  if (e != hipSuccess) {
    m->d_rec = nullptr;
    return fail(kErrHip, hip_msg("wn_mesh_create", e));
  }
  m->d_rec64 = (wn::StripRec64*)((char*)m->d_rec + rec_b);
  m->d_idx = (int32_t*)((char*)m->d_rec + rec_b + rec64_b);
  m->d_flag = (int8_t*)((char*)m->d_rec + rec_b + rec64_b + idx_b);
  m->d_bpairs = (int32_t*)((char*)m->d_rec + rec_b + rec64_b + idx_b + flag_b);
  m->d_bw = (double*)((char*)m->d_rec + rec_b + rec64_b + idx_b + flag_b + bp_b);
  e = hipMemcpy(m->d_idx, idx.data(), idx.size() * sizeof(int32_t), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(m->d_flag, flag.data(), flag.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess && m->nbe > 0)
    e = hipMemcpy(m->d_bpairs, bpairs.data(), bpairs.size() * sizeof(int32_t),
                  hipMemcpyHostToDevice);
  if (e == hipSuccess && m->nbe > 0)
    e = hipMemcpy(m->d_bw, bw.data(), bw.size() * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess && V) {
    std::vector<double> rec((size_t)(4 * m->nrec));
    for (int64_t k = 0; k < m->nrec; ++k) {
      const int64_t v = idx[(size_t)k];
      rec[(size_t)(4 * k)] = V[3 * v];
      rec[(size_t)(4 * k + 1)] = V[3 * v + 1];
      rec[(size_t)(4 * k + 2)] = V[3 * v + 2];
      rec[(size_t)(4 * k + 3)] = (double)flag[(size_t)k];
    }
    e = hipMemcpy(m->d_rec, rec.data(), rec.size() * sizeof(double), hipMemcpyHostToDevice);
  } else if (e == hipSuccess) {
    e = hipMemset(m->d_rec, 0, (size_t)m->nrec * 4 * sizeof(double));
  }
  // FP64 sweep records of the records (zero records give zero sweep records)
  if (e == hipSuccess) e = wn::strip64_prep_launch((const double4*)m->d_rec, m->nrec, m->d_rec64, 0);
  if (e == hipSuccess) e = hipStreamSynchronize(0);
  if (e != hipSuccess) {
    (void)hipFree(m->d_rec);
    m->d_rec = nullptr;
    return fail(kErrHip, hip_msg("wn_mesh_create", e));
  }
  return kOk;
}

std::mutex& mesh_registry_mu() {
  static std::mutex mu;
  return mu;
}

std::set<MeshHandle*>& mesh_registry() {
  static std::set<MeshHandle*> live;
  return live;
}

MeshHandle* lookup_mesh(int64_t handle) {
  std::lock_guard<std::mutex> lock(mesh_registry_mu());
  auto it = mesh_registry().find((MeshHandle*)(intptr_t)handle);
  return it == mesh_registry().end() ? nullptr : *it;
}

// A device-resident distance tree (wn_aabb_*).
struct AabbHandle {
  int64_t nf = 0;
  int64_t nnodes = 0;
  int64_t stack_size = 0;
  size_t bytes = 0;
  void* dev = nullptr;
};

std::mutex& aabb_registry_mu() {
  static std::mutex mu;
  return mu;
}

std::set<AabbHandle*>& aabb_registry() {
  static std::set<AabbHandle*> live;
  return live;
}

AabbHandle* lookup_aabb(int64_t handle) {
  std::lock_guard<std::mutex> lock(aabb_registry_mu());
  auto it = aabb_registry().find((AabbHandle*)(intptr_t)handle);
  return it == aabb_registry().end() ? nullptr : *it;
}

// A hierarchical winding-number tree (wn_fwn_*): host copy always, device
// copy when created with upload != 0.
struct FwnHandle {
  wn::FwnTree tree;
  int64_t nf = 0;
  size_t bytes = 0;
  void* dev = nullptr;
};

std::mutex& fwn_registry_mu() {
  static std::mutex mu;
  return mu;
}

std::set<FwnHandle*>& fwn_registry() {
  static std::set<FwnHandle*> live;
  return live;
}

FwnHandle* lookup_fwn(int64_t handle) {
  std::lock_guard<std::mutex> lock(fwn_registry_mu());
  auto it = fwn_registry().find((FwnHandle*)(intptr_t)handle);
  return it == fwn_registry().end() ? nullptr : *it;
}

// Shared argument checks of the tree entry points; nullptr if fine.
const char* aabb_args(const double* V, const void* F, int64_t nv, int64_t nf, int64_t fdtype,
                      int64_t leaf) {
  if (nv < 0 || nf < 0) return "negative size";
  if (nf > (int64_t)0x7fffffff) return "more than 2^31-1 faces";
  if (fdtype != 0 && fdtype != 1) return "fdtype must be 0 (int64) or 1 (int32)";
  if (leaf < 1) return "leaf_size must be positive";
  if (nf > 0 && (!V || !F)) return "null host buffer";
  for (int64_t i = 0; i < 3 * nf; ++i) {
    const int64_t v = fdtype == 0 ? ((const int64_t*)F)[i] : (int64_t)((const int32_t*)F)[i];
    if (v < 0 || v >= nv) return "face index is outside V";
  }
  return nullptr;
}

}  // namespace

// Scratch of wn_mesh_forward_noalloc: the device memory the strip forward
// needs beyond its inputs and outputs (split-K partials, Morton permutation,
// queries in launch order, radix-sort temporaries), as one file-scope device
// array. Every GPU the library runs on holds its own copy, at its own address,
// until the process ends. In host code `scratch` names a host-side stand-in,
// not device memory: host code uses only the per-device address that
// hipGetSymbolAddress returns (noalloc_scratch below).
constexpr size_t kNoallocScratchBytes = size_t(64) << 20;  // 64 MiB per GPU
static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

namespace {

constexpr int kMaxDevices = 64;

// Device address of `scratch` on `device` (the current device), looked up
// once per device; nullptr with the reason in `why` on failure.
void* noalloc_scratch(int device, std::string& why) {
  static std::mutex mu;
  static void* ptr[kMaxDevices] = {};
  if (device < 0 || device >= kMaxDevices) {
    why = "device ordinal outside the scratch table";
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(mu);
  if (!ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      why = hip_msg("hipGetSymbolAddress(scratch)", e);
      return nullptr;
    }
    ptr[device] = p;
  }
  return ptr[device];
}

// The kernel timer (wn_kernel_timer): event pairs of the timed launches.
struct KernelTimer {
  std::mutex mu;
  bool on = false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> done;
};

KernelTimer& kernel_timer() {
  static KernelTimer t;
  return t;
}

thread_local hipEvent_t t_open = nullptr;  // begin event of this thread's launch

bool capturing(hipStream_t s) {
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  return hipStreamIsCapturing(s, &st) == hipSuccess && st != hipStreamCaptureStatusNone;
}

void timer_clear(KernelTimer& t) {
  for (auto& p : t.done) {
    (void)hipEventDestroy(p.first);
    (void)hipEventDestroy(p.second);
  }
  t.done.clear();
}

}  // namespace

namespace wn {

void timer_begin(hipStream_t s) {
  KernelTimer& t = kernel_timer();
  {
    std::lock_guard<std::mutex> lock(t.mu);
    if (!t.on) return;
  }
  if (capturing(s)) return;
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return;
  if (hipEventRecord(e, s) != hipSuccess) {
    (void)hipEventDestroy(e);
    return;
  }
  t_open = e;
}

void timer_end(hipStream_t s) {
  if (!t_open) return;
  hipEvent_t b = t_open, e = nullptr;
  t_open = nullptr;
  if (hipEventCreate(&e) != hipSuccess || hipEventRecord(e, s) != hipSuccess) {
    (void)hipEventDestroy(b);
    if (e) (void)hipEventDestroy(e);
    return;
  }
  KernelTimer& t = kernel_timer();
  std::lock_guard<std::mutex> lock(t.mu);
  t.done.push_back({b, e});
}

}  // namespace wn

extern "C" {

int64_t wn_kernel_timer(int64_t op, double* out2) {
  KernelTimer& t = kernel_timer();
  std::lock_guard<std::mutex> lock(t.mu);
  if (op == 0 || op == 1) {
    timer_clear(t);
    t.on = op == 1;
    return kOk;
  }
  if (op != 2) return fail(kErrArg, "op must be 0 (off), 1 (on) or 2 (read)");
  if (!out2) return fail(kErrArg, "null output");
  double total = 0.0;
  int64_t n = 0;
  for (auto& p : t.done) {
    float ms = 0.0f;
    if (hipEventSynchronize(p.second) == hipSuccess &&
        hipEventElapsedTime(&ms, p.first, p.second) == hipSuccess) {
      total += ms;
      ++n;
    }
  }
  timer_clear(t);
  out2[0] = (double)n;
  out2[1] = total;
  return kOk;
}

int64_t wn_version(void) { return 2; }

const char* wn_last_error(void) { return g_error.c_str(); }

int64_t wn_device_count(void) { return device_count(); }

int64_t wn_device_free_bytes(void) {
  if (device_count() < 1) return fail(kErrNoDevice, "no HIP device available");
  size_t free_b = 0, total_b = 0;
  hipError_t e = hipMemGetInfo(&free_b, &total_b);
  if (e != hipSuccess) return fail(kErrHip, hip_msg("hipMemGetInfo", e));
  return (int64_t)free_b;
}

// mli_winding_gpu calls with at least this many queries take the FP64 strip
// kernel (the corners' vertex identity recovered from the face SoA by
// coordinate equality, then the host strip route of wn_winding_gpu_mesh,
// which keeps the mesh for the next call): the strip kernel saves ~0.93 ps
// per query-face pair over the FP64 face kernel and the host side costs
// ~0.3 us per face, so a first call breaks even at ~320k queries.
constexpr int64_t kTriStripMinQueries = 262144;

int64_t mli_winding_gpu(const double* tri, const double* q, double* dst, int64_t nf,
                        int64_t nq) {
  if (nq >= kTriStripMinQueries && nf > 0 && nf <= (int64_t)0x7fffffff && tri && q && dst) {
    std::vector<double> V;
    std::vector<int32_t> F;
    const int64_t nv = wn::soa_vertices(tri, nf, V, F);
    return wn_winding_gpu_mesh(V.data(), F.data(), nv, nf, 1, q, dst, nq, wn::kFP64) == kOk ? 1 : 0;
  }
  const wn::FwdConfig cfg = make_config(nf, nq, wn::kFP64, 0, 0, 0, 1.0);
  HostMesh mesh;
  mesh.tri = tri;
  return host_path(mesh, q, dst, nf, nq, cfg) == kOk ? 1 : 0;
}

int64_t wn_winding_gpu_ex(const double* tri, const double* q, double* dst, int64_t nf,
                          int64_t nq, int64_t precision, int64_t qpl, int64_t splits,
                          int64_t sort, double kappa) {
  const wn::FwdConfig cfg = make_config(nf, nq, precision, qpl, splits, sort, kappa);
  HostMesh mesh;
  mesh.tri = tri;
  return host_path(mesh, q, dst, nf, nq, cfg);
}

int64_t wn_winding_gpu_vf(const double* V, const void* F, int64_t nv, int64_t nf,
                          int64_t fdtype, const double* q, double* dst, int64_t nq,
                          int64_t precision) {
  const wn::FwdConfig cfg = make_config(nf, nq, precision, 0, 0, -1, -1.0);
  HostMesh mesh;
  mesh.gather = true;
  mesh.V = V;
  mesh.F = F;
  mesh.nv = nv;
  mesh.fdtype = (int)fdtype;
  return host_path(mesh, q, dst, nf, nq, cfg);
}

int64_t wn_resolve_config(int64_t nf, int64_t nq, int64_t precision, int64_t qpl,
                          int64_t splits, int64_t sort, double kappa, int64_t* out4) {
  const wn::FwdConfig c = make_config(nf, nq, precision, qpl, splits, sort, kappa);
  if (const char* why = wn::fwd_check(nf, nq, c)) return fail(kErrArg, why);
  if (out4) {
    out4[0] = c.qpl;
    out4[1] = c.splits;
    out4[2] = c.sort;
    out4[3] = c.precision;
  }
  return kOk;
}

int64_t wn_forward_workspace_bytes(int64_t nf, int64_t nq, int64_t precision, int64_t qpl,
                                   int64_t splits, int64_t sort) {
  const wn::FwdConfig c = make_config(nf, nq, precision, qpl, splits, sort, 1.0);
  if (const char* why = wn::fwd_check(nf, nq, c)) return fail(kErrArg, why);
  if (c.precision == wn::kFP32C && c.sort && device_count() < 1)
    return fail(kErrNoDevice, "no HIP device available");
  return (int64_t)wn::fwd_workspace_bytes(nf, nq, c);
}

int64_t wn_forward(const double* tri, const double* q, double* dst, int64_t nf, int64_t nq,
                   int64_t precision, int64_t qpl, int64_t splits, int64_t sort, double kappa,
                   void* workspace, int64_t workspace_bytes, void* stream) {
  const wn::FwdConfig c = make_config(nf, nq, precision, qpl, splits, sort, kappa);
  if (const char* why = wn::fwd_check(nf, nq, c)) return fail(kErrArg, why);
  if (nq == 0) return kOk;
  if ((!tri && nf > 0) || !q || !dst) return fail(kErrArg, "null device buffer");
  const size_t need = wn::fwd_workspace_bytes(nf, nq, c);
  if (workspace_bytes < 0 || (size_t)workspace_bytes < need || (need > 0 && !workspace))
    return fail(kErrWorkspace, "workspace smaller than wn_forward_workspace_bytes()");
  hipError_t e = wn::fwd_launch(tri, q, dst, nf, nq, c, workspace, (size_t)workspace_bytes,
                                (hipStream_t)stream);
  if (e != hipSuccess) return fail(kErrHip, hip_msg("wn_forward", e));
  return kOk;
}

int64_t wn_pack_soa(const void* V, const void* F, int64_t nv, int64_t nf, int64_t vdtype,
                    int64_t fdtype, double* tri, int32_t* bad_index, void* stream) {
  if (nv < 0 || nf < 0) return fail(kErrArg, "negative size");
  if ((vdtype != 0 && vdtype != 1) || (fdtype != 0 && fdtype != 1))
    return fail(kErrArg, "vdtype/fdtype must be 0 (64-bit) or 1 (32-bit)");
  if (nf > 0 && (!V || !F || !tri)) return fail(kErrArg, "null device buffer");
  hipError_t e = wn::pack_soa_launch(V, F, nv, nf, (int)vdtype, (int)fdtype, tri, bad_index,
                                     (hipStream_t)stream);
  if (e != hipSuccess) return fail(kErrHip, hip_msg("wn_pack_soa", e));
  return kOk;
}

int64_t wn_backward(const double* tri, const double* q, const double* grad_w, double* grad_q,
                    double* grad_tri, int64_t nf, int64_t nq, int64_t precision, int64_t qpl,
                    int64_t splits, void* stream) {
  if (precision != wn::kFP64) return fail(kErrArg, "backward supports precision 0 (FP64) only");
  const int lanes = (int)(qpl <= 0 ? 4 : qpl);
  int s = (int)splits;
  if (s <= 0) {
    // about 4096 blocks (several rounds of the resident blocks), at least
    // two face tiles per split
    const int64_t qblocks = (nq + 256 * lanes - 1) / (256 * lanes);
    int64_t t = (4096 + qblocks - 1) / (qblocks > 0 ? qblocks : 1);
    const int64_t max_s = nf / (2 * wn::kFaceTile);
    if (t > max_s) t = max_s;
    if (t < 1) t = 1;
    if (t > 4096) t = 4096;
    s = (int)t;
  }
  if (const char* why = wn::bwd_check(nf, nq, lanes, s)) return fail(kErrArg, why);
  if ((nq > 0 && (!q || !grad_w || !grad_q)) || (nf > 0 && (!tri || !grad_tri)))
    return fail(kErrArg, "null device buffer");
  hipError_t e = wn::bwd_launch(tri, q, grad_w, grad_q, grad_tri, nf, nq, lanes, s,
                                (hipStream_t)stream);
  if (e != hipSuccess) return fail(kErrHip, hip_msg("wn_backward", e));
  return kOk;
}

int64_t wn_boundary_edges(const void* F, int64_t nv, int64_t nf, int64_t fdtype, int32_t* edges,
                          double* weight, int64_t capacity) {
  if (nv < 0 || nf < 0) return fail(kErrArg, "negative size");
  if (nv > (int64_t)0x7fffffff) return fail(kErrArg, "more than 2^31-1 vertices");
  if (fdtype != 0 && fdtype != 1) return fail(kErrArg, "fdtype must be 0 (int64) or 1 (int32)");
  if (nf > 0 && !F) return fail(kErrArg, "null host buffer");
  for (int64_t i = 0; i < 3 * nf; ++i) {
    const int64_t v = fdtype == 0 ? ((const int64_t*)F)[i] : (int64_t)((const int32_t*)F)[i];
    if (v < 0 || v >= nv) return fail(kErrArg, "face index is outside V");
  }
  std::vector<int32_t> e;
  std::vector<double> m;
  const int64_t n = wn::boundary_edges(F, nf, (int)fdtype, e, m);
  if (edges && weight && capacity >= n) {
    std::memcpy(edges, e.data(), e.size() * sizeof(int32_t));
    std::memcpy(weight, m.data(), m.size() * sizeof(double));
  }
  return n;
}

int64_t wn_backward_edges(const void* V, int64_t nv, int64_t vdtype, const int32_t* edges,
                          const double* weight, int64_t ne, const double* q,
                          const double* grad_w, double* grad_q, double* grad_V, int64_t nq,
                          int64_t qpl, int64_t splits, void* stream) {
  if (vdtype != 0 && vdtype != 1) return fail(kErrArg, "vdtype must be 0 (float64) or 1 (float32)");
  if (nv < 0) return fail(kErrArg, "negative size");
  const int lanes = (int)(qpl <= 0 ? 4 : qpl);
  int s = (int)splits;
  if (s <= 0) {
    // about 2048 blocks, at least one edge tile per split
    const int64_t qblocks = (nq + 256 * lanes - 1) / (256 * lanes);
    int64_t t = (2048 + qblocks - 1) / (qblocks > 0 ? qblocks : 1);
    const int64_t max_s = (ne + wn::kFaceTile - 1) / wn::kFaceTile;
    if (t > max_s) t = max_s;
    if (t < 1) t = 1;
    if (t > 4096) t = 4096;
    s = (int)t;
  }
  if (const char* why = wn::bwd_edge_check(ne, nq, lanes, s)) return fail(kErrArg, why);
  if ((nq > 0 && (!q || !grad_w || !grad_q)) || (nv > 0 && !grad_V) ||
      (ne > 0 && nq > 0 && (!V || !edges || !weight)))
    return fail(kErrArg, "null device buffer");
  hipError_t e = wn::bwd_edge_launch(V, (int)vdtype, nv, edges, weight, ne, q, grad_w, grad_q,
                                     grad_V, nq, lanes, s, (hipStream_t)stream);
  if (e != hipSuccess) return fail(kErrHip, hip_msg("wn_backward_edges", e));
  return kOk;
}

int64_t wn_batch_create(const int64_t* face_off, const int64_t* query_off, int64_t n_meshes,
                        int64_t precision, int64_t qpl, int64_t chunk_faces, int64_t sort,
                        double kappa) {
  const int lanes = (int)(qpl <= 0 ? 1 : qpl);
  const int64_t chunk = chunk_faces <= 0 ? 2048 : chunk_faces;
  if (const char* why = wn::batch_check(face_off, query_off, n_meshes, (int)precision, lanes, chunk))
    return fail(kErrArg, why);
  if (device_count() < 1) return fail(kErrNoDevice, "no HIP device available");
  const int sorted = sort < 0 ? (precision == wn::kFP32C ? 1 : 0) : (sort ? 1 : 0);
  wn::BatchPlan* plan = nullptr;
  hipError_t e = wn::batch_create(face_off, query_off, n_meshes, (int)precision, lanes, chunk,
                                  sorted, kappa < 0.0 ? 1.0f : (float)kappa, &plan);
  if (e != hipSuccess) return fail(kErrHip, hip_msg("wn_batch_create", e));
  std::lock_guard<std::mutex> lock(batch_registry_mu());
  batch_registry().insert(plan);
  return (int64_t)(intptr_t)plan;
}

int64_t wn_batch_create_strips(const int64_t* rec_off, const int64_t* query_off,
                               int64_t n_meshes, int64_t chunk_tiles, int64_t sort,
                               double kappa) {
  const int64_t chunk = chunk_tiles <= 0 ? 8 : chunk_tiles;
  if (const char* why = wn::batch_strip_check(rec_off, query_off, n_meshes, chunk))
    return fail(kErrArg, why);
  if (device_count() < 1) return fail(kErrNoDevice, "no HIP device available");
  wn::BatchPlan* plan = nullptr;
  hipError_t e = wn::batch_create_strips(rec_off, query_off, n_meshes, chunk,
                                         sort < 0 ? 1 : (sort ? 1 : 0),
                                         kappa < 0.0 ? 1.0f : (float)kappa, wn::kFP32C, &plan);
  if (e != hipSuccess) return fail(kErrHip, hip_msg("wn_batch_create_strips", e));
  std::lock_guard<std::mutex> lock(batch_registry_mu());
  batch_registry().insert(plan);
  return (int64_t)(intptr_t)plan;
}

int64_t wn_batch_create_strips64(const int64_t* rec_off, const int64_t* query_off,
                                 int64_t n_meshes, int64_t chunk_tiles, int64_t sort) {
  const int64_t chunk = chunk_tiles <= 0 ? 8 : chunk_tiles;
  if (const char* why = wn::batch_strip_check(rec_off, query_off, n_meshes, chunk))
    return fail(kErrArg, why);
  if (device_count() < 1) return fail(kErrNoDevice, "no HIP device available");
  wn::BatchPlan* plan = nullptr;
  hipError_t e = wn::batch_create_strips(rec_off, query_off, n_meshes, chunk,
                                         sort < 0 ? 1 : (sort ? 1 : 0), 1.0f, wn::kFP64, &plan);
  if (e != hipSuccess) return fail(kErrHip, hip_msg("wn_batch_create_strips64", e));
  std::lock_guard<std::mutex> lock(batch_registry_mu());
  batch_registry().insert(plan);
  return (int64_t)(intptr_t)plan;
}

int64_t wn_batch_run(int64_t handle, const double* tri, const double* q, double* dst,
                     int64_t use_graph, void* stream) {
  wn::BatchPlan* plan = lookup_batch(handle);
  if (!plan) return fail(kErrArg, "unknown batch handle");
  int64_t info[6];
  wn::batch_info(plan, info);
  if ((info[4] > 0 && !tri) || (info[5] > 0 && (!q || !dst))) return fail(kErrArg, "null device buffer");
  hipError_t e = wn::batch_run(plan, tri, q, dst, use_graph != 0, (hipStream_t)stream);
  if (e != hipSuccess) return fail(kErrHip, hip_msg("wn_batch_run", e));
  return kOk;
}

int64_t wn_batch_set_boundary(int64_t handle, const int32_t* pairs, const double* weight,
                              int64_t ne, const int64_t* edge_off) {
  wn::BatchPlan* plan = lookup_batch(handle);
  if (!plan) return fail(kErrArg, "unknown batch handle");
  if (const char* why = wn::batch_boundary_check(plan, ne, edge_off)) return fail(kErrArg, why);
  if (ne > 0 && (!pairs || !weight)) return fail(kErrArg, "null device buffer");
  hipError_t e = wn::batch_set_boundary(plan, pairs, weight, ne, edge_off);
  if (e != hipSuccess) return fail(kErrHip, hip_msg("wn_batch_set_boundary", e));
  return kOk;
}

int64_t wn_batch_info(int64_t handle, int64_t* out6) {
  wn::BatchPlan* plan = lookup_batch(handle);
  if (!plan) return fail(kErrArg, "unknown batch handle");
  if (!out6) return fail(kErrArg, "null output");
  wn::batch_info(plan, out6);
  return kOk;
}

int64_t wn_batch_destroy(int64_t handle) {
  wn::BatchPlan* plan = nullptr;
  {
    std::lock_guard<std::mutex> lock(batch_registry_mu());
    auto it = batch_registry().find((wn::BatchPlan*)(intptr_t)handle);
    if (it == batch_registry().end()) return fail(kErrArg, "unknown batch handle");
    plan = *it;
    batch_registry().erase(it);
  }
  wn::batch_destroy(plan);
  return kOk;
}

int64_t wn_mesh_records(const double* V, const void* F, int64_t nv, int64_t nf, int64_t fdtype,
                        double* out, int64_t capacity) {
  if (nv < 0 || nf < 0) return fail(kErrArg, "negative size");
  if (nv > (int64_t)0x7fffffff) return fail(kErrArg, "more than 2^31-1 vertices");
  if (fdtype != 0 && fdtype != 1) return fail(kErrArg, "fdtype must be 0 (int64) or 1 (int32)");
  if (nf > 0 && (!V || !F)) return fail(kErrArg, "null host buffer");
  for (int64_t i = 0; i < 3 * nf; ++i) {
    const int64_t v = fdtype == 0 ? ((const int64_t*)F)[i] : (int64_t)((const int32_t*)F)[i];
    if (v < 0 || v >= nv) return fail(kErrArg, "face index is outside V");
  }
  std::vector<double> rec;
  const int64_t n = wn::strip_records(V, F, nf, (int)fdtype, rec);
  if (out && capacity >= n) std::memcpy(out, rec.data(), rec.size() * sizeof(double));
  return n;
}

int64_t wn_mesh_create(const double* V, const void* F, int64_t nv, int64_t nf, int64_t fdtype) {
  if (nv < 0 || nf < 0) return fail(kErrArg, "negative size");
  if (nv > (int64_t)0x7fffffff) return fail(kErrArg, "more than 2^31-1 vertices");
  if (fdtype != 0 && fdtype != 1) return fail(kErrArg, "fdtype must be 0 (int64) or 1 (int32)");
  if (nf > 0 && (!V || !F)) return fail(kErrArg, "null host buffer");
  for (int64_t i = 0; i < 3 * nf; ++i) {
    const int64_t v = fdtype == 0 ? ((const int64_t*)F)[i] : (int64_t)((const int32_t*)F)[i];
    if (v < 0 || v >= nv) return fail(kErrArg, "face index is outside V");
  }
  if (device_count() < 1) return fail(kErrNoDevice, "no HIP device available");
  MeshHandle* m = new MeshHandle();
  m->nf = nf;
  m->nv = nv;
  const int64_t st = mesh_upload(m, V, F, (int)fdtype);
  if (st != kOk) {
    delete m;
    return st;
  }
  std::lock_guard<std::mutex> lock(mesh_registry_mu());
  mesh_registry().insert(m);
  return (int64_t)(intptr_t)m;
}

int64_t wn_mesh_create_topology(const void* F, int64_t nv, int64_t nf, int64_t fdtype) {
  if (nv < 0 || nf < 0) return fail(kErrArg, "negative size");
  if (nv > (int64_t)0x7fffffff) return fail(kErrArg, "more than 2^31-1 vertices");
  if (fdtype != 0 && fdtype != 1) return fail(kErrArg, "fdtype must be 0 (int64) or 1 (int32)");
  if (nf > 0 && !F) return fail(kErrArg, "null host buffer");
  for (int64_t i = 0; i < 3 * nf; ++i) {
    const int64_t v = fdtype == 0 ? ((const int64_t*)F)[i] : (int64_t)((const int32_t*)F)[i];
    if (v < 0 || v >= nv) return fail(kErrArg, "face index is outside V");
  }
  if (device_count() < 1) return fail(kErrNoDevice, "no HIP device available");
  MeshHandle* m = new MeshHandle();
  m->nf = nf;
  m->nv = nv;
  const int64_t st = mesh_upload(m, nullptr, F, (int)fdtype);
  if (st != kOk) {
    delete m;
    return st;
  }
  std::lock_guard<std::mutex> lock(mesh_registry_mu());
  mesh_registry().insert(m);
  return (int64_t)(intptr_t)m;
}

int64_t wn_mesh_set_vertices(int64_t mesh, const void* V, int64_t nv, int64_t vdtype,
                             void* stream) {
  MeshHandle* m = lookup_mesh(mesh);
  if (!m) return fail(kErrArg, "unknown mesh handle");
  if (nv != m->nv) return fail(kErrArg, "vertex count differs from the mesh's");
  if (vdtype != 0 && vdtype != 1) return fail(kErrArg, "vdtype must be 0 (float64) or 1 (float32)");
  if (m->nrec == 0) return kOk;
  if (!V) return fail(kErrArg, "null device buffer");
  hipError_t e = wn::strip_gather_launch(V, (int)vdtype, nv, m->d_idx, m->d_flag, m->nrec,
                                         (double4*)m->d_rec, (hipStream_t)stream);
  if (e == hipSuccess)
    e = wn::strip64_prep_launch((const double4*)m->d_rec, m->nrec, m->d_rec64, (hipStream_t)stream);
  if (e != hipSuccess) return fail(kErrHip, hip_msg("wn_mesh_set_vertices", e));
  return kOk;
}

int64_t wn_mesh_info(int64_t mesh, int64_t* out3) {
  MeshHandle* m = lookup_mesh(mesh);
  if (!m) return fail(kErrArg, "unknown mesh handle");
  if (!out3) return fail(kErrArg, "null output");
  out3[0] = m->nf;
  out3[1] = m->nrec;
  out3[2] = m->nrec / 255;
  return kOk;
}

int64_t wn_mesh_resolve(int64_t mesh, int64_t nq, int64_t splits, int64_t sort, int64_t* out2) {
  MeshHandle* m = lookup_mesh(mesh);
  if (!m) return fail(kErrArg, "unknown mesh handle");
  if (nq < 0 || nq > (int64_t)0x7fffffff) return fail(kErrArg, "query count out of range");
  int64_t s = splits;
  if (s <= 0) {
    static int resident = -1;
    if (resident < 0) resident = device_count() > 0 ? wn::strip_resident_blocks() : 0;
    const int64_t qblocks = std::max<int64_t>(1, (nq + 255) / 256);
    s = best_splits(std::max<int64_t>(m->nrec / 255, 1), qblocks, resident > 0 ? resident : 2048);
  }
  if (s < 1 || s > 65535) return fail(kErrArg, "splits must be in [1, 65535]");
  const int64_t srt = sort < 0 ? (nq >= 4096 ? 1 : 0) : (sort ? 1 : 0);
  if (out2) {
    out2[0] = s;
    out2[1] = srt;
  }
  return kOk;
}

int64_t wn_mesh_workspace_bytes(int64_t mesh, int64_t nq, int64_t splits, int64_t sort) {
  int64_t cfg[2];
  const int64_t st = wn_mesh_resolve(mesh, nq, splits, sort, cfg);
  if (st != kOk) return st;
  if (cfg[1] && device_count() < 1) return fail(kErrNoDevice, "no HIP device available");
  return (int64_t)wn::strip_workspace_bytes(nq, (int)cfg[0], (int)cfg[1],
                                            lookup_mesh(mesh)->nbe);
}

int64_t wn_mesh_forward(int64_t mesh, const double* q, double* dst, int64_t nq, int64_t splits,
                        int64_t sort, double kappa, void* workspace, int64_t workspace_bytes,
                        void* stream) {
  int64_t cfg[2];
  const int64_t st = wn_mesh_resolve(mesh, nq, splits, sort, cfg);
  if (st != kOk) return st;
  if (nq == 0) return kOk;
  if (!q || !dst) return fail(kErrArg, "null device buffer");
  MeshHandle* m = lookup_mesh(mesh);
  const size_t need = wn::strip_workspace_bytes(nq, (int)cfg[0], (int)cfg[1], m->nbe);
  if (workspace_bytes < 0 || (size_t)workspace_bytes < need || (need > 0 && !workspace))
    return fail(kErrWorkspace, "workspace smaller than wn_mesh_workspace_bytes()");
  const wn::RibbonSrc b = m->boundary();
  hipError_t e = wn::strip_launch((const double4*)m->d_rec, m->nrec, q, dst, nq, (int)cfg[0],
                                  (int)cfg[1], kappa < 0.0 ? 1.0f : (float)kappa, workspace,
                                  (size_t)workspace_bytes, (hipStream_t)stream, &b);
  if (e != hipSuccess) return fail(kErrHip, hip_msg("wn_mesh_forward", e));
  return kOk;
}

int64_t wn_mesh_forward_noalloc(int64_t mesh, const double* q, double* dst, int64_t nq,
                                int64_t splits, int64_t sort, double kappa, void* stream) {
  int64_t cfg[2];
  const int64_t st = wn_mesh_resolve(mesh, nq, splits, sort, cfg);
  if (st != kOk) return st;
  if (nq == 0) return kOk;
  if (!q || !dst) return fail(kErrArg, "null device buffer");
  MeshHandle* m = lookup_mesh(mesh);
  int s = (int)cfg[0];
  const int srt = (int)cfg[1];
  if (splits <= 0 && s > 1 &&
      wn::strip_workspace_bytes(nq, s, srt, m->nbe) > kNoallocScratchBytes) {
    // default split count: the largest that fits the scratch (partials take
    // 8 bytes per query and split on top of the split-free workspace)
    const size_t base = wn::strip_workspace_bytes(nq, 1, srt, m->nbe);
    const size_t room = base + 256 < kNoallocScratchBytes ? kNoallocScratchBytes - base - 256 : 0;
    s = (int)std::max<size_t>(1, std::min<size_t>((size_t)s, room / ((size_t)nq * sizeof(double))));
  }
  const size_t need = wn::strip_workspace_bytes(nq, s, srt, m->nbe);
  if (need > kNoallocScratchBytes)
    return fail(kErrWorkspace, "wn_mesh_forward_noalloc: " + std::to_string(nq) + " queries with " +
                                   std::to_string(s) + " splits need " + std::to_string(need) +
                                   " bytes of scratch, more than the " +
                                   std::to_string(kNoallocScratchBytes) +
                                   " bytes of the static scratch; use fewer queries or splits, or "
                                   "wn_mesh_forward with a workspace");
  int device = -1;
  hipError_t e = hipGetDevice(&device);
  if (e != hipSuccess) return fail(kErrHip, hip_msg("hipGetDevice", e));
  std::string why;
  void* ws = noalloc_scratch(device, why);
  if (!ws) return fail(kErrHip, why);
  const wn::RibbonSrc b = m->boundary();
  e = wn::strip_launch((const double4*)m->d_rec, m->nrec, q, dst, nq, s, srt,
                       kappa < 0.0 ? 1.0f : (float)kappa, ws, kNoallocScratchBytes,
                       (hipStream_t)stream, &b);
  if (e != hipSuccess) return fail(kErrHip, hip_msg("wn_mesh_forward_noalloc", e));
  return kOk;
}

int64_t wn_mesh_forward64(int64_t mesh, const double* q, double* dst, int64_t nq,
                          int64_t splits, int64_t sort, void* workspace, int64_t workspace_bytes,
                          void* stream) {
  int64_t cfg[2];
  const int64_t st = wn_mesh_resolve(mesh, nq, splits, sort, cfg);
  if (st != kOk) return st;
  if (nq == 0) return kOk;
  if (!q || !dst) return fail(kErrArg, "null device buffer");
  MeshHandle* m = lookup_mesh(mesh);
  const size_t need = wn::strip_workspace_bytes(nq, (int)cfg[0], (int)cfg[1]);
  if (workspace_bytes < 0 || (size_t)workspace_bytes < need || (need > 0 && !workspace))
    return fail(kErrWorkspace, "workspace smaller than wn_mesh_workspace_bytes()");
  hipError_t e = wn::strip64_launch(m->d_rec64, (const double4*)m->d_rec, m->nrec, q, dst, nq,
                                    (int)cfg[0], (int)cfg[1], workspace, (size_t)workspace_bytes,
                                    (hipStream_t)stream);
  if (e != hipSuccess) return fail(kErrHip, hip_msg("wn_mesh_forward64", e));
  return kOk;
}

int64_t wn_mesh_destroy(int64_t mesh) {
  MeshHandle* m = nullptr;
  {
    std::lock_guard<std::mutex> lock(mesh_registry_mu());
    auto it = mesh_registry().find((MeshHandle*)(intptr_t)mesh);
    if (it == mesh_registry().end()) return fail(kErrArg, "unknown mesh handle");
    m = *it;
    mesh_registry().erase(it);
  }
  if (m->d_rec) {
    (void)hipDeviceSynchronize();
    (void)hipFree(m->d_rec);
  }
  delete m;
  return kOk;
}

// Host-buffer strip path. The strip mesh of the last call is kept on the
// device (HostStripCache, under the host path's mutex) with host copies of
// the V and F it was built from: a call whose F is byte-for-byte the cached
// one skips the host stripifier and the upload (the topology is reused), and
// when V differs only the records are regathered on the device from the new
// V (wn_mesh_set_vertices). The comparison is exact (memcmp), so a hit gives
// the results a fresh mesh would, bit for bit. The queries, output and
// workspace use the host path's stream and grow-only buffer. Same contract
// as wn_winding_gpu_ex: synchronous, returns 0 or a negative status.
int64_t wn_winding_gpu_mesh(const double* V, const void* F, int64_t nv, int64_t nf,
                            int64_t fdtype, const double* q, double* dst, int64_t nq,
                            int64_t precision) {
  if (precision != wn::kFP64 && precision != wn::kFP32C)
    return fail(kErrArg, "precision must be 0 (FP64) or 1 (FP32-compensated)");
  if (nq < 0 || nv < 0 || nf < 0) return fail(kErrArg, "negative size");
  if (nq == 0) return kOk;
  if (!q || !dst) return fail(kErrArg, "null query or output buffer");
  if (fdtype != 0 && fdtype != 1) return fail(kErrArg, "fdtype must be 0 (int64) or 1 (int32)");
  if (nf > 0 && (!V || !F)) return fail(kErrArg, "null host buffer");
  HostLease lease;
  HostPathContext& ctx = lease.ctx();
  HostStripCache& cache = ctx.strips;
  const size_t f_bytes = (size_t)nf * 3 * (fdtype == 0 ? 8 : 4);
  const size_t v_bytes = (size_t)nv * 3 * sizeof(double);
  if (cache.mesh != 0 && !lookup_mesh(cache.mesh)) cache = HostStripCache();  // destroyed elsewhere
  const bool same_f = cache.mesh != 0 && cache.fdtype == fdtype && cache.nv == nv &&
                      cache.nf == nf && (f_bytes == 0 || std::memcmp(cache.F.data(), F, f_bytes) == 0);
  const bool same_v = same_f && (v_bytes == 0 || std::memcmp(cache.V.data(), V, v_bytes) == 0);
  if (!same_f) {
    if (cache.mesh != 0) (void)wn_mesh_destroy(cache.mesh);
    cache = HostStripCache();
    const int64_t mesh = wn_mesh_create(V, F, nv, nf, fdtype);  // validates F against nv
    if (mesh < 0) return mesh;
    cache.mesh = mesh;
    cache.fdtype = (int)fdtype;
    cache.nv = nv;
    cache.nf = nf;
    cache.F.assign((const char*)F, (const char*)F + f_bytes);
    cache.V.assign(V, V + 3 * nv);
  }
  const int64_t mesh = cache.mesh;
  int64_t cfg[2];
  int64_t st = wn_mesh_resolve(mesh, nq, 0, -1, cfg);
  if (st != kOk) return st;
  MeshHandle* m = lookup_mesh(mesh);
  hipError_t e = hipSuccess;
  if (!ctx.ready) {
    e = hipGetDevice(&ctx.device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx.stream, hipStreamNonBlocking);
    if (e == hipSuccess) ctx.ready = true;
  }
  const size_t q_b = wn::align_up((size_t)nq * 3 * sizeof(double));
  const size_t dst_b = wn::align_up((size_t)nq * sizeof(double));
  const size_t ws_b = wn::strip_workspace_bytes(nq, (int)cfg[0], (int)cfg[1],
                                                precision == wn::kFP64 ? 0 : m->nbe);
  const size_t v_b = same_v ? 0 : wn::align_up(v_bytes);
  const size_t need = q_b + dst_b + ws_b + v_b;
  if (e == hipSuccess && need > ctx.cap) {
    if (ctx.buf) {
      (void)hipStreamSynchronize(ctx.stream);
      (void)hipFree(ctx.buf);
      ctx.buf = nullptr;
      ctx.cap = 0;
    }
    e = hipMalloc(&ctx.buf, need);
    if (e == hipSuccess)
      ctx.cap = need;
    else
      ctx.buf = nullptr;
  }
  if (e == hipSuccess) {
    char* base = (char*)ctx.buf;
    double* d_q = (double*)base;
    double* d_dst = (double*)(base + q_b);
    void* d_ws = base + q_b + dst_b;
    if (same_f && !same_v && m->nrec > 0) {
      // new vertices for the cached topology: upload V, regather the records
      double* d_v = (double*)(base + q_b + dst_b + ws_b);
      e = hipMemcpyAsync(d_v, V, v_bytes, hipMemcpyHostToDevice, ctx.stream);
      if (e == hipSuccess && wn_mesh_set_vertices(mesh, d_v, nv, 0, ctx.stream) != kOk)
        e = hipErrorUnknown;
      if (e == hipSuccess) cache.V.assign(V, V + 3 * nv);
    }
    if (e == hipSuccess)
      e = hipMemcpyAsync(d_q, q, (size_t)nq * 3 * sizeof(double), hipMemcpyHostToDevice,
                         ctx.stream);
    if (e == hipSuccess) {
      if (precision == wn::kFP64)
        e = wn::strip64_launch(m->d_rec64, (const double4*)m->d_rec, m->nrec, d_q, d_dst, nq,
                               (int)cfg[0], (int)cfg[1], d_ws, ws_b, ctx.stream);
      else {
        const wn::RibbonSrc b = m->boundary();
        e = wn::strip_launch((const double4*)m->d_rec, m->nrec, d_q, d_dst, nq, (int)cfg[0],
                             (int)cfg[1], 1.0f, d_ws, ws_b, ctx.stream, &b);
      }
    }
    if (e == hipSuccess)
      e = hipMemcpyAsync(dst, d_dst, (size_t)nq * sizeof(double), hipMemcpyDeviceToHost,
                         ctx.stream);
    const hipError_t es = hipStreamSynchronize(ctx.stream);
    if (e == hipSuccess) e = es;
  }
  if (e != hipSuccess) {
    // drop the cached mesh: its records may be half updated
    const std::string why = hip_msg("wn_winding_gpu_mesh", e);
    (void)wn_mesh_destroy(cache.mesh);
    cache = HostStripCache();
    return fail(kErrHip, why);
  }
  return kOk;
}

int64_t wn_release_host_buffers(void) {
  HostPathContext* all = host_contexts();
  for (int k = 0; k < kHostContexts; ++k) {
    HostPathContext& ctx = all[k];
    std::lock_guard<std::mutex> lock(ctx.mu);
    if (ctx.strips.mesh != 0) (void)wn_mesh_destroy(ctx.strips.mesh);
    ctx.strips = HostStripCache();
    if (ctx.buf) {
      (void)hipStreamSynchronize(ctx.stream);
      (void)hipFree(ctx.buf);
      ctx.buf = nullptr;
      ctx.cap = 0;
    }
  }
  return kOk;
}

int64_t wn_scatter_corners(const double* grad_tri, const void* F, int64_t fdtype, int64_t nf,
                           int64_t nv, double* grad_V, void* stream) {
  if (nf < 0 || nv < 0) return fail(kErrArg, "negative size");
  if (fdtype != 0 && fdtype != 1) return fail(kErrArg, "fdtype must be 0 (int64) or 1 (int32)");
  if ((nf > 0 && (!grad_tri || !F)) || (nv > 0 && !grad_V)) return fail(kErrArg, "null device buffer");
  hipError_t e = wn::scatter_corners_launch(grad_tri, F, (int)fdtype, nf, nv, grad_V,
                                            (hipStream_t)stream);
  if (e != hipSuccess) return fail(kErrHip, hip_msg("wn_scatter_corners", e));
  return kOk;
}

int64_t wn_aabb_build(const double* V, const void* F, int64_t nv, int64_t nf, int64_t fdtype,
                      int64_t leaf_size, double* boxes, int64_t* nodes, int64_t* order,
                      int64_t capacity, int64_t* stack_size) {
  const int64_t leaf = leaf_size <= 0 ? 8 : leaf_size;
  if (const char* why = aabb_args(V, F, nv, nf, fdtype, leaf)) return fail(kErrArg, why);
  const int64_t count = wn::aabb_node_count(nf, leaf);
  if (!boxes || !nodes || !order || capacity < count) return count;
  wn::AabbTree t;
  wn::aabb_build(V, F, (int)fdtype, nf, leaf, t);
  std::memcpy(boxes, t.boxes.data(), t.boxes.size() * sizeof(double));
  std::memcpy(nodes, t.nodes.data(), t.nodes.size() * sizeof(int64_t));
  std::memcpy(order, t.order.data(), t.order.size() * sizeof(int64_t));
  if (stack_size) *stack_size = t.stack_size;
  return count;
}

int64_t wn_aabb_create(const double* V, const void* F, int64_t nv, int64_t nf, int64_t fdtype,
                       int64_t leaf_size) {
  const int64_t leaf = leaf_size <= 0 ? 8 : leaf_size;
  if (const char* why = aabb_args(V, F, nv, nf, fdtype, leaf)) return fail(kErrArg, why);
  if (device_count() < 1) return fail(kErrNoDevice, "no HIP device available");
  wn::AabbTree t;
  wn::aabb_build(V, F, (int)fdtype, nf, leaf, t);
  if (t.stack_size > wn::kAabbStack) return fail(kErrArg, "tree deeper than the descent stack");
  AabbHandle* h = new AabbHandle();
  h->nf = nf;
  h->nnodes = (int64_t)t.nodes.size() / 4;
  h->stack_size = t.stack_size;
  h->bytes = wn::aabb_device_bytes(h->nnodes, nf);
  if (h->nnodes > 0) {
    hipStream_t s = nullptr;
    hipError_t e = hipMalloc(&h->dev, h->bytes);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (e == hipSuccess) e = wn::aabb_upload(t, V, F, (int)fdtype, h->dev, s);
    if (s) (void)hipStreamDestroy(s);
    if (e != hipSuccess) {
      if (h->dev) (void)hipFree(h->dev);
      delete h;
      return fail(kErrHip, hip_msg("wn_aabb_create", e));
    }
  }
  std::lock_guard<std::mutex> lock(aabb_registry_mu());
  aabb_registry().insert(h);
  return (int64_t)(intptr_t)h;
}

int64_t wn_aabb_info(int64_t handle, int64_t* out4) {
  AabbHandle* h = lookup_aabb(handle);
  if (!h) return fail(kErrArg, "unknown tree handle");
  if (!out4) return fail(kErrArg, "null output");
  out4[0] = h->nf;
  out4[1] = h->nnodes;
  out4[2] = h->stack_size;
  out4[3] = (int64_t)h->bytes;
  return kOk;
}

int64_t wn_aabb_workspace_bytes(int64_t handle, int64_t np, int64_t sort) {
  if (!lookup_aabb(handle)) return fail(kErrArg, "unknown tree handle");
  if (np < 0 || np > (int64_t)0x7fffffff) return fail(kErrArg, "query count out of range");
  const int srt = sort > 0 ? 1 : 0;  // default: launch order
  if (srt && device_count() < 1) return fail(kErrNoDevice, "no HIP device available");
  return (int64_t)wn::aabb_workspace_bytes(np, srt);
}

int64_t wn_aabb_query(int64_t handle, const double* p, double* sqrd, int64_t* indices,
                      double* closest, int64_t np, int64_t sort, void* workspace,
                      int64_t workspace_bytes, void* stream) {
  AabbHandle* h = lookup_aabb(handle);
  if (!h) return fail(kErrArg, "unknown tree handle");
  if (np < 0 || np > (int64_t)0x7fffffff) return fail(kErrArg, "query count out of range");
  if (np == 0) return kOk;
  if (!p || !sqrd || !indices || !closest) return fail(kErrArg, "null device buffer");
  const int srt = sort > 0 ? 1 : 0;  // default: launch order
  const size_t need = wn::aabb_workspace_bytes(np, srt);
  if (workspace_bytes < 0 || (size_t)workspace_bytes < need || (need > 0 && !workspace))
    return fail(kErrWorkspace, "workspace smaller than wn_aabb_workspace_bytes()");
  hipError_t e = wn::aabb_query_launch(h->dev, h->nnodes, h->nf, p, sqrd, indices, closest, np,
                                       srt, workspace, (size_t)workspace_bytes,
                                       (hipStream_t)stream);
  if (e != hipSuccess) return fail(kErrHip, hip_msg("wn_aabb_query", e));
  return kOk;
}

int64_t wn_aabb_destroy(int64_t handle) {
  AabbHandle* h = nullptr;
  {
    std::lock_guard<std::mutex> lock(aabb_registry_mu());
    auto it = aabb_registry().find((AabbHandle*)(intptr_t)handle);
    if (it == aabb_registry().end()) return fail(kErrArg, "unknown tree handle");
    h = *it;
    aabb_registry().erase(it);
  }
  if (h->dev) {
    (void)hipDeviceSynchronize();
    (void)hipFree(h->dev);
  }
  delete h;
  return kOk;
}

int64_t wn_fwn_create(const double* V, const void* F, int64_t nv, int64_t nf, int64_t fdtype,
                      int64_t leaf_size, int64_t upload) {
  const int64_t leaf = leaf_size <= 0 ? 8 : leaf_size;
  if (const char* why = aabb_args(V, F, nv, nf, fdtype, leaf)) return fail(kErrArg, why);
  if (upload && device_count() < 1) return fail(kErrNoDevice, "no HIP device available");
  FwnHandle* h = new FwnHandle();
  h->nf = nf;
  wn::fwn_build(V, F, (int)fdtype, nf, leaf, h->tree);
  if (h->tree.stack_size > wn::kAabbStack) {
    delete h;
    return fail(kErrArg, "tree deeper than the descent stack");
  }
  if (upload && !h->tree.nodes.empty()) {
    h->bytes = wn::fwn_device_bytes((int64_t)h->tree.nodes.size(), nf);
    hipStream_t s = nullptr;
    hipError_t e = hipMalloc(&h->dev, h->bytes);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (e == hipSuccess) e = wn::fwn_upload(h->tree, h->dev, s);
    if (s) (void)hipStreamDestroy(s);
    if (e != hipSuccess) {
      if (h->dev) (void)hipFree(h->dev);
      delete h;
      return fail(kErrHip, hip_msg("wn_fwn_create", e));
    }
  }
  std::lock_guard<std::mutex> lock(fwn_registry_mu());
  fwn_registry().insert(h);
  return (int64_t)(intptr_t)h;
}

int64_t wn_fwn_info(int64_t handle, int64_t* out4) {
  FwnHandle* h = lookup_fwn(handle);
  if (!h) return fail(kErrArg, "unknown tree handle");
  if (!out4) return fail(kErrArg, "null output");
  out4[0] = h->nf;
  out4[1] = (int64_t)h->tree.nodes.size();
  out4[2] = h->tree.stack_size;
  out4[3] = (int64_t)h->bytes;
  return kOk;
}

int64_t wn_fwn_export(int64_t handle, double* expansions, int64_t* links, int64_t capacity) {
  FwnHandle* h = lookup_fwn(handle);
  if (!h) return fail(kErrArg, "unknown tree handle");
  const int64_t n = (int64_t)h->tree.nodes.size();
  if (!expansions || !links || capacity < n) return n;
  for (int64_t i = 0; i < n; ++i) {
    const wn::FwnNode& e = h->tree.nodes[(size_t)i];
    std::memcpy(expansions + 34 * i, &e.cx, 34 * sizeof(double));
    links[4 * i] = e.left;
    links[4 * i + 1] = e.right;
    links[4 * i + 2] = e.start;
    links[4 * i + 3] = e.count;
  }
  return n;
}

int64_t wn_fwn_query_cpu(int64_t handle, const double* q, double* w, int64_t nq, double beta,
                         int64_t q0, int64_t q1) {
  FwnHandle* h = lookup_fwn(handle);
  if (!h) return fail(kErrArg, "unknown tree handle");
  if (!(beta > 0.0)) return fail(kErrArg, "beta must be positive");
  if (nq < 0 || q0 < 0 || q1 < q0 || q1 > nq) return fail(kErrArg, "bad query span");
  if (q1 > q0 && (!q || !w)) return fail(kErrArg, "null host buffer");
  wn::fwn_query_cpu(h->tree, q, w, beta, q0, q1);
  return kOk;
}

int64_t wn_fwn_workspace_bytes(int64_t handle, int64_t nq, int64_t sort) {
  if (!lookup_fwn(handle)) return fail(kErrArg, "unknown tree handle");
  if (nq < 0 || nq > (int64_t)0x7fffffff) return fail(kErrArg, "query count out of range");
  const int srt = sort < 0 ? (nq >= 4096 ? 1 : 0) : (sort ? 1 : 0);
  if (srt && device_count() < 1) return fail(kErrNoDevice, "no HIP device available");
  return (int64_t)wn::fwn_workspace_bytes(nq, srt);
}

int64_t wn_fwn_query(int64_t handle, const double* q, double* w, int64_t nq, double beta,
                     int64_t sort, void* workspace, int64_t workspace_bytes, void* stream) {
  FwnHandle* h = lookup_fwn(handle);
  if (!h) return fail(kErrArg, "unknown tree handle");
  if (!(beta > 0.0)) return fail(kErrArg, "beta must be positive");
  if (nq < 0 || nq > (int64_t)0x7fffffff) return fail(kErrArg, "query count out of range");
  if (nq == 0) return kOk;
  if (!q || !w) return fail(kErrArg, "null device buffer");
  if (!h->tree.nodes.empty() && !h->dev) return fail(kErrArg, "tree was created without a device copy");
  const int srt = sort < 0 ? (nq >= 4096 ? 1 : 0) : (sort ? 1 : 0);
  const size_t need = wn::fwn_workspace_bytes(nq, srt);
  if (workspace_bytes < 0 || (size_t)workspace_bytes < need || (need > 0 && !workspace))
    return fail(kErrWorkspace, "workspace smaller than wn_fwn_workspace_bytes()");
  hipError_t e = wn::fwn_query_launch(h->dev, (int64_t)h->tree.nodes.size(), q, w, nq, beta, srt,
                                      workspace, (size_t)workspace_bytes, (hipStream_t)stream);
  if (e != hipSuccess) return fail(kErrHip, hip_msg("wn_fwn_query", e));
  return kOk;
}

int64_t wn_fwn_destroy(int64_t handle) {
  FwnHandle* h = nullptr;
  {
    std::lock_guard<std::mutex> lock(fwn_registry_mu());
    auto it = fwn_registry().find((FwnHandle*)(intptr_t)handle);
    if (it == fwn_registry().end()) return fail(kErrArg, "unknown tree handle");
    h = *it;
    fwn_registry().erase(it);
  }
  if (h->dev) {
    (void)hipDeviceSynchronize();
    (void)hipFree(h->dev);
  }
  delete h;
  return kOk;
}

int64_t wn_strip_topology(const void* F, int64_t nv, int64_t nf, int64_t fdtype, int32_t* idx,
                          int8_t* flag, int64_t capacity) {
  if (nv < 0 || nf < 0) return fail(kErrArg, "negative size");
  if (nv > (int64_t)0x7fffffff) return fail(kErrArg, "more than 2^31-1 vertices");
  if (fdtype != 0 && fdtype != 1) return fail(kErrArg, "fdtype must be 0 (int64) or 1 (int32)");
  if (nf > 0 && !F) return fail(kErrArg, "null host buffer");
  for (int64_t i = 0; i < 3 * nf; ++i) {
    const int64_t v = fdtype == 0 ? ((const int64_t*)F)[i] : (int64_t)((const int32_t*)F)[i];
    if (v < 0 || v >= nv) return fail(kErrArg, "face index is outside V");
  }
  std::vector<int32_t> ids;
  std::vector<int8_t> flags;
  const int64_t n = wn::strip_topology(F, nf, (int)fdtype, ids, flags);
  if (idx && flag && capacity >= n) {
    std::memcpy(idx, ids.data(), ids.size() * sizeof(int32_t));
    std::memcpy(flag, flags.data(), flags.size());
  }
  return n;
}

int64_t wn_strip_boundary(const void* F, int64_t nv, int64_t nf, int64_t fdtype,
                          const int32_t* idx, int64_t nrec, int32_t* pairs, double* weight,
                          int64_t capacity) {
  if (nv < 0 || nf < 0 || nrec < 0) return fail(kErrArg, "negative size");
  if (nv > (int64_t)0x7fffffff) return fail(kErrArg, "more than 2^31-1 vertices");
  if (fdtype != 0 && fdtype != 1) return fail(kErrArg, "fdtype must be 0 (int64) or 1 (int32)");
  if ((nf > 0 && !F) || (nrec > 0 && !idx)) return fail(kErrArg, "null host buffer");
  for (int64_t i = 0; i < 3 * nf; ++i) {
    const int64_t v = fdtype == 0 ? ((const int64_t*)F)[i] : (int64_t)((const int32_t*)F)[i];
    if (v < 0 || v >= nv) return fail(kErrArg, "face index is outside V");
  }
  std::vector<int32_t> pr;
  std::vector<double> w;
  const int64_t n = wn::strip_boundary(F, nf, (int)fdtype, nv, idx, nrec, pr, w);
  for (int32_t r : pr)
    if (r < 0) return fail(kErrArg, "a boundary vertex has no record (idx is not F's topology)");
  if (pairs && weight && capacity >= n) {
    std::memcpy(pairs, pr.data(), pr.size() * sizeof(int32_t));
    std::memcpy(weight, w.data(), w.size() * sizeof(double));
  }
  return n;
}

int64_t wn_strip_gather(const void* V, int64_t nv, int64_t vdtype, const int32_t* idx,
                        const int8_t* flag, int64_t nrec, double* rec, void* stream) {
  if (nv < 0 || nrec < 0) return fail(kErrArg, "negative size");
  if (vdtype != 0 && vdtype != 1) return fail(kErrArg, "vdtype must be 0 (float64) or 1 (float32)");
  if (nrec == 0) return kOk;
  if (!V || !idx || !flag || !rec) return fail(kErrArg, "null device buffer");
  hipError_t e = wn::strip_gather_launch(V, (int)vdtype, nv, idx, flag, nrec, (double4*)rec,
                                         (hipStream_t)stream);
  if (e != hipSuccess) return fail(kErrHip, hip_msg("wn_strip_gather", e));
  return kOk;
}

}  // extern "C"
