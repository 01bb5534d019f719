// pcme_api.hip -- host side of libpcme: the extern "C" entry points of
// include/pcme.h, argument checks, device workspaces and kernel launches.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <memory>
#include <mutex>
#include <vector>

#include "pcme_internal.h"

namespace pcme {

static thread_local char g_err[512] = "";

static int fail(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return code;
}

void set_error(const char *msg) { snprintf(g_err, sizeof g_err, "%s", msg); }

#define HIPCHK(call)                                                                 \
  do {                                                                               \
    hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess)                                                            \
      return fail(PCME_EHIP, "%s failed: %s", #call, hipGetErrorString(e_));         \
  } while (0)

struct Plan;

// One per-call workspace: a device buffer for a call's job table (or node
// windows) and the pinned host staging of its upload.  A call takes a
// workspace, uploads into it, launches on the caller's stream and records
// `ev` there after its last kernel.  The next call on the same stream may
// take it again at once (stream order puts its upload after those
// kernels); a call on another stream only once `ev` has completed.  So
// calls on different streams never share a buffer a running kernel reads.
// The staging is rewritten only after the previous upload from it (`cev`)
// has completed.
struct CallWs {
  void *dev = nullptr;
  size_t cap = 0;
  unsigned char *host = nullptr;
  size_t hcap = 0;
  hipEvent_t ev = nullptr;   // after the last kernel of the last call
  hipEvent_t cev = nullptr;  // after the last upload from `host`
  hipStream_t last = nullptr;
  bool recorded = false;
  bool copied = false;
};

// Per-device state: the per-call workspaces, the plan cached by
// pcme_search_frames and the event that orders its launches.
struct Dev {
  int id = -1;
  std::mutex mu;
  int resident_waves[4] = {0, 0, 0, 0};  // full-search / hinted kernel; the same with several references
  int resident_var[8] = {0};       // resident waves of each full-search configuration (kVariants)
  int resident_16[3] = {0, 0, 0};  // resident waves at 16 per CU: hinted, multi-reference, both
  int cus = 1;                     // compute units
  int b_walk_resident = 1;         // resident workgroups of b_walk_kernel
  std::vector<CallWs *> ws;        // per-call workspaces (at most kMaxWs)
  Plan *cached = nullptr;          // plan of the last pcme_search_frames call
  std::vector<pcme_frame_job> cached_jobs;
  uint8_t *scratch = nullptr;      // this device's copy of `scratch` (hipGetSymbolAddress)
  bool scratch_init = false;       // its control block has been cleared once
  SkipTables *skip_tab = nullptr;  // this device's copy of `skip_tables`, filled on first use
  uint32_t *b_stall = nullptr;     // this device's copy of `b_field_stall` (first pcme_b_field call)
};

// Tables of the early P-skip gate; filled on a device by the first
// pcme_skip_gate_batch call there.
static __device__ __align__(256) SkipTables skip_tables;

// Sticky stall word of the B field walkers: set by a launch whose
// dependency wait passed its bound, read and cleared by pcme_b_field_check.
// Host code reaches it through Dev::b_stall.
static __device__ uint32_t b_field_stall;

// Workspace of pcme_search_frames_noalloc: a control block, the job table,
// the tickets and the exchange words of one call, in a file-scope device
// array, so that call allocates nothing.  Each device holds its own copy
// until the process ends; host code reaches it through Dev::scratch (the
// address hipGetSymbolAddress returns on that device), never through the
// name, which in host code is a host-side stand-in.
#define PCME_NOALLOC_BYTES (16u << 20)
#define PCME_NOALLOC_MAX_JOBS 1024
static __device__ __align__(256) uint8_t scratch[PCME_NOALLOC_BYTES];

// Configurations of the full-search kernel the autotuner chooses from
// (search_frames_var_kernel): row walkers per workgroup, the occupancy cap
// its LDS reservation sets, and whether staged LDS regions are kept along a
// row.  Entry 0 is search_frames_kernel, the default.
struct Variant {
  int wpg, wcu;
  bool reuse;
};
static const Variant kVariants[] = {{1, 12, true}, {2, 12, true}, {4, 12, true},
                                    {1, 8, true},  {2, 8, true},  {1, 12, false},
                                    {4, 16, false}, {4, 16, true}};
static const int kNumVariants = (int)(sizeof kVariants / sizeof kVariants[0]);
// A plan with pruner hints, several references or both runs its kernel's
// default configuration or (4, 16, 1), kVariants[k16] (kernel16()).
static const int k16 = 7;
static_assert(k16 < kNumVariants, "k16 names a built configuration");
static_assert(sizeof kVariants / sizeof kVariants[0] <= 8, "Dev::resident_var holds 8 configurations");

static const void *variant_fn(int v) {
  switch (v) {
    case 1: return (const void *)search_frames_var_kernel<2, 12, true>;
    case 2: return (const void *)search_frames_var_kernel<4, 12, true>;
    case 3: return (const void *)search_frames_var_kernel<1, 8, true>;
    case 4: return (const void *)search_frames_var_kernel<2, 8, true>;
    case 5: return (const void *)search_frames_var_kernel<1, 12, false>;
    case 6: return (const void *)search_frames_var_kernel<4, 16, false>;
    case 7: return (const void *)search_frames_var_kernel<4, 16, true>;
    default: return (const void *)search_frames_kernel;
  }
}

// Workspaces per device: with this many in flight on other streams a call
// waits for the oldest one to complete.
static const size_t kMaxWs = 16;

static std::mutex g_mu;
static std::map<int, Dev *> g_devs;

static bool is_gfx950(int d) {
  hipDeviceProp_t p;
  if (hipGetDeviceProperties(&p, d) != hipSuccess) return false;
  return strncmp(p.gcnArchName, "gfx950", 6) == 0;
}

typedef void (*FrameKernel)(const DevJob *, const int2 *, int, uint32_t *, uint32_t);
// The frame-search kernel of a job table: with or without pruner hints, one
// reference or several.
static FrameKernel frame_kernel(bool hinted, bool mr) {
  if (mr) return hinted ? search_frames_mr_hinted_kernel : search_frames_mr_kernel;
  return hinted ? search_frames_hinted_kernel : search_frames_kernel;
}

// The same kinds at 16 waves per CU (4 walkers per workgroup of 256).
static FrameKernel kernel16(bool hinted, bool mr) {
  if (mr) return hinted ? search_frames_mr_hinted16_kernel : search_frames_mr16_kernel;
  return search_frames_hinted16_kernel;
}

// The current device's state; the first call per device validates it.
// Later calls only read the current device id, so entry points stay cheap
// and safe inside a stream capture.
static int get_dev(Dev **out) {
  int d = -1;
  if (hipGetDevice(&d) != hipSuccess || d < 0) return fail(PCME_ENODEV, "no HIP device visible");
  std::lock_guard<std::mutex> g(g_mu);
  auto it = g_devs.find(d);
  if (it != g_devs.end()) {
    *out = it->second;
    return PCME_OK;
  }
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return fail(PCME_ENODEV, "no HIP device visible");
  if (!is_gfx950(d)) return fail(PCME_ENODEV, "device %d is not a gfx950 (MI355X)", d);
  Dev *dv = new Dev();
  dv->id = d;
  hipDeviceProp_t p;
  HIPCHK(hipGetDeviceProperties(&p, d));
  for (int v = 0; v < 4; v++) {
    int per_cu = 0;
    HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, frame_kernel(v & 1, v >= 2), 64, 0));
    if (per_cu < 1) per_cu = 1;
    dv->resident_waves[v] = per_cu * p.multiProcessorCount;
  }
  dv->cus = p.multiProcessorCount;
  for (int k = 0; k < 3; k++) {
    int per_cu = 0;
    HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)kernel16(k != 1, k != 0), 256, 0));
    dv->resident_16[k] = std::max(1, per_cu) * 4 * p.multiProcessorCount;
  }
  for (int v = 0; v < kNumVariants; v++) {
    int per_cu = 0;
    HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, variant_fn(v), 64 * kVariants[v].wpg, 0));
    if (per_cu < 1) per_cu = 1;
    dv->resident_var[v] = per_cu * kVariants[v].wpg * p.multiProcessorCount;
  }
  {
    int per_cu = 0;
    HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)b_walk_kernel, 128, 0));
    dv->b_walk_resident = std::max(1, per_cu) * p.multiProcessorCount;
  }
  HIPCHK(hipGetSymbolAddress((void **)&dv->scratch, HIP_SYMBOL(scratch)));
  if (!dv->scratch) return fail(PCME_EHIP, "hipGetSymbolAddress(scratch) returned null");
  g_devs[d] = dv;
  *out = dv;
  return PCME_OK;
}

// A workspace of at least `bytes` (device) and `hbytes` (pinned host) for a
// call on `stream`; the caller holds dv->mu.  Prefers the workspace last
// used on the same stream, then a free one (its last kernel done), creates
// one while fewer than kMaxWs exist, and else waits for the oldest.
static int take_ws(Dev *dv, hipStream_t stream, size_t bytes, size_t hbytes, CallWs **out) {
  CallWs *pick = nullptr;
  for (CallWs *w : dv->ws)
    if (w->recorded && w->last == stream && w->cap >= bytes && w->hcap >= hbytes) {
      pick = w;
      break;
    }
  if (!pick)
    for (CallWs *w : dv->ws) {
      if (w->recorded && hipEventQuery(w->ev) != hipSuccess) continue;
      if (!pick || (w->cap >= bytes && pick->cap < bytes)) pick = w;
      if (pick->cap >= bytes && pick->hcap >= hbytes) break;
    }
  if (!pick && dv->ws.size() < kMaxWs) {
    pick = new CallWs();
    if (hipEventCreateWithFlags(&pick->ev, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&pick->cev, hipEventDisableTiming) != hipSuccess) {
      if (pick->ev) (void)hipEventDestroy(pick->ev);
      delete pick;
      return fail(PCME_EHIP, "hipEventCreateWithFlags failed");
    }
    dv->ws.push_back(pick);
  }
  if (!pick) {
    pick = dv->ws.front();
    HIPCHK(hipEventSynchronize(pick->ev));
  }
  if (pick->copied) {
    HIPCHK(hipEventSynchronize(pick->cev));
    pick->copied = false;
  }
  if (pick->cap < bytes || pick->hcap < hbytes) {
    // resizing frees buffers: only once the last call's kernels are done
    if (pick->recorded) HIPCHK(hipEventSynchronize(pick->ev));
    pick->recorded = false;
  }
  if (pick->cap < bytes) {
    if (pick->dev) HIPCHK(hipFree(pick->dev));
    pick->dev = nullptr;
    pick->cap = 0;
    if (hipMalloc(&pick->dev, bytes) != hipSuccess) return fail(PCME_ENOMEM, "hipMalloc of %zu bytes failed", bytes);
    pick->cap = bytes;
  }
  if (pick->hcap < hbytes) {
    if (pick->host) HIPCHK(hipHostFree(pick->host));
    pick->host = nullptr;
    pick->hcap = 0;
    HIPCHK(hipHostMalloc((void **)&pick->host, hbytes, hipHostMallocDefault));
    pick->hcap = hbytes;
  }
  *out = pick;
  return PCME_OK;
}

// Copy `bytes` of host memory into the workspace's device buffer on
// `stream`, through its pinned staging.
static int ws_upload(CallWs *w, const void *src, size_t bytes, hipStream_t stream) {
  memcpy(w->host, src, bytes);
  HIPCHK(hipMemcpyAsync(w->dev, w->host, bytes, hipMemcpyHostToDevice, stream));
  HIPCHK(hipEventRecord(w->cev, stream));
  w->copied = true;
  return PCME_OK;
}

// Marks a taken workspace busy until the work enqueued on `stream` by the
// time the guard ends has completed -- on every return path, so a call that
// fails after a launch never frees a buffer a kernel still reads.  Declared
// after the device lock, so it runs while the lock is held.
struct WsGuard {
  CallWs *w = nullptr;
  hipStream_t s = nullptr;
  ~WsGuard() {
    if (w && hipEventRecord(w->ev, s) == hipSuccess) {
      w->recorded = true;
      w->last = s;
    }
  }
};

// slice.rs:52-68
static uint32_t lambda_of(int qp) {
  double l = 0.85 * pow(2.0, ((double)qp - 12.0) / 3.0);
  double v = round(sqrt(l) * 256.0);
  if (v < 16.0) v = 16.0;
  return (uint32_t)v;
}

static int check_params(const pcme_params &p) {
  if (p.mbw < 1 || p.mbh < 1 || p.mbw > 4096 || p.mbh > 4096)
    return fail(PCME_EINVAL, "mbw/mbh out of range (%d x %d)", p.mbw, p.mbh);
  if (p.me_range < 1 || p.me_range > 1024) return fail(PCME_EINVAL, "me_range %d not in [1,1024]", p.me_range);
  if (p.subpel_rounds < 0 || p.subpel_rounds > 8)
    return fail(PCME_EINVAL, "subpel_rounds %d not in [0,8]", p.subpel_rounds);
  if (p.qp < 0 || p.qp > 51) return fail(PCME_EINVAL, "qp %d not in [0,51]", p.qp);
  return PCME_OK;
}

// Bound of every dependency wait (polls of ~1 us); a test hook lowers it.
static uint32_t kSpinLimit = 1u << 22;

// A frame-search plan: the device copy of a job table, its tickets and its
// own exchange words and control words.  Launching a plan issues kernels
// only (no host synchronisation, no copies, no memsets), so it can be
// captured in a hipGraph and replayed.
struct Plan {
  int dev = -1;
  int njobs = 0;
  DevJob *djobs = nullptr;
  int2 *dtickets = nullptr;
  int ntickets = 0;
  uint32_t *xchg = nullptr;
  size_t nwords = 0;
  uint32_t *ctl = nullptr;  // [1] stall flag of the running launch, [2] epoch, [3] clear request,
                            // [4] sticky stall flag (set by any launch, cleared by plan_check),
                            // [8..15] tickets taken per queue, [16..24] queue q's tickets =
                            // [ctl[16+q], ctl[17+q]) (CTL_BYTES)
  int resident = 0;         // resident waves of the search kernel on this device
  bool hinted = false;      // some job carries pruner hints: the hinted kernel
  bool mr = false;          // some job has several references: the multi-reference kernels
  int grid = 0;             // waves launched (persistent grid)
  int variant = 0;          // configuration of the kernel (kVariants; with hints or references: 0 or k16)
  const int *resident_var = nullptr;  // the device's resident waves per configuration
  int resident16 = 0;       // the device's resident waves of the plan's kernel at 16 per CU
  int cus = 1;              // compute units of the device
  hipEvent_t ev = nullptr;  // recorded after each eager launch: the next launch, on any stream, waits
  bool ev_recorded = false;
};

static int check_jobs(const pcme_frame_job *jobs, int njobs) {
  if (njobs < 1 || !jobs) return fail(PCME_EINVAL, "bad job array");
  int rc;
  for (int i = 0; i < njobs; i++) {
    if ((rc = check_params(jobs[i].prm))) return rc;
    if (!jobs[i].src || !jobs[i].ref || !jobs[i].ref_h || !jobs[i].ref_v || !jobs[i].ref_c || !jobs[i].out)
      return fail(PCME_EINVAL, "job %d has a null pointer", i);
    if ((((uintptr_t)jobs[i].src | (uintptr_t)jobs[i].ref | (uintptr_t)jobs[i].ref_h | (uintptr_t)jobs[i].ref_v |
          (uintptr_t)jobs[i].ref_c) & 3u) != 0)
      return fail(PCME_EINVAL, "job %d: plane pointers must be 4-byte aligned", i);
    if (((uintptr_t)jobs[i].hint & 3u) != 0) return fail(PCME_EINVAL, "job %d: hints must be 4-byte aligned", i);
    if (jobs[i].nref < 0 || jobs[i].nref > PCME_MAX_REFS)
      return fail(PCME_EINVAL, "job %d: nref %d not in [0, %d]", i, jobs[i].nref, PCME_MAX_REFS);
    for (int k = 0; k + 1 < jobs[i].nref; k++) {
      const pcme_ref_planes &o = jobs[i].older[k];
      if (!o.luma || !o.h || !o.v || !o.c) return fail(PCME_EINVAL, "job %d: reference %d has a null plane", i, k + 1);
      if ((((uintptr_t)o.luma | (uintptr_t)o.h | (uintptr_t)o.v | (uintptr_t)o.c) & 3u) != 0)
        return fail(PCME_EINVAL, "job %d: reference %d: plane pointers must be 4-byte aligned", i, k + 1);
    }
  }
  return PCME_OK;
}

// The device copy of a frame job (its exchange words are set by the caller).
static DevJob dev_job(const pcme_frame_job &J) {
  DevJob d;
  memset(&d, 0, sizeof d);
  d.src = J.src;
  d.pl[0] = J.ref; d.pl[1] = J.ref_h; d.pl[2] = J.ref_v; d.pl[3] = J.ref_c;
  d.out = J.out;
  d.hint = J.hint;
  d.xchg = nullptr;
  d.mbw = J.prm.mbw;
  d.mbh = J.prm.mbh;
  d.stride = 16 * J.prm.mbw + 2 * PCME_PAD;
  d.me_range = J.prm.me_range;
  d.subpel_rounds = J.prm.subpel_rounds;
  d.partitions = J.prm.partitions ? 1 : 0;
  d.lam = lambda_of(J.prm.qp);
  d.nref = std::max(1, (int)J.nref);
  for (int k = 0; k + 1 < d.nref; k++) {
    d.opl[k][0] = J.older[k].luma; d.opl[k][1] = J.older[k].h;
    d.opl[k][2] = J.older[k].v; d.opl[k][3] = J.older[k].c;
  }
  return d;
}

static void plan_free(Plan *p) {
  if (!p) return;
  if (p->djobs) (void)hipFree(p->djobs);
  if (p->dtickets) (void)hipFree(p->dtickets);
  if (p->xchg) (void)hipFree(p->xchg);
  if (p->ctl) (void)hipFree(p->ctl);
  if (p->ev) (void)hipEventDestroy(p->ev);
  delete p;
}

// Resident waves of the plan's kernel configuration on its device.
// Whether the plan runs the 16-per-CU kernel of its kind (hinted,
// multi-reference or both).
static bool plan_16(const Plan *p) { return (p->hinted || p->mr) && p->variant == k16; }

static int plan_resident(const Plan *p) {
  if (plan_16(p)) return p->resident16;
  return (p->hinted || p->mr) ? p->resident : p->resident_var[p->variant];
}

// The configurations a plan can run: every built one for the full search;
// the default or (4, 16, 1) with pruner hints or several references.
static bool plan_allows(const Plan *p, int v) {
  if (p->hinted || p->mr) return v == 0 || v == k16;
  return v >= 0 && v < kNumVariants;
}

static int plan_create(Dev *dv, const pcme_frame_job *jobs, int njobs, Plan **out) {
  Plan *p = new Plan();
  p->dev = dv->id;
  p->njobs = njobs;
  for (int i = 0; i < njobs; i++) {
    p->hinted |= jobs[i].hint != nullptr;
    p->mr |= jobs[i].nref > 1;
  }
  p->resident = dv->resident_waves[(p->hinted ? 1 : 0) + (p->mr ? 2 : 0)];
  p->resident_var = dv->resident_var;
  p->resident16 = dv->resident_16[p->hinted && p->mr ? 2 : (p->mr ? 1 : 0)];
  p->cus = dv->cus;
  std::vector<DevJob> dj((size_t)njobs);
  size_t words = 0;
  int maxh = 0;
  for (int i = 0; i < njobs; i++) {
    const pcme_frame_job &J = jobs[i];
    dj[(size_t)i] = dev_job(J);
    words += (size_t)J.prm.mbw * J.prm.mbh * 4;
    maxh = std::max(maxh, J.prm.mbh);
  }
  // Eight ticket queues, one per XCD (a wave takes its own XCD's queue
  // first): queue q holds a contiguous run of jobs of about an eighth of
  // the MBs, so neighbouring rows of a frame, and a frame and the next one
  // (whose reference it is), share an L2.  Within a queue tickets are in
  // (row, job) order: a row's dependency (row-1 of the same job, in the same
  // queue) always has a smaller ticket.
  std::vector<int> qof((size_t)njobs);
  {
    size_t total = 0, cum = 0;
    for (int i = 0; i < njobs; i++) total += (size_t)dj[(size_t)i].mbw * dj[(size_t)i].mbh;
    for (int i = 0; i < njobs; i++) {
      const size_t m = (size_t)dj[(size_t)i].mbw * dj[(size_t)i].mbh;
      qof[(size_t)i] = (int)std::min<size_t>(PCME_QUEUES - 1, (cum + m / 2) * PCME_QUEUES / std::max<size_t>(total, 1));
      cum += m;
    }
  }
  std::vector<int2> tk;
  tk.reserve(words / 4);
  uint32_t qoff[PCME_QUEUES + 1];
  for (int q = 0; q < PCME_QUEUES; q++) {
    qoff[q] = (uint32_t)tk.size();
    for (int y = 0; y < maxh; y++)
      for (int i = 0; i < njobs; i++)
        if (qof[(size_t)i] == q && y < dj[(size_t)i].mbh) tk.push_back(make_int2(i, y));
  }
  qoff[PCME_QUEUES] = (uint32_t)tk.size();
  p->ntickets = (int)tk.size();
  p->nwords = words;
  p->grid = std::min(p->ntickets, p->resident);
  if (hipMalloc((void **)&p->xchg, words * 4) != hipSuccess || hipMalloc((void **)&p->ctl, CTL_BYTES) != hipSuccess ||
      hipMalloc((void **)&p->djobs, dj.size() * sizeof(DevJob)) != hipSuccess ||
      hipMalloc((void **)&p->dtickets, tk.size() * sizeof(int2)) != hipSuccess ||
      hipEventCreateWithFlags(&p->ev, hipEventDisableTiming) != hipSuccess) {
    plan_free(p);
    return fail(PCME_ENOMEM, "plan allocation of %zu exchange words failed", words);
  }
  size_t off = 0;
  for (int i = 0; i < njobs; i++) {
    dj[(size_t)i].xchg = p->xchg + off;
    off += (size_t)dj[(size_t)i].mbw * dj[(size_t)i].mbh * 4;
  }
  // ctl[2] = 0: the first launch takes tag 1 and clears the exchange words.
  if (hipMemset(p->ctl, 0, CTL_BYTES) != hipSuccess ||
      hipMemcpy(p->ctl + 16, qoff, sizeof qoff, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(p->djobs, dj.data(), dj.size() * sizeof(DevJob), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(p->dtickets, tk.data(), tk.size() * sizeof(int2), hipMemcpyHostToDevice) != hipSuccess) {
    plan_free(p);
    return fail(PCME_EHIP, "plan upload failed");
  }
  *out = p;
  return PCME_OK;
}

// Launches of one plan share its workspace, so they run one after another:
// an eager launch waits (on the GPU) for the plan's previous eager launch,
// whatever stream either was issued on.  Under stream capture the graph's
// own order applies and no event is used.
static int plan_launch(Plan *p, hipStream_t stream) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  HIPCHK(hipStreamIsCapturing(stream, &cs));
  const bool eager = cs == hipStreamCaptureStatusNone;
  if (eager && p->ev_recorded) HIPCHK(hipStreamWaitEvent(stream, p->ev, 0));
  hipLaunchKernelGGL(prep_kernel, dim3(1), dim3(64), 0, stream, p->ctl);
  HIPCHK(hipGetLastError());
  hipLaunchKernelGGL(clear_kernel, dim3(512), dim3(256), 0, stream, p->ctl, p->xchg, p->nwords);
  HIPCHK(hipGetLastError());
  if (plan_16(p)) {
    hipLaunchKernelGGL(kernel16(p->hinted, p->mr), dim3((unsigned)((p->grid + 3) / 4)), dim3(256), 0, stream,
                       p->djobs, p->dtickets, p->ntickets, p->ctl, kSpinLimit);
  } else if (p->hinted || p->mr) {
    hipLaunchKernelGGL(frame_kernel(p->hinted, p->mr), dim3((unsigned)p->grid), dim3(64), 0, stream, p->djobs,
                       p->dtickets, p->ntickets, p->ctl, kSpinLimit);
  } else {
    const int wpg = kVariants[p->variant].wpg;
    const dim3 g((unsigned)((p->grid + wpg - 1) / wpg)), b((unsigned)(64 * wpg));
    switch (p->variant) {
      case 1: hipLaunchKernelGGL((search_frames_var_kernel<2, 12, true>), g, b, 0, stream, p->djobs, p->dtickets,
                                 p->ntickets, p->ctl, kSpinLimit); break;
      case 2: hipLaunchKernelGGL((search_frames_var_kernel<4, 12, true>), g, b, 0, stream, p->djobs, p->dtickets,
                                 p->ntickets, p->ctl, kSpinLimit); break;
      case 3: hipLaunchKernelGGL((search_frames_var_kernel<1, 8, true>), g, b, 0, stream, p->djobs, p->dtickets,
                                 p->ntickets, p->ctl, kSpinLimit); break;
      case 4: hipLaunchKernelGGL((search_frames_var_kernel<2, 8, true>), g, b, 0, stream, p->djobs, p->dtickets,
                                 p->ntickets, p->ctl, kSpinLimit); break;
      case 5: hipLaunchKernelGGL((search_frames_var_kernel<1, 12, false>), g, b, 0, stream, p->djobs, p->dtickets,
                                 p->ntickets, p->ctl, kSpinLimit); break;
      case 6: hipLaunchKernelGGL((search_frames_var_kernel<4, 16, false>), g, b, 0, stream, p->djobs, p->dtickets,
                                 p->ntickets, p->ctl, kSpinLimit); break;
      case 7: hipLaunchKernelGGL((search_frames_var_kernel<4, 16, true>), g, b, 0, stream, p->djobs, p->dtickets,
                                 p->ntickets, p->ctl, kSpinLimit); break;
      default: hipLaunchKernelGGL(search_frames_kernel, g, b, 0, stream, p->djobs, p->dtickets, p->ntickets, p->ctl,
                                  kSpinLimit); break;
    }
  }
  HIPCHK(hipGetLastError());
  if (eager) {
    HIPCHK(hipEventRecord(p->ev, stream));
    p->ev_recorded = true;
  }
  return PCME_OK;
}

// Waits for `stream` and reports whether any launch of the plan since the
// previous check stalled (the sticky flag ctl[4]), clearing the flag.
static int plan_check(Plan *p, hipStream_t stream) {
  HIPCHK(hipStreamSynchronize(stream));
  if (p->ev_recorded) HIPCHK(hipEventSynchronize(p->ev));
  uint32_t flag = 0;
  HIPCHK(hipMemcpy(&flag, p->ctl + 4, 4, hipMemcpyDeviceToHost));
  if (flag) {
    HIPCHK(hipMemset(p->ctl + 4, 0, 4));
    return fail(PCME_ESTALL, "a wavefront dependency wait exceeded its bound");
  }
  return PCME_OK;
}

static bool same_jobs(const std::vector<pcme_frame_job> &a, const pcme_frame_job *b, int n) {
  return a.size() == (size_t)n && memcmp(a.data(), b, (size_t)n * sizeof(pcme_frame_job)) == 0;
}

static int search_frames_impl(const pcme_frame_job *jobs, int njobs, hipStream_t stream, bool sync) {
  if (njobs == 0) return PCME_OK;
  int rc;
  if ((rc = check_jobs(jobs, njobs))) return rc;
  Dev *dv = nullptr;
  if ((rc = get_dev(&dv))) return rc;
  std::lock_guard<std::mutex> g(dv->mu);
  if (!dv->cached || !same_jobs(dv->cached_jobs, jobs, njobs)) {
    if (dv->cached) {
      HIPCHK(hipDeviceSynchronize());
      plan_free(dv->cached);
      dv->cached = nullptr;
    }
    if ((rc = plan_create(dv, jobs, njobs, &dv->cached))) return rc;
    dv->cached_jobs.assign(jobs, jobs + njobs);
  }
  if ((rc = plan_launch(dv->cached, stream))) return rc;
  return sync ? plan_check(dv->cached, stream) : PCME_OK;
}

// Host-side tables of pcme_search_frames_noalloc, static so the call
// allocates nothing; g_na_mu serialises their use across threads.
static std::mutex g_na_mu;
static DevJob g_na_jobs[PCME_NOALLOC_MAX_JOBS];
static int2 g_na_tickets[PCME_NOALLOC_BYTES / 8];

static size_t up256(size_t v) { return (v + 255) & ~(size_t)255; }

// The ME-only field search of pcme_search_frames_sync with its workspace in
// `scratch`: the same job layout, tickets, queues and kernels as a plan.
static int search_frames_noalloc(const pcme_frame_job *jobs, int njobs, hipStream_t stream) {
  if (njobs == 0) return PCME_OK;
  int rc;
  if ((rc = check_jobs(jobs, njobs))) return rc;
  if (njobs > PCME_NOALLOC_MAX_JOBS)
    return fail(PCME_EINVAL, "pcme_search_frames_noalloc: %d jobs exceed the limit of %d", njobs,
                PCME_NOALLOC_MAX_JOBS);
  size_t words = 0, mbs = 0;
  int maxh = 0, hinted = 0, mr = 0;
  for (int i = 0; i < njobs; i++) {
    words += (size_t)jobs[i].prm.mbw * jobs[i].prm.mbh * 4;
    mbs += (size_t)jobs[i].prm.mbw * jobs[i].prm.mbh;
    maxh = std::max(maxh, jobs[i].prm.mbh);
    hinted |= jobs[i].hint != nullptr;
    mr |= jobs[i].nref > 1;
  }
  size_t ntk = 0;
  for (int i = 0; i < njobs; i++) ntk += (size_t)jobs[i].prm.mbh;
  const size_t off_jobs = up256(CTL_BYTES), off_tk = off_jobs + up256((size_t)njobs * sizeof(DevJob));
  const size_t off_x = off_tk + up256(ntk * sizeof(int2)), need = off_x + words * 4;
  if (need > PCME_NOALLOC_BYTES)
    return fail(PCME_EINVAL,
                "pcme_search_frames_noalloc: %zu macroblocks need %zu bytes of workspace, more than its %u "
                "(about %u 1080p frames)",
                mbs, need, PCME_NOALLOC_BYTES, (unsigned)(PCME_NOALLOC_BYTES / (8160u * 16u + 600u)));
  Dev *dv = nullptr;
  if ((rc = get_dev(&dv))) return rc;
  std::lock_guard<std::mutex> g0(g_na_mu);
  std::lock_guard<std::mutex> g(dv->mu);
  uint8_t *base = dv->scratch;
  uint32_t *ctl = (uint32_t *)base;
  DevJob *djobs = (DevJob *)(base + off_jobs);
  int2 *dtk = (int2 *)(base + off_tk);
  uint32_t *xchg = (uint32_t *)(base + off_x);
  // the job table and tickets as plan_create builds them
  size_t off = 0, total = mbs, cum = 0;
  int qof[PCME_NOALLOC_MAX_JOBS];
  for (int i = 0; i < njobs; i++) {
    const pcme_frame_job &J = jobs[i];
    DevJob &d = g_na_jobs[i];
    d = dev_job(J);
    d.xchg = xchg + off;
    const size_t m = (size_t)J.prm.mbw * J.prm.mbh;
    off += m * 4;
    qof[i] = (int)std::min<size_t>(PCME_QUEUES - 1, (cum + m / 2) * PCME_QUEUES / std::max<size_t>(total, 1));
    cum += m;
  }
  uint32_t qoff[PCME_QUEUES + 1];
  size_t nt = 0;
  for (int q = 0; q < PCME_QUEUES; q++) {
    qoff[q] = (uint32_t)nt;
    for (int y = 0; y < maxh; y++)
      for (int i = 0; i < njobs; i++)
        if (qof[i] == q && y < g_na_jobs[i].mbh) g_na_tickets[nt++] = make_int2(i, y);
  }
  qoff[PCME_QUEUES] = (uint32_t)nt;
  if (!dv->scratch_init) {  // the epoch word starts at zero on each device
    HIPCHK(hipMemsetAsync(ctl, 0, CTL_BYTES, stream));
    dv->scratch_init = true;
  }
  // every call clears its exchange words and the sticky stall flag: the
  // array keeps the previous call's contents, whatever its layout was
  HIPCHK(hipMemsetAsync(ctl + 4, 0, 4, stream));
  HIPCHK(hipMemsetAsync(xchg, 0, words * 4, stream));
  HIPCHK(hipMemcpyAsync(ctl + 16, qoff, sizeof qoff, hipMemcpyHostToDevice, stream));
  HIPCHK(hipMemcpyAsync(djobs, g_na_jobs, (size_t)njobs * sizeof(DevJob), hipMemcpyHostToDevice, stream));
  HIPCHK(hipMemcpyAsync(dtk, g_na_tickets, nt * sizeof(int2), hipMemcpyHostToDevice, stream));
  const int grid = std::min((int)nt, dv->resident_waves[(hinted ? 1 : 0) + (mr ? 2 : 0)]);
  hipLaunchKernelGGL(prep_kernel, dim3(1), dim3(64), 0, stream, ctl);
  HIPCHK(hipGetLastError());
  hipLaunchKernelGGL(clear_kernel, dim3(512), dim3(256), 0, stream, ctl, xchg, words);
  HIPCHK(hipGetLastError());
  hipLaunchKernelGGL(frame_kernel(hinted != 0, mr != 0), dim3((unsigned)grid), dim3(64), 0,
                     stream, djobs, dtk, (int)nt, ctl, kSpinLimit);
  HIPCHK(hipGetLastError());
  // the host tables are reused by the next call: wait here (the call is
  // synchronous like pcme_search_frames_sync) and report a stall
  HIPCHK(hipStreamSynchronize(stream));
  uint32_t flag = 0;
  HIPCHK(hipMemcpy(&flag, ctl + 4, 4, hipMemcpyDeviceToHost));
  if (flag) return fail(PCME_ESTALL, "a wavefront dependency wait exceeded its bound");
  return PCME_OK;
}

}  // namespace pcme

using namespace pcme;

extern "C" {

int pcme_version(void) { return 1; }

const char *pcme_last_error(void) { return g_err; }

int pcme_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  int k = 0;
  for (int i = 0; i < n; i++) k += is_gfx950(i) ? 1 : 0;
  return k;
}

uint32_t pcme_lambda(int32_t qp) {
  if (qp < 0) qp = 0;
  if (qp > 51) qp = 51;
  return lambda_of(qp);
}

void pcme_qpel_table(uint32_t out[4]) {
  out[0] = 0x41111000u;
  out[1] = 0x61312120u;
  out[2] = 0x63333222u;
  out[3] = 0x96939282u;
}

int pcme_pad_planes(const uint8_t *src, int64_t src_pitch, int32_t sstride, int32_t w, int32_t h,
                    uint8_t *dst, int64_t dst_pitch, int32_t dstride, int32_t aw, int32_t ah, int32_t pad,
                    int32_t n, void *stream) {
  if (n < 0 || w < 1 || h < 1 || w > aw || h > ah || sstride < w || pad < 0 || dstride < aw + 2 * pad ||
      !src || !dst)
    return fail(PCME_EINVAL, "pcme_pad_planes: bad geometry");
  if ((dstride & 3) || (dst_pitch & 3) || ((uintptr_t)dst & 3))
    return fail(PCME_EINVAL, "pcme_pad_planes: dst stride/pitch/pointer must be 4-byte aligned");
  if (n == 0) return PCME_OK;
  Dev *dv;
  int rc = get_dev(&dv);
  if (rc) return rc;
  const int rows = ah + 2 * pad;
  dim3 grid((unsigned)((dstride / 4 + 255) / 256), (unsigned)rows, (unsigned)n);
  hipLaunchKernelGGL(pad_planes_kernel, grid, dim3(256), 0, (hipStream_t)stream, src, src_pitch, sstride, w,
                     h, dst, dst_pitch, dstride, rows, pad);
  HIPCHK(hipGetLastError());
  return PCME_OK;
}

int pcme_hpel(const uint8_t *ref, uint8_t *h, uint8_t *v, uint8_t *c, int64_t pitch, int32_t n, int32_t mbw,
              int32_t mbh, void *stream) {
  if (n < 0 || mbw < 1 || mbh < 1 || !ref || !h || !v || !c) return fail(PCME_EINVAL, "pcme_hpel: bad arguments");
  const int stride = 16 * mbw + 2 * PCME_PAD, rows = 16 * mbh + 2 * PCME_PAD;
  if (pitch < (int64_t)stride * rows || (pitch & 3) || (((uintptr_t)h | (uintptr_t)v | (uintptr_t)c) & 3))
    return fail(PCME_EINVAL, "pcme_hpel: pitch smaller than a plane or not 4-byte aligned");
  if (n == 0) return PCME_OK;
  Dev *dv;
  int rc = get_dev(&dv);
  if (rc) return rc;
  dim3 grid((unsigned)((stride + HT_W - 1) / HT_W), (unsigned)((rows + HT_H - 1) / HT_H), (unsigned)n);
  hipLaunchKernelGGL(hpel_kernel, grid, dim3(256), 0, (hipStream_t)stream, ref, h, v, c, pitch, stride, rows);
  HIPCHK(hipGetLastError());
  return PCME_OK;
}

int pcme_luma_planes(const uint8_t *src, int64_t src_pitch, int32_t sstride, int32_t w, int32_t h,
                     uint8_t *full, uint8_t *hp, uint8_t *vp, uint8_t *cp, int64_t pitch, int32_t n, int32_t mbw,
                     int32_t mbh, void *stream) {
  if (n < 0 || mbw < 1 || mbh < 1 || w < 1 || h < 1 || w > 16 * mbw || h > 16 * mbh || sstride < w || !src ||
      !full || !hp || !vp || !cp)
    return fail(PCME_EINVAL, "pcme_luma_planes: bad geometry");
  const int stride = 16 * mbw + 2 * PCME_PAD, rows = 16 * mbh + 2 * PCME_PAD;
  if (pitch < (int64_t)stride * rows || (pitch & 3) ||
      (((uintptr_t)full | (uintptr_t)hp | (uintptr_t)vp | (uintptr_t)cp) & 3))
    return fail(PCME_EINVAL, "pcme_luma_planes: pitch smaller than a plane or not 4-byte aligned");
  if (n == 0) return PCME_OK;
  Dev *dv;
  int rc = get_dev(&dv);
  if (rc) return rc;
  dim3 grid((unsigned)((stride / 4 + 255) / 256), (unsigned)((rows + LP_TH - 1) / LP_TH), (unsigned)n);
  hipLaunchKernelGGL(luma_planes_kernel, grid, dim3(256), 0, (hipStream_t)stream, src, src_pitch, sstride, w, h,
                     full, hp, vp, cp, pitch, stride, rows);
  HIPCHK(hipGetLastError());
  return PCME_OK;
}

static int lp_launch(const LpBatch &b, int64_t blocks, hipStream_t s) {
  hipLaunchKernelGGL(luma_planes_batch_kernel, dim3((unsigned)blocks), dim3(256), 0, s, b);
  HIPCHK(hipGetLastError());
  return PCME_OK;
}

int pcme_luma_planes_batch(const pcme_planes_job *jobs, int32_t njobs, void *stream) {
  if (njobs < 0 || (njobs > 0 && !jobs)) return fail(PCME_EINVAL, "pcme_luma_planes_batch: bad job array");
  for (int i = 0; i < njobs; i++) {  // every job checked before anything is launched
    const pcme_planes_job &J = jobs[i];
    if (J.n < 0 || J.mbw < 1 || J.mbh < 1 || J.w < 1 || J.h < 1 || J.w > 16 * J.mbw || J.h > 16 * J.mbh ||
        J.sstride < J.w || !J.src || !J.full || !J.hp || !J.vp || !J.cp)
      return fail(PCME_EINVAL, "pcme_luma_planes_batch: job %d: bad geometry", i);
    const int64_t stride = 16 * J.mbw + 2 * PCME_PAD, rows = 16 * J.mbh + 2 * PCME_PAD;
    if (J.pitch < stride * rows || (J.pitch & 3) ||
        (((uintptr_t)J.full | (uintptr_t)J.hp | (uintptr_t)J.vp | (uintptr_t)J.cp) & 3))
      return fail(PCME_EINVAL, "pcme_luma_planes_batch: job %d: pitch smaller than a plane or not 4-byte aligned",
                  i);
  }
  Dev *dv;
  int rc = get_dev(&dv);
  if (rc) return rc;
  const hipStream_t s = (hipStream_t)stream;
  LpBatch b{};
  int64_t blocks = 0;
  for (int i = 0; i < njobs; i++) {
    const pcme_planes_job &J = jobs[i];
    if (J.n == 0) continue;
    const int stride = 16 * J.mbw + 2 * PCME_PAD, rows = 16 * J.mbh + 2 * PCME_PAD;
    const int ngx = stride / 4, nbands = (rows + LP_TH - 1) / LP_TH, bpf = (ngx * nbands + 255) / 256;
    if (b.n == LP_MAXJOBS || blocks + (int64_t)bpf * J.n > INT32_MAX) {
      if ((rc = lp_launch(b, blocks, s))) return rc;
      b.n = 0;
      blocks = 0;
    }
    if ((int64_t)bpf * J.n > INT32_MAX) return fail(PCME_EINVAL, "pcme_luma_planes_batch: job %d too large", i);
    b.j[b.n++] = LpJob{J.src, J.full, J.hp, J.vp, J.cp, J.src_pitch, J.pitch, J.sstride, J.w, J.h, stride, rows,
                       ngx, nbands, bpf, (int)blocks, 0};
    blocks += (int64_t)bpf * J.n;
  }
  return b.n ? lp_launch(b, blocks, s) : PCME_OK;
}

int pcme_search_frames(const pcme_frame_job *jobs, int njobs, void *stream) {
  return search_frames_impl(jobs, njobs, (hipStream_t)stream, false);
}

int pcme_search_frames_sync(const pcme_frame_job *jobs, int njobs, void *stream) {
  return search_frames_impl(jobs, njobs, (hipStream_t)stream, true);
}

int pcme_search_frames_noalloc(const pcme_frame_job *jobs, int njobs, void *stream) {
  return search_frames_noalloc(jobs, njobs, (hipStream_t)stream);
}

int pcme_search_frames_check(void *stream) {
  Dev *dv = nullptr;
  int rc;
  if ((rc = get_dev(&dv))) return rc;
  std::lock_guard<std::mutex> g(dv->mu);
  if (!dv->cached) return PCME_OK;
  return plan_check(dv->cached, (hipStream_t)stream);
}

int pcme_plan_create(const pcme_frame_job *jobs, int njobs, pcme_plan **out) {
  if (!out) return fail(PCME_EINVAL, "pcme_plan_create: null out");
  *out = nullptr;
  int rc;
  if ((rc = check_jobs(jobs, njobs))) return rc;
  Dev *dv = nullptr;
  if ((rc = get_dev(&dv))) return rc;
  Plan *p = nullptr;
  if ((rc = plan_create(dv, jobs, njobs, &p))) return rc;
  *out = (pcme_plan *)p;
  return PCME_OK;
}

int pcme_plan_launch(pcme_plan *plan, void *stream) {
  if (!plan) return fail(PCME_EINVAL, "pcme_plan_launch: null plan");
  return plan_launch((Plan *)plan, (hipStream_t)stream);
}

int pcme_plan_check(pcme_plan *plan, void *stream) {
  if (!plan) return fail(PCME_EINVAL, "pcme_plan_check: null plan");
  return plan_check((Plan *)plan, (hipStream_t)stream);
}

int pcme_plan_set_grid(pcme_plan *plan, int32_t waves) {
  if (!plan || waves < 0) return fail(PCME_EINVAL, "pcme_plan_set_grid: bad arguments");
  Plan *p = (Plan *)plan;
  int w = waves == 0 ? plan_resident(p) : waves;
  p->grid = std::max(1, std::min(w, p->ntickets));
  return p->grid;
}

int pcme_plan_get_config(pcme_plan *plan, pcme_tune *out) {
  if (!plan || !out) return fail(PCME_EINVAL, "pcme_plan_get_config: bad arguments");
  const Plan *p = (const Plan *)plan;
  const Variant &v = kVariants[p->variant];
  out->waves = p->grid;
  out->waves_per_group = v.wpg;
  out->waves_per_cu =
      ((p->hinted || p->mr) && !plan_16(p)) ? std::max(1, p->resident / std::max(1, p->cus)) : v.wcu;
  out->reuse_regions = v.reuse ? 1 : 0;
  out->ms = 0.f;
  return PCME_OK;
}

int pcme_plan_set_config(pcme_plan *plan, const pcme_tune *cfg) {
  if (!plan || !cfg || cfg->waves < 0) return fail(PCME_EINVAL, "pcme_plan_set_config: bad arguments");
  Plan *p = (Plan *)plan;
  int v = -1;
  for (int i = 0; i < kNumVariants; i++)
    if (kVariants[i].wpg == cfg->waves_per_group && kVariants[i].wcu == cfg->waves_per_cu &&
        kVariants[i].reuse == (cfg->reuse_regions != 0))
      v = i;
  if (v < 0)
    return fail(PCME_EINVAL, "no kernel configuration with %d waves per group, %d waves per CU, region reuse %d",
                cfg->waves_per_group, cfg->waves_per_cu, cfg->reuse_regions);
  if (!plan_allows(p, v))
    return fail(PCME_EINVAL, "a plan with pruner hints or several references runs the default configuration or "
                             "(4, 16, 1)");
  p->variant = v;
  const int w = cfg->waves == 0 ? plan_resident(p) : cfg->waves;
  p->grid = std::max(1, std::min(w, p->ntickets));
  return PCME_OK;
}

int pcme_plan_autotune(pcme_plan *plan, void *stream, int32_t reps, int32_t *best_waves, float *best_ms) {
  if (!plan || reps < 1) return fail(PCME_EINVAL, "pcme_plan_autotune: bad arguments");
  Plan *p = (Plan *)plan;
  hipStream_t st = (hipStream_t)stream;
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  // Candidate persistent grids: fractions of the resident capacity.  Fewer
  // waves per SIMD give each row walker more issue slots (shorter critical
  // path); more waves hide latency and keep more rows in flight.
  const int fr[] = {16, 14, 12, 10, 8, 6};
  int best = p->grid;
  float bms = 1e30f;
  int rc = PCME_OK;
  for (int f : fr) {
    int w = std::max(1, std::min(p->ntickets, plan_resident(p) * f / 16));
    p->grid = w;
    if ((rc = plan_launch(p, st))) break;  // warm-up
    HIPCHK(hipEventRecord(e0, st));
    for (int r = 0; r < reps && !rc; r++) rc = plan_launch(p, st);
    if (rc) break;
    HIPCHK(hipEventRecord(e1, st));
    HIPCHK(hipEventSynchronize(e1));
    if ((rc = plan_check(p, st))) break;
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    ms /= (float)reps;
    if (ms < bms) {
      bms = ms;
      best = w;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  p->grid = best;
  if (best_waves) *best_waves = best;
  if (best_ms) *best_ms = bms;
  return rc;
}

// Times `reps` launches of the plan as configured (after one untimed).
static int time_plan(Plan *p, hipStream_t st, int reps, hipEvent_t e0, hipEvent_t e1, float *ms) {
  int rc;
  if ((rc = plan_launch(p, st))) return rc;
  HIPCHK(hipEventRecord(e0, st));
  for (int r = 0; r < reps; r++)
    if ((rc = plan_launch(p, st))) return rc;
  HIPCHK(hipEventRecord(e1, st));
  HIPCHK(hipEventSynchronize(e1));
  if ((rc = plan_check(p, st))) return rc;
  HIPCHK(hipEventElapsedTime(ms, e0, e1));
  *ms /= (float)reps;
  return PCME_OK;
}

int pcme_plan_autotune_ex(pcme_plan *plan, void *stream, int32_t reps, pcme_tune *best) {
  if (!plan || reps < 1) return fail(PCME_EINVAL, "pcme_plan_autotune_ex: bad arguments");
  Plan *p = (Plan *)plan;
  hipStream_t st = (hipStream_t)stream;
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  const int fr[] = {16, 14, 12, 10, 8, 6};
  int bv = p->variant, bg = p->grid, rc = PCME_OK;
  float bms = 1e30f;
  for (int v = 0; v < kNumVariants && !rc; v++) {
    if (!plan_allows(p, v)) continue;
    p->variant = v;
    for (int f : fr) {
      p->grid = std::max(1, std::min(p->ntickets, plan_resident(p) * f / 16));
      float ms = 0.f;
      if ((rc = time_plan(p, st, reps, e0, e1, &ms))) break;
      if (ms < bms) {
        bms = ms;
        bv = v;
        bg = p->grid;
      }
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  p->variant = bv;
  p->grid = bg;
  if (best) {
    pcme_plan_get_config(plan, best);
    best->ms = bms;
  }
  return rc;
}

uint32_t pcme_set_spin_limit(uint32_t polls) {
  uint32_t old = kSpinLimit;
  kSpinLimit = polls == 0 ? (1u << 22) : polls;
  return old;
}

void pcme_plan_destroy(pcme_plan *plan) {
  if (!plan) return;
  (void)hipDeviceSynchronize();
  plan_free((Plan *)plan);
}

int pcme_search_batch(const pcme_search_job *jobs, int njobs, pcme_mv *out_mv, uint32_t *out_cost,
                      void *stream) {
  if (njobs < 0 || (njobs > 0 && (!jobs || !out_mv || !out_cost)))
    return fail(PCME_EINVAL, "pcme_search_batch: bad arguments");
  for (int i = 0; i < njobs; i++) {
    const pcme_search_job &J = jobs[i];
    bool shape = (J.w == 16 || J.w == 8) && (J.h == 16 || J.h == 8);
    if (!shape) return fail(PCME_EINVAL, "job %d: block %dx%d not one of 16x16/16x8/8x16/8x8", i, J.w, J.h);
    if (!J.src || !J.ref || !J.ref_h || !J.ref_v || !J.ref_c)
      return fail(PCME_EINVAL, "job %d has a null pointer", i);
    if ((((uintptr_t)J.ref | (uintptr_t)J.ref_h | (uintptr_t)J.ref_v | (uintptr_t)J.ref_c) & 3u) != 0)
      return fail(PCME_EINVAL, "job %d: reference planes must be 4-byte aligned", i);
    if (((uintptr_t)J.src & 3u) || (J.sstride & 3) || J.sstride < J.w)
      return fail(PCME_EINVAL, "job %d: src must be 4-byte aligned with a stride multiple of 4", i);
    if (J.mbw < 1 || J.mbh < 1 || J.bx < 0 || J.by < 0 || J.bx + J.w > 16 * J.mbw || J.by + J.h > 16 * J.mbh)
      return fail(PCME_EINVAL, "job %d: block outside the picture", i);
    if (J.ncands < 0 || J.ncands > 2) return fail(PCME_EINVAL, "job %d: ncands %d not in [0,2]", i, J.ncands);
    if (J.me_range < 0 || J.me_range > 1024 || J.subpel_rounds < 0 || J.subpel_rounds > 8)
      return fail(PCME_EINVAL, "job %d: me_range/subpel_rounds out of range", i);
    if (J.vlim < J.by + J.h + 1 || J.vlim > 16LL * J.mbh + 44)
      return fail(PCME_EINVAL, "job %d: vlim %lld outside [by+h+1, 16*mbh+44]", i, (long long)J.vlim);
    int ax = J.mvp.x < 0 ? -J.mvp.x : J.mvp.x, ay = J.mvp.y < 0 ? -J.mvp.y : J.mvp.y;
    if (ax >= 4096 || ay >= 4096) return fail(PCME_EINVAL, "job %d: mvp out of range", i);
  }
  if (njobs == 0) return PCME_OK;
  Dev *dv;
  int rc = get_dev(&dv);
  if (rc) return rc;
  const size_t bytes = sizeof(pcme_search_job) * (size_t)njobs;
  std::lock_guard<std::mutex> g(dv->mu);
  WsGuard wg;
  wg.s = (hipStream_t)stream;
  CallWs *w = nullptr;
  if ((rc = take_ws(dv, wg.s, bytes, bytes, &w))) return rc;
  wg.w = w;
  if ((rc = ws_upload(w, jobs, bytes, (hipStream_t)stream))) return rc;
  hipLaunchKernelGGL(search_batch_kernel, dim3((unsigned)njobs), dim3(64), 0, (hipStream_t)stream,
                     (const pcme_search_job *)w->dev, njobs, out_mv, out_cost);
  HIPCHK(hipGetLastError());
  return PCME_OK;
}

static bool nb_ok(const pcme_nb_block &b) {
  const int ax = b.mv.x < 0 ? -b.mv.x : b.mv.x, ay = b.mv.y < 0 ? -b.mv.y : b.mv.y;
  return b.ref >= -2 && ax < 4096 && ay < 4096;
}

int pcme_search_part_batch(const pcme_part_job *jobs, int njobs, pcme_part_out *out, void *stream) {
  if (njobs < 0 || (njobs > 0 && (!jobs || !out)) || ((uintptr_t)out & 3u))
    return fail(PCME_EINVAL, "pcme_search_part_batch: bad arguments");
  for (int i = 0; i < njobs; i++) {
    const pcme_part_job &J = jobs[i];
    if (J.part < PCME_P16x16 || J.part > PCME_P8x8) return fail(PCME_EINVAL, "job %d: part %d", i, J.part);
    if (!J.src || !J.ref || !J.ref_h || !J.ref_v || !J.ref_c) return fail(PCME_EINVAL, "job %d has a null pointer", i);
    if ((((uintptr_t)J.ref | (uintptr_t)J.ref_h | (uintptr_t)J.ref_v | (uintptr_t)J.ref_c) & 3u) != 0)
      return fail(PCME_EINVAL, "job %d: reference planes must be 4-byte aligned", i);
    if (((uintptr_t)J.src & 3u) || (J.sstride & 3) || J.sstride < 16)
      return fail(PCME_EINVAL, "job %d: src must be 4-byte aligned with a stride multiple of 4", i);
    if (J.mbw < 1 || J.mbh < 1 || J.mx < 0 || J.my < 0 || J.mx >= J.mbw || J.my >= J.mbh)
      return fail(PCME_EINVAL, "job %d: MB outside the picture", i);
    if (J.me_range < 0 || J.me_range > 1024 || J.subpel_rounds < 0 || J.subpel_rounds > 8)
      return fail(PCME_EINVAL, "job %d: me_range/subpel_rounds out of range", i);
    if (J.vlim < 16LL * J.my + 17 || J.vlim > 16LL * J.mbh + 44)
      return fail(PCME_EINVAL, "job %d: vlim %lld outside [16*my+17, 16*mbh+44]", i, (long long)J.vlim);
    pcme_nb_block sb = {J.seed, 0, {0, 0, 0}};
    bool ok = nb_ok(sb) && nb_ok(J.top_right) && nb_ok(J.top_left);
    for (int k = 0; k < 4; k++) ok = ok && nb_ok(J.left[k]) && nb_ok(J.top[k]);
    if (!ok) return fail(PCME_EINVAL, "job %d: a neighbour ref < -2 or an MV component beyond +-4095", i);
  }
  if (njobs == 0) return PCME_OK;
  Dev *dv;
  int rc = get_dev(&dv);
  if (rc) return rc;
  const size_t bytes = sizeof(pcme_part_job) * (size_t)njobs;
  std::lock_guard<std::mutex> g(dv->mu);
  WsGuard wg;
  wg.s = (hipStream_t)stream;
  CallWs *w = nullptr;
  if ((rc = take_ws(dv, wg.s, bytes, bytes, &w))) return rc;
  wg.w = w;
  if ((rc = ws_upload(w, jobs, bytes, (hipStream_t)stream))) return rc;
  hipLaunchKernelGGL(search_part_batch_kernel, dim3((unsigned)njobs), dim3(64), 0, (hipStream_t)stream,
                     (const pcme_part_job *)w->dev, njobs, out);
  HIPCHK(hipGetLastError());
  return PCME_OK;
}

int pcme_b_analyse_batch(const pcme_b_job *jobs, int njobs, pcme_b_out *out, void *stream) {
  if (njobs < 0 || (njobs > 0 && (!jobs || !out)) || ((uintptr_t)out & 3u))
    return fail(PCME_EINVAL, "pcme_b_analyse_batch: bad arguments");
  for (int i = 0; i < njobs; i++) {
    const pcme_b_job &J = jobs[i];
    if (!J.src || ((uintptr_t)J.src & 3u) || (J.sstride & 3) || J.sstride < 16)
      return fail(PCME_EINVAL, "job %d: src null, not 4-byte aligned or stride not a multiple of 4", i);
    for (int l = 0; l < 2; l++)
      for (int k = 0; k < 4; k++) {
        const uint8_t *pl = l ? J.ref1[k] : J.ref0[k];
        if (!pl || ((uintptr_t)pl & 3u)) return fail(PCME_EINVAL, "job %d: reference plane null or misaligned", i);
      }
    if (J.mbw < 1 || J.mbh < 1 || J.mx < 0 || J.my < 0 || J.mx >= J.mbw || J.my >= J.mbh)
      return fail(PCME_EINVAL, "job %d: MB outside the picture", i);
    if (J.me_range < 0 || J.me_range > 1024 || J.subpel_rounds < 0 || J.subpel_rounds > 8)
      return fail(PCME_EINVAL, "job %d: me_range/subpel_rounds out of range", i);
    if (J.vlim < 16LL * J.my + 17 || J.vlim > 16LL * J.mbh + 44)
      return fail(PCME_EINVAL, "job %d: vlim %lld outside [16*my+17, 16*mbh+44]", i, (long long)J.vlim);
    bool ok = true;
    for (int l = 0; l < 2; l++) {
      for (int k = 0; k < 10; k++) ok = ok && nb_ok(J.nb[l][k]);
      for (int q = 0; q < 4; q++) {
        pcme_nb_block cb = {J.col_mv[l][q], J.col_ref[l][q], {0, 0, 0}};
        ok = ok && nb_ok(cb) && J.col_ref[l][q] >= -1;
      }
    }
    if (!ok) return fail(PCME_EINVAL, "job %d: a neighbour or co-located ref/MV out of range", i);
  }
  if (njobs == 0) return PCME_OK;
  Dev *dv;
  int rc = get_dev(&dv);
  if (rc) return rc;
  const size_t bytes = sizeof(pcme_b_job) * (size_t)njobs;
  std::lock_guard<std::mutex> g(dv->mu);
  WsGuard wg;
  wg.s = (hipStream_t)stream;
  CallWs *w = nullptr;
  if ((rc = take_ws(dv, wg.s, bytes, bytes, &w))) return rc;
  wg.w = w;
  if ((rc = ws_upload(w, jobs, bytes, (hipStream_t)stream))) return rc;
  const pcme_b_job *dj = (const pcme_b_job *)w->dev;
  hipLaunchKernelGGL(b_analyse_16_kernel, dim3((unsigned)njobs), dim3(64), 0, (hipStream_t)stream, dj, njobs, out);
  HIPCHK(hipGetLastError());
  hipLaunchKernelGGL(b_analyse_part_kernel<0>, dim3((unsigned)njobs), dim3(64), 0, (hipStream_t)stream, dj, njobs,
                     out);
  HIPCHK(hipGetLastError());
  hipLaunchKernelGGL(b_analyse_part_kernel<1>, dim3((unsigned)njobs), dim3(64), 0, (hipStream_t)stream, dj, njobs,
                     out);
  HIPCHK(hipGetLastError());
  return PCME_OK;
}

static_assert(sizeof(pcme_b_field_job) == 104, "pcme_b_field_job layout");

static int check_b_field_job(const pcme_b_field_job &F, int i) {
  if (F.mbw < 1 || F.mbh < 1 || F.mbw > 4096 || F.mbh > 4096 || !F.src || !F.col || ((uintptr_t)F.src & 3u) ||
      ((uintptr_t)F.col & 3u))
    return fail(PCME_EINVAL, "pcme_b_fields: picture %d: bad size, source or co-located array", i);
  for (int k = 0; k < 4; k++)
    if (!F.ref0[k] || !F.ref1[k] || (((uintptr_t)F.ref0[k] | (uintptr_t)F.ref1[k]) & 3u))
      return fail(PCME_EINVAL, "pcme_b_fields: picture %d: reference plane null or misaligned", i);
  if (F.qp < 0 || F.qp > 51 || F.me_range < 0 || F.me_range > 1024 || F.subpel_rounds < 0 || F.subpel_rounds > 8)
    return fail(PCME_EINVAL, "pcme_b_fields: picture %d: qp, me_range or subpel_rounds out of range", i);
  return PCME_OK;
}

// Upper bounds of one pcme_b_fields call: pictures, and MBs over all of them.
static const int kMaxBPics = 1024;
static const size_t kMaxBMbs = (size_t)1 << 22;

int pcme_b_fields(const pcme_b_field_job *jobs, int njobs, pcme_b_mb *const *outs, void *stream) {
  if (njobs < 0 || njobs > kMaxBPics || (njobs > 0 && (!jobs || !outs)))
    return fail(PCME_EINVAL, "pcme_b_fields: bad arguments (njobs %d, at most %d)", njobs, kMaxBPics);
  int rc;
  size_t n = 0;
  int maxh = 0;
  for (int i = 0; i < njobs; i++) {
    if (!outs[i] || ((uintptr_t)outs[i] & 7u)) return fail(PCME_EINVAL, "pcme_b_fields: output %d null or misaligned", i);
    if ((rc = check_b_field_job(jobs[i], i))) return rc;
    n += (size_t)jobs[i].mbw * jobs[i].mbh;
    maxh = std::max(maxh, jobs[i].mbh);
  }
  if (n > kMaxBMbs) return fail(PCME_EINVAL, "pcme_b_fields: %zu MBs in one call, at most %zu", n, kMaxBMbs);
  if (njobs == 0) return PCME_OK;
  Dev *dv;
  if ((rc = get_dev(&dv))) return rc;
  // host image: per-MB jobs (raster within each picture) | pictures | tickets
  std::vector<pcme_b_job> bj(n);
  std::vector<BPic> pics((size_t)njobs);
  size_t k0 = 0;
  for (int i = 0; i < njobs; i++) {
    const pcme_b_field_job &F = jobs[i];
    const int mbw = F.mbw, mbh = F.mbh;
    const int stride = 16 * mbw + 2 * PCME_PAD;
    const uint32_t lam = lambda_of(F.qp);
    pics[(size_t)i] = BPic{F.col, outs[i], (int)k0, mbw, mbh, F.partitions ? 1 : 0};
    for (int my = 0; my < mbh; my++)
      for (int mx = 0; mx < mbw; mx++) {
        pcme_b_job &J = bj[k0 + (size_t)my * mbw + mx];
        memset(&J, 0, sizeof J);
        J.src = F.src + (size_t)(PCME_PAD + 16 * my) * stride + PCME_PAD + 16 * mx;
        J.sstride = stride;
        for (int q = 0; q < 4; q++) {
          J.ref0[q] = F.ref0[q];
          J.ref1[q] = F.ref1[q];
        }
        J.mbw = mbw;
        J.mbh = mbh;
        J.mx = mx;
        J.my = my;
        J.lam = lam;
        J.me_range = F.me_range;
        J.subpel_rounds = F.subpel_rounds;
        const int need = 16 * (my + 1) + 96;  // frame.rs:152 with picture.rs:261-265
        J.vlim = need >= 16 * mbh ? 16 * mbh + 44 : need;
      }
    k0 += (size_t)mbw * mbh;
  }
  // (row, picture) order: the row above of a picture is an earlier ticket
  std::vector<int2> tk;
  for (int y = 0; y < maxh; y++)
    for (int i = 0; i < njobs; i++)
      if (y < jobs[i].mbh) tk.push_back(make_int2(i, y));
  const size_t jb = sizeof(pcme_b_job) * n, pb = sizeof(BPic) * (size_t)njobs, tb = sizeof(int2) * tk.size();
  const size_t hb = (jb + pb + 15) / 16 * 16 + tb;  // uploaded: jobs | pictures | tickets
  const size_t po = jb, to = (jb + pb + 15) / 16 * 16;
  const size_t ro = (hb + 255) / 256 * 256, rb = sizeof(pcme_b_out) * n;
  const size_t xo = ro + rb, xwords = n * 8, xb = xwords * 4;
  const size_t co = xo + xb;
  std::vector<uint8_t> img(hb);
  memcpy(img.data(), bj.data(), jb);
  memcpy(img.data() + po, pics.data(), pb);
  memcpy(img.data() + to, tk.data(), tb);
  std::lock_guard<std::mutex> g(dv->mu);
  if (!dv->b_stall) {
    uint32_t *p = nullptr;
    HIPCHK(hipGetSymbolAddress((void **)&p, HIP_SYMBOL(b_field_stall)));
    if (!p) return fail(PCME_EHIP, "hipGetSymbolAddress(b_field_stall) returned null");
    dv->b_stall = p;
  }
  WsGuard wg;
  wg.s = (hipStream_t)stream;
  CallWs *w = nullptr;
  if ((rc = take_ws(dv, wg.s, co + 64, hb, &w))) return rc;
  wg.w = w;
  if ((rc = ws_upload(w, img.data(), hb, wg.s))) return rc;
  uint8_t *base = (uint8_t *)w->dev;
  uint32_t *xw = (uint32_t *)(base + xo);
  uint32_t *ctl = (uint32_t *)(base + co);
  hipLaunchKernelGGL(b_walk_clear_kernel, dim3(64), dim3(256), 0, wg.s, xw, xwords, ctl);
  // as many walkers as are resident at once (or rows)
  const unsigned grid = (unsigned)std::min<size_t>(tk.size(), (size_t)dv->b_walk_resident);
  hipLaunchKernelGGL(b_walk_kernel, dim3(grid), dim3(128), 0, wg.s, (const pcme_b_job *)base,
                     (const BPic *)(base + po), (const int2 *)(base + to), (int)tk.size(), (pcme_b_out *)(base + ro),
                     xw, ctl, dv->b_stall, kSpinLimit);
  HIPCHK(hipGetLastError());
  return PCME_OK;
}

int pcme_b_field(const pcme_b_field_job *job, pcme_b_mb *out, void *stream) {
  if (!job || !out) return fail(PCME_EINVAL, "pcme_b_field: bad arguments");
  pcme_b_mb *const outs[1] = {out};
  return pcme_b_fields(job, 1, outs, stream);
}

int pcme_b_field_check(void *stream) {
  Dev *dv;
  int rc = get_dev(&dv);
  if (rc) return rc;
  std::lock_guard<std::mutex> g(dv->mu);
  if (!dv->b_stall) return PCME_OK;  // no B field launched on this device
  HIPCHK(hipStreamSynchronize((hipStream_t)stream));
  uint32_t v = 0u;
  HIPCHK(hipMemcpy(&v, dv->b_stall, sizeof v, hipMemcpyDeviceToHost));
  if (v == 0u) return PCME_OK;
  const uint32_t zero = 0u;
  HIPCHK(hipMemcpy(dv->b_stall, &zero, sizeof zero, hipMemcpyHostToDevice));
  return fail(PCME_ESTALL, "pcme_b_field: a dependency wait exceeded its bound; the records are incomplete");
}

static_assert(sizeof(pcme_skip_job) == 120 && sizeof(pcme_skip_out) == 16, "pcme_skip_* layout");
static_assert(sizeof(pcme_frame_job) == 184 && offsetof(pcme_frame_job, older) == 88 && sizeof(pcme_mb_out) == 128,
              "pcme_frame_job / pcme_mb_out layout");

int pcme_skip_gate_batch(const pcme_skip_job *jobs, int njobs, pcme_skip_out *out, void *stream) {
  if (njobs < 0 || (njobs > 0 && (!jobs || !out)) || ((uintptr_t)out & 3u))
    return fail(PCME_EINVAL, "pcme_skip_gate_batch: bad arguments");
  for (int i = 0; i < njobs; i++) {
    const pcme_skip_job &J = jobs[i];
    bool ok = J.src_y && J.src_u && J.src_v && J.ref_u && J.ref_v;
    for (int k = 0; k < 4; k++) ok = ok && J.ref[k];
    if (!ok) return fail(PCME_EINVAL, "job %d has a null pointer", i);
    if (J.ystride < 16 || J.cstride < 8) return fail(PCME_EINVAL, "job %d: source strides %d/%d", i, J.ystride, J.cstride);
    if (J.mbw < 1 || J.mbh < 1 || J.mx < 0 || J.my < 0 || J.mx >= J.mbw || J.my >= J.mbh)
      return fail(PCME_EINVAL, "job %d: MB outside the picture", i);
    if (J.qp < 0 || J.qp > 51) return fail(PCME_EINVAL, "job %d: qp %d outside [0, 51]", i, J.qp);
    if (J.flags & ~(PCME_SKIP_T8X8 | PCME_SKIP_ADAPTIVE_T8 | PCME_SKIP_CABAC))
      return fail(PCME_EINVAL, "job %d: unknown flags 0x%x", i, J.flags);
    if (J.chroma_qp_offset < -12 || J.chroma_qp_offset > 12)
      return fail(PCME_EINVAL, "job %d: chroma_qp_offset %d outside [-12, 12]", i, J.chroma_qp_offset);
    if (J.vlim < 16LL * J.my + 17 || J.vlim > 16LL * J.mbh + 44)
      return fail(PCME_EINVAL, "job %d: vlim %lld outside [16*my+17, 16*mbh+44]", i, (long long)J.vlim);
  }
  if (njobs == 0) return PCME_OK;
  Dev *dv;
  int rc = get_dev(&dv);
  if (rc) return rc;
  const size_t bytes = sizeof(pcme_skip_job) * (size_t)njobs;
  std::lock_guard<std::mutex> g(dv->mu);
  if (!dv->skip_tab) {
    SkipTables *dst = nullptr;
    HIPCHK(hipGetSymbolAddress((void **)&dst, HIP_SYMBOL(skip_tables)));
    if (!dst) return fail(PCME_EHIP, "hipGetSymbolAddress(skip_tables) returned null");
    std::unique_ptr<SkipTables> h(new SkipTables());
    build_skip_tables(h.get());
    HIPCHK(hipMemcpy(dst, h.get(), sizeof(SkipTables), hipMemcpyHostToDevice));
    HIPCHK(hipStreamSynchronize(nullptr));  // the copy has landed before any stream reads it
    dv->skip_tab = dst;
  }
  WsGuard wg;
  wg.s = (hipStream_t)stream;
  CallWs *w = nullptr;
  if ((rc = take_ws(dv, wg.s, bytes, bytes, &w))) return rc;
  wg.w = w;
  if ((rc = ws_upload(w, jobs, bytes, (hipStream_t)stream))) return rc;
  hipLaunchKernelGGL(skip_gate_kernel, dim3((unsigned)njobs), dim3(64), 0, (hipStream_t)stream,
                     (const pcme_skip_job *)w->dev, njobs, (const SkipTables *)dv->skip_tab, out);
  HIPCHK(hipGetLastError());
  return PCME_OK;
}

int pcme_pred_frame(const uint8_t *ref, const uint8_t *ref_h, const uint8_t *ref_v, const uint8_t *ref_c,
                    const uint8_t *ref_u, const uint8_t *ref_vv, const pcme_mb_out *mb, int32_t mbw, int32_t mbh,
                    uint8_t *pred_y, uint8_t *pred_u, uint8_t *pred_v, void *stream) {
  if (!ref || !ref_h || !ref_v || !ref_c || !ref_u || !ref_vv || !mb || !pred_y || !pred_u || !pred_v ||
      mbw < 1 || mbh < 1)
    return fail(PCME_EINVAL, "pcme_pred_frame: bad arguments");
  Dev *dv;
  int rc = get_dev(&dv);
  if (rc) return rc;
  hipLaunchKernelGGL(pred_frame_kernel, dim3((unsigned)(mbw * mbh)), dim3(256), 0, (hipStream_t)stream, ref,
                     ref_h, ref_v, ref_c, ref_u, ref_vv, mb, mbw, mbh, pred_y, pred_u, pred_v);
  HIPCHK(hipGetLastError());
  return PCME_OK;
}


// ------------------------------------------------------------- lookahead

static bool a4(const void *p) { return ((uintptr_t)p & 3u) == 0; }

static int check_lowres_frame(const char *fn, int i, const void *p, bool may_be_null) {
  if (!p) return may_be_null ? PCME_OK : fail(PCME_EINVAL, "%s: job %d: null frame", fn, i);
  if (!a4(p)) return fail(PCME_EINVAL, "%s: job %d: frame not 4-byte aligned", fn, i);
  return PCME_OK;
}

static int check_lowres_dims(const char *fn, int i, int w, int h) {
  if (w < 8 || h < 8 || (w & 7) || (h & 7))
    return fail(PCME_EINVAL, "%s: job %d: lowres size %dx%d is not a positive multiple of 8", fn, i, w, h);
  return PCME_OK;
}

int pcme_lowres(const uint8_t *planes, int64_t plane_pitch, int32_t stride, int32_t pad, int32_t w, int32_t h,
                uint8_t *out, int64_t out_pitch, int32_t n, void *stream) {
  if (n < 0 || w < 16 || h < 16 || (w & 15) || (h & 15) || pad < 0 || stride < w + pad || !planes || !out)
    return fail(PCME_EINVAL, "pcme_lowres: bad geometry (w, h multiples of 16, stride >= w + pad)");
  if (n > 1 && (plane_pitch < (int64_t)(h + pad) * stride || out_pitch < (int64_t)(w / 2) * (h / 2)))
    return fail(PCME_EINVAL, "pcme_lowres: plane or lowres pitch smaller than a frame");
  if (!a4(planes) || !a4(out) || ((plane_pitch | out_pitch) & 3) || ((stride | pad) & 3))
    return fail(PCME_EINVAL, "pcme_lowres: planes, out, pitches, stride and pad must be 4-byte aligned");
  if (n == 0) return PCME_OK;
  Dev *dv;
  int rc = get_dev(&dv);
  if (rc) return rc;
  const int w2 = w / 2;
  dim3 grid((unsigned)((w2 / 4 + 255) / 256), (unsigned)(h / 2), (unsigned)n);
  hipLaunchKernelGGL(lowres_kernel, grid, dim3(256), 0, (hipStream_t)stream, planes, plane_pitch, stride, pad, w2,
                     out, out_pitch);
  HIPCHK(hipGetLastError());
  return PCME_OK;
}

int pcme_complexity(const pcme_cplx_job *jobs, int32_t n, void *stream) {
  static const char *fn = "pcme_complexity";
  if (n < 0 || (n > 0 && !jobs)) return fail(PCME_EINVAL, "%s: bad job list", fn);
  int rc;
  for (int i = 0; i < n; i++) {
    const pcme_cplx_job &j = jobs[i];
    if ((rc = check_lowres_dims(fn, i, j.w, j.h))) return rc;
    if ((rc = check_lowres_frame(fn, i, j.cur, false))) return rc;
    if ((rc = check_lowres_frame(fn, i, j.prev, true))) return rc;
    if (!j.sums || ((uintptr_t)j.sums & 7u)) return fail(PCME_EINVAL, "%s: job %d: sums null or not 8-byte aligned", fn, i);
    if (!a4(j.blocks)) return fail(PCME_EINVAL, "%s: job %d: blocks not 4-byte aligned", fn, i);
  }
  if (n == 0) return PCME_OK;
  Dev *dv;
  if ((rc = get_dev(&dv))) return rc;
  for (int c0 = 0; c0 < n; c0 += LA_CHUNK) {
    const int m = std::min(LA_CHUNK, n - c0);
    CplxArgs A;
    memset(&A, 0, sizeof A);
    int maxnb = 0;
    for (int i = 0; i < m; i++) {
      A.j[i] = jobs[c0 + i];
      maxnb = std::max(maxnb, (A.j[i].w / 8) * (A.j[i].h / 8));
    }
    hipLaunchKernelGGL(complexity_clear_kernel, dim3(1), dim3(2 * LA_CHUNK), 0, (hipStream_t)stream, A, m);
    HIPCHK(hipGetLastError());
    const int per_wg = 4 * CPLX_BPW;
    hipLaunchKernelGGL(complexity_kernel, dim3((unsigned)((maxnb + per_wg - 1) / per_wg), (unsigned)m), dim3(256), 0,
                       (hipStream_t)stream, A);
    HIPCHK(hipGetLastError());
  }
  return PCME_OK;
}

int pcme_aq_offsets(const uint8_t *y, int32_t ystride, const uint8_t *u, const uint8_t *v, int32_t cstride,
                    int32_t mbw, int32_t mbh, float strength, float *out, void *stream) {
  if (!y || !u || !v || !out || mbw < 1 || mbh < 1 || ystride < 16 * mbw || cstride < 8 * mbw)
    return fail(PCME_EINVAL, "pcme_aq_offsets: bad arguments");
  if (!a4(y) || !a4(u) || !a4(v) || !a4(out) || (ystride & 3) || (cstride & 3))
    return fail(PCME_EINVAL, "pcme_aq_offsets: planes, strides and out must be 4-byte aligned");
  Dev *dv;
  int rc = get_dev(&dv);
  if (rc) return rc;
  const int n = mbw * mbh;
  hipLaunchKernelGGL(aq_var_kernel, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, (hipStream_t)stream, y, ystride,
                     u, v, cstride, mbw, mbh, out);
  HIPCHK(hipGetLastError());
  hipLaunchKernelGGL(aq_mean_kernel, dim3(1), dim3(64), 0, (hipStream_t)stream, out, n, strength);
  HIPCHK(hipGetLastError());
  return PCME_OK;
}

int pcme_mbtree_analyse(const pcme_tree_job *jobs, int32_t n, void *stream) {
  static const char *fn = "pcme_mbtree_analyse";
  if (n < 0 || (n > 0 && !jobs)) return fail(PCME_EINVAL, "%s: bad job list", fn);
  int rc;
  for (int i = 0; i < n; i++) {
    const pcme_tree_job &j = jobs[i];
    if ((rc = check_lowres_dims(fn, i, j.w, j.h))) return rc;
    if ((rc = check_lowres_frame(fn, i, j.cur, false))) return rc;
    if ((rc = check_lowres_frame(fn, i, j.ref0, true))) return rc;
    if ((rc = check_lowres_frame(fn, i, j.ref1, true))) return rc;
    if (!j.blocks || !a4(j.blocks)) return fail(PCME_EINVAL, "%s: job %d: blocks null or not 4-byte aligned", fn, i);
  }
  if (n == 0) return PCME_OK;
  Dev *dv;
  if ((rc = get_dev(&dv))) return rc;
  for (int c0 = 0; c0 < n; c0 += LA_CHUNK) {
    const int m = std::min(LA_CHUNK, n - c0);
    TreeArgs A;
    memset(&A, 0, sizeof A);
    int maxbh = 0;
    for (int i = 0; i < m; i++) {
      A.j[i] = jobs[c0 + i];
      maxbh = std::max(maxbh, A.j[i].h / 8);
    }
    hipLaunchKernelGGL(tree_kernel, dim3((unsigned)((maxbh + 3) / 4), (unsigned)m), dim3(256), 0,
                       (hipStream_t)stream, A);
    HIPCHK(hipGetLastError());
  }
  return PCME_OK;
}

int pcme_mbtree_propagate(const pcme_tree_node *nodes, int32_t n, int32_t bw, int32_t bh, float *prop,
                          void *stream) {
  static const char *fn = "pcme_mbtree_propagate";
  if (n < 0 || (n > 0 && !nodes)) return fail(PCME_EINVAL, "%s: bad node list", fn);
  if (bw < 1 || bh < 1 || bw > 65535 || bh > 65535 || (int64_t)bw * bh > (1 << 24))
    return fail(PCME_EINVAL, "%s: bad block grid %dx%d", fn, bw, bh);
  for (int i = 0; i < n; i++) {
    if (!nodes[i].blocks || !a4(nodes[i].blocks))
      return fail(PCME_EINVAL, "%s: node %d: blocks null or not 4-byte aligned", fn, i);
    if (nodes[i].ref[0] < -1 || nodes[i].ref[1] < -1)
      return fail(PCME_EINVAL, "%s: node %d: reference index below -1", fn, i);
  }
  if (n > 0 && (!prop || !a4(prop))) return fail(PCME_EINVAL, "%s: prop null or not 4-byte aligned", fn);
  if (n == 0) return PCME_OK;
  Dev *dv;
  int rc = get_dev(&dv);
  if (rc) return rc;
  std::lock_guard<std::mutex> g(dv->mu);
  WsGuard wg;
  wg.s = (hipStream_t)stream;
  CallWs *w = nullptr;
  if ((rc = take_ws(dv, wg.s, (size_t)n * sizeof(int4), 0, &w))) return rc;
  wg.w = w;
  int4 *dwin = (int4 *)w->dev;
  hipStream_t st = (hipStream_t)stream;
  const int nb = bw * bh;
  const size_t total = (size_t)n * (size_t)nb;
  hipLaunchKernelGGL(tree_zero_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, prop, total);
  HIPCHK(hipGetLastError());
  for (int c0 = 0; c0 < n; c0 += LA_CHUNK) {
    const int m = std::min(LA_CHUNK, n - c0);
    TreeNodeArgs A;
    memset(&A, 0, sizeof A);
    for (int i = 0; i < m; i++) A.b[i] = nodes[c0 + i].blocks;
    hipLaunchKernelGGL(tree_window_kernel, dim3((unsigned)m), dim3(256), 0, st, A, nb, dwin + c0);
    HIPCHK(hipGetLastError());
  }
  // Rounds: node i runs once every node that feeds it (a later node whose
  // reference is i) has run -- its own costs are then complete -- and every
  // later node feeding one of its destinations has run, so each cell still
  // receives its adds in the reference's order (nodes last to first).  The
  // highest node not yet run always qualifies.  Within a round no two jobs
  // share a destination, and no job's destination is a node of the round.
  // A single chain runs one node per round; independent queues in one call
  // (refs never crossing) share rounds.
  std::vector<std::vector<int>> feeders((size_t)n);
  for (int j = 0; j < n; j++)
    for (int l = 0; l < 2; l++) {
      const int r = nodes[j].ref[l];
      if (r >= 0 && r < j && (l == 0 || nodes[j].ref[0] != r)) feeders[(size_t)r].push_back(j);
    }
  std::vector<char> done((size_t)n, 0);
  std::vector<int> left((size_t)n, 0);  // feeders of i not yet run
  for (int i = 0; i < n; i++) left[(size_t)i] = (int)feeders[(size_t)i].size();
  // unrun feeders of r above node i: counted from the top as nodes run
  auto ready = [&](int i) {
    if (left[(size_t)i]) return false;
    for (int l = 0; l < 2; l++) {
      const int r = nodes[i].ref[l];
      if (r < 0 || r >= i) continue;
      for (int j : feeders[(size_t)r])
        if (j > i && !done[(size_t)j]) return false;
    }
    return true;
  };
  std::vector<TreePropJob> jobs;
  std::vector<int> round;
  int remaining = n;
  while (remaining > 0) {
    round.clear();
    jobs.clear();
    for (int i = n - 1; i >= 0; i--)
      if (!done[(size_t)i] && ready(i)) round.push_back(i);
    for (int i : round) {
      for (int l = 0; l < 2; l++) {
        const int r = nodes[i].ref[l];
        if (r < 0 || r >= i) continue;
        if (l == 1 && nodes[i].ref[0] == r) continue;  // one job for both lists
        TreePropJob J;
        J.src = nodes[i].blocks;
        J.mine = prop + (size_t)i * nb;
        J.dst = prop + (size_t)r * nb;
        J.win = dwin + i;
        J.mask = (1 << l) | (l == 0 && nodes[i].ref[1] == r ? 2 : 0);
        J.pad_ = 0;
        jobs.push_back(J);
      }
    }
    const unsigned tiles = (unsigned)(((bw + 15) / 16) * ((bh + 15) / 16));
    for (size_t c0 = 0; c0 < jobs.size(); c0 += TREE_JOBS) {
      const size_t m = std::min((size_t)TREE_JOBS, jobs.size() - c0);
      TreePropArgs A;
      memset(&A, 0, sizeof A);
      for (size_t k = 0; k < m; k++) A.j[k] = jobs[c0 + k];
      hipLaunchKernelGGL(tree_prop_kernel, dim3(tiles, (unsigned)m), dim3(256), 0, st, A, bw, bh);
      HIPCHK(hipGetLastError());
    }
    for (int i : round) {
      done[(size_t)i] = 1;
      remaining--;
      for (int l = 0; l < 2; l++) {
        const int r = nodes[i].ref[l];
        if (r >= 0 && r < i && (l == 0 || nodes[i].ref[0] != r)) left[(size_t)r]--;
      }
    }
  }
  return PCME_OK;
}

int pcme_mbtree_offsets(const pcme_tree_block *blocks, const float *prop, int32_t nblocks, float *out,
                        void *stream) {
  if (nblocks < 0 || (nblocks > 0 && (!blocks || !prop || !out)))
    return fail(PCME_EINVAL, "pcme_mbtree_offsets: bad arguments");
  if (nblocks > 0 && (!a4(blocks) || !a4(prop) || !a4(out)))
    return fail(PCME_EINVAL, "pcme_mbtree_offsets: blocks, prop and out must be 4-byte aligned");
  if (nblocks == 0) return PCME_OK;
  Dev *dv;
  int rc = get_dev(&dv);
  if (rc) return rc;
  hipLaunchKernelGGL(tree_offsets_kernel, dim3((unsigned)((nblocks + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, blocks, prop, nblocks, out);
  HIPCHK(hipGetLastError());
  return PCME_OK;
}

}  // extern "C"
