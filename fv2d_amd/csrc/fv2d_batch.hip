// fv2d_batch.hip -- batched serving engine of include/fv2d_amd.h (fv2d_batch_*).
//
// Many independent simulations ("jobs") of one precision share one GPU:
//  - device memory: one arena allocated at creation; a job takes one contiguous range of it
//    (U, Us, Q0 planes with the single-simulation layout, its dt partials, its gravity row)
//    and gets it back when collected. Admission = a free slot + a free range (first fit).
//  - control: one StepCtl per slot on the device. Each step, batch_reduce_kernel reduces
//    every running job's dt partials and runs its loop control (the same advance_control as
//    the single-simulation path), so jobs at different steps, with different dt, advance in
//    the same launches and stop on the device at their own end.
//  - work list: pack_items_kernel lists (slot, strip segment) items per kernel class
//    (reconstruction, Riemann solver, conduction, viscosity, wall/gravity variant); the
//    persistent batch_stage_kernel of each class pulls items from its run of the list.
//  - launches: each kernel class runs its own chain of steps_per_launch steps (reduce, stage
//    1, stage 2) on a stream of its own, forked from and joined back into the batch stream,
//    so that the drain of one class's persistent kernel overlaps the others' work. The
//    whole launch is captured once as a HIP graph and replayed; the graph reads the job
//    table and the work list from device memory, so it stays valid while jobs come and go.
//    It holds no memset nodes: the work-list counters are cleared by the reduce kernels.
//  - results: read back on a second stream, so collecting a finished job overlaps the
//    launches of the others.
#include <hip/hip_runtime.h>

#include <stdio.h>
#include <string.h>

#include <deque>
#include <map>
#include <string>
#include <vector>

#include "../../include/fv2d_amd.h"
#include "fv2d_aux.h"
#include "fv2d_device.h"
#include "fv2d_params.h"

using namespace fv2d;

namespace {

constexpr int MAX_CLASSES = 16;
constexpr int MAX_SLOTS = 1024;  // pack_items_kernel: one thread per slot
constexpr size_t ARENA_ALIGN = 256;

// ------------------------------------------------------------------ device kernels

// One workgroup per slot: reduce the dt partials of each running job of class `cls` and run
// its loop control (the single-simulation reduce_kernel in advance mode). Workgroup 0 also
// clears the class's work-list counters that this step's stage kernels start from.
__global__ void __launch_bounds__(256) batch_reduce_kernel(const BatchJob *jobs, int *counters, int cls) {
  if (blockIdx.x == 0 && threadIdx.x < 2) counters[2 * cls + threadIdx.x] = 0;
  const BatchJob &J = jobs[blockIdx.x];
  if (!J.used || J.cls != cls) return;
  StepCtl *ctl = J.ctl;
  if (!ctl->active) return;  // uniform over the workgroup
  double m = 0.0;
  for (int i = threadIdx.x; i < J.nblocks; i += 256) m = fmax(m, J.partials[i]);
  __shared__ double red[4];
  m = wave_max(m);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x != 0) return;
  const double b = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
  ctl->inv_hyp = b;
  if (ctl->nsteps > 0) ctl->fresh = 0;  // the first step reads Q0, later ones reset(c2p(U))
  advance_control(ctl, b, J.kp, nullptr);
}

// Work list: for each class c, the items (slot, segment) of every used, running job of
// class c, in slot order, at items[ranges[2c] .. ranges[2c] + ranges[2c+1]).
__global__ void __launch_bounds__(1024) pack_items_kernel(const BatchJob *jobs, int nslots, int nclasses,
                                                          BatchItem *items, int *ranges) {
  __shared__ int cnt[MAX_SLOTS];
  const int s = threadIdx.x;
  int base = 0;
  for (int c = 0; c < MAX_CLASSES; ++c) {
    int n = 0;
    if (c < nclasses && s < nslots) {
      const BatchJob &J = jobs[s];
      if (J.used && J.cls == c && J.ctl->active) n = J.nblocks;
    }
    cnt[s] = n;
    __syncthreads();
    for (int off = 1; off < MAX_SLOTS; off <<= 1) {  // inclusive scan
      const int v = (s >= off) ? cnt[s - off] : 0;
      __syncthreads();
      cnt[s] += v;
      __syncthreads();
    }
    const int total = cnt[MAX_SLOTS - 1];
    const int start = base + cnt[s] - n;
    for (int b = 0; b < n; ++b) {
      BatchItem it;
      it.job = s;
      it.blk = b;
      items[start + b] = it;
    }
    if (s == 0) {
      ranges[2 * c] = base;
      ranges[2 * c + 1] = total;
    }
    base += total;
    __syncthreads();
  }
}

__global__ void put_job_kernel(BatchJob *jobs, int slot, BatchJob j) { jobs[slot] = j; }

__global__ void drop_job_kernel(BatchJob *jobs, int slot) {
  jobs[slot].used = 0;
  jobs[slot].ctl->active = 0;
}

// A new job's control block: t = 0, loop to (tend, max_steps), first step from Q0.
__global__ void init_job_ctl_kernel(StepCtl *ctl, double tend, int max_steps) {
  ctl->dt = ctl->dt_dx = ctl->dt_dy = 0.0;
  ctl->t = 0.0;
  ctl->tend = tend;
  ctl->inv_hyp = 0.0;
  ctl->active = 1;
  ctl->nsteps = 0;
  ctl->max_steps = max_steps;
  ctl->fresh = 1;
  ctl->neg_rho = ctl->neg_prs = ctl->nan_cnt = 0ull;
}

// ------------------------------------------------------------------ host side

struct JobHost {
  int64_t id = 0;
  int status = FV2D_JOB_FREE;
  fv2d_params p;
  KParams kp;
  StripGeom g;
  int cls = 0;
  size_t off = 0, bytes = 0;
  void *U = nullptr, *Us = nullptr, *Q0 = nullptr;
  double *partials = nullptr;
  float *grav_row = nullptr;
  uint64_t first_launch = 0;  // the first launch whose status read-back covers this job
};

// Arena layout of one job (offsets from the start of its range).
struct JobLayout {
  size_t state, partials, grav, total;
};

size_t align_up(size_t x) { return (x + ARENA_ALIGN - 1) / ARENA_ALIGN * ARENA_ALIGN; }

// Rows per chunk of a batched job: its own chunk_rows, else 4.
int batch_chunk_rows(const fv2d_params *p) { return p->chunk_rows > 0 ? p->chunk_rows : CHUNK_ROWS_DEFAULT; }

int batch_seg_rows(const fv2d_params *p) {
  if (p->seg_rows > 0) return p->seg_rows;
  // jobs share the chip: full-height segments (4-row prologue per 64 rows) unless the grid
  // is shorter
  const int nw = batch_chunk_rows(p);
  const int s = (p->Ny + nw - 1) / nw * nw;
  return s < 64 ? s : 64;
}

StripGeom batch_geom(const fv2d_params *p) {
  return strip_geom(p->Nx, p->Ny, batch_seg_rows(p), batch_chunk_rows(p));
}

JobLayout job_layout(const fv2d_params *p) {
  const size_t esize = (p->precision == FV2D_FP32) ? 4 : 8;
  const size_t pitch = (size_t)(p->Nx + 63) / 64 * 64;
  const StripGeom g = batch_geom(p);
  JobLayout L;
  L.state = align_up(4 * pitch * (size_t)p->Ny * esize);
  L.partials = align_up(sizeof(double) * (size_t)g.nblocks);
  L.grav = (p->gravity_mode == FV2D_GRAV_ANALYTICAL) ? align_up(sizeof(float) * (size_t)p->Ny) : 0;
  L.total = 3 * L.state + L.partials + L.grav;
  return L;
}

StageClass class_of(const fv2d_params &p, const KParams &kp) {
  StageClass c;
  c.recon = (p.reconstruction == FV2D_PLM) ? PLM : PCM;  // PCM_WB falls through to PCM (Update.h:27-33)
  c.riem = (p.riemann_solver == FV2D_HLL) ? HLL : (p.riemann_solver == FV2D_FSLP) ? FSLP : HLLC;
  c.tc = p.tc_active ? 1 : 0;
  c.visc = p.visc_active ? 1 : 0;
  c.xg = needs_xg(kp) ? 1 : 0;
  c.nw = batch_chunk_rows(&p);
  return c;
}

}  // namespace

struct fv2d_batch {
  fv2d_batch_config cfg;
  int device = 0;
  int ncu = 0;
  hipStream_t stream = nullptr;  // job init, work list, launches
  hipStream_t io = nullptr;      // result read-back
  char *arena = nullptr;
  std::map<size_t, size_t> free_ranges;  // offset -> size, coalesced
  BatchJob *d_jobs = nullptr;
  StepCtl *d_ctl = nullptr;
  BatchItem *d_items = nullptr;
  int *d_ranges = nullptr;
  int *d_counters = nullptr;
  StepCtl *h_ctl = nullptr;      // pinned: status read-back target of each launch
  std::vector<StepCtl> ctl_snap;  // h_ctl as of the last wait
  double *d_upload = nullptr;  // host primitives / gravity rows of a submit
  size_t upload_cap = 0;
  double *d_out = nullptr;  // primitives of a collect
  size_t out_cap = 0;
  double *d_rowsum = nullptr, *d_sums = nullptr;
  int rowsum_cap = 0;
  std::vector<JobHost> slots;
  std::map<int64_t, int> by_id;
  std::deque<int64_t> done;
  std::vector<StageClass> classes;
  std::vector<int> grid;                // persistent workgroups per class
  std::vector<hipStream_t> cls_stream;  // per class: its chain of steps
  std::vector<hipEvent_t> cls_join;     // per class: end of its chain in a launch
  hipEvent_t ev_fork = nullptr;
  int total_items = 0;
  int64_t next_id = 1;
  bool dirty = true;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  int graph_classes = -1;
  hipEvent_t ev_start = nullptr, ev_end = nullptr, ev_ctl = nullptr, ev_upload = nullptr;
  uint64_t launches_queued = 0;   // launch calls so far
  uint64_t launches_waited = 0;   // covered by the last wait
  bool pending = false;
  double last_ms = 0.0;
  std::string err;
};

static thread_local std::string g_batch_err;

#define BFAIL(b, code, ...)                                                                                            \
  do {                                                                                                                 \
    char _m[512];                                                                                                      \
    snprintf(_m, sizeof(_m), __VA_ARGS__);                                                                             \
    (b)->err = _m;                                                                                                     \
    return (code);                                                                                                     \
  } while (0)

#define BHIP(b, call)                                                                                                  \
  do {                                                                                                                 \
    hipError_t _e = (call);                                                                                            \
    if (_e != hipSuccess) BFAIL(b, FV2D_E_HIP, "%s: %s (%s:%d)", #call, hipGetErrorString(_e), __FILE__, __LINE__); \
  } while (0)

static bool arena_alloc(fv2d_batch *b, size_t bytes, size_t &off) {
  for (auto it = b->free_ranges.begin(); it != b->free_ranges.end(); ++it) {
    if (it->second >= bytes) {
      off = it->first;
      const size_t rest = it->second - bytes;
      b->free_ranges.erase(it);
      if (rest) b->free_ranges[off + bytes] = rest;
      return true;
    }
  }
  return false;
}

static void arena_free(fv2d_batch *b, size_t off, size_t bytes) {
  auto it = b->free_ranges.emplace(off, bytes).first;
  auto nx = std::next(it);
  if (nx != b->free_ranges.end() && it->first + it->second == nx->first) {
    it->second += nx->second;
    b->free_ranges.erase(nx);
  }
  if (it != b->free_ranges.begin()) {
    auto pv = std::prev(it);
    if (pv->first + pv->second == it->first) {
      pv->second += it->second;
      b->free_ranges.erase(it);
    }
  }
}

static int class_index(fv2d_batch *b, const StageClass &c) {
  for (size_t i = 0; i < b->classes.size(); ++i)
    if (b->classes[i] == c) return (int)i;
  if ((int)b->classes.size() >= MAX_CLASSES) return -1;
  hipStream_t st = nullptr;
  hipEvent_t ev = nullptr;
  if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) return -1;
  if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
    (void)hipStreamDestroy(st);
    return -1;
  }
  const int occ = (b->cfg.precision == FV2D_FP32) ? batch_occupancy_fp32(c) : batch_occupancy_fp64(c);
  int per_cu = occ;
  if (b->cfg.blocks_per_cu > 0 && b->cfg.blocks_per_cu < occ) per_cu = b->cfg.blocks_per_cu;
  b->classes.push_back(c);
  b->grid.push_back(per_cu * b->ncu);
  b->cls_stream.push_back(st);
  b->cls_join.push_back(ev);
  return (int)b->classes.size() - 1;
}

// One step of the jobs of class c on stream st: reduce + loop control, stage 1, stage 2.
static hipError_t enqueue_class_step(fv2d_batch *b, int c, hipStream_t st) {
  hipLaunchKernelGGL(batch_reduce_kernel, dim3(b->cfg.max_jobs), dim3(256), 0, st, b->d_jobs, b->d_counters, c);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const bool f32 = b->cfg.precision == FV2D_FP32;
  for (int stage = 1; stage <= 2; ++stage) {
    int *counter = b->d_counters + 2 * c + (stage - 1);
    const int *range = b->d_ranges + 2 * c;
    e = f32 ? launch_batch_fp32(b->classes[c], b->d_jobs, b->d_items, range, counter, stage, b->grid[c], st)
            : launch_batch_fp64(b->classes[c], b->d_jobs, b->d_items, range, counter, stage, b->grid[c], st);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

// One launch: the class chains forked from the batch stream and joined back into it.
static hipError_t enqueue_launch(fv2d_batch *b) {
  hipError_t e = hipEventRecord(b->ev_fork, b->stream);
  for (size_t c = 0; c < b->classes.size() && e == hipSuccess; ++c) {
    hipStream_t st = b->cls_stream[c];
    e = hipStreamWaitEvent(st, b->ev_fork, 0);
    for (int k = 0; k < b->cfg.steps_per_launch && e == hipSuccess; ++k) e = enqueue_class_step(b, (int)c, st);
    if (e == hipSuccess) e = hipEventRecord(b->cls_join[c], st);
    if (e == hipSuccess) e = hipStreamWaitEvent(b->stream, b->cls_join[c], 0);
  }
  return e;
}

static int ensure_graph(fv2d_batch *b) {
  if (b->gexec && b->graph_classes == (int)b->classes.size()) return FV2D_OK;
  if (b->gexec) {
    BHIP(b, hipGraphExecDestroy(b->gexec));
    b->gexec = nullptr;
  }
  if (b->graph) {
    BHIP(b, hipGraphDestroy(b->graph));
    b->graph = nullptr;
  }
  BHIP(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  const hipError_t e = enqueue_launch(b);
  hipGraph_t g = nullptr;
  const hipError_t e2 = hipStreamEndCapture(b->stream, &g);
  if (e != hipSuccess) BFAIL(b, FV2D_E_HIP, "graph capture: %s", hipGetErrorString(e));
  if (e2 != hipSuccess) BFAIL(b, FV2D_E_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e2));
  b->graph = g;
  BHIP(b, hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0));
  b->graph_classes = (int)b->classes.size();
  return FV2D_OK;
}

static int grow(fv2d_batch *b, double **buf, size_t &cap, size_t need) {
  if (need <= cap) return FV2D_OK;
  // the buffer may be read by work queued on either stream
  BHIP(b, hipStreamSynchronize(b->stream));
  BHIP(b, hipStreamSynchronize(b->io));
  if (*buf) BHIP(b, hipFree(*buf));
  *buf = nullptr;
  cap = 0;
  BHIP(b, hipMalloc((void **)buf, need));
  cap = need;
  return FV2D_OK;
}

extern "C" {

void fv2d_batch_config_default(fv2d_batch_config *c) {
  if (!c) return;
  memset(c, 0, sizeof *c);
  c->precision = FV2D_FP32;
  c->max_jobs = 256;
  c->arena_bytes = (int64_t)8 << 30;
  c->steps_per_launch = 8;
  c->blocks_per_cu = 0;
  c->max_items = 1 << 18;
  c->use_graph = 1;
}

int64_t fv2d_job_bytes(const fv2d_params *p) {
  if (!p) return -1;
  std::string why;
  if (check_params(p, why)) return -1;
  return (int64_t)job_layout(p).total;
}

const char *fv2d_batch_last_error(const fv2d_batch *b) { return b ? b->err.c_str() : g_batch_err.c_str(); }

void fv2d_batch_destroy(fv2d_batch *b) {
  if (!b) return;
  int cur = 0;
  (void)hipGetDevice(&cur);
  (void)hipSetDevice(b->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  if (b->io) (void)hipStreamSynchronize(b->io);
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  for (hipStream_t st : b->cls_stream) (void)hipStreamSynchronize(st);
  hipEvent_t evs[] = {b->ev_start, b->ev_end, b->ev_ctl, b->ev_upload, b->ev_fork};
  for (hipEvent_t e : evs)
    if (e) (void)hipEventDestroy(e);
  for (hipEvent_t e : b->cls_join) (void)hipEventDestroy(e);
  for (hipStream_t st : b->cls_stream) (void)hipStreamDestroy(st);
  void *bufs[] = {b->arena, b->d_jobs, b->d_ctl, b->d_items, b->d_ranges, b->d_counters,
                  b->d_upload, b->d_out, b->d_rowsum, b->d_sums};
  for (void *x : bufs)
    if (x) (void)hipFree(x);
  if (b->h_ctl) (void)hipHostFree(b->h_ctl);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  if (b->io) (void)hipStreamDestroy(b->io);
  (void)hipSetDevice(cur);
  delete b;
}

int fv2d_batch_create(const fv2d_batch_config *c, fv2d_batch **out) {
  if (!out) {
    g_batch_err = "fv2d_batch_create: out is NULL";
    return FV2D_E_INVALID;
  }
  *out = nullptr;
  if (!c) {
    g_batch_err = "fv2d_batch_create: config is NULL";
    return FV2D_E_INVALID;
  }
  if (c->precision != FV2D_FP32 && c->precision != FV2D_FP64) {
    g_batch_err = "fv2d_batch_create: precision must be FV2D_FP64 or FV2D_FP32";
    return FV2D_E_INVALID;
  }
  if (c->max_jobs < 1 || c->max_jobs > MAX_SLOTS) {
    g_batch_err = "fv2d_batch_create: max_jobs must be in [1, 1024]";
    return FV2D_E_INVALID;
  }
  if (c->arena_bytes < (int64_t)ARENA_ALIGN || c->steps_per_launch < 1 || c->steps_per_launch > 1024 ||
      c->blocks_per_cu < 0 || c->max_items < 1) {
    g_batch_err = "fv2d_batch_create: arena_bytes >= 256, steps_per_launch in [1, 1024], blocks_per_cu >= 0, "
                  "max_items >= 1";
    return FV2D_E_INVALID;
  }
  fv2d_batch *b = new fv2d_batch();
  b->cfg = *c;
  b->cfg.arena_bytes = (int64_t)(((size_t)c->arena_bytes) / ARENA_ALIGN * ARENA_ALIGN);
  b->slots.resize(c->max_jobs);
  b->ctl_snap.assign(c->max_jobs, StepCtl());
  auto fail = [&](const char *what, hipError_t e) {
    char m[512];
    snprintf(m, sizeof m, "fv2d_batch_create: %s: %s", what, hipGetErrorString(e));
    g_batch_err = m;
    fv2d_batch_destroy(b);
    return (e == hipErrorOutOfMemory) ? FV2D_E_NOMEM : FV2D_E_HIP;
  };
  hipError_t e;
  if ((e = hipGetDevice(&b->device)) != hipSuccess) return fail("hipGetDevice", e);
  if ((e = hipDeviceGetAttribute(&b->ncu, hipDeviceAttributeMultiprocessorCount, b->device)) != hipSuccess)
    return fail("CU count", e);
  if ((e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking)) != hipSuccess) return fail("stream", e);
  if ((e = hipStreamCreateWithFlags(&b->io, hipStreamNonBlocking)) != hipSuccess) return fail("io stream", e);
  if ((e = hipMalloc((void **)&b->arena, (size_t)b->cfg.arena_bytes)) != hipSuccess) return fail("arena", e);
  b->free_ranges[0] = (size_t)b->cfg.arena_bytes;
  if ((e = hipMalloc((void **)&b->d_jobs, sizeof(BatchJob) * c->max_jobs)) != hipSuccess) return fail("jobs", e);
  if ((e = hipMalloc((void **)&b->d_ctl, sizeof(StepCtl) * c->max_jobs)) != hipSuccess) return fail("ctl", e);
  if ((e = hipMalloc((void **)&b->d_items, sizeof(BatchItem) * (size_t)c->max_items)) != hipSuccess)
    return fail("work list", e);
  if ((e = hipMalloc((void **)&b->d_ranges, sizeof(int) * 2 * MAX_CLASSES)) != hipSuccess) return fail("ranges", e);
  if ((e = hipMalloc((void **)&b->d_counters, sizeof(int) * 2 * MAX_CLASSES)) != hipSuccess)
    return fail("counters", e);
  if ((e = hipMalloc((void **)&b->d_sums, sizeof(double) * 4)) != hipSuccess) return fail("sums", e);
  if ((e = hipHostMalloc((void **)&b->h_ctl, sizeof(StepCtl) * c->max_jobs, hipHostMallocDefault)) != hipSuccess)
    return fail("pinned status", e);
  memset(b->h_ctl, 0, sizeof(StepCtl) * c->max_jobs);
  // plain memsets: nothing here is captured
  if ((e = hipMemsetAsync(b->d_jobs, 0, sizeof(BatchJob) * c->max_jobs, b->stream)) != hipSuccess)
    return fail("clear jobs", e);
  if ((e = hipMemsetAsync(b->d_ctl, 0, sizeof(StepCtl) * c->max_jobs, b->stream)) != hipSuccess)
    return fail("clear ctl", e);
  if ((e = hipMemsetAsync(b->d_ranges, 0, sizeof(int) * 2 * MAX_CLASSES, b->stream)) != hipSuccess)
    return fail("clear ranges", e);
  if ((e = hipMemsetAsync(b->d_counters, 0, sizeof(int) * 2 * MAX_CLASSES, b->stream)) != hipSuccess)
    return fail("clear counters", e);
  hipEvent_t *evs[] = {&b->ev_start, &b->ev_end, &b->ev_ctl, &b->ev_upload};
  for (hipEvent_t *ev : evs)
    if ((e = hipEventCreate(ev)) != hipSuccess) return fail("event", e);
  if ((e = hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming)) != hipSuccess) return fail("event", e);
  if ((e = hipStreamSynchronize(b->stream)) != hipSuccess) return fail("init", e);
  *out = b;
  return FV2D_OK;
}

int fv2d_batch_submit(fv2d_batch *b, const fv2d_params *p, const fv2d_problem *prob, const double *q, double tend,
                      int32_t max_steps, int64_t *job_id) {
  if (!b) return FV2D_E_INVALID;
  if (!p || !job_id) BFAIL(b, FV2D_E_INVALID, "fv2d_batch_submit: p and job_id must not be NULL");
  if ((prob == nullptr) == (q == nullptr))
    BFAIL(b, FV2D_E_INVALID, "fv2d_batch_submit: give exactly one of prob and q");
  std::string why;
  int rc = check_params(p, why);
  if (rc) BFAIL(b, rc, "fv2d_batch_submit: %s", why.c_str());
  if (p->precision != b->cfg.precision)
    BFAIL(b, FV2D_E_INVALID, "fv2d_batch_submit: job precision %d differs from the batch's %d", p->precision,
          b->cfg.precision);
  if (!(tend == tend)) BFAIL(b, FV2D_E_INVALID, "fv2d_batch_submit: tend is NaN");
  if (max_steps <= 0 && !(tend < FV2D_TEND_NONE))
    BFAIL(b, FV2D_E_INVALID, "fv2d_batch_submit: no end: give tend < 1e300 or max_steps > 0");
  JobHost J;
  J.p = *p;
  const int pitch = (p->Nx + 63) / 64 * 64;
  build_kparams(p, pitch, J.kp);
  J.g = batch_geom(p);
  ProblemArgs pa;
  if (prob && problem_args(*p, J.kp, *prob, pa))
    BFAIL(b, FV2D_E_INVALID, "fv2d_batch_submit: unknown problem kind %d or bad amplitude", prob->kind);
  const int cls = class_index(b, class_of(*p, J.kp));
  if (cls < 0)
    BFAIL(b, FV2D_E_BUSY, "fv2d_batch_submit: no room for another kernel class (%d at most) or its stream",
          MAX_CLASSES);
  int slot = -1;
  for (int s = 0; s < b->cfg.max_jobs; ++s)
    if (b->slots[s].status == FV2D_JOB_FREE) {
      slot = s;
      break;
    }
  if (slot < 0) BFAIL(b, FV2D_E_BUSY, "fv2d_batch_submit: all %d slots are taken", b->cfg.max_jobs);
  if (b->total_items + J.g.nblocks > b->cfg.max_items)
    BFAIL(b, FV2D_E_BUSY, "fv2d_batch_submit: work list full (%d + %d items > %d)", b->total_items, J.g.nblocks,
          b->cfg.max_items);
  const JobLayout L = job_layout(p);
  if (!arena_alloc(b, L.total, J.off))
    BFAIL(b, FV2D_E_BUSY, "fv2d_batch_submit: no free arena range of %zu bytes", L.total);
  BHIP(b, hipSetDevice(b->device));
  J.bytes = L.total;
  J.cls = cls;
  char *base = b->arena + J.off;
  J.U = base;
  J.Us = base + L.state;
  J.Q0 = base + 2 * L.state;
  J.partials = (double *)(base + 3 * L.state);
  J.grav_row = L.grav ? (float *)(base + 3 * L.state + L.partials) : nullptr;
  StepCtl *ctl = b->d_ctl + slot;
  const long long n = (long long)p->Nx * p->Ny;
  const bool f32 = b->cfg.precision == FV2D_FP32;

  // Buffers sized by the largest job so far: grown before any work of this job is queued
  // (growing waits for queued work). Host primitives go through d_upload.
  auto undo = [&]() { arena_free(b, J.off, J.bytes); };
  if ((q && (rc = grow(b, &b->d_upload, b->upload_cap, sizeof(double) * 4 * (size_t)n))) ||
      (rc = grow(b, &b->d_out, b->out_cap, sizeof(double) * 4 * (size_t)n))) {
    undo();
    return rc;
  }
  if (p->Ny > b->rowsum_cap) {
    size_t cap = sizeof(double) * 4 * (size_t)b->rowsum_cap;
    if ((rc = grow(b, &b->d_rowsum, cap, sizeof(double) * 4 * (size_t)p->Ny))) {
      undo();
      return rc;
    }
    b->rowsum_cap = p->Ny;
  }
  if (L.grav) {
    // getAnalyticalGravity (Gravity.h:15-30), through float, as the single-simulation path
    std::vector<float> tab(p->Ny);
    for (int j = 0; j < p->Ny; ++j) {
      const double y = p->ymin + (j + 0.5) * J.kp.dy;
      tab[j] = (float)(p->hot_bubble_g0 * sin(y * M_PI * 2.0 / p->ymax));
    }
    BHIP(b, hipMemcpyAsync(J.grav_row, tab.data(), sizeof(float) * p->Ny, hipMemcpyHostToDevice, b->stream));
    BHIP(b, hipStreamSynchronize(b->stream));
  }
  if (q) {
    BHIP(b, hipMemcpyAsync(b->d_upload, q, sizeof(double) * 4 * n, hipMemcpyHostToDevice, b->stream));
    if (f32)
      hipLaunchKernelGGL(init_state_kernel<float>, dim3(grid_1d(n)), dim3(AUX_THREADS), 0, b->stream, J.kp,
                         b->d_upload, (float *)J.U, (float *)J.Q0);
    else
      hipLaunchKernelGGL(init_state_kernel<double>, dim3(grid_1d(n)), dim3(AUX_THREADS), 0, b->stream, J.kp,
                         b->d_upload, (double *)J.U, (double *)J.Q0);
  } else {
    if (f32)
      hipLaunchKernelGGL(problem_ic_kernel<float>, dim3(grid_1d(n)), dim3(AUX_THREADS), 0, b->stream, J.kp, pa,
                         (float *)J.U, (float *)J.Q0);
    else
      hipLaunchKernelGGL(problem_ic_kernel<double>, dim3(grid_1d(n)), dim3(AUX_THREADS), 0, b->stream, J.kp, pa,
                         (double *)J.U, (double *)J.Q0);
  }
  BHIP(b, hipGetLastError());
  hipLaunchKernelGGL(init_job_ctl_kernel, dim3(1), dim3(1), 0, b->stream, ctl, tend, max_steps);
  if (f32)
    hipLaunchKernelGGL((dt_cells_kernel<float, float>), dim3(J.g.nblocks), dim3(256), 0, b->stream, J.kp,
                       (const float *)J.Q0, 1, J.partials, J.g.nblocks);
  else
    hipLaunchKernelGGL((dt_cells_kernel<double, double>), dim3(J.g.nblocks), dim3(256), 0, b->stream, J.kp,
                       (const double *)J.Q0, 1, J.partials, J.g.nblocks);
  BatchJob dj;
  memset(&dj, 0, sizeof dj);
  dj.kp = J.kp;
  dj.nblocks = J.g.nblocks;
  dj.seg = J.g.seg;
  dj.nsx = J.g.nsx;
  dj.nsy = J.g.nsy;
  dj.U = J.U;
  dj.Us = J.Us;
  dj.Q0 = J.Q0;
  dj.grav_row = J.grav_row;
  dj.ctl = ctl;
  dj.partials = J.partials;
  dj.rk2 = (p->time_stepping == FV2D_TS_RK2) ? 1 : 0;
  dj.used = 1;
  dj.cls = cls;
  hipLaunchKernelGGL(put_job_kernel, dim3(1), dim3(1), 0, b->stream, b->d_jobs, slot, dj);
  BHIP(b, hipGetLastError());
  if (q) {
    BHIP(b, hipEventRecord(b->ev_upload, b->stream));
    BHIP(b, hipEventSynchronize(b->ev_upload));
  }
  J.id = b->next_id++;
  J.status = FV2D_JOB_RUNNING;
  J.first_launch = b->launches_queued + 1;
  b->slots[slot] = J;
  b->by_id[J.id] = slot;
  b->total_items += J.g.nblocks;
  b->dirty = true;
  *job_id = J.id;
  return FV2D_OK;
}

int fv2d_batch_launch(fv2d_batch *b, int32_t launches) {
  if (!b) return FV2D_E_INVALID;
  if (launches < 1) BFAIL(b, FV2D_E_INVALID, "fv2d_batch_launch: launches must be >= 1");
  BHIP(b, hipSetDevice(b->device));
  if (b->dirty) {
    hipLaunchKernelGGL(pack_items_kernel, dim3(1), dim3(MAX_SLOTS), 0, b->stream, b->d_jobs, b->cfg.max_jobs,
                       (int)b->classes.size(), b->d_items, b->d_ranges);
    BHIP(b, hipGetLastError());
    b->dirty = false;
  }
  int rc;
  if (b->cfg.use_graph && (rc = ensure_graph(b))) return rc;
  BHIP(b, hipEventRecord(b->ev_start, b->stream));
  for (int k = 0; k < launches; ++k) {
    if (b->cfg.use_graph) {
      BHIP(b, hipGraphLaunch(b->gexec, b->stream));
    } else {
      BHIP(b, enqueue_launch(b));
    }
  }
  BHIP(b, hipEventRecord(b->ev_end, b->stream));
  BHIP(b, hipMemcpyAsync(b->h_ctl, b->d_ctl, sizeof(StepCtl) * b->cfg.max_jobs, hipMemcpyDeviceToHost, b->stream));
  BHIP(b, hipEventRecord(b->ev_ctl, b->stream));
  b->launches_queued += 1;
  b->pending = true;
  return FV2D_OK;
}

int fv2d_batch_wait(fv2d_batch *b, int32_t *running) {
  if (!b) return FV2D_E_INVALID;
  BHIP(b, hipSetDevice(b->device));
  if (b->pending) {
    BHIP(b, hipEventSynchronize(b->ev_ctl));
    float ms = 0.0f;
    BHIP(b, hipEventElapsedTime(&ms, b->ev_start, b->ev_end));
    b->last_ms = ms;
    b->pending = false;
    b->launches_waited = b->launches_queued;
    memcpy(b->ctl_snap.data(), b->h_ctl, sizeof(StepCtl) * b->cfg.max_jobs);
    for (int s = 0; s < b->cfg.max_jobs; ++s) {
      JobHost &J = b->slots[s];
      if (J.status != FV2D_JOB_RUNNING || J.first_launch > b->launches_waited) continue;
      if (!b->ctl_snap[s].active) {
        J.status = FV2D_JOB_DONE;
        b->done.push_back(J.id);
      }
    }
  }
  if (running) {
    int r = 0;
    for (const JobHost &J : b->slots) r += (J.status == FV2D_JOB_RUNNING) ? 1 : 0;
    *running = r;
  }
  return FV2D_OK;
}

int fv2d_batch_poll(fv2d_batch *b, int64_t *ids, int32_t cap, int32_t *n) {
  if (!b) return FV2D_E_INVALID;
  if (!n || (cap > 0 && !ids)) BFAIL(b, FV2D_E_INVALID, "fv2d_batch_poll: bad arguments");
  int k = 0;
  for (int64_t id : b->done) {
    if (k >= cap) break;
    ids[k++] = id;
  }
  *n = k;
  return FV2D_OK;
}

static void fill_info(const fv2d_batch *b, int slot, fv2d_job_info *info) {
  const JobHost &J = b->slots[slot];
  memset(info, 0, sizeof *info);
  info->id = J.id;
  info->status = J.status;
  if (J.first_launch <= b->launches_waited) {
    const StepCtl &c = b->ctl_snap[slot];
    info->nsteps = c.nsteps;
    info->t = c.t;
    info->dt_last = c.dt;
    info->negative_density = c.neg_rho;
    info->negative_pressure = c.neg_prs;
    info->nan_count = c.nan_cnt;
  }
}

int fv2d_batch_job_info(fv2d_batch *b, int64_t id, fv2d_job_info *info) {
  if (!b) return FV2D_E_INVALID;
  if (!info) BFAIL(b, FV2D_E_INVALID, "fv2d_batch_job_info: info is NULL");
  auto it = b->by_id.find(id);
  if (it == b->by_id.end()) BFAIL(b, FV2D_E_INVALID, "fv2d_batch_job_info: no job %lld", (long long)id);
  fill_info(b, it->second, info);
  return FV2D_OK;
}

static void release(fv2d_batch *b, int slot) {
  JobHost &J = b->slots[slot];
  arena_free(b, J.off, J.bytes);
  b->total_items -= J.g.nblocks;
  b->by_id.erase(J.id);
  for (auto it = b->done.begin(); it != b->done.end(); ++it)
    if (*it == J.id) {
      b->done.erase(it);
      break;
    }
  J = JobHost();
  b->dirty = true;
}

int fv2d_batch_collect(fv2d_batch *b, int64_t id, double *q, double sums[4], fv2d_job_info *info) {
  if (!b) return FV2D_E_INVALID;
  auto it = b->by_id.find(id);
  if (it == b->by_id.end()) BFAIL(b, FV2D_E_INVALID, "fv2d_batch_collect: no job %lld", (long long)id);
  const int slot = it->second;
  JobHost &J = b->slots[slot];
  if (J.status != FV2D_JOB_DONE) BFAIL(b, FV2D_E_STATE, "fv2d_batch_collect: job %lld has not finished", (long long)id);
  BHIP(b, hipSetDevice(b->device));
  const StepCtl &c = b->ctl_snap[slot];
  const bool fresh = c.nsteps == 0;
  // Euler ping-pongs U and Us: after an odd number of steps the state is in Us
  const bool in_us = !fresh && J.p.time_stepping == FV2D_TS_EULER && (c.nsteps & 1);
  const void *src = fresh ? J.Q0 : (in_us ? J.Us : J.U);
  const long long n = (long long)J.p.Nx * J.p.Ny;
  const bool f32 = b->cfg.precision == FV2D_FP32;
  if (q) {
    if (f32)
      hipLaunchKernelGGL(get_prim_kernel<float>, dim3(grid_1d(n)), dim3(AUX_THREADS), 0, b->io, J.kp,
                         (const float *)src, fresh ? 1 : 0, b->d_out);
    else
      hipLaunchKernelGGL(get_prim_kernel<double>, dim3(grid_1d(n)), dim3(AUX_THREADS), 0, b->io, J.kp,
                         (const double *)src, fresh ? 1 : 0, b->d_out);
    BHIP(b, hipGetLastError());
    BHIP(b, hipMemcpyAsync(q, b->d_out, sizeof(double) * 4 * n, hipMemcpyDeviceToHost, b->io));
  }
  if (sums) {
    const void *u = in_us ? J.Us : J.U;
    if (f32)
      hipLaunchKernelGGL(row_sum_kernel<float>, dim3(J.p.Ny, 4), dim3(256), 0, b->io, J.kp, (const float *)u,
                         b->d_rowsum);
    else
      hipLaunchKernelGGL(row_sum_kernel<double>, dim3(J.p.Ny, 4), dim3(256), 0, b->io, J.kp, (const double *)u,
                         b->d_rowsum);
    hipLaunchKernelGGL(total_sum_kernel, dim3(4), dim3(256), 0, b->io, J.kp, b->d_rowsum, b->d_sums);
    BHIP(b, hipGetLastError());
    BHIP(b, hipMemcpyAsync(sums, b->d_sums, sizeof(double) * 4, hipMemcpyDeviceToHost, b->io));
  }
  BHIP(b, hipStreamSynchronize(b->io));
  if (info) fill_info(b, slot, info);
  // the slot's job entry is cleared in stream order, after any launch still in flight
  hipLaunchKernelGGL(drop_job_kernel, dim3(1), dim3(1), 0, b->stream, b->d_jobs, slot);
  BHIP(b, hipGetLastError());
  release(b, slot);
  return FV2D_OK;
}

int fv2d_batch_cancel(fv2d_batch *b, int64_t id) {
  if (!b) return FV2D_E_INVALID;
  auto it = b->by_id.find(id);
  if (it == b->by_id.end()) BFAIL(b, FV2D_E_INVALID, "fv2d_batch_cancel: no job %lld", (long long)id);
  const int slot = it->second;
  BHIP(b, hipSetDevice(b->device));
  hipLaunchKernelGGL(drop_job_kernel, dim3(1), dim3(1), 0, b->stream, b->d_jobs, slot);
  BHIP(b, hipGetLastError());
  // a running job's memory may be in use by queued launches: wait for them
  BHIP(b, hipStreamSynchronize(b->stream));
  release(b, slot);
  return FV2D_OK;
}

int fv2d_batch_last_launch_ms(fv2d_batch *b, double *ms) {
  if (!b) return FV2D_E_INVALID;
  if (!ms) BFAIL(b, FV2D_E_INVALID, "fv2d_batch_last_launch_ms: ms is NULL");
  *ms = b->last_ms;
  return FV2D_OK;
}

}  // extern "C"
