// emg_api.hip -- the extern "C" entry points of libemgpu (include/emgpu.h).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "emg_internal.h"

namespace emg {

static thread_local std::string g_last_error;

int set_error(int code, const std::string &msg) {
  g_last_error = msg;
  // a failed allocation also sets HIP's last error, which the launch checks
  // of later calls read: the failure is reported here, so clear it
  if (code == EMG_ERR_NOMEM) (void)hipGetLastError();
  return code;
}

int hip_check(hipError_t e, const char *what) {
  if (e == hipSuccess) return EMG_OK;
  (void)hipGetLastError();  // reported now, not again by a later launch check
  return set_error(e == hipErrorOutOfMemory ? EMG_ERR_NOMEM : EMG_ERR_HIP,
                   std::string(what) + ": " + hipGetErrorString(e));
}

// ---- kernel timing: event pairs recorded around launches, accumulated lazily
namespace {
struct EvPair {
  hipEvent_t a, b;
  int kind;
};
struct TimerState {
  std::vector<EvPair> pending;
  std::vector<EvPair> free_pairs;
  EvPair open[kTNum];
  bool is_open[kTNum] = {false, false, false, false, false};
};
}  // namespace

static TimerState *tstate(emg_ctx *c);

static void timer_flush(emg_ctx *c) {
  TimerState *t = tstate(c);
  for (EvPair &p : t->pending) {
    float ms = 0.f;
    if (hipEventSynchronize(p.b) == hipSuccess && hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
      c->timers[p.kind].total_ms += ms;
      c->timers[p.kind].last_ms = ms;
      c->timers[p.kind].launches += 1;
    }
    t->free_pairs.push_back(p);
  }
  t->pending.clear();
}

void timer_begin(emg_ctx *c, int kind) {
  if (!c->timing) return;
  TimerState *t = tstate(c);
  if (t->pending.size() >= 512) timer_flush(c);
  EvPair p;
  if (!t->free_pairs.empty()) {
    p = t->free_pairs.back();
    t->free_pairs.pop_back();
  } else {
    if (hipEventCreate(&p.a) != hipSuccess || hipEventCreate(&p.b) != hipSuccess) return;
  }
  p.kind = kind;
  (void)hipEventRecord(p.a, c->stream);
  t->open[kind] = p;
  t->is_open[kind] = true;
}

void timer_end(emg_ctx *c, int kind) {
  if (!c->timing) return;
  TimerState *t = tstate(c);
  if (!t->is_open[kind]) return;
  EvPair p = t->open[kind];
  (void)hipEventRecord(p.b, c->stream);
  t->pending.push_back(p);
  t->is_open[kind] = false;
}

// declared in the other translation units
int init_device_into(emg_ctx *c, State64 &st, const emg_device_species *spec, int nspec,
                     uint64_t seed);
int64_t energy64_part_need(int nspec);

namespace {

constexpr int kFlagPreside = 8;
constexpr int kFlagPid = 16;

// AoS pbuf records -> SoA FP64 state, validating each record
__global__ __launch_bounds__(kBlock) void k_split(const emg_particle *rec, int64_t n, State64 st,
                                                  int64_t dst0, Grid g, int *flag, int need_pid0) {
  int64_t m = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (m >= n) return;
  emg_particle r = rec[m];
  bool ok = (r.x >= 0.0) && (r.x <= g.slx) && (r.y >= 0.0) && (r.y <= g.sly) && (r.z >= 0.0) &&
            (r.z <= g.slz) && (r.vx == r.vx) && (r.vy == r.vy) && (r.vz == r.vz) &&
            (fabs(r.vx) < 1e300) && (fabs(r.vy) < 1e300) && (fabs(r.vz) < 1e300);
  if (!ok) atomicOr(flag, kFlagDomain);
  if (r.preside != 0) atomicOr(flag, kFlagPreside);
  if (need_pid0 && r.pid != 0) atomicOr(flag, kFlagPid);
  int64_t d = dst0 + m;
  st.x[d] = ok ? r.x : 0.0;
  st.y[d] = ok ? r.y : 0.0;
  st.z[d] = ok ? r.z : 0.0;
  st.vx[d] = ok ? r.vx : 0.0;
  st.vy[d] = ok ? r.vy : 0.0;
  st.vz[d] = ok ? r.vz : 0.0;
  st.pid[d] = r.pid;
}

__global__ __launch_bounds__(kBlock) void k_join(State64 st, int64_t src0, int64_t n, int spec,
                                                 emg_particle *rec) {
  int64_t m = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (m >= n) return;
  int64_t s = src0 + m;
  emg_particle r;
  r.x = st.x[s];
  r.y = st.y[s];
  r.z = st.z[s];
  r.vx = st.vx[s];
  r.vy = st.vy[s];
  r.vz = st.vz[s];
  r.pid = st.pid[s];
  r.preside = 0;
  r.nid = 0;
  r.spec = spec;
  rec[m] = r;
}

template <typename T>
int dalloc(T **p, size_t count, const char *what) {
  *p = nullptr;
  if (count == 0) count = 1;
  hipError_t e = hipMalloc((void **)p, sizeof(T) * count);
  if (e != hipSuccess) {
    *p = nullptr;
    return set_error(EMG_ERR_NOMEM, std::string("device allocation failed: ") + what);
  }
  return EMG_OK;
}

void dfree(void *p) {
  if (p) (void)hipFree(p);
}

int alloc_state64(State64 &st, int64_t cap) {
  int rc;
  if ((rc = dalloc(&st.x, cap, "x")) || (rc = dalloc(&st.y, cap, "y")) ||
      (rc = dalloc(&st.z, cap, "z")) || (rc = dalloc(&st.vx, cap, "vx")) ||
      (rc = dalloc(&st.vy, cap, "vy")) || (rc = dalloc(&st.vz, cap, "vz")) ||
      (rc = dalloc(&st.pid, cap, "pid")))
    return rc;
  return EMG_OK;
}

void free_state64(State64 &st) {
  dfree(st.x);
  dfree(st.y);
  dfree(st.z);
  dfree(st.vx);
  dfree(st.vy);
  dfree(st.vz);
  dfree(st.pid);
  st = State64{};
}

// synchronise and translate device flags into an error code
int check_flags(emg_ctx *c) {
  int rc = hip_check(hipMemcpyAsync(c->hflag, c->dflag, sizeof(int), hipMemcpyDeviceToHost, c->stream),
                     "flag read");
  if (rc) return rc;
  rc = hip_check(hipStreamSynchronize(c->stream), "stream synchronize");
  if (rc) return rc;
  int f = *c->hflag;
  if (f == 0) return EMG_OK;
  launch_zero(c->dflag, sizeof(int), c->stream);
  (void)hipStreamSynchronize(c->stream);
  if (f & kFlagSpillLost)
    return set_error(EMG_ERR_CAPACITY, "re-bucketing overflowed the spill region: particles lost");
  if (f & kFlagDomain)
    return set_error(EMG_ERR_DOMAIN, "particle outside [0, L] or non-finite (clamped)");
  if (f & kFlagFar)
    return set_error(EMG_ERR_DOMAIN, "particle moved more than one box length in one step");
  if (f & kFlagPreside)
    return set_error(EMG_ERR_UNSUPPORTED, "only alive records (preside == 0) are supported");
  if (f & kFlagPid)
    return set_error(EMG_ERR_UNSUPPORTED, "EMG_FP32_FIXED does not carry pid: upload pid = 0");
  return set_error(EMG_ERR_HIP, "unknown device flag");
}

}  // namespace
}  // namespace emg

using namespace emg;

struct emg_timer_holder {
  TimerState t;
};

namespace emg {
static TimerState *tstate(emg_ctx *c) {
  // the TimerState lives right after the context in one allocation
  return &reinterpret_cast<emg_timer_holder *>(c + 1)->t;
}
}  // namespace emg

// Staging of emg_get_rho_noalloc: a file-scope device array instead of a
// per-call allocation.  Every GPU that runs the library holds its own copy
// until the process ends; all calls on one GPU share it.  Host code uses the
// device address hipGetSymbolAddress gives for the current device (taken
// once per device, at the first emg_create on it), never the array's name.
namespace emg {
constexpr int64_t kScratchDoubles = int64_t(1) << 24;  // 128 MiB
static __device__ __align__(256) double scratch[kScratchDoubles];
static std::mutex g_scratch_mu;
static std::vector<double *> g_scratch_dev;  // per device id

// the scratch address on the current device (`dev`)
static int scratch_address(int dev, double **out) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (dev < 0) return set_error(EMG_ERR_ARG, "device id");
  if ((size_t)dev >= g_scratch_dev.size()) g_scratch_dev.resize((size_t)dev + 1, nullptr);
  if (!g_scratch_dev[dev]) {
    void *p = nullptr;
    int rc = hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "scratch address");
    if (rc) return rc;
    g_scratch_dev[dev] = static_cast<double *>(p);
  }
  *out = g_scratch_dev[dev];
  return EMG_OK;
}
}  // namespace emg

#define CTX_CHECK(c)                                                   \
  do {                                                                 \
    if (!(c)) return set_error(EMG_ERR_ARG, "null context");           \
    if (hipSetDevice((c)->device) != hipSuccess)                        \
      return set_error(EMG_ERR_HIP, "hipSetDevice failed");            \
  } while (0)

extern "C" {

int emg_last_error(char *buf, size_t len) {
  if (!buf || len == 0) return EMG_ERR_ARG;
  std::snprintf(buf, len, "%s", g_last_error.c_str());
  return EMG_OK;
}

int emg_create(const emg_config *cfg, emg_ctx **out) {
  if (!cfg || !out) return set_error(EMG_ERR_ARG, "null argument");
  *out = nullptr;
  if (cfg->nspec < 1 || cfg->nspec > EMG_MAX_SPECIES)
    return set_error(EMG_ERR_ARG, "nspec must be 1..EMG_MAX_SPECIES");
  for (int d = 0; d < 3; d++)
    if (cfg->n[d] < 2) return set_error(EMG_ERR_ARG, "need at least 2 cells per axis");
  if (cfg->precision != EMG_PLAIN_FP64 && cfg->precision != EMG_FP32_FIXED)
    return set_error(EMG_ERR_ARG, "unknown precision form");
  if (cfg->sort_every < 0) return set_error(EMG_ERR_ARG, "sort_every must be >= 0");
  if (cfg->sort_every != 0 && cfg->precision != EMG_PLAIN_FP64)
    return set_error(EMG_ERR_ARG, "sort_every applies to EMG_PLAIN_FP64 only");
  for (int s = 0; s < cfg->nspec; s++)
    if (cfg->capacity[s] < 0) return set_error(EMG_ERR_ARG, "negative capacity");

  void *mem = ::operator new(sizeof(emg_ctx) + sizeof(emg_timer_holder), std::nothrow);
  if (!mem) return set_error(EMG_ERR_NOMEM, "host allocation failed");
  std::memset(mem, 0, sizeof(emg_ctx));
  emg_ctx *c = static_cast<emg_ctx *>(mem);
  new (reinterpret_cast<emg_timer_holder *>(c + 1)) emg_timer_holder();
  c->cfg = *cfg;
  int rc = hip_check(hipGetDevice(&c->device), "hipGetDevice");
  if (!rc) rc = scratch_address(c->device, &c->scratch);
  if (rc) {
    emg_destroy(c);
    return rc;
  }
  rc = hip_check(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "stream create");
  if (rc) {
    emg_destroy(c);
    return rc;
  }
  c->own_stream = true;
  c->g.nx = cfg->n[0];
  c->g.ny = cfg->n[1];
  c->g.nz = cfg->n[2];
  c->g.slx = (double)cfg->n[0];
  c->g.sly = (double)cfg->n[1];
  c->g.slz = (double)cfg->n[2];
  for (int b = 0; (1 << b) < cfg->n[0]; b++) c->g.cbx = b + 1;
  for (int b = 0; (1 << b) < cfg->n[1]; b++) c->g.cby = b + 1;
  c->ncell = (int64_t)cfg->n[0] * cfg->n[1] * cfg->n[2];
  c->sp.nspec = cfg->nspec;
  int64_t off = 0;
  for (int s = 0; s < cfg->nspec; s++) {
    c->sp.off[s] = off;
    c->sp.cnt[s] = 0;
    c->sp.q[s] = cfg->q[s];
    c->sp.qm[s] = cfg->qm[s];
    off += cfg->capacity[s];
  }
  c->sp.off[cfg->nspec] = off;
  c->cap_total = off;
  int nsp = cfg->nspec;
  if ((rc = dalloc(&c->dflag, 1, "flag")) || (rc = dalloc(&c->ke_out, 3 * nsp, "energy")) ||
      (rc = dalloc((double **)&c->face, (size_t)(face_len(c->g) * nsp), "upper-face charge"))) {
    emg_destroy(c);
    return rc;
  }
  if (hipHostMalloc((void **)&c->hflag, sizeof(int), hipHostMallocDefault) != hipSuccess) {
    emg_destroy(c);
    return set_error(EMG_ERR_NOMEM, "pinned flag");
  }
  launch_zero(c->dflag, sizeof(int), c->stream);
  if (cfg->precision == EMG_PLAIN_FP64) {
    int64_t nn = (int64_t)(cfg->n[0] + 1) * (cfg->n[1] + 1) * (cfg->n[2] + 1);
    if ((rc = alloc_state64(c->s64, c->cap_total)) || (rc = dalloc(&c->node64, 8 * nn, "nodal E/B")) ||
        (rc = dalloc(&c->rho64, c->ncell * nsp, "rho")) ||
        (rc = dalloc(&c->j64, 3 * c->ncell * nsp, "current")) ||
        (rc = dalloc(&c->ke_part, energy64_part_need(nsp), "energy partials"))) {
      emg_destroy(c);
      return rc;
    }
  } else {
    if ((rc = fx_alloc(c, c->cap_total)) || (rc = dalloc(&c->node32, 8 * c->ncell, "nodal E/B")) ||
        (rc = dalloc(&c->rhofx, c->ncell * nsp, "rho")) ||
        (rc = dalloc(&c->jfx, 3 * c->ncell * nsp, "current"))) {
      emg_destroy(c);
      return rc;
    }
  }
  rc = hip_check(hipStreamSynchronize(c->stream), "create");
  if (rc) {
    emg_destroy(c);
    return rc;
  }
  *out = c;
  return EMG_OK;
}

int emg_destroy(emg_ctx *c) {
  if (!c) return EMG_OK;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  free_state64(c->s64);
  free_state64(c->t64);
  free_state64(c->lam);
  for (int k = 0; k < kStateSlots; k++) free_state64(c->slot[k]);
  dfree(c->adj_e);
  dfree(c->adj_g);
  dfree(c->adj_q);
  dfree(c->adj_sum);
  FxBuckets &B = c->fx;
  for (int k = 0; k < 2; k++) {
    dfree(B.st[k].cell);
    dfree(B.st[k].fx);
    dfree(B.st[k].fy);
    dfree(B.st[k].fz);
    dfree(B.st[k].vx);
    dfree(B.st[k].vy);
    dfree(B.st[k].vz);
    dfree(B.st[k].pid);
    dfree(B.start[k]);
    dfree(B.count[k]);
    dfree(B.spill[k]);
    dfree(B.nst[k]);
    dfree(B.cap[k]);
    dfree(B.ob[k]);
    dfree(B.pre[k]);
  }
  dfree(B.stay);
  dfree(B.extra);
  dfree(B.hist);
  dfree(B.cursor);
  dfree(B.shist);
  dfree(B.scur);
  dfree(B.spre);
  dfree(B.spl);
  dfree(c->eb_dev);
  dfree(c->fe_buf);
  dfree(c->sort_rec);
  es_solver_free(c);
  dfree(c->bk_dev);
  dfree(c->node64);
  dfree(c->node32);
  dfree(c->rho64);
  dfree(c->rhofx);
  dfree(c->j64);
  dfree(c->jfx);
  dfree(c->ke_part);
  dfree(c->ke_out);
  dfree(c->face);
  dfree(c->runs[0]);
  dfree(c->runs[1]);
  dfree(c->run_base);
  dfree(c->sort_count);
  dfree(c->sort_cursor);
  dfree(c->sort_start);
  dfree(c->scan_tmp);
  dfree(c->dflag);
  if (c->hflag) (void)hipHostFree(c->hflag);
  TimerState *t = tstate(c);
  for (EvPair &p : t->pending) t->free_pairs.push_back(p);
  for (EvPair &p : t->free_pairs) {
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
  for (int k = 0; k < kTNum; k++)
    if (t->is_open[k]) {
      (void)hipEventDestroy(t->open[k].a);
      (void)hipEventDestroy(t->open[k].b);
    }
  if (c->stream && c->own_stream) (void)hipStreamDestroy(c->stream);
  reinterpret_cast<emg_timer_holder *>(c + 1)->~emg_timer_holder();
  ::operator delete(c);
  return EMG_OK;
}

int emg_set_stream(emg_ctx *c, void *stream) {
  CTX_CHECK(c);
  int rc = hip_check(hipStreamSynchronize(c->stream), "stream synchronize");
  if (rc) return rc;
  if (c->own_stream && stream) {
    (void)hipStreamDestroy(c->stream);
    c->own_stream = false;
  }
  if (stream) {
    c->stream = (hipStream_t)stream;
  } else if (!c->own_stream) {
    rc = hip_check(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "stream create");
    if (rc) return rc;
    c->own_stream = true;
  }
  return EMG_OK;
}

int emg_synchronize(emg_ctx *c) {
  CTX_CHECK(c);
  int rc = hip_check(hipGetLastError(), "kernel launch");
  if (rc) return rc;
  return check_flags(c);
}

int emg_upload_particles(emg_ctx *c, const emg_particle *pbuf, const int64_t *counts, int32_t nspec) {
  CTX_CHECK(c);
  if (nspec != c->sp.nspec) return set_error(EMG_ERR_ARG, "nspec differs from the context");
  if (!counts) return set_error(EMG_ERR_ARG, "null counts");
  int64_t total = 0;
  for (int s = 0; s < nspec; s++) {
    if (counts[s] < 0) return set_error(EMG_ERR_ARG, "negative count");
    total += counts[s];
  }
  if (total > 0 && !pbuf) return set_error(EMG_ERR_ARG, "null pbuf");
  // transbound: keep nid >= 0 in order (host pass; one copy only when needed)
  std::vector<int64_t> kept(nspec, 0);
  std::vector<emg_particle> compact;
  bool need_compact = false;
  for (int64_t m = 0; m < total; m++)
    if (pbuf[m].nid < 0) {
      need_compact = true;
      break;
    }
  const emg_particle *src = pbuf;
  if (need_compact) {
    compact.reserve((size_t)total);
    int64_t m = 0;
    for (int s = 0; s < nspec; s++) {
      for (int64_t k = 0; k < counts[s]; k++, m++)
        if (pbuf[m].nid >= 0) {
          compact.push_back(pbuf[m]);
          kept[s]++;
        }
    }
    src = compact.data();
  } else {
    for (int s = 0; s < nspec; s++) kept[s] = counts[s];
  }
  for (int s = 0; s < nspec; s++)
    if (kept[s] > c->cfg.capacity[s])
      return set_error(EMG_ERR_CAPACITY, "species count exceeds the context capacity");

  State64 tgt = c->s64;
  State64 tmp{};
  int rc;
  if (c->cfg.precision == EMG_FP32_FIXED) {
    if ((rc = alloc_state64(tmp, std::max<int64_t>(c->cap_total, 1)))) return rc;
    tgt = tmp;
  }
  const int64_t chunk = 1 << 22;
  emg_particle *d_rec = nullptr;
  if ((rc = dalloc(&d_rec, (size_t)std::min<int64_t>(chunk, std::max<int64_t>(total, 1)), "staging"))) {
    free_state64(tmp);
    return rc;
  }
  int64_t m = 0;
  for (int s = 0; s < nspec; s++) {
    for (int64_t k = 0; k < kept[s]; k += chunk) {
      int64_t n = std::min(chunk, kept[s] - k);
      rc = hip_check(hipMemcpy(d_rec, src + m + k, sizeof(emg_particle) * n, hipMemcpyHostToDevice),
                     "upload copy");
      if (rc) break;
      hipLaunchKernelGGL(k_split, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                         c->stream, d_rec, n, tgt, c->sp.off[s] + k, c->g, c->dflag,
                         (c->cfg.precision == EMG_FP32_FIXED && !c->cfg.track_pid) ? 1 : 0);
      rc = hip_check(hipStreamSynchronize(c->stream), "upload split");
      if (rc) break;
    }
    if (rc) break;
    m += kept[s];
    c->sp.cnt[s] = kept[s];
  }
  dfree(d_rec);
  if (!rc) rc = check_flags(c);
  if (!rc && c->cfg.precision == EMG_FP32_FIXED) {
    rc = fx_from64(c, tmp);
    if (!rc) rc = check_flags(c);
  }
  free_state64(tmp);
  if (rc) {
    for (int s = 0; s < nspec; s++) c->sp.cnt[s] = 0;
    c->have_particles = false;
    return rc;
  }
  c->have_particles = true;
  c->pushes_since_sort = c->cfg.sort_every;  // unsorted: sort before the first push
  c->moves_after_sort = emg::kUnsorted;
  c->runs_valid = false;
  c->rho_valid = false;
  c->faces_valid = false;
  c->j_valid = false;
  c->ke_valid = false;
  return EMG_OK;
}

int emg_particle_counts(emg_ctx *c, int64_t *counts) {
  CTX_CHECK(c);
  if (!counts) return set_error(EMG_ERR_ARG, "null counts");
  for (int s = 0; s < c->sp.nspec; s++) counts[s] = c->sp.cnt[s];
  return EMG_OK;
}

int emg_download_particles(emg_ctx *c, emg_particle *pbuf, int64_t *counts) {
  CTX_CHECK(c);
  int rc = check_flags(c);
  if (rc) return rc;
  int64_t total = 0;
  for (int s = 0; s < c->sp.nspec; s++) total += c->sp.cnt[s];
  if (counts)
    for (int s = 0; s < c->sp.nspec; s++) counts[s] = c->sp.cnt[s];
  if (total == 0) return EMG_OK;
  if (!pbuf) return set_error(EMG_ERR_ARG, "null pbuf");
  State64 src = c->s64;
  State64 tmp{};
  if (c->cfg.precision == EMG_FP32_FIXED) {
    if ((rc = alloc_state64(tmp, c->cap_total))) return rc;
    rc = fx_to64(c, tmp);
    if (rc) {
      free_state64(tmp);
      return rc;
    }
    src = tmp;
  }
  const int64_t chunk = 1 << 22;
  emg_particle *d_rec = nullptr;
  if ((rc = dalloc(&d_rec, (size_t)std::min<int64_t>(chunk, total), "staging"))) {
    free_state64(tmp);
    return rc;
  }
  int64_t m = 0;
  for (int s = 0; s < c->sp.nspec && !rc; s++) {
    for (int64_t k = 0; k < c->sp.cnt[s]; k += chunk) {
      int64_t n = std::min(chunk, c->sp.cnt[s] - k);
      hipLaunchKernelGGL(k_join, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                         c->stream, src, c->sp.off[s] + k, n, s + 1, d_rec);
      rc = hip_check(hipStreamSynchronize(c->stream), "download join");
      if (rc) break;
      rc = hip_check(hipMemcpy(pbuf + m + k, d_rec, sizeof(emg_particle) * n, hipMemcpyDeviceToHost),
                     "download copy");
      if (rc) break;
    }
    m += c->sp.cnt[s];
  }
  dfree(d_rec);
  free_state64(tmp);
  if (!rc) rc = check_flags(c);
  return rc;
}

int emg_set_fields(emg_ctx *c, const double *eb, const int32_t lo[3], const int32_t hi[3]) {
  CTX_CHECK(c);
  if (!eb || !lo || !hi) return set_error(EMG_ERR_ARG, "null argument");
  int n[3] = {c->g.nx, c->g.ny, c->g.nz};
  int elo[3], ext[3];
  for (int d = 0; d < 3; d++) {
    if (lo[d] > -1 || hi[d] < n[d])
      return set_error(EMG_ERR_ARG, "eb bounds must cover -1..n on every axis");
    elo[d] = lo[d];
    ext[d] = hi[d] - lo[d] + 1;
  }
  size_t bytes = sizeof(double) * 6 * (size_t)ext[0] * ext[1] * ext[2];
  if (bytes > c->eb_dev_bytes) {
    int rc = hip_check(hipStreamSynchronize(c->stream), "stream synchronize");
    if (rc) return rc;
    dfree(c->eb_dev);
    c->eb_dev = nullptr;
    c->eb_dev_bytes = 0;
    if ((rc = dalloc(&c->eb_dev, bytes / sizeof(double), "eb staging"))) return rc;
    c->eb_dev_bytes = bytes;
  }
  int rc = hip_check(hipStreamSynchronize(c->stream), "stream synchronize");
  if (rc) return rc;
  rc = hip_check(hipMemcpy(c->eb_dev, eb, bytes, hipMemcpyHostToDevice), "eb upload");
  if (rc) return rc;
  for (int d = 0; d < 3; d++) {
    c->eb_lo[d] = elo[d];
    c->eb_ext[d] = ext[d];
  }
  launch_relocate64(c, elo, ext);
  c->fields_set = true;
  c->nodes_stale = false;
  return hip_check(hipGetLastError(), "relocate launch");
}

// ---- device-resident fields: the Yee update on the uploaded eb
static int need_eb(emg_ctx *c) {
  if (!c->fields_set || !c->eb_dev) return set_error(EMG_ERR_STATE, "emg_set_fields has not been called");
  return EMG_OK;
}

int emg_bfield(emg_ctx *c) {
  CTX_CHECK(c);
  int rc = need_eb(c);
  if (rc) return rc;
  launch_bfield(c);
  c->nodes_stale = true;
  return hip_check(hipGetLastError(), "bfield launch");
}

int emg_efield(emg_ctx *c, double tcs, double dlt) {
  CTX_CHECK(c);
  int rc = need_eb(c);
  if (rc) return rc;
  if (!c->j_valid) return set_error(EMG_ERR_STATE, "no current deposited (call emg_push_em first)");
  launch_efield(c, tcs, dlt);
  c->nodes_stale = true;
  return hip_check(hipGetLastError(), "efield launch");
}

int emg_grdcrt(emg_ctx *c) {
  CTX_CHECK(c);
  int rc = need_eb(c);
  if (rc) return rc;
  launch_relocate64(c, c->eb_lo, c->eb_ext);
  c->nodes_stale = false;
  return hip_check(hipGetLastError(), "relocate launch");
}

int emg_get_fields(emg_ctx *c, double *eb, const int32_t lo[3], const int32_t hi[3]) {
  CTX_CHECK(c);
  int rc = need_eb(c);
  if (rc) return rc;
  if (!eb || !lo || !hi) return set_error(EMG_ERR_ARG, "null argument");
  for (int d = 0; d < 3; d++)
    if (lo[d] < c->eb_lo[d] || hi[d] > c->eb_lo[d] + c->eb_ext[d] - 1 || hi[d] < lo[d])
      return set_error(EMG_ERR_ARG, "eb bounds outside the uploaded array");
  const int l[3] = {lo[0], lo[1], lo[2]};
  const int ext[3] = {hi[0] - lo[0] + 1, hi[1] - lo[1] + 1, hi[2] - lo[2] + 1};
  const size_t count = (size_t)6 * ext[0] * ext[1] * ext[2];
  double *d_out = nullptr;
  if ((rc = dalloc(&d_out, count, "eb read-out"))) return rc;
  launch_eb_box(c, l, ext, d_out);
  rc = check_flags(c);
  if (!rc) rc = hip_check(hipMemcpy(eb, d_out, sizeof(double) * count, hipMemcpyDeviceToHost), "eb download");
  dfree(d_out);
  return rc;
}

int emg_solve_es(emg_ctx *c, double rhobk) {
  CTX_CHECK(c);
  if (!c->rho_valid) return set_error(EMG_ERR_STATE, "rho is not deposited for the current positions");
  if (!c->eb_dev) {  // no eb uploaded: eb(6, -1:n+1)^3 with B = 0
    int n[3] = {c->g.nx, c->g.ny, c->g.nz};
    size_t count = 6;
    for (int d = 0; d < 3; d++) {
      c->eb_lo[d] = -1;
      c->eb_ext[d] = n[d] + 3;
      count *= (size_t)c->eb_ext[d];
    }
    int rc = dalloc(&c->eb_dev, count, "eb");
    if (rc) return rc;
    c->eb_dev_bytes = count * sizeof(double);
    launch_zero(c->eb_dev, c->eb_dev_bytes, c->stream);
  }
  int rc = solve_es(c, rhobk);
  if (rc) return rc;
  c->fields_set = true;
  c->nodes_stale = false;
  return EMG_OK;
}

int emg_set_background(emg_ctx *c, const double *rhobk) {
  CTX_CHECK(c);
  if (!rhobk) {
    dfree(c->bk_dev);
    c->bk_dev = nullptr;
    return EMG_OK;
  }
  int rc;
  if (!c->bk_dev && (rc = dalloc(&c->bk_dev, (size_t)c->ncell, "background charge"))) return rc;
  rc = hip_check(hipMemcpyAsync(c->bk_dev, rhobk, sizeof(double) * c->ncell,
                                hipMemcpyHostToDevice, c->stream),
                 "background copy");
  if (!rc) rc = hip_check(hipStreamSynchronize(c->stream), "background copy");
  return rc;
}

int emg_get_potential(emg_ctx *c, double *phi) {
  CTX_CHECK(c);
  if (!phi) return set_error(EMG_ERR_ARG, "null argument");
  return read_potential(c, phi);
}

int emg_field_energy(emg_ctx *c, double cs, double *engeb) {
  CTX_CHECK(c);
  int rc = need_eb(c);
  if (rc) return rc;
  if (!engeb) return set_error(EMG_ERR_ARG, "null argument");
  if (!c->fe_buf && (rc = dalloc(&c->fe_buf, (size_t)field_energy_part_need() + 2, "field energy")))
    return rc;
  launch_field_energy(c, cs, c->fe_buf, c->fe_buf + field_energy_part_need());
  rc = check_flags(c);
  if (!rc)
    rc = hip_check(hipMemcpy(engeb, c->fe_buf + field_energy_part_need(), 2 * sizeof(double),
                             hipMemcpyDeviceToHost), "field energy copy");
  return rc;
}

static int need_particles(emg_ctx *c) {
  if (!c->have_particles) return set_error(EMG_ERR_STATE, "no particles uploaded");
  return EMG_OK;
}

static int need_fields(emg_ctx *c) {
  if (!c->fields_set) return set_error(EMG_ERR_STATE, "emg_set_fields has not been called");
  if (c->nodes_stale)
    return set_error(EMG_ERR_STATE, "eb changed on the device since the last relocation: call emg_grdcrt");
  return EMG_OK;
}

// PLAIN_FP64 with sort_every > 0: the cell sort that is due before a push.
// With `carry` (a push that can move the particles into the new order
// itself) a re-sort close to the last sort only counts the cells here and
// *carry is set: the caller launches its push with the re-sort and then
// calls sort64_commit.
static int sort_if_due(emg_ctx *c, bool *carry = nullptr) {
  if (carry) *carry = false;
  if (c->cfg.precision != EMG_PLAIN_FP64 || c->cfg.sort_every <= 0) return EMG_OK;
  if (c->pushes_since_sort < c->cfg.sort_every) {
    c->pushes_since_sort++;
    return EMG_OK;
  }
  int rc;
  if (carry && emg::sort64_fusable(c)) {
    rc = emg::sort64_rank(c);
    if (!rc) *carry = true;
  } else {
    rc = sort64(c);
  }
  if (!rc) c->pushes_since_sort = 1;
  return rc;
}

int emg_positn(emg_ctx *c, int32_t ustep) {
  CTX_CHECK(c);
  int rc = need_particles(c);
  if (rc) return rc;
  if (ustep < 0) return set_error(EMG_ERR_ARG, "ustep must be >= 0");
  if ((rc = sort_if_due(c))) return rc;
  if (c->cfg.precision == EMG_PLAIN_FP64) {
    launch_positn64(c, ustep);
    if (c->moves_after_sort < emg::kUnsorted) c->moves_after_sort++;
  } else {
    launch_fx_step(c, ustep, false, false, false);
  }
  c->rho_valid = false;
  c->faces_valid = false;
  c->j_valid = false;
  return hip_check(hipGetLastError(), "positn launch");
}

int emg_push_es(emg_ctx *c, int32_t ustep) {
  CTX_CHECK(c);
  int rc = need_particles(c);
  if (!rc) rc = need_fields(c);
  if (rc) return rc;
  if (ustep < 0) return set_error(EMG_ERR_ARG, "ustep must be >= 0");
  bool carry;
  if ((rc = sort_if_due(c, &carry))) return rc;
  if (c->cfg.precision == EMG_PLAIN_FP64) {
    launch_push_es64(c, ustep, carry);  // push + charge at the new positions
    // the carried re-sort's buffers are swapped even if its run tables fail
    if (carry) rc = emg::sort64_commit(c);
    if (c->moves_after_sort < emg::kUnsorted) c->moves_after_sort++;
    c->rho_valid = true;
    c->faces_valid = false;
    c->ke_valid = false;
  } else {
    launch_fx_step(c, ustep, true, true, false);
    c->rho_valid = true;
    c->faces_valid = false;
    c->ke_valid = true;  // fused partials of the new velocities
  }
  c->j_valid = false;
  if (rc) return rc;
  return hip_check(hipGetLastError(), "push launch");
}

int emg_deposit_rho(emg_ctx *c, int32_t func) {
  CTX_CHECK(c);
  int rc = need_particles(c);
  if (rc) return rc;
  if (func < 0 || func > 2) return set_error(EMG_ERR_ARG, "func must be 0, 1 or 2");
  if (!c->rho_valid) {  // a push deposits at its new positions itself
    if (c->cfg.precision == EMG_PLAIN_FP64)
      launch_deposit64(c);
    else
      launch_fx_deposit(c);
    c->rho_valid = true;
    c->faces_valid = false;
  }
  return hip_check(hipGetLastError(), "deposit launch");
}

// Read a grid array back: the read-out kernel writes it into a device staging
// buffer (allocated for the call, or `stage` of `stage_cap` doubles), which
// is copied to the host.
static int read_grid(emg_ctx *c, double *host, const int32_t lo[3], const int32_t hi[3], int comps,
                     int species, bool current, double *stage = nullptr, int64_t stage_cap = 0) {
  if (!host || !lo || !hi) return set_error(EMG_ERR_ARG, "null argument");
  int l[3], h[3];
  int64_t n = comps;
  for (int d = 0; d < 3; d++) {
    if (hi[d] < lo[d]) return set_error(EMG_ERR_ARG, "hi < lo");
    l[d] = lo[d];
    h[d] = hi[d];
    n *= (int64_t)(hi[d] - lo[d] + 1);
  }
  double *d_out = stage;
  int rc;
  if (stage) {
    if (n > stage_cap)
      return set_error(EMG_ERR_CAPACITY, "the read-out exceeds the library's static staging "
                                         "array (2^24 values, 128 MiB): read it with emg_get_rho");
  } else if ((rc = dalloc(&d_out, (size_t)n, "read-out"))) {
    return rc;
  }
  if (current)
    launch_read_current(c, species, d_out, l, h);
  else
    launch_read_rho(c, species, d_out, l, h);
  rc = check_flags(c);
  if (!rc)
    rc = hip_check(hipMemcpyAsync(host, d_out, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream),
                   "read-out copy");
  if (!rc) rc = hip_check(hipStreamSynchronize(c->stream), "read-out copy");
  if (!stage) dfree(d_out);
  return rc;
}

int emg_get_rho(emg_ctx *c, double *rho, const int32_t lo[3], const int32_t hi[3]) {
  CTX_CHECK(c);
  if (!c->rho_valid) return set_error(EMG_ERR_STATE, "rho is not deposited for the current positions");
  return read_grid(c, rho, lo, hi, 1, -1, false);
}

int emg_get_rho_noalloc(emg_ctx *c, double *rho, const int32_t lo[3], const int32_t hi[3]) {
  CTX_CHECK(c);
  if (!c->rho_valid) return set_error(EMG_ERR_STATE, "rho is not deposited for the current positions");
  if (!c->scratch) return set_error(EMG_ERR_STATE, "no scratch address for this device");
  return read_grid(c, rho, lo, hi, 1, -1, false, c->scratch, kScratchDoubles);
}

int emg_get_rhodg(emg_ctx *c, int32_t is, double *rho, const int32_t lo[3], const int32_t hi[3]) {
  CTX_CHECK(c);
  if (is < 0 || is >= c->sp.nspec) return set_error(EMG_ERR_ARG, "species out of range");
  if (!c->rho_valid) return set_error(EMG_ERR_STATE, "rho is not deposited for the current positions");
  return read_grid(c, rho, lo, hi, 1, is, false);
}

int emg_energy(emg_ctx *c, double *engk) {
  CTX_CHECK(c);
  if (!engk) return set_error(EMG_ERR_ARG, "null argument");
  int rc = need_particles(c);
  if (rc) return rc;
  if (c->cfg.precision == EMG_PLAIN_FP64)
    launch_energy64(c);
  else if (c->ke_valid)
    fx_energy_from_push(c);
  else
    launch_fx_energy(c);
  rc = check_flags(c);
  if (rc) return rc;
  double sums[3 * EMG_MAX_SPECIES];
  rc = hip_check(hipMemcpy(sums, c->ke_out, sizeof(double) * 3 * c->sp.nspec, hipMemcpyDeviceToHost),
                 "energy copy");
  if (rc) return rc;
  double vol = c->g.slx * c->g.sly * c->g.slz;
  for (int s = 0; s < c->sp.nspec; s++)
    for (int d = 0; d < 3; d++) engk[3 * s + d] = 0.0 + 0.5 * c->cfg.rm[s] * sums[3 * s + d] / vol;
  return EMG_OK;
}

int emg_push_em(emg_ctx *c, int32_t ustep) {
  CTX_CHECK(c);
  int rc = need_particles(c);
  if (!rc) rc = need_fields(c);
  if (rc) return rc;
  if (ustep < 1) return set_error(EMG_ERR_ARG, "ustep must be >= 1");
  bool carry;
  if ((rc = sort_if_due(c, &carry))) return rc;
  if (c->cfg.precision == EMG_PLAIN_FP64) {
    launch_push_em64(c, ustep, carry);
    // the carried re-sort's buffers are swapped even if its run tables fail
    if (carry) rc = emg::sort64_commit(c);
    if (c->moves_after_sort < emg::kUnsorted) c->moves_after_sort++;
    c->ke_valid = false;
  } else {
    launch_fx_step(c, ustep, true, false, true);
    c->ke_valid = true;  // fused partials of the new velocities
  }
  c->j_valid = true;
  c->rho_valid = false;
  c->faces_valid = false;
  if (rc) return rc;
  return hip_check(hipGetLastError(), "push_em launch");
}

int emg_get_current(emg_ctx *c, double *aj, const int32_t lo[3], const int32_t hi[3]) {
  CTX_CHECK(c);
  if (!c->j_valid) return set_error(EMG_ERR_STATE, "no current deposited");
  return read_grid(c, aj, lo, hi, 3, -1, true);
}

int emg_get_ajdg(emg_ctx *c, int32_t is, double *aj, const int32_t lo[3], const int32_t hi[3]) {
  CTX_CHECK(c);
  if (!c->j_valid) return set_error(EMG_ERR_STATE, "no current deposited");
  if (is < 0 || is >= c->sp.nspec) return set_error(EMG_ERR_ARG, "species index");
  return read_grid(c, aj, lo, hi, 3, is, true);
}

// ---- adjoint window (FP64 form)
static int need_fp64(emg_ctx *c) {
  if (c->cfg.precision != EMG_PLAIN_FP64)
    return set_error(EMG_ERR_UNSUPPORTED, "the adjoint window runs in EMG_PLAIN_FP64 contexts");
  return EMG_OK;
}

static int copy_state(const State64 &src, const State64 &dst, int64_t n, hipStream_t st) {
  const void *from[7] = {src.x, src.y, src.z, src.vx, src.vy, src.vz, src.pid};
  void *to[7] = {dst.x, dst.y, dst.z, dst.vx, dst.vy, dst.vz, dst.pid};
  for (int q = 0; q < 7; q++) {
    size_t bytes = (q < 6 ? sizeof(double) : sizeof(int32_t)) * (size_t)n;
    int rc = hip_check(hipMemcpyAsync(to[q], from[q], bytes, hipMemcpyDeviceToDevice, st),
                       "state copy");
    if (rc) return rc;
  }
  return EMG_OK;
}

static int need_adjoint(emg_ctx *c) {
  int rc = need_fp64(c);
  if (rc) return rc;
  if (c->lam.x) return EMG_OK;
  if ((rc = alloc_state64(c->lam, c->cap_total)) ||
      (rc = dalloc(&c->adj_e, 3 * c->ncell, "adjoint field")) ||
      (rc = dalloc(&c->adj_g, c->ncell, "adjoint density")) ||
      (rc = dalloc(&c->adj_q, EMG_MAX_SPECIES, "adjoint factors")) ||
      (rc = dalloc(&c->adj_sum, 3 * EMG_MAX_SPECIES, "adjoint sums")))
    return rc;
  State64 &L = c->lam;
  double *arr[6] = {L.x, L.y, L.z, L.vx, L.vy, L.vz};
  for (double *p : arr) launch_zero(p, sizeof(double) * c->cap_total, c->stream);
  launch_zero(c->adj_e, sizeof(double) * 3 * c->ncell, c->stream);
  return hip_check(hipGetLastError(), "adjoint allocation");
}

int emg_state_save(emg_ctx *c, int32_t slot) {
  CTX_CHECK(c);
  int rc = need_fp64(c);
  if (!rc) rc = need_particles(c);
  if (rc) return rc;
  if (slot < 0 || slot >= kStateSlots) return set_error(EMG_ERR_ARG, "slot must be 0..63");
  if (!c->slot[slot].x && (rc = alloc_state64(c->slot[slot], c->cap_total))) return rc;
  for (int s = 0; s < c->sp.nspec; s++) c->slot_cnt[slot][s] = c->sp.cnt[s];
  return copy_state(c->s64, c->slot[slot], c->cap_total, c->stream);
}

int emg_state_load(emg_ctx *c, int32_t slot) {
  CTX_CHECK(c);
  int rc = need_fp64(c);
  if (rc) return rc;
  if (slot < 0 || slot >= kStateSlots || !c->slot[slot].x)
    return set_error(EMG_ERR_ARG, "slot holds no saved state");
  for (int s = 0; s < c->sp.nspec; s++) c->sp.cnt[s] = c->slot_cnt[slot][s];
  c->have_particles = true;
  c->moves_after_sort = emg::kUnsorted;
  c->runs_valid = false;
  c->rho_valid = c->j_valid = c->ke_valid = false;
  c->faces_valid = false;
  return copy_state(c->slot[slot], c->s64, c->cap_total, c->stream);
}

int emg_adj_reset(emg_ctx *c) {
  CTX_CHECK(c);
  int rc = need_adjoint(c);
  if (rc) return rc;
  State64 &L = c->lam;
  double *arr[6] = {L.x, L.y, L.z, L.vx, L.vy, L.vz};
  for (double *p : arr) launch_zero(p, sizeof(double) * c->cap_total, c->stream);
  return hip_check(hipGetLastError(), "adjoint reset");
}

int emg_adj_step(emg_ctx *c, int32_t ustep) {
  CTX_CHECK(c);
  int rc = need_adjoint(c);
  if (!rc) rc = need_particles(c);
  if (!rc) rc = need_fields(c);
  if (rc) return rc;
  if (ustep < 0) return set_error(EMG_ERR_ARG, "ustep must be >= 0");
  launch_adj_step64(c, ustep);
  return hip_check(hipGetLastError(), "adjoint step launch");
}

int emg_adj_get_field(emg_ctx *c, double *lam_e) {
  CTX_CHECK(c);
  if (!lam_e) return set_error(EMG_ERR_ARG, "null argument");
  int rc = need_adjoint(c);
  if (!rc) rc = check_flags(c);
  if (rc) return rc;
  return hip_check(hipMemcpy(lam_e, c->adj_e, sizeof(double) * 3 * c->ncell, hipMemcpyDeviceToHost),
                   "adjoint field copy");
}

int emg_adj_add_density(emg_ctx *c, const double *g, const double *qfac) {
  CTX_CHECK(c);
  if (!g || !qfac) return set_error(EMG_ERR_ARG, "null argument");
  int rc = need_adjoint(c);
  if (!rc) rc = need_particles(c);
  if (rc) return rc;
  if ((rc = hip_check(hipMemcpyAsync(c->adj_g, g, sizeof(double) * c->ncell, hipMemcpyHostToDevice,
                                     c->stream), "adjoint density upload")) ||
      (rc = hip_check(hipMemcpyAsync(c->adj_q, qfac, sizeof(double) * c->sp.nspec,
                                     hipMemcpyHostToDevice, c->stream), "adjoint factor upload")))
    return rc;
  launch_adj_density64(c, c->adj_q);
  // the host buffers may be reused once this returns
  return hip_check(hipStreamSynchronize(c->stream), "adjoint density");
}

int emg_adj_velocity_sums(emg_ctx *c, double *out) {
  CTX_CHECK(c);
  if (!out) return set_error(EMG_ERR_ARG, "null argument");
  int rc = need_adjoint(c);
  if (!rc) rc = need_particles(c);
  if (rc) return rc;
  launch_adj_vsum(c, c->adj_sum);
  rc = check_flags(c);
  if (rc) return rc;
  return hip_check(hipMemcpy(out, c->adj_sum, sizeof(double) * 3 * c->sp.nspec,
                             hipMemcpyDeviceToHost), "adjoint sums copy");
}

int emg_sort(emg_ctx *c) {
  CTX_CHECK(c);
  int rc = need_particles(c);
  if (rc) return rc;
  if (c->cfg.precision == EMG_PLAIN_FP64) {
    rc = sort64(c);
    if (!rc) c->pushes_since_sort = 0;
  } else {
    rc = fx_rebucket(c);
  }
  return rc;
}

int emg_init_device(emg_ctx *c, const emg_device_species *spec, int32_t nspec, uint64_t seed) {
  CTX_CHECK(c);
  if (!spec || nspec != c->sp.nspec) return set_error(EMG_ERR_ARG, "species table");
  for (int s = 0; s < nspec; s++) {
    if (spec[s].count < 0 || spec[s].count > c->cfg.capacity[s])
      return set_error(EMG_ERR_CAPACITY, "species count exceeds the context capacity");
  }
  for (int s = 0; s < nspec; s++) c->sp.cnt[s] = spec[s].count;
  int rc;
  if (c->cfg.precision == EMG_PLAIN_FP64) {
    rc = init_device_into(c, c->s64, spec, nspec, seed);
  } else {
    State64 tmp{};
    if ((rc = alloc_state64(tmp, c->cap_total))) return rc;
    rc = init_device_into(c, tmp, spec, nspec, seed);
    if (!rc) rc = fx_from64(c, tmp);
    if (!rc) rc = hip_check(hipStreamSynchronize(c->stream), "init");
    free_state64(tmp);
  }
  if (!rc) rc = check_flags(c);
  if (rc) {
    for (int s = 0; s < nspec; s++) c->sp.cnt[s] = 0;
    return rc;
  }
  c->have_particles = true;
  c->pushes_since_sort = c->cfg.sort_every;  // unsorted: sort before the first push
  c->moves_after_sort = emg::kUnsorted;
  c->runs_valid = false;
  c->rho_valid = false;
  c->faces_valid = false;
  c->j_valid = false;
  c->ke_valid = false;
  return EMG_OK;
}

int emg_enable_timing(emg_ctx *c, int32_t on) {
  CTX_CHECK(c);
  if (!on) timer_flush(c);
  c->timing = on != 0;
  return EMG_OK;
}

int emg_kernel_stats(emg_ctx *c, int32_t which, double *total_ms, int64_t *launches) {
  CTX_CHECK(c);
  if (which < 0 || which >= kTNum) return set_error(EMG_ERR_ARG, "kernel kind");
  timer_flush(c);
  if (total_ms) *total_ms = c->timers[which].total_ms;
  if (launches) *launches = c->timers[which].launches;
  return EMG_OK;
}

int emg_reset_kernel_stats(emg_ctx *c) {
  CTX_CHECK(c);
  timer_flush(c);
  for (int k = 0; k < kTNum; k++) {
    c->timers[k].total_ms = 0.0;
    c->timers[k].launches = 0;
    c->timers[k].last_ms = 0.0;
  }
  return EMG_OK;
}

}  // extern "C"
