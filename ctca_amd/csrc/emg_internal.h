// emg_internal.h -- context layout and kernel launchers of libemgpu.
// Not part of the public ABI (include/emgpu.h is).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "emgpu.h"

namespace emg {

constexpr int kBlock = 256;
constexpr int kFixShift = 24;  // rho fixed point: unit CIC weight = 2^24
constexpr double kFixScale = 16777216.0;
constexpr int kJShift = 24;  // current fixed point: 2^24 per unit flux (cell / unit time)
constexpr double kJScale = 16777216.0;
constexpr int kStateSlots = 64;
constexpr int kUnsorted = 1 << 30;  // moves_after_sort of a state in no known order  // saved particle states of the adjoint window

// device error flags
constexpr int kFlagDomain = 1;    // particle outside [0, L] / non-finite
constexpr int kFlagSpillLost = 2; // FP32 re-bucketing lost a particle (capacity)
constexpr int kFlagFar = 4;       // particle moved more than one box length

enum TimerKind { kTPush = 0, kTDeposit = 1, kTSort = 2, kTRelocate = 3, kTEnergy = 4, kTNum = 5 };

struct Grid {
  int nx, ny, nz;
  double slx, sly, slz;
  int cbx, cby;  // FP32 form: bits of ix and iy in the packed cell index (ceil(log2 n))
};

struct SpeciesTab {
  int nspec;
  int64_t off[EMG_MAX_SPECIES + 1];  // segment start in the FP64 state arrays
  int64_t cnt[EMG_MAX_SPECIES];
  double q[EMG_MAX_SPECIES];
  double qm[EMG_MAX_SPECIES];
};

// FP64 (plain) particle state, structure of arrays, species segments at off[]
struct State64 {
  double *x, *y, *z, *vx, *vy, *vz;
  int32_t *pid;
};

// FP32 cell + offset state, structure of arrays, tile-bucketed
struct StateFx {
  int32_t *cell;  // packed ix | iy<<10 | iz<<20
  float *fx, *fy, *fz, *vx, *vy, *vz;
  int32_t *pid;   // nullptr unless the context tracks pid
};

// One mover in a bucket's outbox (32 bytes, two 16-byte halves)
struct alignas(16) FxRec {
  int32_t cell;
  float fx, fy, fz, vx, vy, vz;
  int32_t pid;
};

struct FxBuckets {
  int tile;            // tile edge T
  int ntx, nty, ntz;   // tiles per axis
  int ntiles;          // per species
  int nb;              // nspec * ntiles
  int64_t cap_total;   // slots per buffer (bucket area + spill)
  int64_t area;        // bucket area size (spill region starts here)
  int64_t spill_cap;
  int64_t *start[2];   // [nb + 1] first slot of each bucket
  int32_t *count[2];   // [nb] particles in each bucket
  int32_t *nst[2];     // [nb] of which in the first run (stayers); the rest (direct
                       //      arrivals) sit at the end of the bucket's slots
  int32_t *cap[2];     // [nb] slots owned by each bucket
  int32_t *hist;       // [nb] bucketing scratch: histogram
  int32_t *cursor;     // [nb] bucketing scratch: scatter cursors
  int32_t *spill[2];   // [nspec] spill counters
  int32_t *stay;       // [nb] stayers written by the bucket's workgroup (this step)
  int32_t *extra;      // [nb] direct arrivals reserved by atomics (this step)
  FxRec *ob[2];        // [nb * outbox] movers of each bucket, grouped by direction
  int32_t *pre[2];     // [nb * 28] exclusive prefix of the outbox direction counts
  int32_t *shist;      // [nb] spill particles of the live buffer, by the bucket of their cell
  int32_t *scur;       // [nb] scatter cursors for them
  int64_t *spre;       // [nb + 1] exclusive prefix of shist
  FxRec *spl;          // [nspec * spill_cap] the spill particles sorted by bucket
  StateFx st[2];
  int cur;             // buffer holding the live particles
};

struct Timer {
  hipEvent_t ev[2];
  int pending;
  double total_ms;
  int64_t launches;
  double last_ms;
};

struct EsSolver;  // device field solve (k_poisson.hip)

// A block's run of the FP64 kernels after a sort: len consecutive particles
// of species s from its segment's index beg.  Each species is cut into
// blocks of `run` particles in the sort order (cells in z-major order); a
// block whose particles span fewer than kRunSpan cells is one run, a longer
// one (a sparse region) is cut into np = ceil(span / kRunCells) runs of
// equal cell ranges, so that every run's LDS window covers its cells.
struct Run {
  int64_t beg;
  int32_t len, s;
};
// The FP64 sort key of a position: its cell in z-major order, each index the
// int() of the coordinate clamped into the grid.
__host__ __device__ inline int64_t slab_key_xyz(double x, double y, double z, const Grid &g) {
  int k = (int)z, j = (int)y, i = (int)x;
  k = k < 0 ? 0 : (k >= g.nz ? g.nz - 1 : k);
  j = j < 0 ? 0 : (j >= g.ny ? g.ny - 1 : j);
  i = i < 0 ? 0 : (i >= g.nx ? g.nx - 1 : i);
  return ((int64_t)k * g.ny + j) * g.nx + i;
}

// The active lanes of a wave grouped by equal key, one ballot per distinct
// key (no memory access).  Every lane of the wave must call it (the loop is
// uniform).  lead: the group's first lane; inner: the lane's rank in its
// group, in lane order; size: the group's size, on the lead lane.  A caller
// then sends one atomic per group (the lead lane's) instead of one per lane.
__device__ inline void wave_groups(int64_t key, bool act, int &lead, int &inner, int &size) {
  const int lane = __lane_id();
  unsigned long long todo = __ballot(act);
  lead = lane;
  inner = 0;
  size = 0;
  while (todo) {
    const int l = __builtin_amdgcn_readfirstlane(__ffsll((long long)todo) - 1);
    const int64_t k = (int64_t)(((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(key >> 32), l)
                                 << 32) |
                                (uint32_t)__builtin_amdgcn_readlane((int)key, l));
    const unsigned long long grp = __ballot(act && key == k);
    if (lane == l) size = __popcll(grp);
    if ((grp >> lane) & 1ull) {
      lead = l;
      inner = __builtin_amdgcn_mbcnt_hi((uint32_t)(grp >> 32),
                                        __builtin_amdgcn_mbcnt_lo((uint32_t)grp, 0));
    }
    todo &= ~grp;
  }
}

// Where a push that carries a re-sort writes particle m of species is: slot
// off[is] + start[is nkey + key] - start[is nkey] + rank[m] of t, key being
// the cell of its position before the move (what sort64_rank counted).
struct SortDest {
  State64 t;
  const int64_t *start;
  const int32_t *rank;
  int64_t nkey;
};

constexpr int kRunEs = 4096;   // particles per block of the charge / ES push kernels
constexpr int kRunEm = 1024;   // of the zigzag push
constexpr int kRunSpan = 96;   // cells a block may span and stay whole
constexpr int kRunCells = 64;  // cells per run of a cut block
static_assert(kRunSpan >= kRunCells, "a cut block's runs span over kRunCells / 2 cells");

// Upper-face accumulators of the charge.  The periodic accumulators add the
// weight a particle puts on node n of an axis into node 0; the reference
// keeps it apart until its border exchange and add_boundary_charge, which
// leave the raw deposit of node n in ghost plane -1 and that of node 0 in
// plane n + 1.  A face array per species holds the raw weights of the nodes
// with index n on some axis (filled on demand for a read-out of the ghost
// planes): face X = nodes (nx, b, c), b in 0..ny, c in 0..nz; face Y = nodes
// (a, ny, c), a in 0..nx-1, c in 0..nz; face Z = nodes (a, b, nz), a in
// 0..nx-1, b in 0..ny-1.
__host__ __device__ inline int64_t face_len(const Grid &g) {
  return (int64_t)(g.ny + 1) * (g.nz + 1) + (int64_t)g.nx * (g.nz + 1) + (int64_t)g.nx * g.ny;
}

// face slot of node (a, b, c), a in 0..nx, b in 0..ny, c in 0..nz; -1 when no
// index is n
__host__ __device__ inline int64_t face_index(const Grid &g, int a, int b, int c) {
  if (a == g.nx) return (int64_t)c * (g.ny + 1) + b;
  const int64_t y0 = (int64_t)(g.ny + 1) * (g.nz + 1);
  if (b == g.ny) return y0 + (int64_t)c * g.nx + a;
  if (c == g.nz) return y0 + (int64_t)g.nx * (g.nz + 1) + (int64_t)b * g.nx + a;
  return -1;
}

}  // namespace emg

struct emg_ctx {
  emg_config cfg;
  int device;
  hipStream_t stream;
  bool own_stream;
  emg::Grid g;
  int64_t ncell;
  emg::SpeciesTab sp;
  int64_t cap_total;  // FP64: sum of capacities

  // FP64 state
  emg::State64 s64;
  emg::State64 t64;  // sort scratch (lazy)
  bool have_particles;
  int pushes_since_sort;  // PLAIN_FP64 pushes since the last cell sort (sort_every)
  int moves_after_sort;   // PLAIN_FP64 pushes since the state was last sorted, or kUnsorted

  // FP32 state
  emg::FxBuckets fx;

  // fields
  double *eb_dev;   // eb(6, eb_lo : eb_lo + eb_ext - 1), Fortran layout
  size_t eb_dev_bytes;
  int eb_lo[3], eb_ext[3];
  bool nodes_stale;  // eb changed on the device since the last relocation
  double *fe_buf;    // field-energy partials and result (lazy)
  emg::EsSolver *es; // FFT plans and arrays of the device Poisson solve (lazy)
  double *bk_dev;    // background charge per periodic node (emg_set_background), or null
  double *node64;  // (nx+1)(ny+1)(nz+1) x 8 doubles, nodes 0..n
  float *node32;   // nx ny nz x 8 floats, periodic nodes
  bool fields_set;

  // deposits
  double *rho64;               // nspec x ncell, q-weighted (FP64 form)
  unsigned long long *rhofx;   // nspec x ncell, unit-weight fixed point
  double *j64;                 // nspec x 3 x ncell, unscaled zigzag sums
  unsigned long long *jfx;     // nspec x 3 x ncell fixed point
  bool rho_valid;
  bool j_valid;
  void *face;       // nspec x face_len: upper-face charge (double, or fixed point in FP32_FIXED)
  bool faces_valid; // face holds the deposit of the current positions

  // energy
  double *ke_part;  // partial sums
  int64_t ke_part_n;
  double *ke_out;   // [nspec * 3]
  bool ke_valid;

  // sort scratch
  int32_t *sort_count;   // nspec x ncell (or nb)
  int32_t *sort_cursor;
  int64_t *sort_start;
  int64_t sort_len;
  void *sort_rec;        // cap_total x 64-byte records + cap_total int32 ranks (FP64 sort, lazy)
  emg::Run *runs[2];     // block runs of the last sort: [0] kRunEs, [1] kRunEm particles
  int64_t nruns[2], runs_cap[2];
  int64_t *run_base;     // scan scratch of the run counts (nspec x chunks + 1)
  int64_t run_base_len;
  bool runs_valid;       // the particles are in the order of the sort the runs describe
  int64_t *scan_tmp;
  int64_t scan_tmp_len;

  int *dflag;       // device error flags
  int *hflag;       // pinned host mirror
  double *scratch;  // this device's static staging array (emg_get_rho_noalloc)

  // adjoint window (FP64 form)
  emg::State64 lam;  // lambda_x in x, y, z; lambda_v in vx, vy, vz
  double *adj_e;     // 3 x ncell: transposed gather of qmp * lambda_a
  double *adj_g;     // ncell: uploaded density adjoint
  double *adj_q;     // nspec: its factor per species
  double *adj_sum;   // 3 x nspec: velocity adjoint sums
  emg::State64 slot[emg::kStateSlots];  // saved particle states
  int64_t slot_cnt[emg::kStateSlots][EMG_MAX_SPECIES];

  bool timing;
  emg::Timer timers[emg::kTNum];
};

namespace emg {

// error plumbing (emg_api.hip)
int set_error(int code, const std::string &msg);
int hip_check(hipError_t e, const char *what);

// timers
void timer_begin(emg_ctx *c, int kind);
void timer_end(emg_ctx *c, int kind);

// scan: exclusive prefix sum of n int32 values into int64 out[0..n]; tmp >= scan_tmp_need(n)
int64_t scan_tmp_need(int64_t n);
void exclusive_scan_i32(const int32_t *in, int64_t *out, int64_t n, int64_t *tmp, hipStream_t s);
// two scans in one pass (the sum of `in` must stay below 2^32)
void exclusive_scan2_i32(const int32_t *in, int64_t *out, const int32_t *in2, int64_t *out2,
                         int64_t n, int64_t *tmp, hipStream_t s);

// kernels: FP64 mirror (k_fp64.hip)
void launch_relocate64(emg_ctx *c, int elo[3], int eext[3]);
void launch_positn64(emg_ctx *c, int ustep);
// with resort, the push also moves the particles into the order sort64_rank
// has counted (into t64; sort64_commit then makes it the live state)
void launch_push_es64(emg_ctx *c, int ustep, bool resort);
void launch_deposit64(emg_ctx *c);
void launch_push_em64(emg_ctx *c, int ustep, bool resort);
void launch_faces64(emg_ctx *c);  // c->face (zeroed) += upper-face charge
// device Yee update and field energy on eb_dev (k_yee.hip)
void launch_bfield(emg_ctx *c);
void launch_efield(emg_ctx *c, double tcs, double dlt);
void launch_eb_box(emg_ctx *c, const int lo[3], const int ext[3], double *out_dev);
int64_t field_energy_part_need();
void launch_field_energy(emg_ctx *c, double cs, double *part_dev, double *out_dev);
// device ES field solve (k_poisson.hip): rho of the last deposit + rhobk ->
// phi -> E into eb_dev -> nodes
int solve_es(emg_ctx *c, double rhobk);
int read_potential(emg_ctx *c, double *phi);
void es_solver_free(emg_ctx *c);

// common kernels (k_common.hip)
void launch_relocate32(emg_ctx *c);  // node64-independent: from eb_dev
void launch_energy64(emg_ctx *c);
void launch_read_rho(emg_ctx *c, int species, double *dst_dev, const int lo[3], const int hi[3]);
void launch_read_current(emg_ctx *c, int species, double *dst_dev, const int lo[3], const int hi[3]);
void launch_zero(void *p, size_t bytes, hipStream_t s);
// The FP64 cell sort.  sort64 sorts the live state.  A re-sort close to the
// last sort can ride on the next push instead: sort64_rank counts the
// current cells (each key's start, each particle's rank in its key), the
// push writes every particle to its sorted slot (SortDest), and sort64_commit
// makes that the live state and builds the blocks' run tables of it.
int sort64(emg_ctx *c);
bool sort64_fusable(const emg_ctx *c);
int sort64_rank(emg_ctx *c);
SortDest sort64_dest(const emg_ctx *c);
int sort64_commit(emg_ctx *c);
int init_device(emg_ctx *c, const emg_device_species *spec, int nspec, uint64_t seed);

// adjoint of the FP64 ES push (k_adj.hip)
void launch_adj_step64(emg_ctx *c, int ustep);
void launch_adj_density64(emg_ctx *c, const double *qfac_dev);
void launch_adj_vsum(emg_ctx *c, double *out_dev);

// FP32 + fixed point (k_fx.hip)
int fx_alloc(emg_ctx *c, int64_t ntotal);
int fx_from64(emg_ctx *c, const State64 &src);       // convert + bucket
int fx_to64(emg_ctx *c, State64 &dst);               // convert back (species contiguous)
void launch_fx_step(emg_ctx *c, int ustep, bool push, bool deposit, bool em);
void launch_fx_deposit(emg_ctx *c);
void launch_fx_energy(emg_ctx *c);
void launch_fx_faces(emg_ctx *c);  // c->face (zeroed) += upper-face charge, fixed point
void fx_energy_from_push(emg_ctx *c);
int fx_rebucket(emg_ctx *c);

}  // namespace emg
