// k_common.hip -- kernels shared by both precision forms: fills, the
// exclusive scan used by the sorts, deterministic energy reductions, grid
// read-out into the reference's Fortran halo layouts, the FP64 cell sort and
// the seeded device initialiser.
#include <cmath>

#include "emg_internal.h"

namespace emg {
namespace {

__global__ void k_zero16(uint4 *p, int64_t n16) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = make_uint4(0u, 0u, 0u, 0u);
}

__global__ void k_zero4(uint32_t *p, int64_t n4) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0u;
}

// ---------------------------------------------------------------- scan
constexpr int kScanPer = 16;
constexpr int kScanChunk = kBlock * kScanPer;

__device__ int64_t block_excl_scan(int64_t v, int64_t *total) {
  __shared__ int64_t sh[kBlock];
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int o = 1; o < kBlock; o <<= 1) {
    int64_t t = threadIdx.x >= (unsigned)o ? sh[threadIdx.x - o] : 0;
    __syncthreads();
    sh[threadIdx.x] += t;
    __syncthreads();
  }
  int64_t incl = sh[threadIdx.x];
  *total = sh[kBlock - 1];
  __syncthreads();
  return incl - v;
}

// With a second input in2, each term is in + 2^32 in2: one pass scans both
// (the first sum must stay below 2^32), and k_scan_final splits the result.
// A thread's kScanPer terms, all loaded before any is used (a full run as
// four 16-byte loads per input: base is a multiple of kScanPer and the
// inputs are allocation bases, so they are aligned), zero past n.
__device__ __forceinline__ void scan_load(const int32_t *in, const int32_t *in2, int64_t n,
                                          int64_t base, int64_t t[kScanPer]) {
  static_assert(kScanPer % 4 == 0, "whole 16-byte loads");
  int32_t a[kScanPer], b[kScanPer];
  if (base + kScanPer <= n) {
#pragma unroll
    for (int q = 0; q < kScanPer / 4; q++) {
      const int4 u = reinterpret_cast<const int4 *>(in + base)[q];
      a[4 * q] = u.x;
      a[4 * q + 1] = u.y;
      a[4 * q + 2] = u.z;
      a[4 * q + 3] = u.w;
      int4 w = make_int4(0, 0, 0, 0);
      if (in2) w = reinterpret_cast<const int4 *>(in2 + base)[q];
      b[4 * q] = w.x;
      b[4 * q + 1] = w.y;
      b[4 * q + 2] = w.z;
      b[4 * q + 3] = w.w;
    }
  } else {
#pragma unroll
    for (int k = 0; k < kScanPer; k++) {
      a[k] = base + k < n ? in[base + k] : 0;
      b[k] = in2 && base + k < n ? in2[base + k] : 0;
    }
  }
#pragma unroll
  for (int k = 0; k < kScanPer; k++) t[k] = (int64_t)a[k] + (int64_t)b[k] * 4294967296ll;
}

__global__ __launch_bounds__(kBlock) void k_scan_sums(const int32_t *in, const int32_t *in2,
                                                      int64_t n, int64_t *bsum) {
  int64_t base = (int64_t)blockIdx.x * kScanChunk + (int64_t)threadIdx.x * kScanPer;
  int64_t t[kScanPer];
  scan_load(in, in2, n, base, t);
  int64_t s = 0;
#pragma unroll
  for (int k = 0; k < kScanPer; k++) s += t[k];
  int64_t tot;
  block_excl_scan(s, &tot);
  if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kBlock) void k_scan_top(int64_t *bsum, int64_t nblk) {
  // exclusive scan of the block sums in place, bsum[nblk] = total
  int64_t carry = 0;
  for (int64_t base = 0; base < nblk; base += kBlock) {
    int64_t i = base + threadIdx.x;
    int64_t v = i < nblk ? bsum[i] : 0;
    int64_t tot;
    int64_t ex = block_excl_scan(v, &tot);
    if (i < nblk) bsum[i] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0) bsum[nblk] = carry;
}

__global__ __launch_bounds__(kBlock) void k_scan_final(const int32_t *in, const int32_t *in2,
                                                       int64_t n, const int64_t *bsum,
                                                       int64_t nblk, int64_t *out, int64_t *out2) {
  int64_t base = (int64_t)blockIdx.x * kScanChunk + (int64_t)threadIdx.x * kScanPer;
  int64_t t[kScanPer];
  scan_load(in, in2, n, base, t);
  int64_t s = 0;
#pragma unroll
  for (int k = 0; k < kScanPer; k++) s += t[k];
  int64_t tot;
  int64_t run = block_excl_scan(s, &tot) + bsum[blockIdx.x];
  auto put = [&](int64_t i, int64_t v) {
    if (in2) {
      out[i] = v & 0xffffffffll;
      out2[i] = v >> 32;
    } else {
      out[i] = v;
    }
  };
#pragma unroll
  for (int k = 0; k < kScanPer; k++)
    if (base + k < n) {
      put(base + k, run);
      run += t[k];
    }
  if (blockIdx.x == 0 && threadIdx.x == 0) put(n, bsum[nblk]);
}

// ---------------------------------------------------------------- energy
constexpr int kEnergyBlocks = 512;

__global__ __launch_bounds__(kBlock) void k_energy64_part(State64 s, SpeciesTab sp, double *part) {
  int is = blockIdx.y;
  double a[3] = {0.0, 0.0, 0.0};
  for (int64_t m = (int64_t)blockIdx.x * kBlock + threadIdx.x; m < sp.cnt[is];
       m += (int64_t)gridDim.x * kBlock) {
    int64_t p = sp.off[is] + m;
    double vx = s.vx[p], vy = s.vy[p], vz = s.vz[p];
    a[0] = a[0] + vx * vx;
    a[1] = a[1] + vy * vy;
    a[2] = a[2] + vz * vz;
  }
  __shared__ double red[3][kBlock];
  for (int d = 0; d < 3; d++) red[d][threadIdx.x] = a[d];
  __syncthreads();
  for (int o = kBlock / 2; o > 0; o >>= 1) {
    if (threadIdx.x < (unsigned)o)
      for (int d = 0; d < 3; d++) red[d][threadIdx.x] += red[d][threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x < 3)
    part[3 * ((int64_t)is * gridDim.x + blockIdx.x) + threadIdx.x] = red[threadIdx.x][0];
}

__global__ __launch_bounds__(kBlock) void k_energy64_final(const double *part, int nblk, double *out) {
  int is = blockIdx.x;
  __shared__ double red[3][kBlock];
  double a[3] = {0.0, 0.0, 0.0};
  for (int k = threadIdx.x; k < nblk; k += kBlock)
    for (int d = 0; d < 3; d++) a[d] += part[3 * ((int64_t)is * nblk + k) + d];
  for (int d = 0; d < 3; d++) red[d][threadIdx.x] = a[d];
  __syncthreads();
  for (int o = kBlock / 2; o > 0; o >>= 1) {
    if (threadIdx.x < (unsigned)o)
      for (int d = 0; d < 3; d++) red[d][threadIdx.x] += red[d][threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x < 3) out[3 * is + threadIdx.x] = red[threadIdx.x][0];
}

// ---------------------------------------------------------------- read-out
__device__ __forceinline__ int pmod(int i, int n) {
  int r = i % n;
  return r < 0 ? r + n : r;
}

struct ReadArgs {
  Grid g;
  int nspec;
  int species;  // -1: all species
  double q[EMG_MAX_SPECIES];
  int lo[3], ext[3];
  int64_t ncell;
};

// The raw (unfolded) deposit of species accumulator P at node (a, b, c), a in
// 0..nx, b in 0..ny, c in 0..nz: nodes with an index n are in the face array
// F; any other node is its periodic accumulator less the face nodes folded
// into it (its images with a 0 index replaced by n).
template <class V>
__device__ __forceinline__ V raw_node(const V *P, const V *F, const Grid &g, int a, int b, int c) {
  const int64_t fi = face_index(g, a, b, c);
  if (fi >= 0) return F[fi];
  V v = P[((int64_t)c * g.ny + b) * g.nx + a];
#pragma unroll
  for (int m = 1; m < 8; m++) {
    if (((m & 1) && a != 0) || ((m & 2) && b != 0) || ((m & 4) && c != 0)) continue;
    v = v - F[face_index(g, (m & 1) ? g.nx : a, (m & 2) ? g.ny : b, (m & 4) ? g.nz : c)];
  }
  return v;
}

// The raw nodes whose sum the reference leaves at index t of an axis after
// oh3_exchange_borders(CRH) + add_boundary_charge (ohhelp3.c:1010-1031,
// charge.F90:265-409): the exchange copies node n into ghost -1 and node 0
// into ghost n + 1; the fold then adds ghost -1 into node 0 and ghost n + 1
// into node n, on the z axis over every (x, y) plane index, on y only for z
// in 0..n, on x only for y and z in 0..n.
__device__ __forceinline__ int fold_set(int t, int n, bool fold, int set[2]) {
  set[1] = 0;
  if (t == -1) { set[0] = n; return 1; }
  if (t == n + 1) { set[0] = 0; return 1; }
  if (fold && (t == 0 || t == n)) { set[0] = 0; set[1] = n; return 2; }
  set[0] = t;
  return 1;
}

// the reference's post-fold value at (i, j, k), each in -1..n+1, of species
// accumulator P with faces F
template <class V>
__device__ __forceinline__ V ghost_value(const V *P, const V *F, const Grid &g, int i, int j,
                                         int k) {
  const bool kin = k >= 0 && k <= g.nz, jin = j >= 0 && j <= g.ny, iin = i >= 0 && i <= g.nx;
  int sz[2], sy[2], sx[2];
  const int nz = fold_set(k, g.nz, kin, sz);
  const int ny = fold_set(j, g.ny, jin && kin, sy);
  const int nx = fold_set(i, g.nx, iin && jin && kin, sx);
  V v = 0;
  for (int c = 0; c < nz; c++)
    for (int b = 0; b < ny; b++)
      for (int a = 0; a < nx; a++) v = v + raw_node(P, F, g, sx[a], sy[b], sz[c]);
  return v;
}

// rho(1, lo:hi)^3.  Nodes 0..n: the periodic fold (node n = node 0).  Ghost
// planes -1 and n + 1 (with the faces filled): what the reference's border
// exchange and add_boundary_charge leave there.  Indices beyond -1..n+1, or
// ghosts without faces: periodic images.
__device__ __forceinline__ bool ghost_node(const ReadArgs &r, int i, int j, int k) {
  const bool in = i >= 0 && i <= r.g.nx && j >= 0 && j <= r.g.ny && k >= 0 && k <= r.g.nz;
  const bool ext = i >= -1 && i <= r.g.nx + 1 && j >= -1 && j <= r.g.ny + 1 && k >= -1 &&
                   k <= r.g.nz + 1;
  return !in && ext;
}

__global__ __launch_bounds__(kBlock) void k_read_rho64(ReadArgs r, const double *rho,
                                                      const double *face, int64_t nf,
                                                      double *out) {
  int64_t n = (int64_t)r.ext[0] * r.ext[1] * r.ext[2];
  int64_t o = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (o >= n) return;
  int i = (int)(o % r.ext[0]) + r.lo[0];
  int j = (int)((o / r.ext[0]) % r.ext[1]) + r.lo[1];
  int k = (int)(o / ((int64_t)r.ext[0] * r.ext[1])) + r.lo[2];
  double v = 0.0;
  if (face && ghost_node(r, i, j, k)) {
    for (int s = 0; s < r.nspec; s++) {
      if (r.species >= 0 && s != r.species) continue;
      v = v + ghost_value(rho + (int64_t)s * r.ncell, face + (int64_t)s * nf, r.g, i, j, k);
    }
    out[o] = v;
    return;
  }
  int64_t c = ((int64_t)pmod(k, r.g.nz) * r.g.ny + pmod(j, r.g.ny)) * r.g.nx + pmod(i, r.g.nx);
  if (r.species >= 0) {
    v = rho[(int64_t)r.species * r.ncell + c];
  } else {
    for (int s = 0; s < r.nspec; s++) v = v + rho[(int64_t)s * r.ncell + c];
  }
  out[o] = v;
}

__global__ __launch_bounds__(kBlock) void k_read_rhofx(ReadArgs r, const unsigned long long *rho,
                                                      const unsigned long long *face, int64_t nf,
                                                      double *out) {
  int64_t n = (int64_t)r.ext[0] * r.ext[1] * r.ext[2];
  int64_t o = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (o >= n) return;
  int i = (int)(o % r.ext[0]) + r.lo[0];
  int j = (int)((o / r.ext[0]) % r.ext[1]) + r.lo[1];
  int k = (int)(o / ((int64_t)r.ext[0] * r.ext[1])) + r.lo[2];
  const bool ghost = face && ghost_node(r, i, j, k);
  int64_t c = ((int64_t)pmod(k, r.g.nz) * r.g.ny + pmod(j, r.g.ny)) * r.g.nx + pmod(i, r.g.nx);
  double v = 0.0;
  for (int s = 0; s < r.nspec; s++) {
    if (r.species >= 0 && s != r.species) continue;
    long long w = ghost ? (long long)ghost_value(rho + (int64_t)s * r.ncell, face + (int64_t)s * nf,
                                                 r.g, i, j, k)
                        : (long long)rho[(int64_t)s * r.ncell + c];
    v = v + r.q[s] * ((double)w * (1.0 / kFixScale));
  }
  out[o] = v;
}

// aj(3, lo:hi): sum over species of q(is) * J(is) in species order, each term
// added with one fused multiply-add (the same arithmetic as the device E
// update that consumes the current)
template <bool FIX>
__global__ __launch_bounds__(kBlock) void k_read_current(ReadArgs r, const void *jacc, double *out) {
  int64_t n = (int64_t)r.ext[0] * r.ext[1] * r.ext[2];
  int64_t o = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (o >= n) return;
  int i = (int)(o % r.ext[0]) + r.lo[0];
  int j = (int)((o / r.ext[0]) % r.ext[1]) + r.lo[1];
  int k = (int)(o / ((int64_t)r.ext[0] * r.ext[1])) + r.lo[2];
  int64_t c = ((int64_t)pmod(k, r.g.nz) * r.g.ny + pmod(j, r.g.ny)) * r.g.nx + pmod(i, r.g.nx);
  for (int d = 0; d < 3; d++) {
    double v = 0.0;
    for (int s = 0; s < r.nspec; s++) {
      if (r.species >= 0 && s != r.species) continue;
      int64_t q = ((int64_t)s * 3 + d) * r.ncell + c;
      double js = FIX ? (double)(long long)((const unsigned long long *)jacc)[q] * (1.0 / kJScale)
                      : ((const double *)jacc)[q];
      v = fma(js, r.q[s], v);
    }
    out[3 * o + d] = v;
  }
}

// ---------------------------------------------------------------- FP64 sort
// The reference's comparator (int)(a->z - b->z) (psort-struct.c:15-20) orders
// a species by z slab and leaves the order inside a slab to qsort.  Here a
// counting sort on the cell (int(z), int(y), int(x)): z slabs in order, and
// inside a slab the cells in order, so that the particles of a workgroup
// gather from and deposit to a small box of nodes (the order inside a cell is
// whatever the atomic cursors give).
__device__ __forceinline__ int64_t slab_key(const State64 &s, int64_t m, const Grid &g) {
  return slab_key_xyz(s.x[m], s.y[m], s.z[m], g);
}

// count the particles of each key; the count a particle finds is its rank
// among the key's particles (kept for the scatter).  A re-sort's input is the
// last sort's order moved by a few pushes: a wave holds a few distinct keys,
// the particles that crossed a face interleaved with those that stayed.  With
// `grouped` the wave groups its lanes by key (one ballot per distinct key, no
// memory access), then each group's first lane adds the group's size with one
// atomic -- all groups' atomics in one instruction -- and the lanes rank
// themselves inside their group from the count it returns.  An input in no
// known order (up to 64 keys per wave) takes one atomic per lane.
__global__ __launch_bounds__(kBlock) void k_sort_hist(State64 s, SpeciesTab sp, Grid g, int64_t nkey,
                                                      int32_t *count, int32_t *rank, bool grouped) {
  const int is = blockIdx.y;
  const int64_t m = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const bool act = m < sp.cnt[is];  // the whole wave takes part in the ballots
  const int lane = __lane_id();
  const int64_t key = act ? (int64_t)is * nkey + slab_key(s, m + sp.off[is], g) : -1;
  int lead = lane, inner = 0, size = 1;
  if (grouped) wave_groups(key, act, lead, inner, size);  // uniform over the grid
  int base = 0;
  if (act && lead == lane) base = atomicAdd(&count[key], size);
  if (grouped) base = __shfl(base, lead);
  if (act) rank[m + sp.off[is]] = base + inner;
}

// The destinations of a sort are a random permutation: written straight into
// the seven state arrays, every particle dirties seven cache lines at random.
// Instead each particle goes to its slot as one 64-byte record (one line),
// and a second, coalesced pass splits the records into the arrays.
__global__ __launch_bounds__(kBlock) void k_sort_scatter(State64 s, uint4 *rec, SpeciesTab sp,
                                                         Grid g, int64_t nkey,
                                                         const int64_t *start, const int32_t *rank) {
  int is = blockIdx.y;
  int64_t m = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (m >= sp.cnt[is]) return;
  m += sp.off[is];
  int64_t key = (int64_t)is * nkey + slab_key(s, m, g);
  int64_t d = sp.off[is] + (start[key] - start[(int64_t)is * nkey]) + rank[m];
  auto u2 = [](double a, double b) {
    const unsigned long long ia = (unsigned long long)__double_as_longlong(a),
                             ib = (unsigned long long)__double_as_longlong(b);
    return make_uint4((unsigned)ia, (unsigned)(ia >> 32), (unsigned)ib, (unsigned)(ib >> 32));
  };
  uint4 *r = rec + 4 * d;
  r[0] = u2(s.x[m], s.y[m]);
  r[1] = u2(s.z[m], s.vx[m]);
  r[2] = u2(s.vy[m], s.vz[m]);
  r[3] = make_uint4((unsigned)s.pid[m], 0u, 0u, 0u);
}

// A re-sort a few pushes after the last one is nearly the identity: the lanes
// of a wave land on neighbouring slots, so the particles go straight into the
// seven arrays, coalesced, with no record pass.
__global__ __launch_bounds__(kBlock) void k_sort_scatter_direct(State64 s, State64 t, SpeciesTab sp,
                                                                Grid g, int64_t nkey,
                                                                const int64_t *start,
                                                                const int32_t *rank) {
  int is = blockIdx.y;
  int64_t m = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (m >= sp.cnt[is]) return;
  m += sp.off[is];
  // every load before the first store: the arrays of s and t may alias as
  // far as the compiler knows, so a load after a store would wait for it
  const double x = s.x[m], y = s.y[m], z = s.z[m], vx = s.vx[m], vy = s.vy[m], vz = s.vz[m];
  const int32_t pid = s.pid[m], rk = rank[m];
  int64_t key = (int64_t)is * nkey + slab_key(s, m, g);
  int64_t d = sp.off[is] + (start[key] - start[(int64_t)is * nkey]) + rk;
  t.x[d] = x;
  t.y[d] = y;
  t.z[d] = z;
  t.vx[d] = vx;
  t.vy[d] = vy;
  t.vz[d] = vz;
  t.pid[d] = pid;
}

__global__ __launch_bounds__(kBlock) void k_sort_split(const uint4 *rec, State64 t, SpeciesTab sp) {
  int is = blockIdx.y;
  int64_t d = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (d >= sp.cnt[is]) return;
  d += sp.off[is];
  const uint4 *r = rec + 4 * d;
  auto lo = [](uint4 v) { return __longlong_as_double((long long)(((unsigned long long)v.y << 32) | v.x)); };
  auto hi = [](uint4 v) { return __longlong_as_double((long long)(((unsigned long long)v.w << 32) | v.z)); };
  const uint4 a = r[0], b = r[1], c = r[2], e = r[3];
  t.x[d] = lo(a);
  t.y[d] = hi(a);
  t.z[d] = lo(b);
  t.vx[d] = hi(b);
  t.vy[d] = lo(c);
  t.vz[d] = hi(c);
  t.pid[d] = (int32_t)e.x;
}

// ---------------------------------------------------------------- device init
// SplitMix64 (Steele, Lea, Flood 2014) over a (species, particle, draw)
// counter: a stateless stream, so any particle can be generated independently.
__device__ __forceinline__ uint64_t splitmix(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

__device__ __forceinline__ double u01(uint64_t seed, uint64_t ctr) {
  uint64_t z = splitmix(seed + ctr * 0x9e3779b97f4a7c15ull);
  return ((double)(z >> 11) + 0.5) * (1.0 / 9007199254740992.0);  // (0, 1)
}

struct InitSpec {
  int64_t count;
  double vperp, vpara, drift[3], t[9], ramp;
};

__global__ __launch_bounds__(kBlock) void k_init(State64 st, int64_t off, int is, InitSpec p,
                                                 Grid g, uint64_t seed) {
  int64_t m = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (m >= p.count) return;
  uint64_t base = (((uint64_t)is << 40) + (uint64_t)m) * 8ull;
  double u[7];
  for (int k = 0; k < 7; k++) u[k] = u01(seed, base + k);
  double x = g.slx * u[0], y = g.sly * u[1], z;
  if (p.ramp > 1.0)
    z = g.slz * log1p(u[2] * (p.ramp - 1.0)) / log(p.ramp);
  else
    z = g.slz * u[2];
  if (x >= g.slx) x = 0.0;
  if (y >= g.sly) y = 0.0;
  if (z >= g.slz) z = 0.0;
  double r1 = sqrt(-2.0 * log(u[3])), th = 6.283185307179586 * u[4];
  double r2 = sqrt(-2.0 * log(u[5])), th2 = 6.283185307179586 * u[6];
  double vx = p.vperp * r1 * cos(th), vy = p.vperp * r1 * sin(th), vz = p.vpara * r2 * cos(th2);
  int64_t d = off + m;
  st.x[d] = x;
  st.y[d] = y;
  st.z[d] = z;
  st.vx[d] = p.t[0] * vx + p.t[1] * vy + p.t[2] * vz + p.drift[0];
  st.vy[d] = p.t[3] * vx + p.t[4] * vy + p.t[5] * vz + p.drift[1];
  st.vz[d] = p.t[6] * vx + p.t[7] * vy + p.t[8] * vz + p.drift[2];
  st.pid[d] = 0;
}

dim3 species_grid(const SpeciesTab &sp) {
  int64_t mx = 1;
  for (int s = 0; s < sp.nspec; s++) mx = sp.cnt[s] > mx ? sp.cnt[s] : mx;
  return dim3((unsigned)((mx + kBlock - 1) / kBlock), (unsigned)sp.nspec, 1);
}


// Block runs after a sort.  st(c) = start[s nkey + c] - start[s nkey] is
// the index (in the species' segment) of the first particle of cell c.
struct RunCut {
  const int64_t *st;  // start + s nkey
  int64_t s0, nkey;
  __device__ int64_t at(int64_t c) const { return st[c] - s0; }
  // the cell of particle p: the largest c with st(c) <= p
  __device__ int64_t cell_of(int64_t p) const {
    int64_t lo = 0, hi = nkey - 1;
    while (lo < hi) {
      const int64_t mid = (lo + hi + 1) >> 1;
      if (at(mid) <= p) lo = mid;
      else hi = mid - 1;
    }
    return lo;
  }
};

// Block b of species s holds particles [a, e) = [b run, min((b + 1) run, cnt));
// with cf, cl the cells of its first and last particle, it is one run if
// cl - cf < kRunSpan, else np = ceil((cl - cf + 1) / kRunCells) pieces of
// equal cell ranges (the first from a, the last to e; empty ones dropped).
// visit(beg, end) is called for each run in order; returns the count.
template <typename F>
__device__ int64_t block_runs(const RunCut &rc, int64_t a, int64_t e, F visit) {
  const int64_t cf = rc.cell_of(a), cl = rc.cell_of(e - 1);
  const int64_t span = cl - cf + 1;
  if (span <= kRunSpan) {
    visit(a, e);
    return 1;
  }
  const int64_t np = (span + kRunCells - 1) / kRunCells;
  int64_t k = 0, pa = a;
  for (int64_t q = 1; q <= np; q++) {
    const int64_t pe = q == np ? e : rc.at(cf + q * span / np);
    if (pe > pa) {
      visit(pa, pe);
      k++;
      pa = pe;
    }
  }
  return k;
}

// the number of runs per block (grid: blocks of species blockIdx.y) ...
__global__ void k_run_count(const int64_t *start, int64_t nkey, SpeciesTab sp, int64_t nbmax,
                            int run, int32_t *nrun) {
  const int s = blockIdx.y;
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nbmax) return;
  const int64_t a = b * run, cnt = sp.cnt[s];
  int32_t n = 0;
  if (a < cnt) {
    const RunCut rc{start + (int64_t)s * nkey, start[(int64_t)s * nkey], nkey};
    n = (int32_t)block_runs(rc, a, min(a + run, cnt), [](int64_t, int64_t) {});
  }
  nrun[(int64_t)s * nbmax + b] = n;
}

// ... then, at each block's exclusive prefix, its runs (none written at or
// past cap, which the count bounds; the host checks the total against it)
__global__ void k_run_fill(const int64_t *start, int64_t nkey, SpeciesTab sp, int64_t nbmax,
                           int run, const int64_t *base, Run *runs, int64_t cap) {
  const int s = blockIdx.y;
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nbmax) return;
  const int64_t a = b * run, cnt = sp.cnt[s];
  if (a >= cnt) return;
  const RunCut rc{start + (int64_t)s * nkey, start[(int64_t)s * nkey], nkey};
  int64_t at = base[(int64_t)s * nbmax + b];
  block_runs(rc, a, min(a + run, cnt), [&](int64_t p0, int64_t p1) {
    Run r;
    r.beg = p0;
    r.len = (int32_t)(p1 - p0);
    r.s = s;
    if (at < cap) runs[at] = r;
    at++;
  });
}

}  // namespace

void launch_zero(void *p, size_t bytes, hipStream_t s) {
  if (bytes == 0) return;
  if ((bytes & 15) == 0 && (((uintptr_t)p) & 15) == 0) {
    int64_t n16 = (int64_t)(bytes / 16);
    unsigned grid = (unsigned)std::min<int64_t>((n16 + kBlock - 1) / kBlock, 65536);
    hipLaunchKernelGGL(k_zero16, dim3(grid), dim3(kBlock), 0, s, (uint4 *)p, n16);
  } else {
    int64_t n4 = (int64_t)(bytes / 4);
    unsigned grid = (unsigned)std::min<int64_t>((n4 + kBlock - 1) / kBlock, 65536);
    hipLaunchKernelGGL(k_zero4, dim3(grid), dim3(kBlock), 0, s, (uint32_t *)p, n4);
  }
}

int64_t scan_tmp_need(int64_t n) { return (n + kScanChunk - 1) / kScanChunk + 1; }

static void scan_launch(const int32_t *in, const int32_t *in2, int64_t *out, int64_t *out2,
                        int64_t n, int64_t *tmp, hipStream_t s) {
  int64_t nblk = (n + kScanChunk - 1) / kScanChunk;
  if (nblk == 0) {
    launch_zero(out, sizeof(int64_t), s);
    if (out2) launch_zero(out2, sizeof(int64_t), s);
    return;
  }
  hipLaunchKernelGGL(k_scan_sums, dim3((unsigned)nblk), dim3(kBlock), 0, s, in, in2, n, tmp);
  hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(kBlock), 0, s, tmp, nblk);
  hipLaunchKernelGGL(k_scan_final, dim3((unsigned)nblk), dim3(kBlock), 0, s, in, in2, n, tmp, nblk,
                     out, out2);
}

void exclusive_scan_i32(const int32_t *in, int64_t *out, int64_t n, int64_t *tmp, hipStream_t s) {
  scan_launch(in, nullptr, out, nullptr, n, tmp, s);
}

void exclusive_scan2_i32(const int32_t *in, int64_t *out, const int32_t *in2, int64_t *out2,
                         int64_t n, int64_t *tmp, hipStream_t s) {
  scan_launch(in, in2, out, out2, n, tmp, s);
}

void launch_energy64(emg_ctx *c) {
  timer_begin(c, kTEnergy);
  hipLaunchKernelGGL(k_energy64_part, dim3(kEnergyBlocks, c->sp.nspec), dim3(kBlock), 0, c->stream,
                     c->s64, c->sp, c->ke_part);
  hipLaunchKernelGGL(k_energy64_final, dim3(c->sp.nspec), dim3(kBlock), 0, c->stream, c->ke_part,
                     kEnergyBlocks, c->ke_out);
  timer_end(c, kTEnergy);
}

int64_t energy64_part_need(int nspec) { return (int64_t)kEnergyBlocks * nspec * 3; }

static ReadArgs read_args(emg_ctx *c, int species, const int lo[3], const int hi[3]) {
  ReadArgs r;
  r.g = c->g;
  r.nspec = c->sp.nspec;
  r.species = species;
  for (int s = 0; s < EMG_MAX_SPECIES; s++) r.q[s] = c->cfg.q[s];
  for (int d = 0; d < 3; d++) {
    r.lo[d] = lo[d];
    r.ext[d] = hi[d] - lo[d] + 1;
  }
  r.ncell = c->ncell;
  return r;
}

void launch_read_rho(emg_ctx *c, int species, double *dst, const int lo[3], const int hi[3]) {
  ReadArgs r = read_args(c, species, lo, hi);
  int64_t n = (int64_t)r.ext[0] * r.ext[1] * r.ext[2];
  unsigned grid = (unsigned)((n + kBlock - 1) / kBlock);
  // the reference's ghost planes need the upper-face deposit of the current
  // positions: computed once per deposit, when a read-out first asks for them
  bool ghosts = false;
  for (int d = 0; d < 3; d++) {
    const int nd = d == 0 ? c->g.nx : (d == 1 ? c->g.ny : c->g.nz);
    ghosts = ghosts || (lo[d] <= -1 && hi[d] >= -1) || (lo[d] <= nd + 1 && hi[d] >= nd + 1);
  }
  if (ghosts && c->face && !c->faces_valid) {
    if (c->cfg.precision == EMG_PLAIN_FP64)
      launch_faces64(c);
    else
      launch_fx_faces(c);
    c->faces_valid = true;
  }
  const int64_t nf = face_len(c->g);
  if (c->cfg.precision == EMG_PLAIN_FP64)
    hipLaunchKernelGGL(k_read_rho64, dim3(grid), dim3(kBlock), 0, c->stream, r, c->rho64,
                       ghosts ? (const double *)c->face : nullptr, nf, dst);
  else
    hipLaunchKernelGGL(k_read_rhofx, dim3(grid), dim3(kBlock), 0, c->stream, r, c->rhofx,
                       ghosts ? (const unsigned long long *)c->face : nullptr, nf, dst);
}

void launch_read_current(emg_ctx *c, int species, double *dst, const int lo[3], const int hi[3]) {
  ReadArgs r = read_args(c, species, lo, hi);
  int64_t n = (int64_t)r.ext[0] * r.ext[1] * r.ext[2];
  unsigned grid = (unsigned)((n + kBlock - 1) / kBlock);
  if (c->cfg.precision == EMG_PLAIN_FP64)
    hipLaunchKernelGGL(k_read_current<false>, dim3(grid), dim3(kBlock), 0, c->stream, r,
                       (const void *)c->j64, dst);
  else
    hipLaunchKernelGGL(k_read_current<true>, dim3(grid), dim3(kBlock), 0, c->stream, r,
                       (const void *)c->jfx, dst);
}

// pushes after a sort up to which a re-sort writes the arrays directly
constexpr int kResortDirect = 32;

namespace {

int32_t *rank_buf(const emg_ctx *c) {
  return reinterpret_cast<int32_t *>(static_cast<char *>(c->sort_rec) + 64 * (size_t)c->cap_total);
}

// the sort's buffers (once per context; before the timed region)
int sort64_alloc(emg_ctx *c) {
  int64_t nkey = c->ncell;
  int64_t len = nkey * c->sp.nspec;
  if (c->sort_len < len) {
    if (c->sort_count) (void)hipFree(c->sort_count);
    if (c->sort_cursor) (void)hipFree(c->sort_cursor);
    if (c->sort_start) (void)hipFree(c->sort_start);
    c->sort_count = nullptr;
    c->sort_cursor = nullptr;
    c->sort_start = nullptr;
    c->sort_len = 0;
    if (hipMalloc(&c->sort_count, sizeof(int32_t) * len) != hipSuccess ||
        hipMalloc(&c->sort_cursor, sizeof(int32_t) * len) != hipSuccess ||
        hipMalloc(&c->sort_start, sizeof(int64_t) * (len + 1)) != hipSuccess)
      return set_error(EMG_ERR_NOMEM, "sort scratch");
    c->sort_len = len;
  }
  int64_t need = scan_tmp_need(len);
  if (need > c->scan_tmp_len) {
    if (c->scan_tmp) (void)hipFree(c->scan_tmp);
    if (hipMalloc(&c->scan_tmp, sizeof(int64_t) * need) != hipSuccess)
      return set_error(EMG_ERR_NOMEM, "scan scratch");
    c->scan_tmp_len = need;
  }
  if (!c->t64.x) {
    void *p[7];
    for (int q = 0; q < 7; q++)
      if (hipMalloc(&p[q], (q < 6 ? sizeof(double) : sizeof(int32_t)) * c->cap_total) != hipSuccess) {
        for (int r = 0; r < q; r++) (void)hipFree(p[r]);
        return set_error(EMG_ERR_NOMEM, "sort buffers");
      }
    c->t64 = State64{(double *)p[0], (double *)p[1], (double *)p[2], (double *)p[3],
                     (double *)p[4], (double *)p[5], (int32_t *)p[6]};
  }
  if (!c->sort_rec && hipMalloc(&c->sort_rec, 68 * (size_t)c->cap_total) != hipSuccess) {
    c->sort_rec = nullptr;
    return set_error(EMG_ERR_NOMEM, "sort records");
  }
  return EMG_OK;
}

// the sort's keys and ranks: count each key's particles (a particle's rank
// is the count it finds) and scan the counts into each key's start
int sort64_count(emg_ctx *c) {
  const int64_t nkey = c->ncell;
  const int64_t len = nkey * c->sp.nspec;
  launch_zero(c->sort_count, sizeof(int32_t) * len, c->stream);
  hipLaunchKernelGGL(k_sort_hist, species_grid(c->sp), dim3(kBlock), 0, c->stream, c->s64, c->sp,
                     c->g, nkey, c->sort_count, rank_buf(c), c->moves_after_sort <= kResortDirect);
  exclusive_scan_i32(c->sort_count, c->sort_start, len, c->scan_tmp, c->stream);
  return hip_check(hipGetLastError(), "sort");
}

// the sorted order (in t64) becomes the live state, and the blocks' runs of
// it are built, for both block lengths, from the key starts.  The tables only
// speed the kernels up: one that cannot be allocated is left unbuilt (its
// kernels keep their plain slices) and the sort still succeeds.
int sort64_swap_runs(emg_ctx *c) {
  const int64_t nkey = c->ncell;
  const int64_t len = nkey * c->sp.nspec;
  c->moves_after_sort = 0;
  State64 tmp = c->s64;
  c->s64 = c->t64;
  c->t64 = tmp;
  c->runs_valid = false;
  c->nruns[0] = c->nruns[1] = 0;
  int rc = EMG_OK;
  int64_t mx = 1;
  for (int q = 0; q < c->sp.nspec; q++) mx = c->sp.cnt[q] > mx ? c->sp.cnt[q] : mx;
  for (int k = 0; k < 2 && !rc; k++) {
    const int run = k == 0 ? kRunEs : kRunEm;
    const int64_t nbmax = (mx + run - 1) / run;
    const int64_t nb = nbmax * c->sp.nspec;
    // the counts reuse the sort's histogram (nkey nspec entries), which no
    // later pass reads; with more blocks than cells (over `run` particles
    // per cell) the kernels keep their plain slices
    if (nb > len) break;
    if (c->run_base_len < nb + 1) {
      if (c->run_base) (void)hipFree(c->run_base);
      c->run_base_len = 0;
      if (hipMalloc(&c->run_base, sizeof(int64_t) * (nb + 1)) != hipSuccess) {
        (void)hipGetLastError();
        c->run_base = nullptr;
        break;
      }
      c->run_base_len = nb + 1;
    }
    // a cut block's pieces other than its last end at cell starts more than
    // kRunCells / 2 cells apart (equal ranges of span / np cells, and
    // span / np > span kRunCells / (span + kRunCells) >= kRunCells / 2 for
    // span > kRunSpan >= kRunCells)
    const int64_t cap = nb + (int64_t)c->sp.nspec * (nkey / (kRunCells / 2) + 2) + 1;
    if (c->runs_cap[k] < cap) {
      if (c->runs[k]) (void)hipFree(c->runs[k]);
      c->runs_cap[k] = 0;
      if (hipMalloc(&c->runs[k], sizeof(Run) * cap) != hipSuccess) {
        (void)hipGetLastError();
        c->runs[k] = nullptr;
        break;
      }
      c->runs_cap[k] = cap;
    }
    const dim3 gr((unsigned)((nbmax + kBlock - 1) / kBlock), (unsigned)c->sp.nspec);
    hipLaunchKernelGGL(k_run_count, gr, dim3(kBlock), 0, c->stream, c->sort_start, nkey, c->sp,
                       nbmax, run, c->sort_count);
    exclusive_scan_i32(c->sort_count, c->run_base, nb, c->scan_tmp, c->stream);
    hipLaunchKernelGGL(k_run_fill, gr, dim3(kBlock), 0, c->stream, c->sort_start, nkey, c->sp,
                       nbmax, run, c->run_base, c->runs[k], cap);
    int64_t total = 0;
    rc = hip_check(hipMemcpyAsync(&total, c->run_base + nb, sizeof(int64_t),
                                  hipMemcpyDeviceToHost, c->stream), "sort runs");
    if (!rc) rc = hip_check(hipStreamSynchronize(c->stream), "sort runs");
    c->nruns[k] = total <= cap ? total : 0;
  }
  // a table left unbuilt has no runs: its kernels keep their plain slices
  c->runs_valid = rc == EMG_OK;
  if (rc) return rc;
  return hip_check(hipGetLastError(), "sort");
}

}  // namespace

int sort64(emg_ctx *c) {
  int rc = sort64_alloc(c);
  if (rc) return rc;
  timer_begin(c, kTSort);
  rc = sort64_count(c);
  if (!rc) {
    const int64_t nkey = c->ncell;
    const dim3 grid = species_grid(c->sp);
    if (c->moves_after_sort <= kResortDirect) {
      hipLaunchKernelGGL(k_sort_scatter_direct, grid, dim3(kBlock), 0, c->stream, c->s64, c->t64,
                         c->sp, c->g, nkey, c->sort_start, rank_buf(c));
    } else {
      hipLaunchKernelGGL(k_sort_scatter, grid, dim3(kBlock), 0, c->stream, c->s64,
                         (uint4 *)c->sort_rec, c->sp, c->g, nkey, c->sort_start, rank_buf(c));
      hipLaunchKernelGGL(k_sort_split, grid, dim3(kBlock), 0, c->stream,
                         (const uint4 *)c->sort_rec, c->t64, c->sp);
    }
    rc = sort64_swap_runs(c);
  }
  timer_end(c, kTSort);
  return rc;
}

// A re-sort that the next push carries: the state was sorted a few pushes ago
// (the run tables of that order exist and the direct pass would be taken).
bool sort64_fusable(const emg_ctx *c) {
  return c->moves_after_sort <= kResortDirect && c->t64.x && c->sort_rec && c->runs_valid;
}

int sort64_rank(emg_ctx *c) {
  int rc = sort64_alloc(c);
  if (rc) return rc;
  timer_begin(c, kTSort);
  rc = sort64_count(c);
  timer_end(c, kTSort);
  return rc;
}

SortDest sort64_dest(const emg_ctx *c) {
  return SortDest{c->t64, c->sort_start, rank_buf(c), c->ncell};
}

int sort64_commit(emg_ctx *c) {
  timer_begin(c, kTSort);
  const int rc = sort64_swap_runs(c);
  timer_end(c, kTSort);
  return rc;
}

int init_device_into(emg_ctx *c, State64 &st, const emg_device_species *spec, int nspec,
                     uint64_t seed) {
  for (int s = 0; s < nspec; s++) {
    InitSpec p;
    p.count = spec[s].count;
    p.vperp = spec[s].vth_perp;
    p.vpara = spec[s].vth_para;
    for (int d = 0; d < 3; d++) p.drift[d] = spec[s].drift[d];
    for (int d = 0; d < 9; d++) p.t[d] = spec[s].t[d];
    p.ramp = spec[s].ramp_ratio;
    if (p.count > 0)
      hipLaunchKernelGGL(k_init, dim3((unsigned)((p.count + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                         c->stream, st, c->sp.off[s], s, p, c->g, seed);
  }
  return hip_check(hipGetLastError(), "init_device");
}

}  // namespace emg
