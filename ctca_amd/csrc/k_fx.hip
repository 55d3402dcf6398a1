// k_fx.hip -- the FP32 cell-index + offset form of the particle path with
// deterministic int64 fixed-point charge deposition.
//
// Layout.  Particles of species s live in tile buckets b = s * ntiles + t,
// where a tile is a T^3 block of cells.  Each bucket occupies the slots
// [start[b], start[b] + count[b]) of a structure-of-arrays buffer
// (cell: packed ix | iy << 10 | iz << 20, offsets fx, fy, fz in [0, 1),
// velocities vx, vy, vz).  Two buffers alternate: a step reads buffer cur and
// writes every particle into the bucket of its new tile in buffer 1 - cur, so
// the tile order is kept exactly every step at no extra traffic (each
// particle is read once and written once).  Output buckets get
// count + count/4 + 64 slots; a particle that finds its destination full goes
// to a per-species spill region that the next step processes through the
// global-memory path, so nothing is lost.
//
// Workgroup = one bucket.  The T+1 nodes per axis of the tile's relocated
// fields are staged once into LDS; every particle of the bucket gathers from
// LDS.  Charge goes to an LDS tile of T+3 nodes per axis (one node of halo
// each side, covering moves of up to one cell) as integer CIC weights scaled
// by 2^24, whose eight weights are made to sum to exactly 2^24 (the last one
// takes the remainder), then one global int64 atomic per non-zero LDS node.
// Integer sums are associative, so rho is bitwise independent of particle
// order, workgroup scheduling and launch shape.
#include "emg_internal.h"

namespace emg {
namespace {

constexpr int kSpillBlocksPerSpecies = 8;

__device__ __forceinline__ int wrapc(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

__device__ __forceinline__ int pack_cell(int ix, int iy, int iz) {
  return ix | (iy << kCellBits) | (iz << (2 * kCellBits));
}

struct FxArgs {
  Grid g;
  int ntx, nty, ntz, ntiles, nb, nspec;
  StateFx in, out;
  const int64_t *start_in;
  const int32_t *count_in;
  const int32_t *spill_in;
  const int64_t *start_out;
  int32_t *count_out;
  const int32_t *cap_out;
  int32_t *spill_out;
  int64_t area, spill_cap;
  const float *node32;
  unsigned long long *rho;
  int64_t ncell;
  float qm[EMG_MAX_SPECIES];
  float gstep;
  double *ke_part;  // [(nb + nspec * kSpillBlocksPerSpecies) * 3]
  int *flag;
  int32_t *front, *back;
  struct MoverRec *mv;
  int32_t *mv_count;
  int64_t mv_cap, trash;
};

// a particle that left its bucket: written by the push, placed by k_fx_arrive
struct MoverRec {
  int dest, cell, pid, pad;
  float fx, fy, fz, vx, vy, vz;
};
constexpr int kOutbox = 256;  // movers staged in LDS per workgroup

__device__ __forceinline__ void gather_global(const float *__restrict__ node32, const Grid &g,
                                              int ix, int iy, int iz, float fx, float fy,
                                              float fz, float f[6]) {
  int ix1 = ix + 1 == g.nx ? 0 : ix + 1;
  int iy1 = iy + 1 == g.ny ? 0 : iy + 1;
  int iz1 = iz + 1 == g.nz ? 0 : iz + 1;
  float wx[2] = {1.f - fx, fx}, wy[2] = {1.f - fy, fy}, wz[2] = {1.f - fz, fz};
  int xs[2] = {ix, ix1}, ys[2] = {iy, iy1}, zs[2] = {iz, iz1};
#pragma unroll
  for (int c = 0; c < 6; c++) f[c] = 0.f;
#pragma unroll
  for (int c2 = 0; c2 < 2; c2++)
#pragma unroll
    for (int b2 = 0; b2 < 2; b2++)
#pragma unroll
      for (int a2 = 0; a2 < 2; a2++) {
        float w = wx[a2] * wy[b2] * wz[c2];
        const float4 *p = reinterpret_cast<const float4 *>(
            node32 + 8 * (((int64_t)zs[c2] * g.ny + ys[b2]) * g.nx + xs[a2]));
        float4 A = p[0], B = p[1];
        f[0] += w * A.x;
        f[1] += w * A.y;
        f[2] += w * A.z;
        f[3] += w * A.w;
        f[4] += w * B.x;
        f[5] += w * B.y;
      }
}

template <int T>
__device__ __forceinline__ void gather_lds(const float4 *fA, const float2 *fB, int lx, int ly,
                                           int lz, float fx, float fy, float fz, float f[6]) {
  constexpr int N1 = T + 1;
  float wx[2] = {1.f - fx, fx}, wy[2] = {1.f - fy, fy}, wz[2] = {1.f - fz, fz};
#pragma unroll
  for (int c = 0; c < 6; c++) f[c] = 0.f;
#pragma unroll
  for (int c2 = 0; c2 < 2; c2++)
#pragma unroll
    for (int b2 = 0; b2 < 2; b2++)
#pragma unroll
      for (int a2 = 0; a2 < 2; a2++) {
        float w = wx[a2] * wy[b2] * wz[c2];
        int q = ((lz + c2) * N1 + (ly + b2)) * N1 + (lx + a2);
        float4 A = fA[q];
        float2 B = fB[q];
        f[0] += w * A.x;
        f[1] += w * A.y;
        f[2] += w * A.z;
        f[3] += w * A.w;
        f[4] += w * B.x;
        f[5] += w * B.y;
      }
}

// FP32 Buneman-Boris (same scheme as psolve2.F90:233-257)
__device__ __forceinline__ void boris32(float &vx, float &vy, float &vz, const float f[6],
                                        float qmp) {
  float ex = f[0] * qmp, ey = f[1] * qmp, ez = f[2] * qmp;
  float bx = f[3] * qmp, by = f[4] * qmp, bz = f[5] * qmp;
  float bor = 2.f * __builtin_amdgcn_rcpf(1.f + bx * bx + by * by + bz * bz);
  vx += ex;
  vy += ey;
  vz += ez;
  float tx = vx + vy * bz - vz * by;
  float ty = vy + vz * bx - vx * bz;
  float tz = vz + vx * by - vy * bx;
  vx += bor * (ty * bz - tz * by);
  vy += bor * (tz * bx - tx * bz);
  vz += bor * (tx * by - ty * bx);
  vx += ex;
  vy += ey;
  vz += ez;
}

// move one axis: offset += v * dt, carry whole cells into the index, periodic
__device__ __forceinline__ void move_axis(float &f, int &i, float v, float dt, int n, int *flag) {
  float nf = f + v * dt;
  float fl = floorf(nf);
  int d = (int)fl;
  nf -= fl;
  if (nf >= 1.f) {  // a tiny negative offset rounds to 1 - eps -> 1.0
    nf -= 1.f;
    d += 1;
  }
  int ni = i + d;
  if (ni < 0) ni += n;
  else if (ni >= n) ni -= n;
  if ((unsigned)ni >= (unsigned)n) {  // moved more than one box length, or NaN
    atomicOr(flag, kFlagFar);
    ni = ni < 0 ? 0 : n - 1;
  }
  if (!(nf >= 0.f && nf < 1.f)) nf = 0.f;
  f = nf;
  i = ni;
}

// integer CIC weights summing to exactly 2^24
__device__ __forceinline__ void fixed_weights(float fx, float fy, float fz, int W[8]) {
  float wx[2] = {1.f - fx, fx}, wy[2] = {1.f - fy, fy}, wz[2] = {1.f - fz, fz};
  int sum = 0;
#pragma unroll
  for (int q = 0; q < 7; q++) {
    int a2 = q & 1, b2 = (q >> 1) & 1, c2 = q >> 2;
    W[q] = (int)rintf(wx[a2] * wy[b2] * wz[c2] * 16777216.f);
    sum += W[q];
  }
  W[7] = (1 << kFixShift) - sum;
}

struct Pcl {
  int cell, pid;
  float fx, fy, fz, vx, vy, vz;
};

// one particle's state from a structure-of-arrays buffer; `base` is the
// (uniform) first slot of the range, `i` the 32-bit offset inside it
__device__ __forceinline__ void load_pcl(const StateFx &in, int64_t base, unsigned i, Pcl &p) {
  p.cell = (in.cell + base)[i];
  p.pid = in.pid ? (in.pid + base)[i] : 0;
  p.fx = (in.fx + base)[i];
  p.fy = (in.fy + base)[i];
  p.fz = (in.fz + base)[i];
  p.vx = (in.vx + base)[i];
  p.vy = (in.vy + base)[i];
  p.vz = (in.vz + base)[i];
}

// The common case of every step is handled in straight-line code (LDS
// gather, LDS charge tile, staying in the own bucket); the rare cases (a
// particle outside the staged tile, a move beyond the charge halo, a full
// destination bucket) sit behind wave-uniform branches taken only when some
// lane of the wave needs them.
template <int T, bool PUSH, bool DEPOSIT>
__device__ __forceinline__ void fx_one(const FxArgs &a, int s, int b_own, bool act, Pcl p,
                                       int ox, int oy, int oz, bool lds_ok, int64_t own_start,
                                       int *n_stay, int *n_out, MoverRec *outbox,
                                       const float4 *fA, const float2 *fB,
                                       unsigned long long *rh, double ke[3]) {
  constexpr int N3 = T + 3;
  const Grid &g = a.g;
  int ix = p.cell & kCellMask, iy = (p.cell >> kCellBits) & kCellMask, iz = p.cell >> (2 * kCellBits);
  if ((unsigned)ix >= (unsigned)g.nx || (unsigned)iy >= (unsigned)g.ny ||
      (unsigned)iz >= (unsigned)g.nz) {  // never for a stored particle; keeps indexing in bounds
    if (act) atomicOr(a.flag, kFlagDomain);
    ix = 0;
    iy = 0;
    iz = 0;
  }
  float fx = p.fx, fy = p.fy, fz = p.fz, vx = p.vx, vy = p.vy, vz = p.vz;
  if (PUSH) {
    float f[6];
    int lx = ix - ox, ly = iy - oy, lz = iz - oz;
    bool in_tile = lds_ok && (unsigned)lx < (unsigned)T && (unsigned)ly < (unsigned)T &&
                   (unsigned)lz < (unsigned)T;
    if (!in_tile) {
      lx = 0;
      ly = 0;
      lz = 0;
    }
    gather_lds<T>(fA, fB, lx, ly, lz, fx, fy, fz, f);
    if (__ballot(act && !in_tile)) {
      if (act && !in_tile) gather_global(a.node32, g, ix, iy, iz, fx, fy, fz, f);
    }
    boris32(vx, vy, vz, f, a.qm[s] * a.gstep * 0.5f);
    if (act) {
      ke[0] += (double)vx * (double)vx;
      ke[1] += (double)vy * (double)vy;
      ke[2] += (double)vz * (double)vz;
    }
  }
  move_axis(fx, ix, vx, a.gstep, g.nx, a.flag);
  move_axis(fy, iy, vy, a.gstep, g.ny, a.flag);
  move_axis(fz, iz, vz, a.gstep, g.nz, a.flag);
  if (DEPOSIT) {
    int W[8];
    fixed_weights(fx, fy, fz, W);
    int rx = wrapc(ix - ox + 1, g.nx), ry = wrapc(iy - oy + 1, g.ny), rz = wrapc(iz - oz + 1, g.nz);
    bool in_halo = lds_ok && rx <= T + 1 && ry <= T + 1 && rz <= T + 1;
    if (act && in_halo) {
      int q0 = (rz * N3 + ry) * N3 + rx;
#pragma unroll
      for (int q = 0; q < 8; q++) {
        int a2 = q & 1, b2 = (q >> 1) & 1, c2 = q >> 2;
        atomicAdd(&rh[q0 + (c2 * N3 + b2) * N3 + a2], (unsigned long long)(long long)W[q]);
      }
    }
    if (__ballot(act && !in_halo)) {
      if (act && !in_halo) {
        unsigned long long *r = a.rho + (int64_t)s * a.ncell;
        int xs[2] = {ix, ix + 1 == g.nx ? 0 : ix + 1};
        int ys[2] = {iy, iy + 1 == g.ny ? 0 : iy + 1};
        int zs[2] = {iz, iz + 1 == g.nz ? 0 : iz + 1};
#pragma unroll
        for (int q = 0; q < 8; q++) {
          int a2 = q & 1, b2 = (q >> 1) & 1, c2 = q >> 2;
          atomicAdd(&r[((int64_t)zs[c2] * g.ny + ys[b2]) * g.nx + xs[a2]],
                    (unsigned long long)(long long)W[q]);
        }
      }
    }
  }
  // ---- re-bucketing: destination bucket of the new cell.  Particles that
  // stay take consecutive slots at the front of their own output bucket (LDS
  // counter); the others go to the LDS outbox and k_fx_arrive appends them to
  // their new bucket after every workgroup has finished.  No global atomic
  // with a return value sits in this path, and every lane stores (a lane
  // without a slot stores to the trash slot), so loads of the next particle
  // stay in flight across the whole step.
  int dest = s * a.ntiles + ((iz / T) * a.nty + (iy / T)) * a.ntx + (ix / T);
  bool stay = act && dest == b_own;
  int lane = threadIdx.x & 63;
  uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  int64_t dst = a.trash;
  uint64_t m = __ballot(stay);
  if (m) {
    int leader = __ffsll((unsigned long long)m) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(n_stay, __popcll(m));
    base = __shfl(base, leader);
    if (stay) dst = own_start + base + __popcll(m & lt);
  }
  uint64_t mv = __ballot(act && !stay);
  if (mv) {
    int leader = __ffsll((unsigned long long)mv) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(n_out, __popcll(mv));
    base = __shfl(base, leader);
    if (act && !stay) {
      int k = base + __popcll(mv & lt);
      MoverRec r;
      r.dest = dest;
      r.cell = pack_cell(ix, iy, iz);
      r.pid = p.pid;
      r.pad = 0;
      r.fx = fx;
      r.fy = fy;
      r.fz = fz;
      r.vx = vx;
      r.vy = vy;
      r.vz = vz;
      if (k < kOutbox) {
        outbox[k] = r;
      } else {  // outbox full: straight to the global mover list
        int g2 = atomicAdd(a.mv_count, 1);
        if (g2 < a.mv_cap)
          a.mv[g2] = r;
        else
          atomicOr(a.flag, kFlagSpillLost);
      }
    }
  }
  a.out.cell[dst] = pack_cell(ix, iy, iz);
  if (a.out.pid) a.out.pid[dst] = p.pid;
  a.out.fx[dst] = fx;
  a.out.fy[dst] = fy;
  a.out.fz[dst] = fz;
  a.out.vx[dst] = vx;
  a.out.vy[dst] = vy;
  a.out.vz[dst] = vz;
}

__device__ __forceinline__ void block_reduce3(double ke[3], double *out) {
  __shared__ double red[3][kBlock / 64];
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int d = 0; d < 3; d++) {
    double v = ke[d];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o);
    if (lane == 0) red[d][w] = v;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    double v = 0.0;
    for (int k = 0; k < kBlock / 64; k++) v += red[threadIdx.x][k];
    out[threadIdx.x] = v;
  }
}

// XCD-aware bucket order: consecutive workgroups go round-robin to the 8 XCDs,
// so give XCD x the contiguous bucket range [x nb/8, (x+1) nb/8) (neighbouring
// tiles share field nodes and charge halo nodes in that XCD's L2).
__device__ __forceinline__ int bucket_of_block(int bid, int nb) {
  if ((nb & 7) != 0) return bid;
  return (bid & 7) * (nb >> 3) + (bid >> 3);
}

template <int T, bool PUSH, bool DEPOSIT>
__global__ __launch_bounds__(kBlock) void k_fx_step(FxArgs a) {
  constexpr int N1 = T + 1, N3 = T + 3;
  __shared__ float4 fA[PUSH ? N1 * N1 * N1 : 1];
  __shared__ float2 fB[PUSH ? N1 * N1 * N1 : 1];
  __shared__ unsigned long long rh[DEPOSIT ? N3 * N3 * N3 : 1];
  __shared__ MoverRec outbox[kOutbox];
  __shared__ int n_stay, n_out;
  __shared__ int64_t out_base;
  double ke[3] = {0.0, 0.0, 0.0};
  const int bid = blockIdx.x;
  int part = bid;  // energy partial slot: bucket index, then spill blocks
  int s = 0, b = -1, ox = 0, oy = 0, oz = 0;
  if (threadIdx.x == 0) {
    n_stay = 0;
    n_out = 0;
  }
  if (bid < a.nb) {
    b = bucket_of_block(bid, a.nb);
    part = b;
    s = b / a.ntiles;
    const int t = b - s * a.ntiles;
    ox = (t % a.ntx) * T;
    oy = ((t / a.ntx) % a.nty) * T;
    oz = (t / (a.ntx * a.nty)) * T;
    const int cnt = a.count_in[b];
    const int64_t beg = a.start_in[b];
    if (cnt > 0) {
      if (PUSH) {
        for (int q = threadIdx.x; q < 2 * N1 * N1 * N1; q += kBlock) {
          int nq = q >> 1;
          int lx = nq % N1, ly = (nq / N1) % N1, lz = nq / (N1 * N1);
          int gx = wrapc(ox + lx, a.g.nx), gy = wrapc(oy + ly, a.g.ny), gz = wrapc(oz + lz, a.g.nz);
          const float4 *pp = reinterpret_cast<const float4 *>(
              a.node32 + 8 * (((int64_t)gz * a.g.ny + gy) * a.g.nx + gx));
          if (q & 1) {
            float4 B = pp[1];
            fB[nq] = make_float2(B.x, B.y);
          } else {
            fA[nq] = pp[0];
          }
        }
      }
      if (DEPOSIT)
        for (int q = threadIdx.x; q < N3 * N3 * N3; q += kBlock) rh[q] = 0ull;
      __syncthreads();
      const int64_t own_start = a.start_out[b];
      // software pipelined: the next particle's state is in flight while the
      // current one is pushed (the load index is clamped, so every lane loads)
      Pcl cur, nxt;
      load_pcl(a.in, beg, min((int)threadIdx.x, cnt - 1), cur);
      for (int base = 0; base < cnt; base += kBlock) {
        int inext = base + kBlock + (int)threadIdx.x;
        load_pcl(a.in, beg, (unsigned)min(inext, cnt - 1), nxt);
        fx_one<T, PUSH, DEPOSIT>(a, s, b, base + (int)threadIdx.x < cnt, cur, ox, oy, oz, true,
                                 own_start, &n_stay, &n_out, outbox, fA, fB, rh, ke);
        cur = nxt;
      }
      if (DEPOSIT) {
        __syncthreads();
        unsigned long long *r = a.rho + (int64_t)s * a.ncell;
        for (int q = threadIdx.x; q < N3 * N3 * N3; q += kBlock) {
          unsigned long long v = rh[q];
          if (v != 0ull) {
            int rx = q % N3, ry = (q / N3) % N3, rz = q / (N3 * N3);
            int gx = wrapc(ox - 1 + rx, a.g.nx);
            int gy = wrapc(oy - 1 + ry, a.g.ny);
            int gz = wrapc(oz - 1 + rz, a.g.nz);
            atomicAdd(&r[((int64_t)gz * a.g.ny + gy) * a.g.nx + gx], v);
          }
        }
      }
    }
  } else {
    // spill particles of the previous step: global-memory gather and deposit
    int k = bid - a.nb;
    s = k / kSpillBlocksPerSpecies;
    int sub = k % kSpillBlocksPerSpecies;
    int cnt = a.spill_in[s];
    if (cnt > a.spill_cap) cnt = (int)a.spill_cap;
    int64_t beg = a.area + (int64_t)s * a.spill_cap;
    int rounds = (cnt + kBlock * kSpillBlocksPerSpecies - 1) / (kBlock * kSpillBlocksPerSpecies);
    __syncthreads();
    for (int r = 0; r < rounds; r++) {
      int i = (r * kSpillBlocksPerSpecies + sub) * kBlock + threadIdx.x;
      Pcl cur{};
      if (i < cnt) load_pcl(a.in, beg, (unsigned)i, cur);
      fx_one<T, PUSH, DEPOSIT>(a, s, -1, i < cnt, cur, 0, 0, 0, false, 0, &n_stay, &n_out, outbox,
                               fA, fB, rh, ke);
    }
  }
  // hand the outbox to the global mover list and publish the stayer count
  __syncthreads();
  int nout = min(n_out, kOutbox);
  if (threadIdx.x == 0) {
    if (b >= 0) a.front[b] = n_stay;
    out_base = nout > 0 ? (int64_t)atomicAdd(a.mv_count, nout) : 0;
  }
  __syncthreads();
  for (int q = threadIdx.x; q < nout; q += kBlock) {
    int64_t g2 = out_base + q;
    if (g2 < a.mv_cap)
      a.mv[g2] = outbox[q];
    else
      atomicOr(a.flag, kFlagSpillLost);
  }
  if (PUSH) block_reduce3(ke, a.ke_part + 3 * (int64_t)part);
}

// movers of the step: appended behind the stayers of their destination bucket
// (or, when that is full, to the species' spill region)
__global__ __launch_bounds__(kBlock) void k_fx_arrive(FxArgs a) {
  int n = *a.mv_count;
  if (n > a.mv_cap) n = (int)a.mv_cap;
  for (int k = blockIdx.x * kBlock + threadIdx.x; k < n; k += gridDim.x * kBlock) {
    MoverRec r = a.mv[k];
    int dest = r.dest;
    int slot = a.front[dest] + atomicAdd(&a.back[dest], 1);
    int64_t dst;
    if (slot < a.cap_out[dest]) {
      dst = a.start_out[dest] + slot;
    } else {
      int s = dest / a.ntiles;
      int q = atomicAdd(&a.spill_out[s], 1);
      if (q >= a.spill_cap) {
        atomicOr(a.flag, kFlagSpillLost);
        continue;
      }
      dst = a.area + (int64_t)s * a.spill_cap + q;
    }
    a.out.cell[dst] = r.cell;
    if (a.out.pid) a.out.pid[dst] = r.pid;
    a.out.fx[dst] = r.fx;
    a.out.fy[dst] = r.fy;
    a.out.fz[dst] = r.fz;
    a.out.vx[dst] = r.vx;
    a.out.vy[dst] = r.vy;
    a.out.vz[dst] = r.vz;
  }
}

// standalone charge deposition of the current positions (no move, no bucket change)
template <int T>
__global__ __launch_bounds__(kBlock) void k_fx_deposit(FxArgs a) {
  constexpr int N3 = T + 3;
  __shared__ unsigned long long rh[N3 * N3 * N3];
  const int bid = blockIdx.x;
  const Grid &g = a.g;
  if (bid < a.nb) {
    const int b = bucket_of_block(bid, a.nb);
    const int s = b / a.ntiles, t = b - s * a.ntiles;
    const int ox = (t % a.ntx) * T, oy = ((t / a.ntx) % a.nty) * T, oz = (t / (a.ntx * a.nty)) * T;
    const int cnt = a.count_in[b];
    const int64_t beg = a.start_in[b];
    if (cnt == 0) return;
    for (int q = threadIdx.x; q < N3 * N3 * N3; q += kBlock) rh[q] = 0ull;
    __syncthreads();
    unsigned long long *r = a.rho + (int64_t)s * a.ncell;
    for (int base = 0; base < cnt; base += kBlock) {
      int i = base + threadIdx.x;
      if (i >= cnt) break;
      int64_t src = beg + i;
      int cell = a.in.cell[src];
      int ix = cell & kCellMask, iy = (cell >> kCellBits) & kCellMask, iz = cell >> (2 * kCellBits);
      int W[8];
      fixed_weights(a.in.fx[src], a.in.fy[src], a.in.fz[src], W);
      int rx = wrapc(ix - ox + 1, g.nx), ry = wrapc(iy - oy + 1, g.ny), rz = wrapc(iz - oz + 1, g.nz);
      if (rx <= T + 1 && ry <= T + 1 && rz <= T + 1) {
#pragma unroll
        for (int q = 0; q < 8; q++) {
          int a2 = q & 1, b2 = (q >> 1) & 1, c2 = q >> 2;
          atomicAdd(&rh[((rz + c2) * N3 + (ry + b2)) * N3 + (rx + a2)],
                    (unsigned long long)(long long)W[q]);
        }
      } else {
        int xs[2] = {ix, ix + 1 == g.nx ? 0 : ix + 1};
        int ys[2] = {iy, iy + 1 == g.ny ? 0 : iy + 1};
        int zs[2] = {iz, iz + 1 == g.nz ? 0 : iz + 1};
#pragma unroll
        for (int q = 0; q < 8; q++) {
          int a2 = q & 1, b2 = (q >> 1) & 1, c2 = q >> 2;
          atomicAdd(&r[((int64_t)zs[c2] * g.ny + ys[b2]) * g.nx + xs[a2]],
                    (unsigned long long)(long long)W[q]);
        }
      }
    }
    __syncthreads();
    for (int q = threadIdx.x; q < N3 * N3 * N3; q += kBlock) {
      unsigned long long v = rh[q];
      if (v != 0ull) {
        int rx = q % N3, ry = (q / N3) % N3, rz = q / (N3 * N3);
        int gx = wrapc(ox - 1 + rx, g.nx), gy = wrapc(oy - 1 + ry, g.ny), gz = wrapc(oz - 1 + rz, g.nz);
        atomicAdd(&r[((int64_t)gz * g.ny + gy) * g.nx + gx], v);
      }
    }
  } else {
    int k = bid - a.nb;
    int s = k / kSpillBlocksPerSpecies, sub = k % kSpillBlocksPerSpecies;
    int cnt = a.spill_in[s];
    if (cnt > a.spill_cap) cnt = (int)a.spill_cap;
    int64_t beg = a.area + (int64_t)s * a.spill_cap;
    unsigned long long *r = a.rho + (int64_t)s * a.ncell;
    for (int i = sub * kBlock + threadIdx.x; i < cnt; i += kBlock * kSpillBlocksPerSpecies) {
      int64_t src = beg + i;
      int cell = a.in.cell[src];
      int ix = cell & kCellMask, iy = (cell >> kCellBits) & kCellMask, iz = cell >> (2 * kCellBits);
      int W[8];
      fixed_weights(a.in.fx[src], a.in.fy[src], a.in.fz[src], W);
      int xs[2] = {ix, ix + 1 == g.nx ? 0 : ix + 1};
      int ys[2] = {iy, iy + 1 == g.ny ? 0 : iy + 1};
      int zs[2] = {iz, iz + 1 == g.nz ? 0 : iz + 1};
#pragma unroll
      for (int q = 0; q < 8; q++) {
        int a2 = q & 1, b2 = (q >> 1) & 1, c2 = q >> 2;
        atomicAdd(&r[((int64_t)zs[c2] * g.ny + ys[b2]) * g.nx + xs[a2]],
                  (unsigned long long)(long long)W[q]);
      }
    }
  }
}

// output capacities of the next step: cap = count + count/4 + 64; clears the
// output counters
__global__ void k_fx_caps(const int32_t *count_in, int32_t *cap, int32_t *count_out,
                          int32_t *spill_out, int nb, int nspec) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < nb) {
    int c = count_in[b];
    cap[b] = c + c / 4 + 64;
    count_out[b] = 0;
  }
  if (b < nspec) spill_out[b] = 0;
}

// before a step: arrival counters and the mover list are empty
__global__ void k_fx_step_prep(int32_t *back, int32_t *mv_count, int nb) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < nb) back[b] = 0;
  if (b == 0) *mv_count = 0;
}

// Arrivals into a full bucket still bump its counter (the particle itself
// goes to the spill region), so after a step the counters are clamped back to
// the number of particles actually stored: count <= capacity, spill <= cap.
__global__ void k_fx_clamp(int32_t *count, const int32_t *front, const int32_t *back,
                           const int32_t *cap, int nb, int32_t *spill, int nspec,
                           int32_t spill_cap) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < nb) count[b] = min(front[b] + back[b], cap[b]);
  if (b < nspec) spill[b] = min(spill[b], spill_cap);
}

// ---- bucketing of input ranges (flat species segments, spill regions or
// existing buckets) into the buckets of the other buffer
struct RangeDesc {
  int nrb;     // ranges [0, nrb) are buckets (species r / ntiles)
  int nspec;   // ranges [nrb, nrb + nspec) are per-species segments
  int ntiles;
  int rb;      // blocks per range
  const int64_t *bstart;
  const int32_t *bcount;
  int64_t seg_start[EMG_MAX_SPECIES];
  const int32_t *seg_count_dev;        // device counts (spill) or nullptr
  int64_t seg_count[EMG_MAX_SPECIES];  // host counts (flat)
};

__device__ __forceinline__ void range_of(const RangeDesc &d, int r, int64_t &beg, int64_t &cnt,
                                         int &s) {
  if (r < d.nrb) {
    beg = d.bstart[r];
    cnt = d.bcount[r];
    s = r / d.ntiles;
  } else {
    s = r - d.nrb;
    beg = d.seg_start[s];
    cnt = d.seg_count_dev ? (int64_t)d.seg_count_dev[s] : d.seg_count[s];
  }
}

template <int T>
__device__ __forceinline__ int bucket_of_cell(int cell, int s, int ntx, int nty, int ntiles) {
  int ix = cell & kCellMask, iy = (cell >> kCellBits) & kCellMask, iz = cell >> (2 * kCellBits);
  return s * ntiles + ((iz / T) * nty + (iy / T)) * ntx + (ix / T);
}

template <int T>
__global__ __launch_bounds__(kBlock) void k_fx_hist(RangeDesc d, StateFx in, int ntx, int nty,
                                                    int32_t *hist) {
  int r = blockIdx.x / d.rb, sub = blockIdx.x % d.rb;
  int64_t beg, cnt;
  int s;
  range_of(d, r, beg, cnt, s);
  for (int64_t i = (int64_t)sub * kBlock + threadIdx.x; i < cnt; i += (int64_t)d.rb * kBlock) {
    int b = bucket_of_cell<T>(in.cell[beg + i], s, ntx, nty, d.ntiles);
    atomicAdd(&hist[b], 1);
  }
}

template <int T>
__global__ __launch_bounds__(kBlock) void k_fx_scatter(RangeDesc d, StateFx in, StateFx out,
                                                       int ntx, int nty, const int64_t *start,
                                                       int32_t *cursor) {
  int r = blockIdx.x / d.rb, sub = blockIdx.x % d.rb;
  int64_t beg, cnt;
  int s;
  range_of(d, r, beg, cnt, s);
  for (int64_t i = (int64_t)sub * kBlock + threadIdx.x; i < cnt; i += (int64_t)d.rb * kBlock) {
    int64_t src = beg + i;
    int cell = in.cell[src];
    int b = bucket_of_cell<T>(cell, s, ntx, nty, d.ntiles);
    int64_t dst = start[b] + atomicAdd(&cursor[b], 1);
    out.cell[dst] = cell;
    if (out.pid) out.pid[dst] = in.pid[src];
    out.fx[dst] = in.fx[src];
    out.fy[dst] = in.fy[src];
    out.fz[dst] = in.fz[src];
    out.vx[dst] = in.vx[src];
    out.vy[dst] = in.vy[src];
    out.vz[dst] = in.vz[src];
  }
}

// FP64 -> FP32 cell + offset (flat, species segments at the same offsets)
__global__ __launch_bounds__(kBlock) void k_fx_from64(State64 s64, SpeciesTab sp, StateFx out,
                                                      Grid g, int *flag) {
  int is = blockIdx.y;
  int64_t m = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (m >= sp.cnt[is]) return;
  m += sp.off[is];
  double p[3] = {s64.x[m], s64.y[m], s64.z[m]};
  int n[3] = {g.nx, g.ny, g.nz};
  int c[3];
  float f[3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    double x = p[d];
    if (!(x >= 0.0 && x <= (double)n[d])) {
      atomicOr(flag, kFlagDomain);
      x = 0.0;
    }
    double fl = floor(x);
    int ci = (int)fl;
    float fo = (float)(x - fl);
    if (fo >= 1.f) {
      fo = 0.f;
      ci += 1;
    }
    if (ci >= n[d]) ci -= n[d];
    c[d] = ci;
    f[d] = fo;
  }
  out.cell[m] = pack_cell(c[0], c[1], c[2]);
  if (out.pid) out.pid[m] = s64.pid[m];
  out.fx[m] = f[0];
  out.fy[m] = f[1];
  out.fz[m] = f[2];
  out.vx[m] = (float)s64.vx[m];
  out.vy[m] = (float)s64.vy[m];
  out.vz[m] = (float)s64.vz[m];
}

// bucketed FP32 -> FP64 species-contiguous (dst offset = flat scan of counts)
__global__ __launch_bounds__(kBlock) void k_fx_to64(StateFx in, const int64_t *bstart,
                                                    const int32_t *bcount, const int64_t *flat,
                                                    int ntiles, SpeciesTab sp, State64 out) {
  int b = blockIdx.x;
  int s = b / ntiles;
  int cnt = bcount[b];
  int64_t beg = bstart[b];
  int64_t base = sp.off[s] + (flat[b] - flat[s * ntiles]);
  for (int i = threadIdx.x; i < cnt; i += kBlock) {
    int64_t src = beg + i, dst = base + i;
    int cell = in.cell[src];
    int ix = cell & kCellMask, iy = (cell >> kCellBits) & kCellMask, iz = cell >> (2 * kCellBits);
    out.x[dst] = (double)ix + (double)in.fx[src];
    out.y[dst] = (double)iy + (double)in.fy[src];
    out.z[dst] = (double)iz + (double)in.fz[src];
    out.vx[dst] = (double)in.vx[src];
    out.vy[dst] = (double)in.vy[src];
    out.vz[dst] = (double)in.vz[src];
    out.pid[dst] = in.pid ? in.pid[src] : 0;
  }
}

// kinetic energy of the bucketed FP32 state (when no fused push partials exist)
__global__ __launch_bounds__(kBlock) void k_fx_energy(StateFx in, const int64_t *bstart,
                                                      const int32_t *bcount, int nb,
                                                      const int32_t *spill, int64_t area,
                                                      int64_t spill_cap, double *part) {
  double ke[3] = {0.0, 0.0, 0.0};
  int bid = blockIdx.x;
  int64_t beg;
  int cnt;
  if (bid < nb) {
    beg = bstart[bid];
    cnt = bcount[bid];
  } else {
    int s = bid - nb;
    beg = area + (int64_t)s * spill_cap;
    cnt = spill[s];
  }
  for (int i = threadIdx.x; i < cnt; i += kBlock) {
    float vx = in.vx[beg + i], vy = in.vy[beg + i], vz = in.vz[beg + i];
    ke[0] += (double)vx * vx;
    ke[1] += (double)vy * vy;
    ke[2] += (double)vz * vz;
  }
  block_reduce3(ke, part + 3 * (int64_t)bid);
}

// per-species sum of block partials in a fixed order; partials are indexed by
// bucket (s * ntiles + t) followed by spill_blocks_per_s entries per species
__global__ void k_ke_final(const double *part, int ntiles, int nb, int spill_blocks_per_s,
                           double *out) {
  __shared__ double red[3][kBlock];
  int s = blockIdx.x;
  double v[3] = {0.0, 0.0, 0.0};
  for (int k = threadIdx.x; k < ntiles; k += kBlock)
    for (int d = 0; d < 3; d++) v[d] += part[3 * ((int64_t)s * ntiles + k) + d];
  for (int k = threadIdx.x; k < spill_blocks_per_s; k += kBlock)
    for (int d = 0; d < 3; d++) v[d] += part[3 * ((int64_t)nb + s * spill_blocks_per_s + k) + d];
  for (int d = 0; d < 3; d++) red[d][threadIdx.x] = v[d];
  __syncthreads();
  for (int o = kBlock / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o)
      for (int d = 0; d < 3; d++) red[d][threadIdx.x] += red[d][threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x < 3) out[3 * s + threadIdx.x] = red[threadIdx.x][0];
}

__global__ void k_copy_i32(const int32_t *src, int n, int32_t *dst) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < n) dst[b] = src[b];
}

FxArgs make_args(emg_ctx *c) {
  FxBuckets &B = c->fx;
  FxArgs a;
  a.g = c->g;
  a.ntx = B.ntx;
  a.nty = B.nty;
  a.ntz = B.ntz;
  a.ntiles = B.ntiles;
  a.nb = B.nb;
  a.nspec = c->sp.nspec;
  a.in = B.st[B.cur];
  a.out = B.st[1 - B.cur];
  a.start_in = B.start[B.cur];
  a.count_in = B.count[B.cur];
  a.spill_in = B.spill[B.cur];
  a.start_out = B.start[1 - B.cur];
  a.count_out = B.count[1 - B.cur];
  a.cap_out = B.capb;
  a.spill_out = B.spill[1 - B.cur];
  a.area = B.area;
  a.spill_cap = B.spill_cap;
  a.node32 = c->node32;
  a.rho = c->rhofx;
  a.ncell = c->ncell;
  for (int s = 0; s < EMG_MAX_SPECIES; s++) a.qm[s] = (float)c->cfg.qm[s];
  a.gstep = 0.f;
  a.ke_part = c->ke_part;
  a.flag = c->dflag;
  a.front = B.front;
  a.back = B.back;
  a.mv = (MoverRec *)B.mv;
  a.mv_count = B.mv_count;
  a.mv_cap = B.mv_cap;
  a.trash = B.trash;
  return a;
}

template <int T>
void step_t(emg_ctx *c, const FxArgs &a, bool push, bool deposit, unsigned grid) {
  if (push && deposit)
    hipLaunchKernelGGL((k_fx_step<T, true, true>), dim3(grid), dim3(kBlock), 0, c->stream, a);
  else if (push)
    hipLaunchKernelGGL((k_fx_step<T, true, false>), dim3(grid), dim3(kBlock), 0, c->stream, a);
  else if (deposit)
    hipLaunchKernelGGL((k_fx_step<T, false, true>), dim3(grid), dim3(kBlock), 0, c->stream, a);
  else
    hipLaunchKernelGGL((k_fx_step<T, false, false>), dim3(grid), dim3(kBlock), 0, c->stream, a);
}

// bucket particles described by d (reading buffer 'src') into buffer 1 - src
int bucket_ranges(emg_ctx *c, RangeDesc d, int nr, int rb, int src) {
  FxBuckets &B = c->fx;
  int dst = 1 - src;
  d.rb = rb;
  unsigned grid = (unsigned)((int64_t)nr * rb);
  launch_zero(B.hist, sizeof(int32_t) * B.nb, c->stream);
  switch (B.tile) {
    case 4:
      hipLaunchKernelGGL(k_fx_hist<4>, dim3(grid), dim3(kBlock), 0, c->stream, d, B.st[src], B.ntx,
                         B.nty, B.hist);
      break;
    case 8:
      hipLaunchKernelGGL(k_fx_hist<8>, dim3(grid), dim3(kBlock), 0, c->stream, d, B.st[src], B.ntx,
                         B.nty, B.hist);
      break;
    default: return set_error(EMG_ERR_ARG, "tile must be 4 or 8");
  }
  // capacities from the histogram, start offsets, cursor = 0, spill = 0
  hipLaunchKernelGGL(k_fx_caps, dim3((B.nb + kBlock - 1) / kBlock), dim3(kBlock), 0, c->stream,
                     B.hist, B.capb, B.cursor, B.spill[dst], B.nb, c->sp.nspec);
  exclusive_scan_i32(B.capb, B.start[dst], B.nb, c->scan_tmp, c->stream);
  switch (B.tile) {
    case 4:
      hipLaunchKernelGGL(k_fx_scatter<4>, dim3(grid), dim3(kBlock), 0, c->stream, d, B.st[src],
                         B.st[dst], B.ntx, B.nty, B.start[dst], B.cursor);
      break;
    case 8:
      hipLaunchKernelGGL(k_fx_scatter<8>, dim3(grid), dim3(kBlock), 0, c->stream, d, B.st[src],
                         B.st[dst], B.ntx, B.nty, B.start[dst], B.cursor);
      break;
  }
  hipLaunchKernelGGL(k_copy_i32, dim3((B.nb + kBlock - 1) / kBlock), dim3(kBlock), 0, c->stream,
                     B.cursor, B.nb, B.count[dst]);
  B.cur = dst;
  return hip_check(hipGetLastError(), "bucketing");
}

}  // namespace

int fx_alloc(emg_ctx *c, int64_t ntotal) {
  FxBuckets &B = c->fx;
  int T = c->cfg.tile > 0 ? c->cfg.tile : 4;
  if (T != 4 && T != 8) return set_error(EMG_ERR_ARG, "tile must be 4 or 8");
  if (c->g.nx % T || c->g.ny % T || c->g.nz % T)
    return set_error(EMG_ERR_ARG, "grid must be a multiple of the tile edge");
  if (c->g.nx > (1 << kCellBits) || c->g.ny > (1 << kCellBits) || c->g.nz > (1 << kCellBits))
    return set_error(EMG_ERR_ARG, "FP32_FIXED supports at most 1024 cells per axis");
  B.tile = T;
  B.ntx = c->g.nx / T;
  B.nty = c->g.ny / T;
  B.ntz = c->g.nz / T;
  B.ntiles = B.ntx * B.nty * B.ntz;
  B.nb = B.ntiles * c->sp.nspec;
  B.area = ntotal + ntotal / 4 + 64 * (int64_t)B.nb;
  int64_t maxs = 0;
  for (int s = 0; s < c->sp.nspec; s++) maxs = c->cfg.capacity[s] > maxs ? c->cfg.capacity[s] : maxs;
  B.spill_cap = maxs / 64 + 4096;
  if (B.spill_cap > (1 << 30)) B.spill_cap = 1 << 30;
  B.cap_total = B.area + B.spill_cap * c->sp.nspec + 1;  // + the trash slot
  B.trash = B.cap_total - 1;
  B.mv_cap = ntotal / 4 + 65536;
  for (int k = 0; k < 2; k++) {
    StateFx &S = B.st[k];
    void *p[7];
    for (int q = 0; q < 7; q++) {
      if (hipMalloc(&p[q], sizeof(float) * B.cap_total) != hipSuccess)
        return set_error(EMG_ERR_NOMEM, "FP32 particle buffers");
    }
    S.cell = (int32_t *)p[0];
    S.fx = (float *)p[1];
    S.fy = (float *)p[2];
    S.fz = (float *)p[3];
    S.vx = (float *)p[4];
    S.vy = (float *)p[5];
    S.vz = (float *)p[6];
    S.pid = nullptr;
    if (c->cfg.track_pid && hipMalloc(&S.pid, sizeof(int32_t) * B.cap_total) != hipSuccess)
      return set_error(EMG_ERR_NOMEM, "FP32 pid buffers");
    if (hipMalloc(&B.start[k], sizeof(int64_t) * (B.nb + 1)) != hipSuccess ||
        hipMalloc(&B.count[k], sizeof(int32_t) * B.nb) != hipSuccess ||
        hipMalloc(&B.spill[k], sizeof(int32_t) * EMG_MAX_SPECIES) != hipSuccess)
      return set_error(EMG_ERR_NOMEM, "bucket tables");
    launch_zero(B.count[k], sizeof(int32_t) * B.nb, c->stream);
    launch_zero(B.spill[k], sizeof(int32_t) * EMG_MAX_SPECIES, c->stream);
    launch_zero(B.start[k], sizeof(int64_t) * (B.nb + 1), c->stream);
  }
  if (hipMalloc(&B.front, sizeof(int32_t) * B.nb) != hipSuccess ||
      hipMalloc(&B.back, sizeof(int32_t) * B.nb) != hipSuccess ||
      hipMalloc(&B.mv, sizeof(MoverRec) * B.mv_cap) != hipSuccess ||
      hipMalloc(&B.mv_count, sizeof(int32_t)) != hipSuccess)
    return set_error(EMG_ERR_NOMEM, "mover list");
  if (hipMalloc(&B.capb, sizeof(int32_t) * B.nb) != hipSuccess ||
      hipMalloc(&B.hist, sizeof(int32_t) * B.nb) != hipSuccess ||
      hipMalloc(&B.cursor, sizeof(int32_t) * B.nb) != hipSuccess)
    return set_error(EMG_ERR_NOMEM, "bucket tables");
  int64_t need = scan_tmp_need(B.nb);
  if (need > c->scan_tmp_len) {
    if (c->scan_tmp) (void)hipFree(c->scan_tmp);
    if (hipMalloc(&c->scan_tmp, sizeof(int64_t) * need) != hipSuccess)
      return set_error(EMG_ERR_NOMEM, "scan scratch");
    c->scan_tmp_len = need;
  }
  int64_t kp = (int64_t)(B.nb + c->sp.nspec * kSpillBlocksPerSpecies) * 3;
  if (hipMalloc(&c->ke_part, sizeof(double) * kp) != hipSuccess)
    return set_error(EMG_ERR_NOMEM, "energy partials");
  c->ke_part_n = kp;
  B.cur = 0;
  return EMG_OK;
}

// FP64 flat state (species segments at sp.off) -> bucketed FP32
int fx_from64(emg_ctx *c, const State64 &src) {
  FxBuckets &B = c->fx;
  int flat = 1 - B.cur;
  int64_t mx = 1;
  for (int s = 0; s < c->sp.nspec; s++) mx = c->sp.cnt[s] > mx ? c->sp.cnt[s] : mx;
  hipLaunchKernelGGL(k_fx_from64, dim3((unsigned)((mx + kBlock - 1) / kBlock), c->sp.nspec),
                     dim3(kBlock), 0, c->stream, src, c->sp, B.st[flat], c->g, c->dflag);
  RangeDesc d{};
  d.nrb = 0;
  d.nspec = c->sp.nspec;
  d.ntiles = B.ntiles;
  d.bstart = nullptr;
  d.bcount = nullptr;
  d.seg_count_dev = nullptr;
  for (int s = 0; s < c->sp.nspec; s++) {
    d.seg_start[s] = c->sp.off[s];
    d.seg_count[s] = c->sp.cnt[s];
  }
  return bucket_ranges(c, d, c->sp.nspec, 2048, flat);
}

// bucketed FP32 (with spill) -> compacted buckets (absorbs the spill regions)
int fx_rebucket(emg_ctx *c) {
  FxBuckets &B = c->fx;
  RangeDesc d{};
  d.nrb = B.nb;
  d.nspec = c->sp.nspec;
  d.ntiles = B.ntiles;
  d.bstart = B.start[B.cur];
  d.bcount = B.count[B.cur];
  d.seg_count_dev = B.spill[B.cur];
  for (int s = 0; s < c->sp.nspec; s++) d.seg_start[s] = B.area + (int64_t)s * B.spill_cap;
  return bucket_ranges(c, d, B.nb + c->sp.nspec, 1, B.cur);
}

int fx_to64(emg_ctx *c, State64 &dst) {
  FxBuckets &B = c->fx;
  // spill regions are absorbed first so every particle sits in a bucket
  int rc = fx_rebucket(c);
  if (rc) return rc;
  // flat offsets of the buckets
  exclusive_scan_i32(B.count[B.cur], B.start[1 - B.cur], B.nb, c->scan_tmp, c->stream);
  hipLaunchKernelGGL(k_fx_to64, dim3(B.nb), dim3(kBlock), 0, c->stream, B.st[B.cur],
                     B.start[B.cur], B.count[B.cur], B.start[1 - B.cur], B.ntiles, c->sp, dst);
  return hip_check(hipGetLastError(), "fx_to64");
}

void launch_fx_step(emg_ctx *c, int ustep, bool push, bool deposit, bool em) {
  (void)em;
  FxBuckets &B = c->fx;
  unsigned gb = (unsigned)((B.nb + kBlock - 1) / kBlock);
  hipLaunchKernelGGL(k_fx_caps, dim3(gb), dim3(kBlock), 0, c->stream, B.count[B.cur], B.capb,
                     B.count[1 - B.cur], B.spill[1 - B.cur], B.nb, c->sp.nspec);
  hipLaunchKernelGGL(k_fx_step_prep, dim3(gb), dim3(kBlock), 0, c->stream, B.back, B.mv_count, B.nb);
  exclusive_scan_i32(B.capb, B.start[1 - B.cur], B.nb, c->scan_tmp, c->stream);
  if (deposit) launch_zero(c->rhofx, sizeof(unsigned long long) * c->ncell * c->sp.nspec, c->stream);
  FxArgs a = make_args(c);
  a.gstep = (float)ustep;
  unsigned grid = (unsigned)(B.nb + c->sp.nspec * kSpillBlocksPerSpecies);
  timer_begin(c, kTPush);
  switch (B.tile) {
    case 4: step_t<4>(c, a, push, deposit, grid); break;
    case 8: step_t<8>(c, a, push, deposit, grid); break;
  }
  timer_end(c, kTPush);
  hipLaunchKernelGGL(k_fx_arrive, dim3(512), dim3(kBlock), 0, c->stream, a);
  hipLaunchKernelGGL(k_fx_clamp, dim3(gb), dim3(kBlock), 0, c->stream, B.count[1 - B.cur], B.front,
                     B.back, B.capb, B.nb, B.spill[1 - B.cur], c->sp.nspec, (int32_t)B.spill_cap);
  B.cur = 1 - B.cur;
}

void launch_fx_deposit(emg_ctx *c) {
  FxBuckets &B = c->fx;
  launch_zero(c->rhofx, sizeof(unsigned long long) * c->ncell * c->sp.nspec, c->stream);
  FxArgs a = make_args(c);
  unsigned grid = (unsigned)(B.nb + c->sp.nspec * kSpillBlocksPerSpecies);
  timer_begin(c, kTDeposit);
  switch (B.tile) {
    case 4: hipLaunchKernelGGL(k_fx_deposit<4>, dim3(grid), dim3(kBlock), 0, c->stream, a); break;
    case 8: hipLaunchKernelGGL(k_fx_deposit<8>, dim3(grid), dim3(kBlock), 0, c->stream, a); break;
  }
  timer_end(c, kTDeposit);
}

void launch_fx_energy(emg_ctx *c) {
  FxBuckets &B = c->fx;
  timer_begin(c, kTEnergy);
  hipLaunchKernelGGL(k_fx_energy, dim3(B.nb + c->sp.nspec), dim3(kBlock), 0, c->stream,
                     B.st[B.cur], B.start[B.cur], B.count[B.cur], B.nb, B.spill[B.cur], B.area,
                     B.spill_cap, c->ke_part);
  hipLaunchKernelGGL(k_ke_final, dim3(c->sp.nspec), dim3(kBlock), 0, c->stream, c->ke_part,
                     B.ntiles, B.nb, 1, c->ke_out);
  timer_end(c, kTEnergy);
}

// sum of the fused push partials (spill blocks included)
void fx_energy_from_push(emg_ctx *c) {
  FxBuckets &B = c->fx;
  hipLaunchKernelGGL(k_ke_final, dim3(c->sp.nspec), dim3(kBlock), 0, c->stream, c->ke_part,
                     B.ntiles, B.nb, kSpillBlocksPerSpecies, c->ke_out);
}

}  // namespace emg
