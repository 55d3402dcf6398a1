// k_fx.hip -- the FP32 cell-index + offset form of the particle path with
// deterministic int64 fixed-point charge deposition.
//
// Layout.  Particles of species s live in tile buckets b = s * ntiles + t, a
// tile being a T^3 block of cells.  Structure of arrays per buffer: packed
// cell ix | iy << 10 | iz << 20, offsets fx, fy, fz in [0, 1), velocities.
// Bucket b of a buffer owns the slots [start[b], start[b] + cap[b]) and holds
// count[b] particles in two runs: narr[b] particles that arrived from other
// buckets at the front, count[b] - narr[b] particles that stayed at the back:
//   slot(i) = start + i                      for i <  narr
//   slot(i) = start + cap - count + i        for i >= narr
//
// A step reads buffer cur and writes buffer 1 - cur (cap = count + count/4 +
// 64 per bucket, start = exclusive scan of cap):
//   * one workgroup per bucket stages the T+1 nodes per axis of the bucket's
//     relocated fields into LDS, pushes every particle (FP32 Boris), moves it
//     (cell + offset), deposits its charge into an LDS tile of T+3 nodes per
//     axis as integer CIC weights (2^24 per unit weight, the eight weights of
//     a particle summing to exactly 2^24), and writes it back:
//       - a particle that stays in the tile takes the next slot from the back
//         of its own output bucket (LDS counter, no global atomic);
//       - a particle that changes tile goes to an LDS outbox; at the end of
//         the workgroup the outbox is grouped by destination (the 26
//         neighbouring tiles), one atomic per destination reserves a run at
//         the front of that bucket, and the run is written contiguously;
//   * the charge tile is added to the species' global int64 accumulator with
//     one atomic per non-zero node.  Integer sums are associative, so rho is
//     bitwise independent of particle order, scheduling and launch shape.
// Arrivals beyond a bucket's room (cap - previous count) and particles that
// moved further than a neighbouring tile go through a per-species spill
// region that the next step processes through the global-memory path.
#include "emg_internal.h"

namespace emg {
namespace {

constexpr int kSpillBlocksPerSpecies = 8;
constexpr int kOutbox = 256;   // movers staged in LDS per workgroup
constexpr int kFar = 27;       // outbox code of a destination that is not a neighbour

__device__ __forceinline__ int wrapc(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

__device__ __forceinline__ int pack_cell(int ix, int iy, int iz) {
  return ix | (iy << kCellBits) | (iz << (2 * kCellBits));
}

struct FxArgs {
  Grid g;
  int ntx, nty, ntz, ntiles, nb, nspec;
  StateFx in, out;
  // input buffer
  const int64_t *start_in;
  const int32_t *count_in, *narr_in, *cap_in, *spill_in;
  // output buffer
  const int64_t *start_out;
  const int32_t *cap_out;
  int32_t *front;  // stayers per bucket (written by the bucket's workgroup)
  int32_t *back;   // arrivals per bucket (atomic)
  int32_t *spill_out;
  int64_t area, spill_cap, trash;
  const float *node32;
  unsigned long long *rho;
  int64_t ncell;
  float qm[EMG_MAX_SPECIES];
  float gstep;
  double *ke_part;  // [(nb + nspec * kSpillBlocksPerSpecies) * 3]
  int *flag;
};

struct Pcl {
  int cell, pid;
  float fx, fy, fz, vx, vy, vz;
};

struct OutRec {
  int dest, cell, pid, code, rank;
  float fx, fy, fz, vx, vy, vz;
};

__device__ __forceinline__ void load_pcl(const StateFx &in, int64_t slot, Pcl &p) {
  p.cell = in.cell[slot];
  p.pid = in.pid ? in.pid[slot] : 0;
  p.fx = in.fx[slot];
  p.fy = in.fy[slot];
  p.fz = in.fz[slot];
  p.vx = in.vx[slot];
  p.vy = in.vy[slot];
  p.vz = in.vz[slot];
}

__device__ __forceinline__ void store_pcl(const StateFx &out, int64_t dst, int cell, int pid,
                                          float fx, float fy, float fz, float vx, float vy,
                                          float vz) {
  out.cell[dst] = cell;
  if (out.pid) out.pid[dst] = pid;
  out.fx[dst] = fx;
  out.fy[dst] = fy;
  out.fz[dst] = fz;
  out.vx[dst] = vx;
  out.vy[dst] = vy;
  out.vz[dst] = vz;
}

// rare path (a particle outside its staged tile): one corner at a time
__device__ __forceinline__ void gather_global(const float *__restrict__ node32, const Grid &g,
                                              int ix, int iy, int iz, float fx, float fy,
                                              float fz, float f[6]) {
  int ix1 = ix + 1 == g.nx ? 0 : ix + 1;
  int iy1 = iy + 1 == g.ny ? 0 : iy + 1;
  int iz1 = iz + 1 == g.nz ? 0 : iz + 1;
  for (int c = 0; c < 6; c++) f[c] = 0.f;
#pragma unroll 1
  for (int q = 0; q < 8; q++) {
    int a2 = q & 1, b2 = (q >> 1) & 1, c2 = q >> 2;
    float w = (a2 ? fx : 1.f - fx) * (b2 ? fy : 1.f - fy) * (c2 ? fz : 1.f - fz);
    const float4 *p = reinterpret_cast<const float4 *>(
        node32 + 8 * (((int64_t)(c2 ? iz1 : iz) * g.ny + (b2 ? iy1 : iy)) * g.nx + (a2 ? ix1 : ix)));
    float4 A = p[0], B = p[1];
    f[0] += w * A.x;
    f[1] += w * A.y;
    f[2] += w * A.z;
    f[3] += w * A.w;
    f[4] += w * B.x;
    f[5] += w * B.y;
  }
}

// trilinear gather from the LDS field tile, two corners (one x pair) at a
// time: bounds the registers held by LDS reads
template <int T>
__device__ __forceinline__ void gather_lds(const float4 *fA, const float2 *fB, int lx, int ly,
                                           int lz, float fx, float fy, float fz, float f[6]) {
  constexpr int N1 = T + 1;
  float wx0 = 1.f - fx;
  float wyz[4] = {(1.f - fy) * (1.f - fz), fy * (1.f - fz), (1.f - fy) * fz, fy * fz};
  int q0 = (lz * N1 + ly) * N1 + lx;
#pragma unroll
  for (int c = 0; c < 6; c++) f[c] = 0.f;
#pragma unroll
  for (int r = 0; r < 4; r++) {
    int q = q0 + ((r >> 1) * N1 + (r & 1)) * N1;
    float4 A0 = fA[q], A1 = fA[q + 1];
    float2 B0 = fB[q], B1 = fB[q + 1];
    float w0 = wx0 * wyz[r], w1 = fx * wyz[r];
    f[0] += w0 * A0.x + w1 * A1.x;
    f[1] += w0 * A0.y + w1 * A1.y;
    f[2] += w0 * A0.z + w1 * A1.z;
    f[3] += w0 * A0.w + w1 * A1.w;
    f[4] += w0 * B0.x + w1 * B1.x;
    f[5] += w0 * B0.y + w1 * B1.y;
    __builtin_amdgcn_sched_barrier(0);
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
__device__ __forceinline__ void move_axis(float &f, int &i, float v, float dt, int n, int *flag,
                                          bool act) {
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
    if (act) atomicOr(flag, kFlagFar);
    ni = ni < 0 ? 0 : n - 1;
  }
  if (!(nf >= 0.f && nf < 1.f)) nf = 0.f;
  f = nf;
  i = ni;
}

// integer CIC weights summing to exactly 2^24
__device__ __forceinline__ void fixed_weights(float fx, float fy, float fz, int W[8]) {
  float wx[2] = {1.f - fx, fx};
  float wyz[4] = {(1.f - fy) * (1.f - fz) * 16777216.f, fy * (1.f - fz) * 16777216.f,
                  (1.f - fy) * fz * 16777216.f, fy * fz * 16777216.f};
  int sum = 0;
#pragma unroll
  for (int q = 0; q < 7; q++) {
    W[q] = (int)rintf(wx[q & 1] * wyz[q >> 1]);
    sum += W[q];
  }
  W[7] = (1 << kFixShift) - sum;
}

// destination bucket -> outbox code (26 neighbouring tiles) or kFar
__device__ __forceinline__ int dest_code(int dest, int b_own, int ntx, int nty, int ntz,
                                         int ntiles) {
  if (b_own < 0 || ntx < 3 || nty < 3 || ntz < 3) return kFar;
  if (dest / ntiles != b_own / ntiles) return kFar;
  int t = dest % ntiles, o = b_own % ntiles;
  int dx = t % ntx - o % ntx, dy = (t / ntx) % nty - (o / ntx) % nty,
      dz = t / (ntx * nty) - o / (ntx * nty);
  if (dx == ntx - 1) dx = -1;
  if (dx == 1 - ntx) dx = 1;
  if (dy == nty - 1) dy = -1;
  if (dy == 1 - nty) dy = 1;
  if (dz == ntz - 1) dz = -1;
  if (dz == 1 - ntz) dz = 1;
  if (dx < -1 || dx > 1 || dy < -1 || dy > 1 || dz < -1 || dz > 1) return kFar;
  return (dz + 1) * 9 + (dy + 1) * 3 + (dx + 1);
}

__device__ __forceinline__ int neighbour_bucket(int b_own, int code, int ntx, int nty, int ntz,
                                                int ntiles) {
  int s = b_own / ntiles, o = b_own % ntiles;
  int tx = wrapc(o % ntx + code % 3 - 1, ntx);
  int ty = wrapc((o / ntx) % nty + (code / 3) % 3 - 1, nty);
  int tz = wrapc(o / (ntx * nty) + code / 9 - 1, ntz);
  return s * ntiles + (tz * nty + ty) * ntx + tx;
}

// slot of an arrival: a run at the front of bucket dest, or the spill region
__device__ __forceinline__ int64_t arrival_slot(const FxArgs &a, int dest, int pos) {
  int room = a.cap_out[dest] - a.count_in[dest];
  if (pos < room) return a.start_out[dest] + pos;
  int s = dest / a.ntiles;
  int q = atomicAdd(&a.spill_out[s], 1);
  if (q >= a.spill_cap) {
    atomicOr(a.flag, kFlagSpillLost);
    return -1;
  }
  return a.area + (int64_t)s * a.spill_cap + q;
}

// One particle.  The common case is straight-line code (LDS gather, LDS charge
// tile, staying in the own bucket); the rare cases (outside the staged tile,
// beyond the charge halo, outbox full) sit behind wave-uniform branches taken
// only when some lane of the wave needs them.  Every lane stores exactly once
// (lanes without a slot store to the trash slot), so no path leaves a store
// count unknown to the wait-count analysis and the next particle's loads stay
// in flight.
template <int T, bool PUSH, bool DEPOSIT>
__device__ __forceinline__ void fx_one(const FxArgs &a, int s, int b_own, bool act, Pcl p, int ox,
                                       int oy, int oz, bool lds_ok, int64_t own_end, int *n_stay,
                                       int *n_out, OutRec *outbox, const float4 *fA,
                                       const float2 *fB, unsigned long long *rh, double ke[3]) {
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
    float f[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    int lx = ix - ox, ly = iy - oy, lz = iz - oz;
    bool in_tile = lds_ok && (unsigned)lx < (unsigned)T && (unsigned)ly < (unsigned)T &&
                   (unsigned)lz < (unsigned)T;
    if (!in_tile) {
      lx = 0;
      ly = 0;
      lz = 0;
    }
    if (lds_ok) gather_lds<T>(fA, fB, lx, ly, lz, fx, fy, fz, f);
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
  move_axis(fx, ix, vx, a.gstep, g.nx, a.flag, act);
  move_axis(fy, iy, vy, a.gstep, g.ny, a.flag, act);
  move_axis(fz, iz, vz, a.gstep, g.nz, a.flag, act);
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
  // ---- re-bucketing
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
    if (stay) dst = own_end - 1 - (base + __popcll(m & lt));
  }
  int cell = pack_cell(ix, iy, iz);
  uint64_t mv = __ballot(act && !stay);
  if (mv) {
    int leader = __ffsll((unsigned long long)mv) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(n_out, __popcll(mv));
    base = __shfl(base, leader);
    int k = base + __popcll(mv & lt);
    bool boxed = act && !stay && k < kOutbox;
    if (boxed) {
      OutRec r;
      r.dest = dest;
      r.cell = cell;
      r.pid = p.pid;
      r.code = 0;
      r.rank = 0;
      r.fx = fx;
      r.fy = fy;
      r.fz = fz;
      r.vx = vx;
      r.vy = vy;
      r.vz = vz;
      outbox[k] = r;
    }
    if (__ballot(act && !stay && !boxed)) {  // outbox full: place directly
      if (act && !stay && !boxed) {
        int64_t d2 = arrival_slot(a, dest, atomicAdd(&a.back[dest], 1));
        if (d2 >= 0) store_pcl(a.out, d2, cell, p.pid, fx, fy, fz, vx, vy, vz);
      }
    }
  }
  store_pcl(a.out, dst, cell, p.pid, fx, fy, fz, vx, vy, vz);
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
  __shared__ OutRec outbox[kOutbox];
  __shared__ int n_stay, n_out;
  __shared__ int hist[kFar + 1], runbase[kFar];
  __shared__ int64_t runstart[kFar];
  __shared__ int runroom[kFar];
  double ke[3] = {0.0, 0.0, 0.0};
  const int bid = blockIdx.x;
  int part = bid;  // energy partial slot: bucket index, then spill blocks
  int s = 0, b = -1, ox = 0, oy = 0, oz = 0;
  if (threadIdx.x == 0) {
    n_stay = 0;
    n_out = 0;
  }
  if (threadIdx.x <= kFar) hist[threadIdx.x] = 0;
  if (bid < a.nb) {
    b = bucket_of_block(bid, a.nb);
    part = b;
    s = b / a.ntiles;
    const int t = b - s * a.ntiles;
    ox = (t % a.ntx) * T;
    oy = ((t / a.ntx) % a.nty) * T;
    oz = (t / (a.ntx * a.nty)) * T;
    const int cnt = a.count_in[b];
    const int narr = a.narr_in[b];
    const int64_t beg = a.start_in[b];
    const int64_t gap = a.cap_in[b] - cnt;
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
      const int64_t own_end = a.start_out[b] + a.cap_out[b];
      // software pipelined: the next particle's state is in flight while the
      // current one is pushed (the load index is clamped, so every lane loads)
      Pcl cur, nxt;
      {
        int i = min((int)threadIdx.x, cnt - 1);
        load_pcl(a.in, beg + i + (i >= narr ? gap : 0), cur);
      }
      for (int base = 0; base < cnt; base += kBlock) {
        int i = min(base + kBlock + (int)threadIdx.x, cnt - 1);
        load_pcl(a.in, beg + i + (i >= narr ? gap : 0), nxt);
        fx_one<T, PUSH, DEPOSIT>(a, s, b, base + (int)threadIdx.x < cnt, cur, ox, oy, oz, true,
                                 own_end, &n_stay, &n_out, outbox, fA, fB, rh, ke);
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
      if (i < cnt) load_pcl(a.in, beg + i, cur);
      fx_one<T, PUSH, DEPOSIT>(a, s, -1, i < cnt, cur, 0, 0, 0, false, 0, &n_stay, &n_out, outbox,
                               fA, fB, rh, ke);
    }
  }
  // ---- place the outbox: one reserved run per destination bucket
  __syncthreads();
  const int nout = min(n_out, kOutbox);
  for (int q = threadIdx.x; q < nout; q += kBlock) {
    int c = dest_code(outbox[q].dest, b, a.ntx, a.nty, a.ntz, a.ntiles);
    outbox[q].code = c;
    if (c < kFar) outbox[q].rank = atomicAdd(&hist[c], 1);
  }
  __syncthreads();
  if (threadIdx.x < kFar && hist[threadIdx.x] > 0) {
    int d = neighbour_bucket(b, threadIdx.x, a.ntx, a.nty, a.ntz, a.ntiles);
    runbase[threadIdx.x] = atomicAdd(&a.back[d], hist[threadIdx.x]);
    runstart[threadIdx.x] = a.start_out[d];
    runroom[threadIdx.x] = a.cap_out[d] - a.count_in[d];
  }
  if (threadIdx.x == 0 && b >= 0) a.front[b] = n_stay;
  __syncthreads();
  for (int q = threadIdx.x; q < nout; q += kBlock) {
    OutRec r = outbox[q];
    int64_t dst;
    if (r.code < kFar) {
      int pos = runbase[r.code] + r.rank;
      dst = pos < runroom[r.code] ? runstart[r.code] + pos : arrival_slot(a, r.dest, pos);
    } else {
      dst = arrival_slot(a, r.dest, atomicAdd(&a.back[r.dest], 1));
    }
    if (dst >= 0) store_pcl(a.out, dst, r.cell, r.pid, r.fx, r.fy, r.fz, r.vx, r.vy, r.vz);
  }
  if (PUSH) block_reduce3(ke, a.ke_part + 3 * (int64_t)part);
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
    const int narr = a.narr_in[b];
    const int64_t beg = a.start_in[b];
    const int64_t gap = a.cap_in[b] - cnt;
    if (cnt == 0) return;
    for (int q = threadIdx.x; q < N3 * N3 * N3; q += kBlock) rh[q] = 0ull;
    __syncthreads();
    unsigned long long *r = a.rho + (int64_t)s * a.ncell;
    for (int i = threadIdx.x; i < cnt; i += kBlock) {
      int64_t src = beg + i + (i >= narr ? gap : 0);
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

// output capacities of a step from the input counts; clears the spill counters
__global__ void k_fx_caps(const int32_t *count_in, int32_t *cap_out, int32_t *spill_out, int nb,
                          int nspec) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < nb) {
    int c = count_in[b];
    cap_out[b] = c + c / 4 + 64;
  }
  if (b < nspec) spill_out[b] = 0;
}

__global__ void k_fill_i32(int32_t *p, int n, int v) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < n) p[b] = v;
}

// after a step: the output buckets' run lengths (arrivals that did not fit
// were sent to the spill region), spill counters clamped
__global__ void k_fx_finish(const int32_t *front, const int32_t *back, const int32_t *count_in,
                            const int32_t *cap_out, int32_t *narr_out, int32_t *count_out, int nb,
                            int32_t *spill, int nspec, int32_t spill_cap) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < nb) {
    int na = min(back[b], cap_out[b] - count_in[b]);
    narr_out[b] = na;
    count_out[b] = na + front[b];
  }
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
  const int32_t *bcount, *bnarr, *bcap;
  int64_t seg_start[EMG_MAX_SPECIES];
  const int32_t *seg_count_dev;        // device counts (spill) or nullptr
  int64_t seg_count[EMG_MAX_SPECIES];  // host counts (flat)
};

__device__ __forceinline__ void range_of(const RangeDesc &d, int r, int64_t &beg, int64_t &cnt,
                                         int64_t &narr, int64_t &gap, int &s) {
  if (r < d.nrb) {
    beg = d.bstart[r];
    cnt = d.bcount[r];
    narr = d.bnarr[r];
    gap = d.bcap[r] - cnt;
    s = r / d.ntiles;
  } else {
    s = r - d.nrb;
    beg = d.seg_start[s];
    cnt = d.seg_count_dev ? (int64_t)d.seg_count_dev[s] : d.seg_count[s];
    narr = cnt;
    gap = 0;
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
  int64_t beg, cnt, narr, gap;
  int s;
  range_of(d, r, beg, cnt, narr, gap, s);
  for (int64_t i = (int64_t)sub * kBlock + threadIdx.x; i < cnt; i += (int64_t)d.rb * kBlock) {
    int b = bucket_of_cell<T>(in.cell[beg + i + (i >= narr ? gap : 0)], s, ntx, nty, d.ntiles);
    atomicAdd(&hist[b], 1);
  }
}

template <int T>
__global__ __launch_bounds__(kBlock) void k_fx_scatter(RangeDesc d, StateFx in, StateFx out,
                                                       int ntx, int nty, const int64_t *start,
                                                       int32_t *cursor) {
  int r = blockIdx.x / d.rb, sub = blockIdx.x % d.rb;
  int64_t beg, cnt, narr, gap;
  int s;
  range_of(d, r, beg, cnt, narr, gap, s);
  for (int64_t i = (int64_t)sub * kBlock + threadIdx.x; i < cnt; i += (int64_t)d.rb * kBlock) {
    int64_t src = beg + i + (i >= narr ? gap : 0);
    int cell = in.cell[src];
    int b = bucket_of_cell<T>(cell, s, ntx, nty, d.ntiles);
    int64_t dst = start[b] + atomicAdd(&cursor[b], 1);
    store_pcl(out, dst, cell, in.pid ? in.pid[src] : 0, in.fx[src], in.fy[src], in.fz[src],
              in.vx[src], in.vy[src], in.vz[src]);
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
  store_pcl(out, m, pack_cell(c[0], c[1], c[2]), s64.pid[m], f[0], f[1], f[2], (float)s64.vx[m],
            (float)s64.vy[m], (float)s64.vz[m]);
}

// bucketed FP32 -> FP64 species-contiguous (dst offset = flat scan of counts)
__global__ __launch_bounds__(kBlock) void k_fx_to64(StateFx in, const int64_t *bstart,
                                                    const int32_t *bcount, const int32_t *bnarr,
                                                    const int32_t *bcap, const int64_t *flat,
                                                    int ntiles, SpeciesTab sp, State64 out) {
  int b = blockIdx.x;
  int s = b / ntiles;
  int cnt = bcount[b], narr = bnarr[b];
  int64_t beg = bstart[b], gap = bcap[b] - cnt;
  int64_t base = sp.off[s] + (flat[b] - flat[s * ntiles]);
  for (int i = threadIdx.x; i < cnt; i += kBlock) {
    int64_t src = beg + i + (i >= narr ? gap : 0), dst = base + i;
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
                                                      const int32_t *bcount, const int32_t *bnarr,
                                                      const int32_t *bcap, int nb,
                                                      const int32_t *spill, int64_t area,
                                                      int64_t spill_cap, double *part) {
  double ke[3] = {0.0, 0.0, 0.0};
  int bid = blockIdx.x;
  int64_t beg, gap = 0;
  int cnt, narr;
  if (bid < nb) {
    beg = bstart[bid];
    cnt = bcount[bid];
    narr = bnarr[bid];
    gap = bcap[bid] - cnt;
  } else {
    int s = bid - nb;
    beg = area + (int64_t)s * spill_cap;
    cnt = spill[s];
    narr = cnt;
  }
  for (int i = threadIdx.x; i < cnt; i += kBlock) {
    int64_t src = beg + i + (i >= narr ? gap : 0);
    float vx = in.vx[src], vy = in.vy[src], vz = in.vz[src];
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
  int in = B.cur, out = 1 - B.cur;
  FxArgs a;
  a.g = c->g;
  a.ntx = B.ntx;
  a.nty = B.nty;
  a.ntz = B.ntz;
  a.ntiles = B.ntiles;
  a.nb = B.nb;
  a.nspec = c->sp.nspec;
  a.in = B.st[in];
  a.out = B.st[out];
  a.start_in = B.start[in];
  a.count_in = B.count[in];
  a.narr_in = B.narr[in];
  a.cap_in = B.cap[in];
  a.spill_in = B.spill[in];
  a.start_out = B.start[out];
  a.cap_out = B.cap[out];
  a.front = B.front;
  a.back = B.back;
  a.spill_out = B.spill[out];
  a.area = B.area;
  a.spill_cap = B.spill_cap;
  a.trash = B.trash;
  a.node32 = c->node32;
  a.rho = c->rhofx;
  a.ncell = c->ncell;
  for (int s = 0; s < EMG_MAX_SPECIES; s++) a.qm[s] = (float)c->cfg.qm[s];
  a.gstep = 0.f;
  a.ke_part = c->ke_part;
  a.flag = c->dflag;
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

// bucket the particles described by d (reading buffer 'src') into buffer 1 - src
int bucket_ranges(emg_ctx *c, RangeDesc d, int nr, int rb, int src) {
  FxBuckets &B = c->fx;
  int dst = 1 - src;
  d.rb = rb;
  unsigned grid = (unsigned)((int64_t)nr * rb);
  unsigned gb = (unsigned)((B.nb + kBlock - 1) / kBlock);
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
  hipLaunchKernelGGL(k_fx_caps, dim3(gb), dim3(kBlock), 0, c->stream, B.hist, B.cap[dst],
                     B.spill[dst], B.nb, c->sp.nspec);
  hipLaunchKernelGGL(k_fill_i32, dim3(gb), dim3(kBlock), 0, c->stream, B.cursor, B.nb, 0);
  exclusive_scan_i32(B.cap[dst], B.start[dst], B.nb, c->scan_tmp, c->stream);
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
  // all particles of a bucket form one run from its start
  hipLaunchKernelGGL(k_copy_i32, dim3(gb), dim3(kBlock), 0, c->stream, B.cursor, B.nb, B.count[dst]);
  hipLaunchKernelGGL(k_copy_i32, dim3(gb), dim3(kBlock), 0, c->stream, B.cursor, B.nb, B.narr[dst]);
  B.cur = dst;
  return hip_check(hipGetLastError(), "bucketing");
}

RangeDesc bucket_desc(emg_ctx *c) {
  FxBuckets &B = c->fx;
  RangeDesc d{};
  d.nrb = B.nb;
  d.nspec = c->sp.nspec;
  d.ntiles = B.ntiles;
  d.bstart = B.start[B.cur];
  d.bcount = B.count[B.cur];
  d.bnarr = B.narr[B.cur];
  d.bcap = B.cap[B.cur];
  d.seg_count_dev = B.spill[B.cur];
  for (int s = 0; s < c->sp.nspec; s++) d.seg_start[s] = B.area + (int64_t)s * B.spill_cap;
  return d;
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
        hipMalloc(&B.narr[k], sizeof(int32_t) * B.nb) != hipSuccess ||
        hipMalloc(&B.cap[k], sizeof(int32_t) * B.nb) != hipSuccess ||
        hipMalloc(&B.spill[k], sizeof(int32_t) * EMG_MAX_SPECIES) != hipSuccess)
      return set_error(EMG_ERR_NOMEM, "bucket tables");
    launch_zero(B.count[k], sizeof(int32_t) * B.nb, c->stream);
    launch_zero(B.narr[k], sizeof(int32_t) * B.nb, c->stream);
    launch_zero(B.cap[k], sizeof(int32_t) * B.nb, c->stream);
    launch_zero(B.spill[k], sizeof(int32_t) * EMG_MAX_SPECIES, c->stream);
    launch_zero(B.start[k], sizeof(int64_t) * (B.nb + 1), c->stream);
  }
  if (hipMalloc(&B.front, sizeof(int32_t) * B.nb) != hipSuccess ||
      hipMalloc(&B.back, sizeof(int32_t) * B.nb) != hipSuccess ||
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
  for (int s = 0; s < c->sp.nspec; s++) {
    d.seg_start[s] = c->sp.off[s];
    d.seg_count[s] = c->sp.cnt[s];
  }
  return bucket_ranges(c, d, c->sp.nspec, 2048, flat);
}

// bucketed FP32 (two runs per bucket, spill) -> one run per bucket, no spill
int fx_rebucket(emg_ctx *c) {
  return bucket_ranges(c, bucket_desc(c), c->fx.nb + c->sp.nspec, 1, c->fx.cur);
}

int fx_to64(emg_ctx *c, State64 &dst) {
  FxBuckets &B = c->fx;
  // spill regions are absorbed first so every particle sits in a bucket
  int rc = fx_rebucket(c);
  if (rc) return rc;
  exclusive_scan_i32(B.count[B.cur], B.start[1 - B.cur], B.nb, c->scan_tmp, c->stream);
  hipLaunchKernelGGL(k_fx_to64, dim3(B.nb), dim3(kBlock), 0, c->stream, B.st[B.cur],
                     B.start[B.cur], B.count[B.cur], B.narr[B.cur], B.cap[B.cur],
                     B.start[1 - B.cur], B.ntiles, c->sp, dst);
  return hip_check(hipGetLastError(), "fx_to64");
}

void launch_fx_step(emg_ctx *c, int ustep, bool push, bool deposit, bool em) {
  (void)em;
  FxBuckets &B = c->fx;
  int in = B.cur, out = 1 - B.cur;
  unsigned gb = (unsigned)((B.nb + kBlock - 1) / kBlock);
  hipLaunchKernelGGL(k_fx_caps, dim3(gb), dim3(kBlock), 0, c->stream, B.count[in], B.cap[out],
                     B.spill[out], B.nb, c->sp.nspec);
  hipLaunchKernelGGL(k_fill_i32, dim3(gb), dim3(kBlock), 0, c->stream, B.back, B.nb, 0);
  exclusive_scan_i32(B.cap[out], B.start[out], B.nb, c->scan_tmp, c->stream);
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
  hipLaunchKernelGGL(k_fx_finish, dim3(gb), dim3(kBlock), 0, c->stream, B.front, B.back,
                     B.count[in], B.cap[out], B.narr[out], B.count[out], B.nb, B.spill[out],
                     c->sp.nspec, (int32_t)B.spill_cap);
  B.cur = out;
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
                     B.st[B.cur], B.start[B.cur], B.count[B.cur], B.narr[B.cur], B.cap[B.cur], B.nb,
                     B.spill[B.cur], B.area, B.spill_cap, c->ke_part);
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
