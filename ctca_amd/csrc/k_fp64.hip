// k_fp64.hip -- plain FP64 kernels of the particle path: the drop-in form that
// mirrors the reference arithmetic operation by operation.  Compiled with
// -ffp-contract=off, so every multiply and add is rounded separately in the
// reference's source order (and so bit-identical to a CPU restatement built the
// same way).  One thread per particle, species segments on blockIdx.y, state
// in place (structure of arrays), deposition with FP64 global atomics into
// periodic accumulators (node index mod n == the single-rank fold).
#include "emg_internal.h"

namespace emg {
namespace {

__device__ __forceinline__ int64_t nidx64(const Grid &g, int i, int j, int k) {
  // nodal FP64 field, nodes 0..n per axis
  return ((int64_t)k * (g.ny + 1) + j) * (g.nx + 1) + i;
}

__device__ __forceinline__ int wrapn(int i, int n) {
  // node / cell index in [-n, 2n) -> [0, n)
  return i < 0 ? i + n : (i >= n ? i - n : i);
}

__device__ __forceinline__ int64_t cidx(const Grid &g, int i, int j, int k) {
  return ((int64_t)wrapn(k, g.nz) * g.ny + wrapn(j, g.ny)) * g.nx + wrapn(i, g.nx);
}

// Guard against positions the reference would index out of bounds with:
// anything outside [0, L] or non-finite is clamped and reported.
__device__ __forceinline__ void domain_guard(double &x, double &y, double &z, const Grid &g,
                                             int *flag) {
  bool ok = (x >= 0.0) && (x <= g.slx) && (y >= 0.0) && (y <= g.sly) && (z >= 0.0) &&
            (z <= g.slz);
  if (!ok) {
    atomicOr(flag, kFlagDomain);
    x = (x >= 0.0) ? (x <= g.slx ? x : g.slx) : 0.0;
    y = (y >= 0.0) ? (y <= g.sly ? y : g.sly) : 0.0;
    z = (z >= 0.0) ? (z <= g.slz ? z : g.slz) : 0.0;
  }
}

struct Cic {
  double v1, v2, v3, v4, v5, v6, v7, v8;
  int i, j, k;
};

// CIC weights, src/pcl2field/psolve2.F90:154-183 (charge.F90:160-193).
// xl, yl, zl are the cell-local coordinates, i/j/k the already truncated cell.
__device__ __forceinline__ Cic cic(double xl, double yl, double zl, int i, int j, int k) {
  Cic w;
  w.i = i;
  w.j = j;
  w.k = k;
  double x1 = xl - (double)i, y1 = yl - (double)j, z1 = zl - (double)k;
  double xy1 = x1 * y1, xz1 = x1 * z1, yz1 = y1 * z1;
  double z2 = 1.0 - z1;
  double xz2 = x1 * z2, yz2 = y1 * z2;
  w.v3 = xy1 * z1;
  w.v2 = xz1 - w.v3;
  w.v4 = yz1 - w.v3;
  w.v1 = z1 - xz1 - w.v4;
  w.v7 = xy1 * z2;
  w.v6 = xz2 - w.v7;
  w.v8 = yz2 - w.v7;
  w.v5 = z2 - xz2 - w.v8;
  return w;
}

// psolve2.F90:195-218: eight-node gather of the six nodal components in the
// reference's node order v1..v8 and summation order.
__device__ __forceinline__ void gather6(const double *__restrict__ node, const Grid &g,
                                        const Cic &w, double f[6]) {
  int i1 = min(w.i + 1, g.nx), j1 = min(w.j + 1, g.ny), k1 = min(w.k + 1, g.nz);
  const double *p1 = node + 8 * nidx64(g, w.i, w.j, k1);
  const double *p2 = node + 8 * nidx64(g, i1, w.j, k1);
  const double *p3 = node + 8 * nidx64(g, i1, j1, k1);
  const double *p4 = node + 8 * nidx64(g, w.i, j1, k1);
  const double *p5 = node + 8 * nidx64(g, w.i, w.j, w.k);
  const double *p6 = node + 8 * nidx64(g, i1, w.j, w.k);
  const double *p7 = node + 8 * nidx64(g, i1, j1, w.k);
  const double *p8 = node + 8 * nidx64(g, w.i, j1, w.k);
#pragma unroll
  for (int c = 0; c < 6; c++)
    f[c] = p1[c] * w.v1 + p2[c] * w.v2 + p3[c] * w.v3 + p4[c] * w.v4 + p5[c] * w.v5 +
           p6[c] * w.v6 + p7[c] * w.v7 + p8[c] * w.v8;
}

// Buneman-Boris, psolve2.F90:233-257 (ewmodel = 0: the wave field is +0.0).
__device__ __forceinline__ void boris64(double &vx, double &vy, double &vz, const double f[6],
                                        double qmp) {
  const double ewx = 0.0, ewy = 0.0;
  double eex = (f[0] + ewx) * qmp;
  double eey = (f[1] + ewy) * qmp;
  double eez = f[2] * qmp;
  double bbx = f[3] * qmp, bby = f[4] * qmp, bbz = f[5] * qmp;
  double bor = 2.0 / (1.0 + bbx * bbx + bby * bby + bbz * bbz);
  vx = vx + eex;
  vy = vy + eey;
  vz = vz + eez;
  double vxt = vx + vy * bbz - vz * bby;
  double vyt = vy + vz * bbx - vx * bbz;
  double vzt = vz + vx * bby - vy * bbx;
  vx = vx + bor * (vyt * bbz - vzt * bby);
  vy = vy + bor * (vzt * bbx - vxt * bbz);
  vz = vz + bor * (vxt * bby - vyt * bbx);
  vx = vx + eex;
  vy = vy + eey;
  vz = vz + eez;
}

// psolve2.F90:354-432 / positn.F90:172-254, npbnd = 0: one wrap per axis
__device__ __forceinline__ void wrap_periodic(double &x, double &y, double &z, const Grid &g,
                                              int *flag) {
  if (x < 0.0 || x >= g.slx || y < 0.0 || y >= g.sly || z < 0.0 || z >= g.slz) {
    if (x < 0.0) x = x + g.slx; else if (x >= g.slx) x = x - g.slx;
    if (y < 0.0) y = y + g.sly; else if (y >= g.sly) y = y - g.sly;
    if (z < 0.0) z = z + g.slz; else if (z >= g.slz) z = z - g.slz;
    if (!(x >= 0.0 && x <= g.slx && y >= 0.0 && y <= g.sly && z >= 0.0 && z <= g.slz))
      atomicOr(flag, kFlagFar);
  }
}

// ---------------------------------------------------------------- grdcrt
// grdcrt.F90:158-169 (E) and 216-230 (B); writes nodes 0..n (FP64) and, when
// node32 != nullptr, the periodic FP32 copy (nodes 0..n-1) of the same values.
__global__ __launch_bounds__(kBlock) void k_relocate(const double *__restrict__ eb, int elx,
                                                     int ely, int elz, int exx, int exy,
                                                     double *__restrict__ node64,
                                                     float *__restrict__ node32, Grid g,
                                                     double e0x, double e0y, double e0z,
                                                     double b0x, double b0y, double b0z) {
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t nn = (int64_t)(g.nx + 1) * (g.ny + 1) * (g.nz + 1);
  if (idx >= nn) return;
  int i = (int)(idx % (g.nx + 1));
  int j = (int)((idx / (g.nx + 1)) % (g.ny + 1));
  int k = (int)(idx / ((int64_t)(g.nx + 1) * (g.ny + 1)));
  auto E = [&](int c, int ii, int jj, int kk) -> double {
    return eb[(((int64_t)(kk - elz) * exy + (jj - ely)) * exx + (ii - elx)) * 6 + c];
  };
  double o[6];
  o[0] = 0.5 * (E(0, i, j, k) + E(0, i - 1, j, k)) + e0x;
  o[1] = 0.5 * (E(1, i, j, k) + E(1, i, j - 1, k)) + e0y;
  o[2] = 0.5 * (E(2, i, j, k) + E(2, i, j, k - 1)) + e0z;
  o[3] = 0.25 * (E(3, i, j, k) + E(3, i, j - 1, k) + E(3, i, j, k - 1) + E(3, i, j - 1, k - 1)) + b0x;
  o[4] = 0.25 * (E(4, i, j, k) + E(4, i - 1, j, k) + E(4, i, j, k - 1) + E(4, i - 1, j, k - 1)) + b0y;
  o[5] = 0.25 * (E(5, i, j, k) + E(5, i - 1, j, k) + E(5, i, j - 1, k) + E(5, i - 1, j - 1, k)) + b0z;
  if (node64) {
    double *d = node64 + 8 * idx;
#pragma unroll
    for (int c = 0; c < 6; c++) d[c] = o[c];
    d[6] = 0.0;
    d[7] = 0.0;
  }
  if (node32 && i < g.nx && j < g.ny && k < g.nz) {
    float4 *d = reinterpret_cast<float4 *>(node32 + 8 * (((int64_t)k * g.ny + j) * g.nx + i));
    d[0] = make_float4((float)o[0], (float)o[1], (float)o[2], (float)o[3]);
    d[1] = make_float4((float)o[4], (float)o[5], 0.f, 0.f);
  }
}

// ---------------------------------------------------------------- positn
__global__ __launch_bounds__(kBlock) void k_positn64(State64 s, SpeciesTab sp, Grid g, int ustep,
                                                     int *flag) {
  int is = blockIdx.y;
  int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= sp.cnt[is]) return;
  m += sp.off[is];
  double us = (double)ustep;
  double x = s.x[m], y = s.y[m], z = s.z[m];
  x = x + s.vx[m] * us;
  y = y + s.vy[m] * us;
  z = z + s.vz[m] * us;
  wrap_periodic(x, y, z, g, flag);
  s.x[m] = x;
  s.y[m] = y;
  s.z[m] = z;
}

// ---------------------------------------------------------------- psolve2 + charge
// ---------------------------------------------------------------- charge
// charge.F90:59-83 + distribute_charge 197-210; rho[is][node mod n] += v_k q(is)
// charge.F90:152-210 per particle; the particles of a block are a contiguous
// run of a species (kDepPer per thread).  The sort orders particles by cell
// and they drift slowly, so most of a block's particles lie in a box of
// nodes: the block sums their weights in an LDS window of that box, as int64
// fixed point at 2^48 per unit weight (at most kBlock kDepPer terms of at
// most 2^48 per node, below 2^61), and adds each non-zero window node
// times q(is) to rho with one global FP64 atomic; a particle whose eight
// nodes are not all in the window adds its weights to rho directly.  The
// rounding of a weight (2^-49) is below the FP64 round-off of the sequential
// sum, which the reference's order would give.  The window's shape:
// run_window; the block's particles: block_run.
constexpr int kDepPer = 16;
constexpr int kWinCap = 3840;                    // 30 KB of int64
constexpr int kWinX = 80, kWinY = 8, kWinZ = 6;  // the fixed box
constexpr double kRhoFix = 281474976710656.0;   // 2^48 per unit weight

// window index of node l + 1 on an axis of window extent w (full axis: w = n,
// where node n is node 0); false if it is outside the window
__device__ __forceinline__ bool win_next(int l, int w, int n, int &l1) {
  l1 = l + 1;
  if (l1 < w) return true;
  l1 = 0;
  return w == n && l + 1 == n;
}

// widen one axis of the window by a node on each side (or to the whole axis)
__device__ __forceinline__ void win_grow(int &o, int &w, int n) {
  if (w + 2 >= n) {
    o = 0;
    w = n;
  } else {
    o = o == 0 ? n - 1 : o - 1;
    w += 2;
  }
}

// The LDS window of a block's run of particles a .. b (sorted by cell): nodes
// o .. o + w - 1 per axis (periodic), w <= n, at most cap nodes.  The cells of
// a sorted run go from its first to its last particle's cell in z-major order
// (rows of the y-z plane one after another), so its nodes fit a box of its
// rows plus one, its layers plus one, and in x the row segment it covers (or,
// for a run that passes into the next row, the periodic range from its first
// cell to its last) or whole rows; the rest of the capacity becomes margins
// in z, y and x for the particles that have drifted since the sort.  A run of
// more rows than a layer holds, or a box over the capacity, gets the fixed
// box fx x fy x fz around its middle particle.
__device__ __forceinline__ void run_window(const State64 &s, int64_t a, int64_t b, const Grid &g,
                                           int cap, int fx, int fy, int fz, int &wx, int &wy,
                                           int &wz, int &ox, int &oy, int &oz) {
  auto cell = [&](double v, int n) { return min(max((int)v, 0), n - 1); };
  const int ia = cell(s.x[a], g.nx), ja = cell(s.y[a], g.ny), ka = cell(s.z[a], g.nz);
  const int ib = cell(s.x[b], g.nx), jb = cell(s.y[b], g.ny), kb = cell(s.z[b], g.nz);
  // the run's rows, as consecutive rows of the periodic y-z plane
  const int64_t ra = (int64_t)ka * g.ny + ja, rb = (int64_t)kb * g.ny + jb;
  const int64_t nrow = (rb - ra + (int64_t)g.ny * g.nz) % ((int64_t)g.ny * g.nz) + 1;
  bool boxed = nrow <= g.ny;
  if (boxed) {
    // x: the row segment ia..ib, or (a run into the next row) ia..nx-1 and
    // 0..ib taken as one periodic range; whole rows for longer runs
    if (nrow == 1 && ib >= ia) {
      ox = ia;
      wx = ib - ia + 2;
    } else if (nrow == 2 && ib < ia) {
      ox = ia;
      wx = ib + g.nx - ia + 2;
    } else {
      ox = 0;
      wx = g.nx;
    }
    // y: rows ja .. jb (periodic); z: layers ka .. kb
    oy = ja;
    wy = (int)min<int64_t>(nrow + 1, g.ny);
    oz = ka;
    wz = (kb - ka + g.nz) % g.nz + 2;
    if (wx >= g.nx) {
      ox = 0;
      wx = g.nx;
    }
    if (wy >= g.ny) {
      oy = 0;
      wy = g.ny;
    }
    if (wz >= g.nz) {
      oz = 0;
      wz = g.nz;
    }
    boxed = (int64_t)wx * wy * wz <= cap;
  }
  if (boxed) {  // margins while they fit: z, y, x in turn
    int o[3] = {ox, oy, oz}, w[3] = {wx, wy, wz};
    const int n[3] = {g.nx, g.ny, g.nz};
    for (int d = 2, stuck = 0; stuck < 3; d = d == 0 ? 2 : d - 1) {
      int o2 = o[d], w2 = w[d];
      bool grew = false;
      if (w2 < n[d]) {
        win_grow(o2, w2, n[d]);
        if ((int64_t)w2 * w[(d + 1) % 3] * w[(d + 2) % 3] <= cap) {
          o[d] = o2;
          w[d] = w2;
          grew = true;
        }
      }
      stuck = grew ? 0 : stuck + 1;
    }
    ox = o[0];
    oy = o[1];
    oz = o[2];
    wx = w[0];
    wy = w[1];
    wz = w[2];
  } else {  // the fixed box around the middle particle
    wx = min(fx, g.nx);
    wy = min(fy, g.ny);
    wz = min(fz, g.nz);
    const int64_t mid = a + (b - a) / 2;
    ox = wrapn(cell(s.x[mid], g.nx) - wx / 2 + g.nx, g.nx);
    oy = wrapn(cell(s.y[mid], g.ny) - wy / 2 + g.ny, g.ny);
    oz = wrapn(cell(s.z[mid], g.nz) - wz / 2 + g.nz, g.nz);
  }
}

// The block's run of particles (indices in the species' segment): after a
// sort the run table's entry (at most `per` particles over a bounded range
// of cells), else the per-th slice of species blockIdx.y.  False for a block
// past the end.
__device__ __forceinline__ bool block_run(const SpeciesTab &sp, const Run *runs, int per, int &is,
                                          int64_t &beg, int64_t &end) {
  if (runs) {
    const Run r = runs[blockIdx.x];
    is = r.s;
    beg = r.beg;
    end = r.beg + r.len;
    return true;
  }
  is = blockIdx.y;
  beg = (int64_t)blockIdx.x * per;
  const int64_t cnt = sp.cnt[is];
  if (beg >= cnt) return false;
  end = min(beg + (int64_t)per, cnt);
  return true;
}

// One particle's CIC weights (charge.F90:152-210) at (x, y, z): into the
// block's LDS window when its eight nodes lie in it, else into the species'
// rho r directly.
struct Win {
  int wx, wy, wz, ox, oy, oz;
};

__device__ __forceinline__ void deposit_win(unsigned long long *win, const Win &W, double *r,
                                            const Grid &g, double x, double y, double z,
                                            double qc) {
  int i = (int)x, j = (int)y, k = (int)z;
  Cic w = cic(x, y, z, i, j, k);
  // window coordinates of the cell (cell n, of a position at L, is cell 0's image)
  int li = wrapn(i - W.ox + g.nx, g.nx), lj = wrapn(j - W.oy + g.ny, g.ny),
      lk = wrapn(k - W.oz + g.nz, g.nz);
  int li1, lj1, lk1;
  if (li < W.wx && lj < W.wy && lk < W.wz && win_next(li, W.wx, g.nx, li1) &&
      win_next(lj, W.wy, g.ny, lj1) && win_next(lk, W.wz, g.nz, lk1)) {
    auto at = [&](int a, int b, int c) { return win + (c * W.wy + b) * W.wx + a; };
    auto fx = [](double v) { return (unsigned long long)__double2ll_rn(v * kRhoFix); };
    atomicAdd(at(li, lj, lk1), fx(w.v1));
    atomicAdd(at(li1, lj, lk1), fx(w.v2));
    atomicAdd(at(li1, lj1, lk1), fx(w.v3));
    atomicAdd(at(li, lj1, lk1), fx(w.v4));
    atomicAdd(at(li, lj, lk), fx(w.v5));
    atomicAdd(at(li1, lj, lk), fx(w.v6));
    atomicAdd(at(li1, lj1, lk), fx(w.v7));
    atomicAdd(at(li, lj1, lk), fx(w.v8));
  } else {
    int i1 = i + 1, j1 = j + 1, k1 = k + 1;
    atomicAdd(r + cidx(g, i, j, k1), w.v1 * qc);
    atomicAdd(r + cidx(g, i1, j, k1), w.v2 * qc);
    atomicAdd(r + cidx(g, i1, j1, k1), w.v3 * qc);
    atomicAdd(r + cidx(g, i, j1, k1), w.v4 * qc);
    atomicAdd(r + cidx(g, i, j, k), w.v5 * qc);
    atomicAdd(r + cidx(g, i1, j, k), w.v6 * qc);
    atomicAdd(r + cidx(g, i1, j1, k), w.v7 * qc);
    atomicAdd(r + cidx(g, i, j1, k), w.v8 * qc);
  }
}

// the window's non-zero nodes into rho, one FP64 atomic each (after a barrier)
__device__ __forceinline__ void flush_win(const unsigned long long *win, const Win &W, double *r,
                                          const Grid &g, double qc) {
  for (int q = threadIdx.x; q < W.wx * W.wy * W.wz; q += kBlock) {
    const long long v = (long long)win[q];
    if (v != 0) {
      int li = q % W.wx, lj = (q / W.wx) % W.wy, lk = q / (W.wx * W.wy);
      atomicAdd(r + cidx(g, W.ox + li, W.oy + lj, W.oz + lk), (double)v * (qc / kRhoFix));
    }
  }
}

__global__ __launch_bounds__(kBlock) void k_deposit64(State64 s, SpeciesTab sp, Grid g,
                                                      double *__restrict__ rho, int64_t ncell,
                                                      const Run *__restrict__ runs, int *flag) {
  __shared__ unsigned long long win[kWinCap];
  int is;
  int64_t beg, end;
  if (!block_run(sp, runs, kBlock * kDepPer, is, beg, end)) return;  // uniform over the block
  const int64_t off = sp.off[is];
  // window: nodes o .. o + w - 1 per axis (w <= n, so no node appears twice)
  Win W;
  run_window(s, off + beg, off + end - 1, g, kWinCap, kWinX, kWinY, kWinZ, W.wx, W.wy, W.wz, W.ox,
             W.oy, W.oz);
  for (int q = threadIdx.x; q < W.wx * W.wy * W.wz; q += kBlock) win[q] = 0ull;
  __syncthreads();
  const double qc = sp.q[is];
  double *r = rho + (int64_t)is * ncell;
  for (int64_t m = beg + threadIdx.x; m < end; m += kBlock) {
    double x = s.x[off + m], y = s.y[off + m], z = s.z[off + m];
    domain_guard(x, y, z, g, flag);
    deposit_win(win, W, r, g, x, y, z, qc);
  }
  __syncthreads();
  flush_win(win, W, r, g, qc);
}

// psolve2 and charge in one pass (psolve2.F90:137-433, then charge.F90:
// 59-83): each particle of the block's run is pushed with the reference's
// arithmetic in its order -- int() cell, CIC weights, eight-node gather,
// Buneman-Boris, x += v ustep, periodic wrap -- and its charge deposited at
// the new position exactly as k_deposit64 would.
// The window is shaped from the run's positions before the move (read by
// every thread before any is written: the barrier after zeroing it); the
// particles that move out of it add to rho directly.
// RESORT: the moved particle (and its pid) goes to its sorted slot of dst.t
// instead of back to its own slot; the arithmetic is the same.  Four waves
// per SIMD (at most 128 VGPRs) for both forms: the plain one needs 122, the
// re-sorting one would take 132 and drop to three.
template <bool RESORT>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4)))
void k_push_dep64(State64 s, SpeciesTab sp, const double *__restrict__ node, Grid g, int ustep,
                  double *__restrict__ rho, int64_t ncell, const Run *__restrict__ runs, int *flag,
                  SortDest dst) {
  __shared__ unsigned long long win[kWinCap];
  int is;
  int64_t beg, end;
  if (!block_run(sp, runs, kBlock * kDepPer, is, beg, end)) return;  // uniform over the block
  const int64_t off = sp.off[is];
  const int64_t key0 = (int64_t)is * dst.nkey;
  const int64_t base = RESORT ? off - dst.start[key0] : 0;
  Win W;
  run_window(s, off + beg, off + end - 1, g, kWinCap, kWinX, kWinY, kWinZ, W.wx, W.wy, W.wz, W.ox,
             W.oy, W.oz);
  for (int q = threadIdx.x; q < W.wx * W.wy * W.wz; q += kBlock) win[q] = 0ull;
  __syncthreads();
  const double gustep = (double)ustep;
  const double qmp = sp.qm[is] * gustep * 0.5;
  const double qc = sp.q[is];
  double *r = rho + (int64_t)is * ncell;
  for (int64_t mm = beg + threadIdx.x; mm < end; mm += kBlock) {
    const int64_t m = off + mm;
    double x = s.x[m], y = s.y[m], z = s.z[m];
    double vx = s.vx[m], vy = s.vy[m], vz = s.vz[m];
    int64_t d = m;
    int32_t pid = 0;
    if (RESORT) {
      pid = s.pid[m];
      d = base + dst.start[key0 + slab_key_xyz(x, y, z, g)] + dst.rank[m];
    }
    domain_guard(x, y, z, g, flag);
    int i = (int)x, j = (int)y, k = (int)z;  // int(), psolve2.F90:158-160
    Cic w = cic(x, y, z, i, j, k);
    double f[6];
    gather6(node, g, w, f);
    boris64(vx, vy, vz, f, qmp);
    double xmove = vx * gustep, ymove = vy * gustep, zmove = vz * gustep;
    x = x + xmove;
    y = y + ymove;
    z = z + zmove;
    wrap_periodic(x, y, z, g, flag);
    const State64 &o = RESORT ? dst.t : s;
    o.x[d] = x;
    o.y[d] = y;
    o.z[d] = z;
    o.vx[d] = vx;
    o.vy[d] = vy;
    o.vz[d] = vz;
    if (RESORT) o.pid[d] = pid;
    domain_guard(x, y, z, g, flag);
    deposit_win(win, W, r, g, x, y, z, qc);
  }
  __syncthreads();
  flush_win(win, W, r, g, qc);
}

// ---------------------------------------------------------------- psolve1
// psolve1.F90:146-789: push with floor() cells, then the Umeda zigzag current
// of the two segments b -> relay -> a; J[is][c][cell mod n] += vv * w (q(is)
// is applied at read-out, psolve1.F90:793-802).
struct Relay {
  double la, a, r;
  int ia;
};

__device__ __forceinline__ Relay relay_axis(double xnew, double lb, double cb, int ib, int n,
                                            double sl, bool inside, double &pos) {
  Relay r;
  if (!inside && xnew < 0.0) {  // psolve1.F90:440-447
    r.la = xnew;
    pos = xnew + sl;
    r.ia = -1;
    r.a = -1.0;
    r.r = 0.0;
  } else if (!inside && xnew >= sl) {  // :462-469
    r.la = xnew - 0.0;
    pos = xnew - sl;
    r.ia = n;
    r.a = (double)n;
    r.r = (double)n;
  } else {  // :381-415, :469-478
    r.la = xnew - 0.0;
    pos = xnew;
    r.ia = (int)floor(r.la);
    r.a = (double)r.ia;
    r.r = (ib == r.ia) ? (lb + r.la) * 0.5 : fmax(cb, r.a);
  }
  return r;
}

// The zigzag sums of a block's particles go to an LDS window (three
// components of kEmCap nodes each, shaped from the block's sorted run like the
// charge window, or a fixed box around its middle particle) as int64 fixed
// point at 2^48 per unit, and reach the global FP64 sums once per window
// node.  A contribution is |vv w| <= 4 in the window (larger fluxes go
// straight to the global sums), so a node's window sum of at most
// 2 kBlock kEmPer terms stays below 2^61; the rounding of each term
// (2^-49) is far below the FP64 summation round-off of the reference's order.
constexpr int kEmPer = 4;
constexpr int kEmCap = 1600;                           // 3 x 12.8 KB of int64
constexpr int kEmWinX = 40, kEmWinY = 8, kEmWinZ = 5;  // the fixed box
constexpr double kEmFix = 281474976710656.0;          // 2^48

__device__ __forceinline__ long long em_fix(double v) { return __double2ll_rn(v * kEmFix); }

// RESORT as in k_push_dep64.
template <bool RESORT>
__global__ __launch_bounds__(kBlock) void k_push_em64(State64 s, SpeciesTab sp,
                                                      const double *__restrict__ node, Grid g,
                                                      int ustep, double *__restrict__ jacc,
                                                      int64_t ncell, const Run *__restrict__ runs,
                                                      int *flag, SortDest dst) {
  constexpr int WN = kEmCap;
  __shared__ unsigned long long win[3 * WN];
  int is;
  int64_t beg, end;
  if (!block_run(sp, runs, kBlock * kEmPer, is, beg, end)) return;  // uniform over the block
  const int64_t off = sp.off[is];
  const int64_t key0 = (int64_t)is * dst.nkey;
  const int64_t base = RESORT ? off - dst.start[key0] : 0;
  // window: nodes o .. o + w - 1 per axis (w <= n, so no node appears twice)
  int wx, wy, wz, ox, oy, oz;
  run_window(s, off + beg, off + end - 1, g, WN, kEmWinX, kEmWinY, kEmWinZ, wx, wy, wz, ox, oy,
             oz);
  const int wn = wx * wy * wz;  // nodes per component
  for (int q = threadIdx.x; q < 3 * wn; q += kBlock) win[q] = 0ull;
  __syncthreads();
  double gustep = (double)ustep;
  double rustep = 1.0 / gustep;
  double qmp = sp.qm[is] * gustep * 0.5;
  double *jx = jacc + ((int64_t)is * 3 + 0) * ncell;
  double *jy = jacc + ((int64_t)is * 3 + 1) * ncell;
  double *jz = jacc + ((int64_t)is * 3 + 2) * ncell;
  for (int64_t mm = beg + threadIdx.x; mm < end; mm += kBlock) {
    const int64_t m = off + mm;
    double x = s.x[m], y = s.y[m], z = s.z[m];
    double vx = s.vx[m], vy = s.vy[m], vz = s.vz[m];
    int64_t d = m;
    int32_t pid = 0;
    if (RESORT) {
      pid = s.pid[m];
      d = base + dst.start[key0 + slab_key_xyz(x, y, z, g)] + dst.rank[m];
    }
    domain_guard(x, y, z, g, flag);
    double xlb = x, ylb = y, zlb = z;
    int ib = (int)floor(xlb), jb = (int)floor(ylb), kb = (int)floor(zlb);  // :170-172
    double xb = ib, yb = jb, zb = kb;
    Cic w = cic(xlb, ylb, zlb, ib, jb, kb);
    double f[6];
    gather6(node, g, w, f);
    boris64(vx, vy, vz, f, qmp);
    double xmove = vx * gustep, ymove = vy * gustep, zmove = vz * gustep;
    x = x + xmove;
    y = y + ymove;
    z = z + zmove;
    bool inside = !(x < 0.0 || x >= g.slx || y < 0.0 || y >= g.sly || z < 0.0 || z >= g.slz);
    double px, py, pz;
    Relay rx = relay_axis(x, xlb, xb, ib, g.nx, g.slx, inside, px);
    Relay ry = relay_axis(y, ylb, yb, jb, g.ny, g.sly, inside, py);
    Relay rz = relay_axis(z, zlb, zb, kb, g.nz, g.slz, inside, pz);
    if (!(px >= 0.0 && px <= g.slx && py >= 0.0 && py <= g.sly && pz >= 0.0 && pz <= g.slz))
      atomicOr(flag, kFlagFar);
    const State64 &o = RESORT ? dst.t : s;
    o.x[d] = px;
    o.y[d] = py;
    o.z[d] = pz;
    o.vx[d] = vx;
    o.vy[d] = vy;
    o.vz[d] = vz;
    if (RESORT) o.pid[d] = pid;
    // :717-757
    double xd1 = (rx.r + xlb) * 0.5 - xb, yd1 = (ry.r + ylb) * 0.5 - yb, zd1 = (rz.r + zlb) * 0.5 - zb;
    double vvx1 = (rx.r - xlb) * rustep, vvy1 = (ry.r - ylb) * rustep, vvz1 = (rz.r - zlb) * rustep;
    double vx1w4 = yd1 * zd1, vx1w2 = yd1 - vx1w4, vx1w3 = zd1 - vx1w4, vx1w1 = 1.0 - yd1 - vx1w3;
    double vy1w4 = xd1 * zd1, vy1w2 = xd1 - vy1w4, vy1w3 = zd1 - vy1w4, vy1w1 = 1.0 - xd1 - vy1w3;
    double vz1w4 = xd1 * yd1, vz1w2 = xd1 - vz1w4, vz1w3 = yd1 - vz1w4, vz1w1 = 1.0 - xd1 - vz1w3;
    double xd2 = (rx.la + rx.r) * 0.5 - rx.a, yd2 = (ry.la + ry.r) * 0.5 - ry.a,
           zd2 = (rz.la + rz.r) * 0.5 - rz.a;
    double vvx2 = (rx.la - rx.r) * rustep, vvy2 = (ry.la - ry.r) * rustep,
           vvz2 = (rz.la - rz.r) * rustep;
    double vx2w4 = yd2 * zd2, vx2w2 = yd2 - vx2w4, vx2w3 = zd2 - vx2w4, vx2w1 = 1.0 - yd2 - vx2w3;
    double vy2w4 = xd2 * zd2, vy2w2 = xd2 - vy2w4, vy2w3 = zd2 - vy2w4, vy2w1 = 1.0 - xd2 - vy2w3;
    double vz2w4 = xd2 * yd2, vz2w2 = xd2 - vz2w4, vz2w3 = yd2 - vz2w4, vz2w1 = 1.0 - xd2 - vz2w3;
    // the 12 contributions of each segment (:759-789), in the window when its
    // cell and the next node along every axis are inside
    auto seg = [&](int i, int j, int k, double vvx, double vvy, double vvz, const double wx4[4],
                   const double wy4[4], const double wz4[4]) {
      const int li = wrapn(i - ox, g.nx), lj = wrapn(j - oy, g.ny), lk = wrapn(k - oz, g.nz);
      int li1, lj1, lk1;
      const bool in_win = li < wx && lj < wy && lk < wz && win_next(li, wx, g.nx, li1) &&
                          win_next(lj, wy, g.ny, lj1) && win_next(lk, wz, g.nz, lk1) &&
                          fabs(vvx) <= 4.0 && fabs(vvy) <= 4.0 && fabs(vvz) <= 4.0;
      if (in_win) {
        auto at = [&](int c, int a, int b, int d) { return win + c * wn + (d * wy + b) * wx + a; };
        atomicAdd(at(0, li, lj, lk), (unsigned long long)em_fix(vvx * wx4[0]));
        atomicAdd(at(0, li, lj1, lk), (unsigned long long)em_fix(vvx * wx4[1]));
        atomicAdd(at(0, li, lj, lk1), (unsigned long long)em_fix(vvx * wx4[2]));
        atomicAdd(at(0, li, lj1, lk1), (unsigned long long)em_fix(vvx * wx4[3]));
        atomicAdd(at(1, li, lj, lk), (unsigned long long)em_fix(vvy * wy4[0]));
        atomicAdd(at(1, li1, lj, lk), (unsigned long long)em_fix(vvy * wy4[1]));
        atomicAdd(at(1, li, lj, lk1), (unsigned long long)em_fix(vvy * wy4[2]));
        atomicAdd(at(1, li1, lj, lk1), (unsigned long long)em_fix(vvy * wy4[3]));
        atomicAdd(at(2, li, lj, lk), (unsigned long long)em_fix(vvz * wz4[0]));
        atomicAdd(at(2, li1, lj, lk), (unsigned long long)em_fix(vvz * wz4[1]));
        atomicAdd(at(2, li, lj1, lk), (unsigned long long)em_fix(vvz * wz4[2]));
        atomicAdd(at(2, li1, lj1, lk), (unsigned long long)em_fix(vvz * wz4[3]));
      } else {
        const int i1 = i + 1, j1 = j + 1, k1 = k + 1;
        atomicAdd(jx + cidx(g, i, j, k), vvx * wx4[0]);
        atomicAdd(jx + cidx(g, i, j1, k), vvx * wx4[1]);
        atomicAdd(jx + cidx(g, i, j, k1), vvx * wx4[2]);
        atomicAdd(jx + cidx(g, i, j1, k1), vvx * wx4[3]);
        atomicAdd(jy + cidx(g, i, j, k), vvy * wy4[0]);
        atomicAdd(jy + cidx(g, i1, j, k), vvy * wy4[1]);
        atomicAdd(jy + cidx(g, i, j, k1), vvy * wy4[2]);
        atomicAdd(jy + cidx(g, i1, j, k1), vvy * wy4[3]);
        atomicAdd(jz + cidx(g, i, j, k), vvz * wz4[0]);
        atomicAdd(jz + cidx(g, i1, j, k), vvz * wz4[1]);
        atomicAdd(jz + cidx(g, i, j1, k), vvz * wz4[2]);
        atomicAdd(jz + cidx(g, i1, j1, k), vvz * wz4[3]);
      }
    };
    const double ax1[4] = {vx1w1, vx1w2, vx1w3, vx1w4}, ay1[4] = {vy1w1, vy1w2, vy1w3, vy1w4},
                 az1[4] = {vz1w1, vz1w2, vz1w3, vz1w4};
    const double ax2[4] = {vx2w1, vx2w2, vx2w3, vx2w4}, ay2[4] = {vy2w1, vy2w2, vy2w3, vy2w4},
                 az2[4] = {vz2w1, vz2w2, vz2w3, vz2w4};
    seg(ib, jb, kb, vvx1, vvy1, vvz1, ax1, ay1, az1);
    seg(rx.ia, ry.ia, rz.ia, vvx2, vvy2, vvz2, ax2, ay2, az2);
  }
  __syncthreads();
  for (int q = threadIdx.x; q < 3 * wn; q += kBlock) {
    const long long v = (long long)win[q];
    if (v != 0) {
      const int d = q / wn, r = q - d * wn;
      const int li = r % wx, lj = (r / wx) % wy, lk = r / (wx * wy);
      double *jd = d == 0 ? jx : (d == 1 ? jy : jz);
      atomicAdd(jd + cidx(g, ox + li, oy + lj, oz + lk), (double)v * (1.0 / kEmFix));
    }
  }
}

dim3 species_grid(const SpeciesTab &sp) {
  int64_t mx = 1;
  for (int s = 0; s < sp.nspec; s++) mx = sp.cnt[s] > mx ? sp.cnt[s] : mx;
  return dim3((unsigned)((mx + kBlock - 1) / kBlock), (unsigned)sp.nspec, 1);
}


// The weights that the reference's charge puts on nodes with index n on some
// axis (charge.F90:152-210 with int(): a particle in cell n - 1, or at a
// position of exactly L, reaches them), kept per species in the upper-face
// arrays.  A weight on node n + 1 (only a position of exactly L has one, of
// size zero) is dropped, as the reference's border exchange overwrites that
// node.
__global__ __launch_bounds__(kBlock) void k_faces64(State64 s, SpeciesTab sp, Grid g,
                                                    double *__restrict__ face, int64_t nf) {
  const int is = blockIdx.y;
  int64_t m = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (m >= sp.cnt[is]) return;
  m += sp.off[is];
  double x = s.x[m], y = s.y[m], z = s.z[m];
  x = x >= 0.0 ? (x <= g.slx ? x : g.slx) : 0.0;
  y = y >= 0.0 ? (y <= g.sly ? y : g.sly) : 0.0;
  z = z >= 0.0 ? (z <= g.slz ? z : g.slz) : 0.0;
  const int i = (int)x, j = (int)y, k = (int)z;
  if (i < g.nx - 1 && j < g.ny - 1 && k < g.nz - 1) return;
  const Cic w = cic(x, y, z, i, j, k);
  const double qc = sp.q[is];
  double *f = face + (int64_t)is * nf;
  const int ci[8] = {i, i + 1, i + 1, i, i, i + 1, i + 1, i};
  const int cj[8] = {j, j, j + 1, j + 1, j, j, j + 1, j + 1};
  const int ck[8] = {k + 1, k + 1, k + 1, k + 1, k, k, k, k};
  const double v[8] = {w.v1, w.v2, w.v3, w.v4, w.v5, w.v6, w.v7, w.v8};
#pragma unroll
  for (int q = 0; q < 8; q++) {
    if (ci[q] > g.nx || cj[q] > g.ny || ck[q] > g.nz) continue;
    const int64_t fi = face_index(g, ci[q], cj[q], ck[q]);
    if (fi >= 0) atomicAdd(f + fi, v[q] * qc);
  }
}

}  // namespace

void launch_relocate64(emg_ctx *c, int elo[3], int eext[3]) {
  int64_t nn = (int64_t)(c->g.nx + 1) * (c->g.ny + 1) * (c->g.nz + 1);
  const emg_config &f = c->cfg;
  timer_begin(c, kTRelocate);
  hipLaunchKernelGGL(k_relocate, dim3((unsigned)((nn + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                     c->stream, c->eb_dev, elo[0], elo[1], elo[2], eext[0], eext[1],
                     c->cfg.precision == EMG_PLAIN_FP64 ? c->node64 : nullptr,
                     c->cfg.precision == EMG_FP32_FIXED ? c->node32 : nullptr, c->g, f.e0[0],
                     f.e0[1], f.e0[2], f.b0[0], f.b0[1], f.b0[2]);
  timer_end(c, kTRelocate);
}

void launch_positn64(emg_ctx *c, int ustep) {
  timer_begin(c, kTPush);
  hipLaunchKernelGGL(k_positn64, species_grid(c->sp), dim3(kBlock), 0, c->stream, c->s64, c->sp,
                     c->g, ustep, c->dflag);
  timer_end(c, kTPush);
}

// The grid of a kernel whose blocks take `per` particles each: the run table
// `which` of the last sort when it is valid (one block per run), else
// (slices of `per` particles, species) with runs = nullptr.
static dim3 block_grid(const emg_ctx *c, int per, int which, const Run *&runs) {
  static_assert(kBlock * kDepPer == kRunEs && kBlock * kEmPer == kRunEm,
                "run lengths are the kernels' particles per block");
  if (c->runs_valid && c->nruns[which] > 0) {
    runs = c->runs[which];
    return dim3((unsigned)c->nruns[which]);
  }
  runs = nullptr;
  int64_t mx = 1;
  for (int q = 0; q < c->sp.nspec; q++) mx = c->sp.cnt[q] > mx ? c->sp.cnt[q] : mx;
  return dim3((unsigned)((mx + per - 1) / per), c->sp.nspec);
}

void launch_push_es64(emg_ctx *c, int ustep, bool resort) {
  launch_zero(c->rho64, sizeof(double) * c->ncell * c->sp.nspec, c->stream);
  timer_begin(c, kTPush);
  const Run *runs;
  const dim3 grid = block_grid(c, kBlock * kDepPer, 0, runs);
  if (resort)
    hipLaunchKernelGGL(k_push_dep64<true>, grid, dim3(kBlock), 0, c->stream, c->s64, c->sp,
                       c->node64, c->g, ustep, c->rho64, c->ncell, runs, c->dflag,
                       sort64_dest(c));
  else
    hipLaunchKernelGGL(k_push_dep64<false>, grid, dim3(kBlock), 0, c->stream, c->s64, c->sp,
                       c->node64, c->g, ustep, c->rho64, c->ncell, runs, c->dflag, SortDest{});
  timer_end(c, kTPush);
}

void launch_deposit64(emg_ctx *c) {
  launch_zero(c->rho64, sizeof(double) * c->ncell * c->sp.nspec, c->stream);
  timer_begin(c, kTDeposit);
  const Run *runs;
  const dim3 grid = block_grid(c, kBlock * kDepPer, 0, runs);
  hipLaunchKernelGGL(k_deposit64, grid, dim3(kBlock), 0, c->stream, c->s64, c->sp, c->g, c->rho64,
                     c->ncell, runs, c->dflag);
  timer_end(c, kTDeposit);
}

void launch_push_em64(emg_ctx *c, int ustep, bool resort) {
  launch_zero(c->j64, sizeof(double) * 3 * c->ncell * c->sp.nspec, c->stream);
  timer_begin(c, kTPush);
  const Run *runs;
  const dim3 grid = block_grid(c, kBlock * kEmPer, 1, runs);
  if (resort)
    hipLaunchKernelGGL(k_push_em64<true>, grid, dim3(kBlock), 0, c->stream, c->s64, c->sp,
                       c->node64, c->g, ustep, c->j64, c->ncell, runs, c->dflag, sort64_dest(c));
  else
    hipLaunchKernelGGL(k_push_em64<false>, grid, dim3(kBlock), 0, c->stream, c->s64, c->sp,
                       c->node64, c->g, ustep, c->j64, c->ncell, runs, c->dflag, SortDest{});
  timer_end(c, kTPush);
}

void launch_faces64(emg_ctx *c) {
  const int64_t nf = face_len(c->g);
  launch_zero(c->face, sizeof(double) * nf * c->sp.nspec, c->stream);
  int64_t mx = 1;
  for (int q = 0; q < c->sp.nspec; q++) mx = c->sp.cnt[q] > mx ? c->sp.cnt[q] : mx;
  hipLaunchKernelGGL(k_faces64, dim3((unsigned)((mx + kBlock - 1) / kBlock), c->sp.nspec),
                     dim3(kBlock), 0, c->stream, c->s64, c->sp, c->g, (double *)c->face, nf);
}

}  // namespace emg
