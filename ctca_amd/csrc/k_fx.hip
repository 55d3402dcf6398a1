// k_fx.hip -- the FP32 cell-index + offset form of the particle path with
// deterministic int64 fixed-point charge deposition.
//
// Layout.  Particles of species s live in tile buckets b = s * ntiles + t, a
// tile being a T^3 block of cells.  Structure of arrays per buffer: packed
// cell ix | iy << 10 | iz << 20, offsets fx, fy, fz in [0, 1), velocities.
// Bucket b of a buffer owns the slots [start[b], start[b] + cap[b]) and holds
// count[b] particles in two runs: nst[b] stayers from the front, and the
// direct arrivals (normally none) at the very end:
//   slot(i) = start + i                      for i <  nst
//   slot(i) = start + cap - count + i        for i >= nst
// Besides its slots, every bucket has an outbox: the particles that left the
// bucket's tile for one of the 26 neighbouring tiles in the last step, stored
// contiguously as 32-byte records grouped by direction, with the exclusive
// prefix of the 27 direction counts in pre[b][0..27].  Those particles belong
// to the neighbours: each bucket pulls its arrivals from its 26 neighbours'
// outboxes at the start of the next step.
//
// A step reads buffer cur and writes buffer 1 - cur.  One workgroup per bucket:
//   * stages the T+1 nodes per axis of the relocated fields into LDS;
//   * pushes (FP32 Boris), moves (cell + offset) and deposits (integer CIC
//     weights, 2^24 per unit weight, the eight weights of a particle summing
//     to exactly 2^24, into an LDS tile of T+3 nodes per axis) its own
//     particles and then the arrivals pulled from the neighbours' outboxes;
//   * a particle that stays in the tile takes the next slot of the bucket's
//     first run (LDS counter); a particle that leaves for a neighbouring tile
//     goes to the LDS outbox, which the workgroup writes out grouped by
//     direction at the end -- every global write of the step is a contiguous
//     run, none is a few scattered bytes in another bucket's cache lines;
//   * a particle that moves further than a neighbouring tile (or overflows the
//     outbox) is placed directly: one of kExtra slots at the end of the
//     destination bucket (global atomic), else the species' spill region,
//     which the next step processes through the global-memory path;
//   * the charge tile is added to the species' global int64 accumulator with
//     one atomic per non-zero node.  Integer sums are associative, so rho is
//     bitwise independent of particle order, scheduling and launch shape.
#include "emg_internal.h"

namespace emg {
namespace {

constexpr int kSpillBlocksPerSpecies = 64;
// Movers staged per bucket: the global outboxes hold kObStride records per
// bucket; a step kernel stages up to ob_cap of them in LDS (the current's
// three LDS tiles leave room for 512 at two blocks per CU; the charge kernel
// takes 1024, above the ~490 movers per bucket of a 268M-particle step)
constexpr int kObStride = 1024;
template <int DEP>
constexpr int ob_cap() { return DEP == 2 ? 512 : 1024; }
constexpr int kExtra = 16;    // slots per bucket for direct arrivals
// bucket capacities are multiples of 32 slots, so every bucket starts on a
// 128-byte line of each state array and every wave's load of a round covers
// whole lines: runs that start mid-line made each line straddled by two
// waves' loads be fetched twice (a synthetic stream of the same runs moved
// 18% more bytes and took 40% longer at a 28-byte offset)
constexpr int kAlign = 32;
__host__ __device__ constexpr int64_t align_up(int64_t v) { return (v + kAlign - 1) & ~(int64_t)(kAlign - 1); }
constexpr int kCodes = 27;    // direction codes (dz+1)*9 + (dy+1)*3 + (dx+1); 13 = same tile
constexpr int kFar = 27;      // not a neighbouring tile
constexpr int kPre = 28;      // prefix table entries per bucket
constexpr uint32_t kNoStore = 0x80000000u;  // buffer offset past every bucket: store dropped

__device__ __forceinline__ int wrapc(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// The packed cell index: ix in the low cbx bits, iy in the next cby, iz above
// (the grid's bit widths, ceil(log2 n) for x and y; at most 32 bits in all).
__device__ __forceinline__ int pack_cell(int ix, int iy, int iz, const Grid &g) {
  return ix | (iy << g.cbx) | (iz << (g.cbx + g.cby));
}
__device__ __forceinline__ int cell_ix(int cell, const Grid &g) {
  return cell & ((1 << g.cbx) - 1);
}
__device__ __forceinline__ int cell_iy(int cell, const Grid &g) {
  return (cell >> g.cbx) & ((1 << g.cby) - 1);
}
__device__ __forceinline__ int cell_iz(int cell, const Grid &g) {
  return (int)((unsigned)cell >> (g.cbx + g.cby));
}

struct FxArgs {
  Grid g;
  int ntx, nty, ntz, ntiles, nb, nspec;
  StateFx in, out;
  // input buffer
  const int64_t *start_in;
  const int32_t *count_in, *nst_in, *cap_in, *spill_in, *pre_in;
  const FxRec *ob_in;
  const FxRec *spl;     // the input buffer's spill particles sorted by bucket
  const int32_t *shist; // how many of them per bucket
  const int64_t *spre;  // where each bucket's start
  // output buffer
  const int64_t *start_out;
  const int32_t *cap_out;
  int32_t *stay_out;  // stayers per bucket (written by the bucket's workgroup)
  int32_t *extra;     // direct arrivals per bucket (atomic)
  int32_t *spill_out;
  int32_t *pre_out;
  FxRec *ob_out;
  int64_t area, spill_cap;
  const float *node32;
  unsigned long long *rho;   // charge, nspec x ncell
  unsigned long long *jcur;  // zigzag current, nspec x 3 x ncell
  int64_t ncell;
  float qm[EMG_MAX_SPECIES];
  float gstep, rustep;
  double *ke_part;  // [(nb + nspec * kSpillBlocksPerSpecies) * 3]
  int *flag;
};

struct Pcl {
  int cell, pid;
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

// The hot loop addresses every particle array with one 32-bit byte offset
// (slot * 4, the slot counted from the bucket's first slot) added to the
// array's base shifted to that bucket (uniform over the block), so each load
// needs no per-array 64-bit address arithmetic and the context's slot count
// is not bounded by the 32-bit offsets.
template <class U>
__device__ __forceinline__ U &at(U *base, uint32_t off) {
  return *reinterpret_cast<U *>(reinterpret_cast<char *>(base) + off);
}

template <bool PID>
__device__ __forceinline__ void load_off(const StateFx &in, uint32_t off, Pcl &p) {
  p.cell = at(in.cell, off);
  p.pid = PID ? at(in.pid, off) : 0;
  p.fx = at(in.fx, off);
  p.fy = at(in.fy, off);
  p.fz = at(in.fz, off);
  p.vx = at(in.vx, off);
  p.vy = at(in.vy, off);
  p.vz = at(in.vz, off);
}

// the state arrays shifted to slot `first` (a bucket's first slot)
__device__ __forceinline__ StateFx shifted(const StateFx &s, int64_t first) {
  StateFx r;
  r.cell = s.cell + first;
  r.fx = s.fx + first;
  r.fy = s.fy + first;
  r.fz = s.fz + first;
  r.vx = s.vx + first;
  r.vy = s.vy + first;
  r.vz = s.vz + first;
  r.pid = s.pid ? s.pid + first : nullptr;
  return r;
}

__device__ __forceinline__ void load_rec(const FxRec *ob, int64_t k, Pcl &p) {
  const float4 *q = reinterpret_cast<const float4 *>(ob + k);
  float4 h0 = q[0], h1 = q[1];
  p.cell = __float_as_int(h0.x);
  p.fx = h0.y;
  p.fy = h0.z;
  p.fz = h0.w;
  p.vx = h1.x;
  p.vy = h1.y;
  p.vz = h1.z;
  p.pid = __float_as_int(h1.w);
}

// Masked store of one value per lane through a buffer resource over one
// bucket's slots (nrec bytes): a lane whose offset is kNoStore lies outside
// it and its store is dropped by the hardware range check, so the
// instruction count never depends on the data.
__device__ __forceinline__ void bstore(void *base, int32_t nrec, uint32_t off, uint32_t v) {
  __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc(base, 0, nrec, 0x00020000);
  __builtin_amdgcn_raw_buffer_store_b32(v, r, (int)off, 0, 0);
}

template <bool PID>
__device__ __forceinline__ void store_masked(const StateFx &out, int32_t nrec, uint32_t off,
                                             int cell, int pid, float fx, float fy, float fz,
                                             float vx, float vy, float vz) {
  bstore(out.cell, nrec, off, (uint32_t)cell);
  if (PID) bstore(out.pid, nrec, off, (uint32_t)pid);
  bstore(out.fx, nrec, off, __float_as_uint(fx));
  bstore(out.fy, nrec, off, __float_as_uint(fy));
  bstore(out.fz, nrec, off, __float_as_uint(fz));
  bstore(out.vx, nrec, off, __float_as_uint(vx));
  bstore(out.vy, nrec, off, __float_as_uint(vy));
  bstore(out.vz, nrec, off, __float_as_uint(vz));
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
__device__ __forceinline__ void gather_lds(const float4 *fA, const float2 *fB, int q0, float fx,
                                           float fy, float fz, float f[6]) {
  constexpr int N1 = T + 1;
  float wx0 = 1.f - fx;
  float wyz[4] = {(1.f - fy) * (1.f - fz), fy * (1.f - fz), (1.f - fy) * fz, fy * fz};
#pragma unroll
  for (int r = 0; r < 4; r++) {
    int q = q0 + ((r >> 1) * N1 + (r & 1)) * N1;
    float4 A0 = fA[q], A1 = fA[q + 1];
    float2 B0 = fB[q], B1 = fB[q + 1];
    float w0 = wx0 * wyz[r], w1 = fx * wyz[r];
    if (r == 0) {
      f[0] = w0 * A0.x;
      f[1] = w0 * A0.y;
      f[2] = w0 * A0.z;
      f[3] = w0 * A0.w;
      f[4] = w0 * B0.x;
      f[5] = w0 * B0.y;
    } else {
      f[0] = fmaf(w0, A0.x, f[0]);
      f[1] = fmaf(w0, A0.y, f[1]);
      f[2] = fmaf(w0, A0.z, f[2]);
      f[3] = fmaf(w0, A0.w, f[3]);
      f[4] = fmaf(w0, B0.x, f[4]);
      f[5] = fmaf(w0, B0.y, f[5]);
    }
    f[0] = fmaf(w1, A1.x, f[0]);
    f[1] = fmaf(w1, A1.y, f[1]);
    f[2] = fmaf(w1, A1.z, f[2]);
    f[3] = fmaf(w1, A1.w, f[3]);
    f[4] = fmaf(w1, B1.x, f[4]);
    f[5] = fmaf(w1, B1.y, f[5]);
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

constexpr float kBelowOne = 0x1.fffffep-1f;

// move one axis: offset += v * dt, whole cells carried into the index.  Returns
// the cell displacement; the offset stays in [0, 1) (an offset that rounds up
// to 1.0 is held at the largest float below 1, one ulp from the cell face).
__device__ __forceinline__ int move_offset(float &f, float v, float dt) {
  float nf = fmaf(v, dt, f);
  float fl = floorf(nf);
  f = fminf(nf - fl, kBelowOne);
  return (int)fl;
}

// periodic index i + d for i in [0, n), d in [-n, n]
__device__ __forceinline__ int wrap_add(int i, int d, int n) {
  unsigned x = (unsigned)(i + d + n);
  return (int)min(min(x, x - (unsigned)n), x - 2u * (unsigned)n);
}

// |d| >= n: a move of at least one box length
__device__ __forceinline__ bool far_move(int d, int n) {
  return (unsigned)(d + (n - 1)) > (unsigned)(2 * n - 2);
}

// integer CIC weights: seven truncated products, the eighth takes the
// remainder, so the eight sum to exactly 2^24
__device__ __forceinline__ void fixed_weights(float fx, float fy, float fz, int W[8]) {
  float gx = 1.f - fx;
  float z0 = (1.f - fz) * 16777216.f, z1 = fz * 16777216.f;
  float w00 = (1.f - fy) * z0, w10 = fy * z0, w01 = (1.f - fy) * z1, w11 = fy * z1;
  W[0] = (int)(gx * w00);
  W[1] = (int)(fx * w00);
  W[2] = (int)(gx * w10);
  W[3] = (int)(fx * w10);
  W[4] = (int)(gx * w01);
  W[5] = (int)(fx * w01);
  W[6] = (int)(gx * w11);
  W[7] = (1 << kFixShift) - ((W[0] + W[1] + W[2]) + (W[3] + W[4]) + (W[5] + W[6]));
}

template <int T>
__device__ __forceinline__ int tile_of_cell(int cell, const Grid &g, int ntx, int nty) {
  constexpr int kLog2T = T == 4 ? 2 : 3;
  int ix = cell_ix(cell, g), iy = cell_iy(cell, g), iz = cell_iz(cell, g);
  return ((iz >> kLog2T) * nty + (iy >> kLog2T)) * ntx + (ix >> kLog2T);
}

// the bucket of the same species whose tile lies in direction `code` of b's tile
__device__ __forceinline__ int neighbour_bucket(int b, int code, int ntx, int nty, int ntz,
                                                int ntiles) {
  int s = b / ntiles, o = b % ntiles;
  int tx = wrapc(o % ntx + code % 3 - 1, ntx);
  int ty = wrapc((o / ntx) % nty + (code / 3) % 3 - 1, nty);
  int tz = wrapc(o / (ntx * nty) + code / 9 - 1, ntz);
  return s * ntiles + (tz * nty + ty) * ntx + tx;
}

// place one particle into bucket dest by a global atomic: one of the bucket's
// kExtra end slots, else the species' spill region
__device__ void place_direct(const FxArgs &a, int dest, int cell, int pid, float fx, float fy,
                             float fz, float vx, float vy, float vz) {
  int idx = atomicAdd(&a.extra[dest], 1);
  int64_t slot;
  if (idx < kExtra) {
    slot = a.start_out[dest] + a.cap_out[dest] - 1 - idx;
  } else {
    int s = dest / a.ntiles;
    int q = atomicAdd(&a.spill_out[s], 1);
    if (q >= a.spill_cap) {
      atomicOr(a.flag, kFlagSpillLost);
      return;
    }
    slot = a.area + (int64_t)s * a.spill_cap + q;
  }
  store_pcl(a.out, slot, cell, pid, fx, fy, fz, vx, vy, vz);
}

// Umeda zigzag current of one particle (psolve1.F90:399-415 relay point,
// :717-789 segments and weights), in the frame of the old cell: the particle
// goes from fb (offset in the old cell) to d + fa (d whole cells, fa the new
// offset); the relay point is the midpoint when the cell does not change, else
// the face between the two cells.  Per axis: the offset of each segment's
// midpoint in its own cell and the segment's flux (displacement / ustep).
struct Zig {
  float md1, vv1, md2, vv2;
};

// A particle that leaves the box through a face takes that face as its relay
// point and the ghost cell -1 or n as its second cell (psolve1.F90:440-478),
// whatever the distance moved; d2 returns the second cell relative to ib.
__device__ __forceinline__ Zig zig_axis(float fb, int d, float fa, int ib, int n, float rustep,
                                        int &d2) {
  float xa = (float)d + fa;
  float r;
  if (ib + d < 0) {
    r = (float)(-ib);
    d2 = -1 - ib;
  } else if (ib + d >= n) {
    r = (float)(n - ib);
    d2 = n - ib;
  } else {
    r = d == 0 ? 0.5f * (fb + xa) : (float)max(0, d);
    d2 = d;
  }
  Zig z;
  z.md1 = 0.5f * (r + fb);
  z.vv1 = (r - fb) * rustep;
  z.md2 = 0.5f * (xa + r) - (float)d2;
  z.vv2 = (xa - r) * rustep;
  return z;
}

// the four weights of a flux component over the transverse midpoint offsets
// (u, w): the reference's arithmetic form (psolve1.F90:728-734)
__device__ __forceinline__ void zig_w4(float u, float w, float W[4]) {
  W[3] = u * w;
  W[1] = u - W[3];
  W[2] = w - W[3];
  W[0] = 1.f - u - W[2];
}

// a current contribution rounded to an integer, halves up: floor(v + 0.5) in
// one instruction (v_cvt_rpi_i32_f32; the compiler emits it only under
// fast-math flags, which the rest of this file must not have)
__device__ __forceinline__ int round_half_up(float v) {
  int r;
  asm("v_cvt_rpi_i32_f32 %0, %1" : "=v"(r) : "v"(v));
  return r;
}

// fixed point of a current contribution (kJScale per unit flux; a flux is at
// most a few cells per unit time, far inside the int32 range)
__device__ __forceinline__ unsigned long long jfix(float v) {
  return (unsigned long long)(long long)round_half_up(v * (float)kJScale);
}

// fixed point of a flux already scaled by kJScale times a weight: the same
// value as jfix(flux * w), since scaling by a power of two is exact
__device__ __forceinline__ int jfix_scaled(float fs, float w) { return round_half_up(fs * w); }

// Zigzag current of one particle into the block's current tiles: segment 1
// at tile cell q1, segment 2 at q2.  When both segments lie in one cell (the
// particle did not change cell) the two contributions of each node are summed
// in FP32 (one fma) and rounded to fixed point once, for a single atomic;
// only the lanes whose particle changed cell issue the second segment's
// atomics.
template <int N3>
__device__ __forceinline__ void zig_lds(unsigned long long *jh, int q1, int q2, const Zig &zx,
                                        const Zig &zy, const Zig &zz) {
  constexpr int P = N3 * N3 * N3;
  const bool same = q1 == q2;
  // per component c: the flux (times kJScale) and the two transverse midpoint
  // offsets of each segment, and the node offsets of its four weights
  // (zig_w4 order)
  constexpr float S = (float)kJScale;
  const float f1[3] = {zx.vv1 * S, zy.vv1 * S, zz.vv1 * S},
              f2[3] = {zx.vv2 * S, zy.vv2 * S, zz.vv2 * S};
  const float u1[3] = {zy.md1, zx.md1, zx.md1}, w1[3] = {zz.md1, zz.md1, zy.md1};
  const float u2[3] = {zy.md2, zx.md2, zx.md2}, w2[3] = {zz.md2, zz.md2, zy.md2};
  const int o[3][4] = {{0, N3, N3 * N3, N3 + N3 * N3},
                       {P, P + 1, P + N3 * N3, P + 1 + N3 * N3},
                       {2 * P, 2 * P + 1, 2 * P + N3, 2 * P + 1 + N3}};
#pragma unroll
  for (int c = 0; c < 3; c++) {
    float W[4], U[4];
    zig_w4(u1[c], w1[c], W);
    zig_w4(u2[c], w2[c], U);
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const int v = round_half_up(fmaf(same ? f2[c] : 0.f, U[k], f1[c] * W[k]));
      atomicAdd(&jh[q1 + o[c][k]], (unsigned long long)(long long)v);
    }
  }
  if (!same) {
#pragma unroll
    for (int c = 0; c < 3; c++) {
      float U[4];
      zig_w4(u2[c], w2[c], U);
#pragma unroll
      for (int k = 0; k < 4; k++)
        atomicAdd(&jh[q2 + o[c][k]], (unsigned long long)(long long)jfix_scaled(f2[c], U[k]));
    }
  }
}

__device__ __forceinline__ void zig_seg_global(unsigned long long *j3, int64_t ncell,
                                               const Grid &g, int i, int j, int k, float mx,
                                               float my, float mz, float vx, float vy, float vz) {
  int i1 = i + 1 == g.nx ? 0 : i + 1, j1 = j + 1 == g.ny ? 0 : j + 1, k1 = k + 1 == g.nz ? 0 : k + 1;
  auto at3 = [&](int d, int a, int b, int c) {
    return &j3[(int64_t)d * ncell + ((int64_t)c * g.ny + b) * g.nx + a];
  };
  float W[4];
  zig_w4(my, mz, W);
  atomicAdd(at3(0, i, j, k), jfix(vx * W[0]));
  atomicAdd(at3(0, i, j1, k), jfix(vx * W[1]));
  atomicAdd(at3(0, i, j, k1), jfix(vx * W[2]));
  atomicAdd(at3(0, i, j1, k1), jfix(vx * W[3]));
  zig_w4(mx, mz, W);
  atomicAdd(at3(1, i, j, k), jfix(vy * W[0]));
  atomicAdd(at3(1, i1, j, k), jfix(vy * W[1]));
  atomicAdd(at3(1, i, j, k1), jfix(vy * W[2]));
  atomicAdd(at3(1, i1, j, k1), jfix(vy * W[3]));
  zig_w4(mx, my, W);
  atomicAdd(at3(2, i, j, k), jfix(vz * W[0]));
  atomicAdd(at3(2, i1, j, k), jfix(vz * W[1]));
  atomicAdd(at3(2, i, j1, k), jfix(vz * W[2]));
  atomicAdd(at3(2, i1, j1, k), jfix(vz * W[3]));
}

// One particle.  The common case is straight-line code (LDS gather, LDS charge
// tile, staying in the own bucket or going to the LDS outbox); the rare cases
// (outside the staged tile, beyond the charge halo, a move of a box length or
// more, outbox full) sit behind wave-uniform branches taken only when some
// lane of the wave needs them.  The stayers' stores are masked stores, so the
// common path issues the same memory instructions whatever the data.  `act`
// is false for lanes past the end of a round; full rounds pass a literal true.
// DEP: 0 no deposit, 1 charge (psolve2 + charge), 2 zigzag current (psolve1).
template <int T, bool PUSH, int DEP, bool PID>
__device__ __forceinline__ void fx_one(const FxArgs &a, int s, bool act, const Pcl &p, int ox,
                                       int oy, int oz, int twx, int twy, int twz, bool lds_ok,
                                       const StateFx &outb,
                                       int32_t nrec,
                                       int *n_stay, int *n_out, FxRec *outbox,
                                       unsigned char *ocode, const float4 *fA, const float2 *fB,
                                       unsigned long long *rh, double ke[3]) {
  constexpr int N1 = T + 1, N3 = T + 3;
  const Grid &g = a.g;
  int ix = cell_ix(p.cell, g), iy = cell_iy(p.cell, g), iz = cell_iz(p.cell, g);
  int lx = ix - ox, ly = iy - oy, lz = iz - oz;  // cell in the staged tile
  bool in_tile = lds_ok && max((unsigned)lx, max((unsigned)ly, (unsigned)lz)) < (unsigned)T;
  float fx = p.fx, fy = p.fy, fz = p.fz, vx = p.vx, vy = p.vy, vz = p.vz;
  float f[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  if (PUSH && lds_ok)
    gather_lds<T>(fA, fB, in_tile ? (lz * N1 + ly) * N1 + lx : 0, fx, fy, fz, f);
  if (__ballot(!in_tile)) {
    if (!in_tile) {
      if (ix >= g.nx || iy >= g.ny || iz >= g.nz) {  // never for a stored particle
        if (act) atomicOr(a.flag, kFlagDomain);
        ix = min(ix, g.nx - 1);
        iy = min(iy, g.ny - 1);
        iz = min(iz, g.nz - 1);
      }
      if (PUSH) gather_global(a.node32, g, ix, iy, iz, fx, fy, fz, f);
    }
  }
  if (PUSH) {
    boris32(vx, vy, vz, f, a.qm[s] * a.gstep * 0.5f);
    // kinetic energy in FP64, as energy.F90:38-43 sums it: the square of an
    // FP32 velocity is exact in double, so only the running sum rounds
    const double ax = act ? (double)vx : 0.0, ay = act ? (double)vy : 0.0,
                 az = act ? (double)vz : 0.0;
    ke[0] = fma(ax, ax, ke[0]);
    ke[1] = fma(ay, ay, ke[1]);
    ke[2] = fma(az, az, ke[2]);
  }
  const float fbx = fx, fby = fy, fbz = fz;
  const int ibx = ix, iby = iy, ibz = iz;
  int dx = move_offset(fx, vx, a.gstep);
  int dy = move_offset(fy, vy, a.gstep);
  int dz = move_offset(fz, vz, a.gstep);
  bool far = far_move(dx, g.nx) || far_move(dy, g.ny) || far_move(dz, g.nz);
  if (__ballot(far)) {
    if (far) {  // also a non-finite velocity; the indices stay in range
      if (act) atomicOr(a.flag, kFlagFar);
      dx = max(-g.nx, min(dx, g.nx));
      dy = max(-g.ny, min(dy, g.ny));
      dz = max(-g.nz, min(dz, g.nz));
    }
  }
  // charge tile coordinates: node (rx, ry, rz) of the (T+3)^3 tile is cell
  // (ox - 1 + rx, ...); the particle stays in its tile iff all are in [1, T]
  int rx = lx + dx + 1, ry = ly + dy + 1, rz = lz + dz + 1;
  unsigned rmax = max((unsigned)rx, max((unsigned)ry, (unsigned)rz));
  int rmin = min(rx, min(ry, rz));
  // the tile is twx x twy x twz cells (T, or less for a tile at the grid's
  // upper face when n is not a multiple of T)
  bool stay = act && in_tile && (unsigned)(rx - 1) < (unsigned)twx &&
              (unsigned)(ry - 1) < (unsigned)twy && (unsigned)(rz - 1) < (unsigned)twz;
  ix = wrap_add(ix, dx, g.nx);
  iy = wrap_add(iy, dy, g.ny);
  iz = wrap_add(iz, dz, g.nz);
  // slot of a stayer: the next one of the bucket's first run
  const int lane = __lane_id();
  uint64_t m = __ballot(stay);
  int sbase = 0;
  const int sleader = __ffsll((unsigned long long)m) - 1;
  if (m != 0 && lane == sleader) sbase = atomicAdd(n_stay, __popcll(m));
  if (DEP == 1) {
    int W[8];
    fixed_weights(fx, fy, fz, W);
    bool in_halo = in_tile && rmax <= (unsigned)(T + 1);
    if (act && in_halo) {
      const int q0 = (rz * N3 + ry) * N3 + rx;
#pragma unroll
      for (int q = 0; q < 8; q++)
        atomicAdd(&rh[q0 + ((q >> 2) * N3 + ((q >> 1) & 1)) * N3 + (q & 1)],
                  (unsigned long long)(long long)W[q]);
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
  if (DEP == 2) {
    int d2x, d2y, d2z;
    const Zig zx = zig_axis(fbx, dx, fx, ibx, g.nx, a.rustep, d2x),
              zy = zig_axis(fby, dy, fy, iby, g.ny, a.rustep, d2y),
              zz = zig_axis(fbz, dz, fz, ibz, g.nz, a.rustep, d2z);
    // tile coordinates of the second segment's cell
    const int sx = lx + d2x + 1, sy = ly + d2y + 1, sz = lz + d2z + 1;
    bool in_halo = in_tile && max((unsigned)sx, max((unsigned)sy, (unsigned)sz)) <= (unsigned)(T + 1);
    if (act && in_halo) {
      const int q1 = ((lz + 1) * N3 + (ly + 1)) * N3 + lx + 1, q2 = (sz * N3 + sy) * N3 + sx;
      zig_lds<N3>(rh, q1, q2, zx, zy, zz);
    }
    if (__ballot(act && !in_halo)) {
      if (act && !in_halo) {
        unsigned long long *j3 = a.jcur + (int64_t)s * 3 * a.ncell;
        zig_seg_global(j3, a.ncell, g, ibx, iby, ibz, zx.md1, zy.md1, zz.md1, zx.vv1, zy.vv1,
                       zz.vv1);
        zig_seg_global(j3, a.ncell, g, wrap_add(ibx, max(-g.nx, min(d2x, g.nx)), g.nx),
                       wrap_add(iby, max(-g.ny, min(d2y, g.ny)), g.ny),
                       wrap_add(ibz, max(-g.nz, min(d2z, g.nz)), g.nz), zx.md2, zy.md2, zz.md2,
                       zx.vv2, zy.vv2, zz.vv2);
      }
    }
  }
  const int cell = pack_cell(ix, iy, iz, g);
  uint32_t off = kNoStore;
  if (m != 0) {
    sbase = __builtin_amdgcn_readlane(sbase, sleader);
    int rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                         __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0));
    if (stay) off = (uint32_t)(sbase + rank) << 2;
  }
  const bool mover = act && !stay;
  uint64_t mv = __ballot(mover);
  if (mv != 0) {
    const int leader = __ffsll((unsigned long long)mv) - 1;
    int obase = 0;
    if (lane == leader) obase = atomicAdd(n_out, __popcll(mv));
    obase = __builtin_amdgcn_readlane(obase, leader);
    int k = obase + __builtin_amdgcn_mbcnt_hi((uint32_t)(mv >> 32),
                                              __builtin_amdgcn_mbcnt_lo((uint32_t)mv, 0));
    const bool boxed = mover && lds_ok && k < ob_cap<DEP>();
    if (boxed) {
      FxRec r;
      r.cell = cell;
      r.fx = fx;
      r.fy = fy;
      r.fz = fz;
      r.vx = vx;
      r.vy = vy;
      r.vz = vz;
      r.pid = p.pid;
      outbox[k] = r;
      // direction of the new tile: the tile coordinate of each axis is 0, 1
      // or 2 for the tile before, the own tile or the tile after; moves
      // beyond a neighbouring tile are placed directly at the end
      ocode[k] = (unsigned)rx <= (unsigned)(twx + 1) && (unsigned)ry <= (unsigned)(twy + 1) &&
                         (unsigned)rz <= (unsigned)(twz + 1)
                     ? (unsigned char)(((rz == 0 ? 0 : (rz <= twz ? 1 : 2)) * 3 +
                                        (ry == 0 ? 0 : (ry <= twy ? 1 : 2))) * 3 +
                                       (rx == 0 ? 0 : (rx <= twx ? 1 : 2)))
                     : (unsigned char)kFar;
    }
    if (__ballot(mover && !boxed)) {  // no outbox (spill rounds) or outbox full
      if (mover && !boxed) {
        int dest = s * a.ntiles + tile_of_cell<T>(cell, g, a.ntx, a.nty);
        place_direct(a, dest, cell, p.pid, fx, fy, fz, vx, vy, vz);
      }
    }
  }
  store_masked<PID>(outb, nrec, off, cell, p.pid, fx, fy, fz, vx, vy, vz);
}

template <int NT>
__device__ __forceinline__ void block_reduce3(double ke[3], double *out) {
  __shared__ double red[3][NT / 64];
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
    for (int k = 0; k < NT / 64; k++) v += red[threadIdx.x][k];
    out[threadIdx.x] = v;
  }
}

// XCD-aware bucket order.  Consecutive workgroups go round-robin to the 8
// XCDs, so XCD x takes the workgroups bid = x (mod 8), in order.  Tiles are
// enumerated y-slowest, then z, then x, and XCD x owns the eighth
// [x, x + 1) * ntiles / 8 of them: a y-slab spanning every z, so a density
// that varies along z loads the XCDs evenly and neighbouring tiles share halo
// nodes in that XCD's L2.  Inside its slab an XCD runs all tiles of species
// 0, then of species 1, and so on: the blocks that run side by side on a CU
// then carry similar work.  (With the species of one tile adjacent, the EM
// two-stream's 1.6K-particle beam blocks ran beside 16.4K-particle background
// blocks, and the CU slots they held did little: 6.86 ms against 4.67 ms.)
__device__ __forceinline__ int bucket_of_block(int bid, int nb, int nspec, int ntx, int ntz,
                                               int ntiles) {
  (void)nb;
  int s, v;
  if ((ntiles & 7) == 0) {
    const int per = ntiles >> 3, x = bid & 7, k = bid >> 3;
    s = k / per;
    v = x * per + (k - s * per);
  } else {
    s = bid % nspec;
    v = bid / nspec;
  }
  const int tx = v % ntx, r = v / ntx;
  const int tz = r % ntz, ty = r / ntz;
  return s * ntiles + (tz * (ntiles / (ntx * ntz)) + ty) * ntx + tx;
}

// Arrivals of bucket b: segment c (direction c as seen from the source) of the
// outbox of the neighbour in direction 26 - c.  Wave 0 reads the 26 segment
// bounds and writes their start record and the inclusive prefix of their
// lengths (seg_pre[0] = 0, seg_pre[27] = number of arrivals).
__device__ __forceinline__ void pull_table(const FxArgs &a, int b, int64_t *seg_src,
                                           int *seg_pre) {
  if (threadIdx.x < 64) {
    const int c = threadIdx.x;
    int n_c = 0;
    int64_t src = 0;
    if (c < kCodes && c != 13) {
      int nbk = neighbour_bucket(b, 26 - c, a.ntx, a.nty, a.ntz, a.ntiles);
      int lo = a.pre_in[(int64_t)nbk * kPre + c], hi = a.pre_in[(int64_t)nbk * kPre + c + 1];
      n_c = hi - lo;
      src = (int64_t)nbk * kObStride + lo;
    }
    int v = n_c;
#pragma unroll
    for (int o = 1; o < 32; o <<= 1) {
      int u = __shfl_up(v, o);
      if (c >= o) v += u;
    }
    if (c < kCodes) {
      seg_src[c] = src;
      seg_pre[c + 1] = v;
    }
    if (c == 0) seg_pre[0] = 0;
  }
}

// record index of arrival j: the segment c with seg_pre[c] <= j < seg_pre[c + 1]
__device__ __forceinline__ int64_t arrival_rec(const int64_t *seg_src, const int *seg_pre, int j) {
  int lo = 0, hi = kCodes;
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (seg_pre[mid] <= j) lo = mid;
    else hi = mid;
  }
  return seg_src[lo] + (j - seg_pre[lo]);
}

template <int T, bool PUSH, int DEP, bool PID>
__global__ __launch_bounds__(T == 8 ? 512 : 256) void k_fx_step(FxArgs a) {
  constexpr int NT = T == 8 ? 512 : 256;  // threads: T^3 cells x 64 particles per cell
  constexpr int N1 = T + 1, N3 = T + 3;
  constexpr int kOb = ob_cap<DEP>();
  constexpr int kPer = kOb / NT;
  static_assert(kOb % NT == 0, "the outbox is written kPer records per thread");
  __shared__ float4 fA[PUSH ? N1 * N1 * N1 : 1];
  __shared__ float2 fB[PUSH ? N1 * N1 * N1 : 1];
  constexpr int kTileNodes = DEP == 1 ? N3 * N3 * N3 : (DEP == 2 ? 3 * N3 * N3 * N3 : 1);
  __shared__ unsigned long long rh[kTileNodes];  // charge tile, or the three current tiles
  __shared__ FxRec outbox[kOb];
  __shared__ unsigned char ocode[kOb];
  __shared__ int64_t seg_src[kCodes];
  __shared__ int seg_pre[kPre];
  __shared__ int n_stay, n_out;
  __shared__ int hist[kPre], opre[kPre];
  double ke[3] = {0.0, 0.0, 0.0};  // per lane, FP64
  const int bid = blockIdx.x;
  int part = bid;  // energy partial slot: bucket index, then spill blocks
  int s = 0, b = -1;
  if (threadIdx.x == 0) {
    n_stay = 0;
    n_out = 0;
  }
  if (threadIdx.x < kPre) hist[threadIdx.x] = 0;
  if (bid < a.nb) {
    b = bucket_of_block(bid, a.nb, a.nspec, a.ntx, a.ntz, a.ntiles);
    part = b;
    s = b / a.ntiles;
    const int t = b - s * a.ntiles;
    const int ox = (t % a.ntx) * T, oy = ((t / a.ntx) % a.nty) * T, oz = (t / (a.ntx * a.nty)) * T;
    const int twx = min(T, a.g.nx - ox), twy = min(T, a.g.ny - oy), twz = min(T, a.g.nz - oz);
    const int cnt = a.count_in[b];
    const int nst = a.nst_in[b];
    const uint32_t gap32 = (uint32_t)(a.cap_in[b] - cnt);
    // input and output state seen from the bucket's first slot
    const StateFx inb = shifted(a.in, a.start_in[b]);
    const StateFx outb = shifted(a.out, a.start_out[b]);
    const int32_t nrec = a.cap_out[b] * 4;
    auto off_of = [&](int i) -> uint32_t {
      return ((uint32_t)i + (i >= nst ? gap32 : 0u)) << 2;
    };
    auto ld = [&](int i, Pcl &p) {
      load_off<PID>(inb, off_of(min(i + (int)threadIdx.x, cnt - 1)), p);
    };
    // particle j of the block's list: the own particles, the arrivals pulled
    // from the neighbours' outboxes (needs the pull table), then the previous
    // step's spill particles that lie in this tile; lanes past the list load
    // nothing
    int total = 0, narr = 0;
    const int64_t sp0 = a.spre[b];
    // read before the barrier below: after it the load would stall the block
    const int nspill = a.shist[b];
    auto ldx = [&](int i, Pcl &p) {
      const int j = i + (int)threadIdx.x;
      if (j < cnt) {
        load_off<PID>(inb, off_of(j), p);
      } else if (j < cnt + narr) {
        load_rec(a.ob_in, arrival_rec(seg_src, seg_pre, j - cnt), p);
        if (!PID) p.pid = 0;
      } else if (j < total) {
        load_rec(a.spl, sp0 + (j - cnt - narr), p);
        if (!PID) p.pid = 0;
      }
    };
    auto one = [&](bool act, const Pcl &p) {
      fx_one<T, PUSH, DEP, PID>(a, s, act, p, ox, oy, oz, twx, twy, twz, true, outb, nrec,
                                &n_stay, &n_out,
                                    outbox, ocode, fA, fB, rh, ke);
    };
    // Software pipelined three deep: while a lane pushes particle k, the loads
    // of particles k + 1 and k + 2 (the next rounds of the list) are in
    // flight.  Three state sets rotate through the unrolled body, so no set is
    // copied while its loads are outstanding.  Rounds of own particles only
    // load with clamped indices, so every lane loads; the last rounds, which
    // hold the rest of the own particles and the arrivals, load through the
    // list.  Lanes of the two rounds loaded ahead that fall on arrivals load
    // those when the last rounds begin.
    Pcl A{}, B{}, C{};
    if (cnt > 0) {
      ld(0, A);
      ld(NT, B);
    }
    pull_table(a, b, seg_src, seg_pre);
    if (PUSH) {
      // the field tile: each thread loads whole nodes (E and Bx, then By, Bz),
      // a batch of nodes' loads issued before any is waited for, then stored
      // (one batch of all its nodes; one node at a time in the current
      // kernel, whose registers are fuller)
      constexpr int NN = N1 * N1 * N1;
      constexpr int R = (NN + NT - 1) / NT;
      constexpr int RB = DEP == 2 ? 1 : R;
#pragma unroll
      for (int r0 = 0; r0 < R; r0 += RB) {
        float4 va[RB];
        float2 vb[RB];
#pragma unroll
        for (int r = 0; r < RB; r++) {
          const int nq = (int)threadIdx.x + (r0 + r) * NT;
          if (nq < NN) {
            const int lx = nq % N1, ly = (nq / N1) % N1, lz = nq / (N1 * N1);
            const int gx = wrapc(ox + lx, a.g.nx), gy = wrapc(oy + ly, a.g.ny),
                      gz = wrapc(oz + lz, a.g.nz);
            const float *pn = a.node32 + 8 * (((int64_t)gz * a.g.ny + gy) * a.g.nx + gx);
            va[r] = *reinterpret_cast<const float4 *>(pn);
            vb[r] = *reinterpret_cast<const float2 *>(pn + 4);
          }
        }
#pragma unroll
        for (int r = 0; r < RB; r++) {
          const int nq = (int)threadIdx.x + (r0 + r) * NT;
          if (nq < NN) {
            fA[nq] = va[r];
            fB[nq] = vb[r];
          }
        }
      }
    }
    if (DEP != 0)
      for (int q = threadIdx.x; q < kTileNodes; q += NT) rh[q] = 0ull;
    __syncthreads();
    narr = seg_pre[kCodes];
    total = cnt + narr + nspill;
    const int nfull = (cnt / NT) * NT;
    int base = 0;
    for (; base + 3 * NT <= nfull; base += 3 * NT) {
      ld(base + 2 * NT, C);
      one(true, A);
      ld(base + 3 * NT, A);
      one(true, B);
      ld(base + 4 * NT, B);
      one(true, C);
    }
    // the last rounds: the remaining own particles, then the arrivals and the
    // spill particles
    if (base + 2 * NT > cnt) {  // uniform over the block
      if (base + (int)threadIdx.x >= cnt) ldx(base, A);
      if (base + NT + (int)threadIdx.x >= cnt) ldx(base + NT, B);
    }
    for (; base < total; base += NT) {
      ldx(base + 2 * NT, C);
      one(base + (int)threadIdx.x < total, A);
      A = B;
      B = C;
    }
    if (DEP != 0) {
      // tile node (rx, ry, rz) of component d -> cell (ox - 1 + rx, ...) of
      // the species' charge (d = 0) or current component d
      __syncthreads();
      unsigned long long *r = DEP == 1 ? a.rho + (int64_t)s * a.ncell
                                       : a.jcur + (int64_t)s * 3 * a.ncell;
      for (int q = threadIdx.x; q < kTileNodes; q += NT) {
        unsigned long long v = rh[q];
        if (v != 0ull) {
          int d = q / (N3 * N3 * N3), qq = q - d * (N3 * N3 * N3);
          int rx = qq % N3, ry = (qq / N3) % N3, rz = qq / (N3 * N3);
          int gx = wrapc(ox - 1 + rx, a.g.nx);
          int gy = wrapc(oy - 1 + ry, a.g.ny);
          int gz = wrapc(oz - 1 + rz, a.g.nz);
          atomicAdd(&r[(int64_t)d * a.ncell + ((int64_t)gz * a.g.ny + gy) * a.g.nx + gx], v);
        }
      }
    }
  }
  // ---- write the outbox grouped by direction
  __syncthreads();
  if (b >= 0) {
    const int nout = min(n_out, kOb);
    const int t = b - s * a.ntiles;
    int cq[kPer], rq[kPer];
#pragma unroll
    for (int k = 0; k < kPer; k++) {
      const int q = (int)threadIdx.x + k * NT;
      cq[k] = kFar;
      rq[k] = 0;
      if (q < nout) {
        cq[k] = ocode[q];
        if (cq[k] < kFar) rq[k] = atomicAdd(&hist[cq[k]], 1);
      }
    }
    __syncthreads();
    if (threadIdx.x < 64) {
      const int c = threadIdx.x;
      int v = c < kCodes ? hist[c] : 0;
      int incl = v;
#pragma unroll
      for (int o = 1; o < 32; o <<= 1) {
        int u = __shfl_up(incl, o);
        if (c >= o) incl += u;
      }
      if (c < kPre) {
        int ex = incl - v;  // c = 27: the total
        opre[c] = ex;
        a.pre_out[(int64_t)b * kPre + c] = ex;
      }
    }
    if (threadIdx.x == 0) a.stay_out[b] = n_stay;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kPer; k++) {
      const int q = (int)threadIdx.x + k * NT;
      if (q < nout) {
        const FxRec r = outbox[q];
        if (cq[k] < kFar) {
          a.ob_out[(int64_t)b * kObStride + opre[cq[k]] + rq[k]] = r;
        } else {
          place_direct(a, s * a.ntiles + tile_of_cell<T>(r.cell, a.g, a.ntx, a.nty), r.cell, r.pid,
                       r.fx, r.fy, r.fz, r.vx, r.vy, r.vz);
        }
      }
    }
  }
  if (PUSH) {
    block_reduce3<NT>(ke, a.ke_part + 3 * (int64_t)part);
  }
}

// standalone charge deposition of the current state (no move, no bucket change):
// each bucket's runs plus the arrivals waiting in its neighbours' outboxes
template <int T>
__global__ __launch_bounds__(kBlock) void k_fx_deposit(FxArgs a) {
  constexpr int N3 = T + 3;
  __shared__ unsigned long long rh[N3 * N3 * N3];
  __shared__ int64_t seg_src[kCodes];
  __shared__ int seg_pre[kPre];
  const int bid = blockIdx.x;
  const Grid &g = a.g;
  auto deposit_global = [&](unsigned long long *r, int ix, int iy, int iz, const int W[8]) {
    int xs[2] = {ix, ix + 1 == g.nx ? 0 : ix + 1};
    int ys[2] = {iy, iy + 1 == g.ny ? 0 : iy + 1};
    int zs[2] = {iz, iz + 1 == g.nz ? 0 : iz + 1};
#pragma unroll
    for (int q = 0; q < 8; q++) {
      int a2 = q & 1, b2 = (q >> 1) & 1, c2 = q >> 2;
      atomicAdd(&r[((int64_t)zs[c2] * g.ny + ys[b2]) * g.nx + xs[a2]],
                (unsigned long long)(long long)W[q]);
    }
  };
  if (bid < a.nb) {
    const int b = bucket_of_block(bid, a.nb, a.nspec, a.ntx, a.ntz, a.ntiles);
    const int s = b / a.ntiles, t = b - s * a.ntiles;
    const int ox = (t % a.ntx) * T, oy = ((t / a.ntx) % a.nty) * T, oz = (t / (a.ntx * a.nty)) * T;
    const int cnt = a.count_in[b];
    const int nst = a.nst_in[b];
    const int64_t beg = a.start_in[b];
    const int64_t gap = a.cap_in[b] - cnt;
    pull_table(a, b, seg_src, seg_pre);
    for (int q = threadIdx.x; q < N3 * N3 * N3; q += kBlock) rh[q] = 0ull;
    __syncthreads();
    const int narr = seg_pre[kCodes];
    unsigned long long *r = a.rho + (int64_t)s * a.ncell;
    for (int i = threadIdx.x; i < cnt + narr; i += kBlock) {
      Pcl p;
      if (i < cnt) load_pcl(a.in, beg + i + (i >= nst ? gap : 0), p);
      else load_rec(a.ob_in, arrival_rec(seg_src, seg_pre, i - cnt), p);
      int ix = min(cell_ix(p.cell, g), g.nx - 1), iy = min(cell_iy(p.cell, g), g.ny - 1),
          iz = min(cell_iz(p.cell, g), g.nz - 1);
      int W[8];
      fixed_weights(p.fx, p.fy, p.fz, W);
      int rx = wrapc(ix - ox + 1, g.nx), ry = wrapc(iy - oy + 1, g.ny), rz = wrapc(iz - oz + 1, g.nz);
      if (rx <= T + 1 && ry <= T + 1 && rz <= T + 1) {
#pragma unroll
        for (int q = 0; q < 8; q++) {
          int a2 = q & 1, b2 = (q >> 1) & 1, c2 = q >> 2;
          atomicAdd(&rh[((rz + c2) * N3 + (ry + b2)) * N3 + (rx + a2)],
                    (unsigned long long)(long long)W[q]);
        }
      } else {
        deposit_global(r, ix, iy, iz, W);
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
      Pcl p;
      load_pcl(a.in, beg + i, p);
      int ix = min(cell_ix(p.cell, g), g.nx - 1), iy = min(cell_iy(p.cell, g), g.ny - 1),
          iz = min(cell_iz(p.cell, g), g.nz - 1);
      int W[8];
      fixed_weights(p.fx, p.fy, p.fz, W);
      deposit_global(r, ix, iy, iz, W);
    }
  }
}

// The fixed-point weights that land on nodes with index n on some axis (a
// particle in cell n - 1 of that axis), per species into the upper-face
// arrays: the buckets of the tiles in the last tile layer of an axis (their
// runs and the arrivals waiting in the neighbours' outboxes), and the spill
// particles.  The same integer weights as the periodic deposit, so the
// faces and the periodic accumulator split each particle's weight exactly.
template <int T>
__global__ __launch_bounds__(kBlock) void k_fx_faces(FxArgs a, unsigned long long *face,
                                                     int64_t nf) {
  __shared__ int64_t seg_src[kCodes];
  __shared__ int seg_pre[kPre];
  const int bid = blockIdx.x;
  const Grid &g = a.g;
  auto add = [&](int s, const Pcl &p) {
    const int ix = min(cell_ix(p.cell, g), g.nx - 1), iy = min(cell_iy(p.cell, g), g.ny - 1),
              iz = min(cell_iz(p.cell, g), g.nz - 1);
    if (ix < g.nx - 1 && iy < g.ny - 1 && iz < g.nz - 1) return;
    int W[8];
    fixed_weights(p.fx, p.fy, p.fz, W);
    unsigned long long *f = face + (int64_t)s * nf;
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const int64_t fi = face_index(g, ix + (q & 1), iy + ((q >> 1) & 1), iz + (q >> 2));
      if (fi >= 0) atomicAdd(&f[fi], (unsigned long long)(long long)W[q]);
    }
  };
  if (bid < a.nb) {
    const int b = bid;
    const int s = b / a.ntiles, t = b - s * a.ntiles;
    const int tx = t % a.ntx, ty = (t / a.ntx) % a.nty, tz = t / (a.ntx * a.nty);
    if (tx != a.ntx - 1 && ty != a.nty - 1 && tz != a.ntz - 1) return;  // uniform
    const int cnt = a.count_in[b];
    const int nst = a.nst_in[b];
    const int64_t beg = a.start_in[b];
    const int64_t gap = a.cap_in[b] - cnt;
    pull_table(a, b, seg_src, seg_pre);
    __syncthreads();
    const int narr = seg_pre[kCodes];
    for (int i = threadIdx.x; i < cnt + narr; i += kBlock) {
      Pcl p;
      if (i < cnt) load_pcl(a.in, beg + i + (i >= nst ? gap : 0), p);
      else load_rec(a.ob_in, arrival_rec(seg_src, seg_pre, i - cnt), p);
      add(s, p);
    }
  } else {
    const int k = bid - a.nb;
    const int s = k / kSpillBlocksPerSpecies, sub = k % kSpillBlocksPerSpecies;
    const int cnt = min(a.spill_in[s], (int32_t)a.spill_cap);
    const int64_t beg = a.area + (int64_t)s * a.spill_cap;
    for (int i = sub * kBlock + threadIdx.x; i < cnt; i += kBlock * kSpillBlocksPerSpecies) {
      Pcl p;
      load_pcl(a.in, beg + i, p);
      add(s, p);
    }
  }
}

// kinetic energy of the FP32 state (when no fused push partials exist)
__global__ __launch_bounds__(kBlock) void k_fx_energy(FxArgs a) {
  __shared__ int64_t seg_src[kCodes];
  __shared__ int seg_pre[kPre];
  double ke[3] = {0.0, 0.0, 0.0};
  const int bid = blockIdx.x;
  auto add = [&](const Pcl &p) {
    ke[0] += (double)p.vx * p.vx;
    ke[1] += (double)p.vy * p.vy;
    ke[2] += (double)p.vz * p.vz;
  };
  if (bid < a.nb) {
    const int cnt = a.count_in[bid], nst = a.nst_in[bid];
    const int64_t beg = a.start_in[bid], gap = a.cap_in[bid] - cnt;
    pull_table(a, bid, seg_src, seg_pre);
    __syncthreads();
    const int narr = seg_pre[kCodes];
    for (int i = threadIdx.x; i < cnt + narr; i += kBlock) {
      Pcl p;
      if (i < cnt) load_pcl(a.in, beg + i + (i >= nst ? gap : 0), p);
      else load_rec(a.ob_in, arrival_rec(seg_src, seg_pre, i - cnt), p);
      add(p);
    }
  } else {
    const int s = bid - a.nb;
    const int cnt = min(a.spill_in[s], (int32_t)a.spill_cap);
    const int64_t beg = a.area + (int64_t)s * a.spill_cap;
    for (int i = threadIdx.x; i < cnt; i += kBlock) {
      Pcl p;
      load_pcl(a.in, beg + i, p);
      add(p);
    }
  }
  block_reduce3<kBlock>(ke, a.ke_part + 3 * (int64_t)bid);
}

// output capacities of a step: the bucket's particles, its arrivals waiting in
// the neighbours' outboxes and kExtra direct slots; clears the step counters
// bucket of a stored particle's cell, the cell clamped into the grid
template <int T>
__device__ __forceinline__ int bucket_of_cell(int cell, int s, const Grid &g, int ntiles, int ntx,
                                              int nty) {
  const int ix = min(cell_ix(cell, g), g.nx - 1), iy = min(cell_iy(cell, g), g.ny - 1),
            iz = min(cell_iz(cell, g), g.nz - 1);
  return s * ntiles + tile_of_cell<T>(pack_cell(ix, iy, iz, g), g, ntx, nty);
}

// The spill particles of the input buffer join the lists of the buckets their
// cells lie in: count them per bucket (shist is zero on entry) ...
template <int T>
__global__ void k_fx_spill_hist(StateFx in, const int32_t *spill, int64_t area, int64_t spill_cap,
                                Grid g, int ntiles, int ntx, int nty, int32_t *shist) {
  const int s = blockIdx.y;
  const int64_t cnt = min((int64_t)spill[s], spill_cap);
  const int64_t beg = area + (int64_t)s * spill_cap;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < cnt;
       i += (int64_t)gridDim.x * blockDim.x)
    atomicAdd(&shist[bucket_of_cell<T>(in.cell[beg + i], s, g, ntiles, ntx, nty)], 1);
}

// ... and copy them, as outbox records, to their bucket's range of spl
template <int T>
__global__ void k_fx_spill_scatter(StateFx in, const int32_t *spill, int64_t area,
                                   int64_t spill_cap, Grid g, int ntiles, int ntx, int nty,
                                   const int64_t *spre, int32_t *scur, FxRec *spl) {
  const int s = blockIdx.y;
  const int64_t cnt = min((int64_t)spill[s], spill_cap);
  const int64_t beg = area + (int64_t)s * spill_cap;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < cnt;
       i += (int64_t)gridDim.x * blockDim.x) {
    Pcl p;
    load_pcl(in, beg + i, p);
    const int b = bucket_of_cell<T>(p.cell, s, g, ntiles, ntx, nty);
    FxRec r;
    r.cell = p.cell;
    r.fx = p.fx;
    r.fy = p.fy;
    r.fz = p.fz;
    r.vx = p.vx;
    r.vy = p.vy;
    r.vz = p.vz;
    r.pid = p.pid;
    spl[spre[b] + atomicAdd(&scur[b], 1)] = r;
  }
}

__global__ void k_fx_caps(const int32_t *count_in, const int32_t *pre_in, const int32_t *shist,
                          int32_t *cap_out, int32_t *extra, int32_t *spill_out, int nb, int nspec,
                          int ntx, int nty, int ntz, int ntiles) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < nb) {
    // the 26 neighbours' direction counts, all loaded before any is added
    int lo[kCodes], hi[kCodes];
#pragma unroll
    for (int c = 0; c < kCodes; c++) {
      lo[c] = 0;
      hi[c] = 0;
      if (c != 13) {
        const int n = neighbour_bucket(b, 26 - c, ntx, nty, ntz, ntiles);
        lo[c] = pre_in[(int64_t)n * kPre + c];
        hi[c] = pre_in[(int64_t)n * kPre + c + 1];
      }
    }
    int arr = 0;
#pragma unroll
    for (int c = 0; c < kCodes; c++) arr += hi[c] - lo[c];
    cap_out[b] = (int32_t)align_up(count_in[b] + arr + shist[b] + kExtra);
    extra[b] = 0;
  }
  if (b < nspec) spill_out[b] = 0;
}

// capacities of freshly bucketed particles (no outbox contents)
__global__ void k_fx_caps_hist(const int32_t *hist, int32_t *cap_out, int32_t *spill_out, int nb,
                               int nspec) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < nb) cap_out[b] = (int32_t)align_up(hist[b] + kExtra);
  if (b < nspec) spill_out[b] = 0;
}

__global__ void k_fill_i32(int32_t *p, int64_t n, int v) {
  int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b < n) p[b] = v;
}

// after a step: the output buckets' runs (direct arrivals beyond kExtra went
// to the spill region), spill counters clamped
__global__ void k_fx_finish(const int32_t *stay, const int32_t *extra, int32_t *nst_out,
                            int32_t *count_out, int nb, int32_t *spill, int nspec,
                            int32_t spill_cap, int32_t *shist, int32_t *scur) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < nb) {
    int ne = min(extra[b], kExtra);
    nst_out[b] = stay[b];
    count_out[b] = stay[b] + ne;
    shist[b] = 0;  // for the next step's spill sort
    scur[b] = 0;
  }
  if (b < nspec) spill[b] = min(spill[b], spill_cap);
}

// ---- bucketing of input ranges (existing buckets, flat species segments or
// spill regions, and outboxes) into the buckets of the other buffer
struct RangeDesc {
  int nrb;     // ranges [0, nrb) are buckets (species r / ntiles)
  int nspec;   // ranges [nrb, nrb + nspec) are per-species segments
  int nob;     // ranges [nrb + nspec, nrb + nspec + nob) are bucket outboxes
  int ntiles;
  int rb;      // blocks per range
  const int64_t *bstart;
  const int32_t *bcount, *bnst, *bcap;
  int64_t seg_start[EMG_MAX_SPECIES];
  const int32_t *seg_count_dev;        // device counts (spill) or nullptr
  int64_t seg_count[EMG_MAX_SPECIES];  // host counts (flat)
  const int32_t *pre;
  const FxRec *ob;
};

__device__ __forceinline__ int64_t range_count(const RangeDesc &d, int r, int &s) {
  if (r < d.nrb) {
    s = r / d.ntiles;
    return d.bcount[r];
  }
  if (r < d.nrb + d.nspec) {
    s = r - d.nrb;
    return d.seg_count_dev ? (int64_t)d.seg_count_dev[s] : d.seg_count[s];
  }
  int n = r - d.nrb - d.nspec;
  s = n / d.ntiles;
  return d.pre[(int64_t)n * kPre + kCodes];
}

__device__ __forceinline__ void range_load(const RangeDesc &d, const StateFx &in, int r,
                                           int64_t i, Pcl &p) {
  if (r < d.nrb) {
    int64_t cnt = d.bcount[r];
    load_pcl(in, d.bstart[r] + i + (i >= d.bnst[r] ? d.bcap[r] - cnt : 0), p);
  } else if (r < d.nrb + d.nspec) {
    load_pcl(in, d.seg_start[r - d.nrb] + i, p);
  } else {
    load_rec(d.ob, (int64_t)(r - d.nrb - d.nspec) * kObStride + i, p);
  }
}

// Both passes walk their range a block-wide round at a time (the loop is
// uniform, so every lane of a wave takes part in the grouping).  A range of
// bucket runs or an outbox puts a wave's lanes in one or a few buckets: each
// group of equal buckets takes one atomic (one atomic per lane on one
// counter made a re-bucketing of 268M particles take 145 ms instead of
// 3.9).  The flat segments (an upload, the spill) are in no tile order, and
// their lanes keep an atomic each.
template <int T>
__global__ __launch_bounds__(kBlock) void k_fx_hist(RangeDesc d, StateFx in, Grid g, int ntx,
                                                    int nty, int32_t *hist) {
  int r = blockIdx.x / d.rb, sub = blockIdx.x % d.rb;
  int s;
  int64_t cnt = range_count(d, r, s);
  const bool grouped = r < d.nrb || r >= d.nrb + d.nspec;  // uniform over the block
  for (int64_t i0 = (int64_t)sub * kBlock; i0 < cnt; i0 += (int64_t)d.rb * kBlock) {
    const int64_t i = i0 + threadIdx.x;
    const bool act = i < cnt;
    int b = -1;
    if (act) {
      Pcl p;
      range_load(d, in, r, i, p);
      b = s * d.ntiles + tile_of_cell<T>(p.cell, g, ntx, nty);
    }
    int lead = __lane_id(), inner = 0, size = 1;
    if (grouped) wave_groups(b, act, lead, inner, size);
    if (act && lead == __lane_id()) atomicAdd(&hist[b], size);
  }
}

template <int T>
__global__ __launch_bounds__(kBlock) void k_fx_scatter(RangeDesc d, StateFx in, StateFx out,
                                                       Grid g, int ntx, int nty, const int64_t *start,
                                                       int32_t *cursor) {
  int r = blockIdx.x / d.rb, sub = blockIdx.x % d.rb;
  int s;
  int64_t cnt = range_count(d, r, s);
  const bool grouped = r < d.nrb || r >= d.nrb + d.nspec;  // uniform over the block
  for (int64_t i0 = (int64_t)sub * kBlock; i0 < cnt; i0 += (int64_t)d.rb * kBlock) {
    const int64_t i = i0 + threadIdx.x;
    const bool act = i < cnt;
    Pcl p{};
    int b = -1;
    if (act) {
      range_load(d, in, r, i, p);
      b = s * d.ntiles + tile_of_cell<T>(p.cell, g, ntx, nty);
    }
    int lead = __lane_id(), inner = 0, size = 1;
    if (grouped) wave_groups(b, act, lead, inner, size);
    int base = 0;
    if (act && lead == __lane_id()) base = atomicAdd(&cursor[b], size);
    if (grouped) base = __shfl(base, lead);
    if (act)
      store_pcl(out, start[b] + base + inner, p.cell, p.pid, p.fx, p.fy, p.fz, p.vx, p.vy, p.vz);
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
  store_pcl(out, m, pack_cell(c[0], c[1], c[2], g), s64.pid[m], f[0], f[1], f[2], (float)s64.vx[m],
            (float)s64.vy[m], (float)s64.vz[m]);
}

// bucketed FP32 (one run per bucket) -> FP64 species-contiguous (dst offset =
// flat scan of counts)
__global__ __launch_bounds__(kBlock) void k_fx_to64(StateFx in, const int64_t *bstart,
                                                    const int32_t *bcount, const int64_t *flat,
                                                    int ntiles, SpeciesTab sp, Grid g,
                                                    State64 out) {
  int b = blockIdx.x;
  int s = b / ntiles;
  int cnt = bcount[b];
  int64_t beg = bstart[b];
  int64_t base = sp.off[s] + (flat[b] - flat[s * ntiles]);
  for (int i = threadIdx.x; i < cnt; i += kBlock) {
    int64_t src = beg + i, dst = base + i;
    int cell = in.cell[src];
    int ix = cell_ix(cell, g), iy = cell_iy(cell, g), iz = cell_iz(cell, g);
    out.x[dst] = (double)ix + (double)in.fx[src];
    out.y[dst] = (double)iy + (double)in.fy[src];
    out.z[dst] = (double)iz + (double)in.fz[src];
    out.vx[dst] = (double)in.vx[src];
    out.vy[dst] = (double)in.vy[src];
    out.vz[dst] = (double)in.vz[src];
    out.pid[dst] = in.pid ? in.pid[src] : 0;
  }
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
  a.nst_in = B.nst[in];
  a.cap_in = B.cap[in];
  a.spill_in = B.spill[in];
  a.pre_in = B.pre[in];
  a.ob_in = B.ob[in];
  a.spl = B.spl;
  a.shist = B.shist;
  a.spre = B.spre;
  a.start_out = B.start[out];
  a.cap_out = B.cap[out];
  a.stay_out = B.stay;
  a.extra = B.extra;
  a.spill_out = B.spill[out];
  a.pre_out = B.pre[out];
  a.ob_out = B.ob[out];
  a.area = B.area;
  a.spill_cap = B.spill_cap;
  a.node32 = c->node32;
  a.rho = c->rhofx;
  a.jcur = c->jfx;
  a.ncell = c->ncell;
  for (int s = 0; s < EMG_MAX_SPECIES; s++) a.qm[s] = (float)c->cfg.qm[s];
  a.gstep = 0.f;
  a.rustep = 0.f;
  a.ke_part = c->ke_part;
  a.flag = c->dflag;
  return a;
}

// launched forms: push + charge (psolve2 + charge), push + zigzag current
// (psolve1), move only (positn)
template <int T, bool PID>
void step_tp(emg_ctx *c, const FxArgs &a, bool push, int dep, unsigned grid) {
  if (push && dep == 1)
    hipLaunchKernelGGL((k_fx_step<T, true, 1, PID>), dim3(grid), dim3(T == 8 ? 512 : 256), 0,
                       c->stream, a);
  else if (push && dep == 2)
    hipLaunchKernelGGL((k_fx_step<T, true, 2, PID>), dim3(grid), dim3(T == 8 ? 512 : 256), 0,
                       c->stream, a);
  else
    hipLaunchKernelGGL((k_fx_step<T, false, 0, PID>), dim3(grid), dim3(T == 8 ? 512 : 256), 0,
                       c->stream, a);
}

template <int T>
void step_t(emg_ctx *c, const FxArgs &a, bool push, int dep, unsigned grid) {
  if (a.in.pid)
    step_tp<T, true>(c, a, push, dep, grid);
  else
    step_tp<T, false>(c, a, push, dep, grid);
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
      hipLaunchKernelGGL(k_fx_hist<4>, dim3(grid), dim3(kBlock), 0, c->stream, d, B.st[src], c->g,
                         B.ntx, B.nty, B.hist);
      break;
    case 8:
      hipLaunchKernelGGL(k_fx_hist<8>, dim3(grid), dim3(kBlock), 0, c->stream, d, B.st[src], c->g,
                         B.ntx, B.nty, B.hist);
      break;
    default: return set_error(EMG_ERR_ARG, "tile must be 4 or 8");
  }
  // capacities from the histogram, start offsets, cursor = 0, spill = 0
  hipLaunchKernelGGL(k_fx_caps_hist, dim3(gb), dim3(kBlock), 0, c->stream, B.hist, B.cap[dst],
                     B.spill[dst], B.nb, c->sp.nspec);
  hipLaunchKernelGGL(k_fill_i32, dim3(gb), dim3(kBlock), 0, c->stream, B.cursor, (int64_t)B.nb, 0);
  exclusive_scan_i32(B.cap[dst], B.start[dst], B.nb, c->scan_tmp, c->stream);
  switch (B.tile) {
    case 4:
      hipLaunchKernelGGL(k_fx_scatter<4>, dim3(grid), dim3(kBlock), 0, c->stream, d, B.st[src],
                         B.st[dst], c->g, B.ntx, B.nty, B.start[dst], B.cursor);
      break;
    case 8:
      hipLaunchKernelGGL(k_fx_scatter<8>, dim3(grid), dim3(kBlock), 0, c->stream, d, B.st[src],
                         B.st[dst], c->g, B.ntx, B.nty, B.start[dst], B.cursor);
      break;
  }
  // all particles of a bucket form one run from its start; no outbox contents
  hipLaunchKernelGGL(k_copy_i32, dim3(gb), dim3(kBlock), 0, c->stream, B.cursor, B.nb, B.count[dst]);
  hipLaunchKernelGGL(k_copy_i32, dim3(gb), dim3(kBlock), 0, c->stream, B.cursor, B.nb, B.nst[dst]);
  int64_t npre = (int64_t)B.nb * kPre;
  hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)((npre + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                     c->stream, B.pre[dst], npre, 0);
  B.cur = dst;
  return hip_check(hipGetLastError(), "bucketing");
}

RangeDesc bucket_desc(emg_ctx *c) {
  FxBuckets &B = c->fx;
  RangeDesc d{};
  d.nrb = B.nb;
  d.nspec = c->sp.nspec;
  d.nob = B.nb;
  d.ntiles = B.ntiles;
  d.bstart = B.start[B.cur];
  d.bcount = B.count[B.cur];
  d.bnst = B.nst[B.cur];
  d.bcap = B.cap[B.cur];
  d.seg_count_dev = B.spill[B.cur];
  for (int s = 0; s < c->sp.nspec; s++) d.seg_start[s] = B.area + (int64_t)s * B.spill_cap;
  d.pre = B.pre[B.cur];
  d.ob = B.ob[B.cur];
  return d;
}

}  // namespace

int fx_alloc(emg_ctx *c, int64_t ntotal) {
  FxBuckets &B = c->fx;
  const Grid &g = c->g;
  int T = c->cfg.tile;
  if (T <= 0) {  // default: 8^3-cell tiles where every axis is a multiple of 8
    const bool m8 = g.nx % 8 == 0 && g.ny % 8 == 0 && g.nz % 8 == 0;
    const bool m4 = g.nx % 4 == 0 && g.ny % 4 == 0 && g.nz % 4 == 0;
    const bool ge8 = g.nx >= 8 && g.ny >= 8 && g.nz >= 8;
    T = m8 ? 8 : (m4 ? 4 : (ge8 ? 8 : 4));
  }
  if (T != 4 && T != 8) return set_error(EMG_ERR_ARG, "tile must be 4 or 8");
  // a grid that is not a multiple of T ends in partial tiles; the tile and
  // its halo span at most two periods of each axis
  if (g.nx < T || g.ny < T || g.nz < T)
    return set_error(EMG_ERR_UNSUPPORTED, "FP32_FIXED needs at least `tile` cells per axis "
                                          "(4 with tile 4): use EMG_PLAIN_FP64 for this grid");
  int bz = 0;
  while ((1 << bz) < g.nz) bz++;
  if (g.cbx + g.cby + bz > 32)
    return set_error(EMG_ERR_UNSUPPORTED, "FP32_FIXED packs a cell index in 32 bits: ceil(log2 nx) "
                                          "+ ceil(log2 ny) + ceil(log2 nz) must be at most 32; use "
                                          "EMG_PLAIN_FP64 for this grid");
  B.tile = T;
  B.ntx = (g.nx + T - 1) / T;
  B.nty = (g.ny + T - 1) / T;
  B.ntz = (g.nz + T - 1) / T;
  B.ntiles = B.ntx * B.nty * B.ntz;
  B.nb = B.ntiles * c->sp.nspec;
  // every particle is in a bucket run, an outbox or the spill region, so the
  // buckets' capacities (count + arrivals + kExtra, rounded up to kAlign) sum
  // to at most this
  B.area = align_up(ntotal + (int64_t)(kExtra + kAlign - 1) * B.nb);
  int64_t maxs = 0;
  for (int s = 0; s < c->sp.nspec; s++) maxs = c->cfg.capacity[s] > maxs ? c->cfg.capacity[s] : maxs;
  B.spill_cap = align_up(maxs / 16 + 4096);
  if (B.spill_cap > (1 << 28)) B.spill_cap = 1 << 28;
  B.cap_total = B.area + B.spill_cap * c->sp.nspec;
  // the step's capacity scan packs two sums into one 64-bit term (the first
  // below 2^32); the hot loop's offsets are bucket-relative and do not bound it
  if (B.cap_total >= (int64_t(1) << 32))
    return set_error(EMG_ERR_ARG, "FP32_FIXED holds at most 2^32 particle slots per context");
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
        hipMalloc(&B.nst[k], sizeof(int32_t) * B.nb) != hipSuccess ||
        hipMalloc(&B.cap[k], sizeof(int32_t) * B.nb) != hipSuccess ||
        hipMalloc(&B.spill[k], sizeof(int32_t) * EMG_MAX_SPECIES) != hipSuccess ||
        hipMalloc(&B.pre[k], sizeof(int32_t) * B.nb * kPre) != hipSuccess)
      return set_error(EMG_ERR_NOMEM, "bucket tables");
    if (hipMalloc(&B.ob[k], sizeof(FxRec) * B.nb * kObStride) != hipSuccess)
      return set_error(EMG_ERR_NOMEM, "bucket outboxes");
    launch_zero(B.count[k], sizeof(int32_t) * B.nb, c->stream);
    launch_zero(B.nst[k], sizeof(int32_t) * B.nb, c->stream);
    launch_zero(B.cap[k], sizeof(int32_t) * B.nb, c->stream);
    launch_zero(B.spill[k], sizeof(int32_t) * EMG_MAX_SPECIES, c->stream);
    launch_zero(B.start[k], sizeof(int64_t) * (B.nb + 1), c->stream);
    launch_zero(B.pre[k], sizeof(int32_t) * B.nb * kPre, c->stream);
  }
  if (hipMalloc(&B.stay, sizeof(int32_t) * B.nb) != hipSuccess ||
      hipMalloc(&B.extra, sizeof(int32_t) * B.nb) != hipSuccess ||
      hipMalloc(&B.hist, sizeof(int32_t) * B.nb) != hipSuccess ||
      hipMalloc(&B.cursor, sizeof(int32_t) * B.nb) != hipSuccess ||
      hipMalloc(&B.shist, sizeof(int32_t) * B.nb) != hipSuccess ||
      hipMalloc(&B.scur, sizeof(int32_t) * B.nb) != hipSuccess ||
      hipMalloc(&B.spre, sizeof(int64_t) * (B.nb + 1)) != hipSuccess)
    return set_error(EMG_ERR_NOMEM, "bucket tables");
  if (hipMalloc(&B.spl, sizeof(FxRec) * B.spill_cap * c->sp.nspec) != hipSuccess)
    return set_error(EMG_ERR_NOMEM, "sorted spill buffer");
  launch_zero(B.shist, sizeof(int32_t) * B.nb, c->stream);
  launch_zero(B.scur, sizeof(int32_t) * B.nb, c->stream);
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
  d.nob = 0;
  d.ntiles = B.ntiles;
  for (int s = 0; s < c->sp.nspec; s++) {
    d.seg_start[s] = c->sp.off[s];
    d.seg_count[s] = c->sp.cnt[s];
  }
  return bucket_ranges(c, d, c->sp.nspec, 2048, flat);
}

// bucketed FP32 (runs, outboxes, spill) -> one run per bucket, nothing pending
int fx_rebucket(emg_ctx *c) {
  return bucket_ranges(c, bucket_desc(c), 2 * c->fx.nb + c->sp.nspec, 1, c->fx.cur);
}

int fx_to64(emg_ctx *c, State64 &dst) {
  FxBuckets &B = c->fx;
  int rc = fx_rebucket(c);
  if (rc) return rc;
  exclusive_scan_i32(B.count[B.cur], B.start[1 - B.cur], B.nb, c->scan_tmp, c->stream);
  hipLaunchKernelGGL(k_fx_to64, dim3(B.nb), dim3(kBlock), 0, c->stream, B.st[B.cur],
                     B.start[B.cur], B.count[B.cur], B.start[1 - B.cur], B.ntiles, c->sp, c->g,
                     dst);
  return hip_check(hipGetLastError(), "fx_to64");
}

void launch_fx_step(emg_ctx *c, int ustep, bool push, bool deposit, bool em) {
  // the bucket arrays exist only in the FP32 form: launch nothing otherwise
  if (c->cfg.precision != EMG_FP32_FIXED || c->fx.nb <= 0) return;
  FxBuckets &B = c->fx;
  const int dep = em ? 2 : (deposit ? 1 : 0);
  int in = B.cur, out = 1 - B.cur;
  unsigned gb = (unsigned)((B.nb + kBlock - 1) / kBlock);
  // the input's spill particles, sorted by bucket (the kernels exit at once
  // when there are none)
  const dim3 sg(kSpillBlocksPerSpecies, c->sp.nspec);
  if (B.tile == 4)
    hipLaunchKernelGGL(k_fx_spill_hist<4>, sg, dim3(kBlock), 0, c->stream, B.st[in], B.spill[in],
                       B.area, B.spill_cap, c->g, B.ntiles, B.ntx, B.nty, B.shist);
  else
    hipLaunchKernelGGL(k_fx_spill_hist<8>, sg, dim3(kBlock), 0, c->stream, B.st[in], B.spill[in],
                       B.area, B.spill_cap, c->g, B.ntiles, B.ntx, B.nty, B.shist);
  hipLaunchKernelGGL(k_fx_caps, dim3(gb), dim3(kBlock), 0, c->stream, B.count[in], B.pre[in],
                     B.shist, B.cap[out], B.extra, B.spill[out], B.nb, c->sp.nspec, B.ntx, B.nty,
                     B.ntz, B.ntiles);
  exclusive_scan2_i32(B.cap[out], B.start[out], B.shist, B.spre, B.nb, c->scan_tmp, c->stream);
  if (B.tile == 4)
    hipLaunchKernelGGL(k_fx_spill_scatter<4>, sg, dim3(kBlock), 0, c->stream, B.st[in],
                       B.spill[in], B.area, B.spill_cap, c->g, B.ntiles, B.ntx, B.nty, B.spre,
                       B.scur, B.spl);
  else
    hipLaunchKernelGGL(k_fx_spill_scatter<8>, sg, dim3(kBlock), 0, c->stream, B.st[in],
                       B.spill[in], B.area, B.spill_cap, c->g, B.ntiles, B.ntx, B.nty, B.spre,
                       B.scur, B.spl);
  if (dep == 1) launch_zero(c->rhofx, sizeof(unsigned long long) * c->ncell * c->sp.nspec, c->stream);
  if (dep == 2)
    launch_zero(c->jfx, sizeof(unsigned long long) * 3 * c->ncell * c->sp.nspec, c->stream);
  FxArgs a = make_args(c);
  a.gstep = (float)ustep;
  a.rustep = ustep > 0 ? 1.f / (float)ustep : 0.f;
  unsigned grid = (unsigned)B.nb;
  timer_begin(c, kTPush);
  switch (B.tile) {
    case 4: step_t<4>(c, a, push, dep, grid); break;
    case 8: step_t<8>(c, a, push, dep, grid); break;
  }
  timer_end(c, kTPush);
  hipLaunchKernelGGL(k_fx_finish, dim3(gb), dim3(kBlock), 0, c->stream, B.stay, B.extra,
                     B.nst[out], B.count[out], B.nb, B.spill[out], c->sp.nspec,
                     (int32_t)B.spill_cap, B.shist, B.scur);
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
  FxArgs a = make_args(c);
  timer_begin(c, kTEnergy);
  hipLaunchKernelGGL(k_fx_energy, dim3(B.nb + c->sp.nspec), dim3(kBlock), 0, c->stream, a);
  hipLaunchKernelGGL(k_ke_final, dim3(c->sp.nspec), dim3(kBlock), 0, c->stream, c->ke_part,
                     B.ntiles, B.nb, 1, c->ke_out);
  timer_end(c, kTEnergy);
}

void launch_fx_faces(emg_ctx *c) {
  FxBuckets &B = c->fx;
  const int64_t nf = face_len(c->g);
  launch_zero(c->face, sizeof(unsigned long long) * nf * c->sp.nspec, c->stream);
  FxArgs a = make_args(c);
  unsigned grid = (unsigned)(B.nb + c->sp.nspec * kSpillBlocksPerSpecies);
  switch (B.tile) {
    case 4:
      hipLaunchKernelGGL(k_fx_faces<4>, dim3(grid), dim3(kBlock), 0, c->stream, a,
                         (unsigned long long *)c->face, nf);
      break;
    case 8:
      hipLaunchKernelGGL(k_fx_faces<8>, dim3(grid), dim3(kBlock), 0, c->stream, a,
                         (unsigned long long *)c->face, nf);
      break;
  }
}

// sum of the fused push partials (one per bucket)
void fx_energy_from_push(emg_ctx *c) {
  FxBuckets &B = c->fx;
  hipLaunchKernelGGL(k_ke_final, dim3(c->sp.nspec), dim3(kBlock), 0, c->stream, c->ke_part,
                     B.ntiles, B.nb, 0, c->ke_out);
}

}  // namespace emg
