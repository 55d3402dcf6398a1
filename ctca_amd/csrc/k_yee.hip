// k_yee.hip -- the EM field update on the device, on the library's copy of
// the staggered Yee array eb(6, lo:hi) (the layout emg_set_fields uploads,
// component fastest, ghosts included), for one periodic rank:
//   bfield  (bfield.F90:38-52): B += curl-difference of E (a half step);
//   efield  (efield.F90:55-72): E += tcs curl-difference of B - dlt aj, with
//           aj the folded current of the last emg_push_em;
//   energy  (energy.F90:52-70): the field part engeb.
// Every point of the array, ghosts included, is computed from the periodic
// images of its neighbours: a ghost gets exactly the value of its image,
// which is what the reference's border exchange gives it after the update on
// one periodic rank.  The updates read only the other field, so they run in
// place.  Expression order as in the reference; built without FP contraction.
#include "emg_internal.h"

namespace emg {
namespace {

struct EbView {
  double *eb;
  int lo[3], ext[3];
  Grid g;
  __device__ __forceinline__ int64_t at(int c, int i, int j, int k) const {
    return (((int64_t)(k - lo[2]) * ext[1] + (j - lo[1])) * ext[0] + (i - lo[0])) * 6 + c;
  }
  // periodic image of cell (i, j, k) inside the stored box
  __device__ __forceinline__ int64_t img(int c, int i, int j, int k) const {
    return at(c, pm(i, g.nx), pm(j, g.ny), pm(k, g.nz));
  }
  // index modulo n for the indices the kernels form (a few cells around
  // [0, n)): one compare and add each way, the division only if that misses
  __device__ __forceinline__ static int pm(int i, int n) {
    int r = i < 0 ? i + n : (i >= n ? i - n : i);
    if ((unsigned)r >= (unsigned)n) r = ((i % n) + n) % n;
    return r;
  }
};

constexpr int BX = 3, BY = 4, BZ = 5;  // eb components: EX EY EZ BX BY BZ

// A point's record is 6 doubles (48 bytes, 16-byte aligned): its E or B
// triple is read and written as one 16-byte and one 8-byte access.
struct D3 {
  double x, y, z;
};
__device__ __forceinline__ D3 ld_e(const double *p) {
  const double2 a = *reinterpret_cast<const double2 *>(p);
  return {a.x, a.y, p[2]};
}
__device__ __forceinline__ D3 ld_b(const double *p) {
  const double2 a = *reinterpret_cast<const double2 *>(p + 4);
  return {p[3], a.x, a.y};
}

__global__ __launch_bounds__(kBlock) void k_bfield(EbView v) {
  const int64_t n = (int64_t)v.ext[0] * v.ext[1] * v.ext[2];
  const int64_t o = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (o >= n) return;
  const int i = (int)(o % v.ext[0]) + v.lo[0];
  const int j = (int)((o / v.ext[0]) % v.ext[1]) + v.lo[1];
  const int k = (int)(o / ((int64_t)v.ext[0] * v.ext[1])) + v.lo[2];
  double *p = v.eb + v.at(0, i, j, k);
  const D3 e0 = ld_e(v.eb + v.img(0, i, j, k)), ek = ld_e(v.eb + v.img(0, i, j, k + 1)),
           ej = ld_e(v.eb + v.img(0, i, j + 1, k)), ei = ld_e(v.eb + v.img(0, i + 1, j, k));
  const D3 b = ld_b(p);
  const double bx = b.x + ek.y - e0.y - ej.z + e0.z;
  const double by = b.y + ei.z - e0.z - ek.x + e0.x;
  const double bz = b.z + ej.x - e0.x - ei.y + e0.y;
  p[BX] = bx;
  *reinterpret_cast<double2 *>(p + BY) = make_double2(by, bz);
}

struct CurrentView {
  const void *jacc;  // nspec x 3 x ncell: int64 fixed point (FIX) or FP64 sums
  double q[EMG_MAX_SPECIES];
  int nspec;
  int64_t ncell;
};

// aj(d) of periodic cell c: sum over species of J(is) q(is), in species order,
// each term added with one fused multiply-add (the read-out of
// emg_get_current, so the E update takes exactly the current it reports)
template <bool FIX>
__device__ __forceinline__ double aj_at(const CurrentView &J, int d, int64_t c) {
  double v = 0.0;
  for (int s = 0; s < J.nspec; s++) {
    const int64_t q = ((int64_t)s * 3 + d) * J.ncell + c;
    const double js = FIX ? (double)(long long)((const unsigned long long *)J.jacc)[q] * (1.0 / kJScale)
                          : ((const double *)J.jacc)[q];
    v = fma(js, J.q[s], v);
  }
  return v;
}

template <bool FIX>
__global__ __launch_bounds__(kBlock) void k_efield(EbView v, CurrentView J, double tcs,
                                                   double dlt) {
  const int64_t n = (int64_t)v.ext[0] * v.ext[1] * v.ext[2];
  const int64_t o = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (o >= n) return;
  const int i = (int)(o % v.ext[0]) + v.lo[0];
  const int j = (int)((o / v.ext[0]) % v.ext[1]) + v.lo[1];
  const int k = (int)(o / ((int64_t)v.ext[0] * v.ext[1])) + v.lo[2];
  double *p = v.eb + v.at(0, i, j, k);
  const int64_t c = ((int64_t)EbView::pm(k, v.g.nz) * v.g.ny + EbView::pm(j, v.g.ny)) * v.g.nx +
                    EbView::pm(i, v.g.nx);
  const D3 b0 = ld_b(v.eb + v.img(0, i, j, k)), bj = ld_b(v.eb + v.img(0, i, j - 1, k)),
           bk = ld_b(v.eb + v.img(0, i, j, k - 1)), bi = ld_b(v.eb + v.img(0, i - 1, j, k));
  const D3 e = ld_e(p);
  const double ex = e.x + tcs * (b0.z - bj.z - b0.y + bk.y) - dlt * aj_at<FIX>(J, 0, c);
  const double ey = e.y + tcs * (b0.x - bk.x - b0.z + bi.z) - dlt * aj_at<FIX>(J, 1, c);
  const double ez = e.z + tcs * (b0.y - bi.y - b0.x + bj.x) - dlt * aj_at<FIX>(J, 2, c);
  *reinterpret_cast<double2 *>(p) = make_double2(ex, ey);
  p[2] = ez;
}

// sum E^2 and B^2 over the cells 0..n-1: per-block partials, then one block
// adds them in block order (deterministic)
constexpr int kFeBlocks = 256;

__global__ __launch_bounds__(kBlock) void k_field_energy_part(EbView v, double *part) {
  __shared__ double red[2][kBlock];
  const int64_t n = (int64_t)v.g.nx * v.g.ny * v.g.nz;
  double te = 0.0, tb = 0.0;
  for (int64_t o = (int64_t)blockIdx.x * kBlock + threadIdx.x; o < n;
       o += (int64_t)kFeBlocks * kBlock) {
    const int i = (int)(o % v.g.nx), j = (int)((o / v.g.nx) % v.g.ny),
              k = (int)(o / ((int64_t)v.g.nx * v.g.ny));
    const double *p = v.eb + v.at(0, i, j, k);
    te = te + p[0] * p[0] + p[1] * p[1] + p[2] * p[2];
    tb = tb + p[BX] * p[BX] + p[BY] * p[BY] + p[BZ] * p[BZ];
  }
  red[0][threadIdx.x] = te;
  red[1][threadIdx.x] = tb;
  __syncthreads();
  for (int o = kBlock / 2; o > 0; o >>= 1) {
    if (threadIdx.x < (unsigned)o) {
      red[0][threadIdx.x] += red[0][threadIdx.x + o];
      red[1][threadIdx.x] += red[1][threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = red[0][0];
    part[2 * blockIdx.x + 1] = red[1][0];
  }
}

__global__ void k_field_energy_final(const double *part, double cs, double vol, double *out) {
  if (threadIdx.x != 0) return;
  double te = 0.0, tb = 0.0;
  for (int b = 0; b < kFeBlocks; b++) {
    te += part[2 * b];
    tb += part[2 * b + 1];
  }
  out[0] = 0.5 * te / vol;
  out[1] = 0.5 * cs * tb / vol;
}

// copy the box lo .. lo + ext - 1 of eb (inside the stored bounds) into a
// dense out(6, box)
__global__ __launch_bounds__(kBlock) void k_eb_box(EbView v, int lx, int ly, int lz, int ex, int ey,
                                                   int ez, double *out) {
  const int64_t n = (int64_t)ex * ey * ez;
  const int64_t o = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (o >= n) return;
  const int i = (int)(o % ex) + lx, j = (int)((o / ex) % ey) + ly,
            k = (int)(o / ((int64_t)ex * ey)) + lz;
  const double *p = v.eb + v.at(0, i, j, k);
#pragma unroll
  for (int q = 0; q < 6; q++) out[o * 6 + q] = p[q];
}

EbView view_of(emg_ctx *c) {
  EbView v;
  v.eb = c->eb_dev;
  for (int d = 0; d < 3; d++) {
    v.lo[d] = c->eb_lo[d];
    v.ext[d] = c->eb_ext[d];
  }
  v.g = c->g;
  return v;
}

unsigned blocks_of(const EbView &v) {
  const int64_t n = (int64_t)v.ext[0] * v.ext[1] * v.ext[2];
  return (unsigned)((n + kBlock - 1) / kBlock);
}

}  // namespace

void launch_bfield(emg_ctx *c) {
  const EbView v = view_of(c);
  hipLaunchKernelGGL(k_bfield, dim3(blocks_of(v)), dim3(kBlock), 0, c->stream, v);
}

void launch_efield(emg_ctx *c, double tcs, double dlt) {
  const EbView v = view_of(c);
  CurrentView J;
  const bool fix = c->cfg.precision == EMG_FP32_FIXED;
  J.jacc = fix ? (const void *)c->jfx : (const void *)c->j64;
  for (int s = 0; s < EMG_MAX_SPECIES; s++) J.q[s] = c->cfg.q[s];
  J.nspec = c->sp.nspec;
  J.ncell = c->ncell;
  if (fix)
    hipLaunchKernelGGL(k_efield<true>, dim3(blocks_of(v)), dim3(kBlock), 0, c->stream, v, J, tcs,
                       dlt);
  else
    hipLaunchKernelGGL(k_efield<false>, dim3(blocks_of(v)), dim3(kBlock), 0, c->stream, v, J, tcs,
                       dlt);
}

void launch_eb_box(emg_ctx *c, const int lo[3], const int ext[3], double *out_dev) {
  const EbView v = view_of(c);
  const int64_t n = (int64_t)ext[0] * ext[1] * ext[2];
  hipLaunchKernelGGL(k_eb_box, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                     c->stream, v, lo[0], lo[1], lo[2], ext[0], ext[1], ext[2], out_dev);
}

int64_t field_energy_part_need() { return 2 * kFeBlocks; }

void launch_field_energy(emg_ctx *c, double cs, double *part_dev, double *out_dev) {
  const EbView v = view_of(c);
  hipLaunchKernelGGL(k_field_energy_part, dim3(kFeBlocks), dim3(kBlock), 0, c->stream, v, part_dev);
  const double vol = (double)c->g.slx * c->g.sly * c->g.slz;
  hipLaunchKernelGGL(k_field_energy_final, dim3(1), dim3(64), 0, c->stream, part_dev, cs, vol,
                     out_dev);
}

}  // namespace emg
