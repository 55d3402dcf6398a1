// k_adj.hip -- reverse (adjoint) step of the FP64 electrostatic push, for the
// differentiable window of the torch harness.
//
// Forward step k (emg_push_es in EMG_PLAIN_FP64, psolve2.F90:137-318) maps
// (x_k, v_k) to
//   a      = qmp * E(x_k)                      (trilinear gather of nodal E)
//   v_k+1  = M (v_k + a) + a                   (Boris; M the rotation by b = qmp B)
//   x_k+1  = x_k + ustep * v_k+1               (periodic wrap: a translation)
// With lambda_x, lambda_v the adjoints of (x_k+1, v_k+1), the adjoints of
// (x_k, v_k) are
//   w        = lambda_v + ustep * lambda_x
//   lambda_a = M^T w + w,   lambda_v <- M^T w  (M^T is the rotation by -b)
//   lambda_x <- lambda_x + sum_c lambda_a_c * qmp * grad E_c(x_k)
// and qmp * lambda_a is scattered with the same trilinear weights onto a
// nodal adjoint field (the transpose of the gather), from which the host
// forms the adjoint of rho through the field solve; the contribution of
// x_k to rho is then added by k_adj_density: lambda_x += q * grad g(x_k).
// B enters as a constant (the electrostatic path keeps B static; the
// gradient of a non-uniform B is not part of the step adjoint).
#include "emg_internal.h"

namespace emg {
namespace {

__device__ __forceinline__ int64_t nodal(const Grid &g, int i, int j, int k) {
  return ((int64_t)k * (g.ny + 1) + j) * (g.nx + 1) + i;
}

__device__ __forceinline__ int64_t periodic(const Grid &g, int i, int j, int k) {
  return ((int64_t)(k == g.nz ? 0 : k) * g.ny + (j == g.ny ? 0 : j)) * g.nx + (i == g.nx ? 0 : i);
}

struct Tri {
  int i, j, k;
  double wx[2], wy[2], wz[2];
};

// cell and weights of the reference's trilinear interpolation (int() cells,
// psolve2.F90:154-183), positions in [0, L)
__device__ __forceinline__ Tri tri(double x, double y, double z, const Grid &g) {
  Tri t;
  t.i = min((int)x, g.nx - 1);
  t.j = min((int)y, g.ny - 1);
  t.k = min((int)z, g.nz - 1);
  double x1 = x - t.i, y1 = y - t.j, z1 = z - t.k;
  t.wx[0] = 1.0 - x1;
  t.wx[1] = x1;
  t.wy[0] = 1.0 - y1;
  t.wy[1] = y1;
  t.wz[0] = 1.0 - z1;
  t.wz[1] = z1;
  return t;
}

// gradient of the trilinear interpolant of corner values s[c] (c = a + 2b + 4c')
__device__ __forceinline__ void tri_grad(const Tri &t, const double s[8], double gr[3]) {
  gr[0] = gr[1] = gr[2] = 0.0;
#pragma unroll
  for (int q = 0; q < 8; q++) {
    int a = q & 1, b = (q >> 1) & 1, c = q >> 2;
    double sx = a ? 1.0 : -1.0, sy = b ? 1.0 : -1.0, sz = c ? 1.0 : -1.0;
    gr[0] += s[q] * sx * t.wy[b] * t.wz[c];
    gr[1] += s[q] * t.wx[a] * sy * t.wz[c];
    gr[2] += s[q] * t.wx[a] * t.wy[b] * sz;
  }
}

// rotation u -> u + bor ((u + u x b) x b) of the Boris push; with -b it is
// the transpose of the rotation with b
__device__ __forceinline__ void rotate(double u[3], const double b[3]) {
  double bor = 2.0 / (1.0 + b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
  double tx = u[0] + u[1] * b[2] - u[2] * b[1];
  double ty = u[1] + u[2] * b[0] - u[0] * b[2];
  double tz = u[2] + u[0] * b[1] - u[1] * b[0];
  u[0] += bor * (ty * b[2] - tz * b[1]);
  u[1] += bor * (tz * b[0] - tx * b[2]);
  u[2] += bor * (tx * b[1] - ty * b[0]);
}

__global__ __launch_bounds__(kBlock) void k_adj_step64(State64 s, State64 lam, SpeciesTab sp,
                                                      const double *__restrict__ node, Grid g,
                                                      int ustep, double *__restrict__ adj_e) {
  int is = blockIdx.y;
  int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= sp.cnt[is]) return;
  m += sp.off[is];
  const double gustep = (double)ustep;
  const double qmp = sp.qm[is] * gustep * 0.5;
  Tri t = tri(s.x[m], s.y[m], s.z[m], g);
  // nodal E/B at the eight corners
  double e[8][3], bf[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int q = 0; q < 8; q++) {
    int a = q & 1, b = (q >> 1) & 1, c = q >> 2;
    const double *p = node + 8 * nodal(g, t.i + a, t.j + b, t.k + c);
    double w = t.wx[a] * t.wy[b] * t.wz[c];
    e[q][0] = p[0];
    e[q][1] = p[1];
    e[q][2] = p[2];
    bf[0] += w * p[3];
    bf[1] += w * p[4];
    bf[2] += w * p[5];
  }
  double nb[3] = {-bf[0] * qmp, -bf[1] * qmp, -bf[2] * qmp};
  double lx[3] = {lam.x[m], lam.y[m], lam.z[m]};
  double w3[3] = {lam.vx[m] + gustep * lx[0], lam.vy[m] + gustep * lx[1],
                  lam.vz[m] + gustep * lx[2]};
  double mt[3] = {w3[0], w3[1], w3[2]};
  rotate(mt, nb);
  double la[3] = {mt[0] + w3[0], mt[1] + w3[1], mt[2] + w3[2]};
  // lambda_x += qmp * grad (lambda_a . E)(x_k)
  double sdot[8];
#pragma unroll
  for (int q = 0; q < 8; q++) sdot[q] = la[0] * e[q][0] + la[1] * e[q][1] + la[2] * e[q][2];
  double gr[3];
  tri_grad(t, sdot, gr);
  lam.x[m] = lx[0] + qmp * gr[0];
  lam.y[m] = lx[1] + qmp * gr[1];
  lam.z[m] = lx[2] + qmp * gr[2];
  lam.vx[m] = mt[0];
  lam.vy[m] = mt[1];
  lam.vz[m] = mt[2];
  // transpose of the gather: qmp * lambda_a onto the periodic nodes
#pragma unroll
  for (int q = 0; q < 8; q++) {
    int a = q & 1, b = (q >> 1) & 1, c = q >> 2;
    double w = qmp * t.wx[a] * t.wy[b] * t.wz[c];
    double *d = adj_e + 3 * periodic(g, t.i + a, t.j + b, t.k + c);
    atomicAdd(d + 0, w * la[0]);
    atomicAdd(d + 1, w * la[1]);
    atomicAdd(d + 2, w * la[2]);
  }
}

// lambda_x += qfac[is] * grad g(x) for the trilinear interpolant of the
// periodic nodal scalar g (the adjoint of the particle's CIC charge)
__global__ __launch_bounds__(kBlock) void k_adj_density64(State64 s, State64 lam, SpeciesTab sp,
                                                         Grid g, const double *__restrict__ gn,
                                                         const double *__restrict__ qfac) {
  int is = blockIdx.y;
  double f = qfac[is];
  if (f == 0.0) return;
  int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= sp.cnt[is]) return;
  m += sp.off[is];
  Tri t = tri(s.x[m], s.y[m], s.z[m], g);
  double v[8];
#pragma unroll
  for (int q = 0; q < 8; q++) {
    int a = q & 1, b = (q >> 1) & 1, c = q >> 2;
    v[q] = gn[periodic(g, t.i + a, t.j + b, t.k + c)];
  }
  double gr[3];
  tri_grad(t, v, gr);
  lam.x[m] += f * gr[0];
  lam.y[m] += f * gr[1];
  lam.z[m] += f * gr[2];
}

// per-species sums of lambda_v (block partials added atomically)
__global__ __launch_bounds__(kBlock) void k_adj_vsum(State64 lam, SpeciesTab sp, double *out) {
  __shared__ double red[3][kBlock];
  int is = blockIdx.y;
  int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double v[3] = {0.0, 0.0, 0.0};
  if (m < sp.cnt[is]) {
    m += sp.off[is];
    v[0] = lam.vx[m];
    v[1] = lam.vy[m];
    v[2] = lam.vz[m];
  }
  for (int d = 0; d < 3; d++) red[d][threadIdx.x] = v[d];
  __syncthreads();
  for (int o = kBlock / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o)
      for (int d = 0; d < 3; d++) red[d][threadIdx.x] += red[d][threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x < 3) atomicAdd(&out[3 * is + threadIdx.x], red[threadIdx.x][0]);
}

dim3 species_blocks(const SpeciesTab &sp) {
  int64_t mx = 1;
  for (int s = 0; s < sp.nspec; s++) mx = sp.cnt[s] > mx ? sp.cnt[s] : mx;
  return dim3((unsigned)((mx + kBlock - 1) / kBlock), sp.nspec);
}

}  // namespace

void launch_adj_step64(emg_ctx *c, int ustep) {
  launch_zero(c->adj_e, sizeof(double) * 3 * c->ncell, c->stream);
  hipLaunchKernelGGL(k_adj_step64, species_blocks(c->sp), dim3(kBlock), 0, c->stream, c->s64,
                     c->lam, c->sp, c->node64, c->g, ustep, c->adj_e);
}

void launch_adj_density64(emg_ctx *c, const double *qfac_dev) {
  hipLaunchKernelGGL(k_adj_density64, species_blocks(c->sp), dim3(kBlock), 0, c->stream, c->s64,
                     c->lam, c->sp, c->g, c->adj_g, qfac_dev);
}

void launch_adj_vsum(emg_ctx *c, double *out_dev) {
  launch_zero(out_dev, sizeof(double) * 3 * c->sp.nspec, c->stream);
  hipLaunchKernelGGL(k_adj_vsum, species_blocks(c->sp), dim3(kBlock), 0, c->stream, c->lam, c->sp,
                     out_dev);
}

}  // namespace emg
