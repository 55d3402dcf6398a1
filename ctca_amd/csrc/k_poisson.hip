// k_poisson.hip -- the electrostatic field solve on the device, for one
// periodic rank, so an ES step can stay on the device:
//   rho    = the deposited charge of every species + the background: the
//            per-node array of emg_set_background, if any, then the uniform
//            value (charge.F90:414-449 add_background_charge);
//   phi    : -lap7 phi = rho, spectrally: the 7-point Laplacian has the
//            eigenvalues kappa = sum_d (2 sin(pi k_d / n_d))^2
//            (inifft.F90:390-404), the zero mode is dropped
//            (poissn.F90:398-410); real-to-complex FFTs from hipFFT;
//   E      = -grad phi on the staggered points (esfild.F90:351-367):
//            EX(i) = phi(i) - phi(i+1), etc., written into the device eb
//            (every point, ghosts as periodic images; B untouched);
//   then the relocation to the nodes (grdcrt).
// The FFT round-off differs from FFTW's (which the reference uses); the solve
// is pinned by the spectral closed form and the 7-point residual.
#include <hipfft/hipfft.h>

#include <cmath>
#include <new>
#include <string>

#include "emg_internal.h"

namespace emg {

struct EsSolver {
  hipfftHandle fwd = 0, inv = 0;
  bool have_fwd = false, have_inv = false;
  double *rho = nullptr;    // nz ny nx, then phi in place of it
  double2 *spec = nullptr;  // nz ny (nx/2 + 1)
};

namespace {

__global__ __launch_bounds__(kBlock) void k_rho_total(const double *rho64,
                                                      const unsigned long long *rhofx,
                                                      SpeciesTab sp, int64_t ncell,
                                                      const double *bkarr, double bk,
                                                      double *out) {
  const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (c >= ncell) return;
  double v = 0.0;
  for (int s = 0; s < sp.nspec; s++) {
    if (rho64)
      v = v + rho64[(int64_t)s * ncell + c];
    else
      v = v + sp.q[s] * ((double)(long long)rhofx[(int64_t)s * ncell + c] * (1.0 / kFixScale));
  }
  if (bkarr) v = v + bkarr[c];
  out[c] = v + bk;
}

// spectrum / (N kappa), the zero mode set to 0 (N: the inverse transform is
// unnormalised)
__global__ __launch_bounds__(kBlock) void k_poisson_divide(double2 *spec, Grid g) {
  const int hx = g.nx / 2 + 1;
  const int64_t n = (int64_t)g.nz * g.ny * hx;
  const int64_t o = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (o >= n) return;
  const int kx = (int)(o % hx), ky = (int)((o / hx) % g.ny), kz = (int)(o / ((int64_t)hx * g.ny));
  const double pi = 3.141592653589793;
  const double sx = 2.0 * sin(pi * kx / g.nx), sy = 2.0 * sin(pi * ky / g.ny),
               sz = 2.0 * sin(pi * kz / g.nz);
  const double kappa = sz * sz + sy * sy + sx * sx;
  double2 v = spec[o];
  if (o == 0) {
    v.x = 0.0;
    v.y = 0.0;
  } else {
    const double f = 1.0 / (kappa * ((double)g.nx * g.ny * g.nz));
    v.x *= f;
    v.y *= f;
  }
  spec[o] = v;
}

// index modulo n: one compare and add each way, the division only if needed
__device__ __forceinline__ int pm(int i, int n) {
  int r = i < 0 ? i + n : (i >= n ? i - n : i);
  if ((unsigned)r >= (unsigned)n) r = ((i % n) + n) % n;
  return r;
}

__global__ __launch_bounds__(kBlock) void k_es_efield(const double *phi, double *eb, int lx,
                                                      int ly, int lz, int ex, int ey, int ez,
                                                      Grid g) {
  const int64_t n = (int64_t)ex * ey * ez;
  const int64_t o = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (o >= n) return;
  const int i = (int)(o % ex) + lx, j = (int)((o / ex) % ey) + ly,
            k = (int)(o / ((int64_t)ex * ey)) + lz;
  auto P = [&](int a, int b, int c) {
    return phi[((int64_t)pm(c, g.nz) * g.ny + pm(b, g.ny)) * g.nx + pm(a, g.nx)];
  };
  const double p0 = P(i, j, k);
  double *e = eb + o * 6;
  *reinterpret_cast<double2 *>(e) = make_double2(p0 - P(i + 1, j, k), p0 - P(i, j + 1, k));
  e[2] = p0 - P(i, j, k + 1);
}

}  // namespace

static int fft_check(hipfftResult r, const char *what) {
  if (r == HIPFFT_SUCCESS) return EMG_OK;
  return set_error(EMG_ERR_HIP, std::string(what) + ": hipFFT error " + std::to_string((int)r));
}

void es_solver_free(emg_ctx *c) {
  EsSolver *s = c->es;
  if (!s) return;
  if (s->have_fwd) (void)hipfftDestroy(s->fwd);
  if (s->have_inv) (void)hipfftDestroy(s->inv);
  if (s->rho) (void)hipFree(s->rho);
  if (s->spec) (void)hipFree(s->spec);
  delete s;
  c->es = nullptr;
}

static int es_solver_get(emg_ctx *c, EsSolver **out) {
  if (c->es) {
    *out = c->es;
    return EMG_OK;
  }
  EsSolver *s = new (std::nothrow) EsSolver();
  if (!s) return set_error(EMG_ERR_NOMEM, "field solver");
  c->es = s;
  const int64_t hx = c->g.nx / 2 + 1;
  int rc = EMG_OK;
  if (hipMalloc(&s->rho, sizeof(double) * c->ncell) != hipSuccess ||
      hipMalloc(&s->spec, sizeof(double2) * c->g.nz * c->g.ny * hx) != hipSuccess)
    rc = set_error(EMG_ERR_NOMEM, "field solver arrays");
  if (!rc) rc = fft_check(hipfftPlan3d(&s->fwd, c->g.nz, c->g.ny, c->g.nx, HIPFFT_D2Z), "plan D2Z");
  if (!rc) s->have_fwd = true;
  if (!rc) rc = fft_check(hipfftPlan3d(&s->inv, c->g.nz, c->g.ny, c->g.nx, HIPFFT_Z2D), "plan Z2D");
  if (!rc) s->have_inv = true;
  if (!rc) rc = fft_check(hipfftSetStream(s->fwd, c->stream), "set stream");
  if (!rc) rc = fft_check(hipfftSetStream(s->inv, c->stream), "set stream");
  if (rc) {
    es_solver_free(c);
    return rc;
  }
  *out = s;
  return EMG_OK;
}

int solve_es(emg_ctx *c, double rhobk) {
  EsSolver *s = nullptr;
  int rc = es_solver_get(c, &s);
  if (rc) return rc;
  const unsigned gc = (unsigned)((c->ncell + kBlock - 1) / kBlock);
  const bool fix = c->cfg.precision == EMG_FP32_FIXED;
  hipLaunchKernelGGL(k_rho_total, dim3(gc), dim3(kBlock), 0, c->stream, fix ? nullptr : c->rho64,
                     fix ? c->rhofx : nullptr, c->sp, c->ncell, c->bk_dev, rhobk, s->rho);
  rc = fft_check(hipfftExecD2Z(s->fwd, s->rho, reinterpret_cast<hipfftDoubleComplex *>(s->spec)),
                 "D2Z");
  if (rc) return rc;
  const int64_t ns = (int64_t)c->g.nz * c->g.ny * (c->g.nx / 2 + 1);
  hipLaunchKernelGGL(k_poisson_divide, dim3((unsigned)((ns + kBlock - 1) / kBlock)), dim3(kBlock),
                     0, c->stream, s->spec, c->g);
  rc = fft_check(hipfftExecZ2D(s->inv, reinterpret_cast<hipfftDoubleComplex *>(s->spec), s->rho),
                 "Z2D");
  if (rc) return rc;
  const int64_t ne = (int64_t)c->eb_ext[0] * c->eb_ext[1] * c->eb_ext[2];
  hipLaunchKernelGGL(k_es_efield, dim3((unsigned)((ne + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                     c->stream, s->rho, c->eb_dev, c->eb_lo[0], c->eb_lo[1], c->eb_lo[2],
                     c->eb_ext[0], c->eb_ext[1], c->eb_ext[2], c->g);
  launch_relocate64(c, c->eb_lo, c->eb_ext);
  return hip_check(hipGetLastError(), "field solve launch");
}

int read_potential(emg_ctx *c, double *phi) {
  if (!c->es) return set_error(EMG_ERR_STATE, "no field solve yet");
  int rc = hip_check(hipStreamSynchronize(c->stream), "stream synchronize");
  if (!rc)
    rc = hip_check(hipMemcpy(phi, c->es->rho, sizeof(double) * c->ncell, hipMemcpyDeviceToHost),
                   "potential copy");
  return rc;
}

}  // namespace emg
