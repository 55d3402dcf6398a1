/*
 * oracle.h -- CPU FP64 restatement of the EMSES particle hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (ctca_amd/, include/) may
 * link, load or call this code; only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py use it, as the checker / CPU baseline.
 *
 * Every function restates one routine of the reference (sakatataisei1226/ctca,
 * MPIEMSES3D) for a single periodic MPI rank, in the reference's own operation
 * order, so that a GPU kernel compiled without FP contraction can be compared
 * bit for bit.  Array conventions follow the Fortran arrays they restate:
 * column-major, component index fastest, explicit lower bounds.
 *
 * Parity pins (see tests/test_oracle_*.py):
 *   mt / ranu0 / rann0  : oracle/_ref (reference mt19937arm.c built from its
 *                         own source) bit for bit, plus the published MT19937
 *                         known answer for seed 5489.
 *   inipcl              : moments + draw order against _ref's stream.
 *   grdcrt              : linear fields relocate exactly; brute force.
 *   psolve2 / positn    : E-only closed form, B-only |v| conservation and
 *                         gyro phase, crossed-field drift, kick reversibility,
 *                         CIC partition of unity / linear-field exactness.
 *   charge + fold       : deposit-sum = sum q, hand-computed single particle,
 *                         periodic translation invariance, gather/deposit
 *                         transpose identity.
 *   psolve1 + fold      : discrete continuity rho_new - rho_old = -ustep div J
 *                         (the reference's own rhoupd identity).
 *   energy              : cold-plasma closed form.
 *   field_energy        : single-mode closed forms (A^2 / 4, cs B^2 / 4).
 *   chgntr              : uniform value -sum(q npin)/ncells.
 *   fields.py (Yee)     : div B kept, Gauss's law kept, and a vacuum plane wave
 *                         that follows the Yee dispersion sin(w/2) = 2 cv sin(k/2).
 */
#ifndef EMSES_ORACLE_H
#define EMSES_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* 64-byte particle record, lib/ohhelp/ohhelp-1.1.1/oh_part.h:23-28 */
typedef struct {
  double x, y, z, vx, vy, vz;
  int32_t pid, preside, nid, spec;
} orc_particle;

/* ---- MT19937 + RANU0 / RANN0 (lib/mtarm/mt19937arm.c:57-171,
 *      src/wrapper/wrapper.F90:20-345) ---- */
typedef struct {
  uint32_t s[624];
  int idx;
} orc_mt;

void orc_mt_seed(orc_mt *g, uint32_t seed);
uint32_t orc_mt_u32(orc_mt *g);
double orc_mt_res53(orc_mt *g);
void orc_ranu0(orc_mt *g, double *out, int64_t n);
void orc_rann0(orc_mt *g, double am, double sd, double *out, int64_t n);

/* ---- initial state (src/main/inital.F90:703-737, src/pcl/inipcl.F90:71-358) */
typedef struct {
  int64_t npin;      /* particles of this species                    */
  double peth, path; /* perpendicular / parallel thermal speeds      */
  double spe, speth; /* perpendicular drift speed, angle (deg)       */
  double spa;        /* parallel drift                               */
  double vdri, vdthz, vdthxy; /* rotated drift magnitude and angles  */
  double vdx, vdy, vdz;       /* cartesian drift                      */
} orc_species_init;

/* inital's RNG self test: seed, then one RANU0(1) and one RANN0(1). */
void orc_rng_start(orc_mt *g, uint32_t seed);
/* inipcl for one species, EMSES-U units (before rescal).  t = T(phiz,phixy)
 * row-major 3x3 from inital.F90:116-140. */
void orc_inipcl_species(orc_mt *g, orc_particle *p, const orc_species_init *s,
                        int spec, const int n[3], const double t[9]);

/* ---- grdcrt (src/field/grdcrt.F90:125-230) ----
 * eb(6, elo:ehi, ...) staggered Yee fields; wrk(6, -4:n+4)^3 nodal output,
 * written for nodes 0..n. */
void orc_grdcrt(const double *eb, const int elo[3], const int ehi[3],
                const int n[3], const double e0[3], const double b0[3],
                double *wrk);

/* ---- positn (src/pcl/positn.F90:117-254), periodic branch ---- */
void orc_positn(orc_particle *p, int64_t np, int ustep, const double sl[3]);

/* ---- psolve2 (src/pcl2field/psolve2.F90:102-433), periodic branch ----
 * counts[nspec] are species segment lengths (pbuf layout). */
void orc_psolve2(orc_particle *p, const int64_t *counts, int nspec,
                 const double *qm, const double *wrk, const int n[3],
                 int ustep, const double sl[3]);

/* ---- charge (src/pcl2field/charge.F90:4-210), func=0 ----
 * rho(1, -1:n+1)^3, zeroed by the call. rhodg (optional, may be NULL):
 * rhodg(nspec, -1:n+1)^3 per-species densities (func=2 path). */
void orc_charge(const orc_particle *p, const int64_t *counts, int nspec,
                const double *q, const int n[3], double *rho, double *rhodg);

/* ---- oh3_exchange_borders + add_boundary_charge for one periodic rank
 * (ohhelp3.c:1010-1031, charge.F90:214-412), ncomp interleaved components. */
void orc_fold_rho(double *rho, int ncomp, const int n[3]);

/* ---- add_background_charge (charge.F90:414-449) with chgntr's uniform
 * neutraliser (src/pcl/chgntr.F90:69-190). rhobk(1,-1:n+1)^3 output. */
void orc_chgntr(const double *q, const int64_t *npin, int nspec,
                const int n[3], double *rhobk);
void orc_add_background(double *rho, const double *rhobk, const int n[3]);

/* ---- energy (src/etc/energy.F90:30-71) ----
 * engk[3*nspec] kinetic, engeb[2] field energy (eb as in orc_grdcrt). */
void orc_energy(const orc_particle *p, const int64_t *counts, int nspec,
                const double *rm, const double sl[3], double *engk);
void orc_field_energy(const double *eb, const int elo[3], const int ehi[3],
                      const int n[3], double cs, const double sl[3],
                      double *engeb);

/* ---- psolve1 (src/pcl2field/psolve1.F90:114-817), periodic branch ----
 * aj(3, -4:n+4)^3 accumulated (aj += q * zigzag current per species). */
void orc_psolve1(orc_particle *p, const int64_t *counts, int nspec,
                 const double *qm, const double *q, const double *wrk,
                 const int n[3], int ustep, const double sl[3], double *aj);
/* oh3_exchange_borders(CAJ) + add_boundary_current (psolve1.F90:859-1079) */
void orc_fold_current(double *aj, const int n[3]);
/* rhoupd (src/field/rhoupd.F90:40-49): rho(1,-1:n+1) -= 2 div aj */
void orc_rhoupd(double *rho, const double *aj, const int n[3]);

/* reference-order particle sort comparator result, psort-struct.c:15-20 */
int orc_psort_key_cmp(const orc_particle *a, const orc_particle *b);

#ifdef __cplusplus
}
#endif
#endif
