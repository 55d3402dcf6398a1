/*
 * oracle.c -- CPU FP64 restatement of the EMSES particle hot path.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Built with -ffp-contract=off so
 * that every expression is evaluated in the reference's source order with
 * separately rounded multiplies and adds.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* MT19937 (Matsumoto & Nishimura 1998, as used by lib/mtarm/mt19937arm.c)   */
/* ------------------------------------------------------------------------ */
enum { MT_N = 624, MT_M = 397 };

/* init_genrand, mt19937arm.c:57-71 */
void orc_mt_seed(orc_mt *g, uint32_t seed) {
  g->s[0] = seed;
  for (int i = 1; i < MT_N; i++) {
    uint32_t prev = g->s[i - 1];
    g->s[i] = 1812433253u * (prev ^ (prev >> 30)) + (uint32_t)i;
  }
  g->idx = MT_N;
}

static void mt_twist(orc_mt *g) {
  for (int i = 0; i < MT_N; i++) {
    uint32_t y = (g->s[i] & 0x80000000u) | (g->s[(i + 1) % MT_N] & 0x7fffffffu);
    uint32_t v = g->s[(i + MT_M) % MT_N] ^ (y >> 1);
    if (y & 1u) v ^= 0x9908b0dfu;
    g->s[i] = v;
  }
  g->idx = 0;
}

/* genrand_int32, mt19937arm.c:102-138 */
uint32_t orc_mt_u32(orc_mt *g) {
  if (g->idx >= MT_N) mt_twist(g);
  uint32_t y = g->s[g->idx++];
  y ^= y >> 11;
  y ^= (y << 7) & 0x9d2c5680u;
  y ^= (y << 15) & 0xefc60000u;
  y ^= y >> 18;
  return y;
}

/* genrand_res53, mt19937arm.c:167-171 (the Fortran mtrand_real8_) */
double orc_mt_res53(orc_mt *g) {
  uint32_t a = orc_mt_u32(g) >> 5;
  uint32_t b = orc_mt_u32(g) >> 6;
  return ((double)a * 67108864.0 + (double)b) * (1.0 / 9007199254740992.0);
}

/* RANU0 array form, wrapper.F90:91-125 */
void orc_ranu0(orc_mt *g, double *out, int64_t n) {
  for (int64_t i = 0; i < n; i++) out[i] = orc_mt_res53(g);
}

/* RANN0 array form, Box-Muller pairs, wrapper.F90:211-276 (and the scalar
 * branch 223-231 used when n == 1). */
void orc_rann0(orc_mt *g, double am, double sd, double *out, int64_t n) {
  const double pei2 = 6.28318530717958646;
  if (n == 1) {
    double w1 = orc_mt_res53(g), w2 = orc_mt_res53(g);
    out[0] = am + sd * sqrt(-2.0 * log(w1)) * cos(pei2 * w2);
    return;
  }
  int64_t id = (n % 2 != 0) ? 1 : 0;
  int64_t nn = n - id;
  orc_ranu0(g, out, nn);
  for (int64_t i = 0; i < nn; i += 2) {
    double w1 = sd * sqrt(-2.0 * log(out[i]));
    double w2 = pei2 * out[i + 1];
    out[i] = am + w1 * cos(w2);
    out[i + 1] = am + w1 * sin(w2);
  }
  if (id == 1) {
    double w1 = orc_mt_res53(g), w2 = orc_mt_res53(g);
    out[n - 1] = am + sd * sqrt(-2.0 * log(w1)) * cos(pei2 * w2);
  }
}

/* inital.F90:703-737: seed (myid+1)*10 then one RANU0 and one RANN0 draw */
void orc_rng_start(orc_mt *g, uint32_t seed) {
  double d;
  orc_mt_seed(g, seed);
  orc_ranu0(g, &d, 1);
  orc_rann0(g, 0.0, 1.0, &d, 1);
}

/* inipcl.F90:77-98 (uniform draws), :99-104 (type_rdist=1 positions),
 * :266-320 (Maxwellian velocities, rotation, drifts), :322-326 (tags). */
void orc_inipcl_species(orc_mt *g, orc_particle *p, const orc_species_init *s,
                        int spec, const int n[3], const double t[9]) {
  const double pi = acos(-1.0);
  int64_t np = s->npin;
  double *buf = (double *)malloc(sizeof(double) * (size_t)(np > 0 ? np : 1));
  orc_ranu0(g, buf, np);
  for (int64_t m = 0; m < np; m++) p[m].vx = buf[m];
  orc_ranu0(g, buf, np);
  for (int64_t m = 0; m < np; m++) p[m].vy = buf[m];
  orc_ranu0(g, buf, np);
  for (int64_t m = 0; m < np; m++) p[m].vz = buf[m];
  for (int64_t m = 0; m < np; m++) {
    p[m].x = (double)n[0] * p[m].vx + 0.0;
    p[m].y = (double)n[1] * p[m].vy + 0.0;
    p[m].z = (double)n[2] * p[m].vz + 0.0;
    p[m].nid = 0;
  }
  orc_rann0(g, s->spe * cos(s->speth / 180.0 * pi), s->peth, buf, np);
  for (int64_t m = 0; m < np; m++) p[m].vx = buf[m];
  orc_rann0(g, s->spe * sin(s->speth / 180.0 * pi), s->peth, buf, np);
  for (int64_t m = 0; m < np; m++) p[m].vy = buf[m];
  orc_rann0(g, s->spa, s->path, buf, np);
  for (int64_t m = 0; m < np; m++) p[m].vz = buf[m];
  for (int64_t m = 0; m < np; m++) {
    double vx = p[m].vx, vy = p[m].vy, vz = p[m].vz;
    double vxt = vx * t[0] + vy * t[1] + vz * t[2];
    double vyt = vx * t[3] + vy * t[4] + vz * t[5];
    double vzt = vx * t[6] + vy * t[7] + vz * t[8];
    vxt = vxt + s->vdri * sin(s->vdthz / 180.0 * pi) * cos(s->vdthxy / 180.0 * pi);
    vyt = vyt + s->vdri * sin(s->vdthz / 180.0 * pi) * sin(s->vdthxy / 180.0 * pi);
    vzt = vzt + s->vdri * cos(s->vdthz / 180.0 * pi);
    p[m].vx = vxt + s->vdx;
    p[m].vy = vyt + s->vdy;
    p[m].vz = vzt + s->vdz;
    p[m].spec = spec;
    p[m].preside = 0;
    p[m].pid = 0;
  }
  free(buf);
}

/* ------------------------------------------------------------------------ */
/* array addressing                                                         */
/* ------------------------------------------------------------------------ */
#define WLO 4 /* wrk / aj lower halo: -4 */
static inline int64_t widx(const int n[3], int c, int nc, int i, int j, int k) {
  int64_t nx = n[0] + 9, ny = n[1] + 9;
  return (int64_t)c + nc * ((int64_t)(i + WLO) + nx * ((int64_t)(j + WLO) + ny * (int64_t)(k + WLO)));
}
static inline int64_t ridx(const int n[3], int c, int nc, int i, int j, int k) {
  int64_t nx = n[0] + 3, ny = n[1] + 3;
  return (int64_t)c + nc * ((int64_t)(i + 1) + nx * ((int64_t)(j + 1) + ny * (int64_t)(k + 1)));
}
static inline int64_t eidx(const int lo[3], const int hi[3], int c, int i, int j, int k) {
  int64_t nx = hi[0] - lo[0] + 1, ny = hi[1] - lo[1] + 1;
  return (int64_t)c + 6 * ((int64_t)(i - lo[0]) + nx * ((int64_t)(j - lo[1]) + ny * (int64_t)(k - lo[2])));
}
enum { EX = 0, EY = 1, EZ = 2, BX = 3, BY = 4, BZ = 5 };

/* grdcrt.F90:158-169 and 216-230 */
void orc_grdcrt(const double *eb, const int lo[3], const int hi[3],
                const int n[3], const double e0[3], const double b0[3],
                double *wrk) {
#define EB(c, i, j, k) eb[eidx(lo, hi, (c), (i), (j), (k))]
  for (int k = 0; k <= n[2]; k++)
    for (int j = 0; j <= n[1]; j++)
      for (int i = 0; i <= n[0]; i++) {
        wrk[widx(n, EX, 6, i, j, k)] = 0.5 * (EB(EX, i, j, k) + EB(EX, i - 1, j, k)) + e0[0];
        wrk[widx(n, EY, 6, i, j, k)] = 0.5 * (EB(EY, i, j, k) + EB(EY, i, j - 1, k)) + e0[1];
        wrk[widx(n, EZ, 6, i, j, k)] = 0.5 * (EB(EZ, i, j, k) + EB(EZ, i, j, k - 1)) + e0[2];
      }
  for (int k = 0; k <= n[2]; k++)
    for (int j = 0; j <= n[1]; j++)
      for (int i = 0; i <= n[0]; i++) {
        wrk[widx(n, BX, 6, i, j, k)] =
            0.25 * (EB(BX, i, j, k) + EB(BX, i, j - 1, k) + EB(BX, i, j, k - 1) + EB(BX, i, j - 1, k - 1)) + b0[0];
        wrk[widx(n, BY, 6, i, j, k)] =
            0.25 * (EB(BY, i, j, k) + EB(BY, i - 1, j, k) + EB(BY, i, j, k - 1) + EB(BY, i - 1, j, k - 1)) + b0[1];
        wrk[widx(n, BZ, 6, i, j, k)] =
            0.25 * (EB(BZ, i, j, k) + EB(BZ, i - 1, j, k) + EB(BZ, i, j - 1, k) + EB(BZ, i - 1, j - 1, k)) + b0[2];
      }
#undef EB
}

/* periodic wrap of one coordinate, psolve2.F90:354-432 (npbnd=0 branch) */
static inline void wrap1(double *x, double sl) {
  if (*x < 0.0)
    *x = *x + sl;
  else if (*x >= sl)
    *x = *x - sl;
}

/* positn.F90:134-139 and the periodic branch of 172-254 */
void orc_positn(orc_particle *p, int64_t np, int ustep, const double sl[3]) {
  double us = (double)ustep;
  for (int64_t m = 0; m < np; m++) {
    p[m].x = p[m].x + p[m].vx * us;
    p[m].y = p[m].y + p[m].vy * us;
    p[m].z = p[m].z + p[m].vz * us;
    if (p[m].x < 0.0 || p[m].x >= sl[0] || p[m].y < 0.0 || p[m].y >= sl[1] ||
        p[m].z < 0.0 || p[m].z >= sl[2]) {
      wrap1(&p[m].x, sl[0]);
      wrap1(&p[m].y, sl[1]);
      wrap1(&p[m].z, sl[2]);
    }
  }
}

/* CIC weights, psolve2.F90:154-183 / charge.F90:160-193 */
typedef struct {
  double v1, v2, v3, v4, v5, v6, v7, v8;
  int i, j, k;
} cicw;

static inline cicw cic_weights(double xl, double yl, double zl) {
  cicw w;
  w.i = (int)xl;
  w.j = (int)yl;
  w.k = (int)zl;
  double x1 = xl - (double)w.i, y1 = yl - (double)w.j, z1 = zl - (double)w.k;
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

/* 8-node gather of one nodal component: node order v1..v8 =
 * (i,j,k1),(i1,j,k1),(i1,j1,k1),(i,j1,k1),(i,j,k),(i1,j,k),(i1,j1,k),(i,j1,k) */
static inline double gather(const double *wrk, const int n[3], int c, const cicw *w) {
  int i = w->i, j = w->j, k = w->k, i1 = i + 1, j1 = j + 1, k1 = k + 1;
  return wrk[widx(n, c, 6, i, j, k1)] * w->v1 + wrk[widx(n, c, 6, i1, j, k1)] * w->v2 +
         wrk[widx(n, c, 6, i1, j1, k1)] * w->v3 + wrk[widx(n, c, 6, i, j1, k1)] * w->v4 +
         wrk[widx(n, c, 6, i, j, k)] * w->v5 + wrk[widx(n, c, 6, i1, j, k)] * w->v6 +
         wrk[widx(n, c, 6, i1, j1, k)] * w->v7 + wrk[widx(n, c, 6, i, j1, k)] * w->v8;
}

/* Buneman-Boris velocity update, psolve2.F90:233-257 */
static inline void boris(orc_particle *q, double eex, double eey, double eez,
                         double bbx, double bby, double bbz, double qmp) {
  const double ewx = 0.0, ewy = 0.0; /* ewmodel = 0 */
  eex = (eex + ewx) * qmp;
  eey = (eey + ewy) * qmp;
  eez = eez * qmp;
  bbx = bbx * qmp;
  bby = bby * qmp;
  bbz = bbz * qmp;
  double bor = 2.0 / (1.0 + bbx * bbx + bby * bby + bbz * bbz);
  q->vx = q->vx + eex;
  q->vy = q->vy + eey;
  q->vz = q->vz + eez;
  double vxt = q->vx + q->vy * bbz - q->vz * bby;
  double vyt = q->vy + q->vz * bbx - q->vx * bbz;
  double vzt = q->vz + q->vx * bby - q->vy * bbx;
  q->vx = q->vx + bor * (vyt * bbz - vzt * bby);
  q->vy = q->vy + bor * (vzt * bbx - vxt * bbz);
  q->vz = q->vz + bor * (vxt * bby - vyt * bbx);
  q->vx = q->vx + eex;
  q->vy = q->vy + eey;
  q->vz = q->vz + eez;
}

/* psolve2.F90:137-433 for preside == OH_PCL_ALIVE, no bodies, periodic */
void orc_psolve2(orc_particle *p, const int64_t *counts, int nspec,
                 const double *qm, const double *wrk, const int n[3],
                 int ustep, const double sl[3]) {
  int64_t m0 = 0;
  for (int is = 0; is < nspec; is++) {
    double gustep = (double)ustep;
    double qmp = qm[is] * gustep * 0.5;
    for (int64_t m = m0; m < m0 + counts[is]; m++) {
      orc_particle *q = &p[m];
      if (q->nid == -1) continue;
      cicw w = cic_weights(q->x - 0.0, q->y - 0.0, q->z - 0.0);
      double eex = gather(wrk, n, EX, &w);
      double eey = gather(wrk, n, EY, &w);
      double eez = gather(wrk, n, EZ, &w);
      double bbx = gather(wrk, n, BX, &w);
      double bby = gather(wrk, n, BY, &w);
      double bbz = gather(wrk, n, BZ, &w);
      boris(q, eex, eey, eez, bbx, bby, bbz, qmp);
      double xmove = q->vx * gustep, ymove = q->vy * gustep, zmove = q->vz * gustep;
      q->x = q->x + xmove;
      q->y = q->y + ymove;
      q->z = q->z + zmove;
      if (q->x < 0.0 || q->x >= sl[0] || q->y < 0.0 || q->y >= sl[1] ||
          q->z < 0.0 || q->z >= sl[2]) {
        wrap1(&q->x, sl[0]);
        wrap1(&q->y, sl[1]);
        wrap1(&q->z, sl[2]);
      }
    }
    m0 += counts[is];
  }
}

/* distribute_charge, charge.F90:197-210 */
static inline void distribute(double *rh, const int n[3], int c, int nc, const cicw *w, double qc) {
  int i = w->i, j = w->j, k = w->k, i1 = i + 1, j1 = j + 1, k1 = k + 1;
  rh[ridx(n, c, nc, i, j, k1)] = rh[ridx(n, c, nc, i, j, k1)] + w->v1 * qc;
  rh[ridx(n, c, nc, i1, j, k1)] = rh[ridx(n, c, nc, i1, j, k1)] + w->v2 * qc;
  rh[ridx(n, c, nc, i1, j1, k1)] = rh[ridx(n, c, nc, i1, j1, k1)] + w->v3 * qc;
  rh[ridx(n, c, nc, i, j1, k1)] = rh[ridx(n, c, nc, i, j1, k1)] + w->v4 * qc;
  rh[ridx(n, c, nc, i, j, k)] = rh[ridx(n, c, nc, i, j, k)] + w->v5 * qc;
  rh[ridx(n, c, nc, i1, j, k)] = rh[ridx(n, c, nc, i1, j, k)] + w->v6 * qc;
  rh[ridx(n, c, nc, i1, j1, k)] = rh[ridx(n, c, nc, i1, j1, k)] + w->v7 * qc;
  rh[ridx(n, c, nc, i, j1, k)] = rh[ridx(n, c, nc, i, j1, k)] + w->v8 * qc;
}

/* charge.F90:44-83 (func=0) and 113-126 (func=2, rhodg per species) */
void orc_charge(const orc_particle *p, const int64_t *counts, int nspec,
                const double *q, const int n[3], double *rho, double *rhodg) {
  int64_t nr = (int64_t)(n[0] + 3) * (n[1] + 3) * (n[2] + 3);
  memset(rho, 0, sizeof(double) * (size_t)nr);
  if (rhodg) memset(rhodg, 0, sizeof(double) * (size_t)(nr * nspec));
  int64_t m0 = 0;
  for (int is = 0; is < nspec; is++) {
    for (int64_t m = m0; m < m0 + counts[is]; m++) {
      if (p[m].nid == -1) continue;
      cicw w = cic_weights(p[m].x - 0.0, p[m].y - 0.0, p[m].z - 0.0);
      distribute(rho, n, 0, 1, &w, q[is]);
      if (rhodg) distribute(rhodg, n, is, nspec, &w, q[is]);
    }
    m0 += counts[is];
  }
}

/* One periodic rank: oh3_exchange_borders copies planes between the two ends
 * of the same subdomain (ohhelp3.c:1010-1031 with ctypes from ohinit.F90:111-122),
 * dimension by dimension over the extended range. 'from'/'to' are plane
 * offsets of the downward (lower) and upward (upper) transfers. */
static void exchange_periodic(double *a, int nc, const int n[3], int halo_lo, int halo_hi,
                              int dn_from, int dn_to, int dn_size,
                              int up_from, int up_to, int up_size) {
  int ext[3];
  for (int d = 0; d < 3; d++) ext[d] = n[d] + halo_lo + halo_hi + 1;
  for (int d = 0; d < 3; d++) {
    int64_t stride[3] = {1, ext[0], (int64_t)ext[0] * ext[1]};
    int o1 = (d + 1) % 3, o2 = (d + 2) % 3;
    for (int a2 = 0; a2 < ext[o2]; a2++)
      for (int a1 = 0; a1 < ext[o1]; a1++) {
        int64_t base = a1 * stride[o1] + a2 * stride[o2];
        /* downward: my planes [dn_from, dn_from+size) -> lower neighbour's
         * planes n+dn_to ... ; single rank: receiver is myself */
        for (int s = 0; s < dn_size; s++) {
          int src = dn_from + s, dst = n[d] + dn_to + s;
          for (int c = 0; c < nc; c++)
            a[(base + (int64_t)(dst + halo_lo) * stride[d]) * nc + c] =
                a[(base + (int64_t)(src + halo_lo) * stride[d]) * nc + c];
        }
        for (int s = 0; s < up_size; s++) {
          int src = n[d] + up_from + s, dst = up_to + s;
          for (int c = 0; c < nc; c++)
            a[(base + (int64_t)(dst + halo_lo) * stride[d]) * nc + c] =
                a[(base + (int64_t)(src + halo_lo) * stride[d]) * nc + c];
        }
      }
  }
}

/* add_boundary_charge, charge.F90:214-412, bound(:,:) = 1 (periodic) */
void orc_fold_rho(double *rho, int ncomp, const int n[3]) {
  /* CRH ctypes: downward (0,1,1), upward (0,-1,1) (ohinit.F90:115-116) */
  exchange_periodic(rho, ncomp, n, 1, 1, 0, 1, 1, 0, -1, 1);
  const int sl = -1, su = 1, nl = 1, nu = 1, dl = 0, du = 0;
  int lx = n[0], ly = n[1], lz = n[2];
#define RH(c, i, j, k) rho[ridx(n, (c), ncomp, (i), (j), (k))]
  for (int k = 0; k < nl; k++)
    for (int j = 0; j < ly + (su + nu - sl); j++)
      for (int i = 0; i < lx + (su + nu - sl); i++)
        for (int c = 0; c < ncomp; c++)
          RH(c, sl + i, sl + j, dl + k) = RH(c, sl + i, sl + j, dl + k) + RH(c, sl + i, sl + j, sl + k);
  for (int k = 0; k < nu; k++)
    for (int j = 0; j < ly + (su + nu - sl); j++)
      for (int i = 0; i < lx + (su + nu - sl); i++)
        for (int c = 0; c < ncomp; c++)
          RH(c, sl + i, sl + j, lz + du + k) = RH(c, sl + i, sl + j, lz + du + k) + RH(c, sl + i, sl + j, lz + su + k);
  for (int k = 0; k < lz + (du + nu - dl); k++)
    for (int j = 0; j < nl; j++)
      for (int i = 0; i < lx + (su + nu - sl); i++)
        for (int c = 0; c < ncomp; c++)
          RH(c, sl + i, dl + j, dl + k) = RH(c, sl + i, dl + j, dl + k) + RH(c, sl + i, sl + j, dl + k);
  for (int k = 0; k < lz + (du + nu - dl); k++)
    for (int j = 0; j < nu; j++)
      for (int i = 0; i < lx + (su + nu - sl); i++)
        for (int c = 0; c < ncomp; c++)
          RH(c, sl + i, ly + du + j, dl + k) = RH(c, sl + i, ly + du + j, dl + k) + RH(c, sl + i, ly + su + j, dl + k);
  for (int k = 0; k < lz + (du + nu - dl); k++)
    for (int j = 0; j < ly + (du + nu - dl); j++)
      for (int i = 0; i < nl; i++)
        for (int c = 0; c < ncomp; c++)
          RH(c, dl + i, dl + j, dl + k) = RH(c, dl + i, dl + j, dl + k) + RH(c, sl + i, dl + j, dl + k);
  for (int k = 0; k < lz + (du + nu - dl); k++)
    for (int j = 0; j < ly + (du + nu - dl); j++)
      for (int i = 0; i < nu; i++)
        for (int c = 0; c < ncomp; c++)
          RH(c, lx + du + i, dl + j, dl + k) = RH(c, lx + du + i, dl + j, dl + k) + RH(c, lx + su + i, dl + j, dl + k);
#undef RH
}

/* chgntr.F90:69-80 (rion) and 105-190 (10^3 sub-point CIC spread), no bodies */
void orc_chgntr(const double *q, const int64_t *npin, int nspec,
                const int n[3], double *rhobk) {
  int64_t nr = (int64_t)(n[0] + 3) * (n[1] + 3) * (n[2] + 3);
  memset(rhobk, 0, sizeof(double) * (size_t)nr);
  double rion = 0.0;
  for (int is = 0; is < nspec; is++)
    if (npin[is] >= 0) rion = rion - q[is] * (double)npin[is];
  rion = rion / ((double)n[0] * n[1] * n[2]) * 1.0e-3;
  if (fabs(rion) < 1.0e-10) rion = 0.0;
  if (rion == 0.0) return;
  for (int k = -1; k <= n[2] + 1; k++)
    for (int j = -1; j <= n[1] + 1; j++)
      for (int i = -1; i <= n[0] + 1; i++) {
        if (i == n[0] + 1 || j == n[1] + 1 || k == n[2] + 1) continue;
        for (int kk = 0; kk <= 9; kk++)
          for (int jj = 0; jj <= 9; jj++)
            for (int ii = 0; ii <= 9; ii++) {
              double x1 = ii * 0.1, y1 = jj * 0.1, z1 = kk * 0.1;
              double xy1 = x1 * y1, xz1 = x1 * z1, yz1 = y1 * z1;
              double z2 = 1.0 - z1;
              double xz2 = x1 * z2, yz2 = y1 * z2;
              double v3 = xy1 * z1, v2 = xz1 - v3, v4 = yz1 - v3, v1 = z1 - xz1 - v4;
              double v7 = xy1 * z2, v6 = xz2 - v7, v8 = yz2 - v7, v5 = z2 - xz2 - v8;
              cicw w = {v1, v2, v3, v4, v5, v6, v7, v8, i, j, k};
              /* i1 = n+1 planes exist in rhobk(-1:n+1) */
              distribute(rhobk, n, 0, 1, &w, rion);
            }
      }
}

/* add_background_charge, charge.F90:437-447 (nflag_testp != 1) */
void orc_add_background(double *rho, const double *rhobk, const int n[3]) {
  for (int k = 0; k <= n[2]; k++)
    for (int j = 0; j <= n[1]; j++)
      for (int i = 0; i <= n[0]; i++) {
        int64_t o = ridx(n, 0, 1, i, j, k);
        rho[o] = rho[o] + rhobk[o];
      }
}

/* energy.F90:30-48 */
void orc_energy(const orc_particle *p, const int64_t *counts, int nspec,
                const double *rm, const double sl[3], double *engk) {
  int64_t m0 = 0;
  for (int is = 0; is < nspec; is++) {
    double tkx = 0.0, tky = 0.0, tkz = 0.0;
    for (int64_t m = m0; m < m0 + counts[is]; m++) {
      if (p[m].nid == -1) continue;
      tkx = tkx + p[m].vx * p[m].vx;
      tky = tky + p[m].vy * p[m].vy;
      tkz = tkz + p[m].vz * p[m].vz;
    }
    engk[3 * is + 0] = 0.0 + 0.5 * rm[is] * tkx / (sl[0] * sl[1] * sl[2]);
    engk[3 * is + 1] = 0.0 + 0.5 * rm[is] * tky / (sl[0] * sl[1] * sl[2]);
    engk[3 * is + 2] = 0.0 + 0.5 * rm[is] * tkz / (sl[0] * sl[1] * sl[2]);
    m0 += counts[is];
  }
}

/* energy.F90:52-71 */
void orc_field_energy(const double *eb, const int lo[3], const int hi[3],
                      const int n[3], double cs, const double sl[3],
                      double *engeb) {
  double te = 0.0, tb = 0.0;
  for (int k = 0; k < n[2]; k++)
    for (int j = 0; j < n[1]; j++)
      for (int i = 0; i < n[0]; i++) {
        const double *e = &eb[eidx(lo, hi, 0, i, j, k)];
        te = te + e[EX] * e[EX] + e[EY] * e[EY] + e[EZ] * e[EZ];
        tb = tb + e[BX] * e[BX] + e[BY] * e[BY] + e[BZ] * e[BZ];
      }
  engeb[0] = 0.5 * te / (sl[0] * sl[1] * sl[2]);
  engeb[1] = 0.5 * cs * tb / (sl[0] * sl[1] * sl[2]);
}

/* ------------------------------------------------------------------------ */
/* psolve1: EM push + Umeda zigzag current, psolve1.F90:146-802             */
/* ------------------------------------------------------------------------ */

/* post-move relay point for one axis, psolve1.F90:366-415 / 426-536 */
typedef struct {
  double local_a, a, r;
  int ia;
} relay_t;

void orc_psolve1(orc_particle *p, const int64_t *counts, int nspec,
                 const double *qm, const double *q, const double *wrk,
                 const int n[3], int ustep, const double sl[3], double *aj) {
  int64_t nw = (int64_t)(n[0] + 9) * (n[1] + 9) * (n[2] + 9);
  double *tj = (double *)malloc(sizeof(double) * 3 * (size_t)nw);
  double rustep = 1.0 / (double)ustep;
  int64_t m0 = 0;
  for (int is = 0; is < nspec; is++) {
    memset(tj, 0, sizeof(double) * 3 * (size_t)nw);
    double gustep = (double)ustep;
    double qmp = qm[is] * gustep * 0.5;
#define TJ(c, i, j, k) tj[widx(n, (c), 3, (i), (j), (k))]
    for (int64_t m = m0; m < m0 + counts[is]; m++) {
      orc_particle *pp = &p[m];
      if (pp->nid == -1) continue;
      double xlb = pp->x - 0.0, ylb = pp->y - 0.0, zlb = pp->z - 0.0;
      int ib = (int)floor(xlb), jb = (int)floor(ylb), kb = (int)floor(zlb);
      double xb = ib, yb = jb, zb = kb;
      /* weights from floor() (psolve1.F90:170-198) */
      double x1 = xlb - xb, y1 = ylb - yb, z1 = zlb - zb;
      double xy1 = x1 * y1, xz1 = x1 * z1, yz1 = y1 * z1;
      double z2 = 1.0 - z1;
      double xz2 = x1 * z2, yz2 = y1 * z2;
      cicw w;
      w.v3 = xy1 * z1;
      w.v2 = xz1 - w.v3;
      w.v4 = yz1 - w.v3;
      w.v1 = z1 - xz1 - w.v4;
      w.v7 = xy1 * z2;
      w.v6 = xz2 - w.v7;
      w.v8 = yz2 - w.v7;
      w.v5 = z2 - xz2 - w.v8;
      w.i = ib;
      w.j = jb;
      w.k = kb;
      double eex = gather(wrk, n, EX, &w), eey = gather(wrk, n, EY, &w), eez = gather(wrk, n, EZ, &w);
      double bbx = gather(wrk, n, BX, &w), bby = gather(wrk, n, BY, &w), bbz = gather(wrk, n, BZ, &w);
      boris(pp, eex, eey, eez, bbx, bby, bbz, qmp);
      double xmove = pp->vx * gustep, ymove = pp->vy * gustep, zmove = pp->vz * gustep;
      pp->x = pp->x + xmove;
      pp->y = pp->y + ymove;
      pp->z = pp->z + zmove;

      double lb[3] = {xlb, ylb, zlb};
      double cb[3] = {xb, yb, zb};
      int ibv[3] = {ib, jb, kb};
      double *pos[3] = {&pp->x, &pp->y, &pp->z};
      relay_t r[3];
      int inside = !(pp->x < 0.0 || pp->x >= sl[0] || pp->y < 0.0 || pp->y >= sl[1] ||
                     pp->z < 0.0 || pp->z >= sl[2]);
      for (int d = 0; d < 3; d++) {
        double xv = *pos[d];
        if (!inside && xv < 0.0) { /* :440-447 */
          r[d].local_a = xv;
          *pos[d] = xv + sl[d];
          r[d].ia = -1;
          r[d].a = -1.0;
          r[d].r = 0.0;
        } else if (!inside && xv >= sl[d]) { /* :462-469 */
          r[d].local_a = xv - 0.0;
          *pos[d] = xv - sl[d];
          r[d].ia = n[d];
          r[d].a = (double)n[d];
          r[d].r = (double)n[d];
        } else { /* :381-415 and :469-478 */
          r[d].local_a = xv - 0.0;
          r[d].ia = (int)floor(r[d].local_a);
          r[d].a = r[d].ia;
          if (ibv[d] == r[d].ia)
            r[d].r = (lb[d] + r[d].local_a) * 0.5;
          else
            r[d].r = fmax(cb[d], r[d].a);
        }
      }
      /* :717-757 */
      double xd1 = (r[0].r + xlb) * 0.5 - xb, yd1 = (r[1].r + ylb) * 0.5 - yb, zd1 = (r[2].r + zlb) * 0.5 - zb;
      double vvx1 = (r[0].r - xlb) * rustep, vvy1 = (r[1].r - ylb) * rustep, vvz1 = (r[2].r - zlb) * rustep;
      double vx1w4 = yd1 * zd1, vx1w2 = yd1 - vx1w4, vx1w3 = zd1 - vx1w4, vx1w1 = 1.0 - yd1 - vx1w3;
      double vy1w4 = xd1 * zd1, vy1w2 = xd1 - vy1w4, vy1w3 = zd1 - vy1w4, vy1w1 = 1.0 - xd1 - vy1w3;
      double vz1w4 = xd1 * yd1, vz1w2 = xd1 - vz1w4, vz1w3 = yd1 - vz1w4, vz1w1 = 1.0 - xd1 - vz1w3;
      double xa = r[0].a, ya = r[1].a, za = r[2].a;
      double xd2 = (r[0].local_a + r[0].r) * 0.5 - xa, yd2 = (r[1].local_a + r[1].r) * 0.5 - ya,
             zd2 = (r[2].local_a + r[2].r) * 0.5 - za;
      double vvx2 = (r[0].local_a - r[0].r) * rustep, vvy2 = (r[1].local_a - r[1].r) * rustep,
             vvz2 = (r[2].local_a - r[2].r) * rustep;
      double vx2w4 = yd2 * zd2, vx2w2 = yd2 - vx2w4, vx2w3 = zd2 - vx2w4, vx2w1 = 1.0 - yd2 - vx2w3;
      double vy2w4 = xd2 * zd2, vy2w2 = xd2 - vy2w4, vy2w3 = zd2 - vy2w4, vy2w1 = 1.0 - xd2 - vy2w3;
      double vz2w4 = xd2 * yd2, vz2w2 = xd2 - vz2w4, vz2w3 = yd2 - vz2w4, vz2w1 = 1.0 - xd2 - vz2w3;
      int i1b = ib + 1, j1b = jb + 1, k1b = kb + 1;
      int ia = r[0].ia, ja = r[1].ia, ka = r[2].ia, i1a = ia + 1, j1a = ja + 1, k1a = ka + 1;
      /* :759-789 */
      TJ(0, ib, jb, kb) = TJ(0, ib, jb, kb) + vvx1 * vx1w1;
      TJ(1, ib, jb, kb) = TJ(1, ib, jb, kb) + vvy1 * vy1w1;
      TJ(2, ib, jb, kb) = TJ(2, ib, jb, kb) + vvz1 * vz1w1;
      TJ(0, ib, j1b, kb) = TJ(0, ib, j1b, kb) + vvx1 * vx1w2;
      TJ(1, i1b, jb, kb) = TJ(1, i1b, jb, kb) + vvy1 * vy1w2;
      TJ(2, i1b, jb, kb) = TJ(2, i1b, jb, kb) + vvz1 * vz1w2;
      TJ(0, ib, jb, k1b) = TJ(0, ib, jb, k1b) + vvx1 * vx1w3;
      TJ(1, ib, jb, k1b) = TJ(1, ib, jb, k1b) + vvy1 * vy1w3;
      TJ(2, ib, j1b, kb) = TJ(2, ib, j1b, kb) + vvz1 * vz1w3;
      TJ(0, ib, j1b, k1b) = TJ(0, ib, j1b, k1b) + vvx1 * vx1w4;
      TJ(1, i1b, jb, k1b) = TJ(1, i1b, jb, k1b) + vvy1 * vy1w4;
      TJ(2, i1b, j1b, kb) = TJ(2, i1b, j1b, kb) + vvz1 * vz1w4;
      TJ(0, ia, ja, ka) = TJ(0, ia, ja, ka) + vvx2 * vx2w1;
      TJ(1, ia, ja, ka) = TJ(1, ia, ja, ka) + vvy2 * vy2w1;
      TJ(2, ia, ja, ka) = TJ(2, ia, ja, ka) + vvz2 * vz2w1;
      TJ(0, ia, j1a, ka) = TJ(0, ia, j1a, ka) + vvx2 * vx2w2;
      TJ(1, i1a, ja, ka) = TJ(1, i1a, ja, ka) + vvy2 * vy2w2;
      TJ(2, i1a, ja, ka) = TJ(2, i1a, ja, ka) + vvz2 * vz2w2;
      TJ(0, ia, ja, k1a) = TJ(0, ia, ja, k1a) + vvx2 * vx2w3;
      TJ(1, ia, ja, k1a) = TJ(1, ia, ja, k1a) + vvy2 * vy2w3;
      TJ(2, ia, j1a, ka) = TJ(2, ia, j1a, ka) + vvz2 * vz2w3;
      TJ(0, ia, j1a, k1a) = TJ(0, ia, j1a, k1a) + vvx2 * vx2w4;
      TJ(1, i1a, ja, k1a) = TJ(1, i1a, ja, k1a) + vvy2 * vy2w4;
      TJ(2, i1a, j1a, ka) = TJ(2, i1a, j1a, ka) + vvz2 * vz2w4;
    }
    /* :793-802: aj += wrk(TJ) * q(is) over -1..n+1 */
    double qs = q[is];
    for (int k = -1; k <= n[2] + 1; k++)
      for (int j = -1; j <= n[1] + 1; j++)
        for (int i = -1; i <= n[0] + 1; i++)
          for (int c = 0; c < 3; c++) {
            int64_t o = widx(n, c, 3, i, j, k);
            aj[o] = aj[o] + tj[o] * qs;
          }
#undef TJ
    m0 += counts[is];
  }
  free(tj);
}

/* add_boundary_current, psolve1.F90:859-1079, bounds = 1 (periodic);
 * CAJ ctypes: downward (-1,2,3), upward (-1,-4,3) (ohinit.F90:113-114) */
void orc_fold_current(double *aj, const int n[3]) {
  exchange_periodic(aj, 3, n, 4, 4, -1, 2, 3, -1, -4, 3);
  const int sl = -4, su = 2, nl = 3, nu = 3, dl = -1, du = -1;
  int xu = n[0], yu = n[1], zu = n[2];
#define AJ(c, i, j, k) aj[widx(n, (c), 3, (i), (j), (k))]
  for (int k = 0; k < nl; k++)
    for (int j = 0; j < yu + (su + nu - sl); j++)
      for (int i = 0; i < xu + (su + nu - sl); i++)
        for (int c = 0; c < 3; c++)
          AJ(c, sl + i, sl + j, dl + k) = AJ(c, sl + i, sl + j, dl + k) + AJ(c, sl + i, sl + j, sl + k);
  for (int k = 0; k < nu; k++)
    for (int j = 0; j < yu + (su + nu - sl); j++)
      for (int i = 0; i < xu + (su + nu - sl); i++)
        for (int c = 0; c < 3; c++)
          AJ(c, sl + i, sl + j, zu + du + k) = AJ(c, sl + i, sl + j, zu + du + k) + AJ(c, sl + i, sl + j, zu + su + k);
  for (int k = 0; k < zu + (du + nu - dl); k++)
    for (int j = 0; j < nl; j++)
      for (int i = 0; i < xu + (su + nu - sl); i++)
        for (int c = 0; c < 3; c++)
          AJ(c, sl + i, dl + j, dl + k) = AJ(c, sl + i, dl + j, dl + k) + AJ(c, sl + i, sl + j, dl + k);
  for (int k = 0; k < zu + (du + nu - dl); k++)
    for (int j = 0; j < nu; j++)
      for (int i = 0; i < xu + (su + nu - sl); i++)
        for (int c = 0; c < 3; c++)
          AJ(c, sl + i, yu + du + j, dl + k) = AJ(c, sl + i, yu + du + j, dl + k) + AJ(c, sl + i, yu + su + j, dl + k);
  for (int k = 0; k < zu + (du + nu - dl); k++)
    for (int j = 0; j < yu + (du + nu - dl); j++)
      for (int i = 0; i < nl; i++)
        for (int c = 0; c < 3; c++)
          AJ(c, dl + i, dl + j, dl + k) = AJ(c, dl + i, dl + j, dl + k) + AJ(c, sl + i, dl + j, dl + k);
  for (int k = 0; k < zu + (du + nu - dl); k++)
    for (int j = 0; j < yu + (du + nu - dl); j++)
      for (int i = 0; i < nu; i++)
        for (int c = 0; c < 3; c++)
          AJ(c, xu + du + i, dl + j, dl + k) = AJ(c, xu + du + i, dl + j, dl + k) + AJ(c, xu + su + i, dl + j, dl + k);
#undef AJ
}

/* rhoupd.F90:40-49 */
void orc_rhoupd(double *rho, const double *aj, const int n[3]) {
#define AJ(c, i, j, k) aj[widx(n, (c), 3, (i), (j), (k))]
  for (int k = 0; k <= n[2]; k++)
    for (int j = 0; j <= n[1]; j++)
      for (int i = 0; i <= n[0]; i++) {
        int64_t o = ridx(n, 0, 1, i, j, k);
        rho[o] = rho[o] - (AJ(0, i, j, k) - AJ(0, i - 1, j, k) + AJ(1, i, j, k) - AJ(1, i, j - 1, k) +
                           AJ(2, i, j, k) - AJ(2, i, j, k - 1)) * 2.0;
      }
#undef AJ
}

/* psort-struct.c:15-20: (int)(a->z - b->z) */
int orc_psort_key_cmp(const orc_particle *a, const orc_particle *b) {
  return (int)(a->z - b->z);
}
