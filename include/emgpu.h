/*
 * emgpu.h -- C ABI of the MI355X particle hot path of EMSES (libemgpu.so).
 *
 * One library context owns one GPU-resident particle population (all species
 * of one periodic domain that a single GPU owns) and the grid arrays the path
 * reads and writes.  Every entry point is extern "C", takes plain pointers and
 * sizes, returns 0 (EMG_OK) or a negative EMG_ERR_* code and never aborts; the
 * message of the last error is available from emg_last_error().  Calls that
 * only launch device work (emg_positn, emg_push_es, emg_deposit_rho,
 * emg_push_em, emg_sort, emg_init_device) return once the work is enqueued on
 * the context stream; device-side errors they detect (a particle outside the
 * box, a lost particle) are reported by the next synchronising call
 * (emg_get_*, emg_energy, emg_download_particles, emg_synchronize).  Host
 * buffers passed in may be reused as soon as a call returns.  Contexts are
 * independent; one context must not be used from two host threads at once.
 *
 * The reference interfaces each entry point replaces are cited as
 * path:line in sakatataisei1226/ctca.  The reference passes its state through
 * Fortran module globals (src/main/allcom.F90:33-84); here it is passed
 * explicitly.  Grid arrays cross this boundary in the reference's own Fortran
 * layouts: component index fastest, then i, j, k, with explicit bounds lo..hi.
 */
#ifndef EMGPU_H
#define EMGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EMG_MAX_SPECIES 8

/* return codes */
#define EMG_OK 0
#define EMG_ERR_ARG (-1)         /* invalid argument                           */
#define EMG_ERR_HIP (-2)         /* HIP runtime failure                         */
#define EMG_ERR_NOMEM (-3)       /* device allocation failed                    */
#define EMG_ERR_STATE (-4)       /* call out of order (e.g. push before fields) */
#define EMG_ERR_DOMAIN (-5)      /* particle outside [0, L] or non-finite       */
#define EMG_ERR_UNSUPPORTED (-6) /* record kind / option not supported          */
#define EMG_ERR_CAPACITY (-7)    /* more particles than the context capacity    */

/* precision forms of the path */
#define EMG_PLAIN_FP64 0  /* FP64 state, reference arithmetic, FP64 atomics     */
#define EMG_FP32_FIXED 1  /* FP32 cell index + in-cell offset, int64 fixed point */

/*
 * 64-byte particle record, identical to struct S_particle /
 * type(oh_particle): lib/ohhelp/ohhelp-1.1.1/oh_part.h:23-28,
 * lib/ohhelp/ohhelp-1.1.1/oh_type.F90:15-28.
 */
typedef struct {
  double x, y, z, vx, vy, vz;
  int32_t pid, preside, nid, spec;
} emg_particle;

/*
 * Context configuration.  All physical values are in the reference's S-units
 * after rescal (src/main/rescal.F90:38-193): dr = 1, dt = 2, qm(1) = 1.
 */
typedef struct {
  int32_t n[3];        /* cells per axis nx, ny, nz (slx = nx, ...)          */
  int32_t nspec;       /* species count, <= EMG_MAX_SPECIES                  */
  int32_t precision;   /* EMG_PLAIN_FP64 or EMG_FP32_FIXED                   */
  int32_t tile;        /* FP32_FIXED: tile edge in cells, 4 or 8 (0 = 8 when
                          every axis is a multiple of 8, 4 when a multiple
                          of 4, else 8 if every axis has 8 cells or more,
                          else 4).  Any grid with at least `tile` cells per
                          axis and ceil(log2 nx) + ceil(log2 ny) +
                          ceil(log2 nz) <= 32 (the packed cell index) is
                          taken: tiles at the upper faces may be partial.
                          Others return EMG_ERR_UNSUPPORTED (use
                          EMG_PLAIN_FP64, which takes any grid)              */
  int32_t track_pid;   /* FP32_FIXED: carry pid through every move (+8 B per
                          particle-step); without it download returns pid 0  */
  int32_t sort_every;  /* PLAIN_FP64: re-sort by cell (emg_sort) before every
                          sort_every-th push, and before the first push
                          after an upload; 0 = only when emg_sort is called
                          (the reference's isort cadence).  A sort only
                          permutes the particles.  FP32_FIXED: must be 0
                          (its tile buckets are re-formed every step)       */
  int64_t capacity[EMG_MAX_SPECIES]; /* max particles per species            */
  double q[EMG_MAX_SPECIES];  /* charge per macro particle, allcom q(is)     */
  double qm[EMG_MAX_SPECIES]; /* charge-to-mass ratio, allcom qm(is)         */
  double rm[EMG_MAX_SPECIES]; /* mass, allcom rm(is) (energy only)           */
  double e0[3];        /* uniform applied E added by grdcrt (e0x, e0y, e0z)  */
  double b0[3];        /* uniform applied B added by grdcrt (b0x, b0y, b0z)  */
} emg_config;

typedef struct emg_ctx emg_ctx;

/* Context lifetime.  Replaces the allocation part of ohinit
 * (src/wrapper/ohinit.F90:63-211) for the particle path. */
int emg_create(const emg_config *cfg, emg_ctx **out);
int emg_destroy(emg_ctx *ctx);

/* Copy the message of the last failing call on this thread into buf. */
int emg_last_error(char *buf, size_t len);

/* Optional: run on a caller stream (hipStream_t as void*); NULL = own stream. */
int emg_set_stream(emg_ctx *ctx, void *hip_stream);
int emg_synchronize(emg_ctx *ctx);

/*
 * Particle upload / download: the species-contiguous segments of pbuf,
 * src/pcl2field/psolve2.F90:102-115 (species is occupies counts[is] records in
 * order).  Upload keeps records with nid >= 0 in order and drops nid < 0,
 * which is what the single-rank oh3_transbound compaction does
 * (lib/ohhelp/ohhelp-1.1.1/ohhelp2.c:209-237, 390-470).  Records must be
 * alive (preside == 0, OH_PCL_ALIVE, oh_type.F90:31-34) and inside [0, L].
 * Download writes the records back species-contiguously with spec = is+1,
 * nid = 0, preside = 0 and the uploaded pid.
 */
int emg_upload_particles(emg_ctx *ctx, const emg_particle *pbuf, const int64_t *counts,
                         int32_t nspec);
int emg_download_particles(emg_ctx *ctx, emg_particle *pbuf, int64_t *counts);
int emg_particle_counts(emg_ctx *ctx, int64_t *counts);

/*
 * grdcrt(ps) (src/field/grdcrt.F90:125-230): upload the staggered Yee fields
 * eb(6, lo(1):hi(1), lo(2):hi(2), lo(3):hi(3)) and relocate them to the nodes
 * 0..n on the device, adding e0/b0.  Needs lo <= -1 and hi >= n per axis.
 */
int emg_set_fields(emg_ctx *ctx, const double *eb, const int32_t lo[3], const int32_t hi[3]);

/* positn(ps, ustep) (src/pcl/positn.F90:117-254): x += v * ustep, periodic wrap. */
int emg_positn(emg_ctx *ctx, int32_t ustep);

/*
 * psolve2(ps, ustep) (src/pcl2field/psolve2.F90:4-548): trilinear gather of
 * the relocated E/B, Buneman-Boris velocity update, x += v * ustep, periodic
 * wrap.  The charge deposition at the new positions is done in the same pass
 * in both forms, so the following emg_deposit_rho costs nothing.
 */
int emg_push_es(emg_ctx *ctx, int32_t ustep);

/*
 * charge(ps, func) + oh3_exchange_borders(rho, CRH) + add_boundary_charge
 * (src/pcl2field/charge.F90:4-412, src/main/esses.F90:530-601): CIC charge
 * deposition with the periodic fold, per species on the device.  func selects
 * what the caller reads back, as in the reference: 0 rho (emg_get_rho), 2 rho
 * and the per-species rhodg(is, ...) (emg_get_rhodg), 1 rhodg only -- the
 * caller's rho is left alone (the EM loop calls charge(1, 1) right after
 * rhoupd(1), whose continuity-updated rho must survive, emses.F90:474-476).
 * emg_charge in ctca_amd/fortran/emg_gpu.F90 is that body for the
 * reference's arrays.
 */
int emg_deposit_rho(emg_ctx *ctx, int32_t func);

/* Read the folded rho(1, lo:hi)^3 (Fortran layout) of the last deposit.
 * Nodes 0..n hold the periodic fold (node n equals node 0).  The ghost
 * planes -1 and n + 1 hold what the reference's border exchange and
 * add_boundary_charge leave there (ohhelp3.c:1010-1031, charge.F90:265-409):
 * the raw deposit of node n in plane -1 and of node 0 in plane n + 1, folded
 * along z for y ghosts inside z 0..n and along y and z for x ghosts inside
 * y, z 0..n.  Indices beyond -1..n+1 hold periodic images.  emg_get_rhodg
 * reads species is (0-based), folded the same way (rhodg, CRD). */
int emg_get_rho(emg_ctx *ctx, double *rho, const int32_t lo[3], const int32_t hi[3]);
int emg_get_rhodg(emg_ctx *ctx, int32_t is, double *rho, const int32_t lo[3],
                  const int32_t hi[3]);

/* emg_get_rho without any device allocation: no counterpart in the
 * reference.  For callers that cannot allocate device memory (code that
 * records launches into a HIP graph, tight serving loops, harnesses that
 * forbid hipMalloc).  The same read-out as emg_get_rho, staged through a
 * library-owned static device array of at most 2^24 values (128 MiB; each
 * GPU that runs the library holds one until the process ends):
 * (hi - lo + 1) products beyond that return EMG_ERR_CAPACITY.  The array is
 * shared by every context on a GPU, so calls that overlap on one GPU (two
 * host threads) race; sequential calls do not (each returns after its copy). */
int emg_get_rho_noalloc(emg_ctx *ctx, double *rho, const int32_t lo[3], const int32_t hi[3]);

/* energy (src/etc/energy.F90:30-48): engk[3*is + d] =
 * 0.5 rm(is) sum v_d^2 / (slx sly slz), FP64. */
int emg_energy(emg_ctx *ctx, double *engk);

/*
 * psolve1(ps, ustep, func) (src/pcl2field/psolve1.F90:4-857): the same push
 * with floor() cell indices plus the Umeda (2003) zigzag current of each
 * species, aj += q(is) * J(is), followed by oh3_exchange_borders(aj, CAJ) +
 * add_boundary_current (psolve1.F90:859-1079) as a periodic fold.  The
 * current accumulators are cleared by each call (psuper1 zeroes aj,
 * src/pcl/psuper1.F90:88-91).  In EMG_FP32_FIXED the push, move and zigzag
 * current run as one pass over the tile buckets, the current accumulated per
 * species in int64 fixed point (2^24 per unit flux).
 */
int emg_push_em(emg_ctx *ctx, int32_t ustep);
int emg_get_current(emg_ctx *ctx, double *aj, const int32_t lo[3], const int32_t hi[3]);
/* The current of species is (0-based) alone, q(is) J(is) folded like aj: the
 * per-species diagnostic ajdg(3 is + 1 : 3 is + 3, ...) of psolve1 with
 * func = 1 (psolve1.F90:803-817). */
int emg_get_ajdg(emg_ctx *ctx, int32_t is, double *aj, const int32_t lo[3], const int32_t hi[3]);

/*
 * The EM field update on the library's copy of eb (the array emg_set_fields
 * uploaded, periodic ghosts kept consistent), so an EM step can stay on the
 * device: emg_bfield, emg_grdcrt, emg_push_em, emg_bfield, emg_efield.
 *   emg_bfield  ≙ bfield(ps) (src/field/bfield.F90:38-52): the half-step
 *               B += (EY(k+1) - EY) - (EZ(j+1) - EZ), etc.
 *   emg_efield  ≙ efield(ps) (src/field/efield.F90:55-72): E += tcs (curl B
 *               differences) - dlt aj, aj the folded current of the last
 *               emg_push_em (tcs = 2 cv^2 in S-units, inital.F90:78 and
 *               rescal.F90:132; dlt = 2, or 0 for test-particle runs).
 *   emg_grdcrt  ≙ grdcrt(ps) on the device copy: relocates eb to the nodes
 *               again (emg_push_es / emg_push_em refuse nodes older than the
 *               last field update).
 *   emg_get_fields reads eb(6, lo:hi) back (inside the uploaded bounds).
 *   emg_field_energy ≙ the field part of energy (src/etc/energy.F90:52-70):
 *               engeb[0] = 0.5 sum E^2 / V, engeb[1] = 0.5 cs sum B^2 / V over
 *               the cells 0..n-1.
 */
int emg_bfield(emg_ctx *ctx);
int emg_efield(emg_ctx *ctx, double tcs, double dlt);
int emg_grdcrt(emg_ctx *ctx);
int emg_get_fields(emg_ctx *ctx, double *eb, const int32_t lo[3], const int32_t hi[3]);
int emg_field_energy(emg_ctx *ctx, double cs, double *engeb);

/*
 * The electrostatic field solve on the device, so an ES step can stay there:
 * esfld1 + esfld2 + grdcrt for one periodic rank.  rho = the charge of the
 * last deposit (emg_deposit_rho, or the FP32 push) + the uniform background
 * rhobk per node (add_background_charge, charge.F90:414-449, with chgntr's
 * value); phi solves -lap7 phi = rho with the 7-point Laplacian's eigenvalues
 * sum_d (2 sin(pi k_d / n_d))^2 (inifft.F90:390-404), the zero mode dropped
 * (poissn.F90:398-410), through real-to-complex FFTs (hipFFT, not FFTW: the
 * round-off differs); E = -grad phi on the staggered points (esfild.F90:
 * 351-367) is written into the device eb (B kept; an eb(6, -1:n+1)^3 with
 * B = 0 is created if none was uploaded) and relocated to the nodes.
 * emg_get_potential reads phi(0:nx-1, 0:ny-1, 0:nz-1) of the last solve.
 */
int emg_solve_es(emg_ctx *ctx, double rhobk);

/*
 * A background charge that varies in space: rhobk(1, 0:nx-1, 0:ny-1, 0:nz-1,
 * 3) of the reference (x fastest; node n_d is the periodic image of node 0),
 * added per node by every following emg_solve_es before its uniform rhobk
 * (add_background_charge, charge.F90:440-447; the array is what a snapshot
 * stores as 'rhobk', snap.F90:334-337).  A null pointer removes it.
 */
int emg_set_background(emg_ctx *ctx, const double *rhobk);
int emg_get_potential(emg_ctx *ctx, double *phi);

/*
 * spsort(ps) (src/etc/spsort.F90:4-39 -> lib/psort/psort-struct.c:7-9):
 * reorder each species segment by z slab int(z), the order the reference's
 * comparator (int)(a->z - b->z) produces; inside a slab (which the reference
 * leaves to qsort) the particles are ordered by cell.  In EMG_FP32_FIXED the
 * particles are kept in tile buckets by every step; emg_sort there compacts
 * the buckets and absorbs any spill region.
 */
int emg_sort(emg_ctx *ctx);

/*
 * Seeded synthetic particles generated on the device (large configurations
 * whose host generation would dominate): counter-based SplitMix64 stream,
 * uniform positions (or an exponential density ramp along z), Box-Muller
 * Maxwellian velocities rotated by t (inipcl.F90:310-320 rotation).  Not the
 * reference's MT19937 stream; see emg_host_inipcl for that.
 */
typedef struct {
  int64_t count;
  double vth_perp, vth_para; /* peth, path (S-units)                        */
  double drift[3];           /* added after rotation                        */
  double t[9];               /* rotation T(phiz, phixy), row-major          */
  double ramp_ratio;         /* >1: density ~ exp(z ln(ratio)/Lz); else uniform */
} emg_device_species;

int emg_init_device(emg_ctx *ctx, const emg_device_species *spec, int32_t nspec,
                    uint64_t seed);

/*
 * Host mirror of the reference's seeded initial state: MT19937 seeded with
 * `seed`, the three res53 draws of inital's RNG self test
 * (src/main/inital.F90:703-737), then inipcl's draw order per species
 * (src/pcl/inipcl.F90:77-104, 266-326), in EMSES-U units (before rescal).
 * Writes sum(npin) records to out (species-contiguous).  Host only.
 */
typedef struct {
  int64_t npin;
  double peth, path, spe, speth, spa;
  double vdri, vdthz, vdthxy;
  double vdx, vdy, vdz;
} emg_host_species;

int emg_host_inipcl(const emg_host_species *spec, int32_t nspec, const int32_t n[3],
                    const double t[9], uint32_t seed, emg_particle *out);

/*
 * Adjoint window of the electrostatic push (EMG_PLAIN_FP64 contexts).  No
 * counterpart in the reference: these serve callers that differentiate a
 * window of ES steps (the torch harness ctca_amd/adjoint.py) with the
 * reverse step written by hand.
 *
 * emg_state_save / emg_state_load copy the particle state to / from one of 64
 * device slots (checkpoints and the recomputed states of a window).
 * emg_adj_reset sets the particle adjoints lambda_x = lambda_v = 0.
 * emg_adj_step is the reverse of one emg_push_es(ustep) taken at the current
 * particle state and fields: it maps the adjoints of the state after the step
 * to those of the state before it, except for the state's own effect on the
 * fields, and leaves the adjoint of the nodal E (the transpose of the gather,
 * the particles' qmp * lambda_a) for emg_adj_get_field: lam_e(3, 0:nx-1,
 * 0:ny-1, 0:nz-1), periodic nodes.  emg_adj_add_density adds
 * qfac[is] * grad g(x) to lambda_x, g(0:nx-1, 0:ny-1, 0:nz-1) a periodic nodal
 * scalar: the adjoint of the CIC charge (qfac = q) or of any linear function
 * of the density.  emg_adj_velocity_sums gives sum lambda_v per species and
 * component (out[3 * is + d]).
 */
int emg_state_save(emg_ctx *ctx, int32_t slot);
int emg_state_load(emg_ctx *ctx, int32_t slot);
int emg_adj_reset(emg_ctx *ctx);
int emg_adj_step(emg_ctx *ctx, int32_t ustep);
int emg_adj_get_field(emg_ctx *ctx, double *lam_e);
int emg_adj_add_density(emg_ctx *ctx, const double *g, const double *qfac);
int emg_adj_velocity_sums(emg_ctx *ctx, double *out);

/*
 * Kernel timing with HIP events recorded on the context stream around every
 * launch of a kind (which: 0 push/move, 1 deposit, 2 sort, 3 relocate,
 * 4 energy).  No reference counterpart: the reference times its whole loop
 * with MPI_Wtime (src/main/esses.F90:106-143).  Stats synchronise the stream.
 */
#define EMG_KERNEL_PUSH 0
#define EMG_KERNEL_DEPOSIT 1
#define EMG_KERNEL_SORT 2
#define EMG_KERNEL_RELOCATE 3
#define EMG_KERNEL_ENERGY 4
int emg_enable_timing(emg_ctx *ctx, int32_t on);
int emg_kernel_stats(emg_ctx *ctx, int32_t which, double *total_ms, int64_t *launches);
int emg_reset_kernel_stats(emg_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif
