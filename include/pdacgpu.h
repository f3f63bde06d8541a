/*
 * pdacgpu.h -- C-ABI of the MI355X (gfx950) implementation of PDAC's
 * per-time-step hot path.
 *
 * The reference solver exposes its hot path as argument-less Fortran module
 * subroutines that read and write module-global arrays (tilde.f:125,276;
 * iter.f:38; ygas.f:27; htilde.f:25; ftem.f:19; bdry.f:22), called from the
 * time loop of main_prog (prog.f:143-417).  This header is the drop-in
 * boundary for that loop: one context owns every field in device memory, the
 * host (Fortran through iso_c_binding, C, or Python through ctypes) uploads
 * the state once, calls the same subroutines in the same order, and downloads
 * fields when it writes output or restart files.
 *
 * Conventions
 *   - every function returns int: 0 ok, >0 numerical status (see PDAC_ST_*),
 *     <0 HIP/runtime error (see pdac_last_error());
 *   - host arrays are Fortran column-major (ntot, ncomp): component c of
 *     cell ijk (0-based, i fastest) is at [c*ntot + ijk];
 *   - cell index ijk = (i-1) + (j-1)*nx + (k-1)*nx*ny with 1-based i,j,k as in
 *     the reference (ghost.f:817-864), nx,ny,nz including boundary cells;
 *   - one host thread per context; calls are stream-ordered on the context's
 *     HIP stream and only the calls that return scalars synchronise.
 */
#ifndef PDACGPU_H
#define PDACGPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* particle classes the device kernels are built for (the reference allows
 * up to max_nsolid = 16, dimensions.f:13) */
#define PDAC_MAX_SOLID 4
#define PDAC_MAX_GAS 7

/* numerical status codes (>0) */
#define PDAC_ST_ITER_NOT_CONVERGED 1  /* iter.f:428-450 with cstop */
#define PDAC_ST_EOS_FAILURE 2         /* ftem.f:191-193 */
#define PDAC_ST_BAD_ARGUMENT 3

/* precision forms */
#define PDAC_PREC_FP64 0      /* plain FP64 drop-in */
#define PDAC_PREC_FP32FLUX 1  /* MUSCL/limiter/flux arithmetic in FP32 */

/* Cell orders.  iter: the reference's Gauss-Seidel sweep (iter.f:241) or
 * the two-colour sweep of the GPU.  Explicit step: the reference's serial
 * stage loop (explicit_solver.f:1047-1062, whose gas-enthalpy pressure work
 * reads already updated west/south/bottom neighbours) unless
 * PDAC_ORDER_STAGE asks for every right-hand side of a stage from the
 * stage's starting state. */
#define PDAC_ORDER_LEXICOGRAPHIC 0 /* iter: reference order (CPU only); explicit: reference order */
#define PDAC_ORDER_REDBLACK 1      /* iter: two-colour order (GPU); explicit: reference order */
#define PDAC_ORDER_STAGE 2         /* explicit step only: stage-start order */

typedef struct pdac_dims {
  int32_t job_type;  /* 2 or 3 (control_flags JOB_TYPE_2D/3D) */
  int32_t itc;       /* 1: 2D cylindrical (grid.f:319-340) */
  int32_t nx, ny, nz;/* cells per axis including boundary cells; ny=1 in 2D */
  int32_t nsolid;    /* particle classes (<= PDAC_MAX_SOLID) */
  int32_t ngas;      /* gas species (<= PDAC_MAX_GAS) */
  int32_t precision; /* PDAC_PREC_* */
} pdac_dims;

/* Run-time constants: the &model/&numeric/&particles/&atmosphere namelist
 * values the hot path reads (input.f:159-209). */
typedef struct pdac_params {
  double alpha, alphagrav;          /* implicitness of pressure / gravity */
  double gravx, gravy, gravz;
  double beta;                      /* ultrabeta / beta limiter parameter */
  double omega, delg;               /* SOR relaxation, residual tolerance */
  double rlim, flim;                /* matrix limits (matrix.f, ftem.f) */
  double tref;                      /* reference temperature for icpc<=1 */
  double void_fraction;             /* atmospheric void fraction (inflow) */
  double dk[PDAC_MAX_SOLID];        /* particle diameter [m] */
  double rl[PDAC_MAX_SOLID];        /* particle density */
  double phis[PDAC_MAX_SOLID];      /* sphericity */
  double cps[PDAC_MAX_SOLID];       /* particle specific heat */
  double kap[PDAC_MAX_SOLID];       /* particle conductivity */
  double plim[PDAC_MAX_SOLID];      /* two-phase limit */
  double atm_ygc[PDAC_MAX_GAS];     /* atmospheric mass fractions by type */
  int32_t gas_type[PDAC_MAX_GAS];   /* 1-based reference gas types */
  int32_t muscl, lm, lv;            /* MUSCL on/off, mass / velocity limiter */
  int32_t inmax, maxout;            /* in-cell and outer iteration limits */
  int32_t pmodel, drag_model, icpc, ord;
  int32_t cstop;                    /* stop when iter does not converge */
  int32_t iter_order;               /* PDAC_ORDER_* */
  int32_t time_integration;         /* 0 semi-implicit, 1 explicit (prog.f:192-268) */
  /* viscous stresses and heat conduction (&model, input.f:247-251): the gas
   * stress tensor (viscg, visc.f:150-180) and conduction (hotc.f:12-84),
   * the particle stress tensor (viscs, visc.f:182-291) and conduction, the
   * Coulombic repulsive term (repulsive_model = 1, visc.f:232-289) */
  int32_t gas_viscosity, part_viscosity, repulsive_model;
  double mus[PDAC_MAX_SOLID];       /* particle viscosity (&particles viscosity, input.f:1100) */
  /* gas sub-grid turbulence (&model iturb, cmut, input.f:252-256): 0 off,
   * 1 the classical Smagorinsky model (modturbo = 1; sgsg, turbo.f:145-383,
   * with the length scale of turbulence_setup, turbo.f:44-118); the
   * roughness (iturb = 2), dynamic (modturbo = 2) and particle (iss = 1)
   * models are not built */
  double cmut;
  int32_t iturb;
} pdac_params;

/* Field identifiers for pdac_put/pdac_get.  ncomp is 1 unless noted
 * (S = nsolid, G = ngas, A = (S+1)(S+2)/2 packed interphase entries). */
enum pdac_field {
  PDAC_P = 0, PDAC_EP, PDAC_RGP, PDAC_ROG, PDAC_TG, PDAC_SIEG,
  PDAC_UG, PDAC_VG, PDAC_WG,
  PDAC_RLK,   /* S */
  PDAC_TS,    /* S */
  PDAC_SIES,  /* S */
  PDAC_US, PDAC_VS, PDAC_WS, /* S each */
  PDAC_YGC, PDAC_XGC,        /* G each */
  PDAC_CG, PDAC_CK /* S */, PDAC_MUG, PDAC_KAPG,
  /* step-start (n-level) copies written by fieldn */
  PDAC_PN, PDAC_EPN, PDAC_RGPN, PDAC_SIEGN, PDAC_RLKN /* S */, PDAC_SIESN /* S */,
  PDAC_RUGN, PDAC_RVGN, PDAC_RWGN, PDAC_RUSN, PDAC_RVSN, PDAC_RWSN /* S */,
  PDAC_RGPGCN /* G */,
  /* explicit (tilde) terms and interphase matrices written by tilde/htilde */
  PDAC_RUG, PDAC_RVG, PDAC_RWG, PDAC_RUS, PDAC_RVS, PDAC_RWS /* S */,
  PDAC_APPU, PDAC_APPV, PDAC_APPW /* A */,
  PDAC_RHG, PDAC_RHS /* S */,
  /* mass fluxes of the last iter call */
  PDAC_RGFE, PDAC_RGFN, PDAC_RGFT, PDAC_RSFE, PDAC_RSFN, PDAC_RSFT /* S */,
  PDAC_NFIELDS
};

typedef struct pdac_ctx pdac_ctx;

/* Context lifetime.  device < 0 selects the current device. */
int pdac_create(pdac_ctx **ctx, const pdac_dims *dims, int device);
int pdac_destroy(pdac_ctx *ctx);

/* Grid: cell sizes dx[nx], dy[ny] (ignored in 2D), dz[nz]; cell flags[ntot]
 * (grid.f:45-58 values); atmospheric profile p_atm[nz], t_atm[nz] used by
 * free in/outflow boundaries (inoutflow.f:60-123).  r, rb, inr, inrb and the
 * inverse sizes are derived as in grid_setup (grid.f:284-340). */
int pdac_set_grid(pdac_ctx *ctx, const double *dx, const double *dy,
                  const double *dz, const int32_t *flag,
                  const double *p_atm, const double *t_atm);
int pdac_set_params(pdac_ctx *ctx, const pdac_params *prm);

/* sgsg (turbo.f:145-383), called by main_prog after fieldn or
 * initialize_fieldn when iturb >= 1 (prog.f:207): the Smagorinsky turbulent
 * viscosity mugt = rog * smag * |S| and conductivity kapgt = mugt * cg / 0.5
 * of every computed cell.  A no-op when iturb = 0. */
int pdac_sgsg(pdac_ctx *ctx);

/* Output and restart files without a host synchronisation in the step loop
 * (outp, outp.f:194-384, at every nprint-th step; taperd, io_restart.f:31-70,
 * at every ndump-th): pdac_snapshot_async packs the arrays of one file on the
 * device and queues their download into a staging buffer of the library
 * (device and pinned host memory of pdac_snapshot_size bytes, allocated by
 * the first call of each kind), then returns; the steps queued after it
 * overlap the download.  pdac_snapshot_fetch waits for that download and
 * copies the arrays to host (bytes must be the size).  A second snapshot of
 * a kind before its fetch replaces the first.  Array order and precision:
 *   PDAC_SNAP_OUTPUT   float: p, ug, [vg], wg, tg, xgc(1..ngas), then per
 *                      class rlk/rl, us, [vs], ws, ts (the records of an
 *                      unformatted output.nnnn after its time record)
 *   PDAC_SNAP_RESTART  double: p, ug, [vg], wg, sieg, rlk(1..nsolid), per
 *                      class us, [vs], ws, sies(1..nsolid), ygc(1..ngas),
 *                      rgp, rog, tg, ts(1..nsolid) (the records of pdac.res
 *                      after its two header records)
 * each array ntot values; [vg], [vs] on 3D meshes only. */
#define PDAC_SNAP_OUTPUT 0
#define PDAC_SNAP_RESTART 1
int pdac_snapshot_size(pdac_ctx *ctx, int kind, int64_t *bytes);
int pdac_snapshot_async(pdac_ctx *ctx, int kind);
int pdac_snapshot_fetch(pdac_ctx *ctx, int kind, void *host, int64_t bytes);

/* Immersed boundaries (immb = 1, 3D, semi-implicit): what the reference's
 * set-up leaves in MODULE immersed_boundaries (immb.f:33-60): the forcing
 * points of the x, y and z velocity components (TYPE forcing_point,
 * immb.f:33-45; set_forcing :73-516), the face masks b_e_..b_b_ and ivf_
 * (fill_cells ghost.f:3052-3261, face_fractions ghost.f:3263-3321).
 * nfp[3] point counts; fpi 7 planes of nfp[0]+nfp[1]+nfp[2] int32 (i, j, k
 * 1-based, int, delta_i, delta_j, delta_k; x points, then y, then z, each in
 * the reference's order); fpd 3 planes of doubles (the no-slip point nsl
 * x, y, z); coords x[nx], xb[nx], y[ny], yb[ny], z[nz], zb[nz] of the
 * georeferenced, lifted mesh (compute_UTM_coords topo.f:731-751,
 * set_profile :265-278); bmask 6 planes of ntot int32 (b_e, b_w, b_n, b_s,
 * b_t, b_b); ivf[ntot].  The grid's flags (pdac_set_grid) are those after
 * fill_cells; a grid with immb or filled cells needs this call before any
 * step.  index_q/index_qq and numx/y/z follow local_forcing
 * (ghost.f:2888-3050). */
int pdac_set_ib(pdac_ctx *ctx, const int32_t *nfp, const int32_t *fpi, const double *fpd,
                const double *coords, const int32_t *bmask, const double *ivf);

/* Immersed-boundary set-up on the host (no device call): the forcing points
 * of one velocity component (axis 1..3) of an nx x ny x nz mesh, as
 * set_forcing makes them (immb.f:73-516): first one per cell of ext_cells
 * (0-based cells, ascending: the component lies above the ground next to a
 * filled cell; ext_forcing3d, immb.f:1071-1261), then those of the interior
 * columns in (i, j, k) order (forcing3d, immb.f:913-1069).  ord2d and topo2d
 * hold one value per column at (i-1)*ny + (j-1): the lowest level above the
 * ground and the ground height on the component's grid (interpolate_dem,
 * topo.f:970-1029); cx, cy, cz the grid's coordinates (nx, ny, nz values).
 * Output: *count points, cap at most, as rows of pts (i, j, k, int, di, dj,
 * dk: TYPE forcing_point, immb.f:33-45) and nsl (the no-slip point x, y, z).
 * Columns are split over host threads; the result does not depend on them.
 * PDAC_ST_BAD_ARGUMENT with *count set when more than cap points. */
int pdac_ib_forcing(int32_t nx, int32_t ny, int32_t nz, int32_t axis, const int32_t *ord2d, const double *cx,
                    const double *cy, const double *cz, const double *topo2d, const int64_t *ext_cells,
                    int64_t n_ext, int64_t cap, int32_t *pts, double *nsl, int64_t *count);

/* Immersed-boundary set-up on the host (no device call): fill_cells
 * (ghost.f:3052-3261) in the reference's serial cell order.  For each
 * computed interior cell of flags (ntot values, updated in place), the open
 * faces bd (bits e w t b n s = 1 2 4 8 16 32) and open volume fraction vf;
 * a cell with no open face becomes filled_cell_1 (512).  tx, ty, tc: the
 * ground heights of the x, y and cell-centred grids at (i-1)*ny + (j-1); z,
 * zb: the cell centres and upper faces along z (nz values each). */
int pdac_ib_fill_cells(int32_t nx, int32_t ny, int32_t nz, const double *tx, const double *ty, const double *tc,
                       const double *z, const double *zb, int32_t *flags, int32_t *bd, double *vf);

/* Host <-> device copies of one field; n must equal ntot*ncomp(field). */
int pdac_put(pdac_ctx *ctx, int field, const double *host, int64_t n);
int pdac_get(pdac_ctx *ctx, int field, double *host, int64_t n);
int pdac_field_ncomp(const pdac_ctx *ctx, int field);

/* Derived state after a restart read: molar fractions, void fraction and
 * heat capacities (cnvert itd=2, setup.f:359-392). */
int pdac_cnvert(pdac_ctx *ctx);

/* The hot path, one reference subroutine each. */
int pdac_boundary(pdac_ctx *ctx);                       /* bdry.f:22 */
int pdac_fieldn(pdac_ctx *ctx);                         /* tilde.f:125 */
int pdac_tilde(pdac_ctx *ctx, double dt);               /* tilde.f:276 */
int pdac_iter(pdac_ctx *ctx, double dt, int32_t *nit,
              int32_t *converged);                      /* iter.f:38 */
int pdac_ygas(pdac_ctx *ctx, double dt);                /* ygas.f:27 */
int pdac_htilde(pdac_ctx *ctx, double dt);              /* htilde.f:25 */
int pdac_ftem(pdac_ctx *ctx, double dt, int32_t *eos_fail); /* ftem.f:19 */

/* Reset of the per-step tilde/interphase arrays (allocate_fluxes,
 * tilde.f:56-103); pdac_step calls it once per step. */
int pdac_begin_step(pdac_ctx *ctx);

/* The explicit step (time_integration = 1), explicit_solver.f.  The
 * context needs pdac_params.time_integration = 1 at pdac_set_params, which
 * then allocates the Runge-Kutta workspace.  explicit_iter returns the
 * number of caloric-EOS failures (explicit_solver.f:1217-1219: more than one
 * is PDAC_ST_EOS_FAILURE) and the largest Courant number over the computed
 * cells (compute_courant :2063-2103). */
int pdac_initialize_fieldn(pdac_ctx *ctx, double dt);   /* explicit_solver.f:419 */
int pdac_explicit_iter(pdac_ctx *ctx, double dt, int32_t *eos_fail,
                       double *cfl);                    /* explicit_solver.f:996 */

/* nsteps full steps of main_prog (prog.f:143-329): boundary, then
 * semi-implicit: fieldn and for rk = 1..rungekut with dt/(rungekut+1-rk):
 * tilde, iter, ygas, htilde, ftem; explicit (time_integration = 1):
 * initialize_fieldn and per stage explicit_iter, ygas.  nit_last receives the
 * sweep count of the last iter call, nit_total the sum over all calls (0 for
 * explicit steps). */
int pdac_step(pdac_ctx *ctx, double dt, int32_t rungekut, int32_t nsteps,
              int32_t *nit_last, int64_t *nit_total);

int pdac_sync(pdac_ctx *ctx);
const char *pdac_last_error(void);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): nsteps explicit steps as pdac_step takes them for
 * time_integration = 1 (boundary, initialize_fieldn, per stage explicit_iter
 * and ygas; same results), with the Runge-Kutta workspace in a static device
 * array of PDAC_NOALLOC_BYTES that the library holds on each GPU it runs on
 * until the process ends, instead of the context's own.  The call allocates
 * nothing; the context needs pdac_set_grid and pdac_set_params only, with
 * either time_integration (0 leaves it without a workspace of its own).
 * Meshes whose workspace exceeds the array are refused with
 * PDAC_ST_BAD_ARGUMENT (2D, two classes: about 200,000 cells).  Every call
 * on one GPU shares that GPU's array: calls that overlap on one GPU (from
 * two host threads or on two streams) race.  eos_fail receives the largest
 * caloric-EOS failure count of the steps, cfl the last Courant number. */
#define PDAC_NOALLOC_BYTES 134217728
int pdac_explicit_step_noalloc(pdac_ctx *ctx, double dt, int32_t rungekut, int32_t nsteps,
                               int32_t *eos_fail, double *cfl);

/* Ensembles of independent explicit-step domains of uneven size (no
 * counterpart in the reference, which runs one domain per job).  The members
 * are contexts set up as for pdac_step with time_integration = 1 and the
 * same job_type, nsolid, ngas, precision and device; they keep their own
 * device arrays and stay usable one by one.  pdac_ens_step advances every
 * member nsteps steps of main_prog (boundary, initialize_fieldn, per stage
 * explicit_iter and ygas) with its own dt[m]: one batched launch per kernel
 * covers all members, and a step's launches are captured in a HIP graph on
 * first use and replayed.  eos_fail[m] receives the largest EOS failure
 * count of member m's explicit_iter calls, cfl[m] its last Courant number;
 * PDAC_ST_EOS_FAILURE is returned after the batch when a member exceeded 1.
 * pdac_ens_tune times each kernel class over the tile shapes 256x1, 128x2,
 * 64x4, 32x8 (cells along i x rows) on the members' data and keeps the
 * fastest (tiles[8] receives the choice per class: boundary, xinit, xflux,
 * xrhs, xmass, xsolve, ygas flux, ygas update; us[8*4] the mean times); it
 * leaves every state field unchanged.  Members must outlive the ensemble. */
typedef struct pdac_ens pdac_ens;
int pdac_ens_create(pdac_ens **ens, pdac_ctx *const *members, int count);
int pdac_ens_destroy(pdac_ens *ens);
int pdac_ens_tune(pdac_ens *ens, int32_t reps, int32_t *tiles, double *us);
int pdac_ens_set_tiles(pdac_ens *ens, const int32_t *tiles);
int pdac_ens_step(pdac_ens *ens, const double *dt, int32_t rungekut, int32_t nsteps,
                  int32_t *eos_fail, double *cfl);

/* Diagnostic: momentum fluxes of the last pdac_tilde call (compute_all_
 * fluxes_tilde, tilde.f:850-1326).  comp 0..2 = u,v,w; face 0..2 = e,n,t;
 * out receives ntot*(nsolid+1) values, phase-major. */
int pdac_get_momflux(pdac_ctx *ctx, int comp, int face, double *out);

/* Diagnostic: per-kernel timing with HIP events recorded on the context's
 * stream around each launch (no counterpart in the reference, whose phase
 * timers are prog.f:430-451).  Sweep launches that found the iteration
 * already converged are not counted.  pdac_profile_read synchronises. */
enum pdac_kernel {
  PDAC_K_BOUNDARY = 0, PDAC_K_FIELDN, PDAC_K_MOMFLUX, PDAC_K_TILDE, PDAC_K_PREDICT,
  PDAC_K_GASFLUX, PDAC_K_SNAPSHOT, PDAC_K_SWEEP, PDAC_K_YGAS_FLUX, PDAC_K_YGAS_UPDATE,
  PDAC_K_HTILDE_FLUX, PDAC_K_HTILDE_UPDATE, PDAC_K_FTEM, PDAC_K_FILL, PDAC_K_SWEEP_CHECK,
  PDAC_K_ITERATE, PDAC_K_XINIT, PDAC_K_XFLUX, PDAC_K_XRHS, PDAC_K_XCHAIN, PDAC_K_XMASS,
  PDAC_K_XSOLVE, PDAC_K_STRESS, PDAC_K_COUNT
};
int pdac_profile_enable(pdac_ctx *ctx, int enable);
/* bit k of mask: time kernel k (default: all) */
int pdac_profile_select(pdac_ctx *ctx, uint32_t mask);
int pdac_profile_reset(pdac_ctx *ctx);
int pdac_profile_read(pdac_ctx *ctx, int kernel, int64_t *launches, double *total_ms);
/* cells visited by PDAC_K_SWEEP launches (not converged at visit start)
 * since the last pdac_profile_reset */
int pdac_profile_visits(pdac_ctx *ctx, int64_t *visits);
/* cells whose mass residual PDAC_K_SWEEP launches evaluated (the cells of a
 * colour phase whose inputs changed since their last check) since the last
 * pdac_profile_reset */
int pdac_profile_checks(pdac_ctx *ctx, int64_t *checks);

#ifdef __cplusplus
}
#endif

#endif /* PDACGPU_H */
