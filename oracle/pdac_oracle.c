/*
 * pdac_oracle.c -- CPU restatement (FP64) of PDAC's semi-implicit per-step
 * hot path, written from the reference Fortran.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP implementation.  Only the
 * tests, the smoke check and the CPU-baseline leg of the benchmark load it.
 *
 * Every routine cites the reference file:line it restates.  Arithmetic keeps
 * the reference's operand order (Fortran evaluates a*b*c as (a*b)*c) and its
 * single-precision literals widened to double, so that on identical inputs
 * the results agree with the reference to the last bits; build with
 * -ffp-contract=off.
 *
 * Pinning (what checks each function, beyond retyping it):
 *  - whole steps: golden restart files written by the reference solver
 *    (tests/golden, made by tests/golden/make_golden.py);
 *  - limiters: Sweby closed forms and the TVD region (tests);
 *  - Gauss-Jordan: numpy.linalg.solve on random diagonally dominant systems;
 *  - caloric EOS: round trip h(T) -> T; drag: closed-form Stokes limit.
 *
 * Two orders are available for the pressure-velocity sweep of iter:
 * PDAC_ORDER_LEXICOGRAPHIC is the reference's serial Gauss-Seidel order
 * (iter.f:241); PDAC_ORDER_REDBLACK visits cells of colour (i+j+k) even, then
 * odd, and reads the second-neighbour densities of the MUSCL stencil from a
 * snapshot taken at the start of each colour phase -- the order of the GPU
 * implementation, which it therefore reproduces exactly.
 */
#include "pdac_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define MAXS PDAC_MAX_SOLID
#define MAXP (PDAC_MAX_SOLID + 1)
#define MAXA ((MAXP * (MAXP + 1)) / 2)

/* single-precision literal of the reference widened to double */
#define F(x) ((double)(float)(x))

/* cell flags, grid.f:45-58 */
enum { FLUID = 1, SLIP_WALL = 2, NOSLIP_WALL = 4, BL_CELL = 5, FREE_IO = 8,
       INLET_CELL = 16, ZERO_GRAD = 32, VENT_CELL = 64, DOME_CELL = 129,
       IMMB_CELL = 257, FILLED_CELL_1 = 512, FILLED_CELL_2 = 1024 };

static const double GMW_BASE[7] = {32.0, 28.0, 44.0, 2.0, 18.0, 28.96442, 64.0};
static const double MMUGS[7] = {3.467, 3.798, 3.941, 2.827, 2.641, 3.711, 4.112};
static const double MMUGEK[7] = {106.7, 71.4, 195.2, 59.7, 809.1, 78.6, 335.4};
static const double RGAS = 8.31432, GAMMAAIR = 1.4, TZERO = 0.0, HZEROG = 0.0,
                    HZEROS = 0.0;

struct pdac_cpu_ctx {
  pdac_dims d;
  pdac_params prm;
  int three, nx, ny, nz, ns, ng, np, na;
  long n;
  /* grid, index 0..n+1 (1-based like the reference) */
  double *dx, *dy, *dz, *indx, *indy, *indz, *r, *rb, *inr, *inrb;
  int32_t *flag;
  double *p_atm, *t_atm; /* 1..nz */
  double gmw[7];
  double inrl[MAXS], phi[MAXS];
  double philim[MAXS][MAXS], epsl[MAXS][MAXS], epsu[MAXS][MAXS],
      dkf[MAXS][MAXS];
  double *f[PDAC_NFIELDS];
  /* momentum fluxes: [component u,v,w][face e,n,t][phase] */
  double *mf[3][3];
  /* enthalpy and species fluxes */
  double *egf[3], *esf[3], *yf[3];
  /* viscous stress divergence along x, y, z: np planes (gas, then each class) */
  double *vis[3];
  /* gas viscosity and conductivity with the turbulent parts (mugt, kapgt of
   * MODULE turbulence_model, turbo.f:35-40: zero until sgsg sets them) */
  double *mugt, *kapgt;
  double *conv, *abeta, *appu_old[3];
  double *snap_rgp, *snap_rlk;
  struct xstate *xs; /* explicit Runge-Kutta state, allocated on first use */
  int grid_set, params_set;
  /* immersed boundaries (pdac_cpu_set_ib): forcing points of the x, y, z
   * velocity components (TYPE forcing_point, immb.f:33-45), the forcing
   * point of each cell (numx/y/z), the staggered-grid coordinates, the face
   * masks and 1/vf (immb.f:47-60) and the filled_cell_2 toggles (fc2) */
  int ib;
  int nfp[3];
  int32_t *fpi[3];     /* i, j, k, int, delta_i, delta_j, delta_k: 7 planes of nfp */
  double *fpnsl[3];    /* no-slip point x, y, z: 3 planes of nfp */
  int32_t *fpq[3], *fpqq[3];  /* 0-based interpolation neighbours (-1: none) */
  double *fpvel[3];    /* pseudo-velocity of each phase: np planes of nfp */
  int32_t *num[3];     /* per cell, 1-based forcing point or 0 */
  double *crd[6];      /* x, xb, y, yb, z, zb, 0-based */
  int32_t *bmask;      /* per cell: b_e, b_w, b_n, b_s, b_t, b_b (6 planes) */
  double *ivf;
  uint8_t *fc2;
  int ib_off;          /* tests only: parts of the runtime left out (pdac_cpu_set_ib_off) */
};

static void x_free(struct xstate *x);

static int ncomp_of(const pdac_cpu_ctx *s, int id) {
  switch (id) {
  case PDAC_RLK: case PDAC_TS: case PDAC_SIES: case PDAC_US: case PDAC_VS:
  case PDAC_WS: case PDAC_CK: case PDAC_RLKN: case PDAC_SIESN:
  case PDAC_RUSN: case PDAC_RVSN: case PDAC_RWSN: case PDAC_RUS:
  case PDAC_RVS: case PDAC_RWS: case PDAC_RHS: case PDAC_RSFE:
  case PDAC_RSFN: case PDAC_RSFT:
    return s->ns;
  case PDAC_YGC: case PDAC_XGC: case PDAC_RGPGCN:
    return s->ng;
  case PDAC_APPU: case PDAC_APPV: case PDAC_APPW:
    return s->na;
  default:
    return (id >= 0 && id < PDAC_NFIELDS) ? 1 : 0;
  }
}

#define FLD(id) (s->f[id])
#define IJK(i, j, k) ((long)((i)-1) + (long)((j)-1) * s->nx + (long)((k)-1) * s->nx * s->ny)

static int computed(int32_t fl) { return fl & 1; }
static int wall(int32_t fl) { return fl == SLIP_WALL || fl == NOSLIP_WALL; }
static int wallf(int32_t fl) { return wall(fl) || fl == FILLED_CELL_1; }

/* ------------------------------------------------------------------ */
int pdac_cpu_create(pdac_cpu_ctx **out, const pdac_dims *d) {
  if (!out || !d) return PDAC_ST_BAD_ARGUMENT;
  if (d->nsolid < 1 || d->nsolid > MAXS || d->ngas < 1 || d->ngas > 7 ||
      d->nx < 4 || d->nz < 4 || (d->job_type == 3 && d->ny < 4) ||
      (d->job_type != 2 && d->job_type != 3))
    return PDAC_ST_BAD_ARGUMENT;
  pdac_cpu_ctx *s = (pdac_cpu_ctx *)calloc(1, sizeof(*s));
  s->d = *d;
  s->three = d->job_type == 3;
  s->nx = d->nx; s->ny = s->three ? d->ny : 1; s->nz = d->nz;
  s->ns = d->nsolid; s->ng = d->ngas; s->np = s->ns + 1;
  s->na = (s->np * (s->np + 1)) / 2;
  s->n = (long)s->nx * s->ny * s->nz;
  for (int id = 0; id < PDAC_NFIELDS; ++id)
    s->f[id] = (double *)calloc((size_t)s->n * ncomp_of(s, id), sizeof(double));
  for (int c = 0; c < 3; ++c) {
    for (int fc = 0; fc < 3; ++fc)
      s->mf[c][fc] = (double *)calloc((size_t)s->n * s->np, sizeof(double));
    s->egf[c] = (double *)calloc((size_t)s->n, sizeof(double));
    s->esf[c] = (double *)calloc((size_t)s->n * s->ns, sizeof(double));
    s->yf[c] = (double *)calloc((size_t)s->n * s->ng, sizeof(double));
    s->appu_old[c] = (double *)calloc((size_t)s->n * s->na, sizeof(double));
    s->vis[c] = (double *)calloc((size_t)s->n * s->np, sizeof(double));
  }
  s->conv = (double *)calloc((size_t)s->n, sizeof(double));
  s->mugt = (double *)calloc((size_t)s->n, sizeof(double));
  s->kapgt = (double *)calloc((size_t)s->n, sizeof(double));
  s->abeta = (double *)calloc((size_t)s->n, sizeof(double));
  s->snap_rgp = (double *)calloc((size_t)s->n, sizeof(double));
  s->snap_rlk = (double *)calloc((size_t)s->n * s->ns, sizeof(double));
  s->dx = (double *)calloc(s->nx + 2, sizeof(double));
  s->indx = (double *)calloc(s->nx + 2, sizeof(double));
  s->r = (double *)calloc(s->nx + 2, sizeof(double));
  s->rb = (double *)calloc(s->nx + 2, sizeof(double));
  s->inr = (double *)calloc(s->nx + 2, sizeof(double));
  s->inrb = (double *)calloc(s->nx + 2, sizeof(double));
  s->dy = (double *)calloc(s->ny + 2, sizeof(double));
  s->indy = (double *)calloc(s->ny + 2, sizeof(double));
  s->dz = (double *)calloc(s->nz + 2, sizeof(double));
  s->indz = (double *)calloc(s->nz + 2, sizeof(double));
  s->p_atm = (double *)calloc(s->nz + 2, sizeof(double));
  s->t_atm = (double *)calloc(s->nz + 2, sizeof(double));
  s->flag = (int32_t *)calloc((size_t)s->n, sizeof(int32_t));
  for (int g = 0; g < 7; ++g) s->gmw[g] = GMW_BASE[g] * 1.0e-3; /* setup.f:577 */
  *out = s;
  return 0;
}

int pdac_cpu_destroy(pdac_cpu_ctx *s) {
  if (!s) return 0;
  for (int id = 0; id < PDAC_NFIELDS; ++id) free(s->f[id]);
  for (int c = 0; c < 3; ++c) {
    for (int fc = 0; fc < 3; ++fc) free(s->mf[c][fc]);
    free(s->egf[c]); free(s->esf[c]); free(s->yf[c]); free(s->appu_old[c]); free(s->vis[c]);
  }
  free(s->conv); free(s->abeta); free(s->mugt); free(s->kapgt); free(s->snap_rgp); free(s->snap_rlk);
  free(s->dx); free(s->indx); free(s->r); free(s->rb); free(s->inr);
  free(s->inrb); free(s->dy); free(s->indy); free(s->dz); free(s->indz);
  free(s->p_atm); free(s->t_atm); free(s->flag);
  x_free(s->xs);
  for (int a = 0; a < 3; ++a) {
    free(s->fpi[a]); free(s->fpnsl[a]); free(s->fpq[a]); free(s->fpqq[a]);
    free(s->fpvel[a]); free(s->num[a]);
  }
  for (int a = 0; a < 6; ++a) free(s->crd[a]);
  free(s->bmask); free(s->ivf); free(s->fc2);
  free(s);
  return 0;
}

/* Immersed boundaries: what set_forcing, local_forcing, fill_cells and
 * face_fractions leave (immb.f:73-516, ghost.f:2888-3321).  fpi holds the
 * integer fields of all points (x points, then y, then z) as 7 planes of
 * nfp[0]+nfp[1]+nfp[2] (i, j, k, int, delta_i, delta_j, delta_k), fpd the
 * no-slip points as 3 planes; coords are x, xb (nx each), y, yb (ny),
 * z, zb (nz); bmask the face masks as 6 planes of ntot (e, w, n, s, t, b);
 * ivf 1/vf per cell.  index_q/qq and numx/y/z follow local_forcing
 * (ghost.f:2920-2940: a later point of a cell replaces an earlier one). */
int pdac_cpu_set_ib(pdac_cpu_ctx *s, const int32_t *nfp, const int32_t *fpi, const double *fpd,
                    const double *coords, const int32_t *bmask, const double *ivf) {
  if (!s || !nfp || !s->three) return PDAC_ST_BAD_ARGUMENT;
  long n = s->n;
  int nx = s->nx, ny = s->ny, nz = s->nz;
  long tot = (long)nfp[0] + nfp[1] + nfp[2];
  long base = 0;
  for (int a = 0; a < 3; ++a) {
    int m = nfp[a];
    s->nfp[a] = m;
    s->fpi[a] = (int32_t *)malloc(sizeof(int32_t) * 7 * (size_t)(m + 1));
    s->fpnsl[a] = (double *)malloc(sizeof(double) * 3 * (size_t)(m + 1));
    s->fpq[a] = (int32_t *)malloc(sizeof(int32_t) * (size_t)(m + 1));
    s->fpqq[a] = (int32_t *)malloc(sizeof(int32_t) * (size_t)(m + 1));
    s->fpvel[a] = (double *)calloc((size_t)s->np * (m + 1), sizeof(double));
    s->num[a] = (int32_t *)calloc((size_t)n, sizeof(int32_t));
    for (int f = 0; f < 7; ++f)
      for (int p = 0; p < m; ++p) s->fpi[a][(long)f * m + p] = fpi[(long)f * tot + base + p];
    for (int f = 0; f < 3; ++f)
      for (int p = 0; p < m; ++p) s->fpnsl[a][(long)f * m + p] = fpd[(long)f * tot + base + p];
    for (int p = 0; p < m; ++p) {
      const int32_t *F = s->fpi[a];
      int i = F[p], j = F[m + p], k = F[2 * m + p], it = F[3 * m + p];
      int di = F[4 * m + p], dj = F[5 * m + p], dk = F[6 * m + p];
      if (i < 1 || i > nx || j < 1 || j > ny || k < 1 || k > nz) return PDAC_ST_BAD_ARGUMENT;
      long c = IJK(i, j, k);
      s->num[a][c] = p + 1;
      s->fpq[a][p] = (int32_t)IJK(i + di, j + dj, k + dk);
      s->fpqq[a][p] = (it % 2 != 0 || it == 0) ? (int32_t)IJK(i + 2 * di, j + 2 * dj, k + 2 * dk) : -1;
    }
    base += m;
  }
  const int len[6] = {nx, nx, ny, ny, nz, nz};
  long off = 0;
  for (int a = 0; a < 6; ++a) {
    s->crd[a] = (double *)malloc(sizeof(double) * (size_t)len[a]);
    memcpy(s->crd[a], coords + off, sizeof(double) * (size_t)len[a]);
    off += len[a];
  }
  s->bmask = (int32_t *)malloc(sizeof(int32_t) * 6 * (size_t)n);
  memcpy(s->bmask, bmask, sizeof(int32_t) * 6 * (size_t)n);
  s->ivf = (double *)malloc(sizeof(double) * (size_t)n);
  memcpy(s->ivf, ivf, sizeof(double) * (size_t)n);
  s->fc2 = (uint8_t *)calloc((size_t)n, 1);
  s->ib = 1;
  return 0;
}

/* Tests only: leave parts of the immersed-boundary runtime out, to show that
 * the reference fixtures pin each (bit 0: the boundary forcing, 1: the
 * tilde forcing, 2: the face masks and 1/vf, 3: the momentum stencil
 * masks). */
int pdac_cpu_set_ib_off(pdac_cpu_ctx *s, int bits) {
  if (!s) return PDAC_ST_BAD_ARGUMENT;
  s->ib_off = bits;
  return 0;
}

/* faces, immb.f:1263-1277: the face masks and 1/vf of a cell (1 without
 * immersed boundaries, face_init ghost.f:3322-3345) */
static void cell_faces(const pdac_cpu_ctx *s, long c, double *be, double *bw, double *bn,
                       double *bs, double *bt, double *bb, double *ivf) {
  if (!s->ib || (s->ib_off & 4)) {
    *be = *bw = *bn = *bs = *bt = *bb = 1.0;
    *ivf = 1.0;
    return;
  }
  long n = s->n;
  *be = (double)s->bmask[c]; *bw = (double)s->bmask[n + c];
  *bn = (double)s->bmask[2 * n + c]; *bs = (double)s->bmask[3 * n + c];
  *bt = (double)s->bmask[4 * n + c]; *bb = (double)s->bmask[5 * n + c];
  *ivf = s->ivf[c];
}

/* grid_setup, grid.f:284-340 */
int pdac_cpu_set_grid(pdac_cpu_ctx *s, const double *dx, const double *dy,
                      const double *dz, const int32_t *flag,
                      const double *p_atm, const double *t_atm) {
  int nx = s->nx, ny = s->ny, nz = s->nz;
  for (int i = 1; i <= nx; ++i) s->dx[i] = dx[i - 1];
  for (int k = 1; k <= nz; ++k) s->dz[k] = dz[k - 1];
  for (int j = 1; j <= ny; ++j) s->dy[j] = s->three ? dy[j - 1] : 0.0;
  /* padding cells are never read by computed cells */
  s->dx[0] = s->dx[1]; s->dx[nx + 1] = s->dx[nx];
  s->dz[0] = s->dz[1]; s->dz[nz + 1] = s->dz[nz];
  s->dy[0] = s->dy[1]; s->dy[ny + 1] = s->dy[ny];
  for (int i = 0; i <= nx + 1; ++i) s->indx[i] = 1.0 / s->dx[i];
  for (int k = 0; k <= nz + 1; ++k) s->indz[k] = 1.0 / s->dz[k];
  for (int j = 0; j <= ny + 1; ++j) s->indy[j] = s->three ? 1.0 / s->dy[j] : 0.0;
  if (s->d.itc == 0 || s->three) {
    for (int i = 0; i <= nx + 1; ++i) s->r[i] = s->rb[i] = s->inr[i] = s->inrb[i] = 1.0;
  } else {
    double xb = 0.0, x;
    for (int i = 1; i <= nx; ++i) {
      if (i == 1) { xb = 0.0; x = -0.5 * s->dx[1] + xb; }
      else { xb = xb + s->dx[i]; x = xb - 0.5 * s->dx[i]; }
      s->r[i] = x; s->rb[i] = xb;
    }
    s->inrb[1] = (s->rb[1] == 0.0) ? 1.0e10 : 1.0 / s->rb[1];
    for (int i = 2; i <= nx; ++i) s->inrb[i] = 1.0 / s->rb[i];
    for (int i = 1; i <= nx; ++i) s->inr[i] = 1.0 / s->r[i];
    s->r[0] = s->r[1]; s->rb[0] = s->rb[1]; s->inr[0] = s->inr[1]; s->inrb[0] = s->inrb[1];
    s->r[nx + 1] = s->r[nx]; s->rb[nx + 1] = s->rb[nx];
    s->inr[nx + 1] = s->inr[nx]; s->inrb[nx + 1] = s->inrb[nx];
  }
  memcpy(s->flag, flag, sizeof(int32_t) * (size_t)s->n);
  for (int k = 1; k <= nz; ++k) {
    s->p_atm[k] = p_atm ? p_atm[k - 1] : 0.0;
    s->t_atm[k] = t_atm ? t_atm[k - 1] : 0.0;
  }
  s->grid_set = 1;
  return 0;
}

/* particles_constants_set, particles.f:35-78 */
int pdac_cpu_set_params(pdac_cpu_ctx *s, const pdac_params *prm) {
  s->prm = *prm;
  int ns = s->ns;
  for (int k = 0; k < ns; ++k) {
    s->inrl[k] = 1.0 / prm->rl[k];
    s->phi[k] = F(0.63);
  }
  for (int k = 0; k < ns; ++k)
    for (int kk = 0; kk < ns; ++kk) {
      int k1, k2;
      if (prm->dk[k] >= prm->dk[kk]) { k1 = k; k2 = kk; } else { k1 = kk; k2 = k; }
      double a = sqrt(prm->dk[k2] / prm->dk[k1]);
      s->philim[k][kk] = s->phi[k1] / (s->phi[k1] + (1.0 - s->phi[k1]) * s->phi[k2]);
      s->epsl[k][kk] = (s->phi[k1] - s->phi[k2]) + ((1.0 - a) * (1.0 - s->phi[k1]) * s->phi[k2]);
      double fact = (s->phi[k1] + (1.0 - s->phi[k2]) * s->phi[k1]) / s->phi[k1];
      s->epsl[k][kk] = s->epsl[k][kk] * fact;
      s->epsu[k][kk] = (1.0 - a) * (s->phi[k1] + (1.0 - s->phi[k1]) * s->phi[k2]);
      double sdk = prm->dk[k] + prm->dk[kk];
      double dk3 = prm->dk[k] * prm->dk[k] * prm->dk[k];
      double dkk3 = prm->dk[kk] * prm->dk[kk] * prm->dk[kk];
      s->dkf[k][kk] = (sdk * sdk) / (prm->rl[k] * dk3 + prm->rl[kk] * dkk3);
      s->dkf[kk][k] = s->dkf[k][kk];
    }
  s->params_set = 1;
  return 0;
}

int pdac_cpu_field_ncomp(const pdac_cpu_ctx *s, int id) { return ncomp_of(s, id); }

int pdac_cpu_put(pdac_cpu_ctx *s, int id, const double *h, int64_t n) {
  int nc = ncomp_of(s, id);
  if (nc <= 0 || n != (int64_t)nc * s->n) return PDAC_ST_BAD_ARGUMENT;
  memcpy(s->f[id], h, sizeof(double) * (size_t)n);
  return 0;
}

int pdac_cpu_get(pdac_cpu_ctx *s, int id, double *h, int64_t n) {
  int nc = ncomp_of(s, id);
  if (nc <= 0 || n != (int64_t)nc * s->n) return PDAC_ST_BAD_ARGUMENT;
  memcpy(h, s->f[id], sizeof(double) * (size_t)n);
  return 0;
}

/* ------------------------------------------------------------------ */
/* Neighbour indices.
 * Raw neighbours (myijk) clamp the offset cell into the mesh, which is what
 * cell_neighbours builds for cells with 2 <= i <= nx-1 (ghost.f:530-545).
 * Reflected neighbours (myinds) follow set_myinds (ghost.f:866-2476): a wall
 * or filled_cell_1 neighbour maps to the cell itself, second and third
 * neighbours fall back to nearer ones at walls and next to the boundary. */
typedef struct {
  long c;
  /* raw */
  long ipjk, imjk, ijpk, ijmk, ijkp, ijkm, ippjk, immjk, ijppk, ijmmk, ijkpp, ijkmm;
  long ipjpk, imjpk, ipjmk, imjmk, ipjkp, imjkp, ipjkm, imjkm, ijpkp, ijmkp, ijpkm, ijmkm;
  long ipppjk, ijpppk, ijkppp, ippjpk, ippjkp, ipjppk, ijppkp, ipjkpp, ijpkpp;
  /* reflected */
  long e, w, n, s_, t, b, ee, ww, nn, ss, tt, bb;
  long en, es, et, eb, nt, wn, nb, wt, st;
  long eee, nnn, ttt, een, eet, enn, nnt, ett, ntt;
} nbr;

static long rawn(const pdac_cpu_ctx *s, int i, int j, int k, int di, int dj, int dk) {
  int ii = i + di, jj = j + dj, kk = k + dk;
  if (ii < 1) ii = 1; if (ii > s->nx) ii = s->nx;
  if (jj < 1) jj = 1; if (jj > s->ny) jj = s->ny;
  if (kk < 1) kk = 1; if (kk > s->nz) kk = s->nz;
  return IJK(ii, jj, kk);
}

/* first-level diagonal rule of set_myinds 2D (e.g. ghost.f:990-1000) */
static long refl_diag(const pdac_cpu_ctx *s, long self, long corner, long a, long b) {
  const int32_t *fl = s->flag;
  if (!wall(fl[corner])) return corner;
  int wa = wall(fl[a]), wb = wall(fl[b]);
  if (wa && !wb) return b;
  if (!wa && wb) return a;
  return self;
}

/* second-level diagonal rule of set_myinds 2D (e.g. ghost.f:1095-1113) */
static long refl_diag2(const pdac_cpu_ctx *s, long base, long corner, long a, long b) {
  const int32_t *fl = s->flag;
  if (!wallf(fl[corner])) return corner;
  int wa = wall(fl[a]), wb = wall(fl[b]);
  if (wa && !wb) return b;
  if (!wa && wb) return a;
  return base;
}

static void neighbours(const pdac_cpu_ctx *s, int i, int j, int k, nbr *q) {
  const int32_t *fl = s->flag;
  int nx = s->nx, ny = s->ny, nz = s->nz;
  long c = IJK(i, j, k);
  q->c = c;
  q->ipjk = rawn(s, i, j, k, 1, 0, 0);  q->imjk = rawn(s, i, j, k, -1, 0, 0);
  q->ijkp = rawn(s, i, j, k, 0, 0, 1);  q->ijkm = rawn(s, i, j, k, 0, 0, -1);
  q->ippjk = rawn(s, i, j, k, 2, 0, 0); q->immjk = rawn(s, i, j, k, -2, 0, 0);
  q->ijkpp = rawn(s, i, j, k, 0, 0, 2); q->ijkmm = rawn(s, i, j, k, 0, 0, -2);
  q->ipjkp = rawn(s, i, j, k, 1, 0, 1); q->imjkp = rawn(s, i, j, k, -1, 0, 1);
  q->ipjkm = rawn(s, i, j, k, 1, 0, -1); q->imjkm = rawn(s, i, j, k, -1, 0, -1);
  q->ipppjk = rawn(s, i, j, k, 3, 0, 0); q->ijkppp = rawn(s, i, j, k, 0, 0, 3);
  q->ippjkp = rawn(s, i, j, k, 2, 0, 1); q->ipjkpp = rawn(s, i, j, k, 1, 0, 2);
  if (s->three) {
    q->ijpk = rawn(s, i, j, k, 0, 1, 0);  q->ijmk = rawn(s, i, j, k, 0, -1, 0);
    q->ijppk = rawn(s, i, j, k, 0, 2, 0); q->ijmmk = rawn(s, i, j, k, 0, -2, 0);
    q->ipjpk = rawn(s, i, j, k, 1, 1, 0); q->imjpk = rawn(s, i, j, k, -1, 1, 0);
    q->ipjmk = rawn(s, i, j, k, 1, -1, 0); q->imjmk = rawn(s, i, j, k, -1, -1, 0);
    q->ijpkp = rawn(s, i, j, k, 0, 1, 1); q->ijmkp = rawn(s, i, j, k, 0, -1, 1);
    q->ijpkm = rawn(s, i, j, k, 0, 1, -1); q->ijmkm = rawn(s, i, j, k, 0, -1, -1);
    q->ijpppk = rawn(s, i, j, k, 0, 3, 0); q->ippjpk = rawn(s, i, j, k, 2, 1, 0);
    q->ipjppk = rawn(s, i, j, k, 1, 2, 0); q->ijppkp = rawn(s, i, j, k, 0, 2, 1);
    q->ijpkpp = rawn(s, i, j, k, 0, 1, 2);
  } else {
    q->ijpk = q->ijmk = q->ijppk = q->ijmmk = c;
    q->ipjpk = q->imjpk = q->ipjmk = q->imjmk = c;
    q->ijpkp = q->ijmkp = q->ijpkm = q->ijmkm = c;
    q->ijpppk = q->ippjpk = q->ipjppk = q->ijppkp = q->ijpkpp = c;
  }
  /* first neighbours (ghost.f:963-994, 1556-1590) */
  q->e = wallf(fl[q->ipjk]) ? c : q->ipjk;
  q->w = wallf(fl[q->imjk]) ? c : q->imjk;
  q->t = wallf(fl[q->ijkp]) ? c : q->ijkp;
  q->b = wallf(fl[q->ijkm]) ? c : q->ijkm;
  if (s->three) {
    q->n = wallf(fl[q->ijpk]) ? c : q->ijpk;
    q->s_ = wallf(fl[q->ijmk]) ? c : q->ijmk;
  } else {
    q->n = q->s_ = c;
  }
  /* second neighbours (ghost.f:1060-1093, 1592-1638) */
  q->ee = wallf(fl[q->ippjk]) ? q->ipjk : q->ippjk;
  if (i == nx - 1) q->ee = q->e;
  q->ww = wallf(fl[q->immjk]) ? q->imjk : q->immjk;
  if (i == 2) q->ww = q->w;
  q->tt = wallf(fl[q->ijkpp]) ? q->ijkp : q->ijkpp;
  if (k == nz - 1) q->tt = q->t;
  q->bb = wallf(fl[q->ijkmm]) ? q->ijkm : q->ijkmm;
  if (k == 2) q->bb = q->b;
  if (s->three) {
    /* in 3D the y-direction fallbacks are the reflected first neighbours
     * (ghost.f:1610-1622) */
    q->nn = wallf(fl[q->ijppk]) ? q->n : q->ijppk;
    if (j == ny - 1) q->nn = q->n;
    q->ss = wallf(fl[q->ijmmk]) ? q->s_ : q->ijmmk;
    if (j == 2) q->ss = q->s_;
  } else {
    q->nn = q->ss = c;
  }
  /* third neighbours (ghost.f:1256-1274 in 2D and the 3D analogue) */
  q->eee = wallf(fl[q->ipppjk]) ? q->ippjk : q->ipppjk;
  if (i == nx - 1) q->eee = q->e;
  if (i == nx - 2) q->eee = q->ee;
  q->ttt = wallf(fl[q->ijkppp]) ? q->ijkpp : q->ijkppp;
  if (k == nz - 1) q->ttt = q->t;
  if (k == nz - 2) q->ttt = q->tt;
  if (s->three) {
    q->nnn = wallf(fl[q->ijpppk]) ? q->ijppk : q->ijpppk;
    if (j == ny - 1) q->nnn = q->n;
    if (j == ny - 2) q->nnn = q->nn;
  } else {
    q->nnn = c;
  }
  if (!s->three) {
    /* 2D diagonals (ghost.f:984-1058) and second-level diagonals */
    q->wt = refl_diag(s, c, q->imjkp, q->ijkp, q->imjk);
    q->eb = refl_diag(s, c, q->ipjkm, q->ijkm, q->ipjk);
    q->et = refl_diag(s, c, q->ipjkp, q->ijkp, q->ipjk);
    q->eet = refl_diag2(s, q->ipjk, q->ippjkp, q->ipjkp, q->ippjk);
    if (i == nx - 1) q->eet = q->et;
    q->ett = refl_diag2(s, q->ijkp, q->ipjkpp, q->ipjkp, q->ijkpp);
    if (k == nz - 1) q->ett = q->et;
    q->en = q->es = q->nt = q->wn = q->nb = q->st = c;
    q->een = q->enn = q->nnt = q->ntt = c;
  } else {
    /* 3D diagonals are raw except the bottom ones (ghost.f:1640-1780) */
    q->en = q->ipjpk; q->wn = q->imjpk; q->es = q->ipjmk;
    q->et = q->ipjkp; q->wt = q->imjkp; q->nt = q->ijpkp; q->st = q->ijmkp;
    q->eb = wallf(fl[q->ipjkm]) ? q->ipjk : q->ipjkm;
    q->nb = wallf(fl[q->ijpkm]) ? q->ijpk : q->ijpkm;
    q->eet = q->ippjkp; if (i == nx - 1) q->eet = q->et;
    q->een = q->ippjpk; if (i == nx - 1) q->een = q->en;
    q->nnt = q->ijppkp; if (j == ny - 1) q->nnt = q->nt;
    q->enn = q->ipjppk; if (j == ny - 1) q->enn = q->en;
    q->ett = q->ipjkpp; if (k == nz - 1) q->ett = q->et;
    q->ntt = q->ijpkpp; if (k == nz - 1) q->ntt = q->nt;
  }
}

/* ------------------------------------------------------------------ */
/* stencils (subscr.f:63-210) */
typedef struct {
  double c, e, w, n, s, t, b, ee, ww, nn, ss, tt, bb;
  double en, es, et, eb, nt, wn, nb, wt, st;
  double eee, nnn, ttt, een, eet, enn, nnt, ett, ntt;
} stc;

/* nb (subscr.f:2323-2383): reflected full stencil */
static void g_nb(const pdac_cpu_ctx *s, const nbr *q, const double *a, stc *o) {
  o->c = a[q->c]; o->e = a[q->e]; o->w = a[q->w]; o->t = a[q->t]; o->b = a[q->b];
  o->ee = a[q->ee]; o->tt = a[q->tt]; o->et = a[q->et]; o->eb = a[q->eb];
  o->wt = a[q->wt]; o->eee = a[q->eee]; o->ttt = a[q->ttt];
  o->eet = a[q->eet]; o->ett = a[q->ett];
  if (s->three) {
    o->n = a[q->n]; o->s = a[q->s_]; o->nn = a[q->nn]; o->en = a[q->en];
    o->es = a[q->es]; o->nt = a[q->nt]; o->wn = a[q->wn]; o->nb = a[q->nb];
    o->st = a[q->st]; o->nnn = a[q->nnn]; o->een = a[q->een];
    o->enn = a[q->enn]; o->nnt = a[q->nnt]; o->ntt = a[q->ntt];
  }
}

/* rnb_r (subscr.f:2386-2450): raw full stencil */
static void g_rnb(const pdac_cpu_ctx *s, const nbr *q, const double *a, stc *o) {
  o->c = a[q->c]; o->e = a[q->ipjk]; o->w = a[q->imjk]; o->t = a[q->ijkp];
  o->b = a[q->ijkm]; o->wt = a[q->imjkp]; o->et = a[q->ipjkp]; o->eb = a[q->ipjkm];
  o->ee = a[q->ippjk]; o->tt = a[q->ijkpp]; o->eet = a[q->ippjkp];
  o->ett = a[q->ipjkpp]; o->eee = a[q->ipppjk]; o->ttt = a[q->ijkppp];
  if (s->three) {
    o->n = a[q->ijpk]; o->s = a[q->ijmk]; o->nn = a[q->ijppk];
    o->en = a[q->ipjpk]; o->es = a[q->ipjmk]; o->nt = a[q->ijpkp];
    o->wn = a[q->imjpk]; o->nb = a[q->ijpkm]; o->st = a[q->ijmkp];
    o->nnn = a[q->ijpppk]; o->een = a[q->ippjpk]; o->enn = a[q->ipjppk];
    o->nnt = a[q->ijppkp]; o->ntt = a[q->ijpkpp];
  }
}

/* first_nb / third_nb / first_rnb (subscr.f:2196-2320) */
static void g_first_nb(const pdac_cpu_ctx *s, const nbr *q, const double *a, stc *o) {
  o->c = a[q->c]; o->e = a[q->e]; o->w = a[q->w]; o->t = a[q->t]; o->b = a[q->b];
  if (s->three) { o->n = a[q->n]; o->s = a[q->s_]; }
}
static void g_third_nb(const pdac_cpu_ctx *s, const nbr *q, const double *a, stc *o) {
  o->ee = a[q->ee]; o->ww = a[q->ww]; o->tt = a[q->tt]; o->bb = a[q->bb];
  if (s->three) { o->nn = a[q->nn]; o->ss = a[q->ss]; }
}
static void g_first_rnb(const pdac_cpu_ctx *s, const nbr *q, const double *a, stc *o) {
  o->c = a[q->c]; o->e = a[q->ipjk]; o->w = a[q->imjk]; o->t = a[q->ijkp];
  o->b = a[q->ijkm];
  if (s->three) { o->n = a[q->ijpk]; o->s = a[q->ijmk]; }
}

/* ------------------------------------------------------------------ */
/* limiters, limiters.f:14-56 */
double pdac_cpu_limiter(int lim_type, double beta, double erre) {
  double l, bu, m;
  switch (lim_type) {
  case 1: l = 2.0 * erre / (1.0 + erre); return l > 0.0 ? l : 0.0;
  case 2: m = erre < 1.0 ? erre : 1.0; return m > 0.0 ? m : 0.0;
  case 3: {
    double a = 2.0 * erre < 1.0 ? 2.0 * erre : 1.0;
    double b = erre < 2.0 ? erre : 2.0;
    double r = a > b ? a : b;
    return r > 0.0 ? r : 0.0;
  }
  case 4:
    bu = (1.0 - beta) + beta * erre;
    m = 2.0 * erre;
    if (bu < m) m = bu;
    if (2.0 < m) m = 2.0;
    return m > 0.0 ? m : 0.0;
  case 0: return (1.0 - beta) + beta * erre;
  default: return 0.0;
  }
}
#define LIM(t, e) pdac_cpu_limiter((t), s->prm.beta, (e))

/* ------------------------------------------------------------------ */
/* mole, eosg.f:38-68 */
static void mole(const pdac_cpu_ctx *s, double *xg, const double *yg) {
  double mol = 0.0;
  for (int g = 0; g < s->ng; ++g) mol = mol + yg[g] / s->gmw[s->prm.gas_type[g] - 1];
  double molinv = (fabs(mol) <= 1.0e-10) ? 0.0 : 1.0 / mol;
  for (int g = 0; g < s->ng; ++g) xg[g] = yg[g] / s->gmw[s->prm.gas_type[g] - 1] * molinv;
}

/* thermal_eosg, eosg.f:114-137 */
static double thermal_eosg(const pdac_cpu_ctx *s, double tg, double p, const double *xg) {
  double mg = 0.0;
  for (int g = 0; g < s->ng; ++g) mg = mg + xg[g] * s->gmw[s->prm.gas_type[g] - 1];
  return p / (RGAS * tg) * mg;
}

/* hcapg, hcapgs.f:33-85 */
static void hcapg(const pdac_params *prm, const double *gmw, double tg, double *cpc) {
  double t1 = (prm->icpc <= 1) ? prm->tref : tg;
  double t2 = t1 * t1, t3 = t1 * t2;
  cpc[0] = (2.811e1 - 3.68e-6 * t1 + 1.746e-5 * t2 - 1.065e-8 * t3);
  cpc[1] = (3.115e1 - 1.357e-2 * t1 + 2.680e-5 * t2 - 1.168e-8 * t3);
  cpc[2] = (1.980e1 + 7.344e-2 * t1 - 5.602e-5 * t2 + 1.715e-8 * t3);
  cpc[3] = (2.714e1 + 9.274e-3 * t1 - 1.381e-5 * t2 + 7.645e-9 * t3);
  cpc[4] = (3.224e1 + 1.924e-3 * t1 + 1.055e-5 * t2 - 3.596e-9 * t3);
  cpc[5] = (3.0413e1 - 1.059e-2 * t1 + 2.458e-5 * t2 - 1.135e-8 * t3);
  cpc[6] = (2.385e1 + 6.699e-2 * t1 - 4.961e-5 * t2 + 1.328e-8 * t3);
  for (int g = 0; g < 7; ++g) cpc[g] = cpc[g] / gmw[g];
  if (prm->icpc == 0) { cpc[4] = 1810.0; cpc[5] = 1004.0; }
}

/* hcapg exported for the unit pins: cpc[7] per kilogram for temperature tg
 * under prm->icpc and prm->tref, with the gas molar masses gmw[7] */
void pdac_cpu_hcapg(const pdac_params *prm, const double *gmw, double tg, double *cpc) {
  hcapg(prm, gmw, tg, cpc);
}

/* caloric_eosg, eosg.f:139-223; returns info (0 ok) */
static int caloric_eosg(const pdac_params *prm, const double *gmw, int ng,
                        double sieg, const double *yg, double *tg, double *cg) {
  double cpc[7], cgas = 0.0;
  int info = 1;
  for (int ii = 1; ii <= 2000; ++ii) {
    double tgnn = *tg;
    hcapg(prm, gmw, *tg, cpc);
    cgas = 0.0;
    for (int g = 0; g < ng; ++g) cgas = cpc[prm->gas_type[g] - 1] * yg[g] + cgas;
    *tg = TZERO + (sieg - HZEROG) / cgas;
    if (fabs((tgnn - *tg) / tgnn) <= 1.0e-8) { info = 0; break; }
  }
  *cg = cgas;
  if (*tg <= 0.0) info = 1;
  return info;
}

int pdac_cpu_caloric_eosg(const pdac_params *prm, int ngas, double sieg,
                          const double *ygc, double *tg, double *cg) {
  double gmw[7];
  for (int g = 0; g < 7; ++g) gmw[g] = GMW_BASE[g] * 1.0e-3;
  return caloric_eosg(prm, gmw, ngas, sieg, ygc, tg, cg);
}

/* viscon, visc.f:37-148 */
static void viscon(const pdac_cpu_ctx *s, const double *xgc, double tg, double *mu, double *kap) {
  double mmug[7] = {0}, ckg[7], phij[7][7];
  for (int g = 0; g < s->ng; ++g) {
    int igg = s->prm.gas_type[g] - 1;
    double tst = tg / MMUGEK[igg];
    double om = 1.16145 * pow(tst, -0.14874) + 0.52487 * exp(-0.77320 * tst) +
                2.16178 * exp(-2.43787 * tst);
    double m = 1.0e3 * s->gmw[igg];
    mmug[igg] = 26.69 * sqrt(m * tg) / (MMUGS[igg] * MMUGS[igg] * om);
    mmug[igg] = mmug[igg] * 1.0e-7;
  }
  double t1 = tg, t2 = t1 * tg, t3 = t2 * tg;
  double tr = tg / 132.5;
  double trm1 = 1.0 / tr, trm2 = trm1 * trm1, trm3 = trm2 * trm1, trm4 = trm3 * trm1;
  double trsrt = sqrt(tr);
  ckg[0] = -3.273e-4 + 9.966e-5 * t1 - 3.743e-8 * t2 + 9.732e-12 * t3;
  ckg[1] = 3.919e-4 + 9.816e-5 * t1 - 5.067e-8 * t2 + 1.504e-11 * t3;
  ckg[2] = -7.215e-3 + 8.015e-5 * t1 + 5.477e-9 * t2 - 1.053e-11 * t3;
  ckg[3] = 8.099e-3 + 6.689e-4 * t1 - 4.158e-7 * t2 + 1.562e-10 * t3;
  ckg[4] = 7.341e-3 - 1.013e-5 * t1 + 1.801e-7 * t2 - 9.100e-11 * t3;
  ckg[5] = 25.9778e-3 * (0.2395 * tr + 6.4977e-3 * trsrt + 1.0 - 1.92615 * trm1 +
                         2.00383 * trm2 - 1.07553 * trm3 + 0.229414 * trm4);
  ckg[6] = -8.086e-3 + 6.344e-5 * t1 - 1.382e-8 * t2 + 2.303e-12 * t3;
  for (int g = 0; g < s->ng; ++g) {
    int igg = s->prm.gas_type[g] - 1;
    for (int h = 0; h < s->ng; ++h) {
      int jgg = s->prm.gas_type[h] - 1;
      double aa = s->gmw[igg] / s->gmw[jgg];
      double bb = mmug[igg] / mmug[jgg];
      double cc = 1.0 + sqrt(bb) * pow(aa, -0.25);
      phij[igg][jgg] = cc * cc / sqrt(8.0 * (1.0 + aa));
    }
  }
  double m_ = 0.0, k_ = 0.0;
  for (int g = 0; g < s->ng; ++g) {
    int igg = s->prm.gas_type[g] - 1;
    double sum = 0.0;
    for (int h = 0; h < s->ng; ++h) sum = sum + xgc[h] * phij[igg][s->prm.gas_type[h] - 1];
    double suminv = 1.0 / sum;
    m_ = m_ + xgc[g] * mmug[igg] * suminv;
    k_ = k_ + xgc[g] * ckg[igg] * suminv;
  }
  *mu = m_;
  *kap = k_;
}

/* ------------------------------------------------------------------ */
/* drag, drag.f:11-157 */
static double fmin_f(double a, double b) { return a < b ? a : b; }

double pdac_cpu_kdrags(const pdac_params *prm, int is, double du, double dv,
                       double dw, double ep, double rgp, double rlk, double mug) {
  double vrel = sqrt(du * du + dv * dv + dw * dw);
  double inrl = 1.0 / prm->rl[is];
  double dk = prm->dk[is], phis = prm->phis[is];
  double drag;
  if (ep < 0.8) {
    /* kdragl, drag.f:120-143 (Ergun) */
    double denom = 1.0 / (dk * phis * ep);
    drag = 150.0 * ep * rlk * inrl * mug * denom + 1.75 * rgp * vrel;
    drag = drag * rlk * inrl * denom;
  } else {
    /* kdragg, drag.f:83-118 (Wen-Yu) */
    double eps = rlk * inrl;
    double reynum = rgp * dk * phis * vrel / mug;
    double cd;
    if (reynum <= 1000.0)
      cd = (24.0 / reynum) * (1.0 + F(0.15) * pow(reynum, F(0.687)));
    else
      cd = F(0.43828814);
    cd = fmin_f(cd, 1.0e4);
    drag = 0.75 * cd * eps / (dk * phis) * rgp * vrel;
    drag = drag * pow(ep, F(-2.65)); /* f(ep), drag.f:145-157 */
  }
  return drag;
}

double pdac_cpu_sdrag(const pdac_params *prm, int is, double du, double dv,
                      double dw, double ep, double rgp, double rlk, double mug) {
  /* Syamlal-O'Brien, drag.f:11-58 */
  double vrel = sqrt(du * du + dv * dv + dw * dw);
  double inrl = 1.0 / prm->rl[is];
  double dk = prm->dk[is], phis = prm->phis[is];
  double eps = rlk * inrl;
  double reynum = rgp * dk * phis * vrel / mug;
  double a = pow(ep, F(4.14)), b;
  if (ep <= F(0.85)) b = F(0.8) * pow(ep, F(1.28));
  else b = pow(ep, F(2.65));
  double vr = 0.5 * (a - F(0.06) * reynum +
                     sqrt(F(0.0036) * (reynum * reynum) + F(0.12) * reynum * (2.0 * b - a) + a * a));
  double q = 0.63 + 4.8 / sqrt(reynum);
  double cd = fmin_f(q * q, 1.0e4);
  double drag = 0.75 * cd * eps / (dk * phis) * rgp * vrel;
  return drag / vr;
}

/* ppdrag, drag.f:248-298 (k, kk are 1-based phase indices, both >= 2) */
static double ppdrag(const pdac_cpu_ctx *s, const double *rlkc, int k, int kk) {
  const pdac_params *prm = &s->prm;
  int k1, k2;
  if (prm->dk[k - 2] >= prm->dk[kk - 2]) { k1 = k - 2; k2 = kk - 2; }
  else { k1 = kk - 2; k2 = k - 2; }
  double ep1 = rlkc[k1] * s->inrl[k1];
  double ep2 = rlkc[k2] * s->inrl[k2];
  double epsum = ep1 + ep2;
  if (ep1 > 0.0 && ep2 > 0.0) {
    double xbar = ep1 / epsum, epkl;
    if (xbar <= s->philim[k - 2][kk - 2]) epkl = s->epsl[k - 2][kk - 2] * xbar + s->phi[k2];
    else epkl = s->epsu[k - 2][kk - 2] * (1.0 - xbar) + s->phi[k1];
    double espo = 1.0 / 3.0;
    double m1 = pow(epkl, espo), m2 = pow(epsum, espo);
    double effe = (3.0 * m1 + m2) / (2.0 * (m1 - m2));
    return 1.0 * (1.0 + 1.0) * rlkc[k - 2] * rlkc[kk - 2] * s->dkf[k - 2][kk - 2] * effe;
  }
  return 0.0;
}

/* hvs, hvs.f:9-83 */
static double hvs(const pdac_cpu_ctx *s, double rlk, double rog, double du, double dv,
                  double dw, double mug, double kapg, double cg, int is) {
  if (rlk <= 0.0) return 0.0;
  double dk = s->prm.dk[is];
  double eps = rlk * s->inrl[is];
  double eps2 = eps * eps;
  double asurf = 6.0 * kapg * eps / (dk * dk);
  double vrel = sqrt(du * du + dv * dv + dw * dw);
  double reynum = dk * vrel * rog / mug;
  double pranum = cg * mug / kapg;
  double pranumexp = pow(pranum, 1.0 / 3.0);
  double nus = (2.0 + 5.0 * eps2) * (1.0 + 0.7 * pow(reynum, 0.2) * pranumexp);
  nus = nus + (0.13 + 1.2 * eps2) * pow(reynum, 0.7) * pranumexp;
  return nus * asurf;
}

/* Gauss-Jordan elimination with the two-phase collapse, matrix.f:353-397 */
static void gauss_jordan(int np, const double *plim, double a[MAXP][MAXP], double *b) {
  for (int l = 1; l < np; ++l) {
    if (fabs(a[l][l]) < plim[l - 1]) {
      for (int ll = 0; ll < np; ++ll) {
        a[ll][ll] = a[ll][ll] + a[l][ll];
        a[l][ll] = 0.0;
        a[ll][l] = 0.0;
      }
      b[l] = 0.0;
    }
  }
  for (int l = 0; l < np; ++l) {
    if (a[l][l] != 0.0) {
      double div = 1.0 / a[l][l];
      for (int lj = l + 1; lj < np; ++lj) a[l][lj] = a[l][lj] * div;
      b[l] = b[l] * div;
      a[l][l] = 0.0;
      for (int li = 0; li < np; ++li) {
        double amul = a[li][l];
        for (int lj = l + 1; lj < np; ++lj) a[li][lj] = a[li][lj] - amul * a[l][lj];
        b[li] = b[li] - amul * b[l];
      }
    }
  }
}

void pdac_cpu_gauss_jordan(int np, const double *plim, double *a, double *b) {
  double m[MAXP][MAXP];
  for (int i = 0; i < np; ++i)
    for (int j = 0; j < np; ++j) m[i][j] = a[i * np + j];
  gauss_jordan(np, plim, m, b);
  for (int i = 0; i < np; ++i)
    for (int j = 0; j < np; ++j) a[i * np + j] = m[i][j];
}

#include "pdac_oracle_step.inc"
#include "pdac_oracle_explicit.inc"

/* threads the parallel cell loops run on (OMP_NUM_THREADS) */
int pdac_cpu_threads(void) { return omp_get_max_threads(); }
