// pdac_dev.hpp -- device-side data structures and per-cell building blocks
// of the PDAC hot path on gfx950.
//
// Layout: every field is a plane of ntot doubles per component (Fortran
// column-major (ntot, ncomp)), cell index i fastest, so a wavefront of 64
// consecutive cells reads 512 contiguous bytes of each plane.
//
// Arithmetic keeps the operand order of the reference Fortran and its
// single-precision literals widened to double; the library is compiled with
// -ffp-contract=off so products and sums round separately.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/pdacgpu.h"

namespace pdac {

constexpr int kMaxS = PDAC_MAX_SOLID;  // particle classes supported on device
constexpr int kMaxP = kMaxS + 1;
constexpr int kMaxA = kMaxP * (kMaxP + 1) / 2;

// cell flags, grid.f:45-58
enum : int32_t {
  FLUID = 1, SLIP_WALL = 2, NOSLIP_WALL = 4, BL_CELL = 5, FREE_IO = 8,
  INLET_CELL = 16, ZERO_GRAD = 32, VENT_CELL = 64, DOME_CELL = 129,
  IMMB_CELL = 257, FILLED_CELL_1 = 512, FILLED_CELL_2 = 1024
};

#define FW(x) ((double)(float)(x))

struct Grid {
  const double *dx, *dy, *dz;          // 0..n+1, padded
  const double *indx, *indy, *indz;
  const double *r, *rb, *inr, *inrb;
  const int32_t *flag;
  const uint8_t *reg;                  // 1: the cell's full stencil needs no wall reflection
  const double *p_atm, *t_atm;         // 1..nz
  int nx, ny, nz;
  long n;
  int itc;                             // 1: 2D cylindrical
};

struct Consts {
  double alpha, alphagrav, gravx, gravy, gravz, beta, omega, delg, flim;
  double tref, void_fraction;
  double dk[kMaxS], rl[kMaxS], inrl[kMaxS], phis[kMaxS], cps[kMaxS], plim[kMaxS];
  double phi[kMaxS];
  double philim[kMaxS][kMaxS], epsl[kMaxS][kMaxS], epsu[kMaxS][kMaxS], dkf[kMaxS][kMaxS];
  double gmw[7], atm_ygc[7];
  // pow(gmw_g / gmw_h, -0.25) for gas slots g, h (the mixing rule of
  // viscon), evaluated once on the host: it depends on the species only
  double visw[7][7];
  int gas_type[7];
  int ngas, muscl, lm, lv, inmax, maxout, pmodel, drag_model, icpc, ord;
  // viscous stresses and conduction (gas_viscosity, part_viscosity,
  // repulsive_model) and the particle viscosity and conductivity per class
  int gas_visc, part_visc, repulsive;
  double mus[kMaxS], kap[kMaxS];
  int iturb;  // 1: Smagorinsky sub-grid viscosity and conductivity (mugt, kapgt)
};

// All device arrays of one context.
struct Fields {
  double *f[PDAC_NFIELDS];
  double *mf[3][3];                    // momentum fluxes [u,v,w][e,n,t], (ntot, P)
  double *egf[3], *esf[3], *yf[3];     // enthalpy / species fluxes
  double *conv, *abeta;
  double *appu_old[3];
  double *snap_rgp, *snap_rlk;
  int32_t *ctr;                        // device counters (iter, ftem)
  unsigned long long *visits;          // work-list entries over all sweeps
  int32_t *wl;                         // work list of one colour phase
  // sweep flags (iter), one byte per cell in colour-major order (sweep_slot):
  // sdirty, the cell must be checked again; srefresh, the cell of the other
  // colour in the same pair changed in the previous phase and its snapshot
  // is to be refreshed
  uint8_t *sdirty, *srefresh;
  // immersed boundaries: per cell the face masks b_e..b_b and whether 1/vf
  // differs from 1 (IBM_* bits), and 1/vf; null without immersed boundaries
  const uint8_t *ibm;
  const double *ivf;
  // viscous stress divergence along x, y, z (ntot, P): gas, then each
  // class; null when neither phase is viscous
  double *vis[3];
  // with the sub-grid model: the gas viscosity and conductivity with their
  // turbulent parts (mugt, kapgt) and the squared Smagorinsky length of each
  // cell (smag); null without it
  double *mugt, *kapgt;
  const double *smag;
  // gas-particle drag coefficient of each class (ntot, S): k_drag writes it
  // once per cell for the axis threads of k_tilde
  double *kpgv;
  // particle classes present near each tile of 4x4x4 cells: bit s of ptile
  // is set when a cell of the tile or of a face/edge/corner-adjacent tile
  // holds class s (rlk != 0); ptile_own before that dilation.  tnx, tny:
  // tiles along x and y.  Null: no tile test (every class is computed).
  uint32_t *ptile, *ptile_own;
  int tnx, tny, tnz;
  // the tiles listed for the particle-class flux kernels of a 3D mesh (tlist,
  // *tcount entries): those with a class present near them now or at the
  // consuming kernel's previous call.  Bits 16.. of ptile mark the classes
  // present at that call only, whose fluxes the consumer sets to zero.
  // ptile_mprev, ptile_hprev: the class bits of the momentum and enthalpy
  // flux kernels' previous calls (0xffff: unknown).
  int32_t *tlist, *tcount;
  uint32_t *ptile_mprev, *ptile_hprev;
};

// tile of 4x4x4 cells holding cell (i, j, k), 1-based
__host__ __device__ __forceinline__ long phase_tile(const Fields &F, int i, int j, int k) {
  return (long)((i - 1) >> 2) + (long)F.tnx * ((long)((j - 1) >> 2) + (long)F.tny * (long)((k - 1) >> 2));
}

// cell (i, j, k), 1-based, of lane 0..63 of tile t (x fastest, then y, z)
__device__ __forceinline__ void tile_cell(const Fields &F, long t, int lane, int &i, int &j, int &k) {
  const int tx = (int)(t % F.tnx), ty = (int)((t / F.tnx) % F.tny), tz = (int)(t / ((long)F.tnx * F.tny));
  i = 4 * tx + (lane & 3) + 1;
  j = 4 * ty + ((lane >> 2) & 3) + 1;
  k = 4 * tz + (lane >> 4) + 1;
}

// face masks of a cell (b_e_ .. b_b_, immb.f:56-58): bits of Fields::ibm
enum : uint8_t { IBM_E = 1, IBM_W = 2, IBM_N = 4, IBM_S = 8, IBM_T = 16, IBM_B = 32, IBM_VF = 64 };
constexpr uint8_t IBM_OPEN = 63;  // every face open, 1/vf = 1 (face_init, ghost.f:3322-3345)

// faces, immb.f:1263-1277: the face masks and 1/vf of a cell, as doubles
// (the reference multiplies its integer masks into double fluxes)
struct FaceMask {
  double e, w, n, s, t, b, ivf;
};

__device__ __forceinline__ FaceMask mask_of(uint8_t m, const double *ivf, long c) {
  FaceMask r;
  r.e = (m & IBM_E) ? 1.0 : 0.0;
  r.w = (m & IBM_W) ? 1.0 : 0.0;
  r.n = (m & IBM_N) ? 1.0 : 0.0;
  r.s = (m & IBM_S) ? 1.0 : 0.0;
  r.t = (m & IBM_T) ? 1.0 : 0.0;
  r.b = (m & IBM_B) ? 1.0 : 0.0;
  r.ivf = (m & IBM_VF) ? ivf[c] : 1.0;
  return r;
}

__device__ __forceinline__ FaceMask cell_faces(const Fields &F, long c) {
  return mask_of(F.ibm ? F.ibm[c] : IBM_OPEN, F.ivf, c);
}

// Forcing points of one velocity component on the device (TYPE
// forcing_point, immb.f:33-45, with the geometry of velint3d,
// interp.f:687-781, evaluated once on the host)
struct IbAxis {
  int nfp;
  const int32_t *cell, *q, *qq;  // 0-based cell, first and second interpolation neighbour
  const uint8_t *mode;           // 0: h/zA v(q) (int >= 20), 1: -h/zA v(q), 2: the smaller of two
  const uint8_t *ext;            // int >= 20: homogeneous Neumann values (hn)
  const double *hza, *zbh, *hza2, *zbza;  // h/zA, zB-h, h-zA, zB-zA
  double *vel;                   // pseudo-velocity of each phase: NP planes of nfp
};

struct Ib {
  IbAxis ax[3];
  const int32_t *num[3];  // per cell, 1-based forcing point of each component, 0 none
  uint8_t *fc2;           // filled_cell_2 toggles of hn (interp.f:876-916)
};

// --------------------------------------------------------------------
__host__ __device__ __forceinline__ bool computed(int32_t f) { return f & 1; }
__host__ __device__ __forceinline__ bool is_wall(int32_t f) { return f == SLIP_WALL || f == NOSLIP_WALL; }
__host__ __device__ __forceinline__ bool is_wallf(int32_t f) { return is_wall(f) || f == FILLED_CELL_1; }
__host__ __device__ __forceinline__ int imin(int a, int b) { return a < b ? a : b; }
__host__ __device__ __forceinline__ int imax(int a, int b) { return a > b ? a : b; }

__host__ __device__ __forceinline__ long cell_index(const Grid &g, int i, int j, int k) {
  return (long)(i - 1) + (long)(j - 1) * g.nx + (long)(k - 1) * g.nx * g.ny;
}

// Raw neighbour (the reference's myijk): the offset cell clamped into the
// mesh (ghost.f:530-545, valid for 2 <= i <= nx-1 as all computed cells are).
__host__ __device__ __forceinline__ long rawn(const Grid &g, int i, int j, int k, int di, int dj, int dk) {
  int ii = imin(imax(i + di, 1), g.nx);
  int jj = imin(imax(j + dj, 1), g.ny);
  int kk = imin(imax(k + dk, 1), g.nz);
  return cell_index(g, ii, jj, kk);
}

// Indices of one computed cell's stencil.  Reflected neighbours follow
// set_myinds (ghost.f:866-2476): a wall or filled cell maps to the cell
// itself; second/third neighbours fall back to nearer cells at walls and next
// to the boundary.
template <int D>
struct Nbr {
  long c;
  long ipjk, imjk, ijpk, ijmk, ijkp, ijkm;      // raw first
  long e, w, n, s, t, b;                        // reflected first
  long ee, ww, nn, ss, tt, bb;                  // reflected second
};

template <int D>
__host__ __device__ __forceinline__ void first_nbrs(const Grid &g, int i, int j, int k, Nbr<D> &q) {
  const int32_t *fl = g.flag;
  long c = cell_index(g, i, j, k);
  q.c = c;
  q.ipjk = rawn(g, i, j, k, 1, 0, 0);
  q.imjk = rawn(g, i, j, k, -1, 0, 0);
  q.ijkp = rawn(g, i, j, k, 0, 0, 1);
  q.ijkm = rawn(g, i, j, k, 0, 0, -1);
  q.e = is_wallf(fl[q.ipjk]) ? c : q.ipjk;
  q.w = is_wallf(fl[q.imjk]) ? c : q.imjk;
  q.t = is_wallf(fl[q.ijkp]) ? c : q.ijkp;
  q.b = is_wallf(fl[q.ijkm]) ? c : q.ijkm;
  if (D == 3) {
    q.ijpk = rawn(g, i, j, k, 0, 1, 0);
    q.ijmk = rawn(g, i, j, k, 0, -1, 0);
    q.n = is_wallf(fl[q.ijpk]) ? c : q.ijpk;
    q.s = is_wallf(fl[q.ijmk]) ? c : q.ijmk;
  } else {
    q.ijpk = q.ijmk = q.n = q.s = c;
  }
}

// second neighbours (ghost.f:1060-1093 and 1592-1638)
template <int D>
__host__ __device__ __forceinline__ void second_nbrs(const Grid &g, int i, int j, int k, Nbr<D> &q) {
  const int32_t *fl = g.flag;
  long x;
  x = rawn(g, i, j, k, 2, 0, 0);
  q.ee = (i == g.nx - 1) ? q.e : (is_wallf(fl[x]) ? q.ipjk : x);
  x = rawn(g, i, j, k, -2, 0, 0);
  q.ww = (i == 2) ? q.w : (is_wallf(fl[x]) ? q.imjk : x);
  x = rawn(g, i, j, k, 0, 0, 2);
  q.tt = (k == g.nz - 1) ? q.t : (is_wallf(fl[x]) ? q.ijkp : x);
  x = rawn(g, i, j, k, 0, 0, -2);
  q.bb = (k == 2) ? q.b : (is_wallf(fl[x]) ? q.ijkm : x);
  if (D == 3) {
    x = rawn(g, i, j, k, 0, 2, 0);
    q.nn = (j == g.ny - 1) ? q.n : (is_wallf(fl[x]) ? q.n : x);
    x = rawn(g, i, j, k, 0, -2, 0);
    q.ss = (j == 2) ? q.s : (is_wallf(fl[x]) ? q.s : x);
  } else {
    q.nn = q.ss = q.c;
  }
}

// first-level diagonal of set_myinds 2D (ghost.f:984-1058)
__host__ __device__ __forceinline__ long refl_diag(const int32_t *fl, long self, long corner, long a, long b) {
  if (!is_wall(fl[corner])) return corner;
  bool wa = is_wall(fl[a]), wb = is_wall(fl[b]);
  if (wa && !wb) return b;
  if (!wa && wb) return a;
  return self;
}

// second-level diagonal of set_myinds 2D (ghost.f:1095-1153)
__host__ __device__ __forceinline__ long refl_diag2(const int32_t *fl, long base, long corner, long a, long b) {
  if (!is_wallf(fl[corner])) return corner;
  bool wa = is_wall(fl[a]), wb = is_wall(fl[b]);
  if (wa && !wb) return b;
  if (!wa && wb) return a;
  return base;
}

// ----------------------------------------------------------------------
// limiters, limiters.f:14-56 (T = double for the plain form, float for the
// flux arithmetic of the FP32-flux form)
template <typename T>
__device__ __forceinline__ T limiter(int t, T beta, T erre) {
  T l, m, bu;
  // minmod, the usual choice, tested first
  if (t == 2) {
    m = erre < T(1) ? erre : T(1);
    return m > T(0) ? m : T(0);
  }
  switch (t) {
  case 1: l = T(2) * erre / (T(1) + erre); return l > T(0) ? l : T(0);
  case 2: m = erre < T(1) ? erre : T(1); return m > T(0) ? m : T(0);
  case 3: {
    T a = T(2) * erre < T(1) ? T(2) * erre : T(1);
    T b = erre < T(2) ? erre : T(2);
    T r = a > b ? a : b;
    return r > T(0) ? r : T(0);
  }
  case 4:
    bu = (T(1) - beta) + beta * erre;
    m = T(2) * erre;
    if (bu < m) m = bu;
    if (T(2) < m) m = T(2);
    return m > T(0) ? m : T(0);
  case 0: return (T(1) - beta) + beta * erre;
  default: return T(0);
  }
}

// one MUSCL face increment: ratio of upwind to central gradient, limited,
// times the half-cell increment (flux_m.f:165-184 and its analogues)
template <typename T>
__device__ __forceinline__ T muscl_incr(int lim_type, T beta, bool use_lim, T gradc, T gradup,
                                        T graddn, T cs, T incr_p, T incr_m) {
  T erre = T(0), incr;
  if (cs >= T(0)) { if (gradc != T(0)) erre = gradup / gradc; incr = incr_p; }
  else { if (gradc != T(0)) erre = graddn / gradc; incr = incr_m; }
  T lim = use_lim ? limiter<T>(lim_type, beta, erre) : T(0);
  return lim * gradc * incr;
}

// donor-cell flux of a face: 0.5(cs-|cs|)*down + 0.5(cs+|cs|)*up with the
// reference's left-to-right products
template <typename T>
__device__ __forceinline__ T upw(T cs) { return T(0.5) * (cs - fabs(cs)); }
template <typename T>
__device__ __forceinline__ T dnw(T cs) { return T(0.5) * (cs + fabs(cs)); }

// ----------------------------------------------------------------------
// thermodynamics (eosg.f, hcapgs.f)
__device__ __forceinline__ void mole(const Consts &k, const double *yg, double *xg) {
  double mol = 0.0;
  for (int g = 0; g < k.ngas; ++g) mol = mol + yg[g] / k.gmw[k.gas_type[g] - 1];
  double molinv = (fabs(mol) <= 1.0e-10) ? 0.0 : 1.0 / mol;
  for (int g = 0; g < k.ngas; ++g) xg[g] = yg[g] / k.gmw[k.gas_type[g] - 1] * molinv;
}

// mixture molar mass of thermal_eosg
__device__ __forceinline__ double mix_mass(const Consts &k, const double *xg) {
  double mg = 0.0;
  for (int g = 0; g < k.ngas; ++g) mg = mg + xg[g] * k.gmw[k.gas_type[g] - 1];
  return mg;
}

// thermal_eosg with the mixture molar mass given
__device__ __forceinline__ double thermal_eosg_m(double tg, double p, double mg) {
  return p / (8.31432 * tg) * mg;
}

__device__ __forceinline__ double thermal_eosg(const Consts &k, double tg, double p, const double *xg) {
  return thermal_eosg_m(tg, p, mix_mass(k, xg));
}

// heat capacity of one species (J/(kg K)), hcapgs.f:33-85
__device__ __forceinline__ double hcapg1(const Consts &k, int type0, double tg) {
  double t1 = (k.icpc <= 1) ? k.tref : tg;
  double t2 = t1 * t1, t3 = t1 * t2, c;
  switch (type0) {
  case 0: c = (2.811e1 - 3.68e-6 * t1 + 1.746e-5 * t2 - 1.065e-8 * t3); break;
  case 1: c = (3.115e1 - 1.357e-2 * t1 + 2.680e-5 * t2 - 1.168e-8 * t3); break;
  case 2: c = (1.980e1 + 7.344e-2 * t1 - 5.602e-5 * t2 + 1.715e-8 * t3); break;
  case 3: c = (2.714e1 + 9.274e-3 * t1 - 1.381e-5 * t2 + 7.645e-9 * t3); break;
  case 4: c = (3.224e1 + 1.924e-3 * t1 + 1.055e-5 * t2 - 3.596e-9 * t3); break;
  case 5: c = (3.0413e1 - 1.059e-2 * t1 + 2.458e-5 * t2 - 1.135e-8 * t3); break;
  default: c = (2.385e1 + 6.699e-2 * t1 - 4.961e-5 * t2 + 1.328e-8 * t3); break;
  }
  c = c / k.gmw[type0];
  if (k.icpc == 0) {
    if (type0 == 4) c = 1810.0;
    if (type0 == 5) c = 1004.0;
  }
  return c;
}

// caloric_eosg, eosg.f:139-223: Picard iteration T <- h / cp(T)
__device__ __forceinline__ int caloric_eosg(const Consts &k, double sieg, const double *yg,
                                            double &tg, double &cg) {
  int info = 1;
  double cgas = 0.0;
  for (int ii = 1; ii <= 2000; ++ii) {
    double tgnn = tg;
    cgas = 0.0;
    for (int g = 0; g < k.ngas; ++g) cgas = hcapg1(k, k.gas_type[g] - 1, tg) * yg[g] + cgas;
    tg = 0.0 + (sieg - 0.0) / cgas;
    if (fabs((tgnn - tg) / tgnn) <= 1.0e-8) { info = 0; break; }
  }
  cg = cgas;
  if (tg <= 0.0) info = 1;
  return info;
}

// the same with the species count fixed at compile time, mass fractions in
// registers
template <int NG>
__device__ __forceinline__ int caloric_eosg_n(const Consts &k, double sieg, const double (&yg)[NG],
                                              double &tg, double &cg) {
  int info = 1;
  double cgas = 0.0;
  for (int ii = 1; ii <= 2000; ++ii) {
    double tgnn = tg;
    cgas = 0.0;
#pragma unroll
    for (int g = 0; g < NG; ++g) cgas = hcapg1(k, k.gas_type[g] - 1, tg) * yg[g] + cgas;
    tg = 0.0 + (sieg - 0.0) / cgas;
    if (fabs((tgnn - tg) / tgnn) <= 1.0e-8) { info = 0; break; }
  }
  cg = cgas;
  if (tg <= 0.0) info = 1;
  return info;
}

// Lennard-Jones / Wilke mixture viscosity and conductivity, visc.f:37-148
__device__ __forceinline__ void viscon(const Consts &k, const double *xgc, double tg,
                                       double &mu, double &kap) {
  const double mmugs[7] = {3.467, 3.798, 3.941, 2.827, 2.641, 3.711, 4.112};
  const double mmugek[7] = {106.7, 71.4, 195.2, 59.7, 809.1, 78.6, 335.4};
  double mmug[PDAC_MAX_GAS], ckg[PDAC_MAX_GAS];
  double t1 = tg, t2 = t1 * tg, t3 = t2 * tg;
  double tr = tg / 132.5;
  double trm1 = 1.0 / tr, trm2 = trm1 * trm1, trm3 = trm2 * trm1, trm4 = trm3 * trm1;
  double trsrt = sqrt(tr);
  for (int g = 0; g < k.ngas; ++g) {
    int t = k.gas_type[g] - 1;
    double tst = tg / mmugek[t];
    double om = 1.16145 * pow(tst, -0.14874) + 0.52487 * exp(-0.77320 * tst) +
                2.16178 * exp(-2.43787 * tst);
    double m = 1.0e3 * k.gmw[t];
    mmug[g] = 26.69 * sqrt(m * tg) / (mmugs[t] * mmugs[t] * om);
    mmug[g] = mmug[g] * 1.0e-7;
    switch (t) {
    case 0: ckg[g] = -3.273e-4 + 9.966e-5 * t1 - 3.743e-8 * t2 + 9.732e-12 * t3; break;
    case 1: ckg[g] = 3.919e-4 + 9.816e-5 * t1 - 5.067e-8 * t2 + 1.504e-11 * t3; break;
    case 2: ckg[g] = -7.215e-3 + 8.015e-5 * t1 + 5.477e-9 * t2 - 1.053e-11 * t3; break;
    case 3: ckg[g] = 8.099e-3 + 6.689e-4 * t1 - 4.158e-7 * t2 + 1.562e-10 * t3; break;
    case 4: ckg[g] = 7.341e-3 - 1.013e-5 * t1 + 1.801e-7 * t2 - 9.100e-11 * t3; break;
    case 5:
      ckg[g] = 25.9778e-3 * (0.2395 * tr + 6.4977e-3 * trsrt + 1.0 - 1.92615 * trm1 +
                             2.00383 * trm2 - 1.07553 * trm3 + 0.229414 * trm4);
      break;
    default: ckg[g] = -8.086e-3 + 6.344e-5 * t1 - 1.382e-8 * t2 + 2.303e-12 * t3; break;
    }
  }
  double m_ = 0.0, k_ = 0.0;
  for (int g = 0; g < k.ngas; ++g) {
    int tg_ = k.gas_type[g] - 1;
    double sum = 0.0;
    for (int h = 0; h < k.ngas; ++h) {
      int th = k.gas_type[h] - 1;
      double aa = k.gmw[tg_] / k.gmw[th];
      double bb = mmug[g] / mmug[h];
      double cc = 1.0 + sqrt(bb) * k.visw[g][h];
      double phij = cc * cc / sqrt(8.0 * (1.0 + aa));
      sum = sum + xgc[h] * phij;
    }
    double suminv = 1.0 / sum;
    m_ = m_ + xgc[g] * mmug[g] * suminv;
    k_ = k_ + xgc[g] * ckg[g] * suminv;
  }
  mu = m_;
  kap = k_;
}

// gas-particle drag, drag.f:11-157
__device__ __forceinline__ double kdrags(const Consts &k, int is, double du, double dv, double dw,
                                         double ep, double rgp, double rlk, double mug) {
  double vrel = sqrt(du * du + dv * dv + dw * dw);
  double inrl = k.inrl[is], dk = k.dk[is], phis = k.phis[is];
  double drag;
  if (ep < 0.8) {
    double denom = 1.0 / (dk * phis * ep);
    drag = 150.0 * ep * rlk * inrl * mug * denom + 1.75 * rgp * vrel;
    drag = drag * rlk * inrl * denom;
  } else {
    double eps = rlk * inrl;
    double reynum = rgp * dk * phis * vrel / mug;
    double cd;
    if (reynum <= 1000.0) cd = (24.0 / reynum) * (1.0 + FW(0.15) * pow(reynum, FW(0.687)));
    else cd = FW(0.43828814);
    cd = cd < 1.0e4 ? cd : 1.0e4;
    drag = 0.75 * cd * eps / (dk * phis) * rgp * vrel;
    drag = drag * pow(ep, FW(-2.65));
  }
  return drag;
}

__device__ __forceinline__ double sdrag(const Consts &k, int is, double du, double dv, double dw,
                                        double ep, double rgp, double rlk, double mug) {
  double vrel = sqrt(du * du + dv * dv + dw * dw);
  double inrl = k.inrl[is], dk = k.dk[is], phis = k.phis[is];
  double eps = rlk * inrl;
  double reynum = rgp * dk * phis * vrel / mug;
  double a = pow(ep, FW(4.14)), b;
  if (ep <= FW(0.85)) b = FW(0.8) * pow(ep, FW(1.28));
  else b = pow(ep, FW(2.65));
  double vr = 0.5 * (a - FW(0.06) * reynum +
                     sqrt(FW(0.0036) * (reynum * reynum) + FW(0.12) * reynum * (2.0 * b - a) + a * a));
  double q = 0.63 + 4.8 / sqrt(reynum);
  double cd = q * q;
  cd = cd < 1.0e4 ? cd : 1.0e4;
  double drag = 0.75 * cd * eps / (dk * phis) * rgp * vrel;
  return drag / vr;
}

// particle-particle drag (Syamlal), drag.f:248-298; k, kk 1-based phases >= 2
__device__ __forceinline__ double ppdrag(const Consts &c, const double *rlkc, int k, int kk) {
  int k1, k2;
  if (c.dk[k - 2] >= c.dk[kk - 2]) { k1 = k - 2; k2 = kk - 2; }
  else { k1 = kk - 2; k2 = k - 2; }
  double ep1 = rlkc[k1] * c.inrl[k1];
  double ep2 = rlkc[k2] * c.inrl[k2];
  double epsum = ep1 + ep2;
  if (ep1 > 0.0 && ep2 > 0.0) {
    double xbar = ep1 / epsum, epkl;
    if (xbar <= c.philim[k - 2][kk - 2]) epkl = c.epsl[k - 2][kk - 2] * xbar + c.phi[k2];
    else epkl = c.epsu[k - 2][kk - 2] * (1.0 - xbar) + c.phi[k1];
    double espo = 1.0 / 3.0;
    double m1 = pow(epkl, espo), m2 = pow(epsum, espo);
    double effe = (3.0 * m1 + m2) / (2.0 * (m1 - m2));
    return 1.0 * (1.0 + 1.0) * rlkc[k - 2] * rlkc[kk - 2] * c.dkf[k - 2][kk - 2] * effe;
  }
  return 0.0;
}

// Gunn heat transfer, hvs.f:9-83
__device__ __forceinline__ double hvs(const Consts &k, double rlk, double rog, double du, double dv,
                                      double dw, double mug, double kapg, double cg, int is) {
  if (rlk <= 0.0) return 0.0;
  double dk = k.dk[is];
  double eps = rlk * k.inrl[is];
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

// Gauss-Jordan solve with the two-phase collapse, matrix.f:353-397
template <int NP>
__device__ __forceinline__ void gauss_jordan(const double *plim, double (&a)[NP][NP], double (&b)[NP]) {
#pragma unroll
  for (int l = 1; l < NP; ++l) {
    if (fabs(a[l][l]) < plim[l - 1]) {
#pragma unroll
      for (int ll = 0; ll < NP; ++ll) {
        a[ll][ll] = a[ll][ll] + a[l][ll];
        a[l][ll] = 0.0;
        a[ll][l] = 0.0;
      }
      b[l] = 0.0;
    }
  }
#pragma unroll
  for (int l = 0; l < NP; ++l) {
    if (a[l][l] != 0.0) {
      double div = 1.0 / a[l][l];
#pragma unroll
      for (int lj = l + 1; lj < NP; ++lj) a[l][lj] = a[l][lj] * div;
      b[l] = b[l] * div;
      a[l][l] = 0.0;
#pragma unroll
      for (int li = 0; li < NP; ++li) {
        double amul = a[li][l];
#pragma unroll
        for (int lj = l + 1; lj < NP; ++lj) a[li][lj] = a[li][lj] - amul * a[l][lj];
        b[li] = b[li] - amul * b[l];
      }
    }
  }
}

}  // namespace pdac
