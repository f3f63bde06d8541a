// pdac_kernels.hip -- gfx950 kernels of the PDAC semi-implicit per-step hot
// path and the C-ABI of include/pdacgpu.h.
//
// One thread handles one cell (and, where the work is per phase, one
// phase of one cell).  Writes are race-free by construction: a face flux is
// produced by exactly one cell (the reference's "last writer", see each
// kernel), and the pressure-velocity sweep runs in two colours so that a
// cell's update writes only its own scalars and its own faces.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <string>
#include <utility>
#include <vector>

#include "pdac_dev.hpp"
#include "pdac_flux.hpp"

using namespace pdac;

namespace {

thread_local std::string g_last_error;

int set_error(const char *where, hipError_t e) {
  g_last_error = std::string(where) + ": " + hipGetErrorString(e);
  return -(int)e;
}

#define HIP_RET(call)                                        \
  do {                                                       \
    hipError_t e_ = (call);                                  \
    if (e_ != hipSuccess) return set_error(#call, e_);       \
  } while (0)

constexpr int kBlock = 256;

constexpr unsigned kXcd = 8;  // MI355X: 8 accelerator dies, each with its own L2

// Work-groups are dealt to the XCDs round-robin.  Grids are rounded up to a
// multiple of 8 blocks and block b works on the contiguous range
// (b % 8) * (nblocks / 8) + b / 8, so each XCD sweeps one contiguous slab of
// the mesh and a cell's k+-1 neighbours sit in the same L2.
inline unsigned nblocks(long n) {
  unsigned b = (unsigned)((n + kBlock - 1) / kBlock);
  return (b + kXcd - 1) / kXcd * kXcd;
}

// decode a linear cell index into 1-based (i,j,k)
__device__ __forceinline__ void decode(const Grid &g, long c, int &i, int &j, int &k) {
  if (g.n <= 0x7fffffffL) {
    // 32-bit quotients (a mesh of fewer than 2^31 cells)
    const unsigned cu = (unsigned)c, nx = (unsigned)g.nx, nxy = nx * (unsigned)g.ny;
    const unsigned kq = cu / nxy, r = cu - kq * nxy, jq = r / nx;
    k = (int)kq + 1;
    j = (int)jq + 1;
    i = (int)(r - jq * nx) + 1;
    return;
  }
  long nxy = (long)g.nx * g.ny;
  k = (int)(c / nxy) + 1;
  long r = c - (long)(k - 1) * nxy;
  j = (int)(r / g.nx) + 1;
  i = (int)(r - (long)(j - 1) * g.nx) + 1;
}

__device__ __forceinline__ long cell_id() {
  const unsigned per = gridDim.x / kXcd;
  const unsigned b = blockIdx.x;
  return (long)((b % kXcd) * per + b / kXcd) * blockDim.x + threadIdx.x;
}

// As cell_id for a grid of nblocks(n) * parts blocks that runs 'parts'
// threads per cell: the parts of one block of cells are consecutive blocks
// of the same XCD, so they run close in time and share its L2.
__device__ __forceinline__ long cell_part_id(int parts, int &part) {
  const unsigned per = gridDim.x / (kXcd * parts);
  const unsigned b = blockIdx.x;
  const unsigned t = b / kXcd;
  part = (int)(t % parts);
  return (long)((b % kXcd) * per + t / parts) * blockDim.x + threadIdx.x;
}

// Append 'val' to a device work list from every lane where 'pred' holds: one
// atomic per wavefront, lanes placed in lane order.  Every lane of the
// wavefront must call it.
__device__ __forceinline__ void wl_append(int32_t *cnt, int32_t *list, bool pred, int32_t val) {
  const unsigned long long m = __ballot(pred);
  if (!m) return;
  const int lane = threadIdx.x & 63;
  const int leader = __ffsll((long long)m) - 1;
  int base = 0;
  if (lane == leader) base = atomicAdd(cnt, (int32_t)__popcll(m));
  base = __shfl(base, leader);
  if (pred) list[base + __popcll(m & ((1ull << lane) - 1ull))] = val;
}

// device counters: sweep state of iter (done flag, sweep numbers of the two
// colour launches, the sweep that converged, the first colour phase of the
// call), EOS failures, three pairs of unconverged-cell counts per sweep
// (C_WLP), and for the persistent iteration kernel an arrival counter for its
// grid barriers and an error flag
enum Ctr { C_DONE = 0, C_NIT1 = 1, C_NIT = 2, C_PHB = 3, C_EOS = 4, C_NITF = 5, C_WLP = 8, C_BAR = 14,
           C_ERR = 15, C_N = 16 };

// Whether the iteration of which the host has launched 'gate' sweeps has
// ended, by the test the host applies after each batch of sweeps: a sweep
// found its predecessor converged, the last launched sweep left no cell
// unconverged, or maxout sweeps ran.  The step's kernels after the
// iteration are launched behind each batch with this gate, before the host
// has read the batch's counters: they do nothing unless the batch ended the
// iteration.  gate < 0: no gate.
__device__ __forceinline__ bool iter_ended(const int32_t *ctr, int gate, int maxout) {
  if (gate < 0) return true;
  if (ctr[C_DONE]) return true;
  const int32_t *last = ctr + C_WLP + 2 * (gate % 3);
  if (last[0] + last[1] == 0) return true;
  return gate >= maxout;
}

}  // namespace

// ======================================================================
// boundary, bdry.f:22-726: each computed fluid cell fills the boundary
// neighbours of its six faces in the reference's order (E, W, N, S, T, B).
// Every boundary cell is face-adjacent to at most one computed cell in each
// direction, so no two threads write the same value; each thread reads only
// its own cell and interior cells no thread writes.
template <int D, int NS>
__device__ void inoutflow_d(const Grid &g, const Consts &K, Fields &F, double ugv, const double *usv,
                            long n2, long n1, int k, double gnorm) {
  long n = g.n;
  int sgn = (n2 - n1) >= 0 ? 1 : -1;
  double *tg = F.f[PDAC_TG], *sieg = F.f[PDAC_SIEG], *ygc = F.f[PDAC_YGC], *xgc = F.f[PDAC_XGC];
  double *rog = F.f[PDAC_ROG], *rlk = F.f[PDAC_RLK], *ts = F.f[PDAC_TS], *sies = F.f[PDAC_SIES];
  double *ep = F.f[PDAC_EP], *rgp = F.f[PDAC_RGP], *p = F.f[PDAC_P];
  if (ugv * sgn > 0.0) {
    tg[n2] = tg[n1];
    sieg[n2] = sieg[n1];
    for (int g_ = 0; g_ < K.ngas; ++g_) {
      ygc[g_ * n + n2] = ygc[g_ * n + n1];
      xgc[g_ * n + n2] = xgc[g_ * n + n1];
    }
    rog[n2] = rog[n1];
  } else if (ugv * sgn <= 0.0) {
    tg[n2] = g.t_atm[k];
    double yl[PDAC_MAX_GAS], xl[PDAC_MAX_GAS];
    for (int g_ = 0; g_ < K.ngas; ++g_) {
      yl[g_] = K.atm_ygc[K.gas_type[g_] - 1];
      ygc[g_ * n + n2] = yl[g_];
    }
    mole(K, yl, xl);
    for (int g_ = 0; g_ < K.ngas; ++g_) xgc[g_ * n + n2] = xl[g_];
    rog[n2] = thermal_eosg(K, g.t_atm[k], g.p_atm[k], xl);
  }
  double rls = 0.0;
#pragma unroll
  for (int is = 0; is < NS; ++is) {
    if (usv[is] * sgn > 0.0) {
      rlk[is * n + n2] = rlk[is * n + n1];
      ts[is * n + n2] = ts[is * n + n1];
      sies[is * n + n2] = sies[is * n + n1];
    } else if (usv[is] * sgn <= 0.0) {
      rlk[is * n + n2] = K.rl[is] * (1.0 - K.void_fraction) / (double)NS;
      ts[is * n + n2] = g.t_atm[k];
    }
    rls = rls + rlk[is * n + n2] * K.inrl[is];
  }
  ep[n2] = 1.0 - rls;
  rgp[n2] = rog[n2] * ep[n2];
  double rhomix;
  if (K.pmodel == 1) {
    double sm = rlk[n2];
    for (int is = 1; is < NS; ++is) sm = sm + rlk[is * n + n2];
    rhomix = rgp[n2] + sm;
  } else {
    rhomix = rgp[n2];
  }
  p[n2] = p[n1] + rhomix * gnorm;
}

template <int NS>
__device__ void zerograd_d(const Grid &g, const Consts &K, Fields &F, long n2, long n1) {
  long n = g.n;
  double *rlk = F.f[PDAC_RLK];
  F.f[PDAC_TG][n2] = F.f[PDAC_TG][n1];
  F.f[PDAC_SIEG][n2] = F.f[PDAC_SIEG][n1];
  for (int g_ = 0; g_ < K.ngas; ++g_) {
    F.f[PDAC_YGC][g_ * n + n2] = F.f[PDAC_YGC][g_ * n + n1];
    F.f[PDAC_XGC][g_ * n + n2] = F.f[PDAC_XGC][g_ * n + n1];
  }
  F.f[PDAC_ROG][n2] = F.f[PDAC_ROG][n1];
  double rls = 0.0;
  for (int is = 0; is < NS; ++is) {
    rlk[is * n + n2] = rlk[is * n + n1];
    F.f[PDAC_TS][is * n + n2] = F.f[PDAC_TS][is * n + n1];
    F.f[PDAC_SIES][is * n + n2] = F.f[PDAC_SIES][is * n + n1];
    rls = rls + rlk[is * n + n2] * K.inrl[is];
  }
  F.f[PDAC_EP][n2] = 1.0 - rls;
  F.f[PDAC_RGP][n2] = F.f[PDAC_ROG][n2] * F.f[PDAC_EP][n2];
  F.f[PDAC_P][n2] = F.f[PDAC_P][n1];
}

__device__ __forceinline__ double extrapolate(double f2, double f1, double d12, double dx, int order) {
  double gradient = (f2 - f1) / d12 * (double)order;
  return f2 + gradient * dx;
}

template <int D, int NS>
__device__ __forceinline__ void boundary_cell(const Grid &g, const Consts &K, Fields &F, long c) {
  const int32_t *fl = g.flag;
  int f = fl[c];
  if (!(f == FLUID || f == DOME_CELL || f == BL_CELL)) return;
  int i, j, k;
  decode(g, c, i, j, k);
  long n = g.n;
  double *ug = F.f[PDAC_UG], *vg = F.f[PDAC_VG], *wg = F.f[PDAC_WG];
  double *us = F.f[PDAC_US], *vs = F.f[PDAC_VS], *ws = F.f[PDAC_WS];
  long ipjk = rawn(g, i, j, k, 1, 0, 0), imjk = rawn(g, i, j, k, -1, 0, 0);
  long ijkp = rawn(g, i, j, k, 0, 0, 1), ijkm = rawn(g, i, j, k, 0, 0, -1);
  // quick exit for interior cells: every face neighbour is computed
  bool inner = computed(fl[ipjk]) && computed(fl[imjk]) && computed(fl[ijkp]) && computed(fl[ijkm]);
  long ijpk = c, ijmk = c;
  if (D == 3) {
    ijpk = rawn(g, i, j, k, 0, 1, 0);
    ijmk = rawn(g, i, j, k, 0, -1, 0);
    inner = inner && computed(fl[ijpk]) && computed(fl[ijmk]);
  }
  if (inner) return;
  const long n1 = c;
  double usv[NS];
#define SETS(a, dst, src) _Pragma("unroll") for (int is_ = 0; is_ < NS; ++is_) a[is_ * n + (dst)] = (src)
  // ---- east
  {
    long n2 = ipjk, n0 = imjk;
    switch (fl[n2]) {
    case SLIP_WALL:
      ug[n1] = 0.0; SETS(us, n1, 0.0); ug[n2] = 0.0; SETS(us, n2, 0.0);
      if (D == 3) { vg[n2] = vg[n1]; SETS(vs, n2, vs[is_ * n + n1]); }
      wg[n2] = wg[n1]; SETS(ws, n2, ws[is_ * n + n1]);
      break;
    case NOSLIP_WALL:
      ug[n1] = 0.0; SETS(us, n1, 0.0); ug[n2] = 0.0; SETS(us, n2, 0.0);
      if (!computed(fl[rawn(g, i, j, k, 1, 0, 1)])) { wg[n2] = -wg[n1]; SETS(ws, n2, -ws[is_ * n + n1]); }
      if (D == 3 && !computed(fl[rawn(g, i, j, k, 1, 1, 0)])) { vg[n2] = -vg[n1]; SETS(vs, n2, -vs[is_ * n + n1]); }
      break;
    case FREE_IO: case ZERO_GRAD: {
      ug[n1] = ug[n0]; SETS(us, n1, us[is_ * n + n0]);
      ug[n2] = ug[n1]; SETS(us, n2, us[is_ * n + n1]);
      if (D == 3) { vg[n2] = vg[n1]; SETS(vs, n2, vs[is_ * n + n1]); }
      wg[n2] = wg[n1]; SETS(ws, n2, ws[is_ * n + n1]);
      double gnorm = K.gravx * 0.5 * (g.dx[i + 1] + g.dx[i]);
      if (fl[n2] == ZERO_GRAD) gnorm = 0.0;
      for (int is = 0; is < NS; ++is) usv[is] = us[is * n + n1];
      inoutflow_d<D, NS>(g, K, F, ug[n1], usv, n2, n1, k, gnorm);
      break;
    }
    default: break;
    }
  }
  // ---- west
  {
    long n2 = imjk;
    switch (fl[n2]) {
    case SLIP_WALL:
      ug[n2] = 0.0; SETS(us, n2, 0.0);
      if (D == 3) { vg[n2] = vg[n1]; SETS(vs, n2, vs[is_ * n + n1]); }
      wg[n2] = wg[n1]; SETS(ws, n2, ws[is_ * n + n1]);
      break;
    case NOSLIP_WALL: {
      ug[n2] = 0.0; SETS(us, n2, 0.0);
      bool ext = !computed(fl[rawn(g, i, j, k, -1, 0, 1)]);
      if (ext) { wg[n2] = -wg[n1]; SETS(ws, n2, -ws[is_ * n + n1]); }
      if (D == 3 && ext) { vg[n2] = -vg[n1]; SETS(vs, n2, -vs[is_ * n + n1]); }
      break;
    }
    case FREE_IO: case ZERO_GRAD: {
      ug[n2] = ug[n1]; SETS(us, n2, us[is_ * n + n1]);
      if (D == 3) { vg[n2] = vg[n1]; SETS(vs, n2, vs[is_ * n + n1]); }
      wg[n2] = wg[n1]; SETS(ws, n2, ws[is_ * n + n1]);
      double gnorm = K.gravx * 0.5 * (g.dx[i - 1] + g.dx[i]);
      if (fl[n2] == ZERO_GRAD) gnorm = 0.0;
      for (int is = 0; is < NS; ++is) usv[is] = us[is * n + n2];
      inoutflow_d<D, NS>(g, K, F, ug[n2], usv, n2, n1, k, gnorm);
      break;
    }
    default: break;
    }
  }
  if (D == 3) {
    // ---- north
    {
      long n2 = ijpk, n0 = ijmk;
      switch (fl[n2]) {
      case SLIP_WALL:
        vg[n1] = 0.0; SETS(vs, n1, 0.0); vg[n2] = 0.0; SETS(vs, n2, 0.0);
        ug[n2] = ug[n1]; SETS(us, n2, us[is_ * n + n1]);
        wg[n2] = wg[n1]; SETS(ws, n2, ws[is_ * n + n1]);
        break;
      case NOSLIP_WALL:
        vg[n1] = 0.0; SETS(vs, n1, 0.0); vg[n2] = 0.0; SETS(vs, n2, 0.0);
        if (!computed(fl[rawn(g, i, j, k, 1, 1, 0)])) { ug[n2] = -ug[n1]; SETS(us, n2, -us[is_ * n + n1]); }
        if (!computed(fl[rawn(g, i, j, k, 0, 1, 1)])) { wg[n2] = -wg[n1]; SETS(ws, n2, -ws[is_ * n + n1]); }
        break;
      case FREE_IO: case ZERO_GRAD: {
        vg[n1] = vg[n0]; SETS(vs, n1, vs[is_ * n + n0]);
        vg[n2] = vg[n1]; SETS(vs, n2, vs[is_ * n + n1]);
        ug[n2] = ug[n1]; SETS(us, n2, us[is_ * n + n1]);
        wg[n2] = wg[n1]; SETS(ws, n2, ws[is_ * n + n1]);
        double gnorm = K.gravy * 0.5 * (g.dy[j] + g.dy[j + 1]);
        if (fl[n2] == ZERO_GRAD) gnorm = 0.0;
        for (int is = 0; is < NS; ++is) usv[is] = vs[is * n + n1];
        inoutflow_d<D, NS>(g, K, F, vg[n1], usv, n2, n1, k, gnorm);
        break;
      }
      default: break;
      }
    }
    // ---- south
    {
      long n2 = ijmk;
      switch (fl[n2]) {
      case SLIP_WALL:
        vg[n2] = 0.0; SETS(vs, n2, 0.0);
        ug[n2] = ug[n1]; SETS(us, n2, us[is_ * n + n1]);
        wg[n2] = wg[n1]; SETS(ws, n2, ws[is_ * n + n1]);
        break;
      case NOSLIP_WALL:
        vg[n2] = 0.0; SETS(vs, n2, 0.0);
        if (!computed(fl[rawn(g, i, j, k, 1, -1, 0)])) { ug[n2] = -ug[n1]; SETS(us, n2, -us[is_ * n + n1]); }
        if (!computed(fl[rawn(g, i, j, k, 0, -1, 1)])) { wg[n2] = -wg[n1]; SETS(ws, n2, -ws[is_ * n + n1]); }
        break;
      case FREE_IO: case ZERO_GRAD: {
        vg[n2] = vg[n1]; SETS(vs, n2, vs[is_ * n + n1]);
        ug[n2] = ug[n1]; SETS(us, n2, us[is_ * n + n1]);
        wg[n2] = wg[n1]; SETS(ws, n2, ws[is_ * n + n1]);
        double gnorm = K.gravy * 0.5 * (g.dy[j - 1] + g.dy[j]);
        if (fl[n2] == ZERO_GRAD) gnorm = 0.0;
        for (int is = 0; is < NS; ++is) usv[is] = vs[is * n + n2];
        inoutflow_d<D, NS>(g, K, F, vg[n2], usv, n2, n1, k, gnorm);
        break;
      }
      default: break;
      }
    }
  }
  // ---- top
  {
    long n2 = ijkp;
    switch (fl[n2]) {
    case SLIP_WALL:
      wg[n1] = 0.0; SETS(ws, n1, 0.0);
      ug[n2] = ug[n1]; SETS(us, n2, us[is_ * n + n1]);
      if (D == 3) { vg[n2] = vg[n1]; SETS(vs, n2, vs[is_ * n + n1]); }
      break;
    case NOSLIP_WALL:
      wg[n1] = 0.0; SETS(ws, n1, 0.0);
      if (!computed(fl[rawn(g, i, j, k, 1, 0, 1)])) { ug[n2] = -ug[n1]; SETS(us, n2, -us[is_ * n + n1]); }
      if (D == 3 && !computed(fl[rawn(g, i, j, k, 0, 1, 1)])) { vg[n2] = -vg[n1]; SETS(vs, n2, -vs[is_ * n + n1]); }
      break;
    case FREE_IO: case ZERO_GRAD: {
      double d0 = g.dz[k - 1], d1 = g.dz[k], d2 = g.dz[k + 1];
      long ijkmm = rawn(g, i, j, k, 0, 0, -2);
      ug[n2] = ug[n1]; SETS(us, n2, us[is_ * n + n1]);
      if (D == 3) { vg[n2] = vg[n1]; SETS(vs, n2, vs[is_ * n + n1]); }
      wg[c] = extrapolate(wg[ijkm], wg[ijkmm], d0, d1, K.ord);
      for (int is = 0; is < NS; ++is)
        ws[is * n + c] = extrapolate(ws[is * n + ijkm], ws[is * n + ijkmm], d0, d1, K.ord);
      wg[ijkp] = extrapolate(wg[c], wg[ijkm], d1, d2, K.ord);
      for (int is = 0; is < NS; ++is)
        ws[is * n + ijkp] = extrapolate(ws[is * n + c], ws[is * n + ijkm], d1, d2, K.ord);
      if (fl[n2] == ZERO_GRAD) {
        zerograd_d<NS>(g, K, F, n2, n1);
      } else {
        double gnorm = K.gravz * 0.5 * (g.dz[k] + g.dz[k + 1]);
        for (int is = 0; is < NS; ++is) usv[is] = ws[is * n + n1];
        inoutflow_d<D, NS>(g, K, F, wg[n1], usv, n2, n1, k + 1, gnorm);
      }
      break;
    }
    default: break;
    }
  }
  // ---- bottom
  {
    long n2 = ijkm;
    switch (fl[n2]) {
    case SLIP_WALL:
      wg[n2] = 0.0; SETS(ws, n2, 0.0);
      if (fl[rawn(g, i, j, k, 1, 0, -1)] == SLIP_WALL) { ug[n2] = ug[n1]; SETS(us, n2, us[is_ * n + n1]); }
      if (D == 3 && fl[rawn(g, i, j, k, 0, 1, -1)] == SLIP_WALL) { vg[n2] = vg[n1]; SETS(vs, n2, vs[is_ * n + n1]); }
      break;
    case NOSLIP_WALL:
      wg[n2] = 0.0; SETS(ws, n2, 0.0);
      if (fl[rawn(g, i, j, k, 1, 0, -1)] == NOSLIP_WALL) { ug[n2] = -ug[n1]; SETS(us, n2, -us[is_ * n + n1]); }
      if (D == 3 && fl[rawn(g, i, j, k, 0, 1, -1)] == NOSLIP_WALL) { vg[n2] = -vg[n1]; SETS(vs, n2, -vs[is_ * n + n1]); }
      break;
    case FREE_IO: case ZERO_GRAD: {
      wg[n2] = wg[n1]; SETS(ws, n2, ws[is_ * n + n1]);
      ug[n2] = ug[n1]; SETS(us, n2, us[is_ * n + n1]);
      if (D == 3) { vg[n2] = vg[n1]; SETS(vs, n2, vs[is_ * n + n1]); }
      double gnorm = K.gravz * 0.5 * (g.dz[k - 1] + g.dz[k]);
      if (fl[n2] == ZERO_GRAD) gnorm = 0.0;
      for (int is = 0; is < NS; ++is) usv[is] = ws[is * n + n2];
      inoutflow_d<D, NS>(g, K, F, wg[n2], usv, n2, n1, k - 1, gnorm);
      break;
    }
    default: break;
    }
  }
#undef SETS
}

template <int D, int NS>
__global__ void __launch_bounds__(kBlock) k_boundary(Grid g, Consts K, Fields F) {
  const long c = cell_id();
  if (c < g.n) boundary_cell<D, NS>(g, K, F, c);
}

// hn, interp.f:832-918: homogeneous Neumann values of a filled cell from
// its forcing neighbour (m the forcing cell, nn its neighbour along the
// component); a filled_cell_2 sums halves of its two forcing points
template <int NS>
__device__ __forceinline__ void hn_d(const Grid &g, const Consts &K, Fields &F, const Ib &ib, long m, long nn) {
  const long n = g.n;
  const int32_t fm = g.flag[m], fn = g.flag[nn];
  auto each = [&](auto &&op) {
    for (int id : {PDAC_P, PDAC_EP, PDAC_RGP, PDAC_TG, PDAC_SIEG}) op(F.f[id]);
#pragma unroll
    for (int is = 0; is < NS; ++is) {
      op(F.f[PDAC_RLK] + (long)is * n);
      op(F.f[PDAC_TS] + (long)is * n);
      op(F.f[PDAC_SIES] + (long)is * n);
    }
    for (int g_ = 0; g_ < K.ngas; ++g_) op(F.f[PDAC_YGC] + (long)g_ * n);
  };
  if (fm == FILLED_CELL_1) each([&](double *a) { a[m] = a[nn]; });
  else if (fn == FILLED_CELL_1) each([&](double *a) { a[nn] = a[m]; });
  if (fm == FILLED_CELL_2) {
    if (!ib.fc2[m]) each([&](double *a) { a[m] = 0.5 * a[nn]; });
    else each([&](double *a) { a[m] = a[m] + 0.5 * a[nn]; });
    ib.fc2[m] = ib.fc2[m] ? 0 : 1;
  } else if (fn == FILLED_CELL_2) {
    if (!ib.fc2[nn]) each([&](double *a) { a[nn] = 0.5 * a[m]; });
    else each([&](double *a) { a[nn] = a[nn] + 0.5 * a[m]; });
    ib.fc2[nn] = ib.fc2[nn] ? 0 : 1;
  }
}

// the immersed-boundary part of boundary (bdry.f:103-203): pseudo-velocities
// of the forcing points of an immb or filled cell (velint3d, interp.f:687-781)
template <int D, int NS>
__device__ __forceinline__ void ib_boundary_cell(const Grid &g, const Consts &K, Fields &F, const Ib &ib, long c) {
  const long n = g.n;
  const long nxt[3] = {c + 1, c + (long)g.nx, c + (long)g.nx * g.ny};  // ipjk, ijpk, ijkp
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    if (D == 2 && a == 1) continue;
    const int pt = ib.num[a][c] - 1;
    if (pt < 0) continue;
    const IbAxis &A = ib.ax[a];
    double *vg = F.f[PDAC_UG + a], *vs = F.f[PDAC_US + a];
    const long q = A.q[pt], qq = A.qq[pt];
    const int mode = A.mode[pt];
    const double hza = A.hza[pt];
    auto one = [&](const double *v) {
      if (mode == 0) return hza * v[q];
      if (mode == 1) return -hza * v[q];
      const double v2 = -(A.zbh[pt] * v[q] + A.hza2[pt] * v[qq]) / A.zbza[pt];
      const double v1 = -hza * v[q];
      return fabs(v1) < fabs(v2) ? v1 : v2;
    };
    double vel[NS + 1];
    vel[0] = one(vg);
#pragma unroll
    for (int is = 0; is < NS; ++is) vel[1 + is] = one(vs + (long)is * n);
#pragma unroll
    for (int ph = 0; ph <= NS; ++ph) A.vel[(long)ph * A.nfp + pt] = vel[ph];
    if (A.ext[pt]) hn_d<NS>(g, K, F, ib, c, nxt[a]);
    vg[c] = vel[0];
#pragma unroll
    for (int is = 0; is < NS; ++is) vs[(long)is * n + c] = vel[1 + is];
  }
}

// boundary over a list of cells (one level of the schedule built by the
// host): cells of one level read and write nothing another cell of the
// level writes, and the levels run in order, so the result is the serial
// loop's (bdry.f:77)
template <int D, int NS>
__global__ void __launch_bounds__(kBlock) k_boundary_list(Grid g, Consts K, Fields F, Ib ib, const int32_t *cells,
                                                          int count) {
  const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (t >= count) return;
  const long c = cells[t];
  const int32_t f = g.flag[c];
  if (f == IMMB_CELL || f == FILLED_CELL_1 || f == FILLED_CELL_2) ib_boundary_cell<D, NS>(g, K, F, ib, c);
  else boundary_cell<D, NS>(g, K, F, c);
}

// ======================================================================
// fieldn, tilde.f:125-274: step-start copies, face momenta, viscosity
template <int D, int NS>
__global__ void __launch_bounds__(kBlock) k_fieldn(Grid g, Consts K, Fields F) {
  long c = cell_id();
  if (c >= g.n) return;
  long n = g.n;
  const double *rgp = F.f[PDAC_RGP], *rlk = F.f[PDAC_RLK];
  if (computed(g.flag[c])) {
    int i, j, k;
    decode(g, c, i, j, k);
    Nbr<D> q;
    first_nbrs<D>(g, i, j, k, q);
    const double *dx = g.dx, *dy = g.dy, *dz = g.dz;
    double indxp = 1.0 / (dx[i] + dx[i + 1]), indzp = 1.0 / (dz[k] + dz[k + 1]);
    double rgp_e = (dx[i + 1] * rgp[c] + dx[i] * rgp[q.e]) * indxp;
    double rgp_t = (dz[k + 1] * rgp[c] + dz[k] * rgp[q.t]) * indzp;
    F.f[PDAC_RUGN][c] = rgp_e * F.f[PDAC_UG][c];
    F.f[PDAC_RWGN][c] = rgp_t * F.f[PDAC_WG][c];
    double indyp = 0.0;
    if (D == 3) {
      indyp = 1.0 / (dy[j] + dy[j + 1]);
      double rgp_n = (dy[j + 1] * rgp[c] + dy[j] * rgp[q.n]) * indyp;
      F.f[PDAC_RVGN][c] = rgp_n * F.f[PDAC_VG][c];
    }
#pragma unroll
    for (int is = 0; is < NS; ++is) {
      const double *rl = rlk + is * n;
      double rlk_e = (rl[c] * dx[i + 1] + rl[q.e] * dx[i]) * indxp;
      double rlk_t = (rl[c] * dz[k + 1] + rl[q.t] * dz[k]) * indzp;
      F.f[PDAC_RUSN][is * n + c] = rlk_e * F.f[PDAC_US][is * n + c];
      F.f[PDAC_RWSN][is * n + c] = rlk_t * F.f[PDAC_WS][is * n + c];
      if (D == 3) {
        double rlk_n = (rl[c] * dy[j + 1] + rl[q.n] * dy[j]) * indyp;
        F.f[PDAC_RVSN][is * n + c] = rlk_n * F.f[PDAC_VS][is * n + c];
      }
    }
  }
  F.f[PDAC_PN][c] = F.f[PDAC_P][c];
  F.f[PDAC_EPN][c] = F.f[PDAC_EP][c];
  double rgpn = rgp[c];
  F.f[PDAC_RGPN][c] = rgpn;
  F.f[PDAC_SIEGN][c] = F.f[PDAC_SIEG][c];
  double xl[PDAC_MAX_GAS];
  for (int g_ = 0; g_ < K.ngas; ++g_) {
    F.f[PDAC_RGPGCN][g_ * n + c] = rgpn * F.f[PDAC_YGC][g_ * n + c];
    xl[g_] = F.f[PDAC_XGC][g_ * n + c];
  }
#pragma unroll
  for (int is = 0; is < NS; ++is) {
    F.f[PDAC_RLKN][is * n + c] = rlk[is * n + c];
    F.f[PDAC_SIESN][is * n + c] = F.f[PDAC_SIES][is * n + c];
  }
  double mu, kap;
  viscon(K, xl, F.f[PDAC_TG][c], mu, kap);
  F.f[PDAC_MUG][c] = mu;
  F.f[PDAC_KAPG][c] = kap;
}

__global__ void __launch_bounds__(kBlock) k_fill(double *a, long n, double v) {
  long c = cell_id();
  if (c < n) a[c] = v;
}

__global__ void __launch_bounds__(kBlock) k_copy(double *dst, const double *src, long n) {
  long c = cell_id();
  if (c < n) dst[c] = src[c];
}

// ======================================================================
// tilde (part 1), compute_all_fluxes_tilde, tilde.f:850-1326: momentum
// fluxes of one phase (blockIdx.y) of each computed cell.  E/N/T faces are
// owned by the cell; a W/S/B face is written only when that neighbour is
// not computed (then no other cell writes it).
// 1 for a computed cell whose full momentum-flux stencil (full_idx) lies
// inside the mesh and needs no wall reflection and no mesh-edge special
// case, else 0; 'g' carries no mask, so full_idx takes its general path here
template <int D>
__global__ void __launch_bounds__(kBlock) k_regular(Grid g, uint8_t *reg) {
  const long c = cell_id();
  if (c >= g.n) return;
  uint8_t r = 0;
  if (computed(g.flag[c])) {
    int i, j, k;
    decode(g, c, i, j, k);
    FullIdx<D> x;
    full_idx<D>(g, i, j, k, x);
    // offsets -1..+3 along every axis stay inside the mesh, so each raw
    // index is the cell index plus a fixed offset
    bool same = i >= 2 && i + 3 <= g.nx && k >= 2 && k + 3 <= g.nz && (D == 2 || (j >= 2 && j + 3 <= g.ny));
    same = same && x.Re == x.e && x.Rw == x.w && x.Rn == x.n && x.Rs == x.s && x.Rt == x.t && x.Rb == x.b;
    same = same && x.Ree == x.ee && x.Rnn == x.nn && x.Rtt == x.tt;
    same = same && x.Ren == x.en && x.Ret == x.et && x.Res == x.es && x.Reb == x.eb && x.Rnt == x.nt;
    same = same && x.Rwn == x.wn && x.Rnb == x.nb && x.Rwt == x.wt && x.Rst == x.st;
    same = same && x.Reee == x.eee && x.Rnnn == x.nnn && x.Rttt == x.ttt && x.Reen == x.een;
    same = same && x.Reet == x.eet && x.Renn == x.enn && x.Rnnt == x.nnt && x.Rett == x.ett;
    same = same && x.Rntt == x.ntt;
    r = same ? 1 : 0;
  }
  reg[c] = r;
}

template <int D, typename FT>
__device__ __forceinline__ void momflux_cell(const Grid &g, const Consts &K, const Fields &F, long c,
                                             int fc, int ph, const double *uu, const double *vv,
                                             const double *ww, const double *dd) {
  const long n = g.n;
  int i, j, k;
  decode(g, c, i, j, k);
  FullIdx<D> x;
  full_idx<D>(g, i, j, k, x);
  Stc<FT> u, v, w, dens, sx, sy, sz;
  gather_rnb<D>(x, uu, u);
  gather_rnb<D>(x, ww, w);
  if (D == 3) gather_rnb<D>(x, vv, v);
  if (F.ibm && (fc == BL_CELL || fc == IMMB_CELL)) {
    // velocities of faces inside the topography masked (tilde.f:1072-1118)
    mask_rnb<D>(x, F.ibm, IBM_E, u);
    mask_rnb<D>(x, F.ibm, IBM_T, w);
    if (D == 3) mask_rnb<D>(x, F.ibm, IBM_N, v);
  }
  gather_nb<D>(x, dd, dens);
  interp_x<D>(g, dens, sx, i);
  interp_z<D>(g, dens, sz, k);
  if (D == 3) interp_y(g, dens, sy, j);
  const int32_t *fl = g.flag;
  bool cw = computed(fl[x.w]), cb = computed(fl[x.b]);
  bool cs_ = (D == 3) ? computed(fl[x.s]) : true;
  bool mus = K.muscl > 0 && fc == FLUID;
  long o = (long)ph * n;
  Faces<FT> fu, fv, fw;
  flux_u<D>(g, K, cw, cs_, cb, sx, u, v, w, i, j, k, mus && i != g.nx - 1, fu);
  flux_w<D>(g, K, cw, cs_, cb, sz, u, v, w, i, j, k, mus && k != g.nz - 1, fw);
  FT *m[3][3];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) m[a][b] = (FT *)F.mf[a][b];
  m[0][0][o + c] = fu.fe; m[0][2][o + c] = fu.ft;
  m[2][0][o + c] = fw.fe; m[2][2][o + c] = fw.ft;
  if (fu.has_w) { m[0][0][o + x.w] = fu.fw; m[2][0][o + x.w] = fw.fw; }
  if (fu.has_b) { m[0][2][o + x.b] = fu.fb; m[2][2][o + x.b] = fw.fb; }
  if (D == 3) {
    flux_v(g, K, cw, cs_, cb, sy, u, v, w, i, j, k, mus && j != g.ny - 1, fv);
    m[0][1][o + c] = fu.fn; m[1][0][o + c] = fv.fe; m[1][1][o + c] = fv.fn;
    m[1][2][o + c] = fv.ft; m[2][1][o + c] = fw.fn;
    if (fu.has_w) m[1][0][o + x.w] = fv.fw;
    if (fu.has_b) m[1][2][o + x.b] = fv.fb;
    if (fu.has_s) { m[0][1][o + x.s] = fu.fs; m[1][1][o + x.s] = fv.fs; m[2][1][o + x.s] = fw.fs; }
  }
}

// Classes present near each tile (Fields::ptile), from the particle
// densities of the current state.  A momentum or enthalpy flux of class s at
// a cell reads that class's densities within three cells of it, so a tile
// without the class's bit has every such flux a product with a zero density:
// zero.  Gated like the kernels after the iteration (iter_ended).
__global__ void __launch_bounds__(kBlock) k_ptile_clear(Fields F, long ntiles, const int32_t *ctr, int gate,
                                                        int maxout) {
  if (!iter_ended(ctr, gate, maxout)) return;
  const long t = cell_id();
  if (t == 0) *F.tcount = 0;
  if (t < ntiles) F.ptile_own[t] = 0u;
}

// both consumers' previous class bits made unknown (every class of every
// tile is computed or zeroed at their next call): launched where other
// kernels write the particle flux arrays (the explicit scheme)
__global__ void __launch_bounds__(kBlock) k_ptile_unknown(Fields F, long ntiles) {
  const long t = cell_id();
  if (t >= ntiles) return;
  F.ptile_mprev[t] = 0xffffu;
  F.ptile_hprev[t] = 0xffffu;
}

template <int NS>
__global__ void __launch_bounds__(kBlock) k_ptile_mark(Grid g, Fields F, const int32_t *ctr, int gate, int maxout) {
  if (!iter_ended(ctr, gate, maxout)) return;
  const long c = cell_id();
  if (c >= g.n) return;
  uint32_t bits = 0;
#pragma unroll
  for (int is = 0; is < NS; ++is)
    if (F.f[PDAC_RLK][(long)is * g.n + c] != 0.0) bits |= 1u << is;
  if (!bits) return;
  int i, j, k;
  decode(g, c, i, j, k);
  __hip_atomic_fetch_or(F.ptile_own + phase_tile(F, i, j, k), bits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// prev: the consumer's class bits of its previous call, replaced by this
// call's; tiles with a class now or then go to the list (one atomic per wave)
__global__ void __launch_bounds__(kBlock) k_ptile_dilate(Fields F, uint32_t *prev, const int32_t *ctr, int gate,
                                                         int maxout) {
  if (!iter_ended(ctr, gate, maxout)) return;
  const long t = cell_id();
  const long ntiles = (long)F.tnx * F.tny * F.tnz;
  if (t >= ntiles) return;
  const int tx = (int)(t % F.tnx), ty = (int)((t / F.tnx) % F.tny), tz = (int)(t / ((long)F.tnx * F.tny));
  uint32_t v = 0;
  for (int dz = -1; dz <= 1; ++dz)
    for (int dy = -1; dy <= 1; ++dy)
      for (int dx = -1; dx <= 1; ++dx) {
        const int x = tx + dx, y = ty + dy, z = tz + dz;
        if (x < 0 || y < 0 || z < 0 || x >= F.tnx || y >= F.tny || z >= F.tnz) continue;
        v |= F.ptile_own[x + (long)F.tnx * (y + (long)F.tny * z)];
      }
  const uint32_t gone = prev[t] & ~v & 0xffffu;
  F.ptile[t] = v | (gone << 16);
  prev[t] = v;
  const bool want = (v | gone) != 0u;
  const uint64_t m = __ballot(want);
  if (!m) return;
  const int lane = threadIdx.x & 63, leader = __ffsll((unsigned long long)m) - 1;
  int base = 0;
  if (lane == leader) base = atomicAdd(F.tcount, __popcll(m));
  base = __shfl(base, leader);
  if (want) F.tlist[base + __popcll(m & ((1ull << lane) - 1ull))] = (int32_t)t;
}

// the faces k_momflux writes for one phase of a computed cell, set to zero
// (a class absent near the cell: its fluxes are products with zero densities)
template <int D, typename FT>
__device__ __forceinline__ void momflux_zero(const Grid &g, const Fields &F, long c, int ph, int i, int j, int k) {
  const long o = (long)ph * g.n;
  const int32_t *fl = g.flag;
  const long w = rawn(g, i, j, k, -1, 0, 0), b = rawn(g, i, j, k, 0, 0, -1);
  const long s = D == 3 ? rawn(g, i, j, k, 0, -1, 0) : c;
  const bool hw = !computed(fl[w]), hb = !computed(fl[b]), hs = D == 3 && !computed(fl[s]);
  FT *m[3][3];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int bb = 0; bb < 3; ++bb) m[a][bb] = (FT *)F.mf[a][bb];
  const FT z = FT(0);
  m[0][0][o + c] = z; m[0][2][o + c] = z; m[2][0][o + c] = z; m[2][2][o + c] = z;
  if (hw) { m[0][0][o + w] = z; m[2][0][o + w] = z; }
  if (hb) { m[0][2][o + b] = z; m[2][2][o + b] = z; }
  if (D == 3) {
    m[0][1][o + c] = z; m[1][0][o + c] = z; m[1][1][o + c] = z; m[1][2][o + c] = z; m[2][1][o + c] = z;
    if (hw) m[1][0][o + w] = z;
    if (hb) m[1][2][o + b] = z;
    if (hs) { m[0][1][o + s] = z; m[1][1][o + s] = z; m[2][1][o + s] = z; }
  }
}

// FP32 fluxes: at least three waves per SIMD (168 VGPRs, a few spilled): the
// kernel is bound by the latency of its gathers (4.00 -> 3.72 ms on the
// 256x256x192 cone against two waves at 182 VGPRs).  FP64 fluxes need 238
// VGPRs and keep two waves.
template <int D, int NS, typename FT>
__global__ void __launch_bounds__(kBlock, sizeof(FT) == 4 ? 3 : 2) k_momflux(Grid g, Consts K, Fields F) {
  long c = cell_id();
  if (c >= g.n) return;
  int fc = g.flag[c];
  if (!computed(fc)) return;
  int ph = blockIdx.y;
  const double *uu, *vv, *ww, *dd;
  if (ph == 0) { uu = F.f[PDAC_UG]; vv = F.f[PDAC_VG]; ww = F.f[PDAC_WG]; dd = F.f[PDAC_RGP]; }
  else {
    long o = (long)(ph - 1) * g.n;
    uu = F.f[PDAC_US] + o; vv = F.f[PDAC_VS] + o; ww = F.f[PDAC_WS] + o; dd = F.f[PDAC_RLK] + o;
  }
  if (ph > 0 && F.ptile) {
    int i, j, k;
    decode(g, c, i, j, k);
    if (!((F.ptile[phase_tile(F, i, j, k)] >> (ph - 1)) & 1u)) {
      momflux_zero<D, FT>(g, F, c, ph, i, j, k);
      return;
    }
  }
  momflux_cell<D, FT>(g, K, F, c, fc, ph, uu, vv, ww, dd);
}

// The particle classes' momentum fluxes of a 3D mesh over the listed tiles
// (Fields::tlist), a wave per tile and class (blockIdx.y + 1), a lane per
// cell of the tile: a class present near the tile is computed, one present at
// the previous call only is set to zero (its old values would stand); the
// other classes' fluxes hold the zeros of such a call or of creation.
template <int NS, typename FT>
__global__ void __launch_bounds__(kBlock, sizeof(FT) == 4 ? 3 : 2) k_momflux_tiles(Grid g, Consts K, Fields F) {
  const int ph = blockIdx.y + 1, lane = threadIdx.x & 63;
  const long o = (long)(ph - 1) * g.n;
  const double *uu = F.f[PDAC_US] + o, *vv = F.f[PDAC_VS] + o, *ww = F.f[PDAC_WS] + o, *dd = F.f[PDAC_RLK] + o;
  const int ntl = *F.tcount;
  const long stride = (long)gridDim.x * (kBlock / 64);
  for (long w = (long)blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6); w < ntl; w += stride) {
    const long t = F.tlist[w];
    const uint32_t bits = F.ptile[t];
    const bool on = (bits >> (ph - 1)) & 1u, off = (bits >> (ph + 15)) & 1u;
    if (!on && !off) continue;
    int i, j, k;
    tile_cell(F, t, lane, i, j, k);
    if (i > g.nx || j > g.ny || k > g.nz) continue;
    const long c = cell_index(g, i, j, k);
    const int fc = g.flag[c];
    if (!computed(fc)) continue;
    if (on) momflux_cell<3, FT>(g, K, F, c, fc, ph, uu, vv, ww, dd);
    else momflux_zero<3, FT>(g, F, c, ph, i, j, k);
  }
}

// inter, drag.f:173-246: packed symmetric interphase matrix with zero row sums
template <int NS>
__device__ __forceinline__ void inter_d(const Consts &K, const double *kpgv, double dt,
                                        const double *rlkc, const double (&du_pp)[NS][NS],
                                        const double (&dv_pp)[NS][NS], const double (&dw_pp)[NS][NS],
                                        double *au, double *av, double *aw) {
  constexpr int NP = NS + 1;
  int ks = 0;
#pragma unroll
  for (int k = 1; k <= NP; ++k) {
    if (k == 1) {
      ks = ks + 1;
      au[ks - 1] = 0.0; av[ks - 1] = 0.0; aw[ks - 1] = 0.0;
    } else {
#pragma unroll
      for (int kk = 1; kk <= k; ++kk) {
        ks = ks + 1;
        if (kk == 1) {
          au[ks - 1] = -kpgv[k - 2] * dt; av[ks - 1] = -kpgv[k - 2] * dt; aw[ks - 1] = -kpgv[k - 2] * dt;
        } else if (kk == k) {
          au[ks - 1] = 0.0; av[ks - 1] = 0.0; aw[ks - 1] = 0.0;
        } else {
          double con = ppdrag(K, rlkc, k, kk);
          au[ks - 1] = -(con * du_pp[k - 2][kk - 2]) * dt;
          av[ks - 1] = -(con * dv_pp[k - 2][kk - 2]) * dt;
          aw[ks - 1] = -(con * dw_pp[k - 2][kk - 2]) * dt;
        }
      }
    }
  }
#pragma unroll
  for (int k = 1; k <= NP; ++k) {
    double sx = 0.0, sy = 0.0, sz = 0.0;
#pragma unroll
    for (int kk = 1; kk <= k; ++kk) {
      int i_ = kk + k * (k - 1) / 2 - 1;
      sx = sx + au[i_]; sy = sy + av[i_]; sz = sz + aw[i_];
    }
#pragma unroll
    for (int kk = k; kk <= NP; ++kk) {
      int i_ = k + kk * (kk - 1) / 2 - 1;
      sx = sx + au[i_]; sy = sy + av[i_]; sz = sz + aw[i_];
    }
    int d = k * (k + 1) / 2 - 1;
    au[d] = -sx; av[d] = -sy; aw[d] = -sz;
  }
}

// interphase matrix of one axis (inter, drag.f:173-246): packed symmetric
// entries with zero row sums, from the drag coefficients, the
// particle-particle factors con (packed by pair, ppdrag) and the relative
// particle velocities along the axis.  The entries of each axis depend on
// that axis only, so the three axes of inter_d come out of three calls.
template <int NS>
__device__ __forceinline__ void inter_axis(const double *kpgv, double dt, const double *con,
                                           const double (&d_pp)[NS][NS], double *a) {
  constexpr int NP = NS + 1;
  int ks = 0, pc = 0;
#pragma unroll
  for (int k = 1; k <= NP; ++k) {
    if (k == 1) {
      ks = ks + 1;
      a[ks - 1] = 0.0;
    } else {
#pragma unroll
      for (int kk = 1; kk <= k; ++kk) {
        ks = ks + 1;
        if (kk == 1) {
          a[ks - 1] = -kpgv[k - 2] * dt;
        } else if (kk == k) {
          a[ks - 1] = 0.0;
        } else {
          a[ks - 1] = -(con[pc] * d_pp[k - 2][kk - 2]) * dt;
          ++pc;
        }
      }
    }
  }
#pragma unroll
  for (int k = 1; k <= NP; ++k) {
    double sx = 0.0;
#pragma unroll
    for (int kk = 1; kk <= k; ++kk) sx = sx + a[kk + k * (k - 1) / 2 - 1];
#pragma unroll
    for (int kk = k; kk <= NP; ++kk) sx = sx + a[k + kk * (kk - 1) / 2 - 1];
    a[k * (k + 1) / 2 - 1] = -sx;
  }
}

// tilde (part 2), mesh_loop of tilde.f:418-806: explicit momenta, drag,
// interphase matrix and the (1-alpha) coupling.  Neighbour interphase
// entries come from the copy taken before this call (the serial loop reads
// east/north/top neighbours before it reaches them).  The axes share no
// accumulator: one thread per cell and axis (x, y, z; x, z in 2D), each
// evaluating the drag coefficients, which need every velocity component,
// itself.  The kernel is bound by the latency of its gathers, so three times
// the threads, each with a third of the axis work and fewer registers live
// (three waves per SIMD), keep more of them in flight: 5.96 -> 4.53 ms on the
// 256x256x192 cone.
// the gas-particle drag coefficient of each class at a computed cell
// (kdrags or sdrag, drag.f:11-157, from the cell-centred slip velocities),
// once per cell for the axis threads of k_tilde
template <int D, int NS>
__global__ void __launch_bounds__(kBlock) k_drag(Grid g, Consts K, Fields F) {
  const long c = cell_id();
  if (c >= g.n || !computed(g.flag[c])) return;
  const long n = g.n;
  int i, j, k;
  decode(g, c, i, j, k);
  Nbr<D> q;
  first_nbrs<D>(g, i, j, k, q);
  const double *ug = F.f[PDAC_UG], *vg = F.f[PDAC_VG], *wg = F.f[PDAC_WG];
  const double *us = F.f[PDAC_US], *vs = F.f[PDAC_VS], *ws = F.f[PDAC_WS];
  const double *rgp = F.f[PDAC_RGP], *rlk = F.f[PDAC_RLK], *ep = F.f[PDAC_EP], *mug = F.f[PDAC_MUG];
#pragma unroll
  for (int is = 0; is < NS; ++is) {
    const long os = (long)is * n;
    const double dugs = ((ug[c] - us[os + c]) + (ug[q.imjk] - us[os + q.imjk])) * 0.5;
    const double dwgs = ((wg[c] - ws[os + c]) + (wg[q.ijkm] - ws[os + q.ijkm])) * 0.5;
    const double dvgs = (D == 3) ? ((vg[c] - vs[os + c]) + (vg[q.ijmk] - vs[os + q.ijmk])) * 0.5 : 0.0;
    // with no particles of the class here both laws give the density times
    // nonnegative finite factors: a zero of the density's sign, without the
    // powers and roots (most cells of a plume hold no particles)
    const double rl = rlk[os + c];
    F.kpgv[os + c] = rl == 0.0 ? rl * 0.0
                   : K.drag_model == 1 ? kdrags(K, is, dugs, dvgs, dwgs, ep[c], rgp[c], rl, mug[c])
                                       : sdrag(K, is, dugs, dvgs, dwgs, ep[c], rgp[c], rl, mug[c]);
  }
}

// (VISC: with the viscous stress divergence of k_stress)
template <int D, int NS, typename FT, bool VISC>
__global__ void __launch_bounds__(kBlock, 3) k_tilde(Grid g, Consts K, Fields F, Ib ib, double dt, int old_zero) {
  int part;
  long c = cell_part_id(D == 2 ? 2 : 3, part);
  if (c >= g.n) return;
  if (!computed(g.flag[c])) return;
  constexpr int NP = NS + 1;
  constexpr int NA = NP * (NP + 1) / 2;
  constexpr int NPAIR = NS * (NS - 1) / 2;
  long n = g.n;
  int i, j, k;
  decode(g, c, i, j, k);
  Nbr<D> q;
  first_nbrs<D>(g, i, j, k, q);
  const double *dx = g.dx, *dy = g.dy, *dz = g.dz;
  const double *rgpn = F.f[PDAC_RGPN], *rlkn = F.f[PDAC_RLKN], *pn = F.f[PDAC_PN], *epn = F.f[PDAC_EPN];
  const double *ug = F.f[PDAC_UG], *vg = F.f[PDAC_VG], *wg = F.f[PDAC_WG];
  const double *us = F.f[PDAC_US], *vs = F.f[PDAC_VS], *ws = F.f[PDAC_WS];
  const double *rgp = F.f[PDAC_RGP], *rlk = F.f[PDAC_RLK], *ep = F.f[PDAC_EP];
  const double alpha = K.alpha, ag = K.alphagrav;
  // drag coefficients (k_drag) and particle-particle factors (drag.f:265-295)
  double kpgv[NS], rlkc[NS], con[NPAIR > 0 ? NPAIR : 1];
#pragma unroll
  for (int is = 0; is < NS; ++is) {
    const long os = (long)is * n;
    rlkc[is] = rlk[os + c];
    kpgv[is] = F.kpgv[os + c];
  }
  {
    int pc = 0;
#pragma unroll
    for (int kk0 = 3; kk0 <= NP; ++kk0)
#pragma unroll
      for (int kk = 2; kk < kk0; ++kk) con[pc++] = ppdrag(K, rlkc, kk0, kk);
  }
  double dxi = dx[i], dxip1 = dx[i + 1], dyj = dy[j], dyjp1 = dy[j + 1];
  double dzk = dz[k], dzkp1 = dz[k + 1];
  double indxp = 1.0 / (dx[i] + dx[i + 1]), indyp = 1.0 / (dy[j] + dy[j + 1]);
  double indzp = 1.0 / (dz[k] + dz[k + 1]);
  {
    const int ax = D == 2 ? (part == 0 ? 0 : 2) : part;
    // this axis: neighbour across the + face, the - face raw neighbour, spacing
    const long qn = ax == 0 ? q.e : ax == 1 ? q.n : q.t;
    const long qm = ax == 0 ? q.imjk : ax == 1 ? q.ijmk : q.ijkm;
    const double d0 = ax == 0 ? dxi : ax == 1 ? dyj : dzk;
    const double d1 = ax == 0 ? dxip1 : ax == 1 ? dyjp1 : dzkp1;
    const double ind = ax == 0 ? indxp : ax == 1 ? indyp : indzp;
    const double *vgx = ax == 0 ? ug : ax == 1 ? vg : wg;
    const double *vsx = ax == 0 ? us : ax == 1 ? vs : ws;
    // momentum fluxes of this component (stored in FT), differenced in double
    const FT *m0 = (const FT *)F.mf[ax][0], *m1 = (const FT *)F.mf[ax][1], *m2 = (const FT *)F.mf[ax][2];
    double epn_f = 1.0;
    if (K.pmodel == 1) epn_f = (d0 * epn[qn] + d1 * epn[c]) * ind;
    double t;
    double r0, rs[NS];
    {
      double fx = (double)m0[c] - (double)m0[q.imjk];
      double fz = (double)m2[c] - (double)m2[q.ijkm];
      double fy = (D == 3) ? (double)m1[c] - (double)m1[q.ijmk] : 0.0;
      // the viscous stress divergence (gvisx/y/z, tilde.f:487-506) opens the sum
      t = VISC ? F.vis[ax][c] : 0.0;
      if (ax == 0) {
        t = t - indxp * 2.0 * fx * g.inrb[i];
        t = t + (1.0 - ag) * (dx[i + 1] * rgpn[c] + dx[i] * rgpn[q.e]) * indxp * K.gravx;
        t = t - g.indy[j] * fy;
        t = t - g.indz[k] * fz;
        t = t + (1.0 - alpha) * indxp * 2.0 * epn_f * (pn[c] - pn[q.e]);
        r0 = F.f[PDAC_RUGN][c] + dt * t;
      } else if (ax == 1) {
        t = t + (1.0 - ag) * (dy[j + 1] * rgpn[c] + dy[j] * rgpn[q.n]) * indyp * K.gravy;
        t = t - g.indx[i] * fx;
        t = t - indyp * 2.0 * fy;
        t = t - g.indz[k] * fz;
        t = t + (1.0 - alpha) * indyp * 2.0 * epn_f * (pn[c] - pn[q.n]);
        r0 = F.f[PDAC_RVGN][c] + dt * t;
      } else {
        t = t + (1.0 - ag) * (dz[k + 1] * rgpn[c] + dz[k] * rgpn[q.t]) * indzp * K.gravz;
        t = t - g.indx[i] * fx * g.inr[i];
        t = t - g.indy[j] * fy;
        t = t - indzp * 2.0 * fz;
        t = t + (1.0 - alpha) * indzp * 2.0 * epn_f * (pn[c] - pn[q.t]);
        r0 = F.f[PDAC_RWGN][c] + dt * t;
      }
    }
#pragma unroll
    for (int is = 0; is < NS; ++is) {
      const long o = (long)(is + 1) * n, os = (long)is * n;
      double fx = (double)m0[o + c] - (double)m0[o + q.imjk];
      double fz = (double)m2[o + c] - (double)m2[o + q.ijkm];
      double fy = (D == 3) ? (double)m1[o + c] - (double)m1[o + q.ijmk] : 0.0;
      double epsn_f = 0.0;
      if (K.pmodel == 1) epsn_f = (d0 * rlkn[os + qn] + d1 * rlkn[os + c]) * ind * K.inrl[is];
      const double sink = 0.0;
      t = VISC ? F.vis[ax][o + c] : 0.0;
      if (ax == 0) {
        t = t - indxp * 2.0 * fx * g.inrb[i];
        t = t + (1.0 - ag) * indxp * K.gravx * (dx[i + 1] * rlkn[os + c] + dx[i] * rlkn[os + q.e]);
        t = t - g.indy[j] * fy;
        t = t - g.indz[k] * fz;
        t = t + (1.0 - alpha) * indxp * 2.0 * epsn_f * (pn[c] - pn[q.e]);
        rs[is] = F.f[PDAC_RUSN][os + c] + dt * t;
      } else if (ax == 1) {
        t = t + (1.0 - ag) * indyp * K.gravy * (dy[j + 1] * rlkn[os + c] + dy[j] * rlkn[os + q.n]);
        t = t - g.indx[i] * fx;
        t = t - indyp * 2.0 * fy;
        t = t - g.indz[k] * fz;
        t = t + (1.0 - alpha) * indyp * 2.0 * epsn_f * (pn[c] - pn[q.n]);
        rs[is] = F.f[PDAC_RVSN][os + c] + dt * t;
      } else {
        t = t + (1.0 - ag) * indzp * K.gravz * (rlkn[os + c] * dz[k + 1] + rlkn[os + q.t] * dz[k]);
        t = t - g.indx[i] * fx * g.inr[i];
        t = t - g.indy[j] * fy;
        t = t - indzp * 2.0 * fz;
        t = t + (1.0 - alpha) * indzp * 2.0 * epsn_f * (pn[c] - pn[q.t]);
        rs[is] = F.f[PDAC_RWSN][os + c] + dt * t;
      }
      rs[is] = rs[is] + dt * sink * vsx[os + c];
    }
    // relative particle-particle velocities along the axis (drag.f:216-218)
    // and this axis's interphase matrix
    double d_pp[NS][NS];
#pragma unroll
    for (int a = 0; a < NS; ++a)
#pragma unroll
      for (int b = 0; b < NS; ++b) {
        long oa = (long)a * n, ob = (long)b * n;
        d_pp[a][b] = fabs(vsx[oa + c] - vsx[ob + c] + vsx[oa + qm] - vsx[ob + qm]) * 0.5;
      }
    double am[NA];
    inter_axis<NS>(kpgv, dt, con, d_pp, am);
    double *app = F.f[PDAC_APPU + ax];
#pragma unroll
    for (int a = 0; a < NA; ++a) app[(long)a * n + c] = am[a];
    // (1-alpha) interphase coupling, tilde.f:637-706.  The serial loop reads
    // the east/north/top entries before it reaches those cells: they hold the
    // previous call's values (zero after allocate_fluxes), unless the
    // neighbour is the cell itself.
    const double *oapp = F.appu_old[ax];
    auto AN = [&](int ls) { return qn == c ? am[ls] : (old_zero ? 0.0 : oapp[(long)ls * n + qn]); };
    auto CN = [&](int ls) { return (1.0 - alpha) * (d0 * AN(ls) + d1 * am[ls]) * ind; };
    r0 = r0 - CN(0) * vgx[c];
#pragma unroll
    for (int l = 2; l <= NP; ++l) {
      int ls1 = l * (l - 1) / 2;
      long ol = (long)(l - 2) * n;
#pragma unroll
      for (int ll = 1; ll <= l; ++ll) {
        int ls = ls1 + ll - 1;
        if (ll == 1) {
          r0 = r0 - CN(ls) * vsx[ol + c];
          rs[l - 2] = rs[l - 2] - CN(ls) * vgx[c];
        } else if (ll == l) {
          rs[l - 2] = rs[l - 2] - CN(ls) * vsx[ol + c];
        } else {
          long oll = (long)(ll - 2) * n;
          rs[l - 2] = rs[l - 2] - CN(ls) * vsx[oll + c];
          rs[ll - 2] = rs[ll - 2] - CN(ls) * vsx[ol + c];
        }
      }
    }
    // forcing on the immersed boundary, tilde.f:711-803: the explicit terms
    // of a forcing point's component drive it to the point's pseudo-velocity
    if (ib.num[ax] && g.flag[c] == IMMB_CELL && ib.num[ax][c] != 0) {
      const IbAxis &A = ib.ax[ax];
      const int pt = ib.num[ax][c] - 1;
      const double *pr = F.f[PDAC_P];
      double force = (ind * (d1 * rgp[c] + d0 * rgp[qn]) * A.vel[pt] - r0) / dt;
      double epf = (d0 * ep[qn] + d1 * ep[c]) * ind;
      double presn = -ind * 2.0 * epf * (pr[qn] - pr[c]);
      double dragn = 0.0;
      double drel[NS];
#pragma unroll
      for (int is = 0; is < NS; ++is) {
        const long os = (long)is * n;
        drel[is] = ax == 0 ? ((ug[c] - us[os + c]) + (ug[q.imjk] - us[os + q.imjk])) * 0.5
                 : ax == 1 ? ((vg[c] - vs[os + c]) + (vg[q.ijmk] - vs[os + q.ijmk])) * 0.5
                           : ((wg[c] - ws[os + c]) + (wg[q.ijkm] - ws[os + q.ijkm])) * 0.5;
        dragn = dragn - kpgv[is] * drel[is];
      }
      force = force - presn - dragn;
      r0 = r0 + dt * force;
#pragma unroll
      for (int is = 0; is < NS; ++is) {
        const long os = (long)is * n;
        force = (ind * (d1 * rlk[os + c] + d0 * rlk[os + qn]) * A.vel[(long)(1 + is) * A.nfp + pt] - rs[is]) / dt;
        epf = (d0 * rlk[os + qn] + d1 * rlk[os + c]) * ind * K.inrl[is];
        presn = -ind * 2.0 * epf * (pr[qn] - pr[c]);
        dragn = kpgv[is] * drel[is];
        force = force - presn - dragn;
        rs[is] = rs[is] + dt * force;
      }
    }
    F.f[PDAC_RUG + ax][c] = r0;
#pragma unroll
    for (int is = 0; is < NS; ++is) F.f[PDAC_RUS + ax][(long)is * n + c] = rs[is];
  }
}

#include "pdac_visc.hip.inc"
#include "pdac_iter.hip.inc"
#include "pdac_energy.hip.inc"
#include "pdac_explicit.hip.inc"
#include "pdac_ensemble.hip.inc"
#include "pdac_io.hip.inc"
#include "pdac_api.hip.inc"
#include "pdac_setup.hip.inc"
