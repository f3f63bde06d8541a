// pdac_flux.hpp -- donor-cell + MUSCL face fluxes of the PDAC hot path
// (mass: flux_m.f; momentum: flux_u.f, flux_v.f, flux_w.f; scalars:
// flux_sc.f; staggered densities: interp.f).  Device functions evaluated
// by one thread for one cell and one phase.
#pragma once

#include "pdac_dev.hpp"

namespace pdac {

// a grid array (dx, 1/dx, r, ...) read in the arithmetic type T
template <typename T>
struct GA {
  const double *p;
  __device__ __forceinline__ T operator[](int i) const { return T(p[i]); }
};

// values of one field on the stencil of a cell (subscr.f:63-210 names)
template <typename T>
struct Stc {
  T c, e, w, n, s, t, b, ee, ww, nn, ss, tt, bb;
  T en, es, et, eb, nt, wn, nb, wt, st;
  T eee, nnn, ttt, een, eet, enn, nnt, ett, ntt;
};

// Full stencil indices of a computed cell used by the momentum fluxes:
// raw (myijk) for velocities, reflected (myinds) for densities.
template <int D>
struct FullIdx {
  // raw
  long c, e, w, n, s, t, b, ee, nn, tt, en, et, es, eb, nt, wn, nb, wt, st;
  long eee, nnn, ttt, een, eet, enn, nnt, ett, ntt;
  // reflected
  long Re, Rw, Rn, Rs, Rt, Rb, Ree, Rnn, Rtt, Ren, Ret, Res, Reb, Rnt, Rwn, Rnb, Rwt, Rst;
  long Reee, Rnnn, Rttt, Reen, Reet, Renn, Rnnt, Rett, Rntt;
};

// the stencil of a cell marked regular (k_regular): every index is the cell
// index plus a fixed offset, and every reflected index is the raw one
template <int D>
__device__ __forceinline__ void full_idx_off(const Grid &g, FullIdx<D> &x) {
  const long c = x.c;
  const long sy = D == 3 ? (long)g.nx : 0, sz = (long)g.nx * g.ny;
  x.e = c + 1; x.w = c - 1; x.t = c + sz; x.b = c - sz;
  x.ee = c + 2; x.tt = c + 2 * sz;
  x.et = c + 1 + sz; x.eb = c + 1 - sz; x.wt = c - 1 + sz;
  x.eee = c + 3; x.ttt = c + 3 * sz;
  x.eet = c + 2 + sz; x.ett = c + 1 + 2 * sz;
  if (D == 2) {
    x.n = x.s = x.nn = x.en = x.es = x.nt = x.wn = x.nb = x.st = c;
    x.nnn = x.een = x.enn = x.nnt = x.ntt = c;
  } else {
    x.n = c + sy; x.s = c - sy; x.nn = c + 2 * sy;
    x.en = c + 1 + sy; x.es = c + 1 - sy; x.nt = c + sy + sz; x.wn = c - 1 + sy;
    x.nb = c + sy - sz; x.st = c - sy + sz;
    x.nnn = c + 3 * sy;
    x.een = c + 2 + sy; x.enn = c + 1 + 2 * sy; x.nnt = c + 2 * sy + sz; x.ntt = c + sy + 2 * sz;
  }
  x.Re = x.e; x.Rw = x.w; x.Rn = x.n; x.Rs = x.s; x.Rt = x.t; x.Rb = x.b;
  x.Ree = x.ee; x.Rnn = x.nn; x.Rtt = x.tt;
  x.Ren = x.en; x.Ret = x.et; x.Res = x.es; x.Reb = x.eb; x.Rnt = x.nt; x.Rwn = x.wn;
  x.Rnb = x.nb; x.Rwt = x.wt; x.Rst = x.st;
  x.Reee = x.eee; x.Rnnn = x.nnn; x.Rttt = x.ttt; x.Reen = x.een; x.Reet = x.eet;
  x.Renn = x.enn; x.Rnnt = x.nnt; x.Rett = x.ett; x.Rntt = x.ntt;
}

template <int D>
__device__ __forceinline__ void full_idx(const Grid &g, int i, int j, int k, FullIdx<D> &x) {
  const int32_t *fl = g.flag;
  x.c = cell_index(g, i, j, k);
  if (g.reg && g.reg[x.c]) {
    full_idx_off<D>(g, x);
    return;
  }
  x.e = rawn(g, i, j, k, 1, 0, 0);  x.w = rawn(g, i, j, k, -1, 0, 0);
  x.t = rawn(g, i, j, k, 0, 0, 1);  x.b = rawn(g, i, j, k, 0, 0, -1);
  x.ee = rawn(g, i, j, k, 2, 0, 0); x.tt = rawn(g, i, j, k, 0, 0, 2);
  x.et = rawn(g, i, j, k, 1, 0, 1); x.eb = rawn(g, i, j, k, 1, 0, -1);
  x.wt = rawn(g, i, j, k, -1, 0, 1);
  x.eee = rawn(g, i, j, k, 3, 0, 0); x.ttt = rawn(g, i, j, k, 0, 0, 3);
  x.eet = rawn(g, i, j, k, 2, 0, 1); x.ett = rawn(g, i, j, k, 1, 0, 2);
  long c = x.c;
  x.Re = is_wallf(fl[x.e]) ? c : x.e;
  x.Rw = is_wallf(fl[x.w]) ? c : x.w;
  x.Rt = is_wallf(fl[x.t]) ? c : x.t;
  x.Rb = is_wallf(fl[x.b]) ? c : x.b;
  x.Ree = (i == g.nx - 1) ? x.Re : (is_wallf(fl[x.ee]) ? x.e : x.ee);
  x.Rtt = (k == g.nz - 1) ? x.Rt : (is_wallf(fl[x.tt]) ? x.t : x.tt);
  x.Reee = is_wallf(fl[x.eee]) ? x.ee : x.eee;
  if (i == g.nx - 1) x.Reee = x.Re;
  if (i == g.nx - 2) x.Reee = x.Ree;
  x.Rttt = is_wallf(fl[x.ttt]) ? x.tt : x.ttt;
  if (k == g.nz - 1) x.Rttt = x.Rt;
  if (k == g.nz - 2) x.Rttt = x.Rtt;
  if (D == 2) {
    x.n = x.s = x.nn = x.en = x.es = x.nt = x.wn = x.nb = x.st = c;
    x.nnn = x.een = x.enn = x.nnt = x.ntt = c;
    x.Rn = x.Rs = x.Rnn = x.Ren = x.Res = x.Rnt = x.Rwn = x.Rnb = x.Rst = c;
    x.Rnnn = x.Reen = x.Renn = x.Rnnt = x.Rntt = c;
    x.Rwt = refl_diag(fl, c, x.wt, x.t, x.w);
    x.Reb = refl_diag(fl, c, x.eb, x.b, x.e);
    x.Ret = refl_diag(fl, c, x.et, x.t, x.e);
    x.Reet = (i == g.nx - 1) ? x.Ret : refl_diag2(fl, x.e, x.eet, x.et, x.ee);
    x.Rett = (k == g.nz - 1) ? x.Ret : refl_diag2(fl, x.t, x.ett, x.et, x.tt);
  } else {
    x.n = rawn(g, i, j, k, 0, 1, 0);  x.s = rawn(g, i, j, k, 0, -1, 0);
    x.nn = rawn(g, i, j, k, 0, 2, 0);
    x.en = rawn(g, i, j, k, 1, 1, 0); x.es = rawn(g, i, j, k, 1, -1, 0);
    x.nt = rawn(g, i, j, k, 0, 1, 1); x.wn = rawn(g, i, j, k, -1, 1, 0);
    x.nb = rawn(g, i, j, k, 0, 1, -1); x.st = rawn(g, i, j, k, 0, -1, 1);
    x.nnn = rawn(g, i, j, k, 0, 3, 0);
    x.een = rawn(g, i, j, k, 2, 1, 0); x.enn = rawn(g, i, j, k, 1, 2, 0);
    x.nnt = rawn(g, i, j, k, 0, 2, 1); x.ntt = rawn(g, i, j, k, 0, 1, 2);
    x.Rn = is_wallf(fl[x.n]) ? c : x.n;
    x.Rs = is_wallf(fl[x.s]) ? c : x.s;
    x.Rnn = (j == g.ny - 1) ? x.Rn : (is_wallf(fl[x.nn]) ? x.Rn : x.nn);
    x.Rnnn = is_wallf(fl[x.nnn]) ? x.nn : x.nnn;
    if (j == g.ny - 1) x.Rnnn = x.Rn;
    if (j == g.ny - 2) x.Rnnn = x.Rnn;
    // 3D diagonals are raw except the bottom ones (ghost.f:1640-1780)
    x.Ren = x.en; x.Rwn = x.wn; x.Res = x.es; x.Ret = x.et; x.Rwt = x.wt;
    x.Rnt = x.nt; x.Rst = x.st;
    x.Reb = is_wallf(fl[x.eb]) ? x.e : x.eb;
    x.Rnb = is_wallf(fl[x.nb]) ? x.n : x.nb;
    x.Reet = (i == g.nx - 1) ? x.Ret : x.eet;
    x.Reen = (i == g.nx - 1) ? x.Ren : x.een;
    x.Rnnt = (j == g.ny - 1) ? x.Rnt : x.nnt;
    x.Renn = (j == g.ny - 1) ? x.Ren : x.enn;
    x.Rett = (k == g.nz - 1) ? x.Ret : x.ett;
    x.Rntt = (k == g.nz - 1) ? x.Rnt : x.ntt;
  }
}

// reflected full stencil of a density (nb, subscr.f:2323-2383)
template <int D, typename T>
__device__ __forceinline__ void gather_nb(const FullIdx<D> &x, const double *a, Stc<T> &o) {
  o.c = T(a[x.c]); o.e = T(a[x.Re]); o.w = T(a[x.Rw]); o.t = T(a[x.Rt]); o.b = T(a[x.Rb]);
  o.ee = T(a[x.Ree]); o.tt = T(a[x.Rtt]); o.et = T(a[x.Ret]); o.eb = T(a[x.Reb]); o.wt = T(a[x.Rwt]);
  o.eee = T(a[x.Reee]); o.ttt = T(a[x.Rttt]); o.eet = T(a[x.Reet]); o.ett = T(a[x.Rett]);
  if (D == 3) {
    o.n = T(a[x.Rn]); o.s = T(a[x.Rs]); o.nn = T(a[x.Rnn]); o.en = T(a[x.Ren]); o.es = T(a[x.Res]);
    o.nt = T(a[x.Rnt]); o.wn = T(a[x.Rwn]); o.nb = T(a[x.Rnb]); o.st = T(a[x.Rst]);
    o.nnn = T(a[x.Rnnn]); o.een = T(a[x.Reen]); o.enn = T(a[x.Renn]); o.nnt = T(a[x.Rnnt]);
    o.ntt = T(a[x.Rntt]);
  }
}

// raw stencil of a velocity (rnb_r, subscr.f:2386-2450), entries used by the
// momentum fluxes
template <int D, typename T>
__device__ __forceinline__ void gather_rnb(const FullIdx<D> &x, const double *a, Stc<T> &o) {
  o.c = T(a[x.c]); o.e = T(a[x.e]); o.w = T(a[x.w]); o.t = T(a[x.t]); o.b = T(a[x.b]);
  o.ee = T(a[x.ee]); o.tt = T(a[x.tt]); o.wt = T(a[x.wt]); o.eb = T(a[x.eb]);
  if (D == 3) {
    o.n = T(a[x.n]); o.s = T(a[x.s]); o.nn = T(a[x.nn]); o.es = T(a[x.es]); o.st = T(a[x.st]);
    o.wn = T(a[x.wn]); o.nb = T(a[x.nb]);
  }
}

// maskstencil(u, rnb(mask, b_, ijk)) (subscr.f:1548-1582, tilde.f:1072-1118):
// each raw-neighbour entry of a velocity stencil times that cell's face mask
// (bit 'bit' of ibm), on the entries gather_rnb fills
template <int D, typename T>
__device__ __forceinline__ void mask_rnb(const FullIdx<D> &x, const uint8_t *ibm, uint8_t bit, Stc<T> &o) {
  auto mk = [&](long c) { return (ibm[c] & bit) ? T(1) : T(0); };
  o.c = o.c * mk(x.c); o.e = o.e * mk(x.e); o.w = o.w * mk(x.w); o.t = o.t * mk(x.t); o.b = o.b * mk(x.b);
  o.ee = o.ee * mk(x.ee); o.tt = o.tt * mk(x.tt); o.wt = o.wt * mk(x.wt); o.eb = o.eb * mk(x.eb);
  if (D == 3) {
    o.n = o.n * mk(x.n); o.s = o.s * mk(x.s); o.nn = o.nn * mk(x.nn); o.es = o.es * mk(x.es);
    o.st = o.st * mk(x.st); o.wn = o.wn * mk(x.wn); o.nb = o.nb * mk(x.nb);
  }
}

// ----------------------------------------------------------------------
// staggered densities, interp.f:14-320 (entries the momentum fluxes read)
template <int D, typename T>
__device__ __forceinline__ void interp_x(const Grid &g, const Stc<T> &f, Stc<T> &o, int i) {
  const GA<T> dx{g.dx};
  int ip2 = imin(i + 2, g.nx), ip3 = imin(i + 3, g.nx);
  T indxm = T(1) / (dx[i] + dx[i - 1]), indxp = T(1) / (dx[i] + dx[i + 1]);
  T indxpp = T(1) / (dx[i + 1] + dx[ip2]), indxppp = T(1) / (dx[ip2] + dx[ip3]);
  o.c = (dx[i + 1] * f.c + dx[i] * f.e) * indxp;
  o.e = (dx[ip2] * f.e + dx[i + 1] * f.ee) * indxpp;
  o.t = (dx[i + 1] * f.t + dx[i] * f.et) * indxp;
  o.w = (dx[i] * f.w + dx[i - 1] * f.c) * indxm;
  o.b = (dx[i + 1] * f.b + dx[i] * f.eb) * indxp;
  o.ee = (dx[ip2] * f.eee + dx[ip3] * f.ee) * indxppp;
  o.tt = (dx[i] * f.ett + dx[i + 1] * f.tt) * indxp;
  if (D == 3) {
    o.n = (dx[i + 1] * f.n + dx[i] * f.en) * indxp;
    o.s = (dx[i + 1] * f.s + dx[i] * f.es) * indxp;
    o.nn = (dx[i] * f.enn + dx[i + 1] * f.nn) * indxp;
  }
}

template <typename T>
__device__ __forceinline__ void interp_y(const Grid &g, const Stc<T> &f, Stc<T> &o, int j) {
  const GA<T> dy{g.dy};
  int jp2 = imin(j + 2, g.ny), jp3 = imin(j + 3, g.ny);
  T indym = T(1) / (dy[j] + dy[j - 1]), indyp = T(1) / (dy[j] + dy[j + 1]);
  T indypp = T(1) / (dy[j + 1] + dy[jp2]), indyppp = T(1) / (dy[jp2] + dy[jp3]);
  o.c = (dy[j + 1] * f.c + dy[j] * f.n) * indyp;
  o.e = (dy[j + 1] * f.e + dy[j] * f.en) * indyp;
  o.n = (dy[jp2] * f.n + dy[j + 1] * f.nn) * indypp;
  o.t = (dy[j + 1] * f.t + dy[j] * f.nt) * indyp;
  o.w = (dy[j + 1] * f.w + dy[j] * f.wn) * indyp;
  o.s = (dy[j] * f.s + dy[j - 1] * f.c) * indym;
  o.b = (dy[j + 1] * f.b + dy[j] * f.nb) * indyp;
  o.ee = (dy[j] * f.een + dy[j + 1] * f.ee) * indyp;
  o.nn = (dy[jp2] * f.nnn + dy[jp3] * f.nn) * indyppp;
  o.tt = (dy[j] * f.ntt + dy[j + 1] * f.tt) * indyp;
}

template <int D, typename T>
__device__ __forceinline__ void interp_z(const Grid &g, const Stc<T> &f, Stc<T> &o, int k) {
  const GA<T> dz{g.dz};
  int kp2 = imin(k + 2, g.nz), kp3 = imin(k + 3, g.nz);
  T indzm = T(1) / (dz[k] + dz[k - 1]), indzp = T(1) / (dz[k] + dz[k + 1]);
  T indzpp = T(1) / (dz[k + 1] + dz[kp2]), indzppp = T(1) / (dz[kp2] + dz[kp3]);
  o.c = (dz[k + 1] * f.c + dz[k] * f.t) * indzp;
  o.e = (dz[k + 1] * f.e + dz[k] * f.et) * indzp;
  o.t = (dz[kp2] * f.t + dz[k + 1] * f.tt) * indzpp;
  o.w = (dz[k + 1] * f.w + dz[k] * f.wt) * indzp;
  o.b = (dz[k] * f.b + dz[k - 1] * f.c) * indzm;
  o.ee = (dz[k] * f.eet + dz[k + 1] * f.ee) * indzp;
  o.tt = (dz[kp2] * f.ttt + dz[kp3] * f.tt) * indzppp;
  if (D == 3) {
    o.n = (dz[k + 1] * f.n + dz[k] * f.nt) * indzp;
    o.s = (dz[k + 1] * f.s + dz[k] * f.st) * indzp;
    o.nn = (dz[k] * f.nnt + dz[k + 1] * f.nn) * indzp;
  }
}

// face fluxes of one cell: E, N, T always; W, S, B when set
template <typename T>
struct Faces {
  T fe, fn, ft, fw, fs, fb;
  bool has_w, has_s, has_b;
};

// ----------------------------------------------------------------------
// momentum along x, flux_u.f:41-371
template <int D, typename T>
__device__ __forceinline__ void flux_u(const Grid &g, const Consts &k_, bool cw, bool cs_, bool cb,
                                       const Stc<T> &d, const Stc<T> &u, const Stc<T> &v, const Stc<T> &w,
                                       int i, int j, int k, bool muscl, Faces<T> &o) {
  const GA<T> dx{g.dx};
  T indxp = T(1) / (dx[i] + dx[i + 1]), cs;
  o.has_w = !cw; o.has_s = (D == 3) && !cs_; o.has_b = !cb;
  if (D == 2) {
    const GA<T> r{g.r};
    if (!cw) { cs = T(0.5) * (u.c + u.w); o.fw = upw(cs) * d.c * u.c * r[i] + dnw(cs) * d.w * u.w * r[i]; }
    if (!cb) { cs = (dx[i + 1] * w.b + dx[i] * w.eb) * indxp; o.fb = upw(cs) * d.c * u.c + dnw(cs) * d.b * u.b; }
    cs = T(0.5) * (u.c + u.e);
    o.fe = upw(cs) * d.e * u.e * r[i + 1] + dnw(cs) * d.c * u.c * r[i + 1];
    cs = (dx[i + 1] * w.c + dx[i] * w.e) * indxp;
    o.ft = upw(cs) * d.t * u.t + dnw(cs) * d.c * u.c;
    if (muscl) {
      const GA<T> dz{g.dz}, indx{g.indx};
      int ip2 = imin(i + 2, g.nx), kp2 = imin(k + 2, g.nz);
      T indzm = T(1) / (dz[k] + dz[k - 1]), indzp = T(1) / (dz[k] + dz[k + 1]);
      T indzpp = T(1) / (dz[k + 1] + dz[kp2]);
      T gradc = indx[i + 1] * (d.e * u.e - d.c * u.c);
      T gradw = indx[i] * (d.c * u.c - d.w * u.w);
      T grade = indx[ip2] * (d.ee * u.ee - d.e * u.e);
      cs = T(0.5) * (u.c + u.e);
      o.fe = o.fe + muscl_incr<T>(k_.lv, T(k_.beta), true, gradc, gradw, grade, cs, T(0.5) * dx[i + 1], -T(0.5) * dx[i + 1]) * cs * r[i + 1];
      gradc = (d.t * u.t - d.c * u.c) * T(2) * indzp;
      T gradb = (d.c * u.c - d.b * u.b) * T(2) * indzm;
      T gradt = (d.tt * u.tt - d.t * u.t) * T(2) * indzpp;
      cs = (dx[i + 1] * w.c + dx[i] * w.e) * indxp;
      o.ft = o.ft + muscl_incr<T>(k_.lv, T(k_.beta), true, gradc, gradb, gradt, cs, T(0.5) * dz[k], -T(0.5) * dz[k + 1]) * cs;
    }
  } else {
    if (!cw) { cs = T(0.5) * (u.c + u.w); o.fw = upw(cs) * d.c * u.c + dnw(cs) * d.w * u.w; }
    if (!cs_) { cs = (dx[i + 1] * v.s + dx[i] * v.es) * indxp; o.fs = upw(cs) * d.c * u.c + dnw(cs) * d.s * u.s; }
    if (!cb) { cs = (dx[i + 1] * w.b + dx[i] * w.eb) * indxp; o.fb = upw(cs) * d.c * u.c + dnw(cs) * d.b * u.b; }
    cs = T(0.5) * (u.c + u.e);
    o.fe = upw(cs) * d.e * u.e + dnw(cs) * d.c * u.c;
    cs = (dx[i + 1] * v.c + dx[i] * v.e) * indxp;
    o.fn = upw(cs) * d.n * u.n + dnw(cs) * d.c * u.c;
    cs = (dx[i + 1] * w.c + dx[i] * w.e) * indxp;
    o.ft = upw(cs) * d.t * u.t + dnw(cs) * d.c * u.c;
    if (muscl) {
      const GA<T> dy{g.dy}, dz{g.dz}, indx{g.indx};
      int ip2 = imin(i + 2, g.nx), jp2 = imin(j + 2, g.ny), kp2 = imin(k + 2, g.nz);
      T indym = T(1) / (dy[j] + dy[j - 1]), indyp = T(1) / (dy[j] + dy[j + 1]);
      T indypp = T(1) / (dy[j + 1] + dy[jp2]);
      T indzm = T(1) / (dz[k] + dz[k - 1]), indzp = T(1) / (dz[k] + dz[k + 1]);
      T indzpp = T(1) / (dz[k + 1] + dz[kp2]);
      T gradw = indx[i] * (u.c * d.c - u.w * d.w);
      T gradc = indx[i + 1] * (u.e * d.e - u.c * d.c);
      T grade = indx[ip2] * (u.ee * d.ee - u.e * d.e);
      cs = T(0.5) * (u.c + u.e);
      o.fe = o.fe + muscl_incr<T>(k_.lv, T(k_.beta), true, gradc, gradw, grade, cs, T(0.5) * dx[i + 1], -T(0.5) * dx[i + 1]) * cs;
      T grads = (u.c * d.c - u.s * d.s) * T(2) * indym;
      gradc = (u.n * d.n - u.c * d.c) * T(2) * indyp;
      T gradn = (u.nn * d.nn - u.n * d.n) * T(2) * indypp;
      cs = (dx[i + 1] * v.c + dx[i] * v.e) * indxp;
      o.fn = o.fn + muscl_incr<T>(k_.lv, T(k_.beta), true, gradc, grads, gradn, cs, T(0.5) * dy[j], -T(0.5) * dy[j + 1]) * cs;
      T gradb = (u.c * d.c - u.b * d.b) * T(2) * indzm;
      gradc = (u.t * d.t - u.c * d.c) * T(2) * indzp;
      T gradt = (u.tt * d.tt - u.t * d.t) * T(2) * indzpp;
      cs = (dx[i + 1] * w.c + dx[i] * w.e) * indxp;
      o.ft = o.ft + muscl_incr<T>(k_.lv, T(k_.beta), true, gradc, gradb, gradt, cs, T(0.5) * dz[k], -T(0.5) * dz[k + 1]) * cs;
    }
  }
}

// momentum along y (3D only), flux_v.f:41-222
template <typename T>
__device__ __forceinline__ void flux_v(const Grid &g, const Consts &k_, bool cw, bool cs_, bool cb,
                                       const Stc<T> &d, const Stc<T> &u, const Stc<T> &v, const Stc<T> &w,
                                       int i, int j, int k, bool muscl, Faces<T> &o) {
  const GA<T> dy{g.dy};
  T indyp = T(1) / (dy[j] + dy[j + 1]), cs;
  o.has_w = !cw; o.has_s = !cs_; o.has_b = !cb;
  if (!cw) { cs = (u.wn * dy[j] + u.w * dy[j + 1]) * indyp; o.fw = upw(cs) * d.c * v.c + dnw(cs) * d.w * v.w; }
  if (!cs_) { cs = T(0.5) * (v.c + v.s); o.fs = upw(cs) * d.c * v.c + dnw(cs) * d.s * v.s; }
  if (!cb) { cs = (w.nb * dy[j] + w.b * dy[j + 1]) * indyp; o.fb = upw(cs) * d.c * v.c + dnw(cs) * d.b * v.b; }
  cs = indyp * (u.c * dy[j + 1] + u.n * dy[j]);
  o.fe = upw(cs) * d.e * v.e + dnw(cs) * d.c * v.c;
  cs = T(0.5) * (v.n + v.c);
  o.fn = upw(cs) * d.n * v.n + dnw(cs) * d.c * v.c;
  cs = indyp * (w.c * dy[j + 1] + w.n * dy[j]);
  o.ft = upw(cs) * d.t * v.t + dnw(cs) * d.c * v.c;
  if (muscl) {
    const GA<T> dx{g.dx}, dz{g.dz}, indy{g.indy};
    int ip2 = imin(i + 2, g.nx), jp2 = imin(j + 2, g.ny), kp2 = imin(k + 2, g.nz);
    T indxm = T(1) / (dx[i] + dx[i - 1]), indxp = T(1) / (dx[i] + dx[i + 1]);
    T indxpp = T(1) / (dx[i + 1] + dx[ip2]);
    T indzm = T(1) / (dz[k] + dz[k - 1]), indzp = T(1) / (dz[k] + dz[k + 1]);
    T indzpp = T(1) / (dz[k + 1] + dz[kp2]);
    T gradw = T(2) * indxm * (v.c * d.c - v.w * d.w);
    T gradc = T(2) * indxp * (v.e * d.e - v.c * d.c);
    T grade = T(2) * indxpp * (v.ee * d.ee - v.e * d.e);
    cs = indyp * (u.c * dy[j + 1] + u.n * dy[j]);
    o.fe = o.fe + muscl_incr<T>(k_.lv, T(k_.beta), true, gradc, gradw, grade, cs, T(0.5) * dx[i], -T(0.5) * dx[i + 1]) * cs;
    T grads = (v.c * d.c - v.s * d.s) * indy[j];
    gradc = (v.n * d.n - v.c * d.c) * indy[j + 1];
    T gradn = (v.nn * d.nn - v.n * d.n) * indy[jp2];
    cs = T(0.5) * (v.n + v.c);
    o.fn = o.fn + muscl_incr<T>(k_.lv, T(k_.beta), true, gradc, grads, gradn, cs, T(0.5) * dy[j + 1], -T(0.5) * dy[j + 1]) * cs;
    T gradb = (v.c * d.c - v.b * d.b) * T(2) * indzm;
    gradc = (v.t * d.t - v.c * d.c) * T(2) * indzp;
    T gradt = (v.tt * d.tt - v.t * d.t) * T(2) * indzpp;
    cs = (dy[j + 1] * w.c + dy[j] * w.n) * indyp;
    o.ft = o.ft + muscl_incr<T>(k_.lv, T(k_.beta), true, gradc, gradb, gradt, cs, T(0.5) * dz[k], -T(0.5) * dz[k + 1]) * cs;
  }
}

// momentum along z, flux_w.f:40-356
template <int D, typename T>
__device__ __forceinline__ void flux_w(const Grid &g, const Consts &k_, bool cw, bool cs_, bool cb,
                                       const Stc<T> &d, const Stc<T> &u, const Stc<T> &v, const Stc<T> &w,
                                       int i, int j, int k, bool muscl, Faces<T> &o) {
  const GA<T> dz{g.dz};
  T indzp = T(1) / (dz[k] + dz[k + 1]), cs;
  o.has_w = !cw; o.has_s = (D == 3) && !cs_; o.has_b = !cb;
  if (D == 2) {
    const GA<T> rb{g.rb};
    if (!cw) {
      cs = (u.wt * dz[k] + u.w * dz[k + 1]) * indzp;
      o.fw = T(0.5) * d.w * w.w * (cs + fabs(cs)) * rb[i - 1] + T(0.5) * d.c * w.c * (cs - fabs(cs)) * rb[i - 1];
    }
    if (!cb) {
      cs = T(0.5) * (w.c + w.b);
      o.fb = T(0.5) * d.b * w.b * (cs + fabs(cs)) + T(0.5) * d.c * w.c * (cs - fabs(cs));
    }
    cs = indzp * (u.c * dz[k + 1] + u.t * dz[k]);
    o.fe = T(0.5) * d.c * w.c * (cs + fabs(cs)) * rb[i] + T(0.5) * d.e * w.e * (cs - fabs(cs)) * rb[i];
    cs = T(0.5) * (w.t + w.c);
    o.ft = T(0.5) * d.c * w.c * (cs + fabs(cs)) + T(0.5) * d.t * w.t * (cs - fabs(cs));
    if (muscl) {
      const GA<T> dx{g.dx}, indz{g.indz};
      int ip2 = imin(i + 2, g.nx), kp2 = imin(k + 2, g.nz);
      T indxm = T(1) / (dx[i] + dx[i - 1]), indxp = T(1) / (dx[i] + dx[i + 1]);
      T indxpp = T(1) / (dx[i + 1] + dx[ip2]);
      T gradc = T(2) * indxp * (d.e * w.e - d.c * w.c);
      T gradw = T(2) * indxm * (d.c * w.c - d.w * w.w);
      T grade = T(2) * indxpp * (d.ee * w.ee - d.e * w.e);
      cs = indzp * (u.c * dz[k + 1] + u.t * dz[k]);
      o.fe = o.fe + muscl_incr<T>(k_.lv, T(k_.beta), true, gradc, gradw, grade, cs, T(0.5) * dx[i], -T(0.5) * dx[i + 1]) * cs * rb[i];
      gradc = (d.t * w.t - d.c * w.c) * indz[k + 1];
      T gradb = (d.c * w.c - d.b * w.b) * indz[k];
      T gradt = (d.tt * w.tt - d.t * w.t) * indz[kp2];
      cs = T(0.5) * (w.t + w.c);
      o.ft = o.ft + muscl_incr<T>(k_.lv, T(k_.beta), true, gradc, gradb, gradt, cs, T(0.5) * dz[k + 1], -T(0.5) * dz[k + 1]) * cs;
    }
  } else {
    if (!cw) { cs = (dz[k + 1] * u.w + dz[k] * u.wt) * indzp; o.fw = T(0.5) * d.w * w.w * (cs + fabs(cs)) + T(0.5) * d.c * w.c * (cs - fabs(cs)); }
    if (!cs_) { cs = (dz[k + 1] * v.s + dz[k] * v.st) * indzp; o.fs = T(0.5) * d.s * w.s * (cs + fabs(cs)) + T(0.5) * d.c * w.c * (cs - fabs(cs)); }
    if (!cb) { cs = T(0.5) * (w.b + w.c); o.fb = T(0.5) * d.b * w.b * (cs + fabs(cs)) + T(0.5) * d.c * w.c * (cs - fabs(cs)); }
    cs = (dz[k + 1] * u.c + dz[k] * u.t) * indzp;
    o.fe = T(0.5) * d.c * w.c * (cs + fabs(cs)) + T(0.5) * d.e * w.e * (cs - fabs(cs));
    cs = (dz[k + 1] * v.c + dz[k] * v.t) * indzp;
    o.fn = T(0.5) * d.c * w.c * (cs + fabs(cs)) + T(0.5) * d.n * w.n * (cs - fabs(cs));
    cs = T(0.5) * (w.c + w.t);
    o.ft = T(0.5) * d.c * w.c * (cs + fabs(cs)) + T(0.5) * d.t * w.t * (cs - fabs(cs));
    if (muscl) {
      const GA<T> dx{g.dx}, dy{g.dy}, indz{g.indz};
      int ip2 = imin(i + 2, g.nx), jp2 = imin(j + 2, g.ny), kp2 = imin(k + 2, g.nz);
      T indxm = T(1) / (dx[i] + dx[i - 1]), indxp = T(1) / (dx[i] + dx[i + 1]);
      T indxpp = T(1) / (dx[i + 1] + dx[ip2]);
      T indym = T(1) / (dy[j] + dy[j - 1]), indyp = T(1) / (dy[j] + dy[j + 1]);
      T indypp = T(1) / (dy[j + 1] + dy[jp2]);
      T gradw = T(2) * indxm * (w.c * d.c - w.w * d.w);
      T gradc = T(2) * indxp * (w.e * d.e - w.c * d.c);
      T grade = T(2) * indxpp * (w.ee * d.ee - w.e * d.e);
      cs = (dz[k + 1] * u.c + dz[k] * u.t) * indzp;
      o.fe = o.fe + muscl_incr<T>(k_.lv, T(k_.beta), true, gradc, gradw, grade, cs, T(0.5) * dx[i], -T(0.5) * dx[i + 1]) * cs;
      T grads = T(2) * indym * (w.c * d.c - w.s * d.s);
      gradc = T(2) * indyp * (w.n * d.n - w.c * d.c);
      T gradn = T(2) * indypp * (w.nn * d.nn - w.n * d.n);
      cs = (dz[k + 1] * v.c + dz[k] * v.t) * indzp;
      o.fn = o.fn + muscl_incr<T>(k_.lv, T(k_.beta), true, gradc, grads, gradn, cs, T(0.5) * dy[j], -T(0.5) * dy[j + 1]) * cs;
      T gradb = indz[k] * (w.c * d.c - w.b * d.b);
      gradc = indz[k + 1] * (w.t * d.t - w.c * d.c);
      T gradt = indz[kp2] * (w.tt * d.tt - w.t * d.t);
      cs = T(0.5) * (w.c + w.t);
      o.ft = o.ft + muscl_incr<T>(k_.lv, T(k_.beta), true, gradc, gradb, gradt, cs, T(0.5) * dz[k + 1], -T(0.5) * dz[k + 1]) * cs;
    }
  }
}

}  // namespace pdac
