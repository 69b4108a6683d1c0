// Per-ray building blocks of the range march, compiled for the device (and
// usable on the host): complex helpers, profile interpolation, medium
// selection, range update and output.
//
// Reference lines restated:
//   medium_at     cupyram/profl.py:45-140 (depth interpolation with clamping,
//                 absorbing layer), cupyram/cupyram.py:782-803 (alpw, alpb,
//                 ksqw, ksqb) and cupyram/matrc.py:63-107 (water rows 0..iz,
//                 bottom rows iz+1..nz+1 -> f1, f2, f3, ksq)
//   galerkin      cupyram/fused_kernel.py:20-69 (tridiagonal r1..r3, s1..s3)
//   updat_ray     cupyram/updat.py:38-140, cupyram/cupyram.py:974-999
//   tl_of         cupyram/outpt.py:39-53
#pragma once
#include <cmath>

#include "cpr_internal.h"

template <class R>
struct alignas(2 * sizeof(R)) cx {
  R re, im;
};

template <class R>
CPR_HD inline cx<R> cxm(R a, R b) {
  cx<R> r;
  r.re = a;
  r.im = b;
  return r;
}
template <class R>
CPR_HD inline cx<R> operator+(cx<R> a, cx<R> b) {
  return cxm<R>(a.re + b.re, a.im + b.im);
}
template <class R>
CPR_HD inline cx<R> operator-(cx<R> a, cx<R> b) {
  return cxm<R>(a.re - b.re, a.im - b.im);
}
template <class R>
CPR_HD inline cx<R> operator*(cx<R> a, cx<R> b) {
  return cxm<R>(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
template <class R>
CPR_HD inline cx<R> operator*(R s, cx<R> a) {
  return cxm<R>(s * a.re, s * a.im);
}
// a - b*c
template <class R>
CPR_HD inline cx<R> cfms(cx<R> a, cx<R> b, cx<R> c) {
  return cxm<R>(a.re - b.re * c.re + b.im * c.im, a.im - b.re * c.im - b.im * c.re);
}
// 1/a as conj(a)/|a|^2
template <class R>
CPR_HD inline cx<R> crcp(cx<R> a) {
  const R m = R(1) / (a.re * a.re + a.im * a.im);
  return cxm<R>(a.re * m, -a.im * m);
}
template <class R>
CPR_HD inline R cabs2(cx<R> a) {
  return a.re * a.re + a.im * a.im;
}

// numpy.interp-style clamped linear interpolation: the last knot j with
// x[j] <= z (a duplicated knot resolves to the right segment).
//
// `cursor` (optional) carries the segment from one call to the next for
// nondecreasing z (the rows of a column, top down): the segment is then found
// by walking forward from the previous one instead of a binary search, so a
// column's rows cost one load per knot passed instead of a chain of
// log2(n) dependent loads per row.  Both ways find the same segment, so the
// results are identical.  *cursor < 0 means "not known yet".
CPR_HD inline double interp_clamped(const double *x, const double *y, int n, int ystride, double z,
                                    int *cursor = nullptr) {
  if (z <= x[0]) return y[0];
  if (z >= x[n - 1]) return y[(n - 1) * ystride];
  int lo;
  if (cursor && *cursor >= 0) {
    lo = *cursor;  // x[lo] <= (an earlier z) <= z
    while (lo + 1 < n - 1 && x[lo + 1] <= z) ++lo;
  } else {
    lo = 0;
    int hi = n - 1;
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (x[mid] <= z)
        lo = mid;
      else
        hi = mid;
    }
  }
  if (cursor) *cursor = lo;
  const double x0 = x[lo], x1 = x[lo + 1];
  const double y0 = y[lo * ystride], y1 = y[(lo + 1) * ystride];
  if (x1 == x0) return y0;
  return y0 + (y1 - y0) * ((z - x0) / (x1 - x0));
}

struct MediumRow {
  double f1, f2, f3;
  double kre, kim;
};

constexpr double kEta = 0.018323389971985696;  // 1/(40 pi log10(e))
constexpr double kAttnFloor = 10.0;            // dB/wavelength at the grid bottom

CPR_HD inline double grid_depth(const cpr_ray &P, int row) {
  return row == P.nz + 1 ? P.zmax : static_cast<double>(row) * P.zstep;
}

CPR_HD inline double water_speed(const cpr_env &E, const double *env, int ss, double z, int *cursor = nullptr) {
  return interp_clamped(env + E.z_ss, env + E.cw + ss, E.n_ss, E.n_rp_ss, z, cursor);
}

CPR_HD inline double alpw_at(const cpr_ray &P, const cpr_env &E, const double *env, int ss, double z) {
  return sqrt(water_speed(E, env, ss, z) / P.c0);
}

// Water medium of one grid row (FP64): f1 = 1/alpw, f2 = 1, f3 = alpw,
// ksq = (omega/cw)^2 - k0^2.  `cursor`: see interp_clamped (rows in
// increasing order).
CPR_HD inline MediumRow water_medium_at(const cpr_ray &P, const cpr_env &E, const double *env, int ss, int row,
                                        int *cursor = nullptr) {
  const double z = grid_depth(P, row);
  MediumRow m;
  const double cw = water_speed(E, env, ss, z, cursor);
  const double a = sqrt(cw / P.c0);
  m.f1 = 1.0 / a;
  m.f2 = 1.0;
  m.f3 = a;
  const double ow = P.omega / cw;
  m.kre = ow * ow - P.k0 * P.k0;
  m.kim = 0.0;
  return m;
}

// Bottom medium of one grid row (FP64): f1 = rhob/alpb, f2 = 1/rhob,
// f3 = alpb, ksq = ((omega/cb)(1 + i eta attn))^2 - k0^2 with the absorbing
// layer ramp below the last seabed depth.
CPR_HD inline MediumRow bottom_medium_at(const cpr_ray &P, const cpr_env &E, const double *env, int sb, int row,
                                         int *cursor = nullptr) {
  const double z = grid_depth(P, row);
  MediumRow m;
  const double k0sq = P.k0 * P.k0;
  {
    const double *zsb = env + E.z_sb;
    const int nsb = E.n_sb, st = E.n_rp_sb;
    const double cb = interp_clamped(zsb, env + E.cb + sb, nsb, st, z, cursor);
    const double rh = interp_clamped(zsb, env + E.rhob + sb, nsb, st, z, cursor);
    const double zl = zsb[nsb - 1];
    const double at_last = env[E.attn + sb + (nsb - 1) * st];
    const double z1 = zl + 0.75 * P.lyrw_lambda, z2 = zl + P.lyrw_lambda;
    double at;
    if (z <= zl)
      at = interp_clamped(zsb, env + E.attn + sb, nsb, st, z, cursor);
    else if (z <= z1)
      at = at_last;
    else if (z <= z2)
      at = at_last + (kAttnFloor - at_last) * ((z - z1) / (z2 - z1));
    else
      at = kAttnFloor;
    const double a = sqrt(rh * cb / P.c0);
    m.f1 = rh / a;
    m.f2 = 1.0 / rh;
    m.f3 = a;
    const double tr = P.omega / cb;
    const double ti = tr * (kEta * at);
    m.kre = (tr * tr - ti * ti) - k0sq;
    m.kim = 2.0 * tr * ti;
  }
  return m;
}

// Medium of one grid row for water/bottom selection by iz (FP64).
CPR_HD inline MediumRow medium_at(const cpr_ray &P, const cpr_env &E, const double *env, int ss, int sb, int iz,
                                  int row, int *wcur = nullptr, int *bcur = nullptr) {
  return row <= iz ? water_medium_at(P, E, env, ss, row, wcur) : bottom_medium_at(P, E, env, sb, row, bcur);
}

// Range update of one ray before a march step at range r.  Returns a bit
// mask: 1 = iz changed, 2 = a profile column changed, 4 = Pade set switched.
CPR_HD inline int updat_ray(const cpr_ray &P, const cpr_env &E, const double *env, RayState &st, double r) {
  int flags = 0;
  if (E.n_bt >= 2) {
    const double *bt = env + E.rbzb;
    while (st.bt_ind < E.n_bt - 1 && r >= bt[2 * (st.bt_ind + 1)]) st.bt_ind++;
    const int b0 = st.bt_ind < E.n_bt - 1 ? st.bt_ind : E.n_bt - 2;
    const double r0 = bt[2 * b0], z0 = bt[2 * b0 + 1];
    const double r1 = bt[2 * b0 + 2], z1 = bt[2 * b0 + 3];
    const double z = z0 + ((r + 0.5 * P.dr) - r0) * (z1 - z0) / (r1 - r0);
    int iz = static_cast<int>(floor(z / P.dz));
    iz = iz < 1 ? 1 : iz;
    iz = iz > P.nz - 1 ? P.nz - 1 : iz;
    if (iz != st.iz) flags |= 1;
    st.iz = iz;
  }
  if (E.n_rp_ss >= 2) {
    const double *rp = env + E.rp_ss;
    const int o = st.ss_ind;
    while (st.ss_ind < E.n_rp_ss - 1 && r >= rp[st.ss_ind + 1]) st.ss_ind++;
    if (st.ss_ind != o) flags |= 2;
  }
  if (E.n_rp_sb >= 2) {
    const double *rp = env + E.rp_sb;
    const int o = st.sb_ind;
    while (st.sb_ind < E.n_rp_sb - 1 && r >= rp[st.sb_ind + 1]) st.sb_ind++;
    if (st.sb_ind != o) flags |= 2;
  }
  if (st.pset == 0 && r >= P.rs) {
    st.pset = 2;
    flags |= 4;
  }
  return flags;
}

// Transmission loss of a complex pressure at range r (outpt.py:43-45).
CPR_HD inline double tl_of(double pre, double pim, double r) {
  return -20.0 * log10(sqrt(pre * pre + pim * pim) + 1e-20) + 10.0 * log10(r + 1e-20);
}

// Galerkin coefficients of row i (fused_kernel.py:43-58) from the medium of
// rows i-1, i, i+1: the real stencil c1..c3 plus the wavenumber stencil.
template <class R>
struct Galerkin {
  cx<R> d1, d2, d3;
};

template <class R>
CPR_HD inline Galerkin<R> galerkin(R cfact, R f1c, R f2m, R f2c, R f2p, R f3m, R f3c, R f3p, cx<R> km, cx<R> kc,
                                   cx<R> kp) {
  const R dfact = R(1) / R(12);
  const R c1 = cfact * f1c * (f2m + f2c) * f3m;
  const R c2 = -cfact * f1c * (f2m + R(2) * f2c + f2p) * f3c;
  const R c3 = cfact * f1c * (f2c + f2p) * f3p;
  Galerkin<R> g;
  g.d1 = cxm<R>(c1 + dfact * (km.re + kc.re), dfact * (km.im + kc.im));
  g.d2 = cxm<R>(c2 + dfact * (km.re + R(6) * kc.re + kp.re), dfact * (km.im + R(6) * kc.im + kp.im));
  g.d3 = cxm<R>(c3 + dfact * (kc.re + kp.re), dfact * (kc.im + kp.im));
  return g;
}
