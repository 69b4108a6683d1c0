/*
 * ram_oracle.c -- CPU restatement of the PyRAM / RAM v1.5 split-step Pade
 * range march, used ONLY as test infrastructure (the checker for the HIP
 * march, the golden-vector pin, and the CPU baseline timed by bench.py).
 * Nothing in the shipped library links, loads or calls this file.
 *
 * Precision: complex128 arithmetic and FP64 storage everywhere (C99
 * `double complex`).  One ray per call; orc_run_batch() runs independent rays
 * on OpenMP threads for the CPU baseline.
 *
 * Each function names the reference code it restates (paths relative to the
 * reference repository root):
 *   pade_*        cupyram/pade.py:16-273 (PyRAM.epade and helpers)
 *   ray_profl     cupyram/cupyram.py:728-803 + cupyram/profl.py:8-140
 *   ray_matrc     cupyram/matrc.py:16-107 (medium selection) and
 *                 cupyram/matrc.py:110-234 (Galerkin assembly + PyRAM's
 *                 two-sided decomposition toward row iz+1)
 *   ray_solve     cupyram/solve.py:48-95 (PyRAM solve, eps = 1e-30 adds)
 *   ray_updat     cupyram/updat.py:38-140 + cupyram/cupyram.py:951-999
 *   ray_selfs     cupyram/cupyram.py:1001-1063 (PyRAM self-starter: the
 *                 (1-X)^-1 smoothing solve is applied with np=1, as PyRAM
 *                 does; the reference's fused kernel also sweeps the np-1
 *                 all-zero rows, which are identity solves up to rounding)
 *   ray_outpt     cupyram/outpt.py:11-58 (TL line + TL grid; the CP grid is
 *                 written too -- the reference's kernel drops it, outpt.py:50)
 *   orc_run       cupyram/cupyram.py:273-380 + setup :528-726
 *   orc_run_cells the same march with cell-mean TL on the output grid (no
 *                 reference counterpart: the coarse TL of the surrogate)
 *
 * Parity pins (tests/test_oracle.py): RAM v1.5's tl_ref.line (mean |dTL| <=
 * 0.01 dB), Pade closed forms (Taylor series by series composition, stability
 * zeros), numpy.interp for profiles, brute-force bathymetry indices,
 * identities of the factor solve (identity map, inverse factor, linearity,
 * independence of the meeting row), and a textbook Pekeris normal-mode sum.
 * Cell means are pinned against the TL grid of the same march: with
 * ndr = ndz = 1 every cell is one grid point, and a coarse cell equals the
 * intensity mean of the fine grid points it covers.
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "ram_oracle.h"

typedef double complex cplx;

/* ------------------------------------------------------------------------ */
/* Pade coefficients (pade.py:180-273)                                      */
/* ------------------------------------------------------------------------ */

/* Complex division by Smith's method (R. L. Smith, CACM 5(8), 1962), written
 * out so the result does not depend on which __divdc3 the process binds. */
static cplx cdiv(cplx a, cplx b) {
  const double ar = creal(a), ai = cimag(a), br = creal(b), bi = cimag(b);
  if (fabs(br) >= fabs(bi)) {
    const double t = bi / br, d = br + bi * t;
    return CMPLX((ar + ai * t) / d, (ai - ar * t) / d);
  }
  const double t = br / bi, d = br * t + bi;
  return CMPLX((ar * t + ai) / d, (ai * t - ar) / d);
}

/* Taylor derivatives of g(x) = exp(i sig (sqrt(1+x)-1)) (1+x)^alp (1-x)^(2nu)
 * at x = 0, via g' = g (h1' + h2' + h3') and Leibniz (pade.py:16-43). */
static void pade_derivs(int n, double sig, double alp, int nu,
                        const double *bn /* (n+1)^2 */, cplx *dg /* n+1 */) {
  cplx dh[64];
  cplx h1 = 0.5 * I * sig, h2 = alp, h3 = -2.0 * nu;
  double e1 = -0.5, e2 = -1.0, e3 = -1.0;
  dh[0] = h1 + h2 + h3;
  for (int i = 1; i < n; ++i) {
    h1 = h1 * e1;
    e1 -= 1.0;
    h2 = h2 * e2;
    e2 -= 1.0;
    h3 = -(double)nu * h3 * e3;
    e3 -= 1.0;
    dh[i] = h1 + h2 + h3;
  }
  dg[0] = 1.0;
  dg[1] = dh[0];
  for (int i = 1; i < n; ++i) {
    cplx acc = dh[i];
    for (int j = 0; j < i; ++j) acc += bn[i * (n + 1) + j] * dh[j] * dg[i - j];
    dg[i + 1] = acc;
  }
}

/* Gaussian elimination with partial pivoting on an n x n system stored with
 * row stride n+1 (pade.py:46-91). */
static void pade_gauss(int n, cplx *a, cplx *b) {
  const int ld = n + 1;
  for (int i = 0; i < n; ++i) {
    if (i < n - 1) {
      int ip = i;
      double amp0 = cabs(a[i * ld + i]);
      for (int j = i + 1; j < n; ++j) {
        double amp = cabs(a[j * ld + i]);
        if (amp > amp0) {
          ip = j;
          amp0 = amp;
        }
      }
      if (ip != i) {
        cplx t = b[ip];
        b[ip] = b[i];
        b[i] = t;
        for (int j = i; j <= n; ++j) {
          t = a[ip * ld + j];
          a[ip * ld + j] = a[i * ld + j];
          a[i * ld + j] = t;
        }
      }
    }
    a[i * ld + i] = cdiv(1.0, a[i * ld + i]);
    b[i] *= a[i * ld + i];
    if (i < n - 1) {
      for (int j = i + 1; j <= n; ++j) a[i * ld + j] *= a[i * ld + i];
      for (int k = i + 1; k < n; ++k) {
        b[k] -= a[k * ld + i] * b[i];
        for (int j = i + 1; j < n; ++j) a[k * ld + j] -= a[k * ld + i] * a[i * ld + j];
      }
    }
  }
  for (int i = n - 2; i >= 0; --i)
    for (int j = i + 1; j < n; ++j) b[i] -= a[i * ld + j] * b[j];
}

/* Laguerre iteration on the degree-n polynomial a[0..n] (pade.py:94-149).
 * Returns 0 on success, -1 if 100 iterations pass without convergence. */
static int pade_laguerre(const cplx *a, int n, cplx *zio, double err, int nter) {
  cplx az[64], azz[64];
  cplx z = *zio;
  for (int i = 0; i < n; ++i) az[i] = (double)(i + 1) * a[i + 1];
  for (int i = 0; i < n - 1; ++i) azz[i] = (double)(i + 1) * az[i + 1];
  int iter = 0, jter = 0;
  double adz = INFINITY;
  while (adz > err && iter < nter - 1) {
    cplx p = a[n - 1] + a[n] * z;
    for (int i = n - 2; i >= 0; --i) p = a[i] + z * p;
    if (cabs(p) < 1e-20) break;
    cplx pz = az[n - 2] + az[n - 1] * z;
    for (int i = n - 3; i >= 0; --i) pz = az[i] + z * pz;
    cplx pzz = azz[n - 3] + azz[n - 2] * z;
    for (int i = n - 4; i >= 0; --i) pzz = azz[i] + z * pzz;
    cplx f = cdiv(pz, p);
    cplx g = f * f - cdiv(pzz, p);
    cplx h = csqrt((double)(n - 1) * ((double)n * g - f * f));
    cplx dz = (cabs(f + h) > cabs(f - h)) ? cdiv(-(double)n, f + h) : cdiv(-(double)n, f - h);
    ++iter;
    if (++jter == 9) {
      jter = 0;
      dz *= I;
    }
    z += dz;
    adz = cabs(dz);
    if (iter == 100) return -1;
  }
  *zio = z;
  return 0;
}

/* All roots of a[0..n] by Laguerre + deflation, last two by the quadratic
 * formula (pade.py:152-177).  a is destroyed. */
static int pade_roots(cplx *a, int n, cplx *z) {
  if (n == 1) {
    z[0] = cdiv(-a[0], a[1]);
    return 0;
  }
  if (n != 2) {
    for (int k = n - 1; k >= 2; --k) {
      cplx root = 0.0;
      if (pade_laguerre(a, k + 1, &root, 1e-12, 1000)) return -1;
      if (pade_laguerre(a, k + 1, &root, 0.0, 5)) return -1;
      z[k] = root;
      for (int i = k; i >= 0; --i) a[i] += root * a[i + 1];
      for (int i = 0; i <= k; ++i) a[i] = a[i + 1];
    }
  }
  cplx disc = csqrt(a[1] * a[1] - 4.0 * a[0] * a[2]);
  z[1] = cdiv(0.5 * (-a[1] + disc), a[2]);
  z[0] = cdiv(0.5 * (-a[1] - disc), a[2]);
  return 0;
}

int orc_pade(double freq, double c0, int np_, int ns, double dr, int ip,
             double complex *pd1, double complex *pd2) {
  if (np_ < 1 || np_ > 16) return -2;
  const int n = 2 * np_;
  double bn[(33) * (33)];
  double fact[33];
  cplx a[33 * 33], b[33], dg[33], poly[33], roots[33];
  memset(bn, 0, sizeof(bn));
  for (int i = 0; i < (n + 1) * (n + 1); ++i) a[i] = 0.0;
  const double k0 = 2.0 * M_PI * freq / c0;
  const double sig = k0 * dr;
  const int nu = (ip == 1) ? 0 : 1;
  const double alp = (ip == 1) ? 0.0 : -0.25;
  fact[0] = 1.0;
  for (int i = 1; i < n; ++i) fact[i] = (double)(i + 1) * fact[i - 1];
  for (int i = 0; i <= n; ++i) {
    bn[i * (n + 1) + 0] = 1.0;
    bn[i * (n + 1) + i] = 1.0;
  }
  for (int i = 2; i <= n; ++i)
    for (int j = 1; j < i; ++j) bn[i * (n + 1) + j] = bn[(i - 1) * (n + 1) + j - 1] + bn[(i - 1) * (n + 1) + j];
  pade_derivs(n, sig, alp, nu, bn, dg);
  /* accuracy constraints: numerator - g * denominator = O(x^(n+1)) */
  for (int i = 0; i < n; ++i) b[i] = dg[i + 1];
  for (int i = 0; i < n; ++i) {
    if (2 * i <= n - 1) a[i * (n + 1) + 2 * i] = fact[i];
    for (int j = 0; j <= i; ++j)
      if (2 * j + 1 <= n - 1) a[i * (n + 1) + 2 * j + 1] = -bn[(i + 1) * (n + 1) + j + 1] * fact[j] * dg[i - j];
  }
  /* stability constraints: numerator vanishes at X = -3 (ns>=1), -1.5 (ns>=2) */
  if (ns >= 1) {
    double zp = 1.0;
    b[n - 1] = -1.0;
    for (int j = 0; j < np_; ++j) {
      zp *= -3.0;
      a[(n - 1) * (n + 1) + 2 * j] = zp;
      a[(n - 1) * (n + 1) + 2 * j + 1] = 0.0;
    }
  }
  if (ns >= 2) {
    double zp = 1.0;
    b[n - 2] = -1.0;
    for (int j = 0; j < np_; ++j) {
      zp *= -1.5;
      a[(n - 2) * (n + 1) + 2 * j] = zp;
      a[(n - 2) * (n + 1) + 2 * j + 1] = 0.0;
    }
  }
  pade_gauss(n, a, b);
  poly[0] = 1.0;
  for (int j = 0; j < np_; ++j) poly[j + 1] = b[2 * j];
  if (pade_roots(poly, np_, roots)) return -1;
  for (int j = 0; j < np_; ++j) pd1[j] = cdiv(-1.0, roots[j]);
  poly[0] = 1.0;
  for (int j = 0; j < np_; ++j) poly[j + 1] = b[2 * j + 1];
  if (pade_roots(poly, np_, roots)) return -1;
  for (int j = 0; j < np_; ++j) pd2[j] = cdiv(-1.0, roots[j]);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* The march for one ray                                                     */
/* ------------------------------------------------------------------------ */

typedef struct {
  const orc_ray *in;
  int nz, nzplt, nvr, nvz, nr, np_;
  double k0, omega, eta, rmax, r, zmax, dz, dr;
  int iz, ir, bt_ind, ss_ind, sb_ind, ns, mdr, tlc;
  double dir, rs;
  int n_bt;
  double *rbzb;   /* extended bathymetry, n_bt x 2 */
  double *z_sb;   /* absolute seabed depths */
  double *alpw, *alpb, *rhob, *ksqw_re;
  cplx *ksqw, *ksqb;
  double *f1, *f2, *f3;
  cplx *ksq, *u, *v;
  cplx *r1, *r2, *r3, *s1, *s2, *s3; /* [(nz+2) * np] row-major by row */
  cplx pd1[16], pd2[16];
  /* outputs */
  double *tll, *tlg;
  cplx *cpl, *cpg;
  double *cells; /* [nvz][nvr] sums of |f3 u|^2 / r per output cell, or NULL */
  int *iz_hist;
} ray_state;

/* numpy.interp semantics with clamping: last knot j with x[j] <= z; a
 * duplicate knot resolves to the right segment (profl.py:8-43). */
static double interp_clamped(double z, const double *x, const double *y, int n, int ystride) {
  if (z <= x[0]) return y[0];
  if (z >= x[n - 1]) return y[(n - 1) * ystride];
  int lo = 0, hi = n - 1; /* invariant x[lo] <= z < x[hi] */
  while (hi - lo > 1) {
    int mid = (lo + hi) / 2;
    if (x[mid] <= z)
      lo = mid;
    else
      hi = mid;
  }
  double x0 = x[lo], x1 = x[lo + 1];
  double y0 = y[lo * ystride], y1 = y[(lo + 1) * ystride];
  if (x1 == x0) return y0;
  return y0 + (y1 - y0) * ((z - x0) / (x1 - x0));
}

/* Depth profiles on the z = linspace(0, zmax, nz+2) grid
 * (cupyram.py:728-803, profl.py:45-140). */
static void ray_profl(ray_state *s) {
  const orc_ray *in = s->in;
  const int nz2 = s->nz + 2;
  const double step = s->zmax / (double)(s->nz + 1);
  const double *cwcol = in->cw + s->ss_ind;
  const double *cbcol = in->cb + s->sb_ind;
  const double *rhcol = in->rhob + s->sb_ind;
  const double *atcol = in->attn + s->sb_ind;
  const int nsb = in->n_sb;
  const double zlast = s->z_sb[nsb - 1];
  const double attn_last = atcol[(nsb - 1) * in->n_rp_sb];
  const double lw = in->lyrw * in->lambda;
  const double zl1 = zlast + 0.75 * lw, zl2 = zlast + lw;
  const double attnf = 10.0;
  for (int i = 0; i < nz2; ++i) {
    const double z = (i == nz2 - 1) ? s->zmax : (double)i * step;
    const double cw = interp_clamped(z, in->z_ss, cwcol, in->n_ss, in->n_rp_ss);
    const double cb = interp_clamped(z, s->z_sb, cbcol, nsb, in->n_rp_sb);
    const double rh = interp_clamped(z, s->z_sb, rhcol, nsb, in->n_rp_sb);
    double at;
    if (z <= zlast)
      at = interp_clamped(z, s->z_sb, atcol, nsb, in->n_rp_sb);
    else if (z <= zl1)
      at = attn_last;
    else if (z <= zl2)
      at = attn_last + (attnf - attn_last) * ((z - zl1) / (zl2 - zl1));
    else
      at = attnf;
    s->alpw[i] = sqrt(cw / in->c0);
    s->alpb[i] = sqrt(rh * cb / in->c0);
    s->rhob[i] = rh;
    const double ow = s->omega / cw;
    s->ksqw[i] = ow * ow - s->k0 * s->k0;
    const cplx tb = (s->omega / cb) * (1.0 + I * s->eta * at);
    s->ksqb[i] = tb * tb - s->k0 * s->k0;
  }
}

/* Medium selection by iz and Galerkin assembly + PyRAM decomposition for all
 * np terms (matrc.py:16-107, 110-234; the full-column iz==jz form, which the
 * incremental iz!=jz form reproduces bitwise). */
static void ray_matrc(ray_state *s, int np_) {
  const int nz = s->nz, iz = s->iz;
  for (int i = 0; i < nz + 2; ++i) {
    if (i <= iz) {
      s->f1[i] = 1.0 / s->alpw[i];
      s->f2[i] = 1.0;
      s->f3[i] = s->alpw[i];
      s->ksq[i] = s->ksqw[i];
    } else {
      s->f1[i] = s->rhob[i] / s->alpb[i];
      s->f2[i] = 1.0 / s->rhob[i];
      s->f3[i] = s->alpb[i];
      s->ksq[i] = s->ksqb[i];
    }
  }
  const double a1 = s->k0 * s->k0 / 6.0, a2 = 2.0 * s->k0 * s->k0 / 3.0, a3 = a1;
  const double cfact = 0.5 / (s->dz * s->dz), dfact = 1.0 / 12.0;
  for (int j = 0; j < np_; ++j) {
    /* rows 0 and nz+1 stay zero */
    s->r1[0 * np_ + j] = s->r2[0 * np_ + j] = s->r3[0 * np_ + j] = 0;
    s->s1[0 * np_ + j] = s->s2[0 * np_ + j] = s->s3[0 * np_ + j] = 0;
    s->r1[(nz + 1) * np_ + j] = s->r2[(nz + 1) * np_ + j] = s->r3[(nz + 1) * np_ + j] = 0;
    s->s1[(nz + 1) * np_ + j] = s->s2[(nz + 1) * np_ + j] = s->s3[(nz + 1) * np_ + j] = 0;
  }
  for (int i = 1; i <= nz; ++i) {
    const double c1 = cfact * s->f1[i] * (s->f2[i - 1] + s->f2[i]) * s->f3[i - 1];
    const double c2 = -cfact * s->f1[i] * (s->f2[i - 1] + 2.0 * s->f2[i] + s->f2[i + 1]) * s->f3[i];
    const double c3 = cfact * s->f1[i] * (s->f2[i] + s->f2[i + 1]) * s->f3[i + 1];
    const cplx d1 = c1 + dfact * (s->ksq[i - 1] + s->ksq[i]);
    const cplx d2 = c2 + dfact * (s->ksq[i - 1] + 6.0 * s->ksq[i] + s->ksq[i + 1]);
    const cplx d3 = c3 + dfact * (s->ksq[i] + s->ksq[i + 1]);
    for (int j = 0; j < np_; ++j) {
      s->r1[i * np_ + j] = a1 + s->pd2[j] * d1;
      s->r2[i * np_ + j] = a2 + s->pd2[j] * d2;
      s->r3[i * np_ + j] = a3 + s->pd2[j] * d3;
      s->s1[i * np_ + j] = a1 + s->pd1[j] * d1;
      s->s2[i * np_ + j] = a2 + s->pd1[j] * d2;
      s->s3[i * np_ + j] = a3 + s->pd1[j] * d3;
    }
  }
  /* decomposition: normalise rows 1..iz from the top, nz..iz+2 from the
   * bottom, pivot at iz+1 (matrc.py:212-234) */
  for (int j = 0; j < np_; ++j) {
#define R(a, i) s->a[(i) * np_ + j]
    for (int i = 1; i <= iz; ++i) {
      const cplx rf = 1.0 / (R(r2, i) - R(r1, i) * R(r3, i - 1));
      R(r1, i) *= rf;
      R(r3, i) *= rf;
      R(s1, i) *= rf;
      R(s2, i) *= rf;
      R(s3, i) *= rf;
    }
    for (int i = nz; i >= iz + 2; --i) {
      const cplx rf = 1.0 / (R(r2, i) - R(r3, i) * R(r1, i + 1));
      R(r1, i) *= rf;
      R(r3, i) *= rf;
      R(s1, i) *= rf;
      R(s2, i) *= rf;
      R(s3, i) *= rf;
    }
    cplx p = R(r2, iz + 1);
    p -= R(r1, iz + 1) * R(r3, iz);
    p -= R(r3, iz + 1) * R(r1, iz + 2);
    R(r2, iz + 1) = 1.0 / p;
#undef R
  }
}

/* PyRAM solve: np sequential rational factors (solve.py:48-95). */
static void ray_solve(ray_state *s, int np_) {
  const int nz = s->nz, iz = s->iz;
  const cplx eps = 1e-30;
  cplx *u = s->u, *v = s->v;
  for (int j = 0; j < np_; ++j) {
#define R(a, i) s->a[(i) * np_ + j]
    for (int i = 1; i <= nz; ++i) v[i] = R(s1, i) * u[i - 1] + R(s2, i) * u[i] + R(s3, i) * u[i + 1] + eps;
    for (int i = 2; i <= iz; ++i) v[i] -= R(r1, i) * v[i - 1] + eps;
    for (int i = nz - 1; i >= iz + 2; --i) v[i] -= R(r3, i) * v[i + 1] + eps;
    u[iz + 1] = (v[iz + 1] - R(r1, iz + 1) * v[iz] - R(r3, iz + 1) * v[iz + 2]) * R(r2, iz + 1) + eps;
    for (int i = iz; i >= 0; --i) u[i] = v[i] - R(r3, i) * u[i + 1] + eps;
    for (int i = iz + 2; i <= nz; ++i) u[i] = v[i] - R(r1, i) * u[i - 1] + eps;
#undef R
  }
}

/* TL line at zr (linear in depth between ir and ir+1) and TL/CP grid every
 * ndz rows (outpt.py:11-58). */
static void ray_outpt(ray_state *s) {
  const orc_ray *in = s->in;
  const double eps = 1e-20;
  /* cell means: the field at range r belongs to range bin tlc+1 (the ndr
   * fields ending at the next output); depth bin i holds rows
   * i*ndz+1 .. (i+1)*ndz */
  if (s->cells && s->tlc + 1 < s->nvr) {
    for (int i = 0; i < s->nvz; ++i) {
      double acc = 0.0;
      for (int g = i * in->ndz + 1; g <= (i + 1) * in->ndz; ++g) {
        const cplx p = s->u[g] * s->f3[g];
        acc += creal(p) * creal(p) + cimag(p) * cimag(p);
      }
      s->cells[(size_t)i * s->nvr + s->tlc + 1] += acc / s->r;
    }
  }
  s->mdr += 1;
  if (s->mdr != in->ndr) return;
  s->mdr = 0;
  s->tlc += 1;
  if (s->tlc >= s->nvr) return;
  const int ir = s->ir;
  const cplx ur = (1.0 - s->dir) * s->f3[ir] * s->u[ir] + s->dir * s->f3[ir + 1] * s->u[ir + 1];
  const double temp = 10.0 * log10(s->r + eps);
  s->tll[s->tlc] = -20.0 * log10(cabs(ur) + eps) + temp;
  if (s->cpl) s->cpl[s->tlc] = ur;
  if (s->tlg) {
    for (int i = 0; i < s->nvz; ++i) {
      const int j = (i + 1) * in->ndz;
      const cplx ug = s->u[j] * s->f3[j];
      s->tlg[(size_t)i * s->nvr + s->tlc] = -20.0 * log10(cabs(ug) + eps) + temp;
      if (s->cpg) s->cpg[(size_t)i * s->nvr + s->tlc] = ug;
    }
  }
}

/* Self-starter (cupyram.py:1001-1063). */
static int ray_selfs(ray_state *s) {
  const orc_ray *in = s->in;
  const double si = in->zs / s->dz;
  const int is = (int)floor(si);
  const double dis = si - is;
  const double amp = sqrt(2.0 * M_PI / s->k0);
  s->u[is] = (1.0 - dis) * amp / (s->dz * s->alpw[is]);
  s->u[is + 1] = dis * amp / (s->dz * s->alpw[is + 1]);
  /* divide the delta by (1-X)^2: two (1-X)^-1 solves, np = 1 */
  s->pd1[0] = 0.0;
  s->pd2[0] = -1.0;
  ray_matrc(s, 1);
  ray_solve(s, 1);
  ray_solve(s, 1); /* each solve leaves u[0] = 1e-30 (solve.py:92-93) */
  /* apply (1-X)^2 (1+X)^-1/4 exp(i k0 dr (sqrt(1+X)-1)) */
  if (orc_pade(in->freq, in->c0, s->np_, s->ns, s->dr, 2, s->pd1, s->pd2)) return -1;
  ray_matrc(s, s->np_);
  ray_solve(s, s->np_);
  return 0;
}

/* Range update (updat.py:38-140, cupyram.py:951-999).  Returns 1 when the
 * matrices must be rebuilt. */
static int ray_updat(ray_state *s, int *need_profl) {
  const orc_ray *in = s->in;
  int need = 0;
  *need_profl = 0;
  if (s->n_bt >= 2) {
    while (s->bt_ind < s->n_bt - 1 && s->r >= s->rbzb[2 * (s->bt_ind + 1)]) s->bt_ind++;
    const int jz = s->iz;
    const double rc = s->r + 0.5 * s->dr;
    const double r0 = s->rbzb[2 * s->bt_ind], z0 = s->rbzb[2 * s->bt_ind + 1];
    const double r1 = s->rbzb[2 * (s->bt_ind + 1)], z1 = s->rbzb[2 * (s->bt_ind + 1) + 1];
    const double z = z0 + (rc - r0) * (z1 - z0) / (r1 - r0);
    int iz = (int)floor(z / s->dz);
    if (iz < 1) iz = 1;
    if (iz > s->nz - 1) iz = s->nz - 1;
    s->iz = iz;
    if (iz != jz) need = 1;
  }
  if (in->n_rp_ss >= 2) {
    const int o = s->ss_ind;
    while (s->ss_ind < in->n_rp_ss - 1 && s->r >= in->rp_ss[s->ss_ind + 1]) s->ss_ind++;
    if (s->ss_ind != o) need = *need_profl = 1;
  }
  if (in->n_rp_sb >= 2) {
    const int o = s->sb_ind;
    while (s->sb_ind < in->n_rp_sb - 1 && s->r >= in->rp_sb[s->sb_ind + 1]) s->sb_ind++;
    if (s->sb_ind != o) need = *need_profl = 1;
  }
  if (s->r >= s->rs) {
    s->ns = 0;
    s->rs = s->rmax + s->dr;
    if (orc_pade(in->freq, in->c0, s->np_, s->ns, s->dr, 1, s->pd1, s->pd2)) return -1;
    need = 1;
  }
  return need;
}

int orc_dims(const orc_ray *in, int *nz, int *nvr, int *nvz, int *nsteps) {
  if (in->dz <= 0 || in->dr <= 0 || in->ndr < 1 || in->ndz < 1) return -2;
  const int nz_ = (int)floor(in->zmax / in->dz) - 1;
  const int nzplt = (int)floor(in->zmplt / in->dz);
  const int nvr_ = (int)floor(in->rmax / (in->dr * in->ndr));
  const int nvz_ = (int)floor((double)nzplt / in->ndz);
  const double rmax = nvr_ * in->dr * in->ndr;
  if (nz) *nz = nz_;
  if (nvr) *nvr = nvr_;
  if (nvz) *nvz = nvz_;
  if (nsteps) *nsteps = (int)nearbyint(rmax / in->dr) - 1;
  return 0;
}

static void *xcalloc(size_t n, size_t sz) { return calloc(n ? n : 1, sz); }

static int run_impl(const orc_ray *in, double *tll, double complex *cpl, double *tlg, double complex *cpg,
                    int *iz_hist, double *tlcell, int *trace, double complex *u_final) {
  ray_state s;
  memset(&s, 0, sizeof(s));
  s.in = in;
  s.np_ = in->np;
  s.ns = in->ns;
  s.dz = in->dz;
  s.dr = in->dr;
  s.zmax = in->zmax;
  if (orc_dims(in, &s.nz, &s.nvr, &s.nvz, &s.nr)) return -2;
  s.nzplt = (int)floor(in->zmplt / in->dz);
  if (s.nz < 3 || s.np_ < 1 || s.np_ > 16) return -2;
  s.rmax = s.nvr * in->dr * in->ndr;
  /* rs defaults to rmax + dr (cupyram.py:520); callers pass it resolved */
  s.rs = in->rs;
  const int nz2 = s.nz + 2;
  /* bathymetry extended to rmax with the last depth (cupyram.py:536-568);
   * the extension uses the rmax before truncation to whole output steps */
  s.n_bt = in->n_bt;
  s.rbzb = (double *)xcalloc(2 * (size_t)(in->n_bt + 1), sizeof(double));
  memcpy(s.rbzb, in->rbzb, sizeof(double) * 2 * in->n_bt);
  if (in->rbzb[2 * (in->n_bt - 1)] < in->rmax) {
    s.rbzb[2 * in->n_bt] = in->rmax;
    s.rbzb[2 * in->n_bt + 1] = in->rbzb[2 * (in->n_bt - 1) + 1];
    s.n_bt += 1;
  }
  /* seabed depths are relative to the deepest water-profile point
   * (cupyram.py:580-587) */
  s.z_sb = (double *)xcalloc(in->n_sb, sizeof(double));
  for (int i = 0; i < in->n_sb; ++i) s.z_sb[i] = in->z_sb[i] + in->z_ss[in->n_ss - 1];
  s.eta = 1.0 / (40.0 * M_PI * log10(exp(1.0)));
  s.omega = 2.0 * M_PI * in->freq;
  s.k0 = s.omega / in->c0;
  s.r = in->dr;
  {
    const double ri = in->zr / in->dz;
    s.ir = (int)floor(ri);
    s.dir = ri - s.ir;
  }
  {
    int iz = (int)floor(s.rbzb[1] / in->dz);
    if (iz < 1) iz = 1;
    if (iz > s.nz - 1) iz = s.nz - 1;
    s.iz = iz;
  }
  s.alpw = (double *)xcalloc(nz2, sizeof(double));
  s.alpb = (double *)xcalloc(nz2, sizeof(double));
  s.rhob = (double *)xcalloc(nz2, sizeof(double));
  s.ksqw = (cplx *)xcalloc(nz2, sizeof(cplx));
  s.ksqb = (cplx *)xcalloc(nz2, sizeof(cplx));
  s.f1 = (double *)xcalloc(nz2, sizeof(double));
  s.f2 = (double *)xcalloc(nz2, sizeof(double));
  s.f3 = (double *)xcalloc(nz2, sizeof(double));
  s.ksq = (cplx *)xcalloc(nz2, sizeof(cplx));
  s.u = (cplx *)xcalloc(nz2, sizeof(cplx));
  s.v = (cplx *)xcalloc(nz2, sizeof(cplx));
  const size_t nm = (size_t)nz2 * s.np_;
  s.r1 = (cplx *)xcalloc(nm, sizeof(cplx));
  s.r2 = (cplx *)xcalloc(nm, sizeof(cplx));
  s.r3 = (cplx *)xcalloc(nm, sizeof(cplx));
  s.s1 = (cplx *)xcalloc(nm, sizeof(cplx));
  s.s2 = (cplx *)xcalloc(nm, sizeof(cplx));
  s.s3 = (cplx *)xcalloc(nm, sizeof(cplx));
  s.tll = tll;
  s.cpl = (cplx *)cpl;
  s.tlg = tlg;
  s.cpg = (cplx *)cpg;
  s.cells = tlcell;
  if (tlcell) memset(tlcell, 0, sizeof(double) * (size_t)s.nvz * s.nvr);
  s.mdr = 0;
  s.tlc = -1;
  int rc = 0;
  const int is = (int)floor(in->zs / in->dz);
  if (is + 1 > s.nz + 1 || s.ir + 1 > s.nz + 1 || (s.nvz > 0 && s.nvz * in->ndz > s.nz + 1)) {
    rc = -2;
    goto done;
  }
  ray_profl(&s);
  if (ray_selfs(&s)) {
    rc = -1;
    goto done;
  }
  ray_outpt(&s);
  if (orc_pade(in->freq, in->c0, s.np_, s.ns, s.dr, 1, s.pd1, s.pd2)) {
    rc = -1;
    goto done;
  }
  ray_matrc(&s, s.np_);
  for (int rn = 0; rn < s.nr; ++rn) {
    int need_profl = 0;
    const int need = ray_updat(&s, &need_profl);
    if (need < 0) {
      rc = -1;
      goto done;
    }
    if (iz_hist) iz_hist[rn] = s.iz;
    if (trace) {
      trace[4 * rn + 0] = s.iz;
      trace[4 * rn + 1] = s.ss_ind;
      trace[4 * rn + 2] = s.sb_ind;
      trace[4 * rn + 3] = s.ns;
    }
    if (need_profl) ray_profl(&s);
    if (need) ray_matrc(&s, s.np_);
    ray_solve(&s, s.np_);
    s.r = (rn + 2) * in->dr;
    ray_outpt(&s);
  }
  if (u_final) memcpy(u_final, s.u, sizeof(cplx) * nz2);
  if (tlcell) {
    const double inv_n = 1.0 / ((double)in->ndr * in->ndz);
    for (size_t k = 0; k < (size_t)s.nvz * s.nvr; ++k) tlcell[k] = -10.0 * log10(tlcell[k] * inv_n + 1e-40);
  }
done:
  free(s.rbzb);
  free(s.z_sb);
  free(s.alpw);
  free(s.alpb);
  free(s.rhob);
  free(s.ksqw);
  free(s.ksqb);
  free(s.f1);
  free(s.f2);
  free(s.f3);
  free(s.ksq);
  free(s.u);
  free(s.v);
  free(s.r1);
  free(s.r2);
  free(s.r3);
  free(s.s1);
  free(s.s2);
  free(s.s3);
  return rc;
}

int orc_run(const orc_ray *in, double *tll, double complex *cpl, double *tlg, double complex *cpg,
            int *iz_hist) {
  return run_impl(in, tll, cpl, tlg, cpg, iz_hist, NULL, NULL, NULL);
}

int orc_run_trace(const orc_ray *in, double *tll, int *trace, double complex *u_final) {
  return run_impl(in, tll, NULL, NULL, NULL, NULL, NULL, trace, u_final);
}

int orc_run_cells(const orc_ray *in, double *tll, double *tlcell) {
  return run_impl(in, tll, NULL, NULL, NULL, NULL, tlcell, NULL, NULL);
}

int orc_run_batch(const orc_ray *rays, int nrays, double *const *tll, int nthreads) {
  int bad = 0;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads) reduction(+ : bad)
  for (int b = 0; b < nrays; ++b) bad += (orc_run(&rays[b], tll[b], NULL, NULL, NULL, NULL) != 0);
  return bad;
}

/* Depth profiles alone, for the numpy.interp pin. */
int orc_profiles(const orc_ray *in, int ss_ind, int sb_ind, double *alpw, double *alpb, double *rhob,
                 double complex *ksqw, double complex *ksqb) {
  ray_state s;
  memset(&s, 0, sizeof(s));
  s.in = in;
  s.dz = in->dz;
  s.zmax = in->zmax;
  if (orc_dims(in, &s.nz, NULL, NULL, NULL)) return -2;
  s.ss_ind = ss_ind;
  s.sb_ind = sb_ind;
  s.eta = 1.0 / (40.0 * M_PI * log10(exp(1.0)));
  s.omega = 2.0 * M_PI * in->freq;
  s.k0 = s.omega / in->c0;
  s.z_sb = (double *)xcalloc(in->n_sb, sizeof(double));
  for (int i = 0; i < in->n_sb; ++i) s.z_sb[i] = in->z_sb[i] + in->z_ss[in->n_ss - 1];
  s.alpw = alpw;
  s.alpb = alpb;
  s.rhob = rhob;
  s.ksqw = (cplx *)ksqw;
  s.ksqb = (cplx *)ksqb;
  ray_profl(&s);
  free(s.z_sb);
  return 0;
}

/* One application of the np rational factors to u with a given medium, for
 * the dense-solve pin of the two-sided elimination. */
int orc_apply_factors(int nz, int iz, double dz, double k0, int np_, const double complex *pd1,
                      const double complex *pd2, const double *alpw, const double *alpb, const double *rhob,
                      const double complex *ksqw, const double complex *ksqb, double complex *u) {
  ray_state s;
  memset(&s, 0, sizeof(s));
  const int nz2 = nz + 2;
  s.nz = nz;
  s.iz = iz;
  s.dz = dz;
  s.k0 = k0;
  for (int j = 0; j < np_; ++j) {
    s.pd1[j] = pd1[j];
    s.pd2[j] = pd2[j];
  }
  s.alpw = (double *)alpw;
  s.alpb = (double *)alpb;
  s.rhob = (double *)rhob;
  s.ksqw = (cplx *)ksqw;
  s.ksqb = (cplx *)ksqb;
  s.f1 = (double *)xcalloc(nz2, sizeof(double));
  s.f2 = (double *)xcalloc(nz2, sizeof(double));
  s.f3 = (double *)xcalloc(nz2, sizeof(double));
  s.ksq = (cplx *)xcalloc(nz2, sizeof(cplx));
  s.v = (cplx *)xcalloc(nz2, sizeof(cplx));
  s.u = (cplx *)u;
  const size_t nm = (size_t)nz2 * np_;
  s.r1 = (cplx *)xcalloc(nm, sizeof(cplx));
  s.r2 = (cplx *)xcalloc(nm, sizeof(cplx));
  s.r3 = (cplx *)xcalloc(nm, sizeof(cplx));
  s.s1 = (cplx *)xcalloc(nm, sizeof(cplx));
  s.s2 = (cplx *)xcalloc(nm, sizeof(cplx));
  s.s3 = (cplx *)xcalloc(nm, sizeof(cplx));
  ray_matrc(&s, np_);
  ray_solve(&s, np_);
  free(s.f1);
  free(s.f2);
  free(s.f3);
  free(s.ksq);
  free(s.v);
  free(s.r1);
  free(s.r2);
  free(s.r3);
  free(s.s1);
  free(s.s2);
  free(s.s3);
  return 0;
}
