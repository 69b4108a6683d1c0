// FP64 host computation of the split-step Pade coefficients.
//
// Restates cupyram/pade.py:180-273 (PyRAM.epade): the [np/np] rational
// approximation of g(X) = exp(i k0 dr (sqrt(1+X)-1)) (1+X)^alp (1-X)^(2 nu)
// with ns stability zeros of the numerator at X = -3 (ns >= 1) and X = -1.5
// (ns >= 2).  ip = 1 is the march operator (nu = 0, alp = 0); ip = 2 is the
// self-starter operator (nu = 1, alp = -1/4).  Steps kept alike on purpose
// (they define the coefficients and the order of the factors): Taylor
// derivatives by the Leibniz rule (pade.py:16-43), the accuracy/stability
// linear system (pade.py:229-253), Gaussian elimination with partial pivoting
// (pade.py:46-91), and Laguerre root finding from z = 0 with deflation, the
// last two roots by the quadratic formula (pade.py:94-177).
#include <cmath>
#include <complex>
#include <cstring>
#include <thread>
#include <vector>

#include "cpr_internal.h"

namespace {

using cd = std::complex<double>;

// Complex division by Smith's method (R. L. Smith, CACM 5(8), 1962), written
// out so the result does not depend on the runtime's complex-division helper.
cd cdiv(cd a, cd b) {
  const double ar = a.real(), ai = a.imag(), br = b.real(), bi = b.imag();
  if (std::fabs(br) >= std::fabs(bi)) {
    const double t = bi / br, d = br + bi * t;
    return cd((ar + ai * t) / d, (ai - ar * t) / d);
  }
  const double t = br / bi, d = br * t + bi;
  return cd((ar * t + ai) / d, (ai * t - ar) / d);
}
// Product with the textbook formula (no runtime helper).
cd cmul(cd a, cd b) { return cd(a.real() * b.real() - a.imag() * b.imag(), a.real() * b.imag() + a.imag() * b.real()); }

struct PadeWork {
  int np, n;
  double binom[33][33];
  double fact[33];
  cd A[32][33];
  cd rhs[33];
  cd dg[33];
};

void taylor_of_g(PadeWork &w, double sig, double alp, int nu) {
  const int n = w.n;
  cd dsum[32];
  cd t1(0.0, 0.5 * sig), t2(alp, 0.0), t3(-2.0 * nu, 0.0);
  double x1 = -0.5, x2 = -1.0, x3 = -1.0;
  dsum[0] = t1 + t2 + t3;
  for (int i = 1; i < n; ++i) {
    t1 = t1 * x1;
    x1 -= 1.0;
    t2 = t2 * x2;
    x2 -= 1.0;
    t3 = (-static_cast<double>(nu) * t3) * x3;
    x3 -= 1.0;
    dsum[i] = t1 + t2 + t3;
  }
  w.dg[0] = 1.0;
  w.dg[1] = dsum[0];
  for (int i = 1; i < n; ++i) {
    cd acc = dsum[i];
    for (int j = 0; j < i; ++j) acc += cmul(w.binom[i][j] * dsum[j], w.dg[i - j]);
    w.dg[i + 1] = acc;
  }
}

void solve_pivoted(PadeWork &w) {
  const int n = w.n;
  for (int i = 0; i < n; ++i) {
    if (i < n - 1) {
      int best = i;
      double big = std::abs(w.A[i][i]);
      for (int k = i + 1; k < n; ++k) {
        const double m = std::abs(w.A[k][i]);
        if (m > big) {
          best = k;
          big = m;
        }
      }
      if (best != i) {
        std::swap(w.rhs[best], w.rhs[i]);
        for (int j = i; j <= n; ++j) std::swap(w.A[best][j], w.A[i][j]);
      }
    }
    w.A[i][i] = cdiv(1.0, w.A[i][i]);
    w.rhs[i] = cmul(w.rhs[i], w.A[i][i]);
    if (i < n - 1) {
      for (int j = i + 1; j <= n; ++j) w.A[i][j] = cmul(w.A[i][j], w.A[i][i]);
      for (int k = i + 1; k < n; ++k) {
        w.rhs[k] -= cmul(w.A[k][i], w.rhs[i]);
        for (int j = i + 1; j < n; ++j) w.A[k][j] -= cmul(w.A[k][i], w.A[i][j]);
      }
    }
  }
  for (int i = n - 2; i >= 0; --i)
    for (int j = i + 1; j < n; ++j) w.rhs[i] -= cmul(w.A[i][j], w.rhs[j]);
}

// Laguerre's method on c[0..deg]; false when 100 iterations do not converge.
bool laguerre(const cd *c, int deg, cd &z, double tol, int max_iter) {
  cd d1[33], d2[33];
  for (int i = 0; i < deg; ++i) d1[i] = static_cast<double>(i + 1) * c[i + 1];
  for (int i = 0; i < deg - 1; ++i) d2[i] = static_cast<double>(i + 1) * d1[i + 1];
  int it = 0, rot = 0;
  double step = INFINITY;
  while (step > tol && it < max_iter - 1) {
    cd p = c[deg - 1] + cmul(c[deg], z);
    for (int i = deg - 2; i >= 0; --i) p = c[i] + cmul(z, p);
    if (std::abs(p) < 1e-20) break;
    cd p1 = d1[deg - 2] + cmul(d1[deg - 1], z);
    for (int i = deg - 3; i >= 0; --i) p1 = d1[i] + cmul(z, p1);
    cd p2 = d2[deg - 3] + cmul(d2[deg - 2], z);
    for (int i = deg - 4; i >= 0; --i) p2 = d2[i] + cmul(z, p2);
    const cd f = cdiv(p1, p);
    const cd g = cmul(f, f) - cdiv(p2, p);
    const cd h = std::sqrt(static_cast<double>(deg - 1) * (static_cast<double>(deg) * g - cmul(f, f)));
    cd dz = (std::abs(f + h) > std::abs(f - h)) ? cdiv(-static_cast<double>(deg), f + h)
                                                 : cdiv(-static_cast<double>(deg), f - h);
    ++it;
    if (++rot == 9) {
      rot = 0;
      dz = cd(-dz.imag(), dz.real());
    }
    z += dz;
    step = std::abs(dz);
    if (it == 100) return false;
  }
  return true;
}

bool poly_roots(cd *c, int deg, cd *z) {
  if (deg == 1) {
    z[0] = cdiv(-c[0], c[1]);
    return true;
  }
  for (int k = deg - 1; k >= 2; --k) {
    cd root = 0.0;
    if (!laguerre(c, k + 1, root, 1e-12, 1000)) return false;
    if (!laguerre(c, k + 1, root, 0.0, 5)) return false;
    z[k] = root;
    for (int i = k; i >= 0; --i) c[i] += cmul(root, c[i + 1]);
    for (int i = 0; i <= k; ++i) c[i] = c[i + 1];
  }
  const cd disc = std::sqrt(cmul(c[1], c[1]) - 4.0 * cmul(c[0], c[2]));
  z[1] = cdiv(0.5 * (-c[1] + disc), c[2]);
  z[0] = cdiv(0.5 * (-c[1] - disc), c[2]);
  return true;
}

}  // namespace

int cpr_pade_impl(double freq, double c0, int np_, int ns, double dr, int ip, double *pd1, double *pd2) {
  if (np_ < 1 || np_ > CPR_MAX_NP || !(freq > 0) || !(c0 > 0) || !(dr > 0)) return -2;
  PadeWork w;
  std::memset(static_cast<void *>(&w), 0, sizeof(w));
  w.np = np_;
  w.n = 2 * np_;
  const int n = w.n;
  const double sig = (2.0 * M_PI * freq / c0) * dr;
  const int nu = ip == 1 ? 0 : 1;
  const double alp = ip == 1 ? 0.0 : -0.25;
  w.fact[0] = 1.0;
  for (int i = 1; i < n; ++i) w.fact[i] = static_cast<double>(i + 1) * w.fact[i - 1];
  for (int i = 0; i <= n; ++i) w.binom[i][0] = w.binom[i][i] = 1.0;
  for (int i = 2; i <= n; ++i)
    for (int j = 1; j < i; ++j) w.binom[i][j] = w.binom[i - 1][j - 1] + w.binom[i - 1][j];
  taylor_of_g(w, sig, alp, nu);
  for (int i = 0; i < n; ++i) w.rhs[i] = w.dg[i + 1];
  for (int i = 0; i < n; ++i) {
    if (2 * i <= n - 1) w.A[i][2 * i] = w.fact[i];
    for (int j = 0; j <= i; ++j)
      if (2 * j + 1 <= n - 1) w.A[i][2 * j + 1] = (-w.binom[i + 1][j + 1] * w.fact[j]) * w.dg[i - j];
  }
  const double zeros[2] = {-3.0, -1.5};
  for (int c = 0; c < 2 && c < ns; ++c) {
    const int row = n - 1 - c;
    double zp = 1.0;
    w.rhs[row] = -1.0;
    for (int j = 0; j < np_; ++j) {
      zp *= zeros[c];
      w.A[row][2 * j] = zp;
      w.A[row][2 * j + 1] = 0.0;
    }
  }
  solve_pivoted(w);
  cd poly[33], roots[33];
  for (int which = 0; which < 2; ++which) {
    poly[0] = 1.0;
    for (int j = 0; j < np_; ++j) poly[j + 1] = w.rhs[2 * j + which];
    if (!poly_roots(poly, np_, roots)) return -1;
    double *dst = which == 0 ? pd1 : pd2;
    for (int j = 0; j < np_; ++j) {
      const cd v = cdiv(-1.0, roots[j]);
      dst[2 * j] = v.real();
      dst[2 * j + 1] = v.imag();
    }
  }
  return 0;
}

int cpr_pade_batch_impl(int n, const double *freq, const double *c0, int np_, int ns, const double *dr, int ip,
                        double *out, int nthreads) {
  if (n <= 0) return 0;
  if (nthreads < 1) nthreads = 1;
  if (nthreads > n) nthreads = n;
  std::vector<int> status(nthreads, 0);
  auto work = [&](int t) {
    for (int i = t; i < n; i += nthreads) {
      double *dst = out + static_cast<size_t>(i) * 4 * np_;
      const int rc = cpr_pade_impl(freq[i], c0[i], np_, ns, dr[i], ip, dst, dst + 2 * np_);
      if (rc && !status[t]) status[t] = rc;
    }
  };
  std::vector<std::thread> pool;
  for (int t = 1; t < nthreads; ++t) pool.emplace_back(work, t);
  work(0);
  for (auto &th : pool) th.join();
  for (int t = 0; t < nthreads; ++t)
    if (status[t]) return status[t];
  return 0;
}
