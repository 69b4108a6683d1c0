// Team-of-lanes march for gfx950: the field of a ray stays in registers for
// every range step of a launch (and, for columns up to 12 rows per lane in
// complex64 / 6 in complex128, its Galerkin coefficients too; complex64 lanes
// of 14-24 rows keep most of theirs in LDS).
//
// A team of T lanes owns one ray; lane t holds rows g = 1 + t*M .. (t+1)*M
// (M rows per lane, a template parameter, so every per-row array is fully
// unrolled into VGPRs).  Rows past nz are identity rows (x = 0).  Per Pade
// factor (M + p2 D) u' = (M + p1 D) u:
//   1. each lane forms its RHS rows and eliminates downwards through its rows
//      keeping the coupling to the previous lane's last unknown y_{t-1}
//      ("spike" w): x_k + c'_k x_{k+1} + w_k y_{t-1} = d'_k;
//   2. sweeps upwards to express its rows in y_{t-1} and its own last
//      unknown y_t: x_k = D_k - W_k y_{t-1} - C_k y_t;
//   3. its last row plus the next lane's first row give one tridiagonal
//      equation in (y_{t-1}, y_t, y_{t+1}); the T equations are solved by
//      parallel cyclic reduction with wave shuffles (T = 64), or, for larger
//      teams, by cyclic reduction through LDS down to 64 unknowns, PCR in one
//      wavefront and the reduction undone;
//   4. back-substitutes its rows.
// In complex64 the factor's right-hand side is applied without a product
// with D: with c = (1 - p1/p2) k0^2 / (6 p2) the factor equals
// u' = (p1/p2) u + c y, (M + p2 D) y = (6/k0^2) M u, and the lanes march the
// field divided by the running product of the c's (multiplied back at the
// end of every step and wherever the product would leave [2^-30, 2^30]).
// This solves the same linear system as the reference's top-to-bottom
// Thomas sweep (fused_kernel.py:124-152) in a different elimination order.
// The range update (updat.py:38-140), medium rebuild where the water/bottom
// split moved (matrc.py:63-107, profl.py:45-140) and the TL output
// (outpt.py:11-58) are fused in the same persistent loop over range steps.
#include <hip/hip_runtime.h>

#include <cmath>
#include <type_traits>

#include "march_common.h"

namespace {

// Reciprocal of a positive real: the hardware estimate for FP32 (1 ulp);
// for FP64 the estimate refined by two Newton-Raphson steps
// r <- r + r (1 - x r), each of which squares the relative error.
template <class R>
__device__ inline R rcp_real(R x);
template <>
__device__ inline float rcp_real(float x) {
  return __builtin_amdgcn_rcpf(x);
}
template <>
__device__ inline double rcp_real(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}

// 1/a as conj(a)/|a|^2
template <class R>
__device__ inline cx<R> fast_rcp(cx<R> a) {
  const R m = rcp_real<R>(a.re * a.re + a.im * a.im);
  return cxm<R>(a.re * m, -(a.im * m));
}

// -(a b), written so that the negation lands on FMA source operands
template <class R>
__device__ inline cx<R> cnmul(cx<R> a, cx<R> b) {
  return cxm<R>(a.im * b.im - a.re * b.re, -(a.re * b.im) - a.im * b.re);
}
// a + b c
template <class R>
__device__ inline cx<R> cfma(cx<R> a, cx<R> b, cx<R> c) {
  return cxm<R>(a.re + b.re * c.re - b.im * c.im, a.im + b.re * c.im + b.im * c.re);
}

// Complex a + b c, a - b c and a b with every fused multiply-add written
// out, so that two evaluations of one expression round alike wherever the
// compiler places them: the halo rows a lane carries from factor to factor
// (its neighbours' first and last rows) must equal, bit for bit, the rows
// those neighbours compute, or a launch would not match two shorter ones.
template <class R>
__device__ inline cx<R> xfma(cx<R> a, cx<R> b, cx<R> c) {
  return cxm<R>(fma(-b.im, c.im, fma(b.re, c.re, a.re)), fma(b.im, c.re, fma(b.re, c.im, a.im)));
}
template <class R>
__device__ inline cx<R> xfms(cx<R> a, cx<R> b, cx<R> c) {
  return cxm<R>(fma(b.im, c.im, fma(-b.re, c.re, a.re)), fma(-b.im, c.re, fma(-b.re, c.im, a.im)));
}
template <class R>
__device__ inline cx<R> xmul(cx<R> a, cx<R> b) {
  return cxm<R>(fma(a.re, b.re, -(a.im * b.im)), fma(a.re, b.im, a.im * b.re));
}

// Copy a team-uniform value into scalar registers.
template <class T>
__device__ inline T make_uniform(const T &v) {
  static_assert(sizeof(T) % 4 == 0, "4-byte granularity");
  T r;
  const int *src = reinterpret_cast<const int *>(&v);
  int *dst = reinterpret_cast<int *>(&r);
#pragma unroll
  for (int i = 0; i < static_cast<int>(sizeof(T) / 4); ++i) dst[i] = __builtin_amdgcn_readfirstlane(src[i]);
  return r;
}

template <class R>
__device__ inline R shfl_up_r(R v, int s, int w) {
  return __shfl_up(v, s, w);
}
template <class R>
__device__ inline R shfl_down_r(R v, int s, int w) {
  return __shfl_down(v, s, w);
}

// Team exchange: values from lane t-s ("up") / t+s ("down"); zero outside
// the team.  TW > 0: wave shuffles of width TW.  TW == 0: LDS with one
// barrier per exchange.  LDS layout: two buffers, used alternately, so a
// buffer is written again only after every lane has passed the barrier that
// follows its last reads of it.  A buffer is an array of per-lane slots of
// three complex values (one exchange writes one slot per lane, so reading a
// neighbour's values takes one address), with zeroed margins of kPad slots on
// either side, so that reads of lanes t - 1 and t + 1 need no bounds test.
// The margins are zeroed once per launch (init) and never written.
constexpr int kPad = 32;

template <class R>
struct Slot {
  cx<R> a, c, f;
};

// Reduced system of an LDS-mode team (T > 64 lanes, a multiple of 64):
// cyclic reduction levels k = 0 .. K-1 hold n_k = T / 2^k > 64 unknowns (all
// even counts); level 0's equations go through the next exchange buffer and
// those of levels 1 .. K-1 through the buffer the previous exchange used.
// The solutions of levels 1 .. K need their own LDS arrays, one zero entry
// past the end of each; their complex entries:
__host__ __device__ constexpr int cr_values(int T) {
  int c = 0;
  for (int n = T; n > 64;) {
    n = (n + 1) / 2;
    c += n + 1;
  }
  return c;
}

// LDS bytes of an LDS-mode team of T lanes: the two exchange buffers, the
// reduced system's solution arrays, then the in-wave PCR buffer (64 slots
// with zero margins)
template <class R>
__host__ __device__ constexpr size_t team_lds_bytes(int T) {
  return 2 * static_cast<size_t>(T + 2 * kPad) * sizeof(Slot<R>) + cr_values(T) * sizeof(cx<R>) +
         static_cast<size_t>(64 + 2 * kPad) * sizeof(Slot<R>);
}

// LDS bytes of a block of four 64-lane teams before their coefficient rows:
// the complex64 teams' in-wave PCR buffers (complex128 teams use shuffles)
template <class R>
__host__ __device__ constexpr size_t wave_team_lds_bytes() {
  return sizeof(R) == 4 ? 4 * static_cast<size_t>(64 + 2 * kPad) * sizeof(Slot<R>) : 0;
}

// One elimination step of the reduced system, shared by PCR and cyclic
// reduction: equation (a, 1, c | f) of an unknown with the equations of its
// neighbours at distance s, (am, 1, cm | fm) before and (ap, 1, cp | fp)
// after, gives the equation in the unknowns at distance 2s, normalised to a
// unit diagonal.
template <class R>
__device__ inline void reduce_step(cx<R> &a, cx<R> &c, cx<R> &f, cx<R> am, cx<R> cm, cx<R> fm, cx<R> ap, cx<R> cp,
                                   cx<R> fp) {
  const cx<R> bn = cfms(cfms(cxm<R>(1, 0), a, cm), c, ap);
  const cx<R> an = cnmul(a, am);
  const cx<R> cn = cnmul(c, cp);
  const cx<R> fn = cfms(cfms(f, a, fm), c, fp);
  const cx<R> ib = fast_rcp(bn);
  a = an * ib;
  c = cn * ib;
  f = fn * ib;
}

// PCR over the 64 unknowns of one wavefront (lane t holds equation t),
// exchanging through an LDS buffer of 64 + 2 kPad slots (pb: lane 0's slot)
// used as three arrays a, c, f of 128 complex values, each with kPad-entry
// margins that hold zeros, the values past the ends.  A wavefront's LDS
// operations execute in order, so a lane reads its neighbours' entries right
// after the write of its own; the empty asm statements keep the compiler
// from moving the accesses across one another.  Against wave shuffles, which
// move one dword per instruction: 81.2 instead of 85.4 ms per 100-step
// complex64 Munk launch, 10-21 % on the 4-12-row wavefront teams.
template <class R>
__device__ inline void pcr64_lds(Slot<R> *pb, int t, cx<R> &a, cx<R> &c, cx<R> &f) {
  cx<R> *const A = reinterpret_cast<cx<R> *>(pb - kPad) + kPad;
  cx<R> *const Cc = A + (64 + 2 * kPad);
  cx<R> *const F = Cc + (64 + 2 * kPad);
#pragma unroll
  for (int l = 0; l < 6; ++l) {
    const int s = 1 << l;
    A[t] = a;
    Cc[t] = c;
    F[t] = f;
    asm volatile("" ::: "memory");
    const cx<R> am = A[t - s], cm = Cc[t - s], fm = F[t - s];
    const cx<R> ap = A[t + s], cp = Cc[t + s], fp = F[t + s];
    asm volatile("" ::: "memory");
    reduce_step(a, c, f, am, cm, fm, ap, cp, fp);
  }
}

template <class R, int TW>
struct Team {
  int t, T;
  Slot<R> *lds;  // LDS mode: 2 buffers of (kPad + T + kPad) slots; complex64 wavefront teams: lane 0's slot of the PCR buffer
  int ph;

  // zero both buffers (margins included) and the reduced system's solution
  // arrays; one barrier
  __device__ void init() {
    if constexpr (TW > 0 && sizeof(R) == 4) {
      // a wavefront team's own PCR buffer (lds: lane 0's slot): zeroed once
      for (int i = t; i < TW + 2 * kPad; i += TW) lds[i - kPad] = Slot<R>{cxm<R>(0, 0), cxm<R>(0, 0), cxm<R>(0, 0)};
    }
    if constexpr (TW == 0) {
      const int n = 2 * (T + 2 * kPad);
      for (int i = t; i < n; i += T) lds[i] = Slot<R>{cxm<R>(0, 0), cxm<R>(0, 0), cxm<R>(0, 0)};
      cx<R> *u = reinterpret_cast<cx<R> *>(lds + n);
      for (int i = t; i < cr_values(T); i += T) u[i] = cxm<R>(0, 0);
      Slot<R> *pb = pcr_buf() - kPad;
      for (int i = t; i < 64 + 2 * kPad; i += T) pb[i] = Slot<R>{cxm<R>(0, 0), cxm<R>(0, 0), cxm<R>(0, 0)};
      __syncthreads();
    }
  }

  // LDS mode: solve a_t y_{t-1} + y_t + c_t y_{t+1} = f_t over the team
  // (a_0 = c_{T-1} = 0) and return y_t.  Cyclic reduction halves the system
  // (the unknowns at even positions stay, held compacted by the lanes
  // t < n/2, so whole wavefronts drop out) until 64 unknowns remain;
  // wavefront 0 solves those by PCR (through LDS in complex64, with shuffles in
  // complex128); the levels are then undone
  // (even positions copy, odd ones solve their own equation).  Against PCR
  // over all T lanes: 2K barriers instead of log2 T and a fraction of the
  // VALU work (K = log2 (T / 64) levels).  Called right after an exchange
  // (next3), whose buffer every lane has read before the first barrier here.
  __device__ cx<R> solve_cr(cx<R> a, cx<R> c, cx<R> f, cx<R> &yl, cx<R> &yr) {
    static_assert(TW == 0, "cyclic reduction runs on LDS teams");
    Slot<R> *const L0 = next_buf();                     // level 0
    Slot<R> *const L1 = lds + ph * (T + 2 * kPad) + kPad;  // levels 1 .. K-1, back to back
    cx<R> *const U1 = reinterpret_cast<cx<R> *>(lds + 2 * (T + 2 * kPad));
    const cx<R> z = cxm<R>(0, 0);
    Slot<R> *L = L0;
    int n = T;
    L[t] = Slot<R>{a, c, f};
    while (n > 64) {
      __syncthreads();
      const int n2 = n >> 1;
      Slot<R> *const N = L == L0 ? L1 : L + n;
      if (t < n2) {
        // unknown 2t with its neighbours 2t - 1 (none for t = 0) and 2t + 1
        const Slot<R> o = L[2 * t], p = L[2 * t + 1];
        Slot<R> m = L[t > 0 ? 2 * t - 1 : 0];
        if (t == 0) m = Slot<R>{z, z, z};
        a = o.a;
        c = o.c;
        f = o.f;
        reduce_step(a, c, f, m.a, m.c, m.f, p.a, p.c, p.f);
        if (n2 > 64) N[t] = Slot<R>{a, c, f};
      }
      L = N;
      n = n2;
    }
    cx<R> y = z;
    if (t < 64) {
      // lanes past the last level (T = 448: 56 unknowns) hold y = 0
      if (t >= n) a = c = f = z;
      if constexpr (sizeof(R) == 4) {
        pcr64_lds(pcr_buf(), t, a, c, f);
      } else {
        // complex128: shuffles (the 48-byte LDS slots measured slower)
#pragma unroll
        for (int l = 0; l < 6; ++l) {
          const int s = 1 << l;
          const bool lo = t >= s, hi = t + s < 64;
          const cx<R> a1 = cxm<R>(shfl_up_r(a.re, s, 64), shfl_up_r(a.im, s, 64));
          const cx<R> c1 = cxm<R>(shfl_up_r(c.re, s, 64), shfl_up_r(c.im, s, 64));
          const cx<R> f1 = cxm<R>(shfl_up_r(f.re, s, 64), shfl_up_r(f.im, s, 64));
          const cx<R> a2 = cxm<R>(shfl_down_r(a.re, s, 64), shfl_down_r(a.im, s, 64));
          const cx<R> c2 = cxm<R>(shfl_down_r(c.re, s, 64), shfl_down_r(c.im, s, 64));
          const cx<R> f2 = cxm<R>(shfl_down_r(f.re, s, 64), shfl_down_r(f.im, s, 64));
          reduce_step(a, c, f, lo ? a1 : z, lo ? c1 : z, lo ? f1 : z, hi ? a2 : z, hi ? c2 : z, hi ? f2 : z);
        }
      }
      y = f;
    }
    // undo the levels: level k's equations at L_k, level k+1's solution
    // through U_{k+1} (n_{k+1} entries and a zero one); level 0 separately
    int K = 0;
    for (int m = T; m > 64; m >>= 1) ++K;
    for (int k = K - 1; k >= 1; --k) {
      Slot<R> *Lk = L0;
      cx<R> *U = U1;
      int nk = T;
      for (int i = 0; i < k; ++i) {
        Lk = i == 0 ? L1 : Lk + nk;
        nk >>= 1;
        U += nk + 1;
      }
      const int nk1 = nk >> 1;
      if (t < nk1) U[t] = y;
      __syncthreads();
      if (t < nk) {
        if ((t & 1) == 0) {
          y = U[t >> 1];
        } else {
          const Slot<R> e = Lk[t];
          y = xfms(xfms(e.f, e.a, U[(t - 1) >> 1]), e.c, U[(t + 1) >> 1]);
        }
      }
    }
    // level 0: every lane forms y_t and its neighbours' y_{t-1}, y_{t+1}
    // (zero past the team's ends) by the same expressions their owners use,
    // so no exchange follows
    if (t < T / 2) U1[t] = y;
    __syncthreads();
    const int h = t >> 1;
    if (t & 1) {
      const Slot<R> e = L0[t];
      const cx<R> u0 = U1[h], u1 = U1[h + 1];  // U1[T/2] is the zero entry
      y = xfms(xfms(e.f, e.a, u0), e.c, u1);
      yl = u0;
      yr = u1;
    } else {
      const Slot<R> el = L0[t - 1], er = L0[t + 1];  // L0[-1]: the zero margin
      const cx<R> um = t > 0 ? U1[h - 1] : z, u0 = U1[h], up = U1[h + 1];
      y = u0;
      yl = xfms(xfms(el.f, el.a, um), el.c, u0);
      yr = xfms(xfms(er.f, er.a, u0), er.c, up);
    }
    return y;
  }
  // lane 0's slot of the in-wave PCR buffer
  __device__ Slot<R> *pcr_buf() {
    cx<R> *u = reinterpret_cast<cx<R> *>(lds + 2 * (T + 2 * kPad)) + cr_values(T);
    return reinterpret_cast<Slot<R> *>(u) + kPad;
  }
  // lane 0's slot of the next buffer
  __device__ Slot<R> *next_buf() {
    Slot<R> *b = lds + ph * (T + 2 * kPad) + kPad;
    ph ^= 1;
    return b;
  }

  // y_{t-1} and y_{t+1} of a value y (zero beyond the team's ends)
  __device__ void neighbours(cx<R> v, cx<R> &from_prev, cx<R> &from_next) {
    if constexpr (TW > 0) {
      const cx<R> r = cxm<R>(shfl_up_r(v.re, 1, TW), shfl_up_r(v.im, 1, TW));
      const cx<R> d = cxm<R>(shfl_down_r(v.re, 1, TW), shfl_down_r(v.im, 1, TW));
      from_prev = t >= 1 ? r : cxm<R>(0, 0);
      from_next = t + 1 < T ? d : cxm<R>(0, 0);
    } else {
      Slot<R> *B = next_buf();
      B[t].a = v;
      __syncthreads();
      from_prev = B[t - 1].a;
      from_next = B[t + 1].a;
    }
  }
  // halo: previous lane's `a` and next lane's `b`
  __device__ void halo(cx<R> a, cx<R> b, cx<R> &from_prev, cx<R> &from_next) {
    if constexpr (TW > 0) {
      const cx<R> r = cxm<R>(shfl_up_r(a.re, 1, TW), shfl_up_r(a.im, 1, TW));
      const cx<R> d = cxm<R>(shfl_down_r(b.re, 1, TW), shfl_down_r(b.im, 1, TW));
      from_prev = t >= 1 ? r : cxm<R>(0, 0);
      from_next = t + 1 < T ? d : cxm<R>(0, 0);
    } else {
      Slot<R> *B = next_buf();
      B[t].a = a;
      B[t].c = b;
      __syncthreads();
      from_prev = B[t - 1].a;
      from_next = B[t + 1].c;
    }
  }
  // LX: a complex64 wavefront team exchanges through its PCR buffer (in-order
  // LDS operations of one wavefront; the entries past lane 63 are zero):
  // 1.5-3.5 % faster than shuffles for the 4-12-row lanes, 1.5 % slower for
  // 24 rows
  template <bool LX>
  __device__ void next3(cx<R> a, cx<R> b, cx<R> c, cx<R> &na, cx<R> &nb, cx<R> &nc) {
    if constexpr (TW > 0 && sizeof(R) == 4 && LX) {
      cx<R> *const A = reinterpret_cast<cx<R> *>(lds - kPad) + kPad;
      cx<R> *const Bb = A + (64 + 2 * kPad);
      cx<R> *const Cc = Bb + (64 + 2 * kPad);
      A[t] = a;
      Bb[t] = b;
      Cc[t] = c;
      asm volatile("" ::: "memory");
      na = A[t + 1];
      nb = Bb[t + 1];
      nc = Cc[t + 1];
      asm volatile("" ::: "memory");
    } else if constexpr (TW > 0) {
      const bool ok = t + 1 < T;
      const cx<R> z = cxm<R>(0, 0);
      const cx<R> ra = cxm<R>(shfl_down_r(a.re, 1, TW), shfl_down_r(a.im, 1, TW));
      const cx<R> rb = cxm<R>(shfl_down_r(b.re, 1, TW), shfl_down_r(b.im, 1, TW));
      const cx<R> rc = cxm<R>(shfl_down_r(c.re, 1, TW), shfl_down_r(c.im, 1, TW));
      na = ok ? ra : z;
      nb = ok ? rb : z;
      nc = ok ? rc : z;
    } else {
      Slot<R> *B = next_buf();
      B[t] = Slot<R>{a, b, c};
      __syncthreads();
      const Slot<R> n = B[t + 1];
      na = n.a;
      nb = n.c;
      nc = n.f;
    }
  }
  // one PCR exchange of a wavefront team: (a, c, f) of lanes t-s and t+s
  __device__ void pcr(int s, cx<R> a, cx<R> c, cx<R> f, cx<R> &am, cx<R> &cm, cx<R> &fm, cx<R> &ap, cx<R> &cp,
                      cx<R> &fp) {
    static_assert(TW > 0, "LDS teams solve the reduced system with solve_cr");
    const bool lo = t >= s, hi = t + s < T;
    const cx<R> z = cxm<R>(0, 0);
    const cx<R> a1 = cxm<R>(shfl_up_r(a.re, s, TW), shfl_up_r(a.im, s, TW));
    const cx<R> c1 = cxm<R>(shfl_up_r(c.re, s, TW), shfl_up_r(c.im, s, TW));
    const cx<R> f1 = cxm<R>(shfl_up_r(f.re, s, TW), shfl_up_r(f.im, s, TW));
    const cx<R> a2 = cxm<R>(shfl_down_r(a.re, s, TW), shfl_down_r(a.im, s, TW));
    const cx<R> c2 = cxm<R>(shfl_down_r(c.re, s, TW), shfl_down_r(c.im, s, TW));
    const cx<R> f2 = cxm<R>(shfl_down_r(f.re, s, TW), shfl_down_r(f.im, s, TW));
    am = lo ? a1 : z;
    cm = lo ? c1 : z;
    fm = lo ? f1 : z;
    ap = hi ? a2 : z;
    cp = hi ? c2 : z;
    fp = hi ? f2 : z;
  }
};

template <class R, int M>
struct LaneCol {
  cx<R> u[M];  // field, then d', then D
};

// Elimination arrays of a lane: c' (then C) and the spike w (then W) of
// every row.  Rows k < NR are register arrays.  In LDS teams with more rows
// per lane than the registers hold (complex64 past 24 rows, complex128 past
// 11), rows NR .. M-1 live in the block's LDS as a run of (c', w) pairs per
// lane: otherwise the compiler spills them to private memory, which measured
// up to 2x slower (complex64 32 x 512 lanes 25.2 against 52.9 G cell-steps/s
// with the LDS rows, 28 x 512 37.1 against 55.7 G; complex128 16 x 512 13.7
// against 18.3 G).  A lane touches only its own run; the runs are an odd
// number of 16-byte units apart, so the lanes of a wavefront spread over every
// bank.  Every row index is a constant after unrolling, so each access is one
// LDS instruction at an immediate offset from the lane's base.  (Moving rows
// of the shallower lanes to LDS measured 1-10 % slower.)
template <class R>
struct ElimPair {
  cx<R> c, w;
};

template <class R, int M, int NR>
struct Elim {
  static_assert(NR >= 1 && NR <= M, "at least one register row");
  cx<R> cw[NR], ww[NR];
  ElimPair<R> *tail;  // rows NR .. M-1 of this lane (unused when NR = M)
  __device__ cx<R> c(int k) const { return k < NR ? cw[k] : tail[k - NR].c; }
  __device__ cx<R> w(int k) const { return k < NR ? ww[k] : tail[k - NR].w; }
  __device__ void set(int k, cx<R> cv, cx<R> wv) {
    if (k < NR) {
      cw[k] = cv;
      ww[k] = wv;
    } else {
      tail[k - NR] = ElimPair<R>{cv, wv};
    }
  }
};

// Register rows of the elimination arrays of an LDS team with M rows per
// lane; the other M - NR rows (at most 15 in complex64, 5 in complex128, so
// that a 512-lane team's LDS stays within the CU's 160 KB) go to LDS.
template <class R, int M>
__host__ __device__ constexpr int elim_regs() {
  if (sizeof(R) == 4) return M <= 24 ? M : (M - 15 > 12 ? M - 15 : 12);
  return M <= 11 ? M : M - 5;
}
// LDS bytes per lane of the tail rows: M - NR pairs, padded to an odd number
// of 16-byte units
template <class R, int M>
__host__ __device__ constexpr int elim_tail_lane_bytes() {
  const int units = (M - elim_regs<R, M>()) * static_cast<int>(sizeof(ElimPair<R>)) / 16;
  return units == 0 ? 0 : (units | 1) * 16;
}

// Galerkin coefficients of one row.
template <class R>
struct CoefRow {
  cx<R> d1, d2, d3;
};

// One half (real or imaginary parts) of a row's coefficients.
template <class R>
struct CoefHalf {
  R d1, d2, d3;
};

// Where a lane keeps the coefficients of its rows: VGPRs (DS = false), or a
// global buffer (DS = true, "d-stream": frees 7 VGPRs per row for deep
// columns; the buffer of the rays in flight stays L2/MALL-resident).  The
// buffer holds the real parts [ray][row k][lane], then the imaginary parts in
// the same layout, so that a wavefront reading row k touches one contiguous
// span of each, and a wavefront whose rows all have real coefficients (rows
// above the seabed, see march_team_body) reads only the real half.
template <class R, int M, bool DS, int KC = 0>
struct Coefs {
  CoefRow<R> r[M];
  __device__ CoefRow<R> get(int k) const { return r[k]; }
  __device__ CoefHalf<R> get_re(int k) const { return CoefHalf<R>{r[k].d1.re, r[k].d2.re, r[k].d3.re}; }
  __device__ const Coefs &view() const { return *this; }
};
template <class R, int M, int KC>
struct Coefs<R, M, true, KC> {
  CoefHalf<R> *re, *im;  // team-uniform ray bases: row k of lane t at [k * T + t]
  int T, t;
  R *lc;  // LDS (KC > 0): this lane's real halves d1 d2 d3 of rows 0 .. KC-1
  typedef R __attribute__((address_space(3))) *lreal;
  // Load views: the ray bases copied through an empty asm statement, so the
  // compiler keeps them in scalar registers and forms each row's address
  // there (scalar base + the lane's 32-bit byte offset) instead of hoisting
  // one 64-bit vector address per row out of the step loop and spilling them.
  typedef const char __attribute__((address_space(1))) *gbytes;  // global memory
  typedef const CoefHalf<R> __attribute__((address_space(1))) *ghalf;
  struct View {
    gbytes re, im;
    int64_t row;   // bytes per row: T * sizeof(CoefHalf)
    unsigned off;  // this lane's byte offset: t * sizeof(CoefHalf)
    lreal lc;
    __device__ CoefHalf<R> half(gbytes b, int k) const {
      const ghalf p = reinterpret_cast<ghalf>(b + k * row + off);
      return CoefHalf<R>{p->d1, p->d2, p->d3};
    }
    __device__ CoefHalf<R> real(int k) const {
      if (k < KC) return CoefHalf<R>{lc[3 * k], lc[3 * k + 1], lc[3 * k + 2]};
      return half(re, k);
    }
    __device__ CoefRow<R> get(int k) const {
      const CoefHalf<R> a = real(k), b = half(im, k);
      CoefRow<R> c;
      c.d1 = cxm<R>(a.d1, b.d1);
      c.d2 = cxm<R>(a.d2, b.d2);
      c.d3 = cxm<R>(a.d3, b.d3);
      return c;
    }
    __device__ CoefHalf<R> get_re(int k) const { return real(k); }
  };
  __device__ View view() const {
    gbytes r = (gbytes)(re), i = (gbytes)(im);  // generic to global address space
    asm volatile("" : "+s"(r), "+s"(i));
    return View{r, i, static_cast<int64_t>(T) * static_cast<int64_t>(sizeof(CoefHalf<R>)),
                static_cast<unsigned>(t) * static_cast<unsigned>(sizeof(CoefHalf<R>)), (lreal)(lc)};
  }
  __device__ void put(int k, const CoefRow<R> &c) const {
    (re + k * T)[t] = CoefHalf<R>{c.d1.re, c.d2.re, c.d3.re};
    (im + k * T)[t] = CoefHalf<R>{c.d1.im, c.d2.im, c.d3.im};
    if (k < KC) {
      const lreal l = (lreal)(lc);
      l[3 * k] = c.d1.re;
      l[3 * k + 1] = c.d2.re;
      l[3 * k + 2] = c.d3.re;
    }
  }
  // copy the real halves of rows 0 .. KC-1 from the team buffer to LDS
  __device__ void fill() const {
    const lreal l = (lreal)(lc);
#pragma unroll
    for (int k = 0; k < KC; ++k) {
      const CoefHalf<R> h = (re + k * T)[t];
      l[3 * k] = h.d1;
      l[3 * k + 1] = h.d2;
      l[3 * k + 2] = h.d3;
    }
  }
};

// Rows per lane whose real coefficient halves a team keeps in LDS (streamed
// modes), sized so that the blocks the VGPRs allow per CU still fit its
// 160 KB at the lane counts the tuning table gives each row count (complex64
// 21 x 256 lanes: 19 rows, 78 KB per block; 16 x 128: all 16 rows, 38 KB;
// complex128 64-lane teams: all rows up to 12, 296 B per lane), and the LDS
// bytes per lane KC rows take (an odd number of values of R, so that the
// lanes of a wavefront spread over the banks and every value is aligned to
// its size).  complex128 LDS teams of 8 and 10 rows choose their row count
// at launch (launch_rows).
template <class R, int M>
__host__ __device__ constexpr int coef_cache_rows() {
  if (sizeof(R) == 4) return M <= 16 ? M : (M == 21 ? 19 : (M <= 24 ? 17 : 0));
  return M <= 12 ? M : 0;
}
template <class R>
__host__ __device__ constexpr int coef_cache_lane_bytes(int kc) {
  const int n = kc * 3;  // values of R
  return n == 0 ? 0 : (n | 1) * static_cast<int>(sizeof(R));
}

// Water and bottom media of the lane's rows g0-1 .. g0+M (FP64), kept in a
// per-plan device buffer that persists between launches, laid out
// [ray][q][field][lane] (q = 0 .. M+1 for rows g0-1+q; fields 0-4 the water
// f1 f2 f3 kre kim, 5-9 the bottom ones) so that a team's accesses to one row
// are coalesced.  Both sides are evaluated after the self-starter and at
// profile range switches; when only the seabed row iz moves, the
// coefficients of the rows it touches are re-formed from it without
// interpolating the profiles again, and the TL output reads f3 from it.  (A
// per-lane private copy, filled at every launch, cost about 2.6 ms of scratch
// traffic per launch of the 10,000-ray Munk batch.)
template <int M>
struct MedCache {
  double *p;  // this lane's element of (q = 0, field 0)
  int T;      // lanes per team: the stride between fields
  __device__ double get(int q, int f) const { return p[static_cast<int64_t>(q * 10 + f) * T]; }
  __device__ void set(int q, int f, double v) const { p[static_cast<int64_t>(q * 10 + f) * T] = v; }
};

template <int M>
__device__ inline void put_med(const MedCache<M> &mc, int q, int o, const MediumRow &m) {
  mc.set(q, o + 0, m.f1);
  mc.set(q, o + 1, m.f2);
  mc.set(q, o + 2, m.f3);
  mc.set(q, o + 3, m.kre);
  mc.set(q, o + 4, m.kim);
}

template <int M>
__device__ void fill_media(const MedCache<M> &mc, const cpr_ray &P, const cpr_env &E, const double *env,
                           const RayState &st, int g0) {
  int wcur = -1, bcur = -1;  // profile segments carried down the lane's rows
#pragma unroll 1
  for (int q = 0; q < M + 2; ++q) {
    const int row = g0 - 1 + q;
    if (row <= P.nz + 1) {
      put_med(mc, q, 0, water_medium_at(P, E, env, st.ss_ind, row, &wcur));
      put_med(mc, q, 5, bottom_medium_at(P, E, env, st.sb_ind, row, &bcur));
    } else {
      MediumRow m;
      m.f1 = m.f2 = m.f3 = 1.0;
      m.kre = m.kim = 0.0;
      put_med(mc, q, 0, m);
      put_med(mc, q, 5, m);
    }
  }
}

// Diagonal Galerkin entry of the rows past nz in complex64: |q4 + d2|^2
// overflows FP32, so the reciprocal pivot of such a row is exactly zero and
// the row's elimination values vanish, as the masked pivot of complex128
// makes them (a mask per row would cost a lane-mask register per row).
constexpr float kPadPivot = 1e30f;

// Coefficients of lane row k (global row g0+k) from the cached media, water
// for rows <= iz (matrc.py:63-107) and the Galerkin stencil of
// fused_kernel.py:43-58.  complex64 rows past nz get (0, kPadPivot, 0).
template <class R, int M>
__device__ CoefRow<R> coef_from_cache(const MedCache<M> &mc, int k, int g0, int iz, int nz, double cfact) {
  if (std::is_same<R, float>::value && g0 + k > nz) {
    CoefRow<R> pad;
    pad.d1 = pad.d3 = cxm<R>(0, 0);
    pad.d2 = cxm<R>(static_cast<R>(kPadPivot), 0);
    return pad;
  }
  const double dfact = 1.0 / 12.0;
  const int o0 = (g0 + k - 1 <= iz) ? 0 : 5;
  const int o1 = (g0 + k <= iz) ? 0 : 5;
  const int o2 = (g0 + k + 1 <= iz) ? 0 : 5;
  const double m0[5] = {0.0, mc.get(k, o0 + 1), mc.get(k, o0 + 2), mc.get(k, o0 + 3), mc.get(k, o0 + 4)};
  const double m1[5] = {mc.get(k + 1, o1), mc.get(k + 1, o1 + 1), mc.get(k + 1, o1 + 2), mc.get(k + 1, o1 + 3),
                        mc.get(k + 1, o1 + 4)};
  const double m2[5] = {0.0, mc.get(k + 2, o2 + 1), mc.get(k + 2, o2 + 2), mc.get(k + 2, o2 + 3),
                        mc.get(k + 2, o2 + 4)};
  const double c1 = cfact * m1[0] * (m0[1] + m1[1]) * m0[2];
  const double c2 = -cfact * m1[0] * (m0[1] + 2.0 * m1[1] + m2[1]) * m1[2];
  const double c3 = cfact * m1[0] * (m1[1] + m2[1]) * m2[2];
  CoefRow<R> cr;
  cr.d1 = cxm<R>(static_cast<R>(c1 + dfact * (m0[3] + m1[3])), static_cast<R>(dfact * (m0[4] + m1[4])));
  cr.d2 = cxm<R>(static_cast<R>(c2 + dfact * (m0[3] + 6.0 * m1[3] + m2[3])),
                 static_cast<R>(dfact * (m0[4] + 6.0 * m1[4] + m2[4])));
  cr.d3 = cxm<R>(static_cast<R>(c3 + dfact * (m1[3] + m2[3])), static_cast<R>(dfact * (m1[4] + m2[4])));
  return cr;
}

// f3 (alpw or alpb) of lane row k from the cached media, for the TL output.
template <int M>
__device__ inline double f3_from_cache(const MedCache<M> &mc, int k, int g0, int iz) {
  return mc.get(k + 1, ((g0 + k <= iz) ? 0 : 5) + 2);
}

// Cell-mean intensity: adds |f3 u|^2 * rinv of the lane's rows to the output
// cells of range bin `bin`, depth bin j holding rows j*ndz+1 .. (j+1)*ndz.
// A lane's rows span consecutive depth bins; each partial sum is flushed
// with one atomic add when its bin ends (lanes share bins).
template <class R, int M>
__device__ void add_cells(const MarchArgs &a, const cpr_ray &P, int b, int bin, const cx<R> (&u)[M],
                          const MedCache<M> &mc, int g0, int mv, int iz, double rinv) {
  float *col = a.cells + static_cast<int64_t>(bin) * a.nvz_max * a.ld + b;
  const int ndz = P.ndz;
  int j = (g0 - 1) / ndz;
  int left = ndz - (g0 - 1) % ndz;
  float acc = 0.f;
#pragma unroll
  for (int k = 0; k < M; ++k) {
    if (k < mv) {
      const double f3 = f3_from_cache<M>(mc, k, g0, iz);
      const double i2 = static_cast<double>(u[k].re) * u[k].re + static_cast<double>(u[k].im) * u[k].im;
      acc += static_cast<float>(f3 * f3 * rinv * i2);
    }
    if (--left == 0 || k == M - 1) {
      if (j < P.nvz && acc != 0.f) atomicAdd(col + static_cast<int64_t>(j) * a.ld, acc);
      acc = 0.f;
      ++j;
      left = ndz;
    }
  }
}

// Coefficients a lane keeps in VGPRs are saved to the team buffer at the end
// of a launch and loaded at the start of the next one, in the layout of the
// streamed rows (real halves [ray][k][lane], then the imaginary halves), so a
// launch does not rebuild them from the media.
template <class R, int M>
__device__ void save_coefs(const Coefs<R, M, false> &C, CoefHalf<R> *re, CoefHalf<R> *im, int T, int t) {
#pragma unroll
  for (int k = 0; k < M; ++k) {
    (re + k * T)[t] = CoefHalf<R>{C.r[k].d1.re, C.r[k].d2.re, C.r[k].d3.re};
    (im + k * T)[t] = CoefHalf<R>{C.r[k].d1.im, C.r[k].d2.im, C.r[k].d3.im};
  }
}
template <class R, int M>
__device__ void load_coefs(Coefs<R, M, false> &C, const CoefHalf<R> *re, const CoefHalf<R> *im, int T, int t) {
#pragma unroll
  for (int k = 0; k < M; ++k) {
    const CoefHalf<R> a = (re + k * T)[t], b = (im + k * T)[t];
    C.r[k].d1 = cxm<R>(a.d1, b.d1);
    C.r[k].d2 = cxm<R>(a.d2, b.d2);
    C.r[k].d3 = cxm<R>(a.d3, b.d3);
  }
}

// Re-form the coefficients of lane rows klo..khi (clamped to the lane).
template <class R, int M, bool DS, int KC>
__device__ void rebuild_rows(Coefs<R, M, DS, KC> &C, const MedCache<M> &mc, int g0, int iz, int nz, double cfact,
                             int klo, int khi) {
  klo = klo < 0 ? 0 : klo;
  khi = khi > M - 1 ? M - 1 : khi;
  if constexpr (DS) {
#pragma unroll 1
    for (int k = klo; k <= khi; ++k) C.put(k, coef_from_cache<R, M>(mc, k, g0, iz, nz, cfact));
  } else {
    CoefRow<R> tmp[M];
#pragma unroll 1
    for (int k = klo; k <= khi; ++k) tmp[k] = coef_from_cache<R, M>(mc, k, g0, iz, nz, cfact);
#pragma unroll
    for (int k = 0; k < M; ++k)
      if (k >= klo && k <= khi) C.r[k] = tmp[k];
  }
}

template <class R, int M>
__device__ inline cx<R> pick(const cx<R> (&v)[M], int k) {
  cx<R> r = cxm<R>(0, 0);
#pragma unroll
  for (int q = 0; q < M; ++q)
    if (q == k) r = v[q];
  return r;
}

// Downward elimination with the spike through the lane's rows of one Pade
// factor (q [1 4 1] + D) x = (q [1 4 1] + rho D) u (f = rho), or, SC, in the
// scaled form of k_team_terms: (q [1 4 1] + D) y' = [1 4 1] u' and
// x' = y' + rc u' (f = rc = rho / c), whose right-hand side needs no product
// with D.  REAL: the wavefront's coefficients are real (every row's stencil
// lies in the water, where the wavenumber is real), so the entries q + d
// take real operands and streamed rows read half the bytes.  Rows past nz
// (k >= mv) get a zero reciprocal pivot (complex128: masked; complex64: from
// their kPadPivot diagonal), so c', w and d' vanish there and stay zero
// through the upward sweep and the back substitution.
//
// SC: the lane keeps, per row, e_k = d'_k + rc (u'_k + c'_k u'_{k+1}) for
// k < M - 2, e = d' + rc u' for k = M - 2 and d' for the last row: the upward
// sweep D_k = e_k - c'_k D_{k+1} then yields D_k + rc u'_k (the back
// substitution's x' without the stored u'), and the last row's u' is
// returned in ulast for x'_{M-1} = y'_{M-1} + rc u'_{M-1}.
template <bool REAL, bool SC, class R, int M, class C_t, class X_t>
__device__ inline void eliminate_down(LaneCol<R, M> &L, X_t &X, const C_t &C, cx<R> q,
                                      cx<R> q4, cx<R> f, cx<R> um, cx<R> u_next, int mv, cx<R> &ulast) {
  static_assert(M >= 2, "team lanes hold at least two rows");
  cx<R> cprev = cxm<R>(0, 0), wprev = cxm<R>(0, 0), dprev = cxm<R>(0, 0);
  using Row = std::conditional_t<REAL, CoefHalf<R>, CoefRow<R>>;
  const auto V = C.view();
  auto load = [&](int k) -> Row {
    if constexpr (REAL)
      return V.get_re(k);
    else
      return V.get(k);
  };
  // Coefficients of the next three rows in flight while a row is eliminated
  // (streamed rows come from L2/MALL; one row ahead left their latency
  // exposed: 111.4 -> 96.6 ms per complex64 Munk launch, 232.7 -> 219.0 ms
  // in complex128).  Rows past M reload the last row and go unused.
  Row p0 = load(0);
  Row p1 = load(M > 1 ? 1 : M - 1);
  Row p2 = load(M > 2 ? 2 : M - 1);
#pragma unroll
  for (int k = 0; k < M; ++k) {
    const Row cr = p0;
    p0 = p1;
    p1 = p2;
    if (k + 3 < M) p2 = load(k + 3);
    const cx<R> uc = L.u[k];
    const cx<R> up = (k + 1 < M) ? L.u[k + 1] : u_next;
    cx<R> r1, r2, r3, du;
    if constexpr (REAL) {
      r1 = cxm<R>(q.re + cr.d1, q.im);
      r2 = cxm<R>(q4.re + cr.d2, q4.im);
      r3 = cxm<R>(q.re + cr.d3, q.im);
      if constexpr (!SC)
        du = cxm<R>(cr.d1 * um.re + cr.d2 * uc.re + cr.d3 * up.re, cr.d1 * um.im + cr.d2 * uc.im + cr.d3 * up.im);
    } else {
      r1 = q + cr.d1;
      r2 = q4 + cr.d2;
      r3 = q + cr.d3;
      if constexpr (!SC) du = cfma(cfma(cr.d1 * um, cr.d2, uc), cr.d3, up);
    }
    const cx<R> s = um + up;
    const cx<R> ts = cxm<R>(s.re + R(4) * uc.re, s.im + R(4) * uc.im);
    cx<R> rhs;
    if constexpr (SC)
      rhs = ts;
    else
      rhs = cfma(q * ts, f, du);  // f = rho
    const cx<R> den = k == 0 ? r2 : cfms(r2, r1, cprev);
    R m = rcp_real<R>(den.re * den.re + den.im * den.im);
    if constexpr (!std::is_same<R, float>::value) m = k < mv ? m : R(0);
    const cx<R> inv = cxm<R>(den.re * m, -(den.im * m));
    const cx<R> cc = r3 * inv;
    // first row: the previous unknown is y_{t-1} itself (w = -1, c' = d' = 0)
    const cx<R> wk = k == 0 ? r1 * inv : cnmul(r1 * wprev, inv);
    const cx<R> dk = k == 0 ? rhs * inv : cfms(rhs, r1, dprev) * inv;
    X.set(k, cc, wk);
    if constexpr (!SC)
      L.u[k] = dk;
    else if (k < M - 2)
      L.u[k] = cfma(dk, f, cfma(uc, cc, up));  // f = rc
    else if (k == M - 2)
      L.u[k] = cfma(dk, f, uc);
    else {
      L.u[k] = dk;
      ulast = uc;
    }
    cprev = cc;
    wprev = wk;
    dprev = dk;
    um = uc;
  }
}

template <class R, int M, int TW, bool DS, int NR, int KC>
__device__ void march_team_body(const MarchArgs &a, int b, int t, int T, Slot<R> *lds, R *lc, int s0, int s1) {
  const cpr_ray P = make_uniform(a.rays[b]);
  const cpr_env E = make_uniform(a.envs[b]);
  RayState st = make_uniform(a.state[b]);
  Team<R, TW> team{t, T, lds, 0};
  team.init();
  const int nz = P.nz;
  const int g0 = 1 + t * M;
  int mv = nz - g0 + 1;  // valid rows of this lane
  mv = mv < 0 ? 0 : (mv > M ? M : mv);
  LaneCol<R, M> L;
  Coefs<R, M, DS, KC> C;
  Elim<R, M, NR> X;
  if constexpr (NR < M) {
    // the tail rows follow the team's exchange and reduced-system buffers
    char *const base = reinterpret_cast<char *>(lds) + (team_lds_bytes<R>(T) + 15) / 16 * 16;
    X.tail = reinterpret_cast<ElimPair<R> *>(base + t * elim_tail_lane_bytes<R, M>());
  }
  const double cfact = 0.5 / (P.dz * P.dz);
  CoefHalf<R> *const cre = static_cast<CoefHalf<R> *>(a.dbuf) + static_cast<int64_t>(b) * M * T;
  CoefHalf<R> *const cim = static_cast<CoefHalf<R> *>(a.dbuf) + (static_cast<int64_t>(a.nrays) + b) * M * T;
  if constexpr (DS) {
    C.re = cre;
    C.im = cim;
    C.T = T;
    C.t = t;
    C.lc = lc;
  }
  const MedCache<M> mc{static_cast<double *>(a.mbuf) + static_cast<int64_t>(b) * (M + 2) * 10 * T + t, T};
  // the field in the team layout [ray][k][lane] (rows 1 + t M + k): a team's
  // loads and stores of one k are contiguous
  cx<R> *const ug = static_cast<cx<R> *>(a.u) + static_cast<int64_t>(b) * M * T + t;
#pragma unroll
  for (int k = 0; k < M; ++k) L.u[k] = k < mv ? ug[static_cast<int64_t>(k) * T] : cxm<R>(0, 0);
  // receiver rows: owner lane / slot of rows ir and ir+1
  const double ri = P.zr / P.dz;
  const int ir = static_cast<int>(floor(ri));
  const double dir = ri - ir;
  const int end = s1 < P.nsteps ? s1 : P.nsteps;
  // halo rows: the previous lane's last row and the next lane's first,
  // exchanged once per launch; after each factor they come from its back
  // substitution (the field does not change between steps otherwise)
  cx<R> um, u_next;
  team.halo(L.u[M - 1], L.u[0], um, u_next);
  // The media and the coefficients persist between launches: they are
  // rebuilt in full only after the self-starter (dirty) and at profile
  // switches; VGPR-resident coefficients come back from the team buffer.
  bool full = st.dirty != 0;
  if constexpr (!DS) {
    if (!full && s0 < end) load_coefs<R, M>(C, cre, cim, T, t);
  }
  if constexpr (DS && KC > 0) {
    // LDS coefficient rows from the team buffer (a full rebuild writes them)
    if (!full && s0 < end) C.fill();
  }
  for (int step = s0; step < end; ++step) {
    const double r = (step + 1) * P.dr;
    const int jz = st.iz;
    const int fl = __builtin_amdgcn_readfirstlane(updat_ray(P, E, a.env, st, r));
    st = make_uniform(st);
    // f changes on rows (lo, hi] when iz moves; d changes on rows lo .. hi+1
    const int lo = jz < st.iz ? jz : st.iz, hi = jz > st.iz ? jz : st.iz;
    if (full || (fl & 2)) {
      fill_media<M>(mc, P, E, a.env, st, g0);
      rebuild_rows<R, M, DS>(C, mc, g0, st.iz, nz, cfact, 0, M - 1);
      full = false;
      st.dirty = 0;
    } else if ((fl & 1) && g0 <= hi + 1 && g0 + M - 1 >= lo) {
      rebuild_rows<R, M, DS>(C, mc, g0, st.iz, nz, cfact, lo - g0, hi + 1 - g0);
    }
    // every row stencil of this wavefront above the seabed row: real coefficients
    const bool wreal = __builtin_amdgcn_ballot_w64(g0 + M > st.iz) == 0;
    // Pade factor (A + p2 D) x = (A + p1 D) u with A = (k0^2/6) [1 4 1],
    // divided by p2: (q [1 4 1] + D) x = (q [1 4 1] + rho D) u.  complex64
    // applies it as x = rho u + c y with (q [1 4 1] + D) y = [1 4 1] u,
    // c = (1 - rho) q, on the field scaled by the running product of the c's
    // (x' = y' + rc u', rc = rho / c: 12 fewer VALU instructions per complex
    // row and 6 per real one); complex128 keeps the direct form, whose fewer
    // live values across the reduced system spill less (the scaled form
    // measured 215.8 instead of 196.9 ms per 100-step Munk launch).  The
    // constants and the renormalisation factors rn are formed once per plan
    // in FP64 (k_team_terms) and read here through scalar loads.
    constexpr bool SC = std::is_same<R, float>::value;
    const R *qt = static_cast<const R *>(a.qtab) + (static_cast<int64_t>(b) * 3 + st.pset) * P.np * kTermStride;
    for (int j = 0; j < P.np; ++j) {
      const R *qj = qt + kTermStride * j;
      const cx<R> q = cxm<R>(qj[0], qj[1]);
      const cx<R> q4 = cxm<R>(qj[2], qj[3]);
      const cx<R> rc = SC ? cxm<R>(qj[4], qj[5]) : cxm<R>(qj[8], qj[9]);  // rc, or rho
      // 1. downward elimination with the spike
      cx<R> ulast = cxm<R>(0, 0);
      if (wreal)
        eliminate_down<true, SC>(L, X, C, q, q4, rc, um, u_next, mv, ulast);
      else
        eliminate_down<false, SC>(L, X, C, q, q4, rc, um, u_next, mv, ulast);
      // 2. upward sweep: x_k = D_k - W_k y_{t-1} - C_k y_t for k < M-1
#pragma unroll
      for (int k = M - 3; k >= 0; --k) {
        const cx<R> ck = X.c(k);
        L.u[k] = cfms(L.u[k], ck, L.u[k + 1]);
        X.set(k, cnmul(ck, X.c(k + 1)), cfms(X.w(k), ck, X.w(k + 1)));
      }
      // 3. reduced equation a y_{t-1} + b y_t + c y_{t+1} = f, then PCR.  The
      // next lane's D_0 carries rc u'_0 (its first row's field, u_next here).
      cx<R> nDx, nW, nC;
      team.template next3<!DS>(L.u[0], X.w(0), X.c(0), nDx, nW, nC);
      const cx<R> nD = SC ? cfms(nDx, rc, u_next) : nDx;
      const cx<R> ce = X.c(M - 1);
      cx<R> ea = t > 0 ? X.w(M - 1) : cxm<R>(0, 0);
      const cx<R> eb = cfms(cxm<R>(1, 0), ce, nW);
      cx<R> ec = cnmul(ce, nC);
      cx<R> ef = cfms(L.u[M - 1], ce, nD);
      {
        const cx<R> ib = fast_rcp(eb);
        ea = ea * ib;
        ec = ec * ib;
        ef = ef * ib;
      }
      cx<R> yc, yl, yr;
      if constexpr (TW > 0) {
        // PCR over the 64-lane wavefront, unrolled so that each level's
        // distance s is a constant; complex64 exchanges through the team's
        // LDS buffer (in-order LDS operations of one wavefront, as in
        // solve_cr), complex128 with shuffles
        if constexpr (sizeof(R) == 4) {
          pcr64_lds(team.lds, t, ea, ec, ef);
          yc = ef;
          cx<R> *const A = reinterpret_cast<cx<R> *>(team.lds - kPad) + kPad;
          A[t] = yc;
          asm volatile("" ::: "memory");
          yl = A[t - 1];
          yr = A[t + 1];
          asm volatile("" ::: "memory");
        } else {
#pragma unroll
          for (int l = 0; l < 6; ++l) {
            const int s = 1 << l;
            if (s >= T) break;
            cx<R> am, cm, fm, ap, cp, fp;
            team.pcr(s, ea, ec, ef, am, cm, fm, ap, cp, fp);
            reduce_step(ea, ec, ef, am, cm, fm, ap, cp, fp);
          }
          yc = ef;
          team.neighbours(yc, yl, yr);
        }
      } else {
        yc = team.solve_cr(ea, ec, ef, yl, yr);
      }
      // 4. back substitution (rows past nz come out zero).  The next lane's
      // new first row is formed here from its (D, W, C) of step 3 exactly as
      // that lane forms it, so the next factor needs no halo exchange.
#pragma unroll
      for (int k = 0; k < M - 1; ++k) L.u[k] = xfms(xfms(L.u[k], X.w(k), yl), X.c(k), yc);
      if constexpr (SC) {
        L.u[M - 1] = xfma(yc, rc, ulast);
        um = xfma(yl, rc, um);
      } else {
        L.u[M - 1] = yc;
        um = yl;
      }
      u_next = xfms(xfms(nDx, nW, yc), nC, yr);
      if constexpr (SC) {
        // 5. multiply the scale back in after the marked terms (always the last)
        const cx<R> rn = cxm<R>(qj[6], qj[7]);
        if (rn.re != R(0) || rn.im != R(0)) {
#pragma unroll
          for (int k = 0; k < M; ++k) L.u[k] = xmul(L.u[k], rn);
          um = xmul(um, rn);
          u_next = xmul(u_next, rn);
        }
      }
    }
    if (a.cells) {
      const int bin = st.tlc + 1;
      if (bin < P.nvr) add_cells<R, M>(a, P, b, bin, L.u, mc, g0, mv, st.iz, 1.0 / ((step + 2) * P.dr));
    }
    // output (outpt.py:11-58): complex pressure only; TL is formed by a
    // post-pass over the whole output array
    st.mdr += 1;
    if (st.mdr == P.ndr) {
      st.mdr = 0;
      st.tlc += 1;
      if (st.tlc < P.nvr) {
        const int ta = ir >= 1 ? (ir - 1) / M : 0, ka = ir >= 1 ? (ir - 1) % M : 0;
        const int tb = ir / M, kb = ir % M;
        // only the lanes that own rows ir and ir+1 pick them out
        cx<R> va = cxm<R>(0, 0), vb = cxm<R>(0, 0);
        R fa = R(0), fb = R(0);
        if (t == ta) {
          va = pick<R, M>(L.u, ka);
          fa = static_cast<R>(f3_from_cache<M>(mc, ka, g0, st.iz));
        }
        if (t == tb) {
          vb = pick<R, M>(L.u, kb);
          fb = static_cast<R>(f3_from_cache<M>(mc, kb, g0, st.iz));
        }
        cx<R> ua, ub;
        R f3a, f3b;
        if constexpr (TW > 0) {
          ua = cxm<R>(__shfl(va.re, ta, TW), __shfl(va.im, ta, TW));
          f3a = __shfl(fa, ta, TW);
          const int tbs = tb < T ? tb : 0;
          ub = cxm<R>(__shfl(vb.re, tbs, TW), __shfl(vb.im, tbs, TW));
          f3b = __shfl(fb, tbs, TW);
        } else {
          Slot<R> *B = team.next_buf();
          if (t == ta) {
            B[0].a = va;
            B[0].f = cxm<R>(fa, 0);
          }
          if (t == tb) {
            B[0].c = vb;
            B[1].f = cxm<R>(fb, 0);
          }
          __syncthreads();
          ua = B[0].a;
          ub = B[0].c;
          f3a = B[0].f.re;
          f3b = B[1].f.re;
        }
        if (ir < 1) {
          ua = cxm<R>(0, 0);
          f3a = R(0);
        }
        if (ir + 1 > nz) {
          ub = cxm<R>(0, 0);
          f3b = R(0);
        }
        if (t == 0) {
          const double pre = (1.0 - dir) * static_cast<double>(f3a) * static_cast<double>(ua.re) +
                             dir * static_cast<double>(f3b) * static_cast<double>(ub.re);
          const double pim = (1.0 - dir) * static_cast<double>(f3a) * static_cast<double>(ua.im) +
                             dir * static_cast<double>(f3b) * static_cast<double>(ub.im);
          const int64_t o = static_cast<int64_t>(st.tlc) * a.ld + b;
          a.cpl[2 * o] = pre;
          a.cpl[2 * o + 1] = pim;
        }
        if (a.grids) {
#pragma unroll
          for (int k = 0; k < M; ++k) {
            const int g = g0 + k;
            if (k < mv && g % P.ndz == 0 && g / P.ndz <= P.nvz) {
              const double f3 = f3_from_cache<M>(mc, k, g0, st.iz);
              const int64_t og = (static_cast<int64_t>(st.tlc) * a.nvz_max + (g / P.ndz - 1)) * a.ld + b;
              a.cpg[2 * og] = static_cast<double>(L.u[k].re) * f3;
              a.cpg[2 * og + 1] = static_cast<double>(L.u[k].im) * f3;
            }
          }
        }
      }
    }
  }
#pragma unroll
  for (int k = 0; k < M; ++k)
    if (k < mv) ug[static_cast<int64_t>(k) * T] = L.u[k];
  if constexpr (!DS) {
    if (s0 < end) save_coefs<R, M>(C, cre, cim, T, t);
  }
  if (t == 0) a.state[b] = st;
}

}  // namespace

// Per-term constants of the team march for every ray, Pade set and term,
// formed in FP64 and rounded to the march precision.  The factor
// (A + p2 D) x = (A + p1 D) u with A = (k0^2/6) [1 4 1], divided by p2, is
// (q [1 4 1] + D) x = (q [1 4 1] + rho D) u with q = k0^2 / (6 p2) and
// rho = p1 / p2; its right-hand side matrix is rho (q [1 4 1] + D) +
// (1 - rho) q [1 4 1], so x = rho u + c y with c = (1 - rho) q and
// (q [1 4 1] + D) y = [1 4 1] u.  The march carries the field divided by the
// running product of the c's, x' = y' + (rho / c) u', and multiplies it back
// after the terms marked here (rn != 0: the product since the last such
// term), always after the last one and otherwise before the product would
// leave [2^-30, 2^30].  Stored per term: q, 4 q, rho / c, rn, rho (re, im
// each); complex128 uses q, 4 q and rho (the direct form).
template <class R>
__global__ void k_team_terms(MarchArgs a, R *qtab) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= a.nrays) return;
  const cpr_ray P = a.rays[b];
  const double a1d = P.k0 * P.k0 / 6.0;
  for (int set = 0; set < 3; ++set) {
    const double *pd = a.pade + (static_cast<int64_t>(b) * 3 + set) * 2 * P.np * 2;
    R *o = qtab + (static_cast<int64_t>(b) * 3 + set) * P.np * kTermStride;
    // q, rho and c of term j
    auto term = [&](int j, double &qr, double &qi, double &rr, double &ri, double &cr, double &ci) {
      const double p1r = pd[2 * j], p1i = pd[2 * j + 1];
      const double p2r = pd[2 * (P.np + j)], p2i = pd[2 * (P.np + j) + 1];
      const double n2 = 1.0 / (p2r * p2r + p2i * p2i);
      const double ir = p2r * n2, ii = -p2i * n2;
      qr = a1d * ir;
      qi = a1d * ii;
      rr = p1r * ir - p1i * ii;
      ri = p1r * ii + p1i * ir;
      // a factor with p1 = p2 is the identity (c = 0): a c of 1e-12 q
      // changes x by 1e-12 of the correction term and keeps rho / c finite
      double er = 1.0 - rr, ei = -ri;
      if (er * er + ei * ei < 1e-24) {
        er = 1e-12;
        ei = 0.0;
      }
      cr = er * qr - ei * qi;
      ci = er * qi + ei * qr;
    };
    const double lo2 = 0x1p-60, hi2 = 0x1p60;  // squared bounds of |product|
    double sr = 1.0, si = 0.0;                  // product of c since the last renormalisation
    for (int j = 0; j < P.np; ++j) {
      double qr, qi, rr, ri, cr, ci;
      term(j, qr, qi, rr, ri, cr, ci);
      const double nc = 1.0 / (cr * cr + ci * ci);
      o[kTermStride * j + 0] = static_cast<R>(qr);
      o[kTermStride * j + 1] = static_cast<R>(qi);
      o[kTermStride * j + 2] = static_cast<R>(4.0 * qr);
      o[kTermStride * j + 3] = static_cast<R>(4.0 * qi);
      o[kTermStride * j + 4] = static_cast<R>((rr * cr + ri * ci) * nc);  // rho / c
      o[kTermStride * j + 5] = static_cast<R>((ri * cr - rr * ci) * nc);
      const double tr = sr * cr - si * ci;
      si = sr * ci + si * cr;
      sr = tr;
      bool renorm = j == P.np - 1;
      if (!renorm) {
        double q2r, q2i, r2r, r2i, c2r, c2i;
        term(j + 1, q2r, q2i, r2r, r2i, c2r, c2i);
        const double xr = sr * c2r - si * c2i, xi = sr * c2i + si * c2r;
        const double m2 = xr * xr + xi * xi;
        renorm = !(m2 >= lo2 && m2 <= hi2);
      }
      o[kTermStride * j + 6] = static_cast<R>(renorm ? sr : 0.0);
      o[kTermStride * j + 7] = static_cast<R>(renorm ? si : 0.0);
      o[kTermStride * j + 8] = static_cast<R>(rr);
      o[kTermStride * j + 9] = static_cast<R>(ri);
      if (renorm) {
        sr = 1.0;
        si = 0.0;
      }
    }
  }
}
template __global__ void k_team_terms<float>(MarchArgs, float *);
template __global__ void k_team_terms<double>(MarchArgs, double *);

// Shuffle mode: one 64-lane team per wavefront, four per 256-thread block.
// At least two waves per SIMD (256 VGPRs): a lone wave cannot hide the
// latency of its shuffles and reciprocals (measured 1.6x on 8- and 12-row
// lanes of complex64 against the 1-wave build, which took up to 314 VGPRs).
// complex64 teams solve their reduced system through an LDS buffer of
// 64 + 2 kPad slots each (wave_team_lds_bytes per block).
template <class R, int M, bool DS, int KC>
__global__ void __launch_bounds__(256, 2) k_march_team_w(MarchArgs a, int s0, int s1) {
  extern __shared__ char smem[];
  const int team = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + threadIdx.x / 64);
  if (team >= a.team_nrays) return;
  const int b = __builtin_amdgcn_readfirstlane(a.team_rays[team]);
  Slot<R> *const pb =
      sizeof(R) == 4 ? reinterpret_cast<Slot<R> *>(smem) + (threadIdx.x / 64) * (64 + 2 * kPad) + kPad : nullptr;
  R *const lc = KC > 0 ? reinterpret_cast<R *>(smem + wave_team_lds_bytes<R>() + threadIdx.x * coef_cache_lane_bytes<R>(KC))
                       : nullptr;
  march_team_body<R, M, 64, DS, M, KC>(a, b, threadIdx.x % 64, 64, pb, lc, s0, s1);
}

// LDS mode: one team per block of T = blockDim.x lanes (a multiple of 64),
// with the elimination rows NR .. M-1 and the real coefficient halves of rows
// 0 .. KC-1 in LDS (never both: they start at the same offset).
template <class R, int M, bool DS, int NR, int KC>
__global__ void __launch_bounds__(512) k_march_team_l(MarchArgs a, int s0, int s1) {
  // both would start right after the team's buffers
  static_assert(NR == M || KC == 0, "LDS elimination rows and LDS coefficient rows are not combined");
  extern __shared__ char smem[];
  const int team = blockIdx.x;
  if (team >= a.team_nrays) return;
  const int b = __builtin_amdgcn_readfirstlane(a.team_rays[team]);
  R *const lc = KC > 0 ? reinterpret_cast<R *>(smem + (team_lds_bytes<R>(blockDim.x) + 15) / 16 * 16 +
                                                 threadIdx.x * coef_cache_lane_bytes<R>(KC))
                       : nullptr;
  march_team_body<R, M, 0, DS, NR, KC>(a, b, threadIdx.x, blockDim.x, reinterpret_cast<Slot<R> *>(smem), lc, s0,
                                       s1);
}

namespace {

constexpr size_t kMaxLdsBytes = 160 * 1024;  // LDS of one CU (gfx950)

// An LDS-mode team with the real coefficient halves of rows 0 .. KC-1 in LDS,
// launched only if the block's LDS leaves the CU the blocks the VGPRs allow.
template <class R, int M, bool DS, int KC>
bool launch_cached(const MarchArgs &a, int lanes, size_t base, size_t blocks, int s0, int s1, hipStream_t st) {
  const size_t bytes = (base + 15) / 16 * 16 + static_cast<size_t>(lanes) * coef_cache_lane_bytes<R>(KC);
  if (bytes * blocks > kMaxLdsBytes) return false;
  hipLaunchKernelGGL((k_march_team_l<R, M, DS, M, KC>), dim3(a.team_nrays), dim3(lanes), bytes, st, a, s0, s1);
  return true;
}

template <class R, int M, bool DS>
int launch_rows(const MarchArgs &a, int lanes, int s0, int s1, hipStream_t st) {
  constexpr int KC = DS ? coef_cache_rows<R, M>() : 0;
  if (lanes == 64) {
    const int grid = (a.team_nrays + 3) / 4;
    constexpr size_t lds = wave_team_lds_bytes<R>();
    // two blocks per CU (the VGPRs allow two waves per SIMD)
    constexpr size_t with_cache = lds + 256 * static_cast<size_t>(coef_cache_lane_bytes<R>(KC));
    if constexpr (KC > 0 && 2 * with_cache <= kMaxLdsBytes)
      hipLaunchKernelGGL((k_march_team_w<R, M, DS, KC>), dim3(grid), dim3(256), with_cache, st, a, s0, s1);
    else
      hipLaunchKernelGGL((k_march_team_w<R, M, DS, 0>), dim3(grid), dim3(256), lds, st, a, s0, s1);
  } else {
    if (lanes % 64 || lanes > 512) return -1;
    // The LDS rows of the elimination arrays only where their LDS leaves the
    // CU as many teams as the VGPRs allow (two waves per SIMD: eight waves per
    // CU): a 128-lane complex128 team of 12 rows with them ran 19.7 instead of
    // 23.1 G cell-steps/s (three blocks per CU instead of four).
    constexpr int NR = elim_regs<R, M>();
    const size_t base = team_lds_bytes<R>(lanes);
    const size_t with_tail = (base + 15) / 16 * 16 + static_cast<size_t>(lanes) * elim_tail_lane_bytes<R, M>();
    const int waves = lanes / 64;
    const size_t blocks = waves >= 8 ? 1 : static_cast<size_t>(8 / waves);
    if constexpr (NR < M) {
      if (with_tail * blocks <= kMaxLdsBytes) {
        hipLaunchKernelGGL((k_march_team_l<R, M, DS, NR, 0>), dim3(a.team_nrays), dim3(lanes), with_tail, st, a, s0, s1);
        return hipGetLastError() == hipSuccess ? 0 : -1;
      }
    }
    if constexpr (sizeof(R) == 8 && DS && NR == M && M <= 10) {
      // complex128 8 and 10 rows: 7, 6 or 4 rows, the most that fit (512,
      // 256 and 128 lanes); 10 x 256 lanes ran 29.2 instead of 27.6 G
      // cell-steps/s, 8 x 128 26.2 instead of 24.3 G; 11 x 512 ran 24.1
      // instead of 25.2 G and keeps its coefficients in the team buffer
      if (launch_cached<R, M, DS, (M < 7 ? M : 7)>(a, lanes, base, blocks, s0, s1, st) ||
          launch_cached<R, M, DS, (M < 6 ? M : 6)>(a, lanes, base, blocks, s0, s1, st) ||
          launch_cached<R, M, DS, (M < 4 ? M : 4)>(a, lanes, base, blocks, s0, s1, st))
        return hipGetLastError() == hipSuccess ? 0 : -1;
    } else if constexpr (sizeof(R) == 4 && KC > 0) {
      if (launch_cached<R, M, DS, KC>(a, lanes, base, blocks, s0, s1, st)) return hipGetLastError() == hipSuccess ? 0 : -1;
    }
    if (base > kMaxLdsBytes) return -1;
    hipLaunchKernelGGL((k_march_team_l<R, M, DS, M, 0>), dim3(a.team_nrays), dim3(lanes), base, st, a, s0, s1);
  }
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace

// Rows per lane and where the lane keeps its state: the field is a register
// array, and so are the elimination arrays except for the deep LDS teams'
// tail rows (elim_regs); complex64 keeps the Galerkin coefficients in VGPRs
// up to 12 rows and streams them from a per-team buffer from 14 rows on,
// complex128 up to 6 and from 8 rows on (40 x 512 lanes hold the deepest
// columns, 20,480 rows; complex64 64 x 512 hold 32,768).
bool team_rows_stream(int precision, int rows) { return precision == CPR_C64 ? rows >= 14 : rows >= 8; }

// Rows per lane built for each precision.  Between the powers of two, 14, 20,
// 21 and 28 (complex64) and 10 and 11 (complex128) let a column fill its
// lanes with few padding rows (nz = 5303: 21 x 256 = 5376 rows instead of
// 24 x 256 = 6144), or fill all eight waves of a CU (11 x 512 lanes; complex64
// 14 x 512 lanes for 7,168 rows, 50.9 against 44.7 G cell-steps/s for the
// seven waves of 16 x 448, though 28 x 256 at 58.0 G is the tuned shape).
constexpr int kRows64[] = {4, 8, 12, 14, 16, 20, 21, 24, 28, 32, 40, 48, 64};
constexpr int kRows128[] = {2, 4, 6, 8, 10, 11, 12, 14, 16, 20, 24, 32, 40};

bool team_shape_supported(int precision, int rows, int lanes) {
  const bool lanes_ok = lanes >= 64 && lanes <= 512 && lanes % 64 == 0;
  if (!lanes_ok) return false;
  if (precision == CPR_C64) {
    for (int r : kRows64)
      if (r == rows) return true;
    return false;
  }
  for (int r : kRows128)
    if (r == rows) return true;
  return false;
}

void team_auto_shape(int precision, int nz, int *rows, int *lanes) {
  const int *cand = precision == CPR_C64 ? kRows64 : kRows128;
  const int ncand = precision == CPR_C64 ? static_cast<int>(sizeof(kRows64) / sizeof(int))
                                         : static_cast<int>(sizeof(kRows128) / sizeof(int));
  int best_r = 0, best_l = 0;
  double best_cost = -1;
  for (int i = 0; i < ncand; ++i) {
    const int m = cand[i];
    int l = (nz + m - 1) / m;
    l = l <= 64 ? 64 : (l + 63) / 64 * 64;
    if (l > 512) continue;
    // padded row work, a reduced-system term per lane, and premiums fitted
    // to the autotuner's rates: streamed complex128 coefficients (complex64
    // keeps its streamed rows in LDS), register spills past 24 rows per lane
    // in complex64 (32 in teams of 192 lanes or more, which keep the last
    // elimination rows of such lanes in LDS) or 12 in complex128, and teams
    // of 3 or 5-7 waves, which load the four SIMDs of a CU unevenly (10-25 %
    // slower than 4 or 8 waves); it picks the tuning table's row count at 36
    // of its 38 depths
    const bool c64 = precision == CPR_C64;
    const int m_fit = c64 ? (l >= 192 ? 32 : 24) : 12;
    double prem = (!c64 && team_rows_stream(precision, m)) ? 1.15 : 1.0;
    if (m > m_fit) prem *= std::pow(static_cast<double>(m) / m_fit, c64 ? 3.0 : 1.4);
    if (!c64 && m > 8) prem *= std::pow(m / 8.0, 0.6);  // complex128 lanes past 8 rows already spill
    const int waves = l / 64;
    if (waves == 3 || (waves >= 5 && waves <= 7)) prem *= 1.15;
    const double cost = static_cast<double>(l) * m * prem + 8.0 * l;
    if (best_cost < 0 || cost < best_cost) {
      best_cost = cost;
      best_r = m;
      best_l = l;
    }
  }
  *rows = best_r;
  *lanes = best_l;
}

int64_t team_mbuf_bytes(int rows, int lanes, int nrays) {
  return static_cast<int64_t>(nrays) * (rows + 2) * 10 * lanes * static_cast<int64_t>(sizeof(double));
}

int64_t team_dbuf_bytes(int precision, int rows, int lanes, int nrays) {
  const int64_t row_bytes = precision == CPR_C64 ? sizeof(CoefRow<float>) : sizeof(CoefRow<double>);
  return static_cast<int64_t>(nrays) * rows * lanes * row_bytes;
}


int team_launch(const MarchArgs &a, int precision, int rows, int lanes, int s0, int s1, hipStream_t st) {
  if (precision == CPR_C64) {
    switch (rows) {
      case 4:
        return launch_rows<float, 4, false>(a, lanes, s0, s1, st);
      case 8:
        return launch_rows<float, 8, false>(a, lanes, s0, s1, st);
      case 12:
        return launch_rows<float, 12, false>(a, lanes, s0, s1, st);
      case 14:
        return launch_rows<float, 14, true>(a, lanes, s0, s1, st);
      case 16:
        return launch_rows<float, 16, true>(a, lanes, s0, s1, st);
      case 20:
        return launch_rows<float, 20, true>(a, lanes, s0, s1, st);
      case 21:
        return launch_rows<float, 21, true>(a, lanes, s0, s1, st);
      case 24:
        return launch_rows<float, 24, true>(a, lanes, s0, s1, st);
      case 28:
        return launch_rows<float, 28, true>(a, lanes, s0, s1, st);
      case 32:
        return launch_rows<float, 32, true>(a, lanes, s0, s1, st);
      case 40:
        return launch_rows<float, 40, true>(a, lanes, s0, s1, st);
      case 48:
        return launch_rows<float, 48, true>(a, lanes, s0, s1, st);
      case 64:
        return launch_rows<float, 64, true>(a, lanes, s0, s1, st);
    }
  } else {
    switch (rows) {
      case 2:
        return launch_rows<double, 2, false>(a, lanes, s0, s1, st);
      case 4:
        return launch_rows<double, 4, false>(a, lanes, s0, s1, st);
      case 6:
        return launch_rows<double, 6, false>(a, lanes, s0, s1, st);
      case 8:
        return launch_rows<double, 8, true>(a, lanes, s0, s1, st);
      case 10:
        return launch_rows<double, 10, true>(a, lanes, s0, s1, st);
      case 11:
        return launch_rows<double, 11, true>(a, lanes, s0, s1, st);
      case 12:
        return launch_rows<double, 12, true>(a, lanes, s0, s1, st);
      case 14:
        return launch_rows<double, 14, true>(a, lanes, s0, s1, st);
      case 16:
        return launch_rows<double, 16, true>(a, lanes, s0, s1, st);
      case 20:
        return launch_rows<double, 20, true>(a, lanes, s0, s1, st);
      case 24:
        return launch_rows<double, 24, true>(a, lanes, s0, s1, st);
      case 32:
        return launch_rows<double, 32, true>(a, lanes, s0, s1, st);
      case 40:
        return launch_rows<double, 40, true>(a, lanes, s0, s1, st);
    }
  }
  return -1;
}
