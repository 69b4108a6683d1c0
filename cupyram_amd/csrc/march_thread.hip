// Thread-per-ray march kernels for gfx950.
//
// One lane owns one ray; the column arrays are depth-major with the ray index
// fastest ([row][ld]), so a wavefront of 64 consecutive rays reads and writes
// 64 consecutive elements per row (coalesced).  Each launch marches a range
// of steps with no host involvement: range update (updat), medium
// rematerialisation only where it changed (profl + matrc medium selection),
// np product-form Thomas solves (fused_kernel.py:72-153) and the TL output
// (outpt.py:11-58).
//
// k_selfs runs the PyRAM self-starter (cupyram/cupyram.py:1001-1063) in FP64
// whatever the march precision: a delta at zs, two (1-X)^-1 solves, then the
// ip=2 operator, followed by the first output at r = dr.
#include <hip/hip_runtime.h>

#include "march_common.h"

namespace {

template <class R>
struct Col {
  R *f1, *f2, *f3;
  cx<R> *ksq, *u, *cp, *dp;
  int64_t ld;
  int b;
  __device__ R &F1(int i) const { return f1[i * ld + b]; }
  __device__ R &F2(int i) const { return f2[i * ld + b]; }
  __device__ R &F3(int i) const { return f3[i * ld + b]; }
  __device__ cx<R> &K(int i) const { return ksq[i * ld + b]; }
  __device__ cx<R> &U(int i) const { return u[i * ld + b]; }
  __device__ cx<R> &CP(int i) const { return cp[i * ld + b]; }
  __device__ cx<R> &DP(int i) const { return dp[i * ld + b]; }
};

template <class R>
__device__ Col<R> col_of(const MarchArgs &a, int b) {
  Col<R> c;
  c.f1 = static_cast<R *>(a.f1);
  c.f2 = static_cast<R *>(a.f2);
  c.f3 = static_cast<R *>(a.f3);
  c.ksq = static_cast<cx<R> *>(a.ksq);
  c.u = static_cast<cx<R> *>(a.u);
  c.cp = static_cast<cx<R> *>(a.cp);
  c.dp = static_cast<cx<R> *>(a.dp);
  c.ld = a.ld;
  c.b = b;
  return c;
}

// Write the medium of rows lo..hi (water above iz, bottom below).
template <class R>
__device__ void materialize(const Col<R> &c, const cpr_ray &P, const cpr_env &E, const double *env,
                            const RayState &st, int lo, int hi) {
  int wcur = -1, bcur = -1;  // profile segments carried down the rows
  for (int i = lo; i <= hi; ++i) {
    const MediumRow m = medium_at(P, E, env, st.ss_ind, st.sb_ind, st.iz, i, &wcur, &bcur);
    c.F1(i) = static_cast<R>(m.f1);
    c.F2(i) = static_cast<R>(m.f2);
    c.F3(i) = static_cast<R>(m.f3);
    c.K(i) = cxm<R>(static_cast<R>(m.kre), static_cast<R>(m.kim));
  }
}

// One rational factor: solve (M + p2 D) u' = (M + p1 D) u by a top-to-bottom
// Thomas sweep and back substitution (fused_kernel.py:124-152).
template <class R>
__device__ void factor_solve(const Col<R> &c, int nz, R k0, R cfact, cx<R> p1, cx<R> p2) {
  const R a1 = k0 * k0 / R(6), a2 = R(2) * k0 * k0 / R(3);
  R f2m = c.F2(0), f2c = c.F2(1), f3m = c.F3(0), f3c = c.F3(1);
  cx<R> km = c.K(0), kc = c.K(1), um = c.U(0), uc = c.U(1);
  cx<R> cprev = cxm<R>(0, 0), dprev = cxm<R>(0, 0);
  for (int i = 1; i <= nz; ++i) {
    const R f1c = c.F1(i), f2p = c.F2(i + 1), f3p = c.F3(i + 1);
    const cx<R> kp = c.K(i + 1), up = c.U(i + 1);
    const Galerkin<R> g = galerkin<R>(cfact, f1c, f2m, f2c, f2p, f3m, f3c, f3p, km, kc, kp);
    const cx<R> r1 = cxm<R>(a1, 0) + p2 * g.d1;
    const cx<R> r2 = cxm<R>(a2, 0) + p2 * g.d2;
    const cx<R> r3 = cxm<R>(a1, 0) + p2 * g.d3;
    const cx<R> s1 = cxm<R>(a1, 0) + p1 * g.d1;
    const cx<R> s2 = cxm<R>(a2, 0) + p1 * g.d2;
    const cx<R> s3 = cxm<R>(a1, 0) + p1 * g.d3;
    const cx<R> rhs = s1 * um + s2 * uc + s3 * up;
    const cx<R> inv = crcp(cfms(r2, r1, cprev));
    const cx<R> cc = r3 * inv;
    const cx<R> dd = cfms(rhs, r1, dprev) * inv;
    c.CP(i) = cc;
    c.DP(i) = dd;
    cprev = cc;
    dprev = dd;
    f2m = f2c;
    f2c = f2p;
    f3m = f3c;
    f3c = f3p;
    km = kc;
    kc = kp;
    um = uc;
    uc = up;
  }
  cx<R> un = cxm<R>(0, 0);
  for (int i = nz; i >= 1; --i) {
    un = cfms(c.DP(i), c.CP(i), un);
    c.U(i) = un;
  }
}

// Cell-mean intensity: adds |f3 u|^2 / rf of the field at range rf to the
// output cells of range bin tlc+1 (the ndr fields that end at the next
// output), depth bin j holding rows j*ndz+1 .. (j+1)*ndz.  The lane owns the
// ray, so the sums are plain read-modify-writes.
template <class R>
__device__ void add_cells(const MarchArgs &a, const cpr_ray &P, const RayState &st, const Col<R> &c, double rf) {
  const int bin = st.tlc + 1;
  if (!a.cells || bin >= P.nvr) return;
  float *col = a.cells + static_cast<int64_t>(bin) * a.nvz_max * a.ld + c.b;
  const double rinv = 1.0 / rf;
  for (int j = 0; j < P.nvz; ++j) {
    double acc = 0.0;
    for (int g = j * P.ndz + 1; g <= (j + 1) * P.ndz; ++g) {
      const cx<R> v = c.U(g);
      const double f3 = static_cast<double>(c.F3(g));
      acc += f3 * f3 * (static_cast<double>(v.re) * v.re + static_cast<double>(v.im) * v.im);
    }
    col[static_cast<int64_t>(j) * a.ld] += static_cast<float>(acc * rinv);
  }
}

// Complex pressure at the receiver and on the output grid (outpt.py:39-53);
// the TL of every stored pressure is formed afterwards by k_tl_post.
template <class R>
__device__ void outpt(const MarchArgs &a, const cpr_ray &P, RayState &st, const Col<R> &c) {
  st.mdr += 1;
  if (st.mdr != P.ndr) return;
  st.mdr = 0;
  st.tlc += 1;
  if (st.tlc >= P.nvr) return;
  const double ri = P.zr / P.dz;
  const int ir = static_cast<int>(floor(ri));
  const double dir = ri - ir;
  const cx<R> u0 = c.U(ir), u1 = c.U(ir + 1);
  const double f30 = static_cast<double>(c.F3(ir)), f31 = static_cast<double>(c.F3(ir + 1));
  const double pre = (1.0 - dir) * f30 * static_cast<double>(u0.re) + dir * f31 * static_cast<double>(u1.re);
  const double pim = (1.0 - dir) * f30 * static_cast<double>(u0.im) + dir * f31 * static_cast<double>(u1.im);
  const int64_t o = static_cast<int64_t>(st.tlc) * a.ld + c.b;
  a.cpl[2 * o] = pre;
  a.cpl[2 * o + 1] = pim;
  if (a.grids) {
    for (int i = 0; i < P.nvz; ++i) {
      const int j = (i + 1) * P.ndz;
      const cx<R> uj = c.U(j);
      const double f3 = static_cast<double>(c.F3(j));
      const double gre = static_cast<double>(uj.re) * f3, gim = static_cast<double>(uj.im) * f3;
      const int64_t og = (static_cast<int64_t>(st.tlc) * a.nvz_max + i) * a.ld + c.b;
      a.cpg[2 * og] = gre;
      a.cpg[2 * og + 1] = gim;
    }
  }
}

template <class R>
__device__ inline cx<R> pade_coef(const double *pd, int idx) {
  return cxm<R>(static_cast<R>(pd[2 * idx]), static_cast<R>(pd[2 * idx + 1]));
}

}  // namespace

// FP64 self-starter; the column arrays of `a` are double.
__global__ void k_selfs(MarchArgs a) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= a.nrays) return;
  const cpr_ray P = a.rays[b];
  const cpr_env E = a.envs[b];
  RayState st;
  st.iz = P.iz0;
  st.bt_ind = st.ss_ind = st.sb_ind = 0;
  st.mdr = 0;
  st.tlc = -1;
  st.pset = 0;
  st.dirty = 1;
  const Col<double> c = col_of<double>(a, b);
  materialize(c, P, E, a.env, st, 0, P.nz + 1);
  for (int i = 0; i <= P.nz + 1; ++i) c.U(i) = cxm<double>(0, 0);
  const double si = P.zs / P.dz;
  const int is = static_cast<int>(floor(si));
  const double dis = si - is;
  const double amp = sqrt(2.0 * M_PI / P.k0);
  c.U(is) = cxm<double>((1.0 - dis) * amp / (P.dz * alpw_at(P, E, a.env, 0, grid_depth(P, is))), 0);
  c.U(is + 1) = cxm<double>(dis * amp / (P.dz * alpw_at(P, E, a.env, 0, grid_depth(P, is + 1))), 0);
  const double cfact = 0.5 / (P.dz * P.dz);
  for (int rep = 0; rep < 2; ++rep) {
    factor_solve<double>(c, P.nz, P.k0, cfact, cxm<double>(0, 0), cxm<double>(-1, 0));
    c.U(0) = cxm<double>(0, 0);
  }
  const double *pd = a.pade + (static_cast<int64_t>(b) * 3 + 1) * 2 * P.np * 2;
  for (int j = 0; j < P.np; ++j)
    factor_solve<double>(c, P.nz, P.k0, cfact, pade_coef<double>(pd, j), pade_coef<double>(pd, P.np + j));
  add_cells<double>(a, P, st, c, P.dr);
  outpt<double>(a, P, st, c);
  a.state[b] = st;
}

// Copy the FP64 starter field into the march field of precision R.
template <class R>
__global__ void k_field_from_starter(MarchArgs a, const cx<double> *ud) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= a.nrays) return;
  const int nz = a.rays[b].nz;
  cx<R> *u = static_cast<cx<R> *>(a.u);
  if (a.team_rows > 0) {
    // team layout: row i = 1 + t M + k at [b][k][t]
    const int M = a.team_rows, T = a.team_lanes;
    cx<R> *ub = u + static_cast<int64_t>(b) * M * T;
    for (int i = 1; i <= nz; ++i) {
      const cx<double> v = ud[i * a.ld + b];
      const int t = (i - 1) / M, k = (i - 1) % M;
      ub[static_cast<int64_t>(k) * T + t] = cxm<R>(static_cast<R>(v.re), static_cast<R>(v.im));
    }
    return;
  }
  for (int i = 0; i < nz + 2; ++i) {
    const cx<double> v = ud[i * a.ld + b];
    u[i * a.ld + b] = cxm<R>(static_cast<R>(v.re), static_cast<R>(v.im));
  }
}

template <class R>
__global__ void __launch_bounds__(64) k_march_thread(MarchArgs a, int s0, int s1) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= a.nrays) return;
  const cpr_ray P = a.rays[b];
  const cpr_env E = a.envs[b];
  RayState st = a.state[b];
  const Col<R> c = col_of<R>(a, b);
  const R k0 = static_cast<R>(P.k0);
  const R cfact = static_cast<R>(0.5 / (P.dz * P.dz));
  const int end = s1 < P.nsteps ? s1 : P.nsteps;
  for (int step = s0; step < end; ++step) {
    const double r = (step + 1) * P.dr;
    const int jz = st.iz;
    const int fl = updat_ray(P, E, a.env, st, r);
    if ((fl & 2) || st.dirty) {
      materialize(c, P, E, a.env, st, 0, P.nz + 1);
      st.dirty = 0;
    } else if (fl & 1) {
      const int lo = (jz < st.iz ? jz : st.iz) + 1;
      const int hi = jz > st.iz ? jz : st.iz;
      materialize(c, P, E, a.env, st, lo, hi);
    }
    const double *pd = a.pade + (static_cast<int64_t>(b) * 3 + st.pset) * 2 * P.np * 2;
    for (int j = 0; j < P.np; ++j)
      factor_solve<R>(c, P.nz, k0, cfact, pade_coef<R>(pd, j), pade_coef<R>(pd, P.np + j));
    add_cells<R>(a, P, st, c, (step + 2) * P.dr);
    outpt<R>(a, P, st, c);
  }
  a.state[b] = st;
}

template __global__ void k_march_thread<float>(MarchArgs, int, int);
template __global__ void k_march_thread<double>(MarchArgs, int, int);
template __global__ void k_field_from_starter<float>(MarchArgs, const cx<double> *);
template __global__ void k_field_from_starter<double>(MarchArgs, const cx<double> *);

// TL = -20 log10(|p| + 1e-20) + 10 log10(r + 1e-20) of every stored output
// (outpt.py:43-53); output t of a ray lies at r = (t+1) ndr dr.
__global__ void k_tl_post(MarchArgs a) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  const int t = blockIdx.y;
  if (b >= a.nrays || t >= a.nvr_max) return;
  const cpr_ray &P = a.rays[b];
  if (t >= P.nvr) return;
  const double r = (t + 1) * P.ndr * P.dr;
  const int64_t o = static_cast<int64_t>(t) * a.ld + b;
  a.tll[o] = tl_of(a.cpl[2 * o], a.cpl[2 * o + 1], r);
  if (a.cells) {
    const double inv_n = 1.0 / (static_cast<double>(P.ndr) * P.ndz);
    for (int i = 0; i < P.nvz; ++i) {
      const int64_t og = (static_cast<int64_t>(t) * a.nvz_max + i) * a.ld + b;
      a.tlcell[og] = -10.0 * log10(static_cast<double>(a.cells[og]) * inv_n + 1e-40);
    }
  }
  if (a.grids) {
    for (int i = 0; i < P.nvz; ++i) {
      const int64_t og = (static_cast<int64_t>(t) * a.nvz_max + i) * a.ld + b;
      a.tlg[og] = tl_of(a.cpg[2 * og], a.cpg[2 * og + 1], r);
    }
  }
}
