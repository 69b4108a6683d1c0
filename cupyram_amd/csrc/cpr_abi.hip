// C ABI of libcupyram_hip.so: plan lifetime, uploads, launches, read-back.
// Declared in include/cupyram_hip.h; the reference interface each entry point
// replaces is listed there.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cstdio>
#include <mutex>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "march_common.h"

template <class R>
__global__ void k_march_thread(MarchArgs a, int s0, int s1);
template <class R>
__global__ void k_field_from_starter(MarchArgs a, const cx<double> *ud);
__global__ void k_selfs(MarchArgs a);
__global__ void k_tl_post(MarchArgs a);
int team_launch(const MarchArgs &a, int precision, int rows, int lanes, int s0, int s1, hipStream_t st);
template <class R>
__global__ void k_team_terms(MarchArgs a, R *qtab);
bool team_shape_supported(int precision, int rows, int lanes);
void team_auto_shape(int precision, int nz, int *rows, int *lanes);
int64_t team_dbuf_bytes(int precision, int rows, int lanes, int nrays);
int64_t team_mbuf_bytes(int rows, int lanes, int nrays);

// Ray-major gather of a depth-major output array, for rays b0 .. b0+nb-1:
// source element (t, i, b) at (t * nz + i) * ld + b, destination element
// (b - b0, i, t) at ((b - b0) * nz + i) * nt + t, of type T (double for TL,
// double2 for a complex pressure).  32 x 32 tiles of (t, b) pass through LDS
// so that the reads (along b) and the writes (along t) are both coalesced.
template <class T>
__global__ void __launch_bounds__(256) k_gather_rays(const T *src, T *dst, int64_t ld, int nt, int nz, int b0, int nb) {
  __shared__ T tile[32][33];
  const int tb = blockIdx.x * 32, tt = blockIdx.y * 32, i = blockIdx.z;
  const int lx = threadIdx.x % 32, ly = threadIdx.x / 32;
  for (int r = ly; r < 32; r += 8) {
    const int t = tt + r, b = tb + lx;
    if (t < nt && b < nb) tile[r][lx] = src[(static_cast<int64_t>(t) * nz + i) * ld + b0 + b];
  }
  __syncthreads();
  for (int r = ly; r < 32; r += 8) {
    const int b = tb + r, t = tt + lx;
    if (t < nt && b < nb) dst[(static_cast<int64_t>(b) * nz + i) * nt + t] = tile[lx][r];
  }
}

namespace {

thread_local std::string g_err;

int fail(const std::string &m) {
  g_err = m;
  return -1;
}

#define HIPCHK(expr)                                                                  \
  do {                                                                                \
    hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess) return fail(std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

}  // namespace

int cpr_fail(const std::string &m) { return fail(m); }

struct cpr_plan {
  cpr_config cfg{};
  int nrays = 0, np_ = 0, nz_max = 0, max_steps = 0;
  int64_t ld = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  std::vector<cpr_ray> rays_h;
  cpr_ray *rays_d = nullptr;
  cpr_env *envs_d = nullptr;
  double *env_d = nullptr;
  double *pade_d = nullptr;
  RayState *state_d = nullptr;
  void *col[7] = {};  // f1 f2 f3 ksq u cp dp
  double *tll_d = nullptr, *cpl_d = nullptr, *tlg_d = nullptr, *cpg_d = nullptr;
  float *cells_d = nullptr;
  double *tlcell_d = nullptr;
  int32_t *team_rays_d = nullptr;
  char *dbuf_d = nullptr;     // team coefficients (all team shapes)
  double *mbuf_d = nullptr;   // team media cache
  char *qtab_d = nullptr;     // team per-term constants
  double *stage_d = nullptr;  // staging buffer of the ray-major output reads
  void *starter_cols[7] = {};  // the self-starter's FP64 columns, freed at the next synchronisation
  int64_t stage_n = 0;        // its size in doubles
  int variant = CPR_VARIANT_THREAD;
  int team_rows = 0, team_lanes = 0;
  // hipEvent pairs around march launches not yet read, the time of those
  // already read, and completed events kept for reuse (a caller that marches
  // one step per call records two events per call)
  std::vector<hipEvent_t> events;
  std::vector<hipEvent_t> free_events;
  double done_ms = 0.0;
  int done_launches = 0;
  int64_t bytes = 0;
  hipGraphExec_t graph_exec = nullptr;
  hipGraph_t graph = nullptr;

  MarchArgs args() const {
    MarchArgs a{};
    a.rays = rays_d;
    a.envs = envs_d;
    a.env = env_d;
    a.pade = pade_d;
    a.state = state_d;
    a.nrays = nrays;
    a.np_ = np_;
    a.ld = ld;
    a.nvr_max = cfg.nvr_max;
    a.nvz_max = cfg.nvz_max;
    a.grids = cfg.compute_grids;
    a.f1 = col[0];
    a.f2 = col[1];
    a.f3 = col[2];
    a.ksq = col[3];
    a.u = col[4];
    a.cp = col[5];
    a.dp = col[6];
    a.tll = tll_d;
    a.cpl = cpl_d;
    a.tlg = tlg_d;
    a.cpg = cpg_d;
    a.cells = cells_d;
    a.tlcell = tlcell_d;
    a.team_rays = team_rays_d;
    a.team_nrays = nrays;
    a.dbuf = dbuf_d;
    a.mbuf = mbuf_d;
    a.qtab = qtab_d;
    a.team_rows = variant == CPR_VARIANT_TEAM ? team_rows : 0;
    a.team_lanes = variant == CPR_VARIANT_TEAM ? team_lanes : 0;
    return a;
  }
  size_t real_size() const { return cfg.precision == CPR_C64 ? sizeof(float) : sizeof(double); }
};

namespace {

template <class T>
int dalloc(cpr_plan *p, T **ptr, size_t count) {
  const size_t b = std::max<size_t>(count, 1) * sizeof(T);
  HIPCHK(hipMalloc(reinterpret_cast<void **>(ptr), b));
  HIPCHK(hipMemsetAsync(*ptr, 0, b, p->stream));
  p->bytes += static_cast<int64_t>(b);
  return 0;
}

// Elements of column k: depth-major [nz_max+2][ld]; the field of a team plan
// ([ray][k][lane]) may need more (rows x lanes can exceed nz_max + 2).
size_t column_cells(const cpr_plan *p, int k) {
  size_t cells = static_cast<size_t>(p->nz_max + 2) * static_cast<size_t>(p->ld);
  if (k == 4 && p->variant == CPR_VARIANT_TEAM)
    cells = std::max(cells, static_cast<size_t>(p->nrays) * p->team_rows * p->team_lanes);
  return cells;
}

int alloc_columns(cpr_plan *p, void **col, size_t rs, bool only_u = false) {
  const size_t sizes[7] = {rs, rs, rs, 2 * rs, 2 * rs, 2 * rs, 2 * rs};
  for (int k = 0; k < 7; ++k) {
    if (only_u && k != 4) continue;
    char *q = nullptr;
    if (dalloc(p, &q, column_cells(p, k) * sizes[k])) return -1;
    col[k] = q;
  }
  return 0;
}

void free_columns(cpr_plan *p, void **col, size_t rs) {
  const size_t sizes[7] = {rs, rs, rs, 2 * rs, 2 * rs, 2 * rs, 2 * rs};
  for (int k = 0; k < 7; ++k) {
    if (col[k]) {
      (void)hipFree(col[k]);
      p->bytes -= static_cast<int64_t>(std::max<size_t>(column_cells(p, k) * sizes[k], 1));
      col[k] = nullptr;
    }
  }
}

// The self-starter's FP64 columns (a plan whose march keeps its own
// precision allocates them only for the starter): freed at the plan's next
// synchronisation (cpr_sync, cpr_march_time, the output reads, the next
// cpr_selfs, destroy) instead of right after the starter, so the starter
// needs no host synchronisation.
void release_starter_columns(cpr_plan *p) {
  if (p->starter_cols[0] || p->starter_cols[4]) {
    (void)hipStreamSynchronize(p->stream);
    free_columns(p, p->starter_cols, sizeof(double));
  }
}

int check_ray(const cpr_ray &r, const cpr_env &e, int64_t env_len, int b) {
  char buf[256];
  auto bad = [&](const char *what) {
    std::snprintf(buf, sizeof(buf), "ray %d: %s", b, what);
    return fail(buf);
  };
  if (r.nz < 3) return bad("nz < 3");
  if (r.np < 1 || r.np > CPR_MAX_NP) return bad("np out of range");
  if (!(r.dz > 0) || !(r.dr > 0) || !(r.zstep > 0)) return bad("dz, dr and zstep must be positive");
  if (r.nsteps < 0 || r.ndr < 1 || r.ndz < 1) return bad("bad step counts");
  if (r.iz0 < 1 || r.iz0 > r.nz - 1) return bad("iz0 outside 1..nz-1");
  const int is = static_cast<int>(std::floor(r.zs / r.dz));
  const int ir = static_cast<int>(std::floor(r.zr / r.dz));
  if (is < 0 || is + 1 > r.nz + 1) return bad("source depth outside the grid");
  if (ir < 0 || ir + 1 > r.nz + 1) return bad("receiver depth outside the grid");
  if (r.nvz < 0 || static_cast<int64_t>(r.nvz) * r.ndz > r.nz + 1) return bad("output depths exceed the grid");
  if (e.n_ss < 1 || e.n_rp_ss < 1 || e.n_sb < 1 || e.n_rp_sb < 1 || e.n_bt < 1) return bad("empty profile");
  const int64_t ends[9] = {e.z_ss + e.n_ss,
                           e.rp_ss + e.n_rp_ss,
                           e.cw + static_cast<int64_t>(e.n_ss) * e.n_rp_ss,
                           e.z_sb + e.n_sb,
                           e.rp_sb + e.n_rp_sb,
                           e.cb + static_cast<int64_t>(e.n_sb) * e.n_rp_sb,
                           e.rhob + static_cast<int64_t>(e.n_sb) * e.n_rp_sb,
                           e.attn + static_cast<int64_t>(e.n_sb) * e.n_rp_sb,
                           e.rbzb + 2 * static_cast<int64_t>(e.n_bt)};
  const int64_t starts[9] = {e.z_ss, e.rp_ss, e.cw, e.z_sb, e.rp_sb, e.cb, e.rhob, e.attn, e.rbzb};
  for (int k = 0; k < 9; ++k)
    if (starts[k] < 0 || ends[k] > env_len) return bad("environment offsets outside env_data");
  return 0;
}

// An event for a march launch: a completed one from the pool, or a new one.
int take_event(cpr_plan *p, hipEvent_t *e) {
  if (!p->free_events.empty()) {
    *e = p->free_events.back();
    p->free_events.pop_back();
    return 0;
  }
  HIPCHK(hipEventCreate(e));
  return 0;
}

// Adds the time of every completed launch (in launch order) to the plan's
// total and returns its two events to the pool.
int fold_events(cpr_plan *p) {
  size_t i = 0;
  for (; i + 1 < p->events.size(); i += 2) {
    const hipError_t q = hipEventQuery(p->events[i + 1]);
    if (q == hipErrorNotReady) break;
    if (q != hipSuccess) return fail(std::string("hipEventQuery: ") + hipGetErrorString(q));
    float t = 0;
    HIPCHK(hipEventElapsedTime(&t, p->events[i], p->events[i + 1]));
    p->done_ms += t;
    p->done_launches += 1;
    p->free_events.push_back(p->events[i]);
    p->free_events.push_back(p->events[i + 1]);
  }
  p->events.erase(p->events.begin(), p->events.begin() + static_cast<std::ptrdiff_t>(i));
  return 0;
}

}  // namespace

extern "C" {

const char *cpr_last_error(void) { return g_err.c_str(); }

int cpr_device_count(int *count) {
  int n = 0;
  const hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess) n = 0;
  if (count) *count = n;
  return 0;
}

int cpr_device_info(int device, char *name, int name_len, int64_t *free_bytes, int64_t *total_bytes) {
  int n = 0;
  HIPCHK(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) return fail("cpr_device_info: no such device");
  hipDeviceProp_t prop;
  HIPCHK(hipGetDeviceProperties(&prop, device));
  if (name && name_len > 0) {
    // the marketing name, or the architecture where the runtime leaves it empty
    std::snprintf(name, static_cast<size_t>(name_len), "%s", prop.name[0] ? prop.name : prop.gcnArchName);
  }
  if (free_bytes || total_bytes) {
    int cur = 0;
    HIPCHK(hipGetDevice(&cur));
    HIPCHK(hipSetDevice(device));
    size_t fr = 0, to = 0;
    const hipError_t e = hipMemGetInfo(&fr, &to);
    HIPCHK(hipSetDevice(cur));
    HIPCHK(e);
    if (free_bytes) *free_bytes = static_cast<int64_t>(fr);
    if (total_bytes) *total_bytes = static_cast<int64_t>(to);
  }
  return 0;
}

int cpr_pade(double freq, double c0, int np_, int ns, double dr, int ip, double *pd1, double *pd2) {
  const int rc = cpr_pade_impl(freq, c0, np_, ns, dr, ip, pd1, pd2);
  if (rc == -1) return fail("Laguerre method not converging");
  if (rc) return fail("invalid Pade arguments");
  return 0;
}

int cpr_pade_batch(int n, const double *freq, const double *c0, int np_, int ns, const double *dr, int ip,
                   double *out, int nthreads) {
  const int rc = cpr_pade_batch_impl(n, freq, c0, np_, ns, dr, ip, out, nthreads);
  if (rc == -1) return fail("Laguerre method not converging");
  if (rc) return fail("invalid Pade arguments");
  return 0;
}

int cpr_plan_destroy(cpr_plan *p) {
  if (!p) return 0;
  if (p->stream) (void)hipStreamSynchronize(p->stream);
  if (p->graph_exec) (void)hipGraphExecDestroy(p->graph_exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  for (auto e : p->events) (void)hipEventDestroy(e);
  for (auto e : p->free_events) (void)hipEventDestroy(e);
  free_columns(p, p->starter_cols, sizeof(double));
  free_columns(p, p->col, p->real_size());
  for (void *q : {static_cast<void *>(p->rays_d), static_cast<void *>(p->envs_d), static_cast<void *>(p->env_d),
                  static_cast<void *>(p->pade_d), static_cast<void *>(p->state_d), static_cast<void *>(p->tll_d),
                  static_cast<void *>(p->cpl_d), static_cast<void *>(p->tlg_d), static_cast<void *>(p->cpg_d),
                  static_cast<void *>(p->team_rays_d), static_cast<void *>(p->dbuf_d),
                  static_cast<void *>(p->mbuf_d), static_cast<void *>(p->qtab_d), static_cast<void *>(p->stage_d),
                  static_cast<void *>(p->cells_d),
                  static_cast<void *>(p->tlcell_d)})
    if (q) (void)hipFree(q);
  if (p->own_stream && p->stream) (void)hipStreamDestroy(p->stream);
  delete p;
  return 0;
}

int cpr_plan_create(const cpr_config *cfg, const cpr_ray *rays, const cpr_env *envs, const double *env_data,
                    int64_t env_len, const double *pade, void *stream, cpr_plan **out) {
  if (!cfg || !rays || !envs || !env_data || !pade || !out) return fail("null argument");
  *out = nullptr;
  if (cfg->nrays < 1) return fail("nrays must be >= 1");
  if (cfg->precision != CPR_C64 && cfg->precision != CPR_C128) return fail("precision must be 64 or 128");
  if (cfg->nvr_max < 0 || cfg->nvz_max < 0) return fail("negative output sizes");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail("no HIP device available");
  const int np_ = rays[0].np;
  int nz_max = 0, max_steps = 0;
  for (int b = 0; b < cfg->nrays; ++b) {
    if (check_ray(rays[b], envs[b], env_len, b)) return -1;
    if (rays[b].np != np_) return fail("all rays of a plan must use the same np");
    if (rays[b].nvr > cfg->nvr_max) return fail("ray nvr exceeds nvr_max");
    if ((cfg->compute_grids || cfg->compute_cells) && rays[b].nvz > cfg->nvz_max)
      return fail("ray nvz exceeds nvz_max");
    nz_max = std::max(nz_max, rays[b].nz);
    max_steps = std::max(max_steps, rays[b].nsteps);
  }
  cpr_plan *p = new cpr_plan();
  p->cfg = *cfg;
  p->nrays = cfg->nrays;
  p->np_ = np_;
  p->nz_max = nz_max;
  p->max_steps = max_steps;
  p->ld = (static_cast<int64_t>(cfg->nrays) + 63) / 64 * 64;
  p->rays_h.assign(rays, rays + cfg->nrays);
  auto bail = [&](int rc) {
    std::string m = g_err;
    cpr_plan_destroy(p);
    g_err = m;
    return rc;
  };
  if (stream) {
    p->stream = static_cast<hipStream_t>(stream);
  } else {
    if (hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) != hipSuccess)
      return bail(fail("hipStreamCreate failed"));
    p->own_stream = true;
  }
  // variant choice
  p->variant = cfg->variant;
  if (p->variant == CPR_VARIANT_AUTO) {
    // the register-resident team march when a team shape holds the column
    int rows = 0, lanes = 0;
    team_auto_shape(cfg->precision, nz_max, &rows, &lanes);
    p->variant = (rows > 0 && static_cast<int64_t>(rows) * lanes >= nz_max) ? CPR_VARIANT_TEAM : CPR_VARIANT_THREAD;
  }
  if (p->variant == CPR_VARIANT_TEAM) {
    int rows = cfg->team_rows, lanes = cfg->team_lanes;
    if (rows == 0 || lanes == 0) team_auto_shape(cfg->precision, nz_max, &rows, &lanes);
    if (!team_shape_supported(cfg->precision, rows, lanes))
      return bail(fail("unsupported team shape (rows per lane x lanes per ray)"));
    if (static_cast<int64_t>(rows) * lanes < nz_max) return bail(fail("team shape too small for nz"));
    p->team_rows = rows;
    p->team_lanes = lanes;
  } else if (p->variant != CPR_VARIANT_THREAD) {
    return bail(fail("unknown variant"));
  }
  const int64_t npade = static_cast<int64_t>(cfg->nrays) * 3 * 2 * np_ * 2;
  if (dalloc(p, &p->rays_d, cfg->nrays) || dalloc(p, &p->envs_d, cfg->nrays) || dalloc(p, &p->env_d, env_len) ||
      dalloc(p, &p->pade_d, npade) || dalloc(p, &p->state_d, cfg->nrays))
    return bail(-1);
  if (alloc_columns(p, p->col, p->real_size(), p->variant == CPR_VARIANT_TEAM)) return bail(-1);
  if (p->variant == CPR_VARIANT_TEAM) {
    std::vector<int32_t> ids(cfg->nrays);
    for (int b = 0; b < cfg->nrays; ++b) ids[b] = b;
    if (dalloc(p, &p->team_rays_d, cfg->nrays)) return bail(-1);
    const int64_t db = team_dbuf_bytes(cfg->precision, p->team_rows, p->team_lanes, cfg->nrays);
    if (dalloc(p, &p->dbuf_d, static_cast<size_t>(db))) return bail(-1);
    const int64_t mb = team_mbuf_bytes(p->team_rows, p->team_lanes, cfg->nrays);
    if (dalloc(p, &p->mbuf_d, static_cast<size_t>(mb / sizeof(double)))) return bail(-1);
    // on the plan's stream, behind dalloc's clearing of the same buffer: the
    // stream is non-blocking, so a null-stream copy could land before it
    if (hipMemcpyAsync(p->team_rays_d, ids.data(), sizeof(int32_t) * cfg->nrays, hipMemcpyHostToDevice,
                       p->stream) != hipSuccess ||
        hipStreamSynchronize(p->stream) != hipSuccess)
      return bail(fail("upload failed"));
  }
  const size_t nline = static_cast<size_t>(std::max(cfg->nvr_max, 1)) * p->ld;
  if (dalloc(p, &p->tll_d, nline) || dalloc(p, &p->cpl_d, 2 * nline)) return bail(-1);
  if (cfg->compute_grids) {
    const size_t ngrid = nline * static_cast<size_t>(std::max(cfg->nvz_max, 1));
    if (dalloc(p, &p->tlg_d, ngrid) || dalloc(p, &p->cpg_d, 2 * ngrid)) return bail(-1);
  }
  if (cfg->compute_cells) {
    const size_t ngrid = nline * static_cast<size_t>(std::max(cfg->nvz_max, 1));
    if (dalloc(p, &p->cells_d, ngrid) || dalloc(p, &p->tlcell_d, ngrid)) return bail(-1);
  }
  if (hipMemcpyAsync(p->rays_d, rays, sizeof(cpr_ray) * cfg->nrays, hipMemcpyHostToDevice, p->stream) ||
      hipMemcpyAsync(p->envs_d, envs, sizeof(cpr_env) * cfg->nrays, hipMemcpyHostToDevice, p->stream) ||
      hipMemcpyAsync(p->env_d, env_data, sizeof(double) * env_len, hipMemcpyHostToDevice, p->stream) ||
      hipMemcpyAsync(p->pade_d, pade, sizeof(double) * npade, hipMemcpyHostToDevice, p->stream))
    return bail(fail("upload failed"));
  if (p->variant == CPR_VARIANT_TEAM) {
    // per-term constants of the team march, formed once on the device
    const int64_t nq = static_cast<int64_t>(cfg->nrays) * 3 * np_ * kTermStride;
    if (dalloc(p, &p->qtab_d, static_cast<size_t>(nq) * p->real_size())) return bail(-1);
    const MarchArgs m = p->args();
    const int block = 64, grid = (cfg->nrays + block - 1) / block;
    if (cfg->precision == CPR_C64)
      hipLaunchKernelGGL(k_team_terms<float>, dim3(grid), dim3(block), 0, p->stream, m,
                         reinterpret_cast<float *>(p->qtab_d));
    else
      hipLaunchKernelGGL(k_team_terms<double>, dim3(grid), dim3(block), 0, p->stream, m,
                         reinterpret_cast<double *>(p->qtab_d));
    if (hipGetLastError() != hipSuccess) return bail(fail("k_team_terms launch failed"));
  }
  if (hipStreamSynchronize(p->stream) != hipSuccess) return bail(fail("upload failed"));
  *out = p;
  return 0;
}

int cpr_selfs(cpr_plan *p) {
  if (!p) return fail("null plan");
  if (fold_events(p)) return -1;
  for (auto e : p->events) (void)hipEventDestroy(e);  // launches of an earlier march still in flight
  p->events.clear();
  p->done_ms = 0.0;
  p->done_launches = 0;
  MarchArgs a = p->args();
  void *dcol[7] = {};
  const bool own = !(p->cfg.precision == CPR_C128 && p->variant == CPR_VARIANT_THREAD);
  if (own) {
    release_starter_columns(p);
    if (alloc_columns(p, dcol, sizeof(double))) {
      free_columns(p, dcol, sizeof(double));
      return -1;
    }
    a.f1 = dcol[0];
    a.f2 = dcol[1];
    a.f3 = dcol[2];
    a.ksq = dcol[3];
    a.u = dcol[4];
    a.cp = dcol[5];
    a.dp = dcol[6];
  }
  if (p->cells_d) {
    const size_t ngrid = static_cast<size_t>(std::max(p->cfg.nvr_max, 1)) * p->ld *
                         static_cast<size_t>(std::max(p->cfg.nvz_max, 1));
    HIPCHK(hipMemsetAsync(p->cells_d, 0, ngrid * sizeof(float), p->stream));
  }
  const int block = 64;
  const int grid = (p->nrays + block - 1) / block;
  hipLaunchKernelGGL(k_selfs, dim3(grid), dim3(block), 0, p->stream, a);
  hipError_t err = hipGetLastError();
  if (own) {
    if (err == hipSuccess) {
      MarchArgs m = p->args();
      if (p->cfg.precision == CPR_C64)
        hipLaunchKernelGGL(k_field_from_starter<float>, dim3(grid), dim3(block), 0, p->stream, m,
                           static_cast<const cx<double> *>(dcol[4]));
      else
        hipLaunchKernelGGL(k_field_from_starter<double>, dim3(grid), dim3(block), 0, p->stream, m,
                           static_cast<const cx<double> *>(dcol[4]));
      err = hipGetLastError();
    }
    for (int k = 0; k < 7; ++k) p->starter_cols[k] = dcol[k];
  }
  HIPCHK(err);
  return 0;
}

static int launch_march(cpr_plan *p, int s0, int s1) {
  const MarchArgs a = p->args();
  if (p->variant == CPR_VARIANT_TEAM) {
    if (team_launch(a, p->cfg.precision, p->team_rows, p->team_lanes, s0, s1, p->stream))
      return fail("team march launch failed");
    return 0;
  }
  const int block = 64;
  const int grid = (p->nrays + block - 1) / block;
  if (p->cfg.precision == CPR_C64)
    hipLaunchKernelGGL(k_march_thread<float>, dim3(grid), dim3(block), 0, p->stream, a, s0, s1);
  else
    hipLaunchKernelGGL(k_march_thread<double>, dim3(grid), dim3(block), 0, p->stream, a, s0, s1);
  HIPCHK(hipGetLastError());
  return 0;
}

int cpr_march(cpr_plan *p, int step_begin, int step_end) {
  if (!p) return fail("null plan");
  if (step_begin < 0 || step_end < step_begin) return fail("bad step range");
  step_end = std::min(step_end, p->max_steps);
  if (step_end <= step_begin) return 0;
  hipEvent_t e0, e1;
  if (fold_events(p) || take_event(p, &e0) || take_event(p, &e1)) return -1;
  p->events.push_back(e0);
  p->events.push_back(e1);
  HIPCHK(hipEventRecord(e0, p->stream));
  if (launch_march(p, step_begin, step_end)) return -1;
  HIPCHK(hipEventRecord(e1, p->stream));
  return 0;
}

int cpr_march_graph(cpr_plan *p, int steps_per_launch) {
  if (!p) return fail("null plan");
  if (steps_per_launch < 1) return fail("steps_per_launch must be >= 1");
  if (p->graph_exec) {
    (void)hipGraphExecDestroy(p->graph_exec);
    p->graph_exec = nullptr;
  }
  if (p->graph) {
    (void)hipGraphDestroy(p->graph);
    p->graph = nullptr;
  }
  HIPCHK(hipStreamBeginCapture(p->stream, hipStreamCaptureModeThreadLocal));
  int rc = 0;
  for (int s = 0; s < p->max_steps && !rc; s += steps_per_launch)
    rc = launch_march(p, s, std::min(s + steps_per_launch, p->max_steps));
  hipGraph_t g = nullptr;
  const hipError_t ec = hipStreamEndCapture(p->stream, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (ec != hipSuccess) return fail(std::string("hipStreamEndCapture: ") + hipGetErrorString(ec));
  p->graph = g;
  HIPCHK(hipGraphInstantiate(&p->graph_exec, g, nullptr, nullptr, 0));
  hipEvent_t e0, e1;
  if (fold_events(p) || take_event(p, &e0) || take_event(p, &e1)) return -1;
  p->events.push_back(e0);
  p->events.push_back(e1);
  HIPCHK(hipEventRecord(e0, p->stream));
  HIPCHK(hipGraphLaunch(p->graph_exec, p->stream));
  HIPCHK(hipEventRecord(e1, p->stream));
  return 0;
}

int cpr_sync(cpr_plan *p) {
  if (!p) return fail("null plan");
  HIPCHK(hipStreamSynchronize(p->stream));
  release_starter_columns(p);
  return 0;
}

int cpr_march_time(cpr_plan *p, double *ms, int *nlaunch) {
  if (!p) return fail("null plan");
  HIPCHK(hipStreamSynchronize(p->stream));
  release_starter_columns(p);
  if (fold_events(p)) return -1;
  if (ms) *ms = p->done_ms;
  if (nlaunch) *nlaunch = p->done_launches;
  return 0;
}

namespace {

constexpr int64_t kStageDoubles = int64_t(4) << 20;  // 32 MiB

// Copies the depth-major array src ([nt][nz][ld] elements of E doubles) to
// the ray-major host array dst ([nrays][nz][nt]): the rays are gathered on the
// device in chunks that fill the staging buffer, and each chunk lands in one
// contiguous span of dst.
int read_ray_major(cpr_plan *p, const double *src, double *dst, int nt, int nz, int E) {
  const int64_t per_ray = static_cast<int64_t>(nt) * nz * E;
  if (per_ray == 0 || p->nrays == 0) return 0;
  if (nz > 65535 || (nt + 31) / 32 > 65535) return fail("output grid too large for the gather");
  const int64_t cap = std::max(kStageDoubles, per_ray);  // at least one ray per chunk
  if (p->stage_n < cap) {
    if (p->stage_d) HIPCHK(hipFree(p->stage_d));
    p->bytes -= p->stage_n * static_cast<int64_t>(sizeof(double));
    p->stage_d = nullptr;
    p->stage_n = 0;
    HIPCHK(hipMalloc(reinterpret_cast<void **>(&p->stage_d), cap * sizeof(double)));
    p->stage_n = cap;
    p->bytes += cap * static_cast<int64_t>(sizeof(double));
  }
  const int chunk = static_cast<int>(std::min<int64_t>(p->nrays, p->stage_n / per_ray));
  for (int b0 = 0; b0 < p->nrays; b0 += chunk) {
    const int nb = std::min(chunk, p->nrays - b0);
    const dim3 grid((nb + 31) / 32, (nt + 31) / 32, nz);
    if (E == 1)
      hipLaunchKernelGGL(k_gather_rays<double>, grid, dim3(256), 0, p->stream, src, p->stage_d, p->ld, nt, nz, b0, nb);
    else
      hipLaunchKernelGGL(k_gather_rays<double2>, grid, dim3(256), 0, p->stream, reinterpret_cast<const double2 *>(src),
                         reinterpret_cast<double2 *>(p->stage_d), p->ld, nt, nz, b0, nb);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(dst + b0 * per_ray, p->stage_d, nb * per_ray * sizeof(double), hipMemcpyDeviceToHost,
                          p->stream));
    HIPCHK(hipStreamSynchronize(p->stream));
  }
  return 0;
}

int tl_post(cpr_plan *p) {
  if (p->cfg.nvr_max > 0) {
    const int block = 64;
    hipLaunchKernelGGL(k_tl_post, dim3((p->nrays + block - 1) / block, p->cfg.nvr_max), dim3(block), 0, p->stream,
                       p->args());
    HIPCHK(hipGetLastError());
  }
  return 0;
}

}  // namespace

int cpr_read_outputs(cpr_plan *p, double *tll, double *cpl, double *tlg, double *cpg) {
  if (!p) return fail("null plan");
  if ((tlg || cpg) && !p->cfg.compute_grids) return fail("plan was created without grids");
  if (tl_post(p)) return -1;
  const int nvr = p->cfg.nvr_max, nvz = p->cfg.nvz_max;
  if (tll && read_ray_major(p, p->tll_d, tll, nvr, 1, 1)) return -1;
  if (cpl && read_ray_major(p, p->cpl_d, cpl, nvr, 1, 2)) return -1;
  if (tlg && read_ray_major(p, p->tlg_d, tlg, nvr, nvz, 1)) return -1;
  if (cpg && read_ray_major(p, p->cpg_d, cpg, nvr, nvz, 2)) return -1;
  HIPCHK(hipStreamSynchronize(p->stream));
  release_starter_columns(p);
  return 0;
}

int cpr_read_cells(cpr_plan *p, double *tlc) {
  if (!p || !tlc) return fail("null argument");
  if (!p->cells_d) return fail("plan was created without compute_cells");
  if (tl_post(p)) return -1;
  if (read_ray_major(p, p->tlcell_d, tlc, p->cfg.nvr_max, p->cfg.nvz_max, 1)) return -1;
  HIPCHK(hipStreamSynchronize(p->stream));
  release_starter_columns(p);
  return 0;
}

int cpr_read_field(cpr_plan *p, double *u, int nz_max) {
  if (!p || !u) return fail("null argument");
  if (nz_max < p->nz_max) return fail("nz_max smaller than the plan's");
  HIPCHK(hipStreamSynchronize(p->stream));
  const int64_t ld = p->ld, rows = p->nz_max + 2, n = p->nrays;
  const bool c64 = p->cfg.precision == CPR_C64;
  const bool team = p->variant == CPR_VARIANT_TEAM;
  const int64_t M = p->team_rows, T = p->team_lanes;
  const size_t bytes = column_cells(p, 4) * (c64 ? 8 : 16);
  std::vector<char> buf(bytes);
  HIPCHK(hipMemcpyAsync(buf.data(), p->col[4], bytes, hipMemcpyDeviceToHost, p->stream));
  HIPCHK(hipStreamSynchronize(p->stream));
  std::memset(u, 0, sizeof(double) * 2 * n * (nz_max + 2));
  for (int64_t b = 0; b < n; ++b)
    for (int64_t i = 0; i < rows; ++i) {
      // team plans keep rows 1..nz as [ray][k][lane]; rows 0 and nz+1 are zero
      int64_t e = i * ld + b;
      if (team) {
        if (i < 1 || i > p->rays_h[b].nz) continue;
        e = (b * M + (i - 1) % M) * T + (i - 1) / M;
      }
      double re, im;
      if (c64) {
        const float *f = reinterpret_cast<const float *>(buf.data());
        re = f[2 * e];
        im = f[2 * e + 1];
      } else {
        const double *d = reinterpret_cast<const double *>(buf.data());
        re = d[2 * e];
        im = d[2 * e + 1];
      }
      u[2 * (b * (nz_max + 2) + i)] = re;
      u[2 * (b * (nz_max + 2) + i) + 1] = im;
    }
  return 0;
}

int64_t cpr_plan_device_bytes(cpr_plan *p) { return p ? p->bytes : -1; }

int cpr_team_shape(int precision, int nz, int *rows, int *lanes) {
  if (precision != CPR_C64 && precision != CPR_C128) return fail("precision must be 64 or 128");
  int r = 0, l = 0;
  team_auto_shape(precision, nz, &r, &l);
  if (r <= 0 || static_cast<int64_t>(r) * l < nz) r = l = 0;
  if (rows) *rows = r;
  if (lanes) *lanes = l;
  return 0;
}

int cpr_plan_describe(cpr_plan *p, int b, int *variant, int *team_rows, int *team_lanes) {
  if (!p || b < 0 || b >= p->nrays) return fail("bad plan or ray");
  if (variant) *variant = p->variant;
  if (team_rows) *team_rows = p->team_rows;
  if (team_lanes) *team_lanes = p->team_lanes;
  return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Allocation-free march (cpr_march_noalloc): the thread-per-ray march of a
// plan, with every buffer beyond the caller's inputs and outputs in one
// file-scope device array.  In host code `scratch` names a host-side stand-in,
// not device memory: the array's device address is looked up once per device
// with hipGetSymbolAddress and only that address is used.
// ---------------------------------------------------------------------------
namespace {

constexpr int64_t kScratchBytes = int64_t(64) << 20;  // CPR_NOALLOC_SCRATCH_BYTES
static_assert(kScratchBytes == CPR_NOALLOC_SCRATCH_BYTES, "header and library disagree on the scratch size");

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mu;
std::array<unsigned char *, 64> g_scratch_ptr{};  // device address of `scratch` per device

int scratch_of_current_device(unsigned char **out) {
  int dev = -1;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= static_cast<int>(g_scratch_ptr.size())) return fail("device index beyond the scratch table");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *ptr = nullptr;
    HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return fail("hipGetSymbolAddress returned no address for the scratch array");
    g_scratch_ptr[dev] = static_cast<unsigned char *>(ptr);
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

__global__ void k_fill_f64(double *x, int64_t n, double v) {
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x)
    x[i] = v;
}

int64_t align256(int64_t n) { return (n + 255) / 256 * 256; }

}  // namespace

extern "C" int cpr_march_noalloc(int precision, int nrays, const cpr_ray *rays, const cpr_env *envs,
                                 const double *env_data, int64_t env_len, const double *pade, int nvr_max,
                                 double *tll_d, double *cpl_d, void *stream) {
  if (!rays || !envs || !env_data || !pade || !tll_d || !cpl_d) return fail("null argument");
  if (nrays < 1) return fail("nrays must be >= 1");
  if (precision != CPR_C64 && precision != CPR_C128) return fail("precision must be 64 or 128");
  if (nvr_max < 0 || env_len < 1) return fail("bad sizes");
  const int np_ = rays[0].np;
  int nz_max = 0, max_steps = 0;
  for (int b = 0; b < nrays; ++b) {
    if (check_ray(rays[b], envs[b], env_len, b)) return -1;
    if (rays[b].np != np_) return fail("all rays must use the same np");
    if (rays[b].nvr > nvr_max) return fail("ray nvr exceeds nvr_max");
    nz_max = std::max(nz_max, rays[b].nz);
    max_steps = std::max(max_steps, rays[b].nsteps);
  }
  // scratch layout: ray tables, environment, Pade sets, march state, the
  // march columns of the precision, and (complex64) the FP64 starter columns
  const int64_t ld = nrays;
  const int64_t cells = (static_cast<int64_t>(nz_max) + 2) * ld;
  const int64_t rs = precision == CPR_C64 ? 4 : 8;
  const int64_t col_bytes[7] = {rs, rs, rs, 2 * rs, 2 * rs, 2 * rs, 2 * rs};
  const int64_t npade = static_cast<int64_t>(nrays) * 3 * 2 * np_ * 2;
  int64_t off = 0;
  const int64_t o_rays = off;
  off += align256(nrays * static_cast<int64_t>(sizeof(cpr_ray)));
  const int64_t o_envs = off;
  off += align256(nrays * static_cast<int64_t>(sizeof(cpr_env)));
  const int64_t o_env = off;
  off += align256(env_len * static_cast<int64_t>(sizeof(double)));
  const int64_t o_pade = off;
  off += align256(npade * static_cast<int64_t>(sizeof(double)));
  const int64_t o_state = off;
  off += align256(nrays * static_cast<int64_t>(sizeof(RayState)));
  int64_t o_col[7], o_dcol[7];
  for (int k = 0; k < 7; ++k) {
    o_col[k] = off;
    off += align256(cells * col_bytes[k]);
  }
  const bool own = precision == CPR_C64;  // complex128 starts in its own columns
  for (int k = 0; k < 7; ++k) {
    o_dcol[k] = off;
    if (own) off += align256(cells * col_bytes[k] * 8 / rs);
  }
  if (off > kScratchBytes) {
    char buf[200];
    std::snprintf(buf, sizeof(buf), "batch needs %lld bytes of scratch, more than the %lld of cpr_march_noalloc",
                  static_cast<long long>(off), static_cast<long long>(kScratchBytes));
    return fail(buf);
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail("no HIP device available");
  unsigned char *base = nullptr;
  if (scratch_of_current_device(&base)) return -1;
  hipStream_t st = static_cast<hipStream_t>(stream);
  HIPCHK(hipMemcpyAsync(base + o_rays, rays, sizeof(cpr_ray) * nrays, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(base + o_envs, envs, sizeof(cpr_env) * nrays, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(base + o_env, env_data, sizeof(double) * env_len, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(base + o_pade, pade, sizeof(double) * npade, hipMemcpyHostToDevice, st));
  // the outputs start at zero (entries past a ray's nvr are never written)
  const int64_t nline = static_cast<int64_t>(std::max(nvr_max, 1)) * ld;
  hipLaunchKernelGGL(k_fill_f64, dim3(256), dim3(256), 0, st, tll_d, nline, 0.0);
  hipLaunchKernelGGL(k_fill_f64, dim3(256), dim3(256), 0, st, cpl_d, 2 * nline, 0.0);
  HIPCHK(hipGetLastError());
  MarchArgs a{};
  a.rays = reinterpret_cast<const cpr_ray *>(base + o_rays);
  a.envs = reinterpret_cast<const cpr_env *>(base + o_envs);
  a.env = reinterpret_cast<const double *>(base + o_env);
  a.pade = reinterpret_cast<const double *>(base + o_pade);
  a.state = reinterpret_cast<RayState *>(base + o_state);
  a.nrays = nrays;
  a.np_ = np_;
  a.ld = ld;
  a.nvr_max = nvr_max;
  a.nvz_max = 0;
  a.grids = 0;
  void **cols[7] = {&a.f1, &a.f2, &a.f3, &a.ksq, &a.u, &a.cp, &a.dp};
  for (int k = 0; k < 7; ++k) *cols[k] = base + (own ? o_dcol[k] : o_col[k]);
  a.tll = tll_d;
  a.cpl = cpl_d;
  a.team_rays = nullptr;
  a.team_nrays = 0;
  const int block = 64;
  const int grid = (nrays + block - 1) / block;
  // self-starter in FP64 (cpr_selfs), the march (cpr_march) and the TL
  // post-pass (cpr_read_outputs), with the same kernels as a plan's
  hipLaunchKernelGGL(k_selfs, dim3(grid), dim3(block), 0, st, a);
  HIPCHK(hipGetLastError());
  const cx<double> *ud = static_cast<const cx<double> *>(a.u);
  for (int k = 0; k < 7; ++k) *cols[k] = base + o_col[k];
  if (own) {
    hipLaunchKernelGGL(k_field_from_starter<float>, dim3(grid), dim3(block), 0, st, a, ud);
    HIPCHK(hipGetLastError());
  }
  if (max_steps > 0) {
    if (precision == CPR_C64)
      hipLaunchKernelGGL(k_march_thread<float>, dim3(grid), dim3(block), 0, st, a, 0, max_steps);
    else
      hipLaunchKernelGGL(k_march_thread<double>, dim3(grid), dim3(block), 0, st, a, 0, max_steps);
    HIPCHK(hipGetLastError());
  }
  if (nvr_max > 0) {
    hipLaunchKernelGGL(k_tl_post, dim3(grid, nvr_max), dim3(block), 0, st, a);
    HIPCHK(hipGetLastError());
  }
  return 0;
}
