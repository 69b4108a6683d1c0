// Internal declarations shared by the host and device sources of
// libcupyram_hip.so.
#pragma once
#include <cstddef>
#include <cstdint>
#include <string>

#include "../../include/cupyram_hip.h"

#if defined(__HIPCC__)
#define CPR_HD __host__ __device__
#else
#define CPR_HD
#endif

// Records m as the thread's last error (cpr_last_error) and returns -1.
int cpr_fail(const std::string &m);

int cpr_pade_impl(double freq, double c0, int np_, int ns, double dr, int ip, double *pd1, double *pd2);
int cpr_pade_batch_impl(int n, const double *freq, const double *c0, int np_, int ns, const double *dr, int ip,
                        double *out, int nthreads);

// March state of one ray, kept in device memory between launches.
struct RayState {
  int32_t iz;      // bathymetry row (water rows are 0..iz)
  int32_t bt_ind;  // bathymetry segment
  int32_t ss_ind;  // active water-profile column
  int32_t sb_ind;  // active seabed-profile column
  int32_t mdr;     // steps since the last output
  int32_t tlc;     // index of the last output range
  int32_t pset;    // active Pade set: 0 (ip=1, ns) or 2 (ip=1, ns=0)
  int32_t dirty;   // 1: the materialised medium must be rebuilt in full
};

// Everything a march kernel needs; passed by value.
// values per Pade term in the team march's constant table: q, 4q, rho / c,
// renormalisation factor, rho (re, im each)
constexpr int kTermStride = 10;

struct MarchArgs {
  const cpr_ray *rays;
  const cpr_env *envs;
  const double *env;
  const double *pade;  // [nrays][3][2][np] complex128
  RayState *state;
  int nrays;
  int np_;
  int64_t ld;          // ray stride of the depth-major column arrays
  int nvr_max, nvz_max;
  int grids;
  // column arrays [row][ld]; element type depends on the kernel precision
  void *f1, *f2, *f3, *ksq, *u, *cp, *dp;
  // outputs, depth-major for coalescing: tll/cpl [nvr_max][ld],
  // tlg/cpg [nvr_max][nvz_max][ld]
  double *tll, *tlg;
  double *cpl, *cpg;  // complex128 interleaved
  // cell-mean intensity: sums of |f3 u|^2 / r per output cell
  // [nvr_max][nvz_max][ld] (FP32) and their TL (FP64); null unless requested
  float *cells;
  double *tlcell;
  // team variant
  const int32_t *team_rays;  // rays handled by this launch (team variant)
  int team_nrays;
  void *dbuf;                // Galerkin coefficients of every team row (streamed
                             // by deep teams; saved between launches by the others)
  void *mbuf;                // FP64 water and bottom media of every team row
  const void *qtab;          // team per-term constants [nrays][3 sets][np][kTermStride] in R (k_team_terms)
  // layout of the field u: team_rows = 0: depth-major [row][ld]; > 0: the
  // team layout [ray][k][lane] of team_rows rows x team_lanes lanes
  int team_rows, team_lanes;
};
