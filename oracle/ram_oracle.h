/* ram_oracle.h -- CPU oracle for the split-step Pade range march (test
 * infrastructure only; see ram_oracle.c for the reference lines restated). */
#ifndef RAM_ORACLE_H
#define RAM_ORACLE_H
#include <complex.h>

#ifdef __cplusplus
extern "C" {
#endif

/* One ray, parameters already resolved the way CuPyRAM.get_params/setup do
 * (cupyram/cupyram.py:460-608).  Profiles are row-major [depth][range]. */
typedef struct {
  double freq, zs, zr;
  double c0;     /* this ray's reference sound speed (k0 = 2 pi f / c0) */
  double lambda; /* wavelength used for the absorbing layer width */
  double lyrw;   /* absorbing layer width in wavelengths */
  double dr, dz, zmax, rmax, zmplt, rs;
  int np, ns, ndr, ndz;
  int n_ss, n_rp_ss;
  const double *z_ss, *rp_ss, *cw;
  int n_sb, n_rp_sb;
  const double *z_sb, *rp_sb, *cb, *rhob, *attn;
  int n_bt;
  const double *rbzb; /* [n_bt][2] (range, depth) */
} orc_ray;

int orc_pade(double freq, double c0, int np_, int ns, double dr, int ip, double complex *pd1,
             double complex *pd2);
int orc_dims(const orc_ray *in, int *nz, int *nvr, int *nvz, int *nsteps);
int orc_run(const orc_ray *in, double *tll, double complex *cpl, double *tlg, double complex *cpg,
            int *iz_hist);
/* TL line and cell-mean TL [nvz][nvr]: cell (i, j) = -10 log10 of the mean
 * of |f3 u|^2 / r over the fields at ranges (i ndr, (i+1) ndr] dr and rows
 * (j ndz, (j+1) ndz] (the build's coarse TL; no reference counterpart). */
int orc_run_cells(const orc_ray *in, double *tll, double *tlcell);
/* TL line, the per-step history trace [nsteps][4] = (iz, ss_ind, sb_ind, ns)
 * after each range update, and the field u [nz+2] after the last step. */
int orc_run_trace(const orc_ray *in, double *tll, int *trace, double complex *u_final);
int orc_run_batch(const orc_ray *rays, int nrays, double *const *tll, int nthreads);
int orc_profiles(const orc_ray *in, int ss_ind, int sb_ind, double *alpw, double *alpb, double *rhob,
                 double complex *ksqw, double complex *ksqb);
int orc_apply_factors(int nz, int iz, double dz, double k0, int np_, const double complex *pd1,
                      const double complex *pd2, const double *alpw, const double *alpb, const double *rhob,
                      const double complex *ksqw, const double complex *ksqb, double complex *u);

#ifdef __cplusplus
}
#endif
#endif
