/*
 * cupyram_hip.h -- C ABI of the MI355X (gfx950) split-step Pade range march.
 *
 * Plain C types only: pointers, sizes, int status codes (0 = ok, <0 = error;
 * cpr_last_error() returns the message of the last failure on this thread).
 * No exception crosses the ABI.  The library owns all device memory of a
 * plan; host pointers are borrowed for the duration of a call only.  One
 * plan may be used by one host thread at a time.
 *
 * Reference interfaces replaced (paths relative to the reference repository):
 *   cpr_pade / cpr_pade_batch   cupyram/pade.py:180       compute_pade_coefficients
 *   cpr_plan_create             cupyram/cupyram.py:528    CuPyRAM.setup (device
 *                               buffers, env upload) + cupyram/profl.py:142
 *                               profl_cuda_launcher (profiles, now on device)
 *   cpr_selfs                   cupyram/cupyram.py:1001   CuPyRAM.selfs
 *   cpr_march                   cupyram/fused_kernel.py:157 fused_sum_pade_solve,
 *                               cupyram/matrc.py:237 matrc_cuda_init_profiles,
 *                               cupyram/updat.py:143 updat_indices_cuda,
 *                               cupyram/outpt.py:61 outpt_cuda -- all fused,
 *                               over a range of march steps, with no host sync
 *   cpr_read_outputs            cupyram/cupyram.py:322-333 (D2H of tll/cpl/tlg/cpg)
 */
#ifndef CUPYRAM_HIP_H
#define CUPYRAM_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CPR_MAX_NP 16

/* precision of the march */
enum { CPR_C64 = 64, CPR_C128 = 128 };
/* kernel variant */
enum { CPR_VARIANT_AUTO = 0, CPR_VARIANT_THREAD = 1, CPR_VARIANT_TEAM = 2 };

/* Per-ray march parameters, resolved on the host the way CuPyRAM.get_params
 * and setup do (cupyram/cupyram.py:460-608). */
typedef struct {
  double dz;          /* depth step used for indices (iz, ir, is) */
  double zstep;       /* grid spacing zmax/(nz+1) of linspace(0, zmax, nz+2) */
  double zmax;        /* last grid depth */
  double dr;          /* range step */
  double k0, omega, c0;
  double lyrw_lambda; /* absorbing layer width (m) */
  double rs;          /* range at which the stability constraints switch off */
  double zs, zr;
  int32_t nz;         /* interior rows 1..nz; the column has nz+2 rows */
  int32_t nsteps;     /* march steps after the self-starter */
  int32_t iz0;        /* initial bathymetry row */
  int32_t ndr, ndz;
  int32_t nvr, nvz;   /* output range / depth counts of this ray */
  int32_t np;         /* Pade terms */
} cpr_ray;

/* Per-ray offsets (in doubles) into one packed environment array. */
typedef struct {
  int32_t n_ss, n_rp_ss;  /* water profile: depths x update ranges */
  int32_t n_sb, n_rp_sb;  /* seabed profile */
  int32_t n_bt;           /* bathymetry points (already extended to rmax) */
  int32_t pad_;
  int64_t z_ss, rp_ss, cw;          /* cw is [n_ss][n_rp_ss] */
  int64_t z_sb, rp_sb, cb, rhob, attn; /* z_sb absolute (z_ss[-1] added) */
  int64_t rbzb;                     /* [n_bt][2] */
} cpr_env;

typedef struct {
  int32_t nrays;
  int32_t precision;     /* CPR_C64 or CPR_C128 */
  int32_t variant;       /* CPR_VARIANT_* */
  int32_t compute_grids; /* 1: TL/CP grids every ndz rows */
  int32_t nvr_max, nvz_max;
  int32_t team_rows;     /* rows per lane of the team variant (0 = auto) */
  int32_t team_lanes;    /* lanes per ray of the team variant (0 = auto) */
  int32_t compute_cells; /* 1: cell-mean TL on the output grid (cpr_read_cells) */
} cpr_config;

typedef struct cpr_plan cpr_plan;

const char *cpr_last_error(void);
int cpr_device_count(int *count);
/* Name (NUL-terminated, truncated to name_len bytes) and free / total device
 * memory of HIP device `device`; any output may be NULL.  The batch-scaling
 * CLI's device report and batch sizing; replaces get_gpu_info
 * (cupyram/benchmark.py:22-40: getDeviceProperties and memGetInfo). */
int cpr_device_info(int device, char *name, int name_len, int64_t *free_bytes, int64_t *total_bytes);

/* FP64 Pade coefficients; pd1/pd2 are complex128 as interleaved re,im pairs
 * of length 2*np.  Returns -1 if Laguerre fails to converge. */
int cpr_pade(double freq, double c0, int np_, int ns, double dr, int ip, double *pd1, double *pd2);
/* n independent problems; out is [n][2][np] complex128 (pd1 then pd2). */
int cpr_pade_batch(int n, const double *freq, const double *c0, int np_, int ns, const double *dr,
                   int ip, double *out, int nthreads);

/* Upload a batch.  pade is [nrays][3][2][np] complex128: set 0 = march
 * (ip=1, ns), set 1 = self-starter (ip=2, ns), set 2 = march after rs
 * (ip=1, ns=0).  stream is a hipStream_t (NULL: the plan makes its own). */
int cpr_plan_create(const cpr_config *cfg, const cpr_ray *rays, const cpr_env *envs,
                    const double *env_data, int64_t env_len, const double *pade, void *stream,
                    cpr_plan **out);
int cpr_plan_destroy(cpr_plan *p);
/* Self-starter (FP64) and the first output; resets the march state.
 * Asynchronous on the plan's stream: its FP64 work columns are freed at the
 * plan's next synchronisation (cpr_sync, cpr_march_time, the output reads,
 * the next cpr_selfs or cpr_plan_destroy). */
int cpr_selfs(cpr_plan *p);
/* March steps [step_begin, step_end) of every ray (steps past a ray's own
 * nsteps are skipped).  Asynchronous on the plan's stream. */
int cpr_march(cpr_plan *p, int step_begin, int step_end);
/* Capture the launches of cpr_march(p, 0, nsteps) in chunks of
 * steps_per_launch into a hipGraph and replay it. */
int cpr_march_graph(cpr_plan *p, int steps_per_launch);
int cpr_sync(cpr_plan *p);
/* Time of the march kernels launched since the last cpr_selfs, measured with
 * hipEvents on the plan's stream (ms), and the number of launches. */
int cpr_march_time(cpr_plan *p, double *ms, int *nlaunch);
/* Outputs, ray-major: tll [nrays][nvr_max] f64, cpl [nrays][nvr_max] c128,
 * tlg [nrays][nvz_max][nvr_max] f64, cpg likewise c128.  Any may be NULL. */
int cpr_read_outputs(cpr_plan *p, double *tll, double *cpl, double *tlg, double *cpg);
/* Cell-mean TL, [nrays][nvz_max][nvr_max] f64 (plan created with
 * compute_cells): cell (i, j) holds -10 log10 of the mean of |p|^2 / r over
 * the ndr fields at ranges (i ndr, (i+1) ndr] dr and the ndz rows
 * (j ndz, (j+1) ndz] of each, p = f3 u as in the TL grid.  No counterpart in
 * the reference: this is the coarse TL field the learned surrogate is
 * trained on (the point samples of the TL grid carry the interference
 * pattern, which no smooth predictor reproduces). */
int cpr_read_cells(cpr_plan *p, double *tlc);
/* Field u after the last march step, [nrays][nz_max+2] complex128. */
int cpr_read_field(cpr_plan *p, double *u, int nz_max);
/* Bytes of device memory held by the plan (the staging buffer of the output
 * reads included once the first read has allocated it). */
int64_t cpr_plan_device_bytes(cpr_plan *p);
/* Team shape (rows per lane, lanes per ray) the library would pick for a
 * column of nz rows; rows = lanes = 0 when no register-resident team holds
 * it (such rays use the thread march).  Used by the host to bucket uneven
 * batches (no counterpart in the reference, which pads every ray to the
 * batch maximum, cupyram/cupyram.py:596-597). */
int cpr_team_shape(int precision, int nz, int *rows, int *lanes);
/* Kernel variant and team shape chosen for ray b (for tests/autotuning). */
int cpr_plan_describe(cpr_plan *p, int b, int *variant, int *team_rows, int *team_lanes);

/* Bytes of the device array behind cpr_march_noalloc (one per GPU). */
#define CPR_NOALLOC_SCRATCH_BYTES (64LL << 20)
/* Self-starter, the whole march and the TL line of every ray in one call
 * that allocates no device memory.  No counterpart in the reference: it is
 * for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  It runs the kernels of a
 * thread-variant plan (cpr_plan_create with CPR_VARIANT_THREAD, cpr_selfs,
 * cpr_march over every step, the TL of cpr_read_outputs) and gives the same
 * results; the ray tables, the march state and the column workspace live in
 * one library-owned device array of CPR_NOALLOC_SCRATCH_BYTES per GPU, held
 * from the library's load to the end of the process.  Inputs are host arrays
 * laid out as for cpr_plan_create (copied to the device on `stream`);
 * outputs are the caller's device arrays, range-major: tll_d
 * [nvr_max][nrays] f64 and cpl_d [nvr_max][nrays] complex128 (interleaved
 * re, im); entries past a ray's nvr are set to zero.  Asynchronous on
 * `stream` (a hipStream_t; NULL = the default stream).  Every call on a GPU
 * shares that GPU's array: calls that overlap on one GPU (other streams,
 * other host threads) race.  A batch that needs more than the array holds
 * (about 132 B per row and ray in complex64, 88 B in complex128, plus the
 * tables) is rejected with an error that states its need. */
int cpr_march_noalloc(int precision, int nrays, const cpr_ray *rays, const cpr_env *envs, const double *env_data,
                      int64_t env_len, const double *pade, int nvr_max, double *tll_d, double *cpl_d, void *stream);

/* ------------------------------------------------------------------------
 * Learned TL surrogate (a 1-D Fourier-operator network over depth).  These
 * have no counterpart in the reference: they are the training and inference
 * kernels of the surrogate that answers queries before the PE march does.
 * Device pointers in, asynchronous on `stream` (a hipStream_t, NULL = the
 * default stream); bf16 = IEEE bfloat16, row-major matrices.
 * ------------------------------------------------------------------------ */
enum { CPR_EPI_F32 = 0, CPR_EPI_BF16 = 1, CPR_EPI_GELU_BF16 = 2, CPR_EPI_ATOMIC = 3 };
/* C[b] = alpha op(A[b]) op(B[b]) (+ bias[n]) (+ addend[b]) for b < batch,
 * op(A)[m][k] = ta ? A[k*lda+m] : A[m*lda+k], op(B)[k][n] = tb ? B[n*ldb+k]
 * : B[k*ldb+n] (bf16 operands, FP32 accumulation on MFMA).  Epilogues: f32
 * store, bf16 store, GELU to bf16 (pre-activation stored to `pre` in C's
 * layout when non-NULL), or FP32 atomic add (the only one allowed with
 * splits > 1: K is split over `splits` blocks). */
int cpr_fno_gemm(int ta, int tb, int epi, int batch, int M, int N, int K, float alpha, const void *A, int64_t lda,
                 int64_t sA, const void *B, int64_t ldb, int64_t sB, void *C, int64_t ldc, int64_t sC,
                 const float *bias, const float *addend, int64_t ldd, int64_t sD, float *pre, int splits,
                 void *stream);
/* Truncated depth transform: x [S][Z][C] (bf16, or f32 if x_is_f32) ->
 * out [K][S][2C] bf16 with out[k][s][c] = sum_z bc[k][z] x[s][z][c] and
 * out[k][s][C+c] = sum_z bs[k][z] x[s][z][c].  Z <= 64, C <= 256, K <= 32. */
int cpr_fno_spec_analysis(const void *x, int x_is_f32, void *out, const float *bc, const float *bs, int S, int Z,
                          int C, int K, void *stream);
/* Inverse: y[s][z][c] (+)= sum_k bc[k][z] m[k][s][c] + bs[k][z] m[k][s][C+c];
 * m bf16, y f32. */
int cpr_fno_spec_synth(const void *m, float *y, const float *bc, const float *bs, int S, int Z, int C, int K,
                       int accumulate, void *stream);
/* dx (bf16) = dy * GELU'(pre); pre = NULL gives dx = dy. */
int cpr_fno_gelu_bwd(const float *dy, const float *pre, void *dx, int64_t n, void *stream);
/* Complex mode weights Wr + i Wi [K][C][C] <-> real blocks [[Wr,Wi],[-Wi,Wr]]
 * [K][2C][2C]: expand writes the bf16 block, fold maps a block gradient to
 * the gradients of Wr and Wi. */
int cpr_fno_spec_expand(const float *wr, const float *wi, void *blk, int K, int C, void *stream);
/* The transposed block [K][2C][2C] (blk_t[k][n][m] = blk[k][m][n]): the
 * operand of the forward mode mixing as a product with transposed B. */
int cpr_fno_spec_expand_t(const float *wr, const float *wi, void *blk_t, int K, int C, void *stream);
int cpr_fno_spec_fold(const float *g, float *gr, float *gi, int K, int C, void *stream);
/* out[n] += sum_m x[m][n], x [M][N] bf16 (or f32 if x_is_f32). */
int cpr_fno_colsum(const void *x, int x_is_f32, float *out, int64_t M, int N, void *stream);
/* Backward of y = GELU(pre) . p for a 1-column p [C] (f32): dx [S][Z][C]
 * bf16 = dy[s][z] p[c] GELU'(pre[s][z][c]) with dy [S][Z] bf16, and
 * part [S][C] f32 = the sums of dx over z (bias gradient partials). */
int cpr_fno_outer_gelu_bwd(const void *dy, const float *p, const float *pre, void *dx, float *part, int S, int Z,
                           int C, void *stream);
/* Masked squared error on column 0 of pred [P][ldp]: dpred (bf16, same
 * layout) = scale*mask*(pred-target) in column 0 and 0 elsewhere;
 * loss_sum[0] += sum (mask*(pred-target))^2, loss_sum[1] += sum mask. */
int cpr_fno_mse(const float *pred, int ldp, const float *target, const float *mask, void *dpred, int64_t P,
                float scale, float *loss_sum, void *stream);
/* Adam step `step` (1-based) on FP32 weights w with gradient g, moments m1,
 * m2 and decoupled weight decay wd; refreshes the bf16 copy wb if non-NULL. */
int cpr_fno_adam(float *w, const float *g, float *m1, float *m2, void *wb, int64_t n, float lr, float b1, float b2,
                 float eps, int step, float wd, void *stream);
int cpr_fno_fill(float *x, int64_t n, float v, void *stream);
/* Fused inference layer for Z = 64 depth points, C = F = 128 channels, 16
 * modes; two depth columns per block.  hout [S][64][128] bf16 =
 * act(hin . W + bias + syn(mh)), with wt = W transposed ([out][in] bf16),
 * syn(mh)[s][z][c] = sum_k' bsyn[z][k'] mh(k', s, c) over the 32 real and
 * imaginary mixed modes mh [16][S][256], bsyn [2][64][32] the hi and lo bf16
 * parts of the synthesis matrix, act = GELU; and (ana) xh [16][S][256] =
 * the truncated DFT of hout with ban [2][32][64] (hi, lo parts of the
 * analysis matrix).  pre (FP32 [S][64][128], or NULL) receives the
 * pre-activation hin . W + bias + syn(mh) for the backward pass of training.
 * Supported (syn, act, ana): (0,0,1) lift, (1,1,1) inner layer, (1,1,0) last
 * layer. */
int cpr_fno_layer(int syn, int act, int ana, const void *hin, const void *wt, const float *bias, const void *mh,
                  const void *bsyn, const void *ban, void *hout, void *xh, float *pre, int S, void *stream);
/* Fused backward layer, same shapes as cpr_fno_layer: dout [S][64][128]
 * bf16 = (din . W^T + syn(dxh)) * GELU'(pin), with w = W ([in][out] bf16,
 * the layer's own layout), dxh [16][S][256] the mode-mixing gradients (or
 * NULL: no spectral branch), bsyn the analysis matrix transposed ([2][64][32]
 * hi, lo), pin the FP32 pre-activation of the layer below (NULL: no GELU,
 * the lift); csum [S][128] f32 = the sums of dout over the 64 depths; and
 * (dmh non-NULL) dmh [16][S][256] = the synthesis matrix's transpose applied
 * to dout, with ban [2][32][64] the synthesis matrix (hi, lo).  At least one
 * of dxh and dmh is given. */
int cpr_fno_layer_bwd(const void *din, const void *w, const void *dxh, const void *bsyn, const void *ban,
                      const float *pin, void *dout, void *dmh, float *csum, int S, void *stream);
/* y [S*64] f32 = GELU(h . P1 + c1) . P2 + c2 with p1t = P1 transposed
 * ([out][in] bf16), p2 [128] f32. */
int cpr_fno_project(const void *h, const void *p1t, const float *c1, const float *p2, float c2, float *y, int S,
                    void *stream);
/* The last layer (cpr_fno_layer with syn = act = 1, ana = 0) followed by
 * cpr_fno_project in one kernel: y [S*64] f32 = GELU(h . P1 + c1) . P2 + c2
 * with h = GELU(hin . W + bias + syn(mh)) kept in LDS (inference). */
int cpr_fno_layer_project(const void *hin, const void *wt, const float *bias, const void *mh, const void *bsyn,
                          const void *p1t, const float *c1, const float *p2, float c2, float *y, int S,
                          void *stream);
/* The first layer (cpr_fno_layer with syn = act = 0, ana = 1) formed from
 * the input encoding's structure instead of a product with the encoded
 * input: sample s = q * nr + r (query q, range bin r), depth point z,
 * hout[s][z][c] = la[q][c] + lr[r][c] + lzt[c][z] + lzb[s] lu[c]
 * + (lzv[z] > lzb[s] ? lv[c] : 0), all f32 (la holds the per-query
 * features' share and the bias, lr the range features', lzt the depth
 * features' (channel-major, 16-byte aligned), lzb the local water depth,
 * lzv the depth of each point), and xh its truncated DFT as cpr_fno_layer
 * writes it.  S must be a multiple of nr (inference). */
int cpr_fno_lift(const float *la, const float *lr, const float *lzt, const float *lzb, const float *lzv,
                 const float *lu, const float *lv, int nr, const void *ban, void *hout, void *xh, int S,
                 void *stream);
/* The first layer's modes xh in the table form of cpr_fno_lift, without the
 * layer's output: A = la[q] + lr[r] + lzb[s] lu, xh[k'] = A tsum[k'][0] +
 * bz[k'] + lv tsum[k'][cut(s)] with tsum [2K][65] f32 the suffix sums over
 * depth of the analysis rows (tsum[k'][m] = sum_{z >= m} An[k'][z]), bz
 * [2K][128] f32 the analysis of the depth features' share, cut(s) the number
 * of depth points with lzv <= lzb[s]. */
int cpr_fno_lift_modes(const float *la, const float *lr, const float *lzb, const float *lzv, const float *lu,
                       const float *lv, const float *tsum, const float *bz, int nr, void *xh, int S, void *stream);
/* The second layer (cpr_fno_layer with syn = act = ana = 1) whose pointwise
 * product of the first layer's output is formed from the first layer's
 * tables times its weight (the same table form as cpr_fno_lift, bias
 * separate), so the first layer's output is never written (inference). */
int cpr_fno_layer_lift(const float *la, const float *lr, const float *lzt, const float *lzb, const float *lzv,
                       const float *lu, const float *lv, int nr, const float *bias, const void *mh,
                       const void *bsyn, const void *ban, void *hout, void *xh, int S, void *stream);
int cpr_fno_cast_bf16(const float *x, void *y, int64_t n, void *stream);
/* y [Cc][R] bf16 = x [R][Cc] f32 transposed. */
int cpr_fno_cast_bf16_t(const float *x, void *y, int R, int Cc, void *stream);

#ifdef __cplusplus
}
#endif
#endif
