// Kernels of the learned TL surrogate (a 1-D Fourier-operator network over
// depth), written for gfx950: BF16 operands, FP32 accumulation on MFMA.
//
//   k_gemm          C = alpha * op(A) op(B) [+ bias] [+ addend] [GELU], batched,
//                   optional split-K with FP32 atomics; 128x128 block tiles,
//                   four waves of 64x64, v_mfma_f32_32x32x16_bf16
//   k_spec_analysis modes[k][s][c] = sum_z Bc[k][z] x[s][z][c] (and Bs -> imag)
//   k_spec_synth    y[s][z][c] (+)= sum_k Bc[k][z] m[k][s][c] + Bs[k][z] m[k][s][C+c]
//   k_gelu_bwd, k_outer_gelu_bwd, k_colsum, k_mse, k_adam, k_fill, k_cast
//
// The surrogate has no counterpart in the reference; it is trained only on TL
// fields produced by this repository's march.
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>

#include <cstdint>
#include <string>

#include "cpr_internal.h"

typedef __bf16 bf16;
typedef bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

namespace {

constexpr int BM = 128, BN = 128, BK = 32, LDP = BK + 8;

// tanh form of GELU (Hendrycks & Gimpel 2016), 0.5 x (1 + tanh u) with
// u = k (x + 0.044715 x^3), written through the identity
// 0.5 (1 + tanh u) = s(u) = 1 / (1 + e^(-2u)) and evaluated with the hardware
// exp2 and reciprocal (about 1e-7 relative error, far below the bf16 rounding
// of the activations; e^(-2u) = inf gives s = 0, e^(-2u) = 0 gives s = 1).
__device__ inline float gelu_sig(float x, float &x2) {
  const float k2 = 2.0f * 0.7978845608028654f * 1.4426950408889634f;  // 2 k log2(e)
  x2 = x * x;
  const float v = k2 * (x + 0.044715f * x2 * x);
  return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(-v));
}
__device__ inline float gelu_f(float x) {
  float x2;
  return x * gelu_sig(x, x2);
}
// d/dx [x s] = s + x s (1 - s) 2k (1 + 3 * 0.044715 x^2)
__device__ inline float gelu_grad(float x) {
  const float k = 0.7978845608028654f;
  float x2;
  const float s = gelu_sig(x, x2);
  return s + x * s * (1.0f - s) * 2.0f * k * (1.0f + 3.0f * 0.044715f * x2);
}

struct GemmArgs {
  int M, N, K;
  const bf16 *A;
  int64_t lda, sA;  // batch stride
  const bf16 *B;
  int64_t ldb, sB;
  void *C;
  int64_t ldc, sC;
  const float *bias;    // [N] or null
  const float *addend;  // same layout as C (fp32) or null
  int64_t ldd, sD;
  float *pre;           // fp32 pre-activation out (GELU epilogue) or null
  float alpha;
  int splits;           // split-K factor (atomic epilogue when > 1)
  int kchunk;
  int vec;              // operands allow 16-byte loads: 16-byte aligned bases, ld and batch strides % 8 == 0
  int cvec;             // C allows 16-byte stores (bf16 C, 16-byte aligned base, ldc and sC % 8 == 0)
};

enum { EPI_F32 = 0, EPI_BF16 = 1, EPI_GELU_BF16 = 2, EPI_ATOMIC = 3 };

}  // namespace

template <bool TA, bool TB, int EPI>
__global__ void __launch_bounds__(256) k_gemm(GemmArgs g) {
  // operand tiles during the K loop; the bf16 output tile of the staged
  // epilogue afterwards (BM x BN)
  __shared__ __attribute__((aligned(16))) bf16 smem[BM * BN > (BM + BN) * LDP ? BM * BN : (BM + BN) * LDP];
  bf16 (*As)[LDP] = reinterpret_cast<bf16 (*)[LDP]>(smem);
  bf16 (*Bs)[LDP] = reinterpret_cast<bf16 (*)[LDP]>(smem + BM * LDP);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int z = blockIdx.z;
  const int batch = z / g.splits, split = z % g.splits;
  // Tile order: workgroups are handed to the eight XCDs in turn, each XCD
  // with its own L2.  Within each run of 8 M tiles, the N tiles of one M tile
  // get workgroup ids 8 apart, so they run on one XCD and all but the first
  // read the A tile from its L2; M tiles past the last whole run of 8 keep the
  // plain order.  (Bijective on the tiles of each z slice.)
  int tx = blockIdx.x, ty = blockIdx.y;
  if (gridDim.x > 1) {
    const int nx = gridDim.x;
    const int id = blockIdx.x + nx * blockIdx.y;
    const int grp = id / (8 * nx);
    if ((grp + 1) * 8 <= static_cast<int>(gridDim.y)) {
      const int i = id - grp * 8 * nx;
      ty = grp * 8 + (i & 7);
      tx = i >> 3;
    }
  }
  const int m0 = ty * BM, n0 = tx * BN;
  const bf16 *A = g.A + batch * g.sA;
  const bf16 *B = g.B + batch * g.sB;
  const int kbeg = split * g.kchunk;
  int kend = kbeg + g.kchunk;
  kend = kend < g.K ? kend : g.K;
  const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;
  f32x16 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.0f;
  const bf16 zero = static_cast<bf16>(0.0f);
  // The 16-byte pieces of op(A) and op(B) a thread stages, for the next K
  // tile, are loaded while the current tile is multiplied, so the tile's HBM
  // latency overlaps the MFMA work instead of following the barrier.
  bf16x8 ra[2], rb[2];
  auto vec_a = [&](int k) { return g.vec && k + BK <= kend && m0 + BM <= g.M; };
  auto vec_b = [&](int k) { return g.vec && k + BK <= kend && n0 + BN <= g.N; };
  auto fetch_a = [&](int k) {
    const bf16 *src = !TA ? A + static_cast<int64_t>(m0 + (tid >> 1)) * g.lda + k + (tid & 1) * 16
                          : A + static_cast<int64_t>(k + (tid >> 3)) * g.lda + m0 + (tid & 7) * 16;
    ra[0] = *reinterpret_cast<const bf16x8 *>(src);
    ra[1] = *reinterpret_cast<const bf16x8 *>(src + 8);
  };
  auto fetch_b = [&](int k) {
    const bf16 *src = !TB ? B + static_cast<int64_t>(k + (tid >> 3)) * g.ldb + n0 + (tid & 7) * 16
                          : B + static_cast<int64_t>(n0 + (tid >> 1)) * g.ldb + k + (tid & 1) * 16;
    rb[0] = *reinterpret_cast<const bf16x8 *>(src);
    rb[1] = *reinterpret_cast<const bf16x8 *>(src + 8);
  };
  bool va = vec_a(kbeg), vb = vec_b(kbeg);
  if (va) fetch_a(kbeg);
  if (vb) fetch_b(kbeg);
  for (int k0 = kbeg; k0 < kend; k0 += BK) {
    // stage op(A)[m0..m0+127][k0..k0+31] into As[m][k] and op(B) into Bs[n][k]:
    // whole tiles of 16-byte-loadable operands from the prefetched pieces,
    // the others element by element
    if (!TA) {
      const int r = tid >> 1, kk = (tid & 1) * 16;
      const int m = m0 + r;
      if (va) {
        *reinterpret_cast<bf16x8 *>(&As[r][kk]) = ra[0];
        *reinterpret_cast<bf16x8 *>(&As[r][kk + 8]) = ra[1];
      } else {
#pragma unroll
        for (int i = 0; i < 16; ++i) {
          const int k = k0 + kk + i;
          As[r][kk + i] = (m < g.M && k < kend) ? A[static_cast<int64_t>(m) * g.lda + k] : zero;
        }
      }
    } else {
      const int kk = tid >> 3, mm = (tid & 7) * 16;
      const int k = k0 + kk;
      if (va) {
#pragma unroll
        for (int i = 0; i < 8; ++i) {
          As[mm + i][kk] = ra[0][i];
          As[mm + 8 + i][kk] = ra[1][i];
        }
      } else {
#pragma unroll
        for (int i = 0; i < 16; ++i) {
          const int m = m0 + mm + i;
          As[mm + i][kk] = (m < g.M && k < kend) ? A[static_cast<int64_t>(k) * g.lda + m] : zero;
        }
      }
    }
    if (!TB) {
      const int kk = tid >> 3, nn = (tid & 7) * 16;
      const int k = k0 + kk;
      if (vb) {
#pragma unroll
        for (int i = 0; i < 8; ++i) {
          Bs[nn + i][kk] = rb[0][i];
          Bs[nn + 8 + i][kk] = rb[1][i];
        }
      } else {
#pragma unroll
        for (int i = 0; i < 16; ++i) {
          const int n = n0 + nn + i;
          Bs[nn + i][kk] = (n < g.N && k < kend) ? B[static_cast<int64_t>(k) * g.ldb + n] : zero;
        }
      }
    } else {
      const int r = tid >> 1, kk = (tid & 1) * 16;
      const int n = n0 + r;
      if (vb) {
        *reinterpret_cast<bf16x8 *>(&Bs[r][kk]) = rb[0];
        *reinterpret_cast<bf16x8 *>(&Bs[r][kk + 8]) = rb[1];
      } else {
#pragma unroll
        for (int i = 0; i < 16; ++i) {
          const int k = k0 + kk + i;
          Bs[r][kk + i] = (n < g.N && k < kend) ? B[static_cast<int64_t>(n) * g.ldb + k] : zero;
        }
      }
    }
    __syncthreads();
    const int k1 = k0 + BK;
    const bool na = k1 < kend && vec_a(k1), nb = k1 < kend && vec_b(k1);
    if (na) fetch_a(k1);
    if (nb) fetch_b(k1);
#pragma unroll
    for (int ks = 0; ks < BK; ks += 16) {
      bf16x8 af[2], bfr[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        // rows of LDP = 40 elements (80 bytes): every fragment is 16-byte aligned
        af[i] = *reinterpret_cast<const bf16x8 *>(&As[wm + i * 32 + (lane & 31)][ks + 8 * (lane >> 5)]);
        bfr[i] = *reinterpret_cast<const bf16x8 *>(&Bs[wn + i * 32 + (lane & 31)][ks + 8 * (lane >> 5)]);
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(af[i], bfr[j], acc[i][j], 0, 0, 0);
    }
    __syncthreads();
    va = na;
    vb = nb;
  }
  // epilogue: lane holds column n = lane%32 and rows 8*(r/4) + 4*(lane/32) + r%4
  if (EPI == EPI_BF16 && g.cvec && !g.addend && m0 + BM <= g.M && n0 + BN <= g.N) {
    // whole bf16 tile: through LDS (the K loop ended with a barrier), so
    // that each global store is a 16-byte piece of a 256-byte row run
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int nl = wn + j * 32 + (lane & 31);
        const float bn = g.bias ? g.bias[n0 + nl] : 0.0f;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const int ml = wm + i * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
          smem[ml * BN + nl] = static_cast<bf16>(g.alpha * acc[i][j][r] + bn);
        }
      }
    __syncthreads();
    bf16 *const Cb = static_cast<bf16 *>(g.C) + batch * g.sC + static_cast<int64_t>(m0) * g.ldc + n0;
#pragma unroll
    for (int it = 0; it < BM * BN / 8 / 256; ++it) {
      const int q = tid + 256 * it, ml = q / (BN / 8), n8 = (q % (BN / 8)) * 8;
      *reinterpret_cast<bf16x8 *>(Cb + static_cast<int64_t>(ml) * g.ldc + n8) =
          *reinterpret_cast<const bf16x8 *>(smem + ml * BN + n8);
    }
    return;
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int m = m0 + wm + i * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        const int n = n0 + wn + j * 32 + (lane & 31);
        if (m >= g.M || n >= g.N) continue;
        float v = g.alpha * acc[i][j][r];
        const int64_t oc = batch * g.sC + static_cast<int64_t>(m) * g.ldc + n;
        if (EPI == EPI_ATOMIC) {
          if (split == 0 && g.bias) v += g.bias[n];
          atomicAdd(static_cast<float *>(g.C) + oc, v);
          continue;
        }
        if (g.bias) v += g.bias[n];
        if (g.addend) v += g.addend[batch * g.sD + static_cast<int64_t>(m) * g.ldd + n];
        if (EPI == EPI_F32) {
          static_cast<float *>(g.C)[oc] = v;
        } else if (EPI == EPI_BF16) {
          static_cast<bf16 *>(g.C)[oc] = static_cast<bf16>(v);
        } else {
          if (g.pre) g.pre[batch * g.sC + static_cast<int64_t>(m) * g.ldc + n] = v;
          static_cast<bf16 *>(g.C)[oc] = static_cast<bf16>(gelu_f(v));
        }
      }
}

// modes: out[k][s][c] = sum_z bc[k][z] x[s][z][c], out[k][s][C+c] = sum_z bs[k][z] x[s][z][c]
// x is [S][Z][C] (T = bf16 or float), out [K][S][2C] bf16.  One block of C
// threads per sample; C <= 256, Z <= 64, K <= 32.
template <class T>
__global__ void k_spec_analysis(const T *x, bf16 *out, const float *bc, const float *bs, int S, int Z, int C, int K) {
  __shared__ float sb[2][32][64];
  const int s = blockIdx.x, c = threadIdx.x;
  for (int i = threadIdx.x; i < K * Z; i += blockDim.x) {
    sb[0][i / Z][i % Z] = bc[i];
    sb[1][i / Z][i % Z] = bs[i];
  }
  __syncthreads();
  if (s >= S || c >= C) return;
  float re[32], im[32];
#pragma unroll
  for (int k = 0; k < 32; ++k) re[k] = im[k] = 0.0f;
  for (int zz = 0; zz < Z; ++zz) {
    const float v = static_cast<float>(x[(static_cast<int64_t>(s) * Z + zz) * C + c]);
#pragma unroll
    for (int k = 0; k < 32; ++k) {
      if (k < K) {
        re[k] += sb[0][k][zz] * v;
        im[k] += sb[1][k][zz] * v;
      }
    }
  }
#pragma unroll
  for (int k = 0; k < 32; ++k) {
    if (k < K) {
      const int64_t o = (static_cast<int64_t>(k) * S + s) * 2 * C;
      out[o + c] = static_cast<bf16>(re[k]);
      out[o + C + c] = static_cast<bf16>(im[k]);
    }
  }
}

// y[s][z][c] (+)= sum_k bc[k][z] m[k][s][c] + bs[k][z] m[k][s][C+c]; y fp32.
__global__ void k_spec_synth(const bf16 *m, float *y, const float *bc, const float *bs, int S, int Z, int C, int K,
                             int accumulate) {
  __shared__ float sb[2][32][64];
  const int s = blockIdx.x, c = threadIdx.x;
  for (int i = threadIdx.x; i < K * Z; i += blockDim.x) {
    sb[0][i / Z][i % Z] = bc[i];
    sb[1][i / Z][i % Z] = bs[i];
  }
  __syncthreads();
  if (s >= S || c >= C) return;
  float re[32], im[32];
#pragma unroll
  for (int k = 0; k < 32; ++k) {
    if (k < K) {
      const int64_t o = (static_cast<int64_t>(k) * S + s) * 2 * C;
      re[k] = static_cast<float>(m[o + c]);
      im[k] = static_cast<float>(m[o + C + c]);
    } else {
      re[k] = im[k] = 0.0f;
    }
  }
  for (int zz = 0; zz < Z; ++zz) {
    float v = 0.0f;
#pragma unroll
    for (int k = 0; k < 32; ++k)
      if (k < K) v += sb[0][k][zz] * re[k] + sb[1][k][zz] * im[k];
    const int64_t o = (static_cast<int64_t>(s) * Z + zz) * C + c;
    y[o] = accumulate ? y[o] + v : v;
  }
}

// dx = dy * gelu'(pre) (pre = null: identity), written as bf16 for the next GEMMs
__global__ void k_gelu_bwd(const float *dy, const float *pre, bf16 *dx, int64_t n) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  dx[i] = static_cast<bf16>(pre ? dy[i] * gelu_grad(pre[i]) : dy[i]);
}

// dx[s][z][c] = dy[s][z] p[c] GELU'(pre[s][z][c]) (bf16), the gradient of the
// pre-activation of a GELU layer followed by a 1-column product; part[s][c]
// = sum_z dx[s][z][c] (f32 bias gradient partials).  One block of C threads
// per depth column s.
__global__ void k_outer_gelu_bwd(const bf16 *dy, const float *p, const float *pre, bf16 *dx, float *part, int Z,
                                 int C) {
  const int s = blockIdx.x, c = threadIdx.x;
  if (c >= C) return;
  const float pc = p[c];
  float cs = 0.0f;
  for (int z = 0; z < Z; ++z) {
    const int64_t row = static_cast<int64_t>(s) * Z + z;
    const float v = static_cast<float>(dy[row]) * pc * gelu_grad(pre[row * C + c]);
    dx[row * C + c] = static_cast<bf16>(v);
    cs += v;
  }
  part[static_cast<int64_t>(s) * C + c] = cs;
}

// Spectral weights: complex W[k] = Wr[k] + i Wi[k] ([K][C][C], row = input
// channel) acts on the real row vector [Re | Im] as the block matrix
// [[Wr, Wi], [-Wi, Wr]] ([K][2C][2C]).  expand writes the block (bf16), or
// its transpose [[Wr^T, -Wi^T], [Wi^T, Wr^T]] (TR: the operand layout of a
// product that reads 8 consecutive input channels per lane); fold maps a
// block gradient G to (dWr, dWi) = (G00 + G11, G01 - G10).
template <bool TR>
__global__ void k_spec_expand(const float *wr, const float *wi, bf16 *blk, int K, int C) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t n = static_cast<int64_t>(K) * C * C;
  if (i >= n) return;
  const int k = static_cast<int>(i / (C * C)), r = static_cast<int>((i / C) % C), c = static_cast<int>(i % C);
  bf16 *b = blk + static_cast<int64_t>(k) * 4 * C * C;
  const int L = 2 * C;
  if (TR) {
    b[c * L + r] = static_cast<bf16>(wr[i]);
    b[(C + c) * L + r] = static_cast<bf16>(wi[i]);
    b[c * L + C + r] = static_cast<bf16>(-wi[i]);
    b[(C + c) * L + C + r] = static_cast<bf16>(wr[i]);
  } else {
    b[r * L + c] = static_cast<bf16>(wr[i]);
    b[r * L + C + c] = static_cast<bf16>(wi[i]);
    b[(C + r) * L + c] = static_cast<bf16>(-wi[i]);
    b[(C + r) * L + C + c] = static_cast<bf16>(wr[i]);
  }
}

__global__ void k_spec_fold(const float *g, float *gr, float *gi, int K, int C) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t n = static_cast<int64_t>(K) * C * C;
  if (i >= n) return;
  const int k = static_cast<int>(i / (C * C)), r = static_cast<int>((i / C) % C), c = static_cast<int>(i % C);
  const float *b = g + static_cast<int64_t>(k) * 4 * C * C;
  const int L = 2 * C;
  gr[i] = b[r * L + c] + b[(C + r) * L + C + c];
  gi[i] = b[r * L + C + c] - b[(C + r) * L + c];
}

// out[n] += sum_m x[m][n] (column sums of an M x N bf16 or f32 matrix)
template <class T>
__global__ void k_colsum(const T *x, float *out, int64_t M, int N, int rows_per_block) {
  const int n = threadIdx.x + blockIdx.y * blockDim.x;
  if (n >= N) return;
  const int64_t r0 = static_cast<int64_t>(blockIdx.x) * rows_per_block;
  int64_t r1 = r0 + rows_per_block;
  r1 = r1 < M ? r1 : M;
  float s = 0.0f;
  for (int64_t r = r0; r < r1; ++r) s += static_cast<float>(x[r * N + n]);
  atomicAdd(out + n, s);
}

// Masked mean squared error on column 0 of pred [P][ldp]; grad written to
// dpred (bf16, same layout, other columns zero); loss_sum[0] += sum of squares,
// loss_sum[1] += count.  scale = 2 / (number of valid points).
__global__ void k_mse(const float *pred, int ldp, const float *target, const float *mask, bf16 *dpred, int64_t P,
                      float scale, float *loss_sum) {
  const int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= P) return;
  const float w = mask[p];
  const float d = (pred[p * ldp] - target[p]) * w;
  for (int c = 0; c < ldp; ++c) dpred[p * ldp + c] = static_cast<bf16>(c == 0 ? scale * d : 0.0f);
  atomicAdd(loss_sum, d * d);
  atomicAdd(loss_sum + 1, w);
}

// Adam (Kingma & Ba 2015) on FP32 master weights; refreshes the BF16 copy.
__global__ void k_adam(float *w, const float *g, float *m1, float *m2, bf16 *wb, int64_t n, float lr, float b1,
                       float b2, float eps, float c1, float c2, float wd) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i] + wd * w[i];
  const float a = b1 * m1[i] + (1.0f - b1) * gi;
  const float b = b2 * m2[i] + (1.0f - b2) * gi * gi;
  m1[i] = a;
  m2[i] = b;
  const float wn = w[i] - lr * (a / c1) / (sqrtf(b / c2) + eps);
  w[i] = wn;
  if (wb) wb[i] = static_cast<bf16>(wn);
}

__global__ void k_fill(float *x, int64_t n, float v) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) x[i] = v;
}

__global__ void k_cast_bf16(const float *x, bf16 *y, int64_t n) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) y[i] = static_cast<bf16>(x[i]);
}

// y[c][r] = bf16(x[r][c]) for an R x Cc matrix x (transposed cast)
__global__ void k_cast_bf16_t(const float *x, bf16 *y, int R, int Cc) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= static_cast<int64_t>(R) * Cc) return;
  const int r = static_cast<int>(i / Cc), c = static_cast<int>(i % Cc);
  y[static_cast<int64_t>(c) * R + r] = static_cast<bf16>(x[i]);
}

// ---------------------------------------------------------------------------
// Fused inference kernels (Z = 64 depth points, C = 128 channels, K = 16
// modes, F = 128 input features).  One block of four waves handles two depth
// columns (128 rows); wave w owns column sb = w/2 of the block and output
// channels 64*(w%2) .. +63.  MFMA operand fragments of
// v_mfma_f32_32x32x16_bf16: A row = lane%32 with 8 consecutive k at
// 8*(lane/32); B column = lane%32, same k; accumulator lane holds column
// lane%32, rows 8*(r/4) + 4*(lane/32) + r%4 (r = 0..15).
//
//   k_fno_layer<SYN, ACT, ANA>:
//     acc  = hin[rows] . W                         (W passed as WT[n][k])
//     acc += Bsyn[z][k'] . M[s][k'][c]             (SYN: spectral branch,
//            M the mixed modes [K][S][2C] with k' < K the real and k' >= K
//            the imaginary parts; Bsyn = hi + lo bf16 split of the FP32
//            synthesis matrix, so the product keeps ~16 mantissa bits)
//     hout = ACT ? GELU(acc + bias) : acc + bias   (bf16; acc + bias also to
//            pre in FP32 when given, for the backward pass of training)
//     xh[k'][s][c] = sum_z Ban[k'][z] hout[s][z][c] (ANA: the next layer's
//            truncated DFT, Ban = hi + lo split of the analysis matrix)
//   k_fno_layer<SYN, false, ANA, true> (BWD), one layer of the backward pass:
//     acc  = dpre_l[rows] . Wp^T + A^T(dxh)        (W passed as Wp[n][k]; SYN
//            with Bsyn = the analysis matrix transposed, M = the mixed mode
//            gradients dxh)
//     hout = acc * GELU'(pin)                      (the previous layer's
//            pre-activation gradient, bf16; pin = null: acc, for the lift)
//     csum[s][c] = sum_z hout[s][z][c]            (f32, bias gradient partials)
//     xh   = S^T(hout)                             (ANA with Ban = the
//            synthesis matrix: the previous layer's mode gradients)
//   k_fno_project: y[row] = GELU(h . P1 + c1) . P2 + c2 (f32)
//
// They replace the sequence spec_synth -> gemm(GELU, addend) -> spec_analysis
// of the layer kernels above (and, backward, gemm -> spec_synth -> gelu_bwd
// -> colsum -> spec_analysis): the FP32 spectral field y and the separate
// transforms never reach HBM.
// ---------------------------------------------------------------------------
namespace {

constexpr int FZ = 64, FC = 128, FK2 = 32;  // depth points, channels, 2 x modes
constexpr int HT_LD = FZ + 8;               // padded z stride of the transposed tile

__device__ inline bf16x8 ld8(const bf16 *p) { return *reinterpret_cast<const bf16x8 *>(p); }

struct LayerArgs {
  const bf16 *hin;   // [S][Z][C] (or the network input [S][Z][F], F = C)
  const bf16 *wt;    // [C out][C in]
  const float *bias; // [C]
  const bf16 *mh;    // [K][S][2C] mixed modes (SYN)
  const bf16 *bsyn;  // [2][Z][2K] hi, lo
  const bf16 *ban;   // [2][2K][Z] hi, lo
  bf16 *hout;        // [S][Z][C]
  bf16 *xh;          // [K][S][2C] (ANA)
  float *pre;        // [S][Z][C] FP32 pre-activation acc + bias, or null (training)
  const float *pin;  // BWD: [S][Z][C] FP32 pre-activation the output gradient passes through, or null
  float *csum;       // BWD: [S][C] per-column sums of the output (bias gradient partials)
  // PROJ: the projection fused into the last layer (h never reaches HBM)
  const bf16 *p1t;   // [C out][C in]
  const float *c1;   // [C]
  const float *p2;   // [C]
  float c2;
  float *y;          // [S][Z]
  // LIFT: the first layer formed from the encoding's structure instead of a
  // product with the encoded input: sample s = query q * nr + range bin r,
  // h = la[q] + lr[r] + lz[z] + zb[s] lu + (zv[z] > zb[s]) lv
  const float *la;   // [queries][C] the per-query features' share, plus the bias
  const float *lr;   // [nr][C] the range features' share
  const float *lzt;  // [C][Z] the depth features' share (channel-major)
  const float *lzb;  // [S] local water depth of each sample (encoding units)
  const float *lzv;  // [Z] depth of each point (encoding units)
  const float *lu;   // [C] weight of the local water depth
  const float *lv;   // [C] weight of the below-seabed flag
  int nr;
  int S;
};

// acc[i][j] (+)= hin rows . W for the wave's 64 rows (2 tiles) x 64 channels (2 tiles)
__device__ inline void pointwise_tiles(const bf16 *hrows, const bf16 *wt, int wn, int lane, f32x16 (&acc)[2][2]) {
#pragma unroll
  for (int kk = 0; kk < FC / 16; ++kk) {
    const int k = kk * 16 + 8 * (lane >> 5);
    bf16x8 a[2], b[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) a[i] = ld8(hrows + (i * 32 + (lane & 31)) * FC + k);
#pragma unroll
    for (int j = 0; j < 2; ++j) b[j] = ld8(wt + (wn + j * 32 + (lane & 31)) * FC + k);
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
  }
}

}  // namespace

template <bool SYN, bool ACT, bool ANA, bool BWD = false, bool PROJ = false, bool LIFT = false>
__global__ void __launch_bounds__(256) k_fno_layer(LayerArgs g) {
  static_assert(!(PROJ && (ANA || BWD)), "the projection follows the last forward layer only");
  static_assert(!LIFT || (ANA && !BWD && !PROJ && SYN == ACT),
                "the table form of the pointwise product: the first forward layer, or the second one");
  // row stride of the [sb][z][c] output tile: padded when the projection
  // reads it back as MFMA rows (16-byte reads of 32 rows at one k)
  constexpr int TL = PROJ ? FC + 8 : FC;
  __shared__ __attribute__((aligned(16))) bf16 ms[2][FK2][FC];    // mixed modes of the block's two columns
  __shared__ __attribute__((aligned(16))) bf16 ht[2][FC][HT_LD];  // output tile, transposed (channel-major)
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int sb = wave >> 1, wn = (wave & 1) * 64;
  const int s0 = blockIdx.x * 2, s = s0 + sb;
  const bool live = s < g.S;
  // the wave's A fragments of hin, loaded before the staging of the modes
  // below so that both HBM reads are in flight together
  bf16x8 afr[FC / 16][2];
  if (!LIFT && live) {
    const bf16 *hrows = g.hin + static_cast<int64_t>(s) * FZ * FC;
#pragma unroll
    for (int kk = 0; kk < FC / 16; ++kk)
#pragma unroll
      for (int i = 0; i < 2; ++i) afr[kk][i] = ld8(hrows + (i * 32 + (lane & 31)) * FC + kk * 16 + 8 * (lane >> 5));
  }
  if (SYN) {
    // stage M[s][k'][c] for both columns: 2 x 32 x 128 bf16 in 16-byte pieces
    for (int q = tid; q < 2 * FK2 * (FC / 8); q += 256) {
      const int c8 = (q % (FC / 8)) * 8, kp = (q / (FC / 8)) % FK2, b2 = q / (FC / 8 * FK2);
      const int ss = s0 + b2;
      bf16x8 v;
      if (ss < g.S) {
        const int k = kp % (FK2 / 2), im = kp / (FK2 / 2);
        v = ld8(g.mh + (static_cast<int64_t>(k) * g.S + ss) * 2 * FC + im * FC + c8);
      } else {
#pragma unroll
        for (int e = 0; e < 8; ++e) v[e] = static_cast<bf16>(0.0f);
      }
      *reinterpret_cast<bf16x8 *>(&ms[b2][kp][c8]) = v;
    }
    __syncthreads();
  }
  f32x16 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.0f;
  if constexpr (LIFT) {
    if (live) {
      // rows z = i*32 + 8*(r/4) + 4*(lane/32) + r%4 of channel n: four
      // consecutive depths per (i, r/4), one 16-byte read of the
      // channel-major depth table; the flag from a ballot over the 64 depths
      const int q = s / g.nr, rr = s - q * g.nr;
      const float zb = g.lzb[s];
      const unsigned long long below = __builtin_amdgcn_ballot_w64(g.lzv[lane] > zb);
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int n = wn + j * 32 + (lane & 31);
        const float base = g.la[static_cast<int64_t>(q) * FC + n] + g.lr[static_cast<int64_t>(rr) * FC + n];
        const float u = zb * g.lu[n], v = g.lv[n];
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int r4 = 0; r4 < 4; ++r4) {
            const int z0 = i * 32 + 8 * r4 + 4 * (lane >> 5);
            const float4 t = *reinterpret_cast<const float4 *>(g.lzt + n * FZ + z0);
            const float tz[4] = {t.x, t.y, t.z, t.w};
#pragma unroll
            for (int e = 0; e < 4; ++e)
              acc[i][j][4 * r4 + e] = (base + tz[e]) + u + (((below >> (z0 + e)) & 1ull) ? v : 0.0f);
          }
      }
    }
  } else if (live) {
#pragma unroll
    for (int kk = 0; kk < FC / 16; ++kk) {
      const int k = kk * 16 + 8 * (lane >> 5);
      bf16x8 bw[2];
#pragma unroll
      for (int j = 0; j < 2; ++j) bw[j] = ld8(g.wt + (wn + j * 32 + (lane & 31)) * FC + k);
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(afr[kk][i], bw[j], acc[i][j], 0, 0, 0);
    }
  }
  if (SYN) {
#pragma unroll
    for (int part = 0; part < 2; ++part) {
      const bf16 *bs = g.bsyn + part * FZ * FK2;
#pragma unroll
      for (int ks = 0; ks < FK2 / 16; ++ks) {
        const int kp = ks * 16 + 8 * (lane >> 5);
        bf16x8 a[2], b[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) a[i] = ld8(bs + (i * 32 + (lane & 31)) * FK2 + kp);
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
          for (int e = 0; e < 8; ++e) b[j][e] = ms[sb][kp + e][wn + j * 32 + (lane & 31)];
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j)
            acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
      }
    }
  }
  // epilogue: bias, activation, bf16 output (and its transposed copy for
  // ANA); BWD: the gradient through GELU' of pin, and the column sums.  The
  // bf16 outputs are held in registers and leave through LDS tiles (below),
  // so that every global store is a 16-byte piece of a contiguous run.
  // bf16 outputs held as pairs of consecutive r (packed, 32 VGPRs)
  unsigned hold[2][2][8];
  const bool wpre = !BWD && g.pre && live, rpin = BWD && g.pin && live;
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const int n = wn + j * 32 + (lane & 31);
    const float bn = !BWD && g.bias ? g.bias[n] : 0.0f;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[i][j][r] += bn;
    if (wpre) {
      // FP32 pre-activations (training), one branch for the whole tile
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const int z = i * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
          g.pre[(static_cast<int64_t>(s) * FZ + z) * FC + n] = acc[i][j][r];
        }
    }
    if (rpin) {
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const int z = i * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
          acc[i][j][r] *= gelu_grad(g.pin[(static_cast<int64_t>(s) * FZ + z) * FC + n]);
        }
    }
    float cs = 0.0f;
    // rows r, r+1 of a lane are depths z, z+1: converted as one pair and, for
    // the transposed tile, written as one 4-byte word at a constant offset
    unsigned *const hb = reinterpret_cast<unsigned *>(&ht[sb][n][4 * (lane >> 5)]);
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int rp = 0; rp < 8; ++rp) {
        f32x2 v = {acc[i][j][2 * rp], acc[i][j][2 * rp + 1]};
        if (BWD) {
          cs += v[0] + v[1];
        } else if (ACT) {
          v[0] = gelu_f(v[0]);
          v[1] = gelu_f(v[1]);
        }
        const unsigned pair = __builtin_bit_cast(unsigned, __builtin_convertvector(v, bf16x2));
        hold[j][i][rp] = pair;
        // a column past S is analysed but not stored
        if (ANA) hb[(i * 32 + 8 * (rp >> 1) + 2 * (rp & 1)) / 2] = pair;
      }
    if (BWD) {
      // lanes l and l ^ 32 hold the two z halves of column n
      cs += __shfl_xor(cs, 32, 64);
      if (live && lane < 32) g.csum[static_cast<int64_t>(s) * FC + n] = cs;
    }
  }
  // Output tiles in LDS, written after every wave has finished reading ht
  // (ANA) and ms (SYN): hout of the block's two columns as [sb][z][c] in the
  // ht array (32 of its 36 KB), and the modes xh as [k][sb][re|im][c] in the
  // ms array (16 KB).  In HBM both are contiguous runs: hout[s0 .. s0+1] is
  // 32 KB, and xh[k][s0 .. s0+1] is 1 KB for each mode k.
  bf16 *const htile = &ht[0][0][0];
  bf16 *const xtile = &ms[0][0][0];
  f32x16 a2[2];
  if (ANA) {
    __syncthreads();
    // xh[k'][s][c] = sum_z Ban[k'][z] h[s][z][c]: 32 x 64 (k' x c) per wave
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) a2[j][r] = 0.0f;
#pragma unroll
    for (int part = 0; part < 2; ++part) {
      const bf16 *ba = g.ban + part * FK2 * FZ;
#pragma unroll
      for (int ks = 0; ks < FZ / 16; ++ks) {
        const int z = ks * 16 + 8 * (lane >> 5);
        const bf16x8 a = ld8(ba + (lane & 31) * FZ + z);
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          const bf16x8 b = *reinterpret_cast<const bf16x8 *>(&ht[sb][wn + j * 32 + (lane & 31)][z]);
          a2[j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, a2[j], 0, 0, 0);
        }
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const int n = wn + j * 32 + (lane & 31);
    // per-lane bases, so that every write below takes a constant offset
    unsigned short *const tb = reinterpret_cast<unsigned short *>(htile) + (sb * FZ + 4 * (lane >> 5)) * TL + n;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int r = 0; r < 16; ++r)
        tb[(i * 32 + 8 * (r >> 2) + (r & 3)) * TL] = static_cast<unsigned short>(hold[j][i][r >> 1] >> (16 * (r & 1)));
    if (ANA) {
      // mode k' = 8 (r / 4) + 4 (lane / 32) + r % 4 is k = k' % 16 of part
      // im = k' / 16; the lane's 4 (lane / 32) never carries into bit 4
      bf16 *const xb = xtile + (4 * (lane >> 5) * 4 + sb * 2) * FC + n;
#pragma unroll
      for (int rp = 0; rp < 8; ++rp) {
        const f32x2 v = {a2[j][2 * rp], a2[j][2 * rp + 1]};
        const bf16x2 b = __builtin_convertvector(v, bf16x2);
#pragma unroll
        for (int e = 0; e < 2; ++e) {
          const int c = 8 * ((2 * rp + e) >> 2) + ((2 * rp + e) & 3);
          xb[((c & 15) * 4 + (c >> 4)) * FC] = b[e];
        }
      }
    }
  }
  __syncthreads();
  if constexpr (PROJ) {
    // y[row] = GELU(h . P1 + c1) . P2 + c2 for the block's 128 rows, h read
    // from the LDS tile; wave w takes its column's 64 rows and P1 outputs
    // 64*(w%2) .. +63, the two halves summed through LDS
    f32x16 pacc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int r = 0; r < 16; ++r) pacc[i][j][r] = 0.0f;
#pragma unroll
    for (int kk = 0; kk < FC / 16; ++kk) {
      const int k = kk * 16 + 8 * (lane >> 5);
      bf16x8 a[2], b[2];
#pragma unroll
      for (int i = 0; i < 2; ++i)
        a[i] = *reinterpret_cast<const bf16x8 *>(htile + (sb * FZ + i * 32 + (lane & 31)) * TL + k);
#pragma unroll
      for (int j = 0; j < 2; ++j) b[j] = ld8(g.p1t + (wn + j * 32 + (lane & 31)) * FC + k);
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
          pacc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i], b[j], pacc[i][j], 0, 0, 0);
    }
    float t[2][16];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int r = 0; r < 16; ++r) t[i][r] = 0.0f;
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int n = wn + j * 32 + (lane & 31);
      const float cn = g.c1[n], pn = g.p2[n];
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int r = 0; r < 16; ++r) t[i][r] += gelu_f(pacc[i][j][r] + cn) * pn;
    }
    // sum over the 32 channels held by lanes of the same half-wave
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        float v = t[i][r];
#pragma unroll
        for (int o = 16; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
        t[i][r] = v;
      }
    float *const part = reinterpret_cast<float *>(xtile);  // [sb][channel half][z]
    if ((lane & 31) == 0) {
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int r = 0; r < 16; ++r)
          part[(sb * 2 + (wave & 1)) * FZ + i * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3)] = t[i][r];
    }
    __syncthreads();
    if (tid < 2 * FZ) {
      const int b2 = tid / FZ, z = tid % FZ, ss = s0 + b2;
      if (ss < g.S) g.y[static_cast<int64_t>(ss) * FZ + z] = part[(b2 * 2) * FZ + z] + part[(b2 * 2 + 1) * FZ + z] + g.c2;
    }
    return;
  }
  // hout: 2 x 64 x 128 bf16 = 2048 pieces of 16 bytes, in order
  const int64_t hbase = static_cast<int64_t>(s0) * FZ * FC;
  const int hlive = (g.S - s0 >= 2 ? 2 : 1) * FZ * FC / 8;  // pieces of the live columns
#pragma unroll
  for (int it = 0; it < 8; ++it) {
    const int q = tid + 256 * it;
    if (q < hlive) *reinterpret_cast<bf16x8 *>(g.hout + hbase + q * 8) = *reinterpret_cast<const bf16x8 *>(htile + q * 8);
  }
  if (ANA) {
    // xh: 16 modes x (2 columns x 256) bf16 = 1024 pieces, 64 per mode
#pragma unroll
    for (int it = 0; it < 4; ++it) {
      const int q = tid + 256 * it;
      const int k = q >> 6, w = q & 63;
      if (s0 + (w >> 5) < g.S)
        *reinterpret_cast<bf16x8 *>(g.xh + (static_cast<int64_t>(k) * g.S + s0) * 2 * FC + w * 8) =
            *reinterpret_cast<const bf16x8 *>(xtile + q * 8);
    }
  }
}

// y[row] = GELU(h . P1 + c1) . P2 + c2 for the block's 128 rows.
__global__ void __launch_bounds__(256) k_fno_project(const bf16 *h, const bf16 *p1t, const float *c1, const float *p2,
                                                      float c2, float *y, int S) {
  __shared__ float part[2][2][FZ];  // [column sb][channel half][z]
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int sb = wave >> 1, wh = wave & 1, wn = wh * 64;
  const int s = blockIdx.x * 2 + sb;
  const bool live = s < S;
  f32x16 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.0f;
  if (live) pointwise_tiles(h + static_cast<int64_t>(s) * FZ * FC, p1t, wn, lane, acc);
  // per-lane products with P2, summed over the lane's two channel tiles
  float t[2][16];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int r = 0; r < 16; ++r) t[i][r] = 0.0f;
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const int n = wn + j * 32 + (lane & 31);
    const float cn = c1[n], pn = p2[n];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int r = 0; r < 16; ++r) t[i][r] += gelu_f(acc[i][j][r] + cn) * pn;
  }
  // sum over the 32 channels held by lanes of the same half-wave
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      float v = t[i][r];
#pragma unroll
      for (int o = 16; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
      t[i][r] = v;
    }
  if ((lane & 31) == 0) {
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int r = 0; r < 16; ++r) part[sb][wh][i * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3)] = t[i][r];
  }
  __syncthreads();
  if (tid < 2 * FZ) {
    const int b2 = tid / FZ, z = tid % FZ, ss = blockIdx.x * 2 + b2;
    if (ss < S) y[static_cast<int64_t>(ss) * FZ + z] = part[b2][0][z] + part[b2][1][z] + c2;
  }
}

// ---------------------------------------------------------------------------
// C ABI (declared in include/cupyram_hip.h)
// ---------------------------------------------------------------------------
namespace {
inline unsigned blocks_for(int64_t n, int t) { return static_cast<unsigned>((n + t - 1) / t); }

int launched(const char *name) {
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : cpr_fail(std::string(name) + ": " + hipGetErrorString(e));
}

template <bool TA, bool TB>
int launch_gemm_epi(const GemmArgs &g, int epi, int batch, hipStream_t st) {
  const dim3 grid((g.N + BN - 1) / BN, (g.M + BM - 1) / BM, batch * g.splits);
  switch (epi) {
    case EPI_F32:
      hipLaunchKernelGGL((k_gemm<TA, TB, EPI_F32>), grid, dim3(256), 0, st, g);
      break;
    case EPI_BF16:
      hipLaunchKernelGGL((k_gemm<TA, TB, EPI_BF16>), grid, dim3(256), 0, st, g);
      break;
    case EPI_GELU_BF16:
      hipLaunchKernelGGL((k_gemm<TA, TB, EPI_GELU_BF16>), grid, dim3(256), 0, st, g);
      break;
    case EPI_ATOMIC:
      hipLaunchKernelGGL((k_gemm<TA, TB, EPI_ATOMIC>), grid, dim3(256), 0, st, g);
      break;
    default:
      return cpr_fail("cpr_fno_gemm: unknown epilogue");
  }
  return launched("cpr_fno_gemm");
}
}  // namespace

extern "C" int cpr_fno_gemm(int ta, int tb, int epi, int batch, int M, int N, int K, float alpha, const void *A,
                            int64_t lda, int64_t sA, const void *B, int64_t ldb, int64_t sB, void *C, int64_t ldc,
                            int64_t sC, const float *bias, const float *addend, int64_t ldd, int64_t sD, float *pre,
                            int splits, void *stream) {
  if (M <= 0 || N <= 0 || K <= 0 || batch <= 0 || splits <= 0 || !A || !B || !C) return cpr_fail("cpr_fno_gemm: invalid arguments");
  if (splits > 1 && epi != EPI_ATOMIC) return cpr_fail("cpr_fno_gemm: invalid arguments");
  GemmArgs g;
  g.M = M;
  g.N = N;
  g.K = K;
  g.A = static_cast<const bf16 *>(A);
  g.lda = lda;
  g.sA = sA;
  g.B = static_cast<const bf16 *>(B);
  g.ldb = ldb;
  g.sB = sB;
  g.C = C;
  g.ldc = ldc;
  g.sC = sC;
  g.bias = bias;
  g.addend = addend;
  g.ldd = ldd;
  g.sD = sD;
  g.pre = pre;
  g.alpha = alpha;
  g.splits = splits;
  g.kchunk = ((K + splits - 1) / splits + BK - 1) / BK * BK;
  const auto al16 = [](const void *p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; };
  g.vec = al16(A) && al16(B) && lda % 8 == 0 && ldb % 8 == 0 && sA % 8 == 0 && sB % 8 == 0 && g.kchunk % 8 == 0;
  g.cvec = epi == EPI_BF16 && al16(C) && ldc % 8 == 0 && sC % 8 == 0;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (!ta && !tb) return launch_gemm_epi<false, false>(g, epi, batch, st);
  if (!ta && tb) return launch_gemm_epi<false, true>(g, epi, batch, st);
  if (ta && !tb) return launch_gemm_epi<true, false>(g, epi, batch, st);
  return launch_gemm_epi<true, true>(g, epi, batch, st);
}

extern "C" int cpr_fno_spec_analysis(const void *x, int x_is_f32, void *out, const float *bc, const float *bs, int S,
                                     int Z, int C, int K, void *stream) {
  if (S <= 0 || Z <= 0 || Z > 64 || C <= 0 || C > 256 || K <= 0 || K > 32) return cpr_fail("cpr_fno_spec_analysis: invalid arguments");
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (x_is_f32)
    hipLaunchKernelGGL(k_spec_analysis<float>, dim3(S), dim3(C), 0, st, static_cast<const float *>(x),
                       static_cast<bf16 *>(out), bc, bs, S, Z, C, K);
  else
    hipLaunchKernelGGL(k_spec_analysis<bf16>, dim3(S), dim3(C), 0, st, static_cast<const bf16 *>(x),
                       static_cast<bf16 *>(out), bc, bs, S, Z, C, K);
  return launched("cpr_fno_spec_analysis");
}

extern "C" int cpr_fno_spec_synth(const void *m, float *y, const float *bc, const float *bs, int S, int Z, int C,
                                  int K, int accumulate, void *stream) {
  if (S <= 0 || Z <= 0 || Z > 64 || C <= 0 || C > 256 || K <= 0 || K > 32) return cpr_fail("cpr_fno_spec_synth: invalid arguments");
  hipLaunchKernelGGL(k_spec_synth, dim3(S), dim3(C), 0, static_cast<hipStream_t>(stream),
                     static_cast<const bf16 *>(m), y, bc, bs, S, Z, C, K, accumulate);
  return launched("cpr_fno_spec_synth");
}

extern "C" int cpr_fno_gelu_bwd(const float *dy, const float *pre, void *dx, int64_t n, void *stream) {
  if (n <= 0 || !dy || !dx) return cpr_fail("cpr_fno_gelu_bwd: invalid arguments");
  hipLaunchKernelGGL(k_gelu_bwd, dim3(blocks_for(n, 256)), dim3(256), 0, static_cast<hipStream_t>(stream), dy, pre,
                     static_cast<bf16 *>(dx), n);
  return launched("cpr_fno_gelu_bwd");
}

extern "C" int cpr_fno_spec_expand(const float *wr, const float *wi, void *blk, int K, int C, void *stream) {
  if (K <= 0 || C <= 0) return cpr_fail("cpr_fno_spec_expand: invalid arguments");
  const int64_t n = static_cast<int64_t>(K) * C * C;
  hipLaunchKernelGGL(k_spec_expand<false>, dim3(blocks_for(n, 256)), dim3(256), 0, static_cast<hipStream_t>(stream),
                     wr, wi, static_cast<bf16 *>(blk), K, C);
  return launched("cpr_fno_spec_expand");
}

extern "C" int cpr_fno_spec_expand_t(const float *wr, const float *wi, void *blk_t, int K, int C, void *stream) {
  if (K <= 0 || C <= 0) return cpr_fail("cpr_fno_spec_expand_t: invalid arguments");
  const int64_t n = static_cast<int64_t>(K) * C * C;
  hipLaunchKernelGGL(k_spec_expand<true>, dim3(blocks_for(n, 256)), dim3(256), 0, static_cast<hipStream_t>(stream),
                     wr, wi, static_cast<bf16 *>(blk_t), K, C);
  return launched("cpr_fno_spec_expand_t");
}

extern "C" int cpr_fno_spec_fold(const float *g, float *gr, float *gi, int K, int C, void *stream) {
  if (K <= 0 || C <= 0) return cpr_fail("cpr_fno_spec_fold: invalid arguments");
  const int64_t n = static_cast<int64_t>(K) * C * C;
  hipLaunchKernelGGL(k_spec_fold, dim3(blocks_for(n, 256)), dim3(256), 0, static_cast<hipStream_t>(stream), g, gr,
                     gi, K, C);
  return launched("cpr_fno_spec_fold");
}

extern "C" int cpr_fno_colsum(const void *x, int x_is_f32, float *out, int64_t M, int N, void *stream) {
  if (M <= 0 || N <= 0 || !x || !out) return cpr_fail("cpr_fno_colsum: invalid arguments");
  const int rows = 256;
  const dim3 grid(blocks_for(M, rows), (N + 255) / 256);
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (x_is_f32)
    hipLaunchKernelGGL(k_colsum<float>, grid, dim3(N < 256 ? N : 256), 0, st, static_cast<const float *>(x), out, M,
                       N, rows);
  else
    hipLaunchKernelGGL(k_colsum<bf16>, grid, dim3(N < 256 ? N : 256), 0, st, static_cast<const bf16 *>(x), out, M,
                       N, rows);
  return launched("cpr_fno_colsum");
}

extern "C" int cpr_fno_outer_gelu_bwd(const void *dy, const float *p, const float *pre, void *dx, float *part, int S,
                                      int Z, int C, void *stream) {
  if (S <= 0 || Z <= 0 || C <= 0 || C > 1024 || !dy || !p || !pre || !dx || !part)
    return cpr_fail("cpr_fno_outer_gelu_bwd: invalid arguments");
  hipLaunchKernelGGL(k_outer_gelu_bwd, dim3(S), dim3(C), 0, static_cast<hipStream_t>(stream),
                     static_cast<const bf16 *>(dy), p, pre, static_cast<bf16 *>(dx), part, Z, C);
  return launched("cpr_fno_outer_gelu_bwd");
}

extern "C" int cpr_fno_mse(const float *pred, int ldp, const float *target, const float *mask, void *dpred,
                           int64_t P, float scale, float *loss_sum, void *stream) {
  if (P <= 0 || ldp <= 0) return cpr_fail("cpr_fno_mse: invalid arguments");
  hipLaunchKernelGGL(k_mse, dim3(blocks_for(P, 256)), dim3(256), 0, static_cast<hipStream_t>(stream), pred, ldp,
                     target, mask, static_cast<bf16 *>(dpred), P, scale, loss_sum);
  return launched("cpr_fno_mse");
}

extern "C" int cpr_fno_adam(float *w, const float *g, float *m1, float *m2, void *wb, int64_t n, float lr, float b1,
                            float b2, float eps, int step, float wd, void *stream) {
  if (n <= 0 || step < 1) return cpr_fail("cpr_fno_adam: invalid arguments");
  float c1 = 1.0f, c2 = 1.0f, p1 = 1.0f, p2 = 1.0f;
  for (int i = 0; i < step && i < 100000; ++i) {
    p1 *= b1;
    p2 *= b2;
  }
  c1 = 1.0f - p1;
  c2 = 1.0f - p2;
  hipLaunchKernelGGL(k_adam, dim3(blocks_for(n, 256)), dim3(256), 0, static_cast<hipStream_t>(stream), w, g, m1, m2,
                     static_cast<bf16 *>(wb), n, lr, b1, b2, eps, c1, c2, wd);
  return launched("cpr_fno_adam");
}

extern "C" int cpr_fno_fill(float *x, int64_t n, float v, void *stream) {
  if (n <= 0) return cpr_fail("cpr_fno_fill: invalid arguments");
  hipLaunchKernelGGL(k_fill, dim3(blocks_for(n, 256)), dim3(256), 0, static_cast<hipStream_t>(stream), x, n, v);
  return launched("cpr_fno_fill");
}

extern "C" int cpr_fno_cast_bf16_t(const float *x, void *y, int R, int Cc, void *stream) {
  if (R <= 0 || Cc <= 0 || !x || !y) return cpr_fail("cpr_fno_cast_bf16_t: invalid arguments");
  hipLaunchKernelGGL(k_cast_bf16_t, dim3(blocks_for(static_cast<int64_t>(R) * Cc, 256)), dim3(256), 0,
                     static_cast<hipStream_t>(stream), x, static_cast<bf16 *>(y), R, Cc);
  return launched("cpr_fno_cast_bf16_t");
}

extern "C" int cpr_fno_cast_bf16(const float *x, void *y, int64_t n, void *stream) {
  if (n <= 0) return cpr_fail("cpr_fno_cast_bf16: invalid arguments");
  hipLaunchKernelGGL(k_cast_bf16, dim3(blocks_for(n, 256)), dim3(256), 0, static_cast<hipStream_t>(stream), x,
                     static_cast<bf16 *>(y), n);
  return launched("cpr_fno_cast_bf16");
}

extern "C" int cpr_fno_layer(int syn, int act, int ana, const void *hin, const void *wt, const float *bias,
                             const void *mh, const void *bsyn, const void *ban, void *hout, void *xh, float *pre,
                             int S, void *stream) {
  if (S <= 0 || !hin || !wt || !hout || (syn && (!mh || !bsyn)) || (ana && (!ban || !xh)))
    return cpr_fail("cpr_fno_layer: invalid arguments");
  LayerArgs g{};
  g.hin = static_cast<const bf16 *>(hin);
  g.wt = static_cast<const bf16 *>(wt);
  g.bias = bias;
  g.mh = static_cast<const bf16 *>(mh);
  g.bsyn = static_cast<const bf16 *>(bsyn);
  g.ban = static_cast<const bf16 *>(ban);
  g.hout = static_cast<bf16 *>(hout);
  g.xh = static_cast<bf16 *>(xh);
  g.pre = pre;
  g.pin = nullptr;
  g.csum = nullptr;
  g.S = S;
  const dim3 grid((S + 1) / 2);
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (!syn && !act && ana)
    hipLaunchKernelGGL((k_fno_layer<false, false, true>), grid, dim3(256), 0, st, g);
  else if (syn && act && ana)
    hipLaunchKernelGGL((k_fno_layer<true, true, true>), grid, dim3(256), 0, st, g);
  else if (syn && act && !ana)
    hipLaunchKernelGGL((k_fno_layer<true, true, false>), grid, dim3(256), 0, st, g);
  else
    return cpr_fail("cpr_fno_layer: unsupported (syn, act, ana) combination");
  return launched("cpr_fno_layer");
}

// Modes of the first layer in the table form, without the layer's output:
// with h[s][z][c] = A[s][c] + lz[z][c] + (z >= cut(s)) lv[c] (A = la[q] +
// lr[r] + zb[s] lu, cut(s) the number of depth points at or above the local
// water depth), the truncated DFT sum_z An[k'][z] h[s][z][c] is
// A sig[k'] + bz[k'][c] + lv[c] tsum[k'][cut(s)], where sig[k'] = tsum[k'][0],
// tsum[k'][m] = sum_{z >= m} An[k'][z] and bz = An lz (the analysis matrix An
// in FP32).  One block of 128 threads (one per channel) per sample; xh as
// cpr_fno_layer writes it (bf16 [K][S][2C], real parts then imaginary).
__global__ void __launch_bounds__(128) k_lift_modes(const float *la, const float *lr, const float *lzb,
                                                     const float *lzv, const float *lu, const float *lv,
                                                     const float *tsum, const float *bz, int nr, bf16 *xh, int S) {
  const int s = blockIdx.x, c = threadIdx.x;
  const float zb = lzb[s];
  const int cut = __syncthreads_count(c < FZ && lzv[c < FZ ? c : 0] <= zb);
  const int q = s / nr, r = s - q * nr;
  const float a = la[static_cast<int64_t>(q) * FC + c] + lr[static_cast<int64_t>(r) * FC + c] + zb * lu[c];
  const float v = lv[c];
#pragma unroll 4
  for (int kp = 0; kp < FK2; ++kp) {
    const float x = a * tsum[kp * (FZ + 1)] + bz[kp * FC + c] + v * tsum[kp * (FZ + 1) + cut];
    const int k = kp % (FK2 / 2), im = kp / (FK2 / 2);
    xh[(static_cast<int64_t>(k) * S + s) * 2 * FC + im * FC + c] = static_cast<bf16>(x);
  }
}

// The first layer from the encoding's structure (see LayerArgs::la .. lv),
// with the bias in la, and its truncated DFT: replaces cpr_fno_layer(0, 0, 1,
// ...) on the encoded input, which reads S x 64 x 128 bf16 of it.
extern "C" int cpr_fno_lift(const float *la, const float *lr, const float *lzt, const float *lzb,
                            const float *lzv, const float *lu, const float *lv, int nr, const void *ban, void *hout,
                            void *xh, int S, void *stream) {
  if (S <= 0 || nr <= 0 || S % nr || !la || !lr || !lzt || !lzb || !lzv || !lu || !lv || !ban || !hout || !xh)
    return cpr_fail("cpr_fno_lift: invalid arguments");
  LayerArgs g{};
  g.la = la;
  g.lr = lr;
  g.lzt = lzt;
  g.lzb = lzb;
  g.lzv = lzv;
  g.lu = lu;
  g.lv = lv;
  g.nr = nr;
  g.ban = static_cast<const bf16 *>(ban);
  g.hout = static_cast<bf16 *>(hout);
  g.xh = static_cast<bf16 *>(xh);
  g.S = S;
  hipLaunchKernelGGL((k_fno_layer<false, false, true, false, false, true>), dim3((S + 1) / 2), dim3(256), 0,
                     static_cast<hipStream_t>(stream), g);
  return launched("cpr_fno_lift");
}

// The first layer's modes in the table form (k_lift_modes); tsum [2K][65],
// bz [2K][128] f32 from the analysis matrix and the depth features' share.
extern "C" int cpr_fno_lift_modes(const float *la, const float *lr, const float *lzb, const float *lzv,
                                  const float *lu, const float *lv, const float *tsum, const float *bz, int nr,
                                  void *xh, int S, void *stream) {
  if (S <= 0 || nr <= 0 || S % nr || !la || !lr || !lzb || !lzv || !lu || !lv || !tsum || !bz || !xh)
    return cpr_fail("cpr_fno_lift_modes: invalid arguments");
  hipLaunchKernelGGL(k_lift_modes, dim3(S), dim3(FC), 0, static_cast<hipStream_t>(stream), la, lr, lzb, lzv, lu, lv,
                     tsum, bz, nr, static_cast<bf16 *>(xh), S);
  return launched("cpr_fno_lift_modes");
}

// The second layer with its pointwise product in the table form (the first
// layer's tables multiplied by the layer's weights; bias separate), the
// spectral synthesis of mh, GELU and the next analysis, as cpr_fno_layer.
extern "C" int cpr_fno_layer_lift(const float *la, const float *lr, const float *lzt, const float *lzb,
                                  const float *lzv, const float *lu, const float *lv, int nr, const float *bias,
                                  const void *mh, const void *bsyn, const void *ban, void *hout, void *xh, int S,
                                  void *stream) {
  if (S <= 0 || nr <= 0 || S % nr || !la || !lr || !lzt || !lzb || !lzv || !lu || !lv || !mh || !bsyn || !ban ||
      !hout || !xh)
    return cpr_fail("cpr_fno_layer_lift: invalid arguments");
  LayerArgs g{};
  g.la = la;
  g.lr = lr;
  g.lzt = lzt;
  g.lzb = lzb;
  g.lzv = lzv;
  g.lu = lu;
  g.lv = lv;
  g.nr = nr;
  g.bias = bias;
  g.mh = static_cast<const bf16 *>(mh);
  g.bsyn = static_cast<const bf16 *>(bsyn);
  g.ban = static_cast<const bf16 *>(ban);
  g.hout = static_cast<bf16 *>(hout);
  g.xh = static_cast<bf16 *>(xh);
  g.S = S;
  hipLaunchKernelGGL((k_fno_layer<true, true, true, false, false, true>), dim3((S + 1) / 2), dim3(256), 0,
                     static_cast<hipStream_t>(stream), g);
  return launched("cpr_fno_layer_lift");
}

extern "C" int cpr_fno_layer_project(const void *hin, const void *wt, const float *bias, const void *mh,
                                     const void *bsyn, const void *p1t, const float *c1, const float *p2, float c2,
                                     float *y, int S, void *stream) {
  if (S <= 0 || !hin || !wt || !mh || !bsyn || !p1t || !c1 || !p2 || !y)
    return cpr_fail("cpr_fno_layer_project: invalid arguments");
  LayerArgs g{};
  g.hin = static_cast<const bf16 *>(hin);
  g.wt = static_cast<const bf16 *>(wt);
  g.bias = bias;
  g.mh = static_cast<const bf16 *>(mh);
  g.bsyn = static_cast<const bf16 *>(bsyn);
  g.p1t = static_cast<const bf16 *>(p1t);
  g.c1 = c1;
  g.p2 = p2;
  g.c2 = c2;
  g.y = y;
  g.S = S;
  hipLaunchKernelGGL((k_fno_layer<true, true, false, false, true>), dim3((S + 1) / 2), dim3(256), 0,
                     static_cast<hipStream_t>(stream), g);
  return launched("cpr_fno_layer_project");
}

extern "C" int cpr_fno_layer_bwd(const void *din, const void *w, const void *dxh, const void *bsyn, const void *ban,
                                 const float *pin, void *dout, void *dmh, float *csum, int S, void *stream) {
  const bool syn = dxh != nullptr, ana = dmh != nullptr;
  if (S <= 0 || !din || !w || !dout || !csum || (syn && !bsyn) || (ana && !ban))
    return cpr_fail("cpr_fno_layer_bwd: invalid arguments");
  if (!syn && !ana) return cpr_fail("cpr_fno_layer_bwd: unsupported combination (no dxh and no dmh)");
  LayerArgs g{};
  g.hin = static_cast<const bf16 *>(din);
  g.wt = static_cast<const bf16 *>(w);
  g.bias = nullptr;
  g.mh = static_cast<const bf16 *>(dxh);
  g.bsyn = static_cast<const bf16 *>(bsyn);
  g.ban = static_cast<const bf16 *>(ban);
  g.hout = static_cast<bf16 *>(dout);
  g.xh = static_cast<bf16 *>(dmh);
  g.pre = nullptr;
  g.pin = pin;
  g.csum = csum;
  g.S = S;
  const dim3 grid((S + 1) / 2);
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (syn && ana)
    hipLaunchKernelGGL((k_fno_layer<true, false, true, true>), grid, dim3(256), 0, st, g);
  else if (syn)
    hipLaunchKernelGGL((k_fno_layer<true, false, false, true>), grid, dim3(256), 0, st, g);
  else
    hipLaunchKernelGGL((k_fno_layer<false, false, true, true>), grid, dim3(256), 0, st, g);
  return launched("cpr_fno_layer_bwd");
}

extern "C" int cpr_fno_project(const void *h, const void *p1t, const float *c1, const float *p2, float c2, float *y,
                               int S, void *stream) {
  if (S <= 0 || !h || !p1t || !c1 || !p2 || !y) return cpr_fail("cpr_fno_project: invalid arguments");
  hipLaunchKernelGGL(k_fno_project, dim3((S + 1) / 2), dim3(256), 0, static_cast<hipStream_t>(stream),
                     static_cast<const bf16 *>(h), static_cast<const bf16 *>(p1t), c1, p2, c2, y, S);
  return launched("cpr_fno_project");
}
