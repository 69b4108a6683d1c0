"""Checked launchers of the surrogate kernels (libcupyram_hip.so, cpr_fno_*).

Tensors are torch device tensors (torch provides device memory and the
stream; every operation here runs in the library's own kernels).  Each
launcher verifies on the host that the operand extents cover every index the
kernel will touch before it launches.
"""
from __future__ import annotations

import ctypes as C

import torch

from .. import _abi

EPI_F32, EPI_BF16, EPI_GELU_BF16, EPI_ATOMIC = 0, 1, 2, 3
BK = 32


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _p(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def _need(t, dtype, n, name):
    if not t.is_cuda or not t.is_contiguous():
        raise ValueError(f"{name}: contiguous device tensor expected")
    if t.dtype != dtype:
        raise ValueError(f"{name}: {dtype} expected, got {t.dtype}")
    if t.numel() < n:
        raise ValueError(f"{name}: {t.numel()} elements, kernel reaches {n}")


def _extent(batch, stride, rows, ld, cols):
    """Elements spanned by `batch` matrices of rows x cols (row stride ld)."""
    return (batch - 1) * stride + (rows - 1) * ld + cols


def gemm(a, b, c, M, N, K, *, ta=False, tb=False, epi=EPI_BF16, batch=1, lda=None, ldb=None, ldc=None,
         sa=0, sb=0, sc=0, alpha=1.0, bias=None, addend=None, ldd=None, sd=0, pre=None, splits=1):
    """c[b] = alpha op(a[b]) op(b[b]) (+bias) (+addend), see cpr_fno_gemm."""
    lda = lda if lda is not None else (M if ta else K)
    ldb = ldb if ldb is not None else (K if tb else N)
    ldc = ldc if ldc is not None else N
    ra, ca = (K, M) if ta else (M, K)
    rb, cb = (N, K) if tb else (K, N)
    _need(a, torch.bfloat16, _extent(batch, sa, ra, lda, ca), "gemm A")
    _need(b, torch.bfloat16, _extent(batch, sb, rb, ldb, cb), "gemm B")
    cdt = torch.float32 if epi in (EPI_F32, EPI_ATOMIC) else torch.bfloat16
    _need(c, cdt, _extent(batch, sc, M, ldc, N), "gemm C")
    if bias is not None:
        _need(bias, torch.float32, N, "gemm bias")
    if addend is not None:
        ldd = ldd if ldd is not None else ldc
        _need(addend, torch.float32, _extent(batch, sd, M, ldd, N), "gemm addend")
    if pre is not None:
        _need(pre, torch.float32, _extent(batch, sc, M, ldc, N), "gemm pre")
    if splits > 1 and epi != EPI_ATOMIC:
        raise ValueError("split-K needs the atomic epilogue")
    _abi.check(_abi.lib().cpr_fno_gemm(int(ta), int(tb), epi, batch, M, N, K, float(alpha), _p(a), lda, sa,
                                       _p(b), ldb, sb, _p(c), ldc, sc, _p(bias), _p(addend),
                                       ldd if ldd is not None else 0, sd, _p(pre), int(splits), _stream()))


def splits_for(M, N, K, batch=1, target_blocks=1024):
    """Split-K factor that gives about target_blocks blocks."""
    tiles = -(-M // 128) * -(-N // 128) * batch
    s = max(1, min(target_blocks // max(tiles, 1), K // (4 * BK)))
    return s


def spec_analysis(x, out, bc, bs, S, Z, Cw, K):
    if x.dtype not in (torch.float32, torch.bfloat16):
        raise ValueError("spec_analysis: bf16 or f32 input")
    _need(x, x.dtype, S * Z * Cw, "spec_analysis x")
    _need(out, torch.bfloat16, K * S * 2 * Cw, "spec_analysis out")
    _need(bc, torch.float32, K * Z, "spec_analysis bc")
    _need(bs, torch.float32, K * Z, "spec_analysis bs")
    _abi.check(_abi.lib().cpr_fno_spec_analysis(_p(x), int(x.dtype == torch.float32), _p(out), _p(bc), _p(bs),
                                                S, Z, Cw, K, _stream()))


def spec_synth(m, y, bc, bs, S, Z, Cw, K, accumulate=False):
    _need(m, torch.bfloat16, K * S * 2 * Cw, "spec_synth m")
    _need(y, torch.float32, S * Z * Cw, "spec_synth y")
    _need(bc, torch.float32, K * Z, "spec_synth bc")
    _need(bs, torch.float32, K * Z, "spec_synth bs")
    _abi.check(_abi.lib().cpr_fno_spec_synth(_p(m), _p(y), _p(bc), _p(bs), S, Z, Cw, K, int(accumulate),
                                             _stream()))


def gelu_bwd(dy, pre, dx):
    n = dy.numel()
    _need(dy, torch.float32, n, "gelu_bwd dy")
    if pre is not None:
        _need(pre, torch.float32, n, "gelu_bwd pre")
    _need(dx, torch.bfloat16, n, "gelu_bwd dx")
    _abi.check(_abi.lib().cpr_fno_gelu_bwd(_p(dy), _p(pre), _p(dx), n, _stream()))


def spec_expand(wr, wi, blk, K, Cw, transposed=False):
    """Real block [[Wr, Wi], [-Wi, Wr]] of the complex mode weights, or its
    transpose."""
    _need(wr, torch.float32, K * Cw * Cw, "spec_expand wr")
    _need(wi, torch.float32, K * Cw * Cw, "spec_expand wi")
    _need(blk, torch.bfloat16, K * 4 * Cw * Cw, "spec_expand blk")
    fn = _abi.lib().cpr_fno_spec_expand_t if transposed else _abi.lib().cpr_fno_spec_expand
    _abi.check(fn(_p(wr), _p(wi), _p(blk), K, Cw, _stream()))


def spec_fold(g, gr, gi, K, Cw):
    _need(g, torch.float32, K * 4 * Cw * Cw, "spec_fold g")
    _need(gr, torch.float32, K * Cw * Cw, "spec_fold gr")
    _need(gi, torch.float32, K * Cw * Cw, "spec_fold gi")
    _abi.check(_abi.lib().cpr_fno_spec_fold(_p(g), _p(gr), _p(gi), K, Cw, _stream()))


def colsum(x, out, M, N):
    if x.dtype not in (torch.float32, torch.bfloat16):
        raise ValueError("colsum: bf16 or f32 input")
    _need(x, x.dtype, M * N, "colsum x")
    _need(out, torch.float32, N, "colsum out")
    _abi.check(_abi.lib().cpr_fno_colsum(_p(x), int(x.dtype == torch.float32), _p(out), M, N, _stream()))


def outer_gelu_bwd(dy, p, pre, dx, part, S, Z, Cw):
    """dx = dy p^T * GELU'(pre) (bf16), part[s][c] = sum_z dx[s][z][c] (cpr_fno_outer_gelu_bwd)."""
    _need(dy, torch.bfloat16, S * Z, "outer_gelu_bwd dy")
    _need(p, torch.float32, Cw, "outer_gelu_bwd p")
    _need(pre, torch.float32, S * Z * Cw, "outer_gelu_bwd pre")
    _need(dx, torch.bfloat16, S * Z * Cw, "outer_gelu_bwd dx")
    _need(part, torch.float32, S * Cw, "outer_gelu_bwd part")
    _abi.check(_abi.lib().cpr_fno_outer_gelu_bwd(_p(dy), _p(p), _p(pre), _p(dx), _p(part), int(S), int(Z),
                                                 int(Cw), _stream()))


def mse(pred, ldp, target, mask, dpred, P, scale, loss_sum):
    _need(pred, torch.float32, P * ldp, "mse pred")
    _need(target, torch.float32, P, "mse target")
    _need(mask, torch.float32, P, "mse mask")
    _need(dpred, torch.bfloat16, P * ldp, "mse dpred")
    _need(loss_sum, torch.float32, 2, "mse loss_sum")
    _abi.check(_abi.lib().cpr_fno_mse(_p(pred), ldp, _p(target), _p(mask), _p(dpred), P, float(scale),
                                      _p(loss_sum), _stream()))


def adam(w, g, m1, m2, wb, lr, step, b1=0.9, b2=0.999, eps=1e-8, wd=0.0):
    n = w.numel()
    for t, name in ((w, "w"), (g, "g"), (m1, "m1"), (m2, "m2")):
        _need(t, torch.float32, n, f"adam {name}")
    if wb is not None:
        _need(wb, torch.bfloat16, n, "adam wb")
    _abi.check(_abi.lib().cpr_fno_adam(_p(w), _p(g), _p(m1), _p(m2), _p(wb), n, float(lr), float(b1), float(b2),
                                       float(eps), int(step), float(wd), _stream()))


def fill(x, v=0.0):
    _need(x, torch.float32, x.numel(), "fill")
    _abi.check(_abi.lib().cpr_fno_fill(_p(x), x.numel(), float(v), _stream()))


def cast_bf16(x, y):
    _need(x, torch.float32, x.numel(), "cast x")
    _need(y, torch.bfloat16, x.numel(), "cast y")
    _abi.check(_abi.lib().cpr_fno_cast_bf16(_p(x), _p(y), x.numel(), _stream()))


FUSED_Z, FUSED_C, FUSED_K = 64, 128, 16


def fno_layer(hin, wt, bias, hout, S, *, mh=None, bsyn=None, ban=None, xh=None, act=True, pre=None):
    """Fused layer (cpr_fno_layer): hout = act(hin W + bias [+ syn(mh)]), xh =
    the truncated DFT of hout when ban/xh are given, and the FP32
    pre-activation into `pre` when given (training)."""
    Z, Cw, K2 = FUSED_Z, FUSED_C, 2 * FUSED_K
    syn, ana = mh is not None, xh is not None
    _need(hin, torch.bfloat16, S * Z * Cw, "fno_layer hin")
    _need(wt, torch.bfloat16, Cw * Cw, "fno_layer wt")
    _need(hout, torch.bfloat16, S * Z * Cw, "fno_layer hout")
    if bias is not None:
        _need(bias, torch.float32, Cw, "fno_layer bias")
    if syn:
        _need(mh, torch.bfloat16, FUSED_K * S * 2 * Cw, "fno_layer mh")
        _need(bsyn, torch.bfloat16, 2 * Z * K2, "fno_layer bsyn")
    if ana:
        _need(ban, torch.bfloat16, 2 * K2 * Z, "fno_layer ban")
        _need(xh, torch.bfloat16, FUSED_K * S * 2 * Cw, "fno_layer xh")
    if pre is not None:
        _need(pre, torch.float32, S * Z * Cw, "fno_layer pre")
    _abi.check(_abi.lib().cpr_fno_layer(int(syn), int(act), int(ana), _p(hin), _p(wt), _p(bias), _p(mh),
                                        _p(bsyn), _p(ban), _p(hout), _p(xh), _p(pre), int(S), _stream()))


def fno_layer_bwd(din, w, dout, csum, S, *, dxh=None, bsyn=None, ban=None, dmh=None, pin=None):
    """Fused backward layer (cpr_fno_layer_bwd): dout = (din W^T + syn(dxh)) *
    GELU'(pin), csum = its sums over depth, dmh = its transposed synthesis
    when ban/dmh are given."""
    Z, Cw, K2 = FUSED_Z, FUSED_C, 2 * FUSED_K
    _need(din, torch.bfloat16, S * Z * Cw, "fno_layer_bwd din")
    _need(w, torch.bfloat16, Cw * Cw, "fno_layer_bwd w")
    _need(dout, torch.bfloat16, S * Z * Cw, "fno_layer_bwd dout")
    _need(csum, torch.float32, S * Cw, "fno_layer_bwd csum")
    if dxh is None and dmh is None:
        raise ValueError("fno_layer_bwd: give dxh, dmh or both")
    if dxh is not None:
        _need(dxh, torch.bfloat16, FUSED_K * S * 2 * Cw, "fno_layer_bwd dxh")
        _need(bsyn, torch.bfloat16, 2 * Z * K2, "fno_layer_bwd bsyn")
    if dmh is not None:
        _need(ban, torch.bfloat16, 2 * K2 * Z, "fno_layer_bwd ban")
        _need(dmh, torch.bfloat16, FUSED_K * S * 2 * Cw, "fno_layer_bwd dmh")
    if pin is not None:
        _need(pin, torch.float32, S * Z * Cw, "fno_layer_bwd pin")
    _abi.check(_abi.lib().cpr_fno_layer_bwd(_p(din), _p(w), _p(dxh), _p(bsyn), _p(ban), _p(pin), _p(dout),
                                            _p(dmh), _p(csum), int(S), _stream()))


def fno_project(h, p1t, c1, p2, c2, y, S):
    """y = GELU(h P1 + c1) P2 + c2 (cpr_fno_project)."""
    Z, Cw = FUSED_Z, FUSED_C
    _need(h, torch.bfloat16, S * Z * Cw, "fno_project h")
    _need(p1t, torch.bfloat16, Cw * Cw, "fno_project p1t")
    _need(c1, torch.float32, Cw, "fno_project c1")
    _need(p2, torch.float32, Cw, "fno_project p2")
    _need(y, torch.float32, S * Z, "fno_project y")
    _abi.check(_abi.lib().cpr_fno_project(_p(h), _p(p1t), _p(c1), _p(p2), C.c_float(float(c2)), _p(y), int(S),
                                          _stream()))


def fno_layer_project(hin, wt, bias, mh, bsyn, p1t, c1, p2, c2, y, S):
    """The last fused layer and the projection in one kernel
    (cpr_fno_layer_project): y = GELU(GELU(hin W + bias + syn(mh)) P1 + c1)
    P2 + c2, the layer's output kept on chip."""
    Z, Cw, K2 = FUSED_Z, FUSED_C, 2 * FUSED_K
    _need(hin, torch.bfloat16, S * Z * Cw, "fno_layer_project hin")
    _need(wt, torch.bfloat16, Cw * Cw, "fno_layer_project wt")
    _need(bias, torch.float32, Cw, "fno_layer_project bias")
    _need(mh, torch.bfloat16, FUSED_K * S * 2 * Cw, "fno_layer_project mh")
    _need(bsyn, torch.bfloat16, 2 * Z * K2, "fno_layer_project bsyn")
    _need(p1t, torch.bfloat16, Cw * Cw, "fno_layer_project p1t")
    _need(c1, torch.float32, Cw, "fno_layer_project c1")
    _need(p2, torch.float32, Cw, "fno_layer_project p2")
    _need(y, torch.float32, S * Z, "fno_layer_project y")
    _abi.check(_abi.lib().cpr_fno_layer_project(_p(hin), _p(wt), _p(bias), _p(mh), _p(bsyn), _p(p1t), _p(c1),
                                                _p(p2), C.c_float(float(c2)), _p(y), int(S), _stream()))


def fno_lift(la, lr, lzt, lzb, lzv, lu, lv, nr, ban, hout, xh, S):
    """The first layer from the encoding's structure (cpr_fno_lift): for
    sample s = q * nr + r and depth z, hout = la[q] + lr[r] + lzt[:, z] +
    lzb[s] lu + (lzv[z] > lzb[s]) lv, and xh its truncated DFT."""
    Z, Cw, K2 = FUSED_Z, FUSED_C, 2 * FUSED_K
    if S % nr:
        raise ValueError("fno_lift: the samples are whole queries of nr range bins")
    _need(la, torch.float32, (S // nr) * Cw, "fno_lift la")
    _need(lr, torch.float32, nr * Cw, "fno_lift lr")
    _need(lzt, torch.float32, Cw * Z, "fno_lift lzt")
    _need(lzb, torch.float32, S, "fno_lift lzb")
    _need(lzv, torch.float32, Z, "fno_lift lzv")
    _need(lu, torch.float32, Cw, "fno_lift lu")
    _need(lv, torch.float32, Cw, "fno_lift lv")
    _need(ban, torch.bfloat16, 2 * K2 * Z, "fno_lift ban")
    _need(hout, torch.bfloat16, S * Z * Cw, "fno_lift hout")
    _need(xh, torch.bfloat16, FUSED_K * S * 2 * Cw, "fno_lift xh")
    _abi.check(_abi.lib().cpr_fno_lift(_p(la), _p(lr), _p(lzt), _p(lzb), _p(lzv), _p(lu), _p(lv), int(nr), _p(ban),
                                       _p(hout), _p(xh), int(S), _stream()))


def fno_lift_modes(la, lr, lzb, lzv, lu, lv, tsum, bz, nr, xh, S):
    """The first layer's modes in the table form without the layer's output
    (cpr_fno_lift_modes): xh = la[q] + lr[r] + lzb[s] lu times the analysis
    rows' sums, plus bz, plus lv times the analysis rows' sums over the
    depths below the local water depth (tsum [2K][Z+1] suffix sums)."""
    Z, Cw, K2 = FUSED_Z, FUSED_C, 2 * FUSED_K
    if S % nr:
        raise ValueError("fno_lift_modes: the samples are whole queries of nr range bins")
    _need(la, torch.float32, (S // nr) * Cw, "fno_lift_modes la")
    _need(lr, torch.float32, nr * Cw, "fno_lift_modes lr")
    _need(lzb, torch.float32, S, "fno_lift_modes lzb")
    _need(lzv, torch.float32, Z, "fno_lift_modes lzv")
    _need(lu, torch.float32, Cw, "fno_lift_modes lu")
    _need(lv, torch.float32, Cw, "fno_lift_modes lv")
    _need(tsum, torch.float32, K2 * (Z + 1), "fno_lift_modes tsum")
    _need(bz, torch.float32, K2 * Cw, "fno_lift_modes bz")
    _need(xh, torch.bfloat16, FUSED_K * S * 2 * Cw, "fno_lift_modes xh")
    _abi.check(_abi.lib().cpr_fno_lift_modes(_p(la), _p(lr), _p(lzb), _p(lzv), _p(lu), _p(lv), _p(tsum), _p(bz),
                                             int(nr), _p(xh), int(S), _stream()))


def fno_layer_lift(la, lr, lzt, lzb, lzv, lu, lv, nr, bias, mh, bsyn, ban, hout, xh, S):
    """The second layer with its pointwise product in the table form
    (cpr_fno_layer_lift): hout = GELU(la[q] + lr[r] + lzt[:, z] + lzb[s] lu +
    (lzv[z] > lzb[s]) lv + bias + syn(mh)), xh its truncated DFT."""
    Z, Cw, K2 = FUSED_Z, FUSED_C, 2 * FUSED_K
    if S % nr:
        raise ValueError("fno_layer_lift: the samples are whole queries of nr range bins")
    _need(la, torch.float32, (S // nr) * Cw, "fno_layer_lift la")
    _need(lr, torch.float32, nr * Cw, "fno_layer_lift lr")
    _need(lzt, torch.float32, Cw * Z, "fno_layer_lift lzt")
    _need(lzb, torch.float32, S, "fno_layer_lift lzb")
    _need(lzv, torch.float32, Z, "fno_layer_lift lzv")
    _need(lu, torch.float32, Cw, "fno_layer_lift lu")
    _need(lv, torch.float32, Cw, "fno_layer_lift lv")
    _need(bias, torch.float32, Cw, "fno_layer_lift bias")
    _need(mh, torch.bfloat16, FUSED_K * S * 2 * Cw, "fno_layer_lift mh")
    _need(bsyn, torch.bfloat16, 2 * Z * K2, "fno_layer_lift bsyn")
    _need(ban, torch.bfloat16, 2 * K2 * Z, "fno_layer_lift ban")
    _need(hout, torch.bfloat16, S * Z * Cw, "fno_layer_lift hout")
    _need(xh, torch.bfloat16, FUSED_K * S * 2 * Cw, "fno_layer_lift xh")
    _abi.check(_abi.lib().cpr_fno_layer_lift(_p(la), _p(lr), _p(lzt), _p(lzb), _p(lzv), _p(lu), _p(lv), int(nr),
                                             _p(bias), _p(mh), _p(bsyn), _p(ban), _p(hout), _p(xh), int(S),
                                             _stream()))


def cast_bf16_t(x, y):
    """y [C][R] bf16 = x [R][C] f32 transposed."""
    R, Cc = x.shape
    _need(x, torch.float32, R * Cc, "cast_bf16_t x")
    _need(y, torch.bfloat16, R * Cc, "cast_bf16_t y")
    _abi.check(_abi.lib().cpr_fno_cast_bf16_t(_p(x), _p(y), int(R), int(Cc), _stream()))
