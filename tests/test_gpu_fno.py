"""GPU tests of the surrogate kernels (BF16 MFMA products, depth transforms,
optimiser) and of the network's forward pass and gradients against the
float64 CPU restatement in tests/fno_reference.py."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from cupyram_amd.surrogate import FNO, FNOConfig, dft_matrices  # noqa: E402
from cupyram_amd.surrogate import kernels as k  # noqa: E402

from . import fno_reference as R  # noqa: E402

pytestmark = pytest.mark.gpu


def _bf(t):
    return t.to(torch.bfloat16)


def _rand(g, *shape):
    return torch.randn(*shape, generator=g, dtype=torch.float64)


@pytest.mark.parametrize("ta,tb", [(False, False), (False, True), (True, False), (True, True)])
@pytest.mark.parametrize("M,N,K", [(128, 128, 128), (200, 72, 45), (33, 257, 300), (1, 1, 1), (300, 1, 128)])
def test_gemm_f32_matches_float64(require_gpu, ta, tb, M, N, K):
    g = torch.Generator().manual_seed(M * 1000 + N * 10 + K)
    A = _bf(_rand(g, K, M) if ta else _rand(g, M, K))
    B = _bf(_rand(g, N, K) if tb else _rand(g, K, N))
    opA = (A.T if ta else A).double()
    opB = (B.T if tb else B).double()
    ref = opA @ opB
    C = torch.zeros(M, N, dtype=torch.float32, device="cuda")
    k.gemm(A.cuda(), B.cuda(), C, M, N, K, ta=ta, tb=tb, epi=k.EPI_F32)
    torch.cuda.synchronize()
    bound = (opA.abs() @ opB.abs()) * 1e-5 + 1e-6
    assert torch.all((C.cpu().double() - ref).abs() <= bound)


def test_gemm_epilogues(require_gpu):
    g = torch.Generator().manual_seed(3)
    M, N, K = 130, 140, 96
    A, B = _bf(_rand(g, M, K)), _bf(_rand(g, K, N))
    bias = _rand(g, N).float()
    add = _rand(g, M, N).float()
    acc = A.double() @ B.double() * 0.5 + bias.double() + add.double()
    gelu = torch.nn.functional.gelu(acc, approximate="tanh")
    Cb = torch.empty(M, N, dtype=torch.bfloat16, device="cuda")
    pre = torch.empty(M, N, dtype=torch.float32, device="cuda")
    k.gemm(A.cuda(), B.cuda(), Cb, M, N, K, alpha=0.5, bias=bias.cuda(), addend=add.cuda(), pre=pre,
           epi=k.EPI_GELU_BF16)
    Cf = torch.empty(M, N, dtype=torch.bfloat16, device="cuda")
    k.gemm(A.cuda(), B.cuda(), Cf, M, N, K, alpha=0.5, bias=bias.cuda(), addend=add.cuda(), epi=k.EPI_BF16)
    torch.cuda.synchronize()
    assert torch.allclose(pre.cpu().double(), acc, rtol=1e-5, atol=1e-4)
    # bf16 stores: within one bf16 rounding of the exact value
    assert torch.all((Cb.cpu().double() - gelu).abs() <= gelu.abs() * 2 ** -8 + 1e-4)
    assert torch.all((Cf.cpu().double() - acc).abs() <= acc.abs() * 2 ** -8 + 1e-4)


def test_gemm_batched_strided_and_split_k(require_gpu):
    g = torch.Generator().manual_seed(4)
    nb, M, N, K = 3, 100, 256, 1000
    A, B = _bf(_rand(g, nb, K, M)), _bf(_rand(g, nb, K, N))
    ref = torch.stack([A[b].double().T @ B[b].double() for b in range(nb)])
    C = torch.zeros(nb, M, N, dtype=torch.float32, device="cuda")
    k.gemm(A.cuda(), B.cuda(), C, M, N, K, ta=True, batch=nb, sa=K * M, sb=K * N, sc=M * N, epi=k.EPI_ATOMIC,
           splits=5)
    # an atomic epilogue accumulates into what C holds
    k.gemm(A.cuda(), B.cuda(), C, M, N, K, ta=True, batch=nb, sa=K * M, sb=K * N, sc=M * N, epi=k.EPI_ATOMIC,
           splits=1)
    torch.cuda.synchronize()
    bound = torch.stack([A[b].double().abs().T @ B[b].double().abs() for b in range(nb)]) * 2e-5 + 1e-5
    assert torch.all((C.cpu().double() - 2 * ref).abs() <= bound)


def test_gemm_rejects_out_of_range_operands(require_gpu):
    A = torch.zeros(10, 10, dtype=torch.bfloat16, device="cuda")
    C = torch.zeros(10, 10, dtype=torch.float32, device="cuda")
    with pytest.raises(ValueError):
        k.gemm(A, A, C, 10, 10, 11, epi=k.EPI_F32)
    with pytest.raises(ValueError):
        k.gemm(A, A, C, 10, 10, 10, epi=k.EPI_F32, splits=2)


def test_depth_transforms_match_numpy_fft(require_gpu):
    S, Z, C, K = 7, 64, 128, 16
    ac, as_, sc, ss = [torch.tensor(m, dtype=torch.float32, device="cuda") for m in dft_matrices(Z, K)]
    g = torch.Generator().manual_seed(5)
    x = _bf(_rand(g, S, Z, C))
    X = np.fft.rfft(x.double().numpy(), axis=1)[:, :K, :]                 # [S, K, C]
    out = torch.empty(K, S, 2 * C, dtype=torch.bfloat16, device="cuda")
    k.spec_analysis(x.cuda(), out, ac, as_, S, Z, C, K)
    torch.cuda.synchronize()
    o = out.cpu().double().numpy()
    re, im = o[:, :, :C].transpose(1, 0, 2), o[:, :, C:].transpose(1, 0, 2)
    scale = np.abs(x.double().numpy()).sum(axis=1)[:, None, :]
    assert np.all(np.abs(re - X.real) <= np.abs(X.real) * 2 ** -8 + 1e-5 * scale)
    assert np.all(np.abs(im - X.imag) <= np.abs(X.imag) * 2 ** -8 + 1e-5 * scale)
    # synthesis = irfft of the truncated spectrum
    m = _bf(_rand(g, K, S, 2 * C))
    md = m.double().numpy()
    full = np.zeros((S, Z // 2 + 1, C), complex)
    full[:, :K, :] = (md[:, :, :C] + 1j * md[:, :, C:]).transpose(1, 0, 2)
    y_ref = np.fft.irfft(full, n=Z, axis=1)
    y = torch.full((S, Z, C), 1.0, dtype=torch.float32, device="cuda")
    k.spec_synth(m.cuda(), y, sc, ss, S, Z, C, K, accumulate=True)
    torch.cuda.synchronize()
    np.testing.assert_allclose(y.cpu().double().numpy(), y_ref + 1.0, atol=2e-6)


def test_spectral_block_expand_and_fold(require_gpu):
    K, C = 3, 8
    g = torch.Generator().manual_seed(6)
    wr, wi = _rand(g, K, C, C).float(), _rand(g, K, C, C).float()
    blk = torch.empty(K, 2 * C, 2 * C, dtype=torch.bfloat16, device="cuda")
    blk_t = torch.empty_like(blk)
    k.spec_expand(wr.cuda(), wi.cuda(), blk, K, C)
    k.spec_expand(wr.cuda(), wi.cuda(), blk_t, K, C, transposed=True)
    G = _rand(g, K, 2 * C, 2 * C).float()
    gr = torch.empty(K, C, C, device="cuda")
    gi = torch.empty(K, C, C, device="cuda")
    k.spec_fold(G.cuda(), gr, gi, K, C)
    torch.cuda.synchronize()
    want = torch.cat([torch.cat([wr, wi], 2), torch.cat([-wi, wr], 2)], 1).to(torch.bfloat16)
    assert torch.equal(blk.cpu(), want)
    assert torch.equal(blk_t.cpu(), want.transpose(1, 2))
    assert torch.equal(gr.cpu(), G[:, :C, :C] + G[:, C:, C:])
    assert torch.equal(gi.cpu(), G[:, :C, C:] - G[:, C:, :C])


def test_adam_matches_torch_optim(require_gpu):
    g = torch.Generator().manual_seed(7)
    w0 = _rand(g, 1000).float()
    ref = w0.clone().requires_grad_(True)
    opt = torch.optim.Adam([ref], lr=1e-2, betas=(0.9, 0.999), eps=1e-8, weight_decay=0.01)
    w = w0.cuda()
    m1, m2 = torch.zeros_like(w), torch.zeros_like(w)
    wb = torch.empty(1000, dtype=torch.bfloat16, device="cuda")
    for step in range(1, 4):
        grad = _rand(g, 1000).float()
        ref.grad = grad.clone()
        opt.step()
        k.adam(w, grad.cuda(), m1, m2, wb, 1e-2, step, wd=0.01)
    torch.cuda.synchronize()
    torch.testing.assert_close(w.cpu(), ref.detach(), rtol=1e-5, atol=1e-6)
    assert torch.equal(wb.cpu(), w.cpu().to(torch.bfloat16))


def test_masked_mse(require_gpu):
    g = torch.Generator().manual_seed(8)
    P = 1000
    pred, tgt = _rand(g, P).float(), _rand(g, P).float()
    mask = (torch.rand(P, generator=g) > 0.3).float()
    loss = torch.zeros(2, device="cuda")
    dpred = torch.empty(P, dtype=torch.bfloat16, device="cuda")
    n = float(mask.sum())
    k.mse(pred.cuda(), 1, tgt.cuda(), mask.cuda(), dpred, P, 2.0 / n, loss)
    torch.cuda.synchronize()
    d = (pred.double() - tgt.double()) * mask.double()
    assert abs(float(loss[0]) - float((d ** 2).sum())) <= 1e-4 * float((d ** 2).sum())
    assert float(loss[1]) == n
    assert torch.all((dpred.cpu().double() - 2 * d / n).abs() <= (2 * d / n).abs() * 2 ** -8)


def _small_batch(seed, S=6):
    cfg = FNOConfig()
    g = torch.Generator().manual_seed(seed)
    x = _bf(_rand(g, S, cfg.depth_points, cfg.features) * 0.5)
    tgt = _rand(g, S * cfg.depth_points).float()
    mask = (torch.rand(S * cfg.depth_points, generator=g) > 0.2).float()
    return cfg, x, tgt, mask


def _rel(a, b):
    return float((a.double() - b.double()).norm() / b.double().norm().clamp_min(1e-30))


def test_forward_matches_float64_reference(require_gpu):
    cfg, x, tgt, mask = _small_batch(10)
    fno = FNO(cfg, seed=1)
    pred = fno.forward(x.cuda(), train=True).clone()
    pred_inf = fno.forward(x.cuda(), train=False).clone()
    torch.cuda.synchronize()
    ref = R.forward(R.reference_params(fno), x.double(), cfg)
    err = _rel(pred.cpu(), ref)
    err_inf = _rel(pred_inf.cpu(), ref)
    print(f"forward relative error: training (fused layers) {err:.2e}, layer-by-layer {err_inf:.2e}")
    assert err < 2e-2 and err_inf < 2e-2
    assert _rel(pred.cpu(), pred_inf.cpu()) < 2e-2
    # the training pass keeps every layer's FP32 pre-activation for backward
    for l in range(1, cfg.layers + 1):
        pre = fno._buf(f"pre{l}", (x.shape[0] * cfg.depth_points, cfg.width), torch.float32)
        assert torch.isfinite(pre).all() and float(pre.abs().max()) > 0


@pytest.mark.parametrize("fused,S", [(True, 6), (True, 7), (False, 6)])
def test_gradients_match_float64_autograd(require_gpu, fused, S):
    """Both backward paths (fused backward layers; transforms, products and
    GELU' as separate kernels) against float64 autograd; S = 7 leaves the
    second column of the fused kernels' last block empty."""
    cfg, x, tgt, mask = _small_batch(11, S=S)
    fno = FNO(cfg, seed=2)
    # non-zero biases so that every path carries signal
    for n in fno.names:
        if n.startswith(("b", "c")):
            fno.params[n].normal_(0.0, 0.1)
    fno.refresh()
    pred = fno.forward(x.cuda(), train=True)
    loss, dpred = fno.loss_grad(pred, tgt.cuda(), mask.cuda())
    fno.backward(dpred, fused=fused)
    torch.cuda.synchronize()
    ref_loss, _, ref_grads = R.loss_and_grads(R.reference_params(fno), x.double(), tgt.double(), mask.double(),
                                              cfg)
    gpu_loss = float(loss[0] / loss[1])
    errs = {n: _rel(fno.grads[n].cpu(), ref_grads[n]) for n in fno.names}
    print(f"loss {gpu_loss:.6f} vs {ref_loss:.6f}; gradient relative errors:",
          " ".join(f"{n}={e:.1e}" for n, e in errs.items()))
    assert abs(gpu_loss - ref_loss) <= 1e-2 * ref_loss
    bad = {n: e for n, e in errs.items() if e > 5e-2}
    assert not bad, bad


def test_gradient_matches_finite_differences_of_gpu_loss(require_gpu):
    cfg, x, tgt, mask = _small_batch(12)
    fno = FNO(cfg, seed=3)
    xd, td, md = x.cuda(), tgt.cuda(), mask.cuda()

    def gpu_loss():
        pred = fno.forward(xd, train=False)
        loss, _ = fno.loss_grad(pred, td, md)
        return float(loss[0] / loss[1])

    pred = fno.forward(xd, train=True)
    _, dpred = fno.loss_grad(pred, td, md)
    fno.backward(dpred)
    torch.cuda.synchronize()
    base = gpu_loss()
    # biases enter the epilogues in FP32, so small steps along them are exact
    for name in ("c2", "c1", "bp4", "bp2", "b0"):
        gvec = fno.grads[name].clone()
        v = gvec / gvec.norm()
        eps = 0.02 * base / float(gvec.norm())
        p0 = fno.params[name].clone()
        fno.params[name].copy_(p0 + eps * v)
        lp = gpu_loss()
        fno.params[name].copy_(p0 - eps * v)
        lm = gpu_loss()
        fno.params[name].copy_(p0)
        fd = (lp - lm) / (2 * eps)
        print(f"{name}: finite difference {fd:.5g}, gradient {float(gvec.norm()):.5g}")
        assert abs(fd - float(gvec.norm())) <= 0.05 * float(gvec.norm())


def test_training_reduces_loss_on_a_teacher_network(require_gpu):
    cfg = FNOConfig()
    g = torch.Generator().manual_seed(13)
    x = _bf(_rand(g, 32, cfg.depth_points, cfg.features) * 0.5).cuda()
    teacher = FNO(cfg, seed=20)
    tgt = teacher.forward(x).clone()
    mask = torch.ones_like(tgt)
    student = FNO(cfg, seed=21)
    losses = []
    for _ in range(60):
        pred = student.forward(x, train=True)
        loss, dpred = student.loss_grad(pred, tgt, mask)
        losses.append(float(loss[0] / loss[1]))
        student.backward(dpred)
        student.step(1e-3)
    torch.cuda.synchronize()
    print(f"teacher fit: loss {losses[0]:.4f} -> {losses[-1]:.4f}")
    assert losses[-1] < 0.3 * losses[0]


@pytest.mark.parametrize("S", [6, 7])
def test_fused_inference_matches_float64_reference(require_gpu, S):
    """The fused inference kernels (pointwise + spectral synthesis + GELU +
    next analysis in one layer kernel, fused projection) against the float64
    network and against the unfused path; S = 7 leaves the second column of
    the last block empty."""
    cfg, x, tgt, mask = _small_batch(12, S=S)
    fno = FNO(cfg, seed=3)
    for n in fno.names:
        if n.startswith(("b", "c")):
            fno.params[n].normal_(0.0, 0.1)
    fno.refresh()
    fused = fno.forward_fused(x.cuda()).clone()
    unfused = fno.forward(x.cuda(), train=False).clone()
    torch.cuda.synchronize()
    ref = R.forward(R.reference_params(fno), x.double(), cfg)
    e_ref, e_unf = _rel(fused.cpu(), ref), _rel(fused.cpu(), unfused.cpu())
    print(f"fused vs float64 {e_ref:.2e}, vs unfused {e_unf:.2e}, unfused vs float64 {_rel(unfused.cpu(), ref):.2e}")
    assert e_ref < 2e-2
    assert e_unf < 2e-2


def test_structured_lift_matches_encoded_input(require_gpu):
    """cpr_fno_lift (the first layer from the encoding's per-query, per-range
    and per-depth tables) against its float32 formula and against the first
    layer on the encoded input; cpr_fno_lift_modes (that layer's modes in
    closed form, without its output) against the float64 DFT of the formula;
    and the inference in the table form (forward_fused_lift: modes in closed
    form, the second layer's pointwise product from the tables) and
    forward_fused on the encoded input, both against the float64 network on
    the unrounded encoding; the forms differ by the bf16 rounding of the
    encoded columns and of the first layer's output."""
    import importlib
    E = importlib.import_module("cupyram_amd.surrogate.encode")
    from cupyram_amd.surrogate import dataset as D
    dom = D.DOMAINS["deep_munk"]
    q = D.deep_queries(2, 21)
    cfg = FNOConfig()
    fno = FNO(cfg, seed=7)
    for n in fno.names:
        if n.startswith(("b", "c")):
            fno.params[n].normal_(0.0, 0.1)
    fno.refresh()
    Z, Cw, Km = cfg.depth_points, cfg.width, cfg.modes
    n, nr = len(q), dom.nr
    S = n * nr
    x = E.encode(dom, q, cfg.features)
    inp = E.lift_inputs(dom, q)
    la, lr, lzt, lu, lv = fno.lift_tables(inp)
    h = torch.empty(S * Z, Cw, dtype=torch.bfloat16, device="cuda")
    xh = torch.empty(Km, S, 2 * Cw, dtype=torch.bfloat16, device="cuda")
    k.fno_lift(la, lr, lzt, inp.zb, inp.zv, lu, lv, nr, fno.ban, h, xh, S)
    zb = inp.zb.reshape(n, nr)
    below = (inp.zv[None, None, :] > zb[:, :, None]).float()
    ref = ((la[:, None, None, :] + lr[None, :, None, :]) + lzt.t()[None, None, :, :]
           + zb[:, :, None, None] * lu + below[..., None] * lv).reshape(S * Z, Cw)
    err = (h.float() - ref).abs()
    assert bool((err <= ref.abs() * 2.0 ** -8 + 1e-6).all()), float(err.max())
    # the first layer's modes without its output (cpr_fno_lift_modes) against
    # the float64 truncated DFT of the table form
    tsum, bz = fno._lift_second_tables()[:2]
    xm = torch.empty_like(xh)
    k.fno_lift_modes(la, lr, inp.zb, inp.zv, lu, lv, tsum, bz, nr, xm, S)
    h64 = ((la.double()[:, None, None, :] + lr.double()[None, :, None, :]) + lzt.double().t()[None, None, :, :]
           + zb.double()[:, :, None, None] * lu.double() + below.double()[..., None] * lv.double()).reshape(S, Z, Cw)
    an = torch.cat([fno.ac, fno.as_], 0).double()                  # [2K][Z]
    m64 = torch.einsum("kz,szc->ksc", an, h64)                       # [2K][S][C]
    xm_ref = torch.cat([m64[:Km], m64[Km:]], dim=2)                  # [K][S][2C]: real, then imaginary parts
    e_modes = (xm.double() - xm_ref).abs()
    assert bool((e_modes <= xm_ref.abs() * 2.0 ** -8 + 1e-4 * float(xm_ref.abs().max())).all()), float(e_modes.max())
    h2 = torch.empty_like(h)
    xh2 = torch.empty_like(xh)
    k.fno_layer(x, fno._transposed()["W0"], fno.params["b0"], h2, S, ban=fno.ban, xh=xh2, act=False)
    p_lift = fno.forward_fused_lift(inp).clone()
    p_enc = fno.forward_fused(x).clone()
    torch.cuda.synchronize()
    e_h, e_xh = _rel(h.float(), h2.float()), _rel(xh.float(), xh2.float())
    # the float64 network on the encoding before its bf16 rounding, rebuilt
    # from the tables as test_lift_inputs_reproduce_encode does
    xe = torch.zeros(n, nr, Z, cfg.features, dtype=torch.float64)
    zbc = zb.double().cpu()
    xe[..., :E.N_GLOBAL] = inp.g.double().cpu()[:, None, None, :]
    for j, row in enumerate(inp.fr_rows):
        xe[..., row] = inp.fr.double().cpu()[None, :, None, j]
    for j, row in enumerate(inp.fz_rows):
        xe[..., row] = inp.fz.double().cpu()[None, None, :, j]
    xe[..., E.F_HEIGHT] -= zbc[:, :, None]
    xe[..., E.F_ZREL] -= inp.g.double().cpu()[:, None, None, E.N_GLOBAL - 1]
    xe[..., E.F_ZB] = zbc[:, :, None]
    xe[..., E.F_BELOW] = below.double().cpu()
    ref64 = R.forward(R.reference_params(fno), xe.reshape(S, Z, cfg.features), cfg)
    e_lift, e_enc = _rel(p_lift.cpu(), ref64), _rel(p_enc.cpu(), ref64)
    e_xm = _rel(xm.float(), xh.float())
    print(f"structured vs encoded input: first layer {e_h:.2e}, its modes {e_xh:.2e} (closed form vs the kernel's "
          f"analysis of its bf16 output {e_xm:.2e}); prediction vs the float64 network on the unrounded encoding: "
          f"table form {e_lift:.2e}, encoded input {e_enc:.2e}")
    assert e_xm < 1e-2
    assert e_h < 1e-2 and e_xh < 1e-2
    # both within the fused path's bf16 arithmetic of later layers (observed
    # 2.7e-3 and 2.5e-3): the input rounding is not what limits either
    assert e_lift < 1e-2 and e_enc < 1e-2


@pytest.mark.parametrize("S", [6, 7])
def test_fused_last_layer_and_projection(require_gpu, S):
    """cpr_fno_layer_project (last layer and projection in one kernel, the
    layer's output kept in LDS) equals the last fused layer followed by the
    separate projection kernel up to the bf16 rounding both share."""
    cfg, x, tgt, mask = _small_batch(14, S=S)
    fno = FNO(cfg, seed=5)
    for n in fno.names:
        if n.startswith(("b", "c")):
            fno.params[n].normal_(0.0, 0.1)
    fno.refresh()
    Z, Cw, Km = cfg.depth_points, cfg.width, cfg.modes
    g = torch.Generator().manual_seed(15)
    h = (torch.randn(S * Z, Cw, generator=g) * 0.5).to(torch.bfloat16).cuda()
    mh = (torch.randn(Km, S, 2 * Cw, generator=g) * 0.1).to(torch.bfloat16).cuda()
    wt = fno._transposed()
    L = cfg.layers
    c2 = float(fno.params["c2"].item())
    hl = torch.empty(S * Z, Cw, dtype=torch.bfloat16, device="cuda")
    k.fno_layer(h, wt[f"Wp{L}"], fno.params[f"bp{L}"], hl, S, mh=mh, bsyn=fno.bsyn, act=True)
    y_sep = torch.empty(S * Z, device="cuda")
    k.fno_project(hl, wt["P1"], fno.params["c1"], fno.params["P2"].reshape(-1), c2, y_sep, S)
    y_one = torch.empty(S * Z, device="cuda")
    k.fno_layer_project(h, wt[f"Wp{L}"], fno.params[f"bp{L}"], mh, fno.bsyn, wt["P1"], fno.params["c1"],
                        fno.params["P2"].reshape(-1), c2, y_one, S)
    torch.cuda.synchronize()
    err = (y_one - y_sep).abs().max().item() / y_sep.abs().max().item()
    print(f"S={S}: max |fused - separate| / max |y| = {err:.2e}")
    assert err < 1e-5


def test_fused_inference_follows_training_steps(require_gpu):
    """The fused path rebuilds its transposed weights after optimiser steps."""
    cfg, x, tgt, mask = _small_batch(13)
    fno = FNO(cfg, seed=4)
    a = fno.forward_fused(x.cuda()).clone()
    pred = fno.forward(x.cuda(), train=True)
    _, dpred = fno.loss_grad(pred, tgt.cuda(), mask.cuda())
    fno.backward(dpred)
    fno.step(1e-2)
    b = fno.forward_fused(x.cuda()).clone()
    c = fno.forward(x.cuda(), train=False).clone()
    torch.cuda.synchronize()
    assert not torch.equal(a, b)
    assert _rel(b.cpu(), c.cpu()) < 2e-2
