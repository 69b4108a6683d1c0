"""A 1-D Fourier-operator network over depth, trained and run on the
library's BF16-MFMA kernels.

One sample is one range bin of one query: a depth column of Z points with F
input features per point.  The network is

    h0     = x W0 + b0                                    (lift, F -> C)
    h_l    = GELU(h_{l-1} Wp_l + bp_l + K_l(h_{l-1}))     l = 1..L
    K_l(h) = irfft_Z( rfft_Z(h)[:Km] * W_l )              (W_l complex [Km][C][C])
    q      = GELU(h_L P1 + c1)
    y      = q P2 + c2                                    (C -> 1, TL in normalised dB)

with rfft/irfft the real DFT pair over depth truncated to the first Km
modes (the convention of numpy.fft.rfft / irfft).  Weights are kept in FP32
(master copies and Adam moments) and used as BF16; every product accumulates
in FP32.  Activations are stored as BF16 (inputs of the products) and FP32
(pre-activations for the backward pass).

The network has no counterpart in the reference; it is trained only on TL
fields computed by this library's PE march.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np
import torch

from . import kernels as k


@dataclasses.dataclass(frozen=True)
class FNOConfig:
    features: int = 128   # input features per depth point (padded)
    width: int = 128      # channels C
    modes: int = 16       # retained depth modes Km
    layers: int = 4
    depth_points: int = 64  # Z


def dft_matrices(Z: int, Km: int):
    """Analysis (rfft) and synthesis (irfft) matrices of the truncated real DFT.

    analysis:  X_k = sum_z x_z (cos - i sin)(2 pi k z / Z)      -> (ac, as_) [Km][Z]
    synthesis: x_z = (1/Z)(Re X_0 + 2 sum_{k>=1} Re(X_k e^{+2 pi i k z / Z}))
                                                                -> (sc, ss) [Km][Z]
    """
    z = np.arange(Z)[None, :]
    kk = np.arange(Km)[:, None]
    th = 2.0 * np.pi * kk * z / Z
    ac, as_ = np.cos(th), -np.sin(th)
    w = np.where(kk == 0, 1.0 / Z, 2.0 / Z)
    sc, ss = w * np.cos(th), -w * np.sin(th)
    return ac, as_, sc, ss


class FNO:
    """Parameters, optimiser state and the forward/backward launch sequence."""

    PARAM_ORDER = ("W0", "b0", *[f"{n}{l}" for l in range(1, 9) for n in ("Wr", "Wi", "Wp", "bp")],
                   "P1", "c1", "P2", "c2")

    def __init__(self, cfg: FNOConfig = FNOConfig(), seed: int = 0, device="cuda"):
        if cfg.layers > 8:
            raise ValueError("at most 8 layers")
        self.cfg = cfg
        self.device = torch.device(device)
        F, Cw, Km, L = cfg.features, cfg.width, cfg.modes, cfg.layers
        g = torch.Generator().manual_seed(seed)

        def rnd(*shape, std):
            return (torch.randn(*shape, generator=g, dtype=torch.float64) * std).float()

        p = {"W0": rnd(F, Cw, std=1.0 / math.sqrt(F)), "b0": torch.zeros(Cw)}
        for l in range(1, L + 1):
            p[f"Wr{l}"] = rnd(Km, Cw, Cw, std=0.5 / Cw)
            p[f"Wi{l}"] = rnd(Km, Cw, Cw, std=0.5 / Cw)
            p[f"Wp{l}"] = rnd(Cw, Cw, std=1.0 / math.sqrt(Cw))
            p[f"bp{l}"] = torch.zeros(Cw)
        p["P1"] = rnd(Cw, Cw, std=1.0 / math.sqrt(Cw))
        p["c1"] = torch.zeros(Cw)
        p["P2"] = rnd(Cw, 1, std=1.0 / math.sqrt(Cw))
        p["c2"] = torch.zeros(1)
        self.names = [n for n in self.PARAM_ORDER if n in p]
        self.params = {n: p[n].to(self.device).contiguous() for n in self.names}
        self._alloc_state()
        ac, as_, sc, ss = dft_matrices(cfg.depth_points, Km)
        to = lambda a: torch.tensor(a, dtype=torch.float32, device=self.device).contiguous()  # noqa: E731
        self.ac, self.as_, self.sc, self.ss = to(ac), to(as_), to(sc), to(ss)
        # the fused inference kernels take the transforms as bf16 hi + lo pairs:
        # synthesis [Z][2K] (real-part rows, then imaginary), analysis [2K][Z]
        self.fusable = (cfg.depth_points, cfg.width, cfg.modes, cfg.features) == (64, 128, 16, 128)
        if self.fusable:
            syn = np.concatenate([sc, ss], axis=0).T          # [Z][2K]
            ana = np.concatenate([ac, as_], axis=0)           # [2K][Z]
            self.bsyn = self._hi_lo(syn)
            self.ban = self._hi_lo(ana)
            # FP32 analysis matrix: the first layer's modes in the table form
            self.ana = to(ana)
            # the backward pass applies the transposes: the analysis matrix
            # transposed as its synthesis ([Z][2K]) and the synthesis matrix
            # transposed as its analysis ([2K][Z])
            self.bsyn_t = self._hi_lo(ana.T)
            self.ban_t = self._hi_lo(syn.T)
        self.step_count = 0
        self._bufs = {}
        self.refresh()

    # -- parameters -------------------------------------------------------
    def _alloc_state(self):
        self.grads = {n: torch.zeros_like(t) for n, t in self.params.items()}
        self.m1 = {n: torch.zeros_like(t) for n, t in self.params.items()}
        self.m2 = {n: torch.zeros_like(t) for n, t in self.params.items()}
        self.bf = {n: torch.empty(t.shape, dtype=torch.bfloat16, device=self.device)
                   for n, t in self.params.items() if not n.startswith(("b", "c", "Wr", "Wi"))}
        Km, Cw = self.cfg.modes, self.cfg.width
        self.blk = {l: torch.empty(Km, 2 * Cw, 2 * Cw, dtype=torch.bfloat16, device=self.device)
                    for l in range(1, self.cfg.layers + 1)}
        # the transposed blocks: operand of the forward mode mixing (a product
        # with transposed B stages its tiles with 16-byte loads)
        self.blk_t = {l: torch.empty(Km, 2 * Cw, 2 * Cw, dtype=torch.bfloat16, device=self.device)
                      for l in range(1, self.cfg.layers + 1)}
        self.gblk = {l: torch.zeros(Km, 2 * Cw, 2 * Cw, dtype=torch.float32, device=self.device)
                     for l in range(1, self.cfg.layers + 1)}

    def _hi_lo(self, a):
        """bf16 [2][...]: hi = bf16(a), lo = bf16(a - hi), so hi + lo carries ~16 bits."""
        t = torch.tensor(a, dtype=torch.float32)
        hi = t.to(torch.bfloat16)
        lo = (t - hi.float()).to(torch.bfloat16)
        return torch.stack([hi, lo]).contiguous().to(self.device)

    def refresh(self):
        """Refresh the BF16 copies from the FP32 master weights."""
        for n, t in self.bf.items():
            k.cast_bf16(self.params[n], t)
        self._fused = None  # output bias of the fused projection, read again on next use
        self._lift = None   # tables of the structured first layer, rebuilt on next use
        self._wt = None     # transposed weights of the fused kernels, rebuilt on next use
        self._expand()

    def _expand(self):
        """Real blocks of the spectral weights (and their transposes) from the
        FP32 masters."""
        for l in range(1, self.cfg.layers + 1):
            wr, wi = self.params[f"Wr{l}"], self.params[f"Wi{l}"]
            k.spec_expand(wr, wi, self.blk[l], self.cfg.modes, self.cfg.width)
            k.spec_expand(wr, wi, self.blk_t[l], self.cfg.modes, self.cfg.width, transposed=True)

    def load_state_dict(self, d):
        for n in self.names:
            self.params[n].copy_(d[n].to(self.device, torch.float32))
        self.refresh()

    def n_params(self):
        return sum(t.numel() for t in self.params.values())

    # -- buffers ----------------------------------------------------------
    def _buf(self, name, shape, dtype):
        t = self._bufs.get(name)
        n = int(np.prod(shape))
        if t is None or t.numel() < n or t.dtype != dtype:
            t = torch.empty(n, dtype=dtype, device=self.device)
            self._bufs[name] = t
        return t[:n].view(*shape)

    def release(self):
        self._bufs.clear()

    # -- forward ----------------------------------------------------------
    def _transposed(self):
        """bf16 [out][in] copies of the lift, pointwise and first projection
        weights (the fused kernels' operand layout), rebuilt after a step."""
        if self._wt is None or self._wt[0] != self.step_count:
            wt = {}
            for n in self.bf:
                if n.startswith(("W0", "Wp", "P1")):
                    w = self.params[n]
                    wt[n] = torch.empty(w.shape[1], w.shape[0], dtype=torch.bfloat16, device=self.device)
                    k.cast_bf16_t(w, wt[n])
            self._wt = (self.step_count, wt)
        return self._wt[1]

    def forward(self, x, train: bool = False):
        """x: bf16 [S, Z, F] -> prediction f32 [S*Z] (normalised TL).  With
        train=True the values the backward pass needs are kept; networks of
        the fused shape then run their lift and layers through the fused
        kernels (which also store the pre-activations), the others and
        inference through the layer-by-layer kernels."""
        cfg = self.cfg
        S, Z, F = x.shape
        if Z != cfg.depth_points or F != cfg.features or x.dtype != torch.bfloat16:
            raise ValueError(f"input must be bf16 [S, {cfg.depth_points}, {cfg.features}]")
        x = x.contiguous()
        Cw, Km, L = cfg.width, cfg.modes, cfg.layers
        R = S * Z
        bf, f32 = torch.bfloat16, torch.float32
        self._S = S
        self._x = x
        if train and self.fusable:
            return self._forward_train_fused(x, S)
        h = [self._buf("h0", (R, Cw), bf)]
        k.gemm(x, self.bf["W0"], h[0], R, Cw, F, bias=self.params["b0"], epi=k.EPI_BF16)
        y = self._buf("y", (R, Cw), f32)
        mh = self._buf("mh", (Km, S, 2 * Cw), bf)
        for l in range(1, L + 1):
            xh = self._buf(f"xh{l}" if train else "xh", (Km, S, 2 * Cw), bf)
            k.spec_analysis(h[-1], xh, self.ac, self.as_, S, Z, Cw, Km)
            k.gemm(xh, self.blk_t[l], mh, S, 2 * Cw, 2 * Cw, tb=True, batch=Km, sa=S * 2 * Cw, sb=4 * Cw * Cw,
                   sc=S * 2 * Cw, epi=k.EPI_BF16)
            k.spec_synth(mh, y, self.sc, self.ss, S, Z, Cw, Km)
            hn = self._buf(f"h{l}" if train else f"h{l % 2}", (R, Cw), bf)
            pre = self._buf(f"pre{l}", (R, Cw), f32) if train else None
            k.gemm(h[-1], self.bf[f"Wp{l}"], hn, R, Cw, Cw, bias=self.params[f"bp{l}"], addend=y,
                   pre=pre, epi=k.EPI_GELU_BF16)
            h.append(hn)
        q = self._buf("q", (R, Cw), bf)
        preq = self._buf("preq", (R, Cw), f32) if train else None
        k.gemm(h[-1], self.bf["P1"], q, R, Cw, Cw, bias=self.params["c1"], pre=preq, epi=k.EPI_GELU_BF16)
        pred = self._buf("pred", (R,), f32)
        k.gemm(q, self.bf["P2"], pred, R, 1, Cw, bias=self.params["c2"], epi=k.EPI_F32)
        self._h = h
        return pred

    def _forward_train_fused(self, x, S):
        """The training forward pass on the fused kernels: per layer one
        mode-mixing product and one fused layer (pointwise product, spectral
        synthesis, bias, GELU with its FP32 pre-activation stored, the next
        layer's analysis); the projection layer by layer (its GELU
        pre-activation and output are kept too)."""
        cfg = self.cfg
        Z, Cw, Km, L = cfg.depth_points, cfg.width, cfg.modes, cfg.layers
        R = S * Z
        bf, f32 = torch.bfloat16, torch.float32
        wt = self._transposed()
        h = [self._buf("h0", (R, Cw), bf)]
        xh = self._buf("xh1", (Km, S, 2 * Cw), bf)
        k.fno_layer(x, wt["W0"], self.params["b0"], h[0], S, ban=self.ban, xh=xh, act=False)
        mh = self._buf("mh", (Km, S, 2 * Cw), bf)
        for l in range(1, L + 1):
            k.gemm(xh, self.blk_t[l], mh, S, 2 * Cw, 2 * Cw, tb=True, batch=Km, sa=S * 2 * Cw, sb=4 * Cw * Cw,
                   sc=S * 2 * Cw, epi=k.EPI_BF16)
            last = l == L
            xh = None if last else self._buf(f"xh{l + 1}", (Km, S, 2 * Cw), bf)
            hn = self._buf(f"h{l}", (R, Cw), bf)
            k.fno_layer(h[-1], wt[f"Wp{l}"], self.params[f"bp{l}"], hn, S, mh=mh, bsyn=self.bsyn,
                        ban=None if last else self.ban, xh=xh, act=True, pre=self._buf(f"pre{l}", (R, Cw), f32))
            h.append(hn)
        q = self._buf("q", (R, Cw), bf)
        preq = self._buf("preq", (R, Cw), f32)
        k.gemm(h[-1], self.bf["P1"], q, R, Cw, Cw, bias=self.params["c1"], pre=preq, epi=k.EPI_GELU_BF16)
        pred = self._buf("pred", (R,), f32)
        k.gemm(q, self.bf["P2"], pred, R, 1, Cw, bias=self.params["c2"], epi=k.EPI_F32)
        self._h = h
        return pred

    def forward_fused(self, x):
        """Inference with the fused kernels: per layer one mode-mixing product
        and one fused layer kernel (pointwise product, spectral synthesis,
        bias, GELU, next layer's analysis), then the fused projection.
        x: bf16 [S, Z, F] -> prediction f32 [S*Z] (normalised TL)."""
        cfg = self.cfg
        if not self.fusable:
            raise ValueError("the fused kernels are built for Z=64, width=128, modes=16, features=128")
        S, Z, F = x.shape
        if Z != cfg.depth_points or F != cfg.features or x.dtype != torch.bfloat16:
            raise ValueError(f"input must be bf16 [S, {cfg.depth_points}, {cfg.features}]")
        x = x.contiguous()
        if self._fused is None or self._fused[0] != self.step_count:
            # the output bias as a host float
            self._fused = (self.step_count, float(self.params["c2"].item()))
        wt, c2 = self._transposed(), self._fused[1]
        Cw, Km = cfg.width, cfg.modes
        R = S * Z
        bf = torch.bfloat16
        h = [self._buf("fh0", (R, Cw), bf), self._buf("fh1", (R, Cw), bf)]
        xh = self._buf("fxh", (Km, S, 2 * Cw), bf)
        k.fno_layer(x, wt["W0"], self.params["b0"], h[0], S, ban=self.ban, xh=xh, act=False)
        return self._fused_layers(S, h, xh, wt, c2)

    def lift_tables(self, inp):
        """(la, lr, lzt, lu, lv) of `kernels.fno_lift` for `inp`
        (`encode.LiftInputs`): the first layer is linear in the encoding's
        columns, so its output is the per-query columns' share (with the
        bias), the range bin's, the depth point's, the local water depth's and
        the below-seabed flag's; the weights are the bf16 ones the product
        with the encoded input uses.  All but la depend only on the weights
        and the domain's tables and are kept until either changes."""
        from .encode import F_BELOW, F_HEIGHT, F_ZB, F_ZREL, N_GLOBAL
        c = self._lift
        if c is None or c[0] != self.step_count or c[1] is not inp.fr or c[2] is not inp.fz:
            W = self.bf["W0"].float()                                 # [F, C]
            wg = W[:N_GLOBAL].clone()
            wg[N_GLOBAL - 1] -= W[F_ZREL]    # the offset column holds z - zs
            lr = (inp.fr @ W[inp.fr_rows]).contiguous()
            lzt = (inp.fz @ W[inp.fz_rows]).t().contiguous()
            lu = (W[F_ZB] - W[F_HEIGHT]).contiguous()    # the height column holds z - zb
            lv = W[F_BELOW].contiguous()
            c = self._lift = (self.step_count, inp.fr, inp.fz, wg, lr, lzt, lu, lv)
        la = torch.addmm(self.params["b0"], inp.g, c[3])
        return (la,) + c[4:8]

    def forward_fused_lift(self, inp):
        """`forward_fused` of the encoded queries without the encoded input
        and without the first layer's output: the first layer's modes from
        its tables (`kernels.fno_lift_modes`), the second layer's pointwise
        product from the same tables times its weight
        (`kernels.fno_layer_lift`), the rest as `forward_fused`.  Equal to it
        up to the bf16 rounding of the encoded columns and of the first
        layer's output, which this form skips.  A one-layer network takes the
        first layer's output from `kernels.fno_lift` instead."""
        cfg = self.cfg
        if not self.fusable:
            raise ValueError("the fused kernels are built for Z=64, width=128, modes=16, features=128")
        n, nr = inp.g.shape[0], inp.nr
        S = n * nr
        if inp.zv.numel() != cfg.depth_points:
            raise ValueError(f"the lift inputs must have {cfg.depth_points} depth points")
        if self._fused is None or self._fused[0] != self.step_count:
            self._fused = (self.step_count, float(self.params["c2"].item()))
        wt, c2 = self._transposed(), self._fused[1]
        Cw, Km = cfg.width, cfg.modes
        R = S * cfg.depth_points
        bf = torch.bfloat16
        h = [self._buf("fh0", (R, Cw), bf), self._buf("fh1", (R, Cw), bf)]
        xh = self._buf("fxh", (Km, S, 2 * Cw), bf)
        la, lr, lzt, lu, lv = self.lift_tables(inp)
        if cfg.layers < 2:
            k.fno_lift(la, lr, lzt, inp.zb, inp.zv, lu, lv, nr, self.ban, h[0], xh, S)
            return self._fused_layers(S, h, xh, wt, c2)
        tsum, bz, lr1, lzt1, lu1, lv1, w1 = self._lift_second_tables()
        k.fno_lift_modes(la, lr, inp.zb, inp.zv, lu, lv, tsum, bz, nr, xh, S)
        mh = self._buf("fmh", (Km, S, 2 * Cw), bf)
        k.gemm(xh, self.blk_t[1], mh, S, 2 * Cw, 2 * Cw, tb=True, batch=Km, sa=S * 2 * Cw, sb=4 * Cw * Cw,
               sc=S * 2 * Cw, epi=k.EPI_BF16)
        la1 = la @ w1
        k.fno_layer_lift(la1, lr1, lzt1, inp.zb, inp.zv, lu1, lv1, nr, self.params["bp1"], mh, self.bsyn, self.ban,
                         h[1], xh, S)
        return self._fused_layers(S, h, xh, wt, c2, first=2)

    def _lift_second_tables(self):
        """Tables of the first layer's modes and of the second layer's
        pointwise product in the table form, from the cached first-layer
        tables (call after `lift_tables`): the analysis rows' suffix sums over
        depth [2K][Z+1], the analysis of the depth features' share [2K][C],
        and the range, depth, water-depth and flag tables times the second
        layer's bf16 weight, with that weight [C][C] (for la)."""
        c = self._lift
        if len(c) == 8:
            lr, lzt, lu, lv = c[4:8]
            w1 = self.bf["Wp1"].float()                                # [C in, C out]
            an = self.ana                                             # [2K][Z]
            tsum = torch.zeros(an.shape[0], an.shape[1] + 1, device=self.device)
            tsum[:, :-1] = torch.flip(torch.cumsum(torch.flip(an, [1]), 1), [1])
            bz = an @ lzt.t()
            extra = (tsum.contiguous(), bz.contiguous(), (lr @ w1).contiguous(), (lzt.t() @ w1).t().contiguous(),
                     (lu @ w1).contiguous(), (lv @ w1).contiguous(), w1)
            c = self._lift = c + extra
        return c[8:]

    def _fused_layers(self, S, h, xh, wt, c2, first=1):
        """Layers first .. L and the projection of the fused inference, from
        the output of layer first - 1 in h[(first - 1) % 2] and its modes xh."""
        cfg = self.cfg
        Cw, Km, L = cfg.width, cfg.modes, cfg.layers
        R = S * cfg.depth_points
        bf = torch.bfloat16
        mh = self._buf("fmh", (Km, S, 2 * Cw), bf)
        pred = self._buf("fpred", (R,), torch.float32)
        for l in range(first, L + 1):
            k.gemm(xh, self.blk_t[l], mh, S, 2 * Cw, 2 * Cw, tb=True, batch=Km, sa=S * 2 * Cw, sb=4 * Cw * Cw,
                   sc=S * 2 * Cw, epi=k.EPI_BF16)
            if l < L:
                k.fno_layer(h[(l - 1) % 2], wt[f"Wp{l}"], self.params[f"bp{l}"], h[l % 2], S, mh=mh,
                            bsyn=self.bsyn, ban=self.ban, xh=xh, act=True)
            else:
                # the last layer's output goes straight into the projection
                k.fno_layer_project(h[(l - 1) % 2], wt[f"Wp{l}"], self.params[f"bp{l}"], mh, self.bsyn,
                                    wt["P1"], self.params["c1"], self.params["P2"].reshape(-1), c2, pred, S)
        return pred

    # -- loss and backward -----------------------------------------------
    def loss_grad(self, pred, target, mask):
        """Masked MSE; returns (device loss_sum[2], dpred bf16)."""
        R = pred.numel()
        cnt = float(mask.sum().item())
        loss = self._buf("loss", (2,), torch.float32)
        k.fill(loss)
        dpred = self._buf("dpred", (R,), torch.bfloat16)
        k.mse(pred, 1, target.reshape(-1), mask.reshape(-1), dpred, R, 2.0 / max(cnt, 1.0), loss)
        return loss, dpred

    def backward(self, dpred, fused=None):
        """Accumulate parameter gradients of the last train-mode forward;
        through the fused backward layers when the network's shape has them
        (fused=None or True), else (or with fused=False) layer by layer."""
        if fused and not self.fusable:
            raise ValueError("the fused backward kernels are built for Z=64, width=128, modes=16, features=128")
        cfg = self.cfg
        S, Z, F, Cw, Km, L = self._S, cfg.depth_points, cfg.features, cfg.width, cfg.modes, cfg.layers
        R = S * Z
        bf, f32 = torch.bfloat16, torch.float32
        g = self.grads
        for t in g.values():
            k.fill(t)
        for l in range(1, L + 1):
            k.fill(self.gblk[l])
        h = self._h
        q = self._buf("q", (R, Cw), bf)
        # y = q P2 + c2
        k.gemm(q, dpred, g["P2"], Cw, 1, R, ta=True, epi=k.EPI_ATOMIC, splits=k.splits_for(Cw, 1, R))
        k.colsum(dpred, g["c2"], R, 1)
        if self.fusable if fused is None else fused:
            return self._backward_fused(dpred, S)
        dh = self._buf("dh", (R, Cw), f32)
        k.gemm(dpred, self.bf["P2"], dh, R, Cw, 1, tb=True, epi=k.EPI_F32)
        dpre = self._buf("dpre", (R, Cw), bf)
        k.gelu_bwd(dh, self._buf("preq", (R, Cw), f32), dpre)
        # q = GELU(h_L P1 + c1)
        k.gemm(h[L], dpre, g["P1"], Cw, Cw, R, ta=True, epi=k.EPI_ATOMIC, splits=k.splits_for(Cw, Cw, R))
        k.colsum(dpre, g["c1"], R, Cw)
        k.gemm(dpre, self.bf["P1"], dh, R, Cw, Cw, tb=True, epi=k.EPI_F32)
        dmh = self._buf("dmh", (Km, S, 2 * Cw), bf)
        dxh = self._buf("dxh", (Km, S, 2 * Cw), bf)
        for l in range(L, 0, -1):
            k.gelu_bwd(dh, self._buf(f"pre{l}", (R, Cw), f32), dpre)
            # pointwise branch
            k.gemm(h[l - 1], dpre, g[f"Wp{l}"], Cw, Cw, R, ta=True, epi=k.EPI_ATOMIC,
                   splits=k.splits_for(Cw, Cw, R))
            k.colsum(dpre, g[f"bp{l}"], R, Cw)
            k.gemm(dpre, self.bf[f"Wp{l}"], dh, R, Cw, Cw, tb=True, epi=k.EPI_F32)
            # spectral branch: synthesis^T, mode mixing, analysis^T
            k.spec_analysis(dpre, dmh, self.sc, self.ss, S, Z, Cw, Km)
            xh = self._buf(f"xh{l}", (Km, S, 2 * Cw), bf)
            k.gemm(xh, dmh, self.gblk[l], 2 * Cw, 2 * Cw, S, ta=True, batch=Km, sa=S * 2 * Cw, sb=S * 2 * Cw,
                   sc=4 * Cw * Cw, epi=k.EPI_ATOMIC, splits=k.splits_for(2 * Cw, 2 * Cw, S, batch=Km))
            k.gemm(dmh, self.blk[l], dxh, S, 2 * Cw, 2 * Cw, tb=True, batch=Km, sa=S * 2 * Cw, sb=4 * Cw * Cw,
                   sc=S * 2 * Cw, epi=k.EPI_BF16)
            k.spec_synth(dxh, dh, self.ac, self.as_, S, Z, Cw, Km, accumulate=True)
            k.spec_fold(self.gblk[l], g[f"Wr{l}"], g[f"Wi{l}"], Km, Cw)
        # lift
        k.gelu_bwd(dh, None, dpre)
        k.gemm(self._x, dpre, g["W0"], F, Cw, R, ta=True, epi=k.EPI_ATOMIC, splits=k.splits_for(F, Cw, R))
        k.colsum(dpre, g["b0"], R, Cw)

    def _backward_fused(self, dpred, S):
        """The rest of the backward pass on the fused kernels: the projection's
        GELU gradient in one pass, then per layer the weight-gradient products,
        the mode-mixing gradient and one fused backward layer (pointwise
        product with Wp^T, transposed analysis, GELU' of the layer below, bias
        gradient partials, transposed synthesis for the layer below)."""
        cfg = self.cfg
        Z, Cw, Km, L = cfg.depth_points, cfg.width, cfg.modes, cfg.layers
        R = S * Z
        bf, f32 = torch.bfloat16, torch.float32
        g, h = self.grads, self._h
        part = self._buf("csum", (S, Cw), f32)
        dpre = [self._buf("dpre", (R, Cw), bf), self._buf("dpre2", (R, Cw), bf)]
        dmh = self._buf("dmh", (Km, S, 2 * Cw), bf)
        dxh = self._buf("dxh", (Km, S, 2 * Cw), bf)
        # q = GELU(h_L P1 + c1), y = q P2 + c2
        k.outer_gelu_bwd(dpred, self.params["P2"].reshape(-1), self._buf("preq", (R, Cw), f32), dpre[0], part,
                         S, Z, Cw)
        k.colsum(part, g["c1"], S, Cw)
        k.gemm(h[L], dpre[0], g["P1"], Cw, Cw, R, ta=True, epi=k.EPI_ATOMIC, splits=k.splits_for(Cw, Cw, R))
        k.fno_layer_bwd(dpre[0], self.bf["P1"], dpre[1], part, S, ban=self.ban_t, dmh=dmh,
                        pin=self._buf(f"pre{L}", (R, Cw), f32))
        cur = 1
        for l in range(L, 0, -1):
            d = dpre[cur]
            k.colsum(part, g[f"bp{l}"], S, Cw)
            k.gemm(h[l - 1], d, g[f"Wp{l}"], Cw, Cw, R, ta=True, epi=k.EPI_ATOMIC, splits=k.splits_for(Cw, Cw, R))
            xh = self._buf(f"xh{l}", (Km, S, 2 * Cw), bf)
            k.gemm(xh, dmh, self.gblk[l], 2 * Cw, 2 * Cw, S, ta=True, batch=Km, sa=S * 2 * Cw, sb=S * 2 * Cw,
                   sc=4 * Cw * Cw, epi=k.EPI_ATOMIC, splits=k.splits_for(2 * Cw, 2 * Cw, S, batch=Km))
            k.gemm(dmh, self.blk[l], dxh, S, 2 * Cw, 2 * Cw, tb=True, batch=Km, sa=S * 2 * Cw, sb=4 * Cw * Cw,
                   sc=S * 2 * Cw, epi=k.EPI_BF16)
            below = l > 1
            k.fno_layer_bwd(d, self.bf[f"Wp{l}"], dpre[1 - cur], part, S, dxh=dxh, bsyn=self.bsyn_t,
                            ban=self.ban_t if below else None, dmh=dmh if below else None,
                            pin=self._buf(f"pre{l - 1}", (R, Cw), f32) if below else None)
            cur = 1 - cur
            k.spec_fold(self.gblk[l], g[f"Wr{l}"], g[f"Wi{l}"], Km, Cw)
        # lift (no activation)
        k.gemm(self._x, dpre[cur], g["W0"], cfg.features, Cw, R, ta=True, epi=k.EPI_ATOMIC,
               splits=k.splits_for(cfg.features, Cw, R))
        k.colsum(part, g["b0"], S, Cw)

    def step(self, lr, wd=0.0, b1=0.9, b2=0.999, eps=1e-8):
        """One Adam step on every parameter from the accumulated gradients."""
        self.step_count += 1
        for n in self.names:
            k.adam(self.params[n], self.grads[n], self.m1[n], self.m2[n], self.bf.get(n), lr, self.step_count,
                   b1=b1, b2=b2, eps=eps, wd=0.0 if n.startswith(("b", "c")) else wd)
        self._expand()


def infer(net, x):
    """Inference through the fused kernels when the network's shape has them,
    else the layer-by-layer forward pass."""
    if getattr(net, "fusable", False):
        return net.forward_fused(x)
    return net.forward(x, train=False)
