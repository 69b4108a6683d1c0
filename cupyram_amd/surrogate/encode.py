"""Query parameters of the surrogate and their encoding as network input.

A query is the 101-number description of one TL field:
    ssp     [32]  sound speed (m/s) at 32 evenly spaced depths over [0, ssp_depth]
    bathy   [64]  water depth (m) at 64 evenly spaced ranges over [0, rmax]
    bottom  [3]   three bottom parameters of the domain (see Domain.bottom_names)
    freq          source frequency (Hz)
    zs            source depth (m)
and the answer is the coarse TL (dB) on the output grid of the domain: nr
range bins ending at r_j = (j+1) rmax/nr and nzp depth bins ending at
z_i = (i+1) zplot/nzp, each cell holding -10 log10 of the mean of |p|^2/r
over the march's fields and rows inside it (the 'TL Cells' of a march whose
ndr, ndz make its output grid exactly these bins).  Point samples of the
coherent field (the TL Grid) fluctuate like a Rayleigh amplitude from one
point to the next, about 5.6 dB in standard deviation, which no smooth
predictor follows; the cell means keep the propagation physics (mode
stripping, convergence zones, bottom loss) without that interference noise.

The encoding gives every (range bin, depth point) the 101 normalised query
numbers plus 7 local ones (range, depth, local water depth, height above
the seabed, offset from the source depth, a below-seabed flag, a constant)
and 17 Fourier features of position (log range; sin and cos of pi k r/rmax
and of pi k z/zplot for k = 1, 2, 4, 8), zero-padded to `features` columns;
one network sample is one range bin.  The Fourier features let the first
(linear) layer express structure along range and depth at several scales.
"""
from __future__ import annotations

import dataclasses

import numpy as np
import torch

N_SSP, N_BATHY, N_GLOBAL, N_LOCAL = 32, 64, 101, 7
FOURIER_K = (1, 2, 4, 8)
N_FOURIER = 1 + 4 * len(FOURIER_K)
ENCODING = "v2"   # recorded with trained weights; v2 added the Fourier features


@dataclasses.dataclass(frozen=True)
class Domain:
    name: str
    rmax: float              # range of the last output bin (m)
    zplot: float             # depth of the last output point (m)
    ssp_depth: float         # depth of the last sound-speed knot (m)
    bottom_mean: tuple       # centring of the 3 bottom parameters
    bottom_scale: tuple
    bottom_names: tuple
    freq_lo: float
    freq_hi: float
    nr: int = 200
    nzp: int = 64
    c_ref: float = 1500.0
    c_scale: float = 25.0
    tl_lo: float = 20.0      # TL targets are clipped to [tl_lo, tl_hi] dB
    tl_hi: float = 140.0
    tl_mean: float = 80.0    # normalised target = (TL - tl_mean) / tl_std
    tl_std: float = 20.0
    target: str = "cells"    # "cells" (cell-mean TL) or "points" (TL Grid)

    def ranges(self):
        return (np.arange(self.nr) + 1.0) * (self.rmax / self.nr)

    def depths(self):
        return (np.arange(self.nzp) + 1.0) * (self.zplot / self.nzp)


@dataclasses.dataclass
class Queries:
    ssp: np.ndarray      # [n, 32]
    bathy: np.ndarray    # [n, 64]
    bottom: np.ndarray   # [n, 3]
    freq: np.ndarray     # [n]
    zs: np.ndarray       # [n]

    def __len__(self):
        return len(self.freq)

    def take(self, idx):
        return Queries(self.ssp[idx], self.bathy[idx], self.bottom[idx], self.freq[idx], self.zs[idx])

    def vector(self):
        """The 101 query numbers, [n, 101]."""
        return np.concatenate([self.ssp, self.bathy, self.bottom, self.freq[:, None], self.zs[:, None]], axis=1)


def global_features(dom: Domain, q: Queries, device="cuda"):
    """The 101 normalised query numbers, float32 [n, 101]."""
    dev = torch.device(device)
    t = lambda a: torch.as_tensor(np.asarray(a, np.float32), device=dev)  # noqa: E731
    return torch.cat([
        (t(q.ssp) - dom.c_ref) / dom.c_scale,
        t(q.bathy) / dom.zplot,
        (t(q.bottom) - t(dom.bottom_mean)) / t(dom.bottom_scale),
        ((torch.log(t(q.freq)) - np.log(dom.freq_lo)) / np.log(dom.freq_hi / dom.freq_lo))[:, None],
        (t(q.zs) / dom.zplot)[:, None]], dim=1)


def encode(dom: Domain, q: Queries, features: int = 128, device="cuda"):
    """Network input, bf16 [n*nr, nzp, features].

    Each feature is computed in float32 at its natural size (per query, per
    range bin, per depth point) and rounded to bf16 once; one concatenation
    of their broadcasts writes the bf16 tensor (about 8x less memory traffic
    than filling a float32 tensor slice by slice and casting it)."""
    n = len(q)
    if features < N_GLOBAL + N_LOCAL + N_FOURIER:
        raise ValueError("features too small")
    dev = torch.device(device)
    t = lambda a: torch.as_tensor(np.asarray(a, np.float32), device=dev)  # noqa: E731
    bf = torch.bfloat16
    nr, nz = dom.nr, dom.nzp
    shape = (n, nr, nz)
    g = global_features(dom, q, device)                               # [n, 101]
    r = t(dom.ranges() / dom.rmax)                                    # [nr]
    z = t(dom.depths() / dom.zplot)                                   # [nzp]
    # water depth at each output range, linear between the 64 samples
    pos = r * (N_BATHY - 1)
    i0 = torch.clamp(pos.floor().long(), 0, N_BATHY - 2)
    w = (pos - i0.float())[None, :]
    bz = g[:, N_SSP:N_SSP + N_BATHY]
    zb = bz[:, i0] * (1 - w) + bz[:, i0 + 1] * w                      # [n, nr] (in zplot units)
    zsn = g[:, N_GLOBAL - 1]
    zq = z[None, None, :]
    local = torch.stack([
        r[None, :, None].expand(shape),
        zq.expand(shape),
        zb[:, :, None].expand(shape),
        zq - zb[:, :, None],
        (zq - zsn[:, None, None]).expand(shape),
        (zq > zb[:, :, None]).float(),
        torch.ones(shape, device=dev)], dim=-1).to(bf)               # [n, nr, nzp, 7]
    four_r = [torch.log(r)]
    four_z = []
    for kf in FOURIER_K:
        four_r += [torch.sin(np.pi * kf * r), torch.cos(np.pi * kf * r)]
        four_z += [torch.sin(np.pi * kf * z), torch.cos(np.pi * kf * z)]
    fr = torch.stack(four_r, dim=-1).to(bf)                           # [nr, 9]: log r, (sin, cos) per k
    fz = torch.stack(four_z, dim=-1).to(bf)                           # [nzp, 8]
    # interleave as log r, then per k: sin r, cos r, sin z, cos z
    parts = [g.to(bf)[:, None, None, :].expand(n, nr, nz, N_GLOBAL), local,
             fr[None, :, None, :1].expand(n, nr, nz, 1)]
    for i in range(len(FOURIER_K)):
        parts.append(fr[None, :, None, 1 + 2 * i:3 + 2 * i].expand(n, nr, nz, 2))
        parts.append(fz[None, None, :, 2 * i:2 * i + 2].expand(n, nr, nz, 2))
    pad = features - (N_GLOBAL + N_LOCAL + N_FOURIER)
    if pad:
        parts.append(torch.zeros(1, 1, 1, pad, dtype=bf, device=dev).expand(n, nr, nz, pad))
    return torch.cat(parts, dim=-1).reshape(n * nr, nz, features)


# Rows of the first layer's weight for each column of `encode`: the 101
# query numbers, then the local features, then the Fourier features
F_RANGE, F_DEPTH, F_ZB, F_HEIGHT, F_ZREL, F_BELOW, F_ONE = range(N_GLOBAL, N_GLOBAL + N_LOCAL)
F_LOGR = N_GLOBAL + N_LOCAL
F_SIN_R = [F_LOGR + 1 + 4 * i for i in range(len(FOURIER_K))]
F_COS_R = [f + 1 for f in F_SIN_R]
F_SIN_Z = [f + 2 for f in F_SIN_R]
F_COS_Z = [f + 3 for f in F_SIN_R]


@dataclasses.dataclass
class LiftInputs:
    """What the first layer needs of `encode`'s columns, at their natural
    sizes (float32): the first layer of the network is linear in them, so it
    is formed from per-query, per-range-bin and per-depth-point tables
    (`FNO.forward_fused_lift`) instead of a product with the encoded input.

    g       [n, 101]  the normalised query numbers
    zb      [n * nr]  local water depth of each sample (sample = query, range bin)
    zv      [nzp]     depth of each point (the depth column)
    fr      [nr, 10]  range, log range, and sin, cos of pi k r/rmax
    fr_rows           the first layer's rows of those columns
    fz      [nzp, 12] depth as it enters the depth, height-above-seabed and
                      offset-from-source columns, the constant, and sin, cos
                      of pi k z/zplot
    fz_rows           the first layer's rows of those columns
    The query's water depth enters the water-depth column and, negated, the
    height column; its source depth enters the offset column negated; the
    below-seabed flag is zv > zb."""
    g: torch.Tensor
    zb: torch.Tensor
    zv: torch.Tensor
    fr: torch.Tensor
    fr_rows: list
    fz: torch.Tensor
    fz_rows: list
    nr: int


_DOMAIN_TABLES = {}


def _domain_tables(dom: Domain, device):
    """(r, z, fr, fr_rows, fz, fz_rows) of a domain, built once per device."""
    key = (dom, str(device))
    if key in _DOMAIN_TABLES:
        return _DOMAIN_TABLES[key]
    dev = torch.device(device)
    t = lambda a: torch.as_tensor(np.asarray(a, np.float32), device=dev)  # noqa: E731
    r = t(dom.ranges() / dom.rmax)
    z = t(dom.depths() / dom.zplot)
    fr = [r, torch.log(r)]
    fr_rows = [F_RANGE, F_LOGR]
    fz = [z, z, z, torch.ones_like(z)]
    fz_rows = [F_DEPTH, F_HEIGHT, F_ZREL, F_ONE]
    for i, kf in enumerate(FOURIER_K):
        fr += [torch.sin(np.pi * kf * r), torch.cos(np.pi * kf * r)]
        fr_rows += [F_SIN_R[i], F_COS_R[i]]
        fz += [torch.sin(np.pi * kf * z), torch.cos(np.pi * kf * z)]
        fz_rows += [F_SIN_Z[i], F_COS_Z[i]]
    tables = (r, z.contiguous(), torch.stack(fr, dim=-1).contiguous(), fr_rows,
              torch.stack(fz, dim=-1).contiguous(), fz_rows)
    _DOMAIN_TABLES[key] = tables
    return tables


def lift_inputs(dom: Domain, q: Queries, device="cuda") -> LiftInputs:
    """The tables of `LiftInputs` for the queries `q`, computed exactly as
    `encode` computes the columns (before its bf16 rounding); the range and
    depth tables are the domain's, shared by every call."""
    r, z, fr, fr_rows, fz, fz_rows = _domain_tables(dom, device)
    g = global_features(dom, q, device)
    pos = r * (N_BATHY - 1)
    i0 = torch.clamp(pos.floor().long(), 0, N_BATHY - 2)
    w = (pos - i0.float())[None, :]
    bz = g[:, N_SSP:N_SSP + N_BATHY]
    zb = bz[:, i0] * (1 - w) + bz[:, i0 + 1] * w
    return LiftInputs(g=g.contiguous(), zb=zb.reshape(-1).contiguous(), zv=z, fr=fr, fr_rows=fr_rows, fz=fz,
                      fz_rows=fz_rows, nr=dom.nr)


def normalise_tl(dom: Domain, tl):
    """(target, mask) from TL [.., nr, nzp] (dB): clipped, centred, scaled;
    non-finite points are masked out."""
    tl = torch.as_tensor(tl, dtype=torch.float32)
    mask = torch.isfinite(tl).float()
    tl = torch.nan_to_num(tl, nan=dom.tl_hi, posinf=dom.tl_hi, neginf=dom.tl_lo)
    return (tl.clamp(dom.tl_lo, dom.tl_hi) - dom.tl_mean) / dom.tl_std, mask


def denormalise_tl(dom: Domain, y):
    return y * dom.tl_std + dom.tl_mean
