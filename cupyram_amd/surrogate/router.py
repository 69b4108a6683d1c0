"""Surrogate-first answering of TL queries with PE fallback.

Every query is first answered by the mean of an ensemble of trained
networks.  It is routed to the complex64 PE march instead when

* the ensemble disagrees: the RMS over the output grid of the members'
  standard deviation exceeds `threshold_db`, or
* it is drawn for the spot check (a seeded `spot_fraction` of the queries
  the ensemble accepted) and the PE march finds the ensemble mean more than
  `threshold_db` RMS away from its own field.

The PE march judges the surrogate: the spot-checked queries give the
surrogate's measured error, and when their RMS error exceeds the threshold
the whole batch falls back to the PE march.
"""
from __future__ import annotations

import dataclasses
import math
import time

import numpy as np
import torch

from .encode import Domain, Queries, denormalise_tl, encode
from .fno import infer
from .train import field_rms, pe_fields


@dataclasses.dataclass
class RouterResult:
    tl: torch.Tensor            # [n, nr, nzp] dB, the answer for every query
    routed: np.ndarray          # bool [n], answered by the PE march
    spread_db: np.ndarray       # ensemble disagreement per query
    spot: np.ndarray            # indices of the spot-checked queries
    spot_err_db: np.ndarray     # surrogate error on them, judged by the PE march
    batch_fallback: bool
    seconds: dict

    @property
    def frac_routed(self):
        return float(np.mean(self.routed))


class Router:
    def __init__(self, dom: Domain, members, threshold_db=3.0, spot_fraction=0.01, seed=0, precision="c64",
                 batch=128, log=None, pe=pe_fields, device="cuda"):
        """members: trained networks (anything with forward(x, train=False)
        returning normalised TL [n*nr*nzp]; networks of the fused shape are
        run through forward_fused); pe(dom, queries, precision,
        log=...) -> (TL [n, nr, nzp] dB, seconds, cell-steps) answers the
        routed queries (the complex64 march by default)."""
        if not members:
            raise ValueError("the router needs at least one trained network")
        if not 0.0 <= spot_fraction <= 1.0:
            raise ValueError("spot_fraction must lie in [0, 1]")
        self.dom, self.members = dom, list(members)
        self.threshold, self.spot_fraction, self.seed = float(threshold_db), float(spot_fraction), seed
        self.precision, self.batch, self.log, self.pe = precision, batch, log, pe
        self.device = torch.device(device)

    def _sync(self):
        if self.device.type == "cuda":
            torch.cuda.synchronize()

    def ensemble(self, q: Queries):
        """Ensemble mean TL [n, nr, nzp] and disagreement [n] (dB)."""
        dom, n = self.dom, len(q)
        dev = self.device
        mean = torch.empty(n, dom.nr, dom.nzp, device=dev)
        spread = torch.empty(n, device=dev)
        feats = self.members[0].cfg.features
        for a in range(0, n, self.batch):
            idx = np.arange(a, min(n, a + self.batch))
            x = encode(dom, q.take(idx), feats, device=dev)
            s1 = torch.zeros(len(idx), dom.nr, dom.nzp, device=dev)
            s2 = torch.zeros_like(s1)
            for m in self.members:
                y = denormalise_tl(dom, infer(m, x).view(len(idx), dom.nr, dom.nzp))
                y = y.clamp(dom.tl_lo, dom.tl_hi)
                s1 += y
                s2 += y * y
            mu = s1 / len(self.members)
            var = (s2 / len(self.members) - mu * mu).clamp_min(0.0)
            mean[a:a + len(idx)] = mu
            spread[a:a + len(idx)] = torch.sqrt(var.mean(dim=(1, 2)))
        return mean, spread

    def run(self, q: Queries) -> RouterResult:
        n = len(q)
        t0 = time.perf_counter()
        mean, spread = self.ensemble(q)
        self._sync()
        t_sur = time.perf_counter() - t0
        spread_h = spread.cpu().numpy()
        routed = spread_h > self.threshold
        accepted = np.nonzero(~routed)[0]
        rng = np.random.default_rng(self.seed)
        n_spot = min(len(accepted), int(math.ceil(self.spot_fraction * len(accepted)))) if len(accepted) else 0
        spot = np.sort(rng.choice(accepted, size=n_spot, replace=False)) if n_spot else np.zeros(0, np.int64)
        pe_idx = np.union1d(np.nonzero(routed)[0], spot)
        t1 = time.perf_counter()
        tl = mean
        spot_err = np.zeros(0)
        batch_fallback = False
        if len(pe_idx):
            pe_tl, _, _ = self.pe(self.dom, q.take(pe_idx), self.precision, log=self.log)
            pos = np.searchsorted(pe_idx, spot)
            if len(spot):
                spot_err = field_rms(self.dom, mean[torch.as_tensor(spot, device=self.device)],
                                     pe_tl[torch.as_tensor(pos, device=self.device)]).cpu().numpy()
            tl[torch.as_tensor(pe_idx, device=self.device)] = pe_tl
            routed[spot[spot_err > self.threshold]] = True
            batch_fallback = bool(len(spot_err) and np.sqrt(np.mean(spot_err ** 2)) > self.threshold)
        if batch_fallback:
            rest = np.nonzero(~np.isin(np.arange(n), pe_idx))[0]
            if len(rest):
                rest_tl, _, _ = self.pe(self.dom, q.take(rest), self.precision, log=self.log)
                tl[torch.as_tensor(rest, device=self.device)] = rest_tl
            routed[:] = True
        self._sync()
        t_pe = time.perf_counter() - t1
        return RouterResult(tl=tl, routed=routed, spread_db=spread_h, spot=spot, spot_err_db=spot_err,
                            batch_fallback=batch_fallback,
                            seconds={"surrogate": t_sur, "pe": t_pe, "total": time.perf_counter() - t0})
