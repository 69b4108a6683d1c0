"""Surrogate-first answering of TL queries with PE fallback.

Every query is answered by the mean of an ensemble of trained networks,
unless

* the ensemble disagrees: the RMS over the output grid of the members'
  standard deviation exceeds `threshold_db`; the query goes to the complex64
  PE march; or
* it is drawn for the spot check (a seeded `spot_fraction` of the batch): the
  PE march answers it and judges the ensemble mean on it.

The spot check runs first.  When the RMS error of the ensemble on the spot
queries it would have answered exceeds the threshold, the whole batch falls
back to the PE march and the ensemble is not run on the rest.
"""
from __future__ import annotations

import dataclasses
import math
import time

import numpy as np
import torch

from .encode import Domain, Queries, denormalise_tl, encode, lift_inputs
from .fno import infer
from .train import field_rms, pe_fields


@dataclasses.dataclass
class RouterResult:
    tl: torch.Tensor            # [n, nr, nzp] dB, the answer for every query
    routed: np.ndarray          # bool [n], answered by the PE march
    spread_db: np.ndarray       # ensemble disagreement per query (NaN where a fallback skipped it)
    spot: np.ndarray            # indices of the spot-checked queries
    spot_err_db: np.ndarray     # surrogate error on those the ensemble accepted, judged by the PE march
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
        # members with the fused kernels form their first layer from the
        # encoding's tables (forward_fused_lift); the others read the encoded input
        lifted = [bool(getattr(m, "fusable", False)) for m in self.members]
        for a in range(0, n, self.batch):
            idx = np.arange(a, min(n, a + self.batch))
            sub = q.take(idx)
            inp = lift_inputs(dom, sub, device=dev) if any(lifted) else None
            x = encode(dom, sub, feats, device=dev) if not all(lifted) else None
            s1 = torch.zeros(len(idx), dom.nr, dom.nzp, device=dev)
            s2 = torch.zeros_like(s1)
            for m, lift in zip(self.members, lifted):
                p = m.forward_fused_lift(inp) if lift else infer(m, x)
                y = denormalise_tl(dom, p.view(len(idx), dom.nr, dom.nzp))
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
        dev = self.device
        t0 = time.perf_counter()
        t_sur = t_pe = 0.0
        tl = torch.empty(n, self.dom.nr, self.dom.nzp, device=dev)
        spread_h = np.full(n, np.nan)
        routed = np.zeros(n, bool)
        rng = np.random.default_rng(self.seed)
        n_spot = min(n, int(math.ceil(self.spot_fraction * n)))
        spot = np.sort(rng.choice(n, size=n_spot, replace=False)) if n_spot else np.zeros(0, np.int64)
        spot_err = np.zeros(0)
        batch_fallback = False
        if n_spot:
            # 1. spot check: ensemble and PE march on the sample
            t = time.perf_counter()
            mean_s, spread_s = self.ensemble(q.take(spot))
            self._sync()
            t_sur += time.perf_counter() - t
            t = time.perf_counter()
            pe_s, _, _ = self.pe(self.dom, q.take(spot), self.precision, log=self.log)
            self._sync()
            t_pe += time.perf_counter() - t
            spread_h[spot] = spread_s.cpu().numpy()
            acc = spread_h[spot] <= self.threshold
            err = field_rms(self.dom, mean_s, pe_s).cpu().numpy()
            spot_err = err[acc]
            tl[torch.as_tensor(spot, device=dev)] = pe_s
            # comparisons written as ~(x <= threshold) so that a NaN spread or
            # error (a member that returned NaN) counts as a failure
            routed[spot] = ~acc | ~(err <= self.threshold)
            batch_fallback = bool(len(spot_err) and not np.sqrt(np.mean(spot_err ** 2)) <= self.threshold)
        rest = np.setdiff1d(np.arange(n), spot)
        if len(rest):
            if batch_fallback:
                pe_idx = rest
            else:
                # 2. the ensemble answers the rest, except where it disagrees
                t = time.perf_counter()
                mean_r, spread_r = self.ensemble(q.take(rest))
                self._sync()
                t_sur += time.perf_counter() - t
                spread_h[rest] = spread_r.cpu().numpy()
                tl[torch.as_tensor(rest, device=dev)] = mean_r
                pe_idx = rest[~(spread_h[rest] <= self.threshold)]
            if len(pe_idx):
                t = time.perf_counter()
                pe_r, _, _ = self.pe(self.dom, q.take(pe_idx), self.precision, log=self.log)
                tl[torch.as_tensor(pe_idx, device=dev)] = pe_r
                self._sync()
                t_pe += time.perf_counter() - t
                routed[pe_idx] = True
        if batch_fallback:
            routed[:] = True
        return RouterResult(tl=tl, routed=routed, spread_db=spread_h, spot=spot, spot_err_db=spot_err,
                            batch_fallback=batch_fallback,
                            seconds={"surrogate": t_sur, "pe": t_pe, "total": time.perf_counter() - t0})
