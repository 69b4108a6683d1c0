"""PE-generated training and query sets of the surrogate.

Two query domains, each with a seeded generator and the march setup whose
output grid (the cells of 'TL Cells', or the points of 'TL Grid') is
exactly the domain's output grid:

deep_munk   the 10,000-ray Munk benchmark batch's generator (5000 m Munk
            water, 16 bathymetry knots on 0-50 km from U[3500, 5000] m,
            bottom c/rho/alpha from U[1550, 1900] m/s, U[1.3, 2.0] g/cm3,
            U[0.1, 1.0] dB/lambda, dr 10 m, dz 1 m, np 6), with the sound
            speed perturbed at 32 knots (N(0, 5^2) m/s), freq log-uniform in
            80-125 Hz and zs uniform in 500-1500 m so that every query input
            varies; output grid 200 ranges x 64 depths = every 25th step
            (250 m) and every 78th row (78 m).
wedge       upslope wedge 200 -> 20 m over 10 km, sediment layer 5-30 m thick
            (c U[1550, 1750] m/s, alpha U[0.1, 1.0] dB/lambda, rho from c)
            over a 2000 m/s, 2.2 g/cm3, 0.2 dB/lambda basement, downward-
            refracting water (1520 m/s minus a U[0, 0.1] 1/s gradient, plus
            N(0, 1) m/s at 32 knots), freq log-uniform in 50-1000 Hz, zs
            uniform in 5-50 m; output grid every 50 m x every 4 m (dr, dz
            refined from the default lambda-rules to divide 50 m and 4 m).
            The sediment follows the seabed through 21 range-dependent
            bottom profiles (one every 500 m; each profile's basement top is
            the thickness below the seabed depth at the segment midpoint,
            with a 0.5 m linear transition).
"""
from __future__ import annotations

import math
import sys
import time

import numpy as np

from ..workloads import munk_speed
from .encode import N_BATHY, N_SSP, Domain, Queries

DEEP = Domain(name="deep_munk", rmax=50000.0, zplot=64 * 78.0, ssp_depth=5000.0,
              bottom_mean=(1725.0, 1.65, 0.55), bottom_scale=(100.0, 0.2, 0.26),
              bottom_names=("cb", "rhob", "attn"), freq_lo=80.0, freq_hi=125.0)
WEDGE = Domain(name="wedge", rmax=10000.0, zplot=256.0, ssp_depth=200.0,
               bottom_mean=(17.5, 1650.0, 0.55), bottom_scale=(7.2, 58.0, 0.26),
               bottom_names=("sediment_thickness", "sediment_c", "sediment_attn"),
               freq_lo=50.0, freq_hi=1000.0)
DOMAINS = {d.name: d for d in (DEEP, WEDGE)}


def deep_queries(n, seed):
    rng = np.random.default_rng(seed)
    zk = np.linspace(0.0, DEEP.ssp_depth, N_SSP)
    ssp = munk_speed(zk)[None, :] + rng.normal(0.0, 5.0, size=(n, N_SSP))
    knots_r = np.linspace(0.0, DEEP.rmax, 16)
    depths = rng.uniform(3500.0, 5000.0, size=(n, 16))
    rb = np.linspace(0.0, DEEP.rmax, N_BATHY)
    bathy = np.stack([np.interp(rb, knots_r, d) for d in depths])
    bottom = np.stack([rng.uniform(1550.0, 1900.0, n), rng.uniform(1.3, 2.0, n), rng.uniform(0.1, 1.0, n)], 1)
    freq = np.exp(rng.uniform(np.log(DEEP.freq_lo), np.log(DEEP.freq_hi), n))
    zs = rng.uniform(500.0, 1500.0, n)
    return Queries(ssp, bathy, bottom, freq, zs)


def wedge_queries(n, seed):
    rng = np.random.default_rng(seed)
    zk = np.linspace(0.0, WEDGE.ssp_depth, N_SSP)
    grad = rng.uniform(0.0, 0.1, n)
    ssp = 1520.0 - grad[:, None] * zk[None, :] + rng.normal(0.0, 1.0, size=(n, N_SSP))
    rb = np.linspace(0.0, WEDGE.rmax, N_BATHY)
    bathy = np.tile(200.0 - 180.0 * rb / WEDGE.rmax, (n, 1))
    bottom = np.stack([rng.uniform(5.0, 30.0, n), rng.uniform(1550.0, 1750.0, n), rng.uniform(0.1, 1.0, n)], 1)
    freq = np.exp(rng.uniform(np.log(WEDGE.freq_lo), np.log(WEDGE.freq_hi), n))
    zs = rng.uniform(5.0, 50.0, n)
    return Queries(ssp, bathy, bottom, freq, zs)


def _bathy_table(dom, q, i):
    rb = np.linspace(0.0, dom.rmax, N_BATHY)
    return np.column_stack((rb, q.bathy[i]))


def march_kwargs(dom: Domain, q: Queries, precision="c64"):
    """CuPyRAM keyword arguments that march every query of q onto the
    domain's output grid: 'TL Cells' [n, nzp, nr] (target "cells") or
    'TL Grid' (target "points")."""
    n = len(q)
    zk = np.linspace(0.0, dom.ssp_depth, N_SSP)
    if dom.name == "deep_munk":
        kw = dict(z_sb=np.array([0.0]), rp_sb=np.array([0.0]),
                  cb=[np.array([[v]]) for v in q.bottom[:, 0]], rhob=[np.array([[v]]) for v in q.bottom[:, 1]],
                  attn=[np.array([[v]]) for v in q.bottom[:, 2]],
                  dr=10.0, dz=1.0, ndr=25, ndz=78, zmplt=dom.zplot, rmax=dom.rmax)
    elif dom.name == "wedge":
        dz_def = 0.1 * 1500.0 / q.freq
        dr_def = 6 * 1500.0 / q.freq
        ndz = np.ceil(dom.zplot / dom.nzp / dz_def - 1e-9).astype(np.int64)
        ndr = np.ceil(dom.rmax / dom.nr / dr_def - 1e-9).astype(np.int64)
        dz = dom.zplot / dom.nzp / ndz
        dr = dom.rmax / dom.nr / ndr
        nprof = 21
        rp = np.linspace(0.0, dom.rmax, nprof)
        mid = np.minimum(rp + 0.5 * (rp[1] - rp[0]), dom.rmax)
        rb = np.linspace(0.0, dom.rmax, N_BATHY)
        z_sb, cb, rhob, attn = [], [], [], []
        for i in range(n):
            h, c, a = q.bottom[i]
            rho = 1.5 + 0.5 * (c - 1550.0) / 200.0
            top = np.interp(mid, rb, q.bathy[i]) + h          # basement top of each profile
            # common depth knots of all profiles: every interface and 0.5 m below it
            knots = np.unique(np.concatenate([top, top + 0.5, [300.0]]))
            sed = knots[:, None] <= top[None, :]              # [knot, profile]
            z_sb.append(knots - dom.ssp_depth)                # relative to the deepest sound-speed knot
            cb.append(np.where(sed, c, 2000.0))
            rhob.append(np.where(sed, rho, 2.2))
            attn.append(np.where(sed, a, 0.2))
        kw = dict(z_sb=z_sb, rp_sb=[rp] * n, cb=cb, rhob=rhob, attn=attn,
                  dr=dr, dz=dz, ndr=ndr, ndz=ndz, zmplt=np.full(n, dom.zplot), rmax=np.full(n, dom.rmax))
    else:
        raise ValueError(dom.name)
    kw.update(freq=q.freq, freq_per_ray=True, zs=q.zs, zr=float(min(50.0, dom.ssp_depth)),
              z_ss=zk, rp_ss=np.array([0.0]), cw=[c.reshape(-1, 1) for c in q.ssp],
              rbzb=[_bathy_table(dom, q, i) for i in range(n)], np=6, batch_size=n,
              compute_grids=dom.target == "points", compute_cells=dom.target == "cells", precision=precision)
    return kw
