"""Seeded synthetic workloads of the benchmark configurations.

munk_batch      10,000 rays, Munk profile (axis 1300 m, eps 0.00737) 5000 m
                deep at 100 Hz; piecewise-linear bathymetry with 16 knots on
                0-50 km drawn from U[3500, 5000] m; bottom c, rho, alpha from
                U[1550, 1900] m/s, U[1.3, 2.0] g/cm3, U[0.1, 1.0] dB/lambda;
                dr 10 m, dz 1 m, np 6, zs 1000 m, zr 2500 m (seed 2).
uneven_batch    50,000 rays, depth log-uniform 100-5000 m, freq log-uniform
                25-500 Hz, dz = lambda/10, dr = lambda/2, rmax 20 km, np 6,
                Munk profile scaled to the depth plus 32 knots of N(0, 5^2)
                m/s, zs 25 m, zr 50 m (seed 3).
wedge_queries   upslope wedge 200 -> 20 m over 10 km, sediment layer 5-30 m
                over basement, freq log-uniform 50-1000 Hz (seed 5).
pekeris_c1      the single CPU-runnable Pekeris ray (100 m of 1500 m/s water
                over a 1700 m/s, 1.5 g/cm3, 0.5 dB/lambda bottom, 50 Hz).

Every generator returns the keyword dictionary of ``CuPyRAM(**kw)``.
"""
from __future__ import annotations

import numpy as np


def munk_speed(z, axis=1300.0, eps=0.00737, c_axis=1500.0, scale=1300.0):
    eta = 2.0 * (np.asarray(z, np.float64) - axis) / scale
    return c_axis * (1.0 + eps * (eta + np.exp(-eta) - 1.0))


def munk_batch(nrays=10000, seed=2, rmax=50000.0, np_pade=6):
    rng = np.random.default_rng(seed)
    z_ss = np.linspace(0.0, 5000.0, 51)
    cw = munk_speed(z_ss).reshape(-1, 1)
    knots_r = np.linspace(0.0, 50000.0, 16)
    depths = rng.uniform(3500.0, 5000.0, size=(nrays, 16))
    cb = rng.uniform(1550.0, 1900.0, size=nrays)
    rho = rng.uniform(1.3, 2.0, size=nrays)
    att = rng.uniform(0.1, 1.0, size=nrays)
    rbzb = [np.column_stack((knots_r, depths[i])) for i in range(nrays)]
    return dict(freq=100.0, zs=1000.0, zr=2500.0, z_ss=z_ss, rp_ss=np.array([0.0]), cw=cw,
                z_sb=np.array([0.0]), rp_sb=np.array([0.0]),
                cb=[np.array([[v]]) for v in cb], rhob=[np.array([[v]]) for v in rho],
                attn=[np.array([[v]]) for v in att], rbzb=rbzb,
                rmax=rmax, dr=10.0, dz=1.0, np=np_pade, batch_size=nrays, compute_grids=False)


def uneven_batch(nrays=50000, seed=3, rmax=20000.0, np_pade=6):
    rng = np.random.default_rng(seed)
    depth = np.exp(rng.uniform(np.log(100.0), np.log(5000.0), size=nrays))
    freq = np.exp(rng.uniform(np.log(25.0), np.log(500.0), size=nrays))
    z_ss, cw, c0 = [], [], np.zeros(nrays)
    for i in range(nrays):
        z = np.linspace(0.0, depth[i], 32)
        c = munk_speed(z * (5000.0 / depth[i])) + rng.normal(0.0, 5.0, size=32)
        z_ss.append(z)
        cw.append(c.reshape(-1, 1))
        c0[i] = c.mean()
    lam = c0 / freq
    cb = rng.uniform(1550.0, 1900.0, size=nrays)
    rho = rng.uniform(1.3, 2.0, size=nrays)
    att = rng.uniform(0.1, 1.0, size=nrays)
    return dict(freq=freq, freq_per_ray=True, zs=25.0, zr=50.0,
                z_ss=z_ss, rp_ss=np.array([0.0]), cw=cw, z_sb=np.array([0.0]), rp_sb=np.array([0.0]),
                cb=[np.array([[v]]) for v in cb], rhob=[np.array([[v]]) for v in rho],
                attn=[np.array([[v]]) for v in att],
                rbzb=[np.array([[0.0, d], [rmax, d]]) for d in depth],
                rmax=rmax, dz=lam / 10.0, dr=lam / 2.0, np=np_pade, batch_size=nrays, compute_grids=False)


def pekeris_c1():
    return dict(freq=50.0, zs=25.0, zr=30.0, z_ss=np.array([0.0, 100.0]), rp_ss=np.array([0.0]),
                cw=np.array([[1500.0], [1500.0]]), z_sb=np.array([0.0, 300.0]), rp_sb=np.array([0.0]),
                cb=np.array([[1700.0], [1700.0]]), rhob=np.array([[1.5], [1.5]]),
                attn=np.array([[0.5], [0.5]]), rbzb=np.array([[0.0, 100.0], [5000.0, 100.0]]),
                rmax=5000.0, dr=5.0, dz=0.5, np=8, ns=1, lyrw=10)


def cell_steps(model) -> float:
    """Sum over rays of nz * nsteps of a CuPyRAM instance after setup()."""
    r = model._rays
    return float(np.sum(r["nz"].astype(np.float64) * r["nsteps"].astype(np.float64)))
