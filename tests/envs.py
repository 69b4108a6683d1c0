"""Environments of the reference's tests, restated as keyword dictionaries.

Each helper cites the reference test it mirrors.  ``oracle_for`` builds the
CPU oracle's single-ray input with exactly the per-ray parameters a CuPyRAM
instance resolved (grid, c0, wavelength), so batched runs compare ray by ray.
"""
import numpy as np


def ram_reference_case():
    """RAM v1.5 test case (tests/test_reference.py:20-121)."""
    return dict(freq=50.0, zs=50.0, zr=50.0,
                z_ss=np.array([0.0, 100.0, 400.0]), rp_ss=np.array([0.0, 25000.0]),
                cw=np.array([[1480.0, 1530.0], [1520.0, 1530.0], [1530.0, 1530.0]]),
                z_sb=np.array([0.0]), rp_sb=np.array([0.0]), cb=np.array([[1700.0]]),
                rhob=np.array([[1.5]]), attn=np.array([[0.5]]),
                rbzb=np.array([[0.0, 200.0], [40000.0, 400.0]]),
                rmax=50000.0, dr=500.0, dz=2.0, zmplt=500.0, c0=1600.0)


def pekeris_batching():
    """Pekeris fixture of tests/test_batching.py:10-38."""
    return dict(freq=50.0, zs=100.0, zr=50.0,
                z_ss=np.array([0.0, 200.0]), rp_ss=np.array([0.0]), cw=np.array([[1500.0], [1500.0]]),
                z_sb=np.array([0.0]), rp_sb=np.array([0.0]), cb=np.array([[1700.0]]),
                rhob=np.array([[1.5]]), attn=np.array([[0.5]]),
                rbzb=np.array([[0.0, 200.0], [10000.0, 200.0]]), rmax=10000.0)


def pekeris_phase1():
    """Pekeris test with a step interface (tests/test_pekeris.py:22-162)."""
    depth = 200.0
    return dict(freq=50.0, zs=100.0, zr=50.0,
                z_ss=np.array([0.0, depth]), rp_ss=np.array([0.0]), cw=np.array([[1500.0], [1500.0]]),
                z_sb=np.array([0.0, depth, depth, depth + 100]), rp_sb=np.array([0.0]),
                cb=np.array([[1500.0], [1500.0], [1700.0], [1700.0]]),
                rhob=np.array([[1.0], [1.0], [1.5], [1.5]]),
                attn=np.array([[0.0], [0.0], [0.5], [0.5]]),
                rbzb=np.array([[0.0, depth], [10000.0, depth]]),
                dz=0.1 * 1500 / 50.0, dr=50.0, rmax=10000.0, np=4, ns=1)


def munk_speed(z, c0=1500.0, eps=0.00737, z_axis=1300.0, scale=1300.0):
    eta = 2.0 * (z - z_axis) / scale
    return c0 * (1.0 + eps * (eta + np.exp(-eta) - 1.0))


def munk_phase2(rmax=50000.0):
    """Munk profile test of tests/test_munk.py:47-59,120-193: 25 Hz, axis
    1000 m, 100 sound-speed knots over 5000 m, zs = zr = 1000 m, a 100 m
    1700 m/s bottom layer, dz = 6 m, dr = 50 m, np = 8, ns = 1."""
    z = np.linspace(0.0, 5000.0, 100)
    return dict(freq=25.0, zs=1000.0, zr=1000.0,
                z_ss=z, rp_ss=np.array([0.0]), cw=munk_speed(z, z_axis=1000.0).reshape(-1, 1),
                z_sb=np.array([5000.0, 5100.0]), rp_sb=np.array([0.0]), cb=np.array([[1700.0], [1700.0]]),
                rhob=np.array([[1.8], [1.8]]), attn=np.array([[0.5], [0.5]]),
                rbzb=np.array([[0.0, 5000.0], [rmax, 5000.0]]),
                dz=0.1 * 1500 / 25.0, dr=50.0, rmax=rmax, np=8, ns=1)


def munk_batching(rmax=50000.0):
    """Munk fixture of tests/test_batching.py:184-214 (axis 1300 m, 100 Hz)."""
    z = np.linspace(0, 5000, 51)
    return dict(freq=100.0, zs=1000.0, zr=2500.0,
                z_ss=z, rp_ss=np.array([0.0]), cw=munk_speed(z).reshape(-1, 1),
                z_sb=np.array([0.0]), rp_sb=np.array([0.0]), cb=np.array([[1600.0]]),
                rhob=np.array([[1.5]]), attn=np.array([[0.2]]),
                rbzb=np.array([[0.0, 5000.0], [rmax, 5000.0]]), rmax=rmax)


def wedge(rmax=10000.0):
    """ASA wedge 200 m -> 10 m over 4 km (tests/test_wedge.py:22-105)."""
    r = np.linspace(0, rmax, 200)
    zb = np.where(r <= 4000.0, 200.0 - (r / 4000.0) * 190.0, 10.0)
    zb = np.maximum(zb, 10.0)
    return dict(freq=50.0, zs=5.0, zr=3.0,
                z_ss=np.array([0.0, 200.0]), rp_ss=np.array([0.0]), cw=np.array([[1500.0], [1500.0]]),
                z_sb=np.array([200.0, 300.0]), rp_sb=np.array([0.0]),
                cb=np.array([[1700.0], [1700.0]]), rhob=np.array([[1.8], [1.8]]),
                attn=np.array([[0.5], [0.5]]), rbzb=np.column_stack((r, zb)),
                dz=0.1 * 1500 / 50.0, dr=50.0, rmax=rmax, np=8, ns=1)


def varying_ranges():
    """Ragged per-ray range tables (tests/test_varying_ranges.py:14-106)."""
    z_ss = [np.array([0.0, 100.0, 200.0])] * 3
    rp = [np.array([0.0, 5000.0, 10000.0]), np.array([0.0, 2000.0, 5000.0, 8000.0, 12000.0]),
          np.array([0.0, 3000.0, 7000.0, 11000.0])]
    # vary the columns so that profile switching changes the field
    cw = [np.array([[1500.0 + 5 * k for k in range(len(r))]] * 3) + np.array([[0.0], [10.0], [20.0]])
          for r in rp]
    z_sb = [np.array([0.0, 50.0])] * 3
    cb = [np.full((2, len(r)), 1600.0) + np.arange(len(r)) * 10.0 for r in rp]
    rhob = [np.full((2, len(r)), 1.5) for r in rp]
    attn = [np.full((2, len(r)), 0.5) for r in rp]
    rbzb = [np.array([[0.0, 200.0], [10000.0, 200.0]]), np.array([[0.0, 200.0], [12000.0, 180.0]]),
            np.array([[0.0, 200.0], [11000.0, 150.0]])]
    return dict(freq=50.0, zs=100.0, zr=50.0, z_ss=z_ss, rp_ss=rp, cw=cw, z_sb=z_sb, rp_sb=rp, cb=cb,
                rhob=rhob, attn=attn, rbzb=rbzb, rmax=10000.0, dr=50.0, dz=2.0, batch_size=3)


def deep_and_shallow():
    """Two isovelocity columns in one batch: 5000 m at dz = 0.15 m (about
    35,300 rows, deeper than every register-team shape, so it takes the
    one-lane-per-ray march) and 200 m (a team shape); dz is given per ray so
    that each column gets its own depth grid."""
    depths = [5000.0, 200.0]
    return dict(freq=100.0, zs=50.0, zr=60.0,
                z_ss=[np.array([0.0, d]) for d in depths], rp_ss=[np.array([0.0])] * 2,
                cw=[np.array([[1500.0], [1500.0]])] * 2, z_sb=[np.array([0.0])] * 2, rp_sb=[np.array([0.0])] * 2,
                cb=[np.array([[1700.0]])] * 2, rhob=[np.array([[1.5]])] * 2, attn=[np.array([[0.5]])] * 2,
                rbzb=[np.array([[0.0, d], [1000.0, d]]) for d in depths], rmax=200.0, dr=10.0, dz=np.array([0.15, 0.15]),
                batch_size=2)


def oracle_for(model, env_index=0, freq_index=0):
    """Oracle input for ray (env, freq) of a CuPyRAM instance after setup()."""
    from oracle import ram_oracle as O
    e = model._env[env_index]
    if model._freq_per_ray:
        f = float(model._freqs[env_index])
    else:
        f = float(model._freqs[freq_index])
    c = env_index * model._n_freq + freq_index
    ray = model._rays[c]
    return O.OracleRay(freq=f, zs=float(model._zs_e[env_index]), zr=model._zr, z_ss=e["z_ss"], rp_ss=e["rp_ss"], cw=e["cw"],
                       z_sb=e["z_sb"], rp_sb=e["rp_sb"], cb=e["cb"], rhob=e["rhob"], attn=e["attn"],
                       rbzb=e["rbzb"], c0=float(ray["c0"]), lam=float(ray["lyrw_lambda"]) / model._lyrw,
                       dr=float(ray["dr"]), dz=float(ray["dz"]), zmax=float(ray["zmax"]),
                       rmax=float(np.asarray(model._rmax_input).ravel()[env_index]
                                  if np.ndim(model._rmax_input) else model._rmax_input),
                       zmplt=float(model._zmplt_e[env_index]), rs=float(ray["rs"]), np_pade=model._np, ns=model._ns,
                       ndr=int(model._ndr_e[env_index]), ndz=int(model._ndz_e[env_index]),
                       lyrw=float(model._lyrw))


def tl_stats(a, b):
    d = np.asarray(a) - np.asarray(b)
    d = d[np.isfinite(d)]
    return {"rms": float(np.sqrt(np.mean(d ** 2))), "max": float(np.max(np.abs(d))),
            "mean_abs": float(np.mean(np.abs(d)))}
