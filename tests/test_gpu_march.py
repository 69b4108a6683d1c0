"""GPU parity of the HIP march against the CPU oracle (complex128 PyRAM
restatement), through the C ABI.  Tolerances:
  * complex128 march vs oracle: TL line RMS <= 0.1 dB (the precision gate; the
    observed agreement is ~1e-8 dB);
  * complex64 march vs oracle: TL line RMS <= 0.5 dB;
  * RAM v1.5 golden: mean |dTL| <= 0.01 dB (tests/test_reference.py:28) in
    complex128.
"""
import os

import numpy as np
import pytest

from cupyram_amd import CuPyRAM
from oracle import ram_oracle as O
from tests import envs

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "tl_ref.line")
VARIANTS = ["thread", "team"]


def _run(kw, **extra):
    kw = dict(kw)
    kw.update(extra)
    m = CuPyRAM(**kw)
    return m, m.run()


@pytest.mark.parametrize("variant", VARIANTS)
@pytest.mark.parametrize("precision,tol", [("c128", 0.01), ("c64", 0.05)])
def test_ram_v15_golden(require_gpu, variant, precision, tol):
    ref = np.loadtxt(GOLDEN)
    m, res = _run(envs.ram_reference_case(), precision=precision, variant=variant)
    assert np.array_equal(res["Ranges"], ref[:, 0])
    d = np.abs(res["TL Line"] - ref[:, 1])
    print(f"{precision} {variant}: mean {d.mean():.5f} max {d.max():.5f} dB")
    assert d.mean() <= tol


def _compare_with_oracle(kw, precision, variant, tol_rms, grids=True, **extra):
    m, res = _run(kw, precision=precision, variant=variant, compute_grids=grids, **extra)
    nc = m._total_batch
    tl = np.asarray(res["TL Line"]).reshape(nc, -1)
    worst = 0.0
    for e in range(m._n_env):
        for f in range(m._n_freq):
            c = e * m._n_freq + f
            o = O.run(envs.oracle_for(m, e, f), grids=grids)
            s = envs.tl_stats(tl[c, :o["TL Line"].size], o["TL Line"])
            worst = max(worst, s["rms"])
            assert s["rms"] <= tol_rms, (e, f, s)
            if grids:
                g = np.asarray(res["TL Grid"]).reshape(nc, *np.shape(res["TL Grid"])[-2:])[c]
                sg = envs.tl_stats(np.clip(g, 0, 200), np.clip(o["TL Grid"], 0, 200))
                assert sg["rms"] <= 2 * tol_rms, (e, f, sg)
    return worst


@pytest.mark.parametrize("variant", VARIANTS)
@pytest.mark.parametrize("name", ["pekeris_batching", "pekeris_phase1", "wedge", "varying_ranges"])
def test_c128_matches_oracle(require_gpu, variant, name):
    kw = getattr(envs, name)()
    worst = _compare_with_oracle(kw, "c128", variant, 0.1)
    print(name, variant, "worst rms", worst)
    assert worst < 1e-4  # same algorithm in FP64: rounding-level agreement


@pytest.mark.parametrize("variant", VARIANTS)
@pytest.mark.parametrize("name", ["pekeris_batching", "pekeris_phase1", "wedge", "varying_ranges"])
def test_c64_within_half_db(require_gpu, variant, name):
    kw = getattr(envs, name)()
    worst = _compare_with_oracle(kw, "c64", variant, 0.5)
    print(name, variant, "c64 worst rms", worst)


@pytest.mark.parametrize("variant", VARIANTS)
def test_low_frequency_factor_scales(require_gpu, variant):
    """2 Hz, np 8, dr 2 m: the complex64 team march carries the field divided
    by the running product of its factors' scales c = (1 - p1/p2) k0^2/(6 p2),
    whose product over one range step is about 1e-54 here, far outside FP32;
    the per-plan schedule multiplies it back in after 7 of the 8 terms."""
    kw = dict(freq=2.0, zs=500.0, zr=1000.0,
              z_ss=np.array([0.0, 2000.0]), rp_ss=np.array([0.0]), cw=np.array([[1500.0], [1500.0]]),
              z_sb=np.array([0.0]), rp_sb=np.array([0.0]), cb=np.array([[1700.0]]),
              rhob=np.array([[1.5]]), attn=np.array([[0.5]]),
              rbzb=np.array([[0.0, 2000.0], [4000.0, 2000.0]]),
              dz=5.0, dr=2.0, rmax=3000.0, np=8, ns=1)
    w128 = _compare_with_oracle(kw, "c128", variant, 0.1, grids=False)
    w64 = _compare_with_oracle(kw, "c64", variant, 0.5, grids=False)
    print(variant, f"low frequency: c128 worst rms {w128:.2e} dB, c64 {w64:.2e} dB")
    assert w128 < 1e-4


@pytest.mark.parametrize("variant", VARIANTS)
@pytest.mark.parametrize("precision", ["c128", "c64"])
def test_munk_phase2_no_range_drift(require_gpu, variant, precision):
    """The reference's Munk test (tests/test_munk.py:295-330): complex128 TL
    line within its 0.1 dB of the (Py)RAM reference everywhere (complex64:
    0.5 dB RMS), and no range drift: the mean |dTL| over the last quarter of
    the ranges exceeds that of the first quarter by less than 0.5 dB."""
    m, res = _run(envs.munk_phase2(), precision=precision, variant=variant, compute_grids=False)
    o = O.run(envs.oracle_for(m), grids=False)
    tl = np.asarray(res["TL Line"])
    n = min(tl.size, o["TL Line"].size)
    d = np.abs(tl[:n] - o["TL Line"][:n])
    drift = d[3 * n // 4:].mean() - d[:n // 4].mean()
    print(f"{precision} {variant}: max {d.max():.3e} dB, rms {np.sqrt(np.mean(d * d)):.3e} dB, drift {drift:+.3e} dB")
    if precision == "c128":
        assert d.max() <= 0.1
    else:
        assert np.sqrt(np.mean(d * d)) <= 0.5
    assert abs(drift) < 0.5


@pytest.mark.parametrize("variant", VARIANTS)
def test_munk_short(require_gpu, variant):
    kw = envs.munk_batching(rmax=5000.0)
    _compare_with_oracle(kw, "c128", variant, 0.1, grids=False)
    _compare_with_oracle(kw, "c64", variant, 0.5, grids=False)


@pytest.mark.parametrize("variant", VARIANTS)
def test_identical_rays_identical(require_gpu, variant):
    """tests/test_batching.py:91-105: identical rays agree to < 1e-10 dB."""
    m, res = _run(envs.pekeris_batching(), batch_size=8, variant=variant, precision="c64")
    tl = res["TL Line"]
    assert tl.shape[0] == 8
    for i in range(1, 8):
        assert np.max(np.abs(tl[0] - tl[i])) < 1e-10


@pytest.mark.parametrize("variant", VARIANTS)
def test_batch1_vs_batch8(require_gpu, variant):
    """tests/test_batching.py:64-88: B=1 vs B=8 agree to < 0.01 dB."""
    _, r1 = _run(envs.pekeris_batching(), variant=variant)
    _, r8 = _run(envs.pekeris_batching(), batch_size=8, variant=variant)
    assert r1["TL Line"].ndim == 1 and r8["TL Line"].shape[0] == 8
    assert np.max(np.abs(r1["TL Line"] - r8["TL Line"][0])) < 0.01


@pytest.mark.parametrize("variant", VARIANTS)
@pytest.mark.parametrize("precision", ["c64", "c128"])
def test_graph_and_segments_bitwise(require_gpu, variant, precision):
    """Launching the march in graph-captured segments reproduces one launch bitwise."""
    kw = envs.wedge()
    _, a = _run(kw, precision=precision, variant=variant)
    _, b = _run(kw, precision=precision, variant=variant, steps_per_launch=17)
    assert np.array_equal(a["TL Line"], b["TL Line"])
    assert np.array_equal(a["CP Grid"], b["CP Grid"])


def test_multi_frequency_matches_single(require_gpu):
    """tests/test_multi_frequency.py:98-166: multi-frequency equals per-frequency
    runs on the same grid (< 0.1 dB); the frequency whose wavelength sets the
    absorbing layer shares the whole grid and agrees to rounding."""
    z_ss = np.array([0.0, 100.0])
    env = dict(zs=10.0, zr=50.0, z_ss=z_ss, rp_ss=np.array([0.0]), cw=np.array([[1500.0], [1500.0]]),
               z_sb=np.array([0.0, 10.0]), rp_sb=np.array([0.0]), cb=np.array([[1600.0], [1600.0]]),
               rhob=np.array([[1.5], [1.5]]), attn=np.array([[0.5], [0.5]]),
               rbzb=np.array([[0.0, 100.0], [10000.0, 100.0]]), rmax=1000.0, c0=1500.0)
    freqs = [75.0, 125.0]
    lam = 1500.0 / max(freqs)
    env.update(dr=8 * lam, dz=0.4 * lam)
    _, multi = _run(dict(env, freq=freqs))
    assert multi["TL Line"].shape[0] == 2
    for i, f in enumerate(freqs):
        _, single = _run(dict(env, freq=f))
        d = np.max(np.abs(multi["TL Line"][i] - single["TL Line"]))
        print(f, d)
        assert d < 0.1
        if f == max(freqs):
            assert d < 1e-9


@pytest.mark.parametrize("precision,shape", [("c64", (4, 128)), ("c64", (8, 64)), ("c64", (12, 64)),
                                             ("c64", (16, 64)), ("c64", (24, 64)), ("c64", (32, 64)),
                                             ("c64", (8, 128)), ("c64", (16, 128)),
                                             ("c128", (2, 192)), ("c128", (4, 128)), ("c128", (6, 64)),
                                             ("c128", (8, 64)), ("c128", (12, 64)), ("c128", (4, 192)),
                                             ("c64", (40, 64)), ("c64", (48, 64)), ("c64", (64, 64)),
                                             ("c128", (16, 64)), ("c128", (24, 64)), ("c128", (32, 64)),
                                             ("c128", (40, 64)), ("c64", (20, 64)), ("c64", (21, 64)),
                                             ("c64", (28, 64)), ("c128", (10, 64)), ("c128", (11, 64)),
                                             ("c64", (21, 256)), ("c128", (11, 512)), ("c128", (14, 64)),
                                             ("c128", (20, 64)), ("c64", (21, 128))])
def test_every_team_shape_matches_oracle(require_gpu, precision, shape):
    """Register coefficients (c64 4..12, c128 2..6), streamed coefficients
    (c64 14..64, c128 8..40; past c64 24 and c128 12 rows the wave teams'
    lanes spill), wave-shuffle (64 lanes) and LDS (>64 lanes) teams, on the
    wedge (sloping bathymetry moves the water/bottom split).  c64 21 x 128
    reads its streamed coefficients from the team buffer: its LDS rows would
    leave the CU three blocks instead of four."""
    tol = 0.1 if precision == "c128" else 0.5
    worst = _compare_with_oracle(envs.wedge(), precision, "team", tol, grids=True, team_shape=shape)
    if precision == "c128":
        assert worst < 1e-4


@pytest.mark.parametrize("env", ["munk", "wedge"])
@pytest.mark.parametrize("precision,shape", [("c64", (14, 256)), ("c64", (28, 192)), ("c64", (32, 128)),
                                             ("c64", (40, 256)), ("c64", (64, 192)), ("c128", (12, 320)),
                                             ("c128", (14, 256)), ("c128", (16, 384)), ("c128", (24, 192)),
                                             ("c128", (40, 128)), ("c64", (24, 448))])
def test_lds_tail_rows_match_oracle(require_gpu, env, precision, shape):
    """LDS teams whose deepest elimination rows live in LDS (complex64 past 24
    rows per lane, complex128 past 11, where that LDS leaves the CU its
    blocks: 28 x 192, 40 x 256, 64 x 192; 12 x 320, 16 x 384, 24 x 192), the
    same row counts kept in registers (32 x 128; 14 x 256, 40 x 128), the
    14-row complex64 lanes, and streamed complex64 lanes with their
    coefficient rows in LDS (24 x 448), on a 3,535-row Munk column
    (14 x 256 = 3,584 rows fill the lanes) and on the wedge, whose few valid
    lanes leave the rest of the team as padding."""
    kw = envs.munk_batching(rmax=5000.0) if env == "munk" else envs.wedge(rmax=4000.0)
    tol = 0.1 if precision == "c128" else 0.5
    worst = _compare_with_oracle(kw, precision, "team", tol, grids=env == "wedge", team_shape=shape)
    print(env, precision, shape, "worst rms", worst)
    if precision == "c128":
        assert worst < 1e-4


@pytest.mark.parametrize("env,shape", [("wedge", (8, 128)), ("wedge", (8, 192)), ("wedge", (10, 256)),
                                       ("munk", (8, 512)), ("munk", (10, 384)), ("munk", (10, 512))])
def test_c128_lds_coefficient_rows_match_oracle(require_gpu, env, shape):
    """complex128 LDS teams of 8 and 10 rows per lane keep the real
    coefficient halves of their first rows in LDS: 4 rows at 128 lanes, 6 at
    256, 7 at 192, 384 and 512 (what the block's share of the CU's LDS leaves
    beside the team's buffers); the other rows still come from the team
    buffer.  Wedge (padding lanes) and the 3,535-row Munk column."""
    kw = envs.munk_batching(rmax=5000.0) if env == "munk" else envs.wedge(rmax=4000.0)
    worst = _compare_with_oracle(kw, "c128", "team", 0.1, grids=env == "wedge", team_shape=shape)
    print(env, shape, "worst rms", worst)
    assert worst < 1e-4


def _uneven(n=6, rmax=3000.0, seed=11):
    from cupyram_amd import workloads
    kw = workloads.uneven_batch(nrays=n, seed=seed, rmax=rmax)
    return kw


@pytest.mark.parametrize("precision", ["c64", "c128"])
def test_uneven_batch_per_ray_grids(require_gpu, precision):
    """Per-ray frequency, dz = lambda/10 and dr = lambda/2: rays land in
    different kernel buckets and each matches its own oracle run."""
    kw = _uneven()
    m, res = _run(kw, precision=precision)
    assert len(m.plan_variants) >= 2
    tl = res["TL Line"]
    for e in range(m._n_env):
        o = O.run(envs.oracle_for(m, e, 0), grids=False)
        s = envs.tl_stats(tl[e, :o["TL Line"].size], o["TL Line"])
        assert s["rms"] <= (0.1 if precision == "c128" else 0.5), (e, s)
        assert np.all(np.isnan(tl[e, o["TL Line"].size:]))


def test_bucketing_independent_of_batch_composition(require_gpu):
    """A ray's result does not depend on which other rays share its batch."""
    kw = _uneven(n=6)
    _, full = _run(kw, precision="c64")
    sub = dict(kw)
    keep = [1, 4]
    for k in ("z_ss", "cw", "cb", "rhob", "attn", "rbzb"):
        sub[k] = [kw[k][i] for i in keep]
    for k in ("freq", "dz", "dr"):
        sub[k] = np.asarray(kw[k])[keep]
    sub["batch_size"] = 2
    _, part = _run(sub, precision="c64")
    for j, i in enumerate(keep):
        a, b = full["TL Line"][i], part["TL Line"][j]
        n = np.sum(np.isfinite(b))
        assert np.array_equal(a[:n], b[:n])


def _cells_vs_oracle(kw, precision, variant, **extra):
    m, res = _run(kw, precision=precision, variant=variant, compute_cells=True, **extra)
    nc = m._total_batch
    tlc = np.asarray(res["TL Cells"]).reshape(nc, *np.shape(res["TL Cells"])[-2:])
    worst = 0.0
    for e in range(m._n_env):
        o = O.run_cells(envs.oracle_for(m, e, 0))["TL Cells"]
        g = tlc[e, :o.shape[0], :o.shape[1]]
        s = envs.tl_stats(np.clip(g, 0, 200), np.clip(o, 0, 200))
        worst = max(worst, s["rms"])
    return worst


@pytest.mark.parametrize("variant", VARIANTS)
@pytest.mark.parametrize("precision,tol", [("c128", 1e-6), ("c64", 0.05)])
def test_cells_match_oracle(require_gpu, variant, precision, tol):
    """Cell-mean TL (ranges (i ndr, (i+1) ndr] dr x rows (j ndz, (j+1) ndz])
    on the sloping wedge, team lanes sharing depth bins (atomic flushes) and
    the thread march, vs the oracle's cell means."""
    kw = envs.wedge(rmax=4000.0)
    kw.update(zmplt=150.0, ndr=3, ndz=4)
    worst = _cells_vs_oracle(kw, precision, variant)
    print(precision, variant, "cells worst rms", worst)
    assert worst <= tol


@pytest.mark.parametrize("shape", [(4, 128), (12, 64), (24, 64), (48, 64), (8, 128)])
def test_cells_every_lane_layout(require_gpu, shape):
    """Depth bins of 5 rows straddle lane boundaries differently for each
    rows-per-lane count (register, streamed and LDS teams)."""
    kw = envs.wedge(rmax=2000.0)
    kw.update(zmplt=150.0, ndr=2, ndz=5)
    worst = _cells_vs_oracle(kw, "c64", "team", team_shape=shape)
    assert worst <= 0.05, (shape, worst)


def test_cells_uneven_batch(require_gpu):
    """Per-ray grids in several buckets: each ray's cells vs its oracle run."""
    kw = _uneven(n=6, rmax=2000.0)
    kw.update(ndr=2, ndz=3, zmplt=90.0)
    worst = _cells_vs_oracle(kw, "c64", "auto")
    assert worst <= 0.05, worst


def test_benchmark_cli_batch_run(require_gpu):
    """The batch-scaling CLI's measurement of one small batch (two frequencies)."""
    from cupyram_amd import benchmark as B
    bpr = B.bytes_per_ray([100.0, 150.0], 2000.0, 10.0, 10.0, 4, "c64", probe=32)
    assert bpr > 0
    r = B.run_batch(64, [100.0, 150.0], 2000.0, 10.0, 10.0, 4, "c64")
    assert r["total_fields"] == 128
    assert 0 < r["time_propagation"] < r["time_total"]


def test_many_buckets_every_ray_marched(require_gpu):
    """20,000 sloping-wedge rays (six team buckets, 350 to 2,234 rows, each
    plan on its own non-blocking stream): every ray's cell-mean TL agrees
    with the complex128 march.  Regression test for the team ray list being
    cleared after its upload (a copy not ordered on the plan's stream), which
    left one bucket's rays at their starting field."""
    from cupyram_amd.surrogate import dataset as D
    dom = D.WEDGE
    q = D.wedge_queries(20000, 5)
    tl = {}
    for prec in ("c64", "c128"):
        m = CuPyRAM(**D.march_kwargs(dom, q, precision=prec))
        if prec == "c64":
            m.setup()
            assert len(m._buckets()) >= 5
        tl[prec] = np.clip(np.asarray(m.run()["TL Cells"]), 20.0, 140.0)
    err = np.sqrt(np.nanmean((tl["c64"] - tl["c128"]) ** 2, axis=(1, 2)))
    assert np.all(np.isfinite(err))
    assert err.max() <= 0.5, (int(np.argmax(err)), float(err.max()))


def test_outputs_read_in_chunks(require_gpu):
    """40 rays whose pressure grids (1000 ranges x 400 depths) fill several
    32 MiB staging chunks of the ray-major output read: every output array
    equals, ray by ray and bitwise, that of the same rays in batches of 10."""
    n = 40
    base = envs.pekeris_batching()
    cb = np.linspace(1600.0, 1800.0, n)
    lists = {k: [base[k]] * n for k in ("z_ss", "rp_ss", "cw", "z_sb", "rp_sb", "rhob", "attn", "rbzb")}
    lists["cb"] = [np.array([[c]]) for c in cb]

    def run(sel):
        kw = dict(freq=base["freq"], zs=base["zs"], zr=base["zr"], rmax=5000.0, dr=5.0, dz=0.5, zmplt=200.0,
                  compute_grids=True, batch_size=len(sel), precision="c64")
        kw.update({k: [v[i] for i in sel] for k, v in lists.items()})
        return CuPyRAM(**kw).run()

    full = run(range(n))
    per_ray = 2 * int(np.prod(np.shape(full["CP Grid"])[1:]))   # doubles of one ray's pressure grid
    assert per_ray * n > 3 * (4 << 20)                          # more than three 32 MiB chunks
    for a in range(0, n, 10):
        part = run(range(a, a + 10))
        for key in ("TL Line", "CP Line", "TL Grid", "CP Grid"):
            np.testing.assert_array_equal(np.asarray(full[key])[a:a + 10], np.asarray(part[key]), err_msg=key)


@pytest.mark.parametrize("precision,tol", [("c128", 0.1), ("c64", 0.5)])
def test_munk_full_size_sample(require_gpu, precision, tol):
    """Eight seeded rays of the 10,000-ray Munk workload at its full size
    (nz 5303, 50 km, 4,999 range steps; the headline team shapes) against the
    FP64 oracle run over the same range: TL line RMS within the gate."""
    from cupyram_amd import workloads
    from oracle import ram_oracle as O
    kw = workloads.munk_batch(nrays=10000)
    pick = np.sort(np.random.default_rng(21).choice(10000, size=8, replace=False))
    for k in ("cb", "rhob", "attn", "rbzb"):
        if isinstance(kw[k], list):
            kw[k] = [kw[k][i] for i in pick]
    kw["batch_size"] = len(pick)
    m = CuPyRAM(**kw, precision=precision)
    tl = np.asarray(m.run()["TL Line"])
    assert m.plan_variants[0][1][0] == 2 and int(m._rays["nz"][0]) == 5303
    ref = O.run_batch_tl([envs.oracle_for(m, e, 0) for e in range(len(pick))], nthreads=min(16, os.cpu_count() or 1))
    d = np.concatenate([tl[i, :len(ref[i])] - ref[i] for i in range(len(pick))])
    d = d[np.isfinite(d)]
    assert len(d) > 8 * 4900
    rms = float(np.sqrt(np.mean(d ** 2)))
    print(f"munk full size {precision}: TL line RMS {rms:.3g} dB, max {float(np.max(np.abs(d))):.3g} dB over {len(d)} points")
    assert rms <= tol, (rms, float(np.max(np.abs(d))))


@pytest.mark.parametrize("precision,tol", [("c128", 0.1), ("c64", 0.5)])
def test_column_deeper_than_every_team_shape(require_gpu, precision, tol):
    """A 35,300-row column (one-lane-per-ray march) and a shallow one (team
    march) in one batch, each against the oracle."""
    _compare_with_oracle(envs.deep_and_shallow(), precision, "auto", tol, grids=False)
