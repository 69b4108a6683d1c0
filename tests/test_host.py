"""Host-side logic and the C ABI surface, on the CPU (no GPU calls)."""
import os
import re

import numpy as np
import pytest

from cupyram_amd import AsyncBatcher, CuPyRAM, _abi, compute_pade_coefficients
from cupyram_amd.pade import compute_pade_batch
from oracle import ram_oracle as O
from tests import envs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    header = open(os.path.join(ROOT, "include", "cupyram_hip.h")).read()
    declared = set(re.findall(r"\b(cpr_[a-z0-9_]+)\s*\(", header))
    assert len(declared) >= 15
    lib = _abi.lib()
    for name in sorted(declared):
        assert hasattr(lib, name), name
    assert declared <= set(_abi._SYMBOLS)


def test_reference_import_names():
    """The reference's package and submodule names (cupyram/__init__.py:1-3,
    15-19; the reference's tests import ``from cupyram import CuPyRAM`` and
    ``from cupyram import CuPyRAM, compute_pade_coefficients``) resolve to
    this build's objects."""
    import cupyram
    from cupyram import AsyncBatcher as A, CuPyRAM as M, compute_pade_coefficients as P
    from cupyram.batching import AsyncBatcher as A2
    from cupyram.cupyram import CuPyRAM as M2
    from cupyram.pade import compute_pade_coefficients as P2
    assert M is M2 is CuPyRAM and P is P2 is compute_pade_coefficients and A is A2 is AsyncBatcher
    assert set(cupyram.__all__) == {"CuPyRAM", "compute_pade_coefficients", "AsyncBatcher"}
    pd1, pd2 = P(freq=100.0, c0=1500.0, np_pade=6, ns=1, dr=50.0, ip=1)
    o1, o2 = O.pade(100.0, 1500.0, 6, 1, 50.0, 1)
    np.testing.assert_allclose(pd1, o1, rtol=1e-12)
    np.testing.assert_allclose(pd2, o2, rtol=1e-12)


def test_device_info_without_device_fails_loudly():
    if _abi.device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(_abi.CprError):
        _abi.device_info(0)


def test_plan_create_without_device_fails_loudly():
    if _abi.device_count() > 0:
        pytest.skip("a GPU is visible")
    m = CuPyRAM(**envs.pekeris_batching())
    with pytest.raises(_abi.CprError, match="no HIP device"):
        m.run()


def test_plan_copies_are_ordered_on_the_plan_stream():
    """Plans run on non-blocking streams, which the null stream does not wait
    for: a synchronous hipMemcpy/hipMemset beside the plan's own
    hipMemsetAsync clearing of a buffer can land before or after it.  (An
    upload of the team ray list that lost this race had every team march ray
    0, leaving the other rays at their starting field.)"""
    src = open(os.path.join(ROOT, "cupyram_amd", "csrc", "cpr_abi.hip")).read()
    assert not re.findall(r"\bhipMemcpy\s*\(", src)
    assert not re.findall(r"\bhipMemset\s*\(", src)
    assert "hipMemcpyAsync(p->team_rays_d" in src


@pytest.mark.parametrize("np_pade", [4, 6, 8])
@pytest.mark.parametrize("ns,ip", [(1, 1), (0, 1), (1, 2)])
@pytest.mark.parametrize("freq", [25.0, 50.0, 100.0, 200.0])
@pytest.mark.parametrize("dr", [10.0, 25.0, 50.0, 100.0])
def test_pade_matches_oracle(np_pade, ns, ip, freq, dr):
    """The reference's Pade grid and 1e-8 gate (tests/test_pade_coefficients.py:23-68)."""
    a1, a2 = compute_pade_coefficients(freq, 1500.0, np_pade, ns, dr, ip)
    b1, b2 = O.pade(freq, 1500.0, np_pade, ns, dr, ip)
    assert np.max(np.abs(a1 - b1)) < 1e-8 and np.max(np.abs(a2 - b2)) < 1e-8


def test_pade_batch_equals_scalar():
    f = np.array([25.0, 50.0, 100.0])
    c0 = np.array([1480.0, 1500.0, 1520.0])
    out = compute_pade_batch(f, c0, 6, 1, 10.0, 1, nthreads=2)
    for i in range(3):
        p1, p2 = compute_pade_coefficients(f[i], c0[i], 6, 1, 10.0, 1)
        assert np.array_equal(out[i, 0], p1) and np.array_equal(out[i, 1], p2)


def test_pade_rejects_bad_arguments():
    with pytest.raises(ValueError):
        compute_pade_coefficients(50.0, 1500.0, 0, 1, 10.0, 1)
    with pytest.raises(ValueError):
        compute_pade_coefficients(-5.0, 1500.0, 6, 1, 10.0, 1)


def test_normalize_ragged_lists_pads_with_nan():
    x = CuPyRAM._normalize_to_batch([np.array([1.0, 2.0]), np.array([3.0, 4.0, 5.0])], 2, "x")
    assert x.shape == (2, 3) and np.isnan(x[0, 2]) and x[1, 2] == 5.0
    t = CuPyRAM._normalize_to_batch(np.array([1.0, 2.0]), 3, "t")
    assert t.shape == (3, 2)
    s = CuPyRAM._normalize_to_batch(7.0, 4, "s")
    assert s.shape == (4,)
    with pytest.raises(ValueError, match="list length"):
        CuPyRAM._normalize_to_batch([np.zeros(2)], 2, "bad")


def test_valid_slice():
    a, n = CuPyRAM._get_valid_slice(np.array([1.0, 2.0, np.nan]))
    assert n == 2 and a.tolist() == [1.0, 2.0]
    b = np.full((3, 4), np.nan)
    b[:2, :3] = 1.0
    v, n = CuPyRAM._get_valid_slice(b)
    assert v.shape == (2, 3) and n == 2


def test_input_validation_errors():
    kw = envs.pekeris_batching()
    with pytest.raises(ValueError, match="Source depth"):
        CuPyRAM(**dict(kw, zs=500.0))
    with pytest.raises(ValueError, match="Receiver depth"):
        CuPyRAM(**dict(kw, zr=-1.0))
    with pytest.raises(ValueError, match="inconsistent"):
        CuPyRAM(**dict(kw, cw=np.array([[1500.0, 1500.0], [1500.0, 1500.0]])))
    with pytest.raises(ValueError, match="Max bathy"):
        CuPyRAM(**dict(kw, rbzb=np.array([[0.0, 300.0], [10000.0, 300.0]])))
    with pytest.raises(ValueError, match="np must be"):
        CuPyRAM(**dict(kw, np=40))


def test_default_grid_uses_max_frequency():
    """tests/test_multi_frequency.py:205-230: dr = 60 m, dz = 0.75 m for f = [50, 100, 200]."""
    kw = dict(envs.pekeris_batching(), freq=[50.0, 100.0, 200.0])
    kw.pop("rmax")
    m = CuPyRAM(**kw, rmax=1000.0)
    assert m._dr == pytest.approx(60.0) and m._dz == pytest.approx(0.75)


@pytest.mark.parametrize("name", ["ram_reference_case", "pekeris_batching", "pekeris_phase1", "wedge",
                                  "varying_ranges", "munk_batching"])
def test_setup_grid_matches_oracle(name):
    m = CuPyRAM(**getattr(envs, name)())
    m.setup()
    for e in range(m._n_env):
        nz, nvr, nvz, nsteps = envs.oracle_for(m, e).dims()
        r = m._rays[e]
        assert (r["nz"], r["nvr"], r["nvz"], r["nsteps"]) == (nz, nvr, nvz, nsteps)
    # attributes the reference's tests read after run()/setup()
    assert m.nz == int(m._rays["nz"].max()) and m._dz > 0 and m._dr > 0
    assert len(m.vr) == m._nvr


@pytest.mark.parametrize("name", ["ram_reference_case", "pekeris_batching", "pekeris_phase1", "wedge",
                                  "munk_batching", "pekeris_c1"])
def test_resolved_parameters_match_independent_defaults(name):
    """The product's resolved c0, wavelength, zmax, nz, dr, dz, rs and grid
    counts of a single-ray environment against the oracle's own restatement
    of the defaults (oracle.resolve_ray: c0 = mean of the first sound-speed
    column, lambda = c0/f, dr = np 1500/f, dz = 0.1 1500/f, as in
    cupyram.py:460-608), built from the keyword arguments alone and not from
    the instance."""
    from cupyram_amd import workloads
    kw = workloads.pekeris_c1() if name == "pekeris_c1" else getattr(envs, name)()
    m = CuPyRAM(**kw)
    m.setup()
    ray = O.resolve_ray(**kw)
    r = m._rays[0]
    assert r["c0"] == pytest.approx(ray.c0, rel=1e-14)
    assert r["lyrw_lambda"] == pytest.approx(ray.lyrw * ray.lam, rel=1e-14)
    assert r["zmax"] == pytest.approx(ray.zmax, rel=1e-14)
    assert r["dz"] == pytest.approx(ray.dz, rel=1e-15) and r["dr"] == pytest.approx(ray.dr, rel=1e-15)
    assert r["rs"] == pytest.approx(ray.rs, rel=1e-14)
    assert r["k0"] == pytest.approx(2 * np.pi * ray.freq / ray.c0, rel=1e-14)
    nz, nvr, nvz, nsteps = ray.dims()
    assert (r["nz"], r["nvr"], r["nvz"], r["nsteps"]) == (nz, nvr, nvz, nsteps)
    assert m._zmplt_e[0] == pytest.approx(ray.zmplt)


def test_per_ray_grids():
    kw = envs.varying_ranges()
    kw["dz"] = np.array([1.0, 2.0, 1.5])
    kw["dr"] = np.array([25.0, 50.0, 40.0])
    m = CuPyRAM(**kw)
    m.setup()
    assert m._per_ray_grid
    assert len(set(m._rays["nz"])) == 3
    for e in range(3):
        nz, nvr, nvz, nsteps = envs.oracle_for(m, e).dims()
        r = m._rays[e]
        assert (r["nz"], r["nvr"], r["nsteps"]) == (nz, nvr, nsteps)


def test_setup_runs_once():
    """setup() resolves everything from the constructor's arguments, so a
    second call (run() makes one) keeps the first call's tables: a batch
    prepared on another thread is not set up again."""
    m = CuPyRAM(**envs.varying_ranges())
    m.setup()
    rays, pade = m._rays, m._pade
    m.setup()
    assert m._rays is rays and m._pade is pade
    fresh = CuPyRAM(**envs.varying_ranges())
    fresh.setup()
    np.testing.assert_array_equal(fresh._rays, rays)
    np.testing.assert_array_equal(fresh._pade, pade)


def _square(x):
    return x * x


def test_async_batcher_order_and_prefetch():
    got = list(AsyncBatcher(((i,) for i in range(7)), _square, max_prefetch=3, max_workers=2))
    assert got == [i * i for i in range(7)]
    assert list(AsyncBatcher(iter([]), _square)) == []


def test_shared_object_roundtrip():
    obj = {"a": np.arange(5), "b": "text"}
    name, size = AsyncBatcher.share_object(obj)
    try:
        back = AsyncBatcher.load_shared_object(name, size)
        assert back["b"] == "text" and np.array_equal(back["a"], obj["a"])
    finally:
        AsyncBatcher.cleanup_shared_object(name)
    AsyncBatcher.cleanup_shared_object(name)  # second cleanup is a no-op


def test_benchmark_cli_host_logic():
    """Batch series and synthetic environment of the batch-scaling CLI
    (cupyram/benchmark.py:86-194 semantics)."""
    from cupyram_amd import benchmark as B
    env = B.synthetic_environment(2000.0)
    assert np.array_equal(env["rp_ss"], [0.0, 1000.0, 2000.0])
    assert env["cw"].shape == (6, 3) and np.all(env["cw"] == 1500.0)
    assert np.all(env["rbzb"][:, 1] == 3000.0)
    # plenty of memory: capped at 50,000 rays, 40 % .. 120 %, multiples of 128, no repeats
    s = B.batch_sizes(1e6, 10**12, 1, 9)
    assert s[0] == round(20000 / 128) * 128 and s[-1] == round(60000 / 128) * 128
    assert all(b % 128 == 0 for b in s) and len(s) == len(set(s)) and s == sorted(s)
    # little memory: a linear series from 32
    s = B.batch_sizes(1e6, 300 * 10**6, 1, 9)
    assert s[0] == 32 and s[-1] <= 270
    # frequencies share the memory
    assert B.batch_sizes(1e6, 10**9, 2, 9)[-1] < B.batch_sizes(1e6, 10**9, 1, 9)[-1]


@pytest.mark.parametrize("case", ["varying_ranges", "wedge_queries"])
def test_packed_tables_equal_per_ray_valid_slices(case):
    """The batch-vectorised validation and packing keep, for every ray, exactly
    the tables the per-ray rule (_get_valid_slice, the bathymetry extension to
    rmax) gives, at the offsets the kernels read."""
    if case == "varying_ranges":
        kw = envs.varying_ranges()
    else:
        from cupyram_amd.surrogate import dataset as D
        kw = D.march_kwargs(D.WEDGE, D.wedge_queries(7, 3))
    m = CuPyRAM(**kw)
    m.setup()
    V = CuPyRAM._get_valid_slice
    d = m._env_data
    for b in range(m._n_env):
        rec = m._envs[b * m._n_freq]
        pick = lambda key, n: d[rec[key]:rec[key] + n]  # noqa: E731
        zss = V(np.asarray(kw["z_ss"][b] if isinstance(kw["z_ss"], list) else kw["z_ss"], float))[0]
        assert np.array_equal(pick("z_ss", rec["n_ss"]), zss)
        cw = np.asarray(kw["cw"][b] if isinstance(kw["cw"], list) else kw["cw"], float)
        cw = V(cw if cw.ndim == 2 else cw[:, None])[0]
        assert np.array_equal(pick("cw", cw.size), cw.ravel())
        zsb = V(np.asarray(kw["z_sb"][b] if isinstance(kw["z_sb"], list) else kw["z_sb"], float).ravel())[0]
        assert np.array_equal(pick("z_sb", rec["n_sb"]), zsb + zss[-1])
        for name in ("cb", "rhob", "attn"):
            a = np.asarray(kw[name][b] if isinstance(kw[name], list) else kw[name], float)
            a = V(a if a.ndim == 2 else a[:, None])[0]
            assert np.array_equal(pick(name, a.size), a.ravel())
        bt = V(np.asarray(kw["rbzb"][b] if isinstance(kw["rbzb"], list) else kw["rbzb"], float).reshape(-1, 2))[0]
        rmax = float(np.ravel(kw["rmax"])[b]) if np.ndim(kw["rmax"]) else float(kw["rmax"])
        if bt[-1, 0] < rmax:
            bt = np.vstack([bt, [rmax, bt[-1, 1]]])
        assert rec["n_bt"] == len(bt)
        assert np.array_equal(pick("rbzb", 2 * len(bt)), bt.ravel())


def test_buckets_deepest_first_and_longest_rays_first():
    """Uneven batches: buckets go deepest first, and the rays of a bucket go
    longest (most range steps) first, stably; every ray lands in one bucket."""
    from cupyram_amd import workloads
    m = CuPyRAM(**workloads.uneven_batch(nrays=400, seed=5), precision="c64")
    m.setup()
    buckets = m._buckets()
    nz, ns = m._rays["nz"], m._rays["nsteps"]
    seen = np.concatenate([idx for idx, _, _ in buckets])
    assert np.array_equal(np.sort(seen), np.arange(400))
    deepest = [int(nz[idx].max()) for idx, _, _ in buckets]
    assert deepest == sorted(deepest, reverse=True)
    for idx, shape, variant in buckets:
        assert np.all(np.diff(ns[idx]) <= 0)
        ties = np.diff(ns[idx]) == 0
        assert np.all(np.diff(idx)[ties] > 0)


@pytest.mark.parametrize("precision", ["c64", "c128"])
def test_column_deeper_than_every_team_shape_takes_the_thread_march(precision):
    """A column of about 35,300 rows exceeds the largest team shape of both
    precisions: 'auto' puts it in a one-lane-per-ray bucket beside the team
    bucket of the shallow ray, and variant='team' refuses the batch."""
    m = CuPyRAM(**envs.deep_and_shallow(), precision=precision)
    m.setup()
    assert int(m._rays["nz"].max()) > 32768
    kinds = sorted(v for _, _, v in m._buckets())
    assert kinds == ["team", "thread"]
    t = CuPyRAM(**envs.deep_and_shallow(), precision=precision, variant="team")
    t.setup()
    with pytest.raises(ValueError, match="too deep"):
        t._buckets()


@pytest.mark.parametrize("precision", [64, 128])
def test_tuning_table_and_static_shape_model(precision):
    """The shipped tuning table and the library's static model both give, for
    every column depth they accept, a built row count with a lane count that
    is a multiple of 64 (at most 512) covering the column; past the deepest
    team (complex64 64 x 512 rows, complex128 40 x 512) the static model
    answers (0, 0), the one-lane-per-ray march."""
    from cupyram_amd import autotune
    table = autotune.load_table()
    assert table.get(str(precision)), "tuning table missing"
    deepest = 64 * 512 if precision == 64 else 40 * 512
    for nz in list(range(50, 6000, 37)) + [8000, 12000, 16000, 20480, deepest]:
        for shape in (autotune.tuned_shape(precision, nz), _abi.team_shape(precision, nz)):
            if shape is None:
                continue
            rows, lanes = shape
            assert rows in autotune.ROWS[precision], (nz, shape)
            assert lanes % 64 == 0 and 64 <= lanes <= 512 and rows * lanes >= nz, (nz, shape)
    assert _abi.team_shape(precision, deepest + 1) == (0, 0)


def test_seven_wave_teams_take_512_lanes():
    """Columns whose tuned row count would give 448-lane teams (seven waves
    on a CU's four SIMDs) take 512 lanes: every complex64 row count, and
    complex128 up to 14 rows; deeper complex128 lanes keep 448."""
    from cupyram_amd import autotune
    assert autotune.lanes_for(3584, 8) == 448 and autotune.tuned_shape(128, 3584) == (8, 512)
    assert autotune.lanes_for(8956, 20) == 448 and autotune.tuned_shape(128, 8956) == (20, 448)
    assert autotune.tuned_shape(64, 8956) == (20, 512)


def test_tuning_table_merge_replaces_only_swept_depths():
    """A partial sweep (autotune --nz ... --merge) replaces the entries of its
    precision and depths and keeps the others, sorted by depth."""
    from cupyram_amd import autotune
    old = {"64": [[256, 4], [512, 8], [1024, 16]], "64_mcs": [[256, 4, 1.0], [512, 8, 2.0], [1024, 16, 3.0]],
           "128": [[512, 8]], "128_mcs": [[512, 8, 4.0]]}
    new = {"64": [[512, 12], [2048, 16]], "64_mcs": [[512, 12, 5.0], [2048, 16, 6.0]]}
    out = autotune.merge_tables(old, new)
    assert out["64_mcs"] == [[256, 4, 1.0], [512, 12, 5.0], [1024, 16, 3.0], [2048, 16, 6.0]]
    assert out["64"] == [[256, 4], [512, 12], [1024, 16], [2048, 16]]
    assert out["128"] == old["128"] and out["128_mcs"] == old["128_mcs"]
