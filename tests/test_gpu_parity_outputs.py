"""GPU parity of outputs and paths the basic march tests leave out, each
against the CPU oracle (complex128 PyRAM restatement) through the C ABI:

  * the Pekeris single-ray configuration (100 m of 1500 m/s water over a
    1700 m/s, 1.5 g/cm3, 0.5 dB/lambda bottom, 50 Hz, np 8) in both
    precisions, against the oracle and a textbook normal-mode sum;
  * multi-frequency batches (cupyram.py:240-254, 610-668, 804-848;
    tests/test_multi_frequency.py:98-166) against oracle runs whose grid is
    built here from the reference's rule (lambda = c0 / max f);
  * decimated grids (ndr, ndz > 1; outpt.py:36-53): TL Grid, CP Line and
    CP Grid;
  * the stability-constraint switch at r >= rs (cupyram.py:976-999).

Tolerances: complex128 TL RMS <= 1e-4 dB (same algorithm in FP64: rounding-
level agreement), complex128 pressure within 1e-6 of the field's peak;
complex64 TL RMS <= 0.5 dB.
"""
import numpy as np
import pytest

from cupyram_amd import CuPyRAM, workloads
from oracle import ram_oracle as O
from tests import envs
from tests.test_oracle import pekeris_tl

pytestmark = pytest.mark.gpu
VARIANTS = ["thread", "team"]
TOL = {"c128": 1e-4, "c64": 0.5}


def _stats(a, b):
    return envs.tl_stats(np.clip(a, 0, 200), np.clip(b, 0, 200))


def _check_pressure(cp, ocp, precision, what):
    """Complex pressure: complex128 to 1e-6 of the peak; complex64 through
    its level in dB (TL RMS) only, since phase error accumulates over range
    at FP32 while the level stays within the TL gate."""
    cp, ocp = np.asarray(cp), np.asarray(ocp)
    assert cp.shape == ocp.shape, (what, cp.shape, ocp.shape)
    if precision == "c128":
        err = np.max(np.abs(cp - ocp)) / np.max(np.abs(ocp))
        assert err <= 1e-6, (what, err)
    else:
        lv = lambda p: -20 * np.log10(np.abs(p) + 1e-20)  # noqa: E731
        s = _stats(lv(cp), lv(ocp))
        assert s["rms"] <= 0.5, (what, s)


@pytest.mark.parametrize("variant", VARIANTS)
@pytest.mark.parametrize("precision", ["c128", "c64"])
def test_pekeris_single_ray(require_gpu, variant, precision):
    """The CPU-runnable Pekeris configuration on the HIP march: TL line vs
    the oracle run from the keyword arguments (resolve_ray), and vs the
    normal-mode sum over 1-5 km within the oracle's own 1 dB gate."""
    kw = workloads.pekeris_c1()
    res = CuPyRAM(**kw, precision=precision, variant=variant).run()
    o = O.run(O.resolve_ray(**kw), grids=False)
    assert np.array_equal(res["Ranges"], o["Ranges"])
    s = envs.tl_stats(res["TL Line"], o["TL Line"])
    print(f"pekeris {precision} {variant}: vs oracle rms {s['rms']:.3g} max {s['max']:.3g} dB")
    assert s["rms"] <= TOL[precision], s
    r = res["Ranges"]
    nm = pekeris_tl(r, 50.0, 100.0, 1500.0, 1700.0, 1.0, 1.5, 0.5, 25.0, 30.0)
    sel = r >= 1000.0
    rms_nm = float(np.sqrt(np.mean((res["TL Line"][sel] - nm[sel]) ** 2)))
    print(f"pekeris {precision} {variant}: vs normal modes rms {rms_nm:.3f} dB")
    assert rms_nm <= 1.0


def _multi_freq_env():
    env = dict(zs=10.0, zr=50.0, z_ss=np.array([0.0, 100.0]), rp_ss=np.array([0.0]),
               cw=np.array([[1500.0], [1500.0]]), z_sb=np.array([0.0, 10.0]), rp_sb=np.array([0.0]),
               cb=np.array([[1600.0], [1650.0]]), rhob=np.array([[1.5], [1.5]]), attn=np.array([[0.5], [0.5]]),
               rbzb=np.array([[0.0, 100.0], [600.0, 90.0], [1500.0, 100.0]]), rmax=1500.0)
    return env


@pytest.mark.parametrize("variant", VARIANTS)
@pytest.mark.parametrize("precision", ["c128", "c64"])
def test_multi_frequency_against_oracle(require_gpu, variant, precision):
    """Two frequencies share one grid (the default dr and dz of the highest
    frequency; the absorbing layer 20 lambda with lambda = c0 / max f).  The
    oracle input of each frequency is built here from that rule, not from
    the instance."""
    freqs = [75.0, 125.0]
    env = _multi_freq_env()
    m = CuPyRAM(**dict(env, freq=freqs), precision=precision, variant=variant, np=6)
    res = m.run()
    tl, tlg = res["TL Line"], res["TL Grid"]
    assert tl.shape[0] == 2 and tlg.shape[0] == 2
    fmax = max(freqs)
    c0 = 1500.0
    for i, f in enumerate(freqs):
        ray = O.resolve_ray(**dict(env, freq=f), np=6, dr=6 * 1500.0 / fmax, dz=0.1 * 1500.0 / fmax)
        ray.lam = c0 / fmax
        ray.zmax = float(np.max(env["z_sb"] + env["z_ss"][-1])) + 20 * ray.lam
        o = O.run(ray)
        s = envs.tl_stats(tl[i], o["TL Line"])
        sg = _stats(tlg[i], o["TL Grid"])
        print(f"multi-freq {f} Hz {precision} {variant}: line rms {s['rms']:.3g}, grid rms {sg['rms']:.3g} dB")
        assert s["rms"] <= TOL[precision], (f, s)
        assert sg["rms"] <= 2 * TOL[precision], (f, sg)
        _check_pressure(res["CP Line"][i], o["CP Line"], precision, f"CP Line {f}")
        _check_pressure(res["CP Grid"][i], o["CP Grid"], precision, f"CP Grid {f}")


@pytest.mark.parametrize("variant", VARIANTS)
@pytest.mark.parametrize("precision", ["c128", "c64"])
@pytest.mark.parametrize("ndr,ndz", [(3, 4), (1, 5), (4, 1)])
def test_decimated_grids_against_oracle(require_gpu, variant, precision, ndr, ndz):
    """The sloping wedge with outputs every ndr steps and every ndz rows to
    zmplt = 150 m: TL Line, TL Grid, CP Line and CP Grid vs the oracle."""
    kw = envs.wedge(rmax=4000.0)
    kw.update(zmplt=150.0, ndr=ndr, ndz=ndz)
    res = CuPyRAM(**kw, precision=precision, variant=variant).run()
    o = O.run(O.resolve_ray(**kw))
    assert res["TL Grid"].shape == o["TL Grid"].shape
    assert np.allclose(res["Depths"], o["Depths"]) and np.allclose(res["Ranges"], o["Ranges"])
    s = envs.tl_stats(res["TL Line"], o["TL Line"])
    sg = _stats(res["TL Grid"], o["TL Grid"])
    print(f"grids ndr {ndr} ndz {ndz} {precision} {variant}: line {s['rms']:.3g}, grid {sg['rms']:.3g} dB")
    assert s["rms"] <= TOL[precision] and sg["rms"] <= 2 * TOL[precision]
    _check_pressure(res["CP Line"], o["CP Line"], precision, "CP Line")
    _check_pressure(res["CP Grid"], o["CP Grid"], precision, "CP Grid")


@pytest.mark.parametrize("variant", VARIANTS)
@pytest.mark.parametrize("precision", ["c128", "c64"])
def test_stability_switch_against_oracle(require_gpu, variant, precision):
    """rs = 2000 m on the wedge: the march swaps to the Pade set without
    stability constraints (ns = 0) at the first step with r >= rs.  The
    result matches the oracle and differs from the run without rs."""
    kw = envs.wedge(rmax=4000.0)
    kw.update(zmplt=150.0)
    with_rs = CuPyRAM(**dict(kw, rs=2000.0), precision=precision, variant=variant).run()
    o = O.run(O.resolve_ray(**dict(kw, rs=2000.0)))
    s = envs.tl_stats(with_rs["TL Line"], o["TL Line"])
    sg = _stats(with_rs["TL Grid"], o["TL Grid"])
    print(f"rs switch {precision} {variant}: line {s['rms']:.3g}, grid {sg['rms']:.3g} dB")
    assert s["rms"] <= TOL[precision] and sg["rms"] <= 2 * TOL[precision]
    _check_pressure(with_rs["CP Grid"], o["CP Grid"], precision, "CP Grid")
    if precision == "c128":
        without = CuPyRAM(**kw, precision=precision, variant=variant).run()
        r = with_rs["Ranges"]
        before = r < 2000.0
        assert np.array_equal(with_rs["TL Line"][before], without["TL Line"][before])
        assert np.max(np.abs(with_rs["CP Grid"] - without["CP Grid"])) > 0


def test_reference_import_name_runs_the_golden_case(require_gpu):
    """The reference's own test form (tests/test_reference.py:15-28): import
    through the package name ``cupyram`` and match RAM v1.5's TL line."""
    import os
    from cupyram import CuPyRAM as RefName
    ref = np.loadtxt(os.path.join(os.path.dirname(__file__), "golden", "tl_ref.line"))
    res = RefName(**envs.ram_reference_case()).run()
    assert np.mean(np.abs(res["TL Line"] - ref[:, 1])) <= 0.01


def test_device_info_reports_the_device(require_gpu):
    from cupyram_amd import _abi
    name, free, total = _abi.device_info(0)
    print("device", name, free / 2**30, total / 2**30)
    assert name and 0 < free <= total


@pytest.mark.parametrize("precision,tol", [("c128", 1e-9), ("c64", 2e-3)])
def test_field_read_back_in_both_layouts(require_gpu, precision, tol):
    """cpr_read_field returns the march field u [ray][row] from the thread
    plan's depth-major columns and from the team plan's [ray][k][lane]
    layout alike: the two marches of the wedge agree to rounding (complex128)
    or within complex64's error, row by row, at every ray of the batch."""
    kw = dict(envs.varying_ranges(), compute_grids=False)
    fields = {}
    for variant in ("thread", "team"):
        m = CuPyRAM(**kw, precision=precision, variant=variant)
        m.setup()
        idx = np.arange(m._total_batch)
        plan = m._make_plan(idx, variant=variant)
        try:
            plan.selfs()
            plan.march(0, int(m._rays["nsteps"].max()))
            fields[variant] = plan.read_field(int(m._rays["nz"].max()))
        finally:
            plan.close()
    nz = m._rays["nz"]
    worst = 0.0
    for r in range(len(nz)):
        a, b = fields["thread"][r, 1:nz[r] + 1], fields["team"][r, 1:nz[r] + 1]
        worst = max(worst, float(np.max(np.abs(a - b)) / np.max(np.abs(a))))
        assert np.all(fields["team"][r, nz[r] + 1:] == 0)
    print(f"field {precision}: max relative difference thread vs team {worst:.3g}")
    assert worst <= tol
