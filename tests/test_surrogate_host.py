"""CPU tests of the surrogate's host side: DFT matrices, input encoding and
the PE set-ups of the query domains."""
import numpy as np
import pytest

from cupyram_amd import CuPyRAM
from cupyram_amd.surrogate import dataset as D
from cupyram_amd.surrogate.encode import FOURIER_K, N_FOURIER, N_GLOBAL, N_LOCAL, encode
from cupyram_amd.surrogate.fno import dft_matrices


def test_dft_matrices_match_numpy_rfft_irfft():
    Z, K = 64, 16
    ac, as_, sc, ss = dft_matrices(Z, K)
    rng = np.random.default_rng(0)
    x = rng.normal(size=(5, Z))
    X = np.fft.rfft(x, axis=1)[:, :K]
    np.testing.assert_allclose(x @ ac.T, X.real, atol=1e-12)
    np.testing.assert_allclose(x @ as_.T, X.imag, atol=1e-12)
    M = rng.normal(size=(5, K)) + 1j * rng.normal(size=(5, K))
    full = np.zeros((5, Z // 2 + 1), complex)
    full[:, :K] = M
    y = np.fft.irfft(full, n=Z, axis=1)
    np.testing.assert_allclose(M.real @ sc + M.imag @ ss, y, atol=1e-12)


def test_dft_round_trip_of_band_limited_column():
    Z, K = 64, 16
    ac, as_, sc, ss = dft_matrices(Z, K)
    z = np.arange(Z)
    x = 1.0 + np.cos(2 * np.pi * 3 * z / Z) - 0.5 * np.sin(2 * np.pi * 7 * z / Z)
    re, im = ac @ x, as_ @ x
    np.testing.assert_allclose(re @ sc + im @ ss, x, atol=1e-12)


@pytest.mark.parametrize("name", ["deep_munk", "wedge"])
def test_domain_march_grid_is_the_output_grid(name):
    dom = D.DOMAINS[name]
    gen = D.deep_queries if name == "deep_munk" else D.wedge_queries
    q = gen(40, 7)
    m = CuPyRAM(**D.march_kwargs(dom, q))
    m.setup()
    vr = m.vr if isinstance(m.vr, list) else [m.vr] * len(q)
    vz = m.vz if isinstance(m.vz, list) else [m.vz] * len(q)
    for r, z in zip(vr, vz):
        np.testing.assert_allclose(r, dom.ranges(), rtol=1e-12)
        np.testing.assert_allclose(z, dom.depths(), rtol=1e-12)
    assert np.all(m._rays["nvr"] == dom.nr) and np.all(m._rays["nvz"] == dom.nzp)


def test_generators_are_seeded():
    a, b, c = D.wedge_queries(10, 5), D.wedge_queries(10, 5), D.wedge_queries(10, 6)
    np.testing.assert_array_equal(a.vector(), b.vector())
    assert not np.array_equal(a.vector(), c.vector())
    assert a.vector().shape == (10, N_GLOBAL)
    d = D.deep_queries(1000, 2)
    assert 80 <= d.freq.min() and d.freq.max() <= 125
    assert 3500 <= d.bathy.min() and d.bathy.max() <= 5000


def test_wedge_bottom_profiles_follow_the_seabed():
    q = D.wedge_queries(3, 1)
    kw = D.march_kwargs(D.WEDGE, q)
    for i in range(3):
        h = q.bottom[i, 0]
        z = kw["z_sb"][i] + D.WEDGE.ssp_depth
        rp = kw["rp_sb"][i]
        for p in range(len(rp)):
            mid = min(rp[p] + 250.0, 10000.0)
            seabed = 200.0 - 180.0 * mid / 10000.0
            col = kw["cb"][i][:, p]
            # sediment speed down to seabed + thickness, basement below the 0.5 m transition
            assert np.all(col[z <= seabed + h + 1e-9] == q.bottom[i, 1])
            assert np.all(col[z >= seabed + h + 0.5 - 1e-9] == 2000.0)


def test_encode_local_features():
    torch = pytest.importorskip("torch")
    dom = D.WEDGE
    q = D.wedge_queries(2, 3)
    x = encode(dom, q, device="cpu").to(torch.float32).numpy().reshape(2, dom.nr, dom.nzp, -1)
    assert x.shape[-1] == 128
    r, z = dom.ranges(), dom.depths()
    zb = np.interp(r, np.linspace(0, dom.rmax, 64), q.bathy[0])
    below = (z[None, :] > zb[:, None]).astype(np.float32)
    clear = np.abs(z[None, :] - zb[:, None]) > 1e-3 * dom.zplot   # points on the seabed round either way
    np.testing.assert_array_equal(x[0, :, :, N_GLOBAL + 5][clear], below[clear])
    np.testing.assert_allclose(x[0, :, 0, N_GLOBAL + 0], r / dom.rmax, rtol=1e-2)
    np.testing.assert_allclose(x[0, 0, :, N_GLOBAL + 1], z / dom.zplot, rtol=1e-2)
    o = N_GLOBAL + N_LOCAL
    np.testing.assert_allclose(x[0, :, 0, o], np.log(r / dom.rmax), rtol=1e-2, atol=1e-2)
    for i, kf in enumerate(FOURIER_K):
        np.testing.assert_allclose(x[0, :, 0, o + 1 + 4 * i], np.sin(np.pi * kf * r / dom.rmax), atol=1e-2)
        np.testing.assert_allclose(x[0, 0, :, o + 4 + 4 * i], np.cos(np.pi * kf * z / dom.zplot), atol=1e-2)
    assert np.all(x[..., o + N_FOURIER:] == 0)
    # global part is the same for every point of a query and differs between queries
    assert np.all(x[0, :, :, :N_GLOBAL] == x[0, :1, :1, :N_GLOBAL])
    assert not np.array_equal(x[0, 0, 0, :N_GLOBAL], x[1, 0, 0, :N_GLOBAL])


def _encode_by_slices(dom, q, features=128):
    """The input encoding written feature by feature into a float32 tensor
    and cast once (the straightforward form of encode)."""
    import torch
    from cupyram_amd.surrogate.encode import N_BATHY, N_SSP, global_features
    n = len(q)
    t = lambda a: torch.as_tensor(np.asarray(a, np.float32))  # noqa: E731
    g = global_features(dom, q, "cpu")
    r, z = t(dom.ranges() / dom.rmax), t(dom.depths() / dom.zplot)
    pos = r * (N_BATHY - 1)
    i0 = torch.clamp(pos.floor().long(), 0, N_BATHY - 2)
    w = (pos - i0.float())[None, :]
    bz = g[:, N_SSP:N_SSP + N_BATHY]
    zb = bz[:, i0] * (1 - w) + bz[:, i0 + 1] * w
    zsn = g[:, N_GLOBAL - 1]
    out = torch.zeros(n, dom.nr, dom.nzp, features)
    out[..., :N_GLOBAL] = g[:, None, None, :]
    out[..., N_GLOBAL + 0] = r[None, :, None]
    out[..., N_GLOBAL + 1] = z[None, None, :]
    out[..., N_GLOBAL + 2] = zb[:, :, None]
    out[..., N_GLOBAL + 3] = z[None, None, :] - zb[:, :, None]
    out[..., N_GLOBAL + 4] = z[None, None, :] - zsn[:, None, None]
    out[..., N_GLOBAL + 5] = (z[None, None, :] > zb[:, :, None]).float()
    out[..., N_GLOBAL + 6] = 1.0
    o = N_GLOBAL + N_LOCAL
    out[..., o] = torch.log(r)[None, :, None]
    for i, kf in enumerate(FOURIER_K):
        out[..., o + 1 + 4 * i] = torch.sin(np.pi * kf * r)[None, :, None]
        out[..., o + 2 + 4 * i] = torch.cos(np.pi * kf * r)[None, :, None]
        out[..., o + 3 + 4 * i] = torch.sin(np.pi * kf * z)[None, None, :]
        out[..., o + 4 + 4 * i] = torch.cos(np.pi * kf * z)[None, None, :]
    return out.to(torch.bfloat16).reshape(n * dom.nr, dom.nzp, features)


@pytest.mark.parametrize("name", ["wedge", "deep_munk"])
def test_encode_equals_feature_by_feature_form(name):
    """encode() assembles the bf16 input from per-feature pieces in one
    concatenation; it must equal, bit for bit, the input written feature by
    feature in float32 and cast once."""
    torch = pytest.importorskip("torch")
    dom = D.DOMAINS[name]
    gen = D.wedge_queries if name == "wedge" else D.deep_queries
    q = gen(4, 9)
    assert torch.equal(encode(dom, q, device="cpu"), _encode_by_slices(dom, q))



@pytest.mark.parametrize("name", ["wedge", "deep_munk"])
def test_lift_inputs_reproduce_encode(name):
    """The tables the structured first layer reads (lift_inputs) hold the
    encoded columns: rebuilt column by column from them as their docstring
    states (query numbers; range and depth tables; the water depth in its
    own column, subtracted in the height column; the source depth subtracted
    in the offset column; the flag zv > zb) and cast once, they equal
    encode() bit for bit."""
    torch = pytest.importorskip("torch")
    import importlib
    E = importlib.import_module("cupyram_amd.surrogate.encode")
    dom = D.DOMAINS[name]
    gen = D.wedge_queries if name == "wedge" else D.deep_queries
    q = gen(3, 11)
    inp = E.lift_inputs(dom, q, device="cpu")
    n, nr, nz = len(q), dom.nr, dom.nzp
    zb = inp.zb.reshape(n, nr)
    out = torch.zeros(n, nr, nz, 128)
    out[..., :E.N_GLOBAL] = inp.g[:, None, None, :]
    for j, row in enumerate(inp.fr_rows):
        out[..., row] = inp.fr[None, :, None, j]
    for j, row in enumerate(inp.fz_rows):
        out[..., row] = inp.fz[None, None, :, j]
    out[..., E.F_HEIGHT] = out[..., E.F_HEIGHT] - zb[:, :, None]
    out[..., E.F_ZREL] = out[..., E.F_ZREL] - inp.g[:, None, None, E.N_GLOBAL - 1]
    out[..., E.F_ZB] = zb[:, :, None]
    out[..., E.F_BELOW] = (inp.zv[None, None, :] > zb[:, :, None]).float()
    assert sorted(inp.fr_rows + inp.fz_rows + [E.F_ZB, E.F_BELOW]) == list(range(E.N_GLOBAL, 125))
    assert torch.equal(out.to(torch.bfloat16).reshape(n * nr, nz, 128), encode(dom, q, device="cpu"))
