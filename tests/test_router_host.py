"""CPU tests of the surrogate router: forced ensemble disagreement and forced
surrogate bias above 3 dB must route queries to the PE march (here a stand-in
PE whose answer is known), and accepted queries keep the ensemble mean."""
import types

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from cupyram_amd.surrogate import dataset as D  # noqa: E402
from cupyram_amd.surrogate.encode import global_features  # noqa: E402
from cupyram_amd.surrogate.router import Router  # noqa: E402

DOM = D.WEDGE
FREQ = 99          # column of the normalised log-frequency in the query vector


class Member:
    """Predicts normalised TL = (log-frequency feature) + offset(query)."""

    def __init__(self, offset=0.0, only_above=None):
        self.cfg = types.SimpleNamespace(features=128)
        self.offset, self.only_above = offset, only_above

    def forward(self, x, train=False):
        f = x[..., FREQ].float()                       # [n*nr, nzp]
        off = torch.full_like(f, self.offset)
        if self.only_above is not None:
            off = torch.where(f > self.only_above, off, torch.zeros_like(off))
        return (f + off).reshape(-1)


def truth(q):
    f = global_features(DOM, q, device="cpu")[:, FREQ].to(torch.bfloat16).float()
    return (DOM.tl_mean + DOM.tl_std * f)[:, None, None].expand(len(q), DOM.nr, DOM.nzp).clone()


def make_pe(bias_db=0.0, calls=None):
    def pe(dom, q, precision, log=None):
        if calls is not None:
            calls.append(len(q))
        return truth(q) - bias_db, 0.0, 0.0
    return pe


def test_agreeing_ensemble_is_accepted_and_spot_checked():
    q = D.wedge_queries(300, 11)
    calls = []
    r = Router(DOM, [Member() for _ in range(4)], pe=make_pe(0.5, calls), device="cpu", batch=64)
    res = r.run(q)
    assert not res.routed.any() and not res.batch_fallback
    assert len(res.spot) == 3 and calls == [3]                 # 1 % of the queries
    np.testing.assert_allclose(res.spot_err_db, 0.5, atol=0.02)
    others = np.setdiff1d(np.arange(300), res.spot)
    t = truth(q)
    assert torch.allclose(res.tl[others], t[others], atol=0.1)
    assert torch.allclose(res.tl[res.spot], t[res.spot] - 0.5, atol=1e-6)


def test_disagreement_above_3db_routes_to_pe():
    q = D.wedge_queries(300, 12)
    f = global_features(DOM, q, device="cpu")[:, FREQ].numpy()
    # one member 10 dB off on the upper half of the band: std of {0,0,0,10} = 4.3 dB
    members = [Member(), Member(), Member(), Member(offset=0.5, only_above=0.5)]
    r = Router(DOM, members, pe=make_pe(0.0), device="cpu", batch=64)
    res = r.run(q)
    hi = f > 0.5 + 0.01
    lo = f < 0.5 - 0.01
    assert res.routed[hi].all() and not res.routed[lo].any()
    assert np.all(res.spread_db[hi] > 3.0) and np.all(res.spread_db[lo] < 1e-3)
    t = truth(q)
    assert torch.allclose(res.tl[torch.as_tensor(np.nonzero(hi)[0])], t[torch.as_tensor(np.nonzero(hi)[0])])
    # the spot check judges the ensemble only where it would have answered
    assert len(res.spot_err_db) == int(np.sum(res.spread_db[res.spot] <= 3.0))
    assert not res.batch_fallback


def test_biased_ensemble_fails_the_spot_check_and_the_batch_falls_back():
    q = D.wedge_queries(400, 13)
    calls = []
    # every member agrees on an answer 5 dB too high: no disagreement, but the PE judges it wrong
    r = Router(DOM, [Member(offset=0.25) for _ in range(4)], pe=make_pe(0.0, calls), device="cpu", batch=64)
    res = r.run(q)
    assert res.batch_fallback and res.routed.all()
    assert np.all(res.spot_err_db > 3.0)
    assert sum(calls) == 400 and calls[0] == 4                 # the spot sample first, then the rest
    assert np.isnan(res.spread_db).sum() == 396                # the ensemble skipped the rest
    assert torch.allclose(res.tl, truth(q))


def test_router_rejects_bad_configuration():
    with pytest.raises(ValueError):
        Router(DOM, [], device="cpu")
    with pytest.raises(ValueError):
        Router(DOM, [Member()], spot_fraction=1.5, device="cpu")


def test_no_spot_check_and_single_query():
    """spot_fraction 0: the ensemble answers everything and the march is never
    called; a one-query batch is its own spot sample."""
    q = D.wedge_queries(50, 14)
    calls = []
    r = Router(DOM, [Member() for _ in range(4)], spot_fraction=0.0, pe=make_pe(0.0, calls), device="cpu", batch=64)
    res = r.run(q)
    assert calls == [] and len(res.spot) == 0 and not res.routed.any() and not res.batch_fallback
    assert np.all(np.isfinite(res.spread_db))
    assert torch.allclose(res.tl, truth(q), atol=0.1)
    one = q.take(np.array([3]))
    calls = []
    r = Router(DOM, [Member() for _ in range(4)], pe=make_pe(0.5, calls), device="cpu", batch=64)
    res = r.run(one)
    assert calls == [1] and list(res.spot) == [0] and not res.batch_fallback
    assert torch.allclose(res.tl, truth(one) - 0.5, atol=1e-6)


class NaNMember(Member):
    """Returns NaN where the normalised log-frequency lies above 0.5."""

    def forward(self, x, train=False):
        y = super().forward(x, train)
        f = x[..., FREQ].float().reshape(-1)
        return torch.where(f > 0.5, torch.full_like(y, float("nan")), y)


def test_nan_member_routes_its_queries_to_pe():
    """A member that returns NaN makes the ensemble spread NaN on those
    queries: they count as disagreement (routed to the PE march), never as
    accepted surrogate answers, in the spot sample and in the rest."""
    q = D.wedge_queries(400, 15)
    f = global_features(DOM, q, device="cpu")[:, FREQ].numpy()
    members = [Member(), Member(), Member(), NaNMember()]
    r = Router(DOM, members, spot_fraction=0.05, pe=make_pe(0.0), device="cpu", batch=64)
    res = r.run(q)
    hi = f > 0.5 + 0.01
    lo = f < 0.5 - 0.01
    assert hi.sum() > 50 and lo.sum() > 50
    assert res.routed[hi].all() and not res.routed[lo].any()
    assert not res.batch_fallback
    assert torch.isfinite(res.tl).all()
    t = truth(q)
    assert torch.allclose(res.tl[torch.as_tensor(np.nonzero(hi)[0])], t[torch.as_tensor(np.nonzero(hi)[0])])
    assert len(np.intersect1d(res.spot, np.nonzero(hi)[0])) > 0          # the spot sample saw NaN answers too


def test_nan_spot_error_falls_back():
    """Every member NaN: the spot check cannot judge the ensemble, so no
    surrogate answer survives (all queries answered by the PE)."""
    q = D.wedge_queries(200, 16)

    class AllNaN(Member):
        def forward(self, x, train=False):
            return torch.full_like(super().forward(x, train), float("nan"))

    r = Router(DOM, [AllNaN() for _ in range(4)], pe=make_pe(0.0), device="cpu", batch=64)
    res = r.run(q)
    assert res.routed.all()
    assert torch.isfinite(res.tl).all()
    assert torch.allclose(res.tl, truth(q))
