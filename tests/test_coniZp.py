"""coniZp (every coni PFV of a direction, at any dilation) and its dilation
bound, as the reference's tests/test_dilation_bound.py and
tests/test_coni_zpk.py:172-215 check them.

The bound delta < Q / mu0 is pinned by the known PFVs (the 332 of the
dataset geometries and the paper examples each lie below their direction's
bound), by the definition of mu0 (a nonzero K in Lambda attains it, the
value computed from scratch with the oracle's exact form), by brute force
over a box at low h11, by basis invariance, and by the exact path agreeing
with the library.  coniZp equals coniZpM run at each direction's own bound
(host here, the oracle's points and post-processing on the GPU)."""

import math
from fractions import Fraction

import numpy as np
import pytest

from helpers import as_set, cydata_of, geometries, oracle_pfvs
from oracle import exact
from pfvs_amd import coni, dilation, gpu
from pfvs_amd.cydata import CYData
from pfvs_amd.pvectors import box_enum
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))
PAPER = json.load(open(os.path.join(HERE, "golden", "paper_coni.json")))
GEOMS = geometries()


def _paper_data(e):
    return CYData(e["h21"], e["kappa"], e["c2"], e["H"], coni_curve=e["coni_curve"],
                  coni_cob=e["cob"])


def _known_pfvs():
    """(kappa (coni basis), Q, p_hat, delta) of every dataset and paper PFV."""
    out = []
    for g in GEOMS:
        if not g["K"]:
            continue
        kap = cydata_of(g).kappa_cob
        Q = g["h11"] + g["h21"] + 4
        for K, M in zip(g["K"], g["M"]):
            out.append((kap, Q, *exact.direction_and_delta(kap.tolist(), K, M)))
    for e in PAPER:
        kap = _paper_data(e).kappa_cob
        Q = -int(np.dot(e["K"], e["M"]))
        out.append((kap, Q, *exact.direction_and_delta(kap.tolist(), e["K"], e["M"])))
    return out


KNOWN = _known_pfvs()


def test_known_pfvs_cover_h11_3_to_9():
    assert len(KNOWN) > 300
    assert {k.shape[0] for k, *_ in KNOWN} >= set(range(3, 10))


def test_bound_holds_for_every_known_pfv():
    """delta < Q / mu0 for every known PFV, and the hypotheses hold (the
    bound exists) for every known direction."""
    gpu.require()
    for kap, Q, p_hat, delta in KNOWN:
        b = dilation.coni_dilation_bound(p_hat[1:], kap, Q)
        assert b is not None, p_hat
        assert isinstance(b, Fraction) and delta < b, (p_hat, delta, b)


def test_mu0_is_attained_in_lambda():
    """mu0 is K^T S K at the returned K, which is nonzero and orthogonal to
    p_r; the value computed from the oracle's exact S = -A_rr^-1."""
    gpu.require()
    for kap, _Q, p_hat, _ in KNOWN[::5]:
        mu0, K = dilation.coni_mu0(p_hat[1:], kap)
        assert any(K) and sum(k * int(p) for k, p in zip(K, p_hat[1:])) == 0
        S, s = exact.schur_form(kap.tolist(), p_hat[1:])
        assert s <= 0 and exact.form_value(S, K) == mu0


@pytest.mark.parametrize("h11", [3, 4, 5])
def test_mu0_is_the_box_minimum_at_low_h11(h11):
    """The minimum of K^T S K over a box of Lambda (brute force, the box
    large enough here) equals mu0."""
    gpu.require()
    cases = [c for c in KNOWN if c[0].shape[0] == h11][:6]
    assert cases
    for kap, _Q, p_hat, _ in cases:
        mu0, _ = dilation.coni_mu0(p_hat[1:], kap)
        assert exact.mu0_box(kap.tolist(), p_hat[1:], 6) == mu0


def test_library_and_exact_path_agree():
    """The library's mu0 (host threads) and the exact rational path agree on
    the known directions and on random ones (which often have no bound)."""
    gpu.require()
    rng = np.random.default_rng(3)
    decided = bounded = 0
    for kap, _Q, p_hat, _ in KNOWN[::9]:
        h = kap.shape[0]
        ps = np.vstack([p_hat[1:], rng.integers(-5, 6, size=(6, h - 1))])
        ps = ps[np.any(ps != 0, axis=1)]
        st, num, den, K = dilation._mu0_lib(ps, kap)
        for i, p in enumerate(ps):
            ref = dilation._mu0_exact(p, kap)
            if st[i] < 0:
                continue
            decided += 1
            if st[i] == 1:
                assert ref is None, p
                continue
            bounded += 1
            assert ref is not None and ref[0] == Fraction(int(num[i]), int(den[i])), p
    assert decided > 100 and bounded > 30


def test_library_decides_large_forms():
    """Directions whose form L S has minors beyond int128 (h11 = 7 dataset
    geometries, where the exact positive-definiteness test of the reduced
    form overflows and the proven FP64 factor decides): the library decides
    every one, and its mu0 equals the exact rational path's."""
    gpu.require()
    n = 0
    for data, Q, ps in _cases(10, 7, 7, step=2):
        kap = data.kappa_cob
        st, num, den, K = dilation._mu0_lib(ps, kap)
        assert np.all(st == 0)
        for i in range(0, len(ps), 2):
            ref = dilation._mu0_exact(ps[i], kap)
            assert ref is not None and ref[0] == Fraction(int(num[i]), int(den[i])), ps[i]
            n += 1
    assert n >= 10


def test_basis_change_leaves_mu0():
    """A signed permutation of the non-coni directions applied to kappa and
    p_hat leaves mu0 unchanged."""
    gpu.require()
    rng = np.random.default_rng(0)
    for kap, _Q, p_hat, _ in KNOWN[::25]:
        h = kap.shape[0]
        perm = rng.permutation(h - 1)
        signs = rng.choice([-1, 1], size=h - 1)
        U = np.eye(h, dtype=np.int64)
        U[1:, 1:] = 0
        U[1 + np.arange(h - 1), 1 + perm] = signs
        Ui = np.round(np.linalg.inv(U)).astype(np.int64)
        kap2 = np.einsum("ia,jb,kc,ijk->abc", Ui, Ui, Ui, kap)
        p2 = U @ np.asarray(p_hat, dtype=np.int64)
        assert dilation.coni_mu0(p2[1:], kap2)[0] == dilation.coni_mu0(p_hat[1:], kap)[0]


def test_no_bound_where_the_hypotheses_fail():
    gpu.require()
    h = 4
    assert dilation.coni_dilation_bound([1, 2, 3], np.zeros((h, h, h), np.int64), 100) is None
    kap = np.zeros((h, h, h), dtype=np.int64)             # A_rr = I: S = -I on Lambda
    for i in (1, 2, 3):
        for a, b, c in {(i, i, 1), (i, 1, i), (1, i, i)}:
            kap[a, b, c] = 1
    assert dilation.coni_dilation_bound([1, 0, 0], kap, 100) is None
    assert dilation._mu0_exact([1, 0, 0], kap) is None


def test_bound_ceils_and_threads():
    gpu.require()
    kap, Q, p_hat, _ = KNOWN[40]
    h = kap.shape[0]
    ps = np.vstack([p_hat[1:], np.random.default_rng(4).integers(-5, 6, size=(12, h - 1))])
    ps = ps[np.any(ps != 0, axis=1)]
    want = [0 if b is None else math.ceil(b) for b in dilation.coni_dilation_bounds(ps, kap, Q)]
    for n_jobs in (1, 3):
        assert dilation.coni_dilation_bound_ceils(ps, kap, Q, n_jobs=n_jobs).tolist() == want


def test_grid_up():
    D = np.array([1, 2, 3, 100, 799, 800, 801, 12345, 10**6])
    g = coni._grid_up(D)
    assert np.all(g >= D) and np.all(g <= np.ceil(D * 2 ** (1 / 16)) + 1)


def _cases(n_dirs, lo, hi, step=1):
    """(data, Q, ps): dataset geometries with lo <= h11 <= hi, directions of
    their box (the recorded B, capped at 3) that have a bound."""
    out = []
    for g in GEOMS[::step]:
        if not lo <= g["h11"] <= hi:
            continue
        data = cydata_of(g)
        Q = g["h11"] + g["h21"] + 4
        ps = box_enum(min(g["B"], 3), data.H_cob, device="cpu")[:n_dirs]
        if not len(ps):
            continue
        b = dilation.coni_dilation_bounds(ps, data.kappa_cob, Q)
        ps = ps[[x is not None for x in b]]
        if len(ps):
            out.append((data, Q, ps))
    return out


def test_coniZp_is_zpm_at_each_bound():
    """coniZp (host) returns exactly the union over the directions of
    coniZpM at ceil(bound), each PFV once, all complete."""
    gpu.require()
    n = 0
    for data, Q, ps in _cases(10, 3, 7, step=2):
        Ks, Ms, complete = coni.coniZp(data, ps, Q=Q, device="cpu", n_jobs=2)
        assert complete.all()
        want = set()
        for p, b in zip(ps, dilation.coni_dilation_bounds(ps, data.kappa_cob, Q)):
            K1, M1 = coni.coniZpM(data, p[None], Q=Q, ellipsoid_dilation=math.ceil(b), device="cpu")
            want |= as_set(K1, M1)
        assert as_set(Ks, Ms) == want and len(Ks) == len(want)
        n += len(want)
    assert n > 20


def test_coniZp_order_and_fallback(monkeypatch):
    """A direction without a bound is searched at fallback_dilation and
    reported incomplete; rows come sorted by p-vector, then M, then K."""
    gpu.require()
    data, Q, ps = _cases(10, 5, 5)[1]
    real = dilation.coni_dilation_bound_ceils

    def without_first(ps_, kap, Q_, n_jobs=1):
        b = real(ps_, kap, Q_, n_jobs)
        b[(ps_ == ps[0]).all(axis=1)] = 0
        return b
    monkeypatch.setattr(dilation, "coni_dilation_bound_ceils", without_first)
    Ks, Ms, complete = coni.coniZp(data, ps, Q=Q, fallback_dilation=30, device="cpu")
    assert not complete[0] and complete[1:].all()
    want = set()
    for i, (p, b) in enumerate(zip(ps, real(ps, data.kappa_cob, Q))):
        K1, M1 = coni.coniZpM(data, p[None], Q=Q, ellipsoid_dilation=30 if i == 0 else int(b),
                              device="cpu")
        want |= as_set(K1, M1)
    assert as_set(Ks, Ms) == want and len(Ks) == len(want)
    with pytest.raises(ValueError):
        coni.coniZp(data, ps, Q=Q, fallback_dilation=0, device="cpu")


@pytest.mark.parametrize("e", PAPER[3:5], ids=[e["name"].replace(" ", "_") for e in PAPER[3:5]])
def test_search_past_the_bound_finds_nothing_more(e):
    """For a paper direction: coniZpM at the bound and at twice the bound
    plus 20 find the same PFVs, each below the bound (the paper's PFV
    among them)."""
    gpu.require()
    data = _paper_data(e)
    Q = -int(np.dot(e["K"], e["M"]))
    p = np.asarray(e["p"])[1:]
    b = dilation.coni_dilation_bound(p, data.kappa_cob, Q)
    at = coni.coniZpM(data, p[None], Q=Q, ellipsoid_dilation=float(b), device="cpu")
    big = coni.coniZpM(data, p[None], Q=Q, ellipsoid_dilation=float(2 * b + 20), device="cpu")
    assert as_set(*at) == as_set(*big)
    assert (tuple(e["K"]), tuple(e["M"])) in as_set(*at)
    for K, M in zip(*big):
        ph, delta = exact.direction_and_delta(data.kappa_cob.tolist(), K, M)
        assert list(ph[1:]) == list(map(int, p)) and delta < b


@pytest.mark.gpu
def test_gpu_coniZp_matches_the_oracle(need_gpu):
    """coniZp on the GPU (one batched launch over every direction at its
    own bound) against the oracle's points and exact post-processing at
    each direction's ceil(bound), on dataset geometries h11 3..9 and a
    paper example's scan sample."""
    n = 0
    cases = _cases(16, 3, 9, step=3)
    e = PAPER[4]
    data = _paper_data(e)
    ps = box_enum(e["scan"]["B"], data.H_cob, device="cpu")[::7]
    cases.append((data, -int(np.dot(e["K"], e["M"])), ps))
    for data, Q, ps in cases:
        print(f"h11 {data.h11}: {len(ps)} directions", flush=True)
        Ks, Ms, complete = coni.coniZp(data, ps, Q=Q, device="gpu")
        assert complete.all()
        want = set()
        for p, b in zip(ps, dilation.coni_dilation_bounds(ps, data.kappa_cob, Q)):
            want |= as_set(*oracle_pfvs(data, p[None], Q, math.ceil(b)))
        assert as_set(Ks, Ms) == want and len(Ks) == len(want)
        Kc, Mc, _ = coni.coniZp(data, ps, Q=Q, device="cpu")
        assert np.array_equal(Ks, Kc) and np.array_equal(Ms, Mc)
        n += len(want)
    assert n > 20
