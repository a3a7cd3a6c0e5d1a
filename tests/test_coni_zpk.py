"""coniZpK (the coni PFVs of a direction above a split dilation D0) and the
cost-model routing of coniZp, as the reference's tests/test_coni_zpk.py
checks them:

- coniZpK(D0) together with coniZpM at D0 gives exactly coniZpM at the
  direction's dilation bound (the set every coni PFV of the direction lies
  in), and every PFV coniZpK returns has dilation > D0 by the oracle's
  exact direction_and_delta (an independent computation of delta from
  (K, M));
- the library's short-vector enumeration decides most directions itself
  (status 0), so the first check is not carried by the fallback;
- a direction without a bound raises IncompleteSearchError;
- CostModel interpolates as the reference's (its known values), and
  bound_routing's choice is the brute-force optimum over every threshold;
- coniZp with a cost model that splits directions returns the same PFVs
  as coniZp searching every direction to its bound (host; on the GPU in
  the gpu-marked test);
- the library's points do not depend on its thread count, and a direction
  alone gives its slice of the batch."""

import math

import numpy as np
import pytest

from helpers import as_set
from oracle import exact
from pfvs_amd import coni, dilation, gpu
from pfvs_amd.dilation import CostModel, bound_routing
from pfvs_amd.util import IncompleteSearchError
from test_coniZp import _cases

CASES = _cases(8, 3, 6, step=2)


@pytest.mark.parametrize("D0", [2, 5, 12])
def test_zpk_completes_zpm(D0):
    gpu.require()
    n_above = n_dirs = n_zpk = 0
    for data, Q, ps in CASES:
        kap = data.kappa_cob
        p_full = np.hstack([np.zeros((len(ps), 1), np.int64), ps.astype(np.int64)])
        st = gpu.coni_zpk_points(kap, data.M_lattice(), p_full, Q, 13, D0, threads=2)[4]
        n_zpk += int(np.sum(st == 0))
        Kk, Mk = coni.coniZpK(data, ps, D0, Q=Q, n_jobs=2)
        above = as_set(Kk, Mk)
        assert len(above) == len(Kk)
        full, low = set(), set()
        for p, b in zip(ps, dilation.coni_dilation_bounds(ps, kap, Q)):
            full |= as_set(*coni.coniZpM(data, p[None], Q=Q, ellipsoid_dilation=max(math.ceil(b), D0),
                                         device="cpu"))
            low |= as_set(*coni.coniZpM(data, p[None], Q=Q, ellipsoid_dilation=D0, device="cpu"))
        assert above | low == full, (data.h11, D0)
        assert above <= full
        for K, M in above:
            _, delta = exact.direction_and_delta(kap.tolist(), list(K), list(M))
            assert delta > D0, (K, M, delta)
        n_above += len(above)
        n_dirs += len(ps)
    assert n_zpk >= 0.8 * n_dirs
    assert n_above >= 10


def test_coniZpK_raises_without_a_bound(monkeypatch):
    gpu.require()
    data, Q, ps = CASES[1]
    real = gpu.coni_zpk_points

    def first_undecided(*a, **k):
        M, Kn, q, pidx, st = real(*a, **k)
        keep = pidx != 0
        st = st.copy()
        st[0] = -1
        return M[keep], Kn[keep], q[keep], pidx[keep], st
    monkeypatch.setattr(gpu, "coni_zpk_points", first_undecided)
    monkeypatch.setattr(dilation, "coni_dilation_bounds", lambda ps_, kap, Q_, n_jobs=1: [None] * len(ps_))
    with pytest.raises(IncompleteSearchError):
        coni.coniZpK(data, ps, 3, Q=Q)
    with pytest.raises(ValueError):
        coni.coniZpK(data, ps, 0, Q=Q)


def test_cost_model():
    m = CostModel([100, 400, 1600], [1.0, 2.0, 8.0], [400, 800], [3.0, 2.0], c=0.5)
    assert m.G_at(400) == pytest.approx(2.0)
    assert m.G_at(200) == pytest.approx(2 ** 0.5)               # log-log
    assert m.G_at(6400) == pytest.approx(32.0)                  # the last slope
    assert m.G_at(25) == pytest.approx(0.5)                     # the first slope
    assert m.K_at(800) == 2.0
    with pytest.raises(ValueError):
        CostModel([400, 100], [1, 2], [400], [1])


@pytest.mark.parametrize("shared", [False, True])
def test_bound_routing_is_the_brute_force_optimum(shared):
    rng = np.random.default_rng(3)
    b = np.exp(rng.uniform(0, 8, 40))
    G = lambda D: 0.01 * np.asarray(D, dtype=float) ** 1.3
    K = lambda D0: {50: 4.0, 200: 2.5, 800: 1.0}[int(D0)]
    b_star, D0, t = bound_routing(b, G, K, [50, 200, 800], c=0.1, shared=shared)
    best = math.inf
    for d0 in (50, 200, 800):
        for thr in [-math.inf] + sorted(b):
            to_bound = b <= thr
            t_zpm = (np.sum(G(b[to_bound])) + np.sum(~to_bound) * G(d0)) / len(b)
            t_zpk = np.sum(~to_bound) * K(d0) / len(b) + 0.1
            best = min(best, t_zpm + t_zpk if shared else max(t_zpm, t_zpk))
    assert t == pytest.approx(best)


def _forced_split(D0):
    """A cost model under which every direction with a bound above D0 is
    split at D0 (ZpM steep, ZpK free)."""
    return CostModel([1, 10, 100], [1e-6, 1e-2, 1e2], [D0], [1e-12])


def test_coniZp_split_equals_search_to_the_bound(monkeypatch):
    gpu.require()
    calls = []
    real = gpu.coni_zpk_points

    def counted(*a, **k):
        r = real(*a, **k)
        calls.append(len(r[4]))
        return r
    monkeypatch.setattr(gpu, "coni_zpk_points", counted)
    n = 0
    for data, Q, ps in CASES:
        Ks, Ms, complete = coni.coniZp(data, ps, Q=Q, device="cpu", n_jobs=2, cost_model=_forced_split(3))
        K0, M0, c0 = coni.coniZp(data, ps, Q=Q, device="cpu", n_jobs=2)
        assert np.array_equal(complete, c0)
        assert as_set(Ks, Ms) == as_set(K0, M0) and len(Ks) == len(K0)
        assert np.array_equal(Ks, K0) and np.array_equal(Ms, M0)
        n += len(K0)
    assert sum(calls) >= 20 and n >= 20


@pytest.mark.gpu
def test_gpu_coniZp_split_matches_the_host(need_gpu):
    """coniZp on the GPU with directions split at D0 (coniZpM to D0 in the
    batched launch, coniZpK beside it) against coniZp on the host searching
    every direction to its bound; also with a calibrated cost model."""
    n = 0
    for data, Q, ps in _cases(16, 3, 7, step=2):
        print(f"h11 {data.h11}: {len(ps)} directions", flush=True)
        Kc, Mc, cc = coni.coniZp(data, ps, Q=Q, device="cpu")
        for cm in (_forced_split(3), "calibrate"):
            Ks, Ms, complete = coni.coniZp(data, ps, Q=Q, device="gpu", cost_model=cm)
            assert np.array_equal(complete, cc)
            assert np.array_equal(Ks, Kc) and np.array_equal(Ms, Mc)
        n += len(Kc)
    assert n > 20


def test_zpk_points_do_not_depend_on_threads():
    """pfa_coni_zpk gathers each direction's points and concatenates them in
    direction order: the same arrays on 1 and 4 host threads, and a
    direction alone gives its own slice of the batch."""
    gpu.require()
    data, Q, ps = CASES[2]
    p_full = np.hstack([np.zeros((len(ps), 1), np.int64), ps.astype(np.int64)])
    one = gpu.coni_zpk_points(data.kappa_cob, data.M_lattice(), p_full, Q, 13, 2, threads=1)
    four = gpu.coni_zpk_points(data.kappa_cob, data.M_lattice(), p_full, Q, 13, 2, threads=4)
    for a, b in zip(one, four):
        assert np.array_equal(a, b)
    M, Kn, q, pidx, st = one
    assert np.all(np.diff(pidx) >= 0) and len(q) > 0
    k = int(pidx[len(pidx) // 2])
    Ma, Kna, qa, pa, sa = gpu.coni_zpk_points(data.kappa_cob, data.M_lattice(), p_full[k:k + 1], Q, 13, 2)
    sel = pidx == k
    assert sa[0] == st[k] and np.array_equal(Ma, M[sel]) and np.array_equal(Kna, Kn[sel])
    assert np.array_equal(qa, q[sel]) and np.all(pa == 0)
