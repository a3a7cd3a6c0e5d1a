"""The paper's coni examples (arXiv:2406.13751 sections 5.1-5.5 and C.3,
tests/golden/paper_coni.json, as the reference's
tests/test_paper_examples.py uses them): from the paper's p-vector the
search with Q = -M.K, M0min 13 at the example's minimal dilation returns
the paper's (K, M); a scan of the whole box (the recorded B) finds it too.
CPU for the examples and the small scans; on the GPU the search over every
non-slow scan (up to 227,395 p-vectors), enumerated by the host code."""

import json
import os

import numpy as np
import pytest

from pfvs_amd import coniZpM
from pfvs_amd.cydata import CYData
from pfvs_amd.pvectors import box_enum

HERE = os.path.dirname(os.path.abspath(__file__))
PAPER = json.load(open(os.path.join(HERE, "golden", "paper_coni.json")))


def _data(e):
    return CYData(e["h21"], e["kappa"], e["c2"], e["H"], coni_curve=e["coni_curve"],
                  coni_cob=e["cob"])


def _search(e, data, ps, device):
    Q = -int(np.dot(e["M"], e["K"]))
    return coniZpM(data, ps, Q=Q, M0min=13, ellipsoid_dilation=e["min_dilation"], n_jobs=1,
                   device=device)


def _found(e, Ks, Ms):
    return any(np.array_equal(k, e["K"]) and np.array_equal(m, e["M"]) for k, m in zip(Ks, Ms))


@pytest.mark.parametrize("e", PAPER, ids=[e["name"].replace(" ", "_") for e in PAPER])
def test_example_from_its_p(e):
    data = _data(e)
    Ks, Ms = _search(e, data, np.array([np.asarray(e["p"])[1:]]), "cpu")
    assert _found(e, Ks, Ms)


SMALL_SCANS = [e for e in PAPER if e["scan"]["n_p"] <= 3000]


@pytest.mark.parametrize("e", SMALL_SCANS, ids=[e["name"].replace(" ", "_") for e in SMALL_SCANS])
def test_example_in_small_scan(e):
    data = _data(e)
    ps = box_enum(e["scan"]["B"], data.H_cob, 1, True, device="cpu")
    assert len(ps) == e["scan"]["n_p"]
    Ks, Ms = _search(e, data, ps, "cpu")
    assert _found(e, Ks, Ms)


@pytest.mark.gpu
@pytest.mark.parametrize("e", [e for e in PAPER if not e["scan"]["slow"]],
                         ids=[e["name"].replace(" ", "_") for e in PAPER if not e["scan"]["slow"]])
def test_gpu_example_in_scan(need_gpu, e):
    data = _data(e)
    print(f"{e['name']}: scan B {e['scan']['B']}", flush=True)
    ps = box_enum(e["scan"]["B"], data.H_cob, 1, True, device="cpu")
    assert len(ps) == e["scan"]["n_p"]
    Kg, Mg = _search(e, data, ps, "gpu")
    assert _found(e, Kg, Mg)
    if len(ps) <= 3000:
        Kc, Mc = _search(e, data, ps, "cpu")
        assert np.array_equal(Kg, Kc) and np.array_equal(Mg, Mc)
