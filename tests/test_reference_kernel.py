"""The reference's own C kernel (fpk_enumerate via its tests/c/fp_kernel_cli
driver, built by oracle/ref.py from the reference sources with the image's
GMP) against the golden kernel vectors, the product and the oracle.

- The 812 golden (instance, dilation) cases: the reference's point counts
  equal the golden sets' sizes, and the product's host search returns as
  many points.
- The coniZpM problems of the headline batch and the four-instance mix (the
  oracle's lattice preparation, which builds the same bases as the
  product's): identical point counts, and the product visits at most the
  reference's nodes and at least 99 % of them (the reference prunes in
  float when qmax < 2^22, the product in FP64, which cuts a few more).

Skipped where the reference checkout (and so the binary) is absent."""

import numpy as np
import pytest

from helpers import kernel_problems
from oracle import cpu, ref
from pfvs_amd import gpu

pytestmark = pytest.mark.skipif(not ref.available() and ref.build() is None,
                                reason="the reference's C sources are not here")


def test_golden_cases_point_counts():
    gpu.require()
    probs, exps = kernel_problems()
    res = ref.kernel_counts(probs)
    assert all(r[0] == 0 for r in res)
    assert [r[1] for r in res] == [len(e["pts"]) for e in exps]
    pts, st, _ = gpu.kernel_batch(probs, mode="host")
    assert [len(p[0]) for p in pts] == [r[1] for r in res]


def _pipeline_problems(name, n, seed):
    import bench
    _, _, pf, _, items = bench.workload(name)
    probs = []
    for i in np.random.default_rng(seed).permutation(len(pf))[:n]:
        kap, Mb, Q, qmax = items[i]
        st, _, B, _, mat, H = cpu.build_lattice(kap, Mb, pf[i, :kap.shape[0]])
        if st != 0:
            continue
        d = mat.shape[0]
        rows = [[int(x) for x in r] for r in H if any(int(x) for x in r)]
        small = [r for r in rows if all(abs(x) < 2**63 for x in r)]
        probs.append(dict(mat=mat.tolist(), Q=Q, qmax=qmax, strict=True, linmin=13,
                          linvec=[int(x) for x in B[0]],
                          H=np.array(small, dtype=np.int64).reshape(-1, d),
                          L=[r for r in rows if r not in small]))
    return probs


@pytest.mark.parametrize("name", ["headline", "mixed"])
def test_pipeline_problems_nodes_and_points(name):
    gpu.require()
    probs = _pipeline_problems(name, 400, 1)
    assert len(probs) > 300
    res = ref.kernel_counts(probs)
    pts, st, nodes = gpu.kernel_batch(probs, mode="host")
    assert all(s == 0 for s in st) and all(r[0] == 0 for r in res)
    assert [len(p[0]) for p in pts] == [r[1] for r in res]
    ours, theirs = np.asarray(nodes, dtype=np.int64), np.array([r[2] for r in res])
    assert np.all(ours <= theirs) and np.all(ours >= 0.99 * theirs)
    assert np.mean(ours == theirs) > 0.9
