"""Writes the golden vectors under tests/golden/ from the reference's own
test data (run in a container that has the reference checkout; the outputs
are committed, the reference is never read at test time).

Sources (data only, re-serialized; no code is copied):
  - tests/data/fixtures.json.gz of natemacfadden/pfvs: 219 real coni kernel
    instances (mat, H, linvec, Q, M0min) with exact expected points and q
    at dilations 1/4/13/40, and 71 dataset geometries with their PFVs at
    (B, D); built upstream from the public dataset
    natemacfadden/calabi-yau-coni-pfvs @ 63d2db7942df733e1c92845451f5557cbbd16e87.
  - tests/data/paper_examples.json: the coni examples of arXiv:2406.13751.
  - tests/test_conipfv_kernel.py: the Manwe kernel known answers.
The reference is GPL-3.0-or-later; these data files keep that provenance.
"""

import gzip
import json
import os
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = os.path.dirname(os.path.abspath(__file__))


def main():
    with gzip.open(os.path.join(REF, "tests/data/fixtures.json.gz")) as f:
        fx = json.load(f)
    kernel = [{k: inst[k] for k in ("h11", "polyID", "coniID", "p", "Q", "M0min", "mat", "H",
                                    "linvec", "expected")} for inst in fx["kernel"]]
    with gzip.open(os.path.join(OUT, "kernel_cases.json.gz"), "wt") as f:
        json.dump({"source": fx["source"], "revision": fx["revision"], "kernel": kernel}, f,
                  separators=(",", ":"))
    geoms = [{k: g[k] for k in ("h11", "h21", "polyID", "classID", "coniID", "kappa_coo", "c2",
                                "H", "B", "D", "K", "M")} for g in fx["geometries"]]
    with gzip.open(os.path.join(OUT, "geometries.json.gz"), "wt") as f:
        json.dump({"source": fx["source"], "revision": fx["revision"], "geometries": geoms}, f,
                  separators=(",", ":"))
    with open(os.path.join(REF, "tests/data/paper_examples.json")) as f:
        ex = json.load(f)["examples"]
    coni = [{k: e[k] for k in ("paper", "section", "name", "model", "h11", "h21", "kappa", "c2",
                               "H", "coni_curve", "cob", "M", "K", "p", "p_den", "min_dilation",
                               "scan")} for e in ex if e["kind"] == "coni"]
    with open(os.path.join(OUT, "paper_coni.json"), "w") as f:
        json.dump(coni, f, separators=(",", ":"))
    # Manwe kernel known answers: numbers read from the reference test as data
    src = open(os.path.join(REF, "tests/test_conipfv_kernel.py")).read()
    body = src[src.index("U = np.array(["):src.index("MAX_N_OUT")]
    nums = lambda s: [float(x) for x in re.findall(r"-?\d+\.?\d*(?:e-?\d+)?", s)]
    U = nums(body[body.index("["):body.index("dtype")])
    H = nums(body[body.index("H = np.array("):][12:])
    kat = {"U": [U[i * 7:(i + 1) * 7] for i in range(7)],
           "Q": 162, "linvec": [0, 0, 0, 4, -4, -2, -2], "linmin": 13,
           "H": [[int(x) for x in H[i * 7:(i + 1) * 7]] for i in range(7)],
           "dilations": [1, 2, 10, 20, 10000, 20000, 30000, 40000, 50000],
           "counts": [0, 0, 0, 1, 2, 2, 2, 2, 2]}
    with open(os.path.join(OUT, "manwe_kernel.json"), "w") as f:
        json.dump(kat, f)
    print("wrote", len(kernel), "kernel instances,", len(geoms), "geometries,",
          len(coni), "coni examples")


if __name__ == "__main__":
    main()
