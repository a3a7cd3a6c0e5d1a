"""The reference's own C kernel as a checker -- TEST INFRASTRUCTURE ONLY.

`build()` compiles the reference's standalone driver
tests/c/fp_kernel_cli.c (with pfvs/fp_kernel/fp_kernel.h,
fpk_common.h and fpk_search_impl.h, read where they lie under the
reference checkout) with gcc against the image's GMP into
oracle/_ref/fp_kernel_cli; nothing from the reference is copied.  Where
the checkout is absent (the GPU box) only an already built binary is used.

`kernel_counts(problems)` feeds problems to it (its stdin format,
fp_kernel_cli.c:18-28) and returns per problem (status, points, nodes,
candidates, leaves): the reference's fpk_enumerate, including its float
pruning when qmax < 2^22 (fp_kernel.h:324-373) and its GMP rows.
"""

from __future__ import annotations

import os
import re
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
REF = os.environ.get("PFVS_REFERENCE", "/root/reference")
CLI = os.path.join(HERE, "_ref", "fp_kernel_cli")
GMP_PREFIX = os.environ.get("PFVS_GMP_PREFIX", "/opt/conda")


def build(force: bool = False):
    """Compile the reference's CLI (when its sources are here); returns the
    binary's path or None."""
    src = os.path.join(REF, "tests", "c", "fp_kernel_cli.c")
    inc = os.path.join(REF, "pfvs", "fp_kernel")
    if not os.path.exists(src):
        return CLI if os.path.exists(CLI) else None
    if force or not os.path.exists(CLI) or os.path.getmtime(CLI) < os.path.getmtime(src):
        os.makedirs(os.path.dirname(CLI), exist_ok=True)
        subprocess.run(["gcc", "-std=c11", "-O2", "-I" + inc, "-I" + GMP_PREFIX + "/include", src,
                        "-o", CLI, "-L" + GMP_PREFIX + "/lib", "-Wl,-rpath," + GMP_PREFIX + "/lib",
                        "-lgmp", "-lm"], check=True)
    return CLI


def available() -> bool:
    return os.path.exists(CLI)


_LINE = re.compile(r"status=(-?\d+) n_out=(\d+) nodes=(\d+) cand=(\d+) leaf=(\d+)")


def kernel_counts(problems):
    """[(status, n_out, nodes, cand, leaf)] of the reference's fpk_enumerate
    for problems (dicts: mat, Q, qmax, strict, H rows (any size, Python
    ints), linvec or None, linmin)."""
    lines = []
    for p in problems:
        d = len(p["mat"])
        H = [[int(x) for x in r] for r in p.get("H", [])] + [[int(x) for x in r] for r in p.get("L", [])]
        lin = p.get("linvec")
        lines.append(f"{d} {int(p['Q'])} {int(p['qmax'])} {int(bool(p.get('strict', True)))} {len(H)} "
                     f"{int(lin is not None)} {int(p.get('linmin') or 0)}")
        lines.append(" ".join(str(int(x)) for r in p["mat"] for x in r))
        if H:
            lines.append(" ".join(str(x) for r in H for x in r))
        if lin is not None:
            lines.append(" ".join(str(int(x)) for x in lin))
    out = subprocess.run([CLI], input="\n".join(lines) + "\n", capture_output=True, text=True,
                         check=True).stdout
    res = [tuple(int(g) for g in m.groups()) for m in _LINE.finditer(out)]
    if len(res) != len(problems):
        raise RuntimeError(f"reference CLI answered {len(res)} of {len(problems)} problems")
    return res
