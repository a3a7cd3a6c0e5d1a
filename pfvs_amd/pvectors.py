"""p-vector enumeration: the primitive integer points of the Kahler cone in a
box, the input side of coniZpM (reference: pfvs/pvectors.py:58-89, which
takes it from the latticepts dependency; the same box_enum contract is
called at pfvs/distributed.py:121-139).

The library's enumeration (pfa_box_enum) runs on host threads ("cpu",
"auto") or on the GPU ("gpu", a device index: one GPU thread per unit of
the box), with the same per-unit walk.  Points come in lexicographic order
(p_0 outermost, ascending).
"""

from __future__ import annotations

from . import gpu


def _device(device):
    if device in ("cpu", "auto", -1):
        return -1
    if device == "gpu" or (isinstance(device, int) and device >= 0):
        return 0 if device == "gpu" else int(device)
    raise ValueError(f"device must be 'auto', 'cpu', 'gpu' or a device index, got {device!r}")


def box_enum_status(B: int, H, rhs=1, cap: int = 10**8, primitive: bool = True,
                    device="auto"):
    """latticepts.box_enum's form: (ps, status, total) with ps the first
    min(total, cap) integer p with |p|_inf <= B and H p >= rhs, status 0, or
    1 when more than cap exist."""
    ps, st, total, _ = gpu.box_enum(H, rhs, int(B), primitive, cap, _device(device))
    return ps, st, total


def box_enum(B: int, H, rhs=1, primitive: bool = True, max_pts: int = 10**8, device="auto"):
    """All integer p with |p|_inf <= B and H p >= rhs (rows of H),
    optionally primitive (gcd 1), in lexicographic order; raises if there
    are more than max_pts."""
    ps, st, total = box_enum_status(B, H, rhs, max_pts, primitive, device)
    if st:
        raise RuntimeError(f"box_enum: {total} points exceed max_pts = {max_pts}")
    return ps


def pvecs(data, min_N_pts: int, primitive: bool = True, device="auto"):
    """Primitive cone points in the smallest box holding at least min_N_pts
    of them (pfvs/pvectors.py:58-89)."""
    if min_N_pts <= 0:
        raise ValueError(f"min_N_pts must be > 0, got {min_N_pts}.")
    H = data.H_cob if data.coni else data.H
    B = 1
    while True:
        ps = box_enum(B, H, 1, primitive, device=device)
        if len(ps) >= min_N_pts or B > 64:
            return ps
        B += 1
