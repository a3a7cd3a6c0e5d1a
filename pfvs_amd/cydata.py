"""CYData: the Calabi-Yau data the coniZpM search consumes.

Mirrors the reference's CYData interface (pfvs/cydata.py:36-356) for the
parts on the coniZpM path: the coni change of basis (98-124), `kappa_cob`,
`c2_cob`, `H_cob`, the a-matrix and b-vector (252-303) and `M_lattice()`
(307-356), the lattice of integral M with (a/2) M and (b/24).M integral.
The M-lattice is computed here by successive modular sublattice cuts on
Python ints (pfvs_amd.lattice) instead of the reference's flint dual-lattice
construction; any basis of the same lattice gives the same PFVs.
"""

from __future__ import annotations

import numpy as np

from . import lattice


def _as_int_array(x):
    a = np.asarray(x)
    if a.dtype == object:
        return a
    return a.astype(np.int64)


class CYData:
    """CY data for (coni-)PFV searches: h21, kappa (h11^3), c2, Kahler-cone
    hyperplanes H and, for coni contexts, the conifold curve and a change of
    basis `coni_cob` with cob @ coni_curve = e_0 (pfvs/cydata.py:60-124)."""

    def __init__(self, h21, kappa, c2, H, coni_curve=None, coni_cob=None,
                 vertices="unknown", heights="unknown"):
        self._kappa = _as_int_array(kappa)
        self._c2 = _as_int_array(c2)
        self._H = _as_int_array(H)
        self._h21 = int(h21)
        self._h11 = int(self._kappa.shape[0])
        self._coni = coni_curve is not None
        self._vertices, self._heights = vertices, heights
        self._a = self._b = self._M_lattice = None
        if not self._coni:
            return
        self._coni_curve = np.asarray(coni_curve, dtype=np.int64)
        if coni_cob is None:
            # unimodular U with U @ curve = (g, 0, ...); g must be 1
            U = np.array(lattice.unimodular_for(self._coni_curve.tolist()), dtype=object)
            if int((U @ self._coni_curve.astype(object))[0]) == -1:
                U[0] = -U[0]
            self._cob = U.astype(np.int64)
        else:
            self._cob = np.asarray(coni_cob, dtype=np.int64)
        e0 = np.zeros(self._h11, dtype=np.int64)
        e0[0] = 1
        if not np.array_equal(self._cob @ self._coni_curve, e0):
            raise ValueError("coni_cob is not a valid change of basis: "
                             "cob @ coni_curve must equal (1, 0, ..., 0).")
        cob = self._cob.astype(object)
        kap = self._kappa.astype(object)
        kc = np.tensordot(cob, kap, axes=([1], [0]))
        kc = np.tensordot(kc, cob, axes=([1], [1])).transpose(0, 2, 1)
        kc = np.tensordot(kc, cob, axes=([2], [1]))
        self._kappa_cob = _fit64(kc)
        self._c2_cob = _fit64(cob @ self._c2.astype(object))
        Hc = (self._H.astype(object) @ cob.T)[:, 1:]
        Hc = Hc[np.array([any(int(x) != 0 for x in r) for r in Hc], dtype=bool)]
        self._H_cob = _fit64(Hc)

    # -- properties ---------------------------------------------------------
    def _need_coni(self, name):
        if not self._coni:
            raise AttributeError(f"{name} is only defined for coni CYData")

    @property
    def coni(self) -> bool:
        return self._coni

    @property
    def h11(self) -> int:
        return self._h11

    @property
    def h21(self) -> int:
        return self._h21

    @property
    def kappa(self):
        return self._kappa.copy()

    @property
    def c2(self):
        return self._c2.copy()

    @property
    def H(self):
        return self._H.copy()

    @property
    def coni_curve(self):
        self._need_coni("coni_curve")
        return self._coni_curve.copy()

    @property
    def cob(self):
        self._need_coni("cob")
        return self._cob.copy()

    @property
    def kappa_cob(self):
        self._need_coni("kappa_cob")
        return self._kappa_cob.copy()

    @property
    def c2_cob(self):
        self._need_coni("c2_cob")
        return self._c2_cob.copy()

    @property
    def H_cob(self):
        self._need_coni("H_cob")
        return self._H_cob.copy()

    @property
    def a(self):
        """a_xy = kappa_xxy (x >= y), kappa_xyy (x < y); coni row dropped
        (pfvs/cydata.py:252-283)."""
        if self._a is None:
            k = self._kappa_cob if self._coni else self._kappa
            h = k.shape[0]
            a = np.zeros((h, h), dtype=object)
            for x in range(h):
                for y in range(h):
                    a[x, y] = int(k[x, x, y]) if x >= y else int(k[x, y, y])
            self._a = a[1:] if self._coni else a
        return self._a.copy()

    @property
    def b(self):
        """c2 (non-coni) or c2_cob + 2 e_0 (pfvs/cydata.py:285-303)."""
        if self._b is None:
            b = (self._c2_cob if self._coni else self._c2).astype(object).copy()
            if self._coni:
                b[0] += 2
            self._b = b
        return self._b.copy()

    def M_lattice(self, verify: bool = True) -> np.ndarray:
        """Basis (columns, int64) of {M in Z^h : a M even, b.M = 0 mod 24}
        (pfvs/cydata.py:307-356): the constraints 12 a M = 0 and b.M = 0
        mod 24 are imposed one row at a time."""
        if self._M_lattice is None:
            h = self._h11
            rows = [[int(x) for x in self.b]] + [[12 * int(x) for x in r] for r in self.a]
            B = np.array(lattice.kernel_mod_lattice(rows, 24, h), dtype=object)
            if verify:
                if not all(int(x) % 2 == 0 for x in (self.a @ B).ravel()):
                    raise RuntimeError("M_lattice verification failed: a @ M is not even.")
                if not all(int(x) % 24 == 0 for x in (self.b @ B).ravel()):
                    raise RuntimeError("M_lattice verification failed: b.M is not divisible by 24.")
            self._M_lattice = _fit64(B)
        return self._M_lattice.copy()


def _fit64(A):
    A = np.asarray(A, dtype=object)
    if A.size == 0 or all(-2**63 <= int(x) < 2**63 for x in A.ravel()):
        return np.array(A.tolist(), dtype=np.int64).reshape(A.shape)
    return A
