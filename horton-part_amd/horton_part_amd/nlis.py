"""Non-Linear approximation of Iterative Stockholder (NLIS) on the GPU.

API of ``nlis.py`` in the reference: ``get_nlis_nshell`` (44-56),
``get_initial_nlis_propars`` (59-96) and ``NLISWPart`` (198-367).  Each
pro-atom is a sum of shells ``N g(n, S) exp(-S r^n)`` with
``g(n, S) = n S^(3/n) / (4 pi Gamma(3/n))`` and three parameters (N, S, n)
per shell, the power n fixed.  The weights pass evaluates these shells on the
device; the per-atom fixed point of ``opt_nlis_propars`` (99-195) runs in the
kernel ``k_nlis_inner``.  The device runs NLIS on atom grids (grid_type 1).
"""

import math

import numpy as np

from .mbis import get_nshell
from .part import AbstractISAWPart

__all__ = ["NLISWPart", "get_initial_nlis_propars", "get_nlis_nshell"]


def get_nlis_nshell(number, nshell_dict):
    """Shells of element ``number``: from ``nshell_dict``, else as MBIS counts them."""
    return nshell_dict.get(number, get_nshell(number))


def _power_of(exp_n_dict, number, ishell):
    """The power n of a shell: ``exp_n_dict[(Z, k)]`` when present, else 1."""
    return exp_n_dict[(number, ishell)] if (number, ishell) in exp_n_dict else 1.0


def get_initial_nlis_propars(number, exp_n_dict, nshell_dict, logger=None):
    """(N, S, n) per shell: N = Z / K, S from 2Z down to 0.5 geometrically, n
    from ``exp_n_dict`` (nlis.py:59-96)."""
    nbs = get_nlis_nshell(number, nshell_dict)
    S0 = 2.0 * number
    ratio = (0.5 / S0) ** (1.0 / (nbs - 1)) if nbs > 1 else 1.0
    rows = [(number / nbs, S0 * ratio**k, _power_of(exp_n_dict, number, k)) for k in range(nbs)]
    return np.array(rows, dtype=float).ravel()


def nlis_density(pars, r, nderiv=0):
    """Sum of the shells (N, S, n) of ``pars`` at the radii ``r`` (and the
    radial derivative the reference returns, nlis.py:296-299)."""
    y = np.zeros(len(r), float)
    d = np.zeros(len(r), float)
    for N, S, n in np.asarray(pars, float).reshape(-1, 3):
        f = N * n * S ** (3 / n) * np.exp(-S * r**n) / (4 * np.pi * math.gamma(3.0 / n))
        y += f
        d -= N * S * n * r ** (n - 1) * f
    return (y, d) if nderiv else y


class NLISWPart(AbstractISAWPart):
    """Non-Linear approximation of Iterative Stockholder (nlis.py:198-367)."""

    name = "nlis"
    _nshell_of = staticmethod(get_nlis_nshell)

    def __init__(
        self,
        coordinates,
        numbers,
        pseudo_numbers,
        grid,
        moldens,
        spindens=None,
        lmax=3,
        logger=None,
        threshold=1e-6,
        maxiter=500,
        inner_threshold=1e-8,
        exp_n_dict=1.0,
        nshell_dict=None,
        grid_type=1,
        **kwargs,
    ):
        self._exp_n_dict = exp_n_dict
        self._nshell_dict = nshell_dict or {}
        super().__init__(
            coordinates,
            numbers,
            pseudo_numbers,
            grid,
            moldens,
            spindens,
            lmax=lmax,
            logger=logger,
            threshold=threshold,
            maxiter=maxiter,
            inner_threshold=inner_threshold,
            grid_type=grid_type,
            **kwargs,
        )

    def get_rgrid(self, iatom):
        if self.only_use_molgrid:
            raise NotImplementedError
        return self.get_grid(iatom).rgrid

    def _initial(self, number):
        return get_initial_nlis_propars(number, self._exp_n_dict, self._nshell_dict)

    def _init_propars(self):
        """nlis.py:330-346."""
        per_atom = [self._initial(int(z)) for z in self.numbers]
        self._nshells = [len(p) // 3 for p in per_atom]
        self._ranges = [0] + [int(v) for v in np.cumsum([len(p) for p in per_atom])]
        return np.concatenate(per_atom)

    def get_proatom_rho(self, iatom, propars=None, **kwargs):
        """Pro-atom density and derivative (nlis.py:272-300)."""
        if propars is None:
            propars = self.cache.load("propars")
        r = self.radial_distances[iatom] if self.on_molgrid else self.get_rgrid(iatom).points
        return nlis_density(propars[self._ranges[iatom] : self._ranges[iatom + 1]], r, nderiv=1)

    def _scheme_arrays(self):
        propars = self._init_propars()
        counts = [(self._ranges[a + 1] - self._ranges[a]) // 3 for a in range(self.natom)]
        fn_off = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
        d = self._base_scheme_arrays()
        d.update(
            scheme=3,
            solver=5,
            fn_offsets=fn_off,
            fn_order=np.ascontiguousarray(propars[2::3]),
            inner_maxiter=2000,
            npar=int(propars.size),
        )
        return d

    def _finalize_propars(self):
        """nlis.py:351-367: valence and core split of the last shell."""
        super()._finalize_propars()
        propars = self.cache.load("propars")
        last = [self._ranges[i + 1] for i in range(self.natom)]
        valence_charges = np.array([-propars[e - 3] for e in last])
        valence_widths = np.array([1.0 / propars[e - 2] for e in last])
        self.cache.dump("core_charges", self.cache.load("charges") - valence_charges, tags="o")
        self.cache.dump("valence_charges", valence_charges, tags="o")
        self.cache.dump("valence_widths", valence_widths, tags="o")
