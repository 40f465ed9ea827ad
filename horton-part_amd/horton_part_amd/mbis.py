"""Minimal Basis Iterative Stockholder (MBIS) on the GPU.

API of ``mbis.py`` in the reference: ``get_nshell`` (36-46),
``get_initial_mbis_propars`` (49-78) and ``MBISWPart`` (206-362).  The
per-atom fixed point of ``opt_mbis_propars`` (81-163) runs in the device
kernel ``k_mbis_inner``; the pro-atoms ``N S^3 exp(-S r) / (8 pi)`` in the
fused weights pass.
"""

import numpy as np

from .part import AbstractISAWPart

__all__ = ["MBISWPart", "get_nshell", "get_initial_mbis_propars"]


def get_nshell(number: int):
    """Number of Slater shells of element ``number`` (mbis.py:36-46)."""
    noble = np.array([2, 10, 18, 36, 54, 86, 118])
    return noble.searchsorted(number) + 1


def get_initial_mbis_propars(number: int):
    """Initial (N, S) pairs of element ``number`` (mbis.py:49-78)."""
    nshell = get_nshell(number)
    propars = np.zeros(2 * nshell, float)
    S0 = 2.0 * number
    if nshell > 1:
        alpha = (2.0 / S0) ** (1.0 / (nshell - 1))
    else:
        alpha = 1.0
    nel_in_shell = np.array([2.0, 8.0, 8.0, 18.0, 18.0, 32.0, 32.0])
    for ishell in range(nshell):
        propars[2 * ishell] = nel_in_shell[ishell]
        propars[2 * ishell + 1] = S0 * alpha**ishell
    propars[-2] = number - propars[:-2:2].sum()
    return propars


class MBISWPart(AbstractISAWPart):
    """Minimal Basis Iterative Stockholder (mbis.py:206-362)."""

    name = "mbis"

    def get_rgrid(self, iatom):
        if self.only_use_molgrid:
            raise NotImplementedError
        return self.get_grid(iatom).rgrid

    def _init_propars(self):
        """mbis.py:291-305."""
        self._ranges = [0]
        self._nshells = []
        for iatom in range(self.natom):
            nshell = int(get_nshell(self.numbers[iatom]))
            self._ranges.append(self._ranges[-1] + 2 * nshell)
            self._nshells.append(nshell)
        propars = np.zeros(self._ranges[-1])
        for iatom in range(self.natom):
            propars[self._ranges[iatom] : self._ranges[iatom + 1]] = get_initial_mbis_propars(self.numbers[iatom])
        return propars

    def _scheme_arrays(self):
        fn_off = np.concatenate([[0], np.cumsum([int(get_nshell(z)) for z in self.numbers])]).astype(np.int32)
        d = self._base_scheme_arrays()
        d.update(
            scheme=0,
            solver=0,
            fn_offsets=fn_off,
            inner_maxiter=2000,
            npar=int(2 * fn_off[-1]),
        )
        return d

    def get_proatom_rho(self, iatom, propars=None, **kwargs):
        """Pro-atom density and derivative on the radial grid (mbis.py:234-261)."""
        if propars is None:
            propars = self.cache.load("propars")
        r = self.radial_distances[iatom] if self.on_molgrid else self.get_rgrid(iatom).points
        y = np.zeros(len(r))
        d = np.zeros(len(r))
        my = propars[self._ranges[iatom] : self._ranges[iatom + 1]]
        for ishell in range(self._nshells[iatom]):
            N, S = my[2 * ishell : 2 * ishell + 2]
            f = N * S**3 * np.exp(-S * r) / (8 * np.pi)
            y += f
            d -= S * f
        return y, d

    def _finalize_propars(self):
        """mbis.py:347-362: valence/core split."""
        super()._finalize_propars()
        propars = self.cache.load("propars")
        valence_charges = np.array([-propars[self._ranges[i + 1] - 2] for i in range(self.natom)])
        valence_widths = np.array([1.0 / propars[self._ranges[i + 1] - 1] for i in range(self.natom)])
        self.cache.dump("core_charges", self.cache.load("charges") - valence_charges, tags="o")
        self.cache.dump("valence_charges", valence_charges, tags="o")
        self.cache.dump("valence_widths", valence_widths, tags="o")
