"""Iterative Stockholder Analysis (ISA) on the GPU (isa.py:37-122 in the reference).

The pro-atoms are tabulated: atom a's parameters are its pro-atom's values
at the knots of its radial grid, and the pro-atom is the not-a-knot cubic
spline through them (``get_proatom_spline``, core/stockholder.py:227-264).
Each outer iteration on the device (``hp_part_run``) builds the splines from
the parameters, evaluates them at every molecular point in the weights pass,
and takes as new parameters the spherical averages of the AIM densities
clipped below at 1e-100; the charge is ``Z - integral 4 pi r^2 rhobar``.
"""

import numpy as np

from .hirshfeld import tabulated_scheme
from .part import AbstractISAWPart, ProatomSplines

__all__ = ["ISAWPart"]


class ISAWPart(ProatomSplines, AbstractISAWPart):
    """ISA (isa.py:37-122): atom grids only (grid_type 1, as the reference's
    ``get_proatom_rho`` requires), FP64."""

    name = "is"

    def __init__(self, *args, grid_type=1, **kwargs):
        if grid_type != 1:
            raise NotImplementedError("ISA pro-atoms live on the atoms' radial grids (grid_type 1)")
        super().__init__(*args, grid_type=grid_type, **kwargs)

    def get_rgrid(self, index):
        if self.only_use_molgrid:
            raise NotImplementedError
        return self.get_grid(index).rgrid

    def get_proatom_rho(self, iatom, propars=None, **kwargs):
        """The parameters of atom ``iatom`` and no derivative (isa.py:65-91)."""
        if propars is None:
            propars = self.cache.load("propars")
        if self.on_molgrid:
            raise NotImplementedError
        return propars[self._ranges[iatom] : self._ranges[iatom + 1]], None

    def _init_propars(self):
        """Zeros, one per radial knot of every atom (isa.py:93-100)."""
        self._ranges = [0]
        for index in range(self.natom):
            self._ranges.append(self._ranges[-1] + self.get_rgrid(index).size)
        return np.zeros(self._ranges[-1])

    def _scheme_arrays(self):
        d = tabulated_scheme(self, "isa", [[] for _ in range(self.natom)])
        d["npar"] = int(sum(self.get_rgrid(i).size for i in range(self.natom)))
        for k in ("rec_offsets", "rec_charges", "rec_rho", "rec_coefs"):
            d.pop(k)
        return d

    def _register_lazy(self, dev, rhobar=None):
        """The cache entries of isa.py:108-113: the radial points, the
        clipped spherical averages and the radial weights of the last update."""
        super()._register_lazy(dev, rhobar)
        for a in range(self.natom):
            sph = self.cache.load(f"spherical_average_{a}")
            self.cache.dump(f"spherical_average_{a}", np.clip(sph, 1e-100, np.inf), tags="o")
