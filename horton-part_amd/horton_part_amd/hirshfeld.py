"""Hirshfeld partitioning on the GPU (hirshfeld.py:34-193 in the reference).

Every atom's pro-atom is its element's neutral record from a ``ProAtomDB``
(the spline of ``get_proatom_spline``, core/stockholder.py:227-264, plus
1e-100).  One device weights pass evaluates the splines at every molecular
point and gives the AIM weights and the atom-grid populations
(``hp_part_weights``, ``hp_part_populations``); moments, spin charges and the
dispersion coefficients follow from the same device weights.
"""

import numpy as np

from . import _native
from .cache import just_once
from .part import ProatomSplines, WPart
from .proatomdb import spline_table

__all__ = ["HirshfeldWPart", "check_proatomdb", "do_dispersion", "REFERENCE_C6"]

# Isolated-atom C6 coefficients in atomic units by element (the reference's
# table, hirshfeld.py:59-103: Chu and Dalgarno 2004, hydrogen from Yan et al.
# 1996), scaled by the squared volume ratio in do_dispersion.
REFERENCE_C6 = {
    1: 6.499, 2: 1.42, 3: 1392.0, 4: 227.0, 5: 99.5, 6: 46.6, 7: 24.2, 8: 15.6, 9: 9.52, 10: 6.20,
    11: 1518.0, 12: 626.0, 13: 528.0, 14: 305.0, 15: 185.0, 16: 134.0, 17: 94.6, 18: 64.2, 19: 3923.0,
    20: 2163.0, 21: 1383.0, 22: 1044.0, 23: 832.0, 24: 602.0, 25: 552.0, 26: 482.0, 27: 408.0, 28: 373.0,
    29: 253.0, 30: 284.0, 31: 498.0, 32: 354.0, 33: 246.0, 34: 210.0, 35: 162.0, 36: 130.0, 37: 4769.0,
    38: 3175.0, 49: 779.0, 50: 659.0, 51: 492.0, 52: 445.0, 53: 385.0,
}


def check_proatomdb(numbers, pseudo_numbers, proatomdb):
    """The database's neutral record of each atom must carry the atom's
    pseudo number (hirshfeld.py:34-45)."""
    for i, (number, pseudo) in enumerate(zip(numbers, pseudo_numbers)):
        expected = proatomdb.get_record(number, 0).pseudo_number
        if expected != pseudo:
            raise ValueError(
                "The pseudo number of atom %i does not match with the proatom database (%i!=%i)" % (i, pseudo, expected)
            )


def do_dispersion(part):
    """Atoms-in-molecules C6 coefficients from the AIM volumes (third radial
    moments) relative to the neutral records' (hirshfeld.py:48-124): cache
    entries ``volumes``, ``volume_ratios`` and ``c6s`` (-1 where the element
    has no reference value)."""
    if part.lmax < 3:
        part.logger.warning("Skip computing dispersion coefficients because lmax=%i<3" % part.lmax)
    volumes, new_v = part.cache.load("volumes", alloc=part.natom, tags="o")
    ratios, new_r = part.cache.load("volume_ratios", alloc=part.natom, tags="o")
    c6s, new_c = part.cache.load("c6s", alloc=part.natom, tags="o")
    if new_v or new_r or new_c:
        part.do_moments()
        radial = part.cache.load("radial_moments")
        for i in range(part.natom):
            n = int(part.numbers[i])
            volumes[i] = radial[i, 3]
            ratios[i] = volumes[i] / part.proatomdb.get_record(n, 0).get_moment(3)
            c6s[i] = ratios[i] ** 2 * REFERENCE_C6[n] if n in REFERENCE_C6 else -1


def table_offsets(part):
    """Knot offsets, knots and radial weights of each atom's pro-atom grid."""
    grids = [part.get_rgrid(i) for i in range(part.natom)]
    offsets = np.concatenate([[0], np.cumsum([g.size for g in grids])]).astype(np.int32)
    knots = np.concatenate([g.points for g in grids]).astype(float)
    weights = np.concatenate([g.weights for g in grids]).astype(float)
    return offsets, knots, weights


def tabulated_scheme(part, scheme, records):
    """The ``hp_scheme_desc`` entries of a tabulated scheme: the knots of
    every atom and, per atom, its records as (charge, rho, spline pieces)."""
    offsets, knots, weights = table_offsets(part)
    rec_off, charges, rho, coefs = [0], [], [], []
    for atom_records in records:
        for charge, r, c in atom_records:
            charges.append(int(charge))
            rho.append(np.asarray(r, float))
            coefs.append(np.asarray(c, float).ravel())
        rec_off.append(rec_off[-1] + len(atom_records))
    return {
        "scheme": _native.SCHEME[scheme],
        "solver": 0,
        "grid_type": part.grid_type,
        "threshold": float(getattr(part, "_threshold", 1e-6)),
        "inner_threshold": float(getattr(part, "_threshold", 1e-6)),
        "maxiter": int(getattr(part, "_maxiter", 1)),
        "inner_maxiter": 1,
        "density_cutoff": float(part.density_cutoff),
        "negative_cutoff": float(part.negative_cutoff),
        "population_cutoff": float(part.population_cutoff),
        "cutoff_radius": 0.0,
        "newton_tau": 1.0,
        "newton_linspace": 20,
        "tab_offsets": offsets,
        "tab_knots": knots,
        "tab_rweights": weights,
        "rec_offsets": np.asarray(rec_off, np.int32),
        "rec_charges": np.asarray(charges, np.int32),
        "rec_rho": np.concatenate(rho) if rho else np.zeros(1),
        "rec_coefs": np.concatenate(coefs) if coefs else np.zeros(4),
        "npar": part.natom,
    }


class HirshfeldWPart(ProatomSplines, WPart):
    """Hirshfeld partitioning (hirshfeld.py:127-193) with the weights pass on
    the GPU.  Extra keyword: ``device`` (GPU ordinal, default 0); others that
    the reference ignores (such as ``local``) are ignored."""

    name = "h"

    def __init__(self, coordinates, numbers, pseudo_numbers, grid, moldens, proatomdb, spindens=None, lmax=3,
                 logger=None, grid_type=1, device=0, **kwargs):
        check_proatomdb(numbers, pseudo_numbers, proatomdb)
        self._proatomdb = proatomdb
        self._device = device
        self._dev = None
        super().__init__(coordinates, numbers, pseudo_numbers, grid, moldens, spindens, lmax, logger, grid_type=grid_type)

    @property
    def proatomdb(self):
        return self._proatomdb

    def get_rgrid(self, index):
        return self.proatomdb.get_rgrid(self.numbers[index])

    def get_proatom_rho(self, iatom, *args, **kwargs):
        return self.proatomdb.get_rho(self.numbers[iatom], do_deriv=True)

    def _records(self):
        out = []
        for i in range(self.natom):
            rho, deriv = self.get_proatom_rho(i)
            rho, deriv = self.fix_proatom_rho(i, np.array(rho, float), deriv)
            out.append([(0, rho, spline_table(self.get_rgrid(i).points, rho, deriv))])
        return out

    def _ensure_device(self):
        if self._dev is None:
            scheme = tabulated_scheme(self, "hirshfeld", self._records())
            self._dev = _native.DevicePartition(self._molecule_arrays(), scheme, device=self._device)
        return self._dev

    @just_once
    def do_partitioning(self):
        """update_at_weights (core/stockholder.py:352-384) on the device."""
        dev = self._ensure_device()
        dev.weights(spherical=False)
        for a in range(self.natom):
            n = int(self.grid.size) if self.on_molgrid else int(self.grid.indices[a + 1] - self.grid.indices[a])
            self.cache.add_lazy(f"at_weights_{a}", (lambda a=a, n=n: dev.at_weights(a, n)))
        self.cache.add_lazy("promoldens", dev.promoldens)

    @just_once
    def do_populations(self):
        """core/base.py:277-300 from the device weights pass: atom-grid
        populations (grid types 1, 2) or molecular-grid ones (grid_type 3)."""
        populations, new = self.cache.load("populations", alloc=self.natom, tags="o")
        if new:
            self.do_partitioning()
            pseudo = self.cache.load("pseudo_populations", alloc=self.natom, tags="o")[0]
            pseudo[:] = self._dev.populations()
            populations[:] = pseudo + self.numbers - self.pseudo_numbers

    def _device_moments(self):
        return self._dev.moments(self.lmax)

    @just_once
    def do_dispersion(self):
        do_dispersion(self)

    def close(self):
        if self._dev is not None:
            self._dev.close()
