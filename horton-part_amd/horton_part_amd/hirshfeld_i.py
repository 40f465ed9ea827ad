"""Iterative Hirshfeld (Hirshfeld-I) on the GPU (hirshfeld_i.py:36-176 in the reference).

Each atom's pro-atom interpolates linearly between the database records of
the two integer charges around its current charge q: with i = floor(q) and
x = q - i it is (1 - x) rec(i) + x rec(i + 1) (rec(i) alone when the
pseudo population Z - i is 1 or x is 0).  The outer loop runs on the device
(``hp_part_run``): per iteration the tables are combined from the records,
one weights pass gives every atom's atom-grid population, the charges follow,
and the change is measured on the record grids (core/iterstock.py:32-45).
"""

import numpy as np

from . import _native
from .cache import just_once
from .hirshfeld import check_proatomdb, do_dispersion, tabulated_scheme
from .part import AbstractISAWPart, ProatomSplines
from .proatomdb import spline_table

__all__ = ["HirshfeldIWPart"]


class HirshfeldIWPart(ProatomSplines, AbstractISAWPart):
    """Hirshfeld-I (hirshfeld_i.py:36-176).  Atom grids only (grid_type 1: the
    reference's atom-grid populations do not match molecular-grid weights).
    Runs in FP64."""

    name = "hi"

    def __init__(self, coordinates, numbers, pseudo_numbers, grid, moldens, proatomdb, spindens=None, lmax=3,
                 logger=None, threshold=1e-6, maxiter=500, grid_type=1, device=0, **kwargs):
        check_proatomdb(numbers, pseudo_numbers, proatomdb)
        self._proatomdb = proatomdb
        if grid_type != 1:
            raise NotImplementedError("Hirshfeld-I runs on atom grids (grid_type 1)")
        super().__init__(coordinates, numbers, pseudo_numbers, grid, moldens, spindens, lmax, logger,
                         threshold=threshold, maxiter=maxiter, grid_type=grid_type, device=device)

    @property
    def proatomdb(self):
        return self._proatomdb

    def get_rgrid(self, index):
        return self.proatomdb.get_rgrid(self.numbers[index])

    def get_interpolation_info(self, i, charges=None):
        """(floor of the charge, fractional part) (hirshfeld_i.py:116-122)."""
        if charges is None:
            charges = self.cache.load("charges")
        icharge = int(np.floor(charges[i]))
        return icharge, charges[i] - icharge

    def get_proatom_rho(self, iatom, charges=None, **kwargs):
        """Interpolated record density and derivative (hirshfeld_i.py:124-134)."""
        icharge, x = self.get_interpolation_info(iatom, charges)
        pseudo_pop = self.pseudo_numbers[iatom] - icharge
        number = self.numbers[iatom]
        if pseudo_pop == 1 or x == 0.0:
            return self.proatomdb.get_rho(number, {icharge: 1 - x}, do_deriv=True)
        if pseudo_pop > 1:
            return self.proatomdb.get_rho(number, {icharge: 1 - x, icharge + 1: x}, do_deriv=True)
        raise ValueError("Requesting a pro-atom with a negative (pseudo) population")

    def get_isolated(self, index, charge, grid):
        """One record's spline at ``grid``'s points (hirshfeld_i.py:136-146)."""
        spline = self.proatomdb.get_spline(self.numbers[index], charge)
        out = np.zeros(grid.size)
        self.eval_spline(index, spline, out, grid, "isolated q=%+i" % charge)
        return out

    def eval_proatom(self, index, output, grid):
        """The records' splines combined at ``grid``'s points, plus 1e-100
        (hirshfeld_i.py:148-158)."""
        icharge, x = self.get_interpolation_info(index)
        output[:] = self.get_isolated(index, icharge, grid)
        output *= 1 - x
        pseudo_pop = self.pseudo_numbers[index] - icharge
        if pseudo_pop > 1 and x != 0.0:
            output += self.get_isolated(index, icharge + 1, grid) * x
        elif pseudo_pop <= 0:
            raise ValueError("Requesting a pro-atom with a negative (pseudo) population")
        output += 1e-100

    def _init_propars(self):
        """Charges start at zero: the first iteration is Hirshfeld
        (hirshfeld_i.py:160-164)."""
        return np.zeros(self.natom)

    def _scheme_arrays(self):
        db = self.proatomdb
        records = []
        for i in range(self.natom):
            number = self.numbers[i]
            points = db.get_rgrid(number).points
            atom = []
            for charge in db.get_charges(number):
                r = db.get_record(number, charge)
                atom.append((charge, r.rho, spline_table(points, r.rho, r.deriv)))
            records.append(atom)
        return tabulated_scheme(self, "hi", records)

    def _raise_status(self, status_all, status_last, rc):
        if rc == _native.HP_E_SOLVER and int(np.bitwise_or.reduce(status_last)) & _native.ST_NO_RECORD:
            # the reference raises from get_proatom_rho at these charges
            charges = self._dev.results()[1]
            for i in range(self.natom):
                self.get_proatom_rho(i, charges)
            raise RuntimeError("no pro-atom for an atom's charge")
        super()._raise_status(status_all, status_last, rc)

    @just_once
    def do_dispersion(self):
        do_dispersion(self)
