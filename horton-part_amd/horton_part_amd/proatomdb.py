"""Pro-atom databases for Hirshfeld and Hirshfeld-I (core/proatomdb.py:34-447).

A ``ProAtomRecord`` is the spherically averaged density of one isolated atom
or ion on a radial grid (with its radial derivative when known).  A
``ProAtomDB`` groups records by element and charge; all records of an element
share one radial grid.  Pro-atoms are piecewise cubics through a record's
density: Hermite cubics when the derivative is known, the not-a-knot spline
otherwise, both continued beyond the end knots (``get_spline``,
core/proatomdb.py:378-388).  ``spline_table`` hands the same pieces to the
device in the layout of ``hp_scheme_desc.rec_coefs``.
"""

import logging

import numpy as np
from scipy.interpolate import CubicHermiteSpline, CubicSpline

from .grid import OneDGrid

__all__ = ["ProAtomRecord", "ProAtomDB", "spline_table"]

logger = logging.getLogger(__name__)


def spline_table(points, rho, deriv=None):
    """The pieces of a record's spline, one row ``(c0, c1, c2, c3)`` per knot
    (``c0 d^3 + c1 d^2 + c2 d + c3`` at ``d = r - points[i]``; the last row is
    zero).  Hermite cubics through ``(rho, deriv)``, or the not-a-knot spline
    through ``rho`` when ``deriv`` is None."""
    points = np.asarray(points, float)
    if deriv is None:
        spline = CubicSpline(points, rho)
    else:
        spline = CubicHermiteSpline(points, rho, deriv)
    table = np.zeros((points.size, 4))
    table[:-1] = spline.c.T
    return table


class ProAtomRecord:
    """One isolated-atom density (core/proatomdb.py:34-210).

    ``number`` and integer ``charge`` name the atom or ion; ``rgrid`` is its
    radial grid (``points``, ``weights``, ``integrate``), ``rho`` the density
    there and ``deriv`` its radial derivative or None.  ``pseudo_number``
    defaults to ``number`` (effective core charge otherwise)."""

    def __init__(self, number, charge, energy, rgrid, rho, deriv=None, pseudo_number=None, ipot_energy=None):
        self._number = int(number)
        self._charge = int(charge)
        self._energy = float(energy)
        self._rgrid = rgrid
        self._rho = np.asarray(rho, float)
        self._deriv = None if deriv is None else np.asarray(deriv, float)
        self._pseudo_number = self._number if pseudo_number is None else pseudo_number
        # a one-electron atom's ionisation energy is minus its energy
        self._ipot_energy = -self._energy if self.pseudo_population == 1 else ipot_energy
        self._safe = True

    number = property(lambda self: self._number, doc="The atomic number.")
    pseudo_number = property(lambda self: self._pseudo_number, doc="The effective core charge.")
    charge = property(lambda self: self._charge, doc="The net charge.")
    energy = property(lambda self: self._energy, doc="The total energy.")
    ipot_energy = property(lambda self: self._ipot_energy, doc="The ionisation energy (or None).")
    rho = property(lambda self: self._rho, doc="The density on the radial grid.")
    deriv = property(lambda self: self._deriv, doc="The radial derivative of the density, or None.")
    rgrid = property(lambda self: self._rgrid, doc="The radial grid.")
    safe = property(lambda self: self._safe, doc="False when the record looks bound by its basis set.")

    @property
    def population(self):
        return self._number - self._charge

    @property
    def pseudo_population(self):
        return self._pseudo_number - self._charge

    def update_safe(self, other):
        """A record of the same element with fewer electrons and a lower
        energy marks this one unsafe; the record with one electron less gives
        the ionisation energy (core/proatomdb.py:140-152)."""
        if other.number != self._number:
            return
        if other.population < self.population and other.energy < self.energy:
            self._safe = False
        if other.population == self.population - 1:
            self._ipot_energy = other.energy - self.energy

    def compute_radii(self, populations):
        """Knot indices and radii (linear between knots) at which the running
        integral of ``4 pi r^2 rho`` reaches each population
        (core/proatomdb.py:154-179)."""
        r = self.rgrid.points
        running = np.cumsum(4 * np.pi * r**2 * self.rho * self.rgrid.weights)
        indexes = running.searchsorted(populations)
        radii = []
        for pop, i in zip(populations, indexes):
            if i == len(running):
                radii.append(r[-1])
            else:
                x = (pop - running[i]) / (running[i - 1] - running[i])
                radii.append(x * r[i - 1] + (1 - x) * r[i])
        return indexes, radii

    def get_moment(self, order):
        """``integral rho r^order`` on the radial grid (core/proatomdb.py:181-183)."""
        return self.rgrid.integrate(self.rho, self.rgrid.points**order)

    def chop(self, npoint):
        """Keep the first ``npoint`` knots (core/proatomdb.py:185-190)."""
        self._rho = self._rho[:npoint]
        if self._deriv is not None:
            self._deriv = self._deriv[:npoint]
        self._rgrid = OneDGrid(self._rgrid.points[:npoint], self._rgrid.weights[:npoint])

    def __eq__(self, other):
        same_deriv = (self.deriv is None and other.deriv is None) or (
            self.deriv is not None and other.deriv is not None and bool((self.deriv == other.deriv).all())
        )
        return (
            self.number == other.number
            and self.charge == other.charge
            and self.energy == other.energy
            and bool((self.rho == other.rho).all())
            and same_deriv
            and np.array_equal(self.rgrid.points, other.rgrid.points)
            and np.array_equal(self.rgrid.weights, other.rgrid.weights)
            and self.pseudo_number == other.pseudo_number
            and self.ipot_energy == other.ipot_energy
        )

    def __ne__(self, other):
        return not self.__eq__(other)


class ProAtomDB:
    """Records by (element, charge) (core/proatomdb.py:213-447).  Of several
    records with the same element and charge the lowest in energy is kept;
    records of one element must share their radial grid."""

    def __init__(self, records):
        best = {}
        for r in records:
            key = (r.number, r.charge)
            if key not in best or r.energy < best[key].energy:
                best[key] = r
        self._records = list(best.values())
        self._map = dict(best)
        self._rgrid_map = {}
        for r in self._records:
            rgrid = self._rgrid_map.setdefault(r.number, r.rgrid)
            if not np.allclose(rgrid.points, r.rgrid.points):
                raise ValueError("All proatoms of a given element must have the same radial grid.")
        for number in self.get_numbers():
            for c0 in self.get_charges(number):
                for c1 in self.get_charges(number):
                    self.get_record(number, c0).update_safe(self.get_record(number, c1))

    def get_record(self, number, charge):
        return self._map[(number, charge)]

    def get_numbers(self):
        """Elements in the database, ascending."""
        return sorted(self._rgrid_map)

    def get_charges(self, number, safe=False):
        """Charges of an element's records, descending (only safe ones if asked)."""
        return sorted((r.charge for r in self._records if r.number == number and (r.safe or not safe)), reverse=True)

    def get_rgrid(self, number):
        return self._rgrid_map[number]

    @property
    def size(self):
        return len(self._records)

    def get_rho(self, number, parameters=0, combine="linear", do_deriv=False):
        """Density of one record (``parameters`` an int charge) or of a
        combination ``{charge: coefficient}``, linear or geometric
        (core/proatomdb.py:303-376); with ``do_deriv`` also the radial
        derivative (None when a record has none)."""
        if isinstance(parameters, (int, np.integer)):
            record = self.get_record(number, int(parameters))
            return (record.rho, record.deriv) if do_deriv else record.rho
        if not isinstance(parameters, dict):
            raise TypeError("Could not interpret parameters argument")
        if combine not in ("linear", "geometric"):
            raise ValueError('Combine argument "%s" not supported.' % combine)
        rho = 0.0
        deriv = 0.0 if do_deriv else None
        for charge, coeff in parameters.items():
            if coeff == 0.0:
                continue
            record = self.get_record(number, charge)
            have = do_deriv and record.deriv is not None and deriv is not None
            if combine == "linear":
                rho = rho + coeff * record.rho
                deriv = deriv + coeff * record.deriv if have else None
            else:
                rho = rho + coeff * np.log(record.rho)
                deriv = deriv + coeff * record.deriv / record.rho if have else None
        if combine == "geometric":
            rho = np.exp(rho)
            if do_deriv and deriv is not None:
                deriv = rho * deriv
        if not isinstance(rho, np.ndarray):
            rho = np.zeros_like(self.get_rgrid(number).points)
            if do_deriv:
                deriv = np.zeros_like(self.get_rgrid(number).points)
        return (rho, deriv) if do_deriv else rho

    def get_spline(self, number, parameters=0, combine="linear"):
        """Spline of ``get_rho`` (Hermite when the derivative is known)."""
        rho, deriv = self.get_rho(number, parameters, combine, do_deriv=True)
        points = self.get_rgrid(number).points
        if deriv is None:
            return CubicSpline(points, rho, True)
        return CubicHermiteSpline(points, rho, deriv, True)

    def compact(self, nel_lost):
        """Chop every element's records where its safe records have at most
        ``nel_lost`` electrons left outside (core/proatomdb.py:390-431)."""
        for number in self.get_numbers():
            npoint = 0
            for charge in self.get_charges(number, safe=True):
                r = self.get_record(number, charge)
                nel = r.pseudo_number - charge
                npoint = max(npoint, r.compute_radii([nel - nel_lost])[0][0] + 1)
            for charge in self.get_charges(number):
                self.get_record(number, charge).chop(npoint)
            old = self._rgrid_map[number]
            new = OneDGrid(old.points[:npoint], old.weights[:npoint])
            self._rgrid_map[number] = new
            logger.info("%4i   %5i -> %5i    %10.3e -> %10.3e", number, old.size, new.size, old.points[-1], new.points[-1])

    def normalize(self):
        """Scale every record so that ``rgrid.integrate(rho)`` (the plain
        radial sum, without 4 pi r^2, as core/proatomdb.py:433-447 computes
        it) equals its integer (pseudo) population."""
        for number in self.get_numbers():
            rgrid = self.get_rgrid(number)
            for charge in self.get_charges(number):
                r = self.get_record(number, charge)
                before = rgrid.integrate(r.rho)
                r.rho[:] *= (r.pseudo_number - charge) / before
                logger.info("%4i     %+3i    %15.8e   %15.8e", number, charge, before, rgrid.integrate(r.rho))
