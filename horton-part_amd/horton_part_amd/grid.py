"""Molecular integration grids with the interface the partitioning path uses.

The reference builds its grids with the third-party qc-grid package and uses
only a small surface of it (call sites ``core/base.py:26,264,298,374-402``,
``core/stockholder.py:134,150``, ``core/iterstock.py:41-44``,
``mbis.py:176-184``, ``gisa.py:283-293``; the test set-up is
``tests/test_wpart.py:47-51``).  This module restates that surface:

* ``UniformInteger`` / ``ExpRTransform`` / ``PowerRTransform`` / ``OneDGrid``:
  radial rules.  ``ExpRTransform(rmin, rmax, b)`` maps ``x`` to
  ``rmin*exp(x*ln(rmax/rmin)/b)``; the radii of the reference's water fixture
  agree with it to 2e-15.  ``PowerRTransform(rmin, rmax, b)`` maps ``x`` to
  ``rmin*(x+1)**p`` with ``p = ln(rmax/rmin)/ln(b+1)`` (the radial grids of
  the pro-atom database files, ``tests/common.py:78-110``).
* ``AtomGrid``: an atom-centred product grid with points ``c + r_i u_j`` in
  shell-major order and weights ``w_i * r_i**2 * 4*pi*wl_j`` (``sum wl = 1``).
* ``MolGrid``: the concatenation of atom grids (atom-major) with molecular
  weights ``atomgrid_weight * aim_weight`` where ``aim_weight`` comes from
  Becke's fuzzy-cell partitioning.
* ``BeckeWeights``: A. D. Becke, J. Chem. Phys. 88 (1988) 2547, with the
  heteronuclear size adjustment and the Bragg-Slater radii the reference keeps
  in ``data/constants.yaml`` (``radius_becke``).

``integrate(*f)`` is ``sum_p w_p prod_k f_k(p)``; ``spherical_average(f)`` is a
cubic spline through ``fbar_i = sum_j wl_j f_ij``; ``moments`` returns
Cartesian, pure (real regular solid harmonics, Racah normalised) or radial
moments in HORTON order.
"""

import numpy as np
from scipy.interpolate import CubicSpline

from .lebedev import LEBEDEV_DEGREES, LEBEDEV_NPOINTS, lebedev_rule

__all__ = [
    "OneDGrid",
    "UniformInteger",
    "ExpRTransform",
    "PowerRTransform",
    "AtomGrid",
    "MolGrid",
    "BeckeWeights",
    "LEBEDEV_DEGREES",
    "cartesian_powers",
    "solid_harmonics",
]


class OneDGrid:
    """A one-dimensional quadrature: ``points`` and ``weights``."""

    def __init__(self, points, weights):
        self._points = np.asarray(points, dtype=float)
        self._weights = np.asarray(weights, dtype=float)
        if self._points.shape != self._weights.shape or self._points.ndim != 1:
            raise ValueError("points and weights must be 1D arrays of equal length")

    @property
    def points(self):
        return self._points

    @property
    def weights(self):
        return self._weights

    @property
    def size(self):
        return self._points.size

    def __len__(self):
        return self._points.size

    def integrate(self, *value_arrays):
        """Return ``sum_i w_i prod_k f_k(x_i)``."""
        prod = self._weights
        for v in value_arrays:
            prod = prod * v
        return float(np.sum(prod))


class UniformInteger(OneDGrid):
    """The integers ``0 .. n-1`` with unit weights."""

    def __init__(self, npoints):
        super().__init__(np.arange(npoints, dtype=float), np.ones(npoints))


class ExpRTransform:
    """Exponential radial map ``r(x) = rmin * exp(x * ln(rmax/rmin) / b)``."""

    def __init__(self, rmin, rmax, b=None):
        if rmin <= 0 or rmax <= rmin:
            raise ValueError("ExpRTransform needs 0 < rmin < rmax")
        self.rmin = float(rmin)
        self.rmax = float(rmax)
        self.b = b

    def transform_1d_grid(self, oned):
        b = self.b if self.b is not None else oned.points[-1]
        alpha = np.log(self.rmax / self.rmin) / b
        r = self.rmin * np.exp(oned.points * alpha)
        return OneDGrid(r, oned.weights * r * alpha)


class PowerRTransform:
    """Power radial map ``r(x) = rmin * (x + 1)**p``, ``p = ln(rmax/rmin) / ln(b + 1)``:
    on ``UniformInteger(b + 1)`` the radii run from ``rmin`` to ``rmax``."""

    def __init__(self, rmin, rmax, b=None):
        if rmin <= 0 or rmax <= rmin:
            raise ValueError("PowerRTransform needs 0 < rmin < rmax")
        self.rmin = float(rmin)
        self.rmax = float(rmax)
        self.b = b

    def transform_1d_grid(self, oned):
        b = self.b if self.b is not None else oned.points[-1]
        power = np.log(self.rmax / self.rmin) / np.log(b + 1.0)
        r = self.rmin * np.power(oned.points + 1.0, power)
        return OneDGrid(r, oned.weights * power * self.rmin * np.power(oned.points + 1.0, power - 1.0))


def exp_radial_grid(nrad, rmin=5e-4, rmax=2e1):
    """The reference's default radial grid: ExpRTransform over UniformInteger."""
    return ExpRTransform(rmin, rmax, nrad - 1).transform_1d_grid(UniformInteger(nrad))


def cartesian_powers(lmax):
    """Cartesian exponents ``(nx, ny, nz)`` up to ``lmax`` in HORTON order."""
    out = []
    for l in range(lmax + 1):
        for nx in range(l, -1, -1):
            for ny in range(l - nx, -1, -1):
                out.append((nx, ny, l - nx - ny))
    return out


def solid_harmonics(lmax, d):
    """Real regular solid harmonics up to ``lmax`` at displacement vectors ``d``.

    Racah normalisation (``R_l0 = r^l P_l(cos theta)``), no Condon-Shortley
    phase, order ``c_l0, c_l1, s_l1, c_l2, s_l2, ...`` for each ``l``.
    Built with the standard recurrences for ``r^l P_l^m`` so any ``lmax`` works.
    Returns an array of shape ((lmax+1)**2, npoint).
    """
    d = np.atleast_2d(d)
    x, y, z = d[:, 0], d[:, 1], d[:, 2]
    r2 = x * x + y * y + z * z
    npt = d.shape[0]
    # Q[l][m] = r^l P_l^m(cos t) / (x + i y)^m style helpers via the
    # classic recursion on the unnormalised associated Legendre terms.
    # C[m] + i S[m] = (x + i y)^m
    cm = [np.ones(npt)]
    sm = [np.zeros(npt)]
    for m in range(1, lmax + 1):
        cm.append(cm[-1] * x - sm[-1] * y)
        sm.append(cm[-2] * y + sm[-1] * x)
    # Pbar[l][m] = r^(l-m) * d^m P_l(t)/dt^m  (polynomial in z and r2)
    P = [[None] * (lmax + 1) for _ in range(lmax + 1)]
    for m in range(lmax + 1):
        # start: P[m][m] = (2m-1)!!
        dfac = 1.0
        for k in range(1, 2 * m, 2):
            dfac *= k
        P[m][m] = np.full(npt, dfac)
        if m + 1 <= lmax:
            P[m + 1][m] = (2 * m + 1) * z * P[m][m]
        for l in range(m + 2, lmax + 1):
            P[l][m] = ((2 * l - 1) * z * P[l - 1][m] - (l + m - 1) * r2 * P[l - 2][m]) / (l - m)
    out = np.zeros(((lmax + 1) ** 2, npt))
    i = 0
    from math import factorial

    for l in range(lmax + 1):
        out[i] = P[l][0]
        i += 1
        for m in range(1, l + 1):
            f = np.sqrt(2.0 * factorial(l - m) / factorial(l + m))
            out[i] = f * P[l][m] * cm[m]
            out[i + 1] = f * P[l][m] * sm[m]
            i += 2
    return out


def _moments(points, weights, orders, centers, func_vals, type_mom):
    centers = np.atleast_2d(np.asarray(centers, dtype=float))
    res = []
    for c in centers:
        d = points - c
        wf = weights * func_vals
        if type_mom == "cartesian":
            vals = [np.sum(wf * d[:, 0] ** a * d[:, 1] ** b * d[:, 2] ** cc) for a, b, cc in cartesian_powers(orders)]
        elif type_mom == "pure":
            vals = list(solid_harmonics(orders, d) @ wf)
        elif type_mom == "radial":
            r = np.sqrt(np.sum(d * d, axis=1))
            vals = [np.sum(wf * r**n) for n in range(orders + 1)]
        else:
            raise ValueError(f"unknown type_mom {type_mom}")
        res.append(vals)
    return np.array(res).T


class AtomGrid:
    """Atom-centred product grid: radial rule x one Lebedev rule per shell."""

    def __init__(self, rgrid, *, degrees=None, sizes=None, center=None, rotate=False):
        if rotate:
            raise NotImplementedError("rotated angular grids are not supported")
        self._rgrid = rgrid
        nrad = rgrid.size
        if sizes is None and degrees is None:
            raise ValueError("either sizes or degrees must be given")
        if sizes is None:
            degrees = np.broadcast_to(np.asarray(degrees), (nrad,)) if np.ndim(degrees) == 0 or len(degrees) == 1 else np.asarray(degrees)
            sizes = [LEBEDEV_DEGREES[int(d)] for d in degrees]
        else:
            sizes = list(np.broadcast_to(np.asarray(sizes), (nrad,))) if np.ndim(sizes) == 0 or len(sizes) == 1 else list(sizes)
        if len(sizes) != nrad:
            raise ValueError("one angular size per radial point is required")
        self._sizes = np.array([int(s) for s in sizes])
        self._degrees = np.array([LEBEDEV_NPOINTS[int(s)] for s in self._sizes])
        self._center = np.zeros(3) if center is None else np.asarray(center, dtype=float)
        self._indices = np.concatenate([[0], np.cumsum(self._sizes)])
        r = np.asarray(rgrid.points, dtype=float)
        w = np.asarray(rgrid.weights, dtype=float)
        # per shell: w_i 4 pi r_i^2 (the radial factor of the product weight)
        shell_factor = w * r**2 * 4.0 * np.pi
        if (self._sizes == self._sizes[0]).all():
            # one Lebedev rule for every shell: all shells at once
            u, wl = lebedev_rule(int(self._sizes[0]))
            self._points = (self._center + r[:, None, None] * u[None, :, :]).reshape(-1, 3)
            self._weights = (shell_factor[:, None] * wl[None, :]).ravel()
            self._sph_weights = np.tile(wl, nrad)
        else:
            pts, wts, sph = [], [], []
            for i, n in enumerate(self._sizes):
                u, wl = lebedev_rule(int(n))
                pts.append(self._center + r[i] * u)
                wts.append(shell_factor[i] * wl)
                sph.append(wl)
            self._points = np.concatenate(pts)
            self._weights = np.concatenate(wts)
            self._sph_weights = np.concatenate(sph)

    @property
    def rgrid(self):
        return self._rgrid

    @property
    def center(self):
        return self._center

    @property
    def points(self):
        return self._points

    @property
    def weights(self):
        return self._weights

    @property
    def sph_weights(self):
        """Normalised angular weight of each point within its shell (sum 1)."""
        return self._sph_weights

    @property
    def indices(self):
        """Shell boundaries: shell ``i`` holds points ``indices[i]:indices[i+1]``."""
        return self._indices

    @property
    def degrees(self):
        return self._degrees

    @property
    def sizes(self):
        return self._sizes

    @property
    def l_max(self):
        return int(self._degrees.max())

    @property
    def size(self):
        return self._points.shape[0]

    def integrate(self, *value_arrays):
        prod = self._weights
        for v in value_arrays:
            prod = prod * v
        return float(np.sum(prod))

    def spherical_average(self, func_vals):
        """Cubic spline through the shell averages ``sum_j wl_j f_ij``."""
        func_vals = np.asarray(func_vals)
        if func_vals.shape != (self.size,):
            raise ValueError("func_vals must have one value per grid point")
        avg = np.add.reduceat(func_vals * self._sph_weights, self._indices[:-1])
        return CubicSpline(self._rgrid.points, avg)

    def moments(self, orders, centers, func_vals, type_mom="cartesian"):
        return _moments(self._points, self._weights, orders, centers, func_vals, type_mom)


class BeckeWeights:
    """Becke fuzzy-cell atom weights with size adjustment (order-3 step)."""

    def __init__(self, radii=None, order=3, device=None):
        self._radii = radii
        self._order = order
        self._device = device

    def _radius_table(self, atnums):
        if self._radii is not None:
            return np.array([self._radii[int(z)] for z in atnums], float)
        from .utils import radius_becke

        out = []
        for z in atnums:
            r = radius_becke.get(int(z))
            out.append(1.0 if r is None else r)
        return np.array(out, float)

    def generate_weights(self, points, atcoords, atnums, indices=None, chunk=4096):
        """Weight of each point for the atom that owns it (``indices`` split)."""
        atcoords = np.asarray(atcoords, float)
        natom = len(atcoords)
        if natom == 1:
            return np.ones(len(points))
        radii = self._radius_table(atnums)
        if self._device is not None:
            if self._order != 3 or indices is None:
                raise NotImplementedError("the device Becke weights use order 3 and atom point offsets")
            from . import _native

            return _native.becke_weights(atcoords, radii, points, indices, device=self._device)
        chi = radii[:, None] / radii[None, :]
        u = (chi - 1.0) / (chi + 1.0)
        with np.errstate(divide="ignore", invalid="ignore"):
            a = np.where(np.abs(u * u - 1.0) > 0, u / (u * u - 1.0), 0.0)
        a = np.clip(a, -0.5, 0.5)
        np.fill_diagonal(a, 0.0)
        rij = np.linalg.norm(atcoords[:, None, :] - atcoords[None, :, :], axis=2)
        np.fill_diagonal(rij, 1.0)
        owner = np.zeros(len(points), dtype=np.int64)
        if indices is not None:
            for i in range(natom):
                owner[indices[i] : indices[i + 1]] = i
        out = np.empty(len(points))
        for s in range(0, len(points), chunk):
            p = points[s : s + chunk]
            d = np.linalg.norm(p[:, None, :] - atcoords[None, :, :], axis=2)  # (n, A)
            mu = (d[:, :, None] - d[:, None, :]) / rij[None]
            nu = mu + a[None] * (1.0 - mu * mu)
            f = nu
            for _ in range(self._order):
                f = 1.5 * f - 0.5 * (f * f * f)
            sfun = 0.5 * (1.0 - f)
            idx = np.arange(natom)
            sfun[:, idx, idx] = 1.0
            cell = np.prod(sfun, axis=2)  # (n, A)
            tot = cell.sum(axis=1)
            own = owner[s : s + chunk]
            out[s : s + chunk] = cell[np.arange(len(p)), own] / tot
        return out


class MolGrid:
    """Concatenation of atom grids with Becke-weighted molecular weights."""

    def __init__(self, atnums, atgrids, aim_weights, store=True):
        self._atnums = np.asarray(atnums)
        self._atgrids = list(atgrids) if store else None
        self._atgrids_all = list(atgrids)
        self._indices = np.concatenate([[0], np.cumsum([g.size for g in atgrids])]).astype(np.int64)
        self._points = np.concatenate([g.points for g in atgrids])
        self._atcoords = np.array([g.center for g in atgrids])
        if isinstance(aim_weights, BeckeWeights):
            self._aim_weights = aim_weights.generate_weights(
                self._points, self._atcoords, self._atnums, indices=self._indices
            )
        else:
            self._aim_weights = np.asarray(aim_weights, float)
        self._atweights = np.concatenate([g.weights for g in atgrids])
        self._weights = self._atweights * self._aim_weights

    @classmethod
    def from_size(cls, atnums, atcoords, size, rgrid, aim_weights, rotate=False, store=True):
        atgrids = [AtomGrid(rgrid, sizes=[size], center=c, rotate=rotate) for c in np.asarray(atcoords, float)]
        return cls(atnums, atgrids, aim_weights, store=store)

    @property
    def points(self):
        return self._points

    @property
    def weights(self):
        return self._weights

    @property
    def aim_weights(self):
        return self._aim_weights

    @property
    def atweights(self):
        """Atom-grid weights of every point, before the Becke factor."""
        return self._atweights

    @property
    def indices(self):
        return self._indices

    @property
    def atgrids(self):
        return self._atgrids

    @property
    def atcoords(self):
        return self._atcoords

    @property
    def atnums(self):
        return self._atnums

    @property
    def size(self):
        return self._points.shape[0]

    def get_atomic_grid(self, index):
        return self._atgrids_all[index]

    def integrate(self, *value_arrays):
        prod = self._weights
        for v in value_arrays:
            prod = prod * v
        return float(np.sum(prod))

    def moments(self, orders, centers, func_vals, type_mom="cartesian"):
        return _moments(self._points, self._weights, orders, centers, func_vals, type_mom)
