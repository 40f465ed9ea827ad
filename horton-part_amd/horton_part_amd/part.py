"""Partitioning base classes with the reference's WPart API, on the GPU.

``Part``/``WPart`` restate ``core/base.py:35-692`` (geometry checks, grid
types, cache, ``do_populations``/``do_charges``/``do_moments``/``do_all``);
``AbstractISAWPart`` restates the outer loop of ``core/iterstock.py:61-193``.
The loop itself runs in ``libhortonpart.so`` (``hp_part_run``); this class
prepares the device description of the molecule and scheme, uploads the
initial parameters, and turns the device results, histories and status
bits back into the reference's cache entries, warnings and exceptions.
"""

import logging
import warnings

import numpy as np
from scipy.interpolate import CubicHermiteSpline, CubicSpline

from . import _native
from .cache import Cache, JustOnceClass, just_once
from .utils import DENSITY_CUTOFF, NEGATIVE_CUTOFF, POPULATION_CUTOFF, typecheck_geo

__all__ = ["Part", "WPart", "ProatomSplines", "AbstractISAWPart", "get_ncart_cumul", "get_npure_cumul"]


def get_ncart_cumul(lmax):
    """Number of Cartesian powers up to ``lmax`` (core/base.py:685-687)."""
    return (lmax + 1) * (lmax + 2) * (lmax + 3) // 6


def get_npure_cumul(lmax):
    """Number of pure functions up to ``lmax`` (core/base.py:690-692)."""
    return (lmax + 1) ** 2


class Part(JustOnceClass):
    """Geometry, grid and cache of a density partitioning (core/base.py:35-410)."""

    name = None

    def __init__(self, coordinates, numbers, pseudo_numbers, grid, moldens, spindens, local, lmax, logger, *args, **kwargs):
        super().__init__()
        natom, coordinates, numbers, pseudo_numbers = typecheck_geo(coordinates, numbers, pseudo_numbers)
        self._natom = natom
        self._coordinates = coordinates
        self._numbers = numbers
        self._pseudo_numbers = pseudo_numbers
        self._grid = grid
        self._moldens = moldens
        self._spindens = spindens
        self._local = local
        self._lmax = lmax
        self._cache = Cache()
        self.logger = logger
        if local:
            self._subgrids = self._grid.atgrids

    def __getitem__(self, key):
        return self.cache.load(key)

    def variables_stored_in_cache(self):
        return list(self.cache.iterkeys())

    @property
    def natom(self):
        return self._natom

    @property
    def coordinates(self):
        return self._coordinates

    @property
    def numbers(self):
        return self._numbers

    @property
    def pseudo_numbers(self):
        return self._pseudo_numbers

    @property
    def nelec(self):
        """Number of electrons: molecular-grid integral of the density."""
        return self.grid.integrate(self._moldens)

    @property
    def grid(self):
        return self.get_grid()

    @property
    def local(self):
        return self._local

    @property
    def lmax(self):
        return self._lmax

    @property
    def cache(self):
        return self._cache

    def clear(self):
        JustOnceClass.clear(self)
        self.cache.clear()

    def get_grid(self, index=None):
        if index is None or not self.local:
            return self._grid
        return self._subgrids[index]

    def to_atomic_grid(self, index, data):
        if index is None or not self.local:
            return data
        begin, end = self._grid.indices[index], self._grid.indices[index + 1]
        return data[begin:end]

    def get_moldens(self, index=None, output=None):
        result = self._moldens if (index is None or not self.local) else self.to_atomic_grid(index, self._moldens)
        if output is not None:
            output[:] = result
        return result

    def get_spindens(self, index=None, output=None):
        result = self._spindens if (index is None or not self.local) else self.to_atomic_grid(index, self._spindens)
        if output is not None:
            output[:] = result
        return result

    def get_wcor(self, index):
        return 1.0

    @just_once
    def do_partitioning(self):
        raise NotImplementedError

    @just_once
    def do_populations(self):
        """Atomic populations (core/base.py:277-300): no-op after the ISA loop."""
        populations, new = self.cache.load("populations", alloc=self.natom, tags="o")
        if new:
            self.do_partitioning()
            pseudo = self.cache.load("pseudo_populations", alloc=self.natom, tags="o")[0]
            for index in range(self.natom):
                grid = self.get_grid(index)
                at_weights = self.cache.load(f"at_weights_{index}")
                pseudo[index] = grid.integrate(at_weights, self.get_moldens(index))
            populations[:] = pseudo
            populations += self.numbers - self.pseudo_numbers

    @just_once
    def do_charges(self):
        """Atomic charges (core/base.py:302-311)."""
        charges, new = self._cache.load("charges", alloc=self.natom, tags="o")
        if new:
            self.do_populations()
            charges[:] = self.numbers - self._cache.load("populations")

    @just_once
    def do_spin_charges(self):
        """Spin charges (core/base.py:313-327) from the device AIM weights."""
        if self._spindens is not None:
            spin_charges, _ = self._cache.load("spin_charges", alloc=self.natom, tags="o")
            self.do_partitioning()
            for index in range(self.natom):
                grid = self.get_grid(index)
                at_weights = self.cache.load(f"at_weights_{index}")
                spin_charges[index] = grid.integrate(at_weights, self.get_spindens(index))

    @just_once
    def do_moments(self):
        """AIM Cartesian, pure and radial moments (core/base.py:329-402)."""
        self.do_partitioning()
        cart, pure, rad = self._device_moments()
        self.cache.dump("cartesian_multipoles", cart, tags="o")
        self.cache.dump("pure_multipoles", pure, tags="o")
        self.cache.dump("radial_moments", rad, tags="o")

    def _device_moments(self):
        raise NotImplementedError

    def do_all(self):
        """Run every ``do_*`` method; return the output keys (core/base.py:404-410)."""
        for attr_name in dir(self):
            attr = getattr(self, attr_name)
            if callable(attr) and attr_name.startswith("do_") and attr_name != "do_all":
                attr()
        return list(self.cache.iterkeys(tags="o"))


class WPart(Part):
    """Grid-based partitioning with grid types 1/2/3 (core/base.py:413-635)."""

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
        grid_type=1,
        density_cutoff=DENSITY_CUTOFF,
        negative_cutoff=NEGATIVE_CUTOFF,
        population_cutoff=POPULATION_CUTOFF,
        **kwargs,
    ):
        self._grid_type = grid_type
        assert grid_type in [1, 2, 3]
        self._on_molgrid = grid_type in [2, 3]
        self._only_use_molgrid = grid_type == 3
        local = not self._only_use_molgrid
        if local and grid.atgrids is None:
            raise ValueError("Atomic grids are discarded from molecular grid object, but are needed for local integrations.")
        if logger is None:
            logger = logging.getLogger(self.__class__.__name__)
        super().__init__(coordinates, numbers, pseudo_numbers, grid, moldens, spindens, local, lmax, logger)
        self.history_propars = []
        self.history_charges = []
        self.history_entropies = []
        self.history_changes = []
        self.history_time_update_at_weights = []
        self.history_time_update_propars = []
        self._density_cutoff = density_cutoff
        self._population_cutoff = population_cutoff
        self._negative_cutoff = negative_cutoff

    @property
    def grid_type(self):
        return self._grid_type

    @property
    def on_molgrid(self):
        return self._on_molgrid

    @property
    def only_use_molgrid(self):
        return self._only_use_molgrid

    @property
    def density_cutoff(self):
        return self._density_cutoff

    @property
    def population_cutoff(self):
        return self._population_cutoff

    @property
    def negative_cutoff(self):
        return self._negative_cutoff

    @property
    def radial_distances(self):
        """Distances of every molecular point to every atom (host arrays, on demand)."""
        if not hasattr(self, "_radial_distances_host"):
            self._radial_distances_host = [np.linalg.norm(self.grid.points - c, axis=1) for c in self.coordinates]
        return self._radial_distances_host

    # -- device description ---------------------------------------------------
    def _molecule_arrays(self):
        grid = self.grid
        if getattr(grid, "atgrids", None) is None:
            if not self.only_use_molgrid:
                raise ValueError("Atomic grids are discarded from molecular grid object, but are needed for local integrations.")
            return self._molgrid_block_arrays()
        indices = np.asarray(grid.indices, dtype=np.int64)
        shell_atom, shell_off, shell_r, shell_w, sph, atw = [], [np.zeros(1, np.int64)], [], [], [], []
        for a, atgrid in enumerate(grid.atgrids):
            if atgrid.size != indices[a + 1] - indices[a]:
                raise ValueError("molecular grid is not the concatenation of its atom grids")
            r = atgrid.rgrid.points
            aidx = np.asarray(atgrid.indices, dtype=np.int64)
            shell_atom.append(np.full(len(r), a, np.int32))
            shell_off.append(indices[a] + aidx[1:])
            shell_r.append(r)
            shell_w.append(atgrid.rgrid.weights)
            if hasattr(atgrid, "sph_weights"):
                sph.append(np.asarray(atgrid.sph_weights, float))
            else:
                per_shell = 4 * np.pi * r**2 * atgrid.rgrid.weights
                sph.append(atgrid.weights / np.repeat(per_shell, np.diff(aidx)))
            atw.append(np.asarray(atgrid.weights, float))
        shell_atom = np.concatenate(shell_atom)
        points = np.ascontiguousarray(grid.points, dtype=float)
        return {
            "natom": self.natom,
            "coordinates": np.ascontiguousarray(self.coordinates, dtype=float),
            "numbers": np.ascontiguousarray(self.numbers, dtype=np.int64),
            "pseudo_numbers": np.ascontiguousarray(self.pseudo_numbers, dtype=float),
            "npoint": int(points.shape[0]),
            "points": points,
            "density": np.ascontiguousarray(self._moldens, dtype=float),
            "mol_weights": np.ascontiguousarray(grid.weights, dtype=float),
            "atom_weights": np.ascontiguousarray(np.concatenate(atw)),
            "sph_weights": np.ascontiguousarray(np.concatenate(sph)),
            "nshell": len(shell_atom),
            "shell_atom": shell_atom,
            "shell_offsets": np.concatenate(shell_off),
            "shell_radius": np.ascontiguousarray(np.concatenate(shell_r)),
            "shell_rweight": np.ascontiguousarray(np.concatenate(shell_w)),
        }

    def _molgrid_block_arrays(self):
        """Device description of a molecular grid without atom grids
        (grid_type 3): the points in consecutive blocks of at most 1024, the
        blocks split among the atoms in order.  On the molecular grid a block
        plays the part of a radial shell only as a unit of work: the loop sums
        over all points with the molecular weights, so the split changes no
        result."""
        grid = self.grid
        points = np.ascontiguousarray(grid.points, dtype=float)
        npoint, natom = points.shape[0], self.natom
        size = max(256, -(-npoint // (512 * natom)))
        if size > 1024:
            raise NotImplementedError("molecular grids of more than 512K points per atom need atom grids on the device")
        nblock = max(natom, -(-npoint // size))
        if nblock > npoint:
            raise ValueError("the molecular grid has fewer points than atoms")
        offsets = (np.arange(nblock + 1, dtype=np.int64) * npoint) // nblock
        owner = (np.arange(nblock, dtype=np.int64) * natom) // nblock
        weights = np.ascontiguousarray(grid.weights, dtype=float)
        return {
            "natom": natom,
            "coordinates": np.ascontiguousarray(self.coordinates, dtype=float),
            "numbers": np.ascontiguousarray(self.numbers, dtype=np.int64),
            "pseudo_numbers": np.ascontiguousarray(self.pseudo_numbers, dtype=float),
            "npoint": int(npoint),
            "points": points,
            "density": np.ascontiguousarray(self._moldens, dtype=float),
            "mol_weights": weights,
            "atom_weights": weights,
            "sph_weights": np.zeros(npoint),
            "nshell": int(nblock),
            "shell_atom": owner.astype(np.int32),
            "shell_offsets": offsets,
            "shell_radius": np.zeros(nblock),
            "shell_rweight": np.zeros(nblock),
        }


class ProatomSplines:
    """Pro-atom splines on the host (core/stockholder.py:177-349, 386-398):
    the reference's API to a pro-atom as a radial spline, for schemes whose
    ``get_proatom_rho`` returns values on ``get_rgrid(index)``.  The
    partitioning itself evaluates pro-atoms on the device."""

    def fix_proatom_rho(self, index, rho, deriv):
        """Negative densities become zero and the derivative is dropped
        (core/stockholder.py:201-225)."""
        if rho.min() < 0:
            rgrid = self.get_rgrid(index)
            original = rgrid.integrate(rho)
            rho[rho < 0] = 0.0
            deriv = None
            self.logger.info("                Pro-atom not positive everywhere. Lost %.1e electrons" % (rgrid.integrate(rho) - original))
        return rho, deriv

    def get_proatom_spline(self, index, *args, **kwargs):
        """CubicHermiteSpline through (rho, deriv), or CubicSpline through rho
        when there is no derivative, on the pro-atom's radial grid
        (core/stockholder.py:227-264)."""
        rho, deriv = self.get_proatom_rho(index, *args, **kwargs)
        rho, deriv = self.fix_proatom_rho(index, rho, deriv)
        points = self.get_rgrid(index).points
        if deriv is None:
            return CubicSpline(points, rho, True)
        return CubicHermiteSpline(points, rho, deriv, True)

    def eval_spline(self, index, spline, output, grid, label="noname"):
        """``spline`` at the distances of ``grid.points`` from atom ``index``
        (core/stockholder.py:266-300)."""
        r = np.linalg.norm(grid.points - self.coordinates[index], axis=1)
        output[:] = spline(r)

    def eval_proatom(self, index, output, grid):
        """The pro-atom spline at ``grid``'s points plus 1e-100
        (core/stockholder.py:302-349)."""
        spline = self.get_proatom_spline(index)
        output[:] = 0.0
        self.eval_spline(index, spline, output, grid, label="proatom")
        output += 1e-100
        assert np.isfinite(output).all()

    @just_once
    def do_prosplines(self):
        """Cache the pro-atom density splines (core/stockholder.py:386-398)."""
        self.do_partitioning()
        for index in range(self.natom):
            key = ("spline_prodensity", index)
            if key not in self.cache:
                self.cache.dump(key, self.get_proatom_spline(index), tags="o")


class AbstractISAWPart(WPart):
    """Iterative stockholder outer loop on the GPU (core/iterstock.py:61-193).

    Extra keyword arguments (defaults reproduce the reference):
      precision: "fp64" (default) or "mixed" (FP32 grid arithmetic, then FP64 polish)
      n_polish: minimum FP64 iterations after the mixed stage (default 1)
      mixed_switch: the FP32 stage ends when the change is below
        max(1.01 threshold, mixed_switch) (default 1e-5); its last iteration
        is redone in FP64 and FP64 iterations run until the change is below
        threshold (grid_type 2: max(threshold, mixed_switch))
      mixed_refresh: every this many FP32-stage iterations one runs in FP64
        and refreshes the per-shell correction FP64 - FP32 of the spherical
        averages (default 0: no correction)
      cutoff_radius: pro-atom tails beyond this distance are zero (default None)
      full_sum: evaluate every pro-atom on every point (default False: MBIS
        and GISA skip, per radial shell, atoms whose pro-atom is bounded there
        below 2^-64 of the promolecule, which changes no result beyond
        rounding; LISA always sums every atom)
      device: GPU ordinal (default 0)
    """

    _scheme_name = None
    _solver_name = None

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
        grid_type=1,
        precision="fp64",
        n_polish=1,
        mixed_switch=1e-5,
        mixed_refresh=0,
        cutoff_radius=None,
        full_sum=False,
        device=0,
        **kwargs,
    ):
        self._threshold = threshold
        self._inner_threshold = inner_threshold if inner_threshold < threshold else threshold
        self._maxiter = maxiter
        if precision not in ("fp64", "mixed"):
            raise ValueError("precision must be 'fp64' or 'mixed'")
        self._precision = precision
        self._n_polish = n_polish
        self._mixed_switch = mixed_switch
        self._mixed_refresh = mixed_refresh
        self._cutoff_radius = cutoff_radius
        self._full_sum = bool(full_sum)
        self._device = device
        self._dev = None
        super().__init__(coordinates, numbers, pseudo_numbers, grid, moldens, spindens, lmax, logger, grid_type=grid_type, **kwargs)

    def _scheme_arrays(self):
        """Scheme-specific entries of ``hp_scheme_desc`` (subclasses)."""
        raise NotImplementedError

    def _base_scheme_arrays(self):
        return {
            "grid_type": self.grid_type,
            "threshold": float(self._threshold),
            "inner_threshold": float(self._inner_threshold),
            "maxiter": int(self._maxiter),
            "density_cutoff": float(self.density_cutoff),
            "negative_cutoff": float(self.negative_cutoff),
            "population_cutoff": float(self.population_cutoff),
            "cutoff_radius": float(self._cutoff_radius) if self._cutoff_radius else 0.0,
            "newton_tau": 1.0,
            "newton_linspace": 20,
            "check_mono": 0,
            "mixed_switch": float(self._mixed_switch),
            "mixed_refresh": int(self._mixed_refresh),
            "full_sum": int(self._full_sum),
        }

    def _init_propars(self):
        raise NotImplementedError

    def _ensure_device(self):
        if self._dev is None:
            self._dev = _native.DevicePartition(self._molecule_arrays(), self._scheme_arrays(), device=self._device)
        return self._dev

    def _raise_status(self, status_all, status_last, rc):
        flags = int(np.bitwise_or.reduce(status_all)) if len(status_all) else 0
        if flags & _native.ST_INNER_NOT_CONVERGED:
            self.logger.warning("Inner iteration not converged (atoms %s)." % np.flatnonzero(status_all & 1).tolist())
        if flags & _native.ST_POP_MISMATCH:
            self.logger.warning("The sum of pro-atom parameters is not equal to the atomic population.")
        if flags & _native.ST_NEG_PARS:
            warnings.warn("WARNING: Not all pro-atom parameters are positive!")
        if rc == _native.HP_E_SOLVER:
            last = int(np.bitwise_or.reduce(status_last))
            if last & _native.ST_SINGULAR:
                raise np.linalg.LinAlgError("Matrix is singular.")
            if last & _native.ST_NONFINITE:
                raise ValueError("array must not contain infs or NaNs")
            if last & _native.ST_NEG_DENSITY:
                raise RuntimeError("Negative pro-atom density found!")
            if last & _native.ST_LINESEARCH:
                raise RuntimeError("Line search failed!")
            if last & _native.ST_NEWTON_MAXITER:
                raise RuntimeError("Inner loop: Newton does not converge!")
            if last & _native.ST_NOT_MONOTONIC:
                raise RuntimeError("Pro-atom density should be monotonically decreasing.")
            if last & _native.ST_QP_FAILED:
                raise RuntimeError("QP solver failed.")
            raise RuntimeError("a per-atom solver failed on the device")
        if rc != 0:
            _native.check(rc)

    # -- per-atom solvers that run in the caller ------------------------------
    def _host_solver(self):
        """The caller's per-atom solver (a callable), or None when the device
        solvers run the whole loop (subclasses)."""
        return None

    def _solve_atom_on_host(self, iatom, rho, points, weights, propars):
        """New parameters of atom ``iatom`` from the caller's solver (subclasses)."""
        raise NotImplementedError

    def _host_basis_radii(self, iatom):
        """Radii on which atom ``iatom``'s per-atom problem lives: its radial
        grid (grid_type 1) or its distances to every molecular point."""
        return self.radial_distances[iatom] if self.on_molgrid else self.get_grid(iatom).rgrid.points

    def _run_with_host_solver(self, dev):
        """The outer loop of core/iterstock.py:159-193 with the weights pass on
        the device (``hp_part_weights``) and each atom's update
        (gisa.py:251-318, alisa.py:1304-1335) by the caller's solver: the
        device supplies the spherical averages (grid_type 1) or the AIM weights
        (grid_type 2 and 3) of the pass, the solver returns the parameters."""
        propars = self._init_propars()
        self.cache.dump("propars", propars, tags="o")
        charges = np.zeros(self.natom)
        self.logger.info("Iteration       Change      Entropy")
        shell0 = None
        if not self.on_molgrid:
            shell0 = np.concatenate([[0], np.cumsum([ag.rgrid.size for ag in self.grid.atgrids])])
        counter = 0
        while True:
            counter += 1
            old = propars.copy()
            dev.set_propars(propars)
            rhobar, entropy = dev.weights(spherical=not self.on_molgrid)
            for iatom in range(self.natom):
                sl = slice(self._ranges[iatom], self._ranges[iatom + 1])
                if self.on_molgrid:
                    rho = dev.at_weights(iatom, int(self.grid.size)) * self._moldens
                    points, weights = self.grid.points, self.grid.weights
                else:
                    rgrid = self.get_grid(iatom).rgrid
                    rho = rhobar[shell0[iatom] : shell0[iatom + 1]].copy()
                    points = rgrid.points
                    weights = 4 * np.pi * points**2 * rgrid.weights
                    self.cache.dump(f"radial_points_{iatom}", points, tags="o")
                    self.cache.dump(f"spherical_average_{iatom}", rho, tags="o")
                    self.cache.dump(f"radial_weights_{iatom}", rgrid.weights, tags="o")
                propars[sl] = self._solve_atom_on_host(iatom, rho, points, weights, propars[sl].copy())
                charges[iatom] = self.pseudo_numbers[iatom] - np.einsum("i,i", weights, rho)
            self.history_propars.append(propars.copy())
            self.history_charges.append(charges.copy())
            self.history_entropies.append(entropy)
            change = self.compute_change(propars, old)
            self.history_changes.append(change)
            self.logger.info("%9i   %10.5e   %10.5e" % (counter, change, entropy))
            if change < self._threshold or counter >= self._maxiter:
                break
        self.cache.dump("charges", charges, tags="o")
        self.cache.dump("niter", counter, tags="o")
        self.cache.dump("change", change, tags="o")
        self._inner_niter = None
        self._register_lazy(dev, rhobar=rhobar)
        self._finalize_propars()

    def compute_change(self, propars1, propars2):
        """Root-mean-square change of the pro-atoms (core/iterstock.py:32-45)."""
        msd = 0.0
        for iatom in range(self.natom):
            delta = self.get_proatom_rho(iatom, propars1)[0] - self.get_proatom_rho(iatom, propars2)[0]
            if self.on_molgrid:
                msd += self.grid.integrate(delta, delta)
            else:
                rgrid = self.get_grid(iatom).rgrid
                msd += rgrid.integrate(4 * np.pi * rgrid.points**2, delta, delta)
        return np.sqrt(msd)

    @just_once
    def do_partitioning(self):
        """Outer fixed-point loop (core/iterstock.py:159-193) on the GPU."""
        dev = self._ensure_device()
        if self._host_solver() is not None:
            self._run_with_host_solver(dev)
            return
        propars = self._init_propars()
        dev.set_propars(propars)
        self.logger.info("Iteration       Change      Entropy")
        rc = dev.run(self._precision, self._n_polish)
        inner, last, allst = dev.status()
        self._inner_niter = inner
        if dev.niter > 0:
            hp, hc, he, hg = dev.history()
            self.history_propars = list(hp)
            self.history_charges = list(hc)
            self.history_entropies = list(he)
            self.history_changes = list(hg)
            for i in range(dev.niter):
                self.logger.info("%9i   %10.5e   %10.5e" % (i + 1, hg[i], he[i]))
        self._raise_status(allst, last, rc)
        new_propars, charges, change, _ = dev.results()
        propars[:] = new_propars
        self.cache.dump("propars", propars, tags="o")
        self.cache.dump("charges", charges, tags="o")
        self.cache.dump("niter", dev.niter, tags="o")
        self.cache.dump("change", change, tags="o")
        self._register_lazy(dev)
        self._finalize_propars()

    def _register_lazy(self, dev, rhobar=None):
        grid = self.grid
        for a in range(self.natom):
            size = int(grid.indices[a + 1] - grid.indices[a]) if not self.on_molgrid else int(grid.size)
            self.cache.add_lazy(f"at_weights_{a}", (lambda a=a, size=size: dev.at_weights(a, size)))
        self.cache.add_lazy("promoldens", dev.promoldens)
        if self.on_molgrid:
            return  # spherical averages exist on atom grids only (mbis.py:176-198)
        if rhobar is None:
            rhobar = dev.spherical_averages()
        s = 0
        for a, atgrid in enumerate(grid.atgrids):
            nr = atgrid.rgrid.size
            self.cache.dump(f"radial_points_{a}", atgrid.rgrid.points, tags="o")
            self.cache.dump(f"spherical_average_{a}", rhobar[s : s + nr].copy(), tags="o")
            self.cache.dump(f"radial_weights_{a}", atgrid.rgrid.weights, tags="o")
            s += nr

    def _finalize_propars(self):
        """finalize_propars (core/iterstock.py:48-58)."""
        charges = self.cache.load("charges")
        self.cache.dump("history_propars", np.array(self.history_propars), tags="o")
        self.cache.dump("history_charges", np.array(self.history_charges), tags="o")
        self.cache.dump("history_entropies", np.array(self.history_entropies), tags="o")
        self.cache.dump("history_changes", np.array(self.history_changes), tags="o")
        self.cache.dump("populations", self.numbers - charges, tags="o")
        self.cache.dump("pseudo_populations", self.pseudo_numbers - charges, tags="o")
        dev = self._dev
        t_w, _ = dev.kernel_time(_native.KERNEL_WEIGHTS)
        t_i, _ = dev.kernel_time(_native.KERNEL_INNER)
        self.cache.dump("time_update_at_weights", t_w * 1e-3, tags="o")
        self.cache.dump("time_update_propars", t_i * 1e-3, tags="o")

    def _device_moments(self):
        return self._dev.moments(self.lmax)

    @property
    def inner_niter(self):
        """Per-atom inner iteration counts of the last outer iteration."""
        return getattr(self, "_inner_niter", None)

    def close(self):
        """Release the device memory held by this partitioning."""
        if self._dev is not None:
            self._dev.close()
