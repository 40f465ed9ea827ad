"""Gaussian Iterative Stockholder Analysis (GISA) on the GPU.

API of ``gisa.py`` in the reference: ``init_propars`` (68-88) and
``GaussianISAWPart`` (109-345).  The per-atom quadratic programme of
``opt_propars_qp_interface`` (348-421) -- analytic overlap matrix ``P``,
grid vector ``q`` and population, constraints ``x >= 0`` and
``sum x = pop`` -- is solved on the device by an active-set method in
``k_gisa_qp``; the problem is strictly convex, so its unique optimum is the
one quadprog returns in the reference.
"""

import numpy as np

from .basis import ExpBasisFuncHelper
from .part import AbstractISAWPart

__all__ = ["GaussianISAWPart", "init_propars", "gaussian_overlap_matrix"]


def gaussian_overlap_matrix(alphas):
    """``P`` of gisa.py:389-395 for exponents ``alphas``."""
    alphas = np.asarray(alphas, float)
    P = 2 / np.pi**1.5 * (alphas[:, None] * alphas[None, :]) ** 1.5 / (alphas[:, None] + alphas[None, :]) ** 1.5
    return (P + P.T) / 2


def init_propars(part):
    """Initial coefficients normalised to the molecular-grid electron count (gisa.py:68-88)."""
    part._ranges = [0]
    part._nshells = []
    for iatom in range(part.natom):
        nshell = part.bs_helper.get_nshell(part.numbers[iatom])
        part._ranges.append(part._ranges[-1] + nshell)
        part._nshells.append(nshell)
    propars = np.ones(part._ranges[-1])
    for iatom in range(part.natom):
        inits = np.array(part.bs_helper.get_initial(part.numbers[iatom]), dtype=float)
        inits[inits < 1e-4] = 1e-4
        propars[part._ranges[iatom] : part._ranges[iatom + 1]] = inits / np.sum(inits) * part.pseudo_numbers[iatom]
    propars[:] = propars / np.sum(propars) * part.nelec
    part.initial_propars_modified = propars.copy()
    return propars


class GaussianISAWPart(AbstractISAWPart):
    """Gaussian ISA (gisa.py:109-345) with the QP solved on the GPU."""

    name = "gisa"

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
        radius_cutoff=np.inf,
        solver="quadprog",
        solver_options=None,
        grid_type=1,
        **kwargs,
    ):
        self._solver = solver
        self._solver_options = solver_options or {}
        self._bs_helper = None
        self._ranges = None
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

    @property
    def bs_helper(self):
        if self._bs_helper is None:
            self._bs_helper = ExpBasisFuncHelper.from_function_type()
        return self._bs_helper

    def get_rgrid(self, index):
        if self.only_use_molgrid:
            raise NotImplementedError
        return self.get_grid(index).rgrid

    def get_proatom_rho(self, iatom, propars=None, **kwargs):
        """Pro-atom density on the radial grid (gisa.py:42-65)."""
        if propars is None:
            propars = self.cache.load("propars")
        pops = propars[self._ranges[iatom] : self._ranges[iatom + 1]]
        r = self.radial_distances[iatom] if self.on_molgrid else self.get_rgrid(iatom).points
        return self.bs_helper.compute_proatom_dens(self.numbers[iatom], pops, r, 1)

    def _init_propars(self):
        return init_propars(self)

    def _host_solver(self):
        """A solver callable (gisa.py:321-331) runs in the caller; the device
        supplies each outer iteration's weights pass."""
        return self._solver if callable(self._solver) else None

    def _host_bs_funcs(self, iatom):
        """``bs_funcs_{iatom}`` (gisa.py:91-106): the atom's shells at unit
        population on its radial grid (grid_type 1) or molecular-grid distances."""
        key = f"bs_funcs_{iatom}"
        if key not in self.cache:
            number = self.numbers[iatom]
            r = self._host_basis_radii(iatom)
            rows = [self.bs_helper.compute_proshell_dens(number, k, 1.0, r) for k in range(self.bs_helper.get_nshell(number))]
            self.cache.dump(key, np.array(rows))
        return self.cache.load(key)

    def _solve_atom_on_host(self, iatom, rho, points, weights, propars):
        """GaussianISAWPart._opt_propars with a callable (gisa.py:320-331)."""
        helper = self.bs_helper
        alphas = helper.get_exponent(self.numbers[iatom]) if isinstance(helper, ExpBasisFuncHelper) else None
        return self._solver(self._host_bs_funcs(iatom), rho, propars, points, weights, alphas, self._inner_threshold, **self._solver_options)

    def _solver_code(self):
        if callable(self._solver):
            # the device runs only the weights passes; a linear-coefficient
            # scheme sums every atom (a callable may return any sign)
            return 2
        if self._solver not in ("quadprog",):
            raise NotImplementedError(f"solver {self._solver!r}: the device path solves the QP (quadprog's optimum)")
        return 1

    def _scheme_arrays(self):
        fn_off, alpha, order, norm = self.bs_helper.device_tables(self.numbers)
        solver = self._solver_code()
        d = self._base_scheme_arrays()
        d.update(scheme=1 if solver == 1 else 2, solver=solver, fn_offsets=fn_off, fn_alpha=alpha, fn_order=order, fn_norm=norm, npar=int(fn_off[-1]))
        if solver == 1:
            if not np.all(order == 2.0):
                raise NotImplementedError("the GISA QP uses Gaussian overlaps and needs n = 2 functions")
            blocks = [gaussian_overlap_matrix(alpha[fn_off[a] : fn_off[a + 1]]).ravel() for a in range(self.natom)]
            d["qp_matrix"] = np.ascontiguousarray(np.concatenate(blocks))
            d["inner_maxiter"] = 1
        if callable(self._solver):
            d["inner_maxiter"] = 1
        return d
