"""Global Linear Iterative Stockholder Analysis (gLISA) on the GPU.

API of ``glisa.py`` in the reference: ``GlobalLinearISAWPart`` (55-1028).
gLISA optimises every pro-atom coefficient at once on the molecular grid:
the promolecule ``rho0 = sum_i c_i g_i``, the objective
``sum_p w rho ln(rho / rho0)``, its gradient and its Hessian (``_working_matrix``,
411-479) are evaluated on the device (``hp_part_global_terms``; the Hessian
``G^T diag(w rho / rho0^2) G`` on the matrix cores).  The solvers' own steps
run here on the host, as the reference's do: the self-consistent map
``function_g`` (850-879) from the device gradient, the Newton step
``solve(H, -1 - df)`` and the line search of ``_solver_general_newton``
(617-803), whose pro-atom checks and change use the device pass
``hp_part_global_proatoms`` on the molecular grid (radial grids, for
grid_type 1, are evaluated here: a few hundred values per atom).

Solvers: ``"sc"``, ``"newton"``, ``"m-newton"``, or a callable (called as
the reference calls it, with the solver options only).  The reference's
default ``"cvxopt"`` and its quasi-Newton, DIIS and trust-region solvers
raise ``NotImplementedError``.
"""

import numpy as np
from scipy.linalg import solve

from .alisa import setup_bs_helper
from .basis import ExpBasisFuncHelper
from .cache import just_once
from .gisa import init_propars
from .part import AbstractISAWPart

__all__ = ["GlobalLinearISAWPart", "fix_propars"]

DEVICE_SOLVERS = ("sc", "newton", "m-newton")


def fix_propars(exp_array, propars, delta):
    """Indices kept fixed in a Newton step (utils.py:556-577): from the most
    diffuse function on, those whose coefficient is below 1e-4 in magnitude
    and whose step is negative, up to the first that is not."""
    fixed = []
    for index in np.argsort(exp_array):
        if not (np.abs(propars[index]) < 1e-4 and delta[index] < 0.0):
            break
        fixed.append(index)
    return fixed


class GlobalLinearISAWPart(AbstractISAWPart):
    """gLISA (glisa.py:55-1028) with the global terms on the device."""

    name = "glisa"

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
        solver="cvxopt",
        solver_options=None,
        basis_func="gauss",
        grid_type=1,
        basis_type="analytic",
        **kwargs,
    ):
        self.basis_func = basis_func
        self.basis_type = basis_type
        self._solver = solver
        self._solver_options = dict(solver_options or {})
        self._bs_helper = None
        self._ranges = []
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
            grid_type=grid_type,
            **kwargs,
        )

    @property
    def bs_helper(self):
        return setup_bs_helper(self)

    @property
    def maxiter(self):
        return self._maxiter

    @property
    def threshold(self):
        return self._threshold

    @property
    def mol_pop(self):
        return self.grid.integrate(self._moldens)

    def get_rgrid(self, index):
        if self.only_use_molgrid:
            raise NotImplementedError
        return self.get_grid(index).rgrid

    def get_proatom_rho(self, iatom, propars=None, **kwargs):
        """Pro-atom density and derivative on the radial grid or at the
        molecular points (gisa.py:42-65)."""
        if propars is None:
            propars = self.cache.load("propars")
        pops = propars[self._ranges[iatom] : self._ranges[iatom + 1]]
        r = self.radial_distances[iatom] if self.on_molgrid else self.get_rgrid(iatom).points
        return self.bs_helper.compute_proatom_dens(self.numbers[iatom], pops, r, 1)

    def _init_propars(self):
        return init_propars(self)

    def _scheme_arrays(self):
        fn_off, alpha, order, norm = self.bs_helper.device_tables(self.numbers)
        d = self._base_scheme_arrays()
        # the global problem lives on the molecular grid whatever the grid
        # type (glisa.py:257-270 force_on_molgrid); grid_type 3 only changes
        # where the moments integrate
        d.update(
            scheme=2,
            solver=2,
            grid_type=3 if self.grid_type == 3 else 2,
            fn_offsets=fn_off,
            fn_alpha=alpha,
            fn_order=order,
            fn_norm=norm,
            npar=int(fn_off[-1]),
            inner_maxiter=1,
        )
        return d

    # -- per-step quantities ---------------------------------------------------
    def _check_pars(self, new, old=None, check_mono=False):
        """(pro-atoms valid, change against ``old``) at the coefficients ``new``
        (glisa.py:303-330, iterstock.py:32-45)."""
        if self.on_molgrid:
            mn, msd = self._dev.global_proatoms(new, new if old is None else old)
            return bool((mn >= self.negative_cutoff).all()), float(np.sqrt(msd.sum()))
        valid, msd = True, 0.0
        for iatom in range(self.natom):
            rgrid = self.get_rgrid(iatom)
            rho_new = self.get_proatom_rho(iatom, new)[0]
            if (rho_new < self.negative_cutoff).any():
                valid = False
            if check_mono and (rho_new[:-1] - rho_new[1:] < self.negative_cutoff).any():
                valid = False
            if old is not None:
                delta = rho_new - self.get_proatom_rho(iatom, old)[0]
                msd += rgrid.integrate(4 * np.pi * rgrid.points**2, delta, delta)
        return valid, float(np.sqrt(msd))

    def compute_change(self, propars1, propars2):
        return self._check_pars(propars1, propars2)[1]

    # -- solvers ---------------------------------------------------------------
    def solver_sc(self, niter_print=1):
        """The self-consistent map x_i <- x_i sum_p w rho g_i / rho0(x)
        (glisa.py:805-879), from the device gradient."""
        propars = self.cache.load("propars")
        it = 0
        while True:
            old = propars.copy()
            f, grad, _, _ = self._dev.global_terms(old, nderiv=1)
            propars[:] = -old * grad
            change = self.compute_change(propars, old)
            self.history_entropies.append(f)
            self.history_changes.append(change)
            self.history_propars.append(propars.copy())
            if (it + 1) % niter_print == 0:
                self.logger.info("%9i   %10.5e   %10.5e" % (it + 1, change, f))
            if change < self._threshold:
                break
            it += 1
        self.cache.dump("niter", it + 1, tags="o")
        return propars

    def solver_newton(self, maxiter=100):
        return self._solver_general_newton(maxiter=maxiter, mode="exact")

    def solver_m_newton(self, maxiter=100, linspace_size=40, tau=1.0, linesearch_mode="valid-promol", **kwargs):
        return self._solver_general_newton(
            mode="modified", maxiter=maxiter, linesearch_mode=linesearch_mode, tau=tau, linspace_size=linspace_size, **kwargs
        )

    def _solver_general_newton(self, mode="exact", maxiter=1000, linesearch_mode="valid-promol", tau=1.0, linspace_size=40,
                               check_mono=False, niter_exact_newton=0):
        """Newton steps on the global problem (glisa.py:617-803) with the
        "valid-promol" line search."""
        if mode not in ("exact", "modified"):
            raise NotImplementedError(f"Newton mode {mode!r} is not provided (bfgs needs the quasi-Newton update)")
        if linesearch_mode != "valid-promol":
            raise NotImplementedError("only the 'valid-promol' line search is provided")
        assert tau >= 0
        propars = self.cache.load("propars")
        pop = self.grid.integrate(self._moldens)
        exps = self.bs_helper.get_exponent if isinstance(self.bs_helper, ExpBasisFuncHelper) else None
        for irep in range(maxiter):
            old = propars.copy()
            f, df, hess, min_pro = self._dev.global_terms(propars, nderiv=2)
            try:
                delta = solve(hess, -1 - df, assume_a="sym")
            except np.linalg.LinAlgError as e:
                raise RuntimeError(e)
            if mode == "exact":
                new = propars + delta
            else:
                fixed = np.ones_like(delta)
                if exps is not None:
                    for iatom in range(self.natom):
                        sl = slice(self._ranges[iatom], self._ranges[iatom + 1])
                        for j in fix_propars(exps(self.numbers[iatom]), propars[sl], delta[sl]):
                            fixed[self._ranges[iatom] + j] = 0
                for x in np.linspace(tau, 0, linspace_size, endpoint=False):
                    new = propars + fixed * (x * delta)
                    if self._check_pars(new, check_mono=check_mono and not self.on_molgrid)[0]:
                        break
                else:
                    raise RuntimeError("Line search failed!")
            propars[:] = new
            change = self.compute_change(propars, old)
            self.history_entropies.append(f)
            self.history_propars.append(propars.copy())
            self.history_changes.append(change)
            self.logger.info(f"            {irep + 1:<4}    {change:.5e}    {f:.5e}")
            if change < self._threshold:
                if min_pro < self.negative_cutoff:
                    raise RuntimeError("Negative pro-atom density found!")
                if abs(np.sum(propars) - pop) > self.population_cutoff:
                    self.logger.warning("The sum of pro-atom parameters is not equal to the atomic population.")
                self.cache.dump("niter", irep + 1, tags="o")
                return propars
        raise RuntimeError("Not converged!")

    def _opt_propars(self, **kwargs):
        if callable(self._solver):
            return self._solver(**kwargs)
        if self._solver == "cvxopt":
            raise NotImplementedError(
                "solver 'cvxopt' (the reference's default, an interior-point method of the cvxopt package) is not "
                f"provided: pass solver= one of {list(DEVICE_SOLVERS)}, or a callable"
            )
        if self._solver not in DEVICE_SOLVERS:
            raise NotImplementedError(f"solver {self._solver!r} is not provided; use one of {list(DEVICE_SOLVERS)}")
        return getattr(self, "solver_" + self._solver.replace("-", "_"))(**kwargs)

    @just_once
    def do_partitioning(self):
        """glisa.py:248-281: the global solve, then the weights of the final
        pro-atoms on the molecular grid and the charges from them."""
        dev = self._ensure_device()
        propars = self._init_propars()
        self.cache.dump("propars", propars, tags="o")
        new = self._opt_propars(**self._solver_options)
        propars[:] = new
        dev.set_propars(propars)
        dev.weights(spherical=False)
        charges = self.pseudo_numbers - dev.populations()
        self.cache.dump("charges", charges, tags="o")
        for a in range(self.natom):
            self.cache.add_lazy(f"at_weights_{a}", (lambda a=a: dev.at_weights(a, int(self.grid.size))))
        self.cache.add_lazy("promoldens", dev.promoldens)
        self._inner_niter = None
        self._finalize_propars()
