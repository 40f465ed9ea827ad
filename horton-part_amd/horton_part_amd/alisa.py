"""Linear Iterative Stockholder Analysis (aLISA) on the GPU.

API of ``alisa.py`` in the reference: ``setup_bs_helper`` (1130-1152) and
``LinearISAWPart`` (1155-1335).  Device solvers: ``"sc"`` (solver_sc,
193-291), ``"newton"`` (exact Newton, 543-610) and ``"m-newton"``
(line-searched Newton, 613-680), all through ``_solver_general_newton``
(753-938) semantics.  A solver callable (1323-1335) runs in the caller on
each atom's problem, with the weights pass on the device.  The reference's
API default ``"cvxopt"`` (1219) is an interior-point method of a third-party
package that is not restated here: it stays the default, so that a caller
who omits ``solver=`` gets a ``NotImplementedError`` that says so instead of
another algorithm's numbers.
"""

import numpy as np

from .basis import AnalyticBasisFuncHelper, ExpBasisFuncHelper
from .gisa import GaussianISAWPart

__all__ = ["LinearISAWPart", "setup_bs_helper", "DEVICE_SOLVERS"]

DEVICE_SOLVERS = {"sc": 2, "newton": 3, "m-newton": 4}


def setup_bs_helper(part):
    """Basis helper from a name, a JSON/YAML path or a helper object (alisa.py:1130-1152)."""
    if part._bs_helper is None:
        if isinstance(part.basis_func, str):
            name = part.basis_func.lower()
            if name in ["gauss", "slater"]:
                if part.basis_type != "analytic":
                    raise NotImplementedError("numeric (spline) basis functions are not part of the device path")
                part._bs_helper = ExpBasisFuncHelper.from_function_type(name)
            else:
                part._bs_helper = ExpBasisFuncHelper.from_file(part.basis_func)
        elif isinstance(part.basis_func, AnalyticBasisFuncHelper):
            part._bs_helper = part.basis_func
        else:
            raise NotImplementedError("The type of basis_func should be one of string or class BasisFuncHelper.")
    return part._bs_helper


class LinearISAWPart(GaussianISAWPart):
    """aLISA with per-atom linear coefficients (alisa.py:1155-1335)."""

    name = "lisa"

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
        solver="cvxopt",
        solver_options=None,
        basis_func="gauss",
        basis_type="analytic",
        grid_type=1,
        **kwargs,
    ):
        self.basis_func = basis_func
        self.basis_type = basis_type
        super().__init__(
            coordinates,
            numbers,
            pseudo_numbers,
            grid,
            moldens,
            spindens,
            lmax,
            logger,
            threshold,
            maxiter,
            inner_threshold,
            radius_cutoff,
            solver,
            solver_options,
            grid_type,
            **kwargs,
        )

    @property
    def bs_helper(self):
        return setup_bs_helper(self)

    def _solve_atom_on_host(self, iatom, rho, points, weights, propars):
        """LinearISAWPart._opt_propars with a callable (alisa.py:1304-1335)."""
        return self._solver(
            self._host_bs_funcs(iatom),
            rho,
            propars,
            points,
            weights,
            self._inner_threshold,
            self.logger,
            density_cutoff=self.density_cutoff,
            negative_cutoff=self.negative_cutoff,
            population_cutoff=self.population_cutoff,
            **self._solver_options,
        )

    def _solver_code(self):
        if callable(self._solver):
            return 2
        if self._solver == "cvxopt":
            raise NotImplementedError(
                "solver 'cvxopt' (the reference's default, an interior-point method of the cvxopt package) is not "
                f"provided: pass solver= one of {sorted(DEVICE_SOLVERS)} for the device solvers, or a callable"
            )
        if self._solver not in DEVICE_SOLVERS:
            raise NotImplementedError(f"solver {self._solver!r} is not on the device; use one of {sorted(DEVICE_SOLVERS)}")
        return DEVICE_SOLVERS[self._solver]

    def _scheme_arrays(self):
        d = super()._scheme_arrays()
        if callable(self._solver):
            d["inner_maxiter"] = 1
            return d
        opts = dict(self._solver_options)
        solver = d["solver"]
        if solver == 2:
            d["inner_maxiter"] = int(float(opts.pop("max_niter_inner", 100000)))
        else:
            d["inner_maxiter"] = int(opts.pop("niter", 10000))
            d["newton_tau"] = float(opts.pop("tau", 1.0))
            d["newton_linspace"] = int(opts.pop("linspace_size", 20))
            d["check_mono"] = int(bool(opts.pop("check_mono", False)))
        if opts:
            raise TypeError(f"unknown solver options {sorted(opts)}")
        return d
