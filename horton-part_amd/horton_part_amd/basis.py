"""Exponential basis functions of the GISA and LISA pro-atoms.

The public names are those of the reference's ``core/basis.py``
(``evaluate_function`` 85-186, ``load_params`` 42-82, ``BasisFuncHelper`` and
``AnalyticBasisFuncHelper`` 189-238, ``ExpBasisFuncHelper`` 241-327), with the
same arguments, return shapes and exceptions, because they are the interface
a caller of the scheme classes sees.

A shell is ``c * g(n, alpha) * exp(-alpha r^n)`` with the normalisation
``g(n, alpha) = n alpha^(3/n) / (4 pi Gamma(3/n))`` (``4 pi int r^2 g e^{-alpha
r^n} dr = 1``).  One routine, ``shell_norm``, computes ``g`` for the host
evaluation and for the device tables, so the kernels and the host agree on
the prefactor bit for bit.  The radial derivative returned with
``nderiv=1`` is ``-n r^(n-1) f`` (no factor alpha), which is what the
reference returns.  The tables in ``data/`` are the reference's
``gauss.json`` and ``slater.json`` parameter sets.
"""

import json
import math

import numpy as np

from .utils import DATA_PATH

__all__ = ["evaluate_function", "load_params", "shell_norm", "BasisFuncHelper", "AnalyticBasisFuncHelper", "ExpBasisFuncHelper"]


def shell_norm(order, alpha):
    """``n alpha^(3/n) / (4 pi Gamma(3/n))`` for scalars or arrays of equal shape."""
    n = np.asarray(order, dtype=float)
    a = np.asarray(alpha, dtype=float)
    gam = np.vectorize(math.gamma, otypes=[float])(3.0 / n)
    g = n * a ** (3.0 / n) / (4.0 * np.pi * gam)
    return float(g) if g.ndim == 0 else g


def _check_args(n, alpha, r):
    if (np.asarray(n) <= 0).any():
        raise ValueError("n must be positive")
    if (np.asarray(alpha) < 0).any():
        raise ValueError("alpha must be non-negative")
    if not isinstance(r, np.ndarray):
        raise ValueError("r must be a numpy array")


def evaluate_function(n, population, alpha, r, nderiv=0, axis=None):
    """Shells ``population * g(n, alpha) exp(-alpha r^n)`` on the radii ``r``.

    Scalar ``n`` gives an array shaped like ``r``.  Arrays ``n``, ``alpha``
    and ``population`` (one entry per shell) give one row per shell, summed
    over ``axis`` when it is given.  ``nderiv=1`` also returns the radial
    derivative ``-n r^(n-1) f``; any other order raises NotImplementedError.
    """
    _check_args(n, alpha, r)
    if nderiv not in (0, 1):
        raise NotImplementedError
    per_shell = not np.isscalar(n)
    if per_shell:
        # shells along the first axis, radii along the second
        col = lambda v: np.asarray(v, dtype=float).reshape(-1, 1)  # noqa: E731
        n_, a_, c_, rr = col(n), col(alpha), col(population), np.asarray(r)[None, :]
    else:
        n_, a_, c_, rr = n, alpha, population, r
    rn = rr**n_
    f = c_ * shell_norm(n_, a_) * np.exp(-a_ * rn)
    out = [f]
    if nderiv == 1:
        out.append(-n_ * rr ** (n_ - 1) * f)
    if per_shell and axis is not None:
        out = [np.sum(v, axis=axis) for v in out]
    return out[0] if nderiv == 0 else tuple(out)


def load_params(filename, extension="json"):
    """``(orders, exponents, initials)`` dictionaries keyed by atomic number
    from a file mapping ``Z -> [orders, exponents, initials]`` (JSON or YAML)."""
    kind = extension.lower()
    assert kind in ("json", "yaml"), "Format must be 'json' or 'yaml'."
    with open(filename) as handle:
        if kind == "json":
            table = json.load(handle)
        else:
            import yaml

            table = yaml.safe_load(handle)
    columns = ({}, {}, {})
    for key, rows in table.items():
        for column, row in zip(columns, rows):
            column[int(key)] = np.asarray(row)
    return columns


class BasisFuncHelper:
    """Per-element basis description: initial values and shell densities."""

    def __init__(self, initials, *args, **kwargs):
        self._initials = initials

    @property
    def initials(self):
        return self._initials

    def get_nshell(self, number):
        raise NotImplementedError

    def get_initial(self, number, ishell=None):
        values = self.initials[number]
        return np.asarray(values) if ishell is None else values[ishell]

    def compute_proshell_dens(self, number, ishell, population, points, nderiv=0):
        raise NotImplementedError

    def compute_proatom_dens(self, number, populations, points, nderiv=0):
        """Sum of the element's shells with the given populations (and the
        sum of their derivatives for ``nderiv=1``)."""
        if nderiv not in (0, 1):
            raise RuntimeError("The argument `nderiv` should only be 0 or 1.")
        total = [0.0, 0.0]
        for ishell in range(self.get_nshell(number)):
            res = self.compute_proshell_dens(number, ishell, populations[ishell], points, nderiv)
            parts = (res,) if nderiv == 0 else res
            for j, v in enumerate(parts):
                total[j] = total[j] + v
        return total[0] if nderiv == 0 else (total[0], total[1])


class AnalyticBasisFuncHelper(BasisFuncHelper):
    """Helpers whose shells are closed-form functions of the radius."""


class ExpBasisFuncHelper(AnalyticBasisFuncHelper):
    """Shells ``exp(-alpha r^n)`` with per-element orders, exponents and
    initial populations."""

    def __init__(self, exponents_orders, exponents=None, initials=None):
        super().__init__(initials)
        self._orders = exponents_orders
        self._exponents = exponents

    @property
    def orders(self):
        return self._orders

    @property
    def exponents(self):
        return self._exponents

    def get_nshell(self, number):
        return len(self._exponents[number])

    def get_order(self, number, ishell=None):
        values = self._orders[number]
        return values if ishell is None else values[ishell]

    def get_exponent(self, number, ishell=None):
        values = self._exponents[number]
        return np.asarray(values) if ishell is None else values[ishell]

    def compute_proshell_dens(self, number, ishell, population, points, nderiv=0):
        n = self.get_order(number, ishell)
        alpha = self.get_exponent(number, ishell)
        return evaluate_function(n, population, alpha, points, nderiv)

    def device_tables(self, numbers):
        """The per-atom tables the device kernels take: function offsets
        (``natom + 1``) and per function its exponent, order and ``g``."""
        missing = sorted({int(z) for z in numbers} - {int(z) for z in self._exponents})
        if missing:
            raise KeyError(f"no basis functions for element {missing[0]}")
        alpha = np.concatenate([np.asarray(self._exponents[int(z)], float).ravel() for z in numbers])
        order = np.concatenate([np.asarray(self._orders[int(z)], float).ravel() for z in numbers])
        counts = [self.get_nshell(int(z)) for z in numbers]
        _check_args(order, alpha, np.zeros(0))
        offsets = np.zeros(len(counts) + 1, np.int32)
        np.cumsum(counts, out=offsets[1:])
        norm = np.atleast_1d(shell_norm(order, alpha)) if order.size else np.zeros(0)
        return offsets, alpha, order, norm

    @classmethod
    def from_function_type(cls, func_type="gauss"):
        """The shipped ``gauss`` or ``slater`` parameter set."""
        assert func_type in ("gauss", "slater")
        return cls.from_file(DATA_PATH / f"{func_type}.json")

    @classmethod
    def from_file(cls, filename):
        """Parameters from a JSON or YAML table; elements without initial
        values start from equal populations."""
        kind = "yaml" if str(filename).endswith(".yaml") else "json"
        orders, exponents, initials = load_params(filename, extension=kind)
        for number in exponents.keys() - initials.keys():
            k = len(exponents[number])
            initials[number] = np.full(k, 1.0 / k)
        return cls(orders, exponents, initials)

    from_yaml = from_file
    from_json = from_file

    @classmethod
    def uniform(cls, numbers, exponents, order=1.0, initials=None):
        """The same exponents (and order) for every element in ``numbers``."""
        exps = np.asarray(exponents, float)
        start = np.full(exps.size, 1.0 / exps.size) if initials is None else np.asarray(initials, float)
        elements = sorted({int(z) for z in numbers})
        return cls(
            {z: np.full(exps.size, float(order)) for z in elements},
            {z: exps.copy() for z in elements},
            {z: start.copy() for z in elements},
        )
