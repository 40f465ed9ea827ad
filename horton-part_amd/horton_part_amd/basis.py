"""Exponential basis functions of the GISA and LISA pro-atoms.

Restates ``core/basis.py``: ``evaluate_function`` (85-186) computes
``population * n alpha^(3/n) / (4 pi Gamma(3/n)) * exp(-alpha r^n)`` and its
radial derivative; ``ExpBasisFuncHelper`` (241-327) holds per-element orders,
exponents and initial values loaded from JSON/YAML tables (``load_params``
42-82).  The tables shipped in ``data/`` are the reference's ``gauss.json``
and ``slater.json`` parameter sets (physical constants of the method).
The device kernels evaluate the same functions; this module supplies their
parameters (``norm`` is the prefactor at unit population).
"""

import json
import math

import numpy as np

from .utils import DATA_PATH

__all__ = ["evaluate_function", "load_params", "BasisFuncHelper", "AnalyticBasisFuncHelper", "ExpBasisFuncHelper"]


def load_params(filename, extension="json"):
    """Read ``{Z: [orders, exponents, initials]}`` from a JSON or YAML file."""
    assert extension.lower() in ["json", "yaml"], "Format must be 'json' or 'yaml'."
    with open(filename) as f:
        if extension.lower() == "json":
            data = json.load(f)
        else:
            import yaml

            data = yaml.safe_load(f)
    orders, exps, inits = {}, {}, {}
    for number, values in data.items():
        number = int(number)
        orders[number] = np.asarray(values[0])
        exps[number] = np.asarray(values[1])
        inits[number] = np.asarray(values[2])
    return orders, exps, inits


def evaluate_function(n, population, alpha, r, nderiv=0, axis=None):
    """Normalised ``exp(-alpha r^n)`` shells (vectorised over n/alpha arrays)."""
    if np.any(np.asarray(n) <= 0):
        raise ValueError("n must be positive")
    if np.any(np.asarray(alpha) < 0):
        raise ValueError("alpha must be non-negative")
    if not isinstance(r, np.ndarray):
        raise ValueError("r must be a numpy array")
    from scipy.special import gamma

    prefactor = population * n * alpha ** (3 / n) / (4 * np.pi * gamma(3 / n))
    vec = not np.isscalar(n)
    if vec:
        assert not (np.isscalar(alpha) or np.isscalar(population))
        prefactor = prefactor[:, np.newaxis]
        alpha = alpha[:, np.newaxis]
        r = r[np.newaxis, :]
        n = n[:, np.newaxis]
    f = prefactor * np.exp(-alpha * r**n)
    if nderiv > 0:
        # the reference's derivative has no factor alpha (core/basis.py:175)
        df = -n * r ** (n - 1) * f
    if vec and axis is not None:
        f = np.sum(f, axis=axis)
        if nderiv > 0:
            df = np.sum(df, axis=axis)
    if nderiv == 0:
        return f
    if nderiv == 1:
        return f, df
    raise NotImplementedError


def shell_norm(order, alpha):
    """Prefactor of ``evaluate_function`` at population 1 (same expression)."""
    return 1.0 * order * alpha ** (3 / order) / (4 * np.pi * math.gamma(3 / order))


class BasisFuncHelper:
    """Base of basis function helpers."""

    def __init__(self, initials, *args, **kwargs):
        self._initials = initials

    @property
    def initials(self):
        return self._initials

    def get_nshell(self, number):
        raise NotImplementedError

    def get_initial(self, number, ishell=None):
        return np.asarray(self.initials[number]) if ishell is None else self.initials[number][ishell]

    def compute_proshell_dens(self, number, ishell, population, points, nderiv=0):
        raise NotImplementedError

    def compute_proatom_dens(self, number, populations, points, nderiv=0):
        y = d = 0.0
        for i in range(self.get_nshell(number)):
            res = self.compute_proshell_dens(number, i, populations[i], points, nderiv)
            if nderiv == 0:
                y += res
            else:
                y += res[0]
                d += res[1]
        return y if nderiv == 0 else (y, d)


class AnalyticBasisFuncHelper(BasisFuncHelper):
    """Analytic basis function helper."""


class ExpBasisFuncHelper(AnalyticBasisFuncHelper):
    """Per-element ``exp(-alpha r^n)`` basis (orders, exponents, initials)."""

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
        return len(self.exponents[number])

    def get_order(self, number, ishell=None):
        return self.orders[number] if ishell is None else self.orders[number][ishell]

    def get_exponent(self, number, ishell=None):
        return np.asarray(self.exponents[number]) if ishell is None else self.exponents[number][ishell]

    def compute_proshell_dens(self, number, ishell, population, points, nderiv=0):
        return evaluate_function(
            self.get_order(number, ishell), population, self.get_exponent(number, ishell), points, nderiv
        )

    def device_tables(self, numbers):
        """Per-atom function offsets and per-function (alpha, order, norm)."""
        offsets = [0]
        alpha, order, norm = [], [], []
        for z in numbers:
            z = int(z)
            if z not in self.exponents:
                raise KeyError(f"no basis functions for element {z}")
            for k in range(self.get_nshell(z)):
                a = float(self.get_exponent(z, k))
                n = float(self.get_order(z, k))
                if n <= 0:
                    raise ValueError("n must be positive")
                if a < 0:
                    raise ValueError("alpha must be non-negative")
                alpha.append(a)
                order.append(n)
                norm.append(shell_norm(n, a))
            offsets.append(offsets[-1] + self.get_nshell(z))
        return (np.array(offsets, np.int32), np.array(alpha), np.array(order), np.array(norm))

    @classmethod
    def from_function_type(cls, func_type="gauss"):
        assert func_type in ["gauss", "slater"]
        return cls.from_file(DATA_PATH / f"{func_type}.json")

    @classmethod
    def from_file(cls, filename):
        ext = "yaml" if str(filename).endswith(".yaml") else "json"
        orders, exponents, initials = load_params(filename, extension=ext)
        for number, exps in exponents.items():
            if number not in initials:
                initials[number] = np.ones_like(exps) / len(exps)
        return cls(orders, exponents, initials)

    from_yaml = from_file
    from_json = from_file

    @classmethod
    def uniform(cls, numbers, exponents, order=1.0, initials=None):
        """Same ``exponents`` (and order) for every element in ``numbers``."""
        exponents = np.asarray(exponents, float)
        ords = {int(z): np.full(exponents.size, float(order)) for z in numbers}
        exps = {int(z): exponents.copy() for z in numbers}
        if initials is None:
            inits = {int(z): np.ones(exponents.size) / exponents.size for z in numbers}
        else:
            inits = {int(z): np.asarray(initials, float) for z in numbers}
        return cls(ords, exps, inits)
