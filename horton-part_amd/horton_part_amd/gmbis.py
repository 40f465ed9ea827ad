"""Generalised Minimal Basis Iterative Stockholder (GMBIS) on the GPU.

API of ``gmbis.py`` in the reference: ``get_initial_gmbis_propars`` (36-72)
and ``GMBISWPart`` (75-156): NLIS with MBIS's shell counts and initial
populations (``nlis.py``), exponents from 2Z down to 2 geometrically.
"""

import numpy as np

from .mbis import get_nshell
from .nlis import NLISWPart, _power_of

__all__ = ["GMBISWPart", "get_initial_gmbis_propars"]

_SHELL_POPULATIONS = (2.0, 8.0, 8.0, 18.0, 18.0, 32.0, 32.0)


def get_initial_gmbis_propars(number, exp_n_dict, logger=None):
    """(N, S, n) per MBIS shell; the last population makes the sum Z (gmbis.py:36-72)."""
    nshell = int(get_nshell(number))
    S0 = 2.0 * number
    ratio = (2.0 / S0) ** (1.0 / (nshell - 1)) if nshell > 1 else 1.0
    pops = list(_SHELL_POPULATIONS[:nshell])
    pops[-1] = number - sum(pops[:-1])
    rows = [(pops[k], S0 * ratio**k, _power_of(exp_n_dict, number, k)) for k in range(nshell)]
    return np.array(rows, dtype=float).ravel()


class GMBISWPart(NLISWPart):
    """Generalised MBIS (gmbis.py:75-156)."""

    name = "gmbis"

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
        exp_n_dict=1.0,
        grid_type=1,
        **kwargs,
    ):
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
            exp_n_dict=exp_n_dict,
            nshell_dict={},
            grid_type=grid_type,
            **kwargs,
        )

    def _initial(self, number):
        return get_initial_gmbis_propars(number, self._exp_n_dict)
