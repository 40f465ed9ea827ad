"""Constants, geometry checks, the scheme registry and parameter checks.

Host-side logic of ``utils.py`` in the reference: constants
(``utils.py:50-59`` reading ``data/constants.yaml:27-30``), ``typecheck_geo``
(``utils.py:107-186``), ``wpart_schemes`` (``utils.py:62-104``),
``compute_quantities`` (``utils.py:198-252``) and the validation helpers
(``utils.py:255-553``).  The device path reports the same conditions as
status flags; the scheme classes turn them into these warnings and errors.
"""

import json
import logging
import warnings
from pathlib import Path

import numpy as np

__all__ = [
    "DATA_PATH",
    "DENSITY_CUTOFF",
    "NEGATIVE_CUTOFF",
    "POPULATION_CUTOFF",
    "ANGSTROM",
    "PERIODIC_TABLE",
    "radius_becke",
    "typecheck_geo",
    "wpart_schemes",
    "compute_quantities",
    "check_pro_atom_parameters",
    "check_pars_population",
    "check_dens_negativity",
    "check_pars_negativity",
    "check_dens_monotonicity",
]

DATA_PATH = Path(__file__).resolve().parent / "data"

with open(DATA_PATH / "constants.json") as _f:
    _docu = json.load(_f)

DENSITY_CUTOFF = float(_docu["DENSITY_CUTOFF"])
NEGATIVE_CUTOFF = float(_docu["NEGATIVE_CUTOFF"])
POPULATION_CUTOFF = float(_docu["POPULATION_CUTOFF"])
ANGSTROM = float(_docu["ANGSTROM"])
PERIODIC_TABLE = list(_docu["PERIODIC_TABLE"])
radius_becke = {k: (None if v is None else v * ANGSTROM) for k, v in enumerate(_docu["radius_becke"])}


def wpart_schemes(scheme):
    """Return the partitioning class registered under ``scheme``.

    Only the stockholder schemes of the accelerated path are provided:
    ``"mbis"``, ``"gisa"`` and ``"lisa"``.
    """
    if scheme == "mbis":
        from .mbis import MBISWPart

        return MBISWPart
    if scheme == "gisa":
        from .gisa import GaussianISAWPart

        return GaussianISAWPart
    if scheme == "lisa":
        from .alisa import LinearISAWPart

        return LinearISAWPart
    raise NotImplementedError(f"scheme {scheme!r} is not part of this package")


def typecheck_geo(coordinates=None, numbers=None, pseudo_numbers=None, need_coordinates=True, need_numbers=True, need_pseudo_numbers=True):
    """Validate a molecular geometry; returns ``[natom, coords, numbers, pseudo]``.

    ``numbers`` must be ``np.int64`` and coordinates a float (natom, 3) array,
    otherwise ``TypeError`` (reference ``utils.py:148-174``).
    """
    if coordinates is not None:
        natom = len(coordinates)
    elif numbers is not None:
        natom = len(numbers)
    elif pseudo_numbers is not None:
        natom = len(pseudo_numbers)
    else:
        raise TypeError("At least one argument is required and should not be None")
    if coordinates is None:
        if need_coordinates:
            raise TypeError("Coordinates can not be None.")
    elif coordinates.shape != (natom, 3) or not issubclass(coordinates.dtype.type, float):
        raise TypeError("The argument centers must be a float array with shape (natom,3).")
    if numbers is None:
        if need_numbers:
            raise TypeError("Numbers can not be None.")
    elif numbers.shape != (natom,) or not issubclass(numbers.dtype.type, np.int64):
        raise TypeError("The argument numbers must be a vector with length natom.")
    if pseudo_numbers is None:
        if need_pseudo_numbers:
            pseudo_numbers = numbers.astype(float)
    else:
        if pseudo_numbers.shape != (natom,):
            raise TypeError("The argument pseudo_numbers must be a vector with length natom.")
        if not issubclass(pseudo_numbers.dtype.type, float):
            pseudo_numbers = pseudo_numbers.astype(float)
    result = [natom]
    if need_coordinates:
        result.append(coordinates)
    if need_numbers:
        result.append(numbers)
    if need_pseudo_numbers:
        result.append(pseudo_numbers)
    return result


def compute_quantities(density, pro_atom_params, basis_functions, density_cutoff, do_sick=True, do_ratio=True, do_ln_ratio=True):
    """Pro-shells, pro-atom, sick mask, ratio and log-ratio (host helper)."""
    pro_atom_params = np.asarray(pro_atom_params).flatten()
    pro_shells = basis_functions * pro_atom_params[:, None]
    pro_density = pro_shells.sum(axis=0)
    sick = ratio = ln_ratio = None
    if do_sick:
        sick = (density < density_cutoff) | (pro_density < density_cutoff)
    with np.errstate(all="ignore"):
        if do_ratio:
            ratio = np.divide(density, pro_density, out=np.zeros_like(density), where=~sick)
        if do_ln_ratio:
            ln_ratio = np.log(ratio, out=np.zeros_like(density), where=~sick)
    return pro_shells, pro_density, sick, ratio, ln_ratio


def _emit(info, as_warn, logger):
    if as_warn:
        if logger is not None:
            logger.warning(info)
        else:
            warnings.warn(info)
    else:
        raise RuntimeError(info)


def check_pars_population(propars, ref_pop, as_warn=True, logger=None, population_cutoff=POPULATION_CUTOFF):
    """Warn (or raise) when ``|sum(propars) - ref_pop| > population_cutoff``."""
    diff = float(np.sum(propars) - ref_pop)
    if abs(diff) > population_cutoff:
        level = "WARNING" if as_warn else "ERROR"
        _emit(
            f"{level}: The sum of pro-atom parameters is not equal to reference population.\n"
            f"{level}: The difference is {diff:.5E}",
            as_warn,
            logger,
        )


def check_pars_negativity(pars, as_warn=True, logger=None, negative_cutoff=NEGATIVE_CUTOFF):
    if (np.asarray(pars) < negative_cutoff).any():
        if as_warn:
            _emit("WARNING: Not all pro-atom parameters are positive!", True, logger)
        else:
            raise RuntimeError("Negative pro-atom parameters found!")


def check_dens_negativity(dens, as_warn=False, logger=None, negative_cutoff=NEGATIVE_CUTOFF):
    if (np.asarray(dens) < negative_cutoff).any():
        if as_warn:
            _emit("WARNING: Not all pro-atom density are positive!", True, logger)
        else:
            raise RuntimeError("Negative pro-atom density found!")


def check_dens_monotonicity(dens, as_warn=True, logger=None, negative_cutoff=NEGATIVE_CUTOFF):
    dens = np.asarray(dens)
    if (dens[:-1] - dens[1:] < negative_cutoff).any():
        if as_warn:
            _emit("WARNING: Pro-atom density should be monotonically decreasing.", True, logger)
        else:
            raise RuntimeError("Pro-atom density should be monotonically decreasing.")


def check_pro_atom_parameters(
    pro_atom_params,
    basis_functions=None,
    total_population=None,
    pro_atom_density=None,
    check_monotonicity=True,
    check_negativity=True,
    check_propars_negativity=True,
    logger=None,
    negative_cutoff=NEGATIVE_CUTOFF,
    population_cutoff=POPULATION_CUTOFF,
):
    """Validate pro-atom parameters as the reference does after each solve."""
    pro_atom_params = np.asarray(pro_atom_params)
    if pro_atom_params.ndim != 1:
        raise ValueError("pro_atom_params must be a 1D array")
    if check_propars_negativity:
        check_pars_negativity(pro_atom_params, negative_cutoff=negative_cutoff, logger=logger)
    if basis_functions is not None and pro_atom_density is None:
        pro_atom_density = (np.asarray(basis_functions) * pro_atom_params[:, None]).sum(axis=0)
    if check_negativity and pro_atom_density is not None:
        check_dens_negativity(pro_atom_density, logger=logger, negative_cutoff=negative_cutoff)
    if total_population is not None:
        check_pars_population(pro_atom_params, total_population, logger=logger, population_cutoff=population_cutoff)
    if check_monotonicity and pro_atom_density is not None:
        check_dens_monotonicity(pro_atom_density, as_warn=False, logger=logger, negative_cutoff=negative_cutoff)


def default_logger(name):
    logger = logging.getLogger(name)
    if not logger.handlers:
        logger.addHandler(logging.NullHandler())
    return logger
