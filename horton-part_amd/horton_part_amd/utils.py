"""Constants, geometry checks and the scheme registry.

Host-side logic of ``utils.py`` in the reference: constants
(``utils.py:50-59`` reading ``data/constants.yaml:27-30``), ``typecheck_geo``
(``utils.py:107-186``) and ``wpart_schemes`` (``utils.py:62-104``).  The
parameter checks of ``utils.py:255-553`` run inside the device solvers, which
report their conditions as status bits; the scheme classes turn those into the
reference's warnings and exceptions.
"""

import json
import logging
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

    The stockholder schemes of the accelerated path are provided:
    ``"h"``, ``"hi"``, ``"is"``, ``"mbis"``, ``"gisa"``, ``"lisa"``,
    ``"glisa"``, ``"nlis"`` and ``"gmbis"``.
    """
    if scheme == "h":
        from .hirshfeld import HirshfeldWPart

        return HirshfeldWPart
    if scheme == "hi":
        from .hirshfeld_i import HirshfeldIWPart

        return HirshfeldIWPart
    if scheme == "is":
        from .isa import ISAWPart

        return ISAWPart
    if scheme == "mbis":
        from .mbis import MBISWPart

        return MBISWPart
    if scheme == "gisa":
        from .gisa import GaussianISAWPart

        return GaussianISAWPart
    if scheme == "lisa":
        from .alisa import LinearISAWPart

        return LinearISAWPart
    if scheme == "glisa":
        from .glisa import GlobalLinearISAWPart

        return GlobalLinearISAWPart
    if scheme == "nlis":
        from .nlis import NLISWPart

        return NLISWPart
    if scheme == "gmbis":
        from .gmbis import GMBISWPart

        return GMBISWPart
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


def default_logger(name):
    logger = logging.getLogger(name)
    if not logger.handlers:
        logger.addHandler(logging.NullHandler())
    return logger
