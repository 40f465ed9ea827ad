"""MI355X-native stockholder partitioning (Hirshfeld / Hirshfeld-I / ISA / MBIS / GISA / aLISA / gLISA / NLIS / GMBIS).

Drop-in for the grid-based stockholder loop of horton-part: the scheme
classes keep the reference's ``WPart`` API and run the loop in
``libhortonpart.so`` (hand-written HIP kernels for gfx950).
"""

from .alisa import LinearISAWPart
from .basis import ExpBasisFuncHelper, evaluate_function
from .gisa import GaussianISAWPart
from .glisa import GlobalLinearISAWPart
from .gmbis import GMBISWPart, get_initial_gmbis_propars
from .grid import AtomGrid, BeckeWeights, ExpRTransform, MolGrid, OneDGrid, PowerRTransform, UniformInteger
from .hirshfeld import HirshfeldWPart
from .hirshfeld_i import HirshfeldIWPart
from .isa import ISAWPart
from .mbis import MBISWPart, get_initial_mbis_propars, get_nshell
from .nlis import NLISWPart, get_initial_nlis_propars, get_nlis_nshell
from .proatomdb import ProAtomDB, ProAtomRecord
from .serving import PartitionServer
from .utils import wpart_schemes

__all__ = [
    "HirshfeldWPart",
    "HirshfeldIWPart",
    "ISAWPart",
    "ProAtomDB",
    "ProAtomRecord",
    "MBISWPart",
    "GaussianISAWPart",
    "LinearISAWPart",
    "GlobalLinearISAWPart",
    "NLISWPart",
    "GMBISWPart",
    "get_nlis_nshell",
    "get_initial_nlis_propars",
    "get_initial_gmbis_propars",
    "PartitionServer",
    "wpart_schemes",
    "get_nshell",
    "get_initial_mbis_propars",
    "ExpBasisFuncHelper",
    "evaluate_function",
    "MolGrid",
    "AtomGrid",
    "BeckeWeights",
    "ExpRTransform",
    "PowerRTransform",
    "OneDGrid",
    "UniformInteger",
]
