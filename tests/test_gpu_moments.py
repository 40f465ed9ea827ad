"""AIM moments of any order (core/base.py:329-402) on the device against the
oracle: lmax 0-2 (prefixes of the 40 sums of the lmax = 3 kernel) and 4-6
(the any-order kernel, one lane per quantity), on atom grids (grid types 1)
and on the molecular grid (grid_type 3), for Slater, Gaussian and tabulated
pro-atoms.  Bar: rtol 1e-9 with an absolute floor of 1e-12 times the largest
magnitude of each family (high orders reach 1e3-1e5 on these grids)."""

import sys
import warnings

import numpy as np
import pytest

from conftest import require_gpu

pytestmark = pytest.mark.gpu
warnings.simplefilter("ignore")


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _check(got, expect, what):
    for g, e, name in zip(got, expect, ("cartesian", "pure", "radial")):
        assert g.shape == e.shape, (what, name, g.shape, e.shape)
        scale = max(1.0, float(np.abs(e).max()))
        err = float(np.abs(g - e).max())
        _say(f"  {what} {name}: shape {g.shape} max|d| {err:.2e} (scale {scale:.2e})")
        np.testing.assert_allclose(g, e, rtol=1e-9, atol=1e-12 * scale)


def _oracle_moments(oracle, lmax):
    oracle.lmax = lmax
    return oracle.do_moments()


@pytest.mark.parametrize("grid_type", [1, 3])
def test_moments_any_order_mbis(grid_type):
    require_gpu()
    from horton_part_amd import MBISWPart, workloads
    from oracle.horton_oracle import OracleMBIS

    _say(f"case: MBIS moments lmax 0-6, grid_type {grid_type}, chain 4 atoms seed 11, 40x50, maxiter 12")
    w = workloads.chain(natom=4, seed=11, nrad=40, nang=50)
    kw = dict(maxiter=12, grid_type=grid_type)
    part = MBISWPart(w.coordinates, w.numbers, w.pseudo_numbers, w.grid, w.density, **kw)
    part.do_partitioning()
    oracle = OracleMBIS(w.coordinates, w.numbers, w.pseudo_numbers, w.grid, w.density, **kw)
    oracle.do_partitioning()
    assert part["niter"] == oracle.cache["niter"]
    for lmax in (0, 1, 2, 4, 6):
        _check(part._dev.moments(lmax), _oracle_moments(oracle, lmax), f"lmax {lmax}")
    part.close()


def test_moments_public_api_lmax5():
    """``MBISWPart(..., lmax=5).do_moments()``: the reference's output keys
    with lmax = 5 shapes (56 cartesian, 36 pure, 6 radial)."""
    require_gpu()
    from horton_part_amd import MBISWPart, workloads
    from oracle.horton_oracle import OracleMBIS

    _say("case: MBIS lmax 5 through do_moments, chain 3 atoms seed 5, 36x26, maxiter 10")
    w = workloads.chain(natom=3, seed=5, nrad=36, nang=26)
    part = MBISWPart(w.coordinates, w.numbers, w.pseudo_numbers, w.grid, w.density, maxiter=10, lmax=5)
    part.do_moments()
    oracle = OracleMBIS(w.coordinates, w.numbers, w.pseudo_numbers, w.grid, w.density, maxiter=10, lmax=5)
    oracle.do_partitioning()
    got = (part["cartesian_multipoles"], part["pure_multipoles"], part["radial_moments"])
    assert got[0].shape == (3, 56) and got[1].shape == (3, 36) and got[2].shape == (3, 6)
    _check(got, oracle.do_moments(), "lmax 5")
    part.close()


def test_moments_any_order_gisa_and_hirshfeld():
    """The any-order kernel with Gaussian pro-atoms (GISA) and tabulated ones
    (Hirshfeld on the reference water fixture's pro-atoms)."""
    require_gpu()
    from horton_part_amd import GaussianISAWPart, HirshfeldWPart, workloads
    from oracle.horton_oracle import OracleBasis, OracleGISA, OracleHirshfeld

    from common import load_proatomdb, load_water_fixture, proatom_files
    from oracle import horton_oracle as ho

    _say("case: GISA moments lmax 4, chain 3 atoms seed 6, 36x26, maxiter 10")
    w = workloads.chain(natom=3, seed=6, nrad=36, nang=26)
    part = GaussianISAWPart(w.coordinates, w.numbers, w.pseudo_numbers, w.grid, w.density, maxiter=10)
    part.do_partitioning()
    bs = part.bs_helper
    basis = OracleBasis(bs.orders, bs.exponents, bs.initials)
    oracle = OracleGISA(w.coordinates, w.numbers, w.pseudo_numbers, w.grid, w.density, maxiter=10, basis=basis)
    oracle.do_partitioning()
    _check(part._dev.moments(4), _oracle_moments(oracle, 4), "gisa lmax 4")
    part.close()

    _say("case: Hirshfeld moments lmax 5, reference water fixture")
    coords, nums, pseudo, grid, dens = load_water_fixture()
    db = load_proatomdb([8, 6, 1], 1, -1, "hf_sto3g")
    odb = ho.OracleProAtomDB([ho.load_record_npz(f) for f in proatom_files([8, 6, 1], 1, -1, "hf_sto3g")])
    part = HirshfeldWPart(coords, nums, pseudo, grid, dens, db)
    part.do_partitioning()
    oracle = OracleHirshfeld(coords, nums, pseudo, grid, dens, proatomdb=odb)
    oracle.do_partitioning()
    _check(part._dev.moments(5), _oracle_moments(oracle, 5), "hirshfeld lmax 5")
    part.close()
