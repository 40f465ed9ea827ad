"""Density/partition files and the wavefunction density (host side).

* ``load_fchk`` + ``basis_density`` reproduce the density the reference's
  own water fixture holds (HORTON evaluated it from the same checkpoint file,
  tests/cached/water_sto3g_hf_g03_fchk_exp:5e-4:2e1:120:110.npz) at its
  points;
* the part-gen grid integrates that density to the 10 electrons;
* a density file written by ``save_density`` reads back into the same grid
  and arrays, with the keys of scripts/generate_density.py:232-252.
"""

import os

import numpy as np
import pytest

from horton_part_amd import io

from common import GOLDEN

FCHK = os.path.join(GOLDEN, "water_sto3g_hf_g03.fchk")


def test_fchk_fields():
    f = io.load_fchk(FCHK)
    assert f["Number of electrons"] == 10
    np.testing.assert_array_equal(f["Atomic numbers"], [8, 1, 1])
    assert f["Total SCF Density"].shape == (28,)
    assert f["Shell types"].tolist() == [0, -1, 0, 0]


def test_basis_density_matches_reference_fixture():
    """The reference's water fixture density, at its own points, to 1e-12."""
    d = np.load(os.path.join(GOLDEN, "water_sto3g_hf_exp_120_110.npz"))
    rho = io.basis_density(io.load_fchk(FCHK), d["points"])
    np.testing.assert_allclose(rho, d["dens"], rtol=1e-12, atol=0)


def test_part_gen_grid_integrates_the_electrons():
    f = io.load_fchk(FCHK)
    coords = f["Current cartesian coordinates"].reshape(-1, 3)
    grid = io.part_gen_grid(f["Atomic numbers"], coords, nrad=150, nang=194)
    assert grid.size == 3 * 150 * 194
    assert (grid.weights >= 0).all()
    rg = grid.atgrids[0].rgrid
    assert np.all(np.diff(rg.points) > 0) and rg.points[0] > 1e-4
    # a Gauss-Chebyshev rule of the first kind: sum w / sqrt(1 - x^2) = pi
    rho = io.basis_density(f, grid.points)
    assert grid.integrate(rho) == pytest.approx(10.0, abs=1e-4)


def test_density_file_round_trip(tmp_path):
    f = io.load_fchk(FCHK)
    fn = str(tmp_path / "density.npz")
    grid, rho = io.generate_density(FCHK, fn, nrad=40, nang=50)
    with np.load(fn) as data:
        keys = set(data.files)
    for k in ("atcoords", "atnums", "atcorenums", "points", "weights", "aim_weights", "cellvecs", "nelec", "density",
              "atom_idxs", "atom0/points", "atom2/rgrid/weights", "atom1/shell_idxs"):
        assert k in keys
    g2, data = io.load_density(fn)
    np.testing.assert_array_equal(g2.points, grid.points)
    np.testing.assert_array_equal(g2.weights, grid.weights)
    np.testing.assert_array_equal(data["density"], rho)
    assert g2.integrate(data["density"]) == pytest.approx(grid.integrate(rho), rel=1e-15)
    for a in range(3):
        at = g2.atgrids[a]
        sums = np.add.reduceat(at.sph_weights, at.indices[:-1])
        np.testing.assert_allclose(sums, 1.0, rtol=1e-12)
        np.testing.assert_allclose(at.points, grid.atgrids[a].points)
    assert int(data["nelec"]) == f["Number of electrons"]
