"""Shared helpers of the test-suite."""

import os

import numpy as np

from horton_part_amd.grid import BeckeWeights, MolGrid, exp_radial_grid

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

# tests/test_wpart.py:95 of the reference: MBIS charges of HF/STO-3G water
WATER_MBIS_CHARGES = np.array([-0.61891067, 0.3095756, 0.30932584])


def load_water_fixture():
    """The reference's water fixture (tests/cached/water_sto3g_hf_g03_fchk_exp:5e-4:2e1:120:110.npz)
    on a rebuilt ExpRTransform(5e-4, 20, 120) x Lebedev-110 grid, density reordered
    onto this grid's point order (as tests/common.py:141-153 of the reference does)."""
    d = np.load(os.path.join(GOLDEN, "water_sto3g_hf_exp_120_110.npz"))
    coords, nums, pseudo, dens, points = d["coordinates"], d["numbers"], d["pseudo_numbers"], d["dens"], d["points"]
    grid = MolGrid.from_size(nums, coords, 110, exp_radial_grid(120), BeckeWeights())
    key = lambda P: [tuple(np.round(row, 8)) for row in P]  # noqa: E731
    idx = {row: i for i, row in enumerate(key(points))}
    order = np.array([idx[row] for row in key(grid.points)])
    assert np.abs(points[order] - grid.points).max() < 1e-6
    return coords, nums, pseudo, grid, dens[order]
