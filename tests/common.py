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


def load_msa_fixture():
    """The reference's monosilicic-acid fixture (HF/LANL2DZ, pseudo numbers of
    the LANL2DZ cores; tests/cached/monosilicic_acid_hf_lan_fchk_exp:5e-4:2e1:120:110.npz)
    on the rebuilt grid, density reordered as for water."""
    d = np.load(os.path.join(GOLDEN, "monosilicic_acid_hf_lan_exp_120_110.npz"))
    coords, nums, pseudo, dens, points = d["coordinates"], d["numbers"], d["pseudo_numbers"], d["dens"], d["points"]
    grid = MolGrid.from_size(nums, coords, 110, exp_radial_grid(120), BeckeWeights())
    key = lambda P: [tuple(np.round(row, 8)) for row in P]  # noqa: E731
    idx = {row: i for i, row in enumerate(key(points))}
    order = np.array([idx[row] for row in key(grid.points)])
    assert np.abs(points[order] - grid.points).max() < 1e-6
    return coords, nums, pseudo, grid, dens[order]


def proatom_files(numbers, max_cation, max_anion, level):
    """The reference's atom files of the given elements and charges
    (tests/common.py:64-75 of the reference)."""
    out = []
    for number in numbers:
        for nelec in number - np.arange(max_anion, max_cation + 1):
            path = os.path.join(GOLDEN, "proatoms", "atom_%s_Z%.2i_N%.2i_pow.npz" % (level, number, nelec))
            if os.path.isfile(path):
                out.append(path)
    return out


def load_proatomdb(numbers, max_cation, max_anion, level):
    """A ProAtomDB of the package from the reference's atom files (the
    records of tests/common.py:78-110: PowerRTransform grids)."""
    from horton_part_amd.grid import PowerRTransform, UniformInteger
    from horton_part_amd.proatomdb import ProAtomDB, ProAtomRecord

    records = []
    for path in proatom_files(numbers, max_cation, max_anion, level):
        with np.load(path) as d:
            rmin, rmax, n = d["rgrid"]
            rgrid = PowerRTransform(rmin, rmax, int(n) - 1).transform_1d_grid(UniformInteger(int(n)))
            pseudo = d["pseudo_number"] if "pseudo_number" in d.files else None
            records.append(ProAtomRecord(int(d["number"]), int(d["charge"]), float(d["energy"]), rgrid, d["dens"],
                                         d["deriv"], pseudo_number=pseudo))
    return ProAtomDB(records)


# tests/test_wpart.py:70-95 of the reference (values from HiPart; tolerance 2e-3)
WATER_H_CHARGES = np.array([-0.246171541212, 0.123092011074, 0.123079530138])
WATER_HI_CHARGES = np.array([-0.4214, 0.2107, 0.2107])
WATER_IS_CHARGES = np.array([-0.490017586929, 0.245018706885, 0.244998880045])
# tests/test_wpart.py:139-217 of the reference (tolerance 4e-3)
MSA_H_CHARGES = np.array([0.56175431, -0.30002709, -0.28602105, -0.28335086, -0.26832878, 0.13681904, 0.14535691,
                          0.14206876, 0.15097682])
MSA_HI_CHARGES = np.array([1.14305602, -0.52958298, -0.51787452, -0.51302759, -0.50033981, 0.21958586, 0.23189187,
                           0.22657354, 0.23938904])
MSA_IS_CHARGES = np.array([1.1721364, -0.5799622, -0.5654549, -0.5599638, -0.5444145, 0.2606699, 0.2721848, 0.2664377,
                           0.2783666])
