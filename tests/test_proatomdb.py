"""Host side of the tabulated schemes (CPU): the pro-atom database API
(core/proatomdb.py:34-447), the device tables of its splines, and the
argument checks of the Hirshfeld, Hirshfeld-I and ISA classes."""

import numpy as np
import pytest
from scipy.interpolate import CubicHermiteSpline, CubicSpline

import common as C
from horton_part_amd import HirshfeldIWPart, HirshfeldWPart, ISAWPart, ProAtomDB, ProAtomRecord
from horton_part_amd.grid import OneDGrid, PowerRTransform, UniformInteger
from horton_part_amd.hirshfeld import check_proatomdb
from horton_part_amd.proatomdb import spline_table
from oracle import horton_oracle as ho


def test_power_rgrid_matches_oracle():
    g = PowerRTransform(2.5e-7, 16.3, 33).transform_1d_grid(UniformInteger(34))
    o = ho.power_rgrid(2.5e-7, 16.3, 34)
    np.testing.assert_allclose(g.points, o.points, rtol=1e-15)
    np.testing.assert_allclose(g.weights, o.weights, rtol=1e-14)
    assert g.points[0] == 2.5e-7 and abs(g.points[-1] - 16.3) < 1e-12


def test_database_api():
    db = C.load_proatomdb([8, 6, 1], 1, -1, "hf_sto3g")
    assert db.size == 8 and db.get_numbers() == [1, 6, 8]
    assert db.get_charges(8) == [1, 0, -1] and db.get_charges(1) == [0, -1]
    r = db.get_record(8, 0)
    assert r.population == 8 and r.pseudo_population == 8 and r.pseudo_number == 8
    # ionisation energy from the cation (update_safe), one-electron atom: -E
    assert r.ipot_energy == pytest.approx(db.get_record(8, 1).energy - r.energy)
    assert db.get_record(1, 0).ipot_energy == -db.get_record(1, 0).energy
    # linear combinations and their derivatives
    rho, deriv = db.get_rho(8, {0: 0.25, -1: 0.75}, do_deriv=True)
    np.testing.assert_allclose(rho, 0.25 * r.rho + 0.75 * db.get_record(8, -1).rho, rtol=1e-15)
    np.testing.assert_allclose(deriv, 0.25 * r.deriv + 0.75 * db.get_record(8, -1).deriv, rtol=1e-15)
    geo = db.get_rho(8, {0: 0.5, 1: 0.5}, combine="geometric")
    np.testing.assert_allclose(geo, np.sqrt(r.rho * db.get_record(8, 1).rho), rtol=1e-12)
    with pytest.raises(ValueError):
        db.get_rho(8, {0: 1.0}, combine="harmonic")
    with pytest.raises(TypeError):
        db.get_rho(8, "0")
    with pytest.raises(KeyError):
        db.get_record(8, -2)
    # radial moment and population radii
    g = db.get_rgrid(8)
    assert r.get_moment(2) == pytest.approx(g.integrate(r.rho, g.points**2))
    idx, radii = r.compute_radii([1.0, 7.9])
    assert radii[0] < radii[1] and idx[0] < idx[1]


def test_database_keeps_lowest_energy_and_checks_grids():
    g = OneDGrid(np.array([0.1, 0.5, 1.0, 2.0]), np.ones(4))
    a = ProAtomRecord(1, 0, -0.4, g, np.ones(4))
    b = ProAtomRecord(1, 0, -0.5, g, 2 * np.ones(4))
    db = ProAtomDB([a, b])
    assert db.size == 1 and db.get_record(1, 0) is b
    assert a != b and b == ProAtomRecord(1, 0, -0.5, g, 2 * np.ones(4))
    other = ProAtomRecord(1, -1, -0.1, OneDGrid(np.array([0.1, 0.6, 1.0, 2.0]), np.ones(4)), np.ones(4))
    with pytest.raises(ValueError):
        ProAtomDB([a, other])
    # a bigger anion with a higher energy than the neutral atom is unsafe
    anion = ProAtomRecord(1, -1, -0.1, g, np.ones(4))
    db = ProAtomDB([b, anion])
    assert not db.get_record(1, -1).safe and db.get_charges(1, safe=True) == [0]


def test_normalize_and_compact():
    db = C.load_proatomdb([8], 1, -1, "hf_sto3g")
    n0 = db.get_rgrid(8).size
    db.compact(1e-4)
    assert db.get_rgrid(8).size < n0
    assert all(db.get_record(8, q).rho.size == db.get_rgrid(8).size for q in db.get_charges(8))
    db.normalize()
    g = db.get_rgrid(8)
    for q in db.get_charges(8):
        assert g.integrate(db.get_record(8, q).rho) == pytest.approx(8 - q, rel=1e-14)


def test_spline_tables_match_scipy():
    """The device layout of a record's spline: rows (c0, c1, c2, c3) of scipy's
    PPoly, Hermite with derivatives, not-a-knot without."""
    db = C.load_proatomdb([8], 0, 0, "hf_sto3g")
    r = db.get_record(8, 0)
    x = db.get_rgrid(8).points
    t = spline_table(x, r.rho, r.deriv)
    np.testing.assert_array_equal(t[:-1], CubicHermiteSpline(x, r.rho, r.deriv).c.T)
    assert not t[-1].any()
    t2 = spline_table(x, r.rho)
    np.testing.assert_array_equal(t2[:-1], CubicSpline(x, r.rho).c.T)
    # evaluation in the device's term order, in and beyond the knots
    s = db.get_spline(8, 0)
    rr = np.array([1e-9, 0.3, 2.0, x[-1], 18.0, 25.0])
    i = np.clip(np.searchsorted(x, rr, side="right") - 1, 0, x.size - 2)
    d = rr - x[i]
    z2 = d * d
    v = t[i, 3] + t[i, 2] * d + t[i, 1] * z2 + t[i, 0] * (z2 * d)
    np.testing.assert_array_equal(v, s(rr))


def test_scheme_arguments():
    coords, nums, pseudo, grid, dens = C.load_water_fixture()
    db = C.load_proatomdb([8, 6, 1], 1, -1, "hf_sto3g")
    with pytest.raises(ValueError):
        check_proatomdb(nums, pseudo + np.array([0.0, 0.0, 1.0]), db)
    with pytest.raises(KeyError):
        HirshfeldWPart(coords, np.array([8, 1, 7]), pseudo, grid, dens, db)
    with pytest.raises(NotImplementedError):
        HirshfeldIWPart(coords, nums, pseudo, grid, dens, db, grid_type=2)
    with pytest.raises(NotImplementedError):
        ISAWPart(coords, nums, pseudo, grid, dens, grid_type=3)
