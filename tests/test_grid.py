"""Grid substrate: Lebedev rules, radial rule, atom/molecular grids, Becke.

Pins: exactness of each Lebedev rule on all real spherical harmonics up to
its degree; the reference water fixture's radii and directions; closed-form
integrals of Slater and Gaussian densities; partition of unity of Becke's
cell functions; the spherical-average identities the partitioning relies on.
"""

import numpy as np
import pytest

from horton_part_amd.grid import (
    AtomGrid,
    BeckeWeights,
    ExpRTransform,
    MolGrid,
    UniformInteger,
    cartesian_powers,
    exp_radial_grid,
    solid_harmonics,
)
from horton_part_amd.lebedev import LEBEDEV_NPOINTS, lebedev_rule

from common import GOLDEN


def _real_sph_harm(lmax, u):
    """Real orthonormal spherical harmonics via scipy (independent of the package)."""
    from scipy.special import sph_harm_y

    theta = np.arccos(np.clip(u[:, 2], -1, 1))
    phi = np.arctan2(u[:, 1], u[:, 0])
    out = []
    for l in range(lmax + 1):
        out.append(sph_harm_y(l, 0, theta, phi).real)
        for m in range(1, l + 1):
            y = sph_harm_y(l, m, theta, phi)
            out.append(np.sqrt(2) * y.real)
            out.append(np.sqrt(2) * y.imag)
    return np.array(out)


@pytest.mark.parametrize("npoint", sorted(LEBEDEV_NPOINTS))
def test_lebedev_exactness(npoint):
    deg = LEBEDEV_NPOINTS[npoint]
    u, w = lebedev_rule(npoint)
    assert u.shape == (npoint, 3)
    np.testing.assert_allclose(np.linalg.norm(u, axis=1), 1.0, atol=1e-15)
    assert abs(w.sum() - 1.0) < 1e-14
    assert (w > 0).all()
    # no duplicate points
    assert len({tuple(np.round(p, 12)) for p in u}) == npoint
    Y = _real_sph_harm(deg, u)
    integrals = Y @ w * np.sqrt(4 * np.pi)  # int Y dOmega / sqrt(4 pi) ... = delta_l0
    expected = np.zeros(len(Y))
    expected[0] = 1.0
    np.testing.assert_allclose(integrals, expected, atol=2e-14)
    # and it is not exact one degree higher (the rule is not over-resolved)
    Y2 = _real_sph_harm(deg + 1, u)
    assert np.abs(Y2 @ w)[1:].max() > 1e-10 or deg % 2 == 1


def test_fixture_radii_and_directions():
    """The reference water fixture lies on ExpRTransform(5e-4, 20, 120) x Lebedev-110."""
    d = np.load(f"{GOLDEN}/water_sto3g_hf_exp_120_110.npz")
    pts, coords = d["points"], d["coordinates"]
    rg = exp_radial_grid(120)
    u, _ = lebedev_rule(110)
    for a in range(3):
        seg = pts[a * 13200 : (a + 1) * 13200] - coords[a]
        r = np.linalg.norm(seg, axis=1).reshape(120, 110)
        np.testing.assert_allclose(r, np.repeat(rg.points[:, None], 110, axis=1), rtol=2e-12)
        dirs = seg / np.linalg.norm(seg, axis=1)[:, None]
        for sh in (0, 60, 119):
            s = dirs[sh * 110 : (sh + 1) * 110]
            dmin = np.min(np.linalg.norm(s[:, None, :] - u[None], axis=2), axis=1)
            assert dmin.max() < 1e-9


def test_exp_rtransform():
    rg = ExpRTransform(1e-3, 10.0, 49).transform_1d_grid(UniformInteger(50))
    assert rg.points[0] == pytest.approx(1e-3)
    assert rg.points[-1] == pytest.approx(10.0)
    # int_0^inf 4 pi r^2 exp(-2r) dr = pi (1s Slater normalised to 1 with a^3/8pi, a = 2)
    f = 8.0 / (8 * np.pi) * np.exp(-2 * rg.points)
    assert rg.integrate(4 * np.pi * rg.points**2, f) == pytest.approx(1.0, abs=1e-6)


def test_atom_grid_integrates_slater_and_gaussian():
    rg = exp_radial_grid(120)
    ag = AtomGrid(rg, sizes=[110], center=np.array([0.3, -0.2, 0.1]))
    r = np.linalg.norm(ag.points - ag.center, axis=1)
    for alpha in (1.5, 2.0, 9.0):
        slater = alpha**3 / (8 * np.pi) * np.exp(-alpha * r)
        assert ag.integrate(slater) == pytest.approx(1.0, abs=2e-8)
        gauss = (alpha / np.pi) ** 1.5 * np.exp(-alpha * r**2)
        assert ag.integrate(gauss) == pytest.approx(1.0, abs=2e-8)


def test_spherical_average_identities():
    rg = exp_radial_grid(80)
    ag = AtomGrid(rg, sizes=[194])
    d = ag.points
    r = np.linalg.norm(d, axis=1)
    # spherically symmetric f: the shell average is f(r_i) at the knots
    f = np.exp(-1.3 * r) * (1 + r)
    spl = ag.spherical_average(f)
    np.testing.assert_allclose(spl(rg.points), np.exp(-1.3 * rg.points) * (1 + rg.points), rtol=1e-13)
    # population identity sum_i 4 pi r_i^2 w_i fbar_i = sum_p w_p f_p, any f
    g = f * (1 + 0.3 * d[:, 0] + 0.2 * d[:, 1] * d[:, 2] / (1 + r**2))
    fbar = spl_vals = ag.spherical_average(g)(rg.points)
    assert np.sum(4 * np.pi * rg.points**2 * rg.weights * spl_vals) == pytest.approx(ag.integrate(g), rel=1e-13)
    # a pure l=1 angular dependence averages to zero
    h = np.exp(-r) * d[:, 2]
    assert np.abs(ag.spherical_average(h)(rg.points)).max() < 1e-14
    assert fbar.shape == (80,)


def test_becke_partition_of_unity():
    rng = np.random.default_rng(4)
    coords = rng.uniform(-2, 2, size=(4, 3))
    nums = np.array([1, 6, 8, 7])
    bw = BeckeWeights()
    pts = rng.uniform(-4, 4, size=(300, 3))
    tot = np.zeros(len(pts))
    for a in range(4):
        idx = np.array([0, 0, 0, 0, 0])
        idx[a + 1 :] = len(pts)
        tot += bw.generate_weights(pts, coords, nums, indices=idx)
    np.testing.assert_allclose(tot, 1.0, atol=1e-12)


def test_molgrid_integrates_density():
    rng = np.random.default_rng(2)
    coords = np.array([[0.0, 0.0, 0.0], [1.4, 0.2, -0.3], [-0.5, 1.1, 0.4]])
    nums = np.array([8, 1, 1])
    grid = MolGrid.from_size(nums, coords, 110, exp_radial_grid(100), BeckeWeights())
    rho = np.zeros(grid.size)
    total = 0.0
    for c, z in zip(coords, nums):
        a = rng.uniform(1.5, 6.0)
        rho += z * a**3 / (8 * np.pi) * np.exp(-a * np.linalg.norm(grid.points - c, axis=1))
        total += z
    assert grid.integrate(rho) == pytest.approx(total, rel=1e-5)
    assert grid.indices.tolist() == [0, 11000, 22000, 33000]


def test_solid_harmonics_closed_forms():
    rng = np.random.default_rng(0)
    d = rng.normal(size=(20, 3))
    x, y, z = d.T
    R = solid_harmonics(3, d)
    r2 = x * x + y * y + z * z
    np.testing.assert_allclose(R[0], 1.0)
    np.testing.assert_allclose(R[1:4], [z, x, y])
    np.testing.assert_allclose(R[4], 0.5 * (3 * z * z - r2))
    np.testing.assert_allclose(R[8], np.sqrt(3) * x * y)
    np.testing.assert_allclose(R[15], np.sqrt(5 / 8) * y * (3 * x * x - y * y))
    assert cartesian_powers(2)[4:] == [(2, 0, 0), (1, 1, 0), (1, 0, 1), (0, 2, 0), (0, 1, 1), (0, 0, 2)]


def test_moments_translation_closed_form():
    """Dipole of a displaced spherical density = population x displacement."""
    rg = exp_radial_grid(100)
    ag = AtomGrid(rg, sizes=[110])
    shift = np.array([0.05, -0.02, 0.03])
    r = np.linalg.norm(ag.points - shift, axis=1)
    f = 8 / (8 * np.pi) * np.exp(-2 * r)
    cart = ag.moments(1, np.zeros((1, 3)), f, "cartesian").ravel()
    assert cart[0] == pytest.approx(1.0, abs=1e-7)
    np.testing.assert_allclose(cart[1:4], shift, atol=1e-7)
    pure = ag.moments(1, np.zeros((1, 3)), f, "pure").ravel()
    np.testing.assert_allclose(pure[1:4], shift[[2, 0, 1]], atol=1e-7)
    rad = ag.moments(2, np.zeros((1, 3)), f, "radial").ravel()
    assert rad[0] == pytest.approx(1.0, abs=1e-7)
