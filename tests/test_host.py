"""Host-side logic of the package: cache, type checks, registry, basis
tables, scheme set-up, and the loud failure when the library is missing."""

import os
import subprocess
import sys

import numpy as np
import pytest

from horton_part_amd import GaussianISAWPart, LinearISAWPart, MBISWPart, workloads, wpart_schemes
from horton_part_amd.basis import ExpBasisFuncHelper, evaluate_function, shell_norm
from horton_part_amd.cache import Cache, JustOnceClass, just_once
from horton_part_amd.gisa import gaussian_overlap_matrix
from horton_part_amd.mbis import get_initial_mbis_propars, get_nshell
from horton_part_amd.utils import typecheck_geo
from oracle import horton_oracle as ho


def test_cache_semantics():
    c = Cache()
    a, new = c.load("x", alloc=3, tags="o")
    assert new and a.shape == (3,)
    a[:] = 2.0
    b, new = c.load("x", alloc=3, tags="o")
    assert not new and b is a
    with pytest.raises(TypeError):
        c.load("x", alloc=4, tags="o")
    with pytest.raises(ValueError):
        c.load("x", alloc=3, tags="p")
    c.dump("y", 5, tags="o")
    assert c["y"] if hasattr(c, "__getitem__") else c.load("y") == 5
    assert sorted(c.iterkeys(tags="o")) == ["x", "y"]
    c.clear_item("x")
    assert "x" not in c
    a2, new = c.load("x", alloc=3, tags="o")
    assert new and a2 is a and (a2 == 0).all()
    with pytest.raises(KeyError):
        c.load("missing")
    assert c.load("missing", default=7) == 7
    calls = []
    c.add_lazy("lazy", lambda: calls.append(1) or np.ones(2), tags="o")
    assert "lazy" in c and calls == []
    assert (c.load("lazy") == 1).all() and calls == [1]
    c.load("lazy")
    assert calls == [1]


def test_just_once():
    class A(JustOnceClass):
        def __init__(self):
            super().__init__()
            self.n = 0

        @just_once
        def go(self):
            self.n += 1

    a = A()
    a.go()
    a.go()
    assert a.n == 1 and "go" in a.time_usage
    a.clear()
    a.go()
    assert a.n == 2


def test_typecheck_geo():
    c = np.zeros((2, 3))
    n = np.array([1, 8])
    natom, c2, n2, p2 = typecheck_geo(c, n, None)
    assert natom == 2 and p2.dtype == float
    with pytest.raises(TypeError):
        typecheck_geo(c, n.astype(np.int32), None)
    with pytest.raises(TypeError):
        typecheck_geo(np.zeros((2, 2)), n, None)
    with pytest.raises(TypeError):
        typecheck_geo(c, n, np.zeros(3))


def test_registry_and_init_params():
    assert wpart_schemes("mbis") is MBISWPart
    assert wpart_schemes("gisa") is GaussianISAWPart
    assert wpart_schemes("lisa") is LinearISAWPart
    from horton_part_amd import HirshfeldIWPart, HirshfeldWPart, ISAWPart

    assert wpart_schemes("h") is HirshfeldWPart
    assert wpart_schemes("hi") is HirshfeldIWPart
    assert wpart_schemes("is") is ISAWPart
    with pytest.raises(NotImplementedError):
        wpart_schemes("mulliken")
    for z in (1, 2, 3, 8, 17, 21, 44, 104):
        assert get_nshell(z) == ho.get_nshell(z)
        np.testing.assert_array_equal(get_initial_mbis_propars(z), ho.get_initial_mbis_propars(z))


def test_basis_helper_tables():
    h = ExpBasisFuncHelper.from_function_type("gauss")
    off, alpha, order, norm = h.device_tables(np.array([1, 6, 8]))
    assert off.tolist() == [0, 4, 10, 16]
    np.testing.assert_allclose(alpha[:4], h.get_exponent(1))
    r = np.array([0.3, 1.0])
    for k in range(16):
        np.testing.assert_allclose(norm[k] * np.exp(-alpha[k] * r**order[k]), evaluate_function(order[k], 1.0, alpha[k], r), rtol=1e-15)
    assert shell_norm(1.0, 2.0) == pytest.approx(8.0 / (8 * np.pi))
    assert shell_norm(2.0, 1.0) == pytest.approx(np.pi**-1.5)
    u = ExpBasisFuncHelper.uniform([1, 6], [0.5, 5.0])
    assert u.get_nshell(6) == 2 and (u.get_initial(1) == 0.5).all()


def test_evaluate_function_reference_derivative():
    """core/basis.py:175: the derivative omits the factor alpha."""
    f, df = evaluate_function(2, 1.0, 3.0, np.array([1.0]), 1)
    assert df[0] == pytest.approx(-2 * f[0])


def test_gaussian_overlap_spd():
    P = gaussian_overlap_matrix([220.1, 65.66, 25.98, 1.685, 0.686, 0.2311])
    assert np.allclose(P, P.T)
    assert np.linalg.eigvalsh(P).min() > 0
    np.testing.assert_allclose(P, ho.gaussian_overlap_matrix([220.1, 65.66, 25.98, 1.685, 0.686, 0.2311]))


def test_scheme_descriptions_without_gpu():
    w = workloads.chain(natom=3, seed=3, nrad=20, nang=26)
    p = MBISWPart(w.coordinates, w.numbers, w.pseudo_numbers, w.grid, w.density, local=True)  # unknown kwargs ignored
    md = p._molecule_arrays()
    assert md["nshell"] == 60 and md["shell_offsets"][-1] == w.npoint
    np.testing.assert_allclose(np.add.reduceat(md["sph_weights"], md["shell_offsets"][:-1]), 1.0, atol=1e-14)
    sd = p._scheme_arrays()
    assert sd["npar"] == 2 * sd["fn_offsets"][-1]
    g = GaussianISAWPart(w.coordinates, w.numbers, w.pseudo_numbers, w.grid, w.density)
    gd = g._scheme_arrays()
    assert gd["qp_matrix"].size == sum((gd["fn_offsets"][a + 1] - gd["fn_offsets"][a]) ** 2 for a in range(3))
    lisa = LinearISAWPart(w.coordinates, w.numbers, w.pseudo_numbers, w.grid, w.density, solver="m-newton", solver_options={"tau": 0.5})
    ld = lisa._scheme_arrays()
    assert ld["solver"] == 4 and ld["newton_tau"] == 0.5
    with pytest.raises(NotImplementedError):
        LinearISAWPart(w.coordinates, w.numbers, w.pseudo_numbers, w.grid, w.density, solver="cvxopt")._scheme_arrays()
    with pytest.raises(TypeError):
        MBISWPart(w.coordinates, w.numbers.astype(np.int32), w.pseudo_numbers, w.grid, w.density)


def test_grid_without_atgrids_raises():
    w = workloads.water(nrad=10, nang=6)

    class NoAt:
        atgrids = None
        points = w.grid.points

    with pytest.raises(ValueError):
        MBISWPart(w.coordinates, w.numbers, w.pseudo_numbers, NoAt(), w.density)


def test_missing_library_fails_loudly(tmp_path):
    code = (
        "import sys; sys.path.insert(0, 'horton-part_amd');"
        "from horton_part_amd import _native;"
        "\ntry:\n _native.load()\nexcept RuntimeError as e:\n print('RAISED', e)"
    )
    env = dict(os.environ, HORTONPART_LIB=str(tmp_path / "absent.so"))
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    out = subprocess.run([sys.executable, "-c", code], env=env, cwd=root, capture_output=True, text=True)
    assert "RAISED" in out.stdout


def test_serving_precision_policy():
    """"auto" runs GISA requests mixed and MBIS requests in FP64; the scheme
    arrays carry full_sum (0: tail skipping allowed)."""
    from horton_part_amd.serving import resolve_precision

    assert resolve_precision("auto", 1) == "mixed"
    assert resolve_precision("auto", 0) == "fp64"
    assert resolve_precision("fp64", 1) == "fp64"
    assert resolve_precision("mixed", 0) == "mixed"
    with pytest.raises(ValueError):
        resolve_precision("fp16", 0)
    w = workloads.water(nrad=20, nang=26)
    part = MBISWPart(w.coordinates, w.numbers, w.pseudo_numbers, w.grid, w.density)
    assert part._scheme_arrays()["full_sum"] == 0
    part = MBISWPart(w.coordinates, w.numbers, w.pseudo_numbers, w.grid, w.density, full_sum=True)
    assert part._scheme_arrays()["full_sum"] == 1
