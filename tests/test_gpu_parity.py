"""GPU parity: the HIP path (through libhortonpart.so) against the CPU oracle.

Bar for the plain FP64 path: charges and populations within 1e-10 e, pro-atom
parameters within 1e-9 relative, identical outer iteration counts and
identical per-atom inner iteration counts of the last outer iteration.
The mixed path is held to max |dq| <= 1e-6 e against the FP64 oracle.
"""

import sys
import warnings

import numpy as np
import pytest

from conftest import require_gpu
from horton_part_amd import ExpBasisFuncHelper, GaussianISAWPart, LinearISAWPart, MBISWPart, workloads
from oracle.horton_oracle import OracleBasis, OracleGISA, OracleLISA, OracleMBIS

from common import WATER_MBIS_CHARGES, load_water_fixture

pytestmark = pytest.mark.gpu

warnings.simplefilter("ignore")


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _compare(part, cache, oracle, qtol=1e-10, ptol=1e-9, check_inner=True):
    dq = np.max(np.abs(part["charges"] - cache["charges"]))
    dpop = np.max(np.abs(part["populations"] - cache["populations"]))
    p1, p2 = part["propars"], cache["propars"]
    rel = np.max(np.abs(p1 - p2) / np.maximum(np.abs(p2), 1e-300))
    _say(f"  niter gpu={part['niter']} oracle={cache['niter']} max|dq|={dq:.2e} rel dpar={rel:.2e}")
    if part["niter"] == cache["niter"]:
        hq = np.abs(part["history_charges"] - cache["history_charges"]).max(axis=1)
        first = int(np.argmax(hq > 1e-12)) if (hq > 1e-12).any() else -1
        _say(f"  history max|dq| first>1e-12 at iteration {first}; last values {hq[-3:]}")
        _say(f"  inner niter gpu={list(part.inner_niter)} oracle={[s.get('niter', 1) for s in oracle.inner_stats[-1]]}")
    assert part["niter"] == cache["niter"]
    assert dq <= qtol
    assert dpop <= qtol
    assert rel <= ptol
    if check_inner:
        inner_oracle = [s.get("niter", 1) for s in oracle.inner_stats[-1]]
        assert list(part.inner_niter) == inner_oracle
    np.testing.assert_allclose(part["history_changes"], cache["history_changes"], rtol=1e-6, atol=1e-14)
    np.testing.assert_allclose(part["history_entropies"], cache["history_entropies"], rtol=1e-9, atol=1e-12)


def test_mbis_water_fixture_kat():
    """Reference KAT (tests/test_wpart.py:95-101) through the device path."""
    require_gpu()
    _say("case: MBIS reference water fixture, 120x110, threshold 1e-6")
    coords, nums, pseudo, grid, dens = load_water_fixture()
    part = MBISWPart(coords, nums, pseudo, grid, dens)
    part.do_all()
    oracle = OracleMBIS(coords, nums, pseudo, grid, dens)
    cache = oracle.do_partitioning()
    _compare(part, cache, oracle)
    assert np.abs(part["charges"] - WATER_MBIS_CHARGES).max() < 2e-3
    assert part["charges"] == pytest.approx(part["valence_charges"] + part["core_charges"])
    assert (part["core_charges"] > 0).all()
    assert (part["valence_charges"] < 0).all()
    assert (part["valence_widths"] > 0).all()
    # moments (core/base.py:329-402) against the oracle, and the monopole pin
    cart, pure, rad = oracle.do_moments()
    np.testing.assert_allclose(part["cartesian_multipoles"], cart, rtol=1e-9, atol=1e-11)
    np.testing.assert_allclose(part["pure_multipoles"], pure, rtol=1e-9, atol=1e-11)
    np.testing.assert_allclose(part["radial_moments"], rad, rtol=1e-9, atol=1e-11)
    assert np.abs(part["charges"] - part["cartesian_multipoles"][:, 0]).max() < 1e-3
    assert np.abs(part["charges"] - part["pure_multipoles"][:, 0]).max() < 1e-3
    # per-point weights, promolecule and spherical averages
    for a in range(3):
        np.testing.assert_allclose(part[f"at_weights_{a}"], oracle.cache[f"at_weights_{a}"], rtol=1e-12, atol=1e-15)
        np.testing.assert_allclose(part[f"spherical_average_{a}"], oracle.cache[f"spherical_average_{a}"], rtol=1e-11, atol=1e-14)
    np.testing.assert_allclose(part["promoldens"], oracle.cache["promoldens"], rtol=1e-12, atol=1e-300)


def test_mbis_synthetic_water_degenerate_shells():
    """Water-like workload: nearly equal O shell exponents drive the inner
    solver to its 2000-iteration cap (mbis.py:162 warning path)."""
    require_gpu()
    _say("case: MBIS synthetic water seed 1000, 60x110, threshold 1e-8, maxiter 40")
    w = workloads.water(nrad=60, nang=110)
    kw = dict(threshold=1e-8, maxiter=40)
    part = MBISWPart(w.coordinates, w.numbers, w.pseudo_numbers, w.grid, w.density, **kw)
    part.do_partitioning()
    oracle = OracleMBIS(w.coordinates, w.numbers, w.pseudo_numbers, w.grid, w.density, **kw)
    cache = oracle.do_partitioning()
    _compare(part, cache, oracle, qtol=1e-10, ptol=1e-8)


@pytest.mark.parametrize("nrad", [150, 240, 300])
def test_mbis_solver_layouts(nrad):
    """The MBIS solver's layouts beyond two radial points per lane: one
    wavefront per atom with three (150 points) or four (240) points per lane,
    and one workgroup per atom, one point per lane, beyond 256 (300)."""
    require_gpu()
    _say(f"case: MBIS chain 4 atoms seed 5, {nrad}x26 (solver layout of {nrad} radial points)")
    w = workloads.chain(natom=4, seed=5, nrad=nrad, nang=26)
    kw = dict(maxiter=60)
    part = MBISWPart(w.coordinates, w.numbers, w.pseudo_numbers, w.grid, w.density, **kw)
    part.do_partitioning()
    oracle = OracleMBIS(w.coordinates, w.numbers, w.pseudo_numbers, w.grid, w.density, **kw)
    cache = oracle.do_partitioning()
    _compare(part, cache, oracle)


def test_mbis_chain():
    require_gpu()
    _say("case: MBIS chain 6 atoms seed 7, 50x110")
    w = workloads.chain(natom=6, seed=7, nrad=50, nang=110)
    part = MBISWPart(w.coordinates, w.numbers, w.pseudo_numbers, w.grid, w.density, maxiter=80)
    part.do_partitioning()
    oracle = OracleMBIS(w.coordinates, w.numbers, w.pseudo_numbers, w.grid, w.density, maxiter=80)
    cache = oracle.do_partitioning()
    _compare(part, cache, oracle)


def test_gisa_chain():
    require_gpu()
    _say("case: GISA chain 5 atoms seed 11, 50x110")
    w = workloads.chain(natom=5, seed=11, nrad=50, nang=110)
    part = GaussianISAWPart(w.coordinates, w.numbers, w.pseudo_numbers, w.grid, w.density, maxiter=60)
    part.do_partitioning()
    bs = part.bs_helper
    basis = OracleBasis(bs.orders, bs.exponents, bs.initials)
    oracle = OracleGISA(w.coordinates, w.numbers, w.pseudo_numbers, w.grid, w.density, maxiter=60, basis=basis)
    cache = oracle.do_partitioning()
    _compare(part, cache, oracle, check_inner=False)


def _slater6(numbers):
    alphas = 0.5 * 100.0 ** (np.arange(6) / 5.0)
    return ExpBasisFuncHelper.uniform(numbers, alphas, order=1.0)


def _lisa_pair(w, basis_name, solver, precision="fp64", **kw):
    helper = _slater6(w.numbers) if basis_name == "slater6" else ExpBasisFuncHelper.from_function_type("gauss")
    part = LinearISAWPart(
        w.coordinates, w.numbers, w.pseudo_numbers, w.grid, w.density, solver=solver, basis_func=helper, precision=precision, **kw
    )
    basis = OracleBasis(helper.orders, helper.exponents, helper.initials)
    oracle = OracleLISA(w.coordinates, w.numbers, w.pseudo_numbers, w.grid, w.density, basis=basis, solver=solver, **kw)
    return part, oracle


@pytest.mark.parametrize("basis_name,solver", [("slater6", "sc"), ("gauss", "sc"), ("gauss", "m-newton"), ("slater6", "m-newton")])
def test_lisa_chain(basis_name, solver):
    require_gpu()
    _say(f"case: LISA {solver} basis {basis_name} chain 4 atoms seed 5, 50x110")
    w = workloads.chain(natom=4, seed=5, nrad=50, nang=110)
    part, oracle = _lisa_pair(w, basis_name, solver, maxiter=100)
    part.do_partitioning()
    cache = oracle.do_partitioning()
    _compare(part, cache, oracle, qtol=1e-10, ptol=1e-8)


@pytest.mark.parametrize("basis_name,solver", [("gauss", "m-newton"), ("gauss", "sc")])
def test_lisa_one_iteration(basis_name, solver):
    """One outer iteration from identical parameters: isolates the weights
    pass (spherical averages) from a single per-atom inner solve."""
    require_gpu()
    _say(f"case: LISA {solver} {basis_name} one outer iteration, chain 4 atoms seed 5, 50x110")
    w = workloads.chain(natom=4, seed=5, nrad=50, nang=110)
    part, oracle = _lisa_pair(w, basis_name, solver, maxiter=1)
    part.do_partitioning()
    cache = oracle.do_partitioning()
    for a in range(4):
        g, o = part[f"spherical_average_{a}"], oracle.cache[f"spherical_average_{a}"]
        _say(f"  atom {a}: max rel d(rhobar) = {np.max(np.abs(g - o) / np.maximum(np.abs(o), 1e-300)):.2e}")
    _compare(part, cache, oracle, qtol=1e-10, ptol=1e-8)


def test_lisa_exact_newton_raises_like_reference():
    """Exact Newton diverges on these densities until the Hessian is exactly
    singular; scipy.linalg.solve then raises LinAlgError (alisa.py:924) and
    the device path must raise the same exception."""
    require_gpu()
    _say("case: LISA exact newton, gauss basis, chain 4 atoms seed 5 (expects LinAlgError)")
    w = workloads.chain(natom=4, seed=5, nrad=50, nang=110)
    part, oracle = _lisa_pair(w, "gauss", "newton", maxiter=100)
    with pytest.raises(np.linalg.LinAlgError):
        oracle.do_partitioning()
    with pytest.raises(np.linalg.LinAlgError):
        part.do_partitioning()


def test_mbis_mixed_precision():
    require_gpu()
    _say("case: MBIS mixed precision + FP64 polish, chain 6 atoms seed 7")
    w = workloads.chain(natom=6, seed=7, nrad=50, nang=110)
    part = MBISWPart(w.coordinates, w.numbers, w.pseudo_numbers, w.grid, w.density, maxiter=80, precision="mixed")
    part.do_partitioning()
    oracle = OracleMBIS(w.coordinates, w.numbers, w.pseudo_numbers, w.grid, w.density, maxiter=80)
    cache = oracle.do_partitioning()
    dq = np.max(np.abs(part["charges"] - cache["charges"]))
    _say(f"  mixed niter={part['niter']} oracle niter={cache['niter']} max|dq|={dq:.2e}")
    assert dq <= 1e-6


def test_mbis_cutoff_radius():
    require_gpu()
    _say("case: MBIS with 6-bohr pro-atom cutoff, cluster 12 atoms seed 3, 40x50")
    w = workloads.cluster(natom=12, seed=3, nrad=40, nang=50)
    kw = dict(maxiter=30, cutoff_radius=6.0)
    part = MBISWPart(w.coordinates, w.numbers, w.pseudo_numbers, w.grid, w.density, **kw)
    part.do_partitioning()
    oracle = OracleMBIS(w.coordinates, w.numbers, w.pseudo_numbers, w.grid, w.density, **kw)
    cache = oracle.do_partitioning()
    _compare(part, cache, oracle)


# --- grid_type 2: per-atom solvers over the molecular grid ---------------------


def _run_pair(part, oracle):
    """Run both; if the oracle raises, the device must raise the same type."""
    try:
        cache = oracle.do_partitioning()
    except Exception as exc:  # noqa: BLE001 - compare exception types
        with pytest.raises(type(exc)):
            part.do_partitioning()
        _say(f"  both raised {type(exc).__name__}")
        return None
    part.do_partitioning()
    return cache


@pytest.mark.parametrize(
    "basis_name,solver",
    [("gauss", "sc"), ("slater6", "sc"), ("gauss", "m-newton"), ("slater6", "m-newton"), ("gauss", "newton")],
)
def test_gt2_lisa(basis_name, solver):
    """grid_type 2: LISA solves over the molecular grid; the Newton Hessian
    B diag(w rho_a/pro^2) B^T accumulates on the matrix cores (FP64 MFMA)."""
    require_gpu()
    _say(f"case: grid_type 2 LISA {solver} {basis_name}, chain 3 atoms seed 5, 30x26")
    w = workloads.chain(natom=3, seed=5, nrad=30, nang=26)
    part, oracle = _lisa_pair(w, basis_name, solver, maxiter=60, grid_type=2)
    cache = _run_pair(part, oracle)
    if cache is not None:
        _compare(part, cache, oracle, qtol=1e-10, ptol=1e-8)
        for a in range(3):
            np.testing.assert_allclose(part[f"at_weights_{a}"], oracle.cache[f"at_weights_{a}"], rtol=1e-11, atol=1e-14)


def test_gt2_gisa():
    require_gpu()
    _say("case: grid_type 2 GISA, chain 3 atoms seed 11, 30x26")
    w = workloads.chain(natom=3, seed=11, nrad=30, nang=26)
    part = GaussianISAWPart(w.coordinates, w.numbers, w.pseudo_numbers, w.grid, w.density, maxiter=60, grid_type=2)
    bs = part.bs_helper
    basis = OracleBasis(bs.orders, bs.exponents, bs.initials)
    oracle = OracleGISA(w.coordinates, w.numbers, w.pseudo_numbers, w.grid, w.density, maxiter=60, basis=basis, grid_type=2)
    cache = _run_pair(part, oracle)
    if cache is not None:
        _compare(part, cache, oracle, check_inner=False)


def test_gt2_mbis():
    require_gpu()
    _say("case: grid_type 2 MBIS, chain 3 atoms seed 4, 30x26")
    w = workloads.chain(natom=3, seed=4, nrad=30, nang=26)
    kw = dict(maxiter=25, grid_type=2)
    part = MBISWPart(w.coordinates, w.numbers, w.pseudo_numbers, w.grid, w.density, **kw)
    oracle = OracleMBIS(w.coordinates, w.numbers, w.pseudo_numbers, w.grid, w.density, **kw)
    cache = _run_pair(part, oracle)
    if cache is not None:
        _compare(part, cache, oracle, qtol=1e-10, ptol=1e-8)


def test_gt2_lisa_cluster_multichunk():
    """grid_type 2 Newton on a cluster whose molecular grid spans many point
    chunks, with far-field pro-atom values near zero: the smallest pro-atom
    value (negativity check, alisa.py:889-897) is a minimum over all chunks."""
    require_gpu()
    _say("case: grid_type 2 LISA m-newton slater6, cluster 12 atoms seed 5, 30x50 (18,000 points)")
    w = workloads.cluster(natom=12, seed=5, nrad=30, nang=50)
    part, oracle = _lisa_pair(w, "slater6", "m-newton", maxiter=30, grid_type=2)
    cache = _run_pair(part, oracle)
    if cache is not None:
        _compare(part, cache, oracle, qtol=1e-10, ptol=1e-8)


def test_gt2_lisa_mixed_mfma_fp32():
    """Mixed path on grid_type 2: FP32 pro-atoms and FP32 MFMA Hessian, FP64
    solves, then FP64 polish; charges within 1e-6 e of the FP64 oracle."""
    require_gpu()
    _say("case: grid_type 2 LISA m-newton gauss mixed, chain 3 atoms seed 5, 30x26")
    w = workloads.chain(natom=3, seed=5, nrad=30, nang=26)
    part, oracle = _lisa_pair(w, "gauss", "m-newton", maxiter=60, grid_type=2, precision="mixed")
    part.do_partitioning()
    _, oracle = _lisa_pair(w, "gauss", "m-newton", maxiter=60, grid_type=2)
    cache = oracle.do_partitioning()
    dq = np.max(np.abs(part["charges"] - cache["charges"]))
    _say(f"  mixed niter={part['niter']} oracle niter={cache['niter']} max|dq|={dq:.2e}")
    assert dq <= 1e-6


# --- bounded-tail skipping of the weights pass (full_sum=False, the default) ---


def _evals_done_frac(part):
    wk = part._dev.work()
    done = wk["evals_done_fp64"] + wk["evals_done_fp32"]
    return done / (wk["evals_per_pass"] * part["niter"])


@pytest.mark.parametrize("scheme", ["mbis", "gisa"])
def test_tail_skip_against_oracle(scheme):
    """A 40-atom chain, where the weights pass skips a large share of the
    atom/shell pairs, against the oracle (which sums every atom everywhere)."""
    require_gpu()
    seed = 21 if scheme == "mbis" else 23
    _say(f"case: {scheme} tail skipping, chain 40 atoms seed {seed}, 30x26")
    w = workloads.chain(natom=40, seed=seed, nrad=30, nang=26, device=0)
    args = (w.coordinates, w.numbers, w.pseudo_numbers, w.grid, w.density)
    if scheme == "mbis":
        part = MBISWPart(*args, maxiter=12)
        oracle = OracleMBIS(*args, maxiter=12)
    else:
        part = GaussianISAWPart(*args, maxiter=10)
        bs = part.bs_helper
        oracle = OracleGISA(*args, maxiter=10, basis=OracleBasis(bs.orders, bs.exponents, bs.initials))
    part.do_partitioning()
    frac = _evals_done_frac(part)
    _say(f"  function evaluations performed / algorithmic: {frac:.3f}")
    assert frac < 0.9
    cache = oracle.do_partitioning()
    _compare(part, cache, oracle, check_inner=scheme == "mbis")


@pytest.mark.parametrize("scheme,natom,seed", [("mbis", 120, 5), ("gisa", 80, 9)])
def test_tail_skip_equals_full_sum(scheme, natom, seed):
    """Skipping the bounded tails changes nothing beyond rounding: the same
    iteration count and charges within 1e-12 e of the full sum."""
    require_gpu()
    _say(f"case: {scheme} tail skipping vs full sum, chain {natom} atoms seed {seed}, 40x50")
    w = workloads.chain(natom=natom, seed=seed, nrad=40, nang=50, device=0)
    args = (w.coordinates, w.numbers, w.pseudo_numbers, w.grid, w.density)
    cls = MBISWPart if scheme == "mbis" else GaussianISAWPart
    res = {}
    for full in (True, False):
        part = cls(*args, maxiter=25, full_sum=full)
        part.do_partitioning()
        res[full] = (part["niter"], part["charges"].copy(), part["propars"].copy(), _evals_done_frac(part))
        part.close()
    dq = np.max(np.abs(res[True][1] - res[False][1]))
    rel = np.max(np.abs(res[True][2] - res[False][2]) / np.maximum(np.abs(res[True][2]), 1e-300))
    _say(f"  niter {res[True][0]} / {res[False][0]}; max|dq| {dq:.2e}; rel dpar {rel:.2e}; "
         f"evaluations performed: full {res[True][3]:.3f}, skipping {res[False][3]:.3f}")
    assert res[True][3] == pytest.approx(1.0)
    assert res[False][3] < 0.9
    assert res[True][0] == res[False][0]
    assert dq <= 1e-12
    assert rel <= 1e-10


def test_tail_skip_with_cutoff_radius():
    """Tail skipping on top of the cut-off lists (a 60-atom cluster, 8-bohr
    cut-off): against the oracle with the same cut-off, and against the full
    sum of the candidates."""
    require_gpu()
    _say("case: MBIS 8-bohr cut-off with tail skipping, cluster 60 atoms seed 13, 30x26")
    w = workloads.cluster(natom=60, seed=13, nrad=30, nang=26, device=0)
    args = (w.coordinates, w.numbers, w.pseudo_numbers, w.grid, w.density)
    kw = dict(maxiter=10, cutoff_radius=8.0)
    part = MBISWPart(*args, **kw)
    part.do_partitioning()
    frac = _evals_done_frac(part)
    full = MBISWPart(*args, full_sum=True, **kw)
    full.do_partitioning()
    frac_full = _evals_done_frac(full)
    dq_full = np.max(np.abs(part["charges"] - full["charges"]))
    _say(f"  evaluations performed / algorithmic: skipping {frac:.3f}, full {frac_full:.3f}; max|dq| vs full {dq_full:.2e}")
    assert frac < frac_full
    assert part["niter"] == full["niter"]
    assert dq_full <= 1e-12
    oracle = OracleMBIS(*args, **kw)
    cache = oracle.do_partitioning()
    _compare(part, cache, oracle)
    full.close()


# --- NLIS / GMBIS (nlis.py, gmbis.py) ----------------------------------------------


@pytest.mark.parametrize("scheme", ["nlis", "gmbis"])
def test_nlis_gmbis_chain(scheme):
    """NLIS (4 shells per element, some with n = 2) and GMBIS (MBIS shells,
    n = 1) on a chain against the oracle: shells N g(n, S) exp(-S r^n) in the
    weights pass, the NLIS fixed point in k_nlis_inner."""
    require_gpu()
    from horton_part_amd import GMBISWPart, NLISWPart
    from oracle.horton_oracle import OracleGMBIS, OracleNLIS

    _say(f"case: {scheme} chain 5 atoms seed 17, 50x110")
    w = workloads.chain(natom=5, seed=17, nrad=50, nang=110)
    args = (w.coordinates, w.numbers, w.pseudo_numbers, w.grid, w.density)
    if scheme == "nlis":
        nd = {z: 3 for z in (1, 6, 7, 8)}
        exp_n = {(z, 0): 2.0 for z in (6, 7, 8)}
        part = NLISWPart(*args, maxiter=60, exp_n_dict=exp_n, nshell_dict=nd)
        oracle = OracleNLIS(*args, maxiter=60, exp_n_dict=exp_n, nshell_dict=nd)
    else:
        part = GMBISWPart(*args, maxiter=60, exp_n_dict={})
        oracle = OracleGMBIS(*args, maxiter=60, exp_n_dict={})
    part.do_partitioning()
    cache = oracle.do_partitioning()
    _compare(part, cache, oracle, qtol=1e-10, ptol=1e-8)
    np.testing.assert_allclose(part["valence_charges"], cache["valence_charges"], atol=1e-10)
    np.testing.assert_allclose(part["core_charges"], cache["core_charges"], atol=1e-10)


# --- gLISA (glisa.py): the global problem ---------------------------------------------


def _glisa_pair(w, basis_name, solver, **kw):
    from horton_part_amd import GlobalLinearISAWPart
    from oracle.horton_oracle import OracleGLISA

    helper = _slater6(w.numbers) if basis_name == "slater6" else ExpBasisFuncHelper.from_function_type("gauss")
    part = GlobalLinearISAWPart(w.coordinates, w.numbers, w.pseudo_numbers, w.grid, w.density, solver=solver,
                                basis_func=helper, **kw)
    okw = {k: v for k, v in kw.items() if k in ("threshold", "maxiter", "grid_type")}
    oracle = OracleGLISA(w.coordinates, w.numbers, w.pseudo_numbers, w.grid, w.density,
                         basis=OracleBasis(helper.orders, helper.exponents, helper.initials), solver=solver, **okw)
    return part, oracle


@pytest.mark.parametrize("natom,nrad,nang", [(5, 40, 50), (22, 16, 26), (24, 16, 26)])
def test_glisa_global_terms_against_oracle(natom, nrad, nang):
    """The device objective, gradient and MFMA Hessian of the global problem
    (glisa.py:411-479) against the oracle's at the initial coefficients and at
    perturbed ones.  M = 22 and 112 take the Hessian kernel that forms every
    tile pair of a point chunk in one workgroup (3 and 28 pairs); M = 120
    (padded to 128) the one with a wavefront per tile pair."""
    require_gpu()
    _say(f"case: gLISA global terms, chain {natom} atoms seed 11, {nrad}x{nang} (gauss basis)")
    w = workloads.chain(natom=natom, seed=11, nrad=nrad, nang=nang)
    part, oracle = _glisa_pair(w, "gauss", "sc")
    c = part._init_propars()
    oc = oracle.init_propars()
    np.testing.assert_allclose(c, oc, rtol=1e-14)
    dev = part._ensure_device()
    rng = np.random.default_rng(5)
    for trial in range(2):
        x = c if trial == 0 else c * (1 + 0.3 * rng.uniform(-1, 1, size=c.size))
        f, g, H, mn = dev.global_terms(x, nderiv=2)
        pro = oracle.calc_promol_dens(x)
        fo, go, Ho = oracle.working_matrix(oracle.moldens, pro, 2)
        _say(f"  trial {trial}: |df|={abs(f - fo):.2e} max rel dgrad={np.max(np.abs(g - go) / np.abs(go)):.2e} "
             f"max rel dH={np.max(np.abs(H - Ho) / np.maximum(np.abs(Ho), 1e-300)):.2e}")
        assert f == pytest.approx(fo, rel=1e-12)
        np.testing.assert_allclose(g, go, rtol=1e-12)
        np.testing.assert_allclose(H, Ho, rtol=1e-11, atol=1e-13 * np.abs(Ho).max())
        assert mn == pytest.approx(pro.min(), rel=1e-13)


@pytest.mark.parametrize("grid_type", [1, 2, 3])
def test_glisa_sc(grid_type):
    """gLISA with the self-consistent solver (the reference's tested one,
    tests/scripts/test_partition_density.py:156-199) against the oracle."""
    require_gpu()
    _say(f"case: gLISA sc grid_type {grid_type}, chain 3 atoms seed 5, 30x26")
    w = workloads.chain(natom=3, seed=5, nrad=30, nang=26)
    part, oracle = _glisa_pair(w, "slater6", "sc", grid_type=grid_type)
    part.do_partitioning()
    cache = oracle.do_partitioning()
    dq = np.max(np.abs(part["charges"] - cache["charges"]))
    rel = np.max(np.abs(part["propars"] - cache["propars"]) / np.abs(cache["propars"]))
    _say(f"  niter gpu={part['niter']} oracle={cache['niter']} max|dq|={dq:.2e} rel dpar={rel:.2e}")
    assert part["niter"] == cache["niter"]
    assert dq <= 1e-10
    assert rel <= 1e-8
    np.testing.assert_allclose(part["history_changes"], cache["history_changes"], rtol=1e-6, atol=1e-14)
    np.testing.assert_allclose(part["history_entropies"], cache["history_entropies"], rtol=1e-10)
    for a in range(3):
        np.testing.assert_allclose(part[f"at_weights_{a}"], oracle.cache[f"at_weights_{a}"], rtol=1e-10, atol=1e-14)


def test_glisa_newton_failures_like_reference():
    """m-newton's line search finds no valid step on these inputs, and the
    device path raises the reference's RuntimeError where the oracle does."""
    require_gpu()
    _say("case: gLISA m-newton on the reference water fixture (expects RuntimeError: Line search failed!)")
    coords, nums, pseudo, grid, dens = load_water_fixture()
    from horton_part_amd import GlobalLinearISAWPart
    from oracle.horton_oracle import OracleGLISA

    h = ExpBasisFuncHelper.from_function_type("gauss")
    oracle = OracleGLISA(coords, nums, pseudo, grid, dens, basis=OracleBasis(h.orders, h.exponents, h.initials),
                         solver="m-newton")
    with pytest.raises(RuntimeError, match="Line search failed"):
        oracle.do_partitioning()
    part = GlobalLinearISAWPart(coords, nums, pseudo, grid, dens, solver="m-newton", basis_func=h)
    with pytest.raises(RuntimeError, match="Line search failed"):
        part.do_partitioning()
    assert len(part.history_changes) == len(oracle.history_changes)


@pytest.mark.parametrize("natom", [3, 70])
def test_slater_density_device_matches_host(natom):
    """The device Slater-shell density (hp_slater_density) against the plain
    sum over every atom in numpy.  On the 70-atom chain the workgroups skip
    atoms bounded below 2^-64 of the density over their points' box."""
    require_gpu()
    from horton_part_amd import _native

    _say(f"case: Slater density, chain {natom} atoms seed 9, 20x26 grids")
    w = workloads.chain(natom=natom, seed=9, nrad=20, nang=26, becke=False)
    rng = np.random.default_rng(4)
    nf = rng.integers(1, 4, size=natom)
    off = np.concatenate([[0], np.cumsum(nf)]).astype(np.int32)
    coef = rng.uniform(0.1, 50.0, size=off[-1])
    alpha = rng.uniform(0.8, 8.0, size=off[-1])
    pts = w.grid.points
    dev = _native.slater_density(w.coordinates, off, coef, alpha, pts)
    host = np.zeros(len(pts))
    for a in range(natom):
        r = np.linalg.norm(pts - w.coordinates[a], axis=1)
        for f in range(off[a], off[a + 1]):
            host += coef[f] * np.exp(-alpha[f] * r)
    np.testing.assert_allclose(dev, host, rtol=1e-13, atol=0)


@pytest.mark.parametrize("shape", ["cluster37", "chain60"])
def test_becke_weights_device_matches_host(shape):
    """The device Becke weights against the numpy restatement (every cell
    product in full).  The device visits the atoms nearest-first from each
    point's owner and stops a product below 2^-64 of the running sum: on the
    60-atom chain most products stop after a few factors."""
    require_gpu()
    from horton_part_amd.grid import BeckeWeights

    if shape == "cluster37":
        _say("case: Becke weights, 37-atom cluster seed 5, 20x26 grids")
        w = workloads.cluster(natom=37, seed=5, nrad=20, nang=26, becke=False)
    else:
        _say("case: Becke weights, 60-atom chain seed 9, 20x26 grids")
        w = workloads.chain(natom=60, seed=9, nrad=20, nang=26, becke=False)
    g = w.grid
    host = BeckeWeights().generate_weights(g.points, w.coordinates, w.numbers, indices=g.indices)
    dev = BeckeWeights(device=0).generate_weights(g.points, w.coordinates, w.numbers, indices=g.indices)
    np.testing.assert_allclose(dev, host, rtol=1e-12, atol=1e-15)


def _wide_shell_water(nrad=12):
    """Water on atom grids whose radial shells hold 604 points each: the
    302-point Lebedev rule and the same rule rotated, each at half weight (an
    angular rule as exact as either).  Shells of more than 512 points take a
    weights workgroup each, two points per thread."""
    from horton_part_amd.grid import AtomGrid, BeckeWeights, MolGrid, exp_radial_grid
    from horton_part_amd.lebedev import lebedev_rule

    geo = workloads.water(seed=1000, nrad=4, nang=6)  # the geometry
    coords, numbers = geo.coordinates, geo.numbers
    rg = exp_radial_grid(nrad)
    u, wl = lebedev_rule(302)
    c1, s1, c2, s2 = np.cos(0.3), np.sin(0.3), np.cos(0.7), np.sin(0.7)
    rot = np.array([[c1, -s1, 0.0], [s1, c1, 0.0], [0.0, 0.0, 1.0]]) @ np.array(
        [[c2, 0.0, s2], [0.0, 1.0, 0.0], [-s2, 0.0, c2]]
    )
    u2 = np.concatenate([u, u @ rot.T])
    w2 = np.concatenate([wl, wl]) / 2.0
    r, wr = rg.points, rg.weights
    atgrids = []
    for c in coords:
        g = AtomGrid(rg, sizes=[302], center=c)
        g._points = (c + r[:, None, None] * u2[None, :, :]).reshape(-1, 3)
        g._weights = ((wr * r**2 * 4.0 * np.pi)[:, None] * w2[None, :]).ravel()
        g._sph_weights = np.tile(w2, nrad)
        g._sizes = np.full(nrad, len(w2))
        g._indices = np.concatenate([[0], np.cumsum(g._sizes)])
        atgrids.append(g)
    grid = MolGrid(numbers, atgrids, BeckeWeights())
    rng = np.random.default_rng(7)
    shells = workloads.density_shells(rng, numbers, lambda z: 2 if z == 8 else 1)
    rho = workloads.eval_density(grid.points, coords, shells, rng)
    return coords, numbers, geo.pseudo_numbers, grid, rho


@pytest.mark.parametrize("scheme", ["mbis", "gisa"])
def test_shells_above_512_points(scheme):
    """Radial shells of 604 points against the oracle (FP64)."""
    require_gpu()
    _say(f"case: {scheme} water, 12 radial shells of 604 points per atom (302-point rule and a rotated copy)")
    args = _wide_shell_water()
    assert int(np.diff(args[3].atgrids[0].indices).max()) > 512
    if scheme == "mbis":
        part = MBISWPart(*args, maxiter=60)
        part.do_partitioning()
        oracle = OracleMBIS(*args, maxiter=60)
        cache = oracle.do_partitioning()
        _compare(part, cache, oracle)
    else:
        part = GaussianISAWPart(*args, maxiter=60)
        part.do_partitioning()
        bs = part.bs_helper
        oracle = OracleGISA(*args, maxiter=60, basis=OracleBasis(bs.orders, bs.exponents, bs.initials))
        cache = oracle.do_partitioning()
        _compare(part, cache, oracle, check_inner=False)


def _heavy_chain(nrad):
    """K, O, Rb, Cs, Fr in a row (4, 2, 5, 6 and 7 MBIS shells)."""
    rng = np.random.default_rng(21)
    numbers = np.array([19, 8, 37, 55, 87], dtype=np.int64)
    coords = np.array([[0.0, 0.0, 0.0], [2.6, 0.0, 0.0], [5.8, 0.5, 0.0], [9.5, 0.0, 0.4], [13.4, 0.3, 0.0]])
    return workloads._build("heavy5", coords, numbers, nrad, 26, rng, lambda z: 2)


@pytest.mark.parametrize("nrad", [40, 150])
def test_mbis_four_to_seven_shells(nrad):
    """Atoms of 4-7 MBIS shells: 9-15 sums per inner iteration, four to a
    register in up to four groups, with two (40 radial points) and three
    (150) points per lane; FP64 against the oracle."""
    require_gpu()
    _say(f"case: MBIS K-O-Rb-Cs-Fr chain, {nrad}x26, maxiter 12")
    w = _heavy_chain(nrad)
    args = (w.coordinates, w.numbers, w.pseudo_numbers, w.grid, w.density)
    part = MBISWPart(*args, maxiter=12)
    part.do_partitioning()
    oracle = OracleMBIS(*args, maxiter=12)
    cache = oracle.do_partitioning()
    # several inner solves run to their 2000-iteration cap; the bar is the
    # plain FP64 one (measured: |dq| 3e-14, parameters 2e-10 relative)
    _compare(part, cache, oracle)
