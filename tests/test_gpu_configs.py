"""GPU parity at the benchmark configurations' workloads.

The oracle's results for each case of ``tests/configs.py`` are stored under
``tests/golden/configs/`` (written by ``tests/golden/make_config_fixtures.py``);
these tests rebuild the same seeded inputs on the host, check that they are
the inputs the oracle saw (digest), run the device path and compare:

* plain FP64: identical outer iteration counts and per-atom inner iteration
  counts of the last iteration, charges and populations within 1e-10 e,
  parameters within 1e-9 relative, change and entropy histories;
* mixed precision (FP32 bulk, FP64 sums, solves and polish): max |dq| <= 1e-6 e
  against the FP64 oracle;
* the serving-stream requests through the continuous-batching engine.

The full-size systems (the 200-atom aLISA cluster and the 2000-atom protein
aggregate) are beyond the oracle; they are run at full size and checked by
properties that do not depend on size: the populations add up to the
integrated density, the weights lie in [0, 1], the iteration ends as the
oracle-sized sub-systems show it does.
"""

import os
import sys
import warnings

import numpy as np
import pytest

import configs
from conftest import require_gpu

pytestmark = pytest.mark.gpu
warnings.simplefilter("ignore")

FIXTURES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "configs")
CASES = configs.all_cases()
HAVE = sorted(n for n in CASES if os.path.exists(os.path.join(FIXTURES, n + ".npz")))
SINGLE = [n for n in HAVE if not n.startswith("c4_")]
STREAM = [n for n in HAVE if n.startswith("c4_")]


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _fixture(name):
    return np.load(os.path.join(FIXTURES, name + ".npz"))


def _check_inputs(case, fx):
    np.testing.assert_allclose(case.input_digest(), fx["digest"], rtol=1e-12, atol=0)


def _compare(name, propars, charges, niter, fx, inner=None, qtol=1e-10, ptol=1e-9, hist=None, erel=1e-9):
    dq = float(np.max(np.abs(charges - fx["charges"])))
    rel = float(np.max(np.abs(propars - fx["propars"]) / np.maximum(np.abs(fx["propars"]), 1e-300)))
    _say(f"  {name}: niter gpu={niter} oracle={int(fx['niter'])} max|dq|={dq:.2e} rel dpar={rel:.2e}")
    assert niter == int(fx["niter"])
    assert dq <= qtol
    assert rel <= ptol
    if inner is not None:
        assert list(inner) == list(fx["inner_niter"])
    if hist is not None:
        changes, entropies = hist
        np.testing.assert_allclose(changes, fx["history_changes"], rtol=1e-6, atol=1e-14)
        np.testing.assert_allclose(entropies, fx["history_entropies"], rtol=erel, atol=1e-12)


# Cases whose problem does not fix its solution to 1e-10: the 24-atom aLISA
# cluster's m-newton steps solve Hessians of condition numbers up to ~1e12,
# and the oracle itself, given the density perturbed by 1e-15 relative
# (three draws), moves its charges by up to 1.4e-8 e and its parameters by
# up to 1.3e-3 relative with the same iteration counts.  Its charges are held
# to 5e-8 e (about 3.5x that spread), its entropies to 1e-8 relative (the same
# order as the charges' spread) and its parameters are not compared.
TOLERANCE = {"c2_cluster24_lisa": dict(qtol=5e-8, ptol=np.inf, erel=1e-8)}


@pytest.mark.parametrize("name", SINGLE)
def test_config_fp64(name):
    """The plain FP64 device path against the oracle at the configuration's
    workload (or its oracle-sized sub-system)."""
    require_gpu()
    case = CASES[name]
    _say(f"case: {name} FP64 ({case.cls.__name__}, {case.kw})")
    fx = _fixture(name)
    if "raised" in fx:
        pytest.fail(f"oracle raised {fx['raised']}")
    _check_inputs(case, fx)
    part = case.part()
    part.do_partitioning()
    tol = TOLERANCE.get(name, {})
    _compare(
        name,
        part["propars"],
        part["charges"],
        part["niter"],
        fx,
        inner=part.inner_niter,
        hist=(part["history_changes"], part["history_entropies"]),
        **tol,
    )
    np.testing.assert_allclose(part["populations"], case.workload.numbers - fx["charges"], rtol=0, atol=tol.get("qtol", 1e-10))
    part.close()


MIXED = [n for n in ("c1_chain20_mbis", "c1_chain20_gisa", "c3_protein64_mbis") if n in HAVE]


@pytest.mark.parametrize("name", MIXED)
def test_config_mixed(name):
    """Mixed precision (FP32 exp, promolecule and ratios; FP64 sums, solves
    and polish) within 1e-6 e of the FP64 oracle."""
    require_gpu()
    case = CASES[name]
    _say(f"case: {name} mixed precision")
    fx = _fixture(name)
    _check_inputs(case, fx)
    part = case.part(precision="mixed")
    part.do_partitioning()
    dq = float(np.max(np.abs(part["charges"] - fx["charges"])))
    _say(f"  mixed niter={part['niter']} oracle niter={int(fx['niter'])} max|dq|={dq:.2e}")
    assert dq <= 1e-6
    part.close()


def test_stream_sample_through_batch():
    """Serving-stream requests (MBIS and GISA, 5-30 atoms) admitted together
    into the continuous-batching engine: FP64 results equal the oracle's,
    and the serving policy (GISA mixed, MBIS FP64) is within 1e-6 e."""
    require_gpu()
    from horton_part_amd import PartitionServer

    _say(f"case: {len(STREAM)} serving-stream requests through the batch: {STREAM}")
    fxs = {n: _fixture(n) for n in STREAM}
    for n in STREAM:
        _check_inputs(CASES[n], fxs[n])
    for policy in ("fp64", "auto"):
        with PartitionServer(device=0, point_capacity=4_000_000, atom_capacity=512, shell_capacity=16384,
                             function_capacity=2048, qp_capacity=65536, max_inflight=16,
                             max_molecule_points=1_000_000, ticks_per_step=4) as srv:
            tickets = {}
            for n in STREAM:
                tickets[srv.submit(CASES[n].part(), precision=policy)] = n
            done = {}
            while srv.pending:
                srv.step()
                for d in srv.poll():
                    done[tickets[d.ticket]] = d
        for n in STREAM:
            d, fx = done[n], fxs[n]
            if policy == "fp64":
                _compare(n, d.propars, d.charges, d.niter, fx)
            else:
                dq = float(np.max(np.abs(d.charges - fx["charges"])))
                _say(f"  {n} ({policy}): niter {d.niter} / {int(fx['niter'])} max|dq|={dq:.2e}")
                assert dq <= 1e-6


def test_cluster200_full_size_properties():
    """The 200-atom aLISA cluster (6 Slater functions, m-newton, molecular
    grid, 3.88M points) at full size: on the molecular grid the AIM weights
    of every point add up to one, so the populations add up to the
    integrated density; the iteration ends as on the 24-atom sub-cluster
    (where the oracle agrees with the device): with a change of exactly 0
    once every atom's line search keeps only its zero step."""
    require_gpu()
    from horton_part_amd import ExpBasisFuncHelper, LinearISAWPart, workloads

    _say("case: cluster200 aLISA m-newton grid_type 2, full size (device-built inputs)")
    w = workloads.cluster(natom=200, seed=1002, nrad=100, nang=194, device=0)
    basis = ExpBasisFuncHelper.uniform(w.numbers, configs.SLATER6, order=1.0)
    part = LinearISAWPart(w.coordinates, w.numbers, w.pseudo_numbers, w.grid, w.density, threshold=1e-6, maxiter=200,
                          solver="m-newton", basis_func=basis, grid_type=2)
    part.do_partitioning()
    nelec = w.grid.integrate(w.density)
    pop = part["populations"].sum()
    hc = part["history_changes"]
    _say(f"  niter {part['niter']} last changes {hc[-3:]}; sum populations {pop:.10f} vs integral {nelec:.10f}")
    assert np.isfinite(part["charges"]).all()
    assert abs(pop - nelec) <= 1e-8 * nelec
    for a in (0, 99, 199):
        wa = part[f"at_weights_{a}"]
        assert wa.min() >= 0.0 and wa.max() <= 1.0
    assert part["niter"] < 200 and part["change"] < 1e-6
    part.close()


def test_protein2000_full_size_properties():
    """The 2000-atom protein-like aggregate (MBIS, 12-bohr cut-off, 16.5M
    points) at full size: charges finite and adding up to the noise-level
    difference between the nuclear charge and the density; the change
    history is printed (the 64- and 128-atom sub-systems converge on both
    the oracle and the device)."""
    require_gpu()
    from horton_part_amd import MBISWPart, workloads

    _say("case: protein2000 MBIS 12-bohr cut-off, full size (device-built inputs)")
    w = workloads.cluster(natom=2000, seed=1003, nrad=75, nang=110, probs=configs.PROTEIN_PROBS, becke=False, device=0)
    part = MBISWPart(w.coordinates, w.numbers, w.pseudo_numbers, w.grid, w.density, threshold=1e-6, maxiter=1000,
                     cutoff_radius=12.0)
    part.do_partitioning()
    hc = part["history_changes"]
    q = part["charges"]
    ztot = float(w.numbers.sum())
    _say(f"  niter {part['niter']} change {part['change']:.3e}; min change {hc.min():.3e} at {int(hc.argmin()) + 1}; "
         f"changes at 100/200/400/800: {[float(hc[i]) for i in (99, 199, 399, 799) if i < len(hc)]}; "
         f"sum q {q.sum():.4e} of Z {ztot:.0f}")
    assert np.isfinite(q).all()
    assert abs(q.sum()) <= 1e-3 * ztot
    part.close()


# --- the solver plugin points ---------------------------------------------------


def _gisa_qp_callable(calls):
    from oracle.horton_oracle import opt_propars_qp

    def solver(bs_funcs, rho, propars, points, weights, alphas, threshold, **opts):
        calls.append(len(rho))
        return opt_propars_qp(bs_funcs, rho, weights, alphas)

    return solver


def test_gisa_solver_callable():
    """GaussianISAWPart with a solver callable (gisa.py:320-331): the device
    runs each weights pass, the callable each atom's update.  With the
    oracle's QP as the callable the result equals the oracle's GISA."""
    require_gpu()
    from horton_part_amd import GaussianISAWPart, workloads
    from oracle.horton_oracle import OracleBasis, OracleGISA

    _say("case: GISA with a solver callable, chain 5 atoms seed 11, 50x110")
    w = workloads.chain(natom=5, seed=11, nrad=50, nang=110)
    calls = []
    part = GaussianISAWPart(w.coordinates, w.numbers, w.pseudo_numbers, w.grid, w.density, maxiter=60,
                            solver=_gisa_qp_callable(calls))
    part.do_partitioning()
    bs = part.bs_helper
    oracle = OracleGISA(w.coordinates, w.numbers, w.pseudo_numbers, w.grid, w.density, maxiter=60,
                        basis=OracleBasis(bs.orders, bs.exponents, bs.initials))
    cache = oracle.do_partitioning()
    assert len(calls) == part["niter"] * 5 and set(calls) == {50}
    _compare("gisa callable", part["propars"], part["charges"], part["niter"], dict(cache, inner_niter=None),
             hist=(part["history_changes"], part["history_entropies"]))
    for a in range(5):
        np.testing.assert_allclose(part[f"spherical_average_{a}"], oracle.cache[f"spherical_average_{a}"], rtol=1e-11, atol=1e-14)
        np.testing.assert_allclose(part[f"at_weights_{a}"], oracle.cache[f"at_weights_{a}"], rtol=1e-12, atol=1e-15)


@pytest.mark.parametrize("grid_type", [1, 2])
def test_lisa_solver_callable(grid_type):
    """LinearISAWPart with a solver callable (alisa.py:1323-1335) on atom
    grids and on the molecular grid: the oracle's self-consistent solver as
    the callable gives the oracle's sc result."""
    require_gpu()
    from horton_part_amd import LinearISAWPart, workloads
    from horton_part_amd.basis import ExpBasisFuncHelper
    from oracle.horton_oracle import OracleBasis, OracleLISA, solver_sc

    _say(f"case: aLISA with a solver callable, grid_type {grid_type}")
    w = workloads.chain(natom=3, seed=5, nrad=30, nang=26)
    helper = ExpBasisFuncHelper.from_function_type("gauss")
    seen = {}

    def sc(bs_funcs, rho, propars, points, weights, threshold, logger, density_cutoff, negative_cutoff, population_cutoff):
        seen["shape"] = bs_funcs.shape
        return solver_sc(bs_funcs, rho, propars, weights, threshold, density_cutoff)

    part = LinearISAWPart(w.coordinates, w.numbers, w.pseudo_numbers, w.grid, w.density, maxiter=40, solver=sc,
                          basis_func=helper, grid_type=grid_type)
    part.do_partitioning()
    oracle = OracleLISA(w.coordinates, w.numbers, w.pseudo_numbers, w.grid, w.density, maxiter=40, solver="sc",
                        basis=OracleBasis(helper.orders, helper.exponents, helper.initials), grid_type=grid_type)
    cache = oracle.do_partitioning()
    assert seen["shape"][1] == (w.grid.size if grid_type == 2 else 30)
    _compare("lisa callable", part["propars"], part["charges"], part["niter"], dict(cache, inner_niter=None), ptol=1e-8,
             hist=(part["history_changes"], part["history_entropies"]))


# --- grid_type 3 -----------------------------------------------------------------


class _MolGridOnly:
    """A molecular grid without atom grids (points, weights, integrate)."""

    atgrids = None

    def __init__(self, grid):
        self.points = grid.points
        self.weights = grid.weights
        self.size = grid.size

    def integrate(self, *values):
        out = self.weights.copy()
        for v in values:
            out = out * v
        return out.sum()


@pytest.mark.parametrize("with_atgrids", [True, False])
def test_gt3_mbis(with_atgrids):
    """grid_type 3 (core/base.py:490-499): the molecular-grid loop, and the
    moments of each AIM density over the whole molecular grid (local =
    False).  Without atom grids the device splits the points into blocks."""
    require_gpu()
    from horton_part_amd import MBISWPart, workloads
    from oracle.horton_oracle import OracleMBIS

    _say(f"case: MBIS grid_type 3 ({'with' if with_atgrids else 'without'} atom grids), chain 3 atoms seed 4, 30x26")
    w = workloads.chain(natom=3, seed=4, nrad=30, nang=26)
    grid = w.grid if with_atgrids else _MolGridOnly(w.grid)
    kw = dict(maxiter=25, grid_type=3)
    part = MBISWPart(w.coordinates, w.numbers, w.pseudo_numbers, grid, w.density, **kw)
    part.do_all()
    oracle = OracleMBIS(w.coordinates, w.numbers, w.pseudo_numbers, grid, w.density, **kw)
    cache = oracle.do_partitioning()
    inner = [s.get("niter", 1) for s in oracle.inner_stats[-1]]
    _compare("gt3", part["propars"], part["charges"], part["niter"], dict(cache, inner_niter=np.array(inner)),
             inner=part.inner_niter, ptol=1e-8, hist=(part["history_changes"], part["history_entropies"]))
    cart, pure, rad = oracle.do_moments()
    np.testing.assert_allclose(part["cartesian_multipoles"], cart, rtol=1e-9, atol=1e-11)
    np.testing.assert_allclose(part["pure_multipoles"], pure, rtol=1e-9, atol=1e-11)
    np.testing.assert_allclose(part["radial_moments"], rad, rtol=1e-9, atol=1e-11)
