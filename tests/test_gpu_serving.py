"""Continuous batching (hp_batch_*) on the GPU.

* A request run alone in a batch gives the single-molecule path's result
  (hp_part_run, same precision): same outer iteration count, charges and
  parameters equal to 1e-12 (the kernels see the same data in the same
  order, so the results are expected to be bit-identical).
* Each result is the same alone and in a mix of MBIS/GISA, FP64/mixed
  requests of different sizes that arrive while others are on the GPU and
  wait for slots (to 1e-12 e, with graph replay on and off).
* FP64 batch results equal the CPU oracle to 1e-10 e; mixed results are
  within 1e-6 e of it.
"""

import sys
import warnings

import numpy as np
import pytest

from conftest import require_gpu
from horton_part_amd import GaussianISAWPart, MBISWPart, PartitionServer, _native, workloads
from horton_part_amd.alisa import LinearISAWPart
from oracle.horton_oracle import OracleBasis, OracleGISA, OracleMBIS

pytestmark = pytest.mark.gpu

warnings.simplefilter("ignore")

EXACT = 1e-12

SCHEMES = {"mbis": (MBISWPart, OracleMBIS), "gisa": (GaussianISAWPart, OracleGISA)}


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


_W = {}


def _workload(name):
    if name not in _W:
        if name == "water":
            _W[name] = workloads.water(seed=1000, nrad=40, nang=50)
        elif name == "chain5":
            _W[name] = workloads.chain(natom=5, seed=11, nrad=40, nang=50)
        elif name == "chain8":
            _W[name] = workloads.chain(natom=8, seed=12, nrad=36, nang=26)
        elif name == "chain3":
            _W[name] = workloads.chain(natom=3, seed=13, nrad=40, nang=38)
        elif name == "chain4r":
            # more than 128 radial points: the MBIS solver's workgroup layout
            _W[name] = workloads.chain(natom=4, seed=14, nrad=140, nang=26)
        else:
            raise KeyError(name)
    return _W[name]


def _part(scheme, wname, **kw):
    w = _workload(wname)
    cls = SCHEMES[scheme][0]
    kw.setdefault("threshold", 1e-6)
    kw.setdefault("maxiter", 300)
    return cls(w.coordinates, w.numbers, w.pseudo_numbers, w.grid, w.density, **kw)


def _single(scheme, wname, precision, **kw):
    part = _part(scheme, wname, precision=precision, **kw)
    part.do_partitioning()
    out = (part["charges"].copy(), part["propars"].copy(), int(part["niter"]))
    part.close()
    return out


def _server(**kw):
    kw.setdefault("point_capacity", 400_000)
    kw.setdefault("atom_capacity", 200)
    kw.setdefault("shell_capacity", 4000)
    kw.setdefault("function_capacity", 400)
    kw.setdefault("qp_capacity", 20_000)
    kw.setdefault("max_inflight", 16)
    kw.setdefault("ticks_per_step", 3)
    return PartitionServer(**kw)


@pytest.mark.parametrize("scheme", ["mbis", "gisa"])
@pytest.mark.parametrize("precision,refresh", [("fp64", 0), ("mixed", 0), ("mixed", 3)])
def test_batch_alone_equals_single(scheme, precision, refresh):
    """refresh > 0: the FP32 stage with the FP64 defect correction every
    `refresh` iterations (FP64 and calibrating FP32 passes of a molecule in
    one tick)."""
    require_gpu()
    _say(f"case: batch alone vs hp_part_run, {scheme} {precision} refresh {refresh} chain5")
    q1, p1, n1 = _single(scheme, "chain5", precision, mixed_refresh=refresh)
    with _server() as srv:
        srv.submit(_part(scheme, "chain5", mixed_refresh=refresh), precision=precision)
        (done,) = srv.run_until_done(timeout=120)
    dq = np.max(np.abs(done.charges - q1))
    dp = np.max(np.abs(done.propars - p1))
    _say(f"  niter single={n1} batch={done.niter} event={done.event} max|dq|={dq:.1e} max|dp|={dp:.1e}")
    assert done.event == "converged"
    assert done.niter == n1
    assert dq <= EXACT
    assert dp <= EXACT * max(1.0, np.abs(p1).max())


MIX = [
    ("mbis", "water", "mixed"),
    ("gisa", "chain5", "fp64"),
    ("mbis", "chain8", "mixed"),
    ("gisa", "water", "mixed"),
    ("mbis", "chain5", "fp64"),
    ("gisa", "chain8", "mixed"),
    ("mbis", "chain3", "mixed"),
    ("gisa", "chain3", "fp64"),
    ("mbis", "chain4r", "fp64"),
]


def _alone(scheme, wname, precision, cache={}):
    key = (scheme, wname, precision)
    if key not in cache:
        with _server() as srv:
            srv.submit(_part(scheme, wname), precision=precision)
            (done,) = srv.run_until_done(timeout=120)
        cache[key] = done
    return cache[key]


@pytest.mark.parametrize(
    "graphs,groups,budget",
    [(True, 2, 0), (False, 2, 0), (False, 1, 0), (True, 3, 0), (False, 4, 0), (False, 2, 16), (True, 2, 64)],
)
def test_batch_mix_equals_alone(graphs, groups, budget):
    """Requests arrive while others run and wait for one of 3 slots.  With an
    inner budget the MBIS solves run over several ticks: same results."""
    require_gpu()
    _say(f"case: batch mix of {len(MIX)} requests, 3 slots, graphs={graphs}, groups={groups}, inner budget={budget}")
    with _server(max_inflight=3, use_graphs=graphs, groups=groups, inner_budget=budget) as srv:
        tickets = {}
        done = []
        for i, req in enumerate(MIX):
            tickets[srv.submit(_part(req[0], req[1]), precision=req[2])] = req
            # a few steps between arrivals: later requests join a running batch
            for _ in range(2 if i % 2 else 1):
                srv.step()
                done += srv.poll()
        done += srv.run_until_done(timeout=300)
    assert len(done) == len(MIX)
    worst = 0.0
    for d in done:
        ref = _alone(*tickets[d.ticket])
        dq = np.max(np.abs(d.charges - ref.charges))
        dp = np.max(np.abs(d.propars - ref.propars))
        worst = max(worst, dq)
        _say(f"  {tickets[d.ticket]} niter mix={d.niter} alone={ref.niter} max|dq|={dq:.1e} max|dp|={dp:.1e}")
        assert d.niter == ref.niter
        assert d.event == ref.event
        assert dq <= EXACT
        assert dp <= EXACT * max(1.0, np.abs(ref.propars).max())
    _say(f"  worst max|dq| = {worst:.1e}")


@pytest.mark.parametrize("scheme", ["mbis", "gisa"])
def test_batch_against_oracle(scheme):
    """FP64 requests equal the oracle to 1e-10 e; mixed within 1e-6 e."""
    require_gpu()
    _say(f"case: batch vs oracle, {scheme} water 40x50")
    w = _workload("water")
    ocls = SCHEMES[scheme][1]
    kw = {}
    if scheme == "gisa":
        bs = _part("gisa", "water").bs_helper
        kw["basis"] = OracleBasis(bs.orders, bs.exponents, bs.initials)
    oracle = ocls(w.coordinates, w.numbers, w.pseudo_numbers, w.grid, w.density, threshold=1e-6, maxiter=300, **kw)
    cache = oracle.do_partitioning()
    with _server() as srv:
        t64 = srv.submit(_part(scheme, "water"), precision="fp64")
        t32 = srv.submit(_part(scheme, "water"), precision="mixed")
        done = {d.ticket: d for d in srv.run_until_done(timeout=120)}
    d64, d32 = done[t64], done[t32]
    dq64 = np.max(np.abs(d64.charges - cache["charges"]))
    dq32 = np.max(np.abs(d32.charges - cache["charges"]))
    rel = np.max(np.abs(d64.propars - cache["propars"]) / np.maximum(np.abs(cache["propars"]), 1e-300))
    _say(f"  niter oracle={cache['niter']} fp64={d64.niter} mixed={d32.niter} |dq| fp64={dq64:.1e} mixed={dq32:.1e}")
    assert d64.niter == cache["niter"]
    assert dq64 <= 1e-10
    assert rel <= 1e-9
    assert dq32 <= 1e-6
    # the FP32 stage with the defect correction, in the single-molecule path
    part = _part(scheme, "water", precision="mixed", mixed_refresh=4)
    part.do_partitioning()
    dqc = np.max(np.abs(part["charges"] - cache["charges"]))
    _say(f"  mixed with FP64 correction every 4 iterations: niter={part['niter']} |dq|={dqc:.1e}")
    assert dqc <= 1e-6


def test_batch_rejects_what_it_cannot_run():
    require_gpu()
    _say("case: batch argument checks")
    with _server(atom_capacity=6) as srv:
        # LISA is not a batch scheme
        w = _workload("water")
        lisa = LinearISAWPart(w.coordinates, w.numbers, w.pseudo_numbers, w.grid, w.density, threshold=1e-6, solver="sc")
        with pytest.raises(_native.HPError) as e:
            srv.submit(lisa)
        assert e.value.code == _native.HP_E_UNSUPP
        # larger than the arena (6 atom entries, chain8 needs 9)
        with pytest.raises(_native.HPError) as e:
            srv.submit(_part("mbis", "chain8"))
        assert e.value.code == _native.HP_E_ARG
        # inner settings differ from the batch's
        with pytest.raises(_native.HPError) as e:
            srv.submit(_part("mbis", "water", inner_threshold=1e-10))
        assert e.value.code == _native.HP_E_ARG
        assert srv.pending == 0


def test_batch_serves_2000_atoms():
    """The largest request of the admitted range: a 2000-atom protein-like
    aggregate (16 x 26 points per atom, 832,000 points) served alone equals
    the single-molecule path, and served beside small requests equals itself
    alone (4 outer iterations, FP64; the tail bound culls the far atoms)."""
    require_gpu()
    from configs import PROTEIN_PROBS

    _say("case: batch with a 2000-atom request (protein-like, 16x26, maxiter 4), alone, single path and mix")
    w = workloads.cluster(natom=2000, seed=1003, nrad=16, nang=26, probs=PROTEIN_PROBS, becke=False, device=0)

    def big():
        return MBISWPart(w.coordinates, w.numbers, w.pseudo_numbers, w.grid, w.density, threshold=1e-6, maxiter=4)

    single = big()
    single.do_partitioning()
    q1, p1, n1 = single["charges"].copy(), single["propars"].copy(), int(single["niter"])
    single.close()
    caps = dict(point_capacity=1_000_000, atom_capacity=2100, shell_capacity=33_000, function_capacity=6200,
                max_molecule_points=900_000)
    with _server(**caps) as srv:
        srv.submit(big(), precision="fp64")
        (alone,) = srv.run_until_done(timeout=240)
    dq = np.max(np.abs(alone.charges - q1))
    dp = np.max(np.abs(alone.propars - p1))
    _say(f"  niter single={n1} batch={alone.niter} event={alone.event} max|dq|={dq:.1e} max|dp|={dp:.1e}")
    assert alone.niter == n1
    assert dq <= EXACT
    assert dp <= EXACT * max(1.0, np.abs(p1).max())
    with _server(**caps) as srv:
        tb = srv.submit(big(), precision="fp64")
        srv.submit(_part("mbis", "chain5"), precision="fp64")
        srv.step()
        srv.submit(_part("gisa", "water"), precision="mixed")
        done = {d.ticket: d for d in srv.run_until_done(timeout=240)}
    d = done[tb]
    dq = np.max(np.abs(d.charges - alone.charges))
    _say(f"  in the mix: niter={d.niter} max|dq| vs alone={dq:.1e}")
    assert len(done) == 3
    assert d.niter == alone.niter
    assert dq <= EXACT
    assert np.max(np.abs(d.propars - alone.propars)) <= EXACT * max(1.0, np.abs(alone.propars).max())
