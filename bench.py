"""Benchmark of the stockholder partitioning loop on one MI355X.

Default workload (``--workload serving``): the whole serving stream -- 4096
MBIS/GISA requests of synthetic chain molecules with log-uniform atom counts
in [3, 300], 100 radial x 194 Lebedev points per atom, threshold 1e-6,
Poisson arrivals at 200 requests/s over 20.7 s (seed 1004).  A few host
threads build each request's grid and density (on the GPU) in arrival
order, a bounded window ahead; a request is submitted to the
continuous-batching engine once it has arrived and is built, and the engine
admits it as arena space frees up.  ``value`` is completed requests / (last
completion - first arrival) over the whole stream, with p50/p99 latency
(completion - arrival, queueing included); ``--steps`` splits the stream
into that many equal steps (``ms_per_step`` = makespan / steps).  The plain
FP64 GPU path then runs a systematic sample of the stream one request at a
time, which gives the FP64 rate and max |dq| of the served results (the FP64
path equals the CPU oracle to 1e-10 e in the parity tests).

Single-molecule workloads (``--workload chain20|water|cluster200|...``): one
step = one partition with inputs resident in HBM; ``value`` is the mixed
path's partitions/s with the plain FP64 path timed beside it.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload serving]
"""

import argparse
import json
import os
import platform
import sys
import time

os.environ.setdefault("OMP_NUM_THREADS", "1")
os.environ.setdefault("OPENBLAS_NUM_THREADS", "1")
os.environ.setdefault("MKL_NUM_THREADS", "1")
# The batch engine overlaps one stream per molecule group, and HIP gives each
# stream a hardware queue of its own only while the process has queues to
# spare (GPU_MAX_HW_QUEUES, 4 when unset): ask for 8 before HIP starts, which
# lets the engine run three groups (hortonpart.h, hp_batch_config.groups).
if int(os.environ.get("GPU_MAX_HW_QUEUES", "4")) < 8:
    os.environ["GPU_MAX_HW_QUEUES"] = "8"

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (os.path.join(ROOT, "horton-part_amd"), ROOT):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np  # noqa: E402

METRIC = "partitions/s & p99 latency, mixed vs plain FP64 at max|Δq|≤1e-6 e"
HBM_PEAK_GBPS = 8000.0
# HBM bytes the serving tick's weights launches move per algorithmic byte
# (48 B per point): FETCH_SIZE x 2 + WRITE_SIZE over every k_weights launch of
# a 64-request serving run, each counter scaled by what it reads for a plain
# copy of 256 MiB (profiles/k_weights_serving.md)
K_WEIGHTS_TRAFFIC_RATIO = 1.038


def log(msg):
    print(msg, file=sys.stderr, flush=True)


PROTEIN_PROBS = {"C": 0.31, "H": 0.50, "N": 0.09, "O": 0.09, "S": 0.01}


def build_workload(name, device):
    """(workload, scheme class, keyword arguments, description) of a
    single-molecule benchmark configuration."""
    from horton_part_amd import ExpBasisFuncHelper, GaussianISAWPart, LinearISAWPart, MBISWPart, workloads

    if name == "water":
        w = workloads.water(seed=1000, nrad=100, nang=194, device=device)
        return w, MBISWPart, dict(threshold=1e-8, maxiter=1000), "MBIS, threshold 1e-8"
    if name == "chain20":
        w = workloads.chain(natom=20, seed=1001, nrad=150, nang=302, device=device)
        return w, MBISWPart, dict(threshold=1e-6, maxiter=1000), "MBIS, threshold 1e-6"
    if name == "chain20-gisa":
        w = workloads.chain(natom=20, seed=1001, nrad=150, nang=302, device=device)
        return w, GaussianISAWPart, dict(threshold=1e-6, maxiter=1000), "GISA (gauss.json), threshold 1e-6"
    if name == "cluster200":
        w = workloads.cluster(natom=200, seed=1002, nrad=100, nang=194, device=device)
        alphas = 0.5 * 100.0 ** (np.arange(6) / 5.0)  # geometric 0.5 .. 50 bohr^-1
        basis = ExpBasisFuncHelper.uniform(w.numbers, alphas, order=1.0)
        kw = dict(threshold=1e-6, maxiter=200, solver="m-newton", basis_func=basis, grid_type=2)
        return w, LinearISAWPart, kw, "aLISA, 6 Slater functions (0.5..50), m-newton on the molecular grid, threshold 1e-6"
    if name == "protein2000":
        # Becke weights skipped: on atom grids they enter only the entropy
        w = workloads.cluster(natom=2000, seed=1003, nrad=75, nang=110, probs=PROTEIN_PROBS, becke=False, device=device)
        kw = dict(threshold=1e-6, maxiter=1000, cutoff_radius=12.0)
        return w, MBISWPart, kw, "MBIS, 12-bohr cut-off lists, threshold 1e-6 (molecular weights without Becke factors)"
    raise SystemExit(f"unknown workload {name}")


def make_part(w, cls, kw, precision):
    part = cls(w.coordinates, w.numbers, w.pseudo_numbers, w.grid, w.density, precision=precision, **kw)
    dev = part._ensure_device()
    return part, dev


def time_steps(dev, propars0, precision, steps, warmup):
    lat = []
    out = None
    for i in range(warmup + steps):
        t0 = time.perf_counter()
        dev.set_propars(propars0)
        rc = dev.run(precision, 1)
        out = dev.results()
        t1 = time.perf_counter()
        if rc != 0:
            raise RuntimeError(f"device run failed with {rc}")
        if i >= warmup:
            lat.append(t1 - t0)
    return np.array(lat), out


def _oracle_for(w, cls, kw, maxiter):
    """The oracle object of the same scheme and settings."""
    from horton_part_amd import GaussianISAWPart, LinearISAWPart
    from oracle.horton_oracle import OracleBasis, OracleGISA, OracleLISA, OracleMBIS

    okw = {k: v for k, v in kw.items() if k in ("threshold", "grid_type", "cutoff_radius")}
    okw["maxiter"] = maxiter
    if cls is GaussianISAWPart or cls is LinearISAWPart:
        part = cls(w.coordinates, w.numbers, w.pseudo_numbers, w.grid, w.density, **kw)
        bs = part.bs_helper
        okw["basis"] = OracleBasis(bs.orders, bs.exponents, bs.initials)
        if cls is LinearISAWPart:
            okw["solver"] = kw.get("solver", "sc")
            return OracleLISA(w.coordinates, w.numbers, w.pseudo_numbers, w.grid, w.density, **okw)
        return OracleGISA(w.coordinates, w.numbers, w.pseudo_numbers, w.grid, w.density, **okw)
    return OracleMBIS(w.coordinates, w.numbers, w.pseudo_numbers, w.grid, w.density, **okw)


def cpu_baseline(w, cls, kw, gpu_niter, name, budget_s=20.0, max_point_atoms=2e8):
    """Time the numpy oracle (reference structure, 1 thread) on a bounded
    number of outer iterations and extrapolate to a full partition with the
    GPU's iteration count.  Workloads whose points x atoms exceed
    ``max_point_atoms`` (the oracle stores every atom's distance to every
    point, as the reference does) are sampled on a smaller molecule of the
    same kind and extrapolated per point x atom."""
    import warnings

    from horton_part_amd import workloads

    warnings.simplefilter("ignore")
    sample_w, note = w, "the same workload"
    if w.npoint * w.natom > max_point_atoms:
        # points x atoms of the sample: about max_point_atoms / 4
        nsmall = max(3, int(np.sqrt(max_point_atoms / 4 / (w.npoint / w.natom))))
        nrad = w.grid.atgrids[0].rgrid.size
        nang = w.grid.atgrids[0].size // nrad
        probs = PROTEIN_PROBS if name == "protein2000" else None
        sample_w = workloads.cluster(natom=nsmall, seed=1, nrad=nrad, nang=nang, probs=probs,
                                     becke=name != "protein2000", device=0)
        note = f"a {nsmall}-atom cluster of the same kind ({sample_w.npoint} points)"
    from threadpoolctl import threadpool_limits

    n, used = 1, 0.0
    while True:
        o = _oracle_for(sample_w, cls, kw, n)
        t0 = time.perf_counter()
        with threadpool_limits(1):  # the reference's numpy work runs on one core
            o.do_partitioning()
        dt = time.perf_counter() - t0
        used += dt
        if used + 2 * dt > budget_s or n >= gpu_niter:
            break
        n *= 2
    per_iter = dt / n
    scale = (w.npoint * w.natom) / (sample_w.npoint * sample_w.natom)
    total = per_iter * scale * gpu_niter
    return {
        "value": 1.0 / total,
        "unit": "partitions/s",
        "cores": 1,
        "kind": "port",
        "sample": f"{n} outer iteration(s) of the numpy oracle (reference structure) on {note}: {per_iter:.3f} s/iteration"
        + (f", scaled by points x atoms (x{scale:.1f})" if scale != 1 else "")
        + f", extrapolated to the GPU's {gpu_niter} FP64 iterations per partition",
        "host_cpus": ncpu,
        "cpu_model": platform.processor() or platform.machine(),
    }


# ---------------------------------------------------------------------------
# serving stream
# ---------------------------------------------------------------------------
SERVING_SEED = 1004
SERVING_THRESHOLD = 1e-6
SERVING_MAXITER = 1000
SERVING_GRID = (100, 194)


def _scheme_part(r, w):
    from horton_part_amd import GaussianISAWPart, MBISWPart

    cls = MBISWPart if r["scheme"] == "mbis" else GaussianISAWPart
    return cls(w.coordinates, w.numbers, w.pseudo_numbers, w.grid, w.density, threshold=SERVING_THRESHOLD, maxiter=SERVING_MAXITER)


def _serving_molecule(r, device):
    from horton_part_amd import workloads

    return workloads.chain(natom=r["natom"], seed=r["mol_seed"], nrad=SERVING_GRID[0], nang=SERVING_GRID[1], device=device)


def build_request(r, device):
    """Molecule/scheme arrays of one request, bulk arrays in pinned memory."""
    from horton_part_amd import _native
    from horton_part_amd.serving import request_arrays

    w = _serving_molecule(r, device)
    mol, sch, p0 = request_arrays(_scheme_part(r, w))
    for k in ("points", "density", "mol_weights", "sph_weights"):
        mol[k] = _native.pinned_copy(mol[k])
    return {"req": r, "mol": mol, "sch": sch, "p0": p0}


class StreamBuilder:
    """Builds the requests of a stream in arrival order on a few host threads
    (numpy and the library release the interpreter lock), at most
    ``lookahead`` requests ahead of the ones handed out, so that host memory
    holds a bounded window of the stream."""

    def __init__(self, reqs, device, threads, lookahead):
        from concurrent.futures import ThreadPoolExecutor

        self.reqs = reqs
        self.device = device
        self.pool = ThreadPoolExecutor(max_workers=threads)
        self.lookahead = lookahead
        self.futures = {}
        self.next = 0

    def fill(self, consumed):
        while self.next < len(self.reqs) and self.next < consumed + self.lookahead:
            self.futures[self.next] = self.pool.submit(build_request, self.reqs[self.next], self.device)
            self.next += 1

    def ready(self, i):
        return i in self.futures and self.futures[i].done()

    def take(self, i):
        return self.futures.pop(i).result()

    def close(self):
        self.pool.shutdown(wait=True)


def serve_stream(server, reqs, builder, policy, realtime=True, queue_max=6, progress=None):
    """Submit each request once it has arrived (its Poisson time, relative to
    the first) and is built, keeping the server's host queue short; step the
    server until every request is done.  Returns per request (arrival,
    completion, Completed) in seconds from the first arrival."""
    n = len(reqs)
    t_first = reqs[0]["t"]
    arrival = [r["t"] - t_first for r in reqs]
    done = {}
    live = {}
    i = 0
    builder.fill(0)
    t0 = time.perf_counter()
    next_report = 0.1
    host = {"step_s": 0.0, "poll_s": 0.0, "submit_s": 0.0, "idle_s": 0.0, "steps": 0}
    while len(done) < n:
        now = time.perf_counter() - t0
        ts = time.perf_counter()
        while (i < n and (not realtime or arrival[i] <= now) and builder.ready(i)
               and server.n_queued < queue_max):
            it = builder.take(i)
            server.submit_arrays(it["mol"], it["sch"], it["p0"], precision=policy, n_polish=1, tag=i)
            live[i] = it
            i += 1
            builder.fill(i)
        host["submit_s"] += time.perf_counter() - ts
        if server.pending == 0:
            time.sleep(0.0005)
            host["idle_s"] += time.perf_counter() - ts
            continue
        ts = time.perf_counter()
        server.step()
        tp = time.perf_counter()
        host["step_s"] += tp - ts
        host["steps"] += 1
        for d in server.poll():
            done[d.tag] = (arrival[d.tag], time.perf_counter() - t0, d)
            live.pop(d.tag, None)
        host["poll_s"] += time.perf_counter() - tp
        if progress and len(done) >= next_report * n:
            progress(len(done), n, time.perf_counter() - t0, i)
            next_report += 0.1
    serve_stream.host = host
    return [done[k] for k in range(n)]


def fp64_one_at_a_time(items, device):
    """The plain FP64 GPU path, one request after another (hp_part_run)."""
    from horton_part_amd import _native

    out = []
    for it in items:
        t0 = time.perf_counter()
        dev = _native.DevicePartition(it["mol"], it["sch"], device=device)
        dev.set_propars(it["p0"])
        t1 = time.perf_counter()
        rc = dev.run("fp64", 1)
        propars, charges, change, _ = dev.results()
        t2 = time.perf_counter()
        out.append({"create_s": t1 - t0, "run_s": t2 - t1, "niter": dev.niter, "rc": rc, "charges": charges})
        dev.close()
    return out


def serving_cpu_baseline(items, fp64, budget_s=8.0):
    """The numpy oracle (reference structure, 1 thread) timed per scheme on one
    request of the sample for a bounded number of outer iterations, with the
    time of the weights passes (every pro-atom on every point: per point x
    atom) and of the per-atom updates (per atom) measured apart; the
    sample's CPU time is extrapolated request by request with its FP64
    iteration count."""
    import warnings

    from threadpoolctl import threadpool_limits

    from oracle.horton_oracle import OracleBasis, OracleGISA, OracleMBIS

    warnings.simplefilter("ignore")
    cost = {}
    notes = []
    for scheme in ("mbis", "gisa"):
        cand = [k for k in range(len(items)) if items[k]["req"]["scheme"] == scheme]
        if not cand:
            continue
        j = min(cand, key=lambda k: abs(items[k]["req"]["natom"] - 12))
        r = items[j]["req"]
        w = _serving_molecule(r, None)
        kw = {}
        if scheme == "gisa":
            bs = _scheme_part(r, w).bs_helper
            kw["basis"] = OracleBasis(bs.orders, bs.exponents, bs.initials)
        ocls = OracleMBIS if scheme == "mbis" else OracleGISA
        n = 1
        while True:
            o = ocls(w.coordinates, w.numbers, w.pseudo_numbers, w.grid, w.density, threshold=SERVING_THRESHOLD, maxiter=n, **kw)
            tw = [0.0]
            ta = [0.0]

            def timed(f, acc):
                def g(*a, **k):
                    t = time.perf_counter()
                    out = f(*a, **k)
                    acc[0] += time.perf_counter() - t
                    return out
                return g

            o.update_at_weights = timed(o.update_at_weights, tw)
            o.update_atom = timed(o.update_atom, ta)
            t0 = time.perf_counter()
            with threadpool_limits(1):  # the reference's numpy work runs on one core
                o.do_partitioning()
            dt = time.perf_counter() - t0
            if dt * 2 > budget_s or n >= fp64[j]["niter"]:
                break
            n *= 2
        cost[scheme] = (tw[0] / (n * w.npoint * w.natom), ta[0] / (n * w.natom))
        notes.append(f"{scheme}: request {r['index']} ({r['natom']} atoms, {w.npoint} points), {n} outer iteration(s) "
                     f"in {dt:.1f} s: weights {cost[scheme][0] * 1e12:.1f} ps per point x atom, per-atom updates "
                     f"{cost[scheme][1] * 1e3:.2f} ms per atom")
    total = 0.0
    for it, f in zip(items, fp64):
        cw, ca = cost[it["req"]["scheme"]]
        N, A = it["mol"]["npoint"], it["mol"]["natom"]
        total += f["niter"] * (cw * N * A + ca * A)
    ncpu = len(os.sched_getaffinity(0))
    return {
        "value": len(items) / total,
        "unit": "partitions/s",
        "cores": 1,
        "kind": "port",
        "sample": "the numpy oracle on one request per scheme for a bounded number of outer iterations ("
        + "; ".join(notes) + f"); extrapolated to the {len(items)} sampled requests with their FP64 iteration counts, "
        "per scheme (partitions / summed CPU time, no queueing)",
        "host_cpus": ncpu,
        "cpu_model": platform.processor() or platform.machine(),
        # Requests are independent, so one molecule per core is the CPU's
        # best case: the one-core figure times the host's cores, assuming
        # perfect scaling (not measured).
        "p_process_value": len(items) / total * ncpu,
        "p_process_note": f"one request per core on {ncpu} cores, perfect scaling assumed (extrapolated, not measured)",
    }


def run_serving(args):
    from horton_part_amd import PartitionServer, _native, workloads

    stream = workloads.serving_stream(seed=SERVING_SEED, nreq=args.requests)
    for k, r in enumerate(stream):
        r["index"] = k
    warm = workloads.serving_stream(seed=SERVING_SEED + 1, nreq=max(args.warmup, 1))[: args.warmup]
    for k, r in enumerate(warm):
        r["index"] = -1 - k
    npts = [r["natom"] * SERVING_GRID[0] * SERVING_GRID[1] for r in stream + warm]
    server = PartitionServer(
        device=0,
        point_capacity=args.arena_points,
        atom_capacity=args.arena_points // (SERVING_GRID[0] * SERVING_GRID[1]) + 2 * args.max_inflight,
        shell_capacity=args.arena_points // SERVING_GRID[1] + 2 * args.max_inflight,
        function_capacity=args.arena_points // (SERVING_GRID[0] * SERVING_GRID[1]) * 12,
        qp_capacity=args.arena_points // (SERVING_GRID[0] * SERVING_GRID[1]) * 81,
        max_inflight=args.max_inflight,
        max_molecule_points=max(npts),
        ticks_per_step=args.ticks,
        use_graphs=args.graphs,
        groups=args.groups,
        inner_budget=args.inner_budget,
    )
    if warm:
        log(f"warmup: {len(warm)} request(s) of another stream (seed {SERVING_SEED + 1}) served together, untimed")
        wb = StreamBuilder(warm, 0, args.builders, args.lookahead)
        serve_stream(server, warm, wb, args.serve_precision, realtime=False)
        wb.close()
    st0 = server.stats()
    log(f"serving the stream: {len(stream)} requests (seed {SERVING_SEED}) as they arrive, built on "
        f"{args.builders} host threads (precision: {args.serve_precision})")

    def progress(ndone, n, t, nsub):
        log(f"  {ndone}/{n} done at {t:.1f} s ({nsub} submitted)")

    builder = StreamBuilder(stream, 0, args.builders, args.lookahead)
    res = serve_stream(server, stream, builder, args.serve_precision, progress=progress)
    builder.close()
    st1 = server.stats()
    server.close()
    arrival = np.array([a for a, _, _ in res])
    complete = np.array([c for _, c, _ in res])
    lat = complete - arrival
    makespan = complete.max() - arrival.min()
    value = len(stream) / makespan
    log(f"  {value:.3f} partitions/s, makespan {makespan:.2f} s, p50 {np.percentile(lat, 50):.2f} s, p99 {np.percentile(lat, 99):.2f} s")
    events = {}
    for _, _, d in res:
        events[d.event] = events.get(d.event, 0) + 1

    # the plain FP64 path one request at a time, on a systematic sample of the stream
    sample_idx = list(range(0, len(stream), args.fp64_every))
    log(f"plain FP64 path, one request at a time: every {args.fp64_every}th request ({len(sample_idx)})")
    sb = StreamBuilder([stream[k] for k in sample_idx], 0, args.builders, args.lookahead)
    sb.fill(0)
    items, fp64 = [], []
    for j in range(len(sample_idx)):
        it = sb.take(j)
        sb.fill(j + 1)
        fp64 += fp64_one_at_a_time([it], device=0)
        # keep the sizes, drop the grid arrays
        it["mol"] = {k: v for k, v in it["mol"].items() if not isinstance(v, np.ndarray)}
        items.append(it)
    sb.close()
    run_s = np.array([f["run_s"] for f in fp64])
    create_s = np.array([f["create_s"] for f in fp64])
    fp64_value = len(fp64) / run_s.sum()
    fp64_e2e = len(fp64) / (run_s.sum() + create_s.sum())
    dqs = [float(np.max(np.abs(res[k][2].charges - f["charges"]))) for k, f in zip(sample_idx, fp64)]
    dq = max(dqs)
    over = [{"request": k, "scheme": stream[k]["scheme"], "natom": stream[k]["natom"], "dq": dqs[j],
             "niter": res[k][2].niter, "niter_fp64": fp64[j]["niter"]}
            for j, k in enumerate(sample_idx) if dqs[j] > 1e-6]
    per_sample = [[k, stream[k]["scheme"], stream[k]["natom"], res[k][2].niter, fp64[j]["niter"], dqs[j]]
                  for j, k in enumerate(sample_idx)]
    log(f"  {fp64_value:.3f} partitions/s (run only), {fp64_e2e:.3f} with upload; max|dq| served vs FP64 {dq:.2e}")

    d = {k: st1[k] - st0[k] for k in st1}
    ticks = max(d["ticks"], 1)
    t_w = d["weights_ms"] * 1e-3
    pts = d["points_fp32"] + d["points_fp64"]
    peaks = measured_peaks()
    # the weights pass is bound by its exponentials: the evaluations it
    # performed (after tail skipping) at the probe's rate for the same
    # exp code, per precision, against its measured time
    t_ideal = d["done_fp32"] / peaks["exp_fp32_per_s"] + d["done_fp64"] / peaks["exp_fp64_per_s"]
    exp_rate = (d["done_fp32"] + d["done_fp64"]) / t_w if t_w > 0 else 0.0
    peak_rate = (d["done_fp32"] + d["done_fp64"]) / t_ideal if t_ideal > 0 else 0.0
    bytes_per_tick = 48.0 * pts / ticks
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "partitions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": makespan / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {
            "workload": f"serving stream (seed {SERVING_SEED}): {len(stream)} requests, MBIS/GISA chains with 3-300 atoms, "
            f"{SERVING_GRID[0]}x{SERVING_GRID[1]} grids, Poisson arrivals at 200/s over {stream[-1]['t']:.1f} s, threshold "
            f"{SERVING_THRESHOLD:g}; continuous batching, "
            + {"mixed": "mixed precision + FP64 polish",
               "fp64": "plain FP64",
               "auto": "GISA mixed precision + FP64 polish, MBIS plain FP64"}[args.serve_precision]
            + f"; one step = 1/{args.steps} of the stream",
            "precision_policy": args.serve_precision,
            "requests": len(stream),
            "atoms": int(sum(r["natom"] for r in stream)),
            "points": int(sum(r["natom"] for r in stream) * SERVING_GRID[0] * SERVING_GRID[1]),
            "arena_points": args.arena_points,
            "max_inflight": args.max_inflight,
            "ticks_per_step": args.ticks,
            "graphs": args.graphs,
            "groups": args.groups or (3 if int(os.environ["GPU_MAX_HW_QUEUES"]) >= 8 else 2),
            "hw_queues": int(os.environ["GPU_MAX_HW_QUEUES"]),
            "inner_budget": args.inner_budget,
            "builder_threads": args.builders,
        },
        "p50_ms": float(np.percentile(lat, 50)) * 1e3,
        "p99_ms": float(np.percentile(lat, 99)) * 1e3,
        "makespan_s": makespan,
        "events": events,
        "fp64_one_at_a_time": {
            "value": fp64_value,
            "value_with_upload": fp64_e2e,
            "ms_per_request": float(run_s.mean()) * 1e3,
            "sample": f"every {args.fp64_every}th request of the stream ({len(fp64)} requests), built as served; the "
            "sample mean time per request estimates the stream's",
            "iterations": int(sum(f["niter"] for f in fp64)),
            "errors": int(sum(f["rc"] != 0 for f in fp64)),
        },
        "speedup_vs_fp64": value / fp64_value,
        "max_dq_vs_fp64": dq,
        "max_dq_scope": f"the {len(fp64)} sampled requests: served result vs the FP64 path one at a time",
        "dq_over_1e-6": over,
        "dq_per_sample": {"columns": ["request", "scheme", "natom", "niter", "niter_fp64", "dq"], "rows": per_sample},
        "host_loop": serve_stream.host,
        "batch": {
            "ticks": d["ticks"],
            "molecule_iterations": d["iterations"],
            "tick_ms_mean": d["tick_ms"] / ticks,
            "weights_ms_mean": d["weights_ms"] / ticks,
            "inner_ms_mean": d["inner_ms"] / ticks,
            "gpu_busy_frac": d["tick_ms"] * 1e-3 / makespan,
        },
        "roofline": {
            "kernel": "k_weights (fused pro-atom/promolecule/weights/spherical average), FP32 + FP64 launches of a tick",
            "bound": "valu",
            "achieved": exp_rate / 1e12,
            "peak": peak_rate / 1e12,
            "unit": "Texp/s",
            "frac": t_ideal / t_w if t_w > 0 else 0.0,
            "peak_note": "the exponential throughput of the same exp code with every lane busy and no memory traffic "
            "(hp_exp_rate probe), per precision, weighted by the evaluations each precision performed",
            "traffic": K_WEIGHTS_TRAFFIC_RATIO * bytes_per_tick,
            "traffic_note": "HBM bytes per tick (all weights launches of a tick): the algorithmic bytes_per_tick times "
            "the measured traffic ratio of the rocprofv3 FETCH_SIZE/WRITE_SIZE runs",
            "hbm_gbps": bytes_per_tick / (t_w / ticks) / 1e9 if t_w > 0 else 0.0,
            "hbm_frac": bytes_per_tick / (t_w / ticks) / 1e9 / HBM_PEAK_GBPS if t_w > 0 else 0.0,
            "bytes_per_tick": bytes_per_tick,
            "weights_ms_per_tick": t_w / ticks * 1e3,
            "evals_done_frac": (d["done_fp32"] + d["done_fp64"]) / max(d["evals_fp32"] + d["evals_fp64"], 1.0),
            "fp32_point_frac": d["points_fp32"] / pts if pts else 0.0,
        },
        "measured": peaks,
    }
    if not args.no_cpu:
        log("CPU baseline (numpy oracle, 1 thread, bounded sample)")
        out["cpu_baseline"] = serving_cpu_baseline(items, fp64)
    print(json.dumps(out))


# ---------------------------------------------------------------------------
# tabulated pro-atoms (ISA, Hirshfeld): FP64 only
# ---------------------------------------------------------------------------
TAB_CHAIN_PROBS = {"C": 0.40, "H": 0.45, "O": 0.15}  # elements of the STO-3G pro-atom records


def run_tabulated(args):
    """ISA (``isa-chain20``: 200 outer iterations on the 20-atom chain, 150 x
    302 grids) or Hirshfeld (``h-chain20``: one weights pass with the HF/STO-3G
    neutral records of C, H and O on a C/H/O chain) timed as partitions/s
    with inputs resident; roofline of the tabulated weights pass."""
    import warnings

    from horton_part_amd import HirshfeldWPart, ISAWPart, _native, workloads

    warnings.simplefilter("ignore")
    tests_dir = os.path.join(ROOT, "tests")
    if tests_dir not in sys.path:
        sys.path.insert(0, tests_dir)
    isa = args.workload == "isa-chain20"
    probs = None if isa else TAB_CHAIN_PROBS
    w = workloads.chain(natom=20, seed=1001, nrad=150, nang=302, probs=probs, device=0)
    maxiter = 200
    if isa:
        part = ISAWPart(w.coordinates, w.numbers, w.pseudo_numbers, w.grid, w.density, threshold=1e-12, maxiter=maxiter)
        desc = f"ISA, {maxiter} outer iterations (threshold 1e-12 not reached)"
    else:
        from common import load_proatomdb

        db = load_proatomdb([1, 6, 8], 0, 0, "hf_sto3g")
        part = HirshfeldWPart(w.coordinates, w.numbers, w.pseudo_numbers, w.grid, w.density, db)
        desc = "Hirshfeld, HF/STO-3G neutral records of C, H, O"
    dev = part._ensure_device()
    p0 = np.zeros(dev.npar)
    lat = []
    for i in range(args.warmup + args.steps):
        t0 = time.perf_counter()
        dev.set_propars(p0)
        if isa:
            rc = dev.run("fp64", 1)
            if rc != 0:
                raise RuntimeError(f"device run failed with {rc}")
        else:
            dev.weights(spherical=False)
            dev.populations()
        if i >= args.warmup:
            lat.append(time.perf_counter() - t0)
    lat = np.array(lat)
    ms, n = dev.kernel_time(_native.KERNEL_WEIGHTS)
    wk = dev.work()
    t_w = ms / max(n, 1) * 1e-3
    achieved = wk["bytes_per_pass"] / t_w / 1e9
    evals = w.npoint * w.natom
    peaks = measured_peaks()
    out = {
        "metric": "partitions/s (tabulated pro-atoms, FP64)",
        "value": 1.0 / float(np.mean(lat)),
        "unit": "partitions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": float(np.mean(lat)) * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: {desc}, {w.natom} atoms, {w.npoint} grid points"},
        "niter": dev.niter if isa else 1,
        "roofline": {
            "kernel": "k_table_weights (spline pro-atoms: interval search and cubic per point and atom, promolecule, "
                      "weights, per-shell sums)",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBPS,
            "traffic": None,
            "bytes_per_launch": wk["bytes_per_pass"],
            "launch_us": t_w * 1e6,
            "spline_evals_per_s": evals / t_w,
        },
        "measured": peaks,
    }
    if not args.no_cpu:
        from oracle import horton_oracle as ho

        log("CPU baseline (numpy oracle, 1 thread, bounded sample)")
        from threadpoolctl import threadpool_limits

        with threadpool_limits(1):
            t0 = time.perf_counter()
            if isa:
                ho.OracleISA(w.coordinates, w.numbers, w.pseudo_numbers, w.grid, w.density, maxiter=2).do_partitioning()
            else:
                odb = ho.OracleProAtomDB([ho.load_record_npz(f) for f in __import__("common").proatom_files([1, 6, 8], 0, 0, "hf_sto3g")])
                ho.OracleHirshfeld(w.coordinates, w.numbers, w.pseudo_numbers, w.grid, w.density, proatomdb=odb).do_partitioning()
            dt = time.perf_counter() - t0
        per = dt / 2 * maxiter if isa else dt
        out["cpu_baseline"] = {
            "value": 1.0 / per,
            "unit": "partitions/s",
            "cores": 1,
            "kind": "port",
            "sample": (f"2 outer iterations of the numpy oracle on the same workload ({dt:.2f} s), extrapolated to {maxiter}"
                       if isa else f"the numpy oracle's Hirshfeld pass on the same workload ({dt:.2f} s)"),
        }
    print(json.dumps(out))


# ---------------------------------------------------------------------------
# gLISA: the global terms of one Newton step (objective, gradient, Hessian)
# ---------------------------------------------------------------------------
def run_glisa(args):
    """gLISA's global terms on chain20 (gauss.json basis, 150 x 302 grids):
    one step = one Newton step's objective, gradient and M x M Hessian
    G^T diag(w rho / rho0^2) G over the molecular grid at the initial
    coefficients (``hp_part_global_terms``); roofline of the FP64 MFMA
    Hessian pass against M (M + 1) N algorithmic flops: the symmetric
    product's M (M + 1) / 2 distinct entries, a multiply and an add per entry
    and point (the kernel forms the upper 16 x 16 tiles only), and 2 M N for
    the gradient when the same kernel forms it."""
    import warnings

    from horton_part_amd import GlobalLinearISAWPart, _native, workloads

    warnings.simplefilter("ignore")
    w = workloads.chain(natom=20, seed=1001, nrad=150, nang=302, device=0)
    kw = dict(solver="newton", basis_func="gauss", grid_type=2)
    part = GlobalLinearISAWPart(w.coordinates, w.numbers, w.pseudo_numbers, w.grid, w.density, **kw)
    dev = part._ensure_device()
    p0 = part._init_propars()
    M, N = int(p0.size), int(w.npoint)
    lat = []
    for i in range(args.warmup + args.steps):
        t0 = time.perf_counter()
        dev.global_terms(p0, nderiv=2)
        if i >= args.warmup:
            lat.append(time.perf_counter() - t0)
    lat = np.array(lat)
    ms, n = dev.kernel_time(_native.KERNEL_GLISA_HESS)
    t_h = ms / max(n, 1) * 1e-3
    Mp = (M + 15) // 16 * 16
    T = Mp // 16
    fused = Mp <= 112  # the Hessian kernel also forms the gradient partials
    flops = float(M) * (M + 1) * N + (2.0 * M * N if fused else 0.0)
    executed = 2.0 * 256 * (T * (T + 1) // 2) * N  # the MFMA work of the upper tiles, padding included
    tf = flops / t_h / 1e12
    out = {
        "metric": "gLISA Newton steps/s (objective, gradient, Hessian over the molecular grid, FP64)",
        "value": 1.0 / float(np.mean(lat)),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": float(np.mean(lat)) * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"glisa-chain20: gLISA global terms, gauss.json basis (M = {M} functions), "
                               f"{w.natom} atoms, {N} grid points"},
        "roofline": {
            "kernel": ("k_glisa_hess_lds (every upper 16x16 tile of G^T D G per point chunk on mfma_f64_16x16x4f64, "
                       "rows of G staged in LDS, with the gradient partials E G from the same rows)" if fused else
                       "k_glisa_hess (one wavefront per upper 16x16 tile pair and point chunk, mfma_f64_16x16x4f64)"),
            "bound": "mfma",
            "achieved": tf,
            "peak": 78.6,
            "unit": "TFLOP/s",
            "frac": tf / 78.6,
            "traffic": None,
            "flops_per_launch": flops,
            "flops_note": "M (M + 1) N: the distinct entries of the symmetric M x M product over N points"
                          + (", + 2 M N for the gradient formed in the same launch" if fused else ""),
            "mfma_flops_per_launch": executed,
            "mfma_frac": executed / t_h / 1e12 / 78.6,
            "launch_us": t_h * 1e6,
        },
    }
    if not args.no_cpu:
        from threadpoolctl import threadpool_limits

        from oracle.horton_oracle import OracleBasis, OracleGLISA

        log("CPU baseline (numpy oracle, 1 thread): one working_matrix call")
        bs = part.bs_helper
        o = OracleGLISA(w.coordinates, w.numbers, w.pseudo_numbers, w.grid, w.density, grid_type=2,
                        basis=OracleBasis(bs.orders, bs.exponents, bs.initials))
        q0 = o.init_propars()
        with threadpool_limits(1):
            t0 = time.perf_counter()
            o.working_matrix(o.moldens, o.calc_promol_dens(q0), nderiv=2)
            dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": 1.0 / dt, "unit": "steps/s", "cores": 1, "kind": "port",
                               "sample": f"one objective/gradient/Hessian evaluation of the numpy oracle on the same "
                                         f"workload ({dt:.1f} s)"}
    print(json.dumps(out))


def measured_peaks():
    """Device copy bandwidth and the exp throughput of both precisions."""
    from horton_part_amd import _native

    gbps, _ = _native.stream_copy(1 << 30, 10)
    return {
        "copy_gbps": gbps,
        "exp_fp64_per_s": _native.exp_rate("fp64"),
        "exp_fp32_per_s": _native.exp_rate("mixed"),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="serving: the stream is timed as this many steps of equal size (ms_per_step = makespan / steps); "
                    "single molecules: timed partitions")
    ap.add_argument("--warmup", type=int, default=None,
                    help="serving: requests of another stream served untimed first; single molecules: untimed partitions")
    ap.add_argument("--requests", type=int, default=4096, help="serving: requests of the stream (4096: the whole stream)")
    ap.add_argument("--builders", type=int, default=8, help="serving: host threads that build the requests")
    ap.add_argument("--lookahead", type=int, default=24, help="serving: requests built ahead of submission")
    ap.add_argument("--arena-points", type=int, default=320_000_000,
                    help="serving: grid points resident at once (48 B each: 15.4 GB at the default)")
    ap.add_argument("--max-inflight", type=int, default=256, help="serving: molecules resident at once")
    ap.add_argument("--fp64-every", type=int, default=64, help="serving: FP64 one-at-a-time sample: every n-th request")
    ap.add_argument("--switch-interval", type=float, default=0.0,
                    help="serving: interpreter thread switch interval in seconds (0: the default)")
    ap.add_argument("--workload", default="serving",
                    choices=["serving", "water", "chain20", "chain20-gisa", "cluster200", "protein2000", "isa-chain20",
                             "h-chain20", "glisa-chain20"])
    ap.add_argument("--ticks", type=int, default=8, help="serving: outer iterations per host step")
    ap.add_argument("--groups", type=int, default=0, help="serving: molecule groups interleaved per tick (1-4; 0: 3 with >= 8 hardware queues, else 2)")
    ap.add_argument("--inner-budget", type=int, default=0,
                    help="serving: MBIS inner iterations per atom and launch (0: no limit)")
    ap.add_argument("--graphs", action="store_true", help="serving: replay the ticks of a step as a hipGraph (times whole steps only)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--serve-precision", choices=("mixed", "fp64", "auto"), default="auto",
                    help="serving: precision of the requests (auto: GISA mixed, MBIS FP64)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: run it with --gpus 1 (multi-GPU sharding is a design note, not a run)", file=sys.stderr)
        sys.exit(2)

    from horton_part_amd import _native

    if _native.device_count() < 1:
        raise SystemExit("no GPU visible")
    if args.workload == "serving":
        if args.switch_interval > 0:
            # the loop that steps the batch gets the interpreter back sooner
            # from the threads that build requests
            sys.setswitchinterval(args.switch_interval)
        args.steps = 20 if args.steps is None else args.steps
        args.warmup = 4 if args.warmup is None else args.warmup
        return run_serving(args)
    args.steps = 5 if args.steps is None else args.steps
    args.warmup = 1 if args.warmup is None else args.warmup
    if args.workload in ("isa-chain20", "h-chain20"):
        return run_tabulated(args)
    if args.workload == "glisa-chain20":
        return run_glisa(args)
    log(f"building workload {args.workload} on the GPU")
    t0 = time.perf_counter()
    w, cls, kw, desc = build_workload(args.workload, device=0)
    log(f"  {w.natom} atoms, {w.npoint} points, built in {time.perf_counter() - t0:.1f} s")

    results = {}
    for precision in ("fp64", "mixed"):
        part, dev = make_part(w, cls, kw, precision)
        propars0 = part._init_propars()
        log(f"timing {precision}: warmup {args.warmup}, steps {args.steps}")
        lat, (propars, charges, change, _) = time_steps(dev, propars0, precision, args.steps, args.warmup)
        k = {name: dev.kernel_time(code) for name, code in
             (("weights", _native.KERNEL_WEIGHTS), ("inner", _native.KERNEL_INNER), ("grid", _native.KERNEL_LISA_GRID))}
        results[precision] = {
            "lat": lat,
            "niter": dev.niter,
            "charges": charges,
            "kernels": k,
            "work": dev.work(),
        }
        log(f"  {precision}: {np.mean(lat) * 1e3:.2f} ms/partition, niter {dev.niter}, "
            + ", ".join(f"{n} {ms / max(c, 1) * 1e3:.1f} us/launch" for n, (ms, c) in k.items() if c))
        dev.close()

    fp64, mixed = results["fp64"], results["mixed"]
    dq = float(np.max(np.abs(mixed["charges"] - fp64["charges"])))
    lat = mixed["lat"]
    value = 1.0 / float(np.mean(lat))
    peaks = measured_peaks()

    def per_launch(r, name):
        ms, n = r["kernels"][name]
        return ms / max(n, 1)

    wk = fp64["work"]
    t_w = per_launch(fp64, "weights") * 1e-3
    hbm = wk["bytes_per_pass"] / t_w / 1e9
    # exponentials performed (after culling and tail skipping) per second of
    # the FP64 weights launches, against the exp probe's rate for the same code
    exps = wk["evals_done_fp64"] / (fp64["kernels"]["weights"][0] * 1e-3) / 1e12
    exp_peak = peaks["exp_fp64_per_s"] / 1e12
    roof = {
        "kernel": "k_weights<double> (fused pro-atom/promolecule/weights/spherical average)",
        "bound": "valu",
        "achieved": exps,
        "peak": exp_peak,
        "unit": "Texp/s",
        "frac": exps / exp_peak,
        "peak_note": "the exponential throughput of the same exp code with every lane busy and no memory traffic "
                     "(hp_exp_rate probe)",
        "traffic": wk["bytes_per_pass"] * K_WEIGHTS_TRAFFIC_RATIO,
        "traffic_note": "HBM bytes per launch: the algorithmic 48 B per point times the measured traffic ratio of "
                        "the weights launches",
        "bytes_per_launch": wk["bytes_per_pass"],
        "hbm_gbps": hbm,
        "hbm_frac": hbm / HBM_PEAK_GBPS,
        "launch_us": t_w * 1e6,
        "evals_per_launch": wk["evals_per_pass"],
        "evals_done_frac": wk["evals_done_fp64"] / max(wk["evals_per_pass"] * fp64["niter"], 1.0),
        "copy_frac": hbm / peaks["copy_gbps"],
    }
    if fp64["kernels"]["grid"][1]:
        # molecular-grid Newton passes: the Hessian B diag(w) B^T on MFMA
        ms, n = fp64["kernels"]["grid"]
        tf = wk["hessian_flops"] / (ms * 1e-3) / 1e12 if ms > 0 else 0.0
        roof = {
            "kernel": "k_gt2_pass<double, Newton> (per-atom gradient and Hessian over the molecular grid, FP64 MFMA)",
            "bound": "mfma",
            "achieved": tf,
            "peak": 78.6,
            "unit": "TFLOP/s",
            "frac": tf / 78.6,
            "traffic": None,
            "bound_note": "the Hessian of K = 6 functions fills 6 x 6 of each 16 x 16 tile; SQ counters of this "
                          "workload show the vector ALUs issuing 56% of the pass's time and the matrix pipe busy "
                          "20% (exponentials, divisions and the per-point sums bound it)",
            "hessian_flops": wk["hessian_flops"],
            "hessian_passes": wk["hessian_passes"],
            "pass_launches": n,
            "pass_us": ms / max(n, 1) * 1e3,
            "weights_pass": roof,
        }
        wm = mixed["work"]
        ms32, n32 = mixed["kernels"]["grid"]
        if ms32 > 0 and wm["hessian_flops"] > 0:
            roof["mixed_fp32_mfma_tflops"] = wm["hessian_flops"] / (ms32 * 1e-3) / 1e12
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "partitions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": float(np.mean(lat)) * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {
            "workload": f"{args.workload}: {desc}, {w.natom} atoms, {w.npoint} grid points",
            "natom": w.natom,
            "npoint": int(w.npoint),
            "precision": "mixed (FP32 grid arithmetic, FP64 sums/solves) + FP64 polish",
        },
        "p50_ms": float(np.percentile(lat, 50)) * 1e3,
        "p99_ms": float(np.percentile(lat, 99)) * 1e3,
        "niter_mixed": mixed["niter"],
        "fp64": {
            "value": 1.0 / float(np.mean(fp64["lat"])),
            "ms_per_step": float(np.mean(fp64["lat"])) * 1e3,
            "p99_ms": float(np.percentile(fp64["lat"], 99)) * 1e3,
            "niter": fp64["niter"],
        },
        "max_dq_mixed_vs_fp64": dq,
        "speedup_mixed_vs_fp64": float(np.mean(fp64["lat"]) / np.mean(lat)),
        "kernels_us": {f"{n}_{p}": per_launch(results[p], n) * 1e3 for p in results for n in ("weights", "inner", "grid")
                       if results[p]["kernels"][n][1]},
        "roofline": roof,
        "measured": peaks,
    }
    if not args.no_cpu:
        log("CPU baseline (numpy oracle, 1 thread, bounded sample)")
        out["cpu_baseline"] = cpu_baseline(w, cls, kw, fp64["niter"], args.workload)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
