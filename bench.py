"""Benchmark of the stockholder partitioning loop on one MI355X.

Default workload: MBIS on the 20-atom C/H/N/O chain (seed 1001) with
150 radial x 302 Lebedev points per atom (906,000 points), threshold 1e-6.
One step = one partition: reset the pro-atom parameters to their initial
values, run the outer loop to convergence on the GPU (inputs already resident
in HBM) and read back charges and parameters.  ``value`` is the mixed path's
partitions/s (FP32 pro-atoms/promolecule/ratios, FP64 sums and solves, one
FP64 polish iteration); the plain FP64 path is timed beside it, and
max |dq| between the two is reported (the FP64 path equals the CPU oracle
to 1e-10 e in the parity tests).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload chain20]
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

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (os.path.join(ROOT, "horton-part_amd"), ROOT):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np  # noqa: E402

METRIC = "partitions/s & p99 latency, mixed vs plain FP64 at max|Δq|≤1e-6 e"
HBM_PEAK_GBPS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def build_workload(name, device):
    from horton_part_amd import workloads

    if name == "chain20":
        return workloads.chain(natom=20, seed=1001, nrad=150, nang=302, device=device), "mbis", 1e-6
    if name == "water":
        return workloads.water(seed=1000, nrad=100, nang=194, device=device), "mbis", 1e-8
    if name == "cluster200":
        return workloads.cluster(natom=200, seed=1002, nrad=100, nang=194, device=device), "mbis", 1e-6
    raise SystemExit(f"unknown workload {name}")


def make_part(w, scheme, threshold, precision, maxiter=1000):
    from horton_part_amd import MBISWPart

    part = MBISWPart(w.coordinates, w.numbers, w.pseudo_numbers, w.grid, w.density, threshold=threshold, maxiter=maxiter, precision=precision)
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


def cpu_baseline(w, threshold, gpu_niter, budget_s=20.0):
    """Time the numpy oracle (reference structure, 1 thread) on a bounded
    sample of outer iterations and extrapolate to a full partition."""
    from oracle.horton_oracle import OracleMBIS

    import warnings

    warnings.simplefilter("ignore")
    # time a growing number of outer iterations until the budget is used
    n = 1
    per_iter = None
    used = 0.0
    while True:
        o = OracleMBIS(w.coordinates, w.numbers, w.pseudo_numbers, w.grid, w.density, threshold=threshold, maxiter=n)
        t0 = time.perf_counter()
        o.do_partitioning()
        dt = time.perf_counter() - t0
        used += dt
        per_iter = dt / n
        if used + 2 * dt > budget_s or n >= gpu_niter:
            break
        n *= 2
    total = per_iter * gpu_niter
    return {
        "value": 1.0 / total,
        "unit": "partitions/s",
        "cores": 1,
        "kind": "port",
        "sample": f"{n} outer iteration(s) of the numpy oracle (reference structure) on the same workload, "
        f"{per_iter:.3f} s/iteration, extrapolated to the GPU's {gpu_niter} iterations per partition",
        "host_cpus": len(os.sched_getaffinity(0)),
        "cpu_model": platform.processor() or platform.machine(),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="chain20")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: run it with --gpus 1 (multi-GPU sharding is a design note, not a run)", file=sys.stderr)
        sys.exit(2)

    from horton_part_amd import _native

    if _native.device_count() < 1:
        raise SystemExit("no GPU visible")
    log(f"building workload {args.workload} on the GPU")
    t0 = time.perf_counter()
    w, scheme, threshold = build_workload(args.workload, device=0)
    log(f"  {w.natom} atoms, {w.npoint} points, built in {time.perf_counter() - t0:.1f} s")

    results = {}
    for precision in ("fp64", "mixed"):
        part, dev = make_part(w, scheme, threshold, precision)
        propars0 = part._init_propars()
        log(f"timing {precision}: warmup {args.warmup}, steps {args.steps}")
        lat, (propars, charges, change, _) = time_steps(dev, propars0, precision, args.steps, args.warmup)
        kw_ms, kw_n = dev.kernel_time(_native.KERNEL_WEIGHTS)
        ki_ms, ki_n = dev.kernel_time(_native.KERNEL_INNER)
        results[precision] = {
            "lat": lat,
            "niter": dev.niter,
            "charges": charges,
            "weights_kernel_ms": kw_ms / max(kw_n, 1),
            "inner_kernel_ms": ki_ms / max(ki_n, 1),
            "weights_launches": kw_n,
        }
        log(f"  {precision}: {np.mean(lat) * 1e3:.2f} ms/partition, niter {dev.niter}, "
            f"weights {kw_ms / max(kw_n, 1) * 1e3:.1f} us/launch, inner {ki_ms / max(ki_n, 1) * 1e3:.1f} us/launch")
        dev.close()

    fp64, mixed = results["fp64"], results["mixed"]
    dq = float(np.max(np.abs(mixed["charges"] - fp64["charges"])))
    lat = mixed["lat"]
    value = 1.0 / float(np.mean(lat))

    # roofline of the fused weights pass (FP64 path): algorithmic bytes per
    # launch = 48 B per point (x, y, z, rho, angular weight, molecular weight)
    # + 16 B per radial shell written; exp terms = points x sum_a K_a.
    nshell = sum(g.rgrid.size for g in w.grid.atgrids)
    from horton_part_amd.mbis import get_nshell

    sum_k = int(sum(get_nshell(z) for z in w.numbers))
    bytes_per_launch = 48.0 * w.npoint + 16.0 * nshell
    t_launch = fp64["weights_kernel_ms"] * 1e-3
    achieved = bytes_per_launch / t_launch / 1e9
    exp_terms = float(w.npoint) * sum_k
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
            "workload": f"{args.workload}: MBIS, {w.natom} atoms, {w.npoint} grid points, threshold {threshold:g}",
            "scheme": scheme,
            "natom": w.natom,
            "npoint": int(w.npoint),
            "precision": "mixed (FP32 grid arithmetic, FP64 sums/solves) + 1 FP64 polish iteration",
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
        "kernels_us": {
            "weights_fp64": fp64["weights_kernel_ms"] * 1e3,
            "weights_mixed": mixed["weights_kernel_ms"] * 1e3,
            "inner_fp64": fp64["inner_kernel_ms"] * 1e3,
            "inner_mixed": mixed["inner_kernel_ms"] * 1e3,
        },
        "roofline": {
            "kernel": "k_weights<double> (fused pro-atom/promolecule/weights/spherical average)",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBPS,
            "traffic": None,
            "exp_terms_per_launch": exp_terms,
            "exp_terms_per_s": exp_terms / t_launch,
        },
    }
    if not args.no_cpu:
        log("CPU baseline (numpy oracle, 1 thread, bounded sample)")
        out["cpu_baseline"] = cpu_baseline(w, threshold, fp64["niter"])
    print(json.dumps(out))


if __name__ == "__main__":
    main()
