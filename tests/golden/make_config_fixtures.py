"""Run the oracle on the benchmark-configuration cases and store the results.

Usage (from the repository root)::

    python tests/golden/make_config_fixtures.py [case ...] [--jobs N]

Writes ``tests/golden/configs/<case>.npz`` for every case of
``tests/configs.py`` (or the named ones).  Each file holds the digest of the
inputs the case was run on and the oracle's results: charges, populations,
parameters, iteration count, last change, the per-atom inner iteration
counts of the last outer iteration, and the change and entropy histories.
The oracle is the numpy restatement in ``oracle/horton_oracle.py``; each case
runs in its own process on one thread.
"""

import argparse
import os
import sys
import time
import warnings

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
for p in (os.path.join(ROOT, "horton-part_amd"), ROOT, os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)

OUT = os.path.join(ROOT, "tests", "golden", "configs")


def run_case(name):
    os.environ["OMP_NUM_THREADS"] = "1"
    os.environ["OPENBLAS_NUM_THREADS"] = "1"
    import numpy as np
    from threadpoolctl import threadpool_limits

    import configs

    warnings.simplefilter("ignore")
    case = configs.all_cases()[name]
    t0 = time.time()
    digest = case.input_digest()
    t1 = time.time()
    oracle = case.oracle()
    out = {"digest": digest}
    with threadpool_limits(1):
        try:
            cache = oracle.do_partitioning()
        except Exception as exc:  # the device must raise the same type
            out["raised"] = np.array(type(exc).__name__)
            cache = None
    t2 = time.time()
    if cache is not None:
        inner = [s.get("niter", 1) for s in oracle.inner_stats[-1]]
        out.update(
            charges=cache["charges"],
            populations=cache["populations"],
            propars=cache["propars"],
            niter=np.array(cache["niter"]),
            change=np.array(cache["change"]),
            inner_niter=np.array(inner, dtype=np.int64),
            history_changes=cache["history_changes"],
            history_entropies=cache["history_entropies"],
            history_charges_head=cache["history_charges"][:3],
            history_charges_tail=cache["history_charges"][-3:],
        )
    out["seconds"] = np.array([t1 - t0, t2 - t1])
    os.makedirs(OUT, exist_ok=True)
    np.savez_compressed(os.path.join(OUT, name + ".npz"), **out)
    niter = int(out["niter"]) if "niter" in out else -1
    return f"{name}: niter {niter} build {t1 - t0:.1f} s oracle {t2 - t1:.1f} s" + (f" raised {out['raised']}" if "raised" in out else "")


def main():
    import multiprocessing as mp

    import configs

    ap = argparse.ArgumentParser()
    ap.add_argument("cases", nargs="*")
    ap.add_argument("--jobs", type=int, default=4)
    args = ap.parse_args()
    names = args.cases or list(configs.all_cases())
    with mp.get_context("spawn").Pool(args.jobs) as pool:
        for msg in pool.imap_unordered(run_case, names):
            print(msg, flush=True)


if __name__ == "__main__":
    main()
