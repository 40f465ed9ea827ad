"""Write oracle_golden.json: oracle results on small seeded cases.

Run from the repository root: ``python tests/golden/make_golden.py``.
The cases are the small workloads the tests use; the file pins the oracle
against later changes (the reference-anchored pins are in test_oracle.py).
"""

import json
import os
import sys
import warnings

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
for p in (os.path.join(ROOT, "horton-part_amd"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402

from horton_part_amd import workloads  # noqa: E402
from horton_part_amd.basis import ExpBasisFuncHelper  # noqa: E402
from oracle import horton_oracle as ho  # noqa: E402

CASES = ["mbis_water_30x26", "mbis_chain4_30x26", "gisa_chain3_30x26", "lisa_sc_chain3_30x26", "lisa_mnewton_chain3_30x26"]


def run_case(name):
    warnings.simplefilter("ignore")
    if name == "mbis_water_30x26":
        w = workloads.water(nrad=30, nang=26)
        o = ho.OracleMBIS(w.coordinates, w.numbers, w.pseudo_numbers, w.grid, w.density, threshold=1e-8, maxiter=15)
    elif name == "mbis_chain4_30x26":
        w = workloads.chain(natom=4, seed=21, nrad=30, nang=26)
        o = ho.OracleMBIS(w.coordinates, w.numbers, w.pseudo_numbers, w.grid, w.density, maxiter=30)
    else:
        w = workloads.chain(natom=3, seed=22, nrad=30, nang=26)
        if name.startswith("gisa"):
            h = ExpBasisFuncHelper.from_function_type("gauss")
            basis = ho.OracleBasis(h.orders, h.exponents, h.initials)
            o = ho.OracleGISA(w.coordinates, w.numbers, w.pseudo_numbers, w.grid, w.density, maxiter=30, basis=basis)
        else:
            h = ExpBasisFuncHelper.uniform(w.numbers, 0.5 * 100.0 ** (np.arange(6) / 5.0), order=1.0)
            basis = ho.OracleBasis(h.orders, h.exponents, h.initials)
            solver = "sc" if "_sc_" in name else "m-newton"
            o = ho.OracleLISA(w.coordinates, w.numbers, w.pseudo_numbers, w.grid, w.density, maxiter=30, basis=basis, solver=solver)
    c = o.do_partitioning()
    return {"niter": int(c["niter"]), "charges": c["charges"].tolist(), "propars": c["propars"].tolist()}


if __name__ == "__main__":
    out = {name: run_case(name) for name in CASES}
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "oracle_golden.json"), "w") as f:
        json.dump(out, f, indent=1)
    print({k: v["niter"] for k, v in out.items()})
