"""part-gen -> part-dens on the GPU: the reference's command-line pins.

tests/scripts/test_partition_density.py runs part-gen on
water_sto3g_hf_g03.fchk (150 x 194 grid) and part-dens on the result:
MBIS charges within 1e-2 of [-0.6179280883502, 0.30913844981854,
0.30918595758548] on grid types 1, 2 and 3 (lines 249-251, 254-258), charges
adding up to zero within 1e-3 (MBIS family, 246-247) and 1e-4 (LISA and
gLISA, 198), oxygen negative and hydrogens positive.
"""

import os
import sys
import warnings

import numpy as np
import pytest

from conftest import require_gpu
from common import GOLDEN

pytestmark = pytest.mark.gpu
warnings.simplefilter("ignore")

FCHK = os.path.join(GOLDEN, "water_sto3g_hf_g03.fchk")
MBIS_PIN = np.array([-0.6179280883502, 0.30913844981854, 0.30918595758548])


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def density_file(tmp_path_factory):
    from horton_part_amd import io

    fn = str(tmp_path_factory.mktemp("partgen") / "density.npz")
    io.generate_density(FCHK, fn, nrad=150, nang=194)
    return fn


@pytest.mark.parametrize("grid_type", [1, 2, 3])
def test_part_dens_mbis_water_pin(density_file, grid_type, tmp_path):
    require_gpu()
    from horton_part_amd import io

    _say(f"case: part-dens MBIS water, grid_type {grid_type}")
    fn = str(tmp_path / "part.npz")
    io.partition_density(density_file, fn, type="mbis", grid_type=grid_type)
    out = np.load(fn)
    q = out["charges"]
    _say(f"  charges {q} niter {int(out['niter'])}")
    assert np.abs(q - MBIS_PIN).max() < 1e-2
    assert abs(q.sum()) < 1e-3
    for k in ("natom", "atnums", "type", "charges", "niter", "history_charges", "history_propars",
              "history_entropies", "history_changes", "time_update_at_weights"):
        assert k in out.files


@pytest.mark.parametrize("method,grid_type", [("lisa", 1), ("lisa", 2), ("glisa", 1), ("glisa", 3)])
def test_part_dens_lisa_glisa_water(density_file, method, grid_type, tmp_path):
    require_gpu()
    from horton_part_amd import io

    _say(f"case: part-dens {method} sc water, grid_type {grid_type}")
    fn = str(tmp_path / "part.npz")
    io.partition_density(density_file, fn, type=method, solver="sc", grid_type=grid_type)
    q = np.load(fn)["charges"]
    _say(f"  charges {q}")
    assert q.sum() == pytest.approx(0.0, abs=1e-4)
    assert q[0] < 0 and q[1] > 0 and q[2] > 0
