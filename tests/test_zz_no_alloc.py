"""hp_part_moments_static (no allocation) against hp_part_moments.

The static variant runs the same kernels with their partial sums in a static
device workspace, so its moments equal the allocating entry point's bit for
bit, for lmax 3 (40 sums per row), 1 (a prefix of them) and 5 (one sum per
quantity).  Two molecules in a row, the second smaller, check that the workspace's
leftovers from the first call do not leak into the second; a molecule beyond
the workspace gets a clean HP_E_ARG.
"""

import ctypes
import sys
import warnings

import numpy as np
import pytest

from conftest import require_gpu

pytestmark = pytest.mark.gpu
warnings.simplefilter("ignore")


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _static_moments(dev, natom, lmax=3):
    from horton_part_amd import _native

    cart = np.full((natom, (lmax + 1) * (lmax + 2) * (lmax + 3) // 6), np.nan)
    pure = np.full((natom, (lmax + 1) ** 2), np.nan)
    rad = np.full((natom, lmax + 1), np.nan)
    rc = dev.lib.hp_part_moments_static(dev.handle, lmax, _native.ptr(cart), _native.ptr(pure), _native.ptr(rad))
    return rc, (cart, pure, rad)


@pytest.mark.parametrize("grid_type", [1, 3])
def test_moments_static_equals_allocating(grid_type):
    require_gpu()
    from horton_part_amd import MBISWPart, workloads

    for natom, seed, nrad, nang in ((6, 7, 40, 50), (3, 4, 30, 26)):
        _say(f"case: static moments, grid_type {grid_type}, chain {natom} atoms seed {seed}, {nrad}x{nang}")
        w = workloads.chain(natom=natom, seed=seed, nrad=nrad, nang=nang)
        part = MBISWPart(w.coordinates, w.numbers, w.pseudo_numbers, w.grid, w.density, maxiter=8, grid_type=grid_type)
        part.do_partitioning()
        dev = part._dev
        for lmax in (3, 1, 5):  # the 40-sum kernel, a prefix of it, the any-order kernel
            ref = dev.moments(lmax)
            rc, got = _static_moments(dev, natom, lmax)
            assert rc == 0
            for r, g in zip(ref, got):
                np.testing.assert_array_equal(g, r)
        part.close()


def test_moments_static_rejects_beyond_workspace():
    require_gpu()
    from horton_part_amd import MBISWPart, _native, workloads

    _say("case: static moments beyond the workspace (300 atoms x 100 radial shells)")
    w = workloads.cluster(natom=300, seed=3, nrad=100, nang=6, becke=False, device=0)
    part = MBISWPart(w.coordinates, w.numbers, w.pseudo_numbers, w.grid, w.density, maxiter=1)
    part.do_partitioning()
    rc, _ = _static_moments(part._dev, 300)
    assert rc == _native.HP_E_ARG
    assert "static workspace" in part._dev.lib.hp_last_error().decode()
    part.close()
