"""The C ABI: every function declared in include/hortonpart.h is exported by
libhortonpart.so, and argument validation answers without a GPU."""

import ctypes
import os
import re

import numpy as np
import pytest

from horton_part_amd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "hortonpart.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(hp_\w+)\s*\(", text, re.M)))


def test_header_symbols_exported():
    lib = _native.load()
    names = _declared()
    assert len(names) >= 15
    for name in names:
        assert hasattr(lib, name), name
    assert sorted(_native.EXPORTED_SYMBOLS) == names


def test_abi_version_and_errors():
    lib = _native.load()
    assert lib.hp_abi_version() == 3
    n = ctypes.c_int32(-1)
    lib.hp_device_count(ctypes.byref(n))
    assert n.value >= 0


def _desc(**over):
    natom, nshell, L = 2, 3, 4
    npoint = nshell * L
    arrays = {
        "natom": natom,
        "coordinates": np.zeros((natom, 3)),
        "numbers": np.array([1, 1], np.int64),
        "pseudo_numbers": np.ones(natom),
        "npoint": npoint,
        "points": np.zeros((npoint, 3)),
        "density": np.ones(npoint),
        "mol_weights": np.ones(npoint),
        "atom_weights": np.ones(npoint),
        "sph_weights": np.full(npoint, 0.25),
        "nshell": nshell,
        "shell_atom": np.array([0, 0, 1], np.int32),
        "shell_offsets": np.arange(0, npoint + 1, L, dtype=np.int64),
        "shell_radius": np.ones(nshell),
        "shell_rweight": np.ones(nshell),
    }
    scheme = {
        "scheme": 0,
        "solver": 0,
        "grid_type": 1,
        "fn_offsets": np.array([0, 1, 2], np.int32),
        "threshold": 1e-6,
        "inner_threshold": 1e-8,
        "maxiter": 10,
        "inner_maxiter": 100,
        "density_cutoff": 1e-15,
        "negative_cutoff": -1e-12,
        "population_cutoff": 1e-4,
        "cutoff_radius": 0.0,
        "newton_tau": 1.0,
        "newton_linspace": 20,
        "check_mono": 0,
        "npar": 4,
    }
    for k, v in over.items():
        (scheme if k in scheme else arrays)[k] = v
    return arrays, scheme


@pytest.mark.parametrize(
    "over,code",
    [
        (dict(shell_atom=np.array([1, 0, 0], np.int32)), _native.HP_E_ARG),
        (dict(shell_offsets=np.array([0, 4, 8, 11], np.int64)), _native.HP_E_ARG),
        (dict(grid_type=4), _native.HP_E_ARG),
        (dict(grid_type=2, check_mono=1), _native.HP_E_UNSUPP),
        (dict(grid_type=3, check_mono=1), _native.HP_E_UNSUPP),
        (dict(solver=1), _native.HP_E_ARG),
        (dict(fn_offsets=np.array([0, 8, 9], np.int32)), _native.HP_E_UNSUPP),
        (dict(maxiter=0), _native.HP_E_ARG),
    ],
)
def test_create_rejects_bad_descriptions(over, code):
    """Validation happens before any device call, so it runs on the CPU."""
    arrays, scheme = _desc(**over)
    with pytest.raises(_native.HPError) as exc:
        _native.DevicePartition(arrays, scheme)
    assert exc.value.code == code


_TAB = dict(tab_offsets=np.array([0, 4, 8], np.int32), tab_knots=np.tile([0.1, 0.5, 1.0, 2.0], 2),
            tab_rweights=np.ones(8))
_REC = dict(rec_offsets=np.array([0, 1, 2], np.int32), rec_charges=np.zeros(2, np.int32), rec_rho=np.ones(8),
            rec_coefs=np.zeros(32))


@pytest.mark.parametrize(
    "over,code",
    [
        (dict(scheme=4), _native.HP_E_ARG),  # no tables
        (dict(scheme=4, grid_type=2, **_TAB), _native.HP_E_UNSUPP),  # ISA lives on atom grids
        (dict(scheme=4, **_TAB), _native.HP_E_ARG),  # ISA knots must be the atoms' shells
        (dict(scheme=5, **_TAB), _native.HP_E_ARG),  # Hirshfeld-I without records
        (dict(scheme=5, grid_type=3, **_TAB, **_REC), _native.HP_E_UNSUPP),
        (dict(scheme=5, **dict(_TAB, tab_knots=np.tile([0.1, 0.5, 0.4, 2.0], 2)), **_REC), _native.HP_E_ARG),
        (dict(scheme=6, **_TAB, **dict(_REC, rec_offsets=np.array([0, 2, 2], np.int32))), _native.HP_E_ARG),
        (dict(scheme=7), _native.HP_E_ARG),
    ],
)
def test_create_rejects_bad_tables(over, code):
    """Tabulated pro-atoms (ISA, Hirshfeld-I, Hirshfeld) are validated on the CPU."""
    arrays, scheme = _desc()
    scheme.pop("fn_offsets")
    scheme.update(over)
    with pytest.raises(_native.HPError) as exc:
        _native.DevicePartition(arrays, scheme)
    assert exc.value.code == code


def test_batch_config_validated_without_gpu():
    """hp_batch_create rejects empty capacities before touching a device."""
    lib = _native.load()
    cfg = _native.BatchConfig()
    cfg.point_capacity = 1000
    cfg.atom_capacity = 10
    cfg.shell_capacity = 10
    cfg.function_capacity = 10
    cfg.max_inflight = 0  # no slots
    cfg.qp_capacity = 0
    cfg.max_molecule_points = 1000
    cfg.ticks_per_step = 1
    handle = ctypes.c_void_p()
    rc = lib.hp_batch_create(0, ctypes.byref(cfg), ctypes.byref(handle))
    assert rc == _native.HP_E_ARG
    assert "capacities" in lib.hp_last_error().decode()
    assert not handle.value
    cfg.max_inflight = 8
    cfg.ticks_per_step = 0
    assert lib.hp_batch_create(0, ctypes.byref(cfg), ctypes.byref(handle)) == _native.HP_E_ARG
    # NULL arguments
    assert lib.hp_batch_step(None, None, None) == _native.HP_E_ARG
    n = ctypes.c_int32(0)
    assert lib.hp_batch_poll(None, None, 0, ctypes.byref(n)) == _native.HP_E_ARG
