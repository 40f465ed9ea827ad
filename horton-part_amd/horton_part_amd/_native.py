"""ctypes binding of ``libhortonpart.so`` (declared in ``include/hortonpart.h``).

The library is the only compute path of this package: there is no CPU
fallback.  ``load()`` raises ``RuntimeError`` when the shared library is
missing or fails to load, and the scheme classes call it before any work.
"""

import ctypes
import os
import threading
from pathlib import Path

import numpy as np

__all__ = ["load", "lib_path", "HPError", "EXPORTED_SYMBOLS", "MoleculeDesc", "SchemeDesc", "BatchConfig", "BatchResult"]

HP_E_ARG = -1
HP_E_HIP = -2
HP_E_SOLVER = -3
HP_E_UNSUPP = -4

SCHEME = {"mbis": 0, "gisa": 1, "lisa": 2, "nlis": 3, "isa": 4, "hi": 5, "hirshfeld": 6}
SOLVER = {"mbis": 0, "qp": 1, "sc": 2, "newton": 3, "m-newton": 4}
PREC = {"fp64": 0, "mixed": 1}

ST_INNER_NOT_CONVERGED = 1
ST_POP_MISMATCH = 2
ST_NEG_DENSITY = 4
ST_LINESEARCH = 8
ST_NEWTON_MAXITER = 16
ST_QP_FAILED = 32
ST_NEG_PARS = 64
ST_NOT_MONOTONIC = 128
ST_SINGULAR = 256
ST_NONFINITE = 512
ST_NO_RECORD = 1024

KERNEL_WEIGHTS = 0
KERNEL_INNER = 1
KERNEL_LISA_GRID = 2
KERNEL_GLISA_HESS = 3

EXPORTED_SYMBOLS = [
    "hp_abi_version",
    "hp_last_error",
    "hp_device_count",
    "hp_part_create",
    "hp_part_destroy",
    "hp_part_set_propars",
    "hp_part_weights",
    "hp_part_run",
    "hp_part_results",
    "hp_part_history",
    "hp_part_status",
    "hp_part_spherical_averages",
    "hp_part_at_weights",
    "hp_part_promoldens",
    "hp_part_moments",
    "hp_part_moments_static",
    "hp_part_global_terms",
    "hp_part_global_proatoms",
    "hp_part_populations",
    "hp_part_kernel_time",
    "hp_part_get_work",
    "hp_batch_create",
    "hp_batch_destroy",
    "hp_batch_submit",
    "hp_batch_step",
    "hp_batch_poll",
    "hp_batch_get_stats",
    "hp_host_alloc",
    "hp_host_free",
    "hp_becke_weights",
    "hp_slater_density",
    "hp_stream_copy",
    "hp_exp_rate",
]

BATCH_CONVERGED = 1
BATCH_MAXITER = 2
BATCH_ERROR = 3

_P = ctypes.POINTER
_d = ctypes.c_double
_pd = _P(ctypes.c_double)
_pi32 = _P(ctypes.c_int32)
_pi64 = _P(ctypes.c_int64)


class MoleculeDesc(ctypes.Structure):
    _fields_ = [
        ("natom", ctypes.c_int),
        ("coordinates", _pd),
        ("numbers", _pi64),
        ("pseudo_numbers", _pd),
        ("npoint", ctypes.c_int64),
        ("points", _pd),
        ("density", _pd),
        ("mol_weights", _pd),
        ("atom_weights", _pd),
        ("sph_weights", _pd),
        ("nshell", ctypes.c_int),
        ("shell_atom", _pi32),
        ("shell_offsets", _pi64),
        ("shell_radius", _pd),
        ("shell_rweight", _pd),
    ]


class SchemeDesc(ctypes.Structure):
    _fields_ = [
        ("scheme", ctypes.c_int),
        ("solver", ctypes.c_int),
        ("grid_type", ctypes.c_int),
        ("fn_offsets", _pi32),
        ("fn_alpha", _pd),
        ("fn_order", _pd),
        ("fn_norm", _pd),
        ("qp_matrix", _pd),
        ("threshold", _d),
        ("inner_threshold", _d),
        ("maxiter", ctypes.c_int),
        ("inner_maxiter", ctypes.c_int),
        ("density_cutoff", _d),
        ("negative_cutoff", _d),
        ("population_cutoff", _d),
        ("cutoff_radius", _d),
        ("newton_tau", _d),
        ("newton_linspace", ctypes.c_int),
        ("check_mono", ctypes.c_int),
        ("mixed_switch", _d),
        ("mixed_refresh", ctypes.c_int),
        ("full_sum", ctypes.c_int),
        ("tab_offsets", _pi32),
        ("tab_knots", _pd),
        ("tab_rweights", _pd),
        ("rec_offsets", _pi32),
        ("rec_charges", _pi32),
        ("rec_rho", _pd),
        ("rec_coefs", _pd),
    ]


class PartWork(ctypes.Structure):
    _fields_ = [
        ("points", ctypes.c_int64),
        ("evals_per_pass", _d),
        ("bytes_per_pass", _d),
        ("hessian_passes", ctypes.c_int64),
        ("hessian_flops", _d),
        ("evals_done_fp32", _d),
        ("evals_done_fp64", _d),
    ]


class BatchConfig(ctypes.Structure):
    _fields_ = [
        ("point_capacity", ctypes.c_int64),
        ("atom_capacity", ctypes.c_int32),
        ("shell_capacity", ctypes.c_int32),
        ("function_capacity", ctypes.c_int32),
        ("max_inflight", ctypes.c_int32),
        ("qp_capacity", ctypes.c_int64),
        ("max_molecule_points", ctypes.c_int64),
        ("ticks_per_step", ctypes.c_int32),
        ("use_graphs", ctypes.c_int32),
        ("inner_threshold", _d),
        ("inner_maxiter", ctypes.c_int32),
        ("density_cutoff", _d),
        ("negative_cutoff", _d),
        ("population_cutoff", _d),
        ("groups", ctypes.c_int32),
        ("inner_budget", ctypes.c_int32),
    ]


class BatchResult(ctypes.Structure):
    _fields_ = [
        ("ticket", ctypes.c_int64),
        ("event", ctypes.c_int32),
        ("niter", ctypes.c_int32),
        ("status", ctypes.c_int32),
        ("status_last", ctypes.c_int32),
        ("change", _d),
        ("entropy", _d),
    ]


class BatchStats(ctypes.Structure):
    _fields_ = [
        ("ticks", ctypes.c_int64),
        ("iterations", ctypes.c_int64),
        ("completed", ctypes.c_int64),
        ("tick_ms", _d),
        ("weights_ms", _d),
        ("inner_ms", _d),
        ("points_fp32", _d),
        ("points_fp64", _d),
        ("evals_fp32", _d),
        ("evals_fp64", _d),
        ("done_fp32", _d),
        ("done_fp64", _d),
    ]


class HPError(RuntimeError):
    """Failure reported by the device library (message from hp_last_error)."""

    def __init__(self, code, message):
        super().__init__(f"[hortonpart {code}] {message}")
        self.code = code


def lib_path():
    env = os.environ.get("HORTONPART_LIB")
    if env:
        return Path(env)
    return Path(__file__).resolve().parent / "lib" / "libhortonpart.so"


_LIB = None


def load():
    """Load the device library once; raise ``RuntimeError`` if it is absent."""
    global _LIB
    if _LIB is not None:
        return _LIB
    path = lib_path()
    if not path.exists():
        raise RuntimeError(
            f"libhortonpart.so not found at {path}; build it with `python -c 'import __graft_entry__ as g; g.build()'`"
        )
    try:
        lib = ctypes.CDLL(str(path))
    except OSError as exc:  # pragma: no cover - depends on the ROCm runtime
        raise RuntimeError(f"cannot load {path}: {exc}") from exc
    lib.hp_abi_version.restype = ctypes.c_int
    lib.hp_last_error.restype = ctypes.c_char_p
    lib.hp_device_count.argtypes = [_pi32]
    lib.hp_part_create.argtypes = [ctypes.c_int, _P(MoleculeDesc), _P(SchemeDesc), _P(ctypes.c_void_p)]
    lib.hp_part_destroy.argtypes = [ctypes.c_void_p]
    lib.hp_part_set_propars.argtypes = [ctypes.c_void_p, _pd, ctypes.c_int64]
    lib.hp_part_weights.argtypes = [ctypes.c_void_p, _pd, _pd]
    lib.hp_part_run.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _pi32]
    lib.hp_part_results.argtypes = [ctypes.c_void_p, _pd, _pd, _pd, _pd]
    lib.hp_part_history.argtypes = [ctypes.c_void_p, _pd, _pd, _pd, _pd]
    lib.hp_part_status.argtypes = [ctypes.c_void_p, _pi32, _pi32, _pi32]
    lib.hp_part_spherical_averages.argtypes = [ctypes.c_void_p, _pd]
    lib.hp_part_at_weights.argtypes = [ctypes.c_void_p, ctypes.c_int, _pd, ctypes.c_int64]
    lib.hp_part_promoldens.argtypes = [ctypes.c_void_p, _pd, ctypes.c_int64]
    lib.hp_part_moments.argtypes = [ctypes.c_void_p, ctypes.c_int, _pd, _pd, _pd]
    lib.hp_part_moments_static.argtypes = [ctypes.c_void_p, ctypes.c_int, _pd, _pd, _pd]
    lib.hp_part_kernel_time.argtypes = [ctypes.c_void_p, ctypes.c_int, _pd, _pi64]
    lib.hp_part_global_terms.argtypes = [ctypes.c_void_p, _pd, ctypes.c_int, _pd, _pd, _pd, _pd]
    lib.hp_part_global_proatoms.argtypes = [ctypes.c_void_p, _pd, _pd, _pd, _pd]
    lib.hp_part_populations.argtypes = [ctypes.c_void_p, _pd]
    lib.hp_part_get_work.argtypes = [ctypes.c_void_p, _P(PartWork)]
    lib.hp_batch_create.argtypes = [ctypes.c_int, _P(BatchConfig), _P(ctypes.c_void_p)]
    lib.hp_batch_destroy.argtypes = [ctypes.c_void_p]
    lib.hp_batch_submit.argtypes = [ctypes.c_void_p, _P(MoleculeDesc), _P(SchemeDesc), _pd, ctypes.c_int, ctypes.c_int, _pd, _pd, _pi64]
    lib.hp_batch_step.argtypes = [ctypes.c_void_p, _pi32, _pi32]
    lib.hp_batch_poll.argtypes = [ctypes.c_void_p, _P(BatchResult), ctypes.c_int32, _pi32]
    lib.hp_batch_get_stats.argtypes = [ctypes.c_void_p, _P(BatchStats)]
    lib.hp_host_alloc.argtypes = [ctypes.c_int64, _P(ctypes.c_void_p)]
    lib.hp_stream_copy.argtypes = [ctypes.c_int, ctypes.c_int64, ctypes.c_int, _pd, _pd]
    lib.hp_exp_rate.argtypes = [ctypes.c_int, ctypes.c_int, _pd]
    lib.hp_host_free.argtypes = [ctypes.c_void_p]
    lib.hp_becke_weights.argtypes = [ctypes.c_int, ctypes.c_int, _pd, _pd, ctypes.c_int64, _pd, _pi64, _pd]
    lib.hp_slater_density.argtypes = [ctypes.c_int, ctypes.c_int, _pd, _pi32, _pd, _pd, ctypes.c_int64, _pd, _pd]
    for name in EXPORTED_SYMBOLS:
        if name not in ("hp_abi_version", "hp_last_error"):
            getattr(lib, name).restype = ctypes.c_int
    if lib.hp_abi_version() != 3:
        raise RuntimeError("libhortonpart ABI version mismatch")
    _LIB = lib
    return lib


def check(rc):
    if rc != 0:
        msg = _LIB.hp_last_error().decode() if _LIB is not None else "library not loaded"
        raise HPError(rc, msg)


def ptr(a, ctype=ctypes.c_double):
    """Pointer to a C-contiguous numpy array (None -> NULL)."""
    if a is None:
        return None
    return a.ctypes.data_as(_P(ctype))


def device_count():
    lib = load()
    n = ctypes.c_int32(0)
    rc = lib.hp_device_count(ctypes.byref(n))
    return int(n.value) if rc == 0 else 0


def becke_weights(coordinates, radii, points, atom_offsets, device=0):
    """Becke weights of every point for its owning atom, computed on the GPU."""
    lib = load()
    coordinates = np.ascontiguousarray(coordinates, dtype=float)
    radii = np.ascontiguousarray(radii, dtype=float)
    points = np.ascontiguousarray(points, dtype=float)
    atom_offsets = np.ascontiguousarray(atom_offsets, dtype=np.int64)
    out = np.empty(len(points))
    check(lib.hp_becke_weights(int(device), len(coordinates), ptr(coordinates), ptr(radii), len(points), ptr(points), ptr(atom_offsets, ctypes.c_int64), ptr(out)))
    return out


def slater_density(coordinates, fn_offsets, coef, alpha, points, device=0):
    """sum_a sum_k coef_ak exp(-alpha_ak r_a) at every point, on the GPU."""
    lib = load()
    coordinates = np.ascontiguousarray(coordinates, dtype=float)
    fn_offsets = np.ascontiguousarray(fn_offsets, dtype=np.int32)
    coef = np.ascontiguousarray(coef, dtype=float)
    alpha = np.ascontiguousarray(alpha, dtype=float)
    points = np.ascontiguousarray(points, dtype=float)
    out = np.empty(len(points))
    check(lib.hp_slater_density(int(device), len(coordinates), ptr(coordinates), ptr(fn_offsets, ctypes.c_int32), ptr(coef), ptr(alpha), len(points), ptr(points), ptr(out)))
    return out


class DevicePartition:
    """Owner of one ``hp_part`` handle: a molecule and scheme resident on a GPU."""

    def __init__(self, mol_arrays, scheme_arrays, device=0):
        self.lib = load()
        self._keep = []  # keep host arrays alive during create
        md = MoleculeDesc()
        sd = SchemeDesc()
        self._fill(md, mol_arrays)
        self._fill(sd, scheme_arrays)
        handle = ctypes.c_void_p()
        check(self.lib.hp_part_create(int(device), ctypes.byref(md), ctypes.byref(sd), ctypes.byref(handle)))
        self._keep = []
        self.handle = handle
        self.natom = mol_arrays["natom"]
        self.nshell = mol_arrays["nshell"]
        self.npoint = mol_arrays["npoint"]
        self.npar = scheme_arrays["npar"]
        self.niter = 0

    def _fill(self, desc, arrays):
        for name, ctype in desc._fields_:
            if name not in arrays:
                continue
            v = arrays[name]
            if isinstance(v, np.ndarray):
                v = np.ascontiguousarray(v)
                self._keep.append(v)
                base = ctype._type_
                setattr(desc, name, v.ctypes.data_as(_P(base)))
            else:
                setattr(desc, name, v)

    def close(self):
        if getattr(self, "handle", None) is not None and self.handle.value:
            self.lib.hp_part_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_propars(self, propars):
        propars = np.ascontiguousarray(propars, dtype=float)
        check(self.lib.hp_part_set_propars(self.handle, ptr(propars), propars.size))

    def weights(self, spherical=True):
        """One FP64 weights pass at the current parameters (``hp_part_weights``):
        (spherical averages per radial shell or None, entropy)."""
        rhobar = np.empty(self.nshell) if spherical else None
        entropy = ctypes.c_double()
        check(self.lib.hp_part_weights(self.handle, ptr(rhobar) if spherical else None, ctypes.byref(entropy)))
        return rhobar, entropy.value

    def run(self, precision="fp64", n_polish=1):
        n = ctypes.c_int32(0)
        rc = self.lib.hp_part_run(self.handle, PREC[precision], int(n_polish), ctypes.byref(n))
        self.niter = int(n.value)
        return rc

    def results(self):
        propars = np.empty(self.npar)
        charges = np.empty(self.natom)
        change = ctypes.c_double()
        entropy = ctypes.c_double()
        check(self.lib.hp_part_results(self.handle, ptr(propars), ptr(charges), ctypes.byref(change), ctypes.byref(entropy)))
        return propars, charges, change.value, entropy.value

    def history(self):
        n = self.niter
        hp = np.empty((n, self.npar))
        hc = np.empty((n, self.natom))
        he = np.empty(n)
        hg = np.empty(n)
        check(self.lib.hp_part_history(self.handle, ptr(hp), ptr(hc), ptr(he), ptr(hg)))
        return hp, hc, he, hg

    def status(self):
        it = np.zeros(self.natom, np.int32)
        last = np.zeros(self.natom, np.int32)
        allst = np.zeros(self.natom, np.int32)
        check(self.lib.hp_part_status(self.handle, ptr(it, ctypes.c_int32), ptr(last, ctypes.c_int32), ptr(allst, ctypes.c_int32)))
        return it, last, allst

    def spherical_averages(self):
        out = np.empty(self.nshell)
        check(self.lib.hp_part_spherical_averages(self.handle, ptr(out)))
        return out

    def at_weights(self, atom, size):
        out = np.empty(size)
        check(self.lib.hp_part_at_weights(self.handle, int(atom), ptr(out), size))
        return out

    def promoldens(self):
        out = np.empty(self.npoint)
        check(self.lib.hp_part_promoldens(self.handle, ptr(out), self.npoint))
        return out

    def moments(self, lmax=3):
        lmax = int(lmax)
        if lmax < 0:
            raise ValueError("lmax must be >= 0")
        cart = np.empty((self.natom, (lmax + 1) * (lmax + 2) * (lmax + 3) // 6))
        pure = np.empty((self.natom, (lmax + 1) ** 2))
        rad = np.empty((self.natom, lmax + 1))
        check(self.lib.hp_part_moments(self.handle, int(lmax), ptr(cart), ptr(pure), ptr(rad)))
        return cart, pure, rad

    def global_terms(self, coefs, nderiv=2):
        """gLISA's objective, gradient, Hessian and promolecule minimum at
        ``coefs`` (``hp_part_global_terms``)."""
        coefs = np.ascontiguousarray(coefs, dtype=float)
        m = coefs.size
        f = ctypes.c_double()
        mn = ctypes.c_double()
        grad = np.empty(m) if nderiv >= 1 else None
        hess = np.empty((m, m)) if nderiv == 2 else None
        check(self.lib.hp_part_global_terms(self.handle, ptr(coefs), int(nderiv), ctypes.byref(f), ptr(grad), ptr(hess),
                                            ctypes.byref(mn)))
        return f.value, grad, hess, mn.value

    def global_proatoms(self, coefs_new, coefs_old):
        """Per atom: minimum of its pro-atom at ``coefs_new`` and the squared
        change against ``coefs_old`` on the molecular grid."""
        cn = np.ascontiguousarray(coefs_new, dtype=float)
        co = np.ascontiguousarray(coefs_old, dtype=float)
        mn = np.empty(self.natom)
        msd = np.empty(self.natom)
        check(self.lib.hp_part_global_proatoms(self.handle, ptr(cn), ptr(co), ptr(mn), ptr(msd)))
        return mn, msd

    def populations(self):
        """AIM populations from the last weights pass: on the molecular grid
        (grid_type 2, 3), or on the atom grids for the tabulated schemes."""
        out = np.empty(self.natom)
        check(self.lib.hp_part_populations(self.handle, ptr(out)))
        return out

    def work(self):
        """Algorithmic work counts (``hp_part_get_work``) as a dict."""
        w = PartWork()
        check(self.lib.hp_part_get_work(self.handle, ctypes.byref(w)))
        return {name: getattr(w, name) for name, _ in w._fields_}

    def kernel_time(self, kernel=KERNEL_WEIGHTS):
        ms = ctypes.c_double()
        n = ctypes.c_int64()
        check(self.lib.hp_part_kernel_time(self.handle, int(kernel), ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value


class PinnedArray(np.ndarray):
    """numpy array in page-locked host memory (freed with the array)."""

    def __array_finalize__(self, obj):
        if obj is not None and not hasattr(self, "_owner"):
            self._owner = getattr(obj, "_owner", None)


class _PinnedPool:
    """Page-locked host blocks kept for reuse.  Freeing page-locked memory
    waits for the device to go idle, which would drain the work queued ahead
    of a serving loop whenever a finished request drops its inputs; blocks
    therefore return here (size classes of an eighth of an octave, at least
    4 KiB) and are freed only while more than ``keep_bytes`` sit unused."""

    def __init__(self, keep_bytes=64 << 30):
        self._lock = threading.Lock()
        self._free = {}
        self._idle = 0
        self.keep_bytes = keep_bytes

    @staticmethod
    def size_class(nbytes):
        n = max(int(nbytes), 4096)
        step = 1 << max((n - 1).bit_length() - 3, 12)
        return -(-n // step) * step

    def get(self, nbytes):
        size = self.size_class(nbytes)
        with self._lock:
            blocks = self._free.get(size)
            if blocks:
                self._idle -= size
                return blocks.pop(), size
        p = ctypes.c_void_p()
        check(load().hp_host_alloc(size, ctypes.byref(p)))
        return p, size

    def put(self, ptr, size):
        with self._lock:
            if self._idle + size <= self.keep_bytes:
                self._free.setdefault(size, []).append(ptr)
                self._idle += size
                return
        load().hp_host_free(ptr)


_PINNED_POOL = _PinnedPool()


class _PinnedOwner:
    def __init__(self, nbytes):
        self.ptr, self.size = _PINNED_POOL.get(nbytes)

    def __del__(self):
        try:
            _PINNED_POOL.put(self.ptr, self.size)
        except Exception:
            pass


def pinned_copy(a):
    """A copy of ``a`` in page-locked host memory."""
    a = np.ascontiguousarray(a)
    owner = _PinnedOwner(max(a.nbytes, 1))
    buf = (ctypes.c_char * max(a.nbytes, 1)).from_address(owner.ptr.value)
    out = np.frombuffer(buf, dtype=a.dtype, count=a.size).reshape(a.shape).view(PinnedArray)
    out._owner = owner
    out[...] = a
    return out


def stream_copy(nbytes=1 << 30, reps=20, device=0):
    """(GB/s, ms per copy) of a plain device copy of ``nbytes``."""
    lib = load()
    g = ctypes.c_double()
    ms = ctypes.c_double()
    check(lib.hp_stream_copy(int(device), int(nbytes), int(reps), ctypes.byref(g), ctypes.byref(ms)))
    return g.value, ms.value


def exp_rate(precision="fp64", device=0):
    """Exponentials per second of the weights pass's exp in ``precision``."""
    lib = load()
    r = ctypes.c_double()
    check(lib.hp_exp_rate(int(device), PREC[precision], ctypes.byref(r)))
    return r.value
