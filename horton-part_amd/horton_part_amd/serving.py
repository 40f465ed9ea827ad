"""Continuous-batching partitioning server on one GPU.

The reference runs one ``do_partitioning()`` per WPart object, one molecule
after another (scripts/program.py:120-131).  ``PartitionServer`` keeps many
of those partitionings resident on the device at once (``hp_batch_*`` in
``include/hortonpart.h``): requests are admitted while the GPU works, as
molecule slots and arena space free up; every step advances all molecules
in flight by a few outer iterations with shared kernel launches, and each
molecule leaves on its own convergence test.  Each result is the one
``do_partitioning()`` of the same scheme object gives with the same
precision.

Precision per request: "fp64", "mixed" (FP32 grid arithmetic, FP64 sums,
solves and polish), or "auto": mixed for GISA requests and FP64 for MBIS
requests.  MBIS's inner fixed point turns FP32 noise in the spherical
averages into small jumps of its iteration counts, and on slowly converging
molecules those can move the outer loop's stop by an iteration (charges then
differ from the FP64 result by up to a few 1e-6 e); GISA's QP has no such
switch, and its mixed results stay within 1e-6 e of FP64.

Usage::

    server = PartitionServer(device=0)
    t = server.submit(MBISWPart(...), precision="auto")
    while server.pending:
        server.step()
        for done in server.poll():
            print(done.ticket, done.charges)
"""

import ctypes
import time

import numpy as np

from . import _native
from .utils import DENSITY_CUTOFF, NEGATIVE_CUTOFF, POPULATION_CUTOFF

__all__ = ["PartitionServer", "Completed", "request_arrays"]

_EVENT = {_native.BATCH_CONVERGED: "converged", _native.BATCH_MAXITER: "maxiter", _native.BATCH_ERROR: "error"}


class Completed:
    """A finished request: parameters, charges and how the loop ended."""

    def __init__(self, ticket, tag, propars, charges, res, t_submit, t_done):
        self.ticket = ticket
        self.tag = tag
        self.propars = propars
        self.charges = charges
        self.event = _EVENT.get(res.event, "unknown")
        self.niter = int(res.niter)
        self.status = int(res.status)
        self.status_last = int(res.status_last)
        self.change = float(res.change)
        self.entropy = float(res.entropy)
        self.t_submit = t_submit
        self.t_done = t_done

    @property
    def latency(self):
        return self.t_done - self.t_submit


def request_arrays(part):
    """(molecule arrays, scheme arrays, initial parameters) of a scheme object
    that has not been partitioned yet: what ``submit`` sends to the device."""
    return part._molecule_arrays(), part._scheme_arrays(), np.ascontiguousarray(part._init_propars(), dtype=float)


def resolve_precision(precision, scheme):
    """The precision a request runs in: "auto" is mixed for GISA, FP64 for
    MBIS (``scheme`` is the ``hp_scheme_desc`` scheme number)."""
    if precision == "auto":
        return "mixed" if scheme == _native.SCHEME["gisa"] else "fp64"
    if precision not in _native.PREC:
        raise ValueError(f"unknown precision {precision!r}")
    return precision


class PartitionServer:
    """Owner of one ``hp_batch``: a device arena and its request queue.

    Capacities bound what is resident at once (points, atoms, radial
    shells, pro-atom functions, GISA QP entries, molecules); a request that
    does not fit waits in the queue (first in, first out) until enough
    molecules have left.  The inner-solver settings are the batch's and must
    match every request's.
    """

    def __init__(
        self,
        device=0,
        point_capacity=64_000_000,
        atom_capacity=200_000,
        shell_capacity=2_000_000,
        function_capacity=1_000_000,
        qp_capacity=4_000_000,
        max_inflight=256,
        max_molecule_points=None,
        ticks_per_step=4,
        use_graphs=True,
        inner_threshold=1e-8,
        inner_maxiter=2000,
        density_cutoff=DENSITY_CUTOFF,
        negative_cutoff=NEGATIVE_CUTOFF,
        population_cutoff=POPULATION_CUTOFF,
        groups=0,
        inner_budget=0,
    ):
        self.lib = _native.load()
        cfg = _native.BatchConfig()
        cfg.point_capacity = int(point_capacity)
        cfg.atom_capacity = int(atom_capacity)
        cfg.shell_capacity = int(shell_capacity)
        cfg.function_capacity = int(function_capacity)
        cfg.max_inflight = int(max_inflight)
        cfg.qp_capacity = int(qp_capacity)
        cfg.max_molecule_points = int(max_molecule_points or min(point_capacity, 8_000_000))
        cfg.ticks_per_step = int(ticks_per_step)
        cfg.use_graphs = int(bool(use_graphs))
        cfg.inner_threshold = float(inner_threshold)
        cfg.inner_maxiter = int(inner_maxiter)
        cfg.density_cutoff = float(density_cutoff)
        cfg.negative_cutoff = float(negative_cutoff)
        cfg.population_cutoff = float(population_cutoff)
        cfg.groups = int(groups)
        cfg.inner_budget = int(inner_budget)
        self.config = cfg
        handle = ctypes.c_void_p()
        _native.check(self.lib.hp_batch_create(int(device), ctypes.byref(cfg), ctypes.byref(handle)))
        self.handle = handle
        self._live = {}  # ticket -> (arrays kept alive, outputs, tag, t_submit)
        self.n_inflight = 0
        self.n_queued = 0

    # -- requests ---------------------------------------------------------------
    def submit(self, part, precision="auto", n_polish=1, tag=None):
        """Queue the partitioning of a scheme object (not yet partitioned)."""
        mol, sch, p0 = request_arrays(part)
        return self.submit_arrays(mol, sch, p0, precision=precision, n_polish=n_polish, tag=tag)

    def submit_arrays(self, mol_arrays, scheme_arrays, propars0, precision="auto", n_polish=1, tag=None):
        """Queue a request given as ``hp_molecule_desc``/``hp_scheme_desc``
        arrays (see ``request_arrays``).  The arrays are read when the request
        is admitted; they are kept referenced until its result is polled."""
        precision = resolve_precision(precision, scheme_arrays["scheme"])
        md = _native.MoleculeDesc()
        sd = _native.SchemeDesc()
        keep = []
        _fill(md, mol_arrays, keep)
        _fill(sd, scheme_arrays, keep)
        propars0 = np.ascontiguousarray(propars0, dtype=float)
        out_p = np.empty(propars0.size)
        out_q = np.empty(int(mol_arrays["natom"]))
        keep += [propars0, out_p, out_q, md, sd]
        ticket = ctypes.c_int64(0)
        _native.check(
            self.lib.hp_batch_submit(
                self.handle,
                ctypes.byref(md),
                ctypes.byref(sd),
                _native.ptr(propars0),
                _native.PREC[precision],
                int(n_polish),
                _native.ptr(out_p),
                _native.ptr(out_q),
                ctypes.byref(ticket),
            )
        )
        t = int(ticket.value)
        self._live[t] = (keep, out_p, out_q, tag, time.perf_counter())
        self.n_queued += 1
        return t

    # -- progress -----------------------------------------------------------------
    def step(self):
        """One host step (``hp_batch_step``); returns (in flight, queued)."""
        a = ctypes.c_int32(0)
        q = ctypes.c_int32(0)
        _native.check(self.lib.hp_batch_step(self.handle, ctypes.byref(a), ctypes.byref(q)))
        self.n_inflight, self.n_queued = int(a.value), int(q.value)
        return self.n_inflight, self.n_queued

    def poll(self, max_results=256):
        """Finished requests since the last poll (oldest first)."""
        buf = (_native.BatchResult * max_results)()
        n = ctypes.c_int32(0)
        out = []
        while True:
            _native.check(self.lib.hp_batch_poll(self.handle, buf, max_results, ctypes.byref(n)))
            now = time.perf_counter()
            for i in range(n.value):
                res = buf[i]
                keep, out_p, out_q, tag, t0 = self._live.pop(int(res.ticket))
                out.append(Completed(int(res.ticket), tag, out_p, out_q, res, t0, now))
            if n.value < max_results:
                return out

    def stats(self):
        """Work and GPU time so far (``hp_batch_get_stats``) as a dict."""
        st = _native.BatchStats()
        _native.check(self.lib.hp_batch_get_stats(self.handle, ctypes.byref(st)))
        return {name: getattr(st, name) for name, _ in st._fields_}

    @property
    def pending(self):
        """Requests submitted and not yet polled."""
        return len(self._live)

    def run_until_done(self, timeout=None):
        """Step until every submitted request has finished; returns them."""
        done = []
        t0 = time.perf_counter()
        while self.pending:
            self.step()
            done += self.poll()
            if timeout is not None and time.perf_counter() - t0 > timeout:
                raise TimeoutError(f"{self.pending} request(s) still pending")
        return done

    def close(self):
        if getattr(self, "handle", None) is not None and self.handle.value:
            self.lib.hp_batch_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _fill(desc, arrays, keep):
    for name, ctype in desc._fields_:
        if name not in arrays:
            continue
        v = arrays[name]
        if isinstance(v, np.ndarray):
            v = np.ascontiguousarray(v)
            keep.append(v)
            setattr(desc, name, v.ctypes.data_as(ctypes.POINTER(ctype._type_)))
        else:
            setattr(desc, name, v)
