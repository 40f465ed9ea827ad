"""Density and partition files of the reference's command-line programs.

``part-gen`` (``scripts/generate_density.py:207-252``) writes a molecular
grid and the density on it to an ``.npz`` file; ``part-dens``
(``scripts/partition_density.py:65-85, 242-357``) reads it back, runs a
partitioning scheme and writes the results to another ``.npz`` file.  This
module reads and writes the same keys, so files move between the two
implementations:

density file
    ``atcoords``, ``atnums``, ``atcorenums``, ``points``, ``weights``,
    ``aim_weights``, ``cellvecs``, ``nelec``, ``density``, ``atom_idxs`` and
    per atom ``atom{i}/points``, ``atom{i}/weights``, ``atom{i}/shell_idxs``,
    ``atom{i}/rgrid/points``, ``atom{i}/rgrid/weights``;
partition file
    ``natom``, ``atnums``, ``atcorenums``, ``type``, ``lmax``, ``maxiter``,
    ``threshold``, ``inner_threshold``, ``solver``, ``charges``, ``time``,
    ``time_update_at_weights``, ``time_update_propars``, ``niter`` and the
    ``history_*`` arrays.

The reference computes the density from a wavefunction file with iodata and
gbasis (``generate_density.py:119-204``: ``rho = sum_ab D_ab phi_a phi_b``).
``load_fchk`` and ``basis_density`` restate the part of that used here: a
Gaussian formatted checkpoint file with s, p, sp and Cartesian d shells of
normalised primitives, and its total SCF density matrix.  ``part_gen_grid``
builds the grid ``part-gen`` builds (``generate_density.py:81-116``):
Gauss-Chebyshev points mapped by Becke's transform ``r = R (1 + x)/(1 - x) +
rmin`` (rmin 1e-4, R 1.5) times one Lebedev rule per atom, with Becke weights.
The atom grids of a density file are rebuilt from the points and weights it
stores, so no angular rule has to be regenerated in the same point order.
"""

import time

import numpy as np

from .grid import BeckeWeights, MolGrid, OneDGrid
from .lebedev import lebedev_rule

__all__ = [
    "FileAtomGrid",
    "load_density",
    "save_density",
    "load_fchk",
    "basis_density",
    "part_gen_grid",
    "generate_density",
    "partition_density",
]


class FileAtomGrid:
    """An atom grid given by its points and weights (shell-major), its radial
    rule and its shell boundaries, as a density file stores them."""

    def __init__(self, rgrid, points, weights, shell_idxs, center):
        self._rgrid = rgrid
        self._points = np.ascontiguousarray(points, dtype=float)
        self._weights = np.ascontiguousarray(weights, dtype=float)
        self._indices = np.asarray(shell_idxs, dtype=np.int64)
        self._center = np.asarray(center, dtype=float)
        if self._indices[-1] != len(self._points) or len(self._indices) != rgrid.size + 1:
            raise ValueError("shell_idxs must split the points into one shell per radial point")
        r = rgrid.points
        per_shell = 4 * np.pi * r**2 * rgrid.weights
        with np.errstate(divide="ignore", invalid="ignore"):
            sph = self._weights / np.repeat(per_shell, np.diff(self._indices))
        self._sph_weights = np.where(np.isfinite(sph), sph, 0.0)

    rgrid = property(lambda self: self._rgrid)
    points = property(lambda self: self._points)
    weights = property(lambda self: self._weights)
    indices = property(lambda self: self._indices)
    center = property(lambda self: self._center)
    sph_weights = property(lambda self: self._sph_weights)
    size = property(lambda self: len(self._points))

    def integrate(self, *value_arrays):
        out = self._weights
        for v in value_arrays:
            out = out * v
        return float(np.sum(out))


class _StoredMolGrid:
    """A molecular grid read from a density file: its atom grids, points and
    molecular weights as stored."""

    def __init__(self, atgrids, points, weights, aim_weights, indices, atcoords, atnums):
        self.atgrids = atgrids
        self.points = np.ascontiguousarray(points, dtype=float)
        self.weights = np.ascontiguousarray(weights, dtype=float)
        self.aim_weights = np.asarray(aim_weights, dtype=float)
        self.indices = np.asarray(indices, dtype=np.int64)
        self.atcoords = np.asarray(atcoords, dtype=float)
        self.atnums = np.asarray(atnums)
        self.size = len(self.points)

    def get_atomic_grid(self, index):
        return self.atgrids[index]

    def integrate(self, *value_arrays):
        out = self.weights
        for v in value_arrays:
            out = out * v
        return float(np.sum(out))


def load_density(fn):
    """(grid, data) of a density file (``partition_density.py:65-85``)."""
    with np.load(fn) as f:
        data = {k: f[k] for k in f.files}
    natom = len(data["atnums"])
    atgrids = []
    for i in range(natom):
        rgrid = OneDGrid(data[f"atom{i}/rgrid/points"], data[f"atom{i}/rgrid/weights"])
        atgrids.append(
            FileAtomGrid(rgrid, data[f"atom{i}/points"], data[f"atom{i}/weights"], data[f"atom{i}/shell_idxs"],
                         data["atcoords"][i])
        )
    grid = _StoredMolGrid(atgrids, data["points"], data["weights"], data["aim_weights"], data["atom_idxs"],
                          data["atcoords"], data["atnums"])
    return grid, data


def save_density(fn, grid, density, atcoords, atnums, atcorenums, nelec, **extra):
    """Write a density file with the keys of ``generate_density.py:232-252``."""
    data = {
        "density": np.asarray(density, dtype=float),
        "atcoords": np.asarray(atcoords, dtype=float),
        "atnums": np.asarray(atnums),
        "atcorenums": np.asarray(atcorenums, dtype=float),
        "points": grid.points,
        "weights": grid.weights,
        "aim_weights": grid.aim_weights,
        "cellvecs": np.zeros((0, 3)),
        "nelec": nelec,
        "atom_idxs": np.asarray(grid.indices),
    }
    for i, atgrid in enumerate(grid.atgrids):
        data[f"atom{i}/points"] = atgrid.points
        data[f"atom{i}/weights"] = atgrid.weights
        data[f"atom{i}/shell_idxs"] = np.asarray(atgrid.indices)
        data[f"atom{i}/rgrid/points"] = atgrid.rgrid.points
        data[f"atom{i}/rgrid/weights"] = atgrid.rgrid.weights
    data.update(extra)
    np.savez_compressed(fn, **data)


# ---------------------------------------------------------------------------
# formatted checkpoint files and the density of a Gaussian basis
# ---------------------------------------------------------------------------


def load_fchk(fn):
    """The fields of a Gaussian formatted checkpoint file, keyed by their
    titles: integers and reals as scalars or numpy arrays."""
    fields = {}
    with open(fn) as f:
        lines = f.read().splitlines()
    i = 2
    while i < len(lines):
        line = lines[i]
        i += 1
        if len(line) < 44 or line[43] not in "IRC":
            continue
        title = line[:43].strip()
        kind = line[43]
        rest = line[44:].split()
        conv = int if kind == "I" else (float if kind == "R" else str)
        if rest and rest[0] == "N=":
            n = int(rest[1])
            values = []
            while len(values) < n and i < len(lines):
                values += lines[i].split()
                i += 1
            fields[title] = np.array([conv(v) for v in values[:n]]) if kind != "C" else " ".join(values)
        elif rest:
            fields[title] = conv(rest[0])
    return fields


def _double_factorial(n):
    return 1.0 if n <= 0 else float(np.prod(np.arange(n, 0, -2)))


# Cartesian components in Gaussian's order
_CART = {0: [(0, 0, 0)], 1: [(1, 0, 0), (0, 1, 0), (0, 0, 1)],
         2: [(2, 0, 0), (0, 2, 0), (0, 0, 2), (1, 1, 0), (1, 0, 1), (0, 1, 1)]}


def _fchk_basis(fields):
    """Contracted shells (centre, [(powers, exponents, coefficients)]) in
    basis-function order; shell type -1 is an sp shell."""
    types = np.atleast_1d(fields["Shell types"])
    nprim = np.atleast_1d(fields["Number of primitives per shell"])
    exps = np.atleast_1d(fields["Primitive exponents"])
    coefs = np.atleast_1d(fields["Contraction coefficients"])
    pcoefs = np.atleast_1d(fields.get("P(S=P) Contraction coefficients", np.zeros_like(coefs)))
    centres = np.atleast_1d(fields["Coordinates of each shell"]).reshape(-1, 3)
    funcs = []
    k = 0
    for s, t in enumerate(types):
        sl = slice(k, k + int(nprim[s]))
        k += int(nprim[s])
        if t == -1:
            parts = [(0, coefs[sl]), (1, pcoefs[sl])]
        elif t in (0, 1, 2):
            parts = [(int(t), coefs[sl])]
        else:
            raise NotImplementedError(f"shell type {t}: only s, p, sp and Cartesian d shells are read")
        for l, c in parts:
            for pw in _CART[l]:
                funcs.append((centres[s], pw, exps[sl], c))
    return funcs


def basis_density(fields, points, chunk=20000):
    """rho(p) = sum_ab D_ab phi_a(p) phi_b(p) with the total SCF density
    matrix of the file (``generate_density.py:176-204``); primitives
    normalised as x^a y^b z^c exp(-alpha r^2) with norm
    (2 alpha/pi)^(3/4) (4 alpha)^(l/2) / sqrt((2a-1)!! (2b-1)!! (2c-1)!!)."""
    funcs = _fchk_basis(fields)
    nb = len(funcs)
    packed = np.atleast_1d(fields["Total SCF Density"])
    D = np.zeros((nb, nb))
    D[np.tril_indices(nb)] = packed
    D = D + D.T - np.diag(np.diag(D))
    points = np.asarray(points, dtype=float)
    rho = np.empty(len(points))
    for s in range(0, len(points), chunk):
        p = points[s : s + chunk]
        phi = np.empty((nb, len(p)))
        for b, (c, pw, ex, co) in enumerate(funcs):
            d = p - c
            r2 = np.einsum("ij,ij->i", d, d)
            l = sum(pw)
            norm = (2 * ex / np.pi) ** 0.75 * (4 * ex) ** (l / 2) / np.sqrt(
                _double_factorial(2 * pw[0] - 1) * _double_factorial(2 * pw[1] - 1) * _double_factorial(2 * pw[2] - 1))
            radial = np.exp(-np.outer(r2, ex)) @ (co * norm)
            phi[b] = radial * d[:, 0] ** pw[0] * d[:, 1] ** pw[1] * d[:, 2] ** pw[2]
        rho[s : s + chunk] = np.einsum("ab,ap,bp->p", D, phi, phi, optimize=True)
    return rho


def part_gen_grid(atnums, atcoords, nrad=150, nang=194):
    """The grid of ``generate_density.py:81-116``: nrad Gauss-Chebyshev points
    x_i = cos((2i-1) pi / (2 nrad)) with weights (pi / nrad) sqrt(1 - x_i^2),
    mapped to r = 1.5 (1 + x) / (1 - x) + 1e-4 (dr/dx = 3 / (1 - x)^2), in
    increasing r, times the nang-point Lebedev rule, with Becke weights of
    order 3 (radii of He, Ne, Ar, Kr and Xe set to 0.5, 1, 2, 2.5 and 3.5
    angstrom as there)."""
    from .grid import AtomGrid
    from .utils import ANGSTROM, radius_becke

    i = np.arange(1, nrad + 1)
    x = np.cos((2 * i - 1) * np.pi / (2 * nrad))[::-1]
    w = (np.pi / nrad) * np.sqrt(1 - x**2)
    r = 1.5 * (1 + x) / (1 - x) + 1e-4
    rgrid = OneDGrid(r, w * 3.0 / (1 - x) ** 2)
    radii = {z: (v if v is not None else 1.0) for z, v in radius_becke.items()}
    for z, v in ((2, 0.5), (10, 1.0), (18, 2.0), (36, 2.5), (54, 3.5)):
        radii[z] = v * ANGSTROM
    lebedev_rule(nang)  # rejects unknown sizes early
    atgrids = [AtomGrid(rgrid, sizes=[nang], center=c) for c in np.asarray(atcoords, float)]
    return MolGrid(atnums, atgrids, BeckeWeights(radii=radii, order=3))


def generate_density(fn_in, fn_out, nrad=150, nang=194):
    """``part-gen`` for one formatted checkpoint file: the grid, the density
    on it and the density file (``generate_density.py:207-252``)."""
    fields = load_fchk(fn_in)
    atnums = np.atleast_1d(fields["Atomic numbers"]).astype(np.int64)
    atcoords = np.atleast_1d(fields["Current cartesian coordinates"]).reshape(-1, 3)
    atcorenums = np.atleast_1d(fields.get("Nuclear charges", atnums)).astype(float)
    grid = part_gen_grid(atnums, atcoords, nrad, nang)
    rho = basis_density(fields, grid.points)
    save_density(fn_out, grid, rho, atcoords, atnums, atcorenums, int(fields["Number of electrons"]))
    return grid, rho


# options of part-dens that go to each scheme class (data/keywords.yaml)
_COMMON = ("maxiter", "threshold", "inner_threshold", "lmax", "density_cutoff", "population_cutoff", "negative_cutoff")
_CLASS_ARGS = {
    "mbis": _COMMON + ("grid_type",),
    "gisa": _COMMON + ("solver", "solver_options", "grid_type"),
    "lisa": _COMMON + ("basis_func", "solver", "solver_options", "grid_type", "basis_type"),
    "glisa": ("maxiter", "threshold", "lmax", "basis_func", "solver", "solver_options", "grid_type", "density_cutoff",
              "population_cutoff", "negative_cutoff", "basis_type"),
    "nlis": _COMMON + ("exp_n_dict", "nshell_dict", "grid_type"),
    "gmbis": _COMMON + ("exp_n_dict", "grid_type"),
}
# defaults of data/part-dens.yaml
_DEFAULTS = dict(basis_func="gauss", func_file=None, maxiter=1000, threshold=1e-6, inner_threshold=1e-8, lmax=3,
                 solver="sc", exp_n_dict=None, nshell_dict=None, grid_type=1)


def partition_density(fn_in, fn_out, type="lisa", part_job_type="do_partitioning", **settings):
    """``part-dens`` for one density file (``partition_density.py:242-357``):
    build the scheme object of ``type`` from the file with the options its
    class takes, run ``part_job_type`` and write the partition file.  Returns
    the scheme object."""
    from .utils import wpart_schemes

    if type not in _CLASS_ARGS:
        raise NotImplementedError(f"part-dens type {type!r} is not provided")
    opts = dict(_DEFAULTS)
    opts.update(settings)
    grid, data = load_density(fn_in)
    kwargs = dict(coordinates=data["atcoords"], numbers=data["atnums"].astype(np.int64),
                  pseudo_numbers=data["atcorenums"], grid=grid, moldens=data["density"])
    for k in _CLASS_ARGS[type]:
        if k in opts and opts[k] is not None:
            kwargs[k] = opts[k]
    if "basis_func" in _CLASS_ARGS[type]:
        kwargs["basis_func"] = opts.get("func_file") or opts.get("basis_func")
    if "exp_n_dict" in _CLASS_ARGS[type]:
        # prepare_exp_n_dict, partition_density.py:52-62
        kwargs["exp_n_dict"] = {(int(z), k): float(v) for z, vals in (opts.get("exp_n_dict") or {}).items()
                                for k, v in enumerate(vals)}
    part = wpart_schemes(type)(**kwargs)
    t0 = time.time()
    getattr(part, part_job_type)()
    elapsed = time.time() - t0
    out = {
        "natom": len(data["atnums"]),
        "atnums": data["atnums"],
        "atcorenums": data["atcorenums"],
        "type": type,
        "lmax": opts["lmax"],
        "maxiter": opts["maxiter"],
        "threshold": opts["threshold"],
        "inner_threshold": opts["inner_threshold"] if type != "glisa" else np.nan,
        "solver": opts["solver"],
        "charges": part["charges"],
        "time": elapsed,
        "time_update_at_weights": part["time_update_at_weights"],
        "time_update_propars": part["time_update_propars"],
        "niter": part["niter"],
        "history_charges": part["history_charges"],
        "history_propars": part["history_propars"],
        "history_entropies": part["history_entropies"],
        "history_changes": part["history_changes"],
    }
    np.savez_compressed(fn_out, **out)
    return part
