"""Seeded synthetic molecules, grids and densities of the benchmark workloads.

Every workload is fixed by a ``numpy.random.default_rng`` seed:

* ``water``   O at the origin, H at 0.9572 A along +x, the other H at
  104.52 degrees; 2 Slater density shells on O, 1 on each H.
* ``chain``   N atoms (C/H/N/O with probabilities .35/.45/.10/.10) placed by a
  random walk with steps of 1.0-1.6 A and at least 0.9 A between atoms;
  1-3 density shells per atom.
* ``cluster`` N atoms placed uniformly at random in a cube sized for
  0.1 atoms/A^3 by random sequential addition (>= 0.9 A apart).
* ``serving_stream`` a stream of partitioning requests: Poisson arrivals,
  atom counts log-uniform in [3, 300], MBIS/GISA 50/50, each molecule a
  ``chain`` with its own seed.

The density is ``rho(p) = sum_a sum_k N_ak a_ak^3/(8 pi) exp(-a_ak |p-R_a|)
* (1 + 0.02 u_p)`` with ``a_ak ~ U[1, 20]`` bohr^-1, ``N_ak > 0`` scaled so
that ``sum_k N_ak = Z_a``, and ``u_p ~ U(-1, 1)`` per grid point.  Grids are
ExpRTransform(5e-4, 20 bohr) radial rules times one Lebedev rule per atom,
with Becke molecular weights.
"""

import numpy as np

from .grid import BeckeWeights, MolGrid, exp_radial_grid
from .utils import ANGSTROM

__all__ = ["Workload", "water", "chain", "cluster", "density_shells", "eval_density", "serving_stream"]

Z_OF = {"H": 1, "C": 6, "N": 7, "O": 8, "S": 16}


class Workload:
    """Coordinates, numbers, grid, density (and the density shells used)."""

    def __init__(self, name, coordinates, numbers, grid, density, shells):
        self.name = name
        self.coordinates = coordinates
        self.numbers = numbers
        self.pseudo_numbers = numbers.astype(float)
        self.grid = grid
        self.density = density
        self.shells = shells

    @property
    def natom(self):
        return len(self.numbers)

    @property
    def npoint(self):
        return self.grid.size


def density_shells(rng, numbers, nshell_of):
    """Per atom (N_k, alpha_k) with alpha ~ U[1, 20] and sum N_k = Z."""
    shells = []
    for z in numbers:
        k = nshell_of(int(z))
        alpha = rng.uniform(1.0, 20.0, size=k)
        n = rng.uniform(0.2, 1.0, size=k)
        n = n / n.sum() * float(z)
        shells.append((n, alpha))
    return shells


def eval_density(points, coordinates, shells, rng, chunk=1 << 16):
    """The synthetic density with multiplicative noise (1 + 0.02 u)."""
    rho = np.zeros(len(points))
    for s in range(0, len(points), chunk):
        p = points[s : s + chunk]
        acc = np.zeros(len(p))
        for c, (n, alpha) in zip(coordinates, shells):
            r = np.linalg.norm(p - c, axis=1)
            for nk, ak in zip(n, alpha):
                acc += nk * ak**3 / (8 * np.pi) * np.exp(-ak * r)
        rho[s : s + chunk] = acc
    u = rng.uniform(-1.0, 1.0, size=len(points))
    return rho * (1.0 + 0.02 * u)


def eval_density_device(points, coordinates, shells, rng, device=0):
    """``eval_density`` with the shell sum evaluated on the GPU."""
    from . import _native

    off = np.concatenate([[0], np.cumsum([len(n) for n, _ in shells])]).astype(np.int32)
    coef = np.concatenate([n * a**3 / (8 * np.pi) for n, a in shells])
    alpha = np.concatenate([a for _, a in shells])
    rho = _native.slater_density(coordinates, off, coef, alpha, points, device=device)
    u = rng.uniform(-1.0, 1.0, size=len(points))
    return rho * (1.0 + 0.02 * u)


def _build(name, coords, numbers, nrad, nang, rng, nshell_of, becke=True, device=None):
    rgrid = exp_radial_grid(nrad)
    aim = BeckeWeights(device=device) if becke else np.ones(len(numbers) * nrad * nang)
    grid = MolGrid.from_size(numbers, coords, nang, rgrid, aim)
    shells = density_shells(rng, numbers, nshell_of)
    if device is None:
        rho = eval_density(grid.points, coords, shells, rng)
    else:
        rho = eval_density_device(grid.points, coords, shells, rng, device=device)
    return Workload(name, coords, numbers, grid, rho, shells)


def water(seed=1000, nrad=100, nang=194, device=None):
    """Water-like molecule with 2 shells on O and 1 on each H."""
    rng = np.random.default_rng(seed)
    d = 0.9572 * ANGSTROM
    ang = np.deg2rad(104.52)
    coords = np.array([[0.0, 0.0, 0.0], [d, 0.0, 0.0], [d * np.cos(ang), d * np.sin(ang), 0.0]])
    numbers = np.array([8, 1, 1], dtype=np.int64)
    return _build("water", coords, numbers, nrad, nang, rng, lambda z: 2 if z == 8 else 1, device=device)


def _pick_elements(rng, natom, probs):
    elems = list(probs)
    p = np.array([probs[e] for e in elems])
    return np.array([Z_OF[e] for e in rng.choice(elems, size=natom, p=p / p.sum())], dtype=np.int64)


def chain(natom=20, seed=1001, nrad=150, nang=302, probs=None, becke=True, device=None):
    """Random-walk organic-like molecule."""
    rng = np.random.default_rng(seed)
    probs = probs or {"C": 0.35, "H": 0.45, "N": 0.10, "O": 0.10}
    numbers = _pick_elements(rng, natom, probs)
    pos = np.zeros((natom, 3))
    n = 1
    while n < natom:
        v = rng.normal(size=3)
        v /= np.linalg.norm(v)
        cand = pos[n - 1] + v * rng.uniform(1.0, 1.6)
        # The acceptance test is min_q |cand - q| >= 0.9 with each distance
        # formed by np.linalg.norm.  The squared distances of all accepted
        # atoms at once decide it unless the nearest lies within 1e-12 of
        # the threshold, where the per-atom norms decide as before.
        d = pos[:n] - cand
        m = float(np.min(np.einsum("ij,ij->i", d, d)))
        if m > 0.81 * (1.0 + 1e-12):
            ok = True
        elif m < 0.81 * (1.0 - 1e-12):
            ok = False
        else:
            ok = min(np.linalg.norm(cand - q) for q in pos[:n]) >= 0.9
        if ok:
            pos[n] = cand
            n += 1
    coords = pos * ANGSTROM
    return _build(f"chain{natom}", coords, numbers, nrad, nang, rng, lambda z: int(rng.integers(1, 4)), becke=becke, device=device)


def cluster(natom=200, seed=1002, nrad=100, nang=194, probs=None, becke=True, density_per_a3=0.1, device=None):
    """Uniform random cluster at a fixed number density."""
    rng = np.random.default_rng(seed)
    probs = probs or {"C": 0.35, "H": 0.45, "N": 0.10, "O": 0.10}
    numbers = _pick_elements(rng, natom, probs)
    side = (natom / density_per_a3) ** (1.0 / 3.0)
    pos = []
    while len(pos) < natom:
        cand = rng.uniform(0.0, side, size=3)
        if not pos or np.min(np.linalg.norm(np.array(pos) - cand, axis=1)) >= 0.9:
            pos.append(cand)
    coords = np.array(pos) * ANGSTROM
    return _build(f"cluster{natom}", coords, numbers, nrad, nang, rng, lambda z: int(rng.integers(1, 4)), becke=becke, device=device)


def serving_stream(seed=1004, nreq=4096, rate=200.0, amin=3, amax=300):
    """Requests of the serving stream, in arrival order.

    Each entry: ``t`` arrival time in seconds (Poisson process of ``rate``
    per second, the first at 0), ``natom`` (log-uniform in [amin, amax],
    rounded), ``scheme`` ("mbis" or "gisa" with probability 1/2) and
    ``mol_seed``, the seed of its ``chain`` molecule.
    """
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, size=nreq)
    t = np.concatenate([[0.0], np.cumsum(gaps[1:])])
    natom = np.clip(np.rint(np.exp(rng.uniform(np.log(amin), np.log(amax), size=nreq))), amin, amax).astype(int)
    scheme = np.where(rng.random(nreq) < 0.5, "mbis", "gisa")
    mol_seed = rng.integers(0, 2**31 - 1, size=nreq)
    return [
        {"t": float(t[i]), "natom": int(natom[i]), "scheme": str(scheme[i]), "mol_seed": int(mol_seed[i])}
        for i in range(nreq)
    ]
