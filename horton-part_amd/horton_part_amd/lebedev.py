"""Lebedev-Laikov angular quadrature on the unit sphere.

The reference takes its angular rules from qc-grid (``grid.LEBEDEV_DEGREES``,
used in ``core/stockholder.py:26,134`` and through ``MolGrid.from_size`` in
``tests/test_wpart.py:51``).  qc-grid is not available here, so the rules are
rebuilt from their octahedral orbit structure (V. I. Lebedev and D. N. Laikov,
Doklady Mathematics 59 (1999) 477).  Each rule is a list of orbits:

* ``a1``: the 6 points (+-1, 0, 0) and permutations;
* ``a2``: the 12 points (0, +-1/sqrt2, +-1/sqrt2) and permutations;
* ``a3``: the 8 points (+-1/sqrt3, +-1/sqrt3, +-1/sqrt3);
* ``bk``: 24 points (+-l, +-l, +-m), m = sqrt(1 - 2 l^2), and permutations;
* ``ck``: 24 points (+-p, +-q, 0), q = sqrt(1 - p^2), and permutations;
* ``dk``: 48 points (+-r, +-s, +-t), t = sqrt(1 - r^2 - s^2), all permutations.

The orbit parameters below solve the moment equations
``sum_j w_j x_j^{2a} y_j^{2b} z_j^{2c} = (2a-1)!!(2b-1)!!(2c-1)!!/(2a+2b+2c+1)!!``
for every ``2(a+b+c) <= degree`` to within 1e-16; the test suite re-checks
exactness on all real spherical harmonics up to the rule's degree.  Weights are
normalised to sum to one (qc-grid normalises to 4*pi; the reference only ever
uses the product ``4*pi*w`` through the atom-grid weights).
"""

import itertools

import numpy as np

__all__ = ["LEBEDEV_DEGREES", "LEBEDEV_NPOINTS", "lebedev_rule"]

# (orbit code, parameter 1, parameter 2, weight); weights sum to one.
_RULES = {
    6: (3, [("a1", None, None, 1.0 / 6.0)]),
    14: (5, [("a1", None, None, 0.6666666666666667e-1), ("a3", None, None, 0.7500000000000000e-1)]),
    26: (
        7,
        [
            ("a1", None, None, 0.4761904761904762e-1),
            ("a2", None, None, 0.3809523809523810e-1),
            ("a3", None, None, 0.3214285714285714e-1),
        ],
    ),
    38: (
        9,
        [
            ("a1", None, None, 0.9523809523809524e-2),
            ("a3", None, None, 0.3214285714285714e-1),
            ("ck", 0.4597008433809831, None, 0.2857142857142857e-1),
        ],
    ),
    50: (
        11,
        [
            ("a1", None, None, 0.1269841269841270e-1),
            ("a2", None, None, 0.2257495590828924e-1),
            ("a3", None, None, 0.2109375000000000e-1),
            ("bk", 0.3015113445777636, None, 0.2017333553791887e-1),
        ],
    ),
    110: (
        17,
        [
            ("a1", None, None, 0.3828270494937162e-2),
            ("a3", None, None, 0.9793737512487512e-2),
            ("bk", 0.1851156353447362, None, 0.8211737283191111e-2),
            ("bk", 0.6904210483822922, None, 0.9942814891178103e-2),
            ("bk", 0.3956894730559419, None, 0.9595471336070963e-2),
            ("ck", 0.4783690288121502, None, 0.9694996361663028e-2),
        ],
    ),
    194: (
        23,
        [
            ("a1", None, None, 0.1782340447244611e-2),
            ("a2", None, None, 0.5716905949977102e-2),
            ("a3", None, None, 0.5573383178848738e-2),
            ("bk", 0.6712973442695226, None, 0.5608704082587997e-2),
            ("bk", 0.2892465627575439, None, 0.5158237711805383e-2),
            ("bk", 0.4446933178717437, None, 0.5518771467273614e-2),
            ("bk", 0.1299335447650067, None, 0.4106777028169394e-2),
            ("ck", 0.3457702197611283, None, 0.5051846064614808e-2),
            ("dk", 0.1590417105383530, 0.8360360154824589, 0.5530248916233094e-2),
        ],
    ),
    302: (
        29,
        [
            ("a1", None, None, 0.8545911725128148e-3),
            ("a3", None, None, 0.3599119285025571e-2),
            ("bk", 0.3515640345570105, None, 0.3449788424305883e-2),
            ("bk", 0.6566329410219612, None, 0.3604822601419882e-2),
            ("bk", 0.4729054132581005, None, 0.3576729661743367e-2),
            ("bk", 0.9618308522614784e-1, None, 0.2352101413689164e-2),
            ("bk", 0.2219645236294178, None, 0.3108953122413675e-2),
            ("bk", 0.7011766416089545, None, 0.3650045807677255e-2),
            ("ck", 0.2644152887060663, None, 0.2982344963171804e-2),
            ("ck", 0.5718955891878961, None, 0.3600820932216460e-2),
            ("dk", 0.2510034751770465, 0.8000727494073952, 0.3571540554273387e-2),
            ("dk", 0.1233548532583327, 0.4127724083168531, 0.3392312205006170e-2),
        ],
    ),
}

#: number of points -> polynomial degree that the rule integrates exactly
LEBEDEV_NPOINTS = {n: deg for n, (deg, _) in _RULES.items()}
#: polynomial degree -> number of points (qc-grid's ``LEBEDEV_DEGREES``)
LEBEDEV_DEGREES = {deg: n for n, deg in LEBEDEV_NPOINTS.items()}

_SIGNS3 = list(itertools.product((1.0, -1.0), repeat=3))


def _orbit(code, p1, p2):
    """Return the points of one octahedral orbit, shape (m, 3)."""
    if code == "a1":
        pts = [(s, 0.0, 0.0) for s in (1.0, -1.0)]
        pts += [(0.0, s, 0.0) for s in (1.0, -1.0)]
        pts += [(0.0, 0.0, s) for s in (1.0, -1.0)]
        return np.array(pts)
    if code == "a2":
        v = np.sqrt(0.5)
        pts = []
        for s1, s2 in itertools.product((1.0, -1.0), repeat=2):
            pts += [(0.0, s1 * v, s2 * v), (s1 * v, 0.0, s2 * v), (s1 * v, s2 * v, 0.0)]
        return np.array(pts)
    if code == "a3":
        v = np.sqrt(1.0 / 3.0)
        return np.array([(s1 * v, s2 * v, s3 * v) for s1, s2, s3 in _SIGNS3])
    if code == "bk":
        m = np.sqrt(1.0 - 2.0 * p1 * p1)
        pts = []
        for s1, s2, s3 in _SIGNS3:
            pts += [(s1 * p1, s2 * p1, s3 * m), (s1 * p1, s2 * m, s3 * p1), (s1 * m, s2 * p1, s3 * p1)]
        return np.array(pts)
    if code == "ck":
        q = np.sqrt(1.0 - p1 * p1)
        pts = []
        for s1, s2 in itertools.product((1.0, -1.0), repeat=2):
            pts += [
                (s1 * p1, s2 * q, 0.0),
                (s1 * q, s2 * p1, 0.0),
                (s1 * p1, 0.0, s2 * q),
                (s1 * q, 0.0, s2 * p1),
                (0.0, s1 * p1, s2 * q),
                (0.0, s1 * q, s2 * p1),
            ]
        return np.array(pts)
    if code == "dk":
        t = np.sqrt(1.0 - p1 * p1 - p2 * p2)
        base = (p1, p2, t)
        pts = []
        for perm in itertools.permutations(range(3)):
            v = [base[perm[0]], base[perm[1]], base[perm[2]]]
            for s1, s2, s3 in _SIGNS3:
                pts.append((s1 * v[0], s2 * v[1], s3 * v[2]))
        return np.array(pts)
    raise ValueError(f"unknown orbit code {code}")


_CACHE = {}


def lebedev_rule(npoint):
    """Return ``(points, weights)`` of the Lebedev rule with ``npoint`` points.

    ``points`` has shape (npoint, 3) on the unit sphere and ``weights`` sums to
    one.  Raises ``ValueError`` for sizes without a tabulated rule.
    """
    if npoint not in _RULES:
        raise ValueError(f"No Lebedev rule with {npoint} points; available: {sorted(_RULES)}")
    if npoint not in _CACHE:
        pts, wts = [], []
        for code, p1, p2, w in _RULES[npoint][1]:
            o = _orbit(code, p1, p2)
            pts.append(o)
            wts.append(np.full(len(o), w))
        pts = np.concatenate(pts)
        wts = np.concatenate(wts)
        assert len(pts) == npoint
        pts.setflags(write=False)
        wts.setflags(write=False)
        _CACHE[npoint] = (pts, wts)
    return _CACHE[npoint]
