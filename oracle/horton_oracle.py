"""CPU oracle: a numpy FP64 restatement of horton-part's stockholder loop.

TEST INFRASTRUCTURE ONLY.  Nothing in the product package imports this file;
only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` use it, and only as the checker (or as the timed CPU baseline).

The reference (horton-part 1.1.7) is pure Python and cannot be imported in
this environment (its qc-grid, qpsolvers and cvxopt dependencies are absent),
so this module restates its arithmetic from the source, keeping the
reference's structure: every pro-atom is evaluated on every molecular grid
point, atoms are visited in Python loops, and the inner solvers stop on the
same tests.  Citations are ``file:line`` in the reference's ``src/horton_part``.

It takes grid objects with the qc-grid surface the reference uses
(``points``, ``weights``, ``atgrids``, ``indices``, ``integrate``; atom grids
with ``rgrid``, ``points``, ``weights``, ``indices``, ``sph_weights``).

Pins (see tests/): water MBIS known answer (``tests/test_wpart.py:95-101``),
basis known values (``tests/core/test_basis.py:113-188``), MBIS initial
parameters (``tests/test_mbis.py:26-41``), brute-force QP enumeration,
brute-force Hessian sums, partition of unity, spherical-average identities.
The aLISA ``cvxopt`` solver is not restated (third-party interior-point
iterates): parity unpinned, not provided.
"""

import math
import warnings

import numpy as np
from scipy.interpolate import CubicHermiteSpline, CubicSpline
from scipy.linalg import solve as scipy_solve

DENSITY_CUTOFF = 1e-15  # data/constants.yaml:27
NEGATIVE_CUTOFF = -1e-12  # data/constants.yaml:28
POPULATION_CUTOFF = 1e-4  # data/constants.yaml:29


class OracleWarning(UserWarning):
    pass


# ----------------------------------------------------------------------------
# basis functions: core/basis.py:85-186, 214-232, 293-297
# ----------------------------------------------------------------------------


def evaluate_function(n, population, alpha, r, nderiv=0):
    """``population * n alpha^(3/n) / (4 pi Gamma(3/n)) * exp(-alpha r^n)``.

    core/basis.py:85-186 (scalar ``n``/``alpha`` branch).
    """
    if np.any(np.asarray(n) <= 0):
        raise ValueError("n must be positive")
    if np.any(np.asarray(alpha) < 0):
        raise ValueError("alpha must be non-negative")
    if not isinstance(r, np.ndarray):
        raise ValueError("r must be a numpy array")
    prefactor = population * n * alpha ** (3 / n) / (4 * np.pi * math.gamma(3 / n))
    f = prefactor * np.exp(-alpha * r**n)
    if nderiv == 0:
        return f
    if nderiv == 1:
        return f, -n * r ** (n - 1) * f
    raise NotImplementedError


class OracleBasis:
    """Per-element exponential basis: orders, exponents, initial values."""

    def __init__(self, orders, exponents, initials):
        self.orders = {int(k): np.asarray(v, float) for k, v in orders.items()}
        self.exponents = {int(k): np.asarray(v, float) for k, v in exponents.items()}
        self.initials = {int(k): np.asarray(v, float) for k, v in initials.items()}

    def get_nshell(self, z):
        return len(self.exponents[int(z)])

    def shell(self, z, k, population, r, nderiv=0):
        return evaluate_function(self.orders[int(z)][k], population, self.exponents[int(z)][k], r, nderiv)

    def proatom(self, z, pops, r, nderiv=0):
        """core/basis.py:214-232."""
        y = d = 0.0
        for k in range(self.get_nshell(z)):
            res = self.shell(z, k, pops[k], r, nderiv)
            if nderiv == 0:
                y = y + res
            else:
                y = y + res[0]
                d = d + res[1]
        return y if nderiv == 0 else (y, d)


# ----------------------------------------------------------------------------
# checks: utils.py:198-252, 255-553
# ----------------------------------------------------------------------------


def compute_quantities(density, pars, bs, density_cutoff, do_ratio=True, do_ln_ratio=True):
    """utils.py:198-252."""
    pars = np.asarray(pars).flatten()
    pro_shells = bs * pars[:, None]
    pro = np.einsum("ij->j", pro_shells)
    sick = (density < density_cutoff) | (pro < density_cutoff)
    ratio = ln_ratio = None
    with np.errstate(all="ignore"):
        if do_ratio:
            ratio = np.divide(density, pro, out=np.zeros_like(density), where=~sick)
        if do_ln_ratio:
            ln_ratio = np.log(ratio, out=np.zeros_like(density), where=~sick)
    return pro_shells, pro, sick, ratio, ln_ratio


def check_population(pars, pop, cutoff=POPULATION_CUTOFF):
    """utils.py:426-456: warn when ``|sum(pars) - pop| > cutoff``."""
    if abs(np.sum(pars) - pop) > cutoff:
        warnings.warn("sum of pro-atom parameters differs from the population", OracleWarning)
        return False
    return True


def check_negative_density(pro, negative_cutoff=NEGATIVE_CUTOFF):
    """utils.py:492-520: raise when the pro-atom density is negative."""
    if (pro < negative_cutoff).any():
        raise RuntimeError("Negative pro-atom density found!")


def check_negative_pars(pars, negative_cutoff=NEGATIVE_CUTOFF):
    """utils.py:459-489: warn on negative parameters."""
    if (np.asarray(pars) < negative_cutoff).any():
        warnings.warn("Not all pro-atom parameters are positive", OracleWarning)


def check_monotonic(pro, negative_cutoff=NEGATIVE_CUTOFF):
    """utils.py:523-553 with ``as_warn=False``."""
    if (pro[:-1] - pro[1:] < negative_cutoff).any():
        raise RuntimeError("Pro-atom density should be monotonically decreasing.")


# ----------------------------------------------------------------------------
# MBIS helpers: mbis.py:36-163
# ----------------------------------------------------------------------------


def get_nshell(number):
    """mbis.py:36-46."""
    noble = np.array([2, 10, 18, 36, 54, 86, 118])
    return int(noble.searchsorted(number) + 1)


def get_initial_mbis_propars(number):
    """mbis.py:49-78."""
    nshell = get_nshell(number)
    propars = np.zeros(2 * nshell, float)
    S0 = 2.0 * number
    if nshell > 1:
        alpha = (2.0 / S0) ** (1.0 / (nshell - 1))
    else:
        alpha = 1.0
    nel = np.array([2.0, 8.0, 8.0, 18.0, 18.0, 32.0, 32.0])
    for k in range(nshell):
        propars[2 * k] = nel[k]
        propars[2 * k + 1] = S0 * alpha**k
    propars[-2] = number - propars[:-2:2].sum()
    return propars


def opt_mbis_propars(rho, propars, weights, r, threshold, density_cutoff=1e-15, stats=None):
    """mbis.py:81-163; ``stats`` (dict) receives the iteration count and flags."""
    nshell = len(propars) // 2
    terms = np.zeros((nshell, len(r)), float)
    oldpro = None
    pop = np.einsum("p,p->", weights, rho)
    for irep in range(2000):
        for k in range(nshell):
            N = propars[2 * k]
            S = propars[2 * k + 1]
            terms[k] = N * S**3 * np.exp(-S * r) / (8 * np.pi)
        pro = terms.sum(axis=0)
        sick = (rho < density_cutoff) | (pro < density_cutoff)
        with np.errstate(all="ignore"):
            ratio = rho / pro
        ratio[sick] = 0.0
        terms *= ratio
        for k in range(nshell):
            m0 = np.einsum("p,p->", weights, terms[k])
            m1 = np.einsum("p,p,p->", weights, terms[k], r)
            propars[2 * k] = m0
            propars[2 * k + 1] = 3 * m0 / m1
        if oldpro is None:
            change = 1e100
        else:
            error = oldpro - pro
            change = np.sqrt(np.einsum("p,p,p->", weights, error, error))
        if change < threshold:
            popok = bool(np.isclose(pop, np.sum(propars[0::2]), atol=1e-4))
            if not popok:
                warnings.warn("The sum of propars are not equal to the atomic pop.", OracleWarning)
            if stats is not None:
                stats.update(niter=irep + 1, converged=True, popok=popok)
            return propars
        oldpro = pro
    warnings.warn("MBIS not converged, but still go ahead!", OracleWarning)
    if stats is not None:
        stats.update(niter=2000, converged=False, popok=True)
    return propars


# ----------------------------------------------------------------------------
# GISA QP: gisa.py:348-421, restated as a primal active-set method on
# min 1/2 x'Px + q'x  s.t.  x >= 0, sum(x) = pop   (strictly convex: the
# optimum is unique, so any exact method returns quadprog's answer).
# ----------------------------------------------------------------------------


def _solve_dense(M, b):
    """Gaussian elimination with partial pivoting (row swaps by max |pivot|)."""
    M = M.astype(float).copy()
    b = b.astype(float).copy()
    n = len(b)
    for c in range(n):
        p = c + int(np.argmax(np.abs(M[c:, c])))
        if p != c:
            M[[c, p]] = M[[p, c]]
            b[[c, p]] = b[[p, c]]
        piv = M[c, c]
        for r_ in range(c + 1, n):
            f = M[r_, c] / piv
            if f != 0.0:
                M[r_, c:] -= f * M[c, c:]
                b[r_] -= f * b[c]
    x = np.zeros(n)
    for r_ in range(n - 1, -1, -1):
        x[r_] = (b[r_] - M[r_, r_ + 1 :] @ x[r_ + 1 :]) / M[r_, r_]
    return x


def qp_simplex_active_set(P, q, pop, maxiter=200):
    """Primal active-set QP over the scaled simplex; returns (x, niter)."""
    K = len(q)
    x = np.full(K, pop / K)
    fixed = np.zeros(K, dtype=bool)
    for it in range(maxiter):
        free = np.flatnonzero(~fixed)
        m = len(free)
        M = np.zeros((m + 1, m + 1))
        M[:m, :m] = P[np.ix_(free, free)]
        M[:m, m] = -1.0
        M[m, :m] = 1.0
        rhs = np.concatenate([-q[free], [pop]])
        sol = _solve_dense(M, rhs)
        xs = sol[:m]
        nu = sol[m]
        if np.all(xs >= 0.0):
            x[:] = 0.0
            x[free] = xs
            g = P @ x + q - nu
            mu = np.where(fixed, g, 0.0)
            jmin = int(np.argmin(mu))
            if not fixed.any() or mu[jmin] >= 0.0:
                return x, it + 1
            fixed[jmin] = False
        else:
            alpha = 1.0
            block = -1
            for t, i in enumerate(free):
                if xs[t] < 0.0:
                    a = x[i] / (x[i] - xs[t])
                    if a < alpha:
                        alpha = a
                        block = i
            xn = np.zeros(K)
            xn[free] = x[free] + alpha * (xs - x[free])
            x = xn
            if block >= 0:
                x[block] = 0.0
                fixed[block] = True
    raise RuntimeError("QP active-set did not converge")


def gaussian_overlap_matrix(alphas):
    """gisa.py:389-395: P_kl = 2 pi^-3/2 (a_k a_l)^3/2 / (a_k + a_l)^3/2."""
    a = np.asarray(alphas, float)
    P = 2 / np.pi**1.5 * (a[:, None] * a[None, :]) ** 1.5 / (a[:, None] + a[None, :]) ** 1.5
    return (P + P.T) / 2


def opt_propars_qp(bs_funcs, rho, weights, alphas):
    """gisa.py:348-421 (P analytic, q and pop from the grid)."""
    P = gaussian_overlap_matrix(alphas)
    q = -2 * np.einsum("i,ni,i->n", weights, bs_funcs, rho)
    pop = np.einsum("i,i", weights, rho)
    x, _ = qp_simplex_active_set(P, q, pop)
    check_negative_pars(x)
    check_population(x, pop)
    return x


# ----------------------------------------------------------------------------
# LISA solvers: alisa.py:193-291 (sc), 753-938 (newton / m-newton)
# ----------------------------------------------------------------------------


def solver_sc(bs, rho, propars, weights, threshold, density_cutoff=DENSITY_CUTOFF, max_niter_inner=100000, stats=None):
    """alisa.py:193-291."""
    oldpro = None
    for irep in range(int(max_niter_inner)):
        pro_shells, pro, sick, ratio, _ = compute_quantities(rho, propars, bs, density_cutoff, do_ln_ratio=False)
        propars[:] = np.einsum("p,ip->i", weights, pro_shells * ratio)
        if oldpro is None:
            change = 1e100
        else:
            error = oldpro - pro
            change = np.sqrt(np.einsum("i,i,i", weights, error, error))
        if change < threshold:
            pop = np.einsum("i,i", weights, rho)
            check_population(propars, pop)
            if stats is not None:
                stats.update(niter=irep + 1, converged=True)
            return propars
        oldpro = pro
    warnings.warn("Inner iteration is not converge!", OracleWarning)
    if stats is not None:
        stats.update(niter=int(max_niter_inner), converged=False)
    return propars


def solver_newton(
    bs,
    rho,
    propars,
    weights,
    threshold,
    density_cutoff=DENSITY_CUTOFF,
    negative_cutoff=NEGATIVE_CUTOFF,
    mode="exact",
    tau=1.0,
    linspace_size=20,
    check_mono=False,
    niter=10000,
    stats=None,
):
    """alisa.py:753-938 with ``mode`` "exact" (newton) or "modified" (m-newton)."""
    oldpro = None
    change = 1e100
    pop = np.einsum("i,i", weights, rho)
    for irep in range(niter):
        _, pro, sick, ratio, _ = compute_quantities(rho, propars, bs, density_cutoff)
        if oldpro is not None:
            error = oldpro - pro
            change = np.sqrt(np.einsum("i,i,i", weights, error, error))
        if change < threshold:
            check_negative_density(pro, negative_cutoff)
            check_population(propars, pop)
            if check_mono:
                check_monotonic(pro, negative_cutoff)
            if stats is not None:
                stats.update(niter=irep + 1, converged=True)
            return propars
        integrand = bs * ratio
        with np.errstate(all="ignore"):
            grad_integrand = integrand / pro
        grad_integrand[:, sick] = 0.0
        df = -np.einsum("kp,p->k", integrand, weights)
        hess = np.einsum("kp,jp,p->kj", grad_integrand, bs, weights)
        delta = scipy_solve(hess, -1 - df, assume_a="sym")
        if mode == "exact":
            propars[:] = propars + delta
        else:
            for x in np.linspace(tau, 0, linspace_size):
                new_propars = propars + x * delta
                new_pro = np.einsum("ij->j", bs * new_propars[:, None])
                ok = (new_pro > negative_cutoff).all()
                if check_mono:
                    ok = ok and (new_pro[:-1] - new_pro[1:] > negative_cutoff).all()
                if ok:
                    break
            else:
                raise RuntimeError("Line search failed!")
            propars[:] = new_propars
        oldpro = pro
    raise RuntimeError("Inner loop: Newton does not converge!")


# ----------------------------------------------------------------------------
# the partitioning driver: core/base.py, core/stockholder.py, core/iterstock.py
# ----------------------------------------------------------------------------


class OracleWPart:
    """Shared stockholder machinery (core/stockholder.py:145-175, 352-384)
    and the outer loop (core/iterstock.py:32-58, 159-193)."""

    def __init__(
        self,
        coordinates,
        numbers,
        pseudo_numbers,
        grid,
        moldens,
        threshold=1e-6,
        maxiter=500,
        inner_threshold=1e-8,
        grid_type=1,
        density_cutoff=DENSITY_CUTOFF,
        negative_cutoff=NEGATIVE_CUTOFF,
        population_cutoff=POPULATION_CUTOFF,
        lmax=3,
        cutoff_radius=None,
    ):
        if not issubclass(np.asarray(numbers).dtype.type, np.int64):
            raise TypeError("The argument numbers must be a vector with length natom.")  # utils.py:163
        self.coordinates = np.asarray(coordinates, float)
        self.numbers = numbers
        self.pseudo_numbers = numbers.astype(float) if pseudo_numbers is None else np.asarray(pseudo_numbers, float)
        self.natom = len(numbers)
        self.grid = grid
        self.moldens = moldens
        self.threshold = threshold
        self.inner_threshold = inner_threshold if inner_threshold < threshold else threshold  # iterstock.py:101
        self.maxiter = maxiter
        assert grid_type in (1, 2, 3)
        self.grid_type = grid_type
        self.on_molgrid = grid_type in (2, 3)  # base.py:490-499
        self.only_molgrid = grid_type == 3  # local = False: moments on the molecular grid
        self.density_cutoff = density_cutoff
        self.negative_cutoff = negative_cutoff
        self.population_cutoff = population_cutoff
        self.lmax = lmax
        self.cutoff_radius = cutoff_radius
        self.cache = {}
        self.history_propars = []
        self.history_charges = []
        self.history_entropies = []
        self.history_changes = []
        self.inner_stats = []
        # base.py:630-635: distances of every atom to every molecular point
        self.radial_distances = [np.linalg.norm(grid.points - c, axis=1) for c in self.coordinates]

    # -- grid helpers --------------------------------------------------------
    def to_atomic_grid(self, i, data):
        return data[self.grid.indices[i] : self.grid.indices[i + 1]]  # base.py:623-628

    def rgrid(self, i):
        return self.grid.atgrids[i].rgrid

    @property
    def nelec(self):
        return self.grid.integrate(self.moldens)  # base.py:137-139

    # -- pro-atoms -----------------------------------------------------------
    def eval_proatom(self, i, propars):
        raise NotImplementedError

    def _cull(self, i, values):
        if self.cutoff_radius is not None:
            values = np.where(self.radial_distances[i] <= self.cutoff_radius, values, 0.0)
        return values

    def update_at_weights(self, propars):
        """core/stockholder.py:153-175 and 352-384."""
        promol = np.zeros(self.grid.size)
        proat = []
        for i in range(self.natom):
            work = self._cull(i, self.eval_proatom(i, propars))
            promol += work
            promol += 1e-100
            proat.append(work if self.on_molgrid else self.to_atomic_grid(i, work).copy())
        for i in range(self.natom):
            w = proat[i] / (promol if self.on_molgrid else self.to_atomic_grid(i, promol))
            np.clip(w, 0, 1, out=w)
            self.cache[f"at_weights_{i}"] = w
        self.cache["promoldens"] = promol

    def compute_entropy(self):
        """core/stockholder.py:145-151."""
        rho = self.moldens
        rho0 = self.cache["promoldens"]
        sick = (rho0 < self.density_cutoff) | (rho < self.density_cutoff)
        with np.errstate(all="ignore"):
            ratio = np.divide(rho, rho0, out=np.zeros_like(rho), where=~sick)
            ln_ratio = np.log(ratio, out=np.zeros_like(rho), where=~sick)
        return self.grid.integrate(rho, ln_ratio)

    def get_proatom_rho(self, i, propars):
        raise NotImplementedError

    def compute_change(self, p1, p2):
        """core/iterstock.py:32-45."""
        msd = 0.0
        for i in range(self.natom):
            delta = self.get_proatom_rho(i, p1) - self.get_proatom_rho(i, p2)
            if self.on_molgrid:
                msd += self.grid.integrate(delta, delta)
            else:
                rg = self.rgrid(i)
                msd += rg.integrate(4 * np.pi * rg.points**2, delta, delta)
        return np.sqrt(msd)

    # -- per-atom input to the inner solvers (mbis.py:167-203, gisa.py:257-318)
    def atom_problem(self, i):
        w = self.cache[f"at_weights_{i}"]
        if self.on_molgrid:
            rhoa = w * self.moldens
            return rhoa, self.grid.weights, self.radial_distances[i]
        atgrid = self.grid.atgrids[i]
        dens = self.to_atomic_grid(i, self.moldens)
        f = w * dens
        # spherical average at the knots: sum_j wl_j f_ij (qc-grid AtomGrid.spherical_average)
        rhoa = np.add.reduceat(f * atgrid.sph_weights, atgrid.indices[:-1])
        r = atgrid.rgrid.points
        weights = 4 * np.pi * r**2 * atgrid.rgrid.weights
        self.cache[f"radial_points_{i}"] = r
        self.cache[f"spherical_average_{i}"] = rhoa
        self.cache[f"radial_weights_{i}"] = atgrid.rgrid.weights
        return rhoa, weights, r

    def update_atom(self, i, propars):
        raise NotImplementedError

    def init_propars(self):
        raise NotImplementedError

    def do_partitioning(self):
        """core/iterstock.py:159-193."""
        propars = self.init_propars()
        self.cache["propars"] = propars
        self.cache["charges"] = np.zeros(self.natom)
        counter = 0
        while True:
            counter += 1
            old = propars.copy()
            # _update_propars, iterstock.py:132-149
            self.update_at_weights(propars)
            self.inner_stats.append([])
            for i in range(self.natom):
                self.update_atom(i, propars)
            self.history_propars.append(propars.copy())
            self.history_charges.append(self.cache["charges"].copy())
            self.history_entropies.append(self.compute_entropy())
            change = self.compute_change(propars, old)
            self.history_changes.append(change)
            if change < self.threshold or counter >= self.maxiter:
                break
        # finalize_propars, iterstock.py:48-58
        charges = self.cache["charges"]
        self.cache["niter"] = counter
        self.cache["change"] = change
        self.cache["populations"] = self.numbers - charges
        self.cache["pseudo_populations"] = self.pseudo_numbers - charges
        for k in ("propars", "charges", "entropies", "changes"):
            self.cache[f"history_{k}"] = np.array(getattr(self, f"history_{k}"))
        self.finalize()
        return self.cache

    def finalize(self):
        pass

    # -- moments: core/base.py:329-402 --------------------------------------
    def do_moments(self):
        lmax = self.lmax
        powers = [(nx, ny, l - nx - ny) for l in range(lmax + 1) for nx in range(l, -1, -1) for ny in range(l - nx, -1, -1)]
        cart = np.zeros((self.natom, len(powers)))
        pure = np.zeros((self.natom, (lmax + 1) ** 2))
        rad = np.zeros((self.natom, lmax + 1))
        for i in range(self.natom):
            w = self.cache[f"at_weights_{i}"]
            if self.only_molgrid:
                # local = False (base.py:455-459): the whole molecular grid
                points, weights, aim = self.grid.points, self.grid.weights, self.moldens * w
            else:
                atgrid = self.grid.atgrids[i]
                if w.shape == self.grid.weights.shape and self.on_molgrid:
                    w = self.to_atomic_grid(i, w)
                points, weights = atgrid.points, atgrid.weights
                aim = self.to_atomic_grid(i, self.moldens) * w
            d = points - self.coordinates[i]
            wf = weights * aim
            cart[i] = -np.array([np.sum(wf * d[:, 0] ** a * d[:, 1] ** b * d[:, 2] ** c) for a, b, c in powers])
            cart[i, 0] += self.pseudo_numbers[i]
            pure[i] = -(real_solid_harmonics(lmax, d) @ wf)
            pure[i, 0] += self.pseudo_numbers[i]
            r = np.sqrt(np.sum(d * d, axis=1))
            rad[i] = [np.sum(wf * r**n) for n in range(lmax + 1)]
        self.cache["cartesian_multipoles"] = cart
        self.cache["pure_multipoles"] = pure
        self.cache["radial_moments"] = rad
        return cart, pure, rad


def real_solid_harmonics(lmax, d):
    """Racah-normalised real regular solid harmonics from scipy's complex
    spherical harmonics: ``sqrt(4 pi/(2l+1)) r^l`` times ``Y_l0``,
    ``sqrt2 (-1)^m Re Y_lm`` and ``sqrt2 (-1)^m Im Y_lm`` (the factor
    ``(-1)^m`` removes scipy's Condon-Shortley phase); order c_l0, c_l1,
    s_l1, c_l2, s_l2, ... (HORTON order)."""
    from scipy.special import sph_harm_y

    r = np.sqrt(np.sum(d * d, axis=1))
    theta = np.arccos(np.clip(np.divide(d[:, 2], r, out=np.ones_like(r), where=r > 0), -1.0, 1.0))
    phi = np.arctan2(d[:, 1], d[:, 0])
    out = []
    for l in range(lmax + 1):
        f = np.sqrt(4 * np.pi / (2 * l + 1)) * r**l
        out.append(f * sph_harm_y(l, 0, theta, phi).real)
        for m in range(1, l + 1):
            y = sph_harm_y(l, m, theta, phi)
            out.append(f * np.sqrt(2.0) * (-1) ** m * y.real)
            out.append(f * np.sqrt(2.0) * (-1) ** m * y.imag)
    return np.array(out)


class OracleMBIS(OracleWPart):
    """mbis.py:206-362."""

    def init_propars(self):
        """mbis.py:291-305."""
        self.ranges = [0]
        self.nshells = []
        for z in self.numbers:
            k = get_nshell(int(z))
            self.ranges.append(self.ranges[-1] + 2 * k)
            self.nshells.append(k)
        propars = np.zeros(self.ranges[-1])
        for i, z in enumerate(self.numbers):
            propars[self.ranges[i] : self.ranges[i + 1]] = get_initial_mbis_propars(int(z))
        return propars

    def _rho(self, i, propars, r):
        my = propars[self.ranges[i] : self.ranges[i + 1]]
        y = np.zeros(len(r))
        for k in range(self.nshells[i]):
            N, S = my[2 * k : 2 * k + 2]
            y += N * S**3 * np.exp(-S * r) / (8 * np.pi)
        return y

    def eval_proatom(self, i, propars):
        """mbis.py:263-289 (no 1e-100 here)."""
        return self._rho(i, propars, self.radial_distances[i])

    def get_proatom_rho(self, i, propars):
        """mbis.py:234-261 (density part)."""
        r = self.radial_distances[i] if self.on_molgrid else self.rgrid(i).points
        return self._rho(i, propars, r)

    def update_atom(self, i, propars):
        """mbis.py:167-203."""
        rhoa, weights, r = self.atom_problem(i)
        my = propars[self.ranges[i] : self.ranges[i + 1]]
        stats = {}
        my[:] = opt_mbis_propars(rhoa, my.copy(), weights, r, self.inner_threshold, stats=stats)
        self.inner_stats[-1].append(stats)
        self.cache["charges"][i] = self.pseudo_numbers[i] - np.einsum("p,p->", weights, rhoa)

    def finalize(self):
        """mbis.py:347-362."""
        propars = self.cache["propars"]
        vc = np.array([-propars[self.ranges[i + 1] - 2] for i in range(self.natom)])
        vw = np.array([1.0 / propars[self.ranges[i + 1] - 1] for i in range(self.natom)])
        self.cache["valence_charges"] = vc
        self.cache["valence_widths"] = vw
        self.cache["core_charges"] = self.cache["charges"] - vc


class OracleGISA(OracleWPart):
    """gisa.py:109-345 (solver: the QP)."""

    def __init__(self, *args, basis=None, **kwargs):
        self.basis = basis
        super().__init__(*args, **kwargs)

    def init_propars(self):
        """gisa.py:68-106."""
        self.ranges = [0]
        for z in self.numbers:
            self.ranges.append(self.ranges[-1] + self.basis.get_nshell(int(z)))
        propars = np.ones(self.ranges[-1])
        for i, z in enumerate(self.numbers):
            inits = self.basis.initials[int(z)].copy()
            inits[inits < 1e-4] = 1e-4
            propars[self.ranges[i] : self.ranges[i + 1]] = inits / np.sum(inits) * self.pseudo_numbers[i]
        propars[:] = propars / np.sum(propars) * self.nelec
        self.initial_propars_modified = propars.copy()
        self.bs_funcs = []
        for i, z in enumerate(self.numbers):
            r = self.radial_distances[i] if self.on_molgrid else self.rgrid(i).points
            self.bs_funcs.append(np.array([self.basis.shell(int(z), k, 1.0, r) for k in range(self.basis.get_nshell(int(z)))]))
        return propars

    def eval_proatom(self, i, propars):
        """gisa.py:224-244."""
        pops = propars[self.ranges[i] : self.ranges[i + 1]]
        return self.basis.proatom(int(self.numbers[i]), pops, self.radial_distances[i], 0)

    def get_proatom_rho(self, i, propars):
        """gisa.py:42-65."""
        pops = propars[self.ranges[i] : self.ranges[i + 1]]
        r = self.radial_distances[i] if self.on_molgrid else self.rgrid(i).points
        return self.basis.proatom(int(self.numbers[i]), pops, r, 0)

    def opt(self, i, bs, rhoa, weights, my):
        z = int(self.numbers[i])
        self.inner_stats[-1].append({})
        return opt_propars_qp(bs, rhoa, weights, self.basis.exponents[z])

    def update_atom(self, i, propars):
        """gisa.py:251-318."""
        rhoa, weights, r = self.atom_problem(i)
        my = propars[self.ranges[i] : self.ranges[i + 1]]
        my[:] = self.opt(i, self.bs_funcs[i], rhoa, weights, my.copy())
        self.cache["charges"][i] = self.pseudo_numbers[i] - np.einsum("i,i", weights, rhoa)


class OracleLISA(OracleGISA):
    """alisa.py:1155-1335 with the sc / newton / m-newton solvers."""

    def __init__(self, *args, solver="sc", solver_options=None, **kwargs):
        self.solver = solver
        self.solver_options = dict(solver_options or {})
        super().__init__(*args, **kwargs)

    def opt(self, i, bs, rhoa, weights, my):
        stats = {}
        if self.solver == "sc":
            out = solver_sc(bs, rhoa, my, weights, self.inner_threshold, self.density_cutoff, stats=stats, **self.solver_options)
        elif self.solver in ("newton", "m-newton"):
            out = solver_newton(
                bs,
                rhoa,
                my,
                weights,
                self.inner_threshold,
                self.density_cutoff,
                self.negative_cutoff,
                mode="exact" if self.solver == "newton" else "modified",
                stats=stats,
                **self.solver_options,
            )
        else:
            raise NotImplementedError(self.solver)
        self.inner_stats[-1].append(stats)
        return out


# ----------------------------------------------------------------------------
# NLIS / GMBIS: nlis.py:44-195, 201-367; gmbis.py:36-156
# ----------------------------------------------------------------------------


def nlis_shell_norm(n, S):
    """n S^(3/n) / (4 pi Gamma(3/n)): the unit-population shell prefactor (nlis.py:157)."""
    return n * S ** (3 / n) / (4 * np.pi * math.gamma(3.0 / n))


def get_nlis_nshell(number, nshell_dict):
    """nlis.py:44-56."""
    return nshell_dict.get(number, get_nshell(number))


def get_initial_nlis_propars(number, exp_n_dict, nshell_dict):
    """nlis.py:59-96: (N, S, n) per shell, N = Z/K, S geometric from 2Z to 0.5."""
    nbs = get_nlis_nshell(number, nshell_dict)
    propars = np.ones(3 * nbs, float)
    S0 = 2.0 * number
    alpha = (0.5 / S0) ** (1.0 / (nbs - 1)) if nbs > 1 else 1.0
    for k in range(nbs):
        propars[3 * k] = number / nbs
        propars[3 * k + 1] = S0 * alpha**k
        propars[3 * k + 2] = exp_n_dict[(number, k)] if (number, k) in exp_n_dict else 1.0
    return propars


def get_initial_gmbis_propars(number, exp_n_dict):
    """gmbis.py:36-72: MBIS shell counts and populations, S geometric from 2Z to 2."""
    nshell = get_nshell(number)
    propars = np.zeros(3 * nshell, float)
    S0 = 2.0 * number
    alpha = (2.0 / S0) ** (1.0 / (nshell - 1)) if nshell > 1 else 1.0
    nel = np.array([2.0, 8.0, 8.0, 18.0, 18.0, 32.0, 32.0])
    for k in range(nshell):
        propars[3 * k] = nel[k]
        propars[3 * k + 1] = S0 * alpha**k
        propars[3 * k + 2] = exp_n_dict[(number, k)] if (number, k) in exp_n_dict else 1.0
    propars[-3] = number - propars[:-3:3].sum()
    return propars


def opt_nlis_propars(rho, propars, weights, r, threshold, density_cutoff=1e-15, stats=None):
    """nlis.py:99-195.  Unlike MBIS, the exponent update is 3 / (n m1) with m1
    the weighted moment of the unit-population shell (no factor N)."""
    nshell = len(propars) // 3
    terms = np.zeros((nshell, len(r)), float)
    oldpro = None
    pop = np.einsum("p,p->", weights, rho)
    for irep in range(2000):
        for k in range(nshell):
            S, n = propars[3 * k + 1], propars[3 * k + 2]
            terms[k] = n * S ** (3 / n) * np.exp(-S * r**n) / (4 * np.pi * math.gamma(3.0 / n))
        pro = np.sum(terms * propars[::3, None], axis=0)
        sick = (rho < density_cutoff) | (pro < density_cutoff)
        with np.errstate(all="ignore"):
            ratio = rho / pro
        ratio[sick] = 0.0
        terms_ratio = terms * ratio
        for k in range(nshell):
            N, n = propars[3 * k], propars[3 * k + 2]
            m0 = np.einsum("p,p->", weights, terms_ratio[k] * N)
            m1 = np.einsum("p,p,p->", weights, terms_ratio[k], r**n)
            propars[3 * k] = m0
            propars[3 * k + 1] = 1e-5 if np.isclose(m1, 0.0) else 3 / (m1 * n)
        if oldpro is None:
            change = 1e100
        else:
            error = oldpro - pro
            change = np.sqrt(np.einsum("p,p,p->", weights, error, error))
        if change < threshold:
            popok = bool(np.isclose(pop, np.sum(propars[0::3]), atol=1e-4))
            if not popok:
                warnings.warn("The sum of propars are not equal to the atomic pop.", OracleWarning)
            if stats is not None:
                stats.update(niter=irep + 1, converged=True, popok=popok)
            return propars
        oldpro = pro
    warnings.warn("NLIS not converged, but still go ahead!", OracleWarning)
    if stats is not None:
        stats.update(niter=2000, converged=False, popok=True)
    return propars


class OracleNLIS(OracleWPart):
    """nlis.py:201-367 (exp_n_dict, nshell_dict as NLISWPart takes them)."""

    def __init__(self, *args, exp_n_dict=None, nshell_dict=None, **kwargs):
        self.exp_n_dict = exp_n_dict if exp_n_dict is not None else {}
        self.nshell_dict = nshell_dict or {}
        super().__init__(*args, **kwargs)

    def _initial(self, z):
        return get_initial_nlis_propars(z, self.exp_n_dict, self.nshell_dict)

    def init_propars(self):
        """nlis.py:330-346."""
        per_atom = [self._initial(int(z)) for z in self.numbers]
        self.nshells = [len(p) // 3 for p in per_atom]
        self.ranges = [0] + list(np.cumsum([len(p) for p in per_atom]))
        return np.concatenate(per_atom)

    def _rho(self, i, propars, r):
        my = propars[self.ranges[i] : self.ranges[i + 1]]
        y = np.zeros(len(r))
        for k in range(self.nshells[i]):
            N, S, n = my[3 * k : 3 * k + 3]
            y += N * n * S ** (3 / n) * np.exp(-S * r**n) / (4 * np.pi * math.gamma(3.0 / n))
        return y

    def eval_proatom(self, i, propars):
        """nlis.py:302-328."""
        return self._rho(i, propars, self.radial_distances[i])

    def get_proatom_rho(self, i, propars):
        """nlis.py:272-300 (density part)."""
        r = self.radial_distances[i] if self.on_molgrid else self.rgrid(i).points
        return self._rho(i, propars, r)

    def update_atom(self, i, propars):
        """mbis.py:167-203 with opt_nlis_propars (nlis.py:348-349)."""
        rhoa, weights, r = self.atom_problem(i)
        my = propars[self.ranges[i] : self.ranges[i + 1]]
        stats = {}
        my[:] = opt_nlis_propars(rhoa, my.copy(), weights, r, self.inner_threshold, stats=stats)
        self.inner_stats[-1].append(stats)
        self.cache["charges"][i] = self.pseudo_numbers[i] - np.einsum("p,p->", weights, rhoa)

    def finalize(self):
        """nlis.py:351-367."""
        propars = self.cache["propars"]
        vc = np.array([-propars[self.ranges[i + 1] - 3] for i in range(self.natom)])
        vw = np.array([1.0 / propars[self.ranges[i + 1] - 2] for i in range(self.natom)])
        self.cache["valence_charges"] = vc
        self.cache["valence_widths"] = vw
        self.cache["core_charges"] = self.cache["charges"] - vc


class OracleGMBIS(OracleNLIS):
    """gmbis.py:75-156: NLIS with MBIS shell counts and initial values."""

    def _initial(self, z):
        return get_initial_gmbis_propars(z, self.exp_n_dict)


# ----------------------------------------------------------------------------
# gLISA: glisa.py:55-1028 (the global problem over every coefficient on the
# molecular grid; solvers sc, newton, m-newton)
# ----------------------------------------------------------------------------


def fix_propars(exp_array, propars, delta):
    """utils.py:556-577: from the most diffuse function on, the indices whose
    coefficient is below 1e-4 in magnitude and whose step is negative."""
    fixed = []
    for index in np.argsort(exp_array):
        if abs(propars[index]) < 1e-4 and delta[index] < 0.0:
            fixed.append(index)
        else:
            break
    return fixed


class OracleGLISA(OracleWPart):
    """glisa.py:55-1028.  ``solver`` is "sc", "newton" or "m-newton";
    ``solver_options`` as the reference's solver methods take them."""

    def __init__(self, *args, basis=None, solver="m-newton", solver_options=None, **kwargs):
        self.basis = basis
        self.solver = solver
        self.solver_options = dict(solver_options or {})
        super().__init__(*args, **kwargs)

    def init_propars(self):
        """gisa.py:68-88 and the basis on the molecular grid (glisa.py:250-258)."""
        self.ranges = [0]
        for z in self.numbers:
            self.ranges.append(self.ranges[-1] + self.basis.get_nshell(int(z)))
        propars = np.ones(self.ranges[-1])
        for i, z in enumerate(self.numbers):
            inits = self.basis.initials[int(z)].copy()
            inits[inits < 1e-4] = 1e-4
            propars[self.ranges[i] : self.ranges[i + 1]] = inits / np.sum(inits) * self.pseudo_numbers[i]
        propars[:] = propars / np.sum(propars) * self.nelec
        rows = []
        for i, z in enumerate(self.numbers):
            r = self.radial_distances[i]
            rows += [self.basis.shell(int(z), k, 1.0, r) for k in range(self.basis.get_nshell(int(z)))]
        self.pro_shells = np.array(rows)
        return propars

    def calc_promol_dens(self, propars):
        """glisa.py:346-348 (no floors)."""
        return np.einsum("np,n->p", self.pro_shells, propars)

    def working_matrix(self, rho, rho0, nderiv=0):
        """glisa.py:411-479: objective, gradient and Hessian."""
        sick = (rho < self.density_cutoff) | (rho0 < self.density_cutoff)
        with np.errstate(all="ignore"):
            ratio = np.divide(rho, rho0, out=np.zeros_like(rho), where=~sick)
            ln_ratio = np.log(ratio, out=np.zeros_like(ratio), where=~sick)
        f = self.grid.integrate(rho * ln_ratio)
        if nderiv == 0:
            return f
        M = len(self.pro_shells)
        grad = np.zeros(M)
        hess = np.zeros((M, M))
        w = self.grid.weights
        for i in range(M):
            with np.errstate(all="ignore"):
                dfi = rho * self.pro_shells[i] / rho0
            dfi[sick] = 0.0
            grad[i] = -np.sum(w * dfi)
            if nderiv > 1:
                with np.errstate(all="ignore"):
                    hi = dfi[None, :] * self.pro_shells[i:] / rho0[None, :]
                hi[:, sick] = 0.0
                hess[i, i:] = hi @ w
                hess[i:, i] = hess[i, i:]
        return (f, grad) if nderiv == 1 else (f, grad, hess)

    def get_proatom_rho(self, i, propars):
        """gisa.py:42-65 (glisa.py:149-163)."""
        pops = propars[self.ranges[i] : self.ranges[i + 1]]
        r = self.radial_distances[i] if self.on_molgrid else self.rgrid(i).points
        return self.basis.proatom(int(self.numbers[i]), pops, r, 0)

    def is_promol_valid(self, propars, check_mono=False):
        """glisa.py:303-330."""
        for i in range(self.natom):
            rho0 = self.get_proatom_rho(i, propars)
            if (rho0 < self.negative_cutoff).any():
                return False
            if check_mono and not self.on_molgrid and (rho0[:-1] - rho0[1:] < self.negative_cutoff).any():
                return False
        return True

    def function_g(self, x):
        """glisa.py:850-879."""
        rho = self.moldens
        old_rho0 = self.calc_promol_dens(x)
        sick = (rho < self.density_cutoff) | (old_rho0 < self.density_cutoff)
        new_x = np.zeros_like(x)
        for n in range(len(x)):
            with np.errstate(all="ignore"):
                integrand = rho * self.pro_shells[n] * x[n] / old_rho0
            integrand[sick] = 0.0
            new_x[n] = self.grid.integrate(integrand)
        return new_x

    def solve_sc(self, propars):
        """glisa.py:805-848."""
        counter = 0
        while True:
            old = propars.copy()
            propars[:] = self.function_g(propars)
            change = self.compute_change(propars, old)
            rho0 = self.calc_promol_dens(old)
            self.history_entropies.append(self.working_matrix(self.moldens, rho0))
            self.history_changes.append(change)
            self.history_propars.append(propars.copy())
            counter += 1
            if change < self.threshold:
                return propars, counter

    def solve_newton(self, propars, mode, maxiter=100, linspace_size=40, tau=1.0, check_mono=False):
        """glisa.py:617-803 with linesearch_mode "valid-promol"."""
        rho = self.moldens
        pop = self.grid.integrate(rho)
        for irep in range(maxiter):
            old = propars.copy()
            pro = self.calc_promol_dens(propars)
            f, df, hess = self.working_matrix(rho, pro, 2)
            try:
                delta = scipy_solve(hess, -1 - df, assume_a="sym")
            except np.linalg.LinAlgError as e:
                raise RuntimeError(e)
            if mode == "exact":
                propars[:] = propars + delta
            else:
                fixed = np.ones_like(delta)
                for i, z in enumerate(self.numbers):
                    sl = slice(self.ranges[i], self.ranges[i + 1])
                    for j in fix_propars(self.basis.exponents[int(z)], propars[sl], delta[sl]):
                        fixed[self.ranges[i] + j] = 0
                for x in np.linspace(tau, 0, linspace_size, endpoint=False):
                    new = propars + fixed * (x * delta)
                    if self.is_promol_valid(new, check_mono=check_mono):
                        break
                else:
                    raise RuntimeError("Line search failed!")
                propars[:] = new
            entropy = f  # _compute_entropy(rho, pro) is the objective at pro
            change = self.compute_change(propars, old)
            self.history_entropies.append(entropy)
            self.history_propars.append(propars.copy())
            self.history_changes.append(change)
            if change < self.threshold:
                if (pro < self.negative_cutoff).any():
                    raise RuntimeError("Negative pro-atom density found!")
                check_population(propars, pop, self.population_cutoff)
                return propars, irep + 1
        raise RuntimeError("Not converged!")

    def do_partitioning(self):
        """glisa.py:248-281."""
        propars = self.init_propars()
        if self.solver == "sc":
            propars, niter = self.solve_sc(propars)
        elif self.solver in ("newton", "m-newton"):
            opts = dict(self.solver_options)
            opts.setdefault("maxiter", 100)
            propars, niter = self.solve_newton(propars, "exact" if self.solver == "newton" else "modified", **opts)
        else:
            raise NotImplementedError(self.solver)
        self.cache["propars"] = propars
        self.cache["niter"] = niter
        self.on_molgrid_weights = True
        self._weights_on_molgrid(propars)
        charges = np.array([self.pseudo_numbers[i] - self.grid.integrate(self.cache[f"at_weights_{i}"] * self.moldens)
                            for i in range(self.natom)])
        self.cache["charges"] = charges
        self.cache["populations"] = self.numbers - charges
        self.cache["pseudo_populations"] = self.pseudo_numbers - charges
        self.cache["change"] = self.history_changes[-1]
        for k in ("propars", "entropies", "changes"):
            self.cache[f"history_{k}"] = np.array(getattr(self, f"history_{k}"))
        return self.cache

    def _weights_on_molgrid(self, propars):
        """update_at_weights(force_on_molgrid=True) (stockholder.py:352-384)."""
        promol = np.zeros(self.grid.size)
        proat = []
        for i in range(self.natom):
            work = self.basis.proatom(int(self.numbers[i]), propars[self.ranges[i] : self.ranges[i + 1]],
                                      self.radial_distances[i], 0)
            promol += work
            promol += 1e-100
            proat.append(work)
        for i in range(self.natom):
            w = proat[i] / promol
            np.clip(w, 0, 1, out=w)
            self.cache[f"at_weights_{i}"] = w
        self.cache["promoldens"] = promol


# ----------------------------------------------------------------------------
# tabulated pro-atoms: core/proatomdb.py, hirshfeld.py, hirshfeld_i.py, isa.py
# ----------------------------------------------------------------------------


class OracleRGrid:
    """A radial quadrature with qc-grid's OneDGrid.integrate (sum w prod f)."""

    def __init__(self, points, weights):
        self.points = np.asarray(points, float)
        self.weights = np.asarray(weights, float)
        self.size = self.points.size

    def integrate(self, *f):
        prod = self.weights
        for v in f:
            prod = prod * v
        return float(np.sum(prod))


def power_rgrid(rmin, rmax, npoint):
    """qc-grid PowerRTransform(rmin, rmax, npoint - 1) on UniformInteger(npoint)
    (the record loader, tests/common.py:78-110): r_i = rmin (i + 1)^p with
    p = ln(rmax / rmin) / ln(npoint), weights dr/di."""
    i = np.arange(npoint, dtype=float)
    p = np.log(rmax / rmin) / np.log(npoint)
    return OracleRGrid(rmin * (i + 1.0) ** p, p * rmin * (i + 1.0) ** (p - 1.0))


class OracleRecord:
    """ProAtomRecord's data (core/proatomdb.py:34-83)."""

    def __init__(self, number, charge, energy, rgrid, rho, deriv=None, pseudo_number=None):
        self.number = int(number)
        self.charge = int(charge)
        self.energy = float(energy)
        self.rgrid = rgrid
        self.rho = np.asarray(rho, float)
        self.deriv = None if deriv is None else np.asarray(deriv, float)
        self.pseudo_number = self.number if pseudo_number is None else pseudo_number


class OracleProAtomDB:
    """ProAtomDB.get_record/get_rgrid/get_rho/get_spline
    (core/proatomdb.py:213-388; linear combinations only)."""

    def __init__(self, records):
        best = {}
        for r in records:  # lowest energy per (number, charge), proatomdb.py:228-234
            key = (r.number, r.charge)
            if key not in best or r.energy < best[key].energy:
                best[key] = r
        self.map = best
        self.rgrids = {}
        for r in best.values():
            self.rgrids.setdefault(r.number, r.rgrid)

    def get_record(self, number, charge):
        return self.map[(int(number), int(charge))]

    def get_rgrid(self, number):
        return self.rgrids[int(number)]

    def get_charges(self, number):
        return sorted((c for (n, c) in self.map if n == int(number)), reverse=True)

    def get_rho(self, number, parameters=0, do_deriv=False):
        if isinstance(parameters, int):
            r = self.get_record(number, parameters)
            return (r.rho, r.deriv) if do_deriv else r.rho
        rho, deriv = 0.0, (0.0 if do_deriv else None)
        for charge, coeff in parameters.items():
            if coeff != 0.0:
                r = self.get_record(number, charge)
                rho = rho + coeff * r.rho
                deriv = deriv + coeff * r.deriv if (do_deriv and r.deriv is not None and deriv is not None) else None
        return (rho, deriv) if do_deriv else rho

    def get_spline(self, number, charge):
        rho, deriv = self.get_rho(number, charge, do_deriv=True)
        x = self.get_rgrid(number).points
        return CubicSpline(x, rho, True) if deriv is None else CubicHermiteSpline(x, rho, deriv, True)


def load_record_npz(path):
    """A database record from the reference's atom npz files (tests/common.py:78-110)."""
    with np.load(path) as d:
        rmin, rmax, n = d["rgrid"]
        pseudo = float(d["pseudo_number"]) if "pseudo_number" in d.files else None
        return OracleRecord(int(d["number"]), int(d["charge"]), float(d["energy"]), power_rgrid(rmin, rmax, int(n)),
                            d["dens"], d["deriv"], pseudo_number=pseudo)


class _OracleTabulated(OracleWPart):
    """Shared by the tabulated schemes: pro-atoms are splines of the radial
    distance, evaluated with scipy as the reference does."""

    def spline_proatom(self, i, spline):
        """eval_spline + 1e-100 (core/stockholder.py:266-349)."""
        return spline(self.radial_distances[i]) + 1e-100


class OracleHirshfeld(_OracleTabulated):
    """HirshfeldWPart (hirshfeld.py:127-193): neutral records, one
    update_at_weights, atom-grid populations (core/base.py:277-311)."""

    def __init__(self, *args, proatomdb=None, **kwargs):
        super().__init__(*args, **kwargs)
        self.db = proatomdb
        for i in range(self.natom):  # check_proatomdb, hirshfeld.py:34-45
            if self.db.get_record(self.numbers[i], 0).pseudo_number != self.pseudo_numbers[i]:
                raise ValueError("pseudo number mismatch")

    def rgrid(self, i):
        return self.db.get_rgrid(self.numbers[i])

    def eval_proatom(self, i, propars):
        rho, deriv = self.db.get_rho(self.numbers[i], do_deriv=True)
        rho = np.array(rho, float)
        if rho.min() < 0:  # fix_proatom_rho, core/stockholder.py:201-225
            rho[rho < 0] = 0.0
            deriv = None
        x = self.rgrid(i).points
        spline = CubicSpline(x, rho, True) if deriv is None else CubicHermiteSpline(x, rho, deriv, True)
        return self.spline_proatom(i, spline)

    def do_partitioning(self):
        self.update_at_weights(None)
        pseudo = np.zeros(self.natom)
        for i in range(self.natom):
            w = self.cache[f"at_weights_{i}"]
            if self.only_molgrid:
                pseudo[i] = self.grid.integrate(w, self.moldens)
            else:
                if self.on_molgrid:
                    w = self.to_atomic_grid(i, w)
                at = self.grid.atgrids[i]
                pseudo[i] = at.integrate(w, self.to_atomic_grid(i, self.moldens))
        self.cache["pseudo_populations"] = pseudo
        self.cache["populations"] = pseudo + self.numbers - self.pseudo_numbers
        self.cache["charges"] = self.numbers - self.cache["populations"]
        return self.cache


class OracleHI(_OracleTabulated):
    """HirshfeldIWPart (hirshfeld_i.py:36-176): the parameters are the charges."""

    def __init__(self, *args, proatomdb=None, **kwargs):
        super().__init__(*args, **kwargs)
        self.db = proatomdb

    def rgrid(self, i):
        return self.db.get_rgrid(self.numbers[i])

    def interpolation(self, i, charges):
        icharge = int(np.floor(charges[i]))  # hirshfeld_i.py:116-122
        return icharge, charges[i] - icharge

    def get_proatom_rho(self, i, charges):
        """hirshfeld_i.py:124-134 (density only)."""
        icharge, x = self.interpolation(i, charges)
        pp = self.pseudo_numbers[i] - icharge
        if pp == 1 or x == 0.0:
            return self.db.get_rho(self.numbers[i], {icharge: 1 - x})
        if pp > 1:
            return self.db.get_rho(self.numbers[i], {icharge: 1 - x, icharge + 1: x})
        raise ValueError("Requesting a pro-atom with a negative (pseudo) population")

    def eval_proatom(self, i, charges):
        """The greedy eval_proatom: the records' splines combined after
        evaluation (hirshfeld_i.py:148-158)."""
        icharge, x = self.interpolation(i, charges)
        out = self.db.get_spline(self.numbers[i], icharge)(self.radial_distances[i]) * (1 - x)
        pp = self.pseudo_numbers[i] - icharge
        if pp > 1 and x != 0.0:
            out = out + self.db.get_spline(self.numbers[i], icharge + 1)(self.radial_distances[i]) * x
        elif pp <= 0:
            raise ValueError("Requesting a pro-atom with a negative (pseudo) population")
        return out + 1e-100

    def init_propars(self):
        return np.zeros(self.natom)

    def update_atom(self, i, propars):
        """hirshfeld_i.py:166-172 with compute_pseudo_population (core/base.py:259-264)."""
        at = self.grid.atgrids[i]
        pop = at.integrate(self.cache[f"at_weights_{i}"], self.to_atomic_grid(i, self.moldens))
        propars[i] = self.pseudo_numbers[i] - pop
        self.cache["charges"][i] = propars[i]


class OracleISA(_OracleTabulated):
    """ISAWPart (isa.py:37-122): spline pro-atoms through the clipped
    spherical averages on each atom's radial grid."""

    def __init__(self, *args, **kwargs):
        super().__init__(*args, **kwargs)
        self.ranges = np.concatenate([[0], np.cumsum([self.rgrid(i).size for i in range(self.natom)])])

    def init_propars(self):
        return np.zeros(self.ranges[-1])  # isa.py:93-100

    def get_proatom_rho(self, i, propars):
        return propars[self.ranges[i] : self.ranges[i + 1]]

    def eval_proatom(self, i, propars):
        rho = np.array(self.get_proatom_rho(i, propars), float)
        rho[rho < 0] = 0.0  # fix_proatom_rho
        return self.spline_proatom(i, CubicSpline(self.rgrid(i).points, rho, True))

    def update_atom(self, i, propars):
        """isa.py:102-122."""
        rhoa, _, _ = self.atom_problem(i)
        r = np.clip(self.rgrid(i).points, 1e-100, 1e10)
        sph = np.clip(rhoa, 1e-100, np.inf)
        self.cache[f"spherical_average_{i}"] = sph
        propars[self.ranges[i] : self.ranges[i + 1]] = sph
        self.cache["charges"][i] = self.pseudo_numbers[i] - self.rgrid(i).integrate(4 * np.pi * r**2 * sph)
