"""Pins of the oracle's NITSOL restatement (oracle/bdnitsol.c) and of its
semi-implicit Euler step (integrate_se, bd_solver_m.f90:242-332).

NITSOL is vendored in the reference (src/nitsol) but cannot be run here:
the reference's build needs MKL.  The restatement is pinned by what the
mathematics fixes: linear systems against a direct solve, nonlinear systems
with roots found independently (scalar Newton on a separable system,
scipy's hybrid method on Broyden's tridiagonal function, a published test
problem), the termination codes nitsol.f documents, and the incremental
condition estimate (dlaic1) against exact 2 x 2 singular values and the
bounds it must respect.  The SE step is pinned against a dense Newton solve
of its defining equation x - dt (U(x) + D F(x)) - x0 - w = 0."""
import numpy as np
import pytest
import scipy.optimize

from helpers import fene_wca_chain, oracle_system


def test_linear_system(oracle):
    rng = np.random.default_rng(1)
    n = 40
    A = np.eye(n) * 3 + rng.normal(size=(n, n)) * 0.2
    b = rng.normal(size=n)
    want = np.linalg.solve(A, b)
    for iresup in (0, 1):
        x, iterm, info = oracle.nitsol(lambda v: A @ v - b, np.zeros(n), 1e-11, 1e-14,
                                       [0, 0, 0, 10, 0, 0, iresup, 0, 0, 0])
        assert iterm == 0
        assert np.linalg.norm(A @ x - b) <= 1e-11
        assert np.allclose(x, want, atol=1e-10)
        assert info[0] >= info[4] + 1 and info[3] >= info[4] >= 1


def test_separable_cubic(oracle):
    """F_i = x_i^3 + x_i - c_i: one real root each (monotone), found by
    scalar Newton to machine precision."""
    c = np.linspace(-3.0, 5.0, 25)
    r = np.zeros_like(c)
    for _ in range(60):
        r = r - (r ** 3 + r - c) / (3 * r ** 2 + 1)
    x, iterm, _ = oracle.nitsol(lambda v: v ** 3 + v - c, np.zeros_like(c), 1e-12, 1e-15)
    assert iterm == 0 and np.max(np.abs(x - r)) < 1e-11


@pytest.mark.parametrize("input10", [None, [0, 0, 0, 5, 0, 0, 1, 0, 0, 0], [0, 0, 0, 3, 0, 0, 1, 4, 0, 2],
                                     [0, 0, 0, 20, 0, 0, 0, 1, 0, 1], [0, 0, 0, 20, 0, 0, 0, 0, 0, 3]])
def test_broyden_tridiagonal(oracle, input10):
    """Broyden's tridiagonal function (More, Garbow, Hillstrom 1981, no. 30):
    F_i = (3 - 2 x_i) x_i - x_{i-1} - 2 x_{i+1} + 1, x_0 = x_{n+1} = 0, from
    x = -1; the root from scipy's MINPACK hybrid method."""
    n = 30

    def F(x):
        xp = np.concatenate([[0.0], x, [0.0]])
        return (3 - 2 * x) * x - xp[:-2] - 2 * xp[2:] + 1

    ref = scipy.optimize.fsolve(F, -np.ones(n), xtol=1e-14)
    x, iterm, info = oracle.nitsol(F, -np.ones(n), 1e-12, 1e-15, input10)
    assert iterm == 0
    assert np.max(np.abs(x - ref)) < 1e-10
    assert np.linalg.norm(F(x)) <= 1e-12


def test_termination_codes(oracle):
    def F(x):
        xp = np.concatenate([[0.0], x, [0.0]])
        return (3 - 2 * x) * x - xp[:-2] - 2 * xp[2:] + 1

    x0 = -np.ones(20)
    # nnimax = 1 nonlinear iteration is not enough: iterm 1
    assert oracle.nitsol(F, x0, 1e-14, 0.0, [1, 0, 0, 0, 0, 0, 0, 0, 0, 0])[1] == 1
    # illegal inputs: -k for input(k)
    assert oracle.nitsol(F, x0, 1e-10, 0.0, [-1, 0, 0, 0, 0, 0, 0, 0, 0, 0])[1] == -1
    assert oracle.nitsol(F, x0, 1e-10, 0.0, [0, 0, 0, -2, 0, 0, 0, 0, 0, 0])[1] == -4
    assert oracle.nitsol(F, x0, 1e-10, 0.0, [0, 0, 0, 0, 0, 0, 2, 0, 0, 0])[1] == -7
    assert oracle.nitsol(F, x0, 1e-10, 0.0, [0, 0, 0, 0, 0, 0, 0, 3, 0, 0])[1] == -8
    assert oracle.nitsol(F, x0, 1e-10, 0.0, [0, 0, 0, 0, 0, 0, 0, 0, -2, 0])[1] == -9
    # a failing function: iterm 2
    def bad(x):
        raise ValueError

    assert oracle.nitsol(bad, x0, 1e-10, 0.0)[1] == 2
    # already converged: no iteration
    x, iterm, info = oracle.nitsol(lambda v: v * 0.0, x0, 1e-10, 0.0)
    assert iterm == 0 and info[4] == 0 and info[0] == 1


def test_dlaic1_against_exact_singular_values(oracle):
    rng = np.random.default_rng(2)
    for _ in range(50):
        a, b, c = rng.normal(size=3) * np.array([1.0, 1.0, 1e-3 ** rng.integers(0, 3)])
        R = np.array([[a, b], [0.0, c]])
        sv = np.linalg.svd(R, compute_uv=False)
        # the incremental estimate from a 1 x 1 start is exact for 2 x 2
        big, s1, c1 = oracle.dlaic1(1, [1.0], abs(a), [b], c)
        small, s2, c2 = oracle.dlaic1(2, [1.0], abs(a), [b], c)
        assert np.isclose(big, sv[0], rtol=1e-13) and np.isclose(small, sv[1], rtol=1e-10, atol=1e-300)
        assert np.isclose(s1 * s1 + c1 * c1, 1.0) and np.isclose(s2 * s2 + c2 * c2, 1.0)
        # the returned vector is the singular vector of [[sest, 0], [alpha, gamma]]
        M = np.array([[abs(a), 0.0], [b, c]])
        assert np.isclose(np.linalg.norm(M @ [s1, c1]), big, rtol=1e-12)
        assert np.isclose(np.linalg.norm(M @ [s2, c2]), small, rtol=1e-8, atol=1e-300)
    # larger triangles: estimates stay inside [sigma_min, sigma_max]
    for _ in range(20):
        k = 8
        R = np.triu(rng.normal(size=(k, k))) + np.eye(k) * 2
        big, small = abs(R[0, 0]), abs(R[0, 0])
        xb, xs = np.array([1.0]), np.array([1.0])
        for j in range(1, k):
            nb_, s, c = oracle.dlaic1(1, xb, big, R[:j, j], R[j, j])
            xb = np.concatenate([s * xb, [c]])
            big = nb_
            ns, s, c = oracle.dlaic1(2, xs, small, R[:j, j], R[j, j])
            xs = np.concatenate([s * xs, [c]])
            small = ns
            sv = np.linalg.svd(R[:j + 1, :j + 1], compute_uv=False)
            assert big <= sv[0] * (1 + 1e-12) and small >= sv[-1] * (1 - 1e-12)
            # the estimates are attained: |R^T x| for the tracked vectors
            assert np.isclose(np.linalg.norm(R[:j + 1, :j + 1].T @ xb), big, rtol=1e-10)
            assert np.isclose(np.linalg.norm(R[:j + 1, :j + 1].T @ xs), small, rtol=1e-8)


def _dense_se_step(oracle, so, crd0, w, dt, mob, lhdia):
    """Newton with a central-difference Jacobian on x - dt (D F(x)) - x0 - w."""
    n = crd0.size

    def G(x):
        f, _, ierr = oracle.forces(so, x.reshape(-1, 3), oracle.pairs(so, x.reshape(-1, 3), "DIR"))
        assert ierr == 0
        d = mob @ f.ravel() if lhdia else f.ravel()
        return x - dt * d - crd0.ravel() - w

    x = crd0.ravel().copy()
    for _ in range(30):
        g = G(x)
        if np.linalg.norm(g) < 1e-14:
            break
        J = np.zeros((n, n))
        h = 1e-6
        for k in range(n):
            e = np.zeros(n)
            e[k] = h
            J[:, k] = (G(x + e) - G(x - e)) / (2 * h)
        x = x - np.linalg.solve(J, g)
    return x.reshape(-1, 3)


@pytest.mark.parametrize("lhdia", [False, True])
def test_se_step_solves_its_equation(oracle, lhdia):
    m = fene_wca_chain(8, 4)
    so = oracle_system(oracle, m)
    crd0 = m["crd"] - m["crd"].mean(axis=0)
    dt = 5e-3
    xi = np.random.default_rng(3).standard_normal(24)
    out, used, info = oracle.bd_run(so, crd0, 1, xi, lhdia=lhdia, mob_fctr=oracle.CHOL, tim_stp=dt,
                                    integrator=oracle.SE, se_nlmxitr=50, se_kdmax=20, se_tol=(1e-12, 1e-14))
    assert used == 24 and info.se_iterm == 0 and info.se_nni >= 1
    if lhdia:
        mob = oracle.rpy(crd0)
        bdw = oracle.brn_calc_bdw(oracle.CHOL, mob.copy(), xi.reshape(24, 1))[0]
        w = np.sqrt(2 * dt) * np.ravel(bdw)
        D = np.tril(mob) + np.tril(mob, -1).T
    else:
        w = np.sqrt(2 * dt) * xi
        D = None
    ref = _dense_se_step(oracle, so, crd0, w, dt, D, lhdia)
    ref = ref - ref.mean(axis=0)   # the step's centring of an isolated molecule
    assert np.max(np.abs(out - ref)) < 1e-11


def test_se_approaches_em_for_small_steps(oracle):
    """Without noise, semi-implicit and explicit Euler differ by O(dt^2) per
    step (with noise the gap is O(dt^1.5): the noise moves x by sqrt(dt))."""
    m = fene_wca_chain(20, 6)
    so = oracle_system(oracle, m)
    xi = np.zeros(60)
    gaps = []
    for dt in (1e-3, 5e-4):
        a, _, _ = oracle.bd_run(so, m["crd"], 1, xi, lhdia=False, tim_stp=dt)
        b, _, info = oracle.bd_run(so, m["crd"], 1, xi, lhdia=False, tim_stp=dt, integrator=oracle.SE,
                                   se_nlmxitr=50, se_kdmax=20, se_tol=(1e-13, 1e-15))
        assert info.se_iterm == 0
        gaps.append(np.max(np.abs(a - b)))
    assert 3.0 < gaps[0] / gaps[1] < 5.0   # quadratic in dt
