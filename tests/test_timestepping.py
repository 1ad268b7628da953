"""ode45 (Dormand-Prince 5(4), the time stepper around mobilityProblem) on
the CPU.  Pinned by: the tableau's order conditions and scipy's RK45 (the
same Dormand-Prince pair and continuous extension), exactness on polynomial
solutions, closed-form solutions of linear ODEs, and the CSV trajectory
layout [t; y] of SquiremersGrav.m:69,88-101."""
import importlib

import numpy as np
import pytest

ts = importlib.import_module("stokes-nnkifmm_amd.timestepping")


def test_tableau_order_conditions_and_scipy():
    from scipy.integrate._ivp.rk import RK45
    c, b = ts._C, ts._B
    A = np.zeros((7, 7))
    for i, row in enumerate(ts._A):
        A[i, :len(row)] = row
    np.testing.assert_allclose(A.sum(1), c, atol=1e-15)            # row sums = nodes
    for q in range(1, 6):                                           # bushy-tree conditions to order 5
        assert abs(b @ c ** (q - 1) - 1 / q) < 1e-15, q
    assert abs(b @ (A @ c) - 1 / 6) < 1e-15
    assert abs(b @ (A @ (A @ c)) - 1 / 24) < 1e-15
    np.testing.assert_allclose(ts._P.sum(1), b, atol=1e-15)         # extension at s = 1 is the step
    np.testing.assert_array_equal(ts._P, RK45.P)
    np.testing.assert_allclose(-ts._E, RK45.E, atol=1e-17)         # sign is a convention
    np.testing.assert_allclose(ts._A[6], b[:6], atol=0)              # FSAL


def test_single_step_matches_scipy_dense_output():
    from scipy.integrate import solve_ivp

    def f(t, y):
        return np.array([y[1], -np.sin(y[0]) + 0.1 * np.cos(t)])
    y0, h = np.array([0.7, -0.2]), 0.05
    ynew, K, _ = ts.dp45_step(f, 0.0, y0, h, f(0.0, y0))
    sol = solve_ivp(f, [0, h], y0, method="RK45", first_step=h, max_step=h, rtol=1, atol=1,
                    dense_output=True)
    np.testing.assert_allclose(ynew, sol.y[:, -1], rtol=0, atol=1e-15)
    s = np.array([0.25, 0.5, 0.75])
    np.testing.assert_allclose(ts._interp(y0, h, K, s), sol.sol(s * h).T, rtol=0, atol=1e-15)


def test_polynomial_solution_is_exact():
    # y' = 4 t^3 + 1 has y = t^4 + t + 1: within the order of the step and
    # of its 4th-order continuous extension, so every output point is exact
    t, y = ts.ode45(lambda t, y: np.array([4 * t ** 3 + 1.0]), [0.0, 2.0], [1.0])
    np.testing.assert_allclose(y[:, 0], t ** 4 + t + 1, rtol=1e-13)
    assert len(t) > 5 and np.all(np.diff(t) > 0)


@pytest.mark.parametrize("tol", [1e-3, 1e-6, 1e-9])
def test_linear_systems_against_closed_form(tol):
    lam = np.array([-1.0, -0.3])
    t, y = ts.ode45(lambda t, y: lam * y, [0, 3], [1.0, 2.0], ts.odeset("RelTol", tol, "AbsTol", tol * 1e-3))
    exact = np.array([1.0, 2.0]) * np.exp(np.outer(t, lam))
    assert np.max(np.abs(y - exact)) < 30 * tol
    # rotation (the form of the basis-vector equations db/dt = Omega x b)
    W = np.array([[0, -0.2, 0.5], [0.2, 0, -0.1], [-0.5, 0.1, 0]])
    tt = np.linspace(0, 4, 9)
    t, y = ts.ode45(lambda t, y: W @ y, tt, [1.0, 0, 0], ts.odeset("RelTol", tol, "AbsTol", tol * 1e-3))
    np.testing.assert_array_equal(t, tt)
    from scipy.linalg import expm
    exact = np.stack([expm(W * s) @ [1.0, 0, 0] for s in tt])
    assert np.max(np.abs(y - exact)) < 30 * tol


def test_refine_and_backward_integration():
    f = lambda t, y: -2 * y   # noqa: E731
    t1, y1 = ts.ode45(f, [0, 1], [1.0], ts.odeset("Refine", 1))
    t4, y4 = ts.ode45(f, [0, 1], [1.0])
    assert len(t4) == 4 * (len(t1) - 1) + 1
    np.testing.assert_array_equal(t4[::4], t1)
    np.testing.assert_array_equal(y4[::4], y1)
    tb, yb = ts.ode45(f, [1, 0], [np.exp(-2.0)], ts.odeset("RelTol", 1e-8, "AbsTol", 1e-12))
    assert tb[-1] == 0 and abs(yb[-1, 0] - 1) < 1e-7


def test_output_function_and_csv(tmp_path):
    fn = str(tmp_path / "traj.csv")
    f = lambda t, y: np.array([y[1], -y[0]])   # noqa: E731
    calls = []

    def spy(t, y, flag):
        calls.append(flag)
        return ts.trajectoryOutput(fn)(t, y, flag)
    t, y = ts.ode45(f, [0, 2], [1.0, 0.0], ts.odeset("OutputFcn", spy))
    assert calls[0] == "init" and calls[-1] == "done" and calls.count("") > 1
    data = np.loadtxt(fn, delimiter=",")
    np.testing.assert_array_equal(data[0], t[1:])
    np.testing.assert_array_equal(data[1:], y[1:].T)
    full = str(tmp_path / "full.csv")
    ts.writeTrajectory(full, t, y)
    d2 = np.loadtxt(full, delimiter=",")
    np.testing.assert_array_equal(d2, np.vstack([t[None], y.T]))
    stop = []
    t, y = ts.ode45(f, [0, 2], [1.0, 0.0], ts.odeset("OutputFcn", lambda t, y, fl: stop.append(fl) or len(stop) > 2))
    assert t[-1] < 2


def test_odeset_errors():
    with pytest.raises(ValueError):
        ts.odeset("RelTol")
    with pytest.raises(ValueError):
        ts.odeset("NoSuchOption", 1)
    with pytest.raises(ValueError):
        ts.ode45(lambda t, y: y, [0, 0], [1.0])
    assert ts.odeset("reltol", 1e-5)["RelTol"] == 1e-5
