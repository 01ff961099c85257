"""Pins the oracle's reduced step (psfsolver.cpp:34-59) and reconstruction
(psfsolver.cpp:60) with closed forms and invariants chosen so that a dropped
term, a wrong order of operations or a transposed index fails."""
import numpy as np

import oracle as O


def _rand(K, seed):
    rng = np.random.default_rng(seed)
    C = rng.standard_normal((K, K, K)) * 0.05
    lam = rng.random(K) + 0.1
    g = rng.standard_normal(K) * 0.01
    w = rng.standard_normal(K)
    return C, lam, g, w / np.linalg.norm(w)


def test_contraction_is_the_quadratic_form():
    """y_i = sum_ab w_a C_i(a,b) w_b with C_i stored column-major (C[i, b, a] = C_i(a, b));
    an independent numpy evaluation, and permuting i changes y."""
    C, lam, g, w = _rand(7, 0)
    _, y = O.reduced_step(w, C, lam, g, 0.01, 0.0, False, return_y=True)
    ref = np.einsum("iba,a,b->i", C, w, w)
    assert np.allclose(y, ref, rtol=1e-13, atol=1e-15)
    assert not np.allclose(y, np.einsum("aib,a,b->i", C, w, w))


def test_hand_computed_step_K2():
    """K=2 by hand: y = (w^T C_0 w, w^T C_1 w), then Euler, renormalise, decay, force."""
    w = np.array([0.6, 0.8])
    C = np.zeros((2, 2, 2))
    C[0] = [[1.0, 2.0], [3.0, 4.0]]          # C[0, b, a]
    C[1] = [[0.5, 0.0], [0.0, -1.0]]
    lam = np.array([1.0, 2.0])
    g = np.array([0.1, -0.2])
    dt, nu = 0.1, 0.5
    y0 = 0.6 * 0.6 * 1 + 0.8 * 0.6 * 2 + 0.6 * 0.8 * 3 + 0.8 * 0.8 * 4   # sum_ab w_a C[0,b,a] w_b
    y1 = 0.6 * 0.6 * 0.5 + 0.8 * 0.8 * -1.0
    wn = w + dt * np.array([y0, y1])
    wn = wn * np.sqrt(1.0 / np.dot(wn, wn))
    wn = wn * np.exp(nu * lam * dt)
    wn = wn + dt * g
    out = O.reduced_step(w, C, lam, g, dt, nu, True)
    assert np.allclose(out, wn, rtol=1e-15, atol=1e-16)


def test_viscosity_only_closed_form():
    """C = 0: the renormalisation factor is exactly 1 and w_k(n) = w_k(0) exp(nu lam_k dt)^n."""
    K = 9
    _, lam, g, w0 = _rand(K, 1)
    C = np.zeros((K, K, K))
    w = w0.copy()
    for _ in range(50):
        w = O.reduced_step(w, C, lam, g, 0.02, -0.7, False)
    assert np.allclose(w, w0 * np.exp(-0.7 * lam * 0.02 * 50), rtol=1e-13)


def test_forcing_only_is_linear_in_time():
    K = 5
    _, lam, g, w0 = _rand(K, 2)
    C = np.zeros((K, K, K))
    w = w0.copy()
    for _ in range(100):
        w = O.reduced_step(w, C, lam, g, 0.01, 0.0, True)
    assert np.allclose(w, w0 + 100 * 0.01 * g, rtol=0, atol=1e-14)


def test_energy_preserved_without_forcing():
    C, lam, g, w = _rand(12, 3)
    for _ in range(1000):
        w = O.reduced_step(w, C, lam, g, 0.01, 0.0, False)
    assert abs(np.dot(w, w) - 1.0) < 1e-13


def test_renormalisation_precedes_forcing():
    """|w_{n+1} - dt g| = |w_n| when nu = 0: the energy is restored before forcing is added."""
    C, lam, g, w = _rand(10, 4)
    for _ in range(20):
        wn = O.reduced_step(w, C, lam, g, 0.05, 0.0, True)
        assert abs(np.linalg.norm(wn - 0.05 * g) - np.linalg.norm(w)) < 1e-14
        w = wn


def test_reconstruct_matches_gemv():
    rng = np.random.default_rng(5)
    K, F = 11, 301
    Phi = rng.standard_normal((K, F))
    w = rng.standard_normal(K)
    u = O.reconstruct(Phi, w)
    assert u.dtype == np.float32
    assert np.array_equal(u, (Phi.T @ w).astype(np.float32)) or \
        np.abs(u - (Phi.T @ w)).max() <= np.spacing(np.float32(np.abs(Phi.T @ w).max()))
