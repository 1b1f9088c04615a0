"""Pins for the oracle's Galerkin reduced-order operators (CPU):
experiments/periodic/paper_rollout.py:37-66,166-186 and
experiments/metaball/operators.py:20-60,137-257 restated in
oracle/neb_oracle.py, checked against identities, closed forms and
independent evaluations."""
import math

import numpy as np
import pytest

from oracle import neb_oracle as o


def _fourier_velocity(x, ks):
    """(gu, gv, gw) of the stream functions cos(k.x): u = psi_y, v = -psi_x,
    omega = -lap psi, columns per wavevector."""
    gu, gv, gw = (np.empty((len(x), len(ks))) for _ in range(3))
    for j, (kx, ky) in enumerate(ks):
        ph = kx * x[:, 0] + ky * x[:, 1]
        gu[:, j] = -ky * np.sin(ph)
        gv[:, j] = kx * np.sin(ph)
        gw[:, j] = (kx * kx + ky * ky) * np.cos(ph)
    return gu, gv, gw


def test_rollout_rhs_is_energy_skew_and_sign_sensitive():
    rng = np.random.default_rng(1)
    gu, gv, gw = rng.normal(size=(3, 500, 9))
    a = rng.normal(size=9)
    r = o.rollout_rhs(gu, gv, gw, a, 0.3)
    assert abs(a @ r) < 1e-12 * np.linalg.norm(a) * np.linalg.norm(r)
    # one term with the wrong sign is no longer skew
    u, v, om = gu @ a, gv @ a, gw @ a
    bad = gu.T @ (0.3 * om * v) + gv.T @ (0.3 * om * u)
    assert abs(a @ bad) > 1e-3 * np.linalg.norm(a) * np.linalg.norm(bad)


def test_rollout_rhs_vanishes_on_equal_wavenumber_modes_and_matches_triad():
    """Steady Euler: modes of one |k| interact to nothing; the triad
    psi = A cos x + B cos 2y forces sin x sin 2y at d psi/dt = 6AB/5: with
    u = (psi_y, -psi_x), omega = -lap psi, omega_t = -u . grad omega =
    psi_x omega_y - psi_y omega_x = 6AB sin x sin 2y and omega = 5 psi for
    that mode."""
    x = o.periodic_grid(32)
    w = (2 * np.pi / 32) ** 2
    gu, gv, gw = _fourier_velocity(x, [(1, 0), (0, 1)])
    r = o.rollout_rhs(gu, gv, gw, np.array([0.7, -1.3]), w)
    assert np.abs(r).max() < 1e-12
    # triad: basis cos x, cos 2y, and sin x sin 2y (stream functions)
    ph1, ph2 = x[:, 0], 2 * x[:, 1]
    psi = np.stack([np.cos(ph1), np.cos(ph2), np.sin(ph1) * np.sin(ph2)], 1)
    dpx = np.stack([-np.sin(ph1), 0 * ph1, np.cos(ph1) * np.sin(ph2)], 1)
    dpy = np.stack([0 * ph1, -2 * np.sin(ph2), 2 * np.sin(ph1) * np.cos(ph2)], 1)
    lap = np.array([1.0, 4.0, 5.0])
    gu, gv, gw = dpy, -dpx, psi * lap
    A, B = 0.8, -0.5
    r = o.rollout_rhs(gu, gv, gw, np.array([A, B, 0.0]), w)
    mass = w * (gu.T @ gu + gv.T @ gv)             # diagonal: the modes are orthogonal
    rate = np.linalg.solve(mass, r)                 # d a / dt of the Galerkin system
    assert abs(rate[2] - 6 * A * B / 5) < 1e-12 and np.abs(rate[:2]).max() < 1e-12


def test_etdrk4_exact_for_linear_and_fourth_order():
    lin = np.array([-0.3, -2.0, -7.5])
    c0 = np.array([1.0, -0.5, 2.0])
    c = o.etdrk4(lambda c: 0 * c, lin, c0, 0.05, 40)
    np.testing.assert_allclose(c, np.exp(lin * 2.0) * c0, rtol=1e-12)
    # c' = lam c + c^2: c(t) = lam c0 e^{lam t} / (lam + c0 (1 - e^{lam t}))
    lam, x0, T = np.array([-1.5]), np.array([0.4]), 1.0
    exact = lam * x0 * np.exp(lam * T) / (lam + x0 * (1 - np.exp(lam * T)))
    errs = [abs(o.etdrk4(lambda c: c * c, lam, x0, T / s, s) - exact)[0] for s in (16, 32)]
    assert 14 < errs[0] / errs[1] < 24, errs  # 2^4 = 16 for fourth order (8 for third)


def test_periodic_reduced_basis_orthonormal():
    x = o.periodic_grid(16)
    w = (2 * np.pi / 16) ** 2
    rng = np.random.default_rng(3)
    gu, gv, gw = _fourier_velocity(x, [(1, 0), (0, 1), (1, 1), (2, 1), (1, 2)])
    mix = rng.normal(size=(5, 5))
    bu, bv, bw, vals = o.periodic_reduced_basis(gu @ mix, gv @ mix, gw @ mix, w)
    np.testing.assert_allclose(w * (bu.T @ bu + bv.T @ bv), np.eye(5), atol=1e-10)
    np.testing.assert_allclose(w * (bw.T @ bw), np.diag(vals), atol=1e-9)
    np.testing.assert_allclose(np.sort(vals), [1, 1, 2, 5, 5], rtol=1e-10)
    with pytest.raises(ArithmeticError):
        o.periodic_reduced_basis(np.c_[gu, gu[:, :1]], np.c_[gv, gv[:, :1]], np.c_[gw, gw[:, :1]], w)


def _grid_points(n):
    """The unit periodic cube of the operator derivation (wavenumbers
    2 pi * integer, cell volume 1/n^3), x fastest."""
    ax = np.arange(n) / n
    z, y, xx = np.meshgrid(ax, ax, ax, indexing="ij")
    return xx, y, z


def test_spectral_grid_projection_curl_and_forces():
    n = 16
    tp = 2 * np.pi
    xx, y, z = _grid_points(n)
    normal = np.zeros((n, n, n, 3)); normal[..., 0] = 1.0
    chi = np.zeros((n, n, n))
    g = o.SpectralGrid(normal, chi, n, viscosity=0.01, penalty_time=1.0)
    # gradient field of sin(2 pi x) sin(4 pi y): projected to zero
    grad = np.stack((np.cos(tp * xx) * np.sin(2 * tp * y), 2 * np.sin(tp * xx) * np.cos(2 * tp * y), 0 * z))
    proj, _, _ = g.derived(grad)
    assert np.abs(proj).max() < 1e-12
    # A = (sin 4 pi y, sin 6 pi z, sin 10 pi x): u = curl A exactly, divergence
    # free and kept by the projection (wavenumbers 2, 3, 5 below n / 3)
    A = np.stack((np.sin(2 * tp * y), np.sin(3 * tp * z), np.sin(5 * tp * xx)))
    u = -tp * np.stack((3 * np.cos(3 * tp * z), 5 * np.cos(5 * tp * xx), 2 * np.cos(2 * tp * y)))
    _, cA, _ = g.derived(A)
    np.testing.assert_allclose(cA, u, atol=1e-10)
    proj, cu, lin = g.derived(u)
    np.testing.assert_allclose(proj, u, atol=1e-10)
    # curl u = curl curl A = -lap A for divergence-free A
    np.testing.assert_allclose(cu, tp ** 2 * A * np.array([4.0, 9.0, 25.0])[:, None, None, None], atol=1e-8)
    # viscous force -nu |k|^2 u with chi = 0
    np.testing.assert_allclose(lin, -0.01 * tp ** 2 * u * np.array([9.0, 25.0, 4.0])[:, None, None, None],
                               atol=1e-9)
    # the projected normal penalty -chi/eta (u.n) n: u = (sin 2 pi y, 0, 0), n = e_x, chi = 1
    g1 = o.SpectralGrid(normal, np.ones((n, n, n)), n, viscosity=0.0, penalty_time=2.0)
    f = np.stack((np.sin(tp * y), 0 * y, 0 * y))
    _, _, lin1 = g1.derived(f)
    np.testing.assert_allclose(lin1, -f / 2.0, atol=1e-12)


def test_derive_fields_zero_extends_and_gathers():
    """derive_fields on a fluid subset equals the grid operators applied to
    the zero-extended field, gathered back at the fluid points."""
    n = 8
    rng = np.random.default_rng(6)
    flat = np.sort(rng.choice(n ** 3, 300, replace=False))
    B = rng.normal(size=(300, 3, 2))
    normal = rng.normal(size=(n, n, n, 3))
    chi = (rng.uniform(size=(n, n, n)) > 0.7).astype(float)
    g = o.SpectralGrid(normal, chi, n)
    D = o.derive_fields(B, flat, g)
    u = np.zeros((n ** 3, 3)); u[flat] = B[:, :, 1]
    parts = g.derived(u.reshape(n, n, n, 3).transpose(3, 0, 1, 2))
    for q, part in enumerate(parts):
        np.testing.assert_allclose(D[:, 3 * q:3 * q + 3, 1], part.transpose(1, 2, 3, 0).reshape(-1, 3)[flat],
                                   rtol=1e-13, atol=1e-13)


def test_assemble_operators_against_pointwise_contraction():
    """Q contracted with a, a equals dV sum_p PB_i . ((B a) x (curl B a))
    evaluated pointwise; M and A are the plain weighted products."""
    rng = np.random.default_rng(4)
    N, K = 300, 6
    B = rng.normal(size=(N, 3, K))
    D = rng.normal(size=(N, 9, K))
    dV = 0.01
    M, A, Q = o.assemble_operators(B, D, dV)
    np.testing.assert_allclose(M, dV * np.einsum("pck,pcl->kl", B, B), rtol=1e-12)
    np.testing.assert_allclose(A, dV * np.einsum("pck,pcl->kl", B, D[:, 6:9]), rtol=1e-12)
    a = rng.normal(size=K)
    ua, ca = B @ a, D[:, 3:6] @ a
    direct = dV * np.einsum("pci,pc->i", D[:, 0:3], np.cross(ua, ca))
    jj, kk = np.triu_indices(K)
    np.testing.assert_allclose(Q @ (a[jj] * a[kk]), direct, rtol=1e-11)
