"""Pins for the FP64 oracle: closed forms, finite differences, library routines.

Each test is chosen so that a plausible slip in the oracle (a dropped term, a
wrong sign or index, a transposed operand) fails it.  The reference stores
no arrays for this path: its known-answer tests are restated here --
experiments/2d/_backend/periodic_core/vortex_merger/canonical_gauge/scripts/
verify_analytic.py:48-66,89-97 (Fourier basis Gram = I), tests/
test_metaball_operators.py:22 (blocked Gram = dense), experiments/2d/
check_cpu.py:27-33 (projection identity), tests/test_network_backends.py:74-95
(reverse vs directional FD, first Adam step), experiments/metaball/_backend/
flow/cases/metaball/potential_scalar/frozen/native_cupy_verify.py:49-60
(manufactured curl / vorticity / divergence).
"""
import math

import numpy as np
import pytest

from oracle import neb_oracle as o


def _net(d, F, widths, K, seed=0, act="tanh", omegas=None):
    ncomp = 1 if d == 2 else 3
    sizes = [2 * F * d] + list(widths) + [ncomp * K]
    net = o.init_net(sizes, seed, act, omegas)
    rng = np.random.default_rng(seed + 1)
    for b in net.b:  # nonzero biases exercise the bias-on-value-rows rule
        b[:] = rng.normal(size=b.shape) * 0.3
    return net


def _eval(net, x, F, lo, hi):
    out, _ = o.forward(net, o.feature_jets(x, F, lo, hi))
    return out


def test_pairs_and_channels():
    assert o.pairs(2) == [(0, 0), (0, 1), (1, 1)]
    assert o.pairs(3) == [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
    assert o.channels(2) == 6 and o.channels(3) == 10


def test_feature_jets_closed_form_and_fd():
    rng = np.random.default_rng(3)
    x = rng.uniform(0, 2 * np.pi, (7, 2))
    j = o.feature_jets(x, 3, 0.0, 2 * np.pi)
    # on [0, 2pi) the phase is exactly h*x (HB2P:576-589)
    for a in range(2):
        for h in range(1, 4):
            col = a * 6 + 2 * (h - 1)
            np.testing.assert_allclose(j.value[:, col], np.sin(h * x[:, a]), atol=1e-14)
            np.testing.assert_allclose(j.value[:, col + 1], np.cos(h * x[:, a]), atol=1e-14)
            np.testing.assert_allclose(j.grad[:, col, a], h * np.cos(h * x[:, a]), atol=1e-13)
            np.testing.assert_allclose(j.hess[:, col + 1, 0 if a == 0 else 2],
                                       -h * h * np.cos(h * x[:, a]), atol=1e-12)
    # 3D on [-1,1]: FD of value -> grad and grad -> hess
    x3 = rng.uniform(-1, 1, (5, 3))
    j3 = o.feature_jets(x3, 2, -1.0, 1.0)
    eps = 1e-6
    for a in range(3):
        e = np.zeros(3); e[a] = eps
        jp, jm = o.feature_jets(x3 + e, 2, -1.0, 1.0), o.feature_jets(x3 - e, 2, -1.0, 1.0)
        np.testing.assert_allclose((jp.value - jm.value) / (2 * eps), j3.grad[..., a], atol=2e-7)
        for p, (i, k) in enumerate(o.pairs(3)):
            if k == a:
                np.testing.assert_allclose((jp.grad[..., i] - jm.grad[..., i]) / (2 * eps),
                                           j3.hess[..., p], atol=5e-6)


@pytest.mark.parametrize("d,act", [(2, "tanh"), (3, "tanh"), (2, "sin")])
def test_forward_jets_match_spatial_fd(d, act):
    F = 2
    lo, hi = (0.0, 2 * np.pi) if d == 2 else (-1.0, 1.0)
    omegas = [1.7, 1.3] if act == "sin" else None
    net = _net(d, F, [9, 7], 3, seed=5, act=act, omegas=omegas)
    rng = np.random.default_rng(11)
    x = rng.uniform(lo + 0.1, hi - 0.1, (6, d))
    out = _eval(net, x, F, lo, hi)
    eps = 1e-5
    for a in range(d):
        e = np.zeros(d); e[a] = eps
        op, om = _eval(net, x + e, F, lo, hi), _eval(net, x - e, F, lo, hi)
        np.testing.assert_allclose((op.value - om.value) / (2 * eps), out.grad[..., a], rtol=1e-7, atol=1e-8)
        for p, (i, k) in enumerate(o.pairs(d)):
            if k == a:
                np.testing.assert_allclose((op.grad[..., i] - om.grad[..., i]) / (2 * eps),
                                           out.hess[..., p], rtol=1e-6, atol=1e-7)


def test_independent_scalar_forward():
    """Independent brute-force forward (no jets) for the value channel
    (tests/test_network_backends.py:56-71 pattern)."""
    net = _net(2, 3, [8, 8], 4, seed=2)
    x = np.random.default_rng(0).uniform(0, 2 * np.pi, (5, 2))
    f = np.concatenate([np.stack((np.sin(h * x[:, a]), np.cos(h * x[:, a])), 1)
                        for a in range(2) for h in range(1, 4)], 1)
    h = f
    for W, b in zip(net.W[:-1], net.b[:-1]):
        h = np.tanh(h @ W.T + b)
    ref = h @ net.W[-1].T + net.b[-1]
    np.testing.assert_allclose(_eval(net, x, 3, 0.0, 2 * np.pi).value, ref, rtol=1e-13, atol=1e-13)


def test_activation_reverse_is_transpose():
    rng = np.random.default_rng(4)
    for act in ("tanh", "sin"):
        for d in (2, 3):
            P = d * (d + 1) // 2
            pre = o.Jet(rng.normal(size=(5, 4)), rng.normal(size=(5, 4, d)), rng.normal(size=(5, 4, P)))
            dp = o.Jet(rng.normal(size=(5, 4)), rng.normal(size=(5, 4, d)), rng.normal(size=(5, 4, P)))
            adj = o.Jet(rng.normal(size=(5, 4)), rng.normal(size=(5, 4, d)), rng.normal(size=(5, 4, P)))
            eps = 1e-6
            yp = o.activation_jet(act, o.Jet(pre.value + eps * dp.value, pre.grad + eps * dp.grad, pre.hess + eps * dp.hess))
            ym = o.activation_jet(act, o.Jet(pre.value - eps * dp.value, pre.grad - eps * dp.grad, pre.hess - eps * dp.hess))
            lhs = sum(np.sum(a * (p - m)) for a, p, m in zip(adj.pack(), yp.pack(), ym.pack())) / (2 * eps)
            back = o.activation_jet_reverse(act, pre, adj)
            rhs = sum(np.sum(a * b) for a, b in zip(back.pack(), dp.pack()))
            assert abs(lhs - rhs) < 1e-7 * max(1, abs(rhs))


def test_lift2d_fourier_closed_form():
    """psi = sin(ax+by): u = (psi_y, -psi_x), omega = (a^2+b^2) psi."""
    x = np.random.default_rng(1).uniform(0, 6, (9, 2))
    a, b = 2.0, -3.0
    ph = a * x[:, 0] + b * x[:, 1]
    s, c = np.sin(ph), np.cos(ph)
    jet = o.Jet(s[:, None], np.stack((a * c, b * c), -1)[:, None],
                np.stack((-a * a * s, -a * b * s, -b * b * s), -1)[:, None])
    u, w = o.lift2d(jet)
    np.testing.assert_allclose(u[:, 0, 0], b * c, atol=1e-14)
    np.testing.assert_allclose(u[:, 1, 0], -a * c, atol=1e-14)
    np.testing.assert_allclose(w[:, 0], (a * a + b * b) * s, atol=1e-13)


def test_lift3d_manufactured_curl():
    """A = (sin 2y, sin 3z, sin 5x): exact curl, grad u, div u = 0
    (native_cupy_verify.py:49-60)."""
    rng = np.random.default_rng(2)
    p = rng.uniform(-1, 1, (11, 3))
    x, y, z = p.T
    n = len(p)
    val = np.stack((np.sin(2 * y), np.sin(3 * z), np.sin(5 * x)), 1)
    g = np.zeros((n, 3, 3)); h = np.zeros((n, 3, 6))
    g[:, 0, 1] = 2 * np.cos(2 * y); h[:, 0, 3] = -4 * np.sin(2 * y)
    g[:, 1, 2] = 3 * np.cos(3 * z); h[:, 1, 5] = -9 * np.sin(3 * z)
    g[:, 2, 0] = 5 * np.cos(5 * x); h[:, 2, 0] = -25 * np.sin(5 * x)
    u, du = o.lift3d(o.Jet(val, g, h))
    exact = np.stack((-3 * np.cos(3 * z), -5 * np.cos(5 * x), -2 * np.cos(2 * y)), 1)
    np.testing.assert_allclose(u[:, :, 0], exact, atol=1e-13)
    grad_exact = np.zeros((n, 3, 3))
    grad_exact[:, 0, 2] = 9 * np.sin(3 * z)
    grad_exact[:, 1, 0] = 25 * np.sin(5 * x)
    grad_exact[:, 2, 1] = 4 * np.sin(2 * y)
    np.testing.assert_allclose(du[..., 0], grad_exact, atol=1e-12)
    np.testing.assert_allclose(np.trace(du[..., 0], axis1=1, axis2=2), 0, atol=1e-12)


@pytest.mark.parametrize("d", [2, 3])
def test_lift_reverse_adjoint_identity(d):
    rng = np.random.default_rng(6)
    n, K = 5, 4
    P = d * (d + 1) // 2
    KC = K if d == 2 else 3 * K
    jet = o.Jet(rng.normal(size=(n, KC)), rng.normal(size=(n, KC, d)), rng.normal(size=(n, KC, P)))
    if d == 2:
        u, w = o.lift2d(jet)
        gu, gw = rng.normal(size=u.shape), rng.normal(size=w.shape)
        back = o.lift2d_reverse(gu, gw)
        lhs = np.sum(u * gu) + np.sum(w * gw)
    else:
        u, du = o.lift3d(jet)
        gu, gd = rng.normal(size=u.shape), rng.normal(size=du.shape)
        back = o.lift3d_reverse(gu, gd)
        lhs = np.sum(u * gu) + np.sum(du * gd)
    rhs = sum(np.sum(a * b) for a, b in zip(back.pack(), jet.pack()))
    np.testing.assert_allclose(lhs, rhs, rtol=1e-12)


def test_fourier_basis_gram_identity_and_eigenvalues():
    """verify_analytic.py:48-66,89-97: the normalised Fourier basis on the
    n x n grid has velocity Gram = I and stiffness Gram = diag(kx^2+ky^2)."""
    n, k = 32, 16
    extent = int(np.ceil(np.sqrt(k)))
    wp = [(x, y) for x in range(extent + 1) for y in range(-extent, extent + 1) if x > 0 or (x == 0 and y > 0)]
    wp = np.array(sorted(wp, key=lambda p: (p[0] ** 2 + p[1] ** 2, p[0], p[1]))[:k // 2], dtype=float)
    pts = o.periodic_grid(n)
    ph = pts @ wp.T
    lam = np.repeat(np.sum(wp * wp, 1), 2)
    scale = np.sqrt(2) / (2 * np.pi)
    # psi_k = phi_k / sqrt(lam_k), phi = (cos, sin) pairs
    val = np.stack((np.cos(ph), np.sin(ph)), -1).reshape(n * n, k) * scale
    d1 = np.stack((-np.sin(ph), np.cos(ph)), -1).reshape(n * n, k) * scale
    kx, ky = np.repeat(wp[:, 0], 2), np.repeat(wp[:, 1], 2)
    grad = np.stack((d1 * kx, d1 * ky), -1)
    hess = np.stack((-val * kx * kx, -val * kx * ky, -val * ky * ky), -1)
    s = 1 / np.sqrt(lam)
    jet = o.Jet(val * s, grad * s[:, None], hess * s[:, None])
    u, w = o.lift2d(jet)
    weight = (2 * np.pi / n) ** 2
    M = o.gram(o.field_rows(u), weight, block=97)
    S = o.gram(w, weight)
    np.testing.assert_allclose(M, np.eye(k), atol=1e-12)
    np.testing.assert_allclose(S, np.diag(lam), atol=1e-11)
    vals, _ = o.generalized_eigh(S, M)
    np.testing.assert_allclose(vals, np.sort(lam), rtol=1e-12)


def test_gram_blocked_equals_dense_and_weights():
    rng = np.random.default_rng(17)
    B = rng.normal(size=(137, 5))
    w = rng.uniform(0.1, 1, 137)
    np.testing.assert_allclose(o.gram(B, w, block=31), B.T @ (w[:, None] * B), rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(o.gram(B, 0.25, block=7), 0.25 * np.einsum("ri,rj->ij", B, B), rtol=1e-13, atol=1e-13)


def test_spectral_loss_gradient_fd_and_gate():
    rng = np.random.default_rng(8)
    K = 6
    A = rng.normal(size=(20, K)); M = A.T @ A / 20
    B = rng.normal(size=(30, K)); S = B.T @ B / 30
    L, dM, dS = o.spectral_loss(M, S, lam=1e-2)
    np.testing.assert_allclose(L, np.trace(np.linalg.solve(M, S)) + 1e-2 * np.sum((M - np.eye(K)) ** 2), rtol=1e-12)
    eps = 1e-6
    for _ in range(4):
        E = rng.normal(size=(K, K)); E = E + E.T
        Lp = o.spectral_loss(M + eps * E, S)[0]; Lm = o.spectral_loss(M - eps * E, S)[0]
        np.testing.assert_allclose((Lp - Lm) / (2 * eps), np.sum(dM * E), rtol=1e-6)
        Lp = o.spectral_loss(M, S + eps * E)[0]; Lm = o.spectral_loss(M, S - eps * E)[0]
        np.testing.assert_allclose((Lp - Lm) / (2 * eps), np.sum(dS * E), rtol=1e-6)
    sing = A[:, :3] @ rng.normal(size=(3, K))
    with pytest.raises(FloatingPointError):
        o.spectral_loss(sing.T @ sing, S)
    # the ridge makes the rank-deficient case well posed
    o.spectral_loss(sing.T @ sing, S, ridge=1e-3)


def _step_loss(net, x0, w, d, lam, ridge):
    out, tape = o.forward(net, x0, record=True)
    if d == 2:
        u, om = o.lift2d(out)
        Pm, Ps = o.field_rows(u), om
    else:
        u, du = o.lift3d(out)
        Pm, Ps = o.field_rows(u), o.field_rows(du)
    M, S = o.gram(Pm, w), o.gram(Ps, w)
    L, dM, dS = o.spectral_loss(M, S, lam, ridge)
    return L, (out, tape, Pm, Ps, dM, dS, u)


@pytest.mark.parametrize("d", [2, 3])
def test_spectral_step_reverse_matches_central_fd(d):
    """Config-C analogue at oracle size: handwritten reverse through the
    spectral loss, lifts and second-order jets vs central FD (h=1e-3) along
    random unit parameter directions (tests/test_network_backends.py:74-90)."""
    F, K = 2, 3
    lo, hi = (0.0, 2 * np.pi) if d == 2 else (-1.0, 1.0)
    net = _net(d, F, [10, 8], K, seed=9)
    rng = np.random.default_rng(12)
    x = rng.uniform(lo, hi, (40, d))
    w = 1.0 / len(x)
    x0 = o.feature_jets(x, F, lo, hi)
    lam, ridge = 1e-2, 0.0
    L, (out, tape, Pm, Ps, dM, dS, u) = _step_loss(net, x0, w, d, lam, ridge)
    gPm = o.field_adjoint(Pm, w, dM)
    gPs = o.field_adjoint(Ps, w, dS)
    n = len(x)
    if d == 2:
        adj = o.lift2d_reverse(gPm.reshape(n, 2, K), gPs)
    else:
        adj = o.lift3d_reverse(gPm.reshape(n, 3, K), gPs.reshape(n, 3, 3, K))
    gW, gb = o.backward(net, tape, adj)
    params = net.W + net.b
    grads = gW + gb
    h = 1e-3
    for _ in range(8):
        dirs = [rng.normal(size=p.shape) for p in params]
        nrm = math.sqrt(sum(np.sum(q * q) for q in dirs))
        dirs = [q / nrm for q in dirs]
        exact = sum(np.sum(g * q) for g, q in zip(grads, dirs))
        saved = [p.copy() for p in params]
        for p, q in zip(params, dirs):
            p += h * q
        Lp = _step_loss(net, x0, w, d, lam, ridge)[0]
        for p, s0, q in zip(params, saved, dirs):
            p[...] = s0 - h * q
        Lm = _step_loss(net, x0, w, d, lam, ridge)[0]
        for p, s0 in zip(params, saved):
            p[...] = s0
        fd = (Lp - Lm) / (2 * h)
        assert abs(fd - exact) <= 1e-4 * max(abs(exact), 1e-3 * abs(L)), (fd, exact)


def test_projection_identity_and_dense_solve():
    """experiments/2d/check_cpu.py:27-33 and PP:179-208."""
    rng = np.random.default_rng(77)
    B = rng.normal(size=(137, 8)); T = rng.normal(size=(137, 20)); w = rng.uniform(.1, 1, 137)
    M = o.gram(B, w)
    r = o.project(M, B, T, w, block=31)
    C = np.linalg.solve(M, B.T @ (w[:, None] * T))
    np.testing.assert_allclose(r["coefficients"], C, rtol=1e-12)
    d = np.sum(w[:, None] * (B @ C - T) ** 2, 0)
    np.testing.assert_allclose(d, np.sum(w[:, None] * T * T, 0) - np.sum(C * r["loads"], 0), rtol=1e-12)
    np.testing.assert_allclose(r["direct_relative_L2"], r["Gram_identity_errors"], atol=1e-12)
    assert np.all(r["normal_residuals"] < 1e-12)


def test_adam_first_step_closed_form():
    rng = np.random.default_rng(82)
    ps = [rng.normal(size=(3, 4)), rng.normal(size=4)]
    gs = [rng.normal(size=(3, 4)), rng.normal(size=4)]
    out, norm = o.adam(ps, gs, [np.zeros_like(p) for p in ps], [np.zeros_like(p) for p in ps], 1, 1e-4, clip=1e6)
    for (p1, _, _), p0, g in zip(out, ps, gs):
        np.testing.assert_allclose(p1, p0 - 1e-4 * g / (np.abs(g) + 1e-8), rtol=1e-12, atol=1e-14)
    with pytest.raises(FloatingPointError):
        o.adam(ps, [g * np.nan for g in gs], [np.zeros_like(p) for p in ps], [np.zeros_like(p) for p in ps], 1, 1e-4)


def test_divergence_free_targets_fd():
    """Analytic curl targets vs FD of the same construction; div = 0."""
    rng = np.random.default_rng(5)
    p = rng.uniform(-1, 1, (6, 3))
    u = o.fourier_potential_targets(p, 2, seed=3)
    eps = 1e-5
    div = np.zeros((6, 2))
    for a in range(3):
        e = np.zeros(3); e[a] = eps
        up = o.fourier_potential_targets(p + e, 2, seed=3)
        um = o.fourier_potential_targets(p - e, 2, seed=3)
        div += (up[:, a] - um[:, a]) / (2 * eps)
    np.testing.assert_allclose(div, 0, atol=1e-6 * np.abs(u).max())


def test_blob_domain_points():
    pts, centres = o.blob_domain_points(1000, seed=0, batch=4096)
    assert pts.shape == (1000, 3) and np.all(np.abs(pts) <= 1)
    s = sum(np.exp(-np.sum((pts - c) ** 2, 1) / (2 * 0.35 ** 2)) for c in centres)
    assert np.all(s > 0.5)


def _peel_fields(seed=5, n=400, k=9, rank_ok=True):
    """Random velocity-like columns U [2n, k] and derivative-like columns D."""
    rng = np.random.default_rng(seed)
    U = rng.normal(size=(2 * n, k)) @ np.diag(rng.uniform(0.5, 2.0, k))
    D = U @ rng.normal(size=(k, k)) + 0.3 * rng.normal(size=(2 * n, k))
    return U, D


def test_coefficient_peel_equals_literal_peel():
    """The coefficient form (VCORE:48-69) and the mode-by-mode MGS form
    (VCORE:107-119) compute the same losses and field cotangents:
    GU = w U Cu and GD = w D Cd (VCORE:94-95)."""
    U, D = _peel_fields()
    w = 1.0 / 400
    M, S = o.gram(U, w), o.gram(D, w)
    p = o.coefficient_peel(M, S)
    losses, GU, GD, Q, V, order = o.literal_peel(U, D, w)
    np.testing.assert_array_equal(p["order"], order)
    np.testing.assert_allclose(p["losses"], losses, rtol=1e-11)
    np.testing.assert_allclose(w * U @ p["Cu"], GU, rtol=1e-9, atol=1e-12 * np.abs(GU).max())
    np.testing.assert_allclose(w * D @ p["Cd"], GD, rtol=1e-9, atol=1e-12 * np.abs(GD).max())
    # the peeled modes are U T (columns in sorted order): w-orthonormal, and
    # the same as the literal form's normalised residuals
    np.testing.assert_allclose(U @ p["T"], Q, atol=1e-10)
    np.testing.assert_allclose(o.gram(U @ p["T"], w), np.eye(9), atol=1e-12)
    Ts = p["T"][p["order"]]
    assert np.abs(np.tril(Ts, -1)).max() == 0.0       # upper triangular in sorted order


def test_literal_peel_cotangents_match_fd_with_detached_prefix():
    """Each mode's cotangent is the derivative of its own loss with the
    earlier peeled modes held constant: central FD on that mode's columns."""
    U, D = _peel_fields(seed=8, n=150, k=5)
    w = 1.0 / 150
    losses, GU, GD, Q, V, order = o.literal_peel(U, D, w)
    rng = np.random.default_rng(3)

    def own_loss(r, u, d):
        q, v = Q[:, :r], V[:, :r]
        u, d = u.copy(), d.copy()
        for _ in range(2):
            a = w * (q.T @ u)
            u -= q @ a
            d -= v @ a
        return float(d @ d) / float(u @ u)

    h = 1e-6
    for r, index in enumerate(order):
        du, dd = rng.normal(size=U.shape[0]), rng.normal(size=D.shape[0])
        exact = GU[:, index] @ du + GD[:, index] @ dd
        fd = (own_loss(r, U[:, index] + h * du, D[:, index] + h * dd)
              - own_loss(r, U[:, index] - h * du, D[:, index] - h * dd)) / (2 * h)
        assert abs(fd - exact) <= 1e-6 * max(1.0, abs(exact)), (r, fd, exact)
        assert abs(own_loss(r, U[:, index], D[:, index]) - losses[r]) < 1e-12 * losses[r]


def test_coefficient_peel_fourier_closed_form_and_gates():
    """Orthonormal Fourier stream functions: the peel sorts the modes by
    |k|^2 and returns losses = |k|^2 (the stiffness eigenvalues), T = I
    permuted (verify_analytic.py:89-97 basis).  A repeated mode fails the
    rank gate."""
    n = 32
    x = o.periodic_grid(n)
    w = (2 * np.pi / n) ** 2
    ks = [(3, 1), (1, 0), (0, 2), (2, 2), (1, 1)]
    cols_u, cols_d, lam = [], [], []
    for kx, ky in ks:
        kk = kx * kx + ky * ky
        a = 1.0 / (math.pi * math.sqrt(2.0 * kk))          # unit velocity norm
        ph = kx * x[:, 0] + ky * x[:, 1]
        cols_u.append(np.column_stack((a * ky * np.cos(ph), -a * kx * np.cos(ph))).reshape(-1))
        cols_d.append(a * kk * np.sin(ph))                  # -lap psi for psi = a sin(ph)
        lam.append(kk)
    U, D = np.column_stack(cols_u), np.column_stack(cols_d)
    M, S = o.gram(U, w), o.gram(D, w)
    np.testing.assert_allclose(M, np.eye(5), atol=1e-12)
    p = o.coefficient_peel(M, S)
    np.testing.assert_array_equal(p["order"], np.argsort(lam, kind="stable"))
    np.testing.assert_allclose(p["losses"], np.sort(lam), rtol=1e-12)
    np.testing.assert_allclose(np.abs(p["T"][p["order"]]), np.eye(5), atol=1e-12)
    Ud = np.column_stack((U, U[:, 2]))
    Dd = np.column_stack((D, D[:, 2]))
    with pytest.raises((FloatingPointError, np.linalg.LinAlgError)):
        o.coefficient_peel(o.gram(Ud, w), o.gram(Dd, w))


# --------------------------------------------------------------------------
# 2D causal velocity Rayleigh peel (HB2P:2935-3380)
# --------------------------------------------------------------------------
def _velocity_rows(out):
    n, K = out.value.shape
    grad = out.grad.transpose(0, 2, 1).reshape(2 * n, K)      # rows 2p + (x, y)
    hess = out.hess.transpose(0, 2, 1).reshape(3 * n, K)      # rows 3p + (xx, xy, yy)
    return grad, hess


def test_lift2d_hessian_metric_and_adjoint():
    """The Hessian rows' Gram is the full symmetric Hessian's Frobenius Gram
    (every entry, the xy pair twice), and the reverse is the transpose."""
    rng = np.random.default_rng(31)
    net = _net(2, 2, [12], 5, seed=3)
    x = rng.uniform(0, 2 * np.pi, (40, 2))
    out = _eval(net, x, 2, 0.0, 2 * np.pi)
    u, h = o.lift2d_hessian(out)
    full = np.empty((40, 5, 2, 2))
    full[..., 0, 0], full[..., 1, 1] = out.hess[..., 0], out.hess[..., 2]
    full[..., 0, 1] = full[..., 1, 0] = out.hess[..., 1]
    S_full = np.einsum("nkab,nlab->kl", full, full)
    S_rows = o.field_rows(h).T @ o.field_rows(h)
    np.testing.assert_allclose(S_rows, S_full, rtol=1e-12)
    M = o.field_rows(u).T @ o.field_rows(u)
    np.testing.assert_allclose(M, np.einsum("nka,nla->kl", out.grad, out.grad), rtol=1e-12)
    gu, gh = rng.normal(size=u.shape), rng.normal(size=h.shape)
    adj = o.lift2d_hessian_reverse(gu, gh)
    lhs = np.sum(u * gu) + np.sum(h * gh)
    rhs = np.sum(out.grad * adj.grad) + np.sum(out.hess * adj.hess)
    assert abs(lhs - rhs) < 1e-10 * abs(lhs)


def _velocity_fields(seed=5, n=300, K=7):
    rng = np.random.default_rng(seed)
    net = _net(2, 2, [16, 16], K, seed=seed)
    x = rng.uniform(0, 2 * np.pi, (n, 2))
    return _velocity_rows(_eval(net, x, 2, 0.0, 2 * np.pi))


def test_velocity_peel_equals_cholesky_coefficient_peel():
    """The literal two-pass velocity peel equals the sorted Cholesky peel of
    its Grams (M = w G^T G, S = w H^T diag(1,2,1) H): Rayleigh quotients,
    causal order and the detached-prefix cotangents w G Cu, w metric H Cd."""
    grad, hess = _velocity_fields()
    n = grad.shape[0] // 2
    w = 1.0 / n
    metric = np.tile([1.0, 2.0, 1.0], n)
    lit = o.velocity_rayleigh_peel(grad, hess, w)
    M = w * grad.T @ grad
    S = w * hess.T @ (metric[:, None] * hess)
    cp = o.coefficient_peel(M, S, ortho_tol=1e-3)
    np.testing.assert_array_equal(lit["order"], cp["order"])
    np.testing.assert_allclose(lit["rayleigh"], cp["losses"], rtol=1e-10)
    np.testing.assert_allclose(lit["G_grad"], w * grad @ cp["Cu"], rtol=1e-8, atol=1e-12 * np.abs(lit["G_grad"]).max())
    np.testing.assert_allclose(lit["G_hess"], w * metric[:, None] * (hess @ cp["Cd"]), rtol=1e-8,
                               atol=1e-12 * np.abs(lit["G_hess"]).max())
    assert lit["orth"].max() < 1e-8


def test_velocity_peel_cotangents_match_fd_with_detached_prefix():
    """Each mode's cotangent is the central difference of its own Rayleigh
    quotient with the earlier peeled modes held (fixed order)."""
    grad, hess = _velocity_fields(seed=7, K=5)
    n = grad.shape[0] // 2
    w = 1.0 / n
    base = o.velocity_rayleigh_peel(grad, hess, w, order="fixed")
    rng = np.random.default_rng(2)
    for r in (0, 2, 4):
        dg = rng.normal(size=2 * n)
        dh = rng.normal(size=3 * n)
        h = 1e-6
        vals = []
        for sgn in (1.0, -1.0):
            g2, h2 = grad.copy(), hess.copy()
            g2[:, r] += sgn * h * dg
            h2[:, r] += sgn * h * dh
            vals.append(o.velocity_rayleigh_peel(g2, h2, w, order="fixed")["rayleigh"][r])
        fd = (vals[0] - vals[1]) / (2 * h)
        exact = base["G_grad"][:, r] @ dg + base["G_hess"][:, r] @ dh
        assert abs(fd - exact) <= 1e-6 * max(abs(exact), 1e-8), (r, fd, exact)


def test_velocity_peel_fourier_closed_form():
    """Orthogonal Fourier stream functions on the periodic grid: the peel
    leaves them as they are and the Rayleigh quotients are the sorted |k|^2
    (grad u Frobenius / velocity mass of sin, cos (k.x))."""
    x = o.periodic_grid(32)
    ks = [(1, 0), (0, 2), (1, 1), (2, 1), (3, 0), (1, 2)]
    n = len(x)
    grad = np.zeros((2 * n, len(ks)))
    hess = np.zeros((3 * n, len(ks)))
    for j, (kx, ky) in enumerate(ks):
        ph = kx * x[:, 0] + ky * x[:, 1]
        c, s = np.cos(ph), np.sin(ph)
        grad[0::2, j], grad[1::2, j] = kx * c, ky * c
        hess[0::3, j], hess[1::3, j], hess[2::3, j] = -kx * kx * s, -kx * ky * s, -ky * ky * s
    w = (2 * np.pi / 32) ** 2
    res = o.velocity_rayleigh_peel(grad, hess, w)
    k2 = np.array([kx * kx + ky * ky for kx, ky in ks], dtype=float)
    np.testing.assert_allclose(res["rayleigh"], np.sort(k2), rtol=1e-12)
    np.testing.assert_array_equal(res["order"], np.argsort(k2, kind="stable"))
    np.testing.assert_allclose(res["relative_mass"], 1.0, rtol=1e-12)


def _stream_fields(seed=11, n=300, K=7):
    rng = np.random.default_rng(seed)
    net = _net(2, 2, [16, 16], K, seed=seed)
    x = rng.uniform(0, 2 * np.pi, (n, 2))
    value, g = o.lift2d_stream(_eval(net, x, 2, 0.0, 2 * np.pi))
    return value, o.field_rows(g)


def test_stream_lift_rows():
    """Mass rows are psi itself, stiffness rows (psi_x, psi_y) point-major;
    the lift's transpose is its adjoint."""
    rng = np.random.default_rng(4)
    net = _net(2, 2, [12], 5, seed=4)
    out = _eval(net, rng.uniform(0, 2 * np.pi, (40, 2)), 2, 0.0, 2 * np.pi)
    v, g = o.lift2d_stream(out)
    np.testing.assert_array_equal(v, out.value)
    np.testing.assert_array_equal(o.field_rows(g)[0::2], out.grad[..., 0])
    np.testing.assert_array_equal(o.field_rows(g)[1::2], out.grad[..., 1])
    gv, gg = rng.normal(size=v.shape), rng.normal(size=g.shape)
    back = o.lift2d_stream_reverse(gv, gg)
    lhs = np.sum(gv * v) + np.sum(gg * g)
    rhs = np.sum(back.value * out.value) + np.sum(back.grad * out.grad) + np.sum(back.hess * out.hess)
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_stream_peel_equals_cholesky_coefficient_peel():
    """The literal two-pass streamfunction peel equals the sorted Cholesky
    peel of the centered Grams (M = w Vc^T Vc, S = w G^T G): Rayleigh
    quotients, causal order, masses and the detached-prefix cotangents
    w Vc Cu (already mean-free) and w G Cd."""
    value, grad = _stream_fields()
    n = value.shape[0]
    w = 1.0 / n
    lit = o.streamfunction_rayleigh_peel(value, grad, w)
    vc = value - value.mean(0)
    M = w * vc.T @ vc
    S = w * grad.T @ grad
    cp = o.coefficient_peel(M, S, ortho_tol=1e-3, rel_min=0.0)
    np.testing.assert_array_equal(lit["order"], cp["order"])
    np.testing.assert_allclose(lit["rayleigh"], cp["losses"], rtol=1e-10)
    norm2 = 1.0 / np.diag(cp["T"][cp["order"]]) ** 2
    np.testing.assert_allclose(lit["mass"], norm2, rtol=1e-10)
    np.testing.assert_allclose(lit["projection"], np.diag(M)[cp["order"]] / norm2 - 1.0, rtol=1e-8, atol=1e-12)
    np.testing.assert_allclose(lit["G_value"], w * vc @ cp["Cu"], rtol=1e-8,
                               atol=1e-12 * np.abs(lit["G_value"]).max())
    np.testing.assert_allclose(lit["G_grad"], w * grad @ cp["Cd"], rtol=1e-8, atol=1e-12 * np.abs(lit["G_grad"]).max())
    assert lit["orth"].max() < 1e-8


def test_stream_peel_cotangents_match_fd_with_detached_prefix():
    """Each mode's cotangent (through the centering) is the central
    difference of its own Rayleigh quotient in the raw, uncentered columns,
    with the earlier peeled modes held (fixed order)."""
    value, grad = _stream_fields(seed=12, K=5)
    n = value.shape[0]
    w = 1.0 / n
    base = o.streamfunction_rayleigh_peel(value, grad, w, order="fixed")
    rng = np.random.default_rng(3)
    for r in (0, 2, 4):
        dv = rng.normal(size=n) + 0.7          # a shift too: the centering removes it
        dg = rng.normal(size=2 * n)
        h = 1e-6
        vals = []
        for sgn in (1.0, -1.0):
            v2, g2 = value.copy(), grad.copy()
            v2[:, r] += sgn * h * dv
            g2[:, r] += sgn * h * dg
            vals.append(o.streamfunction_rayleigh_peel(v2, g2, w, order="fixed")["rayleigh"][r])
        fd = (vals[0] - vals[1]) / (2 * h)
        exact = base["G_value"][:, r] @ dv + base["G_grad"][:, r] @ dg
        assert abs(fd - exact) <= 1e-6 * max(abs(exact), 1e-8), (r, fd, exact)


def test_stream_peel_fourier_closed_form():
    """Fourier stream functions plus constants on the periodic grid: the
    centering removes the constants, the peel leaves the modes as they are
    and the Rayleigh quotients are the sorted |k|^2."""
    x = o.periodic_grid(32)
    ks = [(2, 0), (0, 1), (1, 1), (1, 2), (3, 0), (2, 2)]
    n = len(x)
    value = np.zeros((n, len(ks)))
    grad = np.zeros((2 * n, len(ks)))
    for j, (kx, ky) in enumerate(ks):
        ph = kx * x[:, 0] + ky * x[:, 1]
        value[:, j] = np.sin(ph) + 0.5 * (j + 1)
        grad[0::2, j], grad[1::2, j] = kx * np.cos(ph), ky * np.cos(ph)
    w = (2 * np.pi / 32) ** 2
    res = o.streamfunction_rayleigh_peel(value, grad, w)
    k2 = np.array([kx * kx + ky * ky for kx, ky in ks], dtype=float)
    np.testing.assert_allclose(res["rayleigh"], np.sort(k2), rtol=1e-12)
    np.testing.assert_array_equal(res["order"], np.argsort(k2, kind="stable"))
    np.testing.assert_allclose(res["projection"], 0.0, atol=1e-20)
    np.testing.assert_allclose(res["mass"], 2 * np.pi ** 2, rtol=1e-12)


# --------------------------------------------------------------------------
# envelope product (VJ2:47-55)
# --------------------------------------------------------------------------
def _gauss_envelope(x, c=0.3):
    """e = exp(-c |x|^2) with its exact gradient and Hessian pairs."""
    d = x.shape[1]
    e = np.exp(-c * np.sum(x * x, 1))
    g = -2 * c * x * e[:, None]
    h = np.stack([(4 * c * c * x[:, i] * x[:, j] - (2 * c if i == j else 0.0)) * e for i, j in o.pairs(d)], -1)
    return o.Jet(e, g, h)


@pytest.mark.parametrize("d", [2, 3])
def test_envelope_product_matches_spatial_fd(d):
    """The product jets are the jets of the product function: FD of the
    enveloped value -> gradient, of the enveloped gradient -> Hessian."""
    rng = np.random.default_rng(40 + d)
    net = _net(d, 2, [10], 3, seed=d)
    lo, hi = (0.0, 2 * np.pi) if d == 2 else (-1.0, 1.0)
    x = rng.uniform(lo + 0.3, hi - 0.3, (6, d))

    def prod(xx):
        return o.envelope_product(_gauss_envelope(xx), _eval(net, xx, 2, lo, hi))

    y = prod(x)
    eps = 1e-5
    for a in range(d):
        dx = np.zeros(d); dx[a] = eps
        yp, ym = prod(x + dx), prod(x - dx)
        np.testing.assert_allclose((yp.value - ym.value) / (2 * eps), y.grad[..., a], rtol=1e-7, atol=1e-8)
        for p, (i, j) in enumerate(o.pairs(d)):
            if i == a:
                np.testing.assert_allclose((yp.grad[..., j] - ym.grad[..., j]) / (2 * eps), y.hess[..., p],
                                           rtol=1e-6, atol=1e-7)


@pytest.mark.parametrize("d", [2, 3])
def test_envelope_product_reverse_is_transpose(d):
    rng = np.random.default_rng(50 + d)
    n, f, P = 11, 4, d * (d + 1) // 2
    e = o.Jet(rng.normal(size=n), rng.normal(size=(n, d)), rng.normal(size=(n, P)))
    x = o.Jet(rng.normal(size=(n, f)), rng.normal(size=(n, f, d)), rng.normal(size=(n, f, P)))
    a = o.Jet(rng.normal(size=(n, f)), rng.normal(size=(n, f, d)), rng.normal(size=(n, f, P)))
    y, r = o.envelope_product(e, x), o.envelope_product_reverse(e, a)
    lhs = np.sum(y.value * a.value) + np.sum(y.grad * a.grad) + np.sum(y.hess * a.hess)
    rhs = np.sum(x.value * r.value) + np.sum(x.grad * r.grad) + np.sum(x.hess * r.hess)
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)


# --------------------------------------------------------------------------
# per-mode networks
# --------------------------------------------------------------------------
@pytest.mark.parametrize("dim,stiff", [(2, "vorticity"), (2, "hessian"), (3, "grad_u")])
def test_mode_reverse_matches_directional_fd(dim, stiff):
    """mode_reverse's gradient of <dM, M> + <dS, S> for per-mode sine
    networks against central differences of the same objective along random
    directions in every mode's parameters."""
    rng = np.random.default_rng(60 + dim)
    K, F = 3, 1
    sizes = [2 * F * dim, 6, 5, 1 if dim == 2 else 3]
    Ws = [rng.normal(size=(K, fo, fi)) * 0.5 for fi, fo in zip(sizes[:-1], sizes[1:])]
    bs = [rng.normal(size=(K, fo)) * 0.3 for fo in sizes[1:]]
    lo, hi = (0.0, 2 * np.pi) if dim == 2 else (-1.0, 1.0)
    x = rng.uniform(lo, hi, (40, dim))
    x0 = o.feature_jets(x, F, lo, hi)
    w = 0.05
    dM = rng.normal(size=(K, K)); dM = dM + dM.T
    dS = rng.normal(size=(K, K)); dS = dS + dS.T

    def objective(Ws_, bs_):
        U, D, _ = o.mode_fields(o.mode_nets(Ws_, bs_, 1.3, 1.1), x0, dim, stiff)
        return np.sum(dM * o.gram(U, w)) + np.sum(dS * o.gram(D, w))

    nets = o.mode_nets(Ws, bs, 1.3, 1.1)
    U, D, tapes = o.mode_fields(nets, x0, dim, stiff)
    gW, gb = o.mode_reverse(nets, tapes, U, D, w, dM, dS, dim, stiff)
    for _ in range(3):
        dW = [rng.normal(size=W.shape) for W in Ws]
        db = [rng.normal(size=b.shape) for b in bs]
        h = 1e-6
        fp = objective([W + h * a for W, a in zip(Ws, dW)], [b + h * a for b, a in zip(bs, db)])
        fm = objective([W - h * a for W, a in zip(Ws, dW)], [b - h * a for b, a in zip(bs, db)])
        exact = sum(np.sum(g * a) for g, a in zip(gW + gb, dW + db))
        assert abs((fp - fm) / (2 * h) - exact) <= 1e-6 * max(1.0, abs(exact)), ((fp - fm) / (2 * h), exact)


def test_mode_adam_per_mode_clip_and_first_step():
    rng = np.random.default_rng(70)
    K = 4
    Ws = [rng.normal(size=(K, 3, 2))]
    bs = [rng.normal(size=(K, 3))]
    gW = [rng.normal(size=(K, 3, 2))]
    gb = [rng.normal(size=(K, 3))]
    gW[0][2] *= 1e3          # mode 2 beyond the clip
    gb[0][2] *= 1e3
    z = [np.zeros_like(Ws[0])], [np.zeros_like(bs[0])]
    out, norms = o.mode_adam(Ws, bs, gW, gb, z[0], z[0], z[1], z[1], 1, 1e-3, clip=8.0)
    (pW, mW, _), (pb, mb, _) = out
    assert norms[2] > 8.0 and np.all(norms[[0, 1, 3]] < 8.0)
    # clipped mode: first moment = 0.1 * g * clip / norm; unclipped modes: 0.1 * g
    np.testing.assert_allclose(mW[2], 0.1 * gW[0][2] * 8.0 / norms[2], rtol=1e-12)
    np.testing.assert_allclose(mW[0], 0.1 * gW[0][0], rtol=1e-12)
    # first step: p - lr * g / (|g| + eps) elementwise (any positive scale cancels)
    np.testing.assert_allclose(pW, Ws[0] - 1e-3 * gW[0] / (np.abs(gW[0]) + 1e-8 / np.repeat(
        np.minimum(1, 8.0 / norms), 6).reshape(K, 3, 2)), rtol=1e-9)
    with pytest.raises(FloatingPointError):
        o.mode_adam(Ws, bs, [gW[0] * 0], [gb[0] * 0], z[0], z[0], z[1], z[1], 1, 1e-3)
