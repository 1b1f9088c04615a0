"""GPU parity of the HIP path (through the C ABI) against the FP64 oracle.

Sizes are small enough for the oracle to finish in seconds.  Tolerances:
FP64 kernels agree with the oracle to ~1e-10 relative; FP32 to ~1e-5; BF16
(8-bit significand operands, fp32 accumulation) to ~3e-2; the split
BF16 hi+lo Gram to <= 1e-5 relative Frobenius error against the FP64 Gram of
the same fields.
"""
import math

import numpy as np
import pytest
import torch

from oracle import neb_oracle as o

pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():  # CPU-only host: the gpu tier runs on the MI355X box
    pytest.skip("needs a GPU", allow_module_level=True)

import neb_amd as nb  # noqa: E402
from neb_amd import network as nw  # noqa: E402

DEV = "cuda"


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def make(dim, K, widths, F, act="tanh", seed=0, master=torch.float64, omegas=None, bias=True):
    lo, hi = ((0.0, 0.0), (2 * math.pi, 2 * math.pi)) if dim == 2 else ((-1.0,) * 3, (1.0,) * 3)
    cfg = nb.NetConfig(dim=dim, frequencies=F, lo=lo, hi=hi, widths=tuple(widths), modes=K, activation=act,
                       omegas=omegas)
    net = nb.EigenfluidNet(cfg, seed=seed, device=DEV, master_dtype=master)
    if bias:
        rng = np.random.default_rng(seed + 100)
        with torch.no_grad():
            for b in net.b:
                b.copy_(torch.as_tensor(rng.normal(size=b.shape) * 0.2, dtype=master))
        net.touch()
    onet = o.Net([w.detach().cpu().double().numpy() for w in net.W], [b.detach().cpu().double().numpy() for b in net.b],
                 act, list(omegas) if omegas else [1.0] * len(widths))
    return cfg, net, onet


def points(dim, n, seed=1):
    rng = np.random.default_rng(seed)
    if dim == 2:
        return rng.uniform(0, 2 * np.pi, (n, 2))
    return rng.uniform(-1, 1, (n, 3))


TOL = {"f64": 1e-10, "f32": 2e-5, "bf16": 3e-2}


@pytest.mark.parametrize("prec", ["f64", "f32", "bf16"])
@pytest.mark.parametrize("dim", [2, 3])
def test_feature_jets(prec, dim):
    n, F = 77, 3
    cfg, net, _ = make(dim, 4, [16], F)
    x = points(dim, n)
    tape = nw.trunk_forward(net, torch.tensor(x), nw.PRECISIONS[prec])
    X0 = tape.inputs[0].double().cpu().numpy()[:, :, :cfg.features]
    ref = o.feature_jets(x, F, cfg.lo, cfg.hi).pack()
    tol = {"f64": 1e-13, "f32": 1e-6, "bf16": 5e-3}[prec]
    r = rel(X0, ref)
    assert r < tol, r
    assert np.all(tape.inputs[0].double().cpu().numpy()[:, :, cfg.features:] == 0)


@pytest.mark.parametrize("prec", ["f64", "f32", "bf16"])
@pytest.mark.parametrize("dim,act", [(2, "tanh"), (3, "tanh"), (2, "sin")])
def test_forward_output_jets(prec, dim, act):
    n, K, F = 203, 40 if dim == 2 else 12, 3
    omegas = (1.5, 1.2, 1.1) if act == "sin" else None
    cfg, net, onet = make(dim, K, [64, 48, 32], F, act=act, omegas=omegas)
    x = points(dim, n)
    tape = nw.trunk_forward(net, torch.tensor(x), nw.PRECISIONS[prec])
    v, g, h = nw.output_jets(net, tape)
    out, _ = o.forward(onet, o.feature_jets(x, F, cfg.lo, cfg.hi))
    tol = TOL[prec]
    r = rel(v.cpu(), out.value); assert r < tol, r
    r = rel(g.cpu(), out.grad); assert r < tol, r
    r = rel(h.cpu(), out.hess); assert r < tol * (3 if prec == "bf16" else 1), r
    # tape: pre-activation of layer 0
    z0 = tape.pre[0].double().cpu().numpy()[:, :, :64]
    ref_z0 = o.linear_jet(o.feature_jets(x, F, cfg.lo, cfg.hi), onet.W[0], onet.b[0])
    om = onet.omegas[0]
    ref_z0 = o.Jet(om * ref_z0.value, om * ref_z0.grad, om * ref_z0.hess).pack()
    r = rel(z0, ref_z0)
    assert r < tol, r


@pytest.mark.parametrize("prec", ["f64", "f32", "bf16"])
@pytest.mark.parametrize("dim", [2, 3])
def test_output_fields_lift(prec, dim):
    n, F = 150, 2
    K = 36 if dim == 2 else 20
    cfg, net, onet = make(dim, K, [32, 32], F)
    x = points(dim, n, seed=4)
    tape = nw.trunk_forward(net, torch.tensor(x), nw.PRECISIONS[prec])
    f = nw.output_fields(net, tape)
    out, _ = o.forward(onet, o.feature_jets(x, F, cfg.lo, cfg.hi))
    if dim == 2:
        u, w = o.lift2d(out)
        Um, Ds = o.field_rows(u), w
    else:
        u, du = o.lift3d(out)
        Um, Ds = o.field_rows(u), o.field_rows(du)
    tol = TOL[prec]
    r = rel(f.U.cpu(), Um); assert r < tol, r
    r = rel(f.D.cpu(), Ds); assert r < tol * (2 if prec == "bf16" else 1), r
    # component-major layout holds the same rows, permuted
    f1 = nw.output_fields(net, tape, layout=1)
    cm = 2 if dim == 2 else 3
    perm = f1.U.cpu().numpy().reshape(cm, n, K).transpose(1, 0, 2).reshape(n * cm, K)
    assert np.array_equal(perm, f.U.cpu().numpy())


@pytest.mark.parametrize("dim,layout", [(2, 0), (2, 1), (3, 0)])
def test_split_fields_are_the_rounded_fp32_fields(dim, layout):
    """Split planes: hi = bf16(x), lo = bf16(x - hi) of the fp32 field x,
    bit for bit, including ragged K (not a multiple of the 64-feature tile)."""
    n, F = 333, 2
    K = 100 if dim == 2 else 37
    cfg, net, _ = make(dim, K, [32, 32], F)
    tape = nw.trunk_forward(net, torch.tensor(points(dim, n, seed=5)), nw.PRECISIONS["bf16"])
    f = nw.output_fields(net, tape, layout=layout)
    sf = nw.output_fields(net, tape, layout=layout, split=True)
    assert sf.split and not f.split
    for x, hi, lo in ((f.U, sf.U, sf.mass_lo[:, :K]), (f.D, sf.D, sf.stiff_lo[:, :K])):
        h = x.to(torch.bfloat16)
        assert torch.equal(hi, h)
        assert torch.equal(lo, (x - h.float()).to(torch.bfloat16))


@pytest.mark.parametrize("R,K", [(1000, 40), (4099, 130), (20000, 256), (2 * 4096 + 7, 1024)])
def test_split_gram_matches_split_on_load(R, K):
    """The Gram over pre-split planes (asynchronous-copy kernel) forms the
    same three bf16 products of the same operands as the split-on-load
    bf16x3 Gram of the fp32 matrix; only the row splits of the fp32
    partial sums differ, so the two agree far inside the 1e-5 bound both
    keep against FP64.  Ragged R and K exercise partial k-steps and
    columns past the matrix."""
    rng = np.random.default_rng(R + K)
    phi = rng.normal(size=(R, K)) * np.exp(rng.normal(size=K))[None, :]
    ld = nw._ld(K)
    t = torch.zeros((R, ld), dtype=torch.float32, device=DEV)
    t[:, :K] = torch.tensor(phi, dtype=torch.float32)
    hi = t.to(torch.bfloat16)
    lo = (t - hi.float()).to(torch.bfloat16)
    G1 = nw.gram(t, K, 0.37, "bf16x3")
    G2 = nw.gram(hi, K, 0.37, "bf16x3", lo=lo)
    r = rel(G2.cpu(), G1.cpu()); assert r < 1e-6, r
    assert torch.equal(G2, G2.T), "split Gram not exactly symmetric"
    r = rel(G2.cpu(), o.gram(t[:, :K].double().cpu().numpy(), 0.37)); assert r < 1e-5, r
    G3 = nw.gram(hi, K, 1.0, "bf16x3", lo=lo, out=G2.clone(), beta=2.0)
    r = rel(G3.cpu(), (2.0 + 1.0 / 0.37) * G2.cpu()); assert r < 1e-12, r


@pytest.mark.parametrize("rows1,rows2,m,n", [(513, 1200, 70, 33), (4096, 65, 256, 256), (40000, 777, 1024, 256)])
def test_bf16_cross_segments_async(rows1, rows2, m, n):
    """bf16-stored operands take the asynchronous-copy cross kernel: two
    segments with signs +1 / -1, ragged rows (partial k-steps) and column
    counts off the 128 tile; against the FP64 product of the same bf16
    values (bf16 x bf16 products are exact in fp32, the sums round)."""
    rng = np.random.default_rng(rows1 + m)
    X1, Y1 = rng.normal(size=(rows1, m)), rng.normal(size=(rows1, n))
    X2, Y2 = rng.normal(size=(rows2, m)), rng.normal(size=(rows2, n))

    def B(a):
        ld = nw._ld(a.shape[1])
        t = torch.zeros((a.shape[0], ld), dtype=torch.bfloat16, device=DEV)
        t[:, :a.shape[1]] = torch.tensor(a, dtype=torch.float32).to(torch.bfloat16)
        return t
    xs = [B(X1), B(X2)]
    ys = [B(Y1), B(Y2)]
    G = nw.cross([(xs[0], ys[0], rows1, 1.0), (xs[1], ys[1], rows2, -1.0)], m, n, nw.GRAM_MODES["bf16"],
                 scale=2.0).cpu().numpy()
    d = lambda t, c: t[:, :c].double().cpu().numpy()  # noqa: E731
    ref = 2.0 * (d(xs[0], m).T @ d(ys[0], n) - d(xs[1], m).T @ d(ys[1], n))
    r = rel(G, ref)
    assert r < 2e-6, r


def test_split_gram_rejects_other_modes():
    t = torch.zeros((64, 8), dtype=torch.bfloat16, device=DEV)
    with pytest.raises(ValueError):
        nw.gram(t, 8, 1.0, "f32", lo=t)
    with pytest.raises(ValueError):
        nw.gram(t, 8, 1.0, "bf16x3", lo=t, roww=torch.ones(64, device=DEV))


@pytest.mark.parametrize("dim", [2, 3])
def test_reverse_with_split_fields_is_bit_identical(dim):
    """The reverse pass reads the hi plane where it rounded the fp32 fields
    to bf16: the gradients agree bit for bit."""
    F = 2
    K = 40 if dim == 2 else 12
    cfg, net, _ = make(dim, K, [48, 32], F, master=torch.float32)
    x = torch.tensor(points(dim, 500, seed=8))
    tape = nw.trunk_forward(net, x, nw.PRECISIONS["bf16"])
    rng = np.random.default_rng(2)
    dM = rng.normal(size=(K, K)); dM = torch.tensor(dM + dM.T, device=DEV)
    dS = rng.normal(size=(K, K)); dS = torch.tensor(dS + dS.T, device=DEV)
    g1 = nw.backward_fields(net, tape, nw.output_fields(net, tape), dM, dS, 1e-3)
    g2 = nw.backward_fields(net, tape, nw.output_fields(net, tape, split=True), dM, dS, 1e-3)
    assert torch.equal(g1, g2)


@pytest.mark.parametrize("mode,tol", [("f64", 1e-13), ("f32", 2e-6), ("bf16x3", 1e-5), ("bf16", 1e-2)])
@pytest.mark.parametrize("R,K", [(1000, 40), (4099, 130), (20000, 256)])
def test_gram_modes_against_fp64(mode, tol, R, K):
    rng = np.random.default_rng(R + K)
    phi = rng.normal(size=(R, K)) * np.exp(rng.normal(size=K))[None, :]
    dt = torch.float64 if mode == "f64" else torch.float32
    t = torch.tensor(phi, dtype=dt, device=DEV)
    G = nw.gram(t, K, 0.37, mode).cpu().numpy()
    ref = o.gram(t.double().cpu().numpy(), 0.37)
    r = rel(G, ref)
    assert r < tol, r
    assert np.array_equal(G, G.T), "Gram not exactly symmetric"


def test_gram_accumulate_and_row_weights():
    rng = np.random.default_rng(3)
    phi = rng.normal(size=(3000, 64))
    ww = rng.uniform(0.1, 1.0, 3000)
    t = torch.tensor(phi, device=DEV)
    G = nw.gram(t, 64, 2.0, "f64", roww=torch.tensor(ww, device=DEV))
    r = rel(G.cpu(), 2.0 * o.gram(phi, ww)); assert r < 1e-13, r
    nw.gram(t, 64, 1.0, "f64", out=G, beta=1.0)
    r = rel(G.cpu(), 2.0 * o.gram(phi, ww) + o.gram(phi, 1.0)); assert r < 1e-13, r


@pytest.mark.parametrize("mode", ["f64", "f32", "bf16"])
def test_cross_segments_signs(mode):
    rng = np.random.default_rng(9)
    X1, Y1 = rng.normal(size=(513, 70)), rng.normal(size=(513, 33))
    X2, Y2 = rng.normal(size=(1200, 70)), rng.normal(size=(1200, 33))
    dt = torch.float64 if mode == "f64" else torch.float32
    T = lambda a: torch.tensor(a, dtype=dt, device=DEV)  # noqa: E731
    code = nw.GRAM_MODES[mode]
    G = nw.cross([(T(X1), T(Y1), 513, 1.0), (T(X2), T(Y2), 1200, -0.5)], 70, 33, code, scale=3.0).cpu().numpy()
    ref = 3.0 * (X1.T @ Y1 - 0.5 * X2.T @ Y2)
    tol = {"f64": 1e-13, "f32": 2e-6, "bf16": 1e-2}[mode]
    r = rel(G, ref)
    assert r < tol, r


def _oracle_step(onet, x, F, cfg, w, lam, ridge):
    x0 = o.feature_jets(x, F, cfg.lo, cfg.hi)
    out, tape = o.forward(onet, x0, record=True)
    n = len(x)
    K = cfg.modes
    if cfg.dim == 2:
        u, om = o.lift2d(out)
        Pm, Ps = o.field_rows(u), om
    else:
        u, du = o.lift3d(out)
        Pm, Ps = o.field_rows(u), o.field_rows(du)
    M, S = o.gram(Pm, w), o.gram(Ps, w)
    L, dM, dS = o.spectral_loss(M, S, lam, ridge)
    gPm, gPs = o.field_adjoint(Pm, w, dM), o.field_adjoint(Ps, w, dS)
    if cfg.dim == 2:
        adj = o.lift2d_reverse(gPm.reshape(n, 2, K), gPs)
    else:
        adj = o.lift3d_reverse(gPm.reshape(n, 3, K), gPs.reshape(n, 3, 3, K))
    gW, gb = o.backward(onet, tape, adj)
    return L, M, S, gW, gb, dM, dS


def _oracle_reverse(onet, x, F, cfg, w, dM, dS):
    out, tape = o.forward(onet, o.feature_jets(x, F, cfg.lo, cfg.hi), record=True)
    n, K = len(x), cfg.modes
    if cfg.dim == 2:
        u, om = o.lift2d(out)
        gPm, gPs = o.field_adjoint(o.field_rows(u), w, dM), o.field_adjoint(om, w, dS)
        adj = o.lift2d_reverse(gPm.reshape(n, 2, K), gPs)
    else:
        u, du = o.lift3d(out)
        gPm, gPs = o.field_adjoint(o.field_rows(u), w, dM), o.field_adjoint(o.field_rows(du), w, dS)
        adj = o.lift3d_reverse(gPm.reshape(n, 3, K), gPs.reshape(n, 3, 3, K))
    return o.backward(onet, tape, adj)


@pytest.mark.parametrize("dim,act", [(2, "tanh"), (3, "tanh"), (2, "sin")])
def test_spectral_step_fp64_matches_oracle(dim, act):
    """Full FP64 step on the GPU (jets, lift, Gram, objective, reverse) vs
    the oracle."""
    F = 2
    K = 24 if dim == 2 else 8
    omegas = (1.4, 1.2, 1.1) if act == "sin" else None
    cfg, net, onet = make(dim, K, [48, 40, 32], F, act=act, master=torch.float64, omegas=omegas)
    x = points(dim, 300, seed=7)
    w = 1.0 / 300
    L, M, S, gW, gb, _, _ = _oracle_step(onet, x, F, cfg, w, 1e-2, 0.0)
    res = nb.compute(net, torch.tensor(x), w, prec="f64", gram_mode="f64", gate="cholesky")
    tol = 1e-9
    r = rel(res["M"].cpu(), M); assert r < tol, r
    r = rel(res["S"].cpu(), S); assert r < tol, r
    assert abs(float(res["loss"]) - L) < tol * abs(L)
    GW, Gb = net.split(res["grads"])
    for a, b in zip(GW, gW):
        r = rel(a.detach().cpu(), b); assert r < tol, r
    for a, b in zip(Gb[:-1], gb[:-1]):
        r = rel(a.detach().cpu(), b); assert r < tol, r
    assert float(Gb[-1].abs().max()) == 0.0 and np.abs(gb[-1]).max() == 0.0


@pytest.mark.parametrize("prec,tol", [("f32", 2e-5), ("bf16", 4e-2)])
@pytest.mark.parametrize("dim,act", [(2, "tanh"), (3, "tanh"), (2, "sin")])
def test_reverse_pass_reduced_precision(prec, tol, dim, act):
    """FP32 / BF16 reverse pass for given symmetric cotangents dL/dM, dL/dS.
    Random unit-scale cotangents isolate the kernels' rounding; the spectral
    objective's own cotangent M^-1 S M^-1 would amplify any field rounding by
    cond(M) (that amplification is measured by the FP64 step test)."""
    F = 2
    K = 24 if dim == 2 else 8
    omegas = (1.4, 1.2, 1.1) if act == "sin" else None
    cfg, net, onet = make(dim, K, [48, 40, 32], F, act=act, master=torch.float32, omegas=omegas)
    x = points(dim, 300, seed=7)
    w = 1.0 / 300
    rng = np.random.default_rng(21)
    dM = rng.normal(size=(K, K)); dM = dM + dM.T
    dS = rng.normal(size=(K, K)); dS = dS + dS.T
    gW, gb = _oracle_reverse(onet, x, F, cfg, w, dM, dS)
    p = nw.PRECISIONS[prec]
    tape = nw.trunk_forward(net, torch.tensor(x), p)
    fields = nw.output_fields(net, tape)
    g = nw.backward_fields(net, tape, fields, torch.tensor(dM, device=DEV), torch.tensor(dS, device=DEV), w)
    GW, Gb = net.split(g)
    for a, b in zip(GW, gW):
        r = rel(a.detach().cpu(), b); assert r < tol, r
    for a, b in zip(Gb[:-1], gb[:-1]):
        r = rel(a.detach().cpu(), b); assert r < tol, r


def test_fd_directional_fp64_gpu():
    """Config-C style check on the GPU path itself: FP64 handwritten gradient
    vs central FD (h=1e-3) of the GPU FP64 loss along random unit directions."""
    F, K = 2, 16
    cfg, net, _ = make(2, K, [32, 32, 32], F, master=torch.float64)
    x = torch.tensor(points(2, 1024, seed=3))
    w = 4 * math.pi ** 2 / 1024
    res = nb.compute(net, x, w, prec="f64", gram_mode="f64", gate="cholesky")
    g = res["grads"].clone()
    base = net.params.detach().clone()
    rng = np.random.default_rng(5)
    h = 1e-3
    for _ in range(6):
        d = torch.tensor(rng.normal(size=base.numel()), device=DEV)
        d /= d.norm()
        with torch.no_grad():
            net.params.copy_(base + h * d)
        net.touch()
        Lp = float(nb.compute(net, x, w, prec="f64", gram_mode="f64", training=False, gate="cholesky")["loss"])
        with torch.no_grad():
            net.params.copy_(base - h * d)
        net.touch()
        Lm = float(nb.compute(net, x, w, prec="f64", gram_mode="f64", training=False, gate="cholesky")["loss"])
        fd = (Lp - Lm) / (2 * h)
        exact = float(g @ d)
        assert abs(fd - exact) <= 1e-4 * max(abs(exact), 1e-3 * abs(float(res["loss"]))), (fd, exact)
    with torch.no_grad():
        net.params.copy_(base)
    net.touch()


def test_autograd_function_matches_closed_form():
    F, K = 2, 12
    cfg, net, _ = make(2, K, [32, 32], F, master=torch.float64)
    x = torch.tensor(points(2, 256, seed=2))
    w = 1.0 / 256
    net.params.requires_grad_(True)
    M, S = nb.JetGram.apply(net.params, x, net, w, "f64", "f64")
    Mi = torch.linalg.inv(M)
    loss = torch.trace(Mi @ S) + 1e-2 * torch.sum((M - torch.eye(K, dtype=M.dtype, device=DEV)) ** 2)
    g, = torch.autograd.grad(loss, net.params)
    net.params.requires_grad_(False)
    res = nb.compute(net, x, w, prec="f64", gram_mode="f64", gate="cholesky")
    r = rel(g.cpu(), res["grads"].cpu())
    assert r < 1e-9, r


def test_adam_first_step_closed_form():
    cfg, net, _ = make(2, 8, [16], 2, master=torch.float32)
    rng = np.random.default_rng(82)
    g = torch.tensor(rng.normal(size=net.params.numel()), dtype=torch.float32, device=DEV)
    before = net.params.detach().clone()
    norm = net.adam(g, 1e-4, clip=1e6)
    assert abs(norm - float(g.double().norm())) < 1e-6 * norm
    expect = before - 1e-4 * g / (g.abs() + 1e-8)
    assert torch.allclose(net.params, expect, rtol=2e-6, atol=2e-7)
    after = net.params.detach().clone()
    with pytest.raises(FloatingPointError):
        net.adam(g * float("nan"), 1e-4)
    with pytest.raises(FloatingPointError):
        net.adam(torch.zeros_like(g), 1e-4)
    assert torch.equal(net.params, after) and net.epoch == 1  # the device gate skipped both updates


def test_adam_three_steps_match_oracle():
    """Three gated Adam steps against the oracle's (NCC:156-171): the bias
    corrections 1 - b^t for t = 2, 3 and the global-norm clip (active on the
    second step) -- a kernel using b instead of b^t fails here."""
    cfg, net, _ = make(2, 8, [16], 2, master=torch.float64)
    rng = np.random.default_rng(84)
    p = net.params.detach().cpu().numpy().copy()
    m, v = np.zeros_like(p), np.zeros_like(p)
    for t, (scale, clip) in enumerate(((1.0, 1e6), (10.0, 2.0), (0.3, 8.0)), start=1):
        g = rng.normal(size=p.size) * scale
        norm = net.adam(torch.tensor(g, device=DEV), 1e-3, clip=clip)
        (p, m, v), = o.adam([p], [g], [m], [v], t, 1e-3, clip=clip)[0]
        assert abs(norm - float(np.linalg.norm(g))) < 1e-9 * norm
        assert np.allclose(net.params.cpu().numpy(), p, rtol=1e-12, atol=1e-14), t
        assert np.allclose(net.m.cpu().numpy(), m, rtol=1e-12, atol=1e-16), t
        assert np.allclose(net.v.cpu().numpy(), v, rtol=1e-12, atol=1e-18), t
    assert net.epoch == 3


def test_adam_device_gate_skips_update():
    """A failed step gate handed to adam leaves parameters, moments and the
    epoch untouched and raises through on_gate; a passing gate updates
    exactly as without one."""
    cfg, net, _ = make(2, 8, [16], 2, master=torch.float32)
    rng = np.random.default_rng(83)
    g = torch.tensor(rng.normal(size=net.params.numel()), dtype=torch.float32, device=DEV)
    p0, m0, v0 = net.params.clone(), net.m.clone(), net.v.clone()
    calls = []

    def on_gate(v):
        calls.append(v)
        raise FloatingPointError(f"gate {v}")

    bad = torch.tensor([7], dtype=torch.int32, device=DEV)
    with pytest.raises(FloatingPointError, match="gate 7"):
        net.adam(g, 1e-3, gate=bad, on_gate=on_gate)
    assert calls == [7] and net.epoch == 0
    assert torch.equal(net.params, p0) and torch.equal(net.m, m0) and torch.equal(net.v, v0)
    good = torch.zeros(1, dtype=torch.int32, device=DEV)
    n1 = net.adam(g, 1e-3, clip=0.5, gate=good, on_gate=on_gate)
    _, net2, _ = make(2, 8, [16], 2, master=torch.float32)
    n2 = net2.adam(g, 1e-3, clip=0.5)
    assert n1 == n2 and torch.equal(net.params, net2.params) and torch.equal(net.m, net2.m)


def test_deferred_solver_gates():
    """deferred=True gives the same objective and adjoints without a host
    synchronisation and reports the positive-definiteness and condition
    gates through `gate` (raised by raise_gate / check)."""
    rng = np.random.default_rng(5)
    K = 100
    A = rng.normal(size=(400, K)); M0 = A.T @ A / 400
    B = rng.normal(size=(300, K)); S0 = B.T @ B / 300
    Mt, St = torch.tensor(M0, device=DEV), torch.tensor(S0, device=DEV)
    ref = nb.SpectralSolver(1e-2)(Mt, St)
    sd = nb.SpectralSolver(1e-2, deferred=True)
    out = sd(Mt, St)
    assert int(sd.gate.item()) == 0
    sd.check()
    for a, b in zip(out, ref):
        assert torch.equal(a, b)
    Mbad = M0.copy(); Mbad[70, 70] = -1.0
    sd2 = nb.SpectralSolver(1e-2, deferred=True)
    sd2(torch.tensor(Mbad, device=DEV), St)
    assert int(sd2.gate.item()) == 71
    with pytest.raises(FloatingPointError, match=r"pivot 71\)"):
        sd2.check()
    sd3 = nb.SpectralSolver(1e-2, cond_limit=1.0, deferred=True)
    sd3(Mt, St)
    assert int(sd3.gate.item()) == -1
    with pytest.raises(FloatingPointError, match="condition"):
        sd3.check()


@pytest.mark.parametrize("K", [1, 33, 100, 1024])
def test_spectral_term_kernels(K):
    """sym_shift / spectral_terms / mass_adjoint against plain FP64 numpy on
    non-symmetric inputs (ragged K, in-place outputs): symmetrised matrices
    bit for bit, the loss to rounding, the gate values."""
    rng = np.random.default_rng(K)
    M, S, X, C = (rng.normal(size=(K, K)) for _ in range(4))
    lam, ridge = 0.03, 0.25
    I = np.eye(K)
    Mt = torch.tensor(M, device=DEV)
    Mr = nw.sym_shift(Mt, ridge)
    assert np.array_equal(Mr.cpu().numpy(), 0.5 * (M + M.T) + ridge * I)
    A = torch.tensor(S, device=DEV)
    nw.sym_shift(A, 0.0, out=A)  # in place
    assert np.array_equal(A.cpu().numpy(), 0.5 * (S + S.T))
    Xs, Ss = 0.5 * (X + X.T), 0.5 * (S + S.T)
    loss_ref = np.sum(Xs * Ss) + lam * np.sum((M - I) ** 2)
    cond = np.linalg.norm(0.5 * (M + M.T) + ridge * I) * np.linalg.norm(Xs)
    for limit, info, expect in ((cond * 2, None, 0), (cond / 2, None, -1), (cond * 2, 17, 17)):
        Xt = torch.tensor(X, device=DEV)
        it = None if info is None else torch.tensor([info], dtype=torch.int32, device=DEV)
        loss, Xo, Sso, gate = nw.spectral_terms(Mt, torch.tensor(S, device=DEV), Xt, it, lam, ridge, limit)
        assert Xo.data_ptr() == Xt.data_ptr() and np.array_equal(Xt.cpu().numpy(), Xs)
        assert np.array_equal(Sso.cpu().numpy(), Ss)
        assert abs(float(loss) - loss_ref) <= 1e-12 * max(1.0, abs(loss_ref))
        assert int(gate.item()) == expect
    Ct = torch.tensor(C, device=DEV)
    H = nw.H
    H.call("neb_spectral_adjoint", K, Ct.data_ptr(), K, Mt.data_ptr(), K, lam, Ct.data_ptr(), K, nw._stream())
    ref = -0.5 * (C + C.T) + 2.0 * lam * (M - I)
    r = rel(Ct.cpu(), ref); assert r < 1e-15, r
    Xt = torch.tensor(Xs, device=DEV); St = torch.tensor(Ss, device=DEV)
    dM = nw.mass_adjoint(Xt, St, Mt, lam)
    ref = -0.5 * ((Xs @ Ss @ Xs) + (Xs @ Ss @ Xs).T) + 2.0 * lam * (M - I)
    r = rel(dM.cpu(), ref); assert r < 1e-12, r


def test_overlapped_streams_match_sequential():
    """The weight gradients on the side stream, overlapping the jet reverse,
    give the same bits call after call and with or without the
    symmetrisation of symmetric adjoints; field_grams equals field_gram."""
    cfg, net, _ = make(2, 40, [64, 48], 3, master=torch.float32)
    x = torch.tensor(points(2, 3001, seed=9))
    tape = nw.trunk_forward(net, x, nw.PRECISIONS["bf16"])
    f = nw.output_fields(net, tape, split=True)
    M, S = nw.field_grams(f, 0.01, "bf16x3")
    assert torch.equal(M, nw.field_gram(f, "mass", 0.01, "bf16x3"))
    assert torch.equal(S, nw.field_gram(f, "stiff", 0.01, "bf16x3"))
    loss, dM, dS = nb.SpectralSolver(1e-2, 1e-3)(M, S)
    g1 = nw.backward_fields(net, tape, f, dM, dS, 0.01, symmetric=True)
    g2 = nw.backward_fields(net, tape, f, dM, dS, 0.01, symmetric=True)
    g3 = nw.backward_fields(net, tape, f, dM.clone(), dS.clone(), 0.01)
    prod = nw.output_products(net, tape, f)  # queued on the side stream ahead of the reverse
    g4 = nw.backward_fields(net, tape, f, dM, dS, 0.01, symmetric=True, products=prod)
    assert torch.equal(g1, g2) and torch.equal(g1, g3) and torch.equal(g1, g4)
    assert torch.isfinite(g1).all() and float(g1.abs().max()) > 0


def test_projection_against_oracle():
    rng = np.random.default_rng(77)
    B = rng.normal(size=(3001, 24))
    T = rng.normal(size=(3001, 20))
    w = 0.125
    M = o.gram(B, w)
    ref = o.project(M, B, T, w)
    Bt, Tt = torch.tensor(B, device=DEV), torch.tensor(T, device=DEV)
    Mt = nw.gram(Bt, 24, w, "f64")
    res = nb.project(Mt, Bt, Tt, w)
    assert rel(res["coefficients"].cpu(), ref["coefficients"]) < 1e-12
    assert rel(res["direct_relative_L2"].cpu(), ref["direct_relative_L2"]) < 1e-12
    assert float((res["direct_relative_L2"] - res["Gram_identity_errors"]).abs().max()) < 1e-8
    assert float(res["normal_residuals"].max()) < 1e-8


def test_checkpoint_roundtrip(tmp_path):
    cfg, net, _ = make(3, 8, [16, 16], 2, master=torch.float32)
    net.epoch = 3
    p = tmp_path / "ck.npz"
    net.save(p)
    cfg2, net2, _ = make(3, 8, [16, 16], 2, seed=9, master=torch.float32)
    net2.load(p)
    assert net2.epoch == 3 and torch.equal(net2.params, net.params)
    z = np.load(p)
    assert {"W_0", "b_2", "mW_1", "vb_0", "epoch"} <= set(z.files)


def test_chunked_step_equals_single_pass():
    """Two-pass chunked step (Gram accumulation, then recompute + reverse)
    equals the single-pass step; FP64, 3D."""
    F, K = 2, 8
    cfg, net, _ = make(3, K, [32, 32], F, master=torch.float64)
    x = torch.tensor(points(3, 1000, seed=8))
    w = 1.0 / 1000
    ref = nb.compute(net, x, w, prec="f64", gram_mode="f64", gate="cholesky")
    # kept chunk tapes (the default budget) and per-chunk recomputation (budget 0)
    for budget, kept in ((None, True), (0, False)):
        res = nb.chunked_step(net, x, w, prec="f64", gram_mode="f64", chunk=256, tape_budget=budget)
        assert res["tapes_kept"] is kept
        r = rel(res["M"].cpu(), ref["M"].cpu()); assert r < 1e-13, r
        r = rel(res["S"].cpu(), ref["S"].cpu()); assert r < 1e-13, r
        r = rel(res["grads"].cpu(), ref["grads"].cpu()); assert r < 1e-11, r


def test_project_streamed_against_oracle():
    """Streamed two-pass full-mass projection of 20 divergence-free targets
    onto the 3D network basis vs the oracle's dense projection (PP:141-177)."""
    from neb_amd import synthetic as syn
    F, K = 2, 12
    cfg, net, onet = make(3, K, [32, 32], F, master=torch.float64)
    pts, _, _ = syn.blob_domain(2000, seed=3, batch=1 << 14)
    T = syn.curl_targets(pts, 20, seed=4).reshape(-1, 20)       # [3n, 20] point-major
    w = 1.0 / len(pts)
    x = torch.tensor(pts)
    M, ev = nb.fullmass(net, x, w, prec="f64", gram_mode="f64", chunk=512)
    res = nb.project_streamed(net, x, torch.tensor(T, device=DEV), M, w, prec="f64", chunk=512)
    out, _ = o.forward(onet, o.feature_jets(pts, F, cfg.lo, cfg.hi))
    u, _ = o.lift3d(out)
    B = o.field_rows(u)
    Mo = o.gram(B, w)
    r = rel(M.cpu(), Mo); assert r < 1e-12, r
    ref = o.project(Mo, B, T, w)
    r = rel(res["coefficients"].cpu(), ref["coefficients"]); assert r < 1e-9, r
    r = rel(res["direct_relative_L2"].cpu(), ref["direct_relative_L2"]); assert r < 1e-9, r
    assert float(res["normal_residuals"].max()) < 1e-8


def test_project_streamed_split_precision_basis():
    """The projection on the split-precision basis (hi + lo planes joined to
    fp32 by neb_join_bf16) against the oracle's projection on the exact
    basis: coefficients at the basis' FP32-level accuracy, the identity and
    normal gates passing; the join itself is exact."""
    from neb_amd import synthetic as syn
    F, K = 2, 12
    cfg, net, onet = make(3, K, [32, 32], F, master=torch.float32)
    pts, _, _ = syn.blob_domain(3000, seed=3, batch=1 << 14)
    T = syn.curl_targets(pts, 20, seed=4).reshape(-1, 20)
    w = 1.0 / len(pts)
    x = torch.tensor(pts)
    M = nb.spectral.fullmass_fp64(net, x, w, "bf16x3", chunk=1024)
    res = nb.project_streamed(net, x, torch.tensor(T, device=DEV), M, w, prec="bf16x3", chunk=1024)
    out, _ = o.forward(onet, o.feature_jets(pts, F, cfg.lo, cfg.hi))
    u, _ = o.lift3d(out)
    B = o.field_rows(u)
    ref = o.project(o.gram(B, w), B, T, w)
    r = rel(res["coefficients"].cpu(), ref["coefficients"]); assert r < 1e-3, r
    r = rel(res["direct_relative_L2"].cpu(), ref["direct_relative_L2"]); assert r < 1e-5, r
    assert float(res["normal_residuals"].max()) < 1e-8
    # the join: hi + lo of an fp32 split, exactly
    tape = nw.trunk_forward(net, x[:1000], nw.PRECISIONS["bf16x3"])
    f = nw.output_fields(net, tape, split=True)
    J = nw.field_matrix(f, "mass", torch.float32)
    assert torch.equal(J, f.mass[:, :K].float() + f.mass_lo[:, :K].float())
    J64 = nw.field_matrix(f, "stiff", torch.float64)
    assert torch.equal(J64, (f.stiff[:, :K].float() + f.stiff_lo[:, :K].float()).double())


def test_fullmass_rank_gate_rejects_rank_deficient_basis():
    """K modes from a width-4 trunk in 2D: rank(M) <= 8 < K, so the full-mass
    rank gate (experiments/metaball/run.py:108-110) raises."""
    cfg, net, _ = make(2, 16, [4], 2, master=torch.float64)
    x = torch.tensor(points(2, 500))
    with pytest.raises(ValueError):
        nb.fullmass(net, x, 1.0 / 500, prec="f64", gram_mode="f64")


def test_spectral_solver_paths_agree():
    """Newton-Schulz refinement of the previous inverse and the Cholesky path
    give the same objective and adjoints."""
    rng = np.random.default_rng(0)
    K = 64
    A = rng.normal(size=(400, K)); M0 = A.T @ A / 400
    B = rng.normal(size=(300, K)); S = B.T @ B / 300
    Mt, St = torch.tensor(M0, device=DEV), torch.tensor(S, device=DEV)
    solver = nb.SpectralSolver(1e-2)
    solver(Mt, St)
    assert solver.last_path == "spd-inverse"
    M1 = Mt + 1e-4 * torch.tensor(np.eye(K), device=DEV)
    L1, dM1, dS1 = solver(M1, St)
    assert solver.last_path == "newton-schulz"
    L2, dM2, dS2 = nb.spectral_objective(M1, St, 1e-2, gate="cholesky")
    assert abs(float(L1) - float(L2)) < 1e-10 * abs(float(L2))
    r = rel(dM1.cpu(), dM2.cpu()); assert r < 1e-9, r
    r = rel(dS1.cpu(), dS2.cpu()); assert r < 1e-9, r


def test_spectral_solver_backoff_after_failed_warm_start():
    """A warm start that does not converge falls back to the direct inverse
    and the next `backoff` calls skip the warm start; results are unchanged."""
    rng = np.random.default_rng(1)
    K = 48
    A = rng.normal(size=(300, K)); M0 = A.T @ A / 300
    B = rng.normal(size=(200, K)); S = B.T @ B / 200
    Mt, St = torch.tensor(M0, device=DEV), torch.tensor(S, device=DEV)
    solver = nb.SpectralSolver(1e-2, backoff=2)
    solver(Mt, St)
    far = 3.0 * Mt                                   # residual ||I - 3M X|| >> 1/2
    L1, dM1, _ = solver(far, St)
    assert solver.last_path == "spd-inverse" and solver._skip == 2
    near = far + 1e-6 * torch.eye(K, dtype=torch.float64, device=DEV)
    for left in (1, 0):
        L, dM, _ = solver(near, St)
        assert solver.last_path == "spd-inverse" and solver._skip == left
    L, dM, _ = solver(near, St)
    assert solver.last_path == "newton-schulz"
    L2, dM2, _ = nb.spectral_objective(near, St, 1e-2, gate="cholesky")
    assert abs(float(L) - float(L2)) < 1e-10 * abs(float(L2))
    r = rel(dM.cpu(), dM2.cpu()); assert r < 1e-9, r


@pytest.mark.parametrize("K", [5, 64, 70, 100, 130, 300, 1024])
def test_spd_inverse_against_lapack(K):
    from neb_amd import spectral as sp
    rng = np.random.default_rng(K)
    B = rng.normal(size=(K + 40, K))
    A = B.T @ B / K + 1e-3 * np.eye(K)
    X = sp.spd_inverse(torch.tensor(A, device=DEV)).cpu().numpy()
    ref = np.linalg.inv(A)
    r = rel(X, ref)
    assert r < 1e-10 * max(1.0, np.linalg.cond(A) / 1e4), r
    assert np.array_equal(X, X.T)


def test_spd_inverse_rejects_indefinite():
    from neb_amd import spectral as sp
    rng = np.random.default_rng(1)
    B = rng.normal(size=(80, 70))
    A = B.T @ B / 70
    A[40, 40] = -1.0  # breaks positive definiteness at pivot 41 (1-based), inside the first 64-block
    with pytest.raises(FloatingPointError, match=r"pivot 41\)"):
        sp.spd_inverse(torch.tensor(A, device=DEV))
    # the flag is cleared by the next call
    ok = B.T @ B / 70 + np.eye(70)
    sp.spd_inverse(torch.tensor(ok, device=DEV))


def test_spd_inverse_rejects_indefinite_later_block():
    """A nonpositive pivot in the third 64-block: that block's inverse is
    formed inside the sweep of the second block."""
    from neb_amd import spectral as sp
    rng = np.random.default_rng(2)
    B = rng.normal(size=(220, 200))
    A = B.T @ B / 200
    A[150, 150] = -1.0
    with pytest.raises(FloatingPointError, match=r"pivot 151\)"):
        sp.spd_inverse(torch.tensor(A, device=DEV))


@pytest.mark.parametrize("K", [5, 64, 70, 100, 130, 300, 1024])
def test_spd_inverse_chol_against_lapack(K):
    """T = U^-1 for A = U^T U from the sweeps (neb_spd_inverse_chol) against
    LAPACK's Cholesky factor, with the identities that fix it: T upper
    triangular with a positive diagonal and T^T A T = I; the inverse beside
    it is bit-identical to neb_spd_inverse's."""
    from neb_amd import spectral as sp
    rng = np.random.default_rng(K + 7)
    B = rng.normal(size=(K + 40, K))
    A = B.T @ B / K + 1e-3 * np.eye(K)
    At = torch.tensor(A, device=DEV)
    X, T = sp.spd_inverse_chol(At)
    X, T = X.cpu().numpy(), T.cpu().numpy()
    tol = 1e-10 * max(1.0, np.linalg.cond(A) / 1e4)
    assert not np.tril(T, -1).any()
    assert (np.diag(T) > 0).all()
    Tref = np.linalg.inv(np.linalg.cholesky(A).T)
    r = rel(T, Tref); assert r < tol, r
    r = np.abs(T.T @ A @ T - np.eye(K)).max(); assert r < tol, r
    assert np.array_equal(X, sp.spd_inverse(At).cpu().numpy())


def test_spd_inverse_chol_rejects_indefinite():
    from neb_amd import spectral as sp
    rng = np.random.default_rng(3)
    B = rng.normal(size=(220, 200))
    A = B.T @ B / 200
    A[150, 150] = -1.0
    with pytest.raises(FloatingPointError, match=r"pivot 151\)"):
        sp.spd_inverse_chol(torch.tensor(A, device=DEV))


# ---------------------------------------------------------------------------
# coefficient peel objective (MB/velocity/core.py:48-69, 94-95)
# ---------------------------------------------------------------------------
def _oracle_peel_step(onet, x, F, cfg, w):
    x0 = o.feature_jets(x, F, cfg.lo, cfg.hi)
    out, tape = o.forward(onet, x0, record=True)
    n, K = len(x), cfg.modes
    if cfg.dim == 2:
        u, om = o.lift2d(out)
        Pm, Ps = o.field_rows(u), om
    else:
        u, du = o.lift3d(out)
        Pm, Ps = o.field_rows(u), o.field_rows(du)
    M, S = o.gram(Pm, w), o.gram(Ps, w)
    p = o.coefficient_peel(M, S)
    gPm, gPs = w * Pm @ p["Cu"], w * Ps @ p["Cd"]
    if cfg.dim == 2:
        adj = o.lift2d_reverse(gPm.reshape(n, 2, K), gPs)
    else:
        adj = o.lift3d_reverse(gPm.reshape(n, 3, K), gPs.reshape(n, 3, 3, K))
    gW, gb = o.backward(onet, tape, adj)
    return p, M, S, gW, gb


def test_coefficient_peel_gpu_matches_oracle():
    rng = np.random.default_rng(31)
    K = 96
    U = rng.normal(size=(700, K)) @ np.diag(rng.uniform(0.3, 3.0, K))
    D = U @ rng.normal(size=(K, K)) * 0.1 + 0.2 * rng.normal(size=(700, K))
    M, S = o.gram(U, 1 / 350), o.gram(D, 1 / 350)
    ref = o.coefficient_peel(M, S)
    p = nb.coefficient_peel(torch.tensor(M, device=DEV), torch.tensor(S, device=DEV))
    assert np.array_equal(p["order"].cpu().numpy(), ref["order"])
    for key in ("losses", "T", "Cu", "Cd"):
        r = rel(p[key].cpu(), ref[key]); assert r < 1e-10, (key, r)
    assert int(p["gate"].item()) == 0


@pytest.mark.parametrize("dim", [2, 3])
def test_spectral_step_peel_fp64_matches_oracle(dim):
    """Full FP64 step with the coefficient peel objective on the GPU (jets,
    lift, Grams, peel, detached-prefix field cotangents fused into the
    reverse) against the oracle's peel and reverse pass."""
    F = 2
    K = 24 if dim == 2 else 8
    cfg, net, onet = make(dim, K, [48, 40, 32], F, master=torch.float64)
    x = points(dim, 300, seed=7)
    w = 1.0 / 300
    p, M, S, gW, gb = _oracle_peel_step(onet, x, F, cfg, w)
    res = nb.compute(net, torch.tensor(x), w, prec="f64", gram_mode="f64", objective="peel")
    tol = 1e-9
    r = rel(res["M"].cpu(), M); assert r < tol, r
    assert np.array_equal(res["peel"]["order"].cpu().numpy(), p["order"])
    r = rel(res["peel"]["losses"].cpu(), p["losses"]); assert r < tol, r
    GW, Gb = net.split(res["grads"])
    for a, b in zip(GW, gW):
        r = rel(a.detach().cpu(), b); assert r < tol, r
    for a, b in zip(Gb[:-1], gb[:-1]):
        r = rel(a.detach().cpu(), b); assert r < tol, r
    assert float(Gb[-1].abs().max()) == 0.0


def test_chunked_peel_step_equals_single_pass():
    F, K = 2, 8
    cfg, net, _ = make(3, K, [32, 32], F, master=torch.float64)
    x = torch.tensor(points(3, 1000, seed=8))
    w = 1.0 / 1000
    ref = nb.compute(net, x, w, prec="f64", gram_mode="f64", objective="peel")
    res = nb.chunked_step(net, x, w, prec="f64", gram_mode="f64", chunk=256, solver=nb.PeelSolver())
    r = rel(res["grads"].cpu(), ref["grads"].cpu()); assert r < 1e-11, r
    assert abs(float(res["loss"]) - float(ref["loss"])) < 1e-11 * abs(float(ref["loss"]))


def test_peel_reduced_precision_step():
    """FP32 and BF16 kernels under the peel: losses and gradients against the
    FP64 step on the same weights (2D, K = 24 <= width so M has full rank)."""
    F, K = 2, 24
    cfg, net, _ = make(2, K, [48, 40, 32], F, master=torch.float32)
    x = torch.tensor(points(2, 400, seed=3))
    w = 1.0 / 400
    ref = nb.compute(net, x, w, prec="f64", gram_mode="f64", objective="peel")
    for prec, gmode, tol in (("f32", "f32", 1e-3), ("bf16", "bf16x3", 2e-1)):
        res = nb.compute(net, x, w, prec=prec, gram_mode=gmode, objective="peel")
        r = rel(res["peel"]["losses"].cpu(), ref["peel"]["losses"].cpu()); assert r < tol, (prec, r)
        assert torch.isfinite(res["grads"]).all()


def test_peel_gate_skips_adam_update():
    """A rank-deficient mass (two identical modes) fails the peel gate on the
    device: the deferred gate makes adam skip the update and raise."""
    cfg, net, _ = make(2, 8, [16], 2, master=torch.float32)
    rng = np.random.default_rng(4)
    U = rng.normal(size=(300, 8)); U[:, 5] = U[:, 2]
    D = rng.normal(size=(300, 8)); D[:, 5] = D[:, 2]
    solver = nb.PeelSolver(deferred=True)
    solver(torch.tensor(o.gram(U, 0.01), device=DEV), torch.tensor(o.gram(D, 0.01), device=DEV))
    p0 = net.params.clone()
    g = torch.ones_like(net.params)
    with pytest.raises(FloatingPointError, match="gate failed"):
        net.adam(g, 1e-3, gate=solver.gate, on_gate=solver.raise_gate)
    assert torch.equal(net.params, p0) and net.epoch == 0


@pytest.mark.parametrize("dim", [2, 3])
def test_empty_point_set(dim):
    """Zero points: jets, fields and the reverse write nothing; the Grams and
    parameter gradients are sums over no points (zeros), and an accumulating
    Gram keeps beta * G."""
    F, K = 2, 8
    cfg, net, _ = make(dim, K, [32, 32], F, master=torch.float32)
    x = torch.zeros((0, dim))
    for prec, gmode in (("f32", "f32"), ("bf16", "bf16x3"), ("bf16x3", "bf16x3")):
        p = nw.PRECISIONS[prec]
        tape = nw.trunk_forward(net, x, p)
        fields = nw.output_fields(net, tape, split=prec != "f32")
        assert fields.mass.shape[0] == 0 and fields.stiff.shape[0] == 0
        M, S = nw.field_grams(fields, 0.5, gmode)
        assert float(M.abs().max()) == 0.0 and float(S.abs().max()) == 0.0
        dM = torch.eye(K, dtype=torch.float64, device=DEV)
        g = nw.backward_fields(net, tape, fields, dM, 2.0 * dM, 0.5)
        assert float(g.abs().max()) == 0.0
    for mode in ("f32", "f64"):
        dt = torch.float64 if mode == "f64" else torch.float32
        G = torch.full((K, K), 3.0, dtype=torch.float64, device=DEV)
        nw.gram(torch.zeros((0, K), dtype=dt, device=DEV), K, 1.0, mode, out=G, beta=1.0)
        assert float((G - 3.0).abs().max()) == 0.0


@pytest.mark.parametrize("dim", [2, 3])
def test_fp32_fields_bf16_reverse(dim):
    """FP32 jets and fields, split-precision Gram, BF16 reverse on the
    converted tape: Grams at FP32 accuracy against the FP64 oracle (<= 1e-5),
    the reverse pass against the oracle's at the BF16 tolerance."""
    F = 2
    K = 24 if dim == 2 else 8
    cfg, net, onet = make(dim, K, [48, 40, 32], F, master=torch.float32)
    x = points(dim, 300, seed=7)
    w = 1.0 / 300
    L, M, S, *_ = _oracle_step(onet, x, F, cfg, w, 1e-2, 0.0)
    res = nb.compute(net, torch.tensor(x), w, prec="f32", gram_mode="bf16x3", gate="cholesky", reverse_prec="bf16")
    r = rel(res["M"].cpu(), M); assert r < 1e-5, r
    r = rel(res["S"].cpu(), S); assert r < 1e-5, r
    assert torch.isfinite(res["grads"]).all()
    rng = np.random.default_rng(21)
    dM = rng.normal(size=(K, K)); dM = dM + dM.T
    dS = rng.normal(size=(K, K)); dS = dS + dS.T
    gW, gb = _oracle_reverse(onet, x, F, cfg, w, dM, dS)
    tape = nw.trunk_forward(net, torch.tensor(x), nw.PRECISIONS["f32"])
    fields = nw.output_fields(net, tape)
    rt = nw.convert_tape(tape, nw.PRECISIONS["bf16"])
    assert rt.inputs[-1].dtype == torch.bfloat16
    assert torch.equal(rt.pre[0], tape.pre[0].to(torch.bfloat16))
    g = nw.backward_fields(net, rt, fields, torch.tensor(dM, device=DEV), torch.tensor(dS, device=DEV), w)
    GW, Gb = net.split(g)
    for a, b in zip(GW, gW):
        r = rel(a.detach().cpu(), b); assert r < 4e-2, r
    for a, b in zip(Gb[:-1], gb[:-1]):
        r = rel(a.detach().cpu(), b); assert r < 4e-2, r


def test_split_bf16_planes():
    rng = np.random.default_rng(17)
    xs = torch.tensor(rng.normal(size=(37, 53)) * 10.0 ** rng.integers(-6, 6, size=(37, 53)), device=DEV)
    for src in (xs.float(), xs):
        hi = torch.empty((37, 56), dtype=torch.bfloat16, device=DEV)
        lo = torch.empty_like(hi)
        nw.H.call("neb_split_bf16", nw._DT[src.dtype], src.data_ptr(), 37, 53, 53, hi.data_ptr(), lo.data_ptr(), 56,
                  nw._stream())
        assert torch.equal(hi[:, :53], src.float().to(torch.bfloat16))
        err = (hi[:, :53].double() + lo[:, :53].double() - src.double()).abs() / src.double().abs()
        assert float(err.max()) <= 2.0 ** -16, float(err.max())


def test_stage_net_matches_per_layer_staging():
    """neb_stage_net (every layer's hi / lo / transposed bf16 forms in one
    launch) gives the bits of neb_split_bf16 and neb_stage layer by layer,
    padding included, for fp32 and fp64 masters."""
    for master in (torch.float32, torch.float64):
        cfg, net, _ = make(2, 20, [48, 36], 3, master=master)
        sp_ = net.staged_split()
        st = net.staged(nw.PRECISIONS["bf16"])
        s = nw._stream()
        for l, W in enumerate(net.W):
            fo, fi = W.shape
            ld = sp_["W"][l].shape[1]
            hi = torch.zeros((fo, ld), dtype=torch.bfloat16, device=DEV)
            lo = torch.zeros_like(hi)
            nw.H.call("neb_split_bf16", nw._DT[master], W.data_ptr(), fo, fi, fi, hi.data_ptr(), lo.data_ptr(), ld, s)
            assert torch.equal(sp_["W"][l], hi) and torch.equal(sp_["Wlo"][l], lo), l
            assert torch.equal(st["W"][l], hi), l
            if l < len(net.W) - 1:
                ldt = st["WT"][l].shape[1]
                wt = torch.empty((fi, ldt), dtype=torch.bfloat16, device=DEV)
                nw.H.call("neb_stage", nw._DT[master], nw.H.BF16, W.data_ptr(), fo, fi, fi, wt.data_ptr(), ldt, fi,
                          ldt, 1, 1.0, s)
                assert torch.equal(st["WT"][l], wt), l


@pytest.mark.parametrize("dim", [2, 3])
def test_split_precision_forward_matches_oracle(dim):
    """Split-precision jets (bf16 hi / lo planes, three bf16 MFMA products per
    term): fields and Grams at FP32-level accuracy against the FP64 oracle,
    two orders below the bf16 path; the hi planes are the bf16 tape."""
    F = 2
    K = 24 if dim == 2 else 8
    for act, omegas in (("tanh", None), ("sin", (1.4, 1.2, 1.1))):
        cfg, net, onet = make(dim, K, [48, 40, 32], F, act=act, master=torch.float32, omegas=omegas)
        x = points(dim, 300, seed=5)
        w = 1.0 / 300
        out, _ = o.forward(onet, o.feature_jets(x, F, cfg.lo, cfg.hi))
        if dim == 2:
            u, om = o.lift2d(out)
            U64, D64 = o.field_rows(u), om
        else:
            u, du = o.lift3d(out)
            U64, D64 = o.field_rows(u), o.field_rows(du)
        tape = nw.trunk_forward(net, torch.tensor(x), nw.PRECISIONS["bf16x3"])
        assert tape.prec == nw.PRECISIONS["bf16"] and len(tape.inputs_lo) == 4
        f = nw.output_fields(net, tape, split=True)
        U = (f.mass[:, :K].double() + f.mass_lo[:, :K].double()).cpu().numpy()
        D = (f.stiff[:, :K].double() + f.stiff_lo[:, :K].double()).cpu().numpy()
        eu, ed = rel(U, U64), rel(D, D64)
        assert eu < 2e-5 and ed < 2e-5, (act, eu, ed)
        tb = nw.trunk_forward(net, torch.tensor(x), nw.PRECISIONS["bf16"])
        fb = nw.output_fields(net, tb, split=True)
        Ub = (fb.mass[:, :K].double() + fb.mass_lo[:, :K].double()).cpu().numpy()
        assert eu < rel(Ub, U64) / 100, (eu, rel(Ub, U64))
        # the bf16 tape of the reverse: each hi plane is within half a bf16 ulp (2^-8 relative) of hi + lo
        Y = tape.inputs[-1].float() + tape.inputs_lo[-1].float()
        assert float(((tape.inputs[-1].float() - Y).abs() - 2.0 ** -8 * Y.abs()).max()) <= 0.0
        M, S = nw.field_grams(f, w, "bf16x3")
        r = rel(M.cpu(), o.gram(U64, w)); assert r < 1e-5, r
        r = rel(S.cpu(), o.gram(D64, w)); assert r < 1e-5, r


@pytest.mark.parametrize("dim", [2, 3])
def test_split_precision_step(dim):
    """compute(prec="bf16x3"): split-precision forward, split Grams, BF16
    reverse on the hi-plane tape: Grams within 1e-5 of the oracle, gradients
    against the bf16 step on the same fields' cotangents."""
    F = 2
    K = 24 if dim == 2 else 8
    cfg, net, onet = make(dim, K, [48, 40, 32], F, master=torch.float32)
    x = points(dim, 300, seed=7)
    w = 1.0 / 300
    L, M, S, *_ = _oracle_step(onet, x, F, cfg, w, 1e-2, 0.0)
    res = nb.compute(net, torch.tensor(x), w, prec="bf16x3", gram_mode="bf16x3", gate="cholesky")
    r = rel(res["M"].cpu(), M); assert r < 1e-5, r
    r = rel(res["S"].cpu(), S); assert r < 1e-5, r
    # tr(M^-1 S) moves by up to cond(M) times the Grams' relative error
    cond = np.linalg.cond(M)
    assert abs(float(res["loss"]) - L) < 1e-5 * cond * abs(L), (float(res["loss"]), L, cond)
    assert torch.isfinite(res["grads"]).all()
    rng = np.random.default_rng(21)
    dM = rng.normal(size=(K, K)); dM = dM + dM.T
    dS = rng.normal(size=(K, K)); dS = dS + dS.T
    gW, gb = _oracle_reverse(onet, x, F, cfg, w, dM, dS)
    tape = nw.trunk_forward(net, torch.tensor(x), nw.PRECISIONS["bf16x3"])
    fields = nw.output_fields(net, tape, split=True)
    g = nw.backward_fields(net, tape, fields, torch.tensor(dM, device=DEV), torch.tensor(dS, device=DEV), w)
    GW, Gb = net.split(g)
    for a, b in zip(GW, gW):
        r = rel(a.detach().cpu(), b); assert r < 4e-2, r




def test_autograd_function_split_precision():
    """JetGram on the split-precision path: Grams within 1e-5 of the oracle;
    its backward is the handwritten reverse on the hi-plane tape."""
    F, K = 2, 12
    cfg, net, onet = make(2, K, [32, 32], F, master=torch.float32)
    xn = points(2, 256, seed=2)
    x = torch.tensor(xn)
    w = 1.0 / 256
    out, _ = o.forward(onet, o.feature_jets(xn, F, cfg.lo, cfg.hi))
    u, om = o.lift2d(out)
    net.params.requires_grad_(True)
    M, S = nb.JetGram.apply(net.params, x, net, w, "bf16x3", "bf16x3")
    r = rel(M.detach().cpu(), o.gram(o.field_rows(u), w)); assert r < 1e-5, r
    r = rel(S.detach().cpu(), o.gram(om, w)); assert r < 1e-5, r
    rng = np.random.default_rng(4)
    gM = torch.tensor(rng.normal(size=(K, K)), device=DEV); gM = gM + gM.T
    gS = torch.tensor(rng.normal(size=(K, K)), device=DEV); gS = gS + gS.T
    g, = torch.autograd.grad((M, S), net.params, (gM, gS))
    net.params.requires_grad_(False)
    tape = nw.trunk_forward(net, x, nw.PRECISIONS["bf16x3"])
    fields = nw.output_fields(net, tape, split=True)
    ref = nw.backward_fields(net, tape, fields, gM, gS, w)
    r = rel(g.cpu(), ref.cpu()); assert r < 1e-12, r


def test_chunked_split_precision_step_3d():
    """Two-pass chunked step on split-precision jets (3D) against the single
    pass: the same kernels, Grams accumulated over chunks (their FP64 sums in
    another order)."""
    F, K = 2, 8
    cfg, net, _ = make(3, K, [32, 32], F, master=torch.float32)
    x = torch.tensor(points(3, 1000, seed=8))
    w = 1.0 / 1000
    ref = nb.compute(net, x, w, prec="bf16x3", gram_mode="bf16x3", gate="cholesky")
    res = nb.chunked_step(net, x, w, prec="bf16x3", gram_mode="bf16x3", chunk=256)
    r = rel(res["M"].cpu(), ref["M"].cpu()); assert r < 1e-6, r
    r = rel(res["S"].cpu(), ref["S"].cpu()); assert r < 1e-6, r
    r = rel(res["grads"].cpu(), ref["grads"].cpu()); assert r < 1e-2, r


def test_chunked_step_factored_grams():
    """Split-precision 3D chunked step with grams="factored" (no field
    matrices: lift-weighted channel Grams accumulated over the chunks, then
    M = w Wcat Ghat Wcat^T) against the same step with the field Grams and
    the FP64 chunked step: Grams within 3e-5 of FP64, gradient within the
    split step's own error of the field-Gram step."""
    F, K = 2, 48
    cfg, net, _ = make(3, K, [64, 64], F, master=torch.float32)
    x = torch.tensor(points(3, 3000, seed=21))
    w = 1.0 / 3000
    same = nb.SpectralSolver(1e-2, 1e-3)
    r64 = nb.chunked_step(net, x, w, prec="f64", gram_mode="f64", chunk=1024, solver=nb.SpectralSolver(1e-2, 1e-3))
    rf = nb.chunked_step(net, x, w, prec="bf16x3", gram_mode="bf16x3", chunk=1024, solver=same)
    rg = nb.chunked_step(net, x, w, prec="bf16x3", gram_mode="bf16x3", chunk=1024,
                         solver=nb.SpectralSolver(1e-2, 1e-3), grams="factored")
    for key in ("M", "S"):
        r = rel(rg[key].cpu(), r64[key].cpu()); assert r < 3e-5, (key, r)
        r = rel(rg[key].cpu(), rf[key].cpu()); assert r < 3e-5, (key, r)
    r = rel(rg["grads"].cpu(), rf["grads"].cpu()); assert r < 2e-3, r
    g64 = r64["grads"].cpu().double()
    r = rel(rg["grads"].cpu(), g64); assert r < 2e-2, r
    with pytest.raises(ValueError, match="factored"):
        nb.chunked_step(net, x, w, prec="f32", gram_mode="f32", chunk=1024, grams="factored")
