"""GPU parity of the 2D causal velocity Rayleigh peel step (HB2P:2935-3380)
through the C ABI: the Hessian lift (NEB_LIFT_2D_HESSIAN) fields, the peel
and its detached-prefix gradient against the oracle's literal two-pass peel
(oracle.velocity_rayleigh_peel), and the gates."""
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
from neb_amd import spectral as sp  # noqa: E402

DEV = "cuda"


def rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def make(K=12, widths=(40, 32), F=2, master=torch.float64, seed=0):
    cfg = nb.NetConfig(dim=2, frequencies=F, widths=widths, modes=K, stiffness="hessian")
    net = nb.EigenfluidNet(cfg, seed=seed, device=DEV, master_dtype=master)
    rng = np.random.default_rng(seed + 100)
    with torch.no_grad():
        for b in net.b:
            b.copy_(torch.as_tensor(rng.normal(size=b.shape) * 0.2, dtype=master))
    net.touch()
    onet = o.Net([w.detach().cpu().double().numpy() for w in net.W],
                 [b.detach().cpu().double().numpy() for b in net.b], "tanh", [1.0] * len(widths))
    return cfg, net, onet


def oracle_peel(onet, cfg, x, w, order="rayleigh"):
    out, tape = o.forward(onet, o.feature_jets(x, cfg.frequencies, cfg.lo, cfg.hi), record=True)
    n, K = out.value.shape
    grad = out.grad.transpose(0, 2, 1).reshape(2 * n, K)
    hess = out.hess.transpose(0, 2, 1).reshape(3 * n, K)
    res = o.velocity_rayleigh_peel(grad, hess, w, order=order)
    adj = o.Jet(np.zeros((n, K)), res["G_grad"].reshape(n, 2, K).transpose(0, 2, 1),
                res["G_hess"].reshape(n, 3, K).transpose(0, 2, 1))
    gW, gb = o.backward(onet, tape, adj)
    return res, gW, gb, out


@pytest.mark.parametrize("prec", ["f64", "f32", "bf16", "bf16x3"])
def test_hessian_lift_fields(prec):
    cfg, net, onet = make(K=20, master=torch.float32)
    x = np.random.default_rng(3).uniform(0, 2 * np.pi, (257, 2))
    out, _ = o.forward(onet, o.feature_jets(x, cfg.frequencies, cfg.lo, cfg.hi))
    u, h = o.lift2d_hessian(out)
    tape = nw.trunk_forward(net, torch.tensor(x), nw.PRECISIONS[prec])
    f = nw.output_fields(net, tape, split=prec in ("bf16", "bf16x3"))
    assert f.stiff.shape[0] == 3 * 257
    tol = {"f64": 1e-10, "f32": 2e-5, "bf16x3": 2e-5, "bf16": 3e-2}[prec]
    U = nw.field_matrix(f, "mass", torch.float64).cpu()
    D = nw.field_matrix(f, "stiff", torch.float64).cpu()
    r = rel(U, o.field_rows(u)); assert r < tol, r
    r = rel(D, o.field_rows(h)); assert r < tol, r
    # component-major layout holds the same rows
    f1 = nw.output_fields(net, tape, layout=1, split=prec in ("bf16", "bf16x3"))
    D1 = nw.field_matrix(f1, "stiff", torch.float64).cpu().numpy().reshape(3, 257, 20).transpose(1, 0, 2)
    assert np.array_equal(D1.reshape(3 * 257, 20), D.numpy())


@pytest.mark.parametrize("order", ["geometry_rayleigh", "fixed_order"])
def test_velocity_peel_step_fp64_matches_oracle(order):
    cfg, net, onet = make()
    x = np.random.default_rng(7).uniform(0, 2 * np.pi, (400, 2))
    w = 4 * math.pi ** 2 / 400
    res, gW, gb, _ = oracle_peel(onet, cfg, x, w, "rayleigh" if order == "geometry_rayleigh" else "fixed")
    out = sp.velocity_peel_step(net, torch.tensor(x), w, prec="f64", causal_sorting=order)
    assert np.array_equal(out["causal_order"].cpu().numpy(), res["order"])
    r = rel(out["mode_energies"].cpu(), res["rayleigh"]); assert r < 1e-9, r
    assert abs(out["loss"] - float(np.mean(res["rayleigh"]))) < 1e-9 * abs(out["loss"])
    assert abs(out["min_relative_residual_mass"] - res["relative_mass"].min()) < 1e-9
    GW, Gb = net.split(out["grads"])
    for a, b in zip(GW, gW):
        r = rel(a.detach().cpu(), b); assert r < 1e-8, r
    for a, b in zip(Gb[:-1], gb[:-1]):
        r = rel(a.detach().cpu(), b); assert r < 1e-8, r


def test_velocity_peel_step_split_precision_and_adam():
    """The split-precision step (bf16 hi + lo jets and fields, split Grams,
    bf16 reverse) against the oracle, and its gated Adam update."""
    cfg, net, onet = make(master=torch.float32)
    x = np.random.default_rng(8).uniform(0, 2 * np.pi, (512, 2))
    w = 4 * math.pi ** 2 / 512
    res, gW, gb, _ = oracle_peel(onet, cfg, x, w)
    before = net.params.clone()
    out = sp.velocity_peel_step(net, torch.tensor(x), w, prec="bf16x3", lr=1e-4)
    assert np.array_equal(out["causal_order"].cpu().numpy(), res["order"])
    r = rel(out["mode_energies"].cpu(), res["rayleigh"]); assert r < 1e-4, r
    # the bf16 reverse on the hi-plane tape under the peel's cotangents:
    # per layer within 10 %, the whole gradient's direction within 0.5 %
    GW, _ = net.split(out["grads"])
    for a, b in zip(GW, gW):
        r = rel(a.detach().cpu(), b); assert r < 1e-1, r
    g = np.concatenate([a.detach().cpu().double().numpy().ravel() for a in GW])
    ref = np.concatenate([b.ravel() for b in gW])
    cos = float(g @ ref / (np.linalg.norm(g) * np.linalg.norm(ref)))
    assert cos > 0.995, cos
    assert net.epoch == 1 and not torch.equal(net.params, before)


def test_velocity_peel_gate_rejects_repeated_mode_without_update():
    """Two identical candidates: the second's residual mass collapses, the
    relative-mass gate fails on the device and nothing is updated."""
    cfg, net, _ = make(master=torch.float32)
    with torch.no_grad():
        net.W[-1][5].copy_(net.W[-1][2])
        net.b[-1][5] = net.b[-1][2]
    net.touch()
    x = torch.tensor(np.random.default_rng(9).uniform(0, 2 * np.pi, (300, 2)))
    before, m0 = net.params.clone(), net.m.clone()
    with pytest.raises(FloatingPointError):
        sp.velocity_peel_step(net, x, 1.0 / 300, prec="bf16x3", lr=1e-3)
    assert torch.equal(net.params, before) and torch.equal(net.m, m0) and net.epoch == 0
