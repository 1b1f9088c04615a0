"""GPU parity of the 2D Dirichlet streamfunction Rayleigh peel step
(HB2P:3383-3770) through the C ABI: the streamfunction lift
(NEB_LIFT_2D_STREAM) fields with the periodic gauge (neb_center_columns),
the peel and its detached-prefix gradient against the oracle's literal
two-pass peel (oracle.streamfunction_rayleigh_peel), and the gates."""
import math

import numpy as np
import pytest
import torch

from oracle import neb_oracle as o

pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():  # CPU-only host: the gpu tier runs on the MI355X box
    pytest.skip("needs a GPU", allow_module_level=True)

import neb_amd as nb  # noqa: E402
from neb_amd import _hip as H  # noqa: E402
from neb_amd import network as nw  # noqa: E402
from neb_amd import spectral as sp  # noqa: E402

DEV = "cuda"


def rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def make(K=12, widths=(40, 32), F=2, master=torch.float64, seed=0):
    cfg = nb.NetConfig(dim=2, frequencies=F, widths=widths, modes=K, stiffness="stream")
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
    value, g = o.lift2d_stream(out)
    res = o.streamfunction_rayleigh_peel(value, o.field_rows(g), w, order=order)
    adj = o.lift2d_stream_reverse(res["G_value"], res["G_grad"].reshape(n, 2, K))
    gW, gb = o.backward(onet, tape, adj)
    return res, gW, gb, out


@pytest.mark.parametrize("dtype,split", [(torch.float32, False), (torch.float64, False), (torch.bfloat16, True)])
def test_center_columns(dtype, split):
    rows, cols, ld = 1000, 37, 40
    rng = np.random.default_rng(1)
    x = rng.normal(size=(rows, cols)) * 0.3 + rng.normal(size=cols) * 4.0
    mean = torch.tensor(x.mean(0), dtype=torch.float64, device=DEV)
    s = torch.cuda.current_stream().cuda_stream
    if split:
        hi = torch.zeros((rows, ld), dtype=torch.bfloat16, device=DEV)
        lo = torch.zeros_like(hi)
        src = torch.tensor(x, dtype=torch.float32, device=DEV)
        H.call("neb_split_bf16", H.F32, src.data_ptr(), rows, cols, cols, hi.data_ptr(), lo.data_ptr(), ld, s)
        # the planes' own value minus the mean, split again: within the 2^-17
        # relative rounding of the new hi / lo pair
        ref = (hi.double() + lo.double())[:, :cols].cpu().numpy() - x.mean(0)
        H.call("neb_center_columns", H.BF16, hi.data_ptr(), lo.data_ptr(), rows, cols, ld, mean.data_ptr(), s)
        got = (hi.double() + lo.double())[:, :cols].cpu().numpy()
        assert np.all(np.abs(got - ref) <= 2.0 ** -16 * np.abs(ref) + 1e-30)
        assert float(hi[:, cols:].abs().max()) == 0.0           # padding untouched
    else:
        t = torch.zeros((rows, ld), dtype=dtype, device=DEV)
        t[:, :cols] = torch.tensor(x, dtype=dtype)
        ref = t[:, :cols].double().cpu().numpy() - x.mean(0)
        H.call("neb_center_columns", nw._DT[dtype], t.data_ptr(), None, rows, cols, ld, mean.data_ptr(), s)
        tol = 1e-14 if dtype == torch.float64 else 1e-6
        assert np.abs(t[:, :cols].double().cpu().numpy() - ref).max() <= tol * np.abs(ref).max()
        assert float(t[:, cols:].abs().max()) == 0.0
    with pytest.raises(ValueError):
        H.call("neb_center_columns", H.F32, 1, 1, rows, cols, ld, mean.data_ptr(), s)   # lo plane needs bf16


@pytest.mark.parametrize("prec", ["f64", "f32", "bf16", "bf16x3"])
def test_stream_lift_fields_centered(prec):
    cfg, net, onet = make(K=20, master=torch.float32)
    x = np.random.default_rng(3).uniform(0, 2 * np.pi, (257, 2))
    out, _ = o.forward(onet, o.feature_jets(x, cfg.frequencies, cfg.lo, cfg.hi))
    v, g = o.lift2d_stream(out)
    tape = nw.trunk_forward(net, torch.tensor(x), nw.PRECISIONS[prec])
    f = nw.output_fields(net, tape, split=prec in ("bf16", "bf16x3"))
    assert f.mass.shape[0] == 257 and f.stiff.shape[0] == 2 * 257
    tol = {"f64": 1e-10, "f32": 2e-5, "bf16x3": 2e-5, "bf16": 3e-2}[prec]
    U = nw.field_matrix(f, "mass", torch.float64).cpu()
    D = nw.field_matrix(f, "stiff", torch.float64).cpu()
    r = rel(U, v - v.mean(0)); assert r < tol, r
    r = rel(D, o.field_rows(g)); assert r < tol, r
    assert float(U.mean(0).abs().max()) < 1e-6 * float(U.abs().max())
    f1 = nw.output_fields(net, tape, layout=1, split=prec in ("bf16", "bf16x3"))
    D1 = nw.field_matrix(f1, "stiff", torch.float64).cpu().numpy().reshape(2, 257, 20).transpose(1, 0, 2)
    assert np.array_equal(D1.reshape(2 * 257, 20), D.numpy())


@pytest.mark.parametrize("order", ["geometry_rayleigh", "fixed_order"])
def test_stream_peel_step_fp64_matches_oracle(order):
    cfg, net, onet = make()
    x = np.random.default_rng(7).uniform(0, 2 * np.pi, (400, 2))
    w = 4 * math.pi ** 2 / 400
    res, gW, gb, _ = oracle_peel(onet, cfg, x, w, "rayleigh" if order == "geometry_rayleigh" else "fixed")
    out = sp.streamfunction_peel_step(net, torch.tensor(x), w, prec="f64", causal_sorting=order)
    assert np.array_equal(out["causal_order"].cpu().numpy(), res["order"])
    assert np.array_equal(out["causal_order_1based"].cpu().numpy(), res["order"] + 1)
    r = rel(out["mode_energies"].cpu(), res["rayleigh"]); assert r < 1e-9, r
    assert abs(out["loss"] - float(np.mean(res["rayleigh"]))) < 1e-9 * abs(out["loss"])
    assert abs(out["mean_projection_ratio"] - float(np.mean(res["projection"]))) < 1e-8 * max(
        1.0, abs(out["mean_projection_ratio"]))
    r = rel(out["sorting_rayleigh_by_network"].cpu(), res["raw_stiffness"] / res["raw_mass"]); assert r < 1e-10, r
    GW, Gb = net.split(out["grads"])
    for a, b in zip(GW, gW):
        r = rel(a.detach().cpu(), b); assert r < 1e-8, r
    for a, b in zip(Gb[:-1], gb[:-1]):
        r = rel(a.detach().cpu(), b); assert r < 1e-8, r
    # the output bias moves every psi by a constant the gauge removes
    assert float(Gb[-1].abs().max()) == 0.0 and np.abs(gb[-1]).max() < 1e-10 * np.abs(gb[0]).max()


def test_stream_peel_step_split_precision_and_adam():
    """The split-precision step (bf16 hi + lo jets and centered fields, split
    Grams, bf16 reverse) against the oracle, and its gated Adam update."""
    cfg, net, onet = make(master=torch.float32)
    x = np.random.default_rng(8).uniform(0, 2 * np.pi, (512, 2))
    w = 4 * math.pi ** 2 / 512
    res, gW, gb, _ = oracle_peel(onet, cfg, x, w)
    before = net.params.clone()
    out = sp.streamfunction_peel_step(net, torch.tensor(x), w, prec="bf16x3", lr=1e-4)
    assert np.array_equal(out["causal_order"].cpu().numpy(), res["order"])
    # the quotients of this random 12-mode peel move 1e3 times the field
    # error: 2.5e-4 already with FP32 jets, 2.4e-4 split
    r = rel(out["mode_energies"].cpu(), res["rayleigh"]); assert r < 1e-3, r
    # the bf16 reverse on the hi-plane tape: the hidden layers within 20 %
    # (measured 11 % and 14.5 %, FP32 jets 8e-4), the output layer 0.4 %,
    # the whole gradient's direction within 1 % (cosine measured 0.9930)
    GW, _ = net.split(out["grads"])
    for a, b in zip(GW, gW):
        r = rel(a.detach().cpu(), b); assert r < 2e-1, r
    g = np.concatenate([a.detach().cpu().double().numpy().ravel() for a in GW])
    ref = np.concatenate([b.ravel() for b in gW])
    cos = float(g @ ref / (np.linalg.norm(g) * np.linalg.norm(ref)))
    assert cos > 0.99, cos
    assert net.epoch == 1 and not torch.equal(net.params, before)


def test_stream_peel_gate_rejects_repeated_mode_without_update():
    """Two identical candidates: the second's residual mass collapses, the
    mass gate fails on the device and nothing is updated."""
    cfg, net, _ = make(master=torch.float32)
    with torch.no_grad():
        net.W[-1][5].copy_(net.W[-1][2])
    net.touch()
    x = torch.tensor(np.random.default_rng(9).uniform(0, 2 * np.pi, (300, 2)))
    before, m0 = net.params.clone(), net.m.clone()
    with pytest.raises(FloatingPointError):
        sp.streamfunction_peel_step(net, x, 1.0 / 300, prec="bf16x3", lr=1e-3, mass_floor=1e-8)
    assert torch.equal(net.params, before) and torch.equal(net.m, m0) and net.epoch == 0


def test_stream_lift_rejects_factored_reverse():
    cfg = nb.NetConfig(dim=2, frequencies=2, widths=(32, 32), modes=8, stiffness="stream", split_reverse="factored")
    net = nb.EigenfluidNet(cfg, seed=0, device=DEV)
    tape = nw.trunk_forward(net, torch.tensor(np.random.default_rng(0).uniform(0, 6, (64, 2))), nw.PRECISIONS["bf16x3"])
    with pytest.raises(ValueError, match="field-form"):
        nw.factored_reverse_ok(net, tape)
