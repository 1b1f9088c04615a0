"""GPU parity of the envelope product (the output GEMM's epilogue,
neb_jet_output_jets_enveloped) and its transpose (neb_envelope_reverse)
within the jet-level forward / reverse (output_jets, backward_jets; the
analogue of VJ2.forward / VJ2.backward, VJ2:47-82) against the oracle."""
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
TOL = {"f64": 1e-10, "f32": 2e-5, "bf16": 3e-2}
REV = {"f64": 1e-9, "f32": 2e-5, "bf16": 4e-2}


def rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def setup(dim, prec, K=12, n=301, seed=0):
    lo, hi = ((0.0, 0.0), (2 * np.pi, 2 * np.pi)) if dim == 2 else ((-1.0,) * 3, (1.0,) * 3)
    cfg = nb.NetConfig(dim=dim, frequencies=2, lo=lo, hi=hi, widths=(48, 40), modes=K)
    master = torch.float64 if prec == "f64" else torch.float32
    net = nb.EigenfluidNet(cfg, seed=seed, device=DEV, master_dtype=master)
    rng = np.random.default_rng(seed + 9)
    with torch.no_grad():
        for b in net.b:
            b.copy_(torch.as_tensor(rng.normal(size=b.shape) * 0.2, dtype=master))
    net.touch()
    onet = o.Net([w.cpu().double().numpy() for w in net.W], [b.cpu().double().numpy() for b in net.b], "tanh",
                 [1.0, 1.0])
    x = rng.uniform(lo[0], hi[0], (n, dim))
    c = 0.25
    e = np.exp(-c * np.sum(x * x, 1))
    ej = o.Jet(e, -2 * c * x * e[:, None],
               np.stack([(4 * c * c * x[:, i] * x[:, j] - (2 * c if i == j else 0.0)) * e for i, j in o.pairs(dim)],
                        -1))
    env = nw.envelope_jets(torch.tensor(ej.value), torch.tensor(ej.grad), torch.tensor(ej.hess),
                           torch.float64 if prec == "f64" else torch.float32)
    return cfg, net, onet, x, ej, env


@pytest.mark.parametrize("prec", ["f64", "f32", "bf16"])
@pytest.mark.parametrize("dim", [2, 3])
def test_enveloped_output_jets(dim, prec):
    cfg, net, onet, x, ej, env = setup(dim, prec)
    tape = nw.trunk_forward(net, torch.tensor(x), nw.PRECISIONS[prec])
    v, g, h = nw.output_jets(net, tape, envelope=env.to(DEV))
    raw, _ = o.forward(onet, o.feature_jets(x, 2, cfg.lo, cfg.hi))
    ref = o.envelope_product(ej, raw)
    tol = TOL[prec]
    for got, want in ((v, ref.value), (g, ref.grad), (h, ref.hess)):
        r = rel(got.cpu(), want); assert r < tol * (3 if prec == "bf16" else 1), r
    # the identity envelope reproduces the raw jets
    one = torch.zeros_like(env); one[0] = 1.0
    v1, g1, h1 = nw.output_jets(net, tape, envelope=one.to(DEV))
    v0, g0, h0 = nw.output_jets(net, tape)
    assert torch.equal(v1, v0) and torch.equal(g1, g0) and torch.equal(h1, h0)


@pytest.mark.parametrize("prec", ["f64", "f32", "bf16"])
@pytest.mark.parametrize("dim", [2, 3])
def test_backward_jets_with_envelope(dim, prec):
    cfg, net, onet, x, ej, env = setup(dim, prec, seed=1)
    n, nout, P = len(x), cfg.ncomp * cfg.modes, dim * (dim + 1) // 2
    rng = np.random.default_rng(5)
    adj = o.Jet(rng.normal(size=(n, nout)), rng.normal(size=(n, nout, dim)), rng.normal(size=(n, nout, P)))
    _, tape_o = o.forward(onet, o.feature_jets(x, 2, cfg.lo, cfg.hi), record=True)
    for envelope in (None, env):
        a = adj if envelope is None else o.envelope_product_reverse(ej, adj)
        gW, gb = o.backward(onet, tape_o, a)
        tape = nw.trunk_forward(net, torch.tensor(x), nw.PRECISIONS[prec])
        g = nw.backward_jets(net, tape, torch.tensor(adj.value), torch.tensor(adj.grad), torch.tensor(adj.hess),
                             envelope=None if envelope is None else envelope.to(DEV))
        GW, Gb = net.split(g)
        for l, (p_, q) in enumerate(zip(GW, gW)):
            r = rel(p_.cpu(), q); assert r < REV[prec], (l, r)
        for l, (p_, q) in enumerate(zip(Gb, gb)):
            r = rel(p_.cpu(), q); assert r < REV[prec], (l, r)
