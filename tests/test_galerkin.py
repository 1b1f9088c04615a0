"""GPU parity of the Galerkin reduced-order operators (neb_amd.galerkin)
against the oracle (oracle/neb_oracle.py: rollout_rhs, etdrk4,
periodic_reduced_basis, SpectralGrid, derive_fields, assemble_operators),
through the C ABI."""
import math

import numpy as np
import pytest
import torch

from oracle import neb_oracle as o

pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():  # CPU-only host: the gpu tier runs on the MI355X box
    pytest.skip("needs a GPU", allow_module_level=True)

import neb_amd as nb  # noqa: E402
from neb_amd import galerkin as gk  # noqa: E402
from neb_amd import network as nw  # noqa: E402

DEV = "cuda"


def rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("n,K", [(5003, 96), (131, 1030)])
def test_rollout_rhs_kernel(dtype, n, K):
    rng = np.random.default_rng(n + K)
    ld = K + 6
    f = torch.tensor(rng.normal(size=(3, n, ld)), device=DEV).to(dtype)
    gu, gv, gw = (f[i, :, :K] for i in range(3))
    a = rng.normal(size=K)
    w = 0.37
    r = gk.rollout_rhs(gu, gv, gw, torch.tensor(a, device=DEV), w)
    ref = o.rollout_rhs(*(t.double().cpu().numpy() for t in (gu, gv, gw)), a, w)
    err = rel(r.cpu(), ref)
    assert err < 1e-12, err
    # energy skew: a . rhs(a) = 0
    assert abs(float(torch.dot(r, torch.tensor(a, device=DEV)))) < 1e-10 * np.linalg.norm(a) * np.linalg.norm(ref)


def test_periodic_rom_matches_oracle():
    """Reduced basis of the network's velocity / vorticity fields (full-rank
    K = 24 < width), eigenvalues, and a 20-step ETDRK4 rollout of a projected
    Taylor-Green velocity, compared as physical velocity fields (eigenvector
    signs are arbitrary)."""
    K, F = 24, 3
    cfg = nb.NetConfig(dim=2, frequencies=F, widths=(48, 40), modes=K)
    net = nb.EigenfluidNet(cfg, seed=3, device=DEV, master_dtype=torch.float64)
    onet = o.Net([w.cpu().numpy() for w in net.W], [b.cpu().numpy() for b in net.b], "tanh", [1.0, 1.0])
    x = o.periodic_grid(48)
    w = (2 * math.pi / 48) ** 2
    tape = nw.trunk_forward(net, torch.tensor(x), nw.PRECISIONS["f64"])
    rom = gk.PeriodicROM(nw.output_fields(net, tape), w)
    out, _ = o.forward(onet, o.feature_jets(x, F, cfg.lo, cfg.hi))
    u, om = o.lift2d(out)
    bu, bv, bw, vals = o.periodic_reduced_basis(u[:, 0], u[:, 1], om, w)
    r = rel(rom.vals.cpu(), vals); assert r < 1e-9, r
    u0 = np.sin(x[:, 0]) * np.cos(x[:, 1])
    v0 = -np.cos(x[:, 0]) * np.sin(x[:, 1]) + 0.3 * np.sin(2 * x[:, 0])
    c0 = rom.project(torch.tensor(u0, device=DEV), torch.tensor(v0, device=DEV))
    oc0 = w * (bu.T @ u0 + bv.T @ v0)
    assert rel((rom.gu @ c0).cpu(), bu @ oc0) < 1e-9
    assert rom.skew(c0) < 1e-12
    nu, h = 0.01, 0.02
    c = rom.rollout(c0, nu, h, 20)
    oc = o.etdrk4(lambda a: o.rollout_rhs(bu, bv, bw, a, w), -nu * vals, oc0, h, 20)
    r = rel(torch.cat([rom.gu @ c, rom.gv @ c]).cpu(), np.concatenate([bu @ oc, bv @ oc])); assert r < 1e-8, r


def test_periodic_rom_rank_gate():
    """More modes than the trunk spans (K = 40 > 2 * width 16 fields): the
    full-rank gate of paper_rollout.py:180-186 raises."""
    cfg = nb.NetConfig(dim=2, frequencies=2, widths=(8,), modes=40)
    net = nb.EigenfluidNet(cfg, seed=1, device=DEV, master_dtype=torch.float64)
    tape = nw.trunk_forward(net, torch.tensor(o.periodic_grid(24)), nw.PRECISIONS["f64"])
    with pytest.raises(ArithmeticError):
        gk.PeriodicROM(nw.output_fields(net, tape), (2 * math.pi / 24) ** 2)


def test_derive_fields_matches_oracle():
    n, N, K = 16, 900, 5
    rng = np.random.default_rng(11)
    flat = np.sort(rng.choice(n ** 3, N, replace=False))
    B = rng.normal(size=(N, 3, K))
    normal = rng.normal(size=(n, n, n, 3))
    chi = (rng.uniform(size=(n, n, n)) > 0.6).astype(float)
    D = gk.derive_fields(torch.tensor(B, device=DEV), torch.tensor(flat), gk.SpectralGrid(normal, chi, n), block=2)
    ref = o.derive_fields(B, flat, o.SpectralGrid(normal, chi, n))
    r = rel(D.cpu(), ref); assert r < 1e-12, r


def test_assemble_operators_matches_oracle():
    rng = np.random.default_rng(12)
    N, K = 2000, 7
    B = rng.normal(size=(N, 3, K))
    D = rng.normal(size=(N, 9, K))
    dV = 1.0 / 128 ** 3
    M, A, Q = gk.assemble_operators(torch.tensor(B, device=DEV), torch.tensor(D, device=DEV), dV)
    Mo, Ao, Qo = o.assemble_operators(B, D, dV)
    for got, ref in ((M, Mo), (A, Ao), (Q, Qo)):
        r = rel(got.cpu(), ref); assert r < 1e-12, r
