"""Config C at full size on the GPU: the K=1024 differentiation step.

65,536 uniform random points on [0, 2pi)^2, sin/cos features 1..8, 4x256
tanh, K=1024 stream-function modes, FP64 jets, fields, Grams, objective and
reverse; L = tr((M + ridge I)^-1 S) + 1e-2 ||M - I||_F^2 with ridge =
1e-4 mean diag M at the base point (M has rank <= 256 < K).  The handwritten
gradient is checked against central differences (h = 1e-3) along 64 random
unit parameter directions, the reference's directional check
(tests/test_network_backends.py:74-90), to 1e-4.

The loss's own rounding noise is what a central difference divides by h, so
it is pinned first: the product objective's loss against a Cholesky solve of
the same Grams.
"""
import math

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():  # CPU-only host: the gpu tier runs on the MI355X box
    pytest.skip("needs a GPU", allow_module_level=True)

import neb_amd as nb  # noqa: E402
from neb_amd import network as nw  # noqa: E402

DEV = "cuda"


def _loss(net, x, w, ridge):
    return float(nb.compute(net, x, w, prec="f64", gram_mode="f64", ridge=ridge, training=False,
                            gate="cholesky")["loss"])


def test_config_C_fd_h1e3_64_directions():
    n, K = 65536, 1024
    rng = np.random.default_rng(0)
    x = torch.tensor(rng.uniform(0.0, 2 * math.pi, (n, 2)), dtype=torch.float64, device=DEV)
    w = 4 * math.pi ** 2 / n
    cfg = nb.NetConfig(dim=2, frequencies=8, widths=(256,) * 4, modes=K)
    net = nb.EigenfluidNet(cfg, seed=0, device=DEV, master_dtype=torch.float64)
    M0 = nb.compute(net, x, w, prec="f64", gram_mode="f64", training=False, gate="cholesky", ridge=1.0)["M"]
    ridge = 1e-4 * float(torch.diagonal(M0).mean())
    del M0
    res = nb.compute(net, x, w, prec="f64", gram_mode="f64", ridge=ridge, gate="cholesky")
    g, L = res["grads"].clone(), float(res["loss"])
    # noise floor of the loss: the refined sweep inverse against a Cholesky
    # solve on the same Grams (cond(M + ridge I) ~ 1e6)
    M, S = res["M"], res["S"]
    eye = torch.eye(K, dtype=torch.float64, device=DEV)
    Lc = torch.linalg.cholesky(0.5 * (M + M.T) + ridge * eye)
    L_chol = float(torch.trace(torch.cholesky_solve(0.5 * (S + S.T), Lc)) + 1e-2 * ((M - eye) ** 2).sum())
    assert abs(L - L_chol) <= 1e-10 * abs(L), (L, L_chol)
    del res, M, S, Lc
    base = net.params.detach().clone()
    drs = np.random.default_rng(5)
    h = 1e-3
    errs, plain = [], []
    for _ in range(64):
        d = torch.tensor(drs.normal(size=base.numel()), dtype=torch.float64, device=DEV)
        d /= d.norm()
        exact = float(g @ d)
        vals = []
        for sgn in (1.0, -1.0):
            with torch.no_grad():
                net.params.copy_(base + sgn * h * d)
            net.touch()
            vals.append(_loss(net, x, w, ridge))
        fd = (vals[0] - vals[1]) / (2 * h)
        errs.append(abs(fd - exact) / max(abs(exact), 1e-3 * abs(L)))
        plain.append(abs(fd - exact) / abs(exact))
    with torch.no_grad():
        net.params.copy_(base)
    net.touch()
    errs, plain = np.asarray(errs), np.asarray(plain)
    print(f"config C FD h=1e-3: max {errs.max():.2e} median {np.median(errs):.2e}; "
          f"plain relative max {plain.max():.2e}", flush=True)
    assert (errs <= 1e-4).all(), (int((errs <= 1e-4).sum()), float(errs.max()))
    assert np.median(plain) <= 1e-4, float(np.median(plain))
