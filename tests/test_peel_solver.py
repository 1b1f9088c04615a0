"""The coefficient peel (neb_amd.spectral.coefficient_peel,
MB/velocity/core.py:48-69) on small FP64 Grams against the oracle, and its
gate through PeelSolver: the FP64 algebra around the inverse Cholesky
factors of neb_spd_inverse_chol."""
import numpy as np
import pytest
import torch

from neb_amd import spectral as sp
from oracle import neb_oracle as o

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _grams(seed=4, n=300, k=11):
    rng = np.random.default_rng(seed)
    U = rng.normal(size=(2 * n, k)) @ np.diag(rng.uniform(0.3, 3.0, k))
    D = U @ rng.normal(size=(k, k)) + 0.2 * rng.normal(size=(2 * n, k))
    w = 1.0 / n
    return U, D, w, o.gram(U, w), o.gram(D, w)


def test_coefficient_peel_small_matches_oracle():
    U, D, w, M, S = _grams()
    ref = o.coefficient_peel(M, S)
    p = {k: (v.cpu() if torch.is_tensor(v) else v)
         for k, v in sp.coefficient_peel(torch.tensor(M, device=DEV), torch.tensor(S, device=DEV)).items()}
    np.testing.assert_array_equal(p["order"].numpy(), ref["order"])
    for key in ("losses", "T", "Cu", "Cd", "relative_mass"):
        np.testing.assert_allclose(p[key].numpy(), ref[key], rtol=1e-10, atol=1e-12 * np.abs(ref[key]).max())
    assert float(p["loss"]) == pytest.approx(ref["losses"].sum(), rel=1e-12)
    K = M.shape[0]
    np.testing.assert_allclose(p["pair"][:, :K].numpy(), 0.5 * ref["Cu"].T, rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(p["pair"][:, K:].numpy(), 0.5 * ref["Cd"].T, rtol=1e-10, atol=1e-12)
    assert int(p["gate"]) == 0


def test_coefficient_peel_gate():
    U, D, w, M, S = _grams(k=6)
    Ud = np.column_stack((U, U[:, 1] * 2.0))
    Dd = np.column_stack((D, D[:, 1] * 2.0))
    Md, Sd = o.gram(Ud, w), o.gram(Dd, w)
    with pytest.raises(FloatingPointError, match="gate failed"):
        sp.coefficient_peel(torch.tensor(Md, device=DEV), torch.tensor(Sd, device=DEV))
    p = sp.coefficient_peel(torch.tensor(Md, device=DEV), torch.tensor(Sd, device=DEV), deferred=True)
    assert int(p["gate"]) == sp.PEEL_GATE
    solver = sp.PeelSolver(deferred=True)
    solver(torch.tensor(Md, device=DEV), torch.tensor(Sd, device=DEV))
    with pytest.raises(FloatingPointError, match="gate failed"):
        solver.check()
