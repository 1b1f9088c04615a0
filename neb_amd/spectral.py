"""Spectral objective, FP64 orthonormalisation and full-mass projection on
the GPU, plus the training-step driver.

FP64 K x K work (Cholesky, inverse, generalized eigh, mass solve) runs in
torch.linalg on the GPU (rocSOLVER / rocBLAS); every N-sized contraction
runs in libnebhip.so.

Reference surfaces mirrored (paths relative to the reference root;
MB = experiments/metaball/_backend/flow/cases/metaball):
  trace_objective        MB/velocity_trace_jet2.py:107-117
  grams / compute        MB/velocity/core.py:40-46, 79-105
  generalized eigh       experiments/periodic/paper_rollout.py:170-188
  project                experiments/metaball/_backend/exports/basis_projection/
                         postprocessing/paired_projection.py:141-177
"""
from __future__ import annotations

import time
from contextlib import contextmanager

import torch

from . import network as nw


def _sym(A):
    return 0.5 * (A + A.T)


def spectral_objective(M: torch.Tensor, S: torch.Tensor, lam: float = 1e-2, ridge: float = 0.0,
                       cond_limit: float = 1e12, gate: str = "eig"):
    """L = tr((M + ridge I)^-1 S) + lam ||M - I||_F^2 and its closed-form
    adjoints dL/dM = -(M_r^-1 S M_r^-1) + 2 lam (M - I), dL/dS = M_r^-1.

    The gate follows MB/velocity_trace_jet2.py:109-111: a nonfinite,
    nonpositive or too ill-conditioned mass raises FloatingPointError before
    anything is updated.  gate="eig" checks eigenvalues (the reference's
    rule); gate="cholesky" only requires the FP64 Cholesky to succeed.
    Returns (loss: float, dM, dS) with FP64 tensors on M's device."""
    K = M.shape[0]
    eye = torch.eye(K, dtype=torch.float64, device=M.device)
    Mr = _sym(M) + ridge * eye
    if gate == "eig":
        e = torch.linalg.eigvalsh(Mr)
        ok = bool(torch.isfinite(e).all()) and float(e[-1]) > 0 and float(e[0] / e[-1]) >= 1.0 / cond_limit
        if not ok:
            raise FloatingPointError("Physical mass zero/rank deficient/condition exceeds limit; no update")
    L, info = torch.linalg.cholesky_ex(Mr)
    if gate != "eig" and int(info.item()) != 0:
        raise FloatingPointError("mass Cholesky failed; no update")
    inv = _sym(torch.cholesky_inverse(L))
    Ss = _sym(S)
    D = M - eye
    loss = torch.sum(inv * Ss) + lam * torch.sum(D * D)
    cross = inv @ Ss @ inv
    dM = -_sym(cross) + 2.0 * lam * D
    return loss, dM, inv


def generalized_eigh(S: torch.Tensor, M: torch.Tensor, rank_tol: float = 1e-12):
    """FP64 generalized symmetric eigendecomposition S x = lambda M x via
    Cholesky reduction, with the full-rank gate of
    experiments/periodic/paper_rollout.py:180-186 (ArithmeticError when the
    numerical rank of M is below K)."""
    Ms, Ss = _sym(M), _sym(S)
    ev = torch.linalg.eigvalsh(Ms)
    rank = int((ev > ev[-1] * rank_tol).sum())
    if rank != M.shape[0]:
        raise ArithmeticError(f"Full-K rank gate: rank {rank} < K={M.shape[0]}")
    L = torch.linalg.cholesky(Ms)
    Li = torch.linalg.solve_triangular(L, torch.eye(len(M), dtype=M.dtype, device=M.device), upper=False)
    C = _sym(Li @ Ss @ Li.T)
    vals, V = torch.linalg.eigh(C)
    return vals, Li.T @ V


@contextmanager
def timed(record: dict, name: str):
    """Stage timer with device synchronisation on both ends (VCORE:25-30)."""
    torch.cuda.synchronize()
    t = time.perf_counter()
    try:
        yield
    finally:
        torch.cuda.synchronize()
        record[name] = record.get(name, 0.0) + time.perf_counter() - t


def compute(net: nw.EigenfluidNet, points: torch.Tensor, weight: float, prec: str = "bf16",
            gram_mode: str = "bf16x3", lam: float = 1e-2, ridge: float = 0.0, training: bool = True,
            gate: str = "eig", timings: dict | None = None):
    """One spectral step (the analogue of MB/velocity/core.py:79-105
    ``compute``): jets -> lifted fields -> FP64 Grams -> spectral objective ->
    fused reverse.  Returns dict(loss, M, S, grads, fields)."""
    p = nw.PRECISIONS[prec]
    tt = {} if timings is None else timings
    K = net.cfg.modes
    with timed(tt, "neural_forward_analytic_derivatives"):
        tape = nw.trunk_forward(net, points, p)
        fields = nw.output_fields(net, tape)
    with timed(tt, "gram_assembly"):
        M = nw.gram(fields.mass, K, weight, gram_mode)
        S = nw.gram(fields.stiff, K, weight, gram_mode)
    with timed(tt, "spectral_objective"):
        loss, dM, dS = spectral_objective(M, S, lam, ridge, gate=gate)
    out = dict(loss=loss, M=M, S=S, fields=fields, timings=tt)
    if training:
        with timed(tt, "analytic_backward"):
            out["grads"] = nw.backward_fields(net, tape, fields, dM, dS, weight)
    return out


class JetGram(torch.autograd.Function):
    """(params, points) -> (M, S): analytic-jet forward, lift and FP64 Grams;
    the backward is the handwritten reverse pass (no autograd tape through
    the network).  ``params`` must be ``net.params``."""

    @staticmethod
    def forward(ctx, params, points, net, weight, prec, gram_mode):
        p = nw.PRECISIONS[prec]
        tape = nw.trunk_forward(net, points, p)
        fields = nw.output_fields(net, tape)
        K = net.cfg.modes
        M = nw.gram(fields.mass, K, weight, gram_mode)
        S = nw.gram(fields.stiff, K, weight, gram_mode)
        ctx.state = (net, tape, fields, weight)
        return M, S

    @staticmethod
    def backward(ctx, gM, gS):
        net, tape, fields, weight = ctx.state
        K = net.cfg.modes
        z = torch.zeros((K, K), dtype=torch.float64, device=net.device)
        gM = z if gM is None else gM
        gS = z if gS is None else gS
        grads = nw.backward_fields(net, tape, fields, _sym(gM), _sym(gS), weight)
        ctx.state = None
        return grads, None, None, None, None, None


def project(mass: torch.Tensor, basis: torch.Tensor, targets: torch.Tensor, weight: float,
            mode: str = "f64", ridge: float = 0.0):
    """Full-mass projection of targets [R, m] onto the basis fields [R, K]:
    loads = w B^T T (cross kernel), FP64 solve, normal-residual gate, direct
    second-pass residual and the Gram identity (PP:141-177)."""
    K = mass.shape[0]
    mcode = nw.GRAM_MODES[mode]
    R = basis.shape[0]
    loads = nw.cross([(basis, targets, R, 1.0)], K, targets.shape[1], mcode, scale=weight)
    tt = nw.cross([(targets, targets, R, 1.0)], targets.shape[1], targets.shape[1], mcode, scale=weight)
    norms = torch.diagonal(tt).clone()
    Mr = mass + ridge * torch.eye(K, dtype=mass.dtype, device=mass.device)
    coeff = torch.linalg.solve(Mr, loads)
    normal = torch.linalg.norm(Mr @ coeff - loads, dim=0) / torch.linalg.norm(loads, dim=0)
    gram_square = norms - 2 * torch.sum(coeff * loads, 0) + torch.sum(coeff * (mass @ coeff), 0)
    diff = basis[:, :K].to(torch.float64) @ coeff - targets.to(torch.float64)
    direct = weight * torch.sum(diff * diff, 0)
    return dict(loads=loads, target_norms=norms, coefficients=coeff, normal_residuals=normal,
                direct_relative_L2=torch.sqrt(direct / norms),
                Gram_identity_errors=torch.sqrt(torch.clamp(gram_square, min=0) / norms))
