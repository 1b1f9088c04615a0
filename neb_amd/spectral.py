"""Spectral objective, FP64 orthonormalisation and full-mass projection on
the GPU, plus the training-step driver.

FP64 K x K work (Cholesky, inverse, generalized eigh, mass solve) runs in
torch.linalg on the GPU (rocSOLVER / rocBLAS); every N-sized contraction
runs in libnebhip.so.

Reference surfaces mirrored (paths relative to the reference root;
MB = experiments/metaball/_backend/flow/cases/metaball):
  trace_objective        MB/velocity_trace_jet2.py:107-117
  grams / compute        MB/velocity/core.py:40-46, 79-105
  coefficient_peel       MB/velocity/core.py:48-69
  generalized eigh       experiments/periodic/paper_rollout.py:170-188
  project                experiments/metaball/_backend/exports/basis_projection/
                         postprocessing/paired_projection.py:141-177
"""
from __future__ import annotations

import math
import time
from contextlib import contextmanager

import torch

from . import network as nw


def _sym(A):
    return 0.5 * (A + A.T)


def _spd_inverse_async(M: torch.Tensor, out: torch.Tensor | None = None):
    """Enqueue the FP64 SPD inverse (into `out`, a K x K view with unit
    column stride, when given); returns (X, info) without synchronising
    (info[0] != 0 names the first nonpositive pivot, 1-based)."""
    K = M.shape[0]
    A = M.to(torch.float64).contiguous()
    X = torch.empty_like(A) if out is None else out
    info = torch.empty(1, dtype=torch.int32, device=A.device)
    need = nw.H.load().neb_spd_inverse_workspace_size(K)
    ws = nw.workspace(A.device).get(need)
    nw.H.call("neb_spd_inverse", K, A.data_ptr(), A.stride(0), X.data_ptr(), X.stride(0), info.data_ptr(),
              ws.data_ptr(), ws.numel(), nw._stream())
    return X, info


def _spd_inverse_chol_async(M: torch.Tensor):
    """Enqueue neb_spd_inverse_chol: returns (X, T, info) without
    synchronising, X = M^-1 and T = U^-1 for the upper Cholesky factor U of
    M (T^T M T = I, T upper triangular with a positive diagonal); info[0] != 0
    names the first nonpositive pivot (1-based), and X, T are then
    meaningless."""
    K = M.shape[0]
    A = M.to(torch.float64).contiguous()
    X = torch.empty_like(A)
    T = torch.empty_like(A)
    info = torch.empty(1, dtype=torch.int32, device=A.device)
    need = nw.H.load().neb_spd_inverse_workspace_size(K)
    ws = nw.workspace(A.device).get(need)
    nw.H.call("neb_spd_inverse_chol", K, A.data_ptr(), A.stride(0), X.data_ptr(), X.stride(0), T.data_ptr(),
              T.stride(0), info.data_ptr(), ws.data_ptr(), ws.numel(), nw._stream())
    return X, T, info


def spd_inverse_chol(M: torch.Tensor):
    """(M^-1, T) for a symmetric positive definite FP64 matrix on the GPU,
    T = U^-1 the inverse of its upper Cholesky factor (M = U^T U,
    T^T M T = I), both from the blocked Gauss-Jordan sweeps of libnebhip.
    Raises FloatingPointError when M is not positive definite (the
    cholesky gate of velocity/core.py:53-56)."""
    X, T, info = _spd_inverse_chol_async(M)
    bad = int(info.item())
    if bad != 0:
        _raise_not_pd(bad)
    return _sym(X), T


def _refined_inverse(Mr: torch.Tensor, out: torch.Tensor, steps: int):
    """Enqueue the sweep inverse of M_r followed by `steps` Newton-Schulz
    steps X <- X + X (I - M_r X) (two FP64 GEMMs each), written into `out`;
    returns info of the sweep (see _spd_inverse_async).

    The sweep inverse's forward error grows with cond(M_r) faster than a
    backward-stable solve's: at cond(M_r) = 1.3e6 (the K = 1024 objective with
    ridge 1e-4 mean diag M) it moves tr(X S) by 2e-9 relative, noise that a
    central difference of the loss divides by h.  One step squares the
    residual I - M_r X and leaves tr(X S) within 1e-12 of a Cholesky solve."""
    if steps <= 0:
        return _spd_inverse_async(Mr, out=out)[1]
    X, info = _spd_inverse_async(Mr)
    eye = _eye(Mr.shape[0], Mr.device)
    for s in range(steps):
        R = torch.addmm(eye, Mr, X, alpha=-1.0)
        if s == steps - 1:
            out.copy_(torch.addmm(X, X, R))
        else:
            X = torch.addmm(X, X, R)
    return info


def _raise_not_pd(bad: int):
    raise FloatingPointError(f"mass matrix is not positive definite (pivot {bad}); no update")


def spd_inverse(M: torch.Tensor) -> torch.Tensor:
    """FP64 inverse of a symmetric positive definite matrix on the GPU
    (libnebhip blocked Gauss-Jordan sweeps).  Raises FloatingPointError when
    a pivot is not positive, i.e. the matrix is not positive definite (the
    gate of MB/velocity_trace_jet2.py:109-111); no update is made then."""
    X, info = _spd_inverse_async(M)
    bad = int(info.item())
    if bad != 0:
        _raise_not_pd(bad)
    return _sym(X)


def _eye(K: int, device) -> torch.Tensor:
    key = (K, str(device))
    e = _EYES.get(key)
    if e is None:
        e = _EYES[key] = torch.eye(K, dtype=torch.float64, device=device)
    return e


_EYES: dict = {}


def _raise_gate(value: int):
    """FloatingPointError of a failed objective gate value (0 = passed,
    k > 0 = first nonpositive pivot k, -1 = condition limit / nonfinite)."""
    if value > 0:
        _raise_not_pd(value)
    if value != 0:
        raise FloatingPointError("mass condition estimate exceeds limit; no update")


def spectral_objective(M: torch.Tensor, S: torch.Tensor, lam: float = 1e-2, ridge: float = 0.0,
                       cond_limit: float = 1e12, gate: str = "eig", refine: int = 1):
    """L = tr((M + ridge I)^-1 S) + lam ||M - I||_F^2 and its closed-form
    adjoints dL/dM = -(M_r^-1 S M_r^-1) + 2 lam (M - I), dL/dS = M_r^-1.

    The gate follows MB/velocity_trace_jet2.py:109-111: a nonfinite,
    nonpositive or too ill-conditioned mass raises FloatingPointError before
    anything is updated.  gate="eig" checks eigenvalues (the reference's
    rule); gate="cholesky" only requires positive definiteness (every pivot
    of the FP64 inversion positive, as a Cholesky factorization requires) and
    a finite result.  The inverse is the sweep inverse refined by `refine`
    Newton-Schulz steps (_refined_inverse).  Returns (loss, dM, dS) with FP64
    tensors on M's device (loss 0-dim)."""
    Mr = nw.sym_shift(M, ridge)
    if gate == "eig":
        e = torch.linalg.eigvalsh(Mr)
        ok = bool(torch.isfinite(e).all()) and float(e[-1]) > 0 and float(e[0] / e[-1]) >= 1.0 / cond_limit
        if not ok:
            raise FloatingPointError("Physical mass zero/rank deficient/condition exceeds limit; no update")
    K = M.shape[0]
    pair = nw.adjoint_pair(K, M.device)      # [dL/dM | dL/dS]
    X = pair[:, K:]
    info = _refined_inverse(Mr, X, refine)
    loss, X, Ss, g = nw.spectral_terms(M, S, X, info, lam, ridge, math.inf)
    _raise_gate(int(g.item()))
    dM = nw.mass_adjoint(X, Ss, M, lam, out=pair[:, :K])
    return loss, dM, X


class SpectralSolver:
    """Stateful FP64 evaluator of the spectral objective for a training loop.

    Step 1 (or whenever the check below fails) inverts M_r = M + ridge I
    directly (FP64 sweeps, positive-definiteness gate, then `refine`
    Newton-Schulz steps), as spectral_objective does.  Later steps start from the
    previous step's inverse X and refine it with Newton-Schulz iterations
    X <- X (2I - M_r X) (two FP64 GEMMs each, quadratic convergence), which
    is valid while the residual ||I - M_r X||_F < 1; the iteration stops at a
    relative residual <= tol, else the direct inverse runs and the next
    `backoff` steps go straight to it (a rank-deficient mass held up by a
    ridge changes too much per step for the warm start to converge).  Either way the
    returned inverse satisfies ||I - M_r X||_F <= tol, and a condition
    estimate ||M_r||_F ||X||_F above cond_limit raises FloatingPointError
    (the gate of MB/velocity_trace_jet2.py:109-111).

    deferred=True keeps both gates on the device: the call queues its work
    without a host synchronisation and leaves `gate` (one-element int32:
    0 = passed, k > 0 = first nonpositive pivot k, -1 = condition limit or
    nonfinite); the caller hands it to EigenfluidNet.adam(gate=...,
    on_gate=solver.raise_gate), which skips the update on the device and
    raises, or calls check() before using the results."""

    def __init__(self, lam: float = 1e-2, ridge: float = 0.0, cond_limit: float = 1e12, tol: float = 1e-10,
                 max_iter: int = 6, backoff: int = 16, deferred: bool = False, refine: int = 1):
        self.lam, self.ridge, self.cond_limit, self.tol, self.max_iter = lam, ridge, cond_limit, tol, max_iter
        self.refine = refine
        self.backoff = backoff
        self.deferred = deferred
        self.X = None
        self.gate = None
        self.last_path = None
        self._skip = 0       # warm starts left to skip after one that failed

    def _residual(self, Mr, X, eye):
        return float(torch.linalg.norm(eye - Mr @ X))

    def raise_gate(self, value: int):
        """Raise the FloatingPointError of a failed gate value (see `gate`)."""
        _raise_gate(value)

    def check(self):
        """Synchronise and raise if the last deferred call's gate failed."""
        if self.gate is not None:
            self.raise_gate(int(self.gate.item()))

    def __call__(self, M: torch.Tensor, S: torch.Tensor):
        K = M.shape[0]
        Mr = nw.sym_shift(M, self.ridge)
        X = None
        if self._skip > 0:
            self._skip -= 1
        elif self.X is not None and self.X.shape == Mr.shape:
            eye = _eye(K, M.device)
            X = self.X
            ok = False
            for _ in range(self.max_iter):
                R = eye - Mr @ X
                r = float(torch.linalg.norm(R))
                if not r < 0.5:          # also catches NaN
                    break
                if r <= self.tol:
                    ok = True
                    break
                X = X + X @ R            # = X (2I - Mr X)
            if ok:
                self.last_path = "newton-schulz"
            else:
                X = None
                self._skip = self.backoff
        info = None
        pair = nw.adjoint_pair(K, M.device)  # [dL/dM | dL/dS] side by side for the reverse
        if X is None:
            X = pair[:, K:]
            info = _refined_inverse(Mr, X, self.refine)
            self.last_path = "spd-inverse"
        else:
            X = pair[:, K:].copy_(X)
        # X <- sym(X), loss, and the positive-definiteness / condition gates
        # on the device (one fused launch + one reduction block)
        loss, X, Ss, gate = nw.spectral_terms(M, S, X, info, self.lam, self.ridge, self.cond_limit)
        if self.deferred:
            self.gate = gate
        else:
            self.raise_gate(int(gate.item()))  # one host synchronisation for both gates
        self.X = X
        dM = nw.mass_adjoint(X, Ss, M, self.lam, out=pair[:, :K])
        return loss, dM, X


PEEL_GATE = -2      # gate value of a failed coefficient-peel gate


def _raise_peel_gate(value: int):
    if value == PEEL_GATE:
        raise FloatingPointError("Velocity rank/relative mass/orthogonality gate failed before Adam; no update")
    _raise_gate(value)


def coefficient_peel(M: torch.Tensor, S: torch.Tensor, ortho_tol: float = 1e-5, rel_min: float = 1e-4,
                     deferred: bool = False, order: str = "rayleigh", mass_floor: float = 0.0) -> dict:
    """Sorted Cholesky coefficient peel (MB/velocity/core.py:48-69) in FP64
    on the GPU.

    The modes are sorted by their raw Rayleigh quotient diag(S)/diag(M)
    (stable order); T is the inverse upper Cholesky factor of the
    diagonally scaled sorted mass, refined by a second Cholesky pass, so
    that T^T M_sorted T = I (both factors from the FP64 sweeps of
    neb_spd_inverse_chol); loss_r = (T^T S_sorted T)_rr.  Gates (core.py:
    57-61): every loss finite and positive, relative mass
    (1/T_rr)^2 / M_rr >= rel_min, max |T^T M T - I| <= ortho_tol and both
    factorizations positive definite.  They are evaluated on the device
    into `gate` (one int32: 0 = passed, PEEL_GATE = failed); deferred=False
    synchronises and raises FloatingPointError, deferred=True leaves the
    check to EigenfluidNet.adam(gate=..., on_gate=raise_peel_gate).

    order="fixed" keeps the network order (HB2P:2999-3004 "fixed_order");
    mass_floor adds the 2D peel's gates: every raw and residual mass above
    it (HB2P:3137-3153).

    Adjoint (core.py:64-67): with every mode's earlier peeled modes held
    constant, the field cotangents of sum_r loss_r are G_U = w U Cu and
    G_D = w D Cd for M = w U^T U, S = w D^T D.  `pair` holds
    [Cu^T / 2 | Cd^T / 2], the form backward_fields(symmetric=True)
    consumes (it applies field cotangents 2w Phi X^T).

    Returns dict(losses (sorted order), loss (their sum, 0-dim), order,
    T (rows in the original mode order, columns in sorted order), Cu, Cd,
    relative_mass, gram_max, raw, gate, pair)."""
    K = M.shape[0]
    dev = M.device
    M = M.to(torch.float64)
    S = S.to(torch.float64)
    eye = _eye(K, dev)
    raw = torch.diagonal(S) / torch.diagonal(M)
    if order == "rayleigh":
        order = torch.argsort(raw, stable=True)
    elif order == "fixed":
        order = torch.arange(K, device=dev)
    else:
        raise ValueError(f"unknown peel order {order!r}")
    ms = M[order][:, order]
    ss = S[order][:, order]
    scale = torch.diagonal(ms).sqrt()
    # inverse upper Cholesky factors of the scaled sorted mass and of the
    # first pass's Gram (libnebhip sweeps; info: positive definiteness)
    _, T1, info1 = _spd_inverse_chol_async(ms / scale[:, None] / scale[None, :])
    T = T1 / scale[:, None]
    H2 = T.mT @ ms @ T
    _, T2, info2 = _spd_inverse_chol_async(0.5 * (H2 + H2.mT))
    T = T @ T2
    error = (T.mT @ ms @ T - eye).abs().max()
    norm = 1.0 / torch.diagonal(T)
    relative = norm * norm / torch.diagonal(ms)
    F = T.mT @ ss @ T
    losses = torch.diagonal(F).clone()
    ok = (torch.isfinite(losses).all() & (losses > 0).all() & (relative >= rel_min).all()
          & (error <= ortho_tol) & (info1[0] == 0) & (info2[0] == 0) & (norm * norm > mass_floor).all()
          & (torch.diagonal(M) > mass_floor).all() & torch.isfinite(raw).all())
    gate = torch.where(ok, 0, PEEL_GATE).to(torch.int32).reshape(1)
    C = -2.0 * (T * losses[None, :] + T @ torch.triu(F, 1)) / norm[None, :]
    B = 2.0 * T / norm[None, :]
    rows, cols = order[:, None], order[None, :]
    Cu = torch.empty_like(C)
    Cd = torch.empty_like(B)
    Tr = torch.empty_like(T)
    Cu[rows, cols] = C
    Cd[rows, cols] = B
    Tr[order] = T
    pair = nw.adjoint_pair(K, dev)
    pair[:, :K] = 0.5 * Cu.mT
    pair[:, K:] = 0.5 * Cd.mT
    if not deferred:
        _raise_peel_gate(int(gate.item()))
    return dict(losses=losses, loss=losses.sum(), order=order, T=Tr, Cu=Cu, Cd=Cd, relative_mass=relative,
                gram_max=error, raw=raw, gate=gate, pair=pair)


class PeelSolver:
    """The coefficient peel as a step objective: solver(M, S) returns
    (loss, dM, dS) like SpectralSolver, where dM, dS are the halves of the
    peel's cotangent pair (pass them to backward_fields with
    symmetric=True).  deferred=True keeps the gate on the device (`gate`,
    raised by EigenfluidNet.adam through raise_gate); `last` is the peel's
    full result dict."""

    def __init__(self, ortho_tol: float = 1e-5, rel_min: float = 1e-4, deferred: bool = False):
        self.ortho_tol, self.rel_min, self.deferred = ortho_tol, rel_min, deferred
        self.gate = None
        self.last = None
        self.last_path = "coefficient-peel"

    def raise_gate(self, value: int):
        _raise_peel_gate(value)

    def check(self):
        if self.gate is not None:
            self.raise_gate(int(self.gate.item()))

    def __call__(self, M: torch.Tensor, S: torch.Tensor):
        p = coefficient_peel(M, S, self.ortho_tol, self.rel_min, deferred=self.deferred)
        self.gate = p["gate"]
        self.last = p
        K = M.shape[0]
        return p["loss"], p["pair"][:, :K], p["pair"][:, K:]


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
            gate: str = "eig", timings: dict | None = None, objective: str = "trace",
            reverse_prec: str | None = None, peel_options: dict | None = None, grams: str = "fields"):
    """One spectral step (the analogue of MB/velocity/core.py:79-105
    ``compute``): jets -> lifted fields -> FP64 Grams -> spectral objective ->
    fused reverse.  objective="trace" is tr((M + ridge I)^-1 S) +
    lam ||M - I||^2 (MB/velocity_trace_jet2.py:107-117), objective="peel" the
    sorted Cholesky coefficient peel (core.py:48-69; its result dict is
    returned as "peel"; dM, dS are the field-Gram cotangents the reverse
    applies).  reverse_prec (default: prec) runs the reverse pass
    in another operand precision on the same fields and a converted tape:
    prec="f32", gram_mode="bf16x3", reverse_prec="bf16" keeps the fields and
    Grams at FP32 accuracy and runs the reverse on BF16 MFMA; peel_options
    are keyword arguments of coefficient_peel.  grams="factored" (split-
    precision tapes, every lift but the streamfunction one) forms no fields:
    the Grams come from the lift-weighted channel Grams of the last hidden
    jets (nw.factored_field_grams) and the reverse is the factored one on the
    same stacked channels ("fields" is then None).  Returns dict(loss, M, S,
    grads, fields)."""
    if objective not in ("trace", "peel"):
        raise ValueError(f"unknown objective {objective!r}")
    if grams not in ("fields", "factored"):
        raise ValueError(f"unknown grams {grams!r}")
    p = nw.PRECISIONS[prec]
    tt = {} if timings is None else timings
    K = net.cfg.modes
    split = nw.split_fields_for(prec, gram_mode)
    factored = grams == "factored"
    with timed(tt, "neural_forward_analytic_derivatives"):
        tape = nw.trunk_forward(net, points, p)
        if factored:
            if reverse_prec is not None or not nw.factored_grams_ok(net, tape):
                raise ValueError("factored Grams need split-precision tapes of a factored lift (no reverse_prec)")
            fields = None
        else:
            fields = nw.output_fields(net, tape, split=split)
    with timed(tt, "gram_assembly"):
        if factored:
            prod = nw.factored_products(net, tape)
            M, S = nw.factored_field_grams(net, prod.Gh, weight)
        else:
            M, S = nw.field_grams(fields, weight, gram_mode)
    rtape = tape if reverse_prec is None else nw.convert_tape(tape, nw.PRECISIONS[reverse_prec])
    # the output-layer products need only the fields and the tape: they run
    # on the side stream while the FP64 objective holds this one
    if not factored:
        prod = nw.output_products(net, rtape, fields) if training else None
    with timed(tt, "spectral_objective"):
        if objective == "peel":
            peel = coefficient_peel(M, S, **(peel_options or {}))
            loss, dM, dS = peel["loss"], peel["pair"][:, :K], peel["pair"][:, K:]
        else:
            loss, dM, dS = spectral_objective(M, S, lam, ridge, gate=gate)
    out = dict(loss=loss, M=M, S=S, dM=dM, dS=dS, fields=fields, timings=tt)
    if objective == "peel":
        out["peel"] = peel
    if training:
        with timed(tt, "analytic_backward"):
            out["grads"] = nw.backward_fields(net, rtape, fields, dM, dS, weight, symmetric=True, products=prod)
    return out


def velocity_peel_step(net: nw.EigenfluidNet, points: torch.Tensor, weight: float, prec: str = "bf16x3",
                       training: bool = True, causal_sorting: str = "geometry_rayleigh",
                       relative_mass_floor: float = 1e-4, overlap_tolerance: float = 1e-3,
                       mass_floor: float = 1e-12, lr: float | None = None, clip: float = 8.0) -> dict:
    """The 2D causal velocity Rayleigh peel step (HB2P:2935-3380
    ``_velocity_rayleigh_peel_step``) on the shared trunk.

    Velocity u = (psi_y, -psi_x); mass M = w sum |grad psi|^2 and stiffness
    S = w sum (psi_xx^2 + 2 psi_xy^2 + psi_yy^2) as Grams of the lifted
    fields (NetConfig(stiffness="hessian")); the two-pass modified
    Gram-Schmidt peel in the velocity mass metric, every pass's coefficients
    applied to the Hessian rows too, is the sorted Cholesky coefficient peel
    of (M, S) (coefficient_peel; its equality with the literal two-pass peel
    is pinned by the oracle), with the candidates sorted by raw S/M
    ("geometry_rayleigh") or in network order ("fixed_order").  Gates (all on
    the device, a failed one raises FloatingPointError before any update):
    raw and residual masses above mass_floor, relative residual mass >=
    relative_mass_floor, post-peel overlap <= overlap_tolerance, finite
    values.  The gradient is that of the sum of the Rayleigh quotients with
    the earlier (detached) modes held constant; with lr given, one gated Adam
    update follows.  Returns the reference's diagnostics: loss = mean
    Rayleigh quotient, mode_energies (causal order), mode_energies_by_network,
    causal_order, min_relative_residual_mass, max_post_peel_overlap,
    orth_error, peeled_rank, plus grads (training) and the Grams."""
    if net.cfg.dim != 2 or net.cfg.stiffness != "hessian":
        raise ValueError("the velocity Rayleigh peel needs a 2D net with NetConfig(stiffness='hessian')")
    if causal_sorting not in ("fixed_order", "geometry_rayleigh"):
        raise ValueError(f"unsupported causal_sorting {causal_sorting!r}")
    if not 0.0 < relative_mass_floor < 1.0 or not 0.0 < overlap_tolerance < 1.0:
        raise ValueError("relative_mass_floor and overlap_tolerance must be in (0, 1)")
    if points.shape[0] == 0:
        raise ValueError("velocity Rayleigh received no quadrature points")
    opts = dict(ortho_tol=overlap_tolerance, rel_min=relative_mass_floor, mass_floor=mass_floor,
                order="rayleigh" if causal_sorting == "geometry_rayleigh" else "fixed", deferred=True)
    res = compute(net, points, weight, prec=prec, gram_mode="f64" if prec == "f64" else "bf16x3",
                  training=training, objective="peel", peel_options=opts)
    pk = res["peel"]
    K = net.cfg.modes
    if training and lr is not None:
        net.adam(res["grads"], lr, clip=clip, gate=pk["gate"], on_gate=_raise_peel_gate)
    else:
        _raise_peel_gate(int(pk["gate"].item()))
    ray = pk["losses"]
    by_network = torch.empty_like(ray)
    by_network[pk["order"]] = ray
    out = dict(loss=float(ray.mean()), velocity_rayleigh=float(ray.mean()), mode_energies=ray,
               mode_energies_by_network=by_network, causal_order=pk["order"], causal_sorting=causal_sorting,
               min_relative_residual_mass=float(pk["relative_mass"].min()),
               max_post_peel_overlap=float(pk["gram_max"]), orth_error=float(pk["gram_max"]), peeled_rank=K,
               requested_modes=K, peeling_passes_used=2, M=res["M"], S=res["S"], peel=pk)
    if training:
        out["grads"] = res["grads"]
    return out


def streamfunction_peel_step(net: nw.EigenfluidNet, points: torch.Tensor, weight: float, prec: str = "bf16x3",
                             training: bool = True, causal_sorting: str = "geometry_rayleigh",
                             mass_floor: float = 1e-12, lr: float | None = None, clip: float = 8.0) -> dict:
    """The 2D Dirichlet streamfunction Rayleigh peel step (HB2P:3383-3770
    ``_streamfunction_dirichlet_rayleigh_peel_step``, periodic trial space)
    on the shared trunk.

    Jet1 fields of the streamfunction lift (NetConfig(stiffness="stream")):
    psi rows centered over the points (the periodic gauge, HB2P:3475-3479)
    and the gradient rows (psi_x, psi_y); mass M = w sum psi_c^2, stiffness
    S = w sum |grad psi|^2 as Grams of those rows.  The two-pass modified
    Gram-Schmidt peel of the value columns in the w-weighted L2 metric, every
    pass's coefficients applied to the gradient rows too (HB2P:3496-3560), is
    the sorted Cholesky coefficient peel of (M, S) (coefficient_peel; its
    equality with the literal peel is pinned by the oracle), with the
    candidates sorted by raw S/M ("geometry_rayleigh") or in network order
    ("fixed_order").  Gates, on the device (a failed one raises
    FloatingPointError before any update, HB2P:3650-3684): raw and residual
    masses above mass_floor and every quotient finite.  The gradient is that
    of the sum of the Rayleigh quotients with the earlier (detached) modes
    held constant, through the centering (HB2P:3579-3583); with lr given,
    one gated Adam update follows.  Returns the reference's diagnostics
    (HB2P:3719-3770): loss = mean Rayleigh quotient, mode_energies (causal
    order), mode_energies_by_network, causal_order (0- and 1-based),
    sorting_rayleigh_by_network, mean_projection_ratio, orth_error,
    peeled_rank, plus grads (training) and the Grams."""
    if net.cfg.dim != 2 or net.cfg.stiffness != "stream":
        raise ValueError("the streamfunction Rayleigh peel needs a 2D net with NetConfig(stiffness='stream')")
    if causal_sorting not in ("fixed_order", "geometry_rayleigh"):
        raise ValueError(f"unsupported causal_sorting {causal_sorting!r}")
    if points.shape[0] == 0:
        raise ValueError("Dirichlet streamfunction Rayleigh received no points")
    opts = dict(ortho_tol=math.inf, rel_min=0.0, mass_floor=mass_floor,
                order="rayleigh" if causal_sorting == "geometry_rayleigh" else "fixed", deferred=True)
    res = compute(net, points, weight, prec=prec, gram_mode="f64" if prec == "f64" else "bf16x3",
                  training=training, objective="peel", peel_options=opts)
    pk = res["peel"]
    K = net.cfg.modes
    if training and lr is not None:
        net.adam(res["grads"], lr, clip=clip, gate=pk["gate"], on_gate=_raise_stream_gate)
    else:
        _raise_stream_gate(int(pk["gate"].item()))
    ray = pk["losses"]
    by_network = torch.empty_like(ray)
    by_network[pk["order"]] = ray
    # peeled mass norm_r^2 = relative_r * M_rr (sorted); projection ratio
    # |coefficients|^2 / peeled mass = 1 / relative - 1 (HB2P:3546-3550)
    proj = 1.0 / pk["relative_mass"] - 1.0
    proj[0] = 0.0
    out = dict(loss=float(ray.mean()), base_loss=float(ray.mean()), streamfunction_rayleigh=float(ray.mean()),
               mean_energy=float(ray.mean()), max_energy=float(ray.max()), min_energy=float(ray.min()),
               mean_projection_ratio=float(proj.mean()), orth_error=float(pk["gram_max"]), peeled_rank=K,
               requested_modes=K, mode_energies=ray, mode_energies_by_network=by_network,
               causal_sorting=causal_sorting, causal_order=pk["order"], causal_order_1based=pk["order"] + 1,
               sorting_rayleigh_by_network=pk["raw"], spatial_jet_order_used=1, M=res["M"], S=res["S"], peel=pk)
    if training:
        out["grads"] = res["grads"]
    return out


def _raise_stream_gate(value: int):
    if value == PEEL_GATE:
        raise FloatingPointError("streamfunction candidate mass/finiteness gate failed before Adam; no update")
    _raise_gate(value)


# Bytes of chunk tapes a training step may keep between its two passes, and
# the device memory left free beside them for one chunk's fields and reverse
# workspace: 13 split-precision chunks of 131,072 3D points (4 x 256 tanh
# trunk, 10 jet channels) fit a 288 GB MI355X.
CHUNK_TAPE_BUDGET = 192 << 30
CHUNK_TAPE_HEADROOM = 32 << 30


def _device_bytes_available(dev: torch.device) -> int:
    """Free device memory plus what torch's caching allocator holds unused."""
    free, _ = torch.cuda.mem_get_info(dev)
    return int(free + torch.cuda.memory_reserved(dev) - torch.cuda.memory_allocated(dev))


def chunked_step(net: nw.EigenfluidNet, points: torch.Tensor, weight: float, prec: str = "bf16",
                 gram_mode: str = "bf16x3", chunk: int = 131072, solver=None, training: bool = True,
                 timings: dict | None = None, tape_budget: int | None = None, grams: str = "fields"):
    """Spectral step over point chunks, memory-bounded for large point sets
    (the structure of MB/velocity/core.py:79-105: fields -> grams ->
    objective -> recompute + reverse).  Pass 1 evaluates each chunk's jets
    and fields and accumulates the FP64 Grams; pass 2 accumulates the
    parameter gradient chunk by chunk.  Every chunk's jet tape is kept from
    pass 1 for pass 2 when all of them fit `tape_budget` bytes (default
    CHUNK_TAPE_BUDGET, capped at the available device memory less
    CHUNK_TAPE_HEADROOM); otherwise pass 2 recomputes each chunk's tape
    (the reference's per-chunk recomputation, VCORE:92-102: no tape outlives
    its chunk).  Returns dict(loss, M, S, grads, tapes_kept, tape_bytes), the
    last the bytes all chunks' tapes together take.

    grams="factored" (split-precision tapes of a factored lift, the 3D curl
    lift): pass 1 forms no fields; it stacks each chunk's last hidden jet
    channels and accumulates their lift-weighted channel Grams, and the field
    Grams come from those once (nw.factored_field_grams: [3 width]^2
    products instead of [n comp, K] field matrices); pass 2 stacks the
    channels for the factored reverse without forming their Grams again and
    adds the output-weight gradient of the accumulated ones once."""
    p = nw.PRECISIONS[prec]
    tt = {} if timings is None else timings
    K = net.cfg.modes
    n = points.shape[0]
    solver = solver if solver is not None else SpectralSolver()
    split = nw.split_fields_for(prec, gram_mode)
    M = torch.zeros((K, K), dtype=torch.float64, device=net.device)
    S = torch.zeros_like(M)
    starts = list(range(0, n, chunk))
    budget = CHUNK_TAPE_BUDGET if tape_budget is None else int(tape_budget)
    if training and budget > 0 and net.device.type == "cuda":
        budget = min(budget, _device_bytes_available(net.device) - CHUNK_TAPE_HEADROOM)
    tapes = [] if training and budget > 0 else None
    kept_bytes = 0
    if grams not in ("fields", "factored"):
        raise ValueError(f"unknown grams {grams!r}")
    factored = grams == "factored"
    Gh = None
    with timed(tt, "pass1_forward_and_grams"):
        for i, c0 in enumerate(starts):
            tape = nw.trunk_forward(net, points[c0:c0 + chunk], p)
            if factored:
                if not nw.factored_grams_ok(net, tape):
                    raise ValueError("factored Grams need split-precision tapes of a factored lift")
                fp = nw.factored_products(net, tape)
                if Gh is None:
                    Gh = fp.Gh
                else:
                    for acc, g in zip(Gh, fp.Gh):
                        acc.add_(g)
                del fp
            else:
                f = nw.output_fields(net, tape, split=split)
            if i == 0:
                kept_bytes = nw.tape_bytes(tape) * len(starts)
            if tapes is not None:
                if kept_bytes > budget:
                    tapes = None
                else:
                    tapes.append(tape)
            del tape
            if not factored:
                nw.field_grams(f, weight, gram_mode, M=M, S=S, beta=1.0)
                del f
        if factored:
            nw.factored_field_grams(net, Gh, weight, M=M, S=S)
    with timed(tt, "spectral_objective"):
        loss, dM, dS = solver(M, S)
    out = dict(loss=loss, M=M, S=S, timings=tt, tapes_kept=tapes is not None, tape_bytes=kept_bytes)
    if training:
        grads = torch.zeros_like(net.params)
        with timed(tt, "pass2_recompute_and_reverse"):
            for i, c0 in enumerate(starts):
                if tapes is not None:
                    tape, tapes[i] = tapes[i], None     # released once its reverse is queued
                else:
                    tape = nw.trunk_forward(net, points[c0:c0 + chunk], p)
                # the factored reverse of a split-precision tape reads no fields
                f = None if factored or nw.factored_reverse_ok(net, tape) else nw.output_fields(net, tape, split=split)
                prod = nw.factored_products(net, tape, grams=False) if factored else None
                grads += nw.backward_fields(net, tape, f, dM, dS, weight, symmetric=True, products=prod)
                del tape, f, prod
            if factored:   # the output-weight gradient of all chunks' channel Grams
                gW, _ = net.split(grads)
                pair = nw._as_pair(dM, dS)
                if pair is None:
                    pair = torch.cat((dM, dS), 1)
                gW[-1] += nw._factored_output_grad(net, nw.FactoredProducts(None, None, None, Gh), pair,
                                                   weight).to(grads.dtype)
        out["grads"] = grads
    return out


def basis_stream(net: nw.EigenfluidNet, points: torch.Tensor, prec: str = "bf16", chunk: int = 131072,
                 split: bool = False):
    """Yield (first point, Fields) chunk by chunk: the full-grid basis
    evaluation of experiments/metaball/evaluate.py:130-190 without the disk
    round trip.  The split-precision forward (prec="bf16x3") always yields
    split fields."""
    p = nw.PRECISIONS[prec]
    split = split or prec == "bf16x3"
    for c0 in range(0, points.shape[0], chunk):
        tape = nw.trunk_forward(net, points[c0:c0 + chunk], p)
        f = nw.output_fields(net, tape, split=split)
        del tape
        yield c0, f


def fullmass(net: nw.EigenfluidNet, points: torch.Tensor, weight: float, prec: str = "bf16",
             gram_mode: str = "bf16x3", chunk: int = 131072, rank_tol: float = 1e-12):
    """Full-point FP64 mass Gram with the rank gate of
    experiments/metaball/run.py:99-116 (ValueError when rank deficient)."""
    K = net.cfg.modes
    M = torch.zeros((K, K), dtype=torch.float64, device=net.device)
    for _, f in basis_stream(net, points, prec, chunk, split=nw.split_fields_for(prec, gram_mode)):
        nw.field_gram(f, "mass", weight, gram_mode, out=M, beta=1.0)
    ev = torch.linalg.eigvalsh(_sym(M))
    if float(ev[0]) <= rank_tol * float(ev[-1]):
        raise ValueError("Fullmass basis is rank deficient")
    return M, ev


def fullmass_fp64(net: nw.EigenfluidNet, points: torch.Tensor, weight: float, prec: str = "bf16",
                  chunk: int = 131072) -> torch.Tensor:
    """FP64 mass Gram of the basis that project_streamed uses for `prec`
    (the fp32 field matrix widened to FP64 on load; the FP64 fields in the
    FP64 mode), accumulated chunk by chunk: the mass of the projection's Gram
    identity (PP:141-177)."""
    K = net.cfg.modes
    M = torch.zeros((K, K), dtype=torch.float64, device=net.device)
    for _, f in basis_stream(net, points, prec, chunk):
        B = nw.field_matrix(f, "mass", torch.float64 if prec == "f64" else torch.float32)
        nw.gram(B, K, weight, "f64", out=M, beta=1.0)
        del f, B
    return M


def project_streamed(net: nw.EigenfluidNet, points: torch.Tensor, targets: torch.Tensor, mass: torch.Tensor,
                     weight: float, prec: str = "bf16", chunk: int = 131072, ridge: float = 0.0,
                     normal_tol: float = 1e-8, identity_tol: float = 1e-8):
    """Two-pass full-mass projection of targets onto the network basis
    (PP:141-177): pass 1 streams the basis chunk by chunk and accumulates
    loads = w B^T T and the target norms (FP64 cross kernel), the FP64 solve
    (with optional ridge) is gated on the normal residual, pass 2 re-streams
    the basis for the direct residual sum w |B c - T|^2, which must agree with
    the Gram identity.  targets: [n*comp, m] point-major rows, FP64.  The
    basis is the fp32 field matrix of the `prec` forward (split-precision
    fields joined as hi + lo, FP64 fields in the FP64 mode); `mass` must be
    the FP64 mass of that same basis (fullmass_fp64)."""
    K = net.cfg.modes
    m = targets.shape[1]
    comp = 2 if net.cfg.dim == 2 else 3
    loads = torch.zeros((K, m), dtype=torch.float64, device=net.device)
    norms2 = torch.zeros((m, m), dtype=torch.float64, device=net.device)
    for c0, f in basis_stream(net, points, prec, chunk):
        B = nw.field_matrix(f, "mass", torch.float64 if prec == "f64" else torch.float32)
        rows = B.shape[0]
        T = targets[comp * c0: comp * c0 + rows]
        # fp32 basis widened to fp64 on load (fp64 basis in the FP64 mode)
        nw.cross([(B, T, rows, 1.0)], K, m, nw.GRAM_MODES["f64"], scale=weight, out=loads, beta=1.0)
        nw.cross([(T, T, rows, 1.0)], m, m, nw.GRAM_MODES["f64"], scale=weight, out=norms2, beta=1.0)
        del f, B
    norms = torch.diagonal(norms2).clone()
    Mr = _sym(mass) + ridge * torch.eye(K, dtype=torch.float64, device=mass.device)
    coeff = torch.linalg.solve(Mr, loads)
    normal = torch.linalg.norm(Mr @ coeff - loads, dim=0) / torch.linalg.norm(loads, dim=0)
    if not bool(torch.isfinite(coeff).all()) or not bool((normal < normal_tol).all()):
        raise ValueError("fullmass normal-residual gate failed")
    gram_square = norms - 2 * torch.sum(coeff * loads, 0) + torch.sum(coeff * (mass @ coeff), 0)
    direct = torch.zeros(m, dtype=torch.float64, device=net.device)
    for c0, f in basis_stream(net, points, prec, chunk):
        B = nw.field_matrix(f, "mass", torch.float64)
        rows = B.shape[0]
        diff = B @ coeff - targets[comp * c0: comp * c0 + rows]
        direct += weight * torch.sum(diff * diff, 0)
        del f, diff, B
    errors = torch.sqrt(direct / norms)
    identity = torch.sqrt(torch.clamp(gram_square, min=0) / norms)
    if not bool((torch.abs(errors - identity) < identity_tol).all()):
        raise ValueError("Gram/direct identity tolerance failed")
    return dict(loads=loads, target_norms=norms, coefficients=coeff, normal_residuals=normal,
                direct_relative_L2=errors, Gram_identity_errors=identity)


class JetGram(torch.autograd.Function):
    """(params, points) -> (M, S): analytic-jet forward, lift and FP64 Grams;
    the backward is the handwritten reverse pass (no autograd tape through
    the network).  ``params`` must be ``net.params``."""

    @staticmethod
    def forward(ctx, params, points, net, weight, prec, gram_mode):
        p = nw.PRECISIONS[prec]
        tape = nw.trunk_forward(net, points, p)
        fields = nw.output_fields(net, tape, split=nw.split_fields_for(prec, gram_mode))
        M, S = nw.field_grams(fields, weight, gram_mode)
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
