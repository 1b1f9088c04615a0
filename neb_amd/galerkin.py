"""Galerkin reduced-order operators on the network basis: the consumers of
the basis after training.

Reference surfaces (paths relative to the reference root):
  PeriodicROM            experiments/periodic/paper_rollout.py:37-66 (rollout: rhs
                         and ETDRK4), :166-186 (velocity normalisation, full mass,
                         rank gate, generalized eigh, transformed fields)
  derive_fields          experiments/metaball/operators.py:20-60 (Spectral),
                         :137-177 (derive)
  assemble_operators     experiments/metaball/operators.py:62-72 (add),
                         :199-257 (assemble: M, A, packed Q)

N-sized contractions run in libnebhip.so: the rollout right-hand side
(neb_galerkin_rhs2d), the Grams and the cross products (with row weights for
the convective tensor).  The K x K FP64 algebra and the basis transform
([n, K] x [K, K]) use rocBLAS / rocSOLVER through torch; the spectral
derivation on the periodic grid uses rocFFT through torch.fft.
"""
from __future__ import annotations

import math

import torch

from . import _hip as H
from . import network as nw
from . import spectral as sp

_DT = {torch.float32: H.F32, torch.float64: H.F64}


def rollout_rhs(gu: torch.Tensor, gv: torch.Tensor, gw: torch.Tensor, a: torch.Tensor, weight: float,
                out: torch.Tensor | None = None) -> torch.Tensor:
    """rhs = gu^T (w omega v) - gv^T (w omega u) with u = gu a, v = gv a,
    omega = gw a (paper_rollout.py:37-41); fields [n, K] row-major with one
    row stride, fp32 or fp64; a and the result FP64 [K]."""
    n, K = gu.shape
    if gv.shape != gu.shape or gw.shape != gu.shape or gu.stride() != gv.stride() or gu.stride() != gw.stride():
        raise ValueError("gu, gv, gw must share shape and strides")
    if gu.stride(1) != 1 or gu.dtype != gv.dtype or gu.dtype != gw.dtype:
        raise ValueError("fields must be row-major matrices of one dtype")
    a = a.to(torch.float64).contiguous()
    out = torch.empty(K, dtype=torch.float64, device=gu.device) if out is None else out
    ws = nw.workspace(gu.device).get(H.load().neb_galerkin_workspace_size(n, K))
    H.call("neb_galerkin_rhs2d", _DT[gu.dtype], gu.data_ptr(), gv.data_ptr(), gw.data_ptr(), n, K, gu.stride(0),
           a.data_ptr(), float(weight), out.data_ptr(), ws.data_ptr(), ws.numel(), nw._stream())
    return out


def etdrk4_coefficients(linear: torch.Tensor, h: float, roots: int = 32):
    """ETDRK4 coefficients of the diagonal linear part at step h by the
    contour mean over `roots` points (Kassam & Trefethen 2005;
    paper_rollout.py:49-56)."""
    lin = linear.to(torch.float64)
    e = torch.exp(h * lin)
    e2 = torch.exp(0.5 * h * lin)
    r = torch.exp(1j * math.pi * (torch.arange(1, roots + 1, device=lin.device, dtype=torch.float64) - 0.5) / roots)
    lr = h * lin[:, None] + r
    q = h * torch.real(torch.mean((torch.exp(lr / 2) - 1) / lr, -1))
    f1 = h * torch.real(torch.mean((-4 - lr + torch.exp(lr) * (4 - 3 * lr + lr * lr)) / lr ** 3, -1))
    f2 = h * torch.real(torch.mean((2 + lr + torch.exp(lr) * (-2 + lr)) / lr ** 3, -1))
    f3 = h * torch.real(torch.mean((-4 - 3 * lr - lr * lr + torch.exp(lr) * (4 - lr)) / lr ** 3, -1))
    return e, e2, q, f1, f2, f3


class PeriodicROM:
    """Frozen-basis Galerkin model of 2D periodic flow on the network's
    velocity / vorticity fields (paper_rollout.py:166-186): columns
    normalised to unit velocity mass, FP64 mass w(gu^T gu + gv^T gv) and
    stiffness w gw^T gw, the full-rank gate (ArithmeticError), the
    generalized eigenbasis (vals, T), fields transformed by T; then
    coefficient dynamics c' = -nu vals c + rhs(c) advanced by ETDRK4."""

    def __init__(self, fields: nw.Fields, weight: float, dtype=torch.float64, rank_tol: float = 1e-12):
        if fields.layout != 0 or fields.mass.shape[0] != 2 * fields.stiff.shape[0]:
            raise ValueError("PeriodicROM needs point-major 2D velocity / vorticity fields")
        K = fields.K
        U = nw.field_matrix(fields, "mass", torch.float64)
        W = nw.field_matrix(fields, "stiff", torch.float64)
        M = nw.gram(U, K, weight, "f64")
        S = nw.gram(W, K, weight, "f64")
        norm = torch.sqrt(torch.diagonal(M))
        if not bool((norm > 0).all()):
            raise FloatingPointError("a basis mode has zero velocity")
        scale = 1.0 / norm
        Mn = M * scale[:, None] * scale[None, :]
        Sn = S * scale[:, None] * scale[None, :]
        self.mass_condition = float(torch.linalg.cond(0.5 * (Mn + Mn.T)))
        vals, T = sp.generalized_eigh(Sn, Mn, rank_tol)       # raises ArithmeticError below full rank
        T = scale[:, None] * T
        n = W.shape[0]
        self.gu = torch.mm(U.as_strided((n, K), (2 * U.stride(0), 1)), T).to(dtype).contiguous()
        self.gv = torch.mm(U.as_strided((n, K), (2 * U.stride(0), 1), U.storage_offset() + U.stride(0)),
                           T).to(dtype).contiguous()
        self.gw = torch.mm(W, T).to(dtype).contiguous()
        self.vals, self.weight, self.K = vals, float(weight), K

    def rhs(self, a: torch.Tensor) -> torch.Tensor:
        return rollout_rhs(self.gu, self.gv, self.gw, a, self.weight)

    def project(self, u: torch.Tensor, v: torch.Tensor) -> torch.Tensor:
        """Initial coefficients w (gu^T u + gv^T v) of a velocity field on the
        points (paper_rollout.py:39, 190)."""
        u = u.to(self.gu.dtype).reshape(-1, 1)
        v = v.to(self.gu.dtype).reshape(-1, 1)
        return (self.weight * (torch.mm(self.gu.T, u) + torch.mm(self.gv.T, v))).reshape(-1).to(torch.float64)

    def skew(self, a: torch.Tensor) -> float:
        """|a . rhs(a)| / |a|^3: zero for the energy-conserving nonlinearity
        (the check of paper_rollout.py:42-44)."""
        return float(torch.abs(torch.dot(a, self.rhs(a))) / torch.linalg.norm(a) ** 3)

    def rollout(self, c: torch.Tensor, nu: float, h: float, steps: int) -> torch.Tensor:
        """`steps` ETDRK4 steps of c' = -nu vals c + rhs(c) (paper_rollout.py:57-64)."""
        e, e2, q, f1, f2, f3 = etdrk4_coefficients(-nu * self.vals, h)
        c = c.to(torch.float64)
        for _ in range(steps):
            nv = self.rhs(c)
            a = e2 * c + q * nv
            na = self.rhs(a)
            b = e2 * c + q * na
            nb = self.rhs(b)
            z = e2 * a + q * (2 * nb - nv)
            c = e * c + f1 * nv + 2 * f2 * (na + nb) + f3 * self.rhs(z)
        return c


class SpectralGrid:
    """Periodic n^3 operators of the metaball derivation on the GPU
    (operators.py:20-60): 2/3-dealiased Leray projection, curl, viscous and
    normal-penalty forces; FP64 FFTs (rocFFT) over [B, 3, n, n, n] fields
    on the unit cube (wavenumbers 2 pi * integer, x fastest)."""

    def __init__(self, normal, chi, n: int, viscosity: float = 2.5e-4, penalty_time: float = 1e-3, device="cuda"):
        dev = torch.device(device)
        k = 2 * math.pi * torch.fft.fftfreq(n, d=1 / n, dtype=torch.float64, device=dev)
        self.kz, self.ky, self.kx = k[:, None, None], k[None, :, None], k[None, None, :]
        self.k2 = self.kx ** 2 + self.ky ** 2 + self.kz ** 2
        self.safe = self.k2.clone()
        self.safe[0, 0, 0] = 1
        m = torch.fft.fftfreq(n, d=1 / n, dtype=torch.float64, device=dev)
        ok = m.abs() <= n // 3
        self.dealias = ok[:, None, None] & ok[None, :, None] & ok[None, None, :]
        normal = torch.as_tensor(normal, dtype=torch.float64, device=dev)
        chi = torch.as_tensor(chi, dtype=torch.float64, device=dev)
        if tuple(normal.shape) != (n, n, n, 3) or tuple(chi.shape) != (n, n, n) \
                or not bool(torch.isfinite(normal).all()) or not bool(torch.isfinite(chi).all()):
            raise ValueError("complete same-grid normal and solid-fraction arrays required")
        self.normal = normal.permute(3, 0, 1, 2).contiguous()
        self.chi = chi
        self.n, self.nu, self.eta, self.device = n, viscosity, penalty_time, dev

    def _project(self, v):
        dot = (self.kx * v[:, 0] + self.ky * v[:, 1] + self.kz * v[:, 2]) / self.safe
        r = torch.stack((v[:, 0] - self.kx * dot, v[:, 1] - self.ky * dot, v[:, 2] - self.kz * dot), 1)
        r = r * self.dealias
        r[:, :, 0, 0, 0] = 0
        return r

    def _curl(self, v):
        return torch.stack((1j * (self.ky * v[:, 2] - self.kz * v[:, 1]), 1j * (self.kz * v[:, 0] - self.kx * v[:, 2]),
                            1j * (self.kx * v[:, 1] - self.ky * v[:, 0])), 1)

    def derived(self, u):
        """(projected, curl, linear force) of [B, 3, n, n, n] fields."""
        fft = lambda x: torch.fft.fftn(x, dim=(2, 3, 4))  # noqa: E731
        real = lambda x: torch.fft.ifftn(x, dim=(2, 3, 4)).real  # noqa: E731
        v = fft(u)
        viscous = self._project(-self.nu * self.k2 * v)
        un = torch.sum(u * self.normal[None], 1, keepdim=True)
        penalty = self._project(fft(-self.chi[None, None] / self.eta * un * self.normal[None]))
        return real(self._project(v)), real(self._curl(v)), real(viscous + penalty)


def derive_fields(B: torch.Tensor, flat: torch.Tensor, grid: SpectralGrid, block: int = 8) -> torch.Tensor:
    """operators.py:137-177 ``derive``: every mode's basis field B[:, :, k]
    ([N, 3, K] at the fluid points, flat indices into the n^3 grid, x
    fastest) zero-extended to the grid, then (Leray-projected field, curl,
    linear force) gathered back at the fluid points: [N, 9, K] FP64,
    `block` modes per batch of FFTs."""
    N, three, K = B.shape
    if three != 3 or flat.shape != (N,):
        raise ValueError("basis [N, 3, K] and N flat fluid indices required")
    n = grid.n
    flat = flat.to(device=grid.device, dtype=torch.long)
    out = torch.empty((N, 9, K), dtype=torch.float64, device=grid.device)
    for k0 in range(0, K, block):
        kb = min(block, K - k0)
        u = torch.zeros((kb, 3, n ** 3), dtype=torch.float64, device=grid.device)
        u[:, :, flat] = B[:, :, k0:k0 + kb].to(device=grid.device, dtype=torch.float64).permute(2, 1, 0)
        parts = grid.derived(u.view(kb, 3, n, n, n))
        for q, part in enumerate(parts):
            out[:, 3 * q:3 * q + 3, k0:k0 + kb] = part.reshape(kb, 3, n ** 3)[:, :, flat].permute(2, 1, 0)
    return out


_CROSS_TERMS = ((0, 1, 2, 1.0), (0, 2, 1, -1.0), (1, 2, 0, 1.0), (1, 0, 2, -1.0), (2, 0, 1, 1.0), (2, 1, 0, -1.0))


def assemble_operators(B: torch.Tensor, D: torch.Tensor, dV: float):
    """operators.py:199-257 ``assemble`` in FP64 on the GPU: full mass
    M = dV B^T B (Gram kernel), linear operator A = dV B^T L (cross kernel
    over the three components) and the packed convective tensor
    Q[i, (j <= k)] = dV (q[j,i,k] + q[k,i,j]) (j != k), dV q[j,i,j], with
    q[j,i,k] = sum_p PB_i . (B_j x curlB_k): for each output i, the six
    signed terms PB_i^a B_j^b curlB_k^c of the cross product are cross
    products of the B^b and curlB^c columns with row weights PB_i^a.
    B: [N, 3, K], D: [N, 9, K] (projected, curl, linear); pairs in
    numpy.triu_indices(K) order."""
    N, _, K = B.shape
    B = B.to(torch.float64).contiguous()
    D = D.to(torch.float64).contiguous()
    F64 = nw.GRAM_MODES["f64"]
    M = nw.gram(B.view(N * 3, K), K, dV, "f64")
    A = nw.cross([(B[:, c, :], D[:, 6 + c, :], N, 1.0) for c in range(3)], K, K, F64, scale=dV)
    jj, kk = torch.triu_indices(K, K, device=B.device)
    diff = jj != kk
    Q = torch.empty((K, jj.numel()), dtype=torch.float64, device=B.device)
    raw = torch.empty((K, K), dtype=torch.float64, device=B.device)
    weights = D[:, 0:3, :].permute(1, 2, 0).contiguous()            # [3, K, N]: PB_i^a as rows
    for i in range(K):
        for t, (a, b, c, sign) in enumerate(_CROSS_TERMS):
            nw.cross([(B[:, b, :], D[:, 3 + c, :], N, 1.0)], K, K, F64, scale=sign * dV, out=raw,
                     beta=0.0 if t == 0 else 1.0, roww=weights[a, i])
        Q[i] = raw[jj, kk]
        Q[i, diff] += raw[kk[diff], jj[diff]]
    return M, A, Q
