"""FP64 NumPy oracle for the neural eigenfluid network hot path.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this module, and only as the
checker / CPU baseline.  The product path (``neb_amd``) never imports it.

It restates, in plain FP64 array code, the algorithms of the reference backend
(baiming-zhang/NEB; paths relative to the reference root, ``HB2P`` =
experiments/2d/_backend/periodic_core/vortex_merger/src/eigenfluid/gen/
vortex_merger2d/hardcoded_backend.py, ``VJ2`` = experiments/metaball/_backend/
flow/cases/metaball/velocity_trace_jet2.py, ``VCORE`` = .../metaball/velocity/
core.py, ``NCC`` = .../metaball/native_cupy_core.py, ``PP`` =
experiments/metaball/_backend/exports/basis_projection/postprocessing/
paired_projection.py), applied to the shared-trunk tanh network of the
benchmark configurations:

    features (periodic sin/cos jets)  ->  L hidden layers act(omega*(x W^T + b))
    ->  linear output layer  ->  differential lift (2D: u=(psi_y,-psi_x),
    omega=-lap psi;  3D: u = curl A, grad u)  ->  Gram  Phi^T diag(w) Phi
    ->  spectral loss  tr((M+eps I)^-1 S) + lam ||M-I||_F^2  ->  reverse pass.

Jets are compact second-order Taylor jets: value [n, f], grad [n, f, d],
hess [n, f, P] with P = d(d+1)/2 unique Hessian entries in PAIRS order
((0,0),(0,1),(1,1)) in 2D and ((0,0),(0,1),(0,2),(1,1),(1,2),(2,2)) in 3D
(HB2P:154-166, VJ2:9).

Pinning: every function here is checked in tests/test_oracle.py against
closed forms (exact Fourier-basis Gram = I, manufactured curl/Laplacian
fields), central finite differences (spatial jets and the parameter reverse
pass) and library routines (scipy.linalg.eigh, dense solves).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np


def pairs(d: int) -> list[tuple[int, int]]:
    """Unique Hessian index pairs in the reference order (VJ2:9, HB2P:159)."""
    return [(i, j) for i in range(d) for j in range(i, d)]


def channels(d: int) -> int:
    """Packed Jet2 channel count: value + d grads + d(d+1)/2 Hessian entries."""
    return 1 + d + d * (d + 1) // 2


@dataclass
class Jet:
    value: np.ndarray  # [n, f]
    grad: np.ndarray   # [n, f, d]
    hess: np.ndarray   # [n, f, P]

    def pack(self) -> np.ndarray:
        """Channel-major packed operand [C, n, f] (VJ2:20-22, HB2P:1014-1025)."""
        d = self.grad.shape[-1]
        return np.stack([self.value]
                        + [self.grad[..., a] for a in range(d)]
                        + [self.hess[..., p] for p in range(self.hess.shape[-1])])

    @staticmethod
    def unpack(packed: np.ndarray, d: int) -> "Jet":
        P = d * (d + 1) // 2
        return Jet(packed[0], np.stack(list(packed[1:1 + d]), -1),
                   np.stack(list(packed[1 + d:1 + d + P]), -1))


# --------------------------------------------------------------------------
# a1: periodic sin/cos feature jets (HB2P:576-589 / :428-441)
# --------------------------------------------------------------------------
def feature_jets(x: np.ndarray, frequencies: int, lo, hi) -> Jet:
    """Features ordered [for coord: for h in 1..F: (sin, cos)] of
    phase = pi*h*(xn+1), xn = 2(x-lo)/(hi-lo)-1  (HB2P:570-589)."""
    x = np.asarray(x, dtype=np.float64)
    n, d = x.shape
    lo = np.broadcast_to(np.asarray(lo, dtype=np.float64), (d,))
    hi = np.broadcast_to(np.asarray(hi, dtype=np.float64), (d,))
    F = 2 * frequencies * d
    P = d * (d + 1) // 2
    value = np.zeros((n, F))
    grad = np.zeros((n, F, d))
    hess = np.zeros((n, F, P))
    pl = pairs(d)
    col = 0
    for a in range(d):
        scale = 2.0 / (hi[a] - lo[a])
        xn = scale * (x[:, a] - lo[a]) - 1.0
        diag = pl.index((a, a))
        for h in range(1, frequencies + 1):
            k = math.pi * h
            phase = k * (xn + 1.0)
            dp = k * scale
            s, c = np.sin(phase), np.cos(phase)
            value[:, col], grad[:, col, a], hess[:, col, diag] = s, c * dp, -s * dp * dp
            value[:, col + 1], grad[:, col + 1, a], hess[:, col + 1, diag] = c, -s * dp, -c * dp * dp
            col += 2
    return Jet(value, grad, hess)


# --------------------------------------------------------------------------
# activations: f, f', f'', f''' (tanh for the benchmark net; sin = SIREN,
# HB2P:1038-1040, FUSED:3-19)
# --------------------------------------------------------------------------
def activation_derivs(kind: str, u: np.ndarray):
    if kind == "tanh":
        t = np.tanh(u)
        d1 = 1.0 - t * t
        return t, d1, -2.0 * t * d1, -2.0 * d1 * (1.0 - 3.0 * t * t)
    if kind == "sin":
        s, c = np.sin(u), np.cos(u)
        return s, c, -s, -c
    raise ValueError(f"unknown activation {kind!r}")


def envelope_product(e: Jet, x: Jet) -> Jet:
    """Jet2 product y = e x of a per-point scalar envelope jet e (value [n],
    grad [n, d], hess [n, P]) with output jets x ([n, f] ...):
    grad y = e grad x + x grad e, hess_ij y = e hess_ij x + x hess_ij e +
    e_i x_j + e_j x_i (VJ2:47-50 `product`)."""
    ev, eg, eh = e.value[:, None], e.grad[:, None, :], e.hess[:, None, :]
    d = x.grad.shape[-1]
    hess = np.stack([ev * x.hess[..., p] + eh[..., p] * x.value + eg[..., i] * x.grad[..., j]
                     + eg[..., j] * x.grad[..., i] for p, (i, j) in enumerate(pairs(d))], -1)
    return Jet(ev * x.value, ev[..., None] * x.grad + eg * x.value[..., None], hess)


def envelope_product_reverse(e: Jet, a: Jet) -> Jet:
    """Transpose of envelope_product in x (VJ2:51-55 `product_reverse`)."""
    ev, eg, eh = e.value[:, None], e.grad[:, None, :], e.hess[:, None, :]
    d = a.grad.shape[-1]
    av = ev * a.value + np.sum(eg * a.grad, -1) + np.sum(eh * a.hess, -1)
    ag = ev[..., None] * a.grad
    for p, (i, j) in enumerate(pairs(d)):
        ag[..., i] += eg[..., j] * a.hess[..., p]
        ag[..., j] += eg[..., i] * a.hess[..., p]
    return Jet(av, ag, ev[..., None] * a.hess)


def linear_jet(x: Jet, W: np.ndarray, b: np.ndarray | None) -> Jet:
    """Packed-channel linear layer: all channels times W^T, bias on value rows
    only (HB2P:1005-1036, VJ2:26-27)."""
    v = x.value @ W.T
    if b is not None:
        v = v + b
    return Jet(v, _chan_matmul(x.grad, W.T), _chan_matmul(x.hess, W.T))


def _chan_matmul(t: np.ndarray, A: np.ndarray) -> np.ndarray:
    """t [n, i, c] contracted over i with A [i, o] -> [n, o, c] (each tangent
    channel through the same matrix)."""
    return np.swapaxes(np.swapaxes(t, 1, 2) @ A, 1, 2)


def _packed_outer(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """sum over channels and points of a[c, n, o] b[c, n, i] -> [o, i]."""
    return a.reshape(-1, a.shape[-1]).T @ b.reshape(-1, b.shape[-1])


def activation_jet(kind: str, pre: Jet) -> Jet:
    """y=f(u), grad y = f' grad u, hess y = f' hess u + f'' grad u grad u^T
    (HB2P:252-272 unary2, VJ2:31-34)."""
    f0, f1, f2, _ = activation_derivs(kind, pre.value)
    d = pre.grad.shape[-1]
    hess = np.stack([f1 * pre.hess[..., p] + f2 * pre.grad[..., i] * pre.grad[..., j]
                     for p, (i, j) in enumerate(pairs(d))], -1)
    return Jet(f0, f1[..., None] * pre.grad, hess)


def activation_jet_reverse(kind: str, pre: Jet, adj: Jet) -> Jet:
    """Transpose of activation_jet w.r.t. the pre-activation jet
    (HB2P:1147-1174 unary_backward_jet2, VJ2:35-41, FUSED:9-19)."""
    _, f1, f2, f3 = activation_derivs(kind, pre.value)
    d = pre.grad.shape[-1]
    g = pre.grad
    av = adj.value * f1 + f2 * np.sum(adj.grad * g, -1)
    ag = f1[..., None] * adj.grad
    for p, (i, j) in enumerate(pairs(d)):
        ah = adj.hess[..., p]
        av = av + ah * (f2 * pre.hess[..., p] + f3 * g[..., i] * g[..., j])
        ag[..., i] += f2 * ah * g[..., j]
        ag[..., j] += f2 * ah * g[..., i]
    return Jet(av, ag, f1[..., None] * adj.hess)


# --------------------------------------------------------------------------
# network: shared trunk, Xavier-normal init
# --------------------------------------------------------------------------
@dataclass
class Net:
    W: list  # [out, in] per layer (hidden..., output)
    b: list  # [out]
    activation: str = "tanh"
    omegas: list = field(default_factory=list)  # per hidden layer

    @property
    def hidden_layers(self) -> int:
        return len(self.W) - 1


def init_net(sizes: list[int], seed: int, activation: str = "tanh", omegas=None) -> Net:
    """Xavier-normal weights, std = sqrt(2/(fan_in+fan_out)), drawn layer by
    layer from numpy.random.default_rng(seed) as rng.normal(size=(out, in));
    zero biases."""
    rng = np.random.default_rng(seed)
    W, b = [], []
    for fi, fo in zip(sizes[:-1], sizes[1:]):
        W.append(rng.normal(size=(fo, fi)) * math.sqrt(2.0 / (fi + fo)))
        b.append(np.zeros(fo))
    L = len(sizes) - 2
    om = list(omegas) if omegas is not None else [1.0] * L
    return Net(W, b, activation, om)


def forward(net: Net, x0: Jet, record: bool = False):
    """Trunk + linear output layer.  Returns (output jet, tape) where the tape
    holds (layer input jet, scaled pre-activation jet) per hidden layer and
    the last hidden output (HB2P:1802-1880, VJ2:65-76)."""
    x = x0
    tape = []
    for l in range(net.hidden_layers):
        z = linear_jet(x, net.W[l], net.b[l])
        om = net.omegas[l]
        pre = Jet(om * z.value, om * z.grad, om * z.hess)
        if record:
            tape.append((x, pre))
        x = activation_jet(net.activation, pre)
    out = linear_jet(x, net.W[-1], net.b[-1])
    return out, (tape, x)


def backward(net: Net, tape, adj_out: Jet):
    """Handwritten parameter reverse pass through the jets (HB2P:2087-2139,
    VJ2:77-82, linear transpose HB2P:1115-1144)."""
    layers, last = tape
    gW = [None] * len(net.W)
    gb = [None] * len(net.W)
    a = adj_out
    gW[-1] = _packed_outer(a.pack(), last.pack())
    gb[-1] = a.value.sum(0)
    W = net.W[-1]
    a = Jet(a.value @ W, _chan_matmul(a.grad, W), _chan_matmul(a.hess, W))
    for l in range(net.hidden_layers - 1, -1, -1):
        src, pre = layers[l]
        a = activation_jet_reverse(net.activation, pre, a)
        om = net.omegas[l]
        a = Jet(om * a.value, om * a.grad, om * a.hess)
        gW[l] = _packed_outer(a.pack(), src.pack())
        gb[l] = a.value.sum(0)
        W = net.W[l]
        a = Jet(a.value @ W, _chan_matmul(a.grad, W), _chan_matmul(a.hess, W))
    return gW, gb


# --------------------------------------------------------------------------
# a5: differential lifts and their transposes
# --------------------------------------------------------------------------
def lift2d(out: Jet):
    """psi jets [n, K] -> (mass fields [n, 2, K] = (psi_y, -psi_x),
    stiffness field [n, K] = omega = -(psi_xx + psi_yy))  (ELC:35-38,
    HB2P:2530-2546)."""
    u = np.stack((out.grad[..., 1], -out.grad[..., 0]), 1)
    w = -(out.hess[..., 0] + out.hess[..., 2])
    return u, w


def lift2d_reverse(gu: np.ndarray, gw: np.ndarray) -> Jet:
    n, _, K = gu.shape
    grad = np.stack((-gu[:, 1], gu[:, 0]), -1)
    hess = np.stack((-gw, np.zeros_like(gw), -gw), -1)
    return Jet(np.zeros((n, K)), grad, hess)


def lift2d_hessian(out: Jet):
    """psi jets [n, K] -> (mass fields [n, 2, K] = (psi_y, -psi_x), Hessian
    rows [n, 3, K] = (psi_xx, sqrt2 psi_xy, psi_yy)): the velocity mass
    integral psi_x^2 + psi_y^2 and the H1 seminorm
    psi_xx^2 + 2 psi_xy^2 + psi_yy^2 of the causal velocity peel
    (HB2P:2946-2950, the pair metric (1, 2, 1) at HB2P:3022) as Grams of
    these rows."""
    u = np.stack((out.grad[..., 1], -out.grad[..., 0]), 1)
    h = np.stack((out.hess[..., 0], math.sqrt(2.0) * out.hess[..., 1], out.hess[..., 2]), 1)
    return u, h


def lift2d_hessian_reverse(gu: np.ndarray, gh: np.ndarray) -> Jet:
    n, _, K = gu.shape
    grad = np.stack((-gu[:, 1], gu[:, 0]), -1)
    hess = np.stack((gh[:, 0], math.sqrt(2.0) * gh[:, 1], gh[:, 2]), -1)
    return Jet(np.zeros((n, K)), grad, hess)


def velocity_rayleigh_peel(grad: np.ndarray, hess: np.ndarray, w: float, order: str = "rayleigh",
                           passes: int = 2, mass_floor: float = 1e-12):
    """The 2D causal velocity peel, mode by mode (HB2P:3017-3134): raw
    gradient rows [2n, K] (psi_x, psi_y per point), raw Hessian rows [3n, K]
    (psi_xx, psi_xy, psi_yy per point, metric (1, 2, 1)); modes in the stable
    order of raw K/M ("rayleigh") or in network order ("fixed"); each
    candidate's gradient column orthogonalised `passes` times against the
    peeled (normalised) modes in the velocity mass metric, its Hessian column
    with the same coefficients; Rayleigh quotient s / max(m, mass_floor);
    the cotangents of each mode's quotient with the earlier modes held
    constant, reversed through the projections (HB2P:3098-3116).  Returns
    dict(rayleigh (causal order), mass, relative_mass, orth, order,
    G_grad, G_hess (cotangents [2n, K] / [3n, K] in network order), raw_mass,
    raw_stiffness)."""
    n = grad.shape[0] // 2
    K = grad.shape[1]
    metric = np.tile(np.array([1.0, 2.0, 1.0]), n)
    raw_mass = w * np.sum(grad * grad, 0)
    raw_stiff = w * np.sum(metric[:, None] * hess * hess, 0)
    causal = np.argsort(raw_stiff / raw_mass, kind="stable") if order == "rayleigh" else np.arange(K)
    Qg = np.empty_like(grad)
    Qh = np.empty_like(hess)
    Gg = np.zeros_like(grad)
    Gh = np.zeros_like(hess)
    ray, mass, rel, orth = (np.empty(K) for _ in range(4))
    for r, net in enumerate(causal):
        rg = grad[:, net].copy()
        rh = hess[:, net].copy()
        pg, ph = Qg[:, :r], Qh[:, :r]
        for _ in range(passes if r else 0):
            inc = w * (pg.T @ rg)
            rg -= pg @ inc
            rh -= ph @ inc
        m = w * float(rg @ rg)
        st = w * float((metric * rh) @ rh)
        safe = max(m, mass_floor)
        ray[r], mass[r], rel[r] = st / safe, m, m / raw_mass[net]
        ag = (-2.0 * st * w / (safe * safe)) * rg
        ah = (2.0 * w / safe) * metric * rh
        for _ in range(passes if r else 0):
            ca = -(pg.T @ ag + ph.T @ ah)
            ag = ag + w * (pg @ ca)
        Gg[:, net] = ag
        Gh[:, net] = ah
        qg, qh = rg / math.sqrt(safe), rh / math.sqrt(safe)
        Qg[:, r], Qh[:, r] = qg, qh
        o = abs(w * float(qg @ qg) - 1.0)
        if r:
            o = max(o, float(np.max(np.abs(w * (pg.T @ qg)))))
        orth[r] = o
    return dict(rayleigh=ray, mass=mass, relative_mass=rel, orth=orth, order=causal, G_grad=Gg, G_hess=Gh,
                raw_mass=raw_mass, raw_stiffness=raw_stiff)


def lift2d_stream(out: Jet):
    """psi jets [n, K] -> (mass rows [n, K] = psi, stiffness rows [n, 2, K] =
    (psi_x, psi_y)): M = integral psi^2 and K = integral |grad psi|^2 of the
    Dirichlet streamfunction peel (HB2P:3394-3397, the raw value and gradient
    columns of :3461-3474) as Grams of these rows.  Raw: the peel removes
    the mean."""
    return out.value.copy(), np.stack((out.grad[..., 0], out.grad[..., 1]), 1)


def lift2d_stream_reverse(gv: np.ndarray, gg: np.ndarray) -> Jet:
    n, K = gv.shape
    return Jet(gv.copy(), np.stack((gg[:, 0], gg[:, 1]), -1), np.zeros((n, K, 3)))


def streamfunction_rayleigh_peel(value: np.ndarray, grad: np.ndarray, w: float, order: str = "rayleigh",
                                 passes: int = 2, center: bool = True):
    """The 2D Dirichlet streamfunction Rayleigh peel on Jet1 fields, mode by
    mode (HB2P:3455-3583): raw values [n, K] and gradient rows [2n, K]
    (psi_x, psi_y per point); with `center` (the periodic gauge) each value
    column loses its mean over the points first (:3475-3479); modes in the
    stable order of raw K/M ("rayleigh", the geometry order) or in network
    order ("fixed"); each candidate's value column orthogonalised `passes`
    times against the peeled (normalised) modes in the w-weighted L2 metric,
    its gradient column with the same coefficients (:3508-3516); Rayleigh
    quotient K/M of the residual; the cotangents of each mode's quotient with
    the earlier modes held constant, reversed through the projections
    (:3525-3542), then through the centering (:3579-3583).  Returns
    dict(rayleigh, mass, stiffness, projection, orth (causal order), order,
    G_value [n, K], G_grad [2n, K] (network order), raw_mass, raw_stiffness)."""
    n, K = value.shape
    v = value - value.mean(0, keepdims=True) if center else value.copy()
    raw_mass = w * np.sum(v * v, 0)
    raw_stiff = w * np.sum(grad * grad, 0)
    causal = np.argsort(raw_stiff / raw_mass, kind="stable") if order == "rayleigh" else np.arange(K)
    Qv = np.empty_like(v)
    Qg = np.empty_like(grad)
    Gv = np.zeros_like(v)
    Gg = np.zeros_like(grad)
    ray, mass, stiff, proj, orth = (np.empty(K) for _ in range(5))
    for r, net in enumerate(causal):
        rv = v[:, net].copy()
        rg = grad[:, net].copy()
        coeff = np.zeros(r)
        pv, pg = Qv[:, :r], Qg[:, :r]
        for _ in range(passes if r else 0):
            inc = w * (pv.T @ rv)
            coeff += inc
            rv -= pv @ inc
            rg -= pg @ inc
        m = w * float(rv @ rv)
        s = w * float(rg @ rg)
        ray[r], mass[r], stiff[r] = s / m, m, s
        av = (-2.0 * s * w / (m * m)) * rv
        ag = (2.0 * w / m) * rg
        for _ in range(passes if r else 0):
            ca = -(pv.T @ av + pg.T @ ag)
            av = av + w * (pv @ ca)
        Gv[:, net] = av
        Gg[:, net] = ag
        qv, qg = rv / math.sqrt(m), rg / math.sqrt(m)
        Qv[:, r], Qg[:, r] = qv, qg
        o = abs(w * float(qv @ qv) - 1.0)
        if r:
            o = max(o, float(np.max(np.abs(w * (pv.T @ qv)))))
        orth[r] = o
        proj[r] = float(coeff @ coeff) / m if r else 0.0
    if center:
        Gv -= Gv.mean(0, keepdims=True)
    return dict(rayleigh=ray, mass=mass, stiffness=stiff, projection=proj, orth=orth, order=causal, G_value=Gv,
                G_grad=Gg, raw_mass=raw_mass, raw_stiffness=raw_stiff)


def _hid(a: int, b: int) -> int:
    return pairs(3).index(tuple(sorted((a, b))))


def lift3d(out: Jet):
    """Vector potential jets with columns 3k+c -> (u [n, 3, K] = curl A,
    du [n, 3, 3, K] with du[:, a, j] = d_j u_a)  (NCC:97-99, VJ2:93-101)."""
    n, KC = out.value.shape
    K = KC // 3
    g = out.grad.reshape(n, K, 3, 3)   # [n, k, comp, axis]
    h = out.hess.reshape(n, K, 3, 6)   # [n, k, comp, pair]
    u = np.stack((g[:, :, 2, 1] - g[:, :, 1, 2], g[:, :, 0, 2] - g[:, :, 2, 0],
                  g[:, :, 1, 0] - g[:, :, 0, 1]), 1)
    du = np.stack([np.stack([h[:, :, (a + 2) % 3, _hid(j, (a + 1) % 3)]
                             - h[:, :, (a + 1) % 3, _hid(j, (a + 2) % 3)]
                             for j in range(3)], 1) for a in range(3)], 1)
    return u, du


def lift3d_reverse(gu: np.ndarray, gd: np.ndarray) -> Jet:
    """Transpose of lift3d (NCC:101-109 curl_adjoint, VCORE:32-38 lift)."""
    n, _, K = gu.shape
    g = np.zeros((n, K, 3, 3))
    g[:, :, 2, 1] += gu[:, 0]; g[:, :, 1, 2] -= gu[:, 0]
    g[:, :, 0, 2] += gu[:, 1]; g[:, :, 2, 0] -= gu[:, 1]
    g[:, :, 1, 0] += gu[:, 2]; g[:, :, 0, 1] -= gu[:, 2]
    h = np.zeros((n, K, 3, 6))
    for a in range(3):
        for j in range(3):
            h[:, :, (a + 2) % 3, _hid(j, (a + 1) % 3)] += gd[:, a, j]
            h[:, :, (a + 1) % 3, _hid(j, (a + 2) % 3)] -= gd[:, a, j]
    return Jet(np.zeros((n, 3 * K)), g.reshape(n, 3 * K, 3), h.reshape(n, 3 * K, 6))


def field_rows(f: np.ndarray) -> np.ndarray:
    """[n, comp..., K] -> [n*comp, K], point-major, component-minor (VCORE:75-76)."""
    return f.reshape(f.shape[0] * int(np.prod(f.shape[1:-1], dtype=np.int64)), f.shape[-1])


# --------------------------------------------------------------------------
# a6: Gram / stiffness assembly
# --------------------------------------------------------------------------
def gram(Phi: np.ndarray, w, block: int = 3072) -> np.ndarray:
    """Blocked FP64 G = sum_r w_r Phi_r^T Phi_r (VCORE:40-46, metaball/run.py:99-107).
    ``w`` is a scalar or a per-row weight vector."""
    Phi = np.asarray(Phi)
    w = np.broadcast_to(np.asarray(w, dtype=np.float64), (Phi.shape[0],))
    G = np.zeros((Phi.shape[1], Phi.shape[1]))
    for s in range(0, Phi.shape[0], block):
        B = Phi[s:s + block].astype(np.float64)
        G += B.T @ (w[s:s + block, None] * B)
    return G


# --------------------------------------------------------------------------
# a7/a8: spectral objective and its adjoint
# --------------------------------------------------------------------------
def spectral_loss(M: np.ndarray, S: np.ndarray, lam: float = 1e-2, ridge: float = 0.0,
                  cond_limit: float = 1e12):
    """L = tr((M + ridge I)^-1 S) + lam ||M - I||_F^2 with the reference's
    rank/condition gate (VJ2:107-117: nonfinite, nonpositive or condition
    > 1e12 raises FloatingPointError).  Returns (L, dL/dM, dL/dS)."""
    K = len(M)
    Mr = 0.5 * (M + M.T) + ridge * np.eye(K)
    e = np.linalg.eigvalsh(Mr)
    if not np.all(np.isfinite(e)) or e[-1] <= 0 or e[0] / e[-1] < 1.0 / cond_limit:
        raise FloatingPointError("mass matrix nonfinite, rank deficient or condition > limit")
    L = np.linalg.cholesky(Mr)
    inv = np.linalg.solve(L.T, np.linalg.solve(L, np.eye(K)))
    inv = 0.5 * (inv + inv.T)
    Ss = 0.5 * (S + S.T)
    D = M - np.eye(K)
    loss = float(np.trace(inv @ Ss) + lam * np.sum(D * D))
    cross = inv @ Ss @ inv
    dM = -0.5 * (cross + cross.T) + 2.0 * lam * D
    dS = inv
    return loss, dM, dS


def field_adjoint(Phi: np.ndarray, w, dG: np.ndarray) -> np.ndarray:
    """d/dPhi of <dG, Phi^T W Phi> for symmetric dG: 2 W Phi dG (VCORE:94-95)."""
    w = np.asarray(w, dtype=np.float64)
    if w.ndim:
        w = w[:, None]
    return 2.0 * w * (Phi @ dG)


# --------------------------------------------------------------------------
# next row: sorted Cholesky coefficient peel (VCORE:48-69) and its literal
# modified-Gram-Schmidt form (VCORE:107-119)
# --------------------------------------------------------------------------
def coefficient_peel(M: np.ndarray, S: np.ndarray, ortho_tol: float = 1e-5, rel_min: float = 1e-4):
    """Causal peel of the modes sorted by their raw Rayleigh quotient
    diag(S)/diag(M) (stable order, VCORE:49-50): T = inverse upper Cholesky
    factor of the diagonally scaled sorted mass, refined by a second
    Cholesky pass (VCORE:51-56), so that T^T M_sorted T = I; loss_r =
    (T^T S_sorted T)_rr.  Gates (VCORE:57-61): every loss finite and
    positive, relative mass norm_r^2 / M_rr >= rel_min (norm = 1/diag T),
    max |T^T M T - I| <= ortho_tol, else FloatingPointError.

    The detached-prefix adjoint (VCORE:64-67): the field cotangents of
    sum_r loss_r with every mode's earlier (peeled) modes held constant are
    G_U = w U Cu and G_D = w D Cd for Grams M = w U^T U, S = w D^T D.
    Returns dict(losses (sorted order), order, T (rows in original mode
    order, columns in sorted order), Cu, Cd (original order), relative_mass,
    gram_max, raw)."""
    k = len(M)
    raw = np.diag(S) / np.diag(M)
    order = np.argsort(raw, kind="stable")
    ms = M[np.ix_(order, order)]
    ss = S[np.ix_(order, order)]
    scale = np.sqrt(np.diag(ms))
    Lc = np.linalg.cholesky(ms / scale[:, None] / scale[None, :])
    T = np.linalg.inv(Lc.T) / scale[:, None]
    H2 = T.T @ ms @ T
    L2 = np.linalg.cholesky(0.5 * (H2 + H2.T))
    T = T @ np.linalg.inv(L2.T)
    error = float(np.max(np.abs(T.T @ ms @ T - np.eye(k))))
    norm = 1.0 / np.diag(T)
    relative = norm * norm / np.diag(ms)
    F = T.T @ ss @ T
    loss = np.diag(F).copy()
    if not (np.all(np.isfinite(loss) & (loss > 0) & (relative >= rel_min)) and error <= ortho_tol):
        raise FloatingPointError("rank / relative mass / orthogonality gate failed; no update")
    C = -2.0 * (T * loss[None, :] + T @ np.triu(F, 1)) / norm[None, :]
    B = 2.0 * T / norm[None, :]
    Cu = np.empty_like(C)
    Cd = np.empty_like(B)
    Tr = np.empty_like(T)
    Cu[np.ix_(order, order)] = C
    Cd[np.ix_(order, order)] = B
    Tr[order] = T
    return dict(losses=loss, order=order, T=Tr, Cu=Cu, Cd=Cd, relative_mass=relative, gram_max=error, raw=raw)


def literal_peel(U: np.ndarray, D: np.ndarray, w: float):
    """Mode-by-mode form of coefficient_peel (VCORE:107-119), with quadrature
    weight w (the reference's 1/n): modes in the stable order of
    sum D^2 / sum U^2; each mode's velocity column is orthogonalised twice
    against the already peeled (normalised) modes, its D column with the same
    coefficients; loss = s/m; the cotangents of loss w.r.t. that mode's own
    columns, earlier modes held constant, are gd = 2 w d / m and
    gu = -2 s w u / m^2 projected twice off the earlier modes.  Returns
    (losses, GU, GD, Q, V, order)."""
    k = U.shape[1]
    order = np.argsort(np.sum(D * D, 0) / np.sum(U * U, 0), kind="stable")
    Q = np.empty_like(U)
    V = np.empty_like(D)
    GU = np.empty_like(U)
    GD = np.empty_like(D)
    losses = []
    for r, index in enumerate(order):
        u = U[:, index].copy()
        d = D[:, index].copy()
        q, v = Q[:, :r], V[:, :r]
        for _ in range(2):
            a = w * (q.T @ u)
            u -= q @ a
            d -= v @ a
        m = w * float(u @ u)
        s = w * float(d @ d)
        losses.append(s / m)
        gu = -2.0 * s * w * u / (m * m)
        gd = 2.0 * w * d / m
        for _ in range(2):
            gu -= w * (q @ (q.T @ gu + v.T @ gd))
        GU[:, index] = gu
        GD[:, index] = gd
        Q[:, r] = u / math.sqrt(m)
        V[:, r] = d / math.sqrt(m)
    return np.asarray(losses), GU, GD, Q, V, order


# --------------------------------------------------------------------------
# a7: FP64 generalized symmetric eigendecomposition with the rank gate
# (experiments/periodic/paper_rollout.py:180-188)
# --------------------------------------------------------------------------
def generalized_eigh(S: np.ndarray, M: np.ndarray, rank_tol: float = 1e-12):
    import scipy.linalg as la
    Ms = 0.5 * (M + M.T)
    Ss = 0.5 * (S + S.T)
    ev = np.linalg.eigvalsh(Ms)
    rank = int(np.sum(ev > ev[-1] * rank_tol))
    if rank != len(M):
        raise ArithmeticError(f"Full-K rank gate: rank {rank} < K={len(M)}")
    return la.eigh(Ss, Ms)


# --------------------------------------------------------------------------
# a12: full-mass projection (PP:141-177, experiments/2d/check_cpu.py:27-33)
# --------------------------------------------------------------------------
def project(M: np.ndarray, B: np.ndarray, T: np.ndarray, w, block: int = 2048):
    """Two-pass streamed projection of targets T [R, m] onto basis B [R, K]:
    loads = w B^T T, c = M^-1 loads, normal residual gate, then the direct
    residual pass and the Gram identity check."""
    K = M.shape[0]
    R = B.shape[0]
    w = np.broadcast_to(np.asarray(w, dtype=np.float64), (R,))
    loads = np.zeros((K, T.shape[1]))
    norms = np.zeros(T.shape[1])
    for s in range(0, R, block):
        b = B[s:s + block].astype(np.float64)
        t = T[s:s + block].astype(np.float64)
        ww = w[s:s + block, None]
        loads += b.T @ (ww * t)
        norms += np.sum(ww * t * t, 0)
    c = np.linalg.solve(M, loads)
    normal = np.linalg.norm(M @ c - loads, axis=0) / np.linalg.norm(loads, axis=0)
    gram_square = norms - 2 * np.sum(c * loads, 0) + np.sum(c * (M @ c), 0)
    direct = np.zeros(T.shape[1])
    for s in range(0, R, block):
        diff = B[s:s + block].astype(np.float64) @ c - T[s:s + block]
        direct += np.sum(w[s:s + block, None] * diff * diff, 0)
    return dict(loads=loads, target_norms=norms, coefficients=c, normal_residuals=normal,
                direct_relative_L2=np.sqrt(direct / norms),
                Gram_identity_errors=np.sqrt(np.maximum(gram_square, 0) / norms))


# --------------------------------------------------------------------------
# a11: Adam with global-norm clip (NCC:156-171; the shared trunk has one
# parameter group, so the per-mode clip becomes one global clip)
# --------------------------------------------------------------------------
def adam(params, grads, m, v, t: int, lr: float, clip: float = 8.0,
         b1: float = 0.9, b2: float = 0.999, eps: float = 1e-8):
    norm = math.sqrt(sum(float(np.sum(np.asarray(g, dtype=np.float64) ** 2)) for g in grads))
    if not math.isfinite(norm) or norm <= 0:
        raise FloatingPointError("Invalid parameter gradients; no optimizer update")
    scale = min(1.0, clip / max(norm, 1e-30))
    out = []
    for p, g, mm, vv in zip(params, grads, m, v):
        g = g * scale
        mm = b1 * mm + (1 - b1) * g
        vv = b2 * vv + (1 - b2) * g * g
        p = p - lr * (mm / (1 - b1 ** t)) / (np.sqrt(vv / (1 - b2 ** t)) + eps)
        out.append((p, mm, vv))
    return out, norm


# --------------------------------------------------------------------------
# synthetic inputs of the benchmark configurations
# --------------------------------------------------------------------------
def periodic_grid(n: int) -> np.ndarray:
    """n x n grid on [0,2pi)^2, meshgrid indexing='xy' (ELC:26-28)."""
    axis = np.arange(n) * 2 * np.pi / n
    x, y = np.meshgrid(axis, axis, indexing="xy")
    return np.column_stack((x.ravel(), y.ravel()))


def blob_domain_points(count: int, seed: int, blobs: int = 6, sigma: float = 0.35,
                       level: float = 0.5, batch: int = 1 << 20):
    """Rejection-sample ``count`` points of [-1,1]^3 where
    sum_b exp(-|x-c_b|^2/(2 sigma^2)) > level; centres uniform in [-0.6,0.6]^3."""
    rng = np.random.default_rng(seed)
    centres = rng.uniform(-0.6, 0.6, (blobs, 3))
    out = []
    have = 0
    while have < count:
        p = rng.uniform(-1.0, 1.0, (batch, 3))
        s = np.zeros(batch)
        for c in centres:
            s += np.exp(-np.sum((p - c) ** 2, 1) / (2 * sigma * sigma))
        keep = p[s > level]
        out.append(keep)
        have += len(keep)
    return np.concatenate(out)[:count], centres


def fourier_potential_targets(points: np.ndarray, count: int, seed: int, kmax: int = 6,
                              waves: int = 16):
    """Divergence-free targets u = curl A with A a Gaussian-weighted random
    Fourier series (|k| <= kmax), analytic curl as in
    experiments/metaball/large_eddy.py:23-30.  Returns [n, 3, count]."""
    rng = np.random.default_rng(seed)
    n = len(points)
    out = np.zeros((n, 3, count))
    for t in range(count):
        ks = []
        while len(ks) < waves:
            k = rng.integers(-kmax, kmax + 1, 3)
            if 0 < k @ k <= kmax * kmax:
                ks.append(k)
        ks = np.array(ks, dtype=np.float64) * (np.pi / 2)
        amp = rng.normal(size=(waves, 3)) * np.exp(-np.sum(ks * ks, 1) / (2 * (np.pi * kmax / 4) ** 2))[:, None]
        ph = rng.uniform(0, 2 * np.pi, waves)
        arg = points @ ks.T + ph  # [n, waves]
        c = np.cos(arg)
        # A = sum_w amp_w sin(k.x+ph): dA_c/dx_j = sum_w amp_wc k_wj cos(.)
        J = np.einsum("nw,wc,wj->ncj", c, amp, ks)
        out[:, :, t] = np.stack((J[:, 2, 1] - J[:, 1, 2], J[:, 0, 2] - J[:, 2, 0],
                                 J[:, 1, 0] - J[:, 0, 1]), 1)
    return out


# --------------------------------------------------------------------------
# next row: Galerkin reduced-order operators on the network basis
# (experiments/periodic/paper_rollout.py:37-66,166-186;
#  experiments/metaball/operators.py:20-60,137-257)
# --------------------------------------------------------------------------
def periodic_reduced_basis(gu, gv, gw, w: float, rank_tol: float = 1e-12):
    """Velocity-normalised columns, FP64 mass w(gu^T gu + gv^T gv) and
    stiffness w gw^T gw, the full-rank gate, generalized eigh and the
    transformed fields (paper_rollout.py:166-186).  Returns (gu, gv, gw, vals)."""
    import scipy.linalg as la
    norm = np.sqrt(w * np.sum(gu * gu + gv * gv, 0))
    if not np.all(norm > 0):
        raise FloatingPointError("zero-velocity mode")
    gu, gv, gw = gu / norm, gv / norm, gw / norm
    mass = w * (gu.T @ gu + gv.T @ gv)
    stiff = w * (gw.T @ gw)
    mh, sh = 0.5 * (mass + mass.T), 0.5 * (stiff + stiff.T)
    ev = np.linalg.eigvalsh(mh)
    if int(np.sum(ev > ev[-1] * rank_tol)) != len(mh):
        raise ArithmeticError("Full-K rank gate")
    vals, T = la.eigh(sh, mh)
    return gu @ T, gv @ T, gw @ T, vals


def rollout_rhs(gu, gv, gw, a, w: float):
    """Galerkin right-hand side of the 2D vorticity-form Euler nonlinearity
    (paper_rollout.py:37-41): u = gu a, v = gv a, omega = gw a,
    rhs = gu^T (w omega v) - gv^T (w omega u)."""
    u, v, om = gu @ a, gv @ a, gw @ a
    return gu.T @ (w * om * v) - gv.T @ (w * om * u)


def etdrk4_coefficients(linear, h: float, roots: int = 32):
    """Exponential time-differencing RK4 coefficients of the diagonal linear
    part `linear` at step h, by the contour mean over `roots` points
    (Kassam & Trefethen 2005; paper_rollout.py:49-56)."""
    e = np.exp(h * linear)
    e2 = np.exp(0.5 * h * linear)
    r = np.exp(1j * np.pi * (np.arange(1, roots + 1) - 0.5) / roots)
    lr = h * linear[:, None] + r
    q = h * np.real(np.mean((np.exp(lr / 2) - 1) / lr, -1))
    f1 = h * np.real(np.mean((-4 - lr + np.exp(lr) * (4 - 3 * lr + lr * lr)) / lr ** 3, -1))
    f2 = h * np.real(np.mean((2 + lr + np.exp(lr) * (-2 + lr)) / lr ** 3, -1))
    f3 = h * np.real(np.mean((-4 - 3 * lr - lr * lr + np.exp(lr) * (4 - lr)) / lr ** 3, -1))
    return e, e2, q, f1, f2, f3


def etdrk4(rhs, linear, c, h: float, steps: int):
    """`steps` ETDRK4 steps of c' = linear * c + rhs(c) (paper_rollout.py:57-64)."""
    e, e2, q, f1, f2, f3 = etdrk4_coefficients(linear, h)
    for _ in range(steps):
        nv = rhs(c)
        a = e2 * c + q * nv
        na = rhs(a)
        b = e2 * c + q * na
        nb = rhs(b)
        z = e2 * a + q * (2 * nb - nv)
        c = e * c + f1 * nv + 2 * f2 * (na + nb) + f3 * rhs(z)
    return c


class SpectralGrid:
    """Periodic n^3 spectral operators of the metaball operator derivation
    (operators.py:20-60): Leray projection with 2/3 dealiasing, curl,
    viscous and normal-penalty forces, FP64 FFTs over the last three axes of
    [3, n, n, n] fields (x fastest)."""

    def __init__(self, normal, chi, n: int, viscosity: float = 2.5e-4, penalty_time: float = 1e-3):
        k = 2 * np.pi * np.fft.fftfreq(n, d=1 / n)
        self.kz, self.ky, self.kx = k[:, None, None], k[None, :, None], k[None, None, :]
        self.k2 = self.kx ** 2 + self.ky ** 2 + self.kz ** 2
        self.safe = self.k2.copy()
        self.safe[0, 0, 0] = 1
        m = np.fft.fftfreq(n, d=1 / n)
        ok = np.abs(m) <= n // 3
        self.dealias = ok[:, None, None] & ok[None, :, None] & ok[None, None, :]
        self.normal = np.asarray(normal, dtype=np.float64).transpose(3, 0, 1, 2)
        self.chi = np.asarray(chi, dtype=np.float64)
        self.n, self.nu, self.eta = n, viscosity, penalty_time

    def fft(self, u):
        return np.fft.fftn(u, axes=(1, 2, 3))

    def real(self, v):
        return np.fft.ifftn(v, axes=(1, 2, 3)).real

    def project(self, v):
        dot = (self.kx * v[0] + self.ky * v[1] + self.kz * v[2]) / self.safe
        r = np.stack((v[0] - self.kx * dot, v[1] - self.ky * dot, v[2] - self.kz * dot)) * self.dealias
        r[:, 0, 0, 0] = 0
        return r

    def curl(self, v):
        return np.stack((1j * (self.ky * v[2] - self.kz * v[1]), 1j * (self.kz * v[0] - self.kx * v[2]),
                         1j * (self.kx * v[1] - self.ky * v[0])))

    def derived(self, u):
        """(projected, curl, linear force) of one [3, n, n, n] field."""
        v = self.fft(u)
        viscous = self.project(-self.nu * self.k2 * v)
        penalty = self.project(self.fft(-self.chi[None] / self.eta * np.sum(u * self.normal, 0)[None] * self.normal))
        return self.real(self.project(v)), self.real(self.curl(v)), self.real(viscous + penalty)


def derive_fields(B, flat, grid: SpectralGrid):
    """Per mode k: the basis field B[:, :, k] at the fluid points `flat`
    (flat indices of the n^3 grid, x fastest), zero-extended to the grid ->
    (projected, curl, linear) back at the fluid points: [N, 9, K]
    (operators.py:137-177)."""
    N, _, K = B.shape
    n = grid.n
    out = np.empty((N, 9, K))
    for k in range(K):
        u = np.zeros((n ** 3, 3))
        u[flat] = B[:, :, k]
        u = u.reshape(n, n, n, 3).transpose(3, 0, 1, 2)
        parts = grid.derived(u)
        out[:, :, k] = np.concatenate([x.transpose(1, 2, 3, 0).reshape(-1, 3)[flat] for x in parts], 1)
    return out


def assemble_operators(B, D, dV: float):
    """Full-mass M = dV B^T B, linear A = dV B^T L and the packed convective
    tensor Q[i, (j<=k)] = dV (q[j,i,k] + q[k,i,j]) (j != k) or dV q[j,i,j],
    q[j,i,k] = sum_p PB_i . (B_j x curlB_k) (operators.py:199-257; pairs in
    numpy.triu_indices(K) order).  B: [N, 3, K]; D: [N, 9, K] = (projected,
    curl, linear)."""
    N, _, K = B.shape
    b = B.reshape(N * 3, K)
    M = dV * b.T @ b
    A = dV * b.T @ D[:, 6:9, :].reshape(N * 3, K)
    P, C = D[:, 0:3, :], D[:, 3:6, :]
    q = np.zeros((K, K, K))
    for a, bb, c, sign in ((0, 1, 2, 1), (0, 2, 1, -1), (1, 2, 0, 1), (1, 0, 2, -1), (2, 0, 1, 1), (2, 1, 0, -1)):
        q += sign * np.einsum("pj,pi,pk->jik", B[:, bb, :], P[:, a, :], C[:, c, :], optimize=True)
    jj, kk = np.triu_indices(K)
    Q = q[jj, :, kk].T.copy()
    diff = jj != kk
    Q[:, diff] += q[kk[diff], :, jj[diff]].T
    return M, A, dV * Q


# --------------------------------------------------------------------------
# next row: per-mode SIREN networks (HB2P:1590-1880; NCC:111-185)
# --------------------------------------------------------------------------
def mode_nets(Ws, bs, omega_first: float, omega_hidden: float) -> list:
    """One sine Net per mode from mode-stacked weights W_j [K, out, in],
    b_j [K, out] (omega_first on the first layer, omega_hidden after)."""
    L = len(Ws) - 1
    om = [omega_first] + [omega_hidden] * (L - 1)
    return [Net([W[k].astype(np.float64) for W in Ws], [b[k].astype(np.float64) for b in bs], "sin", om)
            for k in range(Ws[0].shape[0])]


def mode_fields(nets: list, x0: Jet, dim: int, stiffness: str = "vorticity"):
    """Field matrices of per-mode networks: column k is mode k's lifted
    fields (point-major rows as field_rows).  Returns (U, D, tapes)."""
    Us, Ds, tapes = [], [], []
    for net in nets:
        out, tape = forward(net, x0, record=True)
        tapes.append(tape)
        if dim == 2 and stiffness == "stream":
            u, d = lift2d_stream(out)       # raw psi (the peel centers it)
            Us.append(u)
            Ds.append(field_rows(d))
            continue
        if dim == 2:
            u, d = lift2d_hessian(out) if stiffness == "hessian" else lift2d(out)
        else:
            u, d = lift3d(out)
        Us.append(field_rows(u))
        Ds.append(field_rows(d) if d.ndim > 2 else d)
    return np.concatenate(Us, 1), np.concatenate(Ds, 1), tapes


def mode_reverse(nets: list, tapes: list, U: np.ndarray, D: np.ndarray, w: float, dM: np.ndarray,
                 dS: np.ndarray, dim: int, stiffness: str = "vorticity"):
    """Per-mode parameter gradients of <dM, w U^T U> + <dS, w D^T D>:
    mode k's field cotangents are columns k of 2w U dM, 2w D dS, pulled back
    through its lift and network.  Returns ([gW_j [K, out, in]], [gb_j])."""
    n = len(tapes[0][1].value)
    gU, gD = field_adjoint(U, w, dM), field_adjoint(D, w, dS)
    gWs, gbs = [], []
    for k, (net, tape) in enumerate(zip(nets, tapes)):
        if dim == 2:
            gu = gU[:, k].reshape(n, 2, 1)
            if stiffness == "hessian":
                adj = lift2d_hessian_reverse(gu, gD[:, k].reshape(n, 3, 1))
            else:
                adj = lift2d_reverse(gu, gD[:, k].reshape(n, 1))
        else:
            adj = lift3d_reverse(gU[:, k].reshape(n, 3, 1), gD[:, k].reshape(n, 3, 3, 1))
        gW, gb = backward(net, tape, adj)
        gWs.append(gW)
        gbs.append(gb)
    L = len(gWs[0])
    return [np.stack([g[j] for g in gWs]) for j in range(L)], [np.stack([g[j] for g in gbs]) for j in range(L)]


def mode_adam(Ws, bs, gWs, gbs, mWs, vWs, mbs, vbs, t: int, lr: float, clip: float = 8.0):
    """Adam with a per-mode norm clip (NCC:156-171): the norm of mode k runs
    over all of its weights and biases; mode k's gradient is scaled by
    min(1, clip / max(norm_k, 1e-30)); nonfinite or zero norms raise."""
    K = Ws[0].shape[0]
    norms = np.sqrt(sum(np.sum(g.reshape(K, -1) ** 2, 1) for g in [*gWs, *gbs]))
    if not np.all(np.isfinite(norms) & (norms > 0)):
        raise FloatingPointError("Invalid parameter gradients; no optimizer update")
    scales = np.minimum(1.0, clip / np.maximum(norms, 1e-30))
    out = []
    for p, m, v, g in zip([*Ws, *bs], [*mWs, *mbs], [*vWs, *vbs], [*gWs, *gbs]):
        g = g * scales.reshape((K,) + (1,) * (g.ndim - 1))
        m = 0.9 * m + 0.1 * g
        v = 0.999 * v + 0.001 * g * g
        p = p - lr * (m / (1 - 0.9 ** t)) / (np.sqrt(v / (1 - 0.999 ** t)) + 1e-8)
        out.append((p, m, v))
    return out, norms
