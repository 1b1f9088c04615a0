"""Shared-trunk eigenfluid network on MI355X: parameters, analytic-jet
forward, Gram assembly and the handwritten reverse pass.

Host-side mirror of the reference backend's model surface
(experiments/metaball/_backend/flow/cases/metaball/native_cupy_core.py:111-185
``Model`` with ``W``/``b`` lists, ``forward``/``backward``/``adam``/
``save``/``load``; velocity_trace_jet2.py ``forward``/``backward``/
``velocity``/``velocity_grad``; velocity/core.py ``fields``/``grams``).  All
arithmetic on the path runs in libnebhip.so kernels; torch only owns device
memory and streams, plus the small FP64 K x K products and factorizations
(rocBLAS / rocSOLVER).

Network: features (periodic sin/cos jets) -> L hidden layers
act(omega_l (x W_l^T + b_l)) -> linear output with K (2D stream function)
or 3K (3D vector potential, rows 3k+c) outputs.  Jets are second order;
channel order value, grads, Hessian pairs (0,0),(0,1),(1,1) [2D] /
(0,0),(0,1),(0,2),(1,1),(1,2),(2,2) [3D].
"""
from __future__ import annotations

import math
from contextlib import nullcontext
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _hip as H

# BF16X3: host-side code of the split-precision forward (bf16 hi / lo planes
# of every jet and weight, three bf16 MFMA products per term); its tape is
# a bf16 tape (the hi planes) with the lo planes beside it
BF16X3 = 3
PRECISIONS = {"bf16": H.BF16, "f32": H.F32, "f64": H.F64, "bf16x3": BF16X3}
_STORE = {H.BF16: torch.bfloat16, H.F32: torch.float32, H.F64: torch.float64}
_FIELD = {H.BF16: torch.float32, H.F32: torch.float32, H.F64: torch.float64}
_CROSS = {H.BF16: H.CROSS_BF16, H.F32: H.CROSS_F32, H.F64: H.CROSS_F64}
GRAM_MODES = {"f32": H.CROSS_F32, "bf16": H.CROSS_BF16, "bf16x3": H.CROSS_BF16X3, "f64": H.CROSS_F64}
_DT = {torch.float32: H.F32, torch.bfloat16: H.BF16, torch.float64: H.F64}


def _ld(k: int) -> int:
    return (k + 7) // 8 * 8


def _stream():
    return torch.cuda.current_stream().cuda_stream


def channels(dim: int) -> int:
    return 1 + dim + dim * (dim + 1) // 2


@dataclass
class NetConfig:
    dim: int = 2
    frequencies: int = 8
    lo: tuple = (0.0, 0.0)
    hi: tuple = (2 * math.pi, 2 * math.pi)
    widths: tuple = (128, 128, 128)
    modes: int = 64
    activation: str = "tanh"
    omegas: tuple | None = None
    # stiffness fields of the 2D lift: "vorticity" (omega = -lap psi, one row
    # per point) or "hessian" (psi_xx, sqrt2 psi_xy, psi_yy: the H1 seminorm
    # of the velocity with the pair metric (1, 2, 1) of the causal velocity
    # peel, HB2P:2935-3060, :3022), both with the velocity (psi_y, -psi_x) as
    # mass fields; or "stream" (the Dirichlet streamfunction peel,
    # HB2P:3383-3474: mass rows psi centered over the points, stiffness rows
    # (psi_x, psi_y)); 3D always uses curl A and grad u (three and nine rows)
    stiffness: str = "vorticity"
    # output-layer reverse of split-precision tapes: "factored" (the K x K
    # contraction in FP64 on the master weights, then width-long split GEMMs
    # on the tape's hi / lo planes; reads no fields), "fields" (the fused
    # lift transpose over the bf16 hi-plane fields), or "auto": factored in
    # 3D, whose ill-conditioned mass cancels the bf16 field form's terms,
    # fields in 2D, where the field form is faster
    split_reverse: str = "auto"

    @property
    def features(self) -> int:
        return 2 * self.frequencies * self.dim

    @property
    def ncomp(self) -> int:
        return 1 if self.dim == 2 else 3

    @property
    def sizes(self) -> list[int]:
        return [self.features, *self.widths, self.ncomp * self.modes]

    @property
    def lift(self) -> int:
        """Lift code of the output-layer entry points (their dim argument)."""
        if self.dim == 3:
            return H.LIFT_3D_CURL
        if self.stiffness == "hessian":
            return H.LIFT_2D_HESSIAN
        if self.stiffness == "stream":
            return H.LIFT_2D_STREAM
        if self.stiffness != "vorticity":
            raise ValueError(f"unknown 2D stiffness {self.stiffness!r}")
        return H.LIFT_2D_VORTICITY

    @property
    def field_rows(self) -> tuple:
        """(mass rows, stiffness rows) per point of the lifted fields."""
        if self.dim == 3:
            return 3, 9
        return {"hessian": (2, 3), "stream": (1, 2)}.get(self.stiffness, (2, 1))

    @property
    def act_code(self) -> int:
        return {"tanh": H.TANH, "sin": H.SIN}[self.activation]

    def omega(self, l: int) -> float:
        return 1.0 if self.omegas is None else float(self.omegas[l])


class Workspace:
    """Grow-only scratch buffer for split reductions (the library allocates
    nothing itself)."""

    def __init__(self, device):
        self.device = device
        self.buf = torch.empty(0, dtype=torch.uint8, device=device)

    def get(self, nbytes: int) -> torch.Tensor:
        nbytes = max(int(nbytes), 256 * 8)
        if self.buf.numel() < nbytes:
            self.buf = torch.empty(nbytes, dtype=torch.uint8, device=self.device)
        return self.buf


_WS: dict = {}
_SIDE: dict = {}


def workspace(device) -> Workspace:
    """Scratch buffer of `device` for the current stream: kernels queued on
    different streams never share one."""
    key = (str(device), torch.cuda.current_stream(device).cuda_stream)
    if key not in _WS:
        _WS[key] = Workspace(device)
    return _WS[key]


def side_stream(device) -> torch.cuda.Stream:
    """Second stream of `device` (default priority) for work that overlaps
    the current stream: the weight gradients beside the jet reverse."""
    key = str(device)
    if key not in _SIDE:
        _SIDE[key] = torch.cuda.Stream(device=device)
    return _SIDE[key]


class EigenfluidNet:
    """Mode-shared trunk with K (2D) or 3K (3D) linear outputs.

    Parameters live in one flat master tensor (fp32, or fp64 for FP64 runs)
    with per-layer views ``W[l]`` [out, in] and ``b[l]`` [out]; Adam moments
    ``m``/``v`` mirror it (the reference's ``mW``/``vW``/``mb``/``vb``,
    NCC:125-128).  Initialization: Xavier-normal, std sqrt(2/(fan_in+fan_out)),
    drawn layer by layer as ``default_rng(seed).normal(size=(out, in))``, zero
    biases.
    """

    def __init__(self, cfg: NetConfig, seed: int = 0, device="cuda", master_dtype=torch.float32):
        self.cfg = cfg
        self.device = torch.device(device)
        self.master_dtype = master_dtype
        sizes = cfg.sizes
        rng = np.random.default_rng(seed)
        host = []
        self._shapes = []
        for fi, fo in zip(sizes[:-1], sizes[1:]):
            host.append((rng.normal(size=(fo, fi)) * math.sqrt(2.0 / (fi + fo))).ravel())
            host.append(np.zeros(fo))
            self._shapes.append((fo, fi))
        flat = np.concatenate(host)
        self.params = torch.tensor(flat, dtype=master_dtype, device=self.device)
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.epoch = 0
        self._views()
        self._staged = {}
        self._split_cache = None
        self.version = 0

    # -- parameter views -------------------------------------------------
    def _views(self):
        self.W, self.b, self._offsets = [], [], []
        off = 0
        for fo, fi in self._shapes:
            self.W.append(self.params[off:off + fo * fi].view(fo, fi))
            wo = off
            off += fo * fi
            self.b.append(self.params[off:off + fo])
            self._offsets.append((wo, off))
            off += fo

    def split(self, flat: torch.Tensor):
        """Views of a flat parameter-shaped tensor as ([W...], [b...])."""
        Ws, bs = [], []
        for (fo, fi), (wo, bo) in zip(self._shapes, self._offsets):
            Ws.append(flat[wo:wo + fo * fi].view(fo, fi))
            bs.append(flat[bo:bo + fo])
        return Ws, bs

    def load_arrays(self, W, b):
        with torch.no_grad():
            for dst, src in zip(self.W + self.b, list(W) + list(b)):
                dst.copy_(torch.as_tensor(np.asarray(src), dtype=self.master_dtype))
        self.touch()

    def touch(self):
        """Mark master weights as changed (drops staged copies)."""
        self.version += 1
        self._staged.clear()

    # -- staging to the MFMA operand precision ----------------------------
    def _stage_bf16(self):
        """Both bf16 staging caches in one launch (neb_stage_net): the split
        planes W hi / lo [out][ld_in] of every layer (staged_split) and, for
        staged(BF16), W = the hi planes and W^T [in][ld_out] of the hidden
        layers."""
        mdt = _DT[self.master_dtype]
        L = len(self.W)
        his, los, wts = [], [], []
        off, rows, cols, ld, ldt = [], [], [], [], []
        for l, W in enumerate(self.W):
            fo, fi = W.shape
            his.append(torch.empty((fo, _ld(fi)), dtype=torch.bfloat16, device=self.device))
            los.append(torch.empty_like(his[-1]))
            wts.append(torch.empty((fi, _ld(fo)), dtype=torch.bfloat16, device=self.device) if l < L - 1 else None)
            off.append(self._offsets[l][0])
            rows.append(fo)
            cols.append(fi)
            ld.append(_ld(fi))
            ldt.append(_ld(fo) if l < L - 1 else 0)
        ptrs = lambda ts: H.ptr_array([t.data_ptr() if t is not None else None for t in ts])  # noqa: E731
        H.call("neb_stage_net", mdt, self.params.data_ptr(), L, H.i64_array(off), H.i64_array(rows), H.i64_array(cols),
               ptrs(his), ptrs(los), H.i64_array(ld), ptrs(wts), H.i64_array(ldt), _stream())
        bias = [b.to(torch.float32).contiguous() for b in self.b]
        self._split_cache = (self.version, {"W": his, "Wlo": los, "b": bias})
        self._staged[(H.BF16, self.version)] = {"W": his, "WT": wts[:-1], "b": bias}

    def staged(self, prec: int):
        key = (prec, self.version)
        if key in self._staged:
            return self._staged[key]
        if prec == H.BF16:
            self._stage_bf16()
            return self._staged[key]
        st = _STORE[prec]
        mdt = _DT[self.master_dtype]
        math_dt = torch.float64 if prec == H.F64 else torch.float32
        out = {"W": [], "WT": [], "b": []}
        s = _stream()
        for l, (W, b) in enumerate(zip(self.W, self.b)):
            fo, fi = W.shape
            ws = torch.empty((fo, _ld(fi)), dtype=st, device=self.device)
            H.call("neb_stage", mdt, prec, W.data_ptr(), fo, fi, fi, ws.data_ptr(), ws.shape[1], fo, ws.shape[1],
                   0, 1.0, s)
            out["W"].append(ws)
            if l < len(self.W) - 1:
                wt = torch.empty((fi, _ld(fo)), dtype=st, device=self.device)
                H.call("neb_stage", mdt, prec, W.data_ptr(), fo, fi, fi, wt.data_ptr(), wt.shape[1], fi,
                       wt.shape[1], 1, 1.0, s)
                out["WT"].append(wt)
            out["b"].append(b.to(math_dt).contiguous())
        self._staged[key] = out
        return out

    def staged_split(self):
        """Master weights as bf16 hi / lo planes ([out][ld_in] each; hi is the
        bf16 staging of staged(BF16)) and fp32 biases, for the split-precision
        forward."""
        if self._split_cache is None or self._split_cache[0] != self.version:
            self._stage_bf16()
        return self._split_cache[1]

    # -- optimizer ---------------------------------------------------------
    def adam(self, grads: torch.Tensor, lr: float, clip: float = 8.0, gate: torch.Tensor | None = None,
             on_gate=None):
        """Adam with global-norm clip (NCC:156-171; one parameter group for
        the shared trunk).  Nonfinite or zero gradients raise
        FloatingPointError before any update.

        The gates are evaluated on the device (neb_adam_gated): the update is
        queued without a host synchronisation and writes nothing when a gate
        fails; one synchronisation afterwards reads the gradient norm and
        `gate` (an optional device int32 flag of an earlier gate of the same
        step, nonzero = failed, e.g. SpectralSolver.gate) and raises.  A
        failed `gate` calls on_gate(value) first, which raises the gate's own
        error.  Returns the gradient norm."""
        dt = _DT[self.master_dtype]
        ws = workspace(self.device).get(256 * 8)
        norm2 = torch.empty(1, dtype=torch.float64, device=self.device)
        s = _stream()
        H.call("neb_sumsq", dt, grads.data_ptr(), grads.numel(), norm2.data_ptr(), ws.data_ptr(), ws.numel(), s)
        t = self.epoch + 1
        if gate is not None and (gate.dtype != torch.int32 or gate.numel() != 1 or gate.device != norm2.device):
            raise ValueError("gate must be a one-element int32 tensor on the parameters' device")
        H.call("neb_adam_gated", dt, self.params.data_ptr(), grads.data_ptr(), self.m.data_ptr(), self.v.data_ptr(),
               self.params.numel(), lr, clip, 0.9, 0.999, 1e-8, t, norm2.data_ptr(),
               gate.data_ptr() if gate is not None else None, s)
        if gate is None:
            n2, gv = float(norm2.item()), 0
        else:
            n2, gv = torch.cat((norm2, gate.to(torch.float64))).tolist()
            gv = int(gv)
        if gv != 0:
            if on_gate is not None:
                on_gate(gv)
            raise FloatingPointError(f"step gate {gv} failed; no optimizer update")
        norm = math.sqrt(n2) if n2 >= 0 else float("nan")
        if not math.isfinite(norm) or norm <= 0:
            raise FloatingPointError("Invalid parameter gradients; no optimizer update")
        self.epoch = t
        self.touch()
        return norm

    # -- checkpoint (NCC:173-185 key layout) ---------------------------------
    def save(self, path):
        arrays = {}
        mW, mb = self.split(self.m)
        vW, vb = self.split(self.v)
        for name, lst in (("W", self.W), ("b", self.b), ("mW", mW), ("vW", vW), ("mb", mb), ("vb", vb)):
            for j, a in enumerate(lst):
                arrays[f"{name}_{j}"] = a.detach().cpu().numpy()
        arrays["epoch"] = np.asarray(self.epoch)
        tmp = str(path) + ".tmp.npz"
        np.savez(tmp, **arrays)
        import os
        os.replace(tmp, path)

    def load(self, path):
        with np.load(path, allow_pickle=False) as z:
            self.epoch = int(z["epoch"])
            mW, mb = self.split(self.m)
            vW, vb = self.split(self.v)
            with torch.no_grad():
                for name, lst in (("W", self.W), ("b", self.b), ("mW", mW), ("vW", vW), ("mb", mb), ("vb", vb)):
                    for j, a in enumerate(lst):
                        src = z[f"{name}_{j}"]
                        if src.shape != tuple(a.shape):
                            raise ValueError(f"{name}_{j}: shape {src.shape} != {tuple(a.shape)}")
                        a.copy_(torch.as_tensor(src, dtype=self.master_dtype))
        self.touch()


# ---------------------------------------------------------------------------
# forward
# ---------------------------------------------------------------------------
@dataclass
class Tape:
    prec: int
    n: int
    inputs: list = field(default_factory=list)   # packed layer inputs [C, n, ld] (X0, Y1, ..., Y_L)
    pre: list = field(default_factory=list)      # packed scaled pre-activations Z_l
    inputs_lo: list | None = None   # split-precision tapes: lo planes of the layer inputs
    pre_lo: list | None = None      # (and of the pre-activations, when kept)


def tape_bytes(tape: Tape) -> int:
    """Device bytes held by a jet tape."""
    ts = [*tape.inputs, *tape.pre, *(tape.inputs_lo or []), *(tape.pre_lo or [])]
    return sum(t.numel() * t.element_size() for t in ts if t is not None)


def trunk_forward(net: EigenfluidNet, x: torch.Tensor, prec: int) -> Tape:
    """Feature jets and all hidden layers; keeps the jet tape."""
    cfg = net.cfg
    if x.ndim != 2 or x.shape[1] != cfg.dim:
        raise ValueError(f"points must be [n, {cfg.dim}]")
    if prec == BF16X3:
        return _trunk_forward_x3(net, x)
    xdt = torch.float64 if prec == H.F64 else torch.float32
    x = x.to(device=net.device, dtype=xdt).contiguous()
    n = x.shape[0]
    C = channels(cfg.dim)
    st = _STORE[prec]
    stg = net.staged(prec)
    s = _stream()
    lo = H.dbl_array([float(v) for v in cfg.lo])
    hi = H.dbl_array([float(v) for v in cfg.hi])
    F = cfg.features
    X0 = torch.empty((C, n, _ld(F)), dtype=st, device=net.device)
    import ctypes
    H.call("neb_feature_jets", prec, cfg.dim, n, x.data_ptr(), cfg.frequencies, ctypes.addressof(lo),
           ctypes.addressof(hi), X0.data_ptr(), X0.shape[2], s)
    tape = Tape(prec, n, [X0], [])
    cur, k = X0, F
    for l, width in enumerate(cfg.widths):
        Z = torch.empty((C, n, _ld(width)), dtype=st, device=net.device)
        Y = torch.empty_like(Z)
        H.call("neb_jet_layer_forward", prec, cfg.dim, cfg.act_code, n, k, width, cur.data_ptr(), cur.shape[2],
               stg["W"][l].data_ptr(), stg["W"][l].shape[1], stg["b"][l].data_ptr(), cfg.omega(l), Z.data_ptr(),
               Y.data_ptr(), Y.shape[2], s)
        tape.pre.append(Z)
        tape.inputs.append(Y)
        cur, k = Y, width
    return tape


def _trunk_forward_x3(net: EigenfluidNet, x: torch.Tensor) -> Tape:
    """Split-precision trunk: fp32 feature jets split into bf16 hi / lo
    planes, then every hidden layer as neb_jet_layer_forward_x3 (three bf16
    MFMA products per term, Z and Y stored as hi / lo planes).  The returned
    tape is the bf16 tape of the hi planes, which the reverse pass reads, with
    the lo planes of the layer inputs in inputs_lo (the pre-activations keep
    their hi plane only: the reverse reads nothing else)."""
    import ctypes
    cfg = net.cfg
    x = x.to(device=net.device, dtype=torch.float32).contiguous()
    n = x.shape[0]
    C = channels(cfg.dim)
    s = _stream()
    lo = H.dbl_array([float(v) for v in cfg.lo])
    hi = H.dbl_array([float(v) for v in cfg.hi])
    F = cfg.features
    X0 = torch.empty((C, n, _ld(F)), dtype=torch.float32, device=net.device)
    H.call("neb_feature_jets", H.F32, cfg.dim, n, x.data_ptr(), cfg.frequencies, ctypes.addressof(lo),
           ctypes.addressof(hi), X0.data_ptr(), X0.shape[2], s)
    Xh = torch.empty(X0.shape, dtype=torch.bfloat16, device=net.device)
    Xl = torch.empty_like(Xh)
    H.call("neb_split_bf16", H.F32, X0.data_ptr(), C * n, F, X0.shape[2], Xh.data_ptr(), Xl.data_ptr(), Xh.shape[2],
           s)
    del X0
    stg = net.staged_split()
    tape = Tape(H.BF16, n, [Xh], [], [Xl], None)
    ch, cl, k = Xh, Xl, F
    for l, width in enumerate(cfg.widths):
        Zh = torch.empty((C, n, _ld(width)), dtype=torch.bfloat16, device=net.device)
        Yh, Yl = torch.empty_like(Zh), torch.empty_like(Zh)
        H.call("neb_jet_layer_forward_x3", cfg.dim, cfg.act_code, n, k, width, ch.data_ptr(), cl.data_ptr(),
               ch.shape[2], stg["W"][l].data_ptr(), stg["Wlo"][l].data_ptr(), stg["W"][l].shape[1],
               stg["b"][l].data_ptr(), cfg.omega(l), Zh.data_ptr(), None, Yh.data_ptr(), Yl.data_ptr(),
               Yh.shape[2], s)
        tape.pre.append(Zh)
        tape.inputs.append(Yh)
        tape.inputs_lo.append(Yl)
        ch, cl, k = Yh, Yl, width
    return tape


def convert_tape(tape: Tape, prec: int) -> Tape:
    """The same jet tape stored in another operand precision (one staging
    launch per tensor, rounding to nearest): an FP32 forward's tape feeds a
    BF16 reverse pass this way."""
    if prec == tape.prec:
        return tape
    st = _STORE[prec]
    s = _stream()

    def conv(t: torch.Tensor) -> torch.Tensor:
        out = torch.empty(t.shape, dtype=st, device=t.device)
        rows, cols = t.shape[0] * t.shape[1], t.shape[2]
        if rows:
            H.call("neb_stage", _DT[t.dtype], prec, t.data_ptr(), rows, cols, cols, out.data_ptr(), cols, rows, cols,
                   0, 1.0, s)
        return out

    return Tape(prec, tape.n, [conv(t) for t in tape.inputs], [conv(t) for t in tape.pre])


def envelope_jets(value: torch.Tensor, grad: torch.Tensor, hess: torch.Tensor, dtype=torch.float32) -> torch.Tensor:
    """Pack the jet of a scalar envelope per point (value [n], grad [n, d],
    hess [n, P] in the pair order) as the [C, n] operand of output_jets /
    backward_jets (fp32; fp64 for FP64 tapes)."""
    return torch.cat([value.reshape(1, -1), grad.T, hess.T], 0).to(dtype).contiguous()


def output_jets(net: EigenfluidNet, tape: Tape, envelope: torch.Tensor | None = None):
    """Output jets (value [n, nout], grad [n, nout, d], hess [n, nout, P])
    -- the analogue of the per-mode field jets of forward_stream_jet2
    (HB2P:1802-1880) / VJ2.forward (VJ2:65-76).  envelope: optional [C, n]
    jet of a per-point envelope e (envelope_jets): the outputs are then the
    products e * raw by the Jet2 product rule, applied in the output GEMM's
    epilogue (VJ2:47-50, NCC:94-95)."""
    cfg = net.cfg
    prec = tape.prec
    stg = net.staged(prec)
    Y = tape.inputs[-1]
    n, k = tape.n, cfg.widths[-1]
    nout = cfg.ncomp * cfg.modes
    C = channels(cfg.dim)
    out = torch.empty((C, n, _ld(nout)), dtype=_FIELD[prec], device=net.device)
    if envelope is not None:
        env = envelope.to(device=net.device, dtype=_FIELD[prec]).contiguous()
        if env.shape != (C, n):
            raise ValueError(f"envelope must be [{C}, {n}]")
        H.call("neb_jet_output_jets_enveloped", prec, cfg.dim, n, k, nout, Y.data_ptr(), Y.shape[2],
               stg["W"][-1].data_ptr(), stg["W"][-1].shape[1], stg["b"][-1].data_ptr(), env.data_ptr(),
               out.data_ptr(), out.shape[2], _stream())
    else:
        H.call("neb_jet_output_jets", prec, cfg.dim, n, k, nout, Y.data_ptr(), Y.shape[2], stg["W"][-1].data_ptr(),
               stg["W"][-1].shape[1], stg["b"][-1].data_ptr(), out.data_ptr(), out.shape[2], _stream())
    out = out[:, :, :nout]
    d = cfg.dim
    return out[0], out[1:1 + d].permute(1, 2, 0), out[1 + d:].permute(1, 2, 0)


@dataclass
class Fields:
    mass: torch.Tensor    # [n*cm, ldK] (view [:, :K]); 2D (u, v), 3D curl A
    stiff: torch.Tensor   # [n*cs, ldK]; 2D omega, 3D grad u (rows 3a+j)
    K: int
    layout: int = 0
    mass_lo: torch.Tensor | None = None   # split fields: mass/stiff hold the bf16 hi
    stiff_lo: torch.Tensor | None = None  # planes, these the lo planes

    @property
    def split(self) -> bool:
        return self.mass_lo is not None

    @property
    def U(self):
        return self.mass[:, :self.K]

    @property
    def D(self):
        return self.stiff[:, :self.K]


def output_fields(net: EigenfluidNet, tape: Tape, layout: int = 0, split: bool = False) -> Fields:
    """Output layer fused with the lift: the field matrices the Gram and the
    spectral objective consume (VCORE:71-77 ``fields``; ELC:35-38).

    split=True (bf16 tapes only) writes each fp32 field value x as two bf16
    planes hi = bf16(x), lo = bf16(x - hi): the operand of the split Gram
    (gram(..., lo=...)), and hi alone is the bf16 operand of the reverse."""
    cfg = net.cfg
    prec = tape.prec
    Y = tape.inputs[-1]
    n, k, K = tape.n, cfg.widths[-1], cfg.modes
    cm, cs = cfg.field_rows
    if tape.inputs_lo is not None:
        # split-precision tape: output layer + lift on split operands
        if not split:
            raise ValueError("a split-precision tape writes split fields (split=True)")
        sw = net.staged_split()
        planes = [torch.empty((r * n, _ld(K)), dtype=torch.bfloat16, device=net.device)
                  for r in (cm, cm, cs, cs)]
        H.call("neb_jet_output_fields_x3", cfg.lift, n, k, K, Y.data_ptr(), tape.inputs_lo[-1].data_ptr(), Y.shape[2],
               sw["W"][-1].data_ptr(), sw["Wlo"][-1].data_ptr(), sw["W"][-1].shape[1],
               *[t.data_ptr() for t in planes], planes[0].shape[1], layout, _stream())
        return _gauge(cfg, Fields(planes[0], planes[2], K, layout, planes[1], planes[3]))
    stg = net.staged(prec)
    if split:
        if prec != H.BF16:
            raise ValueError("split fields need a bf16 tape")
        planes = [torch.empty((r * n, _ld(K)), dtype=torch.bfloat16, device=net.device)
                  for r in (cm, cm, cs, cs)]
        H.call("neb_jet_output_fields_split", cfg.lift, n, k, K, Y.data_ptr(), Y.shape[2],
               stg["W"][-1].data_ptr(), stg["W"][-1].shape[1], *[t.data_ptr() for t in planes],
               planes[0].shape[1], layout, _stream())
        return _gauge(cfg, Fields(planes[0], planes[2], K, layout, planes[1], planes[3]))
    ft = _FIELD[prec]
    mass = torch.empty((cm * n, _ld(K)), dtype=ft, device=net.device)
    stiff = torch.empty((cs * n, _ld(K)), dtype=ft, device=net.device)
    H.call("neb_jet_output_fields", prec, cfg.lift, n, k, K, Y.data_ptr(), Y.shape[2], stg["W"][-1].data_ptr(),
           stg["W"][-1].shape[1], mass.data_ptr(), stiff.data_ptr(), mass.shape[1], layout, _stream())
    return _gauge(cfg, Fields(mass, stiff, K, layout))


def column_mean(x: torch.Tensor, rows: int, cols: int, lo: torch.Tensor | None = None) -> torch.Tensor:
    """FP64 column means of a [rows, ld] matrix (of hi + lo for split
    planes) by the library's split column sums."""
    ws = workspace(x.device).get(H.load().neb_colsum_workspace_size(rows, cols))
    out = torch.empty(cols, dtype=torch.float64, device=x.device)
    for beta, t in ((0.0, x), (1.0, lo)):
        if t is not None:
            H.call("neb_colsum", _DT[t.dtype], t.data_ptr(), rows, cols, t.stride(0), beta, out.data_ptr(),
                   ws.data_ptr(), ws.numel(), _stream())
    return out.div_(max(rows, 1))


def _gauge(cfg: NetConfig, f: Fields) -> Fields:
    """The streamfunction lift's periodic gauge: every mode's psi rows lose
    their mean over the points (HB2P:3475-3479), in place (hi + lo split
    again for split fields).  The Grams and the output-layer reverse then
    read the centered rows; their field cotangents 2w Psi_c dL/dM are
    mean-free already, which is the transpose of the centering
    (HB2P:3579-3583).  Other lifts pass through."""
    if cfg.dim != 2 or cfg.stiffness != "stream":
        return f
    n, K = f.mass.shape[0], f.K
    mu = column_mean(f.mass, n, K, f.mass_lo)
    H.call("neb_center_columns", _DT[f.mass.dtype], f.mass.data_ptr(),
           f.mass_lo.data_ptr() if f.mass_lo is not None else None, n, K, f.mass.stride(0), mu.data_ptr(), _stream())
    return f


# ---------------------------------------------------------------------------
# Gram
# ---------------------------------------------------------------------------
def gram(phi: torch.Tensor, K: int, w: float, mode: str = "bf16x3", out: torch.Tensor | None = None,
         beta: float = 0.0, roww: torch.Tensor | None = None, lo: torch.Tensor | None = None) -> torch.Tensor:
    """G = beta*G + w * Phi[:, :K]^T diag(roww) Phi[:, :K]  in FP64
    (VCORE:40-46 ``grams``; experiments/metaball/run.py:99-116).

    lo: the lo plane of split fields (phi is then the hi plane); only the
    "bf16x3" mode without row weights applies to split operands."""
    if phi.ndim != 2 or phi.stride(1) != 1:
        raise ValueError("Phi must be a row-major 2-D tensor")
    if out is None:
        out = torch.empty((K, K), dtype=torch.float64, device=phi.device)
        beta = 0.0
    m = GRAM_MODES[mode]
    rows = phi.shape[0]
    need = H.load().neb_cross_workspace_size(K, K, rows, 1)
    ws = workspace(phi.device).get(need)
    if lo is not None:
        if mode != "bf16x3" or roww is not None:
            raise ValueError("split operands take the bf16x3 mode without row weights")
        if lo.shape != phi.shape or lo.stride() != phi.stride() or phi.dtype != torch.bfloat16 \
                or lo.dtype != torch.bfloat16:
            raise ValueError("hi and lo planes must be bf16 tensors of one shape and stride")
        H.call("neb_gram_split", phi.data_ptr(), lo.data_ptr(), rows, K, phi.stride(0), float(w), float(beta),
               out.data_ptr(), out.stride(0), ws.data_ptr(), ws.numel(), _stream())
        return out
    rw = 0 if roww is None else roww.to(torch.float64).contiguous().data_ptr()
    H.call("neb_gram", m, phi.data_ptr(), rows, K, phi.stride(0), _DT[phi.dtype], float(w), rw, float(beta),
           out.data_ptr(), out.stride(0), ws.data_ptr(), ws.numel(), _stream())
    return out


def field_matrix(fields: Fields, which: str, dtype=torch.float32) -> torch.Tensor:
    """The mass ("mass") or stiffness ("stiff") field matrix [rows, K] as one
    fp32 (or fp64) tensor: split planes joined as hi + lo (neb_join_bf16),
    other fields converted only when their dtype differs."""
    hi, lo = (fields.mass, fields.mass_lo) if which == "mass" else (fields.stiff, fields.stiff_lo)
    K = fields.K
    if lo is None:
        return hi[:, :K] if hi.dtype == dtype else hi[:, :K].to(dtype)
    out = torch.empty((hi.shape[0], _ld(K)), dtype=dtype, device=hi.device)
    H.call("neb_join_bf16", _DT[dtype], hi.data_ptr(), lo.data_ptr(), hi.shape[0], K, hi.stride(0), out.data_ptr(),
           out.shape[1], _stream())
    return out[:, :K]


def split_fields_for(prec: str, gram_mode: str) -> bool:
    """Whether output_fields should write split planes: the bf16 tape with
    the split-precision Gram (results are bit-identical to fp32 fields), and
    always the split-precision forward."""
    return (prec == "bf16" and gram_mode == "bf16x3") or prec == "bf16x3"


def field_gram(fields: Fields, which: str, w: float, mode: str = "bf16x3", out: torch.Tensor | None = None,
               beta: float = 0.0) -> torch.Tensor:
    """FP64 Gram of the mass ("mass") or stiffness ("stiff") fields, split or not."""
    phi, lo = (fields.mass, fields.mass_lo) if which == "mass" else (fields.stiff, fields.stiff_lo)
    return gram(phi, fields.K, w, mode, out=out, beta=beta, lo=lo)


def field_grams(fields: Fields, w: float, mode: str = "bf16x3", M: torch.Tensor | None = None,
                S: torch.Tensor | None = None, beta: float = 0.0, mass_events=None):
    """(mass Gram, stiffness Gram) as field_gram, one after the other on the
    current stream (run side by side they contend for the same memory path
    and finish later).  mass_events: optional (start, end) events recorded
    around the mass Gram."""
    if mass_events is not None:
        mass_events[0].record()
    M = field_gram(fields, "mass", w, mode, out=M, beta=beta if M is not None else 0.0)
    if mass_events is not None:
        mass_events[1].record()
    S = field_gram(fields, "stiff", w, mode, out=S, beta=beta if S is not None else 0.0)
    return M, S


def cross(segments, m: int, n: int, mode: int, scale: float = 1.0, out=None, beta: float = 0.0,
          roww: torch.Tensor | None = None, block: int | None = None):
    """sum_s sign_s X_s^T Y_s over rows; segments = [(X, Y, rows, sign)] with
    X/Y row-major 2-D tensors (views allowed).  roww: optional FP64 weight
    per row (one segment only).  block: only the diagonal block x block
    blocks of the m x m result (neb_cross_blocks; the entries outside the
    tiles over them are left as they are)."""
    xs = [s[0] for s in segments]
    ys = [s[1] for s in segments]
    if out is None:
        out = torch.empty((m, n), dtype=torch.float64, device=xs[0].device)
        beta = 0.0
    rows = [int(s[2]) for s in segments]
    if block:
        if roww is not None or m != n:
            raise ValueError("a block-diagonal cross has a square result and no row weights")
        need = H.load().neb_cross_blocks_workspace_size(m, block, max(rows))
        ws = workspace(xs[0].device).get(need)
        H.call("neb_cross_blocks", mode, len(segments), H.ptr_array([t.data_ptr() for t in xs]),
               H.ptr_array([t.data_ptr() for t in ys]), H.i64_array([t.stride(0) for t in xs]),
               H.i64_array([t.stride(0) for t in ys]), H.i64_array(rows),
               H.dbl_array([float(s[3]) for s in segments]), _DT[xs[0].dtype], _DT[ys[0].dtype], m, block,
               float(scale), float(beta), out.data_ptr(), out.stride(0), ws.data_ptr(), ws.numel(), _stream())
        return out
    need = H.load().neb_cross_workspace_size(m, n, max(rows), 0)
    ws = workspace(xs[0].device).get(need)
    H.call("neb_cross", mode, len(segments), H.ptr_array([t.data_ptr() for t in xs]),
           H.ptr_array([t.data_ptr() for t in ys]), H.i64_array([t.stride(0) for t in xs]),
           H.i64_array([t.stride(0) for t in ys]), H.i64_array(rows), H.dbl_array([float(s[3]) for s in segments]),
           _DT[xs[0].dtype], _DT[ys[0].dtype], m, n, 0 if roww is None else roww.data_ptr(), 0, float(scale),
           float(beta), out.data_ptr(), out.stride(0), ws.data_ptr(), ws.numel(), _stream())
    return out


def cross_modes(segments, B: int, h: int, scale: float = 1.0, out=None, beta: float = 0.0):
    """Mode-diagonal crosses (neb_cross_modes): out[m, j] = beta out[m, j] +
    scale sum_s sign_s sum_r X_s[r, m] Y_s[r, m h + j], FP64 [B, h];
    segments = [(X, Y, rows, sign)] as for cross (up to 16 per call)."""
    xs = [s[0] for s in segments]
    ys = [s[1] for s in segments]
    if out is None:
        out = torch.empty((B, h), dtype=torch.float64, device=xs[0].device)
        beta = 0.0
    if not out.is_contiguous() or tuple(out.shape) != (B, h) or out.dtype != torch.float64:
        raise ValueError("cross_modes writes a contiguous FP64 [B, h] result")
    rows = [int(s[2]) for s in segments]
    ws = workspace(xs[0].device).get(H.load().neb_cross_modes_workspace_size(B, h, max(rows)))
    H.call("neb_cross_modes", len(segments), H.ptr_array([t.data_ptr() for t in xs]),
           H.ptr_array([t.data_ptr() for t in ys]), H.i64_array([t.stride(0) for t in xs]),
           H.i64_array([t.stride(0) for t in ys]), H.i64_array(rows), H.dbl_array([float(s[3]) for s in segments]),
           _DT[xs[0].dtype], _DT[ys[0].dtype], B, h, float(scale), float(beta), out.data_ptr(), ws.data_ptr(),
           ws.numel(), _stream())
    return out


def _f64_rows(A: torch.Tensor) -> torch.Tensor:
    """A as an FP64 matrix with unit column stride (no copy when it is one)."""
    A = A.to(torch.float64)
    return A if A.stride(1) == 1 else A.contiguous()


def sym_shift(A: torch.Tensor, shift: float = 0.0, out: torch.Tensor | None = None) -> torch.Tensor:
    """(A + A^T) / 2 + shift I in FP64 (one launch; out may be A)."""
    A = _f64_rows(A)
    K = A.shape[0]
    out = torch.empty((K, K), dtype=torch.float64, device=A.device) if out is None else out
    H.call("neb_sym_shift", K, A.data_ptr(), A.stride(0), float(shift), out.data_ptr(), out.stride(0), _stream())
    return out


def spectral_terms(M: torch.Tensor, S: torch.Tensor, X: torch.Tensor, info: torch.Tensor | None, lam: float,
                   ridge: float, cond_limit: float):
    """Objective terms around the inverse X of M_r = sym(M) + ridge I
    (neb_spectral_terms): symmetrises X in place and returns
    (loss, X, sym(S), gate) with the device scalar loss = <X, sym(S)> +
    lam ||M - I||^2 and the device int32 gate (0, first bad pivot, or -1 for
    the condition limit)."""
    M, S = _f64_rows(M), _f64_rows(S)
    K = M.shape[0]
    Ss = torch.empty((K, K), dtype=torch.float64, device=M.device)
    loss = torch.empty((), dtype=torch.float64, device=M.device)
    gate = torch.empty(1, dtype=torch.int32, device=M.device)
    ws = workspace(M.device).get(H.load().neb_spectral_workspace_size(K))
    H.call("neb_spectral_terms", K, M.data_ptr(), M.stride(0), S.data_ptr(), S.stride(0), X.data_ptr(), X.stride(0),
           info.data_ptr() if info is not None else None, float(lam), float(ridge), float(cond_limit),
           Ss.data_ptr(), K, loss.data_ptr(), gate.data_ptr(), ws.data_ptr(), ws.numel(), _stream())
    return loss, X, Ss, gate


def mass_adjoint(X: torch.Tensor, Ss: torch.Tensor, M: torch.Tensor, lam: float,
                 out: torch.Tensor | None = None) -> torch.Tensor:
    """dL/dM = -sym(X Ss X) + 2 lam (M - I) (two FP64 GEMMs + one launch),
    into `out` (a K x K view with unit column stride) when given."""
    M = _f64_rows(M)
    K = M.shape[0]
    C = torch.mm(torch.mm(X, Ss), X)
    out = C if out is None else out
    H.call("neb_spectral_adjoint", K, C.data_ptr(), K, M.data_ptr(), M.stride(0), float(lam), out.data_ptr(),
           out.stride(0), _stream())
    return out


def adjoint_pair(K: int, device) -> torch.Tensor:
    """[K, 2K] FP64 buffer holding dL/dM (columns :K) and dL/dS (columns K:)
    side by side: backward_fields then forms B = W^T [dL/dM | dL/dS] and the
    output-layer gradient [dL/dM | dL/dS] [P_m; P_s] as one GEMM each."""
    return torch.empty((K, 2 * K), dtype=torch.float64, device=device)


def _as_pair(dM: torch.Tensor, dS: torch.Tensor):
    """The [K, 2K] buffer of adjoint_pair when dM and dS are its two halves."""
    K = dM.shape[0]
    if (dM.dtype == dS.dtype == torch.float64 and dM.shape == dS.shape == (K, K) and dM.stride() == (2 * K, 1)
            and dS.stride() == (2 * K, 1) and dS.data_ptr() == dM.data_ptr() + 8 * K):
        return dM.as_strided((K, 2 * K), (2 * K, 1))
    return None


def _rows_view(t: torch.Tensor, start_row: int, step: int) -> torch.Tensor:
    """Every ``step``-th row of a row-major 2-D tensor starting at start_row."""
    return t.as_strided((t.shape[0] // step, t.shape[1]), (t.stride(0) * step, 1),
                        t.storage_offset() + start_row * t.stride(0))


# ---------------------------------------------------------------------------
# reverse pass
# ---------------------------------------------------------------------------
def lift_terms(cfg: NetConfig) -> tuple[list, list]:
    """The lift as signed terms (field row, jet channel, output-weight slice,
    sign) of the mass and the stiffness fields: field row a of a point is
    sum_t sign_t Y[channel_t] W_{slice_t}^T over the terms t with row a, Y
    the packed last hidden jets and W_s the output-weight rows 3k+s (3D) or
    all rows (2D).  2D: u = psi_y, v = -psi_x; omega = -(psi_xx + psi_yy)
    (ELC:35-40) or the Hessian rows (psi_xx, sqrt2 psi_xy, psi_yy)
    (HB2P:3022); 3D: u = curl A and grad u (VJ2:93-106)."""
    if cfg.dim == 3:
        from .lifts import curl_terms
        t = curl_terms()
        return ([(a, 1 + c, s, sg) for (a, c, s, sg) in t if a < 3],
                [(a - 3, 1 + c, s, sg) for (a, c, s, sg) in t if a >= 3])
    if cfg.stiffness == "stream":
        return [(0, 0, 0, 1.0)], [(0, 1, 0, 1.0), (1, 2, 0, 1.0)]
    mass = [(0, 2, 0, 1.0), (1, 1, 0, -1.0)]
    if cfg.stiffness == "hessian":
        return mass, [(0, 3, 0, 1.0), (1, 4, 0, math.sqrt(2.0)), (2, 5, 0, 1.0)]
    return mass, [(0, 3, 0, -1.0), (0, 5, 0, -1.0)]


class FactoredProducts:
    """The last hidden jets of a split-precision tape stacked side by side
    per point (hi / lo bf16 planes [n, nch * width]: the mass channels, then
    the stiffness channels) and, per field group, the lift-weighted channel
    Grams Ghat[s' width + x, s width + y] = sum S[c', c, s', s] Y_c'x^T Y_cy
    over the points (FP64; _factored_ghat); `event` marks their completion on
    the stream that queued them (output_products)."""

    def __init__(self, hi, lo, chans, Gh, event=None):
        self.hi, self.lo, self.chans, self.Gh, self.event = hi, lo, chans, Gh, event


def factored_reverse_ok(net: EigenfluidNet, tape: Tape) -> bool:
    """Whether the reverse of this tape runs factored (tape-only, field-free):
    split-precision tapes with a hidden width that is a multiple of 8, per
    NetConfig.split_reverse."""
    mode = net.cfg.split_reverse
    if mode not in ("auto", "factored", "fields"):
        raise ValueError(f"unknown split_reverse {mode!r}")
    if tape.inputs_lo is None or net.cfg.widths[-1] % 8:
        return False
    if net.cfg.dim == 2 and net.cfg.stiffness == "stream":
        # the centered streamfunction rows need the centered value channel,
        # which the factored algebra does not form
        if mode == "factored":
            raise ValueError("the streamfunction lift takes the field-form reverse (split_reverse='fields')")
        return False
    return net.cfg.dim == 3 if mode == "auto" else mode == "factored"


def factored_grams_ok(net: EigenfluidNet, tape: Tape) -> bool:
    """Whether the field Grams of this tape can come from channel Grams
    (factored_field_grams) and its reverse run factored with them: a
    split-precision tape, a hidden width that is a multiple of 8 and a lift
    whose rows are bias-free combinations of jet channels (every lift but the
    streamfunction lift, whose psi rows carry the output bias and are
    centered)."""
    if tape.inputs_lo is None or net.cfg.widths[-1] % 8:
        return False
    return not (net.cfg.dim == 2 and net.cfg.stiffness == "stream")


def factored_channels(cfg: NetConfig) -> tuple[list, list]:
    """The last hidden jet channels the mass and the stiffness fields read,
    in stacking order."""
    mterms, sterms = lift_terms(cfg)
    return sorted({t[1] for t in mterms}), sorted({t[1] for t in sterms})


def factored_products(net: EigenfluidNet, tape: Tape, grams: bool = True) -> FactoredProducts:
    """Stack the last hidden jet channels the lift reads (neb_stack_channels)
    and form their split-precision channel Grams (the products of the
    output-weight gradient, unweighted); grams=False stacks only (Gh None)."""
    cfg = net.cfg
    width, n = cfg.widths[-1], tape.n
    mch, sch = factored_channels(cfg)
    chans = mch + sch
    Y, Ylo = tape.inputs[-1], tape.inputs_lo[-1]
    ld = len(chans) * width
    hi = torch.empty((n, ld), dtype=torch.bfloat16, device=net.device)
    lo = torch.empty_like(hi)
    H.call("neb_stack_channels", Y.data_ptr(), Ylo.data_ptr(), n, Y.shape[2], width, len(chans), H.int_array(chans),
           hi.data_ptr(), lo.data_ptr(), ld, _stream())
    if not grams:
        return FactoredProducts(hi, lo, (mch, sch), None)
    km = len(mch) * width
    G = [gram(hi[:, :km], km, 1.0, "bf16x3", lo=lo[:, :km]), gram(hi[:, km:], ld - km, 1.0, "bf16x3", lo=lo[:, km:])]
    return FactoredProducts(hi, lo, (mch, sch), _factored_ghat(cfg, G))


def factored_field_grams(net: EigenfluidNet, Gh: list, w: float, M: torch.Tensor | None = None,
                         S: torch.Tensor | None = None):
    """The field Grams from lift-weighted channel Grams (FactoredProducts.Gh,
    summed over any point chunks): every mass and stiffness row of the
    factored lifts is sum_t sg_t Y_c(t) W_s(t)^T over the lift's terms
    (derivatives of the outputs, so no output bias), hence
      M = w Wcat Ghat_mass Wcat^T,   S = w Wcat Ghat_stiff Wcat^T
    with Wcat[k, s width + j] = W_out[ns k + s, j] (FP64; the K x K algebra
    of velocity/core.py:40-46 on [ns width]^2 products instead of [n comp, K]
    field matrices).  Writes into M / S when given; returns (M, S)."""
    cfg = net.cfg
    K, wd = cfg.modes, cfg.widths[-1]
    ns = 3 if cfg.dim == 3 else 1
    Wcat = net.W[-1].to(torch.float64).reshape(K, ns * wd)
    out = []
    for G, dst in zip(Gh, (M, S)):
        R = torch.mm(torch.mm(Wcat, G), Wcat.T)
        R = 0.5 * (R + R.T)
        if dst is None:
            out.append(R.mul_(w))
        else:
            out.append(dst.copy_(R).mul_(w))
    return out[0], out[1]


_FACT_S: dict = {}


def _factored_coeffs(cfg: NetConfig, device) -> list:
    """Per field group (mass, stiffness): S[i', i, s', s] = sum of sg_t sg_t'
    over the lift's term pairs (t, t') of one field row with channels
    (chans[i'], chans[i]) and output-weight slices (s', s)  (FP64, cached)."""
    key = (cfg.dim, cfg.stiffness, str(device))
    if key not in _FACT_S:
        ns = 3 if cfg.dim == 3 else 1
        out = []
        for terms, chans in zip(lift_terms(cfg), factored_channels(cfg)):
            idx = {c: i for i, c in enumerate(chans)}
            S = np.zeros((len(chans), len(chans), ns, ns))
            for (a, c, sl, sg) in terms:
                for (a2, c2, sl2, sg2) in terms:
                    if a2 == a:
                        S[idx[c2], idx[c], sl2, sl] += sg * sg2
            out.append(torch.tensor(S.reshape(len(chans) ** 2, ns * ns), dtype=torch.float64, device=device))
        _FACT_S[key] = out
    return _FACT_S[key]


def _factored_ghat(cfg: NetConfig, G: list) -> list:
    """Ghat_s's = sum_c'c S[c', c, s', s] G[c', c] per field group from the
    channel Grams G ([nch width]^2, FP64), as [ns width, ns width]."""
    wd = cfg.widths[-1]
    ns = 3 if cfg.dim == 3 else 1
    out = []
    for S2, Gg in zip(_factored_coeffs(cfg, G[0].device), G):
        nch = Gg.shape[0] // wd
        Gst = Gg.view(nch, wd, nch, wd).permute(0, 2, 1, 3).reshape(nch * nch, wd * wd)
        out.append(torch.mm(S2.T, Gst).view(ns, ns, wd, wd).permute(0, 2, 1, 3).reshape(ns * wd, ns * wd))
    return out


def _factored_weights(net: EigenfluidNet, pair: torch.Tensor, w: float) -> list:
    """The weight operands of the factored output-layer reverse, per field
    group X (mass with dL/dM, stiffness with dL/dS), in FP64 from the master
    output weights W_s (rows 3k+s in 3D) and the lift coefficients S
    (_factored_coeffs):
      Q_s's = W_s'^T dX^T W_s                       (one [ns width]^2 product)
      R[c', c] = 2w sum_s's S[c', c, s', s] Q_s's   [width, width] blocks
    returned transposed (R^T: the [features][k] operand of adj Y = Y_stack R)."""
    cfg = net.cfg
    K, wd = cfg.modes, cfg.widths[-1]
    ns = 3 if cfg.dim == 3 else 1
    Wcat = net.W[-1].to(torch.float64).reshape(K, ns * wd)        # Wcat[k, s wd + j] = W[ns k + s, j]
    RTs = []
    for S2, dX in zip(_factored_coeffs(cfg, Wcat.device), (pair[:, :K], pair[:, K:])):
        nch = int(round(math.sqrt(S2.shape[0])))
        Q = torch.mm(Wcat.T, torch.mm(dX.T, Wcat))                    # [s' wd + x, s wd + y]
        Qst = Q.view(ns, wd, ns, wd).permute(0, 2, 1, 3).reshape(ns * ns, wd * wd)
        Rq = torch.mm(S2, Qst).view(nch, nch, wd, wd)                 # [c', c, x, y]
        RTs.append((Rq.permute(1, 3, 0, 2).reshape(nch * wd, nch * wd) * (2.0 * w)).contiguous())
    return RTs


def _factored_output_grad(net: EigenfluidNet, fp: FactoredProducts, pair: torch.Tensor, w: float) -> torch.Tensor:
    """Output-weight gradient gW_s = 2w sum_X sum_s' dX W_s' Ghat^X_s's
    ([ncomp K, width], rows ncomp k + s)."""
    cfg = net.cfg
    K, wd = cfg.modes, cfg.widths[-1]
    ns = 3 if cfg.dim == 3 else 1
    Wcat = net.W[-1].to(torch.float64).reshape(K, ns * wd)
    gcat = torch.zeros((K, ns * wd), dtype=torch.float64, device=Wcat.device)
    for Gh, dX in zip(fp.Gh, (pair[:, :K], pair[:, K:])):
        gcat.addmm_(torch.mm(dX, Wcat), Gh, alpha=2.0 * w)
    return gcat.view(K, ns, wd).reshape(ns * K, wd)


def _factored_adjoint(net: EigenfluidNet, tape: Tape, fp: FactoredProducts, RTs: list) -> torch.Tensor:
    """adj Z_L = omega * act-jet^T(Z_L, adj Y): adj Y of each field group is
    the split GEMM Y_stack[:, group] R (neb_gemm_split with R^T staged as
    hi / lo planes), then the activation transpose (neb_jet_act_reverse)."""
    cfg = net.cfg
    width, n = cfg.widths[-1], tape.n
    C = channels(cfg.dim)
    s = _stream()
    ld = fp.hi.shape[1]
    AY = torch.empty((n, ld), dtype=torch.float32, device=net.device)
    c0 = 0
    for RT in RTs:
        kg = RT.shape[0]
        rhi = torch.empty((kg, _ld(kg)), dtype=torch.bfloat16, device=net.device)
        rlo = torch.empty_like(rhi)
        H.call("neb_split_bf16", H.F64, RT.data_ptr(), kg, kg, kg, rhi.data_ptr(), rlo.data_ptr(), rhi.shape[1], s)
        H.call("neb_gemm_split", n, kg, kg, fp.hi[:, c0:].data_ptr(), fp.lo[:, c0:].data_ptr(), ld, rhi.data_ptr(),
               rlo.data_ptr(), rhi.shape[1], AY[:, c0:].data_ptr(), ld, s)
        c0 += kg
    chans = fp.chans[0] + fp.chans[1]
    Z = tape.pre[-1]
    A = torch.empty((C, n, _ld(width)), dtype=_STORE[tape.prec], device=net.device)
    H.call("neb_jet_act_reverse", tape.prec, cfg.dim, cfg.act_code, n, width, AY.data_ptr(), ld, len(chans),
           H.int_array(chans), Z.data_ptr(), Z.shape[2], cfg.omega(len(cfg.widths) - 1), A.data_ptr(), A.shape[2], s)
    return A


def backward_fields(net: EigenfluidNet, tape: Tape, fields: Fields, dM: torch.Tensor, dS: torch.Tensor,
                    w: float, grads: torch.Tensor | None = None, symmetric: bool = False,
                    products: "OutputProducts | None" = None) -> torch.Tensor:
    """Parameter gradient of L(M, S) with M = w U^T U, S = w D^T D, given
    dL/dM and dL/dS (symmetric, FP64 K x K).  Fuses the field adjoint
    2w Phi dL/dG (VCORE:94-95) into the output-layer reverse, then runs the
    handwritten jet reverse through every layer (HB2P:2087-2139, VJ2:77-82).
    symmetric=True uses dM and dS as given, without symmetrising them: the
    field cotangents applied are then 2w U dM^T and 2w D dS^T, which for
    adjoints symmetric by construction (SpectralSolver's) is the gradient of
    L(M, S), and for the coefficient peel's pair [Cu^T/2 | Cd^T/2] its
    detached-prefix cotangents w U Cu, w D Cd; `products` are the
    output-layer products queued earlier by output_products (computed here
    when None).  Split-precision tapes with NetConfig.split_reverse
    "factored" (the 3D default) take the factored output-layer reverse
    (factored_products, _factored_weights, _factored_adjoint,
    _factored_output_grad: the K x K contraction in FP64, then width-long
    split GEMMs on the tape's hi / lo planes); it reads no fields, and
    `fields` may then be None.
    Returns the flat gradient in the master dtype."""
    cfg = net.cfg
    prec = tape.prec
    stg = net.staged(prec)
    d, C, n, K = cfg.dim, channels(cfg.dim), tape.n, cfg.modes
    L = len(cfg.widths)
    width = cfg.widths[-1]
    st = _STORE[prec]
    s = _stream()
    pair = _as_pair(dM, dS) if symmetric else None
    if pair is None:
        pair = adjoint_pair(K, net.device)
        if symmetric:
            pair[:, :K].copy_(dM)
            pair[:, K:].copy_(dS)
        else:
            sym_shift(dM, out=pair[:, :K])
            sym_shift(dS, out=pair[:, K:])
    if grads is None:
        grads = torch.empty_like(net.params)
    gW, gb = net.split(grads)
    mode = _CROSS[prec]
    # The weight gradients (cross products over points) only read the fields,
    # the tape and each layer's adjoint once it exists: they run on the side
    # stream, overlapping the jet reverse on the current stream, which waits
    # for them before returning.
    on_gpu = net.device.type == "cuda"
    main = torch.cuda.current_stream(net.device) if on_gpu else None
    side = side_stream(net.device) if on_gpu else None
    if on_gpu:
        side.wait_stream(main)
    # factored products handed in (the factored field Grams' stacked channels)
    # select the factored reverse whatever NetConfig.split_reverse says
    if isinstance(products, FactoredProducts) or factored_reverse_ok(net, tape):
        fp = products if isinstance(products, FactoredProducts) else None
        if fp is None:
            fp = factored_products(net, tape)
        elif on_gpu and fp.event is not None:
            main.wait_event(fp.event)
            for t in (fp.hi, fp.lo, *(fp.Gh or [])):
                t.record_stream(main)
        RTs = _factored_weights(net, pair, w)
        A = _factored_adjoint(net, tape, fp, RTs)
        if on_gpu:
            side.wait_stream(main)
            for t in (pair, *(fp.Gh or [])):     # read on the side stream after this frame lets them go
                t.record_stream(side)
        with (torch.cuda.stream(side) if on_gpu else nullcontext()):
            if fp.Gh is None:      # stacked channels only: the caller adds the output-weight gradient
                gW[-1].zero_()
            else:
                gW[-1].copy_(_factored_output_grad(net, fp, pair, w))
            gb[-1].zero_()
        del fp, RTs
    else:
        if fields is None or fields.layout != 0:
            raise ValueError("backward_fields expects point-major field matrices")
        Wout = net.W[-1].to(torch.float64)               # [ncomp*K, width]
        if d == 2:
            Bp = torch.mm(Wout.T, pair)                  # [width, 2K] = [W^T dL/dM | W^T dL/dS]
        else:
            Bp = torch.cat([torch.mm(Wout[c::3].T, pair) for c in range(3)], 0)  # [3*width, 2K]
        # 2w B in the operand precision, zero-padded to the row stride (one launch each)
        bm = torch.empty((Bp.shape[0], _ld(K)), dtype=st, device=net.device)
        bs = torch.empty_like(bm)
        for half, dst in ((0, bm), (1, bs)):
            H.call("neb_stage", H.F64, prec, Bp.data_ptr() + 8 * K * half, Bp.shape[0], K, Bp.stride(0),
                   dst.data_ptr(), dst.shape[1], dst.shape[0], dst.shape[1], 0, 2.0 * w, s)
        A = torch.empty((C, n, _ld(width)), dtype=st, device=net.device)
        Z = tape.pre[-1]
        H.call("neb_jet_output_backward", prec, cfg.lift, cfg.act_code, n, width, K, fields.mass.data_ptr(),
               fields.stiff.data_ptr(), fields.mass.shape[1], _DT[fields.mass.dtype], fields.layout, bm.data_ptr(),
               bs.data_ptr(), bm.shape[1], Z.data_ptr(), Z.shape[2], cfg.omega(L - 1), A.data_ptr(), A.shape[2], s)
        with (torch.cuda.stream(side) if on_gpu else nullcontext()):
            if products is None:
                P = _products(net, tape, fields)
            else:
                if on_gpu and products.event is not None:
                    side.wait_event(products.event)
                P = products.P
            _output_layer_grads(d, K, P, pair, w, gW, gb)

    for l in range(L - 1, -1, -1):
        X = tape.inputs[l]
        k = X.shape[2] if l > 0 else cfg.features
        wl = cfg.widths[l]
        if on_gpu:
            side.wait_stream(main)       # A (layer l's adjoint) is complete
            A.record_stream(side)
        with (torch.cuda.stream(side) if on_gpu else nullcontext()):
            A2 = A.view(C * n, A.shape[2])
            X2 = X.view(C * n, X.shape[2])
            G = cross([(A2, X2, C * n, 1.0)], wl, k, mode)
            gW[l].copy_(G)
            cw = workspace(net.device)
            ws = cw.get(H.load().neb_colsum_workspace_size(n, wl))
            db = torch.empty(wl, dtype=torch.float64, device=net.device)
            H.call("neb_colsum", _DT[st], A[0].data_ptr(), n, wl, A.shape[2], 0.0, db.data_ptr(), ws.data_ptr(),
                   ws.numel(), _stream())
            gb[l].copy_(db)
        if l > 0:
            Zp = tape.pre[l - 1]
            An = torch.empty((C, n, _ld(k)), dtype=st, device=net.device)
            WT = stg["WT"][l]
            H.call("neb_jet_layer_backward", prec, d, cfg.act_code, n, k, wl, A.data_ptr(), A.shape[2],
                   WT.data_ptr(), WT.shape[1], Zp.data_ptr(), Zp.shape[2], cfg.omega(l - 1), An.data_ptr(),
                   An.shape[2], s)
            A = An
    if on_gpu:
        main.wait_stream(side)
    return grads


def _output_wt(net: EigenfluidNet, prec: int) -> torch.Tensor:
    """The output weights transposed, [width][ld(nout)] in the operand
    precision (cached per parameter version)."""
    key = ("out_wt", prec, net.version)
    if key not in net._staged:
        W = net.W[-1]
        fo, fi = W.shape
        wt = torch.empty((fi, _ld(fo)), dtype=_STORE[prec], device=net.device)
        H.call("neb_stage", _DT[net.master_dtype], prec, W.data_ptr(), fo, fi, fi, wt.data_ptr(), wt.shape[1], fi,
               wt.shape[1], 1, 1.0, _stream())
        net._staged[key] = wt
    return net._staged[key]


def backward_jets(net: EigenfluidNet, tape: Tape, adj_value: torch.Tensor, adj_grad: torch.Tensor,
                  adj_hess: torch.Tensor, envelope: torch.Tensor | None = None) -> torch.Tensor:
    """Parameter gradient of <adj, output jets> for a cotangent of the output
    jets (value [n, nout], grad [n, nout, d], hess [n, nout, P]): the
    analogue of VJ2.backward (VJ2:77-82) -- the envelope product's transpose
    (neb_envelope_reverse) when an envelope is given, the output layer's
    weight / bias gradient (cross product over all channels, VJ2:28-30) and
    its jet transpose fused with the last activation's
    (neb_jet_layer_backward on the output weights), then the hidden-layer
    reverse as in backward_fields.  Returns the flat gradient."""
    cfg = net.cfg
    prec = tape.prec
    d, C, n = cfg.dim, channels(cfg.dim), tape.n
    nout = cfg.ncomp * cfg.modes
    width = cfg.widths[-1]
    L = len(cfg.widths)
    st = _STORE[prec]
    s = _stream()
    adj = torch.cat([adj_value.reshape(1, n, nout), adj_grad.permute(2, 0, 1), adj_hess.permute(2, 0, 1)], 0)
    adj = adj.to(device=net.device, dtype=_FIELD[prec]).contiguous()
    A_out = torch.empty((C, n, _ld(nout)), dtype=st, device=net.device)
    if envelope is not None:
        env = envelope.to(device=net.device, dtype=_FIELD[prec]).contiguous()
        if env.shape != (C, n):
            raise ValueError(f"envelope must be [{C}, {n}]")
        H.call("neb_envelope_reverse", prec, d, n, nout, env.data_ptr(), adj.data_ptr(), nout, A_out.data_ptr(),
               A_out.shape[2], s)
    elif C * n:
        H.call("neb_stage", _DT[adj.dtype], prec, adj.data_ptr(), C * n, nout, nout, A_out.data_ptr(),
               A_out.shape[2], C * n, A_out.shape[2], 0, 1.0, s)
    grads = torch.empty_like(net.params)
    gW, gb = net.split(grads)
    mode = _CROSS[prec]
    Ylast = tape.inputs[-1]
    gW[-1].copy_(cross([(A_out.view(C * n, -1), Ylast.view(C * n, -1), C * n, 1.0)], nout, width, mode))
    ws = workspace(net.device).get(H.load().neb_colsum_workspace_size(n, nout))
    db = torch.empty(nout, dtype=torch.float64, device=net.device)
    H.call("neb_colsum", _DT[st], A_out[0].data_ptr(), n, nout, A_out.shape[2], 0.0, db.data_ptr(), ws.data_ptr(),
           ws.numel(), s)
    gb[-1].copy_(db)
    wt = _output_wt(net, prec)
    A = torch.empty((C, n, _ld(width)), dtype=st, device=net.device)
    H.call("neb_jet_layer_backward", prec, d, cfg.act_code, n, width, nout, A_out.data_ptr(), A_out.shape[2],
           wt.data_ptr(), wt.shape[1], tape.pre[-1].data_ptr(), tape.pre[-1].shape[2], cfg.omega(L - 1),
           A.data_ptr(), A.shape[2], s)
    stg = net.staged(prec)
    for l in range(L - 1, -1, -1):
        X = tape.inputs[l]
        k = X.shape[2] if l > 0 else cfg.features
        wl = cfg.widths[l]
        gW[l].copy_(cross([(A.view(C * n, A.shape[2]), X.view(C * n, X.shape[2]), C * n, 1.0)], wl, k, mode))
        ws = workspace(net.device).get(H.load().neb_colsum_workspace_size(n, wl))
        db = torch.empty(wl, dtype=torch.float64, device=net.device)
        H.call("neb_colsum", _DT[st], A[0].data_ptr(), n, wl, A.shape[2], 0.0, db.data_ptr(), ws.data_ptr(),
               ws.numel(), s)
        gb[l].copy_(db)
        if l > 0:
            An = torch.empty((C, n, _ld(k)), dtype=st, device=net.device)
            H.call("neb_jet_layer_backward", prec, d, cfg.act_code, n, k, wl, A.data_ptr(), A.shape[2],
                   stg["WT"][l].data_ptr(), stg["WT"][l].shape[1], tape.pre[l - 1].data_ptr(),
                   tape.pre[l - 1].shape[2], cfg.omega(l - 1), An.data_ptr(), An.shape[2], s)
            A = An
    return grads


def _products(net: EigenfluidNet, tape: Tape, fields: Fields) -> list:
    """[P_m; P_s] ([2K, width] FP64; one per output component slice in 3D):
    the field x last-hidden-jet cross products over points of the
    output-layer gradient (VCORE:94-95 through the lift).

    2D: the fields are linear in the output weights W (u = Y_y W^T,
    v = -Y_x W^T, omega = -(Y_xx + Y_yy) W^T), so
        P_m = u^T Y_y - v^T Y_x           = W (Y_x^T Y_x + Y_y^T Y_y)
        P_s = -omega^T (Y_xx + Y_yy)      = W (Y_xx + Y_yy)^T (Y_xx + Y_yy)
    -- 256 x 256 Grams of the last hidden jets over the points (three cross
    launches reading only those jets, a fifth of the bytes of the K-wide
    field products) times the master W in FP64.  3D keeps the field
    products: its 27 channel-pair Grams would read as many bytes."""
    cfg = net.cfg
    d, n, K = cfg.dim, tape.n, cfg.modes
    width = cfg.widths[-1]
    mode = _CROSS[tape.prec]
    Ylast = tape.inputs[-1]
    ch = lambda t, c: t[c]  # noqa: E731  packed channel c as [n, ld]
    U, Dg = fields.mass, fields.stiff
    if fields.layout != 0:
        raise ValueError("the output-layer products expect point-major field matrices")
    if d == 2 and cfg.stiffness == "stream":
        # psi_c = (Y_v - 1 ybar^T) W^T (centered), grad psi = (Y_x, Y_y) W^T:
        #   P_m = psi_c^T Y_v = W (Y_v^T Y_v - n ybar ybar^T),  P_s = W (Y_x^T Y_x + Y_y^T Y_y)
        # (the mean-free Gram cancels the square of the value channel's mean:
        # split tapes take it from the split Gram of hi + lo, whose error stays
        # ~1e-7 of Y_v^T Y_v rather than the bf16 hi plane's 2^-9)
        Ylo = tape.inputs_lo[-1] if tape.inputs_lo is not None else None
        if Ylo is not None:
            Hm = gram(ch(Ylast, 0), width, 1.0, "bf16x3", lo=ch(Ylo, 0))
        else:
            Hm = cross([(ch(Ylast, 0), ch(Ylast, 0), n, 1.0)], width, width, mode)
        yb = column_mean(ch(Ylast, 0), n, width, None if Ylo is None else ch(Ylo, 0))
        Hm.sub_(torch.outer(yb, yb), alpha=float(n))
        Hs = cross([(ch(Ylast, 1), ch(Ylast, 1), n, 1.0), (ch(Ylast, 2), ch(Ylast, 2), n, 1.0)], width, width, mode)
        Wd = net.W[-1].to(torch.float64)
        P = torch.empty((2 * K, width), dtype=torch.float64, device=U.device)
        torch.mm(Wd, Hm, out=P[:K])
        torch.mm(Wd, Hs, out=P[K:])
        return [P]
    if d == 2:
        Hm = cross([(ch(Ylast, 1), ch(Ylast, 1), n, 1.0), (ch(Ylast, 2), ch(Ylast, 2), n, 1.0)], width, width, mode)
        if cfg.stiffness == "hessian":
            # rows (psi_xx, sqrt2 psi_xy, psi_yy): P_s = W (Y_xx^T Y_xx + 2 Y_xy^T Y_xy + Y_yy^T Y_yy)
            Hs = cross([(ch(Ylast, 3), ch(Ylast, 3), n, 1.0), (ch(Ylast, 4), ch(Ylast, 4), n, 1.0),
                        (ch(Ylast, 4), ch(Ylast, 4), n, 1.0), (ch(Ylast, 5), ch(Ylast, 5), n, 1.0)],
                       width, width, mode)
        else:
            Hs = cross([(ch(Ylast, 3), ch(Ylast, 3), n, 1.0), (ch(Ylast, 5), ch(Ylast, 5), n, 1.0)], width, width,
                       mode)
            G35 = cross([(ch(Ylast, 3), ch(Ylast, 5), n, 1.0)], width, width, mode)
            Hs += G35
            Hs += G35.T
        Wd = net.W[-1].to(torch.float64)
        P = torch.empty((2 * K, width), dtype=torch.float64, device=U.device)
        torch.mm(Wd, Hm, out=P[:K])
        torch.mm(Wd, Hs, out=P[K:])
        return [P]
    from .lifts import curl_terms
    comps = [_rows_view(U, a, 3) for a in range(3)]
    grads3 = [_rows_view(Dg, c, 9) for c in range(9)]
    out = []
    for sl in range(3):
        segm = [(comps[a], ch(Ylast, 1 + c), n, sg) for (a, c, w_s, sg) in curl_terms() if a < 3 and w_s == sl]
        segs = [(grads3[a - 3], ch(Ylast, 1 + c), n, sg) for (a, c, w_s, sg) in curl_terms()
                if a >= 3 and w_s == sl]
        P = torch.empty((2 * K, width), dtype=torch.float64, device=U.device)
        cross(segm, K, width, mode, out=P[:K])
        for i in range(0, len(segs), 8):
            cross(segs[i:i + 8], K, width, mode, out=P[K:], beta=1.0 if i else 0.0)
        out.append(P)
    return out


class OutputProducts:
    """Output-layer cross products queued on the side stream (see
    output_products); `event` marks their completion on that stream."""

    def __init__(self, P: list, event):
        self.P, self.event = P, event


_CUS: dict = {}


def _device_cus(dev) -> int:
    key = str(dev)
    if key not in _CUS:
        _CUS[key] = torch.cuda.get_device_properties(dev).multi_processor_count
    return _CUS[key]


def output_products(net: EigenfluidNet, tape: Tape, fields: Fields | None):
    """Queue the output-layer products on the side stream: [P_m; P_s]
    (OutputProducts), or for a split-precision tape the stacked last hidden
    jets and their channel Grams (FactoredProducts).  They depend on the
    fields and the last hidden jets only, not on the objective, so they run
    while the FP64 objective (a chain of small, latency-bound launches) holds
    the current stream; pass the result to backward_fields."""
    dev = net.device
    fact = factored_reverse_ok(net, tape)
    if dev.type != "cuda":
        return factored_products(net, tape) if fact else OutputProducts(_products(net, tape, fields), None)
    main, side = torch.cuda.current_stream(dev), side_stream(dev)
    side.wait_stream(main)
    if fact:
        H.call("neb_set_cross_occupancy", max(1, _device_cus(dev) // 2))
        try:
            with torch.cuda.stream(side):
                fp = factored_products(net, tape)
                fp.event = torch.cuda.Event()
                fp.event.record(side)
        finally:
            H.call("neb_set_cross_occupancy", 0)
        for t in (tape.inputs[-1], tape.inputs_lo[-1]):
            t.record_stream(side)
        return fp
    # the products' launches keep to half the CUs' worth of workgroups: the
    # objective's sweep launches (one workgroup per 64 x 64 tile, 136 at
    # K = 1024) then find whole CUs free.  With one per CU the dispatcher
    # doubles some up, and the first sweep launch waited for the products'
    # kernel to end (196 us instead of 29 us).
    H.call("neb_set_cross_occupancy", max(1, _device_cus(dev) // 2))
    try:
        with torch.cuda.stream(side):
            P = _products(net, tape, fields)
            ev = torch.cuda.Event()
            ev.record(side)
    finally:
        H.call("neb_set_cross_occupancy", 0)
    for t in (fields.mass, fields.stiff, tape.inputs[-1]):
        t.record_stream(side)
    return OutputProducts(P, ev)


def _output_layer_grads(d, K, P, pair, w, gW, gb):
    """Output-layer weight gradient 2w (dL/dM P_m + dL/dS P_s) as one GEMM
    [dL/dM | dL/dS] [P_m; P_s] (the adjoints are symmetric)."""
    if d == 2:
        gW[-1].copy_(torch.addmm(P[0][:K], pair, P[0], beta=0.0, alpha=2.0 * w))
    else:
        for sl in range(3):
            gW[-1][sl::3].copy_(torch.addmm(P[sl][:K], pair, P[sl], beta=0.0, alpha=2.0 * w))
    gb[-1].zero_()
