"""Per-mode SIREN eigenfluid networks: the reference's own architecture.

Each of the K modes is its own small sine network (features -> L hidden
layers sin(omega (x W^T + b)) -> 1 stream function (2D) or 3 vector-potential
components (3D)), with omega_first on the first layer and omega_hidden after
(HB2P:1590-1880 HardcodedModel / HardcodedMode; NCC:111-185 Model; 2D:
hidden 32 x 4, omegas 8 / 30; 3D: hidden 128 x 5, omegas 40 / 36).

On the GPU a block of B modes is one grouped GEMM per layer
(neb_stage_mode_block): the first layer stacks the block's weights by rows
(every mode reads the same features), the hidden and output layers are
block-diagonal (each mode reads its own B*hidden-wide slice).  The block is
then a jet trunk of width B*hidden on the fused row-GEMM jets (every
precision, split operands included), lifts and Grams of neb_amd.network;
its block-diagonal hidden layers run on the _blocks entry points, whose
tiles contract only their own mode's inputs (the zeros off the diagonal are
not read or multiplied), and the weight gradients of those layers are
block-diagonal crosses.  The field matrices of all blocks are written side
by side into the K columns of one [n*comp, K] matrix; the Grams, the
objective (trace, or the reference's coefficient / velocity peel) and the
FP64 algebra are those of the shared trunk.  The reverse forms the field
cotangents of all modes once, then takes each block's tape -- kept from the
fields pass as far as HBM holds the tapes, else recomputed as in the
reference's phase-3 recompute (HB2P:3191-3230) -- and the block's columns of
the cotangents through its output layer, lift and hidden layers.

Parameters are stored mode-stacked, W_j [K, out, in] and b_j [K, out] (the
NCC:173-185 checkpoint layout; arrays() / from_arrays() give the 2D
HardcodedModel key layout of HB2P:1596-1622, 1987-2033).  Adam clips each
mode's gradient by its own norm (NCC:156-171).  With the 2D streamfunction
lift the blocks run on first-order jets (value and gradient channels, the
Jet1 stream of HB2P:1716-1800: the jet kernels' NEB_DIM_2D_JET1 code), half
the channels of the second-order tapes; output_jets evaluates second-order
jets, and forward_stream_jet1 drops their Hessian channels.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import torch

from . import _hip as H
from . import network as nw
from . import spectral as sp


@dataclass
class ModeConfig:
    dim: int = 2
    frequencies: int = 8
    lo: tuple = (0.0, 0.0)
    hi: tuple = (2 * math.pi, 2 * math.pi)
    hidden: int = 32
    hidden_layers: int = 4
    modes: int = 64
    omega_first: float = 8.0
    omega_hidden: float = 30.0
    block: int = 16
    stiffness: str = "vorticity"

    @property
    def features(self) -> int:
        return 2 * self.frequencies * self.dim

    @property
    def ncomp(self) -> int:
        return 1 if self.dim == 2 else 3

    @property
    def sizes(self) -> list[int]:
        return [self.features] + [self.hidden] * self.hidden_layers + [self.ncomp]

    def omega(self, l: int) -> float:
        return self.omega_first if l == 0 else self.omega_hidden

    @property
    def jet_order(self) -> int:
        """1 for the 2D streamfunction lift, whose fields and reverse read
        only the value and gradient channels (the reference evaluates it on
        the Jet1 stream, HB2P:1716-1800); 2 otherwise."""
        return 1 if self.dim == 2 and self.stiffness == "stream" else 2

    def jet_channels(self, order: int | None = None) -> int:
        return 1 + self.dim if (order or self.jet_order) == 1 else nw.channels(self.dim)

    def jet_dim(self, order: int | None = None) -> int:
        """The dim code of the jet-layer kernels for these jets."""
        return H.DIM_2D_JET1 if (order or self.jet_order) == 1 else self.dim

    def lift_code(self, order: int | None = None) -> int:
        """The lift code of the output-layer kernels for these jets."""
        lift = self.block_net_config(8).lift
        return H.LIFT_2D_STREAM_JET1 if (order or self.jet_order) == 1 else lift

    def block_net_config(self, B: int) -> nw.NetConfig:
        """The trunk configuration one block of B modes presents to the jet
        kernels (width B*hidden, B output modes)."""
        return nw.NetConfig(dim=self.dim, frequencies=self.frequencies, lo=self.lo, hi=self.hi,
                            widths=(B * self.hidden,) * self.hidden_layers, modes=B, activation="sin",
                            omegas=tuple(self.omega(l) for l in range(self.hidden_layers)),
                            stiffness=self.stiffness)


# Bytes of block tapes a training step keeps between its forward and its
# reverse, and the device memory left free beside them for the reverse of
# one block (HBM is 288 GB on MI355X: the 2D K = 1024 configuration's tapes
# are 19 GB; the 3D 5 x 128 one's 245 GB stay as far as they fit, the
# remaining blocks are recomputed).
TAPE_BUDGET = 256 << 30
TAPE_HEADROOM = 32 << 30


def siren_init(cfg: ModeConfig, seed: int):
    """Mode-stacked SIREN weights (NCC:113-124): mode i draws from its own
    PCG64(seed + 7919 (i + 1)) stream, layer by layer W then b, uniform in
    +-1/fan_in on the first layer and +-sqrt(6/fan_in)/omega_hidden after,
    rounded to fp32.  Returns ([W_j [K, out, in]], [b_j [K, out]])."""
    gens = [np.random.Generator(np.random.PCG64(seed + 7919 * (i + 1))) for i in range(cfg.modes)]
    sizes = cfg.sizes
    Ws, bs = [], []
    for j, (fi, fo) in enumerate(zip(sizes[:-1], sizes[1:])):
        bound = 1.0 / fi if j == 0 else math.sqrt(6.0 / fi) / cfg.omega_hidden
        Ws.append(np.stack([r.uniform(-bound, bound, (fo, fi)).astype("f4") for r in gens]))
        bs.append(np.stack([r.uniform(-bound, bound, fo).astype("f4") for r in gens]))
    return Ws, bs


class ModeNet:
    """K per-mode sine networks with mode-stacked parameters (see module
    doc).  params is one flat master tensor (fp32, or fp64 for FP64 runs),
    layer by layer W_j [K, out, in] then b_j [K, out]; m / v are the Adam
    moments in the same layout."""

    def __init__(self, cfg: ModeConfig, seed: int = 0, device="cuda", master_dtype=torch.float32):
        if cfg.block % 8 or cfg.block < 8:
            raise ValueError("the mode block must be a multiple of 8 (16-byte aligned field columns)")
        self.cfg = cfg
        self.device = torch.device(device)
        self.master_dtype = master_dtype
        Ws, bs = siren_init(cfg, seed)
        self._shapes = [(W.shape[1], W.shape[2]) for W in Ws]
        flat = np.concatenate([np.concatenate([W.ravel(), b.ravel()]) for W, b in zip(Ws, bs)])
        self.params = torch.tensor(flat, dtype=master_dtype, device=self.device)
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.epoch = 0
        self.version = 0
        self._cache = {}
        self._stage_bufs = {}
        self.W, self.b = self.split(self.params)

    def split(self, flat: torch.Tensor):
        """Views of a parameter-shaped flat tensor as ([W_j], [b_j])."""
        K = self.cfg.modes
        Ws, bs, off = [], [], 0
        for fo, fi in self._shapes:
            Ws.append(flat[off:off + K * fo * fi].view(K, fo, fi))
            off += K * fo * fi
            bs.append(flat[off:off + K * fo].view(K, fo))
            off += K * fo
        return Ws, bs

    def touch(self):
        self.version += 1
        self._cache.clear()

    # -- staging of one mode block ------------------------------------------
    def block_weights(self, m0: int, B: int, prec: str) -> dict:
        """The block's operands: W (and Wlo for split precision) per layer
        as [rows][ld] in the operand precision, WT (the transposes, for the
        reverse) and the biases in the math precision.  The operand buffers
        of a block stay allocated across updates: they start zero, and an
        update rewrites the first layer whole and only the diagonal blocks
        of the block-diagonal ones (neb_stage_mode_block diag = 2)."""
        key = (m0, B, prec, self.version)
        if key in self._cache:
            return self._cache[key]
        split = prec == "bf16x3"
        p = H.BF16 if split else nw.PRECISIONS[prec]
        st = nw._STORE[p]
        mdt = nw._DT[self.master_dtype]
        s = nw._stream()
        bufs = self._stage_bufs.get((m0, B, prec))
        if bufs is None:
            bufs = []
            for j, W in enumerate(self.W):
                fo, fi = W.shape[1:]
                rows, cols = B * fo, (B * fi if j > 0 else fi)
                hi = torch.zeros((rows, nw._ld(cols)), dtype=st, device=self.device)
                bufs.append((hi, torch.zeros_like(hi) if split else None,
                             torch.zeros((cols, nw._ld(rows)), dtype=st, device=self.device)))
            self._stage_bufs[(m0, B, prec)] = bufs
        out = {"W": [], "Wlo": [], "WT": [], "b": []}
        for j, (W, b) in enumerate(zip(self.W, self.b)):
            fo, fi = W.shape[1:]
            diag = 0 if j == 0 else 2
            hi, lo, wt = bufs[j]
            H.call("neb_stage_mode_block", mdt, W.data_ptr(), self.cfg.modes, fo, fi, m0, B, diag, p, hi.data_ptr(),
                   lo.data_ptr() if lo is not None else None, hi.shape[1], wt.data_ptr(), wt.shape[1], s)
            out["W"].append(hi)
            out["Wlo"].append(lo)
            out["WT"].append(wt)
            out["b"].append(b[m0:m0 + B].reshape(-1).to(torch.float64 if p == H.F64 else torch.float32))
        self._cache[key] = out
        return out

    # -- forward -----------------------------------------------------------
    def features(self, x: torch.Tensor, prec: str, order: int | None = None):
        """Feature jets shared by every mode: (hi, lo) planes for split
        precision, else (X0, None); first-order jets (order 1, the default
        for the streamfunction lift) are the first 1 + dim channels."""
        import ctypes
        cfg = self.cfg
        split = prec == "bf16x3"
        p = H.F32 if split else nw.PRECISIONS[prec]
        xdt = torch.float64 if p == H.F64 else torch.float32
        x = x.to(device=self.device, dtype=xdt).contiguous()
        n, F = x.shape[0], cfg.features
        lo = H.dbl_array([float(v) for v in cfg.lo])
        hi = H.dbl_array([float(v) for v in cfg.hi])
        X0 = torch.empty((nw.channels(cfg.dim), n, nw._ld(F)), dtype=nw._STORE[p] if not split else torch.float32,
                         device=self.device)
        H.call("neb_feature_jets", p, cfg.dim, n, x.data_ptr(), cfg.frequencies, ctypes.addressof(lo),
               ctypes.addressof(hi), X0.data_ptr(), X0.shape[2], nw._stream())
        C = cfg.jet_channels(order)
        X0 = X0[:C]
        if not split:
            return X0, None
        Xh = torch.empty(X0.shape, dtype=torch.bfloat16, device=self.device)
        Xl = torch.empty_like(Xh)
        H.call("neb_split_bf16", H.F32, X0.data_ptr(), C * n, F, X0.shape[2], Xh.data_ptr(), Xl.data_ptr(),
               Xh.shape[2], nw._stream())
        return Xh, Xl

    def block_tape(self, feats, m0: int, B: int, prec: str) -> nw.Tape:
        """Hidden jets of modes m0 .. m0+B-1 (the block as one trunk of width
        B*hidden), of the order of the feature jets `feats`."""
        cfg = self.cfg
        bc = cfg.block_net_config(B)
        stg = self.block_weights(m0, B, prec)
        X0, X0lo = feats
        n = X0.shape[1]
        C = X0.shape[0]
        kd = H.DIM_2D_JET1 if C == 1 + cfg.dim else cfg.dim
        s = nw._stream()
        split = prec == "bf16x3"
        p = H.BF16 if split else nw.PRECISIONS[prec]
        st = nw._STORE[p]
        tape = nw.Tape(p, n, [X0], [], [X0lo] if split else None, None)
        cur, cur_lo, k = X0, X0lo, cfg.features
        h = cfg.hidden
        for l in range(cfg.hidden_layers):
            width = B * h
            Z = torch.empty((C, n, nw._ld(width)), dtype=st, device=self.device)
            Y = torch.empty_like(Z)
            # layers after the first are block-diagonal: each tile contracts its own mode's h inputs
            if split:
                Yl = torch.empty_like(Z)
                args = (cur.data_ptr(), cur_lo.data_ptr(), cur.shape[2], stg["W"][l].data_ptr(),
                        stg["Wlo"][l].data_ptr(), stg["W"][l].shape[1], stg["b"][l].data_ptr(), bc.omega(l),
                        Z.data_ptr(), None, Y.data_ptr(), Yl.data_ptr(), Y.shape[2], s)
                if l == 0:
                    H.call("neb_jet_layer_forward_x3", kd, H.SIN, n, k, width, *args)
                else:
                    H.call("neb_jet_layer_forward_x3_blocks", kd, H.SIN, n, width, h, *args)
                tape.inputs_lo.append(Yl)
                cur_lo = Yl
            else:
                args = (cur.data_ptr(), cur.shape[2], stg["W"][l].data_ptr(), stg["W"][l].shape[1],
                        stg["b"][l].data_ptr(), bc.omega(l), Z.data_ptr(), Y.data_ptr(), Y.shape[2], s)
                if l == 0:
                    H.call("neb_jet_layer_forward", p, kd, H.SIN, n, k, width, *args)
                else:
                    H.call("neb_jet_layer_forward_blocks", p, kd, H.SIN, n, width, h, *args)
            tape.pre.append(Z)
            tape.inputs.append(Y)
            cur, k = Y, width
        return tape

    def _block_fields_into(self, tape: nw.Tape, m0: int, B: int, prec: str, planes: list, layout: int = 0):
        """Write the block's lifted fields into columns m0 .. m0+B-1 of the
        full field matrices (planes: mass, mass_lo, stiff, stiff_lo or
        mass, stiff)."""
        cfg = self.cfg
        bc = cfg.block_net_config(B)
        stg = self.block_weights(m0, B, prec)
        Y = tape.inputs[-1]
        n, k = tape.n, B * cfg.hidden
        s = nw._stream()
        col = lambda t: t.data_ptr() + m0 * t.element_size()  # noqa: E731
        if prec == "bf16x3" and layout == 0 and cfg.hidden % 32 == 0:
            # each mode's output rows contract only its own hidden features
            H.call("neb_mode_output_fields_x3", bc.lift, n, cfg.hidden, B, Y.data_ptr(), tape.inputs_lo[-1].data_ptr(),
                   Y.shape[2], stg["W"][-1].data_ptr(), stg["Wlo"][-1].data_ptr(), stg["W"][-1].shape[1],
                   *[col(t) for t in planes], planes[0].shape[1], s)
        elif prec == "bf16x3":
            H.call("neb_jet_output_fields_x3", bc.lift, n, k, B, Y.data_ptr(), tape.inputs_lo[-1].data_ptr(),
                   Y.shape[2], stg["W"][-1].data_ptr(), stg["Wlo"][-1].data_ptr(), stg["W"][-1].shape[1],
                   *[col(t) for t in planes], planes[0].shape[1], layout, s)
        elif len(planes) == 4:
            H.call("neb_jet_output_fields_split", bc.lift, n, k, B, Y.data_ptr(), Y.shape[2], stg["W"][-1].data_ptr(),
                   stg["W"][-1].shape[1], *[col(t) for t in planes], planes[0].shape[1], layout, s)
        else:
            H.call("neb_jet_output_fields", tape.prec, bc.lift, n, k, B, Y.data_ptr(), Y.shape[2],
                   stg["W"][-1].data_ptr(), stg["W"][-1].shape[1], col(planes[0]), col(planes[1]),
                   planes[0].shape[1], layout, s)

    def _blocks(self):
        K, B = self.cfg.modes, self.cfg.block
        return [(m0, min(B, K - m0)) for m0 in range(0, K, B)]

    def fields(self, x: torch.Tensor, prec: str = "bf16x3", split: bool | None = None,
               keep_tapes: bool = False):
        """Lifted fields of all K modes, block by block into one
        [n*comp, K] matrix (split hi / lo planes for the split-precision
        forward, or for the bf16 forward with split=True).  keep_tapes=True
        returns (fields, tapes): the jet tapes of the first blocks, as many as
        fit TAPE_BUDGET bytes and the available device memory less
        TAPE_HEADROOM, kept for backward_fields (tapes[bi] None for a block
        whose tape the reverse recomputes, as the reference does for every
        block, VCORE:92-102)."""
        cfg = self.cfg
        K = cfg.modes
        split = (prec == "bf16x3") if split is None else (split or prec == "bf16x3")
        feats = self.features(x, prec)
        n = feats[0].shape[1]
        cm, cs = cfg.block_net_config(8).field_rows
        if split:
            planes = [torch.empty((r * n, nw._ld(K)), dtype=torch.bfloat16, device=self.device) for r in (cm, cm, cs, cs)]
        else:
            ft = torch.float64 if prec == "f64" else torch.float32
            planes = [torch.empty((r * n, nw._ld(K)), dtype=ft, device=self.device) for r in (cm, cs)]
        tapes = [] if keep_tapes else None
        budget = TAPE_BUDGET
        if keep_tapes and self.device.type == "cuda":
            free, _ = torch.cuda.mem_get_info(self.device)
            avail = free + torch.cuda.memory_reserved(self.device) - torch.cuda.memory_allocated(self.device)
            budget = min(budget, int(avail) - TAPE_HEADROOM)
        kept = 0
        for m0, B in self._blocks():
            tape = self.block_tape(feats, m0, B, prec)
            self._block_fields_into(tape, m0, B, prec, planes)
            if tapes is not None:
                nb = nw.tape_bytes(tape)
                if kept + nb <= budget:
                    tapes.append(tape)
                    kept += nb
                else:
                    tapes.append(None)
                    budget = kept         # later blocks are recomputed too
            del tape
        if split:
            f = nw.Fields(planes[0], planes[2], K, 0, planes[1], planes[3])
        else:
            f = nw.Fields(planes[0], planes[1], K, 0)
        # the streamfunction lift's periodic gauge over all modes' psi rows
        f = nw._gauge(cfg.block_net_config(8), f)
        return (f, tapes) if keep_tapes else f

    def output_jets(self, x: torch.Tensor, prec: str = "f32"):
        """Raw output jets of every mode (value [n, K*ncomp], grad [.., d],
        hess [.., P]; columns m*ncomp + c) -- forward_stream_jet2
        (HB2P:1802-1880) for all modes."""
        cfg = self.cfg
        if prec == "bf16x3":
            raise ValueError("raw output jets come from the bf16 / f32 / f64 forward")
        feats = self.features(x, prec, order=2)
        n = feats[0].shape[1]
        nout = cfg.modes * cfg.ncomp
        C = nw.channels(cfg.dim)
        ft = torch.float64 if prec == "f64" else torch.float32
        out = torch.empty((C, n, nw._ld(nout)), dtype=ft, device=self.device)
        for m0, B in self._blocks():
            tape = self.block_tape(feats, m0, B, prec)
            stg = self.block_weights(m0, B, prec)
            Y = tape.inputs[-1]
            H.call("neb_jet_output_jets", tape.prec, cfg.dim, n, B * cfg.hidden, B * cfg.ncomp, Y.data_ptr(),
                   Y.shape[2], stg["W"][-1].data_ptr(), stg["W"][-1].shape[1], stg["b"][-1].data_ptr(),
                   out.data_ptr() + m0 * cfg.ncomp * out.element_size(), out.shape[2], nw._stream())
        out = out[:, :, :nout]
        d = cfg.dim
        return out[0], out[1:1 + d].permute(1, 2, 0), out[1 + d:].permute(1, 2, 0)

    def forward_stream_jet1(self, x: torch.Tensor, prec: str = "f32"):
        """First-order jets of every mode (value, grad): the Jet1 stream of
        HB2P:1716-1800 (from output_jets, the Hessian channels dropped)."""
        v, g, _ = self.output_jets(x, prec)
        return v, g

    # -- reverse -----------------------------------------------------------
    def backward_fields(self, x: torch.Tensor, fields: nw.Fields, dM: torch.Tensor, dS: torch.Tensor, weight: float,
                        prec: str = "bf16x3", tapes: list | None = None) -> torch.Tensor:
        """Parameter gradient of L(M, S), M = w U^T U, S = w D^T D, for FP64
        dL/dM, dL/dS applied unsymmetrised (field cotangents 2w U dM^T,
        2w D dS^T: the coefficient peel's detached-prefix pair as given).
        The field cotangents are formed once for all modes (two GEMMs over
        the K columns, VCORE:94-95 gu = U Cu / n); then block by block the
        block's tape is recomputed and its columns of the cotangents go
        through the block's output layer and lift (neb_jet_output_backward
        with the block's B columns as the contraction), the output-weight
        products are the cotangent columns against the block's last hidden
        jets, and the hidden reverse keeps the diagonal blocks of each
        weight gradient (HB2P:2087-2139 backward_gradients_block_jet2).
        tapes: the blocks' tapes kept by fields(keep_tapes=True) (consumed:
        each entry is dropped once used); a block without one has its tape
        recomputed."""
        cfg = self.cfg
        K = cfg.modes
        d, C = cfg.jet_dim(), cfg.jet_channels()
        feats = self.features(x, prec)
        n = feats[0].shape[1]
        gdt = torch.float64 if prec == "f64" else torch.float32
        # field cotangents [rows, K] (2w folded in); split planes joined first
        gU = torch.mm(nw.field_matrix(fields, "mass", gdt), dM.to(gdt).T) * (2.0 * weight)
        gD = torch.mm(nw.field_matrix(fields, "stiff", gdt), dS.to(gdt).T) * (2.0 * weight)
        ldg = nw._ld(K)
        GU = torch.zeros((gU.shape[0], ldg), dtype=gdt, device=self.device)
        GD = torch.zeros((gD.shape[0], ldg), dtype=gdt, device=self.device)
        GU[:, :K] = gU
        GD[:, :K] = gD
        del gU, gD
        cot = nw.Fields(GU, GD, K, 0)
        grads = torch.zeros_like(self.params)
        gW, gb = self.split(grads)
        s = nw._stream()
        h, L, nc = cfg.hidden, cfg.hidden_layers, cfg.ncomp
        for bi, (m0, B) in enumerate(self._blocks()):
            tape = tapes[bi] if tapes is not None and tapes[bi] is not None else self.block_tape(feats, m0, B, prec)
            if tapes is not None:
                tapes[bi] = None          # released once the block's reverse is queued
            p = tape.prec
            st = nw._STORE[p]
            mode = nw._CROSS[p]
            bc = cfg.block_net_config(B)
            stg = self.block_weights(m0, B, prec)
            # the block's output weights per component c as a [B h] vector:
            # entry m h + j is mode m's weight W_out[m, c, j] (the one nonzero
            # of that row of the block-diagonal output layer)
            width = B * h
            bm = self.W[-1][m0:m0 + B].permute(1, 0, 2).reshape(nc, width).to(st).contiguous()
            A = torch.empty((C, n, nw._ld(width)), dtype=st, device=self.device)
            Z = tape.pre[-1]
            off = m0 * GU.element_size()
            H.call("neb_jet_output_backward_blocks", p, cfg.lift_code(), H.SIN, n, width, h, GU.data_ptr() + off,
                   GD.data_ptr() + off, ldg, nw._DT[gdt], 0, bm.data_ptr(), bm.data_ptr(), width,
                   Z.data_ptr(), Z.shape[2], bc.omega(L - 1), A.data_ptr(), A.shape[2], s)
            # output weights: the block's cotangent columns against its last hidden jets
            P = self._field_products(tape, cot, m0, B)
            for c in range(nc):   # gW[-1][m0 + m, c] = P[c][m]
                gW[-1][m0:m0 + B, c].copy_(P[c])
            # hidden layers, diagonal blocks of each weight gradient
            for l in range(L - 1, -1, -1):
                X = tape.inputs[l]
                kin = X.shape[2] if l > 0 else cfg.features
                # layer 0 reads the shared features (dense); later layers only their diagonal blocks
                G = nw.cross([(A.view(C * n, A.shape[2]), X.view(C * n, X.shape[2]), C * n, 1.0)], width, kin, mode,
                             block=h if l > 0 else None)
                ws = nw.workspace(self.device).get(H.load().neb_colsum_workspace_size(n, width))
                db = torch.empty(width, dtype=torch.float64, device=self.device)
                H.call("neb_colsum", nw._DT[st], A[0].data_ptr(), n, width, A.shape[2], 0.0, db.data_ptr(),
                       ws.data_ptr(), ws.numel(), s)
                gb[l][m0:m0 + B] = db.view(B, h)
                if l == 0:
                    gW[l][m0:m0 + B].copy_(G.view(B, h, kin))
                else:     # the diagonal h x h blocks: gW[l][m0 + m] = G[m h:(m+1) h, m h:(m+1) h]
                    gW[l][m0:m0 + B].copy_(G.view(B, h, B, h).diagonal(dim1=0, dim2=2).permute(2, 0, 1))
                if l > 0:
                    Zp = tape.pre[l - 1]
                    An = torch.empty((C, n, nw._ld(kin)), dtype=st, device=self.device)
                    WT = stg["WT"][l]
                    H.call("neb_jet_layer_backward_blocks", p, d, H.SIN, n, width, h, A.data_ptr(), A.shape[2],
                           WT.data_ptr(), WT.shape[1], Zp.data_ptr(), Zp.shape[2], bc.omega(l - 1), An.data_ptr(),
                           An.shape[2], s)
                    A = An
            del tape, A
        return grads

    def _field_products(self, tape: nw.Tape, cot: nw.Fields, m0: int, B: int) -> list:
        """Per output component c: [B, h], mode m0+m's output-weight
        products -- the block's field-cotangent column m (mass rows of
        cot.mass, stiffness rows of cot.stiff) against mode m's own h last
        hidden jet features of the channels each field reads (the lift's
        terms), summed over the points (neb_cross_modes: no products against
        the block's other modes)."""
        cfg = self.cfg
        n, h = tape.n, cfg.hidden
        Y = tape.inputs[-1]
        ch = lambda c: Y[c]  # noqa: E731
        col = lambda t: t[:, m0:m0 + B]  # noqa: E731
        U, Dg = cot.mass, cot.stiff
        rv = nw._rows_view

        def run(segs):
            P = nw.cross_modes(segs[:16], B, h)
            for i in range(16, len(segs), 16):
                nw.cross_modes(segs[i:i + 16], B, h, out=P, beta=1.0)
            return P

        if cfg.dim == 2 and cfg.stiffness == "stream":
            # psi rows against the value channel (the cotangent columns are
            # mean-free, so the uncentered channel gives the centered product),
            # gradient rows against the gradient channels
            return [run([(col(U), ch(0), n, 1.0), (col(rv(Dg, 0, 2)), ch(1), n, 1.0),
                         (col(rv(Dg, 1, 2)), ch(2), n, 1.0)])]
        if cfg.dim == 2:
            u, v = rv(U, 0, 2), rv(U, 1, 2)
            segs = [(col(u), ch(2), n, 1.0), (col(v), ch(1), n, -1.0)]
            if cfg.stiffness == "hessian":
                segs += [(col(rv(Dg, 0, 3)), ch(3), n, 1.0), (col(rv(Dg, 2, 3)), ch(5), n, 1.0),
                         (col(rv(Dg, 1, 3)), ch(4), n, math.sqrt(2.0))]
            else:
                segs += [(col(Dg), ch(3), n, -1.0), (col(Dg), ch(5), n, -1.0)]
            return [run(segs)]
        from .lifts import curl_terms
        comps = [rv(U, a, 3) for a in range(3)]
        grads3 = [rv(Dg, c, 9) for c in range(9)]
        out = []
        for sl in range(3):
            segs = [(col(comps[a]), ch(1 + c), n, sg) for (a, c, w_s, sg) in curl_terms() if a < 3 and w_s == sl]
            segs += [(col(grads3[a - 3]), ch(1 + c), n, sg) for (a, c, w_s, sg) in curl_terms()
                     if a >= 3 and w_s == sl]
            out.append(run(segs))
        return out

    # -- the reference's step ---------------------------------------------
    def step(self, x: torch.Tensor, weight: float, prec: str = "bf16x3", objective: str = "peel",
             training: bool = True, peel_options: dict | None = None, lam: float = 1e-2, ridge: float = 0.0) -> dict:
        """Fields -> FP64 Grams -> objective -> recompute + reverse
        (VCORE:79-105 compute on the per-mode networks).  objective="peel":
        the sorted coefficient peel (VCORE:48-69; 2D with
        stiffness="hessian" and the velocity options, HB2P:2935-3380, or
        stiffness="stream" with peel_options rel_min=0, ortho_tol=inf,
        mass_floor: the Dirichlet streamfunction peel, HB2P:3383-3770);
        "trace": tr((M + ridge I)^-1 S) + lam ||M - I||^2."""
        gram_mode = "f64" if prec == "f64" else ("f32" if prec == "f32" else "bf16x3")
        split = prec in ("bf16", "bf16x3")
        fields, tapes = self.fields(x, prec, split=split, keep_tapes=True) if training \
            else (self.fields(x, prec, split=split), None)
        M, S = nw.field_grams(fields, weight, gram_mode)
        K = self.cfg.modes
        out = dict(M=M, S=S, fields=fields)
        if objective == "peel":
            pk = sp.coefficient_peel(M, S, **(peel_options or {}))
            out.update(peel=pk, loss=pk["loss"], losses=pk["losses"])
            dM, dS = pk["pair"][:, :K], pk["pair"][:, K:]
        elif objective == "trace":
            loss, dM, dS = sp.spectral_objective(M, S, lam, ridge, gate="cholesky")
            out["loss"] = loss
        else:
            raise ValueError(f"unknown objective {objective!r}")
        out.update(dM=dM, dS=dS)
        if training:
            out["tapes_kept"] = sum(t is not None for t in tapes)     # blocks whose tape the reverse reuses
            out["grads"] = self.backward_fields(x, fields, dM, dS, weight, prec, tapes=tapes)
        return out

    # -- optimizer ---------------------------------------------------------
    def mode_norms(self, grads: torch.Tensor) -> torch.Tensor:
        """Per-mode gradient norms over every layer's weights and biases."""
        gW, gb = self.split(grads)
        K = self.cfg.modes
        sq = torch.zeros(K, dtype=torch.float64, device=self.device)
        for t in (*gW, *gb):
            sq += t.reshape(K, -1).to(torch.float64).pow(2).sum(1)
        return sq.sqrt()

    def adam(self, grads: torch.Tensor, lr: float, clip: float = 8.0):
        """Adam with a per-mode norm clip (NCC:156-171): mode k's gradient is
        scaled by min(1, clip / |g_k|); a nonfinite or zero norm of any mode
        raises FloatingPointError before any update.  Returns the norms."""
        norms = self.mode_norms(grads)
        if not bool(torch.all(torch.isfinite(norms) & (norms > 0))):
            raise FloatingPointError("Invalid parameter gradients; no optimizer update")
        scales = torch.clamp(clip / torch.clamp(norms, min=1e-30), max=1.0)
        g = grads.clone()
        gW, gb = self.split(g)
        K = self.cfg.modes
        for t in (*gW, *gb):
            t.mul_(scales.to(t.dtype).reshape((K,) + (1,) * (t.dim() - 1)))
        t = self.epoch + 1
        dt = nw._DT[self.master_dtype]
        H.call("neb_adam", dt, self.params.data_ptr(), g.data_ptr(), self.m.data_ptr(), self.v.data_ptr(),
               self.params.numel(), lr, 1.0, 0.9, 0.999, 1e-8, t, nw._stream())
        self.epoch = t
        self.touch()
        return norms

    # -- checkpoints -------------------------------------------------------
    def save(self, path):
        """NPZ with W_j, b_j, mW_j, vW_j, mb_j, vb_j ([K, out, in] / [K, out])
        and epoch (NCC:173-178)."""
        import os
        arrays = {}
        for name, flat in (("", self.params), ("m", self.m), ("v", self.v)):
            Ws, bs = self.split(flat)
            for j, (W, b) in enumerate(zip(Ws, bs)):
                arrays[f"{name}W_{j}"] = W.detach().cpu().numpy()
                arrays[f"{name}b_{j}"] = b.detach().cpu().numpy()
        arrays["epoch"] = np.asarray(self.epoch)
        tmp = str(path) + ".tmp.npz"
        np.savez(tmp, **arrays)
        os.replace(tmp, path)

    def load(self, path):
        with np.load(path, allow_pickle=False) as z:
            self.epoch = int(z["epoch"])
            for name, flat in (("", self.params), ("m", self.m), ("v", self.v)):
                Ws, bs = self.split(flat)
                with torch.no_grad():
                    for j, (W, b) in enumerate(zip(Ws, bs)):
                        for dst, key in ((W, f"{name}W_{j}"), (b, f"{name}b_{j}")):
                            src = z[key]
                            if src.shape != tuple(dst.shape):
                                raise ValueError(f"{key}: shape {src.shape} != {tuple(dst.shape)}")
                            dst.copy_(torch.as_tensor(src, dtype=self.master_dtype))
        self.touch()

    def arrays(self) -> dict:
        """The 2D HardcodedModel key layout (HB2P:1436-1470, 1596-1622):
        m{mode:03d}.l{layer}.weight / .bias, m{mode:03d}.output.weight /
        .bias and the Adam state m{mode:03d}.adam.p{index:02d}.exp_avg /
        .exp_avg_sq / .step (parameters ordered weight, bias per layer)."""
        out = {}
        L = self.cfg.hidden_layers
        P, Mm, Vv = self.split(self.params), self.split(self.m), self.split(self.v)
        for k in range(self.cfg.modes):
            for j in range(L + 1):
                stem = f"m{k:03d}.l{j}" if j < L else f"m{k:03d}.output"
                out[f"{stem}.weight"] = P[0][j][k].detach().cpu().numpy()
                out[f"{stem}.bias"] = P[1][j][k].detach().cpu().numpy()
                for q, (src, dst) in enumerate(((0, "weight"), (1, "bias"))):
                    pi = 2 * j + q
                    out[f"m{k:03d}.adam.p{pi:02d}.exp_avg"] = Mm[src][j][k].detach().cpu().numpy()
                    out[f"m{k:03d}.adam.p{pi:02d}.exp_avg_sq"] = Vv[src][j][k].detach().cpu().numpy()
                    out[f"m{k:03d}.adam.p{pi:02d}.step"] = np.asarray(self.epoch)
        return out

    def from_arrays(self, arrays: dict):
        L = self.cfg.hidden_layers
        P, Mm, Vv = self.split(self.params), self.split(self.m), self.split(self.v)
        steps = set()
        with torch.no_grad():
            for k in range(self.cfg.modes):
                for j in range(L + 1):
                    stem = f"m{k:03d}.l{j}" if j < L else f"m{k:03d}.output"
                    for q, suffix in enumerate(("weight", "bias")):
                        pi = 2 * j + q
                        for dst, key in ((P[q][j][k], f"{stem}.{suffix}"),
                                         (Mm[q][j][k], f"m{k:03d}.adam.p{pi:02d}.exp_avg"),
                                         (Vv[q][j][k], f"m{k:03d}.adam.p{pi:02d}.exp_avg_sq")):
                            src = np.asarray(arrays[key])
                            if src.shape != tuple(dst.shape):
                                raise ValueError(f"{key}: shape {src.shape} != {tuple(dst.shape)}")
                            dst.copy_(torch.as_tensor(src, dtype=self.master_dtype))
                        steps.add(int(arrays[f"m{k:03d}.adam.p{pi:02d}.step"]))
        if len(steps) != 1:
            raise ValueError("per-parameter Adam steps differ; a single shared step is required")
        self.epoch = steps.pop()
        self.touch()

