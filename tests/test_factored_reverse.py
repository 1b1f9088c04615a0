"""The factored output-layer reverse of split-precision tapes.

CPU: the FP64 host algebra (lift_terms, _factored_coeffs, _factored_weights:
the channel blocks R = 2w sum sg sg' W_s'^T dX^T W_s; _factored_ghat,
_factored_output_grad: the output-weight gradient from the channel Grams) against the oracle's unfactored chain -- output layer, lift
(lift2d / lift2d_hessian / lift3d), field adjoint 2w Phi dG, lift transpose
and linear transpose -- on random jets, for every lift.

GPU: the three kernels of the factored path (neb_stack_channels,
neb_gemm_split, neb_jet_act_reverse) against torch / the oracle, and the
bf16x3 reverse at the headline shapes is covered by tests/test_configs.py.
"""
import math
from types import SimpleNamespace

import numpy as np
import pytest
import torch

from neb_amd import network as nw
from oracle import neb_oracle as o


def _oracle_chain(Y: o.Jet, W: np.ndarray, dim: int, stiffness: str, w: float, dM, dS):
    """adj of the last hidden jets (packed [C, n, width]) and the output-weight
    gradient through the oracle's output layer, lift and their transposes."""
    out = o.linear_jet(Y, W, None)
    if dim == 3:
        u, du = o.lift3d(out)
        U, D = o.field_rows(u.transpose(0, 1, 2)), o.field_rows(du)
    elif stiffness == "hessian":
        u, h = o.lift2d_hessian(out)
        U, D = o.field_rows(u), o.field_rows(h)
    else:
        u, om = o.lift2d(out)
        U, D = o.field_rows(u), om
    # field cotangents 2w Phi dG^T (backward_fields(symmetric=True) applies the
    # given adjoints unsymmetrised: the coefficient peel's are not symmetric)
    gU = o.field_adjoint(U, w, dM.T).reshape(u.shape)
    if dim == 3:
        adj = o.lift3d_reverse(gU, o.field_adjoint(D, w, dS.T).reshape(du.shape))
    elif stiffness == "hessian":
        adj = o.lift2d_hessian_reverse(gU, o.field_adjoint(D, w, dS.T).reshape(h.shape))
    else:
        adj = o.lift2d_reverse(gU, o.field_adjoint(D, w, dS.T))
    A = adj.pack()                                   # [C, n, nout]
    adjY = np.stack([a @ W for a in A])              # [C, n, width]
    gW = sum(a.T @ y for a, y in zip(A, Y.pack()))
    return adjY, gW


@pytest.mark.parametrize("dim,stiffness", [(2, "vorticity"), (2, "hessian"), (3, "vorticity")])
def test_factored_algebra_matches_oracle_chain(dim, stiffness):
    rng = np.random.default_rng(21 + dim)
    n, width, K = 37, 16, 12
    ncomp = 3 if dim == 3 else 1
    P = dim * (dim + 1) // 2
    Y = o.Jet(rng.normal(size=(n, width)), rng.normal(size=(n, width, dim)), rng.normal(size=(n, width, P)))
    W = rng.normal(size=(ncomp * K, width)) / math.sqrt(width)
    dM, dS = rng.normal(size=(K, K)), rng.normal(size=(K, K))     # not symmetric
    w = 0.37
    adjY, gW = _oracle_chain(Y, W, dim, stiffness, w, dM, dS)

    cfg = nw.NetConfig(dim=dim, widths=(width,), modes=K, stiffness=stiffness, split_reverse="factored")
    net = SimpleNamespace(cfg=cfg, W=[torch.tensor(W)])
    mch, sch = nw.factored_channels(cfg)
    Yp = Y.pack()
    stack = np.concatenate([Yp[c] for c in mch + sch], 1)
    km = len(mch) * width
    G = [torch.tensor(stack[:, :km].T @ stack[:, :km]), torch.tensor(stack[:, km:].T @ stack[:, km:])]
    fp = nw.FactoredProducts(None, None, (mch, sch), nw._factored_ghat(cfg, G))
    pair = torch.tensor(np.concatenate([dM, dS], 1))
    gout = nw._factored_output_grad(net, fp, pair, w)
    assert np.allclose(gout.numpy(), gW, rtol=1e-11, atol=1e-11 * np.abs(gW).max())
    RTs = nw._factored_weights(net, pair, w)
    AY = np.concatenate([stack[:, :km] @ RTs[0].numpy().T, stack[:, km:] @ RTs[1].numpy().T], 1)
    chans = mch + sch
    for c in range(o.channels(dim)):
        want = adjY[c]
        if c in chans:
            i = chans.index(c)
            got = AY[:, i * width:(i + 1) * width]
            assert np.allclose(got, want, rtol=1e-11, atol=1e-11 * np.abs(adjY).max()), (dim, stiffness, c)
        else:      # channels the lift does not read carry no adjoint
            assert np.abs(want).max() <= 1e-12 * np.abs(adjY).max(), (dim, stiffness, c)


# --------------------------------------------------------------------------
# GPU: the kernels of the factored path
# --------------------------------------------------------------------------
def _bf16_planes(x: torch.Tensor):
    hi = x.to(torch.bfloat16)
    lo = (x - hi.to(x.dtype)).to(torch.bfloat16)
    return hi, lo


@pytest.mark.gpu
def test_factored_kernels_gpu():
    if not torch.cuda.is_available():
        pytest.skip("needs a GPU")
    from neb_amd import _hip as H
    dev = "cuda"
    g = torch.Generator(device="cpu").manual_seed(5)
    C, n, width, ld = 10, 1003, 64, 72
    Yf = torch.randn((C, n, ld), generator=g, dtype=torch.float64).to(dev)
    Yf[:, :, width:] = 0
    Yhi, Ylo = _bf16_planes(Yf)
    chans = [1, 2, 3, 4, 5, 6, 7, 8, 9]
    lds = len(chans) * width
    hi = torch.empty((n, lds), dtype=torch.bfloat16, device=dev)
    lo = torch.empty_like(hi)
    s = torch.cuda.current_stream().cuda_stream
    print("neb_stack_channels", n, width, chans, flush=True)
    H.call("neb_stack_channels", Yhi.data_ptr(), Ylo.data_ptr(), n, ld, width, len(chans), H.int_array(chans),
           hi.data_ptr(), lo.data_ptr(), lds, s)
    torch.cuda.synchronize()
    want_hi = torch.cat([Yhi[c][:, :width] for c in chans], 1)
    want_lo = torch.cat([Ylo[c][:, :width] for c in chans], 1)
    assert torch.equal(hi, want_hi) and torch.equal(lo, want_lo)

    # split GEMM: x (n x k) . w^T (nout x k) with ragged n, k, nout
    for (nn, k, nout) in ((1003, 576, 576), (517, 200, 72), (64, 8, 8)):
        print("neb_gemm_split", nn, k, nout, flush=True)
        X = torch.randn((nn, lds), generator=g, dtype=torch.float64).to(dev)
        Wm = torch.randn((nout, k), generator=g, dtype=torch.float64).to(dev) * 1e3
        xh, xl = _bf16_planes(X)
        wpad = nw._ld(k)
        wh = torch.zeros((nout, wpad), dtype=torch.bfloat16, device=dev)
        wl = torch.zeros_like(wh)
        wh[:, :k], wl[:, :k] = _bf16_planes(Wm)
        out = torch.empty((nn, nw._ld(nout)), dtype=torch.float32, device=dev)
        H.call("neb_gemm_split", nn, k, nout, xh.data_ptr(), xl.data_ptr(), lds, wh.data_ptr(), wl.data_ptr(), wpad,
               out.data_ptr(), out.shape[1], s)
        torch.cuda.synchronize()
        ref = (xh.double() + xl.double())[:, :k] @ (wh.double() + wl.double())[:, :k].T
        err = float((out[:, :nout].double() - ref).norm() / ref.norm())
        assert err < 1e-5, (nn, k, nout, err)     # fp32 accumulation over k

    # activation transpose against the oracle (tanh and sin, 2D and 3D)
    for dim, act, kind in ((3, H.TANH, "tanh"), (2, H.SIN, "sin"), (3, H.SIN, "sin")):
        Cd = o.channels(dim)
        nch = Cd - 1
        print("neb_jet_act_reverse", dim, kind, flush=True)
        Z = (0.5 * torch.randn((Cd, n, ld), generator=g, dtype=torch.float64)).to(dev).to(torch.bfloat16)
        adj = torch.randn((n, nch * width), generator=g, dtype=torch.float32).to(dev)
        out = torch.empty((Cd, n, ld), dtype=torch.bfloat16, device=dev)
        om = 1.7
        H.call("neb_jet_act_reverse", H.BF16, dim, act, n, width, adj.data_ptr(), adj.shape[1], nch,
               H.int_array(list(range(1, Cd))), Z.data_ptr(), ld, om, out.data_ptr(), ld, s)
        torch.cuda.synchronize()
        Zj = o.Jet.unpack(Z[:, :, :width].double().cpu().numpy(), dim)
        A = np.zeros((Cd, n, width))
        A[1:] = adj.double().cpu().numpy().reshape(n, nch, width).transpose(1, 0, 2)
        ref = o.activation_jet_reverse(kind, Zj, o.Jet.unpack(A, dim)).pack() * om
        got = out[:, :, :width].double().cpu().numpy()
        err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
        assert err < 4e-3, (dim, kind, err)
