"""GPU parity of the per-mode SIREN networks (neb_amd.modes: the
reference's own architecture, HB2P:1590-1880 / NCC:111-185) against the
oracle's per-mode networks, through the C ABI: fields of every mode in 2D
(vorticity, Hessian and streamfunction lifts) and 3D, independence of the mode-block
size, the reverse pass, the coefficient-peel step, the per-mode clipped
Adam and both checkpoint layouts."""
import math

import numpy as np
import pytest
import torch

from oracle import neb_oracle as o

pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():  # CPU-only host: the gpu tier runs on the MI355X box
    pytest.skip("needs a GPU", allow_module_level=True)

from neb_amd import modes as md  # noqa: E402
from neb_amd import network as nw  # noqa: E402
from neb_amd import spectral as sp  # noqa: E402

DEV = "cuda"


def rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def make(dim=2, K=24, block=8, hidden=16, layers=2, stiffness="vorticity", master=torch.float64, seed=0):
    if dim == 2:
        cfg = md.ModeConfig(dim=2, frequencies=2, hidden=hidden, hidden_layers=layers, modes=K, omega_first=2.0,
                            omega_hidden=1.5, block=block, stiffness=stiffness)
    else:
        cfg = md.ModeConfig(dim=3, frequencies=2, lo=(-1.0,) * 3, hi=(1.0,) * 3, hidden=hidden, hidden_layers=layers,
                            modes=K, omega_first=2.0, omega_hidden=1.5, block=block)
    net = md.ModeNet(cfg, seed=seed, device=DEV, master_dtype=master)
    return cfg, net


def oracle_of(net):
    Ws = [W.detach().cpu().double().numpy() for W in net.W]
    bs = [b.detach().cpu().double().numpy() for b in net.b]
    return o.mode_nets(Ws, bs, net.cfg.omega_first, net.cfg.omega_hidden)


def points(dim, n, seed):
    rng = np.random.default_rng(seed)
    return rng.uniform(0, 2 * np.pi, (n, 2)) if dim == 2 else rng.uniform(-1, 1, (n, 3))


def field_np(f, which):
    return nw.field_matrix(f, which, torch.float64).cpu().numpy()


def test_siren_init_streams_and_bounds():
    """Mode i's weights come from its own stream: independent of K; every
    layer within its SIREN bound."""
    c8 = md.ModeConfig(modes=8, hidden=16, hidden_layers=2, frequencies=2, block=8)
    c16 = md.ModeConfig(modes=16, hidden=16, hidden_layers=2, frequencies=2, block=8)
    W8, b8 = md.siren_init(c8, 3)
    W16, b16 = md.siren_init(c16, 3)
    for j in range(3):
        assert np.array_equal(W8[j], W16[j][:8]) and np.array_equal(b8[j], b16[j][:8])
        fi = c8.sizes[j]
        bound = 1.0 / fi if j == 0 else math.sqrt(6.0 / fi) / c8.omega_hidden
        assert np.abs(W16[j]).max() <= bound and np.abs(b16[j]).max() <= bound
        assert np.abs(W16[j]).max() > 0.9 * bound


@pytest.mark.parametrize("dim,stiff", [(2, "vorticity"), (2, "hessian"), (3, "grad_u")])
@pytest.mark.parametrize("prec,tol", [("f64", 1e-10), ("bf16x3", 2e-5), ("f32", 2e-5)])
def test_mode_fields_match_oracle(dim, stiff, prec, tol):
    cfg, net = make(dim, K=24 if dim == 2 else 16, stiffness=stiff if dim == 2 else "vorticity",
                    master=torch.float64 if prec == "f64" else torch.float32)
    x = points(dim, 257, 1)
    U, D, _ = o.mode_fields(oracle_of(net), o.feature_jets(x, 2, cfg.lo, cfg.hi), dim,
                            stiff if dim == 2 else "vorticity")
    f = net.fields(torch.tensor(x), prec)
    r = rel(field_np(f, "mass"), U); assert r < tol, r
    r = rel(field_np(f, "stiff"), D); assert r < tol, r


@pytest.mark.parametrize("dim,stiff,hidden,block", [(2, "vorticity", 32, 8), (2, "hessian", 64, 16),
                                                    (3, "grad_u", 128, 8), (3, "grad_u", 32, 16)])
def test_mode_split_fields_kernel_matches_oracle(dim, stiff, hidden, block):
    """Split-precision fields through neb_mode_output_fields_x3 (each mode's
    output rows contract only its own hidden features; hidden a multiple of
    32) against the oracle, on a ragged point count."""
    K = 32 if dim == 2 else 16
    cfg, net = make(dim, K=K, block=block, hidden=hidden, stiffness=stiff if dim == 2 else "vorticity",
                    master=torch.float32)
    x = points(dim, 1001, 11)
    U, D, _ = o.mode_fields(oracle_of(net), o.feature_jets(x, 2, cfg.lo, cfg.hi), dim,
                            stiff if dim == 2 else "vorticity")
    f = net.fields(torch.tensor(x), "bf16x3")
    r = rel(field_np(f, "mass"), U); assert r < 2e-5, r
    r = rel(field_np(f, "stiff"), D); assert r < 2e-5, r


def test_mode_block_size_does_not_change_fields():
    cfg, net8 = make(2, K=24, block=8)
    _, net24 = make(2, K=24, block=24)
    x = torch.tensor(points(2, 300, 2))
    f8, f24 = net8.fields(x, "f64"), net24.fields(x, "f64")
    assert rel(field_np(f8, "mass"), field_np(f24, "mass")) < 1e-14
    assert rel(field_np(f8, "stiff"), field_np(f24, "stiff")) < 1e-14


@pytest.mark.parametrize("dim,stiff", [(2, "vorticity"), (2, "hessian"), (3, "grad_u")])
@pytest.mark.parametrize("prec,tol", [("f64", 1e-9), ("bf16x3", 4e-2)])
def test_mode_reverse_matches_oracle(dim, stiff, prec, tol):
    K = 24 if dim == 2 else 16
    cfg, net = make(dim, K=K, stiffness=stiff if dim == 2 else "vorticity",
                    master=torch.float64 if prec == "f64" else torch.float32)
    x = points(dim, 300, 3)
    w = 1.0 / 300
    rng = np.random.default_rng(9)
    dM = rng.normal(size=(K, K)); dM = dM + dM.T
    dS = rng.normal(size=(K, K)); dS = dS + dS.T
    nets = oracle_of(net)
    st = stiff if dim == 2 else "vorticity"
    U, D, tapes = o.mode_fields(nets, o.feature_jets(x, 2, cfg.lo, cfg.hi), dim, st)
    gW, gb = o.mode_reverse(nets, tapes, U, D, w, dM, dS, dim, st)
    xt = torch.tensor(x)
    f = net.fields(xt, prec, split=prec != "f64")
    g = net.backward_fields(xt, f, torch.tensor(dM, device=DEV), torch.tensor(dS, device=DEV), w, prec)
    GW, Gb = net.split(g)
    for j, (a, b) in enumerate(zip(GW, gW)):
        r = rel(a.cpu(), b); assert r < tol, ("W", j, r)
    for j, (a, b) in enumerate(zip(Gb[:-1], gb[:-1])):
        r = rel(a.cpu(), b); assert r < tol, ("b", j, r)


def test_mode_peel_step_matches_oracle():
    """The 2D velocity peel on per-mode networks (Hessian stiffness): losses,
    order and gradient against the oracle's Cholesky peel of the oracle's
    Grams and the per-mode reverse of its detached-prefix cotangents."""
    K = 16
    cfg, net = make(2, K=K, stiffness="hessian", hidden=16, layers=2)
    x = points(2, 400, 4)
    w = 4 * math.pi ** 2 / 400
    nets = oracle_of(net)
    U, D, tapes = o.mode_fields(nets, o.feature_jets(x, 2, cfg.lo, cfg.hi), 2, "hessian")
    M, S = o.gram(U, w), o.gram(D, w)
    pk = o.coefficient_peel(M, S, ortho_tol=1e-3)
    gW, gb = o.mode_reverse(nets, tapes, U, D, w, pk["Cu"] / 2, pk["Cd"] / 2, 2, "hessian")
    res = net.step(torch.tensor(x), w, prec="f64", objective="peel",
                   peel_options=dict(ortho_tol=1e-3, rel_min=1e-4, mass_floor=1e-12))
    assert np.array_equal(res["peel"]["order"].cpu().numpy(), pk["order"])
    r = rel(res["losses"].cpu(), pk["losses"]); assert r < 1e-9, r
    GW, _ = net.split(res["grads"])
    for j, (a, b) in enumerate(zip(GW, gW)):
        r = rel(a.cpu(), b); assert r < 1e-8, (j, r)


@pytest.mark.parametrize("prec", ["f64", "bf16x3"])
def test_mode_step_partially_kept_tapes(prec, monkeypatch):
    """Block tapes kept from the fields pass (all, the first ones that fit
    the budget, none) give the same gradient bit for bit: the reverse reads
    a kept tape or recomputes the same one."""
    K = 24
    cfg, net = make(2, K=K, block=8, stiffness="hessian", hidden=16, layers=2,
                    master=torch.float64 if prec == "f64" else torch.float32)
    x = torch.tensor(points(2, 300, 5))
    w = 4 * math.pi ** 2 / 300
    opts = dict(ortho_tol=1e-3, rel_min=1e-4, mass_floor=1e-12, deferred=True)
    one = net.block_tape(net.features(x, prec), 0, 8, prec)
    per_block = nw.tape_bytes(one)
    del one
    grads = {}
    for budget, kept in ((md.TAPE_BUDGET, 3), (2 * per_block, 2), (0, 0)):
        monkeypatch.setattr(md, "TAPE_BUDGET", budget)
        res = net.step(x, w, prec=prec, objective="peel", peel_options=opts)
        assert res["tapes_kept"] == kept
        grads[kept] = res["grads"].cpu()
    assert torch.equal(grads[3], grads[0]) and torch.equal(grads[2], grads[0])


@pytest.mark.parametrize("prec,hidden,tol", [("f64", 16, 1e-10), ("bf16x3", 32, 2e-5)])
def test_mode_stream_fields_centered(prec, hidden, tol):
    """Streamfunction lift of per-mode networks (bf16x3 with hidden 32 runs
    neb_mode_output_fields_x3's per-mode contraction): psi rows centered over
    the points, gradient rows, against the oracle."""
    cfg, net = make(2, K=24, stiffness="stream", hidden=hidden, master=torch.float64 if prec == "f64" else torch.float32)
    x = points(2, 301, 5)
    U, D, _ = o.mode_fields(oracle_of(net), o.feature_jets(x, 2, cfg.lo, cfg.hi), 2, "stream")
    f = net.fields(torch.tensor(x), prec)
    r = rel(field_np(f, "mass"), U - U.mean(0)); assert r < tol, r
    r = rel(field_np(f, "stiff"), D); assert r < tol, r


def test_mode_stream_peel_step_matches_oracle():
    """The Dirichlet streamfunction peel on per-mode networks: quotients,
    order and gradient against the oracle's literal peel and the per-mode
    reverse of its cotangents (through the centering)."""
    K = 16
    cfg, net = make(2, K=K, stiffness="stream", hidden=16, layers=2)
    x = points(2, 400, 6)
    n = len(x)
    w = 4 * math.pi ** 2 / n
    nets = oracle_of(net)
    U, D, tapes = o.mode_fields(nets, o.feature_jets(x, 2, cfg.lo, cfg.hi), 2, "stream")
    pk = o.streamfunction_rayleigh_peel(U, D, w)
    gWs = []
    for k, (onet, tape) in enumerate(zip(nets, tapes)):
        adj = o.lift2d_stream_reverse(pk["G_value"][:, k:k + 1], pk["G_grad"][:, k].reshape(n, 2, 1))
        gWs.append(o.backward(onet, tape, adj)[0])
    gW = [np.stack([g[j] for g in gWs]) for j in range(len(gWs[0]))]
    res = net.step(torch.tensor(x), w, prec="f64", objective="peel",
                   peel_options=dict(ortho_tol=math.inf, rel_min=0.0, mass_floor=1e-12))
    assert np.array_equal(res["peel"]["order"].cpu().numpy(), pk["order"])
    r = rel(res["losses"].cpu(), pk["rayleigh"]); assert r < 1e-9, r
    GW, _ = net.split(res["grads"])
    for j, (a, b) in enumerate(zip(GW, gW)):
        r = rel(a.cpu(), b); assert r < 1e-8, (j, r)


def test_mode_adam_matches_oracle_over_three_steps():
    cfg, net = make(2, K=16, master=torch.float64)
    rng = np.random.default_rng(10)
    Ws = [W.cpu().numpy().copy() for W in net.W]
    bs = [b.cpu().numpy().copy() for b in net.b]
    mW = [np.zeros_like(W) for W in Ws]; vW = [np.zeros_like(W) for W in Ws]
    mb = [np.zeros_like(b) for b in bs]; vb = [np.zeros_like(b) for b in bs]
    for t in range(1, 4):
        gW = [rng.normal(size=W.shape) * (50.0 if t == 2 else 1.0) for W in Ws]
        gb = [rng.normal(size=b.shape) for b in bs]
        flat = torch.zeros_like(net.params)
        FW, Fb = net.split(flat)
        for dst, src in zip(FW + Fb, gW + gb):
            dst.copy_(torch.tensor(src))
        norms = net.adam(flat, 1e-3, clip=8.0)
        out, onorms = o.mode_adam(Ws, bs, gW, gb, mW, vW, mb, vb, t, 1e-3, clip=8.0)
        L = len(Ws)
        Ws, bs = [q[0] for q in out[:L]], [q[0] for q in out[L:]]
        mW, mb = [q[1] for q in out[:L]], [q[1] for q in out[L:]]
        vW, vb = [q[2] for q in out[:L]], [q[2] for q in out[L:]]
        assert rel(norms.cpu(), onorms) < 1e-12
        for a, b in zip(net.W + net.b, Ws + bs):
            r = rel(a.cpu(), b); assert r < 1e-12, (t, r)
    assert net.epoch == 3
    with pytest.raises(FloatingPointError):
        net.adam(torch.zeros_like(net.params), 1e-3)
    assert net.epoch == 3


def test_mode_checkpoint_layouts(tmp_path):
    cfg, net = make(2, K=8, master=torch.float32)
    net.adam(torch.randn_like(net.params), 1e-3)
    path = tmp_path / "modes.npz"
    net.save(path)
    with np.load(path, allow_pickle=False) as z:
        assert z["W_0"].shape == (8, cfg.hidden, cfg.features) and z["W_2"].shape == (8, 1, cfg.hidden)
        assert int(z["epoch"]) == 1
    _, other = make(2, K=8, master=torch.float32, seed=5)
    other.load(path)
    assert torch.equal(other.params, net.params) and torch.equal(other.m, net.m) and other.epoch == 1
    arrays = net.arrays()
    assert arrays["m003.l1.weight"].shape == (cfg.hidden, cfg.hidden)
    assert arrays["m003.output.weight"].shape == (1, cfg.hidden)
    assert int(arrays["m007.adam.p05.step"]) == 1
    _, third = make(2, K=8, master=torch.float32, seed=6)
    third.from_arrays(arrays)
    assert torch.equal(third.params, net.params) and torch.equal(third.v, net.v) and third.epoch == 1


@pytest.mark.parametrize("dim,h,B", [(3, 128, 4), (2, 32, 16), (3, 16, 8)])
def test_block_diagonal_entry_points_equal_dense(dim, h, B):
    """The _blocks entry points (each tile contracts its own mode block) on
    block-diagonal weights give the dense entry points' results bit for bit
    (the skipped k-steps only add zeros), in every precision; the
    block-diagonal cross equals the dense cross on its diagonal blocks."""
    from neb_amd import _hip as H
    g = torch.Generator(device="cpu").manual_seed(7 + dim)
    C, n, wd = nw.channels(dim), 777, B * h
    ld = nw._ld(wd)
    Wd = torch.zeros((wd, wd), dtype=torch.float64)
    for m in range(B):
        Wd[m * h:(m + 1) * h, m * h:(m + 1) * h] = torch.randn((h, h), generator=g, dtype=torch.float64) / math.sqrt(h)
    bias = torch.randn(wd, generator=g, dtype=torch.float64)
    X = torch.zeros((C, n, ld), dtype=torch.float64)
    X[:, :, :wd] = torch.randn((C, n, wd), generator=g, dtype=torch.float64)
    s = torch.cuda.current_stream().cuda_stream
    for prec, p in (("bf16", H.BF16), ("f32", H.F32), ("f64", H.F64)):
        st = nw._STORE[p]
        Xd = X.to(DEV, st)
        W = torch.zeros((wd, ld), dtype=st, device=DEV)
        W[:, :wd] = Wd.to(DEV, st)
        WT = torch.zeros((wd, ld), dtype=st, device=DEV)
        WT[:, :wd] = Wd.T.to(DEV, st)
        b = bias.to(DEV, torch.float64 if p == H.F64 else torch.float32)
        outs = []
        for name, extra in (("neb_jet_layer_forward", (wd, wd)), ("neb_jet_layer_forward_blocks", (wd, h))):
            Z = torch.zeros((C, n, ld), dtype=st, device=DEV)
            Y = torch.zeros_like(Z)
            print(name, dim, h, B, prec, flush=True)
            H.call(name, p, dim, H.SIN, n, *extra, Xd.data_ptr(), ld, W.data_ptr(), ld, b.data_ptr(), 1.3, Z.data_ptr(),
                   Y.data_ptr(), ld, s)
            outs.append((Z, Y))
        assert torch.equal(outs[0][0], outs[1][0]) and torch.equal(outs[0][1], outs[1][1]), (prec, "forward")
        Zp = outs[0][0]
        back = []
        for name, extra in (("neb_jet_layer_backward", (wd, wd)), ("neb_jet_layer_backward_blocks", (wd, h))):
            A = torch.zeros((C, n, ld), dtype=st, device=DEV)
            print(name, dim, h, B, prec, flush=True)
            H.call(name, p, dim, H.SIN, n, *extra, Xd.data_ptr(), ld, WT.data_ptr(), ld, Zp.data_ptr(), ld, 0.7,
                   A.data_ptr(), ld, s)
            back.append(A)
        assert torch.equal(back[0], back[1]), (prec, "backward")
        mode = nw._CROSS[p]
        Gd = nw.cross([(Xd.view(C * n, ld), back[0].view(C * n, ld), C * n, 1.0)], wd, wd, mode)
        Gb = nw.cross([(Xd.view(C * n, ld), back[0].view(C * n, ld), C * n, 1.0)], wd, wd, mode, block=h)
        tol = 1e-12 if p == H.F64 else 1e-5       # the row splits differ with the tile count
        for m in range(B):
            sl = slice(m * h, (m + 1) * h)
            assert torch.allclose(Gd[sl, sl], Gb[sl, sl], rtol=tol, atol=tol * float(Gd.abs().max())), (prec, m)
    # split operands
    hi = X.to(torch.bfloat16)
    lo = (X - hi.double()).to(torch.bfloat16)
    whi = torch.zeros((wd, ld), dtype=torch.bfloat16)
    whi[:, :wd] = Wd.to(torch.bfloat16)
    wlo = torch.zeros_like(whi)
    wlo[:, :wd] = (Wd - whi[:, :wd].double()).to(torch.bfloat16)
    hi, lo, whi, wlo = hi.to(DEV), lo.to(DEV), whi.to(DEV), wlo.to(DEV)
    b = bias.to(DEV, torch.float32)
    outs = []
    for name, extra in (("neb_jet_layer_forward_x3", (wd, wd)), ("neb_jet_layer_forward_x3_blocks", (wd, h))):
        Z = torch.zeros((C, n, ld), dtype=torch.bfloat16, device=DEV)
        Yh, Yl = torch.zeros_like(Z), torch.zeros_like(Z)
        print(name, dim, h, B, flush=True)
        H.call(name, dim, H.SIN, n, *extra, hi.data_ptr(), lo.data_ptr(), ld, whi.data_ptr(), wlo.data_ptr(), ld,
               b.data_ptr(), 1.3, Z.data_ptr(), None, Yh.data_ptr(), Yl.data_ptr(), ld, s)
        outs.append((Z, Yh, Yl))
    for a, c in zip(*outs):
        assert torch.equal(a, c), "split forward"


@pytest.mark.parametrize("lift,dim,rm,rs,nc", [(2, 2, 2, 1, 1), (3, 3, 3, 9, 3), (4, 2, 2, 3, 1), (5, 2, 1, 2, 1)])
def test_output_backward_blocks_equals_dense(lift, dim, rm, rs, nc):
    """neb_jet_output_backward_blocks (one weight entry per output feature:
    mode m's output weights against mode m's field cotangents, given as those
    entries) gives neb_jet_output_backward's result on the same
    mode-diagonal weights as a full [ncomp * width][K] matrix:
    exactly the same products for FP32 / FP64, and within the dense path's
    rounding of the FP32 cotangents to bf16 operands for BF16."""
    from neb_amd import _hip as H
    g = torch.Generator(device="cpu").manual_seed(11 + lift)
    h, B, n = 32, 16, 333
    C, wd = nw.channels(dim), B * h
    ld, ldk = nw._ld(wd), nw._ld(B)
    Wt = torch.zeros((nc, B, h, B), dtype=torch.float64)
    Wt.diagonal(dim1=1, dim2=3).copy_(torch.randn((nc, h, B), generator=g, dtype=torch.float64))
    Wt = Wt.reshape(nc * wd, B)
    Z = torch.zeros((C, n, ld), dtype=torch.float64)
    Z[:, :, :wd] = torch.randn((C, n, wd), generator=g, dtype=torch.float64)
    s = torch.cuda.current_stream().cuda_stream
    for prec, p in (("bf16", H.BF16), ("f32", H.F32), ("f64", H.F64)):
        st = nw._STORE[p]
        ft = torch.float64 if p == H.F64 else torch.float32
        GU = torch.zeros((rm * n, ldk), dtype=ft, device=DEV)
        GD = torch.zeros((rs * n, ldk), dtype=ft, device=DEV)
        GU[:, :B] = torch.randn((rm * n, B), generator=g, dtype=torch.float64).to(DEV, ft)
        GD[:, :B] = torch.randn((rs * n, B), generator=g, dtype=torch.float64).to(DEV, ft)
        bm = torch.zeros((nc * wd, ldk), dtype=st, device=DEV)
        bm[:, :B] = Wt.to(DEV, st)
        f = torch.arange(wd)
        vec = torch.stack([Wt[c * wd + f, f // h] for c in range(nc)]).to(DEV, st).contiguous()  # each row's nonzero
        Zd = Z.to(DEV, st)
        outs = []
        for name, kk, wts, ldb in (("neb_jet_output_backward", B, bm, ldk),
                                   ("neb_jet_output_backward_blocks", h, vec, wd)):
            A = torch.zeros((C, n, ld), dtype=st, device=DEV)
            print(name, lift, prec, flush=True)
            H.call(name, p, lift, H.SIN, n, wd, kk, GU.data_ptr(), GD.data_ptr(), ldk, nw._DT[ft], 0,
                   wts.data_ptr(), wts.data_ptr(), ldb, Zd.data_ptr(), ld, 1.7, A.data_ptr(), ld, s)
            outs.append(A[:, :, :wd].double().cpu())
        tol = {"bf16": 1.5e-2, "f32": 1e-6, "f64": 1e-13}[prec]
        r = rel(outs[1], outs[0]); assert r < tol, (prec, r)


@pytest.mark.parametrize("prec", ["bf16", "f32", "f64"])
def test_first_order_layers_equal_second_order_channels(prec):
    """The jet layers on 2D first-order jets (dim code NEB_DIM_2D_JET1:
    value and gradient channels) reproduce the value and gradient channels of
    the second-order layers: the forward (dense and block-diagonal, and the
    split-precision one) never mixes the Hessian channels into them, and the
    reverse with zero Hessian adjoints adds only zeros."""
    from neb_amd import _hip as H
    g = torch.Generator(device="cpu").manual_seed(5)
    p = {"bf16": H.BF16, "f32": H.F32, "f64": H.F64}[prec]
    st = nw._STORE[p]
    n, k, h, B = 513, 64, 32, 4
    wd = B * h
    ld = nw._ld(max(k, wd))
    s = torch.cuda.current_stream().cuda_stream
    X = torch.zeros((6, n, ld), dtype=torch.float64)
    X[:, :, :k] = torch.randn((6, n, k), generator=g, dtype=torch.float64)
    W = torch.zeros((wd, ld), dtype=torch.float64)
    W[:, :k] = torch.randn((wd, k), generator=g, dtype=torch.float64) / math.sqrt(k)
    b = torch.randn(wd, generator=g, dtype=torch.float64).to(DEV, torch.float64 if p == H.F64 else torch.float32)
    Xd, Wd = X.to(DEV, st), W.to(DEV, st)
    outs = {}
    for dim, C in ((2, 6), (H.DIM_2D_JET1, 3)):
        Z = torch.zeros((C, n, ld), dtype=st, device=DEV)
        Y = torch.zeros_like(Z)
        H.call("neb_jet_layer_forward", p, dim, H.SIN, n, k, wd, Xd.data_ptr(), ld, Wd.data_ptr(), ld, b.data_ptr(),
               1.3, Z.data_ptr(), Y.data_ptr(), ld, s)
        outs[dim] = (Z, Y)
    for i in range(2):
        assert torch.equal(outs[2][i][:3], outs[H.DIM_2D_JET1][i]), (prec, "forward", i)
    # block-diagonal hidden layer on those jets, and its reverse
    Wb = torch.zeros((wd, ld), dtype=torch.float64)
    for m in range(B):
        Wb[m * h:(m + 1) * h, m * h:(m + 1) * h] = torch.randn((h, h), generator=g, dtype=torch.float64) / math.sqrt(h)
    Wbd, WbT = Wb.to(DEV, st), torch.zeros((wd, ld), dtype=torch.float64)
    WbT[:, :wd] = Wb[:, :wd].T
    WbT = WbT.to(DEV, st)
    A = torch.zeros((6, n, ld), dtype=torch.float64)
    A[:3, :, :wd] = torch.randn((3, n, wd), generator=g, dtype=torch.float64)
    Ad = A.to(DEV, st)
    res = {}
    for dim, C in ((2, 6), (H.DIM_2D_JET1, 3)):
        Y0 = outs[dim][1]
        Z = torch.zeros((C, n, ld), dtype=st, device=DEV)
        Y = torch.zeros_like(Z)
        H.call("neb_jet_layer_forward_blocks", p, dim, H.SIN, n, wd, h, Y0.data_ptr(), ld, Wbd.data_ptr(), ld,
               b.data_ptr(), 0.9, Z.data_ptr(), Y.data_ptr(), ld, s)
        An = torch.zeros((C, n, ld), dtype=st, device=DEV)
        H.call("neb_jet_layer_backward_blocks", p, dim, H.SIN, n, wd, h, Ad[:C].contiguous().data_ptr(), ld,
               WbT.data_ptr(), ld, outs[dim][0].data_ptr(), ld, 1.3, An.data_ptr(), ld, s)
        res[dim] = (Z, Y, An)
    for i, name in enumerate(("Z", "Y", "reverse")):
        a, c = res[2][i][:3].double().cpu(), res[H.DIM_2D_JET1][i].double().cpu()
        r = rel(c, a); assert r < (1e-14 if p == H.F64 else 1e-6), (prec, name, r)
    if p == H.BF16:   # the split-precision forward
        hi, lo = X.to(torch.bfloat16), (X - X.to(torch.bfloat16).double()).to(torch.bfloat16)
        whi, wlo = W.to(torch.bfloat16), (W - W.to(torch.bfloat16).double()).to(torch.bfloat16)
        hi, lo, whi, wlo = hi.to(DEV), lo.to(DEV), whi.to(DEV), wlo.to(DEV)
        ys = {}
        for dim, C in ((2, 6), (H.DIM_2D_JET1, 3)):
            Z = torch.zeros((C, n, ld), dtype=torch.bfloat16, device=DEV)
            Yh, Yl = torch.zeros_like(Z), torch.zeros_like(Z)
            H.call("neb_jet_layer_forward_x3", dim, H.SIN, n, k, wd, hi[:C].contiguous().data_ptr(),
                   lo[:C].contiguous().data_ptr(), ld, whi.data_ptr(), wlo.data_ptr(), ld, b.data_ptr(), 1.3,
                   Z.data_ptr(), None, Yh.data_ptr(), Yl.data_ptr(), ld, s)
            ys[dim] = (Z, Yh, Yl)
        for i in range(3):
            assert torch.equal(ys[2][i][:3], ys[H.DIM_2D_JET1][i]), ("x3", i)


@pytest.mark.parametrize("xd,yd", [(torch.float32, torch.bfloat16), (torch.float32, torch.float32),
                                   (torch.float64, torch.float64)])
def test_cross_modes_matches_fp64_products(xd, yd):
    """neb_cross_modes: out[m, j] = scale sum_s sign_s sum_r X_s[r, m] Y_s[r, m h + j]
    against the FP64 products of the same stored values, on ragged row
    counts, strided column views and accumulation into out (beta)."""
    g = torch.Generator(device="cpu").manual_seed(3)
    B, h, ld = 12, 16, 200
    segs, ref = [], torch.zeros((B, h), dtype=torch.float64)
    for s, rows in enumerate((777, 500, 777)):
        Xf = torch.randn((rows, 40), generator=g, dtype=torch.float64)
        Yf = torch.randn((rows, ld), generator=g, dtype=torch.float64)
        X = Xf.to(DEV, xd)[:, 5:5 + B]                       # a column view (row stride 40)
        Y = Yf.to(DEV, yd)
        sign = (-1.0) ** s * (1.0 + s)
        segs.append((X, Y, rows, sign))
        Xs, Ys = X.double().cpu(), Y.double().cpu()
        for m in range(B):
            ref[m] += sign * (Xs[:, m:m + 1] * Ys[:, m * h:(m + 1) * h]).sum(0)
    out = nw.cross_modes(segs, B, h, scale=0.5)
    r = rel(out.cpu(), 0.5 * ref); assert r < 1e-12, r
    nw.cross_modes(segs[:1], B, h, out=out, beta=2.0)
    Xs, Ys = segs[0][0].double().cpu(), segs[0][1].double().cpu()
    first = torch.stack([(Xs[:, m:m + 1] * Ys[:, m * h:(m + 1) * h]).sum(0) for m in range(B)])
    r = rel(out.cpu(), ref + first); assert r < 1e-12, r
