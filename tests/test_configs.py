"""Parity at the benchmark configurations' own sizes, through the C ABI.

Every configuration's workload runs here at its real shape on the GPU and is
compared with the FP64 oracle:

  A  2D periodic plain method: 64^2 grid, 32 features, 3x128 tanh, K=64,
     FP64 -- the whole step (fields, Grams, loss, every gradient) against
     the oracle over all 4,096 points.
  B  2D operator grid: 256^2 grid (65,536 points), 4x256 tanh, K=1024 --
     fields on a stratified 2,048-point subset against the oracle, the
     Grams against the FP64 Gram of the same fields, the FP64 generalized
     eigendecomposition against LAPACK on identical FP64 Grams, and the
     reverse pass at full size against the oracle (below); the whole
     headline step (factored Grams) against the FP64 path.
  C  2D differentiation step: tests/test_config_c.py.
  D/E  3D blob domain, 4x256 tanh, K=512 / 1024 vector potential (3K
     outputs), one full 131,072-point chunk of the 1,599,754-point set --
     fields on sampled points against the oracle, split Grams against the
     FP64 Gram of the same fields, the reverse at full chunk size, and the
     two-pass chunked step against the single pass at the chunk size.

Reverse pass at full size: the field cotangent of a Gram adjoint is
2w Phi dL/dG row by row, so zeroing the field rows (and, for the 2D
output-layer products, the last hidden jets) of every point outside a sample
makes the full-size reverse (every tile, every workgroup's walk over several
tiles, the full-width epilogues) equal to the oracle's reverse over the
sampled points alone.
"""
import math

import numpy as np
import pytest
import torch

from oracle import neb_oracle as o

pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():  # CPU-only host: the gpu tier runs on the MI355X box
    pytest.skip("needs a GPU", allow_module_level=True)

import neb_amd as nb  # noqa: E402
from neb_amd import network as nw  # noqa: E402
from neb_amd import spectral as sp  # noqa: E402
from neb_amd import synthetic as syn  # noqa: E402

DEV = "cuda"
TWO_PI = 2 * math.pi
N3D, CHUNK3D = 1_599_754, 131_072


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def cfg_2d(widths, K):
    return nb.NetConfig(dim=2, frequencies=8, widths=tuple(widths), modes=K)


def cfg_3d(K):
    return nb.NetConfig(dim=3, frequencies=6, lo=(-1.0,) * 3, hi=(1.0,) * 3, widths=(256,) * 4, modes=K)


def oracle_net(net):
    return o.Net([w.detach().cpu().double().numpy() for w in net.W],
                 [b.detach().cpu().double().numpy() for b in net.b], net.cfg.activation, [1.0] * len(net.cfg.widths))


def stratified(n, block, seed):
    """One uniformly drawn index per block of `block` consecutive points."""
    rng = np.random.default_rng(seed)
    starts = np.arange(0, n, block)
    return np.minimum(starts + rng.integers(0, block, len(starts)), n - 1)


def rows_of(sub, comp):
    """Point-major field rows p*comp + c of the points `sub`."""
    return (sub[:, None] * comp + np.arange(comp)[None, :]).ravel()


def oracle_fields(onet, cfg, x):
    out, tape = o.forward(onet, o.feature_jets(x, cfg.frequencies, cfg.lo, cfg.hi), record=True)
    if cfg.dim == 2:
        u, om = o.lift2d(out)
        return o.field_rows(u), om, tape
    u, du = o.lift3d(out)
    return o.field_rows(u), o.field_rows(du), tape


def oracle_reverse(onet, cfg, x, w, dM, dS):
    """Oracle parameter gradient of <dM, M> + <dS, S> over the points x."""
    U, D, tape = oracle_fields(onet, cfg, x)
    n, K = len(x), cfg.modes
    gU, gD = o.field_adjoint(U, w, dM), o.field_adjoint(D, w, dS)
    if cfg.dim == 2:
        adj = o.lift2d_reverse(gU.reshape(n, 2, K), gD)
    else:
        adj = o.lift3d_reverse(gU.reshape(n, 3, K), gD.reshape(n, 3, 3, K))
    return o.backward(onet, tape, adj)


def field_values(f, which):
    """FP64 values of the mass or stiffness fields (hi + lo for split planes)."""
    hi, lo = (f.mass, f.mass_lo) if which == "mass" else (f.stiff, f.stiff_lo)
    v = hi[:, :f.K].double()
    return v + lo[:, :f.K].double() if lo is not None else v


def masked(f, sub, dim):
    """Fields (the planes the reverse reads) with every row outside the
    points `sub` zeroed."""
    cm, cs = (2, 1) if dim == 2 else (3, 9)
    out = []
    for t, comp in ((f.mass, cm), (f.stiff, cs)):
        m = torch.zeros_like(t)
        idx = torch.tensor(rows_of(sub, comp), device=t.device)
        m[idx] = t[idx]
        out.append(m)
    return nw.Fields(out[0], out[1], f.K, f.layout)


def masked_tape(tape, sub):
    """The tape with the last hidden jets (hi and lo planes) zeroed outside
    the points `sub` (the 2D output-layer products and the factored
    output-layer reverse read them directly; every other reverse term at an
    unsampled point already carries a zero cotangent)."""
    idx = torch.tensor(sub, device=tape.inputs[-1].device)

    def keep(Y):
        m = torch.zeros_like(Y)
        m[:, idx] = Y[:, idx]
        return m

    lo = None if tape.inputs_lo is None else tape.inputs_lo[:-1] + [keep(tape.inputs_lo[-1])]
    return nw.Tape(tape.prec, tape.n, tape.inputs[:-1] + [keep(tape.inputs[-1])], tape.pre, lo, tape.pre_lo)


def sym_cotangents(K, seed):
    rng = np.random.default_rng(seed)
    dM = rng.normal(size=(K, K)) / math.sqrt(K)
    dS = rng.normal(size=(K, K)) / math.sqrt(K)
    return dM + dM.T, dS + dS.T


def check_grads(net, g, gW, gb, tol, what):
    GW, Gb = net.split(g)
    rw = [rel(a.detach().cpu(), b) for a, b in zip(GW, gW)]
    rb = [rel(a.detach().cpu(), b) for a, b in zip(Gb[:-1], gb[:-1])]
    print(f"{what}: W rel " + " ".join(f"{r:.2e}" for r in rw) + "; b rel " + " ".join(f"{r:.2e}" for r in rb),
          flush=True)
    for l, r in enumerate(rw):
        assert r < tol, (what, "W", l, r)
    for l, r in enumerate(rb):
        assert r < tol, (what, "b", l, r)


# --------------------------------------------------------------------------
# A: 2D periodic plain method, FP64, oracle size -- the whole step
# --------------------------------------------------------------------------
def test_config_A_full_step_fp64():
    cfg = cfg_2d((128, 128, 128), 64)
    net = nb.EigenfluidNet(cfg, seed=0, device=DEV, master_dtype=torch.float64)
    onet = o.init_net(cfg.sizes, 0)
    # initialisation: the same Xavier-normal stream, bit for bit
    for a, b in zip(net.W + net.b, onet.W + onet.b):
        assert np.array_equal(a.cpu().numpy(), b)
    x = o.periodic_grid(64)
    w = (TWO_PI / 64) ** 2
    U, D, tape = oracle_fields(onet, cfg, x)
    M, S = o.gram(U, w), o.gram(D, w)
    L, dM, dS = o.spectral_loss(M, S, 1e-2, 0.0)
    gW, gb = oracle_reverse(onet, cfg, x, w, dM, dS)
    res = nb.compute(net, torch.tensor(x), w, prec="f64", gram_mode="f64", lam=1e-2, gate="eig")
    f = res["fields"]
    assert rel(f.U.cpu(), U) < 1e-12 and rel(f.D.cpu(), D) < 1e-12
    r = rel(res["M"].cpu(), M); assert r < 1e-12, r
    r = rel(res["S"].cpu(), S); assert r < 1e-12, r
    assert abs(float(res["loss"]) - L) < 1e-9 * abs(L), (float(res["loss"]), L)
    check_grads(net, res["grads"], gW, gb, 1e-9, "config A")
    # FP64 generalized eigendecomposition of the full-rank K = 64 pencil
    vals, V = sp.generalized_eigh(res["S"], res["M"])
    ref_vals, _ = o.generalized_eigh(S, M)
    r = rel(vals.cpu(), ref_vals); assert r < 1e-10, r
    Vn = V.cpu().numpy()
    assert np.abs(Vn.T @ M @ Vn - np.eye(64)).max() < 1e-9


# --------------------------------------------------------------------------
# B: 2D operator grid, 65,536 points, 4x256, K=1024
# --------------------------------------------------------------------------
@pytest.fixture(scope="module")
def config_b():
    cfg = cfg_2d((256,) * 4, 1024)
    net = nb.EigenfluidNet(cfg, seed=0, device=DEV, master_dtype=torch.float32)
    x = syn.periodic_grid(256)
    sub = stratified(len(x), 32, seed=11)        # 2,048 points, one per 32-point block
    U, D, _ = oracle_fields(oracle_net(net), cfg, x[sub])
    return cfg, net, x, sub, U, D


FIELD_TOL = {"f64": 1e-10, "f32": 2e-5, "bf16x3": 2e-5, "bf16": 3e-2}


@pytest.mark.parametrize("prec", ["bf16x3", "f32", "bf16", "f64"])
def test_config_B_fields_and_grams(config_b, prec):
    cfg, net, x, sub, U, D = config_b
    K, w = cfg.modes, (TWO_PI / 256) ** 2
    tape = nw.trunk_forward(net, torch.tensor(x), nw.PRECISIONS[prec])
    f = nw.output_fields(net, tape, split=prec in ("bf16", "bf16x3"))
    del tape
    Uv, Dv = field_values(f, "mass"), field_values(f, "stiff")
    ru = rel(Uv[torch.tensor(rows_of(sub, 2), device=DEV)].cpu(), U)
    rd = rel(Dv[torch.tensor(sub, device=DEV)].cpu(), D)
    assert ru < FIELD_TOL[prec] and rd < FIELD_TOL[prec], (prec, ru, rd)
    # Grams of the step's mode against the FP64 Gram of the same fields
    gmode = {"bf16x3": "bf16x3", "bf16": "bf16x3", "f32": "f32", "f64": "f64"}[prec]
    gtol = {"bf16x3": 1e-5, "f32": 2e-6, "f64": 1e-13}[gmode]
    for which, vals in (("mass", Uv), ("stiff", Dv)):
        G = nw.field_gram(f, which, w, gmode)
        G64 = nw.gram(vals.contiguous(), K, w, "f64")
        r = float(torch.linalg.norm(G - G64) / torch.linalg.norm(G64))
        assert r < gtol, (prec, which, r)
    if prec == "bf16x3":
        # the subset's split Gram against the oracle's Gram of the same points
        hi = f.mass[torch.tensor(rows_of(sub, 2), device=DEV)]
        lo = f.mass_lo[torch.tensor(rows_of(sub, 2), device=DEV)]
        Gs = nw.gram(hi, K, w, "bf16x3", lo=lo)
        r = rel(Gs.cpu(), o.gram(U, w)); assert r < 3e-5, r
    del f, Uv, Dv
    torch.cuda.empty_cache()


def test_config_B_generalized_eigh_matches_lapack(config_b):
    """The FP64 generalized eigendecomposition (Cholesky reduction on the GPU)
    against LAPACK (scipy.linalg.eigh) on identical FP64 Grams of the config-B
    step, and the full-rank gate on the rank-deficient mass (rank <= 256 < K)."""
    cfg, net, x, *_ = config_b
    K, w = cfg.modes, (TWO_PI / 256) ** 2
    tape = nw.trunk_forward(net, torch.tensor(x), nw.PRECISIONS["bf16x3"])
    f = nw.output_fields(net, tape, split=True)
    M, S = nw.field_grams(f, w, "bf16x3")
    del tape, f
    Mn, Sn = M.cpu().numpy(), S.cpu().numpy()
    with pytest.raises(ArithmeticError):
        sp.generalized_eigh(S, M)
    with pytest.raises(ArithmeticError):
        o.generalized_eigh(Sn, Mn)
    ridge = 1e-4 * float(np.mean(np.diag(Mn)))
    Mr = Mn + ridge * np.eye(K)
    vals, V = sp.generalized_eigh(S, torch.tensor(Mr, device=DEV))
    ref, _ = o.generalized_eigh(Sn, Mr)
    v = vals.cpu().numpy()
    scale = np.abs(ref).max()
    assert np.abs(v - ref).max() < 1e-9 * scale, np.abs(v - ref).max() / scale
    top = ref[-256:]
    assert np.abs(v[-256:] - top).max() < 1e-8 * np.abs(top).max()
    Vn = V.cpu().numpy()
    Mrs = 0.5 * (Mr + Mr.T)
    assert np.abs(Vn.T @ Mrs @ Vn - np.eye(K)).max() < 1e-7
    resid = np.linalg.norm(0.5 * (Sn + Sn.T) @ Vn - Mrs @ Vn * v[None, :]) / np.linalg.norm(Sn)
    assert resid < 1e-10, resid


REV_TOL = {"f64": 1e-9, "f32": 2e-5, "bf16": 4e-2, "bf16x3": 4e-2}
# the split tape's factored output-layer reverse (FP64 K x K algebra, split GEMMs)
REV_TOL_FACTORED = 1e-2


@pytest.mark.parametrize("prec", ["bf16x3", "bf16x3-factored", "bf16", "f32", "f64"])
def test_config_B_reverse_full_size(prec):
    """The reverse pass over all 65,536 points (4x256, K=1024: every async
    tile walk and full-width epilogue) against the oracle's over a stratified
    2,048-point sample, through fields (and last hidden jets) zeroed outside
    the sample; "bf16x3-factored" takes the factored output-layer reverse
    (the 3D default) on the 2D split tape."""
    cfg = cfg_2d((256,) * 4, 1024)
    if prec == "bf16x3-factored":
        cfg.split_reverse, prec = "factored", "bf16x3"
    master = torch.float64 if prec == "f64" else torch.float32
    net = nb.EigenfluidNet(cfg, seed=0, device=DEV, master_dtype=master)
    rng = np.random.default_rng(3)
    with torch.no_grad():
        for b in net.b:       # nonzero biases: every layer's bias column sum is exercised
            b.copy_(torch.as_tensor(rng.normal(size=b.shape) * 0.1, dtype=master))
    net.touch()
    x = syn.periodic_grid(256)
    w = (TWO_PI / 256) ** 2
    sub = stratified(len(x), 32, seed=12)
    dM, dS = sym_cotangents(cfg.modes, 7)
    gW, gb = oracle_reverse(oracle_net(net), cfg, x[sub], w, dM, dS)
    tape = nw.trunk_forward(net, torch.tensor(x), nw.PRECISIONS[prec])
    f = nw.output_fields(net, tape, split=prec in ("bf16", "bf16x3"))
    g = nw.backward_fields(net, masked_tape(tape, sub), masked(f, sub, cfg.dim), torch.tensor(dM, device=DEV),
                           torch.tensor(dS, device=DEV), w)
    fact = prec == "bf16x3" and cfg.split_reverse == "factored"
    check_grads(net, g, gW, gb, REV_TOL_FACTORED if fact else REV_TOL[prec],
                f"config B reverse {prec}{' factored' if fact else ''}")
    del tape, f, g
    torch.cuda.empty_cache()


def test_config_B_headline_step_factored_grams_vs_fp64():
    """The headline step at full size (65,536-point grid, 4x256 tanh, K=1024,
    split-precision jets, field Grams from the lift-weighted channel Grams,
    factored reverse) against the FP64 path on the same grid, weights and
    ridge: Grams within 1e-5 (the split-precision Gram bound), loss within
    1e-4, whole-step gradient within 5e-3 relative L2 (measured 1.5e-3; the
    objective amplifies the Grams' 7e-6 by up to cond(M + ridge I))."""
    cfg = cfg_2d((256,) * 4, 1024)
    net = nb.EigenfluidNet(cfg, seed=0, device=DEV, master_dtype=torch.float32)
    x = torch.tensor(syn.periodic_grid(256), dtype=torch.float32, device=DEV)
    w = syn.grid_weight(256)

    def trel(a, b):
        return float(torch.linalg.norm(a.double() - b.double()) / torch.linalg.norm(b.double()))

    M0 = nb.compute(net, x, w, prec="f64", gram_mode="f64", training=False, gate="cholesky", ridge=1.0)["M"]
    ridge = 1e-4 * float(torch.diagonal(M0).mean())
    del M0
    ref = nb.compute(net, x, w, prec="f64", gram_mode="f64", ridge=ridge, gate="cholesky")
    res = nb.compute(net, x, w, prec="bf16x3", gram_mode="bf16x3", ridge=ridge, gate="cholesky", grams="factored")
    assert trel(res["M"], ref["M"]) < 1e-5 and trel(res["S"], ref["S"]) < 1e-5
    assert abs(float(res["loss"]) - float(ref["loss"])) < 1e-4 * abs(float(ref["loss"]))
    g, g64 = res["grads"].double(), ref["grads"].double()
    err = float(torch.linalg.norm(g - g64) / torch.linalg.norm(g64))
    assert err < 5e-3, err
    # the same step through the Grams of the K-column fields agrees with it
    fld = nb.compute(net, x, w, prec="bf16x3", gram_mode="bf16x3", ridge=ridge, gate="cholesky")
    assert trel(fld["M"], res["M"]) < 2e-5
    del ref, res, fld
    torch.cuda.empty_cache()


# --------------------------------------------------------------------------
# D / E: 3D blob domain, one full 131,072-point chunk, K = 512 / 1024
# --------------------------------------------------------------------------
@pytest.fixture(scope="module")
def chunk3d():
    # the first rejection batch (2^20 draws) accepts more than 131,072 points,
    # so these are exactly the first chunk of the 1,599,754-point set
    pts, _, _ = syn.blob_domain(CHUNK3D, seed=0)
    return pts


@pytest.mark.parametrize("K", [512, 1024])
def test_config_DE_chunk_fields_and_grams(chunk3d, K):
    cfg = cfg_3d(K)
    net = nb.EigenfluidNet(cfg, seed=0, device=DEV, master_dtype=torch.float32)
    sub = stratified(CHUNK3D, 512, seed=13)      # 256 points
    U, D, _ = oracle_fields(oracle_net(net), cfg, chunk3d[sub])
    w = 1.0 / N3D
    for prec in ("bf16x3", "f64"):
        tape = nw.trunk_forward(net, torch.tensor(chunk3d), nw.PRECISIONS[prec])
        f = nw.output_fields(net, tape, split=prec == "bf16x3")
        del tape
        Uv = field_values(f, "mass")
        ru = rel(Uv[torch.tensor(rows_of(sub, 3), device=DEV)].cpu(), U)
        del Uv
        Dv = field_values(f, "stiff")
        rd = rel(Dv[torch.tensor(rows_of(sub, 9), device=DEV)].cpu(), D)
        del Dv
        assert ru < FIELD_TOL[prec] and rd < FIELD_TOL[prec], (K, prec, ru, rd)
        if prec == "bf16x3":
            for which in ("mass", "stiff"):
                G = nw.field_gram(f, which, w, "bf16x3")
                vals = field_values(f, which).contiguous()
                G64 = nw.gram(vals, K, w, "f64")
                del vals
                r = float(torch.linalg.norm(G - G64) / torch.linalg.norm(G64))
                assert r < 1e-5, (K, which, r)
        del f
        torch.cuda.empty_cache()


@pytest.mark.parametrize("prec", ["bf16x3", "f64"])
def test_config_E_chunk_reverse_full_size(chunk3d, prec):
    """3D reverse over a full 131,072-point chunk (K=1024, 3072 outputs)
    against the oracle over 128 sampled points (fields zeroed elsewhere)."""
    cfg = cfg_3d(1024)
    master = torch.float64 if prec == "f64" else torch.float32
    net = nb.EigenfluidNet(cfg, seed=0, device=DEV, master_dtype=master)
    rng = np.random.default_rng(4)
    with torch.no_grad():
        for b in net.b:
            b.copy_(torch.as_tensor(rng.normal(size=b.shape) * 0.1, dtype=master))
    net.touch()
    sub = stratified(CHUNK3D, 1024, seed=14)
    w = 1.0 / N3D
    dM, dS = sym_cotangents(cfg.modes, 8)
    gW, gb = oracle_reverse(oracle_net(net), cfg, chunk3d[sub], w, dM, dS)
    tape = nw.trunk_forward(net, torch.tensor(chunk3d), nw.PRECISIONS[prec])
    f = nw.output_fields(net, tape, split=prec == "bf16x3")
    g = nw.backward_fields(net, masked_tape(tape, sub), masked(f, sub, cfg.dim), torch.tensor(dM, device=DEV),
                           torch.tensor(dS, device=DEV), w)
    # 3D split tapes take the factored output-layer reverse
    check_grads(net, g, gW, gb, REV_TOL_FACTORED if prec == "bf16x3" else REV_TOL[prec], f"config E reverse {prec}")
    del tape, f, g
    torch.cuda.empty_cache()


@pytest.mark.parametrize("prec,tol_g,tol_grad", [("f64", 1e-12, 1e-9), ("bf16x3", 1e-6, 1e-3)])
def test_config_D_chunked_equals_single_pass(chunk3d, prec, tol_g, tol_grad):
    """The two-pass chunked step at the 131,072-point chunk size over one
    full and one ragged chunk (26,890 points, the size of config E's last
    chunk) against the single pass over the same 157,962 points (K=512,
    full-rank mass, no ridge): Grams, and the gradient of the chunked
    recompute + reverse for the single pass's cotangents (on the split path
    to 1e-3: the reverse's fp32 partial sums over points, of per-point terms
    that largely cancel at these cotangents, run in another order; FP64
    agrees to 1e-9).  In FP64 the whole
    step (its own objective included) agrees too; on split Grams the
    objective's cotangents move by up to cond(M) times the Grams' 1e-7
    difference, which is the objective's conditioning, not the chunking."""
    cfg = cfg_3d(512)
    master = torch.float64 if prec == "f64" else torch.float32
    net = nb.EigenfluidNet(cfg, seed=0, device=DEV, master_dtype=master)
    extra, _, _ = syn.blob_domain(26_890, seed=5)
    x = torch.tensor(np.concatenate([chunk3d, extra]))
    w = 1.0 / N3D
    ref = nb.compute(net, x, w, prec=prec, gram_mode=prec, gate="cholesky")
    refM, refS, refL, refg = ref["M"].clone(), ref["S"].clone(), float(ref["loss"]), ref["grads"].clone()
    pair = torch.cat([ref["dM"], ref["dS"]], 1)
    del ref
    torch.cuda.empty_cache()
    same = lambda M, S: (torch.tensor(refL), pair[:, :512], pair[:, 512:])  # noqa: E731
    res = nb.chunked_step(net, x, w, prec=prec, gram_mode=prec, chunk=CHUNK3D, solver=same)
    r = float(torch.linalg.norm(res["M"] - refM) / torch.linalg.norm(refM)); assert r < tol_g, r
    r = float(torch.linalg.norm(res["S"] - refS) / torch.linalg.norm(refS)); assert r < tol_g, r
    r = float(torch.linalg.norm(res["grads"] - refg) / torch.linalg.norm(refg)); assert r < tol_grad, r
    del res
    if prec == "f64":
        full = nb.chunked_step(net, x, w, prec=prec, gram_mode=prec, chunk=CHUNK3D, solver=nb.SpectralSolver(1e-2))
        assert abs(float(full["loss"]) - refL) < 1e-10 * abs(refL)
        r = float(torch.linalg.norm(full["grads"] - refg) / torch.linalg.norm(refg)); assert r < tol_grad, r
        del full
    torch.cuda.empty_cache()
