"""Benchmark: 2D periodic operator grid, K=1024 spectral training step.

Workload (one step): 256^2 uniform grid on [0,2pi)^2 (65,536 points), sin/cos
features at frequencies 1..8 (32 features), 4 hidden tanh layers of width
256, K=1024 stream-function modes, Xavier-normal weights (seed 0):
second-order jet forward -> lifted fields u=(psi_y,-psi_x) [131,072 x 1024]
and omega=-lap psi [65,536 x 1024] -> FP64 mass and stiffness Grams
Phi^T W Phi (BF16 hi+lo split MFMA) -> FP64 spectral objective
tr((M+eps I)^-1 S) + 1e-2 ||M-I||_F^2 with its closed-form adjoint ->
handwritten reverse pass to all parameters -> Adam.

metric: point*modes/s = N*K / step time (whole job; inputs resident in HBM).
Also reported: Gram TFLOP/s (algorithmic R*K*(K+1)) for the BF16 hi+lo and
FP32 MFMA modes with their relative Frobenius error against the FP64 Gram of
the same fields, the FP32 step, per-stage times, the roofline of the
dominant kernel and the FP64 NumPy oracle timed on the host.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N_GRID = 256
K_MODES = 1024
WIDTHS = (256, 256, 256, 256)
FREQS = 8
PEAK = {"bf16": 2.5e15, "f32": 157.3e12, "f64": 78.6e12}
HBM_PEAK = 8.0e12


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the host oracle baseline")
    ap.add_argument("--no-f32", action="store_true", help="skip the FP32 comparison step")
    ap.add_argument("--no-3d", action="store_true", help="skip the 3D K=1024 (config E) leg of the default run")
    ap.add_argument("--step-only", action="store_true",
                    help="skip the Gram-mode statistics and the FP64 accuracy legs (profiler runs)")
    ap.add_argument("--kernel", default=None,
                    choices=["hidden_fwd", "hidden_bwd", "out_fields", "out_bwd", "gram", "gram_f32in", "cross_dw",
                             "channel_gram",
                             "copy", "hidden_fwd_x3", "out_fields_x3", "products"],
                    help="run one kernel of the config-B step --steps times (for counter profiling)")
    ap.add_argument("--precision", default="bf16x3", choices=["bf16x3", "bf16"],
                    help="config B headline step: bf16x3 = split-precision jets and fields (bf16 hi + lo planes, "
                         "three bf16 MFMA products per term; Grams within 1e-5 of the FP64 path), bf16 = plain "
                         "BF16 jets (Grams within ~5e-3); the other one is reported beside it")
    ap.add_argument("--config", default="B", choices=["B", "D", "E", "modes2d", "modes3d"],
                    help="B: 2D 256^2 grid K=1024 (default); D: 3D blob domain, 1,599,754 points, K=512, "
                         "forward + Grams + FP64 generalized eigendecomposition; E: the same points, K=1024, "
                         "training step + 20-target full-mass projection; modes2d / modes3d: only the "
                         "per-mode SIREN legs of the default run (the reference's 2D periodic potential and 3D "
                         "architecture updates), one JSON line each (profiler runs)")
    return ap.parse_args()


class StageTimer:
    """HIP events on the launch stream (torch's current stream, which every
    libnebhip call uses)."""

    def __init__(self, torch):
        self.torch = torch
        self.events = {}

    def start(self, name):
        e = self.torch.cuda.Event(enable_timing=True)
        e.record()
        self.events.setdefault(name, []).append([e, None])

    def stop(self, name):
        e = self.torch.cuda.Event(enable_timing=True)
        e.record()
        self.events[name][-1][1] = e

    def ms(self):
        return {k: [a.elapsed_time(b) for a, b in v] for k, v in self.events.items()}


# committed rocprofv3 counter summaries (FETCH_SIZE / WRITE_SIZE per launch)
# of the kernel classes of the 2D K=1024 step
PROFILE_OF = {"neb_jet_layer_forward_x3 k=256": "hidden_fwd_x3_k1024_pmc.md",
              "neb_jet_output_fields_x3": "out_fields_x3_k1024_pmc.md",
              "neb_gram_split rows=131072": "gram_split_k1024_pmc.md",
              "neb_jet_output_backward": "out_bwd_k1024_pmc.md",
              "neb_jet_layer_backward k=256": "hidden_bwd_k1024_pmc.md"}
RIDGE_FLOP_PER_BYTE = 2.5e15 / 8.0e12   # bf16 dense MFMA peak / HBM peak
_ES = {0: 4, 1: 2, 2: 8}                 # NEB_F32, NEB_BF16, NEB_F64 element bytes


def launch_work(name, a):
    """(kernel class, algorithmic HBM bytes, useful flops) of one libnebhip
    launch from its arguments (the C ABI order of include/neb_hip.h), or
    (class, None, None) for launches without a roofline figure.  Bytes count
    each operand read and each result written once; flops count the useful
    products (the split-precision forms issue three per product)."""
    if name == "neb_jet_layer_forward_x3":
        dim, n, k, width = a[0], a[2], a[3], a[4]
        C = 6 if dim == 2 else 10
        return f"{name} k={k}", n * C * (k * 4 + width * 6), 2 * C * n * k * width
    if name == "neb_jet_layer_forward":
        prec, dim, n, k, width = a[0], a[1], a[3], a[4], a[5]
        C, es = (6 if dim == 2 else 10), _ES[prec]
        return f"{name} k={k}", n * C * (k + 2 * width) * es, 2 * C * n * k * width
    if name in ("neb_jet_output_fields_x3", "neb_jet_output_fields_split", "neb_jet_output_fields"):
        if name == "neb_jet_output_fields":
            prec, dim, n, k, K = a[0], a[1], a[2], a[3], a[4]
            es_in, es_out = _ES[prec], (8 if prec == 2 else 4)
        else:
            dim, n, k, K = a[0], a[1], a[2], a[3]
            es_in, es_out = (4 if name.endswith("_x3") else 2), 4
        nch, rows, terms = {2: (4, 3, 4), 3: (9, 12, 24), 4: (5, 5, 5)}[dim]
        return name, n * (nch * k * es_in + rows * K * es_out), 2 * terms * n * k * K
    if name == "neb_gram_split":
        rows, K = a[2], a[3]
        return f"{name} rows={rows}", rows * K * 4, rows * K * (K + 1)
    if name == "neb_gram":
        rows, K, dt = a[2], a[3], a[5]
        return f"{name} rows={rows}", rows * K * _ES[dt], rows * K * (K + 1)
    if name == "neb_jet_output_backward":
        prec, dim, n, width, K, fdt = a[0], a[1], a[3], a[4], a[5], a[9]
        C, es = (10 if dim == 3 else 6), _ES[prec]
        rows, terms = {2: (3, 3), 3: (12, 24), 4: (5, 5)}[dim]
        return name, n * (rows * K * _ES[fdt] + 2 * C * width * es), 2 * terms * n * width * K
    if name == "neb_jet_layer_backward":
        prec, dim, n, k, width = a[0], a[1], a[3], a[4], a[5]
        C, es = (6 if dim == 2 else 10), _ES[prec]
        return f"{name} k={k}", n * C * (width + 2 * k) * es, 2 * C * n * k * width
    if name == "neb_cross":
        nseg, xd, yd, m, ncol, sym = a[1], a[8], a[9], a[10], a[11], a[13]
        rows = sum(a[6][i] for i in range(nseg))
        flops = rows * m * (m + 1) if sym else 2 * rows * m * ncol
        return f"{name} {m}x{ncol}", rows * (m * _ES[xd] + ncol * _ES[yd]), flops
    return name, None, None


class LaunchTimer:
    """Launch recorder for neb_amd._hip.set_launch_recorder: HIP events on
    the launching (current) stream around every libnebhip call, with the
    call's kernel class and algorithmic work."""

    def __init__(self, torch):
        self.torch = torch
        self.recs = []

    def __call__(self, name, args):
        e0 = self.torch.cuda.Event(enable_timing=True)
        e0.record()

        def done():
            e1 = self.torch.cuda.Event(enable_timing=True)
            e1.record()
            self.recs.append((launch_work(name, args), e0, e1))
        return done

    def table(self, steps, step_ms):
        """Per kernel class: launches and ms per step, share of the step,
        per-launch time and roofline (bound by the class's arithmetic
        intensity against the bf16 ridge)."""
        out = {}
        for (cls, nbytes, flops), e0, e1 in self.recs:
            r = out.setdefault(cls, dict(launches=0, ms=0.0, bytes=nbytes, flops=flops))
            r["launches"] += 1
            r["ms"] += e0.elapsed_time(e1)
        rows = []
        for cls, r in out.items():
            per = r["ms"] / r["launches"]
            row = dict(kernel=cls, launches_per_step=r["launches"] / steps, ms_per_step=r["ms"] / steps,
                       share_of_step=r["ms"] / steps / step_ms, ms_per_launch=per)
            if r["bytes"]:
                intensity = r["flops"] / r["bytes"]
                if intensity < RIDGE_FLOP_PER_BYTE:
                    bw = r["bytes"] / (per * 1e-3)
                    row.update(bound="hbm", achieved=bw / 1e9, peak=HBM_PEAK / 1e9, unit="GB/s", frac=bw / HBM_PEAK,
                               bytes_per_launch=r["bytes"], flops_per_launch=r["flops"])
                else:
                    fl = r["flops"] / (per * 1e-3)
                    row.update(bound="mfma", achieved=fl / 1e12, peak=PEAK["bf16"] / 1e12, unit="TFLOP/s",
                               frac=fl / PEAK["bf16"], bytes_per_launch=r["bytes"], flops_per_launch=r["flops"])
            rows.append(row)
        rows.sort(key=lambda r: -r["ms_per_step"])
        return rows


def make_step(nb, nw, sp, net, x, w, prec, gram_mode, ridge, timer=None, solver=None, reverse_prec=None,
              grams="fields"):
    """One training step of config B.  grams="factored" (split-precision
    jets): no field matrices -- stacked last-hidden channels, their split
    channel Grams, the field Grams assembled from them in FP64, and the
    factored reverse on the same stacked channels."""
    import torch
    K = net.cfg.modes
    solver = solver if solver is not None else sp.SpectralSolver(1e-2, ridge, deferred=True)

    def mark(fn, name):
        if timer is None:
            return fn()
        timer.start(name)
        r = fn()
        timer.stop(name)
        return r

    def step_factored():
        tape = mark(lambda: nw.trunk_forward(net, x, nw.PRECISIONS[prec]), "trunk_forward")
        fp = mark(lambda: nw.factored_products(net, tape), "channel_grams")
        M, S = mark(lambda: nw.factored_field_grams(net, fp.Gh, w), "field_grams")
        loss, dM, dS = mark(lambda: solver(M, S), "objective_fp64")
        g = mark(lambda: nw.backward_fields(net, tape, None, dM, dS, w, symmetric=True, products=fp), "reverse")
        mark(lambda: net.adam(g, 1e-4, gate=solver.gate, on_gate=solver.raise_gate), "adam")
        return loss

    def step():
        p = nw.PRECISIONS[prec]
        split = nw.split_fields_for(prec, gram_mode)
        tape = mark(lambda: nw.trunk_forward(net, x, p), "trunk_forward")
        fields = mark(lambda: nw.output_fields(net, tape, split=split), "output_fields")
        ev = None
        if timer is not None:
            ev = (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
            timer.events.setdefault("gram_mass_launch", []).append(list(ev))
        M, S = mark(lambda: nw.field_grams(fields, w, gram_mode, mass_events=ev), "grams")
        if reverse_prec is not None:  # FP32 fields and Grams, BF16 reverse on the converted tape
            tape = nw.convert_tape(tape, nw.PRECISIONS[reverse_prec])
        # output-layer products on the side stream, overlapping the FP64 objective
        prod = nw.output_products(net, tape, fields)
        loss, dM, dS = mark(lambda: solver(M, S), "objective_fp64")
        g = mark(lambda: nw.backward_fields(net, tape, fields, dM, dS, w, symmetric=True, products=prod), "reverse")
        # the positive-definiteness / condition gates and the gradient gate are
        # evaluated on the device; adam synchronises once, after queueing
        mark(lambda: net.adam(g, 1e-4, gate=solver.gate, on_gate=solver.raise_gate), "adam")
        return loss

    return step_factored if grams == "factored" else step


def profile_traffic(name):
    """HBM bytes per launch of a kernel from its committed counter summary
    (the `traffic_bytes_per_launch:` line: FETCH_SIZE and WRITE_SIZE of
    rocprofv3 --pmc runs scaled by the copy-kernel calibration); None if the
    summary is absent."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", name)
    try:
        with open(path) as f:
            for line in f:
                if line.strip().startswith("traffic_bytes_per_launch:"):
                    return int(float(line.split(":", 1)[1].strip().split()[0]))
    except OSError:
        return None
    return None


def cpu_baseline(seconds_target=12.0, single_seconds=8.0, batch=1024):
    """FP64 NumPy oracle of the same step on a bounded sample of the grid
    (consecutive `batch`-point batches: large enough that every BLAS call of
    the oracle spreads over all host cores), timed on this host with all of
    its cores and again with BLAS pinned to one thread (the reference's own
    pinning, experiments/metaball/run.py:5-6); scaled to point*modes/s."""
    import numpy as np
    from oracle import neb_oracle as o
    pts = o.periodic_grid(N_GRID)
    sizes = [4 * FREQS] + list(WIDTHS) + [K_MODES]
    net = o.init_net(sizes, 0)
    w = (2 * math.pi / N_GRID) ** 2
    dM = np.eye(K_MODES) * 1e-3  # the objective's K x K algebra is size-independent; fixed cotangents

    def run(seconds):
        t_total, n_done = 0.0, 0
        while True:
            start = n_done % len(pts)
            x = pts[start: start + batch]
            t0 = time.perf_counter()
            x0 = o.feature_jets(x, FREQS, (0.0, 0.0), (2 * math.pi, 2 * math.pi))
            out, tape = o.forward(net, x0, record=True)
            u, om = o.lift2d(out)
            Pm, Ps = o.field_rows(u), om
            o.gram(Pm, w), o.gram(Ps, w)
            gPm, gPs = o.field_adjoint(Pm, w, dM), o.field_adjoint(Ps, w, dM)
            adj = o.lift2d_reverse(gPm.reshape(len(x), 2, K_MODES), gPs)
            o.backward(net, tape, adj)
            t_total += time.perf_counter() - t0
            n_done += len(x)
            if t_total >= seconds:
                return n_done, t_total

    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count()
    threads = None
    n_all, t_all = run(seconds_target)
    single = None
    try:
        from threadpoolctl import threadpool_info, threadpool_limits
        threads = max((d.get("num_threads", 0) for d in threadpool_info()), default=None)
        with threadpool_limits(limits=1):
            n_1, t_1 = run(single_seconds)
        single = n_1 * K_MODES / t_1
    except ImportError:
        pass
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), None)
    except OSError:
        pass
    return dict(value=n_all * K_MODES / t_all, unit="point*modes/s", cores=threads or cores, kind="port",
                sample=f"{n_all} of the 65,536 grid points in {batch}-point batches, full jet fwd + Grams + "
                       f"reverse at K=1024 (FP64 NumPy oracle; K x K objective excluded), {t_all:.1f} s",
                single_thread_value=single, single_thread_sample_s=single_seconds, affinity_cores=cores,
                blas_threads=threads, cpu_model=model)


def bench_kernel(args):
    """Launch one kernel of the config-B step `--steps` times on the real
    shapes (65,536 points, 4x256, K=1024), timed with HIP events; for
    rocprofv3 counter collection.  `copy` is a plain device copy of the mass
    field matrix (512 MiB read + 512 MiB written), the calibration kernel for
    FETCH_SIZE / WRITE_SIZE."""
    import numpy as np
    import torch

    import neb_amd as nb
    from neb_amd import _hip as H
    from neb_amd import network as nw
    from neb_amd import synthetic as syn

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    x = torch.tensor(syn.periodic_grid(N_GRID), dtype=torch.float32, device=dev)
    N, K = x.shape[0], K_MODES
    w = syn.grid_weight(N_GRID)
    cfg = nb.NetConfig(dim=2, frequencies=FREQS, widths=WIDTHS, modes=K)
    net = nb.EigenfluidNet(cfg, seed=0, device=dev)
    p = nw.PRECISIONS["bf16"]
    tape = nw.trunk_forward(net, x, p)
    fields = nw.output_fields(net, tape, split=True)
    stg = net.staged(p)
    s = torch.cuda.current_stream().cuda_stream
    C = 6
    Y, Z = tape.inputs[2], tape.pre[2]
    out = torch.empty_like(Y)
    dM = torch.eye(K, dtype=torch.float64, device=dev) * 1e-3
    Bm = torch.zeros((256, K), dtype=torch.bfloat16, device=dev)
    A = torch.empty_like(Y)
    f32in = nw.output_fields(net, tape) if args.kernel == "gram_f32in" else None
    copy_src = torch.empty(fields.mass.shape, dtype=torch.float32, device=dev) if args.kernel == "copy" else None
    copy_dst = torch.empty_like(copy_src) if copy_src is not None else None
    tx3 = nw.trunk_forward(net, x, nw.PRECISIONS["bf16x3"]) if args.kernel.endswith("_x3") else None
    if tx3 is not None:
        sx3 = net.staged_split()
        O2, O2l = torch.empty_like(Y), torch.empty_like(Y)
    fp = None
    if args.kernel == "channel_gram":   # the headline step's mass channel Gram (65,536 x 512 stacked jets)
        fp = nw.factored_products(net, nw.trunk_forward(net, x, nw.PRECISIONS["bf16x3"]), grams=False)
        km = len(fp.chans[0]) * WIDTHS[-1]

    def run():
        k = args.kernel
        if k == "hidden_fwd":
            H.call("neb_jet_layer_forward", p, 2, H.TANH, N, 256, 256, Y.data_ptr(), 256, stg["W"][2].data_ptr(),
                   256, stg["b"][2].data_ptr(), 1.0, out.data_ptr(), A.data_ptr(), 256, s)
        elif k == "hidden_bwd":
            H.call("neb_jet_layer_backward", p, 2, H.TANH, N, 256, 256, Y.data_ptr(), 256, stg["WT"][2].data_ptr(),
                   256, Z.data_ptr(), 256, 1.0, out.data_ptr(), 256, s)
        elif k == "out_fields":
            nw.output_fields(net, tape, split=True)
        elif k == "out_bwd":
            H.call("neb_jet_output_backward", p, 2, H.TANH, N, 256, K, fields.mass.data_ptr(),
                   fields.stiff.data_ptr(), K, H.BF16, 0, Bm.data_ptr(), Bm.data_ptr(), K, Z.data_ptr(), 256, 1.0,
                   out.data_ptr(), 256, s)
        elif k == "gram":
            nw.field_gram(fields, "mass", w, "bf16x3")
        elif k == "gram_f32in":
            nw.gram(f32in.mass, K, w, "bf16x3")
        elif k == "cross_dw":
            nw.cross([(Y.view(C * N, 256), Z.view(C * N, 256), C * N, 1.0)], 256, 256, nw._CROSS[p])
        elif k == "copy":
            copy_dst.copy_(copy_src)
        elif k == "hidden_fwd_x3":
            H.call("neb_jet_layer_forward_x3", 2, H.TANH, N, 256, 256, tx3.inputs[2].data_ptr(),
                   tx3.inputs_lo[2].data_ptr(), 256, sx3["W"][2].data_ptr(), sx3["Wlo"][2].data_ptr(), 256,
                   sx3["b"][2].data_ptr(), 1.0, out.data_ptr(), None, O2.data_ptr(), O2l.data_ptr(), 256, s)
        elif k == "out_fields_x3":
            nw.output_fields(net, tx3, split=True)
        elif k == "products":
            nw._products(net, tape, fields)
        elif k == "channel_gram":
            nw.gram(fp.hi[:, :km], km, 1.0, "bf16x3", lo=fp.lo[:, :km])

    run()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(args.steps):
        run()
    e1.record()
    torch.cuda.synchronize()
    print(json.dumps({"kernel": args.kernel, "launches": args.steps, "ms_per_launch": e0.elapsed_time(e1) / args.steps,
                      "copy_bytes": 2 * copy_src.numel() * 4 if copy_src is not None else None}), flush=True)


N3D = 1_599_754
CHUNK3D = 131_072


def config_e(args, quick=False):
    """3D blob domain, K=1024 vector potential (the 3D K=1024 target
    workload): the two-pass chunked step (pass 1: jets + lifted fields +
    FP64 Grams per 131,072-point chunk; FP64 objective; pass 2: per-chunk
    recomputation + reverse) + Adam over all 1,599,754 points, in split
    precision (headline), plain BF16 and FP32 kernels; one chunk's field and
    Gram accuracy and the whole step's gradient against the FP64 path; and
    the 20-target FP64 full-mass projection on the split-precision basis
    against the same projection on the FP64 basis.  quick=True (inside the
    default 2D run) times 2 split-precision steps and one of each other."""
    import numpy as np
    import torch

    import neb_amd as nb
    from neb_amd import network as nw
    from neb_amd import spectral as sp
    from neb_amd import synthetic as syn

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    t0 = time.perf_counter()
    pts, _, vol = syn.blob_domain(N3D, seed=0)
    print(f"config E: sampled {N3D} blob-domain points in {time.perf_counter() - t0:.1f} s", file=sys.stderr,
          flush=True)
    x = torch.tensor(pts, dtype=torch.float32, device=dev)
    N, K = N3D, K_MODES
    w = 1.0 / N
    cfg = nb.NetConfig(dim=3, frequencies=6, lo=(-1.0,) * 3, hi=(1.0,) * 3, widths=WIDTHS, modes=K)
    net = nb.EigenfluidNet(cfg, seed=0, device=dev)
    # the 3K=3072-output net spans curls of 3*256 fields: rank(M) <= 768 < K,
    # so the objective runs with a fixed ridge of 1e-4 * mean diag(M_0)
    tape0 = nw.trunk_forward(net, x[:CHUNK3D], nw.PRECISIONS["bf16"])
    f0 = nw.output_fields(net, tape0)
    s0 = nw.output_fields(net, tape0, split=True)
    del tape0
    M0 = nw.gram(f0.mass, K, w * N / CHUNK3D, "bf16x3")
    ridge = 1e-4 * float(torch.diagonal(M0).mean())
    # Gram rate on one chunk's fields (mass R=393,216, grad-u R=1,179,648)
    gram_stats = {}
    for mode, src, gm in (("bf16x3-split", s0, "bf16x3"), ("bf16x3", f0, "bf16x3"), ("f32", f0, "f32")):
        G = nw.field_gram(src, "stiff", w, gm)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        nw.field_gram(src, "mass", w, gm, out=G)
        nw.field_gram(src, "stiff", w, gm, out=G)
        e1.record()
        torch.cuda.synchronize()
        t = e0.elapsed_time(e1) * 1e-3
        gram_stats[mode] = dict(tflops=(12 * CHUNK3D) * K * (K + 1) / t / 1e12, ms=t * 1e3)
    del f0, s0, G, M0

    def grams_for(prec):
        # split-precision tapes: the field Grams from the lift-weighted channel Grams
        return "factored" if prec == "bf16x3" else "fields"

    def step(prec, gmode, solv):
        tt = {}
        res = sp.chunked_step(net, x, w, prec, gmode, CHUNK3D, solv, True, tt, grams=grams_for(prec))
        net.adam(res["grads"], 1e-4, gate=solv.gate, on_gate=solv.raise_gate)
        return res, tt

    def timed_steps(prec, gmode, warm, reps):
        solv = sp.SpectralSolver(1e-2, ridge, deferred=True)
        for _ in range(warm):
            step(prec, gmode, solv)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        stages = {}
        for _ in range(reps):
            res, tt = step(prec, gmode, solv)
            for k, v in tt.items():
                stages[k] = stages.get(k, 0.0) + v / reps * 1e3
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps, stages, res, solv

    base = net.params.detach().clone()

    def reset():
        with torch.no_grad():
            net.params.copy_(base)
            net.m.zero_()
            net.v.zero_()
        net.epoch = 0
        net.touch()

    f32_rate = None
    if not args.no_f32:
        t32, _, _, _ = timed_steps("f32", "f32", 1, 1)
        f32_rate = N * K / t32
        reset()
    # accuracy of one chunk's fields and Grams against the FP64 path
    accuracy = {}
    tape64 = nw.trunk_forward(net, x[:CHUNK3D], nw.PRECISIONS["f64"])
    f64f = nw.output_fields(net, tape64)
    del tape64
    G64 = {wh: nw.field_gram(f64f, wh, w, "f64") for wh in ("mass", "stiff")}
    for pr in ("bf16", "bf16x3"):
        tp = nw.trunk_forward(net, x[:CHUNK3D], nw.PRECISIONS[pr])
        fp = nw.output_fields(net, tp, split=True)
        del tp
        U = nw.field_matrix(fp, "mass", torch.float64)
        acc = dict(velocity_rel_l2=float(torch.linalg.norm(U - f64f.mass[:, :K]) / torch.linalg.norm(f64f.mass[:, :K])))
        del U
        for wh, key in (("mass", "mass_gram_rel_err"), ("stiff", "grad_u_gram_rel_err")):
            G = nw.field_gram(fp, wh, w, "bf16x3")
            acc[key] = float(torch.linalg.norm(G - G64[wh]) / torch.linalg.norm(G64[wh]))
        accuracy[pr] = acc
        del fp
    del f64f, G64
    accuracy["note"] = "first 131,072-point chunk, vs FP64 jets, fields and Grams of the same points and weights"

    # the whole step's gradient against the FP64 path (same points, weights,
    # ridge and base parameters): the split-precision step, and its reverse
    # alone under the FP64 path's cotangents
    grad_err = None
    if not args.no_f32:
        r64 = sp.chunked_step(net, x, w, "f64", "f64", CHUNK3D, sp.SpectralSolver(1e-2, ridge))
        g64 = r64["grads"].double()
        L64 = r64["loss"]
        solv64 = sp.SpectralSolver(1e-2, ridge)
        _, dM64, dS64 = solv64(r64["M"], r64["S"])
        M64, S64 = r64["M"], r64["S"]
        del r64
        rs = sp.chunked_step(net, x, w, "bf16x3", "bf16x3", CHUNK3D, sp.SpectralSolver(1e-2, ridge),
                             grams=grams_for("bf16x3"))
        rr = sp.chunked_step(net, x, w, "bf16x3", "bf16x3", CHUNK3D, lambda M, S: (L64, dM64, dS64),
                             grams=grams_for("bf16x3"))
        step_grams = dict(mass_gram_rel_err=float(torch.linalg.norm(rs["M"] - M64) / torch.linalg.norm(M64)),
                          grad_u_gram_rel_err=float(torch.linalg.norm(rs["S"] - S64) / torch.linalg.norm(S64)),
                          note="the split-precision step's Grams over all points (lift-weighted channel Grams) vs "
                               "the FP64 chunked step's")
        del M64, S64

        def cmp(g):
            g = g.double()
            return dict(rel_l2=float(torch.linalg.norm(g - g64) / torch.linalg.norm(g64)),
                        cosine=float(torch.dot(g, g64) / (torch.linalg.norm(g) * torch.linalg.norm(g64))))
        grad_err = dict(whole_step=cmp(rs["grads"]), reverse_only=cmp(rr["grads"]), step_grams=step_grams,
                        note="split-precision chunked step vs the FP64 chunked step over all points; reverse_only "
                             "= split-precision tape and fields under the FP64 path's cotangents")
        del rs, rr, g64, dM64, dS64

    head = args.precision
    other_prec = "bf16" if head == "bf16x3" else "bf16x3"
    other = None
    if not args.no_f32:
        to, _, _, _ = timed_steps(other_prec, "bf16x3", 1, 1)
        other = dict(point_modes_per_s=N * K / to, ms_per_step=to * 1e3,
                     vs_fp32_step=(N * K / to) / f32_rate if f32_rate else None,
                     accuracy=f"accuracy_vs_fp64['{other_prec}']")
        reset()
    reps = 2 if quick else max(1, min(args.steps, 3))
    warm = 1 if quick else max(1, min(args.warmup, 2))
    dt, stages, res, solver = timed_steps(head, "bf16x3", warm, reps)

    # 20 divergence-free targets projected onto the split-precision basis by
    # the FP64 mass solve (two streamed passes: loads, then the direct
    # residual), and onto the FP64 basis of the same network
    t1 = time.perf_counter()
    tgt = syn.curl_targets(pts, 20, seed=1)
    targets = torch.tensor(tgt.reshape(N * 3, 20), dtype=torch.float64, device=dev)
    del tgt
    gen_s = time.perf_counter() - t1
    projection = dict(targets=20, target_generation_host_s=gen_s, basis=head)
    proj = {}
    for pr in (head, "f64"):
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        # FP64 mass of the projection's basis (the K=1024 basis is rank
        # deficient, so the solve uses the step's ridge)
        M64 = sp.fullmass_fp64(net, x, w, pr, CHUNK3D)
        torch.cuda.synchronize()
        mass_s = time.perf_counter() - t1
        t1 = time.perf_counter()
        proj[pr] = sp.project_streamed(net, x, targets, M64, w, pr, CHUNK3D, ridge=ridge)
        torch.cuda.synchronize()
        if pr == head:
            projection.update(ms=(time.perf_counter() - t1) * 1e3, fp64_mass_ms=mass_s * 1e3)
        del M64
    p, q = proj[head], proj["f64"]
    projection.update(
        normal_residual_max=float(p["normal_residuals"].max()),
        direct_relative_L2_median=float(p["direct_relative_L2"].median()),
        identity_gap_max=float((p["direct_relative_L2"] - p["Gram_identity_errors"]).abs().max()),
        vs_fp64_basis=dict(
            coefficient_rel_err=float(torch.linalg.norm(p["coefficients"] - q["coefficients"])
                                      / torch.linalg.norm(q["coefficients"])),
            loads_rel_err=float(torch.linalg.norm(p["loads"] - q["loads"]) / torch.linalg.norm(q["loads"])),
            residual_max_abs_diff=float((p["direct_relative_L2"] - q["direct_relative_L2"]).abs().max()),
            fp64_basis_residual_median=float(q["direct_relative_L2"].median()),
            note="same network, targets, weights and ridge; the FP64 basis = FP64 jets and fields"))
    del targets, proj, p, q
    return {
        "value": N * K / dt, "unit": "point*modes/s", "ms_per_step": dt * 1e3, "steps": reps, "warmup": warm,
        "config": {"workload": "3D blob domain: 1,599,754 pts, sin/cos freqs 1..6, 4x256 tanh, K=1024 vector "
                               "potential, curl lift, mass+grad-u Gram, 131,072-pt chunks, spectral loss + "
                               "reverse + Adam", "points": N, "modes": K, "chunk": CHUNK3D, "ridge": ridge,
                   "domain_volume_estimate": vol},
        "precision": head, "stages_ms": stages, "tapes_kept": bool(res["tapes_kept"]), "tape_gib": res["tape_bytes"] / 2**30, "gram_one_chunk": gram_stats,
        "fp32_step_point_modes_per_s": f32_rate,
        "vs_fp32_step": (N * K / dt) / f32_rate if f32_rate else None,
        ("bf16_step" if head == "bf16x3" else "split_precision_step"): other,
        "accuracy_vs_fp64": accuracy, "gradient_vs_fp64": grad_err,
        "loss": float(res["loss"]), "objective_path": solver.last_path, "projection": projection,
    }


def bench_3d(args):
    """--config E: the 3D K=1024 step as the headline line."""
    e = config_e(args)
    out = {"metric": "point*modes/s, jet fwd+bwd + Gram spectral step, 3D K=1024", "value": e.pop("value"),
           "unit": e.pop("unit"), "n_gpus": 1, "steps": e.pop("steps"), "warmup": e.pop("warmup"),
           "ms_per_step": e.pop("ms_per_step"), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "bf16", "data": "synthetic: 6-blob domain in [-1,1]^3 (seed 0), Xavier-normal weights seed 0",
           "config": e.pop("config")}
    out.update(e)
    print(json.dumps(out), flush=True)


def bench_3d_eig(args):
    """3D blob domain, K=512 vector potential (1536 outputs, full-rank mass):
    chunked jets + curl-lifted fields + split FP64 Grams over all 1,599,754
    points, then the FP64 generalized eigendecomposition S x = lambda M x."""
    import torch
    import neb_amd as nb
    from neb_amd import network as nw, spectral as sp, synthetic as syn

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    pts, _, vol = syn.blob_domain(N3D, seed=0)
    x = torch.tensor(pts, dtype=torch.float32, device=dev)
    N, K = N3D, 512
    w = 1.0 / N
    cfg = nb.NetConfig(dim=3, frequencies=6, lo=(-1.0,) * 3, hi=(1.0,) * 3, widths=WIDTHS, modes=K)
    net = nb.EigenfluidNet(cfg, seed=0, device=dev)

    def grams(prec, gmode):
        return sp.chunked_step(net, x, w, prec, gmode, CHUNK3D, solver=lambda M, S: (0.0, None, None),
                               training=False, grams="factored" if prec == "bf16x3" else "fields")

    head = args.precision
    other_prec = "bf16" if head == "bf16x3" else "bf16x3"
    for _ in range(max(1, min(args.warmup, 2))):
        warm = grams(head, "bf16x3")
        sp.generalized_eigh(warm["S"], warm["M"])  # first-call library set-up stays out of the timing
    del warm
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = grams(head, "bf16x3")
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    t0 = time.perf_counter()
    vals, _ = sp.generalized_eigh(res["S"], res["M"])
    torch.cuda.synchronize()
    t_eig = time.perf_counter() - t0
    # the sorted Cholesky coefficient peel (the reference's 3D objective) on
    # the same K = 512 Grams: FP64 on the GPU, device-side gate
    peel = {}
    try:
        sp.coefficient_peel(res["M"], res["S"])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        pk = sp.coefficient_peel(res["M"], res["S"], deferred=True)
        torch.cuda.synchronize()
        peel["ms"] = (time.perf_counter() - t0) * 1e3
        peel["gate"] = "passed" if int(pk["gate"].item()) == 0 else "failed"
        peel["orthonormality_max_err"] = float(pk["gram_max"])
    except FloatingPointError as e:
        pk = None
        peel["gate"] = f"failed: {e}"
    ref = grams("f64", "f64")  # FP64 jets, fields and Grams
    v64, _ = sp.generalized_eigh(ref["S"], ref["M"])
    if pk is not None and peel["gate"] == "passed":
        try:
            p64 = sp.coefficient_peel(ref["M"], ref["S"])
            peel["losses_rel_err_vs_fp64_path"] = float(torch.linalg.norm(pk["losses"] - p64["losses"])
                                                        / torch.linalg.norm(p64["losses"]))
            peel["same_order_as_fp64_path"] = bool(torch.equal(pk["order"], p64["order"]))
        except FloatingPointError as e:
            peel["fp64_path_gate"] = f"failed: {e}"
    gram_err = float(torch.linalg.norm(res["M"] - ref["M"]) / torch.linalg.norm(ref["M"]))
    eig_err = float(((vals - v64).abs() / v64.abs()).max())
    # the other jet precision on the same points: time and accuracy
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    oth = grams(other_prec, "bf16x3")
    torch.cuda.synchronize()
    t_oth = time.perf_counter() - t0
    v_oth, _ = sp.generalized_eigh(oth["S"], oth["M"])
    other = dict(forward_fields_grams_ms=t_oth * 1e3,
                 mass_gram_rel_err_vs_fp64_path=float(torch.linalg.norm(oth["M"] - ref["M"])
                                                      / torch.linalg.norm(ref["M"])),
                 eigenvalue_max_rel_err_vs_fp64_path=float(((v_oth - v64).abs() / v64.abs()).max()))
    del oth
    del ref
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    grams("f32", "f32")
    torch.cuda.synchronize()
    t_f32 = time.perf_counter() - t0
    R_m, R_s = 3 * N, 9 * N
    out = {
        "metric": "point*modes/s, jet forward + split-precision Gram assembly + FP64 generalized eigh, 3D K=512",
        "value": N * K / (t_fwd + t_eig), "unit": "point*modes/s", "n_gpus": 1, "steps": 1,
        "warmup": args.warmup, "ms_per_step": (t_fwd + t_eig) * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "bf16",
        "data": "synthetic: 6-blob domain in [-1,1]^3 (seed 0), Xavier-normal weights seed 0",
        "config": {"workload": "3D blob domain: 1,599,754 pts, sin/cos freqs 1..6, 4x256 tanh, K=512 vector "
                               "potential, curl lift, mass + grad-u split Grams in 131,072-pt chunks, FP64 "
                               "generalized eigendecomposition", "points": N, "modes": K, "chunk": CHUNK3D,
                   "domain_volume_estimate": vol},
        "stages_ms": {"forward_fields_grams": t_fwd * 1e3, "generalized_eigh": t_eig * 1e3},
        "gram_useful_tflops": (R_m + R_s) * K * (K + 1) / t_fwd / 1e12,
        "fp32_forward_grams_ms": t_f32 * 1e3,
        "precision": head,
        ("bf16_jets" if head == "bf16x3" else "split_precision_jets"): other,
        "mass_gram_rel_err_vs_fp64_path": gram_err,
        "eigenvalue_max_rel_err_vs_fp64_path": eig_err,
        "eigenvalues_min_max": [float(vals[0]), float(vals[-1])],
        "coefficient_peel": peel,
    }
    print(json.dumps(out), flush=True)


def gram_mode_stats(nw, fields, split, K, w):
    """TFLOP/s and error vs FP64 of the Gram modes on the same fields:
    "bf16x3-split" is the step's path (pre-split bf16 planes), "bf16x3" splits
    the fp32 fields per tile (same bits), "f32" is the FP32 MFMA Gram."""
    import torch
    phi64 = fields.mass[:, :K].double().contiguous()
    G64 = nw.gram(phi64, K, w, "f64")
    del phi64
    gram_stats = {}
    Rm = fields.mass.shape[0]
    flops = Rm * K * (K + 1)
    runs = {"bf16x3-split": lambda out: nw.field_gram(split, "mass", w, "bf16x3", out=out),
            "bf16x3": lambda out: nw.gram(fields.mass, K, w, "bf16x3", out=out),
            "f32": lambda out: nw.gram(fields.mass, K, w, "f32", out=out)}
    results = {}
    for mode, run in runs.items():
        G = run(None)
        results[mode] = G
        err = float(torch.linalg.norm(G - G64) / torch.linalg.norm(G64))
        torch.cuda.synchronize()
        reps = 5
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            run(G)
        e1.record()
        torch.cuda.synchronize()
        t = e0.elapsed_time(e1) / reps * 1e-3
        gram_stats[mode] = dict(tflops=flops / t / 1e12, ms=t * 1e3, rel_err_vs_fp64=err)
    gram_stats["split_vs_split_on_load_rel_diff"] = float(
        torch.linalg.norm(results["bf16x3-split"] - results["bf16x3"]) / torch.linalg.norm(results["bf16x3"]))
    return gram_stats


def fd_check_fp64(nb, sp, cfg, dev, directions=64, hs=(1e-2, 1e-3, 1e-4), seed=0):
    """The K=1024 differentiation check at full size, on the GPU in FP64: the
    handwritten gradient of L = tr((M + ridge I)^-1 S) + 1e-2 ||M - I||^2 over
    65,536 uniform random points against central differences (h = 1e-3)
    along `directions` random unit parameter directions (the reference's
    directional check, tests/test_network_backends.py:74-90), and at h = 1e-2
    and 1e-4 beside it: a central difference errs by O(h^2) (truncation) plus
    (rounding noise of the loss) / h, so the error's trend with h says which
    one limits it.  The ridge is 1e-4 * mean diag(M) at the base point (M has
    rank <= 256 < K, and M + ridge I a large condition number)."""
    import numpy as np
    import torch
    rng = np.random.default_rng(seed)
    n = 65536
    x = torch.tensor(rng.uniform(0.0, 2 * math.pi, (n, 2)), dtype=torch.float64, device=dev)
    w = 4 * math.pi ** 2 / n
    net = nb.EigenfluidNet(cfg, seed=0, device=dev, master_dtype=torch.float64)
    t0 = time.perf_counter()
    M0 = sp.compute(net, x, w, prec="f64", gram_mode="f64", training=False, gate="cholesky", ridge=1.0)["M"]
    ridge = 1e-4 * float(torch.diagonal(M0).mean())
    del M0
    res = sp.compute(net, x, w, prec="f64", gram_mode="f64", ridge=ridge, gate="cholesky")
    g = res["grads"].clone()
    L = float(res["loss"])
    del res
    base = net.params.detach().clone()
    errs = {h: [] for h in hs}
    for _ in range(directions):
        d = torch.tensor(rng.normal(size=base.numel()), dtype=torch.float64, device=dev)
        d /= d.norm()
        exact = float(g @ d)
        for h in hs:
            vals = []
            for sgn in (1.0, -1.0):
                with torch.no_grad():
                    net.params.copy_(base + sgn * h * d)
                net.touch()
                vals.append(float(sp.compute(net, x, w, prec="f64", gram_mode="f64", ridge=ridge, training=False,
                                             gate="cholesky")["loss"]))
            fd = (vals[0] - vals[1]) / (2 * h)
            errs[h].append(abs(fd - exact) / max(abs(exact), 1e-3 * abs(L)))
    with torch.no_grad():
        net.params.copy_(base)
    net.touch()
    per_h = {}
    for h in hs:
        e = np.asarray(errs[h])
        per_h[f"h={h:g}"] = dict(max_rel_err=float(e.max()), median_rel_err=float(np.median(e)),
                                 within_1e_4=int((e <= 1e-4).sum()))
    return dict(points=n, directions=directions, ridge=ridge, loss=L, **per_h, seconds=time.perf_counter() - t0,
                note="|FD - <grad, d>| / max(|<grad, d>|, 1e-3 |L|), FP64 jets, fields, Grams and reverse")


def accuracy_vs_fp64(nw, sp, net, x, K, w, ridge):
    """End-to-end errors of the bf16 and fp32 paths against the FP64 path
    (FP64 jets, fields and Grams) on the same grid and weights."""
    import torch
    dev = x.device
    tape64 = nw.trunk_forward(net, x, nw.PRECISIONS["f64"])
    f64f = nw.output_fields(net, tape64)
    del tape64
    M64 = nw.gram(f64f.mass, K, w, "f64")
    S64 = nw.gram(f64f.stiff, K, w, "f64")
    eye = torch.eye(K, dtype=torch.float64, device=dev)
    lam64, _ = sp.generalized_eigh(S64, M64 + ridge * eye)
    top = WIDTHS[-1]  # rank of M: the physical eigenvalues are the top `top`; the rest are ridge-dominated
    accuracy = {}
    for name, prec, gm in (("bf16", "bf16", "bf16x3"), ("f32", "f32", "f32"), ("f32+bf16x3", "f32", "bf16x3"),
                           ("bf16x3", "bf16x3", "bf16x3")):
        tp = nw.trunk_forward(net, x, nw.PRECISIONS[prec])
        fp = nw.output_fields(net, tp, split=prec in ("bf16", "bf16x3"))
        del tp
        Mp = nw.field_gram(fp, "mass", w, gm)
        Sp = nw.field_gram(fp, "stiff", w, gm)
        U = fp.mass[:, :K].float() + (fp.mass_lo[:, :K].float() if fp.split else 0.0)
        vel = float(torch.linalg.norm(U.double() - f64f.mass[:, :K]) / torch.linalg.norm(f64f.mass[:, :K]))
        lam, _ = sp.generalized_eigh(Sp, Mp + ridge * eye)
        accuracy[name] = dict(velocity_rel_l2=vel,
                              mass_gram_rel_err=float(torch.linalg.norm(Mp - M64) / torch.linalg.norm(M64)),
                              stiffness_gram_rel_err=float(torch.linalg.norm(Sp - S64) / torch.linalg.norm(S64)),
                              eigenvalue_max_rel_err_top=float(((lam[-top:] - lam64[-top:]).abs()
                                                                / lam64[-top:].abs()).max()))
        del fp, U
    # the headline step's Grams: split jets, channel Grams, FP64 assembly
    tp = nw.trunk_forward(net, x, nw.PRECISIONS["bf16x3"])
    prod = nw.factored_products(net, tp)
    Mp, Sp = nw.factored_field_grams(net, prod.Gh, w)
    del tp, prod
    lam, _ = sp.generalized_eigh(Sp, Mp + ridge * eye)
    accuracy["bf16x3_factored_grams"] = dict(
        mass_gram_rel_err=float(torch.linalg.norm(Mp - M64) / torch.linalg.norm(M64)),
        stiffness_gram_rel_err=float(torch.linalg.norm(Sp - S64) / torch.linalg.norm(S64)),
        eigenvalue_max_rel_err_top=float(((lam[-top:] - lam64[-top:]).abs() / lam64[-top:].abs()).max()))
    accuracy["note"] = (f"vs the FP64 path on the same grid and weights; f32+bf16x3 = FP32 jets and fields, "
                        f"split-precision BF16 Grams (the fields and Grams of the fp32_fields_bf16_reverse step); "
                        f"bf16x3 = split-precision jets and fields (bf16 hi + lo planes, three bf16 MFMA products "
                        f"per term) and split Grams; bf16x3_factored_grams = the headline step's Grams "
                        f"(split jets, split channel Grams of the stacked last-hidden jets, FP64 assembly); "
                        f"eigenvalues of S x = lambda (M + ridge I) x, "
                        f"the top {top} (= rank M; the others are set by the ridge); bf16 velocities are the hi + lo "
                        f"planes of the fp32 field values")
    return accuracy


def galerkin_bench(nw, net, x, w):
    """The basis's consumers at the config-B size: the rollout right-hand
    side (neb_galerkin_rhs2d) on the network's FP64 velocity / vorticity
    fields (65,536 points x K = 1024; 5 x n x K x 8 B read per call: an
    HBM-bound pair of passes) and the packed convective tensor of the
    metaball operators (12 N K^3 FP64 flops) on N = 131,072 synthetic points
    at K = 64."""
    import torch

    from neb_amd import galerkin as gk
    dev = x.device
    tape = nw.trunk_forward(net, x, nw.PRECISIONS["f64"])
    f = nw.output_fields(net, tape)
    del tape
    K = f.K
    U = nw.field_matrix(f, "mass", torch.float64)
    n = U.shape[0] // 2
    gu = U.as_strided((n, K), (2 * U.stride(0), 1)).contiguous()
    gv = U.as_strided((n, K), (2 * U.stride(0), 1), U.storage_offset() + U.stride(0)).contiguous()
    gw = nw.field_matrix(f, "stiff", torch.float64).contiguous()
    del f, U
    a = torch.randn(K, dtype=torch.float64, device=dev, generator=torch.Generator(device=dev).manual_seed(0))
    r = gk.rollout_rhs(gu, gv, gw, a, w)
    skew = float(torch.abs(torch.dot(a, r)) / (torch.linalg.norm(a) * torch.linalg.norm(r)))
    torch.cuda.synchronize()
    reps = 20
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        gk.rollout_rhs(gu, gv, gw, a, w, out=r)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    nbytes = 5 * n * K * 8
    rhs = dict(points=n, modes=K, ms_per_call=ms, bytes_per_call=nbytes, achieved_GBps=nbytes / (ms * 1e-3) / 1e9,
               frac_of_hbm=nbytes / (ms * 1e-3) / HBM_PEAK, relative_energy_skew=skew)
    del gu, gv, gw
    N, Kq = 131072, 64
    g = torch.Generator(device=dev).manual_seed(1)
    B = torch.randn((N, 3, Kq), dtype=torch.float64, device=dev, generator=g)
    D = torch.randn((N, 9, Kq), dtype=torch.float64, device=dev, generator=g)
    gk.assemble_operators(B[:4096], D[:4096], 1.0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    gk.assemble_operators(B, D, 1.0 / 128 ** 3)
    torch.cuda.synchronize()
    tq = time.perf_counter() - t0
    flops = 12 * N * Kq ** 3
    quad = dict(points=N, modes=Kq, seconds=tq, tflops=flops / tq / 1e12, frac_of_fp64_peak=flops / tq / PEAK["f64"],
                note="6 K row-weighted FP64 cross products of N x K x K each, plus M and A")
    return dict(rollout_rhs=rhs, convective_tensor=quad)


def modes_bench(nw, with_f32=True):
    """The reference's own architecture at its published K = 1024 shape: K
    per-mode sine networks (hidden 128 x 5, omegas 40 / 36), 3D vector
    potential with second-order jets, 6,144 blob-domain points, one update =
    fields -> FP64 Grams -> sorted coefficient peel -> recompute + reverse ->
    per-mode clipped Adam, mode blocks of 16 (the 3D Velocity step whose stages the reference
    times at 2.698 s forward + 8.952 s backward + 0.266 s Gram / peel on
    unstated NVIDIA hardware, data/paper/initial_capacity_training.json:
    88-92).  Features here are the sin/cos encoding (36 inputs; the
    reference's 15 geometry inputs are not part of this build)."""
    import torch

    from neb_amd import modes as md
    from neb_amd import synthetic as syn
    dev = torch.device("cuda:0")
    pts, _, _ = syn.blob_domain(6144, seed=0)
    x = torch.tensor(pts, dtype=torch.float32, device=dev)
    n, K = 6144, 1024
    w = 1.0 / n
    cfg = md.ModeConfig(dim=3, frequencies=6, lo=(-1.0,) * 3, hi=(1.0,) * 3, hidden=128, hidden_layers=5,
                        modes=K, omega_first=40.0, omega_hidden=36.0, block=16)
    opts = dict(deferred=True)
    out = {}
    # FP64 reference Grams of the same network (fields and Grams only)
    ref = md.ModeNet(cfg, seed=0, device=dev, master_dtype=torch.float64)
    G64 = nw.field_grams(ref.fields(x, "f64"), w, "f64")[0]
    del ref
    torch.cuda.empty_cache()
    for prec, reps in (("bf16x3", 2), ("f32", 1))[:2 if with_f32 else 1]:
        net = md.ModeNet(cfg, seed=0, device=dev, master_dtype=torch.float32)
        res = net.step(x, w, prec=prec, objective="peel", peel_options=opts)     # warm-up: staging, workspaces
        err = float(torch.linalg.norm(res["M"] - G64) / torch.linalg.norm(G64))
        net.adam(res["grads"], 1e-5)                                             # and the optimizer's first launch
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            res = net.step(x, w, prec=prec, objective="peel", peel_options=opts)
            net.adam(res["grads"], 1e-5)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / reps
        out[prec] = dict(ms_per_update=dt * 1e3, updates_timed=reps, point_modes_per_s=n * K / dt,
                         mass_gram_rel_err_vs_fp64=err, tape_blocks_kept=int(res["tapes_kept"]),
                         blocks=K // cfg.block,
                         peel_gate="passed" if int(res["peel"]["gate"].item()) == 0 else "failed")
        del net, res
        torch.cuda.empty_cache()
    if with_f32:
        out["bf16x3"]["vs_f32"] = out["f32"]["ms_per_update"] / out["bf16x3"]["ms_per_update"]
    out["reference_published_s_per_update"] = 2.698 + 8.952 + 0.266
    out["note"] = ("K=1024 per-mode SIRENs 5x128, 3D Jet2, 6,144 points, mode blocks of 16 as grouped GEMMs (each "
                   "tile contracting its own mode's inputs); the published reference figure is another GPU "
                   "(unstated NVIDIA, FP32 CuPy): context only")
    return out


def modes_stream_bench(nw, with_f32=True):
    """The reference's 2D periodic potential configuration
    (experiments/periodic/configs/periodic-potential-k1024-seed0.json): K =
    1024 per-mode sine networks (hidden 32 x 4, sin/cos frequencies 1..6 = 24
    inputs, omegas 8 / 30), 4,096 uniform random points per step on
    [0, 2 pi)^2, Jet1 streamfunction fields with the mean-zero gauge, the
    Dirichlet streamfunction Rayleigh peel (geometry order, two-pass MGS as
    the sorted Cholesky peel, mass floor 1e-14), recompute + reverse and the
    per-mode clipped Adam (clip 5.0), mode blocks of 128 (the reference
    batches 32 modes; 128-mode blocks cut the launches 4x: 47 -> 43 ms per
    step without the optimizer).  FP32 kernels throughout beside it;
    quotient error of each against the FP64 path of the same network."""
    import numpy as np
    import torch

    from neb_amd import modes as md
    dev = torch.device("cuda:0")
    n, K = 4096, 1024
    rng = np.random.default_rng(0)
    x = torch.tensor(rng.uniform(0.0, 2 * math.pi, (n, 2)), dtype=torch.float32, device=dev)
    w = 4 * math.pi ** 2 / n
    cfg = md.ModeConfig(dim=2, frequencies=6, hidden=32, hidden_layers=4, modes=K, omega_first=8.0,
                        omega_hidden=30.0, block=128, stiffness="stream")
    opts = dict(ortho_tol=math.inf, rel_min=0.0, mass_floor=1e-14, deferred=True)
    out = {}
    ref = md.ModeNet(cfg, seed=0, device=dev, master_dtype=torch.float64)
    r64 = ref.step(x, w, prec="f64", objective="peel", peel_options=opts, training=False)
    ray64 = r64["peel"]["losses"]
    del ref, r64
    torch.cuda.empty_cache()
    for prec, reps in (("bf16x3", 5), ("f32", 3))[:2 if with_f32 else 1]:
        net = md.ModeNet(cfg, seed=0, device=dev, master_dtype=torch.float32)
        for it in range(2):     # warm-up: staging, workspaces, kept tapes, the optimizer's first launches
            res = net.step(x, w, prec=prec, objective="peel", peel_options=opts)
            if it == 0:
                err = float(torch.linalg.norm(res["peel"]["losses"] - ray64) / torch.linalg.norm(ray64))
                gate = int(res["peel"]["gate"].item())
            net.adam(res["grads"], 2e-4, clip=5.0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        per = []
        for _ in range(reps):
            t1 = time.perf_counter()
            res = net.step(x, w, prec=prec, objective="peel", peel_options=opts)
            net.adam(res["grads"], 2e-4, clip=5.0)     # reads the gate flag: one synchronisation
            per.append((time.perf_counter() - t1) * 1e3)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / reps
        out[prec] = dict(ms_per_update=dt * 1e3, updates_timed=reps, ms_each=per, point_modes_per_s=n * K / dt,
                         rayleigh_rel_err_vs_fp64=err, mean_rayleigh=float(res["peel"]["losses"].mean()),
                         tape_blocks_kept=int(res["tapes_kept"]), peel_gate="passed" if gate == 0 else "failed")
        del net, res
        torch.cuda.empty_cache()
    if with_f32:
        out["bf16x3"]["vs_f32"] = out["f32"]["ms_per_update"] / out["bf16x3"]["ms_per_update"]
    out["note"] = ("periodic-potential-k1024 shape: K=1024 per-mode SIRENs 4x32, 4,096 points, Jet1 streamfunction "
                   "fields on the first-order (value + gradient) jet kernels, Dirichlet streamfunction peel + per-mode Adam, "
                   "mode blocks of 128; the reference publishes no time for this configuration")
    return out


def gradient_errors(nw, sp, net, x, w, ridge):
    """Gradient of every step variant against the FP64 path's on the same
    points, weights and ridge: relative L2 error and cosine of the full
    parameter gradient, for the whole step (the variant's own fields, Grams,
    objective and reverse) and for its reverse alone (the variant's tape and
    fields under the FP64 path's cotangents dL/dM, dL/dS).  The whole-step
    error includes the objective's amplification of Gram errors (its
    cotangent M_r^-1 S M_r^-1 moves by up to cond(M_r) ~ 1e6 times the Grams'
    relative error); the reverse-only error is the kernels' own rounding."""
    import torch
    res64 = sp.compute(net, x, w, prec="f64", gram_mode="f64", ridge=ridge, gate="cholesky")
    g64 = res64["grads"].double()
    dM, dS = res64["dM"], res64["dS"]

    def cmp(g):
        g = g.double()
        return dict(rel_l2=float(torch.linalg.norm(g - g64) / torch.linalg.norm(g64)),
                    cosine=float(torch.dot(g, g64) / (torch.linalg.norm(g) * torch.linalg.norm(g64))))

    out = {}
    for name, prec, gm, rp, grams in (("bf16x3", "bf16x3", "bf16x3", None, "factored"),
                                      ("bf16x3_field_grams", "bf16x3", "bf16x3", None, "fields"),
                                      ("bf16", "bf16", "bf16x3", None, "fields"),
                                      ("f32+bf16x3", "f32", "bf16x3", "bf16", "fields"),
                                      ("f32", "f32", "f32", None, "fields")):
        res = sp.compute(net, x, w, prec=prec, gram_mode=gm, ridge=ridge, gate="cholesky", reverse_prec=rp,
                         grams=grams)
        whole = cmp(res["grads"])
        del res
        tape = nw.trunk_forward(net, x, nw.PRECISIONS[prec])
        if grams == "factored":
            fields, prod = None, nw.factored_products(net, tape)
        else:
            fields, prod = nw.output_fields(net, tape, split=nw.split_fields_for(prec, gm)), None
        if rp is not None:
            tape = nw.convert_tape(tape, nw.PRECISIONS[rp])
        g = nw.backward_fields(net, tape, fields, dM, dS, w, symmetric=True, products=prod)
        out[name] = dict(whole_step=whole, reverse_only=cmp(g))
        del tape, fields, prod, g
    out["note"] = ("vs the FP64 path's gradient (FP64 jets, fields, Grams, objective and reverse) on the same grid, "
                   "weights and ridge; whole_step = the variant's own step, reverse_only = the variant's tape and "
                   "fields (channel Grams for bf16x3, the headline step) under the FP64 path's cotangents; "
                   "bf16x3_field_grams = split jets with field matrices, split field Grams and the fused reverse")
    del res64
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (single-GPU build: the "
              "point-sharded multi-GPU design is documented, not launched)", file=sys.stderr)
        sys.exit(2)
    if args.config == "E":
        bench_3d(args)
        return
    if args.config == "D":
        bench_3d_eig(args)
        return
    if args.config in ("modes2d", "modes3d"):
        from neb_amd import network as nw
        leg = (modes_stream_bench if args.config == "modes2d" else modes_bench)(nw, with_f32=not args.no_f32)
        print(json.dumps({"config": args.config, **leg}), flush=True)
        return
    if args.kernel:
        bench_kernel(args)
        return
    import numpy as np
    import torch

    import neb_amd as nb
    from neb_amd import network as nw
    from neb_amd import spectral as sp

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    axis = np.arange(N_GRID) * 2 * np.pi / N_GRID
    gx, gy = np.meshgrid(axis, axis, indexing="xy")
    x = torch.tensor(np.column_stack((gx.ravel(), gy.ravel())), dtype=torch.float32, device=dev)
    N = x.shape[0]
    w = (2 * math.pi / N_GRID) ** 2
    cfg = nb.NetConfig(dim=2, frequencies=FREQS, widths=WIDTHS, modes=K_MODES)
    net = nb.EigenfluidNet(cfg, seed=0, device=dev)
    K = K_MODES

    # ridge: the K=1024 / width-256 mass is rank <= 256 by construction; a fixed
    # ridge of 1e-4 * mean diag(M_0) makes the FP64 Cholesky well posed.
    tape = nw.trunk_forward(net, x, nw.PRECISIONS["bf16"])
    fields = nw.output_fields(net, tape)
    split = nw.output_fields(net, tape, split=True)
    M0 = nw.gram(fields.mass, K, w, "bf16x3")
    ridge = 1e-4 * float(torch.diagonal(M0).mean())

    # ---- Gram modes on the same fields: TFLOP/s and error vs FP64 ----------
    # "bf16x3-split": the step's path (pre-split bf16 planes); "bf16x3" splits
    # the fp32 fields per tile; both give the same bits.
    gram_stats, accuracy, fd_check, grad_errors, galerkin, per_mode, per_mode_2d = {}, {}, None, None, None, None, None
    if not args.step_only:
        gram_stats = gram_mode_stats(nw, fields, split, K, w)
        accuracy = accuracy_vs_fp64(nw, sp, net, x, K, w, ridge)
        grad_errors = gradient_errors(nw, sp, net, x, w, ridge)
        galerkin = galerkin_bench(nw, net, x, w)
        per_mode = modes_bench(nw)
        per_mode_2d = modes_stream_bench(nw)
        fd_check = fd_check_fp64(nb, sp, cfg, dev)
    del tape, fields, split

    # ---- FP32 comparison step ------------------------------------------------
    f32_rate = None
    if not args.no_f32:
        step32 = make_step(nb, nw, sp, net, x, w, "f32", "f32", ridge)
        for _ in range(1):
            step32()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        n32 = max(1, min(3, args.steps))
        for _ in range(n32):
            step32()
        torch.cuda.synchronize()
        f32_rate = N * K / ((time.perf_counter() - t0) / n32)

    # ---- FP32 jets / fields, split BF16 Grams, BF16 reverse -----------------
    # (Grams at FP32 accuracy; the reverse on BF16 MFMA from the rounded tape)
    mixed = None
    if not args.no_f32:
        stepm = make_step(nb, nw, sp, net, x, w, "f32", "bf16x3", ridge, reverse_prec="bf16")
        stepm()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        nm = max(1, min(3, args.steps))
        for _ in range(nm):
            stepm()
        torch.cuda.synchronize()
        tm = (time.perf_counter() - t0) / nm
        mixed = dict(point_modes_per_s=N * K / tm, ms_per_step=tm * 1e3,
                     vs_fp32_step=(N * K / tm) / f32_rate if f32_rate else None,
                     accuracy="accuracy_vs_fp64['f32+bf16x3']")

    # ---- the other BF16 step (split-precision or plain jets) ------------------
    head = args.precision
    other_prec = "bf16" if head == "bf16x3" else "bf16x3"
    other = None
    if not args.no_f32:
        stepo = make_step(nb, nw, sp, net, x, w, other_prec, "bf16x3", ridge,
                          grams="factored" if other_prec == "bf16x3" else "fields")
        stepo()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        no = max(1, min(5, args.steps))
        for _ in range(no):
            stepo()
        torch.cuda.synchronize()
        to = (time.perf_counter() - t0) / no
        other = dict(point_modes_per_s=N * K / to, ms_per_step=to * 1e3,
                     vs_fp32_step=(N * K / to) / f32_rate if f32_rate else None,
                     accuracy=f"accuracy_vs_fp64['{other_prec}']")

    # ---- headline step -----------------------------------------------------
    solver = sp.SpectralSolver(1e-2, ridge, deferred=True)
    head_grams = "factored" if head == "bf16x3" else "fields"
    step = make_step(nb, nw, sp, net, x, w, head, "bf16x3", ridge, solver=solver, grams=head_grams)
    for _ in range(args.warmup):
        step()
    timer = StageTimer(torch)
    step_t = make_step(nb, nw, sp, net, x, w, head, "bf16x3", ridge, timer, solver=solver, grams=head_grams)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        loss = step_t()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    stages = {k: float(np.mean(v)) for k, v in timer.ms().items()}
    # per-kernel launch times: the same steps again, with HIP events on the
    # launching stream around every libnebhip call (recording them costs
    # ~10 % of the step, so these steps stay out of `value`)
    launches = LaunchTimer(torch)
    torch.cuda.synchronize()
    nw.H.set_launch_recorder(launches)
    t1 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    dt_rec = (time.perf_counter() - t1) / args.steps
    nw.H.set_launch_recorder(None)
    kernels = launches.table(args.steps, dt_rec * 1e3)

    # The step's split-precision Gram launches (cross kernel + split
    # reduction): algorithmic R*K*(K+1) flops (the symmetric half, multiply +
    # add; the three split products are not credited) over their event-timed
    # average duration -- the channel Grams (R = N points, K = 2 x 256 jet
    # channels) on the factored path, the mass Gram (R = 2N rows, K = 1024)
    # on the field path.
    if "gram_mass_launch" in stages:
        gram_ms = stages["gram_mass_launch"]
        gram_flops = (2 * N) * K * (K + 1)
        gram_kernel, gram_profile = ("neb_gram_split mass Gram: cross_async_kernel<2, 2, 2, 2, false> + "
                                     "cross_reduce_kernel, R=131072, K=1024"), "gram_split_k1024_pmc.md"
    else:
        grow = max((r for r in kernels if r["kernel"].startswith("neb_gram_split")), key=lambda r: r["ms_per_step"])
        gram_ms, gram_flops = grow["ms_per_launch"], grow["flops_per_launch"]
        gram_kernel = (f"{grow['kernel']} channel Gram (K = 512 stacked jet channels): cross_async_kernel + "
                       f"cross_reduce_kernel")
        gram_profile = "gram_split_k512_r65536_pmc.md"
    achieved = gram_flops / (gram_ms * 1e-3)
    gram_roofline = dict(bound="mfma", achieved=achieved / 1e12, peak=PEAK["bf16"] / 1e12, unit="TFLOP/s",
                         frac=achieved / PEAK["bf16"], traffic=profile_traffic(gram_profile),
                         kernel=gram_kernel, launch_ms=gram_ms, flops_per_launch=gram_flops)
    # roofline of the dominant kernel: the kernel class with the largest share
    # of the step's time (event-timed launches of the timed steps above)
    dom = next(r for r in kernels if "bound" in r)
    profile = PROFILE_OF.get(dom["kernel"])
    roofline = dict(bound=dom["bound"], achieved=dom["achieved"], peak=dom["peak"], unit=dom["unit"],
                    frac=dom["frac"], traffic=profile_traffic(profile) if profile else None,
                    kernel=dom["kernel"], share_of_step=dom["share_of_step"], launch_ms=dom["ms_per_launch"],
                    bytes_per_launch=dom["bytes_per_launch"], flops_per_launch=dom["flops_per_launch"],
                    traffic_profile=profile)

    cpu = None if args.no_cpu else cpu_baseline()
    out = {
        "metric": "point*modes/s, jet fwd+bwd + Gram spectral step, 2D K=1024",
        "value": N * K / dt,
        "precision": (f"{head}: " + ("split-precision jets (bf16 hi + lo planes, three bf16 MFMA products per "
                                     "term, fp32 accumulation); the field Grams from split channel Grams of the "
                                     "stacked last-hidden jets, assembled in FP64 (no field matrices), FP64 "
                                     "objective, factored reverse (FP64 output-layer algebra, split GEMM on the "
                                     "tape); Grams within 1e-5 of the FP64 path "
                                     "(accuracy_vs_fp64['bf16x3_factored_grams'])"
                                     if head == "bf16x3" else
                                     "BF16 jets and fields, split bf16x3 Grams, FP64 objective, BF16 reverse "
                                     "(accuracy_vs_fp64['bf16'])")),
        "unit": "point*modes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16",
        "data": "synthetic: 256^2 periodic grid, Xavier-normal weights seed 0",
        "config": {"workload": "2D periodic operator grid: 65,536 pts, sin/cos freqs 1..8, 4x256 tanh, K=1024, "
                               "mass+stiffness Gram, spectral loss + reverse + Adam",
                   "points": N, "modes": K, "widths": list(WIDTHS), "ridge": ridge, "jets": head},
        "stages_ms": stages,
        "gram": gram_stats,
        "fp32_step_point_modes_per_s": f32_rate,
        "fp32_fields_bf16_reverse_step": mixed,
        ("bf16_step" if head == "bf16x3" else "split_precision_step"): other,
        "accuracy_vs_fp64": accuracy,
        "fd_check_fp64": fd_check,
        "loss": float(loss),
        "objective_path": solver.last_path,
        "roofline": roofline,
        "gram_roofline": gram_roofline,
        "kernels": kernels,
        "kernels_note": (f"per kernel class (entry point, inner width or rows): launches, event-timed ms per "
                         f"step and per launch over {args.steps} recorded steps ({dt_rec * 1e3:.3f} ms each), "
                         "share of that step (side-stream kernels overlap the current stream, so shares can sum "
                         "above 1), algorithmic bytes / useful flops per launch and the roofline they imply"),
        "gradient_vs_fp64": grad_errors,
        "galerkin": galerkin,
        "reference_architecture_3d": per_mode,
        "reference_architecture_2d_stream": per_mode_2d,
        "cpu_baseline": cpu,
    }
    if not args.no_3d:
        del net, solver, step, step_t, x
        torch.cuda.empty_cache()
        out["config_E"] = config_e(args, quick=True)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
