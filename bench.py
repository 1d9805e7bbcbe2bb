"""Throughput of FP64 QG ensembles on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--workload W]

Default workload (the headline): 256 independent two-layer QGModels,
nx = ny = 256, dt = 3600 s, FP64, seeded 1e-7 N(0,1) initial PV, stepped with
the fused two-kernel step.  One "step" advances every member by dt.  Prints one
JSON line with the metric (model-days per second summed over members), the HBM
roofline of the fused step measured with HIP events over the timed region, and
the CPU baseline (the NumPy oracle on the host cores, same physics, same run).

Other workloads (--workload):
  uneven     256 two-layer members, nx drawn from {64, 128, 256}, beta, rd,
             delta, rek drawn per member (default_rng(1)), dt = 3600 s
  layered10  64 LayeredModel members, nz = 10, nx = 256, L = 2e6 m,
             exponential stratification, first deformation radius 20 km
  cnn        the learned closure's CNN: forward and training-step throughput
  closed     the learned closure's coarse step: 128 members at nx=64 with the CNN
             forcing inside the fused step (step graphs)
  single512  one two-layer QGModel at nx=512, dt=1800 s (the single-GPU parity
             configuration): run() with fused stretches, a _step_forward loop and
             the oracle on one host core; no graphs, so it can run under a tracer
  dropin     the reference's example runs (five model-years of single QGModels,
             nx = 64 and 256, with and without closures) through the drop-in API,
             beside the wall times the reference's notebook records
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "model-days/sec (ensemble, 1 GPU) at nx=256 2-layer; KE-spectrum err vs hi-res"
MEMBERS, NX, NZ, DT = 256, 256, 2, 3600.0
PHYS = dict(L=1e6, H1=500.0, delta=0.25, U1=0.025, U2=0.0, beta=1.5e-11, rd=15000.0,
            rek=5.787e-7, filterfac=23.6)
HBM_PEAK_GBS = 8000.0


def measured_traffic(radix_row, radix_col, path=None):
    """HBM bytes per fused step from the committed rocprofv3 PMC summary of
    the headline workload (FETCH_SIZE and WRITE_SIZE, copy-calibrated), or
    None when the summary is missing or lacks the tuned kernels."""
    path = path or os.path.join(ROOT, "profiles", "fused_step_qg2_nx256_256members.md")
    try:
        text = open(path).read()
    except OSError:
        return None
    if "## HBM traffic" not in text:
        return None
    text = text.split("## HBM traffic", 1)[1]
    total = 0.0
    for name in (f"k_row_flux<256, {radix_row}, true>",
                 f"k_col_step_reg<256, 2, {radix_col}, true>"):
        for line in text.splitlines():
            cells = [c.strip() for c in line.split("|")]
            if len(cells) > 3 and cells[1] == name:
                try:
                    total += (float(cells[2]) + float(cells[3])) * 2 ** 20
                    break
                except ValueError:
                    return None
        else:
            return None
    return total


def algorithmic_bytes(nz, nx, ny):
    """Per member-step: read qh, dqhdt_p, dqhdt_pp; write qh and the new dqhdt
    (complex128 spectral fields of nz x ny x (nx/2+1))."""
    return 5 * 16 * nz * ny * (nx // 2 + 1)


# ------------------------------------------------------------------ CPU leg
def _cpu_worker(args):
    nsteps, seed, workload = args
    import numpy as np
    from oracle.qg_oracle import OracleLayered, OracleQG
    rng = np.random.default_rng(seed)
    if workload == "layered10":
        from pyqg_amd.constants import exponential_stratification
        H, rho, U = exponential_stratification()
        m = OracleLayered(nz=10, nx=256, H=H, rho=rho, U=U, V=np.zeros(10), L=2e6,
                          beta=1.5e-11, rek=5.787e-7, dt=DT, f=0.0001236812857687059, g=9.81)
        nz, n = 10, 256
    else:
        n = NX if workload == "qg256" else int(rng.choice([64, 128, 256]))
        m = OracleQG(nx=n, dt=DT, **PHYS)
        nz = NZ
    m.q = 1e-7 * rng.standard_normal((nz, n, n))
    m._step_forward()
    t0 = time.perf_counter()
    for _ in range(nsteps):
        m._step_forward()
    return time.perf_counter() - t0


def cpu_baseline(workload="qg256", nsteps=None):
    """The oracle (NumPy FP64, numpy.fft = the reference's no-pyfftw path), one
    member per process, one thread per process, on the host cores."""
    import concurrent.futures as cf
    import multiprocessing as mp
    os.environ["OMP_NUM_THREADS"] = "1"
    os.environ["OPENBLAS_NUM_THREADS"] = "1"
    if nsteps is None:
        nsteps = {"qg256": 500, "uneven": 500, "layered10": 60}[workload]
    # one member per core of the host's share per GPU (16 cores on the
    # MI355X hosts this runs on; the host may show more: affinity and
    # cpu_count are reported beside the result)
    affinity = len(os.sched_getaffinity(0))
    cores = min(affinity, 16)
    with cf.ProcessPoolExecutor(max_workers=cores, mp_context=mp.get_context("spawn")) as ex:
        t0 = time.perf_counter()
        times = list(ex.map(_cpu_worker, [(nsteps, s, workload) for s in range(cores)]))
        wall = time.perf_counter() - t0
    per_member = nsteps * DT / 86400. / (sum(times) / len(times))
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    what = {"qg256": f"the nx={NX} two-layer QG oracle",
            "uneven": "two-layer QG oracle members of nx drawn from {64,128,256}",
            "layered10": "the nz=10 nx=256 LayeredModel oracle"}[workload]
    return {"value": per_member * cores, "unit": "model-days/s", "cores": cores,
            "kind": "port", "affinity_cores": affinity, "cpu_count": os.cpu_count(),
            "cpu_model": model,
            "sample": (f"{cores} members x {nsteps} steps of {what}, one single-threaded "
                       f"process per core ({model}); aggregate = mean per-member rate x "
                       f"cores; {wall:.1f} s wall")}


# ------------------------------------------------------------------ CNN closure
CNN_PEAK_TFLOPS = 2500.0  # MI355X dense BF16 matrix peak


def cnn_flops(channels, h=64, w=64):
    """Multiply-adds x 2 of one sample's forward pass over the real channels."""
    return 2.0 * h * w * 25 * sum(channels[l] * channels[l + 1] for l in range(len(channels) - 1))


def cnn_cpu_baseline(channels, batch=4, reps=3):
    """The same network in FP32 on the host with torch's CPU convolutions
    (circular padding): a port, timed on a small batch."""
    import torch
    torch.manual_seed(0)
    layers = []
    for l in range(len(channels) - 1):
        layers.append(torch.nn.Conv2d(channels[l], channels[l + 1], 5, padding=2,
                                      padding_mode="circular"))
        if l < len(channels) - 2:
            layers.append(torch.nn.ReLU())
    net = torch.nn.Sequential(*layers).eval()
    x = torch.randn(batch, channels[0], 64, 64)
    with torch.no_grad():
        net(x)
        t0 = time.perf_counter()
        for _ in range(reps):
            net(x)
        wall = time.perf_counter() - t0
    return {"value": batch * reps / wall, "unit": "samples/s", "cores": torch.get_num_threads(),
            "kind": "port",
            "sample": f"{reps} forward passes of {batch} samples, torch CPU FP32 conv2d"}


def cnn_library_reference(channels, batch, reps):
    """The same layers through torch.nn.functional.conv2d (MIOpen), BF16,
    channels_last, on inputs padded by two pixels beforehand (no bias, no
    ReLU): the vendor library's time for each convolution on this GPU."""
    import torch
    import torch.nn.functional as F
    torch.backends.cudnn.benchmark = True
    ms = []
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    for l in range(len(channels) - 1):
        x = torch.randn(batch, channels[l], 68, 68, device="cuda", dtype=torch.bfloat16)
        w = torch.randn(channels[l + 1], channels[l], 5, 5, device="cuda", dtype=torch.bfloat16)
        x = x.contiguous(memory_format=torch.channels_last)
        w = w.contiguous(memory_format=torch.channels_last)
        for _ in range(5):
            F.conv2d(x, w)
        torch.cuda.synchronize()
        e0.record()
        for _ in range(reps):
            F.conv2d(x, w)
        e1.record()
        torch.cuda.synchronize()
        ms.append(e0.elapsed_time(e1) / reps)
    return {"kind": "MIOpen through torch.nn.functional.conv2d, BF16 channels_last, "
                    "pre-padded input, no bias or ReLU",
            "layer_ms": ms, "sum_ms": sum(ms)}


def bench_cnn(args):
    """Inference and training throughput of the closure network on libqgcnn."""
    import numpy as np
    import torch
    from pyqg_amd.closure import cnn as C
    cpu = None if args.no_cpu else cnn_cpu_baseline(C.DEFAULT_CHANNELS)
    net = C.CNN(seed=0)
    B = 128
    xp = net.pad_input(torch.randn(B, 64, 64, net.channels[0], device="cuda"))
    st = torch.cuda.current_stream()
    # inference as the closed step runs it: every layer's output preallocated
    # (forward_into), so the launches queue back to back; the allocating
    # forward (net.forward, a new output tensor per layer) is timed beside it
    bufs = net.activation_buffers(B, 64, 64)
    for _ in range(args.warmup):
        net.forward_into(xp, bufs, None)
        net.forward(xp)
    # and at least ~50 ms of back-to-back passes first: the first timed region
    # otherwise catches the clocks still rising (measured 10 % slower)
    for _ in range(max(0, 100 - 2 * args.warmup)):
        net.forward_into(xp, bufs, None)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(st)
    for _ in range(args.steps):
        net.forward_into(xp, bufs, None)
    e1.record(st)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_fwd = e0.elapsed_time(e1) / args.steps
    e0.record(st)
    for _ in range(args.steps):
        net.forward(xp)
    e1.record(st)
    torch.cuda.synchronize()
    ms_fwd_alloc = e0.elapsed_time(e1) / args.steps
    # per-layer times of the same forward pass (events on the launch stream)
    layer_ms = []
    h = xp
    for l in range(net.nlayers):
        last = l == net.nlayers - 1
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        ev[0].record(st)
        for _ in range(args.steps):
            y, _ = net._layer(l, h, relu=not last, out_bf16=True, out_f32=False)
        ev[1].record(st)
        torch.cuda.synchronize()
        layer_ms.append(ev[0].elapsed_time(ev[1]) / args.steps)
        h = y
    fl = cnn_flops(net.channels)
    layer_fl = [B * 2.0 * 64 * 64 * 25 * net.channels[l] * net.channels[l + 1]
                for l in range(net.nlayers)]
    dom = int(np.argmax(layer_fl))
    achieved = layer_fl[dom] / (layer_ms[dom] * 1e-3) / 1e12
    ms_train, train_tflops = None, None
    if not args.cnn_inference_only:
        # training step (batch 64): forward, backward (data and weights), Adam
        Bt = 64
        xt = xp[:Bt].contiguous()
        yt = torch.randn(Bt, 64, 64, net.channels[-1], device="cuda")
        for _ in range(3):
            net.train_step(xt, yt, lr=1e-5)
        torch.cuda.synchronize()
        nt = max(args.steps // 4, 5)
        e0.record(st)
        for _ in range(nt):
            net.train_step(xt, yt, lr=1e-5)
        e1.record(st)
        torch.cuda.synchronize()
        ms_train = e0.elapsed_time(e1) / nt
        train_tflops = 3 * Bt * fl / (ms_train * 1e-3) / 1e12
    library = cnn_library_reference(net.channels, B, args.steps) if args.cnn_library else None

    def layer_kernel(l):
        # the library's dispatch: first- and last-layer kernels where packed;
        # 64- and 32-channel output groups over 32-channel input chunks on the
        # overlapped-staging kernel (64-row images); the per-block kernel else
        if l == 0 and net.first_packed:
            return "k_conv_first<64, 4, 2>"
        if l == net.nlayers - 1 and net.last_packed:
            return "k_conv_last<64, 1>"
        if net.chp[l] % 32 == 0 and net.chp[l + 1] % 64 == 0:
            return "k_conv_fwd_ws<64, 4, 2, 4, true>"
        if net.chp[l] % 32 == 0 and net.chp[l + 1] % 32 == 0:
            return "k_conv_fwd_ws<64, 2, 2, 4, false>"
        mt = 4 if net.chp[l + 1] % 64 == 0 else (2 if net.chp[l + 1] % 32 == 0 else 1)
        return f"k_conv_fwd_mt<64, {mt}, {32 if net.chp[l] % 32 == 0 else 16}, 2>"
    out = {
        "metric": "CNN subgrid-closure inference throughput (64x64 inputs, 8 conv layers)",
        "value": B * args.steps / wall,
        "unit": "samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16",
        "data": "synthetic (random inputs, seeded weights)",
        "config": {"workload": f"CNN closure forward, batch {B}, channels "
                               f"{'-'.join(map(str, net.channels))}, 5x5 circular, NHWC BF16",
                   "batch": B},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": CNN_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / CNN_PEAK_TFLOPS, "traffic": None,
                     "layer_kernels": [layer_kernel(l) for l in range(net.nlayers)],
                     "kernel": (f"{layer_kernel(dom)}: layer {dom + 1} "
                                f"({net.channels[dom]}->{net.channels[dom + 1]} channels), "
                                f"batch {B}"),
                     "algorithmic_flops_per_launch": layer_fl[dom],
                     "kernel_ms": layer_ms[dom], "layer_ms": layer_ms,
                     "forward_ms": ms_fwd,
                     "forward_tflops": B * fl / (ms_fwd * 1e-3) / 1e12,
                     "forward_allocating_ms": ms_fwd_alloc,
                     "train_step_ms_batch64": ms_train, "train_tflops": train_tflops},
        "cpu_baseline": cpu,
    }
    if library is not None:
        out["library_reference"] = library
    print(json.dumps(out))


def bench_closed(args):
    """The learned closure's coarse step: 128 two-layer members at nx=64,
    dt = 4 h, each step the fused row kernel (writing the network input), the
    8-layer network, and the fused column kernel with the network's forcing,
    replayed from step graphs.  The network is untrained (seeded weights):
    the timing does not depend on the values."""
    import numpy as np
    import torch
    from pyqg_amd.closure import cnn as C
    from pyqg_amd.closure import data as D
    from pyqg_amd.closure import param as Pm
    net = C.CNN(seed=0)
    scales = Pm.Scales(np.full(6, 1e-6), np.full(2, 1e-12))
    p = Pm.CNNParameterization(net, scales)
    M, n, dt = 128, 64, 14400.
    ce = Pm.ClosedEnsemble(M, n, dt, D.QG_EDDY, param=p, seed=0)
    plain = Pm.ClosedEnsemble(M, n, dt, D.QG_EDDY, param=None, seed=0)
    out = {}
    for name, e in (("closed", ce), ("plain", plain)):
        # at least ~50 ms of steps first, so that the timed steps do not catch
        # the clocks still rising
        e.step(max(args.warmup, 4, 100 if name == "closed" else 1000))
        e.ens.buckets[0].sync()
        t0 = time.perf_counter()
        e.step(args.steps)
        e.ens.buckets[0].sync()
        out[name] = (time.perf_counter() - t0) * 1e3 / args.steps
    ms = out["closed"]
    print(json.dumps({
        "metric": "learned-closure coarse step (128 members, nx=64, dt=4 h)",
        "value": M * args.steps * dt / 86400. / (ms * 1e-3 * args.steps),
        "unit": "model-days/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": True, "vs_baseline": None, "dtype": "f64+bf16",
        "data": "synthetic (seeded initial PV, seeded untrained network)",
        "config": {"workload": "ClosedEnsemble: 128 two-layer QG members, nx=64, dt=14400 s, "
                               "fused step with the 8-layer CNN forcing", "fused": ce.fused,
                   "plain_ms_per_step": out["plain"],
                   "us_per_member_day": ms * 1e3 / M * 86400. / dt}}))


# ------------------------------------------------------------------ GPU leg
# The reference's own example runs (docs/examples/parameterizations.ipynb:57-58, 88-89,
# 137-138, 165-166): five model-years of a QGModel at dt = 3600 s, diagnostics averaged
# daily after 2.5 years, and the published wall time of each (hardware not stated).
DROPIN_RUNS = [
    ("QGModel nx=64", 64, None, 17.6),
    ("QGModel nx=256", 256, None, 334.0),
    ("QGModel nx=64 + Smagorinsky(constant=0.08)", 64,
     ("Smagorinsky", dict(constant=0.08)), 49.3),
    ("QGModel nx=64 + BackscatterBiharmonic(smag_constant=0.08, back_constant=1.1)", 64,
     ("BackscatterBiharmonic", dict(smag_constant=0.08, back_constant=1.1)), 48.5),
]


def bench_dropin(args):
    """Single models through the drop-in API, exactly as a pyqg user runs them:
    construct, run() to tmax, read q back."""
    import numpy as np
    import pyqg_amd
    year = 24 * 60 * 60 * 360.
    base = dict(dt=3600., tmax=5 * year, tavestart=2.5 * year, twrite=25000, log_level=0)
    runs = []
    for name, nx, param, ref_wall in DROPIN_RUNS:
        kw = dict(base, nx=nx)
        if param is not None:
            kw["parameterization"] = getattr(pyqg_amd, param[0])(**param[1])
        t0 = time.perf_counter()
        m = pyqg_amd.QGModel(**kw)
        t1 = time.perf_counter()
        m.run()
        q = np.asarray(m.q)
        t2 = time.perf_counter()
        if not np.isfinite(q).all():
            raise SystemExit(f"non-finite state after {name}")
        days = base["tmax"] / 86400.
        runs.append({"run": name, "steps": int(m.tc), "model_days": days,
                     "construct_s": t1 - t0, "run_s": t2 - t1,
                     "model_days_per_s": days / (t2 - t1),
                     "reference_wall_s": ref_wall,
                     "reference_model_days_per_s": days / ref_wall,
                     "speedup_vs_reference": ref_wall / (t2 - t1)})
        print(f"{name}: {runs[-1]['run_s']:.2f} s ({runs[-1]['model_days_per_s']:.0f} "
              f"model-days/s), reference {ref_wall} s", file=sys.stderr, flush=True)
    config2 = bench_config2(args)
    heavy = runs[1]
    print(json.dumps({
        "metric": "model-days/s of the reference's example runs through the drop-in QGModel",
        "value": heavy["model_days_per_s"], "unit": "model-days/s", "n_gpus": 1,
        "higher_is_better": True, "vs_baseline": heavy["speedup_vs_reference"],
        "dtype": "f64", "data": "the reference's default initial condition",
        "config": {"workload": "docs/examples/parameterizations.ipynb runs: dt=3600 s, "
                               "tmax=5 years, tavestart=2.5 years, daily diagnostics",
                   "runs": runs, "config2": config2}}))


def bench_config2(args, nsteps=2000, cpu_steps=20):
    """One two-layer member at nx=512, dt=1800 (seeded 1e-7 N(0,1) PV, diagnostics
    off): run() (fused stretches), a plain _step_forward loop (phase by phase), and
    the oracle on one host core for a bounded sample of the same run."""
    import numpy as np
    import pyqg_amd
    phys = dict(PHYS)
    q0 = 1e-7 * np.random.default_rng(0).standard_normal((2, 512, 512))
    out = {"nx": 512, "dt": 1800.0, "steps": nsteps}
    m = pyqg_amd.QGModel(nx=512, dt=1800.0, tmax=nsteps * 1800.0, log_level=0, **phys)
    m.q = q0
    t0 = time.perf_counter()
    m.run()
    np.asarray(m.q)
    out["run_model_days_per_s"] = nsteps * 1800.0 / 86400. / (time.perf_counter() - t0)
    m = pyqg_amd.QGModel(nx=512, dt=1800.0, log_level=0, **phys)
    m.q = q0
    n2 = nsteps // 4
    t0 = time.perf_counter()
    for _ in range(n2):
        m._step_forward()
    np.asarray(m.q)
    out["step_forward_model_days_per_s"] = n2 * 1800.0 / 86400. / (time.perf_counter() - t0)
    if not args.no_cpu:
        from oracle.qg_oracle import OracleQG
        o = OracleQG(nx=512, dt=1800.0, **phys)
        o.q = q0
        o._step_forward()
        t0 = time.perf_counter()
        for _ in range(cpu_steps):
            o._step_forward()
        w = time.perf_counter() - t0
        out["cpu_oracle"] = {"model_days_per_s": cpu_steps * 1800.0 / 86400. / w,
                             "cores": 1, "kind": "port",
                             "sample": f"{cpu_steps} steps of the same run, one process"}
    print(f"config 2 (nx=512): {out}", file=sys.stderr, flush=True)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--members", type=int, default=MEMBERS)
    ap.add_argument("--workload", choices=["qg256", "uneven", "layered10", "cnn", "dropin",
                                           "single512", "closed"],
                    default="qg256")
    ap.add_argument("--cnn-library", action="store_true",
                    help="cnn workload: also time each layer with MIOpen (torch conv2d)")
    ap.add_argument("--cnn-inference-only", action="store_true",
                    help="cnn workload: skip the training-step timing (profiling runs)")
    ap.add_argument("--graph-steps", type=int, default=100,
                    help="steps timed again on the graph-replay path (0: skip)")
    ap.add_argument("--unpacked", action="store_true",
                    help="exchange nx/2 + 1 x-mixed columns per row (packed exchange off)")
    ap.add_argument("--copy-probe", type=int, default=1 << 30,
                    help="bytes moved by the copy kernel that calibrates HBM counters (0: skip)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the ensemble runs on a single MI355X here "
              "(members shard across GPUs with no per-step communication).", file=sys.stderr)
        sys.exit(2)

    if args.workload == "cnn":
        return bench_cnn(args)
    if args.workload == "closed":
        return bench_closed(args)
    if args.workload == "dropin":
        return bench_dropin(args)
    if args.workload == "single512":
        out = bench_config2(args)
        print(json.dumps({
            "metric": "model-days/s of one two-layer QGModel at nx=512 (run(), fused stretches)",
            "value": out["run_model_days_per_s"], "unit": "model-days/s", "n_gpus": 1,
            "higher_is_better": True, "vs_baseline": None, "dtype": "f64",
            "data": "synthetic (seeded 1e-7 N(0,1) initial PV)",
            "config": {"workload": "one QGModel, nx=512, dt=1800 s, diagnostics off", **out}}))
        return None
    cpu = None if args.no_cpu else cpu_baseline(args.workload)

    import numpy as np
    import pyqg_amd

    if args.workload == "qg256":
        M = args.members
        ens = pyqg_amd.Ensemble.qg(np.full(M, NX), dt=DT, **{k: v for k, v in PHYS.items()})
        nxs, nz = np.full(M, NX), NZ
        workload = (f"QGModel two-layer ensemble: {M} members, nx=ny={NX}, dt={DT:g} s, "
                    f"L=1e6 m, rd=15 km, delta=0.25, U1=0.025 m/s, beta=1.5e-11, "
                    f"rek=5.787e-7, FP64, fused 2-kernel step")
    elif args.workload == "uneven":
        M = 256
        rng = np.random.default_rng(1)
        nxs = rng.choice([64, 128, 256], size=M)
        beta = rng.uniform(0.5e-11, 2.5e-11, M)
        rd = rng.uniform(1e4, 2.5e4, M)
        delta = rng.uniform(0.1, 0.5, M)
        rek = rng.uniform(3e-7, 1e-6, M)
        ens = pyqg_amd.Ensemble.qg(nxs, beta=beta, rd=rd, delta=delta, rek=rek, dt=DT)
        nz = NZ
        workload = (f"uneven two-layer QG ensemble: 256 members, nx from {{64,128,256}} "
                    f"({int((nxs == 64).sum())}/{int((nxs == 128).sum())}/"
                    f"{int((nxs == 256).sum())}), per-member beta, rd, delta, rek, dt=3600 s, FP64")
    else:
        M = 64
        H, rho, U = pyqg_amd.constants.exponential_stratification()
        ens = pyqg_amd.Ensemble.layered(M, 256, 10, H, rho, U, L=2e6, dt=DT)
        nxs, nz = np.full(M, 256), 10
        workload = ("LayeredModel ensemble: 64 members, nz=10, nx=ny=256, L=2e6 m, exponential "
                    "stratification (H ~ exp(-z/800 m), 20 km first radius), dt=3600 s, FP64")
    if args.unpacked:
        for b in ens.buckets:
            b.set_packed_exchange(False)
    ens.set_q([1e-7 * np.random.default_rng(s).standard_normal((nz, int(n), int(n)))
               for s, n in enumerate(nxs)])
    ens.step(3)                           # AB start-up, untimed
    tune = ens.autotune(reps=5, cache_path=False)
    ens.step(max(args.warmup - 3, 1))     # warm-up with the tuned kernels, untimed
    for b in ens.buckets:
        b.sync()
        b.timing_enable(True)
    t0 = time.perf_counter()
    ens.step(args.steps)                  # returns after a device synchronize
    wall = time.perf_counter() - t0
    step_ms = 0.0
    per_bucket = []
    for b in ens.buckets:
        ms_row, n_row, ms_col, n_col = b.timing()
        b.timing_enable(False)
        per_bucket.append({"nx": b.nx, "members": b.M, "row_kernel_ms": ms_row / n_row,
                           "col_kernel_ms": ms_col / n_col, "variant": list(b.variant())})
        step_ms += ms_row / n_row + ms_col / n_col

    copy = None
    if args.copy_probe:
        ms = pyqg_amd._lib.copy_probe(args.copy_probe)
        copy = {"bytes": args.copy_probe, "ms": ms,
                "GB/s": 2 * args.copy_probe / (ms * 1e-3) / 1e9}
    pattern = None
    if args.workload == "qg256" and nz == 2:
        # memory-only stand-ins of the two kernels at this shape: the time the
        # fused kernels would take if their transforms were free
        b0 = ens.buckets[0]
        ldk = ((b0.nx // 2 + 1) + 7) // 8 * 8
        ncols = b0.exchange_columns()
        ncols = ncols if ncols % 8 == 0 else ldk  # unpacked rows are written whole
        ms_r, ms_c = pyqg_amd._lib.pattern_probe(b0.ny, ldk, b0.M, nz, ncols=ncols)
        pattern = {"row_ms": ms_r, "col_ms": ms_c, "columns_per_row": ncols,
                   "row_kernel_over_bound": per_bucket[0]["row_kernel_ms"] / ms_r,
                   "col_kernel_over_bound": per_bucket[0]["col_kernel_ms"] / ms_c}

    # the same steps replayed from a hipGraph of two fused steps per bucket
    graph_ms = None
    if args.graph_steps:
        ens.step(2, use_graph=True)       # capture, untimed
        for b in ens.buckets:
            b.sync()
        t0 = time.perf_counter()
        ens.step(args.graph_steps, use_graph=True)
        graph_ms = (time.perf_counter() - t0) * 1e3 / args.graph_steps

    # every member's health (the reference's CFL < 1 assertion, per member)
    cfl = ens.member_cfl()
    if not (np.isfinite(cfl).all() and cfl.max() < 1.0):
        raise SystemExit(f"members failed the health check after the timed steps: "
                         f"{np.nonzero(~(np.isfinite(cfl) & (cfl < 1.0)))[0].tolist()}")

    model_days = M * args.steps * DT / 86400.
    value = model_days / wall
    bytes_step = float(sum(algorithmic_bytes(nz, int(n), int(n)) for n in nxs))
    # With several buckets each runs on its own stream and their kernels
    # overlap, so the sum of the per-kernel event times overstates the step;
    # the step's span is then the timed region's wall clock per step.
    span_ms = step_ms if len(ens.buckets) == 1 else wall * 1e3 / args.steps
    achieved = bytes_step / (span_ms * 1e-3) / 1e9
    traffic, traffic_src = None, None
    if args.workload == "qg256" and M == MEMBERS:
        traffic = measured_traffic(*ens.buckets[0].variant())
        if traffic is not None:
            traffic_src = ("profiles/fused_step_qg2_nx256_256members.md: rocprofv3 --pmc "
                           "FETCH_SIZE and WRITE_SIZE of the same two kernels at this "
                           "workload, copy-calibrated (not measured in this run)")
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "model-days/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded 1e-7 N(0,1) initial PV per member)",
        "config": {"workload": workload, "members": M, "nz": nz,
                   "nx": int(nxs[0]) if args.workload != "uneven" else "64/128/256",
                   "dt": DT},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": traffic_src,
                     "kernel": ("fused step = k_row_flux + k_col_step" if len(ens.buckets) == 1
                                else "fused steps of every bucket, concurrent on their streams: "
                                     "achieved over the timed region's wall clock per step"),
                     "algorithmic_bytes_per_launch": bytes_step,
                     "kernel_ms_per_step": step_ms, "buckets": per_bucket,
                     "copy_kernel": copy, "access_pattern_bound": pattern, "autotune": tune},
        "cpu_baseline": cpu,
        "graph_path": None if graph_ms is None else {
            "steps": args.graph_steps, "ms_per_step": graph_ms,
            "value": M * DT / 86400. / (graph_ms * 1e-3),
            "note": "the same fused steps replayed from one hipGraph of two steps of "
                    "every bucket (wall clock, after the event-timed eager steps)"},
        "health": {"cfl_max": float(cfl.max()), "members_checked": int(cfl.size)},
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
