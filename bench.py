"""Throughput of the FP64 two-layer QG ensemble on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

Workload: 256 independent two-layer QGModels, nx = ny = 256, dt = 3600 s, FP64,
seeded 1e-7 N(0,1) initial PV, stepped with the fused two-kernel step.  One
"step" advances every member by dt.  Prints one JSON line with the metric
(model-days per second summed over members), the HBM roofline of the fused
step measured with HIP events over the timed region, and the CPU baseline
(the NumPy oracle on the host cores, same physics, in the same run).
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


def algorithmic_bytes(nz, nx, ny):
    """Per member-step: read qh, dqhdt_p, dqhdt_pp; write qh and the new dqhdt
    (complex128 spectral fields of nz x ny x (nx/2+1))."""
    return 5 * 16 * nz * ny * (nx // 2 + 1)


# ------------------------------------------------------------------ CPU leg
def _cpu_worker(args):
    nsteps, seed = args
    import numpy as np
    from oracle.qg_oracle import OracleQG
    m = OracleQG(nx=NX, dt=DT, **PHYS)
    m.q = 1e-7 * np.random.default_rng(seed).standard_normal((NZ, NX, NX))
    m._step_forward()
    t0 = time.perf_counter()
    for _ in range(nsteps):
        m._step_forward()
    return time.perf_counter() - t0


def cpu_baseline(nsteps=500):
    """The oracle (NumPy FP64, numpy.fft = the reference's no-pyfftw path), one
    member per process, one thread per process, on the host cores."""
    import concurrent.futures as cf
    import multiprocessing as mp
    os.environ["OMP_NUM_THREADS"] = "1"
    os.environ["OPENBLAS_NUM_THREADS"] = "1"
    cores = min(len(os.sched_getaffinity(0)), 16)
    with cf.ProcessPoolExecutor(max_workers=cores, mp_context=mp.get_context("spawn")) as ex:
        t0 = time.perf_counter()
        times = list(ex.map(_cpu_worker, [(nsteps, s) for s in range(cores)]))
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
    return {"value": per_member * cores, "unit": "model-days/s", "cores": cores,
            "kind": "port",
            "sample": (f"{cores} members x {nsteps} steps of the nx={NX} two-layer QG oracle, "
                       f"one single-threaded process per core ({model}); aggregate = "
                       f"per-member rate x cores; {wall:.1f} s wall")}


# ------------------------------------------------------------------ GPU leg
def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--members", type=int, default=MEMBERS)
    ap.add_argument("--copy-probe", type=int, default=1 << 30,
                    help="bytes moved by the copy kernel that calibrates HBM counters (0: skip)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the ensemble runs on a single MI355X here "
              "(members shard across GPUs with no per-step communication).", file=sys.stderr)
        sys.exit(2)

    cpu = None if args.no_cpu else cpu_baseline()

    import numpy as np
    import pyqg_amd

    M = args.members
    ens = pyqg_amd.Ensemble.qg(np.full(M, NX), dt=DT, **{k: v for k, v in PHYS.items()})
    ens.set_q([1e-7 * np.random.default_rng(s).standard_normal((NZ, NX, NX)) for s in range(M)])
    ens.step(3)                           # AB start-up, untimed
    tune = ens.autotune(reps=5, cache_path=False)
    ens.step(max(args.warmup - 3, 1))     # warm-up with the tuned kernels, untimed
    b = ens.buckets[0]
    b.sync()

    b.timing_enable(True)
    t0 = time.perf_counter()
    ens.step(args.steps)                  # returns after a device synchronize
    wall = time.perf_counter() - t0
    ms_row, n_row, ms_col, n_col = b.timing()
    b.timing_enable(False)

    copy = None
    if args.copy_probe:
        ms = pyqg_amd._lib.copy_probe(args.copy_probe)
        copy = {"bytes": args.copy_probe, "ms": ms,
                "GB/s": 2 * args.copy_probe / (ms * 1e-3) / 1e9}

    qh = ens.get("qh", 0)
    if not np.isfinite(qh).all():
        raise SystemExit("non-finite state after the timed steps")

    model_days = M * args.steps * DT / 86400.
    value = model_days / wall
    row_ms, col_ms = ms_row / n_row, ms_col / n_col
    step_ms = row_ms + col_ms
    bytes_step = algorithmic_bytes(NZ, NX, NX) * M
    achieved = bytes_step / (step_ms * 1e-3) / 1e9
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
        "config": {"workload": f"QGModel two-layer ensemble: {M} members, nx=ny={NX}, "
                               f"dt={DT:g} s, L=1e6 m, rd=15 km, delta=0.25, U1=0.025 m/s, "
                               f"beta=1.5e-11, rek=5.787e-7, FP64, fused 2-kernel step",
                   "members": M, "nx": NX, "nz": NZ, "dt": DT},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "fused step = k_row_flux + k_col_step",
                     "algorithmic_bytes_per_launch": bytes_step,
                     "row_kernel_ms": row_ms, "col_kernel_ms": col_ms,
                     "copy_kernel": copy,
                     "variant": {"row_radix": b.variant()[0], "col_radix": b.variant()[1],
                                 "autotune": list(tune.values())[0]}},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
