"""End-to-end learned-closure experiment (two-layer eddy regime).

    python -m pyqg_amd.closure.experiment [--members 256] [--samples 51200] ...

1. A high-resolution ensemble (nx=256, dt=3600 s) is spun up and sampled; at
   each sample libqgk computes the coarse inputs (q, u, v at nx=64) and the
   subgrid PV forcing, and the time-mean KE spectrum of the high-resolution
   flow is accumulated.
2. The CNN (5x5 convolutions, 128/64/32 channels) is trained on the GPU with
   libqgcnn's BF16 forward/backward kernels; offline skill is measured on
   held-out pairs.
3. Coarse ensembles (nx=64) are run without closure, with the network, and
   with the built-in closures (backscatter, ZB2020, Smagorinsky), started from
   the coarse-grained high-resolution states; their time-mean KE spectra
   (device-accumulated KEspec, calc_ispec per layer) are compared with the
   high-resolution one over the coarse grid's resolved wavenumbers, and the
   cost per simulated day is compared.
Prints one JSON line.
"""
import argparse
import json
import time

import numpy as np

from .. import _lib
import torch  # noqa: E402
from . import cnn as cnn_mod
from . import data as data_mod
from . import param as param_mod
from . import spectra
from . import train as train_mod
from .. import diagnostic_tools as tools


# The two-layer model's time-mean diagnostics compared between the runs (the
# registry entries pyqg's diagnostic_similarities compares, with the
# parameterization terms KEflux and APEflux are summed with).
COMPARED = ["KEspec", "Ensspec", "EKE", "EKEdiss", "KEfrictionspec", "Dissspec", "ENSDissspec",
            "entspec", "ENSflux", "ENSgenspec", "ENSfrictionspec", "APEgenspec", "APEflux",
            "KEflux", "APEgen", "paramspec_APEflux", "paramspec_KEflux"]
# energy-budget terms among them
BUDGET = ["KEflux", "APEflux", "APEgenspec", "KEfrictionspec", "Dissspec", "EKEdiss", "APEgen"]


def run(members=256, samples=51200, spinup_days=1825, every=24, epochs=8, batch=64, lr=1e-3,
        online_members=128, online_days=720, avg_days=360, dt_lo=14400., seed=0, ema=0.998,
        ema_from=0.5, closures=("backscatter", "zb2020", "smagorinsky"), calibrate="none",
        filter_kind="gaussian", stochastic=False, spec_weight=0.0, nets=1, log=print):
    out = {"config": dict(members=members, samples=samples, spinup_days=spinup_days,
                          every=every, epochs=epochs, batch=batch, lr=lr,
                          online_members=online_members, online_days=online_days,
                          avg_days=avg_days, dt_hi=3600., dt_lo=dt_lo, seed=seed, ema=ema,
                          ema_from=ema_from, calibrate=calibrate, filter=filter_kind,
                          stochastic=stochastic, spectral_loss=spec_weight, nets=nets)}
    L = data_mod.QG_EDDY["L"]
    g_hi, g_lo = tools.grid(256, L), tools.grid(64, L)
    # ---------------------------------------------------------- 1. data
    gen = data_mod.PairGenerator(nmembers=members, n_hi=256, n_lo=64, dt=3600., seed=seed,
                                 filter_kind=filter_kind)
    t0 = time.perf_counter()
    gen.spin_up(spinup_days * 24)
    hb = gen.hi.buckets[0]
    hb.sync()
    log(f"spin-up: {spinup_days} days of {members} members in {time.perf_counter() - t0:.1f} s")
    n_hold = max(samples // 10, members)
    xs, ys = [], []
    n = 0
    nsnap = 0
    t_hi_steps = 0.0
    while n < samples + n_hold:
        ts = time.perf_counter()
        gen.hi.step(every)
        hb.sync()
        t_hi_steps += time.perf_counter() - ts
        gen.hi.accumulate_diagnostics(COMPARED)
        x, y = gen.sample()
        xs.append(x)
        ys.append(y)
        n += x.shape[0]
        nsnap += 1
        if nsnap % 50 == 0:
            log(f"sampled {n} pairs")
    e_hi = gen.hi.get_diagnostic("KEspec", mean=True)[0]
    view_hi = gen.hi.mean_model(0)
    x = torch.cat(xs)
    y = torch.cat(ys)
    del xs, ys
    x_tr, y_tr = x[:samples], y[:samples]
    x_te, y_te = x[samples:], y[samples:]
    out["data_seconds"] = time.perf_counter() - t0
    out["hi_res_ms_per_step"] = 1e3 * t_hi_steps / (nsnap * every)
    log(f"data: {x_tr.shape[0]} train + {x_te.shape[0]} held-out pairs in "
        f"{out['data_seconds']:.1f} s")

    # ---------------------------------------------------------- 2. training
    scales = param_mod.Scales.fit(x_tr, y_tr)
    net = cnn_mod.CNN(seed=seed)
    out["n_params"] = net.n_params()
    info = train_mod.train(net, x_tr, y_tr, scales, epochs=epochs, batch=batch, lr=lr,
                           log=log, seed=seed, ema=ema, ema_from=ema_from,
                           spec_weight=spec_weight)
    out["train"] = {"steps": info["steps"], "seconds": info["seconds"],
                    "final_loss": info["history"][-1][1] if info["history"] else None}
    out["offline"] = train_mod.evaluate(net, x_te, y_te, scales)
    amp = train_mod.amplitude(net, x_te, y_te, scales)
    out["offline"]["amplitude"] = amp
    scales_fit = scales
    if calibrate == "variance":
        # restore the forcing's held-out mean square, layer by layer: the MSE fit
        # leaves out the part of the forcing its inputs do not determine, and
        # with it energy the coarse model then lacks
        gain = np.sqrt(np.asarray(amp["target_ms"]) / np.asarray(amp["pred_ms"]))
        scales = param_mod.Scales(scales.in_scale, scales.out_scale * gain)
        out["offline"]["gain"] = gain.tolist()
        log(f"variance calibration gain {gain}")
    noise = None
    if stochastic:
        # the forcing the network leaves out, as an AR(1) process per spectral
        # point with the held-out residual's spectrum and its correlation over
        # one sampling interval (one day) scaled to the coarse step
        amp, rho = train_mod.residual_statistics(net, x_te, y_te, scales, lag=members)
        r1 = float(max(np.mean(rho), 0.0))
        phi = r1 ** (dt_lo / (every * 3600.)) if r1 > 0 else 0.0
        noise = (amp, phi)
        out["offline"]["residual_lag_correlation"] = rho.tolist()
        out["offline"]["noise_phi"] = phi
        log(f"residual: one-interval correlation {rho}, phi per coarse step {phi:.3f}")
    extra = []
    for k in range(1, nets):
        # more networks on the same pairs, other initial weights and batch orders
        nk = cnn_mod.CNN(seed=seed + 1000 * k)
        train_mod.train(nk, x_tr, y_tr, scales_fit, epochs=epochs, batch=batch, lr=lr,
                        log=lambda s: None, seed=seed + 1000 * k, ema=ema, ema_from=ema_from,
                        spec_weight=spec_weight)
        out.setdefault("offline_extra", []).append(train_mod.evaluate(nk, x_te, y_te,
                                                                      scales_fit))
        extra.append(nk)
        log(f"network {k + 1} of {nets}: offline R^2 = {out['offline_extra'][-1]['r2']:.3f}")
    scales_avg = scales_fit
    if extra and calibrate == "variance":
        # the mean of the networks, calibrated as one network would be
        amp_avg = train_mod.amplitude([net] + extra, x_te, y_te, scales_fit)
        gain_avg = np.sqrt(np.asarray(amp_avg["target_ms"]) / np.asarray(amp_avg["pred_ms"]))
        scales_avg = param_mod.Scales(scales_fit.in_scale, scales_fit.out_scale * gain_avg)
        out["offline"]["gain_mean_of_networks"] = gain_avg.tolist()
        log(f"variance calibration gain of the mean of {nets} networks {gain_avg}")
    flop_per_sample = 2 * 64 * 64 * 25 * sum(
        net.chp[l] * net.chp[l + 1] for l in range(net.nlayers))
    out["train"]["tflops"] = 3 * flop_per_sample * info["steps"] * batch / info["seconds"] / 1e12
    log(f"offline R^2 = {out['offline']['r2']:.3f}")
    del x, y, x_tr, y_tr

    # ---------------------------------------------------------- 3. online
    p = param_mod.CNNParameterization(net, scales)
    # coarse-grained high-resolution states as the common start
    _lib.check(_lib.lib.qgk_coarsen(hb.h, gen.lo.buckets[0].h, _lib.FIELD_ID["qh"],
                                    _lib.FIELD_ID["qh"], _lib.ptr(gen.G)))
    qlo = gen.lo.get("q")
    runs = {}
    spectra_lo = {}
    from .. import parameterizations as builtin
    choices = {"backscatter": builtin.BackscatterBiharmonic(), "zb2020": builtin.ZannaBolton2020(),
               "smagorinsky": builtin.Smagorinsky()}
    views = {}
    plan = [("plain", None), ("cnn", p)]
    if extra:
        plan.append(("cnn_avg", param_mod.CNNAverageParameterization([net] + extra,
                                                                     scales_avg)))
    if noise is not None:
        plan.append(("cnn_stochastic", p))
    for name, prm in plan + [(c, choices[c]) for c in closures]:
        ce = param_mod.ClosedEnsemble(online_members, 64, dt_lo, gen.phys, param=prm, seed=seed)
        ce.ens.set_q([qlo[i % members] for i in range(online_members)])
        if name == "cnn_stochastic":
            ce.set_noise(noise[0], noise[1], seed=seed + 1)
        per_day = int(round(86400. / dt_lo))
        nsteps = online_days * per_day
        navg = avg_days * per_day
        cb = ce.ens.buckets[0]
        t1 = time.perf_counter()
        done = 0
        # accumulate KEspec at the start of one step a day over the last avg_days
        for s0 in range(nsteps - navg, nsteps, per_day):
            if s0 > done:
                ce.step(s0 - done)
                done = s0
            ce.ens.accumulate_diagnostics(COMPARED)
        if nsteps > done:
            ce.step(nsteps - done)
        cb.sync()
        wall = time.perf_counter() - t1
        e = ce.ens.get_diagnostic("KEspec", mean=True)[0]
        spectra_lo[name] = e
        views[name] = ce.ens.mean_model(0)
        finite = bool(np.isfinite(ce.ens.get("q", 0)).all())
        runs[name] = {"ke_rmse": spectra.ke_errors(e, g_lo, e_hi, g_hi),
                      "ke_rel_err": spectra.relative_errors(e, g_lo, e_hi, g_hi),
                      "ms_per_step": 1e3 * wall / nsteps, "finite": finite,
                      "total_ke": float(tools.spec_sum(e).sum())}
        log(f"{name}: KE-spectrum rel. error {runs[name]['ke_rel_err']}, "
            f"{runs[name]['ms_per_step']:.2f} ms/step")
    for name in runs:
        if name != "plain":
            runs[name]["similarity_vs_plain"] = spectra.similarity(runs[name]["ke_rmse"],
                                                                  runs["plain"]["ke_rmse"])
            # pyqg's comparison over every compared diagnostic, per layer
            sims = tools.diagnostic_similarities(views[name], view_hi, views["plain"])
            runs[name]["diagnostic_similarities"] = {k: float(v) for k, v in sims.items()}
            budget = [v for k, v in sims.items() if k.rstrip("12") in BUDGET]
            runs[name]["budget_similarity_mean"] = float(np.mean(budget))
    out["online"] = runs
    out["hi_res_total_ke"] = float(tools.spec_sum(e_hi).sum())
    # cost per member and simulated day
    hi_cost = out["hi_res_ms_per_step"] / members * 86400. / 3600.
    lo_cost = runs["cnn"]["ms_per_step"] / online_members * 86400. / dt_lo
    out["cost_ratio_hi_over_cnn"] = hi_cost / lo_cost
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--members", type=int, default=256)
    ap.add_argument("--samples", type=int, default=51200)
    # From 1e-7 noise the baroclinic instability still grows through year 3;
    # the flow reaches its statistically steady state in year 4 or 5.
    ap.add_argument("--spinup-days", type=int, default=1825)
    ap.add_argument("--every", type=int, default=24)
    ap.add_argument("--epochs", type=int, default=8)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--online-members", type=int, default=128)
    ap.add_argument("--online-days", type=int, default=720)
    ap.add_argument("--avg-days", type=int, default=360)
    ap.add_argument("--dt-lo", type=float, default=14400.)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--ema", type=float, default=0.998,
                    help="weight averaging decay over the second half of training (0: off)")
    ap.add_argument("--ema-from", type=float, default=0.5,
                    help="fraction of training after which the weights are averaged")
    ap.add_argument("--closures", default="backscatter,zb2020,smagorinsky",
                    help="built-in closures run online beside the network (comma list)")
    ap.add_argument("--calibrate", choices=["none", "variance"], default="none",
                    help="variance: scale each output layer so the network's held-out "
                         "mean-square forcing equals the target's")
    ap.add_argument("--filter", choices=["gaussian", "sharp"], default="gaussian",
                    help="coarse-graining after the spectral truncation")
    ap.add_argument("--stochastic", action="store_true",
                    help="also run the network with an AR(1) stochastic forcing that has the "
                         "held-out residual's spectrum and time correlation")
    ap.add_argument("--spectral-loss", type=float, default=0.0,
                    help="weight of the log-spectral distance between predicted and target "
                         "forcing added to the MSE in training (0: MSE only)")
    ap.add_argument("--nets", type=int, default=1,
                    help="also train this many networks in all (other seeds) and run the "
                         "mean of their forcings online as cnn_avg")
    ap.add_argument("--out", default=None, help="also append the JSON line to this file")
    a = ap.parse_args()
    res = run(members=a.members, samples=a.samples, spinup_days=a.spinup_days, every=a.every,
              epochs=a.epochs, batch=a.batch, lr=a.lr, online_members=a.online_members,
              online_days=a.online_days, avg_days=a.avg_days, dt_lo=a.dt_lo, seed=a.seed,
              ema=a.ema, ema_from=a.ema_from, closures=[c for c in a.closures.split(",") if c],
              calibrate=a.calibrate, filter_kind=a.filter, stochastic=a.stochastic,
              spec_weight=a.spectral_loss, nets=a.nets,
              log=lambda s: print(s, flush=True))
    line = json.dumps(res)
    print(line)
    if a.out:
        with open(a.out, "a") as f:
            f.write(line + "\n")


if __name__ == "__main__":
    main()
