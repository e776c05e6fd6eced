#!/usr/bin/env python3
"""Throughput of SAM's RK3 pseudo-spectral step on one MI355X.

Workload (default): stratified Boussinesq DNS at 512^3 -- seeded
random-phase solenoidal velocity, E(k) ~ k^4 exp(-2(k/4)^2),
u_rms = 1, T' = 0, N = 4 (grav*lapse/To = 16), nu = kappa = 5e-4, Pr = 1,
RK3 with CFL 0.5, 2/3 dealiasing.  One "step" is one full RK3 step (three
substeps) of the whole 512^3 grid with the state resident in HBM.

  python bench.py [--steps K] [--warmup W] [--n 512] [--no-cpu] [--no-split]

Prints ONE JSON line (metric/value/unit ... roofline, cpu_baseline).
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12          # MI355X HBM3E, bytes/s
METRIC = "RK3 Mpoint-steps/s at 512^3 + %HBM peak; learned vs dm.f SGS step time"


def grid_counts(n, i_strat=1):
    kmax = int((n - 0.1) / 3.0)
    nxm, nym, nzm = kmax + 1, 2 * kmax + 1, 2 * kmax + 1
    S = nxm * nym * nzm
    Z = nxm * nym * n
    Y = nxm * n * n
    lu, lr = (4, 9) if i_strat else (3, 6)
    return dict(S=S, Z=Z, Y=Y, P=n ** 3, lu=lu, lr=lr)


def kernel_bytes(n, bc, i_strat=1, zrhs_fused=False):
    """Algorithmic HBM bytes per launch of each step kernel (fused 3-pass
    model): state/RHS in FP64 (16 B), transform elements bc = 16 or 8 B.
    zrhs_fused: k_zrhs also writes the next substep's z inverse (Lu Z
    fields), which k_zb then does not compute."""
    g = grid_counts(n, i_strat)
    S, Z, Y, lu, lr = g["S"], g["Z"], g["Y"], g["lu"], g["lr"]
    return {
        "zb": lu * S * 16 + lu * Z * bc,
        "yb": lu * Z * bc + lu * Y * bc,
        "xrow": lu * Y * bc + lr * Y * bc,
        "yf": lr * Y * bc + lr * Z * bc,
        "zrhs": lr * Z * bc + 4 * lu * S * 16 + (lu * Z * bc if zrhs_fused else 0),
    }


def zrhs_fused(prof):
    """the timed region's k_zrhs wrote the next z inverse: fewer k_zb than
    k_zrhs launches (k_zb runs only where the previous k_zrhs did not)"""
    return prof.get("zb", (0.0, 0))[1] < prof.get("zrhs", (0.0, 0))[1]


def bytes_moved(n, bc, prof, steps):
    """bytes per step the launched kernels move by the same model: the
    contract's count (step_bytes) less the state reads a fused k_zrhs saves"""
    kb = kernel_bytes(n, bc, zrhs_fused=zrhs_fused(prof))
    return sum(kb[k] * prof.get(k, (0.0, 0))[1] for k in kb) / steps


def step_bytes(n, bc, i_strat=1, nrk=3):
    g = grid_counts(n, i_strat)
    per_sub = 5 * g["lu"] * g["S"] * 16 + 2 * (g["lu"] + g["lr"]) * (g["Z"] + g["Y"]) * bc
    return nrk * per_sub


def make_params(S, n, precision, sgs=0):
    return S.make_params(n, i_strat=1, grav=16.0, to=1.0, lapse=1.0, vis=5e-4, pr=1.0,
                         i_cfl=1, cfl0=0.5, dt0=0.01, precision=precision, sgs=sgs)


def run_gpu(n, precision, steps, warmup, u0, S, keep_state=True):
    sam = S.Sam(make_params(S, n, precision))
    sam.set_state(u0)
    sam.step(warmup)
    sam.sync()
    sam.prof_enable(True)
    sam.sync()
    t0 = time.perf_counter()
    diag = sam.step(steps, want_spectra=True)
    sam.sync()
    t1 = time.perf_counter()
    prof = sam.prof_read()
    sam.prof_enable(False)
    spec = sam.spectra()
    state = sam.get_state() if keep_state else None
    # the median of single steps after the timed block (each synchronised;
    # the headline stays the bracketed K-step mean)
    per = []
    for _ in range(max(5, min(steps, 10))):
        ts = time.perf_counter()
        sam.step(1, want_spectra=True)
        sam.sync()
        per.append(time.perf_counter() - ts)
    sam.close()
    return dict(wall=t1 - t0, diag=diag, prof=prof, spec=spec, state=state, median=float(np.median(per)))


def dominant_roofline(n, bc, prof, steps):
    kb = kernel_bytes(n, bc, zrhs_fused=zrhs_fused(prof))
    best = None
    for name, nbytes in kb.items():
        ms, cnt = prof.get(name, (0.0, 0))
        if cnt == 0:
            continue
        if best is None or ms > best[1]:
            best = (name, ms, cnt, nbytes)
    name, ms, cnt, nbytes = best
    avg_s = ms / cnt / 1e3
    achieved = nbytes / avg_s / 1e9
    per_kernel = {}
    for k, nb in kb.items():
        kms, kc = prof.get(k, (0.0, 0))
        if kc:
            per_kernel[k] = dict(avg_us=round(kms / kc * 1e3, 2), GBps=round(nb / (kms / kc / 1e3) / 1e9, 1),
                                 share=round(kms / sum(v[0] for v in prof.values()), 3))
    return dict(kernel=name, bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK / 1e9, unit="GB/s",
                frac=round(achieved / (HBM_PEAK / 1e9), 4), traffic=None, bytes_per_launch=nbytes,
                avg_launch_us=round(avg_s * 1e6, 2), launches=cnt), per_kernel


def measured_traffic(n, kernel, prec):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary (profiles/step_<n>_hbm_traffic.json), if one exists"""
    path = os.path.join(ROOT, "profiles", f"step_{n}_hbm_traffic.json")
    try:
        with open(path) as fh:
            t = json.load(fh)
        return t["kernels"][f"{kernel}_{prec}"]["hbm_bytes"]
    except (OSError, KeyError, ValueError):
        return None


def les_comparison(S, ic, n_dns, n_les=128, les_steps=2000, refresh=20):
    """Learned closure vs the dynamic model on a 128^3 stratified LES.

    Training: a 512^3 FP32-split DNS of the LES's physics (N = 8, seeded
    k0 = 6 IC, nu = 5e-4) is box-filtered by 4 at four times; the network
    learns the dynamic model's stress and flux at unit coefficient on those
    filtered fields (SAM_LABELS_DYNAMIC), 1200 Adam steps of 2^20 points.
    LES (nu = 1e-5, the DNS IC truncated to 128^3): `les_steps` steps with
    the dynamic model, with the learned closure (C Delta^2 from the dynamic
    procedure every `refresh` steps), with the dynamic model from an IC
    perturbed by 1e-6 (the trajectories' chaotic spread, the floor of any
    comparison), and with the dynamic model whose C Delta^2 is itself
    refreshed every `refresh` steps (the control for that amortisation).
    The learned closure runs twice: with ReLU hidden units (the default) and
    with GELU (tanh form), each network trained on the same samples.
    Whole-run wall time per step; E(k) compared at the end and time-averaged
    over the second half (every 50 steps)."""
    from sam_amd import les as L
    r = n_dns // n_les
    phys_dns = dict(i_strat=1, grav=64.0, to=1.0, lapse=1.0, vis=5e-4, pr=1.0, i_cfl=1, cfl0=0.5, dt0=0.01,
                    precision=S.PREC_SPLIT)
    phys = dict(phys_dns, vis=1e-5)
    u0 = ic.random_solenoidal(n_dns, k0=6.0, seed=20261203, i_strat=1)
    dns = S.Sam(S.make_params(n_dns, **phys_dns))
    dns.set_state(u0)
    mlp = S.Mlp(seed=1, pool_capacity=4 * n_les ** 3)
    mlp.set_labels(S.LABELS_DYNAMIC, 1.0)
    mlp_gelu = S.Mlp(seed=1, pool_capacity=4 * n_les ** 3)
    mlp_gelu.set_activation(S.ACT_GELU)
    mlp_gelu.set_labels(S.LABELS_DYNAMIC, 1.0)
    t0 = time.perf_counter()
    for _ in range(4):
        dns.step(50)
        mlp.add_samples(dns, r)
        mlp_gelu.add_samples(dns, r)
    dns.close()
    t_dns = time.perf_counter() - t0
    t0 = time.perf_counter()
    losses = L.train(mlp, 1200, batch=1 << 20, lr=1e-3, seed=7)
    t_train = time.perf_counter() - t0
    t0 = time.perf_counter()
    losses_gelu = L.train(mlp_gelu, 1200, batch=1 << 20, lr=1e-3, seed=7)
    t_train_gelu = time.perf_counter() - t0
    u_les = L.spectral_truncate(u0, n_les)
    rng = np.random.default_rng(5)
    res = {}
    runs = (("dynamic", S.SGS_DYN, u_les), ("learned", S.SGS_MLP, u_les),
            ("dynamic_perturbed", S.SGS_DYN, u_les * (1 + 1e-6 * rng.standard_normal(u_les.shape))),
            ("dynamic_lagged", S.SGS_DYN, u_les), ("learned_gelu", S.SGS_MLP, u_les))
    half = max(1, les_steps // 2)
    chunk = min(50, max(1, half // 2))    # steps between E(k) samples of the second half
    for name, model, u in runs:
        les = S.Sam(S.make_params(n_les, sgs=model, **phys))
        if model == S.SGS_MLP:
            S.attach_mlp(les, mlp_gelu if name == "learned_gelu" else mlp)
            S.set_mlp_refresh(les, refresh)
        if name == "dynamic_lagged":
            S.set_dyn_refresh(les, refresh)
        les.set_state(u)
        les.sync()
        t0 = time.perf_counter()
        les.step(half)
        acc, ns = 0.0, 0
        for _ in range(half // chunk):
            les.step(chunk, want_spectra=True)
            acc = acc + L.shell_energy(les.spectra())
            ns += 1
        les.sync()
        wall = time.perf_counter() - t0
        check = None
        if model == S.SGS_MLP:
            err, sw = S.sgs_check(les, 1.0)
            check = {"sgs_transfer_rel_err_vs_dynamic": round(float(err), 4), "fell_back": bool(sw)}
        res[name] = dict(ms=wall / (half + chunk * (half // chunk)) * 1e3, final=L.shell_energy(les.spectra()),
                         avg=acc / ns, check=check)
        les.close()

    def err(a, b, lo=1):
        k1 = len(b) - 1
        return round(float(np.linalg.norm(a[lo:k1] - b[lo:k1]) / np.linalg.norm(b[lo:k1])), 4)

    d, m, p, g = res["dynamic"], res["learned"], res["dynamic_perturbed"], res["dynamic_lagged"]
    mg = res["learned_gelu"]
    return {
        "workload": f"stratified LES {n_les}^3, N=8, nu=1e-5, IC: the seeded k0=6 DNS IC truncated, "
                    f"FP32 transforms / FP64 state, {les_steps} steps",
        "dynamic_ms_per_step": round(d["ms"], 3),
        "learned_ms_per_step": round(m["ms"], 3),
        "speedup_learned_vs_dynamic": round(d["ms"] / m["ms"], 3),
        "learned_refresh_steps": refresh,
        "ek_rel_err_learned_vs_dynamic": {
            "time_avg_second_half": err(m["avg"], d["avg"]), "time_avg_second_half_k_ge_8": err(m["avg"], d["avg"], 8),
            "final": err(m["final"], d["final"]), "final_k_ge_8": err(m["final"], d["final"], 8)},
        "ek_rel_err_dynamic_perturbed_vs_dynamic": {
            "time_avg_second_half": err(p["avg"], d["avg"]), "time_avg_second_half_k_ge_8": err(p["avg"], d["avg"], 8),
            "final": err(p["final"], d["final"]), "final_k_ge_8": err(p["final"], d["final"], 8)},
        "a_posteriori_check_at_end": m["check"],
        # control: the dynamic model itself with C Delta^2 refreshed every
        # `refresh` steps (the amortisation the learned closure uses)
        "dynamic_lagged_ms_per_step": round(g["ms"], 3),
        "ek_rel_err_dynamic_lagged_vs_dynamic": {
            "time_avg_second_half": err(g["avg"], d["avg"]), "time_avg_second_half_k_ge_8": err(g["avg"], d["avg"], 8),
            "final": err(g["final"], d["final"]), "final_k_ge_8": err(g["final"], d["final"], 8)},
        # the same closure with GELU hidden units (tanh form)
        "learned_gelu_ms_per_step": round(mg["ms"], 3),
        "speedup_learned_gelu_vs_dynamic": round(d["ms"] / mg["ms"], 3),
        "ek_rel_err_learned_gelu_vs_dynamic": {
            "time_avg_second_half": err(mg["avg"], d["avg"]), "time_avg_second_half_k_ge_8": err(mg["avg"], d["avg"], 8),
            "final": err(mg["final"], d["final"]), "final_k_ge_8": err(mg["final"], d["final"], 8)},
        "a_posteriori_check_at_end_gelu": mg["check"],
        "training": {"samples": mlp.pool_size(), "labels": "dynamic model at unit coefficient",
                     "source": f"{n_dns}^3 DNS box-filtered by {r}, 4 snapshots 50 steps apart",
                     "dns_seconds": round(t_dns, 2), "adam_steps": 1200, "batch": 1 << 20,
                     "seconds": round(t_train, 2), "loss_first": float(losses[0]), "loss_last": float(losses[-1]),
                     "gelu_seconds": round(t_train_gelu, 2), "gelu_loss_last": float(losses_gelu[-1])},
    }


def les_dynamic(S, ic, n, steps, warmup):
    """The stratified LES with the dynamic model (dm.f math, (1,4,1) test
    filter) at n^3: seeded k0 = 8 IC, N = 8, nu = 1e-5, FP32 transforms and
    products, FP64 state; `steps` timed steps after `warmup`.  Reports the
    step time and the split over kernel classes (HIP events)."""
    sam = S.Sam(S.make_params(n, sgs=S.SGS_DYN, i_strat=1, grav=64.0, to=1.0, lapse=1.0, vis=1e-5, pr=1.0,
                              i_cfl=1, cfl0=0.5, dt0=0.01, precision=S.PREC_SPLIT))
    sam.set_state(ic.random_solenoidal(n, k0=8.0, seed=20261203, i_strat=1))
    sam.step(warmup)
    sam.sync()
    sam.prof_enable(True)
    t0 = time.perf_counter()
    d = sam.step(steps)
    sam.sync()
    wall = time.perf_counter() - t0
    prof = sam.prof_read()
    sam.close()
    tot = sum(v[0] for v in prof.values()) or 1.0
    return {
        "workload": f"stratified LES {n}^3 with the dynamic model, k0=8, N=8, nu=1e-5, FP32 transforms / FP64 state",
        "ms_per_step": round(wall / steps * 1e3, 3),
        "value": round(n ** 3 * steps / wall / 1e6, 2),
        "unit": "Mpoint-steps/s",
        "steps": steps,
        "warmup": warmup,
        "c_delta_sq": float(d["c_delta_sq"]),
        "kernel_share": {k: round(v[0] / tot, 3) for k, v in prof.items() if v[1]},
    }


def big_split(S, ic, u_coarse, n_big, steps, warmup):
    """One or more RK3 steps of the FP32-split DNS at n_big^3 (the state of
    the bench run, spectrally padded), timed like the main runs."""
    u = ic.spectral_pad(u_coarse, n_big)
    r = run_gpu(n_big, S.PREC_SPLIT, steps, warmup, u, S, keep_state=False)
    del u
    sb = step_bytes(n_big, 8)
    _, pk = dominant_roofline(n_big, 8, r["prof"], steps)
    return {
        "workload": f"stratified Boussinesq DNS {n_big}^3, FP32-split (FP32 transforms/products, FP64 state)",
        "ms_per_step": round(r["wall"] / steps * 1e3, 3),
        "value": round(n_big ** 3 * steps / r["wall"] / 1e6, 2),
        "unit": "Mpoint-steps/s",
        "step_hbm_frac": round(sb / (r["wall"] / steps) / HBM_PEAK, 4),
        "bytes_per_step": sb,
        "kernels": pk,
        "energy": float(r["spec"]["energy_w"]),
        "steps": steps,
        "warmup": warmup,
    }


def cpu_baseline(n_cpu, steps_cpu, u0_small):
    from oracle import oracle as O
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    o = O.Oracle(O.make_cfg(n_cpu, i_strat=1, grav=16.0, to=1.0, lapse0=1.0, vis=5e-4, pr=1.0,
                            i_cfl=1, cfl0=0.5, dt0=0.01), threads=threads)
    t0 = time.perf_counter()
    o.step(u0_small, steps_cpu)
    dt = time.perf_counter() - t0
    val = n_cpu ** 3 * steps_cpu / dt / 1e6
    model = ""
    try:
        with open("/proc/cpuinfo") as fh:
            model = next((ln.split(":", 1)[1].strip() for ln in fh if ln.startswith("model name")), "")
    except OSError:
        pass
    return dict(value=round(val, 3), unit="Mpoint-steps/s", cores=threads, kind="port", cpu_model=model,
                sample=f"{steps_cpu} RK3 step(s) of the same stratified DNS at {n_cpu}^3 by the C FP64 "
                       f"oracle (OpenMP, {threads} threads), {dt:.1f} s")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-split", action="store_true")
    ap.add_argument("--no-les", action="store_true")
    ap.add_argument("--no-big", action="store_true", help="skip the 1024^3 FP32-split step")
    ap.add_argument("--big-n", type=int, default=1024)
    ap.add_argument("--no-les512", action="store_true", help="skip the 512^3 dynamic-model LES step")
    ap.add_argument("--les-steps", type=int, default=2000,
                    help="steps of each 128^3 LES run (a short count for profiler runs)")
    ap.add_argument("--cpu-n", type=int, default=0, help="CPU baseline grid (default: same n if RAM allows)")
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run the plain copy kernel (1 GiB x 3) for HBM counter calibration")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (the step's multi-GPU slab exchange is design-only); "
              f"got --gpus {args.gpus}", file=sys.stderr)
        sys.exit(2)

    from sam_amd import ic
    from sam_amd import samhip as S
    S.load()

    n = args.n
    log = lambda *a: print(*a, file=sys.stderr, flush=True)
    log(f"[bench] building IC {n}^3")
    u0 = ic.random_solenoidal(n, k0=4.0, seed=20261203, i_strat=1)

    log("[bench] FP64 run")
    r64 = run_gpu(n, S.PREC_FP64, args.steps, args.warmup, u0, S)
    ms_step = r64["wall"] / args.steps * 1e3
    value = n ** 3 * args.steps / r64["wall"] / 1e6
    roof, per_kernel = dominant_roofline(n, 16, r64["prof"], args.steps)
    roof["traffic"] = measured_traffic(n, roof["kernel"], "f64")
    sb = step_bytes(n, 16)
    bm64 = bytes_moved(n, 16, r64["prof"], args.steps)
    step_frac = sb / (r64["wall"] / args.steps) / HBM_PEAK
    log(f"[bench] FP64 {ms_step:.2f} ms/step, {value:.1f} Mpt-steps/s, step HBM frac {step_frac:.3f}")

    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mpoint-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "ms_per_step_median_single": round(r64["median"] * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: seeded random-phase solenoidal IC, E(k)~k^4 exp(-2(k/4)^2), u_rms=1, T'=0",
        "config": {"workload": f"stratified Boussinesq DNS {n}^3, N=4, nu=kappa=5e-4, Pr=1, RK3 CFL 0.5, "
                               f"2/3 dealias, FP64 plain", "n": n, "precision": "fp64"},
        "roofline": roof,
        # contract model bytes (fused 5-kernel model) and the bytes the
        # launched kernels move by the same count (less with the fused z inverse)
        "step_hbm": {"bytes_per_step": sb, "frac": round(step_frac, 4),
                     "achieved_GBps": round(sb / (r64["wall"] / args.steps) / 1e9, 1),
                     "bytes_moved_per_step": int(bm64),
                     "frac_moved": round(bm64 / (r64["wall"] / args.steps) / HBM_PEAK, 4)},
        "kernels": per_kernel,
    }

    # measured copy-kernel bandwidth: the practical HBM ceiling next to the 8 TB/s spec
    cal = S.Sam(make_params(S, 64, S.PREC_FP64))
    cal.device_copy(1 << 30, 2)
    t0 = time.perf_counter()
    cal.device_copy(1 << 30, 10)
    t_copy = time.perf_counter() - t0
    out["hbm_copy"] = {"kernel": "k_copy16, 1 GiB read + 1 GiB written, 10 launches",
                       "GBps": round(10 * (2 << 30) / t_copy / 1e9, 1),
                       "frac_of_peak": round(10 * (2 << 30) / t_copy / HBM_PEAK, 4),
                       "step_frac_of_copy": round(sb / (r64["wall"] / args.steps) / (10 * (2 << 30) / t_copy), 4)}
    log(f"[bench] copy kernel {out['hbm_copy']['GBps']} GB/s")
    if args.calibrate_copy:
        cal.device_copy(1 << 30, 3)
        out["copy_calibration_bytes"] = {"read": 1 << 30, "write": 1 << 30}
    cal.close()

    if not args.no_split:
        log("[bench] FP32-split run")
        rsp = run_gpu(n, S.PREC_SPLIT, args.steps, args.warmup, u0, S)
        ms_sp = rsp["wall"] / args.steps * 1e3
        sb32 = step_bytes(n, 8)
        roof32, pk32 = dominant_roofline(n, 8, rsp["prof"], args.steps)
        roof32["traffic"] = measured_traffic(n, roof32["kernel"], "f32")
        e64, e32 = r64["spec"]["energy_w"], rsp["spec"]["energy_w"]
        d64, d32 = np.sum(r64["spec"]["dk"]), np.sum(rsp["spec"]["dk"])
        ek64, ek32 = r64["spec"]["ek"], rsp["spec"]["ek"]
        out["fp32_split"] = {
            "value": round(n ** 3 * args.steps / rsp["wall"] / 1e6, 2),
            "ms_per_step": round(ms_sp, 3),
            "ms_per_step_median_single": round(rsp["median"] * 1e3, 3),
            "step_hbm_frac": round(sb32 / (rsp["wall"] / args.steps) / HBM_PEAK, 4),
            "step_bytes_moved": int(bytes_moved(n, 8, rsp["prof"], args.steps)),
            "step_hbm_frac_moved": round(bytes_moved(n, 8, rsp["prof"], args.steps) / (rsp["wall"] / args.steps)
                                         / HBM_PEAK, 4),
            "roofline": roof32,
            "kernels": pk32,
            "ke_rel_err_vs_fp64": float(abs(e32 - e64) / abs(e64)),
            "diss_rel_err_vs_fp64": float(abs(d32 - d64) / abs(d64)),
            "ek_rel_l2_err_vs_fp64": float(np.linalg.norm(ek32 - ek64) / np.linalg.norm(ek64)),
            "after_steps": args.warmup + args.steps,
        }
        log(f"[bench] split {ms_sp:.2f} ms/step, KE err {out['fp32_split']['ke_rel_err_vs_fp64']:.2e}")

    if not args.no_les and n % 128 == 0 and n >= 256:
        log("[bench] LES 128^3: learned closure vs dynamic model")
        out["les_128"] = les_comparison(S, ic, n, les_steps=args.les_steps)
        lr = out["les_128"]
        log(f"[bench] LES dyn {lr['dynamic_ms_per_step']} ms, learned {lr['learned_ms_per_step']} ms, "
            f"E(k) err {lr['ek_rel_err_learned_vs_dynamic']}")

    if not args.no_les512 and n >= 512:
        log("[bench] LES 512^3 with the dynamic model")
        out["les_512_dynamic"] = les_dynamic(S, ic, 512, 10, 2)
        log(f"[bench] LES 512^3 dynamic: {out['les_512_dynamic']['ms_per_step']} ms/step")

    if not args.no_big and args.big_n > n:
        log(f"[bench] FP32-split {args.big_n}^3")
        out[f"fp32_split_{args.big_n}"] = big_split(S, ic, r64["state"], args.big_n, 2, 1)
        log(f"[bench] split {args.big_n}^3: {out[f'fp32_split_{args.big_n}']['ms_per_step']} ms/step")

    if not args.no_cpu:
        n_cpu = args.cpu_n or n
        try:
            import psutil
            avail = psutil.virtual_memory().available
        except Exception:
            avail = 64 << 30
        need = 230 * n_cpu ** 3   # oracle working set, bytes (about 28 GB at 512^3)
        while need > 0.6 * avail and n_cpu > 64:
            n_cpu //= 2
            need = 230 * n_cpu ** 3
        log(f"[bench] CPU baseline {n_cpu}^3 x {args.cpu_steps}")
        u_small = u0 if n_cpu == n else ic.random_solenoidal(n_cpu, k0=4.0, seed=20261203, i_strat=1)
        out["cpu_baseline"] = cpu_baseline(n_cpu, args.cpu_steps, u_small)
    else:
        out["cpu_baseline"] = None

    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
