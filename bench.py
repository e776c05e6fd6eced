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


def kernel_bytes(n, bc, i_strat=1, zrhs_fused=False, nrk=3):
    """Algorithmic HBM bytes per launch of each step kernel (fused 3-pass
    model): state/RHS in FP64 (16 B), transform elements bc = 16 or 8 B.
    zrhs: the mean over the nrk launches of a step -- every substep reads the
    state and the fluxes and writes the state and the RHS, and all but the
    first read the previous substep's RHS (the zeta term).  zrhs_fused:
    k_zrhs also writes the next substep's z inverse (Lu Z fields), which k_zb
    then does not compute."""
    g = grid_counts(n, i_strat)
    S, Z, Y, lu, lr = g["S"], g["Z"], g["Y"], g["lu"], g["lr"]
    return {
        "zb": lu * S * 16 + lu * Z * bc,
        "yb": lu * Z * bc + lu * Y * bc,
        "xrow": lu * Y * bc + lr * Y * bc,
        "yf": lr * Y * bc + lr * Z * bc,
        "zrhs": lr * Z * bc + (3 + (nrk - 1) / nrk) * lu * S * 16 + (lu * Z * bc if zrhs_fused else 0),
    }


def kernel_rw(n, bc, i_strat=1, zrhs_fused=False, nrk=3):
    """kernel_bytes split into (bytes read, bytes written) per launch"""
    g = grid_counts(n, i_strat)
    S, Z, Y, lu, lr = g["S"], g["Z"], g["Y"], g["lu"], g["lr"]
    return {
        "zb": (lu * S * 16, lu * Z * bc),
        "yb": (lu * Z * bc, lu * Y * bc),
        "xrow": (lu * Y * bc, lr * Y * bc),
        "yf": (lr * Y * bc, lr * Z * bc),
        "zrhs": (lr * Z * bc + (1 + (nrk - 1) / nrk) * lu * S * 16,
                 2 * lu * S * 16 + (lu * Z * bc if zrhs_fused else 0)),
    }


# (arrays read, arrays written) per launch of the streaming-mix kernel
# (sam_device_stream); a kernel's practical ceiling is the mix nearest its
# own read:write proportions
STREAM_MIXES = ((1, 0), (0, 1), (1, 1), (1, 2), (2, 1), (2, 3), (3, 2), (4, 5), (5, 4), (4, 9))


def nearest_mix(rd, wr):
    def key(m):
        a, b = m
        if a == 0 or b == 0 or rd == 0 or wr == 0:
            return 0.0 if (a == 0) == (rd == 0) and (b == 0) == (wr == 0) else 1e9
        return abs(math.log((a / b) / (rd / wr)))
    return min(STREAM_MIXES, key=key)


def stream_ceilings(S, mixes, nbytes=1 << 30, reps=8, trials=3):
    """sustained GB/s (bytes read + written) of the streaming kernel at each
    read:write mix, nt policy, on this device: the best of `trials` timings of
    `reps` launches (each launch moves (a + b) GiB, so the call overhead is
    well under 1%)"""
    cal = S.Sam(make_params(S, 64, S.PREC_FP64))
    out = {}
    for a, b in mixes:
        cal.device_stream(a, b, nbytes, 1)
        best = 0.0
        for _ in range(trials):
            t0 = time.perf_counter()
            cal.device_stream(a, b, nbytes, reps)
            best = max(best, reps * (a + b) * nbytes / (time.perf_counter() - t0) / 1e9)
        out[(a, b)] = best
    cal.close()
    return out


def with_ceilings(per_kernel, rw, ceil):
    """per-kernel mix, ceiling and fraction of it (in place)"""
    for k, v in per_kernel.items():
        if k not in rw:
            continue
        rd, wr = rw[k]
        m = nearest_mix(rd, wr)
        v["read_write_GB"] = [round(rd / 1e9, 3), round(wr / 1e9, 3)]
        v["mix"] = f"{m[0]}:{m[1]}"
        v["ceiling_GBps"] = round(ceil[m], 1)
        v["frac_of_ceiling"] = round(v["GBps"] / ceil[m], 4)
    return per_kernel


def step_floor_ms(n, bc, prof, steps, ceil):
    """the step's time if every kernel ran at its mix's ceiling (launch
    counts of the run)"""
    rw = kernel_rw(n, bc, zrhs_fused=zrhs_fused(prof))
    t = 0.0
    for k, (rd, wr) in rw.items():
        cnt = prof.get(k, (0.0, 0))[1]
        if cnt:
            t += (rd + wr) / (ceil[nearest_mix(rd, wr)] * 1e9) * cnt
    return t / steps * 1e3


def zrhs_fused(prof):
    """the timed region's k_zrhs wrote the next z inverse: fewer k_zb than
    k_zrhs launches (k_zb runs only where the previous k_zrhs did not)"""
    return prof.get("zb", (0.0, 0))[1] < prof.get("zrhs", (0.0, 0))[1]


def bytes_moved(n, bc, prof, steps):
    """bytes per step the launched kernels move by the same model: the
    contract's count (step_bytes) less the state reads a fused k_zrhs saves
    and the RHS read the first substep does not make"""
    kb = kernel_bytes(n, bc, zrhs_fused=zrhs_fused(prof))
    return sum(kb[k] * prof.get(k, (0.0, 0))[1] for k in kb) / steps


def step_bytes(n, bc, i_strat=1, nrk=3):
    """the contract's model: per substep 5 Lu S 16 + 2 (Lu + Lr)(Z + Y) b_c
    (it counts a read of the previous RHS in every substep)"""
    g = grid_counts(n, i_strat)
    per_sub = 5 * g["lu"] * g["S"] * 16 + 2 * (g["lu"] + g["lr"]) * (g["Z"] + g["Y"]) * bc
    return nrk * per_sub


def step_bytes_algorithmic(n, bc, i_strat=1, nrk=3):
    """the step's algorithmic bytes: the contract's less the RHS read of the
    first substep, which the RK3 update does not make (zeta_1 = 0)"""
    g = grid_counts(n, i_strat)
    return step_bytes(n, bc, i_strat, nrk) - g["lu"] * g["S"] * 16


def les_kernel_bytes(n, bc, bp, i_strat=1, nrk=3, zrhs_fused=False):
    """Algorithmic HBM bytes per launch of the LES step's kernels (the stage
    list of the step with an SGS model; physical fields of bp = 8 or 4 bytes
    per value, P points per field):
      zb, yb      as in the DNS step (k_zb runs every substep unless k_zrhs
                  writes the next z inverse: the persistent form, Nz = 512)
      xb          Lu fields Y -> P                      (x inverse to physical)
      stats       reads Lu P fields                     (vel_max, first substep)
      dmproc      reads u, v, w once                    (fused dynamic procedure)
      sgs         reads Lu P fields, writes Lr P fields (eddy-viscosity or
                  network fluxes)
      xf          Lr fields P -> Y                      (x forward)
      yf          as in the DNS step
      zrhs        as in the DNS step, with the fused z inverse where it runs"""
    g = grid_counts(n, i_strat)
    S, Z, Y, P, lu, lr = g["S"], g["Z"], g["Y"], g["P"], g["lu"], g["lr"]
    kb = kernel_bytes(n, bc, i_strat, zrhs_fused=zrhs_fused, nrk=nrk)
    return {
        "zb": kb["zb"], "yb": kb["yb"],
        "xb": lu * Y * bc + lu * P * bp,
        "stats": lu * P * bp,
        "dmproc": 3 * P * bp,
        "sgs": (lu + lr) * P * bp,
        "xf": lr * P * bp + lr * Y * bc,
        "yf": kb["yf"], "zrhs": kb["zrhs"],
    }


def les_roofline(n, bc, bp, prof, steps):
    """per-kernel achieved bandwidth of an LES run, the dominant kernel's
    roofline and the step's algorithmic bytes (launch counts from the run)"""
    kb = les_kernel_bytes(n, bc, bp, zrhs_fused=zrhs_fused(prof))
    per, tot_bytes = {}, 0.0
    tot_ms = sum(v[0] for v in prof.values()) or 1.0
    best = None
    for k, nb in kb.items():
        ms, cnt = prof.get(k, (0.0, 0))
        if not cnt:
            continue
        tot_bytes += nb * cnt
        avg = ms / cnt / 1e3
        per[k] = dict(avg_us=round(avg * 1e6, 2), GBps=round(nb / avg / 1e9, 1), share=round(ms / tot_ms, 3),
                      launches_per_step=round(cnt / steps, 2))
        if best is None or ms > best[1]:
            best = (k, ms, cnt, nb)
    k, ms, cnt, nb = best
    avg = ms / cnt / 1e3
    roof = dict(kernel=k, bound="hbm", achieved=round(nb / avg / 1e9, 1), peak=HBM_PEAK / 1e9, unit="GB/s",
                frac=round(nb / avg / HBM_PEAK, 4), bytes_per_launch=nb, avg_launch_us=round(avg * 1e6, 2),
                launches=cnt)
    if k == "dmproc":
        # the fused dynamic procedure reads u, v, w once and spends its time on
        # the 21-field stencil and filter in LDS and VALU: HBM is not its bound
        roof["bound"] = "lds/valu"
        roof["note"] = "achieved/frac: HBM bytes (u, v, w once) over the launch time, for reference"
    return roof, per, tot_bytes / steps


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
    """The learned closure against the dynamic model on a 128^3 stratified LES.

    Training data: the stratified DNS of the closure-training workload at
    n_dns^3 (N = 4, seeded k0 = 4 IC, nu = 5e-4, FP32-split), 100 steps to
    develop, then 8 snapshots 50 steps apart box-filtered by n_dns/n_les.
    Three networks learn from the same snapshots, 1200 Adam steps of 2^20
    points each:
      learned_gelu     GELU, the dynamic model's stress at unit coefficient;
                       in the LES scaled by C Delta^2 from the dynamic
                       procedure every `refresh` steps;
      learned_relu     the same with ReLU hidden units;
      learned_gelu_coef GELU, the dynamic model's stress with the coefficient
                       of the dynamic procedure run on each filtered snapshot;
                       the LES uses it with no procedure at all.
    LES (N = 8, nu = 1e-5, the seeded k0 = 6 IC truncated to 128^3):
    `les_steps` steps with the dynamic model (dm.f, the procedure every
    substep), with each network, with the dynamic model from an IC perturbed
    by 1e-6 (the trajectories' chaotic spread, the floor of any comparison),
    and with the dynamic model whose C Delta^2 is refreshed every `refresh`
    steps (the control for the refresh).  Whole-run wall time per step; E(k)
    compared at the end and time-averaged over the second half (every 50
    steps); the per-kernel split and rooflines of the dynamic-model run."""
    from sam_amd import les as L
    r = n_dns // n_les
    phys_dns = dict(i_strat=1, grav=16.0, to=1.0, lapse=1.0, vis=5e-4, pr=1.0, i_cfl=1, cfl0=0.5, dt0=0.01,
                    precision=S.PREC_SPLIT)
    phys = dict(phys_dns, grav=64.0, vis=1e-5)
    nets = {"learned_gelu": (S.ACT_GELU, S.LABELS_DYNAMIC), "learned_relu": (S.ACT_RELU, S.LABELS_DYNAMIC),
            "learned_gelu_coef": (S.ACT_GELU, S.LABELS_DYNAMIC_COEF)}
    nsnap = 8
    mlps = {}
    for name, (act, lab) in nets.items():
        m = S.Mlp(seed=1, pool_capacity=nsnap * n_les ** 3)
        m.set_activation(act)
        m.set_labels(lab, 1.0)
        mlps[name] = m
    dns = S.Sam(S.make_params(n_dns, **phys_dns))
    dns.set_state(ic.random_solenoidal(n_dns, k0=4.0, seed=20261203, i_strat=1))
    t0 = time.perf_counter()
    dns.step(100)
    coefs = []
    for _ in range(nsnap):
        dns.step(50)
        for m in mlps.values():
            m.add_samples(dns, r)
        coefs.append(mlps["learned_gelu_coef"].last_coefficient())
    dns.close()
    t_dns = time.perf_counter() - t0
    train = {}
    for name, m in mlps.items():
        t0 = time.perf_counter()
        losses = L.train(m, 1200, batch=1 << 20, lr=1e-3, seed=7)
        train[name] = dict(seconds=round(time.perf_counter() - t0, 2), loss_first=float(losses[0]),
                           loss_last=float(losses[-1]))
    u_les = L.spectral_truncate(ic.random_solenoidal(n_dns, k0=6.0, seed=20261203, i_strat=1), n_les)
    rng = np.random.default_rng(5)
    runs = [("dynamic", S.SGS_DYN, None, u_les),
            ("dynamic_perturbed", S.SGS_DYN, None, u_les * (1 + 1e-6 * rng.standard_normal(u_les.shape))),
            ("dynamic_lagged", S.SGS_DYN, None, u_les)] + [(name, S.SGS_MLP, name, u_les) for name in nets]
    half = max(1, les_steps // 2)
    chunk = min(50, max(1, half // 2))    # steps between E(k) samples of the second half
    res = {}
    for name, model, net, u in runs:
        les = S.Sam(S.make_params(n_les, sgs=model, **phys))
        if net:
            S.attach_mlp(les, mlps[net])
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
        nsteps = half + chunk * (half // chunk)
        prof = None
        if name == "dynamic":
            # the kernel split from a separate run with HIP events (which add
            # per-launch overhead the timed runs do not carry)
            les.prof_enable(True)
            les.step(200)
            les.sync()
            prof = les.prof_read()
            les.prof_enable(False)
        check = None
        if model == S.SGS_MLP:
            err, sw = S.sgs_check(les, 1.0)
            check = {"sgs_transfer_rel_err_vs_dynamic": round(float(err), 4), "fell_back": bool(sw)}
        res[name] = dict(ms=wall / nsteps * 1e3, final=L.shell_energy(les.spectra()), avg=acc / ns, check=check,
                         prof=prof, steps=nsteps)
        les.close()
    for m in mlps.values():
        m.close()

    def err(a, b, lo=1):
        k1 = len(b) - 1
        return round(float(np.linalg.norm(a[lo:k1] - b[lo:k1]) / np.linalg.norm(b[lo:k1])), 4)

    def errs(x, ref):
        return {"time_avg_second_half": err(x["avg"], ref["avg"]),
                "time_avg_second_half_k_ge_8": err(x["avg"], ref["avg"], 8),
                "final": err(x["final"], ref["final"]), "final_k_ge_8": err(x["final"], ref["final"], 8)}

    d = res["dynamic"]
    roof, per, les_bytes = les_roofline(n_les, 8, 4, d["prof"], 200)
    out = {
        "workload": f"stratified LES {n_les}^3, N=8, nu=1e-5, IC: the seeded k0=6 DNS IC truncated, "
                    f"FP32 transforms / FP64 state, {les_steps} steps",
        "dynamic_ms_per_step": round(d["ms"], 3),
        "dynamic_roofline": roof,
        "dynamic_step_hbm_frac": round(les_bytes / (d["ms"] / 1e3) / HBM_PEAK, 4),
        "dynamic_kernels": per,
        "ek_rel_err_dynamic_perturbed_vs_dynamic": errs(res["dynamic_perturbed"], d),
        "dynamic_lagged_ms_per_step": round(res["dynamic_lagged"]["ms"], 3),
        "ek_rel_err_dynamic_lagged_vs_dynamic": errs(res["dynamic_lagged"], d),
        "learned_refresh_steps": {"learned_gelu": refresh, "learned_relu": refresh, "learned_gelu_coef": None},
        "training": {"source": f"{n_dns}^3 DNS (N=4, k0=4, nu=5e-4) box-filtered by {r}, {nsnap} snapshots "
                               f"50 steps apart after 100", "samples": nsnap * n_les ** 3, "batch": 1 << 20,
                     "adam_steps": 1200, "dns_seconds": round(t_dns, 2),
                     "c_delta_sq_of_snapshots": [float("%.4e" % c) for c in coefs], "nets": train},
    }
    for name in nets:
        x = res[name]
        out[f"{name}_ms_per_step"] = round(x["ms"], 3)
        out[f"speedup_{name}_vs_dynamic"] = round(d["ms"] / x["ms"], 3)
        out[f"speedup_{name}_vs_dynamic_lagged"] = round(res["dynamic_lagged"]["ms"] / x["ms"], 3)
        out[f"ek_rel_err_{name}_vs_dynamic"] = errs(x, d)
        out[f"a_posteriori_check_at_end_{name}"] = x["check"]
    return out


def les_dynamic(S, ic, n, steps, warmup):
    """The stratified LES with the dynamic model (dm.f math, (1,4,1) test
    filter, the procedure fused into one stencil pass) at n^3: seeded k0 = 8
    IC, N = 8, nu = 1e-5, FP32 transforms and products, FP64 state; `steps`
    timed steps after `warmup`.  Reports the step time, the split over kernel
    classes (HIP events), the per-kernel bandwidths by the LES bytes model,
    the dominant kernel's roofline and the step's fraction of HBM peak."""
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
    roof, per, les_bytes = les_roofline(n, 8, 4, prof, steps)
    return {
        "workload": f"stratified LES {n}^3 with the dynamic model, k0=8, N=8, nu=1e-5, FP32 transforms / FP64 state",
        "ms_per_step": round(wall / steps * 1e3, 3),
        "value": round(n ** 3 * steps / wall / 1e6, 2),
        "unit": "Mpoint-steps/s",
        "steps": steps,
        "warmup": warmup,
        "c_delta_sq": float(d["c_delta_sq"]),
        "roofline": roof,
        "bytes_per_step": int(les_bytes),
        "step_hbm_frac": round(les_bytes / (wall / steps) / HBM_PEAK, 4),
        "kernels": per,
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
                     "bytes_algorithmic_per_step": step_bytes_algorithmic(n, 16),
                     "frac_algorithmic": round(step_bytes_algorithmic(n, 16) / (r64["wall"] / args.steps) / HBM_PEAK,
                                               4),
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
    # practical ceilings by read:write mix (the write rate is the lower one)
    ceil = stream_ceilings(S, STREAM_MIXES)
    out["hbm_mix_ceilings"] = {
        "kernel": "k_stream: a GiB read from one array, b GiB written to another per launch, in 16 KiB chunks (nt, 16-byte accesses), best of 3 x 8 launches",
        "GBps_by_read_write": {f"{a}:{b}": round(v, 1) for (a, b), v in ceil.items()}}
    rw64 = kernel_rw(n, 16, zrhs_fused=zrhs_fused(r64["prof"]))
    with_ceilings(out["kernels"], rw64, ceil)
    kmix = nearest_mix(*rw64[roof["kernel"]])
    roof["practical_ceiling"] = {"GBps": round(ceil[kmix], 1), "read_write_mix": f"{kmix[0]}:{kmix[1]}",
                                 "frac": round(roof["achieved"] / ceil[kmix], 4)}
    fl64 = step_floor_ms(n, 16, r64["prof"], args.steps, ceil)
    out["step_hbm"]["floor_ms_at_mix_ceilings"] = round(fl64, 3)
    out["step_hbm"]["frac_of_mix_ceilings"] = round(fl64 / ms_step, 4)
    log(f"[bench] mix ceilings {out['hbm_mix_ceilings']['GBps_by_read_write']}; step floor {fl64:.2f} ms")
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
        rw32 = kernel_rw(n, 8, zrhs_fused=zrhs_fused(rsp["prof"]))
        with_ceilings(pk32, rw32, ceil)
        m32 = nearest_mix(*rw32[roof32["kernel"]])
        roof32["practical_ceiling"] = {"GBps": round(ceil[m32], 1), "read_write_mix": f"{m32[0]}:{m32[1]}",
                                       "frac": round(roof32["achieved"] / ceil[m32], 4)}
        fl32 = step_floor_ms(n, 8, rsp["prof"], args.steps, ceil)
        out["fp32_split"] = {
            "value": round(n ** 3 * args.steps / rsp["wall"] / 1e6, 2),
            "step_floor_ms_at_mix_ceilings": round(fl32, 3),
            "step_frac_of_mix_ceilings": round(fl32 / ms_sp, 4),
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
        log(f"[bench] LES dyn {lr['dynamic_ms_per_step']} ms, lagged {lr['dynamic_lagged_ms_per_step']} ms, "
            f"learned GELU {lr['learned_gelu_ms_per_step']} ms (E(k) {lr['ek_rel_err_learned_gelu_vs_dynamic']}), "
            f"GELU with coefficient {lr['learned_gelu_coef_ms_per_step']} ms "
            f"(E(k) {lr['ek_rel_err_learned_gelu_coef_vs_dynamic']})")

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
