"""GPU parity against the reference itself: libsamhip (through its C ABI)
against tests/golden/ fixtures written by the reference's own compiled code
(whole `sam` runs, rhs.f / vel_max.f / spectra.f / div_check.f outputs, the
transform routines' outputs), and the stage entry points in the reference's
layouts.  FP64 bar 1e-12 relative L2 (the drop-in tolerance)."""
import glob
import json
import os

import numpy as np
import pytest

from sam_amd import ic
from sam_amd import samhip as S

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")
TOL = 1e-12


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def runs():
    g = np.load(os.path.join(GOLD, "sam_runs.npz"))
    return g, json.loads(str(g["cases_json"]))


def sam_of(p, lapse=None, prec=S.PREC_FP64):
    """a handle for an input.dat label set (set_labels.f defaults)"""
    strat = int(p.get("i_strat", 0))
    if lapse is None:
        lapse = p.get("lapse0", 0.0) if strat else 0.0
    return S.Sam(S.make_params(int(p["Nx"]), int(p["Ny"]), int(p["Nz"]), xl=p["xL"], yl=p["yL"], zl=p["zL"],
                               i_strat=strat, nrk_max=int(p.get("nrk_max", 3)), i_cfl=int(p.get("i_cfl", 1)),
                               vis=p["vis"], pr=p.get("Pr", 1.0), grav=p.get("grav", 0.0), to=p.get("To", 1.0),
                               lapse=lapse, cfl0=p.get("cfl0", 1.0), dt0=p.get("dt0", 0.01), precision=prec))


@pytest.mark.parametrize("name", ["tg32", "strat32", "nrk1_16", "nrk2_16", "restart16", "force16"])
def test_steps_match_reference_runs(name):
    """whole runs of the reference program (sam.f:122-326): RK3 from a
    restart and from initial_field.f's fresh stratified start, Euler and RK2
    (init.f:117-134), the restart with lapse = 0, force.f after every step"""
    g, cases = runs()
    p = cases[name]
    s = sam_of(p, lapse=0.0 if name == "restart16" else None)
    u0 = ic.taylor_green(32) if name == "tg32" else g[f"{name}_u0"]
    ns = (1, 10) if name in ("tg32", "strat32") else (5,)
    s.set_state(u0)
    if name == "force16":
        S.set_forcing(s, int(p["k_force"]))
        s.rhs(want_spectra=True)          # sam.f:248-250: the nt = 0 spectrum is force.f's target
    done = 0
    for n in ns:
        d = s.step(n - done)
        done = n
        got = s.get_state()
        assert rel(got, g[f"{name}_u{n}"]) < TOL, (name, n, rel(got, g[f"{name}_u{n}"]))
        t_ref = g[f"{name}_clock{n}"][0]
        assert abs(d["time"] - t_ref) <= 1e-13 * max(1.0, t_ref)
    s.close()


def test_mid_step_velocity_matches_reference():
    """sam_get_state_substep(nrk_max) is the velocity the reference writes to
    vel.NNNNNN inside a step (sam.f:277-284, n_skip_v 2), and the handle
    steps on unchanged after it"""
    g, cases = runs()
    p = cases["midrk16"]
    s = sam_of(p)
    s.set_state(g["midrk16_u0"])
    s.step(2)
    mid = s.get_state_substep(3)
    assert rel(mid, g["midrk16_u2"]) < TOL, rel(mid, g["midrk16_u2"])
    assert rel(s.get_state_substep(1), s.get_state()) == 0.0
    d = s.step(2)
    assert rel(s.get_state(), g["midrk16_u4"]) < TOL
    assert abs(d["time"] - g["midrk16_clock4"][0]) <= 1e-13


@pytest.mark.parametrize("name", ["16u", "32s"])
def test_rhs_spectra_div_match_reference(name):
    """rhs.f (first substep), vel_max.f, spectra.f and div_check.f outputs of
    the reference; the divergence sums also on a field with divergence"""
    f = np.load(os.path.join(GOLD, f"rhs_{name}.npz"))
    n, strat, icfl, seed = (int(v) for v in f["params"])
    s = S.Sam(S.make_params(n, i_strat=strat, grav=16.0, lapse=1.0, vis=5e-3, i_cfl=icfl, cfl0=0.5, dt0=0.01))
    s.set_state(f["u"])
    r, d = s.rhs(want_spectra=True)
    assert rel(r, f["r"]) < TOL
    st = f["stats"]
    lu = s.lu
    np.testing.assert_allclose(d["u_var"][:lu], st[4:4 + lu], rtol=1e-12)
    np.testing.assert_allclose(d["u_max"][:lu], st[8:8 + lu], rtol=1e-12)
    np.testing.assert_allclose([d["energy"], *d["cfl"], d["dt"]], st[12:17], rtol=1e-12)
    sp = s.spectra()
    for k in ("ek", "tk", "dk", "vt", "sample"):
        np.testing.assert_allclose(sp[k], f[k], rtol=1e-11, atol=1e-300)
    ew, drms, dmax, _ = f["scalars"]
    assert abs(sp["energy_w"] - ew) <= 1e-12 * ew
    assert sp["div_max"] < 1e-12 and drms < 1e-12
    dv = s.div_check()
    assert dv["n_div"] == f["div_check"][2]
    # a field with divergence: the state-only outputs of spectra.f and div_check.f
    s.set_state(f["w"])
    s.rhs(want_spectra=True)
    sp = s.spectra()
    for k in ("ek", "dk", "sample"):
        np.testing.assert_allclose(sp[k], f[f"w_{k}"], rtol=1e-11, atol=1e-300)
    ew, drms, dmax, _ = f["w_scalars"]
    assert abs(sp["energy_w"] - ew) <= 1e-12 * ew
    assert abs(sp["div_rms"] - drms) <= 1e-12 * drms
    assert abs(sp["div_max"] - dmax) <= 1e-12 * dmax
    dv = s.div_check()
    ref = f["w_div_check"]
    assert abs(dv["div_max"] - ref[0]) <= 1e-12 * ref[0]
    assert abs(dv["div_rms"] - ref[1]) <= 1e-12 * ref[1]
    assert dv["n_div"] == ref[2] and ref[2] > 0
    s.close()


@pytest.mark.parametrize("name", ["16u", "32s"])
def test_cl_diagnostics_match_reference(name):
    """the Chollet-Lesieur diagnostics (sam_set_cl(h, 1), i_les = 1) of an
    RHS pass: cl_energy.f and spectra.f's dk, vt of the reference's own
    routines on the same state"""
    f = np.load(os.path.join(GOLD, f"rhs_{name}.npz"))
    n, strat, icfl, seed = (int(v) for v in f["params"])
    s = S.Sam(S.make_params(n, i_strat=strat, grav=16.0, lapse=1.0, vis=5e-3, i_cfl=icfl, cfl0=0.5, dt0=0.01))
    S.set_cl(s, S.CL_DIAG)
    s.set_state(f["u"])
    s.rhs(want_spectra=True)
    sp = s.spectra()
    s.close()
    for k in ("ek", "tk", "dk", "vt"):
        np.testing.assert_allclose(sp[k], f[f"cl_{k}"], rtol=1e-11, atol=1e-300)


GOLD_T = [p for p in sorted(glob.glob(os.path.join(GOLD, "transforms_*.npz")))
          if min(int(v) for v in np.load(p)["dims"][:3]) >= 16]


@pytest.mark.parametrize("path", GOLD_T, ids=lambda p: os.path.basename(p))
def test_stage_layouts_match_reference(path):
    """z_trans_b.f, x2z_decomp.f, xy_trans_b.f, xy_trans_f.f, z2x_decomp.f,
    z_trans_f.f outputs of the reference's own routines, each stage entry
    point in the reference's array layout"""
    g = np.load(path)
    nx, ny, nz, nd, strat = (int(v) for v in g["dims"][:5])
    xl, yl, zl = g["box"]
    s = S.Sam(S.make_params(nx, ny, nz, xl=xl, yl=yl, zl=zl, n_dealias=nd, i_strat=strat))
    u = g["u"][: s.nzm]
    assert rel(s.z_trans_b(u), g["z_trans_b"]) < TOL
    assert rel(s.x2z(g["z_trans_b"]), g["x2z"]) < TOL
    assert rel(s.xy_trans_b(g["x2z"]), g["xy_trans_b"]) < TOL
    assert rel(s.xy_trans_f(g["p_in"]), g["xy_trans_f"]) < TOL
    assert rel(s.z2x(g["xy_trans_f"]), g["xy_trans_f"]) == 0.0
    assert rel(s.z_trans_f(g["xy_trans_f"]), g["z_trans_f"]) < TOL
    s.close()


@pytest.mark.parametrize("strat,nrk_max", [(0, 3), (1, 3), (1, 2), (1, 1)])
def test_rk_update_matches_oracle(oracle_mod, strat, nrk_max):
    """sam.f:210-221 and 293-314 as their own entry point, against the oracle"""
    O = oracle_mod
    n = 16
    o = O.Oracle(O.make_cfg(n, i_strat=strat, nrk_max=nrk_max, vis=3e-2, pr=0.7))
    s = S.Sam(S.make_params(n, i_strat=strat, nrk_max=nrk_max, vis=3e-2, pr=0.7))
    rng = np.random.default_rng(5)
    sh = (s.nzm, s.nym, s.nxm, s.lu)
    u, rp, r = (np.asfortranarray(rng.standard_normal(sh) + 1j * rng.standard_normal(sh)) for _ in range(3))
    dt = 0.037
    for nrk in range(1, nrk_max + 1):
        for a, b in ((rp, r), (None, r), (rp, None)):
            got = s.rk_update(nrk, dt, u, r_prev=a, r=b)
            ref = o.rk_update(nrk, dt, u, r_prev=a, r=b)
            assert rel(got, ref) < 1e-14, (nrk, a is None, b is None)
    s.close()


def test_sgs_check_leaves_the_run_unchanged():
    """sam_sgs_check runs three RHS passes on the current state; the state,
    RHS, diagnostics and coefficient counters are restored, so a run with a
    check steps exactly as one without"""
    n = 32
    kw = dict(i_strat=1, grav=64.0, lapse=1.0, vis=1e-5, i_cfl=1, cfl0=0.5, sgs=S.SGS_MLP, pr_t=1.0)
    u0 = ic.random_solenoidal(n, k0=6, seed=41, i_strat=1)
    net = S.Mlp(seed=3, pool_capacity=0)
    net.set_labels(S.LABELS_DYNAMIC)
    runs_ = []
    for check in (False, True):
        s = S.Sam(S.make_params(n, **kw))
        S.attach_mlp(s, net)
        S.set_mlp_refresh(s, 3)
        s.set_state(u0)
        s.step(2)
        if check:
            S.sgs_check(s, 1e30)
        d = s.step(3)
        runs_.append((s.get_state(), d))
        s.close()
    net.close()
    assert rel(runs_[1][0], runs_[0][0]) == 0.0
    assert runs_[1][1]["c_delta_sq"] == runs_[0][1]["c_delta_sq"]
    assert runs_[1][1]["dt"] == runs_[0][1]["dt"]
