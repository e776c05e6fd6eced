"""Pin the oracle to the reference program itself -- CPU only.

tests/golden/sam_runs.npz holds whole runs of the reference's `sam`
(sam.f:122-326 with rhs.f, vel_max.f, spectra.f, the transforms, force.f
and initial_field.f's ICs), compiled and run in the development container by
tests/golden/make_ref_runs.py.  tests/golden/rhs_*.npz hold the reference's
rhs.f, vel_max.f, spectra.f and div_check.f outputs on seeded states
(make_golden.py).  The oracle must reproduce both to 1e-12.
"""
import json
import os

import numpy as np
import pytest

from oracle import oracle as O
from sam_amd import ic

GOLD = os.path.join(os.path.dirname(__file__), "golden")
RUNS = np.load(os.path.join(GOLD, "sam_runs.npz"))
CASES = json.loads(str(RUNS["cases_json"]))
TOL = 1e-12


def rel(a, b):
    return float(np.linalg.norm((np.asarray(a) - np.asarray(b)).ravel()) / max(np.linalg.norm(np.asarray(b).ravel()),
                                                                              1e-300))


def cfg_of(p, lapse=None):
    """oracle configuration of an input.dat label set (set_labels.f defaults)"""
    return O.make_cfg(int(p["Nx"]), int(p["Ny"]), int(p["Nz"]), xl=p["xL"], yl=p["yL"], zl=p["zL"],
                      i_strat=int(p.get("i_strat", 0)), nrk_max=int(p.get("nrk_max", 3)), i_cfl=int(p.get("i_cfl", 1)),
                      vis=p["vis"], pr=p.get("Pr", 1.0), grav=p.get("grav", 0.0), to=p.get("To", 1.0),
                      lapse0=p.get("lapse0", 0.0), lapse=lapse, cfl0=p.get("cfl0", 1.0), dt0=p.get("dt0", 0.01),
                      k_force=int(p.get("k_force", 2)) if p.get("i_force") else -1)


def run_case(name):
    """(oracle, initial state, [(nsteps, expected state, expected time, expected header dt)])"""
    p = CASES[name]
    lapse = 0.0 if name == "restart16" else None       # a restart leaves lapse = 0 (initial_field.f:37)
    o = O.Oracle(cfg_of(p, lapse))
    if name == "tg32":
        u0 = ic.taylor_green(32)
        ns = (1, 10)
        tag = "tg32"
    elif name == "strat32":
        u0 = RUNS["strat32_u0"]
        ns = (1, 10)
        tag = "strat32"
    else:
        u0 = RUNS[f"{name}_u0"]
        ns = (5,)
        tag = name
    return o, u0, [(n, RUNS[f"{tag}_u{n}"], *RUNS[f"{tag}_clock{n}"]) for n in ns]


@pytest.mark.parametrize("name", ["tg32", "strat32", "nrk1_16", "nrk2_16", "restart16", "force16"])
def test_oracle_matches_reference_runs(name):
    o, u0, checks = run_case(name)
    ek0 = None
    if name == "force16":
        # sam.f:248-250: the spectrum of nt = 0 becomes force.f's target ek_0
        r0, _ = o.rhs(u0)
        ek0 = o.spectra(u0, r0)["ek"]
    for n, u_ref, t_ref, dt_hdr in checks:
        u, t, dt, st = o.step(u0, n, ek0=ek0)
        assert rel(u, u_ref) < TOL, (name, n, rel(u, u_ref))
        assert abs(t - t_ref) <= 1e-13 * max(1.0, abs(t_ref))
        # header.* carries the dt that vel_max.f set at the nt_end substep
        _, st_end = o.rhs(u)
        assert abs(st_end[16] - dt_hdr) <= 1e-12 * dt_hdr, (st_end[16], dt_hdr)


def test_reference_runs_are_distinct():
    """the fixtures pin what they claim: a restart really runs with lapse = 0,
    the Euler/RK2 runs differ from RK3, force.f changes the state"""
    o, u0, checks = run_case("restart16")
    o2 = O.Oracle(cfg_of(CASES["restart16"]))               # lapse = lapse0
    assert rel(o2.step(u0, 5)[0], checks[0][1]) > 1e-3
    for name in ("nrk1_16", "nrk2_16"):
        p = dict(CASES[name], nrk_max=3)
        o3 = O.Oracle(cfg_of(p))
        assert rel(o3.step(RUNS[f"{name}_u0"], 5)[0], RUNS[f"{name}_u5"]) > 1e-6
    o, u0, checks = run_case("force16")
    assert rel(o.step(u0, 5)[0], checks[0][1]) > 1e-8


def mid_step_state(o, u, substeps):
    """The velocity sam.f:277-284 writes inside a step: the state after
    `substeps` substeps plus the zeta term of the next one, composed from
    the oracle's RHS and the integrating factor (sam.f:210-221, 293-314)."""
    kx, ky, kz = o.kvectors()
    wx, wy, wz = o.wave
    k2 = ((wx * kx)[None, None, :] ** 2 + (wy * ky)[None, :, None] ** 2 + (wz * kz)[:, None, None] ** 2)
    u = np.array(u, order="F", copy=True)
    r_prev = None
    dt = None
    for m in range(substeps + 1):
        r, st = o.rhs(u)
        if m == 0:
            dt = st[16] if o.cfg.i_cfl == 1 else o.cfg.dt0
        if m > 0:
            u = u + o.zeta[m] * dt * r_prev
        if m == substeps:
            return u
        vis_f = o.cfg.vis * (o.gamma[m] + o.zeta[m]) * dt
        for f in range(o.lu):
            v = vis_f / o.cfg.pr if f == 3 else vis_f
            u[..., f] = (u[..., f] + o.gamma[m] * dt * r[..., f]) * np.exp(-v * k2)
        r_prev = r[..., : o.lu]


def test_mid_step_velocity_file():
    """midrk16: vel.000002 of a run with n_skip_v 2 holds step 2's state
    after its second substep plus the zeta term of the third, the last of
    the three writes at nt = 2; vel.000004 (nt_end) is the clean state"""
    p = CASES["midrk16"]
    o = O.Oracle(cfg_of(p))
    u0 = RUNS["midrk16_u0"]
    u2 = o.step(u0, 2)[0]
    assert rel(o.step(u0, 4)[0], RUNS["midrk16_u4"]) < TOL
    mid = mid_step_state(o, u2, 2)
    assert rel(mid, RUNS["midrk16_u2"]) < TOL, rel(mid, RUNS["midrk16_u2"])
    assert rel(u2, RUNS["midrk16_u2"]) > 1e-3


@pytest.mark.parametrize("name", ["16u", "32s"])
def test_oracle_rhs_spectra_div_vs_reference(name):
    """rhs.f (first substep), vel_max.f, spectra.f and div_check.f of the
    reference on seeded inputs; the divergence sums on a field with
    divergence (spectra.f:55-60,88-89, div_check.f:11-36)"""
    g = np.load(os.path.join(GOLD, f"rhs_{name}.npz"))
    n, strat, icfl, seed = (int(v) for v in g["params"])
    o = O.Oracle(O.make_cfg(n, i_strat=strat, grav=16.0, lapse0=1.0, vis=5e-3, i_cfl=icfl, cfl0=0.5, dt0=0.01))
    u = g["u"]
    r, st = o.rhs(u)
    assert rel(r, g["r"]) < TOL
    ref_st = g["stats"]
    np.testing.assert_allclose(st[4:13], ref_st[4:13], rtol=1e-12)        # u_var, u_max, energy
    np.testing.assert_allclose(st[:4], ref_st[:4], atol=1e-15)            # u_avg (zero mean)
    np.testing.assert_allclose(st[13:16] * st[16], ref_st[13:16], rtol=1e-12)   # cfl_x,y,z
    assert abs(st[16] - ref_st[16]) <= 1e-12 * ref_st[16]
    for tag, uu in (("", u), ("w_", g["w"])):
        sp = o.spectra(uu, g["r"])
        for k in ("ek", "tk", "dk", "vt", "sample"):
            np.testing.assert_allclose(sp[k], g[f"{tag}{k}"], rtol=1e-12, atol=1e-300)
        ew, drms, dmax, _ = g[f"{tag}scalars"]
        assert abs(sp["energy_w"] - ew) <= 1e-12 * ew
        assert abs(sp["div_rms"] - drms) <= 1e-12 * max(drms, 1e-300) + 1e-30
        assert abs(sp["div_max"] - dmax) <= 1e-12 * max(dmax, 1e-300) + 1e-30
        # div_check.f's div_rms and div_max are spectra.f's
        dck = g[f"{tag}div_check"]
        assert abs(dck[0] - dmax) <= 1e-14 * max(dmax, 1e-300) + 1e-30
        assert abs(dck[1] - drms) <= 1e-14 * max(drms, 1e-300) + 1e-30
    assert g["w_scalars"][1] > 1e-3 and g["w_div_check"][2] > 0   # the random field has divergence


@pytest.mark.parametrize("name", ["16u", "32s"])
def test_oracle_cl_vs_reference(name):
    """the Chollet-Lesieur model (i_les = 1): cl_energy.f's band energy and
    spectra.f's dk and vt with nu_t(k) = fac_cl (a + b exp(-c k_cl/|k|))
    (sam.f:128-132, spectra.f:21,39-43, init.f:88-93), against the
    reference's own routines on the same state"""
    g = np.load(os.path.join(GOLD, f"rhs_{name}.npz"))
    n, strat, icfl, seed = (int(v) for v in g["params"])
    o = O.Oracle(O.make_cfg(n, i_strat=strat, grav=16.0, lapse0=1.0, vis=5e-3, i_cfl=icfl, cfl0=0.5, dt0=0.01,
                            i_cl=1))
    e_cl, fac = o.cl_energy(g["u"])
    e_ref, k_cl = g["cl_energy"]
    assert k_cl == int(n / 3.0)
    assert abs(e_cl - e_ref) <= 1e-13 * e_ref
    sp = o.spectra(g["u"], g["r"], fac)
    for k in ("ek", "tk", "dk", "vt"):
        np.testing.assert_allclose(sp[k], g[f"cl_{k}"], rtol=1e-12, atol=1e-300)
    assert np.all(g["cl_vt"][1:][g["sample"][1:] > 0] > 0)            # the model is on in every shell
    assert np.abs(g["cl_dk"] - g["dk"]).max() > 1e-3 * np.abs(g["dk"]).max()


def test_oracle_planes_vs_reference():
    """write_planes.f / get_mean.f (with x_deriv.f, y_deriv.f): the planes
    and means files of the reference's own stratified 32^3 run (n_skip_p 2,
    frames at nt = 0, 2, 4; tests/golden/sam_planes.npz) against the
    oracle's planes of the same states, with the pressure of the previous
    substep (zero at nt = 0)"""
    g = np.load(os.path.join(GOLD, "sam_planes.npz"))
    p = json.loads(str(g["params_json"]))
    n, nfr = p["Nx"], 3
    o = O.Oracle(O.make_cfg(n, i_strat=1, grav=p["grav"], lapse0=p["lapse0"], lapse=0.0, vis=p["vis"], i_cfl=0,
                            dt0=p["dt0"]))
    u0 = g["u0"]
    lu = u0.shape[3]
    nf = 2 * lu
    assert list(g["time"][:, 0]) == [0, 2, 4]
    files = {"xy1": ("xy", 0), "xz1": ("xz", 0), "yz1": ("yz", 0), "yz2": ("yz", 1)}
    for fr, nt in enumerate((0, 2, 4)):
        if nt == 0:
            u, pr = u0, np.zeros(u0.shape[:3], np.complex128, order="F")
        else:
            u, _, _, _, pr = o.step(u0, nt, want_p=True)
        mean, xy, xz, yz = o.planes(u, pr, (p["k_xy_plane1"],), (p["j_xz_plane1"],),
                                    (p["i_yz_plane1"], p["i_yz_plane2"]))
        got = {"xy": xy, "xz": xz, "yz": yz}
        for f, (kind, m) in files.items():
            ref = g[f].reshape(nfr, nf, n, n)[fr]            # Fortran records (n, n) per field
            for q in range(nf):
                a = got[kind][:, :, q, m].T
                assert np.abs(a - ref[q]).max() <= 1e-12 * np.abs(ref[q]).max(), (nt, f, q)
        ref_mean = g["mean"].reshape(nfr, lu, n)[fr]
        assert np.abs(mean.T - ref_mean).max() <= 1e-12 * np.abs(ref_mean).max(), nt
        assert np.abs(ref_mean).max() > 0.1                   # non-trivial horizontal means


def test_divergence_brute_force():
    """spectra.f's divergence sums by brute force on a 16^3 field:
    div_rms = sqrt(sum over stored modes of |k.u|^2) (no kx factor 2),
    div_max = sqrt(max |k.u|^2)"""
    g = np.load(os.path.join(GOLD, "rhs_16u.npz"))
    o = O.Oracle(O.make_cfg(16))
    w = g["w"]
    kx, ky, kz = o.kvectors()
    wx, wy, wz = o.wave
    d2 = []
    for i in range(o.nxm):
        for j in range(o.nym):
            for k in range(o.nzm):
                dv = wx * kx[i] * w[k, j, i, 0] + wy * ky[j] * w[k, j, i, 1] + wz * kz[k] * w[k, j, i, 2]
                d2.append(abs(dv) ** 2)
    d2 = np.array(d2)
    sp = o.spectra(w, g["r"])
    assert abs(sp["div_rms"] - np.sqrt(d2.sum())) <= 1e-13 * np.sqrt(d2.sum())
    assert abs(sp["div_max"] - np.sqrt(d2.max())) <= 1e-13 * np.sqrt(d2.max())


@pytest.mark.parametrize("name", ["16u", "32s"])
def test_oracle_div_check_vs_reference(name):
    g = np.load(os.path.join(GOLD, f"rhs_{name}.npz"))
    n, strat = (int(v) for v in g["params"][:2])
    o = O.Oracle(O.make_cfg(n, i_strat=strat))
    for tag, uu in (("", g["u"]), ("w_", g["w"])):
        ref = g[f"{tag}div_check"]
        got = o.div_check(uu)
        assert abs(got["div_max"] - ref[0]) <= 1e-13 * ref[0] + 1e-30
        assert abs(got["div_rms"] - ref[1]) <= 1e-13 * ref[1] + 1e-30
        assert got["n_div"] == ref[2]


@pytest.mark.parametrize("name", ["strat32", "nrk1_16", "nrk2_16"])
def test_rk_update_composes_the_reference_step(name):
    """one step composed of the oracle's RHS and orc_rk_update (the zeta
    update before each substep's RHS, the gamma/integrating-factor update
    after it, sam.f:169-314) equals the reference's vel.000001 / one step"""
    o, u0, _ = run_case(name)
    u = np.array(u0, order="F", copy=True)
    r_prev, dt = None, None
    for m in range(o.cfg.nrk_max):
        r, st = o.rhs(u)                    # the RHS of the state before the zeta update
        if m == 0:
            dt = st[16] if o.cfg.i_cfl == 1 else o.cfg.dt0
        u = o.rk_update(m + 1, dt, u, r_prev=r_prev, r=r[..., : o.lu])
        r_prev = r[..., : o.lu]
    ref = o.step(u0, 1)[0]
    assert rel(u, ref) < 1e-13, rel(u, ref)
    if name == "strat32":
        assert rel(u, RUNS["strat32_u1"]) < TOL
