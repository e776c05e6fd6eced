"""GPU parity: libsamhip (through its C ABI) against the CPU oracle on the
same inputs.  FP64 bar: relative L2 <= 1e-12 (the drop-in's tolerance),
per stage and after N full RK3 steps.  FP32-split bar: the
error is reported and bounded by FP32 rounding (<= 2e-6 relative per stage).
"""
import numpy as np
import pytest

from sam_amd import ic
from sam_amd import samhip as S

pytestmark = pytest.mark.gpu

TOL64 = 1e-12


def rel_l2(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def pair(oracle_mod, n, ny=None, nz=None, prec=S.PREC_FP64, **kw):
    O = oracle_mod
    cfg = O.make_cfg(n, ny, nz, **kw)
    o = O.Oracle(cfg)
    p = S.make_params(n, ny, nz, xl=cfg.xl, yl=cfg.yl, zl=cfg.zl, n_dealias=cfg.n_dealias,
                      i_strat=cfg.i_strat, nrk_max=cfg.nrk_max, i_cfl=cfg.i_cfl, precision=prec,
                      vis=cfg.vis, pr=cfg.pr, grav=cfg.grav, to=cfg.to, lapse=cfg.lapse,
                      cfl0=cfg.cfl0, dt0=cfg.dt0)
    return o, S.Sam(p)


def rand_spec(o, nvar, seed=0, ldz=None):
    rng = np.random.default_rng(seed)
    sh = (ldz or o.nzm, o.nym, o.nxm, nvar)
    u = rng.standard_normal(sh) + 1j * rng.standard_normal(sh)
    # kx = 0 plane: keep it Hermitian-consistent enough for transforms (not required)
    return np.asfortranarray(u)


GRIDS = [
    dict(n=16),
    dict(n=32, i_strat=1),
    dict(n=16, ny=32, nz=64, xl=3.0, yl=5.0, zl=7.0),
    dict(n=32, n_dealias=0),
    dict(n=32, n_dealias=2, i_strat=1),
    dict(n=64),
]


@pytest.mark.parametrize("g", GRIDS, ids=lambda g: "-".join(f"{k}{v}" for k, v in g.items()))
def test_stage_z_trans_b(oracle_mod, g):
    g = dict(g)
    o, s = pair(oracle_mod, g.pop("n"), **g)
    u = rand_spec(o, o.lu, seed=1)
    ref = o.z_trans_b(u)                              # (Ny_min, Nx_min, Nz, n)
    got = s.z_trans_b(u)                              # the same layout (z_trans_b.f:31-52)
    assert rel_l2(got, ref) < TOL64
    assert rel_l2(s.x2z(ref), o.x2z(ref)) == 0.0     # x2z_decomp.f: pure index map


@pytest.mark.parametrize("g", GRIDS, ids=lambda g: "-".join(f"{k}{v}" for k, v in g.items()))
def test_stage_xy_trans_b(oracle_mod, g):
    g = dict(g)
    o, s = pair(oracle_mod, g.pop("n"), **g)
    zb = rand_spec(o, o.lu, seed=2, ldz=o.nz)
    y = o.x2z(np.transpose(zb, (1, 2, 0, 3)).copy(order="F"))   # (Ny, Nx/2, Nz, n) as x2z_decomp writes it
    ref = o.xy_trans_b(y)
    got = s.xy_trans_b(y)
    assert rel_l2(got, ref) < TOL64


@pytest.mark.parametrize("g", GRIDS, ids=lambda g: "-".join(f"{k}{v}" for k, v in g.items()))
def test_stage_xy_trans_f(oracle_mod, g):
    g = dict(g)
    o, s = pair(oracle_mod, g.pop("n"), **g)
    rng = np.random.default_rng(3)
    p = np.asfortranarray(rng.standard_normal((o.nx, o.ny, o.nz, o.lr)))
    ref = o.xy_trans_f(p)
    got = s.xy_trans_f(p)
    assert rel_l2(got, ref) < TOL64


@pytest.mark.parametrize("g", GRIDS, ids=lambda g: "-".join(f"{k}{v}" for k, v in g.items()))
def test_stage_z_trans_f(oracle_mod, g):
    g = dict(g)
    o, s = pair(oracle_mod, g.pop("n"), **g)
    a = rand_spec(o, o.lr, seed=4, ldz=o.nz)
    ref = o.z_trans_f(a)
    got = s.z_trans_f(a)
    assert rel_l2(got, ref) < TOL64


def test_stage_products_velmax_project(oracle_mod):
    o, s = pair(oracle_mod, 32, i_strat=1, grav=16.0, lapse0=1.0, i_cfl=1, cfl0=0.5)
    rng = np.random.default_rng(5)
    phys = np.asfortranarray(rng.standard_normal((o.nx, o.ny, o.nz, o.lu)))
    assert rel_l2(s.products(phys), o.products(phys)) == 0.0
    st_ref = o.vel_max(phys, 0.01)
    st = s.vel_max(phys, 0.01)
    np.testing.assert_allclose(st, st_ref, rtol=1e-12, atol=1e-15)
    F = rand_spec(o, o.lr, seed=6)
    us = rand_spec(o, o.lu, seed=7)
    assert rel_l2(s.project(F, us), o.project(F, us)) < TOL64


@pytest.mark.parametrize("shape", [(64, 64, 32), (256, 256, 256)], ids=["64x64x32", "256c"])
def test_vel_max_statistics_chunking(oracle_mod, shape):
    """vel_max.f's statistics at sizes where the per-block point chunk and
    the finalize differ (64x64x32: 2048-point chunks, one finalize level;
    256^3: 4096-point chunks, 4096 partial rows, two levels)"""
    o, s = pair(oracle_mod, shape[0], shape[1], shape[2], i_strat=1, grav=16.0, lapse0=1.0, i_cfl=1, cfl0=0.5)
    rng = np.random.default_rng(11)
    phys = np.asfortranarray(rng.standard_normal((o.nx, o.ny, o.nz, o.lu)) * np.array([1.0, 2.0, 0.5, 3.0]))
    st_ref = o.vel_max(phys, 0.01)
    st = s.vel_max(phys, 0.01)
    s.close()
    np.testing.assert_allclose(st, st_ref, rtol=1e-12, atol=1e-15)


@pytest.mark.parametrize("strat", [0, 1])
def test_rhs_full(oracle_mod, strat):
    o, s = pair(oracle_mod, 32, i_strat=strat, grav=16.0, lapse0=1.0, vis=5e-3)
    u = ic.random_solenoidal(32, k0=4, seed=11, i_strat=strat)
    if strat:
        u[:, :, :, 3] = 0.1 * ic.random_solenoidal(32, k0=3, seed=12)[:, :, :, 0]
    s.set_state(u)
    r_ref, st_ref = o.rhs(u)
    r, d = s.rhs(want_spectra=True)
    assert rel_l2(r[..., : o.lu], r_ref[..., : o.lu]) < TOL64
    assert rel_l2(r[..., o.lu], r_ref[..., o.lu]) < TOL64
    np.testing.assert_allclose(d["u_max"][: o.lu], st_ref[8:8 + o.lu], rtol=1e-12)
    np.testing.assert_allclose(d["energy"], st_ref[12], rtol=1e-12)
    sp = s.spectra()
    sp_ref = o.spectra(u, r_ref)
    for key in ("ek", "tk", "dk", "sample"):
        np.testing.assert_allclose(sp[key], sp_ref[key], rtol=1e-11, atol=1e-14)
    np.testing.assert_allclose(sp["energy_w"], sp_ref["energy_w"], rtol=1e-12)
    # divergence sums (spectra.f:55-60,88-89): round-off level for a solenoidal
    # state, so compared on an absolute scale; a state with divergence below
    assert abs(sp["div_rms"] - sp_ref["div_rms"]) < 1e-13 and abs(sp["div_max"] - sp_ref["div_max"]) < 1e-13
    rng = np.random.default_rng(13)
    w = np.asfortranarray(u + 1e-2 * (rng.standard_normal(u.shape) + 1j * rng.standard_normal(u.shape)))
    w[:, :, 0] = u[:, :, 0]
    s.set_state(w)
    r_w, _ = s.rhs(want_spectra=True)
    sp = s.spectra()
    sp_ref = o.spectra(w, r_w)
    assert sp_ref["div_max"] > 1e-3
    np.testing.assert_allclose([sp["div_rms"], sp["div_max"]], [sp_ref["div_rms"], sp_ref["div_max"]], rtol=1e-12)
    dv, dv_ref = s.div_check(), o.div_check(w)
    np.testing.assert_allclose([dv["div_max"], dv["div_rms"]], [dv_ref["div_max"], dv_ref["div_rms"]], rtol=1e-12)
    assert dv["n_div"] == dv_ref["n_div"]


def test_taylor_green_64_steps(oracle_mod):
    """The CPU-runnable workload in full: 64^3 Taylor-Green, nu = 1/400,
    dt = 0.01, 200 steps (t = 2), FP64 parity 1e-12"""
    nsteps = 200
    o, s = pair(oracle_mod, 64, vis=0.0025, dt0=0.01)
    u0 = ic.taylor_green(64)
    s.set_state(u0)
    d = s.step(nsteps)
    got = s.get_state()
    ref, t, dt, st = o.step(u0, nsteps)
    assert rel_l2(got, ref) < TOL64
    assert abs(d["time"] - t) < 1e-12
    np.testing.assert_allclose(d["energy"], st[12], rtol=1e-12)


@pytest.mark.parametrize("g", GRIDS, ids=lambda g: "-".join(f"{k}{v}" for k, v in g.items()))
def test_steps_every_grid(oracle_mod, g):
    """3 full steps on every stage-test grid (non-cubic, unequal lengths,
    n_dealias 0 and 2, Lu = 3 and 4): the step's fused z inverse, whose
    zero band depends on kz_max, against the oracle"""
    g = dict(g)
    n, ny, nz = g.pop("n"), g.pop("ny", None), g.pop("nz", None)
    o, s = pair(oracle_mod, n, ny, nz, vis=5e-3, dt0=0.005, **(dict(grav=16.0, lapse0=1.0) if g.get("i_strat") else {}),
                **g)
    u0 = rand_spec(o, o.lu, seed=11) * 1e-2
    ref = o.step(u0, 3)[0]
    s.set_state(u0)
    s.step(3)
    assert rel_l2(s.get_state(), ref) < TOL64


def test_stratified_cfl_steps(oracle_mod):
    """stratified Boussinesq at 32^3: N = 4 (N^2 = grav*lapse/To = 16), CFL dt"""
    nsteps = 10
    o, s = pair(oracle_mod, 32, i_strat=1, grav=16.0, lapse0=1.0, vis=5e-3, i_cfl=1, cfl0=0.5)
    u0 = ic.random_solenoidal(32, k0=4, seed=20261203, i_strat=1)
    s.set_state(u0)
    d = s.step(nsteps)
    got = s.get_state()
    ref, t, dt, st = o.step(u0, nsteps)
    assert rel_l2(got, ref) < TOL64
    assert abs(d["dt"] - dt) < 1e-12 * dt
    assert abs(d["time"] - t) < 1e-12


@pytest.mark.parametrize("strat", [0, 1])
def test_cfl_short_x_rows(oracle_mod, strat):
    """CFL dt on 16^3 (FP64 x-row blocks of 8 threads, fewer than the 3 Lu
    entries of a statistics row): u_max of every field, dt and 3 steps"""
    o, s = pair(oracle_mod, 16, i_strat=strat, grav=16.0, lapse0=1.0, vis=5e-3, i_cfl=1, cfl0=0.5)
    u0 = ic.random_solenoidal(16, k0=3, seed=17, i_strat=strat)
    s.set_state(u0)
    d = s.step(3)
    ref, t, dt, st = o.step(u0, 3)
    assert rel_l2(s.get_state(), ref) < TOL64
    assert abs(d["dt"] - dt) <= 1e-12 * dt
    np.testing.assert_allclose(d["u_max"][: o.lu], st[8:8 + o.lu], rtol=1e-12)


@pytest.mark.parametrize("prec,n,tol", [(S.PREC_FP64, 128, TOL64), (S.PREC_SPLIT, 64, 1e-5)],
                         ids=["fp64-128", "split-64"])
def test_double_staged_x_rows(oracle_mod, prec, n, tol):
    """grids whose x rows take the double-staged inverse phase (dealiased,
    whole-wavefront columns: FP64 from 128, FP32 from 64): full stratified
    steps against the oracle"""
    nsteps = 2
    o, s = pair(oracle_mod, n, i_strat=1, grav=16.0, lapse0=1.0, vis=5e-3, i_cfl=1, cfl0=0.5, prec=prec)
    u0 = ic.random_solenoidal(n, k0=4, seed=7, i_strat=1)
    u0[..., 3] = 0.1 * ic.random_solenoidal(n, k0=3, seed=8)[..., 0]
    s.set_state(u0)
    d = s.step(nsteps)
    ref, t, dt, st = o.step(u0, nsteps)
    err = rel_l2(s.get_state(), ref)
    assert err < tol, err
    np.testing.assert_allclose(d["energy"], st[12], rtol=10 * tol)


@pytest.mark.parametrize("prec,tol", [(S.PREC_FP64, TOL64), (S.PREC_SPLIT, 1e-5)], ids=["fp64", "split"])
def test_long_x_rows(oracle_mod, prec, tol):
    """nx = 1024 (FP32: each real row as one 512-point complex transform,
    8 rows per block, k_xrow_h) on a thin 1024 x 32 x 32 box: full
    stratified steps against the oracle"""
    nsteps = 2
    o, s = pair(oracle_mod, 1024, 32, 32, xl=64.0, yl=2.0, zl=2.0, i_strat=1, grav=16.0, lapse0=1.0, vis=5e-3,
                i_cfl=1, cfl0=0.5, prec=prec)
    u0 = ic.random_solenoidal(1024, 32, 32, k0=4, seed=9, i_strat=1, xl=64.0, yl=2.0, zl=2.0)
    s.set_state(u0)
    s.step(nsteps)
    ref, t, dt, st = o.step(u0, nsteps)
    err = rel_l2(s.get_state(), ref)
    assert err < tol, err


@pytest.mark.parametrize("prec,tol", [(S.PREC_FP64, TOL64), (S.PREC_SPLIT, 1e-5)], ids=["fp64", "split"])
@pytest.mark.parametrize("shape", [(32, 1024, 32), (32, 32, 1024)], ids=["y1024", "z1024"])
def test_long_y_z_columns_1024(oracle_mod, shape, prec, tol):
    """the 1024-point y and z transforms (16 elements per thread; k_zrhs
    without the fused z inverse) on thin stratified boxes: 2 full steps
    against the oracle, FP64 at 1e-12 and split at 1e-5"""
    nx, ny, nz = shape
    box = dict(xl=nx / 16.0, yl=ny / 16.0, zl=nz / 16.0)
    o, s = pair(oracle_mod, nx, ny, nz, i_strat=1, grav=16.0, lapse0=1.0, vis=5e-3, i_cfl=1, cfl0=0.5, prec=prec,
                **box)
    u0 = ic.random_solenoidal(nx, ny, nz, k0=4, seed=12, i_strat=1, **box)
    u0[..., 3] = 0.1 * ic.random_solenoidal(nx, ny, nz, k0=3, seed=13, **box)[..., 0]
    s.set_state(u0)
    s.step(2)
    ref = o.step(u0, 2)[0]
    err = rel_l2(s.get_state(), ref)
    assert err < tol, err


@pytest.mark.parametrize("ny", [256, 512])
def test_long_y_columns(oracle_mod, ny):
    """long y columns (Ny = 256, 512; 8 z per y-transform tile): full
    stratified FP64 steps on a 32 x Ny x 32 box against the oracle"""
    nsteps = 2
    o, s = pair(oracle_mod, 32, ny, 32, xl=2.0, yl=ny / 16.0, zl=2.0, i_strat=1, grav=16.0, lapse0=1.0, vis=5e-3,
                i_cfl=1, cfl0=0.5, prec=S.PREC_FP64)
    u0 = ic.random_solenoidal(32, ny, 32, k0=4, seed=10, i_strat=1, xl=2.0, yl=ny / 16.0, zl=2.0)
    s.set_state(u0)
    s.step(nsteps)
    ref, t, dt, st = o.step(u0, nsteps)
    err = rel_l2(s.get_state(), ref)
    assert err < 1e-12, err


def test_split_precision_bounded(oracle_mod):
    """FP32-split: transforms/products in FP32, state/RK in FP64"""
    o, s = pair(oracle_mod, 32, i_strat=1, grav=16.0, lapse0=1.0, vis=5e-3, i_cfl=1, cfl0=0.5,
                prec=S.PREC_SPLIT)
    u0 = ic.random_solenoidal(32, k0=4, seed=3, i_strat=1)
    s.set_state(u0)
    s.step(5)
    got = s.get_state()
    ref, t, dt, st = o.step(u0, 5)
    err = rel_l2(got, ref)
    assert err < 1e-5, err


def test_stratified_512_step(oracle_mod):
    """The headline configuration itself: one step of the 512^3 stratified DNS
    (the bench's workload and seeded IC; persistent z kernel, double-staged
    x rows, 512-point y and z transforms) against the oracle.  FP64 1e-12,
    FP32-split 1e-5."""
    kw = dict(i_strat=1, grav=16.0, lapse0=1.0, vis=5e-4, pr=1.0, i_cfl=1, cfl0=0.5)
    u0 = ic.random_solenoidal(512, k0=4, seed=20261203, i_strat=1)
    o, s64 = pair(oracle_mod, 512, **kw)
    ref = o.step(u0, 1)[0]
    s64.set_state(u0)
    s64.step(1)
    assert rel_l2(s64.get_state(), ref) < TOL64
    s64.close()
    _, s32 = pair(oracle_mod, 512, prec=S.PREC_SPLIT, **kw)
    s32.set_state(u0)
    s32.step(1)
    assert rel_l2(s32.get_state(), ref) < 1e-5
    s32.close()


def test_stratified_256_500_steps(oracle_mod):
    """The stratified 256^3 workload in full (seeded IC, N = 4, nu = kappa =
    5e-4, Pr = 1, CFL 0.5, 500 steps): FP64 parity over the first 2 steps
    (1e-12), then 500 steps in FP64 and in the split mode.  The split run's
    kinetic energy is within 1e-5 of FP64, and its dissipation and E(k)
    (relative L2 over the shells) within 1e-4.  Both stay divergence-free."""
    kw = dict(i_strat=1, grav=16.0, lapse0=1.0, vis=5e-4, pr=1.0, i_cfl=1, cfl0=0.5)
    u0 = ic.random_solenoidal(256, k0=4, seed=20261203, i_strat=1)
    o, s64 = pair(oracle_mod, 256, **kw)
    s64.set_state(u0)
    s64.step(2)
    ref = o.step(u0, 2)[0]
    assert rel_l2(s64.get_state(), ref) < TOL64
    s64.step(498, want_spectra=True)
    sp64 = s64.spectra()
    _, s32 = pair(oracle_mod, 256, prec=S.PREC_SPLIT, **kw)
    s32.set_state(u0)
    s32.step(500, want_spectra=True)
    sp32 = s32.spectra()
    ke = abs(sp32["energy_w"] - sp64["energy_w"]) / sp64["energy_w"]
    diss = abs(sp32["dk"].sum() - sp64["dk"].sum()) / abs(sp64["dk"].sum())
    ek = rel_l2(sp32["ek"], sp64["ek"])
    assert ke < 1e-5, ke
    assert diss < 1e-4, diss
    assert ek < 1e-4, ek
    for sp in (sp64, sp32):
        assert sp["div_max"] < 1e-10, sp["div_max"]


# ---------------------------------------------------------------------------
# LES: Smagorinsky (sm.f) and dynamic (dm.f) closures on the collocated grid
# ---------------------------------------------------------------------------
def les_pair(oracle_mod, n, sgs, prec=S.PREC_FP64, **kw):
    O = oracle_mod
    cfg = O.make_cfg(n, sgs=sgs, **kw)
    o = O.Oracle(cfg)
    p = S.make_params(n, xl=cfg.xl, n_dealias=cfg.n_dealias, i_strat=cfg.i_strat, i_cfl=cfg.i_cfl,
                      precision=prec, sgs=sgs, vis=cfg.vis, pr=cfg.pr, grav=cfg.grav, to=cfg.to,
                      lapse=cfg.lapse, cfl0=cfg.cfl0, dt0=cfg.dt0, c_smag=cfg.c_smag, pr_t=cfg.pr_t)
    return o, S.Sam(p)


@pytest.mark.parametrize("sgs", [S.SGS_SMAG, S.SGS_DYN])
def test_eddy_viscosity(oracle_mod, sgs):
    o, s = les_pair(oracle_mod, 32, sgs, i_strat=1, c_smag=0.03)
    u = ic.random_solenoidal(32, k0=6, seed=31, i_strat=1)
    phys = o.to_physical(u)
    vt_ref, cd_ref = o.eddy_visc(phys)
    vt, cd = s.eddy_visc(phys)
    assert abs(cd - cd_ref) <= 1e-11 * abs(cd_ref)
    assert rel_l2(vt, vt_ref) < 1e-11


def test_test_filter(oracle_mod):
    o, s = les_pair(oracle_mod, 32, S.SGS_DYN)
    rng = np.random.default_rng(2)
    f = np.asfortranarray(rng.standard_normal((32, 32, 32)))
    assert rel_l2(s.test_filter(f), o.test_filter(f)) < 1e-14


@pytest.mark.parametrize("sgs", [S.SGS_SMAG, S.SGS_DYN])
def test_les_steps(oracle_mod, sgs):
    o, s = les_pair(oracle_mod, 32, sgs, i_strat=1, grav=64.0, lapse0=1.0, vis=1e-5, i_cfl=1,
                    cfl0=0.5, c_smag=0.03)
    u0 = ic.random_solenoidal(32, k0=6, seed=41, i_strat=1)
    s.set_state(u0)
    d = s.step(5)
    ref, t, dt, st = o.step(u0, 5)
    assert rel_l2(s.get_state(), ref) < 1e-11
    if sgs == S.SGS_DYN:
        assert abs(d["c_delta_sq"] - st[17]) <= 1e-10 * abs(st[17])
