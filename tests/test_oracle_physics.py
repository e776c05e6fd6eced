"""Pin the oracle's RHS, projection, buoyancy, integrating factor and RK3
against closed forms and invariants -- CPU only.

- plane internal gravity wave (gw.tex:111-149): u.grad u = 0, so the discrete
  solution is the RK3 recurrence of a 2x2 oscillator with omega^2 =
  N^2 kh^2/k^2; this fixes the buoyancy sign and size, the lapse term, the
  projection and the RK coefficients (init.f:135-145);
- shear/temperature waves: nonlinear term zero, exact decay exp(-nu k^2 t)
  (and nu/Pr for T') from the integrating factor (sam.f:293-314);
- the 2/3-dealiased nonlinear term conserves energy (sum of tk = 0);
- the projected RHS is divergence free and differs from the unprojected
  flux divergence by a gradient;
- RK3 is third-order accurate;
- vel_max's CFL time step (vel_max.f:36-46) and spectra's shell sums
  (spectra.f:22-67) against direct evaluation.
"""
import numpy as np
import pytest

from oracle import oracle as O
from sam_amd import ic


def gw_state(o, amp_w=0.3, amp_t=0.2):
    """single mode k = (1, 0, 1): u = -w (k.u = 0), T'"""
    u = o.zeros_spec()
    iz = 1
    u[iz, 0, 1, 2] = amp_w * (1 - 0.5j)
    u[iz, 0, 1, 0] = -u[iz, 0, 1, 2]
    u[iz, 0, 1, 3] = amp_t * (0.3 + 1j)
    return u


def rk3_oscillator(q0, A, dt, nsteps, gam, zet):
    q = q0.copy()
    for _ in range(nsteps):
        rprev = None
        for m in range(3):
            r = A @ q
            if m > 0:
                q = q + zet[m] * dt * rprev
            q = q + gam[m] * dt * r
            rprev = r
    return q


def test_gravity_wave_rk3_recurrence():
    grav, to, lapse = 16.0, 1.0, 1.0
    o = O.Oracle(O.make_cfg(16, i_strat=1, grav=grav, to=to, lapse0=lapse, vis=0.0, dt0=0.01))
    u0 = gw_state(o)
    nsteps = 25
    u1, t, dt, _ = o.step(u0, nsteps)
    bz = grav / to
    A = np.array([[0.0, bz * 0.5], [-lapse, 0.0]])   # d(w,T)/dt; projection keeps 1 - kz^2/k^2 = 1/2
    q0 = np.array([u0[1, 0, 1, 2], u0[1, 0, 1, 3]])
    q = rk3_oscillator(q0, A, 0.01, nsteps, o.gamma, o.zeta)
    assert abs(u1[1, 0, 1, 2] - q[0]) < 1e-13
    assert abs(u1[1, 0, 1, 3] - q[1]) < 1e-13
    assert abs(u1[1, 0, 1, 0] + u1[1, 0, 1, 2]) < 1e-13          # stays transverse
    # everything else stays zero
    mask = np.ones(u1.shape, bool)
    mask[1, 0, 1, :] = False
    assert np.abs(u1[mask]).max() < 1e-14
    # and the recurrence approximates the exact oscillation to O(dt^3)
    om = np.sqrt(bz * lapse * 0.5)
    w_exact = q0[0] * np.cos(om * t) + q0[1] * bz * 0.5 / om * np.sin(om * t)
    assert abs(u1[1, 0, 1, 2] - w_exact) < 1e-5


def test_rk3_third_order():
    o = O.Oracle(O.make_cfg(16, i_strat=1, grav=16.0, lapse0=1.0, vis=0.0, dt0=0.04))
    u0 = gw_state(o)
    om = np.sqrt(16.0 * 0.5)
    q0 = np.array([u0[1, 0, 1, 2], u0[1, 0, 1, 3]])
    errs = []
    for k in range(3):
        dt = 0.04 / 2 ** k
        o.cfg.dt0 = dt
        oo = O.Oracle(O.make_cfg(16, i_strat=1, grav=16.0, lapse0=1.0, vis=0.0, dt0=dt))
        u1, t, _, _ = oo.step(u0, 10 * 2 ** k)
        w_exact = q0[0] * np.cos(om * t) + q0[1] * 16.0 * 0.5 / om * np.sin(om * t)
        errs.append(abs(u1[1, 0, 1, 2] - w_exact))
    r1, r2 = errs[0] / errs[1], errs[1] / errs[2]
    assert 7.0 < r1 < 9.0 and 7.0 < r2 < 9.0, errs


def test_viscous_decay_integrating_factor():
    vis, pr = 0.01, 0.7
    o = O.Oracle(O.make_cfg(16, i_strat=1, grav=16.0, lapse0=1.0, vis=vis, pr=pr, dt0=0.05))
    u0 = o.zeros_spec()
    u0[0, 2, 0, 0] = 0.4 - 0.1j            # u = f(y), ky = 2
    u0[0, 2, 0, 0] = u0[0, 2, 0, 0]
    u0[0, o.nym - 2, 0, 0] = np.conj(u0[0, 2, 0, 0])
    u0[3, 0, 0, 3] = 0.2 + 0.05j            # T = g(z), kz = 3
    u0[o.nzm - 3, 0, 0, 3] = np.conj(u0[3, 0, 0, 3])
    u1, t, _, _ = o.step(u0, 8)
    assert abs(u1[0, 2, 0, 0] - u0[0, 2, 0, 0] * np.exp(-vis * 4 * t)) < 1e-15
    assert abs(u1[3, 0, 0, 3] - u0[3, 0, 0, 3] * np.exp(-vis / pr * 9 * t)) < 1e-15


def test_energy_conserving_nonlinear_term():
    o = O.Oracle(O.make_cfg(32, vis=0.0))
    u = ic.random_solenoidal(32, k0=4, seed=2)
    r, _ = o.rhs(u)
    sp = o.spectra(u, r)
    assert abs(np.sum(sp["tk"])) < 1e-13 * np.sum(np.abs(sp["tk"]))


def test_projection_divergence_free_and_gradient():
    o = O.Oracle(O.make_cfg(16, 16, 16, xl=3.0, yl=4.0, zl=5.0, i_strat=1, grav=9.0, lapse0=2.0))
    rng = np.random.default_rng(4)
    F = np.asfortranarray(rng.standard_normal(o.spec_shape(o.lr)) + 1j * rng.standard_normal(o.spec_shape(o.lr)))
    us = np.asfortranarray(rng.standard_normal(o.spec_shape()) + 1j * rng.standard_normal(o.spec_shape()))
    r = o.project(F, us)
    kx, ky, kz = o.kvectors()
    KZ, KY, KX = np.meshgrid(kz * o.wave[2], ky * o.wave[1], kx * o.wave[0], indexing="ij")
    div = KX * r[..., 0] + KY * r[..., 1] + KZ * r[..., 2]
    assert np.abs(div).max() < 1e-12 * np.abs(r[..., :3]).max() * 10
    # unprojected momentum source: -i k_j F_ij + buoyancy (F order uu,uv,vv,uw,vw,ww)
    Fm = [[0, 1, 3], [1, 2, 4], [3, 4, 5]]
    a = [-1j * (KX * F[..., Fm[i][0]] + KY * F[..., Fm[i][1]] + KZ * F[..., Fm[i][2]]) for i in range(3)]
    a[2] = a[2] + o.buoy_fac_z * us[..., 3]
    # r - a = -i k p: parallel to k with the returned pressure
    for i, K in enumerate((KX, KY, KZ)):
        assert np.abs((r[..., i] - a[i]) - (-1j * K * r[..., 4])).max() < 1e-10
    # T equation: -i k.F_T - lapse * w_sav
    aT = -1j * (KX * F[..., 6] + KY * F[..., 7] + KZ * F[..., 8]) - 2.0 * us[..., 2]
    assert np.abs(r[..., 3] - aT).max() < 1e-12


def test_cfl_dt_closed_form():
    o = O.Oracle(O.make_cfg(16, i_cfl=1, cfl0=0.7))
    rng = np.random.default_rng(1)
    phys = np.asfortranarray(rng.standard_normal((16, 16, 16, 3)))
    st = o.vel_max(phys, 0.01)
    umax = np.abs(phys).max(axis=(0, 1, 2))
    cfl = umax * np.array([o.kxm, o.kym, o.kzm]) * 1.0
    assert abs(st[16] - 0.7 * np.sqrt(3.0) / cfl.max()) < 1e-14
    assert abs(st[12] - 0.5 * sum(np.var(phys[..., i]) for i in range(3))) < 1e-13


def test_spectra_brute_force():
    o = O.Oracle(O.make_cfg(16, vis=0.01))
    u = ic.random_solenoidal(16, k0=3, seed=8)
    r, _ = o.rhs(u)
    sp = o.spectra(u, r)
    kx, ky, kz = o.kvectors()
    ek = np.zeros(o.lm)
    for i, a in enumerate(kx):
        fac = 1.0 if a == 0 else 2.0
        for j, b in enumerate(ky):
            for k, c in enumerate(kz):
                s = int(np.sqrt(a * a + b * b + c * c) + 0.5)
                ek[s] += fac * np.sum(np.abs(u[k, j, i, :3]) ** 2)
    assert np.allclose(sp["ek"], ek, rtol=1e-13, atol=1e-16)
    # Parseval: 0.5 sum ek = energy of the physical field
    phys = o.to_physical(u)
    assert abs(sp["energy_w"] - 0.5 * np.mean(np.sum(phys[..., :3] ** 2, axis=-1))) < 1e-13
