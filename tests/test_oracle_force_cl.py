"""Pin the oracle's spectral forcing (force.f) and Chollet-Lesieur model
(cl_energy.f, spectra.f:39-43) -- CPU only.

force.f itself is pinned through the reference's own program: the force16
run of tests/golden/sam_runs.npz (the reference's `sam` built with MPICH by
oracle/Makefile, run with force.f after every step) is matched by the oracle
in test_oracle_refrun.py, and so are cl_energy.f and spectra.f with the
Chollet-Lesieur model (the reference's own routines, tests/golden/rhs_*.npz).
Both are also pinned here by independent evaluations:
- forcing: after force.f the shell energies 2..k_force+1 of u, v, w equal
  the target exactly, every other shell, the mean and T' are untouched, and
  the divergence is unchanged (all components of a mode share one factor);
- cl_energy: the reference walks index ranges k_start..k_stop and mirror
  rows; the test sums the band (k_cl - 1/2)^2 < |k|^2 <= (k_cl + 1/2)^2
  directly over all stored modes;
- spectra with the model: dk and vt by direct summation of the formula;
- the model in the integrating factor (an extension): two shear waves
  u = A1 sin y + A2 sin(k_cl y) have no nonlinear term, so each step
  multiplies each amplitude by exp(-(nu + nu_t(k)) k^2 dt), with nu_t from
  the band energy A2^2/2 at the start of the step.
"""
import numpy as np

from oracle import oracle as O
from sam_amd import ic


def shells(o, u):
    """1-based shell index is = 1 + int(|k| + 1/2) and kx factor per stored mode"""
    kx, ky, kz = o.kvectors()
    ksq = kz[:, None, None] ** 2 + ky[None, :, None] ** 2 + kx[None, None, :] ** 2
    is_ = 1 + np.floor(np.sqrt(ksq) + 0.5).astype(int)
    fac = np.where(kx[None, None, :] == 0, 1.0, 2.0) * np.ones_like(ksq, dtype=float)
    return ksq, is_, fac


def shell_energy(o, u, nshell):
    _, is_, fac = shells(o, u)
    usq = np.sum(np.abs(u[..., :3]) ** 2, axis=3)
    return np.array([np.sum((fac * usq)[is_ == s]) for s in range(1, nshell + 1)])


def random_state(o, seed):
    return ic.random_solenoidal(o.nx, k0=3.0, seed=seed, i_strat=1)


def test_force_restores_the_target_spectrum():
    kf = 3
    o = O.Oracle(O.make_cfg(32, i_strat=1, k_force=kf))
    u = random_state(o, 5)
    # the generator fixes the shell spectrum: a target off by shell-dependent factors
    target = shell_energy(o, u, o.lm) * (1.0 + 0.37 * np.arange(o.lm) ** 0.5)
    out, ekg = o.force(u, target)
    assert np.allclose(ekg, shell_energy(o, u, kf + 1), rtol=1e-13)
    got = shell_energy(o, out, o.lm)
    assert np.allclose(got[1:kf + 1], target[1:kf + 1], rtol=1e-13)
    assert np.array_equal(got[0], shell_energy(o, u, 1)[0])               # the mean is not forced
    _, is_, _ = shells(o, u)
    outside = is_ > kf + 1
    assert np.array_equal(out[outside], u[outside])
    assert np.array_equal(out[..., 3], u[..., 3])                          # T' untouched
    nz = np.abs(u[..., :3]) > 0
    ratio = np.abs(out[..., :3]) / np.where(nz, np.abs(u[..., :3]), 1)
    before = shell_energy(o, u, kf + 1)
    for s in range(2, kf + 2):
        sel = (is_ == s)[..., None] & nz
        assert sel.sum() > 10
        assert np.allclose(ratio[sel], np.sqrt(target[s - 1] / before[s - 1]), rtol=1e-12)
        assert abs(np.sqrt(target[s - 1] / before[s - 1]) - 1) > 0.1


def test_cl_energy_is_the_band_energy():
    for n in (32, 64):
        o = O.Oracle(O.make_cfg(n, i_strat=1, i_cl=1))
        kcl = int(n / 3.0)
        u = random_state(o, 11)
        ksq, _, fac = shells(o, u)
        band = ((kcl - 0.5) ** 2 < ksq) & (ksq <= (kcl + 0.5) ** 2)
        usq = np.sum(np.abs(u[..., :3]) ** 2, axis=3)
        e_ref = np.sum((fac * usq)[band])
        e, f = o.cl_energy(u)
        assert band.sum() > 100
        assert abs(e - e_ref) <= 1e-13 * e_ref
        assert abs(f - np.sqrt(2 * e_ref / kcl)) <= 1e-13 * f


def test_cl_spectra_by_direct_summation():
    o = O.Oracle(O.make_cfg(32, i_strat=1, i_cl=1, vis=1e-3))
    u = random_state(o, 12)
    r, _ = o.rhs(u)
    e, fac_cl = o.cl_energy(u)
    sp = o.spectra(u, r, fac_cl)
    ksq, is_, fac = shells(o, u)
    kx, ky, kz = o.kvectors()
    rk_sq = ksq.astype(float)   # 2 pi box: physical = integer wavenumbers
    kcl = int(32 / 3.0)
    vist = np.where(ksq == 0, 0.0, fac_cl * (0.145 + 5.01 * np.exp(-(3.03 * kcl) / np.sqrt(np.maximum(ksq, 1)))))
    usq = np.sum(np.abs(u[..., :3]) ** 2, axis=3)
    for s in (1, 5, 10, 14):
        m = is_ == s
        dk = np.sum((fac * rk_sq * (1e-3 + vist) * usq)[m])
        vt = np.sum((fac * vist)[m])
        assert abs(sp["dk"][s - 1] - dk) <= 1e-12 * dk
        assert abs(sp["vt"][s - 1] - vt) <= 1e-12 * vt
    plain = O.Oracle(O.make_cfg(32, i_strat=1, vis=1e-3)).spectra(u, r)
    assert np.all(plain["vt"] == 0) and np.all(sp["dk"][1:] > plain["dk"][1:])


def test_cl_in_the_integrating_factor_decays_shear_waves_exactly():
    n, nu, dt, nsteps = 32, 2e-3, 0.01, 6
    kcl = int(n / 3.0)
    o = O.Oracle(O.make_cfg(n, i_cl=2, vis=nu, dt0=dt))
    u = o.zeros_spec()
    a = np.array([0.4, 0.3])
    ks = (1, kcl)
    for amp, k in zip(a, ks):
        u[0, k, 0, 0] = -0.5j * amp          # sin(k y) in u: ky = +k ...
        u[0, o.nym - k, 0, 0] = 0.5j * amp   # ... and ky = -k
    got, _, _, st = o.step(u, nsteps)
    amp = a.copy()
    for _ in range(nsteps):
        e_cl = 2 * (amp[1] / 2) ** 2                  # both ky = +-k_cl modes, kx = 0
        fac_cl = np.sqrt(2 * e_cl / kcl)
        for i, k in enumerate(ks):
            nut = fac_cl * (0.145 + 5.01 * np.exp(-(3.03 * kcl) / k))
            amp[i] *= np.exp(-(nu + nut) * k * k * dt)
    for amp_i, k in zip(amp, ks):
        assert abs(got[0, k, 0, 0] - (-0.5j * amp_i)) <= 1e-12 * amp_i
    assert abs(st[18] - np.sqrt(2 * 2 * (amp_before_last(a, ks, nu, dt, nsteps, kcl) / 2) ** 2 / kcl)) < 1e-12


def amp_before_last(a, ks, nu, dt, nsteps, kcl):
    """the k_cl amplitude at the start of the last step (for fac_cl of that step)"""
    amp = np.array(a, dtype=float)
    for _ in range(nsteps - 1):
        fac_cl = np.sqrt(2 * (2 * (amp[1] / 2) ** 2) / kcl)
        for i, k in enumerate(ks):
            nut = fac_cl * (0.145 + 5.01 * np.exp(-(3.03 * kcl) / k))
            amp[i] *= np.exp(-(nu + nut) * k * k * dt)
    return amp[1]


def test_step_with_forcing_holds_the_low_shells():
    """orc_step_ext: force.f after every step; the low shells end each step
    at the target, the rest evolve freely"""
    kf = 2
    o = O.Oracle(O.make_cfg(32, i_strat=1, grav=16.0, lapse0=1.0, vis=5e-3, dt0=0.01, k_force=kf))
    u0 = random_state(o, 21)
    ek0 = o.spectra(u0, o.rhs(u0)[0])["ek"]
    u, _, _, _ = o.step(u0, 4, ek0=ek0)
    got = shell_energy(o, u, o.lm)
    assert np.allclose(got[1:kf + 1], ek0[1:kf + 1], rtol=1e-12)
    free, _, _, _ = o.step(u0, 4)
    assert not np.allclose(shell_energy(o, free, kf + 1)[1:], ek0[1:kf + 1], rtol=1e-6)
