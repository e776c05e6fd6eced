"""GPU: the learned closure end to end -- DNS -> box-filtered samples ->
training on the GPU -> LES in the loop -- and what it gains over the LES
without a model; the dynamic model as the arbiter with fallback."""
import numpy as np
import pytest

from sam_amd import ic
from sam_amd import les as L
from sam_amd import samhip as S

pytestmark = pytest.mark.gpu

KW = dict(i_strat=1, grav=64.0, lapse=1.0, to=1.0, vis=4e-3, pr=1.0, i_cfl=0, dt0=2e-3,
          precision=S.PREC_SPLIT)


def shells_err(e, ref, k0, k1):
    return float(np.linalg.norm(e[k0:k1] - ref[k0:k1]) / np.linalg.norm(ref[k0:k1]))


@pytest.fixture(scope="module")
def trained():
    n_dns, r = 128, 4
    dns = S.Sam(S.make_params(n_dns, **KW))
    dns.set_state(ic.random_solenoidal(n_dns, k0=6, seed=20261203, i_strat=1))
    dns.step(100)
    mlp = S.Mlp(seed=1, pool_capacity=6 * (n_dns // r) ** 3)
    L.collect_samples(dns, mlp, r, nsnap=6, steps_between=20)
    losses = L.train(mlp, 400, batch=16384, lr=2e-3, seed=3)
    return dns, mlp, losses


def test_training_on_filtered_dns(trained):
    dns, mlp, losses = trained
    assert np.all(np.isfinite(losses))
    assert losses[-20:].mean() < 0.7 * losses[:5].mean(), (losses[:5], losses[-5:])
    # a-priori: the network's tau_12 correlates with the filtered DNS stress
    f, y = mlp.get_samples(0, 8192)
    pred = mlp.infer(f)
    corr = [np.corrcoef(pred[:, k], y[:, k])[0, 1] for k in (3, 4, 5)]
    assert min(corr) > 0.5, corr


def test_learned_closure_gains_over_no_model(trained):
    dns, mlp, _ = trained
    n_les, nsteps = 32, 300
    u_start = dns.get_state()
    ref = dns.step(nsteps, want_spectra=True)
    e_dns = L.shell_energy(dns.spectra())
    errs = {}
    for model in (S.SGS_NONE, S.SGS_MLP, S.SGS_DYN):
        les = S.Sam(S.make_params(n_les, sgs=model, **KW))
        if model == S.SGS_MLP:
            S.attach_mlp(les, mlp)
        les.set_state(L.spectral_truncate(u_start, n_les))
        les.step(nsteps, want_spectra=True)
        e = L.shell_energy(les.spectra())
        assert np.all(np.isfinite(e))
        errs[model] = shells_err(e, e_dns[: len(e)], 2, les.kxm)
    print("E(k) relative error vs DNS: none %.3f  mlp %.3f  dynamic %.3f" %
          (errs[S.SGS_NONE], errs[S.SGS_MLP], errs[S.SGS_DYN]))
    assert errs[S.SGS_MLP] < errs[S.SGS_NONE]


def test_dynamic_model_judges_and_falls_back(trained):
    _, good, _ = trained
    les = S.Sam(S.make_params(32, sgs=S.SGS_MLP, **KW))
    les.set_state(ic.random_solenoidal(32, k0=6, seed=5, i_strat=1))
    S.attach_mlp(les, good)
    err_good, sw = S.sgs_check(les, 1e9)
    assert not sw and np.isfinite(err_good)
    bad = S.Mlp(seed=9)
    p = bad.get_params()
    p[:S.MLP_NPARAM] *= 20.0                     # a badly scaled, untrained network
    bad.set_params(p)
    S.attach_mlp(les, bad)
    err_bad, sw = S.sgs_check(les, 5.0 * max(err_good, 1e-3))
    assert sw and err_bad > err_good
    d = les.step(2)
    assert d["sgs_active"] == S.SGS_DYN


def test_dynamic_labels_are_the_unit_coefficient_dm_stress():
    """SAM_LABELS_DYNAMIC: the labels are tau_ij = -2|S|S_ij (deviatoric) and
    q_j = -(|S|/Pr_t) dT/dx_j of the filtered field's own gradient features"""
    dns = S.Sam(S.make_params(64, **KW))
    dns.set_state(ic.random_solenoidal(64, k0=6, seed=4, i_strat=1))
    dns.step(5)
    mlp = S.Mlp(seed=2, pool_capacity=(64 // 2) ** 3)
    pr_t = 0.7
    mlp.set_labels(S.LABELS_DYNAMIC, pr_t)
    mlp.add_samples(dns, 2)
    f, lab = mlp.get_samples(0, mlp.pool_size())
    g = f[:, :9].astype(np.float64).reshape(-1, 3, 3)
    s_ij = 0.5 * (g + g.transpose(0, 2, 1))
    sm = np.sqrt(2.0 * np.sum(s_ij * s_ij, axis=(1, 2)))
    tr = np.trace(s_ij, axis1=1, axis2=2) / 3.0
    ref = np.stack([-2 * sm * (s_ij[:, 0, 0] - tr), -2 * sm * (s_ij[:, 1, 1] - tr), -2 * sm * (s_ij[:, 2, 2] - tr),
                    -2 * sm * s_ij[:, 0, 1], -2 * sm * s_ij[:, 1, 2], -2 * sm * s_ij[:, 0, 2]]
                   + [-sm * f[:, 9 + j] / pr_t for j in range(3)], axis=1)
    assert np.linalg.norm(lab - ref) < 1e-5 * np.linalg.norm(ref)
    with pytest.raises(S.SamError):
        mlp.set_labels(S.LABELS_FILTERED)          # the pool holds the other kind


def test_unit_coefficient_network_tracks_dynamic_model():
    """a network trained on unit-coefficient dynamic labels, scaled by the
    dynamic procedure's C Delta^2 every 10 steps, reproduces the dynamic LES:
    time-averaged E(k) within 10% over 400 steps at 32^3 (the statistical
    noise of 32^3 shells is a few percent)"""
    n_dns, r = 128, 4
    dns = S.Sam(S.make_params(n_dns, **KW))
    u0 = ic.random_solenoidal(n_dns, k0=6, seed=20261203, i_strat=1)
    dns.set_state(u0)
    mlp = S.Mlp(seed=1, pool_capacity=4 * (n_dns // r) ** 3)
    mlp.set_labels(S.LABELS_DYNAMIC, 1.0)
    L.collect_samples(dns, mlp, r, nsnap=4, steps_between=25)
    losses = L.train(mlp, 600, batch=16384, lr=2e-3, seed=3)
    assert losses[-20:].mean() < 0.05 * losses[:5].mean()
    n_les = 32
    u = L.spectral_truncate(u0, n_les)
    avg = {}
    for model in (S.SGS_DYN, S.SGS_MLP):
        les = S.Sam(S.make_params(n_les, sgs=model, **KW))
        if model == S.SGS_MLP:
            S.attach_mlp(les, mlp)
            S.set_mlp_refresh(les, 10)
        les.set_state(u)
        les.step(200)
        acc = 0.0
        for _ in range(10):
            les.step(20, want_spectra=True)
            acc = acc + L.shell_energy(les.spectra())
        avg[model] = acc / 10
        d = les.step(1)
        assert d["sgs_active"] == model
    k1 = len(avg[S.SGS_DYN]) - 1
    err = shells_err(avg[S.SGS_MLP], avg[S.SGS_DYN], 1, k1)
    print("time-averaged E(k) rel err, unit-coefficient network vs dynamic: %.3f" % err)
    assert err < 0.10, err


def test_lagged_dynamic_coefficient_holds_between_refreshes():
    """sam_set_dyn_refresh(n): the dynamic procedure runs at the first substep
    of every n-th step and C Delta^2 is held in between, so n steps equal the
    Smagorinsky step with c_smag = C Delta^2 / Delta^2"""
    n = 32
    kw = dict(KW, precision=S.PREC_FP64)
    u0 = ic.random_solenoidal(n, k0=6, seed=7, i_strat=1)
    lag = S.Sam(S.make_params(n, sgs=S.SGS_DYN, **kw))
    S.set_dyn_refresh(lag, 4)
    lag.set_state(u0)
    d = lag.step(4)
    cds = d["c_delta_sq"]          # the volume-averaged coefficient may be negative (dm.f does not clip)
    assert cds != 0
    delta_sq = (2 * np.pi / n) ** 2
    smag = S.Sam(S.make_params(n, sgs=S.SGS_SMAG, c_smag=cds / delta_sq, **kw))
    smag.set_state(u0)
    smag.step(4)
    a, b = lag.get_state(), smag.get_state()
    assert np.linalg.norm(a - b) <= 1e-12 * np.linalg.norm(b)
    every = S.Sam(S.make_params(n, sgs=S.SGS_DYN, **kw))
    every.set_state(u0)
    every.step(4)
    assert np.linalg.norm(every.get_state() - b) > 1e-10 * np.linalg.norm(b)


def test_dynamic_coefficient_labels_match_oracle(oracle_mod):
    """SAM_LABELS_DYNAMIC_COEF: the labels are the unit-coefficient labels
    times C Delta^2 of dm.f's dynamic procedure run on the filtered field;
    that coefficient equals the oracle's dm.f coefficient of the same box
    filter of the same DNS field (FP32 filtered field: 2e-3)"""
    O = oracle_mod
    n, r = 64, 2
    kw = dict(KW, precision=S.PREC_FP64)
    dns = S.Sam(S.make_params(n, **kw))
    dns.set_state(ic.random_solenoidal(n, k0=6, seed=4, i_strat=1))
    dns.step(5)
    nets = {}
    for mode in (S.LABELS_DYNAMIC, S.LABELS_DYNAMIC_COEF):
        m = S.Mlp(seed=2, pool_capacity=(n // r) ** 3)
        m.set_labels(mode, 1.0)
        m.add_samples(dns, r)
        nets[mode] = m
    c = nets[S.LABELS_DYNAMIC_COEF].last_coefficient()
    _, unit = nets[S.LABELS_DYNAMIC].get_samples(0, (n // r) ** 3)
    _, coef = nets[S.LABELS_DYNAMIC_COEF].get_samples(0, (n // r) ** 3)
    assert np.linalg.norm(coef - c * unit) <= 1e-6 * np.linalg.norm(coef)
    # the oracle: box filter of the physical field, dm.f on the coarse grid
    o = O.Oracle(O.make_cfg(n, i_strat=1))
    phys = o.to_physical(dns.get_state())
    cx = n // r
    coarse = phys[..., :3].reshape(cx, r, cx, r, cx, r, 3).mean(axis=(1, 3, 5)).astype(np.float32)
    oc = O.Oracle(O.make_cfg(cx, sgs=2))
    _, c_ref = oc.eddy_visc(np.asfortranarray(coarse.astype(np.float64)))
    print(f"C Delta^2 of the filtered field: GPU {c:.6e}, oracle {c_ref:.6e}")
    assert abs(c - c_ref) <= 2e-3 * abs(c_ref)
