"""GPU: the learned closure's BF16-MFMA kernels against the CPU oracle
(oracle/mlp_oracle.py, same BF16 rounding points), its training, the
box-filter sample pipeline and its coupling into the LES step."""
import numpy as np
import pytest

from oracle import mlp_oracle as M
from sam_amd import ic
from sam_amd import samhip as S

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-300))


def random_params(seed, scale=0.25):
    rng = np.random.default_rng(seed)
    p = np.zeros(S.MLP_NPARAM + 32, np.float32)
    p[:S.MLP_NPARAM] = rng.standard_normal(S.MLP_NPARAM) * scale
    p[M.OFF["W1"]:M.OFF["b1"]].reshape(64, 16)[:, 12:] = 0
    p[M.OFF["W4"]:M.OFF["b4"]].reshape(16, 64)[9:] = 0
    p[M.OFF["b4"] + 9:S.MLP_NPARAM] = 0
    p[S.MLP_NPARAM:S.MLP_NPARAM + 12] = rng.uniform(0.5, 2.0, 12)
    p[S.MLP_NPARAM + 16:S.MLP_NPARAM + 25] = rng.uniform(0.5, 2.0, 9)
    return p


ACTS = [(S.ACT_RELU, M.RELU), (S.ACT_GELU, M.GELU)]


@pytest.mark.parametrize("act,oact", ACTS)
def test_inference_matches_oracle(act, oact):
    p = random_params(3)
    m = S.Mlp(seed=1)
    m.set_activation(act)
    m.set_params(p)
    rng = np.random.default_rng(4)
    f = rng.standard_normal((1000, 12)).astype(np.float32)
    got = m.infer(f)
    ref = M.Net(p, bf=True, act=oact).infer(f)
    assert rel(got, ref) < 2e-3, rel(got, ref)
    # and close to the unrounded network (BF16 operands, FP32 accumulate)
    assert rel(got, M.Net(p, bf=False, act=oact).infer(f)) < 5e-2


@pytest.mark.parametrize("act,oact", ACTS)
def test_gradient_matches_oracle(act, oact):
    p = random_params(5)
    m = S.Mlp(seed=1)
    m.set_activation(act)
    m.set_params(p)
    rng = np.random.default_rng(6)
    n = 4096
    f = rng.standard_normal((n, 12)).astype(np.float32)
    y = rng.standard_normal((n, 9)).astype(np.float32)
    loss, g = m.grad(f, y)
    loss_ref, g_ref = M.Net(p, bf=True, act=oact).loss_grad(f, y)
    assert abs(loss - loss_ref) < 1e-4 * loss_ref
    assert rel(g, g_ref) < 5e-3, rel(g, g_ref)
    for k in ("W1", "W2", "W3", "W4", "b1", "b4"):
        sl = slice(M.OFF[k], M.OFF[k] + int(np.prod(M.SHAPES[k])))
        assert rel(g[sl], g_ref[sl]) < 1e-2, k
    loss64, g64 = M.Net(p, bf=False, act=oact).loss_grad(f, y)
    assert rel(g, g64) < 5e-2


def test_activation_choice():
    """the two activations give different networks; unknown values are rejected"""
    p = random_params(3)
    m = S.Mlp(seed=1)
    assert m.activation() == S.ACT_RELU
    m.set_params(p)
    f = np.random.default_rng(4).standard_normal((256, 12)).astype(np.float32)
    relu = m.infer(f)
    m.set_activation(S.ACT_GELU)
    assert m.activation() == S.ACT_GELU
    assert rel(m.infer(f), relu) > 1e-2
    with pytest.raises(S.SamError):
        m.set_activation(2)
    assert m.activation() == S.ACT_GELU


def test_gradient_with_zero_output_columns():
    """a pool without temperature has zero scalar-flux columns: their output
    scale is 0 and their target 0 (not 0/0), the loss and gradient stay finite
    and match the oracle"""
    p = random_params(8)
    p[M.NPARAM + 16 + 6:M.NPARAM + 16 + 9] = 0.0
    p[M.NPARAM + 9:M.NPARAM + 12] = 0.0
    m = S.Mlp(seed=1)
    m.set_params(p)
    rng = np.random.default_rng(9)
    n = 4096
    f = rng.standard_normal((n, 12)).astype(np.float32)
    f[:, 9:] = 0.0
    y = rng.standard_normal((n, 9)).astype(np.float32)
    y[:, 6:] = 0.0
    loss, g = m.grad(f, y)
    loss_ref, g_ref = M.Net(p, bf=True).loss_grad(f, y)
    assert np.isfinite(loss) and np.all(np.isfinite(g))
    assert abs(loss - loss_ref) < 1e-4 * loss_ref
    assert rel(g, g_ref) < 5e-3, rel(g, g_ref)


def test_training_reduces_loss():
    """learn a Smagorinsky-like map tau = -|A| (A + A^T)/2 from gradients"""
    rng = np.random.default_rng(7)
    n = 1 << 16
    A = rng.standard_normal((n, 3, 3)).astype(np.float32)
    Sij = 0.5 * (A + np.transpose(A, (0, 2, 1)))
    mag = np.sqrt(2 * np.sum(Sij ** 2, axis=(1, 2)))
    tau = -mag[:, None, None] * Sij
    tr = np.trace(tau, axis1=1, axis2=2) / 3
    tau = tau - tr[:, None, None] * np.eye(3)
    f = np.concatenate([A.reshape(n, 9), rng.standard_normal((n, 3))], 1).astype(np.float32)
    y = np.stack([tau[:, 0, 0], tau[:, 1, 1], tau[:, 2, 2], tau[:, 0, 1], tau[:, 1, 2], tau[:, 0, 2],
                  -f[:, 9], -f[:, 10], -f[:, 11]], 1).astype(np.float32)
    m = S.Mlp(seed=2, pool_capacity=n)
    m.add_samples_host(f, y)
    m.fit_norms()
    losses = m.train(300, batch=8192, lr=2e-3, seed=1)
    assert losses[-20:].mean() < 0.1 * losses[:5].mean(), (losses[:5], losses[-5:])
    pred = m.infer(f[:4096])
    corr = np.corrcoef(pred[:, 3], y[:4096, 3])[0, 1]
    assert corr > 0.9, corr


def test_box_filter_samples(oracle_mod):
    n, r = 32, 4
    s = S.Sam(S.make_params(n, i_strat=1, grav=16.0, lapse=1.0, vis=5e-3))
    u = ic.random_solenoidal(n, k0=4, seed=9, i_strat=1)
    u[..., 3] = 0.3 * ic.random_solenoidal(n, k0=3, seed=10)[..., 0]
    s.set_state(u)
    m = S.Mlp(seed=0, pool_capacity=(n // r) ** 3)
    m.add_samples(s, r)
    f, lab = m.get_samples(0, (n // r) ** 3)
    o = oracle_mod.Oracle(oracle_mod.make_cfg(n, i_strat=1))
    ph = o.to_physical(u)                                   # (x, y, z, field)
    c = n // r
    blk = lambda a: a.reshape(c, r, c, r, c, r).mean(axis=(1, 3, 5))
    ub = [blk(ph[..., k]) for k in range(4)]
    pairs = [(0, 0), (1, 1), (2, 2), (0, 1), (1, 2), (0, 2)]
    tau = np.stack([blk(ph[..., i] * ph[..., j]) - ub[i] * ub[j] for i, j in pairs], -1)
    tr = (tau[..., 0] + tau[..., 1] + tau[..., 2]) / 3
    tau[..., :3] -= tr[..., None]
    q = np.stack([blk(ph[..., j] * ph[..., 3]) - ub[j] * ub[3] for j in range(3)], -1)
    ref_lab = np.concatenate([tau, q], -1).transpose(2, 1, 0, 3).reshape(-1, 9)   # x fastest
    assert rel(lab, ref_lab) < 1e-5
    dxi = c / (2 * np.pi)
    feats = []
    for k in range(12):
        fld, d = (k // 3, k % 3) if k < 9 else (3, k - 9)
        a = ub[fld]
        feats.append(0.5 * (np.roll(a, -1, d) - np.roll(a, 1, d)) * dxi)
    ref_f = np.stack(feats, -1).transpose(2, 1, 0, 3).reshape(-1, 12)
    assert rel(f, ref_f) < 1e-5


def test_zero_network_les_equals_dns(oracle_mod):
    """a network that outputs zero leaves the fluxes u_i u_j: the LES path
    with the MLP must reproduce the fused DNS step"""
    n = 32
    kw = dict(i_strat=1, grav=16.0, lapse=1.0, vis=5e-3, i_cfl=1, cfl0=0.5)
    u0 = ic.random_solenoidal(n, k0=4, seed=12, i_strat=1)
    dns = S.Sam(S.make_params(n, **kw))
    dns.set_state(u0)
    dns.step(3)
    les = S.Sam(S.make_params(n, sgs=S.SGS_MLP, **kw))
    m = S.Mlp(seed=0)
    p = m.get_params()
    p[:S.MLP_NPARAM] = 0.0
    m.set_params(p)
    S.attach_mlp(les, m)
    les.set_state(u0)
    les.step(3)
    assert rel(les.get_state(), dns.get_state()) < 1e-12


@pytest.mark.parametrize("act,oact", ACTS)
def test_les_rhs_with_network(oracle_mod, act, oact):
    """RHS of the LES step with a random network against the oracle's
    transforms + projection applied to u_i u_j + tau_ij(MLP oracle)"""
    O = oracle_mod
    n = 32
    p = random_params(13, scale=0.1)
    m = S.Mlp(seed=0)
    m.set_activation(act)
    m.set_params(p)
    les = S.Sam(S.make_params(n, sgs=S.SGS_MLP, i_strat=1, grav=16.0, lapse=1.0, vis=5e-3))
    S.attach_mlp(les, m)
    u = ic.random_solenoidal(n, k0=4, seed=14, i_strat=1)
    u[..., 3] = 0.2 * ic.random_solenoidal(n, k0=3, seed=15)[..., 0]
    les.set_state(u)
    r, _ = les.rhs()
    o = O.Oracle(O.make_cfg(n, i_strat=1, grav=16.0, lapse0=1.0, vis=5e-3))
    ph = o.to_physical(u)
    dxi = n / (2 * np.pi)
    feats = np.zeros((n, n, n, 12))
    for k in range(12):
        fld, d = (k // 3, k % 3) if k < 9 else (3, k - 9)
        feats[..., k] = 0.5 * (np.roll(ph[..., fld], -1, d) - np.roll(ph[..., fld], 1, d)) * dxi
    flat = feats.reshape(-1, 12, order="F").astype(np.float32)
    out = M.Net(p, bf=True, act=oact).infer(flat).reshape(n, n, n, 9, order="F")
    prod = o.products(ph)
    for oi, slot in enumerate((0, 2, 5, 1, 4, 3, 6, 7, 8)):
        prod[..., slot] += out[..., oi]
    F = o.z_trans_f(o.xy_trans_f(prod))
    r_ref = o.project(F, u)
    assert rel(r[..., :4], r_ref[..., :4]) < 1e-4, rel(r[..., :4], r_ref[..., :4])
