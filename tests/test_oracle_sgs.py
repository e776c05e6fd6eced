"""Pin the oracle's SGS pieces (sm.f, dm.f, filter.f on the collocated grid)
-- CPU only.  dm.f/sm.f cannot be compiled (they include speciso.h, which
the reference does not hold), so they are pinned by:
- closed forms: the centred-difference strain of a sinusoid, the (1,4,1)/6
  filter's transfer function, and the Germano-Lilly coefficient of a
  two-component field u = a sin(k1 y), v = b cos(k2 y), evaluated point by
  point from the analytic fields (dm.f:131-149,166);
- invariances the coefficient must have: Galilean (u -> u + U0), amplitude
  (u -> lambda u) and a cyclic relabelling of the axes with (u, v, w) and
  the box lengths;
- an independent numpy restatement (np.roll stencils, FFT-diagonalised
  filter).
Mutants of the numpy restatement (wrong alpha^2, sign, ifact, pair map,
unfiltered Leonard term) each fail at least one of these checks, which
shows the checks can tell such mistakes apart."""
import numpy as np

from oracle import oracle as O
from sam_amd import ic


def test_smagorinsky_sinusoid_closed_form():
    n, c = 16, 0.05
    o = O.Oracle(O.make_cfg(n, sgs=1, c_smag=c))
    y = np.arange(n) * 2 * np.pi / n
    phys = np.zeros((n, n, n, 3), order="F")
    phys[:, :, :, 0] = np.sin(2 * y)[None, :, None]
    vt, cd = o.eddy_visc(phys)
    dy = 2 * np.pi / n
    dudy = 2 * np.cos(2 * y) * np.sin(2 * dy) / (2 * dy)      # centred difference of sin(2y)
    smag = np.abs(dudy)                                         # |S| = sqrt(2*2*S12^2), S12 = dudy/2
    assert abs(cd - c * o.delta_sq) < 1e-15
    assert np.abs(vt - c * o.delta_sq * smag[None, :, None]).max() < 1e-14


def test_filter_transfer_function():
    n = 16
    o = O.Oracle(O.make_cfg(n, sgs=2))
    x = np.arange(n) * 2 * np.pi / n
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    f = np.cos(3 * X + 2 * Y - Z)
    g = lambda k: (4 + 2 * np.cos(k * 2 * np.pi / n)) / 6
    out = o.test_filter(np.asfortranarray(f))
    assert np.abs(out - g(3) * g(2) * g(1) * f).max() < 1e-14


def _dm_numpy(phys, n, L=2 * np.pi):
    d = L / n
    u = [phys[..., i] for i in range(3)]
    D = lambda f, ax: (np.roll(f, -1, ax) - np.roll(f, 1, ax)) / (2 * d)
    g = [[D(u[i], j) for j in range(3)] for i in range(3)]
    pairs = [(0, 0), (1, 1), (2, 2), (0, 1), (1, 2), (0, 2)]
    S = [0.5 * (g[i][j] + g[j][i]) for i, j in pairs]
    w = np.array([1, 1, 1, 2, 2, 2.0])
    mag = lambda s: np.sqrt(2 * sum(w[k] * s[k] ** 2 for k in range(6)))
    k = np.fft.fftfreq(n, 1.0 / n)
    tf1 = (4 + 2 * np.cos(2 * np.pi * k / n)) / 6
    TF = tf1[:, None, None] * tf1[None, :, None] * tf1[None, None, :]
    filt = lambda f: np.real(np.fft.ifftn(np.fft.fftn(f) * TF))
    sm = mag(S)
    uh = [filt(x) for x in u]
    sh = [filt(x) for x in S]
    smh = mag(sh)
    M = [-filt(sm * S[k]) + 4.0 * smh * sh[k] for k in range(6)]
    Lm = [filt(u[i] * u[j]) - uh[i] * uh[j] for i, j in pairs]
    num = sum(w[k] * np.sum(Lm[k] * M[k]) for k in range(6))
    den = sum(w[k] * np.sum(M[k] * M[k]) for k in range(6))
    return -0.5 * num / den, sm


def test_dynamic_coefficient_independent_restatement():
    n = 16
    o = O.Oracle(O.make_cfg(n, sgs=2))
    u = ic.random_solenoidal(n, k0=4, seed=77)
    phys = o.to_physical(u)
    vt, cd = o.eddy_visc(phys)
    cd_np, sm = _dm_numpy(phys, n)
    assert abs(cd - cd_np) < 1e-12 * abs(cd_np)
    assert np.abs(vt - cd_np * sm).max() < 1e-12 * np.abs(vt).max()


# ---------------------------------------------------------------------------
# dm.f pins that a misreading of the formula would fail
# ---------------------------------------------------------------------------
PAIRS = [(0, 0), (1, 1), (2, 2), (0, 1), (1, 2), (0, 2)]     # dm.f ip/jp: 11 22 33 12 23 13
IFACT = np.array([1, 1, 1, 2, 2, 2.0])


def _dm_variant(phys, lengths, alpha_sq=4.0, sign=-0.5, ifact=IFACT, pairs=PAIRS, leonard_hat=True):
    """numpy Germano-Lilly C Delta^2 with switchable pieces (the mutants)"""
    n = phys.shape[:3]
    d = [lengths[a] / n[a] for a in range(3)]
    u = [phys[..., i] for i in range(3)]
    D = lambda f, ax: (np.roll(f, -1, ax) - np.roll(f, 1, ax)) / (2 * d[ax])
    g = [[D(u[i], j) for j in range(3)] for i in range(3)]
    S = [0.5 * (g[i][j] + g[j][i]) for i, j in PAIRS]
    mag = lambda s: np.sqrt(2 * sum(ifact[k] * s[k] ** 2 for k in range(6)))

    def filt(f):
        for ax in range(3):
            f = (np.roll(f, 1, ax) + 4 * f + np.roll(f, -1, ax)) / 6.0
        return f

    sm = mag(S)
    uh = [filt(x) for x in u]
    sh = [filt(x) for x in S]
    smh = mag(sh)
    M = [-filt(sm * S[k]) + alpha_sq * smh * sh[k] for k in range(6)]
    if leonard_hat:
        Lm = [filt(u[i] * u[j]) - uh[i] * uh[j] for i, j in pairs]
    else:
        Lm = [filt(u[i] * u[j]) - u[i] * u[j] for i, j in pairs]
    num = sum(ifact[k] * np.sum(Lm[k] * M[k]) for k in range(6))
    den = sum(ifact[k] * np.sum(M[k] * M[k]) for k in range(6))
    return sign * num / den


MUTANTS = {
    "alpha_sq=2": dict(alpha_sq=2.0),
    "sign": dict(sign=0.5),
    "ifact=1": dict(ifact=np.ones(6)),
    "pair 23<->13": dict(pairs=[(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]),
    "pair 12->11": dict(pairs=[(0, 0), (1, 1), (2, 2), (0, 0), (1, 2), (0, 2)]),
    "unfiltered Leonard term": dict(leonard_hat=False),
}

BOX = (2 * np.pi, 3.0, 5.0)


def _oracle_cdsq(phys, lengths):
    n = phys.shape[0]
    o = O.Oracle(O.make_cfg(n, phys.shape[1], phys.shape[2], xl=lengths[0], yl=lengths[1], zl=lengths[2], sgs=2))
    return o.eddy_visc(np.asfortranarray(phys))[1]


def _turbulent_phys(seed=77):
    n = 16
    o = O.Oracle(O.make_cfg(n, xl=BOX[0], yl=BOX[1], zl=BOX[2]))
    u = ic.random_solenoidal(n, k0=4, seed=seed, xl=BOX[0], yl=BOX[1], zl=BOX[2])
    return o.to_physical(u)


def _check_galilean(cdsq):
    phys = _turbulent_phys()
    c0 = cdsq(phys, BOX)
    shifted = phys + np.array([0.7, -1.3, 2.1])[None, None, None, :]
    return abs(cdsq(shifted, BOX) - c0) <= 1e-10 * abs(c0)


def _check_amplitude(cdsq):
    phys = _turbulent_phys()
    c0 = cdsq(phys, BOX)
    return abs(cdsq(3.7 * phys, BOX) - c0) <= 1e-10 * abs(c0)


def _check_permutation(cdsq):
    """new x = old z, new y = old x, new z = old y, with (u,v,w) and lengths"""
    phys = _turbulent_phys()
    c0 = cdsq(phys, BOX)
    rot = np.ascontiguousarray(np.transpose(phys, (2, 0, 1, 3))[..., [2, 0, 1]])
    box = (BOX[2], BOX[0], BOX[1])
    return abs(cdsq(rot, box) - c0) <= 1e-10 * abs(c0)


def _two_mode_closed_form(n, a, k1, p1, b, k2, p2, L=2 * np.pi):
    """C Delta^2 of u = a sin(k1 y + p1), v = b sin(k2 y + p2), w = 0 from
    the analytic fields: the fields are uniform in x and z, so the x and z
    filter passes are the identity and only the y stencil (1,4,1)/6 acts
    (G(k) = (4 + 2 cos k dy)/6 on a mode).  The only strain components are
    S12 = du/dy / 2 and S22 = dv/dy (centred differences), so
    L:M = L22 M22 + 2 L12 M12 (ifact = 2 off the diagonal)."""
    dl = L / n
    y = np.arange(n) * dl
    G = lambda k: (4 + 2 * np.cos(k * dl)) / 6.0
    S12 = lambda yy: 0.5 * a * np.cos(k1 * yy + p1) * np.sin(k1 * dl) / dl
    S22 = lambda yy: b * np.cos(k2 * yy + p2) * np.sin(k2 * dl) / dl
    Smag = lambda yy: np.sqrt(2 * (S22(yy) ** 2 + 2 * S12(yy) ** 2))
    F = lambda h: (h(y - dl) + 4 * h(y) + h(y + dl)) / 6.0             # the stencil on analytic values
    t1, t2 = k1 * y + p1, k2 * y + p2
    uh, vh = G(k1) * a * np.sin(t1), G(k2) * b * np.sin(t2)
    s12h, s22h = G(k1) * S12(y), G(k2) * S22(y)
    smagh = np.sqrt(2 * (s22h ** 2 + 2 * s12h ** 2))
    M22 = -F(lambda yy: Smag(yy) * S22(yy)) + 4.0 * smagh * s22h
    M12 = -F(lambda yy: Smag(yy) * S12(yy)) + 4.0 * smagh * s12h
    # filtered mode products: sin^2 and sin*sin
    vv = 0.5 * b * b * (1 - G(2 * k2) * np.cos(2 * t2))
    uv = 0.5 * a * b * (G(k1 - k2) * np.cos(t1 - t2) - G(k1 + k2) * np.cos(t1 + t2))
    L22 = vv - vh * vh
    L12 = uv - uh * vh
    num = np.sum(L22 * M22 + 2 * L12 * M12)          # M11 = M33 = M23 = M13 = 0
    den = np.sum(M22 * M22 + 2 * M12 * M12)
    return -0.5 * num / den


TWO_MODE = (16, 1.3, 1, 0.4, 0.8, 2, 1.1)   # k1 + k2 odd: L:M does not cancel over y


def _two_mode_phys(n, a, k1, p1, b, k2, p2, L=2 * np.pi):
    y = np.arange(n) * L / n
    phys = np.zeros((n, n, n, 3))
    phys[..., 0] = (a * np.sin(k1 * y + p1))[None, :, None]
    phys[..., 1] = (b * np.sin(k2 * y + p2))[None, :, None]
    return phys


def _check_closed_form(cdsq):
    ref = _two_mode_closed_form(*TWO_MODE)
    got = cdsq(_two_mode_phys(*TWO_MODE), (2 * np.pi,) * 3)
    return abs(got - ref) <= 1e-12 * abs(ref)


CHECKS = [_check_galilean, _check_amplitude, _check_permutation, _check_closed_form]


def test_dynamic_closed_form_two_mode():
    ref = _two_mode_closed_form(*TWO_MODE)
    assert abs(ref) > 1e-4
    got = _oracle_cdsq(_two_mode_phys(*TWO_MODE), (2 * np.pi,) * 3)
    assert abs(got - ref) <= 1e-12 * abs(ref), (got, ref)


def test_dynamic_invariances():
    for chk in CHECKS[:3]:
        assert chk(_oracle_cdsq), chk.__name__


def test_dynamic_checks_catch_mutants():
    """the unmutated restatement passes every check, each mutant fails one"""
    assert all(chk(_dm_variant) for chk in CHECKS)
    for name, kw in MUTANTS.items():
        fn = lambda phys, lengths, kw=kw: _dm_variant(phys, lengths, **kw)
        assert not all(chk(fn) for chk in CHECKS), name
