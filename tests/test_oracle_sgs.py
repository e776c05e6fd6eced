"""Pin the oracle's SGS pieces (sm.f, dm.f, filter.f on the collocated grid)
-- CPU only.  dm.f/sm.f cannot be compiled (they include the missing
speciso.h), so they are pinned by closed forms: centred-difference strain of
a sinusoid, the (1,4,1)/6 filter's transfer function, and an independent
numpy restatement (np.roll stencils, FFT-diagonalised filter) of the
Germano-Lilly coefficient."""
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
