"""Pin the oracle's visualization output (write_planes.f, get_mean.f,
x_deriv.f, y_deriv.f) against closed forms -- CPU only.

Each field is a sum of single Fourier modes, so every plane value, every
derivative and the horizontal-mean profile is known analytically at the
grid points.  This fixes the plane positions (1-based), the field order of
each plane kind, the derivative directions and signs, and the mean.
"""
import numpy as np

from oracle import oracle as O

N = 16
# field -> list of (amplitude, phase, kx, ky, kz); kx > 0 modes are stored
# once (f = 2 Re(c e^{ik.x}), c = A/2 e^{i phase})
MODES = {
    0: [(1.0, 0.0, 2, 3, 1), (0.3, 0.4, 1, -2, 2)],
    1: [(0.5, 0.3, 1, -2, 2)],
    2: [(0.7, -0.2, 3, 1, -1)],
    3: [(0.4, 1.1, 1, 1, 3)],
}
MEAN = {0: (0.2, 2), 3: (0.1, 1)}        # field -> (A, kz): A cos(kz z), kx = ky = 0
P_MODES = [(0.6, 0.5, 2, -1, 1)]


def spectral(o, modes, mean=None):
    a = o.zeros_spec(1)[..., 0]
    for amp, ph, kx, ky, kz in modes:
        a[kz % o.nzm, ky % o.nym, kx] += 0.5 * amp * np.exp(1j * ph)
    if mean:
        amp, kz = mean
        a[kz, 0, 0] += 0.5 * amp
        a[o.nzm - kz, 0, 0] += 0.5 * amp
    return a


def field(modes, mean, x, y, z, d=None):
    """analytic value (d = None) or derivative d in 'x', 'y', 'z'"""
    f = np.zeros(np.broadcast(x, y, z).shape)
    for amp, ph, kx, ky, kz in modes:
        arg = kx * x + ky * y + kz * z + ph
        if d is None:
            f = f + amp * np.cos(arg)
        else:
            k = {"x": kx, "y": ky, "z": kz}[d]
            f = f - amp * k * np.sin(arg)
    if mean:
        amp, kz = mean
        if d is None:
            f = f + amp * np.cos(kz * z)
        elif d == "z":
            f = f - amp * kz * np.sin(kz * z)
    return f


def test_planes_and_means_closed_form():
    o = O.Oracle(O.make_cfg(N, i_strat=1))
    u = np.stack([spectral(o, MODES[n], MEAN.get(n)) for n in range(4)], axis=3)
    u = np.asfortranarray(u)
    p = spectral(o, P_MODES)
    k_xy, j_xz, i_yz = (5, 12), (3,), (1, 9)
    mean, xy, xz, yz = o.planes(u, p, k_xy, j_xz, i_yz)
    g = 2 * np.pi * np.arange(N) / N
    X, Y = np.meshgrid(g, g, indexing="ij")
    ref_f = lambda n, x, y, z, d=None: (field(P_MODES, None, x, y, z, d) if n == "p"
                                        else field(MODES[n], MEAN.get(n), x, y, z, d))
    # xy: u v w T p du/dz dv/dz dT/dz at z = z_k
    for m, k in enumerate(k_xy):
        z = g[k - 1]
        want = [ref_f(0, X, Y, z), ref_f(1, X, Y, z), ref_f(2, X, Y, z), ref_f(3, X, Y, z), ref_f("p", X, Y, z),
                ref_f(0, X, Y, z, "z"), ref_f(1, X, Y, z, "z"), ref_f(3, X, Y, z, "z")]
        for n, w in enumerate(want):
            assert np.abs(xy[:, :, n, m] - w).max() < 1e-12, ("xy", m, n)
    # xz: u v w T p du/dy dw/dy dT/dy at y = y_j
    Xz, Zx = np.meshgrid(g, g, indexing="ij")
    for m, j in enumerate(j_xz):
        y = g[j - 1]
        want = [ref_f(0, Xz, y, Zx), ref_f(1, Xz, y, Zx), ref_f(2, Xz, y, Zx), ref_f(3, Xz, y, Zx),
                ref_f("p", Xz, y, Zx), ref_f(0, Xz, y, Zx, "y"), ref_f(2, Xz, y, Zx, "y"), ref_f(3, Xz, y, Zx, "y")]
        for n, w in enumerate(want):
            assert np.abs(xz[:, :, n, m] - w).max() < 1e-12, ("xz", m, n)
    # yz: u v w T p dv/dx dw/dx dT/dx at x = x_i
    for m, i in enumerate(i_yz):
        x = g[i - 1]
        want = [ref_f(0, x, Xz, Zx), ref_f(1, x, Xz, Zx), ref_f(2, x, Xz, Zx), ref_f(3, x, Xz, Zx),
                ref_f("p", x, Xz, Zx), ref_f(1, x, Xz, Zx, "x"), ref_f(2, x, Xz, Zx, "x"), ref_f(3, x, Xz, Zx, "x")]
        for n, w in enumerate(want):
            assert np.abs(yz[:, :, n, m] - w).max() < 1e-12, ("yz", m, n)
    # horizontal means
    for n in range(4):
        want = MEAN[n][0] * np.cos(MEAN[n][1] * g) if n in MEAN else np.zeros(N)
        assert np.abs(mean[:, n] - want).max() < 1e-13, n


def test_unstratified_plane_fields():
    """Lu = 3: six fields per plane, the derivative slots skip v (xz) / u (yz)"""
    o = O.Oracle(O.make_cfg(N))
    u = np.asfortranarray(np.stack([spectral(o, MODES[n], MEAN.get(n)) for n in range(3)], axis=3))
    p = spectral(o, P_MODES)
    mean, xy, xz, yz = o.planes(u, p, (4,), (7,), (2,))
    assert xy.shape[2] == 6 and xz.shape[2] == 6 and yz.shape[2] == 6 and mean.shape == (N, 3)
    g = 2 * np.pi * np.arange(N) / N
    Xz, Zx = np.meshgrid(g, g, indexing="ij")
    assert np.abs(xz[:, :, 4, 0] - field(MODES[0], MEAN[0], Xz, g[6], Zx, "y")).max() < 1e-12
    assert np.abs(xz[:, :, 5, 0] - field(MODES[2], None, Xz, g[6], Zx, "y")).max() < 1e-12
    assert np.abs(yz[:, :, 4, 0] - field(MODES[1], None, g[1], Xz, Zx, "x")).max() < 1e-12
