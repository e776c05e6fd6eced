"""Seeded initial conditions: solenoidal, Hermitian on kx = 0, u_rms = 1,
and Taylor-Green in SAM's stored form -- CPU only."""
import numpy as np

from oracle import oracle as O
from sam_amd import ic


def test_random_solenoidal_properties():
    n = 32
    u = ic.random_solenoidal(n, k0=4, seed=1, i_strat=1)
    o = O.Oracle(O.make_cfg(n, i_strat=1))
    kx, ky, kz = o.kvectors()
    KZ, KY, KX = np.meshgrid(kz, ky, kx, indexing="ij")
    div = KX * u[..., 0] + KY * u[..., 1] + KZ * u[..., 2]
    assert np.abs(div).max() < 1e-15
    p = u[:, :, 0, :3]
    mir = np.conj(p[(-kz) % o.nzm][:, (-ky) % o.nym])
    assert np.abs(p - mir).max() < 1e-15
    phys = o.to_physical(u)
    urms2 = np.mean(np.sum(phys[..., :3] ** 2, axis=-1)) / 3
    assert abs(urms2 - 1.0) < 1e-12
    assert np.abs(u[..., 3]).max() == 0.0
    # same seed, same field
    assert np.array_equal(u, ic.random_solenoidal(n, k0=4, seed=1, i_strat=1))


def test_taylor_green_physical():
    o = O.Oracle(O.make_cfg(16))
    phys = o.to_physical(ic.taylor_green(16))
    x = np.arange(16) * 2 * np.pi / 16
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    assert np.abs(phys[..., 0] - np.sin(X) * np.cos(Y) * np.cos(Z)).max() < 1e-14
    assert np.abs(phys[..., 1] + np.cos(X) * np.sin(Y) * np.cos(Z)).max() < 1e-14
    assert np.abs(phys[..., 2]).max() < 1e-14


def test_spectral_pad_round_trip():
    """padding to a finer grid keeps every mode (energy) and truncation undoes it"""
    from sam_amd import les as L
    u = ic.random_solenoidal(32, k0=4, seed=3, i_strat=1)
    v = ic.spectral_pad(u, 64)
    d = ic.grid_dims(64, i_strat=1)
    assert v.shape == (d["nzm"], d["nym"], d["nxm"], 4)
    assert np.isclose(np.linalg.norm(v), np.linalg.norm(u), rtol=1e-14, atol=0)
    np.testing.assert_array_equal(L.spectral_truncate(v, 32), u)
