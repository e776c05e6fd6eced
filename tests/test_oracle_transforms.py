"""Pin the oracle's transforms (z_trans_b, x2z, xy_trans_b, xy_trans_f,
z_trans_f) -- CPU only.

Checks: fixtures written by the reference's own compiled routines
(tests/golden/transforms_*.npz, see make_golden.py), the round trip of
tst_transpose.f:98-146, and brute-force direct Fourier sums on small grids.
"""
import glob
import os

import numpy as np
import pytest

from oracle import oracle as O
from sam_amd import ic

GOLD = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "transforms_*.npz")))


def rel(a, b):
    return np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(np.asarray(b).ravel()), 1e-300)


def oracle_for(g):
    nx, ny, nz, nd, st = (int(v) for v in g["dims"][:5])
    xl, yl, zl = g["box"]
    return O.Oracle(O.make_cfg(nx, ny, nz, xl=xl, yl=yl, zl=zl, n_dealias=nd, i_strat=st))


@pytest.mark.parametrize("path", GOLD, ids=lambda p: os.path.basename(p))
def test_golden_reference_transforms(path):
    g = np.load(path)
    o = oracle_for(g)
    assert (o.kxm, o.kym, o.kzm, o.nxm, o.nym, o.nzm) == tuple(int(v) for v in g["dims"][5:])
    u = g["u"]
    zb = o.z_trans_b(np.asfortranarray(u[: o.nzm]))
    assert rel(zb, g["z_trans_b"]) < 1e-14
    y = o.x2z(np.asfortranarray(g["z_trans_b"]))
    assert rel(y[:, : o.nxm], g["x2z"][:, : o.nxm]) == 0.0
    assert rel(o.xy_trans_b(g["x2z"]), g["xy_trans_b"]) < 1e-14
    assert rel(o.xy_trans_f(g["p_in"]), g["xy_trans_f"]) < 1e-14
    assert rel(o.z_trans_f(g["xy_trans_f"]), g["z_trans_f"]) < 1e-14


@pytest.mark.parametrize("n,strat", [(16, 0), (32, 1)])
def test_round_trip_tst_transpose(n, strat):
    """backward then forward transform returns the coefficients
    (tst_transpose.f:124-146 counts |diff| > 1e-15)"""
    o = O.Oracle(O.make_cfg(n, i_strat=strat))
    u = ic.random_solenoidal(n, k0=3, seed=5, i_strat=strat)
    if strat:
        u[..., 3] = u[..., 0]
    phys = o.xy_trans_b(o.x2z(o.z_trans_b(u)))
    back = o.z_trans_f(o.xy_trans_f(phys)[..., : o.lu].copy(order="F"))
    assert np.abs(back - u).max() < 1e-15 * max(1.0, np.abs(u).max() * 10)


def brute_physical(o, u, field):
    """u(x) = sum_{kx=0} c e^{ik.x} + sum_{kx>0} 2 Re(c e^{ik.x}) on the grid"""
    kx, ky, kz = o.kvectors()
    xs = np.arange(o.nx) / o.nx
    ys = np.arange(o.ny) / o.ny
    zs = np.arange(o.nz) / o.nz
    out = np.zeros((o.nx, o.ny, o.nz))
    for i, a in enumerate(kx):
        fac = 1.0 if a == 0 else 2.0
        for j, b in enumerate(ky):
            for k, c in enumerate(kz):
                coef = u[k, j, i, field]
                if coef == 0:
                    continue
                ph = np.exp(2j * np.pi * (a * xs[:, None, None] + b * ys[None, :, None] + c * zs[None, None, :]))
                out += fac * (coef * ph).real
    return out


def test_backward_brute_force():
    o = O.Oracle(O.make_cfg(8, 8, 8, xl=3.0, yl=4.0, zl=5.0))
    u = ic.random_solenoidal(8, k0=2, seed=3, xl=3.0, yl=4.0, zl=5.0)
    phys = o.to_physical(u)
    for f in range(3):
        assert np.abs(phys[..., f] - brute_physical(o, u, f)).max() < 1e-13


def test_forward_brute_force():
    o = O.Oracle(O.make_cfg(8, 16, 8, i_strat=1))
    rng = np.random.default_rng(9)
    p = np.asfortranarray(rng.standard_normal((8, 16, 8, o.lr)))
    got = o.z_trans_f(o.xy_trans_f(p))
    kx, ky, kz = o.kvectors()
    xs, ys, zs = np.arange(8) / 8, np.arange(16) / 16, np.arange(8) / 8
    N = 8 * 16 * 8
    for q in (0, 4, 8):
        for i, a in enumerate(kx):
            for j, b in enumerate(ky):
                for k, c in enumerate(kz):
                    ph = np.exp(-2j * np.pi * (a * xs[:, None, None] + b * ys[None, :, None] + c * zs[None, None, :]))
                    ref = np.sum(p[..., q] * ph) / N
                    assert abs(got[k, j, i, q] - ref) < 1e-14
