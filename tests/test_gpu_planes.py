"""GPU parity of the visualization output (write_planes.f, get_mean.f) against
the oracle: planes of u, v, w, T', the kept pressure and the derivative
fields, and the horizontal means, after a few steps.  FP64 bar 1e-12 (max
abs error over the field's max), split mode 1e-5."""
import numpy as np
import pytest

from sam_amd import ic
from sam_amd import samhip as S

pytestmark = pytest.mark.gpu

N = 32
PHYS = dict(i_strat=1, grav=16.0, to=1.0, vis=5e-3, pr=1.0, i_cfl=0, dt0=0.01)
K_XY, J_XZ, I_YZ = (5, 20), (3,), (1, 32)


def err(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def oracle_planes(oracle_mod, u0, nsteps, i_strat=1):
    o = oracle_mod.Oracle(oracle_mod.make_cfg(N, lapse0=1.0, **dict(PHYS, i_strat=i_strat)))
    u, _, _, _, p = o.step(u0, nsteps, want_p=True)
    return o.planes(u, p, K_XY, J_XZ, I_YZ)


@pytest.mark.parametrize("prec,tol", [(S.PREC_FP64, 1e-12), (S.PREC_SPLIT, 1e-5)])
def test_planes_match_oracle(oracle_mod, prec, tol):
    u0 = ic.random_solenoidal(N, k0=3.0, seed=41, i_strat=1)
    sam = S.Sam(S.make_params(N, lapse=1.0, precision=prec, **PHYS))
    S.set_planes(sam, True)
    sam.set_state(u0)
    sam.step(3)
    got = S.planes(sam, K_XY, J_XZ, I_YZ)
    ref = oracle_planes(oracle_mod, u0, 3)
    for name, g, r in zip(("mean", "xy", "xz", "yz"), got, ref):
        assert g.shape == r.shape, name
        if name == "mean":
            assert err(g, r) < tol, name
            continue
        for n in range(g.shape[2]):
            for m in range(g.shape[3]):
                assert err(g[:, :, n, m], r[:, :, n, m]) < tol, (name, n, m)


def test_unstratified_planes_and_state_unchanged(oracle_mod):
    u0 = ic.random_solenoidal(N, k0=3.0, seed=42, i_strat=0)
    sam = S.Sam(S.make_params(N, **dict(PHYS, i_strat=0)))
    S.set_planes(sam, True)
    sam.set_state(u0)
    sam.step(2)
    before = sam.get_state()
    got = S.planes(sam, K_XY, J_XZ, I_YZ)
    ref = oracle_planes(oracle_mod, u0, 2, i_strat=0)
    assert got[1].shape[2] == 6
    for g, r in zip(got[1:], ref[1:]):
        assert err(g, r) < 1e-12
    assert np.array_equal(sam.get_state(), before)        # output only
    sam.step(1)
    plain = S.Sam(S.make_params(N, **dict(PHYS, i_strat=0)))
    plain.set_state(u0)
    plain.step(3)
    assert np.array_equal(sam.get_state(), plain.get_state())   # keeping p does not change the step


def test_plane_positions_checked():
    sam = S.Sam(S.make_params(N, **PHYS))
    with pytest.raises(S.SamError):
        S.planes(sam, (0,))
    with pytest.raises(S.SamError):
        S.planes(sam, (), (N + 1,))
    with pytest.raises(S.SamError):
        S.planes(sam, tuple(range(1, 11)))
