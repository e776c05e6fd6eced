"""GPU: the DNS step's fused z inverse (k_zrhs writes the next substep's
k_zb output into the Z buffer).  Every other writer of that buffer or of the
state must make the next substep recompute it: a stale or overwritten Z
buffer would feed the wrong state into the step (an O(1) error, far above
the tolerances here).  Each case is compared with a run that does not make
the interleaved call, and the plain run itself with the oracle.  (The
forcing rescales the state inside sam_step; the forcing tests against the
oracle cover that reset.)"""
import numpy as np
import pytest

from sam_amd import ic
from sam_amd import samhip as S

pytestmark = pytest.mark.gpu

N = 32
PHYS = dict(i_strat=1, grav=16.0, to=1.0, lapse=1.0, vis=5e-3, pr=1.0, i_cfl=0, dt0=0.01)


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def run(prec, u0, between=None, nsteps=(2, 2)):
    sam = S.Sam(S.make_params(N, precision=prec, **PHYS))
    sam.set_state(u0)
    sam.step(nsteps[0])
    if between:
        between(sam)
    sam.step(nsteps[1])
    out = sam.get_state()
    sam.close()
    return out


@pytest.mark.parametrize("prec,tol", [(S.PREC_FP64, 1e-13), (S.PREC_SPLIT, 1e-6)])
def test_interleaved_calls_recompute_z(oracle_mod, prec, tol):
    u0 = ic.random_solenoidal(N, k0=3.0, seed=51, i_strat=1)
    plain = run(prec, u0)
    o = oracle_mod.Oracle(oracle_mod.make_cfg(N, lapse0=1.0, **PHYS))
    ref = o.step(u0, 4)[0]
    assert rel(plain, ref) < (1e-12 if prec == S.PREC_FP64 else 1e-5)
    other = ic.random_solenoidal(N, k0=2.0, seed=52, i_strat=1)
    cases = {
        # stage entry points write the Z buffer
        "z_trans_b": lambda s: s.z_trans_b(other),
        "xy_trans_f": lambda s: s.xy_trans_f(np.random.default_rng(1).standard_normal((N, N, N, 9))),
        # the RHS-only pass runs the pipeline without the update
        "rhs": lambda s: s.rhs(),
    }
    for name, fn in cases.items():
        got = run(prec, u0, fn)
        assert rel(got, plain) < tol, (name, rel(got, plain))


def test_set_state_between_steps():
    """a new state after fused steps is transformed afresh: same result as
    a handle that starts from it"""
    u0 = ic.random_solenoidal(N, k0=3.0, seed=53, i_strat=1)
    u1 = ic.random_solenoidal(N, k0=2.0, seed=54, i_strat=1)
    got = run(S.PREC_FP64, u0, lambda s: s.set_state(u1))
    sam = S.Sam(S.make_params(N, **PHYS))
    sam.set_state(u1)
    sam.step(2)
    want = sam.get_state()
    assert np.array_equal(got, want)

