"""CPU checks of bench.py's bytes model: the step's bytes at the grid sizes
worked out by hand for the fused 3-pass step, the read/write split that
selects each kernel's streaming mix, and the step floor built from it (no
GPU)."""
import math

import bench as B


def test_step_bytes_match_the_hand_counts():
    # 512^3 stratified: 112.3 GB/step FP64, 65.7 GB/step FP32-split (fused
    # 3-pass model); 1024^3 FP32-split 526.2 GB/step; 64^3 unstratified FP64
    # 0.16 GB/step
    assert abs(B.step_bytes(512, 16) / 1e9 - 112.3) < 0.05
    assert abs(B.step_bytes(512, 8) / 1e9 - 65.7) < 0.05
    assert abs(B.step_bytes(1024, 8) / 1e9 - 526.2) < 0.05
    assert abs(B.step_bytes(64, 16, i_strat=0) / 1e9 - 0.16) < 0.005


def test_kernel_bytes_sum_to_the_step():
    # five kernels per substep, the zrhs figure averaged over the substeps
    for n, bc in ((512, 16), (512, 8), (256, 16), (1024, 8)):
        kb = B.kernel_bytes(n, bc)
        per_sub = sum(kb.values())
        assert math.isclose(3 * per_sub, B.step_bytes_algorithmic(n, bc), rel_tol=1e-12)


def test_read_write_split_adds_up():
    for n, bc in ((512, 16), (512, 8), (1024, 8), (128, 8)):
        for fused in (False, True):
            kb = B.kernel_bytes(n, bc, zrhs_fused=fused)
            rw = B.kernel_rw(n, bc, zrhs_fused=fused)
            assert set(kb) == set(rw)
            for k in kb:
                assert math.isclose(sum(rw[k]), kb[k], rel_tol=1e-12), k


def test_each_kernel_takes_the_mix_nearest_its_ratio():
    rw = B.kernel_rw(512, 16, zrhs_fused=True)
    got = {k: B.nearest_mix(*v) for k, v in rw.items()}
    # x row: 4 fields in, 9 out; y inverse: 4 Z in (Z/Y = 341/512), 4 Y out;
    # y forward the reverse; z-RHS reads fluxes and state, writes state and
    # the next z inverse
    assert got == {"zb": (2, 3), "yb": (2, 3), "xrow": (4, 9), "yf": (3, 2), "zrhs": (3, 2)}
    assert B.nearest_mix(1, 0) == (1, 0) and B.nearest_mix(0, 5) == (0, 1)


def test_step_floor_uses_each_kernel_at_its_rate():
    rates = {m: 5000.0 for m in B.STREAM_MIXES}   # GB/s
    prof = {"zb": (0.0, 1), "yb": (0.0, 3), "xrow": (0.0, 3), "yf": (0.0, 3), "zrhs": (0.0, 3)}
    fl = B.step_floor_ms(512, 16, prof, 1, rates)
    # every kernel at 5000 GB/s: the floor is the moved bytes over 5 TB/s
    assert math.isclose(fl, B.bytes_moved(512, 16, prof, 1) / 5e12 * 1e3, rel_tol=1e-12)
