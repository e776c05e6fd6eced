"""The Fortran driver (sam_amd/driver) over the C ABI: input.dat parsing and
checks (CPU; check_inputs.f error codes), and a 64^3 Taylor-Green run end to end
from SAM restart files (GPU)."""
import os
import subprocess

import numpy as np
import pytest

from sam_amd import ic
from sam_amd import io as sio

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "sam_amd", "driver", "samhip_driver")

needs_driver = pytest.mark.skipif(not os.path.exists(DRIVER), reason="driver not built (make -C sam_amd/driver)")

TG64 = dict(Nx=64, Ny=64, Nz=64, xL=2 * np.pi, yL=2 * np.pi, zL=2 * np.pi, vis=0.0025,
               n_steps=20, i_restart=1, nt_restart=0, i_cfl=0, dt0=0.01, nrk_max=3,
               n_dealias=1, i_fs=0, n_skip_h=5)


def run(tmp_path, **params):
    sio.write_input(str(tmp_path / "input.dat"), **params)
    return subprocess.run([DRIVER], cwd=tmp_path, capture_output=True, text=True, timeout=90)


@needs_driver
def test_missing_required_input(tmp_path):
    p = dict(TG64)
    del p["vis"]
    r = run(tmp_path, **p)
    assert r.returncode == 6 and "vis" in r.stdout


@needs_driver
def test_dt0_required_when_fixed_dt(tmp_path):
    p = dict(TG64)
    del p["dt0"]
    r = run(tmp_path, **p)
    assert r.returncode == 6 and "dt0" in r.stdout


@needs_driver
def test_invalid_value(tmp_path):
    r = run(tmp_path, **dict(TG64, nrk_max=4))
    assert r.returncode == 9 and "nrk_max" in r.stdout


@needs_driver
def test_malformed_and_bad_label(tmp_path):
    (tmp_path / "input.dat").write_text("Nx 64\nNy 1.2.3x\n")
    r = subprocess.run([DRIVER], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.returncode == 2
    (tmp_path / "input.dat").write_text("N$x 64\n")
    r = subprocess.run([DRIVER], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.returncode == 4


def test_python_parser_multipliers(tmp_path):
    (tmp_path / "in.dat").write_text("# comment\nNX 2k\nvis 2.5m   trailing\nEOF\nNy 3\n")
    v = sio.parse_labels(str(tmp_path / "in.dat"))
    assert v == {"nx": 2000.0, "vis": 2.5e6}


@pytest.mark.gpu
@needs_driver
def test_taylor_green_run_end_to_end(tmp_path, oracle_mod):
    """Taylor-Green 64^3, nu=1/400, dt=0.01, 20 RK3 steps from vel.000000"""
    u0 = ic.taylor_green(64)
    sio.write_vel(str(tmp_path / "vel.000000"), u0)
    sio.write_header(str(tmp_path / "header.000000"), TG64, time=0.0, dt=0.01, nt=0)
    r = run(tmp_path, **TG64)
    assert r.returncode == 0, r.stdout + r.stderr
    o = oracle_mod.Oracle(oracle_mod.make_cfg(64, vis=0.0025, dt0=0.01))
    u = sio.read_vel(str(tmp_path / "vel.000020"), o.nzm, o.nym, o.nxm, o.lu)
    ref, t, dt, st = o.step(u0, 20)
    assert np.linalg.norm(u - ref) / np.linalg.norm(ref) < 1e-12
    hdr = sio.parse_labels(str(tmp_path / "header.000020"))
    assert abs(hdr["time"] - t) < 1e-12 and hdr["nt"] == 20
    e = sio.read_log(str(tmp_path / "energy.out"))
    assert list(e[:, 0]) == [0, 5, 10, 15, 20]
    assert abs(e[0, 2] - 0.125) < 1e-4          # TG energy at t=0 (log precision)


@needs_driver
def test_learned_closure_needs_its_network(tmp_path):
    """i_les 4 reads the trained network from mlp.params before any GPU work"""
    r = run(tmp_path, **dict(TG64, i_les=4))
    assert r.returncode == 6 and "mlp.params" in r.stdout
    (tmp_path / "mlp.params").write_bytes(b"\0" * 400)
    r = run(tmp_path, **dict(TG64, i_les=4))
    assert r.returncode == 10 and "mlp.params" in r.stdout
    r = run(tmp_path, **dict(TG64, i_les=4, mlp_labels=2))
    assert r.returncode == 9 and "mlp_labels" in r.stdout
    r = run(tmp_path, **dict(TG64, i_les=4, mlp_act=2))
    assert r.returncode == 9 and "mlp_act" in r.stdout


@pytest.mark.gpu
@needs_driver
@pytest.mark.parametrize("act", [0, 1])
def test_learned_closure_run_matches_library(tmp_path, act):
    """i_les 4 from SAM's run files gives the state of the same network driven
    through the library directly (unit-coefficient labels, C Delta^2 every 5
    steps), 32^3 Taylor-Green, 12 steps; ReLU (mlp_act 0) and GELU (1)"""
    from sam_amd import les as L
    from sam_amd import samhip as S
    dns = S.Sam(S.make_params(64, vis=1e-3, dt0=0.01))
    dns.set_state(ic.random_solenoidal(64, k0=4, seed=11))
    dns.step(5)
    mlp = S.Mlp(seed=5, pool_capacity=(64 // 2) ** 3)
    mlp.set_activation(act)
    mlp.set_labels(S.LABELS_DYNAMIC, 0.7)
    mlp.add_samples(dns, 2)
    L.train(mlp, 60, batch=4096, lr=2e-3, seed=1)
    params = mlp.get_params()
    params.tofile(str(tmp_path / "mlp.params"))

    run_p = dict(TG64, Nx=32, Ny=32, Nz=32, vis=1e-3, n_steps=12, n_skip_h=4, i_les=4, pr_t=0.7,
                 mlp_labels=1, mlp_refresh=5, mlp_act=act)
    u0 = ic.taylor_green(32)
    sio.write_vel(str(tmp_path / "vel.000000"), u0)
    sio.write_header(str(tmp_path / "header.000000"), run_p, time=0.0, dt=0.01, nt=0)
    r = run(tmp_path, **run_p)
    assert r.returncode == 0, r.stdout + r.stderr

    les = S.Sam(S.make_params(32, vis=1e-3, dt0=0.01, sgs=S.SGS_MLP, pr_t=0.7))
    net = S.Mlp(seed=0)
    net.set_activation(act)
    net.set_labels(S.LABELS_DYNAMIC, 0.7)
    net.set_params(params)
    S.attach_mlp(les, net)
    S.set_mlp_refresh(les, 5)
    les.set_state(u0)
    les.step(12)
    ref = les.get_state()
    d = les.dims
    u = sio.read_vel(str(tmp_path / "vel.000012"), d.nz_min, d.ny_min, d.nx_min, d.lu)
    assert np.linalg.norm(ref - ic.taylor_green(32)) > 1e-6 * np.linalg.norm(ref)
    assert np.linalg.norm(u - ref) <= 1e-14 * np.linalg.norm(ref)


@pytest.mark.gpu
@needs_driver
def test_forcing_and_chollet_lesieur_run_matches_library(tmp_path):
    """i_force 1 / k_force 2 and i_les 1 / cl_couple 1 from SAM's run files:
    the state equals the library's with the same spectra cadence (the
    forcing target is the spectrum of every n_skip_h-th step)"""
    from sam_amd import samhip as S
    run_p = dict(TG64, Nx=32, Ny=32, Nz=32, vis=1e-3, n_steps=10, n_skip_h=5, i_les=1, c_smag=0.1,
                 cl_couple=1, i_force=1, k_force=2)
    u0 = ic.taylor_green(32)
    sio.write_vel(str(tmp_path / "vel.000000"), u0)
    sio.write_header(str(tmp_path / "header.000000"), run_p, time=0.0, dt=0.01, nt=0)
    r = run(tmp_path, **run_p)
    assert r.returncode == 0, r.stdout + r.stderr
    sam = S.Sam(S.make_params(32, vis=1e-3, dt0=0.01))
    S.set_cl(sam, S.CL_COUPLED)
    S.set_forcing(sam, 2)
    sam.set_state(u0)
    for nt in range(10):
        sam.step(1, want_spectra=(nt % 5 == 0))
    ref = sam.get_state()
    d = sam.dims
    u = sio.read_vel(str(tmp_path / "vel.000010"), d.nz_min, d.ny_min, d.nx_min, d.lu)
    assert np.linalg.norm(u - ref) <= 1e-14 * np.linalg.norm(ref)
    plain = S.Sam(S.make_params(32, vis=1e-3, dt0=0.01))
    plain.set_state(u0)
    plain.step(10)
    assert np.linalg.norm(plain.get_state() - ref) > 1e-8 * np.linalg.norm(ref)


@pytest.mark.gpu
@needs_driver
def test_planes_files(tmp_path):
    """n_skip_p 2 over 4 steps: frames at nt = 0, 2, 4 in time.out, mean.out
    and the xy/xz/yz files (write_planes.f layouts); the last frame equals
    the library's planes of the same state"""
    from sam_amd import samhip as S
    n = 32
    run_p = dict(TG64, Nx=n, Ny=n, Nz=n, vis=1e-3, n_steps=4, n_skip_h=2, n_skip_p=2, k_xy_plane1=5,
                 j_xz_plane1=3, i_yz_plane1=7, i_yz_plane2=20)
    u0 = ic.taylor_green(n)
    sio.write_vel(str(tmp_path / "vel.000000"), u0)
    sio.write_header(str(tmp_path / "header.000000"), run_p, time=0.0, dt=0.01, nt=0)
    r = run(tmp_path, **run_p)
    assert r.returncode == 0, r.stdout + r.stderr
    nf, nfr = 6, 3
    t = np.loadtxt(str(tmp_path / "time.out"))
    assert list(t[:, 0]) == [0, 2, 4] and abs(t[2, 1] - 0.04) < 1e-4
    mean = np.fromfile(str(tmp_path / "mean.out")).reshape(nfr, 3, n)
    xy = np.fromfile(str(tmp_path / "xy1.dat")).reshape(nfr, nf, n, n)
    xz = np.fromfile(str(tmp_path / "xz1.dat")).reshape(nfr, nf, n, n)
    yz2 = np.fromfile(str(tmp_path / "yz2.dat")).reshape(nfr, nf, n, n)
    assert not (tmp_path / "xy2.dat").exists()
    sam = S.Sam(S.make_params(n, vis=1e-3, dt0=0.01))
    S.set_planes(sam, True)
    sam.set_state(u0)
    sam.step(4)
    lm, lxy, lxz, lyz = S.planes(sam, (5,), (3,), (7, 20))
    # Fortran records are column-major: (Nx, Ny) per field -> reshape [.., y, x]
    assert np.allclose(xy[2].transpose(0, 2, 1), np.moveaxis(lxy[..., 0], 2, 0), rtol=0, atol=1e-14)
    assert np.allclose(xz[2].transpose(0, 2, 1), np.moveaxis(lxz[..., 0], 2, 0), rtol=0, atol=1e-14)
    assert np.allclose(yz2[2].transpose(0, 2, 1), np.moveaxis(lyz[..., 1], 2, 0), rtol=0, atol=1e-14)
    assert np.allclose(mean[2], lm.T, rtol=0, atol=1e-14)
    hdr = sio.parse_labels(str(tmp_path / "header.000004"))
    assert hdr["n_frame_p"] == 3
    # the VTK converter reads the driver's files
    from sam_amd import planes2vtk as P2V
    out = P2V.convert(str(tmp_path))
    assert len(out) == 4 * 3           # xy1, xz1, yz1, yz2 x 3 frames
    raw = open(str(tmp_path / "xy1_0002.vtk"), "rb").read()
    k = raw.index(b"Scalars u' float\nLOOKUP_TABLE default\n") + len(b"Scalars u' float\nLOOKUP_TABLE default\n")
    vals = np.frombuffer(raw[k:k + 4 * 33 * 33], dtype=">f4").reshape(33, 33)     # [y, x]
    assert np.allclose(vals[:32, :32], xy[2, 0], atol=1e-6)


@pytest.mark.gpu
@needs_driver
def test_planes_files_match_reference(tmp_path):
    """the driver's time.out, mean.out and xy/xz/yz planes files equal the
    reference's own files of the same run (a stratified 32^3 restart,
    n_skip_p 2, 4 steps; tests/golden/sam_planes.npz) at 1e-12, and its log
    files (energy, umax, tstep, variance, spectra) equal the reference's at
    their printed precision, and its header.000004 carries every label of
    the reference's with the same value"""
    import json
    g = np.load(os.path.join(ROOT, "tests", "golden", "sam_planes.npz"))
    p = json.loads(str(g["params_json"]))
    u0 = g["u0"]
    sio.write_vel(str(tmp_path / "vel.000000"), u0)
    sio.write_header(str(tmp_path / "header.000000"), p, time=0.0, dt=0.01, nt=0)
    r = run(tmp_path, **dict(p, n_steps=4))
    assert r.returncode == 0, r.stdout + r.stderr
    t = np.loadtxt(str(tmp_path / "time.out"))
    assert np.allclose(t, g["time"], rtol=0, atol=1e-8)
    for f in ("mean", "xy1", "xz1", "yz1", "yz2"):
        got = np.fromfile(str(tmp_path / (f + (".out" if f == "mean" else ".dat"))), dtype=np.float64)
        ref = g[f]
        assert got.shape == ref.shape, f
        assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max(), (f, np.abs(got - ref).max())
    # the log files (sam.f:252-271, spectra.f:92-111): same lines and layout,
    # numbers equal at their printed precision (5 significant digits)
    for f in ("energy", "umax", "tstep", "variance", "spectra"):
        got = open(str(tmp_path / (f + ".out"))).read().splitlines()
        ref = str(g["log_" + f]).splitlines()
        assert len(got) == len(ref), f
        for a, b in zip(got, ref):
            ta, tb = a.split(), b.split()
            assert len(ta) == len(tb) and len(a) == len(b), (f, a, b)
            for x, y in zip(ta, tb):
                try:
                    xv, yv = float(x), float(y)
                except ValueError:
                    assert x == y, (f, a, b)
                    continue
                assert abs(xv - yv) <= 2e-4 * abs(yv) + 1e-12, (f, a, b)
    # header.000004 (write_header.f): every label of the reference's, same value
    ref_h = tmp_path / "ref_header"
    ref_h.write_text(str(g["log_header"]))
    hr, hg = sio.parse_labels(str(ref_h)), sio.parse_labels(str(tmp_path / "header.000004"))
    missing = [k for k in hr if k not in hg]
    assert not missing, missing
    for k, v in hr.items():
        if isinstance(v, str) or isinstance(hg[k], str):
            assert str(hg[k]) == str(v), (k, hg[k], v)
        else:
            assert abs(hg[k] - v) <= 1e-12 * abs(v) + 1e-300, (k, hg[k], v)


@pytest.mark.gpu
@needs_driver
def test_mid_step_vel_files_match_reference(tmp_path):
    """vel_midstep 1: vel.NNNNNN at nt < nt_end holds what the reference
    writes inside a step (sam.f:277-284); the reference's own run with
    n_skip_v 2 (tests/golden/sam_runs.npz, midrk16) is the expected output"""
    import json
    g = np.load(os.path.join(ROOT, "tests", "golden", "sam_runs.npz"))
    p = dict(json.loads(str(g["cases_json"]))["midrk16"])
    for k in ("i_prob", "flct_u", "flct_t"):
        p.pop(k)
    p.update(i_restart=1, nt_restart=0, n_steps=4, n_skip_v=2, i_lapse=1, vel_midstep=1)
    u0 = g["midrk16_u0"]
    sio.write_vel(str(tmp_path / "vel.000000"), u0)
    sio.write_header(str(tmp_path / "header.000000"), p, time=0.0, dt=0.01, nt=0)
    r = run(tmp_path, **p)
    assert r.returncode == 0, r.stdout + r.stderr
    nzm, nym, nxm = u0.shape[:3]
    for nt in (2, 4):
        u = sio.read_vel(str(tmp_path / f"vel.{nt:06d}"), nzm, nym, nxm, 4)
        ref = g[f"midrk16_u{nt}"]
        assert np.linalg.norm(u - ref) / np.linalg.norm(ref) < 1e-12, nt
