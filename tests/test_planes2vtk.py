"""The VTK converter of the planes files (sam_amd/planes2vtk.py, after
planes2vtk.f) -- CPU only.

Plane files are written in the driver's layout from analytic single-mode
fields; the converted frames are parsed back and every scalar (the fields,
the normal vorticity component and the vorticity magnitude) is compared with
its closed form at the grid points (float32 output)."""
import os

import numpy as np

from sam_amd import io as sio
from sam_amd import planes2vtk as P2V

N = 16
L = 2 * np.pi
# (amplitude, phase, kx, ky, kz) per field: u v w T p
MODES = [(1.0, 0.0, 2, 3, 1), (0.5, 0.3, 1, -2, 2), (0.7, -0.2, 3, 1, -1), (0.4, 1.1, 1, 1, 3),
         (0.6, 0.5, 2, -1, 1)]


def val(q, x, y, z, d=None, scale=1.0):
    a, ph, kx, ky, kz = MODES[q]
    arg = kx * x + ky * y + kz * z + ph
    if d is None:
        return scale * a * np.cos(arg)
    k = {"x": kx, "y": ky, "z": kz}[d]
    return -scale * a * k * np.sin(arg)


def vorticity(x, y, z, s):
    wx = val(2, x, y, z, "y", s) - val(1, x, y, z, "z", s)
    wy = -(val(2, x, y, z, "x", s) - val(0, x, y, z, "z", s))
    wz = val(1, x, y, z, "x", s) - val(0, x, y, z, "y", s)
    return wx, wy, wz


def read_vtk(path):
    raw = open(path, "rb").read()
    head_end = raw.index(b"POINT_DATA")
    npts = int(raw[head_end:raw.index(b"\n", head_end)].split()[1])
    out, pos = {}, raw.index(b"\n", head_end) + 1
    while pos < len(raw):
        line_end = raw.index(b"\n", pos)
        name = raw[pos:line_end].split()[1].decode()
        pos = raw.index(b"\n", line_end + 1) + 1          # LOOKUP_TABLE line
        out[name] = np.frombuffer(raw[pos:pos + 4 * npts], dtype=">f4").astype(np.float64)
        pos += 4 * npts + 1
    return raw[:head_end], out


def test_converted_frames_match_closed_forms(tmp_path):
    g = 2 * np.pi * np.arange(N) / N
    pos = dict(k_xy_plane1=5, j_xz_plane1=3, i_yz_plane1=7, i_yz_plane2=12)
    sio.write_input(str(tmp_path / "input.dat"), Nx=N, Ny=N, Nz=N, xL=L, yL=L, zL=L, vis=1e-3, n_steps=4,
                    i_strat=1, i_restart=1, nt_restart=0, n_skip_p=2, **pos)
    scales = [1.0, 0.5]                    # frame f holds the fields times scales[f]
    (tmp_path / "time.out").write_text("     0  0.0000E+00\n     2  2.0000E-02\n")
    A, B = np.meshgrid(g, g, indexing="ij")
    kinds = {
        "xy1": lambda s: [val(q, A, B, g[4], None, s) for q in range(5)]
                         + [val(q, A, B, g[4], "z", s) for q in (0, 1, 3)],
        "xz1": lambda s: [val(q, A, g[2], B, None, s) for q in range(5)]
                         + [val(q, A, g[2], B, "y", s) for q in (0, 2, 3)],
        "yz1": lambda s: [val(q, g[6], A, B, None, s) for q in range(5)]
                         + [val(q, g[6], A, B, "x", s) for q in (1, 2, 3)],
        "yz2": lambda s: [val(q, g[11], A, B, None, s) for q in range(5)]
                         + [val(q, g[11], A, B, "x", s) for q in (1, 2, 3)],
    }
    for name, fn in kinds.items():
        frames = [np.stack(fn(s), axis=2) for s in scales]
        np.concatenate([f.reshape(-1, order="F") for f in frames]).tofile(str(tmp_path / f"{name}.dat"))

    written = P2V.convert(str(tmp_path))
    assert sorted(os.path.basename(w) for w in written) == sorted(
        f"{k}_{f:04d}.vtk" for k in kinds for f in (0, 1))

    def grid(a):            # (N, N) -> periodic (N+1, N+1), flattened first index fastest
        return P2V.periodic(a).reshape(-1, order="F")

    for f, s in enumerate(scales):
        head, sc = read_vtk(str(tmp_path / f"xy1_{f:04d}.vtk"))
        assert b"DIMENSIONS      17      17 1 \n" in head and b"Z_COORDINATES 1 float" in head
        assert b"t = %s" % (b"%10.4E" % (0.02 * f)) in head
        for q, lab in enumerate(["u'", "v'", "w'", "T'", "p'"]):
            assert np.abs(sc[lab] - grid(val(q, A, B, g[4], None, s))).max() < 2e-6, (f, lab)
        wx, wy, wz = vorticity(A, B, g[4], s)
        assert np.abs(sc["vort_z"] - grid(wz)).max() < 2e-5
        assert np.abs(sc["vort_mag"] - grid(np.sqrt(wx ** 2 + wy ** 2 + wz ** 2))).max() < 2e-5

        head, sc = read_vtk(str(tmp_path / f"xz1_{f:04d}.vtk"))
        assert b"DIMENSIONS      17 1       17\n" in head
        wx, wy, wz = vorticity(A, g[2], B, s)
        assert np.abs(sc["w'"] - grid(val(2, A, g[2], B, None, s))).max() < 2e-6
        assert np.abs(sc["vort_y"] - grid(wy)).max() < 2e-5
        assert np.abs(sc["vort_mag"] - grid(np.sqrt(wx ** 2 + wy ** 2 + wz ** 2))).max() < 2e-5

        for name, i in (("yz1", 6), ("yz2", 11)):
            head, sc = read_vtk(str(tmp_path / f"{name}_{f:04d}.vtk"))
            assert b"DIMENSIONS 1       17      17\n" in head
            wx, wy, wz = vorticity(g[i], A, B, s)
            assert np.abs(sc["p'"] - grid(val(4, g[i], A, B, None, s))).max() < 2e-6
            assert np.abs(sc["vort_x"] - grid(wx)).max() < 2e-5
            assert np.abs(sc["vort_mag"] - grid(np.sqrt(wx ** 2 + wy ** 2 + wz ** 2))).max() < 2e-5

    # --start / --skip select frames as planes2vtk.f does (nt >= start, nt % (skip n_skip_p) == 0)
    for w in written:
        os.remove(w)
    assert sorted(os.path.basename(w) for w in P2V.convert(str(tmp_path), start=1)) == sorted(
        f"{k}_0001.vtk" for k in kinds)
    assert P2V.convert(str(tmp_path), skip=2) and all("_0000" in w for w in P2V.convert(str(tmp_path), skip=2))
