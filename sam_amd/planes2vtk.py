"""Legacy-VTK frames from SAM's planes files (planes2vtk.f).

Run in a run directory holding input.dat, time.out and the xy/xz/yz[m].dat
files that the driver writes:

    python -m sam_amd.planes2vtk [--start NT] [--skip K]

For every plane m of each kind and every frame of time.out with nt >= start
and nt a multiple of skip * n_skip_p it writes xy[m]_NNNN.vtk (NNNN =
nt / n_skip_p), and likewise for xz and yz, as planes2vtk.f:99-441 does:

- a BINARY rectilinear grid (big-endian float32) of (N1+1) x (N2+1) points
  on [-L/2, L/2], the plane's own coordinate at its position, the last
  row and column repeating the first (periodicity4);
- scalars u', v', w', [T'], p' and, from in-plane spectral derivatives
  (deriv1.f / deriv2.f: real FFT per line, kept modes k <= kmax),
  the vorticity component normal to the plane and the vorticity magnitude.
  The plane file's own derivative fields supply the out-of-plane terms.

The reference's rotation of xz planes into earth-fixed coordinates
(planes2vtk.f:70-82) needs the gravity-wave problem's lambda_z, which its
init.f derives for i_prob = 4; that problem setup is not built here.

Differences from planes2vtk.f, which reads time.out once for all planes of
a kind and skips the first time.out line before the xz and yz passes:
here every plane m is converted, and the xz and yz passes start at the
first frame like the xy pass.  The reference reads its frame range from
standard input; here it comes from the options (stop is unused there too).
"""
import argparse
import os
import sys

import numpy as np

from sam_amd import io as sio


def kmax_of(n, n_dealias, xdir):
    """init.f:49-62"""
    if n_dealias == 0 or (n_dealias == 2 and xdir):
        return n // 2 - 1
    return int((n - 0.1) / 3.0)


def deriv(a, axis, kmax, wave):
    """deriv1.f / deriv2.f: d/dx along `axis` of a periodic real plane, modes
    0 < k <= kmax (the mean, modes above kmax and the Nyquist dropped)"""
    n = a.shape[axis]
    f = np.fft.rfft(a, axis=axis)
    k = np.arange(f.shape[axis])
    mult = np.where((k >= 1) & (k <= kmax), 1j * wave * k, 0.0)
    shape = [1] * a.ndim
    shape[axis] = -1
    return np.fft.irfft(f * mult.reshape(shape), n=n, axis=axis)


def periodic(a):
    """periodicity4: append the first row and column"""
    a = np.concatenate([a, a[:1]], axis=0)
    return np.concatenate([a, a[:, :1]], axis=1)


def _f32be(a):
    return np.ascontiguousarray(a, dtype=">f4").tobytes(order="F")


def write_vtk(path, time, dims, coords, scalars):
    """legacy BINARY VTK rectilinear grid with the reference's header text
    ('(i8)' counts; a unit dimension written ' 1 '); coords = 3 arrays,
    scalars = [(label, (N1+1, N2+1) array)]"""
    nl = b"\n"
    i8 = lambda v: b"%8d" % v
    npts = int(np.prod(dims))
    out = [b"# vtk DataFile Version 2.0\nwritten by planes2vtk t =" + (b"%11.4E" % time) + b"\nBINARY\n"]
    out.append(b"DATASET RECTILINEAR_GRID\nDIMENSIONS" + b"".join(i8(d) if d > 1 else b" 1 " for d in dims) + nl)
    for name, c in zip((b"X", b"Y", b"Z"), coords):
        out.append(name + b"_COORDINATES " + (i8(len(c)) if len(c) > 1 else b"1") + b" float\n" + _f32be(c) + nl)
    out.append(b"POINT_DATA" + i8(npts) + nl)
    for label, data in scalars:
        out.append(b"Scalars " + label.encode() + b" float\nLOOKUP_TABLE default\n" + _f32be(data) + nl)
    with open(path, "wb") as fh:
        fh.write(b"".join(out))


def frames(time_path):
    """(nt, time) per time.out line ('#' lines skipped)"""
    out = []
    with open(time_path) as fh:
        for line in fh:
            if not line.strip() or line.lstrip().startswith("#"):
                continue
            t = line.split()
            out.append((int(t[0]), float(t[1].replace("D", "E"))))
    return out


def convert(run_dir=".", start=0, skip=1):
    """write the VTK frames; returns the list of files written"""
    p = dict(sio.DEFAULTS)
    p.update(sio.parse_labels(os.path.join(run_dir, "input.dat")))
    nx, ny, nz = int(p["nx"]), int(p["ny"]), int(p["nz"])
    xl, yl, zl = p["xl"], p["yl"], p["zl"]
    lu = 4 if int(p.get("i_strat", 0)) == 1 else 3
    nf = 2 * lu
    nd = int(p.get("n_dealias", 1))
    kxm, kym, kzm = kmax_of(nx, nd, True), kmax_of(ny, nd, False), kmax_of(nz, nd, False)
    wx, wy, wz = 2 * np.pi / xl, 2 * np.pi / yl, 2 * np.pi / zl
    n_skip_p = int(p.get("n_skip_p", 1e9))
    nt_restart = int(p.get("nt_restart", 0)) if int(p.get("i_restart", 0)) == 1 else 0
    step = skip * n_skip_p
    x = (-0.5 * xl + np.arange(nx + 1) * xl / nx).astype(np.float32)
    y = (-0.5 * yl + np.arange(ny + 1) * yl / ny).astype(np.float32)
    z = (-0.5 * zl + np.arange(nz + 1) * zl / nz).astype(np.float32)
    labels = ["u'", "v'", "w'"] + (["T'"] if lu == 4 else []) + ["p'"]
    planes = {kind: [int(p.get(f"{kind}_plane{m}", 0)) for m in range(1, 10)]
              for kind in ("k_xy", "j_xz", "i_yz")}
    if "n_skip_p" in sio.parse_labels(os.path.join(run_dir, "input.dat")):
        for kind, n in (("k_xy", nz), ("j_xz", ny), ("i_yz", nx)):
            if planes[kind][0] == 0:
                planes[kind][0] = n // 2                      # check_inputs.f:27-44
    kinds = (
        # name, positions, (N1, N2), (k1max, w1), (k2max, w2) of the plane's two directions
        ("xy", planes["k_xy"], (nx, ny), (kxm, wx), (kym, wy)),
        ("xz", planes["j_xz"], (nx, nz), (kxm, wx), (kzm, wz)),
        ("yz", planes["i_yz"], (ny, nz), (kym, wy), (kzm, wz)),
    )
    written = []
    for name, pos, (n1, n2), (k1, w1), (k2, w2) in kinds:
        for m, at in enumerate(pos, start=1):
            if at == 0:
                break
            path = os.path.join(run_dir, f"{name}{m}.dat")
            rec = n1 * n2 * nf
            data = np.fromfile(path, dtype=np.float64)
            for nt, t in frames(os.path.join(run_dir, "time.out")):
                if nt < start or nt % step:
                    continue
                fr = (nt - nt_restart) // n_skip_p
                if (fr + 1) * rec > data.size:
                    continue
                u = data[fr * rec:(fr + 1) * rec].reshape((n1, n2, nf), order="F")
                du = [deriv(u[:, :, q], ax, km, w) for q in range(3)
                      for ax, km, w in ((0, k1, w1), (1, k2, w2))]     # d1 u, d2 u, d1 v, d2 v, d1 w, d2 w
                e1, e2 = u[:, :, lu + 1], u[:, :, lu + 2]                # the file's out-of-plane derivatives
                if name == "xy":        # e1 = du/dz, e2 = dv/dz
                    wxv, wyv, wzv = du[5] - e2, -(du[4] - e1), du[2] - du[1]
                    normal = ("vort_z", wzv)
                elif name == "xz":      # e1 = du/dy, e2 = dw/dy
                    wxv, wyv, wzv = e2 - du[3], -(du[4] - du[1]), du[2] - e1
                    normal = ("vort_y", du[1] - du[4])
                else:                   # e1 = dv/dx, e2 = dw/dx
                    wxv, wyv, wzv = du[4] - du[3], -(e2 - du[1]), e1 - du[0]
                    normal = ("vort_x", wxv)
                scal = [(labels[q], periodic(u[:, :, q])) for q in range(lu + 1)]
                scal.append((normal[0], periodic(normal[1])))
                scal.append(("vort_mag", periodic(np.sqrt(wxv ** 2 + wyv ** 2 + wzv ** 2))))
                out = os.path.join(run_dir, f"{name}{m}_{nt // n_skip_p:04d}.vtk")
                if name == "xy":
                    write_vtk(out, t, (nx + 1, ny + 1, 1), (x, y, z[at - 1:at]), scal)
                elif name == "xz":
                    write_vtk(out, t, (nx + 1, 1, nz + 1), (x, y[at - 1:at], z), scal)
                else:
                    write_vtk(out, t, (1, ny + 1, nz + 1), (x[at - 1:at], y, z), scal)
                written.append(out)
    return written


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--start", type=int, default=0, help="first time step converted")
    ap.add_argument("--skip", type=int, default=1, help="convert every skip-th planes frame")
    ap.add_argument("--dir", default=".")
    a = ap.parse_args(argv)
    for f in convert(a.dir, a.start, a.skip):
        print(f)
    return 0


if __name__ == "__main__":
    sys.exit(main())
