"""SAM's file formats from Python: input.dat, header.NNNNNN, vel.NNNNNN.

- input.dat / header.NNNNNN: one "label value" per line (read_line.f:24-89);
  header files are the list-directed dump of write_header.f:11-33.
- vel.NNNNNN: headerless native complex(8) records of (Nz_min, Ny_min), one
  per (field, kx) at byte offset ((n-1)*Nx_min + i-1)*Nz_min*Ny_min*16
  (write_field.f:20-32), i.e. the array u(Nz_min, Ny_min, Nx_min, Lu) in
  Fortran order.
"""
import os

import numpy as np

# set_labels.f:29-102 order (reals, integers, then the dynamic header block)
INPUT_LABELS = (
    ["xL", "yL", "zL", "c_smag", "dt0", "cfl0", "k_truncate", "z0", "to", "lapse0", "grav", "vis",
     "pr", "amplitude", "lambda_x", "Gam", "omega", "Uo", "shear", "shear_ratio", "flct_u", "flct_t",
     "lambda_z_fs", "Ri_fs",
     "Nx", "Ny", "Nz", "i_restart", "nt_restart", "i_strat", "i_les", "n_steps", "n_skip_h",
     "n_skip_p", "n_skip_v", "n_skip_s"]
    + [f"k_xy_plane{i}" for i in range(1, 10)]
    + [f"j_xz_plane{i}" for i in range(1, 10)]
    + [f"i_yz_plane{i}" for i in range(1, 10)]
    + ["i_stat", "i_cfl", "nrk_max", "n_dealias", "iubc_z", "i_prob", "i_force", "k_force", "i_fs"]
)
DYN_LABELS = ["time", "dt", "t_stat", "nt", "n_frame_p", "n_hist"]
DEFAULTS = {
    "c_smag": 0.01, "dt0": 0.01, "cfl0": 1.0, "k_truncate": 1e6, "z0": 0.0, "to": 1.0, "lapse0": 0.0,
    "grav": 0.0, "pr": 1.0, "shear": 1.0, "shear_ratio": 2.5, "n_skip_h": 1e9, "n_skip_p": 1e9,
    "n_skip_v": 1e9, "n_skip_s": 1e9, "i_cfl": 1, "nrk_max": 3, "n_dealias": 1, "i_prob": 2,
    "k_force": 2, "i_fs": 2,
}


def parse_labels(path):
    """label -> value, following read_line.f (comments '#', blank/eof ends,
    c/k/m multipliers, case-insensitive labels)."""
    out = {}
    with open(path) as fh:
        for line in fh:
            line = line.rstrip("\n")
            if line.startswith("#"):
                continue
            if not line.strip():
                break
            toks = line.split()
            label = toks[0].lower()
            if label == "eof":
                break
            num = toks[1]
            mult = {"c": 1e2, "k": 1e3, "m": 1e6}.get(num[-1], 1.0)
            if mult != 1.0:
                num = num[:-1]
            out[label] = float(num.replace("D", "E").replace("d", "e")) * mult
    return out


def write_input(path, **params):
    """input.dat with the given labels (others take SAM's defaults)."""
    with open(path, "w") as fh:
        for k, v in params.items():
            fh.write(f"{k:<14s}{v!r:>24}\n" if isinstance(v, float) else f"{k:<14s}{v:>24}\n")


def write_header(path, params, time, dt, nt, t_stat=0.0):
    """header.NNNNNN: every input label (defaults filled) + the dynamic block"""
    vals = dict(DEFAULTS)
    vals.update({k.lower(): v for k, v in params.items()})
    with open(path, "w") as fh:
        for lab in INPUT_LABELS:
            fh.write(f" {lab:<12s} {float(vals.get(lab.lower(), 0.0)):.17e}\n")
        for lab, v in zip(DYN_LABELS, (time, dt, t_stat, nt, 0, 0)):
            fh.write(f" {lab:<12s} {float(v):.17e}\n")


def write_vel(path, u):
    """u: (Nz_min, Ny_min, Nx_min, Lu) complex"""
    np.asfortranarray(u, dtype=np.complex128).ravel(order="F").tofile(path)


def read_vel(path, nzm, nym, nxm, lu):
    expect = 16 * nzm * nym * nxm * lu
    size = os.path.getsize(path)
    if size != expect:
        raise ValueError(f"vel file {path}: size {size}, expected {expect} (read_field.f:26-32)")
    return np.fromfile(path, dtype=np.complex128).reshape((nzm, nym, nxm, lu), order="F")


def read_log(path):
    """numeric rows of an (i6,1p11e12.4) log, '#' header lines skipped"""
    rows = []
    with open(path) as fh:
        for line in fh:
            if line.startswith("#") or not line.strip():
                continue
            rows.append([float(t) for t in line.split()])
    return np.array(rows)
