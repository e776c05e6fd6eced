"""Write the golden fixtures of tests/golden/ from the reference's own code.

Runs in the development container only (it needs oracle/_ref/libsamref.so,
which `make -C oracle ref` compiles from /root/reference's fft.f, init.f,
fft_init.f, set_range.f, z_trans_b.f, z_trans_f.f, xy_trans_b.f,
xy_trans_f.f, x2z_decomp.f, z2x_decomp.f, vel_max.f, rhs.f, spectra.f and
div_check.f).  The fixtures are inputs and the reference routines' outputs
on them; tests compare the oracle and the GPU against them.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402


def fftpack_fixture():
    rng = np.random.default_rng(1985)
    out = {}
    for n in (8, 12, 16, 30, 64, 512):
        c = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        r = rng.standard_normal(n)
        out[f"c_in_{n}"] = c
        out[f"cfftf_{n}"] = O.Reference.cfft(c, -1)
        out[f"cfftb_{n}"] = O.Reference.cfft(c, +1)
        out[f"r_in_{n}"] = r
        out[f"rfftf_{n}"] = O.Reference.rfft(r, -1)
        out[f"rfftb_{n}"] = O.Reference.rfft(r, +1)
    np.savez_compressed(os.path.join(HERE, "fftpack.npz"), **out)


def transforms_fixture(name, nx, ny, nz, xl, yl, zl, n_dealias, i_strat, seed):
    R = O.Reference(nx, ny, nz, xl, yl, zl, n_dealias, i_strat)
    rng = np.random.default_rng(seed)
    nv = R.lu
    sh = (R.nze, R.nym, R.nxm, nv)
    u = np.asfortranarray(rng.standard_normal(sh) + 1j * rng.standard_normal(sh))
    u[R.nzm:] = 0.0
    zb = R.z_trans_b(u, nv)                                   # (Ny_min, Nx_min, Nz, n)
    # x2z on one rank: pad y, x modes above Nx_min are unused by xy_trans_b
    nxh = nx // 2
    y = np.zeros((ny, nxh, nz, nv), np.complex128, order="F")
    ky = R.kym
    y[: ky + 1, : R.nxm] = zb[: ky + 1]
    y[ny - ky:, : R.nxm] = zb[R.nym - ky:]
    phys = R.xy_trans_b(y, nv)                                # (Nx, Ny, Nz, n)
    p_in = np.asfortranarray(rng.standard_normal((nx, ny, nz, R.lr)))
    xyf = R.xy_trans_f(p_in, R.lr)                           # (Nz, Ny_min, Nx_min, Lr)
    zf = R.z_trans_f(xyf, R.lr)                               # (Nz_min, Ny_min, Nx_min, Lr)
    dims = np.array([nx, ny, nz, n_dealias, i_strat, R.kxm, R.kym, R.kzm, R.nxm, R.nym, R.nzm])
    np.savez_compressed(os.path.join(HERE, f"transforms_{name}.npz"), dims=dims,
                        box=np.array([xl, yl, zl]), u=u, z_trans_b=zb, x2z=y, xy_trans_b=phys,
                        p_in=p_in, xy_trans_f=xyf, z_trans_f=zf)


def rhs_spectra_fixture(name, n, i_strat, i_cfl, seed):
    """rhs.f (first substep, with vel_max.f), spectra.f and div_check.f of
    the reference on a seeded solenoidal state, spectra.f/div_check.f on a
    state with divergence (a random field), so that div_rms, div_max and
    n_div are not zero, and cl_energy.f + spectra.f with the
    Chollet-Lesieur model (i_les = 1) on the solenoidal state"""
    from sam_amd import ic
    cfg = O.make_cfg(n, i_strat=i_strat, grav=16.0, lapse0=1.0, vis=5e-3, i_cfl=i_cfl, cfl0=0.5, dt0=0.01)
    R = O.ReferencePhys(cfg)
    u = ic.random_solenoidal(n, k0=4, seed=seed, i_strat=i_strat)
    if i_strat:
        u[..., 3] = 0.1 * ic.random_solenoidal(n, k0=3, seed=seed + 1)[..., 0]
    r, st = R.rhs(u)
    sp = R.spectra(u, r)
    dv = R.div_check(u)
    rng = np.random.default_rng(seed + 2)
    sh = (R.nzm, R.nym, R.nxm, R.lu)
    w = np.asfortranarray(rng.standard_normal(sh) + 1j * rng.standard_normal(sh)) * 1e-2
    w[:, :, 0] = u[:, :, 0]          # the kx = 0 plane stays Hermitian
    spw = R.spectra(w, r)
    dvw = R.div_check(w)
    cl = R.cl_spectra(u, r)      # i_les = 1: cl_energy.f, then spectra.f with nu_t(k)
    keys = ("ek", "tk", "dk", "vt", "sample")
    out = dict(params=np.array([n, i_strat, i_cfl, seed]), u=u, r=r, stats=st, w=w)
    for tag, s_, d_ in (("", sp, dv), ("w_", spw, dvw)):
        for k in keys:
            out[f"{tag}{k}"] = s_[k]
        out[f"{tag}scalars"] = np.array([s_["energy_w"], s_["div_rms"], s_["div_max"], s_["n_div"]])
        out[f"{tag}div_check"] = np.array([d_["div_max"], d_["div_rms"], d_["n_div"]])
    for k in ("ek", "tk", "dk", "vt"):
        out[f"cl_{k}"] = cl[k]
    out["cl_energy"] = np.array([cl["e_cl"], cl["k_cl"]])
    np.savez_compressed(os.path.join(HERE, f"rhs_{name}.npz"), **out)


if __name__ == "__main__":
    if not O.ref_available():
        sys.exit("oracle/_ref/libsamref.so missing: run `make -C oracle ref` in the container with /root/reference")
    fftpack_fixture()
    transforms_fixture("16s", 16, 16, 16, 2 * np.pi, 2 * np.pi, 2 * np.pi, 1, 1, 7)
    transforms_fixture("16x8x32", 16, 8, 32, 3.0, 5.0, 7.0, 1, 0, 8)
    transforms_fixture("16d0", 16, 16, 16, 2 * np.pi, 2 * np.pi, 2 * np.pi, 0, 0, 9)
    rhs_spectra_fixture("16u", 16, 0, 0, 21)
    rhs_spectra_fixture("32s", 32, 1, 1, 22)
    print("golden fixtures written to", HERE)
