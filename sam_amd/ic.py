"""Initial conditions in SAM's stored spectral form.

The state is u(Nz_min, Ny_min, Nx_min, Lu), complex, Fortran order, holding
Fourier-series coefficients of u(x) = sum_k u_k exp(i k.x) for kx >= 0 only
(the kx < 0 half follows from Hermitian symmetry), with the forward
normalisation 1/(Nx Ny Nz) of xy_trans_f.f:20 and z_trans_f.f:26.  Wavenumber
maps follow fft_init.f:52-72.

The reference's own IC generator (initial_field.f) uses the compiler's
random_number; configs here enter through vel.000000 files written from these
generators, as SAM's restart path (sam.f:102-105) reads them.
"""
import numpy as np


def kmax(n, n_dealias=1, xdir=False):
    """init.f:49-62"""
    if n_dealias == 0 or (n_dealias == 2 and xdir):
        return n // 2 - 1
    return int((float(n) - 0.1) / 3.0)


def grid_dims(nx, ny=None, nz=None, n_dealias=1, i_strat=0):
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    kxm = kmax(nx, n_dealias, True)
    kym = kmax(ny, n_dealias)
    kzm = kmax(nz, n_dealias)
    lu = 4 if i_strat else 3
    return dict(kxm=kxm, kym=kym, kzm=kzm, nxm=kxm + 1, nym=2 * kym + 1, nzm=2 * kzm + 1, lu=lu,
                lr=9 if i_strat else 6)


def wavenumbers(d):
    kx = np.arange(d["nxm"])
    ky = np.array([j if j <= d["kym"] else j - d["nym"] for j in range(d["nym"])])
    kz = np.array([k if k <= d["kzm"] else k - d["nzm"] for k in range(d["nzm"])])
    return kx, ky, kz


def taylor_green(nx, ny=None, nz=None, n_dealias=1, i_strat=0):
    """u = sin x cos y cos z, v = -cos x sin y cos z, w = 0 (T' = 0) on [0, 2pi)^3:
    u(1,+-1,+-1) = -i/8, v(1,ky=+-1,+-1) = +-i/8, everything else 0."""
    d = grid_dims(nx, ny, nz, n_dealias, i_strat)
    u = np.zeros((d["nzm"], d["nym"], d["nxm"], d["lu"]), np.complex128, order="F")
    for ky in (1, -1):
        for kz in (1, -1):
            iy = ky % d["nym"]
            iz = kz % d["nzm"]
            u[iz, iy, 1, 0] = -0.125j
            u[iz, iy, 1, 1] = 0.125j * ky
    return u


def random_solenoidal(nx, ny=None, nz=None, *, k0=4.0, urms=1.0, seed=20261203, n_dealias=1,
                      i_strat=0, xl=2 * np.pi, yl=None, zl=None):
    """Random-phase divergence-free velocity with E(k) ~ k^4 exp(-2 (k/k0)^2),
    rescaled to u_rms = urms (u_rms^2 = <u.u>/3); T' = 0.

    Amplitudes use the integer-shell magnitude |k| of the (kx,ky,kz) index
    (box 2pi); each mode gets |u_k|^2 = E(|k|)/(4 pi |k|^2) split over a
    random complex direction projected onto k-perp.  Kx = 0 modes are made
    exactly Hermitian (u(0,-ky,-kz) = conj u(0,ky,kz)) and the mean is 0.
    """
    d = grid_dims(nx, ny, nz, n_dealias, i_strat)
    kx, ky, kz = wavenumbers(d)
    yl = xl if yl is None else yl
    zl = xl if zl is None else zl
    KZ, KY, KX = np.meshgrid(kz, ky, kx, indexing="ij")   # shape (nzm, nym, nxm) == Fortran (k,j,i)
    rkx = KX * (2 * np.pi / xl)
    rky = KY * (2 * np.pi / yl)
    rkz = KZ * (2 * np.pi / zl)
    k2 = rkx ** 2 + rky ** 2 + rkz ** 2
    kmag = np.sqrt((KX ** 2 + KY ** 2 + KZ ** 2).astype(np.float64))
    rng = np.random.Generator(np.random.PCG64(seed))
    shape = kmag.shape
    a = rng.standard_normal((3,) + shape) + 1j * rng.standard_normal((3,) + shape)
    with np.errstate(divide="ignore", invalid="ignore"):
        kdota = (rkx * a[0] + rky * a[1] + rkz * a[2]) / np.where(k2 > 0, k2, 1.0)
        a[0] -= rkx * kdota
        a[1] -= rky * kdota
        a[2] -= rkz * kdota
        amp2 = np.where(kmag > 0, kmag ** 4 * np.exp(-2.0 * (kmag / k0) ** 2) / (4 * np.pi * kmag ** 2), 0.0)
        nrm = np.sqrt(np.sum(np.abs(a) ** 2, axis=0))
        scale = np.where(nrm > 0, np.sqrt(amp2) / np.where(nrm > 0, nrm, 1.0), 0.0)
    a *= scale[None]
    # Hermitian symmetry on the kx = 0 plane: mode (ky,kz) <-> (-ky,-kz)
    p = a[:, :, :, 0]
    nym, nzm = d["nym"], d["nzm"]
    iy = np.arange(nym)
    iz = np.arange(nzm)
    miy = (-ky) % nym
    miz = (-kz) % nzm
    keep = (ky[None, :] > 0) | ((ky[None, :] == 0) & (kz[:, None] > 0))   # (nzm, nym)
    mirror = np.conj(p[:, miz][:, :, miy])
    p[:] = np.where(keep[None], p, mirror)
    p[:, 0, 0] = 0.0
    a[:, :, :, 0] = p
    # rescale to u_rms
    factor = np.where(KX > 0, 2.0, 1.0)
    usq = np.sum(factor[None] * np.abs(a) ** 2)
    a *= urms / np.sqrt(usq / 3.0)
    u = np.zeros((d["nzm"], d["nym"], d["nxm"], d["lu"]), np.complex128, order="F")
    for c in range(3):
        u[:, :, :, c] = a[c]
    return u


def spectral_pad(u, n_to, n_dealias=1):
    """Place the modes of u (SAM's stored form on a coarser grid) into the
    stored form of an n_to^3 grid, zero elsewhere -- the refinement direction
    of remesh.f:97-178."""
    nzm_f, nym_f, nxm_f, lu = u.shape
    d = grid_dims(n_to, n_dealias=n_dealias, i_strat=1 if lu == 4 else 0)
    kyf, kzf = (nym_f - 1) // 2, (nzm_f - 1) // 2
    if d["kym"] < kyf or d["kzm"] < kzf or d["nxm"] < nxm_f:
        raise ValueError("target grid is coarser than the source")
    out = np.zeros((d["nzm"], d["nym"], d["nxm"], lu), np.complex128, order="F")
    jy = np.array([j if j <= kyf else j - nym_f for j in range(nym_f)]) % d["nym"]
    jz = np.array([k if k <= kzf else k - nzm_f for k in range(nzm_f)]) % d["nzm"]
    out[np.ix_(jz, jy, np.arange(nxm_f), np.arange(lu))] = u
    return out

