"""ctypes front end of the CPU oracle (``libsam_oracle.so``) and of the
reference harness (``_ref/libsamref.so``).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()``
and ``bench.py``'s cpu_baseline leg, never by the product package.

Arrays follow the reference's Fortran layouts.  numpy arrays are created in
Fortran order so that ``a[k, j, i, n]`` is the reference's ``a(k+1,j+1,i+1,n+1)``.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsam_oracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libsamref.so")

_lib = None
_ref = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


class OrcCfg(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("n_dealias", ctypes.c_int), ("i_strat", ctypes.c_int),
        ("nrk_max", ctypes.c_int), ("i_cfl", ctypes.c_int), ("sgs", ctypes.c_int),
        ("xl", ctypes.c_double), ("yl", ctypes.c_double), ("zl", ctypes.c_double),
        ("vis", ctypes.c_double), ("pr", ctypes.c_double), ("grav", ctypes.c_double),
        ("to", ctypes.c_double), ("lapse0", ctypes.c_double), ("lapse", ctypes.c_double),
        ("cfl0", ctypes.c_double), ("dt0", ctypes.c_double),
        ("c_smag", ctypes.c_double), ("pr_t", ctypes.c_double),
        ("k_force", ctypes.c_int), ("i_cl", ctypes.c_int),
    ]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = ctypes.CDLL(LIB_PATH)
        L.orc_create.restype = ctypes.c_void_p
        L.orc_create.argtypes = [ctypes.POINTER(OrcCfg)]
        L.orc_destroy.argtypes = [ctypes.c_void_p]
        L.orc_dims.argtypes = [ctypes.c_void_p, _ip]
        L.orc_consts.argtypes = [ctypes.c_void_p, _dp]
        L.orc_set_threads.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_cfft.argtypes = [ctypes.c_int, _dp, ctypes.c_int]
        L.orc_rfft.argtypes = [ctypes.c_int, _dp, ctypes.c_int]
        L.orc_z_trans_b.argtypes = [ctypes.c_void_p, _dp, ctypes.c_int, _dp, ctypes.c_int]
        L.orc_x2z.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_int]
        L.orc_xy_trans_b.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_int]
        L.orc_xy_trans_f.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_int]
        L.orc_z_trans_f.argtypes = [ctypes.c_void_p, _dp, ctypes.c_int, ctypes.c_int]
        L.orc_vel_max.argtypes = [ctypes.c_void_p, _dp, ctypes.c_double, _dp]
        L.orc_products.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.orc_project.argtypes = [ctypes.c_void_p, _dp, ctypes.c_int, _dp, _dp]
        L.orc_test_filter.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.orc_eddy_visc.argtypes = [ctypes.c_void_p, _dp, _dp, _dp]
        L.orc_sgs_products.argtypes = [ctypes.c_void_p, _dp, _dp, _dp]
        L.orc_to_physical.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.orc_rhs.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_double, _dp]
        L.orc_step.argtypes = [ctypes.c_void_p, _dp, ctypes.c_int, _dp, _dp]
        L.orc_step_ext.argtypes = [ctypes.c_void_p, _dp, ctypes.c_int, _dp, _dp, _dp, _dp]
        L.orc_planes.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_int, _ip, ctypes.c_int, _ip, ctypes.c_int, _ip,
                                 _dp, _dp, _dp, _dp]
        L.orc_spectra.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_double, _dp]
        L.orc_cl_energy.restype = ctypes.c_double
        L.orc_cl_energy.argtypes = [ctypes.c_void_p, _dp]
        L.orc_cl_fac.restype = ctypes.c_double
        L.orc_cl_fac.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.orc_force.argtypes = [ctypes.c_void_p, _dp, _dp, _dp]
        L.orc_div_check.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.orc_rk_update.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, _dp, _dp, _dp]
        _lib = L
    return _lib


def ref_available():
    return os.path.exists(REF_PATH)


def ref():
    global _ref
    if _ref is None:
        if not ref_available():
            raise RuntimeError("reference harness not built: run `make -C oracle ref`")
        R = ctypes.CDLL(REF_PATH)
        R.ref_setup.argtypes = [ctypes.c_int] * 3 + [ctypes.c_double] * 3 + [ctypes.c_int] * 2 + [_ip]
        for name in ("ref_z_trans_b", "ref_xy_trans_b", "ref_xy_trans_f"):
            getattr(R, name).argtypes = [_dp, _dp, ctypes.c_int]
        R.ref_z_trans_f.argtypes = [_dp, ctypes.c_int]
        R.ref_rfft.argtypes = [ctypes.c_int, _dp, ctypes.c_int]
        R.ref_cfft.argtypes = [ctypes.c_int, _dp, ctypes.c_int]
        R.ref_setup_phys.argtypes = [_ip, _dp, _ip]
        R.ref_rhs.argtypes = [_dp, _dp, _dp]
        R.ref_vel_max.argtypes = [_dp, ctypes.c_double, _dp]
        R.ref_spectra.argtypes = [_dp, _dp, _dp]
        R.ref_div_check.argtypes = [_dp, _dp]
        R.ref_cl_spectra.argtypes = [_dp, _dp, _dp]
        _ref = R
    return _ref


def _p(a):
    assert a.flags["F_CONTIGUOUS"] or a.flags["C_CONTIGUOUS"]
    if a.dtype == np.complex128:
        return a.ctypes.data_as(_dp)
    assert a.dtype == np.float64, a.dtype
    return a.ctypes.data_as(_dp)


def make_cfg(nx, ny=None, nz=None, *, xl=2 * np.pi, yl=None, zl=None, n_dealias=1,
             i_strat=0, nrk_max=3, i_cfl=0, sgs=0, vis=0.0025, pr=1.0, grav=0.0,
             to=1.0, lapse0=0.0, lapse=None, cfl0=0.5, dt0=0.01, c_smag=0.01, pr_t=1.0,
             k_force=-1, i_cl=0):
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    yl = xl if yl is None else yl
    zl = xl if zl is None else zl
    lapse = lapse0 if lapse is None else lapse
    return OrcCfg(nx, ny, nz, n_dealias, i_strat, nrk_max, i_cfl, sgs, xl, yl, zl,
                  vis, pr, grav, to, lapse0, lapse, cfl0, dt0, c_smag, pr_t, k_force, i_cl)


class Oracle:
    """One oracle instance (one grid / parameter set)."""

    def __init__(self, cfg: OrcCfg, threads: int = 0):
        self.L = lib()
        self.cfg = cfg
        self.h = self.L.orc_create(ctypes.byref(cfg))
        if threads:
            self.L.orc_set_threads(self.h, threads)
        d = (ctypes.c_int * 10)()
        self.L.orc_dims(self.h, d)
        (self.kxm, self.kym, self.kzm, self.nxm, self.nym, self.nzm,
         self.lu, self.lr, self.lm, self.nxh) = list(d)
        c = (ctypes.c_double * 13)()
        self.L.orc_consts(self.h, c)
        self.wave = (c[0], c[1], c[2])
        self.gamma = tuple(c[3:6])
        self.zeta = tuple(c[6:9])
        self.adv_cfl_limit = c[9]
        self.buoy_fac_z = c[10]
        self.delta_sq = c[12]
        self.nx, self.ny, self.nz = cfg.nx, cfg.ny, cfg.nz

    def __del__(self):
        try:
            self.L.orc_destroy(self.h)
        except Exception:
            pass

    # ---- shapes -------------------------------------------------------
    def spec_shape(self, nvar=None, ldz=None):
        return (ldz or self.nzm, self.nym, self.nxm, nvar or self.lu)

    def zeros_spec(self, nvar=None, ldz=None):
        return np.zeros(self.spec_shape(nvar, ldz), dtype=np.complex128, order="F")

    def phys_shape(self, nvar):
        return (self.nx, self.ny, self.nz, nvar)

    def kvectors(self):
        """integer wavenumber maps k_x, k_y, k_z (fft_init.f:52-72)"""
        kx = np.arange(self.nxm)
        ky = np.array([j if j <= self.kym else j - self.nym for j in range(self.nym)])
        kz = np.array([k if k <= self.kzm else k - self.nzm for k in range(self.nzm)])
        return kx, ky, kz

    # ---- stages -------------------------------------------------------
    def z_trans_b(self, u, nvar=None):
        nvar = nvar or u.shape[3]
        u = np.asfortranarray(u)
        out = np.zeros((self.nym, self.nxm, self.nz, nvar), np.complex128, order="F")
        self.L.orc_z_trans_b(self.h, _p(u), u.shape[0], _p(out), nvar)
        return out

    def x2z(self, a):
        nvar = a.shape[3]
        a = np.asfortranarray(a)
        out = np.zeros((self.ny, self.nxh, self.nz, nvar), np.complex128, order="F")
        self.L.orc_x2z(self.h, _p(a), _p(out), nvar)
        return out

    def xy_trans_b(self, a):
        nvar = a.shape[3]
        a = np.array(a, order="F", copy=True)
        out = np.zeros((self.nx, self.ny, self.nz, nvar), np.float64, order="F")
        self.L.orc_xy_trans_b(self.h, _p(a), _p(out), nvar)
        return out

    def xy_trans_f(self, p):
        nvar = p.shape[3]
        p = np.array(p, order="F", copy=True)
        out = np.zeros((self.nz, self.nym, self.nxm, nvar), np.complex128, order="F")
        self.L.orc_xy_trans_f(self.h, _p(p), _p(out), nvar)
        return out

    def z_trans_f(self, a):
        """a: (Nz,Ny_min,Nx_min,n) -> (Nz_min,Ny_min,Nx_min,n)"""
        nvar = a.shape[3]
        a = np.array(a, order="F", copy=True)
        self.L.orc_z_trans_f(self.h, _p(a), a.shape[0], nvar)
        return np.asfortranarray(a[: self.nzm])

    def to_physical(self, u):
        u = np.asfortranarray(u)
        out = np.zeros(self.phys_shape(self.lu), np.float64, order="F")
        self.L.orc_to_physical(self.h, _p(u), _p(out))
        return out

    def vel_max(self, phys, dt_in=None):
        st = np.zeros(17)
        phys = np.asfortranarray(phys)
        self.L.orc_vel_max(self.h, _p(phys), self.cfg.dt0 if dt_in is None else dt_in, _p(st))
        return st

    def products(self, phys):
        phys = np.asfortranarray(phys)
        out = np.zeros(self.phys_shape(self.lr), np.float64, order="F")
        self.L.orc_products(self.h, _p(phys), _p(out))
        return out

    def project(self, F, usav):
        F = np.asfortranarray(F)
        usav = np.asfortranarray(usav)
        r = self.zeros_spec(self.lu + 1)
        self.L.orc_project(self.h, _p(F), F.shape[0], _p(usav), _p(r))
        return r

    def rhs(self, u, dt_in=None):
        u = np.asfortranarray(u)
        r = self.zeros_spec(self.lu + 1)
        st = np.zeros(18)
        self.L.orc_rhs(self.h, _p(u), _p(r), self.cfg.dt0 if dt_in is None else dt_in, _p(st))
        return r, st

    def step(self, u, nsteps, time=0.0, dt=None, ek0=None, want_p=False):
        """nsteps RK steps; with cfg.i_cl the Chollet-Lesieur energy of every
        step (st[18] = fac_cl of the last one), with cfg.k_force >= 0 and ek0
        given, force.f after every step toward ek0.  want_p: also return the
        pressure of the last substep (S layout)."""
        u = np.array(u, order="F", copy=True)
        state = np.array([time, self.cfg.dt0 if dt is None else dt])
        st = np.zeros(19)
        if ek0 is None and not want_p:
            self.L.orc_step(self.h, _p(u), nsteps, _p(state), _p(st))
            return u, state[0], state[1], st
        e = np.ascontiguousarray(ek0 if ek0 is not None else np.zeros(self.lm), dtype=np.float64)
        p = self.zeros_spec(1)
        force = ek0 is not None and self.cfg.k_force >= 0
        self.L.orc_step_ext(self.h, _p(u), nsteps, _p(state), _p(st), _p(e) if force else None, _p(p))
        if want_p:
            return u, state[0], state[1], st, p[..., 0]
        return u, state[0], state[1], st

    def planes(self, u, p, k_xy=(), j_xz=(), i_yz=()):
        """write_planes.f / get_mean.f of state u with pressure p (S layout):
        mean (Nz, Lu); xy (Nx, Ny, 2Lu, n_xy); xz (Nx, Nz, 2Lu, n_xz);
        yz (Ny, Nz, 2Lu, n_yz); positions 1-based"""
        u = np.asfortranarray(u)
        p = np.asfortranarray(p.reshape(p.shape[:3]) if p.ndim == 4 else p)
        nf = 2 * self.lu
        ia = lambda v: (ctypes.c_int * max(len(v), 1))(*v)
        mean = np.zeros((self.nz, self.lu), order="F")
        xy = np.zeros((self.nx, self.ny, nf, max(len(k_xy), 1)), order="F")
        xz = np.zeros((self.nx, self.nz, nf, max(len(j_xz), 1)), order="F")
        yz = np.zeros((self.ny, self.nz, nf, max(len(i_yz), 1)), order="F")
        self.L.orc_planes(self.h, _p(u), _p(p), len(k_xy), ia(k_xy), len(j_xz), ia(j_xz), len(i_yz), ia(i_yz),
                          _p(mean), _p(xy), _p(xz), _p(yz))
        return mean, xy[..., :len(k_xy)], xz[..., :len(j_xz)], yz[..., :len(i_yz)]

    def cl_energy(self, u):
        """E_cl of cl_energy.f and fac_cl = sqrt(2 E_cl / k_cl)"""
        u = np.asfortranarray(u)
        e = self.L.orc_cl_energy(self.h, _p(u))
        return e, self.L.orc_cl_fac(self.h, e)

    def force(self, u, ek0):
        """force.f on a copy of u; returns (u, shell energies before)"""
        u = np.array(u, order="F", copy=True)
        ek0 = np.ascontiguousarray(ek0, dtype=np.float64)
        ekg = np.zeros(max(self.cfg.k_force, 0) + 2)
        self.L.orc_force(self.h, _p(u), _p(ek0), _p(ekg))
        return u, ekg[: self.cfg.k_force + 1]

    def spectra(self, u, r, fac_cl=0.0):
        u = np.asfortranarray(u)
        r = np.asfortranarray(r)
        out = np.zeros(5 * self.lm + 3)
        self.L.orc_spectra(self.h, _p(u), _p(r), fac_cl, _p(out))
        lm = self.lm
        return dict(ek=out[:lm], tk=out[lm:2 * lm], dk=out[2 * lm:3 * lm],
                    vt=out[3 * lm:4 * lm], sample=out[4 * lm:5 * lm],
                    energy_w=out[5 * lm], div_rms=out[5 * lm + 1], div_max=out[5 * lm + 2])

    def div_check(self, u):
        """div_check.f: div_max, div_rms, n_div"""
        u = np.asfortranarray(u)
        out = np.zeros(3)
        self.L.orc_div_check(self.h, _p(u), _p(out))
        return dict(div_max=out[0], div_rms=out[1], n_div=out[2])

    def rk_update(self, nrk, dt, u, r_prev=None, r=None):
        """sam.f:210-221 (r_prev) and 293-314 (r) on a copy of u (S layout)"""
        u = np.array(u, order="F", copy=True)
        rp = None if r_prev is None else np.asfortranarray(r_prev[..., : self.lu])
        rr = None if r is None else np.asfortranarray(r[..., : self.lu])
        self.L.orc_rk_update(self.h, nrk, dt, _p(u), _p(rp) if rp is not None else None,
                             _p(rr) if rr is not None else None)
        return u

    def eddy_visc(self, phys):
        phys = np.asfortranarray(phys)
        vt = np.zeros((self.nx, self.ny, self.nz), order="F")
        cd = np.zeros(1)
        self.L.orc_eddy_visc(self.h, _p(phys), _p(vt), _p(cd))
        return vt, cd[0]

    def test_filter(self, f):
        f = np.asfortranarray(f, dtype=np.float64)
        out = np.zeros_like(f, order="F")
        self.L.orc_test_filter(self.h, _p(f), _p(out))
        return out


def cfft(x, sign):
    x = np.array(x, dtype=np.complex128, copy=True)
    lib().orc_cfft(len(x), _p(x), sign)
    return x


def rfft_pk(r, direction):
    r = np.array(r, dtype=np.float64, copy=True)
    lib().orc_rfft(len(r), _p(r), direction)
    return r


class Reference:
    """The reference's own transform routines (oracle/_ref), one grid."""

    def __init__(self, nx, ny, nz, xl, yl, zl, n_dealias=1, i_strat=0):
        self.R = ref()
        d = (ctypes.c_int * 10)()
        self.R.ref_setup(nx, ny, nz, xl, yl, zl, n_dealias, i_strat, d)
        (self.kxm, self.kym, self.kzm, self.nxm, self.nym, self.nzm,
         self.nze, self.lu, self.lr, ierr) = list(d)
        self.nx, self.ny, self.nz = nx, ny, nz

    def z_trans_b(self, u, nvar):
        """u: (Nze,Ny_min,Nx_min,nvar) -> (Ny_min,Nx_min,Nz,nvar)"""
        u = np.array(u, order="F", copy=True)
        assert u.shape[0] == self.nze
        out = np.zeros((self.nym, self.nxm, self.nz, nvar), np.complex128, order="F")
        self.R.ref_z_trans_b(_p(u), _p(out), nvar)
        return out

    def xy_trans_b(self, a, nvar):
        """a: (Ny,Nx/2,Nz,nvar) complex -> (Nx,Ny,Nz,nvar) real"""
        a = np.array(a, order="F", copy=True)
        out = np.zeros((self.nx, self.ny, self.nz, nvar), np.float64, order="F")
        self.R.ref_xy_trans_b(_p(a), _p(out), nvar)
        return out

    def xy_trans_f(self, p, nvar):
        """p: (Nx,Ny,Nz,nvar) real -> (Nz,Ny_min,Nx_min,nvar) complex"""
        p = np.array(p, order="F", copy=True)
        out = np.zeros((self.nz, self.nym, self.nxm, nvar), np.complex128, order="F")
        self.R.ref_xy_trans_f(_p(p), _p(out), nvar)
        return out

    def z_trans_f(self, a, nvar):
        """a: (Nze,Ny_min,Nx_min,nvar), in place; returns the first Nz_min rows"""
        a = np.array(a, order="F", copy=True)
        self.R.ref_z_trans_f(_p(a), nvar)
        return np.asfortranarray(a[: self.nzm])

    @staticmethod
    def rfft(r, direction):
        r = np.array(r, dtype=np.float64, copy=True)
        ref().ref_rfft(len(r), _p(r), direction)
        return r

    @staticmethod
    def cfft(c, direction):
        c = np.array(c, dtype=np.complex128, copy=True)
        ref().ref_cfft(len(c), _p(c), direction)
        return c


class ReferencePhys:
    """The reference's rhs.f, vel_max.f, spectra.f and div_check.f
    (oracle/_ref, linked with MPICH, one rank) for one grid and parameter
    set.  The reference keeps its state in COMMON blocks, so only the most
    recently set up instance is valid."""

    def __init__(self, cfg: OrcCfg):
        self.R = ref()
        ip = (ctypes.c_int * 7)(cfg.nx, cfg.ny, cfg.nz, cfg.n_dealias, cfg.i_strat, cfg.nrk_max, cfg.i_cfl)
        rp = (ctypes.c_double * 10)(cfg.xl, cfg.yl, cfg.zl, cfg.vis, cfg.pr, cfg.grav, cfg.to, cfg.lapse,
                                    cfg.cfl0, cfg.dt0)
        d = (ctypes.c_int * 10)()
        self.R.ref_setup_phys(ip, rp, d)
        (self.kxm, self.kym, self.kzm, self.nxm, self.nym, self.nzm,
         self.nze, self.lu, self.lr, self.lm) = list(d)
        self.nx, self.ny, self.nz = cfg.nx, cfg.ny, cfg.nz

    def _full(self, u, nvar):
        """(Nz_min or Nze, Ny_min, Nx_min, nvar) -> (Nze, ...) zero padded"""
        out = np.zeros((self.nze, self.nym, self.nxm, nvar), np.complex128, order="F")
        out[: min(u.shape[0], self.nze)] = u[: self.nze, ..., :nvar]
        return out

    def rhs(self, u):
        """sam.f's first-substep RHS: r (Nz_min,Ny_min,Nx_min,Lu+1) with the
        pressure last, and the vel_max statistics st[0:17]"""
        uf = self._full(u, self.lu)
        r = np.zeros((self.nze, self.nym, self.nxm, self.lu + 1), np.complex128, order="F")
        st = np.zeros(17)
        self.R.ref_rhs(_p(uf), _p(r), _p(st))
        return np.asfortranarray(r[: self.nzm]), st

    def vel_max(self, phys, dt_in):
        phys = np.asfortranarray(phys, dtype=np.float64)
        st = np.zeros(17)
        self.R.ref_vel_max(_p(phys), dt_in, _p(st))
        return st

    def spectra(self, u, r):
        uf = self._full(u, self.lu)
        rf = self._full(r, self.lr) if r.shape[3] >= self.lr else self._full(
            np.concatenate([r, np.zeros(r.shape[:3] + (self.lr - r.shape[3],), np.complex128)], axis=3), self.lr)
        lm = self.lm
        out = np.zeros(5 * lm + 4)
        self.R.ref_spectra(_p(uf), _p(rf), _p(out))
        return dict(ek=out[:lm], tk=out[lm:2 * lm], dk=out[2 * lm:3 * lm], vt=out[3 * lm:4 * lm],
                    sample=out[4 * lm:5 * lm], energy_w=out[5 * lm], div_rms=out[5 * lm + 1],
                    div_max=out[5 * lm + 2], n_div=out[5 * lm + 3])

    def cl_spectra(self, u, r):
        """cl_energy.f, then spectra.f with the Chollet-Lesieur model (i_les = 1)"""
        uf = self._full(u, self.lu)
        rf = self._full(r, self.lr) if r.shape[3] >= self.lr else self._full(
            np.concatenate([r, np.zeros(r.shape[:3] + (self.lr - r.shape[3],), np.complex128)], axis=3), self.lr)
        lm = self.lm
        out = np.zeros(5 * lm + 6)
        self.R.ref_cl_spectra(_p(uf), _p(rf), _p(out))
        return dict(ek=out[:lm], tk=out[lm:2 * lm], dk=out[2 * lm:3 * lm], vt=out[3 * lm:4 * lm],
                    sample=out[4 * lm:5 * lm], energy_w=out[5 * lm], e_cl=out[5 * lm + 4], k_cl=out[5 * lm + 5])

    def div_check(self, u):
        uf = self._full(u, self.lu)
        out = np.zeros(3)
        self.R.ref_div_check(_p(uf), _p(out))
        return dict(div_max=out[0], div_rms=out[1], n_div=out[2])
