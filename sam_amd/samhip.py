"""ctypes binding of libsamhip (include/samhip.h) and a thin host-side mirror
of the reference's interface for the RK3 step.

The reference's routines take Fortran arrays; here numpy arrays in Fortran
order stand in for them (``a[k, j, i, n]`` is ``a(k+1, j+1, i+1, n+1)``).
Every call goes to the HIP library; there is no CPU fallback.  Loading fails
with an exception when ``libsamhip.so`` has not been built.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libsamhip.so")

SAM_OK = 0
SAM_ERR_REQUIRED = 6
SAM_ERR_VALUE = 9
SAM_ERR_SIZE = 10
SAM_ERR_UNSUPPORTED = 20
SAM_ERR_STATE = 21
SAM_ERR_HIP = 30
SAM_ERR_NOMEM = 31
SAM_ERR_NODEVICE = 32

PREC_FP64 = 0
PREC_SPLIT = 1
SGS_NONE, SGS_SMAG, SGS_DYN, SGS_MLP = 0, 1, 2, 3

# every symbol include/samhip.h declares
EXPORTS = (
    "sam_create", "sam_destroy", "sam_get_dims", "sam_last_error",
    "sam_set_state", "sam_get_state", "sam_get_state_substep", "sam_set_clock",
    "sam_step", "sam_rhs", "sam_get_spectra",
    "sam_z_trans_b", "sam_x2z", "sam_xy_trans_b", "sam_xy_trans_f", "sam_z2x", "sam_z_trans_f",
    "sam_products", "sam_vel_max", "sam_vel_max_noalloc", "sam_project", "sam_rk_update", "sam_div_check",
    "sam_eddy_visc", "sam_test_filter",
    "sam_prof_enable", "sam_prof_read", "sam_sync", "sam_device_copy", "sam_device_stream",
    "sam_mlp_create", "sam_mlp_destroy", "sam_mlp_get_params", "sam_mlp_set_params",
    "sam_mlp_add_samples", "sam_mlp_add_samples_host", "sam_mlp_pool_size", "sam_mlp_get_samples",
    "sam_mlp_clear_pool", "sam_mlp_set_labels", "sam_mlp_last_coefficient", "sam_mlp_set_activation", "sam_mlp_get_activation",
    "sam_mlp_fit_norms", "sam_mlp_train", "sam_mlp_grad", "sam_mlp_infer",
    "sam_attach_mlp", "sam_set_sgs", "sam_sgs_check", "sam_set_mlp_refresh",
    "sam_set_forcing", "sam_set_cl", "sam_set_planes", "sam_planes", "sam_set_dyn_refresh",
)

MLP_NPARAM = 10448
LABELS_FILTERED, LABELS_DYNAMIC, LABELS_DYNAMIC_COEF = 0, 1, 2
# hidden activation of the learned closure (SAM_MLP_ACT_*)
ACT_RELU, ACT_GELU = 0, 1

# kernel classes of sam_prof_read (SAM_K_* in include/samhip.h)
KERNEL_CLASSES = ("zb", "yb", "xrow", "final", "yf", "zrhs", "spec", "xb", "xf", "sgs", "misc", "dmproc", "stats")
K_COUNT = 16


class SamError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"libsamhip error {code}: {msg}")
        self.code = code


class Params(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("n_dealias", ctypes.c_int), ("i_strat", ctypes.c_int), ("nrk_max", ctypes.c_int),
        ("i_cfl", ctypes.c_int), ("iubc_z", ctypes.c_int), ("precision", ctypes.c_int),
        ("sgs", ctypes.c_int),
        ("xl", ctypes.c_double), ("yl", ctypes.c_double), ("zl", ctypes.c_double),
        ("vis", ctypes.c_double), ("pr", ctypes.c_double), ("grav", ctypes.c_double),
        ("to", ctypes.c_double), ("lapse", ctypes.c_double), ("cfl0", ctypes.c_double),
        ("dt0", ctypes.c_double), ("c_smag", ctypes.c_double), ("pr_t", ctypes.c_double),
    ]


class Dims(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in (
        "kx_max", "ky_max", "kz_max", "nx_min", "ny_min", "nz_min", "nze", "lu", "lr", "lm")]


class Diag(ctypes.Structure):
    _fields_ = [
        ("nt", ctypes.c_longlong), ("time", ctypes.c_double), ("dt", ctypes.c_double),
        ("u_avg", ctypes.c_double * 4), ("u_var", ctypes.c_double * 4), ("u_max", ctypes.c_double * 4),
        ("energy", ctypes.c_double), ("cfl_x", ctypes.c_double), ("cfl_y", ctypes.c_double),
        ("cfl_z", ctypes.c_double), ("c_delta_sq", ctypes.c_double), ("sgs_active", ctypes.c_int),
        ("pad_", ctypes.c_int),
    ]

    def as_dict(self):
        return dict(nt=self.nt, time=self.time, dt=self.dt, u_avg=list(self.u_avg),
                    u_var=list(self.u_var), u_max=list(self.u_max), energy=self.energy,
                    cfl=(self.cfl_x, self.cfl_y, self.cfl_z), c_delta_sq=self.c_delta_sq,
                    sgs_active=self.sgs_active)


_lib = None
_dp = ctypes.POINTER(ctypes.c_double)
_hp = ctypes.c_void_p


def lib_path():
    return LIB_PATH


def load():
    """Load libsamhip.so; raises if it is missing (no fallback path exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} not built: run `python -m sam_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    L.sam_create.argtypes = [ctypes.POINTER(Params), ctypes.POINTER(_hp)]
    L.sam_destroy.argtypes = [_hp]
    L.sam_get_dims.argtypes = [_hp, ctypes.POINTER(Dims)]
    L.sam_last_error.restype = ctypes.c_char_p
    L.sam_set_state.argtypes = [_hp, _dp, ctypes.c_int]
    L.sam_get_state.argtypes = [_hp, _dp, ctypes.c_int]
    L.sam_get_state_substep.argtypes = [_hp, ctypes.c_int, _dp, ctypes.c_int]
    L.sam_set_clock.argtypes = [_hp, ctypes.c_double, ctypes.c_double, ctypes.c_longlong]
    L.sam_step.argtypes = [_hp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(Diag)]
    L.sam_rhs.argtypes = [_hp, _dp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(Diag)]
    L.sam_get_spectra.argtypes = [_hp, _dp]
    L.sam_z_trans_b.argtypes = [_hp, _dp, ctypes.c_int, _dp, ctypes.c_int]
    L.sam_x2z.argtypes = [_hp, _dp, _dp, ctypes.c_int]
    L.sam_z2x.argtypes = [_hp, _dp, _dp, ctypes.c_int]
    L.sam_xy_trans_b.argtypes = [_hp, _dp, _dp, ctypes.c_int]
    L.sam_xy_trans_f.argtypes = [_hp, _dp, _dp, ctypes.c_int]
    L.sam_z_trans_f.argtypes = [_hp, _dp, _dp, ctypes.c_int]
    L.sam_products.argtypes = [_hp, _dp, _dp]
    L.sam_vel_max.argtypes = [_hp, _dp, ctypes.c_double, _dp]
    L.sam_vel_max_noalloc.argtypes = [_hp, ctypes.c_void_p, ctypes.c_double, ctypes.c_void_p]
    L.sam_project.argtypes = [_hp, _dp, _dp, _dp]
    L.sam_rk_update.argtypes = [_hp, ctypes.c_int, ctypes.c_double, _dp, _dp, _dp, ctypes.c_int]
    L.sam_div_check.argtypes = [_hp, _dp]
    L.sam_eddy_visc.argtypes = [_hp, _dp, _dp, _dp]
    L.sam_test_filter.argtypes = [_hp, _dp, _dp, ctypes.c_int]
    L.sam_prof_enable.argtypes = [_hp, ctypes.c_int]
    L.sam_prof_read.argtypes = [_hp, _dp, ctypes.POINTER(ctypes.c_longlong)]
    L.sam_sync.argtypes = [_hp]
    L.sam_device_copy.argtypes = [_hp, ctypes.c_longlong, ctypes.c_int]
    L.sam_device_stream.argtypes = [_hp, ctypes.c_int, ctypes.c_int, ctypes.c_longlong, ctypes.c_int]
    _fp = ctypes.POINTER(ctypes.c_float)
    L.sam_mlp_create.argtypes = [ctypes.POINTER(_hp), ctypes.c_uint, ctypes.c_longlong]
    L.sam_mlp_destroy.argtypes = [_hp]
    L.sam_mlp_get_params.argtypes = [_hp, _fp]
    L.sam_mlp_set_params.argtypes = [_hp, _fp]
    L.sam_mlp_add_samples.argtypes = [_hp, _hp, ctypes.c_int]
    L.sam_mlp_add_samples_host.argtypes = [_hp, _fp, _fp, ctypes.c_longlong]
    L.sam_mlp_pool_size.argtypes = [_hp]
    L.sam_mlp_pool_size.restype = ctypes.c_longlong
    L.sam_mlp_clear_pool.argtypes = [_hp]
    L.sam_mlp_get_samples.argtypes = [_hp, ctypes.c_longlong, ctypes.c_longlong, _fp, _fp]
    L.sam_mlp_fit_norms.argtypes = [_hp]
    L.sam_mlp_set_labels.argtypes = [_hp, ctypes.c_int, ctypes.c_double]
    L.sam_mlp_last_coefficient.argtypes = [_hp, _dp]
    L.sam_mlp_set_activation.argtypes = [_hp, ctypes.c_int]
    L.sam_mlp_get_activation.argtypes = [_hp, ctypes.POINTER(ctypes.c_int)]
    L.sam_mlp_train.argtypes = [_hp, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_uint, _fp]
    L.sam_mlp_grad.argtypes = [_hp, _fp, _fp, ctypes.c_int, _fp, ctypes.POINTER(ctypes.c_double)]
    L.sam_mlp_infer.argtypes = [_hp, _fp, _fp, ctypes.c_longlong]
    L.sam_attach_mlp.argtypes = [_hp, _hp]
    L.sam_set_sgs.argtypes = [_hp, ctypes.c_int]
    L.sam_set_mlp_refresh.argtypes = [_hp, ctypes.c_int]
    L.sam_set_dyn_refresh.argtypes = [_hp, ctypes.c_int]
    L.sam_set_forcing.argtypes = [_hp, ctypes.c_int]
    L.sam_set_cl.argtypes = [_hp, ctypes.c_int]
    L.sam_set_planes.argtypes = [_hp, ctypes.c_int]
    _ipt = ctypes.POINTER(ctypes.c_int)
    L.sam_planes.argtypes = [_hp, ctypes.c_int, _ipt, ctypes.c_int, _ipt, ctypes.c_int, _ipt, _dp, _dp, _dp, _dp]
    L.sam_sgs_check.argtypes = [_hp, ctypes.c_double, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)]
    for name in EXPORTS:
        if name in ("sam_last_error", "sam_mlp_pool_size"):
            continue
        getattr(L, name).restype = ctypes.c_int
    _lib = L
    return L


def _chk(rc):
    if rc != SAM_OK:
        raise SamError(rc, load().sam_last_error().decode())


def _ptr(a):
    if a is None:
        return None
    assert a.dtype in (np.float64, np.complex128), a.dtype
    assert a.flags["F_CONTIGUOUS"], "arrays must be Fortran-ordered"
    return a.ctypes.data_as(_dp)


def make_params(nx, ny=None, nz=None, *, xl=2 * np.pi, yl=None, zl=None, n_dealias=1, i_strat=0,
                nrk_max=3, i_cfl=0, precision=PREC_FP64, sgs=SGS_NONE, vis=0.0025, pr=1.0,
                grav=0.0, to=1.0, lapse=0.0, cfl0=0.5, dt0=0.01, c_smag=0.01, pr_t=1.0):
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    yl = xl if yl is None else yl
    zl = xl if zl is None else zl
    return Params(nx, ny, nz, n_dealias, i_strat, nrk_max, i_cfl, 0, precision, sgs,
                  xl, yl, zl, vis, pr, grav, to, lapse, cfl0, dt0, c_smag, pr_t)


class Sam:
    """One device-resident SAM grid (a libsamhip handle)."""

    def __init__(self, params: Params):
        L = load()
        self.L = L
        self.params = params
        h = _hp()
        _chk(L.sam_create(ctypes.byref(params), ctypes.byref(h)))
        self.h = h
        d = Dims()
        _chk(L.sam_get_dims(h, ctypes.byref(d)))
        self.dims = d
        self.kxm, self.kym, self.kzm = d.kx_max, d.ky_max, d.kz_max
        self.nxm, self.nym, self.nzm = d.nx_min, d.ny_min, d.nz_min
        self.lu, self.lr, self.lm = d.lu, d.lr, d.lm
        self.nx, self.ny, self.nz = params.nx, params.ny, params.nz

    def close(self):
        if getattr(self, "h", None):
            self.L.sam_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- state ---------------------------------------------------------
    def set_state(self, u):
        u = np.asfortranarray(u, dtype=np.complex128)
        assert u.shape[1:] == (self.nym, self.nxm, self.lu), u.shape
        _chk(self.L.sam_set_state(self.h, _ptr(u), u.shape[0]))

    def get_state(self, ldz=None):
        ldz = ldz or self.nzm
        u = np.zeros((ldz, self.nym, self.nxm, self.lu), np.complex128, order="F")
        _chk(self.L.sam_get_state(self.h, _ptr(u), ldz))
        return u

    def get_state_substep(self, nrk):
        """the velocity sam.f writes to vel.* inside a step (after the zeta
        update of substep nrk)"""
        u = np.zeros((self.nzm, self.nym, self.nxm, self.lu), np.complex128, order="F")
        _chk(self.L.sam_get_state_substep(self.h, int(nrk), _ptr(u), self.nzm))
        return u

    def set_clock(self, time, dt, nt=0):
        _chk(self.L.sam_set_clock(self.h, time, dt, nt))

    def step(self, nsteps=1, want_spectra=False):
        d = Diag()
        _chk(self.L.sam_step(self.h, nsteps, int(want_spectra), ctypes.byref(d)))
        return d.as_dict()

    def rhs(self, want_spectra=False):
        r = np.zeros((self.nzm, self.nym, self.nxm, self.lu + 1), np.complex128, order="F")
        d = Diag()
        _chk(self.L.sam_rhs(self.h, _ptr(r), self.nzm, int(want_spectra), ctypes.byref(d)))
        return r, d.as_dict()

    def spectra(self):
        out = np.zeros(5 * self.lm + 3)
        _chk(self.L.sam_get_spectra(self.h, out.ctypes.data_as(_dp)))
        lm = self.lm
        return dict(ek=out[:lm], tk=out[lm:2 * lm], dk=out[2 * lm:3 * lm], vt=out[3 * lm:4 * lm],
                    sample=out[4 * lm:5 * lm], energy_w=out[5 * lm], div_rms=out[5 * lm + 1],
                    div_max=out[5 * lm + 2])

    # ---- stages (reference routine parity) ------------------------------
    def z_trans_b(self, u, nvar=None):
        """z_trans_b.f: u (ldz, Ny_min, Nx_min, n) -> (Ny_min, Nx_min, Nz, n)"""
        u = np.asfortranarray(u, dtype=np.complex128)
        nvar = nvar or u.shape[3]
        out = np.zeros((self.nym, self.nxm, self.nz, nvar), np.complex128, order="F")
        _chk(self.L.sam_z_trans_b(self.h, _ptr(u), u.shape[0], _ptr(out), nvar))
        return out

    def x2z(self, a):
        """x2z_decomp.f: (Ny_min, Nx_min, Nz, n) -> (Ny, Nx/2, Nz, n)"""
        a = np.asfortranarray(a, dtype=np.complex128)
        nvar = a.shape[3]
        out = np.zeros((self.ny, self.nx // 2, self.nz, nvar), np.complex128, order="F")
        _chk(self.L.sam_x2z(self.h, _ptr(a), _ptr(out), nvar))
        return out

    def xy_trans_b(self, a):
        """xy_trans_b.f: (Ny, Nx/2, Nz, n) complex -> (Nx, Ny, Nz, n) real"""
        a = np.asfortranarray(a, dtype=np.complex128)
        nvar = a.shape[3]
        out = np.zeros((self.nx, self.ny, self.nz, nvar), np.float64, order="F")
        _chk(self.L.sam_xy_trans_b(self.h, _ptr(a), _ptr(out), nvar))
        return out

    def z2x(self, a):
        """z2x_decomp.f: (Nz, Ny_min, Nx_min, n) -> (Nze, Ny_min, Nx_min, n)"""
        a = np.asfortranarray(a, dtype=np.complex128)
        nvar = a.shape[3]
        out = np.zeros((self.nz, self.nym, self.nxm, nvar), np.complex128, order="F")
        _chk(self.L.sam_z2x(self.h, _ptr(a), _ptr(out), nvar))
        return out

    def xy_trans_f(self, p):
        p = np.asfortranarray(p, dtype=np.float64)
        nvar = p.shape[3]
        out = np.zeros((self.nz, self.nym, self.nxm, nvar), np.complex128, order="F")
        _chk(self.L.sam_xy_trans_f(self.h, _ptr(p), _ptr(out), nvar))
        return out

    def z_trans_f(self, a):
        a = np.asfortranarray(a, dtype=np.complex128)
        nvar = a.shape[3]
        out = np.zeros((self.nzm, self.nym, self.nxm, nvar), np.complex128, order="F")
        _chk(self.L.sam_z_trans_f(self.h, _ptr(a), _ptr(out), nvar))
        return out

    def products(self, phys):
        phys = np.asfortranarray(phys, dtype=np.float64)
        out = np.zeros((self.nx, self.ny, self.nz, self.lr), np.float64, order="F")
        _chk(self.L.sam_products(self.h, _ptr(phys), _ptr(out)))
        return out

    def vel_max(self, phys, dt_in):
        phys = np.asfortranarray(phys, dtype=np.float64)
        st = np.zeros(17)
        _chk(self.L.sam_vel_max(self.h, _ptr(phys), dt_in, st.ctypes.data_as(_dp)))
        return st

    def project(self, F, usav):
        F = np.asfortranarray(F, dtype=np.complex128)
        usav = np.asfortranarray(usav, dtype=np.complex128)
        r = np.zeros((self.nzm, self.nym, self.nxm, self.lu + 1), np.complex128, order="F")
        _chk(self.L.sam_project(self.h, _ptr(F), _ptr(usav), _ptr(r)))
        return r

    def rk_update(self, nrk, dt, u, r_prev=None, r=None):
        """sam.f:210-221 (r_prev) and sam.f:293-314 (r) on a copy of u"""
        u = np.array(u, dtype=np.complex128, order="F", copy=True)
        rp = None if r_prev is None else np.asfortranarray(r_prev[..., : self.lu], dtype=np.complex128)
        rr = None if r is None else np.asfortranarray(r[..., : self.lu], dtype=np.complex128)
        for a in (rp, rr):
            assert a is None or a.shape == u.shape
        _chk(self.L.sam_rk_update(self.h, int(nrk), float(dt), _ptr(u), _ptr(rp), _ptr(rr), u.shape[0]))
        return u

    def div_check(self):
        """div_check.f of the current state: div_max, div_rms, n_div"""
        out = np.zeros(3)
        _chk(self.L.sam_div_check(self.h, out.ctypes.data_as(_dp)))
        return dict(div_max=out[0], div_rms=out[1], n_div=out[2])

    def eddy_visc(self, phys):
        phys = np.asfortranarray(phys, dtype=np.float64)
        vt = np.zeros((self.nx, self.ny, self.nz), np.float64, order="F")
        cd = np.zeros(1)
        _chk(self.L.sam_eddy_visc(self.h, _ptr(phys), _ptr(vt), cd.ctypes.data_as(_dp)))
        return vt, cd[0]

    # ---- measurement ---------------------------------------------------
    def prof_enable(self, on=True):
        _chk(self.L.sam_prof_enable(self.h, int(on)))

    def prof_read(self):
        ms = np.zeros(K_COUNT)
        cnt = (ctypes.c_longlong * K_COUNT)()
        _chk(self.L.sam_prof_read(self.h, ms.ctypes.data_as(_dp), cnt))
        return {name: (float(ms[i]), int(cnt[i])) for i, name in enumerate(KERNEL_CLASSES)}

    def sync(self):
        _chk(self.L.sam_sync(self.h))

    def device_copy(self, nbytes, reps=1):
        _chk(self.L.sam_device_copy(self.h, int(nbytes), int(reps)))

    def device_stream(self, nread, nwrite, nbytes, reps=1):
        _chk(self.L.sam_device_stream(self.h, int(nread), int(nwrite), int(nbytes), int(reps)))

    def test_filter(self, f):
        f = np.asfortranarray(f, dtype=np.float64)
        nvar = 1 if f.ndim == 3 else f.shape[3]
        out = np.zeros_like(f, order="F")
        _chk(self.L.sam_test_filter(self.h, _ptr(f), _ptr(out), nvar))
        return out


def _f32(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


class Mlp:
    """The learned SGS closure (a libsamhip sam_mlp handle)."""

    def __init__(self, seed=0, pool_capacity=0):
        self.L = load()
        h = _hp()
        _chk(self.L.sam_mlp_create(ctypes.byref(h), seed, int(pool_capacity)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.L.sam_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self):
        p = np.zeros(MLP_NPARAM + 32, np.float32)
        _chk(self.L.sam_mlp_get_params(self.h, p.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return p

    def set_params(self, p):
        p, ptr = _f32(p)
        assert p.size == MLP_NPARAM + 32
        _chk(self.L.sam_mlp_set_params(self.h, ptr))

    def add_samples(self, sam, ratio):
        _chk(self.L.sam_mlp_add_samples(self.h, sam.h, int(ratio)))

    def add_samples_host(self, feats, labels):
        f, fp = _f32(feats)
        lab, lp = _f32(labels)
        assert f.shape[1] == 12 and lab.shape[1] == 9 and f.shape[0] == lab.shape[0]
        _chk(self.L.sam_mlp_add_samples_host(self.h, fp, lp, f.shape[0]))

    def pool_size(self):
        return int(self.L.sam_mlp_pool_size(self.h))

    def get_samples(self, offset, n):
        f = np.zeros((n, 12), np.float32)
        lab = np.zeros((n, 9), np.float32)
        _chk(self.L.sam_mlp_get_samples(self.h, offset, n, f.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                        lab.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return f, lab

    def clear_pool(self):
        _chk(self.L.sam_mlp_clear_pool(self.h))

    def set_labels(self, mode, pr_t=1.0):
        """LABELS_FILTERED (subfilter stress of the box filter), LABELS_DYNAMIC
        (dynamic-model stress at unit coefficient) or LABELS_DYNAMIC_COEF (with
        the coefficient of the dynamic procedure on each filtered snapshot);
        set before adding samples"""
        _chk(self.L.sam_mlp_set_labels(self.h, int(mode), ctypes.c_double(pr_t)))

    def last_coefficient(self):
        """C Delta^2 of the last LABELS_DYNAMIC_COEF snapshot"""
        c = ctypes.c_double()
        _chk(self.L.sam_mlp_last_coefficient(self.h, ctypes.byref(c)))
        return c.value

    def set_activation(self, act):
        """ACT_RELU (default) or ACT_GELU; part of the network, set before training"""
        _chk(self.L.sam_mlp_set_activation(self.h, int(act)))

    def activation(self):
        a = ctypes.c_int(-1)
        _chk(self.L.sam_mlp_get_activation(self.h, ctypes.byref(a)))
        return a.value

    def fit_norms(self):
        _chk(self.L.sam_mlp_fit_norms(self.h))

    def train(self, nsteps, batch=1 << 20, lr=1e-3, seed=0):
        losses = np.zeros(max(nsteps, 1), np.float32)
        _chk(self.L.sam_mlp_train(self.h, nsteps, batch, lr, seed,
                                  losses.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return losses[:nsteps]

    def grad(self, feats, labels):
        f, fp = _f32(feats)
        lab, lp = _f32(labels)
        g = np.zeros(MLP_NPARAM, np.float32)
        loss = ctypes.c_double()
        _chk(self.L.sam_mlp_grad(self.h, fp, lp, f.shape[0], g.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                 ctypes.byref(loss)))
        return loss.value, g

    def infer(self, feats):
        f, fp = _f32(feats)
        out = np.zeros((f.shape[0], 9), np.float32)
        _chk(self.L.sam_mlp_infer(self.h, fp, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), f.shape[0]))
        return out


def attach_mlp(sam, mlp):
    _chk(load().sam_attach_mlp(sam.h, mlp.h))
    sam._mlp = mlp          # keep the network alive as long as the grid uses it


def sgs_check(sam, tol):
    """a-posteriori check of the learned closure (the dynamic model as the arbiter);
    returns (err, switched_to_dynamic)"""
    e = ctypes.c_double()
    sw = ctypes.c_int()
    _chk(load().sam_sgs_check(sam.h, tol, ctypes.byref(e), ctypes.byref(sw)))
    return e.value, bool(sw.value)


def set_sgs(sam, model):
    _chk(load().sam_set_sgs(sam.h, int(model)))


def set_forcing(sam, k_force):
    """force.f after every step: shells 2..k_force+1 held at the last
    spectrum (k_force < 0: off)"""
    _chk(load().sam_set_forcing(sam.h, int(k_force)))


CL_OFF, CL_DIAG, CL_COUPLED = 0, 1, 2


def set_cl(sam, mode):
    """Chollet-Lesieur model: CL_DIAG in the spectra (SAM's i_les = 1),
    CL_COUPLED also in the integrating factor"""
    _chk(load().sam_set_cl(sam.h, int(mode)))


def set_planes(sam, keep_pressure=True):
    """keep the last substep's pressure for the planes output"""
    _chk(load().sam_set_planes(sam.h, int(bool(keep_pressure))))


def planes(sam, k_xy=(), j_xz=(), i_yz=()):
    """write_planes.f / get_mean.f of the current state: (mean (Nz, Lu),
    xy (Nx, Ny, 2Lu, n_xy), xz (Nx, Nz, 2Lu, n_xz), yz (Ny, Nz, 2Lu, n_yz));
    positions 1-based"""
    p = sam.params
    nf = 2 * sam.lu
    ia = lambda v: (ctypes.c_int * max(len(v), 1))(*v)
    mean = np.zeros((p.nz, sam.lu), order="F")
    xy = np.zeros((p.nx, p.ny, nf, max(len(k_xy), 1)), order="F")
    xz = np.zeros((p.nx, p.nz, nf, max(len(j_xz), 1)), order="F")
    yz = np.zeros((p.ny, p.nz, nf, max(len(i_yz), 1)), order="F")
    _chk(load().sam_planes(sam.h, len(k_xy), ia(k_xy), len(j_xz), ia(j_xz), len(i_yz), ia(i_yz),
                           mean.ctypes.data_as(_dp), xy.ctypes.data_as(_dp), xz.ctypes.data_as(_dp),
                           yz.ctypes.data_as(_dp)))
    return mean, xy[..., :len(k_xy)], xz[..., :len(j_xz)], yz[..., :len(i_yz)]


def set_dyn_refresh(sam, nsteps):
    """dynamic model with C Delta^2 recomputed every nsteps steps (1: every
    substep, as dm.f)"""
    _chk(load().sam_set_dyn_refresh(sam.h, int(nsteps)))


def set_mlp_refresh(sam, nsteps):
    """steps between dynamic-coefficient updates for a unit-coefficient network"""
    _chk(load().sam_set_mlp_refresh(sam.h, int(nsteps)))
