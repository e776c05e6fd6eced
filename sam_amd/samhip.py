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
    "sam_set_state", "sam_get_state", "sam_set_clock",
    "sam_step", "sam_rhs", "sam_get_spectra",
    "sam_z_trans_b", "sam_xy_trans_b", "sam_xy_trans_f", "sam_z_trans_f",
    "sam_products", "sam_vel_max", "sam_project",
    "sam_eddy_visc", "sam_test_filter",
    "sam_prof_enable", "sam_prof_read", "sam_sync", "sam_device_copy",
)

# kernel classes of sam_prof_read (SAM_K_* in include/samhip.h)
KERNEL_CLASSES = ("zb", "yb", "xrow", "final", "yf", "zrhs", "spec", "xb", "xf", "sgs", "misc")
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
    L.sam_set_clock.argtypes = [_hp, ctypes.c_double, ctypes.c_double, ctypes.c_longlong]
    L.sam_step.argtypes = [_hp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(Diag)]
    L.sam_rhs.argtypes = [_hp, _dp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(Diag)]
    L.sam_get_spectra.argtypes = [_hp, _dp]
    L.sam_z_trans_b.argtypes = [_hp, _dp, ctypes.c_int, _dp, ctypes.c_int]
    L.sam_xy_trans_b.argtypes = [_hp, _dp, _dp, ctypes.c_int]
    L.sam_xy_trans_f.argtypes = [_hp, _dp, _dp, ctypes.c_int]
    L.sam_z_trans_f.argtypes = [_hp, _dp, _dp, ctypes.c_int]
    L.sam_products.argtypes = [_hp, _dp, _dp]
    L.sam_vel_max.argtypes = [_hp, _dp, ctypes.c_double, _dp]
    L.sam_project.argtypes = [_hp, _dp, _dp, _dp]
    L.sam_eddy_visc.argtypes = [_hp, _dp, _dp, _dp]
    L.sam_test_filter.argtypes = [_hp, _dp, _dp, ctypes.c_int]
    L.sam_prof_enable.argtypes = [_hp, ctypes.c_int]
    L.sam_prof_read.argtypes = [_hp, _dp, ctypes.POINTER(ctypes.c_longlong)]
    L.sam_sync.argtypes = [_hp]
    L.sam_device_copy.argtypes = [_hp, ctypes.c_longlong, ctypes.c_int]
    for name in EXPORTS:
        fn = getattr(L, name)
        if fn.restype is None or fn.restype is ctypes.c_int:
            if name != "sam_last_error":
                fn.restype = ctypes.c_int
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
        u = np.asfortranarray(u, dtype=np.complex128)
        nvar = nvar or u.shape[3]
        out = np.zeros((self.nz, self.nym, self.nxm, nvar), np.complex128, order="F")
        _chk(self.L.sam_z_trans_b(self.h, _ptr(u), u.shape[0], _ptr(out), nvar))
        return out

    def xy_trans_b(self, a):
        a = np.asfortranarray(a, dtype=np.complex128)
        nvar = a.shape[3]
        out = np.zeros((self.nx, self.ny, self.nz, nvar), np.float64, order="F")
        _chk(self.L.sam_xy_trans_b(self.h, _ptr(a), _ptr(out), nvar))
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

    def test_filter(self, f):
        f = np.asfortranarray(f, dtype=np.float64)
        nvar = 1 if f.ndim == 3 else f.shape[3]
        out = np.zeros_like(f, order="F")
        _chk(self.L.sam_test_filter(self.h, _ptr(f), _ptr(out), nvar))
        return out
