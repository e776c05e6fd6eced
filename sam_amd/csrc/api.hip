// api.hip -- the extern "C" surface of libsamhip (include/samhip.h).
//
// One handle owns the device-resident state of one grid:
//   U[2]   spectral state, ping-pong          (Nz_min,Ny_min,Nx_min,Lu) c16
//   R      RHS of the last substep (+ pressure) (Nz_min,Ny_min,Nx_min,Lu+1) c16
//   Zb     z-physical work, Lr fields          (Nz,Ny_min,Nx_min,Lr) complex T;
//          between DNS substeps its first Lu fields hold the z inverse of
//          U[cur] (z_ready)
//   Yb     y,z-physical work, Lr fields        (Ny,Nx_min,Nz,Lr) complex T
//   Pb     physical fields (LES path and stage calls only), allocated on demand
// T = double (SAM_PREC_FP64) or float (SAM_PREC_SPLIT).
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "internal.hpp"

using namespace samhip;

static thread_local char g_err[512] = "";

static int set_err(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

#define HIPCHK(call)                                                                                  \
    do {                                                                                              \
        hipError_t e_ = (call);                                                                       \
        if (e_ != hipSuccess)                                                                         \
            return set_err(SAM_ERR_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_),        \
                           __FILE__, __LINE__);                                                       \
    } while (0)

#define RCHK(call)                                                                                    \
    do {                                                                                              \
        int r_ = (call);                                                                              \
        if (r_ != SAM_OK) {                                                                           \
            if (r_ == SAM_ERR_HIP)                                                                    \
                return set_err(r_, "kernel launch failed in %s: %s", #call,                           \
                               hipGetErrorString(hipGetLastError()));                                 \
            return set_err(r_, "%s returned %d", #call, r_);                                          \
        }                                                                                             \
    } while (0)

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
};

struct sam_mlp;
struct sam_handle {
    sam_params p;
    Dims d;
    StepConsts k;
    int dev = 0;
    hipStream_t st = nullptr;
    int prec = 0;
    double *U[2] = {nullptr, nullptr};
    int cur = 0;
    double *R = nullptr;
    void *Zb = nullptr, *Yb = nullptr;
    DevBuf Pb, stage0, stage1, sgsw, snap;
    double *partials = nullptr;
    double *specout = nullptr;
    DevDiag *diag = nullptr;     // [0] step, [1] stage calls
    // force.f: k_force >= 0 scales the low shells back to ek0 after every
    // step; ek0 = ek of the last spectra taken by sam_step / sam_rhs
    int k_force = -1;
    double *ek0 = nullptr, *ekg = nullptr;
    int ek0_valid = 0;
    // planes output: the last substep keeps its pressure (sam.f:148) when on
    int keep_p = 0;
    DevBuf planes_spec, planes_out;
    void *tw[3][2] = {{nullptr, nullptr}, {nullptr, nullptr}, {nullptr, nullptr}};
    int have_spectra = 0;
    int sgs_active = 0;
    sam_mlp *mlp = nullptr;      // network used when p.sgs == SAM_SGS_MLP
    int mlp_refresh = 20;        // steps between dynamic-coefficient updates (unit-coefficient networks)
    int mlp_since = 1 << 30;     // steps since the last update (large: update at the next step)
    int dyn_refresh = 1;         // SAM_SGS_DYN: 1 = dynamic procedure every substep (dm.f), N = every N steps
    // Zb's fields 0..lu-1 hold the z inverse of U[cur] (k_zb's output),
    // written by the last substep's fused k_zrhs; cleared by every other
    // writer of Zb or of the state
    int z_ready = 0;
    int dyn_since = 1 << 30;
    // HIP-event timing of the step's kernels (sam_prof_*)
    bool prof = false;
    std::vector<hipEvent_t> ev_free;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev_pending;
    double prof_ms[SAM_K_COUNT] = {0};
    long long prof_cnt[SAM_K_COUNT] = {0};
};

static hipEvent_t ev_get(sam_handle *h)
{
    hipEvent_t e = nullptr;
    if (!h->ev_free.empty()) {
        e = h->ev_free.back();
        h->ev_free.pop_back();
    } else if (hipEventCreate(&e) != hipSuccess) {
        (void)hipGetLastError();
        e = nullptr;
    }
    return e;
}

// brackets one kernel launch with an event pair when timing is on
struct ProfScope {
    sam_handle *h;
    int id;
    hipEvent_t a = nullptr, b = nullptr;
    ProfScope(sam_handle *h_, int id_) : h(h_), id(id_)
    {
        if (h->prof) {
            a = ev_get(h);
            b = ev_get(h);
            if (a && b) hipEventRecord(a, h->st);
        }
    }
    ~ProfScope()
    {
        if (h->prof && a && b) {
            hipEventRecord(b, h->st);
            h->ev_pending.push_back({id, {a, b}});
        } else {
            if (a) h->ev_free.push_back(a);
            if (b) h->ev_free.push_back(b);
        }
    }
};

// fold finished event pairs into the accumulators (after a stream sync)
static void prof_collect(sam_handle *h)
{
    for (auto &pe : h->ev_pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, pe.second.first, pe.second.second) == hipSuccess) {
            h->prof_ms[pe.first] += ms;
            h->prof_cnt[pe.first] += 1;
        } else {
            (void)hipGetLastError();
        }
        h->ev_free.push_back(pe.second.first);
        h->ev_free.push_back(pe.second.second);
    }
    h->ev_pending.clear();
}

#define PROF(id) ProfScope prof_scope_##__LINE__(h, id)

static bool pow2_in(int n, int lo, int hi) { return n >= lo && n <= hi && (n & (n - 1)) == 0; }

static int kmax_of(int n, int mode, int xdir)
{
    if (mode == 0) return n / 2 - 1;
    if (mode == 2 && xdir) return n / 2 - 1;
    return (int)(((double)n - 0.1) / 3.0);
}

static int ensure(DevBuf &b, size_t bytes)
{
    if (b.bytes >= bytes) return SAM_OK;
    if (b.p) hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
    if (hipMalloc(&b.p, bytes) != hipSuccess) {
        (void)hipGetLastError();
        return set_err(SAM_ERR_NOMEM, "hipMalloc(%zu) failed", bytes);
    }
    b.bytes = bytes;
    return SAM_OK;
}

// exp(-2 pi i k / n), k < n, followed by the half-length table exp(-2 pi i k
// / (n/2)), k < n/2 (the x rows of a 1024 grid run as 512-point transforms)
static int make_twiddles(sam_handle *h, int dir, int n)
{
    const size_t tot = (size_t)n + (size_t)n / 2;
    std::vector<double> td(2 * tot);
    std::vector<float> tf(2 * tot);
    const long double pi = 3.14159265358979323846264338327950288L;
    auto fill = [&](size_t off, int len) {
        for (int k = 0; k < len; k++) {
            const long double a = -2.0L * pi * (long double)k / (long double)len;
            const long double c = cosl(a), s = sinl(a);
            td[2 * (off + k)] = (double)c;
            td[2 * (off + k) + 1] = (double)s;
            tf[2 * (off + k)] = (float)c;
            tf[2 * (off + k) + 1] = (float)s;
        }
    };
    fill(0, n);
    fill(n, n / 2);
    HIPCHK(hipMalloc(&h->tw[dir][0], sizeof(double) * 2 * tot));
    HIPCHK(hipMalloc(&h->tw[dir][1], sizeof(float) * 2 * tot));
    HIPCHK(hipMemcpy(h->tw[dir][0], td.data(), sizeof(double) * 2 * tot, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(h->tw[dir][1], tf.data(), sizeof(float) * 2 * tot, hipMemcpyHostToDevice));
    return SAM_OK;
}

static size_t tsize(const sam_handle *h) { return h->prec ? sizeof(float) : sizeof(double); }

extern "C" const char *sam_last_error(void) { return g_err; }

extern "C" int sam_destroy(sam_handle *h)
{
    if (!h) return SAM_OK;
    if (h->st) hipStreamSynchronize(h->st);
    hipFree(h->U[0]);
    hipFree(h->U[1]);
    hipFree(h->R);
    hipFree(h->Zb);
    hipFree(h->Yb);
    hipFree(h->Pb.p);
    hipFree(h->planes_spec.p);
    hipFree(h->planes_out.p);
    hipFree(h->stage0.p);
    hipFree(h->stage1.p);
    hipFree(h->sgsw.p);
    hipFree(h->snap.p);
    hipFree(h->partials);
    hipFree(h->specout);
    hipFree(h->ek0);
    hipFree(h->ekg);
    hipFree(h->diag);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 2; j++) hipFree(h->tw[i][j]);
    prof_collect(h);
    for (auto e : h->ev_free) hipEventDestroy(e);
    if (h->st) hipStreamDestroy(h->st);
    delete h;
    return SAM_OK;
}

static int validate(const sam_params *p)
{
    if (!p) return set_err(SAM_ERR_VALUE, "null sam_params");
    if (!pow2_in(p->nx, 16, 1024) || !pow2_in(p->ny, 16, 1024) || !pow2_in(p->nz, 16, 1024))
        return set_err(SAM_ERR_UNSUPPORTED, "Nx, Ny, Nz must be powers of two in 16..1024 (got %d %d %d)", p->nx,
                       p->ny, p->nz);
    if (p->n_dealias < 0 || p->n_dealias > 2)
        return set_err(SAM_ERR_VALUE, "n_dealias must be 0, 1 or 2 (got %d)", p->n_dealias);
    if (p->nrk_max < 1 || p->nrk_max > 3) return set_err(SAM_ERR_VALUE, "nrk_max must be 1..3 (got %d)", p->nrk_max);
    if (p->iubc_z != 0) return set_err(SAM_ERR_UNSUPPORTED, "only periodic z (iubc_z = 0) is built");
    if (p->i_strat != 0 && p->i_strat != 1) return set_err(SAM_ERR_VALUE, "i_strat must be 0 or 1");
    if (p->precision != SAM_PREC_FP64 && p->precision != SAM_PREC_SPLIT)
        return set_err(SAM_ERR_VALUE, "precision must be SAM_PREC_FP64 or SAM_PREC_SPLIT");
    if (p->sgs < SAM_SGS_NONE || p->sgs > SAM_SGS_MLP) return set_err(SAM_ERR_VALUE, "unknown sgs model %d", p->sgs);
    if (!(p->xl > 0) || !(p->yl > 0) || !(p->zl > 0)) return set_err(SAM_ERR_VALUE, "box lengths must be > 0");
    if (!(p->vis >= 0)) return set_err(SAM_ERR_VALUE, "vis must be >= 0");
    if (!(p->pr > 0)) return set_err(SAM_ERR_VALUE, "pr must be > 0");
    if (p->i_strat && !(p->to > 0)) return set_err(SAM_ERR_VALUE, "To must be > 0");
    if (p->i_cfl == 1 && !(p->cfl0 > 0)) return set_err(SAM_ERR_VALUE, "cfl0 must be > 0");
    if (p->i_cfl != 1 && !(p->dt0 > 0)) return set_err(SAM_ERR_REQUIRED, "dt0 must be > 0 when i_cfl = 0");
    if (p->sgs != SAM_SGS_NONE && !(p->pr_t > 0)) return set_err(SAM_ERR_VALUE, "pr_t must be > 0");
    return SAM_OK;
}

// Workspace of sam_vel_max_noalloc, one copy per device for the life of
// the process: the statistics reduction's partial rows (one row of 3 Lu <= 12
// doubles per k_stats_phys block, at most one per 2048 points), the rows of
// its second level and the DevDiag it finalises into.  6.3 MB.
constexpr long long NOALLOC_ROWS = 65536;        // grids up to 2^27 points (512^3)
constexpr long long NOALLOC_DIAG = NOALLOC_ROWS * 12 + FINALIZE_ROWS;
constexpr long long NOALLOC_DOUBLES = NOALLOC_DIAG + (long long)(sizeof(DevDiag) / sizeof(double));
static __device__ __align__(256) double scratch[NOALLOC_DOUBLES];
constexpr int NOALLOC_MAX_DEV = 64;
static std::mutex g_scratch_mu;
static double *g_scratch_ptr[NOALLOC_MAX_DEV];   // the array's address on each device

// the device address of `scratch` on device dev (current), looked up once
static int scratch_setup(int dev)
{
    if (dev < 0 || dev >= NOALLOC_MAX_DEV) return set_err(SAM_ERR_NODEVICE, "device %d beyond %d", dev, NOALLOC_MAX_DEV);
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (g_scratch_ptr[dev]) return SAM_OK;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
        (void)hipGetLastError();
        return set_err(SAM_ERR_HIP, "hipGetSymbolAddress(scratch) failed on device %d: %s", dev, hipGetErrorString(e));
    }
    g_scratch_ptr[dev] = (double *)ptr;
    return SAM_OK;
}

extern "C" int sam_create(const sam_params *p, sam_handle **out)
{
    if (!out) return set_err(SAM_ERR_VALUE, "null handle pointer");
    *out = nullptr;
    int rc = validate(p);
    if (rc) return rc;

    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        (void)hipGetLastError();
        return set_err(SAM_ERR_NODEVICE, "no HIP device visible");
    }
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, dev));
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return set_err(SAM_ERR_NODEVICE, "libsamhip is built for gfx950, device is %s", prop.gcnArchName);

    if ((rc = scratch_setup(dev))) return rc;

    sam_handle *h = new sam_handle();
    h->p = *p;
    h->dev = dev;
    h->prec = p->precision;
    Dims &d = h->d;
    d.nx = p->nx;
    d.ny = p->ny;
    d.nz = p->nz;
    d.kxm = kmax_of(p->nx, p->n_dealias, 1);
    d.kym = kmax_of(p->ny, p->n_dealias, 0);
    d.kzm = kmax_of(p->nz, p->n_dealias, 0);
    d.nxm = d.kxm + 1;
    d.nym = 2 * d.kym + 1;
    d.nzm = 2 * d.kzm + 1;
    d.lu = p->i_strat ? 4 : 3;
    d.lr = p->i_strat ? 9 : 6;
    const int ksq = d.kxm * d.kxm + d.kym * d.kym + d.kzm * d.kzm;
    d.lm = (int)(sqrt((double)ksq) + 0.5) + 1;
    d.ss = (long long)d.nxm * d.nym * d.nzm;
    d.zfield = (long long)d.nxm * d.nym * d.nz;
    d.yfield = (long long)d.nz * d.nxm * d.ny;
    d.pfield = (long long)d.nx * d.ny * d.nz;
    const double pi = acos(-1.0);
    d.wave_x = 2.0 * pi / p->xl;
    d.wave_y = 2.0 * pi / p->yl;
    d.wave_z = 2.0 * pi / p->zl;

    StepConsts &k = h->k;
    memset(&k, 0, sizeof k);
    if (p->nrk_max == 1) { k.gamma[0] = 1.0; k.adv_cfl_limit = 1.0; }
    if (p->nrk_max == 2) { k.gamma[0] = 1.0; k.gamma[1] = 0.5; k.zeta[1] = -0.5; k.adv_cfl_limit = 1.0; }
    if (p->nrk_max == 3) {
        k.gamma[0] = 8.0 / 15.0; k.gamma[1] = 5.0 / 12.0; k.gamma[2] = 3.0 / 4.0;
        k.zeta[1] = -17.0 / 60.0; k.zeta[2] = -5.0 / 12.0;
        k.adv_cfl_limit = sqrt(3.0);
    }
    k.vis = p->vis;
    k.pr_inv = 1.0 / p->pr;
    k.buoy_fac_x = 0.0;
    k.buoy_fac_z = p->i_strat ? p->grav / p->to : 0.0;
    k.lapse = p->i_strat ? p->lapse : 0.0;
    k.cfl0 = p->cfl0;
    k.dt0 = p->dt0;
    k.i_cfl = p->i_cfl;
    k.strat = p->i_strat;
    k.cl = 0;                                   // sam_set_cl
    k.cl_a = 0.145;                             // init.f:88-93
    k.cl_b = 5.01;
    k.cl_c = 3.03;
    k.k_cl = (double)(int)(p->nx / 3.0);
    k.pr_t_inv = 1.0 / p->pr_t;

    auto fail = [&](int code) { sam_destroy(h); return code; };
    if (hipStreamCreateWithFlags(&h->st, hipStreamNonBlocking) != hipSuccess)
        return fail(set_err(SAM_ERR_HIP, "hipStreamCreate failed"));
    const size_t ts = tsize(h);
    const size_t su = sizeof(double) * 2 * d.ss * d.lu;
    struct { void **pp; size_t bytes; } allocs[] = {
        {(void **)&h->U[0], su},
        {(void **)&h->U[1], su},
        {(void **)&h->R, sizeof(double) * 2 * d.ss * (d.lu + 1)},
        {&h->Zb, ts * 2 * d.zfield * d.lr},
        {&h->Yb, ts * 2 * d.yfield * d.lr},
        {(void **)&h->partials, sizeof(double) * (size_t)partials_capacity(d)},
        {(void **)&h->specout, sizeof(double) * (5 * d.lm + 8)},
        {(void **)&h->ek0, sizeof(double) * d.lm},
        {(void **)&h->ekg, sizeof(double) * d.lm},
        {(void **)&h->diag, sizeof(DevDiag) * 2},
    };
    for (auto &a : allocs) {
        if (hipMalloc(a.pp, a.bytes) != hipSuccess) {
            (void)hipGetLastError();
            return fail(set_err(SAM_ERR_NOMEM, "hipMalloc(%zu) failed", a.bytes));
        }
    }
    if (hipMemset(h->U[0], 0, su) != hipSuccess || hipMemset(h->R, 0, sizeof(double) * 2 * d.ss * (d.lu + 1)) != hipSuccess)
        return fail(set_err(SAM_ERR_HIP, "hipMemset failed"));
    DevDiag dg[2];
    memset(dg, 0, sizeof dg);
    dg[0].dt = dg[1].dt = p->dt0;
    if (hipMemcpy(h->diag, dg, sizeof dg, hipMemcpyHostToDevice) != hipSuccess)
        return fail(set_err(SAM_ERR_HIP, "hipMemcpy failed"));
    if ((rc = make_twiddles(h, 0, d.nx)) || (rc = make_twiddles(h, 1, d.ny)) || (rc = make_twiddles(h, 2, d.nz)))
        return fail(rc);
    h->sgs_active = p->sgs;
    *out = h;
    return SAM_OK;
}

extern "C" int sam_get_dims(const sam_handle *h, sam_dims *o)
{
    if (!h || !o) return set_err(SAM_ERR_VALUE, "null argument");
    o->kx_max = h->d.kxm;
    o->ky_max = h->d.kym;
    o->kz_max = h->d.kzm;
    o->nx_min = h->d.nxm;
    o->ny_min = h->d.nym;
    o->nz_min = h->d.nzm;
    o->nze = h->d.nz;
    o->lu = h->d.lu;
    o->lr = h->d.lr;
    o->lm = h->d.lm;
    return SAM_OK;
}

// host (ldz, ny_min, nx_min, nvar) <-> device packed (nz_min, ...)
static int h2d_spec(sam_handle *h, double *dst, const double *src, int ldz, int nvar)
{
    const Dims &d = h->d;
    if (ldz < d.nzm) return set_err(SAM_ERR_SIZE, "ldz %d < Nz_min %d", ldz, d.nzm);
    HIPCHK(hipMemcpy2DAsync(dst, sizeof(double) * 2 * d.nzm, src, sizeof(double) * 2 * ldz, sizeof(double) * 2 * d.nzm,
                            (size_t)d.nym * d.nxm * nvar, hipMemcpyHostToDevice, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    return SAM_OK;
}

static int d2h_spec(sam_handle *h, double *dst, const double *src, int ldz, int nvar)
{
    const Dims &d = h->d;
    if (ldz < d.nzm) return set_err(SAM_ERR_SIZE, "ldz %d < Nz_min %d", ldz, d.nzm);
    if (ldz > d.nzm) memset(dst, 0, sizeof(double) * 2 * (size_t)ldz * d.nym * d.nxm * nvar);
    HIPCHK(hipMemcpy2DAsync(dst, sizeof(double) * 2 * ldz, src, sizeof(double) * 2 * d.nzm, sizeof(double) * 2 * d.nzm,
                            (size_t)d.nym * d.nxm * nvar, hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    return SAM_OK;
}

extern "C" int sam_set_state(sam_handle *h, const double *u, int ldz)
{
    if (!h || !u) return set_err(SAM_ERR_VALUE, "null argument");
    h->have_spectra = 0;
    h->z_ready = 0;
    return h2d_spec(h, h->U[h->cur], u, ldz, h->d.lu);
}

extern "C" int sam_get_state(sam_handle *h, double *u, int ldz)
{
    if (!h || !u) return set_err(SAM_ERR_VALUE, "null argument");
    return d2h_spec(h, u, h->U[h->cur], ldz, h->d.lu);
}

extern "C" int sam_set_clock(sam_handle *h, double time, double dt, long long nt)
{
    if (!h) return set_err(SAM_ERR_VALUE, "null handle");
    DevDiag dg;
    HIPCHK(hipMemcpy(&dg, h->diag, sizeof dg, hipMemcpyDeviceToHost));
    dg.time = time;
    dg.dt = dt;
    dg.nt = nt;
    HIPCHK(hipMemcpy(h->diag, &dg, sizeof dg, hipMemcpyHostToDevice));
    return SAM_OK;
}

static int fetch_diag(sam_handle *h, sam_diag *o)
{
    HIPCHK(hipStreamSynchronize(h->st));
    prof_collect(h);
    if (!o) return SAM_OK;
    DevDiag dg;
    HIPCHK(hipMemcpy(&dg, h->diag, sizeof dg, hipMemcpyDeviceToHost));
    memset(o, 0, sizeof *o);
    o->nt = dg.nt;
    o->time = dg.time;
    o->dt = dg.dt;
    for (int i = 0; i < 4; i++) {
        o->u_avg[i] = dg.u_avg[i];
        o->u_var[i] = dg.u_var[i];
        o->u_max[i] = dg.u_max[i];
    }
    o->energy = dg.energy;
    o->cfl_x = dg.cfl_x;
    o->cfl_y = dg.cfl_y;
    o->cfl_z = dg.cfl_z;
    o->c_delta_sq = dg.spare[3];
    o->sgs_active = h->sgs_active;
    return SAM_OK;
}

// physical-space LES fluxes: Y (Lu) -> P -> SGS model -> Q (Lr) -> Y (Lr)
template <typename T> static int les_bufs(sam_handle *h)
{
    const Dims &d = h->d;
    return ensure(h->Pb, sizeof(T) * d.pfield * (d.lu + d.lr));
}

static double delta_sq(const sam_handle *h)
{
    const double dx = h->p.xl / h->p.nx, dy = h->p.yl / h->p.ny, dz = h->p.zl / h->p.nz;
    return pow(dx * dy * dz, 2.0 / 3.0);   // init.f:28
}

// eddy-viscosity coefficient C Delta^2 into dg->c_delta_sq: the dynamic
// procedure of dm.f (dyn) or c_smag Delta^2 of sm.f
template <typename T>
static int sgs_coefficient(sam_handle *h, const T *P, DevDiag *dg, int report, bool dyn)
{
    const Dims &d = h->d;
    if (!dyn) {
        RCHK(launch_dm_coef(h->partials, 0, h->p.c_smag * delta_sq(h), 0, report, dg, h->st));
        return SAM_OK;
    }
    const double dxi = d.nx / h->p.xl, dyi = d.ny / h->p.yl, dzi = d.nz / h->p.zl;   // 1/dx (init.f:22-27)
    int nb = 0;
    RCHK(launch_dm_fused<T>(d, P, dxi, dyi, dzi, h->partials, h->st, &nb));
    RCHK(launch_dm_coef(h->partials, nb, 0.0, 1, report, dg, h->st));
    return SAM_OK;
}

template <typename T> static int mlp_flux(sam_handle *h, const T *P, T *Q);
static bool mlp_unit_coefficient(const sam_mlp *m);

template <typename T> static int les_fluxes(sam_handle *h, int stats)
{
    const Dims &d = h->d;
    const int pi = h->prec;
    if (h->p.sgs == SAM_SGS_MLP && !h->mlp) return set_err(SAM_ERR_STATE, "sgs = MLP but no network attached");
    const bool mlp_dyn = h->p.sgs == SAM_SGS_MLP && h->mlp && mlp_unit_coefficient(h->mlp);
    int rc = les_bufs<T>(h);
    if (rc) return rc;
    T *P = (T *)h->Pb.p, *Q = P + d.pfield * d.lu;
    T *Y = (T *)h->Yb;
    {
        PROF(SAM_K_XB);
        RCHK(launch_xb<T>(d, Y, d.lu, P, h->tw[0][pi], h->st));
    }
    if (stats) {
        PROF(SAM_K_STATS);
        int nb = 0;
        RCHK(launch_stats_phys(d, P, sizeof(T) == sizeof(double), d.lu, h->partials, h->diag, h->st, &nb));
        RCHK(launch_finalize(d, h->k, h->partials, nb, h->diag, 1, h->st));
    }
    // the eddy-viscosity coefficient: when the dynamic procedure runs (every
    // substep for dm.f, at refresh steps for the lagged coefficient and the
    // unit-coefficient network) and for Smagorinsky's fixed value
    bool coef = false, dyn = false;
    if (h->p.sgs == SAM_SGS_MLP) {
        // unit-coefficient network: the dynamic procedure supplies C Delta^2
        // at the first substep of every mlp_refresh-th step
        if (mlp_dyn && stats && ++h->mlp_since >= h->mlp_refresh) {
            coef = dyn = true;
            h->mlp_since = 0;
        }
    } else {
        dyn = h->p.sgs == SAM_SGS_DYN;
        coef = true;
        if (dyn && h->dyn_refresh > 1) {
            // lagged coefficient: the procedure at the first substep of every
            // dyn_refresh-th step, C Delta^2 held in between
            coef = stats && ++h->dyn_since >= h->dyn_refresh;
            if (coef) h->dyn_since = 0;
        }
    }
    if (coef) {
        PROF(SAM_K_DMPROC);
        rc = sgs_coefficient<T>(h, P, h->diag, stats, dyn);
        if (rc) return rc;
    }
    {
        PROF(SAM_K_SGS);
        if (h->p.sgs == SAM_SGS_MLP) {
            rc = mlp_flux<T>(h, P, Q);
            if (rc) return rc;
        } else {
            RCHK(launch_sgs_flux<T>(d, P, h->diag, Q, d.nx / h->p.xl, d.ny / h->p.yl, d.nz / h->p.zl,
                                    1.0 / h->p.pr_t, h->st));
        }
    }
    {
        PROF(SAM_K_XF);
        RCHK(launch_xf<T>(d, Q, d.lr, Y, h->tw[0][pi], h->st));
    }
    return SAM_OK;
}

// Everything up to (and including) the forward y transform: leaves the
// x,y-transformed fluxes of the state U[cur] in Zb (Lr fields, Z layout).
template <typename T> static int fluxes_to_z(sam_handle *h, int stats)
{
    const Dims &d = h->d;
    T *Z = (T *)h->Zb, *Y = (T *)h->Yb;
    const int pi = h->prec;
    if (!h->z_ready) {
        PROF(SAM_K_ZB);
        RCHK(launch_zb<T>(d, h->U[h->cur], d.nzm, d.lu, Z, h->tw[2][pi], h->st));
    }
    h->z_ready = 0;   // Zb is overwritten below
    {
        PROF(SAM_K_YB);
        RCHK(launch_yb<T>(d, Z, d.lu, Y, h->tw[1][pi], h->st));
    }
    if (h->p.sgs == SAM_SGS_NONE) {
        int nb = 0;
        {
            PROF(SAM_K_XROW);
            RCHK(launch_xrow<T>(d, Y, stats, h->partials, h->diag, h->tw[0][pi], h->st, &nb));
        }
        if (stats) {
            PROF(SAM_K_FINAL);
            RCHK(launch_finalize(d, h->k, h->partials, nb, h->diag, 1, h->st));
        }
    } else {
        int rc = les_fluxes<T>(h, stats);
        if (rc) return rc;
    }
    {
        PROF(SAM_K_YF);
        RCHK(launch_yf<T>(d, Y, d.lr, Z, h->tw[1][pi], h->st));
    }
    return SAM_OK;
}

// cl_energy at the first substep of a step (sam.f:128-132)
static int cl_energy(sam_handle *h)
{
    if (!h->k.cl) return SAM_OK;
    PROF(SAM_K_MISC);
    RCHK(launch_cl_energy(h->d, (int)h->k.k_cl, h->U[h->cur], h->partials, h->diag, h->st));
    return SAM_OK;
}

// spectra of U[cur] and R (spectra.f); a spectrum taken at the reference's
// call sites becomes the forcing target ek_0 (sam.f:250-251)
static int spectra(sam_handle *h, bool target)
{
    PROF(SAM_K_SPEC);
    RCHK(launch_spectra(h->d, h->k, h->U[h->cur], h->R, h->partials, h->specout, h->diag, h->st));
    h->have_spectra = 1;
    if (target) {
        HIPCHK(hipMemcpyAsync(h->ek0, h->specout, sizeof(double) * h->d.lm, hipMemcpyDeviceToDevice, h->st));
        h->ek0_valid = 1;
    }
    return SAM_OK;
}

// force.f after a step: shell energies, then the scaling toward ek0.  With
// no spectrum taken yet the current energies become the target.
static int force(sam_handle *h)
{
    if (h->k_force < 0) return SAM_OK;
    h->z_ready = 0;   // the state changes
    PROF(SAM_K_MISC);
    RCHK(launch_force_shells(h->d, h->k_force, h->U[h->cur], h->ekg, h->st));
    if (!h->ek0_valid) {
        HIPCHK(hipMemcpyAsync(h->ek0, h->ekg, sizeof(double) * (h->k_force + 1), hipMemcpyDeviceToDevice, h->st));
        h->ek0_valid = 1;
    }
    RCHK(launch_force_scale(h->d, h->k_force, h->U[h->cur], h->ek0, h->ekg, h->st));
    return SAM_OK;
}

template <typename T> static int substep(sam_handle *h, int m, int spectra_now)
{
    const Dims &d = h->d;
    bool fused = false;
    if (m == 0) {
        int rc = cl_energy(h);
        if (rc) return rc;
    }
    int rc = fluxes_to_z<T>(h, m == 0);
    if (rc) return rc;
    {
        PROF(SAM_K_ZRHS);
        const int mode = (h->keep_p && m == h->p.nrk_max - 1) ? 2 : 0;
        // the DNS substep also writes the next substep's z inverse (Zb in place);
        // the LES substep where the persistent kernel runs (Nz = 512): at 128^3
        // the fused form measured 1-2% slower than k_zb + k_zrhs
        T *zout = (h->p.sgs == SAM_SGS_NONE || d.nz == 512) ? (T *)h->Zb : nullptr;
        RCHK(launch_zrhs<T>(d, h->k, (const T *)h->Zb, h->U[h->cur], h->R, h->U[1 - h->cur], h->diag, m, mode,
                            h->tw[2][h->prec], h->st, zout, &fused));
    }
    if (spectra_now) {
        rc = spectra(h, true);
        if (rc) return rc;
    }
    h->cur = 1 - h->cur;
    h->z_ready = fused;
    return SAM_OK;
}

extern "C" int sam_step(sam_handle *h, int nsteps, int want_spectra, sam_diag *diag)
{
    if (!h) return set_err(SAM_ERR_VALUE, "null handle");
    if (nsteps < 0) return set_err(SAM_ERR_VALUE, "nsteps < 0");
    for (int n = 0; n < nsteps; n++) {
        for (int m = 0; m < h->p.nrk_max; m++) {
            const int spec = want_spectra && n == nsteps - 1 && m == 0;
            int rc = h->prec ? substep<float>(h, m, spec) : substep<double>(h, m, spec);
            if (rc) return rc;
        }
        {
            PROF(SAM_K_MISC);
            RCHK(launch_advance_clock(h->diag, h->st));
        }
        int rc = force(h);   // sam.f:320-322
        if (rc) return rc;
    }
    return fetch_diag(h, diag);
}

// the RHS-only pass; `target`: a spectrum taken here is the forcing target
static int rhs_pass(sam_handle *h, double *r, int ldz, int want_spectra, sam_diag *diag, bool target)
{
    if (!h) return set_err(SAM_ERR_VALUE, "null handle");
    const Dims &d = h->d;
    int rc = cl_energy(h);
    if (rc) return rc;
    rc = h->prec ? fluxes_to_z<float>(h, 1) : fluxes_to_z<double>(h, 1);
    if (rc) return rc;
    if (h->prec)
        RCHK(launch_zrhs<float>(d, h->k, (const float *)h->Zb, h->U[h->cur], h->R, nullptr, h->diag, 0, 1,
                                h->tw[2][1], h->st));
    else
        RCHK(launch_zrhs<double>(d, h->k, (const double *)h->Zb, h->U[h->cur], h->R, nullptr, h->diag, 0, 1,
                                 h->tw[2][0], h->st));
    if (want_spectra) {
        rc = spectra(h, target);
        if (rc) return rc;
    }
    if (r) {
        rc = d2h_spec(h, r, h->R, ldz, d.lu + 1);
        if (rc) return rc;
    }
    return fetch_diag(h, diag);
}

extern "C" int sam_rhs(sam_handle *h, double *r, int ldz, int want_spectra, sam_diag *diag)
{
    return rhs_pass(h, r, ldz, want_spectra, diag, true);
}

// What a call that runs substeps for a look-ahead (sam_get_state_substep,
// sam_sgs_check) must give back: the state, R (with the kept pressure), the
// step diagnostics, the spectra and the counters of the lagged coefficients.
struct Snapshot {
    DevDiag dg;
    int cur, mlp_since, dyn_since, have_spectra, sgs;
};

static int snap_save(sam_handle *h, Snapshot &s)
{
    const Dims &d = h->d;
    const size_t su = sizeof(double) * 2 * d.ss * d.lu, sr = sizeof(double) * 2 * d.ss * (d.lu + 1);
    const size_t ssp = sizeof(double) * (5 * d.lm + 8);
    int rc = ensure(h->snap, su + sr + ssp);
    if (rc) return rc;
    char *b = (char *)h->snap.p;
    HIPCHK(hipMemcpyAsync(b, h->U[h->cur], su, hipMemcpyDeviceToDevice, h->st));
    HIPCHK(hipMemcpyAsync(b + su, h->R, sr, hipMemcpyDeviceToDevice, h->st));
    HIPCHK(hipMemcpyAsync(b + su + sr, h->specout, ssp, hipMemcpyDeviceToDevice, h->st));
    HIPCHK(hipMemcpyAsync(&s.dg, h->diag, sizeof(DevDiag), hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    s.cur = h->cur;
    s.mlp_since = h->mlp_since;
    s.dyn_since = h->dyn_since;
    s.have_spectra = h->have_spectra;
    s.sgs = h->p.sgs;
    return SAM_OK;
}

static int snap_restore(sam_handle *h, const Snapshot &s)
{
    const Dims &d = h->d;
    const size_t su = sizeof(double) * 2 * d.ss * d.lu, sr = sizeof(double) * 2 * d.ss * (d.lu + 1);
    const size_t ssp = sizeof(double) * (5 * d.lm + 8);
    const char *b = (const char *)h->snap.p;
    h->cur = s.cur;
    HIPCHK(hipMemcpyAsync(h->U[h->cur], b, su, hipMemcpyDeviceToDevice, h->st));
    HIPCHK(hipMemcpyAsync(h->R, b + su, sr, hipMemcpyDeviceToDevice, h->st));
    HIPCHK(hipMemcpyAsync(h->specout, b + su + sr, ssp, hipMemcpyDeviceToDevice, h->st));
    HIPCHK(hipMemcpyAsync(h->diag, &s.dg, sizeof(DevDiag), hipMemcpyHostToDevice, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    h->mlp_since = s.mlp_since;
    h->dyn_since = s.dyn_since;
    h->have_spectra = s.have_spectra;
    h->p.sgs = s.sgs;
    h->z_ready = 0;   // Zb was overwritten
    return SAM_OK;
}

// The velocity the reference holds at substep nrk of the coming step, after
// its zeta update (sam.f:210-221): the state sam.f:277-284 writes to vel.*
// for nt < nt_end (the last of the nrk_max writes inside the step wins).
extern "C" int sam_get_state_substep(sam_handle *h, int nrk, double *u, int ldz)
{
    if (!h || !u) return set_err(SAM_ERR_VALUE, "null argument");
    if (nrk < 1 || nrk > h->p.nrk_max) return set_err(SAM_ERR_VALUE, "nrk must be 1..nrk_max (got %d)", nrk);
    if (nrk == 1) return sam_get_state(h, u, ldz);
    if (ldz < h->d.nzm) return set_err(SAM_ERR_SIZE, "ldz %d < Nz_min %d", ldz, h->d.nzm);
    Snapshot s;
    int rc = snap_save(h, s);
    if (rc) return rc;
    for (int m = 0; m < nrk - 1 && !rc; m++) rc = h->prec ? substep<float>(h, m, 0) : substep<double>(h, m, 0);
    if (!rc) rc = ensure(h->stage0, sizeof(double) * 2 * h->d.ss * h->d.lu);
    if (!rc) {
        DevDiag dg;
        double *dU = (double *)h->stage0.p;
        if (hipMemcpyAsync(dU, h->U[h->cur], sizeof(double) * 2 * h->d.ss * h->d.lu, hipMemcpyDeviceToDevice, h->st) !=
                hipSuccess ||
            hipMemcpyAsync(&dg, h->diag, sizeof dg, hipMemcpyDeviceToHost, h->st) != hipSuccess ||
            hipStreamSynchronize(h->st) != hipSuccess)
            rc = set_err(SAM_ERR_HIP, "look-ahead copy failed: %s", hipGetErrorString(hipGetLastError()));
        if (!rc) rc = launch_rk_update(h->d, h->k, nrk - 1, dg.dt, dU, h->R, nullptr, h->st);
        if (!rc) rc = d2h_spec(h, u, dU, ldz, h->d.lu);
    }
    const int rc2 = snap_restore(h, s);
    return rc ? rc : rc2;
}

extern "C" int sam_set_forcing(sam_handle *h, int k_force)
{
    if (!h) return set_err(SAM_ERR_VALUE, "null handle");
    const int kmax = std::max(h->p.nx / 2, std::max(h->p.ny / 2, h->p.nz / 2));
    if (k_force > kmax) return set_err(SAM_ERR_VALUE, "k_force must be <= max(Nx,Ny,Nz)/2 (check_inputs.f:215)");
    if (k_force >= h->d.lm) return set_err(SAM_ERR_VALUE, "k_force beyond the spectral shells");
    h->k_force = k_force < 0 ? -1 : k_force;
    return SAM_OK;
}

extern "C" int sam_set_cl(sam_handle *h, int mode)
{
    if (!h) return set_err(SAM_ERR_VALUE, "null handle");
    if (mode < 0 || mode > 2) return set_err(SAM_ERR_VALUE, "Chollet-Lesieur mode must be 0, 1 or 2");
    const int kcl = (int)h->k.k_cl;
    if (mode && (kcl < 1 || kcl > h->d.kxm || kcl > h->d.kym || kcl > h->d.kzm))
        return set_err(SAM_ERR_UNSUPPORTED, "k_cl = int(Nx/3) lies beyond the stored modes (cl_energy.f would read past u)");
    h->k.cl = mode;
    return SAM_OK;
}

extern "C" int sam_set_planes(sam_handle *h, int keep_pressure)
{
    if (!h) return set_err(SAM_ERR_VALUE, "null handle");
    h->keep_p = keep_pressure ? 1 : 0;
    return SAM_OK;
}

// one batch of planes fields: spectral fields -> z, y, x inverse transforms
// -> physical buffer Pb (nf fields); the Z column (ky = kx = 0) of the
// first Lu fields is the horizontal mean (get_mean.f)
template <typename T> static int planes_batch(sam_handle *h, int batch, int nf, std::vector<cpx<T>> *meancol)
{
    const Dims &d = h->d;
    const int pi = h->prec;
    h->z_ready = 0;
    RCHK(launch_planes_spec(d, batch, h->U[h->cur], h->R, (double *)h->planes_spec.p, h->st));
    RCHK(launch_zb<T>(d, (const double *)h->planes_spec.p, d.nzm, nf, (T *)h->Zb, h->tw[2][pi], h->st));
    if (meancol) {
        meancol->resize((size_t)d.nz * d.lu);
        for (int f = 0; f < d.lu; f++)
            HIPCHK(hipMemcpyAsync(meancol->data() + (size_t)d.nz * f, (const cpx<T> *)h->Zb + (size_t)f * d.zfield,
                                  sizeof(cpx<T>) * d.nz, hipMemcpyDeviceToHost, h->st));
    }
    RCHK(launch_yb<T>(d, (const T *)h->Zb, nf, (T *)h->Yb, h->tw[1][pi], h->st));
    RCHK(launch_xb<T>(d, (const T *)h->Yb, nf, (T *)h->Pb.p, h->tw[0][pi], h->st));
    return SAM_OK;
}

template <typename T>
static int planes_impl(sam_handle *h, int n_xy, const int *k_xy, int n_xz, const int *j_xz, int n_yz, const int *i_yz,
                       double *mean, double *xy, double *xz, double *yz)
{
    const Dims &d = h->d;
    const int lu = d.lu, nf = 2 * lu;
    const size_t pxy = (size_t)d.nx * d.ny, pxz = (size_t)d.nx * d.nz, pyz = (size_t)d.ny * d.nz;
    const size_t oxz = pxy * nf * n_xy, oyz = oxz + pxz * nf * n_xz, ntot = oyz + pyz * nf * n_yz;
    int rc = ensure(h->planes_spec, sizeof(double) * 2 * d.ss * nf);
    if (!rc) rc = ensure(h->Pb, sizeof(T) * d.pfield * nf);
    if (!rc) rc = ensure(h->planes_out, sizeof(double) * (ntot ? ntot : 1));
    if (rc) return rc;
    double *out = (double *)h->planes_out.p;
    std::vector<cpx<T>> col;
    rc = planes_batch<T>(h, 0, nf, &col);
    if (rc) return rc;
    const T *P = (const T *)h->Pb.p;
    for (int m = 0; m < n_xy; m++) RCHK(launch_plane_gather<T>(d, P, 0, k_xy[m] - 1, 0, nf, out + pxy * nf * m, 0, h->st));
    for (int m = 0; m < n_xz; m++)
        RCHK(launch_plane_gather<T>(d, P, 1, j_xz[m] - 1, 0, lu + 1, out + oxz + pxz * nf * m, 0, h->st));
    for (int m = 0; m < n_yz; m++)
        RCHK(launch_plane_gather<T>(d, P, 2, i_yz[m] - 1, 0, lu + 1, out + oyz + pyz * nf * m, 0, h->st));
    if (n_xz || n_yz) {
        rc = planes_batch<T>(h, 1, 2 * (lu - 1), nullptr);
        if (rc) return rc;
        for (int m = 0; m < n_xz; m++)     // d/dy of u, w, [T]
            RCHK(launch_plane_gather<T>(d, P, 1, j_xz[m] - 1, 0, lu - 1, out + oxz + pxz * nf * m, lu + 1, h->st));
        for (int m = 0; m < n_yz; m++)     // d/dx of v, w, [T]
            RCHK(launch_plane_gather<T>(d, P, 2, i_yz[m] - 1, lu - 1, lu - 1, out + oyz + pyz * nf * m, lu + 1, h->st));
    }
    HIPCHK(hipStreamSynchronize(h->st));
    if (n_xy) HIPCHK(hipMemcpy(xy, out, sizeof(double) * pxy * nf * n_xy, hipMemcpyDeviceToHost));
    if (n_xz) HIPCHK(hipMemcpy(xz, out + oxz, sizeof(double) * pxz * nf * n_xz, hipMemcpyDeviceToHost));
    if (n_yz) HIPCHK(hipMemcpy(yz, out + oyz, sizeof(double) * pyz * nf * n_yz, hipMemcpyDeviceToHost));
    if (mean)
        for (int f = 0; f < lu; f++)
            for (int z = 0; z < d.nz; z++) mean[(size_t)d.nz * f + z] = (double)col[(size_t)d.nz * f + z].x;
    return SAM_OK;
}

extern "C" int sam_planes(sam_handle *h, int n_xy, const int *k_xy, int n_xz, const int *j_xz, int n_yz,
                          const int *i_yz, double *mean, double *xy, double *xz, double *yz)
{
    if (!h) return set_err(SAM_ERR_VALUE, "null handle");
    if (n_xy < 0 || n_xz < 0 || n_yz < 0 || n_xy > 9 || n_xz > 9 || n_yz > 9)
        return set_err(SAM_ERR_VALUE, "0..9 planes of each kind (set_labels.f k_xy_plane1..9)");
    if ((n_xy && (!k_xy || !xy)) || (n_xz && (!j_xz || !xz)) || (n_yz && (!i_yz || !yz)))
        return set_err(SAM_ERR_VALUE, "null plane list or output");
    for (int m = 0; m < n_xy; m++)
        if (k_xy[m] < 1 || k_xy[m] > h->d.nz) return set_err(SAM_ERR_VALUE, "k_xy_plane out of 1..Nz");
    for (int m = 0; m < n_xz; m++)
        if (j_xz[m] < 1 || j_xz[m] > h->d.ny) return set_err(SAM_ERR_VALUE, "j_xz_plane out of 1..Ny");
    for (int m = 0; m < n_yz; m++)
        if (i_yz[m] < 1 || i_yz[m] > h->d.nx) return set_err(SAM_ERR_VALUE, "i_yz_plane out of 1..Nx");
    return h->prec ? planes_impl<float>(h, n_xy, k_xy, n_xz, j_xz, n_yz, i_yz, mean, xy, xz, yz)
                   : planes_impl<double>(h, n_xy, k_xy, n_xz, j_xz, n_yz, i_yz, mean, xy, xz, yz);
}

extern "C" int sam_get_spectra(sam_handle *h, double *out)
{
    if (!h || !out) return set_err(SAM_ERR_VALUE, "null argument");
    if (!h->have_spectra) return set_err(SAM_ERR_STATE, "no spectra captured (call sam_step/sam_rhs with want_spectra)");
    HIPCHK(hipStreamSynchronize(h->st));
    HIPCHK(hipMemcpy(out, h->specout, sizeof(double) * (5 * h->d.lm + 3), hipMemcpyDeviceToHost));
    return SAM_OK;
}

// ---------------------------------------------------------------------
// stage entry points
// ---------------------------------------------------------------------
// copy nelem doubles from host into a device buffer of precision T
template <typename T> static int upload_T(sam_handle *h, T *dst, const double *src, long long nelem)
{
    if (sizeof(T) == sizeof(double)) {
        HIPCHK(hipMemcpyAsync(dst, src, sizeof(double) * nelem, hipMemcpyHostToDevice, h->st));
        return SAM_OK;
    }
    int rc = ensure(h->stage1, sizeof(double) * nelem);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(h->stage1.p, src, sizeof(double) * nelem, hipMemcpyHostToDevice, h->st));
    RCHK(launch_convert_to<T>((const double *)h->stage1.p, dst, nelem, h->st));
    return SAM_OK;
}

template <typename T> static int download_T(sam_handle *h, double *dst, const T *src, long long nelem)
{
    if (sizeof(T) == sizeof(double)) {
        HIPCHK(hipMemcpyAsync(dst, src, sizeof(double) * nelem, hipMemcpyDeviceToHost, h->st));
    } else {
        int rc = ensure(h->stage1, sizeof(double) * nelem);
        if (rc) return rc;
        RCHK(launch_convert<T>(src, (double *)h->stage1.p, nelem, h->st));
        HIPCHK(hipMemcpyAsync(dst, h->stage1.p, sizeof(double) * nelem, hipMemcpyDeviceToHost, h->st));
    }
    HIPCHK(hipStreamSynchronize(h->st));
    return SAM_OK;
}

static int check_nvar(const sam_handle *h, int nvar)
{
    if (nvar < 1 || nvar > h->d.lr) return set_err(SAM_ERR_VALUE, "nvar must be 1..%d (got %d)", h->d.lr, nvar);
    return SAM_OK;
}

template <typename T> static int zb_stage(sam_handle *h, const double *u, int ldz, double *out, int nvar)
{
    const Dims &d = h->d;
    h->z_ready = 0;   // Zb is the stage buffer
    int rc = ensure(h->stage0, sizeof(double) * 2 * d.ss * nvar);
    if (!rc) rc = ensure(h->stage1, sizeof(double) * 2 * d.zfield * nvar);
    if (rc) return rc;
    rc = h2d_spec(h, (double *)h->stage0.p, u, ldz, nvar);
    if (rc) return rc;
    RCHK(launch_zb<T>(d, (const double *)h->stage0.p, d.nzm, nvar, (T *)h->Zb, h->tw[2][h->prec], h->st));
    RCHK(launch_z2ref<T>(d, (const T *)h->Zb, nvar, (double *)h->stage1.p, h->st));
    HIPCHK(hipMemcpyAsync(out, h->stage1.p, sizeof(double) * 2 * d.zfield * nvar, hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    return SAM_OK;
}

extern "C" int sam_z_trans_b(sam_handle *h, const double *u, int ldz, double *out, int nvar)
{
    if (!h || !u || !out) return set_err(SAM_ERR_VALUE, "null argument");
    int rc = check_nvar(h, nvar);
    if (rc) return rc;
    return h->prec ? zb_stage<float>(h, u, ldz, out, nvar) : zb_stage<double>(h, u, ldz, out, nvar);
}

// x2z_decomp.f on one rank: (Ny_min,Nx_min,Nz,nvar) -> (Ny,Nx/2,Nz,nvar)
extern "C" int sam_x2z(sam_handle *h, const double *in, double *out, int nvar)
{
    if (!h || !in || !out) return set_err(SAM_ERR_VALUE, "null argument");
    int rc = check_nvar(h, nvar);
    if (rc) return rc;
    const Dims &d = h->d;
    const size_t nin = (size_t)d.zfield * nvar, nout = (size_t)d.ny * (d.nx / 2) * d.nz * nvar;
    rc = ensure(h->stage0, sizeof(double) * 2 * nin);
    if (!rc) rc = ensure(h->stage1, sizeof(double) * 2 * nout);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(h->stage0.p, in, sizeof(double) * 2 * nin, hipMemcpyHostToDevice, h->st));
    RCHK(launch_x2z(d, (const double *)h->stage0.p, nvar, (double *)h->stage1.p, h->st));
    HIPCHK(hipMemcpyAsync(out, h->stage1.p, sizeof(double) * 2 * nout, hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    return SAM_OK;
}

template <typename T> static int xyb_stage(sam_handle *h, const double *in, double *out, int nvar)
{
    const Dims &d = h->d;
    h->z_ready = 0;   // Zb is the stage buffer
    const size_t nin = (size_t)d.ny * (d.nx / 2) * d.nz * nvar;
    int rc = ensure(h->Pb, sizeof(T) * d.pfield * d.lr);
    if (!rc) rc = ensure(h->stage1, sizeof(double) * 2 * nin);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(h->stage1.p, in, sizeof(double) * 2 * nin, hipMemcpyHostToDevice, h->st));
    RCHK(launch_ref2z<T>(d, (const double *)h->stage1.p, nvar, (T *)h->Zb, h->st));
    RCHK(launch_yb<T>(d, (const T *)h->Zb, nvar, (T *)h->Yb, h->tw[1][h->prec], h->st));
    RCHK(launch_xb<T>(d, (const T *)h->Yb, nvar, (T *)h->Pb.p, h->tw[0][h->prec], h->st));
    return download_T<T>(h, out, (const T *)h->Pb.p, d.pfield * nvar);
}

extern "C" int sam_xy_trans_b(sam_handle *h, const double *in, double *out, int nvar)
{
    if (!h || !in || !out) return set_err(SAM_ERR_VALUE, "null argument");
    int rc = check_nvar(h, nvar);
    if (rc) return rc;
    return h->prec ? xyb_stage<float>(h, in, out, nvar) : xyb_stage<double>(h, in, out, nvar);
}

template <typename T> static int xyf_stage(sam_handle *h, const double *in, double *out, int nvar)
{
    const Dims &d = h->d;
    h->z_ready = 0;   // Zb is the stage buffer
    int rc = ensure(h->Pb, sizeof(T) * d.pfield * d.lr);
    if (rc) return rc;
    rc = upload_T<T>(h, (T *)h->Pb.p, in, d.pfield * nvar);
    if (rc) return rc;
    RCHK(launch_xf<T>(d, (const T *)h->Pb.p, nvar, (T *)h->Yb, h->tw[0][h->prec], h->st));
    RCHK(launch_yf<T>(d, (const T *)h->Yb, nvar, (T *)h->Zb, h->tw[1][h->prec], h->st));
    return download_T<T>(h, out, (const T *)h->Zb, 2 * d.zfield * nvar);
}

extern "C" int sam_xy_trans_f(sam_handle *h, const double *in, double *out, int nvar)
{
    if (!h || !in || !out) return set_err(SAM_ERR_VALUE, "null argument");
    int rc = check_nvar(h, nvar);
    if (rc) return rc;
    return h->prec ? xyf_stage<float>(h, in, out, nvar) : xyf_stage<double>(h, in, out, nvar);
}

template <typename T> static int zf_stage(sam_handle *h, const double *in, double *out, int nvar)
{
    const Dims &d = h->d;
    h->z_ready = 0;   // Zb is the stage buffer
    int rc = ensure(h->stage0, sizeof(double) * 2 * d.ss * nvar);
    if (rc) return rc;
    rc = upload_T<T>(h, (T *)h->Zb, in, 2 * d.zfield * nvar);
    if (rc) return rc;
    RCHK(launch_zf<T>(d, (const T *)h->Zb, nvar, (double *)h->stage0.p, h->tw[2][h->prec], h->st));
    HIPCHK(hipMemcpyAsync(out, h->stage0.p, sizeof(double) * 2 * d.ss * nvar, hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    return SAM_OK;
}

extern "C" int sam_z_trans_f(sam_handle *h, const double *in, double *out, int nvar)
{
    if (!h || !in || !out) return set_err(SAM_ERR_VALUE, "null argument");
    int rc = check_nvar(h, nvar);
    if (rc) return rc;
    return h->prec ? zf_stage<float>(h, in, out, nvar) : zf_stage<double>(h, in, out, nvar);
}

// z2x_decomp.f on one rank: (Nz,Ny_min,Nx_min,nvar) -> (Nze,Ny_min,Nx_min,nvar),
// the identity reorder (Nze = Nz), through device memory as the stage calls
extern "C" int sam_z2x(sam_handle *h, const double *in, double *out, int nvar)
{
    if (!h || !in || !out) return set_err(SAM_ERR_VALUE, "null argument");
    int rc = check_nvar(h, nvar);
    if (rc) return rc;
    const size_t bytes = sizeof(double) * 2 * (size_t)h->d.zfield * nvar;
    rc = ensure(h->stage0, bytes);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(h->stage0.p, in, bytes, hipMemcpyHostToDevice, h->st));
    HIPCHK(hipMemcpyAsync(out, h->stage0.p, bytes, hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    return SAM_OK;
}

extern "C" int sam_rk_update(sam_handle *h, int nrk, double dt, double *u, const double *r_prev, const double *r,
                             int ldz)
{
    if (!h || !u) return set_err(SAM_ERR_VALUE, "null argument");
    if (nrk < 1 || nrk > h->p.nrk_max) return set_err(SAM_ERR_VALUE, "nrk must be 1..nrk_max (got %d)", nrk);
    const Dims &d = h->d;
    const size_t fb = sizeof(double) * 2 * d.ss * d.lu;
    int rc = ensure(h->stage0, 3 * fb);
    if (rc) return rc;
    double *dU = (double *)h->stage0.p, *dRp = dU + 2 * d.ss * d.lu, *dR = dRp + 2 * d.ss * d.lu;
    const bool zeta = r_prev && nrk > 1;     // sam.f:212 runs the zeta term for nrk > 1 only
    if ((rc = h2d_spec(h, dU, u, ldz, d.lu))) return rc;
    if (zeta && (rc = h2d_spec(h, dRp, r_prev, ldz, d.lu))) return rc;
    if (r && (rc = h2d_spec(h, dR, r, ldz, d.lu))) return rc;
    RCHK(launch_rk_update(d, h->k, nrk - 1, dt, dU, zeta ? dRp : nullptr, r ? dR : nullptr, h->st));
    return d2h_spec(h, u, dU, ldz, d.lu);
}

extern "C" int sam_div_check(sam_handle *h, double *out)
{
    if (!h || !out) return set_err(SAM_ERR_VALUE, "null argument");
    double *dout = h->specout + 5 * h->d.lm + 3;   // 3 spare doubles past the spectra
    RCHK(launch_div_check(h->d, h->U[h->cur], h->partials, dout, h->st));
    HIPCHK(hipMemcpyAsync(out, dout, sizeof(double) * 3, hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    return SAM_OK;
}

extern "C" int sam_products(sam_handle *h, const double *phys, double *prod)
{
    if (!h || !phys || !prod) return set_err(SAM_ERR_VALUE, "null argument");
    const Dims &d = h->d;
    int rc = ensure(h->stage0, sizeof(double) * d.pfield * (d.lu + d.lr));
    if (rc) return rc;
    double *P = (double *)h->stage0.p, *Q = P + d.pfield * d.lu;
    HIPCHK(hipMemcpyAsync(P, phys, sizeof(double) * d.pfield * d.lu, hipMemcpyHostToDevice, h->st));
    RCHK(launch_products(d, P, 1, Q, h->st));
    HIPCHK(hipMemcpyAsync(prod, Q, sizeof(double) * d.pfield * d.lr, hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    return SAM_OK;
}

extern "C" int sam_vel_max(sam_handle *h, const double *phys, double dt_in, double *stats)
{
    if (!h || !phys || !stats) return set_err(SAM_ERR_VALUE, "null argument");
    const Dims &d = h->d;
    int rc = ensure(h->stage0, sizeof(double) * d.pfield * d.lu);
    if (rc) return rc;
    DevDiag dg;
    memset(&dg, 0, sizeof dg);
    dg.dt = dt_in;
    DevDiag *sd = h->diag + 1;
    HIPCHK(hipMemcpyAsync(sd, &dg, sizeof dg, hipMemcpyHostToDevice, h->st));
    HIPCHK(hipMemcpyAsync(h->stage0.p, phys, sizeof(double) * d.pfield * d.lu, hipMemcpyHostToDevice, h->st));
    int nb = 0;
    RCHK(launch_stats_phys(d, h->stage0.p, 1, d.lu, h->partials, sd, h->st, &nb));
    RCHK(launch_finalize(d, h->k, h->partials, nb, sd, 1, h->st));
    HIPCHK(hipMemcpyAsync(&dg, sd, sizeof dg, hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    for (int i = 0; i < 4; i++) {
        stats[i] = dg.u_avg[i];
        stats[4 + i] = dg.u_var[i];
        stats[8 + i] = dg.u_max[i];
    }
    stats[12] = dg.energy;
    stats[13] = dg.spare[0];
    stats[14] = dg.spare[1];
    stats[15] = dg.spare[2];
    stats[16] = dg.dt;
    return SAM_OK;
}

// sam_vel_max on device buffers with no allocation: the partial rows and the
// DevDiag live in the device array `scratch` (its address from
// hipGetSymbolAddress at sam_create), the kernels are sam_vel_max's with the
// same chunking, so the results equal sam_vel_max's bit for bit.  Enqueued on
// the handle's stream; calls that overlap on one device share the array and
// race.
extern "C" int sam_vel_max_noalloc(sam_handle *h, const double *phys, double dt_in, double *stats)
{
    if (!h || !phys || !stats) return set_err(SAM_ERR_VALUE, "null argument");
    const Dims &d = h->d;
    const long long rows = (d.pfield + 2047) / 2048;   // bounds k_stats_phys's blocks
    if (rows > NOALLOC_ROWS)
        return set_err(SAM_ERR_SIZE,
                       "sam_vel_max_noalloc: %lld points need %lld partial rows, above the workspace's %lld "
                       "(at most %lld points)",
                       d.pfield, rows, NOALLOC_ROWS, NOALLOC_ROWS * 2048);
    double *ws = (h->dev >= 0 && h->dev < NOALLOC_MAX_DEV) ? g_scratch_ptr[h->dev] : nullptr;
    if (!ws) return set_err(SAM_ERR_STATE, "no workspace address for device %d", h->dev);
    DevDiag *dg = reinterpret_cast<DevDiag *>(ws + NOALLOC_DIAG);
    RCHK(launch_diag_seed(dg, dt_in, h->st));
    int nb = 0;
    RCHK(launch_stats_phys(d, phys, 1, d.lu, ws, dg, h->st, &nb));
    RCHK(launch_finalize(d, h->k, ws, nb, dg, 1, h->st, ws + NOALLOC_ROWS * 12));
    RCHK(launch_diag_stats(dg, stats, h->st));
    return SAM_OK;
}

extern "C" int sam_project(sam_handle *h, const double *F, const double *usav, double *r)
{
    if (!h || !F || !usav || !r) return set_err(SAM_ERR_VALUE, "null argument");
    const Dims &d = h->d;
    const size_t fb = sizeof(double) * 2 * d.ss;
    int rc = ensure(h->stage0, fb * (d.lr + d.lu + d.lu + 1));
    if (rc) return rc;
    double *dF = (double *)h->stage0.p;
    double *dU = dF + 2 * d.ss * d.lr;
    double *dR = dU + 2 * d.ss * d.lu;
    HIPCHK(hipMemcpyAsync(dF, F, fb * d.lr, hipMemcpyHostToDevice, h->st));
    HIPCHK(hipMemcpyAsync(dU, usav, fb * d.lu, hipMemcpyHostToDevice, h->st));
    RCHK(launch_project(d, h->k, dF, dU, dR, h->st));
    HIPCHK(hipMemcpyAsync(r, dR, fb * (d.lu + 1), hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    return SAM_OK;
}

template <typename T> static int eddy_stage(sam_handle *h, const double *phys, double *vt, double *cdsq)
{
    const Dims &d = h->d;
    int rc = les_bufs<T>(h);
    if (!rc) rc = ensure(h->sgsw, sizeof(T) * d.pfield);
    if (!rc) rc = ensure(h->stage0, sizeof(double) * d.pfield);
    if (rc) return rc;
    T *P = (T *)h->Pb.p, *smag = (T *)h->sgsw.p;
    rc = upload_T<T>(h, P, phys, d.pfield * d.lu);
    if (rc) return rc;
    DevDiag *sd = h->diag + 1;
    rc = sgs_coefficient<T>(h, P, sd, 1, h->p.sgs == SAM_SGS_DYN);
    if (rc) return rc;
    RCHK(launch_strain_smag<T>(d, P, smag, nullptr, d.nx / h->p.xl, d.ny / h->p.yl, d.nz / h->p.zl, h->st));
    RCHK(launch_nut<T>(d, (const T *)smag, sd, (double *)h->stage0.p, h->st));
    HIPCHK(hipMemcpyAsync(vt, h->stage0.p, sizeof(double) * d.pfield, hipMemcpyDeviceToHost, h->st));
    DevDiag dg;
    HIPCHK(hipMemcpyAsync(&dg, sd, sizeof dg, hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    if (cdsq) *cdsq = dg.spare[3];
    return SAM_OK;
}

extern "C" int sam_eddy_visc(sam_handle *h, const double *phys, double *vt, double *c_delta_sq)
{
    if (!h || !phys || !vt) return set_err(SAM_ERR_VALUE, "null argument");
    if (h->p.sgs != SAM_SGS_SMAG && h->p.sgs != SAM_SGS_DYN)
        return set_err(SAM_ERR_STATE, "handle has no eddy-viscosity model (sgs must be SMAG or DYN)");
    return h->prec ? eddy_stage<float>(h, phys, vt, c_delta_sq) : eddy_stage<double>(h, phys, vt, c_delta_sq);
}

template <typename T> static int filter_stage(sam_handle *h, const double *in, double *out, int nvar)
{
    const Dims &d = h->d;
    int rc = ensure(h->Pb, sizeof(T) * d.pfield * (3 * nvar));
    if (rc) return rc;
    T *A = (T *)h->Pb.p, *B = A + d.pfield * nvar, *tmp = B + d.pfield * nvar;   // tmp: nvar fields
    rc = upload_T<T>(h, A, in, d.pfield * nvar);
    if (rc) return rc;
    RCHK(launch_filter<T>(d, A, B, tmp, nvar, h->st));
    return download_T<T>(h, out, B, d.pfield * nvar);
}

extern "C" int sam_test_filter(sam_handle *h, const double *in, double *out, int nvar)
{
    if (!h || !in || !out) return set_err(SAM_ERR_VALUE, "null argument");
    if (nvar < 1 || nvar > 64) return set_err(SAM_ERR_VALUE, "nvar must be 1..64");
    return h->prec ? filter_stage<float>(h, in, out, nvar) : filter_stage<double>(h, in, out, nvar);
}

// ---------------------------------------------------------------------
// measurement helpers
// ---------------------------------------------------------------------
extern "C" int sam_prof_enable(sam_handle *h, int on)
{
    if (!h) return set_err(SAM_ERR_VALUE, "null handle");
    HIPCHK(hipStreamSynchronize(h->st));
    prof_collect(h);
    h->prof = on != 0;
    return SAM_OK;
}

extern "C" int sam_prof_read(sam_handle *h, double *ms, long long *count)
{
    if (!h || !ms || !count) return set_err(SAM_ERR_VALUE, "null argument");
    HIPCHK(hipStreamSynchronize(h->st));
    prof_collect(h);
    for (int i = 0; i < SAM_K_COUNT; i++) {
        ms[i] = h->prof_ms[i];
        count[i] = h->prof_cnt[i];
        h->prof_ms[i] = 0.0;
        h->prof_cnt[i] = 0;
    }
    return SAM_OK;
}

extern "C" int sam_sync(sam_handle *h)
{
    if (!h) return set_err(SAM_ERR_VALUE, "null handle");
    HIPCHK(hipDeviceSynchronize());
    prof_collect(h);
    return SAM_OK;
}

extern "C" int sam_device_copy(sam_handle *h, long long bytes, int reps)
{
    if (!h || bytes <= 0 || reps < 1) return set_err(SAM_ERR_VALUE, "bad argument");
    const long long n = (bytes + 15) / 16;
    int rc = ensure(h->stage0, (size_t)n * 16);
    if (rc) return rc;
    rc = ensure(h->stage1, (size_t)n * 16);
    if (rc) return rc;
    for (int r = 0; r < reps; r++) RCHK(launch_copy16(h->stage0.p, h->stage1.p, n, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    return SAM_OK;
}

extern "C" int sam_device_stream(sam_handle *h, int nread, int nwrite, long long bytes, int reps)
{
    if (!h || bytes <= 0 || bytes % 16384 || reps < 1) return set_err(SAM_ERR_VALUE, "bad argument");
    if (!stream_mix_supported(nread, nwrite)) return set_err(SAM_ERR_VALUE, "unsupported read:write mix");
    const long long n = bytes / 16;
    int rc = ensure(h->stage0, (size_t)(nread > 0 ? nread : 1) * (size_t)bytes);
    if (rc) return rc;
    rc = ensure(h->stage1, (size_t)(nwrite > 0 ? nwrite : 1) * (size_t)bytes);
    if (rc) return rc;
    for (int r = 0; r < reps; r++) RCHK(launch_stream(nread, nwrite, h->stage0.p, h->stage1.p, n, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    return SAM_OK;
}

// ---------------------------------------------------------------------
// learned closure
// ---------------------------------------------------------------------
struct sam_mlp {
    hipStream_t st = nullptr;
    float *P = nullptr, *norms = nullptr, *m1 = nullptr, *m2 = nullptr;
    void *img = nullptr;        // packed BF16 weight tiles, rebuilt after every parameter change
    int step = 0;
    float *feats = nullptr, *labels = nullptr;
    long long cap = 0, n = 0;
    int label_mode = SAM_LABELS_FILTERED;   // what the pool's labels are (see sam_mlp_set_labels)
    DevBuf cdiag;                           // C Delta^2 of the last SAM_LABELS_DYNAMIC_COEF snapshot
    int activation = SAM_MLP_ACT_RELU;      // hidden activation (sam_mlp_set_activation)
    float pr_t = 1.0f;
    DevBuf gpart, lpart, npart, gout, tmp, lossv;
};

static int mlp_alloc_train(sam_mlp *m, int nch)
{
    int rc = ensure(m->gpart, sizeof(float) * (size_t)MLP_NPARAM * nch);
    if (rc) return rc;
    rc = ensure(m->lpart, sizeof(double) * 1024);
    if (rc) return rc;
    return ensure(m->gout, sizeof(float) * MLP_NPARAM);
}

extern "C" int sam_mlp_create(sam_mlp **out, unsigned seed, long long cap)
{
    if (!out || cap < 0) return set_err(SAM_ERR_VALUE, "bad argument");
    sam_mlp *m = new sam_mlp();
    auto fail = [&](int code) {
        sam_mlp_destroy(m);
        return code;
    };
    if (hipStreamCreateWithFlags(&m->st, hipStreamNonBlocking) != hipSuccess)
        return fail(set_err(SAM_ERR_HIP, "hipStreamCreate failed"));
    // PyTorch-style Linear init: U(-1/sqrt(fan_in), 1/sqrt(fan_in)) for weights and biases
    std::vector<float> P(MLP_NPARAM + 32, 0.f);
    unsigned long long x = 0x9E3779B97F4A7C15ull ^ (unsigned long long)seed;
    auto rnd = [&]() {
        x ^= x << 13; x ^= x >> 7; x ^= x << 17;
        return (double)(x >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
    };
    struct L { int off, boff, out, in, ld, real_in, real_out; } layers[4] = {
        {MLP_OFF_W1, MLP_OFF_B1, 64, 16, 16, MLP_IN, 64}, {MLP_OFF_W2, MLP_OFF_B2, 64, 64, 64, 64, 64},
        {MLP_OFF_W3, MLP_OFF_B3, 64, 64, 64, 64, 64}, {MLP_OFF_W4, MLP_OFF_B4, 16, 64, 64, 64, MLP_OUT}};
    for (auto &l : layers) {
        const double bound = 1.0 / sqrt((double)l.real_in);
        for (int o = 0; o < l.out; o++) {
            for (int i = 0; i < l.in; i++)
                P[l.off + o * l.ld + i] = (o < l.real_out && i < l.real_in) ? (float)(bound * rnd()) : 0.f;
            P[l.boff + o] = o < l.real_out ? (float)(bound * rnd()) : 0.f;
        }
    }
    for (int k = 0; k < 32; k++) P[MLP_NPARAM + k] = (k < MLP_IN || (k >= 16 && k < 16 + MLP_OUT)) ? 1.f : 0.f;
    struct { float **pp; size_t n; } allocs[] = {{&m->P, MLP_NPARAM}, {&m->norms, 32}, {&m->m1, MLP_NPARAM},
                                                 {&m->m2, MLP_NPARAM}};
    for (auto &a : allocs)
        if (hipMalloc(a.pp, sizeof(float) * a.n) != hipSuccess) return fail(set_err(SAM_ERR_NOMEM, "hipMalloc failed"));
    if (hipMalloc(&m->img, MLP_TILE_BYTES) != hipSuccess) return fail(set_err(SAM_ERR_NOMEM, "hipMalloc failed"));
    if (cap > 0) {
        if (hipMalloc(&m->feats, sizeof(float) * 16 * cap) != hipSuccess ||
            hipMalloc(&m->labels, sizeof(float) * 16 * cap) != hipSuccess)
            return fail(set_err(SAM_ERR_NOMEM, "sample pool of %lld points does not fit", cap));
    }
    m->cap = cap;
    if (hipMemcpy(m->P, P.data(), sizeof(float) * MLP_NPARAM, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(m->norms, P.data() + MLP_NPARAM, sizeof(float) * 32, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(m->m1, 0, sizeof(float) * MLP_NPARAM) != hipSuccess ||
        hipMemset(m->m2, 0, sizeof(float) * MLP_NPARAM) != hipSuccess)
        return fail(set_err(SAM_ERR_HIP, "MLP upload failed"));
    if (launch_mlp_pack(m->P, m->img, m->st) != SAM_OK || hipStreamSynchronize(m->st) != hipSuccess)
        return fail(set_err(SAM_ERR_HIP, "weight packing failed"));
    *out = m;
    return SAM_OK;
}

extern "C" int sam_mlp_destroy(sam_mlp *m)
{
    if (!m) return SAM_OK;
    if (m->st) hipStreamSynchronize(m->st);
    hipFree(m->P); hipFree(m->norms); hipFree(m->m1); hipFree(m->m2); hipFree(m->img);
    hipFree(m->feats); hipFree(m->labels);
    for (DevBuf *b : {&m->gpart, &m->lpart, &m->npart, &m->gout, &m->tmp, &m->lossv, &m->cdiag}) hipFree(b->p);
    if (m->st) hipStreamDestroy(m->st);
    delete m;
    return SAM_OK;
}

extern "C" int sam_mlp_get_params(sam_mlp *m, float *params)
{
    if (!m || !params) return set_err(SAM_ERR_VALUE, "null argument");
    HIPCHK(hipStreamSynchronize(m->st));
    HIPCHK(hipMemcpy(params, m->P, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(params + MLP_NPARAM, m->norms, sizeof(float) * 32, hipMemcpyDeviceToHost));
    return SAM_OK;
}

extern "C" int sam_mlp_set_params(sam_mlp *m, const float *params)
{
    if (!m || !params) return set_err(SAM_ERR_VALUE, "null argument");
    HIPCHK(hipStreamSynchronize(m->st));
    HIPCHK(hipMemcpy(m->P, params, sizeof(float) * MLP_NPARAM, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(m->norms, params + MLP_NPARAM, sizeof(float) * 32, hipMemcpyHostToDevice));
    HIPCHK(hipMemset(m->m1, 0, sizeof(float) * MLP_NPARAM));
    HIPCHK(hipMemset(m->m2, 0, sizeof(float) * MLP_NPARAM));
    m->step = 0;
    RCHK(launch_mlp_pack(m->P, m->img, m->st));
    HIPCHK(hipStreamSynchronize(m->st));
    return SAM_OK;
}

extern "C" long long sam_mlp_pool_size(const sam_mlp *m) { return m ? m->n : 0; }

extern "C" int sam_mlp_get_samples(sam_mlp *m, long long off, long long n, float *feats, float *labels)
{
    if (!m || !feats || !labels || off < 0 || n < 0 || off + n > m->n) return set_err(SAM_ERR_VALUE, "bad range");
    std::vector<float> f((size_t)n * 16), l((size_t)n * 16);
    HIPCHK(hipStreamSynchronize(m->st));
    HIPCHK(hipMemcpy(f.data(), m->feats + 16 * off, sizeof(float) * 16 * n, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(l.data(), m->labels + 16 * off, sizeof(float) * 16 * n, hipMemcpyDeviceToHost));
    for (long long p = 0; p < n; p++) {
        for (int j = 0; j < MLP_IN; j++) feats[p * MLP_IN + j] = f[p * 16 + j];
        for (int j = 0; j < MLP_OUT; j++) labels[p * MLP_OUT + j] = l[p * 16 + j];
    }
    return SAM_OK;
}

extern "C" int sam_mlp_clear_pool(sam_mlp *m)
{
    if (!m) return set_err(SAM_ERR_VALUE, "null argument");
    m->n = 0;
    return SAM_OK;
}

template <typename T> static int physical_velocity(sam_handle *h)
{
    const Dims &d = h->d;
    int rc = ensure(h->Pb, sizeof(T) * d.pfield * (d.lu + d.lr));
    if (rc) return rc;
    const int pi = h->prec;
    h->z_ready = 0;
    RCHK(launch_zb<T>(d, h->U[h->cur], d.nzm, d.lu, (T *)h->Zb, h->tw[2][pi], h->st));
    RCHK(launch_yb<T>(d, (const T *)h->Zb, d.lu, (T *)h->Yb, h->tw[1][pi], h->st));
    RCHK(launch_xb<T>(d, (const T *)h->Yb, d.lu, (T *)h->Pb.p, h->tw[0][pi], h->st));
    return SAM_OK;
}

extern "C" int sam_mlp_add_samples(sam_mlp *m, sam_handle *dns, int r)
{
    if (!m || !dns) return set_err(SAM_ERR_VALUE, "null argument");
    const Dims &d = dns->d;
    if (r < 1 || d.nx % r || d.ny % r || d.nz % r || (d.nx / r) < 4 || (d.ny / r) < 4 || (d.nz / r) < 4)
        return set_err(SAM_ERR_VALUE, "filter ratio %d does not divide the grid into >= 4 cells", r);
    const int cx = d.nx / r, cy = d.ny / r, cz = d.nz / r;
    const long long nc = (long long)cx * cy * cz;
    if (m->n + nc > m->cap) return set_err(SAM_ERR_SIZE, "sample pool full (%lld + %lld > %lld)", m->n, nc, m->cap);
    int rc = dns->prec ? physical_velocity<float>(dns) : physical_velocity<double>(dns);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(dns->st));
    rc = ensure(m->tmp, sizeof(float) * nc * 4);
    if (rc) return rc;
    float *coarse = (float *)m->tmp.p;
    if (dns->prec)
        RCHK(launch_box_filter<float>(d, (const float *)dns->Pb.p, r, coarse, m->labels + 16 * m->n, m->st));
    else
        RCHK(launch_box_filter<double>(d, (const double *)dns->Pb.p, r, coarse, m->labels + 16 * m->n, m->st));
    RCHK(launch_coarse_features(coarse, cx, cy, cz, d.lu, (float)(cx / dns->p.xl), (float)(cy / dns->p.yl),
                                (float)(cz / dns->p.zl), m->feats + 16 * m->n, m->st));
    if (m->label_mode == SAM_LABELS_DYNAMIC || m->label_mode == SAM_LABELS_DYNAMIC_COEF) {
        const double *cs = nullptr;
        if (m->label_mode == SAM_LABELS_DYNAMIC_COEF) {
            // dm.f's dynamic procedure on the filtered field itself: C Delta^2 at
            // the coarse grid's Delta, the coefficient an LES of that grid would use
            Dims cd = dns->d;
            cd.nx = cx;
            cd.ny = cy;
            cd.nz = cz;
            cd.pfield = nc;
            if (cx < 16 || cy % 8 || cz < 8)
                return set_err(SAM_ERR_VALUE, "dynamic-coefficient labels need a filtered grid of >= 16 x 8k x 8 points");
            rc = ensure(m->npart, sizeof(double) * std::max<long long>(256 * 32, 2 * dm_fused_rows(cd)));
            if (!rc) rc = ensure(m->cdiag, sizeof(DevDiag));
            if (rc) return rc;
            int nb = 0;
            RCHK(launch_dm_fused<float>(cd, coarse, cx / dns->p.xl, cy / dns->p.yl, cz / dns->p.zl,
                                        (double *)m->npart.p, m->st, &nb));
            RCHK(launch_dm_coef((const double *)m->npart.p, nb, 0.0, 1, 1, (DevDiag *)m->cdiag.p, m->st));
            cs = &((const DevDiag *)m->cdiag.p)->c_delta_sq;
        }
        RCHK(launch_dyn_labels(m->feats + 16 * m->n, nc, 1.0f / m->pr_t, cs, m->labels + 16 * m->n, m->st));
    }
    HIPCHK(hipStreamSynchronize(m->st));
    m->n += nc;
    return SAM_OK;
}

extern "C" int sam_mlp_set_activation(sam_mlp *m, int act)
{
    if (!m) return set_err(SAM_ERR_VALUE, "null argument");
    if (act != SAM_MLP_ACT_RELU && act != SAM_MLP_ACT_GELU) return set_err(SAM_ERR_VALUE, "unknown activation %d", act);
    m->activation = act;
    return SAM_OK;
}

extern "C" int sam_mlp_get_activation(const sam_mlp *m, int *act)
{
    if (!m || !act) return set_err(SAM_ERR_VALUE, "null argument");
    *act = m->activation;
    return SAM_OK;
}

extern "C" int sam_mlp_set_labels(sam_mlp *m, int mode, double pr_t)
{
    if (!m) return set_err(SAM_ERR_VALUE, "null argument");
    if (mode != SAM_LABELS_FILTERED && mode != SAM_LABELS_DYNAMIC && mode != SAM_LABELS_DYNAMIC_COEF)
        return set_err(SAM_ERR_VALUE, "unknown label mode %d", mode);
    if (!(pr_t > 0)) return set_err(SAM_ERR_VALUE, "pr_t must be > 0");
    if (m->n > 0 && mode != m->label_mode) return set_err(SAM_ERR_STATE, "the sample pool holds labels of the other kind");
    m->label_mode = mode;
    m->pr_t = (float)pr_t;
    return SAM_OK;
}

static bool mlp_unit_coefficient(const sam_mlp *m) { return m->label_mode == SAM_LABELS_DYNAMIC; }

extern "C" int sam_mlp_last_coefficient(sam_mlp *m, double *c_delta_sq)
{
    if (!m || !c_delta_sq) return set_err(SAM_ERR_VALUE, "null argument");
    if (m->label_mode != SAM_LABELS_DYNAMIC_COEF || !m->cdiag.p)
        return set_err(SAM_ERR_STATE, "no dynamic-coefficient snapshot has been added");
    DevDiag dg;
    HIPCHK(hipStreamSynchronize(m->st));
    HIPCHK(hipMemcpy(&dg, m->cdiag.p, sizeof dg, hipMemcpyDeviceToHost));
    *c_delta_sq = dg.c_delta_sq;
    return SAM_OK;
}

extern "C" int sam_set_dyn_refresh(sam_handle *h, int nsteps)
{
    if (!h) return set_err(SAM_ERR_VALUE, "null handle");
    if (nsteps < 1) return set_err(SAM_ERR_VALUE, "refresh interval must be >= 1 step");
    h->dyn_refresh = nsteps;
    h->dyn_since = 1 << 30;
    return SAM_OK;
}

extern "C" int sam_set_mlp_refresh(sam_handle *h, int nsteps)
{
    if (!h) return set_err(SAM_ERR_VALUE, "null handle");
    if (nsteps < 1) return set_err(SAM_ERR_VALUE, "refresh interval must be >= 1 step");
    h->mlp_refresh = nsteps;
    h->mlp_since = 1 << 30;
    return SAM_OK;
}

// host [n][k] -> device [n][16] (zero padded)
static int upload_padded(sam_mlp *m, float *dst, const float *src, long long n, int k)
{
    std::vector<float> tmp((size_t)n * 16, 0.f);
    for (long long p = 0; p < n; p++)
        for (int j = 0; j < k; j++) tmp[p * 16 + j] = src[p * k + j];
    HIPCHK(hipMemcpyAsync(dst, tmp.data(), sizeof(float) * 16 * n, hipMemcpyHostToDevice, m->st));
    HIPCHK(hipStreamSynchronize(m->st));
    return SAM_OK;
}

extern "C" int sam_mlp_add_samples_host(sam_mlp *m, const float *feats, const float *labels, long long n)
{
    if (!m || !feats || !labels || n < 0) return set_err(SAM_ERR_VALUE, "bad argument");
    if (m->n + n > m->cap) return set_err(SAM_ERR_SIZE, "sample pool full");
    int rc = upload_padded(m, m->feats + 16 * m->n, feats, n, MLP_IN);
    if (rc) return rc;
    rc = upload_padded(m, m->labels + 16 * m->n, labels, n, MLP_OUT);
    if (rc) return rc;
    m->n += n;
    return SAM_OK;
}

extern "C" int sam_mlp_fit_norms(sam_mlp *m)
{
    if (!m) return set_err(SAM_ERR_VALUE, "null argument");
    if (m->n < 1) return set_err(SAM_ERR_STATE, "empty sample pool");
    int rc = ensure(m->npart, sizeof(double) * 256 * 32);
    if (rc) return rc;
    RCHK(launch_pool_norms(m->feats, m->labels, m->n, (double *)m->npart.p, m->norms, m->st));
    HIPCHK(hipStreamSynchronize(m->st));
    return SAM_OK;
}

// one forward/backward pass over `batch` samples of (feats, labels, npool)
static int mlp_pass(sam_mlp *m, const float *feats, const float *labels, long long npool, long long base,
                    long long pa, long long pb, int batch, double *loss_dst, int apply, float lr)
{
    const int nblk = mlp_step_blocks(batch), nch = nblk;
    int rc = mlp_alloc_train(m, nch);
    if (rc) return rc;
    RCHK(launch_mlp_step(m->img, m->norms, feats, labels, npool, base, pa, pb, batch, (float *)m->gpart.p,
                         (double *)m->lpart.p, m->activation, m->st));
    RCHK(launch_sum_loss((const double *)m->lpart.p, nblk, loss_dst, m->st));
    if (apply) m->step++;
    RCHK(launch_mlp_adam((const float *)m->gpart.p, nch, m->P, m->m1, m->m2, (float *)m->gout.p, lr, m->step > 0 ? m->step : 1,
                         apply, m->st));
    if (apply) RCHK(launch_mlp_pack(m->P, m->img, m->st));
    return SAM_OK;
}

extern "C" int sam_mlp_train(sam_mlp *m, int nsteps, int batch, float lr, unsigned seed, float *losses)
{
    if (!m || nsteps < 0) return set_err(SAM_ERR_VALUE, "bad argument");
    if (batch < 4096 || batch % 4096) return set_err(SAM_ERR_VALUE, "batch must be a positive multiple of 4096");
    if (m->n < 16) return set_err(SAM_ERR_STATE, "sample pool has %lld points", m->n);
    int rc = ensure(m->lossv, sizeof(double) * (nsteps > 0 ? nsteps : 1));
    if (rc) return rc;
    // multiplier coprime with the pool size: a pseudo-random walk through all samples
    static const long long primes[] = {1000003LL, 998244353LL, 1000000007LL, 2147483647LL};
    long long pa = 1;
    for (long long q : primes)
        if (m->n % q) { pa = q; break; }
    if (m->n >= (1LL << 31)) return set_err(SAM_ERR_SIZE, "sample pool above 2^31 points");
    pa %= m->n;
    const long long pb = ((long long)(seed % (unsigned)m->n) * 2654435761LL) % m->n;
    for (int s = 0; s < nsteps; s++) {
        const long long base = ((long long)s * batch + (long long)(seed % 65536u) * 7919LL) % m->n;
        rc = mlp_pass(m, m->feats, m->labels, m->n, base, pa, pb, batch, (double *)m->lossv.p + s, 1, lr);
        if (rc) return rc;
    }
    HIPCHK(hipStreamSynchronize(m->st));
    if (losses && nsteps > 0) {
        std::vector<double> l(nsteps);
        HIPCHK(hipMemcpy(l.data(), m->lossv.p, sizeof(double) * nsteps, hipMemcpyDeviceToHost));
        for (int s = 0; s < nsteps; s++) losses[s] = (float)(l[s] / ((double)batch * MLP_OUT));
    }
    return SAM_OK;
}

extern "C" int sam_mlp_grad(sam_mlp *m, const float *feats, const float *labels, int n, float *grad, double *loss)
{
    if (!m || !feats || !labels || !grad) return set_err(SAM_ERR_VALUE, "null argument");
    if (n < 4096 || n % 4096) return set_err(SAM_ERR_VALUE, "n must be a positive multiple of 4096");
    int rc = ensure(m->tmp, sizeof(float) * 32 * (size_t)n);
    if (rc) return rc;
    float *f = (float *)m->tmp.p, *l = f + 16 * (size_t)n;
    if ((rc = upload_padded(m, f, feats, n, MLP_IN)) || (rc = upload_padded(m, l, labels, n, MLP_OUT))) return rc;
    rc = ensure(m->lossv, sizeof(double));
    if (rc) return rc;
    rc = mlp_pass(m, f, l, n, 0, 1, 0, n, (double *)m->lossv.p, 0, 0.f);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(m->st));
    HIPCHK(hipMemcpy(grad, m->gout.p, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost));
    double ls = 0;
    HIPCHK(hipMemcpy(&ls, m->lossv.p, sizeof(double), hipMemcpyDeviceToHost));
    if (loss) *loss = ls / ((double)n * MLP_OUT);
    return SAM_OK;
}

extern "C" int sam_mlp_infer(sam_mlp *m, const float *feats, float *out, long long n)
{
    if (!m || !feats || !out || n < 0) return set_err(SAM_ERR_VALUE, "bad argument");
    int rc = ensure(m->tmp, sizeof(float) * 32 * (size_t)(n + 16));
    if (rc) return rc;
    float *f = (float *)m->tmp.p, *o = f + 16 * (size_t)(n + 16);
    if ((rc = upload_padded(m, f, feats, n, MLP_IN))) return rc;
    RCHK(launch_mlp_infer(m->P, m->img, m->norms, f, o, n, m->activation, m->st));
    std::vector<float> tmp((size_t)n * 16);
    HIPCHK(hipMemcpyAsync(tmp.data(), o, sizeof(float) * 16 * n, hipMemcpyDeviceToHost, m->st));
    HIPCHK(hipStreamSynchronize(m->st));
    for (long long p = 0; p < n; p++)
        for (int j = 0; j < MLP_OUT; j++) out[p * MLP_OUT + j] = tmp[p * 16 + j];
    return SAM_OK;
}

extern "C" int sam_attach_mlp(sam_handle *h, sam_mlp *m)
{
    if (!h) return set_err(SAM_ERR_VALUE, "null handle");
    h->mlp = m;
    h->mlp_since = 1 << 30;   // a unit-coefficient network gets C Delta^2 at the next step
    return SAM_OK;
}

extern "C" int sam_set_sgs(sam_handle *h, int sgs)
{
    if (!h) return set_err(SAM_ERR_VALUE, "null handle");
    if (sgs < SAM_SGS_NONE || sgs > SAM_SGS_MLP) return set_err(SAM_ERR_VALUE, "unknown sgs model %d", sgs);
    if (sgs == SAM_SGS_MLP && !h->mlp) return set_err(SAM_ERR_STATE, "attach a network first");
    if (sgs != SAM_SGS_NONE && !(h->p.pr_t > 0)) return set_err(SAM_ERR_VALUE, "pr_t must be > 0");
    h->p.sgs = sgs;
    h->sgs_active = sgs;
    return SAM_OK;
}

template <typename T> static int mlp_flux(sam_handle *h, const T *P, T *Q)
{
    const Dims &d = h->d;
    HIPCHK(hipStreamSynchronize(h->mlp->st));   // parameters written on the network's stream
    const double *cs = mlp_unit_coefficient(h->mlp) ? &h->diag->c_delta_sq : nullptr;
    RCHK(launch_les_mlp_flux<T>(d, h->mlp->img, h->mlp->norms, P, Q, (float)(d.nx / h->p.xl), (float)(d.ny / h->p.yl),
                                (float)(d.nz / h->p.zl), cs, h->mlp->activation, h->st));
    return SAM_OK;
}

extern "C" int sam_sgs_check(sam_handle *h, double tol, double *err, int *switched)
{
    if (!h) return set_err(SAM_ERR_VALUE, "null handle");
    if (switched) *switched = 0;
    if (err) *err = 0.0;
    if (h->p.sgs != SAM_SGS_MLP) return SAM_OK;
    const int lm = h->d.lm;
    std::vector<double> spec(5 * lm + 3), tk[3];
    Snapshot s;
    int rc = snap_save(h, s);
    if (rc) return rc;
    // each pass starts from the saved handle: the network's pass is the first
    // substep the run would take next (its C Delta^2 and refresh counter), the
    // dynamic model's pass always runs the procedure
    const int models[3] = {SAM_SGS_NONE, SAM_SGS_DYN, SAM_SGS_MLP};
    for (int i = 0; i < 3 && !rc; i++) {
        h->p.sgs = models[i];
        if (models[i] == SAM_SGS_DYN) h->dyn_since = 1 << 30;
        rc = rhs_pass(h, nullptr, h->d.nzm, 1, nullptr, false);
        if (!rc) rc = sam_get_spectra(h, spec.data());
        if (!rc) tk[i].assign(spec.begin() + lm, spec.begin() + 2 * lm);
        const int rc2 = snap_restore(h, s);
        if (!rc) rc = rc2;
    }
    if (rc) return rc;
    double num = 0.0, den = 0.0;
    for (int k = 0; k < lm; k++) {
        const double td = tk[1][k] - tk[0][k], tm = tk[2][k] - tk[0][k];
        num += (tm - td) * (tm - td);
        den += td * td;
    }
    const double e = den > 0 ? sqrt(num / den) : 0.0;
    if (err) *err = e;
    if (e > tol) {
        h->p.sgs = SAM_SGS_DYN;
        h->sgs_active = SAM_SGS_DYN;
        if (switched) *switched = 1;
    }
    return SAM_OK;
}
