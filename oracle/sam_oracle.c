/*
 * sam_oracle.c -- CPU FP64 restatement of SAM's RK3 pseudo-spectral step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.  The
 * product path (sam_amd/, libsamhip.so) never links, loads or calls it.
 *
 * Every routine restates the algorithm of the reference file cited beside
 * it, in the reference's own array layouts (Fortran column-major, first
 * index fastest, 0-based here), for one MPI rank:
 *
 *   spectral  u(ldz, Ny_min, Nx_min, n)   complex   (ldz = Nze or Nz_min)
 *   z_trans_b output (Ny_min, Nx_min, Nz, n)        complex
 *   x2z output       (Ny, Nx/2, Nz, n)              complex
 *   physical         (Nx, Ny, Nz, n)                real
 *   xy_trans_f out   (Nz, Ny_min, Nx_min, n)        complex
 *
 * FFT conventions follow FFTPACK (fft.f): forward kernel exp(-i..),
 * backward exp(+i..), both unnormalised; real transforms use FFTPACK's
 * half-complex packing r(1)=X0, r(2k)=Re Xk, r(2k+1)=Im Xk, r(N)=X_{N/2}
 * (fft.f:2259-2266, 2377-2390).  The FFT itself is an independent
 * Stockham mixed-radix implementation with long-double twiddle tables.
 *
 * Pinning: FFTPACK known answers (fft.f:2897-3011), the reference's own
 * fft.f / transform routines compiled into oracle/_ref, brute-force DFTs,
 * and closed-form solutions (see tests/test_oracle_*.py).
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cx;

/* ------------------------------------------------------------------ */
/* FFT: Stockham autosort, mixed radix (4,2,3,5,generic), unnormalised  */
/* ------------------------------------------------------------------ */

typedef struct {
    int n, nf, fac[40];
    cx *tw; /* tw[k] = exp(-2 pi i k / n), computed in long double */
} fftplan;

static void plan_init(fftplan *p, int n)
{
    p->n = n;
    p->nf = 0;
    int m = n;
    while (m % 4 == 0) { p->fac[p->nf++] = 4; m /= 4; }
    while (m % 2 == 0) { p->fac[p->nf++] = 2; m /= 2; }
    for (int q = 3; m > 1; q += 2)
        while (m % q == 0) { p->fac[p->nf++] = q; m /= q; }
    p->tw = (cx *)malloc(sizeof(cx) * (size_t)(n > 0 ? n : 1));
    for (int k = 0; k < n; k++) {
        long double a = -2.0L * 3.14159265358979323846264338327950288L * (long double)k / (long double)n;
        p->tw[k] = (double)cosl(a) + I * (double)sinl(a);
    }
}

static void plan_free(fftplan *p) { free(p->tw); p->tw = NULL; }

/* exp(sign * 2 pi i * e / n), e taken modulo n */
static inline cx twid(const fftplan *p, long e, int sign)
{
    cx w = p->tw[e % p->n];
    return sign < 0 ? w : conj(w);
}

/* x: length n in/out; work: length n scratch.  sign=-1 forward, +1 backward. */
static void cfft(const fftplan *p, cx *x, cx *work, int sign)
{
    const int n = p->n;
    if (n <= 1) return;
    cx *a = x, *b = work;
    int l = 1;
    cx v[64], o[64];
    for (int f = 0; f < p->nf; f++) {
        const int q = p->fac[f];
        const int nq = n / q;
        const long step = n / (l * q); /* twiddle exponent unit for this pass */
        for (int j = 0; j < nq; j++) {
            const int jl = j % l;
            for (int r = 0; r < q; r++) {
                cx t = a[j + r * nq];
                if (jl && r) t *= twid(p, (long)r * jl * step, sign);
                v[r] = t;
            }
            if (q == 4) {
                cx s0 = v[0] + v[2], d0 = v[0] - v[2];
                cx s1 = v[1] + v[3], d1 = v[1] - v[3];
                cx jd1 = (sign < 0) ? (cimag(d1) - I * creal(d1)) : (-cimag(d1) + I * creal(d1));
                o[0] = s0 + s1; o[2] = s0 - s1;
                o[1] = d0 + jd1; o[3] = d0 - jd1;
            } else if (q == 2) {
                o[0] = v[0] + v[1]; o[1] = v[0] - v[1];
            } else {
                for (int k = 0; k < q; k++) {
                    cx s = 0;
                    for (int r = 0; r < q; r++)
                        s += v[r] * twid(p, (long)((r * k) % q) * nq, sign);
                    o[k] = s;
                }
            }
            const int idxD = (j / l) * l * q + jl;
            for (int r = 0; r < q; r++) b[idxD + r * l] = o[r];
        }
        cx *t = a; a = b; b = t;
        l *= q;
    }
    if (a != x) memcpy(x, a, sizeof(cx) * (size_t)n);
}

/* FFTPACK-convention real transforms of even length n, via a length-n/2
 * complex transform.  ph: plan of length n/2; pn: plan of length n (twiddles).
 * rfftf: r <- half-complex forward transform of r (fft.f RFFTF semantics).
 * rfftb: r <- real backward transform of half-complex r (fft.f RFFTB). */
static void rfftf_pk(const fftplan *ph, const fftplan *pn, double *r, cx *w1, cx *w2)
{
    const int n = pn->n, h = n / 2;
    for (int j = 0; j < h; j++) w1[j] = r[2 * j] + I * r[2 * j + 1];
    cfft(ph, w1, w2, -1);
    /* X_k = E_k + w^k O_k, E_k=(Z_k+conj Z_{h-k})/2, O_k=(Z_k-conj Z_{h-k})/(2i) */
    for (int k = 0; k <= h; k++) {
        cx zk = w1[k % h], zc = conj(w1[(h - k) % h]);
        cx e = 0.5 * (zk + zc), o = -0.5 * I * (zk - zc);
        cx X = e + twid(pn, k, -1) * o;
        if (k == 0) r[0] = creal(X);
        else if (k == h) r[n - 1] = creal(X);
        else { r[2 * k - 1] = creal(X); r[2 * k] = cimag(X); }
    }
}

static void rfftb_pk(const fftplan *ph, const fftplan *pn, double *r, cx *w1, cx *w2)
{
    const int n = pn->n, h = n / 2;
    /* X_k, k=0..h from half-complex storage */
    for (int k = 0; k < h; k++) {
        cx xk = (k == 0) ? r[0] : (r[2 * k - 1] + I * r[2 * k]);
        int m = h - k;
        cx xm = (m == h) ? r[n - 1] : (m == 0 ? r[0] : (r[2 * m - 1] + I * r[2 * m]));
        cx xmc = conj(xm);
        w1[k] = (xk + xmc) + I * twid(pn, k, +1) * (xk - xmc);
    }
    cfft(ph, w1, w2, +1);
    for (int j = 0; j < h; j++) { r[2 * j] = creal(w1[j]); r[2 * j + 1] = cimag(w1[j]); }
}

/* ------------------------------------------------------------------ */
/* Public FFT entry points (FFTPACK semantics), used to pin the oracle  */
/* ------------------------------------------------------------------ */

void orc_cfft(int n, double *x, int sign)
{
    fftplan p; plan_init(&p, n);
    cx *w = (cx *)malloc(sizeof(cx) * (size_t)n);
    cfft(&p, (cx *)x, w, sign);
    free(w); plan_free(&p);
}

void orc_rfft(int n, double *r, int dir)
{
    fftplan ph, pn; plan_init(&ph, n / 2); plan_init(&pn, n);
    cx *w1 = (cx *)malloc(sizeof(cx) * (size_t)(n / 2 + 1));
    cx *w2 = (cx *)malloc(sizeof(cx) * (size_t)(n / 2 + 1));
    if (dir < 0) rfftf_pk(&ph, &pn, r, w1, w2); else rfftb_pk(&ph, &pn, r, w1, w2);
    free(w1); free(w2); plan_free(&ph); plan_free(&pn);
}

/* ------------------------------------------------------------------ */
/* Configuration and setup (init.f:35-85,97-148; fft_init.f:43-78)      */
/* ------------------------------------------------------------------ */

typedef struct {
    int nx, ny, nz;
    int n_dealias;   /* 0,1,2 (init.f:49-62) */
    int i_strat;     /* 0: Lu=3/Lr=6, 1: Lu=4/Lr=9 */
    int nrk_max;     /* 1,2,3 */
    int i_cfl;       /* 1: dt from CFL (vel_max.f:43-46) */
    int sgs;         /* 0 none, 1 Smagorinsky (sm.f), 2 dynamic (dm.f) */
    double xl, yl, zl;
    double vis, pr, grav, to, lapse0;
    double lapse;    /* value used by rhs.f:133 (lapse0 fresh, 0 after restart) */
    double cfl0, dt0;
    double c_smag;   /* sm.f: nu_t = c_smag * Delta^2 * |S| */
    double pr_t;     /* turbulent Prandtl number for the SGS scalar flux */
    int k_force;     /* force.f after every step when >= 0 (i_force = 1) */
    int i_cl;        /* Chollet-Lesieur: 0 off, 1 diagnostics (i_les = 1), 2 nu_t(k) in the IF */
} orc_cfg;

typedef struct {
    orc_cfg c;
    int kxm, kym, kzm, nxm, nym, nzm, lu, lr, lm, nxh;
    double wave_x, wave_y, wave_z;
    double gam[4], zet[4], adv_cfl_limit;
    double buoy_fac_x, buoy_fac_z, pr_inv;
    double delta_sq;
    int *ky, *kz;          /* integer wavenumber maps, fft_init.f:56-72 */
    double a_cl, b_cl, c_cl; int k_cl;   /* init.f:88-93 */
    fftplan px, pxh, py, pz;
    int nthreads;
} orc;

static int kmax_of(int n, int mode, int dir)
{
    /* init.f:49-62; n_dealias=2 keeps x untruncated */
    if (mode == 0) return n / 2 - 1;
    if (mode == 2 && dir == 0) return n / 2 - 1;
    return (int)(((double)n - 0.1) / 3.0);
}

orc *orc_create(const orc_cfg *cfg)
{
    orc *o = (orc *)calloc(1, sizeof(orc));
    o->c = *cfg;
    o->kxm = kmax_of(cfg->nx, cfg->n_dealias, 0);
    o->kym = kmax_of(cfg->ny, cfg->n_dealias, 1);
    o->kzm = kmax_of(cfg->nz, cfg->n_dealias, 2);
    o->nxm = o->kxm + 1;
    o->nym = 2 * o->kym + 1;
    o->nzm = 2 * o->kzm + 1;
    o->nxh = cfg->nx / 2;
    o->lu = cfg->i_strat ? 4 : 3;
    o->lr = cfg->i_strat ? 9 : 6;
    int ksq = o->kxm * o->kxm + o->kym * o->kym + o->kzm * o->kzm;
    o->lm = (int)(sqrt((double)ksq) + 0.5) + 1; /* sam.f:72-73 */
    const double pi = acos(-1.0);
    o->wave_x = 2.0 * pi / cfg->xl;
    o->wave_y = 2.0 * pi / cfg->yl;
    o->wave_z = 2.0 * pi / cfg->zl;
    for (int i = 0; i < 4; i++) o->gam[i] = o->zet[i] = 0.0;
    if (cfg->nrk_max == 1) { o->gam[0] = 1.0; o->adv_cfl_limit = 1.0; }
    if (cfg->nrk_max == 2) { o->gam[0] = 1.0; o->gam[1] = 0.5; o->zet[1] = -0.5; o->adv_cfl_limit = 1.0; }
    if (cfg->nrk_max == 3) {
        o->gam[0] = 8.0 / 15.0; o->gam[1] = 5.0 / 12.0; o->gam[2] = 3.0 / 4.0;
        o->zet[1] = -17.0 / 60.0; o->zet[2] = -5.0 / 12.0;
        o->adv_cfl_limit = sqrt(3.0);
    }
    o->pr_inv = 1.0 / cfg->pr;
    o->a_cl = 0.145; o->b_cl = 5.01; o->c_cl = 3.03;   /* init.f:88-93 */
    o->k_cl = (int)(cfg->nx / 3.0);
    o->buoy_fac_x = 0.0;
    o->buoy_fac_z = cfg->i_strat ? cfg->grav / cfg->to : 0.0;
    double dx = cfg->xl / cfg->nx, dy = cfg->yl / cfg->ny, dz = cfg->zl / cfg->nz;
    o->delta_sq = pow(dx * dy * dz, 2.0 / 3.0);
    o->ky = (int *)malloc(sizeof(int) * o->nym);
    o->kz = (int *)malloc(sizeof(int) * o->nzm);
    for (int j = 0; j < o->nym; j++) o->ky[j] = (j <= o->kym) ? j : j - o->nym;
    for (int k = 0; k < o->nzm; k++) o->kz[k] = (k <= o->kzm) ? k : k - o->nzm;
    plan_init(&o->px, cfg->nx);
    plan_init(&o->pxh, cfg->nx / 2);
    plan_init(&o->py, cfg->ny);
    plan_init(&o->pz, cfg->nz);
#ifdef _OPENMP
    o->nthreads = omp_get_max_threads();
#else
    o->nthreads = 1;
#endif
    return o;
}

void orc_destroy(orc *o)
{
    if (!o) return;
    free(o->ky); free(o->kz);
    plan_free(&o->px); plan_free(&o->pxh); plan_free(&o->py); plan_free(&o->pz);
    free(o);
}

/* out[0..15]: kxm kym kzm nxm nym nzm lu lr lm nxh */
void orc_dims(const orc *o, int *out)
{
    out[0] = o->kxm; out[1] = o->kym; out[2] = o->kzm;
    out[3] = o->nxm; out[4] = o->nym; out[5] = o->nzm;
    out[6] = o->lu; out[7] = o->lr; out[8] = o->lm; out[9] = o->nxh;
}

/* out: wave_x wave_y wave_z gamma[3] zeta[3] adv_cfl_limit buoy_fac_z pr_inv delta_sq */
void orc_consts(const orc *o, double *out)
{
    out[0] = o->wave_x; out[1] = o->wave_y; out[2] = o->wave_z;
    for (int i = 0; i < 3; i++) { out[3 + i] = o->gam[i]; out[6 + i] = o->zet[i]; }
    out[9] = o->adv_cfl_limit; out[10] = o->buoy_fac_z; out[11] = o->pr_inv; out[12] = o->delta_sq;
}

void orc_set_threads(orc *o, int n)
{
#ifdef _OPENMP
    if (n > 0) { omp_set_num_threads(n); o->nthreads = n; }
#else
    (void)o; (void)n;
#endif
}

/* ------------------------------------------------------------------ */
/* Transforms (periodic case 0)                                         */
/* ------------------------------------------------------------------ */

/* z_trans_b.f:31-52: zero-pad in z, cfftb(Nz), scatter to (Ny_min,Nx_min,Nz,n) */
void orc_z_trans_b(const orc *o, const cx *u, int ldz, cx *out, int nvar)
{
    const int nz = o->c.nz, nym = o->nym, nxm = o->nxm, nzm = o->nzm, kzm = o->kzm;
    const long ncol = (long)nvar * nxm * nym;
#pragma omp parallel
    {
        cx *c = (cx *)malloc(sizeof(cx) * nz), *w = (cx *)malloc(sizeof(cx) * nz);
#pragma omp for schedule(static)
        for (long col = 0; col < ncol; col++) {
            const int iy = (int)(col % nym);
            const int ix = (int)((col / nym) % nxm);
            const int n = (int)(col / ((long)nym * nxm));
            const cx *src = u + (size_t)ldz * (iy + (size_t)nym * (ix + (size_t)nxm * n));
            for (int p = 0; p < nz; p++) {
                if (p <= kzm) c[p] = src[p];
                else if (p >= nz - kzm) c[p] = src[p - nz + nzm];
                else c[p] = 0;
            }
            cfft(&o->pz, c, w, +1);
            for (int z = 0; z < nz; z++)
                out[iy + (size_t)nym * (ix + (size_t)nxm * (z + (size_t)nz * n))] = c[z];
        }
        free(c); free(w);
    }
}

/* x2z_decomp.f:59-152 on one rank: (Ny_min,Nx_min,Nz,n) -> (Ny,Nx/2,Nz,n),
 * zero-padding y; x-modes beyond Nx_min are set to zero here. */
void orc_x2z(const orc *o, const cx *in, cx *out, int nvar)
{
    const int ny = o->c.ny, nz = o->c.nz, nym = o->nym, nxm = o->nxm, nxh = o->nxh, kym = o->kym;
    const long nplane = (long)nvar * nz;
#pragma omp parallel for schedule(static)
    for (long pl = 0; pl < nplane; pl++) {
        for (int ix = 0; ix < nxh; ix++) {
            cx *dst = out + (size_t)ny * (ix + (size_t)nxh * pl);
            if (ix >= nxm) { for (int y = 0; y < ny; y++) dst[y] = 0; continue; }
            const cx *src = in + (size_t)nym * (ix + (size_t)nxm * pl);
            for (int y = 0; y < ny; y++) {
                if (y <= kym) dst[y] = src[y];
                else if (y >= ny - kym) dst[y] = src[y - ny + nym];
                else dst[y] = 0;
            }
        }
    }
}

/* xy_trans_b.f:11-63: cfftb(Ny) per x-mode (in place on `in`), pack the
 * half-complex x row (Re only for kx=0, xy_trans_b.f:28-31), zero the
 * x modes above kx_max, rfftb(Nx) -> physical (Nx,Ny,Nz,n). */
void orc_xy_trans_b(const orc *o, cx *in, double *out, int nvar)
{
    const int nx = o->c.nx, ny = o->c.ny, nz = o->c.nz, nxh = o->nxh, kxm = o->kxm;
    const long nplane = (long)nvar * nz;
#pragma omp parallel
    {
        cx *w = (cx *)malloc(sizeof(cx) * (ny > nx ? ny : nx));
        cx *w2 = (cx *)malloc(sizeof(cx) * (nx / 2 + 1));
        double *row = (double *)malloc(sizeof(double) * nx);
#pragma omp for schedule(static)
        for (long pl = 0; pl < nplane; pl++) {
            cx *pin = in + (size_t)ny * nxh * pl;
            for (int ix = 0; ix <= kxm; ix++) cfft(&o->py, pin + (size_t)ny * ix, w, +1);
            for (int y = 0; y < ny; y++) {
                row[0] = creal(pin[y]);
                for (int ix = 1; ix <= kxm; ix++) {
                    cx v = pin[y + (size_t)ny * ix];
                    row[2 * ix - 1] = creal(v);
                    row[2 * ix] = cimag(v);
                }
                for (int i = 2 * kxm + 1; i < nx; i++) row[i] = 0.0;
                rfftb_pk(&o->pxh, &o->px, row, w, w2);
                memcpy(out + (size_t)nx * (y + (size_t)ny * pl), row, sizeof(double) * nx);
            }
        }
        free(w); free(w2); free(row);
    }
}

/* xy_trans_f.f:20-57: rfftf(Nx) per row (in place on `in`), x1/(NxNy),
 * keep kx <= kx_max, cfftf(Ny), keep |ky| <= ky_max -> (Nz,Ny_min,Nx_min,n) */
void orc_xy_trans_f(const orc *o, double *in, cx *out, int nvar)
{
    const int nx = o->c.nx, ny = o->c.ny, nz = o->c.nz, nxm = o->nxm, nym = o->nym, kym = o->kym;
    const double fn = 1.0 / ((double)nx * (double)ny);
    const long nplane = (long)nvar * nz;
#pragma omp parallel
    {
        cx *wk = (cx *)malloc(sizeof(cx) * (size_t)ny * nxm);
        cx *w = (cx *)malloc(sizeof(cx) * (ny > nx ? ny : nx));
        cx *w2 = (cx *)malloc(sizeof(cx) * (nx / 2 + 1));
#pragma omp for schedule(static)
        for (long pl = 0; pl < nplane; pl++) {
            const int z = (int)(pl % nz), n = (int)(pl / nz);
            double *pin = in + (size_t)nx * ny * pl;
            for (int y = 0; y < ny; y++) {
                double *row = pin + (size_t)nx * y;
                rfftf_pk(&o->pxh, &o->px, row, w, w2);
                wk[y] = row[0] * fn;
                for (int ix = 1; ix < nxm; ix++)
                    wk[y + (size_t)ny * ix] = (row[2 * ix - 1] + I * row[2 * ix]) * fn;
            }
            for (int ix = 0; ix < nxm; ix++) {
                cx *col = wk + (size_t)ny * ix;
                cfft(&o->py, col, w, -1);
                for (int iy = 0; iy < nym; iy++) {
                    int y = (iy <= kym) ? iy : iy + ny - nym;
                    out[z + (size_t)nz * (iy + (size_t)nym * (ix + (size_t)nxm * n))] = col[y];
                }
            }
        }
        free(wk); free(w); free(w2);
    }
}

/* z_trans_f.f:21-31: cfftf(Nz) in place, x1/Nz, truncate to the first
 * Nz_min rows of each column (layout (ldz,Ny_min,Nx_min,n), ldz >= Nz). */
void orc_z_trans_f(const orc *o, cx *u, int ldz, int nvar)
{
    const int nz = o->c.nz, nym = o->nym, nxm = o->nxm, nzm = o->nzm, kzm = o->kzm;
    const double nzinv = 1.0 / (double)nz;
    const long ncol = (long)nvar * nxm * nym;
#pragma omp parallel
    {
        cx *w = (cx *)malloc(sizeof(cx) * nz);
#pragma omp for schedule(static)
        for (long col = 0; col < ncol; col++) {
            cx *c = u + (size_t)ldz * col;
            cfft(&o->pz, c, w, -1);
            for (int k = 0; k <= kzm; k++) c[k] = c[k] * nzinv;
            for (int k = kzm + 1; k < nzm; k++) c[k] = c[k + nz - nzm] * nzinv;
        }
        free(w);
    }
}

/* ------------------------------------------------------------------ */
/* Physical-space pieces: vel_max.f and the products of rhs.f:32-61     */
/* ------------------------------------------------------------------ */

/* stats[0..3] u_avg, [4..7] u_var, [8..11] u_max, [12] energy,
 * [13..15] cfl_x/y/z_by_dt, [16] dt (updated when i_cfl=1; else dt_in). */
void orc_vel_max(const orc *o, const double *phys, double dt_in, double *stats)
{
    const size_t np = (size_t)o->c.nx * o->c.ny * o->c.nz;
    for (int i = 0; i < 17; i++) stats[i] = 0.0;
    for (int n = 0; n < o->lu; n++) {
        const double *f = phys + np * n;
        double s = 0, s2 = 0, mx = 0;
#pragma omp parallel for reduction(+ : s, s2) reduction(max : mx) schedule(static)
        for (size_t i = 0; i < np; i++) {
            double v = f[i];
            s += v; s2 += v * v;
            double a = fabs(v);
            if (a > mx) mx = a;
        }
        const double inv = 1.0 / (double)np;
        stats[n] = s * inv;
        stats[4 + n] = s2 * inv - stats[n] * stats[n];
        stats[8 + n] = mx;
    }
    stats[12] = 0.5 * (stats[4] + stats[5] + stats[6]);
    stats[13] = stats[8] * o->kxm * o->wave_x;
    stats[14] = stats[9] * o->kym * o->wave_y;
    stats[15] = stats[10] * o->kzm * o->wave_z;
    double dt = dt_in;
    if (o->c.i_cfl == 1) {
        double m = stats[13];
        if (stats[14] > m) m = stats[14];
        if (stats[15] > m) m = stats[15];
        dt = o->c.cfl0 * o->adv_cfl_limit / m;
    }
    stats[16] = dt;
}

/* rhs.f:32-61: uu,uv,vv,uw,vw,ww (+uT,vT,wT) */
void orc_products(const orc *o, const double *phys, double *prod)
{
    const size_t np = (size_t)o->c.nx * o->c.ny * o->c.nz;
    const double *u = phys, *v = phys + np, *w = phys + 2 * np, *t = phys + 3 * np;
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < np; i++) {
        prod[i] = u[i] * u[i];
        prod[np + i] = u[i] * v[i];
        prod[2 * np + i] = v[i] * v[i];
        prod[3 * np + i] = u[i] * w[i];
        prod[4 * np + i] = v[i] * w[i];
        prod[5 * np + i] = w[i] * w[i];
        if (o->lu == 4) {
            prod[6 * np + i] = u[i] * t[i];
            prod[7 * np + i] = v[i] * t[i];
            prod[8 * np + i] = w[i] * t[i];
        }
    }
}

/* ------------------------------------------------------------------ */
/* Spectral RHS: divergence, buoyancy, projection (rhs.f:74-142)        */
/* F: (ldf,Ny_min,Nx_min,Lr) forward-transformed fluxes (first Nz_min   */
/* rows valid).  usav: (Nz_min,Ny_min,Nx_min,Lu) state before the zeta  */
/* update (sam.f:175-206).  r: (Nz_min,Ny_min,Nx_min,Lu+1), last = p.   */
/* ------------------------------------------------------------------ */
void orc_project(const orc *o, const cx *F, int ldf, const cx *usav, cx *r)
{
    const int nxm = o->nxm, nym = o->nym, nzm = o->nzm;
    const size_t sf = (size_t)ldf * nym * nxm;       /* field stride in F */
    const size_t ss = (size_t)nzm * nym * nxm;       /* field stride in S layout */
    const int strat = (o->lu == 4);
#pragma omp parallel for schedule(static)
    for (int ix = 0; ix < nxm; ix++) {
        const double rkx = o->wave_x * (double)ix;
        const double rkx2 = rkx * rkx;
        for (int iy = 0; iy < nym; iy++) {
            const double rky = o->wave_y * (double)o->ky[iy];
            const double rk_sq2 = rkx2 + rky * rky;
            for (int k = 0; k < nzm; k++) {
                const double rkz = o->wave_z * (double)o->kz[k];
                double rk_sq = rk_sq2 + rkz * rkz;
                if (rk_sq < 1.0e-10) rk_sq = 1.0e-10;
                const size_t fi = k + (size_t)ldf * (iy + (size_t)nym * ix);
                const size_t si = k + (size_t)nzm * (iy + (size_t)nym * ix);
                cx f[9];
                for (int q = 0; q < o->lr; q++) f[q] = F[fi + sf * q];
                cx r1 = -I * (rkx * f[0] + rky * f[1] + rkz * f[3]);
                cx r2 = -I * (rkx * f[1] + rky * f[2] + rkz * f[4]);
                cx r3 = -I * (rkx * f[3] + rky * f[4] + rkz * f[5]);
                cx r4 = 0;
                if (strat) {
                    const cx tsav = usav[si + 3 * ss], wsav = usav[si + 2 * ss];
                    r1 += o->buoy_fac_x * tsav;
                    r3 += o->buoy_fac_z * tsav;
                    r4 = -I * (rkx * f[6] + rky * f[7] + rkz * f[8]) - o->c.lapse * wsav;
                }
                cx p = -I * (rkx * r1 + rky * r2 + rkz * r3) / rk_sq;
                r1 -= I * rkx * p;
                r2 -= I * rky * p;
                r3 -= I * rkz * p;
                r[si] = r1; r[si + ss] = r2; r[si + 2 * ss] = r3;
                if (strat) r[si + 3 * ss] = r4;
                r[si + (size_t)o->lu * ss] = p;
            }
        }
    }
}

/* ------------------------------------------------------------------ */
/* SGS models on the collocated grid (sm.f, dm.f, filter.f)             */
/* ------------------------------------------------------------------ */

static inline size_t pidx(const orc *o, int x, int y, int z)
{
    const int nx = o->c.nx, ny = o->c.ny, nz = o->c.nz;
    x = (x + nx) % nx; y = (y + ny) % ny; z = (z + nz) % nz;
    return x + (size_t)nx * (y + (size_t)ny * z);
}

/* Strain rate S_n (n: 11,22,33,12,23,13) by centred differences as in
 * sm.f:52-69 / dm.f:80-99, applied to collocated point values. */
static void strain_at(const orc *o, const double *phys, int x, int y, int z, double s[6])
{
    const size_t np = (size_t)o->c.nx * o->c.ny * o->c.nz;
    const double *u = phys, *v = phys + np, *w = phys + 2 * np;
    const double dxi = o->c.nx / o->c.xl, dyi = o->c.ny / o->c.yl, dzi = o->c.nz / o->c.zl;
#define D(f, a, b, c) (f[pidx(o, x + (a), y + (b), z + (c))])
    s[0] = 0.5 * (D(u, 1, 0, 0) - D(u, -1, 0, 0)) * dxi;
    s[1] = 0.5 * (D(v, 0, 1, 0) - D(v, 0, -1, 0)) * dyi;
    s[2] = 0.5 * (D(w, 0, 0, 1) - D(w, 0, 0, -1)) * dzi;
    s[3] = 0.25 * ((D(u, 0, 1, 0) - D(u, 0, -1, 0)) * dyi + (D(v, 1, 0, 0) - D(v, -1, 0, 0)) * dxi);
    s[4] = 0.25 * ((D(v, 0, 0, 1) - D(v, 0, 0, -1)) * dzi + (D(w, 0, 1, 0) - D(w, 0, -1, 0)) * dyi);
    s[5] = 0.25 * ((D(w, 1, 0, 0) - D(w, -1, 0, 0)) * dxi + (D(u, 0, 0, 1) - D(u, 0, 0, -1)) * dzi);
#undef D
}

static const double IFACT[6] = {1, 1, 1, 2, 2, 2};
static const int IPN[6] = {0, 1, 2, 0, 1, 0}, JPN[6] = {0, 1, 2, 1, 2, 2};

static inline double smag_of(const double s[6])
{
    double sum = 0;
    for (int n = 0; n < 6; n++) sum += IFACT[n] * s[n] * s[n];
    return sqrt(2.0 * sum);
}

/* Separable periodic (1,4,1)/6 test filter in x, y and z (filter.f:9-42). */
static void filt3(const orc *o, const double *in, double *out, double *tmp)
{
    const int nx = o->c.nx, ny = o->c.ny, nz = o->c.nz;
    const double w6 = 1.0 / 6.0;
#pragma omp parallel for schedule(static)
    for (int z = 0; z < nz; z++)
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++)
                out[pidx(o, x, y, z)] = (in[pidx(o, x, y, z - 1)] + 4 * in[pidx(o, x, y, z)] + in[pidx(o, x, y, z + 1)]) * w6;
#pragma omp parallel for schedule(static)
    for (int z = 0; z < nz; z++)
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++)
                tmp[pidx(o, x, y, z)] = (out[pidx(o, x, y - 1, z)] + 4 * out[pidx(o, x, y, z)] + out[pidx(o, x, y + 1, z)]) * w6;
#pragma omp parallel for schedule(static)
    for (int z = 0; z < nz; z++)
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++)
                out[pidx(o, x, y, z)] = (tmp[pidx(o, x - 1, y, z)] + 4 * tmp[pidx(o, x, y, z)] + tmp[pidx(o, x + 1, y, z)]) * w6;
}

void orc_test_filter(const orc *o, const double *in, double *out)
{
    const size_t np = (size_t)o->c.nx * o->c.ny * o->c.nz;
    double *tmp = (double *)malloc(sizeof(double) * np);
    filt3(o, in, out, tmp);
    free(tmp);
}

/* Eddy viscosity nu_t.  sgs=1: sm.f:32-101, nu_t = c_smag Delta^2 |S|.
 * sgs=2: dm.f:45-191, Germano-Lilly with volume-averaged coefficient,
 * alpha^2 = 4 (2-Delta test filter), nu_t = C Delta^2 |S|.
 * Returns C Delta^2 (the dm.f coefficient) through *cdsq. */
void orc_eddy_visc(const orc *o, const double *phys, double *vt, double *cdsq)
{
    const int nx = o->c.nx, ny = o->c.ny, nz = o->c.nz;
    const size_t np = (size_t)nx * ny * nz;
    double *smag = (double *)malloc(sizeof(double) * np);
    double *S = NULL;
    if (o->c.sgs == 2) S = (double *)malloc(sizeof(double) * np * 6);
#pragma omp parallel for schedule(static)
    for (int z = 0; z < nz; z++)
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++) {
                double s[6];
                strain_at(o, phys, x, y, z, s);
                size_t i = pidx(o, x, y, z);
                smag[i] = smag_of(s);
                if (S) for (int n = 0; n < 6; n++) S[np * n + i] = s[n];
            }
    double c_delta_sq;
    if (o->c.sgs == 1) {
        c_delta_sq = o->c.c_smag * o->delta_sq;
    } else {
        /* filtered fields: u (3), S (6), |S|S (6), u_i u_j (6) */
        double *tmp = (double *)malloc(sizeof(double) * np);
        double *src = (double *)malloc(sizeof(double) * np);
        double *uh = (double *)malloc(sizeof(double) * np * 3);
        double *sh = (double *)malloc(sizeof(double) * np * 6);
        double *mh = (double *)malloc(sizeof(double) * np * 6);
        double *lh = (double *)malloc(sizeof(double) * np * 6);
        for (int n = 0; n < 3; n++) filt3(o, phys + np * n, uh + np * n, tmp);
        for (int n = 0; n < 6; n++) filt3(o, S + np * n, sh + np * n, tmp);
        for (int n = 0; n < 6; n++) {
#pragma omp parallel for schedule(static)
            for (size_t i = 0; i < np; i++) src[i] = smag[i] * S[np * n + i];
            filt3(o, src, mh + np * n, tmp);
#pragma omp parallel for schedule(static)
            for (size_t i = 0; i < np; i++) src[i] = phys[np * IPN[n] + i] * phys[np * JPN[n] + i];
            filt3(o, src, lh + np * n, tmp);
        }
        double num = 0, den = 0;
        const double alpha_sq = 4.0;
#pragma omp parallel for reduction(+ : num, den) schedule(static)
        for (size_t i = 0; i < np; i++) {
            double s[6];
            for (int n = 0; n < 6; n++) s[n] = sh[np * n + i];
            double sm = smag_of(s);
            for (int n = 0; n < 6; n++) {
                double M = -mh[np * n + i] + alpha_sq * sm * s[n];
                double L = lh[np * n + i] - uh[np * IPN[n] + i] * uh[np * JPN[n] + i];
                num += IFACT[n] * L * M;
                den += IFACT[n] * M * M;
            }
        }
        c_delta_sq = -0.5 * num / den;
        free(tmp); free(src); free(uh); free(sh); free(mh); free(lh);
    }
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < np; i++) vt[i] = c_delta_sq * smag[i];
    if (cdsq) *cdsq = c_delta_sq;
    free(smag); free(S);
}

/* SGS-augmented fluxes: F_ij = u_i u_j - 2 nu_t S_ij;
 * F_Tj = u_j T - (nu_t / Pr_t) dT/dx_j (centred differences). */
void orc_sgs_products(const orc *o, const double *phys, const double *vt, double *prod)
{
    const int nx = o->c.nx, ny = o->c.ny, nz = o->c.nz;
    const size_t np = (size_t)nx * ny * nz;
    orc_products(o, phys, prod);
    const double dxi = nx / o->c.xl, dyi = ny / o->c.yl, dzi = nz / o->c.zl;
    /* product slot of tensor component n (11,22,33,12,23,13) */
    static const int slot[6] = {0, 2, 5, 1, 4, 3};
#pragma omp parallel for schedule(static)
    for (int z = 0; z < nz; z++)
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++) {
                double s[6];
                strain_at(o, phys, x, y, z, s);
                size_t i = pidx(o, x, y, z);
                for (int n = 0; n < 6; n++) prod[np * slot[n] + i] -= 2.0 * vt[i] * s[n];
                if (o->lu == 4) {
                    const double *t = phys + 3 * np;
                    double k = vt[i] / o->c.pr_t;
                    prod[6 * np + i] -= k * 0.5 * (t[pidx(o, x + 1, y, z)] - t[pidx(o, x - 1, y, z)]) * dxi;
                    prod[7 * np + i] -= k * 0.5 * (t[pidx(o, x, y + 1, z)] - t[pidx(o, x, y - 1, z)]) * dyi;
                    prod[8 * np + i] -= k * 0.5 * (t[pidx(o, x, y, z + 1)] - t[pidx(o, x, y, z - 1)]) * dzi;
                }
            }
}

/* ------------------------------------------------------------------ */
/* Whole RHS and RK3 step (sam.f:122-326, rhs.f)                       */
/* ------------------------------------------------------------------ */

typedef struct {
    cx *zb, *yb, *fz;      /* transform work */
    double *phys, *prod, *vt;
    size_t nzb, nyb, nfz, nphys;
} orc_ws;

static void ws_init(const orc *o, orc_ws *w)
{
    const size_t nz = o->c.nz, ny = o->c.ny, nx = o->c.nx;
    w->nzb = (size_t)o->nym * o->nxm * nz * o->lu;
    w->nyb = (size_t)ny * o->nxh * nz * o->lu;
    w->nfz = (size_t)nz * o->nym * o->nxm * o->lr;
    w->nphys = nx * ny * nz;
    w->zb = (cx *)malloc(sizeof(cx) * w->nzb);
    w->yb = (cx *)malloc(sizeof(cx) * w->nyb);
    w->fz = (cx *)malloc(sizeof(cx) * w->nfz);
    w->phys = (double *)malloc(sizeof(double) * w->nphys * o->lu);
    w->prod = (double *)malloc(sizeof(double) * w->nphys * o->lr);
    w->vt = o->c.sgs ? (double *)malloc(sizeof(double) * w->nphys) : NULL;
}

static void ws_free(orc_ws *w)
{
    free(w->zb); free(w->yb); free(w->fz); free(w->phys); free(w->prod); free(w->vt);
}

/* physical velocity (Nx,Ny,Nz,Lu) of the packed spectral state u */
void orc_to_physical(const orc *o, const cx *u, double *phys)
{
    orc_ws w; ws_init(o, &w);
    orc_z_trans_b(o, u, o->nzm, w.zb, o->lu);
    orc_x2z(o, w.zb, w.yb, o->lu);
    orc_xy_trans_b(o, w.yb, phys, o->lu);
    ws_free(&w);
}

/* RHS of rhs.f for the physical velocity held in w->phys; usav is the
 * spectral state before the zeta update.  Writes r (Lu+1 fields, S layout)
 * and, when do_stats, the vel_max statistics (dt updated if i_cfl). */
static void rhs_phys(const orc *o, orc_ws *w, const cx *usav, cx *r, int do_stats, double *stats, double dt_in, double *cdsq)
{
    if (do_stats) orc_vel_max(o, w->phys, dt_in, stats);
    if (o->c.sgs) {
        orc_eddy_visc(o, w->phys, w->vt, cdsq);
        orc_sgs_products(o, w->phys, w->vt, w->prod);
    } else {
        orc_products(o, w->phys, w->prod);
    }
    orc_xy_trans_f(o, w->prod, w->fz, o->lr);
    /* z2x_decomp on one rank is the identity reorder (z2x_decomp.f:44-68) */
    orc_z_trans_f(o, w->fz, o->c.nz, o->lr);
    orc_project(o, w->fz, o->c.nz, usav, r);
}

/* RHS of the packed state u (the nt_end pass of sam.f:288 / a first
 * substep).  r: (Nz_min,Ny_min,Nx_min,Lu+1).  stats as orc_vel_max. */
void orc_rhs(const orc *o, const cx *u, cx *r, double dt_in, double *stats)
{
    orc_ws w; ws_init(o, &w);
    orc_z_trans_b(o, u, o->nzm, w.zb, o->lu);
    orc_x2z(o, w.zb, w.yb, o->lu);
    orc_xy_trans_b(o, w.yb, w.phys, o->lu);
    double cd = 0;
    rhs_phys(o, &w, u, r, 1, stats, dt_in, &cd);
    if (stats) stats[17] = cd;
    ws_free(&w);
}

/* ------------------------------------------------------------------ */
/* Chollet-Lesieur energy and eddy viscosity, spectral forcing           */
/* ------------------------------------------------------------------ */

/* cl_energy.f:8-43: energy of the modes in the shell band around k_cl,
 * walked as the reference does: per (kx, ky) column the kz index range
 * k_start..k_stop (1-based) and its mirror row kk = Nz_min + 2 - k.  For
 * k = 1 the mirror is row Nz_min + 1, past the stored modes: it adds the
 * (zero) padding row of u(Nze, ...). */
double orc_cl_energy(const orc *o, const cx *u)
{
    const int nxm = o->nxm, nym = o->nym, nzm = o->nzm;
    const size_t ss = (size_t)nzm * nym * nxm;
    const double start_sq = ((double)o->k_cl - 0.5) * ((double)o->k_cl - 0.5);
    const double stop_sq = ((double)o->k_cl + 0.5) * ((double)o->k_cl + 0.5);
    double sum_e = 0.0;
    for (int ix = 0; ix < nxm; ix++) {
        const double factor = (ix == 0 || ix == o->c.nx / 2) ? 1.0 : 2.0;
        const int kx2 = ix * ix;
        for (int iy = 0; iy < nym; iy++) {
            const int k_sq2 = kx2 + o->ky[iy] * o->ky[iy];
            const int k_start = (k_sq2 > start_sq) ? 1 : (int)sqrt(start_sq - k_sq2) + 2;
            const int k_stop = (k_sq2 > stop_sq) ? 0 : (int)sqrt(stop_sq - k_sq2) + 1;
            for (int k = k_start; k <= k_stop; k++) {
                const int kk = nzm + 2 - k;
                double usq1 = 0, usq2 = 0;
                for (int n = 0; n < 3; n++) {
                    const cx a = u[(k - 1) + (size_t)nzm * (iy + (size_t)nym * ix) + n * ss];
                    usq1 += creal(a * conj(a));
                    if (kk <= nzm) {
                        const cx b = u[(kk - 1) + (size_t)nzm * (iy + (size_t)nym * ix) + n * ss];
                        usq2 += creal(b * conj(b));
                    }
                }
                sum_e += factor * (usq1 + usq2);
            }
        }
    }
    return sum_e;
}

/* sam.f:131 / spectra.f:21 */
double orc_cl_fac(const orc *o, double e_cl) { return sqrt(2.0 * e_cl / (double)o->k_cl); }

/* spectra.f:39-43: nu_t(k) of the Chollet-Lesieur model, k = |k_int| */
static double cl_vist(const orc *o, double fac_cl, int ksq)
{
    if (ksq == 0) return 0.0;
    return fac_cl * (o->a_cl + o->b_cl * exp(-(o->c_cl * o->k_cl) / sqrt((double)ksq)));
}

/* force.f:12-63: energy of u, v, w in the shells is <= k_force + 1
 * (is = 1 + int(|k| + 0.5), factor 2 for kx > 0) scaled back to ek0 (the
 * spectrum of the last spectra call, ek of spectra.f); shell 1 (the mean)
 * is left alone.  ekg (>= k_force + 1 doubles) receives the energies before
 * the scaling. */
void orc_force(const orc *o, cx *u, const double *ek0, double *ekg)
{
    const int nxm = o->nxm, nym = o->nym, nzm = o->nzm, kf = o->c.k_force;
    const size_t ss = (size_t)nzm * nym * nxm;
    double *work = (double *)calloc((size_t)kf + 2, sizeof(double));
    for (int pass = 0; pass < 2; pass++) {
        for (int ix = 0; ix < nxm; ix++) {
            const double factor = (ix == 0 || ix == o->c.nx / 2) ? 1.0 : 2.0;
            for (int iy = 0; iy < nym; iy++) {
                const int k_sq2 = ix * ix + o->ky[iy] * o->ky[iy];
                for (int k = 0; k < nzm; k++) {
                    const int k_sq = k_sq2 + o->kz[k] * o->kz[k];
                    const int is = 1 + (int)(sqrt((double)k_sq) + 0.5);
                    if (is > kf + 1) continue;
                    const size_t i = k + (size_t)nzm * (iy + (size_t)nym * ix);
                    if (pass == 0) {
                        double usq = 0;
                        for (int n = 0; n < 3; n++) usq += creal(u[i + n * ss] * conj(u[i + n * ss]));
                        work[is - 1] += factor * usq;
                    } else {
                        for (int n = 0; n < 3; n++) u[i + n * ss] *= work[is - 1];
                    }
                }
            }
        }
        if (pass == 0) {
            for (int s = 0; s <= kf; s++) ekg[s] = work[s];
            work[0] = 1.0;
            for (int s = 1; s <= kf; s++) work[s] = sqrt(ek0[s] / ekg[s]);
        }
    }
    free(work);
}

/* nsteps full RK steps of sam.f:122-326 (no I/O, no forcing, no
 * Chollet-Lesieur energy).
 * u: packed state (Nz_min,Ny_min,Nx_min,Lu), updated in place.
 * state[0]=time, state[1]=dt (in/out).  stats: vel_max of the last
 * step's first substep (19 doubles, [17] = C Delta^2 when sgs=2). */
void orc_step(const orc *o, cx *u, int nsteps, double *state, double *stats);

/* orc_step with the reference's per-step extras: cl_energy at the first
 * substep of every step (sam.f:128-132) when i_cl != 0, the
 * Chollet-Lesieur nu_t(k) added to the integrating factor when i_cl = 2
 * (velocity nu + nu_t, scalar nu/Pr + nu_t/Pr_t; an extension: SAM uses
 * the model in its diagnostics only), and force.f after every step
 * (sam.f:320-322) when k_force >= 0, toward ek0.  stats[18] = fac_cl of
 * the last step. */
void orc_step_ext(const orc *o, cx *u, int nsteps, double *state, double *stats, const double *ek0, cx *p_last);

void orc_step(const orc *o, cx *u, int nsteps, double *state, double *stats)
{
    orc_step_ext(o, u, nsteps, state, stats, NULL, NULL);
}

/* p_last (may be NULL): the pressure r(.,.,.,Lu+1) of the last substep,
 * which write_planes.f outputs at the next step (sam.f:136-150). */
void orc_step_ext(const orc *o, cx *u, int nsteps, double *state, double *stats, const double *ek0, cx *p_last)
{
    orc_ws w; ws_init(o, &w);
    const size_t ss = (size_t)o->nzm * o->nym * o->nxm;
    const int lu = o->lu;
    cx *r = (cx *)calloc(ss * (lu + 1), sizeof(cx));
    cx *usav = (cx *)malloc(sizeof(cx) * ss * lu);
    double time = state[0], dt = state[1];
    double st[19] = {0};
    double fac_cl = 0.0;
    double *ekg = (double *)calloc((size_t)(o->c.k_force > 0 ? o->c.k_force : 0) + 2, sizeof(double));
    for (int nt = 0; nt < nsteps; nt++) {
        for (int m = 0; m < o->c.nrk_max; m++) {
            if (m == 0 && o->c.i_cl) fac_cl = orc_cl_fac(o, orc_cl_energy(o, u));
            orc_z_trans_b(o, u, o->nzm, w.zb, lu);
            memcpy(usav, u, sizeof(cx) * ss * lu);
            const double dt_fac2 = o->zet[m] * dt;
            if (m > 0) {
#pragma omp parallel for schedule(static)
                for (size_t i = 0; i < ss * lu; i++) u[i] = u[i] + dt_fac2 * r[i];
            }
            orc_x2z(o, w.zb, w.yb, lu);
            orc_xy_trans_b(o, w.yb, w.phys, lu);
            double cd = 0;
            rhs_phys(o, &w, usav, r, m == 0, st, dt, &cd);
            if (m == 0) { dt = st[16]; st[17] = cd; }
            const double dt_fac1 = o->gam[m] * dt;
            const double dt_fac = dt_fac1 + dt_fac2;
            for (int n = 0; n < lu; n++) {
                double vis_f = o->c.vis * dt_fac;
                if (n == 3) vis_f *= o->pr_inv;
                const double vt_f = (o->c.i_cl == 2 ? dt_fac : 0.0) * (n == 3 ? 1.0 / o->c.pr_t : 1.0);
                cx *un = u + ss * n, *rn = r + ss * n;
#pragma omp parallel for schedule(static)
                for (int ix = 0; ix < o->nxm; ix++) {
                    const double rkx = o->wave_x * (double)ix;
                    const double rkx2 = rkx * rkx;
                    for (int iy = 0; iy < o->nym; iy++) {
                        const double rky = o->wave_y * (double)o->ky[iy];
                        const double rk_sq2 = rkx2 + rky * rky;
                        const int ksq2 = ix * ix + o->ky[iy] * o->ky[iy];
                        for (int k = 0; k < o->nzm; k++) {
                            const double rkz = o->wave_z * (double)o->kz[k];
                            const double rk_sq = rk_sq2 + rkz * rkz;
                            double ve = vis_f;
                            if (vt_f != 0.0) ve += vt_f * cl_vist(o, fac_cl, ksq2 + o->kz[k] * o->kz[k]);
                            const double vf = exp(-ve * rk_sq);
                            const size_t i = k + (size_t)o->nzm * (iy + (size_t)o->nym * ix);
                            un[i] = (un[i] + dt_fac1 * rn[i]) * vf;
                        }
                    }
                }
            }
        }
        time += dt;
        if (o->c.k_force >= 0 && ek0) orc_force(o, u, ek0, ekg);
    }
    state[0] = time; state[1] = dt;
    st[18] = fac_cl;
    if (stats) memcpy(stats, st, sizeof(st));
    if (p_last) memcpy(p_last, r + ss * lu, sizeof(cx) * ss);
    free(r); free(usav); free(ekg);
    ws_free(&w);
}

/* ------------------------------------------------------------------ */
/* Diagnostics: spectra.f:13-111 and div_check.f:11-36                  */
/* fac_cl: sqrt(2 E_cl / k_cl) of the step (used when i_cl != 0).        */
/* out: ek[lm] tk[lm] dk[lm] vt[lm] sample[lm], then energy_w, div_rms,  */
/* div_max (5*lm + 3 doubles).  u and r in S layout (ld = Nz_min).      */
/* ------------------------------------------------------------------ */
void orc_spectra(const orc *o, const cx *u, const cx *r, double fac_cl, double *out)
{
    const int lm = o->lm, nxm = o->nxm, nym = o->nym, nzm = o->nzm;
    const size_t ss = (size_t)nzm * nym * nxm;
    double *acc = (double *)calloc((size_t)lm * 5 + 2, sizeof(double));
    double divs = 0, divmax = 0;
    for (int ix = 0; ix < nxm; ix++) {
        const double factor = (ix == 0 || ix == o->nxh) ? 1.0 : 2.0;
        const int kx2 = ix * ix;
        const double rkx = o->wave_x * ix, rkx2 = rkx * rkx;
        for (int iy = 0; iy < nym; iy++) {
            const int ksq2 = kx2 + o->ky[iy] * o->ky[iy];
            const double rky = o->wave_y * o->ky[iy];
            const double rk_sq2 = rkx2 + rky * rky;
            for (int k = 0; k < nzm; k++) {
                const int ksq = ksq2 + o->kz[k] * o->kz[k];
                const double kmag = sqrt((double)ksq);
                const int is = (int)(kmag + 0.5);
                const double rkz = o->wave_z * o->kz[k];
                const double rk_sq = rk_sq2 + rkz * rkz;
                const size_t i = k + (size_t)nzm * (iy + (size_t)nym * ix);
                const cx u1 = u[i], u2 = u[i + ss], u3 = u[i + 2 * ss];
                const double usq = creal(u1 * conj(u1) + u2 * conj(u2) + u3 * conj(u3));
                const double udu = creal(r[i] * conj(u1) + r[i + ss] * conj(u2) + r[i + 2 * ss] * conj(u3));
                const double vist = o->c.i_cl ? cl_vist(o, fac_cl, ksq) : 0.0;   /* spectra.f:39-43 */
                acc[is] += factor * usq;
                acc[lm + is] += factor * udu;
                acc[2 * lm + is] += factor * rk_sq * (o->c.vis + vist) * usq;
                acc[3 * lm + is] += factor * vist;
                acc[4 * lm + is] += factor;
                const cx div = rkx * u1 + rky * u2 + rkz * u3;
                const double dsq = creal(div * conj(div));
                divs += dsq;
                if (dsq > divmax) divmax = dsq;
            }
        }
    }
    double sek = 0;
    for (int i = 0; i < lm; i++) sek += acc[i];
    memcpy(out, acc, sizeof(double) * 5 * lm);
    out[5 * lm] = 0.5 * sek;
    out[5 * lm + 1] = sqrt(divs);
    out[5 * lm + 2] = sqrt(divmax);
    free(acc);
}

/* div_check.f:11-36 on one rank: out = div_max, div_rms, n_div over the
 * stored modes (no kx factor 2); sums in the reference's loop order. */
void orc_div_check(const orc *o, const cx *u, double *out)
{
    const size_t ss = (size_t)o->nzm * o->nym * o->nxm;
    double s2 = 0, mx = 0, cnt = 0;
    for (int ix = 0; ix < o->nxm; ix++) {
        const double rkx = o->wave_x * ix;
        for (int iy = 0; iy < o->nym; iy++) {
            const double rky = o->wave_y * o->ky[iy];
            for (int k = 0; k < o->nzm; k++) {
                const double rkz = o->wave_z * o->kz[k];
                const size_t i = k + (size_t)o->nzm * (iy + (size_t)o->nym * ix);
                const cx dv = rkx * u[i] + rky * u[i + ss] + rkz * u[i + 2 * ss];
                const double dsq = creal(dv * conj(dv));
                s2 += dsq;
                if (dsq > mx) mx = dsq;
                if (dsq > 1.0e-20) cnt += 1;
            }
        }
    }
    out[0] = sqrt(mx);
    out[1] = sqrt(s2);
    out[2] = cnt;
}

/* The RK update of substep nrk (1-based) on u (S layout): the zeta term
 * u += zeta dt r_prev for nrk > 1 (sam.f:210-221) when r_prev is given, then
 * u = (u + gamma dt r) exp(-vis_f |k|^2) with vis_f = vis (gamma+zeta) dt
 * (/Pr for T') (sam.f:293-314) when r is given. */
void orc_rk_update(const orc *o, int nrk, double dt, cx *u, const cx *r_prev, const cx *r)
{
    const size_t ss = (size_t)o->nzm * o->nym * o->nxm;
    const int m = nrk - 1;
    const double dt_fac2 = o->zet[m] * dt, dt_fac1 = o->gam[m] * dt;
    if (r_prev && nrk > 1)
        for (size_t i = 0; i < ss * o->lu; i++) u[i] = u[i] + dt_fac2 * r_prev[i];
    if (!r) return;
    for (int n = 0; n < o->lu; n++) {
        double vis_f = o->c.vis * (dt_fac1 + dt_fac2);
        if (n == 3) vis_f *= o->pr_inv;
        for (int ix = 0; ix < o->nxm; ix++) {
            const double rkx = o->wave_x * (double)ix;
            for (int iy = 0; iy < o->nym; iy++) {
                const double rky = o->wave_y * (double)o->ky[iy];
                const double rk_sq2 = rkx * rkx + rky * rky;
                for (int k = 0; k < o->nzm; k++) {
                    const double rkz = o->wave_z * (double)o->kz[k];
                    const size_t i = k + (size_t)o->nzm * (iy + (size_t)o->nym * ix) + ss * n;
                    u[i] = (u[i] + dt_fac1 * r[i]) * exp(-vis_f * (rk_sq2 + rkz * rkz));
                }
            }
        }
    }
}

/* ------------------------------------------------------------------ */
/* Visualization output: write_planes.f:18-144 with get_mean.f:8-23    */
/* (periodic case, no jump removal) and x_deriv.f / y_deriv.f           */
/* ------------------------------------------------------------------ */

/* x_deriv.f:7-25: d/dx of an (Nx, Ny) plane, one real FFT per row */
static void plane_x_deriv(const orc *o, const double *u, double *du)
{
    const int nx = o->c.nx, ny = o->c.ny, kxm = o->kxm;
    for (int j = 0; j < ny; j++) {
        double *r = du + (size_t)j * nx;
        for (int i = 0; i < nx; i++) r[i] = u[i + (size_t)j * nx] / nx;
        orc_rfft(nx, r, -1);
        r[0] = 0.0;
        for (int k = 1; k <= kxm; k++) {
            const double rk = k * o->wave_x, re = r[2 * k - 1];
            r[2 * k - 1] = -rk * r[2 * k];
            r[2 * k] = rk * re;
        }
        for (int i = 2 * kxm + 1; i < nx; i++) r[i] = 0.0;
        orc_rfft(nx, r, 1);
    }
}

/* y_deriv.f:35-56: d/dy of an (Nx, Ny) plane, one real FFT per column */
static void plane_y_deriv(const orc *o, const double *u, double *du)
{
    const int nx = o->c.nx, ny = o->c.ny, kym = o->kym;
    double *w = (double *)malloc(sizeof(double) * ny);
    for (int i = 0; i < nx; i++) {
        for (int j = 0; j < ny; j++) w[j] = u[i + (size_t)j * nx] / ny;
        orc_rfft(ny, w, -1);
        w[0] = 0.0;
        for (int k = 1; k <= kym; k++) {
            const double rk = k * o->wave_y, re = w[2 * k - 1];
            w[2 * k - 1] = -rk * w[2 * k];
            w[2 * k] = rk * re;
        }
        for (int j = 2 * kym + 1; j < ny; j++) w[j] = 0.0;
        orc_rfft(ny, w, 1);
        for (int j = 0; j < ny; j++) du[i + (size_t)j * nx] = w[j];
    }
    free(w);
}

/* Planes and means of the state u (S layout) with pressure p (S layout,
 * the r(.,.,.,Lu+1) of the last RHS), as sam.f:145-167 builds them and
 * write_planes.f writes them.  Fields per plane (nf = 2 Lu):
 *   xy: u v w [T] p du/dz dv/dz [dT/dz]   (Nx, Ny, nf) per plane
 *   xz: u v w [T] p du/dy dw/dy [dT/dy]   (Nx, Nz, nf)
 *   yz: u v w [T] p dv/dx dw/dx [dT/dx]   (Ny, Nz, nf)
 * mean: (Nz, Lu), the kx = ky = 0 column through rfftb (get_mean.f case 0).
 * Plane positions are 1-based (k_xy_plane, j_xz_plane, i_yz_plane). */
void orc_planes(const orc *o, const cx *u, const cx *p, int n_xy, const int *k_xy, int n_xz, const int *j_xz,
                int n_yz, const int *i_yz, double *mean, double *xy, double *xz, double *yz)
{
    const int nx = o->c.nx, ny = o->c.ny, nz = o->c.nz, lu = o->lu, nf = 2 * lu;
    const int nxm = o->nxm, nym = o->nym, nzm = o->nzm, kzm = o->kzm;
    const size_t ss = (size_t)nzm * nym * nxm, np = (size_t)nx * ny * nz;
    cx *spec = (cx *)malloc(sizeof(cx) * ss * nf);
    memcpy(spec, u, sizeof(cx) * ss * lu);
    memcpy(spec + ss * lu, p, sizeof(cx) * ss);
    const int dz_src[3] = {0, 1, 3};
    for (int q = 0; q < lu - 1; q++) {      /* sam.f:147-156 */
        const cx *a = u + ss * dz_src[q];
        cx *b = spec + ss * (lu + 1 + q);
        for (int ix = 0; ix < nxm; ix++)
            for (int iy = 0; iy < nym; iy++)
                for (int k = 0; k < nzm; k++) {
                    const size_t i = k + (size_t)nzm * (iy + (size_t)nym * ix);
                    b[i] = I * o->wave_z * o->kz[k] * a[i];
                }
    }
    cx *zb = (cx *)malloc(sizeof(cx) * (size_t)nym * nxm * nz * nf);
    cx *yb = (cx *)malloc(sizeof(cx) * (size_t)ny * o->nxh * nz * nf);
    double *phys = (double *)malloc(sizeof(double) * np * nf);
    orc_z_trans_b(o, spec, nzm, zb, nf);
    orc_x2z(o, zb, yb, nf);
    orc_xy_trans_b(o, yb, phys, nf);

    /* get_mean.f:10-20 */
    double *um = (double *)malloc(sizeof(double) * nz);
    for (int n = 0; n < lu; n++) {
        const cx *a = u + ss * n;
        um[0] = creal(a[0]);
        for (int k = 1; k <= kzm; k++) { um[2 * k - 1] = creal(a[k]); um[2 * k] = cimag(a[k]); }
        for (int k = 2 * kzm + 1; k < nz; k++) um[k] = 0.0;
        orc_rfft(nz, um, 1);
        memcpy(mean + (size_t)nz * n, um, sizeof(double) * nz);
    }
    free(um);

    const size_t pl = (size_t)nx * ny;
    double *du = (double *)malloc(sizeof(double) * pl);
    for (int m = 0; m < n_xy; m++) {          /* write_planes.f:38-48 */
        const int k = k_xy[m] - 1;
        for (int n = 0; n < nf; n++)
            memcpy(xy + pl * (n + (size_t)nf * m), phys + pl * (k + (size_t)nz * n), sizeof(double) * pl);
    }
    for (int m = 0; m < n_xz; m++) {          /* write_planes.f:58-100 */
        const int j = j_xz[m] - 1;
        double *out = xz + (size_t)nx * nz * nf * m;
        for (int n = 0; n <= lu; n++)
            for (int k = 0; k < nz; k++)
                for (int i = 0; i < nx; i++)
                    out[i + (size_t)nx * (k + (size_t)nz * n)] = phys[i + nx * (j + (size_t)ny * (k + (size_t)nz * n))];
        for (int k = 0; k < nz; k++) {
            int L = 1;
            for (int n = 0; n < lu; n++) {
                if (n == 1) continue;
                L++;
                plane_y_deriv(o, phys + pl * (k + (size_t)nz * n), du);
                for (int i = 0; i < nx; i++) out[i + (size_t)nx * (k + (size_t)nz * (lu + L - 1))] = du[i + (size_t)nx * j];
            }
        }
    }
    for (int m = 0; m < n_yz; m++) {          /* write_planes.f:110-144 */
        const int i = i_yz[m] - 1;
        double *out = yz + (size_t)ny * nz * nf * m;
        for (int n = 0; n <= lu; n++)
            for (int k = 0; k < nz; k++)
                for (int j = 0; j < ny; j++)
                    out[j + (size_t)ny * (k + (size_t)nz * n)] = phys[i + nx * (j + (size_t)ny * (k + (size_t)nz * n))];
        for (int k = 0; k < nz; k++) {
            int L = 1;
            for (int n = 0; n < lu; n++) {
                if (n == 0) continue;
                L++;
                plane_x_deriv(o, phys + pl * (k + (size_t)nz * n), du);
                for (int j = 0; j < ny; j++) out[j + (size_t)ny * (k + (size_t)nz * (lu + L - 1))] = du[i + (size_t)nx * j];
            }
        }
    }
    free(du); free(spec); free(zb); free(yb); free(phys);
}
