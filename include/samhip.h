/*
 * samhip.h -- C ABI of libsamhip, the MI355X (gfx950) implementation of
 * SAM's RK3 pseudo-spectral time step.
 *
 * All entry points are extern "C", take plain pointers and sizes, and
 * return an int status (SAM_OK = 0).  Complex arrays are interleaved
 * (re, im) doubles, i.e. Fortran complex(8).  Array shapes are written
 * Fortran-style, first index fastest, exactly as the reference declares
 * them.  Host arrays are copied to/from the device by the call; the step
 * itself keeps its state resident in HBM behind the handle.
 *
 * One handle = one grid = one HIP stream on the current device.  Calls on
 * one handle must come from one host thread at a time.  Every call that
 * returns host data synchronises the handle's stream.
 *
 * Notation: kx_max,ky_max,kz_max per init.f:49-62; Nx_min = kx_max+1,
 * Ny_min = 2*ky_max+1, Nz_min = 2*kz_max+1 (init.f:71-73); Nze = Nz for the
 * periodic z case (init.f:78-80); Lu = 3 (4 if i_strat), Lr = 6 (9).
 */
#ifndef SAMHIP_H
#define SAMHIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------- */
#define SAM_OK               0
#define SAM_ERR_REQUIRED     6   /* a required input is missing (check_inputs.f:222-226) */
#define SAM_ERR_VALUE        9   /* an input has an invalid value (check_inputs.f:227-230) */
#define SAM_ERR_SIZE        10   /* a buffer/file size does not match the grid (read_field.f:26-32) */
#define SAM_ERR_UNSUPPORTED 20   /* valid for the reference, not built here (e.g. iubc_z != 0) */
#define SAM_ERR_STATE       21   /* call out of order (e.g. spectra before any RHS) */
#define SAM_ERR_HIP         30   /* a HIP runtime call failed; see sam_last_error() */
#define SAM_ERR_NOMEM       31   /* device allocation failed */
#define SAM_ERR_NODEVICE    32   /* no usable gfx950 device */

/* precision of the transforms / physical-space products */
#define SAM_PREC_FP64        0   /* plain FP64 drop-in */
#define SAM_PREC_SPLIT       1   /* FP32 transforms+products, FP64 state/RK/reductions */

/* subgrid-scale model coupled into the step */
#define SAM_SGS_NONE         0
#define SAM_SGS_SMAG         1   /* sm.f: nu_t = c_smag * Delta^2 * |S| */
#define SAM_SGS_DYN          2   /* dm.f: Germano-Lilly, volume-averaged coefficient */
#define SAM_SGS_MLP          3   /* learned per-point closure (no reference counterpart) */

/* Run parameters: the subset of input.dat (set_labels.f:29-102) and of the
 * COMMON blocks of sam.h:13-55 that the step reads. */
typedef struct sam_params {
    int    nx, ny, nz;          /* Nx Ny Nz: powers of two, 16..1024 */
    int    n_dealias;           /* 0, 1 (2/3 rule, default) or 2 (init.f:49-62) */
    int    i_strat;             /* 0: u,v,w; 1: + temperature T' (Boussinesq) */
    int    nrk_max;             /* 1, 2 or 3 (init.f:117-145) */
    int    i_cfl;               /* 1: dt = cfl0*adv_cfl_limit/max(u_i,max k_i,max) (vel_max.f:43-46) */
    int    iubc_z;              /* must be 0 (periodic z); others -> SAM_ERR_UNSUPPORTED */
    int    precision;           /* SAM_PREC_* */
    int    sgs;                 /* SAM_SGS_* */
    double xl, yl, zl;          /* box lengths */
    double vis;                 /* kinematic viscosity */
    double pr;                  /* Prandtl number (T uses vis/pr) */
    double grav, to;            /* buoy_fac_z = grav/To (init.f:101-105) */
    double lapse;               /* background gradient in the T equation (rhs.f:133);
                                   lapse0 for a fresh run, 0 reproduces the restart quirk */
    double cfl0, dt0;           /* CFL target / fixed dt */
    double c_smag;              /* Smagorinsky c_smag (sm.f:32) */
    double pr_t;                /* turbulent Prandtl number of the SGS scalar flux */
} sam_params;

typedef struct sam_dims {
    int kx_max, ky_max, kz_max;
    int nx_min, ny_min, nz_min, nze;
    int lu, lr, lm;             /* lm = number of spectral shells (sam.f:72-73) */
} sam_dims;

/* Diagnostics of the most recent first substep (vel_max.f:28-49) plus the
 * run clock.  u_* are per field (u, v, w, T'). */
typedef struct sam_diag {
    long long nt;
    double time, dt;
    double u_avg[4], u_var[4], u_max[4];
    double energy;              /* 0.5*(u_var(1)+u_var(2)+u_var(3)) */
    double cfl_x, cfl_y, cfl_z;
    double c_delta_sq;          /* C*Delta^2 of the dynamic model (dm.f:176), else 0 */
    int    sgs_active;          /* SGS model actually used by the last step */
    int    pad_;
} sam_diag;

typedef struct sam_handle sam_handle;

/* ---- lifecycle ------------------------------------------------------- */
/* Replaces init.f + fft_init.f + the allocations of sam.f:80-98. */
int  sam_create(const sam_params *p, sam_handle **h);
int  sam_destroy(sam_handle *h);
int  sam_get_dims(const sam_handle *h, sam_dims *d);
const char *sam_last_error(void);

/* ---- state (read_field.f / write_field.f record layout) --------------- */
/* u: (ldz, Ny_min, Nx_min, Lu) complex, ldz >= Nz_min (Nz_min = the vel.*
 * file layout, Nze = sam.f's in-memory u).  Rows beyond Nz_min ignored on
 * set, zero-filled on get. */
int  sam_set_state(sam_handle *h, const double *u, int ldz);
int  sam_get_state(sam_handle *h, double *u, int ldz);
/* The velocity of substep nrk (1..nrk_max) of the coming step after its
 * zeta update (sam.f:210-221), which is what sam.f:277-284 writes to vel.*
 * at nt < nt_end (the write inside the step's last substep, nrk = nrk_max,
 * is the one the file keeps).  nrk = 1 is the state itself.  Runs nrk-1
 * substeps ahead and restores the handle (state, RHS, diagnostics). */
int  sam_get_state_substep(sam_handle *h, int nrk, double *u, int ldz);
/* clock carried in header.* (time, dt, nt; write_header.f:19-25) */
int  sam_set_clock(sam_handle *h, double time, double dt, long long nt);

/* ---- the time step (sam.f:122-326) ----------------------------------- */
/* nsteps full RK steps.  diag (may be NULL) receives the statistics of the
 * last step's first substep and the clock after the last step.  With
 * want_spectra != 0 the shell spectra of the last step's first substep
 * (state and RHS at the start of that step, as spectra.f is called in
 * sam.f:248-250) are kept for sam_get_spectra. */
int  sam_step(sam_handle *h, int nsteps, int want_spectra, sam_diag *diag);
/* RHS-only pass of the current state (the nt_end pass, sam.f:288).
 * r (may be NULL): (ldz, Ny_min, Nx_min, Lu+1) complex, last field = p. */
int  sam_rhs(sam_handle *h, double *r, int ldz, int want_spectra, sam_diag *diag);
/* spectra.f:13-111 of the state/RHS pair captured by the last call with
 * want_spectra.  out: ek(lm) tk(lm) dk(lm) vt(lm) sample(lm) energy_w
 * div_rms div_max  (5*lm + 3 doubles). */
int  sam_get_spectra(sam_handle *h, double *out);

/* ---- per-step extras of sam.f ------------------------------------------ */
/* force.f (i_force = 1, sam.f:320-322): after every step of sam_step, the
 * u, v, w Fourier coefficients of shells 2..k_force+1 (1-based shells
 * is = 1 + int(|k| + 0.5)) are scaled by sqrt(ek_0(is)/ek(is)), so the
 * low-shell spectrum stays at ek_0, the ek of the last spectra taken by
 * sam_step / sam_rhs (sam.f:250).  Before any spectrum the energies at the
 * first forced step become ek_0.  k_force < 0 switches forcing off;
 * 0 <= k_force <= max(Nx,Ny,Nz)/2 as check_inputs.f:215 requires. */
int  sam_set_forcing(sam_handle *h, int k_force);
/* Chollet-Lesieur model (i_les = 1): the energy of the band around
 * k_cl = int(Nx/3) is taken at the first substep of every step
 * (cl_energy.f, sam.f:128-132) and nu_t(k) = sqrt(2 E_cl/k_cl) (0.145 +
 * 5.01 exp(-3.03 k_cl/|k|)) (spectra.f:39-43, init.f:88-93) enters
 *   mode 1: the spectra only (dk, vt), as SAM does;
 *   mode 2: also the integrating factor: exp(-(nu + nu_t(k)) k^2 dt) for
 *           u, v, w and exp(-(nu/Pr + nu_t(k)/Pr_t) k^2 dt) for T' (an
 *           extension: SAM never couples the model into the step).
 * mode 0 switches it off.  SAM_ERR_UNSUPPORTED when k_cl exceeds the
 * stored modes of a direction. */
int  sam_set_cl(sam_handle *h, int mode);

/* ---- visualization output (write_planes.f, get_mean.f, x/y_deriv.f) --- */
/* keep_pressure != 0: the last substep of every step keeps its pressure,
 * the r(.,.,.,Lu+1) that write_planes.f outputs at the next planes step
 * (sam.f:136-150).  Costs one field write per step; off by default. */
int  sam_set_planes(sam_handle *h, int keep_pressure);
/* Planes and horizontal means of the current state, as sam.f:145-167 and
 * write_planes.f form them at the first substep of an n_skip_p step
 * (positions 1-based, at most 9 of each kind; nf = 2 Lu fields each):
 *   xy (Nx, Ny, nf, n_xy): u v w [T] p du/dz dv/dz [dT/dz] at z = k_xy(m)
 *   xz (Nx, Nz, nf, n_xz): u v w [T] p du/dy dw/dy [dT/dy] at y = j_xz(m)
 *   yz (Ny, Nz, nf, n_yz): u v w [T] p dv/dx dw/dx [dT/dx] at x = i_yz(m)
 *   mean (Nz, Lu): horizontal mean profiles (get_mean.f, periodic z)
 * p is the kept pressure (zero before any).  Computed in the handle's
 * precision.  mean may be NULL; a kind with n = 0 may pass NULLs. */
int  sam_planes(sam_handle *h, int n_xy, const int *k_xy, int n_xz, const int *j_xz, int n_yz, const int *i_yz,
                double *mean, double *xy, double *xz, double *yz);

/* ---- stage entry points (parity with the reference's routines) -------- */
/* z_trans_b.f:15-52 (case 0, one rank): u (ldz,Ny_min,Nx_min,nvar) ->
 * out (Ny_min,Nx_min,Nz,nvar), the layout z_trans_b hands to x2z_decomp. */
int  sam_z_trans_b(sam_handle *h, const double *u, int ldz, double *out, int nvar);
/* x2z_decomp.f:24-152 on one rank (the all-to-all is a copy): in
 * (Ny_min,Nx_min,Nz,nvar) -> out (Ny,Nx/2,Nz,nvar), y padded with zeros
 * between ky_max and the stored negative ky; x columns >= Nx_min zero. */
int  sam_x2z(sam_handle *h, const double *in, double *out, int nvar);
/* xy_trans_b.f:11-63: in (Ny,Nx/2,Nz,nvar) complex, as x2z_decomp writes it
 * -> physical out (Nx,Ny,Nz,nvar) real.  Reads the stored modes only (the
 * first Nx_min columns and the Ny_min stored y rows): xy_trans_b.f never
 * reads the columns past kx_max, and x2z_decomp leaves the y band zero. */
int  sam_xy_trans_b(sam_handle *h, const double *in, double *out, int nvar);
/* xy_trans_f.f:13-57: physical (Nx,Ny,Nz,nvar) -> (Nz,Ny_min,Nx_min,nvar) */
int  sam_xy_trans_f(sam_handle *h, const double *in, double *out, int nvar);
/* z2x_decomp.f:23-68 on one rank (identity reorder, Nze = Nz):
 * (Nz,Ny_min,Nx_min,nvar) -> (Nze,Ny_min,Nx_min,nvar) */
int  sam_z2x(sam_handle *h, const double *in, double *out, int nvar);
/* z_trans_f.f (case 0): (Nz,Ny_min,Nx_min,nvar) -> (Nz_min,Ny_min,Nx_min,nvar) */
int  sam_z_trans_f(sam_handle *h, const double *in, double *out, int nvar);
/* rhs.f:32-61 products: physical (Nx,Ny,Nz,Lu) -> (Nx,Ny,Nz,Lr) */
int  sam_products(sam_handle *h, const double *phys, double *prod);
/* vel_max.f: stats[0..3] u_avg, [4..7] u_var, [8..11] u_max, [12] energy,
 * [13..15] cfl_{x,y,z}_by_dt, [16] dt (dt_in unless i_cfl=1) */
int  sam_vel_max(sam_handle *h, const double *phys, double dt_in, double *stats);
/* No counterpart in the reference: sam_vel_max for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops).  phys (Nx,Ny,Nz,Lu) FP64 and stats (17 doubles, sam_vel_max's
 * layout) are DEVICE pointers; the call allocates nothing and is enqueued on
 * the handle's stream (sam_sync waits for it).  Its workspace is a
 * file-scope device array of 6.3 MB per device, held until the process ends
 * and shared by every call on that device: calls that overlap on one device
 * race.  Grids above 2^27 points (512^3) return SAM_ERR_SIZE. */
int  sam_vel_max_noalloc(sam_handle *h, const double *phys, double dt_in, double *stats);
/* rhs.f:74-142: F (Nz_min,Ny_min,Nx_min,Lr), usav (Nz_min,Ny_min,Nx_min,Lu)
 * -> r (Nz_min,Ny_min,Nx_min,Lu+1), last field = pressure */
int  sam_project(sam_handle *h, const double *F, const double *usav, double *r);
/* The RK update of substep nrk (1..nrk_max) on u (ldz,Ny_min,Nx_min,Lu):
 *   r_prev != NULL and nrk > 1: u += zeta(nrk) dt r_prev          (sam.f:210-221)
 *   r != NULL: u = (u + gamma(nrk) dt r) exp(-nu_f (gamma+zeta)(nrk) dt |k|^2),
 *              nu_f = vis (vis/Pr for T')                         (sam.f:293-314)
 * r_prev and r are (ldz,Ny_min,Nx_min,Lu); either may be NULL. */
int  sam_rk_update(sam_handle *h, int nrk, double dt, double *u, const double *r_prev, const double *r, int ldz);
/* div_check.f:11-36 of the current state: out[0] div_max = sqrt(max |k.u|^2),
 * out[1] div_rms = sqrt(sum |k.u|^2), out[2] n_div = number of stored modes
 * with |k.u|^2 > 1e-20 (kx >= 0 modes, no conjugate factor). */
int  sam_div_check(sam_handle *h, double *out);

/* ---- subgrid-scale pieces (sm.f, dm.f, filter.f) ---------------------- */
/* eddy viscosity of a physical velocity (Nx,Ny,Nz,Lu) with the handle's
 * sgs model (SMAG or DYN); vt (Nx,Ny,Nz); c_delta_sq may be NULL */
int  sam_eddy_visc(sam_handle *h, const double *phys, double *vt, double *c_delta_sq);
/* (1,4,1)/6 periodic test filter of nvar physical fields (filter.f:9-42) */
int  sam_test_filter(sam_handle *h, const double *in, double *out, int nvar);

/* ---- learned SGS closure (no counterpart in the reference) ------------ */
/* Per-point MLP 12 -> 64 -> 64 -> 64 -> 9 (ReLU or GELU hidden activation),
 * BF16 MFMA / FP32 accumulate.
 * Inputs: du_i/dx_j (9, row-major) and dT'/dx_j (3) by centred differences;
 * outputs: deviatoric tau_11,22,33,12,23,13 and q_1..3; the step then uses
 * F_ij = u_i u_j + tau_ij and F_Tj = u_j T' + q_j. */
typedef struct sam_mlp sam_mlp;
#define SAM_MLP_NPARAM 10448     /* FP32 parameters; + 32 normalisation scales */
int  sam_mlp_create(sam_mlp **m, unsigned seed, long long pool_capacity);
int  sam_mlp_destroy(sam_mlp *m);
/* hidden activation: part of the network.  Set it before training, and
 * before sam_mlp_set_params with parameters trained under it.
 *   SAM_MLP_ACT_RELU  max(x, 0) (default);
 *   SAM_MLP_ACT_GELU  tanh form 0.5 x (1 + tanh(sqrt(2/pi) (x + 0.044715 x^3))). */
#define SAM_MLP_ACT_RELU 0
#define SAM_MLP_ACT_GELU 1
int  sam_mlp_set_activation(sam_mlp *m, int act);
int  sam_mlp_get_activation(const sam_mlp *m, int *act);
/* params: SAM_MLP_NPARAM weights/biases followed by 32 scales */
int  sam_mlp_get_params(sam_mlp *m, float *params);
int  sam_mlp_set_params(sam_mlp *m, const float *params);
/* training samples from the current state of a DNS handle: box filter by
 * `ratio` in each direction and gradient features of the filtered field on
 * the coarse grid; labels as chosen by sam_mlp_set_labels */
int  sam_mlp_add_samples(sam_mlp *m, sam_handle *dns, int ratio);
/* kind of label sam_mlp_add_samples writes (the pool must be empty to change
 * it):
 *   SAM_LABELS_FILTERED  the box filter's subfilter stress and flux,
 *                        tau_ij = (u_i u_j)~ - u~_i u~_j (deviatoric),
 *                        q_j = (u_j T')~ - u~_j T'~ (default);
 *   SAM_LABELS_DYNAMIC   the dynamic model's stress and flux at unit
 *                        coefficient (dm.f/sm.f form, tau_ij = -2|S|S_ij,
 *                        q_j = -(|S|/pr_t) dT'/dx_j) on the filtered field.
 *                        A network trained on these is scaled in the LES by
 *                        C Delta^2 from the dynamic procedure, recomputed
 *                        every sam_set_mlp_refresh steps. */
#define SAM_LABELS_FILTERED 0
#define SAM_LABELS_DYNAMIC  1
/*   SAM_LABELS_DYNAMIC_COEF  the dynamic model's stress and flux with its
 *                        coefficient: C Delta^2 of dm.f's dynamic procedure
 *                        run on each filtered snapshot (Delta = the filtered
 *                        grid's spacing) times the unit-coefficient form.  A
 *                        network trained on these replaces the procedure: the
 *                        LES uses its output directly, with no refresh. */
#define SAM_LABELS_DYNAMIC_COEF 2
int  sam_mlp_set_labels(sam_mlp *m, int mode, double pr_t);
/* C Delta^2 of the last snapshot added under SAM_LABELS_DYNAMIC_COEF */
int  sam_mlp_last_coefficient(sam_mlp *m, double *c_delta_sq);
/* host samples: feats [n][12], labels [n][9] */
int  sam_mlp_add_samples_host(sam_mlp *m, const float *feats, const float *labels, long long n);
long long sam_mlp_pool_size(const sam_mlp *m);
/* copy pool samples [offset, offset+n) to host: feats [n][12], labels [n][9] */
int  sam_mlp_get_samples(sam_mlp *m, long long offset, long long n, float *feats, float *labels);
int  sam_mlp_clear_pool(sam_mlp *m);
/* input/output scales = 1/rms and rms of the pool's columns */
int  sam_mlp_fit_norms(sam_mlp *m);
/* nsteps Adam steps (beta 0.9/0.999) on pseudo-random minibatches of the
 * pool; batch must be a multiple of 4096; losses[nsteps] (may be NULL) */
int  sam_mlp_train(sam_mlp *m, int nsteps, int batch, float lr, unsigned seed, float *losses);
/* loss (mean square, normalised outputs) and its gradient w.r.t. the
 * SAM_MLP_NPARAM parameters for host samples (n multiple of 4096) */
int  sam_mlp_grad(sam_mlp *m, const float *feats, const float *labels, int n, float *grad, double *loss);
/* inference: feats [n][12] -> out [n][9] (physical units, deviatoric tau) */
int  sam_mlp_infer(sam_mlp *m, const float *feats, float *out, long long n);
/* LES handle h uses network m when its sgs model is SAM_SGS_MLP */
int  sam_attach_mlp(sam_handle *h, sam_mlp *m);
/* switch the handle's SGS model between steps (e.g. fall back MLP -> DYN) */
int  sam_set_sgs(sam_handle *h, int sgs);
/* steps between updates of C Delta^2 by the dynamic procedure for a network
 * trained on SAM_LABELS_DYNAMIC labels (default 20; the first step after
 * sam_attach_mlp always updates) */
int  sam_set_mlp_refresh(sam_handle *h, int nsteps);
/* SAM_SGS_DYN with a lagged coefficient (no reference counterpart): the
 * dynamic procedure runs at the first substep of every nsteps-th step and
 * C Delta^2 is held in between; nsteps = 1 (default) runs it every substep
 * as dm.f does.  The control for the learned closure's CDelta^2 refresh. */
int  sam_set_dyn_refresh(sam_handle *h, int nsteps);
/* a-posteriori check of the learned closure on the current state, with the
 * dynamic model as the arbiter: the shell spectra of the SGS energy transfer
 * T_m(k) = tk_m(k) - tk_none(k) (spectra.f's tk of the RHS with model m
 * minus without) are compared; err = ||T_mlp - T_dyn|| / ||T_dyn|| over
 * shells.  When err > tol the handle falls back to SAM_SGS_DYN
 * (*switched = 1).  Costs three RHS evaluations; the state is unchanged. */
int  sam_sgs_check(sam_handle *h, double tol, double *err, int *switched);

/* ---- measurement (no counterpart in the reference, which has no timers) */
/* kernel classes timed with HIP events on the handle's stream */
#define SAM_K_ZB     0   /* z inverse transform            */
#define SAM_K_YB     1   /* y pad + y inverse transform    */
#define SAM_K_XROW   2   /* fused x C2R/products/R2C       */
#define SAM_K_FINAL  3   /* vel_max finalize / dt          */
#define SAM_K_YF     4   /* y forward transform + truncate */
#define SAM_K_ZRHS   5   /* z forward + projection + RK/IF */
#define SAM_K_SPEC   6   /* spectra                        */
#define SAM_K_XB     7   /* x C2R to physical (LES path)   */
#define SAM_K_XF     8   /* x R2C from physical (LES path) */
#define SAM_K_SGS    9   /* SGS fluxes: eddy-viscosity or network (LES path) */
#define SAM_K_MISC  10   /* clock and small kernels        */
#define SAM_K_DMPROC 11  /* dynamic procedure and C Delta^2 (LES path) */
#define SAM_K_STATS 12   /* vel_max statistics of physical fields (LES path) */
#define SAM_K_COUNT 16
/* on != 0: record a HIP event pair around every kernel the step launches */
int  sam_prof_enable(sam_handle *h, int on);
/* accumulated milliseconds and launch counts per kernel class since the
 * last call (arrays of SAM_K_COUNT); resets the accumulators */
int  sam_prof_read(sam_handle *h, double *ms, long long *count);
/* device-wide synchronisation (bench timing brackets) */
int  sam_sync(sam_handle *h);
/* plain device-to-device copy of `bytes` bytes, `reps` times, on the
 * handle's stream: the calibration kernel for HBM counters */
int  sam_device_copy(sam_handle *h, long long bytes, int reps);
/* streaming mix: each launch reads nread x `bytes` bytes from one array and
 * writes nwrite x `bytes` to another (one read and one write stream, 16 KiB
 * chunks in the ratio nread:nwrite; `bytes` a multiple of 16 KiB), `reps`
 * launches, synchronous: the sustained rate of a kernel with the same
 * read:write proportions.  Supported (nread, nwrite): (1,0) (0,1) (1,1)
 * (1,2) (2,1) (2,3) (3,2) (4,5) (5,4) (4,9).  No reference counterpart
 * (measurement only). */
int  sam_device_stream(sam_handle *h, int nread, int nwrite, long long bytes, int reps);

#ifdef __cplusplus
}
#endif
#endif /* SAMHIP_H */
