// internal.hpp -- shared definitions of libsamhip (not part of the C ABI).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "fft_tile.hpp"
#include "../../include/samhip.h"

namespace samhip {

// Grid description shared by every kernel (init.f:47-85, fft_init.f:52-72).
struct Dims {
    int nx, ny, nz;
    int kxm, kym, kzm;
    int nxm, nym, nzm;
    int lu, lr, lm;
    long long ss;            // Nx_min*Ny_min*Nz_min, one spectral field
    long long zfield;        // Nx_min*Ny_min*Nz, one field in the Z layout
    long long yfield;        // Nz*Nx_min*Ny, one field in the Y layout
    long long pfield;        // Nx*Ny*Nz, one physical field
    double wave_x, wave_y, wave_z;
};

// Device-resident diagnostics and clock (vel_max.f outputs, dt, time).
struct DevDiag {
    double time, dt;
    double u_avg[4], u_var[4], u_max[4];
    double energy;
    double cfl_x, cfl_y, cfl_z;
    double c_delta_sq;
    long long nt;
    unsigned long long divmax_bits;    // running max |div u| as IEEE bits
    double fac_cl;                     // sqrt(2 E_cl / k_cl) of the step (sam.f:131)
    double spare[5];
};

// Scalars the RK epilogue needs (init.f:97-148, rhs.f, sam.f:293-314).
struct StepConsts {
    double gamma[4], zeta[4];
    double vis, pr_inv, buoy_fac_x, buoy_fac_z, lapse;
    double cfl0, adv_cfl_limit, dt0;
    int i_cfl, strat;
    // Chollet-Lesieur nu_t(k) = fac_cl (a + b exp(-c k_cl / |k|)) (init.f:88-93,
    // spectra.f:39-43): cl = 0 off, 1 in the spectra only (SAM's i_les = 1),
    // 2 also in the integrating factor (scalar: nu_t / Pr_t)
    int cl;
    double cl_a, cl_b, cl_c, k_cl, pr_t_inv;
};

// XCD-aware tile order.  The dispatcher deals the blocks of a launch out to
// the 8 XCDs round robin (block i runs on XCD i % 8), and each XCD has its own
// L2.  Renumbering so that XCD x takes the contiguous tile range
// [x q + min(x, r), ...) (q = n / 8, r = n % 8, a bijection of [0, n)) keeps
// the neighbouring tiles of one XCD on neighbouring HBM rows.  At 512^3 FP64
// this made k_yb 10% faster (992 -> 896 us), k_zb, k_xrow and k_yf 2-3%.
// The FP32 y-forward transform and k_zrhs ran no faster, and they keep the
// launch order.  The SGS stencil kernels (sgs.hip) use it so that tiles that
// share halo lines run on one XCD, whose L2 then serves the halo.
__device__ __forceinline__ int xcd_tile(int i, int n)
{
    const int q = n / 8, r = n % 8, x = i % 8;
    return x * q + (x < r ? x : r) + i / 8;
}

/* Device buffer layouts (complex elements, first index fastest):
 *   S : (Nz_min, Ny_min, Nx_min, F)   spectral state / RHS, FP64
 *   Z : (Nz, Ny_min, Nx_min, F)       z in physical space, x,y spectral
 *   Y : (Ny, Nx_min, Nz, F)           y,z physical, x spectral (kx <= kx_max)
 *   P : (Nx, Ny, Nz, F) real          physical space
 */

// launchers (kernels.hip); T = float or double; all asynchronous on `st`.
template <typename T> int launch_zb(const Dims &d, const double *S, int ldz, int nvar, T *Z, const void *twz, hipStream_t st);
template <typename T> int launch_yb(const Dims &d, const T *Z, int nvar, T *Y, const void *twy, hipStream_t st);
template <typename T> int launch_yf(const Dims &d, const T *Y, int nvar, T *Z, const void *twy, hipStream_t st);
template <typename T> int launch_xb(const Dims &d, const T *Y, int nvar, T *P, const void *twx, hipStream_t st);
template <typename T> int launch_xf(const Dims &d, const T *P, int nvar, T *Y, const void *twx, hipStream_t st);
template <typename T> int launch_zf(const Dims &d, const T *Z, int nvar, double *S, const void *twz, hipStream_t st);
// fused x-row kernel: inverse x, products, forward x (+ vel_max partials)
template <typename T> int launch_xrow(const Dims &d, T *Y, int want_stats, double *partials, DevDiag *diag,
                                      const void *twx, hipStream_t st, int *nblocks_out);
// fused z-forward + projection + RK/IF update (mode 0; mode 2 also keeps the
// pressure in R's field Lu) or RHS only with the pressure (mode 1).  Zout
// (may equal Z; null = none): in the update modes also write the z inverse
// of the new state there (k_zb's output for the next substep); *fused tells
// whether it was written (the instantiation's columns fit in one wavefront).
template <typename T> int launch_zrhs(const Dims &d, const StepConsts &k, const T *Z, const double *U, double *R,
                                      double *Unew, const DevDiag *diag, int m, int mode, const void *twz,
                                      hipStream_t st, T *Zout = nullptr, bool *fused = nullptr);
int launch_stats_phys(const Dims &d, const void *P, int is_double, int nvar, double *partials, DevDiag *diag,
                      hipStream_t st, int *nblocks_out);
// rows (may be null): 256 x 12 doubles for the second reduction level when
// nblocks >= 8192; null = the tail of a partials buffer of partials_capacity(d)
int launch_finalize(const Dims &d, const StepConsts &k, const double *partials, int nblocks, DevDiag *diag,
                    int update_dt, hipStream_t st, double *rows = nullptr);
constexpr int FINALIZE_ROWS = 256 * 12;
int launch_advance_clock(DevDiag *diag, hipStream_t st);
int launch_project(const Dims &d, const StepConsts &k, const double *F, const double *usav, double *r, hipStream_t st);
int launch_products(const Dims &d, const void *P, int is_double, void *Q, hipStream_t st);
int launch_spectra(const Dims &d, const StepConsts &k, const double *U, const double *R, double *partials,
                   double *out, DevDiag *diag, hipStream_t st);
// cl_energy.f: E_cl of the band around k_cl -> diag->fac_cl
int launch_cl_energy(const Dims &d, int k_cl, const double *U, double *partials, DevDiag *diag, hipStream_t st);
// force.f: shell energies (shells 0..kf) into ekg, then u, v, w of shells 1..kf scaled by sqrt(ek0/ekg)
int launch_force_shells(const Dims &d, int kf, const double *U, double *ekg, hipStream_t st);
int launch_force_scale(const Dims &d, int kf, double *U, const double *ek0, const double *ekg, hipStream_t st);
int launch_strided_copy(const double *src, double *dst, long long n0, long long n1, long long n2, long long n3,
                        long long s0, long long s1, long long s2, long long s3, long long d0, long long d1,
                        long long d2, long long d3, int cplx, hipStream_t st);
template <typename T> int launch_convert(const T *src, double *dst, long long n, hipStream_t st);
template <typename T> int launch_convert_to(const double *src, T *dst, long long n, hipStream_t st);

// stage layouts, RK update, div_check (stage.hip)
template <typename T> int launch_z2ref(const Dims &d, const T *Z, int nvar, double *out, hipStream_t st);
template <typename T> int launch_ref2z(const Dims &d, const double *in, int nvar, T *Z, hipStream_t st);
int launch_x2z(const Dims &d, const double *in, int nvar, double *out, hipStream_t st);
// zeta term with rp (may be null), gamma term + integrating factor with r (may be null); m = nrk - 1
int launch_rk_update(const Dims &d, const StepConsts &k, int m, double dt, double *U, const double *rp,
                     const double *r, hipStream_t st);
long long div_check_rows(const Dims &d);
// a DevDiag set to zero with dt = dt; and vel_max's 17 outputs of a DevDiag
int launch_diag_seed(DevDiag *diag, double dt, hipStream_t st);
int launch_diag_stats(const DevDiag *diag, double *stats, hipStream_t st);
// out (device, 3 doubles): div_max, div_rms, n_div of the state U
int launch_div_check(const Dims &d, const double *U, double *partials, double *out, hipStream_t st);

// blocks of `kernel` (launched with `threads` threads) the current device
// holds at once; cached per kernel and device
long long resident_blocks(const void *kernel, int threads);

// SGS (sgs.hip)
template <typename T> int launch_strain_smag(const Dims &d, const T *P, T *smag, T *S6, double dxi, double dyi,
                                             double dzi, hipStream_t st);
template <typename T> int launch_filter(const Dims &d, const T *in, T *out, T *tmp, int nvar, hipStream_t st);
// dm.f's dynamic procedure in one pass: block sums of L:M and M:M (2 doubles
// per block, at most dm_fused_rows(d) blocks) for launch_dm_coef
template <typename T> int launch_dm_fused(const Dims &d, const T *P, double dxi, double dyi, double dzi,
                                          double *partials, hipStream_t st, int *nblocks_out);
long long dm_fused_rows(const Dims &d);
int launch_dm_coef(const double *partials, int nblocks, double c_fixed, int dynamic, int report, DevDiag *diag,
                   hipStream_t st);
template <typename T> int launch_sgs_flux(const Dims &d, const T *P, const DevDiag *diag, T *Q,
                                          double dxi, double dyi, double dzi, double pr_t_inv, hipStream_t st);
template <typename T> int launch_nut(const Dims &d, const T *smag, const DevDiag *diag, double *vt, hipStream_t st);

// learned closure (mlp.hip).  FP32 parameter block: W1[64][16] b1[64]
// W2[64][64] b2[64] W3[64][64] b3[64] W4[16][64] b4[16]; inputs padded 12->16,
// outputs 9->16.  norms[32]: input scales (16) then output scales (16).
constexpr int MLP_IN = 12, MLP_HID = 64, MLP_OUT = 9;
constexpr int MLP_OFF_W1 = 0;
constexpr int MLP_OFF_B1 = MLP_OFF_W1 + 64 * 16;
constexpr int MLP_OFF_W2 = MLP_OFF_B1 + 64;
constexpr int MLP_OFF_B2 = MLP_OFF_W2 + 64 * 64;
constexpr int MLP_OFF_W3 = MLP_OFF_B2 + 64;
constexpr int MLP_OFF_B3 = MLP_OFF_W3 + 64 * 64;
constexpr int MLP_OFF_W4 = MLP_OFF_B3 + 64;
constexpr int MLP_OFF_B4 = MLP_OFF_W4 + 16 * 64;
constexpr int MLP_NPARAM = MLP_OFF_B4 + 16;
constexpr int MLP_TILE_BYTES = 76 * 64 * 8 + 208 * 4;   // packed BF16 weight tiles + FP32 biases
int launch_mlp_pack(const float *P, void *img, hipStream_t st);
// one fused training pass (mlp_train.hip): forward, loss, backward and the
// weight gradients of `batch` samples; writes mlp_step_blocks(batch)
// gradient rows of MLP_NPARAM floats to part and as many loss partials
int mlp_step_blocks(int batch);
int launch_mlp_step(const void *img, const float *norms, const float *feats, const float *labels, long long npool,
                    long long base, long long pa, long long pb, int batch, float *part, double *loss_part, int act,
                    hipStream_t st);
int launch_mlp_adam(const float *part, int nch, float *P, float *m1, float *m2, float *gout, float lr, int step,
                    int apply, hipStream_t st);
int launch_sum_loss(const double *part, int n, double *out, hipStream_t st);
int launch_mlp_infer(const float *P, const void *img, const float *norms, const float *feats, float *out, long long n,
                     int act, hipStream_t st);
// cscale: device pointer to C Delta^2 for unit-coefficient networks, else null
template <typename T>
int launch_les_mlp_flux(const Dims &d, const void *img, const float *norms, const T *U, T *Q, float dxi, float dyi,
                        float dzi, const double *cscale, int act, hipStream_t st);
// the dynamic model's stress and flux at unit coefficient of the features
// gradients, times *cscale (device) when cscale is not null
int launch_dyn_labels(const float *feats, long long n, float pr_t_inv, const double *cscale, float *labels,
                      hipStream_t st);
template <typename T> int launch_box_filter(const Dims &d, const T *U, int r, float *coarse, float *labels, hipStream_t st);
int launch_coarse_features(const float *C, int nx, int ny, int nz, int lu, float dxi, float dyi, float dzi,
                           float *feats, hipStream_t st);
int launch_pool_norms(const float *feats, const float *labels, long long n, double *part, float *norms, hipStream_t st);

// planes output (write_planes.f): spectral fields of one batch (S layout,
// ld Nz_min) -- batch 0: u v w [T] p d/dz(u v [T]); batch 1: d/dy(u w [T])
// d/dx(v w [T]) -- and the gather of one plane from a physical buffer
int launch_planes_spec(const Dims &d, int batch, const double *U, const double *R, double *out, hipStream_t st);
template <typename T>
int launch_plane_gather(const Dims &d, const T *P, int kind, int pos, int src_f0, int nfld, double *out, int dst_f0,
                        hipStream_t st);

// plain 16-byte-per-thread copy (counter calibration)
int launch_copy16(const void *src, void *dst, long long n16, hipStream_t st);
bool stream_mix_supported(int nr, int nw);
int launch_stream(int nr, int nw, const void *src, void *dst, long long n16, hipStream_t st);

// max number of per-block partial rows the step's reductions may produce
long long partials_capacity(const Dims &d);

} // namespace samhip
