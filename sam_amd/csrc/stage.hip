// stage.hip -- the reference's stage layouts, the RK update and div_check
// as kernels of their own, for the stage entry points of the C ABI.
//
// The step keeps the z-leading Z layout (Nz, Ny_min, Nx_min) between the z
// and y transforms and pads y inside k_yb's loads.  The reference hands
// z_trans_b's output to x2z_decomp as (Ny_min, Nx_min, Nz) and x2z_decomp's
// output to xy_trans_b as (Ny, Nx/2, Nz) (z_trans_b.f:31-52,
// x2z_decomp.f:59-152).  These kernels convert between the two so that each
// stage entry point reads and writes exactly the reference's arrays:
//   k_z2ref     Z (Nz, M, F)            -> (M, Nz, F), M = Ny_min Nx_min
//   k_ref2z     (Ny, Nx/2, Nz, F)       -> Z (Nz, Ny_min, Nx_min, F)
//   k_x2z       (Ny_min, Nx_min, Nz, F) -> (Ny, Nx/2, Nz, F), y zero band
//   k_rk_update the zeta and gamma/integrating-factor updates (sam.f:210-221,293-314)
//   k_div_check div_check.f's sums over the stored modes
#include "internal.hpp"

namespace samhip {

static inline int launched_st()
{
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? SAM_OK : SAM_ERR_HIP;
}

// 32 x 32 tile transpose through LDS (one padding column keeps the 16-byte
// column reads of the tile on distinct banks); 256 threads, 4 rows each
constexpr int TT = 32;

// Z (Nz, M, F) complex T -> out (M, Nz, F) complex double
template <typename T>
__global__ __launch_bounds__(256) void k_z2ref(const cpx<T> *__restrict__ Z, cpx<double> *__restrict__ out, int nz,
                                               long long M)
{
    __shared__ cpx<double> tile[TT][TT + 1];
    const int f = blockIdx.z;
    const long long m0 = (long long)blockIdx.x * TT;
    const int z0 = blockIdx.y * TT;
    const int tx = threadIdx.x % TT, ty = threadIdx.x / TT;
    const cpx<T> *zf = Z + (long long)f * M * nz;
    for (int r = ty; r < TT; r += 256 / TT) {           // r: m within the tile, tx: z
        const long long m = m0 + r;
        const int z = z0 + tx;
        if (m < M && z < nz) {
            const cpx<T> v = zf[m * nz + z];
            tile[r][tx] = {(double)v.x, (double)v.y};
        }
    }
    __syncthreads();
    cpx<double> *of = out + (long long)f * M * nz;
    for (int r = ty; r < TT; r += 256 / TT) {           // r: z within the tile, tx: m
        const int z = z0 + r;
        const long long m = m0 + tx;
        if (m < M && z < nz) of[(long long)z * M + m] = tile[tx][r];
    }
}

// (Ny, Nxh, Nz, F) complex double -> Z (Nz, Ny_min, Nx_min, F) complex T: the
// stored y rows (ky <= ky_max and the ky_max negatives at the top) of the
// first Nx_min x columns
template <typename T>
__global__ __launch_bounds__(256) void k_ref2z(const cpx<double> *__restrict__ in, cpx<T> *__restrict__ Z, Dims d,
                                               int nxh)
{
    __shared__ cpx<double> tile[TT][TT + 1];
    const int f = blockIdx.z;
    const long long M = (long long)d.nym * d.nxm;
    const long long m0 = (long long)blockIdx.x * TT;
    const int z0 = blockIdx.y * TT;
    const int tx = threadIdx.x % TT, ty = threadIdx.x / TT;
    const cpx<double> *inf = in + (long long)f * d.nz * nxh * d.ny;
    for (int r = ty; r < TT; r += 256 / TT) {           // r: z within the tile, tx: m
        const int z = z0 + r;
        const long long m = m0 + tx;
        if (m < M && z < d.nz) {
            const int i = (int)(m / d.nym), jm = (int)(m % d.nym);
            const int j = jm <= d.kym ? jm : jm + d.ny - d.nym;
            tile[r][tx] = inf[((long long)z * nxh + i) * d.ny + j];
        }
    }
    __syncthreads();
    cpx<T> *zf = Z + (long long)f * M * d.nz;
    for (int r = ty; r < TT; r += 256 / TT) {           // r: m within the tile, tx: z
        const long long m = m0 + r;
        const int z = z0 + tx;
        if (m < M && z < d.nz) {
            const cpx<double> v = tile[tx][r];
            zf[m * d.nz + z] = {(T)v.x, (T)v.y};
        }
    }
}

// (Ny_min, Nx_min, Nz, F) -> (Ny, Nxh, Nz, F): x2z_decomp.f:98-150 on one
// rank, with zeros in the truncated y band and in the x columns >= Nx_min
__global__ __launch_bounds__(256) void k_x2z(const cpx<double> *__restrict__ in, cpx<double> *__restrict__ out, Dims d,
                                             int nxh, long long n_out)
{
    const long long o = (long long)blockIdx.x * 256 + threadIdx.x;
    if (o >= n_out) return;
    const int j = (int)(o % d.ny);
    long long rest = o / d.ny;
    const int i = (int)(rest % nxh);
    rest /= nxh;                                         // rest = f Nz + z
    cpx<double> v = {0.0, 0.0};
    const int jtop = d.ny - d.nym;                       // rows j > ky_max + jtop hold the negative ky
    if (i < d.nxm && (j <= d.kym || j > d.kym + jtop)) {
        const int jm = j <= d.kym ? j : j - jtop;
        v = in[(rest * d.nxm + i) * d.nym + jm];
    }
    out[o] = v;
}

// sam.f:210-221 (zeta term, when rp) and sam.f:293-314 (gamma term and
// integrating factor exp(-nu (gamma+zeta) dt k^2), nu/Pr for T', when r) on
// the S layout, with the arithmetic of the fused step kernel
template <int LU>
__global__ __launch_bounds__(256) void k_rk_update(cpx<double> *__restrict__ U, const cpx<double> *__restrict__ rp,
                                                   const cpx<double> *__restrict__ r, Dims d, double dtf2, double dtf1,
                                                   double vf, double pr_inv)
{
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= d.ss) return;
    const int kz = (int)(i % d.nzm);
    const long long c = i / d.nzm;
    const int iy = (int)(c % d.nym), ix = (int)(c / d.nym);
    const int kyw = iy <= d.kym ? iy : iy - d.nym, kzw = kz <= d.kzm ? kz : kz - d.nzm;
    const double rkx = d.wave_x * (double)ix, rky = d.wave_y * (double)kyw, rkz = d.wave_z * (double)kzw;
    const double rk_sq = (rkx * rkx + rky * rky) + rkz * rkz;
#pragma unroll
    for (int n = 0; n < LU; n++) {
        cpx<double> u = U[n * d.ss + i];
        if (rp) u = {u.x + dtf2 * rp[n * d.ss + i].x, u.y + dtf2 * rp[n * d.ss + i].y};
        if (r) {
            const double vis_fac = exp(-((n == 3) ? vf * pr_inv : vf) * rk_sq);
            u = {u.x + dtf1 * r[n * d.ss + i].x, u.y + dtf1 * r[n * d.ss + i].y};
            u = {u.x * vis_fac, u.y * vis_fac};
        }
        U[n * d.ss + i] = u;
    }
}

// div_check.f:11-36 per block: sum |k.u|, sum |k.u|^2, max |k.u|^2 and the
// count of |k.u|^2 > 1e-20 over the stored modes; rows of 4 doubles
constexpr int DIV_CHUNK = 4096;
__global__ __launch_bounds__(256) void k_div_check(const cpx<double> *__restrict__ U, Dims d,
                                                   double *__restrict__ partials)
{
    __shared__ double red[4][256];
    const long long b0 = (long long)blockIdx.x * DIV_CHUNK;
    const long long b1 = b0 + DIV_CHUNK < d.ss ? b0 + DIV_CHUNK : d.ss;
    double s1 = 0, s2 = 0, mx = 0, cnt = 0;
    for (long long i = b0 + threadIdx.x; i < b1; i += 256) {
        const int kz = (int)(i % d.nzm);
        const long long c = i / d.nzm;
        const int iy = (int)(c % d.nym), ix = (int)(c / d.nym);
        const double rkx = d.wave_x * (double)ix;
        const double rky = d.wave_y * (double)(iy <= d.kym ? iy : iy - d.nym);
        const double rkz = d.wave_z * (double)(kz <= d.kzm ? kz : kz - d.nzm);
        const cpx<double> a = U[i], b = U[d.ss + i], e = U[2 * d.ss + i];
        const double dr = rkx * a.x + rky * b.x + rkz * e.x, di = rkx * a.y + rky * b.y + rkz * e.y;
        const double dsq = dr * dr + di * di;
        s1 += sqrt(dsq);
        s2 += dsq;
        mx = fmax(mx, dsq);
        cnt += dsq > 1.0e-20 ? 1.0 : 0.0;
    }
    red[0][threadIdx.x] = s1;
    red[1][threadIdx.x] = s2;
    red[2][threadIdx.x] = mx;
    red[3][threadIdx.x] = cnt;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            red[0][threadIdx.x] += red[0][threadIdx.x + w];
            red[1][threadIdx.x] += red[1][threadIdx.x + w];
            red[2][threadIdx.x] = fmax(red[2][threadIdx.x], red[2][threadIdx.x + w]);
            red[3][threadIdx.x] += red[3][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x < 4) partials[(long long)blockIdx.x * 4 + threadIdx.x] = red[threadIdx.x][0];
}

// the rows in block order by one block: out = div_max, div_rms, n_div
// (div_check.f:34-36; div_avg, the sum of |k.u|, is not an output there)
__global__ __launch_bounds__(256) void k_div_final(const double *__restrict__ partials, int nb, double *__restrict__ out)
{
    __shared__ double red[4][256];
    double s1 = 0, s2 = 0, mx = 0, cnt = 0;
    for (int b = threadIdx.x; b < nb; b += 256) {
        s1 += partials[4 * b];
        s2 += partials[4 * b + 1];
        mx = fmax(mx, partials[4 * b + 2]);
        cnt += partials[4 * b + 3];
    }
    red[0][threadIdx.x] = s1;
    red[1][threadIdx.x] = s2;
    red[2][threadIdx.x] = mx;
    red[3][threadIdx.x] = cnt;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            red[0][threadIdx.x] += red[0][threadIdx.x + w];
            red[1][threadIdx.x] += red[1][threadIdx.x + w];
            red[2][threadIdx.x] = fmax(red[2][threadIdx.x], red[2][threadIdx.x + w]);
            red[3][threadIdx.x] += red[3][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = sqrt(red[2][0]);
        out[1] = sqrt(red[1][0]);
        out[2] = red[3][0];
    }
}

template <typename T> int launch_z2ref(const Dims &d, const T *Z, int nvar, double *out, hipStream_t st)
{
    const long long M = (long long)d.nym * d.nxm;
    dim3 grid((unsigned)((M + TT - 1) / TT), (unsigned)((d.nz + TT - 1) / TT), (unsigned)nvar);
    hipLaunchKernelGGL((k_z2ref<T>), grid, dim3(256), 0, st, (const cpx<T> *)Z, (cpx<double> *)out, d.nz, M);
    return launched_st();
}

template <typename T> int launch_ref2z(const Dims &d, const double *in, int nvar, T *Z, hipStream_t st)
{
    const long long M = (long long)d.nym * d.nxm;
    dim3 grid((unsigned)((M + TT - 1) / TT), (unsigned)((d.nz + TT - 1) / TT), (unsigned)nvar);
    hipLaunchKernelGGL((k_ref2z<T>), grid, dim3(256), 0, st, (const cpx<double> *)in, (cpx<T> *)Z, d, d.nx / 2);
    return launched_st();
}

int launch_x2z(const Dims &d, const double *in, int nvar, double *out, hipStream_t st)
{
    const long long n = (long long)d.ny * (d.nx / 2) * d.nz * nvar;
    hipLaunchKernelGGL(k_x2z, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, (const cpx<double> *)in,
                       (cpx<double> *)out, d, d.nx / 2, n);
    return launched_st();
}

int launch_rk_update(const Dims &d, const StepConsts &k, int m, double dt, double *U, const double *rp,
                     const double *r, hipStream_t st)
{
    const double dtf2 = k.zeta[m] * dt, dtf1 = k.gamma[m] * dt;
    const double vf = k.vis * (dtf1 + dtf2);
    const unsigned nb = (unsigned)((d.ss + 255) / 256);
    if (d.lu == 4)
        hipLaunchKernelGGL(k_rk_update<4>, dim3(nb), dim3(256), 0, st, (cpx<double> *)U, (const cpx<double> *)rp,
                           (const cpx<double> *)r, d, dtf2, dtf1, vf, k.pr_inv);
    else
        hipLaunchKernelGGL(k_rk_update<3>, dim3(nb), dim3(256), 0, st, (cpx<double> *)U, (const cpx<double> *)rp,
                           (const cpx<double> *)r, d, dtf2, dtf1, vf, k.pr_inv);
    return launched_st();
}

long long div_check_rows(const Dims &d) { return (d.ss + DIV_CHUNK - 1) / DIV_CHUNK; }

int launch_div_check(const Dims &d, const double *U, double *partials, double *out, hipStream_t st)
{
    const int nb = (int)div_check_rows(d);
    hipLaunchKernelGGL(k_div_check, dim3((unsigned)nb), dim3(256), 0, st, (const cpx<double> *)U, d, partials);
    hipLaunchKernelGGL(k_div_final, dim3(1), dim3(256), 0, st, (const double *)partials, nb, out);
    return launched_st();
}

__global__ void k_diag_seed(DevDiag *diag, double dt)
{
    double *w = reinterpret_cast<double *>(diag);
    for (int i = threadIdx.x; i < (int)(sizeof(DevDiag) / sizeof(double)); i += 64) w[i] = 0.0;
    __syncthreads();
    if (threadIdx.x == 0) diag->dt = dt;
}

// sam_vel_max's layout: u_avg[4] u_var[4] u_max[4] energy cfl_{x,y,z}/dt dt
__global__ void k_diag_stats(const DevDiag *__restrict__ diag, double *__restrict__ stats)
{
    const int i = threadIdx.x;
    if (i < 4) {
        stats[i] = diag->u_avg[i];
        stats[4 + i] = diag->u_var[i];
        stats[8 + i] = diag->u_max[i];
    } else if (i == 4) {
        stats[12] = diag->energy;
    } else if (i < 8) {
        stats[13 + (i - 5)] = diag->spare[i - 5];
    } else if (i == 8) {
        stats[16] = diag->dt;
    }
}

int launch_diag_seed(DevDiag *diag, double dt, hipStream_t st)
{
    static_assert(sizeof(DevDiag) % sizeof(double) == 0, "DevDiag is whole doubles");
    hipLaunchKernelGGL(k_diag_seed, dim3(1), dim3(64), 0, st, diag, dt);
    return launched_st();
}

int launch_diag_stats(const DevDiag *diag, double *stats, hipStream_t st)
{
    hipLaunchKernelGGL(k_diag_stats, dim3(1), dim3(64), 0, st, diag, stats);
    return launched_st();
}

template int launch_z2ref<double>(const Dims &, const double *, int, double *, hipStream_t);
template int launch_z2ref<float>(const Dims &, const float *, int, double *, hipStream_t);
template int launch_ref2z<double>(const Dims &, const double *, int, double *, hipStream_t);
template int launch_ref2z<float>(const Dims &, const double *, int, float *, hipStream_t);

} // namespace samhip
