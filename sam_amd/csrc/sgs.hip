// sgs.hip -- subgrid-scale closures of the LES step on the collocated grid.
//
//   k_strain      |S| and S_ij by centred differences (sm.f:52-75, dm.f:80-110)
//   k_dm_fields   |S| S_ij and u_i u_j, the fields dm.f filters (dm.f:101-110)
//   k_filter_xy / k_filter_z
//                 periodic (1,4,1)/6 test filter, separable (filter.f:9-42)
//   k_dm_reduce   L_ij, M_ij and the volume sums <L:M>, <M:M> (dm.f:113-147)
//   k_dm_coef     C Delta^2 = -0.5 <L:M>/<M:M> (dm.f:176), or c_smag Delta^2 (sm.f:32)
//   k_sgs_flux    F_ij = u_i u_j - 2 nu_t S_ij, F_Tj = u_j T - nu_t/Pr_t dT/dx_j
//
// Physical fields are (Nx,Ny,Nz,F) real with x fastest; sizes are powers of
// two so periodic neighbours are index masks.
#include "internal.hpp"

namespace samhip {

static inline int launched_sgs()
{
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? SAM_OK : SAM_ERR_HIP;
}

struct Grid3 {
    int nx, ny, nz;
    long long np;
    __device__ __forceinline__ long long at(int x, int y, int z) const
    {
        return (long long)((z & (nz - 1)) * ny + (y & (ny - 1))) * nx + (x & (nx - 1));
    }
};

// S_n, n = 11,22,33,12,23,13 at (x,y,z) from velocity fields u,v,w
template <typename T>
__device__ __forceinline__ void strain_c(const T *__restrict__ P, const Grid3 &g, int x, int y, int z, double dxi,
                                         double dyi, double dzi, double s[6])
{
    const T *u = P, *v = P + g.np, *w = P + 2 * g.np;
    const double ux = ((double)u[g.at(x + 1, y, z)] - (double)u[g.at(x - 1, y, z)]) * dxi;
    const double uy = ((double)u[g.at(x, y + 1, z)] - (double)u[g.at(x, y - 1, z)]) * dyi;
    const double uz = ((double)u[g.at(x, y, z + 1)] - (double)u[g.at(x, y, z - 1)]) * dzi;
    const double vx = ((double)v[g.at(x + 1, y, z)] - (double)v[g.at(x - 1, y, z)]) * dxi;
    const double vy = ((double)v[g.at(x, y + 1, z)] - (double)v[g.at(x, y - 1, z)]) * dyi;
    const double vz = ((double)v[g.at(x, y, z + 1)] - (double)v[g.at(x, y, z - 1)]) * dzi;
    const double wx = ((double)w[g.at(x + 1, y, z)] - (double)w[g.at(x - 1, y, z)]) * dxi;
    const double wy = ((double)w[g.at(x, y + 1, z)] - (double)w[g.at(x, y - 1, z)]) * dyi;
    const double wz = ((double)w[g.at(x, y, z + 1)] - (double)w[g.at(x, y, z - 1)]) * dzi;
    s[0] = 0.5 * ux;
    s[1] = 0.5 * vy;
    s[2] = 0.5 * wz;
    s[3] = 0.25 * (uy + vx);
    s[4] = 0.25 * (vz + wy);
    s[5] = 0.25 * (wx + uz);
}

__device__ __forceinline__ double smag6(const double s[6])
{
    const double sum = s[0] * s[0] + s[1] * s[1] + s[2] * s[2] + 2.0 * (s[3] * s[3] + s[4] * s[4] + s[5] * s[5]);
    return sqrt(2.0 * sum);
}

// |S| (1 field) and S (6 fields)
template <typename T>
__global__ __launch_bounds__(256) void k_strain(const T *__restrict__ P, T *__restrict__ smag, T *__restrict__ S6,
                                                 Grid3 g, double dxi, double dyi, double dzi)
{
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= g.np) return;
    const int lx = __builtin_ctz(g.nx), ly = __builtin_ctz(g.ny);   // power-of-two sizes: no division
    const int x = (int)(i & (g.nx - 1)), y = (int)((i >> lx) & (g.ny - 1)), z = (int)(i >> (lx + ly));
    double s[6];
    strain_c(P, g, x, y, z, dxi, dyi, dzi, s);
    smag[i] = (T)smag6(s);
    if (S6) {
#pragma unroll
        for (int n = 0; n < 6; n++) S6[n * g.np + i] = (T)s[n];
    }
}

template <typename T>
int launch_strain_smag(const Dims &d, const T *P, T *smag, T *S6, double dxi, double dyi, double dzi, hipStream_t st)
{
    Grid3 g{d.nx, d.ny, d.nz, d.pfield};
    hipLaunchKernelGGL((k_strain<T>), dim3((unsigned)((g.np + 255) / 256)), dim3(256), 0, st, P, smag, S6, g, dxi, dyi,
                       dzi);
    return launched_sgs();
}

// |S| S_n (6) and u_i u_j (6, order 11,22,33,12,23,13)
template <typename T>
__global__ __launch_bounds__(256) void k_dm_fields(const T *__restrict__ P, const T *__restrict__ smag,
                                                    const T *__restrict__ S6, T *__restrict__ MS, T *__restrict__ UU,
                                                    long long np)
{
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= np) return;
    const double sm = (double)smag[i];
    const double u = (double)P[i], v = (double)P[np + i], w = (double)P[2 * np + i];
#pragma unroll
    for (int n = 0; n < 6; n++) MS[n * np + i] = (T)(sm * (double)S6[n * np + i]);
    UU[i] = (T)(u * u);
    UU[np + i] = (T)(v * v);
    UU[2 * np + i] = (T)(w * w);
    UU[3 * np + i] = (T)(u * v);
    UU[4 * np + i] = (T)(v * w);
    UU[5 * np + i] = (T)(u * w);
}

template <typename T>
int launch_dm_fields(const Dims &d, const T *P, const T *smag, const T *S6, T *MS, T *UU, hipStream_t st)
{
    hipLaunchKernelGGL((k_dm_fields<T>), dim3((unsigned)((d.pfield + 255) / 256)), dim3(256), 0, st, P, smag, S6, MS,
                       UU, d.pfield);
    return launched_sgs();
}

// (1,4,1)/6 in x and y (within a z plane), then in z; blockIdx.y = field,
// so one launch per direction filters every field
// Block (256 consecutive points of one x-y plane, z + nz * field): x and y
// from a mask and a shift (power-of-two sizes; nx * ny >= 256), no 64-bit
// division per point; periodic neighbours from masks.  Sums in FP64, as
// filter.f.
template <typename T>
__global__ __launch_bounds__(256) void k_filter_xy(const T *__restrict__ in, T *__restrict__ out, Grid3 g, int lx)
{
    const int e = blockIdx.x * 256 + threadIdx.x;
    const int x = e & (g.nx - 1), y = e >> lx;
    const long long plane = (long long)blockIdx.z * g.ny * g.nx;   // field-major: (field, z) planes
    const int xm = (x - 1) & (g.nx - 1), xp = (x + 1) & (g.nx - 1);
    const T *r[3] = {in + plane + (long long)((y - 1) & (g.ny - 1)) * g.nx, in + plane + (long long)y * g.nx,
                     in + plane + (long long)((y + 1) & (g.ny - 1)) * g.nx};
    double acc = 0.0;
#pragma unroll
    for (int b = 0; b < 3; b++) {
        const double row = (double)r[b][xm] + 4.0 * (double)r[b][x] + (double)r[b][xp];
        acc += (b == 1 ? 4.0 : 1.0) * row;
    }
    out[plane + e] = (T)(acc * (1.0 / 36.0));
}

template <typename T>
__global__ __launch_bounds__(256) void k_filter_z(const T *__restrict__ in, T *__restrict__ out, Grid3 g)
{
    const int e = blockIdx.x * 256 + threadIdx.x;
    const int nzf = blockIdx.z, z = nzf & (g.nz - 1);
    const long long pl = (long long)g.ny * g.nx, fbase = (long long)(nzf - z) * pl;   // field start
    const double v = (double)in[fbase + ((z - 1) & (g.nz - 1)) * pl + e] + 4.0 * (double)in[fbase + z * pl + e] +
                     (double)in[fbase + ((z + 1) & (g.nz - 1)) * pl + e];
    out[fbase + z * pl + e] = (T)(v * (1.0 / 6.0));
}

// filters nvar fields in -> out using tmp (nvar fields) as scratch
template <typename T> int launch_filter(const Dims &d, const T *in, T *out, T *tmp, int nvar, hipStream_t st)
{
    Grid3 g{d.nx, d.ny, d.nz, d.pfield};
    if ((long long)d.nx * d.ny % 256) return SAM_ERR_UNSUPPORTED;
    const dim3 grid((unsigned)((long long)d.nx * d.ny / 256), 1, (unsigned)(d.nz * nvar));
    hipLaunchKernelGGL((k_filter_xy<T>), grid, dim3(256), 0, st, in, tmp, g, __builtin_ctz(d.nx));
    hipLaunchKernelGGL((k_filter_z<T>), grid, dim3(256), 0, st, tmp, out, g);
    return launched_sgs();
}

// L:M and M:M sums.  uh: filtered u (3), sh: filtered S (6),
// mh: filtered |S|S (6), lh: filtered u_i u_j (6); alpha^2 = 4.
template <typename T>
__global__ __launch_bounds__(256) void k_dm_reduce(const T *__restrict__ uh, const T *__restrict__ sh,
                                                    const T *__restrict__ mh, const T *__restrict__ lh, long long np,
                                                    long long chunk, double *__restrict__ partials)
{
    __shared__ double red[2][4];
    const long long b0 = (long long)blockIdx.x * chunk;
    const long long b1 = b0 + chunk < np ? b0 + chunk : np;
    double num = 0.0, den = 0.0;
    const int ip[6] = {0, 1, 2, 0, 1, 0}, jp[6] = {0, 1, 2, 1, 2, 2};
    const double ifact[6] = {1, 1, 1, 2, 2, 2};
    for (long long i = b0 + threadIdx.x; i < b1; i += 256) {
        double s[6], u[3];
#pragma unroll
        for (int n = 0; n < 6; n++) s[n] = (double)sh[n * np + i];
#pragma unroll
        for (int n = 0; n < 3; n++) u[n] = (double)uh[n * np + i];
        const double sm = smag6(s);
#pragma unroll
        for (int n = 0; n < 6; n++) {
            const double M = -(double)mh[n * np + i] + 4.0 * sm * s[n];
            const double L = (double)lh[n * np + i] - u[ip[n]] * u[jp[n]];
            num += ifact[n] * L * M;
            den += ifact[n] * M * M;
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        num += __shfl_xor(num, o, 64);
        den += __shfl_xor(den, o, 64);
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[0][w] = num;
        red[1][w] = den;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        partials[2 * (long long)blockIdx.x] = red[0][0] + red[0][1] + red[0][2] + red[0][3];
        partials[2 * (long long)blockIdx.x + 1] = red[1][0] + red[1][1] + red[1][2] + red[1][3];
    }
}

template <typename T>
int launch_dm_reduce(const Dims &d, const T *uh, const T *sh, const T *mh, const T *lh, double *partials,
                     hipStream_t st, int *nblocks_out)
{
    const long long chunk = 4096;
    const unsigned nb = (unsigned)((d.pfield + chunk - 1) / chunk);
    *nblocks_out = (int)nb;
    hipLaunchKernelGGL((k_dm_reduce<T>), dim3(nb), dim3(256), 0, st, uh, sh, mh, lh, d.pfield, chunk, partials);
    return launched_sgs();
}

__global__ __launch_bounds__(256) void k_dm_coef(const double *__restrict__ partials, int nblocks, double c_fixed,
                                                  int dynamic, int report, DevDiag *__restrict__ diag)
{
    __shared__ double a[256], b[256];
    double s0 = 0.0, s1 = 0.0;
    if (dynamic)
        for (int i = threadIdx.x; i < nblocks; i += 256) {
            s0 += partials[2 * i];
            s1 += partials[2 * i + 1];
        }
    a[threadIdx.x] = s0;
    b[threadIdx.x] = s1;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            a[threadIdx.x] += a[threadIdx.x + o];
            b[threadIdx.x] += b[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double c = dynamic ? -0.5 * a[0] / b[0] : c_fixed;
        diag->c_delta_sq = c;                 // used by the flux kernel of this substep
        if (report) diag->spare[3] = c;       // reported value: first substep of the step
    }
}

int launch_dm_coef(const double *partials, int nblocks, double c_fixed, int dynamic, int report, DevDiag *diag,
                   hipStream_t st)
{
    hipLaunchKernelGGL(k_dm_coef, dim3(1), dim3(256), 0, st, partials, nblocks, c_fixed, dynamic, report, diag);
    return launched_sgs();
}

// fluxes with the eddy viscosity nu_t = C Delta^2 |S| (C Delta^2 in diag)
template <typename T>
__global__ __launch_bounds__(256) void k_sgs_flux(const T *__restrict__ P, const DevDiag *__restrict__ diag,
                                                   T *__restrict__ Q, Grid3 g, double dxi, double dyi, double dzi,
                                                   double prt_inv, int lu)
{
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= g.np) return;
    const int lx = __builtin_ctz(g.nx), ly = __builtin_ctz(g.ny);   // power-of-two sizes: no division
    const int x = (int)(i & (g.nx - 1)), y = (int)((i >> lx) & (g.ny - 1)), z = (int)(i >> (lx + ly));
    double s[6];
    strain_c(P, g, x, y, z, dxi, dyi, dzi, s);
    const double nut = diag->c_delta_sq * smag6(s);
    const double u = (double)P[i], v = (double)P[g.np + i], w = (double)P[2 * g.np + i];
    const long long np = g.np;
    Q[i] = (T)(u * u - 2.0 * nut * s[0]);
    Q[np + i] = (T)(u * v - 2.0 * nut * s[3]);
    Q[2 * np + i] = (T)(v * v - 2.0 * nut * s[1]);
    Q[3 * np + i] = (T)(u * w - 2.0 * nut * s[5]);
    Q[4 * np + i] = (T)(v * w - 2.0 * nut * s[4]);
    Q[5 * np + i] = (T)(w * w - 2.0 * nut * s[2]);
    if (lu == 4) {
        const T *t = P + 3 * np;
        const double tt = (double)t[i];
        const double k = nut * prt_inv * 0.5;
        Q[6 * np + i] = (T)(u * tt - k * ((double)t[g.at(x + 1, y, z)] - (double)t[g.at(x - 1, y, z)]) * dxi);
        Q[7 * np + i] = (T)(v * tt - k * ((double)t[g.at(x, y + 1, z)] - (double)t[g.at(x, y - 1, z)]) * dyi);
        Q[8 * np + i] = (T)(w * tt - k * ((double)t[g.at(x, y, z + 1)] - (double)t[g.at(x, y, z - 1)]) * dzi);
    }
}

template <typename T>
int launch_sgs_flux(const Dims &d, const T *P, const T *smag, const DevDiag *diag, T *Q, double dxi, double dyi,
                    double dzi, double pr_t_inv, hipStream_t st)
{
    (void)smag;
    Grid3 g{d.nx, d.ny, d.nz, d.pfield};
    hipLaunchKernelGGL((k_sgs_flux<T>), dim3((unsigned)((g.np + 255) / 256)), dim3(256), 0, st, P, diag, Q, g, dxi,
                       dyi, dzi, pr_t_inv, d.lu);
    return launched_sgs();
}

template <typename T>
__global__ __launch_bounds__(256) void k_nut(const T *__restrict__ smag, const DevDiag *__restrict__ diag,
                                              double *__restrict__ vt, long long np)
{
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i < np) vt[i] = diag->c_delta_sq * (double)smag[i];
}

template <typename T> int launch_nut(const Dims &d, const T *smag, const DevDiag *diag, double *vt, hipStream_t st)
{
    hipLaunchKernelGGL((k_nut<T>), dim3((unsigned)((d.pfield + 255) / 256)), dim3(256), 0, st, smag, diag, vt,
                       d.pfield);
    return launched_sgs();
}

#define SGS_INST(T)                                                                                                   \
    template int launch_strain_smag<T>(const Dims &, const T *, T *, T *, double, double, double, hipStream_t);       \
    template int launch_filter<T>(const Dims &, const T *, T *, T *, int, hipStream_t);                               \
    template int launch_dm_fields<T>(const Dims &, const T *, const T *, const T *, T *, T *, hipStream_t);           \
    template int launch_dm_reduce<T>(const Dims &, const T *, const T *, const T *, const T *, double *, hipStream_t, \
                                     int *);                                                                          \
    template int launch_sgs_flux<T>(const Dims &, const T *, const T *, const DevDiag *, T *, double, double, double, \
                                    double, hipStream_t);                                                 \
    template int launch_nut<T>(const Dims &, const T *, const DevDiag *, double *, hipStream_t);
SGS_INST(double)
SGS_INST(float)

} // namespace samhip
