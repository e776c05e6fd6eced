// sgs.hip -- subgrid-scale closures of the LES step on the collocated grid.
//
//   k_strain      |S| and S_ij by centred differences (sm.f:52-75, dm.f:80-110)
//   k_dm_fused    the dynamic procedure in one pass over u, v, w: strain,
//                 |S| S_ij, u_i u_j, the (1,4,1)/6 test filter of those and
//                 of u and S, L_ij, M_ij and the block sums of L:M, M:M
//                 (dm.f:60-149, filter.f:9-42)
//   k_filter_xy / k_filter_z
//                 periodic (1,4,1)/6 test filter, separable (filter.f:9-42),
//                 for the filter's own entry point
//   k_dm_coef     C Delta^2 = -0.5 <L:M>/<M:M> (dm.f:176), or c_smag Delta^2 (sm.f:32)
//   k_sgs_flux    F_ij = u_i u_j - 2 nu_t S_ij, F_Tj = u_j T - nu_t/Pr_t dT/dx_j
//                 (z-marching tile; k_sgs_flux_pt, one thread per point, for
//                 grids whose fields fit in the caches)
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

// FP32: the hardware square root (v_sqrt_f32, 1 ulp) without the library's
// correctly rounded fix-up sequence; the argument is a sum of squares, >= 0
template <typename C> __device__ __forceinline__ C smag6c(const C s[6])
{
    const C sum = s[0] * s[0] + s[1] * s[1] + s[2] * s[2] + (C)2 * (s[3] * s[3] + s[4] * s[4] + s[5] * s[5]);
    if constexpr (sizeof(C) == 4) return __builtin_amdgcn_sqrtf((C)2 * sum);
    else return sqrt((C)2 * sum);
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

// ---------------------------------------------------------------------
// k_dm_fused: dm.f's dynamic procedure (dm.f:60-149) as one stencil pass.
//
// A block owns an output tile of DMX x DMY points in x, y and ZC planes in
// z, and marches z.  Per plane p:
//   us  (LDS) u, v, w of planes p-1, p, p+1 on the tile with a 2-point halo,
//       a ring of three planes; plane p+2 is loaded into registers while
//       plane p is worked on and stored into the freed slot of p-1;
//   A   threads (row, x segment of 2 outputs) form the 21 fields dm.f
//       filters -- u_i (3), S_n (6), |S| S_n (6), u_i u_j (6), order
//       11,22,33,12,23,13 -- at the segment's 4 points (with the x halo) and
//       their (1,4,1) sums along x in registers, into xf (LDS), on the tile's
//       rows with a 1-row halo;
//   B   thread (x, y) forms the (1,4,1) sums along y, keeps its column's last
//       two x-y-filtered planes in registers, and the sum of three planes in
//       z completes the filter of plane p-1.  Then L_n = (u_i u_j)^ - u^_i u^_j,
//       M_n = -(|S|S)^_n + 4 |S^| S^_n and the sums of ifact_n L_n M_n and
//       ifact_n M_n M_n; the block's sums go to partials (2 doubles).
// u, v, w are read from HBM once per block (plus the halo) and only the sums
// are written: the procedure's ~150 field passes per substep become about
// four.  The stencil and filter arithmetic runs in the step's precision T
// (FP64, or FP32 in the split mode); the volume sums are FP64.
// ---------------------------------------------------------------------
constexpr int DMY = 8, DMF = 21;

template <typename T, int DMX>
__global__ __launch_bounds__(DMX * DMY, sizeof(T) == 4 ? 1 : 1) void k_dm_fused(const T *__restrict__ P, Grid3 g, int zc, double dxi, double dyi,
                                                         double dzi, double *__restrict__ partials)
{
    using C = T;
    constexpr int NT = DMX * DMY;                 // stage B: one thread per output column
    constexpr int UX = DMX + 4, UY = DMY + 4;     // u region: 2-point halo
    constexpr int PY = DMY + 2;                   // x-filtered rows: 1-row halo
    constexpr int NSEG = DMX / 2, NA = NSEG * PY; // stage A items: (row, 2-point x segment)
    static_assert(NA <= NT, "one stage-A item per thread");
    constexpr int NU = 3 * UX * UY;               // u values of one plane
    constexpr int NPRE = (NU + NT - 1) / NT;      // per thread
    // rows padded by one element: stage A's lanes (x segment s, row r) read
    // and write at 2 s + r * stride, and with strides 36 and 32 the two rows
    // of a half-wavefront fell on the same (even) banks, a 2-way conflict on
    // every stage-A LDS access; with 37 and 33 they take even and odd banks
    __shared__ T us[3][3][UY][UX + 1];
    __shared__ T xf[DMF][PY][DMX + 1];
    __shared__ double red[2][NT / 64];
    const int tid = threadIdx.x;
    // tile of this block, dealt to the XCDs as in k_sgs_flux
    const int ntx = g.nx / DMX, nty = g.ny / DMY, t = xcd_tile(blockIdx.x, gridDim.x);
    const int x0 = (t % ntx) * DMX, y0 = (t / ntx % nty) * DMY, z0 = t / (ntx * nty) * zc;
    const int tx = tid % DMX, ty = tid / DMX;     // stage B column
    const int arow = tid / NSEG, aseg = tid % NSEG;   // stage A item
    const C cdx = (C)dxi, cdy = (C)dyi, cdz = (C)dzi;

    // one u plane (global z = zz, periodic) into registers / into a ring slot
    auto fetch = [&](int zz, T (&v)[NPRE]) {
#pragma unroll
        for (int k = 0; k < NPRE; k++) {
            const int e = tid + k * NT;
            T val = (T)0;
            if (e < NU) {
                const int f = e / (UX * UY), r = e % (UX * UY), iy = r / UX, ix = r % UX;
                val = P[f * g.np + g.at(x0 - 2 + ix, y0 - 2 + iy, zz)];
            }
            v[k] = val;
        }
    };
    auto store = [&](int slot, const T (&v)[NPRE]) {
#pragma unroll
        for (int k = 0; k < NPRE; k++) {
            const int e = tid + k * NT;
            if (e < NU) {
                const int f = e / (UX * UY), r = e % (UX * UY);
                us[slot][f][r / UX][r % UX] = v[k];
            }
        }
    };
    {
        T v[NPRE];
        for (int q = 0; q < 3; q++) {               // planes z0-2, z0-1, z0 -> slots 0, 1, 2
            fetch(z0 - 2 + q, v);
            store(q, v);
        }
    }
    __syncthreads();

    // this column's x-y filtered planes: m1 = plane p-1 and the partial z sum
    // zs = (plane p-2) + 4 (plane p-1); plane p completes the sum of p-1
    T zs[DMF], m1[DMF];
#pragma unroll
    for (int f = 0; f < DMF; f++) zs[f] = m1[f] = (T)0;
    double num = 0.0, den = 0.0;
    const int ip[6] = {0, 1, 2, 0, 1, 0}, jp[6] = {0, 1, 2, 1, 2, 2};
    auto plane = [&](const int p) {
        // ring slots of planes p-1, p, p+1
        const int sm1 = (p - 1 - (z0 - 2)) % 3, s0 = (p - (z0 - 2)) % 3, sp1 = (p + 1 - (z0 - 2)) % 3;
        T nxt[NPRE];
        const bool more = p + 2 <= z0 + zc + 1;
        if (more) fetch(p + 2, nxt);

        // stage A: the segment's 4 product points are tile x = 2 aseg - 1 .. 2 aseg + 2
        // (u columns ux = 2 aseg + 1 + j), on row arow - 1 of the tile (uy = arow + 1)
        if (tid < NA) {
            const int uy = arow + 1, ub = 2 * aseg;
            C a[4][3], sn[4][6], smg[4];
#pragma unroll
            for (int f = 0; f < 3; f++) {
                C c[6], yu[4], yd[4], zu[4], zd[4];
#pragma unroll
                for (int k = 0; k < 6; k++) c[k] = us[s0][f][uy][ub + k];
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    yu[j] = us[s0][f][uy + 1][ub + 1 + j];
                    yd[j] = us[s0][f][uy - 1][ub + 1 + j];
                    zu[j] = us[sp1][f][uy][ub + 1 + j];
                    zd[j] = us[sm1][f][uy][ub + 1 + j];
                }
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    a[j][f] = c[j + 1];
                    const C ddx = (c[j + 2] - c[j]) * cdx, ddy = (yu[j] - yd[j]) * cdy, ddz = (zu[j] - zd[j]) * cdz;
                    // S_n from du_f/dx_k: f = 0: 11 12 13, f = 1: 12 22 23, f = 2: 13 23 33
                    if (f == 0) { sn[j][0] = (C)0.5 * ddx; sn[j][3] = ddy; sn[j][5] = ddz; }
                    if (f == 1) { sn[j][3] = (C)0.25 * (sn[j][3] + ddx); sn[j][1] = (C)0.5 * ddy; sn[j][4] = ddz; }
                    if (f == 2) { sn[j][5] = (C)0.25 * (ddx + sn[j][5]); sn[j][4] = (C)0.25 * (sn[j][4] + ddy); sn[j][2] = (C)0.5 * ddz; }
                }
            }
#pragma unroll
            for (int j = 0; j < 4; j++) smg[j] = smag6c(sn[j]);
            const int xo = 2 * aseg;
            auto put = [&](int f, const C (&v)[4]) {
                xf[f][arow][xo] = v[0] + (C)4 * v[1] + v[2];
                xf[f][arow][xo + 1] = v[1] + (C)4 * v[2] + v[3];
            };
#pragma unroll
            for (int f = 0; f < 3; f++) {
                const C v[4] = {a[0][f], a[1][f], a[2][f], a[3][f]};
                put(f, v);
            }
#pragma unroll
            for (int n = 0; n < 6; n++) {
                const C v1[4] = {sn[0][n], sn[1][n], sn[2][n], sn[3][n]};
                put(3 + n, v1);
                const C v2[4] = {smg[0] * sn[0][n], smg[1] * sn[1][n], smg[2] * sn[2][n], smg[3] * sn[3][n]};
                put(9 + n, v2);
                const C v3[4] = {a[0][ip[n]] * a[0][jp[n]], a[1][ip[n]] * a[1][jp[n]], a[2][ip[n]] * a[2][jp[n]],
                                 a[3][ip[n]] * a[3][jp[n]]};
                put(15 + n, v3);
            }
        }
        __syncthreads();
        // the ring slot of plane p-1 is free: plane p+2 goes there
        if (more) store(sm1, nxt);
        // stage B: y sums -> the x-y filtered plane p of this thread's column;
        // plane p - 1 is then filtered: z sum of planes p-2, p-1, p.  The sums
        // are G = 216 x the filtered fields ((1,4,1)/6 in three directions),
        // and L' = 216^2 L, M' = 216^2 M are formed from them: C = -1/2
        // <L:M>/<M:M> is unchanged by the common factor 216^4
        C fl[DMF];
#pragma unroll
        for (int f = 0; f < DMF; f++) {
            const T cur = xf[f][ty][tx] + (C)4 * xf[f][ty + 1][tx] + xf[f][ty + 2][tx];
            fl[f] = zs[f] + cur;
            zs[f] = m1[f] + (C)4 * cur;
            m1[f] = cur;
        }
        if (p >= z0 + 1) {
            const C smh = smag6c(fl + 3);
            C pn = 0, pd = 0;
#pragma unroll
            for (int n = 0; n < 6; n++) {
                const C M = (C)4 * smh * fl[3 + n] - (C)216 * fl[9 + n];
                const C L = (C)216 * fl[15 + n] - fl[ip[n]] * fl[jp[n]];
                const C w = n < 3 ? (C)1 : (C)2;
                pn += w * L * M;
                pd += w * M * M;
            }
            num += (double)pn;     // the volume sums in FP64 (dm.f:145-147)
            den += (double)pd;
        }
        __syncthreads();           // xf is rewritten and the ring read by the next plane
    };
    for (int p = z0 - 1; p <= z0 + zc; p += 3) {
        plane(p);
        if (p + 1 <= z0 + zc) plane(p + 1);
        if (p + 2 <= z0 + zc) plane(p + 2);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        num += __shfl_xor(num, o, 64);
        den += __shfl_xor(den, o, 64);
    }
    if ((tid & 63) == 0) {
        red[0][tid >> 6] = num;
        red[1][tid >> 6] = den;
    }
    __syncthreads();
    if (tid == 0) {
        double a = 0.0, b = 0.0;
        for (int w = 0; w < NT / 64; w++) {
            a += red[0][w];
            b += red[1][w];
        }
        partials[2 * t] = a;      // in tile order: the coefficient sums them in that order
        partials[2 * t + 1] = b;
    }
}

// Planes per block.  A block computes zc + 2 planes of products for zc
// outputs and the device holds R blocks at once, so a launch takes about
// ceil(blocks / R) * (zc + 2) plane times: zc is the power of two in
// [min(8, nz), 64] that minimises that (the smallest on a tie, for balance).
// 128^3 FP32: 16 planes, 512 blocks in one round (8 planes: 1024 blocks in
// two); 512^3: 64 planes.
static int dm_min_planes(const Dims &d) { return d.nz < 8 ? d.nz : 8; }

template <typename T> static int dm_planes(const Dims &d, int dmx)
{
    const void *kern = dmx == 32 ? (const void *)k_dm_fused<T, 32> : (const void *)k_dm_fused<T, 16>;
    const long long res = resident_blocks(kern, dmx * DMY);
    const long long cols = (long long)(d.nx / dmx) * (d.ny / DMY);
    int best = dm_min_planes(d);
    long long best_cost = -1;
    for (int zc = best; zc <= d.nz && zc <= 64; zc *= 2) {
        const long long nblk = cols * (d.nz / zc);
        const long long cost = (nblk + res - 1) / res * (zc + 2);
        if (best_cost < 0 || cost < best_cost) {
            best = zc;
            best_cost = cost;
        }
    }
    return best;
}

// an upper bound on the blocks of launch_dm_fused (for sizing partials)
long long dm_fused_rows(const Dims &d)
{
    const int dmx = d.nx >= 32 ? 32 : 16;
    return (long long)(d.nx / dmx) * (d.ny / DMY) * (d.nz / dm_min_planes(d));
}

template <typename T>
int launch_dm_fused(const Dims &d, const T *P, double dxi, double dyi, double dzi, double *partials, hipStream_t st,
                    int *nblocks_out)
{
    Grid3 g{d.nx, d.ny, d.nz, d.pfield};
    const int dmx = d.nx >= 32 ? 32 : 16;
    if (d.ny % DMY) return SAM_ERR_UNSUPPORTED;
    const int zc = dm_planes<T>(d, dmx);
    const dim3 grid((unsigned)((d.nx / dmx) * (d.ny / DMY) * (d.nz / zc)));
    *nblocks_out = (int)grid.x;
    if (dmx == 32)
        hipLaunchKernelGGL((k_dm_fused<T, 32>), grid, dim3(32 * DMY), 0, st, P, g, zc, dxi, dyi, dzi, partials);
    else
        hipLaunchKernelGGL((k_dm_fused<T, 16>), grid, dim3(16 * DMY), 0, st, P, g, zc, dxi, dyi, dzi, partials);
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

// The same fluxes, one thread per point, the stencil gathered from global
// memory: for small grids (the four fields within 64 MiB, so the neighbour
// reads hit the caches) this beats the z-marching tile below, whose serial
// plane loop leaves too few blocks in flight (128^3: 24 vs 36 us).
template <typename T>
__global__ __launch_bounds__(256) void k_sgs_flux_pt(const T *__restrict__ P, const DevDiag *__restrict__ diag,
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
    const long long np = g.np;
    const double u = (double)P[i], v = (double)P[np + i], w = (double)P[2 * np + i];
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

// fluxes with the eddy viscosity nu_t = C Delta^2 |S| (C Delta^2 in diag):
// F_ij = u_i u_j - 2 nu_t S_ij, F_Tj = u_j T - (nu_t / Pr_t) dT/dx_j, centred
// differences as strain_c.  A block owns a DMX x FY tile in x, y and ZC
// planes in z and marches z: plane z of u, v, w, T with a 1-point x-y halo is
// in LDS (a ring of two planes; plane z+2 is loaded into registers during
// plane z), each thread keeps its own column's values of planes z-1, z, z+1
// in registers for the z differences.  The fields are read from HBM once
// per block plus the x-y halo, instead of once per stencil point.
constexpr int FY = 16;

template <typename T, int DMX>
__global__ __launch_bounds__(DMX * FY) void k_sgs_flux(const T *__restrict__ P, const DevDiag *__restrict__ diag,
                                                        T *__restrict__ Q, Grid3 g, int zc, double dxi, double dyi,
                                                        double dzi, double prt_inv, int lu)
{
    constexpr int NT = DMX * FY, HX = DMX + 2, HY = FY + 2;
    constexpr int NH = 4 * HX * HY;               // values of one plane (4 fields, halo 1)
    constexpr int NPRE = (NH + NT - 1) / NT;
    __shared__ T pl[2][4][HY][HX];
    const int tid = threadIdx.x, tx = tid % DMX, ty = tid / DMX;
    // tile of this block: x fastest, then y, then the z chunk, dealt to the
    // XCDs in contiguous ranges (neighbouring tiles share their halo lines)
    const int ntx = g.nx / DMX, nty = g.ny / FY, t = xcd_tile(blockIdx.x, gridDim.x);
    const int x0 = (t % ntx) * DMX, y0 = (t / ntx % nty) * FY, z0 = t / (ntx * nty) * zc;
    const int nf = lu;                            // fields read: u, v, w (, T)
    const int nh = nf * HX * HY;
    auto fetch = [&](int zz, T (&v)[NPRE]) {
#pragma unroll
        for (int k = 0; k < NPRE; k++) {
            const int e = tid + k * NT;
            T val = (T)0;
            if (e < nh) {
                const int f = e / (HX * HY), r = e % (HX * HY);
                val = P[f * g.np + g.at(x0 - 1 + r % HX, y0 - 1 + r / HX, zz)];
            }
            v[k] = val;
        }
    };
    auto store = [&](int slot, const T (&v)[NPRE]) {
#pragma unroll
        for (int k = 0; k < NPRE; k++) {
            const int e = tid + k * NT;
            if (e < nh) {
                const int f = e / (HX * HY), r = e % (HX * HY);
                pl[slot][f][r / HX][r % HX] = v[k];
            }
        }
    };
    // own column at plane z0 - 1 (registers), planes z0, z0 + 1 (LDS slots 0, 1)
    double below[4];
#pragma unroll
    for (int f = 0; f < 4; f++)
        below[f] = f < nf ? (double)P[f * g.np + g.at(x0 + tx, y0 + ty, z0 - 1)] : 0.0;
    {
        T v[NPRE];
        fetch(z0, v);
        store(0, v);
        fetch(z0 + 1, v);
        store(1, v);
    }
    __syncthreads();
    const double cds = diag->c_delta_sq;
    const int lx = __builtin_ctz(g.nx), ly = __builtin_ctz(g.ny);
    for (int z = z0; z < z0 + zc; z++) {
        const int sc = (z - z0) & 1, sn = sc ^ 1;       // slots of planes z and z + 1
        T nxt[NPRE];
        const bool more = z + 2 <= z0 + zc;
        if (more) fetch(z + 2, nxt);
        const int hx = tx + 1, hy = ty + 1;
        double c[4], ax[4], ay[4], az[4];
#pragma unroll
        for (int f = 0; f < 4; f++) {
            if (f < nf) {
                c[f] = (double)pl[sc][f][hy][hx];
                ax[f] = ((double)pl[sc][f][hy][hx + 1] - (double)pl[sc][f][hy][hx - 1]) * dxi;
                ay[f] = ((double)pl[sc][f][hy + 1][hx] - (double)pl[sc][f][hy - 1][hx]) * dyi;
                az[f] = ((double)pl[sn][f][hy][hx] - below[f]) * dzi;
            } else {
                c[f] = ax[f] = ay[f] = az[f] = 0.0;
            }
        }
        double s[6];
        s[0] = 0.5 * ax[0];
        s[1] = 0.5 * ay[1];
        s[2] = 0.5 * az[2];
        s[3] = 0.25 * (ay[0] + ax[1]);
        s[4] = 0.25 * (az[1] + ay[2]);
        s[5] = 0.25 * (ax[2] + az[0]);
        const double nut = cds * smag6(s);
        const long long i = ((long long)z << (lx + ly)) + ((long long)(y0 + ty) << lx) + (x0 + tx);
        const long long np = g.np;
        const double u = c[0], v = c[1], w = c[2];
        Q[i] = (T)(u * u - 2.0 * nut * s[0]);
        Q[np + i] = (T)(u * v - 2.0 * nut * s[3]);
        Q[2 * np + i] = (T)(v * v - 2.0 * nut * s[1]);
        Q[3 * np + i] = (T)(u * w - 2.0 * nut * s[5]);
        Q[4 * np + i] = (T)(v * w - 2.0 * nut * s[4]);
        Q[5 * np + i] = (T)(w * w - 2.0 * nut * s[2]);
        if (lu == 4) {
            const double tt = c[3], k = nut * prt_inv * 0.5;
            Q[6 * np + i] = (T)(u * tt - k * ax[3]);
            Q[7 * np + i] = (T)(v * tt - k * ay[3]);
            Q[8 * np + i] = (T)(w * tt - k * az[3]);
        }
#pragma unroll
        for (int f = 0; f < 4; f++) below[f] = c[f];
        __syncthreads();                  // plane z is read by every thread
        if (more) store(sc, nxt);         // plane z + 2 into the slot of plane z
        __syncthreads();
    }
}

template <typename T>
int launch_sgs_flux(const Dims &d, const T *P, const DevDiag *diag, T *Q, double dxi, double dyi, double dzi,
                    double pr_t_inv, hipStream_t st)
{
    Grid3 g{d.nx, d.ny, d.nz, d.pfield};
    if ((long long)d.lu * d.pfield * (long long)sizeof(T) <= (64ll << 20) || d.ny % FY) {
        hipLaunchKernelGGL((k_sgs_flux_pt<T>), dim3((unsigned)((g.np + 255) / 256)), dim3(256), 0, st, P, diag, Q, g,
                           dxi, dyi, dzi, pr_t_inv, d.lu);
        return launched_sgs();
    }
    // 64-point rows: the x halo adds two 64-byte sectors to 256 bytes of a
    // row instead of to 128 (FP32)
    const int dmx = d.nx >= 64 ? 64 : d.nx >= 32 ? 32 : 16;
    const long long cols = (long long)(d.nx / dmx) * (d.ny / FY);
    // z chunks until about 2^22 threads (8192 blocks of 512), at least 8 planes
    const long long want = (8192ll * 512) / (dmx * FY);
    int zc = d.nz;
    while (zc > 8 && cols * (d.nz / zc) < want) zc >>= 1;
    const dim3 grid((unsigned)(cols * (d.nz / zc)));
    if (dmx == 64)
        hipLaunchKernelGGL((k_sgs_flux<T, 64>), grid, dim3(64 * FY), 0, st, P, diag, Q, g, zc, dxi, dyi, dzi,
                           pr_t_inv, d.lu);
    else if (dmx == 32)
        hipLaunchKernelGGL((k_sgs_flux<T, 32>), grid, dim3(32 * FY), 0, st, P, diag, Q, g, zc, dxi, dyi, dzi,
                           pr_t_inv, d.lu);
    else
        hipLaunchKernelGGL((k_sgs_flux<T, 16>), grid, dim3(16 * FY), 0, st, P, diag, Q, g, zc, dxi, dyi, dzi,
                           pr_t_inv, d.lu);
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
    template int launch_dm_fused<T>(const Dims &, const T *, double, double, double, double *, hipStream_t, int *); \
    template int launch_sgs_flux<T>(const Dims &, const T *, const DevDiag *, T *, double, double, double,           \
                                    double, hipStream_t);                                                 \
    template int launch_nut<T>(const Dims &, const T *, const DevDiag *, double *, hipStream_t);
SGS_INST(double)
SGS_INST(float)

} // namespace samhip
