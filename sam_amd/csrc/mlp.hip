// mlp.hip -- learned per-point SGS closure: 12 -> 64 -> 64 -> 64 -> 9, ReLU or GELU.
//
// Inputs (12): resolved velocity gradient du_i/dx_j (9, row-major i,j) and
// dT'/dx_j (3), centred differences on the collocated grid (the same
// operator the Smagorinsky/dynamic kernels use).  Outputs (9): deviatoric
// tau_11, tau_22, tau_33, tau_12, tau_23, tau_13 and the scalar flux q_1..3.
// The SGS fluxes entering the step are F_ij = u_i u_j + tau_ij and
// F_Tj = u_j T' + q_j.
//
// Every layer is a BF16 MFMA GEMM with FP32 accumulation
// (v_mfma_f32_16x16x16_bf16, 64 lanes; lane l = (i = l%16, g = l/16)):
//   A (16x16): lane holds A[i][4g..4g+3]     -- weights, out x in
//   B (16x16): lane holds B[4g..4g+3][i]     -- activations, units x points
//   D (16x16): lane holds D[4g..4g+3][i]
// With the points as the N dimension, a layer's D tile is exactly the next
// layer's B operand, so activations stay in registers from layer to layer;
// the weight tiles (pre-arranged per lane, forward and transposed) sit in
// LDS.  One wavefront processes 16 points per tile.
//
// Scale-free form: a point's scaled features x (12) enter as x/|x| and the
// output is multiplied by |x|^2, so the closure is homogeneous of degree 2
// in the resolved gradients (as gradient-model and eddy-viscosity stresses
// are) and holds outside the gradient range it was trained on.
#include "mlp_core.hpp"

namespace samhip {

static inline int launched_mlp()
{
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? SAM_OK : SAM_ERR_HIP;
}

// the packed image: tile t, lane l at index t*64 + l, then the 208 biases;
// built once per parameter update (k_pack_tiles) and copied into LDS by
// each block
__global__ __launch_bounds__(256) void k_pack_tiles(const float *__restrict__ P, v4s *__restrict__ img)
{
    for (int e = blockIdx.x * 256 + threadIdx.x; e < T_COUNT * 64; e += gridDim.x * 256) {
        const int t = e / 64, l = e % 64, i = l & 15, g = l >> 4;
        float w[4];
        for (int kk = 0; kk < 4; kk++) {
            const int k = 4 * g + kk;
            float v;
            if (t < T_W2) v = wparam(P, MLP_OFF_W1, 16, 16 * t + i, k);
            else if (t < T_W3) { const int b = t - T_W2; v = wparam(P, MLP_OFF_W2, 64, 16 * (b / 4) + i, 16 * (b % 4) + k); }
            else if (t < T_W4) { const int b = t - T_W3; v = wparam(P, MLP_OFF_W3, 64, 16 * (b / 4) + i, 16 * (b % 4) + k); }
            else if (t < T_W4T) { const int kb = t - T_W4; v = wparam(P, MLP_OFF_W4, 64, i, 16 * kb + k); }
            else if (t < T_W3T) { const int mb = t - T_W4T; v = wparam(P, MLP_OFF_W4, 64, k, 16 * mb + i); }
            else if (t < T_W2T) { const int b = t - T_W3T; v = wparam(P, MLP_OFF_W3, 64, 16 * (b % 4) + k, 16 * (b / 4) + i); }
            else { const int b = t - T_W2T; v = wparam(P, MLP_OFF_W2, 64, 16 * (b % 4) + k, 16 * (b / 4) + i); }
            w[kk] = v;
        }
        img[e] = bf4(v4f{w[0], w[1], w[2], w[3]});
    }
    float *bimg = reinterpret_cast<float *>(img + T_COUNT * 64);
    const int boff[3] = {MLP_OFF_B1, MLP_OFF_B2, MLP_OFF_B3};
    for (int e = blockIdx.x * 256 + threadIdx.x; e < NB_BIAS; e += gridDim.x * 256)
        bimg[e] = e < BO_4 ? P[boff[e / 64] + e % 64] : P[MLP_OFF_B4 + e - BO_4];
}

// sum the per-chunk gradients (fixed order) and either return them or take an Adam step
// grid: ceil(MLP_NPARAM / 64) blocks x 4 wavefronts; lane -> parameter,
// wavefront w sums rows w, w + 4, ... and the four sums are added in order
__global__ __launch_bounds__(256) void k_mlp_adam(const float *__restrict__ part, int nch, float *__restrict__ P,
                                                   float *__restrict__ m1, float *__restrict__ m2, float *__restrict__ gout,
                                                   float lr, float b1, float b2, float eps, int step, int apply)
{
    __shared__ float ws[4][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int j = blockIdx.x * 64 + lane;
    float ps = 0.f;
    if (j < MLP_NPARAM)
        for (int c = w; c < nch; c += 4) ps += part[(long long)c * MLP_NPARAM + j];
    ws[w][lane] = ps;
    __syncthreads();
    if (w != 0 || j >= MLP_NPARAM) return;
    float gsum = ((ws[0][lane] + ws[1][lane]) + ws[2][lane]) + ws[3][lane];
    // padded rows/columns of W1 (inputs 12..15) and W4/b4 (outputs 9..15) stay zero
    const bool pad = (j >= MLP_OFF_W1 && j < MLP_OFF_B1 && (j - MLP_OFF_W1) % 16 >= MLP_IN) ||
                     (j >= MLP_OFF_W4 && j < MLP_OFF_B4 && (j - MLP_OFF_W4) / 64 >= MLP_OUT) ||
                     (j >= MLP_OFF_B4 && j - MLP_OFF_B4 >= MLP_OUT);
    if (pad) gsum = 0.f;
    if (gout) gout[j] = gsum;
    if (!apply || pad) return;
    const float mm = b1 * m1[j] + (1.f - b1) * gsum;
    const float vv = b2 * m2[j] + (1.f - b2) * gsum * gsum;
    m1[j] = mm;
    m2[j] = vv;
    const float mh = mm / (1.f - powf(b1, (float)step));
    const float vh = vv / (1.f - powf(b2, (float)step));
    P[j] -= lr * mh / (sqrtf(vh) + eps);
}

// the per-block loss partials in a fixed order (strided per thread, then a tree)
__global__ __launch_bounds__(256) void k_sum_loss(const double *__restrict__ part, int n, double *__restrict__ out)
{
    __shared__ double r[256];
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) s += part[i];
    r[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) r[threadIdx.x] += r[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = r[0];
}

// ---------------------------------------------------------------------
// inference on a sample array: feats [n][16] raw -> out [n][16] physical
// ---------------------------------------------------------------------
__device__ __forceinline__ void deviatoric(int g, v4f &o)
{
    if (g == 0) {
        const float tr = (o[0] + o[1] + o[2]) * (1.0f / 3.0f);
        o[0] -= tr;
        o[1] -= tr;
        o[2] -= tr;
    }
}

template <int ACT>
__global__ __launch_bounds__(256) void k_mlp_infer(const float *__restrict__ P, const v4s *__restrict__ img,
                                                    const float *__restrict__ norms, const float *__restrict__ feats,
                                                    float *__restrict__ out, long long n)
{
    __shared__ v4s tiles[T_W4T * 64];
    __shared__ v4f biasv[NB_BIAS / 4];
    copy_consts(img, tiles, T_W4T, biasv);
    __syncthreads();
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, i = lane & 15, g = lane >> 4;
    const long long ntile = (n + 15) / 16;
    for (long long t = (long long)blockIdx.x * 4 + w; t < ntile; t += (long long)gridDim.x * 4) {
        const long long p = t * 16 + i;
        v4f x = {0, 0, 0, 0};
        if (g < 3 && p < n)
            for (int kk = 0; kk < 4; kk++) x[kk] = feats[p * 16 + 4 * g + kk] * norms[4 * g + kk];
        const float r2 = point_r2(x);
        x *= inv_radius(r2);
        const int zo = opaque_zero();
        Acts A;
        mlp_forward<ACT>(reinterpret_cast<const float *>(biasv) + zo, LdsTiles{tiles + zo, lane}, lane, x, A);
        v4f o = A.out;
        for (int r = 0; r < 4; r++) o[r] *= (4 * g + r < MLP_OUT) ? norms[16 + 4 * g + r] * r2 : 0.f;
        deviatoric(g, o);
        if (p < n) *reinterpret_cast<v4f *>(out + p * 16 + 4 * g) = o;
    }
}

// ---------------------------------------------------------------------
// LES: features from the physical field, MLP, SGS fluxes (fused)
// ---------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ float cdiff(const T *__restrict__ f, int nx, int ny, int nz, int x, int y, int z, int dir,
                                       float dxi)
{
    int xp = x, xm = x, yp = y, ym = y, zp = z, zm = z;
    if (dir == 0) { xp = (x + 1) & (nx - 1); xm = (x - 1) & (nx - 1); }
    if (dir == 1) { yp = (y + 1) & (ny - 1); ym = (y - 1) & (ny - 1); }
    if (dir == 2) { zp = (z + 1) & (nz - 1); zm = (z - 1) & (nz - 1); }
    const float a = (float)f[((long long)zp * ny + yp) * nx + xp];
    const float b = (float)f[((long long)zm * ny + ym) * nx + xm];
    return 0.5f * (a - b) * dxi;
}

// slot of output o in the product/flux arrays (uu,uv,vv,uw,vw,ww,uT,vT,wT)
__device__ __forceinline__ int out_slot(int o)
{
    return o == 0 ? 0 : o == 1 ? 2 : o == 2 ? 5 : o == 3 ? 1 : o == 4 ? 4 : o == 5 ? 3 : o;
}

// product of flux slot s (uu, uv, vv, uw, vw, ww, uT, vT, wT): fields a(s) b(s)
__device__ __forceinline__ int slot_a(int s) { return s < 6 ? (s == 0 || s == 1 || s == 3 ? 0 : s == 5 ? 2 : 1) : s - 6; }
__device__ __forceinline__ int slot_b(int s) { return s < 6 ? (s == 0 ? 0 : s <= 2 ? 1 : 2) : 3; }

// the seven values lane (i, g) reads for a tile: field g at its point and the
// point's x, y, z neighbours (periodic)
template <typename T> struct Stencil {
    T c, xp, xm, yp, ym, zp, zm;
};

// the 16 points of a tile share y and z (nx >= 16, tiles aligned to 16):
// their y and z neighbour offsets are wave-uniform and come from the scalar
// unit; only the x neighbours (periodic at x = 0, nx - 1) differ per lane
template <typename T>
__device__ __forceinline__ void stencil_load_tile(const T *__restrict__ F, int lx, int ly, int lz, unsigned t, int i,
                                                  Stencil<T> &s)
{
    const unsigned tu = __builtin_amdgcn_readfirstlane(t);
    const unsigned mx = (1u << lx) - 1, my = (1u << ly) - 1, mz = (1u << lz) - 1;
    const unsigned base = tu * 16, y = (base >> lx) & my, z = base >> (lx + ly);
    const int oyp = (int)((((y + 1) & my) - y) << lx), oym = (int)((((y - 1) & my) - y) << lx);
    const int ozp = (int)((((z + 1) & mz) - z) << (lx + ly)), ozm = (int)((((z - 1) & mz) - z) << (lx + ly));
    const unsigned x = (base & mx) + (unsigned)i, p = base + (unsigned)i;
    const T *Fp = F + p;
    s.c = Fp[0];
    s.xp = Fp[(int)(((x + 1) & mx) - x)];
    s.xm = Fp[(int)(((x - 1) & mx) - x)];
    s.yp = Fp[oyp];
    s.ym = Fp[oym];
    s.zp = Fp[ozp];
    s.zm = Fp[ozm];
}

// Fused LES flux: features (centred differences), the network and the SGS
// fluxes F_ij = u_i u_j + tau_ij, F_Tj = u_j T' + q_j, one 16-point x-segment
// per wavefront and tile.  Lane (i, g) loads field g's 7-point stencil at
// point i and forms that field's 3 gradient features (k = 3g + j); the 12
// features and 4 point values go through a per-wave LDS slot to the MFMA
// layout (lane (i, g) holds features 4g..4g+3).  The next tile's stencil is
// loaded while the current one is computed.
template <typename T, int ACT>
__global__ __launch_bounds__(256) void k_les_mlp_flux(const v4s *__restrict__ img, const float *__restrict__ norms,
                                                       const T *__restrict__ U, T *__restrict__ Q, int lx, int ly,
                                                       int lz, int lu, float dxi, float dyi, float dzi,
                                                       const double *__restrict__ cscale)
{
    __shared__ v4s tiles[T_W4T * 64];
    __shared__ v4f biasv[NB_BIAS / 4];
    __shared__ v4f featv[4][16 * 3];        // per wave: 16 points x 12 features
    __shared__ T ctr[4][16 * 4];            // per wave: 16 points x (u, v, w, T')
    copy_consts(img, tiles, T_W4T, biasv);
    __syncthreads();
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, i = lane & 15, g = lane >> 4;
    const unsigned np = 1u << (lx + ly + lz);
    const unsigned ntile = np >> 4;
    float *feat = reinterpret_cast<float *>(featv[w]);
    T *cw = ctr[w];
    // the forward weight tiles, held in registers for every tile of the wave
    RegTiles wt;
#pragma unroll
    for (int k = 0; k < T_W4T; k++) wt.w[k] = tiles[k * 64 + lane];

    // lane constants
    const bool has_f = g < lu;
    const T *F = U + (size_t)(has_f ? g : 0) * np;
    const float dd[3] = {dxi, dyi, dzi};
    float nsc[3];
#pragma unroll
    for (int j = 0; j < 3; j++) nsc[j] = has_f ? norms[3 * g + j] : 0.f;
    T *qp[4];
    float osc[4];
    // networks trained on unit-coefficient dynamic-model stresses: times C Delta^2
    const float cs = cscale ? (float)*cscale : 1.0f;
    int ca[4], cb[4];
    bool ok[4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const int oi = 4 * g + r, sl = oi < MLP_OUT ? out_slot(oi) : 0;
        ok[r] = oi < MLP_OUT && (sl < 6 || lu == 4);
        osc[r] = ok[r] ? norms[16 + oi] * cs : 0.f;
        qp[r] = Q + (size_t)sl * np;
        ca[r] = 4 * i + slot_a(sl);
        cb[r] = 4 * i + slot_b(sl);
    }

    unsigned t = blockIdx.x * 4 + w;
    const unsigned stride = gridDim.x * 4;
    Stencil<T> cur;
    if (t < ntile) stencil_load_tile(F, lx, ly, lz, t, i, cur);
    for (; t < ntile; t += stride) {
        const unsigned p = t * 16 + i;
        // this lane's 3 features and point value -> LDS
        tile_sync<true>();
        feat[12 * i + 3 * g + 0] = 0.5f * (float)(cur.xp - cur.xm) * dd[0] * nsc[0];
        feat[12 * i + 3 * g + 1] = 0.5f * (float)(cur.yp - cur.ym) * dd[1] * nsc[1];
        feat[12 * i + 3 * g + 2] = 0.5f * (float)(cur.zp - cur.zm) * dd[2] * nsc[2];
        cw[4 * i + g] = has_f ? cur.c : (T)0;
        // prefetch the next tile
        const bool more = t + stride < ntile;
        if (more) stencil_load_tile(F, lx, ly, lz, t + stride, i, cur);
        tile_sync<true>();
        v4f xin = g < 3 ? *reinterpret_cast<const v4f *>(feat + 12 * i + 4 * g) : v4f{0, 0, 0, 0};
        const float r2 = point_r2(xin);
        xin *= inv_radius(r2);

        const int zo = opaque_zero();
        Acts A;
        mlp_forward<ACT>(reinterpret_cast<const float *>(biasv) + zo, wt, lane, xin, A);
        v4f o = A.out;
#pragma unroll
        for (int r = 0; r < 4; r++) o[r] *= osc[r] * r2;
        deviatoric(g, o);
#pragma unroll
        for (int r = 0; r < 4; r++) {
            if (!ok[r]) continue;
            if constexpr (sizeof(T) == 4) qp[r][p] = __builtin_fmaf(cw[ca[r]], cw[cb[r]], o[r]);
            else qp[r][p] = (T)((double)cw[ca[r]] * (double)cw[cb[r]] + (double)o[r]);
        }
    }
}

// ---------------------------------------------------------------------
// training data: box filter of a fine physical field by `r` in each
// direction -> coarse fields and labels tau (deviatoric) / q
// ---------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_box_filter(const T *__restrict__ U, int nx, int ny, int nz, int lu, int r,
                                                     float *__restrict__ coarse, float *__restrict__ labels)
{
    const int cx = nx / r, cy = ny / r, cz = nz / r;
    const long long nc = (long long)cx * cy * cz;
    const long long c = (long long)blockIdx.x * 256 + threadIdx.x;
    if (c >= nc) return;
    const int X = (int)(c % cx), Yy = (int)((c / cx) % cy), Zz = (int)(c / ((long long)cx * cy));
    const long long np = (long long)nx * ny * nz;
    double m[4] = {0, 0, 0, 0}, pp[6] = {0, 0, 0, 0, 0, 0}, pt[3] = {0, 0, 0};
    for (int dz = 0; dz < r; dz++)
        for (int dy = 0; dy < r; dy++)
            for (int dx = 0; dx < r; dx++) {
                const long long q = ((long long)(Zz * r + dz) * ny + (Yy * r + dy)) * nx + (X * r + dx);
                const double u = U[q], v = U[np + q], w = U[2 * np + q], t = lu == 4 ? (double)U[3 * np + q] : 0.0;
                m[0] += u; m[1] += v; m[2] += w; m[3] += t;
                pp[0] += u * u; pp[1] += v * v; pp[2] += w * w; pp[3] += u * v; pp[4] += v * w; pp[5] += u * w;
                pt[0] += u * t; pt[1] += v * t; pt[2] += w * t;
            }
    const double inv = 1.0 / ((double)r * r * r);
    for (int k = 0; k < 4; k++) m[k] *= inv;
    double tau[6] = {pp[0] * inv - m[0] * m[0], pp[1] * inv - m[1] * m[1], pp[2] * inv - m[2] * m[2],
                     pp[3] * inv - m[0] * m[1], pp[4] * inv - m[1] * m[2], pp[5] * inv - m[0] * m[2]};
    const double tr = (tau[0] + tau[1] + tau[2]) / 3.0;
    tau[0] -= tr; tau[1] -= tr; tau[2] -= tr;
    for (int k = 0; k < lu; k++) coarse[k * nc + c] = (float)m[k];
    float *L = labels + c * 16;
    for (int k = 0; k < 6; k++) L[k] = (float)tau[k];
    for (int k = 0; k < 3; k++) L[6 + k] = lu == 4 ? (float)(pt[k] * inv - m[k] * m[3]) : 0.f;
    for (int k = 9; k < 16; k++) L[k] = 0.f;
}

// Dynamic-model labels at unit coefficient from the raw gradient features
// (dm.f/sm.f stress form): tau_ij = -2 |S| S_ij (deviatoric), q_j =
// -(|S|/Pr_t) dT/dx_j, with S_ij = (du_i/dx_j + du_j/dx_i)/2 and |S| =
// sqrt(2 S_ij S_ij) from the same centred differences as the features.  The
// dynamic procedure supplies C Delta^2 at run time.
// cscale (device, may be null): C Delta^2 the labels are multiplied by
__global__ __launch_bounds__(256) void k_dyn_labels(const float *__restrict__ feats, long long n, float pr_t_inv,
                                                     const double *__restrict__ cscale, float *__restrict__ labels)
{
    const long long p = (long long)blockIdx.x * 256 + threadIdx.x;
    if (p >= n) return;
    const float *F = feats + p * 16;
    double g[3][3], S[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) g[i][j] = F[3 * i + j];
    double s2 = 0.0;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            S[i][j] = 0.5 * (g[i][j] + g[j][i]);
            s2 += S[i][j] * S[i][j];
        }
    const double smag = sqrt(2.0 * s2) * (cscale ? *cscale : 1.0);
    const double tr = (S[0][0] + S[1][1] + S[2][2]) / 3.0;
    float *L = labels + p * 16;
    L[0] = (float)(-2.0 * smag * (S[0][0] - tr));
    L[1] = (float)(-2.0 * smag * (S[1][1] - tr));
    L[2] = (float)(-2.0 * smag * (S[2][2] - tr));
    L[3] = (float)(-2.0 * smag * S[0][1]);
    L[4] = (float)(-2.0 * smag * S[1][2]);
    L[5] = (float)(-2.0 * smag * S[0][2]);
    for (int j = 0; j < 3; j++) L[6 + j] = (float)(-smag * (double)F[9 + j] * (double)pr_t_inv);
    for (int k = 9; k < 16; k++) L[k] = 0.f;
}

// 12 gradient features of a coarse field (centred differences)
__global__ __launch_bounds__(256) void k_coarse_features(const float *__restrict__ C, int nx, int ny, int nz, int lu,
                                                          float dxi, float dyi, float dzi, float *__restrict__ feats)
{
    const long long np = (long long)nx * ny * nz;
    const long long p = (long long)blockIdx.x * 256 + threadIdx.x;
    if (p >= np) return;
    const int x = (int)(p % nx), y = (int)((p / nx) % ny), z = (int)(p / ((long long)nx * ny));
    const float dd[3] = {dxi, dyi, dzi};
    float *F = feats + p * 16;
    for (int k = 0; k < 12; k++) {
        const int fld = k < 9 ? k / 3 : 3, dir = k < 9 ? k % 3 : k - 9;
        F[k] = fld < lu ? cdiff(C + fld * np, nx, ny, nz, x, y, z, dir, dd[dir]) : 0.f;
    }
    for (int k = 12; k < 16; k++) F[k] = 0.f;
}

// mean square of each feature and label column (normalisation), per block
__global__ __launch_bounds__(256) void k_pool_moments(const float *__restrict__ feats, const float *__restrict__ labels,
                                                       long long n, double *__restrict__ part)
{
    __shared__ double red[32][8];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    (void)lane;
    double acc[32];
    for (int k = 0; k < 32; k++) acc[k] = 0.0;
    for (long long p = (long long)blockIdx.x * 256 + threadIdx.x; p < n; p += (long long)gridDim.x * 256) {
        for (int k = 0; k < 16; k++) {
            const double a = feats[p * 16 + k], b = labels[p * 16 + k];
            acc[k] += a * a;
            acc[16 + k] += b * b;
        }
    }
    for (int k = 0; k < 32; k++) {
        double v = acc[k];
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
        if ((threadIdx.x & 63) == 0) red[k][w] = v;
    }
    __syncthreads();
    if (threadIdx.x < 32) part[(long long)blockIdx.x * 32 + threadIdx.x] =
        red[threadIdx.x][0] + red[threadIdx.x][1] + red[threadIdx.x][2] + red[threadIdx.x][3];
}

// norms: features scaled by 1/rms, labels by rms (net output * rms = physical)
__global__ void k_set_norms(const double *__restrict__ part, int nb, long long n, float *__restrict__ norms)
{
    const int k = threadIdx.x;
    if (k >= 32) return;
    double s = 0.0;
    for (int b = 0; b < nb; b++) s += part[(long long)b * 32 + k];
    const double rms = sqrt(s / (double)n);
    if (k < 16) norms[k] = (k < MLP_IN && rms > 0) ? (float)(1.0 / rms) : 0.f;
    else norms[k] = (k - 16 < MLP_OUT && rms > 0) ? (float)rms : 0.f;
}

// ---------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------
int launch_mlp_pack(const float *P, void *img, hipStream_t st)
{
    hipLaunchKernelGGL(k_pack_tiles, dim3((T_COUNT * 64 + 255) / 256), dim3(256), 0, st, P, (v4s *)img);
    return launched_mlp();
}

int launch_mlp_adam(const float *part, int nch, float *P, float *m1, float *m2, float *gout, float lr, int step,
                    int apply, hipStream_t st)
{
    hipLaunchKernelGGL(k_mlp_adam, dim3((MLP_NPARAM + 63) / 64), dim3(256), 0, st, part, nch, P, m1, m2, gout, lr,
                       0.9f, 0.999f, 1e-8f, step, apply);
    return launched_mlp();
}

int launch_sum_loss(const double *part, int n, double *out, hipStream_t st)
{
    hipLaunchKernelGGL(k_sum_loss, dim3(1), dim3(256), 0, st, part, n, out);
    return launched_mlp();
}

int launch_mlp_infer(const float *P, const void *img, const float *norms, const float *feats, float *out, long long n,
                     int act, hipStream_t st)
{
    long long nt = (n + 15) / 16;
    unsigned nb = (unsigned)((nt + 3) / 4);
    if (nb > 2048) nb = 2048;
    if (act == SAM_MLP_ACT_GELU)
        hipLaunchKernelGGL(k_mlp_infer<SAM_MLP_ACT_GELU>, dim3(nb), dim3(256), 0, st, P, (const v4s *)img, norms, feats,
                           out, n);
    else
        hipLaunchKernelGGL(k_mlp_infer<SAM_MLP_ACT_RELU>, dim3(nb), dim3(256), 0, st, P, (const v4s *)img, norms, feats,
                           out, n);
    return launched_mlp();
}

int launch_dyn_labels(const float *feats, long long n, float pr_t_inv, const double *cscale, float *labels,
                      hipStream_t st)
{
    hipLaunchKernelGGL(k_dyn_labels, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, feats, n, pr_t_inv, cscale,
                       labels);
    return launched_mlp();
}

template <typename T>
int launch_les_mlp_flux(const Dims &d, const void *img, const float *norms, const T *U, T *Q, float dxi, float dyi,
                        float dzi, const double *cscale, int act, hipStream_t st)
{
    // 32-bit point indices; x-rows of at least one 16-point tile
    if (d.pfield >= (1ll << 31) || d.nx < 16) return SAM_ERR_UNSUPPORTED;
    long long nt = d.pfield / 16;
    unsigned nb = (unsigned)((nt + 3) / 4);
    if (nb > 2048) nb = 2048;
    if (act == SAM_MLP_ACT_GELU)
        hipLaunchKernelGGL((k_les_mlp_flux<T, SAM_MLP_ACT_GELU>), dim3(nb), dim3(256), 0, st, (const v4s *)img, norms, U,
                           Q, __builtin_ctz(d.nx), __builtin_ctz(d.ny), __builtin_ctz(d.nz), d.lu, dxi, dyi, dzi, cscale);
    else
        hipLaunchKernelGGL((k_les_mlp_flux<T, SAM_MLP_ACT_RELU>), dim3(nb), dim3(256), 0, st, (const v4s *)img, norms, U,
                           Q, __builtin_ctz(d.nx), __builtin_ctz(d.ny), __builtin_ctz(d.nz), d.lu, dxi, dyi, dzi, cscale);
    return launched_mlp();
}

template <typename T>
int launch_box_filter(const Dims &d, const T *U, int r, float *coarse, float *labels, hipStream_t st)
{
    const long long nc = (long long)(d.nx / r) * (d.ny / r) * (d.nz / r);
    hipLaunchKernelGGL((k_box_filter<T>), dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, st, U, d.nx, d.ny, d.nz,
                       d.lu, r, coarse, labels);
    return launched_mlp();
}

int launch_coarse_features(const float *C, int nx, int ny, int nz, int lu, float dxi, float dyi, float dzi,
                           float *feats, hipStream_t st)
{
    const long long np = (long long)nx * ny * nz;
    hipLaunchKernelGGL(k_coarse_features, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, st, C, nx, ny, nz, lu, dxi,
                       dyi, dzi, feats);
    return launched_mlp();
}

int launch_pool_norms(const float *feats, const float *labels, long long n, double *part, float *norms, hipStream_t st)
{
    const int nb = 256;
    hipLaunchKernelGGL(k_pool_moments, dim3(nb), dim3(256), 0, st, feats, labels, n, part);
    hipLaunchKernelGGL(k_set_norms, dim3(1), dim3(64), 0, st, part, nb, n, norms);
    return launched_mlp();
}

template int launch_les_mlp_flux<double>(const Dims &, const void *, const float *, const double *, double *, float,
                                         float, float, const double *, int, hipStream_t);
template int launch_les_mlp_flux<float>(const Dims &, const void *, const float *, const float *, float *, float,
                                        float, float, const double *, int, hipStream_t);
template int launch_box_filter<double>(const Dims &, const double *, int, float *, float *, hipStream_t);
template int launch_box_filter<float>(const Dims &, const float *, int, float *, float *, hipStream_t);

} // namespace samhip
