// mlp.hip -- learned per-point SGS closure: 12 -> 64 -> 64 -> 64 -> 9, GELU.
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
#include "internal.hpp"

namespace samhip {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

static inline int launched_mlp()
{
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? SAM_OK : SAM_ERR_HIP;
}

__device__ __forceinline__ short f2bf(float f)
{
    unsigned u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);   // round to nearest even
    return (short)(u >> 16);
}
__device__ __forceinline__ float bf2f(short h) { return __uint_as_float(((unsigned)(unsigned short)h) << 16); }

// a zero the compiler cannot see through: re-basing the LDS weight tiles and
// bias pointers on it each iteration keeps them from being hoisted out of the
// per-tile loops into registers
__device__ __forceinline__ int opaque_zero()
{
    int z = 0;
    asm volatile("" : "+v"(z));
    return z;
}

__device__ __forceinline__ v4s bf4(v4f x) { return v4s{f2bf(x[0]), f2bf(x[1]), f2bf(x[2]), f2bf(x[3])}; }

__device__ __forceinline__ v4f mfma(v4s a, v4s b, v4f c) { return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0); }

// GELU, tanh form, and its derivative
__device__ __forceinline__ float gelu_fwd(float x, float &dg)
{
    const float c = 0.7978845608028654f, a = 0.044715f;
    // tanh(u) = 1 - 2/(1 + exp(2u)); exact limits +-1 for large |u|
    const float t = 1.0f - __fdividef(2.0f, 1.0f + __expf(2.0f * c * (x + a * x * x * x)));
    dg = 0.5f * (1.0f + t) + 0.5f * x * (1.0f - t * t) * c * (1.0f + 3.0f * a * x * x);
    return 0.5f * x * (1.0f + t);
}

// ---- weight tiles in LDS ---------------------------------------------
// tile t (64 lanes x 4 bf16): forward tiles  0..3  W1 (mb),  4..19 W2 (mb*4+kb),
// 20..35 W3, 36..39 W4 (kb); transposed tiles 40..43 W4^T (mb), 44..59 W3^T,
// 60..75 W2^T.
constexpr int T_W1 = 0, T_W2 = 4, T_W3 = 20, T_W4 = 36, T_W4T = 40, T_W3T = 44, T_W2T = 60, T_COUNT = 76;

__device__ __forceinline__ float wparam(const float *P, int off, int ld, int row, int col) { return P[off + row * ld + col]; }

// the packed image: tile t, lane l at index t*64 + l, built once per
// parameter update (k_pack_tiles) and copied into LDS by each block
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
        img[e] = v4s{f2bf(w[0]), f2bf(w[1]), f2bf(w[2]), f2bf(w[3])};
    }
}

// copy the first `ntiles` tiles of the packed image into LDS (16 B per thread)
__device__ __forceinline__ void copy_tiles(const v4s *__restrict__ img, v4s *tiles, int ntiles)
{
    const uint4 *src = reinterpret_cast<const uint4 *>(img);
    uint4 *dst = reinterpret_cast<uint4 *>(tiles);
    for (int e = threadIdx.x; e < ntiles * 32; e += blockDim.x) dst[e] = src[e];
}

__device__ __forceinline__ v4f bias4(const float *__restrict__ P, int off, int u0)
{
    return v4f{P[off + u0], P[off + u0 + 1], P[off + u0 + 2], P[off + u0 + 3]};
}

struct Acts {
    v4s a0;          // inputs (bf16)
    v4f z[3][4];     // pre-activations of the hidden layers
    v4s a[3][4];     // hidden activations (bf16)
    v4f out;         // outputs 4g..4g+3
};

__device__ __forceinline__ v4s gelu_bf4(v4f z)
{
    float d;
    return v4s{f2bf(gelu_fwd(z[0], d)), f2bf(gelu_fwd(z[1], d)), f2bf(gelu_fwd(z[2], d)), f2bf(gelu_fwd(z[3], d))};
}

// one hidden layer 64 -> 64: z = W a + b, tiles tb..tb+15 (mb*4+kb)
__device__ __forceinline__ void fw_hidden(const float *__restrict__ P, const v4s *tiles, int tb, int boff, int lane,
                                          const v4s *ain, v4f *z, v4s *aout)
{
    const int g = lane >> 4;
#pragma unroll
    for (int mb = 0; mb < 4; mb++) {
        v4f acc = bias4(P, boff, 16 * mb + 4 * g);
#pragma unroll
        for (int kb = 0; kb < 4; kb++) acc = mfma(tiles[(tb + mb * 4 + kb) * 64 + lane], ain[kb], acc);
        z[mb] = acc;
        aout[mb] = gelu_bf4(acc);
    }
}

// forward pass of one 16-point tile; x: this lane's 4 normalised inputs
__device__ __forceinline__ void mlp_forward(const float *__restrict__ P, const v4s *tiles, int lane, v4f x, Acts &A)
{
    const int g = lane >> 4;
    A.a0 = bf4(x);
#pragma unroll
    for (int mb = 0; mb < 4; mb++) {
        const v4f z = mfma(tiles[(T_W1 + mb) * 64 + lane], A.a0, bias4(P, MLP_OFF_B1, 16 * mb + 4 * g));
        A.z[0][mb] = z;
        A.a[0][mb] = gelu_bf4(z);
    }
    fw_hidden(P, tiles, T_W2, MLP_OFF_B2, lane, A.a[0], A.z[1], A.a[1]);
    fw_hidden(P, tiles, T_W3, MLP_OFF_B3, lane, A.a[1], A.z[2], A.a[2]);
    v4f o = bias4(P, MLP_OFF_B4, 4 * g);
#pragma unroll
    for (int kb = 0; kb < 4; kb++) o = mfma(tiles[(T_W4 + kb) * 64 + lane], A.a[2][kb], o);
    A.out = o;
}

__device__ __forceinline__ v4f gelu_grad4(v4f z)
{
    v4f d;
    float t;
    gelu_fwd(z[0], t); d[0] = t;
    gelu_fwd(z[1], t); d[1] = t;
    gelu_fwd(z[2], t); d[2] = t;
    gelu_fwd(z[3], t); d[3] = t;
    return d;
}

__device__ __forceinline__ void bw_gelu(const v4f *z, const v4f *da, v4s *dz)
{
#pragma unroll
    for (int mb = 0; mb < 4; mb++) {
        const v4f gd = gelu_grad4(z[mb]);
        dz[mb] = bf4(v4f{da[mb][0] * gd[0], da[mb][1] * gd[1], da[mb][2] * gd[2], da[mb][3] * gd[3]});
    }
}

__device__ __forceinline__ void bw_matvec(const v4s *tiles, int tb, int lane, const v4s *dz, v4f *da)
{
#pragma unroll
    for (int mb = 0; mb < 4; mb++) {
        v4f s = {0, 0, 0, 0};
#pragma unroll
        for (int kb = 0; kb < 4; kb++) s = mfma(tiles[(tb + mb * 4 + kb) * 64 + lane], dz[kb], s);
        da[mb] = s;
    }
}

// backward of one tile from dL/d(out); returns dz of each hidden layer (bf16)
__device__ __forceinline__ void mlp_backward(const v4s *tiles, int lane, const Acts &A, v4f dout, v4s dz[3][4],
                                             v4s &dout_b)
{
    dout_b = bf4(dout);
    v4f da[4];
#pragma unroll
    for (int mb = 0; mb < 4; mb++) da[mb] = mfma(tiles[(T_W4T + mb) * 64 + lane], dout_b, v4f{0, 0, 0, 0});
    bw_gelu(A.z[2], da, dz[2]);
    bw_matvec(tiles, T_W3T, lane, dz[2], da);
    bw_gelu(A.z[1], da, dz[1]);
    bw_matvec(tiles, T_W2T, lane, dz[1], da);
    bw_gelu(A.z[0], da, dz[0]);
}

// ---------------------------------------------------------------------
// training forward+backward over a minibatch of the sample pool
// pool: feats [p][16] raw (12 used), labels [p][16] raw (9 used)
// stores (point-major, bf16): a0 [p][16], a1..a3 [p][64], dz1..dz3 [p][64], dout [p][16]
// ---------------------------------------------------------------------
struct TrainBufs {
    short *a0, *a1, *a2, *a3, *dz1, *dz2, *dz3, *dout;
};

__device__ __forceinline__ void store4(short *base, int ld, int p, int u0, v4s v)
{
    *reinterpret_cast<v4s *>(base + (long long)p * ld + u0) = v;
}

__global__ __launch_bounds__(256) void k_mlp_fb(const float *__restrict__ P, const v4s *__restrict__ img,
                                                 const float *__restrict__ norms,
                                                 const float *__restrict__ feats, const float *__restrict__ labels,
                                                 long long npool, long long base, long long pa, long long pb,
                                                 int batch, TrainBufs tb, double *__restrict__ loss_part)
{
    __shared__ v4s tiles[T_COUNT * 64];
    __shared__ double lred[4];
    copy_tiles(img, tiles, T_COUNT);
    __syncthreads();
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, i = lane & 15, g = lane >> 4;
    const int ntiles = batch / 16;
    double lsum = 0.0;
    for (int t = blockIdx.x * 4 + w; t < ntiles; t += gridDim.x * 4) {
        const int p = t * 16 + i;                                  // position in the batch
        // sample index: pa, base, pb < npool < 2^31, so the product stays below 2^62
        const long long s = (pa * ((base + p) % npool) + pb) % npool;
        v4f x = {0, 0, 0, 0};
        if (g < 3) {
#pragma unroll
            for (int kk = 0; kk < 4; kk++) x[kk] = feats[s * 16 + 4 * g + kk] * norms[4 * g + kk];
        }
        const int zo = opaque_zero();
        const v4s *tl = tiles + zo;
        const float *Pl = P + zo;
        Acts A;
        mlp_forward(Pl, tl, lane, x, A);
        // MSE over the 9 normalised outputs: L = mean over batch*9 of (out - y/os)^2
        v4f dout;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int o = 4 * g + r;
            float d = 0.f;
            if (o < MLP_OUT) {
                const float y = labels[s * 16 + o] / norms[16 + o];
                d = A.out[r] - y;
                lsum += (double)d * d;
            }
            dout[r] = 2.0f * d / (float)(batch * MLP_OUT);
        }
        v4s dz[3][4], dob;
        mlp_backward(tl, lane, A, dout, dz, dob);
        store4(tb.a0, 16, p, 4 * g, A.a0);
        store4(tb.dout, 16, p, 4 * g, dob);
#pragma unroll
        for (int mb = 0; mb < 4; mb++) {
            store4(tb.a1, 64, p, 16 * mb + 4 * g, A.a[0][mb]);
            store4(tb.a2, 64, p, 16 * mb + 4 * g, A.a[1][mb]);
            store4(tb.a3, 64, p, 16 * mb + 4 * g, A.a[2][mb]);
            store4(tb.dz1, 64, p, 16 * mb + 4 * g, dz[0][mb]);
            store4(tb.dz2, 64, p, 16 * mb + 4 * g, dz[1][mb]);
            store4(tb.dz3, 64, p, 16 * mb + 4 * g, dz[2][mb]);
        }
    }
    for (int o = 32; o > 0; o >>= 1) lsum += __shfl_xor(lsum, o, 64);
    if (lane == 0) lred[w] = lsum;
    __syncthreads();
    if (threadIdx.x == 0) loss_part[blockIdx.x] = lred[0] + lred[1] + lred[2] + lred[3];
}

// ---------------------------------------------------------------------
// weight gradients: dW[i][j] = sum_p dz[p][i] a[p][j], db[i] = sum_p dz[p][i]
// (the bias as one more N block of ones).  One wavefront per chunk of points.
// ---------------------------------------------------------------------
template <int MB, int NB>
__device__ void wgrad_layer(const short *__restrict__ dz, int ldz, const short *__restrict__ a, int lda, int p0, int np,
                            float *__restrict__ outW, float *__restrict__ outB, int ldw)
{
    const int lane = threadIdx.x & 63, i = lane & 15, g = lane >> 4;
    v4f acc[MB][NB + 1];
#pragma unroll
    for (int mb = 0; mb < MB; mb++)
#pragma unroll
        for (int nb = 0; nb <= NB; nb++) acc[mb][nb] = v4f{0, 0, 0, 0};
    const short one = (short)0x3f80;
    const v4s ones = {one, one, one, one};
    for (int k0 = p0; k0 < p0 + np; k0 += 16) {
        v4s A[MB], B[NB];
#pragma unroll
        for (int mb = 0; mb < MB; mb++)
            for (int kk = 0; kk < 4; kk++) A[mb][kk] = dz[(long long)(k0 + 4 * g + kk) * ldz + 16 * mb + i];
#pragma unroll
        for (int nb = 0; nb < NB; nb++)
            for (int kk = 0; kk < 4; kk++) B[nb][kk] = a[(long long)(k0 + 4 * g + kk) * lda + 16 * nb + i];
#pragma unroll
        for (int mb = 0; mb < MB; mb++) {
#pragma unroll
            for (int nb = 0; nb < NB; nb++) acc[mb][nb] = mfma(A[mb], B[nb], acc[mb][nb]);
            acc[mb][NB] = mfma(A[mb], ones, acc[mb][NB]);
        }
    }
#pragma unroll
    for (int mb = 0; mb < MB; mb++) {
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int row = 16 * mb + 4 * g + r;
#pragma unroll
            for (int nb = 0; nb < NB; nb++) outW[row * ldw + 16 * nb + i] = acc[mb][nb][r];
            if (i == 0) outB[row] = acc[mb][NB][r];
        }
    }
}

// grid: nchunks blocks x 4 waves; wave w handles layer w of chunk blockIdx.x
__global__ __launch_bounds__(256) void k_mlp_wgrad(TrainBufs tb, int batch, float *__restrict__ part)
{
    const int nch = gridDim.x;
    const int np = batch / nch;
    const int p0 = blockIdx.x * np;
    float *out = part + (long long)blockIdx.x * MLP_NPARAM;
    const int w = threadIdx.x >> 6;
    if (w == 0) wgrad_layer<4, 1>(tb.dz1, 64, tb.a0, 16, p0, np, out + MLP_OFF_W1, out + MLP_OFF_B1, 16);
    if (w == 1) wgrad_layer<4, 4>(tb.dz2, 64, tb.a1, 64, p0, np, out + MLP_OFF_W2, out + MLP_OFF_B2, 64);
    if (w == 2) wgrad_layer<4, 4>(tb.dz3, 64, tb.a2, 64, p0, np, out + MLP_OFF_W3, out + MLP_OFF_B3, 64);
    if (w == 3) wgrad_layer<1, 4>(tb.dout, 16, tb.a3, 64, p0, np, out + MLP_OFF_W4, out + MLP_OFF_B4, 64);
}

// sum the per-chunk gradients (fixed order) and either return them or take an Adam step
__global__ __launch_bounds__(256) void k_mlp_adam(const float *__restrict__ part, int nch, float *__restrict__ P,
                                                   float *__restrict__ m1, float *__restrict__ m2, float *__restrict__ gout,
                                                   float lr, float b1, float b2, float eps, int step, int apply)
{
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= MLP_NPARAM) return;
    float gsum = 0.f;
    for (int c = 0; c < nch; c++) gsum += part[(long long)c * MLP_NPARAM + j];
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

__global__ void k_sum_loss(const double *__restrict__ part, int n, double *__restrict__ out)
{
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int i = 0; i < n; i++) s += part[i];
        *out = s;
    }
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

__global__ __launch_bounds__(256) void k_mlp_infer(const float *__restrict__ P, const v4s *__restrict__ img,
                                                    const float *__restrict__ norms, const float *__restrict__ feats,
                                                    float *__restrict__ out, long long n)
{
    __shared__ v4s tiles[T_W4T * 64];
    copy_tiles(img, tiles, T_W4T);
    __syncthreads();
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, i = lane & 15, g = lane >> 4;
    const long long ntile = (n + 15) / 16;
    for (long long t = (long long)blockIdx.x * 4 + w; t < ntile; t += (long long)gridDim.x * 4) {
        const long long p = t * 16 + i;
        v4f x = {0, 0, 0, 0};
        if (g < 3 && p < n)
            for (int kk = 0; kk < 4; kk++) x[kk] = feats[p * 16 + 4 * g + kk] * norms[4 * g + kk];
        const int zo = opaque_zero();
        Acts A;
        mlp_forward(P + zo, tiles + zo, lane, x, A);
        v4f o = A.out;
        for (int r = 0; r < 4; r++) o[r] *= (4 * g + r < MLP_OUT) ? norms[16 + 4 * g + r] : 0.f;
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

// inputs one lane needs for a tile: the two neighbours of its 4 gradient
// features and the point's u, v, w, T'
template <typename T> struct FluxIn {
    T nb[8];
    T c[4];
};

template <typename T>
__device__ __forceinline__ void flux_load(const T *__restrict__ U, int nx, int ny, int nz, int lu, long long np,
                                          long long p, int g, FluxIn<T> &in)
{
    const int x = (int)(p % nx), y = (int)((p / nx) % ny), z = (int)(p / ((long long)nx * ny));
#pragma unroll
    for (int kk = 0; kk < 4; kk++) {
        const int k = 4 * g + kk;                 // 0..8: du_i/dx_j (i = k/3, j = k%3); 9..11: dT/dx_j
        const int fld = k < 9 ? k / 3 : 3, dir = k < 9 ? k % 3 : k - 9;
        T a = (T)0, b = (T)0;
        if (g < 3 && fld < lu) {
            int xp = x, xm = x, yp = y, ym = y, zp = z, zm = z;
            if (dir == 0) { xp = (x + 1) & (nx - 1); xm = (x - 1) & (nx - 1); }
            if (dir == 1) { yp = (y + 1) & (ny - 1); ym = (y - 1) & (ny - 1); }
            if (dir == 2) { zp = (z + 1) & (nz - 1); zm = (z - 1) & (nz - 1); }
            const T *f = U + fld * np;
            a = f[((long long)zp * ny + yp) * nx + xp];
            b = f[((long long)zm * ny + ym) * nx + xm];
        }
        in.nb[2 * kk] = a;
        in.nb[2 * kk + 1] = b;
    }
    in.c[0] = U[p];
    in.c[1] = U[np + p];
    in.c[2] = U[2 * np + p];
    in.c[3] = lu == 4 ? U[3 * np + p] : (T)0;
}

template <typename T>
__global__ __launch_bounds__(256) void k_les_mlp_flux(const float *__restrict__ Pm, const v4s *__restrict__ img,
                                                       const float *__restrict__ norms, const T *__restrict__ U,
                                                       T *__restrict__ Q, int nx, int ny, int nz, int lu, float dxi,
                                                       float dyi, float dzi)
{
    __shared__ v4s tiles[T_W4T * 64];
    copy_tiles(img, tiles, T_W4T);
    __syncthreads();
    const long long np = (long long)nx * ny * nz;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, i = lane & 15, g = lane >> 4;
    const long long ntile = np / 16;
    const long long stride = (long long)gridDim.x * 4;
    const float dd[3] = {dxi, dyi, dzi};
    long long t = (long long)blockIdx.x * 4 + w;
    FluxIn<T> cur;
    if (t < ntile) flux_load(U, nx, ny, nz, lu, np, t * 16 + i, g, cur);
    for (; t < ntile; t += stride) {
        const long long p = t * 16 + i;
        FluxIn<T> nxt;                      // prefetch the next tile while this one is computed
        const bool more = t + stride < ntile;
        if (more) flux_load(U, nx, ny, nz, lu, np, (t + stride) * 16 + i, g, nxt);
        v4f xin = {0, 0, 0, 0};
#pragma unroll
        for (int kk = 0; kk < 4; kk++) {
            const int k = 4 * g + kk;
            const int dir = k < 9 ? k % 3 : (k < 12 ? k - 9 : 0);
            const float v = 0.5f * (float)(cur.nb[2 * kk] - cur.nb[2 * kk + 1]) * dd[dir];
            xin[kk] = g < 3 ? v * norms[k] : 0.f;
        }
        const int zo = opaque_zero();
        Acts A;
        mlp_forward(Pm + zo, tiles + zo, lane, xin, A);
        v4f o = A.out;
#pragma unroll
        for (int r = 0; r < 4; r++) o[r] *= (4 * g + r < MLP_OUT) ? norms[16 + 4 * g + r] : 0.f;
        deviatoric(g, o);
        const double uu[4] = {(double)cur.c[0], (double)cur.c[1], (double)cur.c[2], (double)cur.c[3]};
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int oi = 4 * g + r;
            if (oi >= MLP_OUT) continue;
            const int slot = out_slot(oi);
            if (slot >= 6 && lu < 4) continue;
            // product of the slot: (uu,uv,vv,uw,vw,ww,uT,vT,wT)
            const int a = slot == 0 ? 0 : slot == 1 ? 0 : slot == 2 ? 1 : slot == 3 ? 0 : slot == 4 ? 1 : slot == 5 ? 2
                        : slot == 6 ? 0 : slot == 7 ? 1 : 2;
            const int b = slot == 0 ? 0 : slot == 1 ? 1 : slot == 2 ? 1 : slot <= 5 ? 2 : 3;
            Q[slot * np + p] = (T)(uu[a] * uu[b] + (double)o[r]);
        }
        if (more) cur = nxt;
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

int launch_mlp_fb(const float *P, const void *img, const float *norms, const float *feats, const float *labels,
                  long long npool, long long base, long long pa, long long pb, int batch, const MlpTrainBufs &b,
                  double *loss_part, int nblk, hipStream_t st)
{
    TrainBufs tb{b.a0, b.a1, b.a2, b.a3, b.dz1, b.dz2, b.dz3, b.dout};
    hipLaunchKernelGGL(k_mlp_fb, dim3(nblk), dim3(256), 0, st, P, (const v4s *)img, norms, feats, labels, npool, base,
                       pa, pb, batch, tb, loss_part);
    return launched_mlp();
}

int launch_mlp_wgrad(const MlpTrainBufs &b, int batch, float *part, int nchunks, hipStream_t st)
{
    TrainBufs tb{b.a0, b.a1, b.a2, b.a3, b.dz1, b.dz2, b.dz3, b.dout};
    hipLaunchKernelGGL(k_mlp_wgrad, dim3(nchunks), dim3(256), 0, st, tb, batch, part);
    return launched_mlp();
}

int launch_mlp_adam(const float *part, int nch, float *P, float *m1, float *m2, float *gout, float lr, int step,
                    int apply, hipStream_t st)
{
    hipLaunchKernelGGL(k_mlp_adam, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, st, part, nch, P, m1, m2, gout, lr,
                       0.9f, 0.999f, 1e-8f, step, apply);
    return launched_mlp();
}

int launch_sum_loss(const double *part, int n, double *out, hipStream_t st)
{
    hipLaunchKernelGGL(k_sum_loss, dim3(1), dim3(64), 0, st, part, n, out);
    return launched_mlp();
}

int launch_mlp_infer(const float *P, const void *img, const float *norms, const float *feats, float *out, long long n,
                     hipStream_t st)
{
    long long nt = (n + 15) / 16;
    unsigned nb = (unsigned)((nt + 3) / 4);
    if (nb > 2048) nb = 2048;
    hipLaunchKernelGGL(k_mlp_infer, dim3(nb), dim3(256), 0, st, P, (const v4s *)img, norms, feats, out, n);
    return launched_mlp();
}

template <typename T>
int launch_les_mlp_flux(const Dims &d, const float *Pm, const void *img, const float *norms, const T *U, T *Q,
                        float dxi, float dyi, float dzi, hipStream_t st)
{
    long long nt = d.pfield / 16;
    unsigned nb = (unsigned)((nt + 3) / 4);
    if (nb > 2048) nb = 2048;
    hipLaunchKernelGGL((k_les_mlp_flux<T>), dim3(nb), dim3(256), 0, st, Pm, (const v4s *)img, norms, U, Q, d.nx, d.ny,
                       d.nz, d.lu, dxi, dyi, dzi);
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

template int launch_les_mlp_flux<double>(const Dims &, const float *, const void *, const float *, const double *,
                                         double *, float, float, float, hipStream_t);
template int launch_les_mlp_flux<float>(const Dims &, const float *, const void *, const float *, const float *,
                                        float *, float, float, float, hipStream_t);
template int launch_box_filter<double>(const Dims &, const double *, int, float *, float *, hipStream_t);
template int launch_box_filter<float>(const Dims &, const float *, int, float *, float *, hipStream_t);

} // namespace samhip
