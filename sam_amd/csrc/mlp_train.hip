// mlp_train.hip -- one training pass of the learned closure as one kernel:
// forward, loss, backward and the weight gradients of a minibatch.
//
// A wavefront takes 16-point tiles of the minibatch.  Forward and backward
// are the per-tile MFMA chains of mlp_core.hpp (activations and
// pre-activation gradients stay in registers, in the layout where the points
// are the MFMA N dimension: lane (i, g) holds units 4g..4g+3 of point i).
// The weight gradient of a layer, dW[o][in] = sum_p dz[p][o] a[p][in], needs
// the points as the K dimension instead (lane (i, g) holds points 4g..4g+3 of
// unit i).  That transpose of a 16x16 block is one MFMA: the block as the A
// operand (A[m][k] = X[k][m]) times the identity gives D = X^T, exact in the
// FP32 accumulator.  The 53 gradient tiles of the network (dW1 4, dW2 16,
// dW3 16, dW4 4, biases 13, the bias as a ones operand) then accumulate in
// registers across the wavefront's tiles; a block's four wavefronts add
// theirs in LDS and write one partial gradient row, and k_mlp_adam sums the
// rows in a fixed order.
//
// The activations and gradients never go to memory: the two-kernel form
// (forward/backward storing BF16 activations, then a gradient GEMM reading
// them back) moved 0.87 GB per 2^20-point batch twice.
//
// This file is built without the VGPR form of the MFMA results, so the
// compiler may keep the 212 accumulator values per lane in AGPRs.
#include "mlp_core.hpp"

namespace samhip {

static inline int launched_train()
{
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? SAM_OK : SAM_ERR_HIP;
}

// 16x16 identity as the B operand: lane (i, g) holds B[4g..4g+3][i]
__device__ __forceinline__ v4s ident_b(int lane)
{
    const int i = lane & 15, g = lane >> 4;
    const short one = (short)0x3f80, zero = 0;
    return v4s{4 * g + 0 == i ? one : zero, 4 * g + 1 == i ? one : zero, 4 * g + 2 == i ? one : zero,
               4 * g + 3 == i ? one : zero};
}

// transpose of a 16x16 BF16 block (units x points -> points in K)
__device__ __forceinline__ v4s tr16(v4s x, v4s id) { return bf4(mfma(x, id, v4f{0, 0, 0, 0})); }

// gradient tile (rows 16 mb + 4g + r, columns 16 nb + i) into the block's
// LDS gradient row: stored by the first wavefront, added by the others
__device__ __forceinline__ void put_tile(float *out, int ld, int mb, int nb, int lane, v4f v, bool add)
{
    const int i = lane & 15, g = lane >> 4;
#pragma unroll
    for (int r = 0; r < 4; r++) {
        float &o = out[(16 * mb + 4 * g + r) * ld + 16 * nb + i];
        o = add ? o + v[r] : v[r];
    }
}
__device__ __forceinline__ void put_bias(float *out, int mb, int lane, v4f v, bool add)
{
    const int i = lane & 15, g = lane >> 4;
    if (i == 0) {
#pragma unroll
        for (int r = 0; r < 4; r++) {
            float &o = out[16 * mb + 4 * g + r];
            o = add ? o + v[r] : v[r];
        }
    }
}

// grid: nblk blocks x 4 wavefronts; wavefront 4 blockIdx.x + w takes tiles
// 4 blockIdx.x + w, + 4 nblk, ... of the batch.  The block's four gradient
// sets are added in wavefront order in LDS and written as gradient row
// blockIdx.x of `part`, with the loss partial loss_part[blockIdx.x].
template <int ACT>
__global__ __launch_bounds__(256, 1) void k_mlp_step(const v4s *__restrict__ img, const float *__restrict__ norms,
                                                     const float *__restrict__ feats, const float *__restrict__ labels,
                                                     long long npool, long long base, long long pa, long long pb,
                                                     int batch, float *__restrict__ part, double *__restrict__ loss_part)
{
    __shared__ v4s tiles[T_COUNT * 64];
    __shared__ v4f biasv[NB_BIAS / 4];
    __shared__ double lred[4];
    __shared__ float grow[MLP_NPARAM];
    copy_consts(img, tiles, T_COUNT, biasv);
    __syncthreads();
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, i = lane & 15, g = lane >> 4;
    const int ntiles = batch / 16;
    const v4s id = ident_b(lane);
    const short one = (short)0x3f80;
    const v4s ones = {one, one, one, one};
    v4f g1[4], g1b[4], g2[4][4], g2b[4], g3[4][4], g3b[4], g4[4], g4b;
    const v4f zero4 = {0, 0, 0, 0};
#pragma unroll
    for (int mb = 0; mb < 4; mb++) {
        g1[mb] = g1b[mb] = g2b[mb] = g3b[mb] = g4[mb] = zero4;
#pragma unroll
        for (int nb = 0; nb < 4; nb++) g2[mb][nb] = g3[mb][nb] = zero4;
    }
    g4b = zero4;
    double lsum = 0.0;
    // a tile's raw features and labels (the sample of lane (i, g): features
    // 4g..4g+3, labels 4g..4g+3); the next tile's are loaded while the current
    // one is computed
    auto fetch = [&](int t, v4f &xr, v4f &yr) {
        const int p = t * 16 + i;                                  // position in the batch
        // sample index: pa, base, pb < npool < 2^31, so the product stays below 2^62
        const long long s = (pa * ((base + p) % npool) + pb) % npool;
        xr = g < 3 ? *reinterpret_cast<const v4f *>(feats + s * 16 + 4 * g) : v4f{0, 0, 0, 0};
        yr = 4 * g < MLP_OUT ? *reinterpret_cast<const v4f *>(labels + s * 16 + 4 * g) : v4f{0, 0, 0, 0};
    };
    const float inv_n = 2.0f / (float)(batch * MLP_OUT);
    v4f xr, yr;
    int t = blockIdx.x * 4 + w;
    if (t < ntiles) fetch(t, xr, yr);
    for (; t < ntiles; t += gridDim.x * 4) {
        v4f x = xr * *reinterpret_cast<const v4f *>(norms + 4 * g);   // zero for g = 3
        const v4f ycur = yr;
        if (t + gridDim.x * 4 < ntiles) fetch(t + gridDim.x * 4, xr, yr);
        const float r2 = point_r2(x), ir = inv_radius(r2), ir2 = ir * ir;
        x *= ir;
        const int zo = opaque_zero();
        const v4s *tl = tiles + zo;
        Acts A;
        mlp_forward<ACT>(reinterpret_cast<const float *>(biasv) + zo, LdsTiles{tl, lane}, lane, x, A);
        // MSE over the 9 normalised outputs: L = mean over batch*9 of (out - y/os)^2
        v4f dout;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int o = 4 * g + r;
            float d = 0.f;
            if (o < MLP_OUT) {
                // scale-free target; a column that is zero over the pool (the scalar
                // flux of a flow without temperature) has scale 0 and target 0
                const float os = norms[16 + o];
                const float y = os > 0.f ? ycur[r] / os * ir2 : 0.f;
                d = A.out[r] - y;
                lsum += (double)d * d;
            }
            dout[r] = d * inv_n;
        }
        v4s dz[3][4], dob;
        mlp_backward<ACT>(tl, lane, A, dout, dz, dob);

        // weight gradients of the tile, points as K
        const v4s ta0 = tr16(A.a0, id), tdo = tr16(dob, id);
        v4s ta[3][4], td[3][4];
#pragma unroll
        for (int l = 0; l < 3; l++)
#pragma unroll
            for (int mb = 0; mb < 4; mb++) {
                ta[l][mb] = tr16(A.a[l][mb], id);
                td[l][mb] = tr16(dz[l][mb], id);
            }
#pragma unroll
        for (int mb = 0; mb < 4; mb++) {
            g1[mb] = mfma(td[0][mb], ta0, g1[mb]);
            g1b[mb] = mfma(td[0][mb], ones, g1b[mb]);
            g2b[mb] = mfma(td[1][mb], ones, g2b[mb]);
            g3b[mb] = mfma(td[2][mb], ones, g3b[mb]);
            g4[mb] = mfma(tdo, ta[2][mb], g4[mb]);
#pragma unroll
            for (int nb = 0; nb < 4; nb++) {
                g2[mb][nb] = mfma(td[1][mb], ta[0][nb], g2[mb][nb]);
                g3[mb][nb] = mfma(td[2][mb], ta[1][nb], g3[mb][nb]);
            }
        }
        g4b = mfma(tdo, ones, g4b);
    }
    // the block's gradient row: every parameter is one (lane, r) of one tile
    for (int q = 0; q < 4; q++) {
        if (w == q) {
            const bool add = q > 0;
#pragma unroll
            for (int mb = 0; mb < 4; mb++) {
                put_tile(grow + MLP_OFF_W1, 16, mb, 0, lane, g1[mb], add);
                put_bias(grow + MLP_OFF_B1, mb, lane, g1b[mb], add);
                put_bias(grow + MLP_OFF_B2, mb, lane, g2b[mb], add);
                put_bias(grow + MLP_OFF_B3, mb, lane, g3b[mb], add);
                put_tile(grow + MLP_OFF_W4, 64, 0, mb, lane, g4[mb], add);
#pragma unroll
                for (int nb = 0; nb < 4; nb++) {
                    put_tile(grow + MLP_OFF_W2, 64, mb, nb, lane, g2[mb][nb], add);
                    put_tile(grow + MLP_OFF_W3, 64, mb, nb, lane, g3[mb][nb], add);
                }
            }
            put_bias(grow + MLP_OFF_B4, 0, lane, g4b, add);
        }
        __syncthreads();
    }
    float *out = part + (long long)blockIdx.x * MLP_NPARAM;
    for (int e = threadIdx.x; e < MLP_NPARAM; e += 256) out[e] = grow[e];
    for (int o = 32; o > 0; o >>= 1) lsum += __shfl_xor(lsum, o, 64);
    if (lane == 0) lred[w] = lsum;
    __syncthreads();
    if (threadIdx.x == 0) loss_part[blockIdx.x] = lred[0] + lred[1] + lred[2] + lred[3];
}

int mlp_step_blocks(int batch)
{
    const int nb = batch / 64;   // 4 tiles of 16 points per block at least
    return nb < 1 ? 1 : nb > 256 ? 256 : nb;
}

int launch_mlp_step(const void *img, const float *norms, const float *feats, const float *labels, long long npool,
                    long long base, long long pa, long long pb, int batch, float *part, double *loss_part, int act,
                    hipStream_t st)
{
    const int nblk = mlp_step_blocks(batch);
    if (act == SAM_MLP_ACT_GELU)
        hipLaunchKernelGGL(k_mlp_step<SAM_MLP_ACT_GELU>, dim3(nblk), dim3(256), 0, st, (const v4s *)img, norms, feats,
                           labels, npool, base, pa, pb, batch, part, loss_part);
    else
        hipLaunchKernelGGL(k_mlp_step<SAM_MLP_ACT_RELU>, dim3(nblk), dim3(256), 0, st, (const v4s *)img, norms, feats,
                           labels, npool, base, pa, pb, batch, part, loss_part);
    return launched_train();
}

} // namespace samhip
