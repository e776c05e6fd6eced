// mlp_core.hpp -- device building blocks of the learned closure, shared by
// the inference/LES kernels (mlp.hip) and the fused training kernel
// (mlp_train.hip): BF16 packing, the MFMA wrappers, the activations, the
// packed weight-tile image and the per-tile forward and backward passes.
// The MFMA register layouts are described at the top of mlp.hip.
#pragma once
#include "internal.hpp"

namespace samhip {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef short v8s __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));

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

typedef __bf16 v2bf __attribute__((ext_vector_type(2)));
__device__ __forceinline__ unsigned pk2(float a, float b) { return __builtin_bit_cast(unsigned, v2bf{(__bf16)a, (__bf16)b}); }
__device__ __forceinline__ v4s bf4(v4f x)
{
    typedef unsigned v2u __attribute__((ext_vector_type(2)));
    return __builtin_bit_cast(v4s, v2u{pk2(x[0], x[1]), pk2(x[2], x[3])});
}

__device__ __forceinline__ v4f mfma(v4s a, v4s b, v4f c) { return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0); }
// Two K blocks of 16 in one gfx950 v_mfma_f32_16x16x32_bf16 (twice the work
// per instruction of the 16x16x16 form).  Lane (i, g) supplies K slots
// 8g..8g+7: the 4 values it holds for block 0 and then the 4 it holds for
// block 1.  The K order inside one MFMA is free as long as A and B agree, so
// the operands are the 16x16x16 lane tiles of the two blocks side by side.
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
__device__ __forceinline__ v4f mfma2(v4s a0, v4s a1, v4s b0, v4s b1, v4f c)
{
    const v8s a = __builtin_shufflevector(a0, a1, 0, 1, 2, 3, 4, 5, 6, 7);
    const v8s b = __builtin_shufflevector(b0, b1, 0, 1, 2, 3, 4, 5, 6, 7);
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(v8bf, a), __builtin_bit_cast(v8bf, b), c, 0, 0, 0);
}

// |x|^2 of a point's scaled features, held 4 per lane in lanes i, i+16, i+32
// (lane i+48 holds zeros); every lane of the point gets the sum
// (gfx950 row swaps: after v_permlane16_swap of two copies of r, the pair
// holds r and its xor-16 partner in every lane, in some order; their sum is
// the same either way, as after v_permlane32_swap for xor 32)
__device__ __forceinline__ float point_r2(v4f x)
{
    float r = x[0] * x[0] + x[1] * x[1] + x[2] * x[2] + x[3] * x[3];
    const auto s16 = __builtin_amdgcn_permlane16_swap(__float_as_uint(r), __float_as_uint(r), false, false);
    r = __uint_as_float(s16[0]) + __uint_as_float(s16[1]);
    const auto s32 = __builtin_amdgcn_permlane32_swap(__float_as_uint(r), __float_as_uint(r), false, false);
    return __uint_as_float(s32[0]) + __uint_as_float(s32[1]);
}
__device__ __forceinline__ float inv_radius(float r2) { return r2 > 0.f ? __builtin_amdgcn_rsqf(r2) : 0.f; }

// Hidden activation, ACT = SAM_MLP_ACT_RELU or SAM_MLP_ACT_GELU.
//
// ReLU: a signed max on the bits (a negative float, sign bit set, is a
// negative integer; fmaxf would also quiet NaNs first), one instruction per
// FP32 value; where the result is rounded to BF16 anyway (act_bf4) it is
// taken after the rounding, two values per instruction.  Derivative 1 for
// x > 0, else 0.
//
// GELU, tanh form: 0.5 x (1 + tanh u) with u = c (x + a x^3), evaluated as
// x * sigmoid(2u) = x / (1 + 2^(-2u log2 e)): one exp2 and one reciprocal,
// quarter-rate instructions.  With 48 activations per lane per 16-point tile
// they take about 40% of the inference kernel's time.  e -> inf for large
// negative x gives exactly -0; e -> 0 gives x.
constexpr float GELU_C = 0.7978845608028654f, GELU_A = 0.044715f;
constexpr float GELU_K1 = -2.0f * GELU_C * 1.4426950408889634f, GELU_K2 = GELU_K1 * GELU_A;
__device__ __forceinline__ float gelu_sig(float x)
{
    const float e = __builtin_amdgcn_exp2f(x * __builtin_fmaf(GELU_K2, x * x, GELU_K1));
    return __builtin_amdgcn_rcpf(1.0f + e);
}
template <int ACT> __device__ __forceinline__ float act(float x)
{
    if constexpr (ACT == SAM_MLP_ACT_GELU) return x * gelu_sig(x);
    else return __int_as_float(max(__float_as_int(x), 0));
}
// GELU derivative: s + g (1 - s) 2c (1 + 3 a x^2), s = sigmoid(2u), g = x s
template <int ACT> __device__ __forceinline__ float act_deriv(float x)
{
    if constexpr (ACT == SAM_MLP_ACT_GELU) {
        const float sg = gelu_sig(x), g = x * sg;
        return __builtin_fmaf(g * (1.0f - sg), __builtin_fmaf(6.0f * GELU_C * GELU_A, x * x, 2.0f * GELU_C), sg);
    } else {
        return x > 0.0f ? 1.0f : 0.0f;
    }
}

// ---- weight tiles in LDS ---------------------------------------------
// tile t (64 lanes x 4 bf16): forward tiles  0..3  W1 (mb),  4..19 W2 (mb*4+kb),
// 20..35 W3, 36..39 W4 (kb); transposed tiles 40..43 W4^T (mb), 44..59 W3^T,
// 60..75 W2^T.
constexpr int T_W1 = 0, T_W2 = 4, T_W3 = 20, T_W4 = 36, T_W4T = 40, T_W3T = 44, T_W2T = 60, T_COUNT = 76;

__device__ __forceinline__ float wparam(const float *P, int off, int ld, int row, int col) { return P[off + row * ld + col]; }

// biases in the packed image, after the tiles (FP32): b1, b2, b3, b4
constexpr int NB_BIAS = 208, BO_1 = 0, BO_2 = 64, BO_3 = 128, BO_4 = 192;

// copy the first `ntiles` tiles of the packed image and the biases into LDS
// (16 B per thread)
__device__ __forceinline__ void copy_consts(const v4s *__restrict__ img, v4s *tiles, int ntiles, v4f *bias)
{
    const uint4 *src = reinterpret_cast<const uint4 *>(img);
    uint4 *dst = reinterpret_cast<uint4 *>(tiles);
    for (int e = threadIdx.x; e < ntiles * 32; e += blockDim.x) dst[e] = src[e];
    const uint4 *bs = reinterpret_cast<const uint4 *>(img + T_COUNT * 64);
    uint4 *bd = reinterpret_cast<uint4 *>(bias);
    for (int e = threadIdx.x; e < NB_BIAS / 4; e += blockDim.x) bd[e] = bs[e];
}

// biases u0..u0+3 of the bias block at `off` (LDS)
__device__ __forceinline__ v4f bias4(const float *bias, int off, int u0)
{
    return *reinterpret_cast<const v4f *>(bias + off + u0);
}

struct Acts {
    v4s a0;          // inputs (bf16)
    v4f z[3][4];     // pre-activations of the hidden layers
    v4s a[3][4];     // hidden activations (bf16)
    v4f out;         // outputs 4g..4g+3
};

// ReLU after the BF16 rounding, on the packed 16-bit values: max(b, 0) as
// signed integers zeroes exactly the negative (sign-bit) values, two per
// instruction; rounding is monotonic, so this equals rounding max(z, 0)
template <int ACT> __device__ __forceinline__ v4s act_bf4(v4f z)
{
    if constexpr (ACT == SAM_MLP_ACT_GELU) {
        return bf4(v4f{act<ACT>(z[0]), act<ACT>(z[1]), act<ACT>(z[2]), act<ACT>(z[3])});
    } else {
        typedef unsigned v2u __attribute__((ext_vector_type(2)));
        auto relu2 = [](unsigned w) {
            unsigned r;
            asm("v_pk_max_i16 %0, %1, 0" : "=v"(r) : "v"(w));
            return r;
        };
        return __builtin_bit_cast(v4s, v2u{relu2(pk2(z[0], z[1])), relu2(pk2(z[2], z[3]))});
    }
}

// Weight-tile sources of the forward pass: tile k of this lane from the LDS
// image, or from registers loaded once per wavefront (forward tiles only).
struct LdsTiles {
    const v4s *t;
    int lane;
    __device__ __forceinline__ v4s operator()(int k) const { return t[k * 64 + lane]; }
};
struct RegTiles {
    v4s w[T_W4T];
    __device__ __forceinline__ v4s operator()(int k) const { return w[k]; }
};

// one hidden layer 64 -> 64: z = W a + b, tiles tb..tb+15 (mb*4+kb); the
// four output blocks' MFMA chains are interleaved (kb outer)
template <int ACT, class TS>
__device__ __forceinline__ void fw_hidden(const float *bias, const TS &tiles, int tb, int boff, int lane,
                                          const v4s *ain, v4f *z, v4s *aout)
{
    const int g = lane >> 4;
    v4f acc[4];
#pragma unroll
    for (int mb = 0; mb < 4; mb++) acc[mb] = bias4(bias, boff, 16 * mb + 4 * g);
#pragma unroll
    for (int kb = 0; kb < 4; kb += 2)
#pragma unroll
        for (int mb = 0; mb < 4; mb++)
            acc[mb] = mfma2(tiles(tb + mb * 4 + kb), tiles(tb + mb * 4 + kb + 1), ain[kb], ain[kb + 1], acc[mb]);
#pragma unroll
    for (int mb = 0; mb < 4; mb++) {
        z[mb] = acc[mb];
        aout[mb] = act_bf4<ACT>(acc[mb]);
    }
}

// forward pass of one 16-point tile; x: this lane's 4 normalised inputs;
// bias in LDS, weight tiles from `tiles`
template <int ACT, class TS>
__device__ __forceinline__ void mlp_forward(const float *bias, const TS &tiles, int lane, v4f x, Acts &A)
{
    const int g = lane >> 4;
    A.a0 = bf4(x);
    v4f z[4];
#pragma unroll
    for (int mb = 0; mb < 4; mb++) z[mb] = mfma(tiles(T_W1 + mb), A.a0, bias4(bias, BO_1, 16 * mb + 4 * g));
#pragma unroll
    for (int mb = 0; mb < 4; mb++) {
        A.z[0][mb] = z[mb];
        A.a[0][mb] = act_bf4<ACT>(z[mb]);
    }
    fw_hidden<ACT, TS>(bias, tiles, T_W2, BO_2, lane, A.a[0], A.z[1], A.a[1]);
    fw_hidden<ACT, TS>(bias, tiles, T_W3, BO_3, lane, A.a[1], A.z[2], A.a[2]);
    v4f o = bias4(bias, BO_4, 4 * g);
#pragma unroll
    for (int kb = 0; kb < 4; kb += 2)
        o = mfma2(tiles(T_W4 + kb), tiles(T_W4 + kb + 1), A.a[2][kb], A.a[2][kb + 1], o);
    A.out = o;
}

template <int ACT> __device__ __forceinline__ v4f act_grad4(v4f z)
{
    return v4f{act_deriv<ACT>(z[0]), act_deriv<ACT>(z[1]), act_deriv<ACT>(z[2]), act_deriv<ACT>(z[3])};
}

template <int ACT> __device__ __forceinline__ void bw_act(const v4f *z, const v4f *da, v4s *dz)
{
#pragma unroll
    for (int mb = 0; mb < 4; mb++) {
        const v4f gd = act_grad4<ACT>(z[mb]);
        dz[mb] = bf4(v4f{da[mb][0] * gd[0], da[mb][1] * gd[1], da[mb][2] * gd[2], da[mb][3] * gd[3]});
    }
}

__device__ __forceinline__ void bw_matvec(const v4s *tiles, int tb, int lane, const v4s *dz, v4f *da)
{
#pragma unroll
    for (int mb = 0; mb < 4; mb++) {
        v4f s = {0, 0, 0, 0};
#pragma unroll
        for (int kb = 0; kb < 4; kb += 2)
            s = mfma2(tiles[(tb + mb * 4 + kb) * 64 + lane], tiles[(tb + mb * 4 + kb + 1) * 64 + lane], dz[kb], dz[kb + 1], s);
        da[mb] = s;
    }
}

// backward of one tile from dL/d(out); returns dz of each hidden layer (bf16)
template <int ACT>
__device__ __forceinline__ void mlp_backward(const v4s *tiles, int lane, const Acts &A, v4f dout, v4s dz[3][4],
                                             v4s &dout_b)
{
    dout_b = bf4(dout);
    v4f da[4];
#pragma unroll
    for (int mb = 0; mb < 4; mb++) da[mb] = mfma(tiles[(T_W4T + mb) * 64 + lane], dout_b, v4f{0, 0, 0, 0});
    bw_act<ACT>(A.z[2], da, dz[2]);
    bw_matvec(tiles, T_W3T, lane, dz[2], da);
    bw_act<ACT>(A.z[1], da, dz[1]);
    bw_matvec(tiles, T_W2T, lane, dz[1], da);
    bw_act<ACT>(A.z[0], da, dz[0]);
}

} // namespace samhip
