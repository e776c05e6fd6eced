// fft_tile.hpp -- batched power-of-two complex FFTs for gfx950, one tile of
// C columns per workgroup, Stockham autosort with radix-8/4/2 passes.
//
// Each thread owns EPT (8 or 16) elements of one column per pass and keeps
// them in VGPRs; passes exchange data through an LDS tile of C*N complex
// elements.  The LDS address of element p of column c is XOR-swizzled within
// 128-byte groups (lds_idx) so that the strided Stockham reads/writes and the
// column-interleaved cooperative loads avoid bank conflicts without padding
// (the tile stays exactly C*N elements).
//
// Sign convention (FFTPACK, fft.f:57-59,171-173): forward = exp(-i..),
// backward = exp(+i..), both unnormalised.  Twiddles come from a table
// tw[k] = exp(-2 pi i k/N) built on the host in long double.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace samhip {

// aligned to its full size so LDS and global accesses use single 8- or 16-byte ops
template <typename T> struct alignas(2 * sizeof(T)) cpx { T x, y; };

template <typename T> __device__ __forceinline__ cpx<T> operator+(cpx<T> a, cpx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T> __device__ __forceinline__ cpx<T> operator-(cpx<T> a, cpx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T> __device__ __forceinline__ cpx<T> cmul(cpx<T> a, cpx<T> b)
{
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// Streaming HBM accesses (non-temporal cache policy, the `nt` bit).  The step
// kernels read and write every HBM line once; with nt on both the loads and
// the stores a plain 2 GiB copy on gfx950 runs 5-7% faster than with the
// default policy (5.81 vs 5.55 TB/s).
template <typename T> __device__ __forceinline__ cpx<T> ld_nt(const cpx<T> *p)
{
    typedef T v2 __attribute__((ext_vector_type(2)));
    const v2 v = __builtin_nontemporal_load(reinterpret_cast<const v2 *>(p));
    return {v.x, v.y};
}
template <typename T> __device__ __forceinline__ void st_nt(cpx<T> *p, cpx<T> a)
{
    typedef T v2 __attribute__((ext_vector_type(2)));
    __builtin_nontemporal_store(v2{a.x, a.y}, reinterpret_cast<v2 *>(p));
}
template <typename T> __device__ __forceinline__ cpx<T> conjg(cpx<T> a) { return {a.x, -a.y}; }
template <typename T> __device__ __forceinline__ cpx<T> scal(cpx<T> a, T s) { return {a.x * s, a.y * s}; }
// a * (S i), S = +1 or -1
template <int S, typename T> __device__ __forceinline__ cpx<T> mul_si(cpx<T> a)
{
    if constexpr (S > 0) return {-a.y, a.x};
    else return {a.y, -a.x};
}

__host__ __device__ constexpr int ilog2(int n) { return n <= 1 ? 0 : 1 + ilog2(n >> 1); }

// Radix schedule, symmetric where possible so that the last pass of one
// transform leaves data exactly where the first pass of the next reads it.
template <int N> struct Sched {
    static constexpr int L = ilog2(N);
    static_assert((1 << L) == N && N >= 16 && N <= 4096, "power of two 16..4096");
    static constexpr int P = (L == 4) ? 2 : (L == 5) ? 2 : (L == 6) ? 2 : (L == 7) ? 3 : (L == 8) ? 3 : (L == 9) ? 3
                           : (L == 10) ? 4 : (L == 11) ? 4 : 4;
    __host__ __device__ static constexpr int radix(int p)
    {
        return (L == 4) ? 4
             : (L == 5) ? (p == 0 ? 8 : 4)
             : (L == 6) ? 8
             : (L == 7) ? (p == 1 ? 2 : 8)
             : (L == 8) ? (p == 1 ? 4 : 8)
             : (L == 9) ? 8
             : (L == 10) ? ((p == 1 || p == 2) ? 4 : 8)
             : (L == 11) ? (p == 1 ? 4 : 8)
             : 8; /* L == 12 */
    }
    __host__ __device__ static constexpr int ns(int p) { return p == 0 ? 1 : ns(p - 1) * radix(p - 1); }
};

// rotate the low B bits of c left by one
template <int B> __device__ __forceinline__ int rotl_bits(int c)
{
    c &= (1 << B) - 1;
    return ((c << 1) & ((1 << B) - 1)) | (c >> (B - 1));
}

// LDS address of element p of column c in a tile with column length N.
// Element p stays in its aligned group of SW elements (SW x element size =
// 128 bytes, all 32 banks); its slot in the group is p XOR a key of the
// higher bits of p and of the column.  The keys were chosen, per element
// size and N, so that the Stockham exchanges of every pass, natural-order
// and transposed tile accesses and the x-row two-row pairs (columns c and
// k, k + 1 in one access group) hit distinct banks (LDS model: 32 banks of
// 4 bytes; an access of W bytes per lane serves 128/W lanes per cycle).
// The column term is rotated so that a group of columns at two adjacent k
// (the x-row pairs) still covers every bank.
template <typename T, int N>
__device__ __forceinline__ int lds_idx(int c, int p)
{
    if constexpr (sizeof(T) == 8) {   // 16-byte elements, 8 per group
        return c * N + (p ^ (((p >> 3) ^ rotl_bits<3>(c)) & 7));
    } else {                          // 8-byte elements, 16 per group
        int key;
        if constexpr (N == 128) key = (p >> 3) ^ ((p >> 5) << 1) ^ rotl_bits<4>(rotl_bits<4>(rotl_bits<4>(c)));
        else if constexpr (N == 256) key = (p >> 2) ^ (p >> 5) ^ rotl_bits<4>(c);
        else if constexpr (N == 512) key = (p >> 3) ^ rotl_bits<4>(c);
        else if constexpr (N == 1024) key = (p >> 4) ^ (p >> 5) ^ rotl_bits<4>(c);
        else key = (p >> 4) ^ c;
        return c * N + (p ^ (key & 15));
    }
}

template <int S, typename T> __device__ __forceinline__ void dft2(cpx<T> &a, cpx<T> &b)
{
    cpx<T> t = a - b;
    a = a + b;
    b = t;
}

template <int S, typename T> __device__ __forceinline__ void dft4(cpx<T> &a, cpx<T> &b, cpx<T> &c, cpx<T> &d)
{
    cpx<T> t0 = a + c, t1 = a - c, t2 = b + d, t3 = mul_si<S>(b - d);
    a = t0 + t2;
    c = t0 - t2;
    b = t1 + t3;
    d = t1 - t3;
}

template <int S, typename T> __device__ __forceinline__ void dft8(cpx<T> *v)
{
    // even/odd split: E = DFT4(v0,v2,v4,v6), O = DFT4(v1,v3,v5,v7)
    cpx<T> e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
    cpx<T> o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
    dft4<S>(e0, e1, e2, e3);
    dft4<S>(o0, o1, o2, o3);
    const T h = (T)0.70710678118654752440084436210484903928483593768847;
    // w^1 = h(1 + S i), w^2 = S i, w^3 = h(-1 + S i); the h scalings are
    // folded into the output sums: e +- h a is one FMA per component
    const cpx<T> a1 = {o1.x - (T)S * o1.y, o1.y + (T)S * o1.x};    // w^1 o1 / h
    const cpx<T> a3 = {o3.x + (T)S * o3.y, o3.y - (T)S * o3.x};    // -w^3 o3 / h
    cpx<T> w2o2 = mul_si<S>(o2);
    v[0] = e0 + o0;
    v[4] = e0 - o0;
    v[1] = {fma(h, a1.x, e1.x), fma(h, a1.y, e1.y)};
    v[5] = {fma(-h, a1.x, e1.x), fma(-h, a1.y, e1.y)};
    v[2] = e2 + w2o2;
    v[6] = e2 - w2o2;
    v[3] = {fma(-h, a3.x, e3.x), fma(-h, a3.y, e3.y)};
    v[7] = {fma(h, a3.x, e3.x), fma(h, a3.y, e3.y)};
}

// LDS ordering between passes.  When every column of the tile belongs to a
// single wavefront (threads per column divide 64), the exchange only has to
// order this wave's own LDS accesses.  The LDS performs one wavefront's
// accesses in issue order, so a read issued after a write sees it without a
// wait (the AMDGPU memory model needs no s_waitcnt for wavefront-scope LDS
// ordering); only the compiler must be kept from moving LDS accesses across,
// and the reads are queued right behind the writes.  Otherwise a workgroup
// barrier for LDS data only (not __syncthreads, which also waits for every
// outstanding global access and so would drain prefetched loads and pending
// stores).
template <bool WAVE> __device__ __forceinline__ void tile_sync()
{
    if constexpr (WAVE) {
        asm volatile("" ::: "memory");
    } else {
        asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
    }
}

// twiddles w^r, r = 0..R-1, of w = exp(S 2 pi i e0 / N) from the table
// tw[k] = exp(-2 pi i k/N): w^1, w^2, w^4 are read (tw_read), the rest
// multiplied (tw_expand)
template <int R, int S, typename T>
__device__ __forceinline__ void tw_read(const cpx<T> *__restrict__ tw, int e0, cpx<T> *w)
{
    auto ld = [&](int e) {
        cpx<T> t = tw[e];
        if constexpr (S > 0) t.y = -t.y;
        return t;
    };
    w[1] = ld(e0);
    if constexpr (R >= 4) w[2] = ld(2 * e0);
    if constexpr (R == 8) w[4] = ld(4 * e0);
}
template <int R, typename T> __device__ __forceinline__ void tw_expand(cpx<T> *w)
{
    if constexpr (R >= 4) w[3] = cmul(w[1], w[2]);
    if constexpr (R == 8) {
        w[5] = cmul(w[1], w[4]);
        w[6] = cmul(w[2], w[4]);
        w[7] = cmul(w[3], w[4]);
    }
}
template <int R, int S, typename T>
__device__ __forceinline__ void twiddles(const cpx<T> *__restrict__ tw, int e0, cpx<T> *w)
{
    tw_read<R, S>(tw, e0, w);
    tw_expand<R>(w);
}

// Butterfly of radix R on v[0..R-1] (inputs in stride order r)
template <int R, int S, typename T> __device__ __forceinline__ void dftR(cpx<T> *v)
{
    if constexpr (R == 8) dft8<S>(v);
    else if constexpr (R == 4) dft4<S>(v[0], v[1], v[2], v[3]);
    else dft2<S>(v[0], v[1]);
}

// Tile FFT engine.  N: length, EPT: elements per thread (8 or 16),
// threads per column TPC = N/EPT.  Thread (c, j0) owns butterflies
// b_m = j0 + m*TPC, m < EPT/R, in each pass.  Register slot s = m*R + r
// holds element b_m + r*N/R at the start of a pass.
template <typename T, int N, int EPT> struct Tile {
    static_assert(EPT == 8 || EPT == 16, "EPT");
    static constexpr int TPC = N / EPT;
    static constexpr bool WAVE = TPC <= 64 && (64 % TPC) == 0;   // column inside one wavefront
    using SC = Sched<N>;

    // k index (element j0 + k*TPC) held in slot s at the start of pass p
    __host__ __device__ static constexpr int kin(int p, int s)
    {
        return (s / SC::radix(p)) + (s % SC::radix(p)) * (EPT / SC::radix(p));
    }
    // position of slot s after the last pass: j0 + kout(s)*TPC
    __host__ __device__ static constexpr int kout(int s) { return kin(SC::P - 1, s); }

    // load the pass-p inputs from LDS
    template <int p> __device__ static __forceinline__ void lds_read(const cpx<T> *lds, int c, int j0, cpx<T> *v)
    {
        constexpr int R = SC::radix(p);
#pragma unroll
        for (int s = 0; s < EPT; s++) {
            const int m = s / R, r = s % R;
            const int b = j0 + m * TPC;
            v[s] = lds[lds_idx<T, N>(c, b + r * (N / R))];
        }
    }

    // per-thread twiddles of one pass: butterfly m, power r
    struct Tw {
        cpx<T> w[8][8];
    };

    // read the pass-p twiddle bases (issued ahead of the pass so the table
    // reads overlap the preceding LDS exchange)
    template <int p, int S> __device__ static __forceinline__ void tw_fetch(const cpx<T> *__restrict__ tw, int j0, Tw &t)
    {
        constexpr int R = SC::radix(p);
        constexpr int Ns = SC::ns(p);
        if constexpr (Ns > 1) {
#pragma unroll
            for (int m = 0; m < EPT / R; m++) {
                const int jl = (j0 + m * TPC) & (Ns - 1);
                tw_read<R, S>(tw, jl * (N / (Ns * R)), t.w[m]);
            }
        }
    }

    // Twiddle table in LDS, laid out without bank conflicts.  The plain
    // table tw[k] = exp(-2 pi i k/N) read at e0, 2 e0, 4 e0 puts the lanes of
    // an access group on one bank in the early passes (e0 = jl N/(Ns R) is a
    // multiple of a power of two).  Here pass p >= 1 stores the bases w^1,
    // w^2 (R >= 4) and w^4 (R = 8) of its Ns twiddle rows as runs of Ns
    // consecutive entries, so consecutive jl read consecutive entries.  At
    // most N/2 entries.
    __host__ __device__ static constexpr int twl_nb(int p) { return SC::radix(p) == 8 ? 3 : SC::radix(p) == 4 ? 2 : 1; }
    __host__ __device__ static constexpr int twl_off(int p)
    {
        return p <= 1 ? 0 : twl_off(p - 1) + twl_nb(p - 1) * SC::ns(p - 1);
    }
    static constexpr int TWL_SIZE = twl_off(SC::P);
    static_assert(TWL_SIZE <= N / 2, "LDS twiddle table fits N/2 entries");
    struct TwLds {
        const cpx<T> *t;
    };
    // fill the LDS table from tw[k] (k < N/2); the caller barriers before use
    template <int p = 1>
    __device__ static __forceinline__ void twl_fill(const cpx<T> *__restrict__ twg, cpx<T> *twl, int tid, int nthreads)
    {
        if constexpr (p < SC::P) {
            constexpr int Ns = SC::ns(p), R = SC::radix(p), off = twl_off(p);
            for (int e = tid; e < twl_nb(p) * Ns; e += nthreads)
                twl[off + e] = twg[(1 << (e / Ns)) * (e % Ns) * (N / (Ns * R))];
            twl_fill<p + 1>(twg, twl, tid, nthreads);
        }
    }
    template <int p, int S> __device__ static __forceinline__ void tw_fetch(TwLds tl, int j0, Tw &t)
    {
        constexpr int R = SC::radix(p);
        constexpr int Ns = SC::ns(p);
        if constexpr (Ns > 1) {
            constexpr int off = twl_off(p);
            auto ld = [&](int e) {
                cpx<T> w = tl.t[e];
                if constexpr (S > 0) w.y = -w.y;
                return w;
            };
#pragma unroll
            for (int m = 0; m < EPT / R; m++) {
                const int jl = (j0 + m * TPC) & (Ns - 1);
                t.w[m][1] = ld(off + jl);
                if constexpr (R >= 4) t.w[m][2] = ld(off + Ns + jl);
                if constexpr (R == 8) t.w[m][4] = ld(off + 2 * Ns + jl);
            }
        }
    }

    // twiddle + butterflies of pass p
    template <int p, int S> __device__ static __forceinline__ void compute(Tw &t, cpx<T> *v)
    {
        constexpr int R = SC::radix(p);
        constexpr int Ns = SC::ns(p);
#pragma unroll
        for (int m = 0; m < EPT / R; m++) {
            if constexpr (Ns > 1) {
                tw_expand<R>(t.w[m]);
#pragma unroll
                for (int r = 1; r < R; r++) v[m * R + r] = cmul(v[m * R + r], t.w[m][r]);
            }
            dftR<R, S>(v + m * R);
        }
    }

    // store the pass-p outputs to LDS
    template <int p> __device__ static __forceinline__ void lds_write(cpx<T> *lds, int c, int j0, const cpx<T> *v)
    {
        constexpr int R = SC::radix(p);
        constexpr int Ns = SC::ns(p);
#pragma unroll
        for (int m = 0; m < EPT / R; m++) {
            const int b = j0 + m * TPC;
            const int idxD = (b / Ns) * Ns * R + (b & (Ns - 1));
#pragma unroll
            for (int r = 0; r < R; r++) lds[lds_idx<T, N>(c, idxD + r * Ns)] = v[m * R + r];
        }
    }

    // passes p..P-1 with the pass-p twiddles already read.  TWPF: read the
    // next pass's twiddles before this pass's LDS exchange (for a twiddle
    // table in global memory, whose latency the exchange then covers); with
    // an LDS table they are read at the pass itself, holding fewer VGPRs.
    template <int S, int p, bool TWPF, typename TWS>
    __device__ static __forceinline__ void run_passes(cpx<T> *lds, TWS tw, int c, int j0, cpx<T> *v, Tw &cur)
    {
        if constexpr (p + 1 < SC::P) {
            Tw nxt;
            if constexpr (TWPF) tw_fetch<p + 1, S>(tw, j0, nxt);
            compute<p, S>(cur, v);
            lds_write<p>(lds, c, j0, v);
            tile_sync<WAVE>();
            lds_read<p + 1>(lds, c, j0, v);
            tile_sync<WAVE>();
            if constexpr (!TWPF) tw_fetch<p + 1, S>(tw, j0, nxt);
            run_passes<S, p + 1, TWPF>(lds, tw, c, j0, v, nxt);
        } else {
            compute<p, S>(cur, v);
        }
    }

    // passes p0..P-1 starting from registers already holding pass-p0
    // inputs; the final pass leaves its outputs in v (positions kout).
    // Requires a __syncthreads() before the call if other threads may still
    // access the LDS tile region this thread's column uses.
    template <int S, int p = 0, bool TWPF = true, typename TWS = const cpx<T> *>
    __device__ static __forceinline__ void run_regs(cpx<T> *lds, TWS tw, int c, int j0, cpx<T> *v)
    {
        Tw cur;
        tw_fetch<p, S>(tw, j0, cur);
        run_passes<S, p, TWPF>(lds, tw, c, j0, v, cur);
    }

    // Twiddles of every pass p >= 1 held in registers across transforms.  A
    // thread's twiddles depend on j0 only, so a kernel that runs many
    // transforms per thread reads and expands them once; the inverse
    // direction multiplies by their conjugates.  Slot s = m*R + r of pass p
    // holds the forward twiddle of butterfly m, power r (r >= 1).
    struct TwRegs {
        cpx<T> w[SC::P][EPT];
    };
    __device__ static __forceinline__ void tw_regs(const cpx<T> *__restrict__ tw, int j0, TwRegs &t)
    {
        tw_regs_pass<1>(tw, j0, t);
    }
    template <int p>
    __device__ static __forceinline__ void tw_regs_pass(const cpx<T> *__restrict__ tw, int j0, TwRegs &t)
    {
        if constexpr (p < SC::P) {
            constexpr int R = SC::radix(p);
            constexpr int Ns = SC::ns(p);
#pragma unroll
            for (int m = 0; m < EPT / R; m++) {
                const int jl = (j0 + m * TPC) & (Ns - 1);
                cpx<T> w[8];
                twiddles<R, -1>(tw, jl * (N / (Ns * R)), w);
#pragma unroll
                for (int r = 1; r < R; r++) t.w[p][m * R + r] = w[r];
            }
            tw_regs_pass<p + 1>(tw, j0, t);
        }
    }
    // a * w (S < 0) or a * conj(w) (S > 0)
    template <int S> __device__ static __forceinline__ cpx<T> twmul(cpx<T> a, cpx<T> w)
    {
        if constexpr (S < 0) return cmul(a, w);
        else return {a.x * w.x + a.y * w.y, a.y * w.x - a.x * w.y};
    }
    template <int p, int S> __device__ static __forceinline__ void compute_r(const TwRegs &t, cpx<T> *v)
    {
        constexpr int R = SC::radix(p);
        constexpr int Ns = SC::ns(p);
#pragma unroll
        for (int m = 0; m < EPT / R; m++) {
            if constexpr (Ns > 1) {
#pragma unroll
                for (int r = 1; r < R; r++) v[m * R + r] = twmul<S>(v[m * R + r], t.w[p][m * R + r]);
            }
            dftR<R, S>(v + m * R);
        }
    }
    // run_regs with the register twiddles
    template <int S, int p = 0>
    __device__ static __forceinline__ void run_regs(cpx<T> *lds, const TwRegs &t, int c, int j0, cpx<T> *v)
    {
        compute_r<p, S>(t, v);
        if constexpr (p + 1 < SC::P) {
            lds_write<p>(lds, c, j0, v);
            tile_sync<WAVE>();
            lds_read<p + 1>(lds, c, j0, v);
            tile_sync<WAVE>();
            run_regs<S, p + 1>(lds, t, c, j0, v);
        }
    }

    // LDS -> registers (pass-0 inputs) -> FFT -> registers (final outputs).
    // Touches only column c of the tile; when WAVE, other waves may still be
    // working on their columns at return, so a caller that next writes other
    // columns must __syncthreads() first.
    template <int S, typename TWS = const cpx<T> *>
    __device__ static __forceinline__ void run_from_lds(cpx<T> *lds, TWS tw, int c, int j0, cpx<T> *v)
    {
        lds_read<0>(lds, c, j0, v);
        tile_sync<WAVE>();
        run_regs<S, 0>(lds, tw, c, j0, v);
    }

    // write the final outputs (from run_*) to LDS in natural order
    __device__ static __forceinline__ void store_natural(cpx<T> *lds, int c, int j0, const cpx<T> *v)
    {
#pragma unroll
        for (int s = 0; s < EPT; s++) lds[lds_idx<T, N>(c, j0 + kout(s) * TPC)] = v[s];
    }

    // reorder final outputs (slots by kout) into pass-0 input slots (kin(0,.))
    __device__ static __forceinline__ void out_to_in(const cpx<T> *vo, cpx<T> *vi)
    {
#pragma unroll
        for (int s = 0; s < EPT; s++) {
#pragma unroll
            for (int t = 0; t < EPT; t++)
                if (kout(t) == kin(0, s)) vi[s] = vo[t];
        }
    }
};

} // namespace samhip
