// Building blocks of the gfx950 3-D FFT: complex arithmetic, register-resident
// small DFTs (radix 2/3/4/8/16) and Stockham autosort stages that exchange data
// through LDS.  The 3-D real transforms used by the solver are assembled from
// these in passes.hpp.
//
// Conventions: DIR = -1 is the forward transform (e^{-i..}), DIR = +1 the
// inverse; both unnormalised, as p3dfft_ftran_r2c / p3dfft_btran_c2r are
// (called at rhs.f90:31-65).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dnsbox {

template <class T>
struct alignas(2 * sizeof(T)) cpx {
    T x, y;
};

template <class T>
__device__ __forceinline__ cpx<T> mk(T a, T b) { return cpx<T>{a, b}; }
template <class T>
__device__ __forceinline__ cpx<T> operator+(cpx<T> a, cpx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <class T>
__device__ __forceinline__ cpx<T> operator-(cpx<T> a, cpx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <class T>
__device__ __forceinline__ cpx<T> operator*(cpx<T> a, cpx<T> b)
{
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <class T>
__device__ __forceinline__ cpx<T> operator*(T s, cpx<T> a) { return {s * a.x, s * a.y}; }
template <class T>
__device__ __forceinline__ cpx<T> conj(cpx<T> a) { return {a.x, -a.y}; }
template <class TO, class TI>
__device__ __forceinline__ cpx<TO> cvt(cpx<TI> a)
{
    return cpx<TO>{TO(a.x), TO(a.y)};
}
// multiply by s*i, s = +-1
template <int S, class T>
__device__ __forceinline__ cpx<T> mul_i(cpx<T> a)
{
    return S > 0 ? cpx<T>{-a.y, a.x} : cpx<T>{a.y, -a.x};
}

// ---------------------------------------------------------------------------
// small DFTs, in place on a[0..R-1]:  a[k] <- sum_r a[r] e^{DIR 2 pi i r k / R}
// ---------------------------------------------------------------------------
template <int R, int DIR, class T>
struct Dft;

template <int DIR, class T>
struct Dft<1, DIR, T> {
    __device__ __forceinline__ static void run(cpx<T>*) {}
};

template <int DIR, class T>
struct Dft<2, DIR, T> {
    __device__ __forceinline__ static void run(cpx<T>* a)
    {
        cpx<T> t = a[0];
        a[0] = t + a[1];
        a[1] = t - a[1];
    }
};

template <int DIR, class T>
struct Dft<3, DIR, T> {
    __device__ __forceinline__ static void run(cpx<T>* a)
    {
        const T c = T(-0.5);
        const T s = T(0.86602540378443864676372317075293618);  // sin(2 pi / 3)
        cpx<T> t = a[1] + a[2];
        cpx<T> d = a[1] - a[2];
        cpx<T> m = a[0] + c * t;
        a[0] = a[0] + t;
        cpx<T> e = mul_i<DIR>(s * d);
        a[1] = m + e;
        a[2] = m - e;
    }
};

template <int DIR, class T>
struct Dft<4, DIR, T> {
    __device__ __forceinline__ static void run(cpx<T>* a)
    {
        cpx<T> s0 = a[0] + a[2], d0 = a[0] - a[2];
        cpx<T> s1 = a[1] + a[3], d1 = mul_i<DIR>(a[1] - a[3]);
        a[0] = s0 + s1;
        a[2] = s0 - s1;
        a[1] = d0 + d1;
        a[3] = d0 - d1;
    }
};

template <int DIR, class T>
struct Dft<8, DIR, T> {
    __device__ __forceinline__ static void run(cpx<T>* a)
    {
        const T h = T(0.70710678118654752440084436210484904);
        cpx<T> e[4] = {a[0], a[2], a[4], a[6]};
        cpx<T> o[4] = {a[1], a[3], a[5], a[7]};
        Dft<4, DIR, T>::run(e);
        Dft<4, DIR, T>::run(o);
        // w8^k = e^{DIR i pi k / 4}
        cpx<T> o1 = h * cpx<T>{o[1].x - T(DIR) * o[1].y, o[1].y + T(DIR) * o[1].x};
        cpx<T> o2 = mul_i<DIR>(o[2]);
        cpx<T> o3 = h * cpx<T>{-o[3].x - T(DIR) * o[3].y, -o[3].y + T(DIR) * o[3].x};
        a[0] = e[0] + o[0];
        a[4] = e[0] - o[0];
        a[1] = e[1] + o1;
        a[5] = e[1] - o1;
        a[2] = e[2] + o2;
        a[6] = e[2] - o2;
        a[3] = e[3] + o3;
        a[7] = e[3] - o3;
    }
};

template <int DIR, class T>
struct Dft<12, DIR, T> {
    __device__ __forceinline__ static void run(cpx<T>* a)
    {
        // 12 = 4 x 3:  n = 3 n1 + n2,  k = k1 + 4 k2
        const T r3 = T(0.86602540378443864676372317075293618);  // sin(pi/3)
        cpx<T> b[3][4];
#pragma unroll
        for (int n2 = 0; n2 < 3; ++n2) {
            cpx<T> t[4] = {a[n2], a[3 + n2], a[6 + n2], a[9 + n2]};
            Dft<4, DIR, T>::run(t);
#pragma unroll
            for (int k1 = 0; k1 < 4; ++k1) b[n2][k1] = t[k1];
        }
        // twiddles w12^{n2 k1}, exponent 0..6: cos/sin of j pi / 6
        const T cs[7] = {T(1), r3, T(0.5), T(0), T(-0.5), -r3, T(-1)};
        const T sn[7] = {T(0), T(0.5), r3, T(1), r3, T(0.5), T(0)};
#pragma unroll
        for (int n2 = 1; n2 < 3; ++n2)
#pragma unroll
            for (int k1 = 1; k1 < 4; ++k1) {
                const int j = n2 * k1;
                cpx<T> w = {cs[j], T(DIR) * sn[j]};
                b[n2][k1] = b[n2][k1] * w;
            }
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1) {
            cpx<T> t[3] = {b[0][k1], b[1][k1], b[2][k1]};
            Dft<3, DIR, T>::run(t);
#pragma unroll
            for (int k2 = 0; k2 < 3; ++k2) a[k1 + 4 * k2] = t[k2];
        }
    }
};

template <int DIR, class T>
struct Dft<16, DIR, T> {
    __device__ __forceinline__ static void run(cpx<T>* a)
    {
        // 16 = 4 x 4:  n = 4 n1 + n2,  k = k1 + 4 k2
        const T c1 = T(0.92387953251128675612818318939678829);  // cos(pi/8)
        const T s1 = T(0.38268343236508977172845998403039887);  // sin(pi/8)
        const T h = T(0.70710678118654752440084436210484904);
        cpx<T> b[4][4];
#pragma unroll
        for (int n2 = 0; n2 < 4; ++n2) {
            cpx<T> t[4] = {a[n2], a[4 + n2], a[8 + n2], a[12 + n2]};
            Dft<4, DIR, T>::run(t);
#pragma unroll
            for (int k1 = 0; k1 < 4; ++k1) b[n2][k1] = t[k1];
        }
        // twiddles w16^{n2 k1}, exponent 0..9:  cos/sin of j pi / 8
        const T cs[10] = {T(1), c1, h, s1, T(0), -s1, -h, -c1, T(-1), -c1};
        const T sn[10] = {T(0), s1, h, c1, T(1), c1, h, s1, T(0), -s1};
#pragma unroll
        for (int n2 = 1; n2 < 4; ++n2)
#pragma unroll
            for (int k1 = 1; k1 < 4; ++k1) {
                const int j = n2 * k1;
                cpx<T> w = {cs[j], T(DIR) * sn[j]};
                b[n2][k1] = b[n2][k1] * w;
            }
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1) {
            cpx<T> t[4] = {b[0][k1], b[1][k1], b[2][k1], b[3][k1]};
            Dft<4, DIR, T>::run(t);
#pragma unroll
            for (int k2 = 0; k2 < 4; ++k2) a[k1 + 4 * k2] = t[k2];
        }
    }
};

// ---------------------------------------------------------------------------
// FFT plans: the factorisation of a length into radices, first stage first.
// ---------------------------------------------------------------------------
template <int... Rs>
struct Plan {
    static constexpr int S = sizeof...(Rs);
    static constexpr int r[S] = {Rs...};
    static constexpr int L = (Rs * ...);
    static constexpr int rmax()
    {
        int m = 1;
        for (int i = 0; i < S; ++i) m = r[i] > m ? r[i] : m;
        return m;
    }
    static constexpr int rmin()
    {
        int m = 1 << 30;
        for (int i = 0; i < S; ++i) m = r[i] < m ? r[i] : m;
        return m;
    }
    static constexpr int radix(int t) { return r[t]; }
    static constexpr int pre(int t)
    {
        int p = 1;
        for (int i = 0; i < t; ++i) p *= r[i];
        return p;
    }
};

// a[r] *= w^{r k}, w = e^{DIR 2 pi i / (p R)}, taken from tw[j] = e^{-2 pi i j / NT}
template <int R, int DIR, class T>
__device__ __forceinline__ void twiddle(cpx<T>* a, int k, int step, const cpx<T>* __restrict__ tw)
{
#pragma unroll
    for (int r = 1; r < R; ++r) {
        cpx<T> w = tw[r * k * step];
        if (DIR > 0) w.y = -w.y;
        a[r] = a[r] * w;
    }
}

// copy the twiddle table of length N into LDS (all threads of the block)
template <int N, class T>
__device__ __forceinline__ const cpx<T>* stage_twiddles(const cpx<T>* __restrict__ g, cpx<T>* s)
{
    for (int i = threadIdx.x; i < N; i += blockDim.x) s[i] = g[i];
    __syncthreads();
    return s;
}

// ---------------------------------------------------------------------------
// Column FFT: transforms along the slow axis of a tile of W columns.  Thread
// (c, g) = (tid % W, tid / W) owns butterflies i = g + j*G of column c, G = L/RMAX.
// Registers a[RMAX] hold, on entry, the stage-0 inputs
//     a[j*R0 + r] = x[g + j*G + r*(L/R0)]
// and, on exit, the outputs
//     a[j*RL + r] = X[g + j*G + r*(L/RL)]
// Intermediate stages go through the LDS tile lds[L][LDW].
// ---------------------------------------------------------------------------
template <class P, int W, int LDW, int DIR, class T>
struct ColFFT {
    static constexpr int L = P::L;
    static constexpr int RM = P::rmax();
    static constexpr int G = L / RM;
    static constexpr int NT = W * G;
    static constexpr bool divisible()
    {
        for (int t = 0; t < P::S; ++t)
            if (RM % P::radix(t) != 0) return false;
        return true;
    }
    static_assert(divisible(), "column plans need every radix to divide the largest one");

    template <int t>
    __device__ __forceinline__ static void stage(cpx<T>* a, cpx<T>* lds, int c, int g,
                                                 const cpx<T>* __restrict__ tw, int twn)
    {
        if constexpr (t < P::S) {
            constexpr int R = P::radix(t);
            constexpr int p = P::pre(t);
            constexpr int BPT = RM / R;
            if constexpr (t > 0) {
#pragma unroll
                for (int j = 0; j < BPT; ++j) {
                    const int i = g + j * G;
#pragma unroll
                    for (int r = 0; r < R; ++r) a[j * R + r] = lds[(i + r * (L / R)) * LDW + c];
                }
                __syncthreads();
#pragma unroll
                for (int j = 0; j < BPT; ++j) {
                    const int i = g + j * G;
                    twiddle<R, DIR, T>(a + j * R, i % p, twn / (p * R), tw);
                }
            }
#pragma unroll
            for (int j = 0; j < BPT; ++j) Dft<R, DIR, T>::run(a + j * R);
            if constexpr (t + 1 < P::S) {
#pragma unroll
                for (int j = 0; j < BPT; ++j) {
                    const int i = g + j * G;
                    const int k = i % p;
                    const int base = (i - k) * R + k;
#pragma unroll
                    for (int r = 0; r < R; ++r) lds[(base + r * p) * LDW + c] = a[j * R + r];
                }
                __syncthreads();
            }
            stage<t + 1>(a, lds, c, g, tw, twn);
        }
    }

    // position along the column of stage-0 input a[q]
    __device__ __forceinline__ static int in_pos(int g, int q)
    {
        constexpr int R0 = P::radix(0);
        return g + (q / R0) * G + (q % R0) * (L / R0);
    }
    // position along the column of final output a[q]
    __device__ __forceinline__ static int out_pos(int g, int q)
    {
        constexpr int RL = P::radix(P::S - 1);
        return g + (q / RL) * G + (q % RL) * (L / RL);
    }

    __device__ __forceinline__ static void run(cpx<T>* a, cpx<T>* lds, int c, int g,
                                               const cpx<T>* __restrict__ tw, int twn)
    {
        stage<0>(a, lds, c, g, tw, twn);
    }

    // F independent column transforms taken stage by stage together (one pair of barriers per
    // stage for all F instead of one per transform): the same arithmetic per transform as run(),
    // with the exchanges of transform f in the LDS tile lds + f L LDW
    template <int F, int t>
    __device__ __forceinline__ static void stage_multi(cpx<T> (*a)[RM], cpx<T>* lds, int c, int g,
                                                       const cpx<T>* __restrict__ tw, int twn)
    {
        if constexpr (t < P::S) {
            constexpr int R = P::radix(t);
            constexpr int p = P::pre(t);
            constexpr int BPT = RM / R;
            if constexpr (t > 0) {
#pragma unroll
                for (int f = 0; f < F; ++f)
#pragma unroll
                    for (int j = 0; j < BPT; ++j) {
                        const int i = g + j * G;
#pragma unroll
                        for (int r = 0; r < R; ++r) a[f][j * R + r] = lds[f * L * LDW + (i + r * (L / R)) * LDW + c];
                    }
                __syncthreads();
#pragma unroll
                for (int f = 0; f < F; ++f)
#pragma unroll
                    for (int j = 0; j < BPT; ++j) {
                        const int i = g + j * G;
                        twiddle<R, DIR, T>(a[f] + j * R, i % p, twn / (p * R), tw);
                    }
            }
#pragma unroll
            for (int f = 0; f < F; ++f)
#pragma unroll
                for (int j = 0; j < BPT; ++j) Dft<R, DIR, T>::run(a[f] + j * R);
            if constexpr (t + 1 < P::S) {
#pragma unroll
                for (int f = 0; f < F; ++f)
#pragma unroll
                    for (int j = 0; j < BPT; ++j) {
                        const int i = g + j * G;
                        const int k = i % p;
                        const int base = (i - k) * R + k;
#pragma unroll
                        for (int r = 0; r < R; ++r) lds[f * L * LDW + (base + r * p) * LDW + c] = a[f][j * R + r];
                    }
                __syncthreads();
            }
            stage_multi<F, t + 1>(a, lds, c, g, tw, twn);
        }
    }
    template <int F>
    __device__ __forceinline__ static void run_multi(cpx<T> (*a)[RM], cpx<T>* lds, int c, int g,
                                                     const cpx<T>* __restrict__ tw, int twn)
    {
        stage_multi<F, 0>(a, lds, c, g, tw, twn);
    }
};

// ---------------------------------------------------------------------------
// Row FFT: NB sequences of length L resident in LDS (sequence s at lds + s*BS,
// element j at offset pad(j) = j + j/PD, or j when PD = 0: one gap every PD elements
// breaks the power-of-two strides of the Stockham stages across LDS banks),
// transformed in place by all NTH threads of the block.
// ---------------------------------------------------------------------------
template <class P, int NB, int BS, int NTH, int DIR, class T, int PD = 0>
struct RowFFT {
    static constexpr int L = P::L;
    __device__ __forceinline__ static constexpr int pad(int j) { return PD ? j + j / PD : j; }

    template <int t>
    __device__ __forceinline__ static void stage(cpx<T>* lds, int tid, const cpx<T>* __restrict__ tw,
                                                 int twn)
    {
        if constexpr (t < P::S) {
            constexpr int R = P::radix(t);
            constexpr int p = P::pre(t);
            constexpr int NBF = NB * (L / R);               // butterflies in the stage
            constexpr int Q = (NBF + NTH - 1) / NTH;        // per thread
            cpx<T> a[Q * R];
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                const int b = tid + q * NTH;
                if (NBF % NTH == 0 || b < NBF) {
                    const int s = b / (L / R), i = b % (L / R);
#pragma unroll
                    for (int r = 0; r < R; ++r) a[q * R + r] = lds[s * BS + pad(i + r * (L / R))];
                }
            }
            __syncthreads();
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                const int b = tid + q * NTH;
                if (NBF % NTH == 0 || b < NBF) {
                    const int s = b / (L / R), i = b % (L / R);
                    const int k = i % p;
                    if constexpr (t > 0) twiddle<R, DIR, T>(a + q * R, k, twn / (p * R), tw);
                    Dft<R, DIR, T>::run(a + q * R);
                    const int base = (i - k) * R + k;
#pragma unroll
                    for (int r = 0; r < R; ++r) lds[s * BS + pad(base + r * p)] = a[q * R + r];
                }
            }
            __syncthreads();
            stage<t + 1>(lds, tid, tw, twn);
        }
    }

    __device__ __forceinline__ static void run(cpx<T>* lds, int tid, const cpx<T>* __restrict__ tw,
                                               int twn)
    {
        stage<0>(lds, tid, tw, twn);
    }
};

}  // namespace dnsbox
