// The middle passes of an RHS for one z-plane per block: the y-inverse FFT of the FI input
// fields, the x pass (c2r -> pointwise op -> r2c) and the y-forward FFT of the FO outputs, with
// the plane held in LDS between them (rhs.f90:30-65 for OpCross; its adjoint for OpVjp).  One
// launch replaces three (y-inverse, x pass, y-forward) and the plane crosses HBM once each way.
// A z-plane of FI fields fits a CU's LDS up to 64^3 (FP32) and 48^3 / 32^3 (FP64, forward / step
// adjoint); the boxes up to 32^3 run it (PlaneGeo::used), larger ones the three passes.
//
// LDS layout of the plane: field f, slot j (the kx index, 0..N/2), element y (0..N-1) at
//   f * FST + j * CST + y,   CST = N + 1, FST = (N/2 + 1) CST,
// so a y-column is contiguous and an x-line is strided by CST.  Every field has N/2 + 1 slots:
// the c2r packing writes the N/2 complex values of the packed line in place of the
// half-spectrum, and the outputs overwrite the first FO input fields slot by slot.
// Phases (a block barrier between each):
//   load    kept rows (PR: ky in Ky) and slots (PR: kx in Kx) of the plane from global memory;
//           the other rows of the loaded slots set to zero
//   y-inv   in-place Stockham FFTs of the loaded columns (radices of Cfg::YP)
//   pack    Z[k] = (X[k] + conj X[M-k]) + i e^{+2 pi i k/N} (X[k] - conj X[M-k]), pairs (k, M-k)
//   x-inv   complex FFTs of length M = N/2 along the slots (radices of Cfg::RP)
//   op      Op on the real and imaginary parts (x = 2m and 2m+1) of every point
//   x-fwd   forward FFTs of the FO outputs
//   unpack  X[k] = 1/2 (Y[k] + conj Y[M-k]) - 1/2 i e^{-2 pi i k/N} (Y[k] - conj Y[M-k]), k <= M
//   y-fwd   forward FFTs of the output slots (PO: kx in Kx), store of the output rows (PO: Ky)
// The arithmetic of each phase is that of the three-pass path (same radices, twiddles and
// pre/post-processing), so the results agree with it to rounding.
#pragma once

#include "box.hpp"

namespace dnsbox {

constexpr int kPlaneThreads = 1024;
constexpr int kPlaneLdsLimit = 160 * 1024;   // bytes of LDS per CU on gfx950

template <int N, class T, class Op>
struct PlaneGeo {
    static constexpr int M = N / 2, H = M + 1;
    static constexpr int CST = N + 1;
    static constexpr int FST = H * CST;
    static constexpr int FI = Op::FI, FO = Op::FO;
    static constexpr int LDS = FI * FST;   // complex elements
    static constexpr long long BYTES = (long long)(LDS + N) * sizeof(cpx<T>) + 4 * 16 * sizeof(double);
    static constexpr bool fits =
        N >= 16 && N <= 64 && BYTES <= kPlaneLdsLimit && FO <= FI && FO > 0 && Op::IN_SPEC && Op::OUT_SPEC;
    // the sizes that run it: measured against the three passes (steps/s of single boxes, both
    // precisions): 16^3 +23 %, 32^3 +6 % (FP64) and +15 % (FP32), 48^3 -1 % (FP64) and +5 % (FP32),
    // 64^3 FP32 -6 %.  The plane blocks are bound by LDS traffic and barriers (one block per CU
    // at 48^3 and 64^3); below 48^3 the launches they save outweigh that.
    static constexpr bool used = fits && N <= 32;
};

// slot of kept column x (PR: 0..KMAX and the Nyquist index; otherwise x itself)
template <int N, bool PR>
struct SlotMap {
    __device__ __forceinline__ static int of(int x) { return PR ? Keep<N>::ix(x) : x; }
};

// y-columns: sequence s = (field s / NS, slot of column s % NS), elements contiguous
template <int N, int FST, int CST, int NS, bool PR>
struct ColSeq {
    __device__ __forceinline__ static int base(int s) { return (s / NS) * FST + SlotMap<N, PR>::of(s % NS) * CST; }
};
// x-lines: sequence s = (field s / N, y = s % N), elements CST apart
template <int N, int FST>
struct LineSeq {
    __device__ __forceinline__ static int base(int s) { return (s / N) * FST + s % N; }
};

// In-place Stockham FFT of NSEQ sequences of length P::L in LDS, element j of sequence s at
// lds[Seq::base(s) + j * ES]; every stage reads its butterflies, waits for the block, writes
// them back in autosort order and waits again.  Twiddles from tw[j] = e^{-2 pi i j / TWN}.
template <class P, int NSEQ, int ES, int NT, int DIR, class T, int TWN, class Seq>
struct LdsFFT {
    static constexpr int L = P::L;
    template <int t>
    __device__ __forceinline__ static void stage(cpx<T>* lds, int tid, const cpx<T>* __restrict__ tw)
    {
        if constexpr (t < P::S) {
            constexpr int R = P::radix(t), p = P::pre(t), B = L / R;
            constexpr int NBF = NSEQ * B, Q = (NBF + NT - 1) / NT;
            cpx<T> a[Q * R];
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                const int b = tid + q * NT;
                if (NBF % NT == 0 || b < NBF) {
                    const int s = b / B, i = b % B;
                    const int base = Seq::base(s);
#pragma unroll
                    for (int r = 0; r < R; ++r) a[q * R + r] = lds[base + (i + r * B) * ES];
                }
            }
            __syncthreads();
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                const int b = tid + q * NT;
                if (NBF % NT == 0 || b < NBF) {
                    const int s = b / B, i = b % B;
                    const int base = Seq::base(s);
                    const int k = i % p;
                    if constexpr (t > 0) twiddle<R, DIR, T>(a + q * R, k, TWN / (p * R), tw);
                    Dft<R, DIR, T>::run(a + q * R);
                    const int o = (i - k) * R + k;
#pragma unroll
                    for (int r = 0; r < R; ++r) lds[base + (o + r * p) * ES] = a[q * R + r];
                }
            }
            __syncthreads();
            stage<t + 1>(lds, tid, tw);
        }
    }
    __device__ __forceinline__ static void run(cpx<T>* lds, int tid, const cpx<T>* __restrict__ tw)
    {
        stage<0>(lds, tid, tw);
    }
};

// c2r packing value of position k (the x pass's pre-processing): A = X[k], Bm = X[M-k]
template <class T>
__device__ __forceinline__ cpx<T> plane_pack(cpx<T> A, cpx<T> Bm, int k, const cpx<T>* __restrict__ tw)
{
    cpx<T> B = conj(Bm);
    if (k == 0) {
        A.y = T(0);
        B.y = T(0);
    }
    cpx<T> w = tw[k];
    w.y = -w.y;
    const cpx<T> fe = A + B, fo = (A - B) * w;
    return fe + mul_i<1>(fo);
}
// r2c post-processing value of position k: zk = Y[k], zmr = Y[M-k] (Y[M] = Y[0])
template <class T>
__device__ __forceinline__ cpx<T> plane_unpack(cpx<T> zk, cpx<T> zmr, int k, const cpx<T>* __restrict__ tw)
{
    const cpx<T> zm = conj(zmr);
    const cpx<T> a = T(0.5) * (zk + zm);
    const cpx<T> b = (T(0.5) * (zk - zm)) * tw[k];
    return a + mul_i<-1>(b);
}

// the plane iz of the scratch X (FI fields fs apart, layout (z, ky, kx)) through the three
// passes, outputs in fields 0..FO-1 of the same plane.  PR: inputs confined to the kept cube
// (rows ky in Ky and slots kx in Kx read); PO: outputs stored at kept rows and slots only.
template <int N, class T, class Op, bool PR, bool PO>
__device__ __forceinline__ void plane_tile(cpx<T>* __restrict__ X, long long fs, int iz, cpx<T>* lds,
                                           const cpx<T>* __restrict__ tw, XRed<T>& red)
{
    using PG = PlaneGeo<N, T, Op>;
    using K = Keep<N>;
    static_assert(PG::fits, "plane does not fit the LDS");
    constexpr int M = PG::M, H = PG::H, CST = PG::CST, FST = PG::FST, FI = PG::FI, FO = PG::FO;
    constexpr int NT = kPlaneThreads;
    constexpr int XC = PR ? K::KX : H, NR = PR ? K::KY : N;    // input slots and rows
    constexpr int XCO = PO ? K::KX : H, NRO = PO ? K::KY : N;  // output slots and rows
    constexpr int KP = M / 2 + 1;                              // (k, M - k) pairs of a line
    const int tid = threadIdx.x;
    const long long pl = (long long)iz * N * H;
    const cpx<T> zero{T(0), T(0)};

    // ---- load (all loads of a thread issued before its LDS writes) ----
    {
        constexpr int NE = FI * NR * XC, Q = (NE + NT - 1) / NT;
        cpx<T> buf[Q];
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int e0 = tid + q * NT;
            const int e = (NE % NT == 0 || e0 < NE) ? e0 : 0;   // the tail re-reads element 0
            const int x = e % XC, r = (e / XC) % NR, f = e / (XC * NR);
            const int iy = PR ? K::iy(r) : r, ix = SlotMap<N, PR>::of(x);
            buf[q] = X[f * fs + pl + iy * H + ix];
        }
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int e = tid + q * NT;
            if (NE % NT == 0 || e < NE) {
                const int x = e % XC, r = (e / XC) % NR, f = e / (XC * NR);
                const int iy = PR ? K::iy(r) : r, ix = SlotMap<N, PR>::of(x);
                lds[f * FST + ix * CST + iy] = buf[q];
            }
        }
        if constexpr (PR) {   // rows outside Ky (KMAX + 1 .. N - KMAX - 1 but the Nyquist row) are zero
            constexpr int NZR = N - NR, NZ = FI * NZR * XC;
            for (int e = tid; e < NZ; e += NT) {
                const int x = e % XC, rr = (e / XC) % NZR, f = e / (XC * NZR);
                int iy = K::KMAX + 1 + rr;
                if (iy >= N / 2) ++iy;
                lds[f * FST + K::ix(x) * CST + iy] = zero;
            }
        }
        __syncthreads();
    }

    // ---- y-inverse of the loaded columns ----
    LdsFFT<typename Cfg<N, T>::YP, FI * XC, 1, NT, +1, T, N, ColSeq<N, FST, CST, XC, PR>>::run(lds, tid, tw);

    // ---- c2r packing, in place, pairs (k, M - k) of every line ----
    for (int e = tid; e < FI * N * KP; e += NT) {
        const int k = e % KP, y = (e / KP) % N, f = e / (KP * N);
        cpx<T>* col = lds + f * FST + y;
        const bool ka = !PR || k <= K::KMAX;                 // X[k], k < M
        const bool kb = !PR || k == 0 || M - k <= K::KMAX;   // X[M - k] (k = 0: the Nyquist slot)
        const cpx<T> A = ka ? col[k * CST] : zero;
        const cpx<T> B = kb ? col[(M - k) * CST] : zero;
        if (k == 0) {
            col[0] = plane_pack(A, B, 0, tw);
        } else if (2 * k == M) {
            col[k * CST] = plane_pack(A, A, k, tw);
        } else {
            col[k * CST] = plane_pack(A, B, k, tw);
            col[(M - k) * CST] = plane_pack(B, A, M - k, tw);
        }
    }
    __syncthreads();

    // ---- x: inverse complex FFTs, pointwise op, forward FFTs ----
    LdsFFT<typename Cfg<N, T>::RP, FI * N, CST, NT, +1, T, N, LineSeq<N, FST>>::run(lds, tid, tw);
    for (int e = tid; e < N * M; e += NT) {
        const int y = e % N, m = e / N;
        T vr[FI], vi[FI], orr[FO], oi[FO];
#pragma unroll
        for (int f = 0; f < FI; ++f) {
            const cpx<T> v = lds[f * FST + m * CST + y];
            vr[f] = v.x;
            vi[f] = v.y;
        }
        Op::apply(vr, orr, red);
        Op::apply(vi, oi, red);
#pragma unroll
        for (int f = 0; f < FO; ++f) lds[f * FST + m * CST + y] = cpx<T>{orr[f], oi[f]};
    }
    __syncthreads();
    LdsFFT<typename Cfg<N, T>::RP, FO * N, CST, NT, -1, T, N, LineSeq<N, FST>>::run(lds, tid, tw);

    // ---- r2c post-processing, in place ----
    for (int e = tid; e < FO * N * KP; e += NT) {
        const int k = e % KP, y = (e / KP) % N, f = e / (KP * N);
        cpx<T>* col = lds + f * FST + y;
        const cpx<T> yk = col[k * CST];
        if (k == 0) {
            col[0] = plane_unpack(yk, yk, 0, tw);
            col[M * CST] = plane_unpack(yk, yk, M, tw);
        } else if (2 * k == M) {
            col[k * CST] = plane_unpack(yk, yk, k, tw);
        } else {
            const cpx<T> ym = col[(M - k) * CST];
            col[k * CST] = plane_unpack(yk, ym, k, tw);
            col[(M - k) * CST] = plane_unpack(ym, yk, M - k, tw);
        }
    }
    __syncthreads();

    // ---- y-forward of the output slots and the store of the output rows ----
    LdsFFT<typename Cfg<N, T>::YP, FO * XCO, 1, NT, -1, T, N, ColSeq<N, FST, CST, XCO, PO>>::run(lds, tid, tw);
    {
        constexpr int NE = FO * NRO * XCO;
        for (int e = tid; e < NE; e += NT) {
            const int x = e % XCO, r = (e / XCO) % NRO, f = e / (XCO * NRO);
            const int iy = PO ? K::iy(r) : r, ix = SlotMap<N, PO>::of(x);
            X[f * fs + pl + iy * H + ix] = lds[f * FST + ix * CST + iy];
        }
    }
}

template <int N, class T, class Op, bool PR, bool PO>
__global__ void __launch_bounds__(kPlaneThreads)
    k_plane(cpx<T>* __restrict__ X, long long fs, const cpx<T>* __restrict__ tw_g, unsigned long long* red_bits)
{
    __shared__ cpx<T> lds[PlaneGeo<N, T, Op>::LDS];
    __shared__ cpx<T> stw_[N];
    __shared__ double wsm[4 * 16];
    const cpx<T>* tw = stage_twiddles<N, T>(tw_g, stw_);
    XRed<T> red;
#pragma unroll
    for (int j = 0; j < 4; ++j) red.m[j] = T(0);
    plane_tile<N, T, Op, PR, PO>(X, fs, blockIdx.x, lds, tw, red);
    if constexpr (Op::NRED > 0) {
        if (red_bits) block_max_atomic<Op::NRED>(red, red_bits, false, wsm);
    }
}

// algorithmic bytes of one plane launch over the N planes: the loaded rows and slots of FI
// fields read, the stored rows and slots of FO fields written
template <int N, class T, class Op, bool PR, bool PO>
constexpr double plane_bytes()
{
    using K = Keep<N>;
    constexpr double in = double(Op::FI) * (PR ? K::KY : N) * (PR ? K::KX : N / 2 + 1);
    constexpr double out = double(Op::FO) * (PO ? K::KY : N) * (PO ? K::KX : N / 2 + 1);
    return double(N) * (in + out) * sizeof(cpx<T>);
}

}  // namespace dnsbox
