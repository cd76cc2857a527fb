// The cross-product x pass (rhs.f90:30-65: six c2r, (omega x u) / N^3 ... scaled inputs, three r2c)
// with the x-lines held in registers: L = N/16 lanes per line (16 at N = 256, 32 at N = 512), 8
// complex values per lane and field, 64/L lines per wavefront, each wavefront on its own lines
// with its own LDS exchange buffer (no block barrier).
//
// A real line of N samples is packed as M = N/2 complex values z[m] = x[2m] + i x[2m+1]. Its
// complex transform of length M runs as three register-local stages with two exchanges through
// the wavefront's LDS buffer (digits k = p + 8q + 32r with p < 8, q < 4, r < R1 = M/32, and
// m = a + R1 b + L c with a < R1, b < 4, c < 8; w = e^{2 pi i / M}):
//   inverse  x[m] = sum_p w8^{cp} w^{p(a+R1 b)} sum_q w4^{bq} w^{8aq} sum_r wR1^{ar} Z[p+8q+32r]
//            stage 1: DFT_R1 over r (lane p + 8h holds q = h + (L/8) j), exchange, stage 2:
//            twiddle, DFT4 over q (lane p + 8h takes a = h, h + L/8), exchange, stage 3:
//            twiddle, DFT8 over p (lane a + R1 b) -> lane l holds x[l + L c], c = 0..7
//   forward  the same stages in reverse order with conjugate twiddles, from x[l + L c] (so the
//            pointwise product and the first forward stage run on registers) to Y[k] at
//            k = l + L s, s = 0..7, the positions the inverse read Z from.
// The exchange layouts (positions in complex elements of the line's buffer)
//   L1 (a, q, p) -> 40a + 8q + p                   between stages 1 and 2
//   L3 (r, c)    -> 17(r >> 1) + 8(r & 1) + c      between stages 2 and 3 (r = a + R1 b, c = p)
// give every 16-lane half of a line's access 16 distinct positions mod 16, i.e. 16 distinct
// pairs of the 32 4-byte LDS banks (no bank conflict), and keep each lane's addresses a per-lane
// base plus immediate offsets.  The c2r/r2c and stage-3 twiddles come from a per-block LDS table; the six stage-2
// twiddles of a lane stay in registers.
// All six inverse transforms of a line end at the same positions in the same lane, so the cross
// product is computed in registers; the outputs are transformed as soon as they are complete
// (n0 after omega_y, n1 and n2 after omega_x) to bound the registers held.
// The c2r pre-processing reads X[k] and X[M-k] straight from global memory, addressed as a
// uniform field base plus one 32-bit lane offset shared by all fields (the next field's loads are
// issued before the current field's transform); the r2c post-processing pairs Y[k] with Y[M-k]
// through the LDS buffer.  Inputs and outputs may be the same scratch fields (an output overwrites
// its field only after every lane of the line has read that field into registers).
#pragma once

#include "box.hpp"

namespace dnsbox {

// sizes with a register-resident cross-product x pass
template <int N, class T>
struct XReg {
    static constexpr bool enabled = (N == 256 || N == 512) && std::is_same<T, float>::value;
};

namespace xreg {

constexpr int kWaves = 4;             // wavefronts per block
constexpr int kThreads = 64 * kWaves;

// the line geometry of size N (see the digit expansion above)
template <int N>
struct Shape {
    static constexpr int M = N / 2;           // complex values per packed line
    static constexpr int L = M / 8;           // lanes per line
    static constexpr int R1 = M / 32;         // radix of stage 1, number of a
    static constexpr int HG = L / 8;          // lane groups h of a line (lane l = p + 8h)
    static constexpr int J = 4 / HG;          // q held per lane in stage 1: q = h + HG j
    static constexpr int LPW = 64 / L;        // lines per wavefront
    static constexpr int LPB = LPW * kWaves;  // lines per block
    static constexpr int ROW = 40 * (R1 - 1) + 32;   // line buffer (complex elements): L1's span
    static_assert(R1 == 2 * HG, "stage 2 gives each lane two values of a");
    static_assert(17 * ((L - 1) >> 1) + 15 < ROW && M <= ROW, "L3 and the mirror fit the line buffer");
};

__device__ __forceinline__ int l1(int a, int q, int p) { return 40 * a + 8 * q + p; }
__device__ __forceinline__ int l3(int r, int c) { return 17 * (r >> 1) + 8 * (r & 1) + c; }

// orders the wavefront's LDS accesses around an exchange (no instruction at wavefront scope; it
// keeps the compiler from moving a lane's LDS access across the exchange)
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <class T>
__device__ __forceinline__ cpx<T> cmul(cpx<T> a, cpx<T> w)
{
    return cpx<T>{a.x * w.x - a.y * w.y, a.x * w.y + a.y * w.x};
}

template <class T>
__device__ __forceinline__ cpx<T> cmulc(cpx<T> a, cpx<T> w)   // a * conj(w)
{
    return cpx<T>{a.x * w.x + a.y * w.y, a.y * w.x - a.x * w.y};
}

// per-block twiddle table (LDS), from the table tw[j] = e^{-2 pi i j / N}:
//   wp[k]      = e^{+2 pi i k / N}, k < M     (c2r pre-processing; the r2c uses the conjugate)
//   w3[p-1][l] = w^{p l}, p = 1..7, l < L     (w = e^{+2 pi i / M})
//   w2[a][q-1] = w^{8 a q}, a < R1, q = 1..3
// The lanes of a line read wp[l + L s] and w3[.][l] at distinct banks and w2 at two per half.
template <int N, class T>
struct TwL {
    using S = Shape<N>;
    cpx<T> wp[S::M];
    cpx<T> w3[7][S::L];
    cpx<T> w2[S::R1][3];
};

template <int N, class T>
__device__ __forceinline__ void fill_tw(const cpx<T>* __restrict__ tw, TwL<N, T>& t, int tid)
{
    using S = Shape<N>;
    auto cj = [&](int j) { const cpx<T> v = tw[j & (N - 1)]; return cpx<T>{v.x, -v.y}; };
    for (int i = tid; i < S::M + 7 * S::L + 3 * S::R1; i += kThreads) {
        if (i < S::M) {
            t.wp[i] = cj(i);
        } else if (i < S::M + 7 * S::L) {
            const int p = (i - S::M) / S::L + 1, l = (i - S::M) % S::L;
            t.w3[p - 1][l] = cj(2 * ((p * l) & (S::M - 1)));   // w^n = conj tw[2n]
        } else {
            const int a = (i - S::M - 7 * S::L) / 3, q = (i - S::M - 7 * S::L) % 3 + 1;
            t.w2[a][q - 1] = cj(2 * ((8 * a * q) & (S::M - 1)));
        }
    }
}

// c2r packing of a half-spectrum line X (row, k <= M): z[s] = Z[k], k = l + L s,
//   Z[k] = (X[k] + conj X[M-k]) + i e^{+2 pi i k/N} (X[k] - conj X[M-k]),
// the imaginary parts of X[0] and X[M] ignored.  PR: X is zero at k > KMAX (k < M).
template <int N, class T, bool PR>
struct Raw {   // X[k] and X[M-k] of the lane's positions k = l + L s, as loaded
    cpx<T> A[8], B[8];
};

// Loads of a pruned line are unconditional: a lane whose position is not kept reads a kept
// position of the same line (KMAX, or the Nyquist entry M) and takes zero, so no load is under a
// lane-dependent branch.
template <class T>
__device__ __forceinline__ cpx<T> ldb(const char* base, unsigned off)
{
    return *reinterpret_cast<const cpx<T>*>(base + off);
}

template <class T>
__device__ __forceinline__ void stb(char* base, unsigned off, cpx<T> v)
{
    *reinterpret_cast<cpx<T>*>(base + off) = v;
}

// Loads of a pruned line are unconditional: a lane whose position is not kept reads a kept
// position of the same line (KMAX, or the Nyquist entry M) and takes zero, so no load is under a
// lane-dependent branch.  base: the field (uniform); line: byte offset of the line in the field.
template <int N, class T, bool PR>
__device__ __forceinline__ void load_raw(const char* base, unsigned line, int l, Raw<N, T, PR>& r)
{
    constexpr int M = N / 2, K = Keep<N>::KMAX, L = Shape<N>::L;
    constexpr unsigned E = sizeof(cpx<T>);
    const cpx<T> zero{T(0), T(0)};
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        const int k = l + L * s;
        // X[k], k in [L s, L s + L - 1]
        if (!PR || L * s + L - 1 <= K) {
            r.A[s] = ldb<T>(base + L * E * s, line + unsigned(l) * E);
        } else if (L * s > K) {
            r.A[s] = zero;
        } else {
            r.A[s] = ldb<T>(base, line + unsigned(min(k, K)) * E);   // zeroed in pack
        }
        // X[M - k], M - k in [M - L s - L + 1, M - L s]; kept: M - k <= K, or k = 0 (Nyquist)
        if (!PR || (s > 0 && M - L * s <= K)) {
            r.B[s] = ldb<T>(base - L * E * s, line + unsigned(M - l) * E);
        } else if (s > 0 && M - L * s - L + 1 > K) {
            r.B[s] = zero;
        } else {
            const bool kb = M - k <= K || k == 0;
            r.B[s] = ldb<T>(base, line + unsigned(kb ? M - k : M) * E);   // zeroed in pack
        }
    }
}

template <int N, class T, bool PR>
__device__ __forceinline__ void pack(const Raw<N, T, PR>& r, int l, const TwL<N, T>& t, cpx<T> (&z)[8])
{
    constexpr int M = N / 2, K = Keep<N>::KMAX, L = Shape<N>::L;
    const cpx<T> zero{T(0), T(0)};
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        const int k = l + L * s;
        cpx<T> a = r.A[s], b = conj(r.B[s]);
        // the clamped loads of load_raw (lane-dependent keep) take zero here, where the data is used
        if (PR && !(L * s + L - 1 <= K) && !(L * s > K)) a = k <= K ? a : zero;
        if (PR && !(s > 0 && M - L * s <= K) && !(s > 0 && M - L * s - L + 1 > K))
            b = (M - k <= K || k == 0) ? b : zero;
        if (s == 0 && l == 0) {
            a.y = T(0);
            b.y = T(0);
        }
        const cpx<T> fe = a + b, fo = cmul(a - b, t.wp[l + L * s]);
        z[s] = fe + mul_i<1>(fo);
    }
}

// inverse complex transform of z[s] = Z[l + L s] -> x[c] = x[l + L c] (unnormalised, e^{+});
// w2 = the lane's stage-2 twiddles w^{8aq}, a = h + HG ai
template <int N, class T>
__device__ __forceinline__ void inverse_line(cpx<T> (&z)[8], cpx<T>* __restrict__ buf, int l, const TwL<N, T>& t,
                                             const cpx<T> (&w2)[2][3])
{
    using S = Shape<N>;
    const int p = l & 7, h = l >> 3;
    // stage 1: DFT_R1 over r of Z[p + 8q + 32r], q = h + HG j  (z index s = j + J r)
#pragma unroll
    for (int j = 0; j < S::J; ++j) {
        cpx<T> v[S::R1];
#pragma unroll
        for (int r = 0; r < S::R1; ++r) v[r] = z[j + S::J * r];
        Dft<S::R1, +1, T>::run(v);
#pragma unroll
        for (int a = 0; a < S::R1; ++a) buf[l1(a, h + S::HG * j, p)] = v[a];
    }
    wave_lds_sync();
    // stage 2: lane (p, h) takes a = h + HG ai; twiddle w^{8aq}, DFT4 over q
    cpx<T> u2[2][4];
#pragma unroll
    for (int ai = 0; ai < 2; ++ai)
#pragma unroll
        for (int q = 0; q < 4; ++q) u2[ai][q] = buf[l1(h + S::HG * ai, q, p)];
    wave_lds_sync();
#pragma unroll
    for (int ai = 0; ai < 2; ++ai) {
#pragma unroll
        for (int q = 1; q < 4; ++q) u2[ai][q] = cmul(u2[ai][q], w2[ai][q - 1]);
        Dft<4, +1, T>::run(u2[ai]);
#pragma unroll
        for (int b = 0; b < 4; ++b) buf[l3(h + S::HG * ai + S::R1 * b, p)] = u2[ai][b];
    }
    wave_lds_sync();
    // stage 3: lane l = a + R1 b; twiddle w^{p l}, DFT8 over p -> x[l + L c]
    z[0] = buf[l3(l, 0)];
#pragma unroll
    for (int q = 1; q < 8; ++q) z[q] = cmul(buf[l3(l, q)], t.w3[q - 1][l]);
    wave_lds_sync();
    Dft<8, +1, T>::run(z);
}

// forward complex transform of y[c] = y[l + L c] -> y[s] = Y[l + L s] (unnormalised, e^{-})
template <int N, class T>
__device__ __forceinline__ void forward_line(cpx<T> (&y)[8], cpx<T>* __restrict__ buf, int l, const TwL<N, T>& t,
                                             const cpx<T> (&w2)[2][3])
{
    using S = Shape<N>;
    const int p = l & 7, h = l >> 3;
    Dft<8, -1, T>::run(y);
    buf[l3(l, 0)] = y[0];
#pragma unroll
    for (int q = 1; q < 8; ++q) buf[l3(l, q)] = cmulc(y[q], t.w3[q - 1][l]);
    wave_lds_sync();
    cpx<T> u2[2][4];
#pragma unroll
    for (int ai = 0; ai < 2; ++ai)
#pragma unroll
        for (int b = 0; b < 4; ++b) u2[ai][b] = buf[l3(h + S::HG * ai + S::R1 * b, p)];
    wave_lds_sync();
#pragma unroll
    for (int ai = 0; ai < 2; ++ai) {
        Dft<4, -1, T>::run(u2[ai]);
        buf[l1(h + S::HG * ai, 0, p)] = u2[ai][0];
#pragma unroll
        for (int q = 1; q < 4; ++q) buf[l1(h + S::HG * ai, q, p)] = cmulc(u2[ai][q], w2[ai][q - 1]);
    }
    wave_lds_sync();
#pragma unroll
    for (int j = 0; j < S::J; ++j) {
        cpx<T> v[S::R1];
#pragma unroll
        for (int a = 0; a < S::R1; ++a) v[a] = buf[l1(a, h + S::HG * j, p)];
        Dft<S::R1, -1, T>::run(v);
#pragma unroll
        for (int r = 0; r < S::R1; ++r) y[j + S::J * r] = v[r];
    }
    wave_lds_sync();
}

// r2c post-processing of Y[k] (k = l + L s in y[s]) and the store of the kept modes:
//   X[k] = 1/2 (Y[k] + conj Y[M-k]) - 1/2 i e^{-2 pi i k/N} (Y[k] - conj Y[M-k]),  k <= M, Y[M] = Y[0]
template <int N, class T, bool PO>
__device__ __forceinline__ void store_unpacked(cpx<T> (&y)[8], cpx<T>* __restrict__ buf, char* base, unsigned line,
                                               int l, const TwL<N, T>& t)
{
    constexpr int M = N / 2, L = Shape<N>::L;
    constexpr int SMAX = PO ? (Keep<N>::KMAX / L + 1) : 8;   // s with l + L s possibly kept
#pragma unroll
    for (int s = 0; s < 8; ++s) buf[l + L * s] = y[s];
    wave_lds_sync();
    cpx<T> zm[SMAX];
#pragma unroll
    for (int s = 0; s < SMAX; ++s) zm[s] = conj(buf[(M - (l + L * s)) & (M - 1)]);
    wave_lds_sync();
#pragma unroll
    for (int s = 0; s < SMAX; ++s) {
        const int k = l + L * s;
        if (!PO || k <= Keep<N>::KMAX) {
            const cpx<T> a = T(0.5) * (y[s] + zm[s]);
            const cpx<T> b = cmulc(T(0.5) * (y[s] - zm[s]), t.wp[k]);
            stb<T>(base + L * sizeof(cpx<T>) * s, line + unsigned(l) * unsigned(sizeof(cpx<T>)), a + mul_i<-1>(b));
        }
    }
    if (l == 0) {   // Nyquist: Y[M] = Y[0], e^{-2 pi i M/N} = -1: X[M] = Re Y[0] - Im Y[0]
        stb<T>(base, line + unsigned(M) * unsigned(sizeof(cpx<T>)), cpx<T>{y[0].x - y[0].y, T(0)});
    }
}

template <int N, class T, bool PR>
__global__ void __launch_bounds__(kThreads)
    k_xcross_r(const cpx<T>* in0, cpx<T>* out0, long long fs, const cpx<T>* __restrict__ tw_g,
               unsigned long long* red_bits)
{
    using S = Shape<N>;
    // line offsets are 32-bit: the caller runs this kernel on fields under 4 GiB (xreg_used)
    constexpr int H = N / 2 + 1;
    __shared__ cpx<T> xbuf[S::LPB * S::ROW];
    __shared__ TwL<N, T> t;
    __shared__ double wsm[4 * 16];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int li = lane / S::L, l = lane % S::L;
    const long long row = (long long)xcd_contiguous(blockIdx.x, gridDim.x) * S::LPB + wave * S::LPW + li;
    cpx<T>* buf = xbuf + (wave * S::LPW + li) * S::ROW;
    fill_tw(tw_g, t, tid);
    __syncthreads();
    cpx<T> w2[2][3];
#pragma unroll
    for (int ai = 0; ai < 2; ++ai)
#pragma unroll
        for (int q = 0; q < 3; ++q) w2[ai][q] = t.w2[(l >> 3) + S::HG * ai][q];
    const char* fin = reinterpret_cast<const char*>(in0);
    char* fout = reinterpret_cast<char*>(out0);
    const long long fsb = fs * (long long)sizeof(cpx<T>);
    const unsigned line = unsigned(row) * unsigned(H * sizeof(cpx<T>));
    XRed<T> red;
    red.m[0] = red.m[1] = red.m[2] = red.m[3] = T(0);

    // order omega_z, v, w, omega_y (n0 complete), u (n1, n2 begun), omega_x (n1, n2 complete):
    // at most five fields are held besides the one in transform.  The scheduling barriers keep
    // each field's loads one transform ahead of its pack (the scheduler would otherwise sink them
    // next to their use), and the pack of u (which waits for every lane's loads of u) ahead of the
    // stores of n0 into field 0.
    cpx<T> oz[8], v[8], w[8], oy[8], n0[8];
    Raw<N, T, PR> raw;
    load_raw(fin + 5 * fsb, line, l, raw);
    __builtin_amdgcn_sched_barrier(0);
    pack(raw, l, t, oz);
    load_raw(fin + fsb, line, l, raw);
    __builtin_amdgcn_sched_barrier(0);
    inverse_line<N>(oz, buf, l, t, w2);
    __builtin_amdgcn_sched_barrier(0);
    pack(raw, l, t, v);
    load_raw(fin + 2 * fsb, line, l, raw);
    __builtin_amdgcn_sched_barrier(0);
    inverse_line<N>(v, buf, l, t, w2);
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        red.m[1] = fmax(red.m[1], fmax(fabs(v[c].x), fabs(v[c].y)));
        n0[c] = cpx<T>{-oz[c].x * v[c].x, -oz[c].y * v[c].y};
    }
    __builtin_amdgcn_sched_barrier(0);
    pack(raw, l, t, w);
    load_raw(fin + 4 * fsb, line, l, raw);
    __builtin_amdgcn_sched_barrier(0);
    inverse_line<N>(w, buf, l, t, w2);
#pragma unroll
    for (int c = 0; c < 8; ++c) red.m[2] = fmax(red.m[2], fmax(fabs(w[c].x), fabs(w[c].y)));
    __builtin_amdgcn_sched_barrier(0);
    pack(raw, l, t, oy);
    load_raw(fin, line, l, raw);   // u: in registers (pack) before n0 overwrites field 0
    __builtin_amdgcn_sched_barrier(0);
    inverse_line<N>(oy, buf, l, t, w2);
#pragma unroll
    for (int c = 0; c < 8; ++c) n0[c] = cpx<T>{n0[c].x + oy[c].x * w[c].x, n0[c].y + oy[c].y * w[c].y};
    cpx<T> u[8];
    __builtin_amdgcn_sched_barrier(0);
    pack(raw, l, t, u);
    __builtin_amdgcn_sched_barrier(0);
    forward_line<N>(n0, buf, l, t, w2);
    store_unpacked<N, T, PR>(n0, buf, fout, line, l, t);
    load_raw(fin + 3 * fsb, line, l, raw);
    __builtin_amdgcn_sched_barrier(0);
    inverse_line<N>(u, buf, l, t, w2);
    cpx<T> n1[8], n2[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        red.m[0] = fmax(red.m[0], fmax(fabs(u[c].x), fabs(u[c].y)));
        n1[c] = cpx<T>{oz[c].x * u[c].x, oz[c].y * u[c].y};
        n2[c] = cpx<T>{-oy[c].x * u[c].x, -oy[c].y * u[c].y};
    }
    cpx<T> ox[8];
    __builtin_amdgcn_sched_barrier(0);
    pack(raw, l, t, ox);
    inverse_line<N>(ox, buf, l, t, w2);
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        n1[c] = cpx<T>{n1[c].x - ox[c].x * w[c].x, n1[c].y - ox[c].y * w[c].y};
        n2[c] = cpx<T>{n2[c].x + ox[c].x * v[c].x, n2[c].y + ox[c].y * v[c].y};
    }
    forward_line<N>(n1, buf, l, t, w2);
    store_unpacked<N, T, PR>(n1, buf, fout + fsb, line, l, t);
    forward_line<N>(n2, buf, l, t, w2);
    store_unpacked<N, T, PR>(n2, buf, fout + 2 * fsb, line, l, t);
    if (red_bits) block_max_atomic<3>(red, red_bits, false, wsm);
}

}  // namespace xreg
}  // namespace dnsbox
