// The cross-product x pass (rhs.f90:30-65: six c2r, (omega x u) / N^3 ... scaled inputs, three r2c)
// with the x-lines held in registers: 16 lanes per line, 8 complex values per lane and field,
// 4 lines per wavefront, each wavefront on its own lines with its own LDS exchange buffer (no
// block barrier).
//
// A real line of N = 256 samples is packed as M = 128 complex values z[m] = x[2m] + i x[2m+1].
// Its complex transform of length 128 runs as three register-local stages with two exchanges
// through the wavefront's LDS buffer (digits k = p + 8q + 32r, m = a + 4b + 16c):
//   inverse  x[a+4b+16c] = sum_p w8^{cp} w^{p(a+4b)} sum_q w4^{bq} w^{8aq} sum_r w4^{ar} Z[p+8q+32r]
//            stage 1: DFT4 over r (lane p + 8h holds q = h, h+2), exchange, stage 2: twiddle,
//            DFT4 over q (lane p + 8h takes a = h, h+2), exchange, stage 3: twiddle, DFT8 over
//            p (lane a + 4b) -> lane l holds x[l + 16c], c = 0..7
//   forward  the same stages in reverse order with conjugate twiddles, from x[l + 16c] (so the
//            pointwise product and the first forward stage run on registers) to Y[k] at
//            k = l + 16s, s = 0..7, the positions the inverse read Z from.
// The exchange layouts (positions in complex elements of the line's buffer)
//   L1 (a, q, p) -> 40a + 8q + p                   between stages 1 and 2
//   L3 (r, c)    -> 17(r >> 1) + 8(r & 1) + c      between stages 2 and 3 (r = a + 4b, c = p)
// give every 16-lane access 16 distinct positions mod 16, i.e. 16 distinct pairs of the 32 4-byte
// LDS banks (no bank conflict), and keep each lane's addresses a per-lane base plus immediate
// offsets.  The c2r/r2c and stage-3 twiddles come from a per-block LDS table; the six stage-2
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
    static constexpr bool enabled = N == 256 && std::is_same<T, float>::value;
};

namespace xreg {

constexpr int kLanes = 16;            // lanes per x-line
constexpr int kLinesPerWave = 4;
constexpr int kWaves = 4;             // wavefronts per block
constexpr int kLinesPerBlock = kLinesPerWave * kWaves;
constexpr int kThreads = 64 * kWaves;
constexpr int kM = 128;               // complex values per packed line
constexpr int kRow = 152;             // line buffer (complex elements): L1 spans 0..151

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

// per-block twiddle table (LDS), from the table tw[j] = e^{-2 pi i j / 256}:
//   wp[k]      = e^{+2 pi i k / 256}, k = 0..127  (c2r pre-processing; the r2c uses the conjugate)
//   w3[p-1][l] = w^{p l}, p = 1..7, l = 0..15     (w = e^{+2 pi i / 128})
//   w2[a][q-1] = w^{8 a q}, a = 0..3, q = 1..3
// The lanes of a line read wp[l + 16s] and w3[.][l] at 16 distinct banks and w2 at two.
template <class T>
struct TwL {
    cpx<T> wp[128];
    cpx<T> w3[7][16];
    cpx<T> w2[4][3];
};

template <class T>
__device__ __forceinline__ void fill_tw(const cpx<T>* __restrict__ tw, TwL<T>& t, int tid)
{
    auto cj = [&](int j) { const cpx<T> v = tw[j & 255]; return cpx<T>{v.x, -v.y}; };
    if (tid < 128) {
        t.wp[tid] = cj(tid);
    } else if (tid < 240) {
        const int p = (tid - 128) / 16 + 1, l = (tid - 128) % 16;
        t.w3[p - 1][l] = cj(2 * ((p * l) & 127));   // w^n = conj tw[2n]
    } else if (tid < 252) {
        const int a = (tid - 240) / 3, q = (tid - 240) % 3 + 1;
        t.w2[a][q - 1] = cj(16 * a * q);
    }
}

// c2r packing of a half-spectrum line X (row, k <= M): z[s] = Z[k], k = l + 16 s,
//   Z[k] = (X[k] + conj X[M-k]) + i e^{+2 pi i k/N} (X[k] - conj X[M-k]),
// the imaginary parts of X[0] and X[M] ignored.  PR: X is zero at k > KMAX (k < M).
template <int N, class T, bool PR>
struct Raw {   // X[k] and X[M-k] of the lane's positions k = l + 16 s, as loaded
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
    constexpr int M = N / 2, K = Keep<N>::KMAX;
    constexpr unsigned E = sizeof(cpx<T>);
    const cpx<T> zero{T(0), T(0)};
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        const int k = l + 16 * s;
        // X[k], k in [16s, 16s + 15]
        if (!PR || 16 * s + 15 <= K) {
            r.A[s] = ldb<T>(base + 16 * E * s, line + unsigned(l) * E);
        } else if (16 * s > K) {
            r.A[s] = zero;
        } else {
            r.A[s] = ldb<T>(base, line + unsigned(min(k, K)) * E);   // zeroed in pack
        }
        // X[M - k], M - k in [M - 16s - 15, M - 16s]; kept: M - k <= K, or k = 0 (Nyquist)
        if (!PR || (s > 0 && M - 16 * s <= K)) {
            r.B[s] = ldb<T>(base - 16 * E * s, line + unsigned(M - l) * E);
        } else if (s > 0 && M - 16 * s - 15 > K) {
            r.B[s] = zero;
        } else {
            const bool kb = M - k <= K || k == 0;
            r.B[s] = ldb<T>(base, line + unsigned(kb ? M - k : M) * E);   // zeroed in pack
        }
    }
}

template <int N, class T, bool PR>
__device__ __forceinline__ void pack(const Raw<N, T, PR>& r, int l, const TwL<T>& t, cpx<T> (&z)[8])
{
    constexpr int M = N / 2, K = Keep<N>::KMAX;
    const cpx<T> zero{T(0), T(0)};
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        const int k = l + 16 * s;
        cpx<T> a = r.A[s], b = conj(r.B[s]);
        // the clamped loads of load_raw (lane-dependent keep) take zero here, where the data is used
        if (PR && !(16 * s + 15 <= K) && !(16 * s > K)) a = k <= K ? a : zero;
        if (PR && !(s > 0 && M - 16 * s <= K) && !(s > 0 && M - 16 * s - 15 > K)) b = (M - k <= K || k == 0) ? b : zero;
        if (s == 0 && l == 0) {
            a.y = T(0);
            b.y = T(0);
        }
        const cpx<T> fe = a + b, fo = cmul(a - b, t.wp[l + 16 * s]);
        z[s] = fe + mul_i<1>(fo);
    }
}

// inverse complex transform of z[s] = Z[l + 16 s] -> x[c] = x[l + 16 c] (unnormalised, e^{+})
template <class T>
__device__ __forceinline__ void inverse128(cpx<T> (&z)[8], cpx<T>* __restrict__ buf, int l, const TwL<T>& t, const cpx<T> (&w2)[2][3])
{
    const int p = l & 7, h = l >> 3;
    // stage 1: DFT4 over r of Z[p + 8q + 32r], q = h + 2j  (z index s = j + 2r)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        cpx<T> v[4] = {z[j], z[j + 2], z[j + 4], z[j + 6]};
        Dft<4, +1, T>::run(v);
#pragma unroll
        for (int a = 0; a < 4; ++a) buf[l1(a, h + 2 * j, p)] = v[a];
    }
    wave_lds_sync();
    // stage 2: lane (p, h) takes a = h + 2ai; twiddle w^{8aq}, DFT4 over q
    cpx<T> u2[2][4];
#pragma unroll
    for (int ai = 0; ai < 2; ++ai)
#pragma unroll
        for (int q = 0; q < 4; ++q) u2[ai][q] = buf[l1(h + 2 * ai, q, p)];
    wave_lds_sync();
#pragma unroll
    for (int ai = 0; ai < 2; ++ai) {
#pragma unroll
        for (int q = 1; q < 4; ++q) u2[ai][q] = cmul(u2[ai][q], w2[ai][q - 1]);
        Dft<4, +1, T>::run(u2[ai]);
#pragma unroll
        for (int b = 0; b < 4; ++b) buf[l3(h + 2 * ai + 4 * b, p)] = u2[ai][b];
    }
    wave_lds_sync();
    // stage 3: lane l = a + 4b; twiddle w^{p l}, DFT8 over p -> x[l + 16 c]
    z[0] = buf[l3(l, 0)];
#pragma unroll
    for (int q = 1; q < 8; ++q) z[q] = cmul(buf[l3(l, q)], t.w3[q - 1][l]);
    wave_lds_sync();
    Dft<8, +1, T>::run(z);
}

// forward complex transform of y[c] = y[l + 16 c] -> y[s] = Y[l + 16 s] (unnormalised, e^{-})
template <class T>
__device__ __forceinline__ void forward128(cpx<T> (&y)[8], cpx<T>* __restrict__ buf, int l, const TwL<T>& t, const cpx<T> (&w2)[2][3])
{
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
        for (int b = 0; b < 4; ++b) u2[ai][b] = buf[l3(h + 2 * ai + 4 * b, p)];
    wave_lds_sync();
#pragma unroll
    for (int ai = 0; ai < 2; ++ai) {
        Dft<4, -1, T>::run(u2[ai]);
        buf[l1(h + 2 * ai, 0, p)] = u2[ai][0];
#pragma unroll
        for (int q = 1; q < 4; ++q) buf[l1(h + 2 * ai, q, p)] = cmulc(u2[ai][q], w2[ai][q - 1]);
    }
    wave_lds_sync();
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        cpx<T> v[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) v[a] = buf[l1(a, h + 2 * j, p)];
        Dft<4, -1, T>::run(v);
#pragma unroll
        for (int r = 0; r < 4; ++r) y[j + 2 * r] = v[r];
    }
    wave_lds_sync();
}

// r2c post-processing of Y[k] (k = l + 16 s in y[s]) and the store of the kept modes:
//   X[k] = 1/2 (Y[k] + conj Y[M-k]) - 1/2 i e^{-2 pi i k/N} (Y[k] - conj Y[M-k]),  k <= M, Y[M] = Y[0]
template <int N, class T, bool PO>
__device__ __forceinline__ void store_unpacked(cpx<T> (&y)[8], cpx<T>* __restrict__ buf, char* base, unsigned line,
                                               int l, const TwL<T>& t)
{
    constexpr int M = N / 2;
    constexpr int SMAX = PO ? (Keep<N>::KMAX / 16 + 1) : 8;   // s with l + 16 s possibly kept
#pragma unroll
    for (int s = 0; s < 8; ++s) buf[l + 16 * s] = y[s];
    wave_lds_sync();
    cpx<T> zm[SMAX];
#pragma unroll
    for (int s = 0; s < SMAX; ++s) zm[s] = conj(buf[(M - (l + 16 * s)) & (M - 1)]);
    wave_lds_sync();
#pragma unroll
    for (int s = 0; s < SMAX; ++s) {
        const int k = l + 16 * s;
        if (!PO || k <= Keep<N>::KMAX) {
            const cpx<T> a = T(0.5) * (y[s] + zm[s]);
            const cpx<T> b = cmulc(T(0.5) * (y[s] - zm[s]), t.wp[k]);
            stb<T>(base + 16 * sizeof(cpx<T>) * s, line + unsigned(l) * unsigned(sizeof(cpx<T>)), a + mul_i<-1>(b));
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
    static_assert(N / 2 == kM, "register x pass: N = 256");
    static_assert((long long)N * N * (N / 2 + 1) * sizeof(cpx<T>) < (1LL << 32), "32-bit offsets within a field");
    constexpr int H = N / 2 + 1;
    __shared__ cpx<T> xbuf[kWaves * kLinesPerWave * kRow];
    __shared__ TwL<T> t;
    __shared__ double wsm[4 * 16];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int li = lane >> 4, l = lane & 15;
    const long long row = (long long)xcd_contiguous(blockIdx.x, gridDim.x) * kLinesPerBlock + wave * kLinesPerWave + li;
    cpx<T>* buf = xbuf + (wave * kLinesPerWave + li) * kRow;
    fill_tw(tw_g, t, tid);
    __syncthreads();
    cpx<T> w2[2][3];
#pragma unroll
    for (int ai = 0; ai < 2; ++ai)
#pragma unroll
        for (int q = 0; q < 3; ++q) w2[ai][q] = t.w2[(l >> 3) + 2 * ai][q];
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
    inverse128(oz, buf, l, t, w2);
    __builtin_amdgcn_sched_barrier(0);
    pack(raw, l, t, v);
    load_raw(fin + 2 * fsb, line, l, raw);
    __builtin_amdgcn_sched_barrier(0);
    inverse128(v, buf, l, t, w2);
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        red.m[1] = fmax(red.m[1], fmax(fabs(v[c].x), fabs(v[c].y)));
        n0[c] = cpx<T>{-oz[c].x * v[c].x, -oz[c].y * v[c].y};
    }
    __builtin_amdgcn_sched_barrier(0);
    pack(raw, l, t, w);
    load_raw(fin + 4 * fsb, line, l, raw);
    __builtin_amdgcn_sched_barrier(0);
    inverse128(w, buf, l, t, w2);
#pragma unroll
    for (int c = 0; c < 8; ++c) red.m[2] = fmax(red.m[2], fmax(fabs(w[c].x), fabs(w[c].y)));
    __builtin_amdgcn_sched_barrier(0);
    pack(raw, l, t, oy);
    load_raw(fin, line, l, raw);   // u: in registers (pack) before n0 overwrites field 0
    __builtin_amdgcn_sched_barrier(0);
    inverse128(oy, buf, l, t, w2);
#pragma unroll
    for (int c = 0; c < 8; ++c) n0[c] = cpx<T>{n0[c].x + oy[c].x * w[c].x, n0[c].y + oy[c].y * w[c].y};
    cpx<T> u[8];
    __builtin_amdgcn_sched_barrier(0);
    pack(raw, l, t, u);
    __builtin_amdgcn_sched_barrier(0);
    forward128(n0, buf, l, t, w2);
    store_unpacked<N, T, PR>(n0, buf, fout, line, l, t);
    load_raw(fin + 3 * fsb, line, l, raw);
    __builtin_amdgcn_sched_barrier(0);
    inverse128(u, buf, l, t, w2);
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
    inverse128(ox, buf, l, t, w2);
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        n1[c] = cpx<T>{n1[c].x - ox[c].x * w[c].x, n1[c].y - ox[c].y * w[c].y};
        n2[c] = cpx<T>{n2[c].x + ox[c].x * v[c].x, n2[c].y + ox[c].y * v[c].y};
    }
    forward128(n1, buf, l, t, w2);
    store_unpacked<N, T, PR>(n1, buf, fout + fsb, line, l, t);
    forward128(n2, buf, l, t, w2);
    store_unpacked<N, T, PR>(n2, buf, fout + 2 * fsb, line, l, t);
    if (red_bits) block_max_atomic<3>(red, red_bits, false, wsm);
}

}  // namespace xreg
}  // namespace dnsbox
