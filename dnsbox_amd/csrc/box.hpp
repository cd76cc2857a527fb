// Kernels of one periodic box on gfx950.
//
// Data layout in HBM (per box, per component):
//   spectral  complex (Nz, Ny, Nx/2+1), kx fastest   (numpy/torch rfftn layout of u[z][y][x])
//   physical  real    (Nz, Ny, Nx),     x  fastest   (== Fortran u(Nx,Ny,Nz) of the reference)
// The reference keeps its spectral arrays z-fastest (p3dfft STRIDE1, rhs.f90:68);
// the host boundary converts (dnsbox_capi.hip).
//
// One RHS (rhs.f90:18-135) is five launches, each one pass over HBM:
//   zinv  : z-inverse FFT of (u, v, w) with the spectral curl fused in the
//           prologue -> six half-spectra (u, v, w, wx, wy, wz), scaled by 1/(NX NY NZ)
//   ycol  : y-inverse FFT of the six fields, in place
//   xpass : per x-line: six c2r -> (omega x u) in registers -> three r2c; the
//           physical fields never reach HBM; max|u_i| for the CFL number is
//           reduced on the way
//   ycol  : y-forward FFT of the three fields, in place
//   zfwd  : z-forward FFT + mask + projection + forcing + the stepper update
//           (IF-Heun predictor / corrector, IF-AB2) in the epilogue
#pragma once

#include <type_traits>

#include "fft_core.hpp"

#ifndef DNSBOX_Z_PREFETCH
#define DNSBOX_Z_PREFETCH 1
#endif
#ifndef DNSBOX_XCD_REMAP
#define DNSBOX_XCD_REMAP 1
#endif
#ifndef DNSBOX_X_PRELOAD
#define DNSBOX_X_PRELOAD 1
#endif

namespace dnsbox {

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own L2.  Give
// each XCD a contiguous range of logical tiles, so that tiles whose rows share cache
// lines (the y pass reads 8-byte-aligned segments of x-lines; a z tile narrower than a
// cache line shares its lines with its neighbours) meet in one L2.
// b: linear block id in [0, nb); returns the logical tile id, a bijection on [0, nb).
__device__ __forceinline__ int xcd_contiguous(int b, int nb)
{
#if DNSBOX_XCD_REMAP
    constexpr int NX = 8;
    const int xcd = b % NX, idx = b / NX, per = nb / NX, rem = nb % NX;
    return xcd * per + (xcd < rem ? xcd : rem) + idx;
#else
    (void)nb;
    return b;
#endif
}

// ---------------------------------------------------------------------------
// device-resident per-box scalars (FP64)
// ---------------------------------------------------------------------------
struct BoxScalars {
    double dt;          // step size of the current step
    double dt_prev;     // step size of the previous step (CFL controller input)
    double t;           // simulated time
    double eband;       // band energy of the RHS input (rhs.f90:137-168)
    double courant;     // Courant number of the last controller evaluation
    double h_prev;      // size of the last step taken, t_n - t_{n-1} (IF-AB2 history spacing)
    unsigned long long maxu[3];   // max |u_i| bits (double, >= 0) from the cross-product pass
    unsigned long long divmax;    // ordered bits of the signed max of div u
    long long nstep;
    double pad1[5];
};

// Box lengths: the wavenumber of index i along axis d is alpha_d kval(i) (variables.f90:226-262,
// alpha_d = 2 pi / L_d); kernels take the integer kval and scale it by a[d] in FP64.
struct BoxPhys {
    double nu, Q, Pin, kcut, kolmo_amp;
    double thr[3];      // (N/2)(2 alpha_d/3): cubic 2/3 rule thresholds      (rhs.f90:80-85)
    double sph_thr2;    // ((N/2) 0.5 alpha_x)^2: spherical rule of the RHS   (rhs.f90:75-78)
    double sph_dthr2;   // ((N/2)(2 alpha_x/3))^2: spherical rule of the dealias (state.f90:186-188)
    double a[3];        // alpha_x, alpha_y, alpha_z (their single-precision values, parameters.f90:18-20)
    double cmin, cmax, dtmax;
    int kolmo_k, bandlim, spherical, parity, adaptive, pad;
    int nyq[3];         // Nyquist index N_d / 2 of each axis (parity = 0 removes those modes)
    int pad2;
};

__host__ __device__ __forceinline__ constexpr int kval(int i, int n)
{
    return i < n / 2 ? i : (i == n / 2 ? 0 : i - n);
}

// |k|^2 of the table wavenumbers (kx, ky, kz) scaled by the box's alphas, in FP64
__device__ __forceinline__ double kk2(int kx, int ky, int kz, const BoxPhys& ph)
{
    const double x = double(kx) * ph.a[0], y = double(ky) * ph.a[1], z = double(kz) * ph.a[2];
    return x * x + y * y + z * z;
}

// rhs.f90:74-88: cubic ".ge." 2/3 rule on the table wavenumbers (Nyquist k = 0),
// spherical option, k = 0; parity = 0 also removes the Nyquist-index modes.
__device__ __forceinline__ bool masked(int ix, int iy, int iz, int kx, int ky, int kz, const BoxPhys& ph)
{
    if (ph.spherical && kk2(kx, ky, kz, ph) >= ph.sph_thr2) return true;
    if (fabs(double(kx) * ph.a[0]) >= ph.thr[0] || fabs(double(ky) * ph.a[1]) >= ph.thr[1] ||
        fabs(double(kz) * ph.a[2]) >= ph.thr[2])
        return true;
    if (kx == 0 && ky == 0 && kz == 0) return true;
    if (!ph.parity && (ix == ph.nyq[0] || iy == ph.nyq[1] || iz == ph.nyq[2])) return true;
    return false;
}

// ---------------------------------------------------------------------------
// Transforms pruned to the 2/3-rule cube.  The mask above zeroes every mode whose table
// wavenumber has |k_i| >= N/3 on some axis (the threshold is (N/2)(2/3) <= KMAX + 1).  A state
// with no energy on those modes keeps none through a step -- the masked nonlinear term is zero
// there and the integrating factor only scales -- so the passes of the step need only the kept
// index range of each axis, |kval(i)| <= KMAX (the Nyquist index, table wavenumber 0, is kept);
// Keep<N> describes one axis of N points.
// A pruned pass loads zeros in place of the other modes and does not store them; it skips only
// exact zeros, so its results agree with the unpruned pass to rounding (the two instantiations
// may contract multiply-adds differently).
//   z passes : columns (iy, ix) over Ky x Kx, state modes over Kz
//   y passes : columns (iz, ix) over all iz x Kx; inverse reads ky in Ky, forward stores ky in Ky
//   x pass   : lines over all (y, z); c2r reads kx in Kx, r2c stores kx in Kx
// ---------------------------------------------------------------------------
template <int N>
struct Keep {
    static constexpr int KMAX = (N - 1) / 3;    // largest integer below N/3
    static constexpr int H = N / 2 + 1;
    static constexpr int KX = KMAX + 2;         // kx indices 0..KMAX and N/2
    static constexpr int KY = 2 * KMAX + 2;     // ky (kz) indices 0..KMAX, N-KMAX..N-1 and N/2
    static_assert(KMAX + 1 < N / 2, "kept ranges must not meet the Nyquist index");
    __device__ __forceinline__ static bool kept(int i) { return abs(kval(i, N)) <= KMAX; }
    __device__ __forceinline__ static int ix(int x) { return x <= KMAX ? x : N / 2; }
    __device__ __forceinline__ static int iy(int y)
    {
        return y <= KMAX ? y : (y <= 2 * KMAX ? y + N - 2 * KMAX - 1 : N / 2);
    }
};

// Grid of a box: NX, NY, NZ points along x, y, z (parameters.f90:12-14).  Each pass transforms
// along one axis: the z passes columns of length NZ over the (ky, kx) columns, the y passes
// columns of length NY over the (kz, kx) columns, the x pass lines of length NX over (y, z).
template <int NX_, int NY_, int NZ_>
struct Dims {
    static constexpr int NX = NX_, NY = NY_, NZ = NZ_;
    static constexpr int H = NX / 2 + 1;                      // kx modes of a half-spectrum row
    static constexpr long long FS = (long long)NZ * NY * H;   // elements of a spectral field
    static constexpr long long NP = (long long)NX * NY * NZ;  // points of a physical field
    static constexpr bool CUBIC = NX == NY && NY == NZ;
    static constexpr int YKCOLS = NZ * Keep<NX>::KX;          // columns of the pruned y passes
    // column iy * H + ix of the full layout is a kept (ky, kx) column of the z passes
    __host__ __device__ __forceinline__ static constexpr bool kept_col(int col)
    {
        if (col < 0 || col >= NY * H) return false;
        const int ky = kval(col / H, NY), x = col % H;
        return ky <= Keep<NY>::KMAX && ky >= -Keep<NY>::KMAX && (x <= Keep<NX>::KMAX || x == NX / 2);
    }
};
template <int N>
using Cube = Dims<N, N, N>;

// Tiles of the pruned z passes: the W-column chunks [W j, W j + W) of the full (ky, kx)
// column index that hold a kept column.  A kept row of the z-plane holds a run of KMAX + 1
// columns and its Nyquist column (the row's last, just before the next row's run); rows are
// NX/2 + 1 columns long, so their runs start at every offset within a cache line.  Tiles that
// follow the chunks keep every row of a tile in one aligned segment of W columns (one 128-byte
// line at W = 16 FP32 columns) at the price of idle lanes (a fill of 85 % at 256^3, W = 16).
template <class Dm, int W>
constexpr bool chunk_kept(int j)
{
    bool any = false;
    for (int c = 0; c < W; ++c) any = any || Dm::kept_col(j * W + c);
    return any;
}
template <class Dm, int W>
constexpr int kept_chunk_count()
{
    int n = 0;
    for (int j = 0; j * W < Dm::NY * Dm::H; ++j) n += chunk_kept<Dm, W>(j) ? 1 : 0;
    return n;
}
template <class Dm, int W>
struct KeptChunks {
    static constexpr int NCH = kept_chunk_count<Dm, W>();
    static_assert(Dm::NY * Dm::H / W < 65536, "chunk index fits 16 bits");
    unsigned short j[NCH];
};
template <class Dm, int W>
constexpr KeptChunks<Dm, W> make_kept_chunks()
{
    KeptChunks<Dm, W> t{};
    int n = 0;
    for (int j = 0; j * W < Dm::NY * Dm::H; ++j)
        if (chunk_kept<Dm, W>(j)) t.j[n++] = (unsigned short)j;
    return t;
}
template <class Dm, int W>
__device__ constexpr KeptChunks<Dm, W> kept_chunks = make_kept_chunks<Dm, W>();

// column of lane c in tile `tile` of a z pass (full layout index iy * H + ix) and whether it
// is one the pass transforms; an idle lane gets its chunk's first column (in range)
template <class Dm, int W, bool PR>
__device__ __forceinline__ int ztile_col(int tile, int c, bool& valid)
{
    if constexpr (PR) {
        const int base = W * int(kept_chunks<Dm, W>.j[tile]);
        valid = Dm::kept_col(base + c);
        return valid ? base + c : base;
    } else {
        const int col = tile * W + c;
        valid = col < Dm::NY * Dm::H;
        return col;
    }
}

__device__ __forceinline__ unsigned long long dbl_bits(double v)
{
    return (unsigned long long)__double_as_longlong(v);
}
__device__ __forceinline__ unsigned long long ordered_bits(double v)
{
    unsigned long long b = dbl_bits(v);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// ---------------------------------------------------------------------------
// per-size configuration
//   ZP, ZW : radices and tile width (columns) of the z passes.  ZP is symmetric
//            (first radix == last radix) so that a thread's output positions of the
//            forward transform are its input positions of the inverse one: the
//            stepper epilogue hands its result to the next z-inverse in registers.
//   YP, YW : the same for the y passes
//   RP     : radices of the length-N/2 complex row transforms (x; real pairs packed)
//   RB     : x-lines per block in the x pass, NTX its threads
//   XPD    : x-pass LDS rows hold element j at j + j/XPD (0: at j)
//   XPAD   : extra complex elements per x-pass LDS row
// The x-pass padding comes from a bank model of the pass's LDS accesses.
// ---------------------------------------------------------------------------
template <int ZW_, int YW_, int RB_, int NTX_, class ZP_, class YP_, class RP_, int XPD_ = 0, int XPAD_ = 0>
struct CfgBase {
    using ZP = ZP_;
    using YP = YP_;
    using RP = RP_;
    static constexpr int ZW = ZW_;
    static constexpr int YW = YW_;
    static constexpr int RB = RB_;
    static constexpr int NTX = NTX_;
    static constexpr int XPD = XPD_;
    static constexpr int XPAD = XPAD_;
};
// CfgDefault<N, T>: the table chosen by the autotuner (tuned_cfg.hpp).  A tuning build
// (-DDNSBOX_TUNE_N=..., -DDNSBOX_TUNE_T=float|double, -DDNSBOX_TUNE_CFG=CfgBase<...>)
// replaces the entry of one size and precision with a candidate.
template <int N, class T>
struct CfgDefault;
}  // namespace dnsbox
#include "tuned_cfg.hpp"
namespace dnsbox {
#ifdef DNSBOX_TUNE_CFG
template <int N, class T>
struct Cfg : std::conditional_t<(N == DNSBOX_TUNE_N && std::is_same<T, DNSBOX_TUNE_T>::value), DNSBOX_TUNE_CFG,
                                CfgDefault<N, T>> {};
#else
template <int N, class T>
struct Cfg : CfgDefault<N, T> {};
#endif

// Column transforms of the z passes taken ZIL at a time, stage by stage (ColFFT::run_multi),
// where the z passes are bound by the latency of their barrier chains rather than by
// bandwidth: the 64^3 FP64 box (the adjoint workload; fewer z tiles than CUs).  ZIL LDS tiles
// per block.
template <class Dm, class T>
struct ZInterleave {
    static constexpr int value = (Dm::CUBIC && Dm::NZ == 64 && std::is_same<T, double>::value) ? 6 : 1;
};

// Launch geometry of the passes of a box of grid Dm: each pass takes the configuration of its
// own transform length (Cfg<NZ> for the z passes, Cfg<NY> for the y passes, Cfg<NX> for the x pass)
template <class Dm, class T>
struct Geo {
    using CZ = Cfg<Dm::NZ, T>;
    using CY = Cfg<Dm::NY, T>;
    using CX = Cfg<Dm::NX, T>;
    using CFZ_I = ColFFT<typename CZ::ZP, CZ::ZW, CZ::ZW, +1, T>;
    using CFZ_F = ColFFT<typename CZ::ZP, CZ::ZW, CZ::ZW, -1, T>;
    static constexpr int NX = Dm::NX, NY = Dm::NY, NZ = Dm::NZ;
    static constexpr int H = Dm::H;                     // kx modes
    static constexpr long long FS = Dm::FS;             // elements per spectral field
    static constexpr int ZW = CZ::ZW, YW = CY::YW;      // columns per tile of the z / y passes
    static constexpr int ZNT = CFZ_I::NT;               // threads of a z pass
    static constexpr int ZRM = CFZ_I::RM;
    static constexpr int ZIL = ZInterleave<Dm, T>::value;
    static constexpr int ZLDS = NZ * ZW * ZIL;          // LDS (complex elements) of a z pass
    static constexpr int YNT = ColFFT<typename CY::YP, YW, YW, +1, T>::NT;
    static constexpr int NTX = CX::NTX;                 // threads of the x pass
    static constexpr int ZCOLS = NY * H;                // z-pass columns (ky, kx)
    static constexpr int YCOLS = NZ * H;                // y-pass columns (kz, kx)
    static constexpr int ZTILES = (ZCOLS + ZW - 1) / ZW;
    static constexpr int YTILES = (YCOLS + YW - 1) / YW;
    static constexpr int XROWS = NY * NZ;
    static constexpr int XBLOCKS = XROWS / CX::RB;
    static_assert(XROWS % CX::RB == 0, "x-lines per block must divide NY NZ");
    static_assert(CZ::ZP::L == NZ && CY::YP::L == NY && CX::RP::L == NX / 2, "plan lengths");
    static_assert(CZ::ZP::radix(0) == CZ::ZP::radix(CZ::ZP::S - 1), "z plans must be symmetric");
};

// ===========================================================================
// z-inverse pass with a fused spectral prologue
// ===========================================================================
// table wavenumbers and indices of one mode
struct ModeK {
    int kx, ky, kz, ix, iy, iz;
};

// Prologue functors: FI input fields -> FO output fields, mode by mode.
// the scaled wavenumbers of a mode in the transform type
template <class T>
struct KScale {
    double a[3];
    __device__ __forceinline__ T x(const ModeK& m) const { return T(double(m.kx) * a[0]); }
    __device__ __forceinline__ T y(const ModeK& m) const { return T(double(m.ky) * a[1]); }
    __device__ __forceinline__ T z(const ModeK& m) const { return T(double(m.kz) * a[2]); }
    __device__ __forceinline__ T d(const ModeK& m, int j) const { return j == 0 ? x(m) : (j == 1 ? y(m) : z(m)); }
    __host__ __device__ static KScale of(const BoxPhys& ph) { return KScale{{ph.a[0], ph.a[1], ph.a[2]}}; }
};

template <class T>
struct ProCurl {  // (u, v, w) -> s (u, v, w, i k x u)      state_derivatives + state_vorticity
    static constexpr int FI = 3, FO = 6;
    T s;
    KScale<T> k;
    __device__ __forceinline__ cpx<T> operator()(int o, const cpx<T>* in, const ModeK& m) const
    {
        const T kx = k.x(m), ky = k.y(m), kz = k.z(m);
        const cpx<T> u = in[0], v = in[1], w = in[2];
        switch (o) {
            case 0: return s * u;
            case 1: return s * v;
            case 2: return s * w;
            case 3: return s * mul_i<1>(ky * w - kz * v);  // wy - vz
            case 4: return s * mul_i<1>(kz * u - kx * w);  // uz - wx
            default: return s * mul_i<1>(kx * v - ky * u); // vx - uy
        }
    }
};
template <class T>
struct ProVelDiv {  // (u, v, w) -> s (u, v, w, i k . u)       io_Courant + state_divergence
    static constexpr int FI = 3, FO = 4;
    T s;
    KScale<T> k;
    __device__ __forceinline__ cpx<T> operator()(int o, const cpx<T>* in, const ModeK& m) const
    {
        if (o < 3) return s * in[o];
        return s * mul_i<1>(k.x(m) * in[0] + k.y(m) * in[1] + k.z(m) * in[2]);
    }
};
template <class T>
struct ProScale {  // (u, v, w) -> s (u, v, w)               state_uhat2u
    static constexpr int FI = 3, FO = 3;
    T s;
    __device__ __forceinline__ cpx<T> operator()(int o, const cpx<T>* in, const ModeK&) const
    {
        return s * in[o];
    }
};

// outputs OB .. OB + K - 1 of zinv_from_regs transformed together (K = min(ZIL, FO - OB)), then
// the next group
template <class Dm, class T, class Pro, int OB>
__device__ __forceinline__ void zinv_groups(const cpx<T> (&U)[Pro::FI][Geo<Dm, T>::ZRM],
                                            cpx<T>* const* __restrict__ out, int col, bool valid, cpx<T>* lds,
                                            const cpx<T>* __restrict__ tw, const Pro& pro, int c, int g, int kx,
                                            int ky, int ix, int iy)
{
    using G = Geo<Dm, T>;
    using CF = typename G::CFZ_I;
    constexpr int RM = G::ZRM, H = G::H;
    constexpr long long ZS = (long long)G::NY * H;
    if constexpr (OB < Pro::FO) {
        constexpr int K = (Pro::FO - OB) < G::ZIL ? (Pro::FO - OB) : G::ZIL;
        cpx<T> a[K][RM];
#pragma unroll
        for (int o = 0; o < K; ++o)
#pragma unroll
            for (int q = 0; q < RM; ++q) {
                cpx<T> v[Pro::FI];
#pragma unroll
                for (int f = 0; f < Pro::FI; ++f) v[f] = U[f][q];
                const int iz = CF::in_pos(g, q);
                a[o][q] = pro(OB + o, v, ModeK{kx, ky, kval(iz, G::NZ), ix, iy, iz});
            }
        CF::template run_multi<K>(a, lds, c, g, tw, G::NZ);
        if (valid) {
#pragma unroll
            for (int o = 0; o < K; ++o)
#pragma unroll
                for (int q = 0; q < RM; ++q) out[OB + o][CF::out_pos(g, q) * ZS + col] = a[o][q];
        }
        zinv_groups<Dm, T, Pro, OB + K>(U, out, col, valid, lds, tw, pro, c, g, kx, ky, ix, iy);
    }
}

// z-inverse of the FO prologue outputs, starting from the FI inputs held in
// registers at the stage-0 positions of the z plan: U[f][q] = in_f(kz = in_pos(g, q)).
template <class Dm, class T, class Pro>
__device__ __forceinline__ void zinv_from_regs(const cpx<T> (&U)[Pro::FI][Geo<Dm, T>::ZRM],
                                               cpx<T>* const* __restrict__ out, int col, bool valid,
                                               cpx<T>* lds, const cpx<T>* __restrict__ tw, const Pro& pro)
{
    using G = Geo<Dm, T>;
    using CF = typename G::CFZ_I;
    constexpr int W = G::ZW, RM = G::ZRM, H = G::H;
    constexpr long long ZS = (long long)G::NY * H;  // stride between kz planes
    const int tid = threadIdx.x;
    const int c = tid % W, g = tid / W;
    const int ix = col % H, iy = col / H;
    const int kx = kval(ix, G::NX), ky = kval(iy, G::NY);
    if constexpr (G::ZIL > 1) {
        zinv_groups<Dm, T, Pro, 0>(U, out, col, valid, lds, tw, pro, c, g, kx, ky, ix, iy);
        return;
    }
#pragma unroll
    for (int o = 0; o < Pro::FO; ++o) {
        cpx<T> a[RM];
#pragma unroll
        for (int q = 0; q < RM; ++q) {
            cpx<T> v[Pro::FI];
#pragma unroll
            for (int f = 0; f < Pro::FI; ++f) v[f] = U[f][q];
            const int iz = CF::in_pos(g, q);
            a[q] = pro(o, v, ModeK{kx, ky, kval(iz, G::NZ), ix, iy, iz});
        }
        CF::run(a, lds, c, g, tw, G::NZ);
        if (valid) {
#pragma unroll
            for (int q = 0; q < RM; ++q) out[o][CF::out_pos(g, q) * ZS + col] = a[q];
        }
    }
}

// PR: pruned to the 2/3-rule cube (Keep): tile columns over Ky x Kx, inputs read over Kz
template <class Dm, class T, class Pro, class TI = T, bool PR = false>
__device__ __forceinline__ void zinv_tile(const cpx<TI>* const* __restrict__ in,
                                          cpx<T>* const* __restrict__ out, int tile,
                                          cpx<T>* lds, const cpx<T>* __restrict__ tw, const Pro& pro)
{
    using G = Geo<Dm, T>;
    using CF = typename G::CFZ_I;
    constexpr int W = G::ZW, RM = G::ZRM, H = G::H;
    constexpr long long ZS = (long long)G::NY * H;
    const int tid = threadIdx.x;
    const int c = tid % W, g = tid / W;
    bool valid;
    const int col = ztile_col<Dm, W, PR>(tile, c, valid);
    cpx<T> U[Pro::FI][RM];
#pragma unroll
    for (int f = 0; f < Pro::FI; ++f)
#pragma unroll
        for (int q = 0; q < RM; ++q) {
            const int iz = CF::in_pos(g, q);
            U[f][q] = valid && (!PR || Keep<G::NZ>::kept(iz)) ? cvt<T>(in[f][iz * ZS + col]) : cpx<T>{T(0), T(0)};
        }
    zinv_from_regs<Dm, T, Pro>(U, out, col, valid, lds, tw, pro);
}

// tiles of the z passes
template <class Dm, class T, bool PR>
constexpr int ztiles()
{
    if constexpr (PR) return KeptChunks<Dm, Geo<Dm, T>::ZW>::NCH;
    else return Geo<Dm, T>::ZTILES;
}
// The pruned y passes tile each z-plane's KX columns separately when that keeps at least 85 %
// of the lanes busy (256^3, 512^3); otherwise (a tile wider than a plane's kept columns, as at
// 32^3 and 48^3) the tiles run over the planes' columns back to back.
template <class Dm, class T>
constexpr bool ytiles_by_plane()
{
    constexpr int W = Geo<Dm, T>::YW, TPP = (Keep<Dm::NX>::KX + W - 1) / W;
    return 100 * Keep<Dm::NX>::KX >= 85 * TPP * W;
}
template <class Dm, class T, bool PR>
constexpr int ytiles()
{
    constexpr int W = Geo<Dm, T>::YW;
    if constexpr (PR && ytiles_by_plane<Dm, T>()) return Dm::NZ * ((Keep<Dm::NX>::KX + W - 1) / W);
    else return ((PR ? Dm::YKCOLS : Geo<Dm, T>::YCOLS) + W - 1) / W;
}

template <class Dm, class T, class Pro, class TI = T, bool PR = false>
__global__ void __launch_bounds__((Geo<Dm, T>::ZNT))
    k_zinv(const cpx<TI>* __restrict__ in0, long long in_fs, cpx<T>* __restrict__ out0,
           long long out_fs, const cpx<T>* __restrict__ tw_g, Pro pro)
{
    using G = Geo<Dm, T>;
    __shared__ cpx<T> lds[G::ZLDS];
    __shared__ cpx<T> stw_[G::NZ];
    const cpx<T>* tw = stage_twiddles<G::NZ, T>(tw_g, stw_);
    const cpx<TI>* in[Pro::FI];
    cpx<T>* out[Pro::FO];
#pragma unroll
    for (int f = 0; f < Pro::FI; ++f) in[f] = in0 + f * in_fs;
#pragma unroll
    for (int f = 0; f < Pro::FO; ++f) out[f] = out0 + f * out_fs;
    zinv_tile<Dm, T, Pro, TI, PR>(in, out, xcd_contiguous(blockIdx.x, gridDim.x), lds, tw, pro);
}

// ===========================================================================
// y pass (plain, in place), one field per blockIdx.y
// ===========================================================================
// PR: pruned (Keep): columns over all iz x Kx (tiled plane by plane where ytiles_by_plane); the
// inverse (DIR = +1) reads ky in Ky only, the forward (DIR = -1) stores ky in Ky only.
// OOP: the tile is read from src (same layout) and written to data; otherwise in place.
template <class Dm, class T, int DIR, bool PR = false, bool OOP = false>
__device__ __forceinline__ void ycol_tile(cpx<T>* __restrict__ data, int tile, cpx<T>* lds,
                                          const cpx<T>* __restrict__ tw, const cpx<T>* __restrict__ src = nullptr)
{
    using G = Geo<Dm, T>;
    using CF = ColFFT<typename G::CY::YP, G::YW, G::YW, DIR, T>;
    constexpr int W = G::YW, RM = CF::RM, H = G::H;
    constexpr int CW = PR ? Keep<G::NX>::KX : H;   // columns per z-plane
    const int tid = threadIdx.x;
    const int c = tid % W, g = tid / W;
    int iz, ix;
    bool valid;
    if constexpr (PR && ytiles_by_plane<Dm, T>()) {
        // tiles within one z-plane (no tile spans two planes' column runs)
        constexpr int TPP = (CW + W - 1) / W;   // tiles per z-plane
        const int x = (tile % TPP) * W + c;
        iz = tile / TPP;
        valid = x < CW;
        ix = Keep<G::NX>::ix(x);
    } else {
        const int col = tile * W + c;
        valid = col < (PR ? Dm::YKCOLS : G::YCOLS);
        ix = PR ? Keep<G::NX>::ix(col % CW) : col % H;
        iz = col / CW;
    }
    cpx<T>* base = data + (long long)iz * G::NY * H + ix;
    const cpx<T>* sbase = OOP ? src + (long long)iz * G::NY * H + ix : base;
    cpx<T> a[RM];
#pragma unroll
    for (int q = 0; q < RM; ++q) {
        const int iy = CF::in_pos(g, q);
        a[q] = valid && (!PR || DIR < 0 || Keep<G::NY>::kept(iy)) ? sbase[iy * H] : cpx<T>{T(0), T(0)};
    }
    CF::run(a, lds, c, g, tw, G::NY);
    if (valid) {
#pragma unroll
        for (int q = 0; q < RM; ++q) {
            const int iy = CF::out_pos(g, q);
            if (!PR || DIR > 0 || Keep<G::NY>::kept(iy)) base[iy * H] = a[q];
        }
    }
}

template <class Dm, class T, int DIR, bool PR = false>
__global__ void __launch_bounds__((Geo<Dm, T>::YNT))
    k_ycol(cpx<T>* __restrict__ data0, long long fs, const cpx<T>* __restrict__ tw_g)
{
    using G = Geo<Dm, T>;
    __shared__ cpx<T> lds[G::NY * G::YW];
    __shared__ cpx<T> stw_[G::NY];
    const cpx<T>* tw = stage_twiddles<G::NY, T>(tw_g, stw_);
    const int L = xcd_contiguous(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
    const int field = L / gridDim.x, tile = L % gridDim.x;
    ycol_tile<Dm, T, DIR, PR>(data0 + field * fs, tile, lds, tw);
}

// k_ycol from src0 into dst0 (fields fs apart in both): the y-inverse that also keeps its
// output apart from the z pass's scratch (the step records it for the step adjoint)
template <class Dm, class T, int DIR, bool PR = false>
__global__ void __launch_bounds__((Geo<Dm, T>::YNT))
    k_ycol_oop(const cpx<T>* __restrict__ src0, cpx<T>* __restrict__ dst0, long long fs,
               const cpx<T>* __restrict__ tw_g)
{
    using G = Geo<Dm, T>;
    __shared__ cpx<T> lds[G::NY * G::YW];
    __shared__ cpx<T> stw_[G::NY];
    const cpx<T>* tw = stage_twiddles<G::NY, T>(tw_g, stw_);
    const int L = xcd_contiguous(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
    const int field = L / gridDim.x, tile = L % gridDim.x;
    ycol_tile<Dm, T, DIR, PR, true>(dst0 + field * fs, tile, lds, tw, src0 + field * fs);
}

// ===========================================================================
// x pass: FI inputs (half-spectrum lines -> c2r, or real lines) -> pointwise op
// in physical space -> FO outputs (r2c half-spectrum lines, or real lines)
// ===========================================================================
template <class T>
struct XRed {  // reductions carried by the x pass (maxima: exact in the transform type)
    T m[4];
};

template <class T>
struct OpCross {  // rhs.f90:39-65: (omega x u); io_Courant's max |u_i| on the way
    static constexpr int FI = 6, FO = 3, NRED = 3;
    static constexpr bool IN_SPEC = true, OUT_SPEC = true;
    __device__ __forceinline__ static void apply(const T* in, T* out, XRed<T>& r)
    {
        const T u = in[0], v = in[1], w = in[2], ox = in[3], oy = in[4], oz = in[5];
        out[0] = oy * w - oz * v;
        out[1] = oz * u - ox * w;
        out[2] = ox * v - oy * u;
        r.m[0] = fmax(r.m[0], fabs(u));
        r.m[1] = fmax(r.m[1], fabs(v));
        r.m[2] = fmax(r.m[2], fabs(w));
    }
};
template <class T>
struct OpVelDivMax {  // max|u|, max|v|, max|w|, signed max(div u); no output
    static constexpr int FI = 4, FO = 0, NRED = 4;
    static constexpr bool IN_SPEC = true, OUT_SPEC = true;
    __device__ __forceinline__ static void apply(const T* in, T*, XRed<T>& r)
    {
        r.m[0] = fmax(r.m[0], fabs(in[0]));
        r.m[1] = fmax(r.m[1], fabs(in[1]));
        r.m[2] = fmax(r.m[2], fabs(in[2]));
        r.m[3] = fmax(r.m[3], in[3]);
    }
};
template <class T, int F, bool INS, bool OUTS>
struct OpCopy {
    static constexpr int FI = F, FO = F, NRED = 0;
    static constexpr bool IN_SPEC = INS, OUT_SPEC = OUTS;
    __device__ __forceinline__ static void apply(const T* in, T* out, XRed<T>&)
    {
#pragma unroll
        for (int f = 0; f < F; ++f) out[f] = in[f];
    }
};

template <int N, class T, class Op>
struct XGeo {
    static constexpr int M = N / 2, H = M + 1;
    static constexpr int XPD = Cfg<N, T>::XPD;
    static constexpr int HP = H + (XPD ? H / XPD + 1 : 0) + Cfg<N, T>::XPAD;
    __device__ __forceinline__ static constexpr int px(int j) { return XPD ? j + j / XPD : j; }
    static constexpr int RB = Cfg<N, T>::RB;
    static constexpr int NTX = Cfg<N, T>::NTX;
    static constexpr int FMAX = Op::FI > Op::FO ? Op::FI : Op::FO;
    static constexpr int LDS = FMAX * RB * HP;
};

// The x pass of one block of RB x-lines in three phases:
//   load    : global -> registers (all loads of the block issued together)
//   stage   : registers -> padded LDS rows
//   compute : c2r pre-processing, row FFTs, pointwise op, row FFTs, r2c post-processing
//             and the global stores
// PR: pruned (Keep): the c2r reads kx in Kx only; PO: the r2c stores kx in Kx only (the forward
// step prunes both; the step adjoint, whose outputs fill the whole spectrum, only its inputs)
template <class Dm, class T, class Op, bool PR = false, bool PO = PR>
struct XPass {
    static constexpr int NX = Dm::NX;
    using X = XGeo<NX, T, Op>;
    static constexpr int M = X::M, H = X::H, HP = X::HP, RB = X::RB, NTX = X::NTX;
    static constexpr int FI = Op::FI, FO = Op::FO;
    static constexpr int NL = FI * RB * H;
    static constexpr int QL = (NL + NTX - 1) / NTX;

    __device__ __forceinline__ static void load(const cpx<T>* __restrict__ in0, long long in_fs, int blk,
                                                cpx<T> (&buf)[QL])
    {
        const int tid = threadIdx.x;
        const long long row0 = (long long)blk * RB;
#pragma unroll
        for (int q = 0; q < QL; ++q) {
            // unconditional loads (the tail re-reads element 0 of the block): a
            // conditionally written buf is kept in scratch memory by the compiler
            const int e0 = tid + q * NTX;
            const int e = (NL % NTX == 0 || e0 < NL) ? e0 : 0;
            const int f = e / (RB * H), rk = e % (RB * H);
            buf[q] = in0[f * in_fs + row0 * H + rk];
        }
    }

    __device__ __forceinline__ static void stage(const cpx<T> (&buf)[QL], cpx<T>* lds)
    {
        const int tid = threadIdx.x;
#pragma unroll
        for (int q = 0; q < QL; ++q) {
            const int e = tid + q * NTX;
            if (NL % NTX == 0 || e < NL) {
                const int f = e / (RB * H), rk = e % (RB * H);
                lds[f * RB * HP + (rk / H) * HP + X::px(rk % H)] = buf[q];
            }
        }
    }

    static constexpr int NI = FI * RB * M;
    static constexpr int QZ = (NI + NTX - 1) / NTX;

    // Z[k] = (X[k] + conj X[M-k]) + i e^{+2 pi i k/N} (X[k] - conj X[M-k]),  k < M;
    // imaginary parts of X[0], X[M] are ignored (c2r convention)
    __device__ __forceinline__ static cpx<T> pre(cpx<T> A, cpx<T> Bm, int k, const cpx<T>* __restrict__ tw)
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

    // the c2r pre-processing applied on the way in: both loads of a Z value (X[k] and its
    // mirror X[M-k], coalesced in reverse) issued together, straight from global memory
    // Fixed-k thread layout: when NTX is a multiple of M and RB of NTX/M, the e-th element of
    // a thread (e = tid + q NTX) is column k = tid % M of line row (tid / M) + (NTX/M) q, so
    // k is the same for all q and every load / store offset is a compile-time constant plus a
    // per-thread base (no index arithmetic per element).
    static constexpr bool FIXED_K = NTX % M == 0 && RB % (NTX / M) == 0;
    static constexpr int FPQ = NTX / M;     // lines advanced per q in the fixed-k layout

    __device__ __forceinline__ static void load_pre(const cpx<T>* __restrict__ in0, long long in_fs, int blk,
                                                    const cpx<T>* __restrict__ tw, cpx<T> (&z)[QZ])
    {
        const int tid = threadIdx.x;
        const long long row0 = (long long)blk * RB;
        cpx<T> A[QZ], B[QZ];
        if constexpr (FIXED_K && NI % NTX == 0) {
            // spectral inputs are whole half-spectrum fields, NZ NY (NX/2+1) elements apart
            constexpr long long FS = Dm::FS;
            const int k = tid % M;
            const cpx<T>* base = in0 + (row0 + tid / M) * H;
            const bool ka = !PR || k <= Keep<NX>::KMAX;                 // X[k] kept
            const bool kb = !PR || k == 0 || M - k <= Keep<NX>::KMAX;   // X[M - k] kept
#pragma unroll
            for (int q = 0; q < QZ; ++q) {
                const int fr = FPQ * q;   // line index of this q relative to the thread's first line
                const cpx<T>* row = base + (fr / RB) * FS + (fr % RB) * H;
                A[q] = ka ? row[k] : cpx<T>{T(0), T(0)};
                B[q] = kb ? row[M - k] : cpx<T>{T(0), T(0)};
            }
#pragma unroll
            for (int q = 0; q < QZ; ++q) z[q] = pre(A[q], B[q], k, tw);
            return;
        }
#pragma unroll
        for (int q = 0; q < QZ; ++q) {
            const int e0 = tid + q * NTX;
            const int e = (NI % NTX == 0 || e0 < NI) ? e0 : 0;   // unconditional, as in load()
            const int fr = e / M, k = e % M;
            const cpx<T>* row = in0 + (fr / RB) * in_fs + (row0 + fr % RB) * H;
            if constexpr (PR) {   // X[k], k < M, is kept for k <= KMAX; X[M - k] for M - k <= KMAX or k = 0
                A[q] = k <= Keep<NX>::KMAX ? row[k] : cpx<T>{T(0), T(0)};
                B[q] = (k == 0 || M - k <= Keep<NX>::KMAX) ? row[M - k] : cpx<T>{T(0), T(0)};
            } else {
                A[q] = row[k];
                B[q] = row[M - k];
            }
        }
#pragma unroll
        for (int q = 0; q < QZ; ++q) {
            const int e0 = tid + q * NTX;
            const int e = (NI % NTX == 0 || e0 < NI) ? e0 : 0;
            z[q] = pre(A[q], B[q], e % M, tw);
        }
    }

    __device__ __forceinline__ static void stage_z(const cpx<T> (&z)[QZ], cpx<T>* lds)
    {
        const int tid = threadIdx.x;
#pragma unroll
        for (int q = 0; q < QZ; ++q) {
            const int e = tid + q * NTX;
            if (NI % NTX == 0 || e < NI) lds[(e / M) * HP + X::px(e % M)] = z[q];
        }
    }

    // on entry the LDS holds the staged input rows (after a barrier) — the pre-processed
    // Z rows when Z_STAGED; leaves the LDS in use until its last read (the caller puts a
    // barrier before the next stage)
    template <bool Z_STAGED = false>
    __device__ __forceinline__ static void compute(cpx<T>* __restrict__ out0, long long out_fs, int blk,
                                                   cpx<T>* lds, const cpx<T>* __restrict__ tw, XRed<T>& red)
    {
        const int tid = threadIdx.x;
        const long long row0 = (long long)blk * RB;
        if constexpr (Op::IN_SPEC && Z_STAGED) {
            RowFFT<typename Cfg<NX, T>::RP, FI * RB, HP, NTX, +1, T, X::XPD>::run(lds, tid, tw, NX);
        } else if constexpr (Op::IN_SPEC) {
            // Z[k] = (X[k] + conj X[M-k]) + i e^{+2 pi i k/N} (X[k] - conj X[M-k]),  k < M;
            // imaginary parts of X[0], X[M] are ignored (c2r convention).
            constexpr int NI = FI * RB * M;
            constexpr int Q = (NI + NTX - 1) / NTX;
            cpx<T> z[Q];
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                const int e = tid + q * NTX;
                if (NI % NTX == 0 || e < NI) {
                    const int fr = e / M, k = e % M;
                    const cpx<T>* row = lds + fr * HP;
                    cpx<T> A = row[X::px(k)], B = conj(row[X::px(M - k)]);
                    if (k == 0) {
                        A.y = T(0);
                        B.y = T(0);
                    }
                    cpx<T> w = tw[k];
                    w.y = -w.y;
                    cpx<T> fe = A + B, fo = (A - B) * w;
                    z[q] = fe + mul_i<1>(fo);
                }
            }
            __syncthreads();
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                const int e = tid + q * NTX;
                if (NI % NTX == 0 || e < NI) lds[(e / M) * HP + X::px(e % M)] = z[q];
            }
            __syncthreads();
            RowFFT<typename Cfg<NX, T>::RP, FI * RB, HP, NTX, +1, T, X::XPD>::run(lds, tid, tw, NX);
        }

        // ---- pointwise op on the physical samples x[2m] = Re z[m], x[2m+1] = Im z[m] ----
        for (int e = tid; e < RB * M; e += NTX) {
            const int r = e / M, m = e % M;
            T vr[FI], vi[FI], orr[FO > 0 ? FO : 1], oi[FO > 0 ? FO : 1];
#pragma unroll
            for (int f = 0; f < FI; ++f) {
                cpx<T> v = lds[f * RB * HP + r * HP + X::px(m)];
                vr[f] = v.x;
                vi[f] = v.y;
            }
            Op::apply(vr, orr, red);
            Op::apply(vi, oi, red);
#pragma unroll
            for (int f = 0; f < FO; ++f) lds[f * RB * HP + r * HP + X::px(m)] = cpx<T>{orr[f], oi[f]};
        }
        if constexpr (FO == 0) return;
        __syncthreads();

        // ---- r2c or real store ----
        if constexpr (FO == 0) {
        } else if constexpr (Op::OUT_SPEC) {
            RowFFT<typename Cfg<NX, T>::RP, FO * RB, HP, NTX, -1, T, X::XPD>::run(lds, tid, tw, NX);
            // X[k] = 1/2 (Z[k] + conj Z[M-k]) - 1/2 i e^{-2 pi i k/N} (Z[k] - conj Z[M-k]), k <= M
            if constexpr (FIXED_K && (FO * RB * M) % NTX == 0) {
                // k < M in the fixed-k layout, then k = M (one entry per line and field)
                constexpr long long FS = Dm::FS;
                const int k = tid % M;
                const bool keep = !PO || k <= Keep<NX>::KMAX;
                if (keep) {
                    const cpx<T> w = tw[k];
                    const int r0 = tid / M;
                    cpx<T>* dst = out0 + (row0 + r0) * H + k;
#pragma unroll
                    for (int q = 0; q < FO * RB * M / NTX; ++q) {
                        const int fr = r0 + FPQ * q;   // line within the FO x RB lines of the block
                        const cpx<T>* row = lds + fr * HP;
                        const cpx<T> zk = row[X::px(k)];
                        const cpx<T> zm = conj(row[X::px(k == 0 ? 0 : M - k)]);
                        const cpx<T> a = T(0.5) * (zk + zm);
                        const cpx<T> b = (T(0.5) * (zk - zm)) * w;
                        dst[((FPQ * q) / RB) * FS + ((FPQ * q) % RB) * H] = a + mul_i<-1>(b);
                    }
                }
                for (int i = tid; i < FO * RB; i += NTX) {   // FO RB may exceed the block (16^3 FP32 VJP)
                    const int f = i / RB, r = i % RB;
                    const cpx<T> z0 = lds[(f * RB + r) * HP + X::px(0)];
                    const cpx<T> zm = conj(z0);
                    const cpx<T> a = T(0.5) * (z0 + zm);
                    const cpx<T> b = (T(0.5) * (z0 - zm)) * tw[M];
                    out0[f * FS + (row0 + r) * H + M] = a + mul_i<-1>(b);
                }
                return;
            }
            for (int e = tid; e < FO * RB * H; e += NTX) {
                const int f = e / (RB * H), rk = e % (RB * H);
                const int r = rk / H, k = rk % H;
                const cpx<T>* row = lds + f * RB * HP + r * HP;
                const cpx<T> zk = row[X::px(k == M ? 0 : k)];
                const cpx<T> zm = conj(row[X::px(k == 0 ? 0 : M - k)]);
                const cpx<T> a = T(0.5) * (zk + zm);
                const cpx<T> b = (T(0.5) * (zk - zm)) * tw[k];
                if (!PO || k <= Keep<NX>::KMAX || k == M) out0[f * out_fs + row0 * H + rk] = a + mul_i<-1>(b);
            }
        } else {
            for (int e = tid; e < FO * RB * M; e += NTX) {
                const int f = e / (RB * M), rm = e % (RB * M);
                out0[f * out_fs + row0 * M + rm] = lds[f * RB * HP + (rm / M) * HP + X::px(rm % M)];
            }
        }
    }
};

template <class Dm, class T, class Op, bool PR = false, bool PO = PR>
__device__ __forceinline__ void xpass_tile(const cpx<T>* __restrict__ in0, long long in_fs,
                                           cpx<T>* __restrict__ out0, long long out_fs, int blk, cpx<T>* lds,
                                           const cpx<T>* __restrict__ tw, XRed<T>& red)
{
    using XP = XPass<Dm, T, Op, PR, PO>;
    using X = typename XP::X;
    static_assert(!PR || (Op::IN_SPEC && DNSBOX_X_PRELOAD), "pruned x-pass input: spectral");
    static_assert(!PO || Op::OUT_SPEC, "pruned x-pass output: spectral");
    if constexpr (Op::IN_SPEC && DNSBOX_X_PRELOAD) {
        cpx<T> z[XP::QZ];
        XP::load_pre(in0, in_fs, blk, tw, z);
        XP::stage_z(z, lds);
        __syncthreads();
        XP::template compute<true>(out0, out_fs, blk, lds, tw, red);
        return;
    } else if constexpr (Op::IN_SPEC) {
        cpx<T> buf[XP::QL];
        XP::load(in0, in_fs, blk, buf);
        XP::stage(buf, lds);
    } else {
        constexpr int M = X::M, HP = X::HP, RB = X::RB, NTX = X::NTX;
        const long long row0 = (long long)blk * RB;
        for (int e = threadIdx.x; e < Op::FI * RB * M; e += NTX) {
            const int f = e / (RB * M), rm = e % (RB * M);
            lds[f * RB * HP + (rm / M) * HP + X::px(rm % M)] = in0[f * in_fs + row0 * M + rm];
        }
    }
    __syncthreads();
    XP::compute(out0, out_fs, blk, lds, tw, red);
}

template <int NRED, class T>
__device__ __forceinline__ void block_max_atomic(XRed<T>& red, unsigned long long* dst_bits,
                                                 bool signed_last, double* wsm)
{
    // dst_bits[j]: max of |.| (plain bits) or, for the last entry when signed_last, ordered bits
    const int lane = threadIdx.x % 64, wid = threadIdx.x / 64, nw = blockDim.x / 64;
#pragma unroll
    for (int j = 0; j < NRED; ++j) {
        double v = wave_max(double(red.m[j]));
        if (lane == 0) wsm[j * 16 + wid] = v;
    }
    __syncthreads();
    if (threadIdx.x < NRED) {
        const int j = threadIdx.x;
        double v = wsm[j * 16];
        for (int i = 1; i < nw; ++i) v = fmax(v, wsm[j * 16 + i]);
        if (signed_last && j == NRED - 1) atomicMax(dst_bits + j, ordered_bits(v));
        else atomicMax(dst_bits + j, dbl_bits(v));
    }
}

template <class Dm, class T, class Op, bool PR = false, bool PO = PR>
__global__ void __launch_bounds__((Cfg<Dm::NX, T>::NTX))
    k_xpass(const cpx<T>* __restrict__ in0, long long in_fs, cpx<T>* __restrict__ out0,
            long long out_fs, const cpx<T>* __restrict__ tw_g, unsigned long long* red_bits,
            int signed_last)
{
    using X = XGeo<Dm::NX, T, Op>;
    __shared__ cpx<T> lds[X::LDS];
    __shared__ cpx<T> stw_[Dm::NX];
    const cpx<T>* tw = stage_twiddles<Dm::NX, T>(tw_g, stw_);
    __shared__ double wsm[4 * 16];
    XRed<T> red;
#pragma unroll
    for (int j = 0; j < 4; ++j) red.m[j] = Op::NRED == 4 && j == 3 ? T(-1e300) : T(0);   // -1e300: -inf in FP32
    xpass_tile<Dm, T, Op, PR, PO>(in0, in_fs, out0, out_fs, blockIdx.x, lds, tw, red);
    if constexpr (Op::NRED > 0) {
        if (red_bits) block_max_atomic<Op::NRED>(red, red_bits, signed_last != 0, wsm);
    }
}

// ===========================================================================
// z-forward pass with the stepper epilogue
// ===========================================================================
enum EpiMode { EPI_RHS = 0, EPI_PRED = 1, EPI_CORR = 2, EPI_AB2 = 3, EPI_STORE = 4, EPI_PREDONLY = 5 };

// S is the element type of the state (the FFT type T, or double for an FP32 box that
// keeps its state in FP64); the state update is computed in S.
template <class T, class S = T>
struct EpiArgs {
    cpx<T>* out;          // EPI_RHS: N(u);  EPI_PRED/PREDONLY: predicted state
    cpx<S>* state;        // EPI_PRED/CORR/AB2: the state (updated in place); EPI_STORE: destination
    cpx<T>* nold;         // EPI_AB2: N_{n-1} (read, then overwritten with N_n); EPI_PRED: N_0 if set
    const void* rhs_in;   // RHS input u (band forcing), of type S if rhs_in_s else T
    int rhs_in_s;
    long long fs;         // field stride of out/state/nold/rhs_in
    double dt_fixed;      // EPI_PREDONLY: step size
    BoxScalars* sc;
    BoxPhys ph;
    cpx<S>* tout;         // EPI_CORR: also receives the new state (null: not stored)
};

// dnsboxheun.f90:227-246 set_time_step, triggered as at :153-157
__device__ __forceinline__ double cfl_dt(double c, double dt, const BoxPhys& ph)
{
    if (c > ph.cmax || (c < ph.cmin && dt < ph.dtmax)) {
        dt = ((ph.cmax + ph.cmin) / (2.0 * c)) * dt;
        if (dt > ph.dtmax) dt = ph.dtmax;
    }
    return dt;
}

// io.f90:664-676: sum_d max|u_d| dt / (L_d / N_d), L_d = 2 pi / alpha_d, from the maxima of the
// current RHS's x pass
template <class Dm>
__device__ __forceinline__ double courant_of(const BoxScalars* sc, const BoxPhys& ph, double dt)
{
    constexpr double tp = 6.283185307179586476925286766559;
    const double m0 = __longlong_as_double((long long)sc->maxu[0]), m1 = __longlong_as_double((long long)sc->maxu[1]),
                 m2 = __longlong_as_double((long long)sc->maxu[2]);
    if constexpr (Dm::CUBIC) {
        const double dx = tp / double(Dm::NX);
        return (m0 * ph.a[0] + m1 * ph.a[1] + m2 * ph.a[2]) * dt / dx;
    } else {
        return m0 * dt / ((tp / ph.a[0]) / double(Dm::NX)) + m1 * dt / ((tp / ph.a[1]) / double(Dm::NY)) +
               m2 * dt / ((tp / ph.a[2]) / double(Dm::NZ));
    }
}

template <class Dm>
__device__ __forceinline__ double step_dt(BoxScalars* sc, const BoxPhys& ph)
{
    double dt = sc->dt_prev;
    if (ph.adaptive) dt = cfl_dt(courant_of<Dm>(sc, ph, dt), dt, ph);
    return dt;
}

// fields FB .. FB + K - 1 of zfwd_regs transformed together (K = min(ZIL, F - FB)), then the next group
template <class Dm, class T, int F, int FB>
__device__ __forceinline__ void zfwd_groups(cpx<T>* lds, int c, int g, const cpx<T>* __restrict__ tw,
                                            cpx<T> (&O)[F][Geo<Dm, T>::ZRM])
{
    using G = Geo<Dm, T>;
    if constexpr (FB < F) {
        constexpr int K = (F - FB) < G::ZIL ? (F - FB) : G::ZIL;
        G::CFZ_F::template run_multi<K>(O + FB, lds, c, g, tw, G::NZ);
        zfwd_groups<Dm, T, F, FB + K>(lds, c, g, tw, O);
    }
}

// forward z FFT of F fields of a column tile; O[f][q] holds mode kz = out_pos(g, q)
template <class Dm, class T, int F>
__device__ __forceinline__ void zfwd_regs(const cpx<T>* __restrict__ in0, long long in_fs, int col,
                                          bool valid, cpx<T>* lds, const cpx<T>* __restrict__ tw,
                                          cpx<T> (&O)[F][Geo<Dm, T>::ZRM])
{
    using G = Geo<Dm, T>;
    using CF = typename G::CFZ_F;
    constexpr int W = G::ZW, RM = G::ZRM, H = G::H;
    constexpr long long ZS = (long long)G::NY * H;
    const int tid = threadIdx.x;
    const int c = tid % W, g = tid / W;
    // issue the loads of all F fields before the first transform
#pragma unroll
    for (int f = 0; f < F; ++f) {
        const cpx<T>* src = in0 + f * in_fs;
#pragma unroll
        for (int q = 0; q < RM; ++q) O[f][q] = valid ? src[CF::in_pos(g, q) * ZS + col] : cpx<T>{T(0), T(0)};
    }
    if constexpr (G::ZIL > 1) {
        zfwd_groups<Dm, T, F, 0>(lds, c, g, tw, O);
    } else {
#pragma unroll
        for (int f = 0; f < F; ++f) CF::run(O[f], lds, c, g, tw, G::NZ);
    }
}

// Stepper epilogue on the forward-transformed n_hat = FFT((omega x u)/(NX NY NZ)) of a column
// tile: mask, pressure projection, band / Kolmogorov forcing (rhs.f90:67-133), then
// the update of MODE.  When NEXT is set, U receives the input of the next RHS at the
// same positions (the predicted state for PRED, the new state for CORR).
// PR: pruned (Keep): state (and predicted-state) modes outside Kz are neither read nor written
template <class Dm, class T, int MODE, bool NEXT, bool PREF = false, class S = T, bool PR = false>
__device__ __forceinline__ void zfwd_epilogue(const cpx<T> (&O)[3][Geo<Dm, T>::ZRM], int col, bool valid,
                                              const EpiArgs<T, S>& ea, cpx<T> (&U)[3][Geo<Dm, T>::ZRM],
                                              const cpx<S> (*Sp)[Geo<Dm, T>::ZRM] = nullptr)
{
    using G = Geo<Dm, T>;
    using CF = typename G::CFZ_F;
    constexpr int W = G::ZW, RM = G::ZRM, H = G::H;
    constexpr long long ZS = (long long)G::NY * H;
    const int g = threadIdx.x / W;
    const int ix = col % H, iy = col / H;
    const int kx = kval(ix, G::NX), ky = kval(iy, G::NY);
    const BoxPhys& ph = ea.ph;
    const T kxs = T(double(kx) * ph.a[0]), kys = T(double(ky) * ph.a[1]);
    double dt = 0.0, hp = 0.0;
    if constexpr (MODE == EPI_PRED || MODE == EPI_AB2) dt = step_dt<Dm>(ea.sc, ph);
    if constexpr (MODE == EPI_AB2) hp = ea.sc->h_prev;
    if constexpr (MODE == EPI_CORR) dt = ea.sc->dt;
    if constexpr (MODE == EPI_PREDONLY) dt = ea.dt_fixed;
    const double n3 = double(G::NX) * double(G::NY) * double(G::NZ);
    double fband = 0.0;
    if (ph.bandlim && MODE != EPI_STORE) fband = ph.Pin / (2.0 * ea.sc->eband);

#pragma unroll
    for (int q = 0; q < RM; ++q) {
        const int iz = CF::out_pos(g, q);
        const int kz = kval(iz, G::NZ);
        const T kzs = T(double(kz) * ph.a[2]);
        const double k2d = kk2(kx, ky, kz, ph);
        const long long idx = (long long)iz * ZS + col;
        cpx<T> nl[3];
        if constexpr (MODE == EPI_STORE) {
            if (valid) {
#pragma unroll
                for (int f = 0; f < 3; ++f) ea.state[f * ea.fs + idx] = cvt<S>(O[f][q]);
            }
            continue;
        } else {
            if (!valid || masked(ix, iy, iz, kx, ky, kz, ph)) {
#pragma unroll
                for (int f = 0; f < 3; ++f) nl[f] = cpx<T>{T(0), T(0)};
            } else {
                // phat = i (k . n) / |k|^2 ;  N = -n - (i k) phat = -n + k (k . n)/|k|^2
                const T inv = T(1) / T(k2d);
                const cpx<T> kn = kxs * O[0][q] + kys * O[1][q] + kzs * O[2][q];
                const cpx<T> pk = inv * kn;
                nl[0] = kxs * pk - O[0][q];
                nl[1] = kys * pk - O[1][q];
                nl[2] = kzs * pk - O[2][q];
                if (ph.bandlim && sqrt(k2d) <= ph.kcut) {  // rhs.f90:109-118
#pragma unroll
                    for (int f = 0; f < 3; ++f) {
                        const cpx<T> ui = ea.rhs_in_s ? cvt<T>(static_cast<const cpx<S>*>(ea.rhs_in)[f * ea.fs + idx])
                                                      : static_cast<const cpx<T>*>(ea.rhs_in)[f * ea.fs + idx];
                        nl[f] = nl[f] + T(fband) * ui;
                    }
                }
                if (ph.kolmo_k > 0 && ix == 0 && iz == 0) {  // f = amp sin(kf y) e_x
                    if (iy == ph.kolmo_k) nl[0].y -= T(0.5 * ph.kolmo_amp * n3);
                    if (iy == G::NY - ph.kolmo_k) nl[0].y += T(0.5 * ph.kolmo_amp * n3);
                }
            }
        }
        if constexpr (MODE == EPI_RHS) {
            if (valid) {
#pragma unroll
                for (int f = 0; f < 3; ++f) ea.out[f * ea.fs + idx] = nl[f];
            }
        } else {
            // rhs.f90:8-16 integrating factor, exponent in FP64; the update in the state type S
            const double e64 = exp((-ph.nu * k2d + ph.Q) * dt);
            const S e = S(e64), sdt = S(dt);
            const bool live = valid && (!PR || Keep<G::NZ>::kept(iz));
            if constexpr (MODE == EPI_PRED) {  // dnsboxheun.f90:88-114
#pragma unroll
                for (int f = 0; f < 3; ++f) {
                    const cpx<S> u = PREF ? Sp[f][q] : (live ? ea.state[f * ea.fs + idx] : cpx<S>{S(0), S(0)});
                    const cpx<S> n = cvt<S>(nl[f]);
                    const cpx<T> up = cvt<T>(e * (u + sdt * n));
                    if (NEXT) U[f][q] = up;
                    if (live) {
                        if (ea.out) ea.out[f * ea.fs + idx] = up;
                        ea.state[f * ea.fs + idx] = e * (u + (sdt * S(0.5)) * n);
                        if (ea.nold) ea.nold[f * ea.fs + idx] = nl[f];
                    }
                }
            } else if constexpr (MODE == EPI_PREDONLY) {  // u* of dnsboxheun.f90:90-91, state untouched
#pragma unroll
                for (int f = 0; f < 3; ++f) {
                    if (!valid) continue;
                    const cpx<S> u = ea.state[f * ea.fs + idx];
                    ea.out[f * ea.fs + idx] = cvt<T>(e * (u + sdt * cvt<S>(nl[f])));
                }
            } else if constexpr (MODE == EPI_CORR) {  // dnsboxheun.f90:120-126
#pragma unroll
                for (int f = 0; f < 3; ++f) {
                    const cpx<S> u = PREF ? Sp[f][q] : (live ? ea.state[f * ea.fs + idx] : cpx<S>{S(0), S(0)});
                    const cpx<S> un = u + (sdt * S(0.5)) * cvt<S>(nl[f]);
                    if (NEXT) U[f][q] = cvt<T>(un);
                    if (live) {
                        ea.state[f * ea.fs + idx] = un;
                        if (ea.tout) ea.tout[f * ea.fs + idx] = un;
                    }
                }
            } else if constexpr (MODE == EPI_AB2) {
                // variable-step IF-AB2 (dnsbox.f90:143-193 with consistent weights), h = dt of this
                // step, h' = t_n - t_{n-1}:
                //   u_{n+1} = E_h u_n + h (1 + h/(2h')) E_h N_n - (h^2/(2h')) E_{h+h'} N_{n-1}
                const double r = dt / (2.0 * hp);
                const S e2 = S(exp((-ph.nu * k2d + ph.Q) * (dt + hp)));
                const S b0 = S(dt * (1.0 + r)) * e, b1 = S(dt * r) * e2;
#pragma unroll
                for (int f = 0; f < 3; ++f) {
                    if (!valid) continue;
                    const cpx<S> u = ea.state[f * ea.fs + idx];
                    const cpx<S> no = cvt<S>(ea.nold[f * ea.fs + idx]);
                    ea.state[f * ea.fs + idx] = e * u + b0 * cvt<S>(nl[f]) - b1 * no;
                    ea.nold[f * ea.fs + idx] = nl[f];
                }
            }
        }
    }
}

// the step size of the step being taken (and the Courant number it came from), written by one
// thread of the pass whose epilogue chose it; no block of that pass reads what it writes
template <class Dm>
__device__ __forceinline__ void pred_scalars(const BoxScalars* /*unused*/, BoxScalars* sc, const BoxPhys& ph)
{
    const double dt = step_dt<Dm>(sc, ph);
    if (ph.adaptive) sc->courant = courant_of<Dm>(sc, ph, sc->dt_prev);
    sc->dt = dt;
}

template <class Dm, class T, int MODE, class S = T>
__device__ __forceinline__ void zfwd_tile(const cpx<T>* __restrict__ in0, long long in_fs, int tile, cpx<T>* lds,
                                          const cpx<T>* tw, const EpiArgs<T, S>& ea)
{
    using G = Geo<Dm, T>;
    const int col = tile * G::ZW + threadIdx.x % G::ZW;
    const bool valid = col < G::ZCOLS;
    cpx<T> O[3][G::ZRM];
    zfwd_regs<Dm, T>(in0, in_fs, col, valid, lds, tw, O);
    zfwd_epilogue<Dm, T, MODE, false, false, S>(O, col, valid, ea, O);
    if constexpr (MODE == EPI_PRED || MODE == EPI_AB2) {
        if (tile == 0 && threadIdx.x == 0) pred_scalars<Dm>(nullptr, ea.sc, ea.ph);
    }
}

template <class Dm, class T, int MODE, class S = T>
__global__ void __launch_bounds__((Geo<Dm, T>::ZNT))
    k_zfwd(const cpx<T>* __restrict__ in0, long long in_fs, const cpx<T>* __restrict__ tw_g,
           EpiArgs<T, S> ea)
{
    using G = Geo<Dm, T>;
    __shared__ cpx<T> lds[G::ZLDS];
    __shared__ cpx<T> stw_[G::NZ];
    const cpx<T>* tw = stage_twiddles<G::NZ, T>(tw_g, stw_);
    zfwd_tile<Dm, T, MODE, S>(in0, in_fs, xcd_contiguous(blockIdx.x, gridDim.x), lds, tw, ea);
}

// Forward z pass + stepper epilogue (PRED or CORR) fused with the z-inverse + curl
// of the next RHS: the next RHS input never makes a round trip through HBM.  The
// input tile (fields 0..2 of X) and the output tile (fields 0..5 of X) cover the
// same columns, so the block reads its whole tile before it writes.
// PR: pruned (Keep): tile columns over Ky x Kx, state modes over Kz
template <class Dm, class T, int MODE, class S = T, bool PR = false>
__device__ __forceinline__ void zfwd_zinv_tile(cpx<T>* __restrict__ X, long long fs, int tile, cpx<T>* lds,
                                               const cpx<T>* tw, const EpiArgs<T, S>& ea, const ProCurl<T>& pro)
{
    using G = Geo<Dm, T>;
    bool valid;
    const int col = ztile_col<Dm, G::ZW, PR>(tile, threadIdx.x % G::ZW, valid);
    cpx<T> O[3][G::ZRM];
#if DNSBOX_Z_PREFETCH
    // the state at this thread's modes (symmetric z plan: output positions = input positions),
    // loaded before the transforms so that its latency overlaps them
    cpx<S> Sp[3][G::ZRM];
    {
        using CF = typename G::CFZ_F;
        constexpr long long ZS = (long long)G::NY * G::H;
        const int g = threadIdx.x / G::ZW;
#pragma unroll
        for (int f = 0; f < 3; ++f)
#pragma unroll
            for (int q = 0; q < G::ZRM; ++q) {
                const int iz = CF::out_pos(g, q);
                Sp[f][q] = valid && (!PR || Keep<G::NZ>::kept(iz)) ? ea.state[f * ea.fs + iz * ZS + col]
                                                                 : cpx<S>{S(0), S(0)};
            }
    }
    zfwd_regs<Dm, T>(X, fs, col, valid, lds, tw, O);
    zfwd_epilogue<Dm, T, MODE, true, true, S, PR>(O, col, valid, ea, O, Sp);
#else
    zfwd_regs<Dm, T>(X, fs, col, valid, lds, tw, O);
    zfwd_epilogue<Dm, T, MODE, true, false, S, PR>(O, col, valid, ea, O);
#endif
    cpx<T>* out[6];
#pragma unroll
    for (int f = 0; f < 6; ++f) out[f] = X + f * fs;
    zinv_from_regs<Dm, T, ProCurl<T>>(O, out, col, valid, lds, tw, pro);
    if constexpr (MODE == EPI_PRED) {
        if (tile == 0 && threadIdx.x == 0) pred_scalars<Dm>(nullptr, ea.sc, ea.ph);
    }
}

// end of step bookkeeping: t += dt, dt_prev = h_prev = dt, clear the CFL maxima
__device__ __forceinline__ void step_end(BoxScalars* sc)
{
    sc->t += sc->dt;
    sc->dt_prev = sc->dt;
    sc->h_prev = sc->dt;
    sc->maxu[0] = sc->maxu[1] = sc->maxu[2] = 0ull;
    sc->nstep += 1;
}

// The corrector pass also ends the step (block 0, first thread): no block of that pass reads
// what step_end writes (the corrector reads dt, which step_end only reads), and the next
// reader of those scalars is a later kernel.
template <class Dm, class T, int MODE, class S = T, bool PR = false>
__global__ void __launch_bounds__((Geo<Dm, T>::ZNT))
    k_zfwd_zinv(cpx<T>* __restrict__ X, long long fs, const cpx<T>* __restrict__ tw_g, EpiArgs<T, S> ea,
                ProCurl<T> pro)
{
    using G = Geo<Dm, T>;
    __shared__ cpx<T> lds[G::ZLDS];
    __shared__ cpx<T> stw_[G::NZ];
    const cpx<T>* tw = stage_twiddles<G::NZ, T>(tw_g, stw_);
    zfwd_zinv_tile<Dm, T, MODE, S, PR>(X, fs, xcd_contiguous(blockIdx.x, gridDim.x), lds, tw, ea, pro);
    if constexpr (MODE == EPI_CORR) {
        if (blockIdx.x == 0 && threadIdx.x == 0) step_end(ea.sc);
    }
}

template <int D = 0>
__global__ void k_step_end(BoxScalars* sc)
{
    if (threadIdx.x == 0) step_end(sc);
}

// ===========================================================================
// spectral reductions: Ekin, dissipation, Eband, EZero/EOne/Esqrt2, shell spectrum
// (state.f90:350-454, rhs.f90:137-168; dissipation by Parseval from the spectrum)
// ===========================================================================
enum { DG_EKIN = 0, DG_DISP, DG_EZERO, DG_EONE, DG_ESQRT2, DG_EBAND, DG_NSCAL };

// dynamic LDS of k_spec_reduce: per wave (4 per block) a sum and a compensation per shell
inline size_t spec_reduce_lds(int nspec) { return size_t(2 * 4 * nspec) * sizeof(double); }

// Shell spectrum and scalar sums of a state, independent of scheduling.  Wave w of the grid takes
// the rows (iz, iy) = w, w + W, ... (W waves in the grid), 64 modes of a row at a time.  Along a
// row the shell index is non-decreasing in ix < NX/2 (|k| grows with kx), so the lanes of one shell
// form one run: a segmented sum over the lanes (fixed order) gives each run's total, and the run's
// last lane adds it to the wave's own bin with Kahan compensation; the Nyquist column (ix = NX/2,
// table kx = 0) is added in a second round.  Each block combines its waves' bins in wave order and
// k_spec_finish sums the blocks in block order.  Scalars: per-lane Kahan sums, reduced in a fixed
// order.  256 threads; dynamic LDS spec_reduce_lds(nspec).
template <class Dm, class T>
__global__ void __launch_bounds__(256)
    k_spec_reduce(const cpx<T>* __restrict__ u0, long long fs, BoxPhys ph, double deltak,
                  int nspec, double* __restrict__ partial /* [gridDim.x][DG_NSCAL + nspec] */)
{
    constexpr int H = Dm::H, NX = Dm::NX, NY = Dm::NY, NZ = Dm::NZ;
    constexpr int NW = 4;             // waves per block
    extern __shared__ double bins[];  // [NW][2][nspec]: sums, compensations
    __shared__ double wsm[DG_NSCAL][NW];
    const int lane = threadIdx.x % 64, wid = threadIdx.x / 64;
    double* bs = bins + wid * 2 * nspec;
    double* bc = bs + nspec;
    for (int i = lane; i < nspec; i += 64) bs[i] = bc[i] = 0.0;
    __builtin_amdgcn_wave_barrier();
    double s[DG_NSCAL], cmp[DG_NSCAL];
#pragma unroll
    for (int j = 0; j < DG_NSCAL; ++j) s[j] = cmp[j] = 0.0;
    const double sqrt2 = 1.4142135623730951;
    const int nwave = gridDim.x * NW;
    for (int row = blockIdx.x * NW + wid; row < NY * NZ; row += nwave) {
        const int iy = row % NY, iz = row / NY;
        const int ky = kval(iy, NY), kz = kval(iz, NZ);
        const cpx<T>* base = u0 + (long long)row * H;
        for (int c0 = 0; c0 < H; c0 += 64) {
            const int ix = c0 + lane;
            const bool valid = ix < H;
            int bin = -1;   // -1: no shell of the spectrum
            double e = 0.0;
            if (valid) {
                const int kx = kval(ix, NX);
#pragma unroll
                for (int f = 0; f < 3; ++f) {
                    const cpx<T> a = base[f * fs + ix];
                    e += double(a.x) * double(a.x) + double(a.y) * double(a.y);
                }
                const double k2 = kk2(kx, ky, kz, ph);
                const double absk = sqrt(k2);
                const double wx = (kx == 0) ? 0.5 : 1.0;
                double v[DG_NSCAL];
                v[DG_EKIN] = wx * e;
                v[DG_DISP] = 2.0 * wx * k2 * e;
                v[DG_EZERO] = absk == 0.0 ? e : 0.0;
                v[DG_EONE] = absk == 1.0 ? e : 0.0;
                v[DG_ESQRT2] = (absk - sqrt2 < 1.0e-9) ? e : 0.0;  // state.f90:383 as written
                v[DG_EBAND] = (absk <= ph.kcut) ? wx * e : 0.0;
#pragma unroll
                for (int j = 0; j < DG_NSCAL; ++j) {  // Kahan
                    const double y = v[j] - cmp[j];
                    const double t = s[j] + y;
                    cmp[j] = (t - s[j]) - y;
                    s[j] = t;
                }
                // state.f90:360-365: nearest shell, the single-precision real(int(.)) test as written
                const int iq = int(absk / deltak);
                const int nk = (absk < double(float(iq)) + 0.5 * deltak) ? iq : iq + 1;
                if (nk < nspec && nk != 0) bin = nk - 1;
            }
#pragma unroll
            for (int round = 0; round < 2; ++round) {
                const bool mine = valid && ((ix == NX / 2) == (round == 1));
                // run key: the bin for this round's lanes, a value of its own for every other lane
                const int key = mine ? bin : -2 - lane;
                double v = mine ? e : 0.0;
#pragma unroll
                for (int d = 1; d < 64; d <<= 1) {   // segmented inclusive sum over equal keys
                    const double vu = __shfl_up(v, d, 64);
                    const int ku = __shfl_up(key, d, 64);
                    if (lane >= d && ku == key) v += vu;
                }
                const int kn = __shfl_down(key, 1, 64);
                if (mine && bin >= 0 && (lane == 63 || kn != key)) {   // the last lane of a run
                    const double y = v - bc[bin];
                    const double t = bs[bin] + y;
                    bc[bin] = (t - bs[bin]) - y;
                    bs[bin] = t;
                }
                __builtin_amdgcn_wave_barrier();
            }
        }
    }
#pragma unroll
    for (int j = 0; j < DG_NSCAL; ++j) {
        const double v = wave_sum(s[j] - cmp[j]);
        if (lane == 0) wsm[j][wid] = v;
    }
    __syncthreads();
    double* out = partial + (long long)blockIdx.x * (DG_NSCAL + nspec);
    if (threadIdx.x < DG_NSCAL) {
        const int j = threadIdx.x;
        out[j] = ((wsm[j][0] + wsm[j][1]) + wsm[j][2]) + wsm[j][3];
    }
    for (int i = threadIdx.x; i < nspec; i += blockDim.x) {
        double acc = 0.0, c = 0.0;
#pragma unroll
        for (int w = 0; w < NW; ++w) {
            const double* wb = bins + w * 2 * nspec;
            const double terms[2] = {wb[i], -wb[nspec + i]};
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                const double y = terms[q] - c;
                const double t = acc + y;
                c = (t - acc) - y;
                acc = t;
            }
        }
        out[DG_NSCAL + i] = acc;
    }
}

// sums the per-block partials in a fixed order (Kahan), scales, writes [DG_NSCAL + nspec]
template <int D = 0>
__global__ void k_spec_finish(const double* __restrict__ partial, int nblk, int nslot,
                              double sm2, double nu, double deltak, double* __restrict__ out,
                              BoxScalars* sc)
{
    for (int j = threadIdx.x; j < nslot; j += blockDim.x) {
        double s = 0.0, c = 0.0;
        for (int b = 0; b < nblk; ++b) {
            const double y = partial[(long long)b * nslot + j] - c;
            const double t = s + y;
            c = (t - s) - y;
            s = t;
        }
        double r = s;
        if (j == DG_EKIN || j == DG_EBAND) r = s * sm2;
        else if (j == DG_DISP) r = s * sm2 * nu;
        else if (j >= DG_NSCAL) r = s * sm2 / deltak;
        out[j] = r;
        if (sc && j == DG_EBAND) sc->eband = r;
    }
}

// band energy of a state, one block over the low modes (rhs.f90:137-168).  The
// table wavenumber decides membership, so with the reference k-table the
// Nyquist index (k = 0) counts as low: per axis the candidate indices are
// |k| <= K plus the Nyquist index.
// candidate indices of the band |k| <= kcut along one axis: |kval| <= K (and the Nyquist index)
template <int N>
__device__ __forceinline__ int band_k(double kcut, double a)
{
    int K = int(kcut / a) + 1;   // |kval| a <= kcut needs |kval| <= kcut / a; one more for rounding
    return K > N / 2 - 1 ? N / 2 - 1 : K;
}
template <int N>
__device__ __forceinline__ int band_idx(int b, int K)
{
    return b <= K ? b : (b < 2 * K + 1 ? N - (b - K) : N / 2);
}

template <class Dm, class T>
__global__ void __launch_bounds__(256)
    k_eband(const cpx<T>* __restrict__ u0, long long fs, BoxPhys ph, double sm2, BoxScalars* sc)
{
    constexpr int H = Dm::H, NX = Dm::NX, NY = Dm::NY, NZ = Dm::NZ;
    const double kcut = ph.kcut;
    const int KX = band_k<NX>(kcut, ph.a[0]), KY = band_k<NY>(kcut, ph.a[1]), KZ = band_k<NZ>(kcut, ph.a[2]);
    const int nxc = KX + 2;          // kx indices 0..KX, NX/2
    const int nyc = 2 * KY + 2;      // ky indices 0..KY, NY-KY..NY-1, NY/2
    const int nzc = 2 * KZ + 2;
    const int tot = nxc * nyc * nzc;
    double s = 0.0;
    for (int e = threadIdx.x; e < tot; e += blockDim.x) {
        const int a = e / (nyc * nzc), b = (e / nzc) % nyc, c = e % nzc;
        const int ix = a <= KX ? a : NX / 2;
        const int iy = band_idx<NY>(b, KY);
        const int iz = band_idx<NZ>(c, KZ);
        const int kx = kval(ix, NX), ky = kval(iy, NY), kz = kval(iz, NZ);
        const double absk = sqrt(kk2(kx, ky, kz, ph));
        if (absk > kcut) continue;
        const long long idx = ((long long)iz * NY + iy) * H + ix;
        double en = 0.0;
#pragma unroll
        for (int f = 0; f < 3; ++f) {
            const cpx<T> v = u0[f * fs + idx];
            en += double(v.x) * double(v.x) + double(v.y) * double(v.y);
        }
        s += (kx == 0 ? 0.5 : 1.0) * en;
    }
    __shared__ double wsm[16];
    const double v = wave_sum(s);
    if (threadIdx.x % 64 == 0) wsm[threadIdx.x / 64] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int i = 0; i < int(blockDim.x / 64); ++i) t += wsm[i];
        sc->eband = t * sm2;
    }
}

// ===========================================================================
// seeded solenoidal initial field (same generator as the CPU oracle)
// ===========================================================================
__device__ __forceinline__ unsigned long long splitmix64(unsigned long long x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
__device__ __forceinline__ unsigned long long hash3(unsigned long long seed, unsigned long long box,
                                                    unsigned long long idx)
{
    return splitmix64(seed ^ splitmix64(box ^ splitmix64(idx)));
}

template <class Dm, class T>
__global__ void __launch_bounds__(256)
    k_init_solenoidal(cpx<T>* __restrict__ u0, long long fs, BoxPhys ph, unsigned long long seed,
                      int box, double k0)
{
    constexpr int H = Dm::H, NX = Dm::NX, NY = Dm::NY, NZ = Dm::NZ;
    constexpr long long TOT = Dm::FS;
    const long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (idx >= TOT) return;
    const int ix = int(idx % H);
    const int iy = int((idx / H) % NY);
    const int iz = int(idx / ((long long)H * NY));
    cpx<T> outv[3] = {{T(0), T(0)}, {T(0), T(0)}, {T(0), T(0)}};
    bool keep = !(ix == NX / 2 || iy == NY / 2 || iz == NZ / 2);
    int kx = kval(ix, NX), ky = kval(iy, NY), kz = kval(iz, NZ);
    if (keep && masked(ix, iy, iz, kx, ky, kz, ph)) keep = false;
    if (keep) {
        bool conjugate = false;
        int cy = iy, cz = iz;
        if (ix == 0 && !(ky > 0 || (ky == 0 && kz > 0))) {
            conjugate = true;
            cy = (NY - iy) % NY;
            cz = (NZ - iz) % NZ;
            ky = -ky;
            kz = -kz;
        }
        // canonical index in the oracle's [kx][ky][kz] ordering
        const unsigned long long cidx = ((unsigned long long)ix * NY + cy) * NZ + cz;
        double ar[3], ai[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const unsigned long long h = hash3(seed, (unsigned long long)box, 3 * cidx + c);
            const double phs = 2.0 * 3.14159265358979323846 * (double(h >> 11) * (1.0 / 9007199254740992.0));
            ar[c] = cos(phs);
            ai[c] = sin(phs);
        }
        const double kk[3] = {double(kx) * ph.a[0], double(ky) * ph.a[1], double(kz) * ph.a[2]};
        const double k2 = kk[0] * kk[0] + kk[1] * kk[1] + kk[2] * kk[2];
        const double kdr = (kk[0] * ar[0] + kk[1] * ar[1] + kk[2] * ar[2]) / k2;
        const double kdi = (kk[0] * ai[0] + kk[1] * ai[1] + kk[2] * ai[2]) / k2;
        double na2 = 0.0;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            ar[c] -= kk[c] * kdr;
            ai[c] -= kk[c] * kdi;
            na2 += ar[c] * ar[c] + ai[c] * ai[c];
        }
        const double na = sqrt(na2);
        if (na >= 1e-300) {
            const double kn = sqrt(k2);
            const double ek = kn * kn * kn * kn * exp(-2.0 * (kn / k0) * (kn / k0));
            const double amp = sqrt(ek / (4.0 * 3.14159265358979323846 * k2)) / na;
#pragma unroll
            for (int c = 0; c < 3; ++c)
                outv[c] = cpx<T>{T(ar[c] * amp), T((conjugate ? -ai[c] : ai[c]) * amp)};
        }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) u0[c * fs + idx] = outv[c];
}

// u *= sqrt(target / Ekin), Ekin read from device memory (no host round trip)
template <class T>
__global__ void k_scale_to_energy(cpx<T>* __restrict__ u0, long long n, const double* ekin,
                                  double target)
{
    const double e = *ekin;
    const T s = T(e > 0.0 ? sqrt(target / e) : 0.0);
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        cpx<T> a = u0[i];
        u0[i] = cpx<T>{s * a.x, s * a.y};
    }
}

// flag |= 1 when any of the three fields holds a nonzero mode outside the kept cube (Keep):
// decides whether a step of a caller's buffer may run the pruned passes
// whether this thread's share (start, start + stride, ...) of the modes of the 3 fields at u0
// outside the kept cube holds a nonzero value
template <class Dm, class T>
__device__ __forceinline__ bool outside_cube_nonzero(const cpx<T>* __restrict__ u0, long long fs, long long start,
                                                     long long stride)
{
    constexpr int H = Dm::H;
    constexpr long long TOT = Dm::FS;
    bool nz = false;
    for (long long idx = start; idx < TOT; idx += stride) {
        const int ix = int(idx % H);
        const int iy = int((idx / H) % Dm::NY);
        const int iz = int(idx / ((long long)H * Dm::NY));
        if (Keep<Dm::NX>::kept(ix) && Keep<Dm::NY>::kept(iy) && Keep<Dm::NZ>::kept(iz)) continue;
#pragma unroll
        for (int f = 0; f < 3; ++f) {
            const cpx<T> a = u0[f * fs + idx];
            nz = nz || a.x != T(0) || a.y != T(0);
        }
    }
    return nz;
}

template <class Dm, class T>
__global__ void __launch_bounds__(256)
    k_outside_cube(const cpx<T>* __restrict__ u0, long long fs, unsigned int* flag)
{
    const bool nz = outside_cube_nonzero<Dm, T>(u0, fs, blockIdx.x * (long long)blockDim.x + threadIdx.x,
                                               (long long)gridDim.x * blockDim.x);
    if (__any(nz) && threadIdx.x % 64 == 0) atomicOr(flag, 1u);
}

// state.f90:120-179 state_project / state.f90:181-209 state_dealias, in place
template <class Dm, class T>
__global__ void __launch_bounds__(256)
    k_project_dealias(cpx<T>* __restrict__ u0, long long fs, BoxPhys ph, int do_dealias,
                      int do_project)
{
    constexpr int H = Dm::H, NX = Dm::NX, NY = Dm::NY, NZ = Dm::NZ;
    constexpr long long TOT = Dm::FS;
    const long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (idx >= TOT) return;
    const int ix = int(idx % H);
    const int iy = int((idx / H) % NY);
    const int iz = int(idx / ((long long)H * NY));
    const int kx = kval(ix, NX), ky = kval(iy, NY), kz = kval(iz, NZ);
    cpx<T> u = u0[idx], v = u0[fs + idx], w = u0[2 * fs + idx];
    const double rx = double(kx) * ph.a[0], ry = double(ky) * ph.a[1], rz = double(kz) * ph.a[2];
    const double k2 = rx * rx + ry * ry + rz * rz;
    if (do_project && !(kx == 0 && ky == 0 && kz == 0)) {
        const T inv = T(1) / T(k2);
        const cpx<T> kd = inv * (T(rx) * u + T(ry) * v + T(rz) * w);
        u = u - T(rx) * kd;
        v = v - T(ry) * kd;
        w = w - T(rz) * kd;
    }
    if (do_dealias) {
        // state.f90:184-197: spherical uses ((NX/2)(2 alpha_x/3))^2 here
        bool m = (ph.spherical && k2 >= ph.sph_dthr2) || fabs(rx) >= ph.thr[0] || fabs(ry) >= ph.thr[1] ||
                 fabs(rz) >= ph.thr[2] || (kx == 0 && ky == 0 && kz == 0) ||
                 (!ph.parity && (ix == NX / 2 || iy == NY / 2 || iz == NZ / 2));
        if (m) u = v = w = cpx<T>{T(0), T(0)};
    }
    u0[idx] = u;
    u0[fs + idx] = v;
    u0[2 * fs + idx] = w;
}

// elementwise helpers of the adjoint-descent step (n complex entries)
// y = a + (h b) 0.5 when HALF, else y = a + h b
template <class T, bool HALF>
__global__ void __launch_bounds__(256)
    k_axpy(cpx<T>* __restrict__ y, const cpx<T>* __restrict__ a, const cpx<T>* __restrict__ b, double h, long long n)
{
    const T hh = T(h);
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        cpx<T> d = hh * b[i];
        if (HALF) d = T(0.5) * d;
        y[i] = a[i] + d;
    }
}

// zero n complex entries (a kernel, not a memset: it may be captured into a graph)
template <class T>
__global__ void __launch_bounds__(256) k_fill_zero(cpx<T>* __restrict__ y, long long n)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        y[i] = cpx<T>{T(0), T(0)};
}

// the clock of a box advanced by h (fictitious time of the adjoint descent)
template <int D = 0>
__global__ void k_add_time(BoxScalars* sc, double h)
{
    if (threadIdx.x == 0) sc->t += h;
}

template <class TO, class TI>
__global__ void k_convert(TO* __restrict__ dst, const TI* __restrict__ src, long long n)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        dst[i] = TO(src[i]);
}

// One spectral field between the reference layout uhat(Nz, Ny, Nx/2+1) (element (ix, iy, iz) at
// (ix Ny + iy) Nz + iz, FP64 complex) and the device layout (at (iz Ny + iy) H + ix, type D):
// TO_DEV reads the reference layout.  A block transposes a 32 x 32 (ix, iz) tile of one iy
// plane through LDS, so that both sides are read and written along their fastest index.
// Grid (ceil(Nz / 32), ceil(H / 32), Ny), 256 threads.
template <class D, bool TO_DEV>
__global__ void __launch_bounds__(256)
    k_ref_layout(const void* __restrict__ src, void* __restrict__ dst, int nx, int ny, int nz)
{
    __shared__ cpx<double> tile[32][33];
    const int h = nx / 2 + 1;
    const int iz0 = blockIdx.x * 32, ix0 = blockIdx.y * 32, iy = blockIdx.z;
    const int tx = threadIdx.x % 32, ty = threadIdx.x / 32;
    // tile[a][b] holds (ix, iz) = (ix0 + a, iz0 + b)
    if constexpr (TO_DEV) {
        const cpx<double>* r = static_cast<const cpx<double>*>(src);
        for (int j = 0; j < 32; j += 8) {
            const int ix = ix0 + ty + j, iz = iz0 + tx;
            if (ix < h && iz < nz) tile[ty + j][tx] = r[((long long)ix * ny + iy) * nz + iz];
        }
        __syncthreads();
        cpx<D>* d = static_cast<cpx<D>*>(dst);
        for (int j = 0; j < 32; j += 8) {
            const int iz = iz0 + ty + j, ix = ix0 + tx;
            if (ix < h && iz < nz) {
                const cpx<double> v = tile[tx][ty + j];
                d[((long long)iz * ny + iy) * h + ix] = cpx<D>{D(v.x), D(v.y)};
            }
        }
    } else {
        const cpx<D>* d = static_cast<const cpx<D>*>(src);
        for (int j = 0; j < 32; j += 8) {
            const int iz = iz0 + ty + j, ix = ix0 + tx;
            if (ix < h && iz < nz) {
                const cpx<D> v = d[((long long)iz * ny + iy) * h + ix];
                tile[tx][ty + j] = cpx<double>{double(v.x), double(v.y)};
            }
        }
        __syncthreads();
        cpx<double>* r = static_cast<cpx<double>*>(dst);
        for (int j = 0; j < 32; j += 8) {
            const int ix = ix0 + ty + j, iz = iz0 + tx;
            if (ix < h && iz < nz) r[((long long)ix * ny + iy) * nz + iz] = tile[ty + j][tx];
        }
    }
}

}  // namespace dnsbox
