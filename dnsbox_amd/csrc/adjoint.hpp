// Backward kernels of one box.
//
// (1) The exact discrete adjoint of the RHS, g = J_N(u)^T lambda, in the Euclidean
//     inner product of the stored half-spectrum coordinates <a, b> = sum Re(conj(a) b)
//     (what torch.autograd propagates).  With u_p = c2r(s u), w_p = c2r(s i k x u),
//     y = w_p x u_p, n = r2c(y), N = -P M n [+ band forcing]:
//         nbar  = -P M lambda,          ybar = c2r(nbar / W)
//         ubar  = ybar x w_p,           wbar = u_p x ybar
//         g     = s W (r2c(ubar) + i k x r2c(wbar))  [+ band terms]
//     W is the Hermitian weight of a stored entry (1 on the kx = 0 and kx = NX/2 planes,
//     2 elsewhere) and (i k x) is Hermitian, so it is its own adjoint.  Five passes, as
//     the forward: z-inverse (9 outputs), y-inverse (9), x (9 -> 6), y-forward (6),
//     z-forward (6) with the step-adjoint epilogue.
// (2) The reference's adjoint-descent RHS, rhsAdj.f90:64-381 (convective form, ".gt."
//     mask, eps at k = 0, H^-1 weighting), with phatadj read as phat.
#pragma once

#include "box.hpp"

namespace dnsbox {

template <class T, int K>
struct Ptrs {
    const cpx<T>* p[K];
};
template <class T, int K>
struct OutPtrs {
    cpx<T>* p[K];
};

__device__ __forceinline__ double herm_w(int ix, int nx) { return (ix == 0 || ix == nx / 2) ? 1.0 : 2.0; }

// ".gt." 2/3 rule of rhsAdj.f90:131-136 (the reference has no Nyquist or k = 0 exclusion
// here; parity = 0 removes the Nyquist-index modes as everywhere else)
__device__ __forceinline__ bool masked_gt(const ModeK& m, const BoxPhys& ph)
{
    if (fabs(double(m.kx) * ph.a[0]) > ph.thr[0] || fabs(double(m.ky) * ph.a[1]) > ph.thr[1] ||
        fabs(double(m.kz) * ph.a[2]) > ph.thr[2])
        return true;
    if (!ph.parity && (m.ix == ph.nyq[0] || m.iy == ph.nyq[1] || m.iz == ph.nyq[2])) return true;
    return false;
}

// ---------------------------------------------------------------------------
// prologues
// ---------------------------------------------------------------------------
template <class T>
struct ProVjpIn {  // (lambda, u) -> (lam_scale (-P M lambda) / W, s u, s i k x u)
    static constexpr int FI = 6, FO = 9;
    T s;
    T lam_scale;
    int nx;      // points along x (the Hermitian weight of the kx = NX/2 plane)
    BoxPhys ph;
    __device__ __forceinline__ cpx<T> operator()(int o, const cpx<T>* in, const ModeK& m) const
    {
        const T kx = T(double(m.kx) * ph.a[0]), ky = T(double(m.ky) * ph.a[1]), kz = T(double(m.kz) * ph.a[2]);
        if (o < 3) {
            if (masked(m.ix, m.iy, m.iz, m.kx, m.ky, m.kz, ph)) return cpx<T>{T(0), T(0)};
            const T inv = T(1) / T(kk2(m.kx, m.ky, m.kz, ph));
            const cpx<T> kl = inv * (kx * in[0] + ky * in[1] + kz * in[2]);
            const T k = o == 0 ? kx : (o == 1 ? ky : kz);
            const T f = lam_scale / T(herm_w(m.ix, nx));
            return f * (k * kl - in[o]);
        }
        const cpx<T> u = in[3], v = in[4], w = in[5];
        switch (o) {
            case 3: return s * u;
            case 4: return s * v;
            case 5: return s * w;
            case 6: return s * mul_i<1>(ky * w - kz * v);
            case 7: return s * mul_i<1>(kz * u - kx * w);
            default: return s * mul_i<1>(kx * v - ky * u);
        }
    }
};

template <class T>
struct ProVjpLam {  // lambda -> lam_scale (-P M lambda) / W: the cotangent part of ProVjpIn alone
    static constexpr int FI = 3, FO = 3;
    T lam_scale;
    int nx;
    BoxPhys ph;
    __device__ __forceinline__ cpx<T> operator()(int o, const cpx<T>* in, const ModeK& m) const
    {
        if (masked(m.ix, m.iy, m.iz, m.kx, m.ky, m.kz, ph)) return cpx<T>{T(0), T(0)};
        const T kx = T(double(m.kx) * ph.a[0]), ky = T(double(m.ky) * ph.a[1]), kz = T(double(m.kz) * ph.a[2]);
        const T inv = T(1) / T(kk2(m.kx, m.ky, m.kz, ph));
        const cpx<T> kl = inv * (kx * in[0] + ky * in[1] + kz * in[2]);
        const T k = o == 0 ? kx : (o == 1 ? ky : kz);
        const T f = lam_scale / T(herm_w(m.ix, nx));
        return f * (k * kl - in[o]);
    }
};

template <class T>
struct ProGrad12 {  // u -> (s u_i, s i k_j u_i):  outputs 3 + 3 i + j = d_j u_i
    static constexpr int FI = 3, FO = 12;
    T s;
    KScale<T> k;
    __device__ __forceinline__ cpx<T> operator()(int o, const cpx<T>* in, const ModeK& m) const
    {
        if (o < 3) return s * in[o];
        const int i = (o - 3) / 3, j = (o - 3) % 3;
        return s * mul_i<1>(k.d(m, j) * in[i]);
    }
};

template <class T>
struct ProAdj12 {  // (u~, u) -> (s u_i, s i k_j u~_i)
    static constexpr int FI = 6, FO = 12;
    T s;
    KScale<T> k;
    __device__ __forceinline__ cpx<T> operator()(int o, const cpx<T>* in, const ModeK& m) const
    {
        if (o < 3) return s * in[3 + o];
        const int i = (o - 3) / 3, j = (o - 3) % 3;
        return s * mul_i<1>(k.d(m, j) * in[i]);
    }
};

// PR: tiles over the kept (ky, kx) columns, inputs read at kept kz only (inputs confined to the
// kept cube); every z of those columns is written
template <class Dm, class T, class Pro, bool PR = false>
__global__ void __launch_bounds__((Geo<Dm, T>::ZNT))
    k_zinv_p(Ptrs<T, Pro::FI> in, cpx<T>* __restrict__ out0, long long out_fs,
             const cpx<T>* __restrict__ tw_g, Pro pro)
{
    using G = Geo<Dm, T>;
    __shared__ cpx<T> lds[G::ZLDS];
    __shared__ cpx<T> stw_[G::NZ];
    const cpx<T>* tw = stage_twiddles<G::NZ, T>(tw_g, stw_);
    cpx<T>* out[Pro::FO];
#pragma unroll
    for (int f = 0; f < Pro::FO; ++f) out[f] = out0 + f * out_fs;
    zinv_tile<Dm, T, Pro, T, PR>(in.p, out, xcd_contiguous(blockIdx.x, gridDim.x), lds, tw, pro);
}

// k_zinv_p over the kept columns (PR) in its first `tiles` blocks; the blocks after them check
// the three fields at chk (field stride chk_fs) and set *viol when one holds a nonzero mode
// outside the kept cube.  The cube-confined VJP runs its first pass with this guard on its
// linearisation state; the flag is read back when the caller withdraws the guarantee.
template <class Dm, class T, class Pro>
__global__ void __launch_bounds__((Geo<Dm, T>::ZNT))
    k_zinv_p_chk(Ptrs<T, Pro::FI> in, cpx<T>* __restrict__ out0, long long out_fs,
                 const cpx<T>* __restrict__ tw_g, Pro pro, const cpx<T>* __restrict__ chk, long long chk_fs,
                 unsigned int* viol, int tiles)
{
    using G = Geo<Dm, T>;
    __shared__ cpx<T> lds[G::ZLDS];
    __shared__ cpx<T> stw_[G::NZ];
    const int b = blockIdx.x;
    if (b >= tiles) {   // a check block (the whole block returns here)
        const long long nthr = (long long)(gridDim.x - tiles) * blockDim.x;
        const bool nz = outside_cube_nonzero<Dm, T>(chk, chk_fs, (long long)(b - tiles) * blockDim.x + threadIdx.x, nthr);
        if (__any(nz) && threadIdx.x % 64 == 0) atomicOr(viol, 1u);
        return;
    }
    const cpx<T>* tw = stage_twiddles<G::NZ, T>(tw_g, stw_);
    cpx<T>* out[Pro::FO];
#pragma unroll
    for (int f = 0; f < Pro::FO; ++f) out[f] = out0 + f * out_fs;
    zinv_tile<Dm, T, Pro, T, true>(in.p, out, xcd_contiguous(b, tiles), lds, tw, pro);
}

// blocks of the confinement check fused into the first pass of a cube-confined VJP
template <class Dm>
constexpr int vjp_check_blocks()
{
    constexpr long long tot = Dm::FS;
    return tot / 256 + 1 < 256 ? int(tot / 256 + 1) : 256;
}

// ---------------------------------------------------------------------------
// physical-space ops of the x pass
// ---------------------------------------------------------------------------
template <class T>
struct OpVjp {  // (ybar, u, w) -> (ybar x w, u x ybar)
    static constexpr int FI = 9, FO = 6, NRED = 0;
    static constexpr bool IN_SPEC = true, OUT_SPEC = true;
    __device__ __forceinline__ static void apply(const T* in, T* out, XRed<T>&)
    {
        const T y0 = in[0], y1 = in[1], y2 = in[2];
        const T u0 = in[3], u1 = in[4], u2 = in[5];
        const T w0 = in[6], w1 = in[7], w2 = in[8];
        out[0] = y1 * w2 - y2 * w1;
        out[1] = y2 * w0 - y0 * w2;
        out[2] = y0 * w1 - y1 * w0;
        out[3] = u1 * y2 - u2 * y1;
        out[4] = u2 * y0 - u0 * y2;
        out[5] = u0 * y1 - u1 * y0;
    }
};

template <class T>
struct OpConv {  // (u, d_j u_i) -> (u . grad) u          rhsAdj.f90:83-122
    static constexpr int FI = 12, FO = 3, NRED = 0;
    static constexpr bool IN_SPEC = true, OUT_SPEC = true;
    __device__ __forceinline__ static void apply(const T* in, T* out, XRed<T>&)
    {
#pragma unroll
        for (int i = 0; i < 3; ++i)
            out[i] = in[0] * in[3 + 3 * i] + in[1] * in[3 + 3 * i + 1] + in[2] * in[3 + 3 * i + 2];
    }
};

template <class T>
struct OpAdjN {  // (u, d_j u~_i) -> sum_j (d_i u~_j + d_j u~_i) u_j      rhsAdj.f90:271-309
    static constexpr int FI = 12, FO = 3, NRED = 0;
    static constexpr bool IN_SPEC = true, OUT_SPEC = true;
    __device__ __forceinline__ static void apply(const T* in, T* out, XRed<T>&)
    {
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            T acc = T(0);
#pragma unroll
            for (int j = 0; j < 3; ++j) acc += (in[3 + 3 * j + i] + in[3 + 3 * i + j]) * in[j];
            out[i] = acc;
        }
    }
};

// ---------------------------------------------------------------------------
// z-forward epilogues
// ---------------------------------------------------------------------------
enum VjpMode { VJP_RHS = 0, VJP_S1 = 1, VJP_S2 = 2 };

template <class T>
struct VjpArgs {
    cpx<T>* out;          // VJP_RHS: g;  VJP_S2: lambda_n
    const cpx<T>* lam;    // VJP_S1: lambda_{n+1}
    cpx<T>* A;            // VJP_S1 writes E (lambda + lambda_*), VJP_S2 reads it
    cpx<T>* Mu;           // VJP_S1 writes E (dt/2 lambda + dt lambda_*)
    const cpx<T>* u;      // state of the linearisation (band forcing)
    const cpx<T>* lamN;   // cotangent of N before lam_scale (band forcing)
    double lam_scale;
    double dt;
    const double* band;   // [Eband(u), sum_band Re(conj(lamN) u)] (lam_scale applied)
    long long fs;
    BoxPhys ph;
};

// the step-adjoint epilogue of zfwd_vjp_tile on the forward-transformed column of a valid lane;
// NEXT: U receives Mu at the kept cube's modes of a kept column (zlive), zero elsewhere
template <class Dm, class T, int MODE, bool NEXT, int UR>
__device__ __forceinline__ void zfwd_vjp_epilogue(const cpx<T> (&O)[6][Geo<Dm, T>::ZRM], int col,
                                                  const VjpArgs<T>& va, cpx<T> (&U)[3][UR], bool zlive)
{
    using G = Geo<Dm, T>;
    using CF = typename G::CFZ_F;
    constexpr int W = G::ZW, RM = G::ZRM, H = G::H;
    constexpr long long ZS = (long long)G::NY * H;
    const int g = threadIdx.x / W;
    const int ix = col % H, iy = col / H;
    const int kx = kval(ix, G::NX), ky = kval(iy, G::NY);
    const BoxPhys& ph = va.ph;
    const T kxs = T(double(kx) * ph.a[0]), kys = T(double(ky) * ph.a[1]);
    const T s = T(1.0 / (double(G::NX) * G::NY * G::NZ));
    const T w = T(herm_w(ix, G::NX));
    double cband = 0.0, fband = 0.0;
    if (ph.bandlim) {
        cband = ph.Pin / (2.0 * va.band[0]);
        fband = -ph.Pin / (2.0 * va.band[0] * va.band[0]) * va.band[1] * 2.0 * double(s) * double(s);
    }
#pragma unroll
    for (int q = 0; q < RM; ++q) {
        const int iz = CF::out_pos(g, q);
        const int kz = kval(iz, G::NZ);
        const T kzs = T(double(kz) * ph.a[2]);
        const double k2d = kk2(kx, ky, kz, ph);
        const long long idx = (long long)iz * ZS + col;
        // g = s W (Ubar + i k x Wbar)
        const cpx<T> a0 = O[3][q], a1 = O[4][q], a2 = O[5][q];
        cpx<T> gg[3];
        gg[0] = (s * w) * (O[0][q] + mul_i<1>(kys * a2 - kzs * a1));
        gg[1] = (s * w) * (O[1][q] + mul_i<1>(kzs * a0 - kxs * a2));
        gg[2] = (s * w) * (O[2][q] + mul_i<1>(kxs * a1 - kys * a0));
        if (ph.bandlim) {
            if (sqrt(k2d) <= ph.kcut) {
                const bool mk = masked(ix, iy, iz, kx, ky, kz, ph);
                const T we = T(kx == 0 ? 0.5 : 1.0);
#pragma unroll
                for (int f = 0; f < 3; ++f) {
                    const cpx<T> uu = va.u[f * va.fs + idx];
                    if (!mk) gg[f] = gg[f] + T(cband * va.lam_scale) * va.lamN[f * va.fs + idx];
                    gg[f] = gg[f] + (T(fband) * we) * uu;
                }
            }
        }
        if constexpr (MODE == VJP_RHS) {
#pragma unroll
            for (int f = 0; f < 3; ++f) va.out[f * va.fs + idx] = gg[f];
        } else if constexpr (MODE == VJP_S1) {
            const T e = T(exp((-ph.nu * k2d + ph.Q) * va.dt));
            const T hdt = T(0.5 * va.dt), tdt = T(va.dt);
#pragma unroll
            for (int f = 0; f < 3; ++f) {
                const cpx<T> l = va.lam[f * va.fs + idx];
                const cpx<T> mu = e * (hdt * l + tdt * gg[f]);
                va.A[f * va.fs + idx] = e * (l + gg[f]);
                va.Mu[f * va.fs + idx] = mu;
                if constexpr (NEXT) U[f][q] = zlive && Keep<G::NZ>::kept(iz) ? mu : cpx<T>{T(0), T(0)};
            }
        } else {
#pragma unroll
            for (int f = 0; f < 3; ++f) {
                const cpx<T> o = va.A[f * va.fs + idx] + gg[f];
                va.out[f * va.fs + idx] = o;
                if constexpr (NEXT) U[f][q] = zlive && Keep<G::NZ>::kept(iz) ? o : cpx<T>{T(0), T(0)};
            }
        }
    }
}

// NEXT: the cotangent this epilogue writes also enters the first pass of the VJP that reads it
// next, the cotangent z-inverse (ProVjpLam over the kept columns, inputs at the kept kz, as in
// the pruned first pass of vjp_rec_passes), whose outputs go to next[0..2] at every z, the way
// zfwd_zinv_tile hands the stepper's result to the next right-hand side: VJP_S1's Mu to VJP_S2
// of the same step, VJP_S2's lambda_n to VJP_S1 of the step before.  The symmetric z plan puts
// the inverse's input positions at the forward's output positions, so the cotangent does not
// make the round trip through HBM before that transform.
template <class Dm, class T, int MODE, bool NEXT = false>
__device__ __forceinline__ void zfwd_vjp_tile(const cpx<T>* __restrict__ X, long long xfs, int tile, cpx<T>* lds,
                                              const cpx<T>* tw, const VjpArgs<T>& va,
                                              cpx<T>* const* next = nullptr, const ProVjpLam<T>& pro = {})
{
    static_assert(!NEXT || MODE != VJP_RHS, "a step VJP hands its cotangent on");
    using G = Geo<Dm, T>;
    using CF = typename G::CFZ_F;
    constexpr int W = G::ZW, RM = G::ZRM, H = G::H;
    constexpr long long ZS = (long long)G::NY * H;
    const int col = tile * W + threadIdx.x % W;
    const int g = threadIdx.x / W;
    const bool valid = col < G::ZCOLS;
    cpx<T> O[6][RM];
    zfwd_regs<Dm, T>(X, xfs, col, valid, lds, tw, O);
    if constexpr (NEXT) {
        const bool zlive = valid && Dm::kept_col(col);
        cpx<T> U[3][RM];
        if (valid) zfwd_vjp_epilogue<Dm, T, MODE, NEXT>(O, col, va, U, zlive);
        else {
#pragma unroll
            for (int f = 0; f < 3; ++f)
#pragma unroll
                for (int q = 0; q < RM; ++q) U[f][q] = cpx<T>{T(0), T(0)};
        }
        // a tile with no kept column (uniform over the block) skips the transform
        if (__syncthreads_or(zlive)) zinv_from_regs<Dm, T, ProVjpLam<T>>(U, next, col, zlive, lds, tw, pro);
    } else {
        if (!valid) return;
        cpx<T> U[3][1];
        zfwd_vjp_epilogue<Dm, T, MODE, false>(O, col, va, U, false);
    }
}

template <class Dm, class T, int MODE>
__global__ void __launch_bounds__((Geo<Dm, T>::ZNT))
    k_zfwd_vjp(const cpx<T>* __restrict__ X, long long xfs, const cpx<T>* __restrict__ tw_g, VjpArgs<T> va)
{
    using G = Geo<Dm, T>;
    __shared__ cpx<T> lds[G::ZLDS];
    __shared__ cpx<T> stw_[G::NZ];
    const cpx<T>* tw = stage_twiddles<G::NZ, T>(tw_g, stw_);
    zfwd_vjp_tile<Dm, T, MODE>(X, xfs, xcd_contiguous(blockIdx.x, gridDim.x), lds, tw, va);
}

// A step VJP's z-forward and epilogue fused with the next VJP's cotangent z-inverse
// (zfwd_vjp_tile with NEXT) in the first `tiles` blocks; the blocks after them check the three
// fields at chk (the next VJP's linearisation state) outside the kept cube, as k_zinv_p_chk does
// in an unfused first pass.
template <class Dm, class T, int MODE>
__global__ void __launch_bounds__((Geo<Dm, T>::ZNT))
    k_zfwd_vjp_next(const cpx<T>* __restrict__ X, long long xfs, const cpx<T>* __restrict__ tw_g, VjpArgs<T> va,
                    cpx<T>* __restrict__ next0, long long next_fs, ProVjpLam<T> pro,
                    const cpx<T>* __restrict__ chk, long long chk_fs, unsigned int* viol, int tiles)
{
    using G = Geo<Dm, T>;
    __shared__ cpx<T> lds[G::ZLDS];
    __shared__ cpx<T> stw_[G::NZ];
    const int b = blockIdx.x;
    if (b >= tiles) {   // a check block (the whole block returns here)
        const long long nthr = (long long)(gridDim.x - tiles) * blockDim.x;
        const bool nz = outside_cube_nonzero<Dm, T>(chk, chk_fs, (long long)(b - tiles) * blockDim.x + threadIdx.x, nthr);
        if (__any(nz) && threadIdx.x % 64 == 0) atomicOr(viol, 1u);
        return;
    }
    const cpx<T>* tw = stage_twiddles<G::NZ, T>(tw_g, stw_);
    cpx<T>* next[3] = {next0, next0 + next_fs, next0 + 2 * next_fs};
    zfwd_vjp_tile<Dm, T, MODE, true>(X, xfs, xcd_contiguous(b, tiles), lds, tw, va, next, pro);
}

// [Eband(u), lam_scale * sum_{|k| <= kcut, unmasked} Re(conj(lamN) u)]  (one block)
template <class Dm, class T>
__global__ void __launch_bounds__(256)
    k_band_vjp_scalars(const cpx<T>* __restrict__ u, const cpx<T>* __restrict__ lamN, long long fs,
                       BoxPhys ph, double lam_scale, double sm2, double* out)
{
    constexpr int H = Dm::H, NX = Dm::NX, NY = Dm::NY, NZ = Dm::NZ;
    const int KX = band_k<NX>(ph.kcut, ph.a[0]), KY = band_k<NY>(ph.kcut, ph.a[1]), KZ = band_k<NZ>(ph.kcut, ph.a[2]);
    const int nxc = KX + 2, nyc = 2 * KY + 2, nzc = 2 * KZ + 2;
    const int tot = nxc * nyc * nzc;
    double se = 0.0, sl = 0.0;
    for (int e = threadIdx.x; e < tot; e += blockDim.x) {
        const int a = e / (nyc * nzc), b = (e / nzc) % nyc, c = e % nzc;
        const int ix = a <= KX ? a : NX / 2;
        const int iy = band_idx<NY>(b, KY);
        const int iz = band_idx<NZ>(c, KZ);
        const int kx = kval(ix, NX), ky = kval(iy, NY), kz = kval(iz, NZ);
        if (sqrt(kk2(kx, ky, kz, ph)) > ph.kcut) continue;
        const long long idx = ((long long)iz * NY + iy) * H + ix;
        const bool mk = masked(ix, iy, iz, kx, ky, kz, ph);
        double en = 0.0, lu = 0.0;
#pragma unroll
        for (int f = 0; f < 3; ++f) {
            const cpx<T> v = u[f * fs + idx];
            const cpx<T> l = lamN[f * fs + idx];
            en += double(v.x) * double(v.x) + double(v.y) * double(v.y);
            lu += double(l.x) * double(v.x) + double(l.y) * double(v.y);
        }
        se += (kx == 0 ? 0.5 : 1.0) * en;
        if (!mk) sl += lu;
    }
    __shared__ double wsm[2][16];
    const double v0 = wave_sum(se), v1 = wave_sum(sl);
    if (threadIdx.x % 64 == 0) {
        wsm[0][threadIdx.x / 64] = v0;
        wsm[1][threadIdx.x / 64] = v1;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0.0, b = 0.0;
        for (int i = 0; i < int(blockDim.x / 64); ++i) {
            a += wsm[0][i];
            b += wsm[1][i];
        }
        out[0] = a * sm2;
        out[1] = b * lam_scale;
    }
}

// rhsAdj epilogues.  STAGE 1: N = -n - k phat (".gt." mask, eps at k = 0), then
// u~ = H (L u + N) (rhsupptilde, hminone).  STAGE 2: out = lin_sign L u~ - n_adj - k phat.
template <class T>
struct AdjArgs {
    cpx<T>* out;
    const cpx<T>* u;     // STAGE 1: the state
    cpx<T>* ut;          // STAGE 1 writes u~, STAGE 2 reads it
    long long fs;
    double lin_sign;
    BoxPhys ph;
};

template <class Dm, class T, int STAGE>
__global__ void __launch_bounds__((Geo<Dm, T>::ZNT))
    k_zfwd_adj(const cpx<T>* __restrict__ X, long long xfs, const cpx<T>* __restrict__ tw_g, AdjArgs<T> aa)
{
    using G = Geo<Dm, T>;
    using CF = typename G::CFZ_F;
    constexpr int W = G::ZW, RM = G::ZRM, H = G::H;
    constexpr long long ZS = (long long)G::NY * H;
    __shared__ cpx<T> lds[G::ZLDS];
    __shared__ cpx<T> stw_[G::NZ];
    const cpx<T>* tw = stage_twiddles<G::NZ, T>(tw_g, stw_);
    const int col = xcd_contiguous(blockIdx.x, gridDim.x) * W + threadIdx.x % W;
    const int g = threadIdx.x / W;
    const bool valid = col < G::ZCOLS;
    cpx<T> O[3][RM];
    zfwd_regs<Dm, T>(X, xfs, col, valid, lds, tw, O);
    if (!valid) return;
    const int ix = col % H, iy = col / H;
    const int kx = kval(ix, G::NX), ky = kval(iy, G::NY);
    const BoxPhys& ph = aa.ph;
#pragma unroll
    for (int q = 0; q < RM; ++q) {
        const int iz = CF::out_pos(g, q);
        const int kz = kval(iz, G::NZ);
        const long long idx = (long long)iz * ZS + col;
        const ModeK m{kx, ky, kz, ix, iy, iz};
        const double k2 = kk2(kx, ky, kz, ph);
        const T kxs = T(double(kx) * ph.a[0]), kys = T(double(ky) * ph.a[1]), kzs = T(double(kz) * ph.a[2]);
        cpx<T> nl[3];
        if (masked_gt(m, ph)) {
#pragma unroll
            for (int f = 0; f < 3; ++f) nl[f] = cpx<T>{T(0), T(0)};
        } else {
            cpx<T> pk = {T(0), T(0)};
            if (k2 > 0) pk = (T(1) / T(k2)) * (kxs * O[0][q] + kys * O[1][q] + kzs * O[2][q]);
            nl[0] = kxs * pk - O[0][q];
            nl[1] = kys * pk - O[1][q];
            nl[2] = kzs * pk - O[2][q];
        }
        const double L = -ph.nu * k2 + ph.Q;
        if constexpr (STAGE == 1) {
            const T hf = T(1.0 / (1.0 + k2));
#pragma unroll
            for (int f = 0; f < 3; ++f)
                aa.ut[f * aa.fs + idx] = hf * (T(L) * aa.u[f * aa.fs + idx] + nl[f]);
        } else {
#pragma unroll
            for (int f = 0; f < 3; ++f)
                aa.out[f * aa.fs + idx] = T(aa.lin_sign * L) * aa.ut[f * aa.fs + idx] + nl[f];
        }
    }
}

}  // namespace dnsbox
