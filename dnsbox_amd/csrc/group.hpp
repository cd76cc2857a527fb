// Many independent boxes of uneven size advanced by one launch per pass.
//
// A group holds boxes of N in {16, 32, 48, 64, 128} (one precision, each with its
// own viscosity, forcing, dt controller and state).  Every pass of the IF-Heun step
// is one launch over the concatenated tiles of all boxes: a block finds its box by
// binary search over the per-box first-tile offsets, then runs the same templated
// tile function the single-box kernels use, selected by a switch on the box size.
// All these sizes use 256-thread blocks in every pass, so one block shape serves
// the whole group; LDS is sized for the largest member.  The step (8 launches) is
// captured once in a hipGraph and replayed.
#pragma once

#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "box_impl.hpp"

namespace dnsbox {

template <class T>
struct GBox {
    int n;
    int zt0, yt0, xb0;  // first global z tile, y tile and x block of this box
    int zt0p, yt0p;     // first z and y tiles of the passes pruned to the 2/3-rule cube
    long long fs;
    cpx<T>* st;          // state advanced in place by the forward step
    cpx<T>* X;           // transform scratch (9 fields)
    const cpx<T>* tw;
    BoxScalars* sc;
    BoxPhys ph;
    // step adjoint: out = (d u_{n+1} / d u_n)^T lam at u, fixed step dt
    const cpx<T>* u;
    const cpx<T>* lam;
    cpx<T>* out;
    cpx<T>* pred;        // u* of the step
    cpx<T>* pout;        // forward step: where the IF-Heun predictor stores u* (null: not stored)
    cpx<T>* tout;        // forward step: where the corrector also stores the new state (null: none)
    cpx<T>* A;           // E (lam + lam_*)
    cpx<T>* Mu;          // E (dt/2 lam + dt lam_*)
    // the step's two record slots of 9 fields (slot r at mix + 9 r fields): the x-pass inputs of the
    // RHS at u (r = 0) and at u* (r = 1) in fields 3..8, written by a recording step and read by the
    // step adjoint, which transforms the cotangent into fields 0..2 (null: no recording)
    cpx<T>* mix;
    double dt;
};

constexpr int kGroupThreads = 256;

template <int N, class T>
constexpr bool group_size_ok()
{
    return Geo<Cube<N>, T>::ZNT == kGroupThreads && Geo<Cube<N>, T>::YNT == kGroupThreads &&
           Cfg<N, T>::NTX == kGroupThreads;
}

template <class T>
struct GroupLds {
    template <int N>
    static constexpr int z() { return Geo<Cube<N>, T>::ZLDS; }
    template <int N>
    static constexpr int y() { return N * Cfg<N, T>::YW; }
    template <int N, class Op>
    static constexpr int x() { return XGeo<N, T, Op>::LDS; }
    static constexpr int mx(int a, int b) { return a > b ? a : b; }
    static constexpr int Z = mx(mx(mx(z<16>(), z<32>()), mx(z<48>(), z<64>())), z<128>());
    static constexpr int Y = mx(mx(mx(y<16>(), y<32>()), mx(y<48>(), y<64>())), y<128>());
    template <class Op>
    static constexpr int X = mx(mx(mx(x<16, Op>(), x<32, Op>()), mx(x<48, Op>(), x<64, Op>())), x<128, Op>());
    static_assert(group_size_ok<16, T>() && group_size_ok<32, T>() && group_size_ok<48, T>() &&
                      group_size_ok<64, T>() && group_size_ok<128, T>(),
                  "group sizes must share one block shape");
};

// box owning global block `blk`, by the first-block offsets selected by `which`
// (0 z tiles, 1 y tiles, 2 x blocks, 3 pruned z tiles, 4 pruned y tiles)
template <class T>
__device__ __forceinline__ int find_box(const GBox<T>* __restrict__ gb, int nbox, int blk, int which)
{
    int lo = 0, hi = nbox - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        const int start = which == 0 ? gb[mid].zt0
                          : which == 1 ? gb[mid].yt0
                          : which == 2 ? gb[mid].xb0
                          : which == 3 ? gb[mid].zt0p
                                       : gb[mid].yt0p;
        if (start <= blk) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

template <class T>
__device__ __forceinline__ const cpx<T>* stage_tw_n(const cpx<T>* __restrict__ g, cpx<T>* s, int n)
{
    for (int i = threadIdx.x; i < n; i += blockDim.x) s[i] = g[i];
    __syncthreads();
    return s;
}

#define DNSBOX_GROUP_SWITCH(n, CALL) \
    switch (n) {                     \
        case 16: CALL(16); break;    \
        case 32: CALL(32); break;    \
        case 48: CALL(48); break;    \
        case 64: CALL(64); break;    \
        default: CALL(128); break;   \
    }

// z-inverse + curl of the state (from_u = 0) or of the step-adjoint point u (from_u = 1)
// PR: the pruned passes (box.hpp Keep; the states must be empty outside the kept cube)
template <class T, bool PR = false>
__global__ void __launch_bounds__(kGroupThreads) k_g_zinv(const GBox<T>* __restrict__ gb, int nbox, int from_u)
{
    __shared__ cpx<T> lds[GroupLds<T>::Z];
    __shared__ cpx<T> stw[128];
    // no XCD remapping across a group: round-robin dispatch spreads every box's tiles over
    // all XCDs, which balances boxes of uneven cost
    const int gt = blockIdx.x;
    const int b = find_box(gb, nbox, gt, PR ? 3 : 0);
    const GBox<T> bx = gb[b];
    const cpx<T>* tw = stage_tw_n(bx.tw, stw, bx.n);
    const int tile = gt - (PR ? bx.zt0p : bx.zt0);
    const cpx<T>* src = from_u ? bx.u : bx.st;
    const cpx<T>* in[3] = {src, src + bx.fs, src + 2 * bx.fs};
    cpx<T>* out[6];
#pragma unroll
    for (int f = 0; f < 6; ++f) out[f] = bx.X + f * bx.fs;
    const ProCurl<T> pro{T(1.0 / (double(bx.n) * bx.n * bx.n)), KScale<T>::of(bx.ph)};
#define DNSBOX_G_ZINV(NN) zinv_tile<Cube<NN>, T, ProCurl<T>, T, PR>(in, out, tile, lds, tw, pro)
    DNSBOX_GROUP_SWITCH(bx.n, DNSBOX_G_ZINV)
#undef DNSBOX_G_ZINV
}

template <class T, int DIR, bool PR = false>
__global__ void __launch_bounds__(kGroupThreads) k_g_ycol(const GBox<T>* __restrict__ gb, int nbox)
{
    __shared__ cpx<T> lds[GroupLds<T>::Y];
    __shared__ cpx<T> stw[128];
    const int L = xcd_contiguous(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
    const int field = L / gridDim.x, gt = L % gridDim.x;
    const int b = find_box(gb, nbox, gt, PR ? 4 : 1);
    const GBox<T> bx = gb[b];
    const cpx<T>* tw = stage_tw_n(bx.tw, stw, bx.n);
    const int tile = gt - (PR ? bx.yt0p : bx.yt0);
    cpx<T>* data = bx.X + field * bx.fs;
#define DNSBOX_G_YCOL(NN) ycol_tile<Cube<NN>, T, DIR, PR>(data, tile, lds, tw)
    DNSBOX_GROUP_SWITCH(bx.n, DNSBOX_G_YCOL)
#undef DNSBOX_G_YCOL
}

// the y pass on the record slots: SRC 1 reads the scratch and writes fields 3.. of slot `slot`
// (the recording step's y-inverse); SRC 2 runs in place on fields 0.. of slot `slot` (the step
// adjoint's cotangent)
template <class T, int DIR, bool PR, int SRC>
__global__ void __launch_bounds__(kGroupThreads) k_g_ycol_rec(const GBox<T>* __restrict__ gb, int nbox, int slot)
{
    __shared__ cpx<T> lds[GroupLds<T>::Y];
    __shared__ cpx<T> stw[128];
    const int L = xcd_contiguous(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
    const int field = L / gridDim.x, gt = L % gridDim.x;
    const int b = find_box(gb, nbox, gt, PR ? 4 : 1);
    const GBox<T> bx = gb[b];
    const cpx<T>* tw = stage_tw_n(bx.tw, stw, bx.n);
    const int tile = gt - (PR ? bx.yt0p : bx.yt0);
    cpx<T>* rec = bx.mix + (9LL * slot + (SRC == 1 ? 3 : 0) + field) * bx.fs;
    const cpx<T>* src = bx.X + field * bx.fs;
#define DNSBOX_G_YCOL(NN)                                                                   \
    if constexpr (SRC == 1) ycol_tile<Cube<NN>, T, DIR, PR, true>(rec, tile, lds, tw, src); \
    else ycol_tile<Cube<NN>, T, DIR, PR>(rec, tile, lds, tw)
    DNSBOX_GROUP_SWITCH(bx.n, DNSBOX_G_YCOL)
#undef DNSBOX_G_YCOL
}

// x pass of Op in place on the scratch; OpCross can track the CFL maxima
template <class T, class Op, bool PR = false, bool PO = PR>
__global__ void __launch_bounds__(kGroupThreads) k_g_xpass(const GBox<T>* __restrict__ gb, int nbox, int track)
{
    __shared__ cpx<T> lds[GroupLds<T>::template X<Op>];
    __shared__ cpx<T> stw[128];
    __shared__ double wsm[4 * 16];
    const int b = find_box(gb, nbox, blockIdx.x, 2);
    const GBox<T> bx = gb[b];
    const cpx<T>* tw = stage_tw_n(bx.tw, stw, bx.n);
    const int blk = blockIdx.x - bx.xb0;
    XRed<T> red;
#pragma unroll
    for (int j = 0; j < 4; ++j) red.m[j] = T(0);
#define DNSBOX_G_X(NN) xpass_tile<Cube<NN>, T, Op, PR, PO>(bx.X, bx.fs, bx.X, bx.fs, blk, lds, tw, red)
    DNSBOX_GROUP_SWITCH(bx.n, DNSBOX_G_X)
#undef DNSBOX_G_X
    if constexpr (Op::NRED == 3) {
        if (track) block_max_atomic<3>(red, bx.sc->maxu, false, wsm);
    }
}

// k_g_xpass reading its inputs from record slot `slot` (SRC 1: the six x-pass inputs in fields
// 3..8, as the recording step; SRC 2: all nine fields, as the step adjoint) into the scratch
template <class T, class Op, bool PR, bool PO, int SRC>
__global__ void __launch_bounds__(kGroupThreads)
    k_g_xpass_rec(const GBox<T>* __restrict__ gb, int nbox, int track, int slot)
{
    __shared__ cpx<T> lds[GroupLds<T>::template X<Op>];
    __shared__ cpx<T> stw[128];
    __shared__ double wsm[4 * 16];
    const int b = find_box(gb, nbox, blockIdx.x, 2);
    const GBox<T> bx = gb[b];
    const cpx<T>* tw = stage_tw_n(bx.tw, stw, bx.n);
    const int blk = blockIdx.x - bx.xb0;
    const cpx<T>* in = bx.mix + (9LL * slot + (SRC == 1 ? 3 : 0)) * bx.fs;
    XRed<T> red;
#pragma unroll
    for (int j = 0; j < 4; ++j) red.m[j] = T(0);
#define DNSBOX_G_X(NN) xpass_tile<Cube<NN>, T, Op, PR, PO>(in, bx.fs, bx.X, bx.fs, blk, lds, tw, red)
    DNSBOX_GROUP_SWITCH(bx.n, DNSBOX_G_X)
#undef DNSBOX_G_X
    if constexpr (Op::NRED == 3) {
        if (track) block_max_atomic<3>(red, bx.sc->maxu, false, wsm);
    }
}

template <class T, int MODE, bool PR = false>
__global__ void __launch_bounds__(kGroupThreads) k_g_zfwd_zinv(const GBox<T>* __restrict__ gb, int nbox)
{
    __shared__ cpx<T> lds[GroupLds<T>::Z];
    __shared__ cpx<T> stw[128];
    // no XCD remapping across a group: round-robin dispatch spreads every box's tiles over
    // all XCDs, which balances boxes of uneven cost
    const int gt = blockIdx.x;
    const int b = find_box(gb, nbox, gt, PR ? 3 : 0);
    const GBox<T> bx = gb[b];
    const cpx<T>* tw = stage_tw_n(bx.tw, stw, bx.n);
    const int tile = gt - (PR ? bx.zt0p : bx.zt0);
    EpiArgs<T, T> ea{};
    ea.out = MODE == EPI_PRED ? bx.pout : nullptr;
    ea.tout = MODE == EPI_CORR ? bx.tout : nullptr;
    ea.state = bx.st;
    ea.rhs_in = bx.st;
    ea.rhs_in_s = 1;
    ea.fs = bx.fs;
    ea.sc = bx.sc;
    ea.ph = bx.ph;
    const ProCurl<T> pro{T(1.0 / (double(bx.n) * bx.n * bx.n)), KScale<T>::of(bx.ph)};
#define DNSBOX_G_Z(NN) zfwd_zinv_tile<Cube<NN>, T, MODE, T, PR>(bx.X, bx.fs, tile, lds, tw, ea, pro)
    DNSBOX_GROUP_SWITCH(bx.n, DNSBOX_G_Z)
#undef DNSBOX_G_Z
    // the corrector also ends each box's step (its first tile; no tile of this pass reads the
    // scalars step_end writes)
    if constexpr (MODE == EPI_CORR) {
        if (tile == 0 && threadIdx.x == 0) step_end(bx.sc);
    }
}

// u* = E (u + dt N(u)) of the step adjoint, into pred
template <class T>
__global__ void __launch_bounds__(kGroupThreads) k_g_zfwd_predonly(const GBox<T>* __restrict__ gb, int nbox)
{
    __shared__ cpx<T> lds[GroupLds<T>::Z];
    __shared__ cpx<T> stw[128];
    // no XCD remapping across a group: round-robin dispatch spreads every box's tiles over
    // all XCDs, which balances boxes of uneven cost
    const int gt = blockIdx.x;
    const int b = find_box(gb, nbox, gt, 0);
    const GBox<T> bx = gb[b];
    const cpx<T>* tw = stage_tw_n(bx.tw, stw, bx.n);
    const int tile = gt - bx.zt0;
    EpiArgs<T, T> ea{};
    ea.out = bx.pred;
    ea.state = const_cast<cpx<T>*>(bx.u);  // read only in this mode
    ea.rhs_in = bx.u;
    ea.rhs_in_s = 0;
    ea.fs = bx.fs;
    ea.dt_fixed = bx.dt;
    ea.sc = bx.sc;
    ea.ph = bx.ph;
#define DNSBOX_G_Z(NN) zfwd_tile<Cube<NN>, T, EPI_PREDONLY, T>(bx.X, bx.fs, tile, lds, tw, ea)
    DNSBOX_GROUP_SWITCH(bx.n, DNSBOX_G_Z)
#undef DNSBOX_G_Z
}

// z-inverse of the VJP prologue: STAGE 1 at u* with cotangent dt/2 lam, STAGE 2 at u
// with cotangent Mu; PR: the kept columns only, inputs read at kept kz (states in the cube),
// and kGroupCheck blocks per box after the `tiles` tile blocks check the box's state outside
// its cube (*viol |= 1 on a nonzero mode there).  The cotangent's three outputs (ProVjpLam) and
// the state's six (ProCurl: s u, s i k x u, the outputs 3..8 of ProVjpIn) are two z-inverses in
// turn, so that three inputs rather than six are held in registers (256 VGPRs and one wave per
// SIMD when all six were held).
constexpr int kGroupCheck = 16;
template <class T, int STAGE, bool PR = false>
__global__ void __launch_bounds__(kGroupThreads)
    k_g_zinv_vjp(const GBox<T>* __restrict__ gb, int nbox, int tiles, unsigned int* viol)
{
    __shared__ cpx<T> lds[GroupLds<T>::Z];
    __shared__ cpx<T> stw[128];
    // no XCD remapping across a group: round-robin dispatch spreads every box's tiles over
    // all XCDs, which balances boxes of uneven cost
    const int gt = blockIdx.x;
    if (PR && gt >= tiles) {   // a check block (the whole block returns here)
        const int cb = gt - tiles, b = cb / kGroupCheck, part = cb % kGroupCheck;
        const GBox<T>& bc = gb[b];
        const cpx<T>* p = STAGE == 1 ? bc.pred : bc.u;
        bool nz = false;
#define DNSBOX_G_OC(NN) nz = outside_cube_nonzero<Cube<NN>, T>(p, bc.fs, part * (long long)blockDim.x + threadIdx.x, \
                                                         (long long)kGroupCheck * blockDim.x)
        DNSBOX_GROUP_SWITCH(bc.n, DNSBOX_G_OC)
#undef DNSBOX_G_OC
        if (__any(nz) && threadIdx.x % 64 == 0) atomicOr(viol, 1u);
        return;
    }
    const int b = find_box(gb, nbox, gt, PR ? 3 : 0);
    const GBox<T> bx = gb[b];
    const cpx<T>* tw = stage_tw_n(bx.tw, stw, bx.n);
    const int tile = gt - (PR ? bx.zt0p : bx.zt0);
    const cpx<T>* l = STAGE == 1 ? bx.lam : bx.Mu;
    const cpx<T>* p = STAGE == 1 ? bx.pred : bx.u;
    const cpx<T>* in_l[3] = {l, l + bx.fs, l + 2 * bx.fs};
    const cpx<T>* in_u[3] = {p, p + bx.fs, p + 2 * bx.fs};
    cpx<T>* out_l[3];
    cpx<T>* out_u[6];
#pragma unroll
    for (int f = 0; f < 3; ++f) out_l[f] = bx.X + f * bx.fs;
#pragma unroll
    for (int f = 0; f < 6; ++f) out_u[f] = bx.X + (3 + f) * bx.fs;
    const double sm = 1.0 / (double(bx.n) * bx.n * bx.n);
    const ProVjpLam<T> pl{T(STAGE == 1 ? 0.5 * bx.dt : 1.0), bx.n, bx.ph};
    const ProCurl<T> pu{T(sm), KScale<T>::of(bx.ph)};
#define DNSBOX_G_ZV(NN)                                                             \
    zinv_tile<Cube<NN>, T, ProVjpLam<T>, T, PR>(in_l, out_l, tile, lds, tw, pl); \
    zinv_tile<Cube<NN>, T, ProCurl<T>, T, PR>(in_u, out_u, tile, lds, tw, pu)
    DNSBOX_GROUP_SWITCH(bx.n, DNSBOX_G_ZV)
#undef DNSBOX_G_ZV
}

// the step adjoint's first pass with recorded x-pass inputs: the cotangent's z-inverse
// (ProVjpLam) into fields 0..2 of the record slot of the stage's RHS (STAGE 1: at u*, slot 1;
// STAGE 2: at u, slot 0); PR: the kept columns, with the kGroupCheck check blocks of k_g_zinv_vjp
template <class T, int STAGE, bool PR = false>
__global__ void __launch_bounds__(kGroupThreads)
    k_g_zinv_vjp_rec(const GBox<T>* __restrict__ gb, int nbox, int tiles, unsigned int* viol)
{
    __shared__ cpx<T> lds[GroupLds<T>::Z];
    __shared__ cpx<T> stw[128];
    const int gt = blockIdx.x;
    if (PR && gt >= tiles) {   // a check block (the whole block returns here)
        const int cb = gt - tiles, b = cb / kGroupCheck, part = cb % kGroupCheck;
        const GBox<T>& bc = gb[b];
        const cpx<T>* p = STAGE == 1 ? bc.pred : bc.u;
        bool nz = false;
#define DNSBOX_G_OC(NN) nz = outside_cube_nonzero<Cube<NN>, T>(p, bc.fs, part * (long long)blockDim.x + threadIdx.x, \
                                                         (long long)kGroupCheck * blockDim.x)
        DNSBOX_GROUP_SWITCH(bc.n, DNSBOX_G_OC)
#undef DNSBOX_G_OC
        if (__any(nz) && threadIdx.x % 64 == 0) atomicOr(viol, 1u);
        return;
    }
    const int b = find_box(gb, nbox, gt, PR ? 3 : 0);
    const GBox<T> bx = gb[b];
    const cpx<T>* tw = stage_tw_n(bx.tw, stw, bx.n);
    const int tile = gt - (PR ? bx.zt0p : bx.zt0);
    const cpx<T>* l = STAGE == 1 ? bx.lam : bx.Mu;
    const cpx<T>* in[3] = {l, l + bx.fs, l + 2 * bx.fs};
    cpx<T>* rec = bx.mix + (STAGE == 1 ? 9 : 0) * bx.fs;
    cpx<T>* out[3] = {rec, rec + bx.fs, rec + 2 * bx.fs};
    const ProVjpLam<T> pl{T(STAGE == 1 ? 0.5 * bx.dt : 1.0), bx.n, bx.ph};
#define DNSBOX_G_ZV(NN) zinv_tile<Cube<NN>, T, ProVjpLam<T>, T, PR>(in, out, tile, lds, tw, pl)
    DNSBOX_GROUP_SWITCH(bx.n, DNSBOX_G_ZV)
#undef DNSBOX_G_ZV
}

// z-forward of the VJP with the step-adjoint epilogue (VJP_S1 / VJP_S2)
template <class T, int MODE>
__global__ void __launch_bounds__(kGroupThreads) k_g_zfwd_vjp(const GBox<T>* __restrict__ gb, int nbox)
{
    __shared__ cpx<T> lds[GroupLds<T>::Z];
    __shared__ cpx<T> stw[128];
    // no XCD remapping across a group: round-robin dispatch spreads every box's tiles over
    // all XCDs, which balances boxes of uneven cost
    const int gt = blockIdx.x;
    const int b = find_box(gb, nbox, gt, 0);
    const GBox<T> bx = gb[b];
    const cpx<T>* tw = stage_tw_n(bx.tw, stw, bx.n);
    const int tile = gt - bx.zt0;
    VjpArgs<T> va{};
    va.out = bx.out;
    va.lam = bx.lam;
    va.A = bx.A;
    va.Mu = bx.Mu;
    va.u = bx.u;
    va.lamN = MODE == VJP_S1 ? bx.lam : bx.Mu;
    va.lam_scale = MODE == VJP_S1 ? 0.5 * bx.dt : 1.0;
    va.dt = bx.dt;
    va.band = nullptr;  // group boxes have no band forcing
    va.fs = bx.fs;
    va.ph = bx.ph;
#define DNSBOX_G_ZF(NN) zfwd_vjp_tile<Cube<NN>, T, MODE>(bx.X, bx.fs, tile, lds, tw, va)
    DNSBOX_GROUP_SWITCH(bx.n, DNSBOX_G_ZF)
#undef DNSBOX_G_ZF
}

// k_g_zfwd_vjp<T, VJP_S1> whose z-forward also runs VJP_S2's first pass (zfwd_vjp_tile with NEXT:
// Mu's z-inverse from registers into fields 0..2 of the record slot at u, the kept columns), with
// the kGroupCheck check blocks per box of that pass (the state u outside the kept cube) after
// the `tiles` tile blocks
template <class T>
__global__ void __launch_bounds__(kGroupThreads)
    k_g_zfwd_vjp_next(const GBox<T>* __restrict__ gb, int nbox, int tiles, unsigned int* viol)
{
    __shared__ cpx<T> lds[GroupLds<T>::Z];
    __shared__ cpx<T> stw[128];
    const int gt = blockIdx.x;
    if (gt >= tiles) {   // a check block (the whole block returns here)
        const int cb = gt - tiles, b = cb / kGroupCheck, part = cb % kGroupCheck;
        const GBox<T>& bc = gb[b];
        bool nz = false;
#define DNSBOX_G_OC(NN) nz = outside_cube_nonzero<Cube<NN>, T>(bc.u, bc.fs, part * (long long)blockDim.x + threadIdx.x, \
                                                         (long long)kGroupCheck * blockDim.x)
        DNSBOX_GROUP_SWITCH(bc.n, DNSBOX_G_OC)
#undef DNSBOX_G_OC
        if (__any(nz) && threadIdx.x % 64 == 0) atomicOr(viol, 1u);
        return;
    }
    const int b = find_box(gb, nbox, gt, 0);
    const GBox<T> bx = gb[b];
    const cpx<T>* tw = stage_tw_n(bx.tw, stw, bx.n);
    const int tile = gt - bx.zt0;
    VjpArgs<T> va{};
    va.lam = bx.lam;
    va.A = bx.A;
    va.Mu = bx.Mu;
    va.u = bx.u;
    va.lamN = bx.lam;
    va.lam_scale = 0.5 * bx.dt;
    va.dt = bx.dt;
    va.band = nullptr;  // group boxes have no band forcing
    va.fs = bx.fs;
    va.ph = bx.ph;
    cpx<T>* next[3] = {bx.mix, bx.mix + bx.fs, bx.mix + 2 * bx.fs};
    const ProVjpLam<T> pl{T(1.0), bx.n, bx.ph};
#define DNSBOX_G_ZF(NN) zfwd_vjp_tile<Cube<NN>, T, VJP_S1, true>(bx.X, bx.fs, tile, lds, tw, va, next, pl)
    DNSBOX_GROUP_SWITCH(bx.n, DNSBOX_G_ZF)
#undef DNSBOX_G_ZF
}

// *flag |= 1 when the state of some box (box blockIdx.y per block row) has a nonzero mode outside
// its kept cube: whether a step of caller buffers may run the pruned passes
template <class T>
__global__ void __launch_bounds__(256) k_g_outside_cube(const GBox<T>* __restrict__ gb, int nbox, unsigned int* flag)
{
    const GBox<T>& bx = gb[blockIdx.y];
    bool nz = false;
#define DNSBOX_G_OC(NN) nz = outside_cube_nonzero<Cube<NN>, T>(bx.st, bx.fs, blockIdx.x * (long long)blockDim.x + threadIdx.x, \
                                                         (long long)gridDim.x * blockDim.x)
    DNSBOX_GROUP_SWITCH(bx.n, DNSBOX_G_OC)
#undef DNSBOX_G_OC
    if (__any(nz) && threadIdx.x % 64 == 0) atomicOr(flag, 1u);
}

inline int group_ztiles(int n, int prec)
{
    int w = 0;
#define DNSBOX_GT(NN) w = prec == 64 ? Cfg<NN, double>::ZW : Cfg<NN, float>::ZW
    DNSBOX_GROUP_SWITCH(n, DNSBOX_GT)
#undef DNSBOX_GT
    return (n * (n / 2 + 1) + w - 1) / w;
}
inline int group_ytiles(int n, int prec)
{
    int w = 0;
#define DNSBOX_GT(NN) w = prec == 64 ? Cfg<NN, double>::YW : Cfg<NN, float>::YW
    DNSBOX_GROUP_SWITCH(n, DNSBOX_GT)
#undef DNSBOX_GT
    return (n * (n / 2 + 1) + w - 1) / w;
}
inline int group_ztiles_pruned(int n, int prec)
{
    int t = 0;
#define DNSBOX_GT(NN) t = prec == 64 ? ztiles<Cube<NN>, double, true>() : ztiles<Cube<NN>, float, true>()
    DNSBOX_GROUP_SWITCH(n, DNSBOX_GT)
#undef DNSBOX_GT
    return t;
}
inline int group_ytiles_pruned(int n, int prec)
{
    int t = 0;
#define DNSBOX_GT(NN) t = prec == 64 ? ytiles<Cube<NN>, double, true>() : ytiles<Cube<NN>, float, true>()
    DNSBOX_GROUP_SWITCH(n, DNSBOX_GT)
#undef DNSBOX_GT
    return t;
}
inline int group_xblocks(int n, int prec)
{
    int rb = 0;
#define DNSBOX_GT(NN) rb = prec == 64 ? Cfg<NN, double>::RB : Cfg<NN, float>::RB
    DNSBOX_GROUP_SWITCH(n, DNSBOX_GT)
#undef DNSBOX_GT
    return n * n / rb;
}

class GroupBase {
   public:
    virtual ~GroupBase() = default;
    std::vector<std::unique_ptr<BoxBase>> boxes;
    hipStream_t stream = nullptr;
    bool vjp_cube = false;            // the caller guarantees the VJPs' states lie in the kept cubes
    // the group form of BoxBase::set_vjp_cube (the flag is the first box's, on the group stream)
    bool set_vjp_cube(bool on)
    {
        unsigned int* viol = boxes[0]->chk + 1;
        if (on) {
            if (!vjp_cube) DNSBOX_HIP(hipMemsetAsync(viol, 0, sizeof(unsigned int), stream));
            vjp_cube = true;
            return false;
        }
        if (!vjp_cube) return false;
        vjp_cube = false;
        unsigned int h = 0;
        DNSBOX_HIP(hipMemcpyAsync(&h, viol, sizeof(h), hipMemcpyDeviceToHost, stream));
        DNSBOX_HIP(hipStreamSynchronize(stream));
        return h != 0;
    }
    bool last_ext_pruned = false;     // the last steps of caller buffers ran the pruned passes
    virtual void step(int nsteps) = 0;
    virtual void step_ext(void* const* states, int nsteps) = 0;
    virtual void step_vjp(const void* const* u, const void* const* lam, void* const* out, const double* dt) = 0;
    virtual void step_ext_record(void* const* states, int nsteps, void* const* traj, void* const* traj_pred) = 0;
    virtual void step_vjp_pred(const void* const* u, const void* const* ustar, const void* const* lam,
                               void* const* out, const double* dt) = 0;
    // step_ext_record that also records the x-pass inputs of the two RHS of each step: box i's
    // step j + 1 in the two 9-field slots at mixed[i] + 18 j fields (fields 3..8 of each)
    virtual void step_ext_record_mixed(void* const* states, int nsteps, void* const* traj, void* const* traj_pred,
                                       void* const* mixed) = 0;
    // step_vjp_pred reading the x-pass inputs of each box's step from its slot pair mixed[i] (18
    // fields, as recorded; fields 0..2 of both slots are overwritten); pruned_rec: the recording ran
    // pruned (the slots hold the kept kx only; needs vjp_cube)
    virtual void step_vjp_mixed(const void* const* u, const void* const* ustar, const void* const* lam,
                                void* const* out, const double* dt, void* const* mixed, bool pruned_rec) = 0;
    virtual void invalidate() = 0;
};

template <class T>
class Group final : public GroupBase {
    using C = cpx<T>;

   public:
    GBox<T>* d_gb = nullptr;          // descriptors of the boxes' own states (read by the graph)
    std::vector<GBox<T>> h_gb;
    int zt = 0, yt = 0, xb = 0;
    int ztp = 0, ytp = 0;             // tiles of the pruned passes
    bool x6_valid = false;
    bool x6_pruned = false;           // the scratch was written by the pruned passes
    int graph_pruned = -1;            // mode the graph was captured in
    hipStream_t gstream = nullptr;
    hipGraphExec_t graph = nullptr;   // graph_steps steps of every box
    int graph_mode = -1;
    int graph_steps = 16;
    // descriptor sets of the device-buffer entry points, keyed by their bytes: the
    // pointers a caller cycles through (checkpoint segments, ping-pong cotangents) are
    // uploaded once and reused
    std::map<std::string, GBox<T>*> dcache;
    static constexpr size_t kMaxCached = 256;

    explicit Group(std::vector<std::unique_ptr<BoxBase>>&& bs)
    {
        boxes = std::move(bs);
        const int prec = sizeof(T) == 8 ? 64 : 32;
        for (auto& b : boxes) {
            b->reserve(9, true);  // VJP scratch and adjoint buffers up front
            GBox<T> g;
            std::memset(&g, 0, sizeof(g));  // descriptor sets are compared by their bytes
            g.n = b->cfg.n;
            g.zt0 = zt;
            g.yt0 = yt;
            g.xb0 = xb;
            g.zt0p = ztp;
            g.yt0p = ytp;
            zt += group_ztiles(g.n, prec);
            yt += group_ytiles(g.n, prec);
            xb += group_xblocks(g.n, prec);
            ztp += group_ztiles_pruned(g.n, prec);
            ytp += group_ytiles_pruned(g.n, prec);
            fill(g, *b);
            h_gb.push_back(g);
        }
        DNSBOX_HIP(hipMalloc(&d_gb, h_gb.size() * sizeof(GBox<T>)));
        DNSBOX_HIP(hipMemcpy(d_gb, h_gb.data(), h_gb.size() * sizeof(GBox<T>), hipMemcpyHostToDevice));
        DNSBOX_HIP(hipStreamCreate(&gstream));
    }
    ~Group() override
    {
        if (graph) (void)hipGraphExecDestroy(graph);
        if (gstream) {
            (void)hipStreamSynchronize(gstream);
            (void)hipStreamDestroy(gstream);
        }
        (void)hipDeviceSynchronize();
        for (auto& kv : dcache) (void)hipFree(kv.second);
        if (d_gb) (void)hipFree(d_gb);
    }

    static void fill(GBox<T>& g, const BoxBase& b)
    {
        g.fs = b.fs;
        g.st = static_cast<C*>(b.state);
        g.X = static_cast<C*>(b.scratch);
        g.tw = static_cast<const C*>(b.tw);
        g.sc = b.sc;
        g.ph = b.cfg.ph;
        g.pred = static_cast<C*>(b.pred);
        g.pout = nullptr;
        g.tout = nullptr;
        g.A = static_cast<C*>(b.adjbuf);
        g.Mu = static_cast<C*>(b.adjbuf) + 3 * b.fs;
    }

    void invalidate() override { x6_valid = false; }

    // a box's buffers move when a single-box entry point grows its scratch (rhsAdj):
    // rewrite the descriptors before the next group launch
    void refresh()
    {
        bool moved = false;
        for (size_t i = 0; i < boxes.size(); ++i) {
            const BoxBase& b = *boxes[i];
            if (h_gb[i].X != b.scratch || h_gb[i].st != b.state || h_gb[i].A != b.adjbuf) {
                fill(h_gb[i], b);
                moved = true;
            }
        }
        if (!moved) return;
        DNSBOX_HIP(hipDeviceSynchronize());
        DNSBOX_HIP(hipMemcpy(d_gb, h_gb.data(), h_gb.size() * sizeof(GBox<T>), hipMemcpyHostToDevice));
        drop_cache();
        x6_valid = false;
    }

    void drop_cache()
    {
        if (dcache.empty()) return;
        DNSBOX_HIP(hipDeviceSynchronize());
        for (auto& kv : dcache) (void)hipFree(kv.second);
        dcache.clear();
    }

    const GBox<T>* descriptors(const std::vector<GBox<T>>& h)
    {
        std::string key(reinterpret_cast<const char*>(h.data()), h.size() * sizeof(GBox<T>));
        auto it = dcache.find(key);
        if (it != dcache.end()) return it->second;
        if (dcache.size() >= kMaxCached) drop_cache();
        GBox<T>* d = nullptr;
        DNSBOX_HIP(hipMalloc(&d, h.size() * sizeof(GBox<T>)));
        DNSBOX_HIP(hipMemcpy(d, h.data(), h.size() * sizeof(GBox<T>), hipMemcpyHostToDevice));
        dcache.emplace(std::move(key), d);
        return d;
    }

    template <bool PR = false>
    void launch_zinv(hipStream_t s, const GBox<T>* d, int from_u)
    {
        hipLaunchKernelGGL((k_g_zinv<T, PR>), dim3(PR ? ztp : zt), dim3(kGroupThreads), 0, s, d, int(h_gb.size()),
                           from_u);
        DNSBOX_LAUNCH_CHECK();
    }

    // one IF-Heun step of every box, its z-inverse + curl already in the scratch
    template <bool PR = false>
    void one_step(hipStream_t s, const GBox<T>* d)
    {
        const int nb = int(h_gb.size());
        const int z = PR ? ztp : zt, y = PR ? ytp : yt;
        hipLaunchKernelGGL((k_g_ycol<T, +1, PR>), dim3(y, 6), dim3(kGroupThreads), 0, s, d, nb);
        hipLaunchKernelGGL((k_g_xpass<T, OpCross<T>, PR>), dim3(xb), dim3(kGroupThreads), 0, s, d, nb, 1);
        hipLaunchKernelGGL((k_g_ycol<T, -1, PR>), dim3(y, 3), dim3(kGroupThreads), 0, s, d, nb);
        hipLaunchKernelGGL((k_g_zfwd_zinv<T, EPI_PRED, PR>), dim3(z), dim3(kGroupThreads), 0, s, d, nb);
        hipLaunchKernelGGL((k_g_ycol<T, +1, PR>), dim3(y, 6), dim3(kGroupThreads), 0, s, d, nb);
        hipLaunchKernelGGL((k_g_xpass<T, OpCross<T>, PR>), dim3(xb), dim3(kGroupThreads), 0, s, d, nb, 0);
        hipLaunchKernelGGL((k_g_ycol<T, -1, PR>), dim3(y, 3), dim3(kGroupThreads), 0, s, d, nb);
        hipLaunchKernelGGL((k_g_zfwd_zinv<T, EPI_CORR, PR>), dim3(z), dim3(kGroupThreads), 0, s, d, nb);
        DNSBOX_LAUNCH_CHECK();
    }
    void one_step(hipStream_t s, const GBox<T>* d, bool pr)
    {
        if (pr) one_step<true>(s, d);
        else one_step<false>(s, d);
    }
    // one_step whose y-inverses write the x-pass inputs into the boxes' record slots (GBox::mix),
    // from which the x passes read them
    template <bool PR>
    void one_step_rec(hipStream_t s, const GBox<T>* d)
    {
        const int nb = int(h_gb.size());
        const int z = PR ? ztp : zt, y = PR ? ytp : yt;
        hipLaunchKernelGGL((k_g_ycol_rec<T, +1, PR, 1>), dim3(y, 6), dim3(kGroupThreads), 0, s, d, nb, 0);
        hipLaunchKernelGGL((k_g_xpass_rec<T, OpCross<T>, PR, PR, 1>), dim3(xb), dim3(kGroupThreads), 0, s, d, nb, 1, 0);
        hipLaunchKernelGGL((k_g_ycol<T, -1, PR>), dim3(y, 3), dim3(kGroupThreads), 0, s, d, nb);
        hipLaunchKernelGGL((k_g_zfwd_zinv<T, EPI_PRED, PR>), dim3(z), dim3(kGroupThreads), 0, s, d, nb);
        hipLaunchKernelGGL((k_g_ycol_rec<T, +1, PR, 1>), dim3(y, 6), dim3(kGroupThreads), 0, s, d, nb, 1);
        hipLaunchKernelGGL((k_g_xpass_rec<T, OpCross<T>, PR, PR, 1>), dim3(xb), dim3(kGroupThreads), 0, s, d, nb, 0, 1);
        hipLaunchKernelGGL((k_g_ycol<T, -1, PR>), dim3(y, 3), dim3(kGroupThreads), 0, s, d, nb);
        hipLaunchKernelGGL((k_g_zfwd_zinv<T, EPI_CORR, PR>), dim3(z), dim3(kGroupThreads), 0, s, d, nb);
        DNSBOX_LAUNCH_CHECK();
    }

    // the boxes' own states run the pruned passes when every box's state allows it
    bool pruned() const
    {
        for (auto& b : boxes)
            if (!b->pruned()) return false;
        return true;
    }

    // IF-Heun steps of every box: the first z-inverse if needed, then graph_steps steps
    // captured once in a hipGraph and replayed, the remainder launched directly
    // (DNSBOX_STEP_GRAPH=0 launches every step directly).  Work runs on the group's own
    // stream, ordered after the caller's stream and before anything it enqueues next.
    void step(int nsteps) override
    {
        if (nsteps <= 0) return;
        refresh();
        hipEvent_t ev;
        DNSBOX_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        DNSBOX_HIP(hipEventRecord(ev, stream));
        DNSBOX_HIP(hipStreamWaitEvent(gstream, ev, 0));
        const bool pr = pruned();
        if (!x6_valid || x6_pruned != pr) {
            if (pr) launch_zinv<true>(gstream, d_gb, 0);
            else launch_zinv<false>(gstream, d_gb, 0);
            x6_valid = true;
            x6_pruned = pr;
        }
        if (graph_mode < 0) {
            const char* e = std::getenv("DNSBOX_STEP_GRAPH");
            graph_mode = (e && std::strcmp(e, "0") == 0) ? 0 : 1;
            const char* k = std::getenv("DNSBOX_GRAPH_STEPS");
            if (k && std::atoi(k) > 0) graph_steps = std::atoi(k);
        }
        if (graph && graph_pruned != int(pr)) {
            DNSBOX_HIP(hipStreamSynchronize(gstream));
            DNSBOX_HIP(hipGraphExecDestroy(graph));
            graph = nullptr;
        }
        if (graph_mode && nsteps >= graph_steps && !graph) {
            hipGraph_t g;
            DNSBOX_HIP(hipStreamBeginCapture(gstream, hipStreamCaptureModeThreadLocal));
            for (int i = 0; i < graph_steps; ++i) one_step(gstream, d_gb, pr);
            DNSBOX_HIP(hipStreamEndCapture(gstream, &g));
            DNSBOX_HIP(hipGraphInstantiate(&graph, g, nullptr, nullptr, 0));
            DNSBOX_HIP(hipGraphDestroy(g));
            graph_pruned = int(pr);
        }
        int done = 0;
        if (graph_mode)
            for (; done + graph_steps <= nsteps; done += graph_steps) DNSBOX_HIP(hipGraphLaunch(graph, gstream));
        for (; done < nsteps; ++done) one_step(gstream, d_gb, pr);
        DNSBOX_HIP(hipEventRecord(ev, gstream));
        DNSBOX_HIP(hipStreamWaitEvent(stream, ev, 0));
        DNSBOX_HIP(hipEventDestroy(ev));
    }

    // whether every caller state of the descriptors d is confined to its kept cube (checked on
    // the device; one 4-byte read back), so that their steps may run the pruned passes
    bool buffers_dealiased(const GBox<T>* d)
    {
        for (auto& b : boxes)
            if (!b->may_prune()) return false;
        unsigned int* flag = boxes[0]->chk;
        const int nb = int(h_gb.size());
        DNSBOX_HIP(hipMemsetAsync(flag, 0, sizeof(unsigned int), stream));
        hipLaunchKernelGGL((k_g_outside_cube<T>), dim3(32, nb), dim3(256), 0, stream, d, nb, flag);
        DNSBOX_LAUNCH_CHECK();
        unsigned int hf = 1;
        DNSBOX_HIP(hipMemcpyAsync(&hf, flag, sizeof(hf), hipMemcpyDeviceToHost, stream));
        DNSBOX_HIP(hipStreamSynchronize(stream));
        return hf == 0;
    }

    // IF-Heun steps of caller-owned states (device buffers of the boxes' shapes), on
    // the caller's stream; each box's step size and clock are its own.  The passes are
    // pruned when every state is confined to its kept cube (checked on the device).
    void step_ext(void* const* states, int nsteps) override
    {
        if (nsteps <= 0) return;
        refresh();
        std::vector<GBox<T>> h = h_gb;
        for (size_t i = 0; i < h.size(); ++i) h[i].st = static_cast<C*>(states[i]);
        const GBox<T>* d = descriptors(h);
        x6_valid = false;
        const bool pr = buffers_dealiased(d);
        last_ext_pruned = pr;
        if (pr) launch_zinv<true>(stream, d, 0);
        else launch_zinv<false>(stream, d, 0);
        for (int i = 0; i < nsteps; ++i) one_step(stream, d, pr);
    }

    // step_ext that also stores box i's state after step j + 1 at traj[i] + j * 3 fields (by the
    // corrector pass) and, when traj_pred is set, that step's predicted state u* at
    // traj_pred[i] + j * 3 fields
    void step_ext_record(void* const* states, int nsteps, void* const* traj, void* const* traj_pred) override
    {
        record(states, nsteps, traj, traj_pred, nullptr);
    }
    void step_ext_record_mixed(void* const* states, int nsteps, void* const* traj, void* const* traj_pred,
                               void* const* mixed) override
    {
        record(states, nsteps, traj, traj_pred, mixed);
    }
    void record(void* const* states, int nsteps, void* const* traj, void* const* traj_pred, void* const* mixed)
    {
        if (nsteps <= 0) return;
        refresh();
        std::vector<GBox<T>> h = h_gb;
        for (size_t i = 0; i < h.size(); ++i) h[i].st = static_cast<C*>(states[i]);
        x6_valid = false;
        const GBox<T>* d0 = descriptors(h);
        const bool pr = buffers_dealiased(d0);
        last_ext_pruned = pr;
        if (pr) {
            // the pruned predictor and corrector store u* and the new state at the kept modes only:
            // the rest of each slot is zero
            for (size_t i = 0; i < h.size(); ++i)
                for (void* const* bufs : {traj_pred, traj}) {
                    if (!bufs) continue;
                    hipLaunchKernelGGL((k_fill_zero<T>), dim3(256), dim3(256), 0, stream, static_cast<C*>(bufs[i]),
                                       (long long)nsteps * 3 * h[i].fs);
                    DNSBOX_LAUNCH_CHECK();
                }
            launch_zinv<true>(stream, d0, 0);
        } else {
            launch_zinv<false>(stream, d0, 0);
        }
        for (int j = 0; j < nsteps; ++j) {
            for (size_t i = 0; i < h.size(); ++i) {
                const long long off = (long long)j * 3 * h[i].fs;
                h[i].pout = traj_pred ? static_cast<C*>(traj_pred[i]) + off : nullptr;
                h[i].tout = static_cast<C*>(traj[i]) + off;
                h[i].mix = mixed ? static_cast<C*>(mixed[i]) + 6 * off : nullptr;   // 18 fields per step
            }
            const GBox<T>* d = descriptors(h);
            if (!mixed) one_step(stream, d, pr);
            else if (pr) one_step_rec<true>(stream, d);
            else one_step_rec<false>(stream, d);
        }
    }

    // out_i = (d u_{n+1} / d u_n)^T lam_i for one fixed-dt IF-Heun step of every box
    void step_vjp(const void* const* u, const void* const* lam, void* const* out, const double* dt) override
    {
        refresh();
        std::vector<GBox<T>> h = h_gb;
        for (size_t i = 0; i < h.size(); ++i) {
            h[i].u = static_cast<const C*>(u[i]);
            h[i].lam = static_cast<const C*>(lam[i]);
            h[i].out = static_cast<C*>(out[i]);
            h[i].dt = dt[i];
        }
        const GBox<T>* d = descriptors(h);
        const int nb = int(h.size());
        hipStream_t s = stream;
        x6_valid = false;
        // u* = E (u + dt N(u))
        launch_zinv(s, d, 1);
        hipLaunchKernelGGL((k_g_ycol<T, +1>), dim3(yt, 6), dim3(kGroupThreads), 0, s, d, nb);
        hipLaunchKernelGGL((k_g_xpass<T, OpCross<T>>), dim3(xb), dim3(kGroupThreads), 0, s, d, nb, 0);
        hipLaunchKernelGGL((k_g_ycol<T, -1>), dim3(yt, 3), dim3(kGroupThreads), 0, s, d, nb);
        hipLaunchKernelGGL((k_g_zfwd_predonly<T>), dim3(zt), dim3(kGroupThreads), 0, s, d, nb);
        DNSBOX_LAUNCH_CHECK();
        // lam_* = J(u*)^T (dt/2 lam); A = E (lam + lam_*), Mu = E (dt/2 lam + dt lam_*)
        vjp_passes<VJP_S1>(s, d);
        // out = A + J(u)^T Mu
        vjp_passes<VJP_S2>(s, d);
    }

    // step_vjp with each box's predicted state u* given (recorded by step_ext_record)
    void step_vjp_pred(const void* const* u, const void* const* ustar, const void* const* lam, void* const* out,
                       const double* dt) override
    {
        refresh();
        std::vector<GBox<T>> h = h_gb;
        for (size_t i = 0; i < h.size(); ++i) {
            h[i].u = static_cast<const C*>(u[i]);
            h[i].pred = const_cast<C*>(static_cast<const C*>(ustar[i]));   // read only by the VJP
            h[i].lam = static_cast<const C*>(lam[i]);
            h[i].out = static_cast<C*>(out[i]);
            h[i].dt = dt[i];
        }
        const GBox<T>* d = descriptors(h);
        x6_valid = false;
        vjp_passes<VJP_S1>(stream, d);
        vjp_passes<VJP_S2>(stream, d);
    }

    void step_vjp_mixed(const void* const* u, const void* const* ustar, const void* const* lam, void* const* out,
                        const double* dt, void* const* mixed, bool pruned_rec) override
    {
        bool pr = vjp_cube;
        for (auto& b : boxes) pr = pr && b->may_prune();
        if (pruned_rec && !pr)
            throw std::invalid_argument("a group step recorded with pruned passes holds its x-pass inputs at the kept "
                                        "kx only: its adjoint needs dnsbox_group_set_vjp_cube(g, 1)");
        refresh();
        std::vector<GBox<T>> h = h_gb;
        for (size_t i = 0; i < h.size(); ++i) {
            h[i].u = static_cast<const C*>(u[i]);
            h[i].pred = const_cast<C*>(static_cast<const C*>(ustar[i]));   // read only by the VJP
            h[i].lam = static_cast<const C*>(lam[i]);
            h[i].out = static_cast<C*>(out[i]);
            h[i].dt = dt[i];
            h[i].mix = static_cast<C*>(mixed[i]);
        }
        const GBox<T>* d = descriptors(h);
        x6_valid = false;
        if (pr) {
            // VJP_S1, whose z-forward also runs VJP_S2's first pass, then the rest of VJP_S2
            const int nb = int(h_gb.size());
            hipLaunchKernelGGL((k_g_zinv_vjp_rec<T, VJP_S1, true>), dim3(ztp + nb * kGroupCheck), dim3(kGroupThreads), 0,
                               stream, d, nb, ztp, boxes[0]->chk + 1);
            vjp_rec_middle<true>(stream, d, 1);
            hipLaunchKernelGGL((k_g_zfwd_vjp_next<T>), dim3(zt + nb * kGroupCheck), dim3(kGroupThreads), 0, stream, d,
                               nb, zt, boxes[0]->chk + 1);
            vjp_rec_middle<true>(stream, d, 0);
            hipLaunchKernelGGL((k_g_zfwd_vjp<T, VJP_S2>), dim3(zt), dim3(kGroupThreads), 0, stream, d, nb);
            DNSBOX_LAUNCH_CHECK();
        } else {
            vjp_rec_passes<VJP_S1, false>(stream, d);
            vjp_rec_passes<VJP_S2, false>(stream, d);
        }
    }
    // the middle passes of a recorded-input VJP on record slot `slot`: y-inverse of the cotangent,
    // x pass on it and the recorded inputs, y-forward of the products
    template <bool PR>
    void vjp_rec_middle(hipStream_t s, const GBox<T>* d, int slot)
    {
        const int nb = int(h_gb.size());
        hipLaunchKernelGGL((k_g_ycol_rec<T, +1, PR, 2>), dim3(PR ? ytp : yt, 3), dim3(kGroupThreads), 0, s, d, nb, slot);
        hipLaunchKernelGGL((k_g_xpass_rec<T, OpVjp<T>, PR, false, 2>), dim3(xb), dim3(kGroupThreads), 0, s, d, nb, 0,
                           slot);
        hipLaunchKernelGGL((k_g_ycol<T, -1>), dim3(yt, 6), dim3(kGroupThreads), 0, s, d, nb);
        DNSBOX_LAUNCH_CHECK();
    }
    // the VJP passes of the recorded-input step adjoint: the cotangent is transformed into fields
    // 0..2 of the stage's record slot (slot 1 at u* for VJP_S1, slot 0 at u for VJP_S2), whose
    // fields 3..8 hold the recorded x-pass inputs the x pass reads beside it
    template <int MODE, bool PR>
    void vjp_rec_passes(hipStream_t s, const GBox<T>* d)
    {
        const int nb = int(h_gb.size());
        const int slot = MODE == VJP_S1 ? 1 : 0;
        hipLaunchKernelGGL((k_g_zinv_vjp_rec<T, MODE, PR>), dim3(PR ? ztp + nb * kGroupCheck : zt),
                           dim3(kGroupThreads), 0, s, d, nb, PR ? ztp : zt, boxes[0]->chk + 1);
        hipLaunchKernelGGL((k_g_ycol_rec<T, +1, PR, 2>), dim3(PR ? ytp : yt, 3), dim3(kGroupThreads), 0, s, d, nb, slot);
        hipLaunchKernelGGL((k_g_xpass_rec<T, OpVjp<T>, PR, false, 2>), dim3(xb), dim3(kGroupThreads), 0, s, d, nb, 0,
                           slot);
        hipLaunchKernelGGL((k_g_ycol<T, -1>), dim3(yt, 6), dim3(kGroupThreads), 0, s, d, nb);
        hipLaunchKernelGGL((k_g_zfwd_vjp<T, MODE>), dim3(zt), dim3(kGroupThreads), 0, s, d, nb);
        DNSBOX_LAUNCH_CHECK();
    }

    // the VJP passes; with vjp_cube (every box's linearisation state confined to its kept cube,
    // as the caller guarantees) those up to the x pass's loads read the kept cube only, and the
    // first pass checks the states outside their cubes on kGroupCheck extra blocks per box
    template <int MODE>
    void vjp_passes(hipStream_t s, const GBox<T>* d)
    {
        bool pr = vjp_cube;
        for (auto& b : boxes) pr = pr && b->may_prune();
        if (pr) vjp_passes_t<MODE, true>(s, d);
        else vjp_passes_t<MODE, false>(s, d);
    }
    template <int MODE, bool PR>
    void vjp_passes_t(hipStream_t s, const GBox<T>* d)
    {
        const int nb = int(h_gb.size());
        hipLaunchKernelGGL((k_g_zinv_vjp<T, MODE, PR>), dim3(PR ? ztp + nb * kGroupCheck : zt), dim3(kGroupThreads), 0,
                           s, d, nb, PR ? ztp : zt, boxes[0]->chk + 1);
        hipLaunchKernelGGL((k_g_ycol<T, +1, PR>), dim3(PR ? ytp : yt, 9), dim3(kGroupThreads), 0, s, d, nb);
        hipLaunchKernelGGL((k_g_xpass<T, OpVjp<T>, PR, false>), dim3(xb), dim3(kGroupThreads), 0, s, d, nb, 0);
        hipLaunchKernelGGL((k_g_ycol<T, -1>), dim3(yt, 6), dim3(kGroupThreads), 0, s, d, nb);
        hipLaunchKernelGGL((k_g_zfwd_vjp<T, MODE>), dim3(zt), dim3(kGroupThreads), 0, s, d, nb);
        DNSBOX_LAUNCH_CHECK();
    }
};

}  // namespace dnsbox
