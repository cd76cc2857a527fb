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

#include <memory>
#include <vector>

#include "box_impl.hpp"

namespace dnsbox {

template <class T>
struct GBox {
    int n;
    int zt0, yt0, xb0;  // first global z tile, y tile and x block of this box
    long long fs;
    cpx<T>* st;
    cpx<T>* X;
    const cpx<T>* tw;
    BoxScalars* sc;
    BoxPhys ph;
};

constexpr int kGroupThreads = 256;

template <int N, class T>
constexpr bool group_size_ok()
{
    return Geo<N, T>::ZNT == kGroupThreads && Geo<N, T>::YNT == kGroupThreads &&
           Cfg<N, T>::NTX == kGroupThreads;
}

template <class T>
struct GroupLds {
    template <int N>
    static constexpr int z() { return N * Cfg<N, T>::ZW; }
    template <int N>
    static constexpr int y() { return N * Cfg<N, T>::YW; }
    template <int N>
    static constexpr int x() { return XGeo<N, T, OpCross<T>>::LDS; }
    static constexpr int mx(int a, int b) { return a > b ? a : b; }
    static constexpr int Z = mx(mx(mx(z<16>(), z<32>()), mx(z<48>(), z<64>())), z<128>());
    static constexpr int Y = mx(mx(mx(y<16>(), y<32>()), mx(y<48>(), y<64>())), y<128>());
    static constexpr int X = mx(mx(mx(x<16>(), x<32>()), mx(x<48>(), x<64>())), x<128>());
    static_assert(group_size_ok<16, T>() && group_size_ok<32, T>() && group_size_ok<48, T>() &&
                      group_size_ok<64, T>() && group_size_ok<128, T>(),
                  "group sizes must share one block shape");
};

// box owning global block `blk`, by the first-block offsets selected by `which`
template <class T>
__device__ __forceinline__ int find_box(const GBox<T>* __restrict__ gb, int nbox, int blk, int which)
{
    int lo = 0, hi = nbox - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        const int start = which == 0 ? gb[mid].zt0 : (which == 1 ? gb[mid].yt0 : gb[mid].xb0);
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

template <class T>
__global__ void __launch_bounds__(kGroupThreads) k_g_zinv(const GBox<T>* __restrict__ gb, int nbox)
{
    __shared__ cpx<T> lds[GroupLds<T>::Z];
    __shared__ cpx<T> stw[128];
    const int b = find_box(gb, nbox, blockIdx.x, 0);
    const GBox<T> bx = gb[b];
    const cpx<T>* tw = stage_tw_n(bx.tw, stw, bx.n);
    const int tile = blockIdx.x - bx.zt0;
    const cpx<T>* in[3] = {bx.st, bx.st + bx.fs, bx.st + 2 * bx.fs};
    cpx<T>* out[6];
#pragma unroll
    for (int f = 0; f < 6; ++f) out[f] = bx.X + f * bx.fs;
    const ProCurl<T> pro{T(1.0 / (double(bx.n) * bx.n * bx.n))};
#define DNSBOX_G_ZINV(NN) zinv_tile<NN, T, ProCurl<T>>(in, out, tile, lds, tw, pro)
    DNSBOX_GROUP_SWITCH(bx.n, DNSBOX_G_ZINV)
#undef DNSBOX_G_ZINV
}

template <class T, int DIR>
__global__ void __launch_bounds__(kGroupThreads) k_g_ycol(const GBox<T>* __restrict__ gb, int nbox)
{
    __shared__ cpx<T> lds[GroupLds<T>::Y];
    __shared__ cpx<T> stw[128];
    const int b = find_box(gb, nbox, blockIdx.x, 1);
    const GBox<T> bx = gb[b];
    const cpx<T>* tw = stage_tw_n(bx.tw, stw, bx.n);
    const int tile = blockIdx.x - bx.yt0;
    cpx<T>* data = bx.X + blockIdx.y * bx.fs;
#define DNSBOX_G_YCOL(NN) ycol_tile<NN, T, DIR>(data, tile, lds, tw)
    DNSBOX_GROUP_SWITCH(bx.n, DNSBOX_G_YCOL)
#undef DNSBOX_G_YCOL
}

template <class T>
__global__ void __launch_bounds__(kGroupThreads) k_g_xcross(const GBox<T>* __restrict__ gb, int nbox, int track)
{
    __shared__ cpx<T> lds[GroupLds<T>::X];
    __shared__ cpx<T> stw[128];
    __shared__ double wsm[4 * 16];
    const int b = find_box(gb, nbox, blockIdx.x, 2);
    const GBox<T> bx = gb[b];
    const cpx<T>* tw = stage_tw_n(bx.tw, stw, bx.n);
    const int blk = blockIdx.x - bx.xb0;
    XRed red;
#pragma unroll
    for (int j = 0; j < 4; ++j) red.m[j] = 0.0;
#define DNSBOX_G_X(NN) xpass_tile<NN, T, OpCross<T>>(bx.X, bx.fs, bx.X, bx.fs, blk, lds, tw, red)
    DNSBOX_GROUP_SWITCH(bx.n, DNSBOX_G_X)
#undef DNSBOX_G_X
    if (track) block_max_atomic<3>(red, bx.sc->maxu, false, wsm);
}

template <class T, int MODE>
__global__ void __launch_bounds__(kGroupThreads) k_g_zfwd_zinv(const GBox<T>* __restrict__ gb, int nbox)
{
    __shared__ cpx<T> lds[GroupLds<T>::Z];
    __shared__ cpx<T> stw[128];
    const int b = find_box(gb, nbox, blockIdx.x, 0);
    const GBox<T> bx = gb[b];
    const cpx<T>* tw = stage_tw_n(bx.tw, stw, bx.n);
    const int tile = blockIdx.x - bx.zt0;
    EpiArgs<T, T> ea{};
    ea.out = nullptr;
    ea.state = bx.st;
    ea.rhs_in = bx.st;
    ea.rhs_in_s = 1;
    ea.fs = bx.fs;
    ea.sc = bx.sc;
    ea.ph = bx.ph;
    const ProCurl<T> pro{T(1.0 / (double(bx.n) * bx.n * bx.n))};
#define DNSBOX_G_Z(NN) zfwd_zinv_tile<NN, T, MODE, T>(bx.X, bx.fs, tile, lds, tw, ea, pro)
    DNSBOX_GROUP_SWITCH(bx.n, DNSBOX_G_Z)
#undef DNSBOX_G_Z
}

template <class T>
__global__ void k_g_step_end(const GBox<T>* __restrict__ gb, int nbox)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbox) return;
    BoxScalars* sc = gb[b].sc;
    sc->t += sc->dt;
    sc->dt_prev = sc->dt;
    sc->maxu[0] = sc->maxu[1] = sc->maxu[2] = 0ull;
    sc->nstep += 1;
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
    virtual void step(int nsteps) = 0;
    virtual void invalidate() = 0;
};

template <class T>
class Group final : public GroupBase {
   public:
    GBox<T>* d_gb = nullptr;
    std::vector<GBox<T>> h_gb;
    int zt = 0, yt = 0, xb = 0;
    bool x6_valid = false;
    hipStream_t gstream = nullptr;
    hipGraphExec_t graph = nullptr;

    explicit Group(std::vector<std::unique_ptr<BoxBase>>&& bs)
    {
        boxes = std::move(bs);
        const int prec = sizeof(T) == 8 ? 64 : 32;
        for (auto& b : boxes) {
            GBox<T> g{};
            g.n = b->cfg.n;
            g.zt0 = zt;
            g.yt0 = yt;
            g.xb0 = xb;
            zt += group_ztiles(g.n, prec);
            yt += group_ytiles(g.n, prec);
            xb += group_xblocks(g.n, prec);
            g.fs = b->fs;
            g.st = static_cast<cpx<T>*>(b->state);
            g.X = static_cast<cpx<T>*>(b->scratch);
            g.tw = static_cast<const cpx<T>*>(b->tw);
            g.sc = b->sc;
            g.ph = b->cfg.ph;
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
        if (d_gb) (void)hipFree(d_gb);
    }

    void invalidate() override { x6_valid = false; }

    void one_step(hipStream_t s)
    {
        const int nb = int(h_gb.size());
        hipLaunchKernelGGL((k_g_ycol<T, +1>), dim3(yt, 6), dim3(kGroupThreads), 0, s, d_gb, nb);
        hipLaunchKernelGGL((k_g_xcross<T>), dim3(xb), dim3(kGroupThreads), 0, s, d_gb, nb, 1);
        hipLaunchKernelGGL((k_g_ycol<T, -1>), dim3(yt, 3), dim3(kGroupThreads), 0, s, d_gb, nb);
        hipLaunchKernelGGL((k_g_zfwd_zinv<T, EPI_PRED>), dim3(zt), dim3(kGroupThreads), 0, s, d_gb, nb);
        hipLaunchKernelGGL((k_g_ycol<T, +1>), dim3(yt, 6), dim3(kGroupThreads), 0, s, d_gb, nb);
        hipLaunchKernelGGL((k_g_xcross<T>), dim3(xb), dim3(kGroupThreads), 0, s, d_gb, nb, 0);
        hipLaunchKernelGGL((k_g_ycol<T, -1>), dim3(yt, 3), dim3(kGroupThreads), 0, s, d_gb, nb);
        hipLaunchKernelGGL((k_g_zfwd_zinv<T, EPI_CORR>), dim3(zt), dim3(kGroupThreads), 0, s, d_gb, nb);
        hipLaunchKernelGGL((k_g_step_end<T>), dim3((nb + 63) / 64), dim3(64), 0, s, d_gb, nb);
    }

    // IF-Heun steps of every box: the first z-inverse if needed, then the captured
    // step replayed.  Work runs on the group's own stream, ordered after the
    // caller's stream and before anything the caller enqueues next.
    void step(int nsteps) override
    {
        if (nsteps <= 0) return;
        hipEvent_t ev;
        DNSBOX_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        DNSBOX_HIP(hipEventRecord(ev, stream));
        DNSBOX_HIP(hipStreamWaitEvent(gstream, ev, 0));
        const int nb = int(h_gb.size());
        if (!x6_valid) {
            hipLaunchKernelGGL((k_g_zinv<T>), dim3(zt), dim3(kGroupThreads), 0, gstream, d_gb, nb);
            DNSBOX_LAUNCH_CHECK();
            x6_valid = true;
        }
        if (!graph) {
            hipGraph_t g;
            DNSBOX_HIP(hipStreamBeginCapture(gstream, hipStreamCaptureModeThreadLocal));
            one_step(gstream);
            DNSBOX_HIP(hipStreamEndCapture(gstream, &g));
            DNSBOX_HIP(hipGraphInstantiate(&graph, g, nullptr, nullptr, 0));
            DNSBOX_HIP(hipGraphDestroy(g));
        }
        for (int i = 0; i < nsteps; ++i) DNSBOX_HIP(hipGraphLaunch(graph, gstream));
        DNSBOX_HIP(hipEventRecord(ev, gstream));
        DNSBOX_HIP(hipStreamWaitEvent(stream, ev, 0));
        DNSBOX_HIP(hipEventDestroy(ev));
    }
};

}  // namespace dnsbox
