// dnsbox_rhs_dev_noalloc: rhs_nonlinear on caller device buffers with no context and no device
// allocation, for callers that cannot allocate (code under HIP graph capture, serving loops).
//
// The transform scratch, the twiddle table and the band-energy scalar live in one file-scope
// device array, `scratch`, of kScratchBytes per GPU (held until the process ends).  Its device
// address is looked up once per device with hipGetSymbolAddress and kept per device; host code
// uses that address, never the array's name.  Every call on a GPU shares that GPU's array, which
// keeps the last call's contents: nothing in it is read before the call writes it (twiddles, band
// energy and the scratch fields are all written first), and calls that overlap on one GPU race.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/dnsbox.h"
#include "box_impl.hpp"
#include "config.hpp"

namespace dnsbox {
void set_last_error(const std::string& msg);   // the message dnsbox_last_error() returns
}

namespace {

using namespace dnsbox;

constexpr int kMaxN = 64;   // largest box: 6 FP64 fields of 64^3 fill the array
constexpr size_t kHead = 256;                                   // BoxScalars
constexpr size_t kTw = 1024;                                    // kMaxN complex doubles
constexpr size_t kFields = 6 * size_t(kMaxN) * kMaxN * (kMaxN / 2 + 1) * 16;
constexpr size_t kScratchBytes = kHead + kTw + kFields;         // 12.98 MB
static_assert(sizeof(BoxScalars) <= kHead, "scalar block fits its slot");

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDev = 64;
std::mutex g_mtx;
void* g_ptr[kMaxDev] = {};   // device address of `scratch` on each device (null: not looked up yet)

// device address of this device's copy of `scratch`, looked up on the first call on the device
void* scratch_ptr(int dev)
{
    std::lock_guard<std::mutex> lock(g_mtx);
    if (!g_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) throw HipError(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_ptr[dev] = p;
    }
    return g_ptr[dev];
}

template <class T>
struct TwTable {
    cpx<T> w[kMaxN];
};

// the twiddle table of the box into the scratch array (passed by value: no host copy, no memset)
template <class T>
__global__ void k_put_twiddles(cpx<T>* __restrict__ dst, TwTable<T> tab, int n)
{
    for (int i = threadIdx.x; i < n; i += blockDim.x) dst[i] = tab.w[i];
}

template <int N, class T>
void rhs_noalloc_t(const BoxPhys& ph, const void* in, void* out, unsigned char* base, hipStream_t s)
{
    static_assert(N <= kMaxN, "box larger than the scratch array");
    static_assert(6 * Geo<Cube<N>, T>::FS * sizeof(cpx<T>) <= kFields, "fields fit the scratch array");
    BoxScalars* sc = reinterpret_cast<BoxScalars*>(base);
    cpx<T>* tw = reinterpret_cast<cpx<T>*>(base + kHead);
    cpx<T>* X = reinterpret_cast<cpx<T>*>(base + kHead + kTw);
    TwTable<T> tab{};
    host_twiddles<N, T>(tab.w);
    hipLaunchKernelGGL((k_put_twiddles<T>), dim3(1), dim3(64), 0, s, tw, tab, N);
    launch_rhs<Cube<N>, T>(s, ph, static_cast<const cpx<T>*>(in), static_cast<cpx<T>*>(out), X, tw, sc);
    DNSBOX_LAUNCH_CHECK();
}

template <class T>
void rhs_noalloc(int n, const BoxPhys& ph, const void* in, void* out, unsigned char* base, hipStream_t s)
{
    switch (n) {
        case 16: rhs_noalloc_t<16, T>(ph, in, out, base, s); break;
        case 32: rhs_noalloc_t<32, T>(ph, in, out, base, s); break;
        case 48: rhs_noalloc_t<48, T>(ph, in, out, base, s); break;
        default: rhs_noalloc_t<64, T>(ph, in, out, base, s); break;
    }
}

int fail(int code, const std::string& msg)
{
    set_last_error(msg);
    return code;
}

}  // namespace

extern "C" {

int dnsbox_rhs_dev_noalloc(const dnsbox_params* p, const void* in, void* out, void* stream)
{
    if (!p || !in || !out) return fail(DNSBOX_E_ARG, "null argument");
    if (!(p->n == 16 || p->n == 32 || p->n == 48 || p->n == 64))
        return fail(DNSBOX_E_ARG, "dnsbox_rhs_dev_noalloc: n must be 16, 32, 48 or 64 (its static scratch holds "
                                  "6 FP64 fields of 64^3, " + std::to_string(kScratchBytes) + " bytes); n = " +
                                      std::to_string(p->n));
    if (p->precision != 32 && p->precision != 64) return fail(DNSBOX_E_ARG, "precision must be 32 or 64");
    if (p->state_precision != 0 && p->state_precision != p->precision)
        return fail(DNSBOX_E_ARG, "dnsbox_rhs_dev_noalloc needs state_precision = precision");
    if (p->bandlim && !(p->kcut > 0.0)) return fail(DNSBOX_E_ARG, "bandlim needs kcut > 0");
    if (p->kolmo_k < 0 || p->kolmo_k >= p->n / 2) return fail(DNSBOX_E_ARG, "kolmo_k out of range");
    if ((p->nx && p->nx != p->n) || (p->ny && p->ny != p->n) || (p->nz && p->nz != p->n))
        return fail(DNSBOX_E_ARG, "nx, ny, nz must be 0 or n (cubic grids)");
    try {
        int dev = 0;
        DNSBOX_HIP(hipGetDevice(&dev));
        if (dev < 0 || dev >= kMaxDev) return fail(DNSBOX_E_ARG, "device index out of range");
        unsigned char* base = static_cast<unsigned char*>(scratch_ptr(dev));
        const BoxConfig bc = make_config(p);
        hipStream_t s = static_cast<hipStream_t>(stream);
        if (p->precision == 64) rhs_noalloc<double>(p->n, bc.ph, in, out, base, s);
        else rhs_noalloc<float>(p->n, bc.ph, in, out, base, s);
        return 0;
    } catch (const HipError& e) {
        return fail(DNSBOX_E_HIP, e.what());
    } catch (const std::exception& e) {
        return fail(DNSBOX_E_ARG, e.what());
    }
}

}  // extern "C"
