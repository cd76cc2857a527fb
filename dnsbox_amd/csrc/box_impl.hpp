// Host-side launchers of one box: owns the device buffers of a box and
// sequences the kernel passes of box.hpp / adjoint.hpp.  Instantiated per
// (grid Dims, FFT type T, state type S): cubic grids in box_f32.hip / box_f64.hip, the
// non-cubic grids of grids.hpp in box_grid_*.hip; the C-ABI (dnsbox_capi.hip) sees only BoxBase.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "adjoint.hpp"
#include "box.hpp"
#include "plane.hpp"
#include "xpass_reg.hpp"

namespace dnsbox {

struct HipError : std::runtime_error {
    explicit HipError(const std::string& s) : std::runtime_error(s) {}
};

#define DNSBOX_HIP(call)                                                                 \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess)                                                            \
            throw ::dnsbox::HipError(std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define DNSBOX_LAUNCH_CHECK() DNSBOX_HIP(hipGetLastError())

struct BoxConfig {
    int n = 32;                  // the cubic size (groups), or the dnsbox_params n of a non-cubic box
    int nx = 32, ny = 32, nz = 32;   // the grid
    int precision = 64;
    int state_precision = 64;
    BoxPhys ph{};
    double deltak = 1.0;
    double dt = 1e-3;
};

constexpr int kNPass = 6;  // passes reported by time_passes (DNSBOX_PASS_* order)

#ifndef DNSBOX_LEGACY_XPASS
#define DNSBOX_LEGACY_XPASS 0
#endif
// 1: the cross-product x pass uses the LDS-staged kernel (k_xpass) at every size
constexpr bool kLegacyXPass = DNSBOX_LEGACY_XPASS != 0;

// the fused plane pass (plane.hpp) for the sizes whose z-plane fits the LDS; DNSBOX_PLANE=0 in the
// environment runs the three separate passes instead (read once per process)
inline bool plane_enabled()
{
    static const bool on = [] {
        const char* e = std::getenv("DNSBOX_PLANE");
        return !(e && std::strcmp(e, "0") == 0);
    }();
    return on;
}
// the plane pass runs cubic boxes only
template <class Dm, class T, class Op>
constexpr bool plane_used()
{
    if constexpr (Dm::CUBIC) return PlaneGeo<Dm::NX, T, Op>::used;
    else return false;
}
template <class Dm, class T, class Op>
inline bool use_plane()
{
    return plane_used<Dm, T, Op>() && plane_enabled();
}
// the register x pass (xpass_reg.hpp) runs the x-lines of its lengths NX: every block takes LPB whole
// lines, and a line's byte offset within a field fits 32 bits
template <class Dm, class T>
constexpr bool xreg_used()
{
    if constexpr (XReg<Dm::NX, T>::enabled && !kLegacyXPass)
        return (Dm::NY * Dm::NZ) % xreg::Shape<Dm::NX>::LPB == 0 &&
               Dm::FS * (long long)sizeof(cpx<T>) < (1LL << 32);
    else return false;
}

// twiddle table tw[j] = e^{-2 pi i j / N} in the transform type, from long-double angles
template <int N, class T>
void host_twiddles(cpx<T>* h)
{
    for (int j = 0; j < N; ++j) {
        long double a = 6.283185307179586476925286766559005768L * (long double)j / (long double)N;
        h[j].x = T(std::cos(a));
        h[j].y = T(-std::sin(a));
    }
}
// the twiddle tables of a box of grid Dm: one table of N entries for a cubic grid, else the
// tables of the x, y and z lengths back to back (offsets TwOff<Dm>)
template <class Dm>
struct TwOff {
    static constexpr int X = 0;
    static constexpr int Y = Dm::CUBIC ? 0 : Dm::NX;
    static constexpr int Z = Dm::CUBIC ? 0 : Dm::NX + Dm::NY;
    static constexpr int LEN = Dm::CUBIC ? Dm::NX : Dm::NX + Dm::NY + Dm::NZ;
};
template <class Dm, class T>
void host_box_twiddles(cpx<T>* h)
{
    host_twiddles<Dm::NX, T>(h + TwOff<Dm>::X);
    if constexpr (!Dm::CUBIC) {
        host_twiddles<Dm::NY, T>(h + TwOff<Dm>::Y);
        host_twiddles<Dm::NZ, T>(h + TwOff<Dm>::Z);
    }
}

// rhs_nonlinear (rhs.f90:18-135) on device buffers of the transform type: out <- N(in), with the
// transform scratch X (6 fields), the twiddles tw and the scalars sc (band energy) given.  The
// five passes of one RHS: z-inverse + curl, y-inverse (6), x cross product, y-forward (3),
// z-forward with the mask / projection / forcing epilogue.  tw: the box's twiddle tables (TwOff).
template <class Dm, class T>
void launch_rhs(hipStream_t s, const BoxPhys& ph, const cpx<T>* in, cpx<T>* out, cpx<T>* X, const cpx<T>* tw,
                BoxScalars* sc)
{
    using G = Geo<Dm, T>;
    using TO = TwOff<Dm>;
    constexpr long long fs = G::FS;
    const double sm = 1.0 / (double(G::NX) * G::NY * G::NZ);
    if (ph.bandlim) hipLaunchKernelGGL((k_eband<Dm, T>), dim3(1), dim3(256), 0, s, in, fs, ph, sm * sm, sc);
    hipLaunchKernelGGL((k_zinv<Dm, T, ProCurl<T>, T, false>), dim3(G::ZTILES), dim3(G::ZNT), 0, s, in, fs, X, fs,
                       tw + TO::Z, ProCurl<T>{T(sm), KScale<T>::of(ph)});
    hipLaunchKernelGGL((k_ycol<Dm, T, +1>), dim3(G::YTILES, 6), dim3(G::YNT), 0, s, X, fs, tw + TO::Y);
    hipLaunchKernelGGL((k_xpass<Dm, T, OpCross<T>>), dim3(G::XBLOCKS), dim3(G::NTX), 0, s, X, fs, X, fs, tw + TO::X,
                       (unsigned long long*)nullptr, 0);
    hipLaunchKernelGGL((k_ycol<Dm, T, -1>), dim3(G::YTILES, 3), dim3(G::YNT), 0, s, X, fs, tw + TO::Y);
    EpiArgs<T, T> ea{};
    ea.out = out;
    ea.rhs_in = in;
    ea.rhs_in_s = 0;
    ea.fs = fs;
    ea.sc = sc;
    ea.ph = ph;
    hipLaunchKernelGGL((k_zfwd<Dm, T, EPI_RHS, T>), dim3(G::ZTILES), dim3(G::ZNT), 0, s, X, fs, tw + TO::Z, ea);
}

// Per-box device buffers and the operations the C-ABI exposes.
class BoxBase {
   public:
    virtual ~BoxBase() = default;

    BoxConfig cfg;
    hipStream_t stream = nullptr;
    int nspec = 0;
    long long fs = 0;        // complex elements per spectral field
    size_t cbytes = 0;       // bytes per complex element of the transforms
    size_t sbytes = 0;       // bytes per complex element of the state
    size_t rbytes = 0;       // bytes per real element of the transforms
    bool uniform = true;     // state type == transform type (device-tensor entry points need it)

    // device buffers
    void* state = nullptr;   // 3 fields (state type)
    void* pred = nullptr;    // 3 fields (IF-Heun predicted state)
    void* scratch = nullptr; // 6 fields (9 for the VJP, 12 for rhsAdj)
    void* nonlin = nullptr;  // 3 fields
    void* nold = nullptr;    // 3 fields (AB2 history), lazily
    void* tw = nullptr;      // twiddle table
    BoxScalars* sc = nullptr;
    double* red_partial = nullptr;
    double* red_out = nullptr;
    unsigned long long* red_bits = nullptr;  // 4 slots
    int red_blocks = 0;
    bool ab2_started = false;
    bool x6_valid = false;   // scratch holds the z-inverse + curl of the current state
    bool x6_pruned = false;  // ... written by the pruned passes (columns outside the kept cube stale)
    // the state has no energy outside the 2/3-rule cube (Keep in box.hpp): true for the zero
    // state, after init_random, state_dealias and every step; set_state checks the field
    bool dealiased = true;
    bool prune = true;       // run the step's passes pruned to the kept cube when dealiased
    // the mask removes every mode outside the kept cube (false only for box lengths whose 2/3-rule
    // threshold falls at an index of the kept range's edge); pruning needs it
    bool prunable = true;
    bool may_prune() const { return prune && prunable; }
    bool pruned() const { return may_prune() && dealiased; }
    void* pred_out_ = nullptr;   // where the IF-Heun step stores u* (null: internal buffer)
    void* traj_out_ = nullptr;   // where the IF-Heun corrector also stores the new state (null: nowhere)
    // two record slots of 9 fields: the step's two RHS write their x-pass inputs (u, v, w and the
    // vorticity after the z- and y-inverse transforms) to fields 3..8 of slot 0 and of slot 1
    void* mixed_out_ = nullptr;
    // the caller guarantees that the states the step / RHS adjoints linearise about have no
    // energy outside the kept cube (dnsbox_set_vjp_cube); their first passes then read the cube only
    bool vjp_cube = false;
    bool last_ext_pruned = false;   // the last step of a caller's buffer ran the pruned passes
    int scratch_fields = 6;
    void* adjbuf = nullptr;  // 9 fields: A, Mu, u~ (adjoint passes), lazily
    void* descbuf = nullptr; // 9 fields: G(u), u*, G(u*) of the adjoint-descent step, lazily
    double* band2 = nullptr; // band-forcing scalars of the VJP
    unsigned int* chk = nullptr;   // [0]: flag of k_outside_cube; [1]: cube-confined VJP guard

    // dnsbox_set_vjp_cube: setting the guarantee clears the device flag that the cube-confined
    // VJPs raise for a linearisation state with energy outside the kept cube; withdrawing it
    // reads the flag back (synchronises the stream) and returns whether such a state was seen
    bool set_vjp_cube(bool on)
    {
        if (on) {
            if (!vjp_cube) DNSBOX_HIP(hipMemsetAsync(chk + 1, 0, sizeof(unsigned int), stream));
            vjp_cube = true;
            return false;
        }
        if (!vjp_cube) return false;
        vjp_cube = false;
        unsigned int h = 0;
        DNSBOX_HIP(hipMemcpyAsync(&h, chk + 1, sizeof(h), hipMemcpyDeviceToHost, stream));
        DNSBOX_HIP(hipStreamSynchronize(stream));
        return h != 0;
    }

    virtual void rhs_state() = 0;                 // nonlin <- N(state)
    virtual void rhs(const void* in, void* out) = 0;
    virtual void heun_step() = 0;
    // nsteps IF-Heun steps; the steady-state step is replayed from a hipGraph
    virtual void heun_steps(int nsteps) = 0;
    virtual void ab2_step() = 0;
    virtual void spec_reduce() = 0;               // state -> red_out[DG_NSCAL + nspec]
    virtual void phys_reduce() = 0;               // state -> red_bits[0..3]
    virtual void init_random(unsigned long long seed, int box, double k0, double energy) = 0;
    virtual void project(int dealias, int proj) = 0;
    virtual void to_physical(void* phys3) = 0;    // state -> device real (3, N^3)
    virtual void from_physical(const void* phys3) = 0;
    // average device time of each step pass over `reps` back-to-back launches (ms),
    // measured with HIP events on the box stream; order of DNSBOX_PASS_*
    virtual void time_passes(int reps, double* ms) = 0;
    // algorithmic HBM bytes of one launch of each step pass (DNSBOX_PASS_* order): every field
    // element a pass must read or write, once, for the passes the next step would run
    virtual void pass_bytes(double* bytes) const = 0;
    // adjoint and caller-buffer entry points (device pointers, internal layout)
    virtual void rhs_vjp(const void* u, const void* lam, void* out) = 0;
    virtual void step_vjp(const void* u, const void* lam, void* lam_out, double dt) = 0;
    virtual void rhs_adj(const void* u, void* out, double lin_sign) = 0;
    // the adjoint-descent loop of dnsboxAdj.f90:76-126 on a device state (null: the box's state):
    // nsteps Heun steps of size dtau in fictitious time on rhsAdj, each followed by state_project
    virtual void adj_descent(void* u, double dtau, int nsteps, double lin_sign) = 0;
    virtual void step_ext(void* st, int nsteps) = 0;
    // step_ext that also stores the state after step j + 1 in traj + j * (3 state fields) and,
    // when traj_pred is set, the predicted state u* of step j + 1 in traj_pred + j * 3 fields
    virtual void step_ext_record(void* st, int nsteps, void* traj, void* traj_pred) = 0;
    // step_vjp at u with the step's predicted state u* = E (u + dt N(u)) given
    virtual void step_vjp_pred(const void* u, const void* ustar, const void* lam, void* lam_out, double dt) = 0;
    // step_ext_record that also records the x-pass inputs of the two RHS of step j + 1 in fields 3..8
    // of the 9-field slots mixed + 2 j and mixed + 2 j + 1 (transform type, internal layout)
    virtual void step_ext_record_mixed(void* st, int nsteps, void* traj, void* traj_pred, void* mixed) = 0;
    // step_vjp_pred with the x-pass inputs of the step's RHS at u and at u* read from the slots
    // mixed_u, mixed_us that recorded them (their fields 0..2 are overwritten); pruned_rec: the
    // recording step ran pruned, so the slots hold the kept kx only and vjp_cube must be set
    virtual void step_vjp_mixed(const void* u, const void* ustar, const void* lam, void* lam_out, double dt,
                                void* mixed_u, void* mixed_us, bool pruned_rec) = 0;
    // step_vjp_mixed over the recorded steps n-1 .. 0 of a segment (u[j], ustar[j], slots
    // mixed[2j], mixed[2j+1]); the result lands in lam_out, lam is work space
    virtual void step_vjp_mixed_seq(const void* const* u, const void* const* ustar, void* lam, void* lam_out,
                                    double dt, void* const* mixed, int n, bool pruned_rec) = 0;
    // FP64 device array -> transform type / state type, and back
    virtual void convert_in(void* dst, const double* src, long long n, bool state_type) = 0;
    virtual void convert_out(double* dst, const void* src, long long n, bool state_type) = 0;
    // one spectral field: the reference layout (FP64 complex, device memory) -> dst in the device
    // layout (state type or transform type), and back
    virtual void from_ref_layout(void* dst, const double* src, bool state_type) = 0;
    virtual void to_ref_layout(double* dst, const void* src, bool state_type) = 0;
    // the state has no energy outside the kept cube (checked on the device; synchronises)
    virtual bool state_confined() = 0;
    // allocate at least nf scratch fields (and the adjoint buffers) now
    virtual void reserve(int nf, bool adj) = 0;
};

BoxBase* make_box_f32(const BoxConfig& c);
BoxBase* make_box_f64(const BoxConfig& c);

template <class Dm, class T, class S = T>
class BoxT final : public BoxBase {
    using G = Geo<Dm, T>;
    static constexpr int NX = Dm::NX, NY = Dm::NY, NZ = Dm::NZ;
    using C = cpx<T>;
    using CS = cpx<S>;
    static constexpr int ZNT = G::ZNT;
    static constexpr int YNT = G::YNT;
    static constexpr bool UNIFORM = std::is_same<S, T>::value;

   public:
    explicit BoxT(const BoxConfig& c)
    {
        cfg = c;
        fs = G::FS;
        cbytes = sizeof(C);
        sbytes = sizeof(CS);
        rbytes = sizeof(T);
        uniform = UNIFORM;
        // variables.f90:112-123: Nspec from the 2/3-rule extent of each axis
        const double* al = cfg.ph.a;
        const double ex = (double(NX) / 2.0) * (2.0 * al[0] / 3.0), ey = (double(NY) / 2.0) * (2.0 * al[1] / 3.0),
                     ez = (double(NZ) / 2.0) * (2.0 * al[2] / 3.0);
        if (cfg.ph.spherical) nspec = int(ex / cfg.deltak) - 1;
        else nspec = int(std::sqrt(ex * ex + ey * ey + ez * ez) / cfg.deltak) - 1;
        if (nspec < 1) nspec = 1;
        // the pruned passes skip |kval| > KMAX; that is exact when the mask removes index KMAX + 1
        // on every axis (|kval| alpha_d >= (N_d/2)(2 alpha_d/3), evaluated as the kernels do)
        const int kmax[3] = {Keep<NX>::KMAX, Keep<NY>::KMAX, Keep<NZ>::KMAX};
        for (int d = 0; d < 3; ++d)
            if (!(std::fabs(double(kmax[d] + 1) * al[d]) >= cfg.ph.thr[d])) prunable = false;
        const size_t fb = size_t(fs) * sizeof(C);
        try {
            DNSBOX_HIP(hipMalloc(&state, 3 * size_t(fs) * sizeof(CS)));
            DNSBOX_HIP(hipMalloc(&pred, 3 * fb));
            DNSBOX_HIP(hipMalloc(&scratch, 6 * fb));
            DNSBOX_HIP(hipMalloc(&nonlin, 3 * fb));
            DNSBOX_HIP(hipMalloc(&tw, TwOff<Dm>::LEN * sizeof(C)));
            DNSBOX_HIP(hipMalloc(&sc, sizeof(BoxScalars)));
            red_blocks = 512;
            DNSBOX_HIP(hipMalloc(&red_partial, size_t(red_blocks) * (DG_NSCAL + nspec) * sizeof(double)));
            DNSBOX_HIP(hipMalloc(&red_out, (DG_NSCAL + nspec) * sizeof(double)));
            DNSBOX_HIP(hipMalloc(&red_bits, 4 * sizeof(unsigned long long)));
            DNSBOX_HIP(hipMalloc(&chk, 2 * sizeof(unsigned int)));
            DNSBOX_HIP(hipMemset(state, 0, 3 * size_t(fs) * sizeof(CS)));
            DNSBOX_HIP(hipMemset(pred, 0, 3 * fb));
            DNSBOX_HIP(hipMemset(nonlin, 0, 3 * fb));
            std::vector<C> h(TwOff<Dm>::LEN);
            host_box_twiddles<Dm, T>(h.data());
            DNSBOX_HIP(hipMemcpy(tw, h.data(), h.size() * sizeof(C), hipMemcpyHostToDevice));
            BoxScalars s{};
            s.dt = s.dt_prev = cfg.dt;
            DNSBOX_HIP(hipMemcpy(sc, &s, sizeof(s), hipMemcpyHostToDevice));
        } catch (...) {
            release();
            throw;
        }
    }
    ~BoxT() override { release(); }

    void release()
    {
        drop_graph();
        if (gstream) {
            (void)hipStreamSynchronize(gstream);
            (void)hipStreamDestroy(gstream);
            gstream = nullptr;
        }
        void* ps[] = {state, pred, scratch, nonlin, nold, tw, sc, red_partial, red_out, red_bits, adjbuf, band2, chk,
                      descbuf};
        for (void* p : ps)
            if (p) (void)hipFree(p);
        state = pred = scratch = nonlin = nold = tw = adjbuf = descbuf = nullptr;
        band2 = nullptr;
        sc = nullptr;
        red_partial = red_out = nullptr;
        red_bits = nullptr;
        chk = nullptr;
    }

    const C* twc() const { return static_cast<const C*>(tw); }
    const C* twx() const { return twc() + TwOff<Dm>::X; }   // x-pass twiddles (length NX)
    const C* twy() const { return twc() + TwOff<Dm>::Y; }   // y passes (NY)
    const C* twz() const { return twc() + TwOff<Dm>::Z; }   // z passes (NZ)
    ProCurl<T> curl() const { return ProCurl<T>{inv_n3(), KScale<T>::of(cfg.ph)}; }
    C* X() const { return static_cast<C*>(scratch); }
    CS* St() const { return static_cast<CS*>(state); }
    C* Pr() const { return static_cast<C*>(pred); }
    static T inv_n3() { return T(1.0 / (double(NX) * NY * NZ)); }

    void require_uniform(const char* what) const
    {
        if (!UNIFORM)
            throw std::invalid_argument(std::string(what) +
                                        " needs the state in the transform precision (state_precision = precision)");
    }

    EpiArgs<T, S> epi() const
    {
        EpiArgs<T, S> ea{};
        ea.fs = fs;
        ea.sc = sc;
        ea.ph = cfg.ph;
        return ea;
    }

    template <class E>
    void eband(const cpx<E>* u)
    {
        if (!cfg.ph.bandlim) return;
        const double sm = 1.0 / (double(NX) * NY * NZ);
        hipLaunchKernelGGL((k_eband<Dm, E>), dim3(1), dim3(256), 0, stream, u, fs, cfg.ph, sm * sm, sc);
        DNSBOX_LAUNCH_CHECK();
    }

    template <class E, bool PR = false>
    void zinv_curl(const cpx<E>* in)
    {
        hipLaunchKernelGGL((k_zinv<Dm, T, ProCurl<T>, E, PR>), dim3(ztiles<Dm, T, PR>()), dim3(ZNT), 0, stream, in,
                           fs, X(), fs, twz(), curl());
        DNSBOX_LAUNCH_CHECK();
    }

    // pass markers of time_passes: an event after each launch of the step, tagged with
    // its DNSBOX_PASS_* id (-1: not a pass); inactive (null) outside time_passes
    std::vector<std::pair<int, hipEvent_t>>* marks = nullptr;
    void mark(int pass)
    {
        if (!marks) return;
        hipEvent_t e;
        DNSBOX_HIP(hipEventCreate(&e));
        marks->emplace_back(pass, e);
        DNSBOX_HIP(hipEventRecord(e, stream));
    }

    // middle passes of an RHS: y-inverse (6), x cross product, y-forward (3).  rec: a 9-field
    // record slot whose fields 3..8 receive the y-inverse's output, which the x pass reads there
    template <bool PR = false>
    void rhs_middle(bool track_cfl, C* rec = nullptr)
    {
        if constexpr (plane_used<Dm, T, OpCross<T>>()) {
            if (plane_enabled() && !rec) {
                hipLaunchKernelGGL((k_plane<NX, T, OpCross<T>, PR, PR>), dim3(NX), dim3(kPlaneThreads), 0, stream, X(),
                                   fs, twc(), track_cfl ? sc->maxu : nullptr);
                DNSBOX_LAUNCH_CHECK();
                mark(2);
                return;
            }
        }
        const C* xin = X();
        if (rec) {
            xin = rec + 3 * fs;
            hipLaunchKernelGGL((k_ycol_oop<Dm, T, +1, PR>), dim3(ytiles<Dm, T, PR>(), 6), dim3(YNT), 0, stream, X(),
                               rec + 3 * fs, fs, twy());
        } else {
            hipLaunchKernelGGL((k_ycol<Dm, T, +1, PR>), dim3(ytiles<Dm, T, PR>(), 6), dim3(YNT), 0, stream, X(), fs,
                               twy());
        }
        DNSBOX_LAUNCH_CHECK();
        mark(1);
        if constexpr (xreg_used<Dm, T>()) {
            hipLaunchKernelGGL((xreg::k_xcross_r<NX, T, PR>), dim3(NY * NZ / xreg::Shape<NX>::LPB),
                               dim3(xreg::kThreads), 0, stream, xin, X(), fs, twx(), track_cfl ? sc->maxu : nullptr);
        } else {
            hipLaunchKernelGGL((k_xpass<Dm, T, OpCross<T>, PR>), dim3(G::XBLOCKS), dim3(G::NTX), 0, stream, xin, fs,
                               X(), fs, twx(), track_cfl ? sc->maxu : nullptr, 0);
        }
        DNSBOX_LAUNCH_CHECK();
        mark(2);
        hipLaunchKernelGGL((k_ycol<Dm, T, -1, PR>), dim3(ytiles<Dm, T, PR>(), 3), dim3(YNT), 0, stream, X(), fs,
                           twy());
        DNSBOX_LAUNCH_CHECK();
        mark(3);
    }

    // five passes of one RHS with an unfused z-forward epilogue of MODE
    template <int MODE, class E>
    void rhs_passes(const cpx<E>* in, const EpiArgs<T, S>& ea, bool track_cfl)
    {
        zinv_curl(in);
        rhs_middle(track_cfl);
        hipLaunchKernelGGL((k_zfwd<Dm, T, MODE, S>), dim3(G::ZTILES), dim3(ZNT), 0, stream, X(), fs, twz(), ea);
        DNSBOX_LAUNCH_CHECK();
    }

    void rhs_state() override
    {
        x6_valid = false;
        eband(St());
        EpiArgs<T, S> ea = epi();
        ea.out = static_cast<C*>(nonlin);
        ea.rhs_in = St();
        ea.rhs_in_s = 1;
        rhs_passes<EPI_RHS>(St(), ea, false);
    }

    void rhs(const void* in, void* out) override
    {
        require_uniform("rhs on device buffers");
        x6_valid = false;
        launch_rhs<Dm, T>(stream, cfg.ph, static_cast<const C*>(in), static_cast<C*>(out), X(), twc(), sc);
        DNSBOX_LAUNCH_CHECK();
    }

    // IF-Heun step (dnsboxheun.f90:84-126) as a pipeline: the forward z pass of each
    // RHS carries the stepper epilogue and the z-inverse + curl of the next RHS.  The passes
    // are pruned to the 2/3-rule cube while the state has no energy outside it.
    void heun_step() override
    {
        if (pruned()) heun_step_t<true>();
        else heun_step_t<false>();
    }

    // the internal predicted-state buffer holds no value outside the kept cube (the pruned
    // predictor writes only the kept modes, and the band energy of u* reads the band's modes)
    bool pred_clean = true;
    void clear_pred()
    {
        hipLaunchKernelGGL((k_fill_zero<T>), dim3(1024), dim3(256), 0, stream, Pr(), 3 * fs);
        DNSBOX_LAUNCH_CHECK();
        pred_clean = true;
    }

    template <bool PR>
    void heun_step_t()
    {
        const ProCurl<T> pro = curl();
        if (cfg.ph.bandlim && !pred_out_) {
            if (PR && !pred_clean) clear_pred();
            if (!PR) pred_clean = false;
        }
        if (!x6_valid || x6_pruned != PR) {
            eband(St());
            zinv_curl<S, PR>(St());
            mark(0);
        }
        C* rec = static_cast<C*>(mixed_out_);
        rhs_middle<PR>(true, rec);
        EpiArgs<T, S> ea = epi();
        // the predicted state is kept for band forcing, or in the caller's buffer when recorded
        C* pst = pred_out_ ? static_cast<C*>(pred_out_) : Pr();
        ea.out = (cfg.ph.bandlim || pred_out_) ? pst : nullptr;
        ea.state = St();
        ea.rhs_in = St();
        ea.rhs_in_s = 1;
        hipLaunchKernelGGL((k_zfwd_zinv<Dm, T, EPI_PRED, S, PR>), dim3(ztiles<Dm, T, PR>()), dim3(ZNT), 0, stream,
                           X(), fs, twz(), ea, pro);
        DNSBOX_LAUNCH_CHECK();
        mark(4);
        eband(pst);
        rhs_middle<PR>(false, rec ? rec + 9 * fs : nullptr);
        EpiArgs<T, S> eb = epi();
        eb.state = St();
        eb.rhs_in = pst;
        eb.rhs_in_s = 0;
        eb.tout = static_cast<CS*>(traj_out_);
        hipLaunchKernelGGL((k_zfwd_zinv<Dm, T, EPI_CORR, S, PR>), dim3(ztiles<Dm, T, PR>()), dim3(ZNT), 0, stream,
                           X(), fs, twz(), eb, pro);
        DNSBOX_LAUNCH_CHECK();
        mark(5);   // the corrector pass also ended the step (t, dt_prev, CFL maxima)
        eband(St());
        mark(-1);
        x6_valid = true;
        x6_pruned = PR;
    }

    // Optional replay of the steady-state IF-Heun steps from a hipGraph (DNSBOX_STEP_GRAPH=1
    // in the environment): graph_steps steps captured once on an internal stream, ordered
    // after the box stream and before its next work with events; the first step after a
    // pipeline restart and the remainder run eagerly.  The graph holds the buffer pointers it
    // was captured with and is re-captured when they change.  Off by default for single
    // boxes: on MI355X the replay measured no faster than stream launches (32^3: 21.0 k vs
    // 22.2 k steps/s, 128^3: 3.02 k vs 3.05 k).
    hipStream_t gstream = nullptr;
    hipGraphExec_t step_graph = nullptr;
    const void* graph_key[4] = {nullptr, nullptr, nullptr, nullptr};
    bool graph_pr = false;  // the captured steps run the pruned passes
    int graph_enabled = -1;
    int graph_steps = 16;   // steps per captured graph (one hipGraphLaunch replays them all)

    void drop_graph()
    {
        if (step_graph) (void)hipGraphExecDestroy(step_graph);
        step_graph = nullptr;
    }

    void heun_steps(int nsteps) override
    {
        if (graph_enabled < 0) {
            const char* e = std::getenv("DNSBOX_STEP_GRAPH");
            graph_enabled = (e && std::strcmp(e, "1") == 0) ? 1 : 0;
            const char* k = std::getenv("DNSBOX_GRAPH_STEPS");
            if (k && std::atoi(k) > 0) graph_steps = std::atoi(k);
        }
        if (nsteps > 0 && (!x6_valid || x6_pruned != pruned())) {
            heun_step();
            --nsteps;
        }
        if (nsteps <= 0) return;
        if (!graph_enabled || marks || nsteps < graph_steps) {
            for (int i = 0; i < nsteps; ++i) heun_step();
            return;
        }
        const void* key[4] = {state, scratch, pred, sc};
        if (step_graph && (std::memcmp(key, graph_key, sizeof(key)) != 0 || graph_pr != pruned())) drop_graph();
        if (!gstream) DNSBOX_HIP(hipStreamCreate(&gstream));
        hipEvent_t ev;
        DNSBOX_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        DNSBOX_HIP(hipEventRecord(ev, stream));
        DNSBOX_HIP(hipStreamWaitEvent(gstream, ev, 0));
        if (!step_graph) {
            hipStream_t own = stream;
            stream = gstream;
            hipGraph_t g = nullptr;
            try {
                DNSBOX_HIP(hipStreamBeginCapture(gstream, hipStreamCaptureModeThreadLocal));
                for (int i = 0; i < graph_steps; ++i) heun_step();
                DNSBOX_HIP(hipStreamEndCapture(gstream, &g));
            } catch (...) {
                stream = own;
                hipGraph_t dummy = nullptr;
                (void)hipStreamEndCapture(gstream, &dummy);
                if (dummy) (void)hipGraphDestroy(dummy);
                (void)hipEventDestroy(ev);
                throw;
            }
            stream = own;
            const hipError_t ie = hipGraphInstantiate(&step_graph, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            DNSBOX_HIP(ie);
            std::memcpy(graph_key, key, sizeof(key));
            graph_pr = pruned();
        }
        for (int i = 0; i + graph_steps <= nsteps; i += graph_steps) DNSBOX_HIP(hipGraphLaunch(step_graph, gstream));
        DNSBOX_HIP(hipEventRecord(ev, gstream));
        DNSBOX_HIP(hipStreamWaitEvent(stream, ev, 0));
        DNSBOX_HIP(hipEventDestroy(ev));
        for (int i = 0; i < nsteps % graph_steps; ++i) heun_step();
    }

    void ab2_step() override
    {
        x6_valid = false;
        if (!nold) DNSBOX_HIP(hipMalloc(&nold, 3 * size_t(fs) * sizeof(C)));
        if (!ab2_started) {
            // first step: IF-Heun, keeping N(u_0) as the AB2 history (dnsbox.f90:87-141)
            eband(St());
            EpiArgs<T, S> ea = epi();
            ea.out = Pr();
            ea.state = St();
            ea.rhs_in = St();
            ea.rhs_in_s = 1;
            ea.nold = static_cast<C*>(nold);
            pred_clean = false;
            rhs_passes<EPI_PRED>(St(), ea, true);
            eband(Pr());
            EpiArgs<T, S> eb = epi();
            eb.state = St();
            eb.rhs_in = Pr();
            eb.rhs_in_s = 0;
            rhs_passes<EPI_CORR>(Pr(), eb, false);
            ab2_started = true;
        } else {
            eband(St());
            EpiArgs<T, S> ea = epi();
            ea.state = St();
            ea.rhs_in = St();
            ea.rhs_in_s = 1;
            ea.nold = static_cast<C*>(nold);
            rhs_passes<EPI_AB2>(St(), ea, true);
        }
        hipLaunchKernelGGL(k_step_end<0>, dim3(1), dim3(64), 0, stream, sc);
        DNSBOX_LAUNCH_CHECK();
    }

    void spec_reduce() override
    {
        const double sm = 1.0 / (double(NX) * NY * NZ);
        const int nslot = DG_NSCAL + nspec;
        hipLaunchKernelGGL((k_spec_reduce<Dm, S>), dim3(red_blocks), dim3(256), spec_reduce_lds(nspec), stream,
                           St(), fs, cfg.ph, cfg.deltak, nspec, red_partial);
        DNSBOX_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_spec_finish<0>, dim3(1), dim3(256), 0, stream, red_partial, red_blocks, nslot,
                           sm * sm, cfg.ph.nu, cfg.deltak, red_out, (BoxScalars*)nullptr);
        DNSBOX_LAUNCH_CHECK();
    }

    void phys_reduce() override
    {
        x6_valid = false;
        DNSBOX_HIP(hipMemsetAsync(red_bits, 0, 4 * sizeof(unsigned long long), stream));
        hipLaunchKernelGGL((k_zinv<Dm, T, ProVelDiv<T>, S>), dim3(G::ZTILES), dim3(ZNT), 0, stream, St(), fs, X(),
                           fs, twz(), ProVelDiv<T>{inv_n3(), KScale<T>::of(cfg.ph)});
        DNSBOX_LAUNCH_CHECK();
        hipLaunchKernelGGL((k_ycol<Dm, T, +1>), dim3(G::YTILES, 4), dim3(YNT), 0, stream, X(), fs, twy());
        DNSBOX_LAUNCH_CHECK();
        hipLaunchKernelGGL((k_xpass<Dm, T, OpVelDivMax<T>>), dim3(G::XBLOCKS), dim3(G::NTX), 0, stream, X(), fs, X(),
                           fs, twx(), red_bits, 1);
        DNSBOX_LAUNCH_CHECK();
    }

    void init_random(unsigned long long seed, int box, double k0, double energy) override
    {
        x6_valid = false;
        dealiased = true;   // the generator leaves every masked mode empty
        hipLaunchKernelGGL((k_init_solenoidal<Dm, S>), dim3((fs + 255) / 256), dim3(256), 0, stream, St(), fs,
                           cfg.ph, seed, box, k0);
        DNSBOX_LAUNCH_CHECK();
        spec_reduce();
        hipLaunchKernelGGL((k_scale_to_energy<S>), dim3(1024), dim3(256), 0, stream, St(), 3 * fs,
                           (const double*)(red_out + DG_EKIN), energy);
        DNSBOX_LAUNCH_CHECK();
    }

    void project(int dealias, int proj) override
    {
        x6_valid = false;
        if (dealias) dealiased = prunable;   // state_dealias zeroes every mode outside the kept cube
        hipLaunchKernelGGL((k_project_dealias<Dm, S>), dim3((fs + 255) / 256), dim3(256), 0, stream, St(), fs,
                           cfg.ph, dealias, proj);
        DNSBOX_LAUNCH_CHECK();
    }

    void to_physical(void* phys3) override
    {
        x6_valid = false;
        hipLaunchKernelGGL((k_zinv<Dm, T, ProScale<T>, S>), dim3(G::ZTILES), dim3(ZNT), 0, stream, St(), fs, X(),
                           fs, twz(), ProScale<T>{inv_n3()});
        DNSBOX_LAUNCH_CHECK();
        hipLaunchKernelGGL((k_ycol<Dm, T, +1>), dim3(G::YTILES, 3), dim3(YNT), 0, stream, X(), fs, twy());
        DNSBOX_LAUNCH_CHECK();
        // real output field f at phys3 + f * NX NY NZ reals
        const long long pf = Dm::NP / 2;  // complex pairs per real field
        hipLaunchKernelGGL((k_xpass<Dm, T, OpCopy<T, 3, true, false>>), dim3(G::XBLOCKS), dim3(G::NTX), 0, stream,
                           X(), fs, static_cast<C*>(phys3), pf, twx(), (unsigned long long*)nullptr, 0);
        DNSBOX_LAUNCH_CHECK();
    }

    void from_physical(const void* phys3) override
    {
        x6_valid = false;
        dealiased = false;   // a physical field may hold energy on every mode
        const long long pf = Dm::NP / 2;
        hipLaunchKernelGGL((k_xpass<Dm, T, OpCopy<T, 3, false, true>>), dim3(G::XBLOCKS), dim3(G::NTX), 0, stream,
                           static_cast<const C*>(phys3), pf, X(), fs, twx(), (unsigned long long*)nullptr, 0);
        DNSBOX_LAUNCH_CHECK();
        hipLaunchKernelGGL((k_ycol<Dm, T, -1>), dim3(G::YTILES, 3), dim3(YNT), 0, stream, X(), fs, twy());
        DNSBOX_LAUNCH_CHECK();
        EpiArgs<T, S> ea = epi();
        ea.state = St();
        hipLaunchKernelGGL((k_zfwd<Dm, T, EPI_STORE, S>), dim3(G::ZTILES), dim3(ZNT), 0, stream, X(), fs, twz(),
                           ea);
        DNSBOX_LAUNCH_CHECK();
    }

    void ensure_scratch(int nf)
    {
        if (scratch_fields >= nf) return;
        DNSBOX_HIP(hipStreamSynchronize(stream));
        DNSBOX_HIP(hipFree(scratch));
        scratch = nullptr;
        DNSBOX_HIP(hipMalloc(&scratch, size_t(nf) * size_t(fs) * sizeof(C)));
        scratch_fields = nf;
        x6_valid = false;
    }
    void ensure_adj()
    {
        if (adjbuf) return;
        DNSBOX_HIP(hipMalloc(&adjbuf, 9 * size_t(fs) * sizeof(C)));
        DNSBOX_HIP(hipMalloc(&band2, 2 * sizeof(double)));
    }
    void reserve(int nf, bool adj) override
    {
        ensure_scratch(nf);
        if (adj) ensure_adj();
    }
    C* adjA() const { return static_cast<C*>(adjbuf); }
    C* adjMu() const { return static_cast<C*>(adjbuf) + 3 * fs; }
    C* adjUt() const { return static_cast<C*>(adjbuf) + 6 * fs; }

    // VJP of N at u with cotangent lam_scale * lam; epilogue MODE of k_zfwd_vjp.  With vjp_cube
    // set (u confined to the kept cube, as the caller guarantees) the passes up to the x pass's
    // loads read the kept cube only (the cotangent enters masked); the products, and so the rest,
    // fill the whole spectrum.  The first pass then also checks u outside the cube on otherwise
    // idle blocks (k_zinv_p_chk), raising the flag read back by set_vjp_cube(false).
    template <int MODE>
    void vjp_passes(const C* u, const C* lam, double lam_scale, VjpArgs<T> va)
    {
        if (may_prune() && vjp_cube) vjp_passes_t<MODE, true>(u, lam, lam_scale, va);
        else vjp_passes_t<MODE, false>(u, lam, lam_scale, va);
    }
    template <int MODE, bool PR>
    void vjp_passes_t(const C* u, const C* lam, double lam_scale, VjpArgs<T> va)
    {
        const double sm = 1.0 / (double(NX) * NY * NZ);
        if (cfg.ph.bandlim) {
            hipLaunchKernelGGL((k_band_vjp_scalars<Dm, T>), dim3(1), dim3(256), 0, stream, u, lam, fs, cfg.ph,
                               lam_scale, sm * sm, band2);
            DNSBOX_LAUNCH_CHECK();
        }
        Ptrs<T, 6> in{{lam, lam + fs, lam + 2 * fs, u, u + fs, u + 2 * fs}};
        ProVjpIn<T> pro{T(sm), T(lam_scale), NX, cfg.ph};
        if constexpr (PR) {
            hipLaunchKernelGGL((k_zinv_p_chk<Dm, T, ProVjpIn<T>>), dim3(ztiles<Dm, T, true>() + vjp_check_blocks<Dm>()),
                               dim3(ZNT), 0, stream, in, X(), fs, twz(), pro, u, fs, chk + 1, ztiles<Dm, T, true>());
        } else {
            hipLaunchKernelGGL((k_zinv_p<Dm, T, ProVjpIn<T>>), dim3(ztiles<Dm, T, false>()), dim3(ZNT), 0, stream, in,
                               X(), fs, twz(), pro);
        }
        DNSBOX_LAUNCH_CHECK();
        if (use_plane<Dm, T, OpVjp<T>>()) {
            if constexpr (plane_used<Dm, T, OpVjp<T>>()) {
                hipLaunchKernelGGL((k_plane<NX, T, OpVjp<T>, PR, false>), dim3(NX), dim3(kPlaneThreads), 0, stream,
                                   X(), fs, twc(), (unsigned long long*)nullptr);
                DNSBOX_LAUNCH_CHECK();
            }
        } else {
            hipLaunchKernelGGL((k_ycol<Dm, T, +1, PR>), dim3(ytiles<Dm, T, PR>(), 9), dim3(YNT), 0, stream, X(), fs,
                               twy());
            DNSBOX_LAUNCH_CHECK();
            hipLaunchKernelGGL((k_xpass<Dm, T, OpVjp<T>, PR, false>), dim3(G::XBLOCKS), dim3(G::NTX), 0, stream, X(),
                               fs, X(), fs, twx(), (unsigned long long*)nullptr, 0);
            DNSBOX_LAUNCH_CHECK();
            hipLaunchKernelGGL((k_ycol<Dm, T, -1>), dim3(G::YTILES, 6), dim3(YNT), 0, stream, X(), fs, twy());
            DNSBOX_LAUNCH_CHECK();
        }
        va.u = u;
        va.lamN = lam;
        va.lam_scale = lam_scale;
        va.band = band2;
        va.fs = fs;
        va.ph = cfg.ph;
        hipLaunchKernelGGL((k_zfwd_vjp<Dm, T, MODE>), dim3(G::ZTILES), dim3(ZNT), 0, stream, X(), fs, twz(), va);
        DNSBOX_LAUNCH_CHECK();
    }

    void rhs_vjp(const void* u, const void* lam, void* out) override
    {
        require_uniform("the RHS adjoint");
        x6_valid = false;
        ensure_scratch(9);
        ensure_adj();
        VjpArgs<T> va{};
        va.out = static_cast<C*>(out);
        vjp_passes<VJP_RHS>(static_cast<const C*>(u), static_cast<const C*>(lam), 1.0, va);
    }

    // lam_out = (d u_{n+1} / d u_n)^T lam for one fixed-dt IF-Heun step from u
    void step_vjp(const void* u_, const void* lam_, void* lam_out, double dt) override
    {
        require_uniform("the step adjoint");
        x6_valid = false;
        ensure_scratch(9);
        ensure_adj();
        const C* u = static_cast<const C*>(u_);
        // u* = E (u + dt N(u))
        eband(u);
        zinv_curl(u);
        rhs_middle(false);
        EpiArgs<T, S> ea = epi();
        ea.out = Pr();
        ea.state = reinterpret_cast<CS*>(const_cast<C*>(u));
        ea.rhs_in = u;
        ea.rhs_in_s = 0;
        ea.dt_fixed = dt;
        pred_clean = false;
        hipLaunchKernelGGL((k_zfwd<Dm, T, EPI_PREDONLY, S>), dim3(G::ZTILES), dim3(ZNT), 0, stream, X(), fs, twz(),
                           ea);
        DNSBOX_LAUNCH_CHECK();
        step_vjp_at(u, Pr(), lam_, lam_out, dt);
    }

    void step_vjp_pred(const void* u_, const void* ustar, const void* lam_, void* lam_out, double dt) override
    {
        require_uniform("the step adjoint");
        x6_valid = false;
        ensure_scratch(9);
        ensure_adj();
        step_vjp_at(static_cast<const C*>(u_), static_cast<const C*>(ustar), lam_, lam_out, dt);
    }

    // the two VJP sequences of one step at u with its predicted state us
    void step_vjp_at(const C* u, const C* us, const void* lam_, void* lam_out, double dt)
    {
        const C* lam = static_cast<const C*>(lam_);
        // lam_* = J(u*)^T (dt/2 lam);  A = E (lam + lam_*),  Mu = E (dt/2 lam + dt lam_*)
        VjpArgs<T> v1{};
        v1.lam = lam;
        v1.A = adjA();
        v1.Mu = adjMu();
        v1.dt = dt;
        vjp_passes<VJP_S1>(us, lam, 0.5 * dt, v1);
        // lam_n = A + J(u)^T Mu
        VjpArgs<T> v2{};
        v2.out = static_cast<C*>(lam_out);
        v2.A = adjA();
        v2.dt = dt;
        vjp_passes<VJP_S2>(u, adjMu(), 1.0, v2);
    }

    // the passes of one VJP of N at u whose x-pass inputs of u (fields 3..8 of `slot`) were recorded
    // by the step that evaluated N(u): only the cotangent is transformed in (fields 0..2 of the slot)
    template <int MODE, bool PR>
    void vjp_rec_passes(const C* u, const C* lam, double lam_scale, VjpArgs<T> va, C* slot)
    {
        vjp_band(u, lam, lam_scale);
        vjp_rec_zinv<PR>(u, lam, lam_scale, slot);
        vjp_rec_middle<PR>(slot);
        vjp_args(va, u, lam, lam_scale);
        hipLaunchKernelGGL((k_zfwd_vjp<Dm, T, MODE>), dim3(G::ZTILES), dim3(ZNT), 0, stream, X(), fs, twz(), va);
        DNSBOX_LAUNCH_CHECK();
    }
    // the band-forcing scalars of a VJP of N at u with cotangent lam_scale * lam (into band2)
    void vjp_band(const C* u, const C* lam, double lam_scale)
    {
        if (!cfg.ph.bandlim) return;
        const double sm = 1.0 / (double(NX) * NY * NZ);
        hipLaunchKernelGGL((k_band_vjp_scalars<Dm, T>), dim3(1), dim3(256), 0, stream, u, lam, fs, cfg.ph, lam_scale,
                           sm * sm, band2);
        DNSBOX_LAUNCH_CHECK();
    }
    void vjp_args(VjpArgs<T>& va, const C* u, const C* lam, double lam_scale) const
    {
        va.u = u;
        va.lamN = lam;
        va.lam_scale = lam_scale;
        va.band = band2;
        va.fs = fs;
        va.ph = cfg.ph;
    }
    // first pass: the cotangent's z-inverse into fields 0..2 of the slot (PR: with the check of u
    // outside the kept cube on otherwise idle blocks)
    template <bool PR>
    void vjp_rec_zinv(const C* u, const C* lam, double lam_scale, C* slot)
    {
        Ptrs<T, 3> in{{lam, lam + fs, lam + 2 * fs}};
        ProVjpLam<T> pro{T(lam_scale), NX, cfg.ph};
        if constexpr (PR) {
            hipLaunchKernelGGL((k_zinv_p_chk<Dm, T, ProVjpLam<T>>), dim3(ztiles<Dm, T, true>() + vjp_check_blocks<Dm>()),
                               dim3(ZNT), 0, stream, in, slot, fs, twz(), pro, u, fs, chk + 1, ztiles<Dm, T, true>());
        } else {
            hipLaunchKernelGGL((k_zinv_p<Dm, T, ProVjpLam<T>>), dim3(ztiles<Dm, T, false>()), dim3(ZNT), 0, stream, in,
                               slot, fs, twz(), pro);
        }
        DNSBOX_LAUNCH_CHECK();
    }
    // y-inverse of the cotangent, x pass on it and the recorded inputs, y-forward of the products
    template <bool PR>
    void vjp_rec_middle(C* slot)
    {
        hipLaunchKernelGGL((k_ycol<Dm, T, +1, PR>), dim3(ytiles<Dm, T, PR>(), 3), dim3(YNT), 0, stream, slot, fs, twy());
        DNSBOX_LAUNCH_CHECK();
        hipLaunchKernelGGL((k_xpass<Dm, T, OpVjp<T>, PR, false>), dim3(G::XBLOCKS), dim3(G::NTX), 0, stream, slot, fs,
                           X(), fs, twx(), (unsigned long long*)nullptr, 0);
        DNSBOX_LAUNCH_CHECK();
        hipLaunchKernelGGL((k_ycol<Dm, T, -1>), dim3(G::YTILES, 6), dim3(YNT), 0, stream, X(), fs, twy());
        DNSBOX_LAUNCH_CHECK();
    }

    void step_vjp_mixed(const void* u_, const void* us_, const void* lam_, void* lam_out, double dt, void* mu_,
                        void* ms_, bool pruned_rec) override
    {
        require_uniform("the step adjoint");
        const bool pr = may_prune() && vjp_cube;
        if (pruned_rec && !pr)
            throw std::invalid_argument("a step recorded with pruned passes holds its x-pass inputs at the kept kx "
                                        "only: its adjoint needs dnsbox_set_vjp_cube(ctx, 1) and pruning enabled");
        x6_valid = false;
        ensure_adj();
        const C* u = static_cast<const C*>(u_);
        const C* us = static_cast<const C*>(us_);
        const C* lam = static_cast<const C*>(lam_);
        // lam_* = J(u*)^T (dt/2 lam);  A = E (lam + lam_*),  Mu = E (dt/2 lam + dt lam_*);  lam_n = A + J(u)^T Mu
        VjpArgs<T> v1{};
        v1.lam = lam;
        v1.A = adjA();
        v1.Mu = adjMu();
        v1.dt = dt;
        VjpArgs<T> v2{};
        v2.out = static_cast<C*>(lam_out);
        v2.A = adjA();
        v2.dt = dt;
        if (pr) {
            step_vjp_rec_pruned(u, us, lam, static_cast<C*>(lam_out), dt, static_cast<C*>(mu_), static_cast<C*>(ms_),
                                false, nullptr, nullptr);
        } else {
            vjp_rec_passes<VJP_S1, false>(us, lam, 0.5 * dt, v1, static_cast<C*>(ms_));
            vjp_rec_passes<VJP_S2, false>(u, adjMu(), 1.0, v2, static_cast<C*>(mu_));
        }
    }

    // One recorded step adjoint with pruned passes: lam -> lam_out at u (slot mu) with predicted
    // state us (slot ms).  VJP_S1's z-forward also runs VJP_S2's first pass (the cotangent Mu's
    // z-inverse into mu, and the check of u outside the kept cube).  first_done: VJP_S1's first
    // pass (lam's z-inverse into ms, the check of us) was run by the step after's last z-forward.
    // ms_next (with us_next): that z-forward here runs the first pass of the step before's VJP_S1
    // on lam_out.  lam_out may be lam: VJP_S2 reads no lam.
    void step_vjp_rec_pruned(const C* u, const C* us, const C* lam, C* lam_out, double dt, C* mu, C* ms,
                             bool first_done, const C* us_next, C* ms_next)
    {
        VjpArgs<T> v1{};
        v1.lam = lam;
        v1.A = adjA();
        v1.Mu = adjMu();
        v1.dt = dt;
        VjpArgs<T> v2{};
        v2.out = lam_out;
        v2.A = adjA();
        v2.dt = dt;
        vjp_band(us, lam, 0.5 * dt);
        if (!first_done) vjp_rec_zinv<true>(us, lam, 0.5 * dt, ms);
        vjp_rec_middle<true>(ms);
        vjp_args(v1, us, lam, 0.5 * dt);
        const int tiles = G::ZTILES, grid = G::ZTILES + vjp_check_blocks<Dm>();
        hipLaunchKernelGGL((k_zfwd_vjp_next<Dm, T, VJP_S1>), dim3(grid), dim3(ZNT), 0, stream, X(), fs, twz(), v1, mu, fs,
                           ProVjpLam<T>{T(1.0), NX, cfg.ph}, u, fs, chk + 1, tiles);
        DNSBOX_LAUNCH_CHECK();
        vjp_band(u, adjMu(), 1.0);
        vjp_rec_middle<true>(mu);
        vjp_args(v2, u, adjMu(), 1.0);
        if (ms_next) {
            hipLaunchKernelGGL((k_zfwd_vjp_next<Dm, T, VJP_S2>), dim3(grid), dim3(ZNT), 0, stream, X(), fs, twz(), v2,
                               ms_next, fs, ProVjpLam<T>{T(0.5 * dt), NX, cfg.ph}, us_next, fs, chk + 1, tiles);
        } else {
            hipLaunchKernelGGL((k_zfwd_vjp<Dm, T, VJP_S2>), dim3(G::ZTILES), dim3(ZNT), 0, stream, X(), fs, twz(), v2);
        }
        DNSBOX_LAUNCH_CHECK();
    }

    // step_vjp_mixed over the recorded steps n-1 .. 0 of a segment (u[j], us[j], slots mixed[2j],
    // mixed[2j+1]); the result lands in out, lam is work space.  With pruned passes each step's
    // last z-forward runs the first pass of the step before.
    void step_vjp_mixed_seq(const void* const* u, const void* const* us, void* lam_, void* out_, double dt,
                            void* const* mixed, int n, bool pruned_rec) override
    {
        require_uniform("the step adjoint");
        const bool pr = may_prune() && vjp_cube;
        if (pruned_rec && !pr)
            throw std::invalid_argument("a step recorded with pruned passes holds its x-pass inputs at the kept kx "
                                        "only: its adjoint needs dnsbox_set_vjp_cube(ctx, 1) and pruning enabled");
        x6_valid = false;
        ensure_adj();
        C* lam = static_cast<C*>(lam_);
        C* out = static_cast<C*>(out_);
        const C* src = lam;
        for (int s = 0; s < n; ++s) {
            const int j = n - 1 - s;
            C* dst = j % 2 == 0 ? out : lam;   // step 0 writes out (the first step works in place when n is even)
            if (pr) {
                step_vjp_rec_pruned(static_cast<const C*>(u[j]), static_cast<const C*>(us[j]), src, dst, dt,
                                    static_cast<C*>(mixed[2 * j]), static_cast<C*>(mixed[2 * j + 1]), s > 0,
                                    j > 0 ? static_cast<const C*>(us[j - 1]) : nullptr,
                                    j > 0 ? static_cast<C*>(mixed[2 * j - 1]) : nullptr);
            } else {
                step_vjp_mixed(u[j], us[j], src, dst, dt, mixed[2 * j], mixed[2 * j + 1], pruned_rec);
            }
            src = dst;
        }
    }

    // rhsAdj.f90:64-381 (phatadj read as phat)
    void rhs_adj(const void* u_, void* out, double lin_sign) override
    {
        require_uniform("rhsAdj");
        x6_valid = false;
        ensure_scratch(12);
        ensure_adj();
        const C* u = static_cast<const C*>(u_);
        const T s = inv_n3();
        AdjArgs<T> aa{};
        aa.out = static_cast<C*>(out);
        aa.u = u;
        aa.ut = adjUt();
        aa.fs = fs;
        aa.lin_sign = lin_sign;
        aa.ph = cfg.ph;
        Ptrs<T, 3> in1{{u, u + fs, u + 2 * fs}};
        hipLaunchKernelGGL((k_zinv_p<Dm, T, ProGrad12<T>>), dim3(G::ZTILES), dim3(ZNT), 0, stream, in1, X(), fs,
                           twz(), ProGrad12<T>{s, KScale<T>::of(cfg.ph)});
        DNSBOX_LAUNCH_CHECK();
        hipLaunchKernelGGL((k_ycol<Dm, T, +1>), dim3(G::YTILES, 12), dim3(YNT), 0, stream, X(), fs, twy());
        DNSBOX_LAUNCH_CHECK();
        hipLaunchKernelGGL((k_xpass<Dm, T, OpConv<T>>), dim3(G::XBLOCKS), dim3(G::NTX), 0, stream, X(), fs, X(), fs,
                           twx(), (unsigned long long*)nullptr, 0);
        DNSBOX_LAUNCH_CHECK();
        hipLaunchKernelGGL((k_ycol<Dm, T, -1>), dim3(G::YTILES, 3), dim3(YNT), 0, stream, X(), fs, twy());
        DNSBOX_LAUNCH_CHECK();
        hipLaunchKernelGGL((k_zfwd_adj<Dm, T, 1>), dim3(G::ZTILES), dim3(ZNT), 0, stream, X(), fs, twz(), aa);
        DNSBOX_LAUNCH_CHECK();
        const C* ut = adjUt();
        Ptrs<T, 6> in2{{ut, ut + fs, ut + 2 * fs, u, u + fs, u + 2 * fs}};
        hipLaunchKernelGGL((k_zinv_p<Dm, T, ProAdj12<T>>), dim3(G::ZTILES), dim3(ZNT), 0, stream, in2, X(), fs,
                           twz(), ProAdj12<T>{s, KScale<T>::of(cfg.ph)});
        DNSBOX_LAUNCH_CHECK();
        hipLaunchKernelGGL((k_ycol<Dm, T, +1>), dim3(G::YTILES, 12), dim3(YNT), 0, stream, X(), fs, twy());
        DNSBOX_LAUNCH_CHECK();
        hipLaunchKernelGGL((k_xpass<Dm, T, OpAdjN<T>>), dim3(G::XBLOCKS), dim3(G::NTX), 0, stream, X(), fs, X(), fs,
                           twx(), (unsigned long long*)nullptr, 0);
        DNSBOX_LAUNCH_CHECK();
        hipLaunchKernelGGL((k_ycol<Dm, T, -1>), dim3(G::YTILES, 3), dim3(YNT), 0, stream, X(), fs, twy());
        DNSBOX_LAUNCH_CHECK();
        hipLaunchKernelGGL((k_zfwd_adj<Dm, T, 2>), dim3(G::ZTILES), dim3(ZNT), 0, stream, X(), fs, twz(), aa);
        DNSBOX_LAUNCH_CHECK();
    }

    // dnsboxAdj.f90:76-126: G = rhsAdj(u); u* = u + dtau G; u <- u + (dtau G) / 2 + (dtau G(u*)) / 2;
    // state_project (:117).  On the box's state the clock advances by dtau per step (time(n+1) =
    // time(n) + dt) and the step pipeline restarts unpruned (rhsAdj's ".gt." rule keeps the
    // |k| = N/3 modes that the forward mask removes).
    void adj_descent(void* u_, double dtau, int nsteps, double lin_sign) override
    {
        require_uniform("the adjoint-descent step");
        if (nsteps <= 0) return;
        const bool own = u_ == nullptr;
        C* u = own ? reinterpret_cast<C*>(St()) : static_cast<C*>(u_);
        if (!descbuf) DNSBOX_HIP(hipMalloc(&descbuf, 9 * size_t(fs) * sizeof(C)));
        C* g0 = static_cast<C*>(descbuf);
        C* us = g0 + 3 * fs;
        C* g1 = us + 3 * fs;
        const long long n3 = 3 * fs;
        const dim3 eg((n3 + 255) / 256 < 2048 ? (n3 + 255) / 256 : 2048);
        for (int i = 0; i < nsteps; ++i) {
            rhs_adj(u, g0, lin_sign);
            hipLaunchKernelGGL((k_axpy<T, false>), eg, dim3(256), 0, stream, us, u, g0, dtau, n3);
            DNSBOX_LAUNCH_CHECK();
            rhs_adj(us, g1, lin_sign);
            hipLaunchKernelGGL((k_axpy<T, true>), eg, dim3(256), 0, stream, u, u, g0, dtau, n3);
            hipLaunchKernelGGL((k_axpy<T, true>), eg, dim3(256), 0, stream, u, u, g1, dtau, n3);
            hipLaunchKernelGGL((k_project_dealias<Dm, T>), dim3((fs + 255) / 256), dim3(256), 0, stream, u, fs,
                               cfg.ph, 0, 1);
            DNSBOX_LAUNCH_CHECK();
        }
        if (own) {
            hipLaunchKernelGGL(k_add_time<0>, dim3(1), dim3(64), 0, stream, sc, dtau * nsteps);
            DNSBOX_LAUNCH_CHECK();
            dealiased = false;
            ab2_started = false;
        }
        x6_valid = false;
    }

    // IF-Heun steps on a caller-owned state buffer
    // whether a caller's state buffer holds no energy outside the kept cube (synchronises)
    bool buffer_dealiased(const CS* st)
    {
        DNSBOX_HIP(hipMemsetAsync(chk, 0, sizeof(unsigned int), stream));
        hipLaunchKernelGGL((k_outside_cube<Dm, S>), dim3((fs + 255) / 256 < 1024 ? (fs + 255) / 256 : 1024), dim3(256),
                           0, stream, st, fs, chk);
        DNSBOX_LAUNCH_CHECK();
        unsigned int h = 1;
        DNSBOX_HIP(hipMemcpyAsync(&h, chk, sizeof(h), hipMemcpyDeviceToHost, stream));
        DNSBOX_HIP(hipStreamSynchronize(stream));
        return h == 0;
    }

    // IF-Heun steps of a caller's buffer, pruned when the buffer allows it (checked on the
    // device); the state after step j is stored at traj + j * 3 fields when traj is set (by the
    // corrector pass, beside the state), the step's predicted state at traj_pred + j * 3 fields
    // when that is set (the pruned step writes only the kept modes, so both are cleared first),
    // and the x-pass inputs of its two RHS in the record slots mixed + (2 j, 2 j + 1) * 9 fields
    void step_buffer(void* st, int nsteps, void* traj, void* traj_pred, void* mixed = nullptr)
    {
        require_uniform("stepping a device buffer");
        if (nsteps <= 0) return;
        void* own = state;
        const bool own_dealiased = dealiased;
        x6_valid = false;
        const size_t sb3 = 3 * size_t(fs) * sizeof(CS);
        try {
            dealiased = may_prune() && buffer_dealiased(static_cast<const CS*>(st));
            last_ext_pruned = dealiased;
            state = st;
            // the pruned steps store u* and the new state at the kept modes only
            for (void* buf : {traj_pred, traj}) {
                if (!buf || !dealiased) continue;
                hipLaunchKernelGGL((k_fill_zero<S>), dim3(1024), dim3(256), 0, stream, static_cast<CS*>(buf),
                                   (long long)nsteps * 3 * fs);
                DNSBOX_LAUNCH_CHECK();
            }
            const size_t slot2 = 18 * size_t(fs) * sizeof(C);   // the two 9-field record slots of a step
            for (int i = 0; i < nsteps; ++i) {
                pred_out_ = traj_pred ? static_cast<char*>(traj_pred) + i * sb3 : nullptr;
                traj_out_ = traj ? static_cast<char*>(traj) + i * sb3 : nullptr;
                mixed_out_ = mixed ? static_cast<char*>(mixed) + i * slot2 : nullptr;
                heun_step();
            }
        } catch (...) {
            state = own;
            pred_out_ = traj_out_ = mixed_out_ = nullptr;
            dealiased = own_dealiased;
            x6_valid = false;
            throw;
        }
        state = own;
        pred_out_ = traj_out_ = mixed_out_ = nullptr;
        dealiased = own_dealiased;
        x6_valid = false;
    }
    void step_ext(void* st, int nsteps) override { step_buffer(st, nsteps, nullptr, nullptr); }
    void step_ext_record(void* st, int nsteps, void* traj, void* traj_pred) override
    {
        step_buffer(st, nsteps, traj, traj_pred);
    }
    void step_ext_record_mixed(void* st, int nsteps, void* traj, void* traj_pred, void* mixed) override
    {
        step_buffer(st, nsteps, traj, traj_pred, mixed);
    }

    void pass_bytes(double* b) const override
    {
        const double cb = double(sizeof(C)), sb = double(sizeof(CS));
        const double ny = double(NY), nz = double(NZ);
        if (pruned()) {
            // kept kx, ky and kz indices; zc kept (ky, kx) columns of the z passes
            const double kx = Keep<NX>::KX, ky = Keep<NY>::KY, kz = Keep<NZ>::KY, zc = kx * ky;
            const double pr = cfg.ph.bandlim ? 3.0 * zc * kz * cb : 0.0;   // predicted state (band forcing)
            b[0] = 3.0 * zc * kz * sb + 6.0 * zc * nz * cb;                // state over Kz -> 6 fields, all z
            b[1] = 6.0 * nz * kx * (ky + ny) * cb;                         // ky in Ky read, all y written
            b[2] = 9.0 * ny * nz * kx * cb;                                // 6 lines read, 3 written, kx in Kx
            b[3] = 3.0 * nz * kx * (ny + ky) * cb;                         // all y read, ky in Ky written
            b[4] = 9.0 * zc * nz * cb + 6.0 * zc * kz * sb + pr;           // 3 + 6 fields all z, state over Kz
            b[5] = 9.0 * zc * nz * cb + 6.0 * zc * kz * sb;
            if constexpr (plane_used<Dm, T, OpCross<T>>()) {
                if (use_plane<Dm, T, OpCross<T>>()) {   // one launch for the three middle passes
                    b[1] = b[3] = 0.0;
                    b[2] = plane_bytes<NX, T, OpCross<T>, true, true>();
                }
            }
        } else {
            const double f = double(G::FS);
            b[0] = f * (6.0 * cb + 3.0 * sb);
            b[1] = f * 12.0 * cb;
            b[2] = f * 9.0 * cb;
            b[3] = f * 6.0 * cb;
            b[4] = f * (9.0 * cb + 6.0 * sb) + (cfg.ph.bandlim ? f * 3.0 * cb : 0.0);
            b[5] = f * (9.0 * cb + 6.0 * sb);
            if constexpr (plane_used<Dm, T, OpCross<T>>()) {
                if (use_plane<Dm, T, OpCross<T>>()) {
                    b[1] = b[3] = 0.0;
                    b[2] = plane_bytes<NX, T, OpCross<T>, false, false>();
                }
            }
        }
    }

    // Average duration of each pass inside real IF-Heun steps, taken on a copy of the state
    // (the box's state and clock are left as they were): an event after every launch of
    // `reps` steps, each pass timed from the previous event.  The z-inverse of the first
    // RHS runs once per pipeline start; it is timed over `reps` back-to-back launches.
    void time_passes(int reps, double* ms) override
    {
        if (reps < 1) reps = 1;
        CS* tmp = nullptr;
        DNSBOX_HIP(hipMalloc(&tmp, 3 * size_t(fs) * sizeof(CS)));
        BoxScalars saved{};
        DNSBOX_HIP(hipStreamSynchronize(stream));
        DNSBOX_HIP(hipMemcpy(&saved, sc, sizeof(saved), hipMemcpyDeviceToHost));
        DNSBOX_HIP(hipMemcpyAsync(tmp, state, 3 * size_t(fs) * sizeof(CS), hipMemcpyDeviceToDevice, stream));
        void* own = state;
        std::vector<std::pair<int, hipEvent_t>> mk;
        double sum[kNPass] = {0};
        int cnt[kNPass] = {0};
        try {
            state = tmp;
            x6_valid = false;
            heun_step();  // warm-up, starts the step pipeline
            marks = &mk;
            mark(-1);
            for (int r = 0; r < reps; ++r) heun_step();
            marks = nullptr;
            mk.emplace_back(-1, nullptr);
            DNSBOX_HIP(hipEventCreate(&mk.back().second));
            // z-inverse + curl, back to back
            DNSBOX_HIP(hipEventRecord(mk.back().second, stream));
            for (int r = 0; r < reps; ++r) {
                if (pruned()) zinv_curl<S, true>(static_cast<const CS*>(tmp));
                else zinv_curl(static_cast<const CS*>(tmp));
            }
            hipEvent_t ez;
            DNSBOX_HIP(hipEventCreate(&ez));
            mk.emplace_back(-2, ez);
            DNSBOX_HIP(hipEventRecord(ez, stream));
            DNSBOX_HIP(hipEventSynchronize(ez));
            for (size_t k = 1; k < mk.size(); ++k) {
                float t = 0.f;
                DNSBOX_HIP(hipEventElapsedTime(&t, mk[k - 1].second, mk[k].second));
                const int p = mk[k].first;
                if (p >= 0) {
                    sum[p] += t;
                    cnt[p] += 1;
                } else if (p == -2) {
                    sum[0] = t;
                    cnt[0] = reps;
                }
            }
        } catch (...) {
            marks = nullptr;
            state = own;
            x6_valid = false;
            for (auto& e : mk)
                if (e.second) (void)hipEventDestroy(e.second);
            (void)hipFree(tmp);
            throw;
        }
        state = own;
        x6_valid = false;
        for (auto& e : mk)
            if (e.second) (void)hipEventDestroy(e.second);
        DNSBOX_HIP(hipMemcpy(sc, &saved, sizeof(saved), hipMemcpyHostToDevice));
        DNSBOX_HIP(hipFree(tmp));
        for (int p = 0; p < kNPass; ++p) ms[p] = cnt[p] ? sum[p] / cnt[p] : 0.0;
    }

    void from_ref_layout(void* dst, const double* src, bool state_type) override
    {
        const dim3 grid((NZ + 31) / 32, (G::H + 31) / 32, NY);
        if (state_type) hipLaunchKernelGGL((k_ref_layout<S, true>), grid, dim3(256), 0, stream, src, dst, NX, NY, NZ);
        else hipLaunchKernelGGL((k_ref_layout<T, true>), grid, dim3(256), 0, stream, src, dst, NX, NY, NZ);
        DNSBOX_LAUNCH_CHECK();
    }
    void to_ref_layout(double* dst, const void* src, bool state_type) override
    {
        const dim3 grid((NZ + 31) / 32, (G::H + 31) / 32, NY);
        if (state_type) hipLaunchKernelGGL((k_ref_layout<S, false>), grid, dim3(256), 0, stream, src, dst, NX, NY, NZ);
        else hipLaunchKernelGGL((k_ref_layout<T, false>), grid, dim3(256), 0, stream, src, dst, NX, NY, NZ);
        DNSBOX_LAUNCH_CHECK();
    }
    bool state_confined() override { return buffer_dealiased(St()); }

    void convert_in(void* dst, const double* src, long long n, bool state_type) override
    {
        if (state_type)
            hipLaunchKernelGGL((k_convert<S, double>), dim3(1024), dim3(256), 0, stream, static_cast<S*>(dst), src, n);
        else
            hipLaunchKernelGGL((k_convert<T, double>), dim3(1024), dim3(256), 0, stream, static_cast<T*>(dst), src, n);
        DNSBOX_LAUNCH_CHECK();
    }
    void convert_out(double* dst, const void* src, long long n, bool state_type) override
    {
        if (state_type)
            hipLaunchKernelGGL((k_convert<double, S>), dim3(1024), dim3(256), 0, stream, dst,
                               static_cast<const S*>(src), n);
        else
            hipLaunchKernelGGL((k_convert<double, T>), dim3(1024), dim3(256), 0, stream, dst,
                               static_cast<const T*>(src), n);
        DNSBOX_LAUNCH_CHECK();
    }
};

// a box of one of the non-cubic grids of grids.hpp (box_grid_*.hip); null for another grid
template <class T, class S>
BoxBase* make_box_grid(const BoxConfig& c);
template <>
BoxBase* make_box_grid<double, double>(const BoxConfig& c);
template <>
BoxBase* make_box_grid<float, float>(const BoxConfig& c);
template <>
BoxBase* make_box_grid<float, double>(const BoxConfig& c);

template <class T, class S>
BoxBase* make_box_t(const BoxConfig& c)
{
    if (!(c.nx == c.ny && c.ny == c.nz)) return make_box_grid<T, S>(c);
#ifdef DNSBOX_ONLY_N
    return c.nx == DNSBOX_ONLY_N ? new BoxT<Cube<DNSBOX_ONLY_N>, T, S>(c) : nullptr;
#endif
    switch (c.nx) {
        case 16: return new BoxT<Cube<16>, T, S>(c);
        case 32: return new BoxT<Cube<32>, T, S>(c);
        case 48: return new BoxT<Cube<48>, T, S>(c);
        case 64: return new BoxT<Cube<64>, T, S>(c);
        case 128: return new BoxT<Cube<128>, T, S>(c);
        case 256: return new BoxT<Cube<256>, T, S>(c);
        case 512: return new BoxT<Cube<512>, T, S>(c);
        default: return nullptr;
    }
}

}  // namespace dnsbox
