// C-ABI of include/dnsbox.h.  Converts between the reference's host layout
// (Fortran uhat(Nz,Ny,Nx/2+1), z fastest) and the device layout (kx fastest),
// maps exceptions to error codes, and never aborts the host program.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/dnsbox.h"
#include "box_impl.hpp"
#include "config.hpp"
#include "grids.hpp"
#include "group.hpp"

using dnsbox::BoxBase;
using dnsbox::BoxConfig;

namespace dnsbox {
GroupBase* make_group(std::vector<std::unique_ptr<BoxBase>>&& boxes, int precision);
}

struct dnsbox_group;

struct dnsbox_ctx {
    dnsbox_params p;
    BoxBase* box = nullptr;
    bool owned = true;              // false: a view of one box of a group
    dnsbox_group* group = nullptr;  // the group the box belongs to
    int device = 0;
};

struct dnsbox_group {
    std::unique_ptr<dnsbox::GroupBase> g;
    std::vector<dnsbox_ctx> views;
};

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

template <class F>
int guarded(F&& f)
{
    try {
        return f();
    } catch (const dnsbox::HipError& e) {
        return fail(DNSBOX_E_HIP, e.what());
    } catch (const std::bad_alloc&) {
        return fail(DNSBOX_E_NOMEM, "host allocation failed");
    } catch (const std::exception& e) {
        return fail(DNSBOX_E_ARG, e.what());
    }
}

bool supported_n(int n)
{
    return n == 16 || n == 32 || n == 48 || n == 64 || n == 128 || n == 256 || n == 512;
}

// device staging buffer of n doubles, reused per context call
struct DevBuf {
    void* p = nullptr;
    explicit DevBuf(size_t bytes) { DNSBOX_HIP(hipMalloc(&p, bytes)); }
    ~DevBuf()
    {
        if (p) (void)hipFree(p);
    }
};

// 3 host spectral fields in the reference layout uhat(Nz, Ny, Nx/2+1) (complex(8)) -> dev, in the
// device layout and the state or transform type: a contiguous copy of each field to the GPU,
// then the layout transpose there (k_ref_layout)
void upload_ref3(dnsbox_ctx* c, void* dev, const double* const* src, long long dstride_bytes, bool state_type)
{
    BoxBase* b = c->box;
    const size_t per = size_t(b->fs) * 2 * sizeof(double);
    DevBuf tmp(per);
    for (int f = 0; f < 3; ++f) {
        DNSBOX_HIP(hipMemcpyAsync(tmp.p, src[f], per, hipMemcpyHostToDevice, b->stream));
        b->from_ref_layout(static_cast<char*>(dev) + f * dstride_bytes, static_cast<const double*>(tmp.p),
                           state_type);
        DNSBOX_HIP(hipStreamSynchronize(b->stream));   // tmp holds the next field next
    }
}

// the reverse of upload_ref3
void download_ref3(dnsbox_ctx* c, const void* dev, double* const* dst, long long sstride_bytes, bool state_type)
{
    BoxBase* b = c->box;
    const size_t per = size_t(b->fs) * 2 * sizeof(double);
    DevBuf tmp(per);
    for (int f = 0; f < 3; ++f) {
        b->to_ref_layout(static_cast<double*>(tmp.p), static_cast<const char*>(dev) + f * sstride_bytes, state_type);
        DNSBOX_HIP(hipMemcpyAsync(dst[f], tmp.p, per, hipMemcpyDeviceToHost, b->stream));
        DNSBOX_HIP(hipStreamSynchronize(b->stream));
    }
}

// copy 3 host FP64 arrays (already in device layout, nper reals each) into dev (context precision)
void upload3(dnsbox_ctx* c, void* dev, const std::vector<double>* h, long long nper, long long dstride_bytes,
             bool state_type)
{
    BoxBase* b = c->box;
    DevBuf tmp(size_t(nper) * sizeof(double));
    for (int f = 0; f < 3; ++f) {
        DNSBOX_HIP(hipMemcpyAsync(tmp.p, h[f].data(), size_t(nper) * sizeof(double), hipMemcpyHostToDevice,
                                  b->stream));
        b->convert_in(static_cast<char*>(dev) + f * dstride_bytes, static_cast<const double*>(tmp.p), nper,
                      state_type);
    }
    DNSBOX_HIP(hipStreamSynchronize(b->stream));
}

void download3(dnsbox_ctx* c, const void* dev, std::vector<double>* h, long long nper,
               long long sstride_bytes, bool state_type)
{
    BoxBase* b = c->box;
    DevBuf tmp(size_t(nper) * sizeof(double));
    for (int f = 0; f < 3; ++f) {
        h[f].resize(nper);
        b->convert_out(static_cast<double*>(tmp.p), static_cast<const char*>(dev) + f * sstride_bytes, nper,
                       state_type);
        DNSBOX_HIP(hipMemcpyAsync(h[f].data(), tmp.p, size_t(nper) * sizeof(double), hipMemcpyDeviceToHost,
                                  b->stream));
        DNSBOX_HIP(hipStreamSynchronize(b->stream));
    }
}

int check_ctx(dnsbox_ctx* c)
{
    if (!c || !c->box) return fail(DNSBOX_E_ARG, "null context");
    return 0;
}

// the box's step pipeline (and its group's) no longer matches its state
void invalidate(dnsbox_ctx* c)
{
    c->box->x6_valid = false;
    if (c->group) c->group->g->invalidate();
}

// the grid of a parameter set: nx, ny, nz, each 0 for n
void grid_of(const dnsbox_params* p, int* g)
{
    g[0] = p->nx > 0 ? p->nx : p->n;
    g[1] = p->ny > 0 ? p->ny : p->n;
    g[2] = p->nz > 0 ? p->nz : p->n;
}

int check_params(const dnsbox_params* p)
{
    if (p->nx < 0 || p->ny < 0 || p->nz < 0) return fail(DNSBOX_E_ARG, "nx, ny, nz must be >= 0 (0 = n)");
    int g[3];
    grid_of(p, g);
    const std::string gs = std::to_string(g[0]) + " x " + std::to_string(g[1]) + " x " + std::to_string(g[2]);
    if (g[0] == g[1] && g[1] == g[2]) {
        if (!supported_n(g[0]))
            return fail(DNSBOX_E_ARG, "unsupported grid " + gs + " (cubic grids: 16, 32, 48, 64, 128, 256, 512)");
    } else if (!dnsbox::grid_listed(g[0], g[1], g[2])) {
        return fail(DNSBOX_E_ARG, "grid nx x ny x nz = " + gs +
                                      " is not built into this library (cubic grids 16 .. 512, and " +
                                      dnsbox::grid_list_text() + ")");
    }
    if (p->precision != 32 && p->precision != 64)
        return fail(DNSBOX_E_ARG, "precision must be 32 or 64");
    if (p->state_precision != 0 && p->state_precision != 32 && p->state_precision != 64)
        return fail(DNSBOX_E_ARG, "state_precision must be 0 (= precision), 32 or 64");
    if (p->state_precision == 32 && p->precision == 64)
        return fail(DNSBOX_E_ARG, "an FP64 box keeps an FP64 state");
    if (p->kolmo_k < 0 || p->kolmo_k >= g[1] / 2) return fail(DNSBOX_E_ARG, "kolmo_k out of range (0 .. ny/2 - 1)");
    if (p->bandlim && !(p->kcut > 0.0)) return fail(DNSBOX_E_ARG, "bandlim needs kcut > 0");
    if (!(p->dt > 0.0)) return fail(DNSBOX_E_ARG, "dt must be > 0");
    if (p->adaptive_dt && !(p->cfl_max >= p->cfl_min && p->cfl_min > 0.0 && p->dt_max > 0.0))
        return fail(DNSBOX_E_ARG, "adaptive dt needs 0 < cfl_min <= cfl_max and dt_max > 0");
    const double al[3] = {p->alpha_x, p->alpha_y, p->alpha_z};
    for (double a : al)
        if (!(a >= 0.0) || !(a < 1e30)) return fail(DNSBOX_E_ARG, "alpha_x/y/z must be >= 0 (0 = 1) and finite");
    return 0;
}

}  // namespace

namespace dnsbox {
void set_last_error(const std::string& msg) { g_err = msg; }
}

extern "C" {

const char* dnsbox_last_error(void) { return g_err.c_str(); }
const char* dnsbox_version(void) { return "dnsbox-mi355x 0.1 (gfx950)"; }

int dnsbox_params_default(dnsbox_params* p)
{
    if (!p) return fail(DNSBOX_E_ARG, "null params");
    std::memset(p, 0, sizeof(*p));
    p->nu = 0.005;       // parameters.f90:20
    p->Q = 0.0;          // :21
    p->Pin = 0.1;        // :23
    p->kcut = 2.0;       // :22
    p->deltak = 1.0;     // :26
    p->cfl_min = 0.15;   // :33
    p->cfl_max = 0.2;    // :34
    p->dt_max = 1.0e-2;  // :35
    p->dt = 1.0e-3;      // variables.f90:9
    p->n = 32;
    p->precision = 64;
    p->bandlim = 1;      // :52
    p->spherical = 0;    // :50
    p->parity_mode = 1;
    p->adaptive_dt = 1;  // tStepFix = .false. (:37)
    p->alpha_x = p->alpha_y = p->alpha_z = 1.0;   // :18-20
    return 0;
}

int dnsbox_create(const dnsbox_params* p, dnsbox_ctx** out)
{
    if (!p || !out) return fail(DNSBOX_E_ARG, "null argument");
    *out = nullptr;
    if (int e = check_params(p)) return e;
    return guarded([&]() -> int {
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
            return fail(DNSBOX_E_HIP, "no HIP device available");
        BoxConfig bc = dnsbox::make_config(p);
        auto c = std::make_unique<dnsbox_ctx>();
        c->p = *p;
        (void)hipGetDevice(&c->device);
        c->box = p->precision == 64 ? dnsbox::make_box_f64(bc) : dnsbox::make_box_f32(bc);
        if (!c->box) return fail(DNSBOX_E_ARG, "unsupported configuration");
        *out = c.release();
        return 0;
    });
}

int dnsbox_destroy(dnsbox_ctx* c)
{
    if (!c) return 0;
    if (!c->owned) return fail(DNSBOX_E_STATE, "a box of a group is destroyed with its group");
    return guarded([&]() -> int {
        if (c->box && c->box->stream) (void)hipStreamSynchronize(c->box->stream);
        else (void)hipDeviceSynchronize();
        delete c->box;
        delete c;
        return 0;
    });
}

int dnsbox_set_stream(dnsbox_ctx* c, void* s)
{
    if (int e = check_ctx(c)) return e;
    c->box->stream = static_cast<hipStream_t>(s);
    return 0;
}

int dnsbox_sync(dnsbox_ctx* c)
{
    if (int e = check_ctx(c)) return e;
    return guarded([&]() -> int {
        DNSBOX_HIP(hipStreamSynchronize(c->box->stream));
        return 0;
    });
}

int dnsbox_nspec(dnsbox_ctx* c)
{
    if (int e = check_ctx(c)) return e;
    return c->box->nspec;
}

int dnsbox_set_state(dnsbox_ctx* c, const double* u, const double* v, const double* w)
{
    if (int e = check_ctx(c)) return e;
    invalidate(c);
    if (!u || !v || !w) return fail(DNSBOX_E_ARG, "null state array");
    return guarded([&]() -> int {
        BoxBase* b = c->box;
        const double* src[3] = {u, v, w};
        upload_ref3(c, b->state, src, b->fs * (long long)b->sbytes, true);
        b->dealiased = b->state_confined();   // the pruned passes rely on it
        b->ab2_started = false;
        return 0;
    });
}

int dnsbox_get_state(dnsbox_ctx* c, double* u, double* v, double* w)
{
    if (int e = check_ctx(c)) return e;
    if (!u || !v || !w) return fail(DNSBOX_E_ARG, "null state array");
    return guarded([&]() -> int {
        BoxBase* b = c->box;
        double* dst[3] = {u, v, w};
        download_ref3(c, b->state, dst, b->fs * (long long)b->sbytes, true);
        return 0;
    });
}

int dnsbox_get_nonlin(dnsbox_ctx* c, double* u, double* v, double* w)
{
    if (int e = check_ctx(c)) return e;
    if (!u || !v || !w) return fail(DNSBOX_E_ARG, "null array");
    return guarded([&]() -> int {
        BoxBase* b = c->box;
        double* dst[3] = {u, v, w};
        download_ref3(c, b->nonlin, dst, b->fs * (long long)b->cbytes, false);
        return 0;
    });
}

int dnsbox_set_state_physical(dnsbox_ctx* c, const double* u, const double* v, const double* w)
{
    if (int e = check_ctx(c)) return e;
    invalidate(c);
    if (!u || !v || !w) return fail(DNSBOX_E_ARG, "null array");
    return guarded([&]() -> int {
        BoxBase* b = c->box;
        const long long n3 = (long long)b->cfg.nx * b->cfg.ny * b->cfg.nz;
        std::vector<double> h[3];
        const double* src[3] = {u, v, w};
        for (int f = 0; f < 3; ++f) h[f].assign(src[f], src[f] + n3);
        DevBuf phys(size_t(3 * n3) * b->rbytes);
        upload3(c, phys.p, h, n3, n3 * (long long)b->rbytes, false);
        b->from_physical(phys.p);
        DNSBOX_HIP(hipStreamSynchronize(b->stream));
        b->ab2_started = false;
        return 0;
    });
}

int dnsbox_get_state_physical(dnsbox_ctx* c, double* u, double* v, double* w)
{
    if (int e = check_ctx(c)) return e;
    invalidate(c);
    if (!u || !v || !w) return fail(DNSBOX_E_ARG, "null array");
    return guarded([&]() -> int {
        BoxBase* b = c->box;
        const long long n3 = (long long)b->cfg.nx * b->cfg.ny * b->cfg.nz;
        DevBuf phys(size_t(3 * n3) * b->rbytes);
        b->to_physical(phys.p);
        std::vector<double> h[3];
        download3(c, phys.p, h, n3, n3 * (long long)b->rbytes, false);
        double* dst[3] = {u, v, w};
        for (int f = 0; f < 3; ++f) std::memcpy(dst[f], h[f].data(), size_t(n3) * sizeof(double));
        return 0;
    });
}

int dnsbox_init_random(dnsbox_ctx* c, uint64_t seed, int box, double k0, double energy)
{
    if (int e = check_ctx(c)) return e;
    invalidate(c);
    if (!(k0 > 0.0) || !(energy >= 0.0)) return fail(DNSBOX_E_ARG, "k0 must be > 0, energy >= 0");
    return guarded([&]() -> int {
        c->box->init_random(seed, box, k0, energy);
        c->box->ab2_started = false;
        return 0;
    });
}

int dnsbox_project(dnsbox_ctx* c, int dealias, int project)
{
    if (int e = check_ctx(c)) return e;
    invalidate(c);
    return guarded([&]() -> int {
        c->box->project(dealias, project);
        return 0;
    });
}

int dnsbox_int_fact(dnsbox_ctx* c, double dt)
{
    if (int e = check_ctx(c)) return e;
    if (!(dt > 0.0)) return fail(DNSBOX_E_ARG, "dt must be > 0");
    return guarded([&]() -> int {
        BoxBase* b = c->box;
        DNSBOX_HIP(hipStreamSynchronize(b->stream));
        dnsbox::BoxScalars s{};
        DNSBOX_HIP(hipMemcpy(&s, b->sc, sizeof(s), hipMemcpyDeviceToHost));
        s.dt = s.dt_prev = dt;
        DNSBOX_HIP(hipMemcpy(b->sc, &s, sizeof(s), hipMemcpyHostToDevice));
        c->p.dt = dt;
        return 0;
    });
}

int dnsbox_rhs_nonlinear(dnsbox_ctx* c)
{
    if (int e = check_ctx(c)) return e;
    invalidate(c);
    return guarded([&]() -> int {
        c->box->rhs_state();
        return 0;
    });
}

int dnsbox_step(dnsbox_ctx* c, int nsteps, int stepper)
{
    if (int e = check_ctx(c)) return e;
    if (nsteps < 0) return fail(DNSBOX_E_ARG, "nsteps < 0");
    if (stepper < 1 || stepper > 3)
        return fail(DNSBOX_E_ARG, "stepper must be 1 (IF-Heun), 2 (IF-AB2) or 3 (IF-Heun + projection)");
    if (c->group) c->group->g->invalidate();
    return guarded([&]() -> int {
        BoxBase* b = c->box;
        if (stepper == 1) {
            b->heun_steps(nsteps);
        } else if (stepper == 2) {
            for (int i = 0; i < nsteps; ++i) b->ab2_step();
        } else {
            for (int i = 0; i < nsteps; ++i) {   // dnsboxband.f90:91-156
                b->heun_step();
                b->project(1, 1);
            }
        }
        return 0;
    });
}

int dnsbox_get_time(dnsbox_ctx* c, double* t, double* dt)
{
    if (int e = check_ctx(c)) return e;
    return guarded([&]() -> int {
        BoxBase* b = c->box;
        DNSBOX_HIP(hipStreamSynchronize(b->stream));
        dnsbox::BoxScalars s{};
        DNSBOX_HIP(hipMemcpy(&s, b->sc, sizeof(s), hipMemcpyDeviceToHost));
        if (t) *t = s.t;
        if (dt) *dt = s.dt_prev;
        return 0;
    });
}

int dnsbox_set_time(dnsbox_ctx* c, double t, double dt)
{
    if (int e = check_ctx(c)) return e;
    if (!(dt > 0.0)) return fail(DNSBOX_E_ARG, "dt must be > 0");
    return guarded([&]() -> int {
        BoxBase* b = c->box;
        DNSBOX_HIP(hipStreamSynchronize(b->stream));
        dnsbox::BoxScalars s{};
        DNSBOX_HIP(hipMemcpy(&s, b->sc, sizeof(s), hipMemcpyDeviceToHost));
        s.t = t;
        s.dt = s.dt_prev = dt;
        DNSBOX_HIP(hipMemcpy(b->sc, &s, sizeof(s), hipMemcpyHostToDevice));
        return 0;
    });
}

int dnsbox_diagnostics(dnsbox_ctx* c, double* diag, double* espec)
{
    if (int e = check_ctx(c)) return e;
    invalidate(c);
    if (!diag) return fail(DNSBOX_E_ARG, "null diag");
    return guarded([&]() -> int {
        BoxBase* b = c->box;
        b->spec_reduce();
        b->phys_reduce();
        std::vector<double> red(dnsbox::DG_NSCAL + b->nspec);
        unsigned long long bits[4];
        dnsbox::BoxScalars s{};
        DNSBOX_HIP(hipMemcpyAsync(red.data(), b->red_out, red.size() * sizeof(double), hipMemcpyDeviceToHost,
                                  b->stream));
        DNSBOX_HIP(hipMemcpyAsync(bits, b->red_bits, sizeof(bits), hipMemcpyDeviceToHost, b->stream));
        DNSBOX_HIP(hipMemcpyAsync(&s, b->sc, sizeof(s), hipMemcpyDeviceToHost, b->stream));
        DNSBOX_HIP(hipStreamSynchronize(b->stream));
        double mx[3];
        for (int i = 0; i < 3; ++i) std::memcpy(&mx[i], &bits[i], sizeof(double));
        unsigned long long ob = bits[3];
        ob = (ob >> 63) ? (ob & 0x7fffffffffffffffull) : ~ob;
        double divmax;
        std::memcpy(&divmax, &ob, sizeof(double));
        const double tp = 6.283185307179586476925286766559;
        const double dt = s.dt_prev;
        const double* al = b->cfg.ph.a;   // io.f90:664-676: dx_d = L_d / N_d, L_d = 2 pi / alpha_d
        const bool cubic = b->cfg.nx == b->cfg.ny && b->cfg.ny == b->cfg.nz;
        diag[DNSBOX_D_EKIN] = red[dnsbox::DG_EKIN];
        diag[DNSBOX_D_DISP] = red[dnsbox::DG_DISP];
        diag[DNSBOX_D_PROD] = 2.0 * c->p.Q * red[dnsbox::DG_EKIN];
        diag[DNSBOX_D_COURANT] = cubic ? (mx[0] * al[0] + mx[1] * al[1] + mx[2] * al[2]) * dt / (tp / double(b->cfg.nx))
                                       : mx[0] * dt / ((tp / al[0]) / double(b->cfg.nx)) +
                                             mx[1] * dt / ((tp / al[1]) / double(b->cfg.ny)) +
                                             mx[2] * dt / ((tp / al[2]) / double(b->cfg.nz));
        diag[DNSBOX_D_DIVMAX] = divmax;
        diag[DNSBOX_D_EZERO] = red[dnsbox::DG_EZERO];
        diag[DNSBOX_D_EONE] = red[dnsbox::DG_EONE];
        diag[DNSBOX_D_ESQRT2] = red[dnsbox::DG_ESQRT2];
        diag[DNSBOX_D_EBAND] = red[dnsbox::DG_EBAND];
        diag[DNSBOX_D_TIME] = s.t;
        diag[DNSBOX_D_DT] = dt;
        if (espec)
            for (int i = 0; i < b->nspec; ++i) espec[i] = red[dnsbox::DG_NSCAL + i];
        return 0;
    });
}

int dnsbox_time_passes(dnsbox_ctx* c, int reps, double* ms)
{
    if (int e = check_ctx(c)) return e;
    invalidate(c);
    if (reps < 1 || !ms) return fail(DNSBOX_E_ARG, "reps >= 1 and an output array are required");
    return guarded([&]() -> int {
        c->box->time_passes(reps, ms);
        return 0;
    });
}

int dnsbox_state_ptr(dnsbox_ctx* c, void** p)
{
    if (int e = check_ctx(c)) return e;
    if (!p) return fail(DNSBOX_E_ARG, "null pointer");
    invalidate(c);
    // the caller may write the state: the step no longer assumes it empty outside the kept
    // cube until dnsbox_project(..., dealias = 1), dnsbox_init_random or dnsbox_set_state
    c->box->dealiased = false;
    *p = c->box->state;
    return 0;
}

int dnsbox_pass_bytes(dnsbox_ctx* c, double* bytes)
{
    if (int e = check_ctx(c)) return e;
    if (!bytes) return fail(DNSBOX_E_ARG, "null pointer");
    c->box->pass_bytes(bytes);
    return 0;
}

int dnsbox_set_pruning(dnsbox_ctx* c, int enable)
{
    if (int e = check_ctx(c)) return e;
    invalidate(c);
    c->box->prune = enable != 0;
    return 0;
}

int dnsbox_set_vjp_cube(dnsbox_ctx* c, int confined)
{
    if (int e = check_ctx(c)) return e;
    return guarded([&]() -> int {
        if (c->box->set_vjp_cube(confined != 0))
            return fail(DNSBOX_E_STATE, "a state differentiated under dnsbox_set_vjp_cube(ctx, 1) had energy outside "
                                        "the kept cube: the adjoints since the guarantee was set are wrong");
        return 0;
    });
}

int dnsbox_get_last_step_pruned(dnsbox_ctx* c, int* pruned)
{
    if (int e = check_ctx(c)) return e;
    if (!pruned) return fail(DNSBOX_E_ARG, "null pointer");
    *pruned = c->box->last_ext_pruned ? 1 : 0;
    return 0;
}

int dnsbox_get_pruning(dnsbox_ctx* c, int* active)
{
    if (int e = check_ctx(c)) return e;
    if (!active) return fail(DNSBOX_E_ARG, "null pointer");
    *active = c->box->pruned() ? 1 : 0;
    return 0;
}

int dnsbox_rhs_dev(dnsbox_ctx* c, const void* in, void* out)
{
    if (int e = check_ctx(c)) return e;
    invalidate(c);
    if (!in || !out) return fail(DNSBOX_E_ARG, "null device pointer");
    return guarded([&]() -> int {
        c->box->rhs(in, out);
        return 0;
    });
}

int dnsbox_step_dev(dnsbox_ctx* c, void* state, int nsteps)
{
    if (int e = check_ctx(c)) return e;
    invalidate(c);
    if (!state || nsteps < 0) return fail(DNSBOX_E_ARG, "null state or nsteps < 0");
    return guarded([&]() -> int {
        c->box->step_ext(state, nsteps);
        return 0;
    });
}

int dnsbox_step_dev_record(dnsbox_ctx* c, void* state, int nsteps, void* traj, void* traj_pred)
{
    if (int e = check_ctx(c)) return e;
    invalidate(c);
    if (!state || !traj || nsteps < 0) return fail(DNSBOX_E_ARG, "null buffer or nsteps < 0");
    if (traj_pred == traj || traj_pred == state) return fail(DNSBOX_E_ARG, "traj_pred must not alias traj or state");
    return guarded([&]() -> int {
        c->box->step_ext_record(state, nsteps, traj, traj_pred);
        return 0;
    });
}

int dnsbox_step_dev_record_mixed(dnsbox_ctx* c, void* state, int nsteps, void* traj, void* traj_pred, void* mixed)
{
    if (int e = check_ctx(c)) return e;
    invalidate(c);
    if (!state || !traj || !traj_pred || !mixed || nsteps < 0)
        return fail(DNSBOX_E_ARG, "null buffer or nsteps < 0");
    if (traj_pred == traj || traj_pred == state || mixed == traj || mixed == traj_pred || mixed == state)
        return fail(DNSBOX_E_ARG, "state, traj, traj_pred and mixed must be distinct buffers");
    if (!c->box->uniform) return fail(DNSBOX_E_ARG, "recording needs state_precision = precision");
    return guarded([&]() -> int {
        c->box->step_ext_record_mixed(state, nsteps, traj, traj_pred, mixed);
        return 0;
    });
}

int dnsbox_step_vjp_mixed_dev(dnsbox_ctx* c, const void* u, const void* ustar, const void* lam, void* lam_out,
                              double dt, void* mixed_u, void* mixed_ustar, int pruned_record)
{
    if (int e = check_ctx(c)) return e;
    invalidate(c);
    if (!u || !ustar || !lam || !lam_out || !mixed_u || !mixed_ustar) return fail(DNSBOX_E_ARG, "null device pointer");
    if (!(dt > 0.0)) return fail(DNSBOX_E_ARG, "dt must be > 0");
    if (lam_out == u || lam_out == lam || lam_out == ustar)
        return fail(DNSBOX_E_ARG, "lam_out must not alias u, ustar or lam");
    if (mixed_u == mixed_ustar) return fail(DNSBOX_E_ARG, "the two record slots must differ");
    return guarded([&]() -> int {
        c->box->step_vjp_mixed(u, ustar, lam, lam_out, dt, mixed_u, mixed_ustar, pruned_record != 0);
        return 0;
    });
}

int dnsbox_step_vjp_mixed_seq_dev(dnsbox_ctx* c, const void* const* u, const void* const* ustar, void* lam,
                                  void* lam_out, double dt, void* const* mixed, int nsteps, int pruned_record)
{
    if (int e = check_ctx(c)) return e;
    invalidate(c);
    if (!u || !ustar || !lam || !lam_out || !mixed) return fail(DNSBOX_E_ARG, "null pointer");
    if (nsteps < 1) return fail(DNSBOX_E_ARG, "nsteps must be >= 1");
    if (!(dt > 0.0)) return fail(DNSBOX_E_ARG, "dt must be > 0");
    if (lam_out == lam) return fail(DNSBOX_E_ARG, "lam_out must not alias lam");
    for (int j = 0; j < nsteps; ++j) {
        if (!u[j] || !ustar[j] || !mixed[2 * j] || !mixed[2 * j + 1]) return fail(DNSBOX_E_ARG, "null device pointer");
        if (u[j] == lam || u[j] == lam_out || ustar[j] == lam || ustar[j] == lam_out)
            return fail(DNSBOX_E_ARG, "lam and lam_out must not alias a state");
        if (mixed[2 * j] == mixed[2 * j + 1]) return fail(DNSBOX_E_ARG, "the two record slots of a step must differ");
    }
    return guarded([&]() -> int {
        c->box->step_vjp_mixed_seq(u, ustar, lam, lam_out, dt, mixed, nsteps, pruned_record != 0);
        return 0;
    });
}

int dnsbox_step_vjp_pred_dev(dnsbox_ctx* c, const void* u, const void* ustar, const void* lam, void* lam_out,
                             double dt)
{
    if (int e = check_ctx(c)) return e;
    invalidate(c);
    if (!u || !ustar || !lam || !lam_out) return fail(DNSBOX_E_ARG, "null device pointer");
    if (!(dt > 0.0)) return fail(DNSBOX_E_ARG, "dt must be > 0");
    if (lam_out == u || lam_out == lam || lam_out == ustar)
        return fail(DNSBOX_E_ARG, "lam_out must not alias u, ustar or lam");
    return guarded([&]() -> int {
        c->box->step_vjp_pred(u, ustar, lam, lam_out, dt);
        return 0;
    });
}

int dnsbox_rhs_vjp_dev(dnsbox_ctx* c, const void* u, const void* lam, void* out)
{
    if (int e = check_ctx(c)) return e;
    invalidate(c);
    if (!u || !lam || !out) return fail(DNSBOX_E_ARG, "null device pointer");
    return guarded([&]() -> int {
        c->box->rhs_vjp(u, lam, out);
        return 0;
    });
}

int dnsbox_step_vjp_dev(dnsbox_ctx* c, const void* u, const void* lam, void* lam_out, double dt)
{
    if (int e = check_ctx(c)) return e;
    invalidate(c);
    if (!u || !lam || !lam_out) return fail(DNSBOX_E_ARG, "null device pointer");
    if (!(dt > 0.0)) return fail(DNSBOX_E_ARG, "dt must be > 0");
    return guarded([&]() -> int {
        c->box->step_vjp(u, lam, lam_out, dt);
        return 0;
    });
}

int dnsbox_rhs_adj_dev(dnsbox_ctx* c, const void* u, void* out, int lin_sign)
{
    if (int e = check_ctx(c)) return e;
    invalidate(c);
    if (!u || !out) return fail(DNSBOX_E_ARG, "null device pointer");
    if (lin_sign != 1 && lin_sign != -1) return fail(DNSBOX_E_ARG, "lin_sign must be +1 or -1");
    return guarded([&]() -> int {
        c->box->rhs_adj(u, out, double(lin_sign));
        return 0;
    });
}

int dnsbox_adj_descent_step_dev(dnsbox_ctx* c, void* u, double dtau, int nsteps, int lin_sign)
{
    if (int e = check_ctx(c)) return e;
    invalidate(c);
    if (!u || nsteps < 0) return fail(DNSBOX_E_ARG, "null device pointer or nsteps < 0");
    if (!(dtau > 0.0)) return fail(DNSBOX_E_ARG, "dtau must be > 0");
    if (lin_sign != 1 && lin_sign != -1) return fail(DNSBOX_E_ARG, "lin_sign must be +1 or -1");
    return guarded([&]() -> int {
        c->box->adj_descent(u, dtau, nsteps, double(lin_sign));
        return 0;
    });
}

int dnsbox_adj_descent_step(dnsbox_ctx* c, double dtau, int nsteps, int lin_sign)
{
    if (int e = check_ctx(c)) return e;
    invalidate(c);
    if (nsteps < 0) return fail(DNSBOX_E_ARG, "nsteps < 0");
    if (!(dtau > 0.0)) return fail(DNSBOX_E_ARG, "dtau must be > 0");
    if (lin_sign != 1 && lin_sign != -1) return fail(DNSBOX_E_ARG, "lin_sign must be +1 or -1");
    return guarded([&]() -> int {
        c->box->adj_descent(nullptr, dtau, nsteps, double(lin_sign));
        return 0;
    });
}

int dnsbox_group_create(const dnsbox_params* params, int nbox, dnsbox_group** out)
{
    if (!params || !out || nbox < 1) return fail(DNSBOX_E_ARG, "null argument or nbox < 1");
    *out = nullptr;
    const int prec = params[0].precision;
    for (int i = 0; i < nbox; ++i) {
        const dnsbox_params* p = params + i;
        if (int e = check_params(p)) return e;
        if (!(p->n == 16 || p->n == 32 || p->n == 48 || p->n == 64 || p->n == 128))
            return fail(DNSBOX_E_ARG, "group boxes need n in {16, 32, 48, 64, 128}");
        int g[3];
        grid_of(p, g);
        if (!(g[0] == p->n && g[1] == p->n && g[2] == p->n))
            return fail(DNSBOX_E_ARG, "group boxes are cubic (nx, ny, nz 0 or n)");
        if (p->precision != prec || (p->state_precision != 0 && p->state_precision != prec))
            return fail(DNSBOX_E_ARG, "group boxes share one precision (state included)");
        if (p->bandlim) return fail(DNSBOX_E_ARG, "group boxes do not support band-limited forcing");
    }
    return guarded([&]() -> int {
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
            return fail(DNSBOX_E_HIP, "no HIP device available");
        std::vector<std::unique_ptr<BoxBase>> boxes;
        for (int i = 0; i < nbox; ++i) {
            BoxConfig bc = dnsbox::make_config(params + i);
            boxes.emplace_back(prec == 64 ? dnsbox::make_box_f64(bc) : dnsbox::make_box_f32(bc));
        }
        auto grp = std::make_unique<dnsbox_group>();
        grp->g.reset(dnsbox::make_group(std::move(boxes), prec));
        grp->views.resize(nbox);
        for (int i = 0; i < nbox; ++i) {
            dnsbox_ctx& v = grp->views[i];
            v.p = params[i];
            v.box = grp->g->boxes[i].get();
            v.owned = false;
            v.group = grp.get();
            (void)hipGetDevice(&v.device);
        }
        *out = grp.release();
        return 0;
    });
}

int dnsbox_group_destroy(dnsbox_group* g)
{
    if (!g) return 0;
    return guarded([&]() -> int {
        (void)hipDeviceSynchronize();
        delete g;
        return 0;
    });
}

int dnsbox_group_box(dnsbox_group* g, int i, dnsbox_ctx** box)
{
    if (!g || !box || i < 0 || i >= int(g->views.size())) return fail(DNSBOX_E_ARG, "bad group or index");
    *box = &g->views[i];
    return 0;
}

int dnsbox_group_size(dnsbox_group* g)
{
    if (!g) return fail(DNSBOX_E_ARG, "null group");
    return int(g->views.size());
}

int dnsbox_group_step(dnsbox_group* g, int nsteps)
{
    if (!g) return fail(DNSBOX_E_ARG, "null group");
    if (nsteps < 0) return fail(DNSBOX_E_ARG, "nsteps < 0");
    return guarded([&]() -> int {
        g->g->step(nsteps);
        for (auto& b : g->g->boxes) b->x6_valid = false;
        return 0;
    });
}

int dnsbox_group_sync(dnsbox_group* g)
{
    if (!g) return fail(DNSBOX_E_ARG, "null group");
    return guarded([&]() -> int {
        DNSBOX_HIP(hipDeviceSynchronize());
        return 0;
    });
}

int dnsbox_group_set_stream(dnsbox_group* g, void* stream)
{
    if (!g) return fail(DNSBOX_E_ARG, "null group");
    g->g->stream = static_cast<hipStream_t>(stream);
    return 0;
}

int dnsbox_group_step_dev(dnsbox_group* g, void* const* states, int nsteps)
{
    if (!g || !states) return fail(DNSBOX_E_ARG, "null argument");
    if (nsteps < 0) return fail(DNSBOX_E_ARG, "nsteps < 0");
    for (size_t i = 0; i < g->views.size(); ++i)
        if (!states[i]) return fail(DNSBOX_E_ARG, "null state pointer");
    return guarded([&]() -> int {
        g->g->step_ext(states, nsteps);
        for (auto& b : g->g->boxes) b->x6_valid = false;
        return 0;
    });
}

int dnsbox_group_step_dev_record(dnsbox_group* g, void* const* states, int nsteps, void* const* traj,
                                 void* const* traj_pred)
{
    if (!g || !states || !traj) return fail(DNSBOX_E_ARG, "null argument");
    if (nsteps < 0) return fail(DNSBOX_E_ARG, "nsteps < 0");
    for (size_t i = 0; i < g->views.size(); ++i) {
        if (!states[i] || !traj[i]) return fail(DNSBOX_E_ARG, "null state or trajectory pointer");
        if (traj_pred && (!traj_pred[i] || traj_pred[i] == traj[i] || traj_pred[i] == states[i]))
            return fail(DNSBOX_E_ARG, "traj_pred[i] must be set and must not alias traj[i] or states[i]");
    }
    return guarded([&]() -> int {
        g->g->step_ext_record(states, nsteps, traj, traj_pred);
        for (auto& b : g->g->boxes) b->x6_valid = false;
        return 0;
    });
}

int dnsbox_group_step_vjp_pred_dev(dnsbox_group* g, const void* const* u, const void* const* ustar,
                                   const void* const* lam, void* const* lam_out, const double* dt)
{
    if (!g || !u || !ustar || !lam || !lam_out || !dt) return fail(DNSBOX_E_ARG, "null argument");
    for (size_t i = 0; i < g->views.size(); ++i) {
        if (!u[i] || !ustar[i] || !lam[i] || !lam_out[i]) return fail(DNSBOX_E_ARG, "null device pointer");
        if (!(dt[i] > 0.0)) return fail(DNSBOX_E_ARG, "dt must be > 0");
        if (lam_out[i] == u[i] || lam_out[i] == lam[i] || lam_out[i] == ustar[i])
            return fail(DNSBOX_E_ARG, "lam_out must not alias u, ustar or lam");
    }
    return guarded([&]() -> int {
        g->g->step_vjp_pred(u, ustar, lam, lam_out, dt);
        for (auto& b : g->g->boxes) b->x6_valid = false;
        return 0;
    });
}

int dnsbox_group_step_dev_record_mixed(dnsbox_group* g, void* const* states, int nsteps, void* const* traj,
                                       void* const* traj_pred, void* const* mixed)
{
    if (!g || !states || !traj || !traj_pred || !mixed) return fail(DNSBOX_E_ARG, "null argument");
    if (nsteps < 0) return fail(DNSBOX_E_ARG, "nsteps < 0");
    for (size_t i = 0; i < g->views.size(); ++i) {
        if (!states[i] || !traj[i] || !traj_pred[i] || !mixed[i])
            return fail(DNSBOX_E_ARG, "null state, trajectory or record pointer");
        if (traj_pred[i] == traj[i] || traj_pred[i] == states[i] || mixed[i] == traj[i] || mixed[i] == traj_pred[i] ||
            mixed[i] == states[i])
            return fail(DNSBOX_E_ARG, "states, traj, traj_pred and mixed of a box must be distinct buffers");
    }
    return guarded([&]() -> int {
        g->g->step_ext_record_mixed(states, nsteps, traj, traj_pred, mixed);
        for (auto& b : g->g->boxes) b->x6_valid = false;
        return 0;
    });
}

int dnsbox_group_step_vjp_mixed_dev(dnsbox_group* g, const void* const* u, const void* const* ustar,
                                    const void* const* lam, void* const* lam_out, const double* dt,
                                    void* const* mixed, int pruned_record)
{
    if (!g || !u || !ustar || !lam || !lam_out || !dt || !mixed) return fail(DNSBOX_E_ARG, "null argument");
    for (size_t i = 0; i < g->views.size(); ++i) {
        if (!u[i] || !ustar[i] || !lam[i] || !lam_out[i] || !mixed[i])
            return fail(DNSBOX_E_ARG, "null device pointer");
        if (!(dt[i] > 0.0)) return fail(DNSBOX_E_ARG, "dt must be > 0");
        if (lam_out[i] == u[i] || lam_out[i] == lam[i] || lam_out[i] == ustar[i])
            return fail(DNSBOX_E_ARG, "lam_out must not alias u, ustar or lam");
    }
    return guarded([&]() -> int {
        g->g->step_vjp_mixed(u, ustar, lam, lam_out, dt, mixed, pruned_record != 0);
        for (auto& b : g->g->boxes) b->x6_valid = false;
        return 0;
    });
}

int dnsbox_group_set_vjp_cube(dnsbox_group* g, int confined)
{
    if (!g) return fail(DNSBOX_E_ARG, "null group");
    return guarded([&]() -> int {
        if (g->g->set_vjp_cube(confined != 0))
            return fail(DNSBOX_E_STATE, "a state differentiated under dnsbox_group_set_vjp_cube(g, 1) had energy "
                                        "outside its kept cube: the adjoints since the guarantee was set are wrong");
        return 0;
    });
}

int dnsbox_group_get_last_step_pruned(dnsbox_group* g, int* pruned)
{
    if (!g || !pruned) return fail(DNSBOX_E_ARG, "null argument");
    *pruned = g->g->last_ext_pruned ? 1 : 0;
    return 0;
}

int dnsbox_group_step_vjp_dev(dnsbox_group* g, const void* const* u, const void* const* lam, void* const* lam_out,
                              const double* dt)
{
    if (!g || !u || !lam || !lam_out || !dt) return fail(DNSBOX_E_ARG, "null argument");
    for (size_t i = 0; i < g->views.size(); ++i) {
        if (!u[i] || !lam[i] || !lam_out[i]) return fail(DNSBOX_E_ARG, "null device pointer");
        if (!(dt[i] > 0.0)) return fail(DNSBOX_E_ARG, "dt must be > 0");
        if (lam_out[i] == u[i] || lam_out[i] == lam[i])
            return fail(DNSBOX_E_ARG, "lam_out must not alias u or lam");
    }
    return guarded([&]() -> int {
        g->g->step_vjp(u, lam, lam_out, dt);
        for (auto& b : g->g->boxes) b->x6_valid = false;
        return 0;
    });
}

}  // extern "C"
