/*
 * CPU FP64 oracle for dnsbox's pseudospectral Navier-Stokes step.
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker; never linked into the product.
 *
 * A plain-C restatement of the reference Fortran (reference files cited per
 * function).  It keeps the reference's structure on purpose: 15 3-D transforms
 * per RHS (9 derivative c2r + 3 velocity c2r + 3 r2c), the reference k-tables
 * with the Nyquist wavenumber set to 0, the cubic ".ge." 2/3 mask, and the
 * reference diagnostics weights.  The FFT is the only part the reference takes
 * from a library (p3dfft over FFTW3, not vendored and unversioned); here it is a
 * recursive mixed-radix Cooley-Tukey transform with long-double twiddles,
 * pinned in tests against numpy.fft and a brute-force DFT.
 *
 * Parity status of each function is recorded in the test files that pin it.
 */
#include "dnsbox_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double _Complex cplx;

#define TWO_PI_L 6.283185307179586476925286766559005768L

/* ------------------------------------------------------------------------- */
/* grid helpers                                                              */
/* ------------------------------------------------------------------------- */

/* variables.f90:226-262 — k(i) = i*(i-1) for i <= N/2, k(N/2+1) = 0 (Nyquist),
 * k(i+1+N/2) = -k(1-i+N/2).  Stored here as the imaginary part (alpha = 1). */
void orc_ktable(int n, double* k)
{
    for (int i = 0; i < n; ++i) {
        if (i < n / 2) k[i] = (double)i;
        else if (i == n / 2) k[i] = 0.0;
        else k[i] = (double)(i - n);
    }
}

/* alpha of axis d (parameters.f90:18-20: real(kind=4), so its single-precision value; 0 = 1) */
static double alpha_of(const orc_params* p, int d)
{
    const double a = d == 0 ? p->alpha_x : (d == 1 ? p->alpha_y : p->alpha_z);
    return a > 0.0 ? (double)(float)a : 1.0;
}

/* 2/3-rule threshold of axis d, (N_d/2)(2 alpha_d/3) (rhs.f90:80-85, state.f90:191-196) */
static double thr_of(const orc_params* p, int d)
{
    const int n = d == 0 ? p->nx : (d == 1 ? p->ny : p->nz);
    return ((double)n / 2.0) * (2.0 * alpha_of(p, d) / 3.0);
}

/* variables.f90:112-123 */
int orc_nspec(const orc_params* p)
{
    const double deltak = p->deltak > 0 ? p->deltak : 1.0;
    const double ax = thr_of(p, 0), ay = thr_of(p, 1), az = thr_of(p, 2);
    if (p->spherical) return (int)(ax / deltak) - 1;
    return (int)(sqrt(ax * ax + ay * ay + az * az) / deltak) - 1;
}

static inline size_t sidx(const orc_params* p, int ix, int iy, int iz)
{
    return ((size_t)ix * p->ny + iy) * p->nz + iz;
}
static inline size_t pidx(const orc_params* p, int x, int y, int z)
{
    return ((size_t)z * p->ny + y) * p->nx + x;
}
static inline int nxh(const orc_params* p) { return p->nx / 2 + 1; }
static inline size_t nspec_elems(const orc_params* p) { return (size_t)nxh(p) * p->ny * p->nz; }
static inline size_t nphys_elems(const orc_params* p) { return (size_t)p->nx * p->ny * p->nz; }

/* rhs.f90:74-88 (and state.f90:184-197): cubic 2/3 truncation with ".ge.",
 * optional spherical truncation at (N/4)^2, and the k = 0 mode.  With the
 * reference k-table the Nyquist index has k = 0 and is therefore never masked;
 * parity_mode = 0 removes the Nyquist-index modes as well. */
int orc_masked(const orc_params* p, int ix, int iy, int iz)
{
    const int nx = p->nx, ny = p->ny, nz = p->nz;
    double rkx = ((ix < nx / 2) ? ix : (ix == nx / 2 ? 0.0 : ix - nx)) * alpha_of(p, 0);
    double rky = ((iy < ny / 2) ? iy : (iy == ny / 2 ? 0.0 : iy - ny)) * alpha_of(p, 1);
    double rkz = ((iz < nz / 2) ? iz : (iz == nz / 2 ? 0.0 : iz - nz)) * alpha_of(p, 2);
    const double tx = thr_of(p, 0), ty = thr_of(p, 1), tz = thr_of(p, 2);
    const double ts = ((double)nx / 2.0) * 0.5 * alpha_of(p, 0);
    if (p->spherical && (rkx * rkx + rky * rky + rkz * rkz >= ts * ts)) return 1;
    if (fabs(rkx) >= tx || fabs(rky) >= ty || fabs(rkz) >= tz) return 1;
    if (rkx == 0.0 && rky == 0.0 && rkz == 0.0) return 1;
    if (!p->parity_mode && (ix == nx / 2 || iy == ny / 2 || iz == nz / 2)) return 1;
    return 0;
}

/* ------------------------------------------------------------------------- */
/* FFT (the reference calls p3dfft/FFTW here: rhs.f90:31-65, state.f90:60-70) */
/* ------------------------------------------------------------------------- */

static int smallest_factor(int n)
{
    if (n % 4 == 0) return 4;
    if (n % 2 == 0) return 2;
    for (int f = 3; f * f <= n; f += 2)
        if (n % f == 0) return f;
    return n;
}

/* tw[j] = exp(sign * 2 pi i j / ntw) */
static cplx* make_twiddles(int n, int sign)
{
    cplx* tw = (cplx*)malloc(sizeof(cplx) * (size_t)n);
    for (int j = 0; j < n; ++j) {
        long double a = TWO_PI_L * (long double)j / (long double)n;
        tw[j] = (double)cosl(a) + (double)(sign * sinl(a)) * I;
    }
    return tw;
}

/* decimation in time: X[k + m q] = sum_r W_n^{r(k + m q)} Y_r[k] */
static void fft_rec(const cplx* in, int is, cplx* out, int n, const cplx* tw, int twstep,
                    int ntw)
{
    if (n == 1) { out[0] = in[0]; return; }
    const int pf = smallest_factor(n), m = n / pf;
    for (int r = 0; r < pf; ++r) fft_rec(in + (size_t)r * is, is * pf, out + (size_t)r * m, m, tw,
                                         twstep * pf, ntw);
    cplx y[64];
    cplx* yy = pf <= 64 ? y : (cplx*)malloc(sizeof(cplx) * pf);
    for (int k = 0; k < m; ++k) {
        for (int r = 0; r < pf; ++r) yy[r] = out[(size_t)r * m + k] * tw[((size_t)r * k * twstep) % ntw];
        for (int q = 0; q < pf; ++q) {
            cplx s = 0;
            for (int r = 0; r < pf; ++r) s += yy[r] * tw[((size_t)m * ((r * q) % pf) * twstep) % ntw];
            out[(size_t)k + (size_t)m * q] = s;
        }
    }
    if (yy != y) free(yy);
}

static void fft_line(cplx* data, int n, const cplx* tw, cplx* buf)
{
    fft_rec(data, 1, buf, n, tw, 1, n);
    memcpy(data, buf, sizeof(cplx) * (size_t)n);
}

void orc_fft1d(double* data, int n, int sign)
{
    cplx* tw = make_twiddles(n, sign);
    cplx* buf = (cplx*)malloc(sizeof(cplx) * (size_t)n);
    fft_line((cplx*)data, n, tw, buf);
    free(buf);
    free(tw);
}

static void set_threads(const orc_params* p)
{
#ifdef _OPENMP
    if (p->nthreads > 0) omp_set_num_threads(p->nthreads);
#else
    (void)p;
#endif
}

/* transform along ky and kz of the half spectrum [kx][ky][kz] */
static void spec_yz(const orc_params* p, cplx* s, int sign)
{
    const int hx = nxh(p), ny = p->ny, nz = p->nz;
    cplx* twy = make_twiddles(ny, sign);
    cplx* twz = make_twiddles(nz, sign);
#pragma omp parallel
    {
        int nmax = ny > nz ? ny : nz;
        cplx* line = (cplx*)malloc(sizeof(cplx) * nmax);
        cplx* buf = (cplx*)malloc(sizeof(cplx) * nmax);
#pragma omp for collapse(2) schedule(static)
        for (int ix = 0; ix < hx; ++ix)
            for (int iz = 0; iz < nz; ++iz) {
                for (int iy = 0; iy < ny; ++iy) line[iy] = s[sidx(p, ix, iy, iz)];
                fft_line(line, ny, twy, buf);
                for (int iy = 0; iy < ny; ++iy) s[sidx(p, ix, iy, iz)] = line[iy];
            }
#pragma omp for collapse(2) schedule(static)
        for (int ix = 0; ix < hx; ++ix)
            for (int iy = 0; iy < ny; ++iy) fft_line(s + sidx(p, ix, iy, 0), nz, twz, buf);
        free(line);
        free(buf);
    }
    free(twy);
    free(twz);
}

/* p3dfft_ftran_r2c: real (Nx,Ny,Nz) -> complex (Nz,Ny,Nx/2+1) STRIDE1, unnormalised */
void orc_r2c(const orc_params* p, const double* phys, double* spec_)
{
    set_threads(p);
    cplx* spec = (cplx*)spec_;
    const int nx = p->nx, ny = p->ny, nz = p->nz, hx = nxh(p);
    cplx* twx = make_twiddles(nx, -1);
#pragma omp parallel
    {
        cplx* line = (cplx*)malloc(sizeof(cplx) * nx);
        cplx* buf = (cplx*)malloc(sizeof(cplx) * nx);
#pragma omp for collapse(2) schedule(static)
        for (int z = 0; z < nz; ++z)
            for (int y = 0; y < ny; ++y) {
                for (int x = 0; x < nx; ++x) line[x] = phys[pidx(p, x, y, z)];
                fft_line(line, nx, twx, buf);
                for (int ix = 0; ix < hx; ++ix) spec[sidx(p, ix, y, z)] = line[ix];
            }
        free(line);
        free(buf);
    }
    free(twx);
    spec_yz(p, spec, -1);
}

/* p3dfft_btran_c2r: complex (Nz,Ny,Nx/2+1) -> real (Nx,Ny,Nz), unnormalised.
 * Imaginary parts of the kx = 0 and kx = Nx/2 entries of each x-line are
 * ignored (the usual c2r convention).  The input is not modified. */
void orc_c2r(const orc_params* p, const double* spec_, double* phys)
{
    set_threads(p);
    const int nx = p->nx, ny = p->ny, nz = p->nz, hx = nxh(p);
    cplx* s = (cplx*)malloc(sizeof(cplx) * nspec_elems(p));
    memcpy(s, spec_, sizeof(cplx) * nspec_elems(p));
    spec_yz(p, s, +1);
    cplx* twx = make_twiddles(nx, +1);
#pragma omp parallel
    {
        cplx* line = (cplx*)malloc(sizeof(cplx) * nx);
        cplx* buf = (cplx*)malloc(sizeof(cplx) * nx);
#pragma omp for collapse(2) schedule(static)
        for (int z = 0; z < nz; ++z)
            for (int y = 0; y < ny; ++y) {
                for (int ix = 0; ix < hx; ++ix) line[ix] = s[sidx(p, ix, y, z)];
                line[0] = creal(line[0]);
                line[nx / 2] = creal(line[nx / 2]);
                for (int ix = 1; ix < nx / 2; ++ix) line[nx - ix] = conj(line[ix]);
                fft_line(line, nx, twx, buf);
                for (int x = 0; x < nx; ++x) phys[pidx(p, x, y, z)] = creal(line[x]);
            }
        free(line);
        free(buf);
    }
    free(twx);
    free(s);
}

/* ------------------------------------------------------------------------- */
/* RHS                                                                        */
/* ------------------------------------------------------------------------- */

typedef struct { double* kx; double* ky; double* kz; } ktabs;

static ktabs ktabs_make(const orc_params* p)
{
    ktabs t;
    t.kx = (double*)malloc(sizeof(double) * p->nx);
    t.ky = (double*)malloc(sizeof(double) * p->ny);
    t.kz = (double*)malloc(sizeof(double) * p->nz);
    orc_ktable(p->nx, t.kx);
    orc_ktable(p->ny, t.ky);
    orc_ktable(p->nz, t.kz);
    /* variables.f90:226-262: k = i (i-1) alpha */
    for (int i = 0; i < p->nx; ++i) t.kx[i] *= alpha_of(p, 0);
    for (int i = 0; i < p->ny; ++i) t.ky[i] *= alpha_of(p, 1);
    for (int i = 0; i < p->nz; ++i) t.kz[i] *= alpha_of(p, 2);
    return t;
}
static void ktabs_free(ktabs* t) { free(t->kx); free(t->ky); free(t->kz); }

/* rhs.f90:137-168 rhs_Eband: (1/N^6) sum_{|k| <= kCutOff} w |u_hat|^2, w = 1/2 where kx = 0
 * (kx-index 0 and the kx-Nyquist index, whose table value is 0). */
double orc_eband(const orc_params* p, const double* uh_, const double* vh_, const double* wh_)
{
    const cplx *uh = (const cplx*)uh_, *vh = (const cplx*)vh_, *wh = (const cplx*)wh_;
    ktabs t = ktabs_make(p);
    const double sm = 1.0 / ((double)p->nx * p->ny * p->nz);
    double e = 0.0;
    for (int ix = 0; ix < nxh(p); ++ix)
        for (int iy = 0; iy < p->ny; ++iy)
            for (int iz = 0; iz < p->nz; ++iz) {
                double absk = sqrt(t.kx[ix] * t.kx[ix] + t.ky[iy] * t.ky[iy] + t.kz[iz] * t.kz[iz]);
                if (absk > p->kcut) continue;
                double sx = (t.kx[ix] == 0.0) ? 0.5 : 1.0;
                size_t m = sidx(p, ix, iy, iz);
                e += creal(conj(uh[m]) * uh[m]) * sx + creal(conj(vh[m]) * vh[m]) * sx +
                     creal(conj(wh[m]) * wh[m]) * sx;
            }
    ktabs_free(&t);
    return e * sm * sm;
}

/* state.f90:456-500 state_derivatives for one component: d/dx_j of f from f_hat */
static void derivs(const orc_params* p, const ktabs* t, const cplx* fh, cplx* tmp, double* dx,
                   double* dy, double* dz)
{
    const double sm = 1.0 / ((double)p->nx * p->ny * p->nz);
    double* outs[3] = {dx, dy, dz};
    for (int d = 0; d < 3; ++d) {
        for (int ix = 0; ix < nxh(p); ++ix)
            for (int iy = 0; iy < p->ny; ++iy)
                for (int iz = 0; iz < p->nz; ++iz) {
                    double kr = d == 0 ? t->kx[ix] : (d == 1 ? t->ky[iy] : t->kz[iz]);
                    size_t m = sidx(p, ix, iy, iz);
                    tmp[m] = fh[m] * (kr * I) * sm;
                }
        orc_c2r(p, (const double*)tmp, outs[d]);
    }
}

/* rhs.f90:18-135 rhs_nonlinear (rotational form), plus the Kolmogorov body force
 * f = (amp sin(kf y), 0, 0) that the workloads add:  f_hat_x(0, +-kf, 0) = -+ i amp N^3/2. */
void orc_rhs_nonlinear(const orc_params* p, const double* uh_, const double* vh_,
                       const double* wh_, double* nu_, double* nv_, double* nw_)
{
    const cplx *uh = (const cplx*)uh_, *vh = (const cplx*)vh_, *wh = (const cplx*)wh_;
    cplx *nu = (cplx*)nu_, *nv = (cplx*)nv_, *nw = (cplx*)nw_;
    const size_t ms = nspec_elems(p), mp = nphys_elems(p);
    const double sm = 1.0 / ((double)p->nx * p->ny * p->nz);
    ktabs t = ktabs_make(p);
    double eband = 0.0;
    if (p->bandlim) eband = orc_eband(p, uh_, vh_, wh_);

    double* d = (double*)malloc(sizeof(double) * mp * 9);
    double *ux = d, *uy = d + mp, *uz = d + 2 * mp, *vx = d + 3 * mp, *vy = d + 4 * mp,
           *vz = d + 5 * mp, *wx = d + 6 * mp, *wy = d + 7 * mp, *wz = d + 8 * mp;
    cplx* tmp = (cplx*)malloc(sizeof(cplx) * ms);
    derivs(p, &t, uh, tmp, ux, uy, uz);
    derivs(p, &t, vh, tmp, vx, vy, vz);
    derivs(p, &t, wh, tmp, wx, wy, wz);
    /* state.f90:502-515 state_vorticity (stored over the first three derivative slots) */
    double* om = (double*)malloc(sizeof(double) * mp * 3);
    double *ox = om, *oy = om + mp, *oz = om + 2 * mp;
    for (size_t i = 0; i < mp; ++i) {
        ox[i] = wy[i] - vz[i];
        oy[i] = uz[i] - wx[i];
        oz[i] = vx[i] - uy[i];
    }
    free(d);
    /* rhs.f90:30-37: velocities with the N^3 factor */
    double* vel = (double*)malloc(sizeof(double) * mp * 3);
    double *ut = vel, *vt = vel + mp, *wt = vel + 2 * mp;
    orc_c2r(p, uh_, ut);
    orc_c2r(p, vh_, vt);
    orc_c2r(p, wh_, wt);
    /* rhs.f90:39-65: (omega x u) / N^3 then r2c */
    double* tr = (double*)malloc(sizeof(double) * mp);
    for (size_t i = 0; i < mp; ++i) tr[i] = (oy[i] * wt[i] - oz[i] * vt[i]) * sm;
    orc_r2c(p, tr, nu_);
    for (size_t i = 0; i < mp; ++i) tr[i] = (oz[i] * ut[i] - ox[i] * wt[i]) * sm;
    orc_r2c(p, tr, nv_);
    for (size_t i = 0; i < mp; ++i) tr[i] = (ox[i] * vt[i] - oy[i] * ut[i]) * sm;
    orc_r2c(p, tr, nw_);
    free(tr);
    free(vel);
    free(om);

    /* rhs.f90:67-133: mask, pressure, projection, band forcing */
    for (int ix = 0; ix < nxh(p); ++ix)
        for (int iy = 0; iy < p->ny; ++iy)
            for (int iz = 0; iz < p->nz; ++iz) {
                size_t m = sidx(p, ix, iy, iz);
                const cplx kx = t.kx[ix] * I, ky = t.ky[iy] * I, kz = t.kz[iz] * I;
                double absk = sqrt(t.kx[ix] * t.kx[ix] + t.ky[iy] * t.ky[iy] + t.kz[iz] * t.kz[iz]);
                if (orc_masked(p, ix, iy, iz)) {
                    nu[m] = 0; nv[m] = 0; nw[m] = 0;
                    continue;
                }
                cplx phat = -1.0 * (kx * nu[m] + ky * nv[m] + kz * nw[m]) / (kx * kx + ky * ky + kz * kz);
                if (p->bandlim && absk <= p->kcut) {
                    const double f = p->Pin / (2.0 * eband);
                    nu[m] = -nu[m] - kx * phat + f * uh[m];
                    nv[m] = -nv[m] - ky * phat + f * vh[m];
                    nw[m] = -nw[m] - kz * phat + f * wh[m];
                } else {
                    nu[m] = -nu[m] - kx * phat;
                    nv[m] = -nv[m] - ky * phat;
                    nw[m] = -nw[m] - kz * phat;
                }
            }
    /* build addition: Kolmogorov forcing */
    if (p->kolmo_k > 0 && p->kolmo_amp != 0.0) {
        const double n3 = (double)p->nx * p->ny * p->nz;
        int ip = p->kolmo_k, im = p->ny - p->kolmo_k;
        if (!orc_masked(p, 0, ip, 0)) nu[sidx(p, 0, ip, 0)] += -0.5 * p->kolmo_amp * n3 * I;
        if (!orc_masked(p, 0, im, 0)) nu[sidx(p, 0, im, 0)] += 0.5 * p->kolmo_amp * n3 * I;
    }
    free(tmp);
    ktabs_free(&t);
}

/* rhs.f90:8-16 rhs_int_fact: exp((nu (kx^2+ky^2+kz^2) + Q) dt) with imaginary k */
void orc_int_fact(const orc_params* p, double dt, double* intfact)
{
    ktabs t = ktabs_make(p);
    for (int ix = 0; ix < nxh(p); ++ix)
        for (int iy = 0; iy < p->ny; ++iy)
            for (int iz = 0; iz < p->nz; ++iz) {
                const cplx kx = t.kx[ix] * I, ky = t.ky[iy] * I, kz = t.kz[iz] * I;
                cplx e = cexp((p->nu * (kx * kx + ky * ky + kz * kz) + p->Q) * dt);
                intfact[sidx(p, ix, iy, iz)] = creal(e);
            }
    ktabs_free(&t);
}

/* dnsboxheun.f90:84-126,144: integrating-factor Heun predictor-corrector, fixed dt */
void orc_heun(const orc_params* p, double* uh_, double* vh_, double* wh_, int nsteps, double dt)
{
    const size_t ms = nspec_elems(p);
    cplx* f[3] = {(cplx*)uh_, (cplx*)vh_, (cplx*)wh_};
    double* itf = (double*)malloc(sizeof(double) * ms);
    orc_int_fact(p, dt, itf);
    cplx* tmp = (cplx*)malloc(sizeof(cplx) * ms * 3);
    cplx* nl = (cplx*)malloc(sizeof(cplx) * ms * 3);
    for (int n = 0; n < nsteps; ++n) {
        orc_rhs_nonlinear(p, uh_, vh_, wh_, (double*)nl, (double*)(nl + ms), (double*)(nl + 2 * ms));
        for (int c = 0; c < 3; ++c)
            for (size_t m = 0; m < ms; ++m) {
                tmp[c * ms + m] = itf[m] * (f[c][m] + dt * nl[c * ms + m]);
                f[c][m] = itf[m] * (f[c][m] + dt * nl[c * ms + m] * 0.5);
            }
        orc_rhs_nonlinear(p, (double*)tmp, (double*)(tmp + ms), (double*)(tmp + 2 * ms), (double*)nl,
                          (double*)(nl + ms), (double*)(nl + 2 * ms));
        for (int c = 0; c < 3; ++c)
            for (size_t m = 0; m < ms; ++m) f[c][m] = f[c][m] + dt * nl[c * ms + m] * 0.5;
    }
    free(nl);
    free(tmp);
    free(itf);
}

/* Integrating-factor AB2 at constant dt.  dnsbox.f90:143-193 codes a variable-step
 * IF-AB2 whose weights sum to 1.5 dt at constant dt (inconsistent); this is the
 * consistent scheme  u_{n+1} = E u_n + dt (1.5 E N_n - 0.5 E^2 N_{n-1}),  E = e^{L dt},
 * started with one IF-Heun step as dnsbox.f90:87-141 does. */
void orc_ab2(const orc_params* p, double* uh_, double* vh_, double* wh_, int nsteps, double dt)
{
    if (nsteps <= 0) return;
    const size_t ms = nspec_elems(p);
    cplx* f[3] = {(cplx*)uh_, (cplx*)vh_, (cplx*)wh_};
    double* itf = (double*)malloc(sizeof(double) * ms);
    orc_int_fact(p, dt, itf);
    cplx* nold = (cplx*)malloc(sizeof(cplx) * ms * 3);
    cplx* nl = (cplx*)malloc(sizeof(cplx) * ms * 3);
    orc_rhs_nonlinear(p, uh_, vh_, wh_, (double*)nold, (double*)(nold + ms), (double*)(nold + 2 * ms));
    orc_heun(p, uh_, vh_, wh_, 1, dt);
    for (int n = 1; n < nsteps; ++n) {
        orc_rhs_nonlinear(p, uh_, vh_, wh_, (double*)nl, (double*)(nl + ms), (double*)(nl + 2 * ms));
        for (int c = 0; c < 3; ++c)
            for (size_t m = 0; m < ms; ++m) {
                const double e = itf[m];
                f[c][m] = e * f[c][m] + dt * (1.5 * e * nl[c * ms + m] - 0.5 * e * e * nold[c * ms + m]);
                nold[c * ms + m] = nl[c * ms + m];
            }
    }
    free(nl);
    free(nold);
    free(itf);
}

/* One variable-step IF-AB2 step of size h after a step of size hp (the history N_{n-1} in nold,
 * replaced by N_n).  dnsbox.f90:143-193 integrates v = e^{-Lt} u with AB2 weights written as
 * 0.5 (2 t_n + dt - 2 t_{n-1}) and 0.5 (t_n - dt - t_{n-1}); the consistent variable-step
 * weights of the same scheme are h (1 + h/2h') and -h^2/2h':
 *   u_{n+1} = E_h u_n + h (1 + h/2h') E_h N_n - (h^2/2h') E_{h+h'} N_{n-1},  E_s = e^{(-nu|k|^2+Q) s} */
static void ab2_step_var(const orc_params* p, cplx* f[3], cplx* nold, double h, double hp)
{
    const size_t ms = nspec_elems(p);
    cplx* nl = (cplx*)malloc(sizeof(cplx) * ms * 3);
    double* e = (double*)malloc(sizeof(double) * ms);
    double* e2 = (double*)malloc(sizeof(double) * ms);
    orc_rhs_nonlinear(p, (const double*)f[0], (const double*)f[1], (const double*)f[2], (double*)nl,
                      (double*)(nl + ms), (double*)(nl + 2 * ms));
    orc_int_fact(p, h, e);
    orc_int_fact(p, h + hp, e2);
    const double r = h / (2.0 * hp);
    for (int c = 0; c < 3; ++c)
        for (size_t m = 0; m < ms; ++m) {
            f[c][m] = e[m] * f[c][m] + h * (1.0 + r) * e[m] * nl[c * ms + m] - h * r * e2[m] * nold[c * ms + m];
            nold[c * ms + m] = nl[c * ms + m];
        }
    free(e2);
    free(e);
    free(nl);
}

/* Variable-step IF-AB2 over the step sizes dts[0..nsteps-1]: the first step is IF-Heun
 * (dnsbox.f90:87-141, keeping N(u_0) as the history), then ab2_step_var with h = dts[n],
 * h' = dts[n-1]. */
void orc_ab2_var(const orc_params* p, double* uh_, double* vh_, double* wh_, int nsteps, const double* dts)
{
    if (nsteps <= 0) return;
    const size_t ms = nspec_elems(p);
    cplx* f[3] = {(cplx*)uh_, (cplx*)vh_, (cplx*)wh_};
    cplx* nold = (cplx*)malloc(sizeof(cplx) * ms * 3);
    orc_rhs_nonlinear(p, uh_, vh_, wh_, (double*)nold, (double*)(nold + ms), (double*)(nold + 2 * ms));
    orc_heun(p, uh_, vh_, wh_, 1, dts[0]);
    for (int n = 1; n < nsteps; ++n) ab2_step_var(p, f, nold, dts[n], dts[n - 1]);
    free(nold);
}

/* orc_ab2_var with the step sizes from the CFL controller applied before every step on the
 * current field (as orc_heun_adaptive does): h_n = set_time_step(Courant(u_n, h_{n-1})). */
void orc_ab2_adaptive(const orc_params* p, double* uh_, double* vh_, double* wh_, int nsteps, double* dt_inout,
                      double cmin, double cmax, double dtmax, double* t_out)
{
    double dt = *dt_inout, t = t_out ? *t_out : 0.0;
    if (nsteps <= 0) return;
    const size_t ms = nspec_elems(p);
    cplx* f[3] = {(cplx*)uh_, (cplx*)vh_, (cplx*)wh_};
    cplx* nold = (cplx*)malloc(sizeof(cplx) * ms * 3);
    dt = orc_cfl_dt(orc_courant(p, uh_, vh_, wh_, dt), dt, cmin, cmax, dtmax);
    orc_rhs_nonlinear(p, uh_, vh_, wh_, (double*)nold, (double*)(nold + ms), (double*)(nold + 2 * ms));
    orc_heun(p, uh_, vh_, wh_, 1, dt);
    t += dt;
    for (int n = 1; n < nsteps; ++n) {
        const double h = orc_cfl_dt(orc_courant(p, uh_, vh_, wh_, dt), dt, cmin, cmax, dtmax);
        ab2_step_var(p, f, nold, h, dt);
        t += h;
        dt = h;
    }
    free(nold);
    *dt_inout = dt;
    if (t_out) *t_out = t;
}

/* state.f90:120-179 state_project (eps = 0 in the forward build) */
void orc_project(const orc_params* p, double* uh_, double* vh_, double* wh_)
{
    cplx *uh = (cplx*)uh_, *vh = (cplx*)vh_, *wh = (cplx*)wh_;
    ktabs t = ktabs_make(p);
    for (int ix = 0; ix < nxh(p); ++ix)
        for (int iy = 0; iy < p->ny; ++iy)
            for (int iz = 0; iz < p->nz; ++iz) {
                const cplx kx = t.kx[ix] * I, ky = t.ky[iy] * I, kz = t.kz[iz] * I;
                if (t.kx[ix] == 0.0 && t.ky[iy] == 0.0 && t.kz[iz] == 0.0) continue;
                size_t m = sidx(p, ix, iy, iz);
                const cplx kk = kx * kx + ky * ky + kz * kz;
                cplx u = uh[m], v = vh[m], w = wh[m];
                uh[m] = u - kx * kx * u / kk - kx * ky * v / kk - kx * kz * w / kk;
                vh[m] = v - ky * kx * u / kk - ky * ky * v / kk - ky * kz * w / kk;
                wh[m] = w - kz * kx * u / kk - kz * ky * v / kk - kz * kz * w / kk;
            }
    ktabs_free(&t);
}

/* state.f90:181-209 state_dealias: cubic .ge. N/3 rule, spherical (N/3)^2, k = 0 */
void orc_dealias(const orc_params* p, double* uh_, double* vh_, double* wh_)
{
    cplx *uh = (cplx*)uh_, *vh = (cplx*)vh_, *wh = (cplx*)wh_;
    ktabs t = ktabs_make(p);
    const double tx = thr_of(p, 0), ty = thr_of(p, 1), tz = thr_of(p, 2);
    for (int ix = 0; ix < nxh(p); ++ix)
        for (int iy = 0; iy < p->ny; ++iy)
            for (int iz = 0; iz < p->nz; ++iz) {
                double kx = t.kx[ix], ky = t.ky[iy], kz = t.kz[iz];
                int m0 = (p->spherical && kx * kx + ky * ky + kz * kz >= tx * tx) ||
                         fabs(kx) >= tx || fabs(ky) >= ty || fabs(kz) >= tz ||
                         (kx == 0.0 && ky == 0.0 && kz == 0.0) ||
                         (!p->parity_mode && (ix == p->nx / 2 || iy == p->ny / 2 || iz == p->nz / 2));
                if (m0) {
                    size_t m = sidx(p, ix, iy, iz);
                    uh[m] = 0; vh[m] = 0; wh[m] = 0;
                }
            }
    ktabs_free(&t);
}

/* io.f90:664-676 io_Courant, evaluated on the physical velocity of the given state */
double orc_courant(const orc_params* p, const double* uh_, const double* vh_, const double* wh_,
                   double dt)
{
    const size_t mp = nphys_elems(p);
    const double sm = 1.0 / ((double)p->nx * p->ny * p->nz);
    double* u = (double*)malloc(sizeof(double) * mp);
    const double* in[3] = {uh_, vh_, wh_};
    const int nn[3] = {p->nx, p->ny, p->nz};
    double c = 0.0;
    for (int d = 0; d < 3; ++d) {
        orc_c2r(p, in[d], u);
        double mx = 0.0;
        for (size_t i = 0; i < mp; ++i) mx = fmax(mx, fabs(u[i] * sm));
        c += mx * dt / ((2.0 * M_PI / alpha_of(p, d)) / (double)nn[d]);   /* dt / (L_d / N_d) */
    }
    free(u);
    return c;
}

/* dnsboxheun.f90:153-157,227-246 set_time_step trigger and update */
double orc_cfl_dt(double courant, double dt, double cmin, double cmax, double dtmax)
{
    if (courant > cmax || (courant < cmin && dt < dtmax)) {
        dt = ((cmax + cmin) / (2.0 * courant)) * dt;
        if (dt > dtmax) dt = dtmax;
    }
    return dt;
}

/* IF-Heun with the CFL controller applied before every step on the current field:
 *   dt_n = set_time_step(Courant(u_n, dt_{n-1}))   (dnsboxheun.f90:153-157, 227-246). */
void orc_heun_adaptive(const orc_params* p, double* uh, double* vh, double* wh, int nsteps,
                       double* dt_inout, double cmin, double cmax, double dtmax, double* t_out)
{
    double dt = *dt_inout, t = t_out ? *t_out : 0.0;
    for (int n = 0; n < nsteps; ++n) {
        double c = orc_courant(p, uh, vh, wh, dt);
        dt = orc_cfl_dt(c, dt, cmin, cmax, dtmax);
        orc_heun(p, uh, vh, wh, 1, dt);
        t += dt;
    }
    *dt_inout = dt;
    if (t_out) *t_out = t;
}

/* io.f90:628-661 io_save_stats + state.f90:350-454,517-544:
 * Ekin (state_kinetic), Disp (state_dissipation, physical, 9 c2r), production 2 Q Ekin,
 * Courant (io_Courant), divMax (state_divergence: signed max of ux+vy+wz),
 * EZero/EOne/Esqrt2 and Espec (state_spectrum, with its weights and the Esqrt2 test
 * "absk - sqrt(2) < 1e-9" exactly as written), Eband (rhs_Eband). */
void orc_diagnostics(const orc_params* p, const double* uh_, const double* vh_,
                     const double* wh_, double dt, double* diag, double* espec)
{
    const cplx* f[3] = {(const cplx*)uh_, (const cplx*)vh_, (const cplx*)wh_};
    const size_t ms = nspec_elems(p), mp = nphys_elems(p);
    const double sm = 1.0 / ((double)p->nx * p->ny * p->nz);
    const double deltak = p->deltak > 0 ? p->deltak : 1.0;
    const int nspec = orc_nspec(p);
    ktabs t = ktabs_make(p);
    double ekin = 0.0, ez = 0.0, e1 = 0.0, es2 = 0.0;
    for (int i = 0; i < nspec; ++i) espec[i] = 0.0;
    for (int ix = 0; ix < nxh(p); ++ix)
        for (int iy = 0; iy < p->ny; ++iy)
            for (int iz = 0; iz < p->nz; ++iz) {
                size_t m = sidx(p, ix, iy, iz);
                double e = 0.0;
                for (int c = 0; c < 3; ++c) e += creal(conj(f[c][m]) * f[c][m]);
                double sx = (t.kx[ix] == 0.0) ? 0.5 : 1.0;
                ekin += e * sx;
                double absk = sqrt(t.kx[ix] * t.kx[ix] + t.ky[iy] * t.ky[iy] + t.kz[iz] * t.kz[iz]);
                int iq = (int)(absk / deltak);
                int nk = (absk < (double)(float)iq + 0.5 * deltak) ? iq : iq + 1;
                if (absk == 0.0) ez += e;
                if (absk == 1.0) e1 += e;
                if (absk - sqrt(2.0) < 1.0e-9) es2 += e;
                if (nk >= nspec || nk == 0) continue;
                espec[nk - 1] += e * (sm * sm) / deltak;
            }
    ekin *= sm * sm;
    /* dissipation: 9 derivative c2r, physical sum */
    double* d = (double*)malloc(sizeof(double) * mp * 3);
    cplx* tmp = (cplx*)malloc(sizeof(cplx) * ms);
    double disp = 0.0;
    double* div = (double*)calloc(mp, sizeof(double));
    for (int c = 0; c < 3; ++c) {
        derivs(p, &t, f[c], tmp, d, d + mp, d + 2 * mp);
        for (size_t i = 0; i < mp * 3; ++i) disp += d[i] * d[i];
        for (size_t i = 0; i < mp; ++i) div[i] += d[c * mp + i];
    }
    disp *= p->nu * sm;
    double divmax = div[0];
    for (size_t i = 1; i < mp; ++i) divmax = fmax(divmax, div[i]);
    free(div);
    free(tmp);
    free(d);
    diag[ORC_EKIN] = ekin;
    diag[ORC_DISP] = disp;
    diag[ORC_PROD] = 2.0 * p->Q * ekin;
    diag[ORC_COURANT] = orc_courant(p, uh_, vh_, wh_, dt);
    diag[ORC_DIVMAX] = divmax;
    diag[ORC_EZERO] = ez;
    diag[ORC_EONE] = e1;
    diag[ORC_ESQRT2] = es2;
    diag[ORC_EBAND] = orc_eband(p, uh_, vh_, wh_);
    ktabs_free(&t);
}

/* ------------------------------------------------------------------------- */
/* initial fields                                                             */
/* ------------------------------------------------------------------------- */

static inline uint64_t splitmix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

/* counter-based generator keyed by (seed, box, counter) */
uint64_t orc_hash(uint64_t seed, uint64_t box, uint64_t idx)
{
    return splitmix64(seed ^ splitmix64(box ^ splitmix64(idx)));
}

static inline int signed_k(int i, int n) { return i <= n / 2 ? i : i - n; }

/* Seeded solenoidal random field (the workloads' initial condition):
 * every retained mode k != 0 off the Nyquist-index planes gets a random complex
 * 3-vector a_c = exp(2 pi i phi_c), phi_c = hash(seed, box, 3*canonical_index + c) / 2^64,
 * projected a - k (k.a)/|k|^2, normalised, scaled by sqrt(E(|k|) / (4 pi |k|^2)),
 * E(k) = k^4 exp(-2 (k/k0)^2); Hermitian symmetry is imposed on the kx = 0 plane;
 * finally the field is rescaled so that the state_kinetic energy equals `energy`.
 * (Replaces the unseeded state.f90:211-285 Yoffe initialisation.) */
void orc_init_solenoidal(const orc_params* p, uint64_t seed, int box, double k0, double energy,
                         double* uh_, double* vh_, double* wh_)
{
    cplx* f[3] = {(cplx*)uh_, (cplx*)vh_, (cplx*)wh_};
    const int nx = p->nx, ny = p->ny, nz = p->nz;
    for (int ix = 0; ix < nxh(p); ++ix)
        for (int iy = 0; iy < ny; ++iy)
            for (int iz = 0; iz < nz; ++iz) {
                size_t m = sidx(p, ix, iy, iz);
                f[0][m] = f[1][m] = f[2][m] = 0;
                if (ix == nx / 2 || iy == ny / 2 || iz == nz / 2) continue;
                if (orc_masked(p, ix, iy, iz)) continue;
                int kx = signed_k(ix, nx), ky = signed_k(iy, ny), kz = signed_k(iz, nz);
                int conjugate = 0;
                int cy = iy, cz = iz;
                if (ix == 0 && !(ky > 0 || (ky == 0 && kz > 0))) {
                    conjugate = 1;
                    cy = (ny - iy) % ny;
                    cz = (nz - iz) % nz;
                    ky = -ky;
                    kz = -kz;
                }
                uint64_t cidx = ((uint64_t)ix * ny + cy) * nz + cz;
                cplx a[3];
                for (int c = 0; c < 3; ++c) {
                    uint64_t h = orc_hash(seed, (uint64_t)box, 3 * cidx + c);
                    double ph = 2.0 * M_PI * ((double)(h >> 11) * (1.0 / 9007199254740992.0));
                    a[c] = cos(ph) + sin(ph) * I;
                }
                const double kpx = kx * alpha_of(p, 0), kpy = ky * alpha_of(p, 1), kpz = kz * alpha_of(p, 2);
                const double k2 = kpx * kpx + kpy * kpy + kpz * kpz;
                cplx kd = (kpx * a[0] + kpy * a[1] + kpz * a[2]) / k2;
                a[0] -= kpx * kd;
                a[1] -= kpy * kd;
                a[2] -= kpz * kd;
                double na = sqrt(creal(conj(a[0]) * a[0] + conj(a[1]) * a[1] + conj(a[2]) * a[2]));
                if (na < 1e-300) continue;
                double kk = sqrt(k2);
                double ek = kk * kk * kk * kk * exp(-2.0 * (kk / k0) * (kk / k0));
                double amp = sqrt(ek / (4.0 * M_PI * k2)) / na;
                for (int c = 0; c < 3; ++c) f[c][m] = conjugate ? conj(a[c] * amp) : a[c] * amp;
            }
    double ekin = 0.0;
    for (int ix = 0; ix < nxh(p); ++ix) {
        double sx = (ix == 0 || ix == nx / 2) ? 0.5 : 1.0;
        for (int iy = 0; iy < ny; ++iy)
            for (int iz = 0; iz < nz; ++iz) {
                size_t m = sidx(p, ix, iy, iz);
                for (int c = 0; c < 3; ++c) ekin += sx * creal(conj(f[c][m]) * f[c][m]);
            }
    }
    const double sm = 1.0 / ((double)nx * ny * nz);
    ekin *= sm * sm;
    const double scale = ekin > 0 ? sqrt(energy / ekin) : 0.0;
    const size_t ms = nspec_elems(p);
    for (int c = 0; c < 3; ++c)
        for (size_t m = 0; m < ms; ++m) f[c][m] *= scale;
}

/* state.f90:287-307 state_init_analytic (Shapiro/Beltrami field), grid from -pi
 * (variables.f90:235-239) */
void orc_init_shapiro(const orc_params* p, double t, double* u, double* v, double* w)
{
    const double factor = sqrt(3.0);
    const double e = exp(-(factor * factor) * t * p->nu);
    for (int k = 0; k < p->nz; ++k)
        for (int j = 0; j < p->ny; ++j)
            for (int i = 0; i < p->nx; ++i) {
                double x = (double)(i - p->nx / 2) * (2.0 * M_PI / alpha_of(p, 0)) / (double)p->nx;
                double y = (double)(j - p->ny / 2) * (2.0 * M_PI / alpha_of(p, 1)) / (double)p->ny;
                double z = (double)(k - p->nz / 2) * (2.0 * M_PI / alpha_of(p, 2)) / (double)p->nz;
                size_t m = pidx(p, i, j, k);
                u[m] = -0.5 * (factor * cos(x) * sin(y) * sin(z) + sin(x) * cos(y) * cos(z)) * e;
                v[m] = 0.5 * (factor * sin(x) * cos(y) * sin(z) - cos(x) * sin(y) * cos(z)) * e;
                w[m] = 1.0 * cos(x) * cos(y) * sin(z) * e;
            }
}

/* state.f90:309-348 state_check_error: sum over components of max |u - u_exact| */
double orc_shapiro_error(const orc_params* p, double t, const double* u, const double* v,
                         const double* w)
{
    const size_t mp = nphys_elems(p);
    double* ex = (double*)malloc(sizeof(double) * mp * 3);
    orc_init_shapiro(p, t, ex, ex + mp, ex + 2 * mp);
    double mu = 0, mv = 0, mw = 0;
    for (size_t i = 0; i < mp; ++i) {
        mu = fmax(mu, fabs(u[i] - ex[i]));
        mv = fmax(mv, fabs(v[i] - ex[mp + i]));
        mw = fmax(mw, fabs(w[i] - ex[2 * mp + i]));
    }
    free(ex);
    return mu + mv + mw;
}

/* ------------------------------------------------------------------------- */
/* adjoints                                                                   */
/* ------------------------------------------------------------------------- */

/* Hermitian weight of a stored half-spectrum entry: the c2r sums every kx plane
 * except kx = 0 and kx = Nx/2 twice. */
static inline double wc(const orc_params* p, int ix) { return (ix == 0 || ix == p->nx / 2) ? 1.0 : 2.0; }

/* Exact discrete vector-Jacobian product of rhs_nonlinear (rotational form, rhs.f90:18-135
 * plus the Kolmogorov body force, which is constant) in the Euclidean inner product of the
 * stored half-spectrum coordinates, <a, b> = sum Re(conj(a) b):
 *     g = J_N(u)^T lambda.
 * Chain (forward: u_t = c2r(u), w = c2r(s i k x u), y = s (w x u_t), n = r2c(y),
 *        N = -P M n [+ Pin/(2 Eband(u)) u on the band]):
 *     nbar = -P M lambda;  ybar = c2r(nbar / W);  wbar = s (u_t x ybar);  utbar = s (ybar x w)
 *     g = W r2c(utbar) + s (i k x)(W r2c(wbar)) [+ band terms],  W = Hermitian weight. */
void orc_rhs_vjp(const orc_params* p, const double* uh_, const double* vh_, const double* wh_,
                 const double* lu_, const double* lv_, const double* lw_, double* gu_, double* gv_,
                 double* gw_)
{
    const cplx* u[3] = {(const cplx*)uh_, (const cplx*)vh_, (const cplx*)wh_};
    const cplx* lam[3] = {(const cplx*)lu_, (const cplx*)lv_, (const cplx*)lw_};
    cplx* g[3] = {(cplx*)gu_, (cplx*)gv_, (cplx*)gw_};
    const size_t ms = nspec_elems(p), mp = nphys_elems(p);
    const double sm = 1.0 / ((double)p->nx * p->ny * p->nz);
    ktabs t = ktabs_make(p);

    /* forward intermediates: u_t (N^3 scale) and the vorticity */
    double* ut = (double*)malloc(sizeof(double) * mp * 3);
    double* om = (double*)malloc(sizeof(double) * mp * 3);
    cplx* tmp = (cplx*)malloc(sizeof(cplx) * ms * 3);
    for (int c = 0; c < 3; ++c) orc_c2r(p, (const double*)u[c], ut + c * mp);
    for (int ix = 0; ix < nxh(p); ++ix)
        for (int iy = 0; iy < p->ny; ++iy)
            for (int iz = 0; iz < p->nz; ++iz) {
                size_t m = sidx(p, ix, iy, iz);
                const double kx = t.kx[ix], ky = t.ky[iy], kz = t.kz[iz];
                tmp[m] = I * (ky * u[2][m] - kz * u[1][m]) * sm;
                tmp[ms + m] = I * (kz * u[0][m] - kx * u[2][m]) * sm;
                tmp[2 * ms + m] = I * (kx * u[1][m] - ky * u[0][m]) * sm;
            }
    for (int c = 0; c < 3; ++c) orc_c2r(p, (const double*)(tmp + c * ms), om + c * mp);

    /* nbar = -P M lambda, divided by the Hermitian weight */
    for (int ix = 0; ix < nxh(p); ++ix)
        for (int iy = 0; iy < p->ny; ++iy)
            for (int iz = 0; iz < p->nz; ++iz) {
                size_t m = sidx(p, ix, iy, iz);
                if (orc_masked(p, ix, iy, iz)) {
                    tmp[m] = tmp[ms + m] = tmp[2 * ms + m] = 0;
                    continue;
                }
                const double k[3] = {t.kx[ix], t.ky[iy], t.kz[iz]};
                const double k2 = k[0] * k[0] + k[1] * k[1] + k[2] * k[2];
                const cplx kl = (k[0] * lam[0][m] + k[1] * lam[1][m] + k[2] * lam[2][m]) / k2;
                for (int c = 0; c < 3; ++c) tmp[c * ms + m] = -(lam[c][m] - k[c] * kl) / wc(p, ix);
            }
    double* yb = (double*)malloc(sizeof(double) * mp * 3);
    for (int c = 0; c < 3; ++c) orc_c2r(p, (const double*)(tmp + c * ms), yb + c * mp);

    /* through y = s (w x u_t) */
    double* utb = (double*)malloc(sizeof(double) * mp * 3);
    double* omb = (double*)malloc(sizeof(double) * mp * 3);
    for (size_t i = 0; i < mp; ++i) {
        const double a[3] = {ut[i], ut[mp + i], ut[2 * mp + i]};
        const double w[3] = {om[i], om[mp + i], om[2 * mp + i]};
        const double y[3] = {yb[i], yb[mp + i], yb[2 * mp + i]};
        /* wbar = s (u_t x ybar) */
        omb[i] = sm * (a[1] * y[2] - a[2] * y[1]);
        omb[mp + i] = sm * (a[2] * y[0] - a[0] * y[2]);
        omb[2 * mp + i] = sm * (a[0] * y[1] - a[1] * y[0]);
        /* utbar = s (ybar x w) */
        utb[i] = sm * (y[1] * w[2] - y[2] * w[1]);
        utb[mp + i] = sm * (y[2] * w[0] - y[0] * w[2]);
        utb[2 * mp + i] = sm * (y[0] * w[1] - y[1] * w[0]);
    }
    cplx* Ub = (cplx*)malloc(sizeof(cplx) * ms * 3);
    cplx* Wb = (cplx*)malloc(sizeof(cplx) * ms * 3);
    for (int c = 0; c < 3; ++c) {
        orc_r2c(p, utb + c * mp, (double*)(Ub + c * ms));
        orc_r2c(p, omb + c * mp, (double*)(Wb + c * ms));
    }
    for (int ix = 0; ix < nxh(p); ++ix)
        for (int iy = 0; iy < p->ny; ++iy)
            for (int iz = 0; iz < p->nz; ++iz) {
                size_t m = sidx(p, ix, iy, iz);
                const double kx = t.kx[ix], ky = t.ky[iy], kz = t.kz[iz];
                const double w = wc(p, ix);
                const cplx a0 = w * Wb[m], a1 = w * Wb[ms + m], a2 = w * Wb[2 * ms + m];
                g[0][m] = w * Ub[m] + sm * I * (ky * a2 - kz * a1);
                g[1][m] = w * Ub[ms + m] + sm * I * (kz * a0 - kx * a2);
                g[2][m] = w * Ub[2 * ms + m] + sm * I * (kx * a1 - ky * a0);
            }
    /* band forcing: N += c(u) u on unmasked |k| <= kcut, c = Pin / (2 Eband(u)) */
    if (p->bandlim) {
        const double eb = orc_eband(p, uh_, vh_, wh_);
        const double cf = p->Pin / (2.0 * eb);
        double lu = 0.0;
        for (int ix = 0; ix < nxh(p); ++ix)
            for (int iy = 0; iy < p->ny; ++iy)
                for (int iz = 0; iz < p->nz; ++iz) {
                    const double absk = sqrt(t.kx[ix] * t.kx[ix] + t.ky[iy] * t.ky[iy] + t.kz[iz] * t.kz[iz]);
                    if (absk > p->kcut || orc_masked(p, ix, iy, iz)) continue;
                    size_t m = sidx(p, ix, iy, iz);
                    for (int c = 0; c < 3; ++c) {
                        g[c][m] += cf * lam[c][m];
                        lu += creal(conj(lam[c][m]) * u[c][m]);
                    }
                }
        /* d Eband / d u = 2 s^2 w_e u on |k| <= kcut (w_e = 1/2 where kx = 0) */
        const double f = -p->Pin / (2.0 * eb * eb) * lu;
        for (int ix = 0; ix < nxh(p); ++ix)
            for (int iy = 0; iy < p->ny; ++iy)
                for (int iz = 0; iz < p->nz; ++iz) {
                    const double absk = sqrt(t.kx[ix] * t.kx[ix] + t.ky[iy] * t.ky[iy] + t.kz[iz] * t.kz[iz]);
                    if (absk > p->kcut) continue;
                    size_t m = sidx(p, ix, iy, iz);
                    const double we = (t.kx[ix] == 0.0) ? 0.5 : 1.0;
                    for (int c = 0; c < 3; ++c) g[c][m] += f * 2.0 * sm * sm * we * u[c][m];
                }
    }
    free(Ub); free(Wb); free(utb); free(omb); free(yb); free(tmp); free(om); free(ut);
    ktabs_free(&t);
}

/* Vector-Jacobian product of one fixed-dt IF-Heun step (dnsboxheun.f90:84-126):
 *   u* = E(u + dt N(u)),  u_{n+1} = E(u + dt/2 N(u)) + dt/2 N(u*)
 *   lam_* = J_N(u*)^T (dt/2 lam);  lam_n = E(lam + lam_*) + J_N(u)^T E(dt/2 lam + dt lam_*).
 * lam (3 components) is overwritten with lam_n. */
void orc_heun_vjp(const orc_params* p, const double* uh, const double* vh, const double* wh,
                  double dt, double* lu_, double* lv_, double* lw_)
{
    const size_t ms = nspec_elems(p);
    cplx* lam[3] = {(cplx*)lu_, (cplx*)lv_, (cplx*)lw_};
    const cplx* u[3] = {(const cplx*)uh, (const cplx*)vh, (const cplx*)wh};
    double* itf = (double*)malloc(sizeof(double) * ms);
    orc_int_fact(p, dt, itf);
    cplx* buf = (cplx*)malloc(sizeof(cplx) * ms * 12);
    cplx *nl = buf, *us = buf + 3 * ms, *a = buf + 6 * ms, *ls = buf + 9 * ms;
    orc_rhs_nonlinear(p, uh, vh, wh, (double*)nl, (double*)(nl + ms), (double*)(nl + 2 * ms));
    for (int c = 0; c < 3; ++c)
        for (size_t m = 0; m < ms; ++m) {
            us[c * ms + m] = itf[m] * (u[c][m] + dt * nl[c * ms + m]);
            a[c * ms + m] = 0.5 * dt * lam[c][m];
        }
    orc_rhs_vjp(p, (double*)us, (double*)(us + ms), (double*)(us + 2 * ms), (double*)a,
                (double*)(a + ms), (double*)(a + 2 * ms), (double*)ls, (double*)(ls + ms),
                (double*)(ls + 2 * ms));
    for (int c = 0; c < 3; ++c)
        for (size_t m = 0; m < ms; ++m) a[c * ms + m] = itf[m] * (0.5 * dt * lam[c][m] + dt * ls[c * ms + m]);
    orc_rhs_vjp(p, uh, vh, wh, (double*)a, (double*)(a + ms), (double*)(a + 2 * ms), (double*)nl,
                (double*)(nl + ms), (double*)(nl + 2 * ms));
    for (int c = 0; c < 3; ++c)
        for (size_t m = 0; m < ms; ++m)
            lam[c][m] = itf[m] * (lam[c][m] + ls[c * ms + m]) + nl[c * ms + m];
    free(buf);
    free(itf);
}

/* rhsAdj.f90:64-381 as written, with one correction: the never-assigned phatadj
 * (rhsAdj.f90:343-348) is read as the phat computed just above it (:335-340).
 *   forward (convective form, ".gt." mask, eps at k = 0)   rhsNonlinear :64-175
 *   u~ = H (L u + N(u)),  H = 1/(1+|k|^2), L = -nu|k|^2 + Q  rhsupptilde  :177-213
 *   n_adj_i = sum_j (d_i u~_j + d_j u~_i) u_j               rhsNonlinearAdj :261-309
 *   rhsadj = L u~ - n_adj - k phat(n_adj)                   :311-381 */
static void grads_of(const orc_params* p, const ktabs* t, const cplx* fh, cplx* tmp, double* d)
{
    derivs(p, t, fh, tmp, d, d + nphys_elems(p), d + 2 * nphys_elems(p));
}

static int masked_gt(const orc_params* p, const ktabs* t, int ix, int iy, int iz)
{
    const double tx = thr_of(p, 0), ty = thr_of(p, 1), tz = thr_of(p, 2);
    if (!p->parity_mode && (ix == p->nx / 2 || iy == p->ny / 2 || iz == p->nz / 2)) return 1;
    return fabs(t->kx[ix]) > tx || fabs(t->ky[iy]) > ty || fabs(t->kz[iz]) > tz;
}

static void project_gt(const orc_params* p, const ktabs* t, cplx* n[3])
{
    for (int ix = 0; ix < nxh(p); ++ix)
        for (int iy = 0; iy < p->ny; ++iy)
            for (int iz = 0; iz < p->nz; ++iz) {
                size_t m = sidx(p, ix, iy, iz);
                if (masked_gt(p, t, ix, iy, iz)) {
                    n[0][m] = n[1][m] = n[2][m] = 0;
                    continue;
                }
                const cplx kx = t->kx[ix] * I, ky = t->ky[iy] * I, kz = t->kz[iz] * I;
                const double eps = (t->kx[ix] == 0.0 && t->ky[iy] == 0.0 && t->kz[iz] == 0.0) ? 1e-13 : 0.0;
                cplx phat = -1.0 * (kx * n[0][m] + ky * n[1][m] + kz * n[2][m]) / (kx * kx + ky * ky + kz * kz + eps);
                n[0][m] = -n[0][m] - kx * phat;
                n[1][m] = -n[1][m] - ky * phat;
                n[2][m] = -n[2][m] - kz * phat;
            }
}

void orc_rhs_adj(const orc_params* p, const double* uh_, const double* vh_, const double* wh_,
                 double* au_, double* av_, double* aw_)
{
    orc_rhs_adj_sign(p, uh_, vh_, wh_, 1.0, au_, av_, aw_);
}

/* orc_rhs_adj with the linear term multiplied by lin_sign: +1 is rhsAdj.f90:356-381; -1 turns
 * L u~ into -L u~, the descent direction -DF^+(u~) of J = 1/2 ||F(u)||^2_{H^-1} */
void orc_rhs_adj_sign(const orc_params* p, const double* uh_, const double* vh_, const double* wh_,
                      double lin_sign, double* au_, double* av_, double* aw_)
{
    const cplx* u[3] = {(const cplx*)uh_, (const cplx*)vh_, (const cplx*)wh_};
    cplx* out[3] = {(cplx*)au_, (cplx*)av_, (cplx*)aw_};
    const size_t ms = nspec_elems(p), mp = nphys_elems(p);
    const double sm = 1.0 / ((double)p->nx * p->ny * p->nz);
    ktabs t = ktabs_make(p);
    cplx* tmp = (cplx*)malloc(sizeof(cplx) * ms);
    double* d = (double*)malloc(sizeof(double) * mp * 9);  /* d[3 i + j] = d_j u_i */
    double* ut = (double*)malloc(sizeof(double) * mp * 3);
    double* tr = (double*)malloc(sizeof(double) * mp);
    cplx* nl = (cplx*)malloc(sizeof(cplx) * ms * 3);
    cplx* uti = (cplx*)malloc(sizeof(cplx) * ms * 3);
    cplx* n[3] = {nl, nl + ms, nl + 2 * ms};
    /* rhsNonlinear (:64-175) */
    for (int c = 0; c < 3; ++c) grads_of(p, &t, u[c], tmp, d + 3 * c * mp);
    for (int c = 0; c < 3; ++c) orc_c2r(p, (const double*)u[c], ut + c * mp);
    for (int c = 0; c < 3; ++c) {
        for (size_t i = 0; i < mp; ++i)
            tr[i] = (ut[i] * d[(3 * c + 0) * mp + i] + ut[mp + i] * d[(3 * c + 1) * mp + i] +
                     ut[2 * mp + i] * d[(3 * c + 2) * mp + i]) * sm;
        orc_r2c(p, tr, (double*)n[c]);
    }
    project_gt(p, &t, n);
    /* rhsupptilde (:177-213), hminone = .true. */
    for (int ix = 0; ix < nxh(p); ++ix)
        for (int iy = 0; iy < p->ny; ++iy)
            for (int iz = 0; iz < p->nz; ++iz) {
                size_t m = sidx(p, ix, iy, iz);
                const cplx kx = t.kx[ix] * I, ky = t.ky[iy] * I, kz = t.kz[iz] * I;
                const cplx kk = kx * kx + ky * ky + kz * kz;
                const double factor = 1.0 / (1.0 - creal(kk));
                for (int c = 0; c < 3; ++c) uti[c * ms + m] = factor * ((p->nu * kk + p->Q) * u[c][m] + n[c][m]);
            }
    /* rhsDerivativespp (:215-259) and rhsNonlinearAdj (:261-309) */
    for (int c = 0; c < 3; ++c) grads_of(p, &t, uti + c * ms, tmp, d + 3 * c * mp);
    for (int i = 0; i < 3; ++i) {
        for (size_t q = 0; q < mp; ++q) {
            double acc = 0.0;
            for (int j = 0; j < 3; ++j)
                acc += d[(3 * j + i) * mp + q] * ut[j * mp + q] + d[(3 * i + j) * mp + q] * ut[j * mp + q];
            tr[q] = acc * sm;
        }
        orc_r2c(p, tr, (double*)n[i]);
    }
    project_gt(p, &t, n);
    /* rhsAdj (:356-381) */
    for (int ix = 0; ix < nxh(p); ++ix)
        for (int iy = 0; iy < p->ny; ++iy)
            for (int iz = 0; iz < p->nz; ++iz) {
                size_t m = sidx(p, ix, iy, iz);
                const cplx kx = t.kx[ix] * I, ky = t.ky[iy] * I, kz = t.kz[iz] * I;
                const cplx L = p->nu * (kx * kx + ky * ky + kz * kz) + p->Q;
                for (int c = 0; c < 3; ++c) out[c][m] = lin_sign * L * uti[c * ms + m] + n[c][m];
            }
    free(uti); free(nl); free(tr); free(ut); free(d); free(tmp);
    ktabs_free(&t);
}

/* dnsboxAdj.f90:76-126, nsteps times: Heun in fictitious time on rhsAdj, then state_project:
 *   G0 = rhsAdj(u) (:79); u* = u + dtau G0 (:83-101); u1 = u + (dtau G0) 0.5 (:87-113);
 *   u = u1 + (dtau rhsAdj(u*)) 0.5 (:105-113); state_project (:117) */
void orc_adj_descent(const orc_params* p, double* uh_, double* vh_, double* wh_, double dtau, int nsteps,
                     double lin_sign)
{
    const size_t ms = nspec_elems(p);
    cplx* f[3] = {(cplx*)uh_, (cplx*)vh_, (cplx*)wh_};
    cplx* g = (cplx*)malloc(sizeof(cplx) * ms * 3);
    cplx* us = (cplx*)malloc(sizeof(cplx) * ms * 3);
    for (int n = 0; n < nsteps; ++n) {
        orc_rhs_adj_sign(p, uh_, vh_, wh_, lin_sign, (double*)g, (double*)(g + ms), (double*)(g + 2 * ms));
        for (int c = 0; c < 3; ++c)
            for (size_t m = 0; m < ms; ++m) {
                us[c * ms + m] = f[c][m] + dtau * g[c * ms + m];
                f[c][m] = f[c][m] + dtau * g[c * ms + m] * 0.5;
            }
        orc_rhs_adj_sign(p, (double*)us, (double*)(us + ms), (double*)(us + 2 * ms), lin_sign, (double*)g,
                         (double*)(g + ms), (double*)(g + 2 * ms));
        for (int c = 0; c < 3; ++c)
            for (size_t m = 0; m < ms; ++m) f[c][m] = f[c][m] + dtau * g[c * ms + m] * 0.5;
        orc_project(p, uh_, vh_, wh_);
    }
    free(us);
    free(g);
}
