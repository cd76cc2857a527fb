/*
 * CPU FP64 oracle for the pseudospectral Navier-Stokes time step of dnsbox.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithm, used by tests/, by __graft_entry__.smoke() and by bench.py's
 * cpu_baseline leg as the checker.  The product (libdnsbox_hip.so) never links,
 * loads or calls it.
 *
 * Layouts follow the reference exactly (one MPI rank, p3dfft STRIDE1):
 *   physical  real    Fortran (Nx, Ny, Nz), x fastest  -> C index [z][y][x]
 *   spectral  complex Fortran (Nz, Ny, Nx/2+1), z fastest -> C index [kx][ky][kz]
 * Transforms are unnormalised, forward sign -1 (u_hat = sum_x u e^{-ik.x}).
 *
 * Every function cites the reference file:line it restates.
 */
#ifndef DNSBOX_ORACLE_H
#define DNSBOX_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    double nu;         /* kinematic viscosity                  parameters.f90:20 */
    double Q;          /* linear forcing multiplier            parameters.f90:21 */
    double Pin;        /* band-limited power input             parameters.f90:23 */
    double kcut;       /* band-forcing cut-off |k|             parameters.f90:22 */
    double kolmo_amp;  /* Kolmogorov forcing amplitude (f = amp sin(kf y) e_x)   */
    double deltak;     /* shell width                          parameters.f90:26 */
    int nx, ny, nz;    /* grid                                 parameters.f90:12-14 */
    int kolmo_k;       /* Kolmogorov forcing wavenumber (0 = off)                */
    int bandlim;       /* band-limited forcing on/off          parameters.f90:52 */
    int spherical;     /* spherical truncation added to cubic  parameters.f90:50 */
    int parity_mode;   /* 1: reference k-table (Nyquist k=0, Nyquist modes kept by the mask);
                          0: Nyquist-index modes removed by the mask as well           */
    int nthreads;      /* OpenMP threads for the transforms (0 = runtime default) */
    double alpha_x, alpha_y, alpha_z;  /* 2 pi / L per axis (0 = 1), used at their single-precision
                                          values as the reference's real(kind=4) parameters
                                          (parameters.f90:18-20)                             */
} orc_params;

/* diagnostics vector layout of orc_diagnostics */
enum { ORC_EKIN = 0, ORC_DISP, ORC_PROD, ORC_COURANT, ORC_DIVMAX, ORC_EZERO, ORC_EONE,
       ORC_ESQRT2, ORC_EBAND, ORC_NDIAG };

int    orc_nspec(const orc_params* p);
void   orc_ktable(int n, double* k);                       /* imaginary part of kx/ky/kz */
void   orc_r2c(const orc_params* p, const double* phys, double* spec);
void   orc_c2r(const orc_params* p, const double* spec, double* phys);
void   orc_fft1d(double* data, int n, int sign);           /* complex, in place, unnormalised */
int    orc_masked(const orc_params* p, int ix, int iy, int iz);
double orc_eband(const orc_params* p, const double* uh, const double* vh, const double* wh);
void   orc_rhs_nonlinear(const orc_params* p, const double* uh, const double* vh, const double* wh,
                         double* nu_, double* nv_, double* nw_);
void   orc_int_fact(const orc_params* p, double dt, double* intfact);
void   orc_heun(const orc_params* p, double* uh, double* vh, double* wh, int nsteps, double dt);
void   orc_ab2(const orc_params* p, double* uh, double* vh, double* wh, int nsteps, double dt);
void   orc_ab2_var(const orc_params* p, double* uh, double* vh, double* wh, int nsteps, const double* dts);
void   orc_ab2_adaptive(const orc_params* p, double* uh, double* vh, double* wh, int nsteps,
                        double* dt_inout, double cmin, double cmax, double dtmax, double* t_out);
void   orc_project(const orc_params* p, double* uh, double* vh, double* wh);
void   orc_dealias(const orc_params* p, double* uh, double* vh, double* wh);
void   orc_diagnostics(const orc_params* p, const double* uh, const double* vh, const double* wh,
                       double dt, double* diag, double* espec);
double orc_courant(const orc_params* p, const double* uh, const double* vh, const double* wh,
                   double dt);
double orc_cfl_dt(double courant, double dt, double cmin, double cmax, double dtmax);
void   orc_heun_adaptive(const orc_params* p, double* uh, double* vh, double* wh, int nsteps,
                         double* dt_inout, double cmin, double cmax, double dtmax, double* t_out);
void   orc_init_solenoidal(const orc_params* p, uint64_t seed, int box, double k0, double energy,
                           double* uh, double* vh, double* wh);
void   orc_init_shapiro(const orc_params* p, double t, double* u, double* v, double* w);
double orc_shapiro_error(const orc_params* p, double t, const double* u, const double* v,
                         const double* w);
uint64_t orc_hash(uint64_t seed, uint64_t box, uint64_t idx);

/* adjoint pieces */
void   orc_rhs_adj(const orc_params* p, const double* uh, const double* vh, const double* wh,
                   double* au, double* av, double* aw);
void   orc_rhs_adj_sign(const orc_params* p, const double* uh, const double* vh, const double* wh,
                        double lin_sign, double* au, double* av, double* aw);
void   orc_adj_descent(const orc_params* p, double* uh, double* vh, double* wh, double dtau, int nsteps,
                       double lin_sign);
void   orc_rhs_vjp(const orc_params* p, const double* uh, const double* vh, const double* wh,
                   const double* lu, const double* lv, const double* lw,
                   double* gu, double* gv, double* gw);
void   orc_heun_vjp(const orc_params* p, const double* uh, const double* vh, const double* wh,
                    double dt, double* lu, double* lv, double* lw);

#ifdef __cplusplus
}
#endif
#endif
