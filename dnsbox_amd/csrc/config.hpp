// dnsbox_params (the C-ABI's plain parameters) -> the device-side box configuration.
#pragma once

#include "../../include/dnsbox.h"
#include "box_impl.hpp"

namespace dnsbox {

inline BoxConfig make_config(const dnsbox_params* p)
{
    BoxConfig bc;
    bc.n = p->n;
    // the grid (parameters.f90:12-14): nx, ny, nz, each 0 for n
    bc.nx = p->nx > 0 ? p->nx : p->n;
    bc.ny = p->ny > 0 ? p->ny : p->n;
    bc.nz = p->nz > 0 ? p->nz : p->n;
    const int nd[3] = {bc.nx, bc.ny, bc.nz};
    bc.precision = p->precision;
    bc.state_precision = p->state_precision == 64 ? 64 : p->precision;
    bc.deltak = p->deltak > 0 ? p->deltak : 1.0;
    bc.dt = p->dt;
    BoxPhys& ph = bc.ph;
    ph.nu = p->nu;
    ph.Q = p->Q;
    ph.Pin = p->Pin;
    ph.kcut = p->kcut;
    ph.kolmo_amp = p->kolmo_amp;
    // alpha_d = 2 pi / L_d, as the reference's real(kind=4) parameters hold it (parameters.f90:18-20)
    const double al[3] = {p->alpha_x, p->alpha_y, p->alpha_z};
    for (int d = 0; d < 3; ++d) ph.a[d] = al[d] > 0.0 ? double(float(al[d])) : 1.0;
    // rhs.f90:74-88 / state.f90:184-197 thresholds, evaluated as the reference writes them
    for (int d = 0; d < 3; ++d) ph.thr[d] = (double(nd[d]) / 2.0) * (2.0 * ph.a[d] / 3.0);
    const double ts = (double(bc.nx) / 2.0) * 0.5 * ph.a[0];
    ph.sph_thr2 = ts * ts;
    ph.sph_dthr2 = ph.thr[0] * ph.thr[0];
    ph.cmin = p->cfl_min;
    ph.cmax = p->cfl_max;
    ph.dtmax = p->dt_max;
    ph.kolmo_k = p->kolmo_amp != 0.0 ? p->kolmo_k : 0;
    ph.bandlim = p->bandlim;
    ph.spherical = p->spherical;
    ph.parity = p->parity_mode;
    ph.adaptive = p->adaptive_dt;
    for (int d = 0; d < 3; ++d) ph.nyq[d] = nd[d] / 2;
    return bc;
}

}  // namespace dnsbox
