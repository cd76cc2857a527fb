// dnsbox_params (the C-ABI's plain parameters) -> the device-side box configuration.
#pragma once

#include "../../include/dnsbox.h"
#include "box_impl.hpp"

namespace dnsbox {

inline BoxConfig make_config(const dnsbox_params* p)
{
    BoxConfig bc;
    bc.n = p->n;
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
    ph.mask_thr = (double(p->n) / 2.0) * (2.0 * 1.0 / 3.0);
    const double ts = (double(p->n) / 2.0) * 0.5 * 1.0;
    ph.sph_thr2 = ts * ts;
    ph.cmin = p->cfl_min;
    ph.cmax = p->cfl_max;
    ph.dtmax = p->dt_max;
    ph.kolmo_k = p->kolmo_amp != 0.0 ? p->kolmo_k : 0;
    ph.bandlim = p->bandlim;
    ph.spherical = p->spherical;
    ph.parity = p->parity_mode;
    ph.adaptive = p->adaptive_dt;
    return bc;
}

}  // namespace dnsbox
