// Boxes on the non-cubic grids of grids.hpp: FFT type double, state type double.
#include "box_impl.hpp"
#include "grids.hpp"

namespace dnsbox {
template <>
BoxBase* make_box_grid<double, double>(const BoxConfig& c)
{
#define DNSBOX_GRID_CASE(a, b, c_) \
    if (c.nx == a && c.ny == b && c.nz == c_) return new BoxT<Dims<a, b, c_>, double, double>(c);
    DNSBOX_GRIDS(DNSBOX_GRID_CASE)
#undef DNSBOX_GRID_CASE
    return nullptr;
}
}  // namespace dnsbox
