// FP32 boxes: FFTs and the nonlinear term in FP32; the integrating-factor
// exponent, the CFL controller and all reductions in FP64 (box.hpp).  The state
// is FP32, or FP64 when state_precision = 64 (the update then runs in FP64 and
// only the transforms see FP32).
#include "box_impl.hpp"

namespace dnsbox {
BoxBase* make_box_f32(const BoxConfig& c)
{
    if (c.state_precision == 64) return make_box_t<float, double>(c);
    return make_box_t<float, float>(c);
}
}  // namespace dnsbox
