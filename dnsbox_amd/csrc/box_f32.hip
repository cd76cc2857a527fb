// FP32 boxes: FFTs and the nonlinear term in FP32; the integrating-factor
// exponent, the CFL controller and all reductions in FP64 (box.hpp).
#include "box_impl.hpp"

namespace dnsbox {
BoxBase* make_box_f32(const BoxConfig& c) { return make_box_t<float>(c); }
}  // namespace dnsbox
