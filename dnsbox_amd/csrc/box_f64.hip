// FP64 boxes: the plain method, compared against the CPU oracle.
#include "box_impl.hpp"

namespace dnsbox {
BoxBase* make_box_f64(const BoxConfig& c) { return make_box_t<double, double>(c); }
}  // namespace dnsbox
