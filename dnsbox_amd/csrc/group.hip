// Box groups (group.hpp): FP32 and FP64 instantiations.
#include "group.hpp"

namespace dnsbox {
GroupBase* make_group(std::vector<std::unique_ptr<BoxBase>>&& boxes, int precision)
{
    if (precision == 64) return new Group<double>(std::move(boxes));
    return new Group<float>(std::move(boxes));
}
}  // namespace dnsbox
