// Explicit instantiation of the extraction pipeline: KIND_FIBER on d=4, m=2.
#include "internal.hpp"

namespace ftk2h_internal {
template int run_pipeline<KIND_FIBER, 4, 2, float>(RunArgs&);
template int run_pipeline<KIND_FIBER, 4, 2, double>(RunArgs&);
}  // namespace ftk2h_internal
