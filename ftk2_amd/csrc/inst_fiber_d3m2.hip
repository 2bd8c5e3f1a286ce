// Explicit instantiation of the extraction pipeline: KIND_FIBER on d=3, m=2.
#include "internal.hpp"

namespace ftk2h_internal {
template int run_pipeline<KIND_FIBER, 3, 2, float>(RunArgs&);
template int run_pipeline<KIND_FIBER, 3, 2, double>(RunArgs&);
}  // namespace ftk2h_internal
