// Explicit instantiation of the extraction pipeline: KIND_CP on d=2, m=1.
#include "internal.hpp"

namespace ftk2h_internal {
template int run_pipeline<KIND_CP, 2, 1, float>(RunArgs&);
template int run_pipeline<KIND_CP, 2, 1, double>(RunArgs&);
}  // namespace ftk2h_internal
