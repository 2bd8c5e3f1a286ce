// Explicit instantiation of the batched pipeline: KIND_CONTOUR, 2, 1.
#include "batch.hpp"

namespace ftk2h_internal {
template int run_batch<KIND_CONTOUR, 2, 1, float>(BatchArgs&, std::vector<ftk2h_result*>&);
template int run_batch<KIND_CONTOUR, 2, 1, double>(BatchArgs&, std::vector<ftk2h_result*>&);
}  // namespace ftk2h_internal
