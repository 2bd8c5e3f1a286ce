// Explicit instantiation of the batched pipeline: KIND_CP, 3, 2.
#include "batch.hpp"

namespace ftk2h_internal {
template int run_batch<KIND_CP, 3, 2, float>(BatchArgs&, std::vector<ftk2h_result*>&);
template int run_batch<KIND_CP, 3, 2, double>(BatchArgs&, std::vector<ftk2h_result*>&);
}  // namespace ftk2h_internal
