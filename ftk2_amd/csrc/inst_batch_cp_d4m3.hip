// Explicit instantiation of the batched pipeline: KIND_CP, 4, 3.
#include "batch.hpp"

namespace ftk2h_internal {
template int run_batch<KIND_CP, 4, 3, float>(BatchArgs&, std::vector<ftk2h_result*>&);
template int run_batch<KIND_CP, 4, 3, double>(BatchArgs&, std::vector<ftk2h_result*>&);
}  // namespace ftk2h_internal
