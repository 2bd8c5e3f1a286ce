// Predicate structs of FTK2 (hguo/ftk2 include/ftk2/core/predicate.hpp:22-130,
// :135-183, :340-389): the same fields, so callers compile unchanged.  The
// predicates are evaluated on the GPU (include/ftk2h.h); these structs carry the
// configuration only.
#pragma once
#include <string>
#include <vector>

namespace ftk2 {

template <int M, typename T = double>
struct Predicate {
  static constexpr int codimension = M;
  double sos_q = 1000000.0;
};

struct AttributeSpec {
  std::string name;
  std::string source;
  std::string type = "scalar";
  int component = -1;
  int slot = -1;
};

template <int M, typename T = double>
struct CriticalPointPredicate : public Predicate<M, T> {
  std::string vector_var_name;
  std::string var_names[M];
  std::string scalar_var_name;
  bool use_multicomponent = true;
  std::vector<AttributeSpec> attributes;
  // Not a field of the reference's predicate: "gradient" (Hessian classification:
  // min / max / saddles) or "flow" (Jacobian classification: source / sink / saddle /
  // spiral), the values of the reference's tracking configuration
  // (high_level/tracking_config.hpp:84-88), sets FeatureElement::type
  // (ftk2h_cp_type in ftk2h.h); empty keeps type = 0 as the reference's engine does.
  std::string field_type;
};

template <typename T = double>
struct ContourPredicate : public Predicate<1, T> {
  std::string var_name;
  T threshold = (T)0.0;
  std::vector<AttributeSpec> attributes;
};

template <typename T = double>
struct FiberPredicate : public Predicate<2, T> {
  std::string var_names[2];
  T thresholds[2] = {(T)0.0, (T)0.0};
  std::vector<AttributeSpec> attributes;
};

}  // namespace ftk2
