// SimplicialEngine<T, Predicate> of FTK2 (hguo/ftk2 include/ftk2/core/engine.hpp:79-99,
// :822-996) with its execute() / get_complex() contract, running on an MI355X
// through the C ABI of include/ftk2h.h (libftk2h.so).  Regular meshes with the
// critical-point, contour and fiber predicates are supported; anything else throws
// (there is no CPU implementation behind this header).
#pragma once
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include <ndarray/ndarray.hh>

#include "../../ftk2h.h"
#include "complex.hpp"
#include "feature.hpp"
#include "mesh.hpp"
#include "predicate.hpp"

namespace ftk2 {

namespace detail {
template <class P>
struct pred_traits {
  static constexpr bool supported = false;
};
template <int M, class T>
struct pred_traits<CriticalPointPredicate<M, T>> {
  static constexpr bool supported = true;
  static constexpr int kind = FTK2H_PRED_CRITICAL_POINT;
};
template <class T>
struct pred_traits<ContourPredicate<T>> {
  static constexpr bool supported = true;
  static constexpr int kind = FTK2H_PRED_CONTOUR;
};
template <class T>
struct pred_traits<FiberPredicate<T>> {
  static constexpr bool supported = true;
  static constexpr int kind = FTK2H_PRED_FIBER;
};

inline void check(int rc) {
  if (rc != FTK2H_OK) throw std::runtime_error(std::string("ftk2h: ") + ftk2h_last_error());
}

// Process-wide context per device.
inline ftk2h_ctx* context(int device) {
  static std::map<int, ftk2h_ctx*> ctxs;
  auto it = ctxs.find(device);
  if (it != ctxs.end()) return it->second;
  ftk2h_ctx* c = nullptr;
  check(ftk2h_ctx_create(device, &c));
  ctxs[device] = c;
  return c;
}
}  // namespace detail

template <typename T, typename PredicateType, typename IDType = uint64_t>
class SimplicialEngine {
  static_assert(std::is_same_v<T, float> || std::is_same_v<T, double>, "T must be float or double");

 public:
  static constexpr int m = PredicateType::codimension;

  SimplicialEngine(std::shared_ptr<Mesh> mesh, PredicateType pred = PredicateType(), int device = 0)
      : mesh_(std::move(mesh)), predicate_(pred), device_(device) {}
  ~SimplicialEngine() { clear_results(); }
  SimplicialEngine(const SimplicialEngine&) = delete;
  SimplicialEngine& operator=(const SimplicialEngine&) = delete;

  void clear_results() {
    if (result_) ftk2h_result_free(result_);
    result_ = nullptr;
  }

  // engine.hpp:96 — extraction over the whole mesh (the C ABI does feed() and the
  // linking of get_complex() in one pass on the device).
  void execute(const std::map<std::string, ftk::ndarray<T>>& data, const std::vector<std::string>& var_names = {}) {
    (void)var_names;  // the predicate's own variable names select the arrays
    clear_results();
    auto reg = std::dynamic_pointer_cast<RegularSimplicialMesh>(mesh_);
    tstride_ = 0;
    if (auto ext = std::dynamic_pointer_cast<ExtrudedSimplicialMesh>(mesh_)) {
      // the layers become one more grid axis; only the vertex ids differ
      auto base = std::dynamic_pointer_cast<RegularSimplicialMesh>(ext->get_base_mesh());
      bool origin = base != nullptr;
      if (base)
        for (auto o : base->get_offset()) origin = origin && o == 0;
      if (!origin)
        throw std::runtime_error("SimplicialEngine: extruded meshes need a RegularSimplicialMesh base without offset");
      auto dims = base->get_local_dims();
      auto gdims = base->get_global_dims();
      dims.push_back(ext->get_n_layers() + 1);
      gdims.push_back(ext->get_n_layers() + 1);
      reg = std::make_shared<RegularSimplicialMesh>(dims, std::vector<uint64_t>{}, gdims);
      tstride_ = ext->get_n_spatial_verts();
    }
    if (!reg)
      throw std::runtime_error("SimplicialEngine: only RegularSimplicialMesh (or extruded over one) is supported on the GPU");
    if constexpr (!detail::pred_traits<PredicateType>::supported) {
      throw std::runtime_error("SimplicialEngine: predicate not supported on the GPU");
    } else {
      const auto& ld = reg->get_local_dims();
      const auto& off = reg->get_offset();
      const auto& gd = reg->get_global_dims();
      const int d = (int)ld.size();
      if (d < 2 || d > 4) throw std::runtime_error("SimplicialEngine: mesh dimension must be 2..4");
      ftk2h_mesh_desc md{};
      md.ndims = d;
      for (int j = 0; j < 4; ++j) {
        md.local_dims[j] = j < d ? ld[j] : 1;
        md.offset[j] = j < d ? off[j] : 0;
        md.global_dims[j] = j < d ? gd[j] : 1;
      }
      md.time_id_stride = tstride_;
      ftk2h_predicate_desc pd{};
      pd.kind = detail::pred_traits<PredicateType>::kind;
      pd.codimension = m;
      pd.dtype = std::is_same_v<T, double> ? FTK2H_F64 : FTK2H_F32;
      pd.sos_q = predicate_.sos_q;
      std::vector<ftk2h_field_view> views;
      ftk2h_field_view scalar{};
      bool has_scalar = false;
      if constexpr (detail::pred_traits<PredicateType>::kind == FTK2H_PRED_CRITICAL_POINT) {
        if (predicate_.use_multicomponent && !predicate_.vector_var_name.empty()) {
          // [M, x, y, (z), t] read at GLOBAL coordinates (engine.hpp:1101-1124)
          const auto& vec = data.at(predicate_.vector_var_name);
          for (int j = 0; j < m; ++j) views.push_back(component_view(vec, j, gd, off, d));
        } else {
          for (int j = 0; j < m; ++j) views.push_back(legacy_view(data.at(predicate_.var_names[j]), reg.get()));
          // scalar interpolation reads global coordinates (predicate.hpp:81-92)
          if (!predicate_.scalar_var_name.empty()) {
            scalar = global_view(data.at(predicate_.scalar_var_name), off, d);
            has_scalar = true;
          }
        }
      } else if constexpr (detail::pred_traits<PredicateType>::kind == FTK2H_PRED_CONTOUR) {
        views.push_back(legacy_view(data.at(predicate_.var_name), reg.get()));
        pd.thresholds[0] = (double)predicate_.threshold;
      } else {
        for (int j = 0; j < 2; ++j) {
          views.push_back(legacy_view(data.at(predicate_.var_names[j]), reg.get()));
          pd.thresholds[j] = (double)predicate_.thresholds[j];
        }
      }
      detail::check(ftk2h_extract_regular(detail::context(device_), &md, &pd, views.data(), (int)views.size(),
                                          has_scalar ? &scalar : nullptr, nullptr, &result_, nullptr));
      add_attributes(data, reg.get(), md);
      if constexpr (detail::pred_traits<PredicateType>::kind == FTK2H_PRED_CRITICAL_POINT) {
        if (!predicate_.field_type.empty()) {
          const int mode = predicate_.field_type == "gradient" ? FTK2H_CLASSIFY_GRADIENT
                           : predicate_.field_type == "flow"   ? FTK2H_CLASSIFY_FLOW
                                                               : 0;
          if (!mode) throw std::invalid_argument("field_type must be 'gradient' or 'flow'");
          detail::check(ftk2h_result_classify(detail::context(device_), result_, &md, views.data(), (int)views.size(),
                                              mode, nullptr));
        }
      }
    }
  }

  // Device payload of finished nodes, lambda and cells beyond which execute_stream
  // moves them to host memory (ftk2h_stream_set_retire_bytes; no counterpart in the
  // reference, which keeps every slab's result in host memory).
  void set_stream_retire_bytes(uint64_t bytes) {
    retire_bytes_ = bytes;
    retire_set_ = true;
  }

  // engine.hpp:214-270 — timesteps read one at a time; each slab [t, t+1] is
  // extracted on the device as it arrives (ftk2h_stream_*), the last timestep and
  // the open union-find frontier stay resident.  The result equals execute() over
  // the whole spacetime mesh.
  void execute_stream(ftk::stream<>& s, const std::vector<std::string>& var_names = {}) {
    (void)var_names;  // the predicate's own variable names select the arrays
    const int n = s.total_timesteps();
    if (n < 2) return;  // as the reference: nothing extracted, earlier results kept
    clear_results();
    if constexpr (!detail::pred_traits<PredicateType>::supported) {
      throw std::runtime_error("SimplicialEngine: predicate not supported on the GPU");
    } else {
      std::vector<std::string> names;
      ftk2h_predicate_desc pd{};
      pd.kind = detail::pred_traits<PredicateType>::kind;
      pd.codimension = m;
      pd.dtype = std::is_same_v<T, double> ? FTK2H_F64 : FTK2H_F32;
      pd.sos_q = predicate_.sos_q;
      if constexpr (detail::pred_traits<PredicateType>::kind == FTK2H_PRED_CRITICAL_POINT) {
        for (int j = 0; j < m; ++j) names.push_back(predicate_.var_names[j]);
      } else if constexpr (detail::pred_traits<PredicateType>::kind == FTK2H_PRED_CONTOUR) {
        names.push_back(predicate_.var_name);
        pd.thresholds[0] = (double)predicate_.threshold;
      } else {
        for (int j = 0; j < 2; ++j) {
          names.push_back(predicate_.var_names[j]);
          pd.thresholds[j] = (double)predicate_.thresholds[j];
        }
      }
      // Attribute recording (engine.hpp:1144-1201) as the reference's slabs allow it:
      // a slab holds only the arrays the predicate reads, plus a CP scalar variable
      // (engine.hpp:218-228), so specs on other sources are skipped, and the values
      // come through get_value at (global - slab offset) coordinates, i.e. each
      // vertex's value at its own timestep.  Those sources' timesteps are kept in
      // host memory and interpolated over the whole spacetime grid after the last
      // slab, which gives the same floats as the per-slab recording.
      std::vector<std::string> slab_vars = names;
      if constexpr (detail::pred_traits<PredicateType>::kind == FTK2H_PRED_CRITICAL_POINT)
        if (!predicate_.scalar_var_name.empty()) slab_vars.push_back(predicate_.scalar_var_name);
      std::vector<const AttributeSpec*> specs;
      std::vector<size_t> spec_src;  // index into slab_vars
      std::vector<char> src_used(slab_vars.size(), 0);
      for (const auto& spec : predicate_.attributes) {
        if (spec.slot < 0 || spec.slot >= 16) continue;
        size_t k = 0;
        while (k < slab_vars.size() && slab_vars[k] != spec.source) ++k;
        if (k == slab_vars.size()) continue;  // not in the slab data: skipped (engine.hpp:1148-1149)
        specs.push_back(&spec);
        spec_src.push_back(k);
        src_used[k] = 1;
      }
      std::vector<std::vector<T>> src_steps(slab_vars.size());
      std::vector<uint64_t> spatial;
      ftk2h_stream* st = nullptr;
      struct Closer {
        ftk2h_stream*& st;
        ~Closer() {
          if (st) ftk2h_stream_close(st);
        }
      } closer{st};
      for (int t = 0; t < n; ++t) {
        auto group = s.read(t);
        std::vector<ftk2h_field_view> views;
        for (const auto& name : names) {
          const auto& a = group->template get_ref<T>(name);
          if (t == 0 && spatial.empty())
            for (size_t j = 0; j < a.nd(); ++j) spatial.push_back(a.dimf(j));
          if (a.nd() != spatial.size()) throw std::runtime_error("execute_stream: timesteps differ in shape");
          ftk2h_field_view v{};
          v.ptr = a.data();
          int64_t str = 1;
          for (size_t j = 0; j < a.nd(); ++j) {
            if (a.dimf(j) != spatial[j]) throw std::runtime_error("execute_stream: timesteps differ in shape");
            v.strides[j] = str;
            str *= (int64_t)a.dimf(j);
          }
          v.on_device = 0;
          views.push_back(v);
        }
        if (t == 0) {
          if (spatial.empty() || spatial.size() > 3) throw std::runtime_error("execute_stream: 1-3 spatial dims");
          // A slab array whose first extent is 1..16 counts as multi-component
          // (engine.hpp:1162); a component or magnitude spec then reads it at global
          // coordinates (engine.hpp:1164-1185), which lie outside a two-timestep slab.
          const bool multi = spatial[0] >= 1 && spatial[0] <= 16;
          for (const auto* spec : specs)
            if (multi && (spec->component >= 0 || spec->type == "magnitude"))
              throw std::runtime_error("execute_stream: attribute '" + spec->source +
                                       "' would read a slab as a multi-component array at global coordinates");
        }
        for (size_t k = 0; k < slab_vars.size(); ++k) {
          if (!src_used[k]) continue;
          const auto& a = group->template get_ref<T>(slab_vars[k]);
          size_t nsp = 1;
          for (auto e : spatial) nsp *= (size_t)e;
          if (a.nelem() != nsp) throw std::runtime_error("execute_stream: attribute source differs in shape");
          src_steps[k].insert(src_steps[k].end(), a.data(), a.data() + nsp);
        }
        if (t == 0) {
          detail::check(ftk2h_stream_open(detail::context(device_), (int)spatial.size(), spatial.data(), &pd,
                                          (int)names.size(), &st));
          if (retire_set_) detail::check(ftk2h_stream_set_retire_bytes(st, retire_bytes_));
        }
        detail::check(ftk2h_stream_append(st, views.data(), nullptr));
      }
      detail::check(ftk2h_stream_finish(st, &result_));
      if (!specs.empty()) {
        const int d = (int)spatial.size() + 1;
        ftk2h_mesh_desc md{};  // the whole spacetime grid
        md.ndims = d;
        for (int j = 0; j < 4; ++j) {
          md.local_dims[j] = j < d - 1 ? spatial[j] : (j == d - 1 ? (uint64_t)n : 1);
          md.global_dims[j] = md.local_dims[j];
        }
        std::vector<ftk2h_attribute_desc> ad;
        for (size_t i = 0; i < specs.size(); ++i) {
          ftk2h_attribute_desc a{};
          a.slot = specs[i]->slot;
          a.ncomp = 1;
          a.comps[0].ptr = src_steps[spec_src[i]].data();
          int64_t str = 1;
          for (int j = 0; j < d; ++j) {
            a.comps[0].strides[j] = str;
            str *= (int64_t)md.local_dims[j];
          }
          a.comps[0].on_device = 0;
          ad.push_back(a);
        }
        detail::check(ftk2h_result_add_attributes(detail::context(device_), result_, &md, ad.data(), (int)ad.size(),
                                                  nullptr));
      }
    }
  }

  // engine.hpp:822 — vertices sorted by simplex, track_id = minimum sorted index of
  // the component, connectivity sorted and unique.
  FeatureComplex get_complex() {
    FeatureComplex cx;
    if (!result_) return cx;
    uint64_t n = 0;
    int dim = 0, nconn = 0;
    detail::check(ftk2h_result_sizes(result_, &n, &dim, &nconn));
    const int nv = dim + 1;
    std::vector<uint64_t> verts(n * nv), track(n);
    std::vector<float> bary(n * nv), scal(n);
    if (n) detail::check(ftk2h_result_copy_nodes(result_, verts.data(), nullptr, bary.data(), scal.data(), track.data()));
    std::vector<float> attrs(n * 16);
    if (n) detail::check(ftk2h_result_copy_attributes(result_, attrs.data()));
    std::vector<uint32_t> types(n);
    if (n) detail::check(ftk2h_result_copy_types(result_, types.data()));
    cx.vertices.resize(n);
    for (uint64_t i = 0; i < n; ++i) {
      FeatureElement& el = cx.vertices[i];
      for (int k = 0; k < 16; ++k) el.attributes[k] = attrs[i * 16 + k];
      el.simplex.dimension = dim;
      for (int k = 0; k < nv; ++k) {
        el.simplex.vertices[k] = verts[i * nv + k];
        el.barycentric_coords[0][k] = bary[i * nv + k];
      }
      el.geometry_type = FeatureGeometryType::Point;
      el.track_id = track[i];
      el.type = types[i];  // 0 unless the predicate asked for a classification
      el.scalar = scal[i];
    }
    for (int c = 0; c < nconn; ++c) {
      int cd = 0;
      uint64_t cells = 0;
      detail::check(ftk2h_result_conn_info(result_, c, &cd, &cells));
      FeatureComplex::SimplexIndices si;
      si.dimension = cd;
      si.indices.resize(cells * (cd + 1));
      if (cells) detail::check(ftk2h_result_copy_conn(result_, c, si.indices.data()));
      cx.connectivity.push_back(std::move(si));
    }
    return cx;
  }

  // Device-side result of the last execute() (for callers that keep working on the GPU).
  const ftk2h_result* result() const { return result_; }

 private:
  ftk2h_field_view legacy_view(const ftk::ndarray<T>& arr, const RegularSimplicialMesh* reg) const {
    // get_value (engine.hpp:999-1009): 1-D arrays by vertex id, else f(local coords)
    ftk2h_field_view v{};
    v.on_device = 0;
    const int d = (int)reg->get_local_dims().size();
    if (arr.nd() == 1) {
      const auto& gd = reg->get_global_dims();
      const auto& off = reg->get_offset();
      int64_t s = 1, base = 0;
      for (int j = 0; j < d; ++j) {
        v.strides[j] = tstride_ && j == d - 1 ? (int64_t)tstride_ : s;  // extruded ids
        base += (int64_t)off[j] * s;
        s *= (int64_t)gd[j];
      }
      v.ptr = arr.data() + base;
    } else {
      int64_t s = 1;
      for (int j = 0; j < d; ++j) {
        v.strides[j] = s;
        s *= (int64_t)arr.dimf(j);
      }
      v.ptr = arr.data();
    }
    return v;
  }
  static ftk2h_field_view global_view(const ftk::ndarray<T>& arr, const std::vector<uint64_t>& off, int d) {
    ftk2h_field_view v{};
    int64_t s = 1, base = 0;
    for (int j = 0; j < d; ++j) {
      v.strides[j] = s;
      base += (int64_t)off[j] * s;
      s *= (int64_t)arr.dimf(j);
    }
    v.ptr = arr.data() + base;
    return v;
  }
  static ftk2h_field_view component_view(const ftk::ndarray<T>& vec, int comp, const std::vector<uint64_t>& gd,
                                         const std::vector<uint64_t>& off, int d) {
    (void)gd;
    ftk2h_field_view v{};
    int64_t s = (int64_t)vec.dimf(0), base = comp;
    for (int j = 0; j < d; ++j) {
      v.strides[j] = s;
      base += (int64_t)off[j] * s;
      s *= (int64_t)vec.dimf(j + 1);
    }
    v.ptr = vec.data() + base;
    return v;
  }

  // Attribute recording (engine.hpp:1144-1201): the array, component and
  // coordinates of each AttributeSpec are chosen as the reference does; the values
  // are interpolated on the device.
  void add_attributes(const std::map<std::string, ftk::ndarray<T>>& data, const RegularSimplicialMesh* reg,
                      const ftk2h_mesh_desc& md) {
    if constexpr (detail::pred_traits<PredicateType>::supported) {
      if (predicate_.attributes.empty()) return;
      const auto& gd = reg->get_global_dims();
      const auto& off = reg->get_offset();
      const int d = (int)gd.size();
      std::vector<ftk2h_attribute_desc> ad;
      for (const auto& spec : predicate_.attributes) {
        if (spec.slot < 0 || spec.slot >= 16) continue;
        auto it = data.find(spec.source);
        if (it == data.end()) continue;
        const auto& arr = it->second;
        // a first extent of 1..16 marks a multi-component array (engine.hpp:1162)
        const bool multi = arr.nd() >= 2 && arr.dimf(0) >= 1 && arr.dimf(0) <= 16;
        ftk2h_attribute_desc a{};
        a.slot = spec.slot;
        if (multi && spec.component >= 0) {
          if ((size_t)spec.component >= arr.dimf(0))
            throw std::out_of_range("attribute component " + std::to_string(spec.component) + " of " + spec.source);
          a.ncomp = 1;
          a.comps[0] = component_view(arr, spec.component, gd, off, d);  // global coordinates
        } else if (multi && spec.type == "magnitude") {
          a.ncomp = (int)arr.dimf(0);
          for (int c = 0; c < a.ncomp; ++c) a.comps[c] = component_view(arr, c, gd, off, d);
        } else {
          a.ncomp = 1;
          a.comps[0] = legacy_view(arr, reg);  // get_value
        }
        ad.push_back(a);
      }
      if (!ad.empty())
        detail::check(ftk2h_result_add_attributes(detail::context(device_), result_, &md, ad.data(), (int)ad.size(),
                                                  nullptr));
    }
  }

  std::shared_ptr<Mesh> mesh_;
  PredicateType predicate_;
  int device_ = 0;
  uint64_t retire_bytes_ = 0;
  bool retire_set_ = false;
  ftk2h_result* result_ = nullptr;
  uint64_t tstride_ = 0;  // id step of the layer axis when mesh_ is extruded
};

}  // namespace ftk2
