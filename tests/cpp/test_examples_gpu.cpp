// End-to-end cases of the reference's tests/test_examples.cpp, compiled against
// the drop-in headers (include/ftk2/core/*.hpp) and run on the GPU through
// libftk2h.so.  Exit code = number of failed checks.
#include <cmath>
#include <cstdio>
#include <map>
#include <vector>

#include <ftk2/core/engine.hpp>

using namespace ftk2;

static int failures = 0;
#define CHECK_EQ(a, b)                                                                      \
  do {                                                                                      \
    auto va = (a);                                                                          \
    auto vb = (b);                                                                          \
    if (!(va == vb)) {                                                                      \
      std::printf("FAILED %s == %s (%lld vs %lld) at line %d\n", #a, #b, (long long)va,   \
                  (long long)vb, __LINE__);                                                 \
      failures++;                                                                           \
    }                                                                                       \
  } while (0)

// vertices sorted by simplex, cells sorted/unique, track ids = component minima
static void check_complex(const FeatureComplex& cx) {
  for (size_t i = 1; i < cx.vertices.size(); ++i) CHECK_EQ(cx.vertices[i - 1].simplex < cx.vertices[i].simplex, true);
  for (size_t i = 0; i < cx.vertices.size(); ++i) CHECK_EQ(cx.vertices[i].track_id <= i, true);
  for (const auto& c : cx.connectivity) {
    const int n = c.dimension + 1;
    for (size_t i = 0; i + n <= c.indices.size(); i += n) {
      for (int q = 1; q < n; ++q) {
        CHECK_EQ(c.indices[i + q - 1] < c.indices[i + q], true);
        CHECK_EQ(cx.vertices[c.indices[i]].track_id, cx.vertices[c.indices[i + q]].track_id);
      }
    }
  }
}

static void levelset_3d() {  // test_examples.cpp:229-257
  const int W = 32, H = 32, Dd = 32, Tn = 10;
  ftk::ndarray<float> s({(size_t)W, (size_t)H, (size_t)Dd, (size_t)Tn});
  for (int t = 0; t < Tn; ++t) {
    const float cx = 16.0f + t * 0.2f;
    for (int z = 0; z < Dd; ++z)
      for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
          const float d = std::sqrt(std::pow(x - cx, 2) + std::pow(y - 16.0f, 2) + std::pow(z - 16.0f, 2));
          s.f(x, y, z, t) = d - 8.0f;
        }
  }
  std::map<std::string, ftk::ndarray<float>> data = {{"S", s}};
  auto mesh = std::make_shared<RegularSimplicialMesh>(std::vector<uint64_t>{32, 32, 32, 10});
  ContourPredicate<float> pred;
  pred.var_name = "S";
  SimplicialEngine<float, ContourPredicate<float>> engine(mesh, pred);
  engine.execute(data, {"S"});
  auto cx = engine.get_complex();
  CHECK_EQ(cx.vertices.size(), (size_t)66418);
  check_complex(cx);
  std::printf("levelset_3d: %zu vertices\n", cx.vertices.size());
}

static void fiber_3d() {  // test_examples.cpp:188-227
  const int W = 32, H = 32, Dd = 32, Tn = 10;
  ftk::ndarray<double> s1({(size_t)W, (size_t)H, (size_t)Dd, (size_t)Tn}), s2({(size_t)W, (size_t)H, (size_t)Dd, (size_t)Tn});
  for (int t = 0; t < Tn; ++t)
    for (int z = 0; z < Dd; ++z)
      for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
          const double a = std::sqrt(std::pow(x - 16.0, 2) + std::pow(y - 16.0, 2) + std::pow(z - 16.0, 2));
          const double b = std::sqrt(std::pow(x - 22.0, 2) + std::pow(y - (16.0 + t * 0.5), 2) + std::pow(z - 16.0, 2));
          s1.f(x, y, z, t) = a - 10.0;
          s2.f(x, y, z, t) = b - 8.0;
        }
  std::map<std::string, ftk::ndarray<double>> data = {{"S1", s1}, {"S2", s2}};
  auto mesh = std::make_shared<RegularSimplicialMesh>(std::vector<uint64_t>{32, 32, 32, 10});
  FiberPredicate<double> pred;
  pred.var_names[0] = "S1";
  pred.var_names[1] = "S2";
  SimplicialEngine<double, FiberPredicate<double>> engine(mesh, pred);
  engine.execute(data, {"S1", "S2"});
  auto cx = engine.get_complex();
  CHECK_EQ(cx.vertices.size(), (size_t)6078);
  check_complex(cx);
  std::printf("fiber_3d: %zu vertices\n", cx.vertices.size());
}

static void critical_point_3d() {  // test_examples.cpp:153-186 (see tests/test_oracle_engine.py)
  const int W = 24, H = 24, Dd = 24, Tn = 8;
  ftk::ndarray<float> u({(size_t)W, (size_t)H, (size_t)Dd, (size_t)Tn}), v(u), w(u);
  for (int t = 0; t < Tn; ++t) {
    const float c = 16.0f + t * 0.1f;
    for (int z = 0; z < Dd; ++z)
      for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
          u.f(x, y, z, t) = (float)x - c;
          v.f(x, y, z, t) = (float)y - 16.0f;
          w.f(x, y, z, t) = (float)z - 16.0f;
        }
  }
  std::map<std::string, ftk::ndarray<float>> data = {{"U", u}, {"V", v}, {"W", w}};
  auto mesh = std::make_shared<RegularSimplicialMesh>(std::vector<uint64_t>{24, 24, 24, 8});
  CriticalPointPredicate<3, float> pred;
  pred.var_names[0] = "U";
  pred.var_names[1] = "V";
  pred.var_names[2] = "W";
  SimplicialEngine<float, CriticalPointPredicate<3, float>> engine(mesh, pred);
  engine.execute(data, {"U", "V", "W"});
  auto cx = engine.get_complex();
  CHECK_EQ(cx.vertices.size(), (size_t)29);
  check_complex(cx);
  std::printf("critical_point_3d: %zu vertices\n", cx.vertices.size());
}

// test_examples.cpp:39-105: CP2D on an offset mesh with a scalar variable.  The
// reference reads its field from hguo/ndarray's synthetic "woven" stream (delta =
// 1/(DT-1)); the woven function restated here is FTK's published one,
// cos(x cos t - y sin t) sin(x sin t + y cos t) on [-7.5, 7.5]^2.
static void critical_point_2d() {
  const int DW = 16, DH = 16, DT = 5;
  ftk::ndarray<double> scalar({(size_t)DW, (size_t)DH, (size_t)DT});
  for (int t = 0; t < DT; ++t)
    for (int y = 0; y < DH; ++y)
      for (int x = 0; x < DW; ++x) {
        const double tt = t * (1.0 / (DT - 1));
        const double X = ((double)x / (DW - 1) - 0.5) * 15.0, Y = ((double)y / (DH - 1) - 0.5) * 15.0;
        scalar.f(x, y, t) = std::cos(X * std::cos(tt) - Y * std::sin(tt)) * std::sin(X * std::sin(tt) + Y * std::cos(tt));
      }
  ftk::ndarray<double> u({(size_t)DW, (size_t)DH, (size_t)DT}), v({(size_t)DW, (size_t)DH, (size_t)DT});
  u.fill(0.0);
  v.fill(0.0);
  for (int t = 0; t < DT; ++t)
    for (int y = 1; y < DH - 1; ++y)
      for (int x = 1; x < DW - 1; ++x) {
        u.f(x, y, t) = (scalar.f(x + 1, y, t) - scalar.f(x - 1, y, t)) / 2.0;
        v.f(x, y, t) = (scalar.f(x, y + 1, t) - scalar.f(x, y - 1, t)) / 2.0;
      }
  std::map<std::string, ftk::ndarray<double>> data = {{"U", u}, {"V", v}, {"Woven", scalar}};
  auto mesh = std::make_shared<RegularSimplicialMesh>(std::vector<uint64_t>{(uint64_t)DW - 2, (uint64_t)DH - 2, (uint64_t)DT},
                                                      std::vector<uint64_t>{1, 1, 0},
                                                      std::vector<uint64_t>{(uint64_t)DW, (uint64_t)DH, (uint64_t)DT});
  CriticalPointPredicate<2, double> cp_pred;
  cp_pred.var_names[0] = "U";
  cp_pred.var_names[1] = "V";
  cp_pred.scalar_var_name = "Woven";
  SimplicialEngine<double, CriticalPointPredicate<2, double>> engine(mesh, cp_pred);
  engine.execute(data, {"U", "V", "Woven"});
  auto cx = engine.get_complex();
  CHECK_EQ(cx.vertices.size(), (size_t)788);
  check_complex(cx);
  std::printf("critical_point_2d: %zu vertices\n", cx.vertices.size());
}

// test_examples.cpp:259-297: streamed execution equals one-shot execution, on the
// woven field of critical_point_2d() (t = k / (Tn - 1)) at contour level 0.5.
static void streaming() {
  const int W = 16, H = 16, Tn = 5;
  auto value = [](int x, int y, int t) {
    const double tt = t * (1.0 / (Tn - 1));
    const double X = ((double)x / (W - 1) - 0.5) * 15.0, Y = ((double)y / (H - 1) - 0.5) * 15.0;
    return std::cos(X * std::cos(tt) - Y * std::sin(tt)) * std::sin(X * std::sin(tt) + Y * std::cos(tt));
  };
  ftk::ndarray<double> scalar({(size_t)W, (size_t)H, (size_t)Tn});
  for (int t = 0; t < Tn; ++t)
    for (int y = 0; y < H; ++y)
      for (int x = 0; x < W; ++x) scalar.f(x, y, t) = value(x, y, t);
  auto mesh = std::make_shared<RegularSimplicialMesh>(std::vector<uint64_t>{(uint64_t)W, (uint64_t)H, (uint64_t)Tn});
  ContourPredicate<double> pred;
  pred.var_name = "scalar";
  pred.threshold = 0.5;
  SimplicialEngine<double, ContourPredicate<double>> engine(mesh, pred);
  engine.execute({{"scalar", scalar}});
  auto one = engine.get_complex();

  ftk::stream<> stream(Tn, [&](int t) {
    auto g = std::make_shared<ftk::ndarray_group>();
    ftk::ndarray<double> a({(size_t)W, (size_t)H});
    for (int y = 0; y < H; ++y)
      for (int x = 0; x < W; ++x) a.f(x, y) = value(x, y, t);
    g->set("scalar", a);
    return g;
  });
  engine.execute_stream(stream);
  auto streamed = engine.get_complex();
  CHECK_EQ(streamed.vertices.size(), one.vertices.size());
  CHECK_EQ(streamed.vertices.size() > 0, true);
  for (size_t i = 0; i < one.vertices.size() && i < streamed.vertices.size(); ++i) {
    CHECK_EQ(streamed.vertices[i].simplex < one.vertices[i].simplex, false);
    CHECK_EQ(one.vertices[i].simplex < streamed.vertices[i].simplex, false);
    CHECK_EQ(streamed.vertices[i].track_id, one.vertices[i].track_id);
  }
  CHECK_EQ(streamed.connectivity.size(), one.connectivity.size());
  for (size_t c = 0; c < one.connectivity.size() && c < streamed.connectivity.size(); ++c)
    CHECK_EQ(streamed.connectivity[c].indices == one.connectivity[c].indices, true);
  check_complex(streamed);
  std::printf("streaming: %zu vertices streamed, %zu one-shot\n", streamed.vertices.size(), one.vertices.size());

  // the same stream with finished state moved to host memory after every slab
  engine.set_stream_retire_bytes(0);
  engine.execute_stream(stream);
  auto retired = engine.get_complex();
  CHECK_EQ(retired.vertices.size(), one.vertices.size());
  for (size_t i = 0; i < one.vertices.size() && i < retired.vertices.size(); ++i) {
    CHECK_EQ(retired.vertices[i].simplex < one.vertices[i].simplex, false);
    CHECK_EQ(one.vertices[i].simplex < retired.vertices[i].simplex, false);
    CHECK_EQ(retired.vertices[i].track_id, one.vertices[i].track_id);
  }
  CHECK_EQ(retired.connectivity.size(), one.connectivity.size());
  for (size_t c = 0; c < one.connectivity.size() && c < retired.connectivity.size(); ++c)
    CHECK_EQ(retired.connectivity[c].indices == one.connectivity[c].indices, true);
  std::printf("streaming_retired: %zu vertices\n", retired.vertices.size());

  // A first extent of 16 makes the reference read each slab as a multi-component
  // array at global coordinates (engine.hpp:1162-1185), outside the slab: refused.
  AttributeSpec comp0;
  comp0.source = "scalar";
  comp0.component = 0;
  comp0.slot = 0;
  pred.attributes = {comp0};
  SimplicialEngine<double, ContourPredicate<double>> eng3(mesh, pred);
  bool threw = false;
  try {
    eng3.execute_stream(stream);
  } catch (const std::runtime_error&) {
    threw = true;
  }
  CHECK_EQ(threw, true);
}

// Attribute recording (engine.hpp:1144-1201) through the engine's AttributeSpec
// rules: a scalar source, one component and the magnitude of a multi-component
// array (read at global coordinates), and a coordinate field.  At critical points
// of (U, V, W) the interpolated components and magnitude vanish; the coordinate
// field interpolates to the feature position.
static void attributes_3d() {
  const int W = 20, H = 20, Dd = 20, Tn = 4;
  ftk::ndarray<double> u({(size_t)W, (size_t)H, (size_t)Dd, (size_t)Tn}), v(u), w(u), xs(u);
  ftk::ndarray<double> vec({3, (size_t)W, (size_t)H, (size_t)Dd, (size_t)Tn});
  for (int t = 0; t < Tn; ++t)
    for (int z = 0; z < Dd; ++z)
      for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
          const double a = x - 9.3 - 0.2 * t, b = y - 10.6 + 0.1 * t, c = z - 9.8;
          u.f(x, y, z, t) = a + 0.1 * b;
          v.f(x, y, z, t) = b - 0.05 * c;
          w.f(x, y, z, t) = c + 0.02 * a;
          xs.f(x, y, z, t) = x;
          vec.f(0, x, y, z, t) = u.f(x, y, z, t);
          vec.f(1, x, y, z, t) = v.f(x, y, z, t);
          vec.f(2, x, y, z, t) = w.f(x, y, z, t);
        }
  std::map<std::string, ftk::ndarray<double>> data = {{"U", u}, {"V", v}, {"W", w}, {"X", xs}, {"VEC", vec}};
  auto mesh = std::make_shared<RegularSimplicialMesh>(std::vector<uint64_t>{(uint64_t)W, (uint64_t)H, (uint64_t)Dd, (uint64_t)Tn});
  CriticalPointPredicate<3, double> pred;
  pred.var_names[0] = "U";
  pred.var_names[1] = "V";
  pred.var_names[2] = "W";
  pred.use_multicomponent = false;
  AttributeSpec s0; s0.source = "W"; s0.slot = 0;
  AttributeSpec s1; s1.source = "VEC"; s1.type = "magnitude"; s1.slot = 1;
  AttributeSpec s2; s2.source = "VEC"; s2.component = 0; s2.slot = 5;
  AttributeSpec s3; s3.source = "X"; s3.slot = 9;
  AttributeSpec s4; s4.source = "missing"; s4.slot = 2;  // skipped, as in the reference
  pred.attributes = {s0, s1, s2, s3, s4};
  SimplicialEngine<double, CriticalPointPredicate<3, double>> engine(mesh, pred);
  engine.execute(data);
  auto cx = engine.get_complex();
  CHECK_EQ(cx.vertices.size() > 0, true);
  for (const auto& el : cx.vertices) {
    CHECK_EQ(std::fabs(el.attributes[0]) < 1e-5f, true);
    double mag = 0;  // sum_i b_i |X_i| (the norm is taken per vertex, then interpolated)
    for (int k = 0; k < 4; ++k) {
      const uint64_t id = el.simplex.vertices[k];
      const int x = (int)(id % W), y = (int)(id / W % H), z = (int)(id / (W * H) % Dd), t = (int)(id / (W * H * Dd));
      const double a = u.f(x, y, z, t), b = v.f(x, y, z, t), c = w.f(x, y, z, t);
      mag += (double)el.barycentric_coords[0][k] * std::sqrt(a * a + b * b + c * c);
    }
    CHECK_EQ(std::fabs(el.attributes[1] - mag) <= 1e-6 * (1 + mag), true);
    CHECK_EQ(std::fabs(el.attributes[5]) < 1e-5f, true);
    CHECK_EQ(el.attributes[2] == 0.0f, true);
    double px = 0;
    for (int k = 0; k < 4; ++k) px += el.barycentric_coords[0][k] * (double)(el.simplex.vertices[k] % W);
    CHECK_EQ(std::fabs(el.attributes[9] - px) < 1e-4, true);
  }
  std::printf("attributes_3d: %zu vertices\n", cx.vertices.size());

  // Streamed (engine.hpp:214-270): the slabs hold only the predicate's arrays, so
  // the specs on U and W are recorded (the same floats as the one-shot pass) and
  // those on X, VEC and a missing source stay 0.
  AttributeSpec s5; s5.source = "U"; s5.slot = 3;
  AttributeSpec s6; s6.source = "W"; s6.slot = 9;  // overwrites the X spec's slot
  pred.attributes = {s0, s1, s2, s3, s4, s5, s6};
  SimplicialEngine<double, CriticalPointPredicate<3, double>> eng2(mesh, pred);
  eng2.execute(data);
  auto one = eng2.get_complex();
  ftk::stream<> stream(Tn, [&](int t) {
    auto g = std::make_shared<ftk::ndarray_group>();
    const char* names[4] = {"U", "V", "W", "X"};
    const ftk::ndarray<double>* src[4] = {&u, &v, &w, &xs};
    for (int k = 0; k < 4; ++k) {
      ftk::ndarray<double> a({(size_t)W, (size_t)H, (size_t)Dd});
      for (int z = 0; z < Dd; ++z)
        for (int y = 0; y < H; ++y)
          for (int x = 0; x < W; ++x) a.f(x, y, z) = src[k]->f(x, y, z, t);
      g->set(names[k], a);
    }
    return g;
  });
  eng2.set_stream_retire_bytes(0);  // finished nodes move to host memory after every slab
  eng2.execute_stream(stream);
  auto streamed = eng2.get_complex();
  CHECK_EQ(streamed.vertices.size(), one.vertices.size());
  for (size_t i = 0; i < one.vertices.size() && i < streamed.vertices.size(); ++i) {
    const auto& a = streamed.vertices[i];
    const auto& b = one.vertices[i];
    CHECK_EQ(a.simplex < b.simplex || b.simplex < a.simplex, false);
    for (int k = 0; k < 16; ++k) {
      const bool slab_source = k == 0 || k == 3 || k == 9;
      CHECK_EQ(a.attributes[k] == (slab_source ? b.attributes[k] : 0.0f), true);
    }
  }
  std::printf("attributes_stream: %zu vertices\n", streamed.vertices.size());
}

// An ExtrudedSimplicialMesh over a regular 2-D base (what the reference's
// FeatureTracker builds, feature_tracker.cpp:454) gives the regular 3-D mesh's
// features with extruded vertex ids (id = spatial + t * n_spatial_verts_).
static void extruded_2d() {
  const int W = 24, H = 20, Tn = 6;
  ftk::ndarray<double> u({(size_t)W, (size_t)H, (size_t)Tn}), v({(size_t)W, (size_t)H, (size_t)Tn});
  for (int t = 0; t < Tn; ++t)
    for (int y = 0; y < H; ++y)
      for (int x = 0; x < W; ++x) {
        u.f(x, y, t) = std::sin(0.37 * x + 0.2 * t) + 0.3 * std::cos(0.21 * y);
        v.f(x, y, t) = std::cos(0.29 * y - 0.15 * t) + 0.3 * std::sin(0.17 * x);
      }
  CriticalPointPredicate<2, double> pred;
  pred.var_names[0] = "u";
  pred.var_names[1] = "v";
  auto base = std::make_shared<RegularSimplicialMesh>(std::vector<uint64_t>{(uint64_t)W, (uint64_t)H});
  auto ext = std::make_shared<ExtrudedSimplicialMesh>(base, Tn - 1);
  CHECK_EQ(ext->get_n_spatial_verts(), (uint64_t)(W + W * H));  // grid_index_to_id({W, H})
  SimplicialEngine<double, CriticalPointPredicate<2, double>> e1(ext, pred);
  e1.execute({{"u", u}, {"v", v}});
  auto cx = e1.get_complex();
  auto reg = std::make_shared<RegularSimplicialMesh>(std::vector<uint64_t>{(uint64_t)W, (uint64_t)H, (uint64_t)Tn});
  SimplicialEngine<double, CriticalPointPredicate<2, double>> e2(reg, pred);
  e2.execute({{"u", u}, {"v", v}});
  auto rx = e2.get_complex();
  CHECK_EQ(cx.vertices.size(), rx.vertices.size());
  CHECK_EQ(cx.vertices.size() > 0, true);
  const uint64_t N = ext->get_n_spatial_verts(), S = (uint64_t)W * H;
  for (size_t i = 0; i < cx.vertices.size() && i < rx.vertices.size(); ++i) {
    for (int k = 0; k <= 2; ++k) {
      const uint64_t id = cx.vertices[i].simplex.vertices[k];
      CHECK_EQ((id % N) + (id / N) * S, rx.vertices[i].simplex.vertices[k]);
    }
    CHECK_EQ(cx.vertices[i].track_id, rx.vertices[i].track_id);
    CHECK_EQ(cx.vertices[i].barycentric_coords[0][0], rx.vertices[i].barycentric_coords[0][0]);
  }
  CHECK_EQ(cx.connectivity.size(), rx.connectivity.size());
  for (size_t c = 0; c < cx.connectivity.size() && c < rx.connectivity.size(); ++c)
    CHECK_EQ(cx.connectivity[c].indices == rx.connectivity[c].indices, true);
  check_complex(cx);
  std::printf("extruded_2d: %zu vertices\n", cx.vertices.size());
}

// The reference's default critical-point input: one multi-component array [M, x, y, t]
// named by vector_var_name (use_multicomponent = true, engine.hpp:1096-1124), equal to
// the legacy separate arrays on the same values; and the opt-in classification
// (field_type, not a reference predicate field) of a moving quadratic potential's
// gradient: every node a minimum (type 1), types 0 without it.
static void multicomponent_and_classification() {
  const int W = 26, H = 22, Tn = 5;
  ftk::ndarray<double> u({(size_t)W, (size_t)H, (size_t)Tn}), v({(size_t)W, (size_t)H, (size_t)Tn});
  ftk::ndarray<double> uv({(size_t)2, (size_t)W, (size_t)H, (size_t)Tn});
  for (int t = 0; t < Tn; ++t)
    for (int y = 0; y < H; ++y)
      for (int x = 0; x < W; ++x) {
        const double a = std::sin(0.37 * x + 0.11 * t) * std::cos(0.29 * y), b = std::cos(0.41 * x) * std::sin(0.33 * y - 0.07 * t);
        u.f(x, y, t) = a;
        v.f(x, y, t) = b;
        uv.f(0, x, y, t) = a;
        uv.f(1, x, y, t) = b;
      }
  std::map<std::string, ftk::ndarray<double>> data = {{"U", u}, {"V", v}, {"UV", uv}};
  auto mesh = std::make_shared<RegularSimplicialMesh>(std::vector<uint64_t>{(uint64_t)W, (uint64_t)H, (uint64_t)Tn});
  CriticalPointPredicate<2, double> legacy;
  legacy.var_names[0] = "U";
  legacy.var_names[1] = "V";
  CriticalPointPredicate<2, double> multi;
  multi.vector_var_name = "UV";
  SimplicialEngine<double, CriticalPointPredicate<2, double>> e1(mesh, legacy), e2(mesh, multi);
  e1.execute(data);
  e2.execute(data);
  auto c1 = e1.get_complex(), c2 = e2.get_complex();
  CHECK_EQ(c1.vertices.size() > 10, true);
  CHECK_EQ(c1.vertices.size(), c2.vertices.size());
  for (size_t i = 0; i < c1.vertices.size() && i < c2.vertices.size(); ++i) {
    CHECK_EQ(c1.vertices[i].simplex == c2.vertices[i].simplex, true);
    CHECK_EQ(c1.vertices[i].track_id, c2.vertices[i].track_id);
    CHECK_EQ(c1.vertices[i].barycentric_coords[0][0] == c2.vertices[i].barycentric_coords[0][0], true);
    CHECK_EQ(c1.vertices[i].type, 0u);
  }
  CHECK_EQ(c1.connectivity.size(), c2.connectivity.size());
  if (!c1.connectivity.empty() && !c2.connectivity.empty())
    CHECK_EQ(c1.connectivity[0].indices == c2.connectivity[0].indices, true);
  check_complex(c2);
  // f = (x - cx)^2 + 1.5 (y - cy)^2 with a moving centre: its gradient's critical
  // points are minima
  ftk::ndarray<double> g({(size_t)2, (size_t)W, (size_t)H, (size_t)Tn});
  for (int t = 0; t < Tn; ++t)
    for (int y = 0; y < H; ++y)
      for (int x = 0; x < W; ++x) {
        g.f(0, x, y, t) = 2.0 * (x - (11.37 + 0.41 * t));
        g.f(1, x, y, t) = 3.0 * (y - (9.23 - 0.27 * t));
      }
  std::map<std::string, ftk::ndarray<double>> gdata = {{"G", g}};
  CriticalPointPredicate<2, double> gp;
  gp.vector_var_name = "G";
  gp.field_type = "gradient";
  SimplicialEngine<double, CriticalPointPredicate<2, double>> e3(mesh, gp);
  e3.execute(gdata);
  auto c3 = e3.get_complex();
  CHECK_EQ(c3.vertices.size() > 0, true);
  for (const auto& el : c3.vertices) CHECK_EQ(el.type, (uint32_t)FTK2H_CP_MINIMUM);
  gp.field_type = "flow";  // the same field read as a flow: a source
  SimplicialEngine<double, CriticalPointPredicate<2, double>> e4(mesh, gp);
  e4.execute(gdata);
  for (const auto& el : e4.get_complex().vertices) CHECK_EQ(el.type, (uint32_t)FTK2H_CP_SOURCE);
  std::printf("multicomponent: %zu vertices, classified: %zu\n", c2.vertices.size(), c3.vertices.size());
}

int main() {
  extruded_2d();
  attributes_3d();
  levelset_3d();
  fiber_3d();
  critical_point_3d();
  critical_point_2d();
  streaming();
  multicomponent_and_classification();
  std::printf("%d failures\n", failures);
  return failures;
}
