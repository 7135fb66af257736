// C++ facade test (include/bvh/collision.hpp), written like the reference's
// tests/collision_object_test.cpp cases it mirrors.  Exit status 0 = pass.
#include <algorithm>
#include <array>
#include <cstdio>
#include <set>
#include <utility>
#include <vector>

#include "bvh/collision.hpp"

// The reference test element: an axis-aligned hexahedron with 8 vertices
// (tests/TestCommon.hpp:48-127).
struct Element {
  std::array<std::array<double, 3>, 8> v;
  std::size_t id;
};

template <>
struct bvh::element_traits<Element> {
  static constexpr int vertices = 8;
  static void vertex(const Element& e, int i, double xyz[3]) {
    for (int c = 0; c < 3; ++c) xyz[c] = e.v[i][c];
  }
  static std::uint64_t global_id(const Element& e) { return e.id; }
};

// build_element_grid, tests/TestCommon.hpp:183-211 (x fastest)
static std::vector<Element> grid(int nx, int ny, int nz, std::size_t base) {
  std::vector<Element> out;
  const double dx = 1.0 / nx, dy = 1.0 / ny, dz = 1.0 / nz;
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const double x = i * dx, y = j * dy, z = k * dz;
        Element e;
        e.id = base + static_cast<std::size_t>(i + j * nx + k * nx * ny);
        e.v = {{{x, y, z}, {x + dx, y, z}, {x + dx, y + dy, z}, {x, y + dy, z},
                {x, y, z + dz}, {x + dx, y, z + dz}, {x + dx, y + dy, z + dz}, {x, y + dy, z + dz}}};
        out.push_back(e);
      }
  return out;
}

static int failures = 0;
#define CHECK(c)                                                      \
  do {                                                                \
    if (!(c)) {                                                       \
      std::fprintf(stderr, "%s:%d CHECK failed: %s\n", __FILE__, __LINE__, #c); \
      ++failures;                                                     \
    }                                                                 \
  } while (0)

// "collision_object narrowphase" (collision_object_test.cpp:376-443), one rank
static void narrowphase_functor_case() {
  bvh::collision_world world(2);
  auto& obj = world.create_collision_object();
  auto& obj2 = world.create_collision_object();
  world.start_iteration();
  auto e1 = grid(1, 1, 1, 0);
  obj.set_entity_data(bvh::span<const Element>(e1.data(), e1.size()), bvh::split_algorithm::clustering);
  obj.init_broadphase();
  auto e2 = grid(2, 3, 2, 0);
  obj2.set_entity_data(bvh::span<const Element>(e2.data(), e2.size()), bvh::split_algorithm::clustering);
  obj2.init_broadphase();
  world.set_narrowphase_functor<Element>([](const bvh::broadphase_collision<Element>& a,
                                            const bvh::broadphase_collision<Element>& b) {
    bvh::narrowphase_result_pair res;
    res.a = bvh::narrowphase_result(sizeof(bvh::detailed_narrowphase_result));
    res.b = bvh::narrowphase_result(sizeof(bvh::detailed_narrowphase_result));
    auto& resa = static_cast<bvh::typed_narrowphase_result<bvh::detailed_narrowphase_result>&>(res.a);
    CHECK(a.object.id() == 0);
    CHECK(b.object.id() == 1);
    CHECK(a.elements.size() == 1);
    for (auto&& e : b.elements)
      resa.emplace_back(bvh::detailed_narrowphase_result{a.meta.global_id(), a.elements[0].id, b.meta.global_id(), e.id});
    return res;
  });
  obj.broadphase(obj2);
  std::vector<bvh::detailed_narrowphase_result> results;
  obj.for_each_result<bvh::detailed_narrowphase_result>([&](const bvh::detailed_narrowphase_result& r) { results.push_back(r); });
  world.finish_iteration();
  std::sort(results.begin(), results.end(), [](const auto& l, const auto& r) { return l.element_q < r.element_q; });
  CHECK(results.size() == 12);
  for (std::size_t i = 0; i < std::min<std::size_t>(results.size(), 12); ++i) CHECK(results[i].element_q == i);
}

// "narrowphase three objects" (collision_object_test.cpp:445-610) with the
// built-in device narrowphase instead of a functor: every element pair of
// different objects overlaps, 1*2 + 1*2 + 2*2 = 8 unordered pairs.
static void three_objects_builtin_case() {
  bvh::collision_world world(2);
  auto& o0 = world.create_collision_object();
  auto& o1 = world.create_collision_object();
  auto& o2 = world.create_collision_object();
  world.start_iteration();
  auto e0 = grid(1, 1, 1, 0), e1 = grid(1, 2, 1, 0), e2 = grid(1, 1, 2, 0);
  o0.set_entity_data(bvh::span<const Element>(e0.data(), e0.size()), bvh::split_algorithm::clustering);
  o1.set_entity_data(bvh::span<const Element>(e1.data(), e1.size()), bvh::split_algorithm::clustering);
  o2.set_entity_data(bvh::span<const Element>(e2.data(), e2.size()), bvh::split_algorithm::clustering);
  for (auto* o : {&o0, &o1, &o2}) o->init_broadphase();
  std::vector<std::pair<int, int>> got;
  auto collect = [&](bvh::collision_object& a, bvh::collision_object& b, int ia, int ib) {
    a.broadphase(b);
    a.for_each_result<bvh::detailed_narrowphase_result>([&](const bvh::detailed_narrowphase_result& r) {
      got.emplace_back(ia * 10 + static_cast<int>(r.element_p), ib * 10 + static_cast<int>(r.element_q));
    });
  };
  collect(o0, o1, 0, 1);
  collect(o1, o2, 1, 2);
  collect(o2, o0, 2, 0);
  world.finish_iteration();
  std::set<std::pair<int, int>> s;
  for (auto p : got) s.insert(p.first < p.second ? p : std::make_pair(p.second, p.first));
  CHECK(got.size() == 8);
  CHECK(s.size() == 8);
}

// local_patches: od patches, sizes sum to n, union of patch k-DOPs = bounds
static void patches_case() {
  bvh::collision_world world(4);
  auto& o = world.create_collision_object();
  auto e = grid(8, 8, 8, 0);
  o.set_entity_data(bvh::span<const Element>(e.data(), e.size()), bvh::split_algorithm::clustering);
  const auto pats = o.local_patches();
  CHECK(pats.size() == 4);
  std::size_t total = 0;
  double mn = 1e300, mx = -1e300;
  for (const auto& p : pats) {
    total += p.size();
    if (!p.empty()) {
      mn = std::min(mn, p.kdop().min(0));
      mx = std::max(mx, p.kdop().max(0));
    }
  }
  CHECK(total == 512);
  CHECK(mn == 0.0 && mx == 1.0);
  CHECK(o.tree_shape().first == 2 * 512 - 1);
  // collision_object_test.cpp:49-81: for_each_tree's tree has one leaf per
  // (non-empty) patch, none for an object without elements
  o.for_each_tree([](const auto& t) {
    CHECK(t.count() == 4);
    CHECK(t.nodes.size() == 2 * 4 - 1);
    CHECK(t.nodes[0].parent_offset == 0 && t.nodes[0].entity_range[0] == 0 && t.nodes[0].entity_range[1] == 4);
  });
  o.set_entity_data(bvh::span<const Element>(e.data(), 0), bvh::split_algorithm::clustering);
  o.for_each_tree([](const auto& t) { CHECK(t.count() == 0 && t.nodes.empty()); });
  o.set_entity_data(bvh::span<const Element>(e.data(), e.size()), bvh::split_algorithm::geom_axis);
  o.for_each_tree([](const auto& t) { CHECK(t.count() == 4); });
}

// KokkosTest.cpp:37-139 (max variant axis, sort_and_sweep_local), rank 0
static void sort_and_sweep_case() {
  using S = bvh::span<const Element>;
  bvh::collision_world world(1);
  auto cube = grid(1, 1, 1, 0);
  const int dims[3][3] = {{12, 1, 1}, {1, 12, 1}, {1, 1, 12}};
  for (int a = 0; a < 3; ++a) {
    auto v = grid(dims[a][0], dims[a][1], dims[a][2], 0);
    CHECK(bvh::max_variant_axis(world, S(v.data(), v.size()), S(cube.data(), cube.size())) == a);
  }
  auto vec = grid(12, 1, 1, 0);
  const S sv(vec.data(), vec.size()), sc(cube.data(), cube.size());
  const auto p1 = bvh::make_patch(world, 0, sv);
  const auto p2 = bvh::make_patch(world, 1, sc);
  CHECK(p1.size() == 12 && p1.kdop().min(0) == 0.0 && p1.kdop().max(0) == 1.0);
  int count = 0;
  std::set<std::size_t> seen;
  bvh::sort_and_sweep_local(world, p1, sv, p2, sc, 0, [&](const Element& x, const Element& y) {
    CHECK(y.id == 0);
    CHECK(x.id < 12);
    seen.insert(x.id);
    ++count;
  });
  CHECK(count == 12 && seen.size() == 12);
  count = 0;
  bvh::sort_and_sweep_local(world, p2, sc, p1, sv, 0, [&](const Element& x, const Element& y) {
    CHECK(x.id == 0);
    CHECK(y.id < 12);
    ++count;
  });
  CHECK(count == 12);
}

int main() {
  try {
    sort_and_sweep_case();
    narrowphase_functor_case();
    three_objects_builtin_case();
    patches_case();
    try {
      bvh::collision_world bad(3);  // od must be a power of 2 (collision_world.cpp:59-65)
      ++failures;
    } catch (const bvh::bvh_error& e) {
      CHECK(e.status == BVH_E_INVALID);
    }
  } catch (const std::exception& e) {
    std::fprintf(stderr, "exception: %s\n", e.what());
    return 2;
  }
  std::printf("facade_test: %s (%d failures)\n", failures ? "FAILED" : "ok", failures);
  return failures ? 1 : 0;
}
