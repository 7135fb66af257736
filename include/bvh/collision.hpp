// bvh/collision.hpp -- C++17 facade of distBVH's collision API over the C ABI
// (include/distbvh.h).  Header only; link against libdistbvh.so.
//
// Names and meaning follow the reference:
//   bvh::collision_world         collision_world.hpp:54-114
//   bvh::collision_object        collision_object.hpp:55-283
//   bvh::split_algorithm         types.hpp:52-57
//   bvh::narrowphase_result(_pair), typed_narrowphase_result<T>,
//   bvh::broadphase_collision<T> collision_query.hpp:50-161
//   bvh::patch<>                 patch.hpp:55-123
//   bvh::sort_and_sweep_local, bvh::max_variant_axis  narrowphase/kokkos.hpp:119-248
// Differences: there is one process and one GPU (no vt); errors throw
// bvh::bvh_error instead of terminating; element types tell the library how
// to bound them through bvh::element_traits (below).
#ifndef BVH_COLLISION_HPP
#define BVH_COLLISION_HPP

#include <array>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <functional>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "../distbvh.h"

namespace bvh {

class bvh_error : public std::runtime_error {
 public:
  bvh_error(const char* fn, bvh_status s)
      : std::runtime_error(std::string(fn) + ": " + bvh_status_string(s) + ": " + bvh_last_error()), status(s) {}
  bvh_status status;
};

namespace detail {
inline void check(bvh_status s, const char* fn) {
  if (s != BVH_OK) throw bvh_error(fn, s);
}
}  // namespace detail
#define BVH_FACADE_CALL(expr) ::bvh::detail::check((expr), #expr)

// ---- minimal span (C++17) ------------------------------------------------------------
template <typename T>
class span {
 public:
  span() = default;
  span(T* p, std::size_t n) : p_(p), n_(n) {}
  template <typename C, typename = decltype(std::declval<C&>().data())>
  span(C& c) : p_(c.data()), n_(c.size()) {}
  T* data() const { return p_; }
  std::size_t size() const { return n_; }
  bool empty() const { return n_ == 0; }
  T& operator[](std::size_t i) const { return p_[i]; }
  T* begin() const { return p_; }
  T* end() const { return p_ + n_; }

 private:
  T* p_ = nullptr;
  std::size_t n_ = 0;
};

enum class split_algorithm { geom_axis = BVH_SPLIT_GEOM_AXIS, ml_geom_axis = BVH_SPLIT_ML_GEOM_AXIS,
                             clustering = BVH_SPLIT_CLUSTERING };

struct world_config {
  std::uint32_t morton_bits = 63;  // 63 (21 bits / axis) or 30 (the reference clustering codes)
  std::uint32_t leaf_size = 1;
  bool fp32_outward = false;
  bool sort_results = false;
  bool timing = false;
  void* stream = nullptr;  // hipStream_t
  int device = 0;
};

// k-DOP as the reference lays it out: 13 (min, max) extents
struct kdop26 {
  double s[26];
  double min(int axis) const { return s[2 * axis]; }
  double max(int axis) const { return s[2 * axis + 1]; }
};

// patch<>, patch.hpp:55-123
template <typename KDop = kdop26>
class patch {
 public:
  patch() = default;
  patch(std::size_t gid, std::size_t size, const KDop& kd, const std::array<double, 3>& c)
      : gid_(gid), size_(size), kdop_(kd), centroid_(c) {}
  std::size_t global_id() const noexcept { return gid_; }
  KDop kdop() const noexcept { return kdop_; }
  std::array<double, 3> centroid() const noexcept { return centroid_; }
  bool empty() const noexcept { return size_ == 0; }
  std::size_t size() const noexcept { return size_; }

 private:
  std::size_t gid_ = static_cast<std::size_t>(-1);
  std::size_t size_ = 0;
  KDop kdop_{};
  std::array<double, 3> centroid_{{0.0, 0.0, 0.0}};
};

// narrowphase_result, collision_query.hpp:66-128: type-erased POD records
class narrowphase_result {
 public:
  narrowphase_result() = default;
  explicit narrowphase_result(std::size_t stride) : stride_(stride) {}
  void* allocate(std::size_t n) {
    n_ += n;
    const std::size_t off = data_.size();
    data_.resize(off + n * stride_, 0);
    return data_.data() + off;
  }
  void append_data(const void* p, std::size_t n) {
    const auto* b = static_cast<const unsigned char*>(p);
    data_.insert(data_.end(), b, b + n * stride_);
    n_ += n;
  }
  const void* at(std::size_t i) const { return data_.data() + i * stride_; }
  void* at(std::size_t i) { return data_.data() + i * stride_; }
  std::size_t stride() const noexcept { return stride_; }
  std::size_t size() const noexcept { return n_; }

 private:
  std::vector<unsigned char> data_;
  std::size_t stride_ = 0;
  std::size_t n_ = 0;
};

// typed_narrowphase_result<T>, collision_query.hpp:130-155
template <typename T>
class typed_narrowphase_result : public narrowphase_result {
 public:
  typed_narrowphase_result() : narrowphase_result(sizeof(T)) {}
  template <typename... Args>
  T& emplace_back(Args&&... a) {
    return *(new (allocate(1)) T(std::forward<Args>(a)...));
  }
  T& operator[](std::size_t i) { return *static_cast<T*>(at(i)); }
};

// narrowphase_result_pair, collision_query.hpp:157-161
struct narrowphase_result_pair {
  narrowphase_result a;
  narrowphase_result b;
};

// The record the built-in narrowphase emits (detailed_narrowphase_result,
// tests/collision_object_test.cpp:318-324).
using detailed_narrowphase_result = bvh_narrowphase_record;

class collision_object;

// broadphase_collision<T>, collision_query.hpp:50-64
template <typename T>
struct broadphase_collision {
  broadphase_collision(collision_object& o, const patch<>& m, std::size_t id, span<const T> e)
      : object(o), meta(m), patch_id(id), elements(e) {}
  collision_object& object;
  patch<> meta;
  std::size_t patch_id;
  span<const T> elements;
};

// How the library bounds an element type.  Specialize for your type with
//   static constexpr int vertices = N;                 // polytope: N vertices
//   static void vertex(const T&, int v, double xyz[3]);
//   static std::uint64_t global_id(const T&);
// for kdop = from_vertices (kdop.hpp:238-264) and centroid = vertex mean, or
//   static constexpr int vertices = 0;                 // any ContactEntity
//   static kdop26 kdop(const T&);  static std::array<double,3> centroid(const T&);
//   static std::uint64_t global_id(const T&);
// (traits.hpp:114-135: the reference's get_entity_kdop / _centroid / _global_id).
template <typename T>
struct element_traits;

namespace detail {
template <typename T>
using functor_t = std::function<narrowphase_result_pair(const broadphase_collision<T>&,
                                                        const broadphase_collision<T>&)>;
struct type_erased_functor {
  virtual ~type_erased_functor() = default;
};
template <typename T>
struct typed_functor : type_erased_functor {
  functor_t<T> f;
};
}  // namespace detail

class collision_world;

class collision_object {
 public:
  collision_object(const collision_object&) = delete;
  collision_object& operator=(const collision_object&) = delete;

  // set_entity_data(View<T*>, split_algorithm), collision_object.hpp:69-84.
  // The elements are read during the call; like the reference
  // (collision_object.cpp:128-129) the object keeps a pointer to them for a
  // user narrowphase functor, so they must outlive the next broadphase.
  template <typename T>
  void set_entity_data(span<const T> elements, split_algorithm alg);

  // Device-resident polytope data (SoA planes, see distbvh.h).
  void set_entity_data_vertices(const double* verts, std::uint32_t verts_per_element, std::size_t n,
                                const std::uint64_t* gids = nullptr, double epsilon = 0.0, bool device = true,
                                split_algorithm alg = split_algorithm::clustering) {
    BVH_FACADE_CALL(bvh_set_entity_data_poly(h_, verts, verts_per_element, n, gids, epsilon,
                                             device ? BVH_MEM_DEVICE : BVH_MEM_HOST, static_cast<std::uint32_t>(alg)));
    elements_ = nullptr;
    n_ = n;
  }

  void init_broadphase() { BVH_FACADE_CALL(bvh_init_broadphase(h_)); }
  void broadphase(collision_object& other);
  void end_phase() { BVH_FACADE_CALL(bvh_end_phase(h_)); }
  int overdecomposition_factor() const noexcept { return static_cast<int>(od_); }
  std::size_t id() const noexcept { return id_; }

  // for_each_result<R>(f), collision_object.hpp:177-187
  template <typename R, typename F>
  void for_each_result(F&& f) {
    if (!functor_results_.empty()) {
      for (const auto& res : functor_results_)
        for (std::size_t i = 0; i < res.size(); ++i) f(*static_cast<const R*>(res.at(i)));
      return;
    }
    static_assert(sizeof(R) == sizeof(bvh_narrowphase_record),
                  "the built-in narrowphase emits 32-byte {patch_p, element_p, patch_q, element_q} records");
    std::uint64_t n = 0;
    BVH_FACADE_CALL(bvh_result_count(h_, &n));
    std::vector<bvh_narrowphase_record> recs(n);
    if (n) BVH_FACADE_CALL(bvh_result_records(h_, recs.data(), n, BVH_MEM_HOST));
    for (const auto& r : recs) {
      R v;
      std::memcpy(static_cast<void*>(&v), &r, sizeof(R));
      f(v);
    }
  }

  // local_patches(), collision_object.hpp:189
  std::vector<patch<>> local_patches() const {
    std::vector<bvh_patch_info> pi(od_);
    BVH_FACADE_CALL(bvh_local_patches(h_, pi.data(), od_));
    std::vector<patch<>> out;
    out.reserve(od_);
    for (const auto& p : pi) {
      kdop26 kd;
      std::memcpy(kd.s, p.kdop, sizeof(kd.s));
      out.emplace_back(p.global_id, p.size, kd, std::array<double, 3>{{p.centroid[0], p.centroid[1], p.centroid[2]}});
    }
    return out;
  }

  // The patch-level tree for_each_tree hands out (snapshot_tree, tree.hpp:
  // 75-452): pre-order nodes (node.hpp:318-322) over the non-empty patches.
  struct patch_tree {
    std::vector<bvh_tree_node> nodes;
    std::vector<std::uint64_t> leafs;  // patch ids, left to right
    std::size_t count() const noexcept { return leafs.size(); }
    bool empty() const noexcept { return leafs.empty(); }
  };

  // for_each_tree(f), collision_object.hpp:164-168: f(const patch_tree &)
  template <typename F>
  void for_each_tree(F&& f) const {
    std::size_t nn = 0, nl = 0;
    BVH_FACADE_CALL(bvh_object_patch_tree(h_, nullptr, 0, &nn, nullptr, 0, &nl));
    patch_tree t;
    t.nodes.resize(nn);
    t.leafs.resize(nl);
    BVH_FACADE_CALL(bvh_object_patch_tree(h_, t.nodes.data(), nn, &nn, t.leafs.data(), nl, &nl));
    std::forward<F>(f)(static_cast<const patch_tree&>(t));
  }

  // the element-level tree's node count and root (bvh_object_tree)
  std::pair<std::size_t, std::size_t> tree_shape() const {
    std::size_t nn = 0, root = 0;
    BVH_FACADE_CALL(bvh_object_tree(h_, nullptr, nullptr, nullptr, 0, &nn, &root));
    return {nn, root};
  }

  bvh_object handle() const noexcept { return h_; }

 private:
  friend class collision_world;
  collision_object(collision_world& w, bvh_object h, std::size_t id, std::size_t od) : world_(&w), h_(h), id_(id), od_(od) {}
  template <typename T>
  void run_functor(collision_object& other, const detail::functor_t<T>& fn);

  collision_world* world_;
  bvh_object h_;
  std::size_t id_;
  std::size_t od_;
  const void* elements_ = nullptr;
  std::size_t n_ = 0;
  std::vector<narrowphase_result> functor_results_;
};

class collision_world {
 public:
  template <typename T>
  using narrowphase_functor = detail::functor_t<T>;

  explicit collision_world(std::size_t od, const world_config& cfg = {}) {
    bvh_world_config c;
    bvh_world_config_init(&c);
    c.morton_bits = cfg.morton_bits;
    c.leaf_size = cfg.leaf_size;
    c.precision = cfg.fp32_outward ? BVH_PREC_FP32_OUTWARD : BVH_PREC_FP64;
    c.flags = (cfg.sort_results ? BVH_FLAG_SORT_RESULTS : 0u) | (cfg.timing ? BVH_FLAG_TIMING : 0u);
    c.stream = cfg.stream;
    c.device = cfg.device;
    BVH_FACADE_CALL(bvh_world_create(od, &c, &h_));
    od_ = od;
  }
  ~collision_world() { bvh_world_destroy(h_); }
  collision_world(const collision_world&) = delete;
  collision_world& operator=(const collision_world&) = delete;

  collision_object& create_collision_object() {
    bvh_object o = nullptr;
    BVH_FACADE_CALL(bvh_create_collision_object(h_, &o));
    objects_.emplace_back(new collision_object(*this, o, objects_.size(), od_));
    return *objects_.back();
  }
  std::size_t num_collision_objects() const noexcept { return objects_.size(); }
  std::size_t overdecomposition_factor() const noexcept { return od_; }

  // set_narrowphase_functor<T>, collision_world.hpp:75-94: the functor runs on
  // the host per broadphase patch pair over the patches' elements.  Without
  // one, the built-in device narrowphase (strict 26-DOP overlap) runs.
  template <typename T>
  void set_narrowphase_functor(narrowphase_functor<T> f) {
    auto tf = std::make_unique<detail::typed_functor<T>>();
    tf->f = std::move(f);
    functor_ = std::move(tf);
  }
  void clear_narrowphase_functor() { functor_.reset(); }

  void start_iteration() { BVH_FACADE_CALL(bvh_start_iteration(h_)); }
  void finish_iteration() { BVH_FACADE_CALL(bvh_finish_iteration(h_)); }
  bvh_world handle() const noexcept { return h_; }

 private:
  friend class collision_object;
  bvh_world h_ = nullptr;
  std::size_t od_ = 1;
  std::vector<std::unique_ptr<collision_object>> objects_;
  std::unique_ptr<detail::type_erased_functor> functor_;
  void (*functor_runner_)(collision_object&, collision_object&, detail::type_erased_functor&) = nullptr;
};

// ---- element ingestion -------------------------------------------------------------------
template <typename T>
void collision_object::set_entity_data(span<const T> el, split_algorithm alg) {
  using tr = element_traits<T>;
  const std::size_t n = el.size();
  std::vector<std::uint64_t> gids(n);
  for (std::size_t i = 0; i < n; ++i) gids[i] = tr::global_id(el[i]);
  if constexpr (tr::vertices > 0) {
    constexpr int nv = tr::vertices;
    std::vector<double> planes(static_cast<std::size_t>(nv) * 3 * n);
    for (std::size_t e = 0; e < n; ++e)
      for (int v = 0; v < nv; ++v) {
        double xyz[3];
        tr::vertex(el[e], v, xyz);
        for (int c = 0; c < 3; ++c) planes[(static_cast<std::size_t>(v) * 3 + c) * n + e] = xyz[c];
      }
    BVH_FACADE_CALL(bvh_set_entity_data_poly(h_, planes.data(), nv, n, gids.data(), 0.0, BVH_MEM_HOST,
                                             static_cast<std::uint32_t>(alg)));
  } else {
    std::vector<double> kd(26 * n), ce(3 * n);
    for (std::size_t e = 0; e < n; ++e) {
      const kdop26 k = tr::kdop(el[e]);
      std::memcpy(&kd[26 * e], k.s, sizeof(k.s));
      const std::array<double, 3> c = tr::centroid(el[e]);
      for (int i = 0; i < 3; ++i) ce[3 * e + i] = c[i];
    }
    BVH_FACADE_CALL(bvh_set_entity_data_snapshot(h_, kd.data(), ce.data(), n, gids.data(), BVH_MEM_HOST,
                                                 static_cast<std::uint32_t>(alg)));
  }
  elements_ = el.data();
  n_ = n;
  world_->functor_runner_ = [](collision_object& a, collision_object& b, detail::type_erased_functor& f) {
    auto* tf = dynamic_cast<detail::typed_functor<T>*>(&f);
    if (!tf) throw std::runtime_error("narrowphase functor element type does not match set_entity_data");
    a.template run_functor<T>(b, tf->f);
  };
}

// The reference's narrowphase: the functor per broadphase patch pair with the
// two patches' elements (collision_object/impl.cpp:269-350); both halves of
// the result land in the initiating object (impl.cpp:328-348).
template <typename T>
void collision_object::run_functor(collision_object& other, const detail::functor_t<T>& fn) {
  std::uint64_t np = 0;
  BVH_FACADE_CALL(bvh_patch_pair_count(h_, &np));
  std::vector<std::uint64_t> pairs(2 * np);
  if (np) BVH_FACADE_CALL(bvh_patch_pairs(h_, pairs.data(), np));
  const auto pa = local_patches();
  const auto pb = other.local_patches();
  const T* ea = static_cast<const T*>(elements_);
  const T* eb = static_cast<const T*>(other.elements_);
  std::vector<std::uint64_t> idx(std::max(n_, other.n_) + 1);
  auto gather = [&](collision_object& o, const T* base, std::size_t patch_id, std::vector<T>& out) {
    std::uint64_t cnt = 0;
    BVH_FACADE_CALL(bvh_patch_elements(o.h_, patch_id, idx.data(), idx.size(), &cnt));
    out.clear();
    out.reserve(cnt);
    for (std::uint64_t i = 0; i < cnt; ++i) out.push_back(base[idx[i]]);
  };
  std::vector<T> va, vb;
  for (std::uint64_t k = 0; k < np; ++k) {
    const std::size_t p = pairs[2 * k], q = pairs[2 * k + 1];
    gather(*this, ea, p, va);
    gather(other, eb, q, vb);
    broadphase_collision<T> first(*this, pa[p], p, span<const T>(va.data(), va.size()));
    broadphase_collision<T> second(other, pb[q], q, span<const T>(vb.data(), vb.size()));
    narrowphase_result_pair r = fn(first, second);
    if (r.a.size() > 0) functor_results_.push_back(std::move(r.a));
    if (r.b.size() > 0) functor_results_.push_back(std::move(r.b));
  }
}

inline void collision_object::broadphase(collision_object& other) {
  functor_results_.clear();
  if (world_->functor_ && &other != this) {
    if (!elements_ || !other.elements_)
      throw std::runtime_error("a narrowphase functor needs host elements (set_entity_data with a span)");
    BVH_FACADE_CALL(bvh_broadphase_patches(h_, other.h_));
    world_->functor_runner_(*this, other, *world_->functor_);
  } else {
    BVH_FACADE_CALL(bvh_broadphase(h_, other.h_));
  }
}

// ---- sort and sweep between two patches (narrowphase/kokkos.hpp:119-248) ---------------
// These run on the world's device; the reference's take no world.
namespace detail {
// the elements' 26-DOPs (n x 26, the build's own values) and centroids (n x 3)
template <typename T>
void element_bounds(collision_world& w, span<const T> el, std::vector<double>& kd, std::vector<double>& ce) {
  using tr = element_traits<T>;
  const std::size_t n = el.size();
  kd.assign(26 * n, 0.0);
  ce.assign(3 * n, 0.0);
  if constexpr (tr::vertices > 0) {
    constexpr int nv = tr::vertices;
    std::vector<double> planes(static_cast<std::size_t>(nv) * 3 * n);
    for (std::size_t e = 0; e < n; ++e)
      for (int v = 0; v < nv; ++v) {
        double xyz[3];
        tr::vertex(el[e], v, xyz);
        for (int c = 0; c < 3; ++c) {
          planes[(static_cast<std::size_t>(v) * 3 + c) * n + e] = xyz[c];
          ce[3 * e + c] += xyz[c];
        }
      }
    for (double& c : ce) c /= nv;
    if (n) BVH_FACADE_CALL(bvh_element_kdops(w.handle(), planes.data(), nv, n, 0.0, BVH_MEM_HOST, kd.data()));
  } else {
    for (std::size_t e = 0; e < n; ++e) {
      const kdop26 k = tr::kdop(el[e]);
      std::memcpy(&kd[26 * e], k.s, sizeof(k.s));
      const std::array<double, 3> c = tr::centroid(el[e]);
      for (int i = 0; i < 3; ++i) ce[3 * e + i] = c[i];
    }
  }
}
}  // namespace detail

// patch(id, span), patch.hpp:67-78: the union of the elements' k-DOPs and
// the mean of their centroids
template <typename T>
patch<> make_patch(collision_world& w, std::size_t gid, span<const T> el) {
  std::vector<double> kd, ce;
  detail::element_bounds(w, el, kd, ce);
  kdop26 u;
  for (int i = 0; i < 13; ++i) {
    u.s[2 * i] = std::numeric_limits<double>::infinity();
    u.s[2 * i + 1] = -std::numeric_limits<double>::infinity();
  }
  std::array<double, 3> c{{0.0, 0.0, 0.0}};
  for (std::size_t e = 0; e < el.size(); ++e) {
    for (int i = 0; i < 13; ++i) {
      const double lo = kd[26 * e + 2 * i], hi = kd[26 * e + 2 * i + 1];
      u.s[2 * i] = (lo < u.s[2 * i]) ? lo : u.s[2 * i];
      u.s[2 * i + 1] = (u.s[2 * i + 1] < hi) ? hi : u.s[2 * i + 1];
    }
    for (int i = 0; i < 3; ++i) c[i] += ce[3 * e + i];
  }
  if (el.size())
    for (double& x : c) x /= static_cast<double>(el.size());
  return patch<>(gid, el.size(), u, c);
}

// max_variant_axis(a, b), narrowphase/kokkos.hpp:119-161: the axis of
// largest variance of both element sets' centroids
template <typename T>
int max_variant_axis(collision_world& w, span<const T> a, span<const T> b) {
  std::vector<double> ka, ca, kb, cb;
  detail::element_bounds(w, a, ka, ca);
  detail::element_bounds(w, b, kb, cb);
  std::int32_t axis = 0;
  BVH_FACADE_CALL(bvh_max_variant_axis(w.handle(), ca.data(), a.size(), cb.data(), b.size(), BVH_MEM_HOST, &axis));
  return axis;
}

// sort_and_sweep_local(pa, a, pb, b, axis, f), narrowphase/kokkos.hpp:164-248:
// f(a[i], b[j]) for every pair whose 26-DOPs overlap, for each element of a
// in the order the reference's sweep finds them (B sorted by its scaled
// minima on `axis`)
template <typename T, typename F>
void sort_and_sweep_local(collision_world& w, const patch<>& pa, span<const T> a, const patch<>& pb, span<const T> b,
                          int axis, F&& f) {
  std::vector<double> ka, ca, kb, cb;
  detail::element_bounds(w, a, ka, ca);
  detail::element_bounds(w, b, kb, cb);
  const kdop26 pka = pa.kdop(), pkb = pb.kdop();
  std::uint64_t np = 0;
  BVH_FACADE_CALL(bvh_sort_and_sweep(w.handle(), pka.s, ka.data(), nullptr, a.size(), pkb.s, kb.data(), nullptr,
                                     b.size(), axis, BVH_MEM_HOST, nullptr, 0, &np));
  std::vector<std::uint64_t> pairs(2 * np);
  if (np)
    BVH_FACADE_CALL(bvh_sort_and_sweep(w.handle(), pka.s, ka.data(), nullptr, a.size(), pkb.s, kb.data(), nullptr,
                                       b.size(), axis, BVH_MEM_HOST, pairs.data(), np, &np));
  for (std::uint64_t k = 0; k < np; ++k) f(a[pairs[2 * k]], b[pairs[2 * k + 1]]);
}

}  // namespace bvh

#endif
