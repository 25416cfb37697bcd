// cxxjij_module.cpp -- the `cxxjij` pybind11 surface for the single-spin-flip
// hot path, on top of the C ABI of include/ojmi.h.
//
// Mirrors the Python-visible names of the reference binding
// (openjij/main.cpp:36-226, openjij/declare.hpp) for:
//   graph:     Graph, Dense, Sparse, CSRSparse         (declare.hpp:43-217)
//   system:    ClassicalIsing_{Dense,Sparse,CSRSparse},
//              TransverseIsing_{Dense,Sparse,CSRSparse},
//              make_classical_ising, make_transverse_ising (declare.hpp:409-438,585-641)
//   algorithm: Algorithm_SingleSpinFlip_run, four overloads per system (declare.hpp:687-793)
//   utility:   ClassicalUpdaterParameter, TransverseFieldUpdaterParameter,
//              ClassicalSchedule, TransverseFieldSchedule, make_*_schedule_list
//              (declare.hpp:795-879, include/openjij/utility/schedule_list.hpp)
//   result:    get_solution (declare.hpp:883-888)
// Every anneal runs on the GPU through ojmi_run; there is no CPU sweep here.
// A callback is honoured by running one sweep per ojmi_run call (the Philox
// stream is indexed by the global sweep number, so the trajectory is the same)
// and calling back after each, as Algorithm::run does
// (include/openjij/algorithm/algorithm.hpp:37-45).
#include <pybind11/functional.h>
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <climits>
#include <cmath>
#include <cstdint>
#include <algorithm>
#include <map>
#include <memory>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "ojmi.h"

namespace py = pybind11;
using namespace pybind11::literals;

namespace {

// ---------------------------------------------------------------------------
// errors: status codes -> the exception types the reference raises
void check(int rc) {
  if (rc == OJMI_OK) return;
  std::string msg = ojmi_last_error();
  if (rc == OJMI_EINVAL) throw std::invalid_argument(msg);
  if (rc == OJMI_ERANGE) throw std::out_of_range(msg);
  throw std::runtime_error(msg);
}

// xorshift128 with the reference's fixed state (include/openjij/utility/random.hpp:39-87);
// drives gen_spin / gen_binary so a seed gives the reference's initial state.
struct Xorshift {
  using result_type = uint_fast32_t;
  static constexpr unsigned min() { return 0u; }
  static constexpr unsigned max() { return UINT_MAX; }
  explicit Xorshift(unsigned s) : w(s) {}
  unsigned operator()() {
    unsigned t = x ^ (x << 11);
    x = y;
    y = z;
    z = w;
    return w = (w ^ (w >> 19)) ^ (t ^ (t >> 8));
  }
  unsigned x = 123456789u, y = 362436069u, z = 521288629u, w;
};

// ---------------------------------------------------------------------------
// schedules (include/openjij/utility/schedule_list.hpp:32-257)
struct ClassicalParam {
  double beta = 0;
};
struct TransverseParam {
  double beta = 0, s = 0;
};
struct ClassicalSchedule {
  ClassicalParam updater_parameter;
  size_t one_mc_step = 0;
};
struct TransverseSchedule {
  TransverseParam updater_parameter;
  size_t one_mc_step = 0;
};

std::vector<ClassicalSchedule> make_classical_schedule_list(double beta_min, double beta_max,
                                                            size_t one_mc_step,
                                                            size_t num_call_updater) {
  double r = std::pow(beta_max / beta_min, 1.0 / static_cast<double>(num_call_updater - 1));
  double b = beta_min;
  std::vector<ClassicalSchedule> out(num_call_updater);
  for (auto& e : out) {
    e.one_mc_step = one_mc_step;
    e.updater_parameter.beta = b;
    b *= r;
  }
  return out;
}

std::vector<TransverseSchedule> make_transverse_field_schedule_list(double beta,
                                                                    size_t one_mc_step,
                                                                    size_t num_call_updater) {
  double ds = 1.0 / static_cast<double>(num_call_updater - 1);
  double s = 0;
  std::vector<TransverseSchedule> out(num_call_updater);
  for (auto& e : out) {
    e.one_mc_step = one_mc_step;
    e.updater_parameter.beta = beta;
    e.updater_parameter.s = s;
    s += ds;
  }
  return out;
}

// ---------------------------------------------------------------------------
// graphs
struct Graph {
  size_t n = 0;
  explicit Graph(size_t num) : n(num) {}
  virtual ~Graph() = default;
  std::vector<int> gen_spin_seed(size_t seed) const {
    Xorshift g((unsigned)seed);
    std::uniform_int_distribution<> uid(0, 1);
    std::vector<int> s(n);
    for (auto& e : s) e = 2 * uid(g) - 1;
    return s;
  }
  std::vector<int> gen_binary_seed(size_t seed) const {
    Xorshift g((unsigned)seed);
    std::uniform_int_distribution<> uid(0, 1);
    std::vector<int> s(n);
    for (auto& e : s) e = uid(g);
    return s;
  }
};

// (n+1)^2 row-major; the upper triangle is authoritative (selfadjointView<Upper>).
struct DenseGraph : Graph {
  std::vector<double> J;
  explicit DenseGraph(size_t num) : Graph(num), J((num + 1) * (num + 1), 0.0) {
    J[num * (num + 1) + num] = 1;
  }
  double& at(size_t i, size_t j) { return J[i * (n + 1) + j]; }
  double at(size_t i, size_t j) const { return J[i * (n + 1) + j]; }

  // include/openjij/graph/dense.hpp:117-140
  void set_interaction_matrix(py::array_t<double, py::array::c_style | py::array::forcecast> m) {
    if (m.ndim() != 2 || m.shape(0) != m.shape(1))
      throw std::runtime_error("interaction.rows() != interaction.cols()");
    if ((size_t)m.shape(0) != n + 1) throw std::runtime_error("invalid matrix size.");
    auto a = m.unchecked<2>();
    for (size_t i = 0; i < n; ++i)
      if (a(i, i) != 0)
        throw std::runtime_error("The diagonal elements of interaction matrix must be zero.");
    if (a(n, n) != 1)
      throw std::runtime_error("The right bottom element of interaction matrix must be unity.");
    for (size_t i = 0; i <= n; ++i)
      for (size_t j = i; j <= n; ++j) {
        at(i, j) = a(i, j);
        at(j, i) = a(i, j);
      }
  }
  std::vector<double> symmetric() const {
    std::vector<double> out(J.size());
    for (size_t i = 0; i <= n; ++i)
      for (size_t j = i; j <= n; ++j) out[i * (n + 1) + j] = out[j * (n + 1) + i] = at(i, j);
    return out;
  }
  // include/openjij/graph/dense.hpp:176-192
  double energy(const std::vector<double>& s) const {
    if (s.size() != n) throw std::out_of_range("Out of range in energy in Dense graph.");
    double e = 0;
    for (size_t i = 0; i < n; ++i) {
      double acc = 0;
      for (size_t j = i + 1; j < n; ++j) acc += at(i, j) * s[j];
      e += s[i] * (acc + at(i, n));
    }
    return e;
  }
  // J(i,j) / h(i) accessors (dense.hpp:211-249)
  double& Jref(size_t i, size_t j) {
    if (i >= n || j >= n) throw std::out_of_range("index out of range");
    if (i != j) return at(std::min(i, j), std::max(i, j));
    return at(i, n);
  }
};

// Upper-triangle CSR of the (n+1) x (n+1) matrix.
struct CSRGraph : Graph {
  std::vector<int64_t> rowptr, col;
  std::vector<double> val;
  explicit CSRGraph(size_t num) : Graph(num) {}
  double energy(const std::vector<double>& s) const {
    if (s.size() != n) throw std::out_of_range("Out of range in energy in CSRSparse graph.");
    double e = 0;
    for (size_t i = 0; i < n; ++i) {
      double acc = 0;
      for (int64_t k = rowptr[i]; k < rowptr[i + 1]; ++k) {
        const int64_t j = col[k];
        if (j > (int64_t)i && j < (int64_t)n) acc += val[k] * s[j];
        else if (j == (int64_t)n) acc += val[k];
      }
      e += s[i] * acc;
    }
    return e;
  }
};

// Adjacency-list graph (include/openjij/graph/sparse.hpp:60-316): J(i,j) and
// h(i) = J(i,i) register an edge in both adjacency lists, at most num_edges
// per node; a coupling whose edge did not fit is stored but, as in the
// reference's release build, not part of the system's matrix.
struct SparseGraph : Graph {
  size_t num_edges;
  std::vector<std::vector<size_t>> adj;
  std::map<std::pair<size_t, size_t>, double> Jm;
  SparseGraph(size_t num, size_t edges) : Graph(num), num_edges(std::min(num, edges)), adj(num) {}
  explicit SparseGraph(size_t num) : SparseGraph(num, num) {}
  void check_index(size_t i) const {
    if (i >= n) throw std::out_of_range("index out of range");
  }
  bool set_adj(size_t from, size_t to) {
    auto& nodes = adj[from];
    if (std::find(nodes.begin(), nodes.end(), to) == nodes.end()) {
      if (nodes.size() >= num_edges) return false;
      nodes.push_back(to);
      set_adj(to, from);
    }
    return true;
  }
  double& J(size_t i, size_t j) {
    check_index(i);
    check_index(j);
    set_adj(i, j);
    return Jm[std::make_pair(std::min(i, j), std::max(i, j))];
  }
  double Jget(size_t i, size_t j) const {
    check_index(i);
    check_index(j);
    auto it = Jm.find(std::make_pair(std::min(i, j), std::max(i, j)));
    if (it == Jm.end()) throw std::out_of_range("map::at");
    return it->second;
  }
  // include/openjij/graph/sparse.hpp:231-250
  double energy(const std::vector<double>& s) const {
    if (s.size() != n) throw std::out_of_range("Out of range in energy in Sparse graph.");
    double e = 0;
    for (size_t i = 0; i < n; ++i)
      for (size_t j : adj[i]) {
        if (i != j) e += 0.5 * Jget(i, j) * s[i] * s[j];
        else e += Jget(i, i) * s[i];
      }
    return e;
  }
  // the system matrix of include/openjij/utility/eigen.hpp:162-191, upper triangle
  CSRGraph csr() const {
    CSRGraph g(n);
    g.rowptr.assign(1, 0);
    for (size_t i = 0; i < n; ++i) {
      std::vector<std::pair<int64_t, double>> row;
      for (size_t j : adj[i]) {
        if (j > i) row.push_back({(int64_t)j, Jget(i, j)});
        else if (j == i) row.push_back({(int64_t)n, Jget(i, i)});
      }
      std::sort(row.begin(), row.end());
      for (auto& p : row) {
        g.col.push_back(p.first);
        g.val.push_back(p.second);
      }
      g.rowptr.push_back((int64_t)g.col.size());
    }
    g.col.push_back((int64_t)n);
    g.val.push_back(1.0);
    g.rowptr.push_back((int64_t)g.col.size());
    return g;
  }
};

std::shared_ptr<CSRGraph> csr_from_scipy(py::object mat) {
  py::object csr = mat.attr("tocsr")();
  py::tuple shape = csr.attr("shape");
  const int64_t rows = shape[0].cast<int64_t>(), cols = shape[1].cast<int64_t>();
  if (rows != cols) throw std::runtime_error("interaction.rows() != interaction.cols()");
  if (rows < 1) throw std::runtime_error("invalid matrix size.");
  auto g = std::make_shared<CSRGraph>((size_t)(rows - 1));
  auto ip = py::array_t<int64_t, py::array::forcecast>(csr.attr("indptr"));
  auto ix = py::array_t<int64_t, py::array::forcecast>(csr.attr("indices"));
  auto dv = py::array_t<double, py::array::forcecast>(csr.attr("data"));
  // keep the upper triangle only, sorted by column (selfadjointView<Upper>)
  std::vector<std::vector<std::pair<int64_t, double>>> up((size_t)rows);
  auto ipu = ip.unchecked<1>();
  auto ixu = ix.unchecked<1>();
  auto dvu = dv.unchecked<1>();
  for (int64_t i = 0; i < rows; ++i)
    for (int64_t k = ipu(i); k < ipu(i + 1); ++k) {
      const int64_t j = ixu(k);
      if (j >= i) up[i].push_back({j, dvu(k)});
    }
  g->rowptr.assign(1, 0);
  for (int64_t i = 0; i < rows; ++i) {
    std::sort(up[i].begin(), up[i].end());
    for (auto& p : up[i]) {
      if (!g->col.empty() && (int64_t)g->rowptr.back() < (int64_t)g->col.size() &&
          g->col.back() == p.first) {
        g->val.back() += p.second;
        continue;
      }
      g->col.push_back(p.first);
      g->val.push_back(p.second);
    }
    g->rowptr.push_back((int64_t)g->col.size());
  }
  return g;
}

py::object csr_to_scipy_symmetric(const CSRGraph& g) {
  std::vector<int64_t> r, c;
  std::vector<double> v;
  for (size_t i = 0; i <= g.n; ++i)
    for (int64_t k = g.rowptr[i]; k < g.rowptr[i + 1]; ++k) {
      r.push_back((int64_t)i);
      c.push_back(g.col[k]);
      v.push_back(g.val[k]);
      if ((size_t)g.col[k] != i) {
        r.push_back(g.col[k]);
        c.push_back((int64_t)i);
        v.push_back(g.val[k]);
      }
    }
  py::module_ sp = py::module_::import("scipy.sparse");
  py::object coo = sp.attr("coo_matrix")(
      py::make_tuple(py::array(py::cast(v)), py::make_tuple(py::array(py::cast(r)), py::array(py::cast(c)))),
      "shape"_a = py::make_tuple(g.n + 1, g.n + 1));
  return coo.attr("tocsr")();
}

// ---------------------------------------------------------------------------
// systems: host state plus a lazily created one-replica GPU context
struct Ctx {
  ojmi_ctx* p = nullptr;
  // the spins the device holds (after the last upload or download): a run
  // whose state is unchanged keeps the device's fields, as the reference
  // system keeps dE from one Algorithm_run to the next
  std::vector<int8_t> held;
  ~Ctx() {
    if (p) ojmi_destroy(p);
  }
};

template <typename G>
struct ClassicalSystem {
  std::shared_ptr<G> graph;
  std::vector<double> spin;  // n+1, dummy last
  size_t num_spins;
  std::shared_ptr<Ctx> ctx;

  // the system keeps its own copy of the couplings, as the reference's
  // `const MatrixXx interaction` member does (classical_ising.hpp:92-97)
  ClassicalSystem(const std::vector<int>& init, std::shared_ptr<G> g)
      : graph(std::make_shared<G>(*g)), num_spins(graph->n) {
    reset_spins(init);
  }
  void reset_spins(const std::vector<int>& init) {
    if (init.size() != num_spins) throw std::out_of_range("init_spin has the wrong size");
    for (int v : init)
      if (v != 1 && v != -1) throw std::invalid_argument("spins must be +1 or -1");
    spin.assign(init.begin(), init.end());
    spin.push_back(1.0);
  }
  ojmi_ctx* context();
};

template <typename G>
struct TransverseSystem {
  std::shared_ptr<G> graph;
  std::vector<std::vector<double>> trotter;  // M columns of n+1 (dummy last)
  size_t num_classical_spins;
  double gamma;
  std::shared_ptr<Ctx> ctx;

  TransverseSystem(const std::vector<std::vector<int>>& init, std::shared_ptr<G> g, double gm)
      : graph(std::make_shared<G>(*g)), num_classical_spins(graph->n), gamma(gm) {
    if (!(init.size() >= 2))
      throw std::invalid_argument("trotter slices must be equal or larger than 2.");
    set(init);
  }
  TransverseSystem(const std::vector<int>& init, std::shared_ptr<G> g, double gm, size_t M)
      : graph(std::make_shared<G>(*g)), num_classical_spins(graph->n), gamma(gm) {
    if (!(M >= 2)) throw std::invalid_argument("trotter slices must be equal or larger than 2.");
    set(std::vector<std::vector<int>>(M, init));
  }
  void set(const std::vector<std::vector<int>>& init) {
    trotter.clear();
    for (auto& col : init) {
      if (col.size() != num_classical_spins)
        throw std::out_of_range("trotter spins have the wrong size");
      std::vector<double> c(col.begin(), col.end());
      for (double v : c)
        if (v != 1 && v != -1) throw std::invalid_argument("spins must be +1 or -1");
      c.push_back(1.0);
      trotter.push_back(std::move(c));
    }
  }
  ojmi_ctx* context();
};

ojmi_ctx* make_ctx(const DenseGraph& g, int64_t M, double gamma) {
  ojmi_ctx* c = nullptr;
  check(ojmi_create_dense((int64_t)g.n, g.J.data(), 1, M, gamma, OJMI_PREC_AUTO, &c));
  return c;
}
ojmi_ctx* make_ctx(const CSRGraph& g, int64_t M, double gamma) {
  ojmi_ctx* c = nullptr;
  check(ojmi_create_csr((int64_t)g.n, g.rowptr.data(), g.col.data(), g.val.data(), 1, M, gamma,
                        OJMI_PREC_AUTO, &c));
  return c;
}

ojmi_ctx* make_ctx(const SparseGraph& g, int64_t M, double gamma) {
  return make_ctx(g.csr(), M, gamma);
}

template <typename G>
ojmi_ctx* ClassicalSystem<G>::context() {
  if (!ctx) {
    ctx = std::make_shared<Ctx>();
    ctx->p = make_ctx(*graph, 1, 0.0);
  }
  return ctx->p;
}
template <typename G>
ojmi_ctx* TransverseSystem<G>::context() {
  if (!ctx) {
    ctx = std::make_shared<Ctx>();
    ctx->p = make_ctx(*graph, (int64_t)trotter.size(), gamma);
  }
  check(ojmi_set_gamma(ctx->p, gamma));
  return ctx->p;
}

// state <-> device
template <typename G>
void push(ClassicalSystem<G>& s, ojmi_ctx* c) {
  std::vector<int8_t> v(s.num_spins);
  const double d = s.spin[s.num_spins];
  for (size_t i = 0; i < s.num_spins; ++i) v[i] = (int8_t)(s.spin[i] * d > 0 ? 1 : -1);
  s.spin[s.num_spins] = 1;
  for (size_t i = 0; i < s.num_spins; ++i) s.spin[i] = v[i];
  if (s.ctx->held != v) {
    check(ojmi_set_spins(c, v.data()));
    s.ctx->held = v;
  }
}
template <typename G>
void pull(ClassicalSystem<G>& s, ojmi_ctx* c) {
  std::vector<int8_t> v(s.num_spins);
  check(ojmi_get_spins(c, v.data()));
  for (size_t i = 0; i < s.num_spins; ++i) s.spin[i] = v[i];
  s.ctx->held = v;
}
template <typename G>
void push(TransverseSystem<G>& s, ojmi_ctx* c) {
  const size_t n = s.num_classical_spins, M = s.trotter.size();
  std::vector<int8_t> v(n * M);
  for (size_t t = 0; t < M; ++t)
    for (size_t i = 0; i < n; ++i) v[t * n + i] = (int8_t)(s.trotter[t][i] > 0 ? 1 : -1);
  if (s.ctx->held != v) {
    check(ojmi_set_spins(c, v.data()));
    s.ctx->held = v;
  }
}
template <typename G>
void pull(TransverseSystem<G>& s, ojmi_ctx* c) {
  const size_t n = s.num_classical_spins, M = s.trotter.size();
  std::vector<int8_t> v(n * M);
  check(ojmi_get_spins(c, v.data()));
  for (size_t t = 0; t < M; ++t)
    for (size_t i = 0; i < n; ++i) s.trotter[t][i] = v[t * n + i];
  s.ctx->held = v;
}

// ---------------------------------------------------------------------------
// Algorithm_SingleSpinFlip_run
template <typename Sys, typename Sched>
void run_impl(Sys& sys, uint64_t seed, const std::vector<Sched>& sched, const py::object& cb) {
  constexpr bool transverse = std::is_same<Sched, TransverseSchedule>::value;
  ojmi_ctx* c = sys.context();
  push(sys, c);
  const size_t L = sched.size();
  std::vector<double> beta(L), s(L);
  std::vector<int64_t> mcs(L);
  for (size_t l = 0; l < L; ++l) {
    beta[l] = sched[l].updater_parameter.beta;
    if constexpr (transverse) s[l] = sched[l].updater_parameter.s;
    mcs[l] = (int64_t)sched[l].one_mc_step;
  }
  if (cb.is_none()) {
    int rc;
    {
      py::gil_scoped_release rel;
      rc = ojmi_run(c, beta.data(), transverse ? s.data() : nullptr, mcs.data(), (int64_t)L,
                    seed, 0);
    }
    check(rc);
    pull(sys, c);
    return;
  }
  uint64_t t = 0;
  const int64_t one = 1;
  for (size_t l = 0; l < L; ++l)
    for (int64_t k = 0; k < mcs[l]; ++k, ++t) {
      int rc;
      {
        py::gil_scoped_release rel;
        rc = ojmi_run(c, &beta[l], transverse ? &s[l] : nullptr, &one, 1, seed, t);
      }
      check(rc);
      pull(sys, c);
      if constexpr (transverse)
        cb(py::cast(&sys, py::return_value_policy::reference), py::make_tuple(beta[l], s[l]));
      else
        cb(py::cast(&sys, py::return_value_policy::reference), beta[l]);
    }
}

uint64_t random_seed() { return (uint64_t)std::random_device{}(); }

template <typename Sys, typename Sched>
void declare_run(py::module_& m) {
  constexpr bool transverse = std::is_same<Sched, TransverseSchedule>::value;
  using Tuple = typename std::conditional<transverse, std::pair<double, double>, double>::type;
  using TupleList = std::vector<std::pair<Tuple, size_t>>;
  auto from_tuples = [](const TupleList& tl) {
    std::vector<Sched> out;
    out.reserve(tl.size());
    for (auto& e : tl) {
      Sched s;
      if constexpr (transverse) {
        s.updater_parameter.beta = e.first.first;
        s.updater_parameter.s = e.first.second;
      } else {
        s.updater_parameter.beta = e.first;
      }
      s.one_mc_step = e.second;
      out.push_back(s);
    }
    return out;
  };
  m.def(
      "Algorithm_SingleSpinFlip_run",
      [](Sys& sys, size_t seed, const std::vector<Sched>& sl, py::object cb) {
        run_impl(sys, (uint64_t)seed, sl, cb);
      },
      "system"_a, "seed"_a, "schedule_list"_a, "callback"_a = py::none());
  m.def(
      "Algorithm_SingleSpinFlip_run",
      [](Sys& sys, const std::vector<Sched>& sl, py::object cb) {
        run_impl(sys, random_seed(), sl, cb);
      },
      "system"_a, "schedule_list"_a, "callback"_a = py::none());
  m.def(
      "Algorithm_SingleSpinFlip_run",
      [from_tuples](Sys& sys, size_t seed, const TupleList& tl, py::object cb) {
        run_impl(sys, (uint64_t)seed, from_tuples(tl), cb);
      },
      "system"_a, "seed"_a, "tuplelist"_a, "callback"_a = py::none());
  m.def(
      "Algorithm_SingleSpinFlip_run",
      [from_tuples](Sys& sys, const TupleList& tl, py::object cb) {
        run_impl(sys, random_seed(), from_tuples(tl), cb);
      },
      "system"_a, "tuplelist"_a, "callback"_a = py::none());
}

// get_solution (include/openjij/result/get_solution.hpp:39-84)
template <typename G>
std::vector<int> solution(ClassicalSystem<G>& s) {
  std::vector<int> out(s.num_spins);
  const double d = s.spin[s.num_spins];
  for (size_t i = 0; i < s.num_spins; ++i) out[i] = (int)(s.spin[i] * d);
  return out;
}
template <typename G>
std::vector<int> solution(TransverseSystem<G>& s) {
  // first slice minimising s^T J_ext s  (= 2 E(s) + 1)
  size_t best = 0;
  double be = 0;
  for (size_t t = 0; t < s.trotter.size(); ++t) {
    std::vector<double> v(s.trotter[t].begin(), s.trotter[t].end() - 1);
    const double e = s.graph->energy(v);
    if (t == 0 || e < be) {
      be = e;
      best = t;
    }
  }
  std::vector<int> out(s.num_classical_spins);
  for (size_t i = 0; i < s.num_classical_spins; ++i) out[i] = (int)s.trotter[best][i];
  return out;
}

py::array_t<double> vec_to_np(const std::vector<double>& v) {
  py::array_t<double> a((py::ssize_t)v.size());
  std::copy(v.begin(), v.end(), a.mutable_data());
  return a;
}

template <typename G>
void declare_systems(py::module_& msys, py::module_& malg, py::module_& mres,
                     const std::string& suffix) {
  using CS = ClassicalSystem<G>;
  using TS = TransverseSystem<G>;
  py::class_<CS>(msys, ("ClassicalIsing" + suffix).c_str(), py::module_local())
      .def(py::init<const std::vector<int>&, std::shared_ptr<G>>(), "init_spin"_a,
           "init_interaction"_a)
      .def("reset_spins", &CS::reset_spins, "init_spin"_a)
      .def_property(
          "spin", [](const CS& s) { return vec_to_np(s.spin); },
          [](CS& s, py::array_t<double, py::array::forcecast> a) {
            if ((size_t)a.size() != s.num_spins + 1)
              throw std::out_of_range("spin has the wrong size");
            std::copy(a.data(), a.data() + a.size(), s.spin.begin());
          })
      .def_property_readonly("interaction",
                             [](const CS& s) -> py::object {
                               if constexpr (std::is_same<G, DenseGraph>::value) {
                                 auto v = s.graph->symmetric();
                                 py::array_t<double> a({s.num_spins + 1, s.num_spins + 1});
                                 std::copy(v.begin(), v.end(), a.mutable_data());
                                 return std::move(a);
                               } else if constexpr (std::is_same<G, SparseGraph>::value) {
                                 return csr_to_scipy_symmetric(s.graph->csr());
                               } else {
                                 return csr_to_scipy_symmetric(*s.graph);
                               }
                             })
      .def_readonly("num_spins", &CS::num_spins);
  msys.def(
      "make_classical_ising",
      [](const std::vector<int>& init, std::shared_ptr<G> g) { return CS(init, g); },
      "init_spin"_a, "init_interaction"_a);

  py::class_<TS>(msys, ("TransverseIsing" + suffix).c_str(), py::module_local())
      .def(py::init<const std::vector<std::vector<int>>&, std::shared_ptr<G>, double>(),
           "init_spin"_a, "init_interaction"_a, "gamma"_a)
      .def(py::init<const std::vector<int>&, std::shared_ptr<G>, double, size_t>(),
           "init_classical_spins"_a, "init_interaction"_a, "gamma"_a, "num_trotter_slices"_a)
      .def(
          "reset_spins",
          [](TS& s, const std::vector<std::vector<int>>& init) {
            if (init.size() != s.trotter.size())
              throw std::out_of_range("wrong number of trotter slices");
            s.set(init);
          },
          "init_trotter_spins"_a)
      .def(
          "reset_spins",
          [](TS& s, const std::vector<int>& cl) {
            s.set(std::vector<std::vector<int>>(s.trotter.size(), cl));
          },
          "classical_spins"_a)
      .def_property(
          "trotter_spins",
          [](const TS& s) {
            const size_t n1 = s.num_classical_spins + 1, M = s.trotter.size();
            py::array_t<double> a({n1, M});
            auto m = a.mutable_unchecked<2>();
            for (size_t t = 0; t < M; ++t)
              for (size_t i = 0; i < n1; ++i) m(i, t) = s.trotter[t][i];
            return a;
          },
          [](TS& s, py::array_t<double, py::array::forcecast> a) {
            const size_t n1 = s.num_classical_spins + 1, M = s.trotter.size();
            if (a.ndim() != 2 || (size_t)a.shape(0) != n1 || (size_t)a.shape(1) != M)
              throw std::out_of_range("trotter_spins has the wrong shape");
            auto m = a.unchecked<2>();
            for (size_t t = 0; t < M; ++t)
              for (size_t i = 0; i < n1; ++i) s.trotter[t][i] = m(i, t);
          })
      .def_property_readonly("interaction",
                             [](const TS& s) -> py::object {
                               if constexpr (std::is_same<G, DenseGraph>::value) {
                                 auto v = s.graph->symmetric();
                                 const size_t n1 = s.num_classical_spins + 1;
                                 py::array_t<double> a({n1, n1});
                                 std::copy(v.begin(), v.end(), a.mutable_data());
                                 return std::move(a);
                               } else if constexpr (std::is_same<G, SparseGraph>::value) {
                                 return csr_to_scipy_symmetric(s.graph->csr());
                               } else {
                                 return csr_to_scipy_symmetric(*s.graph);
                               }
                             })
      .def_readonly("num_classical_spins", &TS::num_classical_spins)
      .def_readwrite("gamma", &TS::gamma);
  msys.def(
      "make_transverse_ising",
      [](const std::vector<std::vector<int>>& init, std::shared_ptr<G> g, double gamma) {
        return TS(init, g, gamma);
      },
      "init_trotter_spins"_a, "init_interaction"_a, "gamma"_a);
  msys.def(
      "make_transverse_ising",
      [](const std::vector<int>& init, std::shared_ptr<G> g, double gamma, size_t M) {
        return TS(init, g, gamma, M);
      },
      "classical_spins"_a, "init_interaction"_a, "gamma"_a, "num_trotter_slices"_a);

  declare_run<CS, ClassicalSchedule>(malg);
  declare_run<TS, TransverseSchedule>(malg);
  mres.def("get_solution", [](CS& s) { return solution(s); }, "system"_a);
  mres.def("get_solution", [](TS& s) { return solution(s); }, "system"_a);
}

// ---------------------------------------------------------------------------
// Batched engine: num_replicas copies annealed in one launch.  This is the
// GPU counterpart of the per-read loop of BaseSampler._cxxjij_sampling
// (openjij/sampler/sampler.py:115-140); it has no single reference function.
struct Engine {
  std::shared_ptr<Ctx> ctx;
  int64_t n = 0, R = 0, M = 1;

  static py::array_t<int8_t> empty_i8(int64_t a, int64_t b, int64_t c) {
    if (c > 0) return py::array_t<int8_t>({a, b, c});
    return py::array_t<int8_t>({a, b});
  }
  void set_spins(py::array_t<int8_t, py::array::c_style | py::array::forcecast> s) {
    if (s.size() != R * M * n) throw std::out_of_range("spins must hold R*M*n entries");
    check(ojmi_set_spins(ctx->p, s.data()));
  }
  py::array_t<int8_t> spins() {
    auto a = M > 1 ? empty_i8(R, M, n) : empty_i8(R, n, 0);
    check(ojmi_get_spins(ctx->p, a.mutable_data()));
    return a;
  }
  void run(py::array_t<double, py::array::forcecast> beta, py::object s,
           py::array_t<int64_t, py::array::forcecast> mcs, uint64_t seed, uint64_t offset) {
    const int64_t L = beta.size();
    if (mcs.size() != L) throw std::invalid_argument("beta and one_mc_step differ in length");
    py::array_t<double, py::array::forcecast> sv;
    const double* sp = nullptr;
    if (!s.is_none()) {
      sv = py::array_t<double, py::array::forcecast>(s);
      if (sv.size() != L) throw std::invalid_argument("beta and s differ in length");
      sp = sv.data();
    }
    int rc;
    {
      py::gil_scoped_release rel;
      rc = ojmi_run(ctx->p, beta.data(), sp, mcs.data(), L, seed, offset);
    }
    check(rc);
  }
  py::array_t<int8_t> solution() {
    auto a = empty_i8(R, n, 0);
    check(ojmi_solution(ctx->p, a.mutable_data()));
    return a;
  }
  py::array_t<double> energies() {
    py::array_t<double> a = M > 1 ? py::array_t<double>({R, M}) : py::array_t<double>(R);
    check(ojmi_energies(ctx->p, a.mutable_data()));
    return a;
  }
  double drift() {
    double d = 0;
    check(ojmi_field_drift(ctx->p, &d));
    return d;
  }
  void set_option(const std::string& k, int64_t v) { check(ojmi_set_option(ctx->p, k.c_str(), v)); }
  void set_gamma(double g) { check(ojmi_set_gamma(ctx->p, g)); }
  py::dict info() {
    ojmi_info i{};
    check(ojmi_get_info(ctx->p, &i));
    py::dict d;
    d["num_spins"] = i.num_spins;
    d["num_replicas"] = i.num_replicas;
    d["num_trotter"] = i.num_trotter;
    d["precision"] = i.precision;
    d["kernel"] = i.kernel;
    d["refresh_every"] = i.refresh_every;
    d["coupling_bytes"] = i.coupling_bytes;
    d["last_sweeps"] = i.last_sweeps;
    d["last_attempts"] = i.last_attempts;
    d["last_kernel_ms"] = i.last_kernel_ms;
    d["last_refresh_ms"] = i.last_refresh_ms;
    d["group"] = i.group;
    d["ring_rows"] = i.ring_rows;
    d["lds_bytes"] = i.lds_bytes;
    d["lds_budget"] = i.lds_budget;
    d["num_problems"] = i.num_problems;
    d["waves_per_replica"] = i.waves_per_replica;
    return d;
  }
};

Engine make_engine_dense(const DenseGraph& g, int64_t R, int64_t M, double gamma, int prec) {
  Engine e;
  e.ctx = std::make_shared<Ctx>();
  check(ojmi_create_dense((int64_t)g.n, g.J.data(), R, M, gamma, prec, &e.ctx->p));
  e.n = (int64_t)g.n;
  e.R = R;
  e.M = M;
  return e;
}
Engine make_engine_csr(const CSRGraph& g, int64_t R, int64_t M, double gamma, int prec) {
  Engine e;
  e.ctx = std::make_shared<Ctx>();
  check(ojmi_create_csr((int64_t)g.n, g.rowptr.data(), g.col.data(), g.val.data(), R, M, gamma,
                        prec, &e.ctx->p));
  e.n = (int64_t)g.n;
  e.R = R;
  e.M = M;
  return e;
}

}  // namespace

PYBIND11_MODULE(cxxjij, m) {
  m.doc() = "MI355X single-spin-flip engine behind OpenJij's cxxjij names";

  py::module_ mg = m.def_submodule("graph", "cxxjij submodule for graph");
  py::class_<Graph, std::shared_ptr<Graph>>(mg, "Graph", py::module_local())
      .def(py::init<size_t>(), "num_spins"_a)
      .def("gen_spin", &Graph::gen_spin_seed, "seed"_a)
      .def("gen_spin", [](const Graph& g) { return g.gen_spin_seed(std::random_device{}()); })
      .def("gen_binary", &Graph::gen_binary_seed, "seed"_a)
      .def("gen_binary", [](const Graph& g) { return g.gen_binary_seed(std::random_device{}()); })
      .def("size", [](const Graph& g) { return g.n; })
      .def("get_num_spins", [](const Graph& g) { return g.n; });

  py::class_<DenseGraph, Graph, std::shared_ptr<DenseGraph>>(mg, "Dense", py::module_local())
      .def(py::init<size_t>(), "num_spins"_a)
      .def(py::init([](const DenseGraph& o) { return std::make_shared<DenseGraph>(o); }),
           "other"_a)
      .def("set_interaction_matrix", &DenseGraph::set_interaction_matrix, "interaction"_a)
      .def(
          "calc_energy",
          [](const DenseGraph& g, const std::vector<double>& s) { return g.energy(s); },
          "spins"_a)
      .def(
          "__setitem__",
          [](DenseGraph& g, std::pair<size_t, size_t> k, double v) { g.Jref(k.first, k.second) = v; },
          "key"_a, "val"_a)
      .def(
          "__getitem__",
          [](DenseGraph& g, std::pair<size_t, size_t> k) { return g.Jref(k.first, k.second); },
          "key"_a)
      .def(
          "__setitem__", [](DenseGraph& g, size_t k, double v) { g.Jref(k, k) = v; }, "key"_a,
          "val"_a)
      .def(
          "__getitem__", [](DenseGraph& g, size_t k) { return g.Jref(k, k); }, "key"_a)
      .def("get_interactions", [](const DenseGraph& g) {
        auto v = g.symmetric();
        py::array_t<double> a({g.n + 1, g.n + 1});
        std::copy(v.begin(), v.end(), a.mutable_data());
        return a;
      });

  py::class_<CSRGraph, Graph, std::shared_ptr<CSRGraph>>(mg, "CSRSparse", py::module_local())
      .def(py::init([](py::object mat) { return csr_from_scipy(mat); }), "interaction"_a)
      .def(
          "calc_energy",
          [](const CSRGraph& g, const std::vector<double>& s) { return g.energy(s); },
          "spins"_a)
      .def("get_interactions", [](const CSRGraph& g) { return csr_to_scipy_symmetric(g); });

  py::class_<SparseGraph, Graph, std::shared_ptr<SparseGraph>>(mg, "Sparse", py::module_local())
      .def(py::init<size_t, size_t>(), "num_spins"_a, "num_edges"_a)
      .def(py::init<size_t>(), "num_spins"_a)
      .def(py::init([](const SparseGraph& o) { return std::make_shared<SparseGraph>(o); }),
           "other"_a)
      .def("adj_nodes",
           [](const SparseGraph& g, size_t i) {
             g.check_index(i);
             return g.adj[i];
           })
      .def("get_num_edges", [](const SparseGraph& g) { return g.num_edges; })
      .def(
          "calc_energy",
          [](const SparseGraph& g, const std::vector<double>& s) { return g.energy(s); },
          "spins"_a)
      .def(
          "__setitem__",
          [](SparseGraph& g, std::pair<size_t, size_t> k, double v) { g.J(k.first, k.second) = v; },
          "key"_a, "val"_a)
      .def(
          "__getitem__",
          [](const SparseGraph& g, std::pair<size_t, size_t> k) { return g.Jget(k.first, k.second); },
          "key"_a)
      .def(
          "__setitem__", [](SparseGraph& g, size_t k, double v) { g.J(k, k) = v; }, "key"_a,
          "val"_a)
      .def(
          "__getitem__", [](const SparseGraph& g, size_t k) { return g.Jget(k, k); }, "key"_a);

  py::module_ mu = m.def_submodule("utility", "cxxjij module for utility");
  py::class_<ClassicalParam>(mu, "ClassicalUpdaterParameter", py::module_local())
      .def(py::init<>())
      .def(py::init([](double b) { return ClassicalParam{b}; }), "beta"_a)
      .def_readwrite("beta", &ClassicalParam::beta)
      .def("__repr__", [](const ClassicalParam& p) {
        return "(beta: " + std::to_string(p.beta) + ")";
      });
  py::class_<TransverseParam>(mu, "TransverseFieldUpdaterParameter", py::module_local())
      .def(py::init<>())
      .def(py::init([](double b, double s) { return TransverseParam{b, s}; }), "beta"_a, "s"_a)
      .def(py::init([](std::pair<double, double> o) { return TransverseParam{o.first, o.second}; }),
           "obj"_a)
      .def_readwrite("beta", &TransverseParam::beta)
      .def_readwrite("s", &TransverseParam::s)
      .def("__repr__", [](const TransverseParam& p) {
        return "(beta: " + std::to_string(p.beta) + ", s: " + std::to_string(p.s) + ")";
      });
  py::class_<ClassicalSchedule>(mu, "ClassicalSchedule", py::module_local())
      .def(py::init<>())
      .def(py::init([](std::pair<ClassicalParam, size_t> o) {
             return ClassicalSchedule{o.first, o.second};
           }),
           "obj"_a)
      .def_readwrite("one_mc_step", &ClassicalSchedule::one_mc_step)
      .def_readwrite("updater_parameter", &ClassicalSchedule::updater_parameter)
      .def("__repr__", [](const ClassicalSchedule& s) {
        return "((beta: " + std::to_string(s.updater_parameter.beta) +
               ") mcs: " + std::to_string(s.one_mc_step) + ")";
      });
  py::class_<TransverseSchedule>(mu, "TransverseFieldSchedule", py::module_local())
      .def(py::init<>())
      .def(py::init([](std::pair<TransverseParam, size_t> o) {
             return TransverseSchedule{o.first, o.second};
           }),
           "obj"_a)
      .def_readwrite("one_mc_step", &TransverseSchedule::one_mc_step)
      .def_readwrite("updater_parameter", &TransverseSchedule::updater_parameter)
      .def("__repr__", [](const TransverseSchedule& s) {
        return "((beta: " + std::to_string(s.updater_parameter.beta) +
               ", s: " + std::to_string(s.updater_parameter.s) +
               ") mcs: " + std::to_string(s.one_mc_step) + ")";
      });
  mu.def("make_classical_schedule_list", &make_classical_schedule_list, "beta_min"_a,
         "beta_max"_a, "one_mc_step"_a, "num_call_updater"_a);
  mu.def("make_transverse_field_schedule_list", &make_transverse_field_schedule_list, "beta"_a,
         "one_mc_step"_a, "num_call_updater"_a);
  mu.def(
      "make_schedule_list",
      [](const std::vector<std::pair<double, size_t>>& tl) {
        std::vector<ClassicalSchedule> out;
        for (auto& e : tl) out.push_back(ClassicalSchedule{ClassicalParam{e.first}, e.second});
        return out;
      },
      "tuplelist"_a);
  mu.def(
      "make_schedule_list",
      [](const std::vector<std::pair<std::pair<double, double>, size_t>>& tl) {
        std::vector<TransverseSchedule> out;
        for (auto& e : tl)
          out.push_back(
              TransverseSchedule{TransverseParam{e.first.first, e.first.second}, e.second});
        return out;
      },
      "tuplelist"_a);

  py::module_ ms = m.def_submodule("system", "cxxjij module for system");
  py::module_ ma = m.def_submodule("algorithm", "cxxjij module for algorithm");
  py::module_ mr = m.def_submodule("result", "cxxjij module for result");
  declare_systems<DenseGraph>(ms, ma, mr, "_Dense");
  declare_systems<SparseGraph>(ms, ma, mr, "_Sparse");
  declare_systems<CSRGraph>(ms, ma, mr, "_CSRSparse");
  py::module_ me = m.def_submodule("engine", "batched GPU engine (num_reads replicas per launch)");
  me.attr("PREC_AUTO") = OJMI_PREC_AUTO;
  me.attr("PREC_FP64") = OJMI_PREC_FP64;
  me.attr("PREC_FP32") = OJMI_PREC_FP32;
  me.attr("PREC_INT32") = OJMI_PREC_INT32;
  me.attr("KERNEL_DENSE_WAVE") = OJMI_KERNEL_DENSE_WAVE;
  me.attr("KERNEL_SQA_WAVE") = OJMI_KERNEL_SQA_WAVE;
  me.attr("KERNEL_CSR_LEVEL") = OJMI_KERNEL_CSR_LEVEL;
  me.attr("KERNEL_DENSE_SHARED") = OJMI_KERNEL_DENSE_SHARED;
  me.attr("KERNEL_DENSE_COLOWN") = OJMI_KERNEL_DENSE_COLOWN;
  me.attr("KERNEL_BATCH_WAVE") = OJMI_KERNEL_BATCH_WAVE;
  me.attr("KERNEL_DENSE_LDS") = OJMI_KERNEL_DENSE_LDS;
  py::class_<Engine>(me, "Engine", py::module_local())
      .def(py::init([](const DenseGraph& g, int64_t R, int64_t M, double gamma, int prec) {
             return make_engine_dense(g, R, M, gamma, prec);
           }),
           "graph"_a, "num_replicas"_a, "num_trotter"_a = 1, "gamma"_a = 0.0,
           "precision"_a = OJMI_PREC_AUTO)
      .def(py::init([](const CSRGraph& g, int64_t R, int64_t M, double gamma, int prec) {
             return make_engine_csr(g, R, M, gamma, prec);
           }),
           "graph"_a, "num_replicas"_a, "num_trotter"_a = 1, "gamma"_a = 0.0,
           "precision"_a = OJMI_PREC_AUTO)
      .def("set_spins", &Engine::set_spins, "spins"_a)
      .def("spins", &Engine::spins)
      .def("run", &Engine::run, "beta"_a, "s"_a, "one_mc_step"_a, "seed"_a, "sweep_offset"_a = 0)
      .def("solution", &Engine::solution)
      .def("energies", &Engine::energies)
      .def("drift", &Engine::drift)
      .def("set_option", &Engine::set_option, "key"_a, "value"_a)
      .def("set_gamma", &Engine::set_gamma, "gamma"_a)
      .def("info", &Engine::info)
      .def_readonly("num_spins", &Engine::n)
      .def_readonly("num_replicas", &Engine::R)
      .def_readonly("num_trotter", &Engine::M);
  m.def("device_count", []() {
    int c = 0;
    ojmi_device_count(&c);
    return c;
  });
}
