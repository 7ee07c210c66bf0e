// Oracle: CPU restatement of ArcTN's numerical contraction path.
//
// TEST INFRASTRUCTURE ONLY. This code is the checker for the HIP executor in
// arctn_amd/; only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
// leg may load it. The product never links, imports or calls it.
//
// Every function restates the reference's algorithm and cites the file:line
// it follows (paths relative to the reference repository root).
//
// Pinning (what checks each part independently of this restatement):
//   permute / sum_axis / trace_pair / diag_pair / select_axis
//       - brute-force index loops on small tensors and NumPy transpose/sum/
//         trace/diagonal/take in tests/test_oracle.py.
//   compile (plan_traces / plan_pairwise / finalize)
//       - known geometry of tests/compiled_pair_steps.rs:125-145 and slot/leg
//         invariants of :38-57, checked in tests/test_oracle.py;
//       - compiled == direct contract_network bit-for-bit (correctness.rs:43-103).
//   execute / contract_network / contract_sliced / output blocks
//       - golden fixtures from the reference's own tests (mpi_cli.rs:72-148,
//         scalar32.rs:99-148, slice_validation.rs:22-31, slice_output.rs:52-70,
//         pybind/tests/test_output_slicing.py:49-65) and naive einsum / NumPy
//         einsum on random small networks.
//   naive_einsum - numpy.einsum(optimize=False).
//   slice_parallel_chunks - the closed-form coverage properties of
//         tests/slice_parallel.rs:97-136.
//   execute_stripped - mantissa * 10^exp == plain result (correctness.rs:75-100)
//         and the 1e200 case of compiled.rs:639-669.
//   adjoint (no reference counterpart; the reference has no gradient code,
//         compiled.rs:7-10) - central finite differences in float64/complex128.
//   GEMM bit pattern: parity unpinned (the reference calls matrixmultiply
//         0.3.10, whose blocking/FMA order no test pins); value parity is pinned
//         by the fixtures above.
#pragma once

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <map>
#include <set>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace orc {

using Leg = uint32_t;
using Path = std::vector<std::pair<size_t, size_t>>;

struct Error : std::runtime_error {
  explicit Error(const std::string& s) : std::runtime_error(s) {}
};

template <class... A>
std::string cat(const A&... a) {
  std::ostringstream o;
  (o << ... << a);
  return o.str();
}

inline std::string vec_str(const std::vector<size_t>& v) {
  std::ostringstream o;
  o << "[";
  for (size_t i = 0; i < v.size(); ++i) o << (i ? ", " : "") << v[i];
  o << "]";
  return o.str();
}

// ---------------------------------------------------------------- scalars ---
// Complex arithmetic follows num-complex 0.4.6 (re = ac - bd, im = ad + bc);
// the oracle is compiled with -ffp-contract=off so no FMA is introduced.
template <class T> struct Num;
template <> struct Num<float> {
  static float zero() { return 0.f; }
  static float one() { return 1.f; }
  static float mul(float a, float b) { return a * b; }
  static double abs(float a) { return std::fabs((double)a); }
  static float from_real(double x) { return (float)x; }
  static float conj(float a) { return a; }
};
template <> struct Num<double> {
  static double zero() { return 0.0; }
  static double one() { return 1.0; }
  static double mul(double a, double b) { return a * b; }
  static double abs(double a) { return std::fabs(a); }
  static double from_real(double x) { return x; }
  static double conj(double a) { return a; }
};
template <class R> struct Num<std::complex<R>> {
  using C = std::complex<R>;
  static C zero() { return C(0, 0); }
  static C one() { return C(1, 0); }
  static C mul(C a, C b) {
    return C(a.real() * b.real() - a.imag() * b.imag(), a.real() * b.imag() + a.imag() * b.real());
  }
  // num-complex `norm` is hypot(re, im) (tensor.rs:317-319).
  static double abs(C a) { return std::hypot((double)a.real(), (double)a.imag()); }
  static C from_real(double x) { return C((R)x, 0); }
  static C conj(C a) { return C(a.real(), -a.imag()); }
};
template <class T> inline T add(T a, T b) { return a + b; }
template <class R> inline std::complex<R> add(std::complex<R> a, std::complex<R> b) {
  return std::complex<R>(a.real() + b.real(), a.imag() + b.imag());
}

// ---------------------------------------------------------------- tensor ----
// src/tensor.rs:11-22
inline size_t checked_numel(const std::vector<size_t>& shape) {
  size_t n = 1;
  for (size_t ax = 0; ax < shape.size(); ++ax) {
    if (shape[ax] == 0) throw Error(cat("shape axis ", ax, " is 0; zero-length axes unsupported"));
    if (n > SIZE_MAX / shape[ax]) throw Error(cat("element count of shape ", vec_str(shape), " overflows"));
    n *= shape[ax];
  }
  return n;
}

template <class T>
struct Tensor {
  std::vector<size_t> shape;
  std::vector<T> data;

  Tensor() = default;
  Tensor(std::vector<size_t> s, std::vector<T> d) : shape(std::move(s)), data(std::move(d)) {
    if (checked_numel(shape) != data.size())
      throw Error(cat("tensor shape ", vec_str(shape), " needs ", checked_numel(shape), " elements, got ", data.size()));
  }
  size_t numel() const { return data.size(); }
  size_t ndim() const { return shape.size(); }

  // src/tensor.rs:440-450
  std::vector<size_t> strides() const {
    std::vector<size_t> s(shape.size(), 1);
    for (size_t i = shape.size(); i-- > 1;) s[i - 1] = s[i] * shape[i];
    return s;
  }

  // src/tensor.rs:453-492: output axis d is input axis axes[d]; identity -> clone.
  Tensor permute(const std::vector<size_t>& axes) const {
    if (axes.size() != ndim()) throw Error("permute: rank mismatch");
    std::vector<bool> seen(ndim(), false);
    for (size_t a : axes) {
      if (a >= ndim() || seen[a]) throw Error("permute: axes are not a permutation");
      seen[a] = true;
    }
    bool ident = true;
    for (size_t i = 0; i < axes.size(); ++i) ident = ident && axes[i] == i;
    if (ident) return *this;
    auto in_strides = strides();
    std::vector<size_t> out_shape(ndim()), step(ndim());
    for (size_t d = 0; d < ndim(); ++d) {
      out_shape[d] = shape[axes[d]];
      step[d] = in_strides[axes[d]];
    }
    Tensor out;
    out.shape = out_shape;
    out.data.reserve(numel());
    std::vector<size_t> idx(ndim(), 0);
    size_t off = 0;
    for (size_t e = 0; e < numel(); ++e) {
      out.data.push_back(data[off]);
      for (size_t d = ndim(); d-- > 0;) {
        idx[d] += 1;
        off += step[d];
        if (idx[d] < out_shape[d]) break;
        off -= step[d] * out_shape[d];
        idx[d] = 0;
      }
    }
    return out;
  }

  // Shared walker for sum_axis / trace_pair (sequential accumulation over `d`
  // entries spaced `inner_step`, starting at each output position's base).
  Tensor reduce_walk(const std::vector<size_t>& drop, size_t d, size_t inner_step, size_t base0) const {
    auto st = strides();
    std::vector<size_t> out_shape, out_steps;
    for (size_t i = 0; i < ndim(); ++i)
      if (std::find(drop.begin(), drop.end(), i) == drop.end()) {
        out_shape.push_back(shape[i]);
        out_steps.push_back(st[i]);
      }
    size_t n_out = 1;
    for (size_t s : out_shape) n_out *= s;
    Tensor out;
    out.shape = out_shape;
    out.data.assign(n_out, Num<T>::zero());
    std::vector<size_t> idx(out_shape.size(), 0);
    size_t base = base0;
    for (size_t o = 0; o < n_out; ++o) {
      T acc = Num<T>::zero();
      size_t off = base;
      for (size_t i = 0; i < d; ++i) {
        acc = add(acc, data[off]);
        off += inner_step;
      }
      out.data[o] = acc;
      for (size_t dd = out_shape.size(); dd-- > 0;) {
        idx[dd] += 1;
        base += out_steps[dd];
        if (idx[dd] < out_shape[dd]) break;
        base -= out_steps[dd] * out_shape[dd];
        idx[dd] = 0;
      }
    }
    return out;
  }

  // src/tensor.rs:593-633
  Tensor sum_axis(size_t ax) const {
    auto st = strides();
    return reduce_walk({ax}, shape[ax], st[ax], 0);
  }

  // src/tensor.rs:503-547 (ax_i < ax_j, equal dims)
  Tensor trace_pair(size_t ax_i, size_t ax_j) const {
    if (!(ax_i < ax_j) || shape[ax_i] != shape[ax_j]) throw Error("trace_pair: bad axes");
    auto st = strides();
    return reduce_walk({ax_i, ax_j}, shape[ax_i], st[ax_i] + st[ax_j], 0);
  }

  // Shared walker for diag_pair / select_axis: pure gather with per-axis steps.
  Tensor gather_walk(const std::vector<size_t>& out_shape, const std::vector<size_t>& step, size_t base) const {
    size_t n_out = 1;
    for (size_t s : out_shape) n_out *= s;
    Tensor out;
    out.shape = out_shape;
    out.data.reserve(n_out);
    std::vector<size_t> idx(out_shape.size(), 0);
    size_t off = base;
    for (size_t o = 0; o < n_out; ++o) {
      out.data.push_back(data[off]);
      for (size_t dd = out_shape.size(); dd-- > 0;) {
        idx[dd] += 1;
        off += step[dd];
        if (idx[dd] < out_shape[dd]) break;
        off -= step[dd] * out_shape[dd];
        idx[dd] = 0;
      }
    }
    return out;
  }

  // src/tensor.rs:550-590: keeps axis ax_i, which advances both source axes.
  Tensor diag_pair(size_t ax_i, size_t ax_j) const {
    if (!(ax_i < ax_j) || shape[ax_i] != shape[ax_j]) throw Error("diag_pair: bad axes");
    auto st = strides();
    std::vector<size_t> out_shape, step;
    for (size_t o = 0; o + 1 < ndim(); ++o) {
      size_t in_ax = o < ax_j ? o : o + 1;
      out_shape.push_back(shape[in_ax]);
      step.push_back(st[in_ax] + (o == ax_i ? st[ax_j] : 0));
    }
    return gather_walk(out_shape, step, 0);
  }

  // src/tensor.rs:636-669
  Tensor select_axis(size_t ax, size_t i) const {
    if (i >= shape[ax]) throw Error("select_axis: index out of range");
    auto st = strides();
    std::vector<size_t> out_shape, step;
    for (size_t d = 0; d < ndim(); ++d)
      if (d != ax) {
        out_shape.push_back(shape[d]);
        step.push_back(st[d]);
      }
    return gather_walk(out_shape, step, i * st[ax]);
  }
};

// ---------------------------------------------------------------- network ---
struct Network {
  std::vector<std::vector<Leg>> inputs;
  std::vector<Leg> output;
  std::map<Leg, size_t> size_dict;

  size_t n_tensors() const { return inputs.size(); }
  size_t dim(Leg l) const {
    auto it = size_dict.find(l);
    if (it == size_dict.end()) throw Error(cat("leg ", l, " has no size_dict dimension"));
    return it->second;
  }
  double log2_dim(Leg l) const { return std::log2((double)dim(l)); }
  std::vector<size_t> shape_of(const std::vector<Leg>& legs) const {
    std::vector<size_t> s;
    for (Leg l : legs) s.push_back(dim(l));
    return s;
  }

  // src/network.rs:171-229
  void validate() const {
    if (inputs.empty()) throw Error("TensorNetwork needs at least one input tensor");
    for (auto& kv : size_dict) {
      if (kv.second == 0) throw Error(cat("leg ", kv.first, " has dimension 0"));
      if (kv.second > (size_t)PTRDIFF_MAX) throw Error(cat("leg ", kv.first, " dimension exceeds isize"));
    }
    std::set<Leg> held;
    for (size_t i = 0; i < inputs.size(); ++i) {
      std::map<Leg, size_t> counts;
      std::vector<size_t> shape;
      for (Leg l : inputs[i]) {
        shape.push_back(dim(l));
        counts[l] += 1;
        held.insert(l);
      }
      for (auto& kv : counts)
        if (kv.second > 2) throw Error(cat("leg ", kv.first, " appears ", kv.second, " times in input ", i));
      checked_numel(shape);
    }
    std::set<Leg> seen;
    std::vector<size_t> oshape;
    for (Leg l : output) {
      if (!seen.insert(l).second) throw Error("output has repeated legs");
      if (!held.count(l)) throw Error(cat("output leg ", l, " is not held by any input"));
      oshape.push_back(dim(l));
    }
    checked_numel(oshape);
    for (auto& kv : size_dict)
      if (!held.count(kv.first)) throw Error(cat("size_dict leg ", kv.first, " is not used by any input tensor"));
  }
};

inline std::vector<Leg> sorted_dedup(std::vector<Leg> v) {
  std::sort(v.begin(), v.end());
  v.erase(std::unique(v.begin(), v.end()), v.end());
  return v;
}

// src/contract.rs:16-40
template <class T>
void validate_shapes(const Network& net, const std::vector<Tensor<T>>& ts) {
  net.validate();
  if (ts.size() != net.n_tensors())
    throw Error(cat("tensor count ", ts.size(), " does not match network ", net.n_tensors()));
  for (size_t i = 0; i < ts.size(); ++i)
    if (ts[i].shape != net.shape_of(net.inputs[i]))
      throw Error(cat("tensor ", i, " shape ", vec_str(ts[i].shape), " does not match legs"));
}

// ---------------------------------------------------------------- GEMM ------
// src/contract.rs:49-100 (geometry) and :102-181 (row-major batched GEMM,
// alpha = 1, beta = 0). Accumulation is sequential in k for every C element.
inline void gemm_geometry(size_t bt, size_t m, size_t k, size_t n) {
  const char* names[4] = {"batch", "m", "k", "n"};
  size_t v[4] = {bt, m, k, n};
  for (int i = 0; i < 4; ++i) {
    if (v[i] == 0) throw Error(cat("GEMM dimension ", names[i], "=0"));
    if (v[i] > (size_t)PTRDIFF_MAX) throw Error(cat("GEMM dimension ", names[i], " exceeds isize"));
  }
  auto mul = [](size_t a, size_t b) {
    if (b && a > (size_t)PTRDIFF_MAX / b) throw Error("GEMM element offset exceeds isize");
    return a * b;
  };
  mul(bt, mul(m, k));
  mul(bt, mul(k, n));
  mul(bt, mul(m, n));
}

template <class T>
std::vector<T> batch_gemm(const std::vector<T>& a, const std::vector<T>& b, size_t bt, size_t m, size_t k, size_t n) {
  gemm_geometry(bt, m, k, n);
  if (a.size() != bt * m * k || b.size() != bt * k * n) throw Error("GEMM operand length mismatch");
  std::vector<T> c(bt * m * n, Num<T>::zero());
  // Column blocks of JB elements keep a block of C's row in L1 while p runs; every
  // C[i,j] still accumulates A[i,p]*B[p,j] for p = 0..k-1 in order, from zero.
  constexpr size_t JB = 4096 / sizeof(T);
  for (size_t ib = 0; ib < bt; ++ib) {
    const T* A = a.data() + ib * m * k;
    const T* B = b.data() + ib * k * n;
    T* C = c.data() + ib * m * n;
    for (size_t j0 = 0; j0 < n; j0 += JB) {
      const size_t j1 = std::min(n, j0 + JB);
      for (size_t i = 0; i < m; ++i) {
        T* Ci = C + i * n;
        for (size_t p = 0; p < k; ++p) {
          const T aip = A[i * k + p];
          const T* Bp = B + p * n;
          for (size_t j = j0; j < j1; ++j) Ci[j] = add(Ci[j], Num<T>::mul(aip, Bp[j]));
        }
      }
    }
  }
  return c;
}

// ---------------------------------------------------------------- compile ---
struct InputPrep {            // src/compiled.rs:34-42
  std::vector<size_t> input_shape;
  std::vector<std::tuple<size_t, size_t, bool>> traces;
  std::vector<Leg> legs;
  std::vector<size_t> shape;
};
struct PairStep {             // src/compiled.rs:79-100
  size_t li, ri, out;
  std::vector<Leg> left_legs, right_legs, result_legs;
  std::vector<size_t> left_shape, right_shape;
  std::vector<size_t> sum_a, sum_b, a_perm, b_perm;
  size_t bt, m, k, n;
  std::vector<size_t> result_shape;
};
struct Finalize {             // src/compiled.rs:176-185
  size_t slot;
  std::vector<Leg> input_legs;
  std::vector<size_t> input_shape, sum, perm, output_shape;
};
struct Compiled {             // src/compiled.rs:223-231
  size_t n_inputs;
  std::vector<InputPrep> preps;
  std::vector<PairStep> steps;
  Finalize fin;
  double log10_flops, log2_peak_size;
};

struct Stats {
  double log10_flops, log2_max_size, log2_max_contraction_size, log2_total_size, log2_read_write, log2_peak_size;
};

inline double logaddexp2(double a, double b) {  // src/path.rs:32-42
  if (a == -INFINITY) return b;
  if (b == -INFINITY) return a;
  double hi = a > b ? a : b, lo = a > b ? b : a;
  return hi + std::log2(1.0 + std::exp2(lo - hi));
}

// src/path.rs:239-434
inline Stats simulate_path(const Network& net, const Path& path, std::vector<std::vector<Leg>>* step_legs_out = nullptr) {
  net.validate();
  size_t n = net.n_tensors();
  {
    auto od = net.output;
    std::sort(od.begin(), od.end());
    if (std::adjacent_find(od.begin(), od.end()) != od.end()) throw Error("output has repeated legs");
  }
  std::set<Leg> in_output(net.output.begin(), net.output.end());
  std::map<Leg, size_t> appearances, refcount;
  for (auto& t : net.inputs)
    for (Leg l : t) appearances[l]++;
  for (Leg l : net.output) appearances[l]++;
  std::vector<std::vector<Leg>> legs;
  std::vector<bool> live;
  std::vector<double> cslot, rwslot, slot_size;
  for (auto& t : net.inputs) {
    auto d = t;
    std::sort(d.begin(), d.end());
    double leaf = 0;
    for (size_t i = 0; i < d.size();) {
      size_t j = i + 1;
      while (j < d.size() && d[j] == d[i]) ++j;
      if (j - i != appearances[d[i]]) leaf += net.log2_dim(d[i]);
      i = j;
    }
    d = sorted_dedup(d);
    for (Leg l : d) refcount[l]++;
    legs.push_back(d);
    live.push_back(true);
    cslot.push_back(leaf);
    double rw = 0;
    for (Leg l : t) rw += net.log2_dim(l);
    rwslot.push_back(rw);
    slot_size.push_back(std::exp2(rw));
  }
  double alive_sum = 0;
  for (double s : slot_size) alive_sum += s;
  double peak = alive_sum;
  double flops = -INFINITY, maxsize = 0, maxc = 0, total = -INFINITY, rwt = -INFINITY;
  if (n == 1) {
    maxsize = 0;
    for (Leg l : net.output) maxsize += net.log2_dim(l);
  }
  for (size_t s = 0; s < path.size(); ++s) {
    size_t a = path[s].first, b = path[s].second;
    if (a == b) throw Error(cat("step ", s, " contracts a tensor with itself"));
    if (a >= legs.size() || b >= legs.size() || !live[a] || !live[b])
      throw Error(cat("step ", s, " references an invalid or consumed tensor"));
    live[a] = live[b] = false;
    std::vector<Leg> uni;
    std::set_union(legs[a].begin(), legs[a].end(), legs[b].begin(), legs[b].end(), std::back_inserter(uni));
    double sl = 0;
    for (Leg l : uni) sl += net.log2_dim(l);
    flops = logaddexp2(flops, sl);
    for (Leg l : legs[a]) refcount[l]--;
    for (Leg l : legs[b]) refcount[l]--;
    std::vector<Leg> res;
    for (Leg l : uni)
      if (refcount[l] > 0 || in_output.count(l)) res.push_back(l);
    for (Leg l : res) refcount[l]++;
    double sz = 0;
    for (Leg l : res) sz += net.log2_dim(l);
    maxsize = std::max(maxsize, sz);
    maxc = std::max(maxc, logaddexp2(logaddexp2(cslot[a], cslot[b]), sz));
    rwt = logaddexp2(rwt, logaddexp2(logaddexp2(rwslot[a], rwslot[b]), sz));
    cslot.push_back(sz);
    rwslot.push_back(sz);
    total = logaddexp2(total, sz);
    double rsize = std::exp2(sz);
    peak = std::max(peak, alive_sum + rsize);
    alive_sum += rsize - slot_size[a] - slot_size[b];
    slot_size.push_back(rsize);
    if (step_legs_out) step_legs_out->push_back(res);
    legs.push_back(res);
    live.push_back(true);
  }
  size_t n_alive = 0, last = 0;
  for (size_t i = 0; i < live.size(); ++i)
    if (live[i]) ++n_alive, last = i;
  if (n_alive != 1) throw Error(cat("incomplete path: ", n_alive, " tensors remain"));
  auto out = sorted_dedup(net.output);
  if (path.empty()) {
    for (Leg l : out)
      if (!std::binary_search(legs[last].begin(), legs[last].end(), l)) throw Error("final legs miss an output leg");
  } else if (legs[last] != out) {
    throw Error("final legs do not match output");
  }
  return Stats{flops * std::log10(2.0), maxsize, maxc, total, rwt, std::log2(peak)};
}

// src/compiled.rs:445-478
inline std::vector<std::tuple<size_t, size_t, bool>> plan_traces(std::vector<Leg>& legs, const std::map<Leg, size_t>& refcount,
                                                                 const std::set<Leg>& in_output) {
  std::vector<std::tuple<size_t, size_t, bool>> ops;
  for (;;) {
    bool found = false;
    size_t di = 0, dj = 0;
    for (size_t i = 0; i < legs.size() && !found; ++i)
      for (size_t j = i + 1; j < legs.size(); ++j)
        if (legs[i] == legs[j]) {
          di = i, dj = j, found = true;
          break;
        }
    if (!found) return ops;
    Leg l = legs[di];
    if (std::count(legs.begin(), legs.end(), l) > 2) throw Error(cat("leg ", l, " appears >2 times in one tensor"));
    bool external = refcount.at(l) > 1 || in_output.count(l);
    if (external) {
      ops.emplace_back(di, dj, true);
      legs.erase(legs.begin() + dj);
    } else {
      ops.emplace_back(di, dj, false);
      legs.erase(legs.begin() + dj);
      legs.erase(legs.begin() + di);
    }
  }
}

inline size_t pos_of(const std::vector<Leg>& v, Leg l) { return (size_t)(std::find(v.begin(), v.end(), l) - v.begin()); }

// src/compiled.rs:481-602
inline std::pair<std::vector<Leg>, PairStep> plan_pairwise(const Network& net, const std::vector<Leg>& la, const std::vector<Leg>& lb,
                                                           size_t li, size_t ri, size_t out,
                                                           const std::map<Leg, size_t>& refcount, const std::set<Leg>& in_output) {
  auto keep = [&](Leg l) { return refcount.at(l) > 0 || in_output.count(l); };
  std::set<Leg> set_a(la.begin(), la.end()), set_b(lb.begin(), lb.end());
  std::vector<Leg> batch, con, free_a, sum_a_legs, free_b, sum_b_legs;
  for (Leg l : la) {
    if (set_b.count(l)) (keep(l) ? batch : con).push_back(l);
    else (keep(l) ? free_a : sum_a_legs).push_back(l);
  }
  for (Leg l : lb)
    if (!set_a.count(l)) (keep(l) ? free_b : sum_b_legs).push_back(l);
  PairStep st;
  st.li = li, st.ri = ri, st.out = out;
  std::vector<Leg> wa = la, wb = lb;
  for (Leg l : sum_a_legs) {
    size_t ax = pos_of(wa, l);
    st.sum_a.push_back(ax);
    wa.erase(wa.begin() + ax);
  }
  for (Leg l : sum_b_legs) {
    size_t ax = pos_of(wb, l);
    st.sum_b.push_back(ax);
    wb.erase(wb.begin() + ax);
  }
  for (auto* g : {&batch, &free_a, &con})
    for (Leg l : *g) st.a_perm.push_back(pos_of(wa, l));
  for (auto* g : {&batch, &con, &free_b})
    for (Leg l : *g) st.b_perm.push_back(pos_of(wb, l));
  auto dprod = [&](const std::vector<Leg>& v) {
    size_t p = 1;
    for (Leg l : v) {
      size_t d = net.dim(l);
      if (p > SIZE_MAX / d) throw Error("compiled GEMM dimension product overflows");
      p *= d;
    }
    return p;
  };
  st.bt = dprod(batch), st.m = dprod(free_a), st.k = dprod(con), st.n = dprod(free_b);
  gemm_geometry(st.bt, st.m, st.k, st.n);
  for (auto* g : {&batch, &free_a, &free_b})
    for (Leg l : *g) st.result_legs.push_back(l);
  st.result_shape = net.shape_of(st.result_legs);
  st.left_legs = la, st.right_legs = lb;
  st.left_shape = net.shape_of(la), st.right_shape = net.shape_of(lb);
  return {st.result_legs, st};
}

// src/compiled.rs:254-366
inline Compiled compile(const Network& net, const Path& path) {
  net.validate();
  {
    auto od = net.output;
    std::sort(od.begin(), od.end());
    if (std::adjacent_find(od.begin(), od.end()) != od.end()) throw Error("output has repeated legs");
  }
  Stats stats = simulate_path(net, path);
  std::set<Leg> in_output(net.output.begin(), net.output.end());
  std::map<Leg, size_t> refcount;
  for (auto& t : net.inputs)
    for (Leg l : sorted_dedup(t)) refcount[l]++;
  Compiled c;
  c.n_inputs = net.n_tensors();
  std::vector<std::vector<Leg>> slots;
  std::vector<bool> live;
  for (auto& legs0 : net.inputs) {
    auto legs = legs0;
    InputPrep p;
    p.traces = plan_traces(legs, refcount, in_output);
    p.input_shape = net.shape_of(legs0);
    p.shape = net.shape_of(legs);
    p.legs = legs;
    c.preps.push_back(p);
    slots.push_back(legs);
    live.push_back(true);
  }
  for (size_t s = 0; s < path.size(); ++s) {
    size_t ia = path[s].first, ib = path[s].second;
    if (ia == ib) throw Error(cat("step ", s, " self contraction"));
    if (ia >= slots.size() || !live[ia]) throw Error(cat("step ", s, " references invalid tensor ", ia));
    if (ib >= slots.size() || !live[ib]) throw Error(cat("step ", s, " references invalid tensor ", ib));
    live[ia] = live[ib] = false;
    auto la = slots[ia], lb = slots[ib];
    for (Leg l : la) refcount[l]--;
    for (Leg l : lb) refcount[l]--;
    auto pr = plan_pairwise(net, la, lb, ia, ib, slots.size(), refcount, in_output);
    for (Leg l : pr.first) refcount[l]++;
    c.steps.push_back(pr.second);
    slots.push_back(pr.first);
    live.push_back(true);
  }
  size_t n_alive = 0, last = 0;
  for (size_t i = 0; i < live.size(); ++i)
    if (live[i]) ++n_alive, last = i;
  if (n_alive != 1) throw Error(cat("incomplete path: ", n_alive, " tensors remain"));
  auto legs = slots[last];
  c.fin.slot = last;
  c.fin.input_legs = legs;
  c.fin.input_shape = net.shape_of(legs);
  for (size_t i = 0; i < legs.size();) {
    if (!in_output.count(legs[i])) {
      c.fin.sum.push_back(i);
      legs.erase(legs.begin() + i);
    } else {
      ++i;
    }
  }
  if (sorted_dedup(legs) != sorted_dedup(net.output)) throw Error("final legs do not match output");
  for (Leg l : net.output) c.fin.perm.push_back(pos_of(legs, l));
  c.fin.output_shape = net.shape_of(net.output);
  c.log10_flops = stats.log10_flops;
  c.log2_peak_size = stats.log2_peak_size;
  return c;
}

// ---------------------------------------------------------------- execute ---
template <class T>
Tensor<T> apply_prep(const InputPrep& p, Tensor<T> t) {  // src/compiled.rs:65-74
  for (auto& op : p.traces) {
    size_t i = std::get<0>(op), j = std::get<1>(op);
    t = std::get<2>(op) ? t.diag_pair(i, j) : t.trace_pair(i, j);
  }
  return t;
}

template <class T>
Tensor<T> apply_step(const PairStep& st, const Tensor<T>& a0, const Tensor<T>& b0) {  // src/compiled.rs:143-171
  if (a0.shape != st.left_shape) throw Error("left operand shape does not match the compiled plan");
  if (b0.shape != st.right_shape) throw Error("right operand shape does not match the compiled plan");
  Tensor<T> a = a0, b = b0;
  for (size_t ax : st.sum_a) a = a.sum_axis(ax);
  for (size_t ax : st.sum_b) b = b.sum_axis(ax);
  a = a.permute(st.a_perm);
  b = b.permute(st.b_perm);
  return Tensor<T>(st.result_shape, batch_gemm(a.data, b.data, st.bt, st.m, st.k, st.n));
}

template <class T>
Tensor<T> apply_finalize(const Finalize& f, Tensor<T> t) {  // src/compiled.rs:211-218
  for (size_t ax : f.sum) t = t.sum_axis(ax);
  return t.permute(f.perm);
}

// src/compiled.rs:383-441. With strip, each step result is divided by its
// largest magnitude and log10 of that magnitude is accumulated.
template <class T>
Tensor<T> execute(const Compiled& c, const std::vector<Tensor<T>>& ts, bool strip, double* exponent) {
  if (ts.size() != c.n_inputs) throw Error(cat("tensor count ", ts.size(), " does not match plan ", c.n_inputs));
  for (size_t i = 0; i < ts.size(); ++i)
    if (ts[i].shape != c.preps[i].input_shape) throw Error(cat("tensor ", i, " shape does not match the compiled plan"));
  std::vector<Tensor<T>> slots;
  std::vector<bool> live;
  for (size_t i = 0; i < ts.size(); ++i) slots.push_back(apply_prep(c.preps[i], ts[i])), live.push_back(true);
  double e = 0.0;
  for (auto& st : c.steps) {
    if (!live[st.li] || !live[st.ri]) throw Error("plan references a consumed tensor");
    live[st.li] = live[st.ri] = false;
    Tensor<T> r = apply_step(st, slots[st.li], slots[st.ri]);
    slots[st.li] = Tensor<T>();
    slots[st.ri] = Tensor<T>();
    if (strip) {
      double fmax = 0.0;
      for (auto& v : r.data) fmax = std::max(fmax, Num<T>::abs(v));
      if (fmax > 0.0 && std::isfinite(fmax)) {
        e += std::log10(fmax);
        T inv = Num<T>::from_real(1.0 / fmax);
        for (auto& v : r.data) v = Num<T>::mul(v, inv);
      }
    }
    slots.push_back(std::move(r));
    live.push_back(true);
  }
  if (exponent) *exponent = e;
  return apply_finalize(c.fin, slots[c.fin.slot]);
}

// src/contract.rs:328-417: the direct executor that re-derives every step.
template <class T>
Tensor<T> contract_network(const Network& net, std::vector<Tensor<T>> ts, const Path& path) {
  validate_shapes(net, ts);
  {
    auto od = net.output;
    std::sort(od.begin(), od.end());
    if (std::adjacent_find(od.begin(), od.end()) != od.end()) throw Error("output has repeated legs");
  }
  std::set<Leg> in_output(net.output.begin(), net.output.end());
  std::map<Leg, size_t> refcount;
  for (auto& t : net.inputs)
    for (Leg l : sorted_dedup(t)) refcount[l]++;
  std::vector<std::vector<Leg>> slegs;
  std::vector<Tensor<T>> sts;
  std::vector<bool> live;
  for (size_t i = 0; i < ts.size(); ++i) {  // resolve_traces, contract.rs:187-223
    auto legs = net.inputs[i];
    auto ops = plan_traces(legs, refcount, in_output);
    Tensor<T> t = std::move(ts[i]);
    for (auto& op : ops) t = std::get<2>(op) ? t.diag_pair(std::get<0>(op), std::get<1>(op)) : t.trace_pair(std::get<0>(op), std::get<1>(op));
    slegs.push_back(legs);
    sts.push_back(std::move(t));
    live.push_back(true);
  }
  for (size_t s = 0; s < path.size(); ++s) {
    size_t ia = path[s].first, ib = path[s].second;
    if (ia == ib) throw Error(cat("step ", s, " self contraction"));
    if (ia >= sts.size() || !live[ia] || ib >= sts.size() || !live[ib]) throw Error(cat("step ", s, " references invalid tensor"));
    live[ia] = live[ib] = false;
    auto la = slegs[ia], lb = slegs[ib];
    for (Leg l : la) refcount[l]--;
    for (Leg l : lb) refcount[l]--;
    // pairwise (contract.rs:226-325): classify legs, sum, permute, GEMM.
    auto keep = [&](Leg l) { return refcount.at(l) > 0 || in_output.count(l); };
    std::set<Leg> set_a(la.begin(), la.end()), set_b(lb.begin(), lb.end());
    std::vector<Leg> batch, con, free_a, sum_a, free_b, sum_b;
    for (Leg l : la) {
      if (set_b.count(l)) (keep(l) ? batch : con).push_back(l);
      else (keep(l) ? free_a : sum_a).push_back(l);
    }
    for (Leg l : lb)
      if (!set_a.count(l)) (keep(l) ? free_b : sum_b).push_back(l);
    Tensor<T> a = std::move(sts[ia]), b = std::move(sts[ib]);
    for (Leg l : sum_a) {
      size_t ax = pos_of(la, l);
      a = a.sum_axis(ax);
      la.erase(la.begin() + ax);
    }
    for (Leg l : sum_b) {
      size_t ax = pos_of(lb, l);
      b = b.sum_axis(ax);
      lb.erase(lb.begin() + ax);
    }
    std::vector<size_t> ap, bp;
    for (auto* g : {&batch, &free_a, &con})
      for (Leg l : *g) ap.push_back(pos_of(la, l));
    for (auto* g : {&batch, &con, &free_b})
      for (Leg l : *g) bp.push_back(pos_of(lb, l));
    a = a.permute(ap);
    b = b.permute(bp);
    auto dp = [&](const std::vector<Leg>& v) {
      size_t p = 1;
      for (Leg l : v) p *= net.dim(l);
      return p;
    };
    std::vector<Leg> rl;
    for (auto* g : {&batch, &free_a, &free_b})
      for (Leg l : *g) rl.push_back(l);
    Tensor<T> r(net.shape_of(rl), batch_gemm(a.data, b.data, dp(batch), dp(free_a), dp(con), dp(free_b)));
    for (Leg l : rl) refcount[l]++;
    slegs.push_back(rl);
    sts.push_back(std::move(r));
    live.push_back(true);
  }
  size_t n_alive = 0, last = 0;
  for (size_t i = 0; i < live.size(); ++i)
    if (live[i]) ++n_alive, last = i;
  if (n_alive != 1) throw Error(cat("incomplete path: ", n_alive, " tensors remain"));
  auto legs = slegs[last];
  Tensor<T> t = std::move(sts[last]);
  for (size_t i = 0; i < legs.size();) {
    if (!in_output.count(legs[i])) {
      t = t.sum_axis(i);
      legs.erase(legs.begin() + i);
    } else {
      ++i;
    }
  }
  if (sorted_dedup(legs) != sorted_dedup(net.output)) throw Error("final legs do not match output");
  std::vector<size_t> axes;
  for (Leg l : net.output) axes.push_back(pos_of(legs, l));
  return t.permute(axes);
}

// ---------------------------------------------------------------- slicing ---
constexpr size_t SLICE_MAX_CHUNKS = 256;          // src/slice.rs:921
constexpr size_t SLICE_PARTIAL_BUDGET = 1u << 22;  // src/slice.rs:924

// src/slice.rs:893-905
inline size_t slice_parallel_chunks(size_t n_slices, size_t out_numel) {
  if (n_slices <= 1) return 1;
  size_t budget = std::max<size_t>(SLICE_PARTIAL_BUDGET / std::max<size_t>(out_numel, 1), 1);
  size_t n0 = std::max<size_t>(std::min(std::min(n_slices, SLICE_MAX_CHUNKS), budget), 1);
  if (n0 <= 1) return 1;
  size_t chunk = (n_slices + n0 - 1) / n0;
  return (n_slices + chunk - 1) / chunk;
}

// src/slice.rs:881-888 (last leg fastest)
inline void decode_slice_index(size_t s, const std::vector<size_t>& dims, std::vector<size_t>& idx) {
  for (size_t d = dims.size(); d-- > 0;) {
    idx[d] = s % dims[d];
    s /= dims[d];
  }
}

// src/slice.rs:146-172
inline void validate_slice_legs(const Network& net, const std::vector<Leg>& sliced) {
  std::set<Leg> seen;
  for (Leg l : sliced) {
    if (!seen.insert(l).second) throw Error(cat("sliced leg ", l, " repeated"));
    auto it = net.size_dict.find(l);
    if (it == net.size_dict.end()) throw Error(cat("sliced leg ", l, " not in size_dict"));
    if (it->second == 0) throw Error(cat("sliced leg ", l, " has dimension 0"));
    bool held = false;
    for (auto& t : net.inputs) held = held || std::count(t.begin(), t.end(), l) > 0;
    if (!held) throw Error(cat("sliced leg ", l, " is not held by any input tensor"));
    for (auto& t : net.inputs)
      if (std::count(t.begin(), t.end(), l) > 1) throw Error(cat("sliced leg ", l, " repeated within one input tensor"));
  }
}

template <class T>
Tensor<T> contract_sliced(const Network& net, const std::vector<Tensor<T>>& ts, const Path& path, const std::vector<Leg>& sliced, int threads);

// src/slice.rs:1070-1173: output-index slices are disjoint blocks.
template <class T>
Tensor<T> contract_output_blocks(const Network& net, const std::vector<Tensor<T>>& ts, const Path& path, const std::vector<Leg>& sliced, int threads) {
  std::vector<Leg> outer, internal;
  for (Leg l : net.output)
    if (std::count(sliced.begin(), sliced.end(), l)) outer.push_back(l);
  for (Leg l : sliced)
    if (!std::count(outer.begin(), outer.end(), l)) internal.push_back(l);
  auto shape = net.shape_of(net.output);
  size_t numel = checked_numel(shape);
  if (numel > (size_t)PTRDIFF_MAX / sizeof(T)) throw Error("full output tensor bytes exceed allocatable capacity");
  Tensor<T> result(shape, std::vector<T>(numel, Num<T>::zero()));
  std::vector<size_t> strides(shape.size(), 1);
  for (size_t ax = shape.size(); ax-- > 1;) strides[ax - 1] = strides[ax] * shape[ax];
  std::vector<size_t> outer_strides(outer.size());
  std::vector<std::pair<size_t, size_t>> kept;
  for (size_t ax = 0; ax < net.output.size(); ++ax) {
    size_t p = pos_of(outer, net.output[ax]);
    if (p < outer.size()) outer_strides[p] = strides[ax];
    else kept.push_back({shape[ax], strides[ax]});
  }
  Network sub = net;
  for (auto& legs : sub.inputs) legs.erase(std::remove_if(legs.begin(), legs.end(), [&](Leg l) { return pos_of(outer, l) < outer.size(); }), legs.end());
  sub.output.erase(std::remove_if(sub.output.begin(), sub.output.end(), [&](Leg l) { return pos_of(outer, l) < outer.size(); }), sub.output.end());
  for (Leg l : outer) sub.size_dict.erase(l);
  auto dims = net.shape_of(outer);
  size_t n_blocks = checked_numel(dims);
  std::vector<size_t> vals(outer.size());
  for (size_t bnum = 0; bnum < n_blocks; ++bnum) {
    decode_slice_index(bnum, dims, vals);
    std::vector<Tensor<T>> bts;
    for (size_t i = 0; i < ts.size(); ++i) {
      Tensor<T> t = ts[i];
      for (size_t ax = net.inputs[i].size(); ax-- > 0;) {
        size_t p = pos_of(outer, net.inputs[i][ax]);
        if (p < outer.size()) t = t.select_axis(ax, vals[p]);
      }
      bts.push_back(std::move(t));
    }
    Tensor<T> block = contract_sliced(sub, bts, path, internal, threads);
    size_t base = 0;
    for (size_t p = 0; p < vals.size(); ++p) base += vals[p] * outer_strides[p];
    for (size_t off = 0; off < block.data.size(); ++off) {
      size_t rem = off, o = base;
      for (size_t q = kept.size(); q-- > 0;) {
        o += (rem % kept[q].first) * kept[q].second;
        rem /= kept[q].first;
      }
      result.data[o] = block.data[off];
    }
  }
  return result;
}

// src/slice.rs:960-1065: internal slices summed in slice order inside
// contiguous chunks, chunk partials merged in chunk order. `threads` only
// changes how chunks are scheduled, never the summation order.
template <class T>
Tensor<T> contract_sliced(const Network& net, const std::vector<Tensor<T>>& ts, const Path& path, const std::vector<Leg>& sliced, int threads) {
  validate_shapes(net, ts);
  validate_slice_legs(net, sliced);
  if (sliced.empty()) return contract_network(net, ts, path);
  for (Leg l : sliced)
    if (std::count(net.output.begin(), net.output.end(), l)) return contract_output_blocks(net, ts, path, sliced, threads);
  std::map<Leg, size_t> spos;
  for (size_t i = 0; i < sliced.size(); ++i) spos[sliced[i]] = i;
  Network sub = net;
  for (auto& legs : sub.inputs) legs.erase(std::remove_if(legs.begin(), legs.end(), [&](Leg l) { return spos.count(l) > 0; }), legs.end());
  for (Leg l : sliced) sub.size_dict.erase(l);
  auto dims = net.shape_of(sliced);
  size_t n_slices = 1;
  for (size_t d : dims) {
    if (n_slices > SIZE_MAX / d) throw Error("slice count overflows");
    n_slices *= d;
  }
  auto run_one = [&](size_t s) {
    std::vector<size_t> idx(sliced.size());
    decode_slice_index(s, dims, idx);
    std::vector<Tensor<T>> st;
    for (size_t i = 0; i < ts.size(); ++i) {
      Tensor<T> t = ts[i];
      for (size_t ax = net.inputs[i].size(); ax-- > 0;) {
        auto it = spos.find(net.inputs[i][ax]);
        if (it != spos.end()) t = t.select_axis(ax, idx[it->second]);
      }
      st.push_back(std::move(t));
    }
    return contract_network(sub, std::move(st), path);
  };
  size_t out_numel = 1;
  for (Leg l : net.output) {
    size_t d = net.dim(l);
    out_numel = (out_numel > SIZE_MAX / d) ? SIZE_MAX : out_numel * d;
  }
  size_t n_chunks = slice_parallel_chunks(n_slices, std::max<size_t>(out_numel, 1));
  size_t chunk = (n_slices + n_chunks - 1) / n_chunks;
  std::vector<Tensor<T>> partial(n_chunks);
  std::vector<std::string> errs(n_chunks);
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads > 0 ? threads : 1) if (n_chunks > 1 && threads > 1)
  for (long c = 0; c < (long)n_chunks; ++c) {
    try {
      size_t lo = c * chunk, hi = std::min((c + 1) * chunk, n_slices);
      Tensor<T> acc = run_one(lo);
      for (size_t s = lo + 1; s < hi; ++s) {
        Tensor<T> part = run_one(s);
        for (size_t e = 0; e < acc.data.size(); ++e) acc.data[e] = add(acc.data[e], part.data[e]);
      }
      partial[c] = std::move(acc);
    } catch (const std::exception& ex) {
      errs[c] = ex.what();
    }
  }
  for (auto& e : errs)
    if (!e.empty()) throw Error(e);
  Tensor<T> acc = std::move(partial[0]);
  for (size_t c = 1; c < n_chunks; ++c)
    for (size_t e = 0; e < acc.data.size(); ++e) acc.data[e] = add(acc.data[e], partial[c].data[e]);
  return acc;
}

// Bounded sample of the sliced loop: internal slices [lo, hi) only, summed with
// the reference's chunking applied to the sample (src/slice.rs:1018-1064).
template <class T>
Tensor<T> contract_slice_range(const Network& net, const std::vector<Tensor<T>>& ts, const Path& path,
                               const std::vector<Leg>& sliced, size_t lo, size_t hi, int threads) {
  validate_shapes(net, ts);
  validate_slice_legs(net, sliced);
  for (Leg l : sliced)
    if (std::count(net.output.begin(), net.output.end(), l)) throw Error("slice range supports internal slices only");
  std::map<Leg, size_t> spos;
  for (size_t i = 0; i < sliced.size(); ++i) spos[sliced[i]] = i;
  Network sub = net;
  for (auto& legs : sub.inputs) legs.erase(std::remove_if(legs.begin(), legs.end(), [&](Leg l) { return spos.count(l) > 0; }), legs.end());
  for (Leg l : sliced) sub.size_dict.erase(l);
  auto dims = net.shape_of(sliced);
  size_t n_slices = 1;
  for (size_t d : dims) n_slices *= d;
  if (!(lo < hi && hi <= n_slices)) throw Error("slice range out of bounds");
  auto run_one = [&](size_t s) {
    std::vector<size_t> idx(sliced.size());
    decode_slice_index(s, dims, idx);
    std::vector<Tensor<T>> st;
    for (size_t i = 0; i < ts.size(); ++i) {
      Tensor<T> t = ts[i];
      for (size_t ax = net.inputs[i].size(); ax-- > 0;) {
        auto it = spos.find(net.inputs[i][ax]);
        if (it != spos.end()) t = t.select_axis(ax, idx[it->second]);
      }
      st.push_back(std::move(t));
    }
    return contract_network(sub, std::move(st), path);
  };
  size_t n = hi - lo;
  size_t out_numel = 1;
  for (Leg l : net.output) out_numel *= net.dim(l);
  size_t n_chunks = slice_parallel_chunks(n, out_numel), chunk = (n + n_chunks - 1) / n_chunks;
  std::vector<Tensor<T>> partial(n_chunks);
  std::vector<std::string> errs(n_chunks);
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads > 0 ? threads : 1) if (n_chunks > 1 && threads > 1)
  for (long c = 0; c < (long)n_chunks; ++c) {
    try {
      size_t a = lo + c * chunk, b = std::min(lo + (c + 1) * chunk, hi);
      Tensor<T> acc = run_one(a);
      for (size_t s = a + 1; s < b; ++s) {
        Tensor<T> part = run_one(s);
        for (size_t e = 0; e < acc.data.size(); ++e) acc.data[e] = add(acc.data[e], part.data[e]);
      }
      partial[c] = std::move(acc);
    } catch (const std::exception& ex) {
      errs[c] = ex.what();
    }
  }
  for (auto& e : errs)
    if (!e.empty()) throw Error(e);
  Tensor<T> acc = std::move(partial[0]);
  for (size_t c = 1; c < n_chunks; ++c)
    for (size_t e = 0; e < acc.data.size(); ++e) acc.data[e] = add(acc.data[e], partial[c].data[e]);
  return acc;
}

// ---------------------------------------------------------------- naive -----
// src/naive.rs:10-76: brute force over the joint index space (<= 2^28).
template <class T>
Tensor<T> naive_einsum(const Network& net, const std::vector<Tensor<T>>& ts) {
  validate_shapes(net, ts);
  std::vector<Leg> all;
  for (auto& t : net.inputs) all.insert(all.end(), t.begin(), t.end());
  all.insert(all.end(), net.output.begin(), net.output.end());
  all = sorted_dedup(all);
  double space = 0;
  for (Leg l : all) space += net.log2_dim(l);
  if (space > 28.0) throw Error("joint index space too large for the naive oracle");
  std::vector<size_t> dims = net.shape_of(all);
  std::vector<std::vector<std::pair<size_t, size_t>>> maps;
  for (size_t i = 0; i < ts.size(); ++i) {
    auto st = ts[i].strides();
    std::vector<std::pair<size_t, size_t>> m;
    for (size_t a = 0; a < net.inputs[i].size(); ++a) m.push_back({pos_of(all, net.inputs[i][a]), st[a]});
    maps.push_back(m);
  }
  auto oshape = net.shape_of(net.output);
  Tensor<T> out(oshape, std::vector<T>(checked_numel(oshape), Num<T>::zero()));
  auto ost = out.strides();
  std::vector<std::pair<size_t, size_t>> omap;
  for (size_t a = 0; a < net.output.size(); ++a) omap.push_back({pos_of(all, net.output[a]), ost[a]});
  size_t total = 1;
  for (size_t d : dims) total *= d;
  std::vector<size_t> idx(all.size(), 0);
  for (size_t e = 0; e < total; ++e) {
    T term = Num<T>::one();
    for (size_t t = 0; t < maps.size(); ++t) {
      size_t off = 0;
      for (auto& ps : maps[t]) off += idx[ps.first] * ps.second;
      term = Num<T>::mul(term, ts[t].data[off]);
    }
    size_t oo = 0;
    for (auto& ps : omap) oo += idx[ps.first] * ps.second;
    out.data[oo] = add(out.data[oo], term);
    for (size_t d = idx.size(); d-- > 0;) {
      if (++idx[d] < dims[d]) break;
      idx[d] = 0;
    }
  }
  return out;
}

// ---------------------------------------------------------------- adjoint ---
// No reference counterpart (the reference exposes stage execution only,
// compiled.rs:7-10). Restates the textbook reverse pass of one compiled pair
// step with PyTorch's complex convention (gradients are dL/d(conj z)):
//   gA2 = gC . B2^H,  gB2 = A2^H . gC   per batch
// then the inverse permutation, and a broadcast over every summed axis. Input
// traces broadcast onto the diagonal, diagonals scatter onto it.
template <class T>
Tensor<T> unsum_axis(const Tensor<T>& g, const std::vector<size_t>& full_shape, size_t ax) {
  // g has full_shape with axis `ax` removed; broadcast back along `ax`.
  Tensor<T> out(full_shape, std::vector<T>(checked_numel(full_shape)));
  auto st = out.strides();
  size_t outer = 1, inner = st[ax], d = full_shape[ax];
  for (size_t i = 0; i < ax; ++i) outer *= full_shape[i];
  for (size_t o = 0; o < outer; ++o)
    for (size_t i = 0; i < d; ++i)
      for (size_t q = 0; q < inner; ++q) out.data[(o * d + i) * inner + q] = g.data[o * inner + q];
  return out;
}

template <class T>
std::vector<size_t> inverse_perm(const std::vector<size_t>& p) {
  std::vector<size_t> inv(p.size());
  for (size_t i = 0; i < p.size(); ++i) inv[p[i]] = i;
  return inv;
}

template <class T>
void step_adjoint(const PairStep& st, const Tensor<T>& a0, const Tensor<T>& b0, const Tensor<T>& gc, Tensor<T>& ga, Tensor<T>& gb) {
  // Forward intermediates.
  std::vector<Tensor<T>> a_chain{a0}, b_chain{b0};
  for (size_t ax : st.sum_a) a_chain.push_back(a_chain.back().sum_axis(ax));
  for (size_t ax : st.sum_b) b_chain.push_back(b_chain.back().sum_axis(ax));
  Tensor<T> a2 = a_chain.back().permute(st.a_perm), b2 = b_chain.back().permute(st.b_perm);
  size_t bt = st.bt, m = st.m, k = st.k, n = st.n;
  std::vector<T> ga2(bt * m * k, Num<T>::zero()), gb2(bt * k * n, Num<T>::zero());
  for (size_t ib = 0; ib < bt; ++ib) {
    const T* A = a2.data.data() + ib * m * k;
    const T* B = b2.data.data() + ib * k * n;
    const T* G = gc.data.data() + ib * m * n;
    for (size_t i = 0; i < m; ++i)
      for (size_t p = 0; p < k; ++p) {
        T acc = Num<T>::zero();
        for (size_t j = 0; j < n; ++j) acc = add(acc, Num<T>::mul(G[i * n + j], Num<T>::conj(B[p * n + j])));
        ga2[ib * m * k + i * k + p] = acc;
      }
    for (size_t p = 0; p < k; ++p)
      for (size_t j = 0; j < n; ++j) {
        T acc = Num<T>::zero();
        for (size_t i = 0; i < m; ++i) acc = add(acc, Num<T>::mul(Num<T>::conj(A[i * k + p]), G[i * n + j]));
        gb2[ib * k * n + p * n + j] = acc;
      }
  }
  std::vector<size_t> a2shape, b2shape;
  for (size_t ax : st.a_perm) a2shape.push_back(a_chain.back().shape[ax]);
  for (size_t ax : st.b_perm) b2shape.push_back(b_chain.back().shape[ax]);
  Tensor<T> g = Tensor<T>(a2shape, ga2).permute(inverse_perm<T>(st.a_perm));
  for (size_t s = st.sum_a.size(); s-- > 0;) g = unsum_axis(g, a_chain[s].shape, st.sum_a[s]);
  ga = g;
  g = Tensor<T>(b2shape, gb2).permute(inverse_perm<T>(st.b_perm));
  for (size_t s = st.sum_b.size(); s-- > 0;) g = unsum_axis(g, b_chain[s].shape, st.sum_b[s]);
  gb = g;
}

template <class T>
Tensor<T> prep_adjoint(const InputPrep& p, const Tensor<T>& gout) {
  // Replay shapes forward, then walk the ops backwards.
  std::vector<std::vector<size_t>> shapes{p.input_shape};
  for (auto& op : p.traces) {
    auto s = shapes.back();
    size_t i = std::get<0>(op), j = std::get<1>(op);
    s.erase(s.begin() + j);
    if (!std::get<2>(op)) s.erase(s.begin() + i);
    shapes.push_back(s);
  }
  Tensor<T> g = gout;
  for (size_t q = p.traces.size(); q-- > 0;) {
    auto& full = shapes[q];
    size_t i = std::get<0>(p.traces[q]), j = std::get<1>(p.traces[q]);
    bool diag = std::get<2>(p.traces[q]);
    Tensor<T> out(full, std::vector<T>(checked_numel(full), Num<T>::zero()));
    auto st = out.strides();
    std::vector<size_t> idx(full.size(), 0);
    for (size_t e = 0; e < out.data.size(); ++e) {
      size_t rem = e;
      for (size_t d = full.size(); d-- > 0;) idx[d] = rem % full[d], rem /= full[d];
      if (idx[i] == idx[j]) {
        size_t go = 0, gs = 1;
        for (size_t d = full.size(); d-- > 0;) {
          if (d == j || (!diag && d == i)) continue;
          go += idx[d] * gs;
          gs *= full[d];
        }
        out.data[e] = g.data[go];
      }
    }
    g = out;
  }
  return g;
}

template <class T>
Tensor<T> finalize_adjoint(const Finalize& f, const Tensor<T>& gout) {
  std::vector<std::vector<size_t>> shapes{f.input_shape};
  for (size_t ax : f.sum) {
    auto s = shapes.back();
    s.erase(s.begin() + ax);
    shapes.push_back(s);
  }
  Tensor<T> g = gout.permute(inverse_perm<T>(f.perm));
  for (size_t q = f.sum.size(); q-- > 0;) g = unsum_axis(g, shapes[q], f.sum[q]);
  return g;
}

// Gradient of the compiled contraction w.r.t. every input, given dL/d(conj out).
template <class T>
std::vector<Tensor<T>> gradient(const Compiled& c, const std::vector<Tensor<T>>& ts, const Tensor<T>& gout) {
  std::vector<Tensor<T>> slots;
  for (size_t i = 0; i < ts.size(); ++i) slots.push_back(apply_prep(c.preps[i], ts[i]));
  for (auto& st : c.steps) slots.push_back(apply_step(st, slots[st.li], slots[st.ri]));
  std::vector<Tensor<T>> g(slots.size());
  std::vector<bool> has(slots.size(), false);
  g[c.fin.slot] = finalize_adjoint(c.fin, gout);
  has[c.fin.slot] = true;
  for (size_t s = c.steps.size(); s-- > 0;) {
    auto& st = c.steps[s];
    if (!has[st.out]) continue;
    Tensor<T> ga, gb;
    step_adjoint(st, slots[st.li], slots[st.ri], g[st.out], ga, gb);
    g[st.li] = ga, has[st.li] = true;
    g[st.ri] = gb, has[st.ri] = true;
  }
  std::vector<Tensor<T>> out;
  for (size_t i = 0; i < ts.size(); ++i) out.push_back(prep_adjoint(c.preps[i], g[i]));
  return out;
}

}  // namespace orc
