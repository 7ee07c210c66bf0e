// Contraction plan of the executor: the compiled per-slice plan plus the slice
// bookkeeping. Host-only integer work, done once per plan.
#pragma once

#include <map>
#include <set>
#include <tuple>

#include "common.hpp"

namespace atn {

struct Net {  // src/network.rs:135-143
  std::vector<std::vector<Leg>> inputs;
  std::vector<Leg> output;
  std::map<Leg, size_t> size_dict;

  size_t dim(Leg l) const {
    auto it = size_dict.find(l);
    if (it == size_dict.end()) throw Error(cat("leg ", l, " has no size_dict dimension"));
    return it->second;
  }
  std::vector<size_t> shape_of(const std::vector<Leg>& legs) const {
    std::vector<size_t> s;
    s.reserve(legs.size());
    for (Leg l : legs) s.push_back(dim(l));
    return s;
  }
  void validate() const;
};

struct InputPrep {  // src/compiled.rs:34-42
  std::vector<size_t> input_shape;
  std::vector<std::tuple<size_t, size_t, bool>> traces;  // (ax_i, ax_j, is_diag)
  std::vector<Leg> legs;
  std::vector<size_t> shape;
  bool identity() const { return traces.empty(); }
};

struct PairStep {  // src/compiled.rs:79-100
  size_t li, ri, out;
  std::vector<Leg> left_legs, right_legs, result_legs;
  std::vector<size_t> left_shape, right_shape, result_shape;
  std::vector<size_t> sum_a, sum_b, a_perm, b_perm;
  size_t bt, m, k, n;
};

struct Finalize {  // src/compiled.rs:176-185
  size_t slot;
  std::vector<Leg> input_legs;
  std::vector<size_t> input_shape, sum, perm, output_shape;
  bool identity() const {
    if (!sum.empty()) return false;
    for (size_t i = 0; i < perm.size(); ++i)
      if (perm[i] != i) return false;
    return true;
  }
};

struct Compiled {
  size_t n_inputs = 0;
  std::vector<InputPrep> preps;
  std::vector<PairStep> steps;
  Finalize fin;
};

struct PathStats {
  double log10_flops, log2_max_size, log2_peak_size;
};

PathStats simulate_path(const Net& net, const Path& path);
Compiled compile(const Net& net, const Path& path);

// Selection of one input's sliced axes for global slice number g:
// value = (g / div) % mod, offset += value * stride (full-tensor elements).
struct SliceSel {
  size_t div, mod, stride, axis;
};

struct Plan {
  Net net;            // the full network
  Path path;
  std::vector<Leg> sliced;
  std::vector<Leg> outer;     // sliced output legs, in net.output order (src/slice.rs:1082-1087)
  std::vector<Leg> internal;  // remaining sliced legs, in `sliced` order (src/slice.rs:1094-1098)
  Net sub;                    // every sliced leg removed from inputs, outer legs from output
  Compiled cp;                // compile(sub, path): the plan shared by every slice
  size_t n_blocks = 1, n_internal = 1, n_slices = 1;
  std::vector<std::vector<SliceSel>> sel;   // per input, empty if it holds no sliced leg
  std::vector<size_t> block_div, block_mod, block_stride;  // output base of block b: sum ((b/div)%mod)*stride
  std::vector<size_t> kept_dims, kept_strides;  // output axes not sliced, in output order
  std::vector<size_t> full_out_shape;
  size_t sub_out_numel = 1, full_out_numel = 1;
  double log10_flops = 0, log2_peak_size = 0;  // stats of the whole (unsliced) network
  double macs_per_slice = 0, mults_per_slice = 0;
  std::vector<bool> varying;  // per slot of cp: depends on the slice index
  size_t n_varying_steps = 0;
  size_t n_chunks = 1;        // reference chunking of internal slices (src/slice.rs:893-905)
};

Plan make_plan(const Net& net, const Path& path, const std::vector<Leg>& sliced);
std::string describe_json(const Plan& p);

size_t slice_parallel_chunks(size_t n_slices, size_t out_numel);

}  // namespace atn
