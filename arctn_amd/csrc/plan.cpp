// Plan compilation (host). Follows src/compiled.rs:254-602 for the per-slice
// plan and src/slice.rs:146-172, 960-1016, 1070-1137 for the slice bookkeeping.
#include "plan.hpp"

#include <algorithm>
#include <cmath>

namespace atn {

static std::vector<Leg> sorted_dedup(std::vector<Leg> v) {
  std::sort(v.begin(), v.end());
  v.erase(std::unique(v.begin(), v.end()), v.end());
  return v;
}

static size_t pos_of(const std::vector<Leg>& v, Leg l) { return (size_t)(std::find(v.begin(), v.end(), l) - v.begin()); }

// src/network.rs:171-229
void Net::validate() const {
  if (inputs.empty()) throw Error("TensorNetwork needs at least one input tensor");
  for (auto& kv : size_dict) {
    if (kv.second == 0) throw Error(cat("leg ", kv.first, " has dimension 0; zero-dimension networks are not supported"));
    if (kv.second > (size_t)PTRDIFF_MAX) throw Error(cat("leg ", kv.first, " dimension exceeds the isize stride range"));
  }
  std::set<Leg> held;
  for (size_t i = 0; i < inputs.size(); ++i) {
    std::map<Leg, size_t> counts;
    std::vector<size_t> shape;
    for (Leg l : inputs[i]) {
      shape.push_back(dim(l));
      counts[l]++;
      held.insert(l);
    }
    for (auto& kv : counts)
      if (kv.second > 2) throw Error(cat("leg ", kv.first, " appears ", kv.second, " times in input tensor ", i, "; at most 2 supported"));
    try {
      checked_numel(shape);
    } catch (const Error& e) {
      throw Error(cat("input tensor ", i, " shape ", vstr(shape), " invalid (元素数溢出): ", e.what()));
    }
  }
  std::set<Leg> seen;
  std::vector<size_t> oshape;
  for (Leg l : output) {
    if (!seen.insert(l).second) throw Error("output contains repeated legs, not supported");
    if (!held.count(l)) throw Error(cat("output leg ", l, " is not held by any input tensor"));
    oshape.push_back(dim(l));
  }
  checked_numel(oshape);
  for (auto& kv : size_dict)
    if (!held.count(kv.first)) throw Error(cat("size_dict leg ", kv.first, " is not used by any input tensor (未被任何输入张量使用)"));
}

// src/path.rs:239-434 (validation + the statistics the executor reports)
PathStats simulate_path(const Net& net, const Path& path) {
  net.validate();
  size_t n = net.inputs.size();
  std::set<Leg> in_output(net.output.begin(), net.output.end());
  std::map<Leg, size_t> refcount;
  std::vector<std::vector<Leg>> legs;
  std::vector<bool> live;
  std::vector<double> slot_size;
  auto log2d = [&](Leg l) { return std::log2((double)net.dim(l)); };
  for (auto& t : net.inputs) {
    auto d = sorted_dedup(t);
    for (Leg l : d) refcount[l]++;
    legs.push_back(d);
    live.push_back(true);
    double s = 0;
    for (Leg l : t) s += log2d(l);
    slot_size.push_back(std::exp2(s));
  }
  double alive = 0;
  for (double s : slot_size) alive += s;
  double peak = alive, flops = -INFINITY, maxsize = 0;
  if (n == 1)
    for (Leg l : net.output) maxsize += log2d(l);
  auto lae2 = [](double a, double b) {
    if (a == -INFINITY) return b;
    if (b == -INFINITY) return a;
    double hi = std::max(a, b), lo = std::min(a, b);
    return hi + std::log2(1.0 + std::exp2(lo - hi));
  };
  for (size_t s = 0; s < path.size(); ++s) {
    size_t a = path[s].first, b = path[s].second;
    if (a == b) throw Error(cat("step ", s, " contracts a tensor with itself (", a, ",", b, ")"));
    if (a >= legs.size() || b >= legs.size() || !live[a] || !live[b])
      throw Error(cat("step ", s, " references an invalid or consumed tensor"));
    live[a] = live[b] = false;
    std::vector<Leg> uni;
    std::set_union(legs[a].begin(), legs[a].end(), legs[b].begin(), legs[b].end(), std::back_inserter(uni));
    double sl = 0;
    for (Leg l : uni) sl += log2d(l);
    flops = lae2(flops, sl);
    for (Leg l : legs[a]) refcount[l]--;
    for (Leg l : legs[b]) refcount[l]--;
    std::vector<Leg> res;
    for (Leg l : uni)
      if (refcount[l] > 0 || in_output.count(l)) res.push_back(l);
    for (Leg l : res) refcount[l]++;
    double sz = 0;
    for (Leg l : res) sz += log2d(l);
    maxsize = std::max(maxsize, sz);
    double rs = std::exp2(sz);
    peak = std::max(peak, alive + rs);
    alive += rs - slot_size[a] - slot_size[b];
    slot_size.push_back(rs);
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
      if (!std::binary_search(legs[last].begin(), legs[last].end(), l)) throw Error("final legs do not contain every output leg");
  } else if (legs[last] != out) {
    throw Error("final legs do not match output");
  }
  return PathStats{flops * std::log10(2.0), maxsize, std::log2(peak)};
}

// src/compiled.rs:445-478
static std::vector<std::tuple<size_t, size_t, bool>> plan_traces(std::vector<Leg>& legs, const std::map<Leg, size_t>& rc,
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
    if (std::count(legs.begin(), legs.end(), l) > 2) throw Error(cat("leg ", l, " appears more than twice in one tensor"));
    if (rc.at(l) > 1 || in_output.count(l)) {
      ops.emplace_back(di, dj, true);  // diagonal keeps axis di
      legs.erase(legs.begin() + dj);
    } else {
      ops.emplace_back(di, dj, false);  // trace removes both axes
      legs.erase(legs.begin() + dj);
      legs.erase(legs.begin() + di);
    }
  }
}

// src/compiled.rs:481-602: leg classes in operand order; A -> [batch, free_a, con],
// B -> [batch, con, free_b]; result [batch, free_a, free_b].
static PairStep plan_pairwise(const Net& net, const std::vector<Leg>& la, const std::vector<Leg>& lb, size_t li, size_t ri,
                              size_t out, const std::map<Leg, size_t>& rc, const std::set<Leg>& in_output) {
  auto keep = [&](Leg l) { return rc.at(l) > 0 || in_output.count(l) > 0; };
  std::set<Leg> sa(la.begin(), la.end()), sb(lb.begin(), lb.end());
  std::vector<Leg> batch, con, free_a, sum_a, free_b, sum_b;
  for (Leg l : la) {
    if (sb.count(l)) (keep(l) ? batch : con).push_back(l);
    else (keep(l) ? free_a : sum_a).push_back(l);
  }
  for (Leg l : lb)
    if (!sa.count(l)) (keep(l) ? free_b : sum_b).push_back(l);
  PairStep st;
  st.li = li, st.ri = ri, st.out = out;
  std::vector<Leg> wa = la, wb = lb;
  for (Leg l : sum_a) {
    size_t ax = pos_of(wa, l);
    st.sum_a.push_back(ax);
    wa.erase(wa.begin() + ax);
  }
  for (Leg l : sum_b) {
    size_t ax = pos_of(wb, l);
    st.sum_b.push_back(ax);
    wb.erase(wb.begin() + ax);
  }
  for (auto* g : {&batch, &free_a, &con})
    for (Leg l : *g) st.a_perm.push_back(pos_of(wa, l));
  for (auto* g : {&batch, &con, &free_b})
    for (Leg l : *g) st.b_perm.push_back(pos_of(wb, l));
  auto dprod = [&](const std::vector<Leg>& v, const char* what) {
    size_t p = 1;
    for (Leg l : v) {
      size_t d = net.dim(l);
      if (p > SIZE_MAX / d) throw Error(cat("compiled GEMM ", what, " dimension product overflows"));
      p *= d;
    }
    return p;
  };
  st.bt = dprod(batch, "batch"), st.m = dprod(free_a, "m"), st.k = dprod(con, "k"), st.n = dprod(free_b, "n");
  // src/contract.rs:49-100: every operand offset must stay inside isize.
  auto mul = [](size_t a, size_t b) {
    if (b && a > (size_t)PTRDIFF_MAX / b) throw Error("GEMM element offset exceeds the isize range");
    return a * b;
  };
  mul(st.bt, mul(st.m, st.k));
  mul(st.bt, mul(st.k, st.n));
  mul(st.bt, mul(st.m, st.n));
  for (auto* g : {&batch, &free_a, &free_b})
    for (Leg l : *g) st.result_legs.push_back(l);
  st.result_shape = net.shape_of(st.result_legs);
  st.left_legs = la, st.right_legs = lb;
  st.left_shape = net.shape_of(la), st.right_shape = net.shape_of(lb);
  return st;
}

// src/compiled.rs:254-366
Compiled compile(const Net& net, const Path& path) {
  simulate_path(net, path);  // validates the network and the path
  std::set<Leg> in_output(net.output.begin(), net.output.end());
  std::map<Leg, size_t> rc;
  for (auto& t : net.inputs)
    for (Leg l : sorted_dedup(t)) rc[l]++;
  Compiled c;
  c.n_inputs = net.inputs.size();
  std::vector<std::vector<Leg>> slots;
  std::vector<bool> live;
  for (auto& legs0 : net.inputs) {
    auto legs = legs0;
    InputPrep p;
    p.traces = plan_traces(legs, rc, in_output);
    p.input_shape = net.shape_of(legs0);
    p.shape = net.shape_of(legs);
    p.legs = legs;
    c.preps.push_back(p);
    slots.push_back(legs);
    live.push_back(true);
  }
  for (size_t s = 0; s < path.size(); ++s) {
    size_t ia = path[s].first, ib = path[s].second;
    live[ia] = live[ib] = false;
    auto la = slots[ia], lb = slots[ib];
    for (Leg l : la) rc[l]--;
    for (Leg l : lb) rc[l]--;
    PairStep st = plan_pairwise(net, la, lb, ia, ib, slots.size(), rc, in_output);
    for (Leg l : st.result_legs) rc[l]++;
    slots.push_back(st.result_legs);
    live.push_back(true);
    c.steps.push_back(std::move(st));
  }
  size_t last = 0;
  for (size_t i = 0; i < live.size(); ++i)
    if (live[i]) last = i;
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
  return c;
}

// src/slice.rs:893-905
size_t slice_parallel_chunks(size_t n_slices, size_t out_numel) {
  if (n_slices <= 1) return 1;
  size_t budget = std::max<size_t>(((size_t)1 << 22) / std::max<size_t>(out_numel, 1), 1);
  size_t n0 = std::max<size_t>(std::min(std::min(n_slices, (size_t)256), budget), 1);
  if (n0 <= 1) return 1;
  size_t chunk = (n_slices + n0 - 1) / n0;
  return (n_slices + chunk - 1) / chunk;
}

Plan make_plan(const Net& net, const Path& path, const std::vector<Leg>& sliced) {
  Plan p;
  p.net = net;
  p.path = path;
  p.sliced = sliced;
  auto st = simulate_path(net, path);  // validates network + path on the full network
  p.log10_flops = st.log10_flops;
  p.log2_peak_size = st.log2_peak_size;
  // src/slice.rs:146-172
  {
    std::set<Leg> seen;
    for (Leg l : sliced) {
      if (!seen.insert(l).second) throw Error(cat("sliced leg ", l, " repeated (重复)"));
      auto it = net.size_dict.find(l);
      if (it == net.size_dict.end()) throw Error(cat("sliced leg ", l, " is not in size_dict"));
      bool held = false;
      for (auto& t : net.inputs) held = held || std::count(t.begin(), t.end(), l) > 0;
      if (!held) throw Error(cat("sliced leg ", l, " is not held by any input tensor"));
      for (auto& t : net.inputs)
        if (std::count(t.begin(), t.end(), l) > 1)
          throw Error(cat("sliced leg ", l, " repeated within one input tensor (同一输入张量中重复)"));
    }
  }
  std::set<Leg> sset(sliced.begin(), sliced.end());
  for (Leg l : net.output)
    if (sset.count(l)) p.outer.push_back(l);
  for (Leg l : sliced)
    if (!std::count(p.outer.begin(), p.outer.end(), l)) p.internal.push_back(l);
  // Slice count overflow contract (src/slice.rs:989-994, 1077-1081).
  for (Leg l : sliced) {
    size_t d = net.dim(l);
    if (p.n_slices > SIZE_MAX / d) throw Error("slice count overflows usize");
    p.n_slices *= d;
  }
  for (Leg l : p.outer) p.n_blocks *= net.dim(l);
  for (Leg l : p.internal) p.n_internal *= net.dim(l);
  p.full_out_shape = net.shape_of(net.output);
  p.full_out_numel = checked_numel(p.full_out_shape);
  if (!p.outer.empty() && p.full_out_numel > (size_t)PTRDIFF_MAX / 16)
    throw Error("full output tensor bytes exceed allocatable capacity (完整输出张量的字节数)");
  // Per-slice network: all sliced legs leave the inputs, outer legs leave the output.
  p.sub = net;
  for (auto& legs : p.sub.inputs)
    legs.erase(std::remove_if(legs.begin(), legs.end(), [&](Leg l) { return sset.count(l) > 0; }), legs.end());
  p.sub.output.erase(std::remove_if(p.sub.output.begin(), p.sub.output.end(), [&](Leg l) { return sset.count(l) > 0; }),
                     p.sub.output.end());
  for (Leg l : sliced) p.sub.size_dict.erase(l);
  p.cp = compile(p.sub, path);
  p.sub_out_numel = checked_numel(p.sub.shape_of(p.sub.output));
  p.n_chunks = slice_parallel_chunks(p.n_internal, p.sub_out_numel);
  // Global slice g = block * n_internal + internal; both decoded last-leg-fastest
  // (src/slice.rs:881-888).
  std::map<Leg, std::pair<size_t, size_t>> div_mod;  // leg -> (div, mod) in g
  {
    size_t div = 1;
    for (size_t q = p.internal.size(); q-- > 0;) {
      size_t d = net.dim(p.internal[q]);
      div_mod[p.internal[q]] = {div, d};
      div *= d;
    }
    div = p.n_internal;
    for (size_t q = p.outer.size(); q-- > 0;) {
      size_t d = net.dim(p.outer[q]);
      div_mod[p.outer[q]] = {div, d};
      div *= d;
    }
  }
  p.sel.resize(net.inputs.size());
  for (size_t i = 0; i < net.inputs.size(); ++i) {
    auto shape = net.shape_of(net.inputs[i]);
    std::vector<size_t> strides(shape.size(), 1);
    for (size_t a = shape.size(); a-- > 1;) strides[a - 1] = strides[a] * shape[a];
    for (size_t a = 0; a < shape.size(); ++a) {
      auto it = div_mod.find(net.inputs[i][a]);
      if (it != div_mod.end()) p.sel[i].push_back(SliceSel{it->second.first, it->second.second, strides[a], a});
    }
  }
  // Output scatter of block results (src/slice.rs:1113-1126, 1157-1170).
  {
    auto& shape = p.full_out_shape;
    std::vector<size_t> strides(shape.size(), 1);
    for (size_t a = shape.size(); a-- > 1;) strides[a - 1] = strides[a] * shape[a];
    for (size_t a = 0; a < shape.size(); ++a) {
      Leg l = net.output[a];
      auto it = div_mod.find(l);
      if (it != div_mod.end()) {
        p.block_div.push_back(it->second.first / p.n_internal);
        p.block_mod.push_back(it->second.second);
        p.block_stride.push_back(strides[a]);
      } else {
        p.kept_dims.push_back(shape[a]);
        p.kept_strides.push_back(strides[a]);
      }
    }
  }
  // Slot dependence on the slice index and per-slice work.
  p.varying.assign(p.cp.n_inputs, false);
  for (size_t i = 0; i < p.cp.n_inputs; ++i) p.varying[i] = !p.sel[i].empty();
  for (auto& s : p.cp.steps) {
    bool v = p.varying[s.li] || p.varying[s.ri];
    p.varying.push_back(v);
    p.n_varying_steps += v;
    p.macs_per_slice += (double)s.bt * (double)s.m * (double)s.k * (double)s.n;
    std::set<Leg> uni(s.left_legs.begin(), s.left_legs.end());
    uni.insert(s.right_legs.begin(), s.right_legs.end());
    double mm = 1;
    for (Leg l : uni) mm *= (double)p.sub.dim(l);
    p.mults_per_slice += mm;
  }
  return p;
}

static std::string jl(const std::vector<Leg>& v) {
  std::string s = "[";
  for (size_t i = 0; i < v.size(); ++i) s += (i ? "," : "") + std::to_string(v[i]);
  return s + "]";
}
static std::string js(const std::vector<size_t>& v) {
  std::string s = "[";
  for (size_t i = 0; i < v.size(); ++i) s += (i ? "," : "") + std::to_string(v[i]);
  return s + "]";
}

std::string describe_json(const Plan& p) {
  const Compiled& c = p.cp;
  std::string s = "{\"n_inputs\":" + std::to_string(c.n_inputs) + ",\"preps\":[";
  for (size_t i = 0; i < c.preps.size(); ++i) {
    auto& q = c.preps[i];
    s += (i ? "," : "") + std::string("{\"input_shape\":") + js(q.input_shape) + ",\"legs\":" + jl(q.legs) +
         ",\"shape\":" + js(q.shape) + ",\"traces\":[";
    for (size_t t = 0; t < q.traces.size(); ++t)
      s += (t ? "," : "") + std::string("[") + std::to_string(std::get<0>(q.traces[t])) + "," +
           std::to_string(std::get<1>(q.traces[t])) + "," + (std::get<2>(q.traces[t]) ? "true" : "false") + "]";
    s += "]}";
  }
  s += "],\"steps\":[";
  for (size_t i = 0; i < c.steps.size(); ++i) {
    auto& t = c.steps[i];
    s += (i ? "," : "") + std::string("{\"li\":") + std::to_string(t.li) + ",\"ri\":" + std::to_string(t.ri) +
         ",\"out\":" + std::to_string(t.out) + ",\"left_legs\":" + jl(t.left_legs) + ",\"right_legs\":" + jl(t.right_legs) +
         ",\"result_legs\":" + jl(t.result_legs) + ",\"left_shape\":" + js(t.left_shape) + ",\"right_shape\":" +
         js(t.right_shape) + ",\"result_shape\":" + js(t.result_shape) + ",\"sum_a\":" + js(t.sum_a) + ",\"sum_b\":" +
         js(t.sum_b) + ",\"a_perm\":" + js(t.a_perm) + ",\"b_perm\":" + js(t.b_perm) + ",\"gemm\":" +
         js({t.bt, t.m, t.k, t.n}) + ",\"varying\":" + (p.varying[t.out] ? "true" : "false") + "}";
  }
  auto& f = c.fin;
  s += "],\"finalize\":{\"slot\":" + std::to_string(f.slot) + ",\"input_legs\":" + jl(f.input_legs) +
       ",\"input_shape\":" + js(f.input_shape) + ",\"sum\":" + js(f.sum) + ",\"perm\":" + js(f.perm) +
       ",\"output_shape\":" + js(f.output_shape) + "},\"outer\":" + jl(p.outer) + ",\"internal\":" + jl(p.internal) +
       ",\"n_slices\":" + std::to_string(p.n_slices) + ",\"n_blocks\":" + std::to_string(p.n_blocks) + ",\"n_chunks\":" +
       std::to_string(p.n_chunks) + "}";
  return s;
}

}  // namespace atn
