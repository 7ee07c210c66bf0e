// C-ABI of the oracle (test infrastructure only; see arctn_oracle.hpp).
//
// Networks arrive as flat arrays (orc_net); tensors as host pointers in
// net.inputs order, row-major, complex values interleaved (re, im).
// Every entry point returns 0 on success or -1 with a message available from
// orc_last_error(); no exception crosses the C boundary.
#include <complex>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <string>

#include "arctn_oracle.hpp"

using namespace orc;

extern "C" {
typedef struct {
  size_t n_inputs;
  const size_t* input_lens;
  const uint32_t* input_legs;
  size_t n_output;
  const uint32_t* output;
  size_t n_dims;
  const uint32_t* dim_legs;
  const size_t* dim_sizes;
  size_t n_steps;
  const size_t* path;  // 2 * n_steps
  size_t n_sliced;
  const uint32_t* sliced;
} orc_net;
}

static thread_local std::string g_err;

static Network to_net(const orc_net* d) {
  Network n;
  size_t off = 0;
  for (size_t i = 0; i < d->n_inputs; ++i) {
    n.inputs.emplace_back(d->input_legs + off, d->input_legs + off + d->input_lens[i]);
    off += d->input_lens[i];
  }
  n.output.assign(d->output, d->output + d->n_output);
  for (size_t i = 0; i < d->n_dims; ++i) n.size_dict[d->dim_legs[i]] = d->dim_sizes[i];
  return n;
}
static Path to_path(const orc_net* d) {
  Path p;
  for (size_t s = 0; s < d->n_steps; ++s) p.push_back({d->path[2 * s], d->path[2 * s + 1]});
  return p;
}
static std::vector<Leg> to_sliced(const orc_net* d) { return std::vector<Leg>(d->sliced, d->sliced + d->n_sliced); }

static int guard(const std::function<void()>& f) {
  try {
    f();
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

template <class T>
static std::vector<Tensor<T>> load_inputs(const Network& net, const void* const* ptrs) {
  std::vector<Tensor<T>> ts;
  for (size_t i = 0; i < net.inputs.size(); ++i) {
    auto shape = net.shape_of(net.inputs[i]);
    size_t n = checked_numel(shape);
    const T* p = static_cast<const T*>(ptrs[i]);
    ts.emplace_back(shape, std::vector<T>(p, p + n));
  }
  return ts;
}

template <class T>
static void store(const Tensor<T>& t, void* out) {
  std::memcpy(out, t.data.data(), t.data.size() * sizeof(T));
}

// dtype codes: 0 f32, 1 f64, 2 complex64, 3 complex128
template <class F>
static void dispatch(int dtype, F&& f) {
  switch (dtype) {
    case 0: f(float{}); break;
    case 1: f(double{}); break;
    case 2: f(std::complex<float>{}); break;
    case 3: f(std::complex<double>{}); break;
    default: throw Error("unsupported dtype code");
  }
}

static std::string legs_json(const std::vector<Leg>& v) {
  std::string s = "[";
  for (size_t i = 0; i < v.size(); ++i) s += (i ? "," : "") + std::to_string(v[i]);
  return s + "]";
}
static std::string sz_json(const std::vector<size_t>& v) {
  std::string s = "[";
  for (size_t i = 0; i < v.size(); ++i) s += (i ? "," : "") + std::to_string(v[i]);
  return s + "]";
}

extern "C" {

const char* orc_last_error(void) { return g_err.c_str(); }
void orc_free(void* p) { std::free(p); }

size_t orc_slice_parallel_chunks(size_t n_slices, size_t out_numel) { return slice_parallel_chunks(n_slices, out_numel); }

// Compiled plan as JSON (mirror of CompiledContraction's accessors).
int orc_compile_describe(const orc_net* d, char** json_out) {
  return guard([&] {
    Compiled c = compile(to_net(d), to_path(d));
    std::string s = "{\"n_inputs\":" + std::to_string(c.n_inputs) + ",\"preps\":[";
    for (size_t i = 0; i < c.preps.size(); ++i) {
      auto& p = c.preps[i];
      s += (i ? "," : "") + std::string("{\"input_shape\":") + sz_json(p.input_shape) + ",\"legs\":" + legs_json(p.legs) +
           ",\"shape\":" + sz_json(p.shape) + ",\"traces\":[";
      for (size_t q = 0; q < p.traces.size(); ++q)
        s += (q ? "," : "") + std::string("[") + std::to_string(std::get<0>(p.traces[q])) + "," + std::to_string(std::get<1>(p.traces[q])) +
             "," + (std::get<2>(p.traces[q]) ? "true" : "false") + "]";
      s += "]}";
    }
    s += "],\"steps\":[";
    for (size_t i = 0; i < c.steps.size(); ++i) {
      auto& t = c.steps[i];
      s += (i ? "," : "") + std::string("{\"li\":") + std::to_string(t.li) + ",\"ri\":" + std::to_string(t.ri) +
           ",\"out\":" + std::to_string(t.out) + ",\"left_legs\":" + legs_json(t.left_legs) + ",\"right_legs\":" + legs_json(t.right_legs) +
           ",\"result_legs\":" + legs_json(t.result_legs) + ",\"left_shape\":" + sz_json(t.left_shape) +
           ",\"right_shape\":" + sz_json(t.right_shape) + ",\"result_shape\":" + sz_json(t.result_shape) +
           ",\"sum_a\":" + sz_json(t.sum_a) + ",\"sum_b\":" + sz_json(t.sum_b) + ",\"a_perm\":" + sz_json(t.a_perm) +
           ",\"b_perm\":" + sz_json(t.b_perm) + ",\"gemm\":" + sz_json({t.bt, t.m, t.k, t.n}) + "}";
    }
    auto& f = c.fin;
    char buf[64];
    s += "],\"finalize\":{\"slot\":" + std::to_string(f.slot) + ",\"input_legs\":" + legs_json(f.input_legs) +
         ",\"input_shape\":" + sz_json(f.input_shape) + ",\"sum\":" + sz_json(f.sum) + ",\"perm\":" + sz_json(f.perm) +
         ",\"output_shape\":" + sz_json(f.output_shape) + "},";
    auto num = [&](double v) {  // JSON has no infinities: an empty path has log10 FLOPs -inf
      if (!std::isfinite(v)) return std::string("null");
      std::snprintf(buf, sizeof buf, "%.17g", v);
      return std::string(buf);
    };
    s += std::string("\"log10_flops\":") + num(c.log10_flops) + ",";
    s += std::string("\"log2_peak_size\":") + num(c.log2_peak_size) + "}";
    *json_out = static_cast<char*>(std::malloc(s.size() + 1));
    std::memcpy(*json_out, s.c_str(), s.size() + 1);
  });
}

int orc_simulate_path(const orc_net* d, double* stats6) {
  return guard([&] {
    Stats s = simulate_path(to_net(d), to_path(d));
    double v[6] = {s.log10_flops, s.log2_max_size, s.log2_max_contraction_size, s.log2_total_size, s.log2_read_write, s.log2_peak_size};
    std::memcpy(stats6, v, sizeof v);
  });
}

// mode: 0 compiled execute, 1 compiled execute_stripped, 2 direct
// contract_network, 3 sliced (d->sliced), 4 naive einsum.
int orc_contract(const orc_net* d, int dtype, int mode, const void* const* inputs, void* out, double* log10_exponent, int threads) {
  return guard([&] {
    Network net = to_net(d);
    dispatch(dtype, [&](auto tag) {
      using T = decltype(tag);
      auto ts = load_inputs<T>(net, inputs);
      Tensor<T> r;
      double e = 0.0;
      switch (mode) {
        case 0: validate_shapes(net, ts); r = execute(compile(net, to_path(d)), ts, false, nullptr); break;
        case 1: validate_shapes(net, ts); r = execute(compile(net, to_path(d)), ts, true, &e); break;
        case 2: r = contract_network(net, ts, to_path(d)); break;
        case 3: r = contract_sliced(net, ts, to_path(d), to_sliced(d), threads); break;
        case 4: r = naive_einsum(net, ts); break;
        default: throw Error("unknown mode");
      }
      if (log10_exponent) *log10_exponent = e;
      store(r, out);
    });
  });
}

// Internal slices [lo, hi) of a sliced contraction, summed (bounded CPU sample).
int orc_contract_slice_range(const orc_net* d, int dtype, size_t lo, size_t hi, const void* const* inputs, void* out,
                             int threads) {
  return guard([&] {
    Network net = to_net(d);
    dispatch(dtype, [&](auto tag) {
      using T = decltype(tag);
      auto ts = load_inputs<T>(net, inputs);
      store(contract_slice_range(net, ts, to_path(d), to_sliced(d), lo, hi, threads), out);
    });
  });
}

// Gradients of the compiled contraction w.r.t. every input, given
// dL/d(conj output). grads[i] receives input i's gradient.
int orc_gradient(const orc_net* d, int dtype, const void* const* inputs, const void* grad_out, void* const* grads) {
  return guard([&] {
    Network net = to_net(d);
    dispatch(dtype, [&](auto tag) {
      using T = decltype(tag);
      auto ts = load_inputs<T>(net, inputs);
      validate_shapes(net, ts);
      Compiled c = compile(net, to_path(d));
      auto oshape = net.shape_of(net.output);
      const T* g = static_cast<const T*>(grad_out);
      Tensor<T> gt(oshape, std::vector<T>(g, g + checked_numel(oshape)));
      auto gs = gradient(c, ts, gt);
      for (size_t i = 0; i < gs.size(); ++i) store(gs[i], grads[i]);
    });
  });
}

// Single-tensor kernels. op: 0 permute(axes=arg), 1 sum_axis(arg[0]),
// 2 trace_pair(arg[0], arg[1]), 3 diag_pair(arg[0], arg[1]),
// 4 select_axis(arg[0], arg[1]). Output shape is implied by the op.
int orc_tensor_op(int dtype, int op, size_t rank, const size_t* shape, const size_t* arg, size_t n_arg, const void* in, void* out) {
  return guard([&] {
    dispatch(dtype, [&](auto tag) {
      using T = decltype(tag);
      std::vector<size_t> sh(shape, shape + rank);
      size_t n = checked_numel(sh);
      const T* p = static_cast<const T*>(in);
      Tensor<T> t(sh, std::vector<T>(p, p + n));
      std::vector<size_t> a(arg, arg + n_arg);
      Tensor<T> r;
      switch (op) {
        case 0: r = t.permute(a); break;
        case 1: r = t.sum_axis(a.at(0)); break;
        case 2: r = t.trace_pair(a.at(0), a.at(1)); break;
        case 3: r = t.diag_pair(a.at(0), a.at(1)); break;
        case 4: r = t.select_axis(a.at(0), a.at(1)); break;
        default: throw Error("unknown op");
      }
      store(r, out);
    });
  });
}

}  // extern "C"
