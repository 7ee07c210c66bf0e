// C ABI (include/arctn_hip.h). No exception crosses this boundary; every
// failure becomes a negative status and a thread-local message.
#include <cstdlib>
#include <memory>
#include <mutex>

#include "../../include/arctn_hip.h"
#include "executor.hpp"

using namespace atn;

struct atn_plan {
  std::shared_ptr<const Plan> p;
};
struct atn_executor {
  std::unique_ptr<Executor> ex;
};

static thread_local std::string g_err;

template <class F>
static int guard(F&& f) {
  try {
    f();
    return ATN_OK;
  } catch (const Error& e) {
    g_err = e.what();
    return e.code < 0 ? e.code : ATN_ERR_INVALID;
  } catch (const std::bad_alloc&) {
    g_err = "host allocation failed";
    return ATN_ERR_INTERNAL;
  } catch (const std::exception& e) {
    g_err = e.what();
    return ATN_ERR_INTERNAL;
  }
}

static Net to_net(const atn_network_desc* d) {
  if (!d) throw Error("network description is NULL");
  Net n;
  size_t off = 0;
  for (size_t i = 0; i < d->n_inputs; ++i) {
    n.inputs.emplace_back(d->input_legs + off, d->input_legs + off + d->input_lens[i]);
    off += d->input_lens[i];
  }
  n.output.assign(d->output, d->output + d->n_output);
  for (size_t i = 0; i < d->n_dims; ++i) {
    if (n.size_dict.count(d->dim_legs[i])) throw Error(cat("size_dict repeats leg ", d->dim_legs[i]));
    n.size_dict[d->dim_legs[i]] = d->dim_sizes[i];
  }
  return n;
}

static ExecOptions to_opts(const atn_exec_options* o) {
  ExecOptions e;
  if (!o) return e;
  if (o->dtype < 0 || o->dtype > 3) throw Error("unsupported dtype code");
  if (o->precision < 0 || o->precision > 5) throw Error("unsupported precision code");
  e.dtype = o->dtype;
  e.precision = o->precision;
  e.batched = o->batched != 0;
  e.use_graph = o->use_graph != 0;
  e.autotune = o->autotune != 0;
  e.strip = o->strip != 0;
  e.max_batch = o->max_batch;
  e.workspace_limit = o->workspace_limit;
  e.slice_begin = o->slice_begin;
  e.slice_end = o->slice_end;
  e.gather = o->permute_b == 0;
  if (o->gemm_kernel < 0 || o->gemm_kernel > 9) throw Error("unknown GEMM kernel");
  e.gemm_kernel = o->gemm_kernel - 1;
  return e;
}

static void to_report(const Report& r, atn_report* out) {
  if (!out) return;
  out->elapsed_ms = r.elapsed_ms;
  out->flops = r.flops;
  out->tflops = r.tflops;
  out->slices_per_s = r.slices_per_s;
  out->log10_exponent = r.log10_exponent;
  out->max_abs_result = r.max_abs_result;
  out->n_slices = r.n_slices;
  out->batch = r.batch;
  out->n_launches = r.n_launches;
  out->workspace_bytes = r.workspace_bytes;
}

static void check_dtype(int dt) {
  if (dt < 0 || dt > 3) throw Error("unsupported dtype code");
}

extern "C" {

int atn_abi_version(void) { return ATN_ABI_VERSION; }
const char* atn_last_error(void) { return g_err.c_str(); }
void atn_free(void* p) { std::free(p); }

int atn_plan_create(const atn_network_desc* desc, atn_plan** out) {
  return guard([&] {
    if (!out) throw Error("out is NULL");
    *out = nullptr;
    Net n = to_net(desc);
    Path path;
    for (size_t s = 0; s < desc->n_steps; ++s) path.push_back({desc->ssa_path[2 * s], desc->ssa_path[2 * s + 1]});
    std::vector<Leg> sliced(desc->sliced_legs, desc->sliced_legs + desc->n_sliced);
    auto p = std::make_shared<Plan>(make_plan(n, path, sliced));
    *out = new atn_plan{p};
  });
}

void atn_plan_destroy(atn_plan* plan) { delete plan; }

int atn_plan_stats(const atn_plan* plan, atn_plan_info* out) {
  return guard([&] {
    if (!plan || !out) throw Error("NULL argument");
    const Plan& p = *plan->p;
    out->n_inputs = p.cp.n_inputs;
    out->n_steps = p.cp.steps.size();
    out->n_slices = p.n_slices;
    out->n_blocks = p.n_blocks;
    out->n_varying_steps = p.n_varying_steps;
    out->log10_flops = p.log10_flops;
    out->log2_peak_size = p.log2_peak_size;
    out->macs_per_slice = p.macs_per_slice;
    out->mults_per_slice = p.mults_per_slice;
  });
}

int atn_plan_describe(const atn_plan* plan, char** json_out) {
  return guard([&] {
    if (!plan || !json_out) throw Error("NULL argument");
    std::string s = describe_json(*plan->p);
    char* buf = static_cast<char*>(std::malloc(s.size() + 1));
    if (!buf) throw std::bad_alloc();
    std::memcpy(buf, s.c_str(), s.size() + 1);
    *json_out = buf;
  });
}

int atn_executor_create(const atn_plan* plan, const atn_exec_options* opts, atn_executor** out) {
  return guard([&] {
    if (!plan || !out) throw Error("NULL argument");
    *out = nullptr;
    auto ex = std::make_unique<Executor>(plan->p, to_opts(opts));
    *out = new atn_executor{std::move(ex)};
  });
}

void atn_executor_destroy(atn_executor* ex) { delete ex; }

int atn_executor_run(atn_executor* ex, const void* const* d_inputs, void* d_out, void* stream, atn_report* report) {
  return guard([&] {
    if (!ex || !d_out) throw Error("NULL argument");
    Report r;
    ex->ex->run(d_inputs, d_out, static_cast<hipStream_t>(stream), report ? &r : nullptr);
    to_report(r, report);
  });
}

int atn_executor_profile(atn_executor* ex, const void* const* d_inputs, void* d_out, void* stream,
                         atn_kernel_stat* stats, size_t max_stats, size_t* n_stats) {
  return guard([&] {
    if (!ex || !d_out || !stats || !n_stats) throw Error("NULL argument");
    std::vector<KernelStat> ks;
    ex->ex->profile(d_inputs, d_out, static_cast<hipStream_t>(stream), ks);
    size_t n = 0;
    for (size_t k = 0; k < ks.size() && n < max_stats; ++k) {
      if (!ks[k].calls) continue;
      atn_kernel_stat& o = stats[n++];
      std::memset(o.name, 0, sizeof o.name);
      std::strncpy(o.name, kernel_kind_name((int)k), sizeof o.name - 1);
      o.calls = ks[k].calls;
      o.launches = ks[k].launches;
      o.ms = ks[k].ms;
      o.flops = ks[k].flops;
      o.bytes = ks[k].bytes;
    }
    *n_stats = n;
  });
}

int atn_execute(const atn_plan* plan, const atn_exec_options* opts, const void* const* d_inputs, void* d_out, void* stream,
                atn_report* report) {
  return guard([&] {
    if (!plan || !d_out) throw Error("NULL argument");
    Executor ex(plan->p, to_opts(opts));
    Report r;
    ex.run(d_inputs, d_out, static_cast<hipStream_t>(stream), report ? &r : nullptr);
    if (!report) ATN_HIP(hipStreamSynchronize(static_cast<hipStream_t>(stream)));  // workspace is freed on return
    to_report(r, report);
  });
}

int atn_prep_forward(const atn_plan* plan, size_t input, int dtype, const void* in, void* out, void* stream) {
  return guard([&] {
    check_dtype(dtype);
    stage_prep_forward(*plan->p, input, dtype, in, out, static_cast<hipStream_t>(stream));
  });
}
int atn_prep_adjoint(const atn_plan* plan, size_t input, int dtype, const void* grad_out, void* grad_in, void* stream) {
  return guard([&] {
    check_dtype(dtype);
    stage_prep_adjoint(*plan->p, input, dtype, grad_out, grad_in, static_cast<hipStream_t>(stream));
  });
}
int atn_step_forward(const atn_plan* plan, size_t step, int dtype, int precision, const void* left, const void* right,
                     void* out, void* stream) {
  return guard([&] {
    check_dtype(dtype);
    stage_step_forward(*plan->p, step, dtype, precision, left, right, out, static_cast<hipStream_t>(stream));
  });
}
int atn_step_adjoint(const atn_plan* plan, size_t step, int dtype, const void* left, const void* right, const void* grad_out,
                     void* grad_left, void* grad_right, void* stream) {
  return guard([&] {
    check_dtype(dtype);
    stage_step_adjoint(*plan->p, step, dtype, left, right, grad_out, grad_left, grad_right, static_cast<hipStream_t>(stream));
  });
}
int atn_finalize_forward(const atn_plan* plan, int dtype, const void* in, void* out, void* stream) {
  return guard([&] {
    check_dtype(dtype);
    stage_finalize_forward(*plan->p, dtype, in, out, static_cast<hipStream_t>(stream));
  });
}
int atn_finalize_adjoint(const atn_plan* plan, int dtype, const void* grad_out, void* grad_in, void* stream) {
  return guard([&] {
    check_dtype(dtype);
    stage_finalize_adjoint(*plan->p, dtype, grad_out, grad_in, static_cast<hipStream_t>(stream));
  });
}

int atn_tensor_op(int dtype, int op, size_t rank, const size_t* shape, const size_t* arg, size_t n_arg, const void* d_in,
                  void* d_out, void* stream) {
  return guard([&] {
    check_dtype(dtype);
    hipStream_t s = static_cast<hipStream_t>(stream);
    std::vector<size_t> sh(shape, shape + rank);
    checked_numel(sh);
    int es = (int)dtype_size(dtype);
    auto need = [&](size_t n) {
      if (n_arg < n) throw Error("missing op arguments");
    };
    auto span = [&](size_t lo, size_t hi) {
      size_t p = 1;
      for (size_t q = lo; q < hi; ++q) p *= sh[q];
      return p;
    };
    switch (op) {
      case 0: {
        std::vector<size_t> perm(arg, arg + n_arg);
        if (perm.size() != rank) throw Error("permute: rank mismatch");
        std::vector<bool> seen(rank, false);
        for (size_t a : perm) {
          if (a >= rank || seen[a]) throw Error("permute: axes are not a permutation");
          seen[a] = true;
        }
        PermPlan pp = plan_permute(sh, perm, 1, es);
        if (pp.kind == 0) launch_copy(prod(sh) * es, d_in, d_out, s);
        else launch_permute(pp, d_in, d_out, s);
        break;
      }
      case 1: {
        need(1);
        if (arg[0] >= rank) throw Error("sum_axis: axis out of range");
        launch_sum_axis(dtype, span(0, arg[0]), sh[arg[0]], span(arg[0] + 1, rank), d_in, d_out, s);
        break;
      }
      case 2:
      case 3: {
        need(2);
        size_t i = arg[0], j = arg[1];
        if (!(i < j && j < rank) || sh[i] != sh[j]) throw Error("trace/diag: bad axes");
        if (op == 2) launch_trace(dtype, span(0, i), sh[i], span(i + 1, j), span(j + 1, rank), d_in, d_out, s);
        else launch_diag(es, span(0, i), sh[i], span(i + 1, j), span(j + 1, rank), d_in, d_out, s);
        break;
      }
      case 4: {
        need(2);
        size_t ax = arg[0], idx = arg[1];
        if (ax >= rank || idx >= sh[ax]) throw Error("select_axis: index out of range");
        GatherArgs g{};
        size_t inner = span(ax + 1, rank), outer = span(0, ax);
        g.n_ax = 0;
        if (outer > 1) g.dims[g.n_ax] = (uint32_t)outer, g.str[g.n_ax] = (int64_t)(sh[ax] * inner), g.n_ax++;
        if (inner > 1) g.dims[g.n_ax] = (uint32_t)inner, g.str[g.n_ax] = 1, g.n_ax++;
        g.inner = outer * inner;
        g.n_sel = 1;
        g.sel_div[0] = 1, g.sel_mod[0] = sh[ax], g.sel_str[0] = (int64_t)inner;
        launch_gather(es, g, idx, 1, d_in, d_out, s);
        break;
      }
      case 5: {  // max magnitude: one double (0 for an empty or all-zero tensor)
        launch_fill_zero(d_out, sizeof(double), s);
        if (prod(sh) > 0) launch_maxabs(dtype, d_in, prod(sh), static_cast<uint64_t*>(d_out), s);
        break;
      }
      default: throw Error("unknown op");
    }
  });
}

}  // extern "C"

// Arguments of atn_gemm / atn_gemm_noalloc, validated before anything is launched.
static GemmArgs gemm_args(int dtype, int precision, int kernel, size_t batch, size_t m, size_t k, size_t n,
                          const void* d_a, size_t stride_a, const void* d_b, size_t stride_b, void* d_c,
                          size_t stride_c) {
  check_dtype(dtype);
  if (!batch || !m || !k || !n) throw Error("GEMM dimensions must be positive");
  if (kernel < -1 || kernel > 8) throw Error("unknown GEMM kernel");
  if (precision < 0 || precision > 5) throw Error("unsupported precision code");
  if (precision == C3M && dtype != C64) throw Error("precision c3m applies to complex64 only");
  GemmArgs g{};
  g.a = d_a, g.b = d_b, g.c = d_c;
  g.m = (int64_t)m, g.k = (int64_t)k, g.n = (int64_t)n;
  g.lda = g.k, g.ldb = g.n, g.ldc = g.n;
  g.nbo = (int64_t)batch, g.nbi = 1;
  g.sao = (int64_t)stride_a, g.sbo = (int64_t)stride_b, g.sco = (int64_t)stride_c;
  g.scratch_bytes = gemm_scratch_bytes(dtype, precision, kernel, g);
  return g;
}

// Scratch of atn_gemm_noalloc: a fixed device array, one copy per GPU, held
// until the process ends. Host code never passes `scratch` itself to a kernel
// (on the host the name is only a stand-in); it uses the device address that
// hipGetSymbolAddress returns for the current GPU, looked up once per device.
static __device__ __align__(256) unsigned char scratch[ATN_NOALLOC_SCRATCH_BYTES];

static std::mutex g_scratch_mu;
static void* g_scratch_addr[64];

static void* noalloc_scratch() {
  int dev = 0;
  ATN_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) throw Error(cat("device ordinal ", dev, " out of range"));
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_addr[dev]) {
    void* p = nullptr;
    ATN_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw Error("hipGetSymbolAddress returned no address for the scratch array", ATN_ERR_DEVICE);
    g_scratch_addr[dev] = p;
  }
  return g_scratch_addr[dev];
}

extern "C" {

int atn_gemm(int dtype, int precision, int kernel, size_t batch, size_t m, size_t k, size_t n, const void* d_a,
             size_t stride_a, const void* d_b, size_t stride_b, void* d_c, size_t stride_c, void* stream) {
  return guard([&] {
    GemmArgs g = gemm_args(dtype, precision, kernel, batch, m, k, n, d_a, stride_a, d_b, stride_b, d_c, stride_c);
    hipStream_t s = static_cast<hipStream_t>(stream);
    void* scr = nullptr;
    ATN_HIP(hipMallocAsync(&scr, g.scratch_bytes, s));
    g.scratch = scr;
    try {
      launch_gemm(dtype, precision, kernel, g, s);
    } catch (...) {
      (void)hipFreeAsync(scr, s);
      throw;
    }
    ATN_HIP(hipFreeAsync(scr, s));
  });
}

int atn_gemm_noalloc(int dtype, int precision, int kernel, size_t batch, size_t m, size_t k, size_t n,
                     const void* d_a, size_t stride_a, const void* d_b, size_t stride_b, void* d_c, size_t stride_c,
                     void* stream) {
  return guard([&] {
    GemmArgs g = gemm_args(dtype, precision, kernel, batch, m, k, n, d_a, stride_a, d_b, stride_b, d_c, stride_c);
    if (g.scratch_bytes > ATN_NOALLOC_SCRATCH_BYTES)
      throw Error(cat("GEMM needs ", g.scratch_bytes, " bytes of scratch; atn_gemm_noalloc has ",
                      (size_t)ATN_NOALLOC_SCRATCH_BYTES));
    g.scratch = noalloc_scratch();
    launch_gemm(dtype, precision, kernel, g, static_cast<hipStream_t>(stream));
  });
}

struct atn_adjoint {
  std::unique_ptr<GradRunner> r;
};

int atn_adjoint_create(const atn_plan* plan, int dtype, int precision, size_t checkpoint_every, atn_adjoint** out) {
  return guard([&] {
    if (!plan || !out) throw Error("NULL argument");
    *out = nullptr;
    check_dtype(dtype);
    if (precision < 0 || precision > 5) throw Error("unsupported precision code");
    auto r = std::make_unique<GradRunner>(plan->p, dtype, precision, checkpoint_every);
    *out = new atn_adjoint{std::move(r)};
  });
}

void atn_adjoint_destroy(atn_adjoint* a) { delete a; }

int atn_adjoint_forward(atn_adjoint* a, const void* const* d_inputs, void* d_out, void* stream) {
  return guard([&] {
    if (!a || !d_out || (!d_inputs && a->r->n_inputs())) throw Error("NULL argument");
    a->r->forward(d_inputs, d_out, static_cast<hipStream_t>(stream));
  });
}

int atn_adjoint_backward(atn_adjoint* a, const void* const* d_inputs, const void* d_grad_out, void* const* d_grad_inputs,
                         void* stream) {
  return guard([&] {
    if (!a || !d_grad_out || ((!d_inputs || !d_grad_inputs) && a->r->n_inputs())) throw Error("NULL argument");
    for (size_t i = 0; i < a->r->n_inputs(); ++i)
      if (!d_inputs[i] || !d_grad_inputs[i]) throw Error(cat("NULL input or gradient pointer ", i));
    a->r->backward(d_inputs, d_grad_out, d_grad_inputs, static_cast<hipStream_t>(stream));
  });
}

struct atn_grouped {
  std::unique_ptr<GroupedExecutor> ex;
};

int atn_grouped_create(const atn_plan* const* plans, size_t n_plans, int dtype, int precision, int strip,
                       atn_grouped** out) {
  return guard([&] {
    if (!out || (n_plans && !plans)) throw Error("NULL argument");
    *out = nullptr;
    check_dtype(dtype);
    if (precision < 0 || precision > 5) throw Error("unsupported precision code");
    std::vector<std::shared_ptr<const Plan>> ps;
    for (size_t i = 0; i < n_plans; ++i) {
      if (!plans[i]) throw Error("NULL plan");
      ps.push_back(plans[i]->p);
    }
    auto ex = std::make_unique<GroupedExecutor>(std::move(ps), dtype, strip != 0, precision);
    *out = new atn_grouped{std::move(ex)};
  });
}

void atn_grouped_destroy(atn_grouped* g) { delete g; }

int atn_grouped_run(atn_grouped* g, const void* const* const* inputs, void* const* outputs, double* log10_exponents,
                    void* stream, atn_report* report) {
  return guard([&] {
    if (!g || (g->ex->n_plans() && (!inputs || !outputs))) throw Error("NULL argument");
    std::vector<const void* const*> ins;
    std::vector<void*> outs;
    for (size_t i = 0; i < g->ex->n_plans(); ++i) {
      if (!inputs[i] || !outputs[i]) throw Error(cat("NULL input array or output of plan ", i));
      ins.push_back(inputs[i]);
      outs.push_back(outputs[i]);
    }
    Report r;
    g->ex->run(ins, outs, log10_exponents, static_cast<hipStream_t>(stream), report ? &r : nullptr);
    to_report(r, report);
  });
}

int atn_grouped_execute(const atn_plan* const* plans, size_t n_plans, int dtype, int precision, int strip,
                        const void* const* const* inputs, void* const* outputs, double* log10_exponents, void* stream,
                        atn_report* report) {
  return guard([&] {
    check_dtype(dtype);
    if (precision < 0 || precision > 5) throw Error("unsupported precision code");
    std::vector<std::shared_ptr<const Plan>> ps;
    std::vector<const void* const*> ins;
    std::vector<void*> outs;
    for (size_t i = 0; i < n_plans; ++i) {
      if (!plans[i]) throw Error("NULL plan");
      ps.push_back(plans[i]->p);
      ins.push_back(inputs[i]);
      outs.push_back(outputs[i]);
    }
    Report r;
    grouped_execute(ps, dtype, strip != 0, precision, ins, outs, log10_exponents, static_cast<hipStream_t>(stream),
                    report ? &r : nullptr);
    to_report(r, report);
  });
}

}  // extern "C"
