// Executors of a Plan on one GPU.
#pragma once

#include <functional>
#include <memory>
#include <mutex>

#include "kernels.hpp"
#include "plan.hpp"

namespace atn {

struct ExecOptions {
  int dtype = C64;
  int precision = NATIVE;
  bool batched = false;
  bool use_graph = false;
  bool autotune = false;
  bool strip = false;
  size_t max_batch = 0;
  size_t workspace_limit = 0;
  // global slice range [slice_begin, slice_end) of the batched form; 0, 0 = all slices
  size_t slice_begin = 0, slice_end = 0;
  // read B operands through gather tables instead of permuting them (where possible)
  bool gather = true;
  // GEMM kernel of every step (GemmKernel); GEMM_AUTO = chosen per step
  int gemm_kernel = -1;
};

struct Report {
  double elapsed_ms = 0, flops = 0, tflops = 0, slices_per_s = 0, log10_exponent = 0, max_abs_result = 0;
  size_t n_slices = 0, batch = 0, n_launches = 0, workspace_bytes = 0;
};

// Where an operand lives at run time.
struct Ref {
  enum Kind { NONE, WS, INPUT, OUTPUT } kind = NONE;
  int id = 0;  // workspace buffer id or input index
};

struct Ctx {
  char* ws = nullptr;
  const std::vector<size_t>* offs = nullptr;
  const void* const* inputs = nullptr;
  void* out = nullptr;
  hipStream_t s = nullptr;
  uint64_t g0 = 0;
  void* ptr(const Ref& r) const {
    switch (r.kind) {
      case Ref::WS: return ws + (*offs)[r.id];
      case Ref::INPUT: return const_cast<void*>(inputs[r.id]);
      case Ref::OUTPUT: return out;
      default: return nullptr;
    }
  }
};

// One enqueued operation and its algorithmic work (FLOPs and HBM bytes of the
// tensors it must read and write), used for per-kernel roofline accounting.
struct Op {
  std::function<void(const Ctx&)> fn;
  double flops = 0, bytes = 0;
  void operator()(const Ctx& c) const { fn(c); }
};

struct KernelStat {
  size_t calls = 0, launches = 0;
  double ms = 0, flops = 0, bytes = 0;
};

// Gather tables of the [con][free] view of an operand of `shape` permuted by
// `perm` (con = the first n_con output axes): rows[p] and cols[j] are element
// offsets into the unpermuted operand, col_run / row_run the lengths of runs of
// consecutive offsets (innermost output axes that stay innermost in memory).
// false when the operand's offsets do not fit int32.
bool gather_tables(const std::vector<size_t>& shape, const std::vector<size_t>& perm, size_t n_con,
                   std::vector<int32_t>& rows, std::vector<int32_t>& cols, int32_t& col_run, int32_t& row_run);

class Executor {
 public:
  Executor(std::shared_ptr<const Plan> plan, const ExecOptions& o);
  ~Executor();
  void run(const void* const* inputs, void* out, hipStream_t s, Report* rep);
  // Eager run with a pair of HIP events around every operation; accumulates
  // per-kernel-kind launches, device time and algorithmic FLOPs/bytes.
  void profile(const void* const* inputs, void* out, hipStream_t s, std::vector<KernelStat>& stats);
  size_t batch() const { return B_; }
  size_t workspace_bytes() const { return ws_bytes_; }

 private:
  struct Buf {
    size_t bytes;
    int first, last;
  };
  void build(size_t B);
  int newbuf(size_t bytes);
  void use(const Ref& r);
  void emit(std::function<void(const Ctx&)> fn, double flops = 0, double bytes = 0) {
    cur_->push_back(Op{std::move(fn), flops, bytes});
    ++t_;
  }
  void gen_steps(bool body, size_t nb, std::vector<Ref>& slots, std::vector<std::vector<size_t>>& shapes);
  void gen_inputs(bool body, size_t nb, std::vector<Ref>& slots, std::vector<std::vector<size_t>>& shapes);
  Ref gen_finalize(size_t nb, const Ref& src, bool to_output);
  // cmax: optional zeroed slot receiving the result's max magnitude (strip, or
  // the FP16X2 scale of a later product); brow / bcol: gather tables of B in
  // consts_ (-1: B is a row-major [k, n] operand); amax / bmax: slots holding the
  // operands' max magnitudes (FP16X2), or null
  // arow / acol: gather tables of A in consts_ (-1: A is a row-major [m, k] operand;
  // set only where the GEMM runs on the simple kernel)
  void gen_gemm(const PairStep& st, size_t nb_outer, bool va, bool vb, const Ref& a, const Ref& b, const Ref& c,
                const Ref* cmax = nullptr, int64_t brow = -1, int64_t bcol = -1, int32_t bcol_run = 1,
                const Ref* amax = nullptr, const Ref* bmax = nullptr, int32_t brow_run = 1, int64_t arow = -1,
                int64_t acol = -1);
  void run_ops(const std::vector<Op>& ops, const Ctx& c);
  void enqueue(const Ctx& c, const std::function<void(const Op&, const Ctx&)>& visit = nullptr);
  void autotune(const Ctx& c, hipStream_t s);

  struct GemmSite {
    GemmArgs g;
    Ref a, b, c;
    int64_t brow = -1, bcol = -1;  // gather tables of B in consts_ (element offsets), or -1
    int32_t bcol_run = 1;
    Ref amax, bmax;  // operand max magnitudes (FP16X2), or NONE
    int64_t arow = -1, acol = -1;  // gather tables of A (simple kernel), or -1
  };
  // FP16X2: step results whose max magnitude a later split product reads as its
  // operand scale (folded into the producing GEMM instead of a separate pass),
  // and the slot holding it per plan slot
  std::vector<char> want_max_;
  std::vector<Ref> slot_max_;
  // Storage order of each plan slot's axes (stored axis d holds the plan's axis
  // slot_order_[s][d]); a step may store its result with the free legs of B in an
  // order chosen for the step that reads it as B. consumer_[s]: (step, 1 if read
  // as B) of slot s, or (-1, 0).
  std::vector<std::vector<size_t>> slot_order_;
  std::vector<std::pair<long, int>> consumer_;
  // Gather tables (int32 element offsets) of B operands read without a permute.
  std::vector<int32_t> const_host_;
  int32_t* consts_ = nullptr;
  // Offsets of the [con][free] view of an operand of `shape` permuted by `perm`
  // (con = the first n_con output axes), appended to const_host_; false when the
  // tables would be too large.
  bool add_gather(const std::vector<size_t>& shape, const std::vector<size_t>& perm, size_t n_con, int64_t& row,
                  int64_t& col, int32_t& col_run, int32_t& row_run);

  std::shared_ptr<const Plan> P_;
  ExecOptions o_;
  int es_;
  size_t B_ = 1;
  size_t g_lo_ = 0, n_run_ = 1;  // slices run: [g_lo_, g_lo_ + n_run_)
  std::vector<Buf> bufs_;
  std::vector<size_t> offs_;
  std::vector<Op> pre_, body_, tail_, post_;
  std::vector<Op>* cur_ = nullptr;
  int t_ = 0;
  int body_begin_ = 0, body_end_ = 0;
  size_t tail_nb_ = 0;
  char* ws_ = nullptr;
  size_t ws_bytes_ = 0;
  int acc_buf_ = -1, expo_buf_ = -1, maxbits_buf_ = -1, scratch_buf_ = -1;
  hipGraphExec_t gexec_ = nullptr;
  hipGraph_t graph_ = nullptr;
  hipStream_t cap_stream_ = nullptr;
  std::vector<const void*> graph_inputs_;
  void* graph_out_ = nullptr;
  size_t launches_per_run_ = 0;
  // gemm kernel choice per GEMM op (autotune)
  std::vector<int> gemm_choice_;
  std::vector<GemmSite> gemm_sites_;
  bool tuned_ = false;
  // tail batch generation replays the body's buffer ids
  std::vector<int> body_ids_;
  size_t replay_pos_ = 0;
  bool replay_ = false, record_ = false;
  // Small GEMMs of the slice-invariant part that run on the simple kernel, grouped
  // per level of the contraction tree into one launch each (their descriptors on
  // the device once the workspace is placed), and the device table of input
  // pointers those launches read (rewritten when a run brings other inputs).
  struct GroupGemm {
    Ref a, b, c;
    int64_t arow, acol, brow, bcol, m, k, n;
    uint64_t block0;
  };
  std::vector<GroupGemm> ggemm_;
  SimpleGemmDesc* ggemm_dev_ = nullptr;
  const void** inputs_dev_ = nullptr;
  std::vector<const void*> inputs_seen_;
  void sync_inputs(const void* const* inputs);
};

// Scratch of one stage operation: bump allocation from a preallocated pool, or,
// without a pool base, stream-ordered hipMallocAsync (freed when the object
// dies); a pool without a base records the operation's peak need instead.
struct ScratchPool {
  char* base = nullptr;
  size_t cap = 0, peak = 0;
};
struct Scratch {
  hipStream_t s;
  ScratchPool* pool;
  size_t used = 0;
  std::vector<void*> ptrs;
  explicit Scratch(hipStream_t st, ScratchPool* p = nullptr) : s(st), pool(p) {}
  Scratch(const Scratch&) = delete;
  Scratch& operator=(const Scratch&) = delete;
  ~Scratch();
  void* get(size_t bytes);
};
// sum_axis over `sums` then permute by `perm` (the GEMM operand of a pair step);
// returns x itself when nothing moves.
const void* stage_operand(Scratch& sc, int dt, const void* x, std::vector<size_t> shape, const std::vector<size_t>& sums,
                          const std::vector<size_t>& perm);
// The adjoint of stage_operand: inverse permutation, then broadcast over each summed axis.
void stage_operand_adjoint(Scratch& sc, int dt, const void* g, std::vector<size_t> shape0,
                           const std::vector<size_t>& sums, const std::vector<size_t>& perm, void* out);

// Stage execution with stream-ordered scratch (src/compiled.rs:60-63, 143-153, 206-209).
void stage_prep_forward(const Plan& p, size_t i, int dt, const void* in, void* out, hipStream_t s);
void stage_prep_adjoint(const Plan& p, size_t i, int dt, const void* g, void* gin, hipStream_t s);
void stage_step_forward(const Plan& p, size_t step, int dt, int prec, const void* l, const void* r, void* out, hipStream_t s);
void stage_step_adjoint(const Plan& p, size_t step, int dt, const void* l, const void* r, const void* g, void* gl, void* gr,
                        hipStream_t s);
void stage_finalize_forward(const Plan& p, int dt, const void* in, void* out, hipStream_t s);
void stage_finalize_adjoint(const Plan& p, int dt, const void* g, void* gin, hipStream_t s);

// Reverse-mode differentiation of an unsliced plan (no reference counterpart):
// forward() computes the output and keeps the checkpoint slots on the device;
// backward() then takes dL/d(conj out) and writes dL/d(conj input) of every
// input, recomputing the intermediates of each checkpoint segment. The inputs
// must be the forward's and unchanged. Tiny steps of one tree level share one
// launch; intermediates live in one planned workspace.
class GradRunner {
 public:
  GradRunner(std::shared_ptr<const Plan> plan, int dt, int prec, size_t checkpoint_every);
  ~GradRunner();
  GradRunner(const GradRunner&) = delete;
  GradRunner& operator=(const GradRunner&) = delete;
  void forward(const void* const* inputs, void* out, hipStream_t s);
  void backward(const void* const* inputs, const void* grad_out, void* const* grad_inputs, hipStream_t s);
  size_t workspace_bytes() const { return ws_bytes_; }
  size_t n_ops() const { return ops_.size(); }
  size_t n_inputs() const { return n_in_; }

 private:
  struct Loc;
  struct Buf {
    size_t bytes;
    int first, last;
  };
  struct GOp {
    enum Kind { PREP, FWD_GROUP, FWD_ONE, FIN, FIN_ADJ, ADJ_GROUP, ADJ_ONE, PREP_ADJ } kind;
    int t;
    size_t index;               // step (FWD_ONE / ADJ_ONE) or input (PREP / PREP_ADJ)
    std::vector<size_t> items;  // steps of a grouped launch
    int vals = 0;               // 0: forward values, 1: backward-pass values
    size_t desc_first = 0, n_desc = 0;
    uint64_t blocks = 0;
  };
  int newbuf(size_t bytes);
  void* at(const Loc& l) const;
  void set_bases(const void* const* inputs, void* const* grad_inputs, void* out, const void* grad_out, hipStream_t s);
  void run_ops(size_t begin, size_t end, hipStream_t s);
  void ensure_pool(hipStream_t s);

  std::shared_ptr<const Plan> P_;
  int dt_, prec_, es_ = 4;
  size_t n_in_ = 0;
  std::vector<Loc> val_f_, val_r_, grad_;
  std::vector<Buf> bufs_;
  std::vector<size_t> offs_;
  std::vector<GOp> ops_;
  size_t fwd_end_ = 0;
  std::vector<int32_t> tab_;
  char* ws_ = nullptr;
  size_t ws_bytes_ = 0;
  void* descs_dev_ = nullptr;
  int32_t* tab_dev_ = nullptr;
  char** bases_dev_ = nullptr;
  char** bases_host_ = nullptr;
  std::vector<char*> bases_;
  size_t n_bases_ = 0;
  ScratchPool pool_;
  bool have_forward_ = false;
  std::vector<const void*> forward_inputs_;
};

// Many independent unsliced plans run in lockstep grouped launches. Planned,
// packed into memory-bounded groups and uploaded once; run() replays them.
class GroupedExecutor {
 public:
  // prec: NATIVE, BF16X3 or FP16X2 (split products on the MFMA-bound steps
  // launched on their own; float32 / complex64 only)
  GroupedExecutor(std::vector<std::shared_ptr<const Plan>> plans, int dt, bool strip, int prec = NATIVE);
  ~GroupedExecutor();
  GroupedExecutor(const GroupedExecutor&) = delete;
  GroupedExecutor& operator=(const GroupedExecutor&) = delete;
  // inputs[p]: device input pointers of plan p; outputs[p]: its output. With
  // strip, log10_exp[p] receives the exponent (result = output * 10^e).
  void run(const std::vector<const void* const*>& inputs, const std::vector<void*>& outputs, double* log10_exp,
           hipStream_t s, Report* rep);
  size_t n_plans() const { return plans_.size(); }
  size_t n_inputs(size_t p) const { return plans_[p]->cp.n_inputs; }

 private:
  struct Group;
  void release();
  std::vector<std::shared_ptr<const Plan>> plans_;
  int dt_;
  bool strip_;
  int prec_;
  std::vector<std::unique_ptr<Group>> groups_;
  char* arena_ = nullptr;
  size_t arena_bytes_ = 0;
};

void grouped_execute(const std::vector<std::shared_ptr<const Plan>>& plans, int dt, bool strip, int prec,
                     const std::vector<const void* const*>& inputs, const std::vector<void*>& outputs, double* log10_exp,
                     hipStream_t s, Report* rep);

}  // namespace atn
