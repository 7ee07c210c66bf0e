/*
 * arctn_hip.h - C ABI of the MI355X (gfx950) executor for ArcTN contraction plans.
 *
 * The executor takes the same plan input as the reference's CPU executor
 * (network legs/dims, SSA path, sliced legs) and runs it on one GPU with
 * hand-written HIP kernels. All tensors handed across this ABI are DEVICE
 * pointers, row-major (C order) in net.inputs leg order, complex values as
 * interleaved (re, im) pairs - the layout of src/tensor.rs:340-344 and of
 * NumPy complex64/complex128.
 *
 * Conventions (the reference engine-interface rule "must not unwind across the
 * C boundary", docs/engine-interface.md):
 *   - every entry point returns ATN_OK (0) or a negative status; the message of
 *     the last failure on the calling thread is atn_last_error();
 *   - all validation happens before any kernel is launched;
 *   - strings returned through char** are released with atn_free();
 *   - plans are immutable after creation and may be shared between threads;
 *     executors are single-threaded objects (one per stream).
 */
#ifndef ARCTN_HIP_H
#define ARCTN_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ATN_ABI_VERSION 1

enum {
  ATN_OK = 0,
  ATN_ERR_INVALID = -1,  /* malformed network, path, slice set, shapes, options */
  ATN_ERR_DEVICE = -2,   /* HIP runtime failure (allocation, launch, sync) */
  ATN_ERR_INTERNAL = -3
};

/* Element types: the four dtypes of the reference (pybind/src/lib.rs:234-239). */
typedef enum { ATN_F32 = 0, ATN_F64 = 1, ATN_C64 = 2, ATN_C128 = 3 } atn_dtype;

/* Arithmetic of the contraction GEMMs.
 *   NATIVE: FP32 MFMA for f32/complex64 (complex as 4M: a real GEMM against the
 *           2x2 real expansion of B), FP64 MFMA for f64/complex128.
 *   C3M:    complex 3M (Gauss) on FP32 MFMA; complex64 only.
 *   BF16X3: complex64/f32 products as hi*hi + hi*lo + lo*hi in BF16 MFMA, FP32 accumulate.
 *   FP16X2: same with FP16 hi/lo halves and per-step power-of-two operand rescaling.
 *   BF16X3_3M / FP16X2_3M: BF16X3 / FP16X2 whose complex64 products run as
 *           Gauss's three real products (9 instead of 12 16-bit products per
 *           complex block) where the 3M split kernel applies (A split by a pass,
 *           B read in complex64 with lanes along columns, 128-row tiles); other
 *           steps, and float32 plans, run as BF16X3 / FP16X2. */
typedef enum {
  ATN_PREC_NATIVE = 0, ATN_PREC_C3M = 1, ATN_PREC_BF16X3 = 2, ATN_PREC_FP16X2 = 3,
  ATN_PREC_BF16X3_3M = 4, ATN_PREC_FP16X2_3M = 5
} atn_precision;

/* Network + plan, as flat arrays (the fields of the execution-plan v2 artifact,
 * src/execution_plan.rs:39-43). Legs are u32 ids (src/network.rs:13). */
typedef struct {
  size_t n_inputs;
  const size_t* input_lens;    /* n_inputs entries: rank of each input */
  const uint32_t* input_legs;  /* concatenated legs, sum(input_lens) entries */
  size_t n_output;
  const uint32_t* output;      /* open legs in output order */
  size_t n_dims;
  const uint32_t* dim_legs;    /* size_dict keys */
  const size_t* dim_sizes;     /* size_dict values */
  size_t n_steps;
  const size_t* ssa_path;      /* 2 * n_steps entries: (left, right) SSA ids */
  size_t n_sliced;
  const uint32_t* sliced_legs; /* may include output legs (output-index slicing) */
} atn_network_desc;

typedef struct atn_plan atn_plan;
typedef struct atn_executor atn_executor;

typedef struct {
  size_t n_inputs;
  size_t n_steps;          /* pair steps of the per-slice plan */
  size_t n_slices;         /* product of all sliced dims (internal x output blocks) */
  size_t n_blocks;         /* output-index blocks (1 without output slicing) */
  size_t n_varying_steps;  /* steps whose operands depend on the slice index */
  double log10_flops;      /* reference convention: sum over steps of prod(union dims), whole network */
  double log2_peak_size;   /* live-set peak of the whole network (src/path.rs:396-399) */
  double macs_per_slice;   /* multiply-adds of the GEMMs of one slice: sum over steps of bt*m*k*n in the
                              plan's scalar type (a plan has no dtype); real FLOPs are 8x this for
                              complex and 2x for real execution */
  double mults_per_slice;  /* scalar multiplications of one slice (reference convention) */
} atn_plan_info;

typedef struct {
  int dtype;            /* atn_dtype */
  int precision;        /* atn_precision */
  int batched;          /* 0: plain per-slice path (one slice, one step at a time);
                           1: all slices of a step in one strided-batched launch,
                              slice-invariant steps hoisted, FP64 cross-slice sum */
  int use_graph;        /* batched only: capture the whole plan in one hipGraph */
  int autotune;         /* time candidate GEMM kernels per step on first use */
  int strip;            /* unsliced plans: power-of-two normalisation after every
                           step, exponent kept in FP64 (execute_stripped) */
  size_t max_batch;     /* slices per batched launch; 0 = largest that fits workspace_limit */
  size_t workspace_limit;  /* bytes; 0 = 60% of free device memory */
  size_t slice_begin;   /* batched form: run global slices [slice_begin, slice_end) only and */
  size_t slice_end;     /* return their partial sum (the share of one device); 0, 0 = all */
  int permute_b;        /* 1: always permute GEMM B operands into [k, n] order; 0: read them
                           through gather tables where the layout allows (no copy) */
  int gemm_kernel;      /* 0: GEMM kernel chosen per step (heuristic, or timed with autotune);
                           k + 1: atn_gemm kernel code k at every step (overrides autotune) */
} atn_exec_options;

typedef struct {
  double elapsed_ms;        /* device time of the execution (hipEvents on the stream) */
  double flops;             /* algorithmic real FLOPs executed (all slices) */
  double tflops;
  double slices_per_s;
  double log10_exponent;    /* strip: result = output * 10^log10_exponent */
  double max_abs_result;
  size_t n_slices;
  size_t batch;             /* slices per batched launch actually used */
  size_t n_launches;        /* kernel launches of one execution */
  size_t workspace_bytes;
} atn_report;

int atn_abi_version(void);
const char* atn_last_error(void);
void atn_free(void* p);

/* Replaces CompiledContraction::compile (src/compiled.rs:254-366) together with
 * the slice bookkeeping of contract_network_sliced (src/slice.rs:960-1016,
 * 1070-1137). Validates the network (src/network.rs:171-229), the path
 * (src/path.rs:239-434) and the slice set (src/slice.rs:146-172). */
int atn_plan_create(const atn_network_desc* desc, atn_plan** out);
void atn_plan_destroy(atn_plan* plan);
int atn_plan_stats(const atn_plan* plan, atn_plan_info* out);
/* JSON mirror of input_preps()/pair_steps()/finalize() of the per-slice plan
 * (src/compiled.rs:241-250, 102-153): slots, legs, shapes, sum axes, perms,
 * (bt, m, k, n). */
int atn_plan_describe(const atn_plan* plan, char** json_out);

/* Executors own device workspace, per-step kernel choices and the captured
 * graph; create one per (plan, options) and reuse it. */
int atn_executor_create(const atn_plan* plan, const atn_exec_options* opts, atn_executor** out);
void atn_executor_destroy(atn_executor* ex);
/* Replaces contract_network_sliced (src/slice.rs:960-1065) and
 * CompiledContraction::execute / execute_stripped (src/compiled.rs:369-381):
 * d_inputs[i] is input i (full, unsliced shape), d_out receives the full output
 * in net.output order. stream may be NULL (default stream). Returns after the
 * work is enqueued and, when report != NULL, completed and timed. */
int atn_executor_run(atn_executor* ex, const void* const* d_inputs, void* d_out, void* stream, atn_report* report);

/* Per-kernel accounting of one eager execution with HIP events around every
 * executor op on `stream` (used for roofline reporting; no reference
 * counterpart). stats[k] for kernel kind k (the kind of an op's last kernel):
 * name, ops, kernel launches, device ms, algorithmic FLOPs and HBM bytes
 * (operands read + results written). *n_stats <= max_stats kinds. */
typedef struct {
  char name[32];
  size_t calls;     /* executor ops (one GEMM, one permute, ...) */
  size_t launches;  /* kernel launches they made (a split-precision GEMM makes several) */
  double ms;
  double flops;
  double bytes;
} atn_kernel_stat;
int atn_executor_profile(atn_executor* ex, const void* const* d_inputs, void* d_out, void* stream,
                         atn_kernel_stat* stats, size_t max_stats, size_t* n_stats);

/* One-shot convenience: create an executor, run, destroy. */
int atn_execute(const atn_plan* plan, const atn_exec_options* opts, const void* const* d_inputs, void* d_out,
                void* stream, atn_report* report);

/* Stage execution of the per-slice plan, the hook the reference exposes for
 * autograd tapes (src/compiled.rs:7-10, 60-63, 143-153, 206-209):
 *   atn_prep_forward:     input i (slot i) <- prep(input)
 *   atn_step_forward:     slot n_inputs+s <- step s (left slot, right slot)
 *   atn_finalize_forward: output <- finalize(slot fin)
 * and their adjoints (no reference counterpart): given dL/d(conj result) they
 * produce dL/d(conj operand) with conjugate-transposed GEMMs, inverse
 * permutations and broadcasts over summed legs. Tensors have the per-slice
 * shapes reported by atn_plan_describe. Scratch memory is stream-ordered. */
int atn_prep_forward(const atn_plan* plan, size_t input, int dtype, const void* in, void* out, void* stream);
int atn_prep_adjoint(const atn_plan* plan, size_t input, int dtype, const void* grad_out, void* grad_in, void* stream);
int atn_step_forward(const atn_plan* plan, size_t step, int dtype, int precision, const void* left, const void* right,
                     void* out, void* stream);
int atn_step_adjoint(const atn_plan* plan, size_t step, int dtype, const void* left, const void* right,
                     const void* grad_out, void* grad_left, void* grad_right, void* stream);
int atn_finalize_forward(const atn_plan* plan, int dtype, const void* in, void* out, void* stream);
int atn_finalize_adjoint(const atn_plan* plan, int dtype, const void* grad_out, void* grad_in, void* stream);

/* Reverse-mode differentiation of an unsliced plan in two calls (no reference
 * counterpart: the reference has no gradient code, and its stage hook,
 * src/compiled.rs:7-10, is what a tape would drive). forward writes the output
 * and keeps the checkpoint slots on the device; backward takes dL/d(conj out)
 * (PyTorch's complex convention) and writes dL/d(conj input) for every input,
 * recomputing the intermediates of each checkpoint segment (checkpoint_every
 * steps in contraction-tree level order; 0 keeps every intermediate). backward
 * needs the inputs of the last forward, unchanged. Pair steps run level by level,
 * the small ones of a level in one launch; precision applies to the large
 * forward products (the adjoint products run in the dtype's native MFMA). */
typedef struct atn_adjoint atn_adjoint;
int atn_adjoint_create(const atn_plan* plan, int dtype, int precision, size_t checkpoint_every, atn_adjoint** out);
void atn_adjoint_destroy(atn_adjoint* a);
int atn_adjoint_forward(atn_adjoint* a, const void* const* d_inputs, void* d_out, void* stream);
int atn_adjoint_backward(atn_adjoint* a, const void* const* d_inputs, const void* d_grad_out, void* const* d_grad_inputs,
                         void* stream);

/* Element-level kernels, exported for parity tests of the data-movement path
 * (src/tensor.rs:453-669): op 0 permute(arg = axes), 1 sum_axis(arg[0]),
 * 2 trace_pair(arg[0], arg[1]), 3 diag_pair(arg[0], arg[1]),
 * 4 select_axis(arg[0], arg[1]), 5 max magnitude (max over the elements of
 * max(|re|, |im|); NaN components never win) as one double in d_out, the value
 * behind the power-of-two rescales of the strip and the FP16 operand scales
 * (no reference counterpart). Device pointers; output shape implied. */
int atn_tensor_op(int dtype, int op, size_t rank, const size_t* shape, const size_t* arg, size_t n_arg,
                  const void* d_in, void* d_out, void* stream);

/* Batched strided GEMM used by the executor, exported for kernel tests:
 * C[b] = A[b] . B[b], row-major, batch strides in elements (0 = broadcast).
 * kernel: -1 automatic choice, 0 one-thread-per-element kernel, 1 MFMA kernel,
 * 2 skinny kernel (m <= 32 with n >= 64, or n <= 32 with m >= 64; else falls back to 0),
 * 3 split-K kernel (workgroups split k; partial sums added in a fixed order),
 * 4 split-precision MFMA kernel on operands pre-split into 16-bit (hi, lo) arrays
 *   by a separate pass (for native precision: 1), 5 / 6 / 7 native MFMA kernel
 *   with a 32 / 64 / 128-row tile (1 picks the row tile from m), 8 the kernel of 4
 *   with only A pre-split and B split as it is staged (B stays in one read; 4 where
 *   B's layout does not allow it). */
int atn_gemm(int dtype, int precision, int kernel, size_t batch, size_t m, size_t k, size_t n, const void* d_a,
             size_t stride_a, const void* d_b, size_t stride_b, void* d_c, size_t stride_c, void* stream);

/* No counterpart in the reference. atn_gemm for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc): same arguments, same kernels, same results. The
 * scratch a kernel choice needs (split-K partial sums, pre-split 16-bit operand
 * halves, FP16X2 operand maxima) comes from a fixed device array of
 * ATN_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends, instead of
 * hipMallocAsync; the call allocates nothing. A GEMM needing more scratch than
 * that fails with ATN_ERR_INVALID before anything is launched. Every call on
 * one GPU shares that GPU's array: calls that overlap in time on one GPU (other
 * streams, other threads) race. */
#define ATN_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
int atn_gemm_noalloc(int dtype, int precision, int kernel, size_t batch, size_t m, size_t k, size_t n,
                     const void* d_a, size_t stride_a, const void* d_b, size_t stride_b, void* d_c, size_t stride_c,
                     void* stream);

/* Grouped execution of many independent unsliced plans (uneven instances):
 * every plan's steps run in lockstep rounds, each round one grouped launch per
 * kernel kind. inputs[p] is the d_inputs array of plan p; outputs[p] its output.
 * With strip != 0 every step result is renormalised by a power of two and
 * log10_exponents[p] receives the exponent (result = output * 10^e). */
int atn_grouped_execute(const atn_plan* const* plans, size_t n_plans, int dtype, int precision, int strip,
                        const void* const* const* inputs, void* const* outputs, double* log10_exponents,
                        void* stream, atn_report* report);

/* The same grouped execution as a reusable object: plans are compiled into
 * grouped launch descriptors, packed into memory-bounded groups and given a
 * device arena once at creation; every run replays them with new input and
 * output pointers. Replaces repeated CompiledContraction::execute /
 * execute_stripped calls (src/compiled.rs:369-381) over many plans. */
typedef struct atn_grouped atn_grouped;
int atn_grouped_create(const atn_plan* const* plans, size_t n_plans, int dtype, int precision, int strip,
                       atn_grouped** out);
void atn_grouped_destroy(atn_grouped* g);
int atn_grouped_run(atn_grouped* g, const void* const* const* inputs, void* const* outputs, double* log10_exponents,
                    void* stream, atn_report* report);

#ifdef __cplusplus
}
#endif
#endif /* ARCTN_HIP_H */
