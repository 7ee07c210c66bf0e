// Host-side interface of the gfx950 kernels.
#pragma once

#include "common.hpp"

namespace atn {

constexpr int MAXR = 40;   // max (merged) rank handled by the generic kernels
constexpr int MAXT = 16;   // max axes per tile group of the LDS permute

// ---- permute (src/tensor.rs:453-492), pure copy, bit-exact ----------------
struct PermTile {
  // Outer (grid) axes: one tile per index.
  int n_r;
  uint32_t r_dims[MAXR];
  int64_t r_in[MAXR], r_out[MAXR];
  // Tile U = G_in u G_out. LDS holds U in output-order row-major layout.
  int n_hin, n_gin, n_hout;
  uint32_t hin_dims[MAXT], gin_dims[MAXT], hout_dims[MAXT];
  int64_t hin_in[MAXT], hout_out[MAXT];
  int32_t hin_lds[MAXT], gin_lds[MAXT];
  uint32_t u_size, lo_in, lo_out, n_hi_in, n_hi_out;
  int32_t sh_in, sh_out;  // log2 of lo_in / lo_out when powers of two, else -1
  uint64_t n_tiles;
};
struct PermGeneric {
  int rank;
  uint32_t dims[MAXR];   // output-order dims
  int64_t in_str[MAXR];  // input stride of each output axis
  uint64_t numel;
};
struct PermPlan {
  int kind = 0;  // 0 identity (alias), 1 LDS tile, 2 generic
  int es = 8;    // element bytes
  size_t numel = 0;
  PermTile tile;
  PermGeneric gen;
  unsigned threads = 256;
  size_t lds = 0;
};
// Permute `shape` by `perm` (output axis d = input axis perm[d]) for `batch`
// consecutive tensors.
PermPlan plan_permute(const std::vector<size_t>& shape, const std::vector<size_t>& perm, size_t batch, int es);
void launch_permute(const PermPlan& p, const void* in, void* out, hipStream_t s);

// ---- reductions / gathers (src/tensor.rs:503-669) --------------------------
// out[o, q] = sum_{i<d} in[o, i, q] (sequential order), view [outer, d, inner]
void launch_sum_axis(int dt, size_t outer, size_t d, size_t inner, const void* in, void* out, hipStream_t s);
// view [A, d, Bm, d, C]; trace -> [A, Bm, C]; diag -> [A, d, Bm, C]
void launch_trace(int dt, size_t A, size_t d, size_t Bm, size_t C, const void* in, void* out, hipStream_t s);
void launch_diag(int es, size_t A, size_t d, size_t Bm, size_t C, const void* in, void* out, hipStream_t s);

// Slice gather: out[b, e] = in[base(g0 + b) + off(e)].
struct GatherArgs {
  int n_ax;
  uint32_t dims[MAXR];
  int64_t str[MAXR];
  int n_sel;
  uint64_t sel_div[MAXR], sel_mod[MAXR];
  int64_t sel_str[MAXR];
  uint64_t inner;
};
void launch_gather(int es, const GatherArgs& a, uint64_t g0, size_t nb, const void* in, void* out, hipStream_t s);

// acc64[blk, e] += sum over slices g in [g0, g0+nb) of block blk of part[g-g0, e]
void launch_reduce_f64(int dt, const void* part, void* acc64, uint64_t g0, size_t nb, size_t n_internal, size_t numel,
                       hipStream_t s);
// out[scatter(blk, e)] = (T) src[blk, e]; src_f64 selects a double accumulator as the source.
struct ScatterArgs {
  int n_blk;
  uint64_t blk_div[MAXR], blk_mod[MAXR];
  int64_t blk_str[MAXR];
  int n_kept;
  uint32_t kept_dims[MAXR];
  int64_t kept_str[MAXR];
};
void launch_scatter(int dt, bool src_f64, const ScatterArgs& a, uint64_t blk0, size_t n_blk, size_t numel,
                    const void* src, void* out, hipStream_t s);
// acc[e] (+)= part[e] in T (first = copy)
void launch_accum(int dt, bool first, const void* part, void* acc, size_t numel, hipStream_t s);
void launch_fill_zero(void* p, size_t bytes, hipStream_t s);
void launch_copy(size_t bytes, const void* in, void* out, hipStream_t s);
// Adjoint helpers: out[o, i, q] = in[o, q] (broadcast of a summed axis), and
// the transpose of trace / diagonal extraction: out (pre-zeroed) [A, d, Bm, d, C]
// receives g on its diagonal; g is [A, Bm, C] (trace) or [A, d, Bm, C] (diag).
void launch_broadcast_axis(int es, size_t outer, size_t d, size_t inner, const void* in, void* out, hipStream_t s);
void launch_diag_scatter(int es, bool trace, size_t A, size_t d, size_t Bm, size_t C, const void* g, void* out, hipStream_t s);

// ---- strip (src/compiled.rs:412-430), power-of-two variant ----------------
// maxbits: one uint64 slot (non-negative double bits); expo: int64 running exponent.
void launch_maxabs(int dt, const void* x, size_t numel, uint64_t* maxbits, hipStream_t s);
// keep > 0: leave x (and expo) unchanged while the max's exponent e lies in
// [-keep, keep] — the power-of-two scale is only needed against over/underflow.
// scaled_max (optional): receives the max of x after the call as double bits.
// pend (optional): x is left as it is; the exponent e of the rescale goes to
// *pend (and is added to expo as usual) for x's reader to divide out
// (GemmArgs::apend / bpend), and scaled_max receives the max of x unchanged.
void launch_scale_pow2(int dt, void* x, size_t numel, const uint64_t* maxbits, int64_t* expo, hipStream_t s,
                       int keep = 0, uint64_t* scaled_max = nullptr, int64_t* pend = nullptr);
// The last step of a stripped run: divide by the exact max magnitude m (Scalar::abs,
// f64) and add log10(m) to log10_exp (src/compiled.rs:412-430); maxbits is scratch.
void launch_strip_exact(int dt, void* x, size_t numel, uint64_t* maxbits, double* log10_exp, hipStream_t s);
// Exponent bound below which intermediate results of a stripped run are left unscaled.
constexpr int STRIP_KEEP = 32;

// ---- GEMM -------------------------------------------------------------------
struct GemmArgs {
  const void* a;
  const void* b;
  void* c;
  int64_t m, k, n;
  int64_t lda, ldb, ldc;       // row strides in elements
  int64_t sao, sai, sbo, sbi, sco, sci;  // batch strides (outer slice batch, inner plan batch)
  int64_t nbo, nbi;
  // Device scratch (or null) of scratch_bytes: the FP16X2 operand max-magnitude
  // bits (first 16 bytes) and split-K partial sums (from byte 256). With less
  // scratch than gemm_scratch_bytes() asks for, split-K uses fewer k ranges.
  void* scratch = nullptr;
  size_t scratch_bytes = 0;
  // Set by launch_gemm for the MFMA kernel: k range length per blockIdx.z (0 =
  // all of k) and the partial-sum buffer when k is split.
  int64_t kc = 0;
  void* part = nullptr;
  // Optional (zeroed) slot receiving max |re|, |im| over C as double bits: the
  // strip's max pass folded into the GEMM epilogue (launch_gemm adds a separate
  // pass after the kernels that do not fold it). C must be contiguous.
  unsigned long long* cmax = nullptr;
  // Optional gather tables: B element (p, j) of a batch entry sits at
  // brow[p] + bcol[j] (elements) instead of p * ldb + j, so B can be read from
  // an operand that was never permuted into [k, n] order. Supported by the
  // simple, streaming, split-K, MFMA and split-precision kernels.
  const int32_t* brow = nullptr;
  const int32_t* bcol = nullptr;
  // bcol holds runs of this many consecutive offsets, each starting at a multiple
  // of the run length (vector loads stay within a run)
  int32_t bcol_run = 1;
  // brow holds runs of this many consecutive offsets, each starting at a multiple
  // of the run length (B's contracted legs innermost in memory)
  int32_t brow_run = 1;
  // B columns as a mixed-radix function of the column index, for operands too
  // large for the int32 gather tables: column
  // j sits at sum_q ((j / (bcd[0] ... bcd[q-1])) % bcd[q]) * bcs[q] elements (group
  // 0 innermost, bcg groups), row p at brow64[p]; read by the streaming and the
  // FP32 MFMA kernels (float32 / complex64)
  int32_t bcg = 0;
  int64_t bcd[6] = {}, bcs[6] = {};
  // division by bcd[q] as a multiply-high (round-up method): q = (h + ((j - h) >> 1))
  // >> (bcl[q] - 1) with h = umulhi(bcm[q], j); bcl[q] = 0 for bcd[q] = 1
  uint32_t bcm[6] = {};
  int32_t bcl[6] = {};
  const int64_t* brow64 = nullptr;
  // FP16X2: max |re|, |im| of A / of B as double bits when already known (folded
  // into the epilogue of the GEMM that produced the operand); null = measured by
  // a pass before the GEMM. An upper bound is enough: it only sets the scale.
  const unsigned long long* amax = nullptr;
  const unsigned long long* bmax = nullptr;
  // Optional pending power-of-two exponents (int64) of A / of B: the operand holds
  // its values times 2^pend (a stripped result whose rescale was left to its
  // reader), and the product is divided by 2^(apend + bpend) in the same exact
  // output scale as the FP16 operand scales. Read by the split-precision kernels
  // only (launch_gemm rejects them elsewhere).
  const int64_t* apend = nullptr;
  const int64_t* bpend = nullptr;
  // Optional gather tables of A, read by the simple kernel only (launch_gemm
  // runs that kernel when they are set): A element (i, p) of a batch entry sits
  // at arow[i] + acol[p] instead of i * lda + p.
  const int32_t* arow = nullptr;
  const int32_t* acol = nullptr;
};
// GEMM_PRESPLIT: split-precision MFMA GEMM whose operands are split into
// 16-bit (hi, lo) arrays by a separate pass (GEMM_MFMA splits them while staging
// each tile); for native precision it is GEMM_MFMA. Its row tile is 32 rows of
// A for m <= 32, else 128.
// GEMM_MFMA_R32 / R64 / R128: the FP32 / FP64 MFMA kernel with that row tile
// whatever the precision mode (GEMM_MFMA picks the row tile from m, or splits
// the operands while staging under split precision; FP64 tiles have at most
// 64 rows).
// GEMM_PRESPLIT_A: the GEMM_PRESPLIT kernel with only A split by a pass; B is
// read in its own precision through the gather tables and split while staging.
enum GemmKernel {
  GEMM_AUTO = -1, GEMM_SIMPLE = 0, GEMM_MFMA = 1, GEMM_SKINNY = 2, GEMM_SPLITK = 3, GEMM_PRESPLIT = 4,
  GEMM_MFMA_R32 = 5, GEMM_MFMA_R64 = 6, GEMM_MFMA_R128 = 7, GEMM_PRESPLIT_A = 8
};
// Returns the kernel used.
int launch_gemm(int dt, int prec, int kernel, const GemmArgs& g, hipStream_t s);
// Device scratch launch_gemm can use for this GEMM (FP16X2 scales, split-K partials).
size_t gemm_scratch_bytes(int dt, int prec, int kernel, const GemmArgs& g);
// The multiply-high constants of GemmArgs::bcm / bcl for dividing 32-bit values by d.
void fast_div_u32(uint32_t d, uint32_t& mul, int32_t& shift);
// Whether the heuristic runs an unbatched m x k x n product on a kernel that reads
// a factored B (GemmArgs::bcg): float32 / complex64 on the FP32 MFMA kernels (not
// the split or 3M forms) or the streaming kernel.
bool gemm_reads_factored(int dt, int prec, int64_t m, int64_t k, int64_t n);
// Whether the heuristic runs an unbatched m x k x n product on the streaming kernel.
bool gemm_auto_streams(int dt, int prec, int64_t m, int64_t k, int64_t n);
// Whether this GEMM runs on the simple kernel under kernel choice `kernel`.
bool gemm_runs_simple(int dt, int prec, int kernel, const GemmArgs& g);
// Whether the heuristic choice for an m x k x n product (element units) runs
// split-precision products under precision prec.
bool split_products_apply(int dt, int prec, int64_t m, int64_t k, int64_t n);
// Many small independent GEMMs (C = A B, one batch entry each, row-major C) in one
// launch of the simple kernel. a, b: workspace byte offsets (>= 0) or -(input
// index + 1); c: a workspace byte offset. A(i, p) = A[arow[i] + acol[p]] or
// A[i * k + p]; B(p, j) = B[brow[p] + bcol[j]] or B[p * n + j] (offsets into the
// gather tables, or -1). block0: first 256-thread block of the GEMM in the launch.
struct SimpleGemmDesc {
  int64_t a, b, c;
  int64_t arow, acol, brow, bcol;
  int64_t m, k, n;
  uint64_t block0;
};
void launch_gemm_simple_group(int dt, const SimpleGemmDesc* d, int n, uint64_t blocks, char* ws,
                              const void* const* inputs, const int32_t* tables, hipStream_t s);
// Conjugating variants for the adjoint: opA/opB: 0 = N, 1 = conj-transpose (as stored).
void launch_gemm_adjoint(int dt, int opa, int opb, const GemmArgs& g, hipStream_t s);

extern thread_local size_t g_launch_count;
// Kind of the last kernel launched on this thread (profiling labels).
enum KernelKind { K_OTHER = 0, K_PERMUTE_TILE, K_PERMUTE_GENERIC, K_SUM, K_TRACE_DIAG, K_GATHER, K_REDUCE, K_SCATTER,
                  K_GEMM_SIMPLE, K_GEMM_SMALL, K_GEMM_MFMA_F32, K_GEMM_MFMA_F64, K_GEMM_MFMA_BF16X3, K_GEMM_MFMA_FP16X2, K_GEMM_MFMA_C3M,
                  K_GEMM_SPLITK, K_STRIP, K_COPY, K_NKINDS };
extern thread_local int g_last_kind;
const char* kernel_kind_name(int k);

}  // namespace atn
