// Grouped execution of many independent unsliced plans (uneven instances).
//
// Each plan keeps the semantics of CompiledContraction::execute /
// execute_stripped (src/compiled.rs:383-441): input preps, pair steps in SSA
// order, optional per-step renormalisation, finalize. The plans advance in
// lockstep by step index: round r runs step r of every plan that has one, and
// each round issues one persistent grouped launch per kernel kind (permutes,
// GEMMs, strip max, strip scale) over all small problems of the round. A
// workgroup walks the round's virtual blocks with a grid stride and finds the
// problem of each one by binary search over the problems' first-block prefix,
// so one launch covers problems of any shapes. Problems above the size limits
// below are launched on their own with the LDS-tiled permute, the MFMA /
// skinny / split-K GEMM kernels and the grid-stride strip kernels, which is
// where the large steps' FLOPs and bytes go.
//
// Device addresses are (base, byte offset) pairs resolved through a per-group
// table of base pointers: base 0 is the group's arena, then every plan's
// inputs, then every plan's output. Plans are packed into groups whose arenas
// fit the workspace budget; groups run one after another.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "executor.hpp"

namespace atn {

namespace {

struct GRef {
  uint32_t base;
  uint64_t off;  // bytes
};
struct GAxis {
  int64_t in_stride;  // input stride (elements) of this output axis
  uint32_t dim;
  uint32_t pad;
};
struct GPermDesc {
  GRef in, out;
  uint64_t numel, block0;
  uint32_t rank, axis0;  // axes [axis0, axis0 + rank) of the round's axis table, output order
};
struct GGemmDesc {
  GRef a, b, c;
  int64_t m, k, n, bt;
  uint64_t block0;
  uint32_t lg_lanes;  // log2 of the lanes that share one output's k loop
};
struct GStripDesc {
  GRef c;
  uint64_t numel, block0;
  uint32_t plan;   // index of the plan within the group (strip slots)
  uint32_t final;  // the plan's last step: always normalise (else only outside 2^+-STRIP_KEEP)
};

constexpr unsigned GRID_CAP = 4096;  // persistent grid: 16 workgroups per CU

__device__ inline char* gres(const GRef& r, char* const* bases) { return bases[r.base] + r.off; }

template <class D>
__device__ inline const D& find_problem(const D* d, int n, uint64_t blk) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (d[mid].block0 <= blk) lo = mid;
    else hi = mid - 1;
  }
  return d[lo];
}

template <class E>
__global__ __launch_bounds__(256) void grouped_permute_kernel(const GPermDesc* __restrict__ d, int n, uint64_t total,
                                                              const GAxis* __restrict__ axes, char* const* bases) {
  for (uint64_t vb = blockIdx.x; vb < total; vb += gridDim.x) {
    const GPermDesc& p = find_problem(d, n, vb);
    uint64_t e = (vb - p.block0) * 256ull + threadIdx.x;
    if (e >= p.numel) continue;
    const E* in = reinterpret_cast<const E*>(gres(p.in, bases));
    E* out = reinterpret_cast<E*>(gres(p.out, bases));
    uint64_t r = e;
    int64_t off = 0;
    for (int q = (int)p.rank - 1; q >= 0; --q) {
      const GAxis& ax = axes[p.axis0 + q];
      uint64_t t = r / ax.dim;
      off += (int64_t)(r - t * ax.dim) * ax.in_stride;
      r = t;
    }
    out[e] = in[off];
  }
}

template <class T> struct GA;
template <> struct GA<float> {
  static __device__ float zero() { return 0.f; }
  static __device__ float mac(float c, float a, float b) { return c + a * b; }
  static __device__ double mag(float a) { return fabs((double)a); }
  static __device__ double abs(float a) { return fabs((double)a); }
  static __device__ float scale(float a, double s) { return a * (float)s; }
};
template <> struct GA<double> {
  static __device__ double zero() { return 0.0; }
  static __device__ double mac(double c, double a, double b) { return c + a * b; }
  static __device__ double mag(double a) { return fabs(a); }
  static __device__ double abs(double a) { return fabs(a); }
  static __device__ double scale(double a, double s) { return a * s; }
};
template <> struct GA<float2> {
  static __device__ float2 zero() { return make_float2(0.f, 0.f); }
  static __device__ float2 mac(float2 c, float2 a, float2 b) {
    return make_float2(c.x + (a.x * b.x - a.y * b.y), c.y + (a.x * b.y + a.y * b.x));
  }
  static __device__ double mag(float2 a) { return fmax(fabs((double)a.x), fabs((double)a.y)); }
  static __device__ double abs(float2 a) { return hypot((double)a.x, (double)a.y); }
  static __device__ float2 scale(float2 a, double s) { return make_float2(a.x * (float)s, a.y * (float)s); }
};
template <> struct GA<double2> {
  static __device__ double2 zero() { return make_double2(0.0, 0.0); }
  static __device__ double2 mac(double2 c, double2 a, double2 b) {
    return make_double2(c.x + (a.x * b.x - a.y * b.y), c.y + (a.x * b.y + a.y * b.x));
  }
  static __device__ double mag(double2 a) { return fmax(fabs(a.x), fabs(a.y)); }
  static __device__ double abs(double2 a) { return hypot(a.x, a.y); }
  static __device__ double2 scale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
};

template <class T> struct GS;
template <> struct GS<float> {
  static __device__ float down(float v, int o, int w) { return __shfl_down(v, o, w); }
  static __device__ float add(float a, float b) { return a + b; }
};
template <> struct GS<double> {
  static __device__ double down(double v, int o, int w) { return __shfl_down(v, o, w); }
  static __device__ double add(double a, double b) { return a + b; }
};
template <> struct GS<float2> {
  static __device__ float2 down(float2 v, int o, int w) { return make_float2(__shfl_down(v.x, o, w), __shfl_down(v.y, o, w)); }
  static __device__ float2 add(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
};
template <> struct GS<double2> {
  static __device__ double2 down(double2 v, int o, int w) { return make_double2(__shfl_down(v.x, o, w), __shfl_down(v.y, o, w)); }
  static __device__ double2 add(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
};

// 2^lg_lanes consecutive lanes per C element: lane l accumulates k indices
// l, l + L, ... in order, then the lanes are added by a fixed shuffle tree.
// A group of L lanes never straddles a wave (L <= 64 and block0 is 256-aligned).
template <class T>
__global__ __launch_bounds__(256) void grouped_gemm_kernel(const GGemmDesc* __restrict__ d, int n, uint64_t total,
                                                           char* const* bases) {
  for (uint64_t vb = blockIdx.x; vb < total; vb += gridDim.x) {
    const GGemmDesc& p = find_problem(d, n, vb);  // uniform across the workgroup
    const int lg = (int)p.lg_lanes, L = 1 << lg;
    const uint64_t t = (vb - p.block0) * 256ull + threadIdx.x;
    const uint64_t e = t >> lg;
    const int lane = (int)(t & (uint64_t)(L - 1));
    const uint64_t mn = (uint64_t)p.m * p.n;
    const bool valid = e < mn * p.bt;
    T acc = GA<T>::zero();
    if (valid) {
      uint64_t z = e / mn, r = e - z * mn, i = r / p.n, j = r - i * p.n;
      const T* A = reinterpret_cast<const T*>(gres(p.a, bases)) + z * p.m * p.k + i * p.k;
      const T* B = reinterpret_cast<const T*>(gres(p.b, bases)) + z * p.k * p.n + j;
      for (int64_t q = lane; q < p.k; q += L) acc = GA<T>::mac(acc, A[q], B[q * p.n]);
    }
    for (int o = L >> 1; o > 0; o >>= 1) acc = GS<T>::add(acc, GS<T>::down(acc, o, L));
    if (valid && lane == 0) reinterpret_cast<T*>(gres(p.c, bases))[e] = acc;
  }
}

template <class T>
__global__ __launch_bounds__(256) void grouped_maxabs_kernel(const GStripDesc* __restrict__ d, int n, uint64_t total,
                                                             char* const* bases, unsigned long long* slots) {
  for (uint64_t vb = blockIdx.x; vb < total; vb += gridDim.x) {
    const GStripDesc& p = find_problem(d, n, vb);  // uniform across the workgroup
    uint64_t e = (vb - p.block0) * 256ull + threadIdx.x;
    __shared__ double wmax[4];
    double m = 0.0;
    // a plan's last step is divided by its exact max magnitude, the others by powers of two
    if (e < p.numel) {
      const T v = reinterpret_cast<const T*>(gres(p.c, bases))[e];
      m = p.final ? GA<T>::abs(v) : GA<T>::mag(v);
    }
    for (int o = 32; o > 0; o >>= 1) {
      double t = __shfl_down(m, o, 64);
      m = t > m ? t : m;  // NaN never wins; a NaN result stays unscaled
    }
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
      m = wmax[0];
      for (int w = 1; w < 4; ++w) m = wmax[w] > m ? wmax[w] : m;
      // the slot only grows: skip the atomic unless this block's max is larger
      // than what the slot holds (a stale read only costs an atomic)
      const unsigned long long b = (unsigned long long)__double_as_longlong(m);
      if (m > 0.0 && b > __hip_atomic_load(&slots[p.plan], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
        atomicMax(&slots[p.plan], b);
    }
    __syncthreads();
  }
}

template <class T>
__global__ __launch_bounds__(256) void grouped_scale_kernel(const GStripDesc* __restrict__ d, int n, uint64_t total,
                                                            char* const* bases, const unsigned long long* slots,
                                                            long long* expo, double* log10_last) {
  for (uint64_t vb = blockIdx.x; vb < total; vb += gridDim.x) {
    const GStripDesc& p = find_problem(d, n, vb);
    double m = __longlong_as_double((long long)slots[p.plan]);
    if (!(m > 0.0) || !isfinite(m)) continue;
    uint64_t e = (vb - p.block0) * 256ull + threadIdx.x;
    T* c = reinterpret_cast<T*>(gres(p.c, bases));
    if (p.final) {  // src/compiled.rs:412-430: x * from_real(1 / m), exponent + log10(m)
      if (vb == p.block0 && threadIdx.x == 0) log10_last[p.plan] += log10(m);
      if (e < p.numel) c[e] = GA<T>::scale(c[e], 1.0 / m);
      continue;
    }
    int ex;
    frexp(m, &ex);
    if (ex >= -STRIP_KEEP && ex <= STRIP_KEEP) continue;
    if (vb == p.block0 && threadIdx.x == 0) expo[p.plan] += ex;
    if (e < p.numel) c[e] = GA<T>::scale(c[e], ldexp(1.0, -ex));
  }
}

__global__ void grouped_zero_kernel(unsigned long long* p, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) p[i] = 0ull;
}

template <class F>
void gdispatch(int dt, F&& f) {
  switch (dt) {
    case F32: f(float{}); break;
    case F64: f(double{}); break;
    case C64: f(float2{}); break;
    case C128: f(double2{}); break;
    default: throw Error("unsupported dtype code");
  }
}

// Problems at or above these sizes are launched on their own.
constexpr uint64_t LARGE_NUMEL = 1ull << 20;  // permute / GEMM output / strip elements
constexpr uint64_t LARGE_MACS = 1ull << 24;   // GEMM multiply-adds
constexpr int64_t LARGE_K = 1 << 14;          // GEMM contraction length (split-K on its own)

// log2 of the lanes per output of a grouped GEMM: about 16 k values per lane.
uint32_t lanes_for(int64_t k) {
  uint32_t lg = 0;
  while (lg < 6 && ((int64_t)16 << lg) < k) ++lg;
  return lg;
}

uint64_t blocks_for(uint64_t n) { return (n + 255) / 256; }

// Reference kinds while one plan is planned: base ARENA with off = buffer id
// (an offset into the plan's arena once placed), INPUT0 + i for input i,
// OUTPUT for the plan's output.
constexpr uint32_t ARENA = 0, INPUT0 = 1, OUTPUT = 0xffffffffu;

struct Single {  // an op launched on its own
  enum Kind { PERMUTE, COPY, GEMM, SUM, DIAG, TRACE, STRIP } kind = PERMUTE;
  GRef in{}, in2{}, out{};
  uint32_t plan = 0;   // GEMM, STRIP: index of the plan within its group
  bool fold = false;   // GEMM: fold the result max into the plan's strip slot
  bool final = false;  // STRIP: the plan's last step (always normalise)
  PermPlan perm;
  GemmArgs gemm{};
  // GEMM: B read through gather tables instead of a permute (offsets into the
  // plan's table vector, then byte offsets into the group's descriptor blob), or -1
  int64_t brow = -1, bcol = -1;
  // GEMM: B read through a factored column offset instead of a permute
  // (gemm.bcg groups) and int64 row offsets: offset into the plan's int64 table
  // vector, then a byte offset into the group's blob, or -1
  int64_t brow64 = -1;
  // GEMM: A read through gather tables by the split pass of a split-precision
  // product instead of a permute (the same offsets), or -1
  int64_t arow = -1, acol = -1;
  // GEMM: operand max slots (FP16 scales) of A / B in the plan's table of result
  // maxima, or -1 (measured by a pass); GEMM without strip, STRIP: the slot that
  // receives this result's max, or -1
  int amax = -1, bmax = -1, max_out = -1;
  // GEMM: pending-exponent slots of A / B (GemmArgs::apend / bpend) in the same
  // table, or -1; STRIP: the slot receiving this result's exponent when its
  // rescale is left to its reader, or -1 (the result is rescaled in place)
  int apend = -1, bpend = -1, pend_out = -1;
  size_t A = 0, d = 0, Bm = 0, C = 0;  // SUM: [A, d, C]; DIAG/TRACE: [A, d, Bm, d, C]; COPY: A bytes
};

struct Round {
  std::vector<GPermDesc> perms;
  std::vector<GAxis> axes;  // axis0 of perms indexes this vector
  std::vector<GGemmDesc> gemms;
  std::vector<GStripDesc> strips;
  std::vector<Single> pre, mid;  // pre: sums, traces and large permutes; mid: large GEMMs
  std::vector<Single> big_strips;  // strips of large results (kind STRIP, `in` the result, A its numel)
};

struct Buf {
  size_t bytes;
  int first, last;
};

struct PlanProg {
  std::vector<Round> rounds;
  size_t arena = 0;  // bytes
  std::vector<int32_t> tables;  // gather tables of the B operands of large GEMMs
  std::vector<int64_t> tables64;  // row offsets of factored B operands
  int n_max = 0;                // result-max slots (large results read by later large GEMMs)
};

// Output-order (dim, input stride) axes of permuting `shape` by `perm`, with
// unit axes dropped and input-adjacent runs merged.
void merged_axes(const std::vector<size_t>& shape, const std::vector<size_t>& perm, std::vector<GAxis>& out) {
  std::vector<int64_t> st(shape.size(), 1);
  for (size_t a = shape.size(); a-- > 1;) st[a - 1] = st[a] * (int64_t)shape[a];
  const size_t first = out.size();
  for (size_t q = 0; q < perm.size(); ++q) {
    size_t a = perm[q];
    if (shape[a] == 1) continue;
    if (out.size() > first && out.back().in_stride == st[a] * (int64_t)shape[a]) {
      out.back().dim *= (uint32_t)shape[a];
      out.back().in_stride = st[a];
    } else {
      out.push_back(GAxis{st[a], (uint32_t)shape[a], 0});
    }
  }
}

// B's free legs (output axes n_con.. of perm) as at most 6 groups of axes that are
// adjacent in memory, innermost first: GemmArgs::bcg / bcd / bcs. false when the
// free legs are not in memory order or form more groups.
bool factor_columns(const std::vector<size_t>& shape, const std::vector<size_t>& perm, size_t n_con, GemmArgs& f) {
  std::vector<int64_t> st(shape.size(), 1);
  for (size_t a = shape.size(); a-- > 1;) st[a - 1] = st[a] * (int64_t)shape[a];
  f.bcg = 0;
  for (size_t d = perm.size(); d-- > n_con;) {
    const size_t a = perm[d];
    if (d + 1 < perm.size() && perm[d + 1] <= a) return false;  // not in memory order
    if (f.bcg && perm[d + 1] == a + 1) {  // adjacent in memory: extends the current group
      f.bcd[f.bcg - 1] *= (int64_t)shape[a];
    } else {
      if (f.bcg == 6) return false;
      f.bcd[f.bcg] = (int64_t)shape[a], f.bcs[f.bcg] = st[a];
      ++f.bcg;
    }
  }
  for (int q = 0; q < f.bcg; ++q) {
    if (f.bcd[q] > (int64_t)UINT32_MAX) return false;
    fast_div_u32((uint32_t)f.bcd[q], f.bcm[q], f.bcl[q]);
  }
  return f.bcg > 0;
}

// Plans one plan's ops into `n_rounds` rounds: 0 preps, 1 + s step s, the last finalize.
PlanProg build_plan(const Plan& P, int dt, bool strip, int prec, size_t n_rounds) {
  const int es = (int)dtype_size(dt);
  const Compiled& C = P.cp;
  PlanProg prog;
  prog.rounds.resize(n_rounds);
  std::vector<Buf> bufs;
  auto newbuf = [&](size_t bytes, int t) {
    bufs.push_back(Buf{std::max<size_t>(bytes, 4), t, t});
    return GRef{ARENA, bufs.size() - 1};
  };
  auto use = [&](const GRef& r, int t) {
    if (r.base == ARENA) bufs[r.off].last = std::max(bufs[r.off].last, t);
  };
  auto sum_axis = [&](Round& R, int t, GRef cur, std::vector<size_t>& shape, size_t ax) {
    Single sg;
    sg.kind = Single::SUM;
    sg.A = 1, sg.d = shape[ax], sg.C = 1;
    for (size_t q = 0; q < ax; ++q) sg.A *= shape[q];
    for (size_t q = ax + 1; q < shape.size(); ++q) sg.C *= shape[q];
    GRef dst = newbuf(sg.A * sg.C * es, t);
    use(cur, t);
    sg.in = cur, sg.out = dst;
    R.pre.push_back(sg);
    shape.erase(shape.begin() + ax);
    return dst;
  };
  // permute `cur` (of `shape`) by `perm` into a new buffer, or into *dst_given
  auto permute = [&](Round& R, int t, GRef cur, const std::vector<size_t>& shape, const std::vector<size_t>& perm,
                     const GRef* dst_given) -> GRef {
    PermPlan pp = plan_permute(shape, perm, 1, es);
    if (pp.kind == 0 && !dst_given) return cur;
    const uint64_t numel = prod(shape);
    GRef dst = dst_given ? *dst_given : newbuf(numel * es, t);
    use(cur, t);
    use(dst, t);
    if (numel >= LARGE_NUMEL) {
      Single sg;
      sg.in = cur, sg.out = dst;
      if (pp.kind == 0) sg.kind = Single::COPY, sg.A = numel * es;
      else sg.kind = Single::PERMUTE, sg.perm = pp;
      R.pre.push_back(sg);
    } else {
      GPermDesc g{};
      g.in = cur, g.out = dst, g.numel = numel;
      g.axis0 = (uint32_t)R.axes.size();
      merged_axes(shape, perm, R.axes);
      g.rank = (uint32_t)(R.axes.size() - g.axis0);
      R.perms.push_back(g);
    }
    return dst;
  };

  std::vector<GRef> slot(C.n_inputs + C.steps.size());
  // result-max slot of each plan slot produced by a large GEMM (stripped runs:
  // the power-of-two rescale writes the max after it), or -1
  std::vector<int> res_max(slot.size(), -1);
  // The step reading each plan slot, and whether a step's product runs on the
  // split-precision kernels (large, launched on its own): a large stripped result
  // read by such a product is not rescaled in place; its exponent goes to a slot
  // (res_pend) that the product divides out with its FP16 operand scales, exactly
  // the values the rescale would have given (every op between, sums and permutes,
  // is linear).
  std::vector<int> reader(slot.size(), -1);
  for (size_t si = 0; si < C.steps.size(); ++si) reader[C.steps[si].li] = reader[C.steps[si].ri] = (int)si;
  auto split_large = [&](const PairStep& st) {
    const uint64_t cn = (uint64_t)st.bt * st.m * st.n;
    const bool lg = cn >= LARGE_NUMEL || cn * st.k >= LARGE_MACS || (int64_t)st.k >= LARGE_K;
    return lg && split_products_apply(dt, prec, (int64_t)st.m, (int64_t)st.k, (int64_t)st.n);
  };
  std::vector<int> res_pend(slot.size(), -1);
  std::vector<std::vector<size_t>> shapes(slot.size());
  for (size_t i = 0; i < C.n_inputs; ++i) {  // round 0: traces / diagonals (src/compiled.rs:401-404)
    GRef cur{INPUT0 + (uint32_t)i, 0};
    auto shape = C.preps[i].input_shape;
    for (auto& op : C.preps[i].traces) {
      size_t ai = std::get<0>(op), aj = std::get<1>(op);
      bool diag = std::get<2>(op);
      Single sg;
      sg.kind = diag ? Single::DIAG : Single::TRACE;
      sg.A = 1, sg.d = shape[ai], sg.Bm = 1, sg.C = 1;
      for (size_t q = 0; q < ai; ++q) sg.A *= shape[q];
      for (size_t q = ai + 1; q < aj; ++q) sg.Bm *= shape[q];
      for (size_t q = aj + 1; q < shape.size(); ++q) sg.C *= shape[q];
      GRef dst = newbuf((diag ? sg.A * sg.d * sg.Bm * sg.C : sg.A * sg.Bm * sg.C) * es, 0);
      use(cur, 0);
      sg.in = cur, sg.out = dst;
      prog.rounds[0].pre.push_back(sg);
      shape.erase(shape.begin() + aj);
      if (!diag) shape.erase(shape.begin() + ai);
      cur = dst;
    }
    slot[i] = cur;
    shapes[i] = shape;
  }
  for (size_t si = 0; si < C.steps.size(); ++si) {  // src/compiled.rs:406-436
    const PairStep& st = C.steps[si];
    const int t = (int)si + 1;
    Round& R = prog.rounds[t];
    auto operand = [&](size_t sl, const std::vector<size_t>& sums, const std::vector<size_t>& perm) {
      GRef cur = slot[sl];
      auto shape = shapes[sl];
      for (size_t ax : sums) cur = sum_axis(R, t, cur, shape, ax);
      return permute(R, t, cur, shape, perm, nullptr);
    };
    const uint64_t cnum = (uint64_t)st.bt * st.m * st.n;
    const bool large = cnum >= LARGE_NUMEL || cnum * st.k >= LARGE_MACS || (int64_t)st.k >= LARGE_K;
    // A large GEMM reads a B that needs only a permutation (no batch or summed
    // legs) through gather tables, with its contracted legs taken in B's memory
    // order (A's contracted legs follow; the result layout does not change)
    std::vector<size_t> a_perm = st.a_perm, b_perm = st.b_perm;
    size_t n_con = 0;  // with no batch legs: the leading entries of b_perm, the trailing ones of a_perm
    for (size_t d = 0; d < b_perm.size(); ++d) {
      const Leg l = st.right_legs[b_perm[d]];
      if (std::count(st.left_legs.begin(), st.left_legs.end(), l) &&
          !std::count(st.result_legs.begin(), st.result_legs.end(), l))
        n_con = d + 1;
    }
    std::vector<int32_t> rows, cols;
    int32_t col_run = 1, row_run = 1;
    GemmArgs fac{};  // factored columns of B (bcg > 0), and their int64 row offsets
    std::vector<int64_t> fac_rows;
    bool gather_b = large && st.bt == 1 && st.sum_b.empty() && plan_permute(shapes[st.ri], b_perm, 1, es).kind != 0;
    if (gather_b) {
      if (n_con > 1) {
        std::vector<size_t> bc(b_perm.begin(), b_perm.begin() + (long)n_con);
        std::sort(bc.begin(), bc.end());
        std::vector<Leg> la = st.left_legs;  // A's legs once its summed legs are gone
        for (size_t ax : st.sum_a) la.erase(la.begin() + (long)ax);
        const size_t nfa = a_perm.size() - n_con;
        for (size_t q = 0; q < n_con; ++q) {
          a_perm[nfa + q] = (size_t)(std::find(la.begin(), la.end(), st.right_legs[bc[q]]) - la.begin());
          b_perm[q] = bc[q];
        }
      }
      // B's columns as a mixed-radix function of the column index (its free legs
      // kept in memory order; rows in int64; no column table, any operand size)
      // where the streaming kernel runs the product with many columns, and where the
      // int32 gather tables (at most 2^26 entries) would not do and the heuristic
      // runs the product on the FP32 MFMA or the streaming kernel (float32 /
      // complex64; the split and 3M forms read tables or row-major B); else B is
      // permuted.
      const bool can_stream = (dt == F32 || dt == C64) && st.m <= (dtype_complex(dt) ? 32u : 64u) && st.n >= 64 &&
                              st.n <= UINT32_MAX;
      const bool table_ok = st.k + st.n <= (1u << 26);
      bool factored = can_stream && st.n > (1u << 20) &&
                      gemm_auto_streams(dt, prec, (int64_t)st.m, (int64_t)st.k, (int64_t)st.n) &&
                      factor_columns(shapes[st.ri], b_perm, n_con, fac);
      gather_b = !factored && table_ok && gather_tables(shapes[st.ri], b_perm, n_con, rows, cols, col_run, row_run);
      if (!factored && !gather_b && gemm_reads_factored(dt, prec, (int64_t)st.m, (int64_t)st.k, (int64_t)st.n))
        factored = factor_columns(shapes[st.ri], b_perm, n_con, fac);
      if (factored) {
        fac_rows.clear();
        const auto& sh = shapes[st.ri];
        std::vector<int64_t> str(sh.size(), 1);
        for (size_t q = sh.size(); q-- > 1;) str[q - 1] = str[q] * (int64_t)sh[q];
        size_t nrow = 1;
        for (size_t d = 0; d < n_con; ++d) nrow *= sh[b_perm[d]];
        for (size_t e = 0; e < nrow; ++e) {
          size_t rem = e;
          int64_t off = 0;
          for (size_t d = n_con; d-- > 0;) {
            off += (int64_t)(rem % sh[b_perm[d]]) * str[b_perm[d]];
            rem /= sh[b_perm[d]];
          }
          fac_rows.push_back(off);
        }
      } else {
        fac.bcg = 0;
      }
      if (!gather_b && !fac.bcg) a_perm = st.a_perm, b_perm = st.b_perm;
    }
    // A of a large split-precision product: read through gather tables by the
    // split pass, which then does the permute's work
    std::vector<int32_t> arows, acols;
    int32_t arun = 1, crun = 1;
    const bool gather_a = large && st.bt == 1 && st.sum_a.empty() && !fac.bcg &&
                          split_products_apply(dt, prec, (int64_t)st.m, (int64_t)st.k, (int64_t)st.n) &&
                          plan_permute(shapes[st.li], a_perm, 1, es).kind != 0 && st.m + st.k <= (1u << 24) &&
                          gather_tables(shapes[st.li], a_perm, a_perm.size() - n_con, arows, acols, crun, arun);
    GRef a = gather_a ? slot[st.li] : operand(st.li, st.sum_a, a_perm);
    GRef b = gather_b || fac.bcg ? slot[st.ri] : operand(st.ri, st.sum_b, b_perm);
    GRef c = newbuf(cnum * es, t);
    use(a, t), use(b, t);
    if (large) {
      Single sg;
      sg.kind = Single::GEMM;
      sg.in = a, sg.in2 = b, sg.out = c;
      GemmArgs& g = sg.gemm;
      g.m = (int64_t)st.m, g.k = (int64_t)st.k, g.n = (int64_t)st.n;
      g.lda = g.k, g.ldb = g.n, g.ldc = g.n;
      g.nbo = 1, g.nbi = (int64_t)st.bt;
      g.sai = g.m * g.k, g.sbi = g.k * g.n, g.sci = g.m * g.n;
      if (gather_a) {
        sg.arow = (int64_t)prog.tables.size();
        prog.tables.insert(prog.tables.end(), arows.begin(), arows.end());
        sg.acol = (int64_t)prog.tables.size();
        prog.tables.insert(prog.tables.end(), acols.begin(), acols.end());
      }
      // operand maxima of earlier large results (a permute or a gather keeps the max)
      if (st.sum_a.empty()) sg.amax = res_max[st.li];
      if (st.sum_b.empty()) sg.bmax = res_max[st.ri];
      sg.apend = res_pend[st.li], sg.bpend = res_pend[st.ri];
      if (fac.bcg) {
        g.bcg = fac.bcg;
        std::copy(fac.bcd, fac.bcd + 6, g.bcd);
        std::copy(fac.bcs, fac.bcs + 6, g.bcs);
        std::copy(fac.bcm, fac.bcm + 6, g.bcm);
        std::copy(fac.bcl, fac.bcl + 6, g.bcl);
        sg.brow64 = (int64_t)prog.tables64.size();
        prog.tables64.insert(prog.tables64.end(), fac_rows.begin(), fac_rows.end());
      }
      if (gather_b) {
        sg.brow = (int64_t)prog.tables.size();
        prog.tables.insert(prog.tables.end(), rows.begin(), rows.end());
        sg.bcol = (int64_t)prog.tables.size();
        prog.tables.insert(prog.tables.end(), cols.begin(), cols.end());
        g.brow_run = row_run, g.bcol_run = col_run;
      }
      R.mid.push_back(sg);
    } else {
      R.gemms.push_back(
          GGemmDesc{a, b, c, (int64_t)st.m, (int64_t)st.k, (int64_t)st.n, (int64_t)st.bt, 0, lanes_for((int64_t)st.k)});
    }
    const bool last = si + 1 == C.steps.size();
    if (strip && cnum >= LARGE_NUMEL) {  // the GEMM (launched on its own) folds the max
      R.mid.back().fold = true;
      Single sg;
      sg.kind = Single::STRIP;
      sg.in = c, sg.A = cnum, sg.final = last;
      if (!last) sg.max_out = res_max[st.out] = prog.n_max++;
      if (!last && reader[st.out] >= 0 && split_large(C.steps[(size_t)reader[st.out]]))
        sg.pend_out = res_pend[st.out] = prog.n_max++;
      R.big_strips.push_back(sg);
    } else if (strip) {
      R.strips.push_back(GStripDesc{c, cnum, 0, 0, last ? 1u : 0u});
    }
    slot[st.out] = c;
    shapes[st.out] = st.result_shape;
  }
  {  // last round: finalize (src/compiled.rs:438-440)
    const Finalize& f = C.fin;
    const int t = (int)n_rounds - 1;
    Round& R = prog.rounds[t];
    GRef cur = slot[f.slot];
    auto shape = f.input_shape;
    for (size_t ax : f.sum) cur = sum_axis(R, t, cur, shape, ax);
    const GRef out{OUTPUT, 0};
    permute(R, t, cur, shape, f.perm, &out);
  }
  // First-fit placement over the plan's round timeline, 256-byte aligned.
  std::vector<size_t> order(bufs.size()), off(bufs.size());
  for (size_t i = 0; i < order.size(); ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) { return bufs[x].first < bufs[y].first; });
  struct Live {
    size_t off, end;
    int last;
  };
  std::vector<Live> live;
  for (size_t id : order) {
    live.erase(std::remove_if(live.begin(), live.end(), [&](const Live& l) { return l.last < bufs[id].first; }), live.end());
    std::sort(live.begin(), live.end(), [](const Live& x, const Live& y) { return x.off < y.off; });
    size_t need = (bufs[id].bytes + 255) / 256 * 256, pos = 0;
    for (auto& l : live) {
      if (pos + need <= l.off) break;
      pos = std::max(pos, l.end);
    }
    off[id] = pos;
    live.push_back(Live{pos, pos + need, bufs[id].last});
    prog.arena = std::max(prog.arena, pos + need);
  }
  auto place = [&](GRef& r) {
    if (r.base == ARENA) r.off = off[r.off];
  };
  for (auto& R : prog.rounds) {
    for (auto& g : R.perms) place(g.in), place(g.out);
    for (auto& g : R.gemms) place(g.a), place(g.b), place(g.c);
    for (auto& g : R.strips) place(g.c);
    for (auto* v : {&R.pre, &R.mid, &R.big_strips})
      for (auto& sg : *v) place(sg.in), place(sg.in2), place(sg.out);
  }
  return prog;
}

struct Launch {  // one round of a group
  size_t perm_off = 0, axes_off = 0, gemm_off = 0, strip_off = 0;  // byte offsets into the descriptor blob
  int n_perm = 0, n_gemm = 0, n_strip = 0;
  uint64_t perm_blocks = 0, gemm_blocks = 0, strip_blocks = 0;
  std::vector<Single> pre, mid, big_strips;
};

}  // namespace

struct GroupedExecutor::Group {
  size_t g0 = 0, g1 = 0;  // plans [g0, g1)
  std::vector<Launch> launches;
  std::vector<char> blob;  // descriptors, strip slots and exponents, GEMM scratch, base table
  size_t slots_off = 0, expo_off = 0, bases_off = 0, gemm_scratch_off = 0, gemm_scratch = 256;
  size_t maxtab_off = 0;  // result maxima (double bits) read as FP16 operand scales
  std::vector<size_t> in_base;
  size_t out_base = 0, nb = 0, arena = 0;
  char* dev = nullptr;
  std::vector<char*> bases;  // last uploaded base table
};

GroupedExecutor::GroupedExecutor(std::vector<std::shared_ptr<const Plan>> plans, int dt, bool strip, int prec)
    : plans_(std::move(plans)), dt_(dt), strip_(strip), prec_(prec) {
  if (prec == C3M && dt != C64) throw Error("precision c3m applies to complex64 only");
  if (prec != NATIVE && !(dt == F32 || dt == C64)) throw Error("split precision applies to float32 / complex64 only");
  const size_t np = plans_.size();
  size_t max_steps = 0;
  for (auto& p : plans_) {
    if (p->n_slices != 1) throw Error("grouped execution takes unsliced plans");
    max_steps = std::max(max_steps, p->cp.steps.size());
  }
  const size_t n_rounds = max_steps + 2;
  std::vector<PlanProg> progs(np);
  for (size_t p = 0; p < np; ++p) progs[p] = build_plan(*plans_[p], dt, strip, prec, n_rounds);

  size_t free_b = 0, total_b = 0;
  ATN_HIP(hipMemGetInfo(&free_b, &total_b));
  const size_t budget = free_b / 10 * 6;
  // Consecutive plans packed into groups whose arenas fit the budget; the
  // groups run one after another in one arena of the largest group's size.
  for (size_t p = 0; p < np;) {
    size_t q = p, sum = 0;
    while (q < np && (q == p || sum + progs[q].arena <= budget)) sum += progs[q++].arena;
    if (sum > budget)
      throw Error(cat("grouped execution: plan ", p, " needs ", sum, " bytes of workspace, over the budget of ", budget));
    auto G = std::make_unique<Group>();
    G->g0 = p, G->g1 = q;
    groups_.push_back(std::move(G));
    p = q;
  }
  try {
    for (auto& Gp : groups_) {
      Group& G = *Gp;
      const size_t ng = G.g1 - G.g0;
      std::vector<size_t> arena_at(ng);
      G.in_base.resize(ng);
      G.nb = 1;
      for (size_t i = 0; i < ng; ++i) {
        arena_at[i] = G.arena;
        G.arena += progs[G.g0 + i].arena;
        G.in_base[i] = G.nb;
        G.nb += plans_[G.g0 + i]->cp.n_inputs;
      }
      G.out_base = G.nb;
      G.nb += ng;
      arena_bytes_ = std::max(arena_bytes_, G.arena);
      auto append = [&](const void* src, size_t bytes) {
        size_t o = (G.blob.size() + 255) / 256 * 256;
        G.blob.resize(o + bytes);
        if (src && bytes) std::memcpy(G.blob.data() + o, src, bytes);
        return o;
      };
      std::vector<int> max_at(ng);  // first result-max slot of each plan in the group's table
      int n_max = 0;
      for (size_t i = 0; i < ng; ++i) max_at[i] = n_max, n_max += progs[G.g0 + i].n_max;
      G.maxtab_off = append(nullptr, (size_t)std::max(n_max, 1) * 8);
      std::vector<size_t> tab_at(ng), tab64_at(ng);  // byte offsets of each plan's tables in the blob
      for (size_t i = 0; i < ng; ++i) {
        const auto& tb = progs[G.g0 + i].tables;
        tab_at[i] = tb.empty() ? 0 : append(tb.data(), tb.size() * sizeof(int32_t));
        const auto& t64 = progs[G.g0 + i].tables64;
        tab64_at[i] = t64.empty() ? 0 : append(t64.data(), t64.size() * sizeof(int64_t));
      }
      G.launches.resize(n_rounds);
      for (size_t r = 0; r < n_rounds; ++r) {
        std::vector<GPermDesc> perms;
        std::vector<GAxis> axes;
        std::vector<GGemmDesc> gemms;
        std::vector<GStripDesc> strips;
        Launch& L = G.launches[r];
        for (size_t i = 0; i < ng; ++i) {
          const Round& R = progs[G.g0 + i].rounds[r];
          auto fix = [&](GRef x) {
            if (x.base == ARENA) x.off += arena_at[i];
            else if (x.base == OUTPUT) x.base = (uint32_t)(G.out_base + i), x.off = 0;
            else x.base = (uint32_t)(G.in_base[i] + (x.base - INPUT0));
            return x;
          };
          for (GPermDesc g : R.perms) {
            g.in = fix(g.in), g.out = fix(g.out), g.axis0 += (uint32_t)axes.size();
            perms.push_back(g);
          }
          axes.insert(axes.end(), R.axes.begin(), R.axes.end());
          for (GGemmDesc g : R.gemms) g.a = fix(g.a), g.b = fix(g.b), g.c = fix(g.c), gemms.push_back(g);
          for (GStripDesc g : R.strips) g.c = fix(g.c), g.plan = (uint32_t)i, strips.push_back(g);
          for (Single sg : R.pre) sg.in = fix(sg.in), sg.out = fix(sg.out), L.pre.push_back(sg);
          for (Single sg : R.mid) {
            sg.in = fix(sg.in), sg.in2 = fix(sg.in2), sg.out = fix(sg.out), sg.plan = (uint32_t)i;
            if (sg.brow >= 0) sg.brow = (int64_t)(tab_at[i] + sg.brow * 4), sg.bcol = (int64_t)(tab_at[i] + sg.bcol * 4);
            if (sg.arow >= 0) sg.arow = (int64_t)(tab_at[i] + sg.arow * 4), sg.acol = (int64_t)(tab_at[i] + sg.acol * 4);
            if (sg.brow64 >= 0) sg.brow64 = (int64_t)(tab64_at[i] + sg.brow64 * 8);
            if (sg.amax >= 0) sg.amax += max_at[i];
            if (sg.bmax >= 0) sg.bmax += max_at[i];
            if (sg.apend >= 0) sg.apend += max_at[i];
            if (sg.bpend >= 0) sg.bpend += max_at[i];
            L.mid.push_back(sg);
          }
          for (Single sg : R.big_strips) {
            sg.in = fix(sg.in), sg.plan = (uint32_t)i;
            if (sg.max_out >= 0) sg.max_out += max_at[i];
            if (sg.pend_out >= 0) sg.pend_out += max_at[i];
            L.big_strips.push_back(sg);
          }
        }
        uint64_t b = 0;
        for (auto& g : perms) g.block0 = b, b += blocks_for(g.numel);
        L.perm_blocks = b, b = 0;
        for (auto& g : gemms) g.block0 = b, b += blocks_for(((uint64_t)g.m * g.n * g.bt) << g.lg_lanes);
        L.gemm_blocks = b, b = 0;
        for (auto& g : strips) g.block0 = b, b += blocks_for(g.numel);
        L.strip_blocks = b;
        L.n_perm = (int)perms.size(), L.n_gemm = (int)gemms.size(), L.n_strip = (int)strips.size();
        if (L.n_perm) {
          L.perm_off = append(perms.data(), perms.size() * sizeof(GPermDesc));
          L.axes_off = append(axes.data(), axes.size() * sizeof(GAxis));
        }
        if (L.n_gemm) L.gemm_off = append(gemms.data(), gemms.size() * sizeof(GGemmDesc));
        if (L.n_strip) L.strip_off = append(strips.data(), strips.size() * sizeof(GStripDesc));
        for (auto& sg : L.mid) G.gemm_scratch = std::max(G.gemm_scratch, gemm_scratch_bytes(dt, prec, GEMM_AUTO, sg.gemm));
      }
      // per plan: max slot, power-of-two exponent, log10 of the last step's exact max
      G.slots_off = append(nullptr, 3 * ng * 8);
      G.expo_off = G.slots_off + ng * 8;
      std::memset(G.blob.data() + G.slots_off, 0, 3 * ng * 8);
      G.gemm_scratch_off = append(nullptr, G.gemm_scratch);
      G.bases_off = append(nullptr, G.nb * sizeof(char*));
      ATN_HIP(hipMalloc((void**)&G.dev, G.blob.size()));
      ATN_HIP(hipMemcpy(G.dev, G.blob.data(), G.blob.size(), hipMemcpyHostToDevice));
      G.bases.assign(G.nb, nullptr);
      G.blob.resize(G.bases_off + G.nb * sizeof(char*));  // keep the base table staging area only
      G.blob.shrink_to_fit();
    }
    ATN_HIP(hipMalloc((void**)&arena_, std::max<size_t>(arena_bytes_, 256)));
  } catch (...) {
    release();
    throw;
  }
}

void GroupedExecutor::release() {
  for (auto& G : groups_)
    if (G->dev) (void)hipFree(G->dev), G->dev = nullptr;
  if (arena_) (void)hipFree(arena_), arena_ = nullptr;
}

GroupedExecutor::~GroupedExecutor() { release(); }

void GroupedExecutor::run(const std::vector<const void* const*>& inputs, const std::vector<void*>& outputs, double* log10_exp,
                          hipStream_t s, Report* rep) {
  const size_t np = plans_.size();
  if (inputs.size() != np || outputs.size() != np) throw Error("grouped execution: one input array and one output per plan");
  const int dt = dt_;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  std::vector<long long> expo_host(np, 0);
  std::vector<double> last_host(np, 0.0);
  const size_t l0 = g_launch_count;
  try {
    if (rep) {
      ATN_HIP(hipEventCreate(&e0));
      ATN_HIP(hipEventCreate(&e1));
      ATN_HIP(hipEventRecord(e0, s));
    }
    for (auto& Gp : groups_) {
      Group& G = *Gp;
      const size_t ng = G.g1 - G.g0;
      std::vector<char*> bases(G.nb, nullptr);
      bases[0] = arena_;
      for (size_t i = 0; i < ng; ++i) {
        for (size_t q = 0; q < plans_[G.g0 + i]->cp.n_inputs; ++q)
          bases[G.in_base[i] + q] = const_cast<char*>(static_cast<const char*>(inputs[G.g0 + i][q]));
        bases[G.out_base + i] = static_cast<char*>(outputs[G.g0 + i]);
      }
      if (bases != G.bases) {
        ATN_HIP(hipStreamSynchronize(s));  // an earlier copy from the staging area has completed
        std::memcpy(G.blob.data() + G.bases_off, bases.data(), G.nb * sizeof(char*));
        ATN_HIP(hipMemcpyAsync(G.dev + G.bases_off, G.blob.data() + G.bases_off, G.nb * sizeof(char*),
                               hipMemcpyHostToDevice, s));
        G.bases = bases;
      }
      char* dev = G.dev;
      char* const* dbases = reinterpret_cast<char* const*>(dev + G.bases_off);
      auto* slots = reinterpret_cast<unsigned long long*>(dev + G.slots_off);
      auto* maxtab = reinterpret_cast<unsigned long long*>(dev + G.maxtab_off);
      auto* expo = reinterpret_cast<long long*>(dev + G.expo_off);
      auto* log10_last = reinterpret_cast<double*>(dev + G.expo_off + ng * 8);
      if (strip_) {  // exponents and log10 parts
        hipLaunchKernelGGL(grouped_zero_kernel, dim3((unsigned)((2 * ng + 255) / 256)), dim3(256), 0, s,
                           reinterpret_cast<unsigned long long*>(expo), (uint64_t)(2 * ng));
        ATN_HIP(hipGetLastError());
        ++g_launch_count;
      }
      auto at = [&](const GRef& x) { return bases[x.base] + x.off; };
      auto run_single = [&](const Single& sg) {
        switch (sg.kind) {
          case Single::PERMUTE: launch_permute(sg.perm, at(sg.in), at(sg.out), s); break;
          case Single::COPY: launch_copy(sg.A, at(sg.in), at(sg.out), s); break;
          case Single::GEMM: {
            GemmArgs g = sg.gemm;
            g.a = at(sg.in), g.b = at(sg.in2), g.c = at(sg.out);
            g.scratch = dev + G.gemm_scratch_off;
            g.scratch_bytes = G.gemm_scratch;
            if (sg.fold) g.cmax = slots + sg.plan;
            if (sg.brow >= 0)
              g.brow = reinterpret_cast<const int32_t*>(dev + sg.brow), g.bcol = reinterpret_cast<const int32_t*>(dev + sg.bcol);
            if (sg.arow >= 0)
              g.arow = reinterpret_cast<const int32_t*>(dev + sg.arow), g.acol = reinterpret_cast<const int32_t*>(dev + sg.acol);
            if (sg.brow64 >= 0) g.brow64 = reinterpret_cast<const int64_t*>(dev + sg.brow64);
            if (sg.amax >= 0) g.amax = maxtab + sg.amax;
            if (sg.bmax >= 0) g.bmax = maxtab + sg.bmax;
            if (sg.apend >= 0) g.apend = reinterpret_cast<const int64_t*>(maxtab + sg.apend);
            if (sg.bpend >= 0) g.bpend = reinterpret_cast<const int64_t*>(maxtab + sg.bpend);
            launch_gemm(dt, prec_, GEMM_AUTO, g, s);
            break;
          }
          case Single::SUM: launch_sum_axis(dt, sg.A, sg.d, sg.C, at(sg.in), at(sg.out), s); break;
          case Single::DIAG: launch_diag((int)dtype_size(dt), sg.A, sg.d, sg.Bm, sg.C, at(sg.in), at(sg.out), s); break;
          case Single::TRACE: launch_trace(dt, sg.A, sg.d, sg.Bm, sg.C, at(sg.in), at(sg.out), s); break;
          case Single::STRIP: break;  // run in the strip phase of its round
        }
      };
      auto grid = [](uint64_t blocks) { return dim3((unsigned)std::min<uint64_t>(blocks, GRID_CAP)); };
      gdispatch(dt, [&](auto tag) {
        using T = decltype(tag);
        for (auto& L : G.launches) {
          const bool strips = L.strip_blocks || !L.big_strips.empty();
          if (strips) {  // strip slots zeroed before the GEMMs that fold their max into them
            hipLaunchKernelGGL(grouped_zero_kernel, dim3((unsigned)((ng + 255) / 256)), dim3(256), 0, s, slots, (uint64_t)ng);
            ATN_HIP(hipGetLastError());
            ++g_launch_count;
          }
          for (auto& sg : L.pre) run_single(sg);
          if (L.perm_blocks) {
            g_last_kind = K_PERMUTE_GENERIC;
            hipLaunchKernelGGL(grouped_permute_kernel<T>, grid(L.perm_blocks), dim3(256), 0, s,
                               reinterpret_cast<const GPermDesc*>(dev + L.perm_off), L.n_perm, L.perm_blocks,
                               reinterpret_cast<const GAxis*>(dev + L.axes_off), dbases);
            ATN_HIP(hipGetLastError());
            ++g_launch_count;
          }
          for (auto& sg : L.mid) run_single(sg);
          if (L.gemm_blocks) {
            g_last_kind = K_GEMM_SIMPLE;
            hipLaunchKernelGGL(grouped_gemm_kernel<T>, grid(L.gemm_blocks), dim3(256), 0, s,
                               reinterpret_cast<const GGemmDesc*>(dev + L.gemm_off), L.n_gemm, L.gemm_blocks, dbases);
            ATN_HIP(hipGetLastError());
            ++g_launch_count;
          }
          if (strips) {
            g_last_kind = K_STRIP;
            if (L.strip_blocks) {
              hipLaunchKernelGGL(grouped_maxabs_kernel<T>, grid(L.strip_blocks), dim3(256), 0, s,
                                 reinterpret_cast<const GStripDesc*>(dev + L.strip_off), L.n_strip, L.strip_blocks, dbases,
                                 slots);
              ++g_launch_count;
            }
            // (the big results' maxima were folded in by their GEMMs)
            if (L.strip_blocks) {
              hipLaunchKernelGGL(grouped_scale_kernel<T>, grid(L.strip_blocks), dim3(256), 0, s,
                                 reinterpret_cast<const GStripDesc*>(dev + L.strip_off), L.n_strip, L.strip_blocks, dbases,
                                 (const unsigned long long*)slots, expo, log10_last);
              ++g_launch_count;
            }
            for (auto& sg : L.big_strips) {
              if (sg.final)
                launch_strip_exact(dt, at(sg.in), sg.A, (uint64_t*)(slots + sg.plan), log10_last + sg.plan, s);
              else
                launch_scale_pow2(dt, at(sg.in), sg.A, (const uint64_t*)(slots + sg.plan), (int64_t*)(expo + sg.plan), s,
                                  STRIP_KEEP, sg.max_out >= 0 ? (uint64_t*)(maxtab + sg.max_out) : nullptr,
                                  sg.pend_out >= 0 ? (int64_t*)(maxtab + sg.pend_out) : nullptr);
            }
            ATN_HIP(hipGetLastError());
          }
        }
      });
      if (strip_ && log10_exp) {
        ATN_HIP(hipMemcpyAsync(expo_host.data() + G.g0, expo, ng * 8, hipMemcpyDeviceToHost, s));
        ATN_HIP(hipMemcpyAsync(last_host.data() + G.g0, log10_last, ng * 8, hipMemcpyDeviceToHost, s));
      }
    }
    if (rep) {
      ATN_HIP(hipEventRecord(e1, s));
      ATN_HIP(hipEventSynchronize(e1));
      float ms = 0;
      ATN_HIP(hipEventElapsedTime(&ms, e0, e1));
      double flops = 0;
      for (auto& p : plans_) flops += (dtype_complex(dt) ? 8.0 : 2.0) * p->macs_per_slice;
      rep->elapsed_ms = ms;
      rep->flops = flops;
      rep->tflops = ms > 0 ? flops / (ms * 1e-3) / 1e12 : 0;
      rep->n_slices = np;
      rep->slices_per_s = ms > 0 ? (double)np / (ms * 1e-3) : 0;
      rep->n_launches = g_launch_count - l0;
      rep->workspace_bytes = arena_bytes_;
      rep->batch = groups_.size();
    }
    if (log10_exp) ATN_HIP(hipStreamSynchronize(s));
  } catch (...) {
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    throw;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (log10_exp)
    for (size_t p = 0; p < np; ++p) log10_exp[p] = strip_ ? (double)expo_host[p] * std::log10(2.0) + last_host[p] : 0.0;
}

void grouped_execute(const std::vector<std::shared_ptr<const Plan>>& plans, int dt, bool strip, int prec,
                     const std::vector<const void* const*>& inputs, const std::vector<void*>& outputs, double* log10_exp,
                     hipStream_t s, Report* rep) {
  GroupedExecutor ex(plans, dt, strip, prec);
  ex.run(inputs, outputs, log10_exp, s, rep);
  ATN_HIP(hipStreamSynchronize(s));  // the arena is freed when ex goes out of scope
}

}  // namespace atn
