// Contraction GEMMs: C[zo, zi] = A[zo, zi] . B[zo, zi], row-major, alpha = 1,
// beta = 0 (src/contract.rs:102-181). Two batch levels: the slice batch (outer,
// stride 0 = operand shared by every slice) and the plan's own batch legs.
//
// Complex on MFMA (4M): with A viewed as a real [m, 2k] matrix and C as a real
// [m, 2n] matrix (interleaved storage, no copies), C = A . Bx where Bx is the
// real [2k, 2n] expansion of B with 2x2 blocks [[br, bi], [-bi, br]]:
//   C[i, 2j]   = sum_p ar*br - ai*bi = Re(C_ij)
//   C[i, 2j+1] = sum_p ar*bi + ai*br = Im(C_ij)
// The expansion is built while staging B into LDS, so HBM traffic is that of
// the complex operands; the MFMA work is 8*m*k*n real FLOPs as for 4M.
#include <algorithm>
#include <type_traits>

#include "kernels.hpp"

namespace atn {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef double v4d __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef _Float16 v8h __attribute__((ext_vector_type(8)));

#define LAUNCH(kernel, grid, block, lds, stream, ...)                            \
  do {                                                                            \
    hipLaunchKernelGGL(kernel, grid, block, lds, stream, __VA_ARGS__);            \
    ATN_HIP(hipGetLastError());                                                   \
    ++g_launch_count;                                                             \
  } while (0)

template <class T> struct CM;
template <> struct CM<float> {
  static __device__ float zero() { return 0.f; }
  static __device__ float mac(float acc, float a, float b) { return acc + a * b; }
  static __device__ float conj(float a) { return a; }
};
template <> struct CM<double> {
  static __device__ double zero() { return 0.0; }
  static __device__ double mac(double acc, double a, double b) { return acc + a * b; }
  static __device__ double conj(double a) { return a; }
};
template <> struct CM<float2> {
  static __device__ float2 zero() { return make_float2(0.f, 0.f); }
  static __device__ float2 mac(float2 acc, float2 a, float2 b) {
    return make_float2(acc.x + (a.x * b.x - a.y * b.y), acc.y + (a.x * b.y + a.y * b.x));
  }
  static __device__ float2 conj(float2 a) { return make_float2(a.x, -a.y); }
};
template <> struct CM<double2> {
  static __device__ double2 zero() { return make_double2(0.0, 0.0); }
  static __device__ double2 mac(double2 acc, double2 a, double2 b) {
    return make_double2(acc.x + (a.x * b.x - a.y * b.y), acc.y + (a.x * b.y + a.y * b.x));
  }
  static __device__ double2 conj(double2 a) { return make_double2(a.x, -a.y); }
};

// Result max-magnitude tracking for the strip (GemmArgs::cmax): each thread
// folds max |re|, |im| of the values it stores, one atomicMax per wave at the
// end. Non-negative doubles order like their bit patterns; NaN never wins.
// Grid bound of kernels that track the result max (persistent loops over
// tiles). The split-precision GEMMs use it only where A is one row tile: with
// several row tiles their full grid holds all row tiles of a range of column
// tiles of one batch entry in flight, which share each B panel in L2, where a
// capped grid spreads them over many batch entries (complex64 m=1024 k=512
// n=8192 x256: 24.2 ms full, 26.9 capped; on one row tile the cap is the faster,
// 6.6 against 7.5 ms on m=16 k=16 n=262144 x256).
constexpr unsigned TRACK_WORKGROUPS = 8192;

// The running max is kept in the value's own precision (float compares for
// float results) and widened to double once, when published.
__device__ inline void fold_max(float& mx, float v) {
  const float a = fabsf(v);
  mx = a > mx ? a : mx;
}
__device__ inline void fold_max(double& mx, double v) {
  const double a = fabs(v);
  mx = a > mx ? a : mx;
}
__device__ inline void fold_max(float& mx, float2 v) { fold_max(mx, v.x), fold_max(mx, v.y); }
__device__ inline void fold_max(double& mx, double2 v) { fold_max(mx, v.x), fold_max(mx, v.y); }
template <class T> struct MaxOf { using R = float; };
template <> struct MaxOf<double> { using R = double; };
template <> struct MaxOf<double2> { using R = double; };
// every lane of the wave calls this
__device__ inline void publish_max(double mx, unsigned long long* cmax) {
  for (int o = 32; o > 0; o >>= 1) {
    const double t = __shfl_down(mx, o, 64);
    mx = t > mx ? t : mx;
  }
  // one slot for the whole grid: read it first and skip the atomic unless this
  // wave's max is larger (the slot only grows, so a stale read is harmless)
  if ((threadIdx.x & 63) == 0 && mx > 0.0) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(mx);
    if (b > __hip_atomic_load(cmax, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMax(cmax, b);
  }
}

// B element (p, j) of one batch entry: row-major with ldb, or through the gather
// tables of an unpermuted operand (GemmArgs::brow / bcol).
__device__ inline int64_t b_index(const GemmArgs& g, int64_t p, int64_t j) {
  return g.bcol ? (int64_t)g.brow[p] + (int64_t)g.bcol[j] : p * g.ldb + j;
}
// Column offset of a factored B (GemmArgs::bcg > 0); the column count is below
// 2^32, so the digits come from 32-bit divisions by multiply-high (divisions by
// a runtime divisor cost the streaming kernel about as much as its multiply-adds).
__device__ inline int64_t fac_col(const GemmArgs& g, int64_t j) {
  uint32_t r = (uint32_t)j;
  int64_t off = 0;
  for (int q = 0; q < g.bcg; ++q) {
    uint32_t t = 0;
    if (g.bcl[q]) {
      const uint32_t h = __umulhi(g.bcm[q], r);
      t = (h + ((r - h) >> 1)) >> (g.bcl[q] - 1);
    }
    off += (int64_t)(r - t * (uint32_t)g.bcd[q]) * g.bcs[q];
    r = t;
  }
  return off;
}

// Granlund-Montgomery round-up constants: l = ceil(log2 d), mul = floor(2^32 (2^l - d)
// / d) + 1 (below 2^32), exact for every 32-bit dividend; d = 1 gets l = 0.
void fast_div_u32(uint32_t d, uint32_t& mul, int32_t& shift) {
  if (d <= 1) {
    mul = 0, shift = 0;
    return;
  }
  int l = 0;
  while (((uint64_t)1 << l) < d) ++l;
  mul = (uint32_t)((((uint64_t)1 << 32) * (((uint64_t)1 << l) - d)) / d + 1);
  shift = l;
}
// The two halves of b_index, for loaders whose column stays fixed across k tiles.
__device__ inline int64_t b_row(const GemmArgs& g, int64_t p) { return g.bcol ? (int64_t)g.brow[p] : p * g.ldb; }
__device__ inline int64_t b_col(const GemmArgs& g, int64_t j) { return g.bcol ? (int64_t)g.bcol[j] : j; }

// ------------------------------------------------------------- simple GEMM --
// One thread per C element, k accumulated in order. Used for skinny and tiny
// steps, which are bandwidth- or launch-bound. OPA/OPB = 1 reads the operand as
// its conjugate transpose (adjoint steps): A is then stored [k, m], B [n, k].
template <class T, int OPA, int OPB>
__global__ __launch_bounds__(256) void gemm_simple_kernel(GemmArgs g) {
  const uint64_t mn = (uint64_t)g.m * g.n;
  const uint64_t nz = (uint64_t)g.nbo * g.nbi;
  for (uint64_t z = blockIdx.y; z < nz; z += gridDim.y) {
    uint64_t zo = z / g.nbi, zi = z - zo * g.nbi;
    const T* A = (const T*)g.a + zo * g.sao + zi * g.sai;
    const T* B = (const T*)g.b + zo * g.sbo + zi * g.sbi;
    T* C = (T*)g.c + zo * g.sco + zi * g.sci;
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < mn; e += (uint64_t)gridDim.x * blockDim.x) {
      uint64_t i = e / g.n, j = e - i * g.n;
      T acc = CM<T>::zero();
      for (int64_t p = 0; p < g.k; ++p) {
        T a = OPA ? CM<T>::conj(A[p * g.lda + i])
                  : A[g.acol ? (int64_t)g.arow[i] + (int64_t)g.acol[p] : (int64_t)i * g.lda + p];
        T b = OPB ? CM<T>::conj(B[j * g.ldb + p]) : B[b_index(g, p, j)];
        acc = CM<T>::mac(acc, a, b);
      }
      C[i * g.ldc + j] = acc;
    }
  }
}

// A group of small independent GEMMs in one launch (a level of a plan's tiny
// slice-invariant steps): each 256-thread block finds its GEMM by binary search
// over the blocks' starting indices, then computes one C element per thread as
// the simple kernel does.
template <class T>
__global__ __launch_bounds__(256) void gemm_simple_group_kernel(const SimpleGemmDesc* __restrict__ d, int n,
                                                                uint64_t blocks, char* ws,
                                                                const void* const* __restrict__ inputs,
                                                                const int32_t* __restrict__ tab) {
  for (uint64_t blk = blockIdx.x; blk < blocks; blk += gridDim.x) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) / 2;
      if (d[mid].block0 <= blk) lo = mid;
      else hi = mid - 1;
    }
    const SimpleGemmDesc& q = d[lo];
    const uint64_t e = (blk - q.block0) * 256 + threadIdx.x;
    if (e >= (uint64_t)(q.m * q.n)) continue;
    const T* A = q.a >= 0 ? (const T*)(ws + q.a) : (const T*)inputs[-q.a - 1];
    const T* B = q.b >= 0 ? (const T*)(ws + q.b) : (const T*)inputs[-q.b - 1];
    T* C = (T*)(ws + q.c);
    const int64_t i = (int64_t)(e / (uint64_t)q.n), j = (int64_t)(e - (uint64_t)i * q.n);
    const int64_t ar = q.acol >= 0 ? (int64_t)tab[q.arow + i] : i * q.k;
    const int64_t bc = q.bcol >= 0 ? (int64_t)tab[q.bcol + j] : j;
    T acc = CM<T>::zero();
    for (int64_t p = 0; p < q.k; ++p) {
      const T a = A[ar + (q.acol >= 0 ? (int64_t)tab[q.acol + p] : p)];
      const T b = B[(q.bcol >= 0 ? (int64_t)tab[q.brow + p] : p * q.n) + bc];
      acc = CM<T>::mac(acc, a, b);
    }
    C[e] = acc;
  }
}

// ----------------------------------------------------------- skinny GEMMs --
// Small m (m <= MM): one thread per output column j keeps all m accumulators,
// so B and C are streamed exactly once with coalesced accesses; A (m x k) is
// wave-uniform and served from the scalar cache. Loads of UN consecutive k
// values are issued before the multiply-adds.
template <class T, int MM>
__global__ __launch_bounds__(256) void gemm_smallm_kernel(GemmArgs g) {
  constexpr int UN = 4;
  const uint64_t nz = (uint64_t)g.nbo * g.nbi;
  for (uint64_t z = blockIdx.y; z < nz; z += gridDim.y) {
    uint64_t zo = z / g.nbi, zi = z - zo * g.nbi;
    const T* A = (const T*)g.a + zo * g.sao + zi * g.sai;
    const T* B = (const T*)g.b + zo * g.sbo + zi * g.sbi;
    T* C = (T*)g.c + zo * g.sco + zi * g.sci;
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < g.n; j += (int64_t)gridDim.x * blockDim.x) {
      T acc[MM];
#pragma unroll
      for (int i = 0; i < MM; ++i) acc[i] = CM<T>::zero();
      int64_t p = 0;
      for (; p + UN <= g.k; p += UN) {
        T b[UN];
#pragma unroll
        for (int u = 0; u < UN; ++u) b[u] = B[(p + u) * g.ldb + j];
#pragma unroll
        for (int u = 0; u < UN; ++u)
#pragma unroll
          for (int i = 0; i < MM; ++i)
            if (i < g.m) acc[i] = CM<T>::mac(acc[i], A[i * g.lda + p + u], b[u]);
      }
      for (; p < g.k; ++p) {
        T b = B[p * g.ldb + j];
#pragma unroll
        for (int i = 0; i < MM; ++i)
          if (i < g.m) acc[i] = CM<T>::mac(acc[i], A[i * g.lda + p], b);
      }
#pragma unroll
      for (int i = 0; i < MM; ++i)
        if (i < g.m) C[i * g.ldc + j] = acc[i];
    }
  }
}

// Streaming small-m GEMM: each thread owns V consecutive columns (16-byte
// loads and stores) and all m rows; KU rows of B are loaded before their
// multiply-adds, so every thread keeps KU 16-byte loads in flight. A (m x k) is
// wave-uniform and read through the scalar cache. B and C are streamed once.
template <class T, int V> struct Vec;
template <> struct Vec<float, 4> { using W = float4; };
template <> struct Vec<float, 2> { using W = float2; };
template <> struct Vec<float, 1> { using W = float; };
template <> struct Vec<float2, 2> { using W = float4; };
template <> struct Vec<float2, 1> { using W = float2; };
template <> struct Vec<double, 2> { using W = double2; };
template <> struct Vec<double, 1> { using W = double; };
template <> struct Vec<double2, 1> { using W = double2; };

// FAC: B through a factored column offset and int64 row offsets (GemmArgs::bcg).
template <class T, int MM, int V, bool FAC>
__global__ __launch_bounds__(256) void gemm_stream_kernel(GemmArgs g) {
  constexpr int KU = 8;
  using W = typename Vec<T, V>::W;
  static_assert(sizeof(W) == V * sizeof(T), "vector width");
  const uint64_t nz = (uint64_t)g.nbo * g.nbi;
  typename MaxOf<T>::R mx = 0;
  for (uint64_t z = blockIdx.y; z < nz; z += gridDim.y) {
    const uint64_t zo = z / g.nbi, zi = z - zo * g.nbi;
    const T* A = (const T*)g.a + zo * g.sao + zi * g.sai;
    const T* B = (const T*)g.b + zo * g.sbo + zi * g.sbi;
    T* C = (T*)g.c + zo * g.sco + zi * g.sci;
    for (int64_t j0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * V; j0 < g.n;
         j0 += (int64_t)gridDim.x * blockDim.x * V) {
      T acc[MM][V];
#pragma unroll
      for (int i = 0; i < MM; ++i)
#pragma unroll
        for (int v = 0; v < V; ++v) acc[i][v] = CM<T>::zero();
      const int64_t cj = FAC ? fac_col(g, j0) : 0;
      auto bat = [&](int64_t p) { return B + (FAC ? g.brow64[p] + cj : b_index(g, p, j0)); };
      int64_t p = 0;
      for (; p + KU <= g.k; p += KU) {
        W b[KU];
#pragma unroll
        for (int u = 0; u < KU; ++u) b[u] = *reinterpret_cast<const W*>(bat(p + u));
#pragma unroll
        for (int u = 0; u < KU; ++u) {
          const T* bu = reinterpret_cast<const T*>(&b[u]);
#pragma unroll
          for (int i = 0; i < MM; ++i)
            if (i < g.m) {
              const T a = A[i * g.lda + p + u];
#pragma unroll
              for (int v = 0; v < V; ++v) acc[i][v] = CM<T>::mac(acc[i][v], a, bu[v]);
            }
        }
      }
      for (; p < g.k; ++p) {
        W bw = *reinterpret_cast<const W*>(bat(p));
        const T* bu = reinterpret_cast<const T*>(&bw);
#pragma unroll
        for (int i = 0; i < MM; ++i)
          if (i < g.m) {
            const T a = A[i * g.lda + p];
#pragma unroll
            for (int v = 0; v < V; ++v) acc[i][v] = CM<T>::mac(acc[i][v], a, bu[v]);
          }
      }
#pragma unroll
      for (int i = 0; i < MM; ++i)
        if (i < g.m) {
          *reinterpret_cast<W*>(C + i * g.ldc + j0) = *reinterpret_cast<const W*>(acc[i]);
          if (g.cmax)
#pragma unroll
            for (int v = 0; v < V; ++v) fold_max(mx, acc[i][v]);
        }
    }
  }
  if (g.cmax) publish_max((double)mx, g.cmax);
}

// Small n (n <= NN): one thread per output row i keeps all n accumulators;
// B (k x n) is wave-uniform.
template <class T, int NN>
__global__ __launch_bounds__(256) void gemm_smalln_kernel(GemmArgs g) {
  constexpr int UN = 4;
  const uint64_t nz = (uint64_t)g.nbo * g.nbi;
  for (uint64_t z = blockIdx.y; z < nz; z += gridDim.y) {
    uint64_t zo = z / g.nbi, zi = z - zo * g.nbi;
    const T* A = (const T*)g.a + zo * g.sao + zi * g.sai;
    const T* B = (const T*)g.b + zo * g.sbo + zi * g.sbi;
    T* C = (T*)g.c + zo * g.sco + zi * g.sci;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < g.m; i += (int64_t)gridDim.x * blockDim.x) {
      T acc[NN];
#pragma unroll
      for (int j = 0; j < NN; ++j) acc[j] = CM<T>::zero();
      const T* Ar = A + i * g.lda;
      int64_t p = 0;
      for (; p + UN <= g.k; p += UN) {
        T a[UN];
#pragma unroll
        for (int u = 0; u < UN; ++u) a[u] = Ar[p + u];
#pragma unroll
        for (int u = 0; u < UN; ++u)
#pragma unroll
          for (int j = 0; j < NN; ++j)
            if (j < g.n) acc[j] = CM<T>::mac(acc[j], a[u], B[(p + u) * g.ldb + j]);
      }
      for (; p < g.k; ++p) {
        T a = Ar[p];
#pragma unroll
        for (int j = 0; j < NN; ++j)
          if (j < g.n) acc[j] = CM<T>::mac(acc[j], a, B[p * g.ldb + j]);
      }
#pragma unroll
      for (int j = 0; j < NN; ++j)
        if (j < g.n) C[i * g.ldc + j] = acc[j];
    }
  }
}

// ------------------------------------------------------------ split-K GEMM --
// Few outputs and a long k (dot-like steps, e.g. the last steps of a closed
// network): a workgroup owns a TM x TN output tile and one k range; its 256
// threads stride over the range, each keeping TM*TN accumulators, and the tile
// is reduced across the workgroup in a fixed order (wave shuffles, then the
// four waves in order). With several k ranges the partial tiles go to scratch
// and a second kernel adds them in range order, so the result does not depend
// on scheduling.
template <class T> struct Shfl;
template <> struct Shfl<float> {
  static __device__ float down(float v, int o) { return __shfl_down(v, o, 64); }
};
template <> struct Shfl<double> {
  static __device__ double down(double v, int o) { return __shfl_down(v, o, 64); }
};
template <> struct Shfl<float2> {
  static __device__ float2 down(float2 v, int o) { return make_float2(__shfl_down(v.x, o, 64), __shfl_down(v.y, o, 64)); }
};
template <> struct Shfl<double2> {
  static __device__ double2 down(double2 v, int o) { return make_double2(__shfl_down(v.x, o, 64), __shfl_down(v.y, o, 64)); }
};
template <class T> __device__ inline T cadd(T a, T b);
template <> __device__ inline float cadd(float a, float b) { return a + b; }
template <> __device__ inline double cadd(double a, double b) { return a + b; }
template <> __device__ inline float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
template <> __device__ inline double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }

struct SplitK {
  int64_t tiles_m, tiles_n, n_tiles;  // n_tiles = nz * tiles_m * tiles_n
  int64_t k_chunk, n_split;
  void* part;  // [n_split][nz * m * n] partial sums, or null when n_split == 1
};

// k split into n_split ranges of k_chunk (the last may be shorter).
struct KSplit {
  int64_t n_split = 1, k_chunk = 0;
};
constexpr size_t MAX_PARTIAL_BYTES = size_t(512) << 20;

// want ranges of at least min_len (rounded to `align`), within `avail` bytes of partials
static KSplit split_k(int64_t k, int64_t want, int64_t min_len, int64_t align, size_t part_bytes, size_t avail) {
  KSplit r;
  want = std::min<int64_t>(want, std::max<int64_t>(1, k / min_len));
  if (part_bytes) want = std::min<int64_t>(want, (int64_t)(avail / part_bytes));
  if (want <= 1) return r;
  r.k_chunk = ((k + want - 1) / want + align - 1) / align * align;
  r.n_split = (k + r.k_chunk - 1) / r.k_chunk;
  if (r.n_split <= 1) r = KSplit{};
  return r;
}

static void splitk_tile(int64_t n, int& tm, int& tn) {
  if (n >= 9) tm = 1, tn = 16;
  else if (n >= 5) tm = 2, tn = 8;
  else if (n >= 3) tm = 4, tn = 4;
  else if (n == 2) tm = 8, tn = 2;
  else tm = 16, tn = 1;
}

static KSplit plan_splitk(int dt, const GemmArgs& g, size_t avail) {
  int tm, tn;
  splitk_tile(g.n, tm, tn);
  const int64_t nz = g.nbo * g.nbi;
  const int64_t n_tiles = nz * ((g.m + tm - 1) / tm) * ((g.n + tn - 1) / tn);
  // enough workgroups to fill the chip, each k range at least 1024 long
  const int64_t want = std::max<int64_t>(1, 2048 / std::max<int64_t>(n_tiles, 1));
  return split_k(g.k, want, 1024, 256, dtype_size(dt) * (size_t)(nz * g.m * g.n), avail);
}

// Row tile of the native MFMA kernel.
static int mfma_rows(int dt, int64_t m) {
  const bool f64 = dt == F64 || dt == C128;
  if (m <= 32) return 32;
  if (m <= 64 || f64) return 64;
  return 128;
}

// Row tile of an MFMA kernel choice (GEMM_MFMA: from m).
static int mfma_rows_for(int dt, int kernel, int64_t m) {
  const bool f64 = dt == F64 || dt == C128;
  switch (kernel) {
    case GEMM_MFMA_R32: return 32;
    case GEMM_MFMA_R64: return 64;
    case GEMM_MFMA_R128: return f64 ? 64 : 128;
    default: return mfma_rows(dt, m);
  }
}

static KSplit plan_mfma_split(int dt, int prec, const GemmArgs& g, size_t avail, int bm) {
  if (prec != NATIVE) return KSplit{};
  const bool cplx = dtype_complex(dt);
  const int64_t bn = (dt == F64 || dt == C128) ? 64 : 128;
  const int64_t N = cplx ? 2 * g.n : g.n;
  const int64_t tiles = g.nbo * g.nbi * ((g.m + bm - 1) / bm) * ((N + bn - 1) / bn);
  if (tiles >= 512 || g.nbo * g.nbi > 65535) return KSplit{};
  // about 1024 workgroups, each k range at least 512 long (a multiple of the k tile)
  const int64_t want = (1024 + tiles - 1) / tiles;
  return split_k(g.k, want, 512, 64, dtype_size(dt) * (size_t)(g.nbo * g.nbi * g.m * g.n), avail);
}

static size_t scratch_avail(const GemmArgs& g) { return g.scratch && g.scratch_bytes > 256 ? g.scratch_bytes - 256 : 0; }

template <class T, int TM, int TN>
__global__ __launch_bounds__(256) void gemm_splitk_kernel(GemmArgs g, SplitK sk) {
  __shared__ T red[4][TM * TN];
  typename MaxOf<T>::R mx = 0;
  const int64_t q0 = (int64_t)blockIdx.y * sk.k_chunk;
  const int64_t q1 = q0 + sk.k_chunk < g.k ? q0 + sk.k_chunk : g.k;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int64_t tile = blockIdx.x; tile < sk.n_tiles; tile += gridDim.x) {
    const int64_t per_z = sk.tiles_m * sk.tiles_n;
    const int64_t z = tile / per_z, r = tile - z * per_z;
    const int64_t i0 = (r / sk.tiles_n) * TM, j0 = (r - (r / sk.tiles_n) * sk.tiles_n) * TN;
    const int64_t zo = z / g.nbi, zi = z - zo * g.nbi;
    const T* A = (const T*)g.a + zo * g.sao + zi * g.sai;
    const T* B = (const T*)g.b + zo * g.sbo + zi * g.sbi;
    T acc[TM][TN];
#pragma unroll
    for (int a = 0; a < TM; ++a)
#pragma unroll
      for (int b = 0; b < TN; ++b) acc[a][b] = CM<T>::zero();
#pragma unroll 2
    for (int64_t q = q0 + threadIdx.x; q < q1; q += 256) {
      T av[TM], bv[TN];
#pragma unroll
      for (int a = 0; a < TM; ++a) av[a] = i0 + a < g.m ? A[(i0 + a) * g.lda + q] : CM<T>::zero();
#pragma unroll
      for (int b = 0; b < TN; ++b) bv[b] = j0 + b < g.n ? B[b_index(g, q, j0 + b)] : CM<T>::zero();
#pragma unroll
      for (int a = 0; a < TM; ++a)
#pragma unroll
        for (int b = 0; b < TN; ++b) acc[a][b] = CM<T>::mac(acc[a][b], av[a], bv[b]);
    }
#pragma unroll
    for (int a = 0; a < TM; ++a)
#pragma unroll
      for (int b = 0; b < TN; ++b) {
        T v = acc[a][b];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v = cadd(v, Shfl<T>::down(v, o));
        if (lane == 0) red[wave][a * TN + b] = v;
      }
    __syncthreads();
    if (threadIdx.x < TM * TN) {
      const int u = threadIdx.x, a = u / TN, b = u - (u / TN) * TN;
      T v = cadd(cadd(cadd(red[0][u], red[1][u]), red[2][u]), red[3][u]);
      if (i0 + a < g.m && j0 + b < g.n) {
        if (sk.n_split == 1) {
          ((T*)g.c)[zo * g.sco + zi * g.sci + (i0 + a) * g.ldc + j0 + b] = v;
          if (g.cmax) fold_max(mx, v);
        } else {
          const int64_t n_out = (int64_t)g.nbo * g.nbi * g.m * g.n;
          ((T*)sk.part)[blockIdx.y * n_out + (z * g.m + i0 + a) * g.n + j0 + b] = v;
        }
      }
    }
    __syncthreads();
  }
  if (g.cmax && sk.n_split == 1) publish_max((double)mx, g.cmax);
}

template <class T>
__global__ __launch_bounds__(256) void splitk_reduce_kernel(GemmArgs g, SplitK sk) {
  const int64_t n_out = (int64_t)g.nbo * g.nbi * g.m * g.n;
  typename MaxOf<T>::R mx = 0;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n_out; e += (int64_t)gridDim.x * blockDim.x) {
    T v = ((const T*)sk.part)[e];
    for (int64_t sp = 1; sp < sk.n_split; ++sp) v = cadd(v, ((const T*)sk.part)[sp * n_out + e]);
    const int64_t mn = g.m * g.n, z = e / mn, r = e - z * mn, i = r / g.n, j = r - i * g.n;
    const int64_t zo = z / g.nbi, zi = z - zo * g.nbi;
    ((T*)g.c)[zo * g.sco + zi * g.sci + i * g.ldc + j] = v;
    if (g.cmax) fold_max(mx, v);
  }
  if (g.cmax) publish_max((double)mx, g.cmax);
}

// Tile order: column tile fastest. (An XCD-aware order - XCD j taking the
// contiguous tile range j/8 of the grid, row tile fastest so the tiles on one
// XCD share B panels in its L2 - cut the split GEMM's L2 refetch of B but ran
// 4% slower on the 30-qubit circuit's largest step, so it is not used; a plain
// row-tile-fastest order in the pre-split GEMM ran its four- and eight-row-tile
// steps 2% and 5% slower.)
__device__ inline void tile_origin(uint64_t t, uint64_t, int, int tiles_n, int bm, int bn, int64_t& m0, int64_t& n0) {
  m0 = (int64_t)(t / (uint64_t)tiles_n) * bm, n0 = (int64_t)(t % (uint64_t)tiles_n) * bn;
}

// --------------------------------------------------------------- MFMA GEMM --
template <class F> struct Mfma;
template <> struct Mfma<float> {
  using V = v4f;
  static __device__ V op(float a, float b, V c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
  static __device__ int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};
template <> struct Mfma<double> {
  using V = v4d;
  static __device__ V op(double a, double b, V c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
  static __device__ int row(int lane, int r) { return 4 * r + (lane >> 4); }
};

// 256 threads = 4 waves in a 2x2 arrangement, wave tile WM x WN built from
// 16x16x4 MFMA tiles. LDS: As[BK][BM+pad] (A transposed so a fragment read is
// 16 consecutive rows), Bs[BK][BN+pad]. Register prefetch of the next K tile
// overlaps global loads with MFMA work.
//   16x16x4 operand layout (measured on gfx950): A frag lane l = A[l%16][l/16],
//   B frag lane l = B[l/16][l%16]; output register r of lane l holds
//   D[4*(l/16) + r][l%16] for the f32 instruction and D[4*r + l/16][l%16] for
//   the f64 instruction.
// TRACK: fold the result max into g.cmax (a separate instantiation, so the
// plain kernel carries no extra live register or epilogue work). GATHER: B is
// read through the gather tables (1) or a factored column offset with int64 row
// offsets (2, GemmArgs::bcg), 0 row-major (separate instantiations: the
// row-offset loads in the k loop cost the row-major kernel a fifth of its rate).
template <class F, bool CPLX, int BM, int BN, int BK, bool TRACK, int GATHER>
__global__ __launch_bounds__(256) void gemm_mfma_kernel(GemmArgs g) {
  constexpr int WM = BM / 2, WN = BN / 2;
  constexpr int TM = WM / 16, TN = WN / 16;
  constexpr int PAD = 64 / (int)sizeof(F);  // row stride = 64 B mod 128 B: fragment reads hit all banks
  constexpr int SA = BM + PAD, SB = BN + PAD;
  constexpr int A_PER = BM * BK / 256;
  constexpr int B_PER = CPLX ? (BK / 2) * (BN / 2) / 256 : BK * BN / 256;
  static_assert(256 % (CPLX ? BN / 2 : BN) == 0, "a thread's B loads share one column");
  static_assert(A_PER >= 1 && B_PER >= 1, "tile too small for 256 threads");
  using V = typename Mfma<F>::V;
  // one LDS array: the A and B tiles, and in the epilogue 32-row slices of the
  // result tile (row stride BN + 16 bytes: the 16-lane groups of one store, whose
  // rows lie 4 or 1 apart, hit distinct banks)
  constexpr int SC = BN + 16 / (int)sizeof(F), EROWS = 32 * SC <= BK * (SA + SB) ? 32 : 16;
  static_assert(EROWS * SC <= BK * (SA + SB) && BM % EROWS == 0, "epilogue slice fits the tile arrays");
  __shared__ __align__(16) F sm[BK * (SA + SB)];
  F* const As = sm;
  F* const Bs = sm + BK * SA;

  const int64_t M = g.m, N = CPLX ? 2 * g.n : g.n, K = CPLX ? 2 * g.k : g.k;
  const int64_t lda = CPLX ? 2 * g.lda : g.lda;  // real units
  const int64_t ldc = CPLX ? 2 * g.ldc : g.ldc;
  const int tiles_n = (int)((N + BN - 1) / BN);
  const int tiles_m = (int)((M + BM - 1) / BM);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm0 = (wave >> 1) * WM, wn0 = (wave & 1) * WN;
  const uint64_t nz = (uint64_t)g.nbo * g.nbi;
  const uint64_t ntile = (uint64_t)tiles_m * tiles_n;
  F mx = 0;
  // k range of this blockIdx.z (real units; kc is a multiple of BK)
  const int64_t kbeg = g.kc ? (int64_t)blockIdx.z * g.kc * (CPLX ? 2 : 1) : 0;
  const int64_t kend = g.kc && kbeg + g.kc * (CPLX ? 2 : 1) < K ? kbeg + g.kc * (CPLX ? 2 : 1) : K;
  // Gather offsets of the B rows a thread loads in a k tile (the same in every
  // tile): those of the first two k tiles read once, the others one k tile
  // ahead of their B loads. Vector loads: a scalar load in flight would hold up
  // the LDS waits of the MFMA loop.
  struct RowOffs {
    typename std::conditional<GATHER == 2, int64_t, int32_t>::type v[B_PER];
  };
  auto rows_at = [&](int64_t k0) {
    RowOffs r{};
    if constexpr (GATHER)
#pragma unroll
      for (int u = 0; u < B_PER; ++u) {
        const int e = tid + 256 * u;
        const int64_t gr = CPLX ? k0 / 2 + e / (BN / 2) : k0 + e / BN;
        r.v[u] = gr < (CPLX ? kend / 2 : kend) ? (GATHER == 2 ? g.brow64[gr] : g.brow[gr]) : 0;
      }
    return r;
  };
  RowOffs rr0{}, rr1{};
  if constexpr (GATHER) rr0 = rows_at(kbeg), rr1 = rows_at(kbeg + BK);

  for (uint64_t z = blockIdx.y; z < nz; z += gridDim.y) {
    uint64_t zo = z / g.nbi, zi = z - zo * g.nbi;
    const F* A = (const F*)g.a + (zo * g.sao + zi * g.sai) * (CPLX ? 2 : 1);
    const F* B = (const F*)g.b + (zo * g.sbo + zi * g.sbi) * (CPLX ? 2 : 1);
    // with split k, partial tiles go to part[blockIdx.z][z][M][N] (real units)
    F* C = g.part ? (F*)g.part + ((uint64_t)blockIdx.z * nz + z) * (uint64_t)(M * N)
                  : (F*)g.c + (zo * g.sco + zi * g.sci) * (CPLX ? 2 : 1);
    const int64_t ldo = g.part ? N : ldc;
    for (uint64_t t = blockIdx.x; t < ntile; t += gridDim.x) {
      int64_t m0, n0;
      tile_origin(t, ntile, tiles_m, tiles_n, BM, BN, m0, n0);
      V acc[TM][TN];
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) acc[i][j] = V{0, 0, 0, 0};

      F ra[A_PER];
      F rb[CPLX ? 2 * B_PER : B_PER];
      // every B element a thread loads lies in one column (256 is a multiple of BN)
      const int64_t bcn = CPLX ? n0 / 2 + tid % (BN / 2) : n0 + tid % BN;
      const bool bcn_ok = bcn < (CPLX ? g.n : N);
      const int64_t bco = !bcn_ok ? 0 : GATHER == 2 ? fac_col(g, bcn) : GATHER ? (int64_t)g.bcol[bcn] : bcn;
      auto load = [&](int64_t k0, const RowOffs& ro) {
#pragma unroll
        for (int u = 0; u < A_PER; ++u) {
          int e = tid + 256 * u, col = e % BK, row = e / BK;
          int64_t gr = m0 + row, gc = k0 + col;
          ra[u] = (gr < M && gc < kend) ? A[gr * lda + gc] : (F)0;
        }
        if constexpr (CPLX) {
#pragma unroll
          for (int u = 0; u < B_PER; ++u) {
            int e = tid + 256 * u, p = e / (BN / 2);
            int64_t gp = k0 / 2 + p;
            bool ok = gp < kend / 2 && bcn_ok;
            const int64_t bi = ok ? ((GATHER ? (int64_t)ro.v[u] : gp * g.ldb) + bco) * 2 : 0;
            rb[2 * u] = ok ? B[bi] : (F)0;
            rb[2 * u + 1] = ok ? B[bi + 1] : (F)0;
          }
        } else {
#pragma unroll
          for (int u = 0; u < B_PER; ++u) {
            int e = tid + 256 * u, row = e / BN;
            int64_t gr = k0 + row;
            rb[u] = (gr < kend && bcn_ok) ? B[(GATHER ? (int64_t)ro.v[u] : gr * g.ldb) + bco] : (F)0;
          }
        }
      };
      auto store = [&]() {
#pragma unroll
        for (int u = 0; u < A_PER; ++u) {
          int e = tid + 256 * u, col = e % BK, row = e / BK;
          As[col * SA + row] = ra[u];
        }
        if constexpr (CPLX) {
#pragma unroll
          for (int u = 0; u < B_PER; ++u) {
            int e = tid + 256 * u, j = e % (BN / 2), p = e / (BN / 2);
            F br = rb[2 * u], bi = rb[2 * u + 1];
            Bs[(2 * p) * SB + 2 * j] = br;
            Bs[(2 * p) * SB + 2 * j + 1] = bi;
            Bs[(2 * p + 1) * SB + 2 * j] = -bi;
            Bs[(2 * p + 1) * SB + 2 * j + 1] = br;
          }
        } else {
#pragma unroll
          for (int u = 0; u < B_PER; ++u) {
            int e = tid + 256 * u, col = e % BN, row = e / BN;
            Bs[row * SB + col] = rb[u];
          }
        }
      };

      const int64_t nk = (kend - kbeg + BK - 1) / BK;
      load(kbeg, rr0);
      RowOffs rn = rr1;  // row offsets of k tile kt + 1
      store();
      __syncthreads();
      for (int64_t kt = 0; kt < nk; ++kt) {
        if (kt + 1 < nk) {
          load(kbeg + (kt + 1) * BK, rn);
          if constexpr (GATHER)
            if (kt + 2 < nk) rn = rows_at(kbeg + (kt + 2) * BK);
        }
        __builtin_amdgcn_s_setprio(1);  // the MFMA run at raised priority (as in gemm_presplit_kernel)
#pragma unroll
        for (int kk = 0; kk < BK; kk += 4) {
          F af[TM], bf[TN];
          const int kr = kk + (lane >> 4), cl = lane & 15;
#pragma unroll
          for (int i = 0; i < TM; ++i) af[i] = As[kr * SA + wm0 + i * 16 + cl];
#pragma unroll
          for (int j = 0; j < TN; ++j) bf[j] = Bs[kr * SB + wn0 + j * 16 + cl];
#pragma unroll
          for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int j = 0; j < TN; ++j) acc[i][j] = Mfma<F>::op(af[i], bf[j], acc[i][j]);
        }
        __builtin_amdgcn_s_setprio(0);
        __syncthreads();
        if (kt + 1 < nk) {
          store();
          __syncthreads();
        }
      }
      const int cl = lane & 15;
      if constexpr (!CPLX) {  // real results: straight from the MFMA layout
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
          for (int j = 0; j < TN; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              const int64_t row = m0 + wm0 + i * 16 + Mfma<F>::row(lane, r), col = n0 + wn0 + j * 16 + cl;
              if (row < M && col < N) {
                C[row * ldo + col] = acc[i][j][r];
                if constexpr (TRACK) fold_max(mx, acc[i][j][r]);
              }
            }
        continue;
      }
      // Complex results: their real and imaginary parts come from different waves,
      // so the MFMA layout alone stores 4-byte values at an 8-byte stride. They go
      // through LDS instead, 32 (or 16) rows of the tile at a time: a row-major
      // slice, then every thread writes 16-byte row pieces (the circuit's
      // contraction 97.75 -> 97.0 ms; real results measured slower this way).
      constexpr int E = 16 / (int)sizeof(F);  // elements per 16-byte piece
      using W = typename std::conditional<sizeof(F) == 4, float4, double2>::type;
      const bool c16 = ((uintptr_t)C & 15) == 0 && ldo % E == 0;
#pragma unroll
      for (int r0 = 0; r0 < BM; r0 += EROWS) {
        __syncthreads();  // the LDS arrays are free (k loop or the previous slice done)
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int tr = wm0 + i * 16 + Mfma<F>::row(lane, r) - r0;
            if (tr >= 0 && tr < EROWS)
#pragma unroll
              for (int j = 0; j < TN; ++j) sm[tr * SC + wn0 + j * 16 + cl] = acc[i][j][r];
          }
        __syncthreads();
        for (int q = tid; q < EROWS * BN / E; q += 256) {
          const int tr = q / (BN / E), ce = (q % (BN / E)) * E;
          const int64_t row = m0 + r0 + tr, col = n0 + ce;
          if (row < M) {
            const F* src = &sm[tr * SC + ce];
            if (c16 && col + E <= N) {
              *reinterpret_cast<W*>(C + row * ldo + col) = *reinterpret_cast<const W*>(src);
              if constexpr (TRACK)
#pragma unroll
                for (int e = 0; e < E; ++e) fold_max(mx, src[e]);
            } else {
              for (int e = 0; e < E; ++e)
                if (col + e < N) {
                  C[row * ldo + col + e] = src[e];
                  if constexpr (TRACK) fold_max(mx, src[e]);
                }
            }
          }
        }
      }
      __syncthreads();  // the next tile stages into the same LDS
    }
  }
  if constexpr (TRACK) publish_max((double)mx, g.cmax);
}

// ------------------------------------------------------- complex 3M GEMM --
// complex64 C = A.B as three real products on the FP32 MFMA (Gauss's 3M):
//   T1 = Ar.Br, T2 = Ai.Bi, T3 = (Ar + Ai).(Br + Bi);  Cr = T1 - T2, Ci = T3 - T1 - T2
// 3 real multiply-adds per complex one instead of the 4 of the 2x2 expansion
// (gemm_mfma_kernel with CPLX). The sums Ar + Ai and Br + Bi are formed in FP32
// while staging; Ci comes out of the cancellation T3 - T1 - T2, so its error
// bound grows with |Ar + Ai||Br + Bi| rather than |Ar||Bi| + |Ai||Br|.
// Tile: BM rows x BN complex columns x BK complex k; 4 waves in a 2x2
// arrangement, three accumulator planes per wave tile. LDS holds the three A
// planes [BK][BM+pad] (transposed) and the three B planes [BK][BN+pad].
// Same batch, split-k (part[z][M][n] complex), gather-table and max-tracking
// conventions as gemm_mfma_kernel.
template <int BM, int BN, int BK, bool TRACK, bool GATHER>
__global__ __launch_bounds__(256) void gemm_c3m_kernel(GemmArgs g) {
  constexpr int WM = BM / 2, WN = BN / 2;
  constexpr int TM = WM / 16, TN = WN / 16;
  constexpr int PAD = 16;
  constexpr int SA = BM + PAD, SB = BN + PAD;
  constexpr int A_PER = BM * BK / 256, B_PER = BK * BN / 256;
  static_assert(256 % BN == 0 && 256 % BK == 0 && A_PER >= 1 && B_PER >= 1, "tile shape");
  using V = v4f;
  __shared__ float As[3][BK * SA];
  __shared__ float Bs[3][BK * SB];

  const int64_t M = g.m, N = g.n, K = g.k;  // complex units
  const int tiles_n = (int)((N + BN - 1) / BN), tiles_m = (int)((M + BM - 1) / BM);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm0 = (wave >> 1) * WM, wn0 = (wave & 1) * WN;
  const uint64_t nz = (uint64_t)g.nbo * g.nbi;
  const uint64_t ntile = (uint64_t)tiles_m * tiles_n;
  const int64_t kbeg = g.kc ? (int64_t)blockIdx.z * g.kc : 0;
  const int64_t kend = g.kc && kbeg + g.kc < K ? kbeg + g.kc : K;
  float mx = 0;

  for (uint64_t z = blockIdx.y; z < nz; z += gridDim.y) {
    uint64_t zo = z / g.nbi, zi = z - zo * g.nbi;
    const float2* A = (const float2*)g.a + (zo * g.sao + zi * g.sai);
    const float2* B = (const float2*)g.b + (zo * g.sbo + zi * g.sbi);
    float2* C = g.part ? (float2*)g.part + ((uint64_t)blockIdx.z * nz + z) * (uint64_t)(M * N)
                       : (float2*)g.c + (zo * g.sco + zi * g.sci);
    const int64_t ldo = g.part ? N : g.ldc;
    for (uint64_t t = blockIdx.x; t < ntile; t += gridDim.x) {
      int64_t m0, n0;
      tile_origin(t, ntile, tiles_m, tiles_n, BM, BN, m0, n0);
      V acc[3][TM][TN];
#pragma unroll
      for (int q = 0; q < 3; ++q)
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
          for (int j = 0; j < TN; ++j) acc[q][i][j] = V{0, 0, 0, 0};
      float2 ra[A_PER], rb[B_PER];
      // a thread's B loads share one column (256 is a multiple of BN)
      const int64_t bcn = n0 + tid % BN;
      const bool bcn_ok = bcn < N;
      const int64_t bco = bcn_ok ? (GATHER ? (int64_t)g.bcol[bcn] : bcn) : 0;
      auto load = [&](int64_t k0) {
#pragma unroll
        for (int u = 0; u < A_PER; ++u) {
          const int e = tid + 256 * u, col = e % BK, row = e / BK;
          const int64_t gr = m0 + row, gc = k0 + col;
          ra[u] = (gr < M && gc < kend) ? A[gr * g.lda + gc] : make_float2(0.f, 0.f);
        }
#pragma unroll
        for (int u = 0; u < B_PER; ++u) {
          const int e = tid + 256 * u, p = e / BN;
          const int64_t gp = k0 + p;
          const bool ok = gp < kend && bcn_ok;
          rb[u] = ok ? B[(GATHER ? (int64_t)g.brow[gp] : gp * g.ldb) + bco] : make_float2(0.f, 0.f);
        }
      };
      auto store = [&]() {
#pragma unroll
        for (int u = 0; u < A_PER; ++u) {
          const int e = tid + 256 * u, col = e % BK, row = e / BK;
          As[0][col * SA + row] = ra[u].x;
          As[1][col * SA + row] = ra[u].y;
          As[2][col * SA + row] = ra[u].x + ra[u].y;
        }
#pragma unroll
        for (int u = 0; u < B_PER; ++u) {
          const int e = tid + 256 * u, p = e / BN, j = e % BN;
          Bs[0][p * SB + j] = rb[u].x;
          Bs[1][p * SB + j] = rb[u].y;
          Bs[2][p * SB + j] = rb[u].x + rb[u].y;
        }
      };
      const int64_t nk = (kend - kbeg + BK - 1) / BK;
      load(kbeg);
      store();
      __syncthreads();
      for (int64_t kt = 0; kt < nk; ++kt) {
        if (kt + 1 < nk) load(kbeg + (kt + 1) * BK);
#pragma unroll
        for (int kk = 0; kk < BK; kk += 4) {
          const int kr = kk + (lane >> 4), cl = lane & 15;
#pragma unroll
          for (int q = 0; q < 3; ++q) {
            float af[TM], bf[TN];
#pragma unroll
            for (int i = 0; i < TM; ++i) af[i] = As[q][kr * SA + wm0 + i * 16 + cl];
#pragma unroll
            for (int j = 0; j < TN; ++j) bf[j] = Bs[q][kr * SB + wn0 + j * 16 + cl];
#pragma unroll
            for (int i = 0; i < TM; ++i)
#pragma unroll
              for (int j = 0; j < TN; ++j) acc[q][i][j] = Mfma<float>::op(af[i], bf[j], acc[q][i][j]);
          }
        }
        __syncthreads();
        if (kt + 1 < nk) {
          store();
          __syncthreads();
        }
      }
      const int cl = lane & 15;
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int64_t row = m0 + wm0 + i * 16 + Mfma<float>::row(lane, r), col = n0 + wn0 + j * 16 + cl;
            if (row < M && col < N) {
              const float t1 = acc[0][i][j][r], t2 = acc[1][i][j][r], t3 = acc[2][i][j][r];
              const float2 v = make_float2(t1 - t2, (t3 - t1) - t2);
              C[row * ldo + col] = v;
              if constexpr (TRACK) fold_max(mx, v);
            }
          }
    }
  }
  if constexpr (TRACK) publish_max((double)mx, g.cmax);
}

// ------------------------------------------------------ split-precision GEMM --
// FP32 (or complex64 via the 2x2 expansion of B) products on the 16-bit MFMA:
// every operand value x is split into hi = round16(x * 2^s) and
// lo = round16(x * 2^s - hi) while staging into LDS, and
//   A.B ~= (Ahi.Bhi + Ahi.Blo + Alo.Bhi) * 2^-(sa+sb)
// accumulated in FP32 (each 16-bit product is exact in FP32; the dropped
// lo.lo term is ~2^-22 relative for FP16 halves, ~2^-16 for BF16 halves).
// BF16 keeps FP32's exponent range, so s = 0; FP16 uses a per-GEMM power-of-
// two scale from the operands' max magnitudes (2^s * max < 2^15).
//   16x16x32 layout: A frag lane l holds A[l%16][8*(l/16) .. +7], B frag lane l
//   holds B[8*(l/16) .. +7][l%16]; output register r of lane l is D[4*(l/16)+r][l%16].
template <bool F16> struct Half;
template <> struct Half<false> {
  using H = __bf16;
  using V = v8bf;
  static __device__ v4f mfma(V a, V b, v4f c) { return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }
};
template <> struct Half<true> {
  using H = _Float16;
  using V = v8h;
  static __device__ v4f mfma(V a, V b, v4f c) { return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0); }
};

template <bool F16>
__device__ inline void split16(float x, typename Half<F16>::H& hi, typename Half<F16>::H& lo) {
  using H = typename Half<F16>::H;
  hi = (H)x;
  lo = (H)(x - (float)hi);
}

__device__ inline int pow2_shift(unsigned long long bits) {
  double m = __longlong_as_double((long long)bits);
  if (!(m > 0.0) || !isfinite(m)) return 0;
  int e;
  frexp(m, &e);  // m = f * 2^e, f in [0.5, 1)
  return 15 - e;
}

// The operands' pending power-of-two exponents (GemmArgs::apend / bpend), divided
// out of the product with the FP16 operand scales.
__device__ inline int pending_exp(const GemmArgs& g) {
  return (g.apend ? (int)*g.apend : 0) + (g.bpend ? (int)*g.bpend : 0);
}

// Staging: A rows are k-contiguous in LDS (ahi/alo[m][k]); B is stored
// transposed (bhi/blo[n][k]) so both MFMA fragments are one 16-byte LDS read.
// VEC: A rows are 16-byte aligned (float4 loads of 4 consecutive k).
template <bool CPLX, bool F16, bool VEC>
__global__ __launch_bounds__(256, 2) void gemm_split_kernel(GemmArgs g) {
  using H = typename Half<F16>::H;
  using V = typename Half<F16>::V;
  constexpr int BM = 128, BN = 128, BK = 64;
  constexpr int WM = 64, WN = 64, TM = 4, TN = 4;
  constexpr int SK = BK + 8;  // row stride in halves (144 B, 16-byte aligned rows)
  extern __shared__ __align__(16) char smem_split[];
  H* ahi = reinterpret_cast<H*>(smem_split);
  H* alo = ahi + BM * SK;
  H* bhi = alo + BM * SK;
  H* blo = bhi + BN * SK;

  const int pe = pending_exp(g);
  float sa = 1.f, sb = 1.f, unscale = ldexpf(1.f, -pe);
  float mx = 0.f;
  if constexpr (F16) {
    int ea = pow2_shift(*g.amax), eb = pow2_shift(*g.bmax);
    sa = ldexpf(1.f, ea), sb = ldexpf(1.f, eb), unscale = ldexpf(1.f, -(ea + eb + pe));
  }
  const int64_t M = g.m, N = CPLX ? 2 * g.n : g.n, K = CPLX ? 2 * g.k : g.k;
  const int64_t lda = CPLX ? 2 * g.lda : g.lda, ldc = CPLX ? 2 * g.ldc : g.ldc;
  const int tiles_n = (int)((N + BN - 1) / BN), tiles_m = (int)((M + BM - 1) / BM);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm0 = (wave >> 1) * WM, wn0 = (wave & 1) * WN;
  const uint64_t nz = (uint64_t)g.nbo * g.nbi, ntile = (uint64_t)tiles_m * tiles_n;
  constexpr int A_Q = BM * BK / 4 / 256;  // float4 groups of A per thread (8)
  // B items per thread: complex -> (complex row pair, complex column); real -> (4 k rows, column)
  constexpr int B_Q = CPLX ? (BK / 4) * (BN / 2) / 256 : (BK / 4) * BN / 256;  // 4 / 8
  static_assert(256 % (CPLX ? BN / 2 : BN) == 0, "a thread's B loads share one column");

  for (uint64_t z = blockIdx.y; z < nz; z += gridDim.y) {
    uint64_t zo = z / g.nbi, zi = z - zo * g.nbi;
    const float* A = (const float*)g.a + (zo * g.sao + zi * g.sai) * (CPLX ? 2 : 1);
    const float* B = (const float*)g.b + (zo * g.sbo + zi * g.sbi) * (CPLX ? 2 : 1);
    float* C = (float*)g.c + (zo * g.sco + zi * g.sci) * (CPLX ? 2 : 1);
    for (uint64_t t = blockIdx.x; t < ntile; t += gridDim.x) {
      int64_t m0, n0;
      tile_origin(t, ntile, tiles_m, tiles_n, BM, BN, m0, n0);
      v4f acc[TM][TN];
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) acc[i][j] = v4f{0, 0, 0, 0};
      float4 ra[A_Q];
      float4 rb[B_Q];
      // every B element a thread loads lies in one column (256 is a multiple of BN)
      const int64_t bcn = CPLX ? n0 / 2 + tid % (BN / 2) : n0 + tid % BN;
      const bool bcn_ok = bcn < (CPLX ? g.n : N);
      const int64_t bco = bcn_ok ? b_col(g, bcn) : 0;
      auto load = [&](int64_t k0) {
#pragma unroll
        for (int u = 0; u < A_Q; ++u) {
          int q = tid + 256 * u, row = q / (BK / 4), c4 = (q % (BK / 4)) * 4;
          int64_t gr = m0 + row, gc = k0 + c4;
          const float* src = A + gr * lda + gc;
          if (VEC && gr < M && gc + 3 < K) {
            ra[u] = *reinterpret_cast<const float4*>(src);
          } else {
            ra[u].x = (gr < M && gc < K) ? src[0] : 0.f;
            ra[u].y = (gr < M && gc + 1 < K) ? src[1] : 0.f;
            ra[u].z = (gr < M && gc + 2 < K) ? src[2] : 0.f;
            ra[u].w = (gr < M && gc + 3 < K) ? src[3] : 0.f;
          }
        }
        if constexpr (CPLX) {
#pragma unroll
          for (int u = 0; u < B_Q; ++u) {  // complex rows p, p+1 of one complex column j
            int q = tid + 256 * u, pp = q / (BN / 2);
            int64_t gp = k0 / 2 + 2 * pp;
            bool ok0 = bcn_ok && gp < g.k, ok1 = bcn_ok && gp + 1 < g.k;
            const float* b0 = ok0 ? B + (b_row(g, gp) + bco) * 2 : B;
            const float* b1 = ok1 ? B + (b_row(g, gp + 1) + bco) * 2 : B;
            rb[u].x = ok0 ? b0[0] : 0.f;
            rb[u].y = ok0 ? b0[1] : 0.f;
            rb[u].z = ok1 ? b1[0] : 0.f;
            rb[u].w = ok1 ? b1[1] : 0.f;
          }
        } else {
#pragma unroll
          for (int u = 0; u < B_Q; ++u) {  // 4 consecutive k rows of one column
            int q = tid + 256 * u, r4 = (q / BN) * 4;
            int64_t gr = k0 + r4;
            bool okc = bcn_ok;
            rb[u].x = (okc && gr < K) ? B[b_row(g, gr) + bco] : 0.f;
            rb[u].y = (okc && gr + 1 < K) ? B[b_row(g, gr + 1) + bco] : 0.f;
            rb[u].z = (okc && gr + 2 < K) ? B[b_row(g, gr + 2) + bco] : 0.f;
            rb[u].w = (okc && gr + 3 < K) ? B[b_row(g, gr + 3) + bco] : 0.f;
          }
        }
      };
      auto put4 = [&](H* hi, H* lo, int off, float x0, float x1, float x2, float x3) {
        H h[4], l[4];
        split16<F16>(x0, h[0], l[0]);
        split16<F16>(x1, h[1], l[1]);
        split16<F16>(x2, h[2], l[2]);
        split16<F16>(x3, h[3], l[3]);
        *reinterpret_cast<uint2*>(hi + off) = *reinterpret_cast<const uint2*>(h);
        *reinterpret_cast<uint2*>(lo + off) = *reinterpret_cast<const uint2*>(l);
      };
      auto store = [&]() {
#pragma unroll
        for (int u = 0; u < A_Q; ++u) {
          int q = tid + 256 * u, row = q / (BK / 4), c4 = (q % (BK / 4)) * 4;
          put4(ahi, alo, row * SK + c4, ra[u].x * sa, ra[u].y * sa, ra[u].z * sa, ra[u].w * sa);
        }
        if constexpr (CPLX) {
#pragma unroll
          for (int u = 0; u < B_Q; ++u) {
            // Bx^T rows: n' = 2j holds (br, -bi) per complex k, n' = 2j+1 holds (bi, br).
            int q = tid + 256 * u, j = q % (BN / 2), pp = q / (BN / 2);
            float br0 = rb[u].x * sb, bi0 = rb[u].y * sb, br1 = rb[u].z * sb, bi1 = rb[u].w * sb;
            put4(bhi, blo, (2 * j) * SK + 4 * pp, br0, -bi0, br1, -bi1);
            put4(bhi, blo, (2 * j + 1) * SK + 4 * pp, bi0, br0, bi1, br1);
          }
        } else {
#pragma unroll
          for (int u = 0; u < B_Q; ++u) {
            int q = tid + 256 * u, col = q % BN, r4 = (q / BN) * 4;
            put4(bhi, blo, col * SK + r4, rb[u].x * sb, rb[u].y * sb, rb[u].z * sb, rb[u].w * sb);
          }
        }
      };
      const int64_t nk = (K + BK - 1) / BK;
      load(0);
      store();
      __syncthreads();
      const int cl = lane & 15;
      for (int64_t kt = 0; kt < nk; ++kt) {
        if (kt + 1 < nk) load((kt + 1) * BK);
#pragma unroll
        for (int ks = 0; ks < BK; ks += 32) {
          const int kg = ks + (lane >> 4) * 8;
          V ah[TM], al[TM];
#pragma unroll
          for (int i = 0; i < TM; ++i) {
            ah[i] = *reinterpret_cast<const V*>(&ahi[(wm0 + i * 16 + cl) * SK + kg]);
            al[i] = *reinterpret_cast<const V*>(&alo[(wm0 + i * 16 + cl) * SK + kg]);
          }
#pragma unroll
          for (int j = 0; j < TN; ++j) {
            V bh = *reinterpret_cast<const V*>(&bhi[(wn0 + j * 16 + cl) * SK + kg]);
            V bl = *reinterpret_cast<const V*>(&blo[(wn0 + j * 16 + cl) * SK + kg]);
#pragma unroll
            for (int i = 0; i < TM; ++i) {
              acc[i][j] = Half<F16>::mfma(al[i], bh, acc[i][j]);
              acc[i][j] = Half<F16>::mfma(ah[i], bl, acc[i][j]);
              acc[i][j] = Half<F16>::mfma(ah[i], bh, acc[i][j]);
            }
          }
        }
        __syncthreads();
        if (kt + 1 < nk) {
          store();
          __syncthreads();
        }
      }
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            int64_t row = m0 + wm0 + i * 16 + 4 * (lane >> 4) + r, col = n0 + wn0 + j * 16 + cl;
            if (row < M && col < N) {
              C[row * ldc + col] = acc[i][j][r] * unscale;
              if (g.cmax) fold_max(mx, acc[i][j][r] * unscale);
            }
          }
    }
  }
  if (g.cmax) publish_max((double)mx, g.cmax);
}

constexpr size_t SPLIT_LDS = 4 * 128 * (64 + 8) * 2;
// Measured on gfx950 (complex64 1024x512x8192 x64, BF16x3): BK 64 with one LDS
// buffer (two workgroups per CU) 302 TFLOP/s algorithmic, BK 32 with two 277.
constexpr int PRESPLIT_BK = 64;
// (hi, lo) halves of a BM x BK tile of A and a 128 x BK tile of B', then the
// raw-B row table (rows words)
constexpr size_t presplit_lds(int bm, int64_t rows = 0) {
  return 2 * (size_t)(bm + 128) * PRESPLIT_BK * 2 + (size_t)(rows + 3) / 4 * 16;
}
// The most raw-B rows a row table holds (16 KiB: 80 KiB of LDS with 128-row
// tiles, two workgroups per CU); longer k runs with pre-split B.
constexpr int64_t PRESPLIT_MAX_TABLE = 4096;

// ------------------------------------------------------------------ dispatch --
// ------------------------------------------------ pre-split operand GEMM --
// Each operand is split once into (hi, lo) 16-bit arrays by a memory-bound
// pass, so the GEMM kernel only copies 16-byte chunks into LDS and issues
// MFMAs (no per-tile conversion work, and the transposition of B happens in
// the pass, not in LDS stores). Layouts, with Kp = K' rounded up to 8 halves
// and zero padded:
//   A: [za][m][Kp]            (complex A read as real [m][2k])
//   B: [zb][n][Kp] transposed (complex: row j holds re, im of B[p][j] at 2p, 2p+1)
// za / zb count the distinct operand matrices of the batch (1 when shared).
struct PreSplit {
  const void *ahi, *alo, *bhi, *blo;
  int64_t kp;
  // split k (pre-split B): blockIdx.z takes halves [z * kc, (z + 1) * kc) of kp and
  // writes its partial product to part[z][batch][m][n] (the result's element type),
  // summed in range order by splitk_reduce_kernel; kc = 0: all of k, into C
  int64_t kc = 0;
  void* part = nullptr;
};

static __host__ __device__ inline int64_t n_mats(int64_t so, int64_t si, int64_t nbo, int64_t nbi) {
  return (so ? nbo : 1) * (si ? nbi : 1);
}

// A -> [za][m][Kp] hi/lo; one thread per 4 consecutive halves of a row.
template <bool CPLX, bool F16>
__global__ __launch_bounds__(256) void split_a_kernel(GemmArgs g, PreSplit ps) {
  using H = typename Half<F16>::H;
  const int64_t K = CPLX ? 2 * g.k : g.k, lda = CPLX ? 2 * g.lda : g.lda, kp = ps.kp;
  const int64_t inner = g.sai ? g.nbi : 1, nm = n_mats(g.sao, g.sai, g.nbo, g.nbi);
  const int64_t per_row = kp / 4, total = nm * g.m * per_row;
  float sa = 1.f;
  if constexpr (F16) sa = ldexpf(1.f, pow2_shift(*g.amax));
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t mat = e / (g.m * per_row), r = e - mat * g.m * per_row, row = r / per_row, c = (r - row * per_row) * 4;
    const int64_t zo = mat / inner, zi = mat - zo * inner;
    const float* base = (const float*)g.a + (zo * g.sao + zi * g.sai) * (CPLX ? 2 : 1);
    const float* src = base + row * lda;
    H h[4], l[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      float x = 0.f;
      if (c + u < K) {
        if (g.acol) {  // A through gather tables (one batch entry): element (row, p) at arow[row] + acol[p]
          const int64_t p = CPLX ? (c + u) >> 1 : c + u;
          const int64_t off = (int64_t)g.arow[row] + (int64_t)g.acol[p];
          x = CPLX ? base[2 * off + ((c + u) & 1)] : base[off];
        } else {
          x = src[c + u];
        }
        x *= sa;
      }
      split16<F16>(x, h[u], l[u]);
    }
    const int64_t off = (mat * g.m + row) * kp + c;
    *reinterpret_cast<uint2*>((H*)ps.ahi + off) = *reinterpret_cast<const uint2*>(h);
    *reinterpret_cast<uint2*>((H*)ps.alo + off) = *reinterpret_cast<const uint2*>(l);
  }
}

// B [k][n] -> transposed [zb][n][Kp] hi/lo through a tk_split_bt (k) x 64 (n) LDS
// tile: 64 halves of k per tile, so that a column's piece of each output array is
// one whole 128-byte line (measured on gfx950: a 64-deep complex64 tile ran the
// k=512 n=8192 x256 pass in 6.3 ms against 4.3 for 32).
constexpr int tk_split_bt(bool cplx) { return cplx ? 32 : 64; }
template <bool CPLX, bool F16>
__global__ __launch_bounds__(256) void split_bt_kernel(GemmArgs g, PreSplit ps) {
  using H = typename Half<F16>::H;
  using E = typename std::conditional<CPLX, float2, float>::type;
  constexpr int TK = tk_split_bt(CPLX), TN = 64, EH = CPLX ? 2 : 1;  // halves per element
  __shared__ E tile[TK][TN + 1];
  const int64_t kp = ps.kp, inner = g.sbi ? g.nbi : 1, nm = n_mats(g.sbo, g.sbi, g.nbo, g.nbi);
  const int64_t tk = (kp / EH + TK - 1) / TK, tn = (g.n + TN - 1) / TN, total = nm * tk * tn;
  float sb = 1.f;
  if constexpr (F16) sb = ldexpf(1.f, pow2_shift(*g.bmax));
  const int t = threadIdx.x;
  for (int64_t b = blockIdx.x; b < total; b += gridDim.x) {
    const int64_t mat = b / (tk * tn), r = b - mat * tk * tn, k0 = (r / tn) * TK, n0 = (r - (r / tn) * tn) * TN;
    const int64_t zo = mat / inner, zi = mat - zo * inner;
    const E* src = (const E*)g.b + zo * g.sbo + zi * g.sbi;
    const int64_t gcn = n0 + t % TN;  // the thread's column in every load below
    const int64_t gco = gcn < g.n ? b_col(g, gcn) : 0;
#pragma unroll
    for (int s = 0; s < TK / 4; ++s) {
      const int kk = t / TN + 4 * s, nn = t % TN;
      const int64_t gk = k0 + kk;
      E v{};
      if (gk < g.k && gcn < g.n) v = src[b_row(g, gk) + gco];
      tile[kk][nn] = v;
    }
    __syncthreads();
#pragma unroll
    for (int w = t; w < TN * (TK / 8); w += 256) {  // item -> (row n, 8 consecutive k elements)
      const int nn = w / (TK / 8), kc = (w % (TK / 8)) * 8;
      const int64_t gn = n0 + nn;
      if (gn < g.n) {
        H h[8 * EH], l[8 * EH];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
          E v = tile[kc + i][nn];
          if constexpr (CPLX) {
            split16<F16>(v.x * sb, h[2 * i], l[2 * i]);
            split16<F16>(v.y * sb, h[2 * i + 1], l[2 * i + 1]);
          } else {
            split16<F16>(v * sb, h[i], l[i]);
          }
        }
        const int64_t col = (k0 + kc) * EH;  // halves
        if (col < kp) {
          const int64_t off = (mat * g.n + gn) * kp + col;
          // kp is a multiple of 8 halves and col of 8: whole 16-byte chunks stay in the row
#pragma unroll
          for (int c = 0; c < EH; ++c)
            if (col + 8 * c < kp) {
              *reinterpret_cast<uint4*>((H*)ps.bhi + off + 8 * c) = *reinterpret_cast<const uint4*>(h + 8 * c);
              *reinterpret_cast<uint4*>((H*)ps.blo + off + 8 * c) = *reinterpret_cast<const uint4*>(l + 8 * c);
            }
        }
      }
    }
    __syncthreads();
  }
}

// Complex B' rows from a compact chunk of (re, im) pairs: row 2j = (re, -im),
// row 2j + 1 = (im, re). 16-bit halves, little-endian pairs in 32-bit words.
__device__ inline uint4 neg_im(uint4 v) { return make_uint4(v.x ^ 0x80000000u, v.y ^ 0x80000000u, v.z ^ 0x80000000u, v.w ^ 0x80000000u); }
__device__ inline uint32_t swap16(uint32_t x) { return (x >> 16) | (x << 16); }
__device__ inline uint4 swap_pairs(uint4 v) { return make_uint4(swap16(v.x), swap16(v.y), swap16(v.z), swap16(v.w)); }

// (hi, lo) 16-bit pairs of two floats, x ~ hi + lo, as packed 32-bit words: one
// packed conversion for hi, the residuals x - hi (exact in FP32), one for lo;
// the same values as split16 on each float.
typedef float v2f __attribute__((ext_vector_type(2)));
template <bool F16> struct HalfPair;
template <> struct HalfPair<false> { typedef __bf16 T __attribute__((ext_vector_type(2))); };
template <> struct HalfPair<true> { typedef _Float16 T __attribute__((ext_vector_type(2))); };
template <bool F16>
__device__ inline void split_pair(v2f x, uint32_t& hi, uint32_t& lo) {
  using P = typename HalfPair<F16>::T;
  const P h = __builtin_convertvector(x, P);
  const P l = __builtin_convertvector(x - __builtin_convertvector(h, v2f), P);
  hi = __builtin_bit_cast(uint32_t, h), lo = __builtin_bit_cast(uint32_t, l);
}

// Buffer resources: a load at a byte offset at or past the resource's size
// returns zeros, so rows past a matrix and the sentinel offsets below read 0
// with no branch per load. Raw B offsets (row + column, bytes) stay below
// RAW_LIMIT; a row past k carries ROW_OOB, a column past n COL_OOB, and any
// sum with a sentinel lands at or past RAW_LIMIT without wrapping 32 bits.
__device__ inline __amdgpu_buffer_rsrc_t buf_rsrc(const void* base, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes, 0x00020000);
}
constexpr uint32_t RAW_LIMIT = 1u << 30, ROW_OOB = 1u << 31, COL_OOB = 1u << 30;
__device__ inline uint32_t tile_bytes(int64_t rows_left, int64_t row_bytes) {
  return (uint32_t)std::min<int64_t>(rows_left * row_bytes, 0x7fffffff);
}

// LDS tiles are 128 rows of BK halves (no padding), one buffer (64 KiB, two
// workgroups per CU); the next k tile waits in registers for a barrier. The
// 16-byte chunk c of row r sits at chunk c ^ f(r), f(r) = r & 7 for 8 chunks per
// row and (r >> 1) & 3 for 4. Measured on gfx950 with a probe of every XOR map of
// the row bits: 16-byte LDS stores are served 8 lanes (128 bytes) at a time, so
// the 8 rows of a store group need distinct chunks, and of the maps that meet
// that, these two also read an MFMA fragment (16 rows, one chunk per 16 lanes)
// at full rate; (r >> 1) & 7, used before, read at full rate but conflicted
// 2-way on the stores (and (r >> 2) & 3 on both).
// Complex B: LDS row r (< 64) holds (re, -im) of compact row r and produces
// output column 2r (real part); row 64 + r holds (im, re) and produces 2r + 1.
// BRAW: B is not pre-split; the FP32 (or complex64) B is read through the gather
// tables and split while staging, so a B element costs one HBM read per 128-row
// tile of A instead of a separate read-split-write pass plus the 16-bit reads.
template <int BK>
__device__ inline int swz(int row, int ch) {
  constexpr int CH = BK / 8;  // 16-byte chunks per row
  const int f = CH == 8 ? (row & 7) : ((row >> 1) & 3);
  return row * BK + (((ch >> 3) ^ f) << 3);
}

// BMODE: 0 B pre-split, 1 raw B with lanes along columns (a chunk is BE rows of
// one column, wave-uniform rows: columns contiguous in memory), 2 raw B with
// lanes along k (chunk q = column q / CH, k chunk q % CH, as for pre-split B: for
// operands whose contracted legs are innermost in memory, GemmArgs::brow_run).
// (Loading the next tile's first k tile during the last k tile's MFMAs, with up
// to four tiles per workgroup, measured the same on the circuit's steps: the
// two workgroups per CU already overlap one's prologue with the other's MFMAs.)
// BM: rows of A per tile. 128: four 64x64 wave tiles (two workgroups per CU);
// 32: four 32x32 wave tiles side by side (three workgroups per CU) for m <= 32,
// where B streams from HBM once and the MFMA work per B byte is a quarter of
// the 128-row tile's. GATHER: raw B read through the gather tables.
// Every operand load is a buffer load at a per-thread offset fixed for the
// whole kernel (A, pre-split B: the k tile goes into the scalar offset) or
// one add away from it (raw B: the row's byte offset from an LDS table filled
// once per workgroup, plus the thread's column offset, fixed per tile).
// Layout of the dynamic LDS: ahi, alo [BM][BK]; bhi, blo [128][BK]; raw B: the
// row table, kp / (CPLX ? 2 : 1) words (presplit_lds).
template <bool CPLX, bool F16, int BK, int BMODE, int BM, bool GATHER>
__global__ __launch_bounds__(256, BM == 32 ? 3 : 2) void gemm_presplit_kernel(GemmArgs g, PreSplit ps) {
  constexpr bool BRAW = BMODE != 0;
  using H = typename Half<F16>::H;
  using V = typename Half<F16>::V;
  static_assert(BK == 64, "k tile (kp is a multiple of 64 halves)");
  static_assert(BM == 128 || BM == 32, "row tile");
  static_assert(!GATHER || BRAW, "pre-split B has no tables");
  constexpr int BN = 128, CH = BK / 8;
  constexpr int WM = BM == 128 ? 64 : 32, WN = WM, TM = WM / 16, TN = WN / 16;
  constexpr int A_Q = BM * CH / 256;                    // 16-byte A chunks per thread
  constexpr int B_Q = (CPLX ? BN / 2 : BN) * CH / 256;  // 16-byte B chunks per thread
  constexpr int TILE_A = BM * BK, TILE_B = BN * BK;     // halves per LDS array
  // raw B with lanes along columns: chunk q of a k tile is (tile column q % BC,
  // k group q / BC), so the lanes of a wave read consecutive columns of one B
  // row; a chunk is BE rows
  constexpr int BC = CPLX ? BN / 2 : BN, BE = CPLX ? 4 : 8;
  static_assert(256 % BC == 0, "a thread's raw B loads share one column");
  using BT = typename std::conditional<CPLX, float2, float>::type;
  extern __shared__ __align__(16) char smem_split[];
  H* const ahi = reinterpret_cast<H*>(smem_split);
  H* const alo = ahi + TILE_A;
  H* const bhi = alo + TILE_A;
  H* const blo = bhi + TILE_B;
  uint32_t* const tab = reinterpret_cast<uint32_t*>(blo + TILE_B);

  const int pe = pending_exp(g);
  float unscale = ldexpf(1.f, -pe), sb = 1.f;
  float mx = 0.f;
  if constexpr (F16) {
    unscale = ldexpf(1.f, -(pow2_shift(*g.amax) + pow2_shift(*g.bmax) + pe));
    sb = ldexpf(1.f, pow2_shift(*g.bmax));
  }
  const int64_t M = g.m, N = CPLX ? 2 * g.n : g.n, kp = ps.kp;
  // k range of this workgroup (split k) and where its product goes
  const int64_t kb = ps.kc ? (int64_t)blockIdx.z * ps.kc : 0;
  const int64_t ke = ps.kc && kb + ps.kc < kp ? kb + ps.kc : kp;
  const bool fold = g.cmax && !ps.part;  // partial products: the reduce kernel folds the max
  const int64_t ldc = ps.part ? N : CPLX ? 2 * g.ldc : g.ldc;
  const int tiles_n = (int)((N + BN - 1) / BN), tiles_m = (int)((M + BM - 1) / BM);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm0 = BM == 128 ? (wave >> 1) * WM : 0, wn0 = BM == 128 ? (wave & 1) * WN : wave * WN;
  const uint64_t nz = (uint64_t)g.nbo * g.nbi, ntile = (uint64_t)tiles_m * tiles_n;
  const int64_t ainner = g.sai ? g.nbi : 1, binner = g.sbi ? g.nbi : 1;
  if constexpr (BRAW) {  // byte offset of every B row (the same in every tile and batch entry)
    const int64_t kr = kp / (CPLX ? 2 : 1);
    for (int64_t r = tid; r < kr; r += 256)
      tab[r] = r < g.k ? (uint32_t)((GATHER ? (int64_t)g.brow[r] : r * g.ldb) * (int64_t)sizeof(BT)) : ROW_OOB;
    __syncthreads();
  }
  // per-thread byte offsets of the 16-byte chunks of a k tile within a row tile
  // ([rows][kp] halves); the k tile's offset goes into the scalar offset
  uint32_t a_off[A_Q], b_off[BRAW ? 1 : B_Q];
#pragma unroll
  for (int u = 0; u < A_Q; ++u) {
    const int q = tid + 256 * u;
    a_off[u] = (uint32_t)(((q / CH) * kp + (q % CH) * 8) * 2);
  }
  if constexpr (!BRAW) {
#pragma unroll
    for (int u = 0; u < B_Q; ++u) {
      const int q = tid + 256 * u;
      b_off[u] = (uint32_t)(((q / CH) * kp + (q % CH) * 8) * 2);
    }
  }

  for (uint64_t z = blockIdx.y; z < nz; z += gridDim.y) {
    const int64_t zo = (int64_t)(z / g.nbi), zi = (int64_t)(z - zo * g.nbi);
    const int64_t amat = (g.sao ? zo : 0) * ainner + (g.sai ? zi : 0);
    const int64_t bmat = (g.sbo ? zo : 0) * binner + (g.sbi ? zi : 0);
    const H* Ah = (const H*)ps.ahi + amat * M * kp;
    const H* Al = (const H*)ps.alo + amat * M * kp;
    const H* Bh = BRAW ? nullptr : (const H*)ps.bhi + bmat * g.n * kp;
    const H* Bl = BRAW ? nullptr : (const H*)ps.blo + bmat * g.n * kp;
    const BT* Braw = (const BT*)g.b + zo * g.sbo + zi * g.sbi;
    float* C = ps.part ? (float*)ps.part + (((int64_t)blockIdx.z * (int64_t)nz + (int64_t)z) * M * N)
                       : (float*)g.c + (zo * g.sco + zi * g.sci) * (CPLX ? 2 : 1);
    const auto rb = buf_rsrc(Braw, RAW_LIMIT);
    for (uint64_t t = blockIdx.x; t < ntile; t += gridDim.x) {
      int64_t m0, n0;
      tile_origin(t, ntile, tiles_m, tiles_n, BM, BN, m0, n0);
      const int64_t nc0 = CPLX ? n0 / 2 : n0;  // first compact column of B
      v4f acc[TM][TN];
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) acc[i][j] = v4f{0, 0, 0, 0};
      // this tile's rows of A and of pre-split B: rows past the end read zeros
      const uint32_t a_bytes = tile_bytes(M - m0, kp * 2);
      const auto rah_ = buf_rsrc(Ah + m0 * kp, a_bytes), ral_ = buf_rsrc(Al + m0 * kp, a_bytes);
      const uint32_t b_bytes = BRAW ? 0u : tile_bytes(g.n - nc0, kp * 2);
      const auto rbh_ = buf_rsrc(BRAW ? (const void*)Braw : (const void*)(Bh + nc0 * kp), b_bytes);
      const auto rbl_ = buf_rsrc(BRAW ? (const void*)Braw : (const void*)(Bl + nc0 * kp), b_bytes);
      // raw B: the byte offset of the thread's column (lanes along columns), or of
      // the column of each chunk (lanes along k)
      uint32_t bco = 0, kco[BMODE == 2 ? B_Q : 1];
      if constexpr (BMODE == 1) {
        const int64_t j = nc0 + tid % BC;
        bco = j < g.n ? (uint32_t)((GATHER ? (int64_t)g.bcol[j] : j) * (int64_t)sizeof(BT)) : COL_OOB;
      }
      if constexpr (BMODE == 2) {
#pragma unroll
        for (int u = 0; u < B_Q; ++u) {
          const int64_t j = nc0 + (tid + 256 * u) / CH;
          kco[u] = j < g.n ? (uint32_t)((int64_t)g.bcol[j] * (int64_t)sizeof(BT)) : COL_OOB;
        }
      }
      uint4 rah[A_Q], ral[A_Q], rbh[BRAW ? 1 : B_Q], rbl[BRAW ? 1 : B_Q];
      BT rbr[BRAW ? B_Q * BE : 1];
      auto load = [&](int64_t k0) {
        const int ks = (int)(k0 * 2);
#pragma unroll
        for (int u = 0; u < A_Q; ++u) {
          rah[u] = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(rah_, a_off[u], ks, 0));
          ral[u] = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(ral_, a_off[u], ks, 0));
        }
        if constexpr (BMODE == 0) {
#pragma unroll
          for (int u = 0; u < B_Q; ++u) {
            rbh[u] = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(rbh_, b_off[u], ks, 0));
            rbl[u] = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(rbl_, b_off[u], ks, 0));
          }
        } else if constexpr (BMODE == 1) {
#pragma unroll
          for (int u = 0; u < B_Q; ++u) {
            // first B row of the chunk (complex rows for complex B); the same on
            // every lane of a wave (BC is a multiple of the wave size)
            const int kc = (tid + 256 * u) / BC;
            const int r0 = (int)((k0 + kc * 8) / (CPLX ? 2 : 1));
            uint32_t ro[BE];
            *reinterpret_cast<uint4*>(ro) = *reinterpret_cast<const uint4*>(tab + r0);
            if constexpr (BE == 8) *reinterpret_cast<uint4*>(ro + 4) = *reinterpret_cast<const uint4*>(tab + r0 + 4);
#pragma unroll
            for (int e = 0; e < BE; ++e) {
              if constexpr (CPLX)
                rbr[u * BE + e] = __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(rb, ro[e] + bco, 0, 0));
              else
                rbr[u * BE + e] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rb, ro[e] + bco, 0, 0));
            }
          }
        } else {
#pragma unroll
          for (int u = 0; u < B_Q; ++u) {
            // BE consecutive k of one column, 32 contiguous bytes (brow_run and k
            // are multiples of BE, chunks start at multiples of BE)
            const int r0 = (int)((k0 + ((tid + 256 * u) % CH) * 8) / (CPLX ? 2 : 1));
            const uint32_t o = tab[r0] + kco[u];
            uint4* dst = reinterpret_cast<uint4*>(&rbr[u * BE]);
            dst[0] = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(rb, o, 0, 0));
            dst[1] = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(rb, o + 16, 0, 0));
          }
        }
      };
      auto store = [&]() {
#pragma unroll
        for (int u = 0; u < A_Q; ++u) {
          const int q = tid + 256 * u, row = q / CH, ch = (q % CH) * 8;
          *reinterpret_cast<uint4*>(ahi + swz<BK>(row, ch)) = rah[u];
          *reinterpret_cast<uint4*>(alo + swz<BK>(row, ch)) = ral[u];
        }
#pragma unroll
        for (int u = 0; u < B_Q; ++u) {
          const int q = tid + 256 * u;
          if constexpr (BRAW) {
            int row, ch;
            if constexpr (BMODE == 2) row = q / CH, ch = (q % CH) * 8;
            else row = q % BC, ch = (q / BC) * 8;
            uint32_t h[4], l[4];
            if constexpr (CPLX) {
              // (re, im) split once; the rows (re, -im) and (im, re) are its halves with
              // the high one negated, and swapped (rounding commutes with negation)
              uint32_t h2[4], l2[4];
#pragma unroll
              for (int e = 0; e < 4; ++e) {
                uint32_t hh, ll;
                split_pair<F16>(v2f{rbr[u * 4 + e].x * sb, rbr[u * 4 + e].y * sb}, hh, ll);
                h[e] = hh ^ 0x80000000u, l[e] = ll ^ 0x80000000u;
                h2[e] = swap16(hh), l2[e] = swap16(ll);
              }
              *reinterpret_cast<uint4*>(bhi + swz<BK>(row, ch)) = *reinterpret_cast<const uint4*>(h);
              *reinterpret_cast<uint4*>(blo + swz<BK>(row, ch)) = *reinterpret_cast<const uint4*>(l);
              *reinterpret_cast<uint4*>(bhi + swz<BK>(row + 64, ch)) = *reinterpret_cast<const uint4*>(h2);
              *reinterpret_cast<uint4*>(blo + swz<BK>(row + 64, ch)) = *reinterpret_cast<const uint4*>(l2);
            } else {
#pragma unroll
              for (int e = 0; e < 4; ++e)
                split_pair<F16>(v2f{rbr[u * 8 + 2 * e] * sb, rbr[u * 8 + 2 * e + 1] * sb}, h[e], l[e]);
              *reinterpret_cast<uint4*>(bhi + swz<BK>(row, ch)) = *reinterpret_cast<const uint4*>(h);
              *reinterpret_cast<uint4*>(blo + swz<BK>(row, ch)) = *reinterpret_cast<const uint4*>(l);
            }
          } else {
            const int row = q / CH, ch = (q % CH) * 8;
            if constexpr (CPLX) {
              *reinterpret_cast<uint4*>(bhi + swz<BK>(row, ch)) = neg_im(rbh[u]);
              *reinterpret_cast<uint4*>(blo + swz<BK>(row, ch)) = neg_im(rbl[u]);
              *reinterpret_cast<uint4*>(bhi + swz<BK>(row + 64, ch)) = swap_pairs(rbh[u]);
              *reinterpret_cast<uint4*>(blo + swz<BK>(row + 64, ch)) = swap_pairs(rbl[u]);
            } else {
              *reinterpret_cast<uint4*>(bhi + swz<BK>(row, ch)) = rbh[u];
              *reinterpret_cast<uint4*>(blo + swz<BK>(row, ch)) = rbl[u];
            }
          }
        }
      };
      const int64_t nk = (ke - kb) / BK;
      const int cl = lane & 15;
      load(kb);
      store();
      __syncthreads();
      for (int64_t kt = 0; kt < nk; ++kt) {
        if (kt + 1 < nk) load(kb + (kt + 1) * BK);
#pragma unroll
        for (int ks = 0; ks < BK; ks += 32) {
          const int kg = ks + (lane >> 4) * 8;
          V ah[TM], al[TM];
#pragma unroll
          for (int i = 0; i < TM; ++i) {
            ah[i] = *reinterpret_cast<const V*>(&ahi[swz<BK>(wm0 + i * 16 + cl, kg)]);
            al[i] = *reinterpret_cast<const V*>(&alo[swz<BK>(wm0 + i * 16 + cl, kg)]);
          }
          V bh[TN], bl[TN];
#pragma unroll
          for (int j = 0; j < TN; ++j) {
            bh[j] = *reinterpret_cast<const V*>(&bhi[swz<BK>(wn0 + j * 16 + cl, kg)]);
            bl[j] = *reinterpret_cast<const V*>(&blo[swz<BK>(wn0 + j * 16 + cl, kg)]);
          }
          // the three products of one accumulator are TM*TN instructions apart,
          // so consecutive MFMAs never wait on each other's result. A wave in its
          // MFMA run issues at raised priority, so the SIMD's other wave (staging,
          // waiting on memory) yields the issue slots to it (gfx950, complex64
          // x64: m=1024 k=512 n=8192 7.44 -> 7.24 ms, m=512 k=512 n=8192 4.04 -> 3.95)
          __builtin_amdgcn_s_setprio(1);
#pragma unroll
          for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int j = 0; j < TN; ++j) acc[i][j] = Half<F16>::mfma(al[i], bh[j], acc[i][j]);
#pragma unroll
          for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int j = 0; j < TN; ++j) acc[i][j] = Half<F16>::mfma(ah[i], bl[j], acc[i][j]);
#pragma unroll
          for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int j = 0; j < TN; ++j) acc[i][j] = Half<F16>::mfma(ah[i], bh[j], acc[i][j]);
        }
        __builtin_amdgcn_s_setprio(0);
        __syncthreads();
        if (kt + 1 < nk) {
          store();
          __syncthreads();
        }
      }
      // Epilogue through LDS (free after the loop's last barrier): the tile goes
      // to a row-major [BM][BN] float array, then every thread writes 16-byte row
      // pieces, so whole cache lines reach HBM (per-lane 4-byte stores at an 8-byte
      // stride wrote complex64 results at about half the device's write rate).
      // The column index is XOR-swizzled by row so that the four 16-lane groups
      // of one store, 4 rows apart, hit distinct banks.
      float* const cs = reinterpret_cast<float*>(smem_split);
      auto cswz = [](int tr) {
        return CPLX ? ((tr >> 2) & 1) | (((tr >> 3) & 1) << 5) : ((tr >> 2) & 3) << 4;
      };
      {
        // rows wm0 + 16 i + 4 (lane / 16) + r: their swizzle depends on the lane only
        const int g4 = lane >> 4;
        const int xr = CPLX ? (g4 & 1) | ((g4 >> 1) << 5) : g4 << 4;
        float* const crow = cs + (wm0 + 4 * g4) * BN;
#pragma unroll
        for (int j = 0; j < TN; ++j) {
          const int lrow = wn0 + j * 16 + cl;  // LDS row of B' = tile column
          const int tc = (CPLX ? 2 * (lrow & 63) + (lrow >> 6) : lrow) ^ xr;
#pragma unroll
          for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int r = 0; r < 4; ++r) crow[(i * 16 + r) * BN + tc] = acc[i][j][r];
        }
      }
      __syncthreads();
      const bool c16 = ((uintptr_t)C & 15) == 0 && (ldc & 3) == 0;
      constexpr int C_Q = BM * BN / 4 / 256;  // 16-byte pieces per thread
#pragma unroll 2
      for (int u = 0; u < C_Q; ++u) {
        const int q = tid + 256 * u, tr = q / (BN / 4), c4 = (q % (BN / 4)) * 4;
        const int sw = cswz(tr);
        float4 v = *reinterpret_cast<const float4*>(&cs[tr * BN + (c4 ^ (sw & ~3))]);
        if (sw & 1) v = make_float4(v.y, v.x, v.w, v.z);
        const int64_t row = m0 + tr, col = n0 + c4;
        if (row < M) {
          v.x *= unscale, v.y *= unscale, v.z *= unscale, v.w *= unscale;
          if (c16 && col + 3 < N) {
            *reinterpret_cast<float4*>(C + row * ldc + col) = v;
            if (fold) fold_max(mx, v.x), fold_max(mx, v.y), fold_max(mx, v.z), fold_max(mx, v.w);
          } else {
            const float e4[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int e = 0; e < 4; ++e)
              if (col + e < N) {
                C[row * ldc + col + e] = e4[e];
                if (fold) fold_max(mx, e4[e]);
              }
          }
        }
      }
      __syncthreads();  // the next tile stages into the same LDS
    }
  }
  if (fold) publish_max((double)mx, g.cmax);
}

// ------------------------------------------- 3M split-precision GEMM --
// complex64 C = A.B as Gauss's three real products on the 16-bit MFMA:
//   T1 = Ar.Br, T2 = Ai.Bi, T3 = (Ar + Ai).(Br + Bi);  Cr = T1 - T2, Ci = T3 - T1 - T2
// each real product as lo.hi + hi.lo + hi.hi of 16-bit halves accumulated in
// FP32: 9 MFMA products per block of complex multiply-adds where the 2x2
// expansion of gemm_presplit_kernel issues 12. A comes split by
// split_a3m_kernel into three planes (Ar, Ai, Ar + Ai) of hi and lo halves,
// [za][3][m][kq] (kq = k rounded up to 32, zero padded); B is read in complex64,
// row-major or through the gather tables, lanes along columns (a thread loads 8
// consecutive complex rows of one column), and split into its three planes while
// staging. FP16 halves use per-operand scales 2^(s-1), one below the 2x2 form's,
// since the sum planes reach twice the operand max.
// Tile: 128 rows x 64 complex columns x 32 complex k (one MFMA k step); 4 waves
// in a 2x2 arrangement, 64 rows x 32 complex columns each, three accumulator
// planes. LDS: A planes [3][128][32] hi and lo, B planes [3][64][32] (B
// transposed) hi and lo, rows of 32 halves whose 16-byte chunk c sits at
// c ^ ((row >> 1) & 3), then the row table of raw B (kq words); the epilogue
// reuses the first 64 KiB as a [128][128] float tile.
constexpr int S3M_BK = 32;
constexpr size_t s3m_lds(int64_t kq) { return (size_t)(3 * 128 + 3 * 64) * S3M_BK * 2 * 2 + (size_t)kq * 4; }
__device__ inline int swz32(int row, int ch) { return row * 32 + ((ch ^ ((row >> 1) & 3)) << 3); }

// A (complex64 [m][k]) -> planes [za][3][m][kq] hi / lo; one thread per 4
// consecutive k of a row (all three planes).
template <bool F16>
__global__ __launch_bounds__(256) void split_a3m_kernel(GemmArgs g, PreSplit ps) {
  using H = typename Half<F16>::H;
  const int64_t kq = ps.kp, inner = g.sai ? g.nbi : 1, nm = n_mats(g.sao, g.sai, g.nbo, g.nbi);
  const int64_t per_row = kq / 4, total = nm * g.m * per_row;
  float sa = 1.f;
  if constexpr (F16) sa = ldexpf(1.f, pow2_shift(*g.amax) - 1);
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t mat = e / (g.m * per_row), r = e - mat * g.m * per_row, row = r / per_row, c = (r - row * per_row) * 4;
    const int64_t zo = mat / inner, zi = mat - zo * inner;
    const float2* base = (const float2*)g.a + zo * g.sao + zi * g.sai;
    H h[3][4], l[3][4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      // A through gather tables (one batch entry): element (row, p) at arow[row] + acol[p]
      const float2 x = c + u >= g.k ? make_float2(0.f, 0.f)
                       : g.acol     ? base[(int64_t)g.arow[row] + (int64_t)g.acol[c + u]]
                                    : base[row * g.lda + c + u];
      const float re = x.x * sa, im = x.y * sa;
      split16<F16>(re, h[0][u], l[0][u]);
      split16<F16>(im, h[1][u], l[1][u]);
      split16<F16>(re + im, h[2][u], l[2][u]);
    }
#pragma unroll
    for (int p = 0; p < 3; ++p) {
      const int64_t off = ((mat * 3 + p) * g.m + row) * kq + c;
      *reinterpret_cast<uint2*>((H*)ps.ahi + off) = *reinterpret_cast<const uint2*>(h[p]);
      *reinterpret_cast<uint2*>((H*)ps.alo + off) = *reinterpret_cast<const uint2*>(l[p]);
    }
  }
}

template <bool F16, bool GATHER>
__global__ __launch_bounds__(256, 2) void gemm_split3m_kernel(GemmArgs g, PreSplit ps) {
  using H = typename Half<F16>::H;
  using V = typename Half<F16>::V;
  constexpr int BM = 128, BNC = 64, BK = S3M_BK, CH = BK / 8;
  constexpr int TM = 4, TN = 2;                     // 16x16 blocks of a wave tile
  constexpr int PA = BM * BK, PB = BNC * BK;        // halves per plane
  constexpr int A_Q = BM * CH / 256;                // 16-byte chunks of one A plane array per thread (2)
  extern __shared__ __align__(16) char smem_split[];
  H* const ahi = reinterpret_cast<H*>(smem_split);  // [3][BM][BK]
  H* const alo = ahi + 3 * PA;
  H* const bhi = alo + 3 * PA;                      // [3][BNC][BK]
  H* const blo = bhi + 3 * PB;
  uint32_t* const tab = reinterpret_cast<uint32_t*>(blo + 3 * PB);

  const int pe = pending_exp(g);
  float unscale = ldexpf(1.f, -pe), sb = 1.f;
  float mx = 0.f;
  if constexpr (F16) {
    const int ea = pow2_shift(*g.amax) - 1, eb = pow2_shift(*g.bmax) - 1;
    unscale = ldexpf(1.f, -(ea + eb + pe));
    sb = ldexpf(1.f, eb);
  }
  const int64_t M = g.m, N = 2 * g.n, kq = ps.kp, ldc = 2 * g.ldc;
  const int tiles_n = (int)((g.n + BNC - 1) / BNC), tiles_m = (int)((M + BM - 1) / BM);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, cl = lane & 15, g4 = lane >> 4;
  const int wm0 = (wave >> 1) * 64, wn0 = (wave & 1) * 32;
  const uint64_t nz = (uint64_t)g.nbo * g.nbi, ntile = (uint64_t)tiles_m * tiles_n;
  const int64_t ainner = g.sai ? g.nbi : 1;
  // byte offset of every B row (the same in every tile and batch entry)
  for (int64_t r = tid; r < kq; r += 256)
    tab[r] = r < g.k ? (uint32_t)((GATHER ? (int64_t)g.brow[r] : r * g.ldb) * 8) : ROW_OOB;
  __syncthreads();
  // per-thread byte offsets of the A chunks within a plane's row tile; B: the
  // thread's column (tid % 64) and group of 8 complex rows (tid / 64, one per wave)
  uint32_t a_off[A_Q];
#pragma unroll
  for (int u = 0; u < A_Q; ++u) {
    const int q = tid + 256 * u;
    a_off[u] = (uint32_t)(((q / CH) * kq + (q % CH) * 8) * 2);
  }
  const int bj = tid % BNC, bgrp = tid / BNC;

  for (uint64_t z = blockIdx.y; z < nz; z += gridDim.y) {
    const int64_t zo = (int64_t)(z / g.nbi), zi = (int64_t)(z - zo * g.nbi);
    const int64_t amat = (g.sao ? zo : 0) * ainner + (g.sai ? zi : 0);
    const float2* Braw = (const float2*)g.b + zo * g.sbo + zi * g.sbi;
    float* C = (float*)g.c + (zo * g.sco + zi * g.sci) * 2;
    const auto rb = buf_rsrc(Braw, RAW_LIMIT);
    for (uint64_t t = blockIdx.x; t < ntile; t += gridDim.x) {
      int64_t m0, nc0;
      tile_origin(t, ntile, tiles_m, tiles_n, BM, BNC, m0, nc0);
      v4f acc[3][TM][TN];
#pragma unroll
      for (int p = 0; p < 3; ++p)
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
          for (int j = 0; j < TN; ++j) acc[p][i][j] = v4f{0, 0, 0, 0};
      // the three A planes of this batch entry as one resource each for hi and lo;
      // plane p's rows of this tile start at scalar byte offset p * plane + tile
      // (rows past m of a plane read the next plane's rows, which only reach
      // result rows past m, never stored; reads past the last plane read zeros)
      const uint32_t plane = (uint32_t)(M * kq * 2), tile = (uint32_t)(m0 * kq * 2);
      const auto rah = buf_rsrc((const H*)ps.ahi + amat * 3 * M * kq, 3 * plane);
      const auto ral = buf_rsrc((const H*)ps.alo + amat * 3 * M * kq, 3 * plane);
      const int64_t jcol = nc0 + bj;
      const uint32_t bco = jcol < g.n ? (uint32_t)((GATHER ? (int64_t)g.bcol[jcol] : jcol) * 8) : COL_OOB;
      uint4 ra_h[3][A_Q], ra_l[3][A_Q];
      float2 rbr[8];
      auto load = [&](int64_t k0) {
#pragma unroll
        for (int p = 0; p < 3; ++p) {
          const int so = (int)(p * plane + tile + (uint32_t)(k0 * 2));
#pragma unroll
          for (int u = 0; u < A_Q; ++u) {
            ra_h[p][u] = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(rah, a_off[u], so, 0));
            ra_l[p][u] = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(ral, a_off[u], so, 0));
          }
        }
        uint32_t ro[8];
        *reinterpret_cast<uint4*>(ro) = *reinterpret_cast<const uint4*>(tab + k0 + 8 * bgrp);
        *reinterpret_cast<uint4*>(ro + 4) = *reinterpret_cast<const uint4*>(tab + k0 + 8 * bgrp + 4);
#pragma unroll
        for (int e = 0; e < 8; ++e)
          rbr[e] = __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(rb, ro[e] + bco, 0, 0));
      };
      auto store = [&]() {
#pragma unroll
        for (int p = 0; p < 3; ++p)
#pragma unroll
          for (int u = 0; u < A_Q; ++u) {
            const int q = tid + 256 * u;
            *reinterpret_cast<uint4*>(ahi + p * PA + swz32(q / CH, q % CH)) = ra_h[p][u];
            *reinterpret_cast<uint4*>(alo + p * PA + swz32(q / CH, q % CH)) = ra_l[p][u];
          }
        // B planes: re, im, re + im of 8 consecutive k of column bj
        uint32_t h[3][4], l[3][4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const float r0 = rbr[2 * e].x * sb, i0 = rbr[2 * e].y * sb;
          const float r1 = rbr[2 * e + 1].x * sb, i1 = rbr[2 * e + 1].y * sb;
          split_pair<F16>(v2f{r0, r1}, h[0][e], l[0][e]);
          split_pair<F16>(v2f{i0, i1}, h[1][e], l[1][e]);
          split_pair<F16>(v2f{r0 + i0, r1 + i1}, h[2][e], l[2][e]);
        }
        const int o = swz32(bj, bgrp);
#pragma unroll
        for (int p = 0; p < 3; ++p) {
          *reinterpret_cast<uint4*>(bhi + p * PB + o) = *reinterpret_cast<const uint4*>(h[p]);
          *reinterpret_cast<uint4*>(blo + p * PB + o) = *reinterpret_cast<const uint4*>(l[p]);
        }
      };
      const int64_t nk = kq / BK;
      load(0);
      store();
      __syncthreads();
      for (int64_t kt = 0; kt < nk; ++kt) {
        if (kt + 1 < nk) load((kt + 1) * BK);
        __builtin_amdgcn_s_setprio(1);  // as in gemm_presplit_kernel
#pragma unroll
        for (int p = 0; p < 3; ++p) {
          V ah[TM], al[TM], bh[TN], bl[TN];
#pragma unroll
          for (int i = 0; i < TM; ++i) {
            ah[i] = *reinterpret_cast<const V*>(&ahi[p * PA + swz32(wm0 + i * 16 + cl, g4)]);
            al[i] = *reinterpret_cast<const V*>(&alo[p * PA + swz32(wm0 + i * 16 + cl, g4)]);
          }
#pragma unroll
          for (int j = 0; j < TN; ++j) {
            bh[j] = *reinterpret_cast<const V*>(&bhi[p * PB + swz32(wn0 + j * 16 + cl, g4)]);
            bl[j] = *reinterpret_cast<const V*>(&blo[p * PB + swz32(wn0 + j * 16 + cl, g4)]);
          }
          // the three products of one accumulator are TM * TN instructions apart
#pragma unroll
          for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int j = 0; j < TN; ++j) acc[p][i][j] = Half<F16>::mfma(al[i], bh[j], acc[p][i][j]);
#pragma unroll
          for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int j = 0; j < TN; ++j) acc[p][i][j] = Half<F16>::mfma(ah[i], bl[j], acc[p][i][j]);
#pragma unroll
          for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int j = 0; j < TN; ++j) acc[p][i][j] = Half<F16>::mfma(ah[i], bh[j], acc[p][i][j]);
        }
        __builtin_amdgcn_s_setprio(0);
        __syncthreads();
        if (kt + 1 < nk) {
          store();
          __syncthreads();
        }
      }
      // Epilogue through LDS (free after the loop's last barrier): (Cr, Ci) of
      // row r, complex column c at float index r * 128 + 2 * (c ^ 16 * ((r >> 2) & 1)),
      // so the two 16-lane groups of a 32-lane store half hit distinct banks;
      // then 16-byte row pieces to C.
      float* const cs = reinterpret_cast<float*>(smem_split);
      {
        const int xr = (g4 & 1) << 4;
        float* const crow = cs + (wm0 + 4 * g4) * 128;
#pragma unroll
        for (int j = 0; j < TN; ++j) {
          const int pc = 2 * ((wn0 + j * 16 + cl) ^ xr);
#pragma unroll
          for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              const float t1 = acc[0][i][j][r], t2 = acc[1][i][j][r], t3 = acc[2][i][j][r];
              *reinterpret_cast<float2*>(&crow[(i * 16 + r) * 128 + pc]) = make_float2(t1 - t2, t3 - t1 - t2);
            }
        }
      }
      __syncthreads();
      const bool c16 = ((uintptr_t)C & 15) == 0 && (ldc & 3) == 0;
      const int64_t n0 = 2 * nc0;  // first float column of the tile
      constexpr int C_Q = BM * 128 / 4 / 256;  // 16-byte pieces per thread
#pragma unroll 2
      for (int u = 0; u < C_Q; ++u) {
        const int q = tid + 256 * u, tr = q / 32, c4 = (q % 32) * 4;
        const int pc = 2 * ((c4 / 2) ^ (((tr >> 2) & 1) << 4));
        float4 v = *reinterpret_cast<const float4*>(&cs[tr * 128 + pc]);
        const int64_t row = m0 + tr, col = n0 + c4;
        if (row < M) {
          v.x *= unscale, v.y *= unscale, v.z *= unscale, v.w *= unscale;
          if (c16 && col + 3 < N) {
            *reinterpret_cast<float4*>(C + row * ldc + col) = v;
            if (g.cmax) fold_max(mx, v.x), fold_max(mx, v.y), fold_max(mx, v.z), fold_max(mx, v.w);
          } else {
            const float e4[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int e = 0; e < 4; ++e)
              if (col + e < N) {
                C[row * ldc + col + e] = e4[e];
                if (g.cmax) fold_max(mx, e4[e]);
              }
          }
        }
      }
      __syncthreads();  // the next tile stages into the same LDS
    }
  }
  if (g.cmax) publish_max((double)mx, g.cmax);
}

template <class F>
static void dispatch_dt(int dt, F&& f) {
  switch (dt) {
    case F32: f(float{}); break;
    case F64: f(double{}); break;
    case C64: f(float2{}); break;
    case C128: f(double2{}); break;
    default: throw Error("unsupported dtype code");
  }
}

static unsigned zgrid(const GemmArgs& g) { return (unsigned)std::min<int64_t>(g.nbo * g.nbi, 65535); }

static void run_simple(int dt, const GemmArgs& g, hipStream_t s, int opa = 0, int opb = 0) {
  if (g.bcg && g.k > 0) throw Error("the simple GEMM kernel does not read a factored B operand");
  dispatch_dt(dt, [&](auto tag) {
    using T = decltype(tag);
    uint64_t mn = (uint64_t)g.m * g.n;
    unsigned gx = (unsigned)std::min<uint64_t>((mn + 255) / 256, 65536);
    uint64_t nz = (uint64_t)g.nbo * g.nbi;
    // keep the total grid bounded for huge batches of tiny problems
    dim3 grid(gx, zgrid(g));
    if (nz * gx > (1u << 22)) grid.y = std::max<unsigned>(1, (1u << 22) / gx);
    g_last_kind = K_GEMM_SIMPLE;
    if (opa == 0 && opb == 0) LAUNCH((gemm_simple_kernel<T, 0, 0>), grid, dim3(256), 0, s, g);
    else if (opa == 1 && opb == 0) LAUNCH((gemm_simple_kernel<T, 1, 0>), grid, dim3(256), 0, s, g);
    else if (opa == 0 && opb == 1) LAUNCH((gemm_simple_kernel<T, 0, 1>), grid, dim3(256), 0, s, g);
    else LAUNCH((gemm_simple_kernel<T, 1, 1>), grid, dim3(256), 0, s, g);
  });
}

template <class F, bool CPLX, int BM, int BN, int BK>
static void run_mfma(int dt, const GemmArgs& g0, hipStream_t s) {
  static_assert(64 % BK == 0, "k ranges are multiples of 64");
  GemmArgs g = g0;
  const KSplit ks = plan_mfma_split(dt, NATIVE, g, scratch_avail(g), BM);
  g.kc = ks.n_split > 1 ? ks.k_chunk : 0;
  g.part = ks.n_split > 1 ? (char*)g.scratch + 256 : nullptr;
  int64_t N = CPLX ? 2 * g.n : g.n;
  uint64_t tiles = (uint64_t)((g.m + BM - 1) / BM) * (uint64_t)((N + BN - 1) / BN);
  dim3 grid((unsigned)std::min<uint64_t>(tiles, 1u << 20), zgrid(g), (unsigned)ks.n_split);
  // tracking the result max: a persistent grid (with several row tiles too: on
  // the uneven 3-regular workload, whose every GEMM folds its max, the full grid
  // ran 6% slower)
  if (g.cmax) grid.x = std::min<unsigned>(grid.x, std::max<unsigned>(1, TRACK_WORKGROUPS / (grid.y * grid.z)));
  g_last_kind = sizeof(F) == 4 ? K_GEMM_MFMA_F32 : K_GEMM_MFMA_F64;
  // with split k the reduce kernel tracks the result max
  const bool track = g.cmax && !g.part;
  if (g.bcg) {
    if constexpr (sizeof(F) == 4) {  // factored B: float32 / complex64 only
      if (track) LAUNCH((gemm_mfma_kernel<F, CPLX, BM, BN, BK, true, 2>), grid, dim3(256), 0, s, g);
      else LAUNCH((gemm_mfma_kernel<F, CPLX, BM, BN, BK, false, 2>), grid, dim3(256), 0, s, g);
    } else {
      throw Error("a factored B operand is read in float32 / complex64 only");
    }
  } else if (g.bcol) {
    if (track) LAUNCH((gemm_mfma_kernel<F, CPLX, BM, BN, BK, true, 1>), grid, dim3(256), 0, s, g);
    else LAUNCH((gemm_mfma_kernel<F, CPLX, BM, BN, BK, false, 1>), grid, dim3(256), 0, s, g);
  } else {
    if (track) LAUNCH((gemm_mfma_kernel<F, CPLX, BM, BN, BK, true, 0>), grid, dim3(256), 0, s, g);
    else LAUNCH((gemm_mfma_kernel<F, CPLX, BM, BN, BK, false, 0>), grid, dim3(256), 0, s, g);
  }
  if (ks.n_split > 1) {
    SplitK sk{};
    sk.n_split = ks.n_split;
    sk.part = g.part;
    const int64_t n_out = g.nbo * g.nbi * g.m * g.n;
    dim3 rg((unsigned)std::min<int64_t>((n_out + 255) / 256, 65536));
    dispatch_dt(dt, [&](auto tag) {
      using T = decltype(tag);
      LAUNCH((splitk_reduce_kernel<T>), rg, dim3(256), 0, s, g, sk);
    });
  }
}

template <int BM>
static void run_c3m_t(const GemmArgs& g0, hipStream_t s) {
  constexpr int BN = 64, BK = 16;
  GemmArgs g = g0;
  const KSplit ks = plan_mfma_split(C64, NATIVE, g, scratch_avail(g), BM);
  g.kc = ks.n_split > 1 ? ks.k_chunk : 0;
  g.part = ks.n_split > 1 ? (char*)g.scratch + 256 : nullptr;
  uint64_t tiles = (uint64_t)((g.m + BM - 1) / BM) * (uint64_t)((g.n + BN - 1) / BN);
  dim3 grid((unsigned)std::min<uint64_t>(tiles, 1u << 20), zgrid(g), (unsigned)ks.n_split);
  if (g.cmax) grid.x = std::min<unsigned>(grid.x, std::max<unsigned>(1, TRACK_WORKGROUPS / (grid.y * grid.z)));
  g_last_kind = K_GEMM_MFMA_C3M;
  const bool track = g.cmax && !g.part;
  if (g.bcol) {
    if (track) LAUNCH((gemm_c3m_kernel<BM, BN, BK, true, true>), grid, dim3(256), 0, s, g);
    else LAUNCH((gemm_c3m_kernel<BM, BN, BK, false, true>), grid, dim3(256), 0, s, g);
  } else {
    if (track) LAUNCH((gemm_c3m_kernel<BM, BN, BK, true, false>), grid, dim3(256), 0, s, g);
    else LAUNCH((gemm_c3m_kernel<BM, BN, BK, false, false>), grid, dim3(256), 0, s, g);
  }
  if (ks.n_split > 1) {
    SplitK sk{};
    sk.n_split = ks.n_split;
    sk.part = g.part;
    const int64_t n_out = g.nbo * g.nbi * g.m * g.n;
    dim3 rg((unsigned)std::min<int64_t>((n_out + 255) / 256, 65536));
    LAUNCH((splitk_reduce_kernel<float2>), rg, dim3(256), 0, s, g, sk);
  }
}

// Complex 3M on the FP32 MFMA: 64-row tiles up to m = 64, else 128.
static void run_c3m(const GemmArgs& g, hipStream_t s) {
  if (g.m <= 64) run_c3m_t<64>(g, s);
  else run_c3m_t<128>(g, s);
}

// The MFMA kernels zero-pad partial tiles; with 32-row tiles they pay off from
// m = 16 (k of 16 real units) or m = 32 (k of 8), measured against the
// streaming kernel on gfx950.
static bool mfma_worth(int dt, const GemmArgs& g) {
  bool cplx = dtype_complex(dt);
  int64_t K = cplx ? 2 * g.k : g.k, N = cplx ? 2 * g.n : g.n;
  return N >= 32 && ((g.m >= 32 && K >= 8) || (g.m >= 16 && K >= 16));
}

template <bool SMALL_M, int X>
static void run_skinny_t(int dt, const GemmArgs& g, hipStream_t s) {
  dispatch_dt(dt, [&](auto tag) {
    using T = decltype(tag);
    int64_t width = SMALL_M ? g.n : g.m;
    unsigned gx = (unsigned)std::min<int64_t>((width + 255) / 256, 65536);
    dim3 grid(gx, zgrid(g));
    if ((uint64_t)grid.y * gx > (1u << 22)) grid.y = std::max<unsigned>(1, (1u << 22) / gx);
    if constexpr (SMALL_M) LAUNCH((gemm_smallm_kernel<T, X>), grid, dim3(256), 0, s, g);
    else LAUNCH((gemm_smalln_kernel<T, X>), grid, dim3(256), 0, s, g);
  });
}

template <class T, int MM, int V>
static void run_stream_t(const GemmArgs& g, hipStream_t s) {
  const int64_t cols = (g.n + V - 1) / V;
  unsigned gx = (unsigned)std::min<int64_t>((cols + 255) / 256, 65536);
  if (g.cmax) gx = std::min<unsigned>(gx, std::max<unsigned>(1, TRACK_WORKGROUPS / zgrid(g)));
  dim3 grid(gx, zgrid(g));
  if ((uint64_t)grid.y * gx > (1u << 22)) grid.y = std::max<unsigned>(1, (1u << 22) / gx);
  if (g.bcg) LAUNCH((gemm_stream_kernel<T, MM, V, true>), grid, dim3(256), 0, s, g);
  else LAUNCH((gemm_stream_kernel<T, MM, V, false>), grid, dim3(256), 0, s, g);
}

// 16-byte column vectors when every row start of B and C is 16-byte aligned.
static int stream_width(int dt, const GemmArgs& g) {
  const int es = (int)dtype_size(dt);
  int v = 16 / es;
  auto ok = [&](int w) {
    const int64_t b = (int64_t)w * es;
    // gathered or factored columns: only within runs of consecutive offsets
    const bool cols = g.bcg ? g.bcs[0] == 1 && g.bcd[0] % w == 0 : g.bcol ? g.bcol_run % w == 0 : g.ldb % w == 0;
    return g.n % w == 0 && cols && g.ldc % w == 0 && g.sbo % w == 0 && g.sbi % w == 0 && g.sco % w == 0 &&
           g.sci % w == 0 && (uintptr_t)g.b % b == 0 && (uintptr_t)g.c % b == 0;
  };
  while (v > 1 && !ok(v)) v /= 2;
  return v;
}

// Small m streamed over n: m <= 64 (real) or m <= 32 (complex), accumulators
// per thread kept at 64 words or fewer.
static bool run_stream(int dt, const GemmArgs& g, hipStream_t s) {
  const bool cplx = dtype_complex(dt);
  if (g.n < 64 || g.m > (cplx ? 32 : 64)) return false;
  g_last_kind = K_GEMM_SMALL;
  const int w = stream_width(dt, g);
  const int mm = g.m <= 4 ? 4 : g.m <= 8 ? 8 : g.m <= 16 ? 16 : g.m <= 32 ? 32 : 64;
  switch (dt) {
    case F32:
      if (mm == 4) { if (w == 4) run_stream_t<float, 4, 4>(g, s); else run_stream_t<float, 4, 1>(g, s); }
      else if (mm == 8) { if (w == 4) run_stream_t<float, 8, 4>(g, s); else run_stream_t<float, 8, 1>(g, s); }
      else if (mm == 16) { if (w == 4) run_stream_t<float, 16, 4>(g, s); else run_stream_t<float, 16, 1>(g, s); }
      else if (mm == 32) { if (w >= 2) run_stream_t<float, 32, 2>(g, s); else run_stream_t<float, 32, 1>(g, s); }
      else run_stream_t<float, 64, 1>(g, s);
      break;
    case C64:
      if (mm == 4) { if (w == 2) run_stream_t<float2, 4, 2>(g, s); else run_stream_t<float2, 4, 1>(g, s); }
      else if (mm == 8) { if (w == 2) run_stream_t<float2, 8, 2>(g, s); else run_stream_t<float2, 8, 1>(g, s); }
      else if (mm == 16) { if (w == 2) run_stream_t<float2, 16, 2>(g, s); else run_stream_t<float2, 16, 1>(g, s); }
      else run_stream_t<float2, 32, 1>(g, s);
      break;
    case F64:
      if (mm == 4) { if (w == 2) run_stream_t<double, 4, 2>(g, s); else run_stream_t<double, 4, 1>(g, s); }
      else if (mm == 8) { if (w == 2) run_stream_t<double, 8, 2>(g, s); else run_stream_t<double, 8, 1>(g, s); }
      else if (mm == 16) { if (w == 2) run_stream_t<double, 16, 2>(g, s); else run_stream_t<double, 16, 1>(g, s); }
      else if (mm == 32) run_stream_t<double, 32, 1>(g, s);
      else run_stream_t<double, 64, 1>(g, s);
      break;
    case C128:
      if (mm == 4) run_stream_t<double2, 4, 1>(g, s);
      else if (mm == 8) run_stream_t<double2, 8, 1>(g, s);
      else if (mm == 16) run_stream_t<double2, 16, 1>(g, s);
      else run_stream_t<double2, 32, 1>(g, s);
      break;
    default: throw Error("unsupported dtype code");
  }
  return true;
}

// 0: not a skinny shape; 1: streaming kernel (folds the result max); 2: smallm / smalln kernel
static int run_skinny(int dt, const GemmArgs& g, hipStream_t s) {
  if (run_stream(dt, g, s)) return 1;
  if (g.bcg) throw Error("a factored B operand needs the streaming kernel (n >= 64, small m)");
  if (g.bcol) return 0;  // the smallm / smalln kernels read B row-major only
  if (g.m <= 32 && g.n >= 64) {
    g_last_kind = K_GEMM_SMALL;
    if (g.m <= 4) run_skinny_t<true, 4>(dt, g, s);
    else if (g.m <= 8) run_skinny_t<true, 8>(dt, g, s);
    else if (g.m <= 16) run_skinny_t<true, 16>(dt, g, s);
    else run_skinny_t<true, 32>(dt, g, s);
    return 2;
  }
  if (g.n <= 32 && g.m >= 64) {
    g_last_kind = K_GEMM_SMALL;
    if (g.n <= 4) run_skinny_t<false, 4>(dt, g, s);
    else if (g.n <= 8) run_skinny_t<false, 8>(dt, g, s);
    else if (g.n <= 16) run_skinny_t<false, 16>(dt, g, s);
    else run_skinny_t<false, 32>(dt, g, s);
    return 2;
  }
  return 0;
}


template <class T, int TM, int TN>
static void run_splitk_t(const GemmArgs& g, const KSplit& ks, hipStream_t s) {
  SplitK sk{};
  const int64_t nz = g.nbo * g.nbi;
  sk.tiles_m = (g.m + TM - 1) / TM, sk.tiles_n = (g.n + TN - 1) / TN;
  sk.n_tiles = nz * sk.tiles_m * sk.tiles_n;
  sk.n_split = ks.n_split;
  sk.k_chunk = ks.n_split > 1 ? ks.k_chunk : g.k;
  sk.part = ks.n_split > 1 ? (char*)g.scratch + 256 : nullptr;
  const int64_t n_out = nz * g.m * g.n;
  dim3 grid((unsigned)std::min<int64_t>(sk.n_tiles, 65536), (unsigned)sk.n_split);
  LAUNCH((gemm_splitk_kernel<T, TM, TN>), grid, dim3(256), 0, s, g, sk);
  if (sk.n_split > 1)
    LAUNCH((splitk_reduce_kernel<T>), dim3((unsigned)std::min<int64_t>((n_out + 255) / 256, 65536)), dim3(256), 0, s, g, sk);
}

static void run_splitk(int dt, const GemmArgs& g, hipStream_t s) {
  g_last_kind = K_GEMM_SPLITK;
  const KSplit ks = plan_splitk(dt, g, scratch_avail(g));
  int tm, tn;
  splitk_tile(g.n, tm, tn);
  dispatch_dt(dt, [&](auto tag) {
    using T = decltype(tag);
    if (tn == 16) run_splitk_t<T, 1, 16>(g, ks, s);
    else if (tn == 8) run_splitk_t<T, 2, 8>(g, ks, s);
    else if (tn == 4) run_splitk_t<T, 4, 4>(g, ks, s);
    else if (tn == 2) run_splitk_t<T, 8, 2>(g, ks, s);
    else run_splitk_t<T, 16, 1>(g, ks, s);
  });
}

// Dot-like shapes: a long k with too few outputs (or skinny columns) to fill the chip.
static bool splitk_worth(const GemmArgs& g) {
  const uint64_t nz = (uint64_t)g.nbo * g.nbi;
  uint64_t par = nz * (uint64_t)g.m * (uint64_t)g.n;
  if (g.m <= 32 && g.n >= 64) par = nz * (uint64_t)g.n;
  else if (g.n <= 32 && g.m >= 64) par = nz * (uint64_t)g.m;
  return g.k >= 256 && par < 32768;
}

// Scratch of the pre-split operands (each half array 256-byte aligned).
struct PreSplitSizes {
  int64_t kp;
  size_t a_half, b_half, total;
};
static PreSplitSizes presplit_sizes(int dt, const GemmArgs& g) {
  PreSplitSizes r;
  const int64_t K = dt == C64 ? 2 * g.k : g.k;
  r.kp = (K + PRESPLIT_BK - 1) / PRESPLIT_BK * PRESPLIT_BK;  // whole k tiles, zero padded
  auto r256 = [](size_t b) { return (b + 255) / 256 * 256; };
  r.a_half = r256((size_t)n_mats(g.sao, g.sai, g.nbo, g.nbi) * (size_t)g.m * (size_t)r.kp * 2);
  r.b_half = r256((size_t)n_mats(g.sbo, g.sbi, g.nbo, g.nbi) * (size_t)g.n * (size_t)r.kp * 2);
  r.total = 2 * r.a_half + 2 * r.b_half;
  return r;
}

// The pre-split kernel reads raw B (GEMM_PRESPLIT_A) when its row table fits
// LDS and its byte offsets stay below 1 GiB (gather tables address one k x n
// batch entry); otherwise B is pre-split as well.
static bool raw_b_fits(int dt, const GemmArgs& g) {
  const int64_t b_span = g.bcol ? g.k * g.n : (g.k - 1) * g.ldb + g.n;
  return presplit_sizes(dt, g).kp / (dt == C64 ? 2 : 1) <= PRESPLIT_MAX_TABLE &&
         b_span * (dt == C64 ? 8 : 4) < (int64_t)(1u << 30);
}

// Split k for the pre-split GEMM with pre-split B: shapes with few output tiles
// and a long k (the last products of closed networks) get about 1024 workgroups,
// each over at least 2048 halves of k; `avail` bytes for the partial products.
static KSplit plan_presplit_split(int dt, const GemmArgs& g, int64_t kp, int bm, size_t avail) {
  const int64_t N = dt == C64 ? 2 * g.n : g.n;
  const int64_t tiles = g.nbo * g.nbi * ((g.m + bm - 1) / bm) * ((N + 127) / 128);
  if (tiles >= 512 || g.nbo * g.nbi > 65535) return KSplit{};
  return split_k(kp, (1024 + tiles - 1) / tiles, 2048, PRESPLIT_BK,
                 dtype_size(dt) * (size_t)(g.nbo * g.nbi * g.m * g.n), avail);
}

// Split-precision products pay off on MFMA-bound shapes only; memory-bound
// skinny steps run the FP32 kernels (same time, full FP32 accuracy).
// Split products pay off with 128-row tiles from m = 64 and with 32-row tiles
// for 8 <= m <= 32 (measured on gfx950 in the 30-qubit circuit's contraction,
// FP16x2 on complex64 x256 read through gather tables, against the 32-row FP32
// MFMA kernel: m=32 k=128 n=131072 9.1 vs 15.9 ms, m=8 k=256 n=32768 3.9 vs
// 7.5 ms; at m=16 k=16 n=262144 the FP32 kernel is the faster, 4.7 vs 5.5 ms).
static bool split_worth(int dt, const GemmArgs& g) {
  const bool cplx = dtype_complex(dt);
  const int64_t K = cplx ? 2 * g.k : g.k, N = cplx ? 2 * g.n : g.n;
  return N >= 128 && K >= 64 && (g.m >= 64 || (g.m >= 8 && g.m <= 32));
}

static int native_mfma_kernel(int dt, int64_t m) {
  const int bm = mfma_rows(dt, m);
  return bm == 32 ? GEMM_MFMA_R32 : bm == 64 ? GEMM_MFMA_R64 : GEMM_MFMA_R128;
}

static int auto_kernel(int dt, int prec, const GemmArgs& g) {
  const bool split_prec = split_precision(prec) && (dt == F32 || dt == C64);
  if (split_prec && split_worth(dt, g)) {
    // Up to eight row tiles of A (32 rows for m <= 32, else 128): A split by a
    // pass, B read in its own precision and split while staging (measured on
    // gfx950, complex64: m=512 k=128 n=32768 x256 20.5 ms against 22.4 for a
    // splitting pass over B and the pre-split GEMM, m=512 k=512 n=8192 15.6
    // against 17.1, m=1024 k=512 n=8192 x256 25.2 against 27.0; the uneven float32
    // workload 571 ms with two row tiles at most, 568 with eight). With more row
    // tiles every tile re-reads and re-splits B, and one splitting pass over B is
    // faster. (Long k, whose raw-B row table does not fit LDS, runs pre-split B.)
    return (g.m + 127) / 128 <= 8 ? GEMM_PRESPLIT_A : GEMM_PRESPLIT;
  }
  if (mfma_worth(dt, g)) return native_mfma_kernel(dt, g.m);
  if (splitk_worth(g)) return GEMM_SPLITK;
  if ((g.m <= (dtype_complex(dt) ? 32 : 64) && g.n >= 64) || (g.n <= 32 && g.m >= 64)) return GEMM_SKINNY;
  return GEMM_SIMPLE;
}

static int launch_gemm_kernels(int dt, int prec, int kernel, const GemmArgs& g, hipStream_t s);

// Kernels that fold the result max into their epilogue (GemmArgs::cmax).
static bool folds_max(int kernel_run) { return kernel_run != GEMM_SIMPLE && kernel_run != -2; }

int launch_gemm(int dt, int prec, int kernel, const GemmArgs& g, hipStream_t s) {
  if (g.cmax) {
    const bool contiguous = g.ldc == g.n && (g.nbi == 1 || g.sci == g.m * g.n) &&
                            (g.nbo == 1 || g.sco == g.nbi * g.m * g.n);
    if (!contiguous) throw Error("GEMM result max tracking needs a contiguous result");
  }
  const int used = launch_gemm_kernels(dt, prec, kernel, g, s);
  if (g.cmax && !folds_max(used)) {
    const int last = g_last_kind;
    launch_maxabs(dt, g.c, (size_t)(g.nbo * g.nbi * g.m * g.n), (uint64_t*)g.cmax, s);
    g_last_kind = last;
  }
  return used < 0 ? GEMM_SKINNY : used;
}

// One pre-split GEMM launch: B pre-split (R = 0), raw B with lanes along columns
// (R = 1, gathered or row-major) or along k (R = 2, gathered).
template <bool C, bool F, int R, int M, bool G>
static void presplit_kernel_launch(const GemmArgs& g, const PreSplit& ps, dim3 grid, hipStream_t s) {
  static bool attr_set = false;
  if (!attr_set) {
    ATN_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&gemm_presplit_kernel<C, F, PRESPLIT_BK, R, M, G>),
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)presplit_lds(M, PRESPLIT_MAX_TABLE)));
    attr_set = true;
  }
  const size_t lds = presplit_lds(M, R ? ps.kp / (C ? 2 : 1) : 0);
  LAUNCH((gemm_presplit_kernel<C, F, PRESPLIT_BK, R, M, G>), grid, dim3(256), lds, s, g, ps);
}
template <bool C, bool F, int R, int M>
static void presplit_gemm(const GemmArgs& g, const PreSplit& ps, dim3 grid, hipStream_t s) {
  if constexpr (R == 1) {
    if (g.bcol) presplit_kernel_launch<C, F, 1, M, true>(g, ps, grid, s);
    else presplit_kernel_launch<C, F, 1, M, false>(g, ps, grid, s);
  } else {
    presplit_kernel_launch<C, F, R, M, R == 2>(g, ps, grid, s);
  }
}

// Scratch of one half array of the 3M form's A planes (256-byte aligned).
static size_t s3m_a_half(const GemmArgs& g) {
  const int64_t kq = (g.k + S3M_BK - 1) / S3M_BK * S3M_BK;
  return ((size_t)n_mats(g.sao, g.sai, g.nbo, g.nbi) * 3 * (size_t)g.m * (size_t)kq * 2 + 255) / 256 * 256;
}

template <bool F, bool G>
static void split3m_launch(const GemmArgs& g, const PreSplit& ps, dim3 grid, hipStream_t s) {
  static bool attr_set = false;
  if (!attr_set) {
    ATN_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&gemm_split3m_kernel<F, G>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)s3m_lds(PRESPLIT_MAX_TABLE)));
    attr_set = true;
  }
  if (ps.kp > PRESPLIT_MAX_TABLE) throw Error("3M split GEMM: k beyond the row table");
  LAUNCH((gemm_split3m_kernel<F, G>), grid, dim3(256), s3m_lds(ps.kp), s, g, ps);
}

static int launch_gemm_kernels(int dt, int prec, int kernel, const GemmArgs& g0, hipStream_t s) {
  GemmArgs g = g0;
  if (kernel == GEMM_AUTO) kernel = auto_kernel(dt, prec, g);
  // factored B: read by the FP32 MFMA kernels (the explicit row tiles) and the streaming kernel
  if (g.bcg && kernel != GEMM_SKINNY && !(kernel >= GEMM_MFMA_R32 && kernel <= GEMM_MFMA_R128))
    kernel = (dt == F32 || dt == C64) && mfma_worth(dt, g) ? native_mfma_kernel(dt, g.m) : GEMM_SKINNY;
  if (g.k == 0) kernel = GEMM_SIMPLE;  // empty sums: every kernel but this one assumes k tiles
  // A through gather tables: read by the simple kernel and by the split passes of
  // the pre-split kernels (one batch entry); any other choice runs the simple kernel
  if (g.acol && !(split_precision(prec) && (dt == F32 || dt == C64) && g.nbo * g.nbi == 1 &&
                  (kernel == GEMM_PRESPLIT || kernel == GEMM_PRESPLIT_A) && g.scratch &&
                  g.scratch_bytes >= gemm_scratch_bytes(dt, prec, kernel, g)))
    kernel = GEMM_SIMPLE;
  if ((g.apend || g.bpend) && g.k > 0 &&
      !(split_precision(prec) && (dt == F32 || dt == C64) &&
        (kernel == GEMM_MFMA || kernel == GEMM_PRESPLIT || kernel == GEMM_PRESPLIT_A)))
    throw Error("pending operand exponents are divided out by the split-precision kernels only");
  if (kernel == GEMM_SPLITK) {
    run_splitk(dt, g, s);
    return GEMM_SPLITK;
  }
  if (kernel == GEMM_SKINNY) {
    const int r = run_skinny(dt, g, s);
    if (r == 1) return GEMM_SKINNY;
    if (r == 2) return -2;  // skinny kernel without the folded max
    kernel = GEMM_SIMPLE;
  }
  if (kernel == GEMM_SIMPLE) {
    run_simple(dt, g, s);
    return GEMM_SIMPLE;
  }
  if (prec == C3M && !g.bcg) {  // complex64 MFMA-shaped steps as three real products
    if (dt != C64) throw Error("precision c3m applies to complex64 only");
    run_c3m(g, s);
    return GEMM_MFMA;
  }
  const bool split_prec = split_precision(prec) && (dt == F32 || dt == C64);
  const int hp = split_halves(prec);  // the half type (3M forms too)
  if ((kernel == GEMM_PRESPLIT || kernel == GEMM_PRESPLIT_A) && !split_prec) kernel = GEMM_MFMA;
  // split precision runs on GEMM_MFMA (split while staging) and GEMM_PRESPLIT;
  // the explicit row-tile kernels are the FP32 / FP64 MFMA kernels in any mode
  if (split_prec && (kernel == GEMM_MFMA || kernel == GEMM_PRESPLIT || kernel == GEMM_PRESPLIT_A)) {
    const PreSplitSizes psz = presplit_sizes(dt, g);
    const bool raw_ok = raw_b_fits(dt, g);
    const bool raw_b = kernel == GEMM_PRESPLIT_A && raw_ok;
    if (kernel == GEMM_PRESPLIT_A && !raw_ok) kernel = GEMM_PRESPLIT;
    const bool pre = kernel != GEMM_MFMA && g.scratch &&
                     g.scratch_bytes >= 256 + (raw_b ? 2 * psz.a_half : psz.total);
    int64_t N = dt == C64 ? 2 * g.n : g.n;
    // 32-row tiles of A for short A (pre-split kernels only)
    const int bm = pre && g.m <= 32 ? 32 : 128;
    uint64_t tiles = (uint64_t)((g.m + bm - 1) / bm) * (uint64_t)((N + 127) / 128);
    dim3 grid((unsigned)std::min<uint64_t>(tiles, 1u << 20), zgrid(g));
    if (g.cmax && g.m <= bm) grid.x = std::min<unsigned>(grid.x, std::max<unsigned>(1, TRACK_WORKGROUPS / grid.y));
    if (hp == FP16X2) {
      if (!g.scratch) throw Error("FP16X2 GEMM needs 16 bytes of device scratch");
      // operand max magnitudes -> per-GEMM power-of-two scales (measured here
      // unless the caller supplies them)
      int64_t na = (g.sao ? g.nbo : 1) * (g.sai ? g.nbi : 1) * g.m * g.k;
      int64_t nb = (g.sbo ? g.nbo : 1) * (g.sbi ? g.nbi : 1) * g.k * g.n;
      unsigned long long* sc = (unsigned long long*)g.scratch;
      if (!g.amax || !g.bmax) launch_fill_zero(g.scratch, 16, s);
      if (!g.amax) launch_maxabs(dt, g.a, (size_t)na, (uint64_t*)sc, s), g.amax = sc;
      if (!g.bmax) launch_maxabs(dt, g.b, (size_t)nb, (uint64_t*)sc + 1, s), g.bmax = sc + 1;
      g_last_kind = K_GEMM_MFMA_FP16X2;
    } else {
      g_last_kind = K_GEMM_MFMA_BF16X3;
    }
    bool f16 = hp == FP16X2, cx = dt == C64;
    if (pre) {  // operands split once into 16-bit (hi, lo) arrays, then the copy-only GEMM
      PreSplit ps;
      char* base = (char*)g.scratch + 256;
      ps.ahi = base, ps.alo = base + psz.a_half;
      ps.bhi = raw_b ? nullptr : base + 2 * psz.a_half, ps.blo = raw_b ? nullptr : base + 2 * psz.a_half + psz.b_half;
      ps.kp = psz.kp;
      const int64_t na = n_mats(g.sao, g.sai, g.nbo, g.nbi) * g.m * psz.kp / 4;
      const int64_t tiles_b = n_mats(g.sbo, g.sbi, g.nbo, g.nbi) * ((psz.kp / (cx ? 2 : 1) + tk_split_bt(cx) - 1) / tk_split_bt(cx)) * ((g.n + 63) / 64);
      const dim3 ga((unsigned)std::min<int64_t>((na + 255) / 256, 65536)), gb((unsigned)std::min<int64_t>(tiles_b, 65536));
#define PRESPLIT_LAUNCH(C_, F_, R_)                                                                        \
  do {                                                                                                     \
    LAUNCH((split_a_kernel<C_, F_>), ga, dim3(256), 0, s, g, ps);                                         \
    if (R_ == 0) LAUNCH((split_bt_kernel<C_, F_>), gb, dim3(256), 0, s, g, ps);                           \
    g_last_kind = F_ ? K_GEMM_MFMA_FP16X2 : K_GEMM_MFMA_BF16X3;                                           \
    if (bm == 32) presplit_gemm<C_, F_, R_, 32>(g, ps, grid, s);                                           \
    else presplit_gemm<C_, F_, R_, 128>(g, ps, grid, s);                                                   \
  } while (0)
#define PRESPLIT_MODES(R_)                                 \
  do {                                                     \
    if (cx && f16) PRESPLIT_LAUNCH(true, true, R_);        \
    else if (cx) PRESPLIT_LAUNCH(true, false, R_);         \
    else if (f16) PRESPLIT_LAUNCH(false, true, R_);        \
    else PRESPLIT_LAUNCH(false, false, R_);                \
  } while (0)
      if (raw_b) {
        // lanes along k where the gathered B has runs of consecutive k offsets
        // (its contracted legs innermost in memory) and short column runs
        const bool lanes_k = g.bcol && g.brow_run % (cx ? 4 : 8) == 0 && g.k % (cx ? 4 : 8) == 0 && g.bcol_run < 16;
        if (split_3m(prec) && cx && bm == 128 && !lanes_k && g.scratch_bytes >= 256 + 2 * s3m_a_half(g) &&
            3 * g.m * ((g.k + S3M_BK - 1) / S3M_BK * S3M_BK) * 2 < ((int64_t)1 << 31)) {
          PreSplit p3;
          const size_t ah = s3m_a_half(g);
          p3.ahi = base, p3.alo = base + ah, p3.bhi = p3.blo = nullptr;
          p3.kp = (g.k + S3M_BK - 1) / S3M_BK * S3M_BK;
          const int64_t na3 = n_mats(g.sao, g.sai, g.nbo, g.nbi) * g.m * p3.kp / 4;
          const dim3 ga3((unsigned)std::min<int64_t>((na3 + 255) / 256, 65536));
          if (f16) {
            LAUNCH((split_a3m_kernel<true>), ga3, dim3(256), 0, s, g, p3);
            g_last_kind = K_GEMM_MFMA_FP16X2;
            if (g.bcol) split3m_launch<true, true>(g, p3, grid, s);
            else split3m_launch<true, false>(g, p3, grid, s);
          } else {
            LAUNCH((split_a3m_kernel<false>), ga3, dim3(256), 0, s, g, p3);
            g_last_kind = K_GEMM_MFMA_BF16X3;
            if (g.bcol) split3m_launch<false, true>(g, p3, grid, s);
            else split3m_launch<false, false>(g, p3, grid, s);
          }
          return GEMM_PRESPLIT_A;
        }
        if (lanes_k)
          PRESPLIT_MODES(2);
        else PRESPLIT_MODES(1);
        return GEMM_PRESPLIT_A;
      }
      // pre-split B: split k over workgroups where the output has few tiles
      const size_t avail = g.scratch_bytes > 256 + psz.total ? g.scratch_bytes - 256 - psz.total : 0;
      const KSplit ks = plan_presplit_split(dt, g, psz.kp, bm, avail);
      if (ks.n_split > 1) {
        ps.kc = ks.k_chunk, ps.part = base + psz.total;
        grid.z = (unsigned)ks.n_split;
      }
      PRESPLIT_MODES(0);
      if (ks.n_split > 1) {
        SplitK sk{};
        sk.n_split = ks.n_split, sk.part = ps.part;
        const int64_t n_out = g.nbo * g.nbi * g.m * g.n;
        const dim3 rg((unsigned)std::min<int64_t>((n_out + 255) / 256, 65536));
        if (cx) LAUNCH((splitk_reduce_kernel<float2>), rg, dim3(256), 0, s, g, sk);
        else LAUNCH((splitk_reduce_kernel<float>), rg, dim3(256), 0, s, g, sk);
      }
#undef PRESPLIT_MODES
#undef PRESPLIT_LAUNCH
      return GEMM_PRESPLIT;
    }
    int64_t cf = dt == C64 ? 2 : 1;
    bool vec = (cf * g.lda) % 4 == 0 && (uintptr_t)g.a % 16 == 0 && (cf * g.sai) % 4 == 0 && (cf * g.sao) % 4 == 0;
    // 72 KiB of dynamic LDS (gfx950 has 160 KiB per CU): raise the per-kernel limit once.
#define SPLIT_LAUNCH(C_, F_, V_)                                                                          \
  do {                                                                                                   \
    static bool attr_set = false;                                                                        \
    if (!attr_set) {                                                                                     \
      ATN_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&gemm_split_kernel<C_, F_, V_>),         \
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)SPLIT_LDS));          \
      attr_set = true;                                                                                   \
    }                                                                                                    \
    LAUNCH((gemm_split_kernel<C_, F_, V_>), grid, dim3(256), SPLIT_LDS, s, g);                            \
  } while (0)
    if (cx && f16) {
      if (vec) SPLIT_LAUNCH(true, true, true);
      else SPLIT_LAUNCH(true, true, false);
    } else if (cx) {
      if (vec) SPLIT_LAUNCH(true, false, true);
      else SPLIT_LAUNCH(true, false, false);
    } else if (f16) {
      if (vec) SPLIT_LAUNCH(false, true, true);
      else SPLIT_LAUNCH(false, true, false);
    } else {
      if (vec) SPLIT_LAUNCH(false, false, true);
      else SPLIT_LAUNCH(false, false, false);
    }
#undef SPLIT_LAUNCH
    return GEMM_MFMA;
  }
  // row tile: 32 or 64 rows for skinny operands, else 128 (FP32) / 64 (FP64)
  const int bm = mfma_rows_for(dt, kernel, g.m);
  switch (dt) {
    case F32:
      if (bm == 32) run_mfma<float, false, 32, 128, 16>(dt, g, s);
      else if (bm == 64) run_mfma<float, false, 64, 128, 16>(dt, g, s);
      else run_mfma<float, false, 128, 128, 16>(dt, g, s);
      break;
    case C64:
      if (bm == 32) run_mfma<float, true, 32, 128, 16>(dt, g, s);
      else if (bm == 64) run_mfma<float, true, 64, 128, 16>(dt, g, s);
      else run_mfma<float, true, 128, 128, 16>(dt, g, s);
      break;
    case F64:
      if (bm == 32) run_mfma<double, false, 32, 64, 16>(dt, g, s);
      else run_mfma<double, false, 64, 64, 16>(dt, g, s);
      break;
    case C128:
      if (bm == 32) run_mfma<double, true, 32, 64, 16>(dt, g, s);
      else run_mfma<double, true, 64, 64, 16>(dt, g, s);
      break;
    default: throw Error("unsupported dtype code");
  }
  return GEMM_MFMA;
}

bool split_products_apply(int dt, int prec, int64_t m, int64_t k, int64_t n) {
  GemmArgs g{};
  g.m = m, g.k = k, g.n = n;
  return split_precision(prec) && (dt == F32 || dt == C64) && split_worth(dt, g);
}

bool gemm_reads_factored(int dt, int prec, int64_t m, int64_t k, int64_t n) {
  GemmArgs g{};
  g.m = m, g.k = k, g.n = n, g.nbo = g.nbi = 1;
  if (!(dt == F32 || dt == C64) || n > (int64_t)UINT32_MAX) return false;
  const int kern = auto_kernel(dt, prec, g);
  if (kern == GEMM_SKINNY) return n >= 64 && m <= (dtype_complex(dt) ? 32 : 64);
  return kern >= GEMM_MFMA_R32 && kern <= GEMM_MFMA_R128 && prec != C3M;
}

bool gemm_auto_streams(int dt, int prec, int64_t m, int64_t k, int64_t n) {
  GemmArgs g{};
  g.m = m, g.k = k, g.n = n, g.nbo = g.nbi = 1;
  return auto_kernel(dt, prec, g) == GEMM_SKINNY && n >= 64 && m <= (dtype_complex(dt) ? 32 : 64);
}

bool gemm_runs_simple(int dt, int prec, int kernel, const GemmArgs& g) {
  if (kernel == GEMM_AUTO) kernel = auto_kernel(dt, prec, g);
  return g.k == 0 || kernel == GEMM_SIMPLE;
}

size_t gemm_scratch_bytes(int dt, int prec, int kernel, const GemmArgs& g) {
  if (kernel == GEMM_AUTO) kernel = auto_kernel(dt, prec, g);
  KSplit ks;
  if (kernel == GEMM_SPLITK) ks = plan_splitk(dt, g, MAX_PARTIAL_BYTES);
  else if (kernel == GEMM_MFMA || kernel >= GEMM_MFMA_R32)
    ks = plan_mfma_split(dt, prec == C3M ? NATIVE : prec, g, MAX_PARTIAL_BYTES,
                         prec == C3M ? (g.m <= 64 ? 64 : 128) : mfma_rows_for(dt, kernel, g.m));
  size_t part = ks.n_split > 1 ? (size_t)ks.n_split * dtype_size(dt) * (size_t)(g.nbo * g.nbi * g.m * g.n) : 0;
  if (split_precision(prec) && (dt == F32 || dt == C64)) {
    if (kernel == GEMM_PRESPLIT) {
      const PreSplitSizes psz = presplit_sizes(dt, g);
      const KSplit ks = plan_presplit_split(dt, g, psz.kp, g.m <= 32 ? 32 : 128, MAX_PARTIAL_BYTES);
      part = psz.total + (ks.n_split > 1 ? (size_t)ks.n_split * dtype_size(dt) * (size_t)(g.nbo * g.nbi * g.m * g.n) : 0);
    }
    if (kernel == GEMM_PRESPLIT_A) {
      if (raw_b_fits(dt, g)) {
        part = 2 * presplit_sizes(dt, g).a_half;
      } else {  // B pre-split as well (GEMM_PRESPLIT)
        const PreSplitSizes psz = presplit_sizes(dt, g);
        const KSplit ks = plan_presplit_split(dt, g, psz.kp, g.m <= 32 ? 32 : 128, MAX_PARTIAL_BYTES);
        part = psz.total + (ks.n_split > 1 ? (size_t)ks.n_split * dtype_size(dt) * (size_t)(g.nbo * g.nbi * g.m * g.n) : 0);
      }
      if (split_3m(prec) && dt == C64 && raw_b_fits(dt, g)) part = std::max(part, 2 * s3m_a_half(g));
    }
  }
  return 256 + part;
}

void launch_gemm_adjoint(int dt, int opa, int opb, const GemmArgs& g, hipStream_t s) { run_simple(dt, g, s, opa, opb); }

void launch_gemm_simple_group(int dt, const SimpleGemmDesc* d, int n, uint64_t blocks, char* ws,
                              const void* const* inputs, const int32_t* tables, hipStream_t s) {
  if (n <= 0 || blocks == 0) return;
  g_last_kind = K_GEMM_SIMPLE;
  const dim3 grid((unsigned)std::min<uint64_t>(blocks, 1u << 16));
  dispatch_dt(dt, [&](auto tag) {
    using T = decltype(tag);
    LAUNCH((gemm_simple_group_kernel<T>), grid, dim3(256), 0, s, d, n, blocks, ws, inputs, tables);
  });
}

}  // namespace atn
