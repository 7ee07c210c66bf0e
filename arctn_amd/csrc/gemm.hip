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

#include "kernels.hpp"

namespace atn {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef double v4d __attribute__((ext_vector_type(4)));

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
        T a = OPA ? CM<T>::conj(A[p * g.lda + i]) : A[i * g.lda + p];
        T b = OPB ? CM<T>::conj(B[j * g.ldb + p]) : B[p * g.ldb + j];
        acc = CM<T>::mac(acc, a, b);
      }
      C[i * g.ldc + j] = acc;
    }
  }
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
template <class F, bool CPLX, int BM, int BN, int BK>
__global__ __launch_bounds__(256) void gemm_mfma_kernel(GemmArgs g) {
  constexpr int WM = BM / 2, WN = BN / 2;
  constexpr int TM = WM / 16, TN = WN / 16;
  constexpr int PAD = 64 / (int)sizeof(F);  // row stride = 64 B mod 128 B: fragment reads hit all banks
  constexpr int SA = BM + PAD, SB = BN + PAD;
  constexpr int A_PER = BM * BK / 256;
  constexpr int B_PER = CPLX ? (BK / 2) * (BN / 2) / 256 : BK * BN / 256;
  static_assert(A_PER >= 1 && B_PER >= 1, "tile too small for 256 threads");
  using V = typename Mfma<F>::V;
  __shared__ F As[BK * SA];
  __shared__ F Bs[BK * SB];

  const int64_t M = g.m, N = CPLX ? 2 * g.n : g.n, K = CPLX ? 2 * g.k : g.k;
  const int64_t lda = CPLX ? 2 * g.lda : g.lda;  // real units
  const int64_t ldc = CPLX ? 2 * g.ldc : g.ldc;
  const int64_t ldb = g.ldb;                      // complex units when CPLX
  const int tiles_n = (int)((N + BN - 1) / BN);
  const int tiles_m = (int)((M + BM - 1) / BM);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm0 = (wave >> 1) * WM, wn0 = (wave & 1) * WN;
  const uint64_t nz = (uint64_t)g.nbo * g.nbi;
  const uint64_t ntile = (uint64_t)tiles_m * tiles_n;

  for (uint64_t z = blockIdx.y; z < nz; z += gridDim.y) {
    uint64_t zo = z / g.nbi, zi = z - zo * g.nbi;
    const F* A = (const F*)g.a + (zo * g.sao + zi * g.sai) * (CPLX ? 2 : 1);
    const F* B = (const F*)g.b + (zo * g.sbo + zi * g.sbi) * (CPLX ? 2 : 1);
    F* C = (F*)g.c + (zo * g.sco + zi * g.sci) * (CPLX ? 2 : 1);
    for (uint64_t t = blockIdx.x; t < ntile; t += gridDim.x) {
      const int64_t m0 = (int64_t)(t / tiles_n) * BM, n0 = (int64_t)(t % tiles_n) * BN;
      V acc[TM][TN];
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) acc[i][j] = V{0, 0, 0, 0};

      F ra[A_PER];
      F rb[CPLX ? 2 * B_PER : B_PER];
      auto load = [&](int64_t k0) {
#pragma unroll
        for (int u = 0; u < A_PER; ++u) {
          int e = tid + 256 * u, col = e % BK, row = e / BK;
          int64_t gr = m0 + row, gc = k0 + col;
          ra[u] = (gr < M && gc < K) ? A[gr * lda + gc] : (F)0;
        }
        if constexpr (CPLX) {
#pragma unroll
          for (int u = 0; u < B_PER; ++u) {
            int e = tid + 256 * u, j = e % (BN / 2), p = e / (BN / 2);
            int64_t gp = k0 / 2 + p, gj = n0 / 2 + j;
            bool ok = gp < g.k && gj < g.n;
            rb[2 * u] = ok ? B[(gp * ldb + gj) * 2] : (F)0;
            rb[2 * u + 1] = ok ? B[(gp * ldb + gj) * 2 + 1] : (F)0;
          }
        } else {
#pragma unroll
          for (int u = 0; u < B_PER; ++u) {
            int e = tid + 256 * u, col = e % BN, row = e / BN;
            int64_t gr = k0 + row, gc = n0 + col;
            rb[u] = (gr < K && gc < N) ? B[gr * ldb + gc] : (F)0;
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

      const int64_t nk = (K + BK - 1) / BK;
      load(0);
      store();
      __syncthreads();
      for (int64_t kt = 0; kt < nk; ++kt) {
        if (kt + 1 < nk) load((kt + 1) * BK);
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
            int64_t row = m0 + wm0 + i * 16 + Mfma<F>::row(lane, r), col = n0 + wn0 + j * 16 + cl;
            if (row < M && col < N) C[row * ldc + col] = acc[i][j][r];
          }
    }
  }
}

// ------------------------------------------------------------------ dispatch --
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
static void run_mfma(const GemmArgs& g, hipStream_t s) {
  int64_t N = CPLX ? 2 * g.n : g.n;
  uint64_t tiles = (uint64_t)((g.m + BM - 1) / BM) * (uint64_t)((N + BN - 1) / BN);
  dim3 grid((unsigned)std::min<uint64_t>(tiles, 1u << 20), zgrid(g));
  g_last_kind = sizeof(F) == 4 ? K_GEMM_MFMA_F32 : K_GEMM_MFMA_F64;
  LAUNCH((gemm_mfma_kernel<F, CPLX, BM, BN, BK>), grid, dim3(256), 0, s, g);
}

static bool mfma_worth(int dt, const GemmArgs& g) {
  bool cplx = dtype_complex(dt);
  int64_t K = cplx ? 2 * g.k : g.k, N = cplx ? 2 * g.n : g.n;
  return g.m >= 32 && N >= 32 && K >= 16;
}

int launch_gemm(int dt, int prec, int kernel, const GemmArgs& g, hipStream_t s) {
  if (kernel == GEMM_AUTO) kernel = mfma_worth(dt, g) ? GEMM_MFMA : GEMM_SIMPLE;
  if (kernel == GEMM_SIMPLE) {
    run_simple(dt, g, s);
    return GEMM_SIMPLE;
  }
  (void)prec;
  switch (dt) {
    case F32: run_mfma<float, false, 128, 128, 16>(g, s); break;
    case C64: run_mfma<float, true, 128, 128, 16>(g, s); break;
    case F64: run_mfma<double, false, 64, 64, 16>(g, s); break;
    case C128: run_mfma<double, true, 64, 64, 16>(g, s); break;
    default: throw Error("unsupported dtype code");
  }
  return GEMM_MFMA;
}

void launch_gemm_adjoint(int dt, int opa, int opb, const GemmArgs& g, hipStream_t s) { run_simple(dt, g, s, opa, opb); }

}  // namespace atn
