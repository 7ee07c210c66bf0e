// Data-movement kernels: LDS-tiled permute, sums, traces, diagonals, slice
// gathers, cross-slice reductions and power-of-two strip. Pure copies and
// sequential-order sums, so results are bit-identical to the reference's
// src/tensor.rs index loops.
#include <algorithm>
#include <numeric>

#include "kernels.hpp"

namespace atn {

thread_local size_t g_launch_count = 0;
thread_local int g_last_kind = K_OTHER;

const char* kernel_kind_name(int k) {
  static const char* names[K_NKINDS] = {"other", "permute_tile", "permute_generic", "sum_axis", "trace_diag", "slice_gather",
                                        "slice_reduce_f64", "output_scatter", "gemm_simple", "gemm_skinny", "gemm_mfma_f32",
                                        "gemm_mfma_f64", "gemm_mfma_bf16x3", "gemm_mfma_fp16x2", "gemm_mfma_c3m",
                                        "gemm_splitk", "strip", "copy"};
  return (k >= 0 && k < K_NKINDS) ? names[k] : "other";
}

template <int ES> struct Bits;
template <> struct Bits<4> { using T = uint32_t; };
template <> struct Bits<8> { using T = uint64_t; };
template <> struct Bits<16> { using T = uint4; };

// mag: max(|re|, |im|) (power-of-two strip scales); abs: the magnitude of
// Scalar::abs (src/tensor.rs:125-127, 173-175, 245-247, 317-319): |x| widened to
// f64, hypot of the widened parts for complex; inv: the reciprocal 1/m as the
// element type's real part (Scalar::from_real(1.0 / factor), compiled.rs:423).
template <class T> struct Arith;
template <> struct Arith<float> {
  using R = float;
  static __device__ float add(float a, float b) { return a + b; }
  static __device__ double mag(float a) { return fabs((double)a); }
  static __device__ double abs(float a) { return fabs((double)a); }
  static __device__ float scale(float a, float s) { return a * s; }
};
template <> struct Arith<double> {
  using R = double;
  static __device__ double add(double a, double b) { return a + b; }
  static __device__ double mag(double a) { return fabs(a); }
  static __device__ double abs(double a) { return fabs(a); }
  static __device__ double scale(double a, double s) { return a * s; }
};
template <> struct Arith<float2> {
  using R = float;
  static __device__ float2 add(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
  static __device__ double mag(float2 a) { return fmax(fabs((double)a.x), fabs((double)a.y)); }
  static __device__ double abs(float2 a) { return hypot((double)a.x, (double)a.y); }
  static __device__ float2 scale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
};
template <> struct Arith<double2> {
  using R = double;
  static __device__ double2 add(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
  static __device__ double mag(double2 a) { return fmax(fabs(a.x), fabs(a.y)); }
  static __device__ double abs(double2 a) { return hypot(a.x, a.y); }
  static __device__ double2 scale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
};

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
template <class F>
static void dispatch_es(int es, F&& f) {
  switch (es) {
    case 4: f(std::integral_constant<int, 4>{}); break;
    case 8: f(std::integral_constant<int, 8>{}); break;
    case 16: f(std::integral_constant<int, 16>{}); break;
    default: throw Error("unsupported element size");
  }
}

static unsigned grid_for(uint64_t work, unsigned threads, unsigned cap = 65536u) {
  uint64_t g = (work + threads - 1) / threads;
  if (g < 1) g = 1;
  return (unsigned)std::min<uint64_t>(g, cap);
}

#define LAUNCH(kernel, grid, block, lds, stream, ...)                            \
  do {                                                                            \
    hipLaunchKernelGGL(kernel, dim3(grid), dim3(block), lds, stream, __VA_ARGS__); \
    ATN_HIP(hipGetLastError());                                                   \
    ++g_launch_count;                                                             \
  } while (0)

// ------------------------------------------------------------------ permute --
// One workgroup per tile of the axis set U. Phase 1 reads U with the input's
// innermost axes (G_in) contiguous across lanes and scatters into LDS laid out
// in output order; phase 2 streams LDS out with the output's innermost axes
// (G_out) contiguous. Offset tables for the non-contiguous part of each phase
// are built once per workgroup in LDS. Every LDS row of lo_out elements is
// padded by one element, so the transposed scatter of phase 1 (lanes writing
// consecutive rows) spreads over the banks instead of hitting one.
template <int ES>
__global__ __launch_bounds__(256) void permute_tile_kernel(PermTile a, const typename Bits<ES>::T* __restrict__ in,
                                                           typename Bits<ES>::T* __restrict__ out) {
  using E = typename Bits<ES>::T;
  extern __shared__ __align__(16) char smem[];
  int64_t* in_hi = reinterpret_cast<int64_t*>(smem);
  int64_t* out_hi = in_hi + a.n_hi_in;
  int32_t* lds_hi = reinterpret_cast<int32_t*>(out_hi + a.n_hi_out);
  int32_t* lds_lo = lds_hi + a.n_hi_in;
  size_t tab_bytes = (size_t)(a.n_hi_in + a.n_hi_out) * 8 + (size_t)(a.n_hi_in + a.lo_in) * 4;
  E* tile = reinterpret_cast<E*>(smem + ((tab_bytes + 15) & ~(size_t)15));

  for (uint32_t h = threadIdx.x; h < a.n_hi_in; h += blockDim.x) {
    uint32_t r = h;
    int64_t o = 0;
    int32_t l = 0;
    for (int d = a.n_hin - 1; d >= 0; --d) {
      uint32_t dim = a.hin_dims[d], q = r / dim, i = r - q * dim;
      r = q;
      o += (int64_t)i * a.hin_in[d];
      l += (int32_t)i * a.hin_lds[d];
    }
    in_hi[h] = o;
    lds_hi[h] = l;
  }
  for (uint32_t h = threadIdx.x; h < a.lo_in; h += blockDim.x) {
    uint32_t r = h;
    int32_t l = 0;
    for (int d = a.n_gin - 1; d >= 0; --d) {
      uint32_t dim = a.gin_dims[d], q = r / dim, i = r - q * dim;
      r = q;
      l += (int32_t)i * a.gin_lds[d];
    }
    lds_lo[h] = l;
  }
  for (uint32_t h = threadIdx.x; h < a.n_hi_out; h += blockDim.x) {
    uint32_t r = h;
    int64_t o = 0;
    for (int d = a.n_hout - 1; d >= 0; --d) {
      uint32_t dim = a.hout_dims[d], q = r / dim, i = r - q * dim;
      r = q;
      o += (int64_t)i * a.hout_out[d];
    }
    out_hi[h] = o;
  }
  __syncthreads();

  for (uint64_t t = blockIdx.x; t < a.n_tiles; t += gridDim.x) {
    int64_t bin = 0, bout = 0;
    if (a.n_tiles <= 0xffffffffull) {
      uint32_t r = (uint32_t)t;
      for (int d = a.n_r - 1; d >= 0; --d) {
        uint32_t dim = a.r_dims[d], q = r / dim, i = r - q * dim;
        r = q;
        bin += (int64_t)i * a.r_in[d];
        bout += (int64_t)i * a.r_out[d];
      }
    } else {
      uint64_t r = t;
      for (int d = a.n_r - 1; d >= 0; --d) {
        uint64_t dim = a.r_dims[d], q = r / dim, i = r - q * dim;
        r = q;
        bin += (int64_t)i * a.r_in[d];
        bout += (int64_t)i * a.r_out[d];
      }
    }
    const E* src = in + bin;
    E* dst = out + bout;
    // Eight independent global loads in flight per lane before the LDS scatter.
    constexpr int UN = 8;
    for (uint32_t j0 = threadIdx.x; j0 < a.u_size; j0 += UN * blockDim.x) {
      E v[UN];
      int32_t d[UN];
#pragma unroll
      for (int q = 0; q < UN; ++q) {
        uint32_t j = j0 + q * blockDim.x;
        if (j < a.u_size) {
          uint32_t hi = a.sh_in >= 0 ? j >> a.sh_in : j / a.lo_in;
          uint32_t lo = j - hi * a.lo_in;
          v[q] = src[in_hi[hi] + lo];
          uint32_t jo = lds_lo[lo] + lds_hi[hi];  // output-order position in U
          d[q] = jo + (a.sh_out >= 0 ? jo >> a.sh_out : jo / a.lo_out);
        }
      }
#pragma unroll
      for (int q = 0; q < UN; ++q)
        if (j0 + q * blockDim.x < a.u_size) tile[d[q]] = v[q];
    }
    __syncthreads();
    for (uint32_t j0 = threadIdx.x; j0 < a.u_size; j0 += UN * blockDim.x) {
#pragma unroll
      for (int q = 0; q < UN; ++q) {
        uint32_t j = j0 + q * blockDim.x;
        if (j < a.u_size) {
          uint32_t hi = a.sh_out >= 0 ? j >> a.sh_out : j / a.lo_out;
          uint32_t lo = j - hi * a.lo_out;
          dst[out_hi[hi] + lo] = tile[j + hi];
        }
      }
    }
    __syncthreads();
  }
}

template <int ES>
__global__ __launch_bounds__(256) void permute_generic_kernel(PermGeneric a, const typename Bits<ES>::T* __restrict__ in,
                                                              typename Bits<ES>::T* __restrict__ out) {
  using E = typename Bits<ES>::T;
  constexpr int UN = 4;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t e0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e0 < a.numel; e0 += UN * stride) {
    E v[UN];
#pragma unroll
    for (int q = 0; q < UN; ++q) {
      uint64_t e = e0 + q * stride;
      if (e < a.numel) {
        uint64_t r = e;
        int64_t off = 0;
        for (int d = a.rank - 1; d >= 0; --d) {
          uint64_t dim = a.dims[d], qq = r / dim;
          off += (int64_t)(r - qq * dim) * a.in_str[d];
          r = qq;
        }
        v[q] = in[off];
      }
    }
#pragma unroll
    for (int q = 0; q < UN; ++q)
      if (e0 + q * stride < a.numel) out[e0 + q * stride] = v[q];
  }
}

PermPlan plan_permute(const std::vector<size_t>& shape0, const std::vector<size_t>& perm0, size_t batch, int es) {
  PermPlan p;
  p.es = es;
  p.numel = batch * prod(shape0);
  // Drop unit axes, then merge input-adjacent axes that stay adjacent in the output.
  std::vector<size_t> keep;  // input axes kept
  for (size_t a = 0; a < shape0.size(); ++a)
    if (shape0[a] != 1) keep.push_back(a);
  std::vector<size_t> perm;  // over kept axes, in output order
  for (size_t d : perm0)
    if (shape0[d] != 1) perm.push_back((size_t)(std::find(keep.begin(), keep.end(), d) - keep.begin()));
  std::vector<size_t> shape;
  for (size_t a : keep) shape.push_back(shape0[a]);
  // groups in output order: runs of consecutive input axes
  std::vector<std::vector<size_t>> groups;
  for (size_t d = 0; d < perm.size(); ++d) {
    if (!groups.empty() && groups.back().back() + 1 == perm[d]) groups.back().push_back(perm[d]);
    else groups.push_back({perm[d]});
  }
  // merged axes, input order: (dim, in_first_axis); output order = order of groups
  size_t ng = groups.size();
  std::vector<size_t> gdim(ng), gfirst(ng);
  for (size_t g = 0; g < ng; ++g) {
    size_t d = 1;
    for (size_t a : groups[g]) d *= shape[a];
    gdim[g] = d;
    gfirst[g] = groups[g][0];
  }
  std::vector<size_t> in_order(ng);  // merged axes sorted by input position
  std::iota(in_order.begin(), in_order.end(), 0);
  std::sort(in_order.begin(), in_order.end(), [&](size_t x, size_t y) { return gfirst[x] < gfirst[y]; });
  bool ident = true;
  for (size_t i = 0; i < ng; ++i) ident = ident && in_order[i] == i;
  if (ident) {
    p.kind = 0;
    return p;
  }
  // Split large merged axes into factors <= 64 so tiles stay small (a reshape
  // of a contiguous group is valid on both sides).
  struct Ax { size_t dim; size_t out_pos; size_t in_pos; };
  std::vector<Ax> axes;  // in input order, each with its output position (fractional order via pairs)
  {
    std::vector<std::vector<size_t>> pieces(ng);  // factors of each merged axis, outer to inner
    for (size_t g = 0; g < ng; ++g) {
      size_t d = gdim[g];
      std::vector<size_t> f;
      while (d > 64) {
        size_t best = 0;
        for (size_t c = 64; c >= 8; --c)
          if (d % c == 0) {
            best = c;
            break;
          }
        if (!best) break;
        f.push_back(best);
        d /= best;
      }
      f.push_back(d);
      std::reverse(f.begin(), f.end());  // outer factor first
      pieces[g] = f;
    }
    // Output order enumerates groups in output order, pieces outer->inner.
    size_t opos = 0;
    std::vector<std::vector<size_t>> out_pos(ng);
    for (size_t g = 0; g < ng; ++g)
      for (size_t q = 0; q < pieces[g].size(); ++q) out_pos[g].push_back(opos++);
    size_t ipos = 0;
    for (size_t gi : in_order)
      for (size_t q = 0; q < pieces[gi].size(); ++q) axes.push_back(Ax{pieces[gi][q], out_pos[gi][q], ipos++});
  }
  size_t r = axes.size();
  // strides
  std::vector<int64_t> in_str(r), out_str(r);
  {
    int64_t s = 1;
    for (size_t a = r; a-- > 0;) {
      in_str[a] = s;
      s *= (int64_t)axes[a].dim;
    }
    std::vector<size_t> by_out(r);
    for (size_t a = 0; a < r; ++a) by_out[axes[a].out_pos] = a;
    s = 1;
    for (size_t d = r; d-- > 0;) {
      out_str[by_out[d]] = s;
      s *= (int64_t)axes[by_out[d]].dim;
    }
  }
  std::vector<size_t> by_out(r);
  for (size_t a = 0; a < r; ++a) by_out[axes[a].out_pos] = a;

  auto generic = [&]() {
    p.kind = 2;
    PermGeneric& g = p.gen;
    if (r + 1 > (size_t)MAXR) throw Error("permute rank exceeds the generic kernel limit");
    g.rank = 0;
    if (batch > 1) {
      g.dims[g.rank] = (uint32_t)batch;
      g.in_str[g.rank] = (int64_t)prod(shape0);
      g.rank++;
    }
    for (size_t d = 0; d < r; ++d) {
      g.dims[g.rank] = (uint32_t)axes[by_out[d]].dim;
      g.in_str[g.rank] = in_str[by_out[d]];
      g.rank++;
    }
    g.numel = p.numel;
    p.threads = 256;
  };

  // Tile budget: the LDS tile holds at most this many elements (32 KiB).
  const size_t umax = (size_t)32768 / (size_t)es;
  for (size_t tmin : {64, 32, 16, 8}) {
    // G_in: innermost input axes with product >= tmin; G_out likewise in output order.
    std::vector<bool> in_gin(r, false), in_gout(r, false);
    size_t prodin = 1, prodout = 1;
    size_t ni = 0, no = 0;  // number of input-/output-innermost axes taken
    for (size_t a = r; a-- > 0 && prodin < tmin;) in_gin[a] = true, prodin *= axes[a].dim, ++ni;
    for (size_t d = r; d-- > 0 && prodout < tmin;) in_gout[by_out[d]] = true, prodout *= axes[by_out[d]].dim, ++no;
    auto usize = [&]() {
      size_t u = 1;
      for (size_t a = 0; a < r; ++a)
        if (in_gin[a] || in_gout[a]) u *= axes[a].dim;
      return u;
    };
    // Grow both groups while the tile fits: longer contiguous runs on both
    // sides and more work per tile (small tiles are dominated by per-tile
    // index decoding and barriers).
    for (bool grew = true; grew;) {
      grew = false;
      if (ni < r) {
        size_t a = r - 1 - ni;
        in_gin[a] = true;
        if (usize() <= umax) prodin *= axes[a].dim, ++ni, grew = true;
        else in_gin[a] = false;
      }
      if (no < r) {
        size_t a = by_out[r - 1 - no];
        bool was = in_gout[a];
        in_gout[a] = true;
        if (usize() <= umax) prodout *= axes[a].dim, ++no, grew = true;
        else in_gout[a] = was;
      }
    }
    size_t u = usize();
    int nu = 0;
    for (size_t a = 0; a < r; ++a)
      if (in_gin[a] || in_gout[a]) ++nu;
    size_t lo_in = prodin, lo_out = prodout;
    size_t n_hi_in = u / lo_in, n_hi_out = u / lo_out;
    size_t tab = (n_hi_in + n_hi_out) * 8 + (n_hi_in + lo_in) * 4;
    size_t lds = ((tab + 15) & ~(size_t)15) + (u + n_hi_out) * (size_t)es;
    int n_gin = 0, n_gout = 0;
    for (size_t a = 0; a < r; ++a) n_gin += in_gin[a], n_gout += in_gout[a];
    if (lds > 64 * 1024 || nu - n_gin > MAXT || n_gin > MAXT || nu - n_gout > MAXT || r - nu + 1 > (size_t)MAXR) continue;
    PermTile& t = p.tile;
    // LDS strides: row-major over U in output order.
    std::vector<int64_t> lds_str(r, 0);
    {
      int64_t s = 1;
      for (size_t d = r; d-- > 0;) {
        size_t a = by_out[d];
        if (in_gin[a] || in_gout[a]) lds_str[a] = s, s *= (int64_t)axes[a].dim;
      }
    }
    t.n_hin = t.n_gin = t.n_hout = 0;
    for (size_t a = 0; a < r; ++a) {  // input order
      if (in_gin[a]) {
        t.gin_dims[t.n_gin] = (uint32_t)axes[a].dim;
        t.gin_lds[t.n_gin] = (int32_t)lds_str[a];
        t.n_gin++;
      } else if (in_gout[a]) {
        t.hin_dims[t.n_hin] = (uint32_t)axes[a].dim;
        t.hin_in[t.n_hin] = in_str[a];
        t.hin_lds[t.n_hin] = (int32_t)lds_str[a];
        t.n_hin++;
      }
    }
    for (size_t d = 0; d < r; ++d) {  // output order
      size_t a = by_out[d];
      if ((in_gin[a] || in_gout[a]) && !in_gout[a]) {
        t.hout_dims[t.n_hout] = (uint32_t)axes[a].dim;
        t.hout_out[t.n_hout] = out_str[a];
        t.n_hout++;
      }
    }
    // Grid axes: batch (outermost) then the remaining axes in output order.
    t.n_r = 0;
    if (batch > 1) {
      t.r_dims[t.n_r] = (uint32_t)batch;
      t.r_in[t.n_r] = (int64_t)prod(shape0);
      t.r_out[t.n_r] = (int64_t)prod(shape0);
      t.n_r++;
    }
    for (size_t d = 0; d < r; ++d) {
      size_t a = by_out[d];
      if (in_gin[a] || in_gout[a]) continue;
      t.r_dims[t.n_r] = (uint32_t)axes[a].dim;
      t.r_in[t.n_r] = in_str[a];
      t.r_out[t.n_r] = out_str[a];
      t.n_r++;
    }
    t.u_size = (uint32_t)u;
    t.lo_in = (uint32_t)lo_in;
    t.lo_out = (uint32_t)lo_out;
    t.n_hi_in = (uint32_t)n_hi_in;
    t.n_hi_out = (uint32_t)n_hi_out;
    auto lg = [](size_t v) { return (v & (v - 1)) == 0 ? (int32_t)__builtin_ctzll(v) : -1; };
    t.sh_in = lg(lo_in);
    t.sh_out = lg(lo_out);
    t.n_tiles = p.numel / u;
    p.kind = 1;
    p.lds = lds;
    p.threads = (unsigned)std::min<size_t>(256, std::max<size_t>(64, (u + 63) / 64 * 64));
    return p;
  }
  generic();
  return p;
}

void launch_permute(const PermPlan& p, const void* in, void* out, hipStream_t s) {
  if (p.kind == 0) {
    if (in != out) ATN_HIP(hipMemcpyAsync(out, in, p.numel * p.es, hipMemcpyDeviceToDevice, s));
    return;
  }
  dispatch_es(p.es, [&](auto esc) {
    constexpr int ES = decltype(esc)::value;
    using E = typename Bits<ES>::T;
    if (p.kind == 1) {
      unsigned grid = (unsigned)std::min<uint64_t>(p.tile.n_tiles, 2048u * 4u);
      g_last_kind = K_PERMUTE_TILE;
      LAUNCH(permute_tile_kernel<ES>, grid, p.threads, p.lds, s, p.tile, (const E*)in, (E*)out);
    } else {
      g_last_kind = K_PERMUTE_GENERIC;
      LAUNCH(permute_generic_kernel<ES>, grid_for((p.gen.numel + 3) / 4, 256), 256, 0, s, p.gen, (const E*)in, (E*)out);
    }
  });
}

// ------------------------------------------------------------- sum / trace ---
template <class T>
__global__ void sum_axis_exact_kernel(const T* __restrict__ in, T* __restrict__ out, uint64_t outer, uint64_t d,
                                      uint64_t inner);

void launch_sum_axis(int dt, size_t outer, size_t d, size_t inner, const void* in, void* out, hipStream_t s) {
  dispatch_dt(dt, [&](auto tag) {
    using T = decltype(tag);
    g_last_kind = K_SUM;
    LAUNCH(sum_axis_exact_kernel<T>, grid_for(outer * inner, 256), 256, 0, s, (const T*)in, (T*)out, (uint64_t)outer,
           (uint64_t)d, (uint64_t)inner);
  });
}

// The reference starts from zero and adds every term (acc = 0; acc += x);
// starting from the first term is identical for IEEE addition since 0 + x == x
// for every x except -0.0, where 0 + (-0.0) = +0.0. Keep the exact form.
template <class T>
__device__ inline T zero_of();
template <> __device__ inline float zero_of<float>() { return 0.f; }
template <> __device__ inline double zero_of<double>() { return 0.0; }
template <> __device__ inline float2 zero_of<float2>() { return make_float2(0.f, 0.f); }
template <> __device__ inline double2 zero_of<double2>() { return make_double2(0.0, 0.0); }

template <class T>
__global__ void sum_axis_exact_kernel(const T* __restrict__ in, T* __restrict__ out, uint64_t outer, uint64_t d,
                                      uint64_t inner) {
  uint64_t n = outer * inner;
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t o = e / inner, q = e - o * inner;
    const T* p = in + o * d * inner + q;
    T acc = zero_of<T>();
    for (uint64_t i = 0; i < d; ++i) acc = Arith<T>::add(acc, p[i * inner]);
    out[e] = acc;
  }
}

template <class T>
__global__ void trace_kernel(const T* __restrict__ in, T* __restrict__ out, uint64_t A, uint64_t d, uint64_t Bm, uint64_t C) {
  uint64_t n = A * Bm * C;
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t c = e % C, t = e / C, b = t % Bm, a = t / Bm;
    uint64_t base = ((a * d * Bm + b) * d) * C + c;
    uint64_t step = (Bm * d + 1) * C;
    T acc = zero_of<T>();
    for (uint64_t x = 0; x < d; ++x) acc = Arith<T>::add(acc, in[base + x * step]);
    out[e] = acc;
  }
}

template <int ES>
__global__ void diag_kernel(const typename Bits<ES>::T* __restrict__ in, typename Bits<ES>::T* __restrict__ out, uint64_t A,
                            uint64_t d, uint64_t Bm, uint64_t C) {
  uint64_t n = A * d * Bm * C;
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t c = e % C, t = e / C, b = t % Bm;
    t /= Bm;
    uint64_t x = t % d, a = t / d;
    out[e] = in[(((a * d + x) * Bm + b) * d + x) * C + c];
  }
}

void launch_trace(int dt, size_t A, size_t d, size_t Bm, size_t C, const void* in, void* out, hipStream_t s) {
  dispatch_dt(dt, [&](auto tag) {
    using T = decltype(tag);
    g_last_kind = K_TRACE_DIAG;
    LAUNCH(trace_kernel<T>, grid_for(A * Bm * C, 256), 256, 0, s, (const T*)in, (T*)out, (uint64_t)A, (uint64_t)d,
           (uint64_t)Bm, (uint64_t)C);
  });
}
void launch_diag(int es, size_t A, size_t d, size_t Bm, size_t C, const void* in, void* out, hipStream_t s) {
  dispatch_es(es, [&](auto esc) {
    constexpr int ES = decltype(esc)::value;
    using E = typename Bits<ES>::T;
    g_last_kind = K_TRACE_DIAG;
    LAUNCH(diag_kernel<ES>, grid_for(A * d * Bm * C, 256), 256, 0, s, (const E*)in, (E*)out, (uint64_t)A, (uint64_t)d,
           (uint64_t)Bm, (uint64_t)C);
  });
}

// ------------------------------------------------------------------ gather ---
template <int ES>
__global__ void gather_kernel(GatherArgs a, uint64_t g0, uint64_t nb, const typename Bits<ES>::T* __restrict__ in,
                              typename Bits<ES>::T* __restrict__ out) {
  uint64_t n = nb * a.inner;
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t b = e / a.inner, r = e - b * a.inner, g = g0 + b;
    int64_t off = 0;
    for (int q = 0; q < a.n_sel; ++q) off += (int64_t)((g / a.sel_div[q]) % a.sel_mod[q]) * a.sel_str[q];
    for (int d = a.n_ax - 1; d >= 0; --d) {
      uint64_t dim = a.dims[d], q = r / dim;
      off += (int64_t)(r - q * dim) * a.str[d];
      r = q;
    }
    out[e] = in[off];
  }
}

void launch_gather(int es, const GatherArgs& a, uint64_t g0, size_t nb, const void* in, void* out, hipStream_t s) {
  dispatch_es(es, [&](auto esc) {
    constexpr int ES = decltype(esc)::value;
    using E = typename Bits<ES>::T;
    g_last_kind = K_GATHER;
    LAUNCH(gather_kernel<ES>, grid_for(nb * a.inner, 256), 256, 0, s, a, g0, (uint64_t)nb, (const E*)in, (E*)out);
  });
}

// --------------------------------------------------------------- reductions --
template <class T> struct Wide;
template <> struct Wide<float> { using W = double; static __device__ W up(float x) { return x; } static __device__ float down(double x) { return (float)x; } };
template <> struct Wide<double> { using W = double; static __device__ W up(double x) { return x; } static __device__ double down(double x) { return x; } };
template <> struct Wide<float2> { using W = double2; static __device__ W up(float2 x) { return make_double2(x.x, x.y); } static __device__ float2 down(double2 x) { return make_float2((float)x.x, (float)x.y); } };
template <> struct Wide<double2> { using W = double2; static __device__ W up(double2 x) { return x; } static __device__ double2 down(double2 x) { return x; } };

// Deterministic: every accumulator element adds the batch's slices of its
// block in slice order.
template <class T>
__global__ void reduce_f64_kernel(const T* __restrict__ part, typename Wide<T>::W* __restrict__ acc, uint64_t g0, uint64_t nb,
                                  uint64_t n_int, uint64_t numel) {
  using W = typename Wide<T>::W;
  uint64_t blk0 = g0 / n_int, blk1 = (g0 + nb - 1) / n_int;
  uint64_t n = (blk1 - blk0 + 1) * numel;
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t bl = blk0 + t / numel, e = t % numel;
    uint64_t lo = bl * n_int > g0 ? bl * n_int : g0, hi = (bl + 1) * n_int < g0 + nb ? (bl + 1) * n_int : g0 + nb;
    W a = acc[bl * numel + e];
    for (uint64_t g = lo; g < hi; ++g) a = Arith<W>::add(a, Wide<T>::up(part[(g - g0) * numel + e]));
    acc[bl * numel + e] = a;
  }
}

void launch_reduce_f64(int dt, const void* part, void* acc64, uint64_t g0, size_t nb, size_t n_internal, size_t numel,
                       hipStream_t s) {
  dispatch_dt(dt, [&](auto tag) {
    using T = decltype(tag);
    uint64_t blocks = (g0 + nb - 1) / n_internal - g0 / n_internal + 1;
    g_last_kind = K_REDUCE;
    LAUNCH(reduce_f64_kernel<T>, grid_for(blocks * numel, 256), 256, 0, s, (const T*)part, (typename Wide<T>::W*)acc64, g0,
           (uint64_t)nb, (uint64_t)n_internal, (uint64_t)numel);
  });
}

template <class T, class S>
__global__ void scatter_kernel(ScatterArgs a, uint64_t blk0, uint64_t n_blk, uint64_t numel, const S* __restrict__ src,
                               T* __restrict__ out) {
  uint64_t n = n_blk * numel;
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t bl = blk0 + t / numel, e = t % numel;
    int64_t off = 0;
    for (int q = 0; q < a.n_blk; ++q) off += (int64_t)((bl / a.blk_div[q]) % a.blk_mod[q]) * a.blk_str[q];
    uint64_t r = e;
    for (int d = a.n_kept - 1; d >= 0; --d) {
      uint64_t dim = a.kept_dims[d], q = r / dim;
      off += (int64_t)(r - q * dim) * a.kept_str[d];
      r = q;
    }
    if constexpr (std::is_same<S, T>::value) out[off] = src[t];
    else out[off] = Wide<T>::down(src[t]);
  }
}

void launch_scatter(int dt, bool src_f64, const ScatterArgs& a, uint64_t blk0, size_t n_blk, size_t numel, const void* src,
                    void* out, hipStream_t s) {
  dispatch_dt(dt, [&](auto tag) {
    using T = decltype(tag);
    using W = typename Wide<T>::W;
    g_last_kind = K_SCATTER;
    if (src_f64)
      LAUNCH((scatter_kernel<T, W>), grid_for(n_blk * numel, 256), 256, 0, s, a, blk0, (uint64_t)n_blk, (uint64_t)numel,
             (const W*)src, (T*)out);
    else
      LAUNCH((scatter_kernel<T, T>), grid_for(n_blk * numel, 256), 256, 0, s, a, blk0, (uint64_t)n_blk, (uint64_t)numel,
             (const T*)src, (T*)out);
  });
}

template <class T>
__global__ void accum_kernel(const T* __restrict__ part, T* __restrict__ acc, uint64_t n, int first) {
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (uint64_t)gridDim.x * blockDim.x)
    acc[e] = first ? part[e] : Arith<T>::add(acc[e], part[e]);
}

void launch_accum(int dt, bool first, const void* part, void* acc, size_t numel, hipStream_t s) {
  dispatch_dt(dt, [&](auto tag) {
    using T = decltype(tag);
    g_last_kind = K_REDUCE;
    LAUNCH(accum_kernel<T>, grid_for(numel, 256), 256, 0, s, (const T*)part, (T*)acc, (uint64_t)numel, first ? 1 : 0);
  });
}

__global__ void fill_zero_kernel(uint32_t* p, uint64_t n) {
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (uint64_t)gridDim.x * blockDim.x) p[e] = 0u;
}

void launch_fill_zero(void* p, size_t bytes, hipStream_t s) {
  if (bytes % 4) throw Error("fill_zero needs a multiple of 4 bytes");
  LAUNCH(fill_zero_kernel, grid_for(bytes / 4, 256), 256, 0, s, (uint32_t*)p, (uint64_t)(bytes / 4));
}

// -------------------------------------------------------------------- strip --
// Power-of-two normalisation: multiplying by 2^-e is exact, so the mantissa
// times 2^exponent reproduces the unstripped value bit for bit (barring
// underflow); the reference divides by the exact max magnitude instead.
// Max magnitude (max(|re|, |im|) for complex elements, the max over every real
// component) as double bits, atomically raised. S: the component type. Components
// go through 16-byte loads (four in flight per thread) where x is 16-byte aligned,
// the tail (or the whole array otherwise) one component at a time; the max is
// taken in S (exact) and the workgroup's max raises *maxbits once.
template <class S>
__global__ __launch_bounds__(256) void maxabs_kernel(const S* __restrict__ x, uint64_t ns, unsigned long long* maxbits) {
  using V = typename std::conditional<sizeof(S) == 4, float4, double2>::type;
  constexpr uint64_t W = 16 / sizeof(S);
  S m = 0;
  auto take = [&](S v) {
    v = fabs(v);
    m = v > m ? v : m;  // NaN never wins; a NaN result stays unscaled
  };
  auto take_v = [&](const V& v) {
    if constexpr (W == 4) take(v.x), take(v.y), take(v.z), take(v.w);
    else take(v.x), take(v.y);
  };
  const uint64_t nv = (uintptr_t)x % 16 == 0 ? ns / W : 0;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x, t0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const V* xv = reinterpret_cast<const V*>(x);
  uint64_t i = t0;
  for (; i + 3 * stride < nv; i += 4 * stride) {
    const V a = xv[i], b = xv[i + stride], c = xv[i + 2 * stride], d = xv[i + 3 * stride];
    take_v(a), take_v(b), take_v(c), take_v(d);
  }
  for (; i < nv; i += stride) take_v(xv[i]);
  for (uint64_t e = nv * W + t0; e < ns; e += stride) take(x[e]);
  for (int o = 32; o > 0; o >>= 1) {
    const S t = __shfl_down(m, o, 64);
    m = t > m ? t : m;
  }
  __shared__ S wmax[4];
  if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < (int)(blockDim.x / 64); ++w) m = wmax[w] > m ? wmax[w] : m;
    if (m > 0) atomicMax(maxbits, (unsigned long long)__double_as_longlong((double)m));
  }
}

// scaled_max (optional): receives the max after the rescale (max * 2^-e, or the
// max itself when nothing is rescaled) as double bits, for a later product's
// FP16 operand scale. pend: record e there for x's reader instead of scaling x
// (scaled_max then receives the max itself).
template <class T>
__global__ void scale_pow2_kernel(T* __restrict__ x, uint64_t n, const unsigned long long* maxbits, long long* expo, int keep,
                                  unsigned long long* scaled_max, long long* pend) {
  const double m = __longlong_as_double((long long)*maxbits);
  const bool first = blockIdx.x == 0 && threadIdx.x == 0;
  int e = 0;
  if (m > 0.0 && isfinite(m)) {
    frexp(m, &e);
    if (keep > 0 && e >= -keep && e <= keep) e = 0;
  }
  if (pend) {
    if (first) {
      *pend = e, *expo += e;
      if (scaled_max) *scaled_max = (unsigned long long)__double_as_longlong(m);
    }
    return;
  }
  if (first && scaled_max) *scaled_max = (unsigned long long)__double_as_longlong(e ? ldexp(m, -e) : m);
  if (e == 0) return;
  if (first) *expo += e;
  const typename Arith<T>::R s = (typename Arith<T>::R)ldexp(1.0, -e);
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    x[i] = Arith<T>::scale(x[i], s);
}

// Exact normalisation of a stripped run's last step (src/compiled.rs:412-430):
// m = max over elements of Scalar::abs (f64), then x *= T::from_real(1 / m) and
// the log10 exponent gains log10(m); nothing changes when m is 0 or not finite.
template <class T>
__global__ void max_abs_exact_kernel(const T* __restrict__ x, uint64_t n, unsigned long long* maxbits) {
  double m = 0.0;
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (uint64_t)gridDim.x * blockDim.x) {
    double v = Arith<T>::abs(x[e]);
    m = v > m ? v : m;
  }
  for (int o = 32; o > 0; o >>= 1) {
    double t = __shfl_down(m, o, 64);
    m = t > m ? t : m;
  }
  if ((threadIdx.x & 63) == 0 && m > 0.0) atomicMax(maxbits, (unsigned long long)__double_as_longlong(m));
}

template <class T>
__global__ void scale_exact_kernel(T* __restrict__ x, uint64_t n, const unsigned long long* maxbits, double* log10_exp) {
  const double m = __longlong_as_double((long long)*maxbits);
  if (!(m > 0.0) || !isfinite(m)) return;
  if (blockIdx.x == 0 && threadIdx.x == 0) *log10_exp += log10(m);
  const typename Arith<T>::R inv = (typename Arith<T>::R)(1.0 / m);
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    x[i] = Arith<T>::scale(x[i], inv);
}

void launch_strip_exact(int dt, void* x, size_t numel, uint64_t* maxbits, double* log10_exp, hipStream_t s) {
  dispatch_dt(dt, [&](auto tag) {
    using T = decltype(tag);
    g_last_kind = K_STRIP;
    LAUNCH(fill_zero_kernel, dim3(1), dim3(64), 0, s, (uint32_t*)maxbits, (uint64_t)2);
    LAUNCH(max_abs_exact_kernel<T>, grid_for(numel, 256, 2048), 256, 0, s, (const T*)x, (uint64_t)numel,
           (unsigned long long*)maxbits);
    LAUNCH(scale_exact_kernel<T>, grid_for(numel, 256), 256, 0, s, (T*)x, (uint64_t)numel,
           (const unsigned long long*)maxbits, log10_exp);
  });
}

void launch_maxabs(int dt, const void* x, size_t numel, uint64_t* maxbits, hipStream_t s) {
  dispatch_dt(dt, [&](auto tag) {
    using T = decltype(tag);
    g_last_kind = K_STRIP;
    using S = typename std::conditional<std::is_same<T, float>::value || std::is_same<T, float2>::value, float, double>::type;
    const uint64_t ns = (uint64_t)numel * (dtype_complex(dt) ? 2 : 1);
    // a quarter of the 16-byte vectors per thread and pass, at most 1024 workgroups
    LAUNCH(maxabs_kernel<S>, grid_for((ns * sizeof(S) / 16 + 3) / 4, 256, 1024), 256, 0, s, (const S*)x, ns,
           (unsigned long long*)maxbits);
  });
}
void launch_scale_pow2(int dt, void* x, size_t numel, const uint64_t* maxbits, int64_t* expo, hipStream_t s, int keep,
                       uint64_t* scaled_max, int64_t* pend) {
  dispatch_dt(dt, [&](auto tag) {
    using T = decltype(tag);
    g_last_kind = K_STRIP;
    // a grid-stride grid of at most 2048 workgroups: most calls find the max in range
    // and return at once, so the launch costs little more than its first wave (one
    // workgroup when the exponent is only recorded)
    LAUNCH(scale_pow2_kernel<T>, pend ? 1u : grid_for(numel, 256, 2048), 256, 0, s, (T*)x, (uint64_t)numel,
           (const unsigned long long*)maxbits, (long long*)expo, keep, (unsigned long long*)scaled_max, (long long*)pend);
  });
}

// ------------------------------------------------------------ copies/adjoint --
template <int ES>
__global__ void copy_kernel(const typename Bits<ES>::T* __restrict__ in, typename Bits<ES>::T* __restrict__ out, uint64_t n) {
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (uint64_t)gridDim.x * blockDim.x) out[e] = in[e];
}

void launch_copy(size_t bytes, const void* in, void* out, hipStream_t s) {
  if (bytes % 4) throw Error("copy needs a multiple of 4 bytes");
  g_last_kind = K_COPY;
  LAUNCH(copy_kernel<4>, grid_for(bytes / 4, 256), 256, 0, s, (const uint32_t*)in, (uint32_t*)out, (uint64_t)(bytes / 4));
}

template <int ES>
__global__ void broadcast_axis_kernel(const typename Bits<ES>::T* __restrict__ in, typename Bits<ES>::T* __restrict__ out,
                                      uint64_t outer, uint64_t d, uint64_t inner) {
  uint64_t n = outer * d * inner;
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t q = e % inner, o = e / (d * inner);
    out[e] = in[o * inner + q];
  }
}

void launch_broadcast_axis(int es, size_t outer, size_t d, size_t inner, const void* in, void* out, hipStream_t s) {
  dispatch_es(es, [&](auto esc) {
    constexpr int ES = decltype(esc)::value;
    using E = typename Bits<ES>::T;
    LAUNCH(broadcast_axis_kernel<ES>, grid_for(outer * d * inner, 256), 256, 0, s, (const E*)in, (E*)out, (uint64_t)outer,
           (uint64_t)d, (uint64_t)inner);
  });
}

template <int ES>
__global__ void diag_scatter_kernel(const typename Bits<ES>::T* __restrict__ g, typename Bits<ES>::T* __restrict__ out, int trace,
                                    uint64_t A, uint64_t d, uint64_t Bm, uint64_t C) {
  uint64_t n = A * d * Bm * C;
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t c = e % C, t = e / C, b = t % Bm;
    t /= Bm;
    uint64_t x = t % d, a = t / d;
    uint64_t src = trace ? (a * Bm + b) * C + c : e;
    out[(((a * d + x) * Bm + b) * d + x) * C + c] = g[src];
  }
}

void launch_diag_scatter(int es, bool trace, size_t A, size_t d, size_t Bm, size_t C, const void* g, void* out, hipStream_t s) {
  dispatch_es(es, [&](auto esc) {
    constexpr int ES = decltype(esc)::value;
    using E = typename Bits<ES>::T;
    LAUNCH(diag_scatter_kernel<ES>, grid_for(A * d * Bm * C, 256), 256, 0, s, (const E*)g, (E*)out, trace ? 1 : 0, (uint64_t)A,
           (uint64_t)d, (uint64_t)Bm, (uint64_t)C);
  });
}

}  // namespace atn
