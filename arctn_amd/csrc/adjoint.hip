// Reverse-mode differentiation of an unsliced plan on the GPU, as two calls:
// forward (value, checkpoints kept on the device) and backward (every input's
// gradient). The reference has no gradient code; its stage hook
// (src/compiled.rs:7-10, 60-63, 143-153, 206-209) is what a tape drives, and
// the per-step adjoint here is the one of stage_step_adjoint:
//   C = A2 . B2 with A2 = perm_a(sum_a(left)), B2 = perm_b(sum_b(right)):
//   dA2 = dC . B2^H,  dB2 = A2^H . dC  (dL/d(conj) convention, as PyTorch),
//   then the inverse permutations and broadcasts over the summed legs.
//
// Schedule. Steps run level by level of the contraction tree (a step's level is
// one more than its operands'). Small steps of one level share one launch of
// grad_group_kernel: one thread per output element, operands read in place
// through offset tables of the permuted views (no permute launches) and, for
// the adjoints, the gradient scattered straight into the operand's layout
// through the same tables. Other steps (large, batched or with summed legs)
// run on their own: MFMA GEMMs for the products, explicit conjugate
// transposes for the adjoint products.
//
// Checkpointing: with checkpoint_every = E > 0 the level-ordered steps form
// segments of E; a slot produced and consumed inside one segment is freed after
// its forward use and recomputed before the segment's adjoints, every other
// slot is a checkpoint kept from the forward pass. E = 0 keeps every slot.
// Device memory is planned offline: every value and gradient buffer has a
// lifetime on the op timeline and buffers are packed first-fit.
#include <algorithm>
#include <climits>
#include <cstring>

#include "executor.hpp"

namespace atn {

namespace {

// One small problem of a grouped launch. kind 0: C[i][j] = sum_p A(i,p) B(p,j);
// kind 1: A-gradient, dst(i,p) = sum_j G[i][j] conj(B(p,j)); kind 2: B-gradient,
// dst(p,j) = sum_i conj(A(i,p)) G[i][j]. A(i,p) = x[a_off + T[arow+i] + T[acol+p]]
// (likewise B with brow/bcol); G and C are contiguous [m][n]; a kind 1 (2)
// result goes to dst[c_off + T[arow+i] + T[acol+p]] (T[brow+p] + T[bcol+j]).
struct GDesc {
  uint32_t kind, a_base, b_base, c_base;
  uint64_t a_off, b_off, c_off;
  int32_t m, k, n, pad;
  int32_t arow, acol, brow, bcol;
  uint64_t block0;
};

template <class T> struct GC;
template <> struct GC<float> {
  static __device__ float zero() { return 0.f; }
  static __device__ float conj(float a) { return a; }
  static __device__ float mac(float c, float a, float b) { return c + a * b; }
};
template <> struct GC<double> {
  static __device__ double zero() { return 0.0; }
  static __device__ double conj(double a) { return a; }
  static __device__ double mac(double c, double a, double b) { return c + a * b; }
};
template <> struct GC<float2> {
  static __device__ float2 zero() { return make_float2(0.f, 0.f); }
  static __device__ float2 conj(float2 a) { return make_float2(a.x, -a.y); }
  static __device__ float2 mac(float2 c, float2 a, float2 b) {
    return make_float2(c.x + (a.x * b.x - a.y * b.y), c.y + (a.x * b.y + a.y * b.x));
  }
};
template <> struct GC<double2> {
  static __device__ double2 zero() { return make_double2(0.0, 0.0); }
  static __device__ double2 conj(double2 a) { return make_double2(a.x, -a.y); }
  static __device__ double2 mac(double2 c, double2 a, double2 b) {
    return make_double2(c.x + (a.x * b.x - a.y * b.y), c.y + (a.x * b.y + a.y * b.x));
  }
};

__device__ inline const GDesc& find_desc(const GDesc* d, int n, uint64_t vb) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (d[mid].block0 <= vb) lo = mid;
    else hi = mid - 1;
  }
  return d[lo];
}

template <class T>
__global__ __launch_bounds__(256) void grad_group_kernel(const GDesc* __restrict__ d, int n, uint64_t total,
                                                         char* const* __restrict__ bases, const int32_t* __restrict__ tab) {
  for (uint64_t vb = blockIdx.x; vb < total; vb += gridDim.x) {
    const GDesc& p = find_desc(d, n, vb);
    const uint64_t e = (vb - p.block0) * 256ull + threadIdx.x;
    const T* A = reinterpret_cast<const T*>(bases[p.a_base]) + p.a_off;
    const T* B = reinterpret_cast<const T*>(bases[p.b_base]) + p.b_off;
    T* C = reinterpret_cast<T*>(bases[p.c_base]) + p.c_off;
    T acc = GC<T>::zero();
    if (p.kind == 0) {
      if (e >= (uint64_t)p.m * p.n) continue;
      const int i = (int)(e / p.n), j = (int)(e % p.n);
      const T* Ai = A + tab[p.arow + i];
      const T* Bj = B + tab[p.bcol + j];
      for (int q = 0; q < p.k; ++q) acc = GC<T>::mac(acc, Ai[tab[p.acol + q]], Bj[tab[p.brow + q]]);
      C[e] = acc;
    } else if (p.kind == 1) {  // A = G [m][n], B = right value
      if (e >= (uint64_t)p.m * p.k) continue;
      const int i = (int)(e / p.k), q = (int)(e % p.k);
      const T* Gi = A + (uint64_t)i * p.n;
      const T* Bq = B + tab[p.brow + q];
      for (int j = 0; j < p.n; ++j) acc = GC<T>::mac(acc, Gi[j], GC<T>::conj(Bq[tab[p.bcol + j]]));
      C[tab[p.arow + i] + tab[p.acol + q]] = acc;
    } else {  // A = left value, B = G [m][n]
      if (e >= (uint64_t)p.k * p.n) continue;
      const int q = (int)(e / p.n), j = (int)(e % p.n);
      const T* Aq = A + tab[p.acol + q];
      for (int i = 0; i < p.m; ++i) acc = GC<T>::mac(acc, GC<T>::conj(Aq[tab[p.arow + i]]), B[(uint64_t)i * p.n + j]);
      C[tab[p.brow + q] + tab[p.bcol + j]] = acc;
    }
  }
}

template <class T>
__global__ void conj_kernel(T* x, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    x[i] = GC<T>::conj(x[i]);
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

constexpr unsigned GRID_CAP = 4096;

// Offsets (elements) of every multi-index over `axes` of a row-major tensor of
// `shape`, last listed axis fastest: the rows or columns of a permuted view.
std::vector<int32_t> view_offsets(const std::vector<size_t>& shape, const std::vector<size_t>& axes) {
  std::vector<size_t> st(shape.size(), 1);
  for (size_t a = shape.size(); a-- > 1;) st[a - 1] = st[a] * shape[a];
  std::vector<int32_t> out{0};
  for (size_t ax : axes) {
    std::vector<int32_t> nx;
    nx.reserve(out.size() * shape[ax]);
    for (int32_t o : out)
      for (size_t v = 0; v < shape[ax]; ++v) nx.push_back((int32_t)(o + v * st[ax]));
    out.swap(nx);
  }
  return out;
}

}  // namespace

// Base table of a run: [0] workspace, [1, 1 + n) inputs, [1 + n, 1 + 2n) input
// gradients, then the output and the output gradient.
struct GradRunner::Loc {
  uint32_t base = 0;
  int buf = -1;  // workspace buffer (base 0), or -1
};

GradRunner::GradRunner(std::shared_ptr<const Plan> plan, int dt, int prec, size_t checkpoint_every)
    : P_(std::move(plan)), dt_(dt), prec_(prec) {
  const Plan& P = *P_;
  if (P.n_slices != 1) throw Error("differentiation needs an unsliced plan");
  if (prec == C3M && dt != C64) throw Error("precision c3m applies to complex64 only");
  if (split_precision(prec) && !(dt == F32 || dt == C64))
    throw Error("split precision applies to float32 / complex64 only");
  es_ = (int)dtype_size(dt);
  const size_t n_in = P.cp.n_inputs, n_st = P.cp.steps.size(), nslot = n_in + n_st;
  n_in_ = n_in;
  // levels and level order
  std::vector<int> level(nslot, 0);
  for (size_t s = 0; s < n_st; ++s) {
    const PairStep& st = P.cp.steps[s];
    level[n_in + s] = 1 + std::max(level[st.li], level[st.ri]);
  }
  std::vector<size_t> order(n_st);
  for (size_t s = 0; s < n_st; ++s) order[s] = s;
  std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) { return level[n_in + x] < level[n_in + y]; });
  std::vector<size_t> pos(n_st), seg(n_st, 0);
  for (size_t q = 0; q < n_st; ++q) pos[order[q]] = q;
  const bool ckpt = checkpoint_every > 0;
  for (size_t s = 0; s < n_st; ++s) seg[s] = ckpt ? pos[s] / checkpoint_every : 0;
  const size_t n_seg = n_st ? seg[order.back()] + 1 : 0;
  // consumer step of every slot (-1: finalize)
  std::vector<long> cons(nslot, -1);
  for (size_t s = 0; s < n_st; ++s) cons[P.cp.steps[s].li] = (long)s, cons[P.cp.steps[s].ri] = (long)s;
  auto internal = [&](size_t x) {  // recomputed in the backward pass
    if (!ckpt || x < n_in) return false;
    const long c = cons[x];
    return c >= 0 && seg[(size_t)c] == seg[x - n_in];
  };
  auto groupable = [&](const PairStep& st) {
    return st.bt == 1 && st.sum_a.empty() && st.sum_b.empty() && st.k <= 4096 && st.m <= 4096 && st.n <= 4096 &&
           (double)st.m * st.k * st.n <= (double)(1 << 20);
  };
  // ---- locations of values and gradients ----
  std::vector<std::vector<size_t>> shape(nslot);
  for (size_t i = 0; i < n_in; ++i) shape[i] = P.cp.preps[i].shape;
  for (size_t s = 0; s < n_st; ++s) shape[n_in + s] = P.cp.steps[s].result_shape;
  auto bytes_of = [&](size_t x) { return std::max<size_t>(prod(shape[x]) * es_, 256); };
  val_f_.assign(nslot, Loc{}), val_r_.assign(nslot, Loc{}), grad_.assign(nslot, Loc{});
  auto& val_f = val_f_;
  auto& val_r = val_r_;
  auto& grad = grad_;
  const uint32_t B_GIN = 1 + (uint32_t)n_in;
  for (size_t i = 0; i < n_in; ++i) {
    if (P.cp.preps[i].identity()) {
      val_f[i].base = 1 + (uint32_t)i;
      grad[i].base = B_GIN + (uint32_t)i;
    } else {
      val_f[i].buf = newbuf(bytes_of(i));
      grad[i].buf = newbuf(bytes_of(i));
    }
    val_r[i] = val_f[i];
  }
  for (size_t x = n_in; x < nslot; ++x) {
    val_f[x].buf = newbuf(bytes_of(x));
    val_r[x] = internal(x) ? Loc{0, newbuf(bytes_of(x))} : val_f[x];
    grad[x].buf = newbuf(bytes_of(x));
  }
  const size_t fin = P.cp.fin.slot;
  // ---- op timeline ----
  int t = 0;
  auto touch = [&](const Loc& l) {
    if (l.buf < 0) return;
    bufs_[l.buf].first = std::min(bufs_[l.buf].first, t);
    bufs_[l.buf].last = std::max(bufs_[l.buf].last, t);
  };
  // Forward prep ops
  for (size_t i = 0; i < n_in; ++i) {
    if (P.cp.preps[i].identity()) continue;
    touch(val_f[i]);
    ops_.push_back(GOp{GOp::PREP, t++, i});
  }
  // Emit the steps of `steps` (level order) as grouped launches per level plus single ops.
  // vals_id: 0 reads forward values, 1 the backward pass's (recomputed or checkpointed) ones
  auto emit_steps = [&](const std::vector<size_t>& steps, int kind, int vals_id) {
    const std::vector<Loc>& vals = vals_id ? val_r : val_f;
    size_t q = 0;
    while (q < steps.size()) {
      const int lv = level[n_in + steps[q]];
      GOp grp{kind == 0 ? GOp::FWD_GROUP : GOp::ADJ_GROUP, t, 0};
      std::vector<size_t> singles;
      for (; q < steps.size() && level[n_in + steps[q]] == lv; ++q) {
        const size_t s = steps[q];
        const PairStep& st = P.cp.steps[s];
        const size_t x = n_in + s;
        if (kind == 0) {
          touch(vals[st.li]), touch(vals[st.ri]), touch(vals[x]);
        } else {
          touch(vals[st.li]), touch(vals[st.ri]), touch(grad[x]), touch(grad[st.li]), touch(grad[st.ri]);
        }
        if (groupable(st)) grp.items.push_back(s);
        else singles.push_back(s);
      }
      if (!grp.items.empty()) {
        grp.vals = vals_id;
        ops_.push_back(std::move(grp));
      }
      for (size_t s : singles) {
        GOp o{kind == 0 ? GOp::FWD_ONE : GOp::ADJ_ONE, t, s};
        o.vals = vals_id;
        ops_.push_back(std::move(o));
      }
      ++t;
    }
  };
  emit_steps(order, 0, 0);
  // finalize forward
  touch(val_f[fin]);
  ops_.push_back(GOp{GOp::FIN, t++, 0});
  fwd_end_ = ops_.size();
  // backward: finalize adjoint
  touch(grad[fin]);
  ops_.push_back(GOp{GOp::FIN_ADJ, t++, 0});
  for (size_t g = n_seg; g-- > 0;) {
    std::vector<size_t> mine, redo;
    for (size_t s : order)
      if (seg[s] == g) {
        mine.push_back(s);
        if (internal(n_in + s)) redo.push_back(s);
      }
    if (!redo.empty()) emit_steps(redo, 0, 1);
    std::vector<size_t> rev(mine.rbegin(), mine.rend());
    emit_steps(rev, 1, 1);
  }
  for (size_t i = 0; i < n_in; ++i) {
    if (P.cp.preps[i].identity()) continue;
    touch(grad[i]);
    ops_.push_back(GOp{GOp::PREP_ADJ, t++, i});
  }
  // (the forward copies of checkpoints stay alive until their adjoint reads them: val_r aliases val_f)
  // ---- placement (first fit over lifetimes) ----
  std::vector<int> idx(bufs_.size());
  for (size_t i = 0; i < idx.size(); ++i) idx[i] = (int)i;
  std::sort(idx.begin(), idx.end(), [&](int x, int y) {
    return bufs_[x].first != bufs_[y].first ? bufs_[x].first < bufs_[y].first : bufs_[x].bytes > bufs_[y].bytes;
  });
  offs_.assign(bufs_.size(), 0);
  struct Live {
    size_t off, end;
    int last;
  };
  std::vector<Live> live;
  size_t total = 256;
  for (int id : idx) {
    const Buf& b = bufs_[id];
    if (b.first == INT_MAX) continue;  // never used
    live.erase(std::remove_if(live.begin(), live.end(), [&](const Live& l) { return l.last < b.first; }), live.end());
    std::sort(live.begin(), live.end(), [](const Live& x, const Live& y) { return x.off < y.off; });
    size_t need = (b.bytes + 255) / 256 * 256, p = 0;
    for (auto& l : live) {
      if (p + need <= l.off) break;
      p = std::max(p, l.end);
    }
    offs_[id] = p;
    live.push_back(Live{p, p + need, b.last});
    total = std::max(total, p + need);
  }
  ws_bytes_ = total;
  // ---- grouped descriptors and offset tables ----
  std::vector<GDesc> descs_;
  for (auto& o : ops_) {
    if (o.kind != GOp::FWD_GROUP && o.kind != GOp::ADJ_GROUP) continue;
    uint64_t blocks = 0;
    o.desc_first = descs_.size();
    for (size_t s : o.items) {
      const PairStep& st = P.cp.steps[s];
      const size_t x = n_in + s;
      // A2 = left permuted to [batch, free_a, con], B2 = right to [batch, con, free_b]
      // (bt == 1: every batch axis has dimension 1 and may join either side)
      size_t nb = 0, nc = 0;
      for (Leg l : st.left_legs) {
        const bool in_r = std::count(st.right_legs.begin(), st.right_legs.end(), l) > 0;
        const bool in_c = std::count(st.result_legs.begin(), st.result_legs.end(), l) > 0;
        nb += in_r && in_c, nc += in_r && !in_c;
      }
      const size_t fa = st.a_perm.size() - nc, cb = nb + nc;
      std::vector<size_t> a_rows(st.a_perm.begin(), st.a_perm.begin() + fa), a_cols(st.a_perm.begin() + fa, st.a_perm.end());
      std::vector<size_t> b_rows(st.b_perm.begin(), st.b_perm.begin() + cb), b_cols(st.b_perm.begin() + cb, st.b_perm.end());
      auto put = [&](const std::vector<int32_t>& v) {
        const int32_t at = (int32_t)tab_.size();
        tab_.insert(tab_.end(), v.begin(), v.end());
        return at;
      };
      GDesc d{};
      d.m = (int32_t)st.m, d.k = (int32_t)st.k, d.n = (int32_t)st.n;
      d.arow = put(view_offsets(st.left_shape, a_rows));
      d.acol = put(view_offsets(st.left_shape, a_cols));
      d.brow = put(view_offsets(st.right_shape, b_rows));
      d.bcol = put(view_offsets(st.right_shape, b_cols));
      const std::vector<Loc>& V = o.vals ? val_r_ : val_f_;
      const Loc& L = V[st.li];
      const Loc& R = V[st.ri];
      auto set = [&](uint32_t& base, uint64_t& off, const Loc& l) { base = l.base, off = l.buf >= 0 ? offs_[l.buf] / es_ : 0; };
      if (o.kind == GOp::FWD_GROUP) {
        d.kind = 0;
        set(d.a_base, d.a_off, L), set(d.b_base, d.b_off, R), set(d.c_base, d.c_off, V[x]);
        d.block0 = blocks;
        blocks += (st.m * st.n + 255) / 256;
        descs_.push_back(d);
      } else {
        GDesc ga = d, gb = d;
        ga.kind = 1;
        set(ga.a_base, ga.a_off, grad_[x]), set(ga.b_base, ga.b_off, R), set(ga.c_base, ga.c_off, grad_[st.li]);
        ga.block0 = blocks;
        blocks += (st.m * st.k + 255) / 256;
        descs_.push_back(ga);
        gb.kind = 2;
        set(gb.a_base, gb.a_off, L), set(gb.b_base, gb.b_off, grad_[x]), set(gb.c_base, gb.c_off, grad_[st.ri]);
        gb.block0 = blocks;
        blocks += (st.k * st.n + 255) / 256;
        descs_.push_back(gb);
      }
    }
    o.n_desc = descs_.size() - o.desc_first;
    o.blocks = blocks;
  }
  n_bases_ = 3 + 2 * n_in;
  // device memory: workspace, descriptors, tables, base table
  ATN_HIP(hipMalloc(&ws_, ws_bytes_));
  if (!descs_.empty()) {
    ATN_HIP(hipMalloc(&descs_dev_, descs_.size() * sizeof(GDesc)));
    ATN_HIP(hipMemcpy(descs_dev_, descs_.data(), descs_.size() * sizeof(GDesc), hipMemcpyHostToDevice));
  }
  if (!tab_.empty()) {
    ATN_HIP(hipMalloc((void**)&tab_dev_, tab_.size() * sizeof(int32_t)));
    ATN_HIP(hipMemcpy(tab_dev_, tab_.data(), tab_.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  }
  ATN_HIP(hipMalloc((void**)&bases_dev_, n_bases_ * sizeof(char*)));
  bases_.assign(n_bases_, nullptr);
  ATN_HIP(hipHostMalloc((void**)&bases_host_, n_bases_ * sizeof(char*)));
}

GradRunner::~GradRunner() {
  if (ws_) (void)hipFree(ws_);
  if (descs_dev_) (void)hipFree(descs_dev_);
  if (tab_dev_) (void)hipFree(tab_dev_);
  if (bases_dev_) (void)hipFree(bases_dev_);
  if (pool_.base) (void)hipFree(pool_.base);
  if (bases_host_) (void)hipHostFree(bases_host_);
}

int GradRunner::newbuf(size_t bytes) {
  bufs_.push_back(Buf{bytes, INT_MAX, -1});
  return (int)bufs_.size() - 1;
}

void GradRunner::set_bases(const void* const* inputs, void* const* grad_inputs, void* out, const void* grad_out,
                           hipStream_t s) {
  std::vector<char*> b(n_bases_, nullptr);
  b[0] = ws_;
  for (size_t i = 0; i < n_in_; ++i) {
    b[1 + i] = const_cast<char*>(static_cast<const char*>(inputs[i]));
    if (grad_inputs) b[1 + n_in_ + i] = static_cast<char*>(grad_inputs[i]);
  }
  b[1 + 2 * n_in_] = static_cast<char*>(out);
  b[2 + 2 * n_in_] = const_cast<char*>(static_cast<const char*>(grad_out));
  bases_ = b;
  // the staging buffer is rewritten only after the previous upload has completed
  ATN_HIP(hipStreamSynchronize(s));
  std::memcpy(bases_host_, b.data(), n_bases_ * sizeof(char*));
  ATN_HIP(hipMemcpyAsync(bases_dev_, bases_host_, n_bases_ * sizeof(char*), hipMemcpyHostToDevice, s));
}

void* GradRunner::at(const Loc& l) const {
  return l.buf >= 0 ? static_cast<void*>(ws_ + offs_[l.buf]) : static_cast<void*>(bases_[l.base]);
}

void GradRunner::run_ops(size_t begin, size_t end, hipStream_t s) {
  const Plan& P = *P_;
  const int dt = dt_, es = es_;
  for (size_t q = begin; q < end; ++q) {
    const GOp& o = ops_[q];
    Scratch sc(s, &pool_);
    switch (o.kind) {
      case GOp::PREP: stage_prep_forward(P, o.index, dt, bases_[1 + o.index], at(val_f_[o.index]), s); break;
      case GOp::PREP_ADJ: stage_prep_adjoint(P, o.index, dt, at(grad_[o.index]), bases_[1 + n_in_ + o.index], s); break;
      case GOp::FIN: stage_finalize_forward(P, dt, at(val_f_[P.cp.fin.slot]), bases_[1 + 2 * n_in_], s); break;
      case GOp::FIN_ADJ: stage_finalize_adjoint(P, dt, bases_[2 + 2 * n_in_], at(grad_[P.cp.fin.slot]), s); break;
      case GOp::FWD_GROUP:
      case GOp::ADJ_GROUP: {
        const dim3 grid((unsigned)std::min<uint64_t>(o.blocks, GRID_CAP));
        gdispatch(dt, [&](auto tag) {
          using T = decltype(tag);
          hipLaunchKernelGGL(grad_group_kernel<T>, grid, dim3(256), 0, s,
                             static_cast<const GDesc*>(descs_dev_) + o.desc_first, (int)o.n_desc, o.blocks, bases_dev_,
                             tab_dev_);
        });
        ATN_HIP(hipGetLastError());
        ++g_launch_count;
        break;
      }
      case GOp::FWD_ONE: {
        const PairStep& st = P.cp.steps[o.index];
        const std::vector<Loc>& V = o.vals ? val_r_ : val_f_;
        const void* a = stage_operand(sc, dt, at(V[st.li]), st.left_shape, st.sum_a, st.a_perm);
        const void* b = stage_operand(sc, dt, at(V[st.ri]), st.right_shape, st.sum_b, st.b_perm);
        GemmArgs g{};
        g.a = a, g.b = b, g.c = at(V[n_in_ + o.index]);
        g.m = st.m, g.k = st.k, g.n = st.n;
        g.lda = g.k, g.ldb = g.n, g.ldc = g.n;
        g.nbo = 1, g.nbi = st.bt;
        g.sai = g.m * g.k, g.sbi = g.k * g.n, g.sci = g.m * g.n;
        g.scratch_bytes = gemm_scratch_bytes(dt, prec_, GEMM_AUTO, g);
        g.scratch = sc.get(g.scratch_bytes);
        launch_gemm(dt, prec_, GEMM_AUTO, g, s);
        break;
      }
      case GOp::ADJ_ONE: {
        // dA2 = G . conj(B2)^T and dB2 = conj(A2)^T . G as MFMA GEMMs over explicit
        // conjugate transposes, then the inverse permutes / broadcasts
        const PairStep& st = P.cp.steps[o.index];
        const std::vector<Loc>& V = o.vals ? val_r_ : val_f_;
        const int64_t bt = st.bt, m = st.m, k = st.k, n = st.n;
        const void* a2 = stage_operand(sc, dt, at(V[st.li]), st.left_shape, st.sum_a, st.a_perm);
        const void* b2 = stage_operand(sc, dt, at(V[st.ri]), st.right_shape, st.sum_b, st.b_perm);
        const void* gc = at(grad_[n_in_ + o.index]);
        auto conj_t = [&](const void* x, int64_t r, int64_t c) {  // [bt][r][c] -> conj [bt][c][r]
          void* y = sc.get(bt * r * c * es);
          PermPlan pp = plan_permute({(size_t)bt, (size_t)r, (size_t)c}, {0, 2, 1}, 1, es);
          if (pp.kind == 0) launch_copy(bt * r * c * es, x, y, s);
          else launch_permute(pp, x, y, s);
          if (dtype_complex(dt))
            gdispatch(dt, [&](auto tag) {
              using T = decltype(tag);
              const uint64_t ne = (uint64_t)(bt * r * c);
              hipLaunchKernelGGL(conj_kernel<T>, dim3((unsigned)std::min<uint64_t>((ne + 255) / 256, 65536)), dim3(256), 0,
                                 s, (T*)y, ne);
              ATN_HIP(hipGetLastError());
              ++g_launch_count;
            });
          return y;
        };
        auto gemm = [&](const void* x, const void* y, void* z, int64_t mm, int64_t kk, int64_t nn) {
          GemmArgs g{};
          g.a = x, g.b = y, g.c = z;
          g.m = mm, g.k = kk, g.n = nn;
          g.lda = kk, g.ldb = nn, g.ldc = nn;
          g.nbo = 1, g.nbi = bt;
          g.sai = mm * kk, g.sbi = kk * nn, g.sci = mm * nn;
          g.scratch_bytes = gemm_scratch_bytes(dt, NATIVE, GEMM_AUTO, g);
          g.scratch = sc.get(g.scratch_bytes);
          launch_gemm(dt, NATIVE, GEMM_AUTO, g, s);
        };
        void* ga2 = sc.get(bt * m * k * es);
        gemm(gc, conj_t(b2, k, n), ga2, m, n, k);
        stage_operand_adjoint(sc, dt, ga2, st.left_shape, st.sum_a, st.a_perm, at(grad_[st.li]));
        void* gb2 = sc.get(bt * k * n * es);
        gemm(conj_t(a2, m, k), gc, gb2, k, m, n);
        stage_operand_adjoint(sc, dt, gb2, st.right_shape, st.sum_b, st.b_perm, at(grad_[st.ri]));
        break;
      }
    }
  }
}

void GradRunner::ensure_pool(hipStream_t s) {
  if (pool_.base) return;
  // a first pass sized the per-op scratch (stream-ordered allocations)
  ATN_HIP(hipStreamSynchronize(s));
  pool_.cap = std::max<size_t>(pool_.peak, 256);
  ATN_HIP(hipMalloc((void**)&pool_.base, pool_.cap));
}

void GradRunner::forward(const void* const* inputs, void* out, hipStream_t s) {
  set_bases(inputs, nullptr, out, nullptr, s);
  run_ops(0, fwd_end_, s);
  have_forward_ = true;
  forward_inputs_.assign(inputs, inputs + n_in_);
}

void GradRunner::backward(const void* const* inputs, const void* grad_out, void* const* grad_inputs, hipStream_t s) {
  if (!have_forward_ || !std::equal(forward_inputs_.begin(), forward_inputs_.end(), inputs))
    throw Error("backward needs the forward pass of the same inputs first");
  set_bases(inputs, grad_inputs, bases_[1 + 2 * n_in_], grad_out, s);
  run_ops(fwd_end_, ops_.size(), s);
  if (!pool_.base) ensure_pool(s);
}

}  // namespace atn
