// Executors of a contraction plan on one GPU.
//
// Plain path (the drop-in for src/slice.rs:960-1065 + src/compiled.rs:383-441):
// one slice at a time, one step at a time: gather the slice's inputs, then per
// step the planned sums, the two permutes and one GEMM; the per-slice results
// are summed in T in the reference's chunk order (src/slice.rs:1018-1064) and
// output blocks are scattered into the full output (src/slice.rs:1139-1171).
//
// Batched path: the per-slice plan is identical for every slice, so
//   * steps whose operands do not depend on the slice index run once (hoisted);
//   * every other step runs for B slices at once: permutes see the slice index
//     as an extra outer axis, GEMMs are strided-batched with stride 0 for the
//     shared operand (or folded into m when only A varies);
//   * slice results are summed into an FP64 accumulator in slice order;
//   * the whole sequence can be captured into one hipGraph.
#include "executor.hpp"

#include <algorithm>
#include <cmath>
#include <map>
#include <set>

namespace atn {

static size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// ------------------------------------------------------------ Executor build --
int Executor::newbuf(size_t bytes) {
  if (replay_) {
    if (replay_pos_ >= body_ids_.size()) throw Error("internal: tail replay out of buffers", -3);
    int id = body_ids_[replay_pos_++];
    if (bufs_[id].bytes < bytes) throw Error("internal: tail buffer larger than body buffer", -3);
    return id;
  }
  bufs_.push_back(Buf{std::max<size_t>(bytes, 4), t_, t_});
  int id = (int)bufs_.size() - 1;
  if (record_) body_ids_.push_back(id);
  return id;
}

void Executor::use(const Ref& r) {
  if (replay_ || r.kind != Ref::WS) return;
  bufs_[r.id].last = std::max(bufs_[r.id].last, t_);
}

static GatherArgs make_gather(const Plan& P, size_t i) {
  GatherArgs g{};
  const auto& legs = P.net.inputs[i];
  auto shape = P.net.shape_of(legs);
  std::vector<size_t> strides(shape.size(), 1);
  for (size_t a = shape.size(); a-- > 1;) strides[a - 1] = strides[a] * shape[a];
  std::set<size_t> sel_axes;
  for (auto& s : P.sel[i]) sel_axes.insert(s.axis);
  // kept axes, merging neighbours that are adjacent in the full tensor
  std::vector<std::pair<size_t, size_t>> ax;  // (dim, stride)
  for (size_t a = 0; a < shape.size(); ++a) {
    if (sel_axes.count(a) || shape[a] == 1) continue;
    if (!ax.empty() && ax.back().second == strides[a] * shape[a]) {
      ax.back().first *= shape[a];
      ax.back().second = strides[a];
    } else {
      ax.push_back({shape[a], strides[a]});
    }
  }
  if (ax.size() > (size_t)MAXR || P.sel[i].size() > (size_t)MAXR) throw Error("gather rank exceeds kernel limit");
  g.n_ax = (int)ax.size();
  g.inner = 1;
  for (size_t q = 0; q < ax.size(); ++q) {
    g.dims[q] = (uint32_t)ax[q].first;
    g.str[q] = (int64_t)ax[q].second;
    g.inner *= ax[q].first;
  }
  g.n_sel = (int)P.sel[i].size();
  for (size_t q = 0; q < P.sel[i].size(); ++q) {
    g.sel_div[q] = P.sel[i][q].div;
    g.sel_mod[q] = P.sel[i][q].mod;
    g.sel_str[q] = (int64_t)P.sel[i][q].stride;
  }
  return g;
}

static ScatterArgs make_scatter(const Plan& P) {
  ScatterArgs a{};
  if (P.block_div.size() > (size_t)MAXR || P.kept_dims.size() > (size_t)MAXR) throw Error("output rank exceeds kernel limit");
  a.n_blk = (int)P.block_div.size();
  for (size_t q = 0; q < P.block_div.size(); ++q) {
    a.blk_div[q] = P.block_div[q];
    a.blk_mod[q] = P.block_mod[q];
    a.blk_str[q] = (int64_t)P.block_stride[q];
  }
  a.n_kept = (int)P.kept_dims.size();
  for (size_t q = 0; q < P.kept_dims.size(); ++q) {
    a.kept_dims[q] = (uint32_t)P.kept_dims[q];
    a.kept_str[q] = (int64_t)P.kept_strides[q];
  }
  return a;
}

void Executor::gen_inputs(bool body, size_t nb, std::vector<Ref>& slots, std::vector<std::vector<size_t>>& shapes) {
  const Plan& P = *P_;
  const int dt = o_.dtype, es = es_;
  for (size_t i = 0; i < P.cp.n_inputs; ++i) {
    bool varying = !P.sel[i].empty();
    if (o_.batched && varying != body) continue;
    size_t nbI = (o_.batched && varying) ? nb : 1;
    Ref ref;
    auto shape = P.cp.preps[i].input_shape;
    if (varying) {
      GatherArgs ga = make_gather(P, i);
      ref = Ref{Ref::WS, newbuf(nbI * ga.inner * es)};
      Ref src{Ref::INPUT, (int)i};
      use(ref);
      emit([=](const Ctx& c) { launch_gather(es, ga, c.g0, nbI, c.ptr(src), c.ptr(ref), c.s); }, 0,
           2.0 * nbI * ga.inner * es);
    } else {
      ref = Ref{Ref::INPUT, (int)i};
    }
    for (auto& op : P.cp.preps[i].traces) {
      size_t ai = std::get<0>(op), aj = std::get<1>(op);
      bool diag = std::get<2>(op);
      size_t A = nbI, d = shape[ai], Bm = 1, C = 1;
      for (size_t q = 0; q < ai; ++q) A *= shape[q];
      for (size_t q = ai + 1; q < aj; ++q) Bm *= shape[q];
      for (size_t q = aj + 1; q < shape.size(); ++q) C *= shape[q];
      size_t out_numel = diag ? A * d * Bm * C : A * Bm * C;
      Ref dst{Ref::WS, newbuf(out_numel * es)};
      Ref src = ref;
      use(src);
      use(dst);
      double by = (double)(A * d * Bm * C + out_numel) * es;
      if (diag)
        emit([=](const Ctx& c) { launch_diag(es, A, d, Bm, C, c.ptr(src), c.ptr(dst), c.s); }, 0, by);
      else
        emit([=](const Ctx& c) { launch_trace(dt, A, d, Bm, C, c.ptr(src), c.ptr(dst), c.s); }, 0, by);
      shape.erase(shape.begin() + aj);
      if (!diag) shape.erase(shape.begin() + ai);
      ref = dst;
    }
    slots[i] = ref;
    shapes[i] = shape;
  }
}

bool gather_tables(const std::vector<size_t>& shape, const std::vector<size_t>& perm, size_t n_con,
                   std::vector<int32_t>& rows, std::vector<int32_t>& cols, int32_t& col_run, int32_t& row_run) {
  const size_t r = shape.size();
  std::vector<int64_t> st(r, 1);
  for (size_t a = r; a-- > 1;) st[a - 1] = st[a] * (int64_t)shape[a];
  if (r && st[0] * (int64_t)shape[0] > INT32_MAX) return false;
  auto table = [&](size_t d0, size_t d1, std::vector<int32_t>& out) {
    size_t len = 1;
    for (size_t d = d0; d < d1; ++d) len *= shape[perm[d]];
    out.assign(len, 0);
    // row-major enumeration over output axes d0..d1-1 (last fastest), the offset
    // carried like an odometer
    std::vector<size_t> idx(d1 - d0, 0);
    int64_t off = 0;
    for (size_t e = 0; e < len; ++e) {
      out[e] = (int32_t)off;
      for (size_t q = d1 - d0; q-- > 0;) {
        const size_t a = perm[d0 + q];
        off += st[a];
        if (++idx[q] < shape[a]) break;
        off -= st[a] * (int64_t)shape[a];
        idx[q] = 0;
      }
    }
  };
  table(0, n_con, rows);
  table(n_con, r, cols);
  // innermost free output axes that are also the innermost input axes: runs of
  // consecutive offsets starting at multiples of the run length
  col_run = 1;
  for (size_t d = r, a = r; d-- > n_con && a-- > 0;) {
    if (perm[d] != a) break;
    col_run *= (int32_t)shape[a];
  }
  // the same for the innermost contracted axes (rows): runs of consecutive row offsets
  row_run = 1;
  for (size_t d = n_con, a = r; d-- > 0 && a-- > 0;) {
    if (perm[d] != a) break;
    row_run *= (int32_t)shape[a];
  }
  return true;
}

bool Executor::add_gather(const std::vector<size_t>& shape, const std::vector<size_t>& perm, size_t n_con, int64_t& row,
                          int64_t& col, int32_t& col_run, int32_t& row_run) {
  size_t nrow = 1, ncol = 1;
  for (size_t d = 0; d < n_con; ++d) nrow *= shape[perm[d]];
  for (size_t d = n_con; d < shape.size(); ++d) ncol *= shape[perm[d]];
  if (const_host_.size() + nrow + ncol > ((size_t)64 << 20)) return false;  // 256 MB of tables at most
  std::vector<int32_t> rows, cols;
  if (!gather_tables(shape, perm, n_con, rows, cols, col_run, row_run)) return false;
  row = (int64_t)const_host_.size();
  const_host_.insert(const_host_.end(), rows.begin(), rows.end());
  col = (int64_t)const_host_.size();
  const_host_.insert(const_host_.end(), cols.begin(), cols.end());
  return true;
}

void Executor::gen_gemm(const PairStep& st, size_t nbo, bool va, bool vb, const Ref& a, const Ref& b, const Ref& c,
                        const Ref* cmax, int64_t brow, int64_t bcol, int32_t bcol_run, const Ref* amax,
                        const Ref* bmax, int32_t brow_run, int64_t arow, int64_t acol) {
  GemmArgs g{};
  g.m = (int64_t)st.m, g.k = (int64_t)st.k, g.n = (int64_t)st.n;
  g.lda = g.k, g.ldb = g.n, g.ldc = g.n;
  g.nbi = (int64_t)st.bt;
  g.sai = g.m * g.k, g.sbi = g.k * g.n, g.sci = g.m * g.n;
  g.nbo = (int64_t)nbo;
  g.sao = va ? (int64_t)st.bt * g.m * g.k : 0;
  g.sbo = vb ? (int64_t)st.bt * g.k * g.n : 0;
  g.sco = (int64_t)st.bt * g.m * g.n;
  if (st.bt == 1 && va && !vb && nbo > 1) {  // A rows of consecutive slices are contiguous, as are C rows
    g.m *= (int64_t)nbo;
    g.nbo = 1;
    g.sao = g.sbo = g.sco = 0;
    g.sai = g.m * g.k, g.sci = g.m * g.n;
  }
  int idx = (int)gemm_choice_.size();
  gemm_choice_.push_back(o_.gemm_kernel);
  g.bcol_run = bcol_run;
  g.brow_run = brow_run;
  const Ref am = amax ? *amax : Ref{}, bm = bmax ? *bmax : Ref{};
  if (amax) use(am);
  if (bmax) use(bm);
  gemm_sites_.push_back(GemmSite{g, a, b, c, brow, bcol, bcol_run, am, bm, arow, acol});
  int dt = o_.dtype, prec = o_.precision;
  use(a);
  use(b);
  use(c);
  double nz = (double)g.nbo * (double)g.nbi, es = (double)es_;
  double flops = (dtype_complex(dt) ? 8.0 : 2.0) * nz * (double)g.m * (double)g.k * (double)g.n;
  double bytes = es * (((g.sao || g.nbo == 1) ? nz : (double)g.nbi) * (double)(g.m * g.k) +
                       ((g.sbo || g.nbo == 1) ? nz : (double)g.nbi) * (double)(g.k * g.n) + nz * (double)(g.m * g.n));
  Ref scr{Ref::WS, scratch_buf_};
  size_t& scr_bytes = bufs_[scratch_buf_].bytes;
  for (int kern : {GEMM_AUTO, GEMM_SIMPLE, GEMM_SKINNY, GEMM_MFMA, GEMM_SPLITK, GEMM_PRESPLIT, GEMM_MFMA_R32,
                   GEMM_MFMA_R64, GEMM_MFMA_R128, GEMM_PRESPLIT_A})  // autotune candidates
    scr_bytes = std::max(scr_bytes, gemm_scratch_bytes(dt, prec, kern, g));
  const int sb = scratch_buf_;
  const Ref mref = cmax ? *cmax : Ref{};
  if (cmax) use(mref);
  emit(
      [=](const Ctx& cx) {
        GemmArgs q = g;
        q.a = cx.ptr(a), q.b = cx.ptr(b), q.c = cx.ptr(c);
        q.scratch = cx.ptr(scr);
        q.scratch_bytes = bufs_[sb].bytes;
        if (mref.kind != Ref::NONE) q.cmax = (unsigned long long*)cx.ptr(mref);
        if (am.kind != Ref::NONE) q.amax = (const unsigned long long*)cx.ptr(am);
        if (bm.kind != Ref::NONE) q.bmax = (const unsigned long long*)cx.ptr(bm);
        if (bcol >= 0) q.brow = consts_ + brow, q.bcol = consts_ + bcol;
        if (acol >= 0) q.arow = consts_ + arow, q.acol = consts_ + acol;
        launch_gemm(dt, prec, gemm_choice_[idx], q, cx.s);
      },
      flops, bytes);
}

void Executor::gen_steps(bool body, size_t nb, std::vector<Ref>& slots, std::vector<std::vector<size_t>>& shapes) {
  const Plan& P = *P_;
  const int dt = o_.dtype, es = es_;
  // The slice-invariant part of the batched form runs level by level of the
  // contraction tree (a step's level: one more than its operands'), and the steps
  // of one level that run on the simple kernel share one grouped launch.
  const bool grouping = o_.batched && !body && !o_.strip;
  std::vector<size_t> order(P.cp.steps.size());
  std::vector<size_t> level(P.cp.n_inputs + P.cp.steps.size(), 0);
  for (size_t si = 0; si < P.cp.steps.size(); ++si) {
    const PairStep& st = P.cp.steps[si];
    level[st.out] = std::max(level[st.li], level[st.ri]) + 1;
    order[si] = si;
  }
  if (grouping)
    std::stable_sort(order.begin(), order.end(),
                     [&](size_t x, size_t y) { return level[P.cp.steps[x].out] < level[P.cp.steps[y].out]; });
  std::vector<GroupGemm> group;
  auto flush = [&]() {
    if (group.empty()) return;
    const size_t first = ggemm_.size();
    uint64_t blocks = 0;
    double flops = 0, bytes = 0;
    const double cf = dtype_complex(dt) ? 8.0 : 2.0;
    for (GroupGemm& q : group) {
      use(q.a);
      use(q.b);
      use(q.c);
      q.block0 = blocks;
      blocks += ((uint64_t)(q.m * q.n) + 255) / 256;
      flops += cf * (double)q.m * (double)q.k * (double)q.n;
      bytes += (double)es * (double)(q.m * q.k + q.k * q.n + q.m * q.n);
      ggemm_.push_back(q);
    }
    const int cnt = (int)group.size();
    emit([=](const Ctx& c) {
      launch_gemm_simple_group(dt, ggemm_dev_ + first, cnt, blocks, c.ws, inputs_dev_, consts_, c.s);
    }, flops, bytes);
    group.clear();
  };
  size_t cur_level = 0;
  for (size_t oi = 0; oi < order.size(); ++oi) {
    const size_t si = order[oi];
    const PairStep& st = P.cp.steps[si];
    bool v = P.varying[st.out];
    if (o_.batched && v != body) continue;
    if (grouping && level[st.out] != cur_level) {
      flush();
      cur_level = level[st.out];
    }
    bool va = o_.batched && P.varying[st.li], vb = o_.batched && P.varying[st.ri];
    size_t nbA = va ? nb : 1, nbB = vb ? nb : 1, nbC = (va || vb) ? nb : 1;
    std::vector<Ref> temps;
    auto chain = [&](Ref cur, std::vector<size_t> shape, const std::vector<size_t>& sums, const std::vector<size_t>& perm,
                     size_t nbX) -> Ref {
      for (size_t ax : sums) {
        size_t outer = nbX, inner = 1, d = shape[ax];
        for (size_t q = 0; q < ax; ++q) outer *= shape[q];
        for (size_t q = ax + 1; q < shape.size(); ++q) inner *= shape[q];
        Ref dst{Ref::WS, newbuf(outer * inner * es)};
        Ref src = cur;
        use(src);
        use(dst);
        emit([=](const Ctx& c) { launch_sum_axis(dt, outer, d, inner, c.ptr(src), c.ptr(dst), c.s); }, 0,
             (double)(outer * inner * (d + 1)) * es);
        temps.push_back(dst);
        shape.erase(shape.begin() + ax);
        cur = dst;
      }
      PermPlan pp = plan_permute(shape, perm, nbX, es);
      if (pp.kind != 0) {
        Ref dst{Ref::WS, newbuf(pp.numel * es)};
        Ref src = cur;
        use(src);
        use(dst);
        emit([=](const Ctx& c) { launch_permute(pp, c.ptr(src), c.ptr(dst), c.s); }, 0, 2.0 * pp.numel * es);
        temps.push_back(dst);
        cur = dst;
      }
      return cur;
    };
    // B read through gather tables instead of being permuted into [con, free_b]
    // (no batch legs, no summed legs): the GEMM kernels address b[p][j] as
    // row_off[p] + col_off[j] of the unpermuted operand. The contracted legs are
    // then taken in B's memory order (A is permuted to match), so where they are
    // B's innermost legs the rows come in runs of consecutive offsets.
    std::vector<size_t> a_perm = st.a_perm, b_perm = st.b_perm;
    // B may be stored in another axis order than the plan's (only slots read as B
    // without summed legs are): its permutation is taken over the stored axes
    const std::vector<size_t>& ob = slot_order_[st.ri];
    std::vector<Leg> rl(ob.size());  // B's legs in storage order
    {
      std::vector<size_t> pos(ob.size());
      for (size_t d = 0; d < ob.size(); ++d) pos[ob[d]] = d, rl[d] = st.right_legs[ob[d]];
      for (size_t& x : b_perm) x = pos[x];
    }
    size_t n_con = 0;  // right legs shared with the left operand and absent from the result
    for (size_t d = 0; d < b_perm.size(); ++d) {
      const Leg l = rl[b_perm[d]];
      const bool in_left = std::count(st.left_legs.begin(), st.left_legs.end(), l) > 0;
      const bool in_res = std::count(st.result_legs.begin(), st.result_legs.end(), l) > 0;
      if (in_left && !in_res) n_con = d + 1;
    }
    const bool gatherable = o_.gather && st.bt == 1 && st.sum_b.empty();
    if (gatherable && n_con > 1) {
      std::vector<size_t> bc(b_perm.begin(), b_perm.begin() + (long)n_con);
      std::sort(bc.begin(), bc.end());
      std::vector<Leg> la = st.left_legs;  // A's legs once its summed legs are gone
      for (size_t ax : st.sum_a) la.erase(la.begin() + (long)ax);
      const size_t nfa = a_perm.size() - n_con;  // a_perm = [free_a, con] without batch legs
      for (size_t t = 0; t < n_con; ++t) {
        const Leg l = rl[bc[t]];
        a_perm[nfa + t] = (size_t)(std::find(la.begin(), la.end(), l) - la.begin());
        b_perm[t] = bc[t];
      }
    }
    // Storage order of the result: [free_a][free_b in B's column order]. When the
    // result is the B of a later step, B's free legs are ordered for both reads:
    // the innermost run of them in B's memory (this step's contiguous columns)
    // stays innermost, and the others go legs contracted by the reader first,
    // its free legs last, so that its columns come in runs too.
    std::vector<size_t>& oc = slot_order_[st.out];
    oc.resize(st.result_legs.size());
    for (size_t d = 0; d < oc.size(); ++d) oc[d] = d;
    const long rd = consumer_[st.out].first;
    if (gatherable && rd >= 0 && consumer_[st.out].second == 1 && P.cp.steps[rd].bt == 1 &&
        P.cp.steps[rd].sum_b.empty()) {
      const size_t nfa = a_perm.size() - n_con;
      std::vector<Leg> res;  // the plan's result order must be [free_a (a_perm)][free_b (b_perm)]
      std::vector<Leg> la = st.left_legs;
      for (size_t ax : st.sum_a) la.erase(la.begin() + (long)ax);
      for (size_t d = 0; d < nfa; ++d) res.push_back(la[a_perm[d]]);
      std::vector<size_t> fr(b_perm.begin() + (long)n_con, b_perm.end());
      for (size_t a : fr) res.push_back(rl[a]);
      if (res == st.result_legs) {
        std::sort(fr.begin(), fr.end());  // memory order
        size_t run = 0;
        while (run < fr.size() && fr[fr.size() - 1 - run] == rl.size() - 1 - run) ++run;
        const PairStep& cs = P.cp.steps[rd];
        auto read_con = [&](Leg l) {
          return std::count(cs.left_legs.begin(), cs.left_legs.end(), l) > 0 &&
                 std::count(cs.result_legs.begin(), cs.result_legs.end(), l) == 0;
        };
        std::vector<size_t> nf;
        for (size_t t = 0; t + run < fr.size(); ++t)
          if (read_con(rl[fr[t]])) nf.push_back(fr[t]);
        for (size_t t = 0; t + run < fr.size(); ++t)
          if (!read_con(rl[fr[t]])) nf.push_back(fr[t]);
        for (size_t t = fr.size() - run; t < fr.size(); ++t) nf.push_back(fr[t]);
        std::copy(nf.begin(), nf.end(), b_perm.begin() + (long)n_con);
        for (size_t t = 0; t < nf.size(); ++t)
          oc[nfa + t] = (size_t)(std::find(st.result_legs.begin(), st.result_legs.end(), rl[nf[t]]) -
                                 st.result_legs.begin());
      }
    }
    // A read through gather tables too (no permute launch) on unbatched steps that
    // run on the simple kernel: the many tiny steps of a plan's slice-invariant part,
    // and every step of the plain form. Not in the batched body, whose tail batch
    // replays the body's buffers and must make the same choices at another batch size.
    int64_t arow = -1, acol = -1;
    Ref a;
    if (gatherable && (!o_.batched || !body) && nbA == 1 && nbC == 1 && st.sum_a.empty() && !o_.strip &&
        plan_permute(shapes[st.li], a_perm, 1, es).kind != 0) {
      GemmArgs t{};
      t.m = (int64_t)st.m, t.k = (int64_t)st.k, t.n = (int64_t)st.n, t.nbo = t.nbi = 1;
      int32_t cr = 1, rr = 1;
      if (gemm_runs_simple(dt, o_.precision, o_.gemm_kernel, t) &&
          add_gather(shapes[st.li], a_perm, a_perm.size() - n_con, arow, acol, cr, rr)) {
        a = slots[st.li];
        use(a);
      } else {
        arow = acol = -1;
      }
    }
    if (acol < 0) a = chain(slots[st.li], shapes[st.li], st.sum_a, a_perm, nbA);
    int64_t brow = -1, bcol = -1;
    int32_t bcol_run = 1, brow_run = 1;
    Ref b;
    if (gatherable && plan_permute(shapes[st.ri], b_perm, 1, es).kind != 0) {
      if (add_gather(shapes[st.ri], b_perm, n_con, brow, bcol, bcol_run, brow_run)) {
        b = slots[st.ri];
        use(b);
      }
    }
    if (bcol < 0) b = chain(slots[st.ri], shapes[st.ri], st.sum_b, b_perm, nbB);
    Ref c{Ref::WS, newbuf(nbC * st.bt * st.m * st.n * es)};
    // operand max magnitudes folded by the producing GEMMs (a permute or a gather
    // keeps the max; a summed leg does not)
    const Ref* am = st.sum_a.empty() && slot_max_[st.li].kind != Ref::NONE ? &slot_max_[st.li] : nullptr;
    const Ref* bm = st.sum_b.empty() && slot_max_[st.ri].kind != Ref::NONE ? &slot_max_[st.ri] : nullptr;
    if (o_.strip) {
      // the GEMM folds the max of its result into mb; the rescale then runs only
      // when that max leaves [2^-STRIP_KEEP, 2^STRIP_KEEP] (always after the last step)
      Ref mb{Ref::WS, maxbits_buf_};
      use(mb);
      emit([=](const Ctx& cx) { launch_fill_zero(cx.ptr(mb), 8, cx.s); });
      gen_gemm(st, nbC, va, vb, a, b, c, &mb, brow, bcol, bcol_run, am, bm, brow_run, arow, acol);
      size_t numel = nbC * st.bt * st.m * st.n;
      Ref ex{Ref::WS, expo_buf_};
      use(c);
      if (si + 1 == P.cp.steps.size()) {
        // last step: the reference's exact division by the max magnitude
        emit([=](const Ctx& cx) {
          launch_strip_exact(dt, cx.ptr(c), numel, (uint64_t*)cx.ptr(mb), (double*)((char*)cx.ptr(ex) + 8), cx.s);
        });
      } else {
        emit([=](const Ctx& cx) {
          launch_scale_pow2(dt, cx.ptr(c), numel, (const uint64_t*)cx.ptr(mb), (int64_t*)cx.ptr(ex), cx.s, STRIP_KEEP);
        });
      }
    } else if (want_max_[st.out]) {
      Ref mc{Ref::WS, newbuf(8)};
      use(mc);
      emit([=](const Ctx& cx) { launch_fill_zero(cx.ptr(mc), 8, cx.s); });
      gen_gemm(st, nbC, va, vb, a, b, c, &mc, brow, bcol, bcol_run, am, bm, brow_run, arow, acol);
      slot_max_[st.out] = mc;
    } else {
      GemmArgs t{};
      t.m = (int64_t)st.m, t.k = (int64_t)st.k, t.n = (int64_t)st.n, t.nbo = t.nbi = 1;
      if (grouping && st.bt == 1 && nbC == 1 && gemm_runs_simple(dt, o_.precision, o_.gemm_kernel, t))
        group.push_back(GroupGemm{a, b, c, arow, acol, brow, bcol, t.m, t.k, t.n, 0});
      else
        gen_gemm(st, nbC, va, vb, a, b, c, nullptr, brow, bcol, bcol_run, am, bm, brow_run, arow, acol);
    }
    slots[st.out] = c;
    shapes[st.out].resize(oc.size());
    for (size_t d = 0; d < oc.size(); ++d) shapes[st.out][d] = st.result_shape[oc[d]];
  }
  flush();
}

Ref Executor::gen_finalize(size_t nb, const Ref& src0, bool to_output) {
  const Plan& P = *P_;
  const int dt = o_.dtype, es = es_;
  const Finalize& f = P.cp.fin;
  auto shape = f.input_shape;
  Ref cur = src0;
  auto sh = shape;
  for (size_t ax : f.sum) sh.erase(sh.begin() + ax);
  PermPlan pp = plan_permute(sh, f.perm, nb, es);
  for (size_t qi = 0; qi < f.sum.size(); ++qi) {
    size_t ax = f.sum[qi];
    size_t outer = nb, inner = 1, d = shape[ax];
    for (size_t q = 0; q < ax; ++q) outer *= shape[q];
    for (size_t q = ax + 1; q < shape.size(); ++q) inner *= shape[q];
    bool last = qi + 1 == f.sum.size() && pp.kind == 0;
    Ref dst = (last && to_output) ? Ref{Ref::OUTPUT, 0} : Ref{Ref::WS, newbuf(outer * inner * es)};
    Ref src = cur;
    use(src);
    use(dst);
    emit([=](const Ctx& c) { launch_sum_axis(dt, outer, d, inner, c.ptr(src), c.ptr(dst), c.s); }, 0,
         (double)(outer * inner * (d + 1)) * es);
    shape.erase(shape.begin() + ax);
    cur = dst;
  }
  size_t numel = nb * prod(shape);
  if (pp.kind != 0) {
    Ref dst = to_output ? Ref{Ref::OUTPUT, 0} : Ref{Ref::WS, newbuf(numel * es)};
    Ref src = cur;
    use(src);
    use(dst);
    emit([=](const Ctx& c) { launch_permute(pp, c.ptr(src), c.ptr(dst), c.s); }, 0, 2.0 * numel * es);
    cur = dst;
  } else if (to_output && cur.kind != Ref::OUTPUT) {
    Ref src = cur;
    use(src);
    emit([=](const Ctx& c) { launch_copy(numel * es, c.ptr(src), c.out, c.s); }, 0, 2.0 * numel * es);
    cur = Ref{Ref::OUTPUT, 0};
  }
  return cur;
}

void Executor::build(size_t B) {
  const Plan& P = *P_;
  const int dt = o_.dtype, es = es_;
  B_ = B;
  bufs_.clear();
  pre_.clear(), body_.clear(), tail_.clear(), post_.clear();
  gemm_choice_.clear();
  gemm_sites_.clear();
  ggemm_.clear();
  const_host_.clear();
  body_ids_.clear();
  replay_ = record_ = false;
  t_ = 0;
  const size_t nslot = P.cp.n_inputs + P.cp.steps.size();
  const bool sliced = P.n_slices > 1;
  const size_t out_numel = P.sub_out_numel;
  cur_ = &pre_;
  scratch_buf_ = newbuf(256);  // FP16X2 scales, split-K partials (grown per GEMM site); lives for the whole run
  if (o_.strip) {
    // expo: int64 power-of-two exponent of the intermediate rescales, then the
    // f64 log10 of the last step's exact max
    maxbits_buf_ = newbuf(8);
    expo_buf_ = newbuf(16);
    Ref ex{Ref::WS, expo_buf_};
    use(ex);
    emit([=](const Ctx& c) { launch_fill_zero(c.ptr(ex), 16, c.s); });
  }
  std::vector<Ref> slots(nslot);
  std::vector<std::vector<size_t>> shapes(nslot);
  ScatterArgs sc = make_scatter(P);
  slot_max_.assign(nslot, Ref{});
  // inputs are stored in the plan's order (as the steps see them)
  slot_order_.assign(nslot, {});
  consumer_.assign(nslot, {-1, 0});
  for (size_t i = 0; i < P.cp.steps.size(); ++i) {
    const PairStep& st = P.cp.steps[i];
    consumer_[st.li] = {(long)i, 0};
    consumer_[st.ri] = {(long)i, 1};
    for (auto [slot, rank] : {std::pair<size_t, size_t>{st.li, st.left_legs.size()}, {st.ri, st.right_legs.size()}})
      if (slot < P.cp.n_inputs) {
        slot_order_[slot].resize(rank);
        for (size_t d = 0; d < rank; ++d) slot_order_[slot][d] = d;
      }
  }
  want_max_.assign(nslot, 0);
  if (split_halves(o_.precision) == FP16X2 && !o_.strip)
    for (const PairStep& st : P.cp.steps)
      if (split_products_apply(dt, o_.precision, (int64_t)st.m, (int64_t)st.k, (int64_t)st.n)) {
        if (st.sum_a.empty()) want_max_[st.li] = 1;
        if (st.sum_b.empty()) want_max_[st.ri] = 1;
      }
  if (o_.batched) {
    gen_inputs(false, 1, slots, shapes);
    gen_steps(false, 1, slots, shapes);
    if (sliced) {
      size_t wbytes = dtype_complex(dt) ? 16 : 8;
      acc_buf_ = newbuf(P.n_blocks * out_numel * wbytes);
      Ref acc{Ref::WS, acc_buf_};
      size_t nbytes = P.n_blocks * out_numel * wbytes;
      use(acc);
      emit([=](const Ctx& c) { launch_fill_zero(c.ptr(acc), nbytes, c.s); });
    }
    int pre_end = t_;
    body_begin_ = t_;
    auto gen_body = [&](size_t nb) {
      std::vector<Ref> sl = slots;
      auto sh = shapes;
      gen_inputs(true, nb, sl, sh);
      gen_steps(true, nb, sl, sh);
      Ref res = gen_finalize(nb, sl[P.cp.fin.slot], !sliced);
      if (sliced) {
        Ref acc{Ref::WS, acc_buf_};
        size_t n_int = P.n_internal;
        use(res);
        use(acc);
        emit([=](const Ctx& c) { launch_reduce_f64(dt, c.ptr(res), c.ptr(acc), c.g0, nb, n_int, out_numel, c.s); }, 0,
             (double)(nb * out_numel * es) + 4.0 * out_numel * es);
      }
    };
    cur_ = &body_;
    record_ = true;
    gen_body(B);
    record_ = false;
    body_end_ = t_;
    tail_nb_ = n_run_ % B;
    if (tail_nb_) {
      cur_ = &tail_;
      replay_ = true;
      replay_pos_ = 0;
      int saved = t_;
      gen_body(tail_nb_);
      t_ = saved;
      replay_ = false;
    }
    // pre buffers read by the body stay alive across all batches
    for (auto& b : bufs_)
      if (b.first < pre_end && b.last >= body_begin_) b.last = std::max(b.last, body_end_);
    cur_ = &post_;
    if (sliced) {
      Ref acc{Ref::WS, acc_buf_};
      size_t nblk = P.n_blocks;
      use(acc);
      emit([=](const Ctx& c) { launch_scatter(dt, true, sc, 0, nblk, out_numel, c.ptr(acc), c.out, c.s); });
    }
  } else {
    body_begin_ = t_;
    cur_ = &body_;
    gen_inputs(true, 1, slots, shapes);
    gen_steps(true, 1, slots, shapes);
    Ref res = gen_finalize(1, slots[P.cp.fin.slot], !sliced);
    if (sliced) {
      size_t nch = P.n_chunks, n_int = P.n_internal;
      size_t chunk_len = (n_int + nch - 1) / nch;
      std::vector<Ref> parts;
      for (size_t c = 0; c < nch; ++c) parts.push_back(Ref{Ref::WS, newbuf(out_numel * es)});
      Ref blk{Ref::WS, newbuf(out_numel * es)};
      use(res);
      for (auto& r : parts) use(r);
      use(blk);
      emit([=](const Ctx& c) {
        uint64_t g = c.g0, b = g / n_int, s = g % n_int, ch = s / chunk_len;
        launch_accum(dt, s % chunk_len == 0, c.ptr(res), c.ptr(parts[ch]), out_numel, c.s);
        if (s == n_int - 1) {
          for (size_t q = 0; q < nch; ++q) launch_accum(dt, q == 0, c.ptr(parts[q]), c.ptr(blk), out_numel, c.s);
          launch_scatter(dt, false, sc, b, 1, out_numel, c.ptr(blk), c.out, c.s);
        }
      });
      // chunk partials and the block accumulator persist across slices
      for (auto& r : parts) bufs_[r.id].first = body_begin_;
      bufs_[blk.id].first = body_begin_;
    }
    body_end_ = t_;
  }
  // strip bookkeeping and GEMM scratch live for the whole run
  for (int id : {maxbits_buf_, expo_buf_})
    if (o_.strip && id >= 0) bufs_[id].first = 0, bufs_[id].last = t_ + 1;
  bufs_[scratch_buf_].first = 0, bufs_[scratch_buf_].last = t_ + 1;
  if (acc_buf_ >= 0 && o_.batched && sliced) bufs_[acc_buf_].last = t_ + 1;
  // Offline first-fit placement of buffers with overlapping lifetimes.
  std::vector<int> order(bufs_.size());
  for (size_t i = 0; i < order.size(); ++i) order[i] = (int)i;
  std::sort(order.begin(), order.end(), [&](int x, int y) {
    if (bufs_[x].first != bufs_[y].first) return bufs_[x].first < bufs_[y].first;
    return bufs_[x].bytes > bufs_[y].bytes;
  });
  offs_.assign(bufs_.size(), 0);
  struct Live {
    size_t off, end;
    int last;
  };
  std::vector<Live> live;
  size_t total = 0;
  for (int id : order) {
    const Buf& b = bufs_[id];
    live.erase(std::remove_if(live.begin(), live.end(), [&](const Live& l) { return l.last < b.first; }), live.end());
    std::sort(live.begin(), live.end(), [](const Live& x, const Live& y) { return x.off < y.off; });
    size_t need = align_up(b.bytes, 256), pos = 0;
    for (auto& l : live) {
      if (pos + need <= l.off) break;
      pos = std::max(pos, l.end);
    }
    offs_[id] = pos;
    live.push_back(Live{pos, pos + need, b.last});
    total = std::max(total, pos + need);
  }
  ws_bytes_ = total;
}

Executor::Executor(std::shared_ptr<const Plan> plan, const ExecOptions& o) : P_(std::move(plan)), o_(o) {
  const Plan& P = *P_;
  es_ = (int)dtype_size(o_.dtype);
  if (o_.precision != NATIVE) {
    if (o_.precision == C3M && o_.dtype != C64) throw Error("precision c3m applies to complex64 only");
    if (split_precision(o_.precision) && !(o_.dtype == C64 || o_.dtype == F32))
      throw Error("split precision applies to float32 / complex64 only");
  }
  if (o_.strip && P.n_slices > 1) throw Error("strip is available for unsliced plans (execute_stripped)");
  if (!o_.batched) o_.use_graph = false;
  g_lo_ = o_.slice_begin;
  const size_t g_hi = o_.slice_end ? o_.slice_end : P.n_slices;
  if (g_lo_ != 0 || g_hi != P.n_slices) {
    if (!o_.batched) throw Error("slice ranges need the batched form");
    if (!(g_lo_ < g_hi && g_hi <= P.n_slices))
      throw Error(cat("slice range [", g_lo_, ", ", g_hi, ") out of bounds for ", P.n_slices, " slices"));
  }
  n_run_ = g_hi - g_lo_;
  size_t limit = o_.workspace_limit;
  if (!limit) {
    size_t fr = 0, tot = 0;
    ATN_HIP(hipMemGetInfo(&fr, &tot));
    limit = (size_t)(0.6 * (double)fr);
  }
  size_t B = 1;
  if (o_.batched) {
    B = n_run_;
    if (o_.max_batch) B = std::min(B, o_.max_batch);
  }
  for (;;) {
    build(B);
    if (ws_bytes_ <= limit || B == 1) break;
    B = (B + 1) / 2;
  }
  if (ws_bytes_ > limit) throw Error(cat("workspace of ", ws_bytes_, " bytes exceeds the limit of ", limit, " bytes"));
  ATN_HIP(hipMalloc(&ws_, std::max<size_t>(ws_bytes_, 256)));
  if (!const_host_.empty()) {
    ATN_HIP(hipMalloc((void**)&consts_, const_host_.size() * sizeof(int32_t)));
    ATN_HIP(hipMemcpy(consts_, const_host_.data(), const_host_.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  }
  if (!ggemm_.empty()) {
    auto where = [&](const Ref& r) -> int64_t {
      if (r.kind == Ref::WS) return (int64_t)offs_[r.id];
      if (r.kind == Ref::INPUT) return -(int64_t)r.id - 1;
      throw Error("internal: grouped GEMM operand outside the workspace and inputs", -3);
    };
    std::vector<SimpleGemmDesc> d(ggemm_.size());
    for (size_t i = 0; i < ggemm_.size(); ++i) {
      const GroupGemm& q = ggemm_[i];
      if (q.c.kind != Ref::WS) throw Error("internal: grouped GEMM result outside the workspace", -3);
      d[i] = SimpleGemmDesc{where(q.a), where(q.b), where(q.c), q.arow, q.acol, q.brow, q.bcol, q.m, q.k, q.n, q.block0};
    }
    ATN_HIP(hipMalloc((void**)&ggemm_dev_, d.size() * sizeof(SimpleGemmDesc)));
    ATN_HIP(hipMemcpy(ggemm_dev_, d.data(), d.size() * sizeof(SimpleGemmDesc), hipMemcpyHostToDevice));
    ATN_HIP(hipMalloc((void**)&inputs_dev_, std::max<size_t>(P.net.inputs.size(), 1) * sizeof(void*)));
  }
  if (o_.use_graph) ATN_HIP(hipStreamCreateWithFlags(&cap_stream_, hipStreamNonBlocking));
}

Executor::~Executor() {
  if (gexec_) (void)hipGraphExecDestroy(gexec_);
  if (graph_) (void)hipGraphDestroy(graph_);
  if (cap_stream_) (void)hipStreamDestroy(cap_stream_);
  if (ws_) (void)hipFree(ws_);
  if (consts_) (void)hipFree(consts_);
  if (ggemm_dev_) (void)hipFree(ggemm_dev_);
  if (inputs_dev_) (void)hipFree(inputs_dev_);
}

void Executor::sync_inputs(const void* const* inputs) {
  if (!inputs_dev_) return;
  const size_t n = P_->net.inputs.size();
  std::vector<const void*> now(inputs, inputs + n);
  if (now == inputs_seen_) return;
  // earlier runs, on any stream, may still read the table
  ATN_HIP(hipDeviceSynchronize());
  ATN_HIP(hipMemcpy(inputs_dev_, now.data(), n * sizeof(void*), hipMemcpyHostToDevice));
  inputs_seen_ = now;
}

void Executor::run_ops(const std::vector<Op>& ops, const Ctx& c) {
  for (auto& op : ops) op(c);
}

void Executor::enqueue(const Ctx& c0, const std::function<void(const Op&, const Ctx&)>& visit) {
  const Plan& P = *P_;
  Ctx c = c0;
  auto run = [&](const std::vector<Op>& ops) {
    if (visit)
      for (auto& op : ops) visit(op, c);
    else
      run_ops(ops, c);
  };
  run(pre_);
  if (o_.batched) {
    size_t nfull = n_run_ / B_;
    for (size_t q = 0; q < nfull; ++q) {
      c.g0 = g_lo_ + q * B_;
      run(body_);
    }
    if (tail_nb_) {
      c.g0 = g_lo_ + nfull * B_;
      run(tail_);
    }
  } else {
    for (size_t g = 0; g < P.n_slices; ++g) {
      c.g0 = g;
      run(body_);
    }
  }
  c.g0 = 0;
  run(post_);
}

void Executor::profile(const void* const* inputs, void* out, hipStream_t s, std::vector<KernelStat>& stats) {
  Ctx c;
  c.ws = ws_;
  c.offs = &offs_;
  c.inputs = inputs;
  c.out = out;
  c.s = s;
  sync_inputs(inputs);
  if (o_.autotune && o_.gemm_kernel == GEMM_AUTO && !tuned_) autotune(c, s);
  struct Rec {
    hipEvent_t a, b;
    int kind;
    size_t launches;
    double flops, bytes;
  };
  std::vector<Rec> recs;
  std::vector<hipEvent_t> pool;
  auto ev = [&]() {
    hipEvent_t e;
    ATN_HIP(hipEventCreate(&e));
    pool.push_back(e);
    return e;
  };
  try {
    enqueue(c, [&](const Op& op, const Ctx& cx) {
      Rec r{ev(), ev(), K_OTHER, 0, op.flops, op.bytes};
      size_t l0 = g_launch_count;
      ATN_HIP(hipEventRecord(r.a, cx.s));
      op(cx);
      ATN_HIP(hipEventRecord(r.b, cx.s));
      r.kind = g_last_kind;
      r.launches = g_launch_count - l0;
      recs.push_back(r);
    });
    ATN_HIP(hipStreamSynchronize(s));
    stats.assign(K_NKINDS, KernelStat{});
    for (auto& r : recs) {
      float ms = 0;
      ATN_HIP(hipEventElapsedTime(&ms, r.a, r.b));
      auto& k = stats[r.kind];
      k.calls += 1;
      k.launches += r.launches;
      k.ms += ms;
      k.flops += r.flops;
      k.bytes += r.bytes;
    }
  } catch (...) {
    for (auto e : pool) (void)hipEventDestroy(e);
    throw;
  }
  for (auto e : pool) (void)hipEventDestroy(e);
}

void Executor::autotune(const Ctx& c, hipStream_t s) {
  // Time the candidate kernels of every GEMM site on its real operands once,
  // keyed by shape, and keep the fastest (the executor's workspace holds
  // operands of the right size; values do not influence timing).
  std::map<std::tuple<int64_t, int64_t, int64_t, int64_t, int64_t, int>, int> best;
  hipEvent_t e0, e1;
  ATN_HIP(hipEventCreate(&e0));
  ATN_HIP(hipEventCreate(&e1));
  for (size_t i = 0; i < gemm_sites_.size(); ++i) {
    auto& site = gemm_sites_[i];
    if (site.acol >= 0) continue;  // A read through gather tables: the simple kernel
    // shared (stride-0) operands change the memory traffic: part of the key
    const int shared = (site.g.sao == 0) | (site.g.sai == 0) << 1 | (site.g.sbo == 0) << 2 | (site.g.sbi == 0) << 3;
    auto key = std::make_tuple(site.g.m, site.g.k, site.g.n, site.g.nbo, site.g.nbi, shared);
    auto it = best.find(key);
    if (it != best.end()) {
      gemm_choice_[i] = it->second;
      continue;
    }
    GemmArgs q = site.g;
    q.a = c.ptr(site.a), q.b = c.ptr(site.b), q.c = c.ptr(site.c);
    q.scratch = c.ws + offs_[scratch_buf_];
    if (site.bcol >= 0) q.brow = consts_ + site.brow, q.bcol = consts_ + site.bcol;
    if (site.amax.kind != Ref::NONE) q.amax = (const unsigned long long*)c.ptr(site.amax);
    if (site.bmax.kind != Ref::NONE) q.bmax = (const unsigned long long*)c.ptr(site.bmax);
    q.scratch_bytes = bufs_[scratch_buf_].bytes;
    float best_ms = 1e30f;
    int choice = GEMM_AUTO;
    const bool split_prec = o_.precision != NATIVE;
    const double macs = (double)q.nbo * (double)q.nbi * (double)q.m * (double)q.n * (double)q.k;
    const double outs = (double)q.nbo * (double)q.nbi * (double)q.m * (double)q.n;
    // the heuristic's own choice first: another kernel replaces it only when at
    // least 3% faster, so near ties do not flip from run to run
    for (int cand : {GEMM_AUTO, GEMM_SIMPLE, GEMM_SKINNY, GEMM_SPLITK, GEMM_PRESPLIT, GEMM_PRESPLIT_A, GEMM_MFMA,
                     GEMM_MFMA_R32, GEMM_MFMA_R64, GEMM_MFMA_R128}) {
      // candidates that cannot win are not timed (the one-thread-per-output kernel
      // on large products, split-K with many outputs, row tiles far from m)
      if (cand == GEMM_SIMPLE && macs > (double)(1 << 26)) continue;
      if (cand == GEMM_SKINNY && q.m > 64 && q.n > 32) continue;
      if (cand == GEMM_SPLITK && (q.k < 256 || outs > (double)(1 << 16))) continue;
      // split products only where they can pay off (the pre-split kernels have
      // 32-row tiles for short A; GEMM_MFMA splits while staging into 128-row tiles)
      if ((cand == GEMM_PRESPLIT || cand == GEMM_PRESPLIT_A) &&
          !split_products_apply(o_.dtype, o_.precision, q.m, q.k, q.n))
        continue;
      if (cand == GEMM_MFMA && (!split_prec || q.m < 64)) continue;
      if (cand == GEMM_MFMA_R32 && q.m > 64) continue;
      if (cand == GEMM_MFMA_R64 && q.m > 256) continue;
      if (cand == GEMM_MFMA_R128 && q.m < 64) continue;
      if ((o_.dtype == F64 || o_.dtype == C128) && cand == GEMM_MFMA_R128) continue;
      launch_gemm(o_.dtype, o_.precision, cand, q, s);
      ATN_HIP(hipEventRecord(e0, s));
      for (int r = 0; r < 3; ++r) launch_gemm(o_.dtype, o_.precision, cand, q, s);
      ATN_HIP(hipEventRecord(e1, s));
      ATN_HIP(hipEventSynchronize(e1));
      float ms = 0;
      ATN_HIP(hipEventElapsedTime(&ms, e0, e1));
      if (cand == GEMM_AUTO) ms *= 0.97f;
      if (ms < best_ms) best_ms = ms, choice = cand;
    }
    gemm_choice_[i] = best[key] = choice;
  }
  ATN_HIP(hipEventDestroy(e0));
  ATN_HIP(hipEventDestroy(e1));
  tuned_ = true;
}

void Executor::run(const void* const* inputs, void* out, hipStream_t s, Report* rep) {
  const Plan& P = *P_;
  Ctx c;
  c.ws = ws_;
  c.offs = &offs_;
  c.inputs = inputs;
  c.out = out;
  c.s = s;
  sync_inputs(inputs);
  if (o_.autotune && o_.gemm_kernel == GEMM_AUTO && !tuned_) autotune(c, s);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (rep) {
    ATN_HIP(hipEventCreate(&e0));
    ATN_HIP(hipEventCreate(&e1));
    ATN_HIP(hipEventRecord(e0, s));
  }
  size_t l0 = g_launch_count;
  if (o_.use_graph) {
    std::vector<const void*> ins(inputs, inputs + P.net.inputs.size());
    if (!gexec_ || ins != graph_inputs_ || out != graph_out_) {
      if (gexec_) (void)hipGraphExecDestroy(gexec_), gexec_ = nullptr;
      if (graph_) (void)hipGraphDestroy(graph_), graph_ = nullptr;
      Ctx cc = c;
      cc.s = cap_stream_;
      size_t lc = g_launch_count;
      ATN_HIP(hipStreamBeginCapture(cap_stream_, hipStreamCaptureModeThreadLocal));
      try {
        enqueue(cc);
      } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(cap_stream_, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
      }
      ATN_HIP(hipStreamEndCapture(cap_stream_, &graph_));
      ATN_HIP(hipGraphInstantiate(&gexec_, graph_, nullptr, nullptr, 0));
      launches_per_run_ = g_launch_count - lc;
      graph_inputs_ = ins;
      graph_out_ = out;
    }
    ATN_HIP(hipGraphLaunch(gexec_, s));
  } else {
    enqueue(c);
    launches_per_run_ = g_launch_count - l0;
  }
  if (rep) {
    ATN_HIP(hipEventRecord(e1, s));
    ATN_HIP(hipEventSynchronize(e1));
    float ms = 0;
    ATN_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    rep->elapsed_ms = ms;
    double c8 = dtype_complex(o_.dtype) ? 8.0 : 2.0;
    rep->flops = c8 * P.macs_per_slice * (double)n_run_;
    rep->tflops = ms > 0 ? rep->flops / (ms * 1e-3) / 1e12 : 0;
    rep->slices_per_s = ms > 0 ? (double)n_run_ / (ms * 1e-3) : 0;
    rep->n_slices = n_run_;
    rep->batch = B_;
    rep->n_launches = launches_per_run_;
    rep->workspace_bytes = ws_bytes_;
    rep->log10_exponent = 0;
    if (o_.strip) {
      struct {
        int64_t pow2;
        double log10_last;
      } e{};
      ATN_HIP(hipMemcpy(&e, ws_ + offs_[expo_buf_], 16, hipMemcpyDeviceToHost));
      rep->log10_exponent = (double)e.pow2 * std::log10(2.0) + e.log10_last;
    }
  }
}

// --------------------------------------------------------------- stages ------
Scratch::~Scratch() {
  for (void* p : ptrs) (void)hipFreeAsync(p, s);
}

void* Scratch::get(size_t bytes) {
  bytes = align_up(std::max<size_t>(bytes, 256), 256);
  if (pool && pool->base) {
    if (used + bytes > pool->cap) throw Error("stage scratch pool too small", -3);
    void* p = pool->base + used;
    used += bytes;
    return p;
  }
  void* p = nullptr;
  ATN_HIP(hipMallocAsync(&p, bytes, s));
  ptrs.push_back(p);
  used += bytes;
  if (pool) pool->peak = std::max(pool->peak, used);
  return p;
}

// sums then permute of one operand; returns the pointer holding the GEMM operand
const void* stage_operand(Scratch& sc, int dt, const void* x, std::vector<size_t> shape, const std::vector<size_t>& sums,
                    const std::vector<size_t>& perm) {
  int es = (int)dtype_size(dt);
  for (size_t ax : sums) {
    size_t outer = 1, inner = 1, d = shape[ax];
    for (size_t q = 0; q < ax; ++q) outer *= shape[q];
    for (size_t q = ax + 1; q < shape.size(); ++q) inner *= shape[q];
    void* dst = sc.get(outer * inner * es);
    launch_sum_axis(dt, outer, d, inner, x, dst, sc.s);
    shape.erase(shape.begin() + ax);
    x = dst;
  }
  PermPlan pp = plan_permute(shape, perm, 1, es);
  if (pp.kind == 0) return x;
  void* dst = sc.get(pp.numel * es);
  launch_permute(pp, x, dst, sc.s);
  return dst;
}

// inverse of `stage_operand`: g has the permuted shape; writes the gradient of x.
void stage_operand_adjoint(Scratch& sc, int dt, const void* g, std::vector<size_t> shape0, const std::vector<size_t>& sums,
                     const std::vector<size_t>& perm, void* out) {
  int es = (int)dtype_size(dt);
  std::vector<std::vector<size_t>> shapes{shape0};
  for (size_t ax : sums) {
    auto s = shapes.back();
    s.erase(s.begin() + ax);
    shapes.push_back(s);
  }
  auto& sh = shapes.back();
  std::vector<size_t> pshape;
  for (size_t ax : perm) pshape.push_back(sh[ax]);
  std::vector<size_t> inv(perm.size());
  for (size_t i = 0; i < perm.size(); ++i) inv[perm[i]] = i;
  PermPlan pp = plan_permute(pshape, inv, 1, es);
  const void* cur = g;
  if (sums.empty()) {
    if (pp.kind == 0) launch_copy(prod(sh) * es, g, out, sc.s);
    else launch_permute(pp, g, out, sc.s);
    return;
  }
  if (pp.kind != 0) {
    void* dst = sc.get(prod(sh) * es);
    launch_permute(pp, g, dst, sc.s);
    cur = dst;
  }
  for (size_t q = sums.size(); q-- > 0;) {
    auto& full = shapes[q];
    size_t ax = sums[q], outer = 1, inner = 1, d = full[ax];
    for (size_t i = 0; i < ax; ++i) outer *= full[i];
    for (size_t i = ax + 1; i < full.size(); ++i) inner *= full[i];
    void* dst = q == 0 ? out : sc.get(prod(full) * es);
    launch_broadcast_axis(es, outer, d, inner, cur, dst, sc.s);
    cur = dst;
  }
}

void stage_prep_forward(const Plan& p, size_t i, int dt, const void* in, void* out, hipStream_t s) {
  if (i >= p.cp.n_inputs) throw Error("input index out of range");
  const InputPrep& q = p.cp.preps[i];
  int es = (int)dtype_size(dt);
  Scratch sc(s);
  auto shape = q.input_shape;
  const void* cur = in;
  if (q.traces.empty()) {
    launch_copy(prod(shape) * es, in, out, s);
    return;
  }
  for (size_t t = 0; t < q.traces.size(); ++t) {
    size_t ai = std::get<0>(q.traces[t]), aj = std::get<1>(q.traces[t]);
    bool diag = std::get<2>(q.traces[t]);
    size_t A = 1, d = shape[ai], Bm = 1, C = 1;
    for (size_t k = 0; k < ai; ++k) A *= shape[k];
    for (size_t k = ai + 1; k < aj; ++k) Bm *= shape[k];
    for (size_t k = aj + 1; k < shape.size(); ++k) C *= shape[k];
    void* dst = t + 1 == q.traces.size() ? out : sc.get((diag ? A * d * Bm * C : A * Bm * C) * es);
    if (diag) launch_diag(es, A, d, Bm, C, cur, dst, s);
    else launch_trace(dt, A, d, Bm, C, cur, dst, s);
    shape.erase(shape.begin() + aj);
    if (!diag) shape.erase(shape.begin() + ai);
    cur = dst;
  }
}

void stage_prep_adjoint(const Plan& p, size_t i, int dt, const void* g, void* gin, hipStream_t s) {
  if (i >= p.cp.n_inputs) throw Error("input index out of range");
  const InputPrep& q = p.cp.preps[i];
  int es = (int)dtype_size(dt);
  Scratch sc(s);
  std::vector<std::vector<size_t>> shapes{q.input_shape};
  for (auto& t : q.traces) {
    auto sh = shapes.back();
    sh.erase(sh.begin() + std::get<1>(t));
    if (!std::get<2>(t)) sh.erase(sh.begin() + std::get<0>(t));
    shapes.push_back(sh);
  }
  if (q.traces.empty()) {
    launch_copy(prod(q.input_shape) * es, g, gin, s);
    return;
  }
  const void* cur = g;
  for (size_t t = q.traces.size(); t-- > 0;) {
    auto& full = shapes[t];
    size_t ai = std::get<0>(q.traces[t]), aj = std::get<1>(q.traces[t]);
    bool diag = std::get<2>(q.traces[t]);
    size_t A = 1, d = full[ai], Bm = 1, C = 1;
    for (size_t k = 0; k < ai; ++k) A *= full[k];
    for (size_t k = ai + 1; k < aj; ++k) Bm *= full[k];
    for (size_t k = aj + 1; k < full.size(); ++k) C *= full[k];
    void* dst = t == 0 ? gin : sc.get(prod(full) * es);
    launch_fill_zero(dst, prod(full) * es, s);
    launch_diag_scatter(es, !diag, A, d, Bm, C, cur, dst, s);
    cur = dst;
  }
}

void stage_step_forward(const Plan& p, size_t step, int dt, int prec, const void* l, const void* r, void* out, hipStream_t s) {
  if (step >= p.cp.steps.size()) throw Error("step index out of range");
  const PairStep& st = p.cp.steps[step];
  Scratch sc(s);
  const void* a = stage_operand(sc, dt, l, st.left_shape, st.sum_a, st.a_perm);
  const void* b = stage_operand(sc, dt, r, st.right_shape, st.sum_b, st.b_perm);
  GemmArgs g{};
  g.a = a, g.b = b, g.c = out;
  g.m = st.m, g.k = st.k, g.n = st.n;
  g.lda = g.k, g.ldb = g.n, g.ldc = g.n;
  g.nbo = 1, g.nbi = st.bt;
  g.sai = g.m * g.k, g.sbi = g.k * g.n, g.sci = g.m * g.n;
  g.scratch_bytes = gemm_scratch_bytes(dt, prec, GEMM_AUTO, g);
  g.scratch = sc.get(g.scratch_bytes);
  launch_gemm(dt, prec, GEMM_AUTO, g, s);
}

void stage_step_adjoint(const Plan& p, size_t step, int dt, const void* l, const void* r, const void* gout, void* gl, void* gr,
                        hipStream_t s) {
  if (step >= p.cp.steps.size()) throw Error("step index out of range");
  const PairStep& st = p.cp.steps[step];
  int es = (int)dtype_size(dt);
  Scratch sc(s);
  const void* a2 = stage_operand(sc, dt, l, st.left_shape, st.sum_a, st.a_perm);
  const void* b2 = stage_operand(sc, dt, r, st.right_shape, st.sum_b, st.b_perm);
  int64_t bt = st.bt, m = st.m, k = st.k, n = st.n;
  if (gl) {
    // gA2[bt, m, k] = G[bt, m, n] . B2^H  (B2 stored [k, n])
    void* ga2 = sc.get(bt * m * k * es);
    GemmArgs g{};
    g.a = gout, g.b = b2, g.c = ga2;
    g.m = m, g.k = n, g.n = k;
    g.lda = n, g.ldb = n, g.ldc = k;
    g.nbo = 1, g.nbi = bt;
    g.sai = m * n, g.sbi = k * n, g.sci = m * k;
    launch_gemm_adjoint(dt, 0, 1, g, s);
    stage_operand_adjoint(sc, dt, ga2, st.left_shape, st.sum_a, st.a_perm, gl);
  }
  if (gr) {
    // gB2[bt, k, n] = A2^H . G  (A2 stored [m, k])
    void* gb2 = sc.get(bt * k * n * es);
    GemmArgs g{};
    g.a = a2, g.b = gout, g.c = gb2;
    g.m = k, g.k = m, g.n = n;
    g.lda = k, g.ldb = n, g.ldc = n;
    g.nbo = 1, g.nbi = bt;
    g.sai = m * k, g.sbi = m * n, g.sci = k * n;
    launch_gemm_adjoint(dt, 1, 0, g, s);
    stage_operand_adjoint(sc, dt, gb2, st.right_shape, st.sum_b, st.b_perm, gr);
  }
}

void stage_finalize_forward(const Plan& p, int dt, const void* in, void* out, hipStream_t s) {
  const Finalize& f = p.cp.fin;
  Scratch sc(s);
  int es = (int)dtype_size(dt);
  const void* x = stage_operand(sc, dt, in, f.input_shape, f.sum, f.perm);
  launch_copy(prod(f.output_shape) * es, x, out, s);
}

void stage_finalize_adjoint(const Plan& p, int dt, const void* g, void* gin, hipStream_t s) {
  const Finalize& f = p.cp.fin;
  Scratch sc(s);
  stage_operand_adjoint(sc, dt, g, f.input_shape, f.sum, f.perm, gin);
}

}  // namespace atn
