// Batched 2-D steps: several simulations of one order advanced by one launch
// per stage of the step.  Every kernel of the 2-D FV + fallback step (the
// device-side loop's dt, the predictor iterations, the final element faces,
// the sp <-> cv transforms, per substep the high-order flux + candidate +
// first-half detection kernel, the smooth-extrema pass, theta, the
// column-strip fallback, the wave-speed maximum and the time advance) is
// launched once for all members: the grid is the concatenation of the
// members' own grids, and a block finds the member it works for in a prefix
// table of the members' block counts (a scan of up to 64 entries loaded at
// once, a binary search above), takes that member's arguments (geometry, parameters,
// buffers) from a device array and runs the kernel's body with its place in
// the member's grid.  Each member therefore computes exactly what its own
// launches compute (same kernels, same blocks, same order of operations), and
// a member whose device loop has finished (its skip flag set) leaves its
// blocks at once.  The operator matrices depend on the order only and are a
// kernel argument.  Included into spd_step.hip.
#pragma once

namespace spd {
namespace batch {

// Member of a batched launch and the block's place in the member's grid of
// gx blocks per row.  Up to 64 members the prefix table is read with
// independent uniform loads (its allocation holds at least 64 entries) and
// scanned; larger batches take a binary search.
__device__ __forceinline__ int member_of(const int* __restrict__ start, int nm) {
  const int b = blockIdx.x;
  if (nm <= 64) {
    int m = 0;
#pragma unroll
    for (int i = 1; i < 64; ++i) {
      const int si = start[i];
      if (i < nm && si <= b) m = i;
    }
    return m;
  }
  int lo = 0, hi = nm - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (start[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}
__device__ __forceinline__ BlockIx place(const int* __restrict__ start, int m, int gx) {
  const int b = (int)blockIdx.x - start[m], nb = start[m + 1] - start[m];
  return BlockIx{b % gx, b / gx, gx, nb / gx};
}

// Per-member arguments of each batched kernel (gx: the member grid's blocks
// per row; 1-D grids have gx = their block count).
struct ARun {
  const double* cmax;
  double h, cfl;
  RunState* st;
  double* dt;
  int* skip;
  int* any_flag;       // zeroed before each substep's detection (null: no global gate)
  const double* partial;
  long long npartial;
  double* cmax_out;
};
template <typename S, typename B> struct APred {
  Geom g;
  Params prm;
  const S* src;
  int Tsrc;
  const double* Usp;
  const double* dt;
  B* out;
  int out_cmp, gx;
};
template <typename B> struct AFace {
  Geom g;
  Params prm;
  const B* src;
  B* Fx;
  B* Fy;
  int gx;
};
struct ATensor {
  Geom g;
  Params prm;
  const double* in;
  double* out;
  double* partial;
  int gx;
};
template <typename B> struct AHoc {
  Geom g;
  Params prm;
  const B* Uad;
  const B* Ffx;
  const B* Ffy;
  const double* Ucv;
  const double* dt;
  double* Uout;
  B* Fhx;
  B* Fhy;
  uint8_t* flags;
  int* any_flag;
  int Er, gx;
};
struct ASed {
  Geom g;
  Params prm;
  const double* Unew;
  uint8_t* flags;
  int gx;
};
struct ATheta {
  Geom g;
  Params prm;
  const uint8_t* flags;
  const double* Ucv;
  const double* Unew;
  const int* any_flag;
  double* theta;
  int gx;
  uint8_t* dirty;  // the member's theta tile marks (FbMarks::dirty)
};
template <typename B> struct AMarch {
  Geom g;
  Params prm;
  const double* Ucv;
  const double* theta;
  const B* Fx;
  const B* Fy;
  const double* dt;
  int MR;
  double* Uout;
  int gx;
};

// one block of 64 threads per member
__global__ void kb_run_dt(const ARun* __restrict__ a) {
  const ARun& A = a[blockIdx.x];
  k_run_dt_body(BlockIx{0, 0, 1, 1}, A.cmax, A.h, A.cfl, A.st, A.dt, A.skip);
}
__global__ void kb_run_advance(const ARun* __restrict__ a) {
  const ARun& A = a[blockIdx.x];
  k_run_advance_body(BlockIx{0, 0, 1, 1}, A.st, A.skip);
}
__global__ void kb_zero_gate(const ARun* __restrict__ a) {
  const ARun& A = a[blockIdx.x];
  if (A.any_flag && !*A.skip) k_zero_counters_body(BlockIx{0, 0, 1, 1}, A.any_flag);
}
__global__ void __launch_bounds__(1024) kb_reduce_max(const ARun* __restrict__ a) {
  const ARun& A = a[blockIdx.x];
  k_reduce_max_body(BlockIx{0, 0, 1, 1}, A.partial, A.npartial, A.cmax_out, A.skip);
}

template <int P, int RS, typename S, typename B, bool FIRST, bool SRC_CMP>
__global__ void __launch_bounds__(256)
kb_pred2d(Ops<P, B> op, const APred<S, B>* __restrict__ a, const int* __restrict__ start, int nm) {
  const int m = member_of(start, nm);
  const APred<S, B>& A = a[m];
  pred2d::k_pred2d_body<P, RS, S, B, FIRST, SRC_CMP>(place(start, m, A.gx), op, A.g, A.prm, A.src,
                                                     A.Tsrc, A.Usp, A.dt, A.out, A.out_cmp);
}
template <int P, int RS, typename B>
__global__ void __launch_bounds__(256)
kb_face_sd2d(Ops<P, B> op, const AFace<B>* __restrict__ a, const int* __restrict__ start, int nm) {
  const int m = member_of(start, nm);
  const AFace<B>& A = a[m];
  k_face_sd2d_body<P, RS, B, B>(place(start, m, A.gx), op, A.g, A.prm, A.src, A.Fx, A.Fy);
}
template <int P, int MODE>
__global__ void __launch_bounds__(256)
kb_tensor2d(Ops<P> op, const ATensor* __restrict__ a, const int* __restrict__ start, int nm) {
  const int m = member_of(start, nm);
  const ATensor& A = a[m];
  k_tensor2d_body<P, MODE>(place(start, m, A.gx), op, A.g, A.prm, A.in, A.out, nullptr, nullptr,
                           A.partial);
}
template <int P, typename B>
__global__ void __launch_bounds__(256)
kb_hocand2d(Ops<P, B> op, int t, double w_i, const AHoc<B>* __restrict__ a,
            const int* __restrict__ start, int nm) {
  const int m = member_of(start, nm);
  const AHoc<B>& A = a[m];
  fv2d::k_hocand2d_body<P, B, true>(place(start, m, A.gx), op, A.g, A.prm, t, A.Uad, A.Ffx, A.Ffy,
                                    A.Ucv, A.dt, w_i, A.Uout, A.Fhx, A.Fhy, A.flags, A.any_flag,
                                    nullptr, A.Er);
}
__global__ void __launch_bounds__(fv2d::DNT)
kb_sed2d(const ASed* __restrict__ a, const int* __restrict__ start, int nm) {
  const int m = member_of(start, nm);
  const ASed& A = a[m];
  fv2d::k_sed2d_body(place(start, m, A.gx), A.g, A.prm, A.Unew, A.flags);
}
__global__ void __launch_bounds__(fv2d::HNT)
kb_theta2d(const ATheta* __restrict__ a, const int* __restrict__ start, int nm) {
  const int m = member_of(start, nm);
  const ATheta& A = a[m];
  fv2d::k_theta2d_body(place(start, m, A.gx), A.g, A.prm, A.flags, A.Ucv, A.Unew, A.any_flag,
                       A.theta, fv2d::FbMarks{nullptr, 1, 1, 0, 0, A.dirty});
}
template <int RS, typename B, bool STORED>
__global__ void __launch_bounds__(fv2d::MW * fv2d::MWAVES)
kb_fbmarch2d(double w_i, const AMarch<B>* __restrict__ a, const int* __restrict__ start, int nm) {
  const int m = member_of(start, nm);
  const AMarch<B>& A = a[m];
  fv2d::k_fbmarch2d_body<RS, B, STORED>(place(start, m, A.gx), A.g, A.prm, A.Ucv, A.theta, A.Fx,
                                        A.Fy, A.dt, w_i, A.MR, A.Uout, nullptr);
}

}  // namespace batch
}  // namespace spd

// Host side: the launch plan of one step of every member, built once per
// batch (and per parity of the FV buffer exchange when the substep count is
// odd), replayed by spd_batch_run_steps.
struct spd_batch {
  int p = 0, precision = 0, T = 0;
  std::vector<spd_ctx*> mem;
  std::vector<double*> usp;
  std::vector<double*> ucv0, unew0;       // the FV buffers at parity 0
  std::vector<void*> dev;                 // device arrays of the plan
  std::vector<std::function<void(hipStream_t)>> plan[2];
  ~spd_batch() {
    for (void* q : dev) (void)hipFree(q);
  }
  // a device copy of host data (at least 256 bytes allocated), freed with the
  // batch
  template <typename X> const X* upload(const std::vector<X>& v, int* rc) {
    void* q = nullptr;
    const size_t bytes = std::max<size_t>(v.size() * sizeof(X), 256);
    if (hipMalloc(&q, bytes) != hipSuccess) {
      *rc = fail(SPD_E_NOMEM, "hipMalloc of a batch launch table");
      return nullptr;
    }
    dev.push_back(q);
    if (!v.empty() && hipMemcpy(q, v.data(), v.size() * sizeof(X), hipMemcpyHostToDevice) != hipSuccess) {
      *rc = fail(SPD_E_HIP, "hipMemcpy of a batch launch table");
      return nullptr;
    }
    return (const X*)q;
  }
};

namespace spd {
namespace batch {

template <int P, typename B> struct Plan2D {
  using OpsB = Ops<P, B>;
  static constexpr int V = 4, T = P + 1;

  // Prefix table of the members' block counts; returns the total.
  static const int* starts(spd_batch* b, const std::vector<long long>& nb, long long* total,
                           int* rc) {
    std::vector<int> st(nb.size() + 1, 0);
    long long acc = 0;
    for (size_t i = 0; i < nb.size(); ++i) {
      st[i] = (int)acc;
      acc += nb[i];
    }
    st[nb.size()] = (int)acc;
    *total = acc;
    return b->upload(st, rc);
  }

  // The plan of one step; `par` selects which of each member's two FV
  // buffers holds the state at the step's start.
  static int build(spd_batch* b, int par) {
    int rc = SPD_OK;
    const int nm = (int)b->mem.size();
    spd_ctx* c0 = b->mem[0];
    const OpsB opb = make_ops<P, B>(*c0->ops);
    const Ops<P> op = make_ops<P>(*c0->ops);
    const int m = c0->cfg.m;
    auto& plan = b->plan[par];
    plan.clear();
    auto pick = [&](int i, bool cv) -> double* {
      return ((par == 0) == cv) ? b->ucv0[i] : b->unew0[i];
    };
    // the FV state entering substep k (the buffers swap after every substep)
    auto ucv = [&](int i, int k) { return pick(i, (k % 2) == 0); };
    auto unew = [&](int i, int k) { return pick(i, (k % 2) != 0); };
    auto dst = [&](int i, int it) -> B* { return (it & 1) ? (B*)b->mem[i]->Uad2 : (B*)b->mem[i]->Uad; };
    auto geom = [&](int i) {
      Geom g = b->mem[i]->g;
      g.skip = b->mem[i]->run_skip;
      return g;
    };

    // device-loop scalars
    std::vector<ARun> run(nm);
    for (int i = 0; i < nm; ++i) {
      spd_ctx* c = b->mem[i];
      run[i] = ARun{c->cmax, c->h_dt, c->cfg.cfl_coeff, c->run, c->run_dt, c->run_skip,
                    c->need_gate ? c->any_flag : nullptr, c->partial, t2d_blocks(c->g, P), c->cmax};
    }
    const ARun* drun = b->upload(run, &rc);
    if (rc) return rc;
    plan.push_back([=](hipStream_t s) {
      hipLaunchKernelGGL(kb_run_dt, dim3(nm), dim3(64), 0, s, drun);
    });

    // predictor iterations (k_pred2d: the first, compact when another follows)
    const int rs_sd = c0->cfg.riemann_sd;
    auto pred = [&](auto first_c, auto cmp_c, auto src_tag, int it) -> int {
      constexpr bool FIRST = decltype(first_c)::value, SRC_CMP = decltype(cmp_c)::value;
      using S = typename decltype(src_tag)::type;
      using L = pred2d::Layout<P, FIRST>;
      std::vector<APred<S, B>> a(nm);
      std::vector<long long> nb(nm);
      for (int i = 0; i < nm; ++i) {
        spd_ctx* c = b->mem[i];
        const int ntx = (c->g.N[0] + L::TX - 1) / L::TX, nty = (c->g.N[1] + L::TY - 1) / L::TY;
        const S* src = FIRST ? (const S*)b->usp[i] : (const S*)dst(i, it - 1);
        const int cmp = (FIRST && m >= 2) ? 1 : 0;
        a[i] = APred<S, B>{geom(i), c->prm, src, FIRST ? 1 : T, b->usp[i], c->run_dt, dst(i, it),
                           cmp, ntx * nty};
        nb[i] = (long long)ntx * nty;
      }
      long long tot = 0;
      const APred<S, B>* da = b->upload(a, &rc);
      const int* st = rc ? nullptr : starts(b, nb, &tot, &rc);
      if (rc) return rc;
      const size_t lds = (size_t)L::TOTAL * sizeof(B);
      plan.push_back([=](hipStream_t s) {
        if (rs_sd != SPD_RS_LLF)
          hipLaunchKernelGGL((kb_pred2d<P, RS_HLLC, S, B, FIRST, SRC_CMP>), dim3((unsigned)tot),
                             dim3(L::NTH), lds, s, opb, da, st, nm);
        else
          hipLaunchKernelGGL((kb_pred2d<P, RS_LLF, S, B, FIRST, SRC_CMP>), dim3((unsigned)tot),
                             dim3(L::NTH), lds, s, opb, da, st, nm);
      });
      return SPD_OK;
    };
    struct TD { using type = double; };
    struct TB { using type = B; };
    using T_ = std::true_type;
    using F_ = std::false_type;
    if ((rc = pred(T_{}, F_{}, TD{}, 0))) return rc;
    for (int it = 1; it < m; ++it) {
      if (it == 1 && m >= 2) rc = pred(F_{}, T_{}, TB{}, it);
      else rc = pred(F_{}, F_{}, TB{}, it);
      if (rc) return rc;
    }
    // final element faces (k_face_sd2d) of the last iterate
    {
      using FT = FaceTile2<P>;
      std::vector<AFace<B>> a(nm);
      std::vector<long long> nb(nm);
      for (int i = 0; i < nm; ++i) {
        spd_ctx* c = b->mem[i];
        const long long tiles =
            (long long)((c->g.N[0] + FT::EX - 1) / FT::EX) * ((c->g.N[1] + FT::EY - 1) / FT::EY);
        nb[i] = T * tiles;
        a[i] = AFace<B>{geom(i), c->prm, dst(i, m - 1), (B*)c->Fface[0], (B*)c->Fface[1],
                        (int)nb[i]};
      }
      long long tot = 0;
      const AFace<B>* da = b->upload(a, &rc);
      const int* st = rc ? nullptr : starts(b, nb, &tot, &rc);
      if (rc) return rc;
      plan.push_back([=](hipStream_t s) {
        if (rs_sd != SPD_RS_LLF)
          hipLaunchKernelGGL((kb_face_sd2d<P, RS_HLLC, B>), dim3((unsigned)tot), dim3(256), 0, s,
                             opb, da, st, nm);
        else
          hipLaunchKernelGGL((kb_face_sd2d<P, RS_LLF, B>), dim3((unsigned)tot), dim3(256), 0, s,
                             opb, da, st, nm);
      });
    }
    // sp -> cv (k_tensor2d<P, 1>) and cv -> sp (k_tensor2d<P, 0>)
    auto tensor = [&](int mode) -> int {
      std::vector<ATensor> a(nm);
      std::vector<long long> nb(nm);
      for (int i = 0; i < nm; ++i) {
        spd_ctx* c = b->mem[i];
        nb[i] = t2d_blocks(c->g, P);
        a[i] = mode == 1 ? ATensor{geom(i), c->prm, b->usp[i], ucv(i, 0), nullptr, (int)nb[i]}
                         : ATensor{geom(i), c->prm, ucv(i, T), b->usp[i], c->partial, (int)nb[i]};
      }
      long long tot = 0;
      const ATensor* da = b->upload(a, &rc);
      const int* st = rc ? nullptr : starts(b, nb, &tot, &rc);
      if (rc) return rc;
      plan.push_back([=](hipStream_t s) {
        if (mode == 1)
          hipLaunchKernelGGL((kb_tensor2d<P, 1>), dim3((unsigned)tot), dim3(256), 0, s, op, da, st,
                             nm);
        else
          hipLaunchKernelGGL((kb_tensor2d<P, 0>), dim3((unsigned)tot), dim3(256), 0, s, op, da, st,
                             nm);
      });
      return SPD_OK;
    };
    if ((rc = tensor(1))) return rc;
    // FV substeps
    const bool gate = c0->need_gate, sed = c0->prm.p > 1 && c0->prm.sed;
    const bool hllc = c0->cfg.riemann_fv != SPD_RS_LLF, stored = c0->g.stored != 0;
    for (int k = 0; k < T; ++k) {
      const double w = c0->ops->w_tp[k];
      if (gate)
        plan.push_back([=](hipStream_t s) {
          hipLaunchKernelGGL(kb_zero_gate, dim3(nm), dim3(64), 0, s, drun);
        });
      {
        using HT = fv2d::HocTile<P>;
        std::vector<AHoc<B>> a(nm);
        std::vector<long long> nb(nm);
        for (int i = 0; i < nm; ++i) {
          spd_ctx* c = b->mem[i];
          const int v = c->launch[SPD_LAUNCH_HOC_EPB];
          const int Eb = (v > 0 && v < HT::E) ? v : HT::E;
          nb[i] = (long long)c->g.N[1] * ((c->g.N[0] + Eb - 1) / Eb);
          a[i] = AHoc<B>{geom(i), c->prm, dst(i, m - 1), (B*)c->Fface[0], (B*)c->Fface[1],
                         ucv(i, k), c->run_dt, unew(i, k), (B*)c->Fho[0], (B*)c->Fho[1],
                         c->flags, c->need_gate ? c->any_flag : nullptr, Eb, (int)nb[i]};
        }
        long long tot = 0;
        const AHoc<B>* da = b->upload(a, &rc);
        const int* st = rc ? nullptr : starts(b, nb, &tot, &rc);
        if (rc) return rc;
        const size_t lds = (size_t)HT::template lds_bytes_det<B>(gate ? 4 : 2);
        plan.push_back([=](hipStream_t s) {
          hipLaunchKernelGGL((kb_hocand2d<P, B>), dim3((unsigned)tot), dim3(256), lds, s, opb, k, w,
                             da, st, nm);
        });
      }
      if (sed) {
        std::vector<ASed> a(nm);
        std::vector<long long> nb(nm);
        for (int i = 0; i < nm; ++i) {
          spd_ctx* c = b->mem[i];
          const int gx = (c->g.Ns[0] + fv2d::DTX - 1) / fv2d::DTX;
          nb[i] = (long long)gx * ((c->g.Ns[1] + fv2d::DTY - 1) / fv2d::DTY);
          a[i] = ASed{geom(i), c->prm, unew(i, k), c->flags, gx};
        }
        long long tot = 0;
        const ASed* da = b->upload(a, &rc);
        const int* st = rc ? nullptr : starts(b, nb, &tot, &rc);
        if (rc) return rc;
        plan.push_back([=](hipStream_t s) {
          hipLaunchKernelGGL(kb_sed2d, dim3((unsigned)tot), dim3(fv2d::DNT), 0, s, da, st, nm);
        });
      }
      {
        std::vector<ATheta> a(nm);
        std::vector<long long> nb(nm);
        for (int i = 0; i < nm; ++i) {
          spd_ctx* c = b->mem[i];
          const int gx = (c->g.Ng[0] + fv2d::HTX - 1) / fv2d::HTX;
          nb[i] = (long long)gx * ((c->g.Ng[1] + fv2d::HTY - 1) / fv2d::HTY);
          a[i] = ATheta{geom(i), c->prm, c->flags, ucv(i, k), unew(i, k),
                        c->any_flag, c->theta, gx, c->theta_dirty};
        }
        long long tot = 0;
        const ATheta* da = b->upload(a, &rc);
        const int* st = rc ? nullptr : starts(b, nb, &tot, &rc);
        if (rc) return rc;
        plan.push_back([=](hipStream_t s) {
          hipLaunchKernelGGL(kb_theta2d, dim3((unsigned)tot), dim3(fv2d::HNT), 0, s, da, st, nm);
        });
      }
      {
        std::vector<AMarch<B>> a(nm);
        std::vector<long long> nb(nm);
        for (int i = 0; i < nm; ++i) {
          spd_ctx* c = b->mem[i];
          const int rows = fv2d::march_rows(c->g.Ns[0], c->g.Ns[1]);
          const int gx = (c->g.Ns[0] + fv2d::MOUT - 1) / fv2d::MOUT;
          nb[i] = (long long)gx *
                  ((c->g.Ns[1] + rows * fv2d::MWAVES - 1) / (rows * fv2d::MWAVES));
          a[i] = AMarch<B>{geom(i), c->prm, ucv(i, k), c->theta, (const B*)c->Fho[0],
                           (const B*)c->Fho[1], c->run_dt, rows, unew(i, k), gx};
        }
        long long tot = 0;
        const AMarch<B>* da = b->upload(a, &rc);
        const int* st = rc ? nullptr : starts(b, nb, &tot, &rc);
        if (rc) return rc;
        plan.push_back([=](hipStream_t s) {
          const dim3 grid((unsigned)tot), blk(fv2d::MW * fv2d::MWAVES);
          if (hllc) {
            if (stored)
              hipLaunchKernelGGL((kb_fbmarch2d<RS_HLLC, B, true>), grid, blk, 0, s, w, da, st, nm);
            else
              hipLaunchKernelGGL((kb_fbmarch2d<RS_HLLC, B, false>), grid, blk, 0, s, w, da, st, nm);
          } else {
            if (stored)
              hipLaunchKernelGGL((kb_fbmarch2d<RS_LLF, B, true>), grid, blk, 0, s, w, da, st, nm);
            else
              hipLaunchKernelGGL((kb_fbmarch2d<RS_LLF, B, false>), grid, blk, 0, s, w, da, st, nm);
          }
        });
      }
    }
    if ((rc = tensor(0))) return rc;
    plan.push_back([=](hipStream_t s) {
      hipLaunchKernelGGL(kb_reduce_max, dim3(nm), dim3(1024), 0, s, drun);
      hipLaunchKernelGGL(kb_run_advance, dim3(nm), dim3(64), 0, s, drun);
    });
    return SPD_OK;
  }
};

}  // namespace batch
}  // namespace spd
