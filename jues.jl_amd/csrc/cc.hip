// Closed-shell coupled-cluster drivers on the device: RCCD (RCCD.jl / mRCCD.jl)
// and RCCSD (mRCCSD.jl / RCCSD.jl) with DIIS, in FP64 or FP32-MFMA contractions.
//
// Layouts (all column-major, FP64 master amplitudes):
//   ijab  t[i,j,a,b]                         (the Julia T2 order)
//   ov    X[i,a,j,b], an (ov) x (ov) matrix  (ring terms)
//   packed T+/T-, V+/V- over pairs i<=j, e<=f (i<j, e<f) for the pp-ladder
//
// RCCD residual (RCCD.jl:242-256 with Wabef eliminated, the v^4 intermediate is
// never formed):
//   R = <ij|ab> + S[ia,jb] + sum_mn t[mn,ab] W'[mn,ij] + sum_ef t[ij,ef] <ab|ef>
//   W' = 2 Wmnij - <mn|ij> = <mn|ij> + sum_ef <mn|ef> t[ij,ef]
//   S = H + H^T with H = U WJ + T WE + (Tx WE)^x (ov layout; the update kernel sums it from
//   the two products' outputs)
//   WJ = K + K U/2 - Kx T/2,  WE = -J + Kx Tx/2 + D/2,  U = 2T - Tx
//   D[kc,jb] = delta_kj Fae[b,c] - delta_cb Fmi[k,j]  (T D + (Tx D)^x + transposes = the
//   Fock-like terms t.Fae^T - Fmi^T.t and their transposes)
// where K[ia,jb] = (ia|jb), Kx[ia,jb] = (ib|ja), J[ia,jb] = (ij|ab),
// T[ia,jb] = t[i,j,a,b], Tx[ia,jb] = t[i,j,b,a] and X^x[ia,jb] = X[ib,ja].
#include "cc.h"
#include "api_util.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>

namespace jues {

void CCDeleter::operator()(CCState* p) const { delete p; }

namespace {

template <typename T>
T* bufp(DevBuf& b, size_t n) {
  return bufp_t<T>(b, n);
}

void build_pairs(jues_ctx* ctx, CCState& S) {
  auto mk = [&](int n, bool strict, DevBuf& buf) -> long {
    std::vector<int2> h;
    for (int q = 0; q < n; ++q)
      for (int p = 0; p < (strict ? q : q + 1); ++p) h.push_back(make_int2(p, q));
    if (h.empty()) h.push_back(make_int2(0, 0));
    buf.alloc(h.size() * sizeof(int2));
    h2d(ctx, buf.p, h.data(), h.size() * sizeof(int2));
    return strict ? (long)n * (n - 1) / 2 : (long)n * (n + 1) / 2;
  };
  S.P.npo = mk(S.o, false, S.pair_buf[0]);
  S.P.nmo = mk(S.o, true, S.pair_buf[1]);
  S.P.npv = mk(S.v, false, S.pair_buf[2]);
  S.P.nmv = mk(S.v, true, S.pair_buf[3]);
  S.P.o_le = S.pair_buf[0].as<int2>();
  S.P.o_lt = S.pair_buf[1].as<int2>();
  S.P.v_le = S.pair_buf[2].as<int2>();
  S.P.v_lt = S.pair_buf[3].as<int2>();
}

// chemist (vv|vv) slabs -> packed V+/V- (FP64); dense route transforms the whole block
void build_vvvv(jues_ctx* ctx, CCState& S, CCInts<double>& I, const double* Cv) {
  const int v = S.v;
  const long v2 = (long)v * v;
  double* Vp = bufp<double>(I.Vp, (size_t)(S.P.npv * S.P.npv));
  double* Vm = bufp<double>(I.Vm, (size_t)std::max(1L, S.P.nmv * S.P.nmv));
  if (ctx->gao.p) {
    DevBuf X((size_t)v2 * v2 * sizeof(double));
    tei_transform_dev<double>(ctx, Cv, v, Cv, v, Cv, v, Cv, v, X.as<double>());
    pack_v<double, double>(ctx, S.P, X.as<double>(), v, 0, v, Vp, Vm);
    sync(ctx);
  } else {
    const long naux = ctx->naux;
    DevBuf Lvv((size_t)v2 * naux * sizeof(double));
    mo_factors_dev<double>(ctx, Cv, v, Cv, v, Lvv.as<double>());
    // slab of nf virtual indices f: X[a,e,b,f] = sum_P Lvv[ae,P] Lvv[bf,P] (~2 GB per slab)
    long nf = std::max(1L, std::min<long>(v, (2L << 30) / (long)(v2 * v * sizeof(double))));
    DevBuf X((size_t)(v2 * v * nf) * sizeof(double));
    for (int f0 = 0; f0 < v; f0 += (int)nf) {
      const int f1 = (int)std::min<long>(v, f0 + nf);
      gemm<double>(ctx, false, true, v2, (long)v * (f1 - f0), naux, 1.0, Lvv.as<double>(), v2, 0,
                   Lvv.as<double>() + (long)v * f0, v2, 0, 0.0, X.as<double>(), v2, 0, 1, "setup_vvvv");
      pack_v<double, double>(ctx, S.P, X.as<double>(), v, f0, f1, Vp, Vm);
    }
    sync(ctx);
  }
}

// MO factors of the (vv|vv) block for the recomputed ladder: L[ae,P] = (Cv^T B_P Cv)[a,e],
// and L[ea,P] (no symmetry of B_P in the AO pair is assumed)
void build_lvv(jues_ctx* ctx, CCInts<double>& I, const double* Cv, int v) {
  const long naux = ctx->naux;
  double* L = bufp<double>(I.Lvv, (size_t)v * v * naux);
  mo_factors_dev<double>(ctx, Cv, v, Cv, v, L);
  double* Lt = bufp<double>(I.Lvvt, (size_t)v * v * naux);
  permute<double, double>(ctx, L, {v, v, naux}, {1, 0, 2}, Lt, 1.0, 0.0, "setup_vvvv");
}

// Whether <ab|ef> is rebuilt every iteration.  Auto: when the packed V+/V- (with its FP32
// or split copy) and the rest of the iteration state would not fit in free device memory.
bool ladder_recompute_mode(jues_ctx* ctx, const CCState& S, int mode, int prec) {
  JUES_REQUIRE(mode >= JUES_LADDER_AUTO && mode <= JUES_LADDER_RECOMPUTE, "cc_setup: bad ladder mode");
  const bool factors = ctx->Bfac.p != nullptr;
  if (mode == JUES_LADDER_STORED) return false;
  if (mode == JUES_LADDER_RECOMPUTE) {
    JUES_REQUIRE(factors, "cc_setup: a recomputed ladder needs factored AO integrals (jues_set_ao_factors)");
    return true;
  }
  if (!factors) return false;
  size_t free_b = 0, total_b = 0;
  JUES_HIP(hipMemGetInfo(&free_b, &total_b));
  const double N2 = (double)S.o * S.o * S.v * S.v;
  const double vv = (double)S.P.npv * S.P.npv + (double)S.P.nmv * S.P.nmv;
  const int dm = S.opts.diis ? S.opts.diis_max : 1;
  const double stored = 8.0 * vv * (prec == JUES_FP64 ? 1.0 : 1.5);
  const double other = 8.0 * N2 * (10 + 2 * dm) + (prec == JUES_FP64 ? 8.0 : 4.0) * N2 * 16 +
                       (prec == JUES_FP16X3_MIXED ? 4.0 * N2 * 9 : 0.0);
  return stored + other > 0.85 * (double)free_b;
}

}  // namespace

// ---------------------------------------------------------------------------
// integral setup
// ---------------------------------------------------------------------------
// FP64 MO integral blocks.  With rebuild set only the blocks of S.i64 are (re)made, from the
// orbitals kept in S; the energy weights and the guess are left alone.
void cc_build_ints(jues_ctx* ctx, CCState& S, const double* Co, const double* Cv, bool rebuild = false) {
  const int o = S.o, v = S.v;
  const long ov = (long)o * v, N2 = ov * ov, oo = (long)o * o;
  auto& I = *S.i64;
  double* Kov = bufp<double>(I.Kov, N2);
  // setup-only blocks, freed on return: (ij|ab) at [i,j,a,b] and (mi|nj) at [m,i,n,j]
  DevBuf Jc_buf(N2 * sizeof(double)), Oc_buf(oo * oo * sizeof(double));
  double* Jc = Jc_buf.as<double>();
  double* Oc = Oc_buf.as<double>();
  if (ctx->gao.p) {
    tei_transform_dev<double>(ctx, Co, o, Cv, v, Co, o, Cv, v, Kov);   // (ia|jb) -> [i,a,j,b]
    tei_transform_dev<double>(ctx, Co, o, Co, o, Cv, v, Cv, v, Jc);    // (ij|ab) -> [i,j,a,b]
    tei_transform_dev<double>(ctx, Co, o, Co, o, Co, o, Co, o, Oc);    // (mi|nj) -> [m,i,n,j]
  } else {
    const long naux = ctx->naux;
    DevBuf Lov(ov * naux * sizeof(double)), Loo(oo * naux * sizeof(double)), Lvv((long)v * v * naux * sizeof(double));
    mo_factors_dev<double>(ctx, Co, o, Cv, v, Lov.as<double>());
    mo_factors_dev<double>(ctx, Co, o, Co, o, Loo.as<double>());
    mo_factors_dev<double>(ctx, Cv, v, Cv, v, Lvv.as<double>());
    gemm<double>(ctx, false, true, ov, ov, naux, 1.0, Lov.as<double>(), ov, 0, Lov.as<double>(), ov, 0, 0.0, Kov, ov,
                 0, 1, "setup_ints");
    gemm<double>(ctx, false, true, oo, (long)v * v, naux, 1.0, Loo.as<double>(), oo, 0, Lvv.as<double>(),
                 (long)v * v, 0, 0.0, Jc, oo, 0, 1, "setup_ints");
    gemm<double>(ctx, false, true, oo, oo, naux, 1.0, Loo.as<double>(), oo, 0, Loo.as<double>(), oo, 0, 0.0, Oc, oo,
                 0, 1, "setup_ints");
    S.k_factored = true;
    if (S.method == 1) {
      // (ij|ka) -> ooov chem [i,j,k,a]; (ia|bc) stays factored: Lov, its two permutations and Lvv
      double* ooov = bufp<double>(I.ooov_c, oo * ov);
      gemm<double>(ctx, false, true, oo, ov, naux, 1.0, Loo.as<double>(), oo, 0, Lov.as<double>(), ov, 0, 0.0, ooov,
                   oo, 0, 1, "setup_ints");
      S.ovvv_factored = true;
      permute<double, double>(ctx, Lov.as<double>(), {o, v, naux}, {1, 0, 2}, bufp<double>(I.LovT, ov * naux), 1.0,
                              0.0, "setup_ints");
      permute<double, double>(ctx, Lov.as<double>(), {o, v, naux}, {0, 2, 1}, bufp<double>(I.Lovp, ov * naux), 1.0,
                              0.0, "setup_ints");
      sync(ctx);
      if (!S.ladder_recompute) I.Lvv = std::move(Lvv);
    }
    sync(ctx);
    I.Lov = std::move(Lov);
  }
  if (ctx->gao.p && S.method == 1) {
    tei_transform_dev<double>(ctx, Co, o, Co, o, Co, o, Cv, v, bufp<double>(I.ooov_c, oo * ov));
    tei_transform_dev<double>(ctx, Co, o, Cv, v, Cv, v, Cv, v, bufp<double>(I.ovvv_c, ov * v * v));
  }
  // derived layouts
  permute<double, double>(ctx, Jc, {o, o, v, v}, {0, 2, 1, 3}, bufp<double>(I.Jov, N2), 1.0, 0.0);
  permute<double, double>(ctx, Oc, {o, o, o, o}, {0, 2, 1, 3}, bufp<double>(I.oooo, oo * oo), 1.0, 0.0);
  permute<double, double>(ctx, Kov, {o, v, o, v}, {0, 2, 1, 3}, bufp<double>(I.Kpp, N2), 1.0, 0.0);
  permute<double, double>(ctx, Kov, {o, v, o, v}, {0, 3, 2, 1}, bufp<double>(I.Ktov, N2), 1.0, 0.0);
  permute<double, double>(ctx, I.Kpp.as<double>(), {o, o, v, v}, {2, 0, 1, 3}, bufp<double>(I.Ka, N2), 1.0, 0.0);
  pack_khh(ctx, S.P, I.Kpp.as<double>(), o, v, bufp<double>(I.KP, S.P.npo * S.P.npv),
           bufp<double>(I.KM, std::max(1L, S.P.nmo * S.P.nmv)));
  // energy weights 2<ij|ab> - <ji|ab>
  if (!rebuild) {
    double* Le = bufp<double>(S.Lener, N2);
    permute<double, double>(ctx, I.Kpp.as<double>(), {o, o, v, v}, {0, 1, 2, 3}, Le, 2.0, 0.0);
    permute<double, double>(ctx, I.Kpp.as<double>(), {o, o, v, v}, {1, 0, 2, 3}, Le, -1.0, 1.0);
  }
  // guess helpers: (ij|ab) in ijab order
  if (!rebuild && S.opts.guess == JUES_GUESS_OVOV) {
    double* g = bufp<double>(S.guess_buf, N2);
    JUES_HIP(hipMemcpyAsync(g, Jc, N2 * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
  }
  if (S.ladder_recompute) build_lvv(ctx, I, Cv, v);
  else build_vvvv(ctx, S, I, Cv);
  sync(ctx);
}

template <typename T>
static void convert_buf(jues_ctx* ctx, DevBuf& src, DevBuf& dst) {
  if (!src.p) return;
  const size_t n = src.bytes / sizeof(double);
  bufp<T>(dst, n);
  convert<double, T>(ctx, src.as<double>(), dst.as<T>(), (long)n);
}

// FP32 copies of the contraction integrals (the FP64 set stays for refinement)
void cc_make_ints32(jues_ctx* ctx, CCState& S) {
  if (S.i32) return;
  S.i32.reset(new CCInts<float>());
  CCInts<double>& a = *S.i64;
  CCInts<float>& b = *S.i32;
  convert_buf<float>(ctx, a.Kov, b.Kov);
  convert_buf<float>(ctx, a.Ktov, b.Ktov);
  convert_buf<float>(ctx, a.Jov, b.Jov);
  convert_buf<float>(ctx, a.Kpp, b.Kpp);
  convert_buf<float>(ctx, a.Ka, b.Ka);
  convert_buf<float>(ctx, a.oooo, b.oooo);
  // Vp/Vm and KP/KM are converted by the first FP32 (non-split) ladder products
  convert_buf<float>(ctx, a.ooov_c, b.ooov_c);
  convert_buf<float>(ctx, a.ovvv_c, b.ovvv_c);
  convert_buf<float>(ctx, a.Lov, b.Lov);
  convert_buf<float>(ctx, a.LovT, b.LovT);
  convert_buf<float>(ctx, a.Lovp, b.Lovp);
  convert_buf<float>(ctx, a.Lvv, b.Lvv);
  convert_buf<float>(ctx, a.Lvvt, b.Lvvt);
  convert_buf<float>(ctx, a.fOO, b.fOO);
  convert_buf<float>(ctx, a.fVV, b.fVV);
  convert_buf<float>(ctx, a.fOV, b.fOV);
  sync(ctx);
}

// ---------------------------------------------------------------------------
// residual
// ---------------------------------------------------------------------------
template <typename T>
CCWork<T>& work(CCState& S);
template <>
CCWork<double>& work<double>(CCState& S) {
  if (!S.w64) S.w64.reset(new CCWork<double>());
  return *S.w64;
}
template <>
CCWork<float>& work<float>(CCState& S) {
  if (!S.w32) S.w32.reset(new CCWork<float>());
  return *S.w32;
}
template <typename T>
CCInts<T>& ints(CCState& S);
template <>
CCInts<double>& ints<double>(CCState& S) { return *S.i64; }
template <>
CCInts<float>& ints<float>(CCState& S) { return *S.i32; }

// ---------------------------------------------------------------------------
// ring and pp-ladder products (FP64, FP32, or split-FP16 in the fast stage)
// ---------------------------------------------------------------------------
namespace {

SplitState& split_state(CCState& S) {
  if (!S.sp) S.sp.reset(new SplitState());
  return *S.sp;
}

// planes of operand X (rows x cols, leading dimension rows; with `trans`, of X^T
// for X stored cols x rows) in `slot`.  Per-iteration slots are split on their
// first use in an iteration, constant slots once.
// Ttov holds the same values as Tov (a permutation of t), so its planes take Tov's scale when
// Tov was split earlier in the same iteration.
template <typename Ti>
const Split16& split_operand(jues_ctx* ctx, CCState& S, int slot, const Ti* X, long rows, long cols,
                             bool trans = false) {
  SplitState& Z = split_state(S);
  SplitState::Entry& e = Z.e[slot];
  const long stamp = slot < SPLIT_CONST ? (long)S.iter : -1;
  // a constant operand, once split, stays valid for the stage (its FP64 source may since have
  // been released to make room)
  if (stamp == -1 && e.stamp == -1 && e.src) return e.buf.s;
  if (e.src != static_cast<const void*>(X) || e.stamp != stamp) {
    const SplitState::Entry& tov = Z.e[SP_TOV];
    const int* same = slot == SP_TTOV && !trans && tov.stamp == stamp && tov.src ? tov.buf.s.sexp : nullptr;
    if (trans) split16t<Ti>(ctx, X, rows, cols, cols, e.buf, "cc_split");
    else split16<Ti>(ctx, X, rows, cols, rows, e.buf, "cc_split", same);
    e.src = X;
    e.stamp = stamp;
  }
  return e.buf.s;
}

}  // namespace

template <>
void ring_mm<double>(jues_ctx* ctx, CCState& S, double alpha, const double* A, int, const double* B, int, double beta,
                     double* C) {
  const long ov = (long)S.o * S.v;
  gemm<double>(ctx, false, false, ov, ov, ov, alpha, A, ov, 0, B, ov, 0, beta, C, ov, 0, 1, "cc_ring");
}

template <>
void ring_mm<float>(jues_ctx* ctx, CCState& S, float alpha, const float* A, int slotA, const float* B, int slotB,
                    float beta, float* C) {
  const long ov = (long)S.o * S.v;
  if (!S.split_on) {
    gemm<float>(ctx, false, false, ov, ov, ov, alpha, A, ov, 0, B, ov, 0, beta, C, ov, 0, 1, "cc_ring");
    return;
  }
  // A is symmetric, so A B = A^T B: its planes serve as the TN kernel's left operand
  const Split16& a = split_operand<float>(ctx, S, slotA, A, ov, ov);
  const Split16& b = split_operand<float>(ctx, S, slotB, B, ov, ov);
  gemm16_tn(ctx, ov, ov, alpha, a, b, beta, C, ov, "cc_ring");
}

template <typename T>
void ring_k_mm(jues_ctx* ctx, CCState& S, T alpha, const T* B, int slotB, T* C) {
  if (!S.k_factored) {
    ring_mm<T>(ctx, S, alpha, ints<T>(S).Kov.template as<T>(), SP_K, B, slotB, T(1), C);
    return;
  }
  const long ov = (long)S.o * S.v, naux = ctx->naux;
  const T* L = ints<T>(S).Lov.template as<T>();
  T* UL = bufp<T>(work<T>(S).UL, (size_t)(ov * naux));
  if constexpr (sizeof(T) == 4) {
    if (S.split_on) {
      // UL = B^T Lov (B symmetric: its planes are the TN kernel's left operand, shared with the
      // later B WJ product); then C += alpha (Lov^T)^T (UL^T), both k = naux operands transposed
      SplitState& Z = split_state(S);
      const Split16& b = split_operand<float>(ctx, S, slotB, B, ov, ov);
      const Split16& l = split_operand<double>(ctx, S, SP_LOV, S.i64->Lov.as<double>(), ov, naux);
      gemm16_tn(ctx, ov, naux, 1.0f, b, l, 0.0f, UL, ov, "cc_ring_lr");
      const Split16& lt = split_operand<double>(ctx, S, SP_LOVT, S.i64->Lov.as<double>(), naux, ov, true);
      split16t<float>(ctx, UL, naux, ov, ov, Z.ult, "cc_split");
      gemm16_tn(ctx, ov, ov, alpha, lt, Z.ult.s, 1.0f, C, ov, "cc_ring_lr");
      return;
    }
  }
  gemm<T>(ctx, false, false, ov, naux, ov, T(1), B, ov, 0, L, ov, 0, T(0), UL, ov, 0, 1, "cc_ring_lr");
  gemm<T>(ctx, false, true, ov, ov, naux, alpha, L, ov, 0, UL, ov, 0, T(1), C, ov, 0, 1, "cc_ring_lr");
}

template void ring_k_mm<double>(jues_ctx*, CCState&, double, const double*, int, double*);
template void ring_k_mm<float>(jues_ctx*, CCState&, float, const float*, int, float*);

template <typename T>
void ovvv_z_slab(jues_ctx* ctx, CCState& S, int m0, int cnt, const T* LovT, const T* Lvv, const T* Tpp, T* Zs,
                 T* Zsp, T* Z) {
  const int o = S.o, v = S.v;
  const long ov = (long)o * v, vv = (long)v * v, oo = (long)o * o, naux = ctx->naux, nr = (long)v * cnt;
  bool split = false;
  if constexpr (sizeof(T) == 4) split = S.split_on;
  if constexpr (sizeof(T) == 4) {
    if (split) {
      // Zs = (LovT slab) Lvv^T from the factors' planes (k = P), split once per stage
      const Split16& lat = split_operand<double>(ctx, S, SP_LOVAT, S.i64->LovT.as<double>(), naux, ov, true);
      const Split16& lvt = split_operand<double>(ctx, S, SP_LVVT, S.i64->Lvv.as<double>(), naux, vv, true);
      gemm16_tn(ctx, nr, vv, 1.0f, split16_cols(lat, (long)v * m0, nr), lvt, 0.0f, Zs, nr, "cc_ovvv_gen");
    }
  }
  if (!split)
    gemm<T>(ctx, false, true, nr, vv, naux, T(1), LovT + (long)v * m0, ov, 0, Lvv, vv, 0, T(0), Zs, nr, 0, 1,
            "cc_ovvv_gen");
  permute<T, T>(ctx, Zs, {v, cnt, v, v}, {3, 0, 2, 1}, Zsp, 1.0, 0.0, "cc_ovvv_gen");
  T* Zc = Z + oo * v * m0;
  if constexpr (sizeof(T) == 4) {
    if (split) {
      // tau's transposed planes (k = (ef)) once per iteration, the slab's planes every slab
      SplitState& Zt = split_state(S);
      if (Zt.tppt_src != static_cast<const void*>(Tpp) || Zt.tppt_stamp != (long)S.iter) {
        split16t<float>(ctx, Tpp, vv, oo, oo, Zt.tppt, "cc_split");
        Zt.tppt_src = Tpp;
        Zt.tppt_stamp = S.iter;
      }
      split16<float>(ctx, Zsp, vv, nr, vv, Zt.zsp, "cc_split");
      gemm16_tn(ctx, oo, nr, 1.0f, Zt.tppt.s, Zt.zsp.s, 0.0f, Zc, oo, "cc_ovvv");
      return;
    }
  }
  gemm<T>(ctx, false, false, oo, nr, vv, T(1), Tpp, oo, 0, Zsp, vv, 0, T(0), Zc, oo, 0, 1, "cc_ovvv");
}

template void ovvv_z_slab<double>(jues_ctx*, CCState&, int, int, const double*, const double*, const double*,
                                  double*, double*, double*);

template <typename T>
void fac_tov_lov(jues_ctx* ctx, CCState& S, const T* Tov, T* W1) {
  const long ov = (long)S.o * S.v, naux = ctx->naux;
  if constexpr (sizeof(T) == 4) {
    if (S.split_on) {
      // Tov is symmetric: its ring planes serve as the TN kernel's left operand
      const Split16& t = split_operand<float>(ctx, S, SP_TOV, Tov, ov, ov);
      const Split16& l = split_operand<double>(ctx, S, SP_LOV, S.i64->Lov.as<double>(), ov, naux);
      gemm16_tn(ctx, ov, naux, 1.0f, t, l, 0.0f, W1, ov, "cc_ovvv_fac");
      return;
    }
  }
  gemm<T>(ctx, false, false, ov, naux, ov, T(1), Tov, ov, 0, ints<T>(S).Lov.template as<T>(), ov, 0, T(0), W1, ov,
          0, 1, "cc_ovvv_fac");
}

template <typename T>
void fac_lov_y2(jues_ctx* ctx, CCState& S, const T* Y2, bool y2_right, T* C) {
  const long ov = (long)S.o * S.v, naux = ctx->naux;
  if constexpr (sizeof(T) == 4) {
    if (S.split_on) {
      SplitState& Z = split_state(S);
      if (Z.y2t_src != static_cast<const void*>(Y2) || Z.y2t_stamp != (long)S.iter) {
        split16t<float>(ctx, Y2, naux, ov, ov, Z.y2t, "cc_split");
        Z.y2t_src = Y2;
        Z.y2t_stamp = S.iter;
      }
      const Split16& lt = split_operand<double>(ctx, S, SP_LOVT, S.i64->Lov.as<double>(), naux, ov, true);
      if (y2_right) gemm16_tn(ctx, ov, ov, 1.0f, lt, Z.y2t.s, 1.0f, C, ov, "cc_ovvv_fac");
      else gemm16_tn(ctx, ov, ov, 1.0f, Z.y2t.s, lt, 1.0f, C, ov, "cc_ovvv_fac");
      return;
    }
  }
  const T* L = ints<T>(S).Lov.template as<T>();
  if (y2_right) gemm<T>(ctx, false, true, ov, ov, naux, T(1), L, ov, 0, Y2, ov, 0, T(1), C, ov, 0, 1, "cc_ovvv_fac");
  else gemm<T>(ctx, false, true, ov, ov, naux, T(1), Y2, ov, 0, L, ov, 0, T(1), C, ov, 0, 1, "cc_ovvv_fac");
}

template <typename T>
void fac_y3_lvv(jues_ctx* ctx, CCState& S, const T* Y3, T* Q) {
  const long oo = (long)S.o * S.o, vv = (long)S.v * S.v, naux = ctx->naux;
  if constexpr (sizeof(T) == 4) {
    if (S.split_on) {
      SplitState& Z = split_state(S);
      split16t<float>(ctx, Y3, naux, oo, oo, Z.y3t, "cc_split");
      const Split16& lvt = split_operand<double>(ctx, S, SP_LVVT, S.i64->Lvv.as<double>(), naux, vv, true);
      gemm16_tn(ctx, oo, vv, 1.0f, Z.y3t.s, lvt, 0.0f, Q, oo, "cc_ovvv_fac");
      return;
    }
  }
  gemm<T>(ctx, false, true, oo, vv, naux, T(1), Y3, oo, 0, ints<T>(S).Lvv.template as<T>(), vv, 0, T(0), Q, oo, 0,
          1, "cc_ovvv_fac");
}

template void fac_tov_lov<double>(jues_ctx*, CCState&, const double*, double*);
template void fac_tov_lov<float>(jues_ctx*, CCState&, const float*, float*);
template void fac_lov_y2<double>(jues_ctx*, CCState&, const double*, bool, double*);
template void fac_lov_y2<float>(jues_ctx*, CCState&, const float*, bool, float*);
template void fac_y3_lvv<double>(jues_ctx*, CCState&, const double*, double*);
template void fac_y3_lvv<float>(jues_ctx*, CCState&, const float*, float*);
template void ovvv_z_slab<float>(jues_ctx*, CCState&, int, int, const float*, const float*, const float*, float*,
                                 float*, float*);

// The ladder with <ab|ef> rebuilt from the factors every iteration (DF-RCCD.jl:198-215
// contracts bvv[a,e,Q] bvv[b,f,Q] the same way): for a slab of f, X[e,a,b,f] =
// sum_P L[ea,P] L[bf,P] = (ae|bf); its rows of V+/V- (pairs e<=f / e<f with f in the slab)
// are a contiguous K-range of the ladder product, accumulated into Lp / Lm.  V+/V- never exist.
template <typename T>
static void ladder_recompute_mm(jues_ctx* ctx, CCState& S, const double* amp, const double* t1, T* Tp, T* Tm, T* Lp,
                                T* Lm) {
  const PairTables& P = S.P;
  const int v = S.v;
  const long v2 = (long)v * v, naux = ctx->naux;
  const bool minus = P.nmo > 0 && P.nmv > 0;
  CCWork<T>& W = work<T>(S);
  pack_t<T>(ctx, P, amp, S.o, v, Tp, Tm, t1);
  // slabs of about 4 GiB of (ae|bf) (JUES_LADDER_SLAB_BYTES overrides, e.g. to test many slabs)
  long slab_bytes = 4L << 30;
  if (const char* env = std::getenv("JUES_LADDER_SLAB_BYTES")) slab_bytes = std::max(1L, std::atol(env));
  const long nf = std::max(1L, std::min<long>(v, slab_bytes / (long)(v2 * v * sizeof(T))));
  const long rows_p = (long)std::min<long>(v, nf) * v;  // bound on the packed rows of one slab
  T* X = bufp<T>(W.Xs, (size_t)(v2 * v * nf));
  T* Vps = bufp<T>(W.Vps, (size_t)(rows_p * P.npv));
  T* Vms = bufp<T>(W.Vms, (size_t)std::max(1L, rows_p * P.nmv));
  const T* L = ints<T>(S).Lvv.template as<T>();
  const T* Lt = ints<T>(S).Lvvt.template as<T>();
  bool split = false;
  if constexpr (sizeof(T) == 4) split = S.split_on;
  bool first_m = true;
  for (int f0 = 0; f0 < v; f0 += (int)nf) {
    const int f1 = (int)std::min<long>(v, f0 + nf);
    const long r0p = (long)f0 * (f0 + 1) / 2, nrp = (long)f1 * (f1 + 1) / 2 - r0p;
    const long r0m = (long)f0 * (f0 - 1) / 2, nrm = (long)f1 * (f1 - 1) / 2 - r0m;
    const long ncol = (long)v * (f1 - f0);
    if constexpr (sizeof(T) == 4) {
      if (split) {
        SplitState& Z = split_state(S);
        // X[(ea),(bf)] from the planes (k = P) of L[ea,P] and of the slab's columns (bf) of L[bf,P]
        const Split16& lt = split_operand<double>(ctx, S, SP_LVVT, S.i64->Lvv.as<double>(), naux, v2, true);
        const Split16& ltt = split_operand<double>(ctx, S, SP_LVVTT, S.i64->Lvvt.as<double>(), naux, v2, true);
        gemm16_tn(ctx, v2, ncol, 1.0f, ltt, split16_cols(lt, (long)v * f0, ncol), 0.0f, X, v2, "cc_vvvv_gen");
        pack_v_ea<float>(ctx, P, X, v, f0, f1, Vps, Vms);
        split16t<float>(ctx, Tp + r0p * P.npo, nrp, P.npo, P.npo, Z.tpc, "cc_split");
        split16<float>(ctx, Vps, nrp, P.npv, nrp, Z.vps, "cc_split");
        gemm16_tn(ctx, P.npo, P.npv, 1.0f, Z.tpc.s, Z.vps.s, f0 == 0 ? 0.0f : 1.0f, Lp, P.npo, "cc_pp_ladder");
        if (minus && nrm > 0) {
          split16t<float>(ctx, Tm + r0m * P.nmo, nrm, P.nmo, P.nmo, Z.tmc, "cc_split");
          split16<float>(ctx, Vms, nrm, P.nmv, nrm, Z.vms, "cc_split");
          gemm16_tn(ctx, P.nmo, P.nmv, 1.0f, Z.tmc.s, Z.vms.s, first_m ? 0.0f : 1.0f, Lm, P.nmo, "cc_pp_ladder");
          first_m = false;
        }
        continue;
      }
    }
    gemm<T>(ctx, false, true, v2, ncol, naux, T(1), Lt, v2, 0, L + (long)v * f0, v2, 0, T(0), X, v2, 0, 1,
            "cc_vvvv_gen");
    pack_v_ea<T>(ctx, P, X, v, f0, f1, Vps, Vms);
    gemm<T>(ctx, false, false, P.npo, P.npv, nrp, T(1), Tp + r0p * P.npo, P.npo, 0, Vps, nrp, 0,
            f0 == 0 ? T(0) : T(1), Lp, P.npo, 0, 1, "cc_pp_ladder");
    if (minus && nrm > 0) {
      gemm<T>(ctx, false, false, P.nmo, P.nmv, nrm, T(1), Tm + r0m * P.nmo, P.nmo, 0, Vms, nrm, 0,
              first_m ? T(0) : T(1), Lm, P.nmo, 0, 1, "cc_pp_ladder");
      first_m = false;
    }
  }
}

template <typename T>
void ladder_mm(jues_ctx* ctx, CCState& S, const double* amp, T* Tp, T* Tm, T* Lp, T* Lm, const double* t1) {
  if (S.ladder_recompute) {
    ladder_recompute_mm<T>(ctx, S, amp, t1, Tp, Tm, Lp, Lm);
    return;
  }
  const PairTables& P = S.P;
  const bool minus = P.nmo > 0 && P.nmv > 0;
  if constexpr (sizeof(T) == 4) {
    if (S.split_on) {
      // Lp[(ij),(ab)] = sum_(ef) Tp[(ij),(ef)] Vp[(ef),(ab)]: the planes of Tp^T are the
      // k-contiguous left operand; Vp's planes come straight from the FP64 integrals
      pack_t<float>(ctx, P, amp, S.o, S.v, Tp, Tm, t1);
      const Split16& tp = split_operand<float>(ctx, S, SP_TPT, Tp, P.npv, P.npo, true);
      const Split16& vp = split_operand<double>(ctx, S, SP_VP, S.i64->Vp.as<double>(), P.npv, P.npv);
      gemm16_tn(ctx, P.npo, P.npv, 1.0f, tp, vp, 0.0f, Lp, P.npo, "cc_pp_ladder");
      if (minus) {
        const Split16& tm = split_operand<float>(ctx, S, SP_TMT, Tm, P.nmv, P.nmo, true);
        const Split16& vm = split_operand<double>(ctx, S, SP_VM, S.i64->Vm.as<double>(), P.nmv, P.nmv);
        gemm16_tn(ctx, P.nmo, P.nmv, 1.0f, tm, vm, 0.0f, Lm, P.nmo, "cc_pp_ladder");
      }
      return;
    }
    CCInts<float>& b = *S.i32;
    if (!b.Vp.p) {
      convert_buf<float>(ctx, S.i64->Vp, b.Vp);
      convert_buf<float>(ctx, S.i64->Vm, b.Vm);
    }
  }
  CCInts<T>& I = ints<T>(S);
  pack_t<T>(ctx, P, amp, S.o, S.v, Tp, Tm, t1);
  // (one launch per product: at over six rounds of workgroups each, a shared launch measured no faster)
  gemm<T>(ctx, false, false, P.npo, P.npv, P.npv, T(1), Tp, P.npo, 0, I.Vp.template as<T>(), P.npv, 0, T(0), Lp,
          P.npo, 0, 1, "cc_pp_ladder");
  if (minus)
    gemm<T>(ctx, false, false, P.nmo, P.nmv, P.nmv, T(1), Tm, P.nmo, 0, I.Vm.template as<T>(), P.nmv, 0, T(0), Lm,
            P.nmo, 0, 1, "cc_pp_ladder");
}

// With P+ = (pairs k<=l), P- = (pairs k<l) and the pair conventions of Tp / Tm:
//   W+^T[(ij),(kl)] = W'[kl,ij] + W'[lk,ij] (k<l), W'[kk,ij] (k=l) = O+ + sum_(c<=d) Tp[(ij),(cd)] KP[(kl),(cd)]
//   W-^T[(ij),(kl)] = W'[kl,ij] - W'[lk,ij]                        = O- + sum_(c<d)  Tm[(ij),(cd)] KM[(kl),(cd)]
//   Lp[(ij),(ab)] += 1/2 sum_(kl in P+) W+^T[(ij),(kl)] Tq[(kl),(ab)]   (Tq: Tp with (a,a) columns doubled)
//   Lm[(ij),(ab)] += 1/2 sum_(kl in P-) W-^T[(ij),(kl)] Tm[(kl),(ab)]
// since W'[kl,ij] = W'[lk,ji] and t[kl,ab] = t[lk,ba]: a quarter of the unpacked products' flops.
template <typename T>
void hh_mm(jues_ctx* ctx, CCState& S, const T* Wp, const T* Tp, const T* Tm, T* Lp, T* Lm) {
  const PairTables& P = S.P;
  const bool minus = P.nmo > 0 && P.nmv > 0;
  CCWork<T>& W = work<T>(S);
  T* Wpt = bufp<T>(W.Wpt, P.npo * P.npo);
  T* Wmt = bufp<T>(W.Wmt, std::max(1L, P.nmo * P.nmo));
  T* Tq = bufp<T>(W.Tq, P.npo * P.npv);
  pack_whh<T>(ctx, P, Wp, S.o, Wpt, Wmt);
  pair_diag2<T>(ctx, P, Tp, Tq);
  if constexpr (sizeof(T) == 4) {
    if (S.split_on) {
      // the TN kernel's operands are planes of the k-contiguous transposes
      const CCInts<double>& I64 = *S.i64;
      const Split16& tp = split_operand<float>(ctx, S, SP_TPT, Tp, P.npv, P.npo, true);
      const Split16& kp = split_operand<double>(ctx, S, SP_KPT, I64.KP.as<double>(), P.npv, P.npo, true);
      gemm16_tn(ctx, P.npo, P.npo, 1.0f, tp, kp, 1.0f, Wpt, P.npo, "cc_hh");
      const Split16& wp = split_operand<float>(ctx, S, SP_WPT, Wpt, P.npo, P.npo, true);
      const Split16& tq = split_operand<float>(ctx, S, SP_TQ, Tq, P.npo, P.npv);
      gemm16_tn(ctx, P.npo, P.npv, 0.5f, wp, tq, 1.0f, Lp, P.npo, "cc_hh");
      if (minus) {
        const Split16& tmt = split_operand<float>(ctx, S, SP_TMT, Tm, P.nmv, P.nmo, true);
        const Split16& km = split_operand<double>(ctx, S, SP_KMT, I64.KM.as<double>(), P.nmv, P.nmo, true);
        gemm16_tn(ctx, P.nmo, P.nmo, 1.0f, tmt, km, 1.0f, Wmt, P.nmo, "cc_hh");
        const Split16& wm = split_operand<float>(ctx, S, SP_WMT, Wmt, P.nmo, P.nmo, true);
        const Split16& tm = split_operand<float>(ctx, S, SP_TM, Tm, P.nmo, P.nmv);
        gemm16_tn(ctx, P.nmo, P.nmv, 0.5f, wm, tm, 1.0f, Lm, P.nmo, "cc_hh");
      }
      return;
    }
    CCInts<float>& b = *S.i32;
    if (!b.KP.p) {
      convert_buf<float>(ctx, S.i64->KP, b.KP);
      convert_buf<float>(ctx, S.i64->KM, b.KM);
    }
  }
  CCInts<T>& I = ints<T>(S);
  // the plus and minus products share each launch: the W' builds have fewer tiles than one
  // round of resident workgroups
  if (minus) {
    gemm_pair<T>(ctx, true, P.npo, P.npo, P.npv, Tp, P.npo, I.KP.template as<T>(), P.npo, Wpt, P.npo, P.nmo, P.nmo,
                 P.nmv, Tm, P.nmo, I.KM.template as<T>(), P.nmo, Wmt, P.nmo, T(1), T(1), "cc_hh");
    gemm_pair<T>(ctx, false, P.npo, P.npv, P.npo, Wpt, P.npo, Tq, P.npo, Lp, P.npo, P.nmo, P.nmv, P.nmo, Wmt, P.nmo,
                 Tm, P.nmo, Lm, P.nmo, T(0.5), T(1), "cc_hh");
  } else {
    gemm<T>(ctx, false, true, P.npo, P.npo, P.npv, T(1), Tp, P.npo, 0, I.KP.template as<T>(), P.npo, 0, T(1), Wpt,
            P.npo, 0, 1, "cc_hh");
    gemm<T>(ctx, false, false, P.npo, P.npv, P.npo, T(0.5), Wpt, P.npo, 0, Tq, P.npo, 0, T(1), Lp, P.npo, 0, 1,
            "cc_hh");
  }
}

template void hh_mm<double>(jues_ctx*, CCState&, const double*, const double*, const double*, double*, double*);
template void hh_mm<float>(jues_ctx*, CCState&, const float*, const float*, const float*, float*, float*);

template void ladder_mm<double>(jues_ctx*, CCState&, const double*, double*, double*, double*, double*,
                                const double*);
template void ladder_mm<float>(jues_ctx*, CCState&, const double*, float*, float*, float*, float*, const double*);

// Doubles-residual contractions that depend only on the doubles amplitudes `tau2`
// (t2 for RCCD; the CCSD residual passes its own amplitudes and adds T1 terms).
// Produces W.S (ov, symmetrised) and W.Lp / W.Lm (both ladders, packed).
template <typename T>
void rccd_contractions(jues_ctx* ctx, CCState& S, const double* t2) {
  const int o = S.o, v = S.v;
  const long ov = (long)o * v, N2 = ov * ov, oo = (long)o * o, vv = (long)v * v;
  CCInts<T>& I = ints<T>(S);
  CCWork<T>& W = work<T>(S);
  const std::vector<long> d4 = {o, o, v, v};
  T* Tpp = bufp<T>(W.Tpp, N2);
  T* Upp = bufp<T>(W.Upp, N2);
  T* Tov = bufp<T>(W.Tov, N2);
  T* Ttov = bufp<T>(W.Ttov, N2);
  T* Uov = bufp<T>(W.Uov, N2);
  T* Ua = bufp<T>(W.Ua, N2);
  T* WJ = bufp<T>(W.WJ, N2);
  T* WE = bufp<T>(W.WE, N2);
  T* Y = bufp<T>(W.Y, N2);
  T* H = bufp<T>(W.H, N2);
  T* Wp = bufp<T>(W.Wp, oo * oo);
  T* Fae = bufp<T>(W.Fae, vv);
  T* Fmi = bufp<T>(W.Fmi, oo);
  T* Tp = bufp<T>(W.Tp, S.P.npo * S.P.npv);
  T* Tm = bufp<T>(W.Tm, std::max(1L, S.P.nmo * S.P.nmv));
  T* Lp = bufp<T>(W.Lp, S.P.npo * S.P.npv);
  T* Lm = bufp<T>(W.Lm, std::max(1L, S.P.nmo * S.P.nmv));
  const T one = T(1), zero = T(0), half = T(0.5);

  // amplitudes in the compute type and the layouts the GEMMs read
  // Tpp = t, Upp = 2t - t^(ij), Tov = T[i,a,j,b], Ttov = t[i,j,b,a] at [i,a,j,b], Uov = U[i,a,j,b]
  rccd_layouts<T>(ctx, t2, o, v, Tpp, Upp, Tov, Ttov, Uov);
  permute<T, T>(ctx, Upp, d4, {2, 0, 1, 3}, Ua, 1.0, 0.0, "cc_prep");     // U[a,m,n,f]

  // Fae[a,e] = -sum U[a,(mnf)] K[e,(mnf)];  Fmi[m,i] = sum K[m,(nef)] U[i,(nef)], kept transposed
  // (Fmi^T[i,m], the product of two row-major views)
  gemm<T>(ctx, false, true, v, v, oo * v, T(-1), Ua, v, 0, I.Ka.template as<T>(), v, 0, zero, Fae, v, 0, 1, "cc_fock");
  gemm<T>(ctx, false, true, o, o, ov * v, one, Upp, o, 0, I.Kpp.template as<T>(), o, 0, zero, Fmi, o, 0, 1, "cc_fock");
  // particle-particle ladder on the packed (anti)symmetric parts, then the hole-hole
  // ladder (W' = oooo + K t^T) added into the same packed sums
  ladder_mm<T>(ctx, S, t2, Tp, Tm, Lp, Lm);
  JUES_HIP(hipMemcpyAsync(Wp, I.oooo.p, oo * oo * sizeof(T), hipMemcpyDeviceToDevice, ctx->stream));
  hh_mm<T>(ctx, S, Wp, Tp, Tm, Lp, Lm);
  // ring terms (every left operand is a symmetric (ov) x (ov) matrix)
  const T* K = I.Kov.template as<T>();
  const T* Kx = I.Ktov.template as<T>();
  JUES_HIP(hipMemcpyAsync(WJ, K, N2 * sizeof(T), hipMemcpyDeviceToDevice, ctx->stream));
  ring_k_mm<T>(ctx, S, half, Uov, SP_UOV, WJ);
  ring_mm<T>(ctx, S, -half, Kx, SP_KX, Tov, SP_TOV, one, WJ);
  permute<T, T>(ctx, I.Jov.template as<T>(), {ov, ov}, {0, 1}, WE, -1.0, 0.0, "cc_prep");
  ring_mm<T>(ctx, S, half, Kx, SP_KX, Ttov, SP_TTOV, one, WE);
  // Fock-like terms t.Fae^T - Fmi^T.t, carried by WE: with D[kc,jb] = delta_kj Fae[b,c] -
  // delta_cb Fmi[k,j], T D and (Tx D)^x each give half of them, up to a transpose that
  // S = H + H^T absorbs
  fock_dress<T>(ctx, WE, Fae, Fmi, o, v, half);
  ring_mm<T>(ctx, S, one, Tov, SP_TOV, WE, SP_WE, zero, H);
  ring_mm<T>(ctx, S, one, Uov, SP_UOV, WJ, SP_WJ, one, H);
  // exchange ring term t[m,i,b,e] WmBEj[m,a,e,j] = (Tx WE)[ib,ja]
  ring_mm<T>(ctx, S, one, Ttov, SP_TTOV, WE, SP_WE, zero, Y);
  // an FP64 iteration that anchors a defect correction keeps WJ, WE and W' (before WJ is reused)
  if constexpr (sizeof(T) == 8)
    if (S.dc && S.dc->capture) dc_capture(ctx, S, WJ, WE, W.Wpt.template as<T>(), W.Wmt.template as<T>());
  // S = H + Y^x + (H + Y^x)^T: summed by the update kernel, or here (into WJ) when its LDS slab
  // would not fit
  if (!fused_sym(o, sizeof(T))) {
    permute<T, T>(ctx, Y, {o, v, o, v}, {0, 3, 2, 1}, H, 1.0, 1.0, "cc_sym");
    JUES_HIP(hipMemcpyAsync(WJ, H, N2 * sizeof(T), hipMemcpyDeviceToDevice, ctx->stream));
    permute<T, T>(ctx, H, {ov, ov}, {1, 0}, WJ, 1.0, 1.0, "cc_sym");
  }
  W.out_H = H;
  W.out_Y = Y;
  W.out_S = WJ;
}

template void rccd_contractions<double>(jues_ctx*, CCState&, const double*);
template void rccd_contractions<float>(jues_ctx*, CCState&, const double*);

// ---------------------------------------------------------------------------
// defect correction (RCCD): the residual's change from the anchor t0, formed from d = t - t0
// ---------------------------------------------------------------------------
void dc_capture(jues_ctx* ctx, CCState& S, const double* WJ, const double* WE, const double* Wpt,
                const double* Wmt) {
  DCState& D = *S.dc;
  const long ov = (long)S.o * S.v, N2 = ov * ov;
  const PairTables& P = S.P;
  auto put = [&](DevBuf& dst, const double* src, long n) {
    if (D.inner_prec == JUES_FP64) {
      double* p = bufp<double>(dst, (size_t)n);
      JUES_HIP(hipMemcpyAsync(p, src, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    } else {
      convert<double, float>(ctx, src, bufp<float>(dst, (size_t)n), n);
    }
  };
  put(D.WJ0, WJ, N2);
  put(D.WE0, WE, N2);
  put(D.Wpt0, Wpt, P.npo * P.npo);
  if (P.nmo > 0 && P.nmv > 0) put(D.Wmt0, Wmt, P.nmo * P.nmo);
}

namespace {
// Hole-hole part of the bracket: t W'(t) - t0 W'0 = W'0 d + dW' t with dW' = <kl|cd> d[ij,cd], on
// the packed pairs of hh_mm (W'0 is the anchor's W+^T / W-^T).  Tpd / Tmd: packed d (from ladder_mm).
template <typename T>
void hh_delta_mm(jues_ctx* ctx, CCState& S, const T* Tpd, const T* Tmd, const double* t2, T* Lp, T* Lm) {
  const PairTables& P = S.P;
  const bool minus = P.nmo > 0 && P.nmv > 0;
  CCWork<T>& W = work<T>(S);
  DCState& D = *S.dc;
  T* dWpt = bufp<T>(W.Wpt, P.npo * P.npo);
  T* dWmt = bufp<T>(W.Wmt, std::max(1L, P.nmo * P.nmo));
  T* Tqd = bufp<T>(W.Tq, P.npo * P.npv);
  T* Tpt = bufp<T>(D.Tpt, P.npo * P.npv);
  T* Tmt = bufp<T>(D.Tmt, std::max(1L, P.nmo * P.nmv));
  T* Tqt = bufp<T>(D.Tqt, P.npo * P.npv);
  pack_t<T>(ctx, P, t2, S.o, S.v, Tpt, Tmt);
  pair_diag2<T>(ctx, P, Tpd, Tqd);
  pair_diag2<T>(ctx, P, Tpt, Tqt);
  if constexpr (sizeof(T) == 4) {
    CCInts<float>& b = *S.i32;
    if (!b.KP.p) {
      convert_buf<float>(ctx, S.i64->KP, b.KP);
      convert_buf<float>(ctx, S.i64->KM, b.KM);
    }
  }
  // (plain products in the stage's type also in the split stage: a few percent of the iteration)
  CCInts<T>& I = ints<T>(S);
  gemm<T>(ctx, false, true, P.npo, P.npo, P.npv, T(1), Tpd, P.npo, 0, I.KP.template as<T>(), P.npo, 0, T(0), dWpt,
          P.npo, 0, 1, "cc_hh");
  gemm<T>(ctx, false, false, P.npo, P.npv, P.npo, T(0.5), D.Wpt0.as<T>(), P.npo, 0, Tqd, P.npo, 0, T(1), Lp, P.npo,
          0, 1, "cc_hh");
  gemm<T>(ctx, false, false, P.npo, P.npv, P.npo, T(0.5), dWpt, P.npo, 0, Tqt, P.npo, 0, T(1), Lp, P.npo, 0, 1,
          "cc_hh");
  if (minus) {
    gemm<T>(ctx, false, true, P.nmo, P.nmo, P.nmv, T(1), Tmd, P.nmo, 0, I.KM.template as<T>(), P.nmo, 0, T(0), dWmt,
            P.nmo, 0, 1, "cc_hh");
    gemm<T>(ctx, false, false, P.nmo, P.nmv, P.nmo, T(0.5), D.Wmt0.as<T>(), P.nmo, 0, Tmd, P.nmo, 0, T(1), Lm,
            P.nmo, 0, 1, "cc_hh");
    gemm<T>(ctx, false, false, P.nmo, P.nmv, P.nmo, T(0.5), dWmt, P.nmo, 0, Tmt, P.nmo, 0, T(1), Lm, P.nmo, 0, 1,
            "cc_hh");
  }
}
}  // namespace

// The products of rccd_contractions with d in place of t and the anchor's intermediates (the
// split into products is the one documented with DCState); out_H / out_Y / out_S and Lp / Lm
// then hold the bracket's parts, which the update kernel adds to R0.
template <typename T>
void rccd_delta_contractions(jues_ctx* ctx, CCState& S, const double* t2) {
  const int o = S.o, v = S.v;
  const long ov = (long)o * v, N2 = ov * ov, oo = (long)o * o, vv = (long)v * v;
  CCInts<T>& I = ints<T>(S);
  CCWork<T>& W = work<T>(S);
  DCState& D = *S.dc;
  const std::vector<long> d4 = {o, o, v, v};
  double* d = bufp<double>(D.d, N2);
  JUES_HIP(hipMemcpyAsync(d, t2, N2 * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
  axpby<double>(ctx, N2, -1.0, D.t0.as<double>(), 1.0, d);
  T* Tpp = bufp<T>(W.Tpp, N2);
  T* Upp = bufp<T>(W.Upp, N2);
  T* Tov = bufp<T>(W.Tov, N2);
  T* Ttov = bufp<T>(W.Ttov, N2);
  T* Uov = bufp<T>(W.Uov, N2);
  T* Ua = bufp<T>(W.Ua, N2);
  T* WJ = bufp<T>(W.WJ, N2);
  T* WE = bufp<T>(W.WE, N2);
  T* Y = bufp<T>(W.Y, N2);
  T* H = bufp<T>(W.H, N2);
  T* Fae = bufp<T>(W.Fae, vv);
  T* Fmi = bufp<T>(W.Fmi, oo);
  T* Tp = bufp<T>(W.Tp, S.P.npo * S.P.npv);
  T* Tm = bufp<T>(W.Tm, std::max(1L, S.P.nmo * S.P.nmv));
  T* Lp = bufp<T>(W.Lp, S.P.npo * S.P.npv);
  T* Lm = bufp<T>(W.Lm, std::max(1L, S.P.nmo * S.P.nmv));
  T* Tt = bufp<T>(D.Tt, N2);
  T* Ttt = bufp<T>(D.Ttt, N2);
  T* Ut = bufp<T>(D.Ut, N2);
  const T* WJ0 = D.WJ0.as<T>();
  const T* WE0 = D.WE0.as<T>();
  const T one = T(1), zero = T(0), half = T(0.5);

  // layouts of d (as rccd_contractions makes those of t) and the ring layouts of t
  rccd_layouts<T>(ctx, d, o, v, Tpp, Upp, Tov, Ttov, Uov);
  rccd_layouts<T>(ctx, t2, o, v, nullptr, nullptr, Tt, Ttt, Ut);
  permute<T, T>(ctx, Upp, d4, {2, 0, 1, 3}, Ua, 1.0, 0.0, "cc_prep");
  // Fae, Fmi are linear in t: their values at d
  gemm<T>(ctx, false, true, v, v, oo * v, T(-1), Ua, v, 0, I.Ka.template as<T>(), v, 0, zero, Fae, v, 0, 1, "cc_fock");
  gemm<T>(ctx, false, true, o, o, ov * v, one, Upp, o, 0, I.Kpp.template as<T>(), o, 0, zero, Fmi, o, 0, 1, "cc_fock");
  // ladders: pp-ladder of d (linear), hole-hole W'0 d + dW' t
  ladder_mm<T>(ctx, S, d, Tp, Tm, Lp, Lm);
  hh_delta_mm<T>(ctx, S, Tp, Tm, t2, Lp, Lm);
  // dWJ = K U(d)/2 - Kx T(d)/2,  dWE = Kx Tx(d)/2 + D(Fae(d), Fmi(d))/2
  const T* Kx = I.Ktov.template as<T>();
  ring_mm<T>(ctx, S, -half, Kx, SP_KX, Tov, SP_TOV, zero, WJ);
  ring_k_mm<T>(ctx, S, half, Uov, SP_UOV, WJ);
  ring_mm<T>(ctx, S, half, Kx, SP_KX, Ttov, SP_TTOV, zero, WE);
  fock_dress<T>(ctx, WE, Fae, Fmi, o, v, half);
  // dH = T(d) WE0 + U(d) WJ0 + T(t) dWE + U(t) dWJ,  dY = Tx(d) WE0 + Tx(t) dWE
  ring_mm<T>(ctx, S, one, Tov, SP_TOV, WE0, SP_WE0, zero, H);
  ring_mm<T>(ctx, S, one, Uov, SP_UOV, WJ0, SP_WJ0, one, H);
  ring_mm<T>(ctx, S, one, Tt, SP_TOV_T, WE, SP_WE, one, H);
  ring_mm<T>(ctx, S, one, Ut, SP_UOV_T, WJ, SP_WJ, one, H);
  ring_mm<T>(ctx, S, one, Ttov, SP_TTOV, WE0, SP_WE0, zero, Y);
  ring_mm<T>(ctx, S, one, Ttt, SP_TTOV_T, WE, SP_WE, one, Y);
  if (!fused_sym(o, sizeof(T))) {
    permute<T, T>(ctx, Y, {o, v, o, v}, {0, 3, 2, 1}, H, 1.0, 1.0, "cc_sym");
    JUES_HIP(hipMemcpyAsync(WJ, H, N2 * sizeof(T), hipMemcpyDeviceToDevice, ctx->stream));
    permute<T, T>(ctx, H, {ov, ov}, {1, 0}, WJ, 1.0, 1.0, "cc_sym");
  }
  W.out_H = H;
  W.out_Y = Y;
  W.out_S = WJ;
}

template void rccd_delta_contractions<double>(jues_ctx*, CCState&, const double*);
template void rccd_delta_contractions<float>(jues_ctx*, CCState&, const double*);

// ---------------------------------------------------------------------------
// iteration
// ---------------------------------------------------------------------------
static void diis_begin(CCState& S, DiisOrder& ord, int& slot) {
  const int dmax = S.opts.diis ? S.opts.diis_max : 1;
  if ((int)S.active.size() >= dmax) {
    slot = S.active.front();
    S.active.erase(S.active.begin());
  } else {
    slot = 0;
    while (std::find(S.active.begin(), S.active.end(), slot) != S.active.end()) ++slot;
  }
  ord.nold = (int)S.active.size();
  for (int k = 0; k < ord.nold; ++k) ord.old_slots[k] = S.active[k];
  ord.newest = slot;
}

template <typename T>
void cc_iteration(jues_ctx* ctx, CCState& S, jues_iter_info* info) {
  const int o = S.o, v = S.v;
  const long N2 = (long)o * o * v * v, Nd = S.Nd;
  hipEvent_t e0 = S.ev[0], e1 = S.ev[1];
  JUES_HIP(hipEventRecord(e0, ctx->stream));

  DiisOrder ord{};
  int slot = 0;
  diis_begin(S, ord, slot);
  double* vals = S.vals.as<double>();
  double* errs = S.errs.as<double>();
  double* t = S.t.as<double>();
  double* tn = vals + (long)slot * Nd;
  double* en = errs + (long)slot * Nd;

  const long v2 = (long)v * v;
  const int nb1 = S.method == 1 ? 64 : 0;
  const long PB = v2 + nb1;
  double* part = static_cast<double*>(scratch(ctx, 13, sizeof(double) * (size_t)(PB * (2 + MAX_DIIS))));

  const T* R1extra = nullptr;
  DCState* dc = S.dc.get();
  const bool dc_on = dc && dc->on && S.method == 0;
  const bool dc_cap = dc && dc->capture && S.method == 0 && sizeof(T) == 8;
  if (dc_cap)
    JUES_HIP(hipMemcpyAsync(bufp<double>(dc->t0, N2), t, N2 * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
  if (S.method == 0) {
    if (dc_on) rccd_delta_contractions<T>(ctx, S, t);
    else rccd_contractions<T>(ctx, S, t);
  } else {
    rccsd_contractions<T>(ctx, S, t, t + N2);
    R1extra = nullptr;
  }
  CCWork<T>& W = work<T>(S);
  UpdateArgs<T> A{};
  A.o = o;
  A.v = v;
  A.eps = S.eps.as<double>();
  // a correction iteration adds its bracket to the anchor's numerator R0 in place of <ij|ab>
  A.Kpp = dc_on ? dc->R0.as<double>() : S.i64->Kpp.as<double>();
  A.R_out = dc_cap ? bufp<double>(dc->R0, N2) : nullptr;
  A.Lener = S.Lener.as<double>();
  if (fused_sym(o, sizeof(T))) {
    A.H = W.out_H;
    A.Y = W.out_Y;
  } else {
    A.S = W.out_S;
  }
  A.Lp = W.Lp.template as<T>();
  A.Lm = W.Lm.template as<T>();
  A.extra = S.method == 1 ? W.X2.template as<T>() : nullptr;
  A.t1n = nullptr;
  A.lambda = S.lambda;
  A.npo = S.P.npo;
  A.nmo = S.P.nmo;
  A.t_old = t;
  A.t_new = tn;
  A.e_new = en;
  A.nold = ord.nold;
  for (int k = 0; k < ord.nold; ++k) A.e_old[k] = errs + (long)ord.old_slots[k] * Nd;
  A.part = part;
  A.part_stride = PB;
  A.part_off = 0;
  if (S.method == 1) {
    // T1 first: the T2 kernel's energy term needs the new T1
    T1UpdateArgs B{};
    B.o = o;
    B.v = v;
    B.eps = S.eps.as<double>();
    B.R1 = S.R1.as<double>();
    B.Lener1 = S.fock_on ? S.fov2.as<double>() : nullptr;
    B.lambda = S.lambda;
    B.t_old = t + N2;
    B.t_new = tn + N2;
    B.e_new = en + N2;
    B.nold = ord.nold;
    for (int k = 0; k < ord.nold; ++k) B.e_old[k] = errs + (long)ord.old_slots[k] * Nd + N2;
    B.part = part;
    B.part_stride = PB;
    B.part_off = v2;
    t1_update(ctx, B, nb1);
    A.t1n = tn + N2;
  }
  t2_update<T>(ctx, A);
  double* small = S.small.as<double>();
  reduce_rows(ctx, part, PB, PB, 2 + ord.nold, small);
  S.active.push_back(slot);
  ord.m = (int)S.active.size();
  for (int k = 0; k < ord.m; ++k) ord.slots[k] = S.active[k];
  diis_solve(ctx, small, ord);
  diis_extrap(ctx, vals, Nd, ord, small, Nd, t);
  JUES_HIP(hipEventRecord(e1, ctx->stream));
  double h[2];
  d2h(ctx, h, small, sizeof(h));
  float ms = 0.f;
  JUES_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)R1extra;

  S.iter += 1;
  info->iter = S.iter;
  info->stage = sizeof(T) == 4 ? (S.split_on ? 2 : 1) : 0;
  info->energy = h[0];
  info->delta_e = S.have_prev ? h[0] - S.E_prev : h[0];
  info->rnorm = std::sqrt(std::max(0.0, h[1]));
  info->diis_size = ord.m;
  info->ms = ms;
  info->coupling = S.lambda;
  S.E_prev = h[0];
  S.have_prev = true;
  if (!std::isfinite(h[0]) || !std::isfinite(h[1]))
    throw Error(JUES_NONFINITE, "non-finite energy or residual at CC iteration " + std::to_string(S.iter));
}

template void cc_iteration<double>(jues_ctx*, CCState&, jues_iter_info*);
template void cc_iteration<float>(jues_ctx*, CCState&, jues_iter_info*);

// ---------------------------------------------------------------------------
// setup / drivers
// ---------------------------------------------------------------------------
namespace {
// t1[i,a] = f_ia / (eps_i - eps_a): the T1 guess of a Fock-driven CCSD (AutoRCCSD.jl:249)
__global__ __launch_bounds__(256) void t1_fock_guess_kernel(const double* __restrict__ fov,
                                                            const double* __restrict__ eps, int o, int v,
                                                            double* __restrict__ t1) {
  const long x = blockIdx.x * 256L + threadIdx.x;
  if (x >= (long)o * v) return;
  const int i = (int)(x % o), a = (int)(x / o);
  t1[x] = fov[x] / (eps[i] - eps[o + a]);
}
}  // namespace

void cc_setup(jues_ctx* ctx, int method, int prec, const jues_cc_opts* opts_in, const double* Cocc, int o,
              const double* Cvir, int v, int n, const double* eps, const double* T1_init, const double* T2_init,
              const double* fock) {
  JUES_REQUIRE(method == 0 || method == 1, "cc_setup: method must be 0 (RCCD) or 1 (RCCSD)");
  JUES_REQUIRE(prec == JUES_FP64 || prec == JUES_FP32_MIXED || prec == JUES_FP16X3_MIXED, "cc_setup: bad precision");
  JUES_REQUIRE(Cocc && Cvir && eps && o > 0 && v > 0, "cc_setup: bad arguments");
  JUES_REQUIRE(n == ctx->n && n > 0, "cc_setup: n does not match the AO integrals");
  jues_cc_opts opts;
  jues_cc_default_opts(&opts);
  if (opts_in) opts = *opts_in;
  JUES_REQUIRE(opts.diis_max >= 1 && opts.diis_max <= MAX_DIIS, "cc_setup: diis_max must be in 1..8");
  JUES_REQUIRE(opts.fcn >= 0 && opts.fcn < o, "cc_setup: bad fcn");
  JUES_REQUIRE(opts.guess >= 0 && opts.guess <= 3, "cc_setup: bad guess");
  JUES_REQUIRE(opts.guess != JUES_GUESS_USER || T2_init, "cc_setup: guess=USER needs T2");
  JUES_REQUIRE(opts.continuation >= 1 && std::isfinite(opts.coupling), "cc_setup: bad continuation/coupling");
  JUES_REQUIRE(opts.refine == JUES_REFINE_RESTART || opts.refine == JUES_REFINE_DEFECT, "cc_setup: bad refine");
  JUES_REQUIRE(!fock || (method == 1 && opts.coupling == 1.0 && opts.continuation == 1),
               "cc_setup: a Fock matrix needs CCSD at coupling 1 without continuation");
  sync(ctx);
  ctx->cc.reset();
  ctx->cc.reset(new CCState());
  CCState& S = *ctx->cc;
  S.method = method;
  S.prec = prec;
  S.opts = opts;
  S.lambda = opts.coupling / opts.continuation;
  const int fc = opts.fcn;
  S.o = o - fc;
  S.v = v;
  S.n = n;
  const int oo_ = S.o;
  const long N2 = (long)oo_ * oo_ * v * v, N1 = (long)oo_ * v;
  S.Nd = N2 + (method == 1 ? N1 : 0);
  for (auto& e : S.ev) JUES_HIP(hipEventCreate(&e));
  // orbital data (frozen core dropped)
  DeviceArray<double> co(ctx, Cocc + (size_t)fc * n, (size_t)n * oo_), cv(ctx, Cvir, (size_t)n * v);
  std::vector<double> e(eps + fc, eps + o + v);
  S.eps.alloc(e.size() * sizeof(double));
  h2d(ctx, S.eps.p, e.data(), e.size() * sizeof(double));
  build_pairs(ctx, S);
  S.ladder_recompute = ladder_recompute_mode(ctx, S, opts.ladder, prec);
  S.i64.reset(new CCInts<double>());
  cc_build_ints(ctx, S, co.p, cv.p);
  S.Co = std::move(co.buf);  // kept to rebuild released FP64 blocks
  S.Cv = std::move(cv.buf);
  if (method == 1) cc_build_ccsd_ints(ctx, S);
  if (fock) {
    // fock is the (o + v) x (o + v) MO Fock matrix over [Cocc | Cvir]; frozen core dropped
    const int m = o + v;
    std::vector<double> foo((size_t)oo_ * oo_), fvv((size_t)v * v), fov((size_t)oo_ * v), fov2((size_t)oo_ * v);
    for (int i = 0; i < oo_; ++i)
      for (int k = 0; k < oo_; ++k) foo[k + (size_t)oo_ * i] = k == i ? 0.0 : fock[(fc + k) + (size_t)m * (fc + i)];
    for (int a = 0; a < v; ++a)
      for (int e2 = 0; e2 < v; ++e2) fvv[a + (size_t)v * e2] = a == e2 ? 0.0 : fock[(o + a) + (size_t)m * (o + e2)];
    for (int a = 0; a < v; ++a)
      for (int i = 0; i < oo_; ++i) {
        fov[i + (size_t)oo_ * a] = fock[(fc + i) + (size_t)m * (o + a)];
        fov2[i + (size_t)oo_ * a] = 2.0 * fov[i + (size_t)oo_ * a];
      }
    S.i64->fOO.alloc(foo.size() * sizeof(double));
    S.i64->fVV.alloc(fvv.size() * sizeof(double));
    S.i64->fOV.alloc(fov.size() * sizeof(double));
    S.fov2.alloc(fov2.size() * sizeof(double));
    h2d(ctx, S.i64->fOO.p, foo.data(), foo.size() * sizeof(double));
    h2d(ctx, S.i64->fVV.p, fvv.data(), fvv.size() * sizeof(double));
    h2d(ctx, S.i64->fOV.p, fov.data(), fov.size() * sizeof(double));
    h2d(ctx, S.fov2.p, fov2.data(), fov2.size() * sizeof(double));
    S.fock_on = true;
  }
  // amplitudes
  S.t.alloc(S.Nd * sizeof(double));
  double* t = S.t.as<double>();
  fill<double>(ctx, t, S.Nd, 0.0);
  // first-order amplitudes of the equations being iterated, whose integrals carry lambda
  switch (opts.guess) {
    case JUES_GUESS_ZERO: break;
    case JUES_GUESS_MP2:
      JUES_HIP(hipMemcpyAsync(t, S.i64->Kpp.p, N2 * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
      divide_d2(ctx, t, S.eps.as<double>(), S.o, S.v, S.lambda);
      if (S.fock_on) {
        hipLaunchKernelGGL(t1_fock_guess_kernel, dim3((unsigned)((N1 + 255) / 256)), dim3(256), 0, ctx->stream,
                           S.i64->fOV.as<double>(), S.eps.as<double>(), S.o, S.v, t + N2);
        JUES_HIP(hipGetLastError());
      }
      break;
    case JUES_GUESS_OVOV:
      JUES_HIP(hipMemcpyAsync(t, S.guess_buf.p, N2 * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
      divide_d2(ctx, t, S.eps.as<double>(), S.o, S.v, S.lambda);
      break;
    case JUES_GUESS_USER:
      h2d(ctx, t, T2_init, N2 * sizeof(double));
      if (method == 1 && T1_init) h2d(ctx, t + N2, T1_init, N1 * sizeof(double));
      break;
  }
  S.guess_buf.release();
  const int dm = opts.diis ? opts.diis_max : 1;
  S.vals.alloc((size_t)dm * S.Nd * sizeof(double));
  S.errs.alloc((size_t)dm * S.Nd * sizeof(double));
  S.small.alloc(SMALL_SIZE * sizeof(double));
  fill<double>(ctx, S.small.as<double>(), SMALL_SIZE, 0.0);
  if (prec != JUES_FP64) cc_make_ints32(ctx, S);
  S.E_prev = cc_energy(ctx, S);
  S.have_prev = true;
  S.iter = 0;
  sync(ctx);
}

double cc_energy(jues_ctx* ctx, CCState& S) {
  const long N2 = (long)S.o * S.o * S.v * S.v;
  double E = dot<double, double>(ctx, S.Lener.as<double>(), S.t.as<double>(), N2);
  if (S.method == 1) E += cc_energy_t1(ctx, S);
  if (S.fock_on) E += dot<double, double>(ctx, S.fov2.as<double>(), S.t.as<double>() + N2, (long)S.o * S.v);
  return S.lambda * E;
}

// In a fast stage the FP64 integral blocks are read only by the update kernel (<ij|ab>) once
// the stage's FP32 copies and split planes exist, so when device memory is short (below 15% of
// the device free after the stage's first iteration, or JUES_THIN_FP64_INTS=1) the others are
// released; the next stage switch rebuilds them from the orbitals kept in S.
static void thin_fp64_ints(jues_ctx* ctx, CCState& S) {
  if (S.i64_thin) return;
  size_t free_b = 0, total_b = 0;
  JUES_HIP(hipMemGetInfo(&free_b, &total_b));
  const char* env = std::getenv("JUES_THIN_FP64_INTS");
  const bool force = env && std::atoi(env) != 0;
  if (!force && (double)free_b >= 0.15 * (double)total_b) return;
  sync(ctx);
  CCInts<double>& I = *S.i64;
  for (DevBuf* b : {&I.Kov, &I.Ktov, &I.Jov, &I.Ka, &I.oooo, &I.Vp, &I.Vm, &I.KP, &I.KM, &I.ooov_c, &I.ovvv_c,
                    &I.Lov, &I.LovT, &I.Lovp, &I.Lvv, &I.Lvvt})
    b->release();
  S.i64_thin = true;
}

static void ensure_fp64_ints(jues_ctx* ctx, CCState& S) {
  if (!S.i64_thin) return;
  cc_build_ints(ctx, S, S.Co.as<double>(), S.Cv.as<double>(), true);
  S.i64_thin = false;
}

void cc_run(jues_ctx* ctx, CCState& S, int stage_prec, int nsteps, bool stop_on_conv, jues_iter_info* last,
            bool* converged, jues_iter_cb cb, void* user, double fast_r_conv) {
  *converged = false;
  // the fast stage's test (step below a target, or stalled); with fast_r_conv >= 0 (the inner
  // solves of a defect correction) in every precision, with that target
  const bool fast_test = stage_prec != JUES_FP64 || fast_r_conv >= 0.0;
  const double fast_target = fast_r_conv >= 0.0 ? fast_r_conv : S.opts.fp32_r_conv;
  // Only one stage's buffers are held at a time (at o=80, v=420 the FP64 and FP32 work sets
  // together do not fit in one GPU); a later switch rebuilds what it needs.
  if (stage_prec == JUES_FP64) {
    if (S.w32 || S.sp || S.i32) {
      sync(ctx);
      S.w32.reset();
      S.sp.reset();
      S.i32.reset();
    }
  } else if (S.w64) {
    sync(ctx);
    S.w64.reset();
  }
  // a stage switch rebuilds released FP64 blocks (the new stage's copies and splits are made from them)
  if (stage_prec != S.last_stage) ensure_fp64_ints(ctx, S);
  S.last_stage = stage_prec;
  if (stage_prec != JUES_FP64) cc_make_ints32(ctx, S);
  S.split_on = stage_prec == JUES_FP16X3_MIXED;
  int stall = 0;
  double best = 1e300;
  static const char* stage_name[3] = {"fp64", "fp32", "fp16x3"};
  for (int k = 0; k < nsteps; ++k) {
    jues_iter_info info{};
    if (stage_prec != JUES_FP64) {
      cc_iteration<float>(ctx, S, &info);
      thin_fp64_ints(ctx, S);
    } else {
      cc_iteration<double>(ctx, S, &info);
    }
    if (cb) cb(user, &info);
    if (S.opts.verbose)
      std::fprintf(stderr, "  cc iter %3d %s E=% .15f dE=% .3e |dt|=%.3e diis=%d %.1f ms\n", info.iter,
                   stage_name[info.stage], info.energy, info.delta_e, info.rnorm, info.diis_size, info.ms);
    *last = info;
    if (!fast_test) {
      if (std::fabs(info.delta_e) < S.opts.e_conv && info.rnorm < S.opts.r_conv) {
        *converged = true;
        if (stop_on_conv) break;
      }
    } else {
      // fast stage: done when the step is below the target or stops shrinking
      if (info.rnorm < best * 0.9) { best = info.rnorm; stall = 0; }
      else ++stall;
      if (info.rnorm < fast_target || stall >= 4) {
        *converged = true;
        if (stop_on_conv) break;
      }
    }
  }
}

void cc_reset_diis(CCState& S) { S.active.clear(); }

// Defect correction: each FP64 iteration (DIIS reset, so a Jacobi step) anchors R0, t0 and the
// intermediates; unless it meets e_conv / r_conv, the correction equation is then iterated in
// inner_prec (with DIIS) until its step is below max(r_conv / 10, kappa ||dt||_anchor) or stalls,
// and the next FP64 iteration starts from that solution.  Each anchor's step measures the FP64
// residual itself.  kappa is about the factor by which one anchor can shrink the error, set by the
// inner products' rounding (relative to |d|): 1e-5 for FP32 MFMA, 1e-3 for split-FP16, whose
// products err by up to ~30x more; iterating the inner equation below that buys nothing.
// Returns the number of FP64 iterations run.
int cc_refine_dc(jues_ctx* ctx, CCState& S, int inner_prec, int max_anchors, int inner_maxit,
                 jues_iter_info* last, bool* converged, jues_iter_cb cb, void* user) {
  JUES_REQUIRE(S.method == 0, "defect-correction refinement is implemented for RCCD");
  JUES_REQUIRE(inner_prec == JUES_FP64 || inner_prec == JUES_FP32_MIXED || inner_prec == JUES_FP16X3_MIXED,
               "defect correction: bad inner precision");
  JUES_REQUIRE(max_anchors >= 1 && inner_maxit >= 0, "defect correction: bad iteration limits");
  *converged = false;
  S.dc.reset(new DCState());
  DCState& D = *S.dc;
  D.inner_prec = inner_prec;
  struct Clear {
    CCState& S;
    ~Clear() {
      if (S.dc) S.dc->capture = S.dc->on = false;
    }
  } clear{S};
  const double kappa = inner_prec == JUES_FP16X3_MIXED ? 1e-3 : (inner_prec == JUES_FP32_MIXED ? 1e-5 : 0.0);
  int k = 0;
  while (k < max_anchors) {
    cc_reset_diis(S);
    D.capture = true;
    bool c = false;
    cc_run(ctx, S, JUES_FP64, 1, false, last, &c, cb, user);
    D.capture = false;
    ++k;
    if (c) {
      *converged = true;
      break;
    }
    if (k == max_anchors || inner_maxit == 0) continue;
    cc_reset_diis(S);
    D.on = true;
    jues_iter_info inner{};
    bool ci = false;
    cc_run(ctx, S, inner_prec, inner_maxit, true, &inner, &ci, cb, user,
           std::max(0.1 * S.opts.r_conv, kappa * last->rnorm));
    D.on = false;
  }
  // the anchor's numerator, amplitudes and intermediates (about 20 GB at o=60, v=340) are not
  // kept once the FP64 iterations end
  sync(ctx);
  S.dc.reset();
  return k;
}

}  // namespace jues

using namespace jues;

namespace {

jues_status cc_solve(jues_ctx* ctx, int method, int prec, const jues_cc_opts* opts, const double* Cocc, int o,
                     const double* Cvir, int v, int n, const double* eps, double* E, double* T1_inout,
                     double* T2_inout, jues_iter_cb cb, void* user, const double* fock = nullptr) {
  JUES_API_BEGIN(ctx)
  JUES_REQUIRE(E, "null energy pointer");
  cc_setup(ctx, method, prec, opts, Cocc, o, Cvir, v, n, eps, T1_inout, T2_inout, fock);
  CCState& S = *ctx->cc;
  jues_iter_info last{};
  bool conv = false;
  const int nl = S.opts.continuation;
  for (int k = 1; k <= nl; ++k) {
    if (k > 1) {
      // next coupling, warm-started from the previous solution
      S.lambda = S.opts.coupling * k / nl;
      cc_reset_diis(S);
      S.E_prev = cc_energy(ctx, S);
    }
    if (prec == JUES_FP64) {
      cc_run(ctx, S, JUES_FP64, S.opts.maxit, true, &last, &conv, cb, user);
    } else {
      cc_run(ctx, S, prec, S.opts.maxit, true, &last, &conv, cb, user);
      if (k == nl) {
        cc_reset_diis(S);
        if (S.opts.refine == JUES_REFINE_DEFECT && method == 0)
          cc_refine_dc(ctx, S, prec, S.opts.refine_maxit, S.opts.maxit, &last, &conv, cb, user);
        else
          cc_run(ctx, S, JUES_FP64, S.opts.refine_maxit, true, &last, &conv, cb, user);
      }
    }
    if (!conv && k < nl) {
      ctx->err = "CC continuation did not converge at coupling " + std::to_string(S.lambda);
      *E = cc_energy(ctx, S);
      return JUES_NOT_CONVERGED;
    }
  }
  *E = cc_energy(ctx, S);
  const long N2 = (long)S.o * S.o * S.v * S.v;
  if (T2_inout) d2h(ctx, T2_inout, S.t.p, N2 * sizeof(double));
  if (T1_inout && method == 1) d2h(ctx, T1_inout, S.t.as<double>() + N2, (size_t)S.o * S.v * sizeof(double));
  if (!std::isfinite(*E)) throw Error(JUES_NONFINITE, "non-finite CC energy");
  if (!conv) {
    ctx->err = "CC equations did not converge in the iteration limit";
    return JUES_NOT_CONVERGED;
  }
  JUES_API_END(ctx)
}

}  // namespace

extern "C" {

jues_status jues_rccd(jues_ctx* ctx, int prec, const jues_cc_opts* opts, const double* Cocc, int o, const double* Cvir,
                      int v, int n, const double* eps, double* E, double* T2_inout, jues_iter_cb cb, void* user) {
  return cc_solve(ctx, 0, prec, opts, Cocc, o, Cvir, v, n, eps, E, nullptr, T2_inout, cb, user);
}

jues_status jues_rccsd(jues_ctx* ctx, int prec, const jues_cc_opts* opts, const double* Cocc, int o,
                       const double* Cvir, int v, int n, const double* eps, double* E, double* T1_inout,
                       double* T2_inout, jues_iter_cb cb, void* user) {
  return cc_solve(ctx, 1, prec, opts, Cocc, o, Cvir, v, n, eps, E, T1_inout, T2_inout, cb, user);
}

jues_status jues_rccsd_fock(jues_ctx* ctx, int prec, const jues_cc_opts* opts, const double* Cocc, int o,
                            const double* Cvir, int v, int n, const double* fock, double* E, double* T1_inout,
                            double* T2_inout, jues_iter_cb cb, void* user) {
  if (!ctx) return JUES_BAD_ARG;
  if (!fock || o <= 0 || v <= 0) {
    ctx->err = "jues_rccsd_fock: bad arguments";
    return JUES_BAD_ARG;
  }
  // the resolvents take the Fock diagonal (AutoRCCSD.jl:220-222,243-246)
  const int m = o + v;
  std::vector<double> diag((size_t)m);
  for (int p = 0; p < m; ++p) diag[p] = fock[p + (size_t)m * p];
  return cc_solve(ctx, 1, prec, opts, Cocc, o, Cvir, v, n, diag.data(), E, T1_inout, T2_inout, cb, user, fock);
}

jues_status jues_cc_setup(jues_ctx* ctx, int method, int prec, const jues_cc_opts* opts, const double* Cocc, int o,
                          const double* Cvir, int v, int n, const double* eps, const double* T1_init,
                          const double* T2_init) {
  JUES_API_BEGIN(ctx)
  cc_setup(ctx, method, prec, opts, Cocc, o, Cvir, v, n, eps, T1_init, T2_init, nullptr);
  JUES_API_END(ctx)
}

jues_status jues_cc_iterate(jues_ctx* ctx, int stage_prec, int nsteps, int stop_on_conv, jues_iter_info* info,
                            int* converged) {
  JUES_API_BEGIN(ctx)
  JUES_REQUIRE(ctx->cc, "jues_cc_iterate: call jues_cc_setup first");
  JUES_REQUIRE(stage_prec == JUES_FP64 || stage_prec == JUES_FP32_MIXED || stage_prec == JUES_FP16X3_MIXED,
               "jues_cc_iterate: bad precision");
  jues_iter_info last{};
  bool conv = false;
  cc_run(ctx, *ctx->cc, stage_prec, nsteps, stop_on_conv != 0, &last, &conv, nullptr, nullptr);
  if (info) *info = last;
  if (converged) *converged = conv ? 1 : 0;
  JUES_API_END(ctx)
}

jues_status jues_cc_refine(jues_ctx* ctx, int inner_prec, int max_anchors, int inner_maxit, jues_iter_info* info,
                           int* converged, int* n_fp64) {
  JUES_API_BEGIN(ctx)
  JUES_REQUIRE(ctx->cc, "jues_cc_refine: call jues_cc_setup first");
  jues_iter_info last{};
  bool conv = false;
  const int n = cc_refine_dc(ctx, *ctx->cc, inner_prec, max_anchors, inner_maxit, &last, &conv, nullptr, nullptr);
  if (info) *info = last;
  if (converged) *converged = conv ? 1 : 0;
  if (n_fp64) *n_fp64 = n;
  JUES_API_END(ctx)
}

jues_status jues_cc_reset_diis(jues_ctx* ctx) {
  JUES_API_BEGIN(ctx)
  JUES_REQUIRE(ctx->cc, "no CC state");
  cc_reset_diis(*ctx->cc);
  JUES_API_END(ctx)
}

jues_status jues_cc_get(jues_ctx* ctx, double* E, double* T1_out, double* T2_out) {
  JUES_API_BEGIN(ctx)
  JUES_REQUIRE(ctx->cc, "no CC state");
  CCState& S = *ctx->cc;
  if (E) *E = cc_energy(ctx, S);
  const long N2 = (long)S.o * S.o * S.v * S.v;
  if (T2_out) d2h(ctx, T2_out, S.t.p, N2 * sizeof(double));
  if (T1_out && S.method == 1) d2h(ctx, T1_out, S.t.as<double>() + N2, (size_t)S.o * S.v * sizeof(double));
  JUES_API_END(ctx)
}

}  // extern "C"
