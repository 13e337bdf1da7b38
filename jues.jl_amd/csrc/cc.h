// Coupled-cluster state shared by cc.hip (RCCD, drivers) and ccsd.hip (RCCSD terms).
#pragma once

#include "cc_kernels.h"
#include "jues_internal.h"

namespace jues {

// MO integral blocks in one precision.  ov layouts are [i,a,j,b]; chemist blocks
// keep the tei_transform order.
template <typename T>
struct CCInts {
  DevBuf Kov;     // (ia|jb) at [i,a,j,b]
  DevBuf Ktov;    // (ib|ja) at [i,a,j,b]
  DevBuf Jov;     // (ij|ab) at [i,a,j,b]
  DevBuf Kpp;     // <ij|ab> = (ia|jb) at [i,j,a,b]
  DevBuf Ka;      // <mn|ef> at [e,m,n,f]
  DevBuf oooo;    // <mn|ij> at [m,n,i,j]
  DevBuf Vp, Vm;  // packed pp-ladder integrals
  DevBuf KP, KM;  // packed hole-hole ladder integrals (pair sums / differences of <kl|cd>)
  DevBuf ooov_c;  // (ij|ka) at [i,j,k,a]      (RCCSD)
  DevBuf ovvv_c;  // (ia|bc) at [i,a,b,c]      (RCCSD with dense AO integrals)
  // RCCSD with AO factors: (ia|bc) = sum_P Lov[i,a,P] Lvv[b,c,P] is never stored; the
  // o v^3 terms contract these factors (LovT[a,i,P] = Lov[i,a,P], Lovp[i,P,a] = Lov[i,a,P])
  DevBuf Lov, LovT, Lovp;
  DevBuf Lvv;     // MO factors L[ae,P] (v^2 x naux) when the ladder is recomputed or (ia|bc) factored
  DevBuf Lvvt;    // the same with a and e exchanged: L[ea,P]
  // off-diagonal Fock blocks of a non-canonical reference (AutoRCCSD.jl:215-232):
  // f_mi at [m,i], f_ae at [a,e] (diagonals zero), f_ia at [i,a]
  DevBuf fOO, fVV, fOV;
};

template <typename T>
struct CCWork {
  DevBuf Tpp, Upp, Tov, Ttov, Uov, Ua, Tp, Tm, WJ, WE, Y, H, Wp, Fae, Fmi, Lp, Lm;
  DevBuf Wpt, Wmt, Tq;  // hole-hole ladder on packed pairs
  DevBuf Xs, Vps, Vms;  // recomputed ladder: (ae|bf) slab and its packed V+/V- rows
  // RCCSD
  DevBuf t1, Fme, X2, Z, Q, n1, U2, Tau, FmiT;
  // RCCSD o v^3 terms through the factors: B1[m,P,e] / Y2[m,e,P] = sum_f t1[m,f] Lvv[e,f,P],
  // Y3[m,j,P] = sum_f Lov[m,f,P] t1[j,f], W1[i,e,P], W2[i,f,P], g[P]; Zs / Zsp: slabs of (mf|ae)
  DevBuf B1, Y2, Y3, W1, W2, gP, Zs, Zsp;
  DevBuf UL;  // U Lov (ov x naux): the factored K U ring product
  // where the last residual evaluation left H, Y (ov layout) and, when the update kernel does
  // not symmetrise them itself, S = H + Y^x + (H + Y^x)^T (the buffers may be shared)
  T* out_H = nullptr;
  T* out_Y = nullptr;
  T* out_S = nullptr;
};

// Split-FP16 operand planes of the fast stage.  Slots below SPLIT_CONST hold
// per-iteration operands (re-split on their first use in an iteration); the
// others hold integral matrices, split once.
// The _T slots hold the full amplitudes' layouts beside those of the correction d in a
// defect-correction iteration; SP_WE0 / SP_WJ0 hold the anchor's intermediates.
// SP_LOV / SP_LOVT hold the ov x naux factor Lov and its transpose (the factored K U product);
// SP_LOVAT the transpose of LovT[(a,i),P] (RCCSD's (mf|ae) slabs).
enum SplitSlot {
  SP_TOV, SP_UOV, SP_TTOV, SP_WE, SP_WJ, SP_U2, SP_Q, SP_TPT, SP_TMT, SP_TQ, SP_TM, SP_WPT, SP_WMT,
  SP_TOV_T, SP_UOV_T, SP_TTOV_T,
  SPLIT_CONST = 16, SP_K = SPLIT_CONST, SP_KX, SP_VP, SP_VM, SP_KPT, SP_KMT, SP_LVVT, SP_LVVTT, SP_WE0, SP_WJ0,
  SP_LOV, SP_LOVT, SP_LOVAT, SPLIT_SLOTS
};
struct SplitState {
  struct Entry {
    Split16Buf buf;
    const void* src = nullptr;
    long stamp = -2;
  };
  Entry e[SPLIT_SLOTS];
  Split16Buf vps, vms, tpc, tmc;  // per-slab planes of the recomputed ladder (never cached)
  Split16Buf ult;                 // (U Lov)^T of the factored K U product (never cached)
  Split16Buf tppt, zsp;           // RCCSD's tau^T (once per iteration) and one (ef) x (a m) slab
  const void* tppt_src = nullptr;
  long tppt_stamp = -2;
  Split16Buf y2t, y3t;            // RCCSD's Y2^T (once per iteration, two products) and Y3^T
  const void* y2t_src = nullptr;
  long y2t_stamp = -2;
};

// Defect-correction refinement of RCCD.  An FP64 iteration at the anchor t0 keeps its numerator
// R0 = <ij|ab> + S(t0) + L(t0) and its intermediates WJ0, WE0 and the packed hole-hole W'0; the
// iterations in between solve D t = lambda (R0 + [R(t) - R(t0)]) with the bracket formed from
// d = t - t0 (every product has d, or a difference linear in d, as one factor):
//   dWJ = K U(d)/2 - Kx T(d)/2,  dWE = Kx Tx(d)/2 + D(Fae(d), Fmi(d))/2
//   dH = T(d) WE0 + T(t) dWE + U(d) WJ0 + U(t) dWJ,   dY = Tx(d) WE0 + Tx(t) dWE
//   dL = pp-ladder(d) + W'0 d + dW' t,  dW'[kl,ij] = sum_cd <kl|cd> d[ij,cd]
// so that rounding in the inner stage's type scales with |d| rather than |t|.
struct DCState {
  DevBuf t0, R0, d;             // FP64: anchor amplitudes, its numerator, t - t0
  DevBuf WJ0, WE0, Wpt0, Wmt0;  // anchor intermediates in the inner stage's type
  DevBuf Tt, Ttt, Ut;           // ov layouts of the full amplitudes t
  DevBuf Tpt, Tmt, Tqt;         // packed pairs of t (Tqt: Tpt with the (a,a) columns doubled)
  int inner_prec = -1;          // the type the anchor was captured in
  bool capture = false;         // the next FP64 iteration records an anchor
  bool on = false;              // iterations solve the correction equation
};

struct CCState {
  int method = 0;  // 0 RCCD, 1 RCCSD
  int prec = 0;
  jues_cc_opts opts{};
  int o = 0, v = 0, n = 0;
  long Nd = 0;     // amplitude vector length: o^2 v^2 (+ o v for RCCSD, T1 stored after T2)
  PairTables P;
  DevBuf pair_buf[4];
  DevBuf eps;      // FP64 orbital energies (frozen core removed)
  DevBuf Lener;    // FP64 2<ij|ab> - <ji|ab>
  DevBuf Lener1;   // unused (T1 energy enters through the T2 kernel)
  DevBuf guess_buf;
  DevBuf t;        // FP64 master amplitudes [t2 | t1]
  DevBuf vals, errs, small;  // DIIS storage
  std::vector<int> active;   // DIIS slots, oldest first
  DevBuf R1;       // FP64 T1 residual numerator (RCCSD)
  bool ovvv_factored = false;  // RCCSD: the (ia|bc) terms through the factors (no o v^3 block)
  bool k_factored = false;     // K = (ia|jb) = Lov Lov^T with Lov kept: K U as two naux-wide products
  DevBuf Co, Cv;               // device orbitals (frozen core removed), to rebuild FP64 blocks
  bool i64_thin = false;       // FP64 blocks other than <ij|ab> released during a fast stage
  int last_stage = -1;         // precision of the last cc_run stage
  std::unique_ptr<CCInts<double>> i64;
  std::unique_ptr<CCInts<float>> i32;
  std::unique_ptr<CCWork<double>> w64;
  std::unique_ptr<CCWork<float>> w32;
  std::unique_ptr<SplitState> sp;
  std::unique_ptr<DCState> dc;
  bool split_on = false;  // the FP32 stage runs its ring / pp-ladder products on split-FP16 MFMA
  bool ladder_recompute = false;  // <ab|ef> rebuilt from factors every iteration instead of stored
  hipEvent_t ev[2] = {nullptr, nullptr};
  double lambda = 1.0;  // coupling of the equations being iterated
  double E_prev = 0;
  bool have_prev = false;
  bool fock_on = false;  // Fock-driven CCSD: the fOO/fVV/fOV terms and 2 f_ia t_ia in the energy
  DevBuf fov2;           // FP64 2 f_ia (the T1 energy weight)
  int iter = 0;
  ~CCState() {
    for (auto& e : ev)
      if (e) (void)hipEventDestroy(e);
  }
};

template <typename T>
CCWork<T>& work(CCState& S);
template <typename T>
CCInts<T>& ints(CCState& S);
template <typename T>
T* bufp_t(DevBuf& b, size_t n) {
  if (b.bytes < n * sizeof(T)) b.alloc(n * sizeof(T));
  return b.as<T>();
}

// Ring product C = alpha A B + beta C over (ov) x (ov) matrices whose left
// operand is symmetric; slotA/slotB name the operands for the split-FP16 stage.
template <typename T>
void ring_mm(jues_ctx* ctx, CCState& S, T alpha, const T* A, int slotA, const T* B, int slotB, T beta, T* C);
// C += alpha K B with K = (ia|jb) at [i,a,j,b] and B a symmetric (ov) x (ov) matrix: the full
// ring product, or with factored integrals (K = Lov Lov^T) C += alpha Lov (B Lov)^T, two
// products of 2 (ov)^2 naux flops instead of one of 2 (ov)^3
template <typename T>
void ring_k_mm(jues_ctx* ctx, CCState& S, T alpha, const T* B, int slotB, T* C);
// Hole-hole ladder sum_kl W'[kl,ij] t[kl,ab], W' = Wp + <kl|cd> t[ij,cd], on packed pairs,
// added into the pp-ladder's Lp / Lm (run after ladder_mm, whose Tp / Tm it reads).
// Wp: [m,n,i,j], oooo plus any T1 terms.
template <typename T>
void hh_mm(jues_ctx* ctx, CCState& S, const T* Wp, const T* Tp, const T* Tm, T* Lp, T* Lm);
// Packed pp-ladder Lp/Lm from FP64 amplitudes `amp` (ijab), or of tau = amp + t1 t1 when t1
// (o x v) is given; Tp/Tm are workspaces.
template <typename T>
void ladder_mm(jues_ctx* ctx, CCState& S, const double* amp, T* Tp, T* Tm, T* Lp, T* Lm, const double* t1 = nullptr);

template <typename T>
void rccd_contractions(jues_ctx* ctx, CCState& S, const double* t2);
template <typename T>
void rccsd_contractions(jues_ctx* ctx, CCState& S, const double* t2, const double* t1);
template <typename T>
void rccd_delta_contractions(jues_ctx* ctx, CCState& S, const double* t2);
// One slab of occupied m (m0 .. m0 + cnt - 1) of RCCSD's Z[(ij),(a,m)] = sum_ef tau[ij,ef] (mf|ae) with
// (ia|bc) factored: Zs[(f,m'),(a,e)] = sum_P LovT[(f,m),P] Lvv[(ae),P], permuted to Zsp[(e,f),(a,m')],
// times Tpp = tau [(ij),(ef)] into Z's columns (a,m) of the slab.  Split-FP16 in the split stage.
template <typename T>
void ovvv_z_slab(jues_ctx* ctx, CCState& S, int m0, int cnt, const T* LovT, const T* Lvv, const T* Tpp, T* Zs,
                 T* Zsp, T* Z);
// RCCSD's (ia|bc) factor products (split-FP16 in the split stage): W1 = Tov Lov (ov x naux);
// C += Lov Y2^T (y2_right) or C += Y2 Lov^T (Y2: ov x naux); Q = Y3 Lvv^T (Y3: o^2 x naux)
template <typename T>
void fac_tov_lov(jues_ctx* ctx, CCState& S, const T* Tov, T* W1);
template <typename T>
void fac_lov_y2(jues_ctx* ctx, CCState& S, const T* Y2, bool y2_right, T* C);
template <typename T>
void fac_y3_lvv(jues_ctx* ctx, CCState& S, const T* Y3, T* Q);
// the anchor's intermediates, in the inner stage's type (called by the FP64 contractions)
void dc_capture(jues_ctx* ctx, CCState& S, const double* WJ, const double* WE, const double* Wpt,
                const double* Wmt);
template <typename T>
void cc_iteration(jues_ctx* ctx, CCState& S, jues_iter_info* info);
void cc_run(jues_ctx* ctx, CCState& S, int stage_prec, int nsteps, bool stop_on_conv, jues_iter_info* last,
            bool* converged, jues_iter_cb cb, void* user, double fast_r_conv = -1.0);
void cc_reset_diis(CCState& S);
// defect-correction refinement (RCCD): at most max_anchors FP64 iterations with inner solves between
int cc_refine_dc(jues_ctx* ctx, CCState& S, int inner_prec, int max_anchors, int inner_maxit,
                 jues_iter_info* last, bool* converged, jues_iter_cb cb, void* user);
void cc_build_ccsd_ints(jues_ctx* ctx, CCState& S);
double cc_energy(jues_ctx* ctx, CCState& S);
double cc_energy_t1(jues_ctx* ctx, CCState& S);

}  // namespace jues
