// RCCSD residual on the device (CoupledCluster/mRCCSD.jl:118-241, and the
// RCCSD.jl:187-289 variant of Fae).
//
// The doubles residual keeps the RCCD structure of cc.hip with
//   tau  = t2 + t1 t1    in the hole-hole and particle-particle ladders,
//   tau' = t2 + t1 t1/2  in Fae / Fmi,
//   WJ   = K + K (U - 2 tt^x)/2 - Kx T/2 + T1 terms,
//   WE   = -J + Kx (Tx + 2 tt^x)/2 + T1 terms + D/2, tt^x[ia,jb] = t1[ib] t1[ja],
//   D[kc,jb] = delta_kj Fae[b,c] - delta_cb Fmi[k,j] (the Fock-like terms, as in RCCD),
// Wabef's T1 terms become  -sum_m t1[mb] Z[ij,ma] + (ia<->jb)  with
// Z[ij,ma] = sum_ef tau[ij,ef] (mf|ae); the remaining T1 terms are small
// contractions done by the TTGT engine.  Integral labels below: occupied
// letters i j m n, virtual letters a b e f; ooov_c[i,j,k,a] = (ij|ka),
// ovvv_c[i,a,b,c] = (ia|bc).
#include "cc.h"

#include <algorithm>
#include <cstring>

namespace jues {
namespace {

template <typename T>
__global__ __launch_bounds__(256) void tau_kernel(const double* __restrict__ t2, const double* __restrict__ t1, int o,
                                                  int v, double* __restrict__ tau, double* __restrict__ taup,
                                                  T* __restrict__ ttx) {
  const long oo = (long)o * o, ov = (long)o * v, n = oo * v * v;
  for (long x = blockIdx.x * 256L + threadIdx.x; x < n; x += (long)gridDim.x * 256L) {
    const long ij = x % oo, ab = x / oo;
    const int i = (int)(ij % o), j = (int)(ij / o), a = (int)(ab % v), b = (int)(ab / v);
    const double tt = t1[i + (long)o * a] * t1[j + (long)o * b];
    tau[x] = t2[x] + tt;
    taup[x] = t2[x] + 0.5 * tt;
    // ttx at ov index [i,a,j,b] = t1[i,b] t1[j,a]
    ttx[i + (long)o * a + ov * (j + (long)o * b)] = (T)(t1[i + (long)o * b] * t1[j + (long)o * a]);
  }
}

int g256(long n) { return (int)std::max(1L, std::min((n + 255) / 256, 131072L)); }

}  // namespace

void cc_build_ccsd_ints(jues_ctx* ctx, CCState& S) {
  const long N1 = (long)S.o * S.v;
  S.R1.alloc(N1 * sizeof(double));
  (void)ctx;
}

double cc_energy_t1(jues_ctx* ctx, CCState& S) {
  // sum_ijab Lener[ijab] t1[ia] t1[jb]
  const int o = S.o, v = S.v;
  const long N2 = (long)o * o * v * v;
  const double* t1 = S.t.as<double>() + N2;
  double* tau = bufp_t<double>(S.tau64, N2);
  double* taup = bufp_t<double>(S.taup64, N2);
  double* ttx = bufp_t<double>(S.ttx64, N2);
  DevBuf zero(N2 * sizeof(double));
  fill<double>(ctx, zero.as<double>(), N2, 0.0);
  hipLaunchKernelGGL((tau_kernel<double>), dim3(g256(N2)), dim3(256), 0, ctx->stream, zero.as<double>(), t1, o, v, tau,
                     taup, ttx);
  JUES_HIP(hipGetLastError());
  return dot<double, double>(ctx, S.Lener.as<double>(), tau, N2);
}

template <typename T>
void rccsd_contractions(jues_ctx* ctx, CCState& S, const double* t2, const double* t1_64) {
  const int o = S.o, v = S.v;
  const long ov = (long)o * v, N2 = ov * ov, oo = (long)o * o, vv = (long)v * v;
  CCInts<T>& I = ints<T>(S);
  CCWork<T>& W = work<T>(S);
  const std::vector<long> d4 = {o, o, v, v};
  auto TN = [&](const void* p, const char* idx) {
    Tens<T> x;
    x.p = const_cast<T*>(static_cast<const T*>(p));
    x.idx = idx;
    for (const char* c = idx; *c; ++c) x.dims.push_back(std::strchr("ijmn", *c) ? (long)o : (long)v);
    return x;
  };
  T* Tpp = bufp_t<T>(W.Tpp, N2);   // tau
  T* Upp = bufp_t<T>(W.Upp, N2);   // 2 tau' - tau'^(ij)
  T* Tov = bufp_t<T>(W.Tov, N2);   // t2 [i,a,j,b]
  T* Ttov = bufp_t<T>(W.Ttov, N2); // t2[i,j,b,a] at [i,a,j,b]
  T* Uov = bufp_t<T>(W.Uov, N2);   // 2 t2 - t2^x (ov)
  T* Ua = bufp_t<T>(W.Ua, N2);
  T* WJ = bufp_t<T>(W.WJ, N2);
  T* WE = bufp_t<T>(W.WE, N2);
  T* Y = bufp_t<T>(W.Y, N2);
  T* H = bufp_t<T>(W.H, N2);
  T* Wp = bufp_t<T>(W.Wp, oo * oo);
  T* Fae = bufp_t<T>(W.Fae, vv);
  T* Fmi = bufp_t<T>(W.Fmi, oo);
  T* Fme = bufp_t<T>(W.Fme, ov);
  T* t1 = bufp_t<T>(W.t1, ov);
  T* n1 = bufp_t<T>(W.n1, ov);
  T* X2 = bufp_t<T>(W.X2, N2);
  T* U2 = bufp_t<T>(W.U2, N2);
  T* Q = bufp_t<T>(W.Q, std::max(N2, oo * ov));
  T* Z = bufp_t<T>(W.Z, oo * ov);
  T* Tp = bufp_t<T>(W.Tp, S.P.npo * S.P.npv);
  T* Tm = bufp_t<T>(W.Tm, std::max(1L, S.P.nmo * S.P.nmv));
  T* Lp = bufp_t<T>(W.Lp, S.P.npo * S.P.npv);
  T* Lm = bufp_t<T>(W.Lm, std::max(1L, S.P.nmo * S.P.nmv));
  const T one = T(1), zero = T(0), half = T(0.5);
  const bool mrccsd = S.opts.ccsd_eqs == JUES_CCSD_MRCCSD;

  // ---- amplitudes in the compute type -------------------------------------------
  double* tau = bufp_t<double>(S.tau64, N2);
  double* taup = bufp_t<double>(S.taup64, N2);
  T* ttx = U2;  // tt^x [i,a,j,b] (overwritten below once used)
  hipLaunchKernelGGL((tau_kernel<T>), dim3(g256(N2)), dim3(256), 0, ctx->stream, t2, t1_64, o, v, tau, taup, ttx);
  JUES_HIP(hipGetLastError());
  convert<double, T>(ctx, t1_64, t1, ov);
  permute<double, T>(ctx, tau, d4, {0, 1, 2, 3}, Tpp, 1.0, 0.0, "cc_prep");
  permute<double, T>(ctx, taup, d4, {0, 1, 2, 3}, Upp, 2.0, 0.0, "cc_prep");
  permute<double, T>(ctx, taup, d4, {1, 0, 2, 3}, Upp, -1.0, 1.0, "cc_prep");
  permute<T, T>(ctx, Upp, d4, {2, 0, 1, 3}, Ua, 1.0, 0.0, "cc_prep");
  permute<double, T>(ctx, t2, d4, {0, 2, 1, 3}, Tov, 1.0, 0.0, "cc_prep");
  permute<double, T>(ctx, t2, d4, {0, 3, 1, 2}, Ttov, 1.0, 0.0, "cc_prep");
  permute<T, T>(ctx, Tov, {ov, ov}, {0, 1}, Uov, 2.0, 0.0, "cc_prep");
  permute<T, T>(ctx, Ttov, {ov, ov}, {0, 1}, Uov, -1.0, 1.0, "cc_prep");

  // ---- Fae, Fmi, Fme --------------------------------------------------------------
  gemm<T>(ctx, false, true, v, v, oo * v, T(-1), Ua, v, 0, I.Ka.template as<T>(), v, 0, zero, Fae, v, 0, 1, "cc_fock");
  gemm<T>(ctx, false, true, o, o, ov * v, one, I.Kpp.template as<T>(), o, 0, Upp, o, 0, zero, Fmi, o, 0, 1, "cc_fock");
  const T* ovvv = I.ovvv_c.template as<T>();
  const T* ooov = I.ooov_c.template as<T>();
  const T* Kov = I.Kov.template as<T>();
  const T* Jov = I.Jov.template as<T>();
  // Fae[a,e] += sum_mf t1[mf] (2(mf|ae) - X),  X = (ma|ef) (mRCCSD.jl:140) or (me|af) (RCCSD.jl:189)
  contract<T>(ctx, 2.0, TN(t1, "mf"), TN(ovvv, "mfae"), 1.0, TN(Fae, "ae"), "cc_t1");
  contract<T>(ctx, -1.0, TN(t1, "mf"), TN(ovvv, mrccsd ? "maef" : "meaf"), 1.0, TN(Fae, "ae"), "cc_t1");
  // Fmi[m,i] += sum_ne t1[ne] (2(mi|ne) - (ni|me))
  contract<T>(ctx, 2.0, TN(t1, "ne"), TN(ooov, "mine"), 1.0, TN(Fmi, "mi"), "cc_t1");
  contract<T>(ctx, -1.0, TN(t1, "ne"), TN(ooov, "nime"), 1.0, TN(Fmi, "mi"), "cc_t1");
  // Fme[m,e] = sum_nf t1[nf] (2(me|nf) - (ne|mf))
  contract<T>(ctx, 2.0, TN(t1, "nf"), TN(Kov, "menf"), 0.0, TN(Fme, "me"), "cc_t1");
  contract<T>(ctx, -1.0, TN(t1, "nf"), TN(Kov, "nemf"), 1.0, TN(Fme, "me"), "cc_t1");

  // ---- off-diagonal Fock terms of a non-canonical reference (AutoRCCSD.jl:76-81,127-132):
  // Fae += f_ae - t1[ma] f_me / 2, Fmi += f_mi + t1[ie] f_me / 2, Fme += f_me
  if (S.fock_on) {
    const T* fOO = I.fOO.template as<T>();
    const T* fVV = I.fVV.template as<T>();
    const T* fOV = I.fOV.template as<T>();
    axpby<T>(ctx, vv, 1.0, fVV, 1.0, Fae);
    contract<T>(ctx, -0.5, TN(t1, "ma"), TN(fOV, "me"), 1.0, TN(Fae, "ae"), "cc_fock_t1");
    axpby<T>(ctx, oo, 1.0, fOO, 1.0, Fmi);
    contract<T>(ctx, 0.5, TN(t1, "ie"), TN(fOV, "me"), 1.0, TN(Fmi, "mi"), "cc_fock_t1");
    axpby<T>(ctx, ov, 1.0, fOV, 1.0, Fme);
  }

  // ---- T1 residual (mRCCSD.jl:202-212) ----------------------------------------------
  contract<T>(ctx, 1.0, TN(t1, "ie"), TN(Fae, "ae"), 0.0, TN(n1, "ia"), "cc_t1");
  contract<T>(ctx, -1.0, TN(t1, "ma"), TN(Fmi, "mi"), 1.0, TN(n1, "ia"), "cc_t1");
  contract<T>(ctx, 1.0, TN(Fme, "me"), TN(Uov, "iame"), 1.0, TN(n1, "ia"), "cc_t1");
  contract<T>(ctx, 2.0, TN(t1, "me"), TN(Kov, "iame"), 1.0, TN(n1, "ia"), "cc_t1");
  contract<T>(ctx, -1.0, TN(t1, "me"), TN(Jov, "maie"), 1.0, TN(n1, "ia"), "cc_t1");
  contract<T>(ctx, -2.0, TN(Tov, "mane"), TN(ooov, "mine"), 1.0, TN(n1, "ia"), "cc_t1");
  contract<T>(ctx, 1.0, TN(Tov, "mane"), TN(ooov, "nime"), 1.0, TN(n1, "ia"), "cc_t1");
  contract<T>(ctx, 2.0, TN(Tov, "iemf"), TN(ovvv, "mfae"), 1.0, TN(n1, "ia"), "cc_t1");
  contract<T>(ctx, -1.0, TN(Tov, "iemf"), TN(ovvv, "meaf"), 1.0, TN(n1, "ia"), "cc_t1");
  if (S.fock_on) axpby<T>(ctx, ov, 1.0, I.fOV.template as<T>(), 1.0, n1);  // f_ia
  convert<T, double>(ctx, n1, S.R1.as<double>(), ov);

  // ---- Fb = Fae - t1.Fme/2, Fj = Fmi + t1.Fme/2 ------------------------------------
  contract<T>(ctx, -0.5, TN(t1, "mb"), TN(Fme, "me"), 1.0, TN(Fae, "be"), "cc_t1");
  contract<T>(ctx, 0.5, TN(t1, "je"), TN(Fme, "me"), 1.0, TN(Fmi, "mj"), "cc_t1");

  // ---- particle-particle ladder on tau ------------------------------------------------
  ladder_mm<T>(ctx, S, tau, Tp, Tm, Lp, Lm);
  // ---- hole-hole ladder: W' = oooo + T1 terms + <mn|ef> tau[ij,ef], added into Lp / Lm ----
  JUES_HIP(hipMemcpyAsync(Wp, I.oooo.p, oo * oo * sizeof(T), hipMemcpyDeviceToDevice, ctx->stream));
  contract<T>(ctx, 1.0, TN(t1, "je"), TN(ooov, "mine"), 1.0, TN(Wp, "mnij"), "cc_t1");
  contract<T>(ctx, 1.0, TN(t1, "ie"), TN(ooov, "njme"), 1.0, TN(Wp, "mnij"), "cc_t1");
  hh_mm<T>(ctx, S, Wp, Tp, Tm, Lp, Lm);
  // Wabef's T1 part: X2 = Lz + Lz^P,  Lz[ijab] = -sum_m t1[mb] Z[ij,ma],  Z = tau . (mf|ae)
  contract<T>(ctx, 1.0, TN(Tpp, "ijef"), TN(ovvv, "mfae"), 0.0, TN(Z, "ijma"), "cc_ovvv");
  contract<T>(ctx, -1.0, TN(t1, "mb"), TN(Z, "ijma"), 0.0, TN(Q, "ijab"), "cc_t1");
  permute<T, T>(ctx, Q, d4, {0, 1, 2, 3}, X2, 1.0, 0.0, "cc_sym");
  permute<T, T>(ctx, Q, d4, {1, 0, 3, 2}, X2, 1.0, 1.0, "cc_sym");

  // ---- ring terms -----------------------------------------------------------------------
  const T* Kx = I.Ktov.template as<T>();
  // U2 = U - 2 tt^x (U2 holds tt^x on entry); Q = Tx + 2 tt^x
  permute<T, T>(ctx, Ttov, {ov, ov}, {0, 1}, Q, 1.0, 0.0, "cc_prep");
  permute<T, T>(ctx, U2, {ov, ov}, {0, 1}, Q, 2.0, 1.0, "cc_prep");
  permute<T, T>(ctx, Uov, {ov, ov}, {0, 1}, U2, 1.0, -2.0, "cc_prep");
  JUES_HIP(hipMemcpyAsync(WJ, Kov, N2 * sizeof(T), hipMemcpyDeviceToDevice, ctx->stream));
  // every left operand of the ring products is a symmetric (ov) x (ov) matrix
  ring_mm<T>(ctx, S, half, Kov, SP_K, U2, SP_U2, one, WJ);
  ring_mm<T>(ctx, S, -half, Kx, SP_KX, Tov, SP_TOV, one, WJ);
  contract<T>(ctx, 1.0, TN(t1, "jf"), TN(ovvv, "mebf"), 1.0, TN(WJ, "mejb"), "cc_t1");
  contract<T>(ctx, -1.0, TN(t1, "nb"), TN(ooov, "njme"), 1.0, TN(WJ, "mejb"), "cc_t1");
  permute<T, T>(ctx, Jov, {ov, ov}, {0, 1}, WE, -1.0, 0.0, "cc_prep");
  ring_mm<T>(ctx, S, half, Kx, SP_KX, Q, SP_Q, one, WE);
  contract<T>(ctx, -1.0, TN(t1, "jf"), TN(ovvv, "mfbe"), 1.0, TN(WE, "mejb"), "cc_t1");
  contract<T>(ctx, 1.0, TN(t1, "nb"), TN(ooov, "mjne"), 1.0, TN(WE, "mejb"), "cc_t1");
  // Fock-like terms with Fb / Fj, carried by WE as in RCCD: WE += D/2 with
  // D[kc,jb] = delta_kj Fb[b,c] - delta_cb Fj[k,j]
  T* FmiT = bufp_t<T>(W.FmiT, oo);
  permute<T, T>(ctx, Fmi, {o, o}, {1, 0}, FmiT, 1.0, 0.0, "cc_fock");
  fock_dress<T>(ctx, WE, Fae, FmiT, o, v, half);
  ring_mm<T>(ctx, S, one, Tov, SP_TOV, WE, SP_WE, zero, H);
  ring_mm<T>(ctx, S, one, Uov, SP_UOV, WJ, SP_WJ, one, H);
  // exchange ring term t[m,i,b,e] WmBEj[m,a,e,j] = (Tx WE)[ib,ja]
  ring_mm<T>(ctx, S, one, Ttov, SP_TTOV, WE, SP_WE, zero, Y);
  // T1^2 terms: -t1[ie] t1[ma] (me|jb) and -t1[ie] t1[mb] (mj|ae)
  contract<T>(ctx, 1.0, TN(t1, "ie"), TN(Kov, "mejb"), 0.0, TN(Q, "imjb"), "cc_t1");
  contract<T>(ctx, -1.0, TN(t1, "ma"), TN(Q, "imjb"), 1.0, TN(H, "iajb"), "cc_t1");
  contract<T>(ctx, 1.0, TN(t1, "ie"), TN(Jov, "maje"), 0.0, TN(Q, "imaj"), "cc_t1");
  contract<T>(ctx, -1.0, TN(t1, "mb"), TN(Q, "imaj"), 1.0, TN(H, "iajb"), "cc_t1");
  // T1-linear terms: t1[ie] (jb|ae) - t1[ma] (mi|jb)
  contract<T>(ctx, 1.0, TN(t1, "ie"), TN(ovvv, "jbae"), 1.0, TN(H, "iajb"), "cc_t1");
  contract<T>(ctx, -1.0, TN(t1, "ma"), TN(ooov, "mijb"), 1.0, TN(H, "iajb"), "cc_t1");
  // S = H + Y^x + (H + Y^x)^T: summed by the update kernel, or here (into WJ) when its LDS slab
  // would not fit
  if (!fused_sym(o, sizeof(T))) {
    permute<T, T>(ctx, Y, {o, v, o, v}, {0, 3, 2, 1}, H, 1.0, 1.0, "cc_sym");
    JUES_HIP(hipMemcpyAsync(WJ, H, N2 * sizeof(T), hipMemcpyDeviceToDevice, ctx->stream));
    permute<T, T>(ctx, H, {ov, ov}, {1, 0}, WJ, 1.0, 1.0, "cc_sym");
  }
}

template void rccsd_contractions<double>(jues_ctx*, CCState&, const double*, const double*);
template void rccsd_contractions<float>(jues_ctx*, CCState&, const double*, const double*);

}  // namespace jues
