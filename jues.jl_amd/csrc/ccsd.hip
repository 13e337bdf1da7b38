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
//
// With AO factors the o v^3 block (ia|bc) = sum_P Lov[i,a,P] Lvv[b,c,P] is never stored
// (at o=80, v=420 it would be 47 GB in FP64, and each TTGT contraction with it a second
// permuted copy): every term that reads it contracts the factors instead, through the
// small intermediates B1[m,P,e] = Y2[m,e,P] = sum_f t1[m,f] Lvv[e,f,P] and
// Y3[m,j,P] = sum_f Lov[m,f,P] t1[j,f], and Z is built from slabs of (mf|ae) generated
// for a few m at a time (2 o v^3 naux flops per iteration).
#include "cc.h"

#include <algorithm>
#include <cstring>

namespace jues {
namespace {

// tau = t2 + t1 t1 into Tpp, 2 tau' - tau'^(ij) (tau' = t2 + t1 t1 / 2) into Upp, both [i,j,a,b],
// and tt^x[i,a,j,b] = t1[i,b] t1[j,a] (ov layout) into ttx, in the compute type
template <typename T>
__global__ __launch_bounds__(256) void ccsd_amps_kernel(const double* __restrict__ t2, const double* __restrict__ t1,
                                                        int o, int v, T* __restrict__ Tpp, T* __restrict__ Upp,
                                                        T* __restrict__ ttx) {
  const long oo = (long)o * o, ov = (long)o * v, n = oo * v * v;
  for (long x = blockIdx.x * 256L + threadIdx.x; x < n; x += (long)gridDim.x * 256L) {
    const long ij = x % oo, ab = x / oo;
    const int i = (int)(ij % o), j = (int)(ij / o), a = (int)(ab % v), b = (int)(ab / v);
    const double tt = t1[i + (long)o * a] * t1[j + (long)o * b];
    const double tt_ji = t1[j + (long)o * a] * t1[i + (long)o * b];
    const double t2ji = t2[j + (long)o * i + oo * ab];
    Tpp[x] = (T)(t2[x] + tt);
    Upp[x] = (T)(2.0 * (t2[x] + 0.5 * tt) - (t2ji + 0.5 * tt_ji));
    ttx[i + (long)o * a + ov * (j + (long)o * b)] = (T)tt_ji;
  }
}

// per (a,b): sum_ij Lener[i,j,a,b] t1[i,a] t1[j,b]
__global__ __launch_bounds__(256) void energy_t1_kernel(const double* __restrict__ Lener, const double* __restrict__ t1,
                                                        int o, int v, double* __restrict__ part) {
  __shared__ double sh[4];
  const long oo = (long)o * o;
  const int ab = blockIdx.x, a = ab % v, b = ab / v;
  const double* L = Lener + oo * ab;
  double s = 0.0;
  for (long ij = threadIdx.x; ij < oo; ij += 256) {
    const int i = (int)(ij % o), j = (int)(ij / o);
    s += L[ij] * t1[i + (long)o * a] * t1[j + (long)o * b];
  }
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) part[ab] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

__global__ __launch_bounds__(256) void sum_parts_kernel(const double* __restrict__ part, long count,
                                                        double* __restrict__ out) {
  __shared__ double sh[256];
  double s = 0.0;
  for (long i = threadIdx.x; i < count; i += 256) s += part[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = sh[0];
}

int g256(long n) { return (int)std::max(1L, std::min((n + 255) / 256, 131072L)); }

}  // namespace

void cc_build_ccsd_ints(jues_ctx* ctx, CCState& S) {
  const long N1 = (long)S.o * S.v;
  S.R1.alloc(N1 * sizeof(double));
  (void)ctx;
}

double cc_energy_t1(jues_ctx* ctx, CCState& S) {
  // sum_ijab Lener[ijab] t1[ia] t1[jb], fixed-order partials per (a,b)
  const int o = S.o, v = S.v;
  const long N2 = (long)o * o * v * v, vv = (long)v * v;
  const double* t1 = S.t.as<double>() + N2;
  double* part = static_cast<double*>(scratch(ctx, 7, sizeof(double) * (size_t)(vv + 1)));
  hipLaunchKernelGGL(energy_t1_kernel, dim3((unsigned)vv), dim3(256), 0, ctx->stream, S.Lener.as<double>(), t1, o, v,
                     part);
  hipLaunchKernelGGL(sum_parts_kernel, dim3(1), dim3(256), 0, ctx->stream, part, vv, part + vv);
  JUES_HIP(hipGetLastError());
  double E = 0.0;
  d2h(ctx, &E, part + vv, sizeof(double));
  return E;
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
  const bool fac = S.ovvv_factored;
  T* Tpp = bufp_t<T>(W.Tpp, N2);   // tau
  T* Upp = bufp_t<T>(W.Upp, N2);   // 2 tau' - tau'^(ij)
  T* Tov = bufp_t<T>(W.Tov, N2);   // t2 [i,a,j,b]
  T* Ttov = bufp_t<T>(W.Ttov, N2); // t2[i,j,b,a] at [i,a,j,b]
  T* Uov = bufp_t<T>(W.Uov, N2);   // 2 t2 - t2^x (ov)
  T* Ua = bufp_t<T>(W.Ua, N2);
  // WJ and WE are built after Upp and Ua are last read (Fmi, Fae), Y after Tpp (Z): with
  // factored (ia|bc) they share those buffers
  T* WJ = fac ? Upp : bufp_t<T>(W.WJ, N2);
  T* WE = fac ? Ua : bufp_t<T>(W.WE, N2);
  T* Y = fac ? Tpp : bufp_t<T>(W.Y, N2);
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
  const long naux = ctx->naux;

  // ---- amplitudes in the compute type -------------------------------------------
  T* ttx = U2;  // tt^x [i,a,j,b] (overwritten below once used)
  convert<double, T>(ctx, t1_64, t1, ov);
  {
    KTimer k(ctx, "cc_prep", 0.0, (double)N2 * (16.0 + 3.0 * sizeof(T)));
    hipLaunchKernelGGL((ccsd_amps_kernel<T>), dim3(g256(N2)), dim3(256), 0, ctx->stream, t2, t1_64, o, v, Tpp, Upp,
                       ttx);
    JUES_HIP(hipGetLastError());
  }
  permute<T, T>(ctx, Upp, d4, {2, 0, 1, 3}, Ua, 1.0, 0.0, "cc_prep");
  permute<double, T>(ctx, t2, d4, {0, 2, 1, 3}, Tov, 1.0, 0.0, "cc_prep");
  permute<double, T>(ctx, t2, d4, {0, 3, 1, 2}, Ttov, 1.0, 0.0, "cc_prep");
  permute<T, T>(ctx, Tov, {ov, ov}, {0, 1}, Uov, 2.0, 0.0, "cc_prep");
  permute<T, T>(ctx, Ttov, {ov, ov}, {0, 1}, Uov, -1.0, 1.0, "cc_prep");

  // factors of (ia|bc) and their contractions with t1
  const T* Lov = fac ? I.Lov.template as<T>() : nullptr;    // [m,e,P]
  const T* LovT = fac ? I.LovT.template as<T>() : nullptr;  // [e,m,P]
  const T* Lovp = fac ? I.Lovp.template as<T>() : nullptr;  // [m,P,e]
  const T* Lvv = fac ? I.Lvv.template as<T>() : nullptr;    // [a,e,P]
  T *B1 = nullptr, *Y2 = nullptr, *Y3 = nullptr;
  if (fac) {
    B1 = bufp_t<T>(W.B1, ov * naux);
    Y2 = bufp_t<T>(W.Y2, ov * naux);
    Y3 = bufp_t<T>(W.Y3, oo * naux);
    // B1[m,P,e] = Y2[m,e,P] = sum_f t1[m,f] Lvv[e,f,P] (batched over P; B1's rows interleave
    // the batches through ldc = o naux), Y3[m,j,P] = sum_f Lov[m,f,P] t1[j,f]
    gemm<T>(ctx, false, true, o, v, v, one, t1, o, 0, Lvv, v, vv, zero, B1, o * naux, o, (int)naux, "cc_ovvv_fac");
    gemm<T>(ctx, false, true, o, v, v, one, t1, o, 0, Lvv, v, vv, zero, Y2, o, ov, (int)naux, "cc_ovvv_fac");
    gemm<T>(ctx, false, true, o, o, v, one, Lov, o, ov, t1, o, 0, zero, Y3, o, oo, (int)naux, "cc_ovvv_fac");
  }

  // ---- Fae, Fmi, Fme --------------------------------------------------------------
  gemm<T>(ctx, false, true, v, v, oo * v, T(-1), Ua, v, 0, I.Ka.template as<T>(), v, 0, zero, Fae, v, 0, 1, "cc_fock");
  gemm<T>(ctx, false, true, o, o, ov * v, one, I.Kpp.template as<T>(), o, 0, Upp, o, 0, zero, Fmi, o, 0, 1, "cc_fock");
  const T* ovvv = I.ovvv_c.template as<T>();
  const T* ooov = I.ooov_c.template as<T>();
  const T* Kov = I.Kov.template as<T>();
  const T* Jov = I.Jov.template as<T>();
  // Fae[a,e] += sum_mf t1[mf] (2(mf|ae) - X),  X = (ma|ef) (mRCCSD.jl:140) or (me|af) (RCCSD.jl:189)
  if (fac) {
    // 2 sum_mf t1[mf] (mf|ae) = 2 sum_P g[P] Lvv[(ae),P] with g[P] = sum_(mf) t1[(mf)] Lov[(mf),P];
    // (ma|ef) term: sum_(mP) Lovp[(mP),a] B1[(mP),e]; (me|af) term: sum_(mP) B1[(mP),a] Lovp[(mP),e]
    T* g = bufp_t<T>(W.gP, naux);
    gemm<T>(ctx, true, false, naux, 1, ov, one, Lov, ov, 0, t1, ov, 0, zero, g, naux, 0, 1, "cc_ovvv_fac");
    gemm<T>(ctx, false, false, vv, 1, naux, T(2), Lvv, vv, 0, g, naux, 0, one, Fae, vv, 0, 1, "cc_ovvv_fac");
    gemm<T>(ctx, true, false, v, v, o * naux, T(-1), mrccsd ? Lovp : B1, o * naux, 0, mrccsd ? B1 : Lovp, o * naux, 0,
            one, Fae, v, 0, 1, "cc_ovvv_fac");
  } else {
    contract<T>(ctx, 2.0, TN(t1, "mf"), TN(ovvv, "mfae"), 1.0, TN(Fae, "ae"), "cc_t1");
    contract<T>(ctx, -1.0, TN(t1, "mf"), TN(ovvv, mrccsd ? "maef" : "meaf"), 1.0, TN(Fae, "ae"), "cc_t1");
  }
  // Fmi[m,i] += sum_ne t1[ne] (2(mi|ne) - (ni|me))
  contract<T>(ctx, 2.0, TN(t1, "ne"), TN(ooov, "mine"), 1.0, TN(Fmi, "mi"), "cc_t1");
  contract<T>(ctx, -1.0, TN(t1, "ne"), TN(ooov, "nime"), 1.0, TN(Fmi, "mi"), "cc_t1");
  // Fme[m,e] = sum_nf t1[nf] (2(me|nf) - (ne|mf)), (ne|mf) = (mf|ne) = Kx[m,e,n,f]
  const T* Kx = I.Ktov.template as<T>();
  contract<T>(ctx, 2.0, TN(t1, "nf"), TN(Kov, "menf"), 0.0, TN(Fme, "me"), "cc_t1");
  contract<T>(ctx, -1.0, TN(t1, "nf"), TN(Kx, "menf"), 1.0, TN(Fme, "me"), "cc_t1");

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
  // (Ua is free from the Fae product until WE is built: the permuted copy of (mi|ae) goes
  // there; t2[m,n,a,e] = t2[n,m,e,a] is read as Tov[n,e,m,a], no copy)
  contract<T>(ctx, 1.0, TN(t1, "ie"), TN(Fae, "ae"), 0.0, TN(n1, "ia"), "cc_t1");
  contract<T>(ctx, -1.0, TN(t1, "ma"), TN(Fmi, "mi"), 1.0, TN(n1, "ia"), "cc_t1");
  contract<T>(ctx, 1.0, TN(Fme, "me"), TN(Uov, "iame"), 1.0, TN(n1, "ia"), "cc_t1");
  contract<T>(ctx, 2.0, TN(t1, "me"), TN(Kov, "iame"), 1.0, TN(n1, "ia"), "cc_t1");
  contract<T>(ctx, -1.0, TN(t1, "me"), TN(Jov, "maie"), 1.0, TN(n1, "ia"), "cc_t1", Ua, N2);
  contract<T>(ctx, -2.0, TN(Tov, "nema"), TN(ooov, "mine"), 1.0, TN(n1, "ia"), "cc_t1");
  contract<T>(ctx, 1.0, TN(Tov, "nema"), TN(ooov, "nime"), 1.0, TN(n1, "ia"), "cc_t1");
  if (fac) {
    // 2 sum_emf T[ie,mf] (mf|ae): W1[(ie),P] = T[(ie),(mf)] Lov[(mf),P], n1 += 2 W1[i,(eP)] Lvv[a,(eP)];
    // - sum_emf T[ie,mf] (me|af): W2[i,f,P] = sum_(em) T[i,(em),f] LovT[(em),P] (batched over f and
    // written straight into [i,f,P]), n1 -= W2[i,(fP)] Lvv[a,(fP)]
    T* W1 = bufp_t<T>(W.W1, ov * naux);
    T* W2 = bufp_t<T>(W.W2, ov * naux);
    fac_tov_lov<T>(ctx, S, Tov, W1);
    gemm<T>(ctx, false, true, o, v, v * naux, T(2), W1, o, 0, Lvv, v, 0, one, n1, o, 0, 1, "cc_ovvv_fac");
    gemm<T>(ctx, false, false, o, naux, ov, one, Tov, o, oo * v, LovT, ov, 0, zero, W2, ov, o, v, "cc_ovvv_fac");
    gemm<T>(ctx, false, true, o, v, v * naux, T(-1), W2, o, 0, Lvv, v, 0, one, n1, o, 0, 1, "cc_ovvv_fac");
  } else {
    contract<T>(ctx, 2.0, TN(Tov, "iemf"), TN(ovvv, "mfae"), 1.0, TN(n1, "ia"), "cc_t1");
    contract<T>(ctx, -1.0, TN(Tov, "iemf"), TN(ovvv, "meaf"), 1.0, TN(n1, "ia"), "cc_t1");
  }
  if (S.fock_on) axpby<T>(ctx, ov, 1.0, I.fOV.template as<T>(), 1.0, n1);  // f_ia
  convert<T, double>(ctx, n1, S.R1.as<double>(), ov);

  // ---- Fb = Fae - t1.Fme/2, Fj = Fmi + t1.Fme/2 ------------------------------------
  contract<T>(ctx, -0.5, TN(t1, "mb"), TN(Fme, "me"), 1.0, TN(Fae, "be"), "cc_t1");
  contract<T>(ctx, 0.5, TN(t1, "je"), TN(Fme, "me"), 1.0, TN(Fmi, "mj"), "cc_t1");

  // ---- particle-particle ladder on tau = t2 + t1 t1 (packed on the fly) ----------------
  ladder_mm<T>(ctx, S, t2, Tp, Tm, Lp, Lm, t1_64);
  // ---- hole-hole ladder: W' = oooo + T1 terms + <mn|ef> tau[ij,ef], added into Lp / Lm ----
  JUES_HIP(hipMemcpyAsync(Wp, I.oooo.p, oo * oo * sizeof(T), hipMemcpyDeviceToDevice, ctx->stream));
  contract<T>(ctx, 1.0, TN(t1, "je"), TN(ooov, "mine"), 1.0, TN(Wp, "mnij"), "cc_t1");
  contract<T>(ctx, 1.0, TN(t1, "ie"), TN(ooov, "njme"), 1.0, TN(Wp, "mnij"), "cc_t1");
  hh_mm<T>(ctx, S, Wp, Tp, Tm, Lp, Lm);
  // Wabef's T1 part: X2 = Lz + Lz^P,  Lz[ijab] = -sum_m t1[mb] Z[ij,ma],  Z = tau . (mf|ae)
  if (fac) {
    // Z[i,j,a,m] over slabs of m: Zs[(f,m'),(a,e)] = sum_P LovT[(f,m),P] Lvv[(ae),P] = (mf|ae),
    // Zsp[(e,f),(a,m')] its permutation, Z[(ij),(a,m)] = Tpp[(ij),(ef)] Zsp; then
    // Lz[(ija),b] = -sum_m Z[(ija),m] t1[m,b]
    const long per_m = (long)v * vv;
    const long ms = std::max(1L, std::min<long>(o, (2L << 30) / (per_m * (long)sizeof(T))));
    T* Zs = bufp_t<T>(W.Zs, (size_t)(ms * per_m));
    T* Zsp = bufp_t<T>(W.Zsp, (size_t)(ms * per_m));
    for (int m0 = 0; m0 < o; m0 += (int)ms) {
      const int cnt = (int)std::min<long>(ms, o - m0);
      ovvv_z_slab<T>(ctx, S, m0, cnt, LovT, Lvv, Tpp, Zs, Zsp, Z);
    }
    gemm<T>(ctx, false, false, oo * v, v, o, T(-1), Z, oo * v, 0, t1, o, 0, zero, Q, oo * v, 0, 1, "cc_t1");
  } else {
    contract<T>(ctx, 1.0, TN(Tpp, "ijef"), TN(ovvv, "mfae"), 0.0, TN(Z, "ijma"), "cc_ovvv");
    contract<T>(ctx, -1.0, TN(t1, "mb"), TN(Z, "ijma"), 0.0, TN(Q, "ijab"), "cc_t1");
  }
  permute<T, T>(ctx, Q, d4, {0, 1, 2, 3}, X2, 1.0, 0.0, "cc_sym");
  permute<T, T>(ctx, Q, d4, {1, 0, 3, 2}, X2, 1.0, 1.0, "cc_sym");

  // ---- ring terms -----------------------------------------------------------------------
  // U2 = U - 2 tt^x (U2 holds tt^x on entry); Q = Tx + 2 tt^x
  permute<T, T>(ctx, Ttov, {ov, ov}, {0, 1}, Q, 1.0, 0.0, "cc_prep");
  permute<T, T>(ctx, U2, {ov, ov}, {0, 1}, Q, 2.0, 1.0, "cc_prep");
  permute<T, T>(ctx, Uov, {ov, ov}, {0, 1}, U2, 1.0, -2.0, "cc_prep");
  JUES_HIP(hipMemcpyAsync(WJ, Kov, N2 * sizeof(T), hipMemcpyDeviceToDevice, ctx->stream));
  // every left operand of the ring products is a symmetric (ov) x (ov) matrix
  ring_k_mm<T>(ctx, S, half, U2, SP_U2, WJ);
  ring_mm<T>(ctx, S, -half, Kx, SP_KX, Tov, SP_TOV, one, WJ);
  // WJ[me,jb] += sum_f t1[jf] (me|bf)
  if (fac) fac_lov_y2<T>(ctx, S, Y2, true, WJ);
  else contract<T>(ctx, 1.0, TN(t1, "jf"), TN(ovvv, "mebf"), 1.0, TN(WJ, "mejb"), "cc_t1");
  // (H is free until the ring products: the [b,j,m,e] product of the ooov terms goes there)
  contract<T>(ctx, -1.0, TN(t1, "nb"), TN(ooov, "njme"), 1.0, TN(WJ, "mejb"), "cc_t1", H, N2);
  permute<T, T>(ctx, Jov, {ov, ov}, {0, 1}, WE, -1.0, 0.0, "cc_prep");
  ring_mm<T>(ctx, S, half, Kx, SP_KX, Q, SP_Q, one, WE);
  // WE[me,jb] -= sum_f t1[jf] (mf|be): with factors, Q[(mj),(be)] = sum_P Y3[(mj),P] Lvv[(be),P]
  // (Q is free once the ring product above has read it) permuted into WE
  if (fac) {
    fac_y3_lvv<T>(ctx, S, Y3, Q);
    permute<T, T>(ctx, Q, d4, {0, 3, 1, 2}, WE, -1.0, 1.0, "cc_ovvv_fac");
  } else {
    contract<T>(ctx, -1.0, TN(t1, "jf"), TN(ovvv, "mfbe"), 1.0, TN(WE, "mejb"), "cc_t1");
  }
  contract<T>(ctx, 1.0, TN(t1, "nb"), TN(ooov, "mjne"), 1.0, TN(WE, "mejb"), "cc_t1", H, N2);
  // Fock-like terms with Fb / Fj, carried by WE as in RCCD: WE += D/2 with
  // D[kc,jb] = delta_kj Fb[b,c] - delta_cb Fj[k,j]
  T* FmiT = bufp_t<T>(W.FmiT, oo);
  permute<T, T>(ctx, Fmi, {o, o}, {1, 0}, FmiT, 1.0, 0.0, "cc_fock");
  fock_dress<T>(ctx, WE, Fae, FmiT, o, v, half);
  ring_mm<T>(ctx, S, one, Tov, SP_TOV, WE, SP_WE, zero, H);
  ring_mm<T>(ctx, S, one, Uov, SP_UOV, WJ, SP_WJ, one, H);
  // exchange ring term t[m,i,b,e] WmBEj[m,a,e,j] = (Tx WE)[ib,ja]
  ring_mm<T>(ctx, S, one, Ttov, SP_TTOV, WE, SP_WE, zero, Y);
  // T1^2 terms: -t1[ie] t1[ma] (me|jb) and -t1[ie] t1[mb] (mj|ae); (me|jb) = K[j,b,m,e] is read
  // in place, and the products permuted into H go through U2, free once WJ is built
  contract<T>(ctx, 1.0, TN(t1, "ie"), TN(Kov, "jbme"), 0.0, TN(Q, "imjb"), "cc_t1");
  contract<T>(ctx, -1.0, TN(t1, "ma"), TN(Q, "imjb"), 1.0, TN(H, "iajb"), "cc_t1", U2, N2);
  contract<T>(ctx, 1.0, TN(t1, "ie"), TN(Jov, "maje"), 0.0, TN(Q, "imaj"), "cc_t1");
  contract<T>(ctx, -1.0, TN(t1, "mb"), TN(Q, "imaj"), 1.0, TN(H, "iajb"), "cc_t1", U2, N2);
  // T1-linear terms: t1[ie] (jb|ae) - t1[ma] (mi|jb); with factors the first is Y2[(ia),P] Lov[(jb),P]
  if (fac) fac_lov_y2<T>(ctx, S, Y2, false, H);
  else contract<T>(ctx, 1.0, TN(t1, "ie"), TN(ovvv, "jbae"), 1.0, TN(H, "iajb"), "cc_t1");
  contract<T>(ctx, -1.0, TN(t1, "ma"), TN(ooov, "mijb"), 1.0, TN(H, "iajb"), "cc_t1", U2, N2);
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

template void rccsd_contractions<double>(jues_ctx*, CCState&, const double*, const double*);
template void rccsd_contractions<float>(jues_ctx*, CCState&, const double*, const double*);

}  // namespace jues
