// Hand-written adjoint of the AO->MO transform and the RMP2 energy.
//
// E = sum K_iajb (2 K_iajb - K_ibja) / D_ijab with K = (ia|jb) (RMP2.jl:33-37 on
// the block of Transformation.jl:39-93).  With every entry of C and g
// independent, G_iajb = dE/dK_iajb = 2 (2 K_iajb - K_ibja) / D_ijab and
//   dE/dg[mu,nu,lam,sig] = sum G_iajb Co[mu,i] Cv[nu,a] Co[lam,j] Cv[sig,b],
//   dE/dC = [dC1 + dC3 | dC2 + dC4] for the four positions C enters, each g with the other
//   three positions transformed, contracted with G:
//     dC1[mu,i]  = sum (mu a|j sig) G[i,a,j,b] Cv[sig,b]     dC2[nu,a]  = sum (i nu|j b) G[i,a,j,b]
//     dC3[lam,j] = sum (i nu|lam b) G[i,a,j,b] Cv[nu,a]      dC4[sig,b] = sum (i a|j sig) G[i,a,j,b]
//   (i nu|lam b) and (i nu|j b) are intermediates of the forward (ia|jb) transform, kept for
//   this; (mu a|j sig) is one more half-transform and (i a|j sig) its occupied quarter
//   (mp2_grad_dev).  Nothing is stored between calls.
// No symmetry of g is assumed.
//
// Factor route ((mu nu|lam sig) = sum_P B[mu,nu,P] B[lam,sig,P]): K = L L^T with
// L[i,a,P] = (Co^T B_P Cv)[i,a], so K and G are symmetric (ov) x (ov) matrices and
//   dL = dE/dL = (G + G^T) L = 2 G L,
//   dCo[mu,i] = sum_{a,P} dL[i,a,P] (B_P Cv)[mu,a],  dCv[nu,a] = sum_{i,P} dL[i,a,P] (B_P^T Co)[nu,i];
// no n^2 ov half-transform is formed (no symmetry of B_P in (mu,nu) is assumed).
//
// Orbital energies: with W = K (2K - K_ibja) / D^2 every eps enters D through two slots,
//   dE/de_k = -(sum_{a,jb} W[k,a,jb] + sum_{ia,b} W[ia,k,b])          (occupied k)
//   dE/de_{o+c} = sum_{i,jb} W[i,c,jb] + sum_{ia,j} W[ia,j,c]          (virtual c),
// i.e. r = row sums + column sums of W as an (ov) x (ov) matrix, then sums of r over a or i.
#include "api_util.h"
#include "jues_internal.h"

#include <algorithm>

namespace jues {
namespace {

// G[i,a,j,b] = 2(2K - Kx)/D and per-column energy partials (one block per (j,b)); with W
// non-null also W = K(2K - Kx)/D^2 and its column sums wcol[jb]
template <typename T>
__global__ __launch_bounds__(256) void mp2_g_kernel(const T* __restrict__ K, const double* __restrict__ eps, int o,
                                                    int v, T* __restrict__ G, double* __restrict__ part,
                                                    double* __restrict__ W, double* __restrict__ wcol) {
  __shared__ double sh[8];
  const int ov = o * v;
  const int jb = blockIdx.x;
  const int j = jb % o, b = jb / o;
  const double ejb = eps[j] - eps[o + b];
  const long col = (long)jb * ov;
  double s = 0.0, sw = 0.0;
  for (int ia = threadIdx.x; ia < ov; ia += 256) {
    const int i = ia % o, a = ia / o;
    const double k = (double)K[col + ia];
    const double kx = (double)K[(long)(i + o * b) + (long)ov * (j + o * a)];
    const double d = eps[i] - eps[o + a] + ejb;
    s += k * (2.0 * k - kx) / d;
    G[col + ia] = (T)(2.0 * (2.0 * k - kx) / d);
    if (W) {
      const double w = k * (2.0 * k - kx) / (d * d);
      W[col + ia] = w;
      sw += w;
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    s += __shfl_down(s, off, 64);
    sw += __shfl_down(sw, off, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    sh[threadIdx.x >> 6] = s;
    sh[4 + (threadIdx.x >> 6)] = sw;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    part[jb] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
    if (W) wcol[jb] = (sh[4] + sh[5]) + (sh[6] + sh[7]);
  }
}

__global__ __launch_bounds__(256) void sum_fixed_kernel2(const double* __restrict__ part, int count,
                                                         double* __restrict__ out) {
  __shared__ double sh[256];
  double s = 0.0;
  for (int i = threadIdx.x; i < count; i += 256) s += part[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = sh[0];
}

// rowpart[c, ia] = sum over the columns jb of chunk c of W[ia, jb] (coalesced along ia)
__global__ __launch_bounds__(256) void mp2_w_rowsum_kernel(const double* __restrict__ W, long ov, long chunk,
                                                           double* __restrict__ rowpart) {
  const long ia = blockIdx.x * 256L + threadIdx.x;
  if (ia >= ov) return;
  const long jb0 = blockIdx.y * chunk, jb1 = std::min(ov, jb0 + chunk);
  double s = 0.0;
  for (long jb = jb0; jb < jb1; ++jb) s += W[ia + ov * jb];
  rowpart[blockIdx.y * ov + ia] = s;
}

// r[ia] = wcol[ia] + sum_c rowpart[c, ia]  (fixed order)
__global__ __launch_bounds__(256) void mp2_w_r_kernel(const double* __restrict__ rowpart, int nchunk,
                                                      const double* __restrict__ wcol, long ov,
                                                      double* __restrict__ r) {
  const long ia = blockIdx.x * 256L + threadIdx.x;
  if (ia >= ov) return;
  double s = 0.0;
  for (int c = 0; c < nchunk; ++c) s += rowpart[c * ov + ia];
  r[ia] = wcol[ia] + s;
}

// deps[k] = -sum_a r[k,a] (occupied k), deps[o+c] = sum_i r[i,c] (virtual c)
__global__ __launch_bounds__(256) void mp2_deps_kernel(const double* __restrict__ r, int o, int v,
                                                       double* __restrict__ deps) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= o + v) return;
  double s = 0.0;
  if (p < o) {
    for (int a = 0; a < v; ++a) s += r[p + (long)o * a];
    deps[p] = -s;
  } else {
    const long c = p - o;
    for (int i = 0; i < o; ++i) s += r[i + o * c];
    deps[p] = s;
  }
}

// G = dE/dK from K, the energy into *E (host) and, with deps non-null, dE/deps (device, o+v)
template <typename T>
void mp2_g(jues_ctx* ctx, const T* K, const double* eps, int o, int v, T* G, double* E, double* deps) {
  const long ov = (long)o * v, N = ov * ov;
  DevBuf part((ov + 1) * sizeof(double));
  DevBuf Wb(deps ? N * sizeof(double) : 0), wcol(deps ? ov * sizeof(double) : 0);
  {
    KTimer t(ctx, "mp2_grad_g", 0.0, (double)N * (sizeof(T) * 3 + (deps ? 8 : 0)));
    hipLaunchKernelGGL((mp2_g_kernel<T>), dim3((unsigned)ov), dim3(256), 0, ctx->stream, K, eps, o, v, G,
                       part.as<double>(), deps ? Wb.as<double>() : nullptr, deps ? wcol.as<double>() : nullptr);
    hipLaunchKernelGGL(sum_fixed_kernel2, dim3(1), dim3(256), 0, ctx->stream, part.as<double>(), (int)ov,
                       part.as<double>() + ov);
    JUES_HIP(hipGetLastError());
  }
  if (deps) {
    KTimer t(ctx, "mp2_grad_eps", 0.0, (double)N * 8);
    const long chunk = 128, nchunk = (ov + chunk - 1) / chunk;
    DevBuf rowpart((size_t)(nchunk * ov) * sizeof(double)), r(ov * sizeof(double));
    const unsigned gx = (unsigned)((ov + 255) / 256);
    hipLaunchKernelGGL(mp2_w_rowsum_kernel, dim3(gx, (unsigned)nchunk), dim3(256), 0, ctx->stream, Wb.as<double>(),
                       ov, chunk, rowpart.as<double>());
    hipLaunchKernelGGL(mp2_w_r_kernel, dim3(gx), dim3(256), 0, ctx->stream, rowpart.as<double>(), (int)nchunk,
                       wcol.as<double>(), ov, r.as<double>());
    hipLaunchKernelGGL(mp2_deps_kernel, dim3((unsigned)((o + v + 255) / 256)), dim3(256), 0, ctx->stream,
                       r.as<double>(), o, v, deps);
    JUES_HIP(hipGetLastError());
    sync(ctx);
  }
  d2h(ctx, E, part.as<double>() + ov, sizeof(double));
}

// Factor route of the adjoint: dE/dC from L = Co^T B_P Cv without any n^2 ov intermediate.
template <typename T>
void mp2_grad_factors_dev(jues_ctx* ctx, const T* Co, const T* Cv, int o, int v, const double* eps, double* E, T* dC,
                          double* deps) {
  const long n = ctx->n, naux = ctx->naux;
  const long ov = (long)o * v, N = ov * ov;
  const T* B = ao_factors<T>(ctx);
  const T one = T(1), zero = T(0);
  DevBuf Lb(ov * naux * sizeof(T)), dLb(ov * naux * sizeof(T));
  T* L = Lb.as<T>();
  T* dL = dLb.as<T>();
  mo_factors_dev<T>(ctx, Co, o, Cv, v, L);
  {
    DevBuf Kb(N * sizeof(T)), Gb(N * sizeof(T));
    T* K = Kb.as<T>();
    T* G = Gb.as<T>();
    gemm<T>(ctx, false, true, ov, ov, naux, one, L, ov, 0, L, ov, 0, zero, K, ov, 0, 1, "mp2grad_k");
    mp2_g<T>(ctx, K, eps, o, v, G, E, deps);
    Kb.release();
    // dL = (G + G^T) L = 2 G L (G is symmetric: K = L L^T is)
    gemm<T>(ctx, false, false, ov, naux, ov, T(2), G, ov, 0, L, ov, 0, zero, dL, ov, 0, 1, "mp2grad_dl");
  }
  Lb.release();
  T* dCo = dC;          // n x o (leading dimension n)
  T* dCv = dC + n * o;  // n x v
  {
    // Z[mu,a,P] = (B_P Cv)[mu,a];  dCo = Z (n x v naux) . dL^T (v naux x o)
    DevBuf Z(n * v * naux * sizeof(T));
    gemm<T>(ctx, false, false, n, v, n, one, B, n, n * n, Cv, n, 0, zero, Z.as<T>(), n, n * v, (int)naux,
            "mp2grad_bc");
    gemm<T>(ctx, false, true, n, o, v * naux, one, Z.as<T>(), n, 0, dL, o, 0, zero, dCo, n, 0, 1, "mp2grad");
  }
  {
    // Wt[nu,i,P] = (B_P^T Co)[nu,i];  dLp[i,P,a] = dL[i,a,P];  dCv = Wt (n x o naux) . dLp (o naux x v)
    DevBuf Wt(n * o * naux * sizeof(T)), dLp(ov * naux * sizeof(T));
    gemm<T>(ctx, true, false, n, o, n, one, B, n, n * n, Co, n, 0, zero, Wt.as<T>(), n, n * o, (int)naux,
            "mp2grad_bc");
    permute<T, T>(ctx, dL, {o, v, naux}, {0, 2, 1}, dLp.as<T>(), 1.0, 0.0, "mp2grad_perm");
    gemm<T>(ctx, false, false, n, v, o * naux, one, Wt.as<T>(), n, 0, dLp.as<T>(), o * naux, 0, zero, dCv, n, 0, 1,
            "mp2grad");
  }
}

// Dense route of the adjoint (the AO tensor g is on the device); dg may be null.
//
// E depends on C at four positions of (ia|jb) = sum g[mu,nu,lam,sig] Co[mu,i] Cv[nu,a] Co[lam,j]
// Cv[sig,b].  The forward transform runs here with two of its intermediates kept:
//   Z1[i,nu,lam,sig] = sum_mu Co[mu,i] g            X2[i,nu,lam,b] = sum_sig Z1 Cv[sig,b]
//   X3[i,nu,j,b]     = sum_lam X2 Co[lam,j]         K[i,a,j,b]     = sum_nu X3 Cv[nu,a]
// (the last two batched, so nothing n^2-sized is permuted).  X2 (positions 1 and 4 transformed)
// gives dE/dC at position 3 and Y3[nu,i,b,j] = X3 permuted (positions 1, 3, 4) at position 2;
// positions 1 and 4 take one more half-transform, Y[mu,a,j,sig] = (mu a|j sig):
//   pos 1: dCo[mu,i]  += sum_{a,j,sig} Y[mu,a,j,sig] T1[i,a,j,sig],  T1 = sum_b G[i,a,j,b] Cv[sig,b]
//   pos 2: dCv[nu,a]   = sum_{i,b,j} Y3[nu,i,b,j] G[i,a,j,b]
//   pos 3: dCo[lam,j]  = sum_b sum_{i,nu} X2[i,nu,lam,b] Gt[i,nu,j,b],  Gt = sum_a G[i,a,j,b] Cv[nu,a]
//   pos 4: dCv[sig,b] += sum_{i,a,j} Yt[i,a,j,sig] G[i,a,j,b],     Yt = sum_mu Co[mu,i] Y
// dg (FP64 in both forms) may be null.
template <typename T>
void mp2_grad_dev(jues_ctx* ctx, const T* Co, const T* Cv, int o, int v, const double* eps, double* E, T* dC,
                  double* dg, double* deps) {
  const long n = ctx->n, n2 = n * n, n3 = n2 * n;
  const long ov = (long)o * v, N = ov * ov, on = (long)o * n, nov = n * ov;
  const T one = T(1), zero = T(0);
  const T* g = ao_dense<T>(ctx);
  T* dCo = dC;           // n x o  (columns 0..o-1 of dC, leading dimension n)
  T* dCv = dC + n * o;   // n x v
  DevBuf Gb(N * sizeof(T)), X2b(nov * n * sizeof(T)), Y3b(nov * o * sizeof(T));
  T* G = Gb.as<T>();
  T* X2 = X2b.as<T>();
  T* Y3 = Y3b.as<T>();
  {
    // forward (ia|jb) with sig contracted second (the (ia|jb) transform contracts lam second, which
    // costs less, but position 3 needs X2[i,nu,lam,b] = (i nu|lam b)), then lam and nu contracted in
    // place, batched, to X3[i,nu,j,b] = (i nu|j b) and K[i,a,j,b]
    {
      DevBuf Z1(o * n3 * sizeof(T));
      gemm<T>(ctx, true, false, o, n3, n, one, Co, n, 0, g, n, 0, zero, Z1.as<T>(), o, 0, 1, "xform_q1");
      gemm<T>(ctx, false, false, o * n2, v, n, one, Z1.as<T>(), o * n2, 0, Cv, n, 0, zero, X2, o * n2, 0, 1,
              "xform_q2");
    }
    DevBuf X3(nov * o * sizeof(T)), Kb(N * sizeof(T));
    gemm<T>(ctx, false, false, on, o, n, one, X2, on, on * n, Co, n, 0, zero, X3.as<T>(), on, on * o, v, "xform_q3");
    gemm<T>(ctx, false, false, o, v, n, one, X3.as<T>(), o, on, Cv, n, 0, zero, Kb.as<T>(), o, ov, (int)(o * v),
            "xform_q4");
    // Y3[nu,i,b,j] = (i nu|j b), position 2's operand (an n o^2 v permutation, not n^2 o v)
    permute<T, T>(ctx, X3.as<T>(), {(long)o, n, (long)o, (long)v}, {1, 0, 3, 2}, Y3, 1.0, 0.0, "mp2grad_s2");
    X3.release();
    mp2_g<T>(ctx, Kb.as<T>(), eps, o, v, G, E, deps);
  }
  // position 2: Gp2[i,b,j,a] = G[i,a,j,b]; dCv = Y3[nu,(i,b,j)] Gp2[(i,b,j),a]
  {
    DevBuf Gp2(N * sizeof(T));
    permute<T, T>(ctx, G, {(long)o, (long)v, (long)o, (long)v}, {0, 3, 2, 1}, Gp2.as<T>(), 1.0, 0.0, "mp2grad_s2");
    gemm<T>(ctx, false, false, n, v, ov * o, one, Y3, n, 0, Gp2.as<T>(), ov * o, 0, zero, dCv, n, 0, 1, "mp2grad_s2");
  }
  Y3b.release();
  // position 3: Gt[i,nu,j,b] (batched over jb; also the first step of the dE/dg back-transform);
  // dCo[lam,j] = sum_b X2[(i,nu),lam,b]^T Gt[(i,nu),j,b], batched over b and summed in a fixed order
  DevBuf Gt(on * ov * sizeof(T));
  gemm<T>(ctx, false, true, o, n, v, one, G, o, ov, Cv, n, 0, zero, Gt.as<T>(), o, on, (int)ov, "mp2grad_s3");
  {
    DevBuf W(n * ov * sizeof(T));
    gemm_batch_sum<T>(ctx, true, false, n, o, on, one, X2, on, on * n, Gt.as<T>(), on, on * o, zero, dCo, n, v,
                      W.as<T>(), "mp2grad_s3");
  }
  X2b.release();
  // Y[mu,a,j,sig] = (mu a|j sig): X[mu,nu,j,sig] = sum_lam g[mu,nu,lam,sig] Co[lam,j] (batched over sig),
  // then Y[mu,a,(j,sig)] = sum_nu X[mu,nu,(j,sig)] Cv[nu,a] (batched over (j,sig))
  DevBuf Yb(nov * n * sizeof(T));
  T* Y = Yb.as<T>();
  {
    DevBuf X(n3 * o * sizeof(T));
    gemm<T>(ctx, false, false, n2, o, n, one, g, n2, n3, Co, n, 0, zero, X.as<T>(), n2, n2 * o, (int)n,
            "mp2grad_half_x");
    gemm<T>(ctx, false, false, n, v, n, one, X.as<T>(), n, n2, Cv, n, 0, zero, Y, n, n * v, (int)on,
            "mp2grad_half_y");
  }
  // position 1: T1[(i,a,j),sig] = sum_b G[(i,a,j),b] Cv[sig,b]; dCo += Y[mu,(a,j,sig)] T1[i,(a,j,sig)]^T
  {
    DevBuf T1(ov * on * sizeof(T));
    gemm<T>(ctx, false, true, ov * o, n, v, one, G, ov * o, 0, Cv, n, 0, zero, T1.as<T>(), ov * o, 0, 1,
            "mp2grad_s1");
    gemm<T>(ctx, false, true, n, o, nov, one, Y, n, 0, T1.as<T>(), o, 0, one, dCo, n, 0, 1, "mp2grad_s1");
  }
  // position 4: Yt[i,(a,j,sig)] = sum_mu Co[mu,i] Y[mu,(a,j,sig)]; dCv += Yt[(i,a,j),sig]^T G[(i,a,j),b]
  {
    DevBuf Yt(ov * on * sizeof(T));
    gemm<T>(ctx, true, false, o, nov, n, one, Co, n, 0, Y, n, 0, zero, Yt.as<T>(), o, 0, 1, "mp2grad_s4");
    Yb.release();
    gemm<T>(ctx, true, false, n, v, ov * o, one, Yt.as<T>(), ov * o, 0, G, ov * o, 0, one, dCv, n, 0, 1,
            "mp2grad_s4");
  }
  if (dg) {
    // back-transform G -> dE/dg with the occupied index i contracted last, so that the n^4 result
    // comes out of one GEMM already in [mu,nu,lam,sig] order (no permutation pass):
    //   Gt[(i,nu),j,b]   = sum_a G[i,a,j,b] Cv[nu,a]             (position 3)
    //   P2[(i,nu),lam,b] = sum_j Gt[(i,nu),j,b] Co[lam,j]        (batched over b)
    //   Q[(i,nu,lam),sig] = sum_b P2[(i,nu,lam),b] Cv[sig,b]
    //   dg[mu,(nu,lam,sig)] = sum_i Co[mu,i] Q[i,(nu,lam,sig)]
    DevBuf P2(on * n * v * sizeof(T));
    gemm<T>(ctx, false, true, on, n, o, one, Gt.as<T>(), on, on * o, Co, n, 0, zero, P2.as<T>(), on, on * n, v,
            "mp2grad_back_p2");
    Gt.release();
    DevBuf Q(on * n2 * sizeof(T));
    gemm<T>(ctx, false, true, on * n, n, v, one, P2.as<T>(), on * n, 0, Cv, n, 0, zero, Q.as<T>(), on * n, 0, 1,
            "mp2grad_back_q");
    P2.release();
    if constexpr (sizeof(T) == sizeof(double)) {
      gemm<double>(ctx, false, false, n, n3, o, 1.0, Co, n, 0, Q.as<double>(), o, 0, 0.0, dg, n, 0, 1,
                   "mp2grad_back_dg");
    } else if (!gemm32_out64(ctx, n, n3, o, Co, n, Q.as<float>(), o, dg, n, "mp2grad_back_dg")) {
      // the last product writes the n^4 result once, in FP64: FP32 MFMA work with an FP64 epilogue
      // (above), else FP64 on FP64 copies of Q and Co (an FP32 result and a conversion would write
      // and re-read an FP32 copy of the n^4 tensor)
      DevBuf Q64(on * n2 * sizeof(double)), Co64((long)n * o * sizeof(double));
      convert<float, double>(ctx, Q.as<float>(), Q64.as<double>(), on * n2);
      convert<float, double>(ctx, Co, Co64.as<double>(), (long)n * o);
      Q.release();
      gemm<double>(ctx, false, false, n, n3, o, 1.0, Co64.as<double>(), n, 0, Q64.as<double>(), o, 0, 0.0, dg, n,
                   0, 1, "mp2grad_back_dg");
    }
  }
}

}  // namespace
}  // namespace jues

using namespace jues;

extern "C" jues_status jues_rmp2_grad(jues_ctx* ctx, int prec, const double* C, int o, int n, const double* eps,
                                      double* E, double* dE_dC, double* dE_dg, double* dE_deps) {
  JUES_API_BEGIN(ctx)
  JUES_REQUIRE(C && eps && E && o > 0 && o < n, "jues_rmp2_grad: bad arguments");
  JUES_REQUIRE(n == ctx->n && n > 0, "jues_rmp2_grad: n does not match the AO integrals");
  JUES_REQUIRE(!dE_dg || ctx->gao.p, "jues_rmp2_grad: dE/dg needs dense AO integrals");
  JUES_REQUIRE(prec == JUES_FP64 || prec == JUES_FP32_MIXED, "jues_rmp2_grad: bad precision");
  const int v = n - o;
  DeviceArray<double> e(ctx, eps, (size_t)n);
  DevBuf de(dE_deps ? (size_t)n * sizeof(double) : 0);
  double* deps = dE_deps ? de.as<double>() : nullptr;
  const size_t nn = (size_t)n * n, n4 = nn * nn;
  // dE/dg needs the dense route; otherwise factors take the factor route
  const bool factor_route = !ctx->gao.p;
  auto run = [&](auto zero) {
    using T = decltype(zero);
    DeviceArray<T> c(ctx, C, nn);
    DevBuf dC(nn * sizeof(T));
    DevBuf dg(dE_dg ? n4 * sizeof(double) : 0);  // dE/dg is formed in FP64 in both forms
    if (factor_route)
      mp2_grad_factors_dev<T>(ctx, c.p, c.p + (size_t)n * o, o, v, e.p, E, dC.as<T>(), deps);
    else
      mp2_grad_dev<T>(ctx, c.p, c.p + (size_t)n * o, o, v, e.p, E, dC.as<T>(), dE_dg ? dg.as<double>() : nullptr,
                      deps);
    // results back in FP64 (an FP32 stage converts on the device first)
    auto fetch = [&](double* host, const DevBuf& dev, size_t count) {
      if constexpr (sizeof(T) == sizeof(double)) {
        d2h(ctx, host, dev.p, count * sizeof(double));
      } else {
        DevBuf d64(count * sizeof(double));
        convert<T, double>(ctx, dev.as<T>(), d64.as<double>(), (long)count);
        d2h(ctx, host, d64.p, count * sizeof(double));
      }
    };
    if (dE_dC) fetch(dE_dC, dC, nn);
    if (dE_dg) d2h(ctx, dE_dg, dg.p, n4 * sizeof(double));
  };
  if (prec == JUES_FP64) run(0.0);
  else run(0.0f);
  if (dE_deps) d2h(ctx, dE_deps, deps, (size_t)n * sizeof(double));
  if (!std::isfinite(*E)) throw Error(JUES_NONFINITE, "jues_rmp2_grad: non-finite energy (a zero denominator?)");
  JUES_API_END(ctx)
}
