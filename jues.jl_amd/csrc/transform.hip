// AO -> MO four-index transformation (Backend/Transformation.jl:39-93) and the
// RMP2 energy reduction (MollerPlesset/RMP2.jl:27-41).
//
// Dense route: four quarter transforms, each one MFMA GEMM (plain or batched)
// over the full AO tensor.  The outer index (mu or sigma) with the smaller target
// dimension is contracted first.  If the inner index next to the other outer one
// has a smaller target than it (lambda for mu first, as d3 = o < d4 = v in
// (ia|jb)), that inner index is contracted second, batched, which makes the
// second quarter 2 n^3 d1 d3 instead of 2 n^3 d1 d4; otherwise the other outer
// index is, and the last two quarters contract nu and lambda in place (the
// smaller of d2, d3 first).  Every order writes the result in [p, q, r, s]
// order from its last GEMM, with no permutation pass.  Factor route:
// (pq|rs) = sum_P (C1^T B_P C2)_pq (C3^T B_P C4)_rs, two batched GEMMs per side
// and one GEMM over P; the n^4 AO tensor is never formed.
#include "jues_internal.h"

#include <algorithm>

namespace jues {

template <>
const double* ao_dense<double>(jues_ctx* ctx) {
  JUES_REQUIRE(ctx->gao.p, "no dense AO integrals set");
  return ctx->gao.as<double>();
}
template <>
const float* ao_dense<float>(jues_ctx* ctx) {
  JUES_REQUIRE(ctx->gao.p, "no dense AO integrals set");
  if (!ctx->gao32.p) {
    const long nn = (long)ctx->n * ctx->n * ctx->n * ctx->n;
    ctx->gao32.alloc(sizeof(float) * nn);
    convert<double, float>(ctx, ctx->gao.as<double>(), ctx->gao32.as<float>(), nn);
  }
  return ctx->gao32.as<float>();
}
template <>
const double* ao_factors<double>(jues_ctx* ctx) {
  JUES_REQUIRE(ctx->Bfac.p, "no AO factors set");
  return ctx->Bfac.as<double>();
}
template <>
const float* ao_factors<float>(jues_ctx* ctx) {
  JUES_REQUIRE(ctx->Bfac.p, "no AO factors set");
  if (!ctx->Bfac32.p) {
    const long nn = (long)ctx->n * ctx->n * ctx->naux;
    ctx->Bfac32.alloc(sizeof(float) * nn);
    convert<double, float>(ctx, ctx->Bfac.as<double>(), ctx->Bfac32.as<float>(), nn);
  }
  return ctx->Bfac32.as<float>();
}

template <typename T>
void mo_factors_dev(jues_ctx* ctx, const T* C1, int d1, const T* C2, int d2, T* L) {
  const long n = ctx->n, naux = ctx->naux;
  T* Y = static_cast<T*>(scratch(ctx, 4, sizeof(T) * (size_t)(n * std::min(d1, d2) * naux)));
  mo_factors_dev<T>(ctx, C1, d1, C2, d2, L, Y);
}

template <typename T>
void mo_factors_dev(jues_ctx* ctx, const T* C1, int d1, const T* C2, int d2, T* L, T* Y) {
  const long n = ctx->n, naux = ctx->naux;
  const T* B = ao_factors<T>(ctx);
  // The smaller of d1, d2 is contracted first: 2 n^2 min(d1,d2) naux + 2 n d1 d2 naux flops.
  if (d1 <= d2) {
    // Y[nu, p, P] = B_P^T C1 ; L[p, q, P] = Y_P^T C2
    gemm<T>(ctx, true, false, n, d1, n, T(1), B, n, n * n, C1, n, 0, T(0), Y, n, n * d1, (int)naux,
            "xform_factor");
    gemm<T>(ctx, true, false, d1, d2, n, T(1), Y, n, n * d1, C2, n, 0, T(0), L, d1, (long)d1 * d2, (int)naux,
            "xform_factor");
  } else {
    // Y[mu, q, P] = B_P C2 ; L[p, q, P] = C1^T Y_P
    gemm<T>(ctx, false, false, n, d2, n, T(1), B, n, n * n, C2, n, 0, T(0), Y, n, n * d2, (int)naux,
            "xform_factor");
    gemm<T>(ctx, true, false, d1, d2, n, T(1), C1, n, 0, Y, n, n * d2, T(0), L, d1, (long)d1 * d2, (int)naux,
            "xform_factor");
  }
}

template <typename T>
void tei_transform_dev(jues_ctx* ctx, const T* C1, int d1, const T* C2, int d2, const T* C3, int d3,
                       const T* C4, int d4, T* out) {
  const long n = ctx->n;
  JUES_REQUIRE(n > 0, "no AO integrals set");
  if (ctx->gao.p) {
    const T* G = ao_dense<T>(ctx);
    const long n2 = n * n, n3 = n2 * n;
    const long d12 = (long)d1 * d2;
    if (d1 <= d4 && d3 < d4) {
      // mu first, then lam (a smaller target than sig's), batched over sig: the second quarter
      // costs 2 n^3 d1 d3 instead of 2 n^3 d1 d4, and each product writes its result in the
      // order the next one reads
      //   X1[p, (nu,lam,sig)] = sum_mu C1[mu, p] g[mu, (nu,lam,sig)]
      //   X2[(p,nu), r, sig]  = sum_lam X1[(p,nu), lam, sig] C3[lam, r]   (batched over sig)
      //   X3[(p,nu,r), s]     = sum_sig X2[(p,nu,r), sig] C4[sig, s]
      //   out[p, q, (r,s)]    = sum_nu X3[p, nu, (r,s)] C2[nu, q]        (batched over (r,s))
      const long pn = (long)d1 * n, pnr = pn * d3;
      T* X1 = static_cast<T*>(scratch(ctx, 0, sizeof(T) * (size_t)(d1 * n3)));
      gemm<T>(ctx, true, false, d1, n3, n, T(1), C1, n, 0, G, n, 0, T(0), X1, d1, 0, 1, "xform_q1");
      T* X2 = static_cast<T*>(scratch(ctx, 1, sizeof(T) * (size_t)(pnr * n)));
      gemm<T>(ctx, false, false, pn, d3, n, T(1), X1, pn, pn * n, C3, n, 0, T(0), X2, pn, pnr, (int)n, "xform_q2");
      T* X3 = static_cast<T*>(scratch(ctx, 2, sizeof(T) * (size_t)(pnr * d4)));
      gemm<T>(ctx, false, false, pnr, d4, n, T(1), X2, pnr, 0, C4, n, 0, T(0), X3, pnr, 0, 1, "xform_q3");
      gemm<T>(ctx, false, false, d1, d2, n, T(1), X3, d1, pn, C2, n, 0, T(0), out, d1, d12, d3 * d4, "xform_q4");
    } else if (d1 > d4 && d2 < d1) {
      // the mirror image: sig first, then nu (a smaller target than mu's), batched over (lam,s)
      //   Y1[(mu,nu,lam), s] = sum_sig g[(mu,nu,lam), sig] C4[sig, s]
      //   Y2[mu, q, (lam,s)] = sum_nu Y1[mu, nu, (lam,s)] C2[nu, q]      (batched over (lam,s))
      //   X3[p, (q,lam,s)]   = sum_mu C1[mu, p] Y2[mu, (q,lam,s)]
      //   out[(p,q), r, s]   = sum_lam X3[(p,q), lam, s] C3[lam, r]      (batched over s)
      T* Y1 = static_cast<T*>(scratch(ctx, 0, sizeof(T) * (size_t)(n3 * d4)));
      gemm<T>(ctx, false, false, n3, d4, n, T(1), G, n3, 0, C4, n, 0, T(0), Y1, n3, 0, 1, "xform_q1");
      T* Y2 = static_cast<T*>(scratch(ctx, 1, sizeof(T) * (size_t)(n * d2 * n * d4)));
      gemm<T>(ctx, false, false, n, d2, n, T(1), Y1, n, n2, C2, n, 0, T(0), Y2, n, n * d2, (int)(n * d4), "xform_q2");
      T* X3 = static_cast<T*>(scratch(ctx, 2, sizeof(T) * (size_t)(d12 * n * d4)));
      gemm<T>(ctx, true, false, d1, d2 * n * d4, n, T(1), C1, n, 0, Y2, n, 0, T(0), X3, d1, 0, 1, "xform_q3");
      gemm<T>(ctx, false, false, d12, d3, n, T(1), X3, d12, d12 * n, C3, n, 0, T(0), out, d12, d12 * d3, d4,
              "xform_q4");
    } else {
      // the two outer indices first (the one with the smaller target first), to X2[p, nu, lam, s]
      T* X2 = static_cast<T*>(scratch(ctx, 1, sizeof(T) * (size_t)(d1 * n2 * d4)));
      if (d1 <= d4) {
        T* X1 = static_cast<T*>(scratch(ctx, 0, sizeof(T) * (size_t)(d1 * n3)));
        gemm<T>(ctx, true, false, d1, n3, n, T(1), C1, n, 0, G, n, 0, T(0), X1, d1, 0, 1, "xform_q1");
        gemm<T>(ctx, false, false, d1 * n2, d4, n, T(1), X1, d1 * n2, 0, C4, n, 0, T(0), X2, d1 * n2, 0, 1,
                "xform_q2");
      } else {
        T* Y1 = static_cast<T*>(scratch(ctx, 0, sizeof(T) * (size_t)(n3 * d4)));
        gemm<T>(ctx, false, false, n3, d4, n, T(1), G, n3, 0, C4, n, 0, T(0), Y1, n3, 0, 1, "xform_q1");
        gemm<T>(ctx, true, false, d1, n2 * d4, n, T(1), C1, n, 0, Y1, n, 0, T(0), X2, d1, 0, 1, "xform_q2");
      }
      // then nu and lam where they lie, as batched products, so that the result is written in
      // [p, q, r, s] order with no permutation pass
      if (d3 <= d2) {
        // X3[(p,nu), r, s] = sum_lam X2[(p,nu), lam, s] C3[lam, r]      (batched over s)
        // out[p, q, (r,s)] = sum_nu X3[p, nu, (r,s)] C2[nu, q]         (batched over (r,s))
        T* X3 = static_cast<T*>(scratch(ctx, 2, sizeof(T) * (size_t)((long)d1 * n * d3 * d4)));
        gemm<T>(ctx, false, false, (long)d1 * n, d3, n, T(1), X2, (long)d1 * n, (long)d1 * n2, C3, n, 0, T(0), X3,
                (long)d1 * n, (long)d1 * n * d3, d4, "xform_q3");
        gemm<T>(ctx, false, false, d1, d2, n, T(1), X3, d1, (long)d1 * n, C2, n, 0, T(0), out, d1, d12, d3 * d4,
                "xform_q4");
      } else {
        // X3[p, q, (lam,s)] = sum_nu X2[p, nu, (lam,s)] C2[nu, q]      (batched over (lam,s))
        // out[(p,q), r, s] = sum_lam X3[(p,q), lam, s] C3[lam, r]      (batched over s)
        T* X3 = static_cast<T*>(scratch(ctx, 2, sizeof(T) * (size_t)(d12 * n * d4)));
        gemm<T>(ctx, false, false, d1, d2, n, T(1), X2, d1, (long)d1 * n, C2, n, 0, T(0), X3, d1, d12,
                (int)(n * d4), "xform_q3");
        gemm<T>(ctx, false, false, d12, d3, n, T(1), X3, d12, d12 * n, C3, n, 0, T(0), out, d12, d12 * d3, d4,
                "xform_q4");
      }
    }
  } else {
    const long naux = ctx->naux;
    JUES_REQUIRE(ctx->Bfac.p, "no AO integrals set");
    T* L = static_cast<T*>(scratch(ctx, 5, sizeof(T) * (size_t)((long)d1 * d2 * naux)));
    mo_factors_dev<T>(ctx, C1, d1, C2, d2, L);
    T* R = L;  // (pq|pq)-shaped blocks such as (ia|jb): both sides are the same factors
    if (C1 != C3 || C2 != C4 || d1 != d3 || d2 != d4) {
      R = static_cast<T*>(scratch(ctx, 6, sizeof(T) * (size_t)((long)d3 * d4 * naux)));
      mo_factors_dev<T>(ctx, C3, d3, C4, d4, R);
    }
    const long m = (long)d1 * d2, nn = (long)d3 * d4;
    gemm<T>(ctx, false, true, m, nn, naux, T(1), L, m, 0, R, nn, 0, T(0), out, m, 0, 1, "xform_factor_gemm");
  }
}

namespace {

// One block per (j,b) column of K[i,a,j,b]; each thread walks (i,a) with stride 256.
template <typename T>
__global__ __launch_bounds__(256) void mp2_energy_kernel(const T* __restrict__ K, const double* __restrict__ eps,
                                                         int o, int v, double* __restrict__ part) {
  __shared__ double sh[4];
  const int ov = o * v;
  const int jb = blockIdx.x;
  const int j = jb % o, b = jb / o;
  const double ejb = eps[j] - eps[o + b];
  const long col = (long)jb * ov;
  double s = 0.0;
  for (int ia = threadIdx.x; ia < ov; ia += 256) {
    const int i = ia % o, a = ia / o;
    const double k = (double)K[col + ia];
    const double kx = (double)K[(long)(i + o * b) + (long)ov * (j + o * a)];  // (ib|ja)
    s += k * (2.0 * k - kx) / (eps[i] - eps[o + a] + ejb);
  }
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) part[jb] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

__global__ __launch_bounds__(256) void sum_fixed_kernel(const double* __restrict__ part, int count,
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

// UMP2 blocks (UMP2.jl:30-77).  One block per (J,B) column of K[i,a,J,B] = (ia|JB).
// same != 0 (alpha-alpha or beta-beta, K square): accumulates (ia|jb)((ia|jb) - (ib|ja)) / D,
// half of which is the sum over i > j, a > b of <ij||ab>^2 / D; otherwise (ia|JB)^2 / D.
template <typename T>
__global__ __launch_bounds__(256) void ump2_energy_kernel(const T* __restrict__ K, const double* __restrict__ eA,
                                                          const double* __restrict__ eB, int oA, int vA, int oB,
                                                          int same, double* __restrict__ part) {
  __shared__ double sh[4];
  const int ovA = oA * vA;
  const int jb = blockIdx.x;
  const int j = jb % oB, b = jb / oB;
  const double ejb = eB[j] - eB[oB + b];
  const long col = (long)jb * ovA;
  double s = 0.0;
  for (int ia = threadIdx.x; ia < ovA; ia += 256) {
    const int i = ia % oA, a = ia / oA;
    const double k = (double)K[col + ia];
    const double d = eA[i] - eA[oA + a] + ejb;
    if (same) {
      const double kx = (double)K[(long)(i + oA * b) + (long)ovA * (j + oA * a)];  // (ib|ja)
      s += k * (k - kx) / d;
    } else {
      s += k * k / d;
    }
  }
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) part[jb] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

}  // namespace

template <typename T>
double ump2_block_dev(jues_ctx* ctx, const T* K, const double* eA, const double* eB, int oA, int vA, int oB, int vB,
                      bool same) {
  const int cols = oB * vB;
  double* part = static_cast<double*>(scratch(ctx, 14, sizeof(double) * (size_t)(cols + 1)));
  {
    KTimer t(ctx, "ump2_energy", 4.0 * oA * vA * (double)cols, (double)oA * vA * cols * sizeof(T) * (same ? 2 : 1));
    hipLaunchKernelGGL((ump2_energy_kernel<T>), dim3(cols), dim3(256), 0, ctx->stream, K, eA, eB, oA, vA, oB,
                       same ? 1 : 0, part);
    hipLaunchKernelGGL(sum_fixed_kernel, dim3(1), dim3(256), 0, ctx->stream, part, cols, part + cols);
    JUES_HIP(hipGetLastError());
  }
  double E = 0.0;
  d2h(ctx, &E, part + cols, sizeof(double));
  return same ? 0.5 * E : E;
}

template <typename T>
double mp2_energy_dev(jues_ctx* ctx, const T* K, const double* eps, int o, int v) {
  const int ov = o * v;
  double* part = static_cast<double*>(scratch(ctx, 14, sizeof(double) * (size_t)(ov + 1)));
  {
    KTimer t(ctx, "mp2_energy", 6.0 * ov * ov, (double)ov * ov * sizeof(T) * 2);
    hipLaunchKernelGGL((mp2_energy_kernel<T>), dim3(ov), dim3(256), 0, ctx->stream, K, eps, o, v, part);
    hipLaunchKernelGGL(sum_fixed_kernel, dim3(1), dim3(256), 0, ctx->stream, part, ov, part + ov);
    JUES_HIP(hipGetLastError());
  }
  double E = 0.0;
  d2h(ctx, &E, part + ov, sizeof(double));
  return E;
}

template void mo_factors_dev<double>(jues_ctx*, const double*, int, const double*, int, double*);
template void mo_factors_dev<float>(jues_ctx*, const float*, int, const float*, int, float*);
template void mo_factors_dev<double>(jues_ctx*, const double*, int, const double*, int, double*, double*);
template void mo_factors_dev<float>(jues_ctx*, const float*, int, const float*, int, float*, float*);
template void tei_transform_dev<double>(jues_ctx*, const double*, int, const double*, int, const double*, int,
                                        const double*, int, double*);
template void tei_transform_dev<float>(jues_ctx*, const float*, int, const float*, int, const float*, int,
                                       const float*, int, float*);
template double mp2_energy_dev<double>(jues_ctx*, const double*, const double*, int, int);
template double mp2_energy_dev<float>(jues_ctx*, const float*, const double*, int, int);
template double ump2_block_dev<double>(jues_ctx*, const double*, const double*, const double*, int, int, int, int,
                                       bool);
template double ump2_block_dev<float>(jues_ctx*, const float*, const double*, const double*, int, int, int, int,
                                      bool);

}  // namespace jues
