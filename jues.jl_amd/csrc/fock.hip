// Open-shell MP2 and the MO Fock matrix: MollerPlesset/UMP2.jl:12-79 and
// IntegralTransformation.get_fock (Backend/IntegralTransformation.jl:119-141).
#include "jues_internal.h"
#include "api_util.h"

#include <cstring>

using namespace jues;

namespace jues {
namespace {

template <typename T>
double ump2_dev(jues_ctx* ctx, const double* Cao, int na, const double* Cav, int va, const double* Cbo, int nb,
                const double* Cbv, int vb, const double* epsa, const double* epsb) {
  const int n = ctx->n;
  DeviceArray<T> cao(ctx, Cao, (size_t)n * na), cav(ctx, Cav, (size_t)n * va), cbo(ctx, Cbo, (size_t)n * nb),
      cbv(ctx, Cbv, (size_t)n * vb);
  DeviceArray<double> ea(ctx, epsa, (size_t)na + va), eb(ctx, epsb, (size_t)nb + vb);
  double E = 0.0;
  {
    DevBuf K(sizeof(T) * (size_t)na * va * na * va);
    tei_transform_dev<T>(ctx, cao.p, na, cav.p, va, cao.p, na, cav.p, va, K.as<T>());       // (ia|jb)
    E += ump2_block_dev<T>(ctx, K.as<T>(), ea.p, ea.p, na, va, na, va, true);
  }
  {
    DevBuf K(sizeof(T) * (size_t)na * va * nb * vb);
    tei_transform_dev<T>(ctx, cao.p, na, cav.p, va, cbo.p, nb, cbv.p, vb, K.as<T>());       // (ia|JB)
    E += ump2_block_dev<T>(ctx, K.as<T>(), ea.p, eb.p, na, va, nb, vb, false);
  }
  if (nb > 0 && vb > 0) {
    DevBuf K(sizeof(T) * (size_t)nb * vb * nb * vb);
    tei_transform_dev<T>(ctx, cbo.p, nb, cbv.p, vb, cbo.p, nb, cbv.p, vb, K.as<T>());       // (IA|JB)
    E += ump2_block_dev<T>(ctx, K.as<T>(), eb.p, eb.p, nb, vb, nb, vb, true);
  }
  return E;
}

}  // namespace
}  // namespace jues

extern "C" {

jues_status jues_ump2(jues_ctx* ctx, int prec, const double* Cao, int na, const double* Cav, int va,
                      const double* Cbo, int nb, const double* Cbv, int vb, int n, const double* epsa,
                      const double* epsb, double* E) {
  JUES_API_BEGIN(ctx)
  JUES_REQUIRE(Cao && Cav && epsa && epsb && E && (Cbo || nb == 0) && (Cbv || vb == 0),
               "jues_ump2: null pointer");
  JUES_REQUIRE(na > 0 && va > 0 && nb >= 0 && vb >= 0, "jues_ump2: bad orbital counts");
  JUES_REQUIRE(n == ctx->n && n > 0, "jues_ump2: n does not match the AO integrals");
  JUES_REQUIRE(prec == JUES_FP64 || prec == JUES_FP32_MIXED, "jues_ump2: bad precision");
  *E = prec == JUES_FP64 ? ump2_dev<double>(ctx, Cao, na, Cav, va, Cbo, nb, Cbv, vb, epsa, epsb)
                         : ump2_dev<float>(ctx, Cao, na, Cav, va, Cbo, nb, Cbv, vb, epsa, epsb);
  if (!std::isfinite(*E)) throw Error(JUES_NONFINITE, "jues_ump2: non-finite energy (a zero denominator?)");
  JUES_API_END(ctx)
}

// f[p,q] = (C^T (h + 2 J[D] - K[D]) C)[p,q] with D = Co Co^T:
//   dense:   J[mu,nu] = sum g[mu,nu,lam,sig] D[lam,sig],  K[mu,nu] = sum g[mu,lam,nu,sig] D[lam,sig];
//   factors: J = sum_P B_P tr(B_P D),                     K = sum_P B_P D B_P^T.
jues_status jues_get_fock(jues_ctx* ctx, const double* hao, const double* C, int m, const double* Co, int o, int n,
                          double* f) {
  JUES_API_BEGIN(ctx)
  JUES_REQUIRE(hao && C && Co && f && m > 0 && o >= 0, "jues_get_fock: bad arguments");
  JUES_REQUIRE(n == ctx->n && n > 0, "jues_get_fock: n does not match the AO integrals");
  const long nn = (long)n * n;
  DeviceArray<double> h(ctx, hao, (size_t)nn), c(ctx, C, (size_t)n * m), co(ctx, Co, (size_t)n * o);
  DevBuf D(sizeof(double) * nn), F(sizeof(double) * nn), X(sizeof(double) * (size_t)n * m);
  gemm<double>(ctx, false, true, n, n, o, 1.0, co.p, n, 0, co.p, n, 0, 0.0, D.as<double>(), n, 0, 1, "fock");
  JUES_HIP(hipMemcpyAsync(F.p, h.p, sizeof(double) * nn, hipMemcpyDeviceToDevice, ctx->stream));
  auto TN = [&](const double* p, const char* idx, std::vector<long> dims) {
    Tens<double> t;
    t.p = const_cast<double*>(p);
    t.idx = idx;
    t.dims = std::move(dims);
    return t;
  };
  const long nl = n;
  if (ctx->gao.p) {
    const double* g = ctx->gao.as<double>();
    contract<double>(ctx, 2.0, TN(g, "mnls", {nl, nl, nl, nl}), TN(D.as<double>(), "ls", {nl, nl}), 1.0,
                     TN(F.as<double>(), "mn", {nl, nl}), "fock");
    contract<double>(ctx, -1.0, TN(g, "mlns", {nl, nl, nl, nl}), TN(D.as<double>(), "ls", {nl, nl}), 1.0,
                     TN(F.as<double>(), "mn", {nl, nl}), "fock");
  } else {
    JUES_REQUIRE(ctx->Bfac.p, "jues_get_fock: no AO integrals set");
    const long na = ctx->naux;
    const double* B = ctx->Bfac.as<double>();
    DevBuf vP(sizeof(double) * (size_t)na), Y(sizeof(double) * (size_t)nn * na);
    contract<double>(ctx, 1.0, TN(B, "lsP", {nl, nl, na}), TN(D.as<double>(), "ls", {nl, nl}), 0.0,
                     TN(vP.as<double>(), "P", {na}), "fock");
    contract<double>(ctx, 2.0, TN(B, "mnP", {nl, nl, na}), TN(vP.as<double>(), "P", {na}), 1.0,
                     TN(F.as<double>(), "mn", {nl, nl}), "fock");
    contract<double>(ctx, 1.0, TN(B, "mlP", {nl, nl, na}), TN(D.as<double>(), "ls", {nl, nl}), 0.0,
                     TN(Y.as<double>(), "msP", {nl, nl, na}), "fock");
    contract<double>(ctx, -1.0, TN(Y.as<double>(), "msP", {nl, nl, na}), TN(B, "nsP", {nl, nl, na}), 1.0,
                     TN(F.as<double>(), "mn", {nl, nl}), "fock");
  }
  // f = C^T F C
  gemm<double>(ctx, false, false, n, m, n, 1.0, F.as<double>(), n, 0, c.p, n, 0, 0.0, X.as<double>(), n, 0, 1,
               "fock");
  DevBuf fo(sizeof(double) * (size_t)m * m);
  gemm<double>(ctx, true, false, m, m, n, 1.0, c.p, n, 0, X.as<double>(), n, 0, 0.0, fo.as<double>(), m, 0, 1,
               "fock");
  d2h(ctx, f, fo.p, sizeof(double) * (size_t)m * m);
  JUES_API_END(ctx)
}

}  // extern "C"
