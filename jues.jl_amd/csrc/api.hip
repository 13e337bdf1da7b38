// C ABI entry points (include/jues.h): context, AO integrals, tei_transform,
// get_eri, RMP2 and the GEMM test hook.  Coupled-cluster entry points live in cc.hip.
#include "jues_internal.h"
#include "api_util.h"

#include <cstring>

using namespace jues;

extern "C" {

jues_status jues_create(jues_ctx** out, int device) {
  if (!out) return JUES_BAD_ARG;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) return JUES_HIP_ERROR;
  if (hipSetDevice(device) != hipSuccess) return JUES_HIP_ERROR;
  try {
    noalloc_attach(device);  // the allocation-free entry point's scratch address on this device
  } catch (...) {
    (void)hipGetLastError();
    return JUES_HIP_ERROR;
  }
  jues_ctx* c = new jues_ctx();
  c->device = device;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return JUES_HIP_ERROR;
  }
  *out = c;
  return JUES_OK;
}

void jues_destroy(jues_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  ctx->cc.reset();
  release_scratch(ctx);
  for (auto& p : ctx->pending) { (void)hipEventDestroy(p.a); (void)hipEventDestroy(p.b); }
  for (auto e : ctx->event_pool) (void)hipEventDestroy(e);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

const char* jues_last_error(const jues_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

size_t jues_device_bytes(const jues_ctx*) { return device_bytes_in_use(); }

jues_status jues_release(jues_ctx* ctx) {
  JUES_API_BEGIN(ctx)
  sync(ctx);
  ctx->cc.reset();
  release_scratch(ctx);
  ctx->gao.release();
  ctx->gao32.release();
  ctx->Bfac.release();
  ctx->Bfac32.release();
  ctx->n = ctx->naux = 0;
  JUES_API_END(ctx)
}

jues_status jues_set_ao_eri(jues_ctx* ctx, const double* gao, int n) {
  JUES_API_BEGIN(ctx)
  JUES_REQUIRE(gao && n > 0, "jues_set_ao_eri: bad arguments");
  sync(ctx);
  ctx->cc.reset();
  ctx->Bfac.release();
  ctx->Bfac32.release();
  ctx->gao32.release();
  const size_t nn = (size_t)n * n * n * n;
  ctx->gao.alloc(nn * sizeof(double));
  h2d(ctx, ctx->gao.p, gao, nn * sizeof(double));
  ctx->n = n;
  ctx->naux = 0;
  JUES_API_END(ctx)
}

jues_status jues_set_ao_factors(jues_ctx* ctx, const double* B, int n, int naux) {
  JUES_API_BEGIN(ctx)
  JUES_REQUIRE(B && n > 0 && naux > 0, "jues_set_ao_factors: bad arguments");
  sync(ctx);
  ctx->cc.reset();
  ctx->gao.release();
  ctx->gao32.release();
  ctx->Bfac32.release();
  const size_t nn = (size_t)n * n * naux;
  ctx->Bfac.alloc(nn * sizeof(double));
  h2d(ctx, ctx->Bfac.p, B, nn * sizeof(double));
  ctx->n = n;
  ctx->naux = naux;
  JUES_API_END(ctx)
}

// b[mu,nu,Q] = sum_P pqP[mu,nu,P] Jpqh[P,Q] (DF.jl:52-58) as one FP64 MFMA GEMM of
// (n^2 x naux) by (naux x naux); the result becomes the context's AO factors.
jues_status jues_set_df_integrals(jues_ctx* ctx, const double* pqP, const double* Jpqh, int n, int naux) {
  JUES_API_BEGIN(ctx)
  JUES_REQUIRE(pqP && Jpqh && n > 0 && naux > 0, "jues_set_df_integrals: bad arguments");
  sync(ctx);
  ctx->cc.reset();
  ctx->gao.release();
  ctx->gao32.release();
  ctx->Bfac32.release();
  ctx->Bfac.release();
  const long nn = (long)n * n;
  DeviceArray<double> three(ctx, pqP, (size_t)nn * naux), metric(ctx, Jpqh, (size_t)naux * naux);
  ctx->Bfac.alloc((size_t)nn * naux * sizeof(double));
  gemm<double>(ctx, false, false, nn, naux, naux, 1.0, three.p, nn, 0, metric.p, naux, 0, 0.0,
               ctx->Bfac.as<double>(), nn, 0, 1, "df_bmunu");
  sync(ctx);
  ctx->n = n;
  ctx->naux = naux;
  JUES_API_END(ctx)
}

jues_status jues_get_ao_factors(jues_ctx* ctx, double* B, int n, int naux) {
  JUES_API_BEGIN(ctx)
  JUES_REQUIRE(B, "jues_get_ao_factors: null pointer");
  JUES_REQUIRE(ctx->Bfac.p && n == ctx->n && naux == ctx->naux, "jues_get_ao_factors: no factors of that shape");
  d2h(ctx, B, ctx->Bfac.p, (size_t)n * n * naux * sizeof(double));
  JUES_API_END(ctx)
}

jues_status jues_tei_transform(jues_ctx* ctx, int prec, const double* C1, int d1, const double* C2, int d2,
                               const double* C3, int d3, const double* C4, int d4, int n, double* out) {
  JUES_API_BEGIN(ctx)
  JUES_REQUIRE(C1 && C2 && C3 && C4 && out, "jues_tei_transform: null pointer");
  JUES_REQUIRE(d1 > 0 && d2 > 0 && d3 > 0 && d4 > 0, "jues_tei_transform: bad dimensions");
  JUES_REQUIRE(n == ctx->n && n > 0, "jues_tei_transform: n does not match the AO integrals");
  JUES_REQUIRE(prec == JUES_FP64 || prec == JUES_FP32_MIXED, "jues_tei_transform: bad precision");
  const size_t total = (size_t)d1 * d2 * d3 * d4;
  if (prec == JUES_FP64) {
    DeviceArray<double> c1(ctx, C1, (size_t)n * d1), c2(ctx, C2, (size_t)n * d2), c3(ctx, C3, (size_t)n * d3),
        c4(ctx, C4, (size_t)n * d4);
    DevBuf o(total * sizeof(double));
    tei_transform_dev<double>(ctx, c1.p, d1, c2.p, d2, c3.p, d3, c4.p, d4, o.as<double>());
    d2h(ctx, out, o.p, total * sizeof(double));
  } else {
    DeviceArray<float> c1(ctx, C1, (size_t)n * d1), c2(ctx, C2, (size_t)n * d2), c3(ctx, C3, (size_t)n * d3),
        c4(ctx, C4, (size_t)n * d4);
    DevBuf o(total * sizeof(float)), o64(total * sizeof(double));
    tei_transform_dev<float>(ctx, c1.p, d1, c2.p, d2, c3.p, d3, c4.p, d4, o.as<float>());
    convert<float, double>(ctx, o.as<float>(), o64.as<double>(), (long)total);
    d2h(ctx, out, o64.p, total * sizeof(double));
  }
  JUES_API_END(ctx)
}

jues_status jues_get_eri(jues_ctx* ctx, int prec, const char* eri_string, const char* notation, const double* Cao,
                         int nalpha, const double* Cav, int nvira, const double* Cbo, int nbeta, const double* Cbv,
                         int nvirb, int n, int fcn, double* out) {
  JUES_API_BEGIN(ctx)
  JUES_REQUIRE(eri_string && std::strlen(eri_string) == 4,
               std::string("Invalid string given to get_eri: ") + (eri_string ? eri_string : "(null)"));
  const std::string note = notation ? notation : "phys";
  JUES_REQUIRE(note == "phys" || note == "chem", "get_eri: notation must be \"phys\" or \"chem\"");
  JUES_REQUIRE(out, "get_eri: null output");
  JUES_REQUIRE(n == ctx->n && n > 0, "get_eri: n does not match the AO integrals");
  JUES_REQUIRE(prec == JUES_FP64 || prec == JUES_FP32_MIXED, "get_eri: bad precision");
  JUES_REQUIRE(fcn >= 0, "get_eri: bad fcn");
  std::string s(eri_string);
  for (char c : s) JUES_REQUIRE(std::strchr("oOvV", c), "Invalid string given to get_eri: " + s);
  std::string chem = s;
  if (note == "phys") chem = std::string{s[0], s[2], s[1], s[3]};
  // IntegralTransformation.jl:52-68: O -> Cao (nalpha), o -> Cbo (nbeta), V -> Cav (nvira),
  // v -> Cbv (nvirb); fcn drops the first fcn columns of either occupied block
  const double* Cs[4];
  int ds[4];
  for (int k = 0; k < 4; ++k) {
    const char c = chem[k];
    const double* C = c == 'O' ? Cao : c == 'o' ? Cbo : c == 'V' ? Cav : Cbv;
    const int d = c == 'O' ? nalpha : c == 'o' ? nbeta : c == 'V' ? nvira : nvirb;
    const bool occ = c == 'o' || c == 'O';
    JUES_REQUIRE(C && d > 0, std::string("get_eri: no orbitals for '") + c + "'");
    JUES_REQUIRE(!occ || fcn < d, "get_eri: bad fcn");
    Cs[k] = occ ? C + (size_t)fcn * n : C;
    ds[k] = occ ? d - fcn : d;
  }
  const size_t total = (size_t)ds[0] * ds[1] * ds[2] * ds[3];
  DevBuf r64(total * sizeof(double));
  if (prec == JUES_FP64) {
    DeviceArray<double> c1(ctx, Cs[0], (size_t)n * ds[0]), c2(ctx, Cs[1], (size_t)n * ds[1]),
        c3(ctx, Cs[2], (size_t)n * ds[2]), c4(ctx, Cs[3], (size_t)n * ds[3]);
    DevBuf o4(total * sizeof(double));
    tei_transform_dev<double>(ctx, c1.p, ds[0], c2.p, ds[1], c3.p, ds[2], c4.p, ds[3], o4.as<double>());
    if (note == "phys")
      permute<double, double>(ctx, o4.as<double>(), {ds[0], ds[1], ds[2], ds[3]}, {0, 2, 1, 3}, r64.as<double>(),
                              1.0, 0.0);
    else
      JUES_HIP(hipMemcpyAsync(r64.p, o4.p, total * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
  } else {
    DeviceArray<float> c1(ctx, Cs[0], (size_t)n * ds[0]), c2(ctx, Cs[1], (size_t)n * ds[1]),
        c3(ctx, Cs[2], (size_t)n * ds[2]), c4(ctx, Cs[3], (size_t)n * ds[3]);
    DevBuf o4(total * sizeof(float));
    tei_transform_dev<float>(ctx, c1.p, ds[0], c2.p, ds[1], c3.p, ds[2], c4.p, ds[3], o4.as<float>());
    if (note == "phys")
      permute<float, double>(ctx, o4.as<float>(), {ds[0], ds[1], ds[2], ds[3]}, {0, 2, 1, 3}, r64.as<double>(),
                             1.0, 0.0);
    else
      convert<float, double>(ctx, o4.as<float>(), r64.as<double>(), (long)total);
  }
  d2h(ctx, out, r64.p, total * sizeof(double));
  JUES_API_END(ctx)
}

jues_status jues_rmp2(jues_ctx* ctx, int prec, const double* Cocc, int o, const double* Cvir, int v, int n,
                      const double* eps, double* E) {
  JUES_API_BEGIN(ctx)
  JUES_REQUIRE(Cocc && Cvir && eps && E && o > 0 && v > 0, "jues_rmp2: bad arguments");
  JUES_REQUIRE(n == ctx->n && n > 0, "jues_rmp2: n does not match the AO integrals");
  DeviceArray<double> e(ctx, eps, (size_t)o + v);
  const size_t total = (size_t)o * v * o * v;
  if (prec == JUES_FP64) {
    DeviceArray<double> co(ctx, Cocc, (size_t)n * o), cv(ctx, Cvir, (size_t)n * v);
    DevBuf K(total * sizeof(double));
    tei_transform_dev<double>(ctx, co.p, o, cv.p, v, co.p, o, cv.p, v, K.as<double>());
    *E = mp2_energy_dev<double>(ctx, K.as<double>(), e.p, o, v);
  } else {
    DeviceArray<float> co(ctx, Cocc, (size_t)n * o), cv(ctx, Cvir, (size_t)n * v);
    DevBuf K(total * sizeof(float));
    tei_transform_dev<float>(ctx, co.p, o, cv.p, v, co.p, o, cv.p, v, K.as<float>());
    *E = mp2_energy_dev<float>(ctx, K.as<float>(), e.p, o, v);
  }
  if (!std::isfinite(*E)) throw Error(JUES_NONFINITE, "jues_rmp2: non-finite energy (a zero denominator?)");
  JUES_API_END(ctx)
}

jues_status jues_gemm_test(jues_ctx* ctx, int prec, char transa, char transb, int M, int N, int K, double alpha,
                           const double* A, int lda, const double* B, int ldb, double beta, double* C, int ldc) {
  JUES_API_BEGIN(ctx)
  const bool ta = transa == 'T' || transa == 't', tb = transb == 'T' || transb == 't';
  JUES_REQUIRE(M >= 0 && N >= 0 && K >= 0 && A && B && C, "jues_gemm_test: bad arguments");
  const size_t na = (size_t)lda * (ta ? M : K), nb = (size_t)ldb * (tb ? K : N), nc = (size_t)ldc * N;
  if (prec == JUES_FP64) {
    DeviceArray<double> a(ctx, A, na), b(ctx, B, nb), c(ctx, C, nc);
    gemm<double>(ctx, ta, tb, M, N, K, alpha, a.p, lda, 0, b.p, ldb, 0, beta, c.p, ldc, 0, 1);
    d2h(ctx, C, c.p, nc * sizeof(double));
  } else if (prec == JUES_FP16X3_MIXED) {
    JUES_REQUIRE(K > 0 && M > 0 && N > 0, "jues_gemm_test: split-FP16 GEMM needs M, N, K > 0");
    DeviceArray<float> a(ctx, A, na), b(ctx, B, nb), c(ctx, C, nc);
    // the kernel reads k-contiguous planes: op(A)^T and op(B), transposed by the split where needed
    Split16Buf sa, sb;
    if (ta) split16<float>(ctx, a.p, K, M, lda, sa);
    else split16t<float>(ctx, a.p, K, M, lda, sa);
    if (!tb) split16<float>(ctx, b.p, K, N, ldb, sb);
    else split16t<float>(ctx, b.p, K, N, ldb, sb);
    gemm16_tn(ctx, M, N, (float)alpha, sa.s, sb.s, (float)beta, c.p, ldc);
    DevBuf c64(nc * sizeof(double));
    convert<float, double>(ctx, c.p, c64.as<double>(), (long)nc);
    d2h(ctx, C, c64.p, nc * sizeof(double));
  } else {
    DeviceArray<float> a(ctx, A, na), b(ctx, B, nb), c(ctx, C, nc);
    gemm<float>(ctx, ta, tb, M, N, K, (float)alpha, a.p, lda, 0, b.p, ldb, 0, (float)beta, c.p, ldc, 0, 1);
    DevBuf c64(nc * sizeof(double));
    convert<float, double>(ctx, c.p, c64.as<double>(), (long)nc);
    d2h(ctx, C, c64.p, nc * sizeof(double));
  }
  JUES_API_END(ctx)
}

jues_status jues_gemm_pair_test(jues_ctx* ctx, int prec, char transb, int M0, int N0, int K0, const double* A0,
                                int lda0, const double* B0, int ldb0, double* C0, int ldc0, int M1, int N1, int K1,
                                const double* A1, int lda1, const double* B1, int ldb1, double* C1, int ldc1,
                                double alpha, double beta) {
  JUES_API_BEGIN(ctx)
  const bool tb = transb == 'T' || transb == 't';
  JUES_REQUIRE(M0 >= 0 && N0 >= 0 && K0 >= 0 && M1 >= 0 && N1 >= 0 && K1 >= 0 && A0 && B0 && C0 && A1 && B1 && C1,
               "jues_gemm_pair_test: bad arguments");
  JUES_REQUIRE(prec == JUES_FP64 || prec == JUES_FP32_MIXED, "jues_gemm_pair_test: FP64 or FP32 only");
  const size_t na0 = (size_t)lda0 * K0, nb0 = (size_t)ldb0 * (tb ? K0 : N0), nc0 = (size_t)ldc0 * N0;
  const size_t na1 = (size_t)lda1 * K1, nb1 = (size_t)ldb1 * (tb ? K1 : N1), nc1 = (size_t)ldc1 * N1;
  auto run = [&](auto zero) {
    using T = decltype(zero);
    DeviceArray<T> a0(ctx, A0, na0), b0(ctx, B0, nb0), c0(ctx, C0, nc0);
    DeviceArray<T> a1(ctx, A1, na1), b1(ctx, B1, nb1), c1(ctx, C1, nc1);
    gemm_pair<T>(ctx, tb, M0, N0, K0, a0.p, lda0, b0.p, ldb0, c0.p, ldc0, M1, N1, K1, a1.p, lda1, b1.p, ldb1, c1.p,
                 ldc1, (T)alpha, (T)beta);
    for (int q = 0; q < 2; ++q) {
      const T* c = q ? c1.p : c0.p;
      const size_t n = q ? nc1 : nc0;
      DevBuf c64(std::max<size_t>(1, n) * sizeof(double));
      convert<T, double>(ctx, c, c64.as<double>(), (long)n);
      d2h(ctx, q ? C1 : C0, c64.p, n * sizeof(double));
    }
  };
  if (prec == JUES_FP64) run(0.0);
  else run(0.0f);
  JUES_API_END(ctx)
}

int jues_kernel_stats(jues_ctx* ctx, const char** names, double* ms, double* calls, double* flops, double* bytes,
                      int cap) {
  if (!ctx) return 0;
  try {
    (void)hipSetDevice(ctx->device);
    flush_timers(ctx);
  } catch (...) {
    return 0;
  }
  int k = 0;
  for (auto& kv : ctx->kstats) {
    if (k >= cap) break;
    if (names) names[k] = kv.first.c_str();
    if (ms) ms[k] = kv.second.ms;
    if (calls) calls[k] = kv.second.calls;
    if (flops) flops[k] = kv.second.flops;
    if (bytes) bytes[k] = kv.second.bytes;
    ++k;
  }
  return k;
}

void jues_kernel_stats_reset(jues_ctx* ctx, int enable) {
  if (!ctx) return;
  try {
    (void)hipSetDevice(ctx->device);
    flush_timers(ctx);
  } catch (...) {
  }
  ctx->kstats.clear();
  ctx->timing = enable != 0;
}

void jues_cc_default_opts(jues_cc_opts* o) {
  if (!o) return;
  o->maxit = 100;
  o->e_conv = 1e-10;
  o->r_conv = 1e-9;
  o->diis = 1;
  o->diis_max = 5;
  o->guess = JUES_GUESS_MP2;
  o->ccsd_eqs = JUES_CCSD_MRCCSD;
  o->fcn = 0;
  o->fp32_r_conv = 1e-5;
  o->refine_maxit = 60;
  o->verbose = 0;
  o->coupling = 1.0;
  o->continuation = 1;
  o->ladder = JUES_LADDER_AUTO;
  o->refine = JUES_REFINE_RESTART;
}

}  // extern "C"
