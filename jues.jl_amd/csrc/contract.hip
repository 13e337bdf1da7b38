// Two-operand tensor contraction on the device by transpose-transpose-GEMM-transpose.
//
// C[idxC] = alpha * sum_K A[idxA] B[idxB] + beta * C.  Letters present in A and
// B but not in C are contracted; every letter of C comes from exactly one of A, B.
// Of the 16 ways to order the free and contracted groups (and to produce C or
// C^T from the GEMM) the one that permutes the fewest bytes is chosen, so
// operands already laid out as a matrix go straight to the MFMA GEMM.
#include "jues_internal.h"

#include <algorithm>
#include <limits>

namespace jues {
namespace {

struct Plan {
  std::string fa, fb, kk;   // free-A order, free-B order, contracted order
  bool swapC;               // GEMM produces C^T (rows = free-B)
  long cost;
};

long dim_of(char c, const std::string& idx, const std::vector<long>& dims) {
  const size_t p = idx.find(c);
  return p == std::string::npos ? -1 : dims[p];
}

std::vector<int> perm_to(const std::string& from, const std::string& to) {
  // out dim k = in dim perm[k] with out order `to`, in order `from`
  std::vector<int> p(to.size());
  for (size_t k = 0; k < to.size(); ++k) p[k] = (int)from.find(to[k]);
  return p;
}

}  // namespace

template <typename Ti, typename To>
void add_perm(jues_ctx* ctx, double alpha, const Tens<Ti>& A, double beta, const Tens<To>& C, const char* tag) {
  JUES_REQUIRE(A.idx.size() == C.idx.size() && A.idx.size() == A.dims.size() && C.idx.size() == C.dims.size(),
               "add_perm: rank mismatch");
  for (size_t k = 0; k < C.idx.size(); ++k)
    JUES_REQUIRE(dim_of(C.idx[k], A.idx, A.dims) == C.dims[k], "add_perm: index mismatch " + A.idx + "->" + C.idx);
  permute<Ti, To>(ctx, A.p, A.dims, perm_to(A.idx, C.idx), C.p, alpha, beta, tag);
}

template <typename T>
void contract(jues_ctx* ctx, double alpha, const Tens<T>& A, const Tens<T>& B, double beta, const Tens<T>& C,
              const char* tag, T* work, long work_elems) {

  JUES_REQUIRE(A.idx.size() == A.dims.size() && B.idx.size() == B.dims.size() && C.idx.size() == C.dims.size(),
               "contract: rank mismatch");
  std::string fa_A, fb_B, kA, kB;  // groups in operand order
  for (char c : A.idx) {
    const bool inB = B.idx.find(c) != std::string::npos, inC = C.idx.find(c) != std::string::npos;
    JUES_REQUIRE(inB != inC, std::string("contract: unsupported index ") + c);
    if (inB) {
      JUES_REQUIRE(dim_of(c, B.idx, B.dims) == dim_of(c, A.idx, A.dims), std::string("contract: dim mismatch ") + c);
      kA += c;
    } else {
      JUES_REQUIRE(dim_of(c, C.idx, C.dims) == dim_of(c, A.idx, A.dims), std::string("contract: dim mismatch ") + c);
      fa_A += c;
    }
  }
  for (char c : B.idx) {
    const bool inA = A.idx.find(c) != std::string::npos, inC = C.idx.find(c) != std::string::npos;
    JUES_REQUIRE(inA != inC, std::string("contract: unsupported index ") + c);
    if (inA) kB += c;
    else {
      JUES_REQUIRE(dim_of(c, C.idx, C.dims) == dim_of(c, B.idx, B.dims), std::string("contract: dim mismatch ") + c);
      fb_B += c;
    }
  }
  JUES_REQUIRE(fa_A.size() + fb_B.size() == C.idx.size(), "contract: output indices");
  std::string fa_C, fb_C;
  for (char c : C.idx) (A.idx.find(c) != std::string::npos ? fa_C : fb_C) += c;

  long M = 1, N = 1, K = 1;
  for (char c : fa_A) M *= dim_of(c, A.idx, A.dims);
  for (char c : fb_B) N *= dim_of(c, B.idx, B.dims);
  for (char c : kA) K *= dim_of(c, A.idx, A.dims);
  if (M == 0 || N == 0) return;

  // choose the plan permuting the fewest elements
  Plan best{};
  best.cost = std::numeric_limits<long>::max();
  for (int ik = 0; ik < 2; ++ik)
    for (int ia = 0; ia < 2; ++ia)
      for (int ib = 0; ib < 2; ++ib)
        for (int sc = 0; sc < 2; ++sc) {
          Plan p{ia ? fa_C : fa_A, ib ? fb_C : fb_B, ik ? kB : kA, sc == 1, 0};
          const bool a_ok = A.idx == p.fa + p.kk || A.idx == p.kk + p.fa;
          const bool b_ok = B.idx == p.kk + p.fb || B.idx == p.fb + p.kk;
          const bool c_ok = sc ? C.idx == p.fb + p.fa : C.idx == p.fa + p.fb;
          p.cost = (a_ok ? 0 : A.size()) + (b_ok ? 0 : B.size()) + (c_ok ? 0 : 2 * C.size());
          if (p.cost < best.cost) best = p;
        }
  const Plan& p = best;
  const std::string direct = p.swapC ? p.fb + p.fa : p.fa + p.fb;
  const bool a_perm = !(A.idx == p.fa + p.kk || A.idx == p.kk + p.fa);
  const bool b_perm = !(B.idx == p.kk + p.fb || B.idx == p.fb + p.kk);
  const bool c_direct = C.idx == direct;
  // the largest permuted copy (operand or product) goes to the caller's workspace when it
  // fits there, the others to the context's scratch slots
  const long sz[3] = {a_perm ? A.size() : 0, b_perm ? B.size() : 0, c_direct ? 0 : C.size()};
  int to_work = -1;
  for (int q = 0; q < 3; ++q)
    if (work && sz[q] > 0 && sz[q] <= work_elems && (to_work < 0 || sz[q] > sz[to_work])) to_work = q;
  auto temp = [&](int q, int slot) -> T* {
    if (q == to_work) return work;
    return static_cast<T*>(scratch(ctx, slot, sizeof(T) * (size_t)sz[q]));
  };

  // operand A as op(A) (M x K)
  const T* Ap = A.p;
  bool ta = false;
  if (A.idx == p.fa + p.kk) ta = false;
  else if (A.idx == p.kk + p.fa) ta = true;
  else {
    T* tmp = temp(0, 10);
    permute<T, T>(ctx, A.p, A.dims, perm_to(A.idx, p.fa + p.kk), tmp, 1.0, 0.0, "contract_perm");
    Ap = tmp;
    ta = false;
  }
  const T* Bp = B.p;
  bool tb = false;
  if (B.idx == p.kk + p.fb) tb = false;
  else if (B.idx == p.fb + p.kk) tb = true;
  else {
    T* tmp = temp(1, 11);
    permute<T, T>(ctx, B.p, B.dims, perm_to(B.idx, p.kk + p.fb), tmp, 1.0, 0.0, "contract_perm");
    Bp = tmp;
    tb = false;
  }
  const long lda = ta ? K : M, ldb = tb ? N : K;
  T* Cp = C.p;
  if (!c_direct) Cp = temp(2, 12);
  const T al = c_direct ? (T)alpha : T(1), be = c_direct ? (T)beta : T(0);
  if (!p.swapC) {
    gemm<T>(ctx, ta, tb, M, N, K, al, Ap, lda, 0, Bp, ldb, 0, be, Cp, M, 0, 1, tag);
  } else {
    // C^T (N x M) = op(B)^T op(A)^T
    gemm<T>(ctx, !tb, !ta, N, M, K, al, Bp, ldb, 0, Ap, lda, 0, be, Cp, N, 0, 1, tag);
  }
  if (!c_direct) {
    Tens<T> tmpT;
    tmpT.p = Cp;
    tmpT.idx = direct;
    for (char c : direct) tmpT.dims.push_back(dim_of(c, C.idx, C.dims));
    add_perm<T, T>(ctx, alpha, tmpT, beta, C, "contract_perm");
  }
}

template void contract<double>(jues_ctx*, double, const Tens<double>&, const Tens<double>&, double,
                               const Tens<double>&, const char*, double*, long);
template void contract<float>(jues_ctx*, double, const Tens<float>&, const Tens<float>&, double,
                              const Tens<float>&, const char*, float*, long);
template void add_perm<double, double>(jues_ctx*, double, const Tens<double>&, double, const Tens<double>&,
                                       const char*);
template void add_perm<float, float>(jues_ctx*, double, const Tens<float>&, double, const Tens<float>&,
                                     const char*);
template void add_perm<double, float>(jues_ctx*, double, const Tens<double>&, double, const Tens<float>&,
                                      const char*);
template void add_perm<float, double>(jues_ctx*, double, const Tens<float>&, double, const Tens<double>&,
                                      const char*);

}  // namespace jues
