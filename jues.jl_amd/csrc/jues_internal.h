// Internal declarations shared by the libjues_hip translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/jues.h"

namespace jues {

struct Error : std::exception {
  jues_status st;
  std::string msg;
  Error(jues_status s, std::string m) : st(s), msg(std::move(m)) {}
  const char* what() const noexcept override { return msg.c_str(); }
};

#define JUES_HIP(x)                                                                    \
  do {                                                                                 \
    hipError_t e_ = (x);                                                               \
    if (e_ != hipSuccess)                                                              \
      throw ::jues::Error(JUES_HIP_ERROR, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define JUES_REQUIRE(cond, msg)                                          \
  do {                                                                   \
    if (!(cond)) throw ::jues::Error(JUES_BAD_ARG, std::string(msg));    \
  } while (0)

// Owning device allocation.  Allocation failure raises JUES_OOM.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  explicit DevBuf(size_t nbytes) { alloc(nbytes); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) { release(); p = o.p; bytes = o.bytes; o.p = nullptr; o.bytes = 0; }
    return *this;
  }
  ~DevBuf() { release(); }
  void alloc(size_t nbytes);
  void release();
  template <typename T> T* as() const { return static_cast<T*>(p); }
};

size_t device_bytes_in_use();

// Allocation-free scope: while one is active on the calling thread, DevBuf::alloc and
// scratch() raise instead of allocating, and gemm() avoids its workspace paths (split-K,
// the transposed copy of a short A).  Used by the entry points that must not allocate.
struct NoAllocScope {
  bool prev;
  NoAllocScope();
  ~NoAllocScope();
};
bool no_alloc_active();
// looks up the device address of the allocation-free entry point's static scratch array on
// `device` (current) once; called by jues_create (noalloc.hip)
void noalloc_attach(int device);

struct KStat {
  double ms = 0, calls = 0, flops = 0, bytes = 0;
};

struct CCState;  // cc.hip
struct CCDeleter {
  void operator()(CCState* p) const;  // defined in cc.hip
};

}  // namespace jues

struct jues_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  // AO integrals: exactly one of gao (n^4) / Bfac (n*n*naux) is set.
  int n = 0, naux = 0;
  jues::DevBuf gao;
  jues::DevBuf Bfac;
  jues::DevBuf gao32, Bfac32;  // FP32 copies, made on first FP32 use
  // kernel timing (HIP events on `stream`)
  bool timing = false;
  std::map<std::string, jues::KStat> kstats;
  struct Pending { std::string name; hipEvent_t a, b; double flops, bytes; };
  std::vector<Pending> pending;
  std::vector<hipEvent_t> event_pool;
  std::unique_ptr<jues::CCState, jues::CCDeleter> cc;
  jues_ctx();
  ~jues_ctx();
};

namespace jues {

// Scoped kernel timer: records events around the enclosed launches when timing is on.
struct KTimer {
  jues_ctx* ctx;
  bool on;
  size_t slot = 0;
  KTimer(jues_ctx* c, const char* name, double flops, double bytes);
  ~KTimer();
};
void flush_timers(jues_ctx* ctx);  // synchronises and folds pending events into kstats

// ---- GEMM (gemm.hip) -------------------------------------------------------
// Column-major C[M,N] = alpha*op(A)*op(B) + beta*C, batched with element strides.
template <typename T>
void gemm(jues_ctx* ctx, bool ta, bool tb, long M, long N, long K, T alpha, const T* A, long lda,
          long sA, const T* B, long ldb, long sB, T beta, T* C, long ldc, long sC, int batch = 1,
          const char* tag = nullptr);

// C = A B in FP32 MFMA work with the FP64 result written directly (NN, beta = 0, one product), on
// the high-occupancy FP32 kernel; returns false, launching nothing, where that kernel does not apply.
bool gemm32_out64(jues_ctx* ctx, long M, long N, long K, const float* A, long lda, const float* B, long ldb,
                  double* C, long ldc, const char* tag = nullptr);

// C = alpha sum_z op(A_z) op(B_z) + beta C over a batch (strides sA, sB): the batch's products are
// written to work (batch * M * N elements) and added in a fixed order, so the result is the same
// on every call.
template <typename T>
void gemm_batch_sum(jues_ctx* ctx, bool ta, bool tb, long M, long N, long K, T alpha, const T* A, long lda, long sA,
                    const T* B, long ldb, long sB, T beta, T* C, long ldc, int batch, T* work,
                    const char* tag = nullptr);

// Two independent products C_i = alpha A_i op(B_i) + beta C_i (A_i m-contiguous, op the same
// for both) in one launch, so that the second fills the first one's last partial round of
// workgroups; falls back to two gemm calls where the high-occupancy kernels do not apply.
template <typename T>
void gemm_pair(jues_ctx* ctx, bool tb, long M0, long N0, long K0, const T* A0, long lda0, const T* B0, long ldb0,
               T* C0, long ldc0, long M1, long N1, long K1, const T* A1, long lda1, const T* B1, long ldb1, T* C1,
               long ldc1, T alpha, T beta, const char* tag = nullptr);

// C = alpha (A0 op(B0) + A1 op(B1)) + beta C: one product over two contraction segments that
// share the leading dimensions and batch strides, so C is read and written once.  One launch
// when K0 and K1 are multiples of 16 and the high-occupancy kernels apply, else two GEMMs.
template <typename T>
void gemm_ksum(jues_ctx* ctx, bool tb, long M, long N, long K0, const T* A0, const T* B0, long K1, const T* A1,
               const T* B1, long lda, long ldb, long sA, long sB, T alpha, T beta, T* C, long ldc, long sC,
               int batch, const char* tag = nullptr);

// ---- split-FP16 MFMA GEMM (gemm16.hip) ---------------------------------------
// An FP32 matrix X (rows x cols) as two FP16 planes hi + lo = s X, column-major and
// zero-padded to rows_p (multiple of 32) x cols_p (>= cols + 256), interleaved per
// 32-row block: column c, block b holds 32 hi then 32 lo halves at p + c*2*rows_p + 64b.
// `sexp` is a device int holding k, s = 2^k.
struct Split16 {
  _Float16* p = nullptr;
  int* sexp = nullptr;
  long rows = 0, cols = 0, rows_p = 0, cols_p = 0;
};
struct Split16Buf {
  DevBuf mem;
  Split16 s;
  void ensure(long rows, long cols);  // (re)allocates when the padded size grows
};
// columns [c0, c0 + nc) of a split operand (its planes, padding beyond them included)
inline Split16 split16_cols(const Split16& s, long c0, long nc) {
  Split16 r = s;
  r.p = s.p + c0 * 2 * s.rows_p;
  r.cols = nc;
  r.cols_p = s.cols_p - c0;
  return r;
}
// planes of X; the scale is found on the device (no host synchronisation), or taken from
// sexp_of when X is known to hold the same values as the operand that scale belongs to
template <typename Ti>
void split16(jues_ctx* ctx, const Ti* X, long rows, long cols, long ldx, Split16Buf& out,
             const char* tag = nullptr, const int* sexp_of = nullptr);
// planes (rows x cols) of X^T, X stored cols x rows with leading dimension ldx
template <typename Ti>
void split16t(jues_ctx* ctx, const Ti* X, long rows, long cols, long ldx, Split16Buf& out,
              const char* tag = nullptr);
// C[m,n] = alpha sum_k A(k,m) B(k,n) + beta C[m,n], K = A.rows = B.rows, FP32 C
void gemm16_tn(jues_ctx* ctx, long M, long N, float alpha, const Split16& A, const Split16& B, float beta,
               float* C, long ldc, const char* tag = nullptr);

// ---- elementwise / data movement (kernels.hip) -------------------------------
// out = alpha * permutedims(in, perm) + beta * out   (perm 0-based; out dim k = in dim perm[k])
template <typename Ti, typename To>
void permute(jues_ctx* ctx, const Ti* in, const std::vector<long>& in_dims,
             const std::vector<int>& perm, To* out, double alpha, double beta,
             const char* tag = nullptr);
template <typename Ti, typename To>
void convert(jues_ctx* ctx, const Ti* in, To* out, long n);
template <typename T>
void fill(jues_ctx* ctx, T* out, long n, double value);
// y = a*x + b*y
template <typename T>
void axpby(jues_ctx* ctx, long n, double a, const T* x, double b, T* y);
// deterministic FP64 dot of two vectors (any element types)
template <typename Ta, typename Tb>
double dot(jues_ctx* ctx, const Ta* x, const Tb* y, long n);
double max_abs(jues_ctx* ctx, const double* x, long n);

// ---- contraction engine (einsum.cpp) -----------------------------------------
template <typename T>
struct Tens {
  T* p = nullptr;
  std::string idx;         // one letter per dimension, first letter fastest
  std::vector<long> dims;
  long size() const { long s = 1; for (long d : dims) s *= d; return s; }
};
// C[idxC] = alpha * sum A[idxA] B[idxB] + beta * C  (any index order; TTGT on the device).
// A permuted copy of an operand or of the product goes to `work` (work_elems elements) when
// given and large enough, else to the context's scratch.
template <typename T>
void contract(jues_ctx* ctx, double alpha, const Tens<T>& A, const Tens<T>& B, double beta,
              const Tens<T>& C, const char* tag = nullptr, T* work = nullptr, long work_elems = 0);
// C[idxC] = alpha * P(A) + beta * C
template <typename Ti, typename To>
void add_perm(jues_ctx* ctx, double alpha, const Tens<Ti>& A, double beta, const Tens<To>& C,
              const char* tag = nullptr);

// scratch pool owned by the context for temporaries of the contraction engine
void* scratch(jues_ctx* ctx, int slot, size_t bytes);

void release_scratch(jues_ctx* ctx);

// ---- integral transformation (transform.hip) -------------------------------------
// AO integrals in precision T (FP32 copies are made on first use and cached).
template <typename T> const T* ao_dense(jues_ctx* ctx);
template <typename T> const T* ao_factors(jues_ctx* ctx);
// out[p,q,r,s] (chemist, d1 x d2 x d3 x d4) from device C_k (n x d_k)
template <typename T>
void tei_transform_dev(jues_ctx* ctx, const T* C1, int d1, const T* C2, int d2, const T* C3, int d3,
                       const T* C4, int d4, T* out);
// MO factors L[p,q,P] = (C1^T B_P C2)[p,q]  (needs factors)
template <typename T>
void mo_factors_dev(jues_ctx* ctx, const T* C1, int d1, const T* C2, int d2, T* L);
// the same with the half-transformed factors in `work` (n * min(d1, d2) * naux elements)
template <typename T>
void mo_factors_dev(jues_ctx* ctx, const T* C1, int d1, const T* C2, int d2, T* L, T* work);
// RMP2 energy from K[i,a,j,b] = (ia|jb) (device, o x v x o x v) and eps (device FP64, n)
template <typename T>
double mp2_energy_dev(jues_ctx* ctx, const T* K, const double* eps, int o, int v);
// One UMP2 spin block from K[i,a,J,B] = (ia|JB) (oA vA x oB vB) and the two spins' orbital
// energies (occupied first); same = alpha-alpha or beta-beta (K square, eA == eB).
template <typename T>
double ump2_block_dev(jues_ctx* ctx, const T* K, const double* eA, const double* eB, int oA, int vA, int oB, int vB,
                      bool same);

// ---- host <-> device helpers ---------------------------------------------------
void h2d(jues_ctx* ctx, void* dst, const void* src, size_t bytes);
void d2h(jues_ctx* ctx, void* dst, const void* src, size_t bytes);
void sync(jues_ctx* ctx);

}  // namespace jues
