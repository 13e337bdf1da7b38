// Split-FP16 MFMA GEMM for gfx950 ("fp16x3"): FP32-class products on the FP16
// matrix cores.
//
// An FP32 operand X is stored as two FP16 planes, hi = fp16(s X) and
// lo = fp16(s X - hi), with one power-of-two scale s per operand chosen so that
// max |s X| < 2^15 (kept as its integer exponent and applied with ldexp, so no
// scale or descale factor can overflow).  hi + lo carries 22 significant bits, and
//     a b ~ (a_hi b_hi + a_hi b_lo + a_lo b_hi) / (s_a s_b)
// drops only the lo*lo term (relative 2^-22).  The three products run on
// v_mfma_f32_16x16x32_f16 with FP32 accumulation.  A split GEMM costs three FP16
// MFMA GEMMs: at the 2.5 PFLOP/s dense FP16 peak that is 833 TFLOP/s of
// FP32-equivalent work, 5.3x the 157.3 TFLOP/s FP32 MFMA peak.
//
// Planes are column-major and k-contiguous, stored interleaved: for column c and
// k-block b (32 consecutive k), 32 hi halves then 32 lo halves, so that one
// k-slice of one column is a single 128-byte line.  They are padded with zeros to
// a multiple of 32 rows and 256 columns, so the GEMM loads no partial tiles and
// needs no bounds logic except on the output.  The GEMM is the "TN" form
// C[m,n] = alpha sum_k A(k,m) B(k,n) + beta C[m,n]; a symmetric left operand
// (every amplitude and integral matrix of the ring terms) is its own transpose,
// so C = A B needs no data movement.
//
// Tile: 256x128 per 512-thread workgroup, 8 wavefronts of 64x64 (4x4 MFMA blocks
// of 16x16), k-slices of 32 staged through LDS (double-buffered, 48 KB per
// buffer) with two further slices in flight in registers.  LDS rows are 128
// bytes (hi then lo) with 16-byte chunks XOR-swizzled by row, so both the
// global-to-LDS stores and the fragment reads are bank-conflict free.  Among
// 128x128/4-wave, 256x128/4-wave, 128x256, 256x256, 32x32x16 MFMA, LDS-DMA with
// three LDS stages and split (non-interleaved) planes, this was fastest on the
// ring (20400^3) and ladder (1830 x 57970 x 57970) products.
#include "jues_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace jues {
namespace {

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef float fvec4 __attribute__((ext_vector_type(4)));
typedef unsigned int u4 __attribute__((ext_vector_type(4)));

// Workgroup tile BM x BN over WM x WN wavefronts (each (BM/WM) x (BN/WN)), k-slices of 32.
constexpr int BM = 256, BN = 128, WM = 4, WN = 2, NT = 64 * WM * WN, BK = 32, NXCD = 8, GROUP_M = 8;
constexpr int PAD_C = 256;  // column padding of every plane (>= BM, BN)
constexpr int kBufferDword3 = 0x00020000;  // gfx9-family raw buffer descriptor word 3

__device__ inline __amdgpu_buffer_rsrc_t make_rsrc(const void* base) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, 0x7fffffff, kBufferDword3);
}

// byte offset of 16-byte chunk q (0-3 hi, 4-7 lo; k = 8(q%4)..+7) of LDS row r in a
// slice of 128-byte rows
__device__ inline int lds_off(int r, int q) { return r * 128 + ((q ^ (r & 7)) << 4); }

__device__ inline void tile_coords(int pid, int nwg, int tiles_m, int tiles_n, int& tm, int& tn) {
  // XCD-aware bijective remap, then grouped ordering (as the FP32/FP64 GEMM)
  const int per = (nwg + NXCD - 1) / NXCD;
  int tall = nwg % NXCD;
  if (tall == 0) tall = NXCD;
  const int xcd = pid % NXCD, loc = pid / NXCD;
  const int p2 = xcd < tall ? xcd * per + loc : tall * per + (xcd - tall) * (per - 1) + loc;
  const int group_size = GROUP_M * tiles_n;
  const int gid = p2 / group_size;
  const int first_m = gid * GROUP_M;
  const int gsz = min(tiles_m - first_m, GROUP_M);
  const int r = p2 - gid * group_size;
  tm = first_m + r % gsz;
  tn = r / gsz;
}

// KIND separates kernel symbols for profiles (0 generic, 1 CC ring, 2 CC pp-ladder, 3 CC hole-hole).
// blockIdx.z selects a chunk of kchunk rows of K (a multiple of BK; the last chunk may be
// shorter) whose product goes to C + z sC: a split-K launch writes partial products.
template <int KIND>
__global__ __launch_bounds__(NT, 1) void gemm16_kernel(int M, int N, int Kp, float alpha,
                                                       const _Float16* __restrict__ A, long lda2,
                                                       const int* __restrict__ ea,
                                                       const _Float16* __restrict__ B, long ldb2,
                                                       const int* __restrict__ eb, float beta,
                                                       float* __restrict__ C, long ldc, int tiles_m, int tiles_n,
                                                       int kchunk, long sC) {
  constexpr int PA = BM * 128, PB = BN * 128;         // bytes of one slice of each operand
  constexpr int CA = BM * 8 / NT, CB = BN * 8 / NT;   // 16-byte chunks per thread per slice
  constexpr int IM = BM / WM / 16, JN = BN / WN / 16;
  static_assert(CA * NT == BM * 8 && CB * NT == BN * 8, "tile does not divide into chunks");
  __shared__ __attribute__((aligned(16))) char lds[2][PA + PB];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave % WM, wn = wave / WM;
  int tm, tn;
  tile_coords(blockIdx.x, gridDim.x, tiles_m, tiles_n, tm, tn);
  const int m0 = tm * BM, n0 = tn * BN;

  // chunk q = tid + NT p of a slice is tile column q/8, chunk q%8 of its 128-byte line
  int voffA[CA], loffA[CA], voffB[CB], loffB[CB];
#pragma unroll
  for (int p = 0; p < CA; ++p) {
    const int q = tid + NT * p;
    voffA[p] = (int)(2 * ((q & 7) * 8 + (long)(q >> 3) * lda2));
    loffA[p] = lds_off(q >> 3, q & 7);
  }
#pragma unroll
  for (int p = 0; p < CB; ++p) {
    const int q = tid + NT * p;
    voffB[p] = (int)(2 * ((q & 7) * 8 + (long)(q >> 3) * ldb2));
    loffB[p] = PA + lds_off(q >> 3, q & 7);
  }
  // this chunk's k-range starts 2 kz0 halves into every column (k-blocks of 32 hi + 32 lo)
  const int kz0 = (int)blockIdx.z * kchunk;
  const _Float16* aT = A + (long)m0 * lda2 + 2L * kz0;
  const _Float16* bT = B + (long)n0 * ldb2 + 2L * kz0;
  C += (long)blockIdx.z * sC;

  // Two register slots hold the next two k-slices in flight; the loop is unrolled by
  // two so that slot indices are compile-time constants.  Loads past the last slice
  // re-read the last slice (in bounds, never used), which keeps the body free of
  // branches so the scheduler can interleave loads, LDS traffic and MFMAs.  (Macros,
  // not lambdas: lambdas over the register arrays cost spills.)
  const int nk = min(kchunk, Kp - kz0) / BK;
  u4 sa0[CA], sb0[CB], sa1[CA], sb1[CB];
#define GLOAD(SA, SB, KT)                                                                                 \
  {                                                                                                       \
    const int k0_ = min((KT), nk - 1) * 2 * BK;                                                           \
    const auto rA_ = make_rsrc(aT + k0_), rB_ = make_rsrc(bT + k0_);                                      \
    _Pragma("unroll") for (int p = 0; p < CA; ++p)                                                        \
        SA[p] = __builtin_bit_cast(u4, __builtin_amdgcn_raw_buffer_load_b128(rA_, voffA[p], 0, 0));      \
    _Pragma("unroll") for (int p = 0; p < CB; ++p)                                                        \
        SB[p] = __builtin_bit_cast(u4, __builtin_amdgcn_raw_buffer_load_b128(rB_, voffB[p], 0, 0));      \
  }
#define LSTORE(BUF, SA, SB)                                                                               \
  {                                                                                                       \
    _Pragma("unroll") for (int p = 0; p < CA; ++p) *reinterpret_cast<u4*>(&lds[BUF][loffA[p]]) = SA[p];  \
    _Pragma("unroll") for (int p = 0; p < CB; ++p) *reinterpret_cast<u4*>(&lds[BUF][loffB[p]]) = SB[p];  \
  }

  fvec4 acc[IM][JN];
#pragma unroll
  for (int i = 0; i < IM; ++i)
#pragma unroll
    for (int j = 0; j < JN; ++j) acc[i][j] = fvec4{0, 0, 0, 0};

  // fragment rows: lane%16 within a 16-row block, k-chunk lane/16 (hi), lane/16 + 4 (lo)
  const int fr = lane & 15, fc = lane >> 4;
  int aoff[IM], boff[JN];
#pragma unroll
  for (int i = 0; i < IM; ++i) aoff[i] = lds_off(wm * (BM / WM) + i * 16 + fr, fc);
#pragma unroll
  for (int j = 0; j < JN; ++j) boff[j] = PA + lds_off(wn * (BN / WN) + j * 16 + fr, fc);
  // the lo chunk of a fragment is its hi chunk's slot with bit 2 flipped: offset ^ 64
#define COMPUTE(BUF)                                                                                      \
  {                                                                                                       \
    h8 ah[IM], al[IM], bh[JN], bl[JN];                                                                    \
    _Pragma("unroll") for (int i = 0; i < IM; ++i) ah[i] = *reinterpret_cast<const h8*>(&lds[BUF][aoff[i]]); \
    _Pragma("unroll") for (int j = 0; j < JN; ++j) bh[j] = *reinterpret_cast<const h8*>(&lds[BUF][boff[j]]); \
    _Pragma("unroll") for (int i = 0; i < IM; ++i) al[i] = *reinterpret_cast<const h8*>(&lds[BUF][aoff[i] ^ 64]); \
    _Pragma("unroll") for (int j = 0; j < JN; ++j) bl[j] = *reinterpret_cast<const h8*>(&lds[BUF][boff[j] ^ 64]); \
    /* operands swapped (B^T A^T) so that a lane's accumulator column is an m index */                  \
    _Pragma("unroll") for (int i = 0; i < IM; ++i) _Pragma("unroll") for (int j = 0; j < JN; ++j)        \
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(bh[j], ah[i], acc[i][j], 0, 0, 0);             \
    _Pragma("unroll") for (int i = 0; i < IM; ++i) _Pragma("unroll") for (int j = 0; j < JN; ++j) {      \
      acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(bl[j], ah[i], acc[i][j], 0, 0, 0);               \
      acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(bh[j], al[i], acc[i][j], 0, 0, 0);               \
    }                                                                                                     \
  }
  // Schedule of one half-iteration: all fragment reads first, then the MFMAs with the
  // next slice's global loads and LDS stores spread between them.
#define SCHEDULE()                                                                                        \
  {                                                                                                       \
    constexpr int NR = 2 * (IM + JN), NM = 3 * IM * JN, NS = CA + CB, per = NM / (2 * NS);               \
    __builtin_amdgcn_sched_group_barrier(0x100, NR, 0); /* DS reads */                                   \
    _Pragma("unroll") for (int r = 0; r < NS; ++r) {                                                      \
      __builtin_amdgcn_sched_group_barrier(0x008, per, 0); /* MFMA */                                    \
      __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);   /* VMEM read */                               \
    }                                                                                                     \
    _Pragma("unroll") for (int r = 0; r < NS; ++r) {                                                      \
      __builtin_amdgcn_sched_group_barrier(0x008, per, 0);                                               \
      __builtin_amdgcn_sched_group_barrier(0x200, 1, 0); /* DS write */                                  \
    }                                                                                                     \
    __builtin_amdgcn_sched_group_barrier(0x008, NM - per * 2 * NS, 0);                                   \
  }

  // prologue: slice 0 -> LDS buffer 0, slice 1 -> slot 1, slice 2 -> slot 0
  GLOAD(sa0, sb0, 0);
  LSTORE(0, sa0, sb0);
  GLOAD(sa1, sb1, 1);
  GLOAD(sa0, sb0, 2);
  __syncthreads();
  for (int kt = 0; kt < nk; kt += 2) {
    // buffer 0 holds slice kt; slot 1 holds kt+1, slot 0 holds kt+2
    COMPUTE(0);
    LSTORE(1, sa1, sb1);
    GLOAD(sa1, sb1, kt + 3);
    SCHEDULE();
    __syncthreads();
    if (kt + 1 >= nk) break;
    COMPUTE(1);
    LSTORE(0, sa0, sb0);
    GLOAD(sa0, sb0, kt + 4);
    SCHEDULE();
    __syncthreads();
  }

  const int descale = -(ea[0] + eb[0]);
  // acc[i][j][r] is C(m, n) with m = lane%16 of block i, n = 4(lane/16) + r of block j
  if (beta != 0.0f && m0 + BM <= M && n0 + BN <= N) {
    // interior tile accumulating into C: all of the lane's C loads first, then the updates
    // (one load -> store round trip per element would serialise the epilogue)
    float cv[IM][JN][4];
#pragma unroll
    for (int i = 0; i < IM; ++i)
#pragma unroll
      for (int j = 0; j < JN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r)
          cv[i][j][r] = C[(long)(m0 + wm * (BM / WM) + i * 16 + fr) +
                          (long)(n0 + wn * (BN / WN) + j * 16 + 4 * fc + r) * ldc];
#pragma unroll
    for (int i = 0; i < IM; ++i)
#pragma unroll
      for (int j = 0; j < JN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r)
          C[(long)(m0 + wm * (BM / WM) + i * 16 + fr) + (long)(n0 + wn * (BN / WN) + j * 16 + 4 * fc + r) * ldc] =
              alpha * ldexpf(acc[i][j][r], descale) + beta * cv[i][j][r];
    return;
  }
#pragma unroll
  for (int i = 0; i < IM; ++i) {
    const int m = m0 + wm * (BM / WM) + i * 16 + fr;
    if (m >= M) continue;
#pragma unroll
    for (int j = 0; j < JN; ++j) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int n = n0 + wn * (BN / WN) + j * 16 + 4 * fc + r;
        if (n < N) {
          const long off = (long)m + (long)n * ldc;
          float v = alpha * ldexpf(acc[i][j][r], descale);
          if (beta != 0.0f) v += beta * C[off];
          C[off] = v;
        }
      }
    }
  }
}

#undef GLOAD
#undef LSTORE
#undef COMPUTE
#undef SCHEDULE

// C = beta C + sum_s W[s] (fixed order), W[s] is M x N with leading dimension M
__global__ __launch_bounds__(256) void splitk16_reduce_kernel(const float* __restrict__ W, int S, long M, long N,
                                                              float* __restrict__ C, long ldc, float beta) {
  const long total = M * N;
  for (long e = blockIdx.x * 256L + threadIdx.x; e < total; e += (long)gridDim.x * 256L) {
    float acc = 0.f;
    for (int q = 0; q < S; ++q) acc += W[(long)q * total + e];
    const long off = e % M + (e / M) * ldc;
    C[off] = beta != 0.f ? acc + beta * C[off] : acc;
  }
}

constexpr int MAX_PART = 1024;

// 16 bytes of Ti, loaded by one global_load_dwordx4
template <typename Ti>
struct alignas(16) Vec16 {
  Ti x[16 / sizeof(Ti)];
};

// per-block maxima of |X|.  A contiguous X (ldx == rows) is swept flat, with 16-byte loads
// when `vec` (X 16-byte aligned), four in flight per thread; otherwise each block takes whole
// columns.  No per-element index division.
template <typename Ti>
__global__ __launch_bounds__(256) void maxabs_kernel(const Ti* __restrict__ X, long rows, long cols, long ldx,
                                                      int vec, double* __restrict__ part) {
  Ti m = Ti(0);
  const long stride = (long)gridDim.x * 256L;
  if (ldx == rows) {
    const long total = rows * cols;
    long done = 0;
    if (vec) {
      constexpr int W = 16 / sizeof(Ti);
      const long nv = total / W;
      const Vec16<Ti>* Xv = reinterpret_cast<const Vec16<Ti>*>(X);
      long e = blockIdx.x * 256L + threadIdx.x;
      for (; e + 3 * stride < nv; e += 4 * stride) {
        Vec16<Ti> q[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) q[u] = Xv[e + u * stride];
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
          for (int w = 0; w < W; ++w) m = fmax(m, fabs(q[u].x[w]));
      }
      for (; e < nv; e += stride) {
        const Vec16<Ti> q = Xv[e];
#pragma unroll
        for (int w = 0; w < W; ++w) m = fmax(m, fabs(q.x[w]));
      }
      done = nv * W;
    }
    for (long e = done + blockIdx.x * 256L + threadIdx.x; e < total; e += stride) m = fmax(m, fabs(X[e]));
  } else {
    for (long c = blockIdx.x; c < cols; c += gridDim.x)
      for (long r = threadIdx.x; r < rows; r += 256) m = fmax(m, fabs(X[r + c * ldx]));
  }
  __shared__ double red[256];
  red[threadIdx.x] = (double)m;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// s = 2^k, k = 15 - e with max = f 2^e, f in [0.5, 1): max * s < 2^15.  A zero or
// non-finite maximum gives k = 0 (non-finite values then reach the result).
__global__ __launch_bounds__(256) void split_scale_kernel(const double* __restrict__ part, int nb,
                                                           int* __restrict__ sexp) {
  __shared__ double red[256];
  double m = 0.0;
  for (int i = threadIdx.x; i < nb; i += 256) m = fmax(m, part[i]);
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double mx = red[0];
    int k = 0;
    if (mx > 0.0 && isfinite(mx)) {
      int e;
      (void)frexp(mx, &e);
      k = 15 - e;
    }
    sexp[0] = k;
  }
}

// interleaved planes from X (rows x cols, leading dimension ldx); each thread writes
// 8 consecutive rows of one column (rows_p is a multiple of 32, so they share a k-block)
template <typename Ti>
__global__ __launch_bounds__(256) void split_kernel(const Ti* __restrict__ X, long rows, long cols, long ldx,
                                                     long rows_p, long cols_p, const int* __restrict__ sexp,
                                                     _Float16* __restrict__ P) {
  const int k = sexp[0];
  const long r8 = rows_p / 8, total = r8 * cols_p;
  for (long e = blockIdx.x * 256L + threadIdx.x; e < total; e += (long)gridDim.x * 256L) {
    const long rq = e % r8, c = e / r8;
    h8 vh, vl;
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      const long r = rq * 8 + t;
      const Ti x = (r < rows && c < cols) ? (Ti)ldexp(X[r + c * ldx], k) : Ti(0);
      const _Float16 h = (_Float16)(float)x;
      vh[t] = h;
      vl[t] = (_Float16)(float)(x - (Ti)h);
    }
    const long off = c * 2 * rows_p + (rq / 4) * 64 + (rq % 4) * 8;
    *reinterpret_cast<h8*>(P + off) = vh;
    *reinterpret_cast<h8*>(P + off + 32) = vl;
  }
}

// the same planes for tall operands (rows_p / 8 >= 192): each block takes whole columns, each
// thread 8 consecutive rows of one, read with 16-byte loads when `vec` (X 16-byte aligned and
// ldx * sizeof(Ti) a multiple of 16) and the 8 rows are inside X.  Columns cols..cols_p-1 and
// rows rows..rows_p-1 are written as zeros.
template <typename Ti>
__global__ __launch_bounds__(256) void split_cols_kernel(const Ti* __restrict__ X, long rows, long cols, long ldx,
                                                          long rows_p, long cols_p, int vec,
                                                          const int* __restrict__ sexp, _Float16* __restrict__ P) {
  constexpr int W = 16 / sizeof(Ti);
  const int k = sexp[0];
  const int r8 = (int)(rows_p / 8);
  for (long c = blockIdx.x; c < cols_p; c += gridDim.x) {
    _Float16* Pc = P + c * 2 * rows_p;
    for (int rq = threadIdx.x; rq < r8; rq += 256) {
      const long r0 = 8L * rq;
      Ti x[8];
      if (c < cols && vec && r0 + 8 <= rows) {
        const Vec16<Ti>* src = reinterpret_cast<const Vec16<Ti>*>(X + c * ldx + r0);
#pragma unroll
        for (int u = 0; u < 8 / W; ++u) {
          const Vec16<Ti> q = src[u];
#pragma unroll
          for (int w = 0; w < W; ++w) x[u * W + w] = q.x[w];
        }
      } else {
#pragma unroll
        for (int t = 0; t < 8; ++t) x[t] = (c < cols && r0 + t < rows) ? X[c * ldx + r0 + t] : Ti(0);
      }
      h8 vh, vl;
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        const Ti y = (Ti)ldexp(x[t], k);
        const _Float16 h = (_Float16)(float)y;
        vh[t] = h;
        vl[t] = (_Float16)(float)(y - (Ti)h);
      }
      const long off = (rq / 4) * 64 + (rq % 4) * 8;
      *reinterpret_cast<h8*>(Pc + off) = vh;
      *reinterpret_cast<h8*>(Pc + off + 32) = vl;
    }
  }
}

// planes of X^T: X is stored cols x rows (leading dimension ldx), the planes are
// rows x cols.  64 x 64 tiles pass through LDS so that both the reads of X (along
// its columns' contiguous index) and the 16-byte plane stores are coalesced.
template <typename Ti>
__global__ __launch_bounds__(256) void split_t_kernel(const Ti* __restrict__ X, long rows, long cols, long ldx,
                                                       long rows_p, const int* __restrict__ sexp,
                                                       _Float16* __restrict__ P) {
  __shared__ Ti tile[64][65];  // [r][c]
  const long r0 = blockIdx.x * 64L, c0 = blockIdx.y * 64L;
  const int t = threadIdx.x;
#pragma unroll 4
  for (int i = 0; i < 16; ++i) {
    const int rr = t / 64 + 4 * i, cc = t % 64;
    const long r = r0 + rr, c = c0 + cc;
    tile[rr][cc] = (r < rows && c < cols) ? X[c + r * ldx] : Ti(0);
  }
  __syncthreads();
  const int k = sexp[0];
#pragma unroll
  for (int p = 0; p < 2; ++p) {
    const int q = t + 256 * p;
    const int rq = q % 8, cc = q / 8;
    const long r = r0 + 8 * rq;
    if (r >= rows_p) continue;
    h8 vh, vl;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const Ti x = (Ti)ldexp(tile[8 * rq + e][cc], k);
      const _Float16 h = (_Float16)(float)x;
      vh[e] = h;
      vl[e] = (_Float16)(float)(x - (Ti)h);
    }
    const long off = (c0 + cc) * 2 * rows_p + (r / 32) * 64 + (r % 32);
    *reinterpret_cast<h8*>(P + off) = vh;
    *reinterpret_cast<h8*>(P + off + 32) = vl;
  }
}

int grid_for(long n, int cap) { return (int)std::max(1L, std::min((n + 255) / 256, (long)cap)); }
int aligned16(const void* p) { return reinterpret_cast<uintptr_t>(p) % 16 == 0; }

}  // namespace

void Split16Buf::ensure(long rows, long cols) {
  JUES_REQUIRE(rows > 0 && cols > 0, "split16: empty matrix");
  // 256 spare zero columns: a column view (split16_cols) near the end still has whole tiles
  const long rp = (rows + BK - 1) / BK * BK, cp = (cols + PAD_C - 1) / PAD_C * PAD_C + PAD_C;
  const size_t planes = 2 * (size_t)rp * cp * sizeof(_Float16);
  const size_t need = planes + 256;
  if (mem.bytes < need) {
    mem.release();
    mem.alloc(need);
  }
  char* base = mem.as<char>();
  s.p = reinterpret_cast<_Float16*>(base);
  s.sexp = reinterpret_cast<int*>(base + planes);
  s.rows = rows;
  s.cols = cols;
  s.rows_p = rp;
  s.cols_p = cp;
}

template <typename Ti>
void split16(jues_ctx* ctx, const Ti* X, long rows, long cols, long ldx, Split16Buf& out, const char* tag,
             const int* sexp_of) {
  JUES_REQUIRE(ldx >= rows, "split16: ldx < rows");
  out.ensure(rows, cols);
  const Split16& s = out.s;
  const long n = rows * cols;
  KTimer t(ctx, tag ? tag : "split16", 0.0, (sexp_of ? 1.0 : 2.0) * sizeof(Ti) * n + 4.0 * s.rows_p * s.cols_p);
  if (sexp_of) {
    // the caller vouches that X holds the same values as the operand whose scale this is
    JUES_HIP(hipMemcpyAsync(s.sexp, sexp_of, sizeof(int), hipMemcpyDeviceToDevice, ctx->stream));
  } else {
    double* part = static_cast<double*>(scratch(ctx, 3, MAX_PART * sizeof(double)));
    const int nb = grid_for(n, MAX_PART);
    hipLaunchKernelGGL((maxabs_kernel<Ti>), dim3(nb), dim3(256), 0, ctx->stream, X, rows, cols, ldx, aligned16(X),
                       part);
    hipLaunchKernelGGL(split_scale_kernel, dim3(1), dim3(256), 0, ctx->stream, part, nb, s.sexp);
  }
  if (s.rows_p / 8 >= 192) {
    const int vec = aligned16(X) && (ldx * (long)sizeof(Ti)) % 16 == 0;
    hipLaunchKernelGGL((split_cols_kernel<Ti>), dim3((unsigned)std::min(s.cols_p, 65536L)), dim3(256), 0, ctx->stream,
                       X, rows, cols, ldx, s.rows_p, s.cols_p, vec, s.sexp, s.p);
  } else {
    hipLaunchKernelGGL((split_kernel<Ti>), dim3(grid_for(s.rows_p / 8 * s.cols_p, 131072)), dim3(256), 0, ctx->stream,
                       X, rows, cols, ldx, s.rows_p, s.cols_p, s.sexp, s.p);
  }
  JUES_HIP(hipGetLastError());
}

template <typename Ti>
void split16t(jues_ctx* ctx, const Ti* X, long rows, long cols, long ldx, Split16Buf& out, const char* tag) {
  JUES_REQUIRE(ldx >= cols, "split16t: ldx < cols");
  out.ensure(rows, cols);
  const Split16& s = out.s;
  double* part = static_cast<double*>(scratch(ctx, 3, MAX_PART * sizeof(double)));
  const long n = rows * cols;
  KTimer t(ctx, tag ? tag : "split16", 0.0, 2.0 * sizeof(Ti) * n + 4.0 * s.rows_p * s.cols_p);
  const int nb = grid_for(n, MAX_PART);
  // the stored matrix is cols x rows
  hipLaunchKernelGGL((maxabs_kernel<Ti>), dim3(nb), dim3(256), 0, ctx->stream, X, cols, rows, ldx, aligned16(X), part);
  hipLaunchKernelGGL(split_scale_kernel, dim3(1), dim3(256), 0, ctx->stream, part, nb, s.sexp);
  const dim3 grid((unsigned)((s.rows_p + 63) / 64), (unsigned)(s.cols_p / 64));
  hipLaunchKernelGGL((split_t_kernel<Ti>), grid, dim3(256), 0, ctx->stream, X, rows, cols, ldx, s.rows_p, s.sexp,
                     s.p);
  JUES_HIP(hipGetLastError());
}

void gemm16_tn(jues_ctx* ctx, long M, long N, float alpha, const Split16& A, const Split16& B, float beta, float* C,
               long ldc, const char* tag) {
  if (M <= 0 || N <= 0) return;
  JUES_REQUIRE(A.rows == B.rows && A.rows_p == B.rows_p, "gemm16: operands differ in K");
  JUES_REQUIRE(M <= A.cols && N <= B.cols, "gemm16: M or N exceeds the operand");
  JUES_REQUIRE(ldc >= M, "gemm16: ldc < M");
  JUES_REQUIRE(A.rows_p % BK == 0, "gemm16: planes not padded");
  JUES_REQUIRE((M + BM - 1) / BM * BM <= A.cols_p && (N + BN - 1) / BN * BN <= B.cols_p, "gemm16: tiles exceed planes");
  // 32-bit per-lane byte offsets inside one tile's columns
  JUES_REQUIRE(4.0 * (double)std::max(A.rows_p, B.rows_p) * std::max(BM, BN) < 2147483647.0, "gemm16: K too large");
  const long tiles_m = (M + BM - 1) / BM, tiles_n = (N + BN - 1) / BN;
  JUES_REQUIRE(tiles_m * tiles_n < (1L << 31), "gemm16: too many tiles");
  const double flops = 2.0 * M * N * A.rows;
  const double bytes = 4.0 * ((double)A.rows * M + (double)A.rows * N) + (beta != 0.0f ? 8.0 : 4.0) * M * N;
  const char* name = tag ? tag : "gemm16";
  const int kind = !tag ? 0
                        : (std::strcmp(tag, "cc_ring") == 0          ? 1
                           : std::strcmp(tag, "cc_pp_ladder") == 0 ? 2
                           : std::strcmp(tag, "cc_hh") == 0        ? 3
                                                                   : 0);
  // Fewer tiles than two rounds of workgroups (one per CU) and a long K: split K into S
  // chunks in one launch, partial products to a workspace, summed into C in a fixed order.  A
  // long-K product whose tiles fill their last round poorly (under 80% overall) is split into
  // the fewest chunks that fill the rounds, when the partial products fit in 2 GiB.
  const long Kp = A.rows_p, tiles = tiles_m * tiles_n;
  long S = tiles < 256 ? std::min(Kp / 1024, (512 + tiles - 1) / tiles) : 1;
  auto round_fill = [&](long s) {
    const long t = tiles * s;
    return (double)t / (double)(((t + 255) / 256) * 256);
  };
  if (tiles >= 256 && Kp >= 32768 && round_fill(1) < 0.8 && !no_alloc_active()) {
    S = 1;
    for (long s = 2; s <= 16; ++s)
      if (round_fill(s) > round_fill(S) + 0.02 && 4.0 * s * M * N <= double(2L << 30)) S = s;
  }
  long kchunk = Kp;
  if (S >= 2) {
    kchunk = ((Kp + S - 1) / S + BK - 1) / BK * BK;
    S = (Kp + kchunk - 1) / kchunk;
  } else {
    S = 1;
  }
  float* Cw = C;
  long ldw = ldc, sW = 0;
  float bw = beta;
  if (S > 1) {
    Cw = static_cast<float*>(scratch(ctx, 17, sizeof(float) * (size_t)S * M * N));
    ldw = M;
    sW = M * N;
    bw = 0.0f;
  }
  {
    KTimer timer(ctx, name, flops, bytes);
    dim3 grid((unsigned)tiles, 1, (unsigned)S);
#define JUES_G16(K_)                                                                                             \
  hipLaunchKernelGGL((gemm16_kernel<K_>), grid, dim3(NT), 0, ctx->stream, (int)M, (int)N, (int)Kp, alpha, A.p,    \
                     2 * A.rows_p, A.sexp, B.p, 2 * B.rows_p, B.sexp, bw, Cw, ldw, (int)tiles_m, (int)tiles_n,     \
                     (int)kchunk, sW)
    if (kind == 1) JUES_G16(1);
    else if (kind == 2) JUES_G16(2);
    else if (kind == 3) JUES_G16(3);
    else JUES_G16(0);
#undef JUES_G16
    JUES_HIP(hipGetLastError());
  }
  if (S > 1) {
    KTimer t(ctx, name, 0.0, 4.0 * (S + (beta != 0.0f ? 2 : 1)) * M * N);
    hipLaunchKernelGGL(splitk16_reduce_kernel, dim3((unsigned)std::min((M * N + 255) / 256, 65536L)), dim3(256), 0,
                       ctx->stream, Cw, (int)S, M, N, C, ldc, beta);
    JUES_HIP(hipGetLastError());
  }
}

template void split16<float>(jues_ctx*, const float*, long, long, long, Split16Buf&, const char*, const int*);
template void split16<double>(jues_ctx*, const double*, long, long, long, Split16Buf&, const char*, const int*);
template void split16t<float>(jues_ctx*, const float*, long, long, long, Split16Buf&, const char*);
template void split16t<double>(jues_ctx*, const double*, long, long, long, Split16Buf&, const char*);

}  // namespace jues
