// LDS-tiled MFMA GEMM for gfx950 (FP64 and FP32), column-major, batched.
//
// One 256-thread workgroup (4 wavefronts of 64) computes a 128x128 tile of C.
// Each wavefront owns a 64x64 sub-tile as a 4x4 grid of 16x16 MFMA blocks
// (v_mfma_f64_16x16x4f64 / v_mfma_f32_16x16x4f32).  The MFMA is issued with the
// operands swapped (B^T * A^T) so that each lane's accumulator column is an m
// index: lanes 0..15 then hold 16 consecutive rows of C and the epilogue stores
// are 16-element contiguous runs along the leading dimension.
//
// The K loop stages BK-deep slices of A and B through LDS; the next slice's
// global loads are issued before the MFMAs of the current one so the load
// latency overlaps the matrix work.  Workgroup ids are remapped so that each of
// the 8 XCDs (each with its own L2) receives a contiguous range of tiles, and
// tiles are walked in groups of 8 tile-rows to reuse B columns within an L2.
#include "jues_internal.h"

#include <algorithm>
#include <cstring>

namespace jues {
namespace {

typedef double dvec4 __attribute__((ext_vector_type(4)));
typedef float fvec4 __attribute__((ext_vector_type(4)));

template <typename T> struct Mfma;
template <> struct Mfma<double> {
  typedef dvec4 acc;
  static constexpr int BK = 16;
  // accumulator element r of lane l is row (l/16) + 4r of the 16x16 block
  __device__ static inline int row(int lane, int r) { return (lane >> 4) + 4 * r; }
  __device__ static inline acc op(double a, double b, acc c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
};
template <> struct Mfma<float> {
  typedef fvec4 acc;
  static constexpr int BK = 32;
  // accumulator element r of lane l is row 4(l/16) + r of the 16x16 block
  __device__ static inline int row(int lane, int r) { return 4 * (lane >> 4) + r; }
  __device__ static inline acc op(float a, float b, acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
};

constexpr int BM = 128, BN = 128, NT = 256, PAD = 1, NXCD = 8, GROUP_M = 8;

__device__ inline void tile_coords(int pid, int nwg, int tiles_m, int tiles_n, int& tm, int& tn) {
  // XCD-aware remap (bijective for any nwg): consecutive hardware ids go to
  // different XCDs; give each XCD a contiguous block of logical tile ids.
  const int per = (nwg + NXCD - 1) / NXCD;
  int tall = nwg % NXCD;
  if (tall == 0) tall = NXCD;
  const int xcd = pid % NXCD, loc = pid / NXCD;
  const int p2 = xcd < tall ? xcd * per + loc : tall * per + (xcd - tall) * (per - 1) + loc;
  // grouped ordering: GROUP_M tile-rows share each sweep over tile-columns
  const int group_size = GROUP_M * tiles_n;
  const int gid = p2 / group_size;
  const int first_m = gid * GROUP_M;
  const int gsz = min(tiles_m - first_m, GROUP_M);
  const int r = p2 - gid * group_size;
  tm = first_m + r % gsz;
  tn = r / gsz;
}

// KIND only separates kernel symbols (0 generic, 1 CC ring, 2 CC pp-ladder) so that
// profiles attribute time to the coupled-cluster hot spots.
template <typename T, bool TA, bool TB, int KIND>
__global__ __launch_bounds__(256, 2) void gemm_kernel(int M, int N, int K, T alpha,
                                                      const T* __restrict__ A, long lda, long sA,
                                                      const T* __restrict__ B, long ldb, long sB,
                                                      T beta, T* __restrict__ C, long ldc, long sC,
                                                      int tiles_m, int tiles_n) {
  constexpr int BK = Mfma<T>::BK;
  constexpr int LA = BM * BK / NT;  // A elements loaded per thread per slice
  constexpr int LB = BN * BK / NT;
  __shared__ T As[BK][BM + PAD];
  __shared__ T Bs[BK][BN + PAD];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  A += (long)blockIdx.z * sA;
  B += (long)blockIdx.z * sB;
  C += (long)blockIdx.z * sC;

  int tm, tn;
  tile_coords(blockIdx.x, gridDim.x, tiles_m, tiles_n, tm, tn);
  const int m0 = tm * BM, n0 = tn * BN;

  T ra[LA], rb[LB];
  auto load = [&](int k0) {
#pragma unroll
    for (int r = 0; r < LA; ++r) {
      const int e = tid + NT * r;
      const int m = TA ? e / BK : e % BM;
      const int k = TA ? e % BK : e / BM;
      const int gm = m0 + m, gk = k0 + k;
      T val = T(0);
      if (gm < M && gk < K) val = TA ? A[(long)gk + (long)gm * lda] : A[(long)gm + (long)gk * lda];
      ra[r] = val;
    }
#pragma unroll
    for (int r = 0; r < LB; ++r) {
      const int e = tid + NT * r;
      const int n = TB ? e % BN : e / BK;
      const int k = TB ? e / BN : e % BK;
      const int gn = n0 + n, gk = k0 + k;
      T val = T(0);
      if (gn < N && gk < K) val = TB ? B[(long)gn + (long)gk * ldb] : B[(long)gk + (long)gn * ldb];
      rb[r] = val;
    }
  };
  auto store = [&]() {
#pragma unroll
    for (int r = 0; r < LA; ++r) {
      const int e = tid + NT * r;
      const int m = TA ? e / BK : e % BM;
      const int k = TA ? e % BK : e / BM;
      As[k][m] = ra[r];
    }
#pragma unroll
    for (int r = 0; r < LB; ++r) {
      const int e = tid + NT * r;
      const int n = TB ? e % BN : e / BK;
      const int k = TB ? e / BN : e % BK;
      Bs[k][n] = rb[r];
    }
  };

  typename Mfma<T>::acc acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = typename Mfma<T>::acc{0, 0, 0, 0};

  const int nk = (K + BK - 1) / BK;
  load(0);
  for (int kt = 0; kt < nk; ++kt) {
    __syncthreads();
    store();
    __syncthreads();
    if (kt + 1 < nk) load((kt + 1) * BK);
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      const int kr = kk + (lane >> 4);
      T af[4], bf[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) af[i] = As[kr][wm * 64 + i * 16 + (lane & 15)];
#pragma unroll
      for (int j = 0; j < 4; ++j) bf[j] = Bs[kr][wn * 64 + j * 16 + (lane & 15)];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = Mfma<T>::op(bf[j], af[i], acc[i][j]);
    }
  }

  // acc[i][j][r] holds C(m, n) with m = lane%16 of block i, n = Mfma<T>::row(lane, r) of block j.
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int m = m0 + wm * 64 + i * 16 + (lane & 15);
    if (m >= M) continue;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int n = n0 + wn * 64 + j * 16 + Mfma<T>::row(lane, r);
        if (n < N) {
          const long off = (long)m + (long)n * ldc;
          T v = alpha * acc[i][j][r];
          if (beta != T(0)) v += beta * C[off];
          C[off] = v;
        }
      }
    }
  }
}

template <typename T>
__global__ void scale_kernel(T* C, long M, long N, long ldc, T beta) {
  const long total = M * N;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
    const long m = e % M, n = e / M;
    C[m + n * ldc] = beta == T(0) ? T(0) : beta * C[m + n * ldc];
  }
}

// C = beta*C + sum_s W[s] (fixed order), W[s] is M x N with leading dimension M.
template <typename T>
__global__ void splitk_reduce_kernel(const T* __restrict__ W, int S, long M, long N, T* __restrict__ C, long ldc,
                                     T beta) {
  const long total = M * N;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
    double acc = 0.0;
    for (int q = 0; q < S; ++q) acc += (double)W[(long)q * total + e];
    const long m = e % M, n = e / M;
    const long off = m + n * ldc;
    if (beta != T(0)) acc += (double)beta * (double)C[off];
    C[off] = (T)acc;
  }
}

}  // namespace

template <typename T>
static void gemm_impl(jues_ctx* ctx, bool ta, bool tb, long M, long N, long K, T alpha, const T* A, long lda,
                      long sA, const T* B, long ldb, long sB, T beta, T* C, long ldc, long sC, int batch,
                      const char* tag, bool allow_split) {
  if (M <= 0 || N <= 0 || batch <= 0) return;
  JUES_REQUIRE(M < (1L << 31) && N < (1L << 31) && K < (1L << 31), "gemm: dimension too large");
  JUES_REQUIRE(ldc >= M, "gemm: ldc < M");
  JUES_REQUIRE(lda >= (ta ? K : M) || K == 0, "gemm: lda too small");
  JUES_REQUIRE(ldb >= (tb ? N : K) || K == 0, "gemm: ldb too small");
  const char* name = tag ? tag : (sizeof(T) == 8 ? "gemm_f64" : "gemm_f32");
  const int kind = !tag ? 0 : (std::strcmp(tag, "cc_ring") == 0 ? 1 : (std::strcmp(tag, "cc_pp_ladder") == 0 ? 2 : 0));
  if (K <= 0) {
    KTimer t(ctx, name, 0, 0);
    for (int b = 0; b < batch; ++b)
      hipLaunchKernelGGL(scale_kernel<T>, dim3(1024), dim3(256), 0, ctx->stream, C + (long)b * sC, M,
                         N, ldc, beta);
    JUES_HIP(hipGetLastError());
    return;
  }
  const long tiles_m = (M + BM - 1) / BM, tiles_n = (N + BN - 1) / BN;
  JUES_REQUIRE(tiles_m * tiles_n < (1L << 31), "gemm: too many tiles");
  // Few output tiles and a long K (the Fock-like terms): split K over workgroups,
  // write partial products to a workspace and reduce them in a fixed order.
  if (allow_split && batch == 1 && tiles_m * tiles_n < 128 && K >= 4096) {
    const long want = std::max(2L, std::min(K / 1024, 512L / (tiles_m * tiles_n)));
    const int BKs = sizeof(T) == 8 ? 16 : 32;
    long kc = (K + want - 1) / want;
    kc = (kc + BKs - 1) / BKs * BKs;
    const int S = (int)((K + kc - 1) / kc);
    T* Wk = static_cast<T*>(scratch(ctx, 9, sizeof(T) * (size_t)S * M * N));
    const long sa = ta ? kc : kc * lda, sb = tb ? kc * ldb : kc;
    if (S > 1)
      gemm_impl<T>(ctx, ta, tb, M, N, kc, alpha, A, lda, sa, B, ldb, sb, T(0), Wk, M, M * N, S - 1, tag, false);
    const long klast = K - (long)(S - 1) * kc;
    gemm_impl<T>(ctx, ta, tb, M, N, klast, alpha, A + (long)(S - 1) * sa, lda, 0, B + (long)(S - 1) * sb, ldb, 0,
                 T(0), Wk + (long)(S - 1) * M * N, M, 0, 1, tag, false);
    KTimer t(ctx, name, 0.0, (double)sizeof(T) * S * M * N);
    const long tot = M * N;
    const int nb = (int)std::max(1L, std::min((tot + 255) / 256, 65536L));
    hipLaunchKernelGGL((splitk_reduce_kernel<T>), dim3(nb), dim3(256), 0, ctx->stream, Wk, S, M, N, C, ldc, beta);
    JUES_HIP(hipGetLastError());
    return;
  }
  const double flops = 2.0 * M * N * K * batch;
  const double bytes = (double)sizeof(T) * ((double)M * K + (double)K * N + (beta != T(0) ? 2.0 : 1.0) * M * N) * batch;
  KTimer timer(ctx, name, flops, bytes);
  for (int b0 = 0; b0 < batch; b0 += 65535) {
    const int nb = std::min(batch - b0, 65535);
    dim3 grid((unsigned)(tiles_m * tiles_n), 1, nb);
    const T* Ab = A + (long)b0 * sA;
    const T* Bb = B + (long)b0 * sB;
    T* Cb = C + (long)b0 * sC;
#define JUES_GEMM_LAUNCH(TA_, TB_, K_)                                                             \
  hipLaunchKernelGGL((gemm_kernel<T, TA_, TB_, K_>), grid, dim3(NT), 0, ctx->stream, (int)M, (int)N, (int)K, \
                     alpha, Ab, lda, sA, Bb, ldb, sB, beta, Cb, ldc, sC, (int)tiles_m, (int)tiles_n)
    if (!ta && !tb) {
      if (kind == 1) JUES_GEMM_LAUNCH(false, false, 1);
      else if (kind == 2) JUES_GEMM_LAUNCH(false, false, 2);
      else JUES_GEMM_LAUNCH(false, false, 0);
    } else if (!ta && tb) JUES_GEMM_LAUNCH(false, true, 0);
    else if (ta && !tb) JUES_GEMM_LAUNCH(true, false, 0);
    else JUES_GEMM_LAUNCH(true, true, 0);
#undef JUES_GEMM_LAUNCH
    JUES_HIP(hipGetLastError());
  }
}

template <typename T>
void gemm(jues_ctx* ctx, bool ta, bool tb, long M, long N, long K, T alpha, const T* A, long lda, long sA,
          const T* B, long ldb, long sB, T beta, T* C, long ldc, long sC, int batch, const char* tag) {
  gemm_impl<T>(ctx, ta, tb, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, batch, tag, true);
}

template void gemm<double>(jues_ctx*, bool, bool, long, long, long, double, const double*, long, long,
                           const double*, long, long, double, double*, long, long, int, const char*);
template void gemm<float>(jues_ctx*, bool, bool, long, long, long, float, const float*, long, long,
                          const float*, long, long, float, float*, long, long, int, const char*);

}  // namespace jues
