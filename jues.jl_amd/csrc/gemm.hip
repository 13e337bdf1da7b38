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

constexpr int NT = 256, NXCD = 8, GROUP_M = 8;

// Raw buffer loads: the 64-bit base lives in a scalar resource descriptor, the
// per-lane part is one 32-bit byte offset and the per-element part a scalar
// offset, so an interior tile's loads hold no 64-bit address registers.
constexpr int kBufferDword3 = 0x00020000;  // gfx9-family raw buffer descriptor word 3
__device__ inline __amdgpu_buffer_rsrc_t make_rsrc(const void* base) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, 0x7fffffff, kBufferDword3);
}
template <typename T> __device__ inline T bload(__amdgpu_buffer_rsrc_t r, int voff, int soff);
template <> __device__ inline double bload<double>(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}
template <> __device__ inline float bload<float>(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}

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

// As tile_coords over nwg workgroup ids of which only the first tiles_m * tiles_n carry a
// tile (a launch padded to a multiple of NXCD); false for the padding ids.
__device__ inline bool tile_coords_bounded(int pid, int nwg, int tiles_m, int tiles_n, int& tm, int& tn) {
  const int per = (nwg + NXCD - 1) / NXCD;
  int tall = nwg % NXCD;
  if (tall == 0) tall = NXCD;
  const int xcd = pid % NXCD, loc = pid / NXCD;
  const int p2 = xcd < tall ? xcd * per + loc : tall * per + (xcd - tall) * (per - 1) + loc;
  if (p2 >= tiles_m * tiles_n) return false;
  const int group_size = GROUP_M * tiles_n;
  const int gid = p2 / group_size;
  const int first_m = gid * GROUP_M;
  const int gsz = min(tiles_m - first_m, GROUP_M);
  const int r = p2 - gid * group_size;
  tm = first_m + r % gsz;
  tn = r / gsz;
  return true;
}

// One product of a high-occupancy kernel launch: C = alpha A op(B) + beta C, batched over
// blockIdx.z with strides sA, sB, sC.  A launch carries one or two of them (two
// independent products sharing the grid, so that neither leaves a partly idle last round).
template <typename T>
struct GemmSide {
  const T* A;
  const T* B;
  T* C;
  long lda, ldb, ldc, sA, sB, sC;
  int M, N, K, tiles_m, tiles_n;
  int Klast;  // K of the last batch entry (a split-K launch's shorter last chunk), else K
  // optional second contraction segment: C += alpha A1 op(B1) over K1 more columns (same lda,
  // ldb and batch strides); with K1 > 0, K and K1 are multiples of 16
  const T* A1;
  const T* B1;
  int K1;
  // FP32 products that write an FP64 result (the OUT64 instantiation; C is not used there)
  double* C64 = nullptr;
};

// Shared-memory layout of one operand slice.  When the operand's contiguous
// global dimension is the row index (m, or n for B) the slice is stored k-major,
// [k][m], padded so that the four k rows an MFMA fragment read touches start 16
// (FP32) / 32 (FP64) banks apart; when k is contiguous it is stored m-major,
// [m][k] with an odd padded row, so that both the coalesced stores and the
// fragment reads are bank-conflict free.
template <typename T, bool KMAJOR, int ROWS, int BK>
struct Slice {
  static constexpr int PADR = sizeof(T) == 8 ? 16 : 16;  // k-major row padding (elements)
  static constexpr int PADK = 1;                          // m-major row padding (elements)
  static constexpr int SIZE = KMAJOR ? BK * (ROWS + PADR) : ROWS * (BK + PADK);
  __device__ static inline int idx(int r, int k) { return KMAJOR ? k * (ROWS + PADR) + r : r * (BK + PADK) + k; }
};

// KIND only separates kernel symbols (0 generic, 1 CC ring, 2 CC pp-ladder) so that
// profiles attribute time to the coupled-cluster hot spots.
// BM x BN workgroup tile over WM x (4/WM) wavefronts of 64x64: 128x128 (2x2) in general; 64x256
// (1x4) and 256x64 (4x1) for skinny products (M or N <= 64 against a long other side), where a
// 128-wide tile would compute mostly padding.
template <typename T, bool TA, bool TB, int KIND, int BM = 128, int BN = 128, int WM = 2>
__global__ __launch_bounds__(256, (BM == 128 && BN == 128) ? 2 : 1) void gemm_kernel(int M, int N, int K, T alpha,
                                                      const T* __restrict__ A, long lda, long sA,
                                                      const T* __restrict__ B, long ldb, long sB,
                                                      T beta, T* __restrict__ C, long ldc, long sC,
                                                      int tiles_m, int tiles_n, int fast_ok) {
  constexpr int BK = Mfma<T>::BK;
  constexpr int LA = BM * BK / NT;  // A elements loaded per thread per slice
  constexpr int LB = BN * BK / NT;
  // A is m-contiguous unless transposed; B is k-contiguous unless transposed.
  using SA = Slice<T, !TA, BM, BK>;
  using SB = Slice<T, TB, BN, BK>;
  __shared__ T As[2][SA::SIZE];
  __shared__ T Bs[2][SB::SIZE];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  static_assert(BM == 64 * WM && BN == 64 * (4 / WM), "tile does not match the wave layout");
  const int wm = wave % WM, wn = wave / WM;
  A += (long)blockIdx.z * sA;
  B += (long)blockIdx.z * sB;
  C += (long)blockIdx.z * sC;

  int tm, tn;
  tile_coords(blockIdx.x, gridDim.x, tiles_m, tiles_n, tm, tn);
  const int m0 = tm * BM, n0 = tn * BN;

  // element e of a thread's share: the contiguous global dimension varies fastest
  auto a_rk = [](int e, int& m, int& k) {
    if (TA) { k = e % BK; m = e / BK; } else { m = e % BM; k = e / BM; }
  };
  auto b_rk = [](int e, int& n, int& k) {
    if (TB) { n = e % BN; k = e / BN; } else { k = e % BK; n = e / BK; }
  };

  T ra[LA], rb[LB];
  // Interior tiles and full k-slices load without any bounds logic: element r of
  // a thread's share sits at p0 + k0*kstride + r*rstride.  Edge tiles and the
  // last partial k-slice clamp rows/columns into the matrix (those values only
  // reach output elements that are never stored) and zero k >= K with a select,
  // so no load is ever predicated by a branch.
  // per-lane byte offsets inside a tile's k-slice and per-element scalar strides
  const int a_lane = (int)(sizeof(T) * (TA ? (tid % BK) + (long)(tid / BK) * lda : (tid % BM) + (long)(tid / BM) * lda));
  const int b_lane = (int)(sizeof(T) * (TB ? (tid % BN) + (long)(tid / BN) * ldb : (tid % BK) + (long)(tid / BK) * ldb));
  const int a_rs = (int)(sizeof(T) * (TA ? (long)(NT / BK) * lda : (long)(NT / BM) * lda));
  const int b_rs = (int)(sizeof(T) * (TB ? (long)(NT / BN) * ldb : (long)(NT / BK) * ldb));
  const bool interior = fast_ok && (m0 + BM <= M) && (n0 + BN <= N);
  auto load = [&](int k0) {
    if (interior && k0 + BK <= K) {
      const auto rA = make_rsrc(TA ? A + k0 + (long)m0 * lda : A + m0 + (long)k0 * lda);
      const auto rB = make_rsrc(TB ? B + n0 + (long)k0 * ldb : B + k0 + (long)n0 * ldb);
#pragma unroll
      for (int r = 0; r < LA; ++r) ra[r] = bload<T>(rA, a_lane, r * a_rs);
#pragma unroll
      for (int r = 0; r < LB; ++r) rb[r] = bload<T>(rB, b_lane, r * b_rs);
    } else {
#pragma unroll
      for (int r = 0; r < LA; ++r) {
        int m, k;
        a_rk(tid + NT * r, m, k);
        const int gm = min(m0 + m, M - 1), gk = k0 + k, gkc = min(gk, K - 1);
        const T val = TA ? A[(long)gkc + (long)gm * lda] : A[(long)gm + (long)gkc * lda];
        ra[r] = gk < K ? val : T(0);
      }
#pragma unroll
      for (int r = 0; r < LB; ++r) {
        int n, k;
        b_rk(tid + NT * r, n, k);
        const int gn = min(n0 + n, N - 1), gk = k0 + k, gkc = min(gk, K - 1);
        const T val = TB ? B[(long)gn + (long)gkc * ldb] : B[(long)gkc + (long)gn * ldb];
        rb[r] = gk < K ? val : T(0);
      }
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int r = 0; r < LA; ++r) {
      int m, k;
      a_rk(tid + NT * r, m, k);
      As[buf][SA::idx(m, k)] = ra[r];
    }
#pragma unroll
    for (int r = 0; r < LB; ++r) {
      int n, k;
      b_rk(tid + NT * r, n, k);
      Bs[buf][SB::idx(n, k)] = rb[r];
    }
  };

  typename Mfma<T>::acc acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = typename Mfma<T>::acc{0, 0, 0, 0};

  const int fr = lane & 15, fk = lane >> 4;
  T af[2][4], bf[2][4];
  auto frags = [&](int buf, int kk, int slot) {
#pragma unroll
    for (int i = 0; i < 4; ++i) af[slot][i] = As[buf][SA::idx(wm * 64 + i * 16 + fr, kk + fk)];
#pragma unroll
    for (int j = 0; j < 4; ++j) bf[slot][j] = Bs[buf][SB::idx(wn * 64 + j * 16 + fr, kk + fk)];
  };

  const int nk = (K + BK - 1) / BK;
  load(0);
  store(0);
  __syncthreads();
  // the last slice is peeled off, so that the loop body is one block without conditions
  auto compute = [&](int cur) {
    if constexpr (sizeof(T) == 4) {
      // FP32: the 32-cycle MFMAs cannot hide a fragment read, so fragments for
      // step s+1 are fetched while the MFMAs of step s issue
      frags(cur, 0, 0);
#pragma unroll
      for (int s = 0; s < BK / 4; ++s) {
        if (s + 1 < BK / 4) frags(cur, (s + 1) * 4, (s + 1) & 1);
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
          for (int j = 0; j < 4; ++j) acc[i][j] = Mfma<T>::op(bf[s & 1][j], af[s & 1][i], acc[i][j]);
      }
    } else {
      // FP64: sixteen 64-cycle MFMAs per step cover the fragment latency
#pragma unroll
      for (int s = 0; s < BK / 4; ++s) {
        frags(cur, s * 4, 0);
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
          for (int j = 0; j < 4; ++j) acc[i][j] = Mfma<T>::op(bf[0][j], af[0][i], acc[i][j]);
      }
    }
  };
  for (int kt = 0; kt + 1 < nk; ++kt) {
    const int cur = kt & 1;
    load((kt + 1) * BK);
    compute(cur);
    store(cur ^ 1);
    __syncthreads();
  }
  compute((nk - 1) & 1);

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

// FP32 C = alpha A op(B) + beta C with A m-contiguous and B k-contiguous (NN) or
// n-contiguous (NT), 128x128 tile, 4 wavefronts of 64x64.
// The K slice is 16 deep, so a workgroup needs 32 KB of LDS and at most 128 VGPRs per
// lane: four workgroups (16 wavefronts) share a CU and cover each other's load and
// barrier latency (two workgroups of 32-deep slices reach 73% MFMA-busy, this form ~93% with the
// last slice peeled off the main loop, ~89% with a conditional loop body).
// LDS holds each slice as 64-byte rows of 16 k-values per m (A) or n (B); the 16-byte
// chunk q of row r sits at chunk q ^ ((r >> 2) & 3), so slices are written with 16-byte
// stores and fragments read with 16-byte loads.  Lane l of a 16x16x4 MFMA supplies
// k = 4 (l / 16) + s at step s = 0..3 (the same permutation for both operands): one
// 16-byte read per 16-row block.  B's four fragment blocks are held for the slice and A's
// are streamed one block ahead of the MFMAs that use them.  Rows of the last tile row or
// column beyond M or N load a clamped row and are not stored.  A K tail (16 not dividing K)
// is the first slice: zero-padded, loaded element by element once in the prologue.  TA (A
// k-contiguous, TN / TT) stages A like a k-contiguous B; SEG instantiations add a second contraction
// segment (A1, B1 over K1 columns).
typedef float f4v __attribute__((ext_vector_type(4)));

__device__ inline int g32_chunk(int q, int r) { return q ^ ((r >> 2) & 3); }

// Workgroups [0, nt0) run product p0 (nt0 padded up to a multiple of NXCD when a second
// product follows), the rest p1.  OUT64: C = alpha A op(B) is written to the FP64 array C64
// (beta = 0).
template <int KIND, bool TB, bool TA = false, bool SEG = false, bool OUT64 = false>
__global__ __launch_bounds__(256, 4) void gemm32_nn_kernel(GemmSide<float> p0, GemmSide<float> p1, int nt0,
                                                           float alpha, float beta) {
  constexpr int BM = 128, BN = 128, BK = 16;
  __shared__ __attribute__((aligned(16))) float As[2][BM * BK];
  __shared__ __attribute__((aligned(16))) float Bs[2][BN * BK];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const bool second = (int)blockIdx.x >= nt0;
  const GemmSide<float> g = second ? p1 : p0;
  const int M = g.M, N = g.N, K = (int)blockIdx.z == (int)gridDim.z - 1 ? g.Klast : g.K;
  const long lda = g.lda, ldb = g.ldb, ldc = g.ldc;
  const float* __restrict__ A = g.A + (long)blockIdx.z * g.sA;
  const float* __restrict__ B = g.B + (long)blockIdx.z * g.sB;
  float* __restrict__ C = OUT64 ? nullptr : g.C + (long)blockIdx.z * g.sC;
  // a second contraction segment (A1, B1 over K1 columns) only in the SEG instantiations
  const int K1 = SEG ? g.K1 : 0;
  const float* A1 = K1 > 0 ? g.A1 + (long)blockIdx.z * g.sA : A;
  const float* B1 = K1 > 0 ? g.B1 + (long)blockIdx.z * g.sB : B;
  int tm, tn;
  if (!tile_coords_bounded(second ? (int)blockIdx.x - nt0 : (int)blockIdx.x, second ? (int)gridDim.x - nt0 : nt0,
                           g.tiles_m, g.tiles_n, tm, tn))
    return;
  const int m0 = tm * BM, n0 = tn * BN;

  // A share: row am, 4-k chunks ac0 and ac0 + 2; B share: rows bn and bn + 64, chunk bc
  // (ac0 is uniform within a wavefront: in a scalar register the per-k offsets are scalar too)
  const int am = tid & 127, ac0 = __builtin_amdgcn_readfirstlane(tid >> 7);
  const int bn = tid >> 2, bc = tid & 3;
  float ra[8];
  f4v rat[2];  // TA (A k-contiguous): shared like a k-contiguous B, rows bn and bn + 64, chunk bc
  f4v rb[2];
  // rows past the edge of M or N load the last row instead (their results are not stored).
  // B k-contiguous (TB false): 16-byte loads of 4 k; B n-contiguous (TB true, op(B) = B^T):
  // loaded like A, one dword per k with row n = am of the tile.
  const int a_voff = 4 * (min(m0 + am, M - 1) - m0);
  const int bt_voff = 4 * (min(n0 + am, N - 1) - n0);
  const int b_voff0 = 4 * (4 * bc + (min(n0 + bn, N - 1) - n0) * (int)ldb);
  const int b_voff1 = 4 * (4 * bc + (min(n0 + bn + 64, N - 1) - n0) * (int)ldb);
  // TA: k-contiguous rows addressed from two bases, rows m0.. and m1 = min(m0 + 64, M - 1).., so that a
  // byte offset spans at most 63 rows (the same for B in the TA instantiations)
  const int m1 = min(m0 + 64, M - 1), n1 = min(n0 + 64, N - 1);
  const int at_voff0 = 4 * (4 * bc + (min(m0 + bn, M - 1) - m0) * (int)lda);
  const int at_voff1 = 4 * (4 * bc + (min(m0 + bn + 64, M - 1) - m1) * (int)lda);
  const int b2_voff1 = 4 * (4 * bc + (min(n0 + bn + 64, N - 1) - n1) * (int)ldb);
  auto load_seg = [&](const float* Ab, const float* Bb, int k0) {
    if constexpr (TA) {
      const auto rA0 = make_rsrc(Ab + k0 + (long)m0 * lda);
      const auto rA1 = make_rsrc(Ab + k0 + (long)m1 * lda);
      rat[0] = __builtin_bit_cast(f4v, __builtin_amdgcn_raw_buffer_load_b128(rA0, at_voff0, 0, 0));
      rat[1] = __builtin_bit_cast(f4v, __builtin_amdgcn_raw_buffer_load_b128(rA1, at_voff1, 0, 0));
    } else {
      const auto rA = make_rsrc(Ab + m0 + (long)k0 * lda);
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int u = 0; u < 4; ++u)
          ra[4 * j + u] = bload<float>(rA, a_voff, 4 * (4 * (ac0 + 2 * j) + u) * (int)lda);
    }
    if constexpr (TB) {
      const auto rB = make_rsrc(Bb + n0 + (long)k0 * ldb);
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int u = 0; u < 4; ++u)
          rb[j][u] = bload<float>(rB, bt_voff, 4 * (4 * (ac0 + 2 * j) + u) * (int)ldb);
    } else if constexpr (TA) {
      const auto rB0 = make_rsrc(Bb + k0 + (long)n0 * ldb);
      const auto rB1 = make_rsrc(Bb + k0 + (long)n1 * ldb);
      rb[0] = __builtin_bit_cast(f4v, __builtin_amdgcn_raw_buffer_load_b128(rB0, b_voff0, 0, 0));
      rb[1] = __builtin_bit_cast(f4v, __builtin_amdgcn_raw_buffer_load_b128(rB1, b2_voff1, 0, 0));
    } else {
      const auto rB = make_rsrc(Bb + k0 + (long)n0 * ldb);
      rb[0] = __builtin_bit_cast(f4v, __builtin_amdgcn_raw_buffer_load_b128(rB, b_voff0, 0, 0));
      rb[1] = __builtin_bit_cast(f4v, __builtin_amdgcn_raw_buffer_load_b128(rB, b_voff1, 0, 0));
    }
  };
  // full slice s of the contraction: segment 0's K / BK slices, then segment 1's
  const int nk0 = K / BK;
  auto load = [&](int s) {
    if constexpr (SEG) {
      if (s < nk0) load_seg(A, B, s * BK);
      else load_seg(A1, B1, (s - nk0) * BK);
    } else {
      load_seg(A, B, s * BK);
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      if constexpr (TA) {
        const int m = bn + 64 * j;
        *reinterpret_cast<f4v*>(&As[buf][m * BK + 4 * g32_chunk(bc, m)]) = rat[j];
      } else {
        const int c = ac0 + 2 * j;
        *reinterpret_cast<f4v*>(&As[buf][am * BK + 4 * g32_chunk(c, am)]) =
            f4v{ra[4 * j], ra[4 * j + 1], ra[4 * j + 2], ra[4 * j + 3]};
      }
    }
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      if constexpr (TB) {
        const int c = ac0 + 2 * j;
        *reinterpret_cast<f4v*>(&Bs[buf][am * BK + 4 * g32_chunk(c, am)]) = rb[j];
      } else {
        const int n = bn + 64 * j;
        *reinterpret_cast<f4v*>(&Bs[buf][n * BK + 4 * g32_chunk(bc, n)]) = rb[j];
      }
    }
  };

  fvec4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = fvec4{0, 0, 0, 0};

  const int fr = lane & 15, fk = lane >> 4;
  // slices: the K tail (columns 16 floor(K/16) .. K-1, zero-padded) first when there is one, then the
  // full slices from k = 0; the tail is loaded once, in the prologue
  const int K16 = K / BK * BK, tail = K > K16 ? 1 : 0;
  const int nk = K16 / BK + tail + K1 / BK;
  if (tail) {
    if constexpr (TA) {
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int gk = K16 + 4 * bc + u;
          const int m = min(m0 + bn + 64 * j, M - 1);
          const float val = A[(long)min(gk, K - 1) + (long)m * lda];
          rat[j][u] = gk < K ? val : 0.f;
        }
    } else {
      const float* Ar = A + m0 + (a_voff >> 2);
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int gk = K16 + 4 * (ac0 + 2 * j) + u;
          const float val = Ar[(long)min(gk, K - 1) * lda];
          ra[4 * j + u] = gk < K ? val : 0.f;
        }
    }
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        float val;
        if constexpr (TB) {
          const int gk = K16 + 4 * (ac0 + 2 * j) + u;
          val = gk < K ? B[n0 + (bt_voff >> 2) + (long)min(gk, K - 1) * ldb] : 0.f;
        } else {
          const int gk = K16 + 4 * bc + u;
          const int n = min(n0 + bn + 64 * j, N - 1);
          val = gk < K ? B[(long)min(gk, K - 1) + (long)n * ldb] : 0.f;
        }
        rb[j][u] = val;
      }
  } else {
    load(0);
  }
  store(0);
  __syncthreads();
  auto compute = [&](int cur) {
    f4v bf[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int r = wn * 64 + j * 16 + fr;
      bf[j] = *reinterpret_cast<const f4v*>(&Bs[cur][r * BK + 4 * g32_chunk(fk, r)]);
    }
    f4v a0 = *reinterpret_cast<const f4v*>(&As[cur][(wm * 64 + fr) * BK + 4 * g32_chunk(fk, wm * 64 + fr)]);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      f4v a1;
      if (i + 1 < 4) {
        const int r = wm * 64 + (i + 1) * 16 + fr;
        a1 = *reinterpret_cast<const f4v*>(&As[cur][r * BK + 4 * g32_chunk(fk, r)]);
      }
#pragma unroll
      for (int s2 = 0; s2 < 4; ++s2)
#pragma unroll
        for (int j = 0; j < 4; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(bf[j][s2], a0[s2], acc[i][j], 0, 0, 0);
      if (i + 1 < 4) a0 = a1;
    }
  };
  // the last slice is peeled off, so that the loop body is one block without conditions (a schedule
  // that keeps the matrix pipe about 5% busier on the ring products than the conditional form)
  for (int kt = 0; kt + 1 < nk; ++kt) {
    const int cur = kt & 1;
    load(kt + 1 - tail);
    compute(cur);
    store(cur ^ 1);
    __syncthreads();
  }
  compute((nk - 1) & 1);

  if constexpr (OUT64) {
    double* __restrict__ C64 = g.C64 + (long)blockIdx.z * g.sC;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int m = m0 + wm * 64 + i * 16 + (lane & 15);
      if (m >= M) continue;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int n = n0 + wn * 64 + j * 16 + 4 * (lane >> 4) + r;
          if (n < N) C64[(long)m + (long)n * ldc] = (double)(alpha * acc[i][j][r]);
        }
      }
    }
    return;
  }
  if (beta != 0.f && m0 + BM <= M && n0 + BN <= N) {
    // interior tile accumulating into C: per 16-row block, all of the lane's C loads before
    // its stores (an element-wise load -> store round trip would serialise the epilogue)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const long m = m0 + wm * 64 + i * 16 + (lane & 15);
      float cv[4][4];
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) cv[j][r] = C[m + (long)(n0 + wn * 64 + j * 16 + 4 * (lane >> 4) + r) * ldc];
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r)
          C[m + (long)(n0 + wn * 64 + j * 16 + 4 * (lane >> 4) + r) * ldc] = alpha * acc[i][j][r] + beta * cv[j][r];
    }
    return;
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int m = m0 + wm * 64 + i * 16 + (lane & 15);
    if (m >= M) continue;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int n = n0 + wn * 64 + j * 16 + 4 * (lane >> 4) + r;
        if (n < N) {
          const long off = (long)m + (long)n * ldc;
          float v = alpha * acc[i][j][r];
          if (beta != 0.f) v += beta * C[off];
          C[off] = v;
        }
      }
    }
  }
}

// FP64 C = alpha A op(B) + beta C with A m-contiguous and B k-contiguous (NN) or
// n-contiguous (NT): 128x128 tile over 8 wavefronts of 64x32, so the accumulators take 64
// VGPRs and a lane stays within 128 -- two 512-thread workgroups (16 wavefronts) per CU.
// K slices are 16 doubles deep, held in LDS as 128-byte rows per m (A) or n (B) with the
// 16-byte chunk q of row r at chunk q ^ (r & 7); lane l supplies k = 4 (l / 16) + s at MFMA
// step s = 0..3 for both operands (two 16-byte fragment reads per 16-row block).  B's two
// blocks are held for the slice and A's four streamed one ahead.  Rows past M or N load a
// clamped row and are not stored; a K tail is the first slice, zero-padded, loaded in the prologue.
typedef double d2v __attribute__((ext_vector_type(2)));

__device__ inline int g64_chunk(int q, int r) { return q ^ (r & 7); }

template <int KIND, bool TB, bool TA = false, bool SEG = false>
__global__ __launch_bounds__(512, 4) void gemm64_kernel(GemmSide<double> p0, GemmSide<double> p1, int nt0,
                                                        double alpha, double beta) {
  constexpr int BM = 128, BN = 128, BK = 16;
  __shared__ __attribute__((aligned(16))) double As[2][BM * BK];
  __shared__ __attribute__((aligned(16))) double Bs[2][BN * BK];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;  // 2 x 4 wavefronts of 64 x 32
  const bool second = (int)blockIdx.x >= nt0;
  const GemmSide<double> g = second ? p1 : p0;
  const int M = g.M, N = g.N, K = (int)blockIdx.z == (int)gridDim.z - 1 ? g.Klast : g.K;
  const long lda = g.lda, ldb = g.ldb, ldc = g.ldc;
  const double* __restrict__ A = g.A + (long)blockIdx.z * g.sA;
  const double* __restrict__ B = g.B + (long)blockIdx.z * g.sB;
  double* __restrict__ C = g.C + (long)blockIdx.z * g.sC;
  // a second contraction segment (A1, B1 over K1 columns) only in the SEG instantiations
  const int K1 = SEG ? g.K1 : 0;
  const double* A1 = K1 > 0 ? g.A1 + (long)blockIdx.z * g.sA : A;
  const double* B1 = K1 > 0 ? g.B1 + (long)blockIdx.z * g.sB : B;
  int tm, tn;
  if (!tile_coords_bounded(second ? (int)blockIdx.x - nt0 : (int)blockIdx.x, second ? (int)gridDim.x - nt0 : nt0,
                           g.tiles_m, g.tiles_n, tm, tn))
    return;
  const int m0 = tm * BM, n0 = tn * BN;
  // A (and n-contiguous B) share: row am, k = 4 kq + u; k-contiguous B share: rows bn, bn + 64, chunk bc
  // (a k-contiguous A, TA, is shared like a k-contiguous B)
  const int am = tid & 127, kq = __builtin_amdgcn_readfirstlane(tid >> 7);
  const int bn = tid >> 3, bc = tid & 7;
  double ra[4];
  d2v rat[2];
  d2v rb[2];
  const int a_voff = 8 * (min(m0 + am, M - 1) - m0);
  const int bt_voff = 8 * (min(n0 + am, N - 1) - n0);
  const int b_voff0 = 8 * (2 * bc + (min(n0 + bn, N - 1) - n0) * (int)ldb);
  const int b_voff1 = 8 * (2 * bc + (min(n0 + bn + 64, N - 1) - n0) * (int)ldb);
  // TA: k-contiguous rows addressed from two bases, rows m0.. and m1 = min(m0 + 64, M - 1).., so that a
  // byte offset spans at most 63 rows (the same for B in the TA instantiations)
  const int m1 = min(m0 + 64, M - 1), n1 = min(n0 + 64, N - 1);
  const int at_voff0 = 8 * (2 * bc + (min(m0 + bn, M - 1) - m0) * (int)lda);
  const int at_voff1 = 8 * (2 * bc + (min(m0 + bn + 64, M - 1) - m1) * (int)lda);
  const int b2_voff1 = 8 * (2 * bc + (min(n0 + bn + 64, N - 1) - n1) * (int)ldb);
  auto load_seg = [&](const double* Ab, const double* Bb, int k0) {
    if constexpr (TA) {
      const auto rA0 = make_rsrc(Ab + k0 + (long)m0 * lda);
      const auto rA1 = make_rsrc(Ab + k0 + (long)m1 * lda);
      rat[0] = __builtin_bit_cast(d2v, __builtin_amdgcn_raw_buffer_load_b128(rA0, at_voff0, 0, 0));
      rat[1] = __builtin_bit_cast(d2v, __builtin_amdgcn_raw_buffer_load_b128(rA1, at_voff1, 0, 0));
    } else {
      const auto rA = make_rsrc(Ab + m0 + (long)k0 * lda);
#pragma unroll
      for (int u = 0; u < 4; ++u) ra[u] = bload<double>(rA, a_voff, 8 * (4 * kq + u) * (int)lda);
    }
    if constexpr (TB) {
      const auto rB = make_rsrc(Bb + n0 + (long)k0 * ldb);
#pragma unroll
      for (int u = 0; u < 4; ++u) rb[u >> 1][u & 1] = bload<double>(rB, bt_voff, 8 * (4 * kq + u) * (int)ldb);
    } else if constexpr (TA) {
      const auto rB0 = make_rsrc(Bb + k0 + (long)n0 * ldb);
      const auto rB1 = make_rsrc(Bb + k0 + (long)n1 * ldb);
      rb[0] = __builtin_bit_cast(d2v, __builtin_amdgcn_raw_buffer_load_b128(rB0, b_voff0, 0, 0));
      rb[1] = __builtin_bit_cast(d2v, __builtin_amdgcn_raw_buffer_load_b128(rB1, b2_voff1, 0, 0));
    } else {
      const auto rB = make_rsrc(Bb + k0 + (long)n0 * ldb);
      rb[0] = __builtin_bit_cast(d2v, __builtin_amdgcn_raw_buffer_load_b128(rB, b_voff0, 0, 0));
      rb[1] = __builtin_bit_cast(d2v, __builtin_amdgcn_raw_buffer_load_b128(rB, b_voff1, 0, 0));
    }
  };
  // full slice s of the contraction: segment 0's K / BK slices, then segment 1's
  const int nk0 = K / BK;
  auto load = [&](int s) {
    if constexpr (SEG) {
      if (s < nk0) load_seg(A, B, s * BK);
      else load_seg(A1, B1, (s - nk0) * BK);
    } else {
      load_seg(A, B, s * BK);
    }
  };
  auto store = [&](int buf) {
    if constexpr (TA) {
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int m = bn + 64 * j;
        *reinterpret_cast<d2v*>(&As[buf][m * BK + 2 * g64_chunk(bc, m)]) = rat[j];
      }
    } else {
      *reinterpret_cast<d2v*>(&As[buf][am * BK + 2 * g64_chunk(2 * kq, am)]) = d2v{ra[0], ra[1]};
      *reinterpret_cast<d2v*>(&As[buf][am * BK + 2 * g64_chunk(2 * kq + 1, am)]) = d2v{ra[2], ra[3]};
    }
    if constexpr (TB) {
      *reinterpret_cast<d2v*>(&Bs[buf][am * BK + 2 * g64_chunk(2 * kq, am)]) = rb[0];
      *reinterpret_cast<d2v*>(&Bs[buf][am * BK + 2 * g64_chunk(2 * kq + 1, am)]) = rb[1];
    } else {
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int n = bn + 64 * j;
        *reinterpret_cast<d2v*>(&Bs[buf][n * BK + 2 * g64_chunk(bc, n)]) = rb[j];
      }
    }
  };

  dvec4 acc[4][2];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = dvec4{0, 0, 0, 0};
  const int fr = lane & 15, fk = lane >> 4;
  // the K tail first (zero-padded, loaded once in the prologue), then the full slices from k = 0
  const int K16 = K / BK * BK, tail = K > K16 ? 1 : 0;
  const int nk = K16 / BK + tail + K1 / BK;
  if (tail) {
    if constexpr (TA) {
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int u = 0; u < 2; ++u) {
          const int gk = K16 + 2 * bc + u;
          const int m = min(m0 + bn + 64 * j, M - 1);
          const double val = A[(long)min(gk, K - 1) + (long)m * lda];
          rat[j][u] = gk < K ? val : 0.0;
        }
    } else {
      const double* Ar = A + m0 + (a_voff >> 3);
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int gk = K16 + 4 * kq + u;
        const double val = Ar[(long)min(gk, K - 1) * lda];
        ra[u] = gk < K ? val : 0.0;
      }
    }
    if constexpr (TB) {
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int gk = K16 + 4 * kq + u;
        const double val = B[n0 + (bt_voff >> 3) + (long)min(gk, K - 1) * ldb];
        rb[u >> 1][u & 1] = gk < K ? val : 0.0;
      }
    } else {
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int u = 0; u < 2; ++u) {
          const int gk = K16 + 2 * bc + u;
          const int n = min(n0 + bn + 64 * j, N - 1);
          const double val = B[(long)min(gk, K - 1) + (long)n * ldb];
          rb[j][u] = gk < K ? val : 0.0;
        }
    }
  } else {
    load(0);
  }
  store(0);
  __syncthreads();
  auto compute = [&](int cur) {
    d2v bf[2][2];
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int r = wn * 32 + j * 16 + fr;
      bf[j][0] = *reinterpret_cast<const d2v*>(&Bs[cur][r * BK + 2 * g64_chunk(2 * fk, r)]);
      bf[j][1] = *reinterpret_cast<const d2v*>(&Bs[cur][r * BK + 2 * g64_chunk(2 * fk + 1, r)]);
    }
    d2v a0[2], a1[2];
    {
      const int r = wm * 64 + fr;
      a0[0] = *reinterpret_cast<const d2v*>(&As[cur][r * BK + 2 * g64_chunk(2 * fk, r)]);
      a0[1] = *reinterpret_cast<const d2v*>(&As[cur][r * BK + 2 * g64_chunk(2 * fk + 1, r)]);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      if (i + 1 < 4) {
        const int r = wm * 64 + (i + 1) * 16 + fr;
        a1[0] = *reinterpret_cast<const d2v*>(&As[cur][r * BK + 2 * g64_chunk(2 * fk, r)]);
        a1[1] = *reinterpret_cast<const d2v*>(&As[cur][r * BK + 2 * g64_chunk(2 * fk + 1, r)]);
      }
#pragma unroll
      for (int s2 = 0; s2 < 4; ++s2)
#pragma unroll
        for (int j = 0; j < 2; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(bf[j][s2 >> 1][s2 & 1], a0[s2 >> 1][s2 & 1], acc[i][j],
                                                           0, 0, 0);
      if (i + 1 < 4) { a0[0] = a1[0]; a0[1] = a1[1]; }
    }
  };
  // the last slice is peeled off, so that the loop body is one block without conditions
  for (int kt = 0; kt + 1 < nk; ++kt) {
    const int cur = kt & 1;
    load(kt + 1 - tail);
    compute(cur);
    store(cur ^ 1);
    __syncthreads();
  }
  compute((nk - 1) & 1);

  // acc[i][j][r]: C(m, n) with m = lane % 16 of block i, n = lane / 16 + 4 r of block j
  if (beta != 0.0 && m0 + BM <= M && n0 + BN <= N) {
    // interior tile accumulating into C: per 16-row block, the loads before the stores
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const long m = m0 + wm * 64 + i * 16 + (lane & 15);
      double cv[2][4];
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) cv[j][r] = C[m + (long)(n0 + wn * 32 + j * 16 + (lane >> 4) + 4 * r) * ldc];
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r)
          C[m + (long)(n0 + wn * 32 + j * 16 + (lane >> 4) + 4 * r) * ldc] = alpha * acc[i][j][r] + beta * cv[j][r];
    }
    return;
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int m = m0 + wm * 64 + i * 16 + (lane & 15);
    if (m >= M) continue;
#pragma unroll
    for (int j = 0; j < 2; ++j) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int n = n0 + wn * 32 + j * 16 + (lane >> 4) + 4 * r;
        if (n < N) {
          const long off = (long)m + (long)n * ldc;
          double v = alpha * acc[i][j][r];
          if (beta != 0.0) v += beta * C[off];
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

// FP32 NN C = alpha A B + beta C for a short, wide product (M <= 64): a (16 IM)x256 tile over four
// wavefronts of (16 IM)x64 side by side (gemm32_nn_kernel's wavefront tile for IM = 4), four
// workgroups per CU.  K must be a multiple of 16.  Slices, swizzle and edge clamping as in
// gemm32_nn_kernel.
template <int IM>
__global__ __launch_bounds__(256, 4) void gemm32_sk_kernel(GemmSide<float> g, float alpha, float beta) {
  constexpr int BM = 16 * IM, BN = 256, BK = 16;
  __shared__ __attribute__((aligned(16))) float As[2][BM * BK];
  __shared__ __attribute__((aligned(16))) float Bs[2][BN * BK];
  const int tid = threadIdx.x, lane = tid & 63, wn = tid >> 6;
  const int M = g.M, N = g.N, K = g.K;
  const long lda = g.lda, ldb = g.ldb, ldc = g.ldc;
  const float* __restrict__ A = g.A + (long)blockIdx.z * g.sA;
  const float* __restrict__ B = g.B + (long)blockIdx.z * g.sB;
  float* __restrict__ C = g.C + (long)blockIdx.z * g.sC;
  int tm, tn;
  if (!tile_coords_bounded((int)blockIdx.x, (int)gridDim.x, g.tiles_m, g.tiles_n, tm, tn)) return;
  const int m0 = tm * BM, n0 = tn * BN;
  // A share: row am = lane (lanes below BM), 4-k chunk ac = wavefront; B share: rows bn + 64 j
  // (j < 4), 16-byte chunk bc
  const int am = lane, ac = __builtin_amdgcn_readfirstlane(wn);
  const bool a_share = lane < BM;
  const int bn = tid >> 2, bc = tid & 3;
  // A (at most 64 rows, L2-resident) is loaded one slice ahead.  B streams from HBM: for IM <= 3
  // it is loaded two slices ahead into two register stages, so that its loads have two slices of
  // MFMA work to arrive; IM = 4 has the registers for one stage.
  constexpr int NSB = IM <= 3 ? 2 : 1;
  float ra[4];
  f4v rb[NSB][4];
  const int a_voff = 4 * (min(m0 + am, M - 1) - m0);
  const int nlim = N - 1 - n0, ldbi = (int)ldb;
  const float* At = A + m0;
  const float* Bt = B + (long)n0 * ldb;
  auto load_a = [&](int k0) {
    if (a_share) {
      const auto rA = make_rsrc(At + (long)k0 * lda);
#pragma unroll
      for (int u = 0; u < 4; ++u) ra[u] = bload<float>(rA, a_voff, 4 * (4 * ac + u) * (int)lda);
    }
  };
  auto load_b = [&](int k0, f4v (&r)[4]) {
    const auto rB = make_rsrc(Bt + k0);
#pragma unroll
    for (int j = 0; j < 4; ++j)
      r[j] = __builtin_bit_cast(
          f4v, __builtin_amdgcn_raw_buffer_load_b128(rB, 4 * (4 * bc + min(bn + 64 * j, nlim) * ldbi), 0, 0));
  };
  auto store = [&](int buf, const f4v (&r)[4]) {
    if (a_share) *reinterpret_cast<f4v*>(&As[buf][am * BK + 4 * g32_chunk(ac, am)]) = f4v{ra[0], ra[1], ra[2], ra[3]};
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int n = bn + 64 * j;
      *reinterpret_cast<f4v*>(&Bs[buf][n * BK + 4 * g32_chunk(bc, n)]) = r[j];
    }
  };

  fvec4 acc[IM][4];
#pragma unroll
  for (int i = 0; i < IM; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = fvec4{0, 0, 0, 0};
  const int fr = lane & 15, fk = lane >> 4;
  const int nk = K / BK;
  auto compute = [&](int cur) {
    f4v bf[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int r = wn * 64 + j * 16 + fr;
      bf[j] = *reinterpret_cast<const f4v*>(&Bs[cur][r * BK + 4 * g32_chunk(fk, r)]);
    }
#pragma unroll
    for (int i = 0; i < IM; ++i) {
      const int r = i * 16 + fr;
      const f4v a0 = *reinterpret_cast<const f4v*>(&As[cur][r * BK + 4 * g32_chunk(fk, r)]);
#pragma unroll
      for (int s2 = 0; s2 < 4; ++s2)
#pragma unroll
        for (int j = 0; j < 4; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(bf[j][s2], a0[s2], acc[i][j], 0, 0, 0);
    }
  };
  if constexpr (NSB == 2) {
    // On entry to slice t (t even in the loop): LDS buffer 0 holds slice t, rb[1] holds slice
    // t + 1 (in flight) and rb[0] is free.  A is issued before B each time, so that waiting for A
    // leaves the newer B loads in flight.  Two slices per trip, so that the stages are named
    // statically; the last one to three slices are peeled off, so that the loop body has no
    // conditions.
    load_a(0);
    load_b(0, rb[0]);
    if (nk > 1) load_b(BK, rb[1]);
    store(0, rb[0]);
    __syncthreads();
    int t = 0;
    for (; t + 3 < nk; t += 2) {
      load_a((t + 1) * BK);
      load_b((t + 2) * BK, rb[0]);
      compute(0);
      store(1, rb[1]);
      __syncthreads();
      load_a((t + 2) * BK);
      load_b((t + 3) * BK, rb[1]);
      compute(1);
      store(0, rb[0]);
      __syncthreads();
    }
    if (t + 1 < nk) load_a((t + 1) * BK);
    if (t + 2 < nk) load_b((t + 2) * BK, rb[0]);
    compute(0);
    if (t + 1 < nk) {
      store(1, rb[1]);
      __syncthreads();
      if (t + 2 < nk) load_a((t + 2) * BK);
      compute(1);
      if (t + 2 < nk) {
        store(0, rb[0]);
        __syncthreads();
        compute(0);
      }
    }
  } else {
    load_a(0);
    load_b(0, rb[0]);
    store(0, rb[0]);
    __syncthreads();
    // the last slice is peeled off, so that the loop body is one block without conditions
    for (int kt = 0; kt + 1 < nk; ++kt) {
      const int cur = kt & 1;
      load_a((kt + 1) * BK);
      load_b((kt + 1) * BK, rb[0]);
      compute(cur);
      store(cur ^ 1, rb[0]);
      __syncthreads();
    }
    compute((nk - 1) & 1);
  }

#pragma unroll
  for (int i = 0; i < IM; ++i) {
    const int m = m0 + i * 16 + fr;
    if (m >= M) continue;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int n = n0 + wn * 64 + j * 16 + 4 * fk + r;
        if (n < N) {
          const long off = (long)m + (long)n * ldc;
          float v = alpha * acc[i][j][r];
          if (beta != 0.f) v += beta * C[off];
          C[off] = v;
        }
      }
    }
  }
}

// FP64 NN C = alpha A B + beta C for a short, wide product (M <= 64, e.g. the first quarter
// of the AO->MO transform, M = o): a (16 IM)x256 tile over eight wavefronts of (16 IM)x32 side by
// side (the FP64 kernel's wavefront tile for IM = 4), two workgroups per CU.  K must be a
// multiple of 16.  Slices, swizzle and edge clamping as in gemm64_kernel.
template <int IM>
__global__ __launch_bounds__(512, 4) void gemm64_sk_kernel(GemmSide<double> g, double alpha, double beta) {
  constexpr int BM = 16 * IM, BN = 256, BK = 16;
  __shared__ __attribute__((aligned(16))) double As[2][BM * BK];
  __shared__ __attribute__((aligned(16))) double Bs[2][BN * BK];
  const int tid = threadIdx.x, lane = tid & 63, wn = tid >> 6;
  const int M = g.M, N = g.N, K = g.K;
  const long lda = g.lda, ldb = g.ldb, ldc = g.ldc;
  const double* __restrict__ A = g.A + (long)blockIdx.z * g.sA;
  const double* __restrict__ B = g.B + (long)blockIdx.z * g.sB;
  double* __restrict__ C = g.C + (long)blockIdx.z * g.sC;
  int tm, tn;
  if (!tile_coords_bounded((int)blockIdx.x, (int)gridDim.x, g.tiles_m, g.tiles_n, tm, tn)) return;
  const int m0 = tm * BM, n0 = tn * BN;
  // A share: row am = lane (lanes below BM), 16-byte chunk ka = wavefront (k = 2 ka, 2 ka + 1);
  // B share: rows bn + 64 j, chunk bc
  const int am = lane, ka = __builtin_amdgcn_readfirstlane(wn);
  const bool a_share = lane < BM;
  const int bn = tid >> 3, bc = tid & 7;
  // A (at most 64 rows, L2-resident) is loaded one slice ahead.  B streams from HBM: for IM <= 3
  // it is loaded two slices ahead into two register stages, so that its loads have two slices of
  // MFMA work to arrive (K is short in these products); IM = 4 has the registers for one stage.
  constexpr int NSB = IM <= 3 ? 2 : 1;
  double ra[2];
  d2v rb[NSB][4];
  const int a_voff = 8 * (min(m0 + am, M - 1) - m0);
  // B rows past N load row N - 1 (uniform limit; the offsets are recomputed per load rather
  // than held in registers)
  const int nlim = N - 1 - n0, ldbi = (int)ldb;
  const double* At = A + m0;
  const double* Bt = B + (long)n0 * ldb;
  auto load_a = [&](int k0) {
    if (a_share) {
      const auto rA = make_rsrc(At + (long)k0 * lda);
#pragma unroll
      for (int u = 0; u < 2; ++u) ra[u] = bload<double>(rA, a_voff, 8 * (2 * ka + u) * (int)lda);
    }
  };
  auto load_b = [&](int k0, d2v (&r)[4]) {
    const auto rB = make_rsrc(Bt + k0);
#pragma unroll
    for (int j = 0; j < 4; ++j)
      r[j] = __builtin_bit_cast(
          d2v, __builtin_amdgcn_raw_buffer_load_b128(rB, 8 * (2 * bc + min(bn + 64 * j, nlim) * ldbi), 0, 0));
  };
  auto store = [&](int buf, const d2v (&r)[4]) {
    if (a_share) *reinterpret_cast<d2v*>(&As[buf][am * BK + 2 * g64_chunk(ka, am)]) = d2v{ra[0], ra[1]};
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int n = bn + 64 * j;
      *reinterpret_cast<d2v*>(&Bs[buf][n * BK + 2 * g64_chunk(bc, n)]) = r[j];
    }
  };

  dvec4 acc[IM][2];
#pragma unroll
  for (int i = 0; i < IM; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = dvec4{0, 0, 0, 0};
  const int fr = lane & 15, fk = lane >> 4;
  const int nk = K / BK;
  auto compute = [&](int cur) {
    d2v bf[2][2];
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int r = wn * 32 + j * 16 + fr;
      bf[j][0] = *reinterpret_cast<const d2v*>(&Bs[cur][r * BK + 2 * g64_chunk(2 * fk, r)]);
      bf[j][1] = *reinterpret_cast<const d2v*>(&Bs[cur][r * BK + 2 * g64_chunk(2 * fk + 1, r)]);
    }
    // (A's fragments are read as they are used: with four wavefronts per SIMD the LDS latency
    // is hidden, and the registers a one-ahead copy would take are not there)
#pragma unroll
    for (int i = 0; i < IM; ++i) {
      const int r = i * 16 + fr;
      d2v a0[2];
      a0[0] = *reinterpret_cast<const d2v*>(&As[cur][r * BK + 2 * g64_chunk(2 * fk, r)]);
      a0[1] = *reinterpret_cast<const d2v*>(&As[cur][r * BK + 2 * g64_chunk(2 * fk + 1, r)]);
#pragma unroll
      for (int s2 = 0; s2 < 4; ++s2)
#pragma unroll
        for (int j = 0; j < 2; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(bf[j][s2 >> 1][s2 & 1], a0[s2 >> 1][s2 & 1], acc[i][j],
                                                           0, 0, 0);
    }
  };
  if constexpr (NSB == 2) {
    // On entry to slice t (t even in the loop): LDS buffer 0 holds slice t, rb[1] holds slice
    // t + 1 (in flight) and rb[0] is free.  A is issued before B each time, so that waiting for A
    // leaves the newer B loads in flight.  Two slices per trip, so that the stages are named
    // statically; the last one to three slices are peeled off, so that the loop body has no
    // conditions.
    load_a(0);
    load_b(0, rb[0]);
    if (nk > 1) load_b(BK, rb[1]);
    store(0, rb[0]);
    __syncthreads();
    int t = 0;
    for (; t + 3 < nk; t += 2) {
      load_a((t + 1) * BK);
      load_b((t + 2) * BK, rb[0]);
      compute(0);
      store(1, rb[1]);
      __syncthreads();
      load_a((t + 2) * BK);
      load_b((t + 3) * BK, rb[1]);
      compute(1);
      store(0, rb[0]);
      __syncthreads();
    }
    if (t + 1 < nk) load_a((t + 1) * BK);
    if (t + 2 < nk) load_b((t + 2) * BK, rb[0]);
    compute(0);
    if (t + 1 < nk) {
      store(1, rb[1]);
      __syncthreads();
      if (t + 2 < nk) load_a((t + 2) * BK);
      compute(1);
      if (t + 2 < nk) {
        store(0, rb[0]);
        __syncthreads();
        compute(0);
      }
    }
  } else {
    load_a(0);
    load_b(0, rb[0]);
    store(0, rb[0]);
    __syncthreads();
    // the last slice is peeled off, so that the loop body is one block without conditions
    for (int kt = 0; kt + 1 < nk; ++kt) {
      const int cur = kt & 1;
      load_a((kt + 1) * BK);
      load_b((kt + 1) * BK, rb[0]);
      compute(cur);
      store(cur ^ 1, rb[0]);
      __syncthreads();
    }
    compute((nk - 1) & 1);
  }

  // acc[i][j][r]: C(m, n) with m = 16 i + lane % 16, n = n0 + 32 wn + 16 j + lane / 16 + 4 r
#pragma unroll
  for (int i = 0; i < IM; ++i) {
    const int m = m0 + i * 16 + fr;
    if (m >= M) continue;
#pragma unroll
    for (int j = 0; j < 2; ++j) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int n = n0 + wn * 32 + j * 16 + fk + 4 * r;
        if (n < N) {
          const long off = (long)m + (long)n * ldc;
          double v = alpha * acc[i][j][r];
          if (beta != 0.0) v += beta * C[off];
          C[off] = v;
        }
      }
    }
  }
}

// FP64 NN C = alpha A B + beta C for a tall, narrow product (N <= 64, M large, e.g. the MP2
// adjoint's g Co half-transform step, N = o): a 256x(16 IN) tile over eight wavefronts of
// 32x(16 IN) stacked in m, two workgroups per CU.  A (m-contiguous, the streamed operand) is
// staged as in gemm64_kernel, B (k-contiguous, at most 64 rows) as gemm64_sk_kernel's B.
// K must be a multiple of 16.
template <int IN>
__global__ __launch_bounds__(512, 4) void gemm64_tall_kernel(GemmSide<double> g, double alpha, double beta) {
  constexpr int BM = 256, BN = 16 * IN, BK = 16;
  __shared__ __attribute__((aligned(16))) double As[2][BM * BK];
  __shared__ __attribute__((aligned(16))) double Bs[2][BN * BK];
  const int tid = threadIdx.x, lane = tid & 63, wm = tid >> 6;
  const int M = g.M, N = g.N, K = g.K;
  const long lda = g.lda, ldb = g.ldb, ldc = g.ldc;
  const double* __restrict__ A = g.A + (long)blockIdx.z * g.sA;
  const double* __restrict__ B = g.B + (long)blockIdx.z * g.sB;
  double* __restrict__ C = g.C + (long)blockIdx.z * g.sC;
  int tm, tn;
  if (!tile_coords_bounded((int)blockIdx.x, (int)gridDim.x, g.tiles_m, g.tiles_n, tm, tn)) return;
  const int m0 = tm * BM, n0 = tn * BN;
  // A share: row am, k = 8 kq + u (u = 0..7); B share (threads below 8 BN): row bn, 16-byte chunk bc
  const int am = tid & 255, kq = __builtin_amdgcn_readfirstlane(tid >> 8);
  const int bn = tid >> 3, bc = tid & 7;
  const bool b_share = tid < 8 * BN;
  double ra[8];
  d2v rb;
  const int a_voff = 8 * (min(m0 + am, M - 1) - m0);
  const int b_voff = 8 * (2 * bc + (min(n0 + bn, N - 1) - n0) * (int)ldb);
  auto load = [&](int k0) {
    const auto rA = make_rsrc(A + m0 + (long)k0 * lda);
#pragma unroll
    for (int u = 0; u < 8; ++u) ra[u] = bload<double>(rA, a_voff, 8 * (8 * kq + u) * (int)lda);
    if (b_share) {
      const auto rB = make_rsrc(B + k0 + (long)n0 * ldb);
      rb = __builtin_bit_cast(d2v, __builtin_amdgcn_raw_buffer_load_b128(rB, b_voff, 0, 0));
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int q = 0; q < 4; ++q)
      *reinterpret_cast<d2v*>(&As[buf][am * BK + 2 * g64_chunk(4 * kq + q, am)]) = d2v{ra[2 * q], ra[2 * q + 1]};
    if (b_share) *reinterpret_cast<d2v*>(&Bs[buf][bn * BK + 2 * g64_chunk(bc, bn)]) = rb;
  };

  dvec4 acc[2][IN];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < IN; ++j) acc[i][j] = dvec4{0, 0, 0, 0};
  const int fr = lane & 15, fk = lane >> 4;
  const int nk = K / BK;
  load(0);
  store(0);
  __syncthreads();
  // the last slice is peeled off, so that the loop body is one block without conditions
  auto compute = [&](int cur) {
    // A's two blocks are held for the slice, B's blocks streamed one ahead of their MFMAs (for
    // IN = 4 read as they are used: the one-ahead copy would not fit in 128 VGPRs)
    constexpr bool ahead = IN < 4;
    d2v af[2][2], b0[2], b1[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const int r = wm * 32 + i * 16 + fr;
      af[i][0] = *reinterpret_cast<const d2v*>(&As[cur][r * BK + 2 * g64_chunk(2 * fk, r)]);
      af[i][1] = *reinterpret_cast<const d2v*>(&As[cur][r * BK + 2 * g64_chunk(2 * fk + 1, r)]);
    }
    if (ahead) {
      b0[0] = *reinterpret_cast<const d2v*>(&Bs[cur][fr * BK + 2 * g64_chunk(2 * fk, fr)]);
      b0[1] = *reinterpret_cast<const d2v*>(&Bs[cur][fr * BK + 2 * g64_chunk(2 * fk + 1, fr)]);
    }
#pragma unroll
    for (int j = 0; j < IN; ++j) {
      if (!ahead) {
        const int r = j * 16 + fr;
        b0[0] = *reinterpret_cast<const d2v*>(&Bs[cur][r * BK + 2 * g64_chunk(2 * fk, r)]);
        b0[1] = *reinterpret_cast<const d2v*>(&Bs[cur][r * BK + 2 * g64_chunk(2 * fk + 1, r)]);
      } else if (j + 1 < IN) {
        const int r = (j + 1) * 16 + fr;
        b1[0] = *reinterpret_cast<const d2v*>(&Bs[cur][r * BK + 2 * g64_chunk(2 * fk, r)]);
        b1[1] = *reinterpret_cast<const d2v*>(&Bs[cur][r * BK + 2 * g64_chunk(2 * fk + 1, r)]);
      }
#pragma unroll
      for (int s2 = 0; s2 < 4; ++s2)
#pragma unroll
        for (int i = 0; i < 2; ++i)
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(b0[s2 >> 1][s2 & 1], af[i][s2 >> 1][s2 & 1], acc[i][j],
                                                           0, 0, 0);
      if (ahead && j + 1 < IN) { b0[0] = b1[0]; b0[1] = b1[1]; }
    }
  };
  for (int kt = 0; kt + 1 < nk; ++kt) {
    const int cur = kt & 1;
    load((kt + 1) * BK);
    compute(cur);
    store(cur ^ 1);
    __syncthreads();
  }
  compute((nk - 1) & 1);

  // acc[i][j][r]: C(m, n) with m = m0 + 32 wm + 16 i + lane % 16, n = n0 + 16 j + lane / 16 + 4 r
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const int m = m0 + wm * 32 + i * 16 + fr;
    if (m >= M) continue;
#pragma unroll
    for (int j = 0; j < IN; ++j) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int n = n0 + j * 16 + fk + 4 * r;
        if (n < N) {
          const long off = (long)m + (long)n * ldc;
          double v = alpha * acc[i][j][r];
          if (beta != 0.0) v += beta * C[off];
          C[off] = v;
        }
      }
    }
  }
}

// FP32 NN C = alpha A B + beta C for a tall, narrow product (N <= 64, M large): a 256x(16 IN) tile over
// four wavefronts of 64x(16 IN) stacked in m, four workgroups per CU.  A (m-contiguous) is staged as in
// gemm32_nn_kernel, each thread holding one row's four 16-byte k-chunks; B (k-contiguous, at most 64
// rows) as gemm32_nn_kernel's B.  K must be a multiple of 16.
template <int IN>
__global__ __launch_bounds__(256, 4) void gemm32_tall_kernel(GemmSide<float> g, float alpha, float beta) {
  constexpr int BM = 256, BN = 16 * IN, BK = 16;
  __shared__ __attribute__((aligned(16))) float As[2][BM * BK];
  __shared__ __attribute__((aligned(16))) float Bs[2][BN * BK];
  const int tid = threadIdx.x, lane = tid & 63, wm = tid >> 6;
  const int M = g.M, N = g.N, K = g.K;
  const long lda = g.lda, ldb = g.ldb, ldc = g.ldc;
  const float* __restrict__ A = g.A + (long)blockIdx.z * g.sA;
  const float* __restrict__ B = g.B + (long)blockIdx.z * g.sB;
  float* __restrict__ C = g.C + (long)blockIdx.z * g.sC;
  int tm, tn;
  if (!tile_coords_bounded((int)blockIdx.x, (int)gridDim.x, g.tiles_m, g.tiles_n, tm, tn)) return;
  const int m0 = tm * BM, n0 = tn * BN;
  // A share: row am = tid, k = 4 c + u (c, u = 0..3); B share (threads below 4 BN): row bn, chunk bc
  const int am = tid;
  const int bn = tid >> 2, bc = tid & 3;
  const bool b_share = tid < 4 * BN;
  float ra[16];
  f4v rb;
  const int a_voff = 4 * (min(m0 + am, M - 1) - m0);
  const int b_voff = 4 * (4 * bc + (min(n0 + bn, N - 1) - n0) * (int)ldb);
  auto load = [&](int k0) {
    const auto rA = make_rsrc(A + m0 + (long)k0 * lda);
#pragma unroll
    for (int q = 0; q < 16; ++q) ra[q] = bload<float>(rA, a_voff, 4 * q * (int)lda);
    if (b_share) {
      const auto rB = make_rsrc(B + k0 + (long)n0 * ldb);
      rb = __builtin_bit_cast(f4v, __builtin_amdgcn_raw_buffer_load_b128(rB, b_voff, 0, 0));
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int c = 0; c < 4; ++c)
      *reinterpret_cast<f4v*>(&As[buf][am * BK + 4 * g32_chunk(c, am)]) =
          f4v{ra[4 * c], ra[4 * c + 1], ra[4 * c + 2], ra[4 * c + 3]};
    if (b_share) *reinterpret_cast<f4v*>(&Bs[buf][bn * BK + 4 * g32_chunk(bc, bn)]) = rb;
  };

  fvec4 acc[4][IN];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < IN; ++j) acc[i][j] = fvec4{0, 0, 0, 0};
  const int fr = lane & 15, fk = lane >> 4;
  const int nk = K / BK;
  load(0);
  store(0);
  __syncthreads();
  // B's blocks are held for the slice, A's streamed one ahead of their MFMAs
  auto compute = [&](int cur) {
    f4v bf[IN];
#pragma unroll
    for (int j = 0; j < IN; ++j) {
      const int r = j * 16 + fr;
      bf[j] = *reinterpret_cast<const f4v*>(&Bs[cur][r * BK + 4 * g32_chunk(fk, r)]);
    }
    f4v a0 = *reinterpret_cast<const f4v*>(&As[cur][(wm * 64 + fr) * BK + 4 * g32_chunk(fk, wm * 64 + fr)]);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      f4v a1;
      if (i + 1 < 4) {
        const int r = wm * 64 + (i + 1) * 16 + fr;
        a1 = *reinterpret_cast<const f4v*>(&As[cur][r * BK + 4 * g32_chunk(fk, r)]);
      }
#pragma unroll
      for (int s2 = 0; s2 < 4; ++s2)
#pragma unroll
        for (int j = 0; j < IN; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(bf[j][s2], a0[s2], acc[i][j], 0, 0, 0);
      if (i + 1 < 4) a0 = a1;
    }
  };
  // the last slice is peeled off, so that the loop body is one block without conditions
  for (int kt = 0; kt + 1 < nk; ++kt) {
    const int cur = kt & 1;
    load((kt + 1) * BK);
    compute(cur);
    store(cur ^ 1);
    __syncthreads();
  }
  compute((nk - 1) & 1);

  // acc[i][j][r]: C(m, n) with m = m0 + 64 wm + 16 i + lane % 16, n = n0 + 16 j + 4 (lane / 16) + r
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int m = m0 + wm * 64 + i * 16 + fr;
    if (m >= M) continue;
#pragma unroll
    for (int j = 0; j < IN; ++j) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int n = n0 + j * 16 + 4 * fk + r;
        if (n < N) {
          const long off = (long)m + (long)n * ldc;
          float v = alpha * acc[i][j][r];
          if (beta != 0.f) v += beta * C[off];
          C[off] = v;
        }
      }
    }
  }
}

static inline long tiles_m128(long x) { return (x + 127) / 128; }

// Whether a product can run on the high-occupancy kernels: 128x128 tiles, at least one full
// k-slice, and a tile's k-slice addressable with 32-bit byte offsets from one buffer base:
// an m-contiguous A (n-contiguous B) spans 16 columns of lda (ldb); a k-contiguous B spans 128 rows of
// ldb; in the TN / TT instantiations a k-contiguous A, and its k-contiguous B, span 64 rows of lda /
// ldb from each of two bases.
template <typename T>
static bool occ_kernel_ok(bool ta, bool tb, long K, long lda, long ldb) {
  if (K < 16) return false;
  const auto fits = [](long rows, long ld) { return (long)sizeof(T) * rows * (ld + 1) < (1L << 31); };
  const bool a_ok = ta ? fits(64, lda) : fits(16, lda);
  const bool b_ok = tb ? fits(16, ldb) : fits(ta ? 64 : 128, ldb);
  return a_ok && b_ok;
}

template <typename T>
static GemmSide<T> make_side(long M, long N, long K, const T* A, long lda, long sA, const T* B, long ldb, long sB,
                             T* C, long ldc, long sC) {
  return GemmSide<T>{A, B, C, lda, ldb, ldc, sA, sB, sC, (int)M, (int)N, (int)K, (int)tiles_m128(M),
                     (int)tiles_m128(N), (int)K, nullptr, nullptr, 0};
}

// One launch of the high-occupancy kernel over p0 (and p1 when `two`), nb batch entries.
template <typename T>
static void occ_launch(jues_ctx* ctx, bool tb, int kind, GemmSide<T> p0, GemmSide<T> p1, bool two, int nb, T alpha,
                       T beta, bool ta = false, bool seg = false) {
  const long t0 = (long)p0.tiles_m * p0.tiles_n;
  const long nt0 = two ? (t0 + NXCD - 1) / NXCD * NXCD : t0;
  const long total = two ? nt0 + (long)p1.tiles_m * p1.tiles_n : t0;
  JUES_REQUIRE(total < (1L << 31), "gemm: too many tiles");
  if (!two) p1 = p0;
  const dim3 grid((unsigned)total, 1, nb);
  if constexpr (sizeof(T) == 4) {
#define JUES_GEMM32(K_, TB_, TA_, SEG_)                                                                 \
  hipLaunchKernelGGL((gemm32_nn_kernel<K_, TB_, TA_, SEG_>), grid, dim3(NT), 0, ctx->stream, p0, p1, (int)nt0, \
                     (float)alpha, (float)beta)
    if (seg) {
      if (tb) JUES_GEMM32(0, true, false, true);
      else JUES_GEMM32(0, false, false, true);
    } else if (ta) {
      if (tb) JUES_GEMM32(0, true, true, false);
      else JUES_GEMM32(0, false, true, false);
    } else if (tb) JUES_GEMM32(0, true, false, false);
    else if (kind == 1) JUES_GEMM32(1, false, false, false);
    else if (kind == 2) JUES_GEMM32(2, false, false, false);
    else JUES_GEMM32(0, false, false, false);
#undef JUES_GEMM32
  } else {
#define JUES_GEMM64(K_, TB_, TA_, SEG_)                                                               \
  hipLaunchKernelGGL((gemm64_kernel<K_, TB_, TA_, SEG_>), grid, dim3(512), 0, ctx->stream, p0, p1, (int)nt0, \
                     (double)alpha, (double)beta)
    if (seg) {
      if (tb) JUES_GEMM64(0, true, false, true);
      else JUES_GEMM64(0, false, false, true);
    } else if (ta) {
      if (tb) JUES_GEMM64(0, true, true, false);
      else JUES_GEMM64(0, false, true, false);
    } else if (tb) JUES_GEMM64(0, true, false, false);
    else if (kind == 1) JUES_GEMM64(1, false, false, false);
    else if (kind == 2) JUES_GEMM64(2, false, false, false);
    else JUES_GEMM64(0, false, false, false);
#undef JUES_GEMM64
  }
  JUES_HIP(hipGetLastError());
}

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
  // tile shape: 0 = 128x128, 1 = 64x256 (skinny M), 2 = 256x64 (skinny N)
  // (the ring and ladder kernel symbols, KIND 1 and 2, exist in the 128x128 shape only: a skinny
  // product under those tags runs on the KIND 0 kernel of its shape)
  // (a skinny product with K <= 128 is bound by streaming op(B) and C: the 128x128 kernel, whose
  // extra MFMA work on the clamped rows costs nothing there, serves it)
  // (a batched short product, M <= 64 with at least 1024 workgroups of 256 columns -- 256 when K is
  // long -- such as the 40 x 216 blocks of the MP2 adjoint's second half-transform, takes the
  // skinny-M tile too: its 16-row blocks waste less of the MFMA work than a 128-row tile)
  const bool short_batched = M <= 64 && !ta && !tb && K > 128 && K % 16 == 0 && N >= 128 &&
                             (long)batch * ((N + 255) / 256) >= (K >= 2048 ? 256 : 1024);
  const int shape = (M <= 64 && N >= 8192 && (ta || K > 128)) || short_batched ? 1
                    : (N <= 64 && M >= 8192 && (ta || K > 128)) ? 2
                                                                 : 0;
  const long BM = shape == 1 ? 64 : shape == 2 ? 256 : 128, BN = shape == 1 ? 256 : shape == 2 ? 64 : 128;
  const long tiles_m = (M + BM - 1) / BM, tiles_n = (N + BN - 1) / BN;
  JUES_REQUIRE(tiles_m * tiles_n < (1L << 31), "gemm: too many tiles");
  // Few output tiles and a long K (the Fock-like terms): split K over workgroups,
  // write partial products to a workspace and reduce them in a fixed order.  The chunks fill
  // the chip's resident workgroup slots (four FP32 or two FP64 high-occupancy workgroups per CU).
  // A long-K product whose tiles fill their last round of slots poorly (under 80% overall, such as
  // the o^2 x v.m x v^2 products of RCCSD's (ia|bc) terms) is split into the fewest chunks that
  // fill the rounds, when the partial products fit in 2 GiB.
  const long slots = sizeof(T) == 4 ? 1024 : 512;
  const long tiles = tiles_m * tiles_n;
  auto round_fill = [&](long s) {
    const long t = tiles * s;
    return (double)t / (double)(((t + slots - 1) / slots) * slots);
  };
  long s_fill = 1;
  if (allow_split && batch == 1 && shape == 0 && tiles >= 128 && K >= 32768 && !no_alloc_active() &&
      round_fill(1) < 0.8) {
    for (long s = 2; s <= 16; ++s)
      if (round_fill(s) > round_fill(s_fill) + 0.02 && (double)s * M * N * sizeof(T) <= double(2L << 30)) s_fill = s;
  }
  if (allow_split && batch == 1 && (tiles < 128 || s_fill > 1) && K >= 4096 && !no_alloc_active()) {
    const long want = s_fill > 1 ? s_fill : std::max(2L, std::min(K / 1024, slots / tiles));
    const int BKs = sizeof(T) == 8 ? 16 : 32;
    long kc = (K + want - 1) / want;
    kc = (kc + BKs - 1) / BKs * BKs;
    const int S = (int)((K + kc - 1) / kc);
    T* Wk = static_cast<T*>(scratch(ctx, 9, sizeof(T) * (size_t)S * M * N));
    const long sa = ta ? kc : kc * lda, sb = tb ? kc * ldb : kc;
    const long klast = K - (long)(S - 1) * kc;
    if (shape == 0 && occ_kernel_ok<T>(ta, tb, std::min(kc, klast), lda, ldb)) {
      // all chunks in one launch, the last one shorter
      KTimer t1(ctx, name, 2.0 * M * N * K, (double)sizeof(T) * ((double)M * K + (double)K * N + (double)S * M * N));
      GemmSide<T> p = make_side<T>(M, N, kc, A, lda, sa, B, ldb, sb, Wk, M, M * N);
      p.Klast = (int)klast;
      occ_launch<T>(ctx, tb, kind, p, p, false, S, alpha, T(0), ta);
    } else {
      if (S > 1)
        gemm_impl<T>(ctx, ta, tb, M, N, kc, alpha, A, lda, sa, B, ldb, sb, T(0), Wk, M, M * N, S - 1, tag, false);
      gemm_impl<T>(ctx, ta, tb, M, N, klast, alpha, A + (long)(S - 1) * sa, lda, 0, B + (long)(S - 1) * sb, ldb, 0,
                   T(0), Wk + (long)(S - 1) * M * N, M, 0, 1, tag, false);
    }
    KTimer t(ctx, name, 0.0, (double)sizeof(T) * S * M * N);
    const long tot = M * N;
    const int nb = (int)std::max(1L, std::min((tot + 255) / 256, 65536L));
    hipLaunchKernelGGL((splitk_reduce_kernel<T>), dim3(nb), dim3(256), 0, ctx->stream, Wk, S, M, N, C, ldc, beta);
    JUES_HIP(hipGetLastError());
    return;
  }
  const double flops = 2.0 * M * N * K * batch;
  const double bytes = (double)sizeof(T) * ((double)M * K + (double)K * N + (beta != T(0) ? 2.0 : 1.0) * M * N) * batch;
  // Short, wide NN products (M <= 64): gemm32_sk_kernel / gemm64_sk_kernel, 64x256 tiles; a
  // transposed A (at most 64 x K, the transform's orbital block) is transposed first
  if (shape == 1 && ta && !tb && lda == K && batch == 1 && K % 16 == 0 && !no_alloc_active() &&
      (long)sizeof(T) * (256 + 64) * (std::max(M, ldb) + 1) < (1L << 31)) {
    T* At = static_cast<T*>(scratch(ctx, 18, sizeof(T) * (size_t)(M * K)));
    permute<T, T>(ctx, A, {K, M}, {1, 0}, At, 1.0, 0.0, name);
    gemm_impl<T>(ctx, false, tb, M, N, K, alpha, At, M, 0, B, ldb, sB, beta, C, ldc, sC, 1, tag, allow_split);
    return;
  }
  if constexpr (sizeof(T) == 4) {
    if (shape == 1 && !ta && !tb && K % 16 == 0 && 4L * (256 + 64) * (std::max(lda, ldb) + 1) < (1L << 31)) {
      KTimer timer(ctx, name, flops, bytes);
      for (int b0 = 0; b0 < batch; b0 += 65535) {
        const int nb = std::min(batch - b0, 65535);
        GemmSide<float> p = make_side<float>(M, N, K, reinterpret_cast<const float*>(A + (long)b0 * sA), lda, sA,
                                             reinterpret_cast<const float*>(B + (long)b0 * sB), ldb, sB,
                                             reinterpret_cast<float*>(C + (long)b0 * sC), ldc, sC);
        const int bm = M <= 48 ? 48 : 64;  // 16-row blocks: 3 or 4
        p.tiles_m = (int)((M + bm - 1) / bm);
        p.tiles_n = (int)((N + 255) / 256);
        const dim3 grid((unsigned)((long)p.tiles_m * p.tiles_n), 1, nb);
        if (bm == 48)
          hipLaunchKernelGGL(gemm32_sk_kernel<3>, grid, dim3(256), 0, ctx->stream, p, (float)alpha, (float)beta);
        else
          hipLaunchKernelGGL(gemm32_sk_kernel<4>, grid, dim3(256), 0, ctx->stream, p, (float)alpha, (float)beta);
        JUES_HIP(hipGetLastError());
      }
      return;
    }
  }
  if constexpr (sizeof(T) == 8) {
    if (shape == 1 && !ta && !tb && K % 16 == 0 && 8L * (256 + 64) * (std::max(lda, ldb) + 1) < (1L << 31)) {
      KTimer timer(ctx, name, flops, bytes);
      for (int b0 = 0; b0 < batch; b0 += 65535) {
        const int nb = std::min(batch - b0, 65535);
        GemmSide<double> p = make_side<double>(M, N, K, reinterpret_cast<const double*>(A + (long)b0 * sA), lda, sA,
                                               reinterpret_cast<const double*>(B + (long)b0 * sB), ldb, sB,
                                               reinterpret_cast<double*>(C + (long)b0 * sC), ldc, sC);
        const int bm = M <= 48 ? 48 : 64;  // 16-row blocks: 3 or 4
        p.tiles_m = (int)((M + bm - 1) / bm);
        p.tiles_n = (int)((N + 255) / 256);
        const dim3 grid((unsigned)((long)p.tiles_m * p.tiles_n), 1, nb);
        if (bm == 48)
          hipLaunchKernelGGL(gemm64_sk_kernel<3>, grid, dim3(512), 0, ctx->stream, p, (double)alpha, (double)beta);
        else
          hipLaunchKernelGGL(gemm64_sk_kernel<4>, grid, dim3(512), 0, ctx->stream, p, (double)alpha, (double)beta);
        JUES_HIP(hipGetLastError());
      }
      return;
    }
  }
  // Tall, narrow NN products (N <= 64): gemm64_tall_kernel / gemm32_tall_kernel, 256x48 or 256x64
  // tiles (a k-slice of A spans 16 lda, B's 64 rows 64 ldb, within 32-bit byte offsets)
  if constexpr (sizeof(T) == 4) {
    if (shape == 2 && !ta && !tb && K % 16 == 0 && 4L * 16 * (lda + 1) < (1L << 31) &&
        4L * 64 * (ldb + 1) < (1L << 31)) {
      KTimer timer(ctx, name, flops, bytes);
      for (int b0 = 0; b0 < batch; b0 += 65535) {
        const int nb = std::min(batch - b0, 65535);
        GemmSide<float> p = make_side<float>(M, N, K, reinterpret_cast<const float*>(A + (long)b0 * sA), lda, sA,
                                             reinterpret_cast<const float*>(B + (long)b0 * sB), ldb, sB,
                                             reinterpret_cast<float*>(C + (long)b0 * sC), ldc, sC);
        const int bn = N <= 48 ? 48 : 64;  // 16-column blocks: 3 or 4
        p.tiles_m = (int)((M + 255) / 256);
        p.tiles_n = (int)((N + bn - 1) / bn);
        const dim3 grid((unsigned)((long)p.tiles_m * p.tiles_n), 1, nb);
        if (bn == 48)
          hipLaunchKernelGGL(gemm32_tall_kernel<3>, grid, dim3(256), 0, ctx->stream, p, (float)alpha, (float)beta);
        else
          hipLaunchKernelGGL(gemm32_tall_kernel<4>, grid, dim3(256), 0, ctx->stream, p, (float)alpha, (float)beta);
        JUES_HIP(hipGetLastError());
      }
      return;
    }
  }
  if constexpr (sizeof(T) == 8) {
    if (shape == 2 && !ta && !tb && K % 16 == 0 && 8L * 16 * (lda + 1) < (1L << 31) &&
        8L * 64 * (ldb + 1) < (1L << 31)) {
      KTimer timer(ctx, name, flops, bytes);
      for (int b0 = 0; b0 < batch; b0 += 65535) {
        const int nb = std::min(batch - b0, 65535);
        GemmSide<double> p = make_side<double>(M, N, K, reinterpret_cast<const double*>(A + (long)b0 * sA), lda, sA,
                                               reinterpret_cast<const double*>(B + (long)b0 * sB), ldb, sB,
                                               reinterpret_cast<double*>(C + (long)b0 * sC), ldc, sC);
        const int bn = N <= 48 ? 48 : 64;  // 16-column blocks: 3 or 4
        p.tiles_m = (int)((M + 255) / 256);
        p.tiles_n = (int)((N + bn - 1) / bn);
        const dim3 grid((unsigned)((long)p.tiles_m * p.tiles_n), 1, nb);
        if (bn == 48)
          hipLaunchKernelGGL(gemm64_tall_kernel<3>, grid, dim3(512), 0, ctx->stream, p, (double)alpha, (double)beta);
        else
          hipLaunchKernelGGL(gemm64_tall_kernel<4>, grid, dim3(512), 0, ctx->stream, p, (double)alpha, (double)beta);
        JUES_HIP(hipGetLastError());
      }
      return;
    }
  }
  // FP32 NN and NT: gemm32_nn_kernel (four workgroups per CU); FP64: gemm64_kernel (two
  // 512-thread workgroups per CU)
  if (shape == 0 && occ_kernel_ok<T>(ta, tb, K, lda, ldb)) {
    KTimer timer(ctx, name, flops, bytes);
    for (int b0 = 0; b0 < batch; b0 += 65535) {
      const int nb = std::min(batch - b0, 65535);
      const GemmSide<T> p = make_side<T>(M, N, K, A + (long)b0 * sA, lda, sA, B + (long)b0 * sB, ldb, sB,
                                         C + (long)b0 * sC, ldc, sC);
      occ_launch<T>(ctx, tb, kind, p, p, false, nb, alpha, beta, ta);
    }
    return;
  }
  KTimer timer(ctx, name, flops, bytes);
  // the buffer-load path addresses a tile's k-slice with 32-bit byte offsets
  const long span = (long)sizeof(T) * (std::max<long>(BM, BN) + 64) * (std::max(lda, ldb) + 1);
  const int fast_ok = span < (1L << 31) ? 1 : 0;
  for (int b0 = 0; b0 < batch; b0 += 65535) {
    const int nb = std::min(batch - b0, 65535);
    dim3 grid((unsigned)(tiles_m * tiles_n), 1, nb);
    const T* Ab = A + (long)b0 * sA;
    const T* Bb = B + (long)b0 * sB;
    T* Cb = C + (long)b0 * sC;
#define JUES_GEMM_LAUNCH_S(TA_, TB_, K_, BM_, BN_, WM_)                                                  \
  hipLaunchKernelGGL((gemm_kernel<T, TA_, TB_, K_, BM_, BN_, WM_>), grid, dim3(NT), 0, ctx->stream, (int)M, (int)N, \
                     (int)K, alpha, Ab, lda, sA, Bb, ldb, sB, beta, Cb, ldc, sC, (int)tiles_m, (int)tiles_n, fast_ok)
#define JUES_GEMM_LAUNCH(TA_, TB_, K_)                                          \
  do {                                                                          \
    if (shape == 1) JUES_GEMM_LAUNCH_S(TA_, TB_, K_, 64, 256, 1);               \
    else if (shape == 2) JUES_GEMM_LAUNCH_S(TA_, TB_, K_, 256, 64, 4);          \
    else JUES_GEMM_LAUNCH_S(TA_, TB_, K_, 128, 128, 2);                         \
  } while (0)
    if (!ta && !tb) {
      if (kind == 1 && shape == 0) JUES_GEMM_LAUNCH_S(false, false, 1, 128, 128, 2);
      else if (kind == 2 && shape == 0) JUES_GEMM_LAUNCH_S(false, false, 2, 128, 128, 2);
      else JUES_GEMM_LAUNCH(false, false, 0);
    } else if (!ta && tb) JUES_GEMM_LAUNCH(false, true, 0);
    else if (ta && !tb) JUES_GEMM_LAUNCH(true, false, 0);
    else JUES_GEMM_LAUNCH(true, true, 0);
#undef JUES_GEMM_LAUNCH
#undef JUES_GEMM_LAUNCH_S
    JUES_HIP(hipGetLastError());
  }
}

template <typename T>
void gemm(jues_ctx* ctx, bool ta, bool tb, long M, long N, long K, T alpha, const T* A, long lda, long sA,
          const T* B, long ldb, long sB, T beta, T* C, long ldc, long sC, int batch, const char* tag) {
  gemm_impl<T>(ctx, ta, tb, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, batch, tag, true);
}

bool gemm32_out64(jues_ctx* ctx, long M, long N, long K, const float* A, long lda, const float* B, long ldb,
                  double* C, long ldc, const char* tag) {
  if (M <= 0 || N <= 0 || K <= 0 || !occ_kernel_ok<float>(false, false, K, lda, ldb)) return false;
  JUES_REQUIRE(lda >= M && ldb >= K && ldc >= M, "gemm32_out64: leading dimension too small");
  GemmSide<float> p = make_side<float>(M, N, K, A, lda, 0, B, ldb, 0, nullptr, ldc, 0);
  p.C64 = C;
  const long tiles = (long)p.tiles_m * p.tiles_n;
  JUES_REQUIRE(tiles < (1L << 31), "gemm: too many tiles");
  KTimer timer(ctx, tag ? tag : "gemm_f32", 2.0 * M * N * K, 4.0 * ((double)M * K + (double)K * N) + 8.0 * M * N);
  hipLaunchKernelGGL((gemm32_nn_kernel<0, false, false, false, true>), dim3((unsigned)tiles, 1, 1), dim3(NT), 0,
                     ctx->stream, p, p, (int)tiles, 1.0f, 0.0f);
  JUES_HIP(hipGetLastError());
  return true;
}

template <typename T>
void gemm_batch_sum(jues_ctx* ctx, bool ta, bool tb, long M, long N, long K, T alpha, const T* A, long lda, long sA,
                    const T* B, long ldb, long sB, T beta, T* C, long ldc, int batch, T* work, const char* tag) {
  JUES_REQUIRE(batch >= 1 && M > 0 && N > 0, "gemm_batch_sum: empty product");
  const char* name = tag ? tag : (sizeof(T) == 8 ? "gemm_f64" : "gemm_f32");
  gemm_impl<T>(ctx, ta, tb, M, N, K, alpha, A, lda, sA, B, ldb, sB, T(0), work, M, M * N, batch, tag, false);
  KTimer t(ctx, name, 0.0, (double)sizeof(T) * ((double)batch + (beta != T(0) ? 2.0 : 1.0)) * M * N);
  const long tot = M * N;
  hipLaunchKernelGGL((splitk_reduce_kernel<T>), dim3((unsigned)std::max(1L, std::min((tot + 255) / 256, 65536L))),
                     dim3(256), 0, ctx->stream, work, batch, M, N, C, ldc, beta);
  JUES_HIP(hipGetLastError());
}

template <typename T>
void gemm_pair(jues_ctx* ctx, bool tb, long M0, long N0, long K0, const T* A0, long lda0, const T* B0, long ldb0,
               T* C0, long ldc0, long M1, long N1, long K1, const T* A1, long lda1, const T* B1, long ldb1, T* C1,
               long ldc1, T alpha, T beta, const char* tag) {
  const bool skinny0 = (M0 <= 64 && N0 >= 8192) || (N0 <= 64 && M0 >= 8192);
  const bool skinny1 = (M1 <= 64 && N1 >= 8192) || (N1 <= 64 && M1 >= 8192);
  const bool empty0 = M0 <= 0 || N0 <= 0, empty1 = M1 <= 0 || N1 <= 0;
  if (empty0 || empty1 || skinny0 || skinny1 || !occ_kernel_ok<T>(false, tb, K0, lda0, ldb0) ||
      !occ_kernel_ok<T>(false, tb, K1, lda1, ldb1)) {
    gemm<T>(ctx, false, tb, M0, N0, K0, alpha, A0, lda0, 0, B0, ldb0, 0, beta, C0, ldc0, 0, 1, tag);
    gemm<T>(ctx, false, tb, M1, N1, K1, alpha, A1, lda1, 0, B1, ldb1, 0, beta, C1, ldc1, 0, 1, tag);
    return;
  }
  JUES_REQUIRE(M0 < (1L << 31) && N0 < (1L << 31) && K0 < (1L << 31) && M1 < (1L << 31) && N1 < (1L << 31) &&
                   K1 < (1L << 31),
               "gemm: dimension too large");
  JUES_REQUIRE(ldc0 >= M0 && ldc1 >= M1, "gemm: ldc < M");
  JUES_REQUIRE(lda0 >= M0 && lda1 >= M1, "gemm: lda too small");
  JUES_REQUIRE(ldb0 >= (tb ? N0 : K0) && ldb1 >= (tb ? N1 : K1), "gemm: ldb too small");
  const char* name = tag ? tag : (sizeof(T) == 8 ? "gemm_f64" : "gemm_f32");
  const int kind = !tag ? 0 : (std::strcmp(tag, "cc_ring") == 0 ? 1 : (std::strcmp(tag, "cc_pp_ladder") == 0 ? 2 : 0));
  const double sz = sizeof(T), bf = beta != T(0) ? 2.0 : 1.0;
  const double flops = 2.0 * M0 * N0 * K0 + 2.0 * M1 * N1 * K1;
  GemmSide<T> p0 = make_side<T>(M0, N0, K0, A0, lda0, 0, B0, ldb0, 0, C0, ldc0, 0);
  GemmSide<T> p1 = make_side<T>(M1, N1, K1, A1, lda1, 0, B1, ldb1, 0, C1, ldc1, 0);
  // Fewer tiles than about two rounds of resident workgroups (1024 FP32, 512 FP64) and long
  // K: split both K ranges into S chunks (one launch; the last chunk of each shorter), write
  // the partial products to a workspace and add them to C0 / C1 in a fixed order.
  const long tiles = (long)p0.tiles_m * p0.tiles_n + (long)p1.tiles_m * p1.tiles_n;
  const long slots = sizeof(T) == 4 ? 1024 : 512;
  const long S = std::min(std::min(K0, K1) / 1024, (2 * slots + tiles - 1) / tiles);
  if (tiles < slots && S >= 2) {
    const long kc0 = (K0 + S - 1) / S, kc1 = (K1 + S - 1) / S;
    const long n0 = M0 * N0, n1 = M1 * N1;
    T* Wk = static_cast<T*>(scratch(ctx, 16, sizeof(T) * (size_t)S * (n0 + n1)));
    {
      KTimer timer(ctx, name, flops, sz * ((double)M0 * K0 + (double)K0 * N0 + (double)M1 * K1 + (double)K1 * N1 +
                                           (double)S * (n0 + n1)));
      p0 = make_side<T>(M0, N0, kc0, A0, lda0, kc0 * lda0, B0, ldb0, tb ? kc0 * ldb0 : kc0, Wk, M0, n0);
      p1 = make_side<T>(M1, N1, kc1, A1, lda1, kc1 * lda1, B1, ldb1, tb ? kc1 * ldb1 : kc1, Wk + S * n0, M1, n1);
      p0.Klast = (int)(K0 - (S - 1) * kc0);
      p1.Klast = (int)(K1 - (S - 1) * kc1);
      occ_launch<T>(ctx, tb, kind, p0, p1, true, (int)S, alpha, T(0));
    }
    KTimer t(ctx, name, 0.0, sz * (S + bf) * (n0 + n1));
    hipLaunchKernelGGL((splitk_reduce_kernel<T>), dim3((unsigned)std::min((n0 + 255) / 256, 65536L)), dim3(256), 0,
                       ctx->stream, Wk, (int)S, M0, N0, C0, ldc0, beta);
    hipLaunchKernelGGL((splitk_reduce_kernel<T>), dim3((unsigned)std::min((n1 + 255) / 256, 65536L)), dim3(256), 0,
                       ctx->stream, Wk + S * n0, (int)S, M1, N1, C1, ldc1, beta);
    JUES_HIP(hipGetLastError());
    return;
  }
  const double bytes = sz * ((double)M0 * K0 + (double)K0 * N0 + bf * M0 * N0 + (double)M1 * K1 + (double)K1 * N1 +
                             bf * M1 * N1);
  KTimer timer(ctx, name, flops, bytes);
  occ_launch<T>(ctx, tb, kind, p0, p1, true, 1, alpha, beta);
}

template <typename T>
void gemm_ksum(jues_ctx* ctx, bool tb, long M, long N, long K0, const T* A0, const T* B0, long K1, const T* A1,
               const T* B1, long lda, long ldb, long sA, long sB, T alpha, T beta, T* C, long ldc, long sC, int batch,
               const char* tag) {
  const bool skinny = (M <= 64 && N >= 8192) || (N <= 64 && M >= 8192);
  if (M <= 0 || N <= 0 || batch <= 0) return;
  if (skinny || K0 < 16 || K1 < 16 || K0 % 16 != 0 || K1 % 16 != 0 || batch > 65535 ||
      !occ_kernel_ok<T>(false, tb, std::max(K0, K1), lda, ldb)) {
    gemm<T>(ctx, false, tb, M, N, K0, alpha, A0, lda, sA, B0, ldb, sB, beta, C, ldc, sC, batch, tag);
    gemm<T>(ctx, false, tb, M, N, K1, alpha, A1, lda, sA, B1, ldb, sB, T(1), C, ldc, sC, batch, tag);
    return;
  }
  JUES_REQUIRE(M < (1L << 31) && N < (1L << 31) && K0 + K1 < (1L << 31), "gemm: dimension too large");
  JUES_REQUIRE(ldc >= M && lda >= M && ldb >= (tb ? N : std::max(K0, K1)), "gemm: leading dimension too small");
  const char* name = tag ? tag : (sizeof(T) == 8 ? "gemm_f64" : "gemm_f32");
  const double K = (double)(K0 + K1);
  KTimer timer(ctx, name, 2.0 * M * N * K * batch,
               (double)sizeof(T) * ((double)M * K + K * N + (beta != T(0) ? 2.0 : 1.0) * M * N) * batch);
  GemmSide<T> p = make_side<T>(M, N, K0, A0, lda, sA, B0, ldb, sB, C, ldc, sC);
  p.A1 = A1;
  p.B1 = B1;
  p.K1 = (int)K1;
  occ_launch<T>(ctx, tb, 0, p, p, false, batch, alpha, beta, false, true);
}

template void gemm_ksum<double>(jues_ctx*, bool, long, long, long, const double*, const double*, long, const double*,
                                const double*, long, long, long, long, double, double, double*, long, long, int,
                                const char*);
template void gemm_ksum<float>(jues_ctx*, bool, long, long, long, const float*, const float*, long, const float*,
                               const float*, long, long, long, long, float, float, float*, long, long, int,
                               const char*);

template void gemm_pair<double>(jues_ctx*, bool, long, long, long, const double*, long, const double*, long, double*,
                                long, long, long, long, const double*, long, const double*, long, double*, long,
                                double, double, const char*);
template void gemm_pair<float>(jues_ctx*, bool, long, long, long, const float*, long, const float*, long, float*, long,
                               long, long, long, const float*, long, const float*, long, float*, long, float, float,
                               const char*);

template void gemm_batch_sum<double>(jues_ctx*, bool, bool, long, long, long, double, const double*, long, long,
                                     const double*, long, long, double, double*, long, int, double*, const char*);
template void gemm_batch_sum<float>(jues_ctx*, bool, bool, long, long, long, float, const float*, long, long,
                                    const float*, long, long, float, float*, long, int, float*, const char*);
template void gemm<double>(jues_ctx*, bool, bool, long, long, long, double, const double*, long, long,
                           const double*, long, long, double, double*, long, long, int, const char*);
template void gemm<float>(jues_ctx*, bool, bool, long, long, long, float, const float*, long, long,
                          const float*, long, long, float, float*, long, long, int, const char*);

}  // namespace jues
