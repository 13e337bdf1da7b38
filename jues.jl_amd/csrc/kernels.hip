// Bandwidth-bound data movement: tiled N-d permutation with scaling/accumulation,
// conversions, fills, axpby and deterministic FP64 reductions.
#include "jues_internal.h"

#include <algorithm>
#include <cstring>

namespace jues {
namespace {

constexpr int PT = 32;       // permutation tile edge
constexpr int PMAXD = 6;

struct PermDesc {
  int nd;                    // number of dimensions (>= 2 after normalisation)
  long d[PMAXD];             // input dims
  long istr[PMAXD];          // input strides
  long ostr[PMAXD];          // output stride of each input dim
  int a, b;                  // tiled dims: a = input-fastest (0), b = output-fastest or dim 1
  int outer[PMAXD];          // remaining dims
  int nouter;
  long tiles_a, tiles_b;
};

// One 256-thread block moves a 32x32 tile spanned by input dims a (stride 1 in
// the input) and b; the tile passes through LDS so that the writes run along
// whichever of a/b is fastest in the output.
template <typename Ti, typename To>
__global__ __launch_bounds__(256) void permute_kernel(const Ti* __restrict__ in, To* __restrict__ out,
                                                      PermDesc P, double alpha, double beta,
                                                      int out_fast_is_a) {
  __shared__ double tile[PT][PT + 1];
  long bid = blockIdx.x;
  const long ta = bid % P.tiles_a;
  bid /= P.tiles_a;
  const long tb = bid % P.tiles_b;
  bid /= P.tiles_b;
  long ibase = 0, obase = 0;
  for (int q = 0; q < P.nouter; ++q) {
    const int dd = P.outer[q];
    const long x = bid % P.d[dd];
    bid /= P.d[dd];
    ibase += x * P.istr[dd];
    obase += x * P.ostr[dd];
  }
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  const long a0 = ta * PT, b0 = tb * PT;
  const long na = P.d[P.a], nb = P.d[P.b];
#pragma unroll
  for (int r = 0; r < PT; r += 8) {
    const long ia = a0 + tx, ib = b0 + ty + r;
    double v = 0.0;
    if (ia < na && ib < nb) v = (double)in[ibase + ia * P.istr[P.a] + ib * P.istr[P.b]];
    tile[ty + r][tx] = v;
  }
  __syncthreads();
  if (out_fast_is_a) {
#pragma unroll
    for (int r = 0; r < PT; r += 8) {
      const long ia = a0 + tx, ib = b0 + ty + r;
      if (ia < na && ib < nb) {
        const long off = obase + ia * P.ostr[P.a] + ib * P.ostr[P.b];
        double v = alpha * tile[ty + r][tx];
        if (beta != 0.0) v += beta * (double)out[off];
        out[off] = (To)v;
      }
    }
  } else {
#pragma unroll
    for (int r = 0; r < PT; r += 8) {
      const long ib = b0 + tx, ia = a0 + ty + r;
      if (ia < na && ib < nb) {
        const long off = obase + ia * P.ostr[P.a] + ib * P.ostr[P.b];
        double v = alpha * tile[tx][ty + r];
        if (beta != 0.0) v += beta * (double)out[off];
        out[off] = (To)v;
      }
    }
  }
}

template <typename Ti, typename To>
__global__ void convert_kernel(const Ti* __restrict__ in, To* __restrict__ out, long n) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x)
    out[e] = (To)in[e];
}

template <typename T>
__global__ void fill_kernel(T* __restrict__ out, long n, T value) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x)
    out[e] = value;
}

template <typename T>
__global__ void axpby_kernel(long n, double a, const T* __restrict__ x, double b, T* __restrict__ y) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x)
    y[e] = (T)(a * (double)x[e] + (b == 0.0 ? 0.0 : b * (double)y[e]));
}

constexpr int RED_BLOCKS = 1024;

__device__ inline double block_sum(double v, double* sh) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s += sh[i];
  return s;  // valid in thread 0
}

template <typename Ta, typename Tb>
__global__ __launch_bounds__(256) void dot_partial_kernel(const Ta* __restrict__ x, const Tb* __restrict__ y,
                                                          long n, double* __restrict__ part) {
  __shared__ double sh[4];
  double s = 0.0;
  for (long e = blockIdx.x * 256L + threadIdx.x; e < n; e += (long)gridDim.x * 256L)
    s += (double)x[e] * (double)y[e];
  s = block_sum(s, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ __launch_bounds__(256) void maxabs_partial_kernel(const double* __restrict__ x, long n,
                                                             double* __restrict__ part) {
  __shared__ double sh[256];
  double s = 0.0;
  for (long e = blockIdx.x * 256L + threadIdx.x; e < n; e += (long)gridDim.x * 256L) s = fmax(s, fabs(x[e]));
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// Deterministic fixed-order reduction of `count` partial sums into out[0].
__global__ __launch_bounds__(256) void reduce_partials_kernel(const double* __restrict__ part, int count,
                                                              double* __restrict__ out, int mode) {
  __shared__ double sh[256];
  double s = 0.0;
  for (int i = threadIdx.x; i < count; i += 256) s = mode ? fmax(s, part[i]) : s + part[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      sh[threadIdx.x] = mode ? fmax(sh[threadIdx.x], sh[threadIdx.x + w]) : sh[threadIdx.x] + sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = sh[0];
}

int grid_for(long n) { return (int)std::max(1L, std::min((n + 255) / 256, 65536L)); }

}  // namespace

template <typename Ti, typename To>
void permute(jues_ctx* ctx, const Ti* in, const std::vector<long>& in_dims, const std::vector<int>& perm,
             To* out, double alpha, double beta, const char* tag) {
  const int nd0 = (int)in_dims.size();
  JUES_REQUIRE(nd0 >= 1 && nd0 <= PMAXD && (int)perm.size() == nd0, "permute: bad rank");
  long total = 1;
  for (long d : in_dims) total *= d;
  if (total == 0) return;
  // drop unit dimensions (keeps the permutation of the others)
  std::vector<long> d;
  std::vector<int> map(nd0, -1);
  for (int j = 0; j < nd0; ++j)
    if (in_dims[j] != 1) { map[j] = (int)d.size(); d.push_back(in_dims[j]); }
  std::vector<int> p;
  for (int k = 0; k < nd0; ++k)
    if (map[perm[k]] >= 0) p.push_back(map[perm[k]]);
  while (d.size() < 2) { d.push_back(1); p.push_back((int)p.size()); }
  PermDesc P{};
  P.nd = (int)d.size();
  long s = 1;
  for (int j = 0; j < P.nd; ++j) { P.d[j] = d[j]; P.istr[j] = s; s *= d[j]; }
  s = 1;
  for (int k = 0; k < P.nd; ++k) { P.ostr[p[k]] = s; s *= d[p[k]]; }
  P.a = 0;
  P.b = p[0] != 0 ? p[0] : 1;
  P.nouter = 0;
  for (int j = 0; j < P.nd; ++j)
    if (j != P.a && j != P.b) P.outer[P.nouter++] = j;
  P.tiles_a = (P.d[P.a] + PT - 1) / PT;
  P.tiles_b = (P.d[P.b] + PT - 1) / PT;
  long nblocks = P.tiles_a * P.tiles_b;
  for (int q = 0; q < P.nouter; ++q) nblocks *= P.d[P.outer[q]];
  const int out_fast_is_a = (p[0] == 0) ? 1 : 0;
  KTimer t(ctx, tag ? tag : "permute", 0.0,
           (double)total * (sizeof(Ti) + sizeof(To) * (beta != 0.0 ? 2 : 1)));
  hipLaunchKernelGGL((permute_kernel<Ti, To>), dim3((unsigned)nblocks), dim3(256), 0, ctx->stream, in, out, P,
                     alpha, beta, out_fast_is_a);
  JUES_HIP(hipGetLastError());
}

template <typename Ti, typename To>
void convert(jues_ctx* ctx, const Ti* in, To* out, long n) {
  if (n <= 0) return;
  KTimer t(ctx, "convert", 0.0, (double)n * (sizeof(Ti) + sizeof(To)));
  hipLaunchKernelGGL((convert_kernel<Ti, To>), dim3(grid_for(n)), dim3(256), 0, ctx->stream, in, out, n);
  JUES_HIP(hipGetLastError());
}

template <typename T>
void fill(jues_ctx* ctx, T* out, long n, double value) {
  if (n <= 0) return;
  hipLaunchKernelGGL((fill_kernel<T>), dim3(grid_for(n)), dim3(256), 0, ctx->stream, out, n, (T)value);
  JUES_HIP(hipGetLastError());
}

template <typename T>
void axpby(jues_ctx* ctx, long n, double a, const T* x, double b, T* y) {
  if (n <= 0) return;
  KTimer t(ctx, "axpby", 0.0, (double)n * sizeof(T) * 3);
  hipLaunchKernelGGL((axpby_kernel<T>), dim3(grid_for(n)), dim3(256), 0, ctx->stream, n, a, x, b, y);
  JUES_HIP(hipGetLastError());
}

template <typename Ta, typename Tb>
double dot(jues_ctx* ctx, const Ta* x, const Tb* y, long n) {
  double* part = static_cast<double*>(scratch(ctx, 15, (RED_BLOCKS + 1) * sizeof(double)));
  const int nb = (int)std::min<long>(RED_BLOCKS, std::max(1L, (n + 255) / 256));
  hipLaunchKernelGGL((dot_partial_kernel<Ta, Tb>), dim3(nb), dim3(256), 0, ctx->stream, x, y, n, part);
  hipLaunchKernelGGL(reduce_partials_kernel, dim3(1), dim3(256), 0, ctx->stream, part, nb, part + RED_BLOCKS, 0);
  JUES_HIP(hipGetLastError());
  double r = 0.0;
  d2h(ctx, &r, part + RED_BLOCKS, sizeof(double));
  return r;
}

double max_abs(jues_ctx* ctx, const double* x, long n) {
  double* part = static_cast<double*>(scratch(ctx, 15, (RED_BLOCKS + 1) * sizeof(double)));
  const int nb = (int)std::min<long>(RED_BLOCKS, std::max(1L, (n + 255) / 256));
  hipLaunchKernelGGL(maxabs_partial_kernel, dim3(nb), dim3(256), 0, ctx->stream, x, n, part);
  hipLaunchKernelGGL(reduce_partials_kernel, dim3(1), dim3(256), 0, ctx->stream, part, nb, part + RED_BLOCKS, 1);
  JUES_HIP(hipGetLastError());
  double r = 0.0;
  d2h(ctx, &r, part + RED_BLOCKS, sizeof(double));
  return r;
}

#define JUES_PERM_INST(TI, TO)                                                                          \
  template void permute<TI, TO>(jues_ctx*, const TI*, const std::vector<long>&, const std::vector<int>&, \
                                TO*, double, double, const char*);
JUES_PERM_INST(double, double)
JUES_PERM_INST(float, float)
JUES_PERM_INST(double, float)
JUES_PERM_INST(float, double)
template void convert<double, float>(jues_ctx*, const double*, float*, long);
template void convert<float, double>(jues_ctx*, const float*, double*, long);
template void convert<double, double>(jues_ctx*, const double*, double*, long);
template void fill<double>(jues_ctx*, double*, long, double);
template void fill<float>(jues_ctx*, float*, long, double);
template void axpby<double>(jues_ctx*, long, double, const double*, double, double*);
template void axpby<float>(jues_ctx*, long, double, const float*, double, float*);
template double dot<double, double>(jues_ctx*, const double*, const double*, long);
template double dot<float, float>(jues_ctx*, const float*, const float*, long);
template double dot<double, float>(jues_ctx*, const double*, const float*, long);

}  // namespace jues
