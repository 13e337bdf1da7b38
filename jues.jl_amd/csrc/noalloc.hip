// Allocation-free RMP2 energy (jues_rmp2_noalloc; no reference counterpart): the FP64 factor
// route of jues_rmp2 for callers that cannot allocate device memory (code under HIP graph
// capture, serving loops, harnesses that forbid hipMalloc).
//
// Every device intermediate lives in `scratch`, one file-scope device array that each GPU
// the library runs on holds (at its own address) until the process ends.  Host code never
// passes the array's name to a kernel or a copy: jues_create looks up the array's device
// address once per device (hipGetSymbolAddress, with that device current) and every use
// goes through that address.  Calls on one GPU share the array and leave their contents in
// it: every region below is written before it is read, so nothing needs clearing, and
// calls that overlap on one GPU race.
//
// Layout (FP64, each region starting on a 256-byte boundary):
//   Co (n x o), Cv (n x v), eps (o + v), part (o v + 1 energy partials and the sum),
//   Y  (half-transformed factors, later L^T[(b,j),P] = L[(j,b),P]), L[(i,a),P],
//   slab K[i,a,(b,j)] = (ia|jb) for nj occupied j at a time (o v^2 per j).
// (ia|jb) and its exchange (ib|ja) share a slab, so the o^2 v^2 block is never formed.
#include "api_util.h"
#include "jues_internal.h"

#include <algorithm>
#include <mutex>

static __device__ __align__(256) unsigned char scratch[JUES_NOALLOC_SCRATCH_BYTES];

namespace jues {
namespace {

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[MAX_DEVICES] = {};  // device address of `scratch` on each device

// one block per column (b, jj) of the slab; part[j + o b] as jues_rmp2's per-(j,b) partials
__global__ __launch_bounds__(256) void mp2_slab_energy_kernel(const double* __restrict__ Ks,
                                                              const double* __restrict__ eps, int o, int v, int j0,
                                                              double* __restrict__ part) {
  __shared__ double sh[4];
  const int ov = o * v;
  const int col = blockIdx.x;
  const int b = col % v, jj = col / v, j = j0 + jj;
  const double ejb = eps[j] - eps[o + b];
  const double* K = Ks + (long)col * ov;           // K[., ., j, b]
  const double* Kj = Ks + (long)jj * v * ov;       // the slab columns of this j
  double s = 0.0;
  for (int ia = threadIdx.x; ia < ov; ia += 256) {
    const int i = ia % o, a = ia / o;
    const double k = K[ia];
    const double kx = Kj[(long)a * ov + i + o * b];  // (ib|ja): column (a, jj), row (i, b)
    s += k * (2.0 * k - kx) / (eps[i] - eps[o + a] + ejb);
  }
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) part[j + o * b] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

__global__ __launch_bounds__(256) void sum_partials_kernel(const double* __restrict__ part, int count,
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

size_t region(size_t doubles) { return (doubles * sizeof(double) + 255) / 256 * 256; }

}  // namespace

void noalloc_attach(int device) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  JUES_REQUIRE(device >= 0 && device < MAX_DEVICES, "allocation-free scratch: device index out of range");
  if (g_scratch_ptr[device]) return;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(::scratch));
  if (e != hipSuccess || !p)
    throw Error(JUES_HIP_ERROR, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_scratch_ptr[device] = p;
}

static void* scratch_address(int device) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  void* p = device >= 0 && device < MAX_DEVICES ? g_scratch_ptr[device] : nullptr;
  JUES_REQUIRE(p, "allocation-free scratch: no address for this device (create the context with jues_create)");
  return p;
}

}  // namespace jues

using namespace jues;

extern "C" jues_status jues_rmp2_noalloc(jues_ctx* ctx, int prec, const double* Cocc, int o, const double* Cvir,
                                         int v, int n, const double* eps, double* E) {
  JUES_API_BEGIN(ctx)
  JUES_REQUIRE(Cocc && Cvir && eps && E && o > 0 && v > 0, "jues_rmp2_noalloc: bad arguments");
  JUES_REQUIRE(prec == JUES_FP64, "jues_rmp2_noalloc: FP64 only");
  JUES_REQUIRE(n == ctx->n && n > 0, "jues_rmp2_noalloc: n does not match the AO integrals");
  JUES_REQUIRE(ctx->Bfac.p, "jues_rmp2_noalloc: needs AO factors (jues_set_ao_factors or jues_set_df_integrals)");
  const long naux = ctx->naux, ov = (long)o * v;
  const size_t rCo = region((size_t)n * o), rCv = region((size_t)n * v), rE = region((size_t)o + v),
               rP = region((size_t)ov + 1),
               rY = region((size_t)std::max<long>((long)n * std::min(o, v), ov) * naux), rL = region((size_t)ov * naux),
               rSlab = region((size_t)ov * v);
  const size_t fixed = rCo + rCv + rE + rP + rY + rL;
  if (fixed + rSlab > JUES_NOALLOC_SCRATCH_BYTES)
    throw Error(JUES_BAD_ARG, "jues_rmp2_noalloc: n=" + std::to_string(n) + " o=" + std::to_string(o) +
                                  " v=" + std::to_string(v) + " naux=" + std::to_string(naux) + " needs " +
                                  std::to_string(fixed + rSlab) + " bytes of scratch; the maximum is " +
                                  std::to_string((size_t)JUES_NOALLOC_SCRATCH_BYTES));
  const int nj = (int)std::min<size_t>((size_t)o, (JUES_NOALLOC_SCRATCH_BYTES - fixed) / region((size_t)ov * v));
  char* base = static_cast<char*>(scratch_address(ctx->device));
  double* Co = reinterpret_cast<double*>(base);
  double* Cv = reinterpret_cast<double*>(base + rCo);
  double* ep = reinterpret_cast<double*>(base + rCo + rCv);
  double* part = reinterpret_cast<double*>(base + rCo + rCv + rE);
  double* Y = reinterpret_cast<double*>(base + rCo + rCv + rE + rP);
  double* L = reinterpret_cast<double*>(base + rCo + rCv + rE + rP + rY);
  double* Ks = reinterpret_cast<double*>(base + fixed);
  // no timer events and no context scratch while the call runs: any device allocation raises
  const bool timing = ctx->timing;
  ctx->timing = false;
  try {
    NoAllocScope guard;
    h2d(ctx, Co, Cocc, (size_t)n * o * sizeof(double));
    h2d(ctx, Cv, Cvir, (size_t)n * v * sizeof(double));
    h2d(ctx, ep, eps, ((size_t)o + v) * sizeof(double));
    // L[i,a,P] = (Co^T B_P Cv)[i,a] (the half-transform in Y), then L^T[(b,j),P] into Y
    mo_factors_dev<double>(ctx, Co, o, Cv, v, L, Y);
    double* Lt = Y;
    permute<double, double>(ctx, L, {o, v, naux}, {1, 0, 2}, Lt, 1.0, 0.0, "mp2_noalloc");
    for (int j0 = 0; j0 < o; j0 += nj) {
      const int cnt = std::min(nj, o - j0);
      // K[(i,a),(b,jj)] = sum_P L[(i,a),P] L[(j0+jj,b),P]
      gemm<double>(ctx, false, true, ov, (long)v * cnt, naux, 1.0, L, ov, 0, Lt + (long)v * j0, ov, 0, 0.0, Ks, ov,
                   0, 1, "xform_factor_gemm");
      hipLaunchKernelGGL(mp2_slab_energy_kernel, dim3((unsigned)(v * cnt)), dim3(256), 0, ctx->stream, Ks, ep, o, v,
                         j0, part);
      JUES_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(256), 0, ctx->stream, part, (int)ov, part + ov);
    JUES_HIP(hipGetLastError());
    d2h(ctx, E, part + ov, sizeof(double));
  } catch (...) {
    ctx->timing = timing;
    throw;
  }
  ctx->timing = timing;
  if (!std::isfinite(*E)) throw Error(JUES_NONFINITE, "jues_rmp2_noalloc: non-finite energy (a zero denominator?)");
  JUES_API_END(ctx)
}
