// Perturbative triples (T) on the device: PerturbativeTriples.compute_pT
// (CoupledCluster/PerturbativeTriples.jl:35-138) and its call after the Fock-driven CCSD
// (AutoRCCSD.jl:294-300).
//
// For each occupied triple i >= j >= k the v^3 intermediate W[a,b,c]
// (PerturbativeTriples.jl:105-111) is assembled in HBM by six MFMA GEMMs over
// K = v + o: each pairs one Vvvvo term (K = v) with the Vvooo term (K = o) that
// has the same output permutation and the same occupied pairs.  The operand
// copies (PTOperands) are arranged so that every GEMM is either plain or batched
// over c with its output in W[a,b,c] order (a fastest): W is accumulated in place
// with coalesced stores and no permutation pass.
//
// The energy kernel then visits ordered 8x8x8 virtual blocks A >= B >= C.  A
// workgroup stages the six permuted blocks of W (and V = W + T1 x Vvovo terms,
// PerturbativeTriples.jl:112) in LDS, evaluates the bracket of :118-125 for
// a >= b >= c with weight 1/(1 + d_ab + d_bc) and divides by D (:121).  Every
// triple's sum is reduced in a fixed order, so the result is bitwise reproducible.
#include "jues_internal.h"
#include "api_util.h"

#include <algorithm>
#include <cstdio>

namespace jues {
namespace {

constexpr int TE = 8;  // energy tile edge

// (p q | r s) = K[p*ko1 + q*kv1 + r*ko2 + s*kv2] for occupied p, r and virtual q, s
struct KView {
  const double* p;
  long ko1, kv1, ko2, kv2;
};

// Kc[q,r,p,s] = (p q | s r) (v x v block per occupied pair), t1t[a,i] = T1[i,a]
template <typename T>
__global__ __launch_bounds__(512) void pt_energy_kernel(const T* __restrict__ W, int v,
                                                        const double* __restrict__ Kc,
                                                        const double* __restrict__ t1t, int o,
                                                        const double* __restrict__ fo,
                                                        const double* __restrict__ fv, int i, int j, int k,
                                                        const int* __restrict__ blocks,
                                                        double* __restrict__ part) {
  __shared__ double Ws[6][512];
  __shared__ double Vs[6][512];
  __shared__ double red[8];
  const int t = threadIdx.x;
  const int x = t & 7, y = (t >> 3) & 7, z = t >> 6;
  const int A = blocks[3 * blockIdx.x], B = blocks[3 * blockIdx.x + 1], C = blocks[3 * blockIdx.x + 2];
  // tiles in the order (A,B,C) (A,C,B) (B,A,C) (B,C,A) (C,A,B) (C,B,A)
  const int P1[6] = {A, A, B, B, C, C}, P2[6] = {B, C, A, C, A, B}, P3[6] = {C, B, C, A, B, A};
  const long vv = (long)v * v;
  const double* Kjk = Kc + (j + (long)o * k) * vv;  // (j q | k r) at q + v r
  const double* Kik = Kc + (i + (long)o * k) * vv;
  const double* Kij = Kc + (i + (long)o * j) * vv;
  const double* t1i = t1t + (long)v * i;
  const double* t1j = t1t + (long)v * j;
  const double* t1k = t1t + (long)v * k;
#pragma unroll
  for (int s = 0; s < 6; ++s) {
    const int p = TE * P1[s] + x, q = TE * P2[s] + y, r = TE * P3[s] + z;
    double w = 0.0, vval = 0.0;
    if (p < v && q < v && r < v) {
      w = (double)W[p + (long)v * q + vv * r];
      vval = w + Kjk[q + (long)v * r] * t1i[p] + Kik[p + (long)v * r] * t1j[q] + Kij[p + (long)v * q] * t1k[r];
    }
    Ws[s][t] = w;
    Vs[s][t] = vval;
  }
  __syncthreads();
  const int a = TE * A + x, b = TE * B + y, c = TE * C + z;
  double e = 0.0;
  if (a < v && b < v && c < v && a >= b && b >= c) {
    // element (a,b,c) and its five permutations, read from the tile that holds each
    const int l0 = x + 8 * y + 64 * z;   // W[a,b,c]  in (A,B,C)
    const int l1 = x + 8 * z + 64 * y;   // W[a,c,b]  in (A,C,B)
    const int l2 = y + 8 * x + 64 * z;   // W[b,a,c]  in (B,A,C)
    const int l3 = y + 8 * z + 64 * x;   // W[b,c,a]  in (B,C,A)
    const int l4 = z + 8 * x + 64 * y;   // W[c,a,b]  in (C,A,B)
    const int l5 = z + 8 * y + 64 * x;   // W[c,b,a]  in (C,B,A)
    const double wabc = Ws[0][l0], wacb = Ws[1][l1], wbac = Ws[2][l2], wbca = Ws[3][l3], wcab = Ws[4][l4],
                 wcba = Ws[5][l5];
    const double vabc = Vs[0][l0], vacb = Vs[1][l1], vbac = Vs[2][l2], vbca = Vs[3][l3], vcab = Vs[4][l4],
                 vcba = Vs[5][l5];
    const double X = wabc * vabc + wacb * vacb + wbac * vbac + wbca * vbca + wcab * vcab + wcba * vcba;
    const double Y = vabc + vbca + vcab;
    const double Z = vacb + vbac + vcba;
    const double E = (Y - 2.0 * Z) * (wabc + wbca + wcab) + (Z - 2.0 * Y) * (wacb + wbac + wcba) + 3.0 * X;
    const double D = fo[i] + fo[j] + fo[k] - fv[a] - fv[b] - fv[c];
    e = E / (D * (1.0 + (a == b) + (b == c)));
  }
  for (int off = 32; off > 0; off >>= 1) e += __shfl_down(e, off, 64);
  if ((t & 63) == 0) red[t >> 6] = e;
  __syncthreads();
  if (t == 0) {
    double s = 0.0;
    for (int w = 0; w < 8; ++w) s += red[w];
    part[blockIdx.x] = s;
  }
}

// et[0] = weight * sum(part[0..n)), summed in a fixed order by one workgroup
__global__ __launch_bounds__(256) void pt_reduce_kernel(const double* __restrict__ part, int n, double weight,
                                                        double* __restrict__ et) {
  __shared__ double sh[256];
  double s = 0.0;
  for (int q = threadIdx.x; q < n; q += 256) s += part[q];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) et[0] = weight * sh[0];
}

// Operands of the (T) kernels in precision T (FP64 or FP32); K, T1 and the orbital
// energies stay FP64.  Each K = o term of W shares its output layout and its pair
// indices with one K = v term, so the integral and amplitude copies carry both,
// side by side along the contraction index d' in [0, v + o):
//   Pe [a,b,d',i] = Vvvvo[b,d,a,i] (d' = d < v),   -T2[i,l,a,b] (d' = v + l)
//   P2e[a,b,d',i] = Vvvvo[a,d,b,i],                 -T2[i,l,b,a]
//   Te [c,d',k,j] = T2[k,j,c,d],                     Vvooo[c,k,j,l]
// with d' padded by zero columns to a multiple of 16 (pt_ve), so that the two GEMMs with
// one output layout run as one product over two contraction segments.
template <typename T>
struct PTOperands {
  DevBuf Pe, P2e, Te, W;
  DevBuf Kc, t1t;  // compact (p q | s r) blocks and T1 transposed, FP64
  const double* fo = nullptr;
  const double* fv = nullptr;
};

// Kc[q,r,p,s] = (p q | s r) from any 4-index layout described by a KView, and t1t = T1^T.
template <typename T>
void pt_pack_k(jues_ctx* ctx, PTOperands<T>& S, const KView& K, const double* t1, int o, int v) {
  // the view's strides name the layout: (ia|jb) with i, a, j, b strides ko1, kv1, ko2, kv2
  std::vector<long> dims(4);
  std::vector<int> perm(4);
  // order the four axes by stride to recover the stored layout, then permute to [q, r, p, s]
  struct Ax { long stride; long dim; int role; };  // role 0 = p, 1 = q, 2 = s (second occ), 3 = r (second vir)
  std::vector<Ax> ax = {{K.ko1, o, 0}, {K.kv1, v, 1}, {K.ko2, o, 2}, {K.kv2, v, 3}};
  std::sort(ax.begin(), ax.end(), [](const Ax& x, const Ax& y) { return x.stride < y.stride; });
  for (int d = 0; d < 4; ++d) dims[d] = ax[d].dim;
  const int want[4] = {1, 3, 0, 2};  // output axes: q, r, p, s
  for (int d = 0; d < 4; ++d)
    for (int e = 0; e < 4; ++e)
      if (ax[e].role == want[d]) perm[d] = e;
  S.Kc.alloc(sizeof(double) * (size_t)o * o * v * v);
  permute<double, double>(ctx, K.p, dims, perm, S.Kc.template as<double>(), 1.0, 0.0, "pt_prep");
  S.t1t.alloc(sizeof(double) * (size_t)o * v);
  permute<double, double>(ctx, t1, {o, v}, {1, 0}, S.t1t.template as<double>(), 1.0, 0.0, "pt_prep");
}

// dst rows of `width` bytes every `dpitch` bytes <- src rows every `spitch` bytes
static void copy2d(jues_ctx* ctx, void* dst, size_t dpitch, const void* src, size_t spitch, size_t width,
                   size_t rows) {
  JUES_HIP(hipMemcpy2DAsync(dst, dpitch, src, spitch, width, rows, hipMemcpyDeviceToDevice, ctx->stream));
}

// contraction length of the packed (T) operands: v + o rounded up to a multiple of 16
static inline long pt_ve(int o, int v) { return ((long)v + o + 15) / 16 * 16; }

// W[a,b,c] (PerturbativeTriples.jl:105-111) for one triple (i, j, k): six terms, each a
// Vvvvo term plus the Vvooo term with the same output layout, as three products of two
// contraction segments each (W written once and updated twice).
template <typename T>
void pt_build_W(jues_ctx* ctx, const PTOperands<T>& S, int o, int v, int i, int j, int k) {
  const long vv = (long)v * v, ve = pt_ve(o, v), blkP = vv * ve, blkT = (long)v * ve;
  const T* Pe = S.Pe.template as<T>();
  const T* P2e = S.P2e.template as<T>();
  const T* Te = S.Te.template as<T>();
  T* W = S.W.template as<T>();
  auto P = [&](int x) { return Pe + x * blkP; };
  auto P2 = [&](int x) { return P2e + x * blkP; };
  auto Tq = [&](int p, int q) { return Te + (long)(p + (long)o * q) * blkT; };
  const T one = T(1);
  // ijk abc + jik bac: W[(a,b),c] = Pe_i[(a,b),d'] Te_kj[c,d'] + P2e_j[(a,b),d'] Te_ki[c,d']
  gemm_ksum<T>(ctx, true, vv, v, ve, P(i), Tq(k, j), ve, P2(j), Tq(k, i), vv, v, 0, 0, one, T(0), W, vv, 0, 1, "pt_w");
  // ikj acb + kij cab: W[a,b,c] += Pe_i[a,c,d'] Te_jk[b,d'] + P2e_k[a,c,d'] Te_ji[b,d']  (batched over c)
  gemm_ksum<T>(ctx, true, v, v, ve, P(i), Tq(j, k), ve, P2(k), Tq(j, i), vv, v, v, 0, one, one, W, v, vv, v, "pt_w");
  // kji cba + jki bca: W[a,(b,c)] += Te_ij[a,d'] P2e_k[(b,c),d'] + Te_ik[a,d'] Pe_j[(b,c),d']
  gemm_ksum<T>(ctx, true, v, vv, ve, Tq(i, j), P2(k), ve, Tq(i, k), P(j), v, vv, 0, 0, one, one, W, v, 0, 1, "pt_w");
}

// The amplitude columns of Pe/P2e and the amplitude half of Te, from T2 (device FP64).
template <typename T>
void pt_pack_amplitudes(jues_ctx* ctx, PTOperands<T>& S, const double* T2d, int o, int v) {
  const long vv = (long)v * v, ve = pt_ve(o, v);
  const size_t es = sizeof(T);
  DevBuf tmp(es * (size_t)o * o * vv);
  // -T2[i,l,a,b] as [a,b,l,i] -> Pe columns v..v+o
  permute<double, T>(ctx, T2d, {o, o, v, v}, {2, 3, 1, 0}, tmp.as<T>(), -1.0, 0.0, "pt_prep");
  copy2d(ctx, S.Pe.template as<T>() + vv * v, es * vv * ve, tmp.p, es * vv * o, es * vv * o, (size_t)o);
  // -T2[i,l,b,a] as [a,b,l,i] -> P2e columns v..v+o
  permute<double, T>(ctx, T2d, {o, o, v, v}, {3, 2, 1, 0}, tmp.as<T>(), -1.0, 0.0, "pt_prep");
  copy2d(ctx, S.P2e.template as<T>() + vv * v, es * vv * ve, tmp.p, es * vv * o, es * vv * o, (size_t)o);
  // T2[k,j,c,d] as [c,d,k,j] -> Te columns 0..v
  permute<double, T>(ctx, T2d, {o, o, v, v}, {2, 3, 0, 1}, tmp.as<T>(), 1.0, 0.0, "pt_prep");
  copy2d(ctx, S.Te.template as<T>(), es * v * ve, tmp.p, es * vv, es * vv, (size_t)o * o);
  sync(ctx);
}

template <typename T>
double pt_run(jues_ctx* ctx, PTOperands<T>& S, const double* T2d, int o, int v, int verbose) {
  const long vv = (long)v * v;
  pt_pack_amplitudes<T>(ctx, S, T2d, o, v);
  S.W.alloc(sizeof(T) * (size_t)vv * v);
  // ordered virtual blocks A >= B >= C
  const int nbk = (v + TE - 1) / TE;
  std::vector<int> hb;
  for (int A = 0; A < nbk; ++A)
    for (int B = 0; B <= A; ++B)
      for (int C = 0; C <= B; ++C) { hb.push_back(A); hb.push_back(B); hb.push_back(C); }
  const int nblk = (int)(hb.size() / 3);
  DevBuf blocks(hb.size() * sizeof(int));
  h2d(ctx, blocks.p, hb.data(), hb.size() * sizeof(int));
  const long ntrip = (long)o * (o + 1) * (o + 2) / 6;
  DevBuf part(sizeof(double) * (size_t)nblk), et(sizeof(double) * (size_t)ntrip);
  long tix = 0;
  for (int i = 0; i < o; ++i) {
    if (verbose) {  // PerturbativeTriples.jl:81
      sync(ctx);
      std::fprintf(stderr, "  (T) occupied orbital %d of %d\n", i + 1, o);
    }
    for (int j = 0; j <= i; ++j)
      for (int k = 0; k <= j; ++k, ++tix) {
        pt_build_W<T>(ctx, S, o, v, i, j, k);
        const double weight = 2.0 - (i == j) - (j == k);
        KTimer tm(ctx, "pt_energy", 0.0, (double)sizeof(T) * vv * v);
        hipLaunchKernelGGL((pt_energy_kernel<T>), dim3(nblk), dim3(512), 0, ctx->stream, S.W.template as<T>(), v,
                           S.Kc.template as<double>(), S.t1t.template as<double>(), o, S.fo, S.fv, i, j, k,
                           blocks.as<int>(), part.as<double>());
        hipLaunchKernelGGL(pt_reduce_kernel, dim3(1), dim3(256), 0, ctx->stream, part.as<double>(), nblk, weight,
                           et.as<double>() + tix);
        JUES_HIP(hipGetLastError());
      }
  }
  std::vector<double> het((size_t)ntrip);
  d2h(ctx, het.data(), et.p, het.size() * sizeof(double));
  double Et = 0.0;
  for (double e : het) Et += e;
  return Et;
}

// Integral columns of Pe/P2e (from P[a,b,d,i] / P2[a,b,d,i], written by `make` into a
// v^3 o scratch) and the Vvooo half of Te (from Vq[c,l,k,j] = Vvooo[c,k,j,l]).
template <typename T, typename MakeP>
void pt_pack_integrals(jues_ctx* ctx, PTOperands<T>& S, int o, int v, MakeP make, const T* Vq) {
  const long vv = (long)v * v, ve = pt_ve(o, v);
  const size_t es = sizeof(T);
  S.Pe.alloc(es * (size_t)vv * ve * o);
  S.P2e.alloc(es * (size_t)vv * ve * o);
  S.Te.alloc(es * (size_t)v * ve * o * o);
  if (ve > (long)v + o) {  // the padding columns d' >= v + o stay zero
    fill<T>(ctx, S.Pe.template as<T>(), vv * ve * o, 0.0);
    fill<T>(ctx, S.P2e.template as<T>(), vv * ve * o, 0.0);
    fill<T>(ctx, S.Te.template as<T>(), (long)v * ve * o * o, 0.0);
  }
  {
    DevBuf tmp(es * (size_t)vv * v * o);
    make(tmp.as<T>(), 0);
    copy2d(ctx, S.Pe.p, es * vv * ve, tmp.p, es * vv * v, es * vv * v, (size_t)o);
    make(tmp.as<T>(), 1);
    copy2d(ctx, S.P2e.p, es * vv * ve, tmp.p, es * vv * v, es * vv * v, (size_t)o);
    sync(ctx);
  }
  copy2d(ctx, S.Te.template as<T>() + vv, es * v * ve, Vq, es * v * o, es * v * o, (size_t)o * o);
  sync(ctx);
}

template <typename T>
double pt_from_arrays(jues_ctx* ctx, const double* T1, const double* T2, const double* Vvvvo, const double* Vvooo,
                      const double* Vvovo, const double* fo, const double* fv, int o, int v, int verbose) {
  PTOperands<T> S;
  DeviceArray<double> t1(ctx, T1, (size_t)o * v), t2(ctx, T2, (size_t)o * o * v * v),
      k4(ctx, Vvovo, (size_t)v * o * v * o), efo(ctx, fo, (size_t)o), efv(ctx, fv, (size_t)v);
  {
    DeviceArray<double> vvvo(ctx, Vvvvo, (size_t)o * v * v * v), vooo(ctx, Vvooo, (size_t)v * o * o * o);
    DevBuf vq(sizeof(T) * (size_t)v * o * o * o);
    // Vq[c,l,k,j] = Vvooo[c,k,j,l]
    permute<double, T>(ctx, vooo.p, {v, o, o, o}, {0, 3, 1, 2}, vq.as<T>(), 1.0, 0.0, "pt_prep");
    // P[a,b,d,i] = Vvvvo[b,d,a,i],  P2[a,b,d,i] = Vvvvo[a,d,b,i]
    pt_pack_integrals<T>(ctx, S, o, v, [&](T* dst, int which) {
      permute<double, T>(ctx, vvvo.p, {v, v, v, o}, which == 0 ? std::vector<int>{2, 0, 1, 3}
                                                                : std::vector<int>{0, 2, 1, 3},
                         dst, 1.0, 0.0, "pt_prep");
    }, vq.as<T>());
  }
  // Vvovo[a,i,b,j] = (ia|jb)
  pt_pack_k<T>(ctx, S, KView{k4.p, (long)v, 1L, (long)v * o * v, (long)v * o}, t1.p, o, v);
  S.fo = efo.p;
  S.fv = efv.p;
  return pt_run<T>(ctx, S, t2.p, o, v, verbose);
}

template <typename T>
double pt_from_orbitals(jues_ctx* ctx, const double* Co, int o, const double* Cv, int v, const double* eps,
                        const double* T1, const double* T2, int verbose) {
  PTOperands<T> S;
  const int n = ctx->n;
  DeviceArray<double> co(ctx, Co, (size_t)n * o), cv(ctx, Cv, (size_t)n * v), e(ctx, eps, (size_t)o + v),
      t1(ctx, T1, (size_t)o * v), t2(ctx, T2, (size_t)o * o * v * v);
  DevBuf ovov(sizeof(double) * (size_t)o * v * o * v);
  tei_transform_dev<double>(ctx, co.p, o, cv.p, v, co.p, o, cv.p, v, ovov.as<double>());   // (ia|jb) [i,a,j,b]
  {
    DevBuf ovvv(sizeof(double) * (size_t)o * v * v * v), ooov(sizeof(double) * (size_t)o * o * o * v);
    DevBuf vq(sizeof(T) * (size_t)v * o * o * o);
    tei_transform_dev<double>(ctx, co.p, o, cv.p, v, cv.p, v, cv.p, v, ovvv.as<double>());   // (ia|bc)
    tei_transform_dev<double>(ctx, co.p, o, co.p, o, co.p, o, cv.p, v, ooov.as<double>());   // (jk|ia)
    // Vvooo[c,k,j,l] = (jl|kc) = ooov[j,l,k,c]:  Vq[c,l,k,j] = ooov[j,l,k,c]
    permute<double, T>(ctx, ooov.as<double>(), {o, o, o, v}, {3, 1, 2, 0}, vq.as<T>(), 1.0, 0.0, "pt_prep");
    // Vvvvo[b,d,a,i] = (ia|db):  P[a,b,d,i] = ovvv[i,a,d,b],  P2[a,b,d,i] = ovvv[i,b,d,a]
    pt_pack_integrals<T>(ctx, S, o, v, [&](T* dst, int which) {
      permute<double, T>(ctx, ovvv.as<double>(), {o, v, v, v}, which == 0 ? std::vector<int>{1, 3, 2, 0}
                                                                         : std::vector<int>{3, 1, 2, 0},
                         dst, 1.0, 0.0, "pt_prep");
    }, vq.as<T>());
  }
  pt_pack_k<T>(ctx, S, KView{ovov.as<double>(), 1L, (long)o, (long)o * v, (long)o * v * o}, t1.p, o, v);
  S.fo = e.p;
  S.fv = e.p + o;
  return pt_run<T>(ctx, S, t2.p, o, v, verbose);
}

}  // namespace
}  // namespace jues

using namespace jues;

extern "C" {

jues_status jues_compute_pT(jues_ctx* ctx, int prec, const double* T1, const double* T2, const double* Vvvvo,
                            const double* Vvooo, const double* Vvovo, const double* fo, const double* fv, int o,
                            int v, int verbose, double* Et) {
  JUES_API_BEGIN(ctx)
  JUES_REQUIRE(T1 && T2 && Vvvvo && Vvooo && Vvovo && fo && fv && Et, "jues_compute_pT: null pointer");
  JUES_REQUIRE(o > 0 && v > 0, "jues_compute_pT: bad dimensions");
  JUES_REQUIRE(prec == JUES_FP64 || prec == JUES_FP32_MIXED, "jues_compute_pT: precision must be FP64 or FP32_MIXED");
  *Et = prec == JUES_FP64 ? pt_from_arrays<double>(ctx, T1, T2, Vvvvo, Vvooo, Vvovo, fo, fv, o, v, verbose)
                          : pt_from_arrays<float>(ctx, T1, T2, Vvvvo, Vvooo, Vvovo, fo, fv, o, v, verbose);
  if (!std::isfinite(*Et)) throw Error(JUES_NONFINITE, "jues_compute_pT: non-finite energy");
  JUES_API_END(ctx)
}

jues_status jues_rccsd_pT(jues_ctx* ctx, int prec, const double* Cocc, int o, const double* Cvir, int v, int n,
                          const double* eps, const double* T1, const double* T2, int verbose, double* Et) {
  JUES_API_BEGIN(ctx)
  JUES_REQUIRE(Cocc && Cvir && eps && T1 && T2 && Et && o > 0 && v > 0, "jues_rccsd_pT: bad arguments");
  JUES_REQUIRE(n == ctx->n && n > 0, "jues_rccsd_pT: n does not match the AO integrals");
  JUES_REQUIRE(prec == JUES_FP64 || prec == JUES_FP32_MIXED, "jues_rccsd_pT: precision must be FP64 or FP32_MIXED");
  *Et = prec == JUES_FP64 ? pt_from_orbitals<double>(ctx, Cocc, o, Cvir, v, eps, T1, T2, verbose)
                          : pt_from_orbitals<float>(ctx, Cocc, o, Cvir, v, eps, T1, T2, verbose);
  if (!std::isfinite(*Et)) throw Error(JUES_NONFINITE, "jues_rccsd_pT: non-finite energy");
  JUES_API_END(ctx)
}

}  // extern "C"
