// Bandwidth-bound kernels of the coupled-cluster iteration: symmetric packing of
// the particle-particle ladder, the fused residual-assembly / Jacobi update /
// energy / DIIS-dot kernel, the deterministic partial-sum reduction, the on-device
// bordered DIIS solve and the extrapolation.
#include "cc_kernels.h"

namespace jues {
namespace {

// pair (p<=q) index p + q(q+1)/2 and strict pair (p<q) index p + q(q-1)/2
__device__ inline long tri_le(int p, int q) { return (long)p + (long)q * (q + 1) / 2; }
__device__ inline long tri_lt(int p, int q) { return (long)p + (long)q * (q - 1) / 2; }

// Tp[(i<=j),(e<=f)] = t[ijef] + t[ijfe] (e<f), t[ijee] (e=f)
// Tm[(i<j),(e<f)]   = t[ijef] - t[ijfe]
// With t1 non-null the packed amplitudes are those of tau = t + t1 t1,
// tau[ijef] = t[ijef] + t1[i,e] t1[j,f] (the CCSD ladders), formed on the fly.
// With ROWDIAG the single term is taken on the row diagonal instead: for the hole-hole
// ladder integrals K[kl,cd] = <kl|cd>, KP[(k<=l),(c<=d)] = K[klcd] + K[kldc] (k<l),
// K[kkcd] (k=l) and KM[(k<l),(c<d)] = K[klcd] - K[kldc].
template <typename T, bool ROWDIAG = false>
__global__ __launch_bounds__(256) void pack_t_kernel(const double* __restrict__ t, const double* __restrict__ t1,
                                                     int o, int v,
                                                     const int2* __restrict__ pairs_o_le, const int2* __restrict__ pairs_v_le,
                                                     long npo, long npv, T* __restrict__ Tp,
                                                     const int2* __restrict__ pairs_o_lt, const int2* __restrict__ pairs_v_lt,
                                                     long nmo, long nmv, T* __restrict__ Tm) {
  const long oo = (long)o * o;
  const long tot_p = npo * npv, tot_m = nmo * nmv;
  for (long x = blockIdx.x * 256L + threadIdx.x; x < tot_p + tot_m; x += (long)gridDim.x * 256L) {
    const bool plus = x < tot_p;
    const long y = plus ? x : x - tot_p;
    const long nr = plus ? npo : nmo;
    const long r = y % nr, c = y / nr;
    const int2 ij = plus ? pairs_o_le[r] : pairs_o_lt[r], ef = plus ? pairs_v_le[c] : pairs_v_lt[c];
    const long base = ij.x + (long)o * ij.y;
    double a = t[base + oo * (ef.x + (long)v * ef.y)];
    double b = t[base + oo * (ef.y + (long)v * ef.x)];
    if (t1) {
      a += t1[ij.x + (long)o * ef.x] * t1[ij.y + (long)o * ef.y];
      b += t1[ij.x + (long)o * ef.y] * t1[ij.y + (long)o * ef.x];
    }
    if (plus) Tp[y] = (T)((ROWDIAG ? ij.x == ij.y : ef.x == ef.y) ? a : a + b);
    else Tm[y] = (T)(a - b);
  }
}

// From a chemist slab X[a,e,b,f'] (f' = f - f0, f in [f0,f1)) build
// Vp[(e<=f),(a<=b)] = ((ae|bf) + (af|be))/2 and Vm[(e<f),(a<b)] = ((ae|bf) - (af|be))/2,
// using (af|be) = X[b,e,a,f'].
// Row rr of the packed matrix goes to Vp[(rr - rowp0) + ldp * c] (Vm likewise): the full
// matrix with rowp0 = 0, ldp = npv, or a slab holding only this call's rows.
template <typename Ts, typename Td>
__global__ __launch_bounds__(256) void pack_v_kernel(const Ts* __restrict__ X, int v, int f0, int f1,
                                                     const int2* __restrict__ pairs_v_le, long npv, Td* __restrict__ Vp,
                                                     long rowp0, long ldp, const int2* __restrict__ pairs_v_lt,
                                                     long nmv, Td* __restrict__ Vm, long rowm0, long ldm) {
  // rows: e<=f with f in [f0,f1): q from f0(f0+1)/2 to f1(f1+1)/2
  const long r0p = (long)f0 * (f0 + 1) / 2, r1p = (long)f1 * (f1 + 1) / 2;
  const long r0m = (long)f0 * (f0 - 1) / 2, r1m = (long)f1 * (f1 - 1) / 2;
  const long nrp = r1p - r0p, nrm = r1m - r0m;
  const long tot_p = nrp * npv, tot_m = nrm * nmv;
  const long v1 = v, v2 = v1 * v, v3 = v2 * v;
  for (long x = blockIdx.x * 256L + threadIdx.x; x < tot_p + tot_m; x += (long)gridDim.x * 256L) {
    if (x < tot_p) {
      const long rr = r0p + x % nrp, c = x / nrp;
      const int2 ef = pairs_v_le[rr], ab = pairs_v_le[c];
      const long fl = ef.y - f0;
      const double p = (double)X[ab.x + v1 * ef.x + v2 * ab.y + v3 * fl];
      const double q = (double)X[ab.y + v1 * ef.x + v2 * ab.x + v3 * fl];
      Vp[(rr - rowp0) + ldp * c] = (Td)(0.5 * (p + q));
    } else {
      const long y = x - tot_p;
      const long rr = r0m + y % nrm, c = y / nrm;
      const int2 ef = pairs_v_lt[rr], ab = pairs_v_lt[c];
      const long fl = ef.y - f0;
      const double p = (double)X[ab.x + v1 * ef.x + v2 * ab.y + v3 * fl];
      const double q = (double)X[ab.y + v1 * ef.x + v2 * ab.x + v3 * fl];
      Vm[(rr - rowm0) + ldm * c] = (Td)(0.5 * (p - q));
    }
  }
}

// W[m,n,i,j] (o^4, holding <mn|ij> plus any T1 terms) on packed pairs, transposed:
// WPT[(i<=j),(k<=l)] = W[klij] + W[lkij] (k<l), W[kkij] (k=l);  WMT[(i<j),(k<l)] = W[klij] - W[lkij]
template <typename T>
__global__ __launch_bounds__(256) void pack_whh_kernel(const T* __restrict__ W, int o, const int2* __restrict__ o_le,
                                                       long npo, T* __restrict__ WPT, const int2* __restrict__ o_lt,
                                                       long nmo, T* __restrict__ WMT) {
  const long oo = (long)o * o;
  const long tot_p = npo * npo, tot_m = nmo * nmo;
  for (long x = blockIdx.x * 256L + threadIdx.x; x < tot_p + tot_m; x += (long)gridDim.x * 256L) {
    const bool plus = x < tot_p;
    const long y = plus ? x : x - tot_p, n = plus ? npo : nmo;
    const long r = y % n, c = y / n;
    const int2 ij = plus ? o_le[r] : o_lt[r], kl = plus ? o_le[c] : o_lt[c];
    const long col = oo * (ij.x + (long)o * ij.y);
    const double a = (double)W[kl.x + (long)o * kl.y + col];
    const double b = (double)W[kl.y + (long)o * kl.x + col];
    if (plus) WPT[y] = (T)(kl.x == kl.y ? a : a + b);
    else WMT[y] = (T)(a - b);
  }
}

// Tq = Tp with its (a,a) columns doubled (exact)
template <typename T>
__global__ __launch_bounds__(256) void pair_diag2_kernel(const T* __restrict__ Tp, long npo, long npv,
                                                         const int2* __restrict__ v_le, T* __restrict__ Tq) {
  for (long x = blockIdx.x * 256L + threadIdx.x; x < npo * npv; x += (long)gridDim.x * 256L) {
    const int2 ab = v_le[x / npo];
    const T t = Tp[x];
    Tq[x] = ab.x == ab.y ? t + t : t;
  }
}

// WE[(k,c),(j,b)] += s (delta_kj Fae[b,c] - delta_cb Fmi[k,j]) on the ov layout (row k + o c,
// column j + o b), with Fmi held transposed (FmiT[j,k] = Fmi[k,j]).  Threads x < o v^2 take
// the (k,c,b) entries with j = k (and the Fmi part where c = b); the rest take (k, j != k, b)
// with c = b, so every entry has one writer.
template <typename T>
__global__ __launch_bounds__(256) void fock_dress_kernel(T* __restrict__ WE, const T* __restrict__ Fae,
                                                         const T* __restrict__ FmiT, int o, int v, T s) {
  const long ov = (long)o * v, n1 = ov * v, n2 = (long)o * (o - 1) * v;
  for (long x = blockIdx.x * 256L + threadIdx.x; x < n1 + n2; x += (long)gridDim.x * 256L) {
    if (x < n1) {
      const int k = (int)(x % o);
      const long cb = x / o;
      const int c = (int)(cb % v), b = (int)(cb / v);
      T d = Fae[b + (long)v * c];
      if (c == b) d -= FmiT[k + (long)o * k];
      WE[k + (long)o * c + ov * (k + (long)o * b)] += s * d;
    } else {
      const long y = x - n1;
      const int k = (int)(y % o);
      const long jb = y / o;
      const int jj = (int)(jb % (o - 1)), b = (int)(jb / (o - 1));
      const int j = jj < k ? jj : jj + 1;
      WE[k + (long)o * b + ov * (j + (long)o * b)] -= s * FmiT[j + (long)o * k];
    }
  }
}

// The RCCD residual's amplitude layouts from t2 (FP64, [i,j,a,b]) in one pass:
// Tpp = t, Upp = 2t - t^(ij) ([i,j,a,b]); Tov[i,a,j,b] = t[ijab]; Uov[i,a,j,b] = U[ijab];
// Ttov[i,b,j,a] = t[ijab] (that is t[i,j,b,a] at [i,a,j,b]).  Threads run i fastest, so
// every write is a contiguous run of o; the exchanged read t[jiab] stays in one o x o slab.
template <typename T>
__global__ __launch_bounds__(256) void rccd_layouts_kernel(const double* __restrict__ t, int o, int v,
                                                           T* __restrict__ Tpp, T* __restrict__ Upp,
                                                           T* __restrict__ Tov, T* __restrict__ Ttov,
                                                           T* __restrict__ Uov) {
  const long oo = (long)o * o, ov = (long)o * v, n = oo * v * v;
  for (long x = blockIdx.x * 256L + threadIdx.x; x < n; x += (long)gridDim.x * 256L) {
    const long ij = x % oo, ab = x / oo;
    const int i = (int)(ij % o), j = (int)(ij / o);
    const int a = (int)(ab % v), b = (int)(ab / v);
    const double tv = t[x];
    const double u = 2.0 * tv - t[j + (long)o * i + oo * ab];
    if (Tpp) Tpp[x] = (T)tv;
    if (Upp) Upp[x] = (T)u;
    const long y = i + (long)o * a + ov * (j + (long)o * b);
    Tov[y] = (T)tv;
    Uov[y] = (T)u;
    Ttov[i + (long)o * b + ov * (j + (long)o * a)] = (T)tv;
  }
}

// The same layouts with the o x o slab t[:,:,a,b] of one (a,b) pair staged in LDS (rows padded to
// o + 1), so that the exchanged read t[j,i,a,b] comes from LDS and every global read is
// contiguous.  Dynamic LDS: o (o + 1) doubles.
template <typename T>
__global__ __launch_bounds__(256) void rccd_layouts_slab_kernel(const double* __restrict__ t, int o, int v,
                                                                T* __restrict__ Tpp, T* __restrict__ Upp,
                                                                T* __restrict__ Tov, T* __restrict__ Ttov,
                                                                T* __restrict__ Uov) {
  extern __shared__ double slab[];  // slab[i + (o + 1) j] = t[i,j,a,b]
  const int oo = o * o, ld = o + 1;
  const long ov = (long)o * v, vv = (long)v * v;
  for (long ab = blockIdx.x; ab < vv; ab += gridDim.x) {
    const int a = (int)(ab % v), b = (int)(ab / v);
    const double* src = t + ab * oo;
    __syncthreads();  // the previous pair's readers are done with the slab
    for (int q = threadIdx.x; q < oo; q += 256) slab[q % o + ld * (q / o)] = src[q];
    __syncthreads();
    for (int q = threadIdx.x; q < oo; q += 256) {
      const int i = q % o, j = q / o;
      const double tv = slab[i + ld * j];
      const double u = 2.0 * tv - slab[j + ld * i];
      if (Tpp) Tpp[ab * oo + q] = (T)tv;
      if (Upp) Upp[ab * oo + q] = (T)u;
      const long y = i + (long)o * a + ov * (j + (long)o * b);
      Tov[y] = (T)tv;
      Uov[y] = (T)u;
      Ttov[i + (long)o * b + ov * (j + (long)o * a)] = (T)tv;
    }
  }
}

__device__ inline double block_sum256(double v, double* sh) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  return (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

// One block per (a,b): assemble R[ijab] = K[ijab] + S[ia,jb] + L+ + L- (both ladders),
// t_new = R/D, e = t_new - t_old; write t_new and e into the DIIS slot; block
// partial sums of E = sum Lener*t_new, e.e and e.e_k for the stored errors.
template <typename TR>
__global__ __launch_bounds__(256) void t2_update_kernel(UpdateArgs<TR> A) {
  __shared__ double sh[4];
  const int o = A.o, v = A.v;
  const int ab = blockIdx.x;
  const int a = ab % v, b = ab / v;
  const long oo = (long)o * o, ov = (long)o * v;
  const double dab = -A.eps[o + a] - A.eps[o + b];
  const int sab = a < b ? 1 : (a > b ? -1 : 0);
  const long qp = a <= b ? (long)a + (long)b * (b + 1) / 2 : (long)b + (long)a * (a + 1) / 2;
  const long qm = a < b ? (long)a + (long)b * (b - 1) / 2 : (long)b + (long)a * (a - 1) / 2;
  double sE = 0.0, sEE = 0.0;
  double sd[MAX_DIIS];
#pragma unroll
  for (int k = 0; k < MAX_DIIS; ++k) sd[k] = 0.0;
  // fused symmetrisation: the transposed terms H[jb,ia] + Y[ja,ib] run along j in memory, so
  // they are read as contiguous j-runs into an LDS slab[j + (o + 1) i] and taken from there
  extern __shared__ __attribute__((aligned(16))) unsigned char dyn_lds[];
  TR* slab = reinterpret_cast<TR*>(dyn_lds);
  const int ld = o + 1;
  if (A.H) {
    for (int q = threadIdx.x; q < (int)oo; q += 256) {
      const int j = q % o, i = q / o;
      slab[j + ld * i] = A.H[(long)j + (long)o * b + ov * ((long)i + (long)o * a)] +
                         A.Y[(long)j + (long)o * a + ov * ((long)i + (long)o * b)];
    }
    __syncthreads();
  }
  for (int ij = threadIdx.x; ij < (int)oo; ij += 256) {
    const int i = ij % o, j = ij / o;
    const long x = (long)ij + oo * ab;
    double R = A.Kpp[x];
    if (A.S) R += (double)A.S[(long)i + (long)o * a + ov * ((long)j + (long)o * b)];
    if (A.H)
      R += (double)A.H[(long)i + (long)o * a + ov * ((long)j + (long)o * b)] +
           (double)A.Y[(long)i + (long)o * b + ov * ((long)j + (long)o * a)] + (double)slab[j + ld * i];
    if (A.Lp) {
      const long pp = i <= j ? (long)i + (long)j * (j + 1) / 2 : (long)j + (long)i * (i + 1) / 2;
      R += (double)A.Lp[pp + A.npo * qp];
      const int sij = i < j ? 1 : (i > j ? -1 : 0);
      if (sij != 0 && sab != 0) {
        const long pm = i < j ? (long)i + (long)j * (j - 1) / 2 : (long)j + (long)i * (i - 1) / 2;
        R += (double)(sij * sab) * (double)A.Lm[pm + A.nmo * qm];
      }
    }
    if (A.extra) R += (double)A.extra[x];
    if (A.R_out) A.R_out[x] = R;
    const double tn = A.lambda * R / (A.eps[i] + A.eps[j] + dab);
    const double e = tn - A.t_old[x];
    A.t_new[x] = tn;
    A.e_new[x] = e;
    const double tt = A.t1n ? A.t1n[i + (long)o * a] * A.t1n[j + (long)o * b] : 0.0;
    sE += A.lambda * A.Lener[x] * (tn + tt);
    sEE += e * e;
#pragma unroll
    for (int k = 0; k < MAX_DIIS; ++k)
      if (k < A.nold) sd[k] += e * A.e_old[k][x];
  }
  const long PB = A.part_stride;
  double r = block_sum256(sE, sh);
  if (threadIdx.x == 0) A.part[0 * PB + A.part_off + blockIdx.x] = r;
  r = block_sum256(sEE, sh);
  if (threadIdx.x == 0) A.part[1 * PB + A.part_off + blockIdx.x] = r;
  for (int k = 0; k < A.nold; ++k) {
    r = block_sum256(sd[k], sh);
    if (threadIdx.x == 0) A.part[(2 + k) * PB + A.part_off + blockIdx.x] = r;
  }
}

// T1 part of the update: t1_new = R1/D_ia, error and partials (RCCSD).
__global__ __launch_bounds__(256) void t1_update_kernel(T1UpdateArgs A) {
  __shared__ double sh[4];
  const long n1 = (long)A.o * A.v;
  double sE = 0.0, sEE = 0.0, sd[MAX_DIIS];
#pragma unroll
  for (int k = 0; k < MAX_DIIS; ++k) sd[k] = 0.0;
  for (long x = blockIdx.x * 256L + threadIdx.x; x < n1; x += (long)gridDim.x * 256L) {
    const int i = (int)(x % A.o), a = (int)(x / A.o);
    const double tn = A.lambda * A.R1[x] / (A.eps[i] - A.eps[A.o + a]);
    const double e = tn - A.t_old[x];
    A.t_new[x] = tn;
    A.e_new[x] = e;
    sE += A.Lener1 ? A.Lener1[x] * tn : 0.0;
    sEE += e * e;
#pragma unroll
    for (int k = 0; k < MAX_DIIS; ++k)
      if (k < A.nold) sd[k] += e * A.e_old[k][x];
  }
  const long PB = A.part_stride;
  double r = block_sum256(sE, sh);
  if (threadIdx.x == 0) A.part[0 * PB + A.part_off + blockIdx.x] = r;
  r = block_sum256(sEE, sh);
  if (threadIdx.x == 0) A.part[1 * PB + A.part_off + blockIdx.x] = r;
  for (int k = 0; k < A.nold; ++k) {
    r = block_sum256(sd[k], sh);
    if (threadIdx.x == 0) A.part[(2 + k) * PB + A.part_off + blockIdx.x] = r;
  }
}

// Fixed-order sums of nq partial arrays; out[q] = sum_b part[q*PB + b].
__global__ __launch_bounds__(256) void reduce_rows_kernel(const double* __restrict__ part, long PB, long count,
                                                          double* __restrict__ out) {
  __shared__ double sh[256];
  const int q = blockIdx.x;
  double s = 0.0;
  for (long b = threadIdx.x; b < count; b += 256) s += part[q * PB + b];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[q] = sh[0];
}

// Store the new B row/column and solve the bordered Pulay system (mRCCD.jl:185-198)
// in FP64 with partial pivoting.  small layout: [0]=E, [1]=ee, [2..2+nold)=dots,
// B at SMALL_B (MAX_DIIS^2, by slot), coefficients at SMALL_C (MAX_DIIS, by order).
__global__ void diis_solve_kernel(double* small, DiisOrder ord) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double* B = small + SMALL_B;
  const int s = ord.newest;
  B[s * MAX_DIIS + s] = small[1];
  for (int k = 0; k < ord.nold; ++k) {
    const int sk = ord.old_slots[k];
    B[s * MAX_DIIS + sk] = small[2 + k];
    B[sk * MAX_DIIS + s] = small[2 + k];
  }
  const int m = ord.m;
  double* c = small + SMALL_C;
  if (m == 1) { c[0] = 1.0; small[SMALL_FLAG] = 0.0; return; }
  double M[(MAX_DIIS + 1) * (MAX_DIIS + 1)];
  double r[MAX_DIIS + 1];
  const int N = m + 1;
  double mx = 0.0;
  for (int p = 0; p < m; ++p)
    for (int q = 0; q < m; ++q) mx = fmax(mx, fabs(B[ord.slots[p] * MAX_DIIS + ord.slots[q]]));
  if (!(mx > 0.0)) mx = 1.0;
  for (int p = 0; p < N; ++p) {
    for (int q = 0; q < N; ++q) {
      double val;
      if (p < m && q < m) val = B[ord.slots[p] * MAX_DIIS + ord.slots[q]] / mx;
      else if (p == m && q == m) val = 0.0;
      else val = -1.0;
      M[p * N + q] = val;
    }
    r[p] = p == m ? -1.0 : 0.0;
  }
  bool ok = true;
  for (int k = 0; k < N; ++k) {
    int piv = k;
    for (int p = k + 1; p < N; ++p)
      if (fabs(M[p * N + k]) > fabs(M[piv * N + k])) piv = p;
    if (!(fabs(M[piv * N + k]) > 1e-300)) { ok = false; break; }
    if (piv != k) {
      for (int q = 0; q < N; ++q) { const double tmp = M[k * N + q]; M[k * N + q] = M[piv * N + q]; M[piv * N + q] = tmp; }
      const double tmp = r[k]; r[k] = r[piv]; r[piv] = tmp;
    }
    for (int p = k + 1; p < N; ++p) {
      const double f = M[p * N + k] / M[k * N + k];
      for (int q = k; q < N; ++q) M[p * N + q] -= f * M[k * N + q];
      r[p] -= f * r[k];
    }
  }
  if (ok) {
    for (int k = N - 1; k >= 0; --k) {
      double acc = r[k];
      for (int q = k + 1; q < N; ++q) acc -= M[k * N + q] * r[q];
      r[k] = acc / M[k * N + k];
    }
    for (int p = 0; p < m; ++p) ok = ok && isfinite(r[p]);
  }
  if (ok) {
    for (int p = 0; p < m; ++p) c[p] = r[p];
    small[SMALL_FLAG] = 0.0;
  } else {  // singular subspace: take the newest vector (a plain Jacobi step)
    for (int p = 0; p < m; ++p) c[p] = p == m - 1 ? 1.0 : 0.0;
    small[SMALL_FLAG] = 1.0;
  }
}

// t[x] = sum_k c_k vals[slot_k][x]
__global__ __launch_bounds__(256) void diis_extrap_kernel(const double* __restrict__ vals, long stride, DiisOrder ord,
                                                          const double* __restrict__ small, long n, double* __restrict__ t) {
  double c[MAX_DIIS];
#pragma unroll
  for (int k = 0; k < MAX_DIIS; ++k) c[k] = k < ord.m ? small[SMALL_C + k] : 0.0;
  for (long x = blockIdx.x * 256L + threadIdx.x; x < n; x += (long)gridDim.x * 256L) {
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < MAX_DIIS; ++k)
      if (k < ord.m) s += c[k] * vals[(long)ord.slots[k] * stride + x];
    t[x] = s;
  }
}

// t[ijab] /= D[ijab]
__global__ __launch_bounds__(256) void divide_d2_kernel(double* __restrict__ t, const double* __restrict__ eps, int o,
                                                        int v, double scale) {
  const long oo = (long)o * o, n = oo * v * v;
  for (long x = blockIdx.x * 256L + threadIdx.x; x < n; x += (long)gridDim.x * 256L) {
    const long ij = x % oo, ab = x / oo;
    const int i = (int)(ij % o), j = (int)(ij / o), a = (int)(ab % v), b = (int)(ab / v);
    t[x] = scale * t[x] / (eps[i] + eps[j] - eps[o + a] - eps[o + b]);
  }
}

// As pack_v_kernel, from a slab generated with rows in (e,a) order: X[e,a,b,f'] = (ae|bf),
// so that (ae|bf) = X[e,a,b,f'] and (af|be) = (be|af) = X[e,b,a,f'] are both contiguous in e
// and threads that run over e also write consecutive packed rows (pairs e<=f of one f).
template <typename Ts, typename Td>
__global__ __launch_bounds__(256) void pack_v_ea_kernel(const Ts* __restrict__ X, int v, int f0, int f1,
                                                        const int2* __restrict__ pairs_v_le, long npv,
                                                        Td* __restrict__ Vp, const int2* __restrict__ pairs_v_lt,
                                                        long nmv, Td* __restrict__ Vm) {
  const long r0p = (long)f0 * (f0 + 1) / 2, r1p = (long)f1 * (f1 + 1) / 2;
  const long r0m = (long)f0 * (f0 - 1) / 2, r1m = (long)f1 * (f1 - 1) / 2;
  const long nrp = r1p - r0p, nrm = r1m - r0m;
  const long tot_p = nrp * npv, tot_m = nrm * nmv;
  const long v1 = v, v2 = v1 * v, v3 = v2 * v;
  for (long x = blockIdx.x * 256L + threadIdx.x; x < tot_p + tot_m; x += (long)gridDim.x * 256L) {
    if (x < tot_p) {
      const long rl = x % nrp, c = x / nrp;
      const int2 ef = pairs_v_le[r0p + rl], ab = pairs_v_le[c];
      const long base = ef.x + v3 * (ef.y - f0);
      const double p = (double)X[base + v1 * ab.x + v2 * ab.y];
      const double q = (double)X[base + v1 * ab.y + v2 * ab.x];
      Vp[rl + nrp * c] = (Td)(0.5 * (p + q));
    } else {
      const long y = x - tot_p;
      const long rl = y % nrm, c = y / nrm;
      const int2 ef = pairs_v_lt[r0m + rl], ab = pairs_v_lt[c];
      const long base = ef.x + v3 * (ef.y - f0);
      const double p = (double)X[base + v1 * ab.x + v2 * ab.y];
      const double q = (double)X[base + v1 * ab.y + v2 * ab.x];
      Vm[rl + nrm * c] = (Td)(0.5 * (p - q));
    }
  }
}

int grid256(long n) { return (int)std::max(1L, std::min((n + 255) / 256, 131072L)); }

}  // namespace

template <typename T>
void pack_t(jues_ctx* ctx, const PairTables& P, const double* t, int o, int v, T* Tp, T* Tm, const double* t1) {
  const long tot = P.npo * P.npv + P.nmo * P.nmv;
  KTimer k(ctx, "cc_pack_t", 0.0, (double)tot * (16.0 + sizeof(T)));
  hipLaunchKernelGGL((pack_t_kernel<T>), dim3(grid256(tot)), dim3(256), 0, ctx->stream, t, t1, o, v, P.o_le, P.v_le,
                     P.npo, P.npv, Tp, P.o_lt, P.v_lt, P.nmo, P.nmv, Tm);
  JUES_HIP(hipGetLastError());
}

void pack_khh(jues_ctx* ctx, const PairTables& P, const double* K, int o, int v, double* KP, double* KM) {
  const long tot = P.npo * P.npv + P.nmo * P.nmv;
  KTimer k(ctx, "cc_pack_k", 0.0, (double)tot * 24.0);
  hipLaunchKernelGGL((pack_t_kernel<double, true>), dim3(grid256(tot)), dim3(256), 0, ctx->stream, K, nullptr, o, v,
                     P.o_le, P.v_le, P.npo, P.npv, KP, P.o_lt, P.v_lt, P.nmo, P.nmv, KM);
  JUES_HIP(hipGetLastError());
}

template <typename T>
void pack_whh(jues_ctx* ctx, const PairTables& P, const T* W, int o, T* WPT, T* WMT) {
  const long tot = P.npo * P.npo + P.nmo * P.nmo;
  KTimer k(ctx, "cc_hh_pack", 0.0, (double)tot * 3.0 * sizeof(T));
  hipLaunchKernelGGL((pack_whh_kernel<T>), dim3(grid256(tot)), dim3(256), 0, ctx->stream, W, o, P.o_le, P.npo, WPT,
                     P.o_lt, P.nmo, WMT);
  JUES_HIP(hipGetLastError());
}

template <typename T>
void pair_diag2(jues_ctx* ctx, const PairTables& P, const T* Tp, T* Tq) {
  const long tot = P.npo * P.npv;
  KTimer k(ctx, "cc_hh_pack", 0.0, (double)tot * 2.0 * sizeof(T));
  hipLaunchKernelGGL((pair_diag2_kernel<T>), dim3(grid256(tot)), dim3(256), 0, ctx->stream, Tp, P.npo, P.npv, P.v_le,
                     Tq);
  JUES_HIP(hipGetLastError());
}

template <typename T>
void fock_dress(jues_ctx* ctx, T* WE, const T* Fae, const T* FmiT, int o, int v, T s) {
  const long n = (long)o * v * v + (long)o * (o - 1) * v;
  KTimer k(ctx, "cc_fock", 0.0, (double)n * 3.0 * sizeof(T));
  hipLaunchKernelGGL((fock_dress_kernel<T>), dim3(grid256(n)), dim3(256), 0, ctx->stream, WE, Fae, FmiT, o, v, s);
  JUES_HIP(hipGetLastError());
}

template <typename T>
void rccd_layouts(jues_ctx* ctx, const double* t, int o, int v, T* Tpp, T* Upp, T* Tov, T* Ttov, T* Uov) {
  const long n = (long)o * o * v * v;
  const size_t slab = (size_t)o * (o + 1) * sizeof(double);
  // t is read once through the LDS slab, twice (direct and exchanged) otherwise
  const double outs = 3.0 + (Tpp ? 1.0 : 0.0) + (Upp ? 1.0 : 0.0);
  KTimer k(ctx, "cc_prep", 0.0, (double)n * ((slab <= 65536 ? 8.0 : 16.0) + outs * sizeof(T)));
  if (slab <= 65536) {
    const long vv = (long)v * v;
    hipLaunchKernelGGL((rccd_layouts_slab_kernel<T>), dim3((unsigned)std::min(vv, 1L << 20)), dim3(256), slab,
                       ctx->stream, t, o, v, Tpp, Upp, Tov, Ttov, Uov);
  } else {
    hipLaunchKernelGGL((rccd_layouts_kernel<T>), dim3(grid256(n)), dim3(256), 0, ctx->stream, t, o, v, Tpp, Upp, Tov,
                       Ttov, Uov);
  }
  JUES_HIP(hipGetLastError());
}

template <typename Ts, typename Td>
void pack_v(jues_ctx* ctx, const PairTables& P, const Ts* X, int v, int f0, int f1, Td* Vp, Td* Vm, bool slab) {
  const long r0p = (long)f0 * (f0 + 1) / 2, r0m = (long)f0 * (f0 - 1) / 2;
  const long nrp = (long)f1 * (f1 + 1) / 2 - r0p;
  const long nrm = (long)f1 * (f1 - 1) / 2 - r0m;
  const long tot = nrp * P.npv + nrm * P.nmv;
  KTimer k(ctx, "cc_pack_v", 0.0, (double)tot * (2 * sizeof(Ts) + sizeof(Td)));
  hipLaunchKernelGGL((pack_v_kernel<Ts, Td>), dim3(grid256(tot)), dim3(256), 0, ctx->stream, X, v, f0, f1, P.v_le,
                     P.npv, Vp, slab ? r0p : 0L, slab ? std::max(1L, nrp) : P.npv, P.v_lt, P.nmv, Vm,
                     slab ? r0m : 0L, slab ? std::max(1L, nrm) : P.nmv);
  JUES_HIP(hipGetLastError());
}

template <typename TR>
void t2_update(jues_ctx* ctx, const UpdateArgs<TR>& A) {
  const long v2 = (long)A.v * A.v;
  const long n2 = (long)A.o * A.o * v2;
  // S (or H and Y, each read twice), Lp, Lm and extra in the residual's compute type
  KTimer k(ctx, "cc_update", 0.0, (double)n2 * (8.0 * (5 + A.nold + (A.R_out ? 1 : 0)) + (A.H ? 7 : 4) * sizeof(TR)));
  size_t lds = 0;
  if (A.H) {
    JUES_REQUIRE(A.Y && !A.S && fused_sym(A.o, sizeof(TR)), "t2_update: fused symmetrisation misconfigured");
    lds = (size_t)A.o * (A.o + 1) * sizeof(TR);
  }
  hipLaunchKernelGGL((t2_update_kernel<TR>), dim3((unsigned)v2), dim3(256), lds, ctx->stream, A);
  JUES_HIP(hipGetLastError());
}

void t1_update(jues_ctx* ctx, const T1UpdateArgs& A, int nblocks) {
  hipLaunchKernelGGL(t1_update_kernel, dim3(nblocks), dim3(256), 0, ctx->stream, A);
  JUES_HIP(hipGetLastError());
}

void reduce_rows(jues_ctx* ctx, const double* part, long PB, long count, int nq, double* out) {
  hipLaunchKernelGGL(reduce_rows_kernel, dim3(nq), dim3(256), 0, ctx->stream, part, PB, count, out);
  JUES_HIP(hipGetLastError());
}

void diis_solve(jues_ctx* ctx, double* small, const DiisOrder& ord) {
  hipLaunchKernelGGL(diis_solve_kernel, dim3(1), dim3(64), 0, ctx->stream, small, ord);
  JUES_HIP(hipGetLastError());
}

void diis_extrap(jues_ctx* ctx, const double* vals, long stride, const DiisOrder& ord, const double* small, long n,
                 double* t) {
  KTimer k(ctx, "cc_diis_extrap", 0.0, (double)n * 8.0 * (ord.m + 1));
  hipLaunchKernelGGL(diis_extrap_kernel, dim3(grid256(n)), dim3(256), 0, ctx->stream, vals, stride, ord, small, n, t);
  JUES_HIP(hipGetLastError());
}

void divide_d2(jues_ctx* ctx, double* t, const double* eps, int o, int v, double scale) {
  const long n = (long)o * o * v * v;
  hipLaunchKernelGGL(divide_d2_kernel, dim3(grid256(n)), dim3(256), 0, ctx->stream, t, eps, o, v, scale);
  JUES_HIP(hipGetLastError());
}

template void fock_dress<double>(jues_ctx*, double*, const double*, const double*, int, int, double);
template void fock_dress<float>(jues_ctx*, float*, const float*, const float*, int, int, float);
template void rccd_layouts<double>(jues_ctx*, const double*, int, int, double*, double*, double*, double*, double*);
template void rccd_layouts<float>(jues_ctx*, const double*, int, int, float*, float*, float*, float*, float*);
template void pack_whh<double>(jues_ctx*, const PairTables&, const double*, int, double*, double*);
template void pack_whh<float>(jues_ctx*, const PairTables&, const float*, int, float*, float*);
template void pair_diag2<double>(jues_ctx*, const PairTables&, const double*, double*);
template void pair_diag2<float>(jues_ctx*, const PairTables&, const float*, float*);
template void pack_t<double>(jues_ctx*, const PairTables&, const double*, int, int, double*, double*, const double*);
template void pack_t<float>(jues_ctx*, const PairTables&, const double*, int, int, float*, float*, const double*);
template <typename T>
void pack_v_ea(jues_ctx* ctx, const PairTables& P, const T* X, int v, int f0, int f1, T* Vp, T* Vm) {
  const long nrp = (long)f1 * (f1 + 1) / 2 - (long)f0 * (f0 + 1) / 2;
  const long nrm = (long)f1 * (f1 - 1) / 2 - (long)f0 * (f0 - 1) / 2;
  const long tot = nrp * P.npv + nrm * P.nmv;
  KTimer k(ctx, "cc_pack_v", 0.0, (double)tot * 3 * sizeof(T));
  hipLaunchKernelGGL((pack_v_ea_kernel<T, T>), dim3(grid256(tot)), dim3(256), 0, ctx->stream, X, v, f0, f1, P.v_le,
                     P.npv, Vp, P.v_lt, P.nmv, Vm);
  JUES_HIP(hipGetLastError());
}

template void pack_v_ea<double>(jues_ctx*, const PairTables&, const double*, int, int, int, double*, double*);
template void pack_v_ea<float>(jues_ctx*, const PairTables&, const float*, int, int, int, float*, float*);
template void pack_v<double, double>(jues_ctx*, const PairTables&, const double*, int, int, int, double*, double*,
                                     bool);
template void pack_v<double, float>(jues_ctx*, const PairTables&, const double*, int, int, int, float*, float*, bool);
template void pack_v<float, float>(jues_ctx*, const PairTables&, const float*, int, int, int, float*, float*, bool);
template void t2_update<double>(jues_ctx*, const UpdateArgs<double>&);
template void t2_update<float>(jues_ctx*, const UpdateArgs<float>&);

}  // namespace jues
