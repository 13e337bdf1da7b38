// Declarations of the coupled-cluster bandwidth-bound kernels (cc_kernels.hip).
#pragma once

#include "jues_internal.h"

namespace jues {

constexpr int MAX_DIIS = 8;
// layout of the small FP64 device array used by the update / DIIS kernels
constexpr int SMALL_B = 16;                               // MAX_DIIS x MAX_DIIS Gram matrix, by slot
constexpr int SMALL_C = SMALL_B + MAX_DIIS * MAX_DIIS;   // coefficients, by age order
constexpr int SMALL_FLAG = SMALL_C + MAX_DIIS;          // 1 if the last solve was singular
constexpr int SMALL_SIZE = SMALL_FLAG + 8;

// (p,q) pair tables for the packed ladder: *_le lists p<=q, *_lt lists p<q, both
// in column-major triangular order (index p + q(q+1)/2, resp. p + q(q-1)/2).
struct PairTables {
  int2 *o_le = nullptr, *o_lt = nullptr, *v_le = nullptr, *v_lt = nullptr;
  long npo = 0, nmo = 0, npv = 0, nmv = 0;
};

template <typename TR>
struct UpdateArgs {
  int o, v;
  const double* eps;
  const double* Kpp;    // <ij|ab> (FP64, ijab)
  const double* Lener;  // 2<ij|ab> - <ji|ab> (FP64, ijab)
  const TR* S;          // ring + Fock-like terms, symmetrised, ov layout [i,a,j,b]; or null with:
  const TR* H;          //   H and Y (ov layout): S[ia,jb] = H[ia,jb] + Y[ib,ja] + H[jb,ia] + Y[ja,ib],
  const TR* Y;          //   summed by the update kernel itself (fused_sym)
  const TR* Lp;         // packed ladders (pp + hh), symmetric part
  const TR* Lm;         // packed ladders (pp + hh), antisymmetric part
  const TR* extra;      // any further ijab-ordered residual terms (RCCSD), may be null
  const double* t1n;    // new T1 (o x v) for the CCSD energy term, null for RCCD
  double* R_out;        // if set, receives the numerator R (FP64, ijab; before lambda and D)
  double lambda;        // coupling: t_new = lambda R / D, E = lambda sum Lener tau
  long npo, nmo;
  const double* t_old;
  double* t_new;
  double* e_new;
  const double* e_old[MAX_DIIS];
  int nold;
  double* part;         // partial sums: part[q*part_stride + part_off + block]
  long part_stride, part_off;
};

struct T1UpdateArgs {
  int o, v;
  double lambda;
  const double* eps;
  const double* R1;     // T1 residual numerator (FP64, o x v)
  const double* Lener1; // may be null
  const double* t_old;
  double* t_new;
  double* e_new;
  const double* e_old[MAX_DIIS];
  int nold;
  double* part;
  long part_stride, part_off;
};

struct DiisOrder {
  int m;                    // number of active pairs
  int slots[MAX_DIIS];      // oldest -> newest
  int newest;               // slot just written
  int nold;                 // number of other active slots
  int old_slots[MAX_DIIS];  // their slots, in the order their dots were accumulated
};

// packed T+/T- of t, or of tau = t + t1 t1 (tau[ijef] = t[ijef] + t1[i,e] t1[j,f]) when t1 is given
template <typename T>
void pack_t(jues_ctx* ctx, const PairTables& P, const double* t, int o, int v, T* Tp, T* Tm,
            const double* t1 = nullptr);
template <typename Ts, typename Td>
void pack_v(jues_ctx* ctx, const PairTables& P, const Ts* X, int v, int f0, int f1, Td* Vp, Td* Vm,
            bool slab = false);
// hole-hole ladder integrals on pairs: KP[(k<=l),(c<=d)], KM[(k<l),(c<d)] from K[k,l,c,d] = <kl|cd>
void pack_khh(jues_ctx* ctx, const PairTables& P, const double* K, int o, int v, double* KP, double* KM);
// W[m,n,i,j] -> WPT[(i<=j),(k<=l)], WMT[(i<j),(k<l)] (the kl-pair sum / difference, transposed)
template <typename T>
void pack_whh(jues_ctx* ctx, const PairTables& P, const T* W, int o, T* WPT, T* WMT);
// RCCD amplitude layouts from FP64 t2 [i,j,a,b]: Tpp = t, Upp = 2t - t^(ij), Tov[i,a,j,b] = t,
// Ttov[i,a,j,b] = t[i,j,b,a], Uov[i,a,j,b] = U[i,j,a,b]; Tpp and Upp may be null
template <typename T>
void rccd_layouts(jues_ctx* ctx, const double* t, int o, int v, T* Tpp, T* Upp, T* Tov, T* Ttov, T* Uov);
// WE[(k,c),(j,b)] += s (delta_kj Fae[b,c] - delta_cb Fmi[k,j]) (ov layout; FmiT[j,k] = Fmi[k,j])
template <typename T>
void fock_dress(jues_ctx* ctx, T* WE, const T* Fae, const T* FmiT, int o, int v, T s);
// Tq = Tp with its (a,a) columns doubled
template <typename T>
void pair_diag2(jues_ctx* ctx, const PairTables& P, const T* Tp, T* Tq);
// slab-local V+/V- rows (pairs with f in [f0,f1)) from X[e,a,b,f-f0] = (ae|bf)
template <typename T>
void pack_v_ea(jues_ctx* ctx, const PairTables& P, const T* X, int v, int f0, int f1, T* Vp, T* Vm);
template <typename TR>
void t2_update(jues_ctx* ctx, const UpdateArgs<TR>& A);
// whether the update kernel sums S = H + H^T + Y^x + (Y^x)^T itself: its o x (o + 1) LDS slab
// of element size `elem` must fit in 48 KiB
inline bool fused_sym(int o, size_t elem) { return (size_t)o * (o + 1) * elem <= 48 * 1024; }
void t1_update(jues_ctx* ctx, const T1UpdateArgs& A, int nblocks);
void reduce_rows(jues_ctx* ctx, const double* part, long PB, long count, int nq, double* out);
void diis_solve(jues_ctx* ctx, double* small, const DiisOrder& ord);
void diis_extrap(jues_ctx* ctx, const double* vals, long stride, const DiisOrder& ord, const double* small, long n,
                 double* t);
void divide_d2(jues_ctx* ctx, double* t, const double* eps, int o, int v, double scale);

}  // namespace jues
