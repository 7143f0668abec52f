// admx_qp.hip -- per-column reduction of the pass partials fused with the
// batched K-dim QPs and the projections:
//   update_q!: Xtz .*= -1; per individual create_tableau!(simplex=true) +
//              quadratic_program!(c=1); q_next = q + delta; project_q!
//              (src/algorithms_inner.jl:199-227, src/qp.jl:8-76, 171-317,
//              src/projections.jl:1-63)
//   update_p!: per SNP create_tableau!(simplex=false) + quadratic_program!(c=0);
//              p_next = p + delta; project_p!  (src/algorithms_inner.jl:275-305)
//
// One lane owns one problem.  Its (K+c+1)^2 sweep tableau sits in LDS,
// interleaved across the workgroup (element e of lane t at e*NT + t), so every
// LDS access of a wave touches consecutive addresses; the step, the parameters,
// the initial diagonal and the swept column live in registers (every loop over
// them is unrolled, so their indices are compile-time).  The active-set loop is the reference's
// Jennrich-Sampson method step for step; arithmetic is evaluated without FMA
// contraction, as Julia evaluates it outside @turbo.
#include <type_traits>

#include "admx_common.hpp"
#include "admx_internal.hpp"
#include "admx_project.hpp"

namespace admx {

template <typename T, int K, int C, int NT>
struct QPLayout {
  static constexpr int SZ = K + C + 1;
  static constexpr int TAB = 0;
  static constexpr int TOTAL = SZ * SZ;
};

// Lanes per workgroup of the QP kernels: the LDS footprint grows as K^2.
template <int K> struct QPLanes { static constexpr int N = K <= 12 ? 64 : 32; };

#define QP_TB(r, c) L[((r) + SZ * (c)) * NT + t]

// sweep!(A, k, tmp, inverse)  (src/qp.jl:297-317).  The saved column lives in
// registers; both loops are unrolled, so the SZ^2 read-modify-writes of the
// tableau have compile-time LDS offsets and are issued back to back (their
// arithmetic, element by element, is the reference's: A[j,i] -= (p tmp[i]) tmp[j]).
template <typename T, int K, int C, int NT>
__device__ __forceinline__ void qp_sweep(T* __restrict__ L, int t, int k, bool inverse) {
#pragma clang fp contract(off)
  using LY = QPLayout<T, K, C, NT>;
  constexpr int SZ = K + C + 1;
  const T pk = (T)1 / QP_TB(k, k);
  T tmp[SZ];
#pragma unroll
  for (int i = 0; i < SZ; ++i) tmp[i] = QP_TB(i, k);
#pragma unroll
  for (int i = 0; i < SZ; ++i) {
    QP_TB(i, k) = 0;
    QP_TB(k, i) = 0;
  }
  const T sg = inverse ? (T)1 : (T)-1;
#pragma unroll
  for (int i = 0; i < SZ; ++i) tmp[i] = (i == k) ? sg : tmp[i];
#pragma unroll
  for (int i = 0; i < SZ; ++i) {
    const T pv = pk * tmp[i];
#pragma unroll
    for (int j = 0; j < SZ; ++j) QP_TB(j, i) = QP_TB(j, i) - pv * tmp[j];
  }
}

// quadratic_program!(delta, tableau, par, pmin=0, pmax=1, p=K, c)  (src/qp.jl:171-292)
template <typename T, int K, int C, int NT>
__device__ int qp_solve(T* __restrict__ L, int t, const T (&par)[K], T (&del)[K]) {
#pragma clang fp contract(off)
  using LY = QPLayout<T, K, C, NT>;
  constexpr int SZ = LY::SZ, E = SZ - 1;
  const T small = (T)1e-5, tol = (T)1e-8;
  T dg[K];  // the diagonal before any sweep
#pragma unroll
  for (int i = 0; i < K; ++i) {
    del[i] = 0;
    dg[i] = QP_TB(i, i);
  }
  // Sweep the parameters (rejected when a pivot is not positive or has lost
  // all but tol of its initial value), then the constraint (rejected when its
  // pivot is not negative).  One sweep instance serves every step (rolled loop;
  // dg by an unrolled select, no register indexing).
#pragma unroll 1
  for (int i = 0; i < K + C; ++i) {
    const T di = QP_TB(i, i);
    if (i < K) {
      T d0 = dg[0];
#pragma unroll
      for (int k = 1; k < K; ++k) d0 = (i == k) ? dg[k] : d0;
      if (d0 <= (T)0 || di < d0 * tol) return 0;
    } else if (di >= (T)0) {
      return 0;
    }
    qp_sweep<T, K, C, NT>(L, t, i, false);
  }
  uint32_t swept = (K >= 32) ? 0xffffffffu : ((1u << K) - 1u);
  for (int iteration = 1; iteration <= 1000; ++iteration) {
    T a = 1;
#pragma unroll
    for (int i = 0; i < K; ++i) {
      if ((swept >> i) & 1u) {
        const T ui = QP_TB(i, E);
        T ai;
        if (ui > (T)0) ai = (T)1 - par[i] - del[i];
        else ai = (T)0 - par[i] - del[i];
        if ((T)fabs((double)ui) > (T)1e-10) {
          const T r = ai / ui;
          a = r < a ? r : a;
        }
      }
    }
#pragma unroll
    for (int i = 0; i < K; ++i) {
      if ((swept >> i) & 1u) {
        const T ui = QP_TB(i, E);
        del[i] = del[i] + a * ui;
        const T nu = ((T)1 - a) * ui;
        QP_TB(i, E) = nu;
        QP_TB(E, i) = nu;
      }
    }
    // a swept parameter that is critical is inverse-swept; otherwise an unswept
    // one that violates the KKT conditions is swept; otherwise solved
    int pick = -1;
#pragma unroll
    for (int i = 0; i < K; ++i) {
      const T xi = par[i] + del[i];
      const bool critical = ((T)0 >= xi - small) || ((T)1 <= xi + small);
      if (pick < 0 && ((swept >> i) & 1u) && fabs((double)QP_TB(i, i)) > 1e-10 && critical) pick = i;
    }
    const bool inverse = pick >= 0;
    if (!inverse) {
#pragma unroll
      for (int i = 0; i < K; ++i) {
        const T ui = QP_TB(i, E);
        const T xi = par[i] + del[i];
        const bool violation = (ui > (T)0 && (T)0 >= xi - small) || (ui < (T)0 && (T)1 <= xi + small);
        if (pick < 0 && !((swept >> i) & 1u) && violation) pick = i;
      }
      if (pick < 0) return iteration;
    }
    qp_sweep<T, K, C, NT>(L, t, pick, inverse);
    swept ^= 1u << pick;  // inverse: the bit was set; forward: it was clear
  }
  return 0;
}

// Load a full K x K Hessian + gradient into the tableau of lane t, then solve.
// create_tableau!  (src/qp.jl:8-76).  The Gerschgorin term mu of qp.jl:54-57 is
// computed through a mul! whose output aliases its input; BLAS gemv with
// beta = 0 clears y before reading x, so the reference executes mu = 0.
template <typename T, int K, int C, int NT>
__device__ __forceinline__ int qp_from_hessian(T* __restrict__ L, int t, const T (&Hf)[K * K], const T (&xtz)[K],
                                               const T (&par)[K], T (&del)[K]) {
#pragma clang fp contract(off)
  using LY = QPLayout<T, K, C, NT>;
  constexpr int SZ = LY::SZ;
#pragma unroll
  for (int c = 0; c < SZ; ++c)
#pragma unroll
    for (int r = 0; r < SZ; ++r) {
      T v = 0;
      if (r < K && c < K) v = Hf[r + K * c];
      QP_TB(r, c) = v;
    }
#pragma unroll
  for (int i = 0; i < K; ++i) {
    QP_TB(SZ - 1, i) = xtz[i];  // -r with r = -Xtz
    QP_TB(i, SZ - 1) = xtz[i];
  }
  if constexpr (C == 1) {
#pragma unroll
    for (int i = 0; i < K; ++i) {
      QP_TB(K, i) = 1;
      QP_TB(i, K) = 1;
    }
    T sx = 0;
#pragma unroll
    for (int i = 0; i < K; ++i) sx += par[i];
    QP_TB(K + 1, K) = (T)1 - sx;
    QP_TB(K, K + 1) = (T)1 - sx;
  }
  return qp_solve<T, K, C, NT>(L, t, par, del);
}

// ------------------------------------------------------------------ Q side finalize
// Partials of type PT are summed, and the rank-1 form recombined, in the wider
// of T and PT (an FP32 context's passes run in FP64); the QP runs in T.
// Rank-1 (TRICK) partials carry W = sum (a+b) p p' (upper triangle), hb = sum b p,
// sb = sum b and sd = sum d; with q the pass's own state (x_in),
//   XtX = W + sb 11' - 1 hb' - hb 1',  Xtz = W q - hb + sd 1
// (sum c p + d (1-p) over the same genotypes, by c = a s, d = b t, s = q.p).
template <typename T, int K, bool TRICK, bool DO_QP, typename PT = T>
__global__ __launch_bounds__(QPLanes<K>::N) void k_fin_q_update(FinArgs f) {
  using AT = typename std::conditional<(sizeof(PT) > sizeof(T)), PT, T>::type;
  constexpr int NT = QPLanes<K>::N;
  constexpr int NTRI = Tri<K>::N;
  constexpr int NA = TRICK ? NTRI + K + 2 : NTRI + K;
  using LY = QPLayout<T, K, 1, NT>;
  __shared__ T lds[DO_QP ? LY::TOTAL * NT : 1];
  const int t = threadIdx.x;
  const int64_t col = (int64_t)blockIdx.x * NT + t;
  if (col >= f.n) return;
  const PT* __restrict__ part = static_cast<const PT*>(f.part);
  AT acc[NA];
#pragma unroll
  for (int a = 0; a < NA; ++a) acc[a] = 0;
  for (int c = 0; c < f.nchunk; ++c) {
    const PT* src = part + (int64_t)c * NA * f.n + col;
#pragma unroll
    for (int a = 0; a < NA; ++a) acc[a] += (AT)src[(int64_t)a * f.n];
  }
  const int64_t g = f.off + col;
  T Hf[K * K], xtz[K];
  if constexpr (TRICK) {
    const T* __restrict__ qin = static_cast<const T*>(f.x_in) + (int64_t)K * g;
    AT qv[K], wq[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      qv[k] = (AT)qin[k];
      wq[k] = 0;
    }
    const AT sb = acc[NTRI + K], sd = acc[NTRI + K + 1];
    int n = 0;
#pragma unroll
    for (int k = 0; k < K; ++k)
#pragma unroll
      for (int k2 = k; k2 < K; ++k2, ++n) {
        const AT w = acc[n];
        // (W q)_k summed over ascending columns: entries (k', k), k' < k, come from earlier rows
        wq[k] = __builtin_fma(w, qv[k2], wq[k]);
        if (k2 != k) wq[k2] = __builtin_fma(w, qv[k], wq[k2]);
        const AT h = w + sb - acc[NTRI + k] - acc[NTRI + k2];
        Hf[k2 + K * k] = (T)h;
        Hf[k + K * k2] = (T)h;
      }
#pragma unroll
    for (int k = 0; k < K; ++k) xtz[k] = (T)(wq[k] - acc[NTRI + k] + sd);
  } else {
    int n = 0;
#pragma unroll
    for (int k = 0; k < K; ++k)
#pragma unroll
      for (int k2 = k; k2 < K; ++k2, ++n) {
        Hf[k2 + K * k] = (T)acc[n];
        Hf[k + K * k2] = (T)acc[n];
      }
#pragma unroll
    for (int k = 0; k < K; ++k) xtz[k] = (T)acc[NTRI + k];
  }
  if (f.XtX) {
    T* X = static_cast<T*>(f.XtX) + (int64_t)K * K * g;
#pragma unroll
    for (int e = 0; e < K * K; ++e) X[e] = f.accumulate ? X[e] + Hf[e] : Hf[e];
  }
  if (f.Xtz) {
    T* Z = static_cast<T*>(f.Xtz) + (int64_t)K * g;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const T v = f.neg_xtz ? -xtz[k] : xtz[k];
      Z[k] = f.accumulate ? Z[k] + v : v;
    }
  }
  if constexpr (DO_QP) {
    const T* __restrict__ xin = static_cast<const T*>(f.x_in) + (int64_t)K * g;
    T par[K];
#pragma unroll
    for (int k = 0; k < K; ++k) par[k] = xin[k];
    T* L = lds;
    T del[K];
    const int rc = qp_from_hessian<T, K, 1, NT>(L, t, Hf, xtz, par, del);
    if (f.status) f.status[g] = rc;
    T v[K];
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = par[k] + del[k];
    project_simplex<T, K>(v);
    T* xo = static_cast<T*>(f.x_out) + (int64_t)K * g;
#pragma unroll
    for (int k = 0; k < K; ++k) xo[k] = v[k];
  }
}

// em_q! finalize (src/algorithms_inner.jl:76-89): q_next = q .* E / 2J, or the
// tile loop's q_next += q .* E.  The FP64 partials (every context's passes run
// in FP64) carry sum (c-d) p_k and sum d, K+1 per chunk; the sums are formed in
// FP64 and the result is stored in T.
template <typename T, int K>
__global__ __launch_bounds__(256) void k_fin_q_em(FinArgs f) {
  const int64_t col = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (col >= f.n) return;
  const double* __restrict__ part = static_cast<const double*>(f.part);
  constexpr int NA = K + 1;
  double acc[K], dsum = 0;
#pragma unroll
  for (int k = 0; k < K; ++k) acc[k] = 0;
  for (int c = 0; c < f.nchunk; ++c) {
    const double* src = part + (int64_t)c * NA * f.n + col;
#pragma unroll
    for (int k = 0; k < K; ++k) acc[k] += src[(int64_t)k * f.n];
    dsum += src[(int64_t)K * f.n];
  }
#pragma unroll
  for (int k = 0; k < K; ++k) acc[k] += dsum;
  const int64_t g = f.off + col;
  const T* q = static_cast<const T*>(f.x_in) + (int64_t)K * g;
  T* o = static_cast<T*>(f.x_out) + (int64_t)K * g;
  const T div = (T)(2 * f.J);
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const T v = q[k] * (T)acc[k];
    o[k] = f.accumulate ? o[k] + v : v / div;
  }
}

// ------------------------------------------------------------------ P side finalize
template <typename T, int K, bool DO_QP, typename PT = T>
__global__ __launch_bounds__(QPLanes<K>::N) void k_fin_p_update(FinArgs f) {
  using AT = typename std::conditional<(sizeof(PT) > sizeof(T)), PT, T>::type;
  constexpr int NT = QPLanes<K>::N;
  constexpr int NTRI = Tri<K>::N;
  constexpr int NA = NTRI + K;
  using LY = QPLayout<T, K, 0, NT>;
  __shared__ T lds[DO_QP ? LY::TOTAL * NT : 1];
  const int t = threadIdx.x;
  const int64_t col = (int64_t)blockIdx.x * NT + t;
  if (col >= f.n) return;
  const PT* __restrict__ part = static_cast<const PT*>(f.part);
  AT acc[NA];
#pragma unroll
  for (int a = 0; a < NA; ++a) acc[a] = 0;
  for (int c = 0; c < f.nchunk; ++c) {
    const PT* src = part + (int64_t)c * NA * f.n + col;
#pragma unroll
    for (int a = 0; a < NA; ++a) acc[a] += (AT)src[(int64_t)a * f.n];
  }
  T Hf[K * K], xtz[K];
  {
    int n = 0;
#pragma unroll
    for (int k = 0; k < K; ++k)
#pragma unroll
      for (int k2 = k; k2 < K; ++k2, ++n) {
        Hf[k2 + K * k] = (T)acc[n];
        Hf[k + K * k2] = (T)acc[n];
      }
#pragma unroll
    for (int k = 0; k < K; ++k) xtz[k] = (T)acc[NTRI + k];
  }
  const int64_t g = f.off + col;
  if (f.XtX) {
    T* X = static_cast<T*>(f.XtX) + (int64_t)K * K * g;
#pragma unroll
    for (int e = 0; e < K * K; ++e) X[e] = f.accumulate ? X[e] + Hf[e] : Hf[e];
  }
  if (f.Xtz) {
    T* Z = static_cast<T*>(f.Xtz) + (int64_t)K * g;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const T v = f.neg_xtz ? -xtz[k] : xtz[k];
      Z[k] = f.accumulate ? Z[k] + v : v;
    }
  }
  if constexpr (DO_QP) {
    const T* __restrict__ xin = static_cast<const T*>(f.x_in) + (int64_t)K * g;
    T par[K];
#pragma unroll
    for (int k = 0; k < K; ++k) par[k] = xin[k];
    T* L = lds;
    T del[K];
    const int rc = qp_from_hessian<T, K, 0, NT>(L, t, Hf, xtz, par, del);
    if (f.status) f.status[g] = rc;
    T* xo = static_cast<T*>(f.x_out) + (int64_t)K * g;
#pragma unroll
    for (int k = 0; k < K; ++k) xo[k] = project_box<T>(par[k] + del[k]);
  }
}

// em_p! finalize (src/algorithms_inner.jl:119-137): p_tmp = p .* A, p_next =
// (1-p) .* B, then p_next = p_tmp / (p_tmp + p_next); or the tile loop's +=.
template <typename T, int K>
__global__ __launch_bounds__(256) void k_fin_p_em(FinArgs f) {
  const int64_t col = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (col >= f.n) return;
  const double* __restrict__ part = static_cast<const double*>(f.part);  // FP64 partials (FP64 passes)
  double A[K], B[K];
#pragma unroll
  for (int k = 0; k < K; ++k) A[k] = B[k] = 0;
  for (int c = 0; c < f.nchunk; ++c) {
    const double* src = part + (int64_t)c * 2 * K * f.n + col;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      A[k] += src[(int64_t)k * f.n];
      B[k] += src[(int64_t)(K + k) * f.n];
    }
  }
  const int64_t g = f.off + col;
  const T* p = static_cast<const T*>(f.x_in) + (int64_t)K * g;
  T* o = static_cast<T*>(f.x_out) + (int64_t)K * g;
  if (f.accumulate) {
    T* pt = static_cast<T*>(f.aux) + (int64_t)K * g;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      pt[k] = pt[k] + p[k] * (T)A[k];
      o[k] = o[k] + ((T)1 - p[k]) * (T)B[k];
    }
  } else {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const T pt = p[k] * (T)A[k], pn = ((T)1 - p[k]) * (T)B[k];
      o[k] = pt / (pt + pn);
    }
  }
}

// ------------------------------------------------------------------ standalone batched QP
// One problem per lane from host-shaped inputs: XtX K x K x n, Xtz (already
// negated, as the reference stores it after `Xtz .*= -1`), x K x n.
template <typename T, int K, int C>
__global__ __launch_bounds__(QPLanes<K>::N) void k_qp_batch(int64_t n, const T* __restrict__ XtX,
                                                            const T* __restrict__ Xtz_neg, const T* __restrict__ x,
                                                            T* __restrict__ delta, int32_t* __restrict__ status) {
  constexpr int NT = QPLanes<K>::N;
  using LY = QPLayout<T, K, C, NT>;
  __shared__ T lds[LY::TOTAL * NT];
  const int t = threadIdx.x;
  const int64_t col = (int64_t)blockIdx.x * NT + t;
  if (col >= n) return;
  T Hf[K * K], xtz[K], par[K];
#pragma unroll
  for (int e = 0; e < K * K; ++e) Hf[e] = XtX[(int64_t)K * K * col + e];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    xtz[k] = -Xtz_neg[(int64_t)K * col + k];
    par[k] = x[(int64_t)K * col + k];
  }
  T* L = lds;
  T del[K];
  const int rc = qp_from_hessian<T, K, C, NT>(L, t, Hf, xtz, par, del);
  if (status) status[col] = rc;
#pragma unroll
  for (int k = 0; k < K; ++k) delta[(int64_t)K * col + k] = del[k];
}

// QP + step + projection of the reduced Hessians: the second half of the
// finalize when the QP tableau's LDS footprint would leave a fused reduce + QP
// kernel at one or two waves per CU.  Reads XtX (full K x K) and Xtz as the
// reduction stored them (negated when neg_xtz), so the tableau, and every
// result, equals the fused kernel's bit for bit.
template <typename T, int K, int C>
__global__ __launch_bounds__(QPLanes<K>::N) void k_qp_apply(FinArgs f, const T* __restrict__ X,
                                                            const T* __restrict__ Z) {
  constexpr int NT = QPLanes<K>::N;
  using LY = QPLayout<T, K, C, NT>;
  __shared__ T lds[LY::TOTAL * NT];
  const int t = threadIdx.x;
  const int64_t col = (int64_t)blockIdx.x * NT + t;
  if (col >= f.n) return;
  const int64_t g = f.off + col;
  const T* __restrict__ xin = static_cast<const T*>(f.x_in) + (int64_t)K * g;
  T Hf[K * K], xtz[K], par[K];
#pragma unroll
  for (int e = 0; e < K * K; ++e) Hf[e] = X[(int64_t)K * K * g + e];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const T z = Z[(int64_t)K * g + k];
    xtz[k] = f.neg_xtz ? -z : z;
    par[k] = xin[k];
  }
  T* L = lds;
  T del[K];
  const int rc = qp_from_hessian<T, K, C, NT>(L, t, Hf, xtz, par, del);
  if (f.status) f.status[g] = rc;
  T* xo = static_cast<T*>(f.x_out) + (int64_t)K * g;
  if constexpr (C == 1) {
    T v[K];
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = par[k] + del[k];
    project_simplex<T, K>(v);
#pragma unroll
    for (int k = 0; k < K; ++k) xo[k] = v[k];
  } else {
#pragma unroll
    for (int k = 0; k < K; ++k) xo[k] = project_box<T>(par[k] + del[k]);
  }
}

// Split the finalize into a reduction kernel and k_qp_apply from K = 9
// (simplex) / 10 (box).  The fused kernel holds two or three 64-problem
// workgroups per CU (the tableau's LDS: 50 KiB per workgroup at K = 8 simplex,
// 100 KiB at K = 12), so it reads the partials at low occupancy; the reduction
// kernel reads them at full occupancy and k_qp_apply reads the sums.  I = J = 100,000, K = 12: fin_q
// 12.1 -> 2.4 ms, fin_p 10 -> 1.3 ms; I = 200,000, J = 50,000, K = 10, P side
// (84 chunks x 65 partials per SNP): 1.68 ms fused against 0.28 split; at
// K = 8 (c3: 40 chunks x 44 partials) the fused kernel is faster, 0.73 vs
// 0.86 ms.
template <typename T, int K, int C>
constexpr bool qp_split() {
  return C == 1 ? K >= 9 : K >= 10;
}

// Launch the split path if it applies; false leaves the fused kernel to the caller.
template <typename T, int K, int C, typename RED>
static bool fin_split(const FinArgs& f, hipStream_t s, RED reduce_kernel) {
  if (!qp_split<T, K, C>() || f.accumulate) return false;
  void* X = f.XtX ? f.XtX : f.XtX_scratch;
  void* Z = f.Xtz ? f.Xtz : f.Xtz_scratch;
  if (!X || !Z) return false;
  constexpr int NT = QPLanes<K>::N;
  FinArgs r = f;
  r.XtX = X;
  r.Xtz = Z;
  hipLaunchKernelGGL(reduce_kernel, dim3((unsigned)((f.n + NT - 1) / NT)), dim3(NT), 0, s, r);
  hipLaunchKernelGGL((k_qp_apply<T, K, C>), dim3((unsigned)((f.n + NT - 1) / NT)), dim3(NT), 0, s, r,
                     static_cast<const T*>(X), static_cast<const T*>(Z));
  return true;
}

template <typename T, int K>
__global__ __launch_bounds__(256) void k_project_q(int64_t n, T* __restrict__ b) {
  const int64_t col = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (col >= n) return;
  T v[K];
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = b[(int64_t)K * col + k];
  project_simplex<T, K>(v);
#pragma unroll
  for (int k = 0; k < K; ++k) b[(int64_t)K * col + k] = v[k];
}

template <typename T>
__global__ __launch_bounds__(256) void k_project_p(int64_t n, T* __restrict__ b) {
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < n; e += (int64_t)gridDim.x * 256)
    b[e] = project_box<T>(b[e]);
}

// ------------------------------------------------------------------ launchers
static inline unsigned nblk(int64_t n, int nt) { return (unsigned)((n + nt - 1) / nt); }

// The genotype passes of every context run in FP64 and write FP64 partials in
// the rank-1 Q-side layout; T is the context's state and QP type.
template <typename T, int K>
static void fin_q_update_k(const FinArgs& f, bool do_qp, hipStream_t s) {
  constexpr int NT = QPLanes<K>::N;
  if (do_qp && fin_split<T, K, 1>(f, s, k_fin_q_update<T, K, true, false, double>)) return;
  if (do_qp) hipLaunchKernelGGL((k_fin_q_update<T, K, true, true, double>), dim3(nblk(f.n, NT)), dim3(NT), 0, s, f);
  else hipLaunchKernelGGL((k_fin_q_update<T, K, true, false, double>), dim3(nblk(f.n, NT)), dim3(NT), 0, s, f);
}
template <typename T, int K>
static void fin_p_update_k(const FinArgs& f, bool do_qp, hipStream_t s) {
  constexpr int NT = QPLanes<K>::N;
  if (do_qp && fin_split<T, K, 0>(f, s, k_fin_p_update<T, K, false, double>)) return;
  if (do_qp) hipLaunchKernelGGL((k_fin_p_update<T, K, true, double>), dim3(nblk(f.n, NT)), dim3(NT), 0, s, f);
  else hipLaunchKernelGGL((k_fin_p_update<T, K, false, double>), dim3(nblk(f.n, NT)), dim3(NT), 0, s, f);
}
template <typename T, int K>
static void fin_q_em_k(const FinArgs& f, hipStream_t s) {
  hipLaunchKernelGGL((k_fin_q_em<T, K>), dim3(nblk(f.n, 256)), dim3(256), 0, s, f);
}
template <typename T, int K>
static void fin_p_em_k(const FinArgs& f, hipStream_t s) {
  hipLaunchKernelGGL((k_fin_p_em<T, K>), dim3(nblk(f.n, 256)), dim3(256), 0, s, f);
}
template <typename T, int K>
static void qp_batch_k(bool simplex, int64_t n, const void* XtX, const void* Xtz, const void* x, void* delta,
                       int32_t* status, hipStream_t s) {
  constexpr int NT = QPLanes<K>::N;
  if (simplex)
    hipLaunchKernelGGL((k_qp_batch<T, K, 1>), dim3(nblk(n, NT)), dim3(NT), 0, s, n, (const T*)XtX, (const T*)Xtz,
                       (const T*)x, (T*)delta, status);
  else
    hipLaunchKernelGGL((k_qp_batch<T, K, 0>), dim3(nblk(n, NT)), dim3(NT), 0, s, n, (const T*)XtX, (const T*)Xtz,
                       (const T*)x, (T*)delta, status);
}
template <typename T, int K>
static void project_q_k(int64_t n, void* b, hipStream_t s) {
  hipLaunchKernelGGL((k_project_q<T, K>), dim3(nblk(n, 256)), dim3(256), 0, s, n, (T*)b);
}

template <typename T> void launch_fin_q_update(int K, const FinArgs& f, bool do_qp, hipStream_t s) {
  ADMX_DISPATCH_K(K, (fin_q_update_k<T, KK>(f, do_qp, s)));
}
template <typename T> void launch_fin_p_update(int K, const FinArgs& f, bool do_qp, hipStream_t s) {
  ADMX_DISPATCH_K(K, (fin_p_update_k<T, KK>(f, do_qp, s)));
}
template <typename T> void launch_fin_q_em(int K, const FinArgs& f, hipStream_t s) {
  ADMX_DISPATCH_K(K, (fin_q_em_k<T, KK>(f, s)));
}
template <typename T> void launch_fin_p_em(int K, const FinArgs& f, hipStream_t s) {
  ADMX_DISPATCH_K(K, (fin_p_em_k<T, KK>(f, s)));
}
template <typename T>
void launch_qp_batch(int K, bool simplex, int64_t n, const void* XtX, const void* Xtz_neg, const void* x, void* delta,
                     int32_t* status, hipStream_t s) {
  ADMX_DISPATCH_K(K, (qp_batch_k<T, KK>(simplex, n, XtX, Xtz_neg, x, delta, status, s)));
}
template <typename T> void launch_project_q(int K, int64_t n, void* b, hipStream_t s) {
  ADMX_DISPATCH_K(K, (project_q_k<T, KK>(n, b, s)));
}
template <typename T> void launch_project_p(int64_t n, void* b, hipStream_t s) {
  const int64_t nb = (n + 255) / 256;
  hipLaunchKernelGGL((k_project_p<T>), dim3((unsigned)(nb < 65536 ? (nb > 0 ? nb : 1) : 65536)), dim3(256), 0, s, n,
                     (T*)b);
}

#define ADMX_INST(T)                                                                                         \
  template void launch_fin_q_update<T>(int, const FinArgs&, bool, hipStream_t);                               \
  template void launch_fin_p_update<T>(int, const FinArgs&, bool, hipStream_t);                               \
  template void launch_fin_q_em<T>(int, const FinArgs&, hipStream_t);                                         \
  template void launch_fin_p_em<T>(int, const FinArgs&, hipStream_t);                                         \
  template void launch_qp_batch<T>(int, bool, int64_t, const void*, const void*, const void*, void*, int32_t*, \
                                   hipStream_t);                                                             \
  template void launch_project_q<T>(int, int64_t, void*, hipStream_t);                                        \
  template void launch_project_p<T>(int64_t, void*, hipStream_t);
ADMX_INST(ADMX_T)

#ifdef ADMX_WITH_MIXED
// FP64 finalize of FP32 partials (mixed-precision passes, direct Hessian form).
template <int K>
static void fin_q_mixed_k(const FinArgs& f, hipStream_t s) {
  constexpr int NT = QPLanes<K>::N;
  if (fin_split<double, K, 1>(f, s, k_fin_q_update<double, K, false, false, float>)) return;
  hipLaunchKernelGGL((k_fin_q_update<double, K, false, true, float>), dim3(nblk(f.n, NT)), dim3(NT), 0, s, f);
}
template <int K>
static void fin_p_mixed_k(const FinArgs& f, hipStream_t s) {
  constexpr int NT = QPLanes<K>::N;
  if (fin_split<double, K, 0>(f, s, k_fin_p_update<double, K, false, float>)) return;
  hipLaunchKernelGGL((k_fin_p_update<double, K, true, float>), dim3(nblk(f.n, NT)), dim3(NT), 0, s, f);
}
void launch_fin_q_update_mixed(int K, const FinArgs& f, hipStream_t s) { ADMX_DISPATCH_K(K, (fin_q_mixed_k<KK>(f, s))); }
void launch_fin_p_update_mixed(int K, const FinArgs& f, hipStream_t s) { ADMX_DISPATCH_K(K, (fin_p_mixed_k<KK>(f, s))); }
#endif

}  // namespace admx
