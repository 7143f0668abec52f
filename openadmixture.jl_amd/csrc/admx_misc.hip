// admx_misc.hip -- the smaller device kernels around the passes:
//   * fixed-order reduction of per-block log-likelihood partials;
//   * ZAL quasi-Newton extrapolation (src/quasi_newton.jl:5-31) with the
//     projections of src/algorithms_outer.jl:141-142 fused into its last step;
//   * the QN accept rule (src/algorithms_outer.jl:150-156) on device;
//   * BED decode and missing / allele counts (HBM-bound);
//   * the synthetic-cohort genotype generator.
#include "admx_common.hpp"
#include "admx_internal.hpp"
#include "admx_project.hpp"

namespace admx {

// ------------------------------------------------------------------ reductions
__global__ __launch_bounds__(256) void k_reduce_sum(const double* __restrict__ parts, int64_t n,
                                                    double* __restrict__ out, int accumulate) {
  __shared__ double sh[256];
  double s = 0.0;
  for (int64_t e = threadIdx.x; e < n; e += 256) s += parts[e];
  const double tot = block_sum<256>(s, sh);
  if (threadIdx.x == 0) *out = accumulate ? *out + tot : tot;
}

void launch_reduce_sum(const double* parts, int64_t nparts, double* out, int accumulate, hipStream_t s) {
  hipLaunchKernelGGL(k_reduce_sum, dim3(1), dim3(256), 0, s, parts, nparts, out, accumulate);
}

__global__ __launch_bounds__(256) void k_reduce_sum_if(const double* __restrict__ parts, int64_t n,
                                                       double* __restrict__ out, const int* __restrict__ flag,
                                                       int want) {
  if ((*flag != 0) != (want != 0)) return;  // uniform
  __shared__ double sh[256];
  double s = 0.0;
  for (int64_t e = threadIdx.x; e < n; e += 256) s += parts[e];
  const double tot = block_sum<256>(s, sh);
  if (threadIdx.x == 0) *out = tot;
}

void launch_reduce_sum_if(const double* parts, int64_t nparts, double* out, const int* flag, int want,
                          hipStream_t s) {
  hipLaunchKernelGGL(k_reduce_sum_if, dim3(1), dim3(256), 0, s, parts, nparts, out, flag, want);
}

// Block sums of NP per-thread values of a 256-thread block, stored to out[0, NP):
// a fixed shuffle tree per wave, then the four wave sums in wave order.
template <int NP>
__device__ __forceinline__ void block_store_sums(const double (&acc)[NP], double* sh, double* __restrict__ out) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int a = 0; a < NP; ++a) {
    double x = acc[a];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
    if (lane == 0) sh[wave * NP + a] = x;
  }
  __syncthreads();
  for (int a = threadIdx.x; a < NP; a += 256) out[a] = ((sh[a] + sh[NP + a]) + sh[2 * NP + a]) + sh[3 * NP + a];
}

// ------------------------------------------------------------------ ZAL quasi-Newton
// update_UV!: U[:,col] = x_next - x, V[:,col] = x_next2 - x_next, and the
// partial inner products A[a,b] = U[:,a]'(U[:,b] - V[:,b]) and
// r[a] = U[:,a]'(x - x_next) over the first m columns, accumulated in double.
template <typename T>
__global__ __launch_bounds__(256) void k_zal_dots(int64_t N, const T* __restrict__ x, const T* __restrict__ xn,
                                                  const T* __restrict__ xn2, T* __restrict__ U, T* __restrict__ V,
                                                  int col, int m, double* __restrict__ part) {
  constexpr int NP = kMaxQ * kMaxQ + kMaxQ;
  __shared__ double sh[4 * NP];
  double acc[NP];
#pragma unroll
  for (int a = 0; a < NP; ++a) acc[a] = 0.0;
  for (int64_t n = (int64_t)blockIdx.x * 256 + threadIdx.x; n < N; n += (int64_t)gridDim.x * 256) {
    const T xv = x[n], xnv = xn[n], xn2v = xn2[n];
    const T u = xnv - xv, v = xn2v - xnv;
    U[n + N * col] = u;
    V[n + N * col] = v;
    T uu[kMaxQ], vv[kMaxQ];
#pragma unroll
    for (int a = 0; a < kMaxQ; ++a) {
      if (a < m) {
        uu[a] = (a == col) ? u : U[n + N * a];
        vv[a] = (a == col) ? v : V[n + N * a];
      } else {
        uu[a] = 0;
        vv[a] = 0;
      }
    }
    const T dx = xv - xnv;
#pragma unroll
    for (int a = 0; a < kMaxQ; ++a) {
      acc[kMaxQ * kMaxQ + a] += (double)uu[a] * (double)dx;
#pragma unroll
      for (int b = 0; b < kMaxQ; ++b) acc[a + kMaxQ * b] += (double)uu[a] * (double)(uu[b] - vv[b]);
    }
  }
  block_store_sums<NP>(acc, sh, part + (int64_t)blockIdx.x * NP);
}

// Sum of the NP per-block partials of a dots kernel over nblocks blocks into
// tot (shared), by a 256-thread block in a fixed order: wave w owns the values
// a = w, w + 4, ..., its lane l adds blocks l, l + 64, ... of all of them at
// once (every load independent), and a fixed shuffle tree combines the lanes.
template <int NP>
__device__ __forceinline__ void sum_partials(const double* __restrict__ part, int nblocks, double* tot) {
  constexpr int NV = (NP + 3) / 4;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double acc[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) acc[v] = 0.0;
  for (int b = lane; b < nblocks; b += 64) {
#pragma unroll
    for (int v = 0; v < NV; ++v)
      if (wave + 4 * v < NP) acc[v] += part[(int64_t)b * NP + wave + 4 * v];
  }
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    double x = acc[v];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
    if (lane == 0 && wave + 4 * v < NP) tot[wave + 4 * v] = x;
  }
  __syncthreads();
}

// Sum the block partials (sum_partials), then coef = inv(A) * r with the
// inverse by Gauss-Jordan elimination with partial pivoting (Julia's inv).
__global__ __launch_bounds__(256) void k_zal_solve(const double* __restrict__ part, int nblocks, int m,
                                                   double* __restrict__ coef) {
  constexpr int NP = kMaxQ * kMaxQ + kMaxQ;
  __shared__ double tot[NP];
  sum_partials<NP>(part, nblocks, tot);
  if (threadIdx.x != 0) return;
  double W[kMaxQ][2 * kMaxQ];
  const int w = 2 * m;
  for (int r = 0; r < m; ++r)
    for (int c = 0; c < m; ++c) {
      W[r][c] = tot[r + kMaxQ * c];
      W[r][m + c] = (r == c) ? 1.0 : 0.0;
    }
  for (int c = 0; c < m; ++c) {
    int piv = c;
    for (int r = c + 1; r < m; ++r)
      if (fabs(W[r][c]) > fabs(W[piv][c])) piv = r;
    if (piv != c)
      for (int t = 0; t < w; ++t) {
        const double tt = W[c][t];
        W[c][t] = W[piv][t];
        W[piv][t] = tt;
      }
    const double inv = 1.0 / W[c][c];
    for (int t = 0; t < w; ++t) W[c][t] *= inv;
    for (int r = 0; r < m; ++r)
      if (r != c) {
        const double f = W[r][c];
        if (f != 0.0)
          for (int t = 0; t < w; ++t) W[r][t] -= f * W[c][t];
      }
  }
  for (int a = 0; a < m; ++a) {
    double s = 0.0;
    for (int b = 0; b < m; ++b) s += W[a][m + b] * tot[kMaxQ * kMaxQ + b];
    coef[a] = s;
  }
}

// x_tmp = x_next - V coef, then project_q! on the Q columns ...
template <typename T, int K>
__global__ __launch_bounds__(256) void k_zal_apply_q(int64_t I, int64_t N, const T* __restrict__ xn,
                                                     const T* __restrict__ V, const double* __restrict__ coef, int m,
                                                     T* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= I) return;
  T cf[kMaxQ];
#pragma unroll
  for (int a = 0; a < kMaxQ; ++a) cf[a] = (a < m) ? (T)coef[a] : (T)0;
  T v[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int64_t n = (int64_t)K * i + k;
    T s = 0;
    for (int a = 0; a < m; ++a) s += V[n + N * a] * cf[a];
    v[k] = xn[n] - s;
  }
  project_simplex<T, K>(v);
#pragma unroll
  for (int k = 0; k < K; ++k) out[(int64_t)K * i + k] = v[k];
}

// ... and project_p! on the P entries.
template <typename T>
__global__ __launch_bounds__(256) void k_zal_apply_p(int64_t n0, int64_t N, const T* __restrict__ xn,
                                                     const T* __restrict__ V, const double* __restrict__ coef, int m,
                                                     T* __restrict__ out) {
  T cf[kMaxQ];
#pragma unroll
  for (int a = 0; a < kMaxQ; ++a) cf[a] = (a < m) ? (T)coef[a] : (T)0;
  for (int64_t n = n0 + (int64_t)blockIdx.x * 256 + threadIdx.x; n < N; n += (int64_t)gridDim.x * 256) {
    T s = 0;
    for (int a = 0; a < m; ++a) s += V[n + N * a] * cf[a];
    out[n] = project_box<T>(xn[n] - s);
  }
}

template <typename T, int K>
static void zal_apply_q_k(int64_t I, int64_t N, const void* xn, const void* V, const double* coef, int m, void* out,
                          hipStream_t s) {
  hipLaunchKernelGGL((k_zal_apply_q<T, K>), dim3((unsigned)((I + 255) / 256)), dim3(256), 0, s, I, N, (const T*)xn,
                     (const T*)V, coef, m, (T*)out);
}

static unsigned grid_for(int64_t n, int64_t cap) {
  int64_t b = (n + 255) / 256;
  if (b < 1) b = 1;
  return (unsigned)(b < cap ? b : cap);
}

template <typename T>
void launch_zal(int K, int64_t I, int64_t J, const void* x, const void* xn, const void* xn2, void* U, void* V,
                void* x_out, int col, int m, int Qcols, double* scratch, int64_t scratch_len, hipStream_t s) {
  (void)Qcols;
  const int64_t N = (int64_t)K * (I + J);
  constexpr int NP = kMaxQ * kMaxQ + kMaxQ;
  int64_t nb = grid_for(N, 512);
  if ((nb * NP + kMaxQ) > scratch_len) nb = (scratch_len - kMaxQ) / NP;
  double* part = scratch;
  double* coef = scratch + nb * NP;
  hipLaunchKernelGGL((k_zal_dots<T>), dim3((unsigned)nb), dim3(256), 0, s, N, (const T*)x, (const T*)xn,
                     (const T*)xn2, (T*)U, (T*)V, col, m, part);
  hipLaunchKernelGGL(k_zal_solve, dim3(1), dim3(256), 0, s, part, (int)nb, m, coef);
  ADMX_DISPATCH_K(K, (zal_apply_q_k<T, KK>(I, N, xn, V, coef, m, x_out, s)));
  const int64_t np = N - (int64_t)K * I;
  hipLaunchKernelGGL((k_zal_apply_p<T>), dim3(grid_for(np, 4096)), dim3(256), 0, s, (int64_t)K * I, N,
                     (const T*)xn, (const T*)V, coef, m, (T*)x_out);
}

// ------------------------------------------------------------------ LBQN quasi-Newton
// update_UV_LBQN! (src/quasi_newton.jl:15-23) after the host has shifted the
// history columns right: U[:,0] = x_next - x, V[:,0] = (x_next2 - x_next) - U[:,0],
// and the partial inner products G[a,b] = V[:,a]'V[:,b] and r[a] = V[:,a]'U[:,0]
// over the first m columns, in double.
template <typename T>
__global__ __launch_bounds__(256) void k_lbqn_dots(int64_t N, const T* __restrict__ x, const T* __restrict__ xn,
                                                   const T* __restrict__ xn2, T* __restrict__ U, T* __restrict__ V,
                                                   int m, double* __restrict__ part) {
  constexpr int NP = kMaxQ * kMaxQ + kMaxQ;
  __shared__ double sh[4 * NP];
  double acc[NP];
#pragma unroll
  for (int a = 0; a < NP; ++a) acc[a] = 0.0;
  for (int64_t n = (int64_t)blockIdx.x * 256 + threadIdx.x; n < N; n += (int64_t)gridDim.x * 256) {
    const T xv = x[n], xnv = xn[n], xn2v = xn2[n];
    const T u = xnv - xv;
    T v = xn2v - xnv;
    v = v - u;
    U[n] = u;
    V[n] = v;
    T vv[kMaxQ];
#pragma unroll
    for (int a = 0; a < kMaxQ; ++a) vv[a] = (a == 0) ? v : ((a < m) ? V[n + N * a] : (T)0);
#pragma unroll
    for (int a = 0; a < kMaxQ; ++a) {
      acc[kMaxQ * kMaxQ + a] += (double)vv[a] * (double)u;
#pragma unroll
      for (int b = 0; b < kMaxQ; ++b) acc[a + kMaxQ * b] += (double)vv[a] * (double)vv[b];
    }
  }
  block_store_sums<NP>(acc, sh, part + (int64_t)blockIdx.x * NP);
}

// The two-loop recursion of update_QN_LBQN! (src/quasi_newton.jl:37-58) in
// Gram form: with x_q^(i) = u - sum_{k<i} alpha_k V_k,
//   alpha_i = (r_i - sum_{k<i} alpha_k G[i,k]) / G[i,i],  gamma = r_0 / G[0,0],
// and x_next = x - gamma u + sum_i gamma alpha_i V_i - sum_i alpha_i U_i:
// coef[0..kMaxQ) multiplies U, coef[kMaxQ..2kMaxQ) multiplies V.
__global__ __launch_bounds__(256) void k_lbqn_solve(const double* __restrict__ part, int nblocks, int m,
                                                    double* __restrict__ coef) {
  constexpr int NP = kMaxQ * kMaxQ + kMaxQ;
  __shared__ double tot[NP];
  sum_partials<NP>(part, nblocks, tot);
  if (threadIdx.x != 0) return;
  const double* G = tot;
  const double* r = tot + kMaxQ * kMaxQ;
  double alpha[kMaxQ];
  for (int i = 0; i < m; ++i) {
    double d = r[i];
    for (int k = 0; k < i; ++k) d -= alpha[k] * G[i + kMaxQ * k];
    alpha[i] = (1.0 / G[i + kMaxQ * i]) * d;
  }
  const double gamma = r[0] / G[0];
  for (int a = 0; a < kMaxQ; ++a) {
    coef[a] = (a < m) ? -alpha[a] - (a == 0 ? gamma : 0.0) : 0.0;
    coef[kMaxQ + a] = (a < m) ? gamma * alpha[a] : 0.0;
  }
}

// x_tmp = x + U cu + V cv, then project_q! on the Q columns ...
template <typename T, int K>
__global__ __launch_bounds__(256) void k_lbqn_apply_q(int64_t I, int64_t N, const T* __restrict__ x,
                                                      const T* __restrict__ U, const T* __restrict__ V,
                                                      const double* __restrict__ coef, int m, T* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= I) return;
  T v[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int64_t n = (int64_t)K * i + k;
    T s = 0;
    for (int a = 0; a < m; ++a) s += U[n + N * a] * (T)coef[a] + V[n + N * a] * (T)coef[kMaxQ + a];
    v[k] = x[n] + s;
  }
  project_simplex<T, K>(v);
#pragma unroll
  for (int k = 0; k < K; ++k) out[(int64_t)K * i + k] = v[k];
}

// ... and project_p! on the P entries.
template <typename T>
__global__ __launch_bounds__(256) void k_lbqn_apply_p(int64_t n0, int64_t N, const T* __restrict__ x,
                                                      const T* __restrict__ U, const T* __restrict__ V,
                                                      const double* __restrict__ coef, int m, T* __restrict__ out) {
  for (int64_t n = n0 + (int64_t)blockIdx.x * 256 + threadIdx.x; n < N; n += (int64_t)gridDim.x * 256) {
    T s = 0;
    for (int a = 0; a < m; ++a) s += U[n + N * a] * (T)coef[a] + V[n + N * a] * (T)coef[kMaxQ + a];
    out[n] = project_box<T>(x[n] + s);
  }
}

template <typename T, int K>
static void lbqn_apply_q_k(int64_t I, int64_t N, const void* x, const void* U, const void* V, const double* coef,
                           int m, void* out, hipStream_t s) {
  hipLaunchKernelGGL((k_lbqn_apply_q<T, K>), dim3((unsigned)((I + 255) / 256)), dim3(256), 0, s, I, N, (const T*)x,
                     (const T*)U, (const T*)V, coef, m, (T*)out);
}

// The history columns are shifted by the caller (device copies) before this.
template <typename T>
void launch_lbqn(int K, int64_t I, int64_t J, const void* x, const void* xn, const void* xn2, void* U, void* V,
                 void* x_out, int m, double* scratch, int64_t scratch_len, hipStream_t s) {
  const int64_t N = (int64_t)K * (I + J);
  constexpr int NP = kMaxQ * kMaxQ + kMaxQ;
  int64_t nb = grid_for(N, 512);
  if ((nb * NP + 2 * kMaxQ) > scratch_len) nb = (scratch_len - 2 * kMaxQ) / NP;
  double* part = scratch;
  double* coef = scratch + nb * NP;
  hipLaunchKernelGGL((k_lbqn_dots<T>), dim3((unsigned)nb), dim3(256), 0, s, N, (const T*)x, (const T*)xn,
                     (const T*)xn2, (T*)U, (T*)V, m, part);
  hipLaunchKernelGGL(k_lbqn_solve, dim3(1), dim3(256), 0, s, part, (int)nb, m, coef);
  ADMX_DISPATCH_K(K, (lbqn_apply_q_k<T, KK>(I, N, x, U, V, coef, m, x_out, s)));
  const int64_t np = N - (int64_t)K * I;
  hipLaunchKernelGGL((k_lbqn_apply_p<T>), dim3(grid_for(np, 4096)), dim3(256), 0, s, (int64_t)K * I, N,
                     (const T*)x, (const T*)U, (const T*)V, coef, m, (T*)x_out);
}

template void launch_lbqn<double>(int, int64_t, int64_t, const void*, const void*, const void*, void*, void*, void*,
                                  int, double*, int64_t, hipStream_t);
template void launch_lbqn<float>(int, int64_t, int64_t, const void*, const void*, const void*, void*, void*, void*,
                                 int, double*, int64_t, hipStream_t);

// ------------------------------------------------------------------ accept rule
template <typename T>
__global__ __launch_bounds__(256) void k_accept(int64_t N, T* __restrict__ x, const T* __restrict__ x_tmp,
                                                const T* __restrict__ x_next2, double* __restrict__ lls) {
  const bool take_qn = lls[LL_PREV] < lls[LL_QN];
  const T* src = take_qn ? x_tmp : x_next2;
  for (int64_t n = (int64_t)blockIdx.x * 256 + threadIdx.x; n < N; n += (int64_t)gridDim.x * 256) x[n] = src[n];
  if (blockIdx.x == 0 && threadIdx.x == 0) lls[LL_NEW] = take_qn ? lls[LL_QN] : lls[LL_BASIC];
}

template <typename T>
void launch_accept(int64_t N, void* x, const void* x_tmp, const void* x_next2, double* lls, hipStream_t s) {
  hipLaunchKernelGGL((k_accept<T>), dim3(grid_for(N, 4096)), dim3(256), 0, s, N, (T*)x, (const T*)x_tmp,
                     (const T*)x_next2, lls);
}

template void launch_zal<double>(int, int64_t, int64_t, const void*, const void*, const void*, void*, void*, void*,
                                 int, int, int, double*, int64_t, hipStream_t);
template void launch_zal<float>(int, int64_t, int64_t, const void*, const void*, const void*, void*, void*, void*,
                                int, int, int, double*, int64_t, hipStream_t);
template void launch_accept<double>(int64_t, void*, const void*, const void*, double*, hipStream_t);
template void launch_accept<float>(int64_t, void*, const void*, const void*, double*, hipStream_t);

// ------------------------------------------------------------------ BED decode / counts
// Dense int8 genotypes (I x J column-major, -1 = missing), one lane per BED byte.
__global__ __launch_bounds__(256) void k_decode(const uint8_t* __restrict__ bed, int64_t pitch, int64_t I, int64_t J,
                                                int8_t* __restrict__ g) {
  const int64_t nb = (I + 3) >> 2;
  const int64_t b = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (b >= nb) return;
  for (int64_t j = blockIdx.y; j < J; j += gridDim.y) {
    const uint32_t byte = bed[pitch * j + b];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int64_t i = 4 * b + t;
      if (i < I) {
        const uint32_t c = (byte >> (2 * t)) & 3u;
        g[i + I * j] = (int8_t)(c == 1u ? -1 : code_g(c));
      }
    }
  }
}

void launch_decode(const uint8_t* bed, int64_t pitch, int64_t I, int64_t J, int8_t* g, hipStream_t s) {
  const int64_t nb = (I + 3) >> 2;
  dim3 grid((unsigned)((nb + 255) / 256), (unsigned)(J < 65535 ? J : 65535));
  hipLaunchKernelGGL(k_decode, grid, dim3(256), 0, s, bed, pitch, I, J, g);
}

// Valid-individual mask (2 bits per individual) of 32-bit word d of a column.
__device__ __forceinline__ uint32_t valid_mask(int64_t d, int64_t I) {
  const int64_t rem = I - 16 * d;
  if (rem >= 16) return 0xffffffffu;
  if (rem <= 0) return 0u;
  return (1u << (2 * rem)) - 1u;
}

// Per-SNP nonmissing count and A2 allele count: one wave per SNP column,
// 16 genotypes per 32-bit load, popcounts, wave reduction.
__global__ __launch_bounds__(256) void k_counts_snp(const uint8_t* __restrict__ bed, int64_t pitch, int64_t I,
                                                    int64_t J, unsigned long long* __restrict__ snp_nm,
                                                    unsigned long long* __restrict__ snp_alt) {
  const int lane = threadIdx.x & 63;
  const int64_t j = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (j >= J) return;
  const uint32_t* col = reinterpret_cast<const uint32_t*>(bed + pitch * j);
  const int64_t ndw = (I + 15) >> 4;
  unsigned long long nm = 0, alt = 0;
  for (int64_t d = lane; d < ndw; d += 64) {
    const uint32_t vm = valid_mask(d, I);
    const uint32_t w = col[d];
    const uint32_t lo = w & 0x55555555u & vm, hi = (w >> 1) & 0x55555555u & vm;
    const uint32_t nvalid = __popc(vm) >> 1;
    nm += nvalid - __popc(lo & ~hi);
    alt += __popc(hi) + __popc(hi & lo);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    nm += __shfl_xor(nm, off, 64);
    alt += __shfl_xor(alt, off, 64);
  }
  if (lane == 0) {
    snp_nm[j] = nm;
    snp_alt[j] = alt;
  }
}

// Per-individual nonmissing count: one lane per 32-bit word position (16
// individuals) streaming a chunk of SNPs, integer atomics (exact) at the end.
__global__ __launch_bounds__(256) void k_counts_ind(const uint8_t* __restrict__ bed, int64_t pitch, int64_t I,
                                                    int64_t J, int64_t jchunk, unsigned long long* __restrict__ ind_nm) {
  const int64_t d = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t ndw = (I + 15) >> 4;
  if (d >= ndw) return;
  const int64_t jb = (int64_t)blockIdx.y * jchunk;
  const int64_t je = (jb + jchunk < J) ? jb + jchunk : J;
  uint32_t cnt[16];
#pragma unroll
  for (int t = 0; t < 16; ++t) cnt[t] = 0;
  for (int64_t j = jb; j < je; ++j) {
    const uint32_t w = *reinterpret_cast<const uint32_t*>(bed + pitch * j + 4 * d);
    const uint32_t miss = w & ~(w >> 1) & 0x55555555u;
#pragma unroll
    for (int t = 0; t < 16; ++t) cnt[t] += 1u - ((miss >> (2 * t)) & 1u);
  }
#pragma unroll
  for (int t = 0; t < 16; ++t) {
    const int64_t i = 16 * d + t;
    if (i < I && je > jb) atomicAdd(&ind_nm[i], (unsigned long long)cnt[t]);
  }
}

void launch_counts_snp(const uint8_t* bed, int64_t pitch, int64_t I, int64_t J, unsigned long long* snp_nm,
                       unsigned long long* snp_alt, hipStream_t s) {
  hipLaunchKernelGGL(k_counts_snp, dim3((unsigned)((J + 3) / 4)), dim3(256), 0, s, bed, pitch, I, J, snp_nm, snp_alt);
}

void launch_counts(const uint8_t* bed, int64_t pitch, int64_t I, int64_t J, unsigned long long* snp_nm,
                   unsigned long long* snp_alt, unsigned long long* ind_nm, hipStream_t s) {
  hipLaunchKernelGGL(k_counts_snp, dim3((unsigned)((J + 3) / 4)), dim3(256), 0, s, bed, pitch, I, J, snp_nm, snp_alt);
  const int64_t ndw = (I + 15) >> 4;
  int64_t nchunk = (J + 255) / 256;
  if (nchunk > 1024) nchunk = 1024;
  if (nchunk < 1) nchunk = 1;
  const int64_t jchunk = (J + nchunk - 1) / nchunk;
  nchunk = (J + jchunk - 1) / jchunk;
  if (nchunk < 1) nchunk = 1;
  hipLaunchKernelGGL(k_counts_ind, dim3((unsigned)((ndw + 255) / 256), (unsigned)nchunk), dim3(256), 0, s, bed, pitch,
                     I, J, jchunk, ind_nm);
}

// SnpLinAlg's mean genotype of SNP j over its observed entries, in double,
// stored as T (0 for a SNP with no observed entry).
template <typename T>
__global__ __launch_bounds__(256) void k_snp_mean(const unsigned long long* __restrict__ nm,
                                                  const unsigned long long* __restrict__ alt, int64_t J,
                                                  T* __restrict__ mu) {
  const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (j < J) mu[j] = nm[j] ? (T)((double)alt[j] / (double)nm[j]) : (T)0;
}

template <typename T>
void launch_snp_mean(const unsigned long long* nm, const unsigned long long* alt, int64_t J, T* mu, hipStream_t s) {
  hipLaunchKernelGGL((k_snp_mean<T>), dim3((unsigned)((J + 255) / 256)), dim3(256), 0, s, nm, alt, J, mu);
}
template void launch_snp_mean<double>(const unsigned long long*, const unsigned long long*, int64_t, double*,
                                      hipStream_t);
template void launch_snp_mean<float>(const unsigned long long*, const unsigned long long*, int64_t, float*,
                                     hipStream_t);

// ------------------------------------------------------------------ synthetic genotypes
__device__ __forceinline__ uint64_t smix(uint64_t z) {
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
__device__ __forceinline__ double u01(uint64_t h) { return (double)(h >> 11) * 0x1.0p-53; }

// g_ij ~ Binomial(2, q_i . p_j) from two hashed uniforms, missing with
// probability miss_rate from a third; one lane per BED byte (4 individuals).
__global__ __launch_bounds__(256) void k_gen_bed(const double* __restrict__ Qt, const double* __restrict__ Pt,
                                                 int64_t I, int64_t j0, int64_t J, int K, uint64_t base,
                                                 double miss_rate, uint8_t* __restrict__ bed, int64_t pitch) {
  const int64_t b = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (b >= pitch) return;
  for (int64_t j = j0 + blockIdx.y; j < J; j += gridDim.y) {
    uint32_t byte = 0;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int64_t i = 4 * b + t;
      if (i < I) {
        double s = Qt[(int64_t)K * i] * Pt[(int64_t)K * j];
        for (int k = 1; k < K; ++k) s = __builtin_fma(Qt[k + (int64_t)K * i], Pt[k + (int64_t)K * j], s);
        const uint64_t ctr = (uint64_t)j * (uint64_t)I + (uint64_t)i;
        const double u1 = u01(smix(smix(3 * ctr + 0) ^ base));
        const double u2 = u01(smix(smix(3 * ctr + 1) ^ base));
        const double u3 = u01(smix(smix(3 * ctr + 2) ^ base));
        const int g = (u1 < s) + (u2 < s);
        uint32_t code = g == 0 ? 0u : (g == 1 ? 2u : 3u);
        if (u3 < miss_rate) code = 1u;
        byte |= code << (2 * t);
      }
    }
    bed[pitch * j + b] = (uint8_t)byte;
  }
}

void launch_gen_bed_rows(const double* Qt, const double* Pt, int64_t I, int64_t j0, int64_t j1, int K,
                         uint64_t seed, double miss_rate, uint8_t* bed, int64_t pitch, hipStream_t s) {
  uint64_t base = seed ^ 0x5851F42D4C957F2DULL;
  // smix on the host: identical arithmetic to the device/CPU definition.
  base += 0x9E3779B97F4A7C15ULL;
  base = (base ^ (base >> 30)) * 0xBF58476D1CE4E5B9ULL;
  base = (base ^ (base >> 27)) * 0x94D049BB133111EBULL;
  base = base ^ (base >> 31);
  const int64_t rows = j1 - j0;
  if (rows <= 0) return;
  dim3 grid((unsigned)((pitch + 255) / 256), (unsigned)(rows < 65535 ? rows : 65535));
  hipLaunchKernelGGL(k_gen_bed, grid, dim3(256), 0, s, Qt, Pt, I, j0, j1, K, base, miss_rate, bed, pitch);
}

void launch_gen_bed(const double* Qt, const double* Pt, int64_t I, int64_t J, int K, uint64_t seed, double miss_rate,
                    uint8_t* bed, int64_t pitch, hipStream_t s) {
  launch_gen_bed_rows(Qt, Pt, I, 0, J, K, seed, miss_rate, bed, pitch, s);
}

// ------------------------------------------------------------------ log-likelihood state check
// The FP64 log-likelihood pass multiplies s^g (1-s)^(2-g) into a product; the
// reference sums m [g log s + (2-g) log(1-s)] (src/loops.jl:517-539), whose
// terms 0 log 0 are NaN when s reaches exactly 0 or 1 (FRAPPE EM drives p to 0
// or 1 on a SNP whose observed genotypes are all 0 or all 2).  The pass takes
// the term-by-term form whenever this check cannot rule that out.
__global__ void k_flag_clear(int* flag) { *flag = 0; }

// The product form needs 0 < s < 1 for every s = q.p: q >= 0 with column sums
// in [1e-200, 1 + eps] and p in [1e-12, 1 - delta] give s <= (1 + eps)(1 - delta)
// < 1 for eps < delta.  FP64 states: eps = 1e-13, delta = 1e-12.  FP32 states
// (an FP32 context's q and p, widened): a projected column's FP32 sum is 1 only
// to a few 1e-7 (up to K = 16 roundings), so eps = 4e-6 and delta = 8e-6 (the
// FP32 projection's bound 1 - 1e-5 is 1.0014e-5 from 1).
__global__ __launch_bounds__(256) void k_ll_state_check(int K, const double* __restrict__ q, int64_t I,
                                                        const double* __restrict__ p, int64_t J, double eps,
                                                        double delta, int* flag) {
  bool bad = false;
  const int64_t stride = (int64_t)gridDim.x * 256;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < I; i += stride) {
    double sum = 0.0;
    for (int k = 0; k < K; ++k) {
      const double v = q[(int64_t)K * i + k];
      bad |= !(v >= 0.0 && v <= 1e300);  // also NaN
      sum += v;
    }
    bad |= !(sum >= 1e-200 && sum <= 1.0 + eps);
  }
  const int64_t np = (int64_t)K * J;
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < np; e += stride) {
    const double v = p[e];
    bad |= !(v >= 1e-12 && v <= 1.0 - delta);
  }
  if (__syncthreads_or(bad) && threadIdx.x == 0) atomicOr(flag, 1);
}

void launch_ll_state_check(int K, const double* q, int64_t I, const double* p, int64_t J, int* flag, hipStream_t s,
                           bool f32_state) {
  hipLaunchKernelGGL(k_flag_clear, dim3(1), dim3(1), 0, s, flag);
  const int64_t n = std::max<int64_t>(I, (int64_t)K * J);
  const unsigned blocks = (unsigned)std::min<int64_t>(std::max<int64_t>((n + 255) / 256, 1), 1024);
  const double eps = f32_state ? 4e-6 : 1e-13, delta = f32_state ? 8e-6 : 1e-12;
  hipLaunchKernelGGL(k_ll_state_check, dim3(blocks), dim3(256), 0, s, K, q, I, p, J, eps, delta, flag);
}

// FP64 copy of an FP32 array (an FP32 context's q or p for the FP64 passes).
__global__ __launch_bounds__(256) void k_widen(const float* __restrict__ x, int64_t n, double* __restrict__ y) {
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < n; e += (int64_t)gridDim.x * 256) y[e] = (double)x[e];
}

void launch_widen(const float* x, int64_t n, double* y, hipStream_t s) {
  if (n <= 0) return;
  const int64_t blocks = std::min<int64_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_widen, dim3((unsigned)blocks), dim3(256), 0, s, x, n, y);
}

// BED rows as a .bed file holds them (ibytes per row, back to back) into the
// device layout (pitch per row, bytes past ibytes zero); rows from `real` on
// are the missing code.  One lane per 32-bit word of the destination.
__global__ __launch_bounds__(256) void k_repitch(const uint8_t* __restrict__ raw, int64_t ibytes, int64_t real,
                                                 int64_t rows, uint8_t* __restrict__ dst, int64_t pitch) {
  const int64_t wpr = pitch >> 2, n = rows * wpr;
  for (int64_t w = (int64_t)blockIdx.x * 256 + threadIdx.x; w < n; w += (int64_t)gridDim.x * 256) {
    const int64_t r = w / wpr, b0 = (w - r * wpr) * 4;
    uint32_t v = 0x55555555u;
    if (r < real) {
      const uint8_t* __restrict__ src = raw + r * ibytes;
      v = 0;
#pragma unroll
      for (int e = 0; e < 4; ++e)
        if (b0 + e < ibytes) v |= (uint32_t)src[b0 + e] << (8 * e);
    }
    reinterpret_cast<uint32_t*>(dst + r * pitch)[b0 >> 2] = v;
  }
}

void launch_repitch(const uint8_t* raw, int64_t ibytes, int64_t real, int64_t rows, uint8_t* dst, int64_t pitch,
                    hipStream_t s) {
  const int64_t n = rows * (pitch >> 2);
  if (n <= 0) return;
  const int64_t blocks = std::min<int64_t>((n + 255) / 256, 65536);
  hipLaunchKernelGGL(k_repitch, dim3((unsigned)blocks), dim3(256), 0, s, raw, ibytes, real, rows, dst, pitch);
}

// ------------------------------------------------------------------ cross-validation folds
// Fold of genotype entry (i, j) for an nfolds-fold holdout: the entry counter
// j*I + i hashed with the fold seed, mapped to [0, nfolds) by a multiply-shift
// of the top 32 hash bits.  Every entry belongs to exactly one fold.
__device__ __forceinline__ int fold_of(uint64_t ctr, uint64_t base, int nfolds) {
  return (int)(((smix(ctr ^ base) >> 32) * (uint64_t)nfolds) >> 32);
}

// dst = src with every entry of fold `fold` set to missing (01); fold < 0 is a
// plain copy.  One lane per BED byte (4 individuals); both BEDs share the pitch.
__global__ __launch_bounds__(256) void k_mask_fold(const uint8_t* __restrict__ src, uint8_t* __restrict__ dst,
                                                   int64_t pitch, int64_t I, int64_t J, int fold, int nfolds,
                                                   uint64_t base) {
  const int64_t b = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (b >= pitch) return;
  for (int64_t j = blockIdx.y; j < J; j += gridDim.y) {
    uint32_t byte = src[pitch * j + b];
    if (fold >= 0) {
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int64_t i = 4 * b + t;
        if (i < I && fold_of((uint64_t)j * (uint64_t)I + (uint64_t)i, base, nfolds) == fold)
          byte = (byte & ~(3u << (2 * t))) | (1u << (2 * t));
      }
    }
    dst[pitch * j + b] = (uint8_t)byte;
  }
}

// Held-out binomial deviance of the fit (q, p) on the observed entries of fold
// `fold` of the unmasked BED:
//   dev = 2 [ g log(g / 2s) + (2-g) log((2-g) / 2(1-s)) ],  0 log 0 = 0,
// summed in FP64 with the entry count.  One lane per individual, a block per
// (256 individuals, SNP chunk); per-block partials [2][nblocks].
__global__ __launch_bounds__(256) void k_heldout_dev(const uint8_t* __restrict__ bed, int64_t pitch, int64_t I,
                                                     int64_t J, int K, const double* __restrict__ q,
                                                     const double* __restrict__ p, int64_t jchunk, int fold,
                                                     int nfolds, uint64_t base, double* __restrict__ part) {
  __shared__ double red[256];
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t jb = (int64_t)blockIdx.y * jchunk;
  const int64_t je = (jb + jchunk < J) ? jb + jchunk : J;
  double dev = 0.0, cnt = 0.0;
  if (i < I) {
    const uint32_t sh = (uint32_t)(i & 3) * 2u;
    for (int64_t j = jb; j < je; ++j) {
      const uint32_t c = (bed[pitch * j + (i >> 2)] >> sh) & 3u;
      if (c == 1u || fold_of((uint64_t)j * (uint64_t)I + (uint64_t)i, base, nfolds) != fold) continue;
      double s = 0.0;
      for (int k = 0; k < K; ++k) s = __builtin_fma(q[k + (int64_t)K * i], p[k + (int64_t)K * j], s);
      const int g = code_g(c);
      double d = 0.0;
      if (g > 0) d += (double)g * log((double)g / (2.0 * s));
      if (g < 2) d += (double)(2 - g) * log((double)(2 - g) / (2.0 * (1.0 - s)));
      dev += 2.0 * d;
      cnt += 1.0;
    }
  }
  const int64_t nb = (int64_t)gridDim.x * gridDim.y;
  const int64_t blk = (int64_t)blockIdx.y * gridDim.x + blockIdx.x;
  const double tdev = block_sum<256>(dev, red);
  const double tcnt = block_sum<256>(cnt, red);
  if (threadIdx.x == 0) {
    part[blk] = tdev;
    part[nb + blk] = tcnt;
  }
}

static uint64_t fold_base(uint64_t seed) {
  uint64_t z = seed ^ 0x2545F4914F6CDD1DULL;  // smix on the host, as on the device
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

void launch_mask_fold(const uint8_t* src, uint8_t* dst, int64_t pitch, int64_t I, int64_t J, int fold, int nfolds,
                      uint64_t seed, hipStream_t s) {
  dim3 grid((unsigned)((pitch + 255) / 256), (unsigned)(J < 65535 ? J : 65535));
  hipLaunchKernelGGL(k_mask_fold, grid, dim3(256), 0, s, src, dst, pitch, I, J, fold, nfolds, fold_base(seed));
}

int64_t heldout_dev_blocks(int64_t I, int64_t J, int64_t* jchunk) {
  const int64_t bx = (I + 255) / 256;
  int64_t ny = (16384 + bx - 1) / bx;
  if (ny > J) ny = J;
  if (ny < 1) ny = 1;
  *jchunk = (J + ny - 1) / ny;
  ny = (J + *jchunk - 1) / *jchunk;
  return bx * ny;
}

void launch_heldout_dev(const uint8_t* bed, int64_t pitch, int64_t I, int64_t J, int K, const double* q,
                        const double* p, int fold, int nfolds, uint64_t seed, double* part, hipStream_t s) {
  int64_t jchunk = 0;
  const int64_t nb = heldout_dev_blocks(I, J, &jchunk);
  const int64_t bx = (I + 255) / 256;
  dim3 grid((unsigned)bx, (unsigned)(nb / bx));
  hipLaunchKernelGGL(k_heldout_dev, grid, dim3(256), 0, s, bed, pitch, I, J, K, q, p, jchunk, fold, nfolds,
                     fold_base(seed), part);
}

}  // namespace admx
