// admx_mlp.hip -- the learned warm start: genotype-projection features and a
// small FP32 multilayer perceptron (forward, backward, Adam) on the device.
//
// Features (per individual i, over the first Jp SNPs of a reference panel):
//   f_i = (1/sqrt(Jp)) sum_{j observed} (g_ij - mu_j) * isd_j * R[j, :]      (64 values)
// with the panel's mean mu_j = 2 f_j and isd_j = 1/sqrt(2 f_j (1-f_j)), R a
// fixed random Jp x 64 matrix; missing entries contribute 0.
//
// Network (row-major activations, batch rows):
//   H1 = relu(X W1' + b1), H2 = relu(H1 W2' + b2), Z = H2 W3' + b3, Y = softmax(Z)
// with W_l stored out x in.  Loss: soft-target cross entropy
//   L = -(1/B) sum_b sum_k T[b,k] log Y[b,k],  dZ = (Y - T)/B.
// Parameters, gradients and Adam moments live in one flat float vector laid
// out [W1, b1, W2, b2, W3, b3].  Every reduction is a fixed-order tree, so
// training is bitwise reproducible.
#include "admx_common.hpp"
#include "admx_internal.hpp"

namespace admx {

// ------------------------------------------------------------------ features
// F = Z R over the panel SNPs, Z[i][j] = (g_ij - mu_j) isd_j (0 if missing):
// a GEMM on the matrix cores in FP32 (v_mfma_f32_32x32x2f32: products exact in
// FP32, FP32 accumulation).  A wave owns 64 individuals (two 32-row tiles) and
// all 64 features (two 32-column tiles) and walks its SNP chunk two SNPs per
// product: A[i][k] (lane i + 32k) is Z of individual i at SNP j + k, B[k][n]
// (lane n + 32k) is R[j + k][n], D[i][n] in lane n + 32((i/4)%2), register
// (i%4) + 4(i/8).  The workgroup's four waves share the chunk's R rows through
// LDS, 16 SNPs (4 KB) per block, double-buffered with one barrier per block;
// each lane's genotype bytes of the next block are loaded while the current
// one is multiplied.  Partial feature vectors per chunk [chunk][i][64], reduced
// in chunk order by k_feat_reduce.
typedef float f32x16 __attribute__((ext_vector_type(16)));
__global__ __launch_bounds__(256) void k_features(const uint8_t* __restrict__ bed, int64_t pitch, int64_t I,
                                                  int64_t Jp, const float* __restrict__ R,
                                                  const float* __restrict__ mu, const float* __restrict__ isd,
                                                  int64_t jchunk, float* __restrict__ part) {
  __shared__ __align__(16) float rs[2][16][kFeat];
  const int t = threadIdx.x, lane = t & 63, h = lane >> 5, cn = lane & 31;
  const int64_t i0 = ((int64_t)blockIdx.x * 4 + (t >> 6)) * 64;
  const int64_t jb = (int64_t)blockIdx.y * jchunk;
  const int64_t je = (jb + jchunk < Jp) ? jb + jchunk : Jp;
  // genotype byte addresses of the lane's two individuals (clamped past I: their rows are not written)
  const uint8_t* __restrict__ g0 = bed + ((i0 + cn < I ? i0 + cn : I - 1) >> 2);
  const uint8_t* __restrict__ g1 = bed + ((i0 + 32 + cn < I ? i0 + 32 + cn : I - 1) >> 2);
  const uint32_t sh0 = 2u * (uint32_t)((i0 + cn) & 3), sh1 = 2u * (uint32_t)((i0 + 32 + cn) & 3);
  // R block of SNPs [j, j+16): thread t moves 16 bytes of row j + t/16 (rows past Jp clamped; their Z is 0)
  auto rload = [&](int64_t j) {
    const int64_t r = (j + (t >> 4) < Jp) ? j + (t >> 4) : Jp - 1;
    return *reinterpret_cast<const float4*>(R + r * kFeat + 4 * (t & 15));
  };
  // genotype bytes of SNPs j + 2u + h, u = 0..7 (rows up to je + 14 <= J + 14; the BED holds J + 16 rows)
  uint32_t b0[8], b1[8];
  auto gload = [&](int64_t j) {
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      b0[u] = g0[pitch * (j + 2 * u + h)];
      b1[u] = g1[pitch * (j + 2 * u + h)];
    }
  };
  f32x16 acc[2][2] = {{(f32x16){}, (f32x16){}}, {(f32x16){}, (f32x16){}}};
  float4 rn = rload(jb);
  *reinterpret_cast<float4*>(&rs[0][t >> 4][4 * (t & 15)]) = rn;
  gload(jb);
  __syncthreads();
  int buf = 0;
  for (int64_t j = jb; j < je; j += 16) {
    uint32_t c0[8], c1[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      c0[u] = b0[u];
      c1[u] = b1[u];
    }
    const bool more = j + 16 < je;
    if (more) {
      rn = rload(j + 16);
      gload(j + 16);
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int64_t jj = j + 2 * u + h;
      const bool live = jj < je;
      const int64_t jc = live ? jj : je - 1;
      const float m = mu[jc], sd = isd[jc];
      const uint32_t k0 = (c0[u] >> sh0) & 3u, k1 = (c1[u] >> sh1) & 3u;
      const float z0 = (live && k0 != 1u) ? ((float)code_g(k0) - m) * sd : 0.f;
      const float z1 = (live && k1 != 1u) ? ((float)code_g(k1) - m) * sd : 0.f;
#pragma unroll
      for (int nt = 0; nt < 2; ++nt) {
        const float bv = rs[buf][2 * u + h][32 * nt + cn];
        acc[0][nt] = __builtin_amdgcn_mfma_f32_32x32x2f32(z0, bv, acc[0][nt], 0, 0, 0);
        acc[1][nt] = __builtin_amdgcn_mfma_f32_32x32x2f32(z1, bv, acc[1][nt], 0, 0, 0);
      }
    }
    if (more) *reinterpret_cast<float4*>(&rs[buf ^ 1][t >> 4][4 * (t & 15)]) = rn;
    __syncthreads();
    buf ^= 1;
  }
#pragma unroll
  for (int T = 0; T < 2; ++T)
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int64_t i = i0 + 32 * T + 8 * (r >> 2) + 4 * h + (r & 3);
      if (i < I) {
        float* __restrict__ out = part + ((int64_t)blockIdx.y * I + i) * kFeat + cn;
        out[0] = acc[T][0][r];
        out[32] = acc[T][1][r];
      }
    }
}

__global__ __launch_bounds__(256) void k_feat_reduce(const float* __restrict__ part, int64_t n, int nchunk,
                                                     float scale, float* __restrict__ F) {
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < n; e += (int64_t)gridDim.x * 256) {
    float s = 0.f;
    for (int c = 0; c < nchunk; ++c) s += part[(int64_t)c * n + e];
    F[e] = s * scale;
  }
}

// ------------------------------------------------------------------ FP32 GEMM
// C[M x N] = op(A) op(B) (+ bias[n]) (relu) (* (mask > 0)) (+= C if accumulate),
// row-major; op(A) is M x Kd (A is Kd x M when TA), op(B) is Kd x N (B is N x Kd
// when TB).  64 x 64 tile per 256-lane block, Kd in steps of 16 through LDS;
// each wave multiplies one 32 x 32 quarter of the tile on the matrix cores
// (v_mfma_f32_32x32x2f32: FP32 products, FP32 accumulation), A[i][k] in lane
// i + 32k, B[k][n] in lane n + 32k, D[i][n] in lane n + 32((i/4)%2), register
// (i%4) + 4(i/8).
template <bool TA, bool TB>
__global__ __launch_bounds__(256) void k_sgemm(int M, int N, int Kd, const float* __restrict__ A, int lda,
                                               const float* __restrict__ B, int ldb, float* __restrict__ C, int ldc,
                                               const float* __restrict__ bias, int relu,
                                               const float* __restrict__ mask, int accumulate) {
  __shared__ float As[16][64 + 4];
  __shared__ float Bs[16][64 + 4];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, h = lane >> 5, cn = lane & 31;
  const int wm = 32 * (w >> 1), wn = 32 * (w & 1);  // the wave's quarter of the tile
  const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
  f32x16 acc = (f32x16){};
  for (int k0 = 0; k0 < Kd; k0 += 16) {
    // 64 x 16 tiles of op(A) and op(B): 1024 values, 4 per lane; consecutive
    // lanes walk the dimension that is contiguous in memory (k for A and for a
    // transposed B, m for a transposed A, n for B), so every load is coalesced
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int idx = threadIdx.x + 256 * r;
      const int mm = TA ? (idx & 63) : (idx >> 4), ka = TA ? (idx >> 6) : (idx & 15);  // A tile as [m][k]
      const int gm = m0 + mm, gka = k0 + ka;
      float av = 0.f;
      if (gm < M && gka < Kd) av = TA ? A[(int64_t)gka * lda + gm] : A[(int64_t)gm * lda + gka];
      As[ka][mm] = av;
      const int nn = TB ? (idx >> 4) : (idx & 63), kb = TB ? (idx & 15) : (idx >> 6);  // B tile as [k][n]
      const int gn = n0 + nn, gkb = k0 + kb;
      float bv = 0.f;
      if (gn < N && gkb < Kd) bv = TB ? B[(int64_t)gn * ldb + gkb] : B[(int64_t)gkb * ldb + gn];
      Bs[kb][nn] = bv;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < 16; kk += 2)
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(As[kk + h][wm + cn], Bs[kk + h][wn + cn], acc, 0, 0, 0);
    __syncthreads();
  }
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int gm = m0 + wm + 8 * (r >> 2) + 4 * h + (r & 3), gn = n0 + wn + cn;
    if (gm >= M || gn >= N) continue;
    float val = acc[r];
    if (bias) val += bias[gn];
    if (relu) val = val > 0.f ? val : 0.f;
    if (mask && !(mask[(int64_t)gm * ldc + gn] > 0.f)) val = 0.f;
    float* cp = C + (int64_t)gm * ldc + gn;
    *cp = accumulate ? *cp + val : val;
  }
}

// db[n] = sum_b dZ[b, n] in a fixed order: a block owns 32 columns, thread
// (r, c) sums rows r, r + 8, r + 16, ... of column c (coalesced 128-byte rows,
// 8 independent load streams per column instead of one), then the 8 slices
// are added in slice order through LDS.
__global__ __launch_bounds__(256) void k_colsum(const float* __restrict__ X, int M, int N, float* __restrict__ out) {
  __shared__ float part[8][32];
  const int c = threadIdx.x & 31, r = threadIdx.x >> 5;
  const int n = blockIdx.x * 32 + c;
  float s = 0.f;
  if (n < N)
    for (int m = r; m < M; m += 8) s += X[(int64_t)m * N + n];
  part[r][c] = s;
  __syncthreads();
  if (r == 0 && n < N) {
    float t = part[0][c];
#pragma unroll
    for (int k = 1; k < 8; ++k) t += part[k][c];
    out[n] = t;
  }
}

// Softmax rows of Z (M x K, K <= 16) into Y; with targets T, dZ = (Y - T)/M and
// per-row loss -sum_k T log Y into rowloss.
__global__ __launch_bounds__(256) void k_softmax_ce(const float* __restrict__ Z, int M, int K,
                                                    const float* __restrict__ T, float* __restrict__ Y,
                                                    float* __restrict__ dZ, float* __restrict__ rowloss) {
  const int m = blockIdx.x * 256 + threadIdx.x;
  if (m >= M) return;
  const float* z = Z + (int64_t)m * K;
  float mx = z[0];
  for (int k = 1; k < K; ++k) mx = fmaxf(mx, z[k]);
  float e[16], s = 0.f;
  for (int k = 0; k < K; ++k) {
    e[k] = expf(z[k] - mx);
    s += e[k];
  }
  const float is = 1.f / s, ls = logf(s);
  float loss = 0.f;
  for (int k = 0; k < K; ++k) {
    const float y = e[k] * is;
    if (Y) Y[(int64_t)m * K + k] = y;
    if (T) {
      const float t = T[(int64_t)m * K + k];
      if (dZ) dZ[(int64_t)m * K + k] = (y - t) / (float)M;
      if (t > 0.f) loss -= t * (z[k] - mx - ls);
    }
  }
  if (rowloss) rowloss[m] = loss;
}

// Gather rows idx[0..n) of X (ld columns) into Y.
__global__ __launch_bounds__(256) void k_gather_rows(const float* __restrict__ X, int ld,
                                                     const int32_t* __restrict__ idx, int n, float* __restrict__ Y) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= (int64_t)n * ld) return;
  const int64_t r = e / ld, c = e - r * ld;
  Y[e] = X[(int64_t)idx[r] * ld + c];
}

// Adam (Kingma & Ba 2015) with bias correction: m = b1 m + (1-b1) g,
// v = b2 v + (1-b2) g^2, w -= lr * (m/c1) / (sqrt(v/c2) + eps).
__global__ __launch_bounds__(256) void k_adam(int64_t n, float* __restrict__ w, const float* __restrict__ g,
                                              float* __restrict__ m, float* __restrict__ v, float lr, float b1,
                                              float b2, float c1, float c2, float eps) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= n) return;
  const float gv = g[e];
  const float mv = b1 * m[e] + (1.f - b1) * gv;
  const float vv = b2 * v[e] + (1.f - b2) * gv * gv;
  m[e] = mv;
  v[e] = vv;
  w[e] -= lr * (mv / c1) / (sqrtf(vv / c2) + eps);
}

// Row-major I x K probabilities -> the FP64 q block of x (identical layout:
// q[k,i] at k + K*i).
__global__ __launch_bounds__(256) void k_probs_to_q(const float* __restrict__ Y, int64_t n, double* __restrict__ q) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e < n) q[e] = (double)Y[e];
}

// p[k,j] = observed allele frequency of SNP j, alt / max(2 nonmissing, 1),
// clipped to [1e-5, 1 - 1e-5], for every k (the warm start's first p).
// Also the log-likelihood of every SNP at p_kj = f_j for all k, which any q on
// the simplex gives (s_ij = q_i . p_j = f_j): over SNP j's nonmissing
// genotypes, sum g log f + (2 - g) log(1 - f) = alt log f + (2 nm - alt) log(1 - f).
// One partial per workgroup, summed in a fixed order by the caller.
__global__ __launch_bounds__(256) void k_freq_to_p(const unsigned long long* __restrict__ nm,
                                                   const unsigned long long* __restrict__ alt, int64_t J, int K,
                                                   double* __restrict__ p, double* __restrict__ ll_part) {
  __shared__ double sh[256];
  const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
  double term = 0.0;
  if (j < J) {
    const unsigned long long d = 2ull * nm[j];
    double f = (double)alt[j] / (double)(d > 0 ? d : 1ull);
    f = fmin(fmax(f, 1e-5), 1.0 - 1e-5);
    for (int k = 0; k < K; ++k) p[(int64_t)K * j + k] = f;
    const double a = (double)alt[j], b = (double)(d - alt[j]);
    term = a * log(f) + b * log(1.0 - f);
  }
  if (ll_part) {
    const double tot = block_sum<256>(term, sh);
    if (threadIdx.x == 0) ll_part[blockIdx.x] = tot;
  }
}

// ------------------------------------------------------------------ launchers
static unsigned blocks_for(int64_t n) { return (unsigned)((n + 255) / 256); }

void launch_freq_to_p(const unsigned long long* nm, const unsigned long long* alt, int64_t J, int K, double* p,
                      double* ll_part, hipStream_t s) {
  hipLaunchKernelGGL(k_freq_to_p, dim3(blocks_for(J)), dim3(256), 0, s, nm, alt, J, K, p, ll_part);
}

void launch_features(const uint8_t* bed, int64_t pitch, int64_t I, int64_t Jp, const float* R, const float* mu,
                     const float* isd, int nchunk, int64_t jchunk, float* part, float* F, hipStream_t s) {
  dim3 grid((unsigned)((I + 255) / 256), (unsigned)nchunk);
  hipLaunchKernelGGL(k_features, grid, dim3(256), 0, s, bed, pitch, I, Jp, R, mu, isd, jchunk, part);
  const int64_t n = I * kFeat;
  const unsigned nb = blocks_for(n) < 8192u ? blocks_for(n) : 8192u;
  hipLaunchKernelGGL(k_feat_reduce, dim3(nb), dim3(256), 0, s, part, n, nchunk, 1.0f / sqrtf((float)Jp), F);
}

void launch_sgemm(bool ta, bool tb, int M, int N, int Kd, const float* A, int lda, const float* B, int ldb, float* C,
                  int ldc, const float* bias, bool relu, const float* mask, bool accumulate, hipStream_t s) {
  dim3 grid((unsigned)((N + 63) / 64), (unsigned)((M + 63) / 64));
  if (!ta && tb)
    hipLaunchKernelGGL((k_sgemm<false, true>), grid, dim3(256), 0, s, M, N, Kd, A, lda, B, ldb, C, ldc, bias,
                       (int)relu, mask, (int)accumulate);
  else if (!ta && !tb)
    hipLaunchKernelGGL((k_sgemm<false, false>), grid, dim3(256), 0, s, M, N, Kd, A, lda, B, ldb, C, ldc, bias,
                       (int)relu, mask, (int)accumulate);
  else if (ta && !tb)
    hipLaunchKernelGGL((k_sgemm<true, false>), grid, dim3(256), 0, s, M, N, Kd, A, lda, B, ldb, C, ldc, bias,
                       (int)relu, mask, (int)accumulate);
  else
    hipLaunchKernelGGL((k_sgemm<true, true>), grid, dim3(256), 0, s, M, N, Kd, A, lda, B, ldb, C, ldc, bias,
                       (int)relu, mask, (int)accumulate);
}

void launch_colsum(const float* X, int M, int N, float* out, hipStream_t s) {
  hipLaunchKernelGGL(k_colsum, dim3((unsigned)((N + 31) / 32)), dim3(256), 0, s, X, M, N, out);
}

void launch_softmax_ce(const float* Z, int M, int K, const float* T, float* Y, float* dZ, float* rowloss,
                       hipStream_t s) {
  hipLaunchKernelGGL(k_softmax_ce, dim3(blocks_for(M)), dim3(256), 0, s, Z, M, K, T, Y, dZ, rowloss);
}

void launch_gather_rows(const float* X, int ld, const int32_t* idx, int n, float* Y, hipStream_t s) {
  hipLaunchKernelGGL(k_gather_rows, dim3(blocks_for((int64_t)n * ld)), dim3(256), 0, s, X, ld, idx, n, Y);
}

void launch_adam(int64_t n, float* w, const float* g, float* m, float* v, float lr, float b1, float b2, float c1,
                 float c2, float eps, hipStream_t s) {
  hipLaunchKernelGGL(k_adam, dim3(blocks_for(n)), dim3(256), 0, s, n, w, g, m, v, lr, b1, b2, c1, c2, eps);
}

void launch_probs_to_q(const float* Y, int64_t n, double* q, hipStream_t s) {
  hipLaunchKernelGGL(k_probs_to_q, dim3(blocks_for(n)), dim3(256), 0, s, Y, n, q);
}

}  // namespace admx
