// admx_project.hpp -- device projections onto the probability simplex and
// onto the box [pc, 1-pc] (src/projections.jl), shared by the QP step kernels
// and the quasi-Newton step kernels.
#pragma once
#include "admx_common.hpp"

namespace admx {

// project_q!(b; pseudocount=1e-5) on one column held in registers
// (src/projections.jl:27-63).  Only the sorted values enter the threshold, so
// an odd-even transposition network replaces the reference's index quicksort.
template <typename T, int K>
__device__ __forceinline__ void project_simplex(T (&v)[K]) {
#pragma clang fp contract(off)
  const T pc = (T)1e-5;
  const T tau = (T)1 - (T)K * pc;
  T s[K];
#pragma unroll
  for (int i = 0; i < K; ++i) {
    v[i] -= pc;
    s[i] = v[i];
  }
#pragma unroll
  for (int pass = 0; pass < K; ++pass) {
#pragma unroll
    for (int i = pass & 1; i + 1 < K; i += 2) {
      const T hi = s[i] < s[i + 1] ? s[i + 1] : s[i];
      const T lo = s[i] < s[i + 1] ? s[i] : s[i + 1];
      s[i] = hi;
      s[i + 1] = lo;
    }
  }
  T tsum = 0, tmax = 0;
  bool found = false;
#pragma unroll
  for (int i = 0; i + 1 < K; ++i) {
    if (!found) {
      tsum += s[i];
      const T tm = (tsum - tau) / (T)(i + 1);
      if (tm >= s[i + 1]) {
        found = true;
        tmax = tm;
      }
    }
  }
  if (!found) tmax = (tsum + s[K - 1] - tau) / (T)K;
#pragma unroll
  for (int i = 0; i < K; ++i) {
    const T d = v[i] - tmax;
    v[i] = (d > (T)0 ? d : (T)0) + pc;
  }
}

// project_p!(b; pseudocount=1e-5)  (src/projections.jl:69-74)
template <typename T>
__device__ __forceinline__ T project_box(T b) {
  const T pc = (T)1e-5;
  const T v = b > pc ? b : pc;
  const T hi = (T)1 - pc;
  return v < hi ? v : hi;
}

}  // namespace admx
