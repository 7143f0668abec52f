// admx_noalloc.hip -- the log-likelihood pass for callers that cannot allocate
// device memory (HIP graph capture, serving loops): every input and the output
// come from the caller, and the pass's scratch (one partial sum per
// (256-individual block, SNP chunk) plus the state-check flag) lives in a
// fixed file-scope device array instead of the context's grown buffers.
//
// The kernels and the chunking are those of admx_loglikelihood /
// admx_loglikelihood_loop on a fresh context (default waves per pass), so the
// result is the same double bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

#include "admixhip.h"
#include "admx_host.hpp"
#include "admx_internal.hpp"

using namespace admx;
using namespace admx_host;

// Partial sums in doubles [0, ADMX_NOALLOC_LL_PARTS), then the int flag of the
// log-likelihood state check.  Each GPU holds its own copy for the life of
// the process.
static __device__ __align__(256) double scratch[ADMX_NOALLOC_LL_PARTS + 32];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
double* g_ptr[kMaxDevices] = {};  // device address of `scratch` on each device (nullptr: not resolved yet)

// Waves per pass of a fresh context (admx_ctx::target_waves) and the Q-side
// chunk count it gives (qside_chunks in admx_capi.cpp).
constexpr int64_t kDefaultWaves = 65536;
int ll_chunks(int64_t I, int64_t J, int64_t* chunk) {
  const int64_t waves = (I + 63) / 64;
  int64_t n = (kDefaultWaves + waves - 1) / std::max<int64_t>(waves, 1);
  n = std::max<int64_t>(1, std::min<int64_t>(n, std::max<int64_t>(1, J / 64)));
  *chunk = std::max<int64_t>(1, (J + n - 1) / n);
  return (int)std::max<int64_t>(1, (J + *chunk - 1) / *chunk);
}

}  // namespace

namespace admx {

// Resolves the device address of `scratch` on the current device, once per
// device.  Called when a context is created and on the entry point's first
// call on a device.
int noalloc_setup(int device, double** out) {
  if (device < 0 || device >= kMaxDevices) return fail(ADMX_EINVAL, "device index beyond the supported 64");
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(ADMX_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_ptr[device] = static_cast<double*>(p);
  }
  if (out) *out = g_ptr[device];
  return ADMX_OK;
}

}  // namespace admx

extern "C" int admx_loglikelihood_noalloc(const uint8_t* bed, int64_t pitch, int64_t I, int64_t J, int64_t K,
                                          const double* q, const double* p, double* out, void* stream) {
  return guarded([&]() -> int {
    if (!bed || !q || !p || !out) return bad("null pointer");
    if (I < 1 || J < 1) return bad("I and J must be >= 1");
    if (K < 1 || K > kMaxK) return bad("K must be in 1..16");
    if (pitch < (I + 3) / 4 || pitch % 16 != 0) return bad("pitch must be a multiple of 16 and >= ceil(I/4)");
    int64_t chunk = 0;
    const int nchunk = ll_chunks(I, J, &chunk);
    const int64_t nbx = (I + 255) / 256;
    if (nbx * nchunk > ADMX_NOALLOC_LL_PARTS)
      return bad("cohort too large for the fixed scratch of admx_loglikelihood_noalloc "
                 "(ceil(I/256) x SNP chunks > ADMX_NOALLOC_LL_PARTS)");
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    double* sp = nullptr;
    CHK(noalloc_setup(dev, &sp));
    hipStream_t s = static_cast<hipStream_t>(stream);
    int* flag = reinterpret_cast<int*>(sp + ADMX_NOALLOC_LL_PARTS);
    // the flag is cleared by a kernel (k_flag_clear) inside the check
    launch_ll_state_check((int)K, q, I, p, J, flag, s);
    PassArgs a{bed, pitch, I, q, p, 0, I, 0, J, chunk, sp};
    a.strict = flag;
    launch_qside<double>(MODE_LL, (int)K, a, nchunk, s);
    launch_reduce_sum(sp, nbx * nchunk, out, 0, s);
    return check_launch();
  });
}
