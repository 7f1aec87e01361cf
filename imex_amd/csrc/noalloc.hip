// noalloc.hip — the divergence metric of device-resident faces without any allocation.
//
// imex_div_b_noalloc is imex_div_b (cal_divergence_error, imex_subsystems.cpp:630-660) for callers
// that may not allocate: code under HIP graph capture, tight serving loops.  It takes the caller's
// device face arrays and writes the metric to the caller's device memory.  Its one piece of scratch
// memory, the per-block partial sums, is the file-scope array `scratch` below, resolved to a device
// address once per device on the first call there; every call on that device shares it and keeps the
// last call's partials, so calls that overlap on one device (two streams, or two host threads) race.
#include <algorithm>
#include <mutex>
#include <string>
#include "kernels.hpp"

namespace {

constexpr int NOALLOC_BLOCKS = 256;        // partials: one per block, the launch shape of k_divb
constexpr int NOALLOC_MAX_SIDE = 32768;    // largest nx or ny accepted (int cell indices)
constexpr int NOALLOC_MAX_DEVICES = 64;

static __device__ __align__(256) double scratch[NOALLOC_BLOCKS];

std::mutex g_mu;
double* g_scratch[NOALLOC_MAX_DEVICES] = {};

__global__ void __launch_bounds__(256) k_divb_noalloc(const double* __restrict__ bx, const double* __restrict__ by,
                                                      int nx, int NX, long long N, double dx, double dy,
                                                      double* __restrict__ part) {
  __shared__ double sh[16];
  double acc = divb_thread_sum(bx, by, nx, NX, N, dx, dy, blockIdx.x, gridDim.x);
  double s = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// The partials summed in block order by one thread: the order imex_div_b sums them in.
__global__ void k_divb_noalloc_fin(const double* __restrict__ part, int nb, double* __restrict__ out) {
  if (threadIdx.x != 0) return;
  double s = 0.0;
  for (int b = 0; b < nb; b++) s += part[b];
  out[0] = sqrt(s);
}

// The device address of `scratch` on the current device (resolved once per device).
int scratch_address(double** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= NOALLOC_MAX_DEVICES)
    return imex_internal_error(IMEX_ERR_HIP, "imex_div_b_noalloc: no current HIP device");
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return imex_internal_error(IMEX_ERR_HIP, std::string("imex_div_b_noalloc: hipGetSymbolAddress: ") + hipGetErrorString(e));
    g_scratch[dev] = static_cast<double*>(p);
  }
  *out = g_scratch[dev];
  return 0;
}

}  // namespace

extern "C" int imex_div_b_noalloc(const imex_params* p, const double* bx, const double* by, double* out, void* stream) {
  if (!p || !bx || !by || !out) return imex_internal_error(IMEX_ERR_ARG, "imex_div_b_noalloc: null argument");
  if (p->nx < 1 || p->ny < 1) return imex_internal_error(IMEX_ERR_ARG, "imex_div_b_noalloc: empty grid");
  if (p->nx > NOALLOC_MAX_SIDE || p->ny > NOALLOC_MAX_SIDE)
    return imex_internal_error(IMEX_ERR_ARG, "imex_div_b_noalloc: grid larger than 32768 cells per side");
  if (!(p->x1 > p->x0) || !(p->y1 > p->y0)) return imex_internal_error(IMEX_ERR_ARG, "imex_div_b_noalloc: empty domain");
  double* part = nullptr;
  int rc = scratch_address(&part);
  if (rc) return rc;
  const long long N = (long long)p->nx * p->ny;
  const int nb = (int)std::min<long long>((N + 255) / 256, NOALLOC_BLOCKS);
  const double dx = (p->x1 - p->x0) / p->nx, dy = (p->y1 - p->y0) / p->ny;
  hipStream_t st = static_cast<hipStream_t>(stream);
  k_divb_noalloc<<<nb, 256, 0, st>>>(bx, by, p->nx, p->nx + 2 * GC, N, dx, dy, part);
  k_divb_noalloc_fin<<<1, 64, 0, st>>>(part, nb, out);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return imex_internal_error(IMEX_ERR_HIP, std::string("imex_div_b_noalloc: launch: ") + hipGetErrorString(e));
  return 0;
}
