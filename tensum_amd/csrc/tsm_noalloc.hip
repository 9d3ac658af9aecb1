// tsm_noalloc.hip -- the fixed scratch buffer of tsm_probe_density_noalloc and
// its point-location kernel.
//
// Callers that cannot allocate device memory (code under HIP graph capture,
// tight serving loops) use a file-scope device array instead of per-call
// hipMalloc: every GPU the library runs on holds one copy of `scratch` until
// the process ends.  Host code never passes the array's name to a kernel or a
// copy (in host code it is a stand-in, not a device address): the address is
// looked up once per device with hipGetSymbolAddress, with that device
// current, and kept per device.  Every call on a GPU shares its array, so
// calls that overlap on one GPU race; the array keeps the last call's contents.
#include <hip/hip_runtime.h>

#include <mutex>

#include "tsm_device.h"

namespace tsm {

constexpr size_t SCRATCH_BYTES = size_t(8) << 20;
static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

namespace {
constexpr int MAX_DEV = 64;
std::mutex g_mu;
void* g_addr[MAX_DEV] = {};
}  // namespace

size_t noalloc_scratch_bytes() { return SCRATCH_BYTES; }

// Device address of the scratch array on the current device (looked up on the
// first call per device).  Returns 0 or a hipError_t.
int noalloc_scratch(void** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return (int)e;
  if (dev < 0 || dev >= MAX_DEV) return (int)hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_addr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return (int)e;
    g_addr[dev] = p;
  }
  *out = g_addr[dev];
  return 0;
}

// Containing triangle of each probe point (tsm_probe_density and its
// no-allocation variant): the first cell, in cell order, whose barycentric
// coordinates are all >= -1e-12.  One workgroup per probe scans the cells; cell[p] = the cell (-1 if none holds
// the point), pc[3p..] = its corner vertex ids (0 if none), pw[3p..] = weights.
__global__ __launch_bounds__(256) void k_probe_locate(DevGeo G, const double* px, const double* py,
                                                      int* pc, double* pw, int* cell) {
  __shared__ int best[4];
  const int p = blockIdx.x;
  const double X = px[p], Y = py[p];
  int found = 0x7fffffff;
  for (int k = threadIdx.x; k < G.nc; k += blockDim.x) {
    const int a = G.cv[3 * k], b = G.cv[3 * k + 1], d = G.cv[3 * k + 2];
    const double x0 = G.x[a], y0 = G.y[a], x1 = G.x[b], y1 = G.y[b], x2 = G.x[d], y2 = G.y[d];
    const double det = (x1 - x0) * (y2 - y0) - (x2 - x0) * (y1 - y0);
    const double l1 = ((X - x0) * (y2 - y0) - (x2 - x0) * (Y - y0)) / det;
    const double l2 = ((x1 - x0) * (Y - y0) - (X - x0) * (y1 - y0)) / det;
    const double l0 = 1.0 - l1 - l2;
    const double mn = fmin(l0, fmin(l1, l2));
    if (mn >= -1e-12) { found = k; break; }  // later cells of this lane are larger
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) found = min(found, __shfl_xor(found, o, 64));
  if ((threadIdx.x & 63) == 0) best[threadIdx.x >> 6] = found;
  __syncthreads();
  if (threadIdx.x != 0) return;
  for (int w = 1; w < (int)(blockDim.x >> 6); ++w) found = min(found, best[w]);
  if (found == 0x7fffffff) {
    for (int j = 0; j < 3; ++j) { pc[3 * p + j] = 0; pw[3 * p + j] = 0.0; }
    cell[p] = -1;
    return;
  }
  const int k = found;
  cell[p] = k;
  const int a = G.cv[3 * k], b = G.cv[3 * k + 1], d = G.cv[3 * k + 2];
  const double x0 = G.x[a], y0 = G.y[a], x1 = G.x[b], y1 = G.y[b], x2 = G.x[d], y2 = G.y[d];
  const double det = (x1 - x0) * (y2 - y0) - (x2 - x0) * (y1 - y0);
  const double l1 = ((X - x0) * (y2 - y0) - (x2 - x0) * (Y - y0)) / det;
  const double l2 = ((x1 - x0) * (Y - y0) - (X - x0) * (y1 - y0)) / det;
  const double l0 = 1.0 - l1 - l2;
  pc[3 * p] = a; pc[3 * p + 1] = b; pc[3 * p + 2] = d;
  pw[3 * p] = l0; pw[3 * p + 1] = l1; pw[3 * p + 2] = l2;
}

void launch_probe_locate(const DevGeo& G, int np, const double* px, const double* py, int* pc,
                         double* pw, int* cell, hipStream_t st) {
  hipLaunchKernelGGL(k_probe_locate, dim3(np), dim3(256), 0, st, G, px, py, pc, pw, cell);
}

}  // namespace tsm
