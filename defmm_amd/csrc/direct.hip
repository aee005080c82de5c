// FP64 direct sum on the GPU: the check of the reference's CheckFULL
// (include/interface.hpp:178-221) on sampled targets,
//   p_i = sum_j exp(i kappa r_ij) / (4 pi r_ij) q_j,  terms with r_ij^2 < 1e-16 skipped,
// and its variant for callers that cannot allocate device memory.
//
// Few targets against many sources (2,000 sampled targets, 10^6 sources) give
// a one-thread-per-target kernel only a handful of blocks, so the sources are
// split into parts: block (b, k) sums part k of the sources for 256 targets
// (compensated), writing one partial per (part, target) to a scratch buffer;
// a second kernel adds the partials of each target in part order
// (compensated).  Targets are processed in tiles of kTileTargets, so the
// scratch never exceeds kPartsMax x kTileTargets partials whatever the input
// size; the part count depends only on the number of sources, so both entry
// points return bit-identical results.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/defmm_amd.h"
#include "errors.hpp"

namespace dfm {

constexpr int kPartsMax = 64;          // source parts
constexpr int kTileTargets = 4096;     // targets per tile
constexpr int kSourcesPerPart = 2048;  // parts = ceil(ns / 2048), at most kPartsMax
constexpr int kMaxDevices = 64;

// Scratch of defmm_direct_noalloc: partial sums of one tile, 4 MiB, one
// copy per GPU (held until the process ends).  Host code must not pass the
// array's name to a kernel (it is a host-side stand-in): it uses the device
// address obtained from hipGetSymbolAddress.
static __device__ __align__(256) double2 scratch[kPartsMax * kTileTargets];

static int directParts(int64_t ns) {
  int64_t p = (ns + kSourcesPerPart - 1) / kSourcesPerPart;
  return (int)(p < 1 ? 1 : p > kPartsMax ? kPartsMax : p);
}

__global__ __launch_bounds__(256) void k_direct_part(const double* __restrict__ src, int ns,
                                                     const double2* __restrict__ q,
                                                     const double* __restrict__ trg, int nt,
                                                     double kappa, int chunk,
                                                     double2* __restrict__ partial) {
  __shared__ double sx[256], sy[256], sz[256];
  __shared__ double2 sq[256];
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int part = blockIdx.y;
  const bool act = i < nt;
  double xt = 0., yt = 0., zt = 0.;
  if (act) {
    xt = trg[3 * (size_t)i];
    yt = trg[3 * (size_t)i + 1];
    zt = trg[3 * (size_t)i + 2];
  }
  const int jb = part * chunk;
  const int je = min(ns, jb + chunk);
  double pr = 0., pi = 0., cr = 0., ci = 0.;
  for (int j0 = jb; j0 < je; j0 += 256) {
    const int n = min(256, je - j0);
    __syncthreads();
    if ((int)threadIdx.x < n) {
      const size_t j = (size_t)j0 + threadIdx.x;
      sx[threadIdx.x] = src[3 * j];
      sy[threadIdx.x] = src[3 * j + 1];
      sz[threadIdx.x] = src[3 * j + 2];
      sq[threadIdx.x] = q[j];
    }
    __syncthreads();
    if (act)
      for (int j = 0; j < n; j++) {
        const double x = xt - sx[j], y = yt - sy[j], z = zt - sz[j];
        const double R = x * x + y * y + z * z;
        if (R < 1.e-16) continue;
        const double r = sqrt(R);
        const double K = 0.079577471545947667884 / r;
        double si, co;
        sincos(kappa * r, &si, &co);
        co *= K;
        si *= K;
        const double tr = co * sq[j].x - si * sq[j].y - cr;
        const double s1 = pr + tr;
        cr = (s1 - pr) - tr;
        pr = s1;
        const double ti = co * sq[j].y + si * sq[j].x - ci;
        const double s2 = pi + ti;
        ci = (s2 - pi) - ti;
        pi = s2;
      }
  }
  if (act) partial[(size_t)part * nt + i] = make_double2(pr, pi);
}

__global__ __launch_bounds__(256) void k_direct_reduce(const double2* __restrict__ partial,
                                                       int parts, int nt,
                                                       double2* __restrict__ out) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= nt) return;
  double pr = 0., pi = 0., cr = 0., ci = 0.;
  for (int k = 0; k < parts; k++) {
    const double2 v = partial[(size_t)k * nt + i];
    const double tr = v.x - cr;
    const double s1 = pr + tr;
    cr = (s1 - pr) - tr;
    pr = s1;
    const double ti = v.y - ci;
    const double s2 = pi + ti;
    ci = (s2 - pi) - ti;
    pi = s2;
  }
  out[i] = make_double2(pr, pi);
}

// Direct sum of device arrays on stream s with a partial-sum buffer of at
// least directParts(ns) * min(nt, kTileTargets) elements.
static int directRun(const double* src, int64_t ns, const double2* q, const double* trg,
                     int64_t nt, double kappa, double2* p, double2* partial, hipStream_t s) {
  const int parts = directParts(ns);
  const int chunk = (int)((ns + parts - 1) / parts);
  for (int64_t t0 = 0; t0 < nt; t0 += kTileTargets) {
    const int n = (int)(nt - t0 < kTileTargets ? nt - t0 : kTileTargets);
    const unsigned nb = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(k_direct_part, dim3(nb, (unsigned)parts), dim3(256), 0, s, src, (int)ns,
                       q, trg + 3 * t0, n, kappa, chunk, partial);
    hipLaunchKernelGGL(k_direct_reduce, dim3(nb), dim3(256), 0, s, (const double2*)partial,
                       parts, n, p + t0);
  }
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    g_err = std::string("direct sum launch: ") + hipGetErrorString(e);
    return DEFMM_EHIP;
  }
  return DEFMM_OK;
}

// Device address of this GPU's scratch array, looked up once per device.
static int scratchOf(int dev, double2** out) {
  static std::mutex mu;
  static void* addr[kMaxDevices] = {};
  if (dev < 0 || dev >= kMaxDevices) {
    g_err = "device index beyond the scratch table";
    return DEFMM_EARG;
  }
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      g_err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
      return DEFMM_EHIP;
    }
    addr[dev] = ptr;
  }
  *out = (double2*)addr[dev];
  return DEFMM_OK;
}

}  // namespace dfm

extern "C" {

int defmm_direct(const double* src_xyz, int64_t n_src, const double* q, const double* trg_xyz,
                 int64_t n_trg, double kappa, double* p) {
  using namespace dfm;
  if (!src_xyz || !q || !trg_xyz || !p || n_src < 1 || n_trg < 1 || n_src > 2147483647 ||
      n_trg > 2147483647) {
    g_err = "invalid argument";
    return DEFMM_EARG;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) { g_err = "no HIP device"; return DEFMM_ENODEVICE; }
  void *ds = nullptr, *dq = nullptr, *dt = nullptr, *dp = nullptr, *dw = nullptr;
  int rc = DEFMM_OK;
  auto ck = [&](hipError_t e, const char* what) {
    if (e != hipSuccess) (void)hipGetLastError();  // consumed here, not at the next check
    if (e != hipSuccess && rc == DEFMM_OK) {
      g_err = std::string(what) + ": " + hipGetErrorString(e);
      rc = e == hipErrorOutOfMemory ? DEFMM_EMEMORY : DEFMM_EHIP;
    }
  };
  const int64_t tile = n_trg < kTileTargets ? n_trg : kTileTargets;
  ck(hipMalloc(&ds, n_src * 24), "hipMalloc");
  if (!rc) ck(hipMalloc(&dq, n_src * 16), "hipMalloc");
  if (!rc) ck(hipMalloc(&dt, n_trg * 24), "hipMalloc");
  if (!rc) ck(hipMalloc(&dp, n_trg * 16), "hipMalloc");
  if (!rc) ck(hipMalloc(&dw, (size_t)directParts(n_src) * tile * 16), "hipMalloc");
  if (!rc) ck(hipMemcpy(ds, src_xyz, n_src * 24, hipMemcpyHostToDevice), "hipMemcpy");
  if (!rc) ck(hipMemcpy(dq, q, n_src * 16, hipMemcpyHostToDevice), "hipMemcpy");
  if (!rc) ck(hipMemcpy(dt, trg_xyz, n_trg * 24, hipMemcpyHostToDevice), "hipMemcpy");
  if (!rc)
    rc = directRun((const double*)ds, n_src, (const double2*)dq, (const double*)dt, n_trg, kappa,
                   (double2*)dp, (double2*)dw, 0);
  if (!rc) ck(hipMemcpy(p, dp, n_trg * 16, hipMemcpyDeviceToHost), "hipMemcpy");
  for (void* m : {ds, dq, dt, dp, dw})
    if (m) (void)hipFree(m);
  return rc;
}

int defmm_direct_noalloc(const void* src_dev, int64_t n_src, const void* q_dev,
                         const void* trg_dev, int64_t n_trg, double kappa, void* p_dev,
                         void* stream) {
  using namespace dfm;
  if (!src_dev || !q_dev || !trg_dev || !p_dev) {
    g_err = "null argument";
    return DEFMM_EARG;
  }
  if (n_src < 1 || n_trg < 1 || n_src > 2147483647 || n_trg > 2147483647) {
    g_err = "n_src and n_trg must be in 1 .. 2^31-1";
    return DEFMM_EARG;
  }
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) {
    g_err = "hipGetDevice failed";
    return DEFMM_EHIP;
  }
  double2* ws = nullptr;
  int rc = scratchOf(dev, &ws);
  if (rc) return rc;
  return directRun((const double*)src_dev, n_src, (const double2*)q_dev, (const double*)trg_dev,
                   n_trg, kappa, (double2*)p_dev, ws, (hipStream_t)stream);
}

}  // extern "C"
