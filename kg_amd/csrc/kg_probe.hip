// kg_probe.hip -- HBM bandwidth probe: what plain streaming kernels move on
// this GPU, as the ceiling the stage kernels are compared with.  Measurement
// only (no counterpart in the reference).
//
// Kinds (FP64 arrays of n doubles, grid-stride loop, 16-byte accesses):
//   0 read    acc += a[i]            (n x 8 B read)
//   1 write   d[i] = c               (n x 8 B written)
//   2 copy    d[i] = a[i]            (read + write)
//   3 mix     d[i] = a[i] + b[i] + c[i]   (three reads, one write: the
//                                          proportion of an RK stage launch)
//   4, 5      copy and mix with non-temporal loads and stores
#include <hip/hip_runtime.h>

#include <cstdio>
#include <string>

#include "../../include/kg.h"

namespace {

thread_local std::string g_probe_err;

template <int KIND>
__global__ void __launch_bounds__(256) probe_kernel(const double2* __restrict__ a, const double2* __restrict__ b,
                                                     const double2* __restrict__ c, double2* __restrict__ d,
                                                     int64_t n2) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (KIND == 0) {
    double2 acc = {0.0, 0.0};
    for (; i < n2; i += stride) {
      const double2 v = a[i];
      acc.x += v.x;
      acc.y += v.y;
    }
    // never true for the probe's data (all ones); keeps the loads alive
    if (acc.x + acc.y == -1.0 && (int64_t)threadIdx.x < n2) d[threadIdx.x] = acc;
  } else if (KIND == 1) {
    const double2 v = {1.0, 1.0};
    for (; i < n2; i += stride) d[i] = v;
  } else if (KIND == 2) {
    for (; i < n2; i += stride) d[i] = a[i];
  } else if (KIND == 3) {
    for (; i < n2; i += stride) {
      const double2 x = a[i], y = b[i], z = c[i];
      d[i] = double2{x.x + y.x + z.x, x.y + y.y + z.y};
    }
  } else if (KIND == 4) {
    for (; i < n2; i += stride) {
      double2 v;
      v.x = __builtin_nontemporal_load(&a[i].x);
      v.y = __builtin_nontemporal_load(&a[i].y);
      __builtin_nontemporal_store(v.x, &d[i].x);
      __builtin_nontemporal_store(v.y, &d[i].y);
    }
  } else {
    for (; i < n2; i += stride) {
      const double x0 = __builtin_nontemporal_load(&a[i].x), x1 = __builtin_nontemporal_load(&a[i].y);
      const double y0 = __builtin_nontemporal_load(&b[i].x), y1 = __builtin_nontemporal_load(&b[i].y);
      const double z0 = __builtin_nontemporal_load(&c[i].x), z1 = __builtin_nontemporal_load(&c[i].y);
      __builtin_nontemporal_store(x0 + y0 + z0, &d[i].x);
      __builtin_nontemporal_store(x1 + y1 + z1, &d[i].y);
    }
  }
}

__global__ void probe_fill(double* p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 1.0;
}

hipError_t launch(int kind, int blocks, const double* a, const double* b, const double* c, double* d, int64_t n2,
                  hipStream_t s) {
  const double2 *a2 = (const double2*)a, *b2 = (const double2*)b, *c2 = (const double2*)c;
  double2* d2 = (double2*)d;
  switch (kind) {
    case 0: probe_kernel<0><<<blocks, 256, 0, s>>>(a2, b2, c2, d2, n2); break;
    case 1: probe_kernel<1><<<blocks, 256, 0, s>>>(a2, b2, c2, d2, n2); break;
    case 2: probe_kernel<2><<<blocks, 256, 0, s>>>(a2, b2, c2, d2, n2); break;
    case 3: probe_kernel<3><<<blocks, 256, 0, s>>>(a2, b2, c2, d2, n2); break;
    case 4: probe_kernel<4><<<blocks, 256, 0, s>>>(a2, b2, c2, d2, n2); break;
    default: probe_kernel<5><<<blocks, 256, 0, s>>>(a2, b2, c2, d2, n2); break;
  }
  return hipGetLastError();
}

}  // namespace

extern "C" {

const char* kg_probe_last_error(void) { return g_probe_err.c_str(); }

kg_status kg_hbm_probe(int32_t device, int32_t kind, int64_t n, int32_t blocks, int32_t reps, double* ms_per_launch) {
  if (kind < 0 || kind > 5 || n < 2 || (n & 1) || blocks < 1 || reps < 1 || !ms_per_launch) {
    g_probe_err = "kg_hbm_probe: kind 0..5, even n >= 2, blocks >= 1, reps >= 1";
    return KG_ERR_INVALID;
  }
  hipError_t e = hipSetDevice(device);
  double* buf[4] = {nullptr, nullptr, nullptr, nullptr};
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  float ms = 0.f;
  for (int q = 0; q < 4 && e == hipSuccess; ++q) e = hipMalloc(&buf[q], (size_t)n * sizeof(double));
  if (e == hipSuccess) e = hipStreamCreate(&s);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  for (int q = 0; q < 4 && e == hipSuccess; ++q) {
    probe_fill<<<1024, 256, 0, s>>>(buf[q], n);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = launch(kind, blocks, buf[0], buf[1], buf[2], buf[3], n / 2, s);  // warm-up
  if (e == hipSuccess) e = hipEventRecord(e0, s);
  for (int r = 0; r < reps && e == hipSuccess; ++r) e = launch(kind, blocks, buf[0], buf[1], buf[2], buf[3], n / 2, s);
  if (e == hipSuccess) e = hipEventRecord(e1, s);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (s) (void)hipStreamDestroy(s);
  for (int q = 0; q < 4; ++q)
    if (buf[q]) (void)hipFree(buf[q]);
  if (e != hipSuccess) {
    g_probe_err = std::string("kg_hbm_probe: ") + hipGetErrorString(e);
    return KG_ERR_HIP;
  }
  *ms_per_launch = (double)ms / reps;
  return KG_OK;
}

}  // extern "C"
