// Achievable HBM bandwidth on the device the library runs on, STREAM style:
// copy (c = a) and triad (c = a + s b) over FP64 arrays far larger than the
// 256 MB Infinity Cache, with hand-written kernels: 16-byte accesses, UNR
// independent loads in flight per lane.  Two layouts: blocks dealt over the
// whole array in a grid-stride loop, and one contiguous chunk per block
// (each block streams its own stretch of every array, which keeps DRAM
// pages open: the faster layout on MI355X).  Plain and streaming
// (nontemporal) accesses, 256- and 1024-thread blocks and several grid
// sizes are timed; the best of each kernel is reported.  A measurement aid
// with no counterpart in the reference: it gives the solver kernels' rates a
// ceiling measured on the same GPU.
#include <algorithm>

#include "tl_internal.hpp"

namespace {

using tl::S;

typedef double dv2 __attribute__((ext_vector_type(2)));  // 16 bytes
constexpr int UNR = 4;  // 16-byte vectors per lane per trip

template <bool NT>
__device__ inline dv2 ld2(const dv2 *p) {
  if (NT) return __builtin_nontemporal_load(p);
  return *p;
}

template <bool NT>
__device__ inline void st2(dv2 *p, dv2 v) {
  if (NT) __builtin_nontemporal_store(v, p);
  else *p = v;
}

// c = a (TRIAD = 0) or c = a + s b (TRIAD = 1) over n2 dv2 elements.
// CHUNK = 0: grid-stride; CHUNK = 1: block b covers elements
// [b chunk, (b + 1) chunk), chunk = ceil(n2 / blocks).
template <bool NT, int TRIAD, int CHUNK>
__global__ __launch_bounds__(1024) void k_stream(const dv2 *__restrict__ a,
                                                 const dv2 *__restrict__ b,
                                                 dv2 *__restrict__ c, long n2, double s) {
  long i, stride;
  if (CHUNK) {
    const long chunk = (n2 + gridDim.x - 1) / gridDim.x;
    const long lo = (long)blockIdx.x * chunk;
    const long hi = lo + chunk < n2 ? lo + chunk : n2;
    a += lo;
    b += lo;
    c += lo;
    n2 = hi > lo ? hi - lo : 0;
    stride = blockDim.x;
    i = threadIdx.x;
  } else {
    stride = (long)gridDim.x * blockDim.x;
    i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  }
  // whole trips: UNR vectors per lane, the loads issued before the stores
  for (; i + (UNR - 1) * stride < n2; i += UNR * stride) {
    dv2 va[UNR], vb[UNR];
#pragma unroll
    for (int u = 0; u < UNR; ++u) {
      va[u] = ld2<NT>(a + i + u * stride);
      if (TRIAD) vb[u] = ld2<NT>(b + i + u * stride);
    }
#pragma unroll
    for (int u = 0; u < UNR; ++u) {
      dv2 v = va[u];
      if (TRIAD) {
        v += s * vb[u];
      }
      st2<NT>(c + i + u * stride, v);
    }
  }
  for (; i < n2; i += stride) {
    dv2 v = ld2<NT>(a + i);
    if (TRIAD) {
      v += s * ld2<NT>(b + i);
    }
    st2<NT>(c + i, v);
  }
}

__global__ void k_init(double *__restrict__ p, long n, double v) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x)
    p[i] = v + 1e-9 * (double)(i & 1023);
}

}  // namespace

extern "C" int tl_hbm_stream(long n, int reps, double *copy_gbs, double *triad_gbs) {
  TL_REQUIRE_INIT();
  if (n < 1024 || (n & 1) || reps < 1 || !copy_gbs || !triad_gbs)
    return tl::fail(TL_ERR_ARG, "tl_hbm_stream: need an even n >= 1024, reps >= 1, outputs");
  tl::State &s = S();
  double *buf[3] = {nullptr, nullptr, nullptr};
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = TL_OK;
  auto cleanup = [&]() {
    (void)hipStreamSynchronize(s.stream);
    for (double *p : buf)
      if (p) (void)hipFree(p);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
  };
  for (double *&p : buf) {
    const hipError_t e = hipMalloc((void **)&p, sizeof(double) * (size_t)n);
    if (e != hipSuccess) {
      p = nullptr;
      rc = tl::hip_fail(e, "tl_hbm_stream: hipMalloc");
      cleanup();
      return rc;
    }
    hipLaunchKernelGGL(k_init, dim3(4096), dim3(256), 0, s.stream, p, n, 1.0);
  }
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    cleanup();
    return tl::fail(TL_ERR_HIP, "tl_hbm_stream: hipEventCreate");
  }
  const dv2 *a = (const dv2 *)buf[0], *b = (const dv2 *)buf[1];
  dv2 *c = (dv2 *)buf[2];
  const long n2 = n / 2;
  typedef void (*KS)(const dv2 *, const dv2 *, dv2 *, long, double);
  // [triad][nt][chunk]
  static const KS ks[2][2][2] = {
      {{k_stream<false, 0, 0>, k_stream<false, 0, 1>}, {k_stream<true, 0, 0>, k_stream<true, 0, 1>}},
      {{k_stream<false, 1, 0>, k_stream<false, 1, 1>}, {k_stream<true, 1, 0>, k_stream<true, 1, 1>}}};
  const int grids[] = {s.cus * 4, s.cus * 16, s.cus * 64};
  const int threads[] = {256, 1024};
  double best[2] = {0.0, 0.0};
  for (int triad = 0; triad < 2; ++triad)
    for (int nt = 0; nt < 2; ++nt)
      for (int chunk = 0; chunk < 2; ++chunk)
        for (int tb : threads)
          for (int g : grids) {
            const KS k = ks[triad][nt][chunk];
            hipLaunchKernelGGL(k, dim3(g), dim3(tb), 0, s.stream, a, b, c, n2, 3.0);  // warm-up
            (void)hipEventRecord(e0, s.stream);
            for (int r = 0; r < reps; ++r)
              hipLaunchKernelGGL(k, dim3(g), dim3(tb), 0, s.stream, a, b, c, n2, 3.0);
            (void)hipEventRecord(e1, s.stream);
            if (hipEventSynchronize(e1) != hipSuccess) {
              cleanup();
              return tl::fail(TL_ERR_HIP, "tl_hbm_stream: kernel");
            }
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            const double bytes = (triad ? 24.0 : 16.0) * (double)n * reps;
            best[triad] = std::max(best[triad], bytes / (ms * 1e-3) / 1e9);
          }
  if (hipGetLastError() != hipSuccess) rc = tl::fail(TL_ERR_HIP, "tl_hbm_stream: launch");
  cleanup();
  *copy_gbs = best[0];
  *triad_gbs = best[1];
  return rc;
}
