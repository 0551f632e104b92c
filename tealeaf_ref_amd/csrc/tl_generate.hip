// Set-up on the device: the reference's initialise_chunk + generate_chunk
// for one chunk (kernels/initialise_chunk_kernel.f90:26-132,
// kernels/generate_chunk_kernel.f90:30-142).  Every coordinate is rounded
// exactly as the reference computes it on the host (separate multiply and
// add, correctly rounded square root), so which cells a state covers is
// decided bit for bit as there.
#include <string>
#include <vector>

#include "tl_internal.hpp"

using namespace tl;

namespace {

constexpr int MAX_STATES = 256;

struct DevState {
  int geometry;
  double density, energy, xmin, xmax, ymin, ymax, radius;
};

// vertex coordinate x_min + d*(j-1) (initialise_chunk_kernel.f90:56-63)
__device__ inline double vertex(double x0, double d, int j) {
  return __dadd_rn(x0, __dmul_rn(d, (double)(j - 1)));
}
// cell centre 0.5*(x_j + x_{j+1}) (:65-72)
__device__ inline double centre(double x0, double d, int j) {
  return __dmul_rn(0.5, __dadd_rn(vertex(x0, d, j), vertex(x0, d, j + 1)));
}

__global__ void k_generate(int nx, int ny, int h, long pitch, double xmin, double ymin,
                           double dx, double dy, int nstates, const DevState *__restrict__ st,
                           double *__restrict__ density, double *__restrict__ energy0,
                           double *__restrict__ u) {
  const long w = (long)nx + 2 * h, n = w * (ny + 2L * h);
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x) {
    const int j = (int)(i % w) + 1 - h, k = (int)(i / w) + 1 - h;  // reference indices
    double rho = st[0].density, e = st[0].energy;
    // coordinate arrays of the reference run from -1 (cells) / -1 (vertices)
    const int jc = min(max(j, -1), nx + 2), kc = min(max(k, -1), ny + 2);
    const int jv = min(max(j, -1), nx + 3), kv = min(max(k, -1), ny + 3);
    for (int s = 1; s < nstates; ++s) {
      const DevState &S = st[s];
      bool hit = false;
      if (S.geometry == TL_GEOM_RECTANGLE) {
        if (j >= 1 && j <= nx && k >= 1 && k <= ny)
          hit = vertex(xmin, dx, j + 1) >= S.xmin && vertex(xmin, dx, j) < S.xmax &&
                vertex(ymin, dy, k + 1) >= S.ymin && vertex(ymin, dy, k) < S.ymax;
      } else if (S.geometry == TL_GEOM_CIRCLE) {
        const double ax = __dsub_rn(centre(xmin, dx, jc), S.xmin);
        const double ay = __dsub_rn(centre(ymin, dy, kc), S.ymin);
        const double rad = __dsqrt_rn(__dadd_rn(__dmul_rn(ax, ax), __dmul_rn(ay, ay)));
        hit = rad <= S.radius;
      } else if (S.geometry == TL_GEOM_POINT) {
        hit = vertex(xmin, dx, jv) == S.xmin && vertex(ymin, dy, kv) == S.ymin;
      }
      if (hit) {
        rho = S.density;
        e = S.energy;
      }
    }
    const long at = (long)(k - 1) * pitch + (j - 1);
    density[at] = rho;
    energy0[at] = e;
    u[at] = e * rho;
  }
}

__global__ void k_fill_field(double *__restrict__ a, long n, double v) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x)
    a[i] = v;
}

}  // namespace

extern "C" int tl_generate(double xmin, double ymin, double dx, double dy, int nstates,
                           const tl_state *states) {
  TL_REQUIRE_INIT();
  State &s = S();
  if (!states || nstates < 1 || nstates > MAX_STATES)
    return fail(TL_ERR_ARG, "tl_generate: need 1 to " + std::to_string(MAX_STATES) + " states");
  if (!(dx > 0.0) || !(dy > 0.0)) return fail(TL_ERR_ARG, "tl_generate: bad cell size");
  std::vector<DevState> hs(nstates);
  for (int i = 0; i < nstates; ++i) {
    const tl_state &t = states[i];  // a state of no known geometry covers nothing
    hs[i] = DevState{t.geometry, t.density, t.energy, t.xmin, t.xmax, t.ymin, t.ymax, t.radius};
  }
  DevState *dst = nullptr;
  TL_HIP(hipMalloc((void **)&dst, sizeof(DevState) * nstates));
  hipError_t e = hipMemcpyAsync(dst, hs.data(), sizeof(DevState) * nstates, hipMemcpyHostToDevice,
                                s.stream);
  if (e == hipSuccess) {
    const long n = ((long)s.nx + 2L * s.h) * ((long)s.ny + 2L * s.h);
    const int nb = (int)((n + BLOCK - 1) / BLOCK < 8192 ? (n + BLOCK - 1) / BLOCK : 8192);
    hipLaunchKernelGGL(k_generate, dim3(nb), dim3(BLOCK), 0, s.stream, s.nx, s.ny, s.h, s.pitch,
                       xmin, ymin, dx, dy, nstates, dst, s.density.p11, s.energy0.p11,
                       s.u[s.ucur].p11);
    const long nv = (long)(s.field_bytes / sizeof(double));
    hipLaunchKernelGGL(k_fill_field, dim3(4096), dim3(BLOCK), 0, s.stream, s.volume.base, nv,
                       dx * dy);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(s.stream);
  }
  (void)hipFree(dst);
  if (e != hipSuccess) return hip_fail(e, "tl_generate");
  return TL_OK;
}
