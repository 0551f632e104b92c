// Adjoint of the implicit step.  No counterpart in the reference.
//
// One step of the reference solves A(rho) u_n = b_n with b_1 = rho * e_0 and
// b_n = rho * (u_{n-1} / rho) = u_{n-1} afterwards (tea_leaf_common.f90:67-72
// and tea_leaf_finalise).  A = I + rx Lx(Kx) + ry Ly(Ky) is symmetric, with
// face coefficients K = (1/w_a + 1/w_b) / 2 of the two cells of a face,
// w = rho or 1/rho (kernels/tea_leaf_common_kernel.f90:75-101).  For a
// functional J(u_N), reverse mode gives, from g_N = dJ/du_N:
//   lambda_n = A^-1 g_n,   g_{n-1} = lambda_n,
//   dJ/dK_f  = -r_f (lambda_a - lambda_b)(u_a - u_b)   (r_f = rx or ry),
//   dJ/dlog(rho_c) = sum over the faces f of c of dJ/dK_f dK_f/dlog(rho_c)
//                    + [n = 1] lambda_1 rho_c e_0,c,
// with dK_f/dlog(rho_c) = -1/(2 rho_c) for w = rho and rho_c/2 for w = 1/rho.
// External faces carry no dependence (zeroed, or reflective with a zero
// difference).  The solves reuse the forward loop kernels; this file holds
// the device-resident snapshot slots (checkpoints, seeds, gradient) and the
// face-gradient kernel.
#include <string>

#include "tl_internal.hpp"
#include "tl_stencil.hpp"

namespace tl {
dev::Geo geo(int cols_per_tile);
}
using namespace tl;
using dev::Geo;

namespace {

int check_slot(int slot) {
  State &s = S();
  if (slot < 0 || slot >= (int)s.snaps.size())
    return fail(TL_ERR_ARG, "snapshot slot " + std::to_string(slot) + " not reserved");
  return TL_OK;
}

// cell (1,1) of a snapshot slot, laid out like a field
double *snap11(int slot) {
  State &s = S();
  return s.snaps[slot] + (long)s.h * s.pitch + PADL;
}

__global__ __launch_bounds__(BLOCK) void k_fill(double *__restrict__ a, long n, double v) {
  for (long i = blockIdx.x * (long)BLOCK + threadIdx.x; i < n; i += (long)gridDim.x * BLOCK)
    a[i] = v;
}

// grad_c += sum_f dJ/dK_f dK_f/dlog(rho_c) (+ lambda_c rho_c e0_c)
__global__ __launch_bounds__(BLOCK) void k_face_grad(Geo G, const double *__restrict__ lam,
                                                      const double *__restrict__ u,
                                                      const double *__restrict__ rho,
                                                      const double *__restrict__ e0,
                                                      double *__restrict__ grad, double rx,
                                                      double ry, int recip) {
  const long P = G.pitch;
  dev::TileIter it(G.tiles, G.blocks);
  for (long t = it.t; t < it.t_end; ++t) {
    const dev::Tile tt = dev::tile_of(G, t);
    if (tt.x >= G.nx) continue;
    const int x = tt.x;
    for (int y = tt.y0; y < tt.y1; ++y) {
      const long i = (long)y * P + x;
      const double lc = lam[i], uc = u[i], rc = rho[i];
      const double dk = recip ? 0.5 * rc : -0.5 / rc;  // dK_f / dlog(rho_c)
      double s = 0.0;
      if (x > 0) s += rx * (lc - lam[i - 1]) * (uc - u[i - 1]);
      if (x + 1 < G.nx) s += rx * (lc - lam[i + 1]) * (uc - u[i + 1]);
      if (y > 0) s += ry * (lc - lam[i - P]) * (uc - u[i - P]);
      if (y + 1 < G.ny) s += ry * (lc - lam[i + P]) * (uc - u[i + P]);
      double g = -s * dk;
      if (e0) g += lc * rc * e0[i];
      grad[i] += g;
    }
  }
}

}  // namespace

extern "C" int tl_snapshot_reserve(int count) {
  TL_REQUIRE_INIT();
  State &s = S();
  if (count < 0) return fail(TL_ERR_ARG, "tl_snapshot_reserve: negative count");
  while ((int)s.snaps.size() < count) {
    double *q = nullptr;
    TL_HIP(hipMalloc((void **)&q, s.field_bytes));
    s.snaps.push_back(q);
    TL_HIP(hipMemsetAsync(q, 0, s.field_bytes, s.stream));
  }
  return TL_OK;
}

extern "C" int tl_snapshot_save(int slot, int field_id) {
  TL_REQUIRE_INIT();
  int rc = check_slot(slot);
  if (rc) return rc;
  State &s = S();
  Field *f = field_by_id(field_id);
  if (!f) return fail(TL_ERR_ARG, "unknown field id " + std::to_string(field_id));
  TL_HIP(hipMemcpyAsync(s.snaps[slot], f->base, s.field_bytes, hipMemcpyDeviceToDevice,
                        s.stream));
  return TL_OK;
}

extern "C" int tl_snapshot_load(int slot, int field_id) {
  TL_REQUIRE_INIT();
  int rc = check_slot(slot);
  if (rc) return rc;
  State &s = S();
  Field *f = field_by_id(field_id);
  if (!f) return fail(TL_ERR_ARG, "unknown field id " + std::to_string(field_id));
  TL_HIP(hipMemcpyAsync(f->base, s.snaps[slot], s.field_bytes, hipMemcpyDeviceToDevice,
                        s.stream));
  return TL_OK;
}

extern "C" int tl_snapshot_fill(int slot, double value) {
  TL_REQUIRE_INIT();
  int rc = check_slot(slot);
  if (rc) return rc;
  State &s = S();
  const long n = (long)(s.field_bytes / sizeof(double));
  const long nb = (n + BLOCK - 1) / BLOCK < 4096 ? (n + BLOCK - 1) / BLOCK : 4096;
  hipLaunchKernelGGL(k_fill, dim3((int)nb), dim3(BLOCK), 0, s.stream, s.snaps[slot], n, value);
  TL_HIP(hipGetLastError());
  return TL_OK;
}

// Host <-> slot copies of the cells inside both halos (as tl_upload).
static int copy_slot(int slot, double *host, int host_halo, bool up) {
  TL_REQUIRE_INIT();
  int rc = check_slot(slot);
  if (rc) return rc;
  return copy_host(snap11(slot), host, host_halo, up);
}

extern "C" int tl_snapshot_upload(int slot, const double *host, int host_halo) {
  return copy_slot(slot, const_cast<double *>(host), host_halo, true);
}

extern "C" int tl_snapshot_download(int slot, double *host, int host_halo) {
  return copy_slot(slot, host, host_halo, false);
}

extern "C" int tl_snapshot_upload_device(int slot, const double *dev, long ld) {
  TL_REQUIRE_INIT();
  int rc = check_slot(slot);
  if (rc) return rc;
  return copy_interior_device(snap11(slot), const_cast<double *>(dev), ld, true,
                              "tl_snapshot_upload_device");
}

extern "C" int tl_snapshot_download_device(int slot, double *dev, long ld) {
  TL_REQUIRE_INIT();
  int rc = check_slot(slot);
  if (rc) return rc;
  return copy_interior_device(snap11(slot), dev, ld, false, "tl_snapshot_download_device");
}

extern "C" int tl_adjoint_face_grad(int lambda_field, int u_slot, int e0_slot, int grad_slot) {
  TL_REQUIRE_INIT();
  State &s = S();
  int rc;
  if ((rc = check_slot(u_slot)) || (rc = check_slot(grad_slot))) return rc;
  if (e0_slot >= 0 && (rc = check_slot(e0_slot))) return rc;
  if (!s.zb_all) return fail(TL_ERR_ARG, "tl_adjoint_face_grad: needs all faces external");
  Field *lam = field_by_id(lambda_field);
  if (!lam) return fail(TL_ERR_ARG, "unknown field id " + std::to_string(lambda_field));
  const Geo G = geo(BLOCK);
  hipLaunchKernelGGL(k_face_grad, dim3(G.blocks), dim3(BLOCK), 0, s.stream, G, lam->p11,
                     snap11(u_slot), s.density.p11, e0_slot >= 0 ? snap11(e0_slot) : nullptr,
                     snap11(grad_slot), s.rx, s.ry,
                     s.coefficient == TL_RECIP_CONDUCTIVITY ? 1 : 0);
  TL_HIP(hipGetLastError());
  return TL_OK;
}
