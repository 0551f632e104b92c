// Lifecycle, field storage, host<->device copies, reduction finishing and
// kernel timing of libtealeaf_hip.so.
#include <algorithm>
#include <cstdio>
#include <cstring>

#include "tl_internal.hpp"
#include "tl_stencil.hpp"

namespace tl {

State &S() {
  static State s;
  return s;
}

int fail(int code, const std::string &msg) {
  S().err = msg;
  return code;
}

int hip_fail(hipError_t e, const char *what) {
  S().err = std::string(what) + ": " + hipGetErrorString(e);
  return TL_ERR_HIP;
}

Grid grid_for(int nx, int ny, int cols_per_tile, int target_tiles, int max_rows) {
  Grid g;
  g.tiles_x = (nx + cols_per_tile - 1) / cols_per_tile;
  long r = ((long)g.tiles_x * ny) / target_tiles;
  if (r < 1) r = 1;
  if (r > max_rows) r = max_rows;
  g.rows_per_tile = (int)r;
  g.tiles_y = (ny + g.rows_per_tile - 1) / g.rows_per_tile;
  g.tiles = (long)g.tiles_x * g.tiles_y;
  long b = g.tiles < MAX_GRID ? g.tiles : MAX_GRID;
  if (b >= NUM_XCD) b -= b % NUM_XCD;
  g.blocks = (int)b;
  return g;
}

// --------------------------------------------------------------------------
// kernel timing
// --------------------------------------------------------------------------
// Adds the pending events' durations to the per-class totals.  Launches
// queued past a loop's exit return at once: a launch shorter than a quarter
// of the median of its class among the pending events is one of those and is
// counted apart (prof_idle), so the totals describe launches that did the
// work.  (The median, not the maximum: a launch that the GPU reached before
// the host had queued it measures the host's gap too.)
static void prof_resolve() {
  State &s = S();
  (void)hipStreamSynchronize(s.stream);
  const size_t n = s.ev_pending.size();
  std::vector<float> ms(n, -1.f);
  std::vector<float> by_cls[8];
  for (size_t i = 0; i < n; ++i) {
    const auto &ev = s.ev_pending[i];
    if (hipEventElapsedTime(&ms[i], ev.a, ev.b) != hipSuccess) ms[i] = -1.f;
    else by_cls[ev.cls].push_back(ms[i]);
  }
  float cut[8];
  for (int c = 0; c < 8; ++c) {
    std::vector<float> &v = by_cls[c];
    cut[c] = 0.f;
    if (v.empty()) continue;
    std::nth_element(v.begin(), v.begin() + v.size() / 2, v.end());
    cut[c] = 0.25f * v[v.size() / 2];
  }
  for (size_t i = 0; i < n; ++i) {
    const auto &ev = s.ev_pending[i];
    if (ms[i] >= 0.f) {
      if (ms[i] < cut[ev.cls]) {
        s.prof_idle[ev.cls] += 1;
      } else {
        s.prof_ms[ev.cls] += ms[i];
        s.prof_count[ev.cls] += 1;
      }
    }
    s.ev_pool.push_back(ev.a);
    s.ev_pool.push_back(ev.b);
  }
  s.ev_pending.clear();
}

void prof_begin(int cls, hipEvent_t *a) {
  State &s = S();
  *a = nullptr;
  if (!s.prof || cls < 0 || cls >= 8) return;
  if ((s.prof_seq[cls]++ % s.prof_period) != 0) return;
  if (s.ev_pool.empty()) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return;
    s.ev_pool.push_back(e);
  }
  *a = s.ev_pool.back();
  s.ev_pool.pop_back();
  (void)hipEventRecord(*a, s.stream);
}

void prof_arm(int cls, hipEvent_t *a) {
  State &s = S();
  *a = nullptr;
  if (!s.prof || cls < 0 || cls >= 8) return;
  if ((s.prof_seq[cls]++ % s.prof_period) != 0) return;
  hipEvent_t e[2] = {nullptr, nullptr};
  for (hipEvent_t &x : e) {
    if (s.ev_pool.empty()) {
      if (hipEventCreate(&x) != hipSuccess) {
        for (hipEvent_t y : e)
          if (y) s.ev_pool.push_back(y);
        return;
      }
    } else {
      x = s.ev_pool.back();
      s.ev_pool.pop_back();
    }
  }
  s.arm_a = e[0];
  s.arm_b = e[1];
  s.arm_fired = false;
  *a = e[0];
}

void prof_end(int cls, hipEvent_t a) {
  State &s = S();
  if (!s.prof || !a) return;
  if (a == s.arm_a) {  // a prof_arm region: its launch recorded both events
    const hipEvent_t b = s.arm_b;
    const bool fired = s.arm_fired;
    s.arm_a = s.arm_b = nullptr;
    s.arm_fired = false;
    if (!fired) {  // nothing was launched
      s.ev_pool.push_back(a);
      s.ev_pool.push_back(b);
      return;
    }
    s.ev_pending.push_back({a, b, cls});
    if (s.ev_pending.size() >= 4096) prof_resolve();
    return;
  }
  hipEvent_t b;
  if (s.ev_pool.empty()) {
    if (hipEventCreate(&b) != hipSuccess) return;
  } else {
    b = s.ev_pool.back();
    s.ev_pool.pop_back();
  }
  (void)hipEventRecord(b, s.stream);
  s.ev_pending.push_back({a, b, cls});
  // keep the number of outstanding events bounded
  if (s.ev_pending.size() >= 4096) prof_resolve();
}

static void prof_drain() {
  if (!S().ev_pending.empty()) prof_resolve();
}

// --------------------------------------------------------------------------
// reduction finish: one block sums the partials of each requested slot
// --------------------------------------------------------------------------
__global__ __launch_bounds__(BLOCK) void k_finish(const double *__restrict__ partials, int n,
                                                  const int4 slots, int nslots,
                                                  double *__restrict__ out) {
  __shared__ double lds[BLOCK / WAVE];
  const int sl[4] = {slots.x, slots.y, slots.z, slots.w};
  for (int i = 0; i < nslots; ++i) {
    const double v = dev::sum_partials(partials + (long)sl[i] * MAX_GRID, n, lds);
    if (threadIdx.x == 0) out[i] = v;
    __syncthreads();
  }
}

int finish_sums(const int *slots, int nslots, int n, double *out) {
  State &s = S();
  int4 sl = make_int4(slots[0], nslots > 1 ? slots[1] : 0, nslots > 2 ? slots[2] : 0,
                      nslots > 3 ? slots[3] : 0);
  hipLaunchKernelGGL(k_finish, dim3(1), dim3(BLOCK), 0, s.stream, s.partials, n, sl, nslots,
                     s.dscal);
  TL_HIP(hipGetLastError());
  TL_HIP(hipMemcpyAsync(s.hscal, s.dscal, sizeof(double) * nslots, hipMemcpyDeviceToHost,
                        s.stream));
  TL_HIP(hipStreamSynchronize(s.stream));
  for (int i = 0; i < nslots; ++i) out[i] = s.hscal[i];
  return TL_OK;
}

}  // namespace tl

using namespace tl;

// --------------------------------------------------------------------------
// C ABI: lifecycle
// --------------------------------------------------------------------------
static int alloc_field(Field &f) {
  State &s = S();
  TL_HIP(hipMalloc((void **)&f.base, s.field_bytes));
  TL_HIP(hipMemsetAsync(f.base, 0, s.field_bytes, s.stream));
  f.p11 = f.base + (long)s.h * s.pitch + PADL;
  return TL_OK;
}

static void free_field(Field &f) {
  if (f.base) (void)hipFree(f.base);
  f.base = f.p11 = nullptr;
}

static Field *all_fields(int i) {
  State &s = S();
  Field *list[] = {&s.density, &s.energy0, &s.energy1, &s.u0, &s.r, &s.r_store, &s.Mi,
                   &s.w, &s.z, &s.utemp, &s.rtemp, &s.Di, &s.Kx, &s.Ky, &s.volume,
                   &s.cp, &s.bfp, &s.u[0], &s.u[1], &s.p[0], &s.p[1], &s.sd[0], &s.sd[1]};
  const int n = sizeof(list) / sizeof(list[0]);
  return i < n ? list[i] : nullptr;
}

extern "C" int tl_finalize(void) {
  State &s = S();
  if (!s.init) return TL_OK;
  if (s.stream) (void)hipStreamSynchronize(s.stream);
  for (int i = 0; Field *f = all_fields(i); ++i) free_field(*f);
  if (s.partials) (void)hipFree(s.partials);
  if (s.dscal) (void)hipFree(s.dscal);
  if (s.hscal) (void)hipHostFree(s.hscal);
  if (s.ctl) (void)hipFree(s.ctl);
  if (s.hctl) (void)hipHostFree(s.hctl);
  if (s.hist) (void)hipFree(s.hist);
  for (hipEvent_t e : s.loop_ev)
    if (e) (void)hipEventDestroy(e);
  if (s.f32) (void)hipFree(s.f32);
  if (s.stage) (void)hipFree(s.stage);
  for (double *q : s.snaps) (void)hipFree(q);
  for (auto &ev : s.ev_pending) {
    (void)hipEventDestroy(ev.a);
    (void)hipEventDestroy(ev.b);
  }
  for (auto e : s.ev_pool) (void)hipEventDestroy(e);
  if (s.stream) (void)hipStreamDestroy(s.stream);
  s = State();
  return TL_OK;
}

extern "C" int tl_init(int nx, int ny, int halo, int device) {
  State &s = S();
  if (s.init) return fail(TL_ERR_STATE, "tl_init called twice without tl_finalize");
  if (nx < 1 || ny < 1 || halo < 1 || halo > PADL)
    return fail(TL_ERR_ARG, "tl_init: need nx, ny >= 1 and 1 <= halo <= 32");
  int ndev = 0;
  TL_HIP(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) return fail(TL_ERR_ARG, "tl_init: no such HIP device");
  TL_HIP(hipSetDevice(device));
  TL_HIP(hipDeviceGetAttribute(&s.cus, hipDeviceAttributeMultiprocessorCount, device));
  s.nx = nx;
  s.ny = ny;
  s.h = halo;
  s.device = device;
  // two spare columns past the halo: paired loads may touch column nx+1
  s.pitch = ((long)PADL + nx + halo + 2 + 31) / 32 * 32;
  s.rows = (long)ny + 2L * halo;
  s.field_bytes = sizeof(double) * (size_t)(s.pitch * s.rows + 64);
  TL_HIP(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking));
  s.init = true;  // so that tl_finalize cleans up on a failure below
  int rc;
  for (int i = 0; Field *f = all_fields(i); ++i)
    if ((rc = alloc_field(*f)) != TL_OK) {
      std::string m = s.err;
      tl_finalize();
      return fail(rc, m);
    }
  TL_HIP(hipMalloc((void **)&s.partials, sizeof(double) * MAX_GRID * NPART));
  TL_HIP(hipMemsetAsync(s.partials, 0, sizeof(double) * MAX_GRID * NPART, s.stream));
  TL_HIP(hipMalloc((void **)&s.dscal, sizeof(double) * 64));
  TL_HIP(hipHostMalloc((void **)&s.hscal, sizeof(double) * 64, hipHostMallocDefault));
  // the fused loops' control block, its host copies and the loop driver's batch events
  // (set up here so that a loop call allocates nothing but its history)
  TL_HIP(hipMalloc(&s.ctl, 64));
  TL_HIP(hipMemsetAsync(s.ctl, 0, 64, s.stream));
  TL_HIP(hipHostMalloc(&s.hctl, 256, hipHostMallocDefault));
  TL_HIP(hipEventCreateWithFlags(&s.loop_ev[0], hipEventDisableTiming));
  TL_HIP(hipEventCreateWithFlags(&s.loop_ev[1], hipEventDisableTiming));
  TL_HIP(hipStreamSynchronize(s.stream));
  return TL_OK;
}

extern "C" const char *tl_last_error(void) { return S().err.c_str(); }

extern "C" int tl_set_options(int coefficient, int preconditioner, int reflective) {
  TL_REQUIRE_INIT();
  State &s = S();
  if (coefficient != TL_CONDUCTIVITY && coefficient != TL_RECIP_CONDUCTIVITY)
    return fail(TL_ERR_ARG, "tea_leaf: unknown coefficient option");
  if (preconditioner < TL_PREC_NONE || preconditioner > TL_PREC_JAC_BLOCK)
    return fail(TL_ERR_ARG, "tl_set_options: invalid preconditioner type");
  if (preconditioner == TL_PREC_JAC_BLOCK && s.h > 1)
    return fail(TL_ERR_ARG,
                "read_input: Unable to use nonstandard halo depth with block jacobi preconditioner");
  s.coefficient = coefficient;
  s.precond = preconditioner;
  s.reflective = reflective ? 1 : 0;
  return TL_OK;
}

extern "C" int tl_set_ppcg_active(int active) {
  TL_REQUIRE_INIT();
  S().ppcg_active = active ? 1 : 0;
  return TL_OK;
}

extern "C" int tl_synchronize(void) {
  TL_REQUIRE_INIT();
  TL_HIP(hipStreamSynchronize(S().stream));
  return TL_OK;
}

namespace tl {
Field *field_by_id(int id) {
  State &s = S();
  switch (id) {
    case TL_FIELD_DENSITY: return &s.density;
    case TL_FIELD_ENERGY0: return &s.energy0;
    case TL_FIELD_ENERGY1: return &s.energy1;
    case TL_FIELD_U: return &s.u[s.ucur];
    case TL_FIELD_P: return &s.p[s.pcur];
    case TL_FIELD_SD: return &s.sd[s.sdcur];
    case TL_FIELD_R: return &s.r;
    case TL_FIELD_Z: return &s.z;
    case TL_FIELD_KX: return &s.Kx;
    case TL_FIELD_KY: return &s.Ky;
    case TL_FIELD_DI: return &s.Di;
    case TL_FIELD_U0: return &s.u0;
    case TL_FIELD_W: return &s.w;
    case TL_FIELD_R_STORE: return &s.r_store;
    case TL_FIELD_MI: return &s.Mi;
    case TL_FIELD_UTEMP: return &s.utemp;
    case TL_FIELD_RTEMP: return &s.rtemp;
    case TL_FIELD_VOLUME: return &s.volume;
    case TL_FIELD_CP: return &s.cp;
    case TL_FIELD_BFP: return &s.bfp;
    default: return nullptr;
  }
}
}  // namespace tl

namespace tl {
// The device staging buffer, grown to n doubles; null if that fails.
static double *staging(size_t n) {
  State &s = S();
  if (s.stage_len < n) {
    if (s.stage) (void)hipFree(s.stage);
    s.stage = nullptr;
    s.stage_len = 0;
    if (hipMalloc((void **)&s.stage, n * sizeof(double)) != hipSuccess) {
      (void)hipGetLastError();  // clear it: the caller copies without staging
      s.stage = nullptr;
      return nullptr;
    }
    s.stage_len = n;
  }
  return s.stage;
}

// A 2D copy straight between pageable host memory and a pitched device array
// moves about 3 GB/s on MI355X; one contiguous copy over PCIe to or from the
// staging buffer plus a 2D copy on the device is several times faster.  A
// download stages only when the host rows are contiguous (the copy back then
// writes no host cell outside the rectangle).
int copy_host(double *p11, double *host, int host_halo, bool up) {
  State &s = S();
  if (!host) return fail(TL_ERR_ARG, "null host pointer");
  if (host_halo < 0) return fail(TL_ERR_ARG, "negative host halo");
  const int hh = host_halo < s.h ? host_halo : s.h;
  const long hw = (long)s.nx + 2L * host_halo;           // host row length
  const long cols = (long)s.nx + 2L * hh, rows = (long)s.ny + 2L * hh;
  double *hp = host + (long)(host_halo - hh) + (long)(host_halo - hh) * hw;
  double *dp = p11 - hh - (long)hh * s.pitch;
  const size_t span = (size_t)((rows - 1) * hw + cols);  // host doubles from hp to the last cell
  double *st = (up || hw == cols) ? staging(span) : nullptr;
  if (up) {
    if (st) {
      TL_HIP(hipMemcpyAsync(st, hp, span * sizeof(double), hipMemcpyHostToDevice, s.stream));
      TL_HIP(hipMemcpy2DAsync(dp, s.pitch * sizeof(double), st, hw * sizeof(double),
                              cols * sizeof(double), rows, hipMemcpyDeviceToDevice, s.stream));
    } else {
      TL_HIP(hipMemcpy2DAsync(dp, s.pitch * sizeof(double), hp, hw * sizeof(double),
                              cols * sizeof(double), rows, hipMemcpyHostToDevice, s.stream));
    }
  } else if (st) {
    TL_HIP(hipMemcpy2DAsync(st, cols * sizeof(double), dp, s.pitch * sizeof(double),
                            cols * sizeof(double), rows, hipMemcpyDeviceToDevice, s.stream));
    TL_HIP(hipMemcpyAsync(hp, st, span * sizeof(double), hipMemcpyDeviceToHost, s.stream));
  } else {
    TL_HIP(hipMemcpy2DAsync(hp, hw * sizeof(double), dp, s.pitch * sizeof(double),
                            cols * sizeof(double), rows, hipMemcpyDeviceToHost, s.stream));
  }
  TL_HIP(hipStreamSynchronize(s.stream));
  return TL_OK;
}
}  // namespace tl

// Copy the rectangle of cells inside both halos.
static int copy_field(int id, double *host, int host_halo, bool up) {
  TL_REQUIRE_INIT();
  Field *f = field_by_id(id);
  if (!f) return fail(TL_ERR_ARG, "unknown field id " + std::to_string(id));
  return copy_host(f->p11, host, host_halo, up);
}

extern "C" int tl_upload(int field_id, const double *host, int host_halo) {
  return copy_field(field_id, const_cast<double *>(host), host_halo, true);
}

extern "C" int tl_download(int field_id, double *host, int host_halo) {
  return copy_field(field_id, host, host_halo, false);
}

namespace tl {
// An ny x nx array in device memory of the library's GPU with row stride ld
// (in doubles): checked on the host, so that a host pointer or another GPU's
// memory is refused instead of being handed to a copy.
int check_device_array(const double *p, long ld, const char *who) {
  State &s = S();
  if (!p) return fail(TL_ERR_ARG, std::string(who) + ": null device pointer");
  if (ld < s.nx) return fail(TL_ERR_ARG, std::string(who) + ": row stride below nx");
  hipPointerAttribute_t a;
  const hipError_t e = hipPointerGetAttributes(&a, p);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return fail(TL_ERR_ARG, std::string(who) + ": not device memory");
  }
  if (a.type != hipMemoryTypeDevice || a.device != s.device)
    return fail(TL_ERR_ARG, std::string(who) + ": not device memory of the library's GPU");
  return TL_OK;
}

// Interior cells of a device-side field base (pitched, halo offset) <->
// a dense ny x nx device array with row stride ld; synchronous.
int copy_interior_device(double *p11, double *dev, long ld, bool up, const char *who) {
  State &s = S();
  int rc = check_device_array(dev, ld, who);
  if (rc) return rc;
  if (up)
    TL_HIP(hipMemcpy2DAsync(p11, s.pitch * sizeof(double), dev, ld * sizeof(double),
                            (size_t)s.nx * sizeof(double), s.ny, hipMemcpyDeviceToDevice,
                            s.stream));
  else
    TL_HIP(hipMemcpy2DAsync(dev, ld * sizeof(double), p11, s.pitch * sizeof(double),
                            (size_t)s.nx * sizeof(double), s.ny, hipMemcpyDeviceToDevice,
                            s.stream));
  TL_HIP(hipStreamSynchronize(s.stream));
  return TL_OK;
}
}  // namespace tl

extern "C" int tl_upload_device(int field_id, const double *dev, long ld) {
  TL_REQUIRE_INIT();
  Field *f = field_by_id(field_id);
  if (!f) return fail(TL_ERR_ARG, "unknown field id " + std::to_string(field_id));
  return copy_interior_device(f->p11, const_cast<double *>(dev), ld, true, "tl_upload_device");
}

extern "C" int tl_download_device(int field_id, double *dev, long ld) {
  TL_REQUIRE_INIT();
  Field *f = field_by_id(field_id);
  if (!f) return fail(TL_ERR_ARG, "unknown field id " + std::to_string(field_id));
  return copy_interior_device(f->p11, dev, ld, false, "tl_download_device");
}

// --------------------------------------------------------------------------
// C ABI: profiling
// --------------------------------------------------------------------------
extern "C" int tl_profile_enable(int on) {
  TL_REQUIRE_INIT();
  if (on < 0) return fail(TL_ERR_ARG, "tl_profile_enable: negative period");
  S().prof = on != 0;
  S().prof_period = on > 1 ? on : 1;
  return TL_OK;
}

extern "C" int tl_device_bytes(double *bytes) {
  TL_REQUIRE_INIT();
  if (!bytes) return fail(TL_ERR_ARG, "null pointer");
  const State &s = S();
  double b = 0.0;
  for (int i = 0; all_fields(i); ++i) b += (double)s.field_bytes;
  b += (double)s.snaps.size() * (double)s.field_bytes;
  b += (double)s.f32_bytes;
  b += sizeof(double) * (double)MAX_GRID * NPART;
  if (s.hist) b += sizeof(double) * 4.0 * (double)s.hist_len;
  b += sizeof(double) * (double)s.stage_len;
  *bytes = b;
  return TL_OK;
}

extern "C" int tl_profile_get(int cls, long *launches, double *total_ms) {
  TL_REQUIRE_INIT();
  if (cls < 0 || cls >= 8) return fail(TL_ERR_ARG, "bad kernel class");
  prof_drain();
  if (launches) *launches = S().prof_count[cls];
  if (total_ms) *total_ms = S().prof_ms[cls];
  return TL_OK;
}

extern "C" int tl_profile_idle(int cls, long *launches) {
  TL_REQUIRE_INIT();
  if (cls < 0 || cls >= 8) return fail(TL_ERR_ARG, "bad kernel class");
  prof_drain();
  if (launches) *launches = S().prof_idle[cls];
  return TL_OK;
}

extern "C" int tl_profile_reset(void) {
  TL_REQUIRE_INIT();
  prof_drain();
  for (int i = 0; i < 8; ++i) {
    S().prof_count[i] = 0;
    S().prof_ms[i] = 0.0;
    S().prof_seq[i] = 0;
    S().prof_idle[i] = 0;
  }
  return TL_OK;
}
