// Internal definitions shared by the translation units of libtealeaf_hip.so.
//
// Device layout of every field: rows of `pitch` doubles, 256-byte aligned;
// interior cell (j,k) (1-based, as in the reference) lives at
//   base + (k-1+h)*pitch + PADL + (j-1)
// so each interior row starts on a 256-byte boundary and halo cells j<1 / j>nx
// and k<1 / k>ny are addressable up to depth h.  Kernels take a pointer to cell
// (1,1) and index it with 0-based (x, y) = (j-1, k-1).
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/tealeaf_hip.h"

namespace tl {

constexpr int PADL = 32;          // doubles before interior column 1 (256 B)
constexpr int BLOCK = 256;        // threads per block for streaming kernels
constexpr int WAVE = 64;          // CDNA wavefront width
constexpr int MAX_GRID = 2048;    // max blocks of a reduction kernel (= partials)
constexpr int NUM_XCD = 8;        // MI355X: 8 XCDs, blocks dealt round robin

// Boundary-read modes of the 5-point operator.
//  MIRROR: a neighbour outside the interior reads the mirrored interior cell,
//          which equals the reflective halo the reference refreshes before
//          the read, and is multiplied by a zero face coefficient otherwise.
//  PHYS:   read the stored halo cell (the reference's exact state, including
//          halos it leaves stale).
//  ZERO:   read 0 (the un = vector_r halo of the Jacobi kernel, which the
//          reference never writes).
enum Bnd { MIRROR = 0, PHYS = 1, ZERO = 2 };

struct Field {
  double *base = nullptr;  // allocation
  double *p11 = nullptr;   // cell (1,1)
};

struct State {
  bool init = false;
  int nx = 0, ny = 0, h = 1, device = 0;
  int cus = 0;             // compute units of the device
  long pitch = 0;          // doubles per row
  long rows = 0;           // ny + 2h
  size_t field_bytes = 0;
  hipStream_t stream = nullptr;

  // options
  int coefficient = TL_CONDUCTIVITY;
  int precond = TL_PREC_NONE;
  int reflective = 0;
  int ppcg_active = 0;
  int zb_all = 1;          // last tl_common_init had all four faces external
  double rx = 0.0, ry = 0.0;

  // fields (two buffers for the ones that are ping-ponged)
  Field density, energy0, energy1, u0, r, r_store, Mi, w, z, utemp, rtemp, Di,
      Kx, Ky, volume, cp, bfp;
  Field u[2], p[2], sd[2];
  int ucur = 0, pcur = 0, sdcur = 0;
  int walk_dir = 1;        // row direction of the next Chebyshev walk

  // reductions: partial sums per block, and small device scalars
  double *partials = nullptr;  // NPART arrays of MAX_GRID doubles
  double *dscal = nullptr;     // scratch scalars (results of finish kernels)
  double *hscal = nullptr;     // pinned host mirror of dscal
  void *ctl = nullptr;         // loop control block (device)
  void *hctl = nullptr;        // pinned host copies of the control block
  hipEvent_t loop_ev[2] = {nullptr, nullptr};  // batch markers of the loop driver
  double *hist = nullptr;      // loop history: rro, alpha, beta, error (4 x hist_len)
  long hist_len = 0;

  // FP32 fields of the mixed-precision solver (allocated on first use):
  // rows of pitch32 floats, cell (1,1) at f32 + k*field32 + pitch32 + PADL32
  float *f32 = nullptr;
  long pitch32 = 0;
  size_t field32 = 0;          // floats per field
  size_t f32_bytes = 0;        // the whole FP32 allocation

  // device staging buffer of host copies (allocated on first use)
  double *stage = nullptr;
  size_t stage_len = 0;       // doubles

  // snapshot slots (checkpoints and work arrays of the adjoint), each laid
  // out exactly like a field allocation
  std::vector<double *> snaps;

  // profiling
  bool prof = false;
  int prof_period = 1;      // time every prof_period-th launch of each class
  long prof_seq[8] = {0};   // launches seen per class while enabled
  struct Ev { hipEvent_t a, b; int cls; };
  std::vector<Ev> ev_pending;
  std::vector<hipEvent_t> ev_pool;
  long prof_count[8] = {0};
  double prof_ms[8] = {0};
  long prof_idle[8] = {0};  // timed launches that returned at once
  hipEvent_t arm_a = nullptr, arm_b = nullptr;  // events of a prof_arm region
  bool arm_fired = false;                       // ... taken by its launch

  std::string err;
};

State &S();

// partial-sum slots
enum Part { P0 = 0, P1, P2, P3, P4, P5, NPART };
inline double *part(int slot) { return S().partials + (long)slot * MAX_GRID; }

// set the error string and return the code
int fail(int code, const std::string &msg);
int hip_fail(hipError_t e, const char *what);

#define TL_HIP(call)                                         \
  do {                                                       \
    hipError_t e__ = (call);                                 \
    if (e__ != hipSuccess) return ::tl::hip_fail(e__, #call); \
  } while (0)

#define TL_REQUIRE_INIT()                                             \
  do {                                                                \
    if (!::tl::S().init) return ::tl::fail(TL_ERR_STATE, "tl_init has not been called"); \
  } while (0)

// launch geometry for streaming kernels over the nx*ny interior
struct Grid {
  int tiles_x = 0, tiles_y = 0, rows_per_tile = 1;
  int blocks = 0;          // launched blocks (<= MAX_GRID, multiple of NUM_XCD)
  long tiles = 0;
};
// Tiles of cols_per_tile columns and as many rows (<= max_rows) as give
// about target_tiles tiles; at most MAX_GRID blocks walk them.
Grid grid_for(int nx, int ny, int cols_per_tile = BLOCK, int target_tiles = MAX_GRID,
              int max_rows = 32);

// profiling helpers (no-ops unless enabled).  prof_begin / prof_end time a
// region between two marker events on the stream; prof_arm / prof_end time
// the one kernel of the region launched through klaunch(), whose dispatch
// records its own start and end (no marker packets around it).
void prof_begin(int cls, hipEvent_t *a);
void prof_arm(int cls, hipEvent_t *a);
void prof_end(int cls, hipEvent_t a);

// Launch on the library's stream; inside a prof_arm region the launch takes
// the region's events.
template <typename... F, typename... A>
inline void klaunch(void (*k)(F...), dim3 grid, dim3 block, A... args) {
  State &s = S();
  if (s.prof && s.arm_a && !s.arm_fired) {
    s.arm_fired = true;
    hipExtLaunchKernelGGL(k, grid, block, 0, s.stream, s.arm_a, s.arm_b, 0, args...);
  } else {
    hipLaunchKernelGGL(k, grid, block, 0, s.stream, args...);
  }
}

// finish a reduction: sum `n` partials of `slot` (deterministically) into
// dscal[idx]; returns after copying it to the host (synchronises)
int finish_sums(const int *slots, int nslots, int n, double *out);

// field of a TL_FIELD_* id (the current buffer of ping-ponged fields)
Field *field_by_id(int id);

// synchronous copy of the cells inside both halos between a caller's host
// array (halo host_halo) and the pitched device array whose cell (1,1) is
// at p11 (tl_runtime.hip)
int copy_host(double *p11, double *host, int host_halo, bool up);

// checks of a caller's ny x nx device array (row stride ld) and a
// synchronous copy of the interior cells between it and p11 (tl_runtime.hip)
int check_device_array(const double *p, long ld, const char *who);
int copy_interior_device(double *p11, double *dev, long ld, bool up, const char *who);

// Chebyshev iterations on the vector walk (tl_loops.hip): 1 = done,
// 0 = not applicable here, < 0 = -(error code)
int cheby_run_fused(const double *ch_alphas, const double *ch_betas, int step_first, int count,
                    int want_norm, double *norm);

}  // namespace tl

// ---------------------------------------------------------------------------
// device helpers
// ---------------------------------------------------------------------------
namespace tl {
namespace dev {

// The value of the next lower (lane - 1) or next higher (lane + 1) lane of
// the wave; lane 0 (lane 63) gets its own value back, as __shfl_up /
// __shfl_down(v, 1) give it.  TL_DPP = 1 moves it with a DPP wave shift
// (v_mov_b32 with wave_shr:1 / wave_shl:1, one VALU instruction per 32 bits:
// lane i takes lane i - 1 / i + 1, the lane without a source keeps `old`),
// TL_DPP = 0 through the LDS crossbar (ds_bpermute): the default; the DPP
// form measured within 0.5 % of it on every pass at 4096^2 (builds whose
// set-up kernels round alike, since the values in the fields change the
// passes' times).
#ifndef TL_DPP
#define TL_DPP 0
#endif
template <int CTRL>
__device__ inline unsigned dpp32(unsigned v) {
  return (unsigned)__builtin_amdgcn_update_dpp((int)v, (int)v, CTRL, 0xF, 0xF, false);
}
template <int CTRL, typename T>
__device__ inline T dpp_shift(T v) {
  static_assert(sizeof(T) == 4 || sizeof(T) == 8, "32- or 64-bit values");
  if constexpr (sizeof(T) == 8) {
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = dpp32<CTRL>((unsigned)b), hi = dpp32<CTRL>((unsigned)(b >> 32));
    return __builtin_bit_cast(T, (unsigned long long)lo | ((unsigned long long)hi << 32));
  } else {
    return __builtin_bit_cast(T, dpp32<CTRL>(__builtin_bit_cast(unsigned, v)));
  }
}
template <typename T>
__device__ inline T from_lower_lane(T v) {
#if TL_DPP
  return dpp_shift<0x138>(v);  // wave_shr:1
#else
  return __shfl_up(v, 1, WAVE);
#endif
}
template <typename T>
__device__ inline T from_upper_lane(T v) {
#if TL_DPP
  return dpp_shift<0x130>(v);  // wave_shl:1
#else
  return __shfl_down(v, 1, WAVE);
#endif
}

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int o = WAVE / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
  return v;
}

// Deterministic block sum: fixed lane tree per wave, then waves in order.
// Every thread returns the same total.  `lds` holds BLOCK/WAVE doubles.
__device__ inline double block_sum(double v, double *lds) {
  const int lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
  v = wave_sum(v);
  __syncthreads();
  if (lane == 0) lds[wid] = v;
  __syncthreads();
  double t = 0.0;
#pragma unroll
  for (int i = 0; i < BLOCK / WAVE; ++i) t += lds[i];
  return t;
}

// Sum of n partials, identical bits in every block of every launch.
__device__ inline double sum_partials(const double *__restrict__ p, int n, double *lds) {
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += BLOCK) s += p[i];
  return block_sum(s, lds);
}

// Tile walk shared by the streaming kernels: blocks are remapped so that the
// blocks of one XCD own a contiguous band of tiles (halo rows stay in one L2).
struct TileIter {
  long t, t_end;
  __device__ TileIter(long tiles, int blocks) {
    const int b = blockIdx.x;
    const bool remap = blocks > 0 && blocks % NUM_XCD == 0;
    if (blocks < 0) blocks = -blocks;  // negative count: plain round-robin order
    const int lb = remap ? (b % NUM_XCD) * (blocks / NUM_XCD) + b / NUM_XCD : b;
    const long base = tiles / blocks, rem = tiles % blocks;
    t = lb * base + (lb < rem ? lb : rem);
    t_end = t + base + (lb < rem ? 1 : 0);
  }
};

}  // namespace dev
}  // namespace tl
