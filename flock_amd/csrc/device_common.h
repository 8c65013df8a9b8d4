// device_common.h -- shared device-side types and wavefront/workgroup primitives for
// the gfx950 transportation-simplex engine.  64-lane wavefronts throughout.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace flock {

constexpr int kWave = 64;

// Global-memory (address space 1) pointers: loads and stores through them compile to
// global_load/global_store with a scalar base instead of 64-bit flat addressing.
#define FLOCK_GLOBAL __attribute__((address_space(1)))
// LDS (address space 3) pointers: ds_read/ds_write instead of flat accesses for data
// that is known to live in LDS.
#define FLOCK_LDS __attribute__((address_space(3)))
template <class T>
__device__ __forceinline__ FLOCK_GLOBAL T* gl(T* p) { return (FLOCK_GLOBAL T*)p; }
typedef int32_t i32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// Values that are the same for the whole workgroup, moved to scalar registers.
__device__ __forceinline__ int32_t uniform(int32_t x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ uint64_t uniform64(uint64_t x) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)x);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(x >> 32));
  return ((uint64_t)hi << 32) | lo;
}
template <class T>
__device__ __forceinline__ T* uniform_ptr(T* p) {
  const uint64_t v = (uint64_t)p;
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return (T*)(((uint64_t)hi << 32) | lo);
}

constexpr int32_t kBig = 2147483647;      // INT_MAX sentinel of the Vogel two-min scan
// Cost written into the padding cells of every row (columns n .. ld-1) at the start of a
// solve, with their duals held at 0: the range contract ((2(m+n)+1) * max cost < 2^31,
// costs >= 0) bounds every dual below 2^30, so a padding cell prices 2^30 - u > 0.
constexpr int32_t kPadCost = 1 << 30;

__device__ __forceinline__ int32_t imin(int32_t a, int32_t b) { return a < b ? a : b; }
constexpr uint64_t kNoKey = ~0ull;

// Per-instance descriptor in device memory.  Inputs are written by the host before a
// launch; outputs (objective .. status) and the basis/dual arrays are written by the
// engine.  Matrices are row-major with a padded pitch (multiple of 4 int32 = 16 bytes):
//   cost  : m rows x ld  (cell (i,j) at cost[i*ld + j])
//   costT : n rows x ldT (the same matrix transposed, filled by the engine itself)
struct DevInstance {
  const int32_t* cost;
  int32_t* costT;
  const int32_t* supply;
  const int32_t* demand;
  int32_t m, n, ld, ldT;
  int32_t* basis_row;       // m+n-1, by basis slot
  int32_t* basis_col;
  int32_t* basis_flow;
  int32_t* dual;            // m+n: u (rows) then v (columns), u[0] = 0
  int32_t* scratch;         // global scratch for the tree build when LDS is short
  int32_t* state;           // resumable tree state (server steps); may be null
  int64_t objective;
  int64_t iterations;
  int32_t status;           // flock_status
  int32_t phase;            // 0 = fresh, 1 = simplex in progress (state valid), 2 = done
  int32_t n_flows;          // basic arcs with positive flow
  int32_t mode;             // switch rule: 0 multi-pivot rounds, 1 switched to Dantzig
  int64_t sweeps;           // full reduced-cost sweeps performed
  // device wall-clock ticks spent per phase (thread 0's view, summed over steps):
  // 0 transpose+Vogel, 1 tree build + duals, 2 full sweeps, 3 candidate re-pricing,
  // 4 pivot cycle discovery + leaving arc, 5 pivot flow update + re-rooting
  int64_t ticks[6];
  uint64_t t_start;         // device wall clock when the instance's solve began / ended
  uint64_t t_done;
};

__device__ __forceinline__ uint64_t now_ticks() { return wall_clock64(); }

struct EngineParams {
  int64_t max_iterations;   // <= 0: 50*V + 100000
  int32_t initial;          // 0 Vogel, 1 NWC
  int32_t pricing;          // 0 Dantzig, 1 candidate list, 2 multi-pivot, 4 switch (multi, then Dantzig)
  int32_t step_pivots;      // > 0: stop after this many pivots and save state (server)
  int32_t minor_limit;      // candidate list: minor iterations per sweep
  int32_t lds_bytes;        // dynamic LDS bytes available to each workgroup
  int32_t prio_count;       // the first prio_count queue positions (largest work first) run at raised wave priority
  int32_t tree_global;      // large-instance path: the basis tree lives in the state buffer (global memory)
  int32_t pad_;
  // time slicing (server): when slice_ticks > 0, no instance is started and no pivot is
  // begun later than slice_ticks device-clock ticks after *clock0 (the launch stamp the
  // reset kernel writes); the kernel turns that into `deadline`
  uint64_t slice_ticks;
  const uint64_t* clock0;
  uint64_t deadline;
  // time slicing: the number of groups of this launch inside a phase that cannot stop at
  // the deadline (a fresh instance's Vogel and tree build).  Past the deadline, groups
  // that are pivoting keep pivoting while it is non-zero (the launch cannot end before
  // those phases do), so the cores do not idle; no instance starts after the deadline,
  // so the count only falls.  Null: stop at the deadline.
  int32_t* busy;
};

// ---------------------------------------------------------------------------
// reductions
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t shfl_xor_u64(uint64_t x, int mask) {
  uint32_t lo = (uint32_t)x, hi = (uint32_t)(x >> 32);
  lo = __shfl_xor(lo, mask, kWave);
  hi = __shfl_xor(hi, mask, kWave);
  return ((uint64_t)hi << 32) | lo;
}

// DPP lane moves (GFX9 data-parallel primitives).  row_shr:k (0x110 + k) moves lane
// i - k of the same 16-lane row into lane i; row_bcast:15 (0x142) moves lane 15 of row r
// into every lane of row r + 1; row_bcast:31 (0x143) moves lane 31 into rows 2 and 3.
// Lanes without a source, or in rows outside kRowMask, receive `old`.
template <int kCtrl, int kRowMask>
__device__ __forceinline__ uint64_t dpp_u64(uint64_t old, uint64_t src) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp((int)(uint32_t)old, (int)(uint32_t)src, kCtrl, kRowMask,
                                                            0xf, false);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp((int)(uint32_t)(old >> 32), (int)(uint32_t)(src >> 32),
                                                            kCtrl, kRowMask, 0xf, false);
  return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ uint64_t readlane_u64(uint64_t x, int lane) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)x, lane);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(x >> 32), lane);
  return ((uint64_t)hi << 32) | lo;
}
// Wavefront reduction of an associative, commutative op with identity `idn`: row_shr
// 1, 2, 4, 8 leave each row's total in its lane 15 (the two windows merged at each step
// are disjoint), the two broadcasts gather the row totals into lane 63, and readlane
// hands lane 63's value to every lane (a wave-uniform result).  Every lane of the
// wavefront must be active.
template <class Op>
__device__ __forceinline__ uint64_t wave_reduce_u64(uint64_t x, uint64_t idn, Op op) {
  x = op(x, dpp_u64<0x111, 0xf>(idn, x));
  x = op(x, dpp_u64<0x112, 0xf>(idn, x));
  x = op(x, dpp_u64<0x114, 0xf>(idn, x));
  x = op(x, dpp_u64<0x118, 0xf>(idn, x));
  x = op(x, dpp_u64<0x142, 0xa>(idn, x));
  x = op(x, dpp_u64<0x143, 0xc>(idn, x));
  return readlane_u64(x, 63);
}
// 32-bit wavefront min / sum, same DPP pattern on one register.
template <int kCtrl, int kRowMask>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t old, uint32_t src) {
  return (uint32_t)__builtin_amdgcn_update_dpp((int)old, (int)src, kCtrl, kRowMask, 0xf, false);
}
template <class Op>
__device__ __forceinline__ uint32_t wave_reduce_u32(uint32_t x, uint32_t idn, Op op) {
  x = op(x, dpp_u32<0x111, 0xf>(idn, x));
  x = op(x, dpp_u32<0x112, 0xf>(idn, x));
  x = op(x, dpp_u32<0x114, 0xf>(idn, x));
  x = op(x, dpp_u32<0x118, 0xf>(idn, x));
  x = op(x, dpp_u32<0x142, 0xa>(idn, x));
  x = op(x, dpp_u32<0x143, 0xc>(idn, x));
  return (uint32_t)__builtin_amdgcn_readlane((int)x, 63);
}
struct MaxU32 { __device__ __forceinline__ uint32_t operator()(uint32_t a, uint32_t b) const { return b > a ? b : a; } };
struct AddU32 { __device__ __forceinline__ uint32_t operator()(uint32_t a, uint32_t b) const { return a + b; } };
// Max and sum have identity 0, which is what a DPP move with bound_ctrl writes into lanes
// without a source, so the compiler folds each move into the operation (v_max_u32_dpp,
// one instruction per step).  Min is taken as the complement of the max of complements.
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t x) { return wave_reduce_u32(x, 0u, MaxU32{}); }
__device__ __forceinline__ uint32_t wave_min_u32(uint32_t x) { return ~wave_max_u32(~x); }
__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t x) { return wave_reduce_u32(x, 0u, AddU32{}); }

struct AddU64 { __device__ __forceinline__ uint64_t operator()(uint64_t a, uint64_t b) const { return a + b; } };

// 64-bit max as two 32-bit passes: the high words' maximum, then the largest low word
// among the lanes holding it (unsigned 64-bit order is (high, low) order); min through
// complements.  Both passes use the folded 32-bit form above.
__device__ __forceinline__ uint64_t wave_max_u64(uint64_t x) {
  const uint32_t hi = wave_max_u32((uint32_t)(x >> 32));
  const uint32_t lo = wave_max_u32((uint32_t)(x >> 32) == hi ? (uint32_t)x : 0u);
  return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ uint64_t wave_min_u64(uint64_t x) { return ~wave_max_u64(~x); }
__device__ __forceinline__ int64_t wave_sum_i64(int64_t x) {
  return (int64_t)wave_reduce_u64((uint64_t)x, 0ull, AddU64{});
}

// Workgroup reductions.  `red` is an LDS array of at least T/64 + 1 entries.  The
// result slot is written only after a barrier that every wave passes after reading the
// previous result, so back-to-back calls on the same `red` are safe.
template <int T>
__device__ __forceinline__ uint64_t block_min_u64(uint64_t x, uint64_t* red) {
  constexpr int W = T / kWave;
  x = wave_min_u64(x);
  const int w = threadIdx.x / kWave, l = threadIdx.x % kWave;
  if (l == 0) red[w] = x;
  __syncthreads();
  if (w == 0) {
    uint64_t y = l < W ? red[l] : kNoKey;
    y = wave_min_u64(y);
    if (l == 0) red[W] = y;
  }
  __syncthreads();
  return red[W];
}

template <int T>
__device__ __forceinline__ uint64_t block_max_u64(uint64_t x, uint64_t* red) {
  constexpr int W = T / kWave;
  x = wave_max_u64(x);
  const int w = threadIdx.x / kWave, l = threadIdx.x % kWave;
  if (l == 0) red[w] = x;
  __syncthreads();
  if (w == 0) {
    uint64_t y = l < W ? red[l] : 0ull;
    y = wave_max_u64(y);
    if (l == 0) red[W] = y;
  }
  __syncthreads();
  return red[W];
}

template <int T>
__device__ __forceinline__ int64_t block_sum_i64(int64_t x, uint64_t* red) {
  constexpr int W = T / kWave;
  x = wave_sum_i64(x);
  const int w = threadIdx.x / kWave, l = threadIdx.x % kWave;
  if (l == 0) red[w] = (uint64_t)x;
  __syncthreads();
  if (w == 0) {
    int64_t y = l < W ? (int64_t)red[l] : 0;
    y = wave_sum_i64(y);
    if (l == 0) red[W] = (uint64_t)y;
  }
  __syncthreads();
  return (int64_t)red[W];
}

// Workgroup min / max with one barrier: every wavefront writes its result into `buf`
// (at least T/64 entries) and every lane then reads all of them.  A buffer must not be
// written again before every lane has read it: callers alternate buffers so that
// another barrier always separates two uses of the same one.
template <int T>
__device__ __forceinline__ uint64_t block_min1_u64(uint64_t x, uint64_t* buf) {
  constexpr int W = T / kWave;
  x = wave_min_u64(x);
  if (threadIdx.x % kWave == 0) buf[threadIdx.x / kWave] = x;
  __syncthreads();
  uint64_t y = buf[0];
#pragma unroll
  for (int k = 1; k < W; ++k) y = buf[k] < y ? buf[k] : y;
  return y;
}
template <int T>
__device__ __forceinline__ uint64_t block_max1_u64(uint64_t x, uint64_t* buf) {
  constexpr int W = T / kWave;
  x = wave_max_u64(x);
  if (threadIdx.x % kWave == 0) buf[threadIdx.x / kWave] = x;
  __syncthreads();
  uint64_t y = buf[0];
#pragma unroll
  for (int k = 1; k < W; ++k) y = buf[k] > y ? buf[k] : y;
  return y;
}

// (value, index) packed so that unsigned 64-bit order is (signed value, index) order.
__device__ __forceinline__ uint64_t pack_vi(int32_t v, uint32_t idx) {
  return ((uint64_t)((uint32_t)v ^ 0x80000000u) << 32) | idx;
}
__device__ __forceinline__ int32_t key_val(uint64_t k) { return (int32_t)((uint32_t)(k >> 32) ^ 0x80000000u); }
__device__ __forceinline__ uint32_t key_idx(uint64_t k) { return (uint32_t)k; }

}  // namespace flock
