// CPU stand-in for the device-side HIP API used by flock_amd/csrc/engine.h, so that
// the engine's device code can be compiled for the host and run under
// AddressSanitizer / ThreadSanitizer.  TEST INFRASTRUCTURE ONLY: built by
// tests/emu/Makefile into tests/emu/_build/, never part of the product library.
//
// Execution model: one OS thread per lane.  Every lane knows its thread index, its
// workgroup's barrier and its wavefront's barrier and exchange buffer.  __syncthreads
// waits on the workgroup barrier; wavefront shuffles go through the exchange buffer
// between two wavefront barriers; the wavefront-scope fence/barrier pair that the
// engine uses as its wavefront sync point also waits on the wavefront barrier (lanes
// are NOT otherwise kept in lockstep, which is stricter than the hardware).
//
// With EMU_FIBERS defined, lanes are cooperative fibers on one OS thread instead
// (ucontext; a lane runs until it blocks on a barrier), which is deterministic and much
// faster; the sanitizer builds use real threads so ThreadSanitizer sees the lanes.
#pragma once
#include <atomic>
#include <barrier>
#include <chrono>
#include <cstdint>
#include <cstring>
#include <cstdio>
#include <cstdlib>
#include <vector>

#define __device__
#define __host__
#define __global__
#define __forceinline__ inline
#define __shared__

namespace emu {
struct Dim3 { unsigned x, y, z; };
// A barrier that also checks that every participant arrived from the same call site;
// a mismatch means the lanes' synchronisation sequences diverged (a bug in the code
// under test), which is reported and aborts the process.
#ifdef EMU_FIBERS
void fiber_block_until(const int* gen, int seen);
#endif
struct TagBarrier {
  explicit TagBarrier(int n) : count(n), tags(n, 0), bar(n, Check{this}) {}
#ifdef EMU_FIBERS
  int arrived = 0, generation = 0;
  void wait(int who, int tag) {
    tags[who] = tag;
    if (++arrived == count) {
      Check{this}();
      arrived = 0;
      ++generation;
      return;
    }
    fiber_block_until(&generation, generation);
  }
#else
  void wait(int who, int tag) { tags[who] = tag; bar.arrive_and_wait(); }
#endif
  struct Check {
    TagBarrier* self;
    void operator()() noexcept {
      for (int i = 1; i < self->count; ++i)
        if (self->tags[i] != self->tags[0]) {
          std::fprintf(stderr, "emu: barrier mismatch: lane 0 at %d, lane %d at %d\n", self->tags[0], i, self->tags[i]);
          std::abort();
        }
    }
  };
  int count;
  std::vector<int> tags;
  std::barrier<Check> bar;
};
struct Lane {
  unsigned tid;                        // threadIdx.x
  TagBarrier* block;                   // workgroup barrier
  TagBarrier* wave;                    // wavefront barrier
  uint64_t* xchg;                      // 64-entry exchange buffer of the wavefront
};
#ifdef EMU_FIBERS
inline Lane lane_ctx;  // the running fiber's lane (set by the scheduler on every switch)
#else
inline thread_local Lane lane_ctx;
#endif
inline Dim3 thread_idx() { return Dim3{lane_ctx.tid, 0, 0}; }
}  // namespace emu

#define threadIdx (::emu::thread_idx())

#define __syncthreads() (emu::lane_ctx.block->wait((int)emu::lane_ctx.tid, __LINE__))
inline void __threadfence_block() { std::atomic_thread_fence(std::memory_order_seq_cst); }
inline void __threadfence() { std::atomic_thread_fence(std::memory_order_seq_cst); }

inline uint64_t emu_exchange(uint64_t v, unsigned src) {
  const unsigned lane = emu::lane_ctx.tid % 64;
  emu::lane_ctx.xchg[lane] = v;
  emu::lane_ctx.wave->wait((int)lane, -1);
  const uint64_t r = emu::lane_ctx.xchg[src % 64];
  emu::lane_ctx.wave->wait((int)lane, -2);
  return r;
}
inline uint32_t __shfl_xor(uint32_t v, int mask, int width = 64) {
  (void)width;
  return (uint32_t)emu_exchange(v, (emu::lane_ctx.tid % 64) ^ (unsigned)mask);
}
inline int __shfl(int v, int src, int width = 64) {
  (void)width;
  return (int)(uint32_t)emu_exchange((uint32_t)v, (unsigned)src);
}
// 64-lane ballot: every lane publishes its predicate, then reads all 64
inline uint64_t __ballot(int pred) {
  const unsigned lane = emu::lane_ctx.tid % 64;
  emu::lane_ctx.xchg[lane] = pred ? 1 : 0;
  emu::lane_ctx.wave->wait((int)lane, -3);
  uint64_t m = 0;
  for (int i = 0; i < 64; ++i) m |= (emu::lane_ctx.xchg[i] ? 1ull : 0ull) << i;
  emu::lane_ctx.wave->wait((int)lane, -4);
  return m;
}
// DPP moves used by the wavefront reductions: row_shr:k (0x111..0x11f), row_bcast:15
// (0x142), row_bcast:31 (0x143).  A lane whose row or bank is masked off, or that has no
// source lane, receives `old` (bound_ctrl false); other controls abort.
namespace emu {
inline int update_dpp(int old, int src, int ctrl, int row_mask, int bank_mask, bool bound_ctrl) {
  const int lane = (int)(lane_ctx.tid % 64), row = lane / 16, idx = lane % 16;
  int from = -1;
  if (ctrl >= 0x111 && ctrl <= 0x11f) { if (idx >= ctrl - 0x110) from = lane - (ctrl - 0x110); }
  else if (ctrl == 0x142) { if (row > 0) from = 16 * row - 1; }
  else if (ctrl == 0x143) { if (row > 1) from = 31; }
  else { std::fprintf(stderr, "emu: unsupported DPP control 0x%x\n", ctrl); std::abort(); }
  const int got = (int)(uint32_t)emu_exchange((uint32_t)src, from < 0 ? (unsigned)lane : (unsigned)from);
  if (!((row_mask >> row) & 1) || !((bank_mask >> (idx / 4)) & 1)) return old;
  if (from < 0) return bound_ctrl ? 0 : old;
  return got;
}
}  // namespace emu
#define __builtin_amdgcn_update_dpp(o, s, c, r, b, bc) (emu::update_dpp((o), (s), (c), (r), (b), (bc)))
#define __builtin_amdgcn_readlane(x, l) ((int)(uint32_t)emu_exchange((uint32_t)(x), (unsigned)(l)))
inline int __ffsll(unsigned long long x) { return __builtin_ffsll((long long)x); }
inline int __popcll(unsigned long long x) { return __builtin_popcountll(x); }

#define __builtin_amdgcn_readfirstlane(x) (x)
inline int __clz(int x) { return x == 0 ? 32 : __builtin_clz((unsigned)x); }
#define __builtin_amdgcn_fence(order, scope) std::atomic_thread_fence(std::memory_order_seq_cst)
#define __builtin_amdgcn_wave_barrier() (emu::lane_ctx.wave->wait((int)(emu::lane_ctx.tid % 64), __LINE__))

inline uint64_t wall_clock64() {
  return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
             std::chrono::steady_clock::now().time_since_epoch()).count() / 10;
}

template <class T>
inline T atomicAdd(T* p, T v) { return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST); }

// device-scope atomics and the sleep hint (spin loops yield the lane)
#define __HIP_MEMORY_SCOPE_AGENT 3
#define __hip_atomic_load(p, order, scope) __atomic_load_n((p), __ATOMIC_SEQ_CST)
#define __hip_atomic_store(p, v, order, scope) __atomic_store_n((p), (v), __ATOMIC_SEQ_CST)
#define __hip_atomic_fetch_add(p, v, order, scope) __atomic_fetch_add((p), (v), __ATOMIC_SEQ_CST)
namespace emu {
void lane_yield();
}
#define __builtin_amdgcn_s_sleep(n) (emu::lane_yield())
// wave issue priority: no scheduling effect on the CPU
#define __builtin_amdgcn_s_setprio(n) ((void)0)
template <class T>
inline T atomicExch(T* p, T v) { return __atomic_exchange_n(p, v, __ATOMIC_SEQ_CST); }
template <class T>
inline T atomicMin(T* p, T v) { return __atomic_fetch_min(p, v, __ATOMIC_SEQ_CST); }
