// Host driver that runs the device engine (flock_amd/csrc/engine.h) on the CPU through
// the stand-in tests/emu/hip/hip_runtime.h.  TEST INFRASTRUCTURE ONLY.
//
// One instance, one group of G lanes (one OS thread per lane).  Every buffer is its own
// exact-size heap allocation (the LDS region too), so AddressSanitizer reports any
// out-of-bounds access of the engine; ThreadSanitizer builds report data races between
// lanes.  The C entry point mirrors what k_solve_queue does for one group.
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <thread>
#include <vector>
#ifdef EMU_FIBERS
#include <ucontext.h>
#endif

#include "../../flock_amd/csrc/engine.h"

using namespace flock;

#ifdef EMU_FIBERS
// ---- cooperative fibers: one per lane, round-robin, switching only at barriers ----
namespace emu {
struct Fiber {
  ucontext_t ctx;
  std::vector<char> stack;
  Lane lane;
  const int* wait_gen = nullptr;
  int wait_seen = 0;
  bool done = false;
  std::function<void()> body;
};
static ucontext_t g_sched;
static std::vector<Fiber*> g_fibers;
static int g_cur = -1;

static void fiber_entry(unsigned lo, unsigned hi) {
  Fiber* f = (Fiber*)(((uintptr_t)hi << 32) | lo);
  f->body();
  f->done = true;
  swapcontext(&f->ctx, &g_sched);
}

void fiber_block_until(const int* gen, int seen) {
  Fiber* f = g_fibers[g_cur];
  f->wait_gen = gen;
  f->wait_seen = seen;
  swapcontext(&f->ctx, &g_sched);
}

static long long g_yields = 0;
void lane_yield() {
  if (++g_yields > 400000000LL) { std::fprintf(stderr, "emu: lanes spin without progress (livelock)\n"); std::abort(); }
  Fiber* f = g_fibers[g_cur];
  swapcontext(&f->ctx, &g_sched);  // stays runnable
}

static void run_fibers(std::vector<Fiber*>& fs) {
  g_yields = 0;
  g_fibers = fs;
  for (;;) {
    bool any = false, progressed = false;
    for (size_t i = 0; i < fs.size(); ++i) {
      Fiber* f = fs[i];
      if (f->done) continue;
      any = true;
      if (f->wait_gen && *f->wait_gen == f->wait_seen) continue;  // still blocked
      f->wait_gen = nullptr;
      g_cur = (int)i;
      lane_ctx = f->lane;
      swapcontext(&g_sched, &f->ctx);
      progressed = true;
    }
    if (!any) break;
    if (!progressed) { std::fprintf(stderr, "emu: all lanes blocked (deadlock)\n"); std::abort(); }
  }
}
}  // namespace emu
#else
namespace emu {
void lane_yield() { std::this_thread::yield(); }
}  // namespace emu
#endif

namespace {

#ifdef EMU_FIBERS
template <int G>
void run_group(DevInstance& D, char* region, size_t region_bytes, const EngineParams& P) {
  emu::TagBarrier block(G);
  constexpr int W = G / 64;
  std::vector<emu::TagBarrier*> waves;
  std::vector<std::vector<uint64_t>> xchg(W, std::vector<uint64_t>(64));
  for (int w = 0; w < W; ++w) waves.push_back(new emu::TagBarrier(64));
  std::vector<emu::Fiber*> fs;
  for (int t = 0; t < G; ++t) {
    emu::Fiber* f = new emu::Fiber();
    f->stack.resize(1 << 18);
    f->lane = emu::Lane{(unsigned)t, &block, waves[t / 64], xchg[t / 64].data()};
    f->body = [&D, region, region_bytes, &P] {
      if (P.pricing == 2 || P.pricing == 4) solve_instance<G, true>(D, region, region_bytes, P);
      else solve_instance<G, false>(D, region, region_bytes, P);
    };
    getcontext(&f->ctx);
    f->ctx.uc_stack.ss_sp = f->stack.data();
    f->ctx.uc_stack.ss_size = f->stack.size();
    f->ctx.uc_link = nullptr;
    const uintptr_t p = (uintptr_t)f;
    makecontext(&f->ctx, (void (*)())emu::fiber_entry, 2, (unsigned)(p & 0xffffffffu), (unsigned)(p >> 32));
    fs.push_back(f);
  }
  emu::run_fibers(fs);
  for (auto* f : fs) delete f;
  for (auto* b : waves) delete b;
}
#else
template <int G>
void run_group(DevInstance& D, char* region, size_t region_bytes, const EngineParams& P) {
  emu::TagBarrier block(G);
  constexpr int W = G / 64;
  std::vector<emu::TagBarrier*> waves;
  std::vector<std::vector<uint64_t>> xchg(W, std::vector<uint64_t>(64));
  for (int w = 0; w < W; ++w) waves.push_back(new emu::TagBarrier(64));
  std::vector<std::thread> th;
  for (int t = 0; t < G; ++t) {
    th.emplace_back([&, t] {
      emu::lane_ctx = emu::Lane{(unsigned)t, &block, waves[t / 64], xchg[t / 64].data()};
      if (P.pricing == 2 || P.pricing == 4) solve_instance<G, true>(D, region, region_bytes, P);
      else solve_instance<G, false>(D, region, region_bytes, P);
    });
  }
  for (auto& x : th) x.join();
  for (auto* b : waves) delete b;
}
#endif

template <class T>
T* exact(size_t count) {
  T* p = (T*)std::malloc(sizeof(T) * (count ? count : 1));
  std::memset(p, 0xA5, sizeof(T) * (count ? count : 1));  // poison: reads of unset data show up
  return p;
}

// One launch of the persistent work-queue kernel: `wgs` workgroups of T lanes, all of
// them running at the same time (fibers of one scheduler, or one OS thread per lane),
// each executing the kernel body queue_worker<T, G, multi> with its own LDS block.
template <int T, int G>
void run_launch(DevInstance* insts, const int32_t* order, int32_t count, int32_t* head, const EngineParams& P,
                int wgs, size_t lds_bytes) {
  struct Wg {
    emu::TagBarrier* block;
    std::vector<emu::TagBarrier*> waves;
    std::vector<std::vector<uint64_t>> xchg;
    char* smem;
  };
  constexpr int W = T / 64;
  std::vector<Wg> g(wgs);
  for (auto& w : g) {
    w.block = new emu::TagBarrier(T);
    for (int k = 0; k < W; ++k) w.waves.push_back(new emu::TagBarrier(64));
    w.xchg.assign(W, std::vector<uint64_t>(64));
    w.smem = exact<char>(lds_bytes);
  }
#ifdef EMU_FIBERS
  std::vector<emu::Fiber*> fs;
  for (int b = 0; b < wgs; ++b)
    for (int t = 0; t < T; ++t) {
      emu::Fiber* f = new emu::Fiber();
      f->stack.resize(1 << 18);
      f->lane = emu::Lane{(unsigned)t, g[b].block, g[b].waves[t / 64], g[b].xchg[t / 64].data()};
      char* smem = g[b].smem;
      f->body = [=, &P] {
        if (P.pricing == 2 || P.pricing == 4) queue_worker<T, G, true>(insts, order, count, head, P, smem);
        else queue_worker<T, G, false>(insts, order, count, head, P, smem);
      };
      getcontext(&f->ctx);
      f->ctx.uc_stack.ss_sp = f->stack.data();
      f->ctx.uc_stack.ss_size = f->stack.size();
      f->ctx.uc_link = nullptr;
      const uintptr_t p = (uintptr_t)f;
      makecontext(&f->ctx, (void (*)())emu::fiber_entry, 2, (unsigned)(p & 0xffffffffu), (unsigned)(p >> 32));
      fs.push_back(f);
    }
  emu::run_fibers(fs);
  for (auto* f : fs) delete f;
#else
  std::vector<std::thread> th;
  for (int b = 0; b < wgs; ++b)
    for (int t = 0; t < T; ++t)
      th.emplace_back([&, b, t] {
        emu::lane_ctx = emu::Lane{(unsigned)t, g[b].block, g[b].waves[t / 64], g[b].xchg[t / 64].data()};
        if (P.pricing == 2 || P.pricing == 4) queue_worker<T, G, true>(insts, order, count, head, P, g[b].smem);
        else queue_worker<T, G, false>(insts, order, count, head, P, g[b].smem);
      });
  for (auto& x : th) x.join();
#endif
  for (auto& w : g) {
    delete w.block;
    for (auto* b : w.waves) delete b;
    std::free(w.smem);
  }
}

// One group of G lanes running `body` (fibers of one scheduler, or one OS thread per lane).
template <int G>
void run_lanes(const std::function<void()>& body) {
  emu::TagBarrier block(G);
  constexpr int W = G / 64;
  std::vector<emu::TagBarrier*> waves;
  std::vector<std::vector<uint64_t>> xchg(W, std::vector<uint64_t>(64));
  for (int w = 0; w < W; ++w) waves.push_back(new emu::TagBarrier(64));
#ifdef EMU_FIBERS
  std::vector<emu::Fiber*> fs;
  for (int t = 0; t < G; ++t) {
    emu::Fiber* f = new emu::Fiber();
    f->stack.resize(1 << 18);
    f->lane = emu::Lane{(unsigned)t, &block, waves[t / 64], xchg[t / 64].data()};
    f->body = body;
    getcontext(&f->ctx);
    f->ctx.uc_stack.ss_sp = f->stack.data();
    f->ctx.uc_stack.ss_size = f->stack.size();
    f->ctx.uc_link = nullptr;
    const uintptr_t p = (uintptr_t)f;
    makecontext(&f->ctx, (void (*)())emu::fiber_entry, 2, (unsigned)(p & 0xffffffffu), (unsigned)(p >> 32));
    fs.push_back(f);
  }
  emu::run_fibers(fs);
  for (auto* f : fs) delete f;
#else
  std::vector<std::thread> th;
  for (int t = 0; t < G; ++t)
    th.emplace_back([&, t] {
      emu::lane_ctx = emu::Lane{(unsigned)t, &block, waves[t / 64], xchg[t / 64].data()};
      body();
    });
  for (auto& x : th) x.join();
#endif
  for (auto* b : waves) delete b;
}

}  // namespace

// A whole launch class: `count` instances (host arrays, row-major, back to back)
// solved by `wgs` workgroups of T lanes, G per instance, with a per-group LDS region of
// `region_bytes` bytes (as the runtime sizes it: the class maximum).  All workgroups
// run together and pop instances from one queue head.  Results per
// instance: objective, iterations, status.
extern "C" int emu_queue(int count, const int32_t* ms, const int32_t* ns, const int32_t* const* costs,
                         const int32_t* const* supplies, const int32_t* const* demands, int T, int G, int wgs,
                         int pricing, int64_t region_bytes, int64_t* objective, int64_t* iterations,
                         int32_t* status) {
  std::vector<DevInstance> D(count);
  std::vector<void*> bufs;
  for (int k = 0; k < count; ++k) {
    const int m = ms[k], n = ns[k], V = m + n;
    const int ld = (n + 3) & ~3, ldT = (m + 3) & ~3;
    int32_t* c = exact<int32_t>((size_t)m * ld);
    for (int i = 0; i < m; ++i)
      for (int j = 0; j < ld; ++j) c[(size_t)i * ld + j] = j < n ? costs[k][(size_t)i * n + j] : 0;
    int32_t* s = exact<int32_t>(m);
    int32_t* d = exact<int32_t>(n);
    std::memcpy(s, supplies[k], 4 * (size_t)m);
    std::memcpy(d, demands[k], 4 * (size_t)n);
    DevInstance& X = D[k];
    std::memset(&X, 0, sizeof X);
    X.cost = c; X.costT = exact<int32_t>((size_t)n * ldT); X.supply = s; X.demand = d;
    X.m = m; X.n = n; X.ld = ld; X.ldT = ldT;
    X.basis_row = exact<int32_t>(V - 1); X.basis_col = exact<int32_t>(V - 1); X.basis_flow = exact<int32_t>(V - 1);
    X.dual = exact<int32_t>(V); X.scratch = exact<int32_t>(aux_ints(m, n)); X.state = nullptr;
    for (void* p : {(void*)c, (void*)X.costT, (void*)s, (void*)d, (void*)X.basis_row, (void*)X.basis_col,
                    (void*)X.basis_flow, (void*)X.dual, (void*)X.scratch})
      bufs.push_back(p);
  }
  std::vector<int32_t> order(count);
  for (int k = 0; k < count; ++k) order[k] = k;
  int32_t head = 0;
  EngineParams P;
  std::memset(&P, 0, sizeof P);
  P.pricing = pricing; P.minor_limit = 64; P.lds_bytes = (int32_t)region_bytes;
  const size_t lds = (size_t)region_bytes * (size_t)(T / G);
  if (T == 256 && G == 64) run_launch<256, 64>(D.data(), order.data(), count, &head, P, wgs, lds);
  else if (T == 128 && G == 64) run_launch<128, 64>(D.data(), order.data(), count, &head, P, wgs, lds);
  else if (T == 256 && G == 256) run_launch<256, 256>(D.data(), order.data(), count, &head, P, wgs, lds);
  else if (T == 1024 && G == 1024) run_launch<1024, 1024>(D.data(), order.data(), count, &head, P, wgs, lds);
  else return -1;
  for (int k = 0; k < count; ++k) { objective[k] = D[k].objective; iterations[k] = D[k].iterations; status[k] = D[k].status; }
  for (void* p : bufs) std::free(p);
  return 0;
}

extern "C" int emu_solve(int m, int n, const int32_t* cost, const int32_t* supply, const int32_t* demand, int group,
                         int groups_per_wg, int pricing, int initial, int64_t max_it, int step_pivots,
                         int32_t* brow_out, int32_t* bcol_out, int32_t* bflow_out, int32_t* dual_out,
                         int64_t* objective, int64_t* iterations, int32_t* status, int32_t* steps_out) {
  const int V = m + n;
  const int ld = (n + 3) & ~3, ldT = (m + 3) & ~3;
  int32_t* c = exact<int32_t>((size_t)m * ld);
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < ld; ++j) c[(size_t)i * ld + j] = j < n ? cost[(size_t)i * n + j] : 0;
  int32_t* cT = exact<int32_t>((size_t)n * ldT);
  int32_t* s = exact<int32_t>(m);
  int32_t* d = exact<int32_t>(n);
  std::memcpy(s, supply, 4 * (size_t)m);
  std::memcpy(d, demand, 4 * (size_t)n);
  int32_t* br = exact<int32_t>(V - 1);
  int32_t* bc = exact<int32_t>(V - 1);
  int32_t* bf = exact<int32_t>(V - 1);
  int32_t* du = exact<int32_t>(V);
  int32_t* scr = exact<int32_t>(aux_ints(m, n));
  int32_t* st = exact<int32_t>(7 * (size_t)V);
  DevInstance D;
  std::memset(&D, 0, sizeof D);
  D.cost = c; D.costT = cT; D.supply = s; D.demand = d; D.m = m; D.n = n; D.ld = ld; D.ldT = ldT;
  D.basis_row = br; D.basis_col = bc; D.basis_flow = bf; D.dual = du; D.scratch = scr; D.state = st;
  EngineParams P;
  std::memset(&P, 0, sizeof P);
  // step_pivots < 0: time slices of -step_pivots clock ticks instead of a pivot cap
  const uint64_t slice_ticks = step_pivots < 0 ? (uint64_t)(-(int64_t)step_pivots) : 0;
  P.max_iterations = max_it; P.initial = initial; P.pricing = pricing; P.step_pivots = step_pivots > 0 ? step_pivots : 0;
  P.minor_limit = 64;
  bool aux_in_lds = false;
  const size_t cap = (size_t)160 * 1024 / (size_t)groups_per_wg;
  const size_t region_bytes = align16(group_lds_bytes(m, n, cap, &aux_in_lds));
  P.lds_bytes = (int32_t)region_bytes;
  char* region = exact<char>(region_bytes);
  int steps = 0;
  do {
    if (slice_ticks) P.deadline = wall_clock64() + slice_ticks;
    if (group == 64) run_group<64>(D, region, region_bytes, P);
    else if (group == 256) run_group<256>(D, region, region_bytes, P);
    else return -1;
    ++steps;
  } while (D.phase != 2 && steps < 100000);
  std::memcpy(brow_out, br, 4 * (size_t)(V - 1));
  std::memcpy(bcol_out, bc, 4 * (size_t)(V - 1));
  std::memcpy(bflow_out, bf, 4 * (size_t)(V - 1));
  std::memcpy(dual_out, du, 4 * (size_t)V);
  *objective = D.objective; *iterations = D.iterations; *status = D.status; *steps_out = steps;
  for (void* p : {(void*)c, (void*)cT, (void*)s, (void*)d, (void*)br, (void*)bc, (void*)bf, (void*)du, (void*)scr,
                  (void*)st, (void*)region})
    std::free(p);
  return aux_in_lds ? 1 : 0;
}

extern "C" int64_t emu_group_bytes(int m, int n, int64_t cap) {
  bool in_lds = false;
  return (int64_t)align16(group_lds_bytes(m, n, (size_t)cap, &in_lds));
}

// The large-instance path as the runtime launches it: setup, then rounds of (sweep tile
// kernels, one per tile of `vpt` vectors, then one pivot round) until the instance is
// done.  Setup and rounds run on one group of 256 lanes here.  tree_global: 0 the tree
// in LDS (saved to the state buffer between rounds), 1 the tree in the state buffer for
// the whole solve with Vogel's line state in LDS, 2 both in global memory (the LDS region
// holds only the group's scalars).  Returns the rounds run.
extern "C" int emu_large(int m, int n, const int32_t* cost, const int32_t* supply, const int32_t* demand, int pricing,
                         int vpt_override, int tree_global, int32_t* brow_out, int32_t* bcol_out, int32_t* bflow_out,
                         int32_t* dual_out, int64_t* objective, int64_t* iterations, int32_t* status) {
  const bool setup_only = pricing < 0;  // stop after the set-up kernel: the basis out is Vogel's
  if (setup_only) pricing = 0;
  const int V = m + n;
  const int ld = (n + 3) & ~3, ldT = (m + 3) & ~3;
  int32_t* c = exact<int32_t>((size_t)m * ld);
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < ld; ++j) c[(size_t)i * ld + j] = j < n ? cost[(size_t)i * n + j] : 0;
  DevInstance D;
  std::memset(&D, 0, sizeof D);
  D.cost = c; D.costT = exact<int32_t>((size_t)n * ldT);
  int32_t* s = exact<int32_t>(m); std::memcpy(s, supply, 4 * (size_t)m);
  int32_t* d = exact<int32_t>(n); std::memcpy(d, demand, 4 * (size_t)n);
  D.supply = s; D.demand = d; D.m = m; D.n = n; D.ld = ld; D.ldT = ldT;
  D.basis_row = exact<int32_t>(V - 1); D.basis_col = exact<int32_t>(V - 1); D.basis_flow = exact<int32_t>(V - 1);
  D.dual = exact<int32_t>(V); D.scratch = exact<int32_t>(aux_ints(m, n)); D.state = exact<int32_t>(7 * (size_t)V);
  EngineParams P;
  std::memset(&P, 0, sizeof P);
  P.pricing = pricing; P.minor_limit = 64;
  bool in_lds = false;
  size_t region_bytes = align16(group_lds_bytes(m, n, (size_t)160 * 1024, &in_lds));
  if (tree_global) {
    P.tree_global = 1;
    region_bytes = align16(shared_bytes() + (tree_global == 1 ? vogel_lds_bytes(m, n) : 0));
  }
  P.lds_bytes = (int32_t)region_bytes;
  char* region = exact<char>(region_bytes);
  const uint32_t nvec = (uint32_t)m * ((uint32_t)ld >> 2);
  const uint32_t vpt = vpt_override > 0 ? (uint32_t)vpt_override : std::max<uint32_t>(1024, (nvec + 255) / 256);
  const int nblocks = (int)((nvec + vpt - 1) / vpt);
  uint64_t* keys = exact<uint64_t>(nblocks);
  int32_t* cands = exact<int32_t>((size_t)3 * kTileCands * nblocks);
  Shared* S = exact<Shared>(1);
  // the grid kernels before setup: transpose (done here on the host) and the sorted lines
  int32_t* cT = const_cast<int32_t*>((const int32_t*)D.costT);
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < n; ++j) cT[(size_t)j * ldT + i] = c[(size_t)i * ld + j];
  int16_t* books = exact<int16_t>(2 * (size_t)m * n);
  int npow = 1;
  while (npow < std::max(m, n)) npow <<= 1;
  uint32_t* bkeys = exact<uint32_t>(npow);
  Ctx Cx;
  Cx.cost = gl((const int32_t*)c); Cx.costT = gl(cT); Cx.m = m; Cx.n = n; Cx.V = V; Cx.ld = ld; Cx.ldT = ldT;
  for (int L = 0; L < V; ++L) run_lanes<256>([&, L] { line_book<256>(Cx, L, bkeys, gl(books)); });
  run_lanes<256>([&] { large_setup<256>(D, region, region_bytes, P, gl((const int16_t*)books)); });
  int rounds = 0;
  while (!setup_only && D.phase == 1 && rounds < 1000000) {
    const SweepArgs A = sweep_args(D, V, keys, cands, vpt, &D.phase);
    for (int b = 0; b < nblocks; ++b) run_lanes<256>([&, b] { sweep_tile<256>(A, (uint32_t)b, *S); });
    run_lanes<256>([&] {
      if (pricing == 2 || pricing == 4) large_round<256, true>(D, region, region_bytes, P, keys, cands, nblocks);
      else large_round<256, false>(D, region, region_bytes, P, keys, cands, nblocks);
    });
    ++rounds;
  }
  std::memcpy(brow_out, D.basis_row, 4 * (size_t)(V - 1));
  std::memcpy(bcol_out, D.basis_col, 4 * (size_t)(V - 1));
  std::memcpy(bflow_out, D.basis_flow, 4 * (size_t)(V - 1));
  std::memcpy(dual_out, D.dual, 4 * (size_t)V);
  *objective = D.objective; *iterations = D.iterations; *status = D.status;
  for (void* p : {(void*)c, (void*)D.costT, (void*)s, (void*)d, (void*)D.basis_row, (void*)D.basis_col,
                  (void*)D.basis_flow, (void*)D.dual, (void*)D.scratch, (void*)D.state, (void*)region, (void*)keys,
                  (void*)cands, (void*)S, (void*)books, (void*)bkeys})
    std::free(p);
  return rounds;
}
