// server.hip -- continuous-batching solver service (flock_server_* in flock_amd.h).
//
// Requests are copied in at submit time and queued.  The device works in time slices:
//   1. admission: queued requests take a free slot while their device footprint fits
//      the HBM budget (one arena of exactly the budget, first-fit allocation).  Their
//      data goes through a pinned staging buffer on a copy stream, so admission overlaps
//      with the slice that is running;
//   2. a slice is one persistent work-queue launch over every admitted, unfinished
//      request (oldest first).  No request is started after the slice's deadline, and
//      no pivot is begun after it once no request of the slice is still in its Vogel or
//      tree build (those cannot stop, and the launch lasts until they end, so the other
//      requests keep pivoting meanwhile); an unfinished request saves its basis tree to
//      its slot and resumes from it in the next slice, so costs, trees and duals stay
//      resident;
//   3. after a slice, finished requests are gathered into one contiguous device buffer
//      (one small kernel), copied back in one transfer, and their slots and memory freed.
// The next slice is launched before poll returns, so the device keeps working while the
// caller handles results and submits new requests; a poll that finds the slice still
// running admits queued requests at once (their copies overlap the slice) and returns
// without waiting when max_wait_ms is 0.  A handle is driven by one host thread at a
// time.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <map>
#include <vector>

#include "../../include/flock_amd.h"
#include "engine.h"
#include "runtime_internal.h"

using flock::DevInstance;

// Copies the basis (row, col, flow per slot) and duals of finished slots into one
// contiguous buffer: block k handles slots[list[k]] and writes at out + offs[k].
__global__ void k_gather(const DevInstance* slots, const int32_t* list, const int64_t* offs, int32_t count,
                         int32_t* out) {
  const int k = blockIdx.x;
  if (k >= count) return;
  const DevInstance& D = slots[list[k]];
  const int V = D.m + D.n;
  int32_t* o = out + offs[k];
  for (int i = threadIdx.x; i < V - 1; i += blockDim.x) {
    o[i] = D.basis_row[i];
    o[(V - 1) + i] = D.basis_col[i];
    o[2 * (V - 1) + i] = D.basis_flow[i];
  }
  for (int i = threadIdx.x; i < V; i += blockDim.x) o[3 * (V - 1) + i] = D.dual[i];
}

namespace flock {

namespace {

double now_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

size_t al(size_t x) { return (x + 255) & ~(size_t)255; }

struct Request {
  int64_t ticket;
  int m, n;
  std::vector<int32_t> cost, supply, demand;
};

struct Slot {
  bool used = false;
  bool fresh = false;   // admitted while a slice was running: joins the next slice
  int64_t ticket = -1;
  int m = 0, n = 0, ld = 0, ldT = 0;
  size_t off = 0, bytes = 0;
  size_t cost, costT, supply, demand, brow, bcol, bflow, dual, scratch, state;
};

struct Done {
  int64_t ticket;
  int m, n;
  int64_t objective, iterations, sweeps;
  int32_t status;
  double solve_ms;
  double phase_us[6];
  std::vector<int32_t> brow, bcol, bflow, dual;
};

size_t request_bytes(int m, int n, int* ld, int* ldT) {
  const int V = m + n;
  *ld = (n + 3) & ~3;
  *ldT = (m + 3) & ~3;
  return al(4 * (size_t)m * *ld) + al(4 * (size_t)n * *ldT) + al(4 * (size_t)m) + al(4 * (size_t)n) +
         4 * al(4 * (size_t)V) + al(4 * aux_ints(m, n)) + al(4 * 7 * (size_t)V);
}

constexpr size_t kStagingBytes = (size_t)256 << 20;
constexpr int kDefaultSliceUs = 4000;

}  // namespace

struct Server {
  flock_server_config cfg;
  flock_options opt;
  EngineParams P;
  void* arena = nullptr;
  std::map<size_t, size_t> free_list;  // offset -> size
  DevInstance* d_slots = nullptr;
  int32_t* d_order[2] = {nullptr, nullptr};
  int32_t* d_heads = nullptr;
  uint64_t* d_clock0 = nullptr;
  std::vector<DevInstance> h_slots;
  std::vector<Slot> slots;
  std::deque<Request> queue;
  std::deque<Done> done;
  std::vector<int32_t> order_h[2];   // sources of the order uploads (kept alive)
  int cur = 0;                       // order buffer of the running slice
  hipStream_t compute = nullptr, copy = nullptr;
  hipEvent_t ev_start = nullptr, ev_end = nullptr, ev_copied = nullptr;
  bool running = false, copies_pending = false;
  double slice_ms = 0;
  char* staging = nullptr;           // pinned
  size_t staging_used = 0;
  int32_t* d_gather = nullptr;       // finished results, gathered
  size_t gather_cap = 0;             // int32 entries
  int32_t* h_gather = nullptr;       // pinned
  size_t h_gather_cap = 0;
  int32_t* d_glist = nullptr;
  int64_t* d_goffs = nullptr;
  std::vector<int32_t> glist_h;
  std::vector<int64_t> goffs_h;
  int64_t next_ticket = 0;
  int32_t in_flight = 0;
  int64_t slices = 0;

  bool arena_alloc(size_t bytes, size_t* off) {
    for (auto it = free_list.begin(); it != free_list.end(); ++it) {
      if (it->second < bytes) continue;
      *off = it->first;
      const size_t rest = it->second - bytes;
      const size_t roff = it->first + bytes;
      free_list.erase(it);
      if (rest) free_list[roff] = rest;
      return true;
    }
    return false;
  }
  void arena_free(size_t off, size_t bytes) {
    auto it = free_list.emplace(off, bytes).first;
    auto nx = std::next(it);
    if (nx != free_list.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_list.erase(nx);
    }
    if (it != free_list.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_list.erase(it);
      }
    }
  }

  // Stage `bytes` of host data in pinned memory and queue its copy to device `dst`.
  bool stage_copy(void* dst, const void* src, size_t bytes) {
    std::memcpy(staging + staging_used, src, bytes);
    const bool ok = hipMemcpyAsync(dst, staging + staging_used, bytes, hipMemcpyHostToDevice, copy) == hipSuccess;
    staging_used += (bytes + 255) & ~(size_t)255;
    return ok;
  }

  // Admit queued requests (first come first served) while a slot, arena memory and
  // staging space are free.  Copies run on the copy stream; the next slice waits for them.
  int admit() {
    if (queue.empty()) return FLOCK_OK;
    if (copies_pending) {  // the staging buffer is reused only after its copies finished
      if (hipStreamSynchronize(copy) != hipSuccess) return set_error(FLOCK_E_DEVICE, "copy stream sync failed");
      copies_pending = false;
    }
    staging_used = 0;
    char* b = (char*)arena;
    int admitted = 0;
    while (!queue.empty()) {
      int s = -1;
      for (int k = 0; k < (int)slots.size(); ++k)
        if (!slots[k].used) { s = k; break; }
      if (s < 0) break;
      Request& r = queue.front();
      const size_t stage_need = al(4 * (size_t)r.m * r.n) + al(4 * (size_t)r.m) + al(4 * (size_t)r.n) +
                                al(sizeof(DevInstance));
      if (staging_used + stage_need > kStagingBytes) break;  // next round
      int ld, ldT;
      const size_t bytes = request_bytes(r.m, r.n, &ld, &ldT);
      size_t off;
      if (!arena_alloc(bytes, &off)) break;  // first come first served: wait for memory
      Slot& sl = slots[s];
      sl.used = true; sl.fresh = running; sl.ticket = r.ticket; sl.m = r.m; sl.n = r.n; sl.ld = ld; sl.ldT = ldT;
      sl.off = off; sl.bytes = bytes;
      const int V = r.m + r.n;
      size_t o = off;
      sl.cost = o; o += al(4 * (size_t)r.m * ld);
      sl.costT = o; o += al(4 * (size_t)r.n * ldT);
      sl.supply = o; o += al(4 * (size_t)r.m);
      sl.demand = o; o += al(4 * (size_t)r.n);
      sl.brow = o; o += al(4 * (size_t)V);
      sl.bcol = o; o += al(4 * (size_t)V);
      sl.bflow = o; o += al(4 * (size_t)V);
      sl.dual = o; o += al(4 * (size_t)V);
      sl.scratch = o; o += al(4 * aux_ints(r.m, r.n));
      sl.state = o;
      // the cost rows are staged packed and spread to the padded layout by the copy
      std::memcpy(staging + staging_used, r.cost.data(), 4 * (size_t)r.m * r.n);
      if (hipMemcpy2DAsync(b + sl.cost, 4 * (size_t)ld, staging + staging_used, 4 * (size_t)r.n, 4 * (size_t)r.n,
                           r.m, hipMemcpyHostToDevice, copy) != hipSuccess)
        return set_error(FLOCK_E_DEVICE, "server upload failed");
      staging_used += al(4 * (size_t)r.m * r.n);
      if (!stage_copy(b + sl.supply, r.supply.data(), 4 * (size_t)r.m) ||
          !stage_copy(b + sl.demand, r.demand.data(), 4 * (size_t)r.n))
        return set_error(FLOCK_E_DEVICE, "server upload failed");
      DevInstance& D = h_slots[s];
      std::memset(&D, 0, sizeof D);
      D.cost = (const int32_t*)(b + sl.cost);
      D.costT = (int32_t*)(b + sl.costT);
      D.supply = (const int32_t*)(b + sl.supply);
      D.demand = (const int32_t*)(b + sl.demand);
      D.m = r.m; D.n = r.n; D.ld = ld; D.ldT = ldT;
      D.basis_row = (int32_t*)(b + sl.brow);
      D.basis_col = (int32_t*)(b + sl.bcol);
      D.basis_flow = (int32_t*)(b + sl.bflow);
      D.dual = (int32_t*)(b + sl.dual);
      D.scratch = (int32_t*)(b + sl.scratch);
      D.state = (int32_t*)(b + sl.state);
      D.phase = 0;
      if (!stage_copy(d_slots + s, &D, sizeof D)) return set_error(FLOCK_E_DEVICE, "server descriptor upload failed");
      queue.pop_front();
      ++in_flight;
      ++admitted;
    }
    if (admitted) {
      if (hipEventRecord(ev_copied, copy) != hipSuccess) return set_error(FLOCK_E_DEVICE, "event record failed");
      copies_pending = true;
    }
    return FLOCK_OK;
  }

  // Launch one slice over every admitted, unfinished slot (oldest ticket first).
  int launch_slice() {
    const int nb = cur ^ 1;
    std::vector<int32_t>& order = order_h[nb];
    order.clear();
    int maxm = 1, maxn = 1;
    size_t lds = 0;
    for (int s = 0; s < (int)slots.size(); ++s) {
      if (!slots[s].used) continue;
      slots[s].fresh = false;
      order.push_back(s);
    }
    if (order.empty()) return FLOCK_OK;
    std::sort(order.begin(), order.end(), [&](int a, int b) { return slots[a].ticket < slots[b].ticket; });
    int bigV = 0;
    for (int s : order)
      if (slots[s].m + slots[s].n > bigV) { bigV = slots[s].m + slots[s].n; maxm = slots[s].m; maxn = slots[s].n; }
    // one engine configuration per slice, chosen for its largest request
    const int cfg = pick_cfg(maxm, maxn, &opt, (int64_t)order.size());
    for (int s : order) {
      bool dummy;
      lds = std::max(lds, cfg_group_bytes(cfg, slots[s].m, slots[s].n, &dummy));
    }
    if (copies_pending && hipStreamWaitEvent(compute, ev_copied, 0) != hipSuccess)
      return set_error(FLOCK_E_DEVICE, "stream wait failed");
    if (hipMemcpyAsync(d_order[nb], order.data(), 4 * order.size(), hipMemcpyHostToDevice, compute) != hipSuccess)
      return set_error(FLOCK_E_DEVICE, "order upload failed");
    // heads[0]: the slice's queue head; heads[1]: its count of groups in Vogel / tree build
    int st = launch_reset(d_slots, 0, d_heads, 2, compute, d_clock0);
    if (st) return st;
    if (hipEventRecord(ev_start, compute) != hipSuccess) return set_error(FLOCK_E_DEVICE, "event record failed");
    EngineParams Q = P;
    Q.slice_ticks = (uint64_t)((cfg_slice_us() > 0 ? cfg_slice_us() : kDefaultSliceUs) * lib_ticks_per_us());
    Q.clock0 = d_clock0;
    Q.busy = d_heads + 1;
    st = launch_queue_on(cfg, d_slots, d_order[nb], (int32_t)order.size(), d_heads, Q, lds, 0, compute);
    if (st) return st;
    if (hipEventRecord(ev_end, compute) != hipSuccess) return set_error(FLOCK_E_DEVICE, "event record failed");
    cur = nb;
    running = true;
    ++slices;
    return FLOCK_OK;
  }
  int cfg_slice_us() const { return cfg.slice_us; }

  // Wait for the running slice and move its finished requests to `done`.
  int finish_slice() {
    if (!running) return FLOCK_OK;
    if (hipEventSynchronize(ev_end) != hipSuccess) return set_error(FLOCK_E_DEVICE, "slice failed");
    running = false;
    float ms = 0;
    hipEventElapsedTime(&ms, ev_start, ev_end);
    slice_ms = ms;
    int hi = 0;
    for (int s = 0; s < (int)slots.size(); ++s)
      if (slots[s].used && !slots[s].fresh) hi = s + 1;
    if (hi == 0) return FLOCK_OK;
    if (hipMemcpyAsync(h_slots.data(), d_slots, sizeof(DevInstance) * (size_t)hi, hipMemcpyDeviceToHost, compute) !=
            hipSuccess ||
        hipStreamSynchronize(compute) != hipSuccess)
      return set_error(FLOCK_E_DEVICE, "descriptor download failed");
    glist_h.clear();
    goffs_h.clear();
    int64_t total = 0;
    for (int s = 0; s < hi; ++s) {
      const Slot& sl = slots[s];
      if (!sl.used || sl.fresh || h_slots[s].phase != 2) continue;
      glist_h.push_back(s);
      goffs_h.push_back(total);
      total += 4 * (int64_t)(sl.m + sl.n) - 3;
    }
    if (glist_h.empty()) return FLOCK_OK;
    int st = ensure_gather((size_t)total, glist_h.size());
    if (st) return st;
    const int32_t cnt = (int32_t)glist_h.size();
    if (hipMemcpyAsync(d_glist, glist_h.data(), 4 * glist_h.size(), hipMemcpyHostToDevice, compute) != hipSuccess ||
        hipMemcpyAsync(d_goffs, goffs_h.data(), 8 * goffs_h.size(), hipMemcpyHostToDevice, compute) != hipSuccess)
      return set_error(FLOCK_E_DEVICE, "gather list upload failed");
    hipLaunchKernelGGL(k_gather, dim3(cnt), dim3(256), 0, compute, (const DevInstance*)d_slots, (const int32_t*)d_glist,
                       (const int64_t*)d_goffs, cnt, d_gather);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(h_gather, d_gather, 4 * (size_t)total, hipMemcpyDeviceToHost, compute) != hipSuccess ||
        hipStreamSynchronize(compute) != hipSuccess)
      return set_error(FLOCK_E_DEVICE, "result gather failed");
    for (int k = 0; k < cnt; ++k) {
      const int s = glist_h[k];
      Slot& sl = slots[s];
      const DevInstance& D = h_slots[s];
      const int V = sl.m + sl.n;
      const int32_t* g = h_gather + goffs_h[k];
      Done d;
      d.ticket = sl.ticket; d.m = sl.m; d.n = sl.n;
      d.objective = D.objective; d.iterations = D.iterations; d.status = D.status; d.solve_ms = slice_ms;
      for (int q = 0; q < 6; ++q) d.phase_us[q] = (double)D.ticks[q] / lib_ticks_per_us();
      d.sweeps = D.sweeps;
      d.brow.assign(g, g + (V - 1));
      d.bcol.assign(g + (V - 1), g + 2 * (V - 1));
      d.bflow.assign(g + 2 * (V - 1), g + 3 * (V - 1));
      d.dual.assign(g + 3 * (V - 1), g + 3 * (V - 1) + V);
      done.push_back(std::move(d));
      arena_free(sl.off, sl.bytes);
      sl.used = false;
      --in_flight;
    }
    return FLOCK_OK;
  }

  int ensure_gather(size_t ints, size_t nlist) {
    if (ints > gather_cap) {
      if (d_gather) hipFree(d_gather);
      const size_t cap = std::max(ints, gather_cap * 2);
      if (hipMalloc(&d_gather, 4 * cap) != hipSuccess) { d_gather = nullptr; gather_cap = 0; return set_error(FLOCK_E_DEVICE, "gather hipMalloc failed"); }
      gather_cap = cap;
    }
    if (ints > h_gather_cap) {
      if (h_gather) hipHostFree(h_gather);
      const size_t cap = std::max(ints, h_gather_cap * 2);
      if (hipHostMalloc((void**)&h_gather, 4 * cap, hipHostMallocDefault) != hipSuccess) {
        h_gather = nullptr; h_gather_cap = 0;
        return set_error(FLOCK_E_DEVICE, "gather hipHostMalloc failed");
      }
      h_gather_cap = cap;
    }
    (void)nlist;
    return FLOCK_OK;
  }
};

}  // namespace flock

using flock::Server;

extern "C" {

int flock_server_create(const flock_server_config* cfg, const flock_options* opt, flock_server** out) {
  if (!out) return flock::set_error(FLOCK_E_INVALID, "null output");
  *out = nullptr;
  int st = flock::ensure_init(-1);
  if (st) return st;
  Server* s = new Server();
  std::memset(&s->cfg, 0, sizeof s->cfg);
  if (cfg) s->cfg = *cfg;
  if (s->cfg.hbm_budget_bytes <= 0) s->cfg.hbm_budget_bytes = (int64_t)8 << 30;
  if (s->cfg.max_slots <= 0) s->cfg.max_slots = 2048;
  if (s->cfg.slice_us < 0) s->cfg.slice_us = 0;
  flock_default_options(&s->opt);
  if (opt) s->opt = *opt;
  s->P = flock::make_params(opt);
  s->P.step_pivots = s->cfg.step_pivots > 0 ? s->cfg.step_pivots : 0;
  if (s->cfg.minor_limit > 0) s->P.minor_limit = s->cfg.minor_limit;
  const size_t budget = (size_t)s->cfg.hbm_budget_bytes;
  const int S = s->cfg.max_slots;
  bool ok = hipMalloc(&s->arena, budget) == hipSuccess;
  ok = ok && hipMalloc(&s->d_slots, sizeof(flock::DevInstance) * (size_t)S) == hipSuccess;
  ok = ok && hipMalloc(&s->d_order[0], sizeof(int32_t) * (size_t)S) == hipSuccess;
  ok = ok && hipMalloc(&s->d_order[1], sizeof(int32_t) * (size_t)S) == hipSuccess;
  ok = ok && hipMalloc(&s->d_heads, sizeof(int32_t) * flock::kMaxClasses) == hipSuccess;
  ok = ok && hipMalloc(&s->d_clock0, sizeof(uint64_t)) == hipSuccess;
  ok = ok && hipMalloc(&s->d_glist, sizeof(int32_t) * (size_t)S) == hipSuccess;
  ok = ok && hipMalloc(&s->d_goffs, sizeof(int64_t) * (size_t)S) == hipSuccess;
  ok = ok && hipHostMalloc((void**)&s->staging, flock::kStagingBytes, hipHostMallocDefault) == hipSuccess;
  ok = ok && hipStreamCreateWithFlags(&s->compute, hipStreamNonBlocking) == hipSuccess;
  ok = ok && hipStreamCreateWithFlags(&s->copy, hipStreamNonBlocking) == hipSuccess;
  ok = ok && hipEventCreate(&s->ev_start) == hipSuccess && hipEventCreate(&s->ev_end) == hipSuccess;
  ok = ok && hipEventCreateWithFlags(&s->ev_copied, hipEventDisableTiming) == hipSuccess;
  ok = ok && hipMemset(s->d_slots, 0, sizeof(flock::DevInstance) * (size_t)S) == hipSuccess;
  if (!ok) {
    flock_server_destroy(reinterpret_cast<flock_server*>(s));
    return flock::set_error(FLOCK_E_DEVICE, "server set-up failed (budget %zu bytes)", budget);
  }
  s->free_list[0] = budget;
  s->h_slots.assign(S, flock::DevInstance{});
  s->slots.assign(S, flock::Slot{});
  *out = reinterpret_cast<flock_server*>(s);
  return FLOCK_OK;
}

int flock_server_submit(flock_server* hs, const flock_instance* inst, int64_t* ticket) {
  Server* s = reinterpret_cast<Server*>(hs);
  if (!s || !ticket) return flock::set_error(FLOCK_E_INVALID, "null server/ticket");
  int st = flock::validate_instance(inst);
  if (st) return st;
  int ld, ldT;
  const size_t bytes = flock::request_bytes(inst->m, inst->n, &ld, &ldT);
  if ((int64_t)bytes > s->cfg.hbm_budget_bytes)
    return flock::set_error(FLOCK_E_BUDGET, "request needs %zu bytes, budget %lld", bytes,
                            (long long)s->cfg.hbm_budget_bytes);
  if (4 * (size_t)inst->m * inst->n + 4 * (size_t)(inst->m + inst->n) + 4096 > flock::kStagingBytes)
    return flock::set_error(FLOCK_E_BUDGET, "request exceeds the %zu-byte staging buffer", flock::kStagingBytes);
  if (s->queue.size() >= (size_t)1 << 20) return flock::set_error(FLOCK_E_BUSY, "server queue full");
  flock::Request r;
  r.ticket = s->next_ticket++;
  r.m = inst->m; r.n = inst->n;
  r.cost.assign(inst->cost, inst->cost + (size_t)inst->m * inst->n);
  r.supply.assign(inst->supply, inst->supply + inst->m);
  r.demand.assign(inst->demand, inst->demand + inst->n);
  s->queue.push_back(std::move(r));
  *ticket = s->next_ticket - 1;
  return FLOCK_OK;
}

int flock_server_poll(flock_server* hs, double max_wait_ms, int32_t cap, int64_t* tickets, flock_result* res,
                      int32_t* n_done) {
  Server* s = reinterpret_cast<Server*>(hs);
  if (!s || !n_done) return flock::set_error(FLOCK_E_INVALID, "null server/n_done");
  *n_done = 0;
  const double t0 = flock::now_ms();
  int st;
  for (;;) {
    if (s->running) {
      const hipError_t q = hipEventQuery(s->ev_end);
      if (q == hipErrorNotReady) {
        // the slice is still on the device: admit queued requests now (their copies
        // overlap it; they join the next slice), and return unless asked to wait
        if ((st = s->admit())) return st;
        if (!s->done.empty() || flock::now_ms() - t0 >= max_wait_ms) break;
      } else if (q != hipSuccess) {
        return flock::set_error(FLOCK_E_DEVICE, "slice failed: %s", hipGetErrorString(q));
      }
    }
    if ((st = s->finish_slice())) return st;  // waits for the slice when it is still running
    if ((st = s->admit())) return st;
    if ((st = s->launch_slice())) return st;  // runs while the caller handles results
    if (!s->done.empty() || !s->running || flock::now_ms() - t0 >= max_wait_ms) break;
  }
  int32_t k = 0;
  while (k < cap && !s->done.empty()) {
    flock::Done& d = s->done.front();
    if (tickets) tickets[k] = d.ticket;
    if (res) {
      flock_result& r = res[k];
      r.objective = d.objective; r.iterations = d.iterations; r.status = d.status; r.solve_ms = d.solve_ms;
      r.sweeps = d.sweeps;
      for (int q = 0; q < 6; ++q) r.phase_us[q] = d.phase_us[q];
      r.start_us = 0; r.done_us = 0;
      const int V = d.m + d.n;
      int nf = 0;
      for (int q = 0; q < V - 1; ++q) {
        if (d.bflow[q] == 0) continue;
        if (r.flows) r.flows[nf] = flock_flow{d.brow[q], d.bcol[q], d.bflow[q]};
        ++nf;
      }
      r.n_flows = nf;
      if (r.basis_row) std::memcpy(r.basis_row, d.brow.data(), 4 * (size_t)(V - 1));
      if (r.basis_col) std::memcpy(r.basis_col, d.bcol.data(), 4 * (size_t)(V - 1));
      if (r.basis_flow) std::memcpy(r.basis_flow, d.bflow.data(), 4 * (size_t)(V - 1));
      if (r.u) std::memcpy(r.u, d.dual.data(), 4 * (size_t)d.m);
      if (r.v) std::memcpy(r.v, d.dual.data() + d.m, 4 * (size_t)d.n);
    }
    s->done.pop_front();
    ++k;
  }
  *n_done = k;
  return FLOCK_OK;
}

int32_t flock_server_pending(const flock_server* hs) {
  const Server* s = reinterpret_cast<const Server*>(hs);
  if (!s) return 0;
  return (int32_t)(s->queue.size() + s->in_flight + s->done.size());
}

void flock_server_destroy(flock_server* hs) {
  Server* s = reinterpret_cast<Server*>(hs);
  if (!s) return;
  if (s->compute) { hipStreamSynchronize(s->compute); hipStreamDestroy(s->compute); }
  if (s->copy) { hipStreamSynchronize(s->copy); hipStreamDestroy(s->copy); }
  if (s->ev_start) hipEventDestroy(s->ev_start);
  if (s->ev_end) hipEventDestroy(s->ev_end);
  if (s->ev_copied) hipEventDestroy(s->ev_copied);
  if (s->arena) hipFree(s->arena);
  if (s->d_slots) hipFree(s->d_slots);
  for (int k = 0; k < 2; ++k)
    if (s->d_order[k]) hipFree(s->d_order[k]);
  if (s->d_heads) hipFree(s->d_heads);
  if (s->d_clock0) hipFree(s->d_clock0);
  if (s->d_glist) hipFree(s->d_glist);
  if (s->d_goffs) hipFree(s->d_goffs);
  if (s->d_gather) hipFree(s->d_gather);
  if (s->h_gather) hipHostFree(s->h_gather);
  if (s->staging) hipHostFree(s->staging);
  delete s;
}

}  // extern "C"
