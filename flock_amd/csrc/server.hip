// server.hip -- continuous-batching solver service (flock_server_* in flock_amd.h).
//
// Requests are copied in at submit time and queued.  Each scheduling step
//   1. admits queued requests, first come first served, while a device slot is free
//      and the request's device footprint fits the HBM budget (one device arena of
//      exactly the budget, first-fit allocation);
//   2. launches the per-instance engine over every admitted, unfinished request, one
//      persistent work-queue launch per kernel size class on its own stream, with a
//      pivot budget per request (the engine saves the basis tree to its slot's device
//      state and resumes from it on the next step);
//   3. collects finished requests and frees their slots and memory.
// A request that arrives while others are on the GPU joins at the next step, so new
// work overlaps with work already in flight; costs, trees and duals stay resident on
// the device between steps.  A handle is driven by one host thread at a time.
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

}  // namespace

struct Server {
  flock_server_config cfg;
  flock_options opt;
  EngineParams P;
  void* arena = nullptr;
  std::map<size_t, size_t> free_list;  // offset -> size
  DevInstance* d_slots = nullptr;
  int32_t* d_order = nullptr;
  int32_t* d_heads = nullptr;
  std::vector<DevInstance> h_slots;
  std::vector<Slot> slots;
  std::deque<Request> queue;
  std::deque<Done> done;
  std::vector<int32_t> step_order;
  static constexpr int kStreams = 6;
  hipStream_t streams[kStreams] = {};
  hipEvent_t ev_start = nullptr, ev_end = nullptr;
  int64_t next_ticket = 0;
  int32_t in_flight = 0;

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

  int admit() {
    char* b = (char*)arena;
    while (!queue.empty()) {
      int s = -1;
      for (int k = 0; k < (int)slots.size(); ++k)
        if (!slots[k].used) { s = k; break; }
      if (s < 0) break;
      Request& r = queue.front();
      Slot& sl = slots[s];
      int ld, ldT;
      const size_t bytes = request_bytes(r.m, r.n, &ld, &ldT);
      size_t off;
      if (!arena_alloc(bytes, &off)) break;  // first come first served: wait for memory
      sl.used = true; sl.ticket = r.ticket; sl.m = r.m; sl.n = r.n; sl.ld = ld; sl.ldT = ldT;
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
      hipStream_t st = streams[0];
      if (hipMemcpy2DAsync(b + sl.cost, 4 * (size_t)ld, r.cost.data(), 4 * (size_t)r.n, 4 * (size_t)r.n, r.m,
                           hipMemcpyHostToDevice, st) != hipSuccess ||
          hipMemcpyAsync(b + sl.supply, r.supply.data(), 4 * (size_t)r.m, hipMemcpyHostToDevice, st) != hipSuccess ||
          hipMemcpyAsync(b + sl.demand, r.demand.data(), 4 * (size_t)r.n, hipMemcpyHostToDevice, st) != hipSuccess)
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
      if (hipMemcpyAsync(d_slots + s, &D, sizeof D, hipMemcpyHostToDevice, st) != hipSuccess)
        return set_error(FLOCK_E_DEVICE, "server descriptor upload failed");
      queue.pop_front();
      ++in_flight;
    }
    // the copies above come from pageable host memory; finish them before the request
    // buffers are released and before the step reads them
    if (hipStreamSynchronize(streams[0]) != hipSuccess) return set_error(FLOCK_E_DEVICE, "server sync failed");
    return FLOCK_OK;
  }

  int step(double* step_ms) {
    // order arrays per kernel class over admitted, unfinished slots (a member: the
    // upload below is asynchronous)
    std::vector<int32_t>& order = step_order;
    order.clear();
    struct Cl { int cfg, begin, count; size_t lds; };
    std::vector<Cl> cls;
    for (int c = 0; c < num_cfgs(); ++c)
      for (int b = 0; b < kSizeBuckets; ++b) {
        Cl cl{c, (int)order.size(), 0, 0};
        for (int s = 0; s < (int)slots.size(); ++s) {
          if (!slots[s].used) continue;
          if (pick_cfg(slots[s].m, slots[s].n, &opt, in_flight) != c || class_bucket(c, slots[s].m + slots[s].n) != b) continue;
          order.push_back(s);
          bool dummy;
          cl.lds = std::max(cl.lds, cfg_group_bytes(c, slots[s].m, slots[s].n, &dummy));
        }
        cl.count = (int)order.size() - cl.begin;
        if (cl.count) cls.push_back(cl);
      }
    *step_ms = 0;
    if (order.empty()) return FLOCK_OK;
    if (hipMemcpyAsync(d_order, order.data(), 4 * order.size(), hipMemcpyHostToDevice, streams[0]) != hipSuccess)
      return set_error(FLOCK_E_DEVICE, "order upload failed");
    int st = launch_reset(d_slots, 0, d_heads, kMaxClasses, streams[0]);
    if (st) return st;
    if (hipEventRecord(ev_start, streams[0]) != hipSuccess) return set_error(FLOCK_E_DEVICE, "event record failed");
    // classes launch one after another on one stream (concurrent launches of different
    // engine configurations are not used)
    int h = 0;
    for (const Cl& cl : cls) {
      st = launch_queue_on(cl.cfg, d_slots, d_order + cl.begin, cl.count, d_heads + h, P, cl.lds, 0, streams[0]);
      if (st) return st;
      ++h;
    }
    if (hipEventRecord(ev_end, streams[0]) != hipSuccess || hipEventSynchronize(ev_end) != hipSuccess)
      return set_error(FLOCK_E_DEVICE, "event failed");
    float ms = 0;
    hipEventElapsedTime(&ms, ev_start, ev_end);
    *step_ms = ms;
    return FLOCK_OK;
  }

  int collect(double step_ms) {
    if (hipMemcpy(h_slots.data(), d_slots, sizeof(DevInstance) * h_slots.size(), hipMemcpyDeviceToHost) != hipSuccess)
      return set_error(FLOCK_E_DEVICE, "descriptor download failed");
    char* b = (char*)arena;
    for (int s = 0; s < (int)slots.size(); ++s) {
      Slot& sl = slots[s];
      if (!sl.used) continue;
      const DevInstance& D = h_slots[s];
      if (D.phase != 2) continue;
      const int V = sl.m + sl.n;
      Done d;
      d.ticket = sl.ticket; d.m = sl.m; d.n = sl.n;
      d.objective = D.objective; d.iterations = D.iterations; d.status = D.status; d.solve_ms = step_ms;
      for (int k = 0; k < 6; ++k) d.phase_us[k] = (double)D.ticks[k] / lib_ticks_per_us();
      d.sweeps = D.sweeps;
      d.brow.resize(V - 1); d.bcol.resize(V - 1); d.bflow.resize(V - 1); d.dual.resize(V);
      if (V > 1 && (hipMemcpy(d.brow.data(), b + sl.brow, 4 * (size_t)(V - 1), hipMemcpyDeviceToHost) != hipSuccess ||
                    hipMemcpy(d.bcol.data(), b + sl.bcol, 4 * (size_t)(V - 1), hipMemcpyDeviceToHost) != hipSuccess ||
                    hipMemcpy(d.bflow.data(), b + sl.bflow, 4 * (size_t)(V - 1), hipMemcpyDeviceToHost) != hipSuccess))
        return set_error(FLOCK_E_DEVICE, "result download failed");
      if (hipMemcpy(d.dual.data(), b + sl.dual, 4 * (size_t)V, hipMemcpyDeviceToHost) != hipSuccess)
        return set_error(FLOCK_E_DEVICE, "dual download failed");
      done.push_back(std::move(d));
      arena_free(sl.off, sl.bytes);
      sl.used = false;
      --in_flight;
    }
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
  if (s->cfg.step_pivots <= 0) s->cfg.step_pivots = 512;
  flock_default_options(&s->opt);
  if (opt) s->opt = *opt;
  s->P = flock::make_params(opt);
  s->P.step_pivots = s->cfg.step_pivots;
  const size_t budget = (size_t)s->cfg.hbm_budget_bytes;
  if (hipMalloc(&s->arena, budget) != hipSuccess) { delete s; return flock::set_error(FLOCK_E_DEVICE, "arena hipMalloc(%zu) failed", budget); }
  s->free_list[0] = budget;
  const int S = s->cfg.max_slots;
  if (hipMalloc(&s->d_slots, sizeof(flock::DevInstance) * (size_t)S) != hipSuccess ||
      hipMalloc(&s->d_order, sizeof(int32_t) * (size_t)S) != hipSuccess ||
      hipMalloc(&s->d_heads, sizeof(int32_t) * flock::kMaxClasses) != hipSuccess) {
    delete s;
    return flock::set_error(FLOCK_E_DEVICE, "server hipMalloc failed");
  }
  s->h_slots.assign(S, flock::DevInstance{});
  s->slots.assign(S, flock::Slot{});
  for (int k = 0; k < Server::kStreams; ++k)
    if (hipStreamCreateWithFlags(&s->streams[k], hipStreamNonBlocking) != hipSuccess) {
      delete s;
      return flock::set_error(FLOCK_E_DEVICE, "stream create failed");
    }
  if (hipEventCreate(&s->ev_start) != hipSuccess || hipEventCreate(&s->ev_end) != hipSuccess) {
    delete s;
    return flock::set_error(FLOCK_E_DEVICE, "event create failed");
  }
  if (hipMemset(s->d_slots, 0, sizeof(flock::DevInstance) * (size_t)S) != hipSuccess) {
    delete s;
    return flock::set_error(FLOCK_E_DEVICE, "memset failed");
  }
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
  while (s->done.empty()) {
    int st = s->admit();
    if (st) return st;
    if (s->in_flight == 0) break;
    double ms = 0;
    st = s->step(&ms);
    if (st) return st;
    st = s->collect(ms);
    if (st) return st;
    if (flock::now_ms() - t0 >= max_wait_ms) break;
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
  for (int k = 0; k < Server::kStreams; ++k)
    if (s->streams[k]) { hipStreamSynchronize(s->streams[k]); hipStreamDestroy(s->streams[k]); }
  if (s->ev_start) hipEventDestroy(s->ev_start);
  if (s->ev_end) hipEventDestroy(s->ev_end);
  if (s->arena) hipFree(s->arena);
  if (s->d_slots) hipFree(s->d_slots);
  if (s->d_order) hipFree(s->d_order);
  if (s->d_heads) hipFree(s->d_heads);
  delete s;
}

}  // extern "C"
