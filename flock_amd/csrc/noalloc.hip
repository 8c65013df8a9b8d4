// noalloc.hip -- flock_solve_noalloc: the plain solver without any device allocation.
//
// Every device buffer of the solve (instance descriptor, work order, queue head, the
// padded cost matrix and its transpose, supplies, demands, basis, duals, the engine's
// global scratch and the compacted flow list) is carved out of one file-scope device
// array, `scratch`, of kNoAllocBytes.  Host code never names that array directly: at
// set-up, once per device and with that device current, hipGetSymbolAddress gives its
// address on that device, and only that address is passed to copies and kernels.
// The array keeps the last call's contents; the reset kernel clears what the solve
// needs at zero (queue head, per-instance status and counters) and the descriptor is
// uploaded whole each call.  Calls on one device share the array, so they are
// serialised by a lock inside the library (calls overlapping on one device from
// several host threads of one process wait for each other).
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>

#include "../../include/flock_amd.h"
#include "engine.h"
#include "runtime_internal.h"

using flock::DevInstance;

namespace {

constexpr size_t kNoAllocBytes = (size_t)48 << 20;  // covers 2000 x 2000 and more
constexpr int kMaxDevices = 64;

static __device__ __align__(256) char scratch[kNoAllocBytes];

// Compacted flow list: every basic arc with positive flow, in basis-slot order, as
// (source, destination, qty) triples; the count goes to *count.  One workgroup.
__global__ __launch_bounds__(1024) void k_flow_list(const int32_t* brow, const int32_t* bcol, const int32_t* bflow,
                                                     int32_t slots, int32_t* out, int32_t* count) {
  __shared__ int32_t part[1024];
  const int t = threadIdx.x;
  const int per = (slots + 1023) / 1024;
  const int b = t * per, e = b + per < slots ? b + per : slots;
  int c = 0;
  for (int k = b; k < e; ++k) c += bflow[k] != 0;
  part[t] = c;
  __syncthreads();
  for (int s = 1; s < 1024; s <<= 1) {  // inclusive scan (Hillis-Steele)
    const int v = t >= s ? part[t - s] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int at = part[t] - c;
  for (int k = b; k < e; ++k) {
    if (bflow[k] == 0) continue;
    out[3 * at] = brow[k]; out[3 * at + 1] = bcol[k]; out[3 * at + 2] = bflow[k];
    ++at;
  }
  if (t == 1023) *count = part[1023];
}

struct DeviceScratch {
  char* ptr = nullptr;          // address of `scratch` on this device
  hipEvent_t e0 = nullptr, e1 = nullptr;
};
DeviceScratch g_dev[kMaxDevices];
std::mutex g_mu;

size_t al(size_t x) { return (x + 255) & ~(size_t)255; }

// Set-up for the current device: the array's device address and two timing events.
int setup_device(int dev) {
  if (dev < 0 || dev >= kMaxDevices) return flock::set_error(FLOCK_E_DEVICE, "device %d out of range", dev);
  DeviceScratch& d = g_dev[dev];
  if (d.ptr) return FLOCK_OK;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return flock::set_error(FLOCK_E_DEVICE, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
  if (hipEventCreate(&d.e0) != hipSuccess || hipEventCreate(&d.e1) != hipSuccess)
    return flock::set_error(FLOCK_E_DEVICE, "event set-up failed");
  d.ptr = (char*)p;
  return FLOCK_OK;
}

}  // namespace

namespace flock {
// Called by flock_init (and by the first flock_solve_noalloc on a device).
int noalloc_setup() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return set_error(FLOCK_E_DEVICE, "hipGetDevice failed");
  std::lock_guard<std::mutex> lk(g_mu);
  return setup_device(dev);
}
}  // namespace flock

extern "C" int flock_solve_noalloc(const flock_instance* inst, const flock_options* opt, flock_result* res) {
  if (!res) return flock::set_error(FLOCK_E_INVALID, "null result");
  res->status = FLOCK_E_INVALID;
  int st = flock::validate_instance(inst);
  if (st) { res->status = st; return st; }
  if ((st = flock::ensure_init(-1))) { res->status = st; return st; }
  const int m = inst->m, n = inst->n, V = m + n;
  const int ld = (n + 3) & ~3, ldT = (m + 3) & ~3;
  // layout inside the array (offsets in bytes, 256-byte aligned)
  size_t off = 0;
  const size_t oDesc = off; off += al(sizeof(DevInstance));
  const size_t oOrder = off; off += al(sizeof(int32_t));
  const size_t oHead = off; off += al(sizeof(int32_t) * flock::kMaxClasses);
  const size_t oT0 = off; off += al(sizeof(uint64_t));
  const size_t oCount = off; off += al(sizeof(int32_t));
  const size_t oCost = off; off += al(4 * (size_t)m * ld);
  const size_t oCostT = off; off += al(4 * (size_t)n * ldT);
  const size_t oSup = off; off += al(4 * (size_t)m);
  const size_t oDem = off; off += al(4 * (size_t)n);
  const size_t oRow = off; off += al(4 * (size_t)V);
  const size_t oCol = off; off += al(4 * (size_t)V);
  const size_t oFlow = off; off += al(4 * (size_t)V);
  const size_t oDual = off; off += al(4 * (size_t)V);
  const size_t oAux = off; off += al(4 * flock::aux_ints(m, n));
  const size_t oList = off; off += al(12 * (size_t)V);
  if (off > kNoAllocBytes) {
    res->status = FLOCK_E_TOO_LARGE;
    return flock::set_error(FLOCK_E_TOO_LARGE,
                            "flock_solve_noalloc: %d x %d needs %zu bytes of its fixed %zu-byte device scratch "
                            "array (about 8 bytes per cell); use flock_solve",
                            m, n, off, kNoAllocBytes);
  }
  std::lock_guard<std::mutex> lk(g_mu);
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) { res->status = FLOCK_E_DEVICE; return flock::set_error(FLOCK_E_DEVICE, "hipGetDevice failed"); }
  if ((st = setup_device(dev))) { res->status = st; return st; }
  DeviceScratch& d = g_dev[dev];
  char* b = d.ptr;
  hipStream_t sm = flock::lib_stream();
  DevInstance D;
  std::memset(&D, 0, sizeof D);
  D.cost = (const int32_t*)(b + oCost);
  D.costT = (int32_t*)(b + oCostT);
  D.supply = (const int32_t*)(b + oSup);
  D.demand = (const int32_t*)(b + oDem);
  D.m = m; D.n = n; D.ld = ld; D.ldT = ldT;
  D.basis_row = (int32_t*)(b + oRow);
  D.basis_col = (int32_t*)(b + oCol);
  D.basis_flow = (int32_t*)(b + oFlow);
  D.dual = (int32_t*)(b + oDual);
  D.scratch = (int32_t*)(b + oAux);
  D.state = nullptr;
  const int32_t zero = 0;
#define NA_TRY(expr)                                                                              \
  do {                                                                                            \
    hipError_t e_ = (expr);                                                                       \
    if (e_ != hipSuccess) {                                                                       \
      res->status = FLOCK_E_DEVICE;                                                               \
      return flock::set_error(FLOCK_E_DEVICE, "%s failed: %s", #expr, hipGetErrorString(e_));      \
    }                                                                                             \
  } while (0)
  NA_TRY(hipMemcpy2DAsync(b + oCost, 4 * (size_t)ld, inst->cost, 4 * (size_t)n, 4 * (size_t)n, m,
                          hipMemcpyHostToDevice, sm));
  NA_TRY(hipMemcpyAsync(b + oSup, inst->supply, 4 * (size_t)m, hipMemcpyHostToDevice, sm));
  NA_TRY(hipMemcpyAsync(b + oDem, inst->demand, 4 * (size_t)n, hipMemcpyHostToDevice, sm));
  NA_TRY(hipMemcpyAsync(b + oDesc, &D, sizeof D, hipMemcpyHostToDevice, sm));
  NA_TRY(hipMemcpyAsync(b + oOrder, &zero, sizeof zero, hipMemcpyHostToDevice, sm));
  flock::EngineParams P = flock::make_params(opt);
  const int cfg = flock::pick_cfg(m, n, opt, 1);
  bool auxLds = false;
  const size_t lds = flock::cfg_group_bytes(cfg, m, n, &auxLds);
  NA_TRY(hipEventRecord(d.e0, sm));
  DevInstance* dDesc = (DevInstance*)(b + oDesc);
  int32_t* dHead = (int32_t*)(b + oHead);
  if ((st = flock::launch_reset(dDesc, 1, dHead, 1, sm, (uint64_t*)(b + oT0)))) { res->status = st; return st; }
  if ((st = flock::launch_queue_on(cfg, dDesc, (const int32_t*)(b + oOrder), 1, dHead, P, lds, 0, sm))) {
    res->status = st;
    return st;
  }
  NA_TRY(hipEventRecord(d.e1, sm));
  hipLaunchKernelGGL(k_flow_list, dim3(1), dim3(1024), 0, sm, (const int32_t*)(b + oRow), (const int32_t*)(b + oCol),
                     (const int32_t*)(b + oFlow), (int32_t)(V - 1), (int32_t*)(b + oList), (int32_t*)(b + oCount));
  NA_TRY(hipGetLastError());
  DevInstance out;
  NA_TRY(hipMemcpyAsync(&out, b + oDesc, sizeof out, hipMemcpyDeviceToHost, sm));
  int32_t nflows = 0;
  NA_TRY(hipMemcpyAsync(&nflows, b + oCount, sizeof nflows, hipMemcpyDeviceToHost, sm));
  NA_TRY(hipStreamSynchronize(sm));
  if (res->basis_row) NA_TRY(hipMemcpyAsync(res->basis_row, b + oRow, 4 * (size_t)(V - 1), hipMemcpyDeviceToHost, sm));
  if (res->basis_col) NA_TRY(hipMemcpyAsync(res->basis_col, b + oCol, 4 * (size_t)(V - 1), hipMemcpyDeviceToHost, sm));
  if (res->basis_flow) NA_TRY(hipMemcpyAsync(res->basis_flow, b + oFlow, 4 * (size_t)(V - 1), hipMemcpyDeviceToHost, sm));
  if (res->u) NA_TRY(hipMemcpyAsync(res->u, b + oDual, 4 * (size_t)m, hipMemcpyDeviceToHost, sm));
  if (res->v) NA_TRY(hipMemcpyAsync(res->v, b + oDual + 4 * (size_t)m, 4 * (size_t)n, hipMemcpyDeviceToHost, sm));
  if (res->flows && nflows > 0)
    NA_TRY(hipMemcpyAsync(res->flows, b + oList, 12 * (size_t)nflows, hipMemcpyDeviceToHost, sm));
  NA_TRY(hipStreamSynchronize(sm));
  float ms = 0;
  NA_TRY(hipEventElapsedTime(&ms, d.e0, d.e1));
#undef NA_TRY
  res->objective = out.objective;
  res->iterations = out.iterations;
  res->status = out.status;
  res->sweeps = out.sweeps;
  res->n_flows = nflows;
  res->solve_ms = ms;
  const double tpu = flock::lib_ticks_per_us();
  for (int k = 0; k < 6; ++k) res->phase_us[k] = (double)out.ticks[k] / tpu;
  res->start_us = 0;
  res->done_us = (double)(out.t_done - out.t_start) / tpu;
  return res->status;
}
