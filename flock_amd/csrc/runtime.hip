// runtime.hip -- host side of the C ABI (include/flock_amd.h): validation, device
// memory layout, kernel selection and launches.  One library stream per process, at
// the default priority.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/flock_amd.h"
#include "engine.h"
#include "runtime_internal.h"

using namespace flock;

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
// Persistent work-queue kernel (body: queue_worker in engine.h).  The workgroup-per-
// instance configurations of up to 512 threads are held to 168 VGPRs, i.e. three waves
// per SIMD; the few values that spill are stored at kernel entry and reloaded once per
// instance, outside the pivot loop.
template <int T, int G, bool M>
struct QueueWaves { static constexpr int value = (G == T && T <= 512 && !M) ? 3 : 1; };
template <int T, int G, bool M>
__global__ __launch_bounds__(T) __attribute__((amdgpu_waves_per_eu(QueueWaves<T, G, M>::value))) void k_solve_queue(DevInstance* insts, const int32_t* order, int32_t count,
                                                   int32_t* head, EngineParams P) {
  extern __shared__ __align__(16) char smem[];
  queue_worker<T, G, M>(insts, order, count, head, P, smem);
}

// Large-instance path (bodies in engine.h): setup and rounds run on one workgroup of
// T = 512 or 1024 threads, the sweep on a grid of 256-thread workgroups, one matrix tile
// each.
template <int T>
__global__ __launch_bounds__(T) void k_large_setup(DevInstance* insts, const int16_t* books, EngineParams P) {
  extern __shared__ __align__(16) char smem[];
  large_setup<T>(insts[0], smem, (size_t)P.lds_bytes, P, gl(books));
}
// Transpose of the cost matrix, one 64x64 tile per workgroup staged through LDS.
__global__ __launch_bounds__(256) void k_large_transpose(const DevInstance* insts) {
  __shared__ int32_t tile[64 * 65];
  const DevInstance& D = insts[0];
  const int m = D.m, n = D.n, ld = D.ld, ldT = D.ldT;
  const int32_t* cost = D.cost;
  int32_t* costT = D.costT;
  const int bi = (int)blockIdx.y * 64, bj = (int)blockIdx.x * 64;
  const int tx = (int)threadIdx.x % 64, ty = (int)threadIdx.x / 64;
  for (int r = ty; r < 64; r += 4) {
    const int i = bi + r, j = bj + tx;
    tile[r * 65 + tx] = (i < m && j < n) ? cost[(size_t)i * ld + j] : 0;
  }
  __syncthreads();
  for (int r = ty; r < 64; r += 4) {
    const int j = bj + r, i = bi + tx;
    if (j < n && i < m) costT[(size_t)j * ldT + i] = tile[tx * 65 + r];
  }
}
// Vogel books: one workgroup sorts one line (rows, then columns).
__global__ __launch_bounds__(256) void k_large_books(const DevInstance* insts, int16_t* books) {
  __shared__ uint32_t keys[16384];  // a line of up to 16384 entries (m+n <= 16383)
  const DevInstance& D = insts[0];
  Ctx C;
  C.cost = gl(uniform_ptr(D.cost)); C.costT = gl(uniform_ptr(D.costT));
  C.m = uniform(D.m); C.n = uniform(D.n); C.V = C.m + C.n; C.ld = uniform(D.ld); C.ldT = uniform(D.ldT);
  line_book<256>(C, (int)blockIdx.x, keys, gl(books));
}
template <int T, bool M>
__global__ __launch_bounds__(T) void k_large_round(DevInstance* insts, const uint64_t* keys, const int32_t* cands,
                                                   int32_t nblocks, EngineParams P) {
  extern __shared__ __align__(16) char smem[];
  large_round<T, M>(insts[0], smem, (size_t)P.lds_bytes, P, keys, cands, nblocks);
}
__global__ __launch_bounds__(256) void k_large_sweep(SweepArgs A) {
  __shared__ Shared S;
  sweep_tile<256>(A, blockIdx.x, S);
}

// The launch-to-completion floor of a grid: the same grid as the sweep, no work.
__global__ __launch_bounds__(256) void k_empty_grid() {}

__global__ void k_reset(DevInstance* insts, int32_t count, int32_t* heads, int32_t nheads, uint64_t* t0) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) { insts[i].phase = 0; insts[i].status = 0; insts[i].iterations = 0; insts[i].sweeps = 0; }
  if (i < nheads) heads[i] = 0;
  if (i == 0 && t0) *t0 = wall_clock64();
}

namespace flock {

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
static thread_local std::string g_err;

int set_error(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIP_TRY(expr)                                                                       \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      return set_error(FLOCK_E_DEVICE, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                       __FILE__, __LINE__);                                                 \
  } while (0)

// ---------------------------------------------------------------------------
// device context
// ---------------------------------------------------------------------------
static std::mutex g_mu;
static bool g_init = false;
static int g_device = 0;
static int g_cus = 0;
static size_t g_lds_max = 0;
static hipStream_t g_stream = nullptr;
static double g_ticks_per_us = 100.0;

int ensure_init(int device) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_init) return FLOCK_OK;
  int ndev = 0;
  HIP_TRY(hipGetDeviceCount(&ndev));
  if (ndev <= 0) return set_error(FLOCK_E_DEVICE, "no HIP device");
  if (device < 0) HIP_TRY(hipGetDevice(&device));
  HIP_TRY(hipSetDevice(device));
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, device));
  g_device = device;
  g_cus = prop.multiProcessorCount;
  g_lds_max = prop.maxSharedMemoryPerMultiProcessor ? prop.maxSharedMemoryPerMultiProcessor : 65536;
  if (g_lds_max > 160 * 1024) g_lds_max = 160 * 1024;
  HIP_TRY(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
  int wclk_khz = 0;
  if (hipDeviceGetAttribute(&wclk_khz, hipDeviceAttributeWallClockRate, device) == hipSuccess && wclk_khz > 0)
    g_ticks_per_us = wclk_khz / 1000.0;
  g_init = true;
  return FLOCK_OK;
}

hipStream_t lib_stream() { return g_stream; }
double lib_ticks_per_us() { return g_ticks_per_us; }
int lib_cus() { return g_cus; }
size_t lib_lds_max() { return g_lds_max; }

// ---------------------------------------------------------------------------
// validation (exactness contract of the int32 engine)
// ---------------------------------------------------------------------------
int validate_instance(const flock_instance* in, int max_v) {
  if (!in || !in->cost || !in->supply || !in->demand) return set_error(FLOCK_E_INVALID, "null instance pointer");
  const int64_t m = in->m, n = in->n;
  if (m < 1 || n < 1) return set_error(FLOCK_E_INVALID, "m=%lld n=%lld must be >= 1", (long long)m, (long long)n);
  if (m + n > max_v)
    return set_error(FLOCK_E_TOO_LARGE, "m+n=%lld exceeds %d", (long long)(m + n), max_v);
  int64_t ss = 0, sd = 0;
  for (int64_t i = 0; i < m; ++i) {
    if (in->supply[i] < 0) return set_error(FLOCK_E_INVALID, "negative supply at %lld", (long long)i);
    ss += in->supply[i];
  }
  for (int64_t j = 0; j < n; ++j) {
    if (in->demand[j] < 0) return set_error(FLOCK_E_INVALID, "negative demand at %lld", (long long)j);
    sd += in->demand[j];
  }
  if (ss != sd) return set_error(FLOCK_E_UNBALANCED, "sum(supply)=%lld != sum(demand)=%lld", (long long)ss, (long long)sd);
  if (ss > 2147483647LL) return set_error(FLOCK_E_RANGE, "total supply %lld exceeds int32", (long long)ss);
  int32_t cmax = 0;
  const int64_t cells = m * n;
  for (int64_t k = 0; k < cells; ++k) {
    const int32_t c = in->cost[k];
    if (c < 0) return set_error(FLOCK_E_RANGE, "negative cost at %lld", (long long)k);
    cmax = c > cmax ? c : cmax;
  }
  // |u|,|v| <= (m+n-1)*cmax and |r| <= (2(m+n)+1)*cmax must fit int32.
  if ((2 * (m + n) + 1) * (int64_t)cmax >= 2147483647LL)
    return set_error(FLOCK_E_RANGE, "cost range %d too large for exact int32 duals at m+n=%lld", cmax,
                     (long long)(m + n));
  return FLOCK_OK;
}

// ---------------------------------------------------------------------------
// kernel configurations
// ---------------------------------------------------------------------------
struct KernelCfg {
  int threads;  // workgroup size
  int group;    // threads per instance (64 = one wavefront per instance)
  int max_v;    // largest m+n the configuration accepts
  const void* fn;        // Dantzig and candidate-list rules
  const void* fn_multi;  // multi-pivot rule
};

static const KernelCfg kCfgs[] = {
#define FLOCK_CFG(T, G, MAXV) {T, G, MAXV, (const void*)k_solve_queue<T, G, false>, (const void*)k_solve_queue<T, G, true>}
    FLOCK_CFG(256, 64, 2048),
    FLOCK_CFG(128, 64, 2048),
    FLOCK_CFG(64, 64, 2048),
    FLOCK_CFG(256, 256, kMaxVertices),
    FLOCK_CFG(512, 512, kMaxVertices),
    FLOCK_CFG(1024, 1024, kMaxVertices),
#undef FLOCK_CFG
};
constexpr int kNumCfgs = (int)(sizeof(kCfgs) / sizeof(kCfgs[0]));

// size-class policy (tuned on MI355X; see autotune): instances with m+n up to
// g_wave_max_v get one wavefront each, packed g_wave_pack per workgroup; larger ones
// get a whole workgroup of 256 threads up to g_wg256_max_v, else 1024 threads.
static int g_wave_max_v = 0;
static int g_wave_pack = 4;
static int g_wg256_max_v = 2000;
static int g_minor_limit = 64;
// per mille of each launch's instances, largest first, whose workgroups run at raised
// wave priority
static int g_prio_permille = 0;
// large-instance path: 16-byte cost vectors per thread in one tile of the sweep grid
// (256 threads per tile), i.e. the tile shape; tiles = ceil(vectors / (256 * this))
static int g_sweep_vectors = 4;
static int g_large_threads = 512;  // large-instance setup / round workgroup (512 or 1024)
static int g_large_minor_limit = 64;  // large-instance path's candidate minor limit (0: g_minor_limit)
static int g_few_256_max_v = 256;
static int g_few_512_max_v = 1000;
// plain solves of matrices with at least this many cells take the graph path
static int g_graph_min_cells = 250000;

int num_cfgs() { return kNumCfgs; }

int size_bucket(int V) {
  if (V <= 128) return 0;
  if (V <= 256) return 1;
  if (V <= 512) return 2;
  if (V <= 1024) return 3;
  if (V <= 2048) return 4;
  return 5;
}
// LDS is sized per launch for its largest instance; wavefront configurations pack
// several groups into one workgroup, so they are split by size bucket to keep small
// instances dense.  Workgroup configurations launch once per configuration.
int class_bucket(int cfg, int V) { return kCfgs[cfg].group == 64 ? size_bucket(V) : 0; }
int cfg_threads(int c) { return kCfgs[c].threads; }
int cfg_group(int c) { return kCfgs[c].group; }

void set_policy(int wave_max_v, int wave_pack, int wg256_max_v) {
  if (wave_max_v >= 0) g_wave_max_v = std::min(wave_max_v, 2048);
  if (wave_pack == 1 || wave_pack == 2 || wave_pack == 4) g_wave_pack = wave_pack;
  if (wg256_max_v >= 0) g_wg256_max_v = wg256_max_v;
}
void get_policy(int* wave_max_v, int* wave_pack, int* wg256_max_v) {
  *wave_max_v = g_wave_max_v; *wave_pack = g_wave_pack; *wg256_max_v = g_wg256_max_v;
}

// The 256- and 512-thread engines need the tree and the pivot's snapshot in LDS.
static bool snap_fits_lds(int m, int n) {
  return shared_bytes() + tree_lds_bytes(m, n) + 4 * aux_ints(m, n) <= g_lds_max;
}
static int pick_cfg_raw(int m, int n, const flock_options* opt, int64_t count);
int pick_cfg(int m, int n, const flock_options* opt, int64_t count) {
  const int c = pick_cfg_raw(m, n, opt, count);
  return (c == 3 || c == 4) && !snap_fits_lds(m, n) ? 5 : c;
}
static int pick_cfg_raw(int m, int n, const flock_options* opt, int64_t count) {
  const int V = m + n;
  const int threads = opt ? opt->threads : 0;   // threads per instance
  const int pack = opt ? opt->team : 0;         // instances per workgroup for wave groups
  if (threads == 64 || (threads == 0 && V <= g_wave_max_v)) {
    if (V > 2048) return 5;
    const int p = pack > 0 ? pack : g_wave_pack;
    return p >= 4 ? 0 : (p >= 2 ? 1 : 2);
  }
  if (threads == 256) return 3;
  if (threads == 512) return 4;
  if (threads == 1024) return 5;
  if (count <= g_cus) return V <= g_few_256_max_v ? 3 : (V <= g_few_512_max_v ? 4 : 5);
  return V <= g_wg256_max_v ? 3 : 5;
}

size_t cfg_group_bytes(int cfg, int m, int n, bool* aux_in_lds) {
  const int gpw = kCfgs[cfg].threads / kCfgs[cfg].group;
  return align16(group_lds_bytes(m, n, g_lds_max / gpw, aux_in_lds));
}

int launch_queue_on(int cfg, DevInstance* d_insts, const int32_t* d_order, int32_t count, int32_t* d_head,
                    const EngineParams& P0, size_t group_bytes, int grid_cap, hipStream_t stream) {
  EngineParams P = P0;
  P.lds_bytes = (int32_t)group_bytes;
  P.prio_count = (int32_t)(((int64_t)count * g_prio_permille + 999) / 1000);
  const KernelCfg& K = kCfgs[cfg];
  const int gpw = K.threads / K.group;
  const size_t lds = group_bytes * gpw;
  if (lds > g_lds_max) return set_error(FLOCK_E_TOO_LARGE, "LDS need %zu exceeds %zu", lds, g_lds_max);
  const void* fn = (P.pricing == 2 || P.pricing == 4) ? K.fn_multi : K.fn;
  if (lds > 65536) HIP_TRY(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  int occ = 0;
  HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fn, K.threads, lds));
  if (occ < 1) occ = 1;
  int grid = std::min<int64_t>((int64_t)(count + gpw - 1) / gpw, (int64_t)g_cus * occ);
  if (grid_cap > 0) grid = std::min(grid, grid_cap);
  if (grid < 1) grid = 1;
  void* args[] = {&d_insts, &d_order, &count, &d_head, &P};
  HIP_TRY(hipLaunchKernel(fn, dim3(grid), dim3(K.threads), args, lds, stream));
  return FLOCK_OK;
}

int launch_queue(int cfg, DevInstance* d_insts, const int32_t* d_order, int32_t count, int32_t* d_head,
                 const EngineParams& P, size_t lds, int grid_cap) {
  return launch_queue_on(cfg, d_insts, d_order, count, d_head, P, lds, grid_cap, g_stream);
}

int launch_reset(DevInstance* d_insts, int32_t count, int32_t* d_heads, int32_t nheads, hipStream_t stream,
                 uint64_t* d_t0) {
  const int n = std::max(count, nheads);
  k_reset<<<(n + 255) / 256, 256, 0, stream ? stream : g_stream>>>(d_insts, count, d_heads, nheads, d_t0);
  HIP_TRY(hipGetLastError());
  return FLOCK_OK;
}

EngineParams make_params(const flock_options* opt) {
  EngineParams P;
  std::memset(&P, 0, sizeof P);
  flock_options o;
  flock_default_options(&o);
  if (opt) o = *opt;
  P.max_iterations = o.max_iterations;
  P.initial = o.initial;
  P.pricing = o.pricing == FLOCK_PRICE_AUTO ? FLOCK_PRICE_CANDIDATES : o.pricing;
  P.step_pivots = 0;
  P.minor_limit = g_minor_limit;
  return P;
}

// ---------------------------------------------------------------------------
// batch layout in one device allocation
// ---------------------------------------------------------------------------
static size_t al(size_t x) { return (x + 255) & ~(size_t)255; }

int Batch::layout(int32_t count, const flock_instance* inst, bool with_state) {
  n = count;
  hdesc.assign(count, DevInstance{});
  offs.assign(count, Offs{});
  size_t off = al(sizeof(DevInstance) * (size_t)count);
  off += al(sizeof(int32_t) * (size_t)count);  // order
  off += al(sizeof(int32_t) * kMaxClasses);    // queue heads
  off += al(sizeof(uint64_t));                 // device clock at the batch launch
  for (int k = 0; k < count; ++k) {
    const int m = inst[k].m, nn = inst[k].n, V = m + nn;
    const int ld = (nn + 3) & ~3, ldT = (m + 3) & ~3;
    Offs& o = offs[k];
    o.m = m; o.n = nn; o.ld = ld; o.ldT = ldT;
    o.cost = off; off += al(4 * (size_t)m * ld);
    o.costT = off; off += al(4 * (size_t)nn * ldT);
    o.supply = off; off += al(4 * (size_t)m);
    o.demand = off; off += al(4 * (size_t)nn);
    o.brow = off; off += al(4 * (size_t)V);
    o.bcol = off; off += al(4 * (size_t)V);
    o.bflow = off; off += al(4 * (size_t)V);
    o.dual = off; off += al(4 * (size_t)V);
    o.scratch = off; off += al(4 * aux_ints(m, nn));
    o.state = with_state ? off : 0; if (with_state) off += al(4 * 7 * (size_t)V);
    o.books = with_state ? off : 0; if (with_state) off += al(4 * (size_t)m * nn);
  }
  bytes = off;
  return FLOCK_OK;
}

int Batch::alloc() {
  if (bytes > cap) {
    if (base) hipFree(base);
    base = nullptr; cap = 0;
    HIP_TRY(hipMalloc(&base, bytes));
    cap = bytes;
  }
  d_insts = (DevInstance*)base;
  d_order = (int32_t*)((char*)base + al(sizeof(DevInstance) * (size_t)n));
  d_heads = (int32_t*)((char*)d_order + al(sizeof(int32_t) * (size_t)n));
  d_t0 = (uint64_t*)((char*)d_heads + al(sizeof(int32_t) * kMaxClasses));
  for (int k = 0; k < n; ++k) {
    const Offs& o = offs[k];
    DevInstance& D = hdesc[k];
    char* b = (char*)base;
    D.cost = (const int32_t*)(b + o.cost);
    D.costT = (int32_t*)(b + o.costT);
    D.supply = (const int32_t*)(b + o.supply);
    D.demand = (const int32_t*)(b + o.demand);
    D.m = o.m; D.n = o.n; D.ld = o.ld; D.ldT = o.ldT;
    D.basis_row = (int32_t*)(b + o.brow);
    D.basis_col = (int32_t*)(b + o.bcol);
    D.basis_flow = (int32_t*)(b + o.bflow);
    D.dual = (int32_t*)(b + o.dual);
    D.scratch = (int32_t*)(b + o.scratch);
    D.state = o.state ? (int32_t*)(b + o.state) : nullptr;
    D.objective = 0; D.iterations = 0; D.status = 0; D.phase = 0; D.n_flows = 0; D.sweeps = 0;
  }
  return FLOCK_OK;
}

int Batch::upload(const flock_instance* inst, char* pinned, size_t pinned_cap) {
  for (int k = 0; k < n; ++k) {
    const Offs& o = offs[k];
    char* b = (char*)base;
    const size_t rows = 4 * (size_t)o.m * o.ld;
    if (pinned && n == 1 && rows <= pinned_cap) {  // padded rows staged in pinned memory: one fast copy
      for (int i = 0; i < o.m; ++i)
        std::memcpy(pinned + 4 * (size_t)i * o.ld, inst[k].cost + (size_t)i * o.n, 4 * (size_t)o.n);
      HIP_TRY(hipMemcpyAsync(b + o.cost, pinned, rows, hipMemcpyHostToDevice, g_stream));
    } else {
      HIP_TRY(hipMemcpy2DAsync(b + o.cost, 4 * (size_t)o.ld, inst[k].cost, 4 * (size_t)o.n, 4 * (size_t)o.n, o.m,
                               hipMemcpyHostToDevice, g_stream));
    }
    HIP_TRY(hipMemcpyAsync(b + o.supply, inst[k].supply, 4 * (size_t)o.m, hipMemcpyHostToDevice, g_stream));
    HIP_TRY(hipMemcpyAsync(b + o.demand, inst[k].demand, 4 * (size_t)o.n, hipMemcpyHostToDevice, g_stream));
  }
  HIP_TRY(hipMemcpyAsync(d_insts, hdesc.data(), sizeof(DevInstance) * (size_t)n, hipMemcpyHostToDevice, g_stream));
  // Work order: largest estimated work first (pivots ~ m+n, each sweeping m*n cells).
  std::vector<int32_t> order(n);
  for (int k = 0; k < n; ++k) order[k] = k;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    const double wa = (double)offs[a].m * offs[a].n * (offs[a].m + offs[a].n);
    const double wb = (double)offs[b].m * offs[b].n * (offs[b].m + offs[b].n);
    return wa > wb;
  });
  sorted_order = order;
  planned_for = -1;
  return FLOCK_OK;
}

// Split the work order into per-configuration classes (contiguous in d_order).
int Batch::plan(const flock_options* opt) {
  const int key = opt ? (opt->threads * 16 + opt->team) : 0;
  int wm, wp, w2;
  get_policy(&wm, &wp, &w2);
  const int64_t pkey = (((int64_t)key * 1000003 + wm * 31 + wp * 7 + w2) * 1000003 + g_few_256_max_v) * 1000003 +
                       g_few_512_max_v;
  if (planned_for == pkey) return FLOCK_OK;
  classes.clear();
  std::vector<int32_t>& grouped = class_order;  // kept alive: the upload below is asynchronous
  grouped.clear();
  // classes = (configuration, size bucket): LDS is sized per launch for its largest
  // instance, so instances are bucketed by m+n to keep small ones densely packed
  for (int c = 0; c < num_cfgs(); ++c) {
    for (int b = 0; b < kSizeBuckets; ++b) {
      Class cl;
      cl.cfg = c; cl.begin = (int)grouped.size(); cl.lds = 0;
      bool dummy;
      for (int k : sorted_order) {
        if (pick_cfg(offs[k].m, offs[k].n, opt, n) != c || class_bucket(c, offs[k].m + offs[k].n) != b) continue;
        grouped.push_back(k);
        cl.lds = std::max(cl.lds, cfg_group_bytes(c, offs[k].m, offs[k].n, &dummy));
      }
      cl.count = (int)grouped.size() - cl.begin;
      if (cl.count) classes.push_back(cl);
    }
  }
  HIP_TRY(hipMemcpyAsync(d_order, grouped.data(), sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice, g_stream));
  planned_for = pkey;
  return FLOCK_OK;
}

// Classes launch one after another on the library stream (one kernel at a time).
int Batch::run(const EngineParams& P, int grid_cap) {
  if (classes.size() > (size_t)kMaxClasses) return set_error(FLOCK_E_INVALID, "too many launch classes");
  int st = launch_reset(d_insts, n, d_heads, kMaxClasses, nullptr, d_t0);
  if (st) return st;
  int h = 0;
  for (const Class& cl : classes) {
    st = launch_queue(cl.cfg, d_insts, d_order + cl.begin, cl.count, d_heads + h, P, cl.lds, grid_cap);
    if (st) return st;
    ++h;
  }
  return FLOCK_OK;
}

int Batch::download_desc() {
  HIP_TRY(hipMemcpyAsync(hdesc.data(), d_insts, sizeof(DevInstance) * (size_t)n, hipMemcpyDeviceToHost, g_stream));
  HIP_TRY(hipMemcpyAsync(&h_t0, d_t0, sizeof(uint64_t), hipMemcpyDeviceToHost, g_stream));
  HIP_TRY(hipStreamSynchronize(g_stream));
  return FLOCK_OK;
}

int Batch::fill_result(int k, flock_result* r) {
  const DevInstance& D = hdesc[k];
  const Offs& o = offs[k];
  const int V = o.m + o.n;
  r->objective = D.objective;
  r->iterations = D.iterations;
  r->status = D.status;
  r->sweeps = D.sweeps;
  for (int k = 0; k < 6; ++k) r->phase_us[k] = (double)D.ticks[k] / g_ticks_per_us;
  r->start_us = (double)(int64_t)(D.t_start - h_t0) / g_ticks_per_us;
  r->done_us = (double)(int64_t)(D.t_done - h_t0) / g_ticks_per_us;
  r->n_flows = 0;
  char* b = (char*)base;
  std::vector<int32_t> br, bc, bf;
  const bool need_basis = r->flows || r->basis_row || r->basis_col || r->basis_flow;
  if (need_basis) {
    br.resize(V - 1); bc.resize(V - 1); bf.resize(V - 1);
    if (V > 1) {
      HIP_TRY(hipMemcpyAsync(br.data(), b + o.brow, 4 * (size_t)(V - 1), hipMemcpyDeviceToHost, g_stream));
      HIP_TRY(hipMemcpyAsync(bc.data(), b + o.bcol, 4 * (size_t)(V - 1), hipMemcpyDeviceToHost, g_stream));
      HIP_TRY(hipMemcpyAsync(bf.data(), b + o.bflow, 4 * (size_t)(V - 1), hipMemcpyDeviceToHost, g_stream));
    }
  }
  if (r->u) HIP_TRY(hipMemcpyAsync(r->u, b + o.dual, 4 * (size_t)o.m, hipMemcpyDeviceToHost, g_stream));
  if (r->v) HIP_TRY(hipMemcpyAsync(r->v, b + o.dual + 4 * (size_t)o.m, 4 * (size_t)o.n, hipMemcpyDeviceToHost, g_stream));
  HIP_TRY(hipStreamSynchronize(g_stream));
  if (need_basis) {
    if (r->basis_row) std::memcpy(r->basis_row, br.data(), 4 * (size_t)(V - 1));
    if (r->basis_col) std::memcpy(r->basis_col, bc.data(), 4 * (size_t)(V - 1));
    if (r->basis_flow) std::memcpy(r->basis_flow, bf.data(), 4 * (size_t)(V - 1));
    int nf = 0;
    for (int s = 0; s < V - 1; ++s) {
      if (bf[s] == 0) continue;
      if (r->flows) r->flows[nf] = flock_flow{br[s], bc[s], bf[s]};
      ++nf;
    }
    r->n_flows = nf;
  } else {
    r->n_flows = D.n_flows;
  }
  return FLOCK_OK;
}

Batch::~Batch() {
  if (base) hipFree(base);
}

}  // namespace flock

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
namespace flock {

// ---------------------------------------------------------------------------
// Graph path for one large instance: setup kernel, then a captured graph of
// kRoundsPerGraph x (sweep grid, pivot round) launched until the instance is done.
// Two graphs are kept in flight; after each one the instance's phase is copied to
// pinned host memory and checked when the graph behind it has been queued.
// ---------------------------------------------------------------------------
constexpr int kRoundsPerGraph = 32;
constexpr int kLargeTiles = 1024;   // sweep workgroups at most
// setup / round workgroup: 1024 threads, or 512 (g_large_threads) for instances whose
// tree stays in LDS and whose m+n is at most 2048 (four vertices per rank in a pivot)
constexpr int kLargeSmallMaxV = 2048;

struct LargeRunner {
  uint64_t* keys = nullptr;
  int32_t* cands = nullptr;
  int32_t* h_phase = nullptr;       // pinned, two slots
  hipEvent_t ev[2] = {nullptr, nullptr};
  hipGraphExec_t exec = nullptr;
  // what the captured graph was built for
  DevInstance* g_insts = nullptr;
  int g_nblocks = 0;
  int g_threads = 0;
  SweepArgs g_A{};
  EngineParams g_P{};

  int init() {
    if (keys) return FLOCK_OK;
    HIP_TRY(hipMalloc(&keys, sizeof(uint64_t) * kLargeTiles));
    HIP_TRY(hipMalloc(&cands, sizeof(int32_t) * 3 * kTileCands * kLargeTiles));
    HIP_TRY(hipHostMalloc((void**)&h_phase, sizeof(int32_t) * 2, hipHostMallocDefault));
    for (auto& e : ev) HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    return FLOCK_OK;
  }

  int graph(DevInstance* d_insts, int nblocks, const SweepArgs& A, const EngineParams& P, size_t lds, int threads) {
    if (exec && g_insts == d_insts && g_nblocks == nblocks && g_threads == threads &&
        std::memcmp(&g_A, &A, sizeof A) == 0 && std::memcmp(&g_P, &P, sizeof P) == 0)
      return FLOCK_OK;
    if (exec) { hipGraphExecDestroy(exec); exec = nullptr; }
    hipGraph_t gr = nullptr;
    HIP_TRY(hipStreamBeginCapture(g_stream, hipStreamCaptureModeThreadLocal));
    for (int k = 0; k < kRoundsPerGraph; ++k) {
      hipLaunchKernelGGL(k_large_sweep, dim3(nblocks), dim3(256), 0, g_stream, A);
      const bool multi = P.pricing == 2 || P.pricing == 4;
      const void* fn = threads == 512 ? (multi ? (const void*)k_large_round<512, true> : (const void*)k_large_round<512, false>)
                                      : (multi ? (const void*)k_large_round<1024, true> : (const void*)k_large_round<1024, false>);
      const uint64_t* kp = keys;
      const int32_t* cp = cands;
      int32_t nb = nblocks;
      void* args[] = {(void*)&d_insts, (void*)&kp, (void*)&cp, (void*)&nb, (void*)&P};
      (void)hipLaunchKernel(fn, dim3(1), dim3(threads), args, lds, g_stream);
    }
    const hipError_t ce = hipStreamEndCapture(g_stream, &gr);
    if (ce != hipSuccess) return set_error(FLOCK_E_DEVICE, "graph capture failed: %s", hipGetErrorString(ce));
    const hipError_t ie = hipGraphInstantiate(&exec, gr, nullptr, nullptr, 0);
    hipGraphDestroy(gr);
    if (ie != hipSuccess) { exec = nullptr; return set_error(FLOCK_E_DEVICE, "graph instantiate failed"); }
    g_insts = d_insts; g_nblocks = nblocks; g_A = A; g_P = P; g_threads = threads;
    return FLOCK_OK;
  }

  // Sweep grid of an m x ld matrix: tiles of vpt 16-byte vectors (g_sweep_vectors per
  // thread of a 256-thread tile, more when that would exceed kLargeTiles workgroups).
  static void sweep_grid(const Batch::Offs& o, uint32_t* vpt, int* nblocks) {
    const uint32_t nvec = (uint32_t)o.m * ((uint32_t)o.ld >> 2);
    *vpt = std::max<uint32_t>((uint32_t)g_sweep_vectors * 256, (nvec + kLargeTiles - 1) / kLargeTiles);
    *nblocks = (int)((nvec + *vpt - 1) / *vpt);
  }

  // Solve instance 0 of batch b (uploaded, with state storage) on the library stream.
  int run(Batch& b, const EngineParams& P0) {
    int st = init();
    if (st) return st;
    const Batch::Offs& o = b.offs[0];
    uint32_t vpt;
    int nblocks;
    sweep_grid(o, &vpt, &nblocks);
    bool dummy;
    size_t lds = align16(group_lds_bytes(o.m, o.n, g_lds_max, &dummy));
    EngineParams P = P0;
    P.tree_global = 0;
    if (lds > g_lds_max) {
      // the tree does not fit LDS: it lives in the state buffer; Vogel's line state stays
      // in LDS when it fits, else it goes to global memory too
      const size_t vog = shared_bytes() + vogel_lds_bytes(o.m, o.n);
      lds = align16(vog <= g_lds_max ? vog : shared_bytes());
      P.tree_global = 1;
    }
    const int threads = (g_large_threads == 512 && !P.tree_global && o.m + o.n <= kLargeSmallMaxV) ? 512 : 1024;
    const void* fsetup = threads == 512 ? (const void*)k_large_setup<512> : (const void*)k_large_setup<1024>;
    if (lds > 65536) {
      const void* fns[] = {(const void*)k_large_round<512, false>, (const void*)k_large_round<512, true>,
                           (const void*)k_large_round<1024, false>, (const void*)k_large_round<1024, true>};
      for (const void* f : fns) HIP_TRY(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    }
    // the set-up kernel runs alone: it takes the LDS that Vogel's register mode needs for
    // its line state and tracked keys when that is more than the rounds' and still fits
    size_t slds = lds;
    {
      const size_t want = align16(shared_bytes() + vogel_reg_lds_bytes(o.m, o.n) + vogel_trk_bytes(o.m, o.n));
      if (want > slds && want <= g_lds_max) slds = want;
    }
    if (slds > 65536)
      HIP_TRY(hipFuncSetAttribute(fsetup, hipFuncAttributeMaxDynamicSharedMemorySize, (int)slds));
    P.lds_bytes = (int32_t)lds;
    if (g_large_minor_limit > 0) P.minor_limit = g_large_minor_limit;
    if ((st = launch_reset(b.d_insts, 1, b.d_heads, 1, nullptr, b.d_t0))) return st;
    // Vogel's sorted lines: transpose and sort on the whole device, then one workgroup
    const int16_t* books = nullptr;
    if (P.initial != 1) {
      int16_t* bk = (int16_t*)((char*)b.base + o.books);
      hipLaunchKernelGGL(k_large_transpose, dim3((o.n + 63) / 64, (o.m + 63) / 64), dim3(256), 0, g_stream,
                         (const DevInstance*)b.d_insts);
      hipLaunchKernelGGL(k_large_books, dim3(o.m + o.n), dim3(256), 0, g_stream, (const DevInstance*)b.d_insts, bk);
      HIP_TRY(hipGetLastError());
      books = bk;
    }
    {
      DevInstance* di = b.d_insts;
      EngineParams Ps = P;
      Ps.lds_bytes = (int32_t)slds;
      void* args[] = {(void*)&di, (void*)&books, (void*)&Ps};
      HIP_TRY(hipLaunchKernel(fsetup, dim3(1), dim3(threads), args, slds, g_stream));
    }
    SweepArgs A = sweep_args(b.hdesc[0], o.m + o.n, keys, cands, vpt, &b.d_insts[0].phase);
    if ((st = graph(b.d_insts, nblocks, A, P, lds, threads))) return st;
    // every round pivots or finishes, so maxIt + 1 rounds always suffice
    const int64_t maxIt = P.max_iterations > 0 ? P.max_iterations : 50LL * (o.m + o.n) + 100000;
    const int64_t maxGraphs = (maxIt + 1) / kRoundsPerGraph + 3;
    const int32_t* d_phase = &b.d_insts[0].phase;
    auto queue_one = [&](int slot) -> int {
      HIP_TRY(hipGraphLaunch(exec, g_stream));
      HIP_TRY(hipMemcpyAsync(&h_phase[slot], d_phase, sizeof(int32_t), hipMemcpyDeviceToHost, g_stream));
      HIP_TRY(hipEventRecord(ev[slot], g_stream));
      return FLOCK_OK;
    };
    int64_t queued = 0;
    if ((st = queue_one(0))) return st;
    ++queued;
    for (int64_t k = 0;; ++k) {
      if (queued < maxGraphs) {
        if ((st = queue_one((int)(queued & 1)))) return st;
        ++queued;
      }
      HIP_TRY(hipEventSynchronize(ev[k & 1]));
      if (h_phase[k & 1] == 2) break;
      if (k + 1 >= queued) return set_error(FLOCK_E_DEVICE, "graph path did not finish within its round budget");
    }
    HIP_TRY(hipStreamSynchronize(g_stream));
    return FLOCK_OK;
  }
};

LargeRunner* large_runner() {
  static thread_local LargeRunner* r = new LargeRunner();  // process lifetime
  return r;
}

int graph_min_cells() { return g_graph_min_cells; }

// ---------------------------------------------------------------------------
// Plain solve of one instance on the per-instance engine, with the fewest host-device
// transfers: the descriptor, work order, queue head, supplies, demands and padded cost
// rows are staged in pinned memory and go over in one copy; one launch solves; the
// descriptor and the basis and duals come back in one copy.  Device layout (256-byte
// aligned): [descriptor][order][head][basis row][col][flow][duals][supply][demand][cost]
// [cost transposed][scratch].
// ---------------------------------------------------------------------------
struct PlainArena {
  char* dev = nullptr;
  char* host = nullptr;  // pinned
  size_t dev_cap = 0, host_cap = 0;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  ~PlainArena() {}  // process lifetime (thread_local pointer, never freed)

  int solve(const flock_instance* inst, const flock_options* opt, flock_result* res) {
    const int m = inst->m, n = inst->n, V = m + n;
    const int ld = (n + 3) & ~3, ldT = (m + 3) & ~3;
    size_t off = 0;
    const size_t oDesc = off; off += al(sizeof(DevInstance));
    const size_t oOrder = off; off += al(sizeof(int32_t));
    const size_t oHead = off; off += al(sizeof(int32_t));
    const size_t oRow = off; off += al(4 * (size_t)V);
    const size_t oCol = off; off += al(4 * (size_t)V);
    const size_t oFlow = off; off += al(4 * (size_t)V);
    const size_t oDual = off; off += al(4 * (size_t)V);
    const size_t outEnd = off;
    const size_t oSup = off; off += al(4 * (size_t)m);
    const size_t oDem = off; off += al(4 * (size_t)n);
    const size_t oCost = off; off += al(4 * (size_t)m * ld);
    const size_t inEnd = off;
    const size_t oCostT = off; off += al(4 * (size_t)n * ldT);
    const size_t oAux = off; off += al(4 * aux_ints(m, n));
    if (off > dev_cap) {
      if (dev) hipFree(dev);
      dev = nullptr; dev_cap = 0;
      HIP_TRY(hipMalloc(&dev, off));
      dev_cap = off;
    }
    if (inEnd > host_cap) {
      if (host) hipHostFree(host);
      host = nullptr; host_cap = 0;
      HIP_TRY(hipHostMalloc((void**)&host, inEnd, hipHostMallocDefault));
      host_cap = inEnd;
    }
    if (!e0) { HIP_TRY(hipEventCreate(&e0)); HIP_TRY(hipEventCreate(&e1)); }
    DevInstance D;
    std::memset(&D, 0, sizeof D);
    D.cost = (const int32_t*)(dev + oCost);
    D.costT = (int32_t*)(dev + oCostT);
    D.supply = (const int32_t*)(dev + oSup);
    D.demand = (const int32_t*)(dev + oDem);
    D.m = m; D.n = n; D.ld = ld; D.ldT = ldT;
    D.basis_row = (int32_t*)(dev + oRow);
    D.basis_col = (int32_t*)(dev + oCol);
    D.basis_flow = (int32_t*)(dev + oFlow);
    D.dual = (int32_t*)(dev + oDual);
    D.scratch = (int32_t*)(dev + oAux);
    D.state = nullptr;
    std::memcpy(host + oDesc, &D, sizeof D);
    std::memset(host + oOrder, 0, oRow - oOrder);  // order[0] = 0, queue head = 0
    std::memcpy(host + oSup, inst->supply, 4 * (size_t)m);
    std::memcpy(host + oDem, inst->demand, 4 * (size_t)n);
    if (ld == n) {
      std::memcpy(host + oCost, inst->cost, 4 * (size_t)m * n);
    } else {
      for (int i = 0; i < m; ++i)
        std::memcpy(host + oCost + 4 * (size_t)i * ld, inst->cost + (size_t)i * n, 4 * (size_t)n);
    }
    hipStream_t sm = g_stream;
    // the output region goes over too (as garbage): the inputs then travel in one copy
    HIP_TRY(hipMemcpyAsync(dev, host, inEnd, hipMemcpyHostToDevice, sm));
    const EngineParams P = make_params(opt);
    const int cfg = pick_cfg(m, n, opt, 1);
    bool auxLds = false;
    const size_t lds = cfg_group_bytes(cfg, m, n, &auxLds);
    HIP_TRY(hipEventRecord(e0, sm));
    int st = launch_queue_on(cfg, (DevInstance*)(dev + oDesc), (const int32_t*)(dev + oOrder), 1,
                             (int32_t*)(dev + oHead), P, lds, 0, sm);
    if (st) return st;
    HIP_TRY(hipEventRecord(e1, sm));
    HIP_TRY(hipMemcpyAsync(host, dev, outEnd, hipMemcpyDeviceToHost, sm));
    HIP_TRY(hipStreamSynchronize(sm));
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    const DevInstance& R = *(const DevInstance*)(host + oDesc);
    const int32_t* br = (const int32_t*)(host + oRow);
    const int32_t* bc = (const int32_t*)(host + oCol);
    const int32_t* bf = (const int32_t*)(host + oFlow);
    const int32_t* du = (const int32_t*)(host + oDual);
    res->objective = R.objective;
    res->iterations = R.iterations;
    res->status = R.status;
    res->sweeps = R.sweeps;
    res->solve_ms = ms;
    for (int k = 0; k < 6; ++k) res->phase_us[k] = (double)R.ticks[k] / g_ticks_per_us;
    res->start_us = 0;
    res->done_us = (double)(int64_t)(R.t_done - R.t_start) / g_ticks_per_us;
    if (res->basis_row) std::memcpy(res->basis_row, br, 4 * (size_t)(V - 1));
    if (res->basis_col) std::memcpy(res->basis_col, bc, 4 * (size_t)(V - 1));
    if (res->basis_flow) std::memcpy(res->basis_flow, bf, 4 * (size_t)(V - 1));
    if (res->u) std::memcpy(res->u, du, 4 * (size_t)m);
    if (res->v) std::memcpy(res->v, du + m, 4 * (size_t)n);
    int nf = 0;
    for (int s = 0; s < V - 1; ++s) {
      if (bf[s] == 0) continue;
      if (res->flows) res->flows[nf] = flock_flow{br[s], bc[s], bf[s]};
      ++nf;
    }
    res->n_flows = nf;
    return FLOCK_OK;
  }
};

}  // namespace flock

extern "C" {

const char* flock_last_error(void) { return flock::g_err.c_str(); }

int flock_init(int device) {
  int st = flock::ensure_init(device);
  // set-up of flock_solve_noalloc on this device (its scratch array's address)
  if (!st) st = flock::noalloc_setup();
  return st;
}

void flock_default_options(flock_options* opt) {
  std::memset(opt, 0, sizeof *opt);
  opt->initial = FLOCK_INIT_VOGEL;
  opt->pricing = FLOCK_PRICE_DANTZIG;
  opt->max_iterations = 0;
}

int flock_solve(const flock_instance* inst, const flock_options* opt, flock_result* res) {
  if (!res) return flock::set_error(FLOCK_E_INVALID, "null result");
  res->status = FLOCK_E_INVALID;
  int st = flock::validate_instance(inst, flock::kMaxVerticesLarge);
  if (st) { res->status = st; return st; }
  st = flock::ensure_init(-1);
  if (st) { res->status = st; return st; }
  // one reusable device arena per host thread; never freed (process lifetime)
  static thread_local flock::Batch* bp = new flock::Batch();
  flock::Batch& b = *bp;
  // large matrices: graph of sweep-grid and pivot-round kernels (needs the state buffer);
  // instances whose tree cannot live in one workgroup's LDS always take it
  const int gmin = flock::graph_min_cells();
  const bool large = (inst->m + inst->n > flock::kMaxVertices) ||
                     (gmin > 0 && (!opt || opt->threads == 0) && (int64_t)inst->m * inst->n >= gmin);
  if (!large) {  // per-instance engine: one copy in, one launch, one copy out
    static thread_local flock::PlainArena* pa = new flock::PlainArena();
    st = pa->solve(inst, opt, res);
    if (st) { res->status = st; return st; }
    return res->status;
  }
  // large matrices go over from pinned staging (one per host thread, grown as needed)
  static thread_local char* pinned = nullptr;
  static thread_local size_t pinned_cap = 0;
  const size_t need = 4 * (size_t)inst->m * (size_t)((inst->n + 3) & ~3);
  if (need > pinned_cap) {
    if (pinned) hipHostFree(pinned);
    pinned = nullptr; pinned_cap = 0;
    HIP_TRY(hipHostMalloc((void**)&pinned, need, hipHostMallocDefault));
    pinned_cap = need;
  }
  if ((st = b.layout(1, inst, large)) || (st = b.alloc()) || (st = b.upload(inst, pinned, pinned_cap))) {
    res->status = st;
    return st;
  }
  EngineParams P = flock::make_params(opt);
  static thread_local hipEvent_t e0 = nullptr, e1 = nullptr;
  if (!e0) { HIP_TRY(hipEventCreate(&e0)); HIP_TRY(hipEventCreate(&e1)); }
  HIP_TRY(hipEventRecord(e0, flock::lib_stream()));
  st = flock::large_runner()->run(b, P);
  if (st) { res->status = st; return st; }
  HIP_TRY(hipEventRecord(e1, flock::lib_stream()));
  if ((st = b.download_desc())) { res->status = st; return st; }
  float ms = 0;
  HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
  st = b.fill_result(0, res);
  res->solve_ms = ms;
  if (st) { res->status = st; return st; }
  return res->status;
}

int flock_batch_create(int32_t count, const flock_instance* inst, flock_batch** out) {
  if (!out || count < 1 || !inst) return flock::set_error(FLOCK_E_INVALID, "bad batch arguments");
  *out = nullptr;
  for (int k = 0; k < count; ++k) {
    int st = flock::validate_instance(&inst[k]);
    if (st) return flock::set_error(st, "instance %d: %s", k, flock_last_error());
  }
  int st = flock::ensure_init(-1);
  if (st) return st;
  flock::Batch* b = new flock::Batch();
  if ((st = b->layout(count, inst, false)) || (st = b->alloc()) || (st = b->upload(inst))) { delete b; return st; }
  HIP_TRY(hipStreamSynchronize(flock::lib_stream()));
  *out = reinterpret_cast<flock_batch*>(b);
  return FLOCK_OK;
}

int flock_batch_run(flock_batch* bb, const flock_options* opt) {
  flock::Batch* b = reinterpret_cast<flock::Batch*>(bb);
  if (!b) return flock::set_error(FLOCK_E_INVALID, "null batch");
  EngineParams P = flock::make_params(opt);
  int grid_cap = 0;
  int st0 = b->plan(opt);
  if (st0) return st0;
  if (!b->ev0) { HIP_TRY(hipEventCreate(&b->ev0)); HIP_TRY(hipEventCreate(&b->ev1)); }
  HIP_TRY(hipEventRecord(b->ev0, flock::lib_stream()));
  int st = b->run(P, grid_cap);
  if (st) return st;
  HIP_TRY(hipEventRecord(b->ev1, flock::lib_stream()));
  HIP_TRY(hipEventSynchronize(b->ev1));
  float ms = 0;
  HIP_TRY(hipEventElapsedTime(&ms, b->ev0, b->ev1));
  b->last_ms = ms;
  return FLOCK_OK;
}

double flock_batch_last_ms(const flock_batch* bb) {
  const flock::Batch* b = reinterpret_cast<const flock::Batch*>(bb);
  return b ? b->last_ms : -1.0;
}

int flock_batch_results(flock_batch* bb, flock_result* res) {
  flock::Batch* b = reinterpret_cast<flock::Batch*>(bb);
  if (!b || !res) return flock::set_error(FLOCK_E_INVALID, "null batch/result");
  int st = b->download_desc();
  if (st) return st;
  for (int k = 0; k < b->n; ++k) {
    st = b->fill_result(k, &res[k]);
    if (st) return st;
  }
  return FLOCK_OK;
}

void flock_batch_free(flock_batch* bb) {
  flock::Batch* b = reinterpret_cast<flock::Batch*>(bb);
  if (!b) return;
  if (b->ev0) { hipEventDestroy(b->ev0); hipEventDestroy(b->ev1); }
  delete b;
}

int flock_batch_solve(int32_t count, const flock_instance* inst, const flock_options* opt, flock_result* res) {
  flock_batch* b = nullptr;
  int st = flock_batch_create(count, inst, &b);
  if (st) return st;
  st = flock_batch_run(b, opt);
  if (!st) st = flock_batch_results(b, res);
  if (!st) for (int k = 0; k < count; ++k) res[k].solve_ms = flock_batch_last_ms(b);
  flock_batch_free(b);
  return st;
}

int flock_tuning_get(flock_tuning* t) {
  if (!t) return flock::set_error(FLOCK_E_INVALID, "null tuning");
  std::memset(t, 0, sizeof *t);
  flock::get_policy(&t->wave_max_v, &t->wave_pack, &t->wg256_max_v);
  t->minor_limit = flock::g_minor_limit;
  t->few_256_max_v = flock::g_few_256_max_v;
  t->few_512_max_v = flock::g_few_512_max_v;
  t->team_min_cells = 0;
  t->prio_permille = flock::g_prio_permille;
  t->sweep_vectors = flock::g_sweep_vectors;
  t->large_threads = flock::g_large_threads;
  t->large_minor_limit = flock::g_large_minor_limit;
  t->graph_min_cells = flock::g_graph_min_cells;
  return FLOCK_OK;
}

int flock_tuning_set(const flock_tuning* t) {
  if (!t) return flock::set_error(FLOCK_E_INVALID, "null tuning");
  if (t->wave_pack != 1 && t->wave_pack != 2 && t->wave_pack != 4)
    return flock::set_error(FLOCK_E_INVALID, "wave_pack must be 1, 2 or 4");
  if (t->wave_max_v < 0 || t->wave_max_v > 2048 || t->wg256_max_v < 0 || t->minor_limit < 1 ||
      t->few_256_max_v < 0 || t->few_512_max_v < 0 || t->team_min_cells < 0 || t->graph_min_cells < 0 ||
      t->prio_permille < 0 || t->prio_permille > 1000 || t->sweep_vectors < 1 || t->sweep_vectors > 64 ||
      (t->large_threads != 512 && t->large_threads != 1024) || t->large_minor_limit < 0)
    return flock::set_error(FLOCK_E_INVALID, "tuning values out of range");
  if (t->team_min_cells != 0)
    return flock::set_error(FLOCK_E_INVALID, "team sweeps are not part of this library (team_min_cells must be 0)");
  flock::set_policy(t->wave_max_v, t->wave_pack, t->wg256_max_v);
  flock::g_minor_limit = t->minor_limit;
  flock::g_few_256_max_v = t->few_256_max_v;
  flock::g_few_512_max_v = t->few_512_max_v;
  flock::g_graph_min_cells = t->graph_min_cells;
  flock::g_prio_permille = t->prio_permille;
  flock::g_sweep_vectors = t->sweep_vectors;
  flock::g_large_threads = t->large_threads;
  flock::g_large_minor_limit = t->large_minor_limit;
  return FLOCK_OK;
}

int flock_autotune(int32_t count, const flock_instance* inst, const flock_options* opt, flock_tuning* out,
                   double* best_ms) {
  if (!out || !best_ms) return flock::set_error(FLOCK_E_INVALID, "null output");
  flock_batch* b = nullptr;
  int st = flock_batch_create(count, inst, &b);
  if (st) return st;
  std::vector<flock_result> res(count);
  std::vector<int64_t> ref;
  flock_tuning best;
  flock_tuning_get(&best);
  const flock_tuning start = best;
  // one measurement: warm-up + timed launch; objectives must equal the first candidate's
  auto measure = [&](const flock_tuning& t, double* ms) -> int {
    int e = flock_tuning_set(&t);
    if (e) return e;
    if ((e = flock_batch_run(b, opt)) || (e = flock_batch_run(b, opt))) return e;
    *ms = flock_batch_last_ms(b);
    std::memset(res.data(), 0, sizeof(flock_result) * res.size());
    if ((e = flock_batch_results(b, res.data()))) return e;
    for (int k = 0; k < count; ++k) {
      if (res[k].status != FLOCK_OK)
        return flock::set_error(FLOCK_E_DEVICE, "autotune: instance %d status %d", k, res[k].status);
      if ((int)ref.size() < count) ref.push_back(res[k].objective);
      else if (ref[k] != res[k].objective)
        return flock::set_error(FLOCK_E_DEVICE, "autotune: objective of instance %d differs between policies", k);
    }
    return FLOCK_OK;
  };
  double bms = 0;
  st = measure(best, &bms);
  const bool candidates = opt && (opt->pricing == FLOCK_PRICE_CANDIDATES || opt->pricing == FLOCK_PRICE_AUTO);
  auto consider = [&](flock_tuning t) -> int {
    double ms = 0;
    const int e = measure(t, &ms);
    if (e) return e;
    if (ms < bms) { bms = ms; best = t; }
    return FLOCK_OK;
  };
  if (!st && candidates)
    for (int ml : {16, 24, 32, 48, 128, 256}) {
      flock_tuning t = best; t.minor_limit = ml;
      if ((st = consider(t))) break;
    }
  // instances per workgroup: the smallest instances on one wavefront each, 1, 2 or 4 per
  // workgroup (their classes launch before the workgroup classes, one after another)
  for (int wv : {64, 128, 256})
    for (int pack : {1, 2, 4}) {
      if (st) break;
      flock_tuning t = best; t.wave_max_v = wv; t.wave_pack = pack;
      st = consider(t);
    }
  if (!st)
    for (int wg : {400, 600, 1200, 8000}) {
      flock_tuning t = best; t.wg256_max_v = wg;
      if ((st = consider(t))) break;
    }
  // raised wave priority for the largest instances (the launch's critical path)
  if (!st)
    for (int pm : {10, 30, 100}) {
      flock_tuning t = best; t.prio_permille = pm;
      if ((st = consider(t))) break;
    }
  flock_batch_free(b);
  if (st) { flock_tuning_set(&start); return st; }
  flock_tuning_set(&best);
  *out = best;
  *best_ms = bms;
  return FLOCK_OK;
}

int flock_autotune_large(const flock_instance* inst, const flock_options* opt, flock_tuning* out, double* best_ms) {
  if (!out || !best_ms) return flock::set_error(FLOCK_E_INVALID, "null output");
  flock_tuning best;
  flock_tuning_get(&best);
  const flock_tuning start = best;
  int64_t ref = 0;
  bool have_ref = false;
  double bms = 1e300;
  // one measurement: warm-up + timed solve; the objective must equal the first one's
  auto measure = [&](const flock_tuning& t, double* ms) -> int {
    int e = flock_tuning_set(&t);
    if (e) return e;
    flock_result r;
    double best_run = 1e300;
    for (int rep = 0; rep < 2; ++rep) {
      std::memset(&r, 0, sizeof r);
      if ((e = flock_solve(inst, opt, &r)) != FLOCK_OK) return e;
      best_run = std::min(best_run, r.solve_ms);
    }
    if (!have_ref) { ref = r.objective; have_ref = true; }
    else if (r.objective != ref)
      return flock::set_error(FLOCK_E_DEVICE, "autotune: objective differs between large-path policies");
    *ms = best_run;
    return FLOCK_OK;
  };
  int st = FLOCK_OK;
  for (int v : {1, 2, 4, 8}) {  // vectors per thread per sweep tile
    flock_tuning t = start; t.sweep_vectors = v;
    double ms = 0;
    if ((st = measure(t, &ms))) break;
    if (ms < bms) { bms = ms; best = t; }
  }
  for (int th : {512, 1024}) {  // set-up / round workgroup, at the best tile shape
    if (st || th == best.large_threads) continue;
    flock_tuning t = best; t.large_threads = th;
    double ms = 0;
    if ((st = measure(t, &ms))) break;
    if (ms < bms) { bms = ms; best = t; }
  }
  if (opt && (opt->pricing == FLOCK_PRICE_CANDIDATES || opt->pricing == FLOCK_PRICE_AUTO))
    for (int ml : {32, 64, 128}) {  // the path's candidate minor limit
      if (st || ml == best.large_minor_limit) continue;
      flock_tuning t = best; t.large_minor_limit = ml;
      double ms = 0;
      if ((st = measure(t, &ms))) break;
      if (ms < bms) { bms = ms; best = t; }
    }
  if (st) { flock_tuning_set(&start); return st; }
  flock_tuning_set(&best);
  *out = best;
  *best_ms = bms;
  return FLOCK_OK;
}

int flock_measure_sweep(const flock_instance* inst, int32_t reps, double* sweep_ms, double* empty_ms,
                        int32_t* grid) {
  if (!sweep_ms || !empty_ms || reps < 1) return flock::set_error(FLOCK_E_INVALID, "bad measurement arguments");
  int st = flock::validate_instance(inst);
  if (st) return st;
  if ((st = flock::ensure_init(-1))) return st;
  flock::LargeRunner* lr = flock::large_runner();
  if ((st = lr->init())) return st;
  flock::Batch b;
  if ((st = b.layout(1, inst, true)) || (st = b.alloc())) return st;
  // the sweep kernel reads the duals from the saved tree state and runs while phase == 1
  b.hdesc[0].phase = 1;
  const flock::Batch::Offs& o = b.offs[0];
  HIP_TRY(hipMemsetAsync((char*)b.base + o.state, 0, 4 * 7 * (size_t)(o.m + o.n), flock::lib_stream()));
  if ((st = b.upload(inst))) return st;
  uint32_t vpt;
  int nblocks;
  flock::LargeRunner::sweep_grid(o, &vpt, &nblocks);
  if (grid) *grid = nblocks;
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  hipStream_t sm = flock::lib_stream();
  const SweepArgs A = sweep_args(b.hdesc[0], o.m + o.n, lr->keys, lr->cands, vpt, &b.d_insts[0].phase);
  auto timed = [&](bool sweep, double* out) -> int {
    for (int w = 0; w < 3; ++w) {  // warm-up
      if (sweep) hipLaunchKernelGGL(k_large_sweep, dim3(nblocks), dim3(256), 0, sm, A);
      else hipLaunchKernelGGL(k_empty_grid, dim3(nblocks), dim3(256), 0, sm);
    }
    HIP_TRY(hipEventRecord(e0, sm));
    for (int k = 0; k < reps; ++k) {
      if (sweep) hipLaunchKernelGGL(k_large_sweep, dim3(nblocks), dim3(256), 0, sm, A);
      else hipLaunchKernelGGL(k_empty_grid, dim3(nblocks), dim3(256), 0, sm);
    }
    HIP_TRY(hipEventRecord(e1, sm));
    HIP_TRY(hipEventSynchronize(e1));
    HIP_TRY(hipGetLastError());
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    *out = (double)ms / reps;
    return FLOCK_OK;
  };
  st = timed(true, sweep_ms);
  if (!st) st = timed(false, empty_ms);
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (!st) HIP_TRY(hipStreamSynchronize(sm));
  return st;
}

int flock_scale_float_costs(int64_t count, const float* costs, int32_t* out, int32_t* scale) {
  if (!costs || !out) return flock::set_error(FLOCK_E_INVALID, "null pointer");
  for (int k = 0; k <= 6; ++k) {
    const double f = std::pow(10.0, k);
    bool ok = true;
    for (int64_t i = 0; i < count && ok; ++i) {
      const double x = (double)costs[i] * f;
      const double r = std::nearbyint(x);
      if (!(std::fabs(r) < 2147483647.0)) { ok = false; break; }
      // exact when the decimal N/10^k parses back to the very same float
      ok = (float)(r / f) == costs[i];
      if (ok) out[i] = (int32_t)r;
    }
    if (ok) { if (scale) *scale = k; return FLOCK_OK; }
  }
  return flock::set_error(FLOCK_E_NONINTEGRAL, "costs have no exact decimal scale <= 10^6");
}

}  // extern "C"
