// runtime_internal.h -- host-side structures shared by runtime.hip and server.hip.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/flock_amd.h"
#include "device_common.h"

namespace flock {

constexpr int kMaxVertices = 8000;  // m+n of the per-instance engines: the tree (20 B/vertex) lives in 160 KB of LDS
// m+n of plain solves on the large-instance path, whose tree can live in global memory;
// bounded by the engine's index widths (Vogel's 14-bit line field in its argmax key)
constexpr int kMaxVerticesLarge = 16383;
constexpr int kSizeBuckets = 6;     // m+n buckets: <=128, <=256, <=512, <=1024, <=2048, more
constexpr int kMaxClasses = 40;     // (configuration, size bucket) launch classes

int set_error(int code, const char* fmt, ...);
int ensure_init(int device);
hipStream_t lib_stream();
double lib_ticks_per_us();
int lib_cus();
size_t lib_lds_max();
int validate_instance(const flock_instance* in, int max_v = kMaxVertices);
int num_cfgs();
int size_bucket(int V);
int class_bucket(int cfg, int V);
int cfg_threads(int c);
int cfg_group(int c);
int pick_cfg(int m, int n, const flock_options* opt, int64_t count);
size_t cfg_group_bytes(int cfg, int m, int n, bool* aux_in_lds);
void set_policy(int wave_max_v, int wave_pack, int wg256_max_v);
void get_policy(int* wave_max_v, int* wave_pack, int* wg256_max_v);
int launch_queue(int cfg, DevInstance* d_insts, const int32_t* d_order, int32_t count, int32_t* d_head,
                 const EngineParams& P, size_t lds, int grid_cap);
int launch_queue_on(int cfg, DevInstance* d_insts, const int32_t* d_order, int32_t count, int32_t* d_head,
                    const EngineParams& P, size_t group_bytes, int grid_cap, hipStream_t stream);
int launch_reset(DevInstance* d_insts, int32_t count, int32_t* d_heads, int32_t nheads, hipStream_t stream,
                 uint64_t* d_t0 = nullptr);
EngineParams make_params(const flock_options* opt);
int noalloc_setup();  // noalloc.hip: per-device set-up of flock_solve_noalloc
int graph_min_cells();
struct LargeRunner;
LargeRunner* large_runner();

// A set of instances resident in one device allocation.
struct Batch {
  struct Offs {
    int m, n, ld, ldT;
    size_t cost, costT, supply, demand, brow, bcol, bflow, dual, scratch, state;
    size_t books;  // sorted lines for Vogel (with state only)
  };
  struct Class {
    int cfg, begin, count;
    size_t lds;
  };
  int n = 0;
  size_t bytes = 0, cap = 0;
  void* base = nullptr;
  DevInstance* d_insts = nullptr;
  int32_t* d_order = nullptr;
  int32_t* d_heads = nullptr;
  uint64_t* d_t0 = nullptr;
  uint64_t h_t0 = 0;
  std::vector<DevInstance> hdesc;
  std::vector<Offs> offs;
  std::vector<Class> classes;
  std::vector<int32_t> sorted_order;
  std::vector<int32_t> class_order;
  int64_t planned_for = -1;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double last_ms = 0;

  int layout(int32_t count, const flock_instance* inst, bool with_state);
  int alloc();
  // pinned: optional pinned staging of at least 4*m*ld bytes for a one-instance batch
  // (the copy is asynchronous: the caller keeps it until the stream has synchronised)
  int upload(const flock_instance* inst, char* pinned = nullptr, size_t pinned_cap = 0);
  int plan(const flock_options* opt);
  int run(const EngineParams& P, int grid_cap);
  int download_desc();
  int fill_result(int k, flock_result* r);
  ~Batch();
};

}  // namespace flock
