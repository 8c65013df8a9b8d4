// engine.h -- a group of threads solves one balanced transportation instance end to end:
//   transpose -> Vogel (or north-west corner) -> basis tree in LDS (level-synchronous
//   BFS over CSR, which also yields the duals) -> loop { reduced-cost sweep fused with
//   argmin -> stepping-stone pivot with parallel cycle discovery and parallel subtree
//   re-rooting } -> basis, duals, objective.
//
// A group is either one 64-lane wavefront (G = 64: several instances per workgroup,
// lockstep execution, no workgroup barriers) or the whole workgroup (G = blockDim).
//
// Exact integer arithmetic throughout (int32 costs, flows, duals; int64 objective).
// The rules follow the reference's sequential path so that, with Dantzig pricing, the
// pivot sequence is the reference's (see the per-function citations).
//
// Tree representation (rooted at vertex 0 = row 0, the reference's BFS root,
// DUAL_solver.cu:44-46).  Vertices: rows 0..m-1, columns m..m+n-1.  For every
// non-root vertex x the basic arc (x, parent[x]) is stored at x:
//   parent[x], depth[x], slot[x] (basis slot of the arc), flow[x]
// plus a preorder numbering pre[x] and subtree size size[x], so that
//   y is an ancestor-or-self of x  <=>  pre[y] <= pre[x] < pre[y] + size[y].
// That O(1) test is what makes cycle discovery and subtree identification
// embarrassingly parallel over vertices.
#pragma once
#include "device_common.h"

namespace flock {

// ---------------------------------------------------------------------------
// groups
// ---------------------------------------------------------------------------
template <int G>
__device__ __forceinline__ int grank() { return (int)(threadIdx.x % G); }

template <int G>
__device__ __forceinline__ void gsync() {
  if constexpr (G == kWave) {
    // one wavefront executes in lockstep; this orders its LDS/global accesses for the
    // compiler and the memory model at wavefront scope
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  } else {
    __syncthreads();
  }
}

// Scalars shared by one group; they sit at the start of the group's LDS region.
struct Shared {
  uint64_t red[17];
  int64_t tk[6];      // phase tick counters (rank 0 updates them)
  int32_t cnt;
  int32_t pend_row, pend_col;  // Vogel: kinds of line the next step rescans
  int32_t qtail;
  int32_t status;
  int32_t nlevels;
  int32_t idx;
  int32_t stop;       // deadline reached (rank 0 decides, the group follows)
  int32_t pad0_[2];
  int32_t cnt2;       // Vogel: column refresh list counter (cnt: rows)
  int32_t lastx;      // Vogel: the line covered by the previous step
  int32_t dlo[2];     // pivot: least depth of a cycle vertex (atomicMin), by pivot parity
  int32_t dsel;       // pivot: parity slot of dlo used by the next pivot
  int32_t pad_[5];
  uint64_t red1[3][16];  // one-barrier reductions (group_min1 / group_max1), by buffer
};

template <int G>
__device__ __forceinline__ uint64_t group_min_u64(uint64_t x, Shared& S) {
  if constexpr (G == kWave) return wave_min_u64(x);
  else return block_min_u64<G>(x, S.red);
}
template <int G>
__device__ __forceinline__ uint64_t group_max_u64(uint64_t x, Shared& S) {
  if constexpr (G == kWave) return wave_max_u64(x);
  else return block_max_u64<G>(x, S.red);
}
// One-barrier variants for the solve loop.  Buffer use: 0 candidate re-pricing and
// Vogel's penalty argmax, 1 the pivot's leaving arc, 2 the full sweep; consecutive
// reductions on one buffer always have another barrier between them.
template <int G, int kBuf>
__device__ __forceinline__ uint64_t group_min1_u64(uint64_t x, Shared& S) {
  if constexpr (G == kWave) return wave_min_u64(x);
  else return block_min1_u64<G>(x, S.red1[kBuf]);
}
template <int G, int kBuf>
__device__ __forceinline__ uint64_t group_max1_u64(uint64_t x, Shared& S) {
  if constexpr (G == kWave) return wave_max_u64(x);
  else return block_max1_u64<G>(x, S.red1[kBuf]);
}
template <int G>
__device__ __forceinline__ int64_t group_sum_i64(int64_t x, Shared& S) {
  if constexpr (G == kWave) return wave_sum_i64(x);
  else return block_sum_i64<G>(x, S.red);
}

// ---------------------------------------------------------------------------
// LDS layout of one group.
//   [Shared][tree: parent depth (int16) | pre,size interleaved (int16 pairs) | slot vat (int16)
//    | flow (int32) | du | dv]
//   (vat[p]: the vertex at pre-order position p)
//   [aux: tree-build scratch, later the pivot's path snapshot]
// The Vogel state aliases tree+aux (Vogel finishes before the tree is built).  When
// tree+aux does not fit, aux lives in the instance's global scratch instead.
// ---------------------------------------------------------------------------
struct TreeLds {
  int16_t *parent, *depth, *ps, *slot, *vat;  // ps: pre-order number and subtree size, interleaved
  int32_t *flow, *du, *dv;
  __device__ __forceinline__ int16_t& pre(int x) const { return ps[2 * x]; }
  __device__ __forceinline__ int16_t& size(int x) const { return ps[2 * x + 1]; }
  // both in one 32-bit load: pre | size << 16
  __device__ __forceinline__ int32_t pre_size(int x) const {
    int32_t v;
    __builtin_memcpy(&v, ps + 2 * x, 4);
    return v;
  }
};

__host__ __device__ inline size_t align16(size_t x) { return (x + 15) & ~(size_t)15; }

__host__ __device__ inline size_t shared_bytes() { return align16(sizeof(Shared)); }

__host__ __device__ inline size_t tree_lds_bytes(int m, int n) {
  const int V = m + n;
  return 4 * align16(2 * (size_t)V) + 2 * align16(4 * (size_t)V) + align16(4 * (size_t)((m + 3) & ~3)) +
         align16(4 * (size_t)((n + 3) & ~3));
}
// Vogel state without its tracked keys: cover flags, residuals, penalties, scan list,
// tracked count/flag, book position
__host__ __device__ inline size_t vogel_lds_bytes(int m, int n);
// the register mode keeps no per-line penalty array (the penalties live in the ranks'
// argmax keys), so its line state is one V-word array smaller
__host__ __device__ inline size_t vogel_reg_lds_bytes(int m, int n) {
  return vogel_lds_bytes(m, n) - align16(4 * (size_t)(m + n));
}
__host__ __device__ inline size_t vogel_lds_bytes(int m, int n) {
  const int V = m + n;
  return align16((size_t)((m + 3) & ~3)) + align16((size_t)((n + 3) & ~3)) + 2 * align16(4 * (size_t)V) +
         2 * align16(2 * (size_t)V) + align16((size_t)V);
}
// aux region in int32 units: max(tree-build scratch, path snapshot)
//   build: start[V+1], vis[V], lvl[V+2], adj[2V] (int16), queue[V] (int16)
//   snapshot: for each side of the cycle, vertex, slot (int16), pre | size << 16 and flow (int32)
//   of its vertices by depth, V entries each
__host__ __device__ inline size_t aux_ints(int m, int n) {
  const size_t V = (size_t)(m + n);
  const size_t build = (V + 1) + V + (V + 2) + V + (V + 1) / 2 + 8;
  const size_t snap = 2 * (2 * ((V + 1) / 2 + 4) + 2 * (V + 4));
  return build > snap ? build : snap;
}
// bytes of one group's LDS region (aux in LDS if that fits `cap`, else in global)
__host__ __device__ inline size_t group_lds_bytes(int m, int n, size_t cap, bool* aux_in_lds) {
  const size_t base = shared_bytes();
  const size_t tree = tree_lds_bytes(m, n);
  const size_t vog = vogel_lds_bytes(m, n);
  const size_t tile = 64 * 65 * 4;
  size_t full = base + tree + 4 * aux_ints(m, n);
  full = full > base + vog ? full : base + vog;
  full = full > base + tile ? full : base + tile;
  if (full <= cap) { if (aux_in_lds) *aux_in_lds = true; return full; }
  if (aux_in_lds) *aux_in_lds = false;
  size_t part = base + tree;
  part = part > base + vog ? part : base + vog;
  part = part > base + tile ? part : base + tile;
  return part;
}

__device__ inline TreeLds carve_tree(char* base, int m, int n) {
  const int V = m + n;
  TreeLds t;
  size_t off = 0;
  t.parent = (int16_t*)(base + off); off += align16(2 * (size_t)V);
  t.depth = (int16_t*)(base + off); off += align16(2 * (size_t)V);
  t.ps = (int16_t*)(base + off); off += align16(4 * (size_t)V);
  t.slot = (int16_t*)(base + off); off += align16(2 * (size_t)V);
  t.vat = (int16_t*)(base + off); off += align16(2 * (size_t)V);
  t.flow = (int32_t*)(base + off); off += align16(4 * (size_t)V);
  t.du = (int32_t*)(base + off); off += align16(4 * (size_t)((m + 3) & ~3));
  t.dv = (int32_t*)(base + off);
  return t;
}

// Cycle snapshot used by the pivot (generic pointers: LDS or global aux): side 0 holds
// the cycle vertices that are ancestors of the entering row, side 1 those of the
// entering column, each indexed by depth.
// Addresses are computed from the side (no runtime-indexed pointer arrays, which would
// live in scratch memory).
template <class I32, class I16>
struct SnapT {
  I32* base;
  int h;     // int32 words per int16 array
  int w;     // int32 words per int32 array
  int side;  // int32 words per side
  __device__ __forceinline__ I16* vert(int k) const { return (I16*)(base + k * side); }
  __device__ __forceinline__ I16* slot(int k) const { return (I16*)(base + k * side + h); }
  // pre-order number | subtree size << 16 (one load per interval)
  __device__ __forceinline__ I32* ps(int k) const { return base + k * side + 2 * h; }
  __device__ __forceinline__ I32* flow(int k) const { return base + k * side + 2 * h + w; }
};
__device__ __forceinline__ int32_t pack_ps(int pre, int size) { return (int32_t)((uint32_t)(uint16_t)pre | ((uint32_t)size << 16)); }
__device__ __forceinline__ int ps_pre(int32_t v) { return (int)((uint32_t)v & 0xffffu); }
__device__ __forceinline__ int ps_size(int32_t v) { return (int)((uint32_t)v >> 16); }
// generic pointers (aux in LDS or in global scratch) and LDS-typed pointers (aux known
// to be in LDS: ds_read/ds_write instead of flat accesses)
typedef SnapT<int32_t, int16_t> Snap;
typedef SnapT<FLOCK_LDS int32_t, FLOCK_LDS int16_t> SnapL;
template <class SN>
__device__ inline SN carve_snap(int32_t* aux, int V) {
  SN s;
  s.base = (decltype(s.base))aux;
  s.h = (V + 1) / 2 + 4;
  s.w = V + 4;
  s.side = 2 * s.h + 2 * s.w;
  return s;
}

typedef const FLOCK_GLOBAL int32_t gcint;
typedef FLOCK_GLOBAL int32_t gint;

struct Ctx {
  gcint* cost;
  gint* costT;
  int32_t m, n, V, ld, ldT;
};

__device__ __forceinline__ int32_t cost_at(const Ctx& C, int i, int j) { return C.cost[(uint32_t)(i * C.ld + j)]; }

// ---------------------------------------------------------------------------
// transpose: costT[j*ldT + i] = cost[i*ld + j], 64x64 tiles staged through LDS
// (reads and writes both coalesced).  Uses 64*65*4 bytes at `tile`.  Each rank's loads
// of a tile are issued together (clamped addresses, unconditional), so a tile costs one
// global round trip instead of one per row.  The tiles are software-pipelined: the
// loads of tile k+1 are issued before the stores of tile k, so they are not queued
// behind those stores (the vector memory counter counts loads and stores in issue
// order) and their latency overlaps tile k's LDS reads and stores.
// ---------------------------------------------------------------------------
template <int G>
__device__ void transpose(const Ctx& C, int32_t* tile) {
  const int rk = grank<G>();
  const int tx = rk % 64, ty = rk / 64;
  constexpr int R = G / 64;                // rows of a tile per pass of the group
  constexpr int PER = 64 / R;              // tile rows per rank
  constexpr int CH = PER < 16 ? PER : 16;  // loads in flight per rank (one chunk)
  constexpr int NCH = PER / CH;            // chunks per tile
  const int mLast = C.m - 1, nLast = C.n - 1;
  const int tilesJ = (C.n + 63) >> 6, units = ((C.m + 63) >> 6) * tilesJ * NCH;
  int32_t v[CH];
  auto load_chunk = [&](int q) {  // chunk q % NCH of tile q / NCH
    const int k = q / NCH, c = q % NCH;
    const int bi = (k / tilesJ) << 6, bj = (k % tilesJ) << 6;
    const int j = bj + tx < nLast ? bj + tx : nLast;
#pragma unroll
    for (int u = 0; u < CH; ++u) {
      const int i = bi + ty + (c * CH + u) * R;
      v[u] = C.cost[(uint32_t)((i < mLast ? i : mLast) * C.ld + j)];
    }
  };
  auto unit = [&](int q) {
    const int k = q / NCH, c = q % NCH;
#pragma unroll
    for (int u = 0; u < CH; ++u) tile[(ty + (c * CH + u) * R) * 65 + tx] = v[u];
    if (c < NCH - 1) {  // uniform: more rows of this tile to load first
      load_chunk(q + 1);
      return;
    }
    gsync<G>();
    load_chunk(q + 1 < units ? q + 1 : q);  // unconditional (the last one reloads), so the wait is exact
    const int bi = (k / tilesJ) << 6, bj = (k % tilesJ) << 6;
    // Ranks past the matrix's last row or column store the edge element (the value the
    // edge rank stores at the same address), so every store is unconditional and the
    // compiler can wait for the next chunk's loads without waiting for these stores.
    const int i = bi + tx < mLast ? bi + tx : mLast;
    for (int c2 = 0; c2 < NCH; ++c2) {
      int32_t w[CH];
      int32_t jc[CH];
#pragma unroll
      for (int u = 0; u < CH; ++u) {
        const int jj = bj + ty + (c2 * CH + u) * R;
        jc[u] = jj < nLast ? jj : nLast;
        w[u] = tile[(i - bi) * 65 + (jc[u] - bj)];
      }
#pragma unroll
      for (int u = 0; u < CH; ++u) C.costT[(uint32_t)(jc[u] * C.ldT + i)] = w[u];
    }
    gsync<G>();
  };
  // The first unit is peeled, so every entry to the loop has the same memory operations
  // in flight (next loads, then this unit's stores) and the wait for the loads is exact.
  load_chunk(0);
  unit(0);
  for (int q = 1; q < units; ++q) unit(q);
}

// ---------------------------------------------------------------------------
// Vogel's approximation (restates IBFS_vogel.cu:546-769, the host variant):
//  * penalty of a line = 2nd smallest - smallest uncovered cost (kBig sentinels,
//    :588-606); rows are refreshed only after a row was eliminated and columns only
//    after a column was eliminated (prev_row/prev_col, :588-642), so one kind can be
//    stale -- restated as is;
//  * argmax over [rows | columns] keeps the LAST maximum (>=, :654);
//  * the min-cost cell of the chosen line is the FIRST minimum (<, :673, :723);
//  * res_d > res_s eliminates the row, otherwise the column (:690-703), except that
//    the last uncovered column is never eliminated (the reference is undefined there).
// Parallel form.  Every line tracks its kTrack smallest uncovered keys (cost << b |
// cross index, b = index bits of the cross kind), ascending, found by one wavefront
// scanning the line (or reading its sorted book).  Covering a line only removes keys,
// so the uncovered tracked keys stay the smallest uncovered keys of the line: the two
// smallest (the penalty) and the smallest (the chosen cell) come from LDS alone until
// fewer than two tracked keys are left of a line that had more cells, which is then
// scanned again.  A step is one pass in which every rank refreshes the penalties of
// its own lines (pending kind) and forms their argmax keys -- (penalty + 1, line,
// first uncovered tracked cell) -- one group max, and rank 0 eliminating; lines that
// need a scan raise a flag bit in their key, and the step then scans them and takes
// the max again.
// ---------------------------------------------------------------------------
constexpr int kTrack = 4;
constexpr uint8_t kTrackFull = 0x80;  // nt flag: the tracked keys were all of the line's uncovered cells

__device__ __forceinline__ int index_bits(int len) {
  int b = 0;
  while ((1 << b) < len) ++b;
  return b;
}
// penalty from the two smallest keys (~0u: absent, worth kBig)
__device__ __forceinline__ int32_t key_penalty(uint32_t k1, uint32_t k2, int b) {
  const int64_t v1 = k1 == ~0u ? (int64_t)kBig : (int64_t)(k1 >> b);
  const int64_t v2 = k2 == ~0u ? (int64_t)kBig : (int64_t)(k2 >> b);
  return (int32_t)(v2 - v1);
}

__device__ __forceinline__ void track_insert(uint32_t k, uint32_t* l) {
  const bool c0 = k < l[0], c1 = k < l[1], c2 = k < l[2], c3 = k < l[3];
  l[3] = c2 ? l[2] : (c3 ? k : l[3]);
  l[2] = c1 ? l[1] : (c2 ? k : l[2]);
  l[1] = c0 ? l[0] : (c1 ? k : l[1]);
  l[0] = c0 ? k : l[0];
}

// Line scan by one wavefront: the kTrack smallest uncovered keys of the line in
// ascending order into out[0..kTrack) (lane 0 writes), their count and full flag into
// *nt; returns the two smallest in k1/k2 (~0u when absent), uniform over the wave.
// 16-byte loads (the line starts 16-byte aligned and is padded to a multiple of 4),
// four cover flags read as one word, U vectors in flight per lane.
template <class TRK>
__device__ __forceinline__ void wave_line_track(gcint* data, int len, const uint8_t* crossCov, int lane, int b,
                                                TRK* out, uint8_t* nt, uint32_t& k1, uint32_t& k2) {
  uint32_t l[kTrack] = {~0u, ~0u, ~0u, ~0u};
  uint32_t cnt = 0;
  constexpr int U = 4;
  if (len <= U * kWave) {  // short line: one scalar load per lane and slot
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int j = lane + u * kWave;
      const int32_t c = data[j < len ? j : 0];
      const bool live = j < len && !crossCov[j < len ? j : 0];
      track_insert(live ? (((uint32_t)c << b) | (uint32_t)j) : ~0u, l);  // ~0u inserts nothing
      cnt += live ? 1u : 0u;
    }
  }
  const int nvec = len > U * kWave ? (len + 3) >> 2 : 0;
  for (int v0 = lane; v0 < nvec; v0 += U * kWave) {
    i32x4 c[U];
    uint32_t f[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int v = v0 + u * kWave;
      const int vv = v < nvec ? v : 0;
      c[u] = *(const FLOCK_GLOBAL i32x4*)(data + 4 * vv);
      f[u] = *(const uint32_t*)(crossCov + 4 * vv);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int v = v0 + u * kWave;
      const int j0 = v < nvec ? 4 * v : len;  // out-of-range vectors never count
      const int32_t cc[4] = {c[u].x, c[u].y, c[u].z, c[u].w};
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int j = j0 + e;
        const bool live = j < len && ((f[u] >> (8 * e)) & 0xffu) == 0;
        track_insert(live ? (((uint32_t)cc[e] << b) | (uint32_t)j) : ~0u, l);
        cnt += live ? 1u : 0u;
      }
    }
  }
  const uint32_t total = wave_sum_u32(cnt);
  // the kTrack smallest of the wave: every lane holds its own kTrack smallest, so
  // kTrack rounds of (wave min, the owner pops it) find them in order (keys are unique)
  uint32_t got[kTrack];
#pragma unroll
  for (int e = 0; e < kTrack; ++e) {
    const uint32_t mn = wave_min_u32(l[0]);
    got[e] = mn;
    if (l[0] == mn && mn != ~0u) { l[0] = l[1]; l[1] = l[2]; l[2] = l[3]; l[3] = ~0u; }
  }
  if (lane == 0) {
#pragma unroll
    for (int e = 0; e < kTrack; ++e) out[e] = got[e];
    *nt = (uint8_t)((total < (uint32_t)kTrack ? total : (uint32_t)kTrack) | (total <= (uint32_t)kTrack ? kTrackFull : 0));
  }
  k1 = got[0]; k2 = got[1];
}

// ---------------------------------------------------------------------------
// Sorted lines ("books") for Vogel on large instances: line L's cross indices in
// increasing (cost, index) order -- row L at books + L*n, column q at books + m*n + q*m
// (int16).  The smallest uncovered keys of a line are then its first uncovered book
// entries, and entries before the first one stay covered for good, so each line
// keeps a read position that only moves forward.  One workgroup sorts one line: keys
// cost << b | index (b = index bits; cost * 2^b < 2^31 by the range contract) in LDS,
// bitonic.
// ---------------------------------------------------------------------------
typedef const FLOCK_GLOBAL int16_t gcshort;

template <int G>
__device__ void line_book(const Ctx& C, int L, uint32_t* keys, FLOCK_GLOBAL int16_t* books) {
  const int m = C.m, n = C.n, tid = grank<G>();
  const bool row = L < m;
  gcint* data = row ? C.cost + (uint32_t)(L * C.ld) : C.costT + (uint32_t)((L - m) * C.ldT);
  FLOCK_GLOBAL int16_t* book = books + (row ? (size_t)L * n : (size_t)m * n + (size_t)(L - m) * m);
  const int len = row ? n : m;
  const int b = index_bits(len);
  const int npow = 1 << b;
  for (int i = tid; i < npow; i += G) keys[i] = i < len ? ((uint32_t)data[i] << b) | (uint32_t)i : 0xffffffffu;
  gsync<G>();
  for (int k = 2; k <= npow; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = tid; i < npow; i += G) {
        const int p = i ^ j;
        if (p > i) {
          const uint32_t a = keys[i], c = keys[p];
          if ((a > c) == ((i & k) == 0)) { keys[i] = c; keys[p] = a; }
        }
      }
      gsync<G>();
    }
  const uint32_t mask = (1u << b) - 1u;
  for (int i = tid; i < len; i += G) book[i] = (int16_t)(keys[i] & mask);
}

// Book read by one wavefront: the first kTrack uncovered entries from the read position
// *pos on (64 entries per ballot), as keys, with count and full flag; *pos moves to the
// first uncovered entry.  k1/k2 as for wave_line_track.
template <class TRK>
__device__ __forceinline__ void wave_book_track(gcshort* book, gcint* data, int len, int16_t* pos,
                                                const uint8_t* crossCov, int lane, int b, TRK* out, uint8_t* nt,
                                                uint32_t& k1, uint32_t& k2) {
  int at = *pos, found = 0, first = -1;
  uint32_t r0 = ~0u, r1 = ~0u;  // this lane's key if it holds rank 0 / rank 1
  while (at < len && found <= kTrack) {
    const int q = at + lane;
    const int idx = q < len ? book[q] : 0;
    // every lane loads its entry's cost with the entry's cover flag (one global round trip
    // after the book read, instead of a second one after the ballot)
    const int32_t cst = data[idx];
    const bool live = q < len && !crossCov[idx];
    const uint64_t mask = __ballot(live);
    if (first < 0 && mask) first = at + __ffsll((unsigned long long)mask) - 1;
    if (live) {
      const int rank = found + __popcll(mask & ((1ull << lane) - 1ull));
      const uint32_t key = ((uint32_t)cst << b) | (uint32_t)idx;
      if (rank < kTrack) out[rank] = key;
      if (rank == 0) r0 = key;
      if (rank == 1) r1 = key;
    }
    found += __popcll(mask);
    at += kWave;
  }
  if (lane == 0) {
    *pos = (int16_t)(first >= 0 ? first : len);
    *nt = (uint8_t)((found < kTrack ? found : kTrack) | (found <= kTrack ? kTrackFull : 0));
  }
  k1 = wave_min_u32(r0);
  k2 = wave_min_u32(r1);
}

// Vogel's LDS state (aliases the tree and aux regions): cover flags, residuals,
// penalties, the scan list, per-line tracked count/flag and book position; the tracked
// keys (kTrack per line, 16 bytes) follow in LDS when they fit, else they live in the
// instance's global scratch.
__host__ __device__ inline size_t vogel_trk_bytes(int m, int n) { return align16(16 * (size_t)(m + n)); }

// KL > 0 (register mode): each rank keeps its (at most KL) lines' tracked keys, counts,
// dirty flags and argmax keys in registers, so a step touches LDS only for the few lines
// whose tracked keys met the line covered last (dirty lines) and for scans; the tracked
// keys in `trk` are read only after the first pass and after scans, so they may live in
// global memory.  KL = 0: every step walks every line's state in LDS.
constexpr int kRegLines = 4;       // lines per rank of the work-queue engines
constexpr int kRegLinesWide = 8;   // lines per rank of the large-instance set-up kernel
template <int G, class TRK, int KL>
__device__ void vogel_run(const Ctx& C, gcint* supply, gcint* demand, char* base, TRK* trk, Shared& S, gint* brow,
                          gint* bcol, gint* bflow, gcshort* books) {
  const int m = C.m, n = C.n, V = C.V, tid = grank<G>();
  const int lane = tid % kWave, wave = tid / kWave;
  constexpr int W = G / kWave;
  constexpr bool kReg = KL > 0;
  constexpr int kRL = KL > 0 ? KL : 1;  // register slots (one unused slot when KL = 0)
  size_t off = 0;
  uint8_t* covR = (uint8_t*)(base + off); off += align16((size_t)((m + 3) & ~3));
  uint8_t* covC = (uint8_t*)(base + off); off += align16((size_t)((n + 3) & ~3));
  int32_t* res = (int32_t*)(base + off); off += align16(4 * (size_t)V);
  // penalties are kept in LDS by the scan mode only (the register mode's are in the keys)
  int32_t* pen = kReg ? nullptr : (int32_t*)(base + off);
  if (!kReg) off += align16(4 * (size_t)V);
  int16_t* list = (int16_t*)(base + off); off += align16(2 * (size_t)V);
  uint8_t* nt = (uint8_t*)(base + off); off += align16((size_t)V);
  int16_t* bpos = (int16_t*)(base + off);  // book read positions
  const int bRow = index_bits(n), bCol = index_bits(m);  // rows track column indices and vice versa

  for (int L = tid; L < V; L += G) {
    res[L] = L < m ? supply[L] : demand[L - m];
    if constexpr (!kReg) pen[L] = 0;
    nt[L] = 0; bpos[L] = 0;  // with books nothing is tracked: every line is read at step 0
    if (L < m) covR[L] = 0; else covC[L - m] = 0;
  }
  if (!books) {
    // First pass without books: the kTrack smallest keys of every line, one rank per
    // line, read from the opposite-major copy (row L is column L of the transpose), so
    // the lanes of a wavefront read consecutive addresses and each rank keeps its own
    // top keys with no cross-lane reduction.  Every cell is uncovered at step 0.
    constexpr int U = 16;  // loads in flight per rank
    for (int L = tid; L < V; L += G) {
      const bool row = L < m;
      gcint* lbase = row ? (gcint*)C.costT + L : C.cost + (L - m);
      const uint32_t stride = (uint32_t)(row ? C.ldT : C.ld);
      const int len = row ? n : m;
      const int b = row ? bRow : bCol;
      uint32_t l[kTrack] = {~0u, ~0u, ~0u, ~0u};
      for (int k0 = 0; k0 < len; k0 += U) {
        int32_t c[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int k = k0 + u < len ? k0 + u : len - 1;
          c[u] = lbase[(uint32_t)k * stride];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int k = k0 + u;
          track_insert(k < len ? (((uint32_t)c[u] << b) | (uint32_t)k) : ~0u, l);  // ~0u inserts nothing
        }
      }
#pragma unroll
      for (int e = 0; e < kTrack; ++e) trk[4 * L + e] = l[e];
      nt[L] = (uint8_t)((len < kTrack ? len : kTrack) | (len <= kTrack ? kTrackFull : 0));
    }
  }
  if (tid == 0) { S.pend_row = 1; S.pend_col = 1; S.cnt = 0; S.lastx = -1; }
  int covCols = 0;  // columns covered so far (only rank 0 eliminates, so only its copy counts)
  // register state of this rank's lines L = tid + u*G (kReg): tracked keys, argmax key,
  // and one word of flags -- tracked count (bits 0-2), full (3), dirty (4), covered (5),
  // listed for a scan (6)
  constexpr uint32_t fFull = 8, fDirty = 16, fCov = 32, fNeed = 64;
  uint32_t rk[kRL][4];
  uint32_t rf[kRL];
  uint64_t rkey[kRL];
  if constexpr (kReg) {
#pragma unroll
    for (int u = 0; u < kRL; ++u) {
      const int L = tid + u * G;
      rk[u][0] = rk[u][1] = rk[u][2] = rk[u][3] = ~0u;
      rf[u] = fDirty | (L >= V ? fCov : 0u);
      rkey[u] = (uint64_t)(uint32_t)L << 18;
      if (!books && L < V) {  // the first pass tracked this rank's line: clean, with its key
        const u32x4 tk = *(const u32x4*)(trk + 4 * L);
        const uint8_t ntl = nt[L];
        const int cnt = ntl & 0x7f;
        const int b = L < m ? bRow : bCol;
        rk[u][0] = tk.x; rk[u][1] = tk.y; rk[u][2] = tk.z; rk[u][3] = tk.w;
        rf[u] = (uint32_t)cnt | ((ntl & kTrackFull) ? fFull : 0u);
        const uint32_t k1 = cnt > 0 ? tk.x : ~0u, k2 = cnt > 1 ? tk.y : ~0u;
        rkey[u] = ((uint64_t)(uint32_t)(key_penalty(k1, k2, b) + 1) << 32) | ((uint64_t)(uint32_t)L << 18) |
                  (k1 != ~0u ? ((1u << 17) | (k1 & ((1u << b) - 1u))) : 0u);
      }
    }
  }
  gsync<G>();

  for (int step = 0; step < V - 1; ++step) {
    const bool pendR = S.pend_row != 0, pendC = S.pend_col != 0;
    uint64_t best = 0;
    if constexpr (kReg) {
      // the line covered by the previous step: its own rank marks it, ranks whose lines
      // tracked it mark those dirty and move their first uncovered cell on if it was it
      const int X = S.lastx;
      const bool xRow = X >= 0 && X < m;
      const uint32_t xIdx = X < 0 ? 0xffffu : (uint32_t)(xRow ? X : X - m);
      // the rank index, opaque to the compiler here, so line numbers and the values
      // derived from them are recomputed each step instead of being kept (spilled)
      // across the whole Vogel loop
      int tidr = tid;
      asm volatile("" : "+v"(tidr));
      const int wave0 = uniform(tid & ~(kWave - 1));  // this wavefront's first rank
#pragma unroll
      for (int u = 0; u < kRL; ++u) {
        if (wave0 + u * G >= V) continue;  // uniform: the wavefront has no line in slot u
        const int L = tidr + u * G;
        const bool row = L < m;
        const int b = row ? bRow : bCol;
        const uint32_t mask = (1u << b) - 1u;
        const bool isX = L == X;
        rf[u] |= isX ? fCov : 0u;
        rkey[u] = isX ? ((uint64_t)(uint32_t)L << 18) : rkey[u];  // penalty -1
        const bool live = (rf[u] & fCov) == 0;
        const int cnt = (int)(rf[u] & 7u);
        bool hit = false;
#pragma unroll
        for (int e = 0; e < 4; ++e) hit = hit | ((e < cnt) & ((rk[u][e] & mask) == xIdx));
        hit = hit & live & (X >= 0) & (row != xRow);
        rf[u] |= hit ? fDirty : 0u;
        const bool refreshL = live & ((rf[u] & fDirty) != 0) & (row ? pendR : pendC);
        const bool jcHit = hit & (((uint32_t)rkey[u] & 0x3ffffu) == ((1u << 17) | xIdx));
        if (refreshL | jcHit) {  // few lines per step: walk the tracked keys against the flags
          const uint8_t* cross = row ? covC : covR;
          bool unc[4];
#pragma unroll
          for (int e = 0; e < 4; ++e) unc[e] = (e < cnt) & (cross[e < cnt ? (rk[u][e] & mask) : 0] == 0);
          const uint32_t* kk = rk[u];
          const uint32_t k1 = unc[0] ? kk[0] : unc[1] ? kk[1] : unc[2] ? kk[2] : unc[3] ? kk[3] : ~0u;
          const uint32_t a3 = unc[3] ? kk[3] : ~0u;
          const uint32_t a2 = unc[2] ? kk[2] : a3;
          const uint32_t a1 = unc[1] ? kk[1] : a2;
          const uint32_t k2 = unc[0] ? a1 : unc[1] ? a2 : unc[2] ? a3 : ~0u;
          const uint32_t jbits = k1 == ~0u ? 0u : ((1u << 17) | (k1 & mask));
          if (refreshL) {
            if ((k2 != ~0u) | ((rf[u] & fFull) != 0)) {
              rkey[u] = ((uint64_t)(uint32_t)(key_penalty(k1, k2, b) + 1) << 32) | ((uint64_t)(uint32_t)L << 18) | jbits;
              rf[u] &= ~fDirty;
            } else {
              rf[u] |= fNeed;
              list[atomicAdd(&S.cnt, 1)] = (int16_t)L;
            }
          } else {  // penalty stays (stale kind); the first uncovered cell moves on
            rkey[u] = (rkey[u] & ~0x3ffffull) | jbits;
          }
        }
        best = ((rf[u] & fNeed) == 0 && L < V && rkey[u] > best) ? rkey[u] : best;
      }
    } else {
      constexpr bool refresh = true;
      // kVB lines per rank at a time, every LDS load of a batch issued before first use
      constexpr int kVB = 4;
      for (int L0 = tid; L0 < V; L0 += kVB * G) {
        uint32_t key[kVB][4];
        uint8_t ntv[kVB], cov[kVB], fl[kVB][4];
        int32_t pv[kVB];
#pragma unroll
        for (int u = 0; u < kVB; ++u) {  // lines past V re-read this rank's first line L0
          const int L = L0 + u * G;
          const bool in = L < V;
          const int LL = in ? L : L0;
          ntv[u] = nt[LL];
          const u32x4 tk = *(const u32x4*)(trk + 4 * LL);
          key[u][0] = tk.x; key[u][1] = tk.y; key[u][2] = tk.z; key[u][3] = tk.w;
          const uint8_t cv = LL < m ? covR[LL] : covC[LL - m];
          cov[u] = in ? cv : 1;
          pv[u] = pen[LL];
        }
#pragma unroll
        for (int u = 0; u < kVB; ++u) {
          const int L = L0 + u * G;
          const bool row = L < m;
          const uint8_t* cross = row ? covC : covR;
          const uint32_t mask = (1u << (row ? bRow : bCol)) - 1u;
          const int cnt = ntv[u] & 0x7f;
#pragma unroll
          for (int e = 0; e < 4; ++e) fl[u][e] = cross[e < cnt ? (key[u][e] & mask) : 0];
        }
        // branch-free except the scan list append: the two smallest uncovered tracked
        // keys by selects, every in-range line's penalty rewritten
#pragma unroll
        for (int u = 0; u < kVB; ++u) {
          const int L = L0 + u * G;
          const bool row = L < m;
          const int b = row ? bRow : bCol;
          const uint32_t mask = (1u << b) - 1u;
          const int cnt = ntv[u] & 0x7f;
          bool unc[4];
#pragma unroll
          for (int e = 0; e < 4; ++e) unc[e] = (e < cnt) & (fl[u][e] == 0);
          const uint32_t* kk = key[u];
          const uint32_t k1 = unc[0] ? kk[0] : unc[1] ? kk[1] : unc[2] ? kk[2] : unc[3] ? kk[3] : ~0u;
          const uint32_t a3 = unc[3] ? kk[3] : ~0u;
          const uint32_t a2 = unc[2] ? kk[2] : a3;
          const uint32_t a1 = unc[1] ? kk[1] : a2;
          const uint32_t k2 = unc[0] ? a1 : unc[1] ? a2 : unc[2] ? a3 : ~0u;
          const bool refreshL = refresh & !cov[u] & (row ? pendR : pendC);
          const bool resolved = (k2 != ~0u) | ((ntv[u] & kTrackFull) != 0);
          const bool need = refreshL & !resolved;
          const int32_t p = refreshL ? (resolved ? key_penalty(k1, k2, b) : 0) : pv[u];
          if (L < V) pen[L] = p;
          if (need) list[atomicAdd(&S.cnt, 1)] = (int16_t)L;
          const bool jv = k1 != ~0u;
          const uint64_t k = ((uint64_t)(uint32_t)(p + 1) << 32) | ((uint64_t)(uint32_t)L << 18) |
                             (jv ? ((1u << 17) | (k1 & mask)) : 0u);
          best = (L < V && !need && k > best) ? k : best;
        }
      }
    }
    {
      best = group_max1_u64<G, 0>(best, S);
      // lines to scan (the count was published by the atomics before the max's barrier);
      // uniform: every rank reads the same count
      const int cnt = uniform(S.cnt);
      if (cnt == 0) goto decided;
      // one wavefront per listed line; each wavefront's max key of its lines is combined
      // with the max above through buffer 1 (idle outside pivots)
      uint64_t wbest = 0;
      for (int e = wave; e < cnt; e += W) {
        const int L = list[e];
        if (L < 0 || L >= V) continue;  // guard: list entries are lines
        const bool row = L < m;
        gcint* data = row ? C.cost + (uint32_t)(L * C.ld) : C.costT + (uint32_t)((L - m) * C.ldT);
        const int len = row ? n : m;
        const int b = row ? bRow : bCol;
        const uint8_t* cross = row ? covC : covR;
        uint32_t k1, k2;
        if (books)
          wave_book_track(books + (row ? (size_t)L * n : (size_t)m * n + (size_t)(L - m) * m), data, len, &bpos[L],
                          cross, lane, b, trk + 4 * L, &nt[L], k1, k2);
        else
          wave_line_track(data, len, cross, lane, b, trk + 4 * L, &nt[L], k1, k2);
        const int32_t p = key_penalty(k1, k2, b);
        if constexpr (!kReg) {
          if (lane == 0) pen[L] = p;
        }
        const uint64_t k = ((uint64_t)(uint32_t)(p + 1) << 32) | ((uint64_t)(uint32_t)L << 18) |
                           (k1 != ~0u ? ((1u << 17) | (k1 & ((1u << b) - 1u))) : 0u);
        wbest = k > wbest ? k : wbest;
      }
      if (lane == 0) S.red1[1][wave] = wbest;
      gsync<G>();
#pragma unroll
      for (int w = 0; w < W; ++w) best = S.red1[1][w] > best ? S.red1[1][w] : best;
      if (tid == 0) S.cnt = 0;  // read above, before the barrier; next used after this step's last one
      if constexpr (kReg) {  // scanned lines: fresh tracked keys (all uncovered) back to registers
        int tidr = tid;  // opaque, as in the refresh above
        asm volatile("" : "+v"(tidr));
#pragma unroll
        for (int u = 0; u < kRL; ++u) {
          if ((rf[u] & fNeed) == 0) continue;
          const int L = tidr + u * G;
          const bool row = L < m;
          const int b = row ? bRow : bCol;
          const u32x4 tk = *(const u32x4*)(trk + 4 * L);
          const uint8_t ntl = nt[L];
          rk[u][0] = tk.x; rk[u][1] = tk.y; rk[u][2] = tk.z; rk[u][3] = tk.w;
          const int cnt = ntl & 0x7f;
          rf[u] = (uint32_t)cnt | ((ntl & kTrackFull) ? fFull : 0u);  // clean, not covered, not listed
          const uint32_t k1 = cnt > 0 ? tk.x : ~0u, k2 = cnt > 1 ? tk.y : ~0u;
          rkey[u] = ((uint64_t)(uint32_t)(key_penalty(k1, k2, b) + 1) << 32) | ((uint64_t)(uint32_t)L << 18) |
                    (k1 != ~0u ? ((1u << 17) | (k1 & ((1u << b) - 1u))) : 0u);
        }
      }
    }
  decided:
    const int Lc = (int)((best >> 18) & 0x3fffu);
    int jc;
    if ((best >> 17) & 1u) {
      jc = (int)(best & 0x1ffffu);  // the line's first uncovered tracked key: its first minimum
    } else {  // nothing uncovered is tracked: search the line itself
      const uint8_t* cross = Lc < m ? covC : covR;
      gcint* data = Lc < m ? C.cost + (uint32_t)(Lc * C.ld) : C.costT + (uint32_t)((Lc - m) * C.ldT);
      const int len = Lc < m ? n : m;
      uint64_t mk = kNoKey;
      if (books) {  // first uncovered book entry from the line's read position on
        gcshort* book = books + (Lc < m ? (size_t)Lc * n : (size_t)m * n + (size_t)(Lc - m) * m);
        for (int at = bpos[Lc]; at < len; at += G) {
          const int q = at + tid;
          uint64_t f = kNoKey;
          if (q < len) { const int j = book[q]; if (!cross[j]) f = (uint64_t)q; }
          f = group_min_u64<G>(f, S);
          if (f != kNoKey) { const int j = book[(int)f]; mk = pack_vi(data[j], (uint32_t)j); break; }
        }
      } else {
        for (int j = tid; j < len; j += G)
          if (!cross[j]) { const uint64_t k = pack_vi(data[j], (uint32_t)j); mk = k < mk ? k : mk; }
        mk = group_min_u64<G>(mk, S);
      }
      if (mk == kNoKey) {  // no uncovered cross line: cannot happen for a balanced instance
        if (tid == 0) S.status = 4;
        gsync<G>();
        return;
      }
      jc = (int)key_idx(mk);
      // every rank has read the cover flags for the search before rank 0 changes them
      gsync<G>();
    }
    // rank 0 eliminates; nobody else reads the state it changes before the barrier
    if (tid == 0) {
      const int r = Lc < m ? Lc : jc, q = Lc < m ? jc : Lc - m;
      const int32_t rs = res[r], rd = res[m + q];
      bool coverRow = rd > rs;
      if (!coverRow && covCols == n - 1) coverRow = true;
      int32_t f;
      if (coverRow) {
        f = rs; res[m + q] = rd - rs; res[r] = 0; covR[r] = 1;
        if constexpr (!kReg) pen[r] = -1;
      } else {
        f = rd; res[r] = rs - rd; res[m + q] = 0; covC[q] = 1;
        if constexpr (!kReg) pen[m + q] = -1;
        ++covCols;
      }
      S.pend_row = coverRow ? 1 : 0;
      S.pend_col = coverRow ? 0 : 1;
      S.lastx = coverRow ? r : m + q;
      brow[step] = r; bcol[step] = q; bflow[step] = f;
    }
    gsync<G>();
  }
}

// books: sorted lines (large instances) or null (scan lines).  scratch: the instance's
// global scratch (at least vogel_trk_bytes), used for the tracked keys when they do not
// fit in LDS after the rest of the state (`avail` bytes at `base`).
// gstate: global memory of at least vogel_lds_bytes for the line state when even that
// does not fit the `avail` LDS bytes at `base` (large-instance path with the tree in
// global memory); null otherwise.
// kWide: also instantiate the register mode with kRegLinesWide lines per rank (the
// large-instance set-up kernel, which has registers to spare), for m+n up to
// kRegLinesWide * G, with the tracked keys in LDS or in global scratch.
template <int G, bool kWide = false>
__device__ void vogel(const Ctx& C, gcint* supply, gcint* demand, char* base, size_t avail, Shared& S, gint* brow,
                      gint* bcol, gint* bflow, int32_t* scratch, gcshort* books = nullptr, char* gstate = nullptr) {
  const size_t need = vogel_lds_bytes(C.m, C.n);
  const size_t needR = vogel_reg_lds_bytes(C.m, C.n);
  const size_t trkBytes = vogel_trk_bytes(C.m, C.n);
  typedef FLOCK_LDS uint32_t LdsTrk;
  typedef FLOCK_GLOBAL uint32_t GlTrk;
  const bool regLines = C.V <= kRegLines * G;
  const bool regWide = kWide && C.V <= kRegLinesWide * G;
  if ((regLines || regWide) && needR + trkBytes <= avail) {  // register mode, tracked keys in LDS
    LdsTrk* trk = (LdsTrk*)(base + needR);
    if (regLines) {
      vogel_run<G, LdsTrk, kRegLines>(C, supply, demand, base, trk, S, brow, bcol, bflow, books);
    } else {
      if constexpr (kWide) vogel_run<G, LdsTrk, kRegLinesWide>(C, supply, demand, base, trk, S, brow, bcol, bflow, books);
    }
  } else if (!regLines && !regWide && need + trkBytes <= avail) {  // scan mode, tracked keys in LDS
    vogel_run<G, LdsTrk, 0>(C, supply, demand, base, (LdsTrk*)(base + need), S, brow, bcol, bflow, books);
  } else if (need <= avail) {
    GlTrk* trk = (GlTrk*)scratch;
    if (kWide && C.V <= kRegLinesWide * G) {
      if constexpr (kWide) vogel_run<G, GlTrk, kRegLinesWide>(C, supply, demand, base, trk, S, brow, bcol, bflow, books);
    } else {
      vogel_run<G, GlTrk, 0>(C, supply, demand, base, trk, S, brow, bcol, bflow, books);
    }
  } else if (gstate) {
    vogel_run<G, GlTrk, 0>(C, supply, demand, gstate, (GlTrk*)scratch, S, brow, bcol, bflow, books);
  } else {  // the host sizes LDS so that this cannot happen; report instead of overrunning it
    if (grank<G>() == 0) S.status = 6;
    gsync<G>();
  }
}

// North-west corner rule (IBFS_nwc.cu:3-34; a tie moves to the next row, :17-23, except
// on the last row).  Sequential, m+n-1 steps.
template <int G>
__device__ inline void nwc(const Ctx& C, gcint* supply, gcint* demand, gint* brow, gint* bcol, gint* bflow) {
  if (grank<G>() != 0) return;
  int i = 0, j = 0;
  int64_t rs = supply[0], rd = demand[0];
  for (int k = 0; k < C.V - 1; ++k) {
    brow[k] = i; bcol[k] = j;
    if ((rd >= rs && i < C.m - 1) || j == C.n - 1) {
      bflow[k] = (int32_t)rs; rd -= rs; ++i; rs = i < C.m ? supply[i] : 0;
    } else {
      bflow[k] = (int32_t)rd; rs -= rd; ++j; rd = j < C.n ? demand[j] : 0;
    }
  }
}

__device__ __forceinline__ int32_t dual_of(const TreeLds& t, int m, int x) { return x < m ? t.du[x] : t.dv[x - m]; }
__device__ __forceinline__ void set_dual(const TreeLds& t, int m, int x, int32_t v) {
  if (x < m) t.du[x] = v; else t.dv[x - m] = v;
}

// ---------------------------------------------------------------------------
// Basis tree from the m+n-1 basic cells: CSR adjacency (degree count, scan, fill --
// the reference's determine_length / inclusive_scan / fill_Ea, parallel_kernels.cu:
// 205-256, 294-319), then a level-synchronous BFS from vertex 0 that sets parent,
// depth, slot, flow and the duals (x_child = c(row,col) - x_parent, DUAL_solver.cu:
// 52-62), then subtree sizes bottom-up and preorder numbers top-down, level by level.
// Returns false if the cells do not form a spanning tree.
// ---------------------------------------------------------------------------
template <int G>
__device__ bool build_tree(const Ctx& C, gcint* brow, gcint* bcol, gcint* bflow, const TreeLds& t, int32_t* scr,
                           Shared& S) {
  const int m = C.m, V = C.V, tid = grank<G>();
  int32_t* start = scr;                    // V+1
  int32_t* vis = start + V + 1;            // V  (fill cursor, then visited flag, then size)
  int32_t* lvl = vis + V;                  // V+2
  int16_t* adj = (int16_t*)(lvl + V + 2);  // 2V
  int16_t* queue = adj + 2 * V;            // V
  for (int x = tid; x <= V; x += G) start[x] = 0;
  gsync<G>();
  int badCell = 0;
  for (int k = tid; k < V - 1; k += G) {
    const int r = brow[k], c = bcol[k];
    if (r < 0 || r >= m || c < 0 || c >= C.n) { badCell = 1; continue; }  // guard
    atomicAdd(&start[r + 1], 1);
    atomicAdd(&start[m + c + 1], 1);
  }
  if (badCell) atomicAdd(&S.status, 1);
  gsync<G>();
  if (S.status != 0) return false;
  if (tid == 0) {  // exclusive scan of the degrees
    int32_t run = 0;
    for (int x = 1; x <= V; ++x) { run += start[x]; start[x] = run; }
  }
  gsync<G>();
  for (int x = tid; x < V; x += G) vis[x] = start[x];
  gsync<G>();
  for (int k = tid; k < V - 1; k += G) {
    const int r = brow[k], c = bcol[k];
    adj[atomicAdd(&vis[r], 1)] = (int16_t)k;
    adj[atomicAdd(&vis[m + c], 1)] = (int16_t)k;
  }
  gsync<G>();
  for (int x = tid; x < V; x += G) { vis[x] = 0; t.parent[x] = -1; t.depth[x] = 0; }
  gsync<G>();
  if (tid == 0) {
    vis[0] = 1; queue[0] = 0; t.slot[0] = -1; t.flow[0] = 0; set_dual(t, m, 0, 0);
    lvl[0] = 0; lvl[1] = 1; S.qtail = 1; S.nlevels = 1;
  }
  gsync<G>();
  int d = 0;
  for (;;) {
    const int b = lvl[d], e = lvl[d + 1];
    if (b == e) break;  // uniform: written before the sync
    for (int q = b + tid; q < e; q += G) {
      const int x = queue[q];
      const int px = t.parent[x];
      const int32_t dx = dual_of(t, m, x);
      for (int a = start[x]; a < start[x + 1]; ++a) {
        const int k = adj[a];
        const int r = brow[k], c = bcol[k];
        const int y = x < m ? m + c : r;
        if (y == px) continue;
        if (atomicExch(&vis[y], 1) != 0) continue;  // a cycle in the cell set
        t.parent[y] = (int16_t)x; t.depth[y] = (int16_t)(d + 1); t.slot[y] = (int16_t)k;
        t.flow[y] = bflow[k];
        set_dual(t, m, y, cost_at(C, r, c) - dx);
        queue[atomicAdd(&S.qtail, 1)] = (int16_t)y;
      }
    }
    gsync<G>();
    if (tid == 0) { lvl[d + 2] = S.qtail; S.nlevels = d + 2; }
    gsync<G>();
    ++d;
  }
  const int nl = S.nlevels;  // levels 0..nl-2 are non-empty
  const bool ok = S.qtail == V;
  gsync<G>();
  if (!ok) return false;
  for (int x = tid; x < V; x += G) vis[x] = 1;
  gsync<G>();
  for (int dd = nl - 2; dd >= 1; --dd) {  // subtree sizes, deepest level first
    for (int q = lvl[dd] + tid; q < lvl[dd + 1]; q += G) {
      const int x = queue[q];
      atomicAdd(&vis[t.parent[x]], vis[x]);
    }
    gsync<G>();
  }
  if (tid == 0) t.pre(0) = 0;
  gsync<G>();
  for (int dd = 0; dd < nl - 1; ++dd) {  // preorder, children in CSR order
    for (int q = lvl[dd] + tid; q < lvl[dd + 1]; q += G) {
      const int x = queue[q];
      const int px = t.parent[x];
      int run = t.pre(x) + 1;
      for (int a = start[x]; a < start[x + 1]; ++a) {
        const int k = adj[a];
        const int y = x < m ? m + bcol[k] : brow[k];
        if (y == px) continue;
        t.pre(y) = (int16_t)run;
        run += vis[y];
      }
    }
    gsync<G>();
  }
  for (int x = tid; x < V; x += G) t.size(x) = (int16_t)vis[x];
  gsync<G>();
  return true;
}

// ---------------------------------------------------------------------------
// Reduced-cost sweep fused with the argmin (computeReducedCosts + thrust::min_element,
// parallel_kernels.cu:329-353 and uv_model_parallel.cu:169; host scan :184-194):
// r_ij = c_ij - u_i - v_j over all cells, minimum r, ties -> lowest row-major index.
// The padded matrix is walked as a flat array of 16-byte vectors; rank t takes
// vectors t, t+G, t+2G, ... in increasing order, kSweepUnroll loads in flight at a time.
// The rank's vectors are dealt round-robin to K slots (slot = position in the unrolled
// batch mod K); a strict `<` keeps each slot's first minimum, so the least (r, index)
// key over the slots is the rank's, and the group min over ranks finishes the argmin.
// Each slot's best cell is also a pricing candidate (K per rank).
// ---------------------------------------------------------------------------
// 16-byte loads in flight per rank: 8 for workgroup groups, 4 for wavefront groups
template <int G>
struct SweepUnroll { static constexpr int value = G == kWave ? 4 : 8; };
// candidates per rank: 8 for workgroup groups (2048 at 256 threads, 8192 at 1024; more
// candidates measured fewer sweeps and pivots at every group size: at 1024 threads an
// 1800x1800 instance needs 1645 sweeps instead of 7692 with one per rank), 4 for
// wavefront groups
template <int G>
struct CandsPerRank { static constexpr int value = G == kWave ? 4 : 8; };

struct Cand {
  int32_t row, col, cost;
};
// K candidate cells of one rank: row << 16 | col (-1: none) and cost
template <int K>
struct CandSet {
  int32_t rc[K];
  int32_t cost[K];
  __device__ __forceinline__ void clear() {
#pragma unroll
    for (int k = 0; k < K; ++k) { rc[k] = -1; cost[k] = 0; }
  }
  __device__ __forceinline__ Cand get(int k) const {
    Cand c;
    c.row = rc[k] < 0 ? -1 : (rc[k] >> 16);
    c.col = rc[k] < 0 ? 0 : (rc[k] & 0xffff);
    c.cost = cost[k];
    return c;
  }
  __device__ __forceinline__ void set(int k, const Cand& c) {
    rc[k] = c.row < 0 ? -1 : ((c.row << 16) | c.col);
    cost[k] = c.cost;
  }
};

template <int G, int K>
__device__ uint64_t sweep_argmin(const Ctx& C, const TreeLds& t, Shared& S, CandSet<K>& mine) {
  const uint32_t vpr = (uint32_t)C.ld >> 2;
  const uint32_t nvec = (uint32_t)C.m * vpr;
  // q / vpr as (q * magic) >> 40 with magic = floor(2^40 / vpr) + 1: the error term is
  // below q / 2^40 < 2^-16 while the fractional part of q / vpr is at most
  // 1 - 1/vpr, so the quotient is exact for q < 2^24 and vpr <= 2^11 (m*n <= 2^26).
  const uint64_t magic = (0x10000000000ull / vpr) + 1;
  const uint32_t ld = (uint32_t)C.ld;
  constexpr int kSweepUnroll = SweepUnroll<G>::value;
  static_assert(kSweepUnroll % K == 0, "slots must divide the unrolled batch");
  // per vector only its minimum reduced cost is compared (strict <, so the slot's first
  // minimum vector wins); the winning cell inside that vector is recovered once at the end
  int32_t bestR[K];
  uint32_t bestQ[K];
#pragma unroll
  for (int k = 0; k < K; ++k) { bestR[k] = kBig; bestQ[k] = 0xffffffffu; }
  for (uint32_t q0 = (uint32_t)grank<G>(); q0 < nvec; q0 += kSweepUnroll * G) {
    i32x4 cv[kSweepUnroll];
    uint32_t rw[kSweepUnroll], c4[kSweepUnroll];
    // addresses first, loads unconditional (out-of-range slots re-read vector 0), so all
    // kSweepUnroll loads are in flight together
#pragma unroll
    for (int k = 0; k < kSweepUnroll; ++k) {
      const uint32_t q = q0 + (uint32_t)k * G;
      const uint32_t qq = q < nvec ? q : nvec - 1;  // slots past the matrix repeat the last vector
      const uint32_t r = (uint32_t)(((uint64_t)qq * magic) >> 40);
      rw[k] = r; c4[k] = qq - r * vpr;
      cv[k] = *(const FLOCK_GLOBAL i32x4*)(C.cost + (r * ld + 4 * c4[k]));
    }
#pragma unroll
    for (int k = 0; k < kSweepUnroll; ++k) {
      const uint32_t q = q0 + (uint32_t)k * G;
      const uint32_t qq = q < nvec ? q : nvec - 1;
      const i32x4 vv = *(const i32x4*)(t.dv + 4 * c4[k]);
      const int32_t u = t.du[rw[k]];
      // min over the vector of c - v, then - u once (the row is the vector's); padding
      // cells (cost kPadCost, dual 0) price positive and never hold a negative minimum.
      // A repeated last vector has the same (r, index) keys as the original.
      const int32_t mn = imin(imin(cv[k].x - vv.x, cv[k].y - vv.y), imin(cv[k].z - vv.z, cv[k].w - vv.w)) - u;
      const bool better = mn < bestR[k % K];
      bestR[k % K] = better ? mn : bestR[k % K];
      bestQ[k % K] = better ? qq : bestQ[k % K];
    }
  }
  // recover each slot's cell (its vector reloaded; the K loads in flight together)
  i32x4 cv[K];
  uint32_t rr[K], cc4[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const uint32_t q = bestQ[k] != 0xffffffffu ? bestQ[k] : 0u;
    rr[k] = (uint32_t)(((uint64_t)q * magic) >> 40);
    cc4[k] = q - rr[k] * vpr;
    cv[k] = *(const FLOCK_GLOBAL i32x4*)(C.cost + (rr[k] * ld + 4 * cc4[k]));
  }
  uint64_t key = kNoKey;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const bool none = bestQ[k] == 0xffffffffu;
    const uint32_t r = rr[k], c = cc4[k];
    const int32_t u = t.du[r];
    const i32x4 vv = *(const i32x4*)(t.dv + 4 * c);
    const int32_t cc[4] = {cv[k].x, cv[k].y, cv[k].z, cv[k].w};
    const int32_t dv[4] = {vv.x, vv.y, vv.z, vv.w};
    int e = 3;
#pragma unroll
    for (int j = 3; j >= 0; --j)  // first cell at the minimum
      e = (4 * (int)c + j < C.n && cc[j] - u - dv[j] == bestR[k]) ? j : e;
    const int32_t ce = e == 0 ? cc[0] : e == 1 ? cc[1] : e == 2 ? cc[2] : cc[3];
    const uint32_t col = 4 * c + (uint32_t)e;
    mine.rc[k] = none ? -1 : (int32_t)((r << 16) | col);
    mine.cost[k] = none ? 0 : ce;
    const uint64_t kk = none ? kNoKey : pack_vi(bestR[k], r * (uint32_t)C.n + col);
    key = kk < key ? kk : key;
  }
  return group_min1_u64<G, 2>(key, S);
}

// Re-price a rank's candidates with the current duals (costs kept in registers): their
// least (r, row-major index) key.
template <int K>
__device__ __forceinline__ uint64_t reprice_set(const Ctx& C, const TreeLds& t, const CandSet<K>& cs) {
  uint64_t key = kNoKey;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int32_t rc = cs.rc[k];
    const int row = rc < 0 ? 0 : (rc >> 16), col = rc < 0 ? 0 : (rc & 0xffff);
    const int32_t r = cs.cost[k] - t.du[row] - t.dv[col];
    // an empty slot ORs in all ones (kNoKey) instead of branching
    const uint64_t kk = pack_vi(r, (uint32_t)row * (uint32_t)C.n + (uint32_t)col) | (rc < 0 ? kNoKey : 0ull);
    key = kk < key ? kk : key;
  }
  return key;
}

// Re-price a candidate with the current duals (cost kept in a register).
__device__ __forceinline__ uint64_t reprice(const Ctx& C, const TreeLds& t, const Cand& c) {
  if (c.row < 0) return kNoKey;
  const int32_t r = c.cost - t.du[c.row] - t.dv[c.col];
  return pack_vi(r, (uint32_t)c.row * (uint32_t)C.n + (uint32_t)c.col);
}

// ---------------------------------------------------------------------------
// One tile of a full sweep (the large-instance path's sweep grid): vectors [tile*vpt,
// (tile+1)*vpt) of the padded matrix against the duals in global memory (u then v).
// Writes the tile's argmin key (r, row-major index) to keys[tile] and, per wavefront,
// its best cell (row, col, cost) to cands[3*(tile*W + wave)] (row -1: none).
// Everything comes from kernel arguments, so a rank issues its cost vectors, their row
// duals, their four column duals and the phase word together (one round trip); as in
// sweep_argmin, a vector is compared by its least reduced cost and the winning cell is
// recovered once at the end.  When the phase word is not 1 (instance done) nothing is
// written.
// ---------------------------------------------------------------------------
// candidates a 256-thread sweep tile writes: one per 16-lane row of each wavefront
constexpr int kTileCands = 16;
struct MinU64Row { __device__ __forceinline__ uint64_t operator()(uint64_t a, uint64_t b) const { return b < a ? b : a; } };
struct SweepArgs {
  const int32_t* cost;   // m x ld, padded rows
  const int32_t* dual;   // m + n: u then v
  const int32_t* phase;  // instance phase word
  uint64_t* keys;
  int32_t* cands;
  int32_t m, n, ld;
  uint32_t vpt;
};

template <int G>
__device__ void sweep_tile(const SweepArgs& A, uint32_t tile, Shared& S) {
  const int tid = grank<G>(), lane = tid % kWave, wave = tid / kWave;
  static_assert(G / kWave * 4 == kTileCands, "a tile's candidates: four rows per wavefront");
  constexpr int U = 4;  // vectors in flight per rank
  gcint* cost = gl(A.cost);
  gcint* dual = gl(A.dual);
  const int m = A.m, n = A.n, nLast = A.n - 1;
  const uint32_t vpr = (uint32_t)A.ld >> 2;
  const uint32_t nvec = (uint32_t)m * vpr;
  const uint64_t magic = (0x10000000000ull / vpr) + 1;  // exact quotient, as in sweep_argmin
  const uint32_t ld = (uint32_t)A.ld;
  const uint32_t v0 = tile * A.vpt;
  const uint32_t v1 = v0 + A.vpt < nvec ? v0 + A.vpt : nvec;
  const int32_t ph = *gl(A.phase);
  int32_t bestR = kBig;
  uint32_t bestQ = 0xffffffffu;
  for (uint32_t q0 = v0 + (uint32_t)tid; q0 < v1; q0 += U * G) {
    i32x4 cv[U];
    int32_t uu[U], vv[U][4];
    uint32_t c4[U];
#pragma unroll
    for (int k = 0; k < U; ++k) {
      const uint32_t q = q0 + (uint32_t)k * G;
      const uint32_t qq = q < v1 ? q : v0;  // out-of-range slots re-read the tile's first vector
      const uint32_t r = (uint32_t)(((uint64_t)qq * magic) >> 40);
      c4[k] = qq - r * vpr;
      cv[k] = *(const FLOCK_GLOBAL i32x4*)(cost + (r * ld + 4 * c4[k]));
      uu[k] = dual[r];
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int j = 4 * (int)c4[k] + e;
        vv[k][e] = dual[m + (j < nLast ? j : nLast)];
      }
    }
#pragma unroll
    for (int k = 0; k < U; ++k) {
      const uint32_t q = q0 + (uint32_t)k * G;
      const int lim = q < v1 ? n - 4 * (int)c4[k] : 0;  // cells past the row end or the tile never win
      const int32_t r0 = lim > 0 ? cv[k].x - uu[k] - vv[k][0] : kBig;
      const int32_t r1 = lim > 1 ? cv[k].y - uu[k] - vv[k][1] : kBig;
      const int32_t r2 = lim > 2 ? cv[k].z - uu[k] - vv[k][2] : kBig;
      const int32_t r3 = lim > 3 ? cv[k].w - uu[k] - vv[k][3] : kBig;
      const int32_t mn = imin(imin(r0, r1), imin(r2, r3));
      const bool better = mn < bestR;
      bestR = better ? mn : bestR;
      bestQ = better ? q : bestQ;
    }
  }
  if (ph != 1) return;  // uniform over the grid
  // the rank's best cell: its vector reloaded, the first cell at the minimum
  const bool none = bestQ == 0xffffffffu;
  const uint32_t qb = none ? v0 : bestQ;
  const uint32_t rb = (uint32_t)(((uint64_t)qb * magic) >> 40);
  const uint32_t cb = qb - rb * vpr;
  const i32x4 c = *(const FLOCK_GLOBAL i32x4*)(cost + (rb * ld + 4 * cb));
  const int32_t ub = dual[rb];
  const int32_t cc[4] = {c.x, c.y, c.z, c.w};
  int e = 3;
#pragma unroll
  for (int j = 3; j >= 0; --j) {
    const int col = 4 * (int)cb + j;
    e = (col < n && cc[j] - ub - dual[m + (col < nLast ? col : nLast)] == bestR) ? j : e;
  }
  const int32_t ce = e == 0 ? cc[0] : e == 1 ? cc[1] : e == 2 ? cc[2] : cc[3];
  const uint32_t colb = 4 * cb + (uint32_t)e;
  const uint64_t key = none ? kNoKey : pack_vi(bestR, rb * (uint32_t)n + colb);
  // one candidate per 16-lane row of each wavefront (kTileCands per tile): the row's
  // least key by DPP row shifts lands in the row's lane 15
  uint64_t rk = key;
  rk = MinU64Row{}(rk, dpp_u64<0x111, 0xf>(kNoKey, rk));
  rk = MinU64Row{}(rk, dpp_u64<0x112, 0xf>(kNoKey, rk));
  rk = MinU64Row{}(rk, dpp_u64<0x114, 0xf>(kNoKey, rk));
  rk = MinU64Row{}(rk, dpp_u64<0x118, 0xf>(kNoKey, rk));
  const int row16 = lane >> 4;
  const uint64_t r0 = readlane_u64(rk, 15), r1 = readlane_u64(rk, 31), r2 = readlane_u64(rk, 47),
                 r3 = readlane_u64(rk, 63);
  const uint64_t rowMin = row16 == 0 ? r0 : row16 == 1 ? r1 : row16 == 2 ? r2 : r3;
  int32_t* cw = A.cands + 3 * ((size_t)tile * kTileCands + (size_t)wave * 4 + (size_t)row16);
  if (rowMin == kNoKey) {
    if ((lane & 15) == 0) cw[0] = -1;
  } else if (key == rowMin) {  // the one lane holding the row's best cell (keys are unique)
    cw[0] = (int32_t)rb; cw[1] = (int32_t)colb; cw[2] = ce;
  }
  const uint64_t wk = r0 < r1 ? (r0 < r2 ? (r0 < r3 ? r0 : r3) : (r2 < r3 ? r2 : r3))
                              : (r1 < r2 ? (r1 < r3 ? r1 : r3) : (r2 < r3 ? r2 : r3));
  const uint64_t bk = group_min_u64<G>(wk, S);
  if (tid == 0) A.keys[tile] = bk;
}

// ---------------------------------------------------------------------------
// One stepping-stone pivot on entering cell (ei, ej) with reduced cost r < 0
// (perform_a_sequencial_pivot / execute_pivot_on_host / do_flow_adjustment_on_host,
// PIVOT_uv.cu:156-308, 318-401):
//  A. cycle discovery, parallel over vertices: x is on the tree path between row I and
//     column J iff it is an ancestor-or-self of exactly one of them.  The arc (x,
//     parent x) carries -theta when x has the type of the path end it hangs from (rows
//     on I's side, columns on J's side).  Leaving arc = the FIRST minimum flow among
//     the -theta arcs in path order J -> apex -> I (PIVOT_uv.cu:280-296), found with
//     one group min over (flow, side, depth rank, vertex) keys.
//  B. flows +-theta on the path; the chain a = P[0] .. P[k] = q from the leaving arc's
//     lower vertex a down to the entering endpoint q inside subtree(a) is recorded with
//     a snapshot of its old preorder number, size, arc slot and (updated) arc flow.
//  C. subtree(a) is cut, re-rooted at q and hung below the other endpoint o, as o's
//     first child.  Every vertex's new preorder number, depth, size, parent, arc
//     slot/flow and dual follow from its own old values, the snapshot and a few scalars,
//     so all vertices update in place in parallel (the re-rooted subtree's preorder is
//     the concatenation of the old intervals of P[k], P[k-1] minus P[k], ..., P[0]
//     minus P[1]).  The entering arc takes the leaving arc's slot (PIVOT_uv.cu:248-261);
//     duals in the subtree shift by -/+ r so that the entering arc becomes tight.
// ---------------------------------------------------------------------------
// kR: vertices each rank handles per pass, their LDS loads issued together.
// kWideWin: the re-root's windowed form takes twice the slots per rank where kR is below
// four (the graph path's round kernel and the work-queue kernels below 1024 threads; the
// 1024-thread work-queue kernel keeps kR, as it would spill more in its other phases).
template <int G, int kR, class SN, bool kWideWin = false>
__device__ void pivot_pass(const Ctx& C, const TreeLds& t, const SN& sn, Shared& S, int ei, int ej, int32_t r) {
  const int m = C.m, V = C.V, tid = grank<G>();
  const uint64_t t0 = now_ticks();
  const int I = ei, J = m + ej;
  // every scalar the in-place pass needs is read here or taken from the snapshot: once
  // the group minimum is known, ranks go straight on to rewrite their own vertices
  const int preI = t.pre(I), preJ = t.pre(J), depI = t.depth[I], depJ = t.depth[J];
  const int sel = S.dsel;
  // ---- A: leaving arc, and every cycle vertex snapshotted by side and depth (the least
  // cycle depth, one below the apex, goes to S.dlo[sel])
  uint64_t best = kNoKey;
  // with several vertices per rank, every LDS load of a pass is issued before the first
  // use (no per-vertex branch waits on a load), so one wait covers all kR vertices
  for (int x0 = tid; x0 < V; x0 += kR * G) {
    int px[kR], sx[kR], dx[kR], slx[kR];
    int32_t fx[kR];
#pragma unroll
    for (int u = 0; u < kR; ++u) {
      const int x = x0 + u * G;
      const int xx = x < V ? x : 0;
      const int32_t psv = t.pre_size(xx);
      px[u] = ps_pre(psv);
      const int sv = ps_size(psv);
      sx[u] = x < V ? sv : 0;  // size 0: never an ancestor
      if constexpr (kR > 1) { dx[u] = t.depth[xx]; fx[u] = t.flow[xx]; slx[u] = t.slot[xx]; }
    }
#pragma unroll
    for (int u = 0; u < kR; ++u) {
      const int x = x0 + u * G;
      const bool aI = (uint32_t)(preI - px[u]) < (uint32_t)sx[u];  // px <= preI < px + sx
      const bool aJ = (uint32_t)(preJ - px[u]) < (uint32_t)sx[u];
      if (aI != aJ) {
        // one vertex per rank: only cycle vertices load the rest (fewer LDS reads)
        if constexpr (kR == 1) { dx[u] = t.depth[x]; fx[u] = t.flow[x]; slx[u] = t.slot[x]; }
        const int d = dx[u];
        const int sd = aI ? 0 : 1;
        sn.vert(sd)[d] = (int16_t)x; sn.ps(sd)[d] = pack_ps(px[u], sx[u]);
        sn.slot(sd)[d] = (int16_t)slx[u]; sn.flow(sd)[d] = fx[u];
        atomicMin(&S.dlo[sel], d);
        const uint32_t lowbits =
            aI ? ((1u << 31) | ((uint32_t)d << 16) | (uint32_t)x) : (((32767u - (uint32_t)d) << 16) | (uint32_t)x);
        const uint64_t k = aI == (x < m) ? (((uint64_t)(uint32_t)fx[u] << 32) | lowbits) : kNoKey;
        best = k < best ? k : best;
      }
    }
  }
  best = group_min1_u64<G, 1>(best, S);  // its barrier also publishes the snapshot
  if (best == kNoKey) {  // no blocking arc: the tree is corrupt; report instead of pivoting
    if (tid == 0) S.status = 2;
    gsync<G>();
    return;
  }
  const int a = (int)(best & 0xffffu);
  const int32_t theta = (int32_t)(best >> 32);
  const bool sideI = (best >> 31) & 1u;
  const int dA = sideI ? (int)((best >> 16) & 0x7fffu) : 32767 - (int)((best >> 16) & 0x7fffu);
  const int o = sideI ? J : I;
  const uint64_t t1 = now_ticks();
  if (tid == 0) S.tk[4] += (int64_t)(t1 - t0);
  // the chain P[0] = a .. P[k] = q on q's side: snapshot entries dA .. dA + k
  const int qs = sideI ? 0 : 1;
  const int32_t psA = sn.ps(qs)[dA];
  const int A = ps_pre(psA), s = ps_size(psA), slotA = sn.slot(qs)[dA];
  const int k = (sideI ? depI : depJ) - dA;
  const int B = sideI ? preJ : preI;
  const int dO = sideI ? depJ : depI;
  const auto cVert = sn.vert(qs) + dA;
  const auto cPS = sn.ps(qs) + dA;
  const auto cSlot = sn.slot(qs) + dA;
  const auto cFlow = sn.flow(qs) + dA;
  // ---- C (in place): flows +-theta on the cycle, subtree(a) re-rooted at q below o
  const int newStart = B < A ? B + 1 : B - s + 1;
  const int32_t dRow = sideI ? r : -r;   // q = I: u += r, v -= r ; q = J: u -= r, v += r
  const int32_t dCol = -dRow;
  // binary search for the deepest chain vertex P[tt] whose old interval holds a position:
  // the candidate range [lo, lo + len) shrinks by the same scalar half in every lane, so
  // a probe lo + half never passes k; "P[c] holds p" is true for c <= tt and false
  // after, and false for every c >= 1 outside subtree(a), where lo stays 0
  const int len0 = uniform(k + 1);
  // window: subtree(a) and the pre-order positions it moves across (the same in every
  // lane; readfirstlane makes the branch below, which holds a barrier, scalar).  The
  // windowed form keeps three values per window slot; with kWideWin it takes twice the
  // dense form's vertices per rank where those are fewer than four (one below 2G
  // vertices, two on 1024-thread groups): every window then fits one pass up to 2G
  // vertices, and up to 4G at 1024 threads
  const int wsize = uniform(B < A ? A + s - 1 - B : B - A + 1);
  constexpr int kRW = kWideWin && kR < 4 ? 2 * kR : kR;
  if (wsize <= kRW * G) {
    // Windowed form, whenever this pivot's window fits one pass (whatever m+n).  Only
    // the pre-order window [wlo, whi] (subtree(a) and the vertices it moves across)
    // changes pre-order number, depth or dual, and only cycle vertices change size,
    // flow, parent or slot.  Window positions are visited through vat; cycle vertices
    // through the snapshot (depths dLca+1 .. depI on I's side, .. depJ on J's side).
    const int dLca = S.dlo[sel] - 1;
    const int n0 = depI - dLca, nc = n0 + depJ - dLca;
    // window slot w = rank + u*G: the subtree's positions first (w < s: A + w), then the
    // positions it moves across (B < A: B+1 .. A-1; else A+s .. B, i.e. A + w again)
    const int outBase = B < A ? B + 1 - s : A;
    const int shift = B < A ? s : -s;  // every window position outside subtree(a) moves
    const int depBase = dO + 1 + k - dA;  // new depth of subtree vertices: depBase + old - 2 tt
    const int dvIdx = (int)(t.dv - t.du) - m;  // column x's dual at t.du[x + dvIdx]
    // the rank index, opaque here: slot numbers rank + u*G are recomputed (one add)
    // instead of being hoisted to the kernel start and kept in scratch
    int tidr = tid;
    asm volatile("" : "+v"(tidr));
    int xs[kRW], np[kRW], lo[kRW];
#pragma unroll
    for (int u = 0; u < kRW; ++u) {
      const int w = tidr + u * G;
      const int p = (w < s ? A : outBase) + w;
      const bool valid = w < wsize;
      // slots past the window read position 0, the root, and keep it complemented: the
      // root is never in the window (it has no parent arc), so no rank rewrites its
      // depth or dual in this pass; vat itself is rewritten only after the barrier below
      const int x = t.vat[valid ? p : 0];
      xs[u] = valid ? x : ~x;
      lo[u] = 0;
    }
    // every slot searches (a slot outside the subtree keeps lo = 0): skipping slots by a
    // uniform test costs more than the searches it saves
    for (int len = len0; len > 1;) {
      const int half = len >> 1;
      int32_t v[kRW];
#pragma unroll
      for (int u = 0; u < kRW; ++u) v[u] = cPS[lo[u] + half];
#pragma unroll
      for (int u = 0; u < kRW; ++u) {
        const int p = A + tidr + u * G;
        const bool anc = (uint32_t)(p - ps_pre(v[u])) < (uint32_t)ps_size(v[u]);
        lo[u] = anc ? lo[u] + half : lo[u];
      }
      len -= half;
    }
#pragma unroll
    for (int u = 0; u < kRW; ++u) {
      const int w = tidr + u * G;
      const bool inSub = w < s;
      const int p = (inSub ? A : outBase) + w;
      const int tt = lo[u];
      const int x = xs[u] < 0 ? ~xs[u] : xs[u];
      const int dx = t.depth[x], di = x < m ? x : x + dvIdx;
      const int32_t du = t.du[di];
      // P[tt] \ P[tt+1] follows P[tt+1]'s interval; for tt = k (P[k] = q) entry t1 = k
      // gives p >= p1, so the same formula leaves q's interval in place
      const int t1 = tt < k ? tt + 1 : k;
      const int32_t v0 = cPS[tt], v1 = cPS[t1];
      const int p1 = ps_pre(v1), sz1 = ps_size(v1);
      const int nrel = p - ps_pre(v0) + (p < p1 ? sz1 : 0);
      np[u] = inSub ? newStart + nrel : p + shift;
      const int ndep = inSub ? depBase + dx - 2 * tt : dx;
      const int32_t ndual = du + (inSub ? (x < m ? dRow : dCol) : 0);
      if (xs[u] >= 0) {
        t.pre(x) = (int16_t)np[u]; t.depth[x] = (int16_t)ndep;
        t.du[di] = ndual;
      }
    }
    // cycle vertices: sizes on the paths to o and a, flows +-theta, the chain's arcs
    for (int i = tid; i < nc; i += G) {
      const int sd = i < n0 ? 0 : 1;
      const int d = dLca + 1 + (i < n0 ? i : i - n0);
      const int32_t psx = sn.ps(sd)[d];
      const int x = sn.vert(sd)[d], px = ps_pre(psx), sx = ps_size(psx);
      const int32_t fx = sn.flow(sd)[d];
      const bool chain = sd == qs && d >= dA;
      if (chain) {
        const int tt = d - dA;
        if (tt == k) {
          t.size(x) = (int16_t)s; t.parent[x] = (int16_t)o; t.flow[x] = theta; t.slot[x] = (int16_t)slotA;
        } else {
          const int up = cVert[tt + 1];
          const bool upMinus = sideI ? (up < m) : (up >= m);
          t.size(x) = (int16_t)(s - ps_size(cPS[tt + 1])); t.parent[x] = (int16_t)up;
          t.flow[x] = cFlow[tt + 1] + (upMinus ? -theta : theta); t.slot[x] = cSlot[tt + 1];
        }
      } else {
        const bool ancO = (uint32_t)(B - px) < (uint32_t)sx;
        const bool ancPa = px < A && A < px + sx;  // proper ancestor of a
        t.size(x) = (int16_t)(sx + ((int)ancO - (int)ancPa) * s);
        t.flow[x] = fx + (((sd == 0) == (x < m)) ? -theta : theta);
      }
    }
    gsync<G>();  // publishes every tree field; later readers need nothing below
    // vat and the depth-minimum slots are next read by the next pivot (after its phase
    // A barrier; every caller also has a re-pricing or sweep barrier between pivots), so
    // no barrier closes this pass
#pragma unroll
    for (int u = 0; u < kRW; ++u)
      if (xs[u] >= 0) t.vat[np[u]] = (int16_t)xs[u];
    if (tid == 0) { S.dlo[sel ^ 1] = 0x7fffffff; S.dsel = sel ^ 1; }
    if (tid == 0) S.tk[5] += (int64_t)(now_ticks() - t1);
    return;
  }
  for (int x0 = tid; x0 < V; x0 += kR * G) {
    int px[kR], sx[kR], dx[kR], lo[kR];
    int32_t fl[kR], du[kR];
#pragma unroll
    for (int u = 0; u < kR; ++u) {
      const int x = x0 + u * G;
      // ranks past the last vertex re-read their own first vertex (x0 < V): every other
      // vertex is being rewritten by its own rank in this pass
      const bool in = x < V;
      const int xx = in ? x : x0;
      const int32_t psv = t.pre_size(xx);
      const int pv = ps_pre(psv), sv = ps_size(psv);
      px[u] = in ? pv : -1;
      sx[u] = in ? sv : 0;
      dx[u] = t.depth[xx];
      fl[u] = t.flow[xx];
      du[u] = xx < m ? t.du[xx] : t.dv[xx - m];
    }
#pragma unroll
    for (int u = 0; u < kR; ++u) lo[u] = 0;
    // deepest chain vertex P[tt] that is an ancestor-or-self of x (nested intervals),
    // the kR searches' loads issued together each round; tail ranks (px = -1) hold no
    // interval, so their lo stays 0
    for (int len = len0; len > 1;) {
      const int half = len >> 1;
      int32_t v[kR];
#pragma unroll
      for (int u = 0; u < kR; ++u) v[u] = cPS[lo[u] + half];
#pragma unroll
      for (int u = 0; u < kR; ++u) {
        const bool anc = (uint32_t)(px[u] - ps_pre(v[u])) < (uint32_t)ps_size(v[u]);
        lo[u] = anc ? lo[u] + half : lo[u];
      }
      len -= half;
    }
    // chain entries tt and tt + 1 (clamped to k) for every vertex, loaded together (the
    // arc fields of entry tt + 1 only on the chain itself)
    int cp0[kR], cp1[kR], cs1[kR];
#pragma unroll
    for (int u = 0; u < kR; ++u) {
      const int tt = lo[u], t1 = tt < k ? tt + 1 : k;
      const int32_t v0 = cPS[tt], v1 = cPS[t1];
      cp0[u] = ps_pre(v0); cp1[u] = ps_pre(v1); cs1[u] = ps_size(v1);
    }
    // every field of the vertex is rewritten (unchanged ones with their old value), so
    // the only branches are the tail guard and the chain's arc update
#pragma unroll
    for (int u = 0; u < kR; ++u) {
      const int x = x0 + u * G;
      const bool inSub = A <= px[u] && px[u] < A + s;
      const int tt = lo[u];
      const int py = cp0[u];
      const bool last = tt == k;
      const int sz1 = cs1[u];
      const int rel = px[u] - py;  // position inside P[tt]'s old interval
      const int nrel = last ? rel : sz1 + rel - (px[u] < cp1[u] ? 0 : sz1);
      const bool chain = inSub & (px[u] == py);  // x is P[tt]: its arc becomes the old arc above P[tt+1]
      const int shifted = B < A ? ((B < px[u] && px[u] < A) ? px[u] + s : px[u])
                                : ((A + s <= px[u] && px[u] <= B) ? px[u] - s : px[u]);
      const int npre = inSub ? newStart + nrel : shifted;
      const int ndep = inSub ? dO + 1 + (k - tt) + (dx[u] - (dA + tt)) : dx[u];
      const bool ancO = (uint32_t)(B - px[u]) < (uint32_t)sx[u];
      const bool ancPa = px[u] < A && A < px[u] + sx[u];  // proper ancestor of a
      const int outDelta = ((int)ancO - (int)ancPa) * s;
      const int chainSize = s - (last ? 0 : sz1);
      const int otherSize = sx[u] + (inSub ? 0 : outDelta);
      const int nsize = chain ? chainSize : otherSize;
      const bool aI = (uint32_t)(preI - px[u]) < (uint32_t)sx[u];  // px <= preI < px + sx
      const bool aJ = (uint32_t)(preJ - px[u]) < (uint32_t)sx[u];
      const bool minus = aI == (x < m);  // the arc to the parent carries -theta
      int32_t nflow = fl[u] + (aI != aJ ? (minus ? -theta : theta) : 0);
      const int32_t ndual = du[u] + (inSub ? (x < m ? dRow : dCol) : 0);
      if (x < V) {
        if (chain) {
          int up, sl;
          if (last) {
            up = o; sl = slotA; nflow = theta;
          } else {
            up = cVert[tt + 1];
            const bool upMinus = sideI ? (up < m) : (up >= m);
            sl = cSlot[tt + 1];
            nflow = cFlow[tt + 1] + (upMinus ? -theta : theta);
          }
          t.parent[x] = (int16_t)up; t.slot[x] = (int16_t)sl;
        }
        t.pre(x) = (int16_t)npre; t.depth[x] = (int16_t)ndep; t.size(x) = (int16_t)nsize; t.flow[x] = nflow;
        *(x < m ? &t.du[x] : &t.dv[x - m]) = ndual;
        t.vat[npre] = (int16_t)x;  // nobody reads vat in this pass
      }
    }
  }
  if (tid == 0) { S.dlo[sel ^ 1] = 0x7fffffff; S.dsel = sel ^ 1; }
  gsync<G>();
  if (tid == 0) S.tk[5] += (int64_t)(now_ticks() - t1);
}

template <int G, bool kWideWin = false, class SN>
__device__ __forceinline__ void pivot(const Ctx& C, const TreeLds& t, const SN& sn, Shared& S, int ei, int ej,
                                      int32_t r) {
  // 1024-thread groups take two vertices per rank and pass at most (their registers are
  // capped at 128)
  constexpr int kRBig = G >= 1024 ? 2 : 4;
  if (C.V > 2 * G) pivot_pass<G, kRBig, SN, kWideWin>(C, t, sn, S, ei, ej, r);
  else pivot_pass<G, 1, SN, kWideWin>(C, t, sn, S, ei, ej, r);
}

// ---------------------------------------------------------------------------
// Multi-pivot rounds (the reference's parallel stepping stone, SS_METHOD::
// perform_a_parallel_pivot, PIVOT_ss.cu:294-580, and its conflict check
// check_pivot_feasibility, parallel_kernels.cu:620-660).  Every rank holds one candidate
// entering cell from the last full sweep.  In a round, ranks whose candidate prices
// negative with the current duals walk its cycle (the tree path row -> apex <- column,
// by parent pointers and depths) and claim each arc on it (an arc is named by its child
// vertex) with an atomicMin of a priority key (more negative reduced cost first, then
// rank); a candidate whose second walk finds every claim still its own is accepted.
// Accepted cycles share no arc, so pivoting them one after another leaves every other
// accepted cycle, its reduced cost and its theta unchanged; they pivot best first.  The
// best candidate always wins.  Rounds repeat on the same candidates (re-priced) up to
// minor_limit times or until none prices negative; then the caller sweeps again.
// `claim` needs V words (the pivot's snapshot region, free between pivots).  Returns
// the pivots made; stops at the iteration budget.
// ---------------------------------------------------------------------------
template <int G>
__device__ __attribute__((noinline)) int multi_pivot(const Ctx& C, const TreeLds& t, const Snap& sn, uint32_t* claim, Shared& S,
                           const Cand& cand, int64_t iter, int64_t maxIt, int rounds) {
  const int m = C.m, V = C.V, tid = grank<G>();
  constexpr uint32_t kRange = (1u << 22) - 1;
  const int I = cand.row, J = m + cand.col;
  const uint32_t mine = (uint32_t)cand.row * (uint32_t)C.n + (uint32_t)cand.col;
  bool spent = cand.row < 0;  // this rank's candidate has entered the basis
  int done = 0;
  for (int round = 0; round < rounds && iter + done < maxIt; ++round) {
    int32_t r = 0;
    if (!spent) r = cand.cost - t.du[cand.row] - t.dv[cand.col];
    const bool active = !spent && r < 0;
    const uint32_t mag = (uint32_t)(-(int64_t)r) < kRange ? (uint32_t)(-(int64_t)r) : kRange;
    const uint32_t key = active ? ((kRange - mag) << 10) | ((uint32_t)tid & 1023u) : 0xffffffffu;
    for (int x = tid; x < V; x += G) claim[x] = 0xffffffffu;
    gsync<G>();
    if (active) {
      int x = I, y = J;
      while (t.depth[x] > t.depth[y]) { atomicMin(&claim[x], key); x = t.parent[x]; }
      while (t.depth[y] > t.depth[x]) { atomicMin(&claim[y], key); y = t.parent[y]; }
      while (x != y) { atomicMin(&claim[x], key); atomicMin(&claim[y], key); x = t.parent[x]; y = t.parent[y]; }
    }
    gsync<G>();
    bool accepted = false;
    if (active) {
      bool own = true;
      int x = I, y = J;
      while (own && t.depth[x] > t.depth[y]) { own = claim[x] == key; x = t.parent[x]; }
      while (own && t.depth[y] > t.depth[x]) { own = claim[y] == key; y = t.parent[y]; }
      while (own && x != y) { own = claim[x] == key && claim[y] == key; x = t.parent[x]; y = t.parent[y]; }
      accepted = own;
    }
    gsync<G>();  // the claim words are the pivot's snapshot region from here on
    int inRound = 0;
    while (iter + done < maxIt) {
      const uint64_t k = group_min_u64<G>(accepted ? reprice(C, t, cand) : kNoKey, S);
      if (k == kNoKey || key_val(k) >= 0) break;
      const uint32_t cell = key_idx(k);
      if (accepted && cell == mine) { accepted = false; spent = true; }
      const int ei = (int)(cell / (uint32_t)C.n), ej = (int)(cell - (uint32_t)ei * (uint32_t)C.n);
      pivot<G>(C, t, sn, S, ei, ej, key_val(k));
      if (S.status != 0) return done;
      ++done; ++inRound;
    }
    if (inRound == 0) break;  // nothing prices negative any more (uniform)
  }
  return done;
}

// ---------------------------------------------------------------------------
// Save / restore the tree between scheduling steps (server mode).
// state layout (int32): parent, depth, pre, size, slot, flow, dual  (7 x V)
// ---------------------------------------------------------------------------
// Pivot set-up after the tree is built or loaded: vat (vertex at each pre-order
// position) and the cycle-depth minima.
template <int G>
__device__ void tree_index(const Ctx& C, const TreeLds& t, Shared& S) {
  for (int x = grank<G>(); x < C.V; x += G) t.vat[t.pre(x)] = (int16_t)x;
  if (grank<G>() == 0) { S.dlo[0] = 0x7fffffff; S.dlo[1] = 0x7fffffff; S.dsel = 0; }
  gsync<G>();
}

template <int G>
__device__ void save_state(const Ctx& C, const TreeLds& t, gint* st) {
  const int V = C.V;
  for (int x = grank<G>(); x < V; x += G) {
    st[x] = t.parent[x]; st[V + x] = t.depth[x]; st[2 * V + x] = t.pre(x); st[3 * V + x] = t.size(x);
    st[4 * V + x] = t.slot[x]; st[5 * V + x] = t.flow[x]; st[6 * V + x] = dual_of(t, C.m, x);
  }
}
// Duals only, at the state's dual array (6V..7V), for the sweep grid (tree in global).
template <int G>
__device__ void save_duals(const Ctx& C, const TreeLds& t, gint* st) {
  const int V = C.V;
  for (int x = grank<G>(); x < V; x += G) st[6 * V + x] = dual_of(t, C.m, x);
}
template <int G>
__device__ void load_state(const Ctx& C, const TreeLds& t, gcint* st) {
  const int V = C.V;
  for (int x = grank<G>(); x < V; x += G) {
    t.parent[x] = (int16_t)st[x]; t.depth[x] = (int16_t)st[V + x]; t.pre(x) = (int16_t)st[2 * V + x];
    t.size(x) = (int16_t)st[3 * V + x]; t.slot[x] = (int16_t)st[4 * V + x]; t.flow[x] = st[5 * V + x];
    set_dual(t, C.m, x, st[6 * V + x]);
  }
  gsync<G>();
}

// Final outputs: basis by slot, duals, objective, count of positive flows.
template <int G>
__device__ void write_outputs(const Ctx& C, const TreeLds& t, DevInstance& D, Shared& S) {
  const int m = C.m, V = C.V;
  gint* dual = gl(uniform_ptr(D.dual));
  gint* brow = gl(uniform_ptr(D.basis_row));
  gint* bcol = gl(uniform_ptr(D.basis_col));
  gint* bflow = gl(uniform_ptr(D.basis_flow));
  int64_t obj = 0, nz = 0;
  for (int x = grank<G>(); x < V; x += G) {
    dual[x] = dual_of(t, m, x);
    if (x == 0) continue;
    const int p = t.parent[x];
    const int r = x < m ? x : p, c = (x < m ? p : x) - m;
    const int k = t.slot[x];
    if (k < 0 || k >= V - 1 || r < 0 || r >= m || c < 0 || c >= C.n) continue;  // corrupt tree guard
    brow[k] = r; bcol[k] = c; bflow[k] = t.flow[x];
    obj += (int64_t)t.flow[x] * (int64_t)cost_at(C, r, c);
    nz += t.flow[x] != 0;
  }
  obj = group_sum_i64<G>(obj, S);
  nz = group_sum_i64<G>(nz, S);
  if (grank<G>() == 0) { D.objective = obj; D.n_flows = (int32_t)nz; }
}

// ---------------------------------------------------------------------------
// Solve (or advance) one instance with one group.  `region` is the group's LDS region
// of `region_bytes` bytes, starting with its Shared block.
// ---------------------------------------------------------------------------
// kMulti: this instantiation carries the multi-pivot rule (kernels launched for it);
// the others leave its code out, which keeps their register count down.
template <int G, bool kMulti>
__device__ void solve_instance(DevInstance& D, char* region, size_t region_bytes, const EngineParams& P) {
  if (D.phase == 2) return;
  Shared& S = *(Shared*)region;
  char* smem = region + shared_bytes();
  const size_t avail = region_bytes - shared_bytes();
  Ctx C;
  C.cost = gl(uniform_ptr(D.cost)); C.costT = gl(uniform_ptr(D.costT));
  C.m = uniform(D.m); C.n = uniform(D.n); C.V = C.m + C.n; C.ld = uniform(D.ld); C.ldT = uniform(D.ldT);
  gcint* supply = gl(uniform_ptr(D.supply));
  gcint* demand = gl(uniform_ptr(D.demand));
  gint* brow = gl(uniform_ptr(D.basis_row));
  gint* bcol = gl(uniform_ptr(D.basis_col));
  gint* bflow = gl(uniform_ptr(D.basis_flow));
  const int m = C.m, n = C.n, V = C.V;
  const TreeLds t = carve_tree(smem, m, n);
  const size_t treeBytes = tree_lds_bytes(m, n);
  int32_t* aux = (treeBytes + 4 * aux_ints(m, n) <= avail) ? (int32_t*)(smem + treeBytes)
                                                          : uniform_ptr(D.scratch);
  const bool auxLds = treeBytes + 4 * aux_ints(m, n) <= avail;
  const Snap sn = carve_snap<Snap>(aux, V);
  const SnapL snL = carve_snap<SnapL>(aux, V);
  // 256- and 512-thread engines keep the pivot's snapshot in LDS only (one pivot code
  // path, fewer live registers); the host gives them only instances whose tree and
  // snapshot fit their LDS
  constexpr bool kSnapLdsOnly = (G == 256 || G == 512) && !kMulti;
  if constexpr (kSnapLdsOnly) {
    if (!auxLds) {
      if (grank<G>() == 0) { D.status = 6; D.phase = 2; D.iterations = 0; }
      return;
    }
  }
  int64_t iter = D.iterations;
  int64_t sweeps = D.sweeps;
  int32_t mode = D.phase == 0 ? 0 : D.mode;  // switch rule: 1 once multi-pivoting stopped
  const bool lead = grank<G>() == 0;
  if (lead) {
    S.status = 0;
    for (int k = 0; k < 6; ++k) S.tk[k] = D.phase == 0 ? 0 : D.ticks[k];
    if (D.phase == 0) D.t_start = now_ticks();
  }
  uint64_t tp = now_ticks();
  gsync<G>();
  if (D.phase == 0) {
    iter = 0; sweeps = 0;
    // Vogel and the tree build cannot stop at a slice deadline: count this group as busy
    // meanwhile (relaxed device-scope counter; it only decides when pivoting groups stop)
    if (P.busy && lead) __hip_atomic_fetch_add(P.busy, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (C.ld > n) {  // padding cells of every row price positive in the sweep (kPadCost)
      const int padw = C.ld - n;
      FLOCK_GLOBAL int32_t* cw = (FLOCK_GLOBAL int32_t*)C.cost;  // the library's own copy
      for (int e = grank<G>(); e < m * padw; e += G) cw[(uint32_t)((e / padw) * C.ld + n + e % padw)] = kPadCost;
    }
    if (P.initial == 1) {
      nwc<G>(C, supply, demand, brow, bcol, bflow);
    } else {
      transpose<G>(C, (int32_t*)smem);
      __threadfence_block();
      gsync<G>();
      vogel<G>(C, supply, demand, smem, avail, S, brow, bcol, bflow, uniform_ptr(D.scratch));
    }
    __threadfence_block();
    gsync<G>();
    if (S.status != 0) {
      if (grank<G>() == 0) { D.status = S.status; D.phase = 2; D.iterations = 0; }
      if (P.busy && lead) __hip_atomic_fetch_add(P.busy, -1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return;
    }
    { const uint64_t t2 = now_ticks(); if (lead) S.tk[0] += (int64_t)(t2 - tp); tp = t2; }
    const bool ok = build_tree<G>(C, brow, bcol, bflow, t, aux, S);
    { const uint64_t t2 = now_ticks(); if (lead) S.tk[1] += (int64_t)(t2 - tp); tp = t2; }
    if (P.busy && lead) __hip_atomic_fetch_add(P.busy, -1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (!ok) {
      if (grank<G>() == 0) { D.status = 2; D.phase = 2; D.iterations = 0; }
      return;
    }
  } else {
    load_state<G>(C, t, gl(uniform_ptr(D.state)));
  }
  for (int j = n + grank<G>(); j < C.ld; j += G) t.dv[j] = 0;  // padding duals (published below)
  tree_index<G>(C, t, S);
  const int64_t maxIt = P.max_iterations > 0 ? P.max_iterations : 50LL * V + 100000;
  int32_t status = 0;
  bool finished = false;
  int64_t stepDone = 0;
  constexpr int K = kMulti ? 1 : CandsPerRank<G>::value;
  CandSet<K> cands;
  cands.clear();
  int minor = 0;
  bool haveList = false;
  for (;;) {
    uint64_t key;
    tp = now_ticks();
    if (P.pricing == 1 && haveList && minor < P.minor_limit) {
      key = group_min1_u64<G, 0>(reprice_set(C, t, cands), S);
      if (lead) S.tk[3] += (int64_t)(now_ticks() - tp);
      if (key == kNoKey || key_val(key) >= 0) { haveList = false; continue; }
      ++minor;
    } else {
      key = sweep_argmin<G, K>(C, t, S, cands);
      if (lead) S.tk[2] += (int64_t)(now_ticks() - tp);
      ++sweeps;
      haveList = P.pricing == 1;
      minor = 0;
      if (key == kNoKey || key_val(key) >= 0) { finished = true; break; }
    }
    if (iter >= maxIt) { status = 1; finished = true; break; }
    if (P.step_pivots > 0 && stepDone >= P.step_pivots) break;
    if (P.deadline && stepDone > 0) {  // one decision for the whole group; >= 1 pivot per call
      if (lead)
        S.stop = now_ticks() >= P.deadline &&
                 (!P.busy || __hip_atomic_load(P.busy, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0);
      gsync<G>();
      if (S.stop) break;
    }
    if constexpr (kMulti) {  // multi-pivot: every accepted disjoint cycle of this sweep
      if (mode == 0) {
        // the switch rule makes one claim round per sweep and leaves multi-pivoting for
        // good once a round pivots fewer than 10 times (switch_model_parallel.cu:182-184)
        const bool sw = P.pricing == 4;
        const int done = multi_pivot<G>(C, t, sn, (uint32_t*)aux, S, cands.get(0), iter, maxIt,
                                        sw ? 1 : P.minor_limit);
        if (S.status != 0) { status = S.status; finished = true; break; }
        iter += done; stepDone += done;
        if (sw && done < 10) mode = 1;
        continue;
      }
    }
    const uint32_t cell = key_idx(key);
    const int ei = (int)(cell / (uint32_t)n), ej = (int)(cell - (uint32_t)ei * (uint32_t)n);
    if (kSnapLdsOnly || auxLds) pivot<G, (G < 1024)>(C, t, snL, S, ei, ej, key_val(key));
    else pivot<G, (G < 1024)>(C, t, sn, S, ei, ej, key_val(key));
    if (S.status != 0) { status = S.status; finished = true; break; }
    ++iter; ++stepDone;
  }
  if (finished) {
    write_outputs<G>(C, t, D, S);
    if (grank<G>() == 0) { D.status = status; D.phase = 2; D.iterations = iter; D.sweeps = sweeps; D.mode = mode; }
  } else {
    save_state<G>(C, t, gl(uniform_ptr(D.state)));
    if (grank<G>() == 0) { D.phase = 1; D.iterations = iter; D.sweeps = sweeps; D.mode = mode; }
  }
  if (lead) {
    for (int k = 0; k < 6; ++k) D.ticks[k] = S.tk[k];
    if (finished) D.t_done = now_ticks();
  }
  gsync<G>();
}

// ---------------------------------------------------------------------------
// Large-instance path: the solve loop split into kernels that communicate only through
// global memory between launches (the host captures them in a graph):
//   setup  (one workgroup): transpose, initial basis, tree; tree and duals -> state
//   sweep  (a grid, one tile of the matrix per workgroup; sweep_tile): tile argmin key
//          and per-wavefront best cells, duals read from the saved state
//   round  (one workgroup): state -> LDS; argmin over the tile keys; one pivot (Dantzig)
//          or a pivot plus re-priced candidate pivots (candidate list); state <- LDS,
//          or the outputs when no reduced cost is negative.
// D.phase is 1 while the loop runs and 2 when done; later kernels return at once.
// ---------------------------------------------------------------------------
template <int G>
__device__ void large_setup(DevInstance& D, char* region, size_t region_bytes, const EngineParams& P,
                            gcshort* books) {
  Shared& S = *(Shared*)region;
  char* smem = region + shared_bytes();
  const size_t avail = region_bytes - shared_bytes();
  Ctx C;
  C.cost = gl(uniform_ptr(D.cost)); C.costT = gl(uniform_ptr(D.costT));
  C.m = uniform(D.m); C.n = uniform(D.n); C.V = C.m + C.n; C.ld = uniform(D.ld); C.ldT = uniform(D.ldT);
  gint* brow = gl(uniform_ptr(D.basis_row));
  gint* bcol = gl(uniform_ptr(D.basis_col));
  gint* bflow = gl(uniform_ptr(D.basis_flow));
  // tree_global: the tree lives in the instance's state buffer for the whole solve
  // (instances whose tree does not fit LDS); otherwise in LDS, saved to state per round
  const bool treeGlobal = P.tree_global != 0;
  char* gtree = (char*)uniform_ptr(D.state);
  const TreeLds t = carve_tree(treeGlobal ? gtree : smem, C.m, C.n);
  const size_t treeBytes = tree_lds_bytes(C.m, C.n);
  int32_t* aux = (!treeGlobal && treeBytes + 4 * aux_ints(C.m, C.n) <= avail) ? (int32_t*)(smem + treeBytes)
                                                                              : uniform_ptr(D.scratch);
  const bool lead = grank<G>() == 0;
  if (lead) {
    S.status = 0;
    for (int k = 0; k < 6; ++k) S.tk[k] = 0;
    D.t_start = now_ticks();
  }
  uint64_t tp = now_ticks();
  gsync<G>();
  if (P.initial == 1) {
    nwc<G>(C, gl(uniform_ptr(D.supply)), gl(uniform_ptr(D.demand)), brow, bcol, bflow);
  } else {
    if (!books) {  // with books the grid kernels have transposed the matrix already
      transpose<G>(C, (int32_t*)smem);
      __threadfence_block();
      gsync<G>();
    }
    vogel<G, true>(C, gl(uniform_ptr(D.supply)), gl(uniform_ptr(D.demand)), smem, avail, S, brow, bcol, bflow,
             uniform_ptr(D.scratch), books, treeGlobal ? gtree : nullptr);
  }
  __threadfence_block();
  gsync<G>();
  bool ok = S.status == 0;
  { const uint64_t t2 = now_ticks(); if (lead) S.tk[0] += (int64_t)(t2 - tp); tp = t2; }
  if (ok) ok = build_tree<G>(C, brow, bcol, bflow, t, aux, S);
  { const uint64_t t2 = now_ticks(); if (lead) S.tk[1] += (int64_t)(t2 - tp); tp = t2; }
  if (ok) {
    if (treeGlobal) {  // the tree stays where it is; the sweep grid reads the duals
      tree_index<G>(C, t, S);
      save_duals<G>(C, t, gl(uniform_ptr(D.state)));
    } else {
      save_state<G>(C, t, gl(uniform_ptr(D.state)));
    }
  }
  if (lead) {
    for (int k = 0; k < 6; ++k) D.ticks[k] = S.tk[k];
    D.iterations = 0; D.sweeps = 0; D.mode = 0;
    if (ok) { D.status = 0; D.phase = 1; }
    else { D.status = S.status != 0 ? S.status : 2; D.phase = 2; D.t_done = now_ticks(); }
  }
}

template <int G, bool kMulti>
__device__ void large_round(DevInstance& D, char* region, size_t region_bytes, const EngineParams& P,
                            const uint64_t* keys, const int32_t* cands, int nblocks) {
  if (D.phase != 1) return;
  Shared& S = *(Shared*)region;
  char* smem = region + shared_bytes();
  const size_t avail = region_bytes - shared_bytes();
  Ctx C;
  C.cost = gl(uniform_ptr(D.cost)); C.costT = gl(uniform_ptr(D.costT));
  C.m = uniform(D.m); C.n = uniform(D.n); C.V = C.m + C.n; C.ld = uniform(D.ld); C.ldT = uniform(D.ldT);
  const int n = C.n, tid = grank<G>();
  const bool treeGlobal = P.tree_global != 0;  // see large_setup
  const TreeLds t = carve_tree(treeGlobal ? (char*)uniform_ptr(D.state) : smem, C.m, n);
  const size_t treeBytes = tree_lds_bytes(C.m, n);
  const bool auxLds = !treeGlobal && treeBytes + 4 * aux_ints(C.m, C.n) <= avail;
  int32_t* aux = auxLds ? (int32_t*)(smem + treeBytes) : uniform_ptr(D.scratch);
  const Snap sn = carve_snap<Snap>(aux, C.V);
  const SnapL snL = carve_snap<SnapL>(aux, C.V);
  const bool lead = tid == 0;
  if (lead) {
    S.status = 0;
    for (int k = 0; k < 6; ++k) S.tk[k] = D.ticks[k];
  }
  if (treeGlobal) {  // the tree (vat included) persists in global memory between rounds
    if (tid == 0) { S.dlo[0] = 0x7fffffff; S.dlo[1] = 0x7fffffff; S.dsel = 0; }
    gsync<G>();
  } else {
    load_state<G>(C, t, gl(uniform_ptr(D.state)));
    tree_index<G>(C, t, S);
  }
  uint64_t key = kNoKey;
  for (int b = tid; b < nblocks; b += G) { const uint64_t k = keys[b]; key = k < key ? k : key; }
  key = group_min_u64<G>(key, S);
  // candidate list: the tiles' candidates dealt round-robin to the ranks; the i-th one a
  // rank is dealt goes to its slot i % K, and each slot keeps the one that prices lowest
  // with the current duals (the candidate rule re-prices the whole set in its minor
  // iterations; the multi-pivot rules take the rank's best)
  constexpr int K = CandsPerRank<G>::value;
  CandSet<K> cs;
  cs.clear();
  Cand cand;
  cand.row = -1; cand.col = 0; cand.cost = 0;
  if (P.pricing >= 1) {
    uint64_t bestk[K];
#pragma unroll
    for (int k = 0; k < K; ++k) bestk[k] = kNoKey;
    uint64_t bestc = kNoKey;
    int slot = 0;
    for (int q = tid; q < nblocks * kTileCands; q += G) {
      const int32_t* c = cands + 3 * q;
      Cand x;
      x.row = c[0];
      x.col = x.row < 0 ? 0 : c[1];
      x.cost = x.row < 0 ? 0 : c[2];
      const uint64_t k = x.row < 0 ? kNoKey : reprice(C, t, x);
#pragma unroll
      for (int e = 0; e < K; ++e)
        if (e == slot && k < bestk[e]) { bestk[e] = k; cs.set(e, x); }
      if (k < bestc) { bestc = k; cand = x; }
      slot = slot + 1 < K ? slot + 1 : 0;
    }
  }
  // multi-pivot rules: rank 0 holds the sweep's argmin cell, so every round pivots at
  // least once whatever tiles the ranks' candidates came from (the grid can have more
  // tiles than a rank per wavefront candidate)
  if ((P.pricing == 2 || P.pricing == 4) && tid == 0 && key != kNoKey && key_val(key) < 0) {
    const uint32_t cell = key_idx(key);
    cand.row = (int)(cell / (uint32_t)n);
    cand.col = (int)(cell - (uint32_t)cand.row * (uint32_t)n);
    cand.cost = cost_at(C, cand.row, cand.col);
  }
  int64_t iter = D.iterations;
  const int64_t maxIt = P.max_iterations > 0 ? P.max_iterations : 50LL * C.V + 100000;
  int32_t status = 0;
  bool finished = key == kNoKey || key_val(key) >= 0;
  bool multiDone = false;
  int32_t mode = D.mode;  // switch rule: 1 once multi-pivoting stopped
  if constexpr (kMulti) {  // multi-pivot rounds on the tiles' candidates
    if (!finished && mode == 0) {
      multiDone = true;
      if (iter >= maxIt) { status = 1; finished = true; }
      else {
        const bool sw = P.pricing == 4;
        const int done = multi_pivot<G>(C, t, sn, (uint32_t*)aux, S, cand, iter, maxIt, sw ? 1 : P.minor_limit);
        iter += done;
        if (S.status != 0) { status = S.status; finished = true; }
        if (sw && done < 10) mode = 1;
      }
    }
  }
  if (!finished && !multiDone) {
    int minor = 0;
    for (;;) {
      if (iter >= maxIt) { status = 1; finished = true; break; }
      const uint32_t cell = key_idx(key);
      const int ei = (int)(cell / (uint32_t)n), ej = (int)(cell - (uint32_t)ei * (uint32_t)n);
      if (auxLds) pivot<G, true>(C, t, snL, S, ei, ej, key_val(key));
      else pivot<G, true>(C, t, sn, S, ei, ej, key_val(key));
      if (S.status != 0) { status = S.status; finished = true; break; }
      ++iter;
      if (P.pricing != 1 || ++minor >= P.minor_limit) break;  // Dantzig: one pivot per sweep
      const uint64_t tp = now_ticks();
      key = group_min1_u64<G, 0>(reprice_set(C, t, cs), S);
      if (lead) S.tk[3] += (int64_t)(now_ticks() - tp);
      if (key == kNoKey || key_val(key) >= 0) break;  // list exhausted: next full sweep
    }
  }
  if (finished) {
    write_outputs<G>(C, t, D, S);
  } else if (treeGlobal) {
    gsync<G>();  // the last pivot's tree writes before the duals are read back
    save_duals<G>(C, t, gl(uniform_ptr(D.state)));
  } else {
    save_state<G>(C, t, gl(uniform_ptr(D.state)));
  }
  if (lead) {
    for (int k = 0; k < 6; ++k) D.ticks[k] = S.tk[k];
    D.iterations = iter;
    D.mode = mode;
    D.sweeps = D.sweeps + 1;
    if (finished) { D.status = status; D.phase = 2; D.t_done = now_ticks(); }
  }
}

// Arguments of the sweep grid for instance D (duals at its state's dual array).
__host__ __device__ inline SweepArgs sweep_args(const DevInstance& D, int V, uint64_t* keys, int32_t* cands,
                                                uint32_t vpt, const int32_t* phase) {
  SweepArgs A;
  A.cost = D.cost; A.dual = D.state + 6 * (size_t)V; A.phase = phase; A.keys = keys; A.cands = cands;
  A.m = D.m; A.n = D.n; A.ld = D.ld; A.vpt = vpt;
  return A;
}

// ---------------------------------------------------------------------------
// Persistent work-queue worker (the body of k_solve_queue for workgroup `block` of
// `grid`): every group of G threads (one wavefront or the whole workgroup) pops
// instance indices from `head` and solves them one after another.  With a slice
// deadline no instance is started after it.
// ---------------------------------------------------------------------------
template <int T, int G, bool kMulti>
__device__ void queue_worker(DevInstance* insts, const int32_t* order, int32_t count, int32_t* head,
                             const EngineParams& P, char* smem) {
  const size_t region_bytes = (size_t)P.lds_bytes;
  char* region = smem + (size_t)(threadIdx.x / G) * region_bytes;
  Shared& S = *(Shared*)region;
  EngineParams Q = P;
  Q.deadline = P.slice_ticks > 0 ? uniform64(*P.clock0 + P.slice_ticks) : 0;
  for (;;) {
    int idx;
    if constexpr (G == kWave) {
      int v = 0;
      if (grank<G>() == 0) v = (Q.deadline && now_ticks() >= Q.deadline) ? count : atomicAdd(head, 1);
      idx = __shfl(v, 0, kWave);
    } else {
      if (threadIdx.x == 0) S.idx = (Q.deadline && now_ticks() >= Q.deadline) ? count : atomicAdd(head, 1);
      __syncthreads();
      idx = S.idx;
      __syncthreads();
    }
    if (idx >= count) break;
    // the largest instances set the launch's makespan: their waves get issue priority
    // over co-resident workgroups' (uniform per group)
    if (idx < Q.prio_count) __builtin_amdgcn_s_setprio(3);
    else __builtin_amdgcn_s_setprio(0);
    solve_instance<G, kMulti>(insts[order[idx]], region, region_bytes, Q);
  }
}

}  // namespace flock
