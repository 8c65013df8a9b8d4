// transport_oracle.cpp -- CPU restatement of the reference's sequential
// transportation-simplex path, in exact int64 arithmetic.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
// product (flock_amd/) never links, calls or falls back to it.
//
// What it restates (reference = researchgroup-zx93920/flock, "pure_sequencial"
// configuration: BFS_METHOD "vam_host", REDUCED_COST_MODE "sequencial",
// SEQ_CYCLE_SEARCH "dfs"):
//   * oracle_vogel      <- find_vogel_bfs_sequencial
//                          src/parallel_simplex_method/IBFS_vogel.cu:546-769
//   * oracle_nwc        <- find_nw_corner_bfs_seq
//                          src/parallel_simplex_method/IBFS_nwc.cu:3-34
//   * tree_duals        <- UV_METHOD::find_dual_using_host_bfs
//                          src/parallel_simplex_method/parallel_uv_method/DUAL_solver.cu:34-63
//   * sweep (in solve)  <- uvModel_parallel::get_reduced_costs, sequencial branch
//                          src/parallel_simplex_method/parallel_uv_method/uv_model_parallel.cu:178-200
//   * pivot (in solve)  <- UV_METHOD::perform_a_sequencial_pivot + execute_pivot_on_host
//                          + do_flow_adjustment_on_host
//                          src/parallel_simplex_method/parallel_uv_method/PIVOT_uv.cu:156-308,318-401
//   * loop + objective  <- uvModel_parallel::execute
//                          src/parallel_simplex_method/parallel_uv_method/uv_model_parallel.cu:227-395
//
// Deliberate, documented deviations from the float reference (all exact-integer):
//   1. Costs, flows and duals are integers (int64 here).  Zero flows stay 0 instead of
//      the 1e-6 perturbation (parallel_kernels.cu:59-63); the zero-flow arc stays basic.
//   2. Optimality test is exact `r < 0` instead of `r < 0 && |r| > 0.01`
//      (PIVOT_uv.cu:323); on integer costs r <= -1 whenever it is negative.
//   3. Vogel penalties are exact integer differences.  A line with one uncovered cell
//      has penalty INT32_MAX - c (the reference computes (float)INT_MAX - c, which
//      float-rounds and, for c < 64, overflows the int conversion).
//   4. Vogel tie res_d == res_s eliminates the column (IBFS_vogel.cu:690-703) EXCEPT when
//      that column is the last uncovered one; then the row is eliminated.  The reference
//      leaves that case undefined (it reads an uninitialised index on the next step).
//   5. The iteration budget is an option; running out is reported as a status, never
//      as "optimal" (the reference stops silently at MAX_ITERATIONS, uv_model_parallel.cu:301).
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>
#include <chrono>

namespace {

constexpr int64_t kBig = 2147483647;  // INT_MAX sentinel of IBFS_vogel.cu:590-591

enum Status : int {
  OK = 0,
  E_ITERATION_LIMIT = 1,
  E_FOREST = 2,
  E_UNBALANCED = 3,
  E_INVALID = 4,
};

struct Basis {
  // Slot k holds basic cell (row[k], col[k]) with flow[k]; slots are numbered in
  // Vogel allocation order (flows[counter], IBFS_vogel.cu:693,702) and an entering arc
  // re-uses the slot of the leaving arc (PIVOT_uv.cu:254-261).
  std::vector<int32_t> row, col;
  std::vector<int64_t> flow;
};

// ---------------------------------------------------------------------------
// Vogel's approximation, host variant (IBFS_vogel.cu:546-769).
// Penalties of one kind (rows / columns) are recomputed only after a line of that kind
// was eliminated (prev_row / prev_col, :588-642), so the other kind's penalties can be
// stale -- restated as is.  Argmax uses >= (last maximum wins, :654); the min-cost cell
// uses strict < (first minimum, :673 and :723).
// ---------------------------------------------------------------------------
int vogel(int m, int n, const int32_t* c, const int32_t* s, const int32_t* d, Basis& B) {
  const int V = m + n;
  std::vector<int64_t> rs(s, s + m), rd(d, d + n);
  std::vector<char> rowCov(m, 0), colCov(n, 0);
  std::vector<int64_t> diff(V, 0);
  int covRows = 0, covCols = 0;
  bool prevRow = true, prevCol = true;
  B.row.assign(V - 1, 0); B.col.assign(V - 1, 0); B.flow.assign(V - 1, 0);

  while (covRows + covCols < V - 1) {
    if (prevRow) {
      for (int i = 0; i < m; ++i) {
        if (rowCov[i]) { diff[i] = -1; continue; }
        int64_t t1 = kBig, t2 = kBig;
        const int32_t* ci = c + (size_t)i * n;
        for (int j = 0; j < n; ++j) {
          if (colCov[j]) continue;
          int64_t e = ci[j];
          if (e <= t1) { t2 = t1; t1 = e; } else if (e <= t2) { t2 = e; }
        }
        diff[i] = t2 - t1;
      }
      prevRow = false;
    }
    if (prevCol) {
      for (int j = 0; j < n; ++j) {
        if (colCov[j]) { diff[m + j] = -1; continue; }
        int64_t t1 = kBig, t2 = kBig;
        for (int i = 0; i < m; ++i) {
          if (rowCov[i]) continue;
          int64_t e = c[(size_t)i * n + j];
          if (e <= t1) { t2 = t1; t1 = e; } else if (e <= t2) { t2 = e; }
        }
        diff[m + j] = t2 - t1;
      }
      prevCol = false;
    }
    int64_t best = INT64_MIN; int arg = -1;
    for (int k = 0; k < V; ++k) if (diff[k] >= best) { best = diff[k]; arg = k; }

    const int counter = covRows + covCols;
    int r, q;  // chosen cell (row r, column q)
    if (arg >= m) {
      q = arg - m; r = -1; int64_t ct = kBig;
      for (int i = 0; i < m; ++i) {
        int64_t e = c[(size_t)i * n + q];
        if (e < ct && !rowCov[i]) { ct = e; r = i; }
      }
    } else {
      r = arg; q = -1; int64_t ct = kBig;
      const int32_t* cr = c + (size_t)r * n;
      for (int j = 0; j < n; ++j) {
        int64_t e = cr[j];
        if (e < ct && !colCov[j]) { ct = e; q = j; }
      }
    }
    if (r < 0 || q < 0) return E_INVALID;  // unreachable with the last-line rule below
    B.row[counter] = r; B.col[counter] = q;
    bool coverRow = rd[q] > rs[r];
    if (!coverRow && covCols == n - 1) coverRow = true;  // deviation 4 (last column)
    if (coverRow) {
      B.flow[counter] = rs[r]; rd[q] -= rs[r]; rs[r] = 0;
      rowCov[r] = 1; prevRow = true; covRows++;
    } else {
      B.flow[counter] = rd[q]; rs[r] -= rd[q]; rd[q] = 0;
      colCov[q] = 1; prevCol = true; covCols++;
    }
  }
  return OK;
}

// North-west corner rule (IBFS_nwc.cu:3-34): walk from cell (0,0); demand >= supply
// moves to the next row (:17-23), otherwise to the next column.  On the last row a tie
// moves to the next column instead, so exactly m+n-1 cells are allocated (the reference
// stops early there and leaves the remaining slots unset).
int nwc(int m, int n, const int32_t* s, const int32_t* d, Basis& B) {
  const int V = m + n;
  std::vector<int64_t> rs(s, s + m), rd(d, d + n);
  B.row.assign(V - 1, 0); B.col.assign(V - 1, 0); B.flow.assign(V - 1, 0);
  int i = 0, j = 0;
  for (int k = 0; k < V - 1; ++k) {
    if (i >= m || j >= n) return E_INVALID;
    B.row[k] = i; B.col[k] = j;
    if ((rd[j] >= rs[i] && i < m - 1) || j == n - 1) {
      B.flow[k] = rs[i]; rd[j] -= rs[i]; rs[i] = 0; ++i;
    } else {
      B.flow[k] = rd[j]; rs[i] -= rd[j]; rd[j] = 0; ++j;
    }
  }
  return OK;
}

// Spanning-tree bookkeeping rebuilt each iteration: adjacency by slot, then a BFS from
// vertex 0 (the reference's root, DUAL_solver.cu:44-46) giving parent slot, depth and
// duals.  Vertex ids: rows 0..m-1, columns m..m+n-1.
struct Tree {
  std::vector<int32_t> adjStart, adj;          // CSR: slot ids per vertex
  std::vector<int32_t> parentSlot, parentVtx, depth, queue;
  std::vector<int64_t> dual;                   // u (rows) then v (columns)
};

int tree_duals(int m, int n, const int32_t* c, const Basis& B, Tree& T) {
  const int V = m + n;
  T.adjStart.assign(V + 1, 0);
  for (int k = 0; k < V - 1; ++k) { T.adjStart[B.row[k] + 1]++; T.adjStart[m + B.col[k] + 1]++; }
  for (int v = 0; v < V; ++v) T.adjStart[v + 1] += T.adjStart[v];
  T.adj.assign(2 * (V - 1), 0);
  std::vector<int32_t> fill(T.adjStart.begin(), T.adjStart.end() - 1);
  for (int k = 0; k < V - 1; ++k) { T.adj[fill[B.row[k]]++] = k; T.adj[fill[m + B.col[k]]++] = k; }
  T.parentSlot.assign(V, -1); T.parentVtx.assign(V, -1); T.depth.assign(V, -1);
  T.dual.assign(V, 0); T.queue.assign(V, 0);
  int head = 0, tail = 0;
  T.queue[tail++] = 0; T.depth[0] = 0;
  while (head < tail) {
    int p = T.queue[head++];
    for (int e = T.adjStart[p]; e < T.adjStart[p + 1]; ++e) {
      int k = T.adj[e];
      int ch = (p < m) ? m + B.col[k] : B.row[k];
      if (T.depth[ch] >= 0) continue;
      T.depth[ch] = T.depth[p] + 1; T.parentSlot[ch] = k; T.parentVtx[ch] = p;
      // x_child = c(row,col) - x_parent  (DUAL_solver.cu:58-60)
      T.dual[ch] = (int64_t)c[(size_t)B.row[k] * n + B.col[k]] - T.dual[p];
      T.queue[tail++] = ch;
    }
  }
  return tail == V ? OK : E_FOREST;
}

}  // namespace

extern "C" {

struct oracle_options {
  int64_t max_iterations;   // <= 0: 50*(m+n) + 100000
  int32_t initial;          // 0 = Vogel (vam_host), 1 = north-west corner
  int32_t record_limit;     // entering cells recorded into trace_* (0 = none)
};

struct oracle_result {
  int64_t objective;
  int64_t iterations;       // pivots performed (the reference's totalIterations counts
                            // the final optimality check too: iterations + 1)
  int32_t status;
  int32_t nonzero_flows;    // basic arcs with positive flow (the reference's active_flows)
  double  solve_ms;         // Vogel + loop, as totalSolveTime (uv_model_parallel.cu:237-369)
  double  vogel_ms;
};

// Vogel BFS alone: rows/cols/flows of the m+n-1 basic cells in allocation order.
int oracle_vogel(int m, int n, const int32_t* cost, const int32_t* supply, const int32_t* demand,
                 int32_t* rows, int32_t* cols, int64_t* flows) {
  if (m < 1 || n < 1) return E_INVALID;
  Basis B;
  int st = vogel(m, n, cost, supply, demand, B);
  if (st) return st;
  std::memcpy(rows, B.row.data(), sizeof(int32_t) * (m + n - 1));
  std::memcpy(cols, B.col.data(), sizeof(int32_t) * (m + n - 1));
  std::memcpy(flows, B.flow.data(), sizeof(int64_t) * (m + n - 1));
  return OK;
}

int oracle_nwc(int m, int n, const int32_t* supply, const int32_t* demand,
               int32_t* rows, int32_t* cols, int64_t* flows) {
  if (m < 1 || n < 1) return E_INVALID;
  Basis B;
  int st = nwc(m, n, supply, demand, B);
  if (st) return st;
  std::memcpy(rows, B.row.data(), sizeof(int32_t) * (m + n - 1));
  std::memcpy(cols, B.col.data(), sizeof(int32_t) * (m + n - 1));
  std::memcpy(flows, B.flow.data(), sizeof(int64_t) * (m + n - 1));
  return OK;
}

// Full sequential solve.  Outputs (each may be null): basis in slot order, duals u,v,
// and the first record_limit entering cells (row*n+col) with their reduced costs.
int oracle_solve(int m, int n, const int32_t* cost, const int32_t* supply, const int32_t* demand,
                 const oracle_options* opt, oracle_result* res,
                 int32_t* basis_rows, int32_t* basis_cols, int64_t* basis_flows,
                 int64_t* u, int64_t* v, int64_t* trace_cell, int64_t* trace_r) {
  std::memset(res, 0, sizeof(*res));
  if (m < 1 || n < 1) { res->status = E_INVALID; return E_INVALID; }
  int64_t ss = 0, sd = 0;
  for (int i = 0; i < m; ++i) { if (supply[i] < 0) { res->status = E_INVALID; return E_INVALID; } ss += supply[i]; }
  for (int j = 0; j < n; ++j) { if (demand[j] < 0) { res->status = E_INVALID; return E_INVALID; } sd += demand[j]; }
  if (ss != sd) { res->status = E_UNBALANCED; return E_UNBALANCED; }
  const int V = m + n;
  int64_t maxIt = (opt && opt->max_iterations > 0) ? opt->max_iterations : 50LL * V + 100000;
  int recLimit = opt ? opt->record_limit : 0;

  auto t0 = std::chrono::steady_clock::now();
  Basis B;
  int st = (opt && opt->initial == 1) ? nwc(m, n, supply, demand, B) : vogel(m, n, cost, supply, demand, B);
  auto t1 = std::chrono::steady_clock::now();
  res->vogel_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  if (st) { res->status = st; return st; }

  Tree T;
  std::vector<int32_t> pathI, pathJ, path;
  int64_t it = 0;
  st = OK;
  for (;;) {
    st = tree_duals(m, n, cost, B, T);
    if (st) break;
    // Row-major scan, strict < keeps the first minimum (uv_model_parallel.cu:184-194).
    // The per-row minimum is found first so the inner loop vectorises; the first index
    // is located only in a row that beats the running best.
    int64_t minR = kBig; int pr = -1, pc = -1;
    const int64_t* vv = T.dual.data() + m;
    for (int i = 0; i < m; ++i) {
      const int32_t* ci = cost + (size_t)i * n;
      const int64_t ui = T.dual[i];
      int64_t rowMin = INT64_MAX;
      for (int j = 0; j < n; ++j) { int64_t r = (int64_t)ci[j] - ui - vv[j]; rowMin = r < rowMin ? r : rowMin; }
      if (rowMin < minR) {
        for (int j = 0; j < n; ++j) {
          if ((int64_t)ci[j] - ui - vv[j] == rowMin) { minR = rowMin; pr = i; pc = j; break; }
        }
      }
    }
    if (!(minR < 0)) break;  // optimal (PIVOT_uv.cu:323, exact)
    if (it >= maxIt) { st = E_ITERATION_LIMIT; break; }
    if (it < recLimit && trace_cell) { trace_cell[it] = (int64_t)pr * n + pc; trace_r[it] = minR; }

    // Tree path from column vertex J = m+pc to row vertex pr (PIVOT_uv.cu:338-356; the
    // path in a tree is unique, so walking parents to the common ancestor gives the same
    // vertex sequence as the reference's DFS backtracker).
    const int J = m + pc, I = pr;
    pathJ.clear(); pathI.clear();
    int a = J, b = I;
    while (T.depth[a] > T.depth[b]) { pathJ.push_back(a); a = T.parentVtx[a]; }
    while (T.depth[b] > T.depth[a]) { pathI.push_back(b); b = T.parentVtx[b]; }
    while (a != b) { pathJ.push_back(a); a = T.parentVtx[a]; pathI.push_back(b); b = T.parentVtx[b]; }
    // Slots along the path J -> apex -> I, in order.
    path.clear();
    for (int x : pathJ) path.push_back(T.parentSlot[x]);
    for (int k = (int)pathI.size() - 1; k >= 0; --k) path.push_back(T.parentSlot[pathI[k]]);
    if (path.empty()) { st = E_FOREST; break; }
    // Leaving arc: first strict minimum flow over odd edges (1st, 3rd, ... from the
    // column side), PIVOT_uv.cu:280-296.
    int64_t theta = INT64_MAX; int leave = -1;
    for (size_t e = 0; e < path.size(); e += 2) {
      int k = path[e];
      if (B.flow[k] < theta) { theta = B.flow[k]; leave = k; }
    }
    // Flow adjustment (-1)^i * theta along the path (PIVOT_uv.cu:236-244), then the
    // entering arc takes the leaving arc's slot with flow theta (:248-261).
    for (size_t e = 0; e < path.size(); ++e) B.flow[path[e]] += (e % 2 == 0) ? -theta : theta;
    B.row[leave] = pr; B.col[leave] = pc; B.flow[leave] = theta;
    ++it;
  }
  auto t2 = std::chrono::steady_clock::now();
  res->solve_ms = std::chrono::duration<double, std::milli>(t2 - t0).count();
  res->iterations = it;
  res->status = st;
  int64_t obj = 0; int nz = 0;
  for (int k = 0; k < V - 1; ++k) {
    obj += B.flow[k] * (int64_t)cost[(size_t)B.row[k] * n + B.col[k]];
    nz += B.flow[k] != 0;
  }
  res->objective = obj;
  res->nonzero_flows = nz;
  if (basis_rows) std::memcpy(basis_rows, B.row.data(), sizeof(int32_t) * (V - 1));
  if (basis_cols) std::memcpy(basis_cols, B.col.data(), sizeof(int32_t) * (V - 1));
  if (basis_flows) std::memcpy(basis_flows, B.flow.data(), sizeof(int64_t) * (V - 1));
  if ((u || v) && tree_duals(m, n, cost, B, T) == OK) {
    if (u) for (int i = 0; i < m; ++i) u[i] = T.dual[i];
    if (v) for (int j = 0; j < n; ++j) v[j] = T.dual[m + j];
  }
  return st;
}

// Optimality certificate, independent of how the basis was found:
//   bit 0: a flow is negative             bit 1: a row sum differs from its supply
//   bit 2: a column sum differs            bit 3: the m+n-1 cells are not a spanning tree
//   bit 4: some reduced cost is negative   bit 5: given duals disagree with the tree's
// Returns the bit mask (0 = certified optimal) and writes the objective.
int oracle_certify(int m, int n, const int32_t* cost, const int32_t* supply, const int32_t* demand,
                   int nb, const int32_t* rows, const int32_t* cols, const int64_t* flows,
                   const int64_t* u_in, const int64_t* v_in, int64_t* objective) {
  const int V = m + n;
  int bad = 0;
  std::vector<int64_t> rsum(m, 0), csum(n, 0);
  int64_t obj = 0;
  for (int k = 0; k < nb; ++k) {
    if (rows[k] < 0 || rows[k] >= m || cols[k] < 0 || cols[k] >= n) { *objective = 0; return 1 << 3; }
    if (flows[k] < 0) bad |= 1;
    rsum[rows[k]] += flows[k]; csum[cols[k]] += flows[k];
    obj += flows[k] * (int64_t)cost[(size_t)rows[k] * n + cols[k]];
  }
  for (int i = 0; i < m; ++i) if (rsum[i] != supply[i]) bad |= 2;
  for (int j = 0; j < n; ++j) if (csum[j] != demand[j]) bad |= 4;
  *objective = obj;
  if (nb != V - 1) return bad | 8;
  Basis B;
  B.row.assign(rows, rows + nb); B.col.assign(cols, cols + nb); B.flow.assign(flows, flows + nb);
  Tree T;
  if (tree_duals(m, n, cost, B, T) != OK) return bad | 8;  // V-1 edges + connected => tree
  for (int i = 0; i < m && !(bad & 16); ++i)
    for (int j = 0; j < n; ++j)
      if ((int64_t)cost[(size_t)i * n + j] - T.dual[i] - T.dual[m + j] < 0) { bad |= 16; break; }
  if (u_in && v_in) {
    for (int i = 0; i < m; ++i) if (u_in[i] != T.dual[i]) bad |= 32;
    for (int j = 0; j < n; ++j) if (v_in[j] != T.dual[m + j]) bad |= 32;
  }
  return bad;
}

}  // extern "C"
