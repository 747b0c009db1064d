// Seeded synthetic sparsity patterns for the symbolic-factorisation engine.
//
// The reference reads an already-reordered pattern dumped by SuperLU_DIST from three ASCII
// files (gSoFa_SOAComparison/symbfact_warp_centric_new_scheduling.cuh:2567-2616, README.md:27-30).
// No such dump exists at the workload sizes used here, so this file produces the patterns
// itself: 2-D / 3-D stencils, nested-dissection ordering by recursive coordinate bisection,
// Erdos-Renyi graphs in reverse Cuthill-McKee order, and seeded entry drops that make the
// pattern structurally unsymmetric. Every random decision is a pure function of
// (seed, indices) through splitmix64, so a pattern is identical whatever order it is built in.
// Structure only: there are no numeric values anywhere.

#include "gsofa_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

namespace gsofa {

uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

// Counter-based draw: a hash of (seed, a, b); independent of evaluation order.
static inline uint64_t draw(uint64_t seed, uint64_t a, uint64_t b) {
  return splitmix64(splitmix64(seed ^ (a * 0xD1B54A32D192ED03ull)) ^ (b * 0xABC98388FB8FAC03ull));
}

// True with probability ppm / 1e6.
static inline bool bernoulli_ppm(uint64_t h, uint32_t ppm) {
  return (h % 1000000ull) < ppm;
}

namespace {

struct Grid {
  int dims;       // 2 or 3
  int64_t w[3];   // extents along x, y, z
  int64_t n;      // number of points kept: the first n points in row-major order
  void coords(int64_t p, int64_t c[3]) const {
    c[0] = p % w[0];
    int64_t q = p / w[0];
    c[1] = q % w[1];
    c[2] = (dims == 3) ? q / w[1] : 0;
  }
  int64_t index(const int64_t c[3]) const { return c[0] + w[0] * (c[1] + w[1] * c[2]); }
};

// Recursive coordinate bisection. `pts` holds natural (row-major) indices in ascending order.
// Split the longest bounding-box axis at its midpoint plane: points below go left, points on
// the plane form the separator, points above go right. Numbering: left subtree, right subtree,
// then the separator (natural order). Sets of at most `leaf` points are numbered in natural
// order and recorded as ND leaves.
struct NDBuilder {
  const Grid& g;
  int64_t leaf;
  std::vector<int32_t>& perm;      // natural index -> new id
  std::vector<int32_t>& leaf_of;   // new id -> leaf index, or -1 for separator vertices
  std::vector<int64_t>& leaf_lo;   // leaf index -> first new id
  std::vector<int64_t>& leaf_hi;   // leaf index -> one past last new id
  int64_t next = 0;

  void number_natural(const std::vector<int64_t>& pts, int32_t leaf_id) {
    for (int64_t p : pts) {
      perm[p] = (int32_t)next;
      leaf_of[next] = leaf_id;
      ++next;
    }
  }

  void run(std::vector<int64_t>&& pts) {
    // Explicit stack of work items: either "split this set" or "number this separator".
    struct Item { std::vector<int64_t> pts; bool is_sep; };
    std::vector<Item> stack;
    stack.push_back({std::move(pts), false});
    while (!stack.empty()) {
      Item it = std::move(stack.back());
      stack.pop_back();
      if (it.pts.empty()) continue;
      if (it.is_sep) { number_natural(it.pts, -1); continue; }
      if ((int64_t)it.pts.size() <= leaf) {
        int32_t id = (int32_t)leaf_lo.size();
        leaf_lo.push_back(next);
        number_natural(it.pts, id);
        leaf_hi.push_back(next);
        continue;
      }
      int64_t lo[3] = {INT64_MAX, INT64_MAX, INT64_MAX}, hi[3] = {INT64_MIN, INT64_MIN, INT64_MIN};
      int64_t c[3];
      for (int64_t p : it.pts) {
        g.coords(p, c);
        for (int d = 0; d < 3; ++d) { lo[d] = std::min(lo[d], c[d]); hi[d] = std::max(hi[d], c[d]); }
      }
      int axis = 0;
      for (int d = 1; d < g.dims; ++d)
        if (hi[d] - lo[d] > hi[axis] - lo[axis]) axis = d;
      const int64_t mid = lo[axis] + (hi[axis] - lo[axis]) / 2;
      std::vector<int64_t> L, S, R;
      for (int64_t p : it.pts) {
        g.coords(p, c);
        if (c[axis] < mid) L.push_back(p);
        else if (c[axis] == mid) S.push_back(p);
        else R.push_back(p);
      }
      // LIFO: push in reverse of the numbering order (left, right, separator).
      stack.push_back({std::move(S), true});
      stack.push_back({std::move(R), false});
      stack.push_back({std::move(L), false});
    }
  }
};

void build_csr_from_lists(int64_t n, std::vector<std::vector<int32_t>>& rows, HostMatrix& M) {
  M.n = (int32_t)n;
  M.beg.resize(n);
  M.end.resize(n);
  int64_t nnz = 0;
  for (auto& r : rows) {
    std::sort(r.begin(), r.end());
    r.erase(std::unique(r.begin(), r.end()), r.end());
    nnz += (int64_t)r.size();
  }
  M.adj.resize(nnz);
  int64_t pos = 0;
  for (int64_t i = 0; i < n; ++i) {
    M.beg[i] = (int32_t)pos;
    std::copy(rows[i].begin(), rows[i].end(), M.adj.begin() + pos);
    pos += (int64_t)rows[i].size();
    M.end[i] = (int32_t)pos;
  }
}

}  // namespace

int generate_stencil(const GenParams& p, HostMatrix& M) {
  Grid g;
  g.dims = p.dims;
  g.w[0] = p.nx;
  g.w[1] = p.ny;
  g.w[2] = (p.dims == 3) ? p.nz : 1;
  const int64_t full = g.w[0] * g.w[1] * g.w[2];
  g.n = (p.n_keep > 0 && p.n_keep < full) ? p.n_keep : full;
  if (g.n <= 0 || g.n > INT32_MAX) return GSOFA_ERR_INVALID_ARG;
  const int64_t n = g.n;

  std::vector<int32_t> perm(n), leaf_of(n, -1);
  std::vector<int64_t> leaf_lo, leaf_hi;
  if (p.order == GSOFA_ORDER_ND) {
    std::vector<int64_t> pts(n);
    std::iota(pts.begin(), pts.end(), 0);
    NDBuilder b{g, p.nd_leaf > 0 ? p.nd_leaf : 64, perm, leaf_of, leaf_lo, leaf_hi};
    b.run(std::move(pts));
  } else {
    std::iota(perm.begin(), perm.end(), 0);
  }

  // Stencil offsets: 5/9-point in 2-D, 7/27-point in 3-D.
  std::vector<std::array<int, 3>> offs;
  for (int dz = -1; dz <= 1; ++dz)
    for (int dy = -1; dy <= 1; ++dy)
      for (int dx = -1; dx <= 1; ++dx) {
        if (dx == 0 && dy == 0 && dz == 0) continue;
        if (g.dims == 2 && dz != 0) continue;
        const int manhattan = std::abs(dx) + std::abs(dy) + std::abs(dz);
        const bool full_box = (g.dims == 2 && p.points == 9) || (g.dims == 3 && p.points == 27);
        if (!full_box && manhattan != 1) continue;
        offs.push_back({dx, dy, dz});
      }

  std::vector<std::vector<int32_t>> rows(n);
  int64_t c[3], d[3];
  for (int64_t q = 0; q < n; ++q) {
    const int32_t i = perm[q];
    rows[i].push_back(i);  // diagonal always present
    g.coords(q, c);
    for (auto& o : offs) {
      bool inside = true;
      for (int k = 0; k < 3; ++k) {
        d[k] = c[k] + o[k];
        if (d[k] < 0 || d[k] >= g.w[k]) inside = false;
      }
      if (!inside) continue;
      const int64_t r = g.index(d);
      if (r >= n) continue;  // truncated away
      const int32_t j = perm[r];
      const bool lower = j < i;
      bool eligible = (p.drop_side == GSOFA_DROP_ANY) || (p.drop_side == GSOFA_DROP_LOWER && lower) ||
                      (p.drop_side == GSOFA_DROP_UPPER && !lower);
      if (eligible && p.drop_ppm > 0 && bernoulli_ppm(draw(p.seed, (uint64_t)i, (uint64_t)j), p.drop_ppm))
        continue;
      rows[i].push_back(j);
    }
  }
  // Extra entries: row r with r % extra_every == 0 inside an ND leaf gets one entry (r, c),
  // c uniform over the other vertices of that leaf.
  if (p.extra_every > 0 && p.order == GSOFA_ORDER_ND) {
    for (int64_t r = 0; r < n; r += p.extra_every) {
      const int32_t lf = leaf_of[r];
      if (lf < 0) continue;
      const int64_t lo = leaf_lo[lf], sz = leaf_hi[lf] - lo;
      if (sz < 2) continue;
      const uint64_t h = draw(p.seed ^ 0xE17A0000ull, (uint64_t)r, 0x7u);
      int64_t col = lo + (int64_t)(h % (uint64_t)(sz - 1));
      if (col >= r) ++col;  // skip the diagonal
      rows[r].push_back((int32_t)col);
    }
  }
  build_csr_from_lists(n, rows, M);
  return GSOFA_OK;
}

// Reverse Cuthill-McKee over a symmetric adjacency (no self loops). Components are taken in
// order of their lowest-degree unvisited vertex (ties: lowest index); inside a component the
// start is a pseudo-peripheral vertex found by repeated BFS (George-Liu); neighbours are
// visited by increasing degree, ties by index. Returns perm: old index -> new index.
static std::vector<int32_t> rcm_order(int64_t n, const std::vector<int64_t>& ptr, const std::vector<int32_t>& nb) {
  std::vector<int32_t> deg(n);
  for (int64_t v = 0; v < n; ++v) deg[v] = (int32_t)(ptr[v + 1] - ptr[v]);
  std::vector<int32_t> by_deg(n);
  std::iota(by_deg.begin(), by_deg.end(), 0);
  std::stable_sort(by_deg.begin(), by_deg.end(), [&](int32_t a, int32_t b) { return deg[a] < deg[b]; });

  std::vector<int32_t> order;
  order.reserve(n);
  std::vector<char> visited(n, 0);
  std::vector<int32_t> level(n, -1), touched;
  auto bfs_last_level = [&](int32_t s, int32_t& ecc) -> int32_t {
    // BFS from s inside its component; returns a min-degree vertex of the last level.
    touched.clear();
    std::vector<int32_t> q{s};
    level[s] = 0;
    touched.push_back(s);
    size_t h = 0;
    while (h < q.size()) {
      int32_t v = q[h++];
      for (int64_t k = ptr[v]; k < ptr[v + 1]; ++k) {
        int32_t w = nb[k];
        if (level[w] < 0) { level[w] = level[v] + 1; q.push_back(w); touched.push_back(w); }
      }
    }
    ecc = level[q.back()];
    int32_t best = -1;
    for (int32_t v : q)
      if (level[v] == ecc && (best < 0 || deg[v] < deg[best] || (deg[v] == deg[best] && v < best))) best = v;
    for (int32_t v : touched) level[v] = -1;
    return best;
  };
  for (int32_t seed_v : by_deg) {
    if (visited[seed_v]) continue;
    int32_t start = seed_v, ecc = 0;
    int32_t cand = bfs_last_level(start, ecc);
    for (int it = 0; it < 8; ++it) {
      int32_t ecc2 = 0;
      int32_t nxt = bfs_last_level(cand, ecc2);
      if (ecc2 <= ecc) break;
      start = cand;
      ecc = ecc2;
      cand = nxt;
    }
    start = cand;
    size_t h = order.size();
    order.push_back(start);
    visited[start] = 1;
    std::vector<int32_t> tmp;
    while (h < order.size()) {
      int32_t v = order[h++];
      tmp.clear();
      for (int64_t k = ptr[v]; k < ptr[v + 1]; ++k)
        if (!visited[nb[k]]) { visited[nb[k]] = 1; tmp.push_back(nb[k]); }
      std::sort(tmp.begin(), tmp.end(), [&](int32_t a, int32_t b) {
        return deg[a] != deg[b] ? deg[a] < deg[b] : a < b;
      });
      order.insert(order.end(), tmp.begin(), tmp.end());
    }
  }
  std::vector<int32_t> perm(n);
  for (int64_t k = 0; k < n; ++k) perm[order[k]] = (int32_t)(n - 1 - k);
  return perm;
}

int generate_erdos_renyi(const GenParams& p, HostMatrix& M) {
  const int64_t n = p.n_keep;
  if (n <= 1 || n > INT32_MAX) return GSOFA_ERR_INVALID_ARG;
  // m undirected edges drawn uniformly (u != v); duplicates collapse.
  const double avg = p.avg_degree_x100 / 100.0;
  const int64_t m = (int64_t)std::llround(avg * (double)n / 2.0);
  std::vector<std::pair<int32_t, int32_t>> edges;
  edges.reserve(2 * m);
  for (int64_t e = 0; e < m; ++e) {
    uint64_t h = draw(p.seed, 0xE0E0ull, (uint64_t)e);
    int32_t u = (int32_t)(h % (uint64_t)n);
    int32_t v = (int32_t)((h >> 32) % (uint64_t)(n - 1));
    if (v >= u) ++v;
    edges.push_back({u, v});
    edges.push_back({v, u});
  }
  std::sort(edges.begin(), edges.end());
  edges.erase(std::unique(edges.begin(), edges.end()), edges.end());
  std::vector<int64_t> ptr(n + 1, 0);
  std::vector<int32_t> nb(edges.size());
  for (auto& e : edges) ptr[e.first + 1]++;
  for (int64_t v = 0; v < n; ++v) ptr[v + 1] += ptr[v];
  for (size_t k = 0; k < edges.size(); ++k) nb[k] = edges[k].second;  // edges sorted by (u, v)

  std::vector<int32_t> perm;
  if (p.order == GSOFA_ORDER_RCM) perm = rcm_order(n, ptr, nb);
  else { perm.resize(n); std::iota(perm.begin(), perm.end(), 0); }

  std::vector<std::vector<int32_t>> rows(n);
  for (int64_t u = 0; u < n; ++u) {
    const int32_t i = perm[u];
    rows[i].push_back(i);
    for (int64_t k = ptr[u]; k < ptr[u + 1]; ++k) {
      const int32_t j = perm[nb[k]];
      const bool lower = j < i;
      bool eligible = (p.drop_side == GSOFA_DROP_ANY) || (p.drop_side == GSOFA_DROP_LOWER && lower) ||
                      (p.drop_side == GSOFA_DROP_UPPER && !lower);
      if (eligible && p.drop_ppm > 0 && bernoulli_ppm(draw(p.seed, (uint64_t)i, (uint64_t)j), p.drop_ppm))
        continue;
      rows[i].push_back(j);
    }
  }
  build_csr_from_lists(n, rows, M);
  return GSOFA_OK;
}

int generate(const GenParams& p, HostMatrix& M) {
  switch (p.kind) {
    case GSOFA_GEN_STENCIL: return generate_stencil(p, M);
    case GSOFA_GEN_ERDOS_RENYI: return generate_erdos_renyi(p, M);
    default: return GSOFA_ERR_INVALID_ARG;
  }
}

// The workload presets. Seeds are 0x5EED0000 + workload number (1..5); batch and stream members
// derive their own seed from (workload seed, member index).
static double log_uniform(uint64_t h, double lo, double hi) {
  const double u = (double)(h >> 11) * (1.0 / 9007199254740992.0);
  return std::exp(std::log(lo) + u * (std::log(hi) - std::log(lo)));
}

int preset_params(int workload, int64_t member, GenParams& p) {
  p = GenParams{};
  p.seed = 0x5EED0000ull + (uint64_t)workload;
  switch (workload) {
    case GSOFA_WL_GRID2D_4096:  // 64x64 5-point, natural order, 20% of off-diagonals dropped
      p.kind = GSOFA_GEN_STENCIL; p.dims = 2; p.points = 5; p.nx = 64; p.ny = 64;
      p.order = GSOFA_ORDER_NATURAL; p.drop_side = GSOFA_DROP_ANY; p.drop_ppm = 200000;
      return GSOFA_OK;
    case GSOFA_WL_GRID3D_110592:  // 48^3 7-point, ND, 15% of upper off-diagonals dropped
      p.kind = GSOFA_GEN_STENCIL; p.dims = 3; p.points = 7; p.nx = p.ny = p.nz = 48;
      p.order = GSOFA_ORDER_ND; p.nd_leaf = 64; p.drop_side = GSOFA_DROP_UPPER; p.drop_ppm = 150000;
      return GSOFA_OK;
    case GSOFA_WL_GRID2D_659033:  // 812-wide 9-point truncated to 659,033, ND, 10% lower dropped, extras
      p.kind = GSOFA_GEN_STENCIL; p.dims = 2; p.points = 9; p.nx = 812; p.ny = 812; p.n_keep = 659033;
      p.order = GSOFA_ORDER_ND; p.nd_leaf = 64; p.drop_side = GSOFA_DROP_LOWER; p.drop_ppm = 100000;
      p.extra_every = 8;
      return GSOFA_OK;
    case GSOFA_WL_BATCH256:
    case GSOFA_WL_STREAM: {
      const bool batch = workload == GSOFA_WL_BATCH256;
      const uint64_t ms = splitmix64(p.seed ^ (0x9E37ull * (uint64_t)(member + 1)));
      const double lo = batch ? 2000.0 : 1000.0, hi = batch ? 200000.0 : 500000.0;
      const int64_t n = (int64_t)std::llround(log_uniform(splitmix64(ms ^ 1), lo, hi));
      p.seed = ms;
      p.drop_ppm = 100000;
      if ((member & 1) == 0) {
        // ND stencil; alternates 2-D 9-point and 3-D 7-point between stencil members.
        p.kind = GSOFA_GEN_STENCIL; p.order = GSOFA_ORDER_ND; p.nd_leaf = 64; p.drop_side = GSOFA_DROP_LOWER;
        if (((member >> 1) & 1) == 0) {
          p.dims = 2; p.points = 9;
          p.nx = (int64_t)std::ceil(std::sqrt((double)n)); p.ny = p.nx;
        } else {
          p.dims = 3; p.points = 7;
          int64_t w = (int64_t)std::ceil(std::cbrt((double)n));
          while (w * w * w < n) ++w;
          p.nx = p.ny = p.nz = w;
        }
        p.n_keep = n;
      } else {
        p.kind = GSOFA_GEN_ERDOS_RENYI; p.order = GSOFA_ORDER_RCM; p.n_keep = n; p.avg_degree_x100 = 600;
        p.drop_side = GSOFA_DROP_ANY;
      }
      return GSOFA_OK;
    }
    default:
      return GSOFA_ERR_INVALID_ARG;
  }
}

}  // namespace gsofa
