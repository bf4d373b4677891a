// Host symbolic analysis: ordering, elimination tree, column counts and the
// supernodal multifrontal plan.  Runs once per setup() on the fixed KKT
// pattern (the role of qdldl_ldlt.cpp:22-56 in the reference).
#include "symbolic.hpp"

#include <algorithm>
#include <cstdint>
#include <numeric>
#include <stdexcept>

namespace cx {

namespace {

// Full symmetric adjacency (no diagonal) of the lower pattern, in labels given by `lab`
// (identity when empty).
void build_adjacency(const LowerPattern& K, const std::vector<int>& iperm, std::vector<int>& ptr,
                     std::vector<int>& idx) {
  const int N = K.N;
  ptr.assign(N + 1, 0);
  for (int c = 0; c < N; ++c)
    for (int k = K.colptr[c]; k < K.colptr[c + 1]; ++k) {
      int r = K.rowind[k];
      if (r == c) continue;
      int a = iperm.empty() ? r : iperm[r];
      int b = iperm.empty() ? c : iperm[c];
      ptr[a + 1]++;
      ptr[b + 1]++;
    }
  for (int i = 0; i < N; ++i) ptr[i + 1] += ptr[i];
  idx.assign(ptr[N], 0);
  std::vector<int> fill(ptr.begin(), ptr.end() - 1);
  for (int c = 0; c < N; ++c)
    for (int k = K.colptr[c]; k < K.colptr[c + 1]; ++k) {
      int r = K.rowind[k];
      if (r == c) continue;
      int a = iperm.empty() ? r : iperm[r];
      int b = iperm.empty() ? c : iperm[c];
      idx[fill[a]++] = b;
      idx[fill[b]++] = a;
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// Approximate minimum degree (quotient graph, element absorption, mass
// elimination, supervariable detection, approximate external degrees).
// ---------------------------------------------------------------------------
std::vector<int> amd_order(const LowerPattern& K) {
  const int N = K.N;
  std::vector<int> order;
  if (N == 0) return order;
  std::vector<int> aptr, aidx;
  build_adjacency(K, {}, aptr, aidx);

  enum : char { VAR = 0, ELEM = 1, DEAD = 2 };
  std::vector<std::vector<int>> A(N), E(N), Le(N);
  for (int i = 0; i < N; ++i) {
    A[i].assign(aidx.begin() + aptr[i], aidx.begin() + aptr[i + 1]);
    std::sort(A[i].begin(), A[i].end());
    A[i].erase(std::unique(A[i].begin(), A[i].end()), A[i].end());
  }
  std::vector<char> st(N, VAR);
  std::vector<int> nv(N, 1), degree(N), rep(N, -1), elw(N, 0), tmpdeg(N, 0);
  std::vector<int64_t> w(N, 0), hashv(N, 0);
  std::vector<int> mark(N, 0), mark2(N, 0);
  int stamp = 0, stamp2 = 0;
  int64_t wflg = 1;
  // degree buckets
  std::vector<int> head(N + 1, -1), nxt(N, -1), prv(N, -1);
  int mindeg = 0;
  auto bucket_insert = [&](int i, int d) {
    prv[i] = -1;
    nxt[i] = head[d];
    if (head[d] != -1) prv[head[d]] = i;
    head[d] = i;
    if (d < mindeg) mindeg = d;
  };
  auto bucket_remove = [&](int i) {
    int d = degree[i];
    if (prv[i] != -1)
      nxt[prv[i]] = nxt[i];
    else
      head[d] = nxt[i];
    if (nxt[i] != -1) prv[nxt[i]] = prv[i];
    prv[i] = nxt[i] = -1;
  };
  for (int i = 0; i < N; ++i) {
    degree[i] = (int)A[i].size();
    bucket_insert(i, degree[i]);
  }
  mindeg = 0;
  int nel = 0;
  std::vector<int> Lp, keep;
  order.reserve(N);
  while (nel < N) {
    while (mindeg <= N && head[mindeg] == -1) ++mindeg;
    if (mindeg > N) throw std::runtime_error("amd: empty degree lists");
    const int p = head[mindeg];
    bucket_remove(p);
    // --- new element Lp = union of absorbed elements and p's variables ---
    ++stamp;
    mark[p] = stamp;
    Lp.clear();
    for (int e : E[p]) {
      if (st[e] != ELEM) continue;
      for (int i : Le[e])
        if (st[i] == VAR && mark[i] != stamp) mark[i] = stamp, Lp.push_back(i);
      st[e] = DEAD;
      rep[e] = p;
      std::vector<int>().swap(Le[e]);
    }
    for (int i : A[p])
      if (st[i] == VAR && mark[i] != stamp) mark[i] = stamp, Lp.push_back(i);
    std::vector<int>().swap(E[p]);
    std::vector<int>().swap(A[p]);
    st[p] = ELEM;
    nel += nv[p];
    int degme = 0;
    for (int i : Lp) degme += nv[i], bucket_remove(i);
    // --- |Le \ Lp| for every element touching Lp ---
    for (int i : Lp)
      for (int e : E[i]) {
        if (st[e] != ELEM || e == p) continue;
        if (w[e] < wflg) w[e] = wflg + elw[e];
        w[e] -= nv[i];
      }
    // --- degree update, element pruning/absorption, mass elimination ---
    keep.clear();
    for (int i : Lp) {
      int dext = 0;
      int64_t h = 0;
      std::vector<int>& Ei = E[i];
      size_t k = 0;
      for (int e : Ei) {
        if (st[e] != ELEM || e == p) continue;
        int64_t ext = w[e] - wflg;
        if (ext > 0) {
          dext += (int)ext;
          Ei[k++] = e;
          h += e;
        } else {  // Le subset of Lp: aggressive absorption into p
          st[e] = DEAD;
          rep[e] = p;
          std::vector<int>().swap(Le[e]);
        }
      }
      Ei.resize(k);
      Ei.push_back(p);
      std::vector<int>& Ai = A[i];
      k = 0;
      for (int j : Ai) {
        if (st[j] != VAR || mark[j] == stamp) continue;  // in Lp (or p): reached through p
        Ai[k++] = j;
        dext += nv[j];
        h += j;
      }
      Ai.resize(k);
      if (Ei.size() == 1 && Ai.empty()) {  // only adjacent to p: eliminate together with p
        st[i] = DEAD;
        rep[i] = p;
        nv[p] += nv[i];
        degme -= nv[i];
        nel += nv[i];
        std::vector<int>().swap(E[i]);
        continue;
      }
      tmpdeg[i] = dext;
      hashv[i] = h;
      keep.push_back(i);
    }
    // --- supervariable detection among the surviving Lp members ---
    std::sort(keep.begin(), keep.end(), [&](int a, int b) {
      return hashv[a] != hashv[b] ? hashv[a] < hashv[b] : a < b;
    });
    for (size_t a0 = 0; a0 < keep.size();) {
      size_t a1 = a0 + 1;
      while (a1 < keep.size() && hashv[keep[a1]] == hashv[keep[a0]]) ++a1;
      for (size_t x = a0; x < a1; ++x) {
        int a = keep[x];
        if (st[a] != VAR) continue;
        bool marked = false;
        for (size_t y = x + 1; y < a1; ++y) {
          int b = keep[y];
          if (st[b] != VAR || E[b].size() != E[a].size() || A[b].size() != A[a].size()) continue;
          if (!marked) {
            ++stamp2;
            for (int e : E[a]) mark2[e] = stamp2;
            for (int j : A[a]) mark2[j] = stamp2;
            marked = true;
          }
          bool same = true;
          for (int e : E[b])
            if (mark2[e] != stamp2) { same = false; break; }
          if (same)
            for (int j : A[b])
              if (mark2[j] != stamp2) { same = false; break; }
          if (!same) continue;
          nv[a] += nv[b];
          st[b] = DEAD;
          rep[b] = a;
          std::vector<int>().swap(E[b]);
          std::vector<int>().swap(A[b]);
        }
      }
      a0 = a1;
    }
    // --- finalize the element and re-insert its variables ---
    elw[p] = degme;
    std::vector<int>& lp = Le[p];
    lp.clear();
    for (int i : keep) {
      if (st[i] != VAR) continue;
      lp.push_back(i);
      int d1 = degree[i] + degme - nv[i];
      int d2 = tmpdeg[i] + degme - nv[i];
      int d = std::min(d1, d2);
      d = std::min(d, N - nel - nv[i]);
      if (d < 0) d = 0;
      degree[i] = d;
      bucket_insert(i, d);
    }
    lp.shrink_to_fit();
    wflg += (int64_t)N + 1;
    order.push_back(p);
  }
  // --- expand supervariables / mass-eliminated nodes after their pivot ---
  std::vector<int> pos(N, -1);
  for (size_t k = 0; k < order.size(); ++k) pos[order[k]] = (int)k;
  std::vector<int> root(N, -1);
  std::vector<int> cnt(order.size() + 1, 0);
  for (int v = 0; v < N; ++v) {
    int r = v;
    while (pos[r] < 0) r = rep[r];
    root[v] = pos[r];
    cnt[pos[r] + 1]++;
  }
  for (size_t k = 0; k < order.size(); ++k) cnt[k + 1] += cnt[k];
  std::vector<int> perm(N);
  std::vector<int> fill(cnt.begin(), cnt.end() - 1);
  for (size_t k = 0; k < order.size(); ++k) perm[fill[k]++] = order[k];
  for (int v = 0; v < N; ++v)
    if (pos[v] < 0) perm[fill[root[v]]++] = v;
  return perm;
}

// ---------------------------------------------------------------------------
std::vector<int> etree_of(const LowerPattern& K, const std::vector<int>& perm) {
  const int N = K.N;
  std::vector<int> iperm(N);
  for (int k = 0; k < N; ++k) iperm[perm[k]] = k;
  // upper lists: for each new column k, neighbours i < k
  std::vector<int> ptr(N + 1, 0), idx;
  for (int c = 0; c < N; ++c)
    for (int k = K.colptr[c]; k < K.colptr[c + 1]; ++k) {
      int r = K.rowind[k];
      if (r == c) continue;
      ptr[std::max(iperm[r], iperm[c]) + 1]++;
    }
  for (int i = 0; i < N; ++i) ptr[i + 1] += ptr[i];
  idx.resize(ptr[N]);
  std::vector<int> fill(ptr.begin(), ptr.end() - 1);
  for (int c = 0; c < N; ++c)
    for (int k = K.colptr[c]; k < K.colptr[c + 1]; ++k) {
      int r = K.rowind[k];
      if (r == c) continue;
      int a = iperm[r], b = iperm[c];
      idx[fill[std::max(a, b)]++] = std::min(a, b);
    }
  std::vector<int> parent(N, -1), anc(N, -1);
  for (int k = 0; k < N; ++k)
    for (int t = ptr[k]; t < ptr[k + 1]; ++t) {
      int i = idx[t];
      while (i != -1 && i < k) {
        int nx = anc[i];
        anc[i] = k;
        if (nx == -1) parent[i] = k;
        i = nx;
      }
    }
  return parent;
}

std::vector<int> postorder(const std::vector<int>& parent, const std::vector<int>& key) {
  const int N = (int)parent.size();
  std::vector<int> head(N, -1), next(N, -1), post;
  post.reserve(N);
  if (key.empty()) {
    for (int j = N - 1; j >= 0; --j)
      if (parent[j] != -1) next[j] = head[parent[j]], head[parent[j]] = j;
  } else {
    // children of every node linked in ascending (key, index) order: the child
    // with the largest key is visited last and ends up adjacent to its parent
    std::vector<int> order(N);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return key[a] > key[b]; });
    for (int j : order)
      if (parent[j] != -1) next[j] = head[parent[j]], head[parent[j]] = j;
  }
  std::vector<int> stack;
  for (int r = 0; r < N; ++r) {
    if (parent[r] != -1) continue;
    stack.push_back(r);
    while (!stack.empty()) {
      int p = stack.back();
      int c = head[p];
      if (c == -1) {
        stack.pop_back();
        post.push_back(p);
      } else {
        head[p] = next[c];
        stack.push_back(c);
      }
    }
  }
  return post;
}

std::vector<int> column_counts(const LowerPattern& K, const std::vector<int>& perm,
                               const std::vector<int>& parent, const std::vector<int>& post) {
  const int N = K.N;
  std::vector<int> iperm(N);
  for (int k = 0; k < N; ++k) iperm[perm[k]] = k;
  std::vector<int> aptr, aidx;
  build_adjacency(K, iperm, aptr, aidx);
  std::vector<int> first(N, -1), maxfirst(N, -1), prevleaf(N, -1), anc(N), delta(N, 0);
  for (int k = 0; k < N; ++k) {
    int j = post[k];
    delta[j] = (first[j] == -1) ? 1 : 0;
    for (; j != -1 && first[j] == -1; j = parent[j]) first[j] = k;
  }
  std::iota(anc.begin(), anc.end(), 0);
  for (int k = 0; k < N; ++k) {
    int j = post[k];
    if (parent[j] != -1) delta[parent[j]]--;
    for (int t = aptr[j]; t < aptr[j + 1]; ++t) {
      int i = aidx[t];
      if (i <= j || first[j] <= maxfirst[i]) continue;  // j is not a leaf of row subtree i
      maxfirst[i] = first[j];
      int jprev = prevleaf[i];
      prevleaf[i] = j;
      delta[j]++;
      if (jprev != -1) {  // subsequent leaf: subtract overlap at the least common ancestor
        int q = jprev;
        while (q != anc[q]) q = anc[q];
        for (int s = jprev; s != q;) {
          int sp = anc[s];
          anc[s] = q;
          s = sp;
        }
        delta[q]--;
      }
    }
    if (parent[j] != -1) anc[j] = parent[j];
  }
  std::vector<int> cc(delta);
  for (int k = 0; k < N; ++k) {
    int j = post[k];
    if (parent[j] != -1) cc[parent[j]] += cc[j];
  }
  return cc;
}

// ---------------------------------------------------------------------------
namespace {

// Supernodal plan for a fixed elimination order perm (new -> old) that lists
// every column after its elimination-tree descendants; parent and cc are the
// etree and the column counts of L in the new labels.  force (new labels,
// optional): a node whose first column is marked is merged into its parent
// whatever its zero count.
SupernodalPlan build_plan(const LowerPattern& K, const std::vector<int>& perm,
                          const std::vector<int>& parent, const std::vector<int>& cc,
                          const RelaxParams& rp, const std::vector<char>* force = nullptr) {
  SupernodalPlan P;
  const int N = K.N;
  P.N = N;
  P.nnzL_exact = 0;
  for (int j = 0; j < N; ++j) P.nnzL_exact += cc[j];

  // fundamental supernodes
  std::vector<int> nchild(N, 0);
  for (int j = 0; j < N; ++j)
    if (parent[j] != -1) nchild[parent[j]]++;
  std::vector<int> ffirst;
  for (int j = 0; j < N; ++j) {
    bool join = j > 0 && parent[j - 1] == j && cc[j - 1] == cc[j] + 1 && nchild[j] == 1;
    if (!join) ffirst.push_back(j);
  }
  const int nsf = (int)ffirst.size();
  ffirst.push_back(N);

  // relaxed amalgamation: merge the immediately preceding child into its parent
  struct Node { int first, last, ncol, nrow; double zeros; };
  std::vector<Node> stack;
  auto storage = [](double k, double r) { return k * r - k * (k - 1) / 2; };
  for (int f = 0; f < nsf; ++f) {
    Node s{ffirst[f], ffirst[f + 1] - 1, ffirst[f + 1] - ffirst[f], cc[ffirst[f]], 0.0};
    while (!stack.empty()) {
      const Node& c = stack.back();
      int pc = parent[c.last];
      if (pc < s.first || pc > s.last) break;  // preceding node is not a child of s
      int k = c.ncol + s.ncol;
      int r = c.ncol + s.nrow;
      if (k > rp.max_merge_cols) break;
      double sm = storage(k, r);
      double actual = (storage(c.ncol, c.nrow) - c.zeros) + (storage(s.ncol, s.nrow) - s.zeros);
      double z = sm - actual;
      double frac = sm > 0 ? z / sm : 0.0;
      bool ok = k <= rp.nrelax0 || (k <= rp.nrelax1 && frac < rp.zrelax0) ||
                (k <= rp.nrelax2 && frac < rp.zrelax1) || frac < rp.zrelax2 ||
                (force && (*force)[c.first]);
      if (!ok) break;
      s = Node{c.first, s.last, k, r, z};
      stack.pop_back();
    }
    stack.push_back(s);
  }
  P.ns = (int)stack.size();
  P.sn_first.resize(P.ns);
  P.sn_ncol.resize(P.ns);
  P.col_sn.assign(N, -1);
  for (int s = 0; s < P.ns; ++s) {
    P.sn_first[s] = stack[s].first;
    P.sn_ncol[s] = stack[s].ncol;
    for (int j = stack[s].first; j <= stack[s].last; ++j) P.col_sn[j] = s;
  }
  // supernode tree
  P.sn_parent.assign(P.ns, -1);
  for (int s = 0; s < P.ns; ++s) {
    int last = P.sn_first[s] + P.sn_ncol[s] - 1;
    if (parent[last] != -1) P.sn_parent[s] = P.col_sn[parent[last]];
  }
  P.child_ptr.assign(P.ns + 1, 0);
  for (int s = 0; s < P.ns; ++s)
    if (P.sn_parent[s] != -1) P.child_ptr[P.sn_parent[s] + 1]++;
  for (int s = 0; s < P.ns; ++s) P.child_ptr[s + 1] += P.child_ptr[s];
  P.child.resize(P.child_ptr[P.ns]);
  {
    std::vector<int> fill(P.child_ptr.begin(), P.child_ptr.end() - 1);
    for (int s = 0; s < P.ns; ++s)
      if (P.sn_parent[s] != -1) P.child[fill[P.sn_parent[s]]++] = s;
  }
  // front row structures (children before parents: postorder)
  std::vector<int> iperm(N);
  for (int k = 0; k < N; ++k) iperm[perm[k]] = k;
  std::vector<int> aptr, aidx;
  build_adjacency(K, iperm, aptr, aidx);
  std::vector<int> mark(N, -1);
  P.row_ptr.assign(P.ns + 1, 0);
  P.rows.clear();
  std::vector<int> extra;
  for (int s = 0; s < P.ns; ++s) {
    int f = P.sn_first[s], nc = P.sn_ncol[s], l = f + nc;
    extra.clear();
    for (int j = f; j < l; ++j)
      for (int t = aptr[j]; t < aptr[j + 1]; ++t) {
        int i = aidx[t];
        if (i >= l && mark[i] != s) mark[i] = s, extra.push_back(i);
      }
    for (int t = P.child_ptr[s]; t < P.child_ptr[s + 1]; ++t) {
      int c = P.child[t];
      for (int q = P.row_ptr[c] + P.sn_ncol[c]; q < P.row_ptr[c + 1]; ++q) {
        int i = P.rows[q];
        if (i >= l && mark[i] != s) mark[i] = s, extra.push_back(i);
      }
    }
    std::sort(extra.begin(), extra.end());
    for (int j = f; j < l; ++j) P.rows.push_back(j);
    P.rows.insert(P.rows.end(), extra.begin(), extra.end());
    P.row_ptr[s + 1] = (int)P.rows.size();
  }
  // levels
  P.level.assign(P.ns, 0);
  P.nlevels = 0;
  for (int s = 0; s < P.ns; ++s) {
    int lv = 0;
    for (int t = P.child_ptr[s]; t < P.child_ptr[s + 1]; ++t) lv = std::max(lv, P.level[P.child[t]] + 1);
    P.level[s] = lv;
    P.nlevels = std::max(P.nlevels, lv + 1);
  }
  P.level_ptr.assign(P.nlevels + 1, 0);
  for (int s = 0; s < P.ns; ++s) P.level_ptr[P.level[s] + 1]++;
  for (int l = 0; l < P.nlevels; ++l) P.level_ptr[l + 1] += P.level_ptr[l];
  P.level_sn.resize(P.ns);
  {
    std::vector<int> fill(P.level_ptr.begin(), P.level_ptr.end() - 1);
    for (int s = 0; s < P.ns; ++s) P.level_sn[fill[P.level[s]]++] = s;
  }
  // statistics
  P.nnzL = 0;
  P.flops = 0;
  P.max_front = 0;
  for (int s = 0; s < P.ns; ++s) {
    double nc = P.sn_ncol[s], nr = P.nr(s);
    P.nnzL += nc * nr - nc * (nc - 1) / 2;
    for (int t = 0; t < P.sn_ncol[s]; ++t) {
      double c = nr - t;
      P.flops += c * (c - 1);
    }
    P.max_front = std::max(P.max_front, P.nr(s));
  }
  P.perm = perm;
  P.iperm = iperm;
  return P;
}

// Thin fronts (a few pivot columns over a large update matrix, such as a
// short separator beside a dense one) cost a memory-bound trailing pass over
// their whole update matrix and an extend-add of it into the parent, and
// relaxed amalgamation can absorb a front only when it immediately precedes
// its parent, which a dense separator's own chain usually does instead.  This
// order puts every such front (and, transitively, thin fronts absorbed into
// it) right before its parent's columns, after the subtrees of all the
// non-absorbed children.  Each column still follows its descendants, so the
// fill is that of the first order; amalgamation then merges the thin fronts
// into the parent (explicit zeros in their columns instead of an update
// matrix of their own; force marks their columns in the new order).  Returns
// false when no front qualifies.
bool thin_front_order(const SupernodalPlan& P, const RelaxParams& rp, std::vector<int>& perm2,
                      std::vector<char>& force) {
  const int ns = P.ns;
  std::vector<char> absorb(ns, 0);
  int nabs = 0;
  for (int s = 0; s < ns; ++s) {
    const int p = P.sn_parent[s];
    if (p < 0) continue;
    const int nc = P.sn_ncol[s], nu = P.nr(s) - nc;
    if ((nc <= rp.thin_ncol || 64L * nc <= nu) && nu >= rp.thin_min_update && 2L * nu >= P.nr(p))
      absorb[s] = 1, ++nabs;
  }
  if (nabs == 0) return false;
  perm2.clear();
  perm2.reserve(P.N);
  force.clear();
  force.reserve(P.N);
  // A frame emits the subtree of s: the non-absorbed children's subtrees of
  // every member of s's cluster (s and the fronts absorbed into it,
  // descendants first), then the members' columns.
  struct Frame {
    std::vector<int> mem;
    size_t mi = 0;
    int ci = 0;
  };
  std::vector<Frame> st;
  std::vector<std::pair<int, int>> dfs;  // (node, next child slot)
  auto push = [&](int s) {
    Frame f;
    dfs.assign(1, {s, P.child_ptr[s]});
    while (!dfs.empty()) {
      const int v = dfs.back().first;
      int c = dfs.back().second;
      while (c < P.child_ptr[v + 1] && !absorb[P.child[c]]) ++c;
      if (c < P.child_ptr[v + 1]) {
        dfs.back().second = c + 1;
        const int a = P.child[c];
        dfs.push_back({a, P.child_ptr[a]});
      } else {
        f.mem.push_back(v);
        dfs.pop_back();
      }
    }
    f.ci = P.child_ptr[f.mem[0]];
    st.push_back(std::move(f));
  };
  for (int r = 0; r < ns; ++r) {
    if (P.sn_parent[r] >= 0) continue;
    push(r);
    while (!st.empty()) {
      Frame& f = st.back();
      if (f.mi < f.mem.size()) {
        const int m = f.mem[f.mi];
        if (f.ci < P.child_ptr[m + 1]) {
          const int q = P.child[f.ci++];
          if (!absorb[q]) push(q);
          continue;
        }
        if (++f.mi < f.mem.size()) f.ci = P.child_ptr[f.mem[f.mi]];
        continue;
      }
      for (int m : f.mem)
        for (int j = P.sn_first[m]; j < P.sn_first[m] + P.sn_ncol[m]; ++j) {
          perm2.push_back(P.perm[j]);
          force.push_back(absorb[m]);
        }
      st.pop_back();
    }
  }
  return (int)perm2.size() == P.N;
}

}  // namespace

SupernodalPlan plan_supernodes(const LowerPattern& K, const RelaxParams& rp) {
  const int N = K.N;
  if (N == 0) {
    SupernodalPlan P;
    P.N = 0;
    return P;
  }
  std::vector<int> perm0 = amd_order(K);
  std::vector<int> par0 = etree_of(K, perm0);
  // Column counts do not depend on which postorder is used, so count once in a
  // plain postorder, then visit children by ascending count: the child whose
  // front is closest to its parent's sits next to it, where relaxed
  // amalgamation can absorb it (dense separators then form one wide front).
  std::vector<int> cc0 = column_counts(K, perm0, par0, postorder(par0));
  std::vector<int> post = postorder(par0, cc0);
  // relabel so that the elimination tree is postordered
  std::vector<int> perm(N);
  for (int k = 0; k < N; ++k) perm[k] = perm0[post[k]];
  std::vector<int> ipost(N);
  for (int k = 0; k < N; ++k) ipost[post[k]] = k;
  std::vector<int> parent(N);
  for (int k = 0; k < N; ++k) parent[k] = par0[post[k]] == -1 ? -1 : ipost[par0[post[k]]];
  std::vector<int> cc(N);
  for (int k = 0; k < N; ++k) cc[k] = cc0[post[k]];
  SupernodalPlan P = build_plan(K, perm, parent, cc, rp);
  std::vector<int> perm2;
  std::vector<char> force;
  if (rp.thin_ncol > 0 && thin_front_order(P, rp, perm2, force)) {
    std::vector<int> par2 = etree_of(K, perm2);
    std::vector<int> cc2 = column_counts(K, perm2, par2, postorder(par2));
    SupernodalPlan P2 = build_plan(K, perm2, par2, cc2, rp, &force);
    if (P2.nnzL_exact == P.nnzL_exact) return P2;  // an equivalent order: same fill
  }
  return P;
}

}  // namespace cx
