// Host-side setup of libpdm: scene construction in engine order, nested
// dissection ordering, supernodal symbolic analysis and the numeric
// multifrontal Cholesky of K_s = M/h^2 + sum_i w_i G_i G_i^T (scalar N x N;
// the reference's 3N x 3N H2ML equals K_s (x) I3, World.h:1089-1139).
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <map>
#include <numeric>

#include "pdm_internal.h"

namespace pdm {

namespace {

inline double det3(const double* a) {
  return a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
         a[2] * (a[3] * a[7] - a[4] * a[6]);
}

// Dm columns p1-p0, p2-p0, p3-p0 (row-major 3x3).
void dm_of(const double* X, int a, int b, int c, int d, double* Dm) {
  for (int r = 0; r < 3; ++r) {
    Dm[r * 3 + 0] = X[3 * b + r] - X[3 * a + r];
    Dm[r * 3 + 1] = X[3 * c + r] - X[3 * a + r];
    Dm[r * 3 + 2] = X[3 * d + r] - X[3 * a + r];
  }
}

// Orientation fix (Dfobj.h:474-487), volume and the 4x3 stencil
// G = [-colsum(invDm); invDm] so that F[:,j] = sum_k x_k G[k][j].
void tet_geometry(const double* X, int* v, double* vol, double* G) {
  double Dm[9];
  dm_of(X, v[0], v[1], v[2], v[3], Dm);
  if (det3(Dm) < 0) {
    std::swap(v[1], v[2]);
    dm_of(X, v[0], v[1], v[2], v[3], Dm);
  }
  double d = det3(Dm);
  *vol = d / 6.0;
  double inv[9];
  inv[0] = (Dm[4] * Dm[8] - Dm[5] * Dm[7]) / d;
  inv[1] = (Dm[2] * Dm[7] - Dm[1] * Dm[8]) / d;
  inv[2] = (Dm[1] * Dm[5] - Dm[2] * Dm[4]) / d;
  inv[3] = (Dm[5] * Dm[6] - Dm[3] * Dm[8]) / d;
  inv[4] = (Dm[0] * Dm[8] - Dm[2] * Dm[6]) / d;
  inv[5] = (Dm[2] * Dm[3] - Dm[0] * Dm[5]) / d;
  inv[6] = (Dm[3] * Dm[7] - Dm[4] * Dm[6]) / d;
  inv[7] = (Dm[1] * Dm[6] - Dm[0] * Dm[7]) / d;
  inv[8] = (Dm[0] * Dm[4] - Dm[1] * Dm[3]) / d;
  for (int j = 0; j < 3; ++j) {
    G[0 * 3 + j] = -(inv[0 * 3 + j] + inv[1 * 3 + j] + inv[2 * 3 + j]);
    for (int m = 0; m < 3; ++m) G[(m + 1) * 3 + j] = inv[m * 3 + j];
  }
}

// Recursive coordinate-bisection nested dissection: split the longest axis at
// the median, take the boundary layer of the upper half as the vertex
// separator, order halves first and the separator last.
void nd_rec(std::vector<int>& verts, const double* xyz, const std::vector<std::vector<int>>& adj,
            std::vector<int>& mark, int& stamp, std::vector<int>& order) {
  const size_t kLeaf = 48;
  if (verts.size() <= kLeaf) {
    order.insert(order.end(), verts.begin(), verts.end());
    return;
  }
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int v : verts)
    for (int r = 0; r < 3; ++r) {
      lo[r] = std::min(lo[r], xyz[3 * v + r]);
      hi[r] = std::max(hi[r], xyz[3 * v + r]);
    }
  int ax = 0;
  for (int r = 1; r < 3; ++r)
    if (hi[r] - lo[r] > hi[ax] - lo[ax]) ax = r;
  std::vector<std::pair<double, int>> key(verts.size());
  for (size_t i = 0; i < verts.size(); ++i) key[i] = {xyz[3 * verts[i] + ax], verts[i]};
  std::sort(key.begin(), key.end());
  size_t mid = key.size() / 2;
  double cut = key[mid].first;
  // A: coordinate < cut ; B: coordinate >= cut
  ++stamp;
  int stampA = stamp;
  std::vector<int> A, B;
  for (auto& kv : key) {
    if (kv.first < cut) {
      A.push_back(kv.second);
      mark[kv.second] = stampA;
    } else {
      B.push_back(kv.second);
    }
  }
  if (A.empty() || B.empty()) {  // degenerate split: fall back to index halves
    A.assign(verts.begin(), verts.begin() + verts.size() / 2);
    B.assign(verts.begin() + verts.size() / 2, verts.end());
    ++stamp;
    stampA = stamp;
    for (int v : A) mark[v] = stampA;
  }
  std::vector<int> S, B2;
  for (int v : B) {
    bool touches = false;
    for (int u : adj[v])
      if (mark[u] == stampA) {
        touches = true;
        break;
      }
    (touches ? S : B2).push_back(v);
  }
  std::sort(S.begin(), S.end(), [&](int a, int b) {
    for (int r = 0; r < 3; ++r)
      if (xyz[3 * a + r] != xyz[3 * b + r]) return xyz[3 * a + r] < xyz[3 * b + r];
    return a < b;
  });
  nd_rec(A, xyz, adj, mark, stamp, order);
  nd_rec(B2, xyz, adj, mark, stamp, order);
  order.insert(order.end(), S.begin(), S.end());
}

}  // namespace

void nd_order(int N, const double* xyz, const std::vector<std::vector<int>>& adj, std::vector<int>& order) {
  order.clear();
  order.reserve(N);
  std::vector<int> mark(N, 0);
  int stamp = 0;
  // connected components are dissected independently
  std::vector<int> comp(N, -1);
  int nc = 0;
  for (int s = 0; s < N; ++s) {
    if (comp[s] >= 0) continue;
    std::vector<int> q{s};
    comp[s] = nc;
    for (size_t h = 0; h < q.size(); ++h)
      for (int u : adj[q[h]])
        if (comp[u] < 0) {
          comp[u] = nc;
          q.push_back(u);
        }
    std::sort(q.begin(), q.end());
    nd_rec(q, xyz, adj, mark, stamp, order);
    ++nc;
  }
}

// Faces (s0,s1,s2), (s0,s1,s3), (s0,s2,s3), (s1,s2,s3) of every tet's sorted
// vertex ids, in tet order; a face that no other tet shares is a contour, and
// it is turned outward: with n = (q1 - q0) x (q2 - q1), q0 and q1 swap when n
// points towards the owning tet's centre. The reference finds unshared faces
// by an all-pairs scan; a sort of the 4T keys finds the same set.
std::vector<int> contour_faces(const double* xrest, const int* tets, int T) {
  const int fc[4][3] = {{0, 1, 2}, {0, 1, 3}, {0, 2, 3}, {1, 2, 3}};
  std::vector<std::array<int, 3>> tri(4 * (size_t)T);
  for (int t = 0; t < T; ++t) {
    int sv[4] = {tets[4 * t], tets[4 * t + 1], tets[4 * t + 2], tets[4 * t + 3]};
    std::sort(sv, sv + 4);
    for (int k = 0; k < 4; ++k) tri[4 * (size_t)t + k] = {sv[fc[k][0]], sv[fc[k][1]], sv[fc[k][2]]};
  }
  std::vector<int> ord(tri.size());
  std::iota(ord.begin(), ord.end(), 0);
  std::sort(ord.begin(), ord.end(), [&](int a, int b) { return tri[a] < tri[b] || (tri[a] == tri[b] && a < b); });
  std::vector<char> uniq(tri.size(), 0);
  for (size_t k = 0; k < ord.size(); ++k) {
    bool same_prev = k > 0 && tri[ord[k - 1]] == tri[ord[k]];
    bool same_next = k + 1 < ord.size() && tri[ord[k + 1]] == tri[ord[k]];
    uniq[ord[k]] = !(same_prev || same_next);
  }
  std::vector<int> out;
  for (size_t i = 0; i < tri.size(); ++i) {
    if (!uniq[i]) continue;
    std::array<int, 3> f = tri[i];
    const int* tv = tets + 4 * (i / 4);
    double c[3], q0[3], e1[3], e2[3];
    for (int r = 0; r < 3; ++r) {
      c[r] = 0.25 * (xrest[3 * tv[0] + r] + xrest[3 * tv[1] + r] + xrest[3 * tv[2] + r] + xrest[3 * tv[3] + r]);
      q0[r] = xrest[3 * f[0] + r];
      e1[r] = xrest[3 * f[1] + r] - q0[r];
      e2[r] = xrest[3 * f[2] + r] - xrest[3 * f[1] + r];
    }
    double n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
    if ((c[0] - q0[0]) * n[0] + (c[1] - q0[1]) * n[1] + (c[2] - q0[2]) * n[2] > 0.0) std::swap(f[0], f[1]);
    out.insert(out.end(), f.begin(), f.end());
  }
  return out;
}

// Leaves hold at most 4 triangles; a node splits its triangles at the median
// centroid along its box's longest axis (ties by triangle id, so the tree
// depends only on the input).
void build_tri_bvh(HostScene& S) {
  int F = (int)(S.tris.size() / 9);
  S.bvh_box.clear();
  S.bvh_a.clear();
  S.bvh_b.clear();
  S.bvh_tri.resize(F);
  std::iota(S.bvh_tri.begin(), S.bvh_tri.end(), 0);
  if (F == 0) return;
  auto cen = [&](int t, int ax) {
    const double* q = &S.tris[9 * (size_t)t];
    return q[ax] + q[3 + ax] + q[6 + ax];
  };
  struct Item { int node, lo, hi; };
  std::vector<Item> st;
  auto new_node = [&]() {
    S.bvh_box.insert(S.bvh_box.end(), 6, 0.0);
    S.bvh_a.push_back(0);
    S.bvh_b.push_back(0);
    return (int)S.bvh_a.size() - 1;
  };
  st.push_back({new_node(), 0, F});
  while (!st.empty()) {
    Item it = st.back();
    st.pop_back();
    double* bx = &S.bvh_box[6 * (size_t)it.node];
    for (int r = 0; r < 3; ++r) {
      bx[r] = 1e300;
      bx[3 + r] = -1e300;
    }
    for (int k = it.lo; k < it.hi; ++k) {
      const double* q = &S.tris[9 * (size_t)S.bvh_tri[k]];
      for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) {
          bx[r] = std::min(bx[r], q[3 * c + r]);
          bx[3 + r] = std::max(bx[3 + r], q[3 * c + r]);
        }
    }
    if (it.hi - it.lo <= 4) {
      S.bvh_a[it.node] = -1 - it.lo;
      S.bvh_b[it.node] = it.hi - it.lo;
      continue;
    }
    int ax = 0;
    for (int r = 1; r < 3; ++r)
      if (bx[3 + r] - bx[r] > bx[3 + ax] - bx[ax]) ax = r;
    int mid = (it.lo + it.hi) / 2;
    std::nth_element(S.bvh_tri.begin() + it.lo, S.bvh_tri.begin() + mid, S.bvh_tri.begin() + it.hi, [&](int a, int b) {
      double ca = cen(a, ax), cb = cen(b, ax);
      return ca < cb || (ca == cb && a < b);
    });
    int l = new_node(), r = new_node();
    S.bvh_a[it.node] = l;
    S.bvh_b[it.node] = r;
    st.push_back({l, it.lo, mid});
    st.push_back({r, mid, it.hi});
  }
}

void build_self_mesh(HostScene& S, const int* tets, int T) {
  const int fc[4][3] = {{0, 1, 2}, {0, 1, 3}, {0, 2, 3}, {1, 2, 3}};
  std::vector<std::array<int, 3>> tri(4 * (size_t)T);
  for (int t = 0; t < T; ++t) {
    int sv[4] = {tets[4 * t], tets[4 * t + 1], tets[4 * t + 2], tets[4 * t + 3]};
    std::sort(sv, sv + 4);
    for (int k = 0; k < 4; ++k) tri[4 * (size_t)t + k] = {sv[fc[k][0]], sv[fc[k][1]], sv[fc[k][2]]};
  }
  std::vector<int> ord(tri.size());
  std::iota(ord.begin(), ord.end(), 0);
  std::sort(ord.begin(), ord.end(), [&](int a, int b) { return tri[a] < tri[b] || (tri[a] == tri[b] && a < b); });
  std::vector<char> uniq(tri.size(), 0);
  for (size_t k = 0; k < ord.size(); ++k) {
    bool sp = k > 0 && tri[ord[k - 1]] == tri[ord[k]], sn = k + 1 < ord.size() && tri[ord[k + 1]] == tri[ord[k]];
    uniq[ord[k]] = !(sp || sn);
  }
  S.sc_vtx.clear();
  S.sc_face.clear();
  std::vector<int> cidx(S.N, -1);  // caller vertex -> contact index
  for (size_t i = 0; i < tri.size(); ++i) {
    if (!uniq[i]) continue;
    for (int k = 0; k < 3; ++k) {
      int v = tri[i][k];
      if (cidx[v] < 0) {
        cidx[v] = (int)S.sc_vtx.size();
        S.sc_vtx.push_back(S.iperm[v]);
      }
      S.sc_face.push_back(cidx[v]);
    }
  }
  S.sc_of_v.assign(S.N, -1);
  for (size_t a = 0; a < S.sc_vtx.size(); ++a) S.sc_of_v[S.sc_vtx[a]] = (int)a;
  std::vector<int> fe(S.sc_face.size());
  for (size_t q = 0; q < fe.size(); ++q) fe[q] = S.sc_vtx[S.sc_face[q]];
  vertex_faces(S.N, fe, S.scf_ptr, S.scf_q);
}

void vertex_faces(int N, const std::vector<int>& cf, std::vector<int>& vfptr, std::vector<int>& vfq) {
  vfptr.assign(N + 1, 0);
  for (int v : cf) ++vfptr[v + 1];
  for (int i = 0; i < N; ++i) vfptr[i + 1] += vfptr[i];
  vfq.assign(cf.size(), 0);
  std::vector<int> fill(vfptr.begin(), vfptr.end() - 1);
  for (size_t q = 0; q < cf.size(); ++q) vfq[fill[cf[q]]++] = (int)q;  // q = 3 * face + corner, ascending
}

std::vector<int> surface_vertices(int N, const std::vector<int>& tv, int T) {
  std::map<std::array<int, 3>, int> cnt;
  const int fc[4][3] = {{0, 1, 2}, {0, 1, 3}, {0, 2, 3}, {1, 2, 3}};
  for (int t = 0; t < T; ++t) {
    int s[4] = {tv[4 * t], tv[4 * t + 1], tv[4 * t + 2], tv[4 * t + 3]};
    std::sort(s, s + 4);
    for (auto& f : fc) cnt[{s[f[0]], s[f[1]], s[f[2]]}]++;
  }
  std::vector<char> on(N, 0);
  for (auto& kv : cnt)
    if (kv.second == 1)
      for (int k = 0; k < 3; ++k) on[kv.first[k]] = 1;
  std::vector<int> out;
  for (int i = 0; i < N; ++i)
    if (on[i]) out.push_back(i);
  return out;
}

// Elimination tree (Liu) of a symmetric pattern given as adjacency lists.
static void etree_of(int n, const std::vector<std::vector<int>>& adj, std::vector<int>& parent) {
  parent.assign(n, -1);
  std::vector<int> anc(n, -1);
  for (int j = 0; j < n; ++j) {
    for (int i : adj[j]) {
      if (i >= j) continue;
      int r = i;
      while (anc[r] >= 0 && anc[r] != j) {
        int nx = anc[r];
        anc[r] = j;
        r = nx;
      }
      if (anc[r] < 0) {
        anc[r] = j;
        parent[r] = j;
      }
    }
  }
}

bool build_scene(HostScene& S, int N, const double* xrest, int T, const int* tets, int mass_same, double total_mass,
                 std::string& err) {
  if (N <= 0 || T <= 0) {
    err = "empty mesh";
    return false;
  }
  for (int i = 0; i < 4 * T; ++i)
    if (tets[i] < 0 || tets[i] >= N) {
      err = "tet index out of range";
      return false;
    }
  // vertex graph
  std::vector<std::vector<int>> adj(N);
  for (int t = 0; t < T; ++t)
    for (int a = 0; a < 4; ++a)
      for (int b = 0; b < 4; ++b)
        if (a != b) adj[tets[4 * t + a]].push_back(tets[4 * t + b]);
  for (auto& l : adj) {
    std::sort(l.begin(), l.end());
    l.erase(std::unique(l.begin(), l.end()), l.end());
  }
  std::vector<int> nd;
  nd_order(N, xrest, adj, nd);
  // etree postorder of the ND-permuted pattern
  std::vector<int> inv(N);
  for (int i = 0; i < N; ++i) inv[nd[i]] = i;
  std::vector<std::vector<int>> padj(N);
  for (int i = 0; i < N; ++i) {
    for (int u : adj[nd[i]]) padj[i].push_back(inv[u]);
    std::sort(padj[i].begin(), padj[i].end());
  }
  std::vector<int> par;
  etree_of(N, padj, par);
  std::vector<std::vector<int>> kids(N);
  std::vector<int> roots;
  for (int j = 0; j < N; ++j) (par[j] < 0 ? roots : kids[par[j]]).push_back(j);
  std::vector<int> post;
  post.reserve(N);
  {
    std::vector<std::pair<int, size_t>> st;
    for (int r : roots) {
      st.push_back({r, 0});
      while (!st.empty()) {
        auto& top = st.back();
        if (top.second < kids[top.first].size()) {
          int c = kids[top.first][top.second++];
          st.push_back({c, 0});
        } else {
          post.push_back(top.first);
          st.pop_back();
        }
      }
    }
  }
  S.N = N;
  S.T = T;
  S.perm.resize(N);
  S.iperm.resize(N);
  for (int i = 0; i < N; ++i) S.perm[i] = nd[post[i]];
  for (int i = 0; i < N; ++i) S.iperm[S.perm[i]] = i;
  S.xrest.resize(3 * (size_t)N);
  for (int i = 0; i < N; ++i)
    for (int r = 0; r < 3; ++r) S.xrest[3 * i + r] = xrest[3 * S.perm[i] + r];
  // tets in engine order; stable order by smallest vertex for locality
  std::vector<int> tord(T);
  std::iota(tord.begin(), tord.end(), 0);
  auto tmin = [&](int t) {
    int m = N;
    for (int k = 0; k < 4; ++k) m = std::min(m, S.iperm[tets[4 * t + k]]);
    return m;
  };
  std::vector<int> tkey(T);
  for (int t = 0; t < T; ++t) tkey[t] = tmin(t);
  std::stable_sort(tord.begin(), tord.end(), [&](int a, int b) { return tkey[a] < tkey[b]; });
  // keep the caller's tet numbering for parameters: E/nu arrays are permuted
  // by tet_order (engine tet i = caller tet tet_order[i]).
  S.tv.resize(4 * (size_t)T);
  S.tG.resize(12 * (size_t)T);
  S.tvol.resize(T);
  std::vector<double> mcaller(N, 0.0);
  double totvol = 0;
  for (int i = 0; i < T; ++i) {
    int t = tord[i];
    int v[4];
    for (int k = 0; k < 4; ++k) v[k] = tets[4 * t + k];
    double vol, G[12];
    tet_geometry(xrest, v, &vol, G);
    if (!(vol > 0)) {
      err = "degenerate tet";
      return false;
    }
    totvol += vol;
    for (int k = 0; k < 4; ++k) mcaller[tets[4 * t + k]] += 0.25 * vol;  // Dfobj.h:506-512
    for (int k = 0; k < 4; ++k) S.tv[4 * i + k] = S.iperm[v[k]];
    for (int q = 0; q < 12; ++q) S.tG[12 * (size_t)i + q] = G[q];
    S.tvol[i] = vol;
  }
  S.mass.resize(N);
  for (int i = 0; i < N; ++i)
    S.mass[i] = mass_same ? total_mass / totvol * mcaller[S.perm[i]] : total_mass / N;  // Dfobj.h:515-525
  S.E.assign(T, 2e5);
  S.nu_ctor.assign(T, 0.4);
  S.tet_order = tord;
  S.contact_vtx = surface_vertices(N, S.tv, T);
  return true;
}

// Several bodies in one world: body b's mass is lumped on its own tets
// (Dfobj.h:506-525), m_v = sum over the body's tets at v of Mass_b vol_t / (4 V_b),
// or Mass_b / N_b per vertex without "Is Mass Same".
bool set_body_masses(HostScene& S, const int* tets, int nb, const int* body_of_tet, const double* body_mass,
                     int mass_same, std::string& err) {
  if (!body_of_tet || !body_mass) {
    err = "body arrays missing";
    return false;
  }
  std::vector<double> vol_b(nb, 0.0), mc(S.N, 0.0);
  std::vector<int> vbody(S.N, -1), nv_b(nb, 0);
  for (int e = 0; e < S.T; ++e) {
    int t = S.tet_order[e], b = body_of_tet[t];
    if (b < 0 || b >= nb) {
      err = "body index out of range";
      return false;
    }
    vol_b[b] += S.tvol[e];
    for (int k = 0; k < 4; ++k) vbody[tets[4 * t + k]] = b;
  }
  for (int b = 0; b < nb; ++b)
    if (!(vol_b[b] > 0) || !(body_mass[b] > 0)) {
      err = "body without volume or mass";
      return false;
    }
  for (int i = 0; i < S.N; ++i)
    if (vbody[i] >= 0) nv_b[vbody[i]]++;
  for (int e = 0; e < S.T; ++e) {
    int t = S.tet_order[e], b = body_of_tet[t];
    for (int k = 0; k < 4; ++k) mc[tets[4 * t + k]] += 0.25 * S.tvol[e] * body_mass[b] / vol_b[b];
  }
  for (int i = 0; i < S.N; ++i) {
    int c = S.perm[i];
    S.mass[i] = mass_same ? mc[c] : (vbody[c] >= 0 ? body_mass[vbody[c]] / nv_b[vbody[c]] : 0.0);
  }
  return true;
}

void update_weights(HostScene& S) {
  if (S.nu_fwd_t.size() != (size_t)S.T) S.nu_fwd_t.assign(S.T, S.nu_fwd);
  S.wr.resize(S.T);
  S.wv.resize(S.T);
  for (int t = 0; t < S.T; ++t) {
    double mMu = S.E[t] / (1.0 + S.nu_fwd_t[t]);  // CorotateFEMConstraint.h:69; pydifem.cc:217
    S.wr[t] = mMu * S.tvol[t];
    S.wv[t] = mMu * S.tvol[t] * S.nu_ctor[t];
  }
}

void assemble_L(const HostScene& S, Csr& L) {
  int N = S.N;
  std::vector<std::map<int, double>> rows(N);
  for (int t = 0; t < S.T; ++t) {
    const double* G = &S.tG[12 * (size_t)t];
    double w = S.wr[t] + S.wv[t];
    for (int a = 0; a < 4; ++a)
      for (int b = 0; b < 4; ++b) {
        double gg = G[3 * a] * G[3 * b] + G[3 * a + 1] * G[3 * b + 1] + G[3 * a + 2] * G[3 * b + 2];
        rows[S.tv[4 * t + a]][S.tv[4 * t + b]] += w * gg;
      }
  }
  for (int m = 0; m < S.M; ++m)
    for (int a = 0; a < 4; ++a)
      for (int b = 0; b < 4; ++b) rows[S.mv[4 * m + a]][S.mv[4 * m + b]] += S.mw[m] * S.mg[4 * m + a] * S.mg[4 * m + b];
  for (size_t k = 0; k < S.att_idx.size(); ++k) rows[S.att_idx[k]][S.att_idx[k]] += S.att_w[k];
  for (int i = 0; i < N; ++i) rows[i][i] += 0.0;  // structural diagonal
  L.n = N;
  L.ptr.assign(N + 1, 0);
  L.col.clear();
  L.val.clear();
  for (int i = 0; i < N; ++i) {
    for (auto& kv : rows[i]) {
      L.col.push_back(kv.first);
      L.val.push_back(kv.second);
    }
    L.ptr[i + 1] = (int)L.col.size();
  }
}

// Supernodal symbolic factorisation. The engine order is a postorder of the
// elimination tree, so parent(j) > j and subtrees are contiguous.
bool symbolic(const Csr& A, Factor& F, std::string& err) {
  int n = A.n;
  std::vector<std::vector<int>> adj(n);
  for (int i = 0; i < n; ++i)
    for (int k = A.ptr[i]; k < A.ptr[i + 1]; ++k) adj[i].push_back(A.col[k]);
  std::vector<int> par;
  etree_of(n, adj, par);
  for (int j = 0; j < n; ++j)
    if (par[j] >= 0 && par[j] <= j) {
      err = "ordering is not a postorder";
      return false;
    }
  std::vector<int> nkids(n, 0);
  for (int j = 0; j < n; ++j)
    if (par[j] >= 0) nkids[par[j]]++;
  // column structures (rows > j) by merging children
  std::vector<std::vector<int>> cs(n);
  std::vector<std::vector<int>> kids(n);
  for (int j = 0; j < n; ++j)
    if (par[j] >= 0) kids[par[j]].push_back(j);
  for (int j = 0; j < n; ++j) {
    std::vector<int> s;
    for (int i : adj[j])
      if (i > j) s.push_back(i);
    for (int c : kids[j])
      for (int r : cs[c])
        if (r > j) s.push_back(r);
    std::sort(s.begin(), s.end());
    s.erase(std::unique(s.begin(), s.end()), s.end());
    cs[j] = std::move(s);
  }
  // fundamental supernodes
  std::vector<int> first{0};
  for (int j = 1; j < n; ++j) {
    bool merge = par[j - 1] == j && nkids[j] == 1 && cs[j - 1].size() == cs[j].size() + 1;
    if (!merge) first.push_back(j);
  }
  first.push_back(n);
  int ns = (int)first.size() - 1;
  // relaxed amalgamation of small supernodes into their parent when the
  // child is the parent's last child (contiguous columns)
  std::vector<int> sn_of(n);
  for (int s = 0; s < ns; ++s)
    for (int j = first[s]; j < first[s + 1]; ++j) sn_of[j] = s;
  std::vector<int> sfirst(first), slast(ns);
  for (int s = 0; s < ns; ++s) slast[s] = first[s + 1] - 1;
  // rows of supernode s: its columns + cs[last column]... merged structure
  std::vector<std::vector<int>> srows(ns);
  for (int s = 0; s < ns; ++s) {
    std::vector<int>& r = srows[s];
    for (int j = first[s]; j < first[s + 1]; ++j) r.push_back(j);
    for (int i : cs[first[s + 1] - 1]) r.push_back(i);
    // include rows of earlier columns (fundamental: identical apart from own cols)
    std::vector<int> extra;
    for (int j = first[s]; j < first[s + 1]; ++j)
      for (int i : cs[j])
        if (i >= first[s + 1]) extra.push_back(i);
    r.insert(r.end(), extra.begin(), extra.end());
    std::sort(r.begin(), r.end());
    r.erase(std::unique(r.begin(), r.end()), r.end());
  }
  std::vector<int> spar(ns, -1);
  for (int s = 0; s < ns; ++s) {
    int pj = par[first[s + 1] - 1];
    spar[s] = pj < 0 ? -1 : sn_of[pj];
  }
  // amalgamation pass (children before parents in index order)
  std::vector<int> alive(ns, 1), mergedInto(ns, -1);
  const int kMaxW = 64;
  for (int s = 0; s < ns; ++s) {
    int p = spar[s];
    if (p < 0 || !alive[s]) continue;
    // follow merges of p
    while (mergedInto[p] >= 0) p = mergedInto[p];
    int ws = slast[s] - sfirst[s] + 1;
    int wp = slast[p] - sfirst[p] + 1;
    if (slast[s] + 1 != sfirst[p]) continue;  // not adjacent
    if (ws + wp > kMaxW) continue;
    // merged rows
    std::vector<int> mr;
    std::set_union(srows[s].begin(), srows[s].end(), srows[p].begin(), srows[p].end(), std::back_inserter(mr));
    int64_t before = (int64_t)srows[s].size() * ws + (int64_t)srows[p].size() * wp;
    // lower-trapezoid counts are what matter; approximate with m*w
    int64_t after = (int64_t)mr.size() * (ws + wp);
    bool small = ws + wp <= 16;
    if (after <= before * (small ? 1.6 : 1.15)) {
      srows[p] = std::move(mr);
      sfirst[p] = sfirst[s];
      alive[s] = 0;
      mergedInto[s] = p;
      srows[s].clear();
    }
  }
  // compact
  std::vector<int> newid(ns, -1);
  int ns2 = 0;
  for (int s = 0; s < ns; ++s)
    if (alive[s]) newid[s] = ns2++;
  F.n = n;
  F.ns = ns2;
  F.first.assign(ns2 + 1, 0);
  F.rowptr.assign(ns2 + 1, 0);
  F.lptr.assign(ns2 + 1, 0);
  F.parent.assign(ns2, -1);
  F.rows.clear();
  F.max_m = F.max_w = 0;
  for (int s = 0; s < ns; ++s) {
    if (!alive[s]) continue;
    int id = newid[s];
    F.first[id] = sfirst[s];
    int p = spar[s];
    while (p >= 0 && mergedInto[p] >= 0) p = mergedInto[p];
    F.parent[id] = p < 0 ? -1 : newid[p];
    F.rows.insert(F.rows.end(), srows[s].begin(), srows[s].end());
    F.rowptr[id + 1] = (int64_t)F.rows.size();
    int w = slast[s] - sfirst[s] + 1;
    int m = (int)srows[s].size();
    F.lptr[id + 1] = F.lptr[id] + (int64_t)m * w;
    F.max_m = std::max(F.max_m, m);
    F.max_w = std::max(F.max_w, w);
  }
  F.first[ns2] = n;
  F.nnz = F.lptr[ns2];
  F.dptr.assign(ns2 + 1, 0);
  for (int s = 0; s < ns2; ++s) {
    int64_t w = F.first[s + 1] - F.first[s];
    F.dptr[s + 1] = F.dptr[s] + w * w;
  }
  // children lists
  F.child_ptr.assign(ns2 + 1, 0);
  for (int s = 0; s < ns2; ++s)
    if (F.parent[s] >= 0) F.child_ptr[F.parent[s] + 1]++;
  for (int s = 0; s < ns2; ++s) F.child_ptr[s + 1] += F.child_ptr[s];
  F.child.assign(F.child_ptr[ns2], 0);
  {
    std::vector<int> fill(F.child_ptr.begin(), F.child_ptr.end() - 1);
    for (int s = 0; s < ns2; ++s)
      if (F.parent[s] >= 0) F.child[fill[F.parent[s]]++] = s;
  }
  // relative indices of below rows in the parent's row list, update offsets
  F.relptr.assign(ns2 + 1, 0);
  F.uptr.assign(ns2 + 1, 0);
  F.rel.clear();
  F.flops_factor = 0;
  for (int s = 0; s < ns2; ++s) {
    int w = F.first[s + 1] - F.first[s];
    int64_t r0 = F.rowptr[s], r1 = F.rowptr[s + 1];
    int m = (int)(r1 - r0);
    int p = F.parent[s];
    if (p >= 0) {
      const int* pr = &F.rows[F.rowptr[p]];
      int pm = (int)(F.rowptr[p + 1] - F.rowptr[p]);
      for (int64_t k = r0 + w; k < r1; ++k) {
        const int* it = std::lower_bound(pr, pr + pm, F.rows[k]);
        if (it == pr + pm || *it != F.rows[k]) {
          err = "supernode structure not nested in parent";
          return false;
        }
        F.rel.push_back((int)(it - pr));
      }
    } else if (m > w) {
      err = "root supernode has below rows";
      return false;
    }
    F.relptr[s + 1] = (int64_t)F.rel.size();
    F.uptr[s + 1] = F.uptr[s] + (m - w);
    F.flops_factor += (double)w * w * w / 3.0 + (double)(m - w) * w * w + (double)(m - w) * (m - w) * w;
  }
  // levels by height (leaves = 0)
  std::vector<int> height(ns2, 0);
  for (int s = 0; s < ns2; ++s)
    if (F.parent[s] >= 0) height[F.parent[s]] = std::max(height[F.parent[s]], height[s] + 1);
  int nlev = 0;
  for (int s = 0; s < ns2; ++s) nlev = std::max(nlev, height[s] + 1);
  F.level_ptr.assign(nlev + 1, 0);
  for (int s = 0; s < ns2; ++s) F.level_ptr[height[s] + 1]++;
  for (int l = 0; l < nlev; ++l) F.level_ptr[l + 1] += F.level_ptr[l];
  F.level_sn.assign(ns2, 0);
  {
    std::vector<int> fill(F.level_ptr.begin(), F.level_ptr.end() - 1);
    // within a level, largest fronts first (better load balance)
    std::vector<int> ord(ns2);
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) {
      return (F.rowptr[a + 1] - F.rowptr[a]) > (F.rowptr[b + 1] - F.rowptr[b]);
    });
    for (int s : ord) F.level_sn[fill[height[s]]++] = s;
  }
  return true;
}

// Multifrontal numeric factorisation of A + diag(diag_add) (A = L_s CSR).
// Returns false if a pivot is not positive.
bool numeric(const Csr& A, const std::vector<double>& diag_add, Factor& F, std::string& err) {
  int ns = F.ns;
  F.L.assign(F.nnz, 0.0);
  F.Dinv.assign(F.dptr[ns], 0.0);
  std::vector<std::vector<double>> U(ns);  // update matrices (m-w)^2, row-major lower
  std::vector<int> pos(F.n, -1);
  for (int s = 0; s < ns; ++s) {
    int c0 = F.first[s], w = F.first[s + 1] - c0;
    const int* rows = &F.rows[F.rowptr[s]];
    int m = (int)(F.rowptr[s + 1] - F.rowptr[s]);
    int mb = m - w;
    double* Ls = &F.L[F.lptr[s]];  // column-major m x w
    for (int k = 0; k < m; ++k) pos[rows[k]] = k;
    // assemble original entries of columns c0..c0+w-1 (rows >= column)
    for (int j = 0; j < w; ++j) {
      int col = c0 + j;
      for (int q = A.ptr[col]; q < A.ptr[col + 1]; ++q) {
        int r = A.col[q];
        if (r < col) continue;
        Ls[(size_t)j * m + pos[r]] += A.val[q];
      }
      Ls[(size_t)j * m + j] += diag_add[col];
    }
    std::vector<double> Us((size_t)mb * mb, 0.0);
    // extend-add children
    for (int q = F.child_ptr[s]; q < F.child_ptr[s + 1]; ++q) {
      int c = F.child[q];
      int wc = F.first[c + 1] - F.first[c];
      int mc = (int)(F.rowptr[c + 1] - F.rowptr[c]);
      int mbc = mc - wc;
      const int* rel = &F.rel[F.relptr[c]];
      std::vector<double>& Uc = U[c];
      for (int a = 0; a < mbc; ++a) {
        int pa = rel[a];
        for (int b = 0; b <= a; ++b) {
          int pb = rel[b];
          double v = Uc[(size_t)a * mbc + b];
          // pa >= pb since rows are sorted and rel is increasing
          if (pb < w)
            Ls[(size_t)pb * m + pa] += v;
          else
            Us[(size_t)(pa - w) * mb + (pb - w)] += v;
        }
      }
      std::vector<double>().swap(Uc);
    }
    // dense partial Cholesky of the first w columns
    for (int j = 0; j < w; ++j) {
      double* cj = Ls + (size_t)j * m;
      for (int k = 0; k < j; ++k) {
        const double* ck = Ls + (size_t)k * m;
        double f = ck[j];
        if (f == 0.0) continue;
        for (int i = j; i < m; ++i) cj[i] -= f * ck[i];
      }
      double d = cj[j];
      if (!(d > 0.0)) {
        err = "non-positive pivot";
        return false;
      }
      d = std::sqrt(d);
      cj[j] = d;
      double inv = 1.0 / d;
      for (int i = j + 1; i < m; ++i) cj[i] *= inv;
    }
    // inverse of the diagonal triangle L_ss (row-major w x w, zero above the
    // diagonal): the device solves apply it as one dense product per supernode
    {
      double* D = &F.Dinv[F.dptr[s]];
      for (int j = 0; j < w; ++j)
        for (int r = 0; r < w; ++r) {
          if (r < j) {
            D[(size_t)r * w + j] = 0.0;
            continue;
          }
          double acc = (r == j) ? 1.0 : 0.0;
          for (int q = j; q < r; ++q) acc -= Ls[(size_t)q * m + r] * D[(size_t)q * w + j];
          D[(size_t)r * w + j] = acc / Ls[(size_t)r * m + r];
        }
    }
    // update matrix U_s = Us - L21 L21^T (lower)
    if (mb > 0) {
      for (int k = 0; k < w; ++k) {
        const double* ck = Ls + (size_t)k * m + w;
        for (int a = 0; a < mb; ++a) {
          double f = ck[a];
          if (f == 0.0) continue;
          double* ua = &Us[(size_t)a * mb];
          for (int b = 0; b <= a; ++b) ua[b] -= f * ck[b];
        }
      }
      U[s].swap(Us);
    }
    for (int k = 0; k < m; ++k) pos[rows[k]] = -1;
  }
  return true;
}

void build_corners(const HostScene& S, Corners& C) {
  int N = S.N;
  C.vptr.assign(N + 1, 0);
  for (int t = 0; t < S.T; ++t)
    for (int k = 0; k < 4; ++k) C.vptr[S.tv[4 * t + k] + 1]++;
  for (int i = 0; i < N; ++i) C.vptr[i + 1] += C.vptr[i];
  C.nslot = C.vptr[N];
  std::vector<int> fill(C.vptr.begin(), C.vptr.end() - 1);
  C.tet_slot.assign(4 * (size_t)S.T, 0);
  for (int t = 0; t < S.T; ++t)
    for (int k = 0; k < 4; ++k) C.tet_slot[4 * t + k] = fill[S.tv[4 * t + k]]++;
}

void build_strips(const HostScene& S, const Corners& C, Strips& P) {
  const int T = S.T, N = S.N, L = kStripLen, K = kStripSlots;
  // vertex -> tets
  std::vector<int> vtp(N + 1, 0), vtl(4 * (size_t)T);
  for (int q = 0; q < 4 * T; ++q) vtp[S.tv[q] + 1]++;
  for (int i = 0; i < N; ++i) vtp[i + 1] += vtp[i];
  {
    std::vector<int> fill(vtp.begin(), vtp.end() - 1);
    for (int q = 0; q < 4 * T; ++q) vtl[fill[S.tv[q]]++] = q / 4;
  }
  // Greedy strips: start at the first unused tet (engine order), then extend
  // to the unused tet sharing the most vertices with the last two tets
  // (lowest id on ties; the next unused tet when no neighbour is left).
  std::vector<char> used(T, 0);
  std::vector<int> score(T, 0), touched;
  P.tet.clear();
  int next0 = 0;
  while (true) {
    while (next0 < T && used[next0]) ++next0;
    if (next0 >= T) break;
    int cur = next0;
    for (int i = 0; i < L; ++i) {
      used[cur] = 1;
      P.tet.push_back(cur);
      if (i + 1 == L) break;
      int prev = i > 0 ? P.tet[P.tet.size() - 2] : -1;
      int best = -1;
      for (int k = 0; k < 4; ++k) {
        int v = S.tv[4 * cur + k];
        for (int a = vtp[v]; a < vtp[v + 1]; ++a) {
          int u = vtl[a];
          if (used[u]) continue;
          if (score[u] == 0) touched.push_back(u);
          score[u] += 2;  // shares v with the current tet
        }
      }
      if (prev >= 0)
        for (int u : touched)
          for (int k = 0; k < 4; ++k)
            for (int m = 0; m < 4; ++m)
              if (S.tv[4 * u + k] == S.tv[4 * prev + m]) score[u] += 1;
      for (int u : touched)
        if (best < 0 || score[u] > score[best] || (score[u] == score[best] && u < best)) best = u;
      for (int u : touched) score[u] = 0;
      touched.clear();
      if (best < 0) {
        while (next0 < T && used[next0]) ++next0;
        if (next0 >= T) break;
        best = next0;
      }
      cur = best;
    }
    while (P.tet.size() % L) P.tet.push_back(-1);
  }
  P.nstrip = (int)(P.tet.size() / L);
  P.code.assign(P.tet.size(), 0);
  P.fbase.assign(P.nstrip, 0);
  // Register slots per strip (farthest next use is evicted); a flush ends the
  // sum of a vertex in its slot. Flushes are recorded as (vertex, order).
  struct Flush {
    int v, pos, slot;
    std::vector<int> q;
  };
  std::vector<Flush> fl;
  fl.reserve(T);
  for (int s = 0; s < P.nstrip; ++s) {
    P.fbase[s] = (int)fl.size();
    const int* st = &P.tet[(size_t)s * L];
    int len = 0;
    while (len < L && st[len] >= 0) ++len;
    int sv[kStripSlots];
    std::vector<int> sq[kStripSlots];
    for (int k = 0; k < K; ++k) sv[k] = -1;
    auto next_use = [&](int v, int from) {
      for (int i = from; i < len; ++i)
        for (int k = 0; k < 4; ++k)
          if (S.tv[4 * st[i] + k] == v) return i;
      return L + 1;
    };
    auto flush = [&](int pos, int k) {
      P.code[(size_t)s * L + pos] |= 1 << (12 + k);
      fl.push_back(Flush{sv[k], pos, k, std::move(sq[k])});
      sq[k].clear();
      sv[k] = -1;
    };
    for (int i = 0; i < len; ++i) {
      const int t = st[i];
      int slot[4];
      for (int c = 0; c < 4; ++c) {
        const int v = S.tv[4 * t + c];
        int k = 0;
        while (k < K && sv[k] != v) ++k;
        if (k == K) {
          k = 0;
          while (k < K && sv[k] >= 0) ++k;
        }
        if (k == K) {  // evict the vertex used farthest ahead (never by this tet)
          int far = -1, fu = -1;
          for (int e = 0; e < K; ++e) {
            int u = next_use(sv[e], i);
            if (u > fu) fu = u, far = e;
          }
          flush(i - 1, far);  // i > 0: a slot is only full after a tet
          k = far;
        }
        sv[k] = v;
        slot[c] = k;
        sq[k].push_back(C.tet_slot[4 * t + c]);
      }
      P.code[(size_t)s * L + i] |= slot[0] | slot[1] << 3 | slot[2] << 6 | slot[3] << 9;
      for (int k = 0; k < K; ++k)
        if (sv[k] >= 0 && next_use(sv[k], i + 1) > L) flush(i, k);
    }
    // the kernel flushes position by position, slots in ascending order
    std::stable_sort(fl.begin() + P.fbase[s], fl.end(), [](const Flush& a, const Flush& b) {
      return a.pos < b.pos || (a.pos == b.pos && a.slot < b.slot);
    });
  }
  // strip sums numbered vertex by vertex (flush order within a vertex)
  const int nf = (int)fl.size();
  P.nsum = nf;
  P.ssptr.assign(N + 1, 0);
  for (const Flush& f : fl) P.ssptr[f.v + 1]++;
  for (int i = 0; i < N; ++i) P.ssptr[i + 1] += P.ssptr[i];
  std::vector<int> fill(P.ssptr.begin(), P.ssptr.end() - 1);
  P.sum_of.assign(nf, 0);
  std::vector<int> owner(nf);
  for (int e = 0; e < nf; ++e) {
    P.sum_of[e] = fill[fl[e].v]++;
    owner[P.sum_of[e]] = e;
  }
  P.fptr.assign(nf + 1, 0);
  P.fq.clear();
  P.fq.reserve(4 * (size_t)T);
  for (int p = 0; p < nf; ++p) {
    const Flush& f = fl[owner[p]];
    P.fq.insert(P.fq.end(), f.q.begin(), f.q.end());
    P.fptr[p + 1] = (int)P.fq.size();
  }
}

// Per-segment tet selection of Muscle::Initialize (Muscle.h:109-215):
// tets the segment crosses, tets whose centre lies within exp(-d/0.005) > 0.2,
// plus the tets containing the segment's start and end (added again).
namespace {
inline double svol(const double* p0, const double* p1, const double* p2, const double* p3) {
  double a[3], b[3], c[3];
  for (int r = 0; r < 3; ++r) {
    a[r] = p1[r] - p0[r];
    b[r] = p2[r] - p0[r];
    c[r] = p3[r] - p0[r];
  }
  return ((a[1] * b[2] - a[2] * b[1]) * c[0] + (a[2] * b[0] - a[0] * b[2]) * c[1] + (a[0] * b[1] - a[1] * b[0]) * c[2]) /
         6.0;
}
bool seg_crosses_tri(const double* p0, const double* p1, const double* p2, const double* s, const double* e) {
  if (svol(s, p0, p1, p2) * svol(e, p0, p1, p2) < 0.0) {
    double a = svol(s, e, p0, p1), b = svol(s, e, p1, p2), c = svol(s, e, p2, p0);
    return (a > 0 && b > 0 && c > 0) || (a < 0 && b < 0 && c < 0);
  }
  return false;
}
double abs_det_from(const double* q, const double* p1, const double* p2, const double* p3) {
  double D[9];
  for (int r = 0; r < 3; ++r) {
    D[r * 3 + 0] = p1[r] - q[r];
    D[r * 3 + 1] = p2[r] - q[r];
    D[r * 3 + 2] = p3[r] - q[r];
  }
  return std::fabs(det3(D));
}
inline double dist3(const double* a, const double* b) {
  return std::sqrt((a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]) + (a[2] - b[2]) * (a[2] - b[2]));
}
}  // namespace

int select_muscle_tets(const double* X, const int* tets, int T, const double* start, const double* end,
                       std::vector<int>& out_tet, std::vector<double>& out_w, const std::vector<int>* cand) {
  out_tet.clear();
  out_w.clear();
  double seg = dist3(start, end);
  double dir[3] = {(end[0] - start[0]) / seg, (end[1] - start[1]) / seg, (end[2] - start[2]) / seg};
  const int nj = cand ? (int)cand->size() : T;
  for (int q = 0; q < nj; ++q) {
    const int j = cand ? (*cand)[q] : q;
    int v[4] = {tets[4 * j], tets[4 * j + 1], tets[4 * j + 2], tets[4 * j + 3]};
    double Dm[9];
    dm_of(X, v[0], v[1], v[2], v[3], Dm);
    if (det3(Dm) < 0) std::swap(v[1], v[2]);
    const double *p0 = &X[3 * v[0]], *p1 = &X[3 * v[1]], *p2 = &X[3 * v[2]], *p3 = &X[3 * v[3]];
    double c[3], proj[3];
    for (int r = 0; r < 3; ++r) c[r] = 0.25 * (p0[r] + p1[r] + p2[r] + p3[r]);
    double t = (c[0] - start[0]) * dir[0] + (c[1] - start[1]) * dir[1] + (c[2] - start[2]) * dir[2];
    for (int r = 0; r < 3; ++r) proj[r] = start[r] + t * dir[r];
    double d = (std::max(dist3(proj, start), dist3(proj, end)) / seg < 1.0) ? dist3(proj, c) : 1e5;
    d = std::min(d, std::min(dist3(c, start), dist3(c, end)));
    double wgt = std::exp(-std::fabs(d) / 0.005);
    bool crosses = seg_crosses_tri(p0, p1, p2, start, end) || seg_crosses_tri(p0, p1, p3, start, end) ||
                   seg_crosses_tri(p0, p2, p3, start, end) || seg_crosses_tri(p1, p2, p3, start, end);
    if (crosses || wgt > 0.2) {
      out_tet.push_back(j);
      out_w.push_back(wgt);
    }
    double D = std::fabs(det3(Dm));
    double bs = (abs_det_from(start, p1, p2, p3) + abs_det_from(start, p2, p3, p0) + abs_det_from(start, p3, p0, p1) +
                 abs_det_from(start, p0, p1, p2)) / D;
    if (std::fabs(bs - 1.0) < 1e-4) {
      out_tet.push_back(j);
      out_w.push_back(wgt);
    }
    double be = (abs_det_from(end, p1, p2, p3) + abs_det_from(end, p2, p3, p0) + abs_det_from(end, p3, p0, p1) +
                 abs_det_from(end, p0, p1, p2)) / D;
    if (std::fabs(be - 1.0) < 1e-4) {
      out_tet.push_back(j);
      out_w.push_back(wgt);
    }
  }
  return (int)out_tet.size();
}

// LinearMuscleConstraint data per selected tet (LinearMuscleConstraint.h:89-110,
// 150-168): stencil g = [-(a+b+c), a, b, c] with (a,b,c) = invDm f, weight
// vol * k_muscle * exp-weight.
// Uniform grid over tet bounding boxes: candidate tets of a segment in
// increasing index order. A tet can be selected only if its box, grown by
// kReach, meets the segment's box: crossing needs the boxes to meet, a weight
// above 0.2 needs the centroid within 0.005*ln 5 of the segment, and the
// end-point containment test allows 1e-4 of relative volume slack (a sliver
// of the tet's size). The candidates then go through the same test in the
// same order as the full scan, so the selection is unchanged.
struct TetGrid {
  double lo[3], cell;
  int n[3];
  std::vector<int> start, items;
  std::vector<double> box;  // 6 per tet
  std::vector<int> stamp;
  int cur = 0;
  static constexpr double kReach = 0.005 * 1.6094379124341003 + 1e-6;
  void build(const double* X, const int* tets, int T) {
    box.resize(6 * (size_t)T);
    double hi[3] = {-1e300, -1e300, -1e300}, ext = 0;
    for (int r = 0; r < 3; ++r) lo[r] = 1e300;
    for (int j = 0; j < T; ++j) {
      double* b = &box[6 * (size_t)j];
      for (int r = 0; r < 3; ++r) {
        b[r] = 1e300;
        b[3 + r] = -1e300;
      }
      for (int k = 0; k < 4; ++k)
        for (int r = 0; r < 3; ++r) {
          double x = X[3 * tets[4 * j + k] + r];
          b[r] = std::min(b[r], x);
          b[3 + r] = std::max(b[3 + r], x);
        }
      double diag = std::sqrt((b[3] - b[0]) * (b[3] - b[0]) + (b[4] - b[1]) * (b[4] - b[1]) + (b[5] - b[2]) * (b[5] - b[2]));
      for (int r = 0; r < 3; ++r) {
        b[r] -= kReach + 1e-3 * diag;
        b[3 + r] += kReach + 1e-3 * diag;
        lo[r] = std::min(lo[r], b[r]);
        hi[r] = std::max(hi[r], b[3 + r]);
        ext = std::max(ext, b[3 + r] - b[r]);
      }
    }
    double dom = std::max(hi[0] - lo[0], std::max(hi[1] - lo[1], hi[2] - lo[2]));
    cell = std::max(ext, dom / 96.0);
    for (int r = 0; r < 3; ++r) n[r] = std::max(1, (int)std::ceil((hi[r] - lo[r]) / cell) + 1);
    size_t nc = (size_t)n[0] * n[1] * n[2];
    std::vector<int> cnt(nc + 1, 0);
    auto each = [&](int j, auto&& f) {
      const double* b = &box[6 * (size_t)j];
      int a0[3], a1[3];
      for (int r = 0; r < 3; ++r) {
        a0[r] = std::min(n[r] - 1, std::max(0, (int)((b[r] - lo[r]) / cell)));
        a1[r] = std::min(n[r] - 1, std::max(0, (int)((b[3 + r] - lo[r]) / cell)));
      }
      for (int x = a0[0]; x <= a1[0]; ++x)
        for (int y = a0[1]; y <= a1[1]; ++y)
          for (int z = a0[2]; z <= a1[2]; ++z) f(((size_t)x * n[1] + y) * n[2] + z);
    };
    for (int j = 0; j < T; ++j) each(j, [&](size_t c) { cnt[c + 1]++; });
    for (size_t c = 0; c < nc; ++c) cnt[c + 1] += cnt[c];
    start = cnt;
    items.resize(cnt[nc]);
    for (int j = 0; j < T; ++j) each(j, [&](size_t c) { items[cnt[c]++] = j; });
    stamp.assign(T, 0);
  }
  void query(const double* a, const double* b, std::vector<int>& out) {
    out.clear();
    ++cur;
    double q0[3], q1[3];
    int c0[3], c1[3];
    for (int r = 0; r < 3; ++r) {
      q0[r] = std::min(a[r], b[r]);
      q1[r] = std::max(a[r], b[r]);
      c0[r] = std::min(n[r] - 1, std::max(0, (int)((q0[r] - lo[r]) / cell)));
      c1[r] = std::min(n[r] - 1, std::max(0, (int)((q1[r] - lo[r]) / cell)));
    }
    for (int x = c0[0]; x <= c1[0]; ++x)
      for (int y = c0[1]; y <= c1[1]; ++y)
        for (int z = c0[2]; z <= c1[2]; ++z) {
          size_t c = ((size_t)x * n[1] + y) * n[2] + z;
          for (int i = start[c]; i < start[c + 1]; ++i) {
            int j = items[i];
            if (stamp[j] == cur) continue;
            stamp[j] = cur;
            const double* bb = &box[6 * (size_t)j];
            bool meet = true;
            for (int r = 0; r < 3; ++r) meet &= bb[r] <= q1[r] && bb[3 + r] >= q0[r];
            if (meet) out.push_back(j);
          }
        }
    std::sort(out.begin(), out.end());
  }
};

bool add_muscles(HostScene& S, const double* X, const int* tets, int n_seg, const double* seg_start,
                 const double* seg_end, const int* seg_group, int n_groups, double k_muscle, std::string& err) {
  S.mv.clear();
  S.mg.clear();
  S.mw.clear();
  S.mgroup.clear();
  S.n_groups = n_groups;
  std::vector<int> sel, mtet;
  std::vector<double> w;
  std::vector<int> etet(S.T);
  for (int e = 0; e < S.T; ++e) etet[S.tet_order[e]] = e;
  TetGrid grid;
  std::vector<int> cand;
  if (n_seg > 0) grid.build(X, tets, S.T);
  for (int s = 0; s < n_seg; ++s) {
    if (seg_group[s] < 0 || seg_group[s] >= n_groups) {
      err = "muscle segment group out of range";
      return false;
    }
    const double* st = seg_start + 3 * s;
    const double* en = seg_end + 3 * s;
    double f[3] = {en[0] - st[0], en[1] - st[1], en[2] - st[2]};
    double fl = std::sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
    if (!(fl > 0)) {
      err = "zero-length muscle segment";
      return false;
    }
    for (int r = 0; r < 3; ++r) f[r] /= fl;
    grid.query(st, en, cand);
    select_muscle_tets(X, tets, S.T, st, en, sel, w, &cand);
    for (size_t k = 0; k < sel.size(); ++k) {
      int v[4] = {tets[4 * sel[k]], tets[4 * sel[k] + 1], tets[4 * sel[k] + 2], tets[4 * sel[k] + 3]};
      double vol, G[12];
      tet_geometry(X, v, &vol, G);
      double a[3];
      for (int m = 0; m < 3; ++m) a[m] = G[(m + 1) * 3 + 0] * f[0] + G[(m + 1) * 3 + 1] * f[1] + G[(m + 1) * 3 + 2] * f[2];
      for (int q = 0; q < 4; ++q) S.mv.push_back(S.iperm[v[q]]);
      S.mg.push_back(-(a[0] + a[1] + a[2]));
      S.mg.push_back(a[0]);
      S.mg.push_back(a[1]);
      S.mg.push_back(a[2]);
      S.mw.push_back(vol * k_muscle * w[k]);
      S.mgroup.push_back(seg_group[s]);
      mtet.push_back(etet[sel[k]]);
    }
  }
  S.M = (int)S.mw.size();
  // Store the constraints grouped by engine tet (segment order kept within a
  // tet), each in its tet's corner order, so that the element kernels add a
  // tet's muscle forces into the tet's own corner slots.
  std::vector<int> ord(S.M);
  for (int m = 0; m < S.M; ++m) ord[m] = m;
  std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return mtet[a] < mtet[b]; });
  std::vector<int> mv(4 * (size_t)S.M), mgroup(S.M);
  std::vector<double> mg(4 * (size_t)S.M), mw(S.M);
  S.mptr.assign(S.T + 1, 0);
  for (int i = 0; i < S.M; ++i) {
    int m = ord[i], e = mtet[m];
    S.mptr[e + 1]++;
    for (int q = 0; q < 4; ++q) {
      int k = 0;
      while (k < 4 && S.tv[4 * (size_t)e + k] != S.mv[4 * (size_t)m + q]) ++k;
      if (k == 4) {
        err = "muscle constraint does not match its tet";
        return false;
      }
      mv[4 * (size_t)i + k] = S.mv[4 * (size_t)m + q];
      mg[4 * (size_t)i + k] = S.mg[4 * (size_t)m + q];
    }
    mw[i] = S.mw[m];
    mgroup[i] = S.mgroup[m];
  }
  for (int e = 0; e < S.T; ++e) S.mptr[e + 1] += S.mptr[e];
  S.mv.swap(mv);
  S.mg.swap(mg);
  S.mw.swap(mw);
  S.mgroup.swap(mgroup);
  return true;
}

}  // namespace pdm
