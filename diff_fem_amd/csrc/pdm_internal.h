// Internal host-side structures of libpdm (not part of the C ABI).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace pdm {

// Mesh + material + stepping parameters after setup, in the engine's own
// vertex order (nested-dissection permutation of the caller's order).
struct HostScene {
  int N = 0;                       // vertices
  int T = 0;                       // tets
  std::vector<int> tet_order;      // engine tet i = caller tet tet_order[i]
  int M = 0;                       // muscle constraints
  std::vector<int> perm;           // perm[new] = caller index
  std::vector<int> iperm;          // iperm[caller] = new index
  std::vector<double> xrest;       // 3N, engine order
  std::vector<double> mass;        // N
  // tets (engine vertex ids, oriented)
  std::vector<int> tv;             // 4T
  std::vector<double> tG;          // 12T: G[k][j] at 12t + 3k + j
  std::vector<double> tvol;        // T
  std::vector<double> E;           // T
  std::vector<double> nu_ctor;     // T
  double nu_fwd = 0.4;
  std::vector<double> nu_fwd_t;    // T: nu entering mMu per engine tet
  std::vector<double> wr, wv;      // T
  // muscles
  std::vector<int> mv;             // 4M
  std::vector<double> mg;          // 4M
  std::vector<double> mw;          // M (k * weight * vol)
  std::vector<int> mgroup;         // M
  std::vector<int> mptr;           // T+1: muscles of engine tet e are [mptr[e], mptr[e+1])
  int n_groups = 0;
  // attachments (engine ids)
  std::vector<int> att_idx;
  std::vector<double> att_w;
  // contact candidates (engine ids) and obstacles
  std::vector<int> contact_vtx;
  std::vector<double> planes;      // 6 per plane
  std::vector<double> spheres;     // 4 per sphere
  std::vector<double> tris;        // 9 per static obstacle triangle
  // bounding-volume hierarchy over tris: node k has box bvh_box[6k..6k+5]
  // (min xyz, max xyz); inner node: bvh_a = left child, bvh_b = right child;
  // leaf: bvh_a = -1 - first, bvh_b = count, its triangles bvh_tri[first..]
  std::vector<double> bvh_box;
  std::vector<int> bvh_a, bvh_b, bvh_tri;
  // underwater drag: outward-oriented surface faces (engine ids, 3 per face,
  // in the reference's face order) and, per vertex, its incident faces as
  // 3 * face + corner in ascending face order (vfptr: N+1)
  int underwater = 0;
  std::vector<int> cf;
  // self contact: surface vertices in the reference's contact numbering
  // (first appearance over the sorted surface triangles), surface triangles
  // in contact indices, and per engine vertex its incident self faces as
  // 3 * face + corner (ascending)
  int self_collision = 0;
  double self_mu = 1.0;
  std::vector<int> sc_vtx, sc_face, sc_of_v, scf_ptr, scf_q;
  std::vector<int> vfptr, vfq;
  // stepping
  double h = 0.005, damping = 0.9999;
  double g[3] = {0, 0, 0};
  int gravity_on = 1;
  int max_iter = 30, max_iter1 = 30;
  int contact = 0;
  double alpha = 1.0, mu = 1.0;
  int outer_cap = 16;
};

// Scalar sparse matrix, CSR, engine order.
struct Csr {
  int n = 0;
  std::vector<int> ptr, col;
  std::vector<double> val;
};

// Supernodal Cholesky K_s = L L^T (engine order is already fill-reducing and
// postordered, so supernodes are contiguous column ranges).
struct Factor {
  int n = 0, ns = 0;
  std::vector<int> first;          // ns+1: first column of each supernode
  std::vector<int64_t> rowptr;     // ns+1: row structure offsets
  std::vector<int> rows;           // rows of supernode s (first w are its own columns)
  std::vector<int64_t> lptr;       // ns+1: offset of the m x w column-major block
  std::vector<double> L;           // values
  std::vector<int> parent;         // ns
  std::vector<int> child_ptr, child;   // children lists (ns+1, ...), ascending order
  std::vector<int64_t> relptr;     // per supernode: offset into rel for its below rows
  std::vector<int> rel;            // position of each below row in the parent's row list
  std::vector<int64_t> uptr;       // ns+1: offset (in rows) of each update block
  std::vector<int> level_ptr, level_sn;  // forward levels by height (leaves first)
  std::vector<int64_t> dptr;       // ns+1: offset of the supernode's w x w diagonal inverse
  std::vector<double> Dinv;        // inv(L_ss) per supernode (row-major, zero above the diagonal)
  int max_m = 0, max_w = 0;
  int64_t nnz = 0;                 // stored entries of L (incl. supernode padding)
  double flops_factor = 0;         // numeric factorisation flops
  double reg_used = 0;
};

// Incident-corner lists for the deterministic tet->vertex gather.
struct Corners {
  int nslot = 0;
  std::vector<int> vptr;           // N+1
  std::vector<int> tet_slot;       // 4T: slot of (tet, corner)
};

// Tet strips of the wave-uniform local step. A wave walks the kStripLen tets
// of a strip (face neighbours where the mesh allows) and keeps the running
// corner-force sums of kStripSlots vertices in registers; when a vertex leaves
// them (no later use in the strip, or evicted for a vertex needed sooner) its
// sum is written once, as one "strip sum". A vertex's strip sums are stored
// contiguously (vptr-like ssptr), so the per-vertex gathers read them as they
// read corner slots, about a quarter as many on a structured mesh. The
// dense element kernel writes corner slots instead and k_strip_fold forms the
// same sums from them in the same order, so a vertex's force does not depend
// on which variant ran.
constexpr int kStripLen = 12;
constexpr int kStripSlots = 6;
struct Strips {
  int nstrip = 0, nsum = 0;
  std::vector<int> tet;     // nstrip * kStripLen tet ids, -1 past a short strip's end
  std::vector<int> code;    // per position: the corners' slots (3 bits each) | flush mask << 12
  std::vector<int> fbase;   // nstrip: first entry of each strip's flushes in sum_of
  std::vector<int> sum_of;  // per flush, in strip order (position, then slot): its strip sum
  std::vector<int> ssptr;   // N + 1: strip sums of each vertex
  std::vector<int> fptr;    // nsum + 1: corner slots summed into each strip sum
  std::vector<int> fq;      // 4T corner slots, in accumulation order
};
void build_strips(const HostScene& S, const Corners& C, Strips& P);

// Build the engine scene from caller arrays (see pdm.h for the meaning).
bool build_scene(HostScene& S, int N, const double* xrest, int T, const int* tets, int mass_same, double total_mass,
                 std::string& err);
void nd_order(int N, const double* xyz, const std::vector<std::vector<int>>& adj, std::vector<int>& order);
void update_weights(HostScene& S);
bool set_body_masses(HostScene& S, const int* tets, int nb, const int* body_of_tet, const double* body_mass,
                     int mass_same, std::string& err);
void assemble_L(const HostScene& S, Csr& L);  // elastic + muscle + attachment (no mass)
bool symbolic(const Csr& A, Factor& F, std::string& err);
bool numeric(const Csr& A, const std::vector<double>& diag_add, Factor& F, std::string& err);
void build_corners(const HostScene& S, Corners& C);
// cand: optional increasing list of candidate tets (all tets when null)
int select_muscle_tets(const double* x_caller, const int* tets_caller, int T, const double* start, const double* end,
                       std::vector<int>& out_tet, std::vector<double>& out_w, const std::vector<int>* cand = nullptr);
bool add_muscles(HostScene& S, const double* x_caller, const int* tets_caller, int n_seg, const double* seg_start,
                 const double* seg_end, const int* seg_group, int n_groups, double k_muscle, std::string& err);
// Outward-oriented surface faces of a tet mesh (caller numbering), in the
// order Dfobj::Initialize lists its contours (Dfobj.h:497-596).
std::vector<int> contour_faces(const double* xrest, const int* tets, int T);
// Surface mesh for self contact (Dfobj.h:497-566, World.h:540-558), from the
// caller's tets; fills S.sc_* in engine vertex ids.
void build_self_mesh(HostScene& S, const int* tets, int T);
// Median-split bounding-volume hierarchy of the obstacle triangles (S.tris).
void build_tri_bvh(HostScene& S);
// Incident faces of every vertex (3 * face + corner, ascending face order).
void vertex_faces(int N, const std::vector<int>& cf, std::vector<int>& vfptr, std::vector<int>& vfq);
std::vector<int> surface_vertices(int N, const std::vector<int>& tv, int T);

}  // namespace pdm
