// pd_oracle.cpp — CPU FP64 restatement of diff_fem's corotated-tet projective
// dynamics step and of its gradient.
//
// TEST INFRASTRUCTURE ONLY. This file is the checker: tests, the smoke check and
// the benchmark's CPU-baseline leg load it; the product library never does.
//
// It restates, in plain C++ with AoS state x[3i+c] like the reference:
//   * Dfobj::Initialize tet loop (sim/Dfobj.h:455-529): orientation fix i1<->i2,
//     vol = det(Dm)/6, invDm, volume-lumped masses scaled to the total mass.
//   * CorotateFEMConstraint (sim/fem/Constraint/CorotateFEMConstraint.h:62-77,
//     85-102, 185-205, 221-392): mMu = E/(1+nu), weights mMu*vol and mMu*vol*nu,
//     polar factor with the det<0 column flip, 5-step volume projection.
//   * LinearMuscleConstraint (sim/fem/Constraint/LinearMuscleConstraint.h:89-271)
//     and the per-segment tet selection of Muscle::Initialize (sim/Muscle.h:109-215).
//   * AttachmentConstraint (sim/fem/Constraint/AttachmentConstraint.h:70-99).
//   * World::PreComputation / FactorizeLDLT (sim/fem/World.h:1089-1139, 1212-1232)
//     as the scalar N x N system K_s (H2ML = K_s (x) I3), solved by a skyline
//     Cholesky under reverse Cuthill-McKee ordering, with the reference's
//     regularisation ladder (reg = 1e-5, 1e-4, ...).
//   * Environment::Step gravity kick (sim/Environment.h:228-259),
//     World::TimeStepping (sim/fem/World.h:673-720) and the PD local/global loop
//     with its 5e-7 early exit (sim/fem/World.h:835-919, 1601-1670).
//   * The contact phase (sim/fem/World.h:921-1023, 1317-1436;
//     sim/utils/collisionBrutal.h:281-426, 484-503) against analytic static
//     obstacles (plane, sphere) using the vertex-vs-static-face CCD rule of
//     sim/utils/arcsim/collision.cpp:127-278 (crossing time or end-of-step
//     distance below thickness 1e-4; normal opposed to the motion).
//   * The gradient the reference obtains by CppAD taping (python/pydifem.cc:
//     261-301): here a hand-derived reverse sweep over the recorded iterates,
//     control flow (iteration counts) frozen as executed.
//
// Parity pins live in tests/ (closed forms, numpy SVD, brute-force 3N assembly,
// dense solves, invariants, central finite differences; the muscle target, the
// muscle tet selection, the contact Jacobi loop and the regularisation ladder
// by hand arithmetic, closed forms of their rules and independent computations
// in tests/test_oracle_pins.py; sphere-contact gradients, whose normal moves
// with the detected positions, against central differences in
// tests/test_oracle.py).
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <complex>
#include <cstring>
#include <map>
#include <string>
#include <unordered_map>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr double kEps = 5e-7;          // World.h:214
constexpr double kThickness = 1e-4;    // arcsim collision.hpp:51

thread_local std::string g_err;

// ---------------------------------------------------------------- 3x3 helpers
// Matrices are row-major a[r*3+c].
inline double det3(const double* a) {
  return a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
         a[2] * (a[3] * a[7] - a[4] * a[6]);
}
inline void inv3(const double* a, double* o) {
  double d = det3(a);
  o[0] = (a[4] * a[8] - a[5] * a[7]) / d;
  o[1] = (a[2] * a[7] - a[1] * a[8]) / d;
  o[2] = (a[1] * a[5] - a[2] * a[4]) / d;
  o[3] = (a[5] * a[6] - a[3] * a[8]) / d;
  o[4] = (a[0] * a[8] - a[2] * a[6]) / d;
  o[5] = (a[2] * a[3] - a[0] * a[5]) / d;
  o[6] = (a[3] * a[7] - a[4] * a[6]) / d;
  o[7] = (a[1] * a[6] - a[0] * a[7]) / d;
  o[8] = (a[0] * a[4] - a[1] * a[3]) / d;
}
inline void matmul3(const double* a, const double* b, double* o) {
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c)
      o[r * 3 + c] = a[r * 3 + 0] * b[0 * 3 + c] + a[r * 3 + 1] * b[1 * 3 + c] + a[r * 3 + 2] * b[2 * 3 + c];
}
inline void matmul3_bt(const double* a, const double* b, double* o) {  // a * b^T
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c)
      o[r * 3 + c] = a[r * 3 + 0] * b[c * 3 + 0] + a[r * 3 + 1] * b[c * 3 + 1] + a[r * 3 + 2] * b[c * 3 + 2];
}
inline void matmul3_at(const double* a, const double* b, double* o) {  // a^T * b
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c)
      o[r * 3 + c] = a[0 * 3 + r] * b[0 * 3 + c] + a[1 * 3 + r] * b[1 * 3 + c] + a[2 * 3 + r] * b[2 * 3 + c];
}

// One-sided (Hestenes) Jacobi SVD of a 3x3 matrix: F = U diag(S) V^T with
// S >= 0 sorted descending, U and V orthogonal. An independent route to the
// same factors Eigen::JacobiSVD returns (CorotateFEMConstraint.h:191); the
// quantities used downstream (U V^T and U diag(.) V^T) do not depend on the
// sign/ordering freedom of the SVD for nonsingular F.
void svd3(const double* F, double* U, double* S, double* V) {
  double A[9];
  std::memcpy(A, F, sizeof(A));
  for (int i = 0; i < 9; ++i) V[i] = (i % 4 == 0) ? 1.0 : 0.0;
  const int pairs[3][2] = {{0, 1}, {0, 2}, {1, 2}};
  for (int sweep = 0; sweep < 60; ++sweep) {
    double off = 0.0;
    for (auto& pq : pairs) {
      int p = pq[0], q = pq[1];
      double alpha = 0, beta = 0, gamma = 0;
      for (int r = 0; r < 3; ++r) {
        alpha += A[r * 3 + p] * A[r * 3 + p];
        beta += A[r * 3 + q] * A[r * 3 + q];
        gamma += A[r * 3 + p] * A[r * 3 + q];
      }
      double scale = std::sqrt(alpha * beta);
      if (scale == 0.0 || std::fabs(gamma) <= 1e-300) continue;
      double rel = std::fabs(gamma) / scale;
      if (rel <= 1e-17) continue;
      off = std::max(off, rel);
      double zeta = (beta - alpha) / (2.0 * gamma);
      double t = (zeta >= 0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
      double c = 1.0 / std::sqrt(1.0 + t * t);
      double s = c * t;
      for (int r = 0; r < 3; ++r) {
        double ap = A[r * 3 + p], aq = A[r * 3 + q];
        A[r * 3 + p] = c * ap - s * aq;
        A[r * 3 + q] = s * ap + c * aq;
        double vp = V[r * 3 + p], vq = V[r * 3 + q];
        V[r * 3 + p] = c * vp - s * vq;
        V[r * 3 + q] = s * vp + c * vq;
      }
    }
    if (off <= 1e-16) break;
  }
  double sig[3];
  for (int c = 0; c < 3; ++c)
    sig[c] = std::sqrt(A[0 * 3 + c] * A[0 * 3 + c] + A[1 * 3 + c] * A[1 * 3 + c] + A[2 * 3 + c] * A[2 * 3 + c]);
  int ord[3] = {0, 1, 2};
  std::sort(ord, ord + 3, [&](int a, int b) { return sig[a] > sig[b]; });
  double A2[9], V2[9];
  for (int k = 0; k < 3; ++k) {
    S[k] = sig[ord[k]];
    for (int r = 0; r < 3; ++r) {
      A2[r * 3 + k] = A[r * 3 + ord[k]];
      V2[r * 3 + k] = V[r * 3 + ord[k]];
    }
  }
  std::memcpy(V, V2, sizeof(V2));
  double tiny = 1e-300 + 1e-15 * S[0];
  for (int k = 0; k < 3; ++k) {
    if (S[k] > tiny) {
      for (int r = 0; r < 3; ++r) U[r * 3 + k] = A2[r * 3 + k] / S[k];
    } else {
      // complete an orthonormal basis (rank-deficient F)
      double u[3] = {0, 0, 0};
      if (k == 2) {
        u[0] = U[1] * U[5] - U[2] * U[4];
        u[1] = U[2] * U[3] - U[0] * U[5];
        u[2] = U[0] * U[4] - U[1] * U[3];
      } else {
        // pick the axis least aligned with the previous columns
        for (int ax = 0; ax < 3; ++ax) {
          double e[3] = {0, 0, 0};
          e[ax] = 1;
          for (int j = 0; j < k; ++j) {
            double dot = e[0] * U[0 * 3 + j] + e[1] * U[1 * 3 + j] + e[2] * U[2 * 3 + j];
            for (int r = 0; r < 3; ++r) e[r] -= dot * U[r * 3 + j];
          }
          double n = std::sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
          if (n > 0.5) {
            for (int r = 0; r < 3; ++r) u[r] = e[r] / n;
            break;
          }
        }
      }
      for (int r = 0; r < 3; ++r) U[r * 3 + k] = u[r];
    }
  }
}

// 5-step linearised volume projection, CorotateFEMConstraint.h:362-374.
void volume_projection(const double* S, double* D) {
  D[0] = D[1] = D[2] = 0.0;
  for (int it = 0; it < 5; ++it) {
    double a0 = S[0] + D[0], a1 = S[1] + D[1], a2 = S[2] + D[2];
    double CD = a0 * a1 * a2 - 1.0;
    double g0 = a1 * a2, g1 = a0 * a2, g2 = a0 * a1;
    double gg = g0 * g0 + g1 * g1 + g2 * g2;
    double coef = (g0 * D[0] + g1 * D[1] + g2 * D[2] - CD) / gg;
    D[0] = coef * g0;
    D[1] = coef * g1;
    D[2] = coef * g2;
  }
}

// Local projection of one corotated tet at deformation gradient F:
// md (polar factor, third column negated when det F < 0) and
// md_volume = U diag(S+D) V^T (CorotateFEMConstraint.h:352-377).
void corotate_project(const double* F, double* md, double* mdv) {
  double U[9], S[3], V[9];
  svd3(F, U, S, V);
  matmul3_bt(U, V, md);
  if (det3(F) < 0) {
    md[2] = -md[2];
    md[5] = -md[5];
    md[8] = -md[8];
  }
  double D[3];
  volume_projection(S, D);
  double US[9];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) US[r * 3 + c] = U[r * 3 + c] * (S[c] + D[c]);
  matmul3_bt(US, V, mdv);
}

// Jacobian-transpose of phi(S) = S + D(S) (the 5-step projection), by forward
// mode over the 3 input directions: returns J[i][j] = d phi_i / d S_j.
// Also the divided differences rdd = ((phi_0-phi_1)/(S_0-S_1),
// (phi_0-phi_2)/(S_0-S_2), (phi_1-phi_2)/(S_1-S_2)) of phi = S + D, without
// dividing by S_i - S_j: with a = S + D the current iterate and k the third
// index, g_i - g_j = -a_k (a_i - a_j), so the new difference is
// (S_i - S_j)(1 - coef a_k r) where r is the current ratio (1 at D = 0). The
// ratio is exact also where S_i = S_j (it is then the derivative limit).
void volume_projection_jac(const double* S, double* Dout, double J[3][3], double* rdd) {
  double D[3] = {0, 0, 0};
  double dD[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};  // dD[i][j] = dD_i/dS_j
  double r[3] = {1.0, 1.0, 1.0};
  for (int it = 0; it < 5; ++it) {
    double a[3] = {S[0] + D[0], S[1] + D[1], S[2] + D[2]};
    double da[3][3];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) da[i][j] = (i == j ? 1.0 : 0.0) + dD[i][j];
    double CD = a[0] * a[1] * a[2] - 1.0;
    double g[3] = {a[1] * a[2], a[0] * a[2], a[0] * a[1]};
    double dCD[3], dg[3][3];
    for (int j = 0; j < 3; ++j) {
      dCD[j] = da[0][j] * a[1] * a[2] + a[0] * da[1][j] * a[2] + a[0] * a[1] * da[2][j];
      dg[0][j] = da[1][j] * a[2] + a[1] * da[2][j];
      dg[1][j] = da[0][j] * a[2] + a[0] * da[2][j];
      dg[2][j] = da[0][j] * a[1] + a[0] * da[1][j];
    }
    double gg = g[0] * g[0] + g[1] * g[1] + g[2] * g[2];
    double num = g[0] * D[0] + g[1] * D[1] + g[2] * D[2] - CD;
    double coef = num / gg;
    double dcoef[3];
    for (int j = 0; j < 3; ++j) {
      double dnum = dg[0][j] * D[0] + g[0] * dD[0][j] + dg[1][j] * D[1] + g[1] * dD[1][j] + dg[2][j] * D[2] +
                    g[2] * dD[2][j] - dCD[j];
      double dgg = 2.0 * (g[0] * dg[0][j] + g[1] * dg[1][j] + g[2] * dg[2][j]);
      dcoef[j] = (dnum * gg - num * dgg) / (gg * gg);
    }
    r[0] = 1.0 - coef * a[2] * r[0];
    r[1] = 1.0 - coef * a[1] * r[1];
    r[2] = 1.0 - coef * a[0] * r[2];
    double nD[3], ndD[3][3];
    for (int i = 0; i < 3; ++i) {
      nD[i] = coef * g[i];
      for (int j = 0; j < 3; ++j) ndD[i][j] = dcoef[j] * g[i] + coef * dg[i][j];
    }
    std::memcpy(D, nD, sizeof(D));
    std::memcpy(dD, ndD, sizeof(dD));
  }
  for (int i = 0; i < 3; ++i) {
    Dout[i] = D[i];
    rdd[i] = r[i];
    for (int j = 0; j < 3; ++j) J[i][j] = (i == j ? 1.0 : 0.0) + dD[i][j];
  }
}

// Vector-Jacobian product of the corotated projection:
//   Fbar = wr * (d md/dF)^T [Zr] + wv * (d mdv/dF)^T [Zv]
// using the SVD-based derivative of spectral matrix functions
// G(F) = U diag(phi(S)) V^T: with P = U^T dF V,
//   (U^T dG V)_ii = sum_j dphi_i/dS_j P_jj
//   (U^T dG V)_ij = s_ij (phi_i-phi_j)/(S_i-S_j) + a_ij (phi_i+phi_j)/(S_i+S_j), i != j,
// (the divided difference from volume_projection_jac, exact near S_i = S_j),
// s = sym part, a = skew part of P. The off-diagonal map is self-adjoint.
void corotate_vjp(const double* F, double wr, double wv, const double* Z, double* Fbar) {
  double U[9], S[3], V[9];
  svd3(F, U, S, V);
  bool flip = det3(F) < 0;
  double D[3], J[3][3], rdd[3];
  volume_projection_jac(S, D, J, rdd);
  double phi[3] = {S[0] + D[0], S[1] + D[1], S[2] + D[2]};
  // Rotation part: md = (U V^T) * Df, Df = diag(1,1,-1) if flipped.
  // d md = dR * Df  ->  adjoint input Zr = Z * Df.
  double Zr[9];
  std::memcpy(Zr, Z, sizeof(Zr));
  if (flip) {
    Zr[2] = -Zr[2];
    Zr[5] = -Zr[5];
    Zr[8] = -Zr[8];
  }
  double t1[9], Hr[9], Hv[9];
  matmul3_at(U, Zr, t1);
  matmul3(t1, V, Hr);  // U^T Zr V
  matmul3_at(U, Z, t1);
  matmul3(t1, V, Hv);  // U^T Z V
  double Out[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  // diagonal: rotation part has phi = 1 (zero Jacobian); volume part J^T
  for (int i = 0; i < 3; ++i) {
    double acc = 0;
    for (int j = 0; j < 3; ++j) acc += J[j][i] * Hv[j * 3 + j];
    Out[i * 3 + i] = wv * acc;
  }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      if (i == j) continue;
      double ssum = S[i] + S[j];
      double beta_r = ssum > 0 ? 2.0 / ssum : 0.0;
      double alpha_v, beta_v;
      alpha_v = rdd[i + j - 1];  // pair (0,1) -> 0, (0,2) -> 1, (1,2) -> 2
      beta_v = ssum > 0 ? (phi[i] + phi[j]) / ssum : 0.0;
      // Forward map on the (ij, ji) pair is [[(a+b)/2, (a-b)/2], [(a-b)/2, (a+b)/2]]
      // (symmetric), so the adjoint is the same map.
      double hr_ij = Hr[i * 3 + j], hr_ji = Hr[j * 3 + i];
      double hv_ij = Hv[i * 3 + j], hv_ji = Hv[j * 3 + i];
      double rot = 0.5 * (0.0 + beta_r) * hr_ij + 0.5 * (0.0 - beta_r) * hr_ji;
      double vol = 0.5 * (alpha_v + beta_v) * hv_ij + 0.5 * (alpha_v - beta_v) * hv_ji;
      Out[i * 3 + j] = wr * rot + wv * vol;
    }
  // Fbar = U Out V^T
  double t2[9];
  matmul3(U, Out, t2);
  matmul3_bt(t2, V, Fbar);
}

struct Tet {
  int v[4];
  double vol;
  double G[4][3];  // F[:,j] = sum_k x_k * G[k][j]
};

struct Muscle {
  int v[4];
  double g[4];   // F f = sum_k x_k g[k]
  double w;      // k_muscle * weight * vol
  int group;     // activation group
};

struct Plane { double p[3], n[3]; };
struct Sphere { double c[3], r; };

// sph >= 0: sphere obstacle, normal n = sg p/|p| with p = x0 + t (x1 - x0) - c,
// x1 the contact-phase position P_det (P_0 = PD result, P_{j+1} = output of
// taped contact iteration j), t a crossing root of |p| = r (root) or 1.
struct Impact {
  int vtx; int obs; double n[3];
  int sph = -1, det = 0;
  double t = 1.0;
  bool root = false;
};

// Skyline Cholesky of the scalar K_s in RCM order.
struct Skyline {
  int n = 0;
  std::vector<int> perm, iperm;  // perm[new] = old
  std::vector<int> first;        // first column of row i (new order)
  std::vector<int64_t> start;    // offset of row i
  std::vector<double> val;       // L entries, row i: columns first[i]..i
  bool factor(const std::vector<std::map<int, double>>& K, double reg) {
    for (int i = 0; i < n; ++i) {
      int oi = perm[i];
      double* Li = &val[start[i]];
      int fi = first[i];
      std::fill(Li, Li + (i - fi + 1), 0.0);
      for (auto& kv : K[oi]) {
        int j = iperm[kv.first];
        if (j <= i) Li[j - fi] += kv.second;
      }
      Li[i - fi] += reg;
      for (int j = fi; j < i; ++j) {
        const double* Lj = &val[start[j]];
        int fj = first[j];
        int k0 = std::max(fi, fj);
        double s = Li[j - fi];
        for (int k = k0; k < j; ++k) s -= Li[k - fi] * Lj[k - fj];
        Li[j - fi] = s / Lj[j - fj];
      }
      double s = Li[i - fi];
      for (int k = fi; k < i; ++k) s -= Li[k - fi] * Li[k - fi];
      if (!(s > 0.0)) return false;
      Li[i - fi] = std::sqrt(s);
    }
    return true;
  }
  // Solve K x = b in place on a 3-component AoS vector (one column per component).
  void solve3(double* x) const {
    std::vector<double> y(3 * (size_t)n);
    for (int i = 0; i < n; ++i)
      for (int c = 0; c < 3; ++c) y[3 * i + c] = x[3 * perm[i] + c];
    for (int i = 0; i < n; ++i) {
      const double* Li = &val[start[i]];
      int fi = first[i];
      for (int c = 0; c < 3; ++c) {
        double s = y[3 * i + c];
        for (int k = fi; k < i; ++k) s -= Li[k - fi] * y[3 * k + c];
        y[3 * i + c] = s / Li[i - fi];
      }
    }
    for (int i = n - 1; i >= 0; --i) {
      const double* Li = &val[start[i]];
      int fi = first[i];
      for (int c = 0; c < 3; ++c) {
        y[3 * i + c] /= Li[i - fi];
        double yi = y[3 * i + c];
        for (int k = fi; k < i; ++k) y[3 * k + c] -= Li[k - fi] * yi;
      }
    }
    for (int i = 0; i < n; ++i)
      for (int c = 0; c < 3; ++c) x[3 * perm[i] + c] = y[3 * i + c];
  }
};

// One executed Jacobi iteration of the contact phase, as taped for the
// reverse sweep: input ori, output x_it, whether it opened a pass, and the
// winning contact basis + pre-projection rhs of every contact vertex.
struct ContactIter {
  std::vector<double> ori, xit;
  bool first = false;
  std::vector<int> cv;        // vertices with an impact
  std::vector<double> B, f;   // 9 and 3 per cv
  std::vector<int> sph, det;  // per cv: sphere index (-1: other obstacle), detection position
  std::vector<double> st;     // per cv: sphere contact time t
  std::vector<char> sroot;    // per cv: t is a crossing root
  // self contact: whether a group acted, each group as executed (face corners,
  // basis, approaching members, final stick flags) and the forces it read
  // (rhs + obstacle reactions, 3N)
  struct Group {
    int fv[3];
    double B[9];
    std::vector<int> vl;
    std::vector<char> stick;
  };
  bool grp = false;
  std::vector<Group> groups;
  std::vector<double> F;
};

struct StepTape {
  std::vector<double> xn, vstar;           // start-of-step state (3N)
  std::vector<double> vn;                  // start-of-step velocity (3N)
  std::vector<std::vector<double>> ys;     // PD iterates y_0..y_K
  std::vector<double> act;                 // activations used this step
  bool contact_active = false;
  std::vector<ContactIter> citers;         // executed contact iterations
  bool final_latest = true;                // result = x_it of the last iteration (else its input)
  bool self_on = false;                    // self contact handled (forces carried between iterations)
};

struct Scene {
  int N = 0;
  std::vector<double> xrest;  // 3N
  std::vector<double> mass;   // N
  std::vector<Tet> tets;
  std::vector<double> E;      // per tet
  std::vector<double> nu_ctor;  // per tet (volume-term weight factor)
  std::vector<double> nu_fwd_t;  // per tet: nu entering mMu = E/(1+nu)
  std::vector<double> wr, wv;   // per tet weights
  std::vector<int> attach_idx;
  std::vector<double> attach_w;
  std::vector<Muscle> muscles;
  int n_groups = 0;
  // sim
  double h = 0.01, damping = 0.9999;
  double g[3] = {0, 0, 0};
  int gravity_on = 1;
  int max_iter = 30, max_iter1 = 30;
  int contact = 0;
  double alpha = 1.0, mu = 1.0;
  // self contact ("Self-collision", World.h:1246-1315, 1446-1586): the surface
  // mesh in the reference's contact numbering (vertices in order of first
  // appearance over the sorted surface triangles, Dfobj.h:497-566)
  int self_collision = 0;
  double self_mu = 1.0;                       // "World Friction"
  std::vector<int> sc_vtx;                    // contact index -> vertex
  std::vector<std::array<int, 3>> sc_face;    // surface triangles, contact indices
  int outer_cap = 16;  // detection passes with a Jacobi loop (the reference has no cap)
  // Self-contact impact normals held at recorded values (a checker's switch:
  // the reference computes them from plain doubles, so its taped gradient
  // treats them as constants; central differences with the normals replayed
  // measure that derivative). 0 off, 1 record, 2 replay; keyed by (step of
  // orc_rollout, detection pass, vertex, face corners).
  int sn_mode = 0, cur_step = 0;
  std::map<std::array<int, 6>, std::array<double, 3>> sn_tab;
  std::vector<Plane> planes;
  std::vector<Sphere> spheres;
  std::vector<std::array<double, 9>> tris;  // static obstacle triangles
  std::vector<int> contact_vtx;  // surface vertices
  // underwater drag (Environment.h:162-168, 236-250; Dfobj.h:846-994)
  int underwater = 0;
  std::vector<std::array<int, 3>> contours;  // outward surface faces
  // factor
  std::vector<std::map<int, double>> K;  // scalar K_s rows (original order)
  std::vector<double> Kdiag;
  Skyline sky;
  double reg_used = 0.0;
  bool factored = false;
  // tape
  std::vector<StepTape> tape;
  std::vector<int> iters_pd, iters_contact, outer_count;
};

void build_tet(const double* xr, const int* t, Tet& out) {
  int i0 = t[0], i1 = t[1], i2 = t[2], i3 = t[3];
  auto Dm_of = [&](int a, int b, int c, int d, double* Dm) {
    for (int r = 0; r < 3; ++r) {
      Dm[r * 3 + 0] = xr[3 * b + r] - xr[3 * a + r];
      Dm[r * 3 + 1] = xr[3 * c + r] - xr[3 * a + r];
      Dm[r * 3 + 2] = xr[3 * d + r] - xr[3 * a + r];
    }
  };
  double Dm[9];
  Dm_of(i0, i1, i2, i3, Dm);
  if (det3(Dm) < 0) {  // Dfobj.h:474-487
    std::swap(i1, i2);
    Dm_of(i0, i1, i2, i3, Dm);
  }
  out.v[0] = i0; out.v[1] = i1; out.v[2] = i2; out.v[3] = i3;
  out.vol = det3(Dm) / 6.0;
  double iD[9];
  inv3(Dm, iD);
  for (int j = 0; j < 3; ++j) {
    out.G[0][j] = -(iD[0 * 3 + j] + iD[1 * 3 + j] + iD[2 * 3 + j]);
    for (int m = 0; m < 3; ++m) out.G[m + 1][j] = iD[m * 3 + j];
  }
}

inline void tet_F(const Tet& t, const double* x, double* F) {
  for (int r = 0; r < 3; ++r)
    for (int j = 0; j < 3; ++j) {
      double s = 0;
      for (int k = 0; k < 4; ++k) s += x[3 * t.v[k] + r] * t.G[k][j];
      F[r * 3 + j] = s;
    }
}

void update_weights(Scene& s) {
  size_t T = s.tets.size();
  s.wr.resize(T);
  s.wv.resize(T);
  for (size_t t = 0; t < T; ++t) {
    double mMu = s.E[t] / (1.0 + s.nu_fwd_t[t]);  // CorotateFEMConstraint.h:69, pydifem.cc:217
    s.wr[t] = mMu * s.tets[t].vol;           // :247
    s.wv[t] = mMu * s.tets[t].vol * s.nu_ctor[t];  // :267
  }
}

void assemble_K(Scene& s) {
  int N = s.N;
  s.K.assign(N, {});
  double ih2 = 1.0 / (s.h * s.h);
  for (int i = 0; i < N; ++i) s.K[i][i] += s.mass[i] * ih2;
  for (size_t t = 0; t < s.tets.size(); ++t) {
    const Tet& te = s.tets[t];
    double w = s.wr[t] + s.wv[t];
    for (int a = 0; a < 4; ++a)
      for (int b = 0; b < 4; ++b) {
        double gg = te.G[a][0] * te.G[b][0] + te.G[a][1] * te.G[b][1] + te.G[a][2] * te.G[b][2];
        s.K[te.v[a]][te.v[b]] += w * gg;
      }
  }
  for (const Muscle& m : s.muscles)
    for (int a = 0; a < 4; ++a)
      for (int b = 0; b < 4; ++b) s.K[m.v[a]][m.v[b]] += m.w * m.g[a] * m.g[b];
  for (size_t k = 0; k < s.attach_idx.size(); ++k) s.K[s.attach_idx[k]][s.attach_idx[k]] += s.attach_w[k];
  s.Kdiag.assign(N, 0.0);
  for (int i = 0; i < N; ++i) s.Kdiag[i] = s.K[i][i];
}

void rcm_order(const Scene& s, std::vector<int>& perm) {
  int N = s.N;
  std::vector<int> deg(N);
  for (int i = 0; i < N; ++i) deg[i] = (int)s.K[i].size();
  std::vector<int> order;
  order.reserve(N);
  std::vector<char> seen(N, 0);
  auto bfs_last = [&](int root, std::vector<int>* levels_out) {
    std::vector<int> lev(N, -1), q{root};
    lev[root] = 0;
    for (size_t h = 0; h < q.size(); ++h) {
      int u = q[h];
      for (auto& kv : s.K[u])
        if (lev[kv.first] < 0) {
          lev[kv.first] = lev[u] + 1;
          q.push_back(kv.first);
        }
    }
    if (levels_out) *levels_out = lev;
    return q.back();
  };
  for (int comp = 0; comp < N; ++comp) {
    if (seen[comp]) continue;
    int root = comp;
    for (int k = 0; k < 3; ++k) root = bfs_last(root, nullptr);
    std::vector<int> q{root};
    seen[root] = 1;
    for (size_t hd = 0; hd < q.size(); ++hd) {
      int u = q[hd];
      std::vector<int> nb;
      for (auto& kv : s.K[u])
        if (!seen[kv.first]) {
          seen[kv.first] = 1;
          nb.push_back(kv.first);
        }
      std::sort(nb.begin(), nb.end(), [&](int a, int b) { return deg[a] != deg[b] ? deg[a] < deg[b] : a < b; });
      for (int v : nb) q.push_back(v);
    }
    for (int v : q) order.push_back(v);
  }
  perm.assign(order.rbegin(), order.rend());
}

bool factorize(Scene& s) {
  update_weights(s);
  assemble_K(s);
  Skyline& L = s.sky;
  L.n = s.N;
  rcm_order(s, L.perm);
  L.iperm.assign(s.N, 0);
  for (int i = 0; i < s.N; ++i) L.iperm[L.perm[i]] = i;
  L.first.assign(s.N, 0);
  L.start.assign(s.N + 1, 0);
  for (int i = 0; i < s.N; ++i) {
    int f = i;
    for (auto& kv : s.K[L.perm[i]]) f = std::min(f, L.iperm[kv.first]);
    L.first[i] = f;
    L.start[i + 1] = L.start[i] + (i - f + 1);
  }
  L.val.assign(L.start[s.N], 0.0);
  // FactorizeLDLT regularisation ladder, World.h:1212-1232
  double reg = 1e-6;
  bool ok = L.factor(s.K, 0.0);
  s.reg_used = 0.0;
  int tries = 0;
  while (!ok && tries < 40) {
    reg *= 10;
    ok = L.factor(s.K, reg);
    s.reg_used = reg;
    ++tries;
  }
  s.factored = ok;
  if (!ok) g_err = "factorization failed";
  return ok;
}

// d-term: accumulate J*d into out (3N), evaluated at positions x.
void add_Jd(const Scene& s, const double* x, const double* act, double* out) {
  for (size_t t = 0; t < s.tets.size(); ++t) {
    const Tet& te = s.tets[t];
    double F[9], md[9], mdv[9], P[9];
    tet_F(te, x, F);
    corotate_project(F, md, mdv);
    for (int i = 0; i < 9; ++i) P[i] = s.wr[t] * md[i] + s.wv[t] * mdv[i];
    for (int k = 0; k < 4; ++k)
      for (int r = 0; r < 3; ++r)
        out[3 * te.v[k] + r] += P[r * 3 + 0] * te.G[k][0] + P[r * 3 + 1] * te.G[k][1] + P[r * 3 + 2] * te.G[k][2];
  }
  for (const Muscle& m : s.muscles) {
    double a = act ? act[m.group] : 0.0;
    double Ff[3] = {0, 0, 0};
    for (int k = 0; k < 4; ++k)
      for (int r = 0; r < 3; ++r) Ff[r] += x[3 * m.v[k] + r] * m.g[k];
    double sc = m.w * (1.0 - a);  // p0 = (1-a) F f, LinearMuscleConstraint.h:269
    for (int k = 0; k < 4; ++k)
      for (int r = 0; r < 3; ++r) out[3 * m.v[k] + r] += sc * Ff[r] * m.g[k];
  }
  for (size_t k = 0; k < s.attach_idx.size(); ++k) {
    int i = s.attach_idx[k];
    for (int r = 0; r < 3; ++r) out[3 * i + r] += s.attach_w[k] * s.xrest[3 * i + r];
  }
}

// out += scale * L_full x   (L_full includes attachments; excludes masses)
void add_Lx(const Scene& s, const double* x, double scale, double* out) {
  for (size_t t = 0; t < s.tets.size(); ++t) {
    const Tet& te = s.tets[t];
    double F[9];
    tet_F(te, x, F);
    double w = s.wr[t] + s.wv[t];
    for (int k = 0; k < 4; ++k)
      for (int r = 0; r < 3; ++r)
        out[3 * te.v[k] + r] +=
            scale * w * (F[r * 3 + 0] * te.G[k][0] + F[r * 3 + 1] * te.G[k][1] + F[r * 3 + 2] * te.G[k][2]);
  }
  for (const Muscle& m : s.muscles) {
    double Ff[3] = {0, 0, 0};
    for (int k = 0; k < 4; ++k)
      for (int r = 0; r < 3; ++r) Ff[r] += x[3 * m.v[k] + r] * m.g[k];
    for (int k = 0; k < 4; ++k)
      for (int r = 0; r < 3; ++r) out[3 * m.v[k] + r] += scale * m.w * Ff[r] * m.g[k];
  }
  for (size_t k = 0; k < s.attach_idx.size(); ++k) {
    int i = s.attach_idx[k];
    for (int r = 0; r < 3; ++r) out[3 * i + r] += scale * s.attach_w[k] * x[3 * i + r];
  }
}

// out += scale * offdiag(K_s) (x) I3 applied to y
void add_Cy(const Scene& s, const double* y, double scale, double* out) {
  for (int i = 0; i < s.N; ++i)
    for (auto& kv : s.K[i]) {
      if (kv.first == i) continue;
      for (int r = 0; r < 3; ++r) out[3 * i + r] += scale * kv.second * y[3 * kv.first + r];
    }
}

// Static-obstacle CCD, restating arcsim collision_test for a vertex against a
// static face (collision.cpp:211-278): candidate times are the crossing roots
// in [0,1] then t = 1; an impact needs |signed distance| < thickness there.
// The contact normal opposes the vertex motion (collision.cpp:269-273).
bool detect_plane(const Plane& pl, const double* x0, const double* x1, double* n) {
  double ds = 0, de = 0, vn = 0;
  for (int r = 0; r < 3; ++r) {
    ds += (x0[r] - pl.p[r]) * pl.n[r];
    de += (x1[r] - pl.p[r]) * pl.n[r];
    vn += (x1[r] - x0[r]) * pl.n[r];
  }
  bool hit = false;
  if (ds != de) {
    double t = ds / (ds - de);
    if (t >= 0 && t <= 1) hit = true;  // distance at the crossing is 0
  }
  if (!hit && std::fabs(de) < kThickness) hit = true;
  if (!hit) return false;
  double sg = vn > 0 ? -1.0 : 1.0;
  for (int r = 0; r < 3; ++r) n[r] = sg * pl.n[r];
  return true;
}

bool detect_sphere(const Sphere& sp, const double* x0, const double* x1, double* n, double* t_out = nullptr,
                   bool* root_out = nullptr) {
  double d[3], e[3];
  for (int r = 0; r < 3; ++r) {
    d[r] = x1[r] - x0[r];
    e[r] = x0[r] - sp.c[r];
  }
  double a = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
  double b = 2 * (d[0] * e[0] + d[1] * e[1] + d[2] * e[2]);
  double c = e[0] * e[0] + e[1] * e[1] + e[2] * e[2] - sp.r * sp.r;
  double ts[3];
  int nt = 0;
  if (a > 0) {
    double disc = b * b - 4 * a * c;
    if (disc >= 0) {
      double sq = std::sqrt(disc);
      double q = -0.5 * (b + (b >= 0 ? sq : -sq));
      double r1 = q / a, r2 = q != 0 ? c / q : r1;
      if (r1 > r2) std::swap(r1, r2);
      ts[nt++] = r1;
      ts[nt++] = r2;
    }
  }
  ts[nt++] = 1.0;
  for (int k = 0; k < nt; ++k) {
    double t = ts[k];
    if (t < 0 || t > 1) continue;
    double p[3];
    for (int r = 0; r < 3; ++r) p[r] = x0[r] + t * d[r] - sp.c[r];
    double len = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
    double dist = len - sp.r;
    if (std::fabs(dist) < kThickness && len > 0) {
      double nf[3] = {p[0] / len, p[1] / len, p[2] / len};
      double vn = nf[0] * d[0] + nf[1] * d[1] + nf[2] * d[2];
      double sg = vn > 0 ? -1.0 : 1.0;
      for (int r = 0; r < 3; ++r) n[r] = sg * nf[r];
      if (t_out) *t_out = t;
      if (root_out) *root_out = k < nt - 1;
      return true;
    }
  }
  return false;
}

// computeBasis, collisionBrutal.h:408-426: columns [n, t, b].
// Vertex-vs-static-triangle test of ArcSim (collision.cpp:127-138, 211-278)
// with the face at rest: the kDOP-18 box filter in single precision
// (bvh.cpp:54-96), then the coplanarity time (the only root of the degenerate
// cubic) and t = 1; a hit lies within the thickness of the face's plane and
// inside the face, or at t = 1 within the thickness of a corner or an edge
// line. n = +-(unit(y1-y0) x unit(y2-y0)), turned against the motion.
static bool kdop18_overlap(const double* a0, const double* a1, const double* tri) {
  auto dists = [](const double* p, float* d) {
    float x = (float)p[0], y = (float)p[1], z = (float)p[2];
    d[0] = x; d[1] = y; d[2] = z;
    d[3] = x + y; d[4] = x + z; d[5] = y + z;
    d[6] = x - y; d[7] = x - z; d[8] = y - z;
  };
  float A[2][9], B[3][9], amin[9], amax[9], bmin[9], bmax[9];
  dists(a1, A[0]);
  dists(a0, A[1]);
  for (int k = 0; k < 3; ++k) dists(tri + 3 * k, B[k]);
  for (int i = 0; i < 9; ++i) {
    amin[i] = std::min(A[0][i], A[1][i]);
    amax[i] = std::max(A[0][i], A[1][i]);
    bmin[i] = std::min(std::min(B[0][i], B[1][i]), B[2][i]);
    bmax[i] = std::max(std::max(B[0][i], B[1][i]), B[2][i]);
  }
  double d = (double)(float)kThickness;
  for (int i = 0; i < 9; ++i) {
    double m = i < 3 ? d : std::sqrt(2.0) * d;
    float lo = (float)(bmin[i] - m), hi = (float)(bmax[i] + m);
    if (amin[i] > hi || amax[i] < lo) return false;
  }
  return true;
}

static double triple(const double* a, const double* b, const double* c) {
  double bc[3] = {b[1] * c[2] - b[2] * c[1], b[2] * c[0] - b[0] * c[2], b[0] * c[1] - b[1] * c[0]};
  return a[0] * bc[0] + a[1] * bc[1] + a[2] * bc[2];
}

bool detect_tri(const double* tri, const double* x0, const double* x1, double* n) {
  if (!kdop18_overlap(x0, x1, tri)) return false;
  const double* Y[3] = {tri, tri + 3, tri + 6};
  double mv[3], rel[3][3], V[3];
  for (int r = 0; r < 3; ++r) {
    mv[r] = x1[r] - x0[r];
    V[r] = 0.0 - mv[r];
    for (int k = 0; k < 3; ++k) rel[k][r] = Y[k][r] - x0[r];
  }
  double u1[3], u2[3];
  for (int r = 0; r < 3; ++r) {
    u1[r] = Y[1][r] - Y[0][r];
    u2[r] = Y[2][r] - Y[0][r];
  }
  double n1 = std::sqrt(u1[0] * u1[0] + u1[1] * u1[1] + u1[2] * u1[2]);
  double n2 = std::sqrt(u2[0] * u2[0] + u2[1] * u2[1] + u2[2] * u2[2]);
  for (int r = 0; r < 3; ++r) {
    u1[r] /= n1;
    u2[r] /= n2;
  }
  double fc[3] = {u1[1] * u2[2] - u1[2] * u2[1], u1[2] * u2[0] - u1[0] * u2[2], u1[0] * u2[1] - u1[1] * u2[0]};
  double fcc = fc[0] * fc[0] + fc[1] * fc[1] + fc[2] * fc[2];
  if (!(fcc >= 1e-6)) return false;
  double fl = std::sqrt(fcc), un[3] = {fc[0] / fl, fc[1] / fl, fc[2] / fl};
  double c0 = triple(rel[0], rel[1], rel[2]);
  double c1 = triple(V, rel[1], rel[2]) + triple(rel[0], V, rel[2]) + triple(rel[0], rel[1], V);
  std::vector<double> ts;
  if (c1 != 0.0) ts.push_back(-c0 / c1);
  ts.push_back(1.0);
  for (double t : ts) {
    if (t < 0 || t > 1) continue;
    double p[3], q[3][3];
    for (int r = 0; r < 3; ++r) p[r] = x0[r] + t * mv[r];
    for (int k = 0; k < 3; ++k)
      for (int r = 0; r < 3; ++r) q[k][r] = Y[k][r] - p[r];
    double hgt = 0;
    for (int r = 0; r < 3; ++r) hgt += (p[r] - Y[0][r]) * un[r];
    double b0 = triple(q[1], q[2], un), b1 = triple(q[2], q[0], un), b2 = triple(q[0], q[1], un);
    double w[3] = {-b0 / (b0 + b1 + b2), -b1 / (b0 + b1 + b2), -b2 / (b0 + b1 + b2)};
    bool inside = std::min(std::min(-w[0], -w[1]), -w[2]) >= -1e-6;
    if (t == 1.0) {
      for (int k = 0; k < 3; ++k) {
        double dv[3];
        for (int r = 0; r < 3; ++r) dv[r] = p[r] - Y[k][r];
        if (std::sqrt(dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2]) < kThickness) inside = true;
      }
      for (int k = 0; k < 3; ++k) {
        const double *A = Y[k], *B = Y[(k + 1) % 3];
        double pa[3], ab[3];
        for (int r = 0; r < 3; ++r) {
          pa[r] = p[r] - A[r];
          ab[r] = B[r] - A[r];
        }
        double dp = pa[0] * ab[0] + pa[1] * ab[1] + pa[2] * ab[2], ll = ab[0] * ab[0] + ab[1] * ab[1] + ab[2] * ab[2];
        double e[3];
        for (int r = 0; r < 3; ++r) e[r] = pa[r] - dp * ab[r] / ll;
        if (std::sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]) < kThickness) inside = true;
      }
    }
    double rv[3];
    for (int r = 0; r < 3; ++r) rv[r] = w[0] * V[r] + w[1] * V[r] + w[2] * V[r];
    double sg = (un[0] * rv[0] + un[1] * rv[1] + un[2] * rv[2]) > 0 ? -1.0 : 1.0;
    if (std::fabs(hgt) < kThickness && inside) {
      for (int r = 0; r < 3; ++r) n[r] = sg * fc[r];
      return true;
    }
  }
  return false;
}

// ---------------------------------------------------------------- self contact
// Coplanarity times of the moving vertex-face test (collision.cpp:211-278
// takes the roots of this cubic in [0, 1] in ascending order): the real roots
// in [0, 1] of f(t) = a3 t^3 + a2 t^2 + a1 t + a0. f is monotone between its
// stationary points; each monotone piece of [0, 1] with a sign change holds
// one root, bracketed and bisected to a width of 2^-40 (< 1e-12, the stated
// tolerance), reported as the bracket's midpoint. Zeros at piece ends are
// roots (reported once); f identically zero gives t = 0.
static int cubic_roots01(double a3, double a2, double a1, double a0, double t[3]) {
  auto f = [&](double x) { return ((a3 * x + a2) * x + a1) * x + a0; };
  if (a3 == 0.0 && a2 == 0.0 && a1 == 0.0 && a0 == 0.0) {
    t[0] = 0.0;
    return 1;
  }
  std::vector<double> cut{0.0};
  std::vector<double> st;  // stationary points: roots of 3 a3 t^2 + 2 a2 t + a1
  const double A = 3.0 * a3, B = 2.0 * a2, c = a1;
  if (A == 0.0) {
    if (B != 0.0) st.push_back(-c / B);
  } else {
    const double disc = B * B - 4.0 * A * c;
    if (disc > 0.0) {
      const double q = -0.5 * (B + std::copysign(std::sqrt(disc), B));
      st.push_back(q / A);
      if (q != 0.0) st.push_back(c / q);
    }
  }
  std::sort(st.begin(), st.end());
  for (double x : st)
    if (x > 0.0 && x < 1.0 && x != cut.back()) cut.push_back(x);
  cut.push_back(1.0);
  int n = 0;
  for (size_t k = 0; k + 1 < cut.size(); ++k) {
    double lo = cut[k], hi = cut[k + 1];
    const double flo = f(lo), fhi = f(hi);
    double r;
    if (flo == 0.0) {
      r = lo;
    } else if (fhi == 0.0) {
      r = hi;
    } else if ((flo < 0.0) != (fhi < 0.0)) {
      for (int it = 0; it < 40; ++it) {
        const double mid = 0.5 * (lo + hi), fm = f(mid);
        if (fm == 0.0) {
          lo = hi = mid;
          break;
        }
        if ((fm < 0.0) == (flo < 0.0))
          lo = mid;
        else
          hi = mid;
      }
      r = 0.5 * (lo + hi);
    } else {
      continue;
    }
    if (n == 0 || r != t[n - 1]) t[n++] = r;
  }
  return n;
}

// ArcSim's kDOP-18 box test (single precision) between the swept vertex and
// the swept face (both end points of all three corners), margin `thick`.
static bool kdop18_overlap_pts(const double* const* A, int na, const double* const* Bp, int nb, double thick) {
  auto dists = [](const double* p, float* d) {
    float x = (float)p[0], y = (float)p[1], z = (float)p[2];
    d[0] = x; d[1] = y; d[2] = z;
    d[3] = x + y; d[4] = x + z; d[5] = y + z;
    d[6] = x - y; d[7] = x - z; d[8] = y - z;
  };
  float amin[9], amax[9], bmin[9], bmax[9], q[9];
  for (int i = 0; i < 9; ++i) {
    amin[i] = bmin[i] = 3.0e38f;
    amax[i] = bmax[i] = -3.0e38f;
  }
  for (int k = 0; k < na; ++k) {
    dists(A[k], q);
    for (int i = 0; i < 9; ++i) { amin[i] = std::min(amin[i], q[i]); amax[i] = std::max(amax[i], q[i]); }
  }
  for (int k = 0; k < nb; ++k) {
    dists(Bp[k], q);
    for (int i = 0; i < 9; ++i) { bmin[i] = std::min(bmin[i], q[i]); bmax[i] = std::max(bmax[i], q[i]); }
  }
  double d = (double)(float)thick;
  for (int i = 0; i < 9; ++i) {
    double m = i < 3 ? d : std::sqrt(2.0) * d;
    float lo = (float)(bmin[i] - m), hi = (float)(bmax[i] + m);
    if (amin[i] > hi || amax[i] < lo) return false;
  }
  return true;
}

struct SelfImpact { int a, f[3]; double t, d, n[3]; };

// Moving vertex vs moving face (collision.cpp:127-278), mult = 2 as
// find_face_impacts uses for self contact. p0/p1: vertex at step start / now;
// y0[k]/y1[k]: face corner k at step start / now.
bool vf_moving(const double* p0, const double* p1, const double* const* y0, const double* const* y1, SelfImpact& im) {
  const double mult = 2.0, thk = kThickness * mult;
  const double* A[2] = {p1, p0};
  const double* Bp[6] = {y1[0], y0[0], y1[1], y0[1], y1[2], y0[2]};
  if (!kdop18_overlap_pts(A, 2, Bp, 6, kThickness * mult)) return false;
  double v0[3], x[3][3], v[3][3];
  for (int r = 0; r < 3; ++r) {
    v0[r] = p1[r] - p0[r];
    for (int k = 0; k < 3; ++k) {
      x[k][r] = y0[k][r] - p0[r];
      v[k][r] = (y1[k][r] - y0[k][r]) - v0[r];
    }
  }
  double a0 = triple(x[0], x[1], x[2]);
  double a1 = triple(v[0], x[1], x[2]) + triple(x[0], v[1], x[2]) + triple(x[0], x[1], v[2]);
  double a2 = triple(x[0], v[1], v[2]) + triple(v[0], x[1], v[2]) + triple(v[0], v[1], x[2]);
  double a3 = triple(v[0], v[1], v[2]);
  double ts[10];
  int nsol = cubic_roots01(a3, a2, a1, a0, ts);
  ts[nsol++] = 1;
  for (int i = 0; i < nsol; i++) {
    double t = ts[i];
    if (t < 0 || t > 1) continue;
    double P[3], Y[3][3];
    for (int r = 0; r < 3; ++r) {
      P[r] = p0[r] + t * (p1[r] - p0[r]);
      for (int k = 0; k < 3; ++k) Y[k][r] = y0[k][r] + t * (y1[k][r] - y0[k][r]);
    }
    // signed_vf_distance
    double e1[3], e2[3];
    for (int r = 0; r < 3; ++r) { e1[r] = Y[1][r] - Y[0][r]; e2[r] = Y[2][r] - Y[0][r]; }
    double l1 = std::sqrt(e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2]);
    double l2 = std::sqrt(e2[0] * e2[0] + e2[1] * e2[1] + e2[2] * e2[2]);
    for (int r = 0; r < 3; ++r) { e1[r] /= l1; e2[r] /= l2; }
    double nn[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
    double n2 = nn[0] * nn[0] + nn[1] * nn[1] + nn[2] * nn[2];
    if (!(n2 >= 1e-6)) continue;  // infinite distance: never a hit
    double nl = std::sqrt(n2);
    for (int r = 0; r < 3; ++r) nn[r] /= nl;
    double h = 0;
    for (int r = 0; r < 3; ++r) h += (P[r] - Y[0][r]) * nn[r];
    double q[3][3];
    for (int k = 0; k < 3; ++k)
      for (int r = 0; r < 3; ++r) q[k][r] = Y[k][r] - P[r];
    double b0 = triple(q[1], q[2], nn), b1 = triple(q[2], q[0], nn), b2 = triple(q[0], q[1], nn);
    double w[3] = {-b0 / (b0 + b1 + b2), -b1 / (b0 + b1 + b2), -b2 / (b0 + b1 + b2)};
    bool inside = std::min(std::min(-w[0], -w[1]), -w[2]) >= -1e-6;
    if (t == 1) {
      for (int k = 0; k < 3; ++k) {
        double dv[3] = {P[0] - Y[k][0], P[1] - Y[k][1], P[2] - Y[k][2]};
        inside = inside || std::sqrt(dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2]) < thk;
      }
      for (int k = 0; k < 3; ++k) {
        const double *Aa = Y[k], *Bb = Y[(k + 1) % 3];
        double pa[3], ab[3];
        for (int r = 0; r < 3; ++r) { pa[r] = P[r] - Aa[r]; ab[r] = Bb[r] - Aa[r]; }
        double dp = pa[0] * ab[0] + pa[1] * ab[1] + pa[2] * ab[2], ll = ab[0] * ab[0] + ab[1] * ab[1] + ab[2] * ab[2];
        double e[3];
        for (int r = 0; r < 3; ++r) e[r] = pa[r] - dp * ab[r] / ll;
        inside = inside || std::sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]) < thk;
      }
    }
    double rv[3];
    for (int r = 0; r < 3; ++r) rv[r] = w[0] * v[0][r] + w[1] * v[1][r] + w[2] * v[2][r];
    // normal from the step-start corners, turned against the relative motion
    double s1[3], s2[3];
    for (int r = 0; r < 3; ++r) { s1[r] = y0[1][r] - y0[0][r]; s2[r] = y0[2][r] - y0[0][r]; }
    double m1 = std::sqrt(s1[0] * s1[0] + s1[1] * s1[1] + s1[2] * s1[2]);
    double m2 = std::sqrt(s2[0] * s2[0] + s2[1] * s2[1] + s2[2] * s2[2]);
    for (int r = 0; r < 3; ++r) { s1[r] /= m1; s2[r] /= m2; }
    double c[3] = {s1[1] * s2[2] - s1[2] * s2[1], s1[2] * s2[0] - s1[0] * s2[2], s1[0] * s2[1] - s1[1] * s2[0]};
    double sg = (nn[0] * rv[0] + nn[1] * rv[1] + nn[2] * rv[2]) > 0 ? -1.0 : 1.0;
    if (std::fabs(h) < thk && inside) {
      im.t = t;
      im.d = std::fabs(h);
      for (int r = 0; r < 3; ++r) im.n[r] = sg * c[r];
      return true;
    }
  }
  return false;
}

// Stick/slide friction among a self-contact group (World::do_solve_friction,
// World.h:1246-1315), in the group's local basis.
void solve_group_friction(const double* p0, double m0, double* r0, const std::vector<std::array<double, 3>>& ps,
                          const std::vector<double>& ms, std::vector<std::array<double, 3>>& rs,
                          const std::vector<char>& stick, double mu) {
  double pst[3] = {p0[0], p0[1], p0[2]}, mst = m0;
  int n = (int)stick.size();
  for (int k = 0; k < n; ++k)
    if (stick[k]) {
      for (int r = 0; r < 3; ++r) pst[r] += ps[k][r];
      mst += ms[k];
    }
  double psl[3] = {0, 0, 0}, msl = 0, fn = 0;
  for (int k = 0; k < n; ++k)
    if (!stick[k]) {
      for (int r = 0; r < 3; ++r) psl[r] += ps[k][r];
      msl += ms[k];
      fn += rs[k][0];
    }
  double v[3] = {pst[0] / mst, pst[1] / mst, pst[2] / mst};
  r0[1] = v[1] * m0 - p0[1];
  r0[2] = v[2] * m0 - p0[2];
  for (int k = 0; k < n; ++k)
    if (stick[k]) {
      rs[k][1] = v[1] * ms[k] - ps[k][1];
      rs[k][2] = v[2] * ms[k] - ps[k][2];
    }
  if (msl > 0) {
    double rel[3];
    for (int r = 0; r < 3; ++r) rel[r] = psl[r] / msl - pst[r] / mst;
    double tl = std::sqrt(rel[1] * rel[1] + rel[2] * rel[2]);
    for (int r = 0; r < 3; ++r) rel[r] /= tl;
    double mult = mu * fn / (msl + mst);
    r0[1] += rel[1] * mult * m0;
    r0[2] += rel[2] * mult * m0;
    for (int k = 0; k < n; ++k)
      if (stick[k]) {
        rs[k][1] += rel[1] * mult * ms[k];
        rs[k][2] += rel[2] * mult * ms[k];
      }
    double av[3] = {(p0[0] + r0[0]) / m0, (p0[1] + r0[1]) / m0, (p0[2] + r0[2]) / m0};
    for (int k = 0; k < n; ++k)
      if (!stick[k]) {
        double rv[3];
        for (int r = 0; r < 3; ++r) rv[r] = ps[k][r] / ms[k] - av[r];
        double t2 = std::sqrt(rv[1] * rv[1] + rv[2] * rv[2]);
        for (int r = 0; r < 3; ++r) rv[r] /= t2;
        double m = -mu * rs[k][0] * ms[k] / (ms[k] + mst);
        rs[k][1] = m * rv[1];
        rs[k][2] = m * rv[2];
      }
  }
}

struct SelfGroup {
  int f[3];               // face, contact indices
  double n[3];            // normal of its first vertex's impact
  std::vector<int> verts; // contact indices, ascending
};

// getSelfCollisionsInfo (collisionBrutal.h:643-700): per vertex the closest
// impact of the list (smaller t; among t = 1, smaller d), then vertices
// grouped by the face they hit, in ascending vertex order.
std::vector<SelfGroup> self_groups(const std::vector<SelfImpact>& imps) {
  std::map<int, const SelfImpact*> closest;
  for (const SelfImpact& im : imps) {
    auto it = closest.find(im.a);
    if (it == closest.end()) {
      closest[im.a] = &im;
      continue;
    }
    const SelfImpact* old = it->second;
    if (im.t < old->t) {
      it->second = &im;
      continue;
    }
    if (old->t == 1 && old->d > im.d) it->second = &im;
  }
  std::vector<SelfGroup> out;
  for (auto& kv : closest) {
    const SelfImpact* im = kv.second;
    size_t g = 0;
    for (; g < out.size(); ++g)
      if (out[g].f[0] == im->f[0] && out[g].f[1] == im->f[1] && out[g].f[2] == im->f[2]) break;
    if (g < out.size()) {
      out[g].verts.push_back(kv.first);
    } else {
      SelfGroup sg;
      for (int k = 0; k < 3; ++k) sg.f[k] = im->f[k];
      for (int r = 0; r < 3; ++r) sg.n[r] = im->n[r];
      sg.verts.push_back(kv.first);
      out.push_back(sg);
    }
  }
  return out;
}

void contact_basis(const double* n, double* B) {
  double t[3] = {0, n[2], -n[1]};  // n x (1,0,0)
  double tn = std::sqrt(t[1] * t[1] + t[2] * t[2]);
  if (tn < 1e-15) {  // n x (0,1,0)
    t[0] = -n[2]; t[1] = 0; t[2] = n[0];
    tn = std::sqrt(t[0] * t[0] + t[2] * t[2]);
  }
  for (int r = 0; r < 3; ++r) t[r] /= tn;
  double b[3] = {n[1] * t[2] - n[2] * t[1], n[2] * t[0] - n[0] * t[2], n[0] * t[1] - n[1] * t[0]};
  double bn = std::sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
  double nn = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
  for (int r = 0; r < 3; ++r) {
    B[r * 3 + 0] = n[r] / nn;
    B[r * 3 + 1] = t[r];
    B[r * 3 + 2] = b[r] / bn;
  }
}

// Cone projection of World::CollisionHandling (World.h:1363-1436): the rhs of a
// contact vertex is replaced by the reaction B r.
void project_contact(const double* B, double mu, const double* f_in, double* out) {
  double fl[3];
  for (int c = 0; c < 3; ++c) fl[c] = B[0 * 3 + c] * f_in[0] + B[1 * 3 + c] * f_in[1] + B[2 * 3 + c] * f_in[2];
  if (fl[0] > 0.0) {  // take-off: unchanged
    for (int r = 0; r < 3; ++r) out[r] = f_in[r];
    return;
  }
  double rr[3] = {-fl[0], -fl[1], -fl[2]};
  double rt = std::sqrt(rr[1] * rr[1] + rr[2] * rr[2]);
  if (rt > mu * rr[0]) {
    rr[1] *= mu * rr[0] / rt;
    rr[2] *= mu * rr[0] / rt;
  }
  for (int r = 0; r < 3; ++r) out[r] = B[r * 3 + 0] * rr[0] + B[r * 3 + 1] * rr[1] + B[r * 3 + 2] * rr[2];
}

// ------------------------------------------------------------ underwater drag
// Contours of Dfobj::Initialize (Dfobj.h:497-596): the sorted-vertex faces
// (s0,s1,s2), (s0,s1,s3), (s0,s2,s3), (s1,s2,s3) of every tet, in tet order,
// that no other tet shares, turned so that n = (q1-q0) x (q2-q1) points away
// from the owning tet's centre.
std::vector<std::array<int, 3>> build_contours(const Scene& s) {
  std::vector<std::array<int, 3>> tri;
  for (const Tet& t : s.tets) {
    int sv[4] = {t.v[0], t.v[1], t.v[2], t.v[3]};
    std::sort(sv, sv + 4);
    tri.push_back({sv[0], sv[1], sv[2]});
    tri.push_back({sv[0], sv[1], sv[3]});
    tri.push_back({sv[0], sv[2], sv[3]});
    tri.push_back({sv[1], sv[2], sv[3]});
  }
  std::map<std::array<int, 3>, int> cnt;
  for (auto& f : tri) cnt[f]++;
  std::vector<std::array<int, 3>> out;
  const double* X = s.xrest.data();
  for (size_t i = 0; i < tri.size(); ++i) {
    if (cnt[tri[i]] != 1) continue;
    std::array<int, 3> f = tri[i];
    const Tet& t = s.tets[i / 4];
    double c[3], n[3], a[3], b[3];
    for (int r = 0; r < 3; ++r) {
      c[r] = 0.25 * (X[3 * t.v[0] + r] + X[3 * t.v[1] + r] + X[3 * t.v[2] + r] + X[3 * t.v[3] + r]);
      a[r] = X[3 * f[1] + r] - X[3 * f[0] + r];
      b[r] = X[3 * f[2] + r] - X[3 * f[1] + r];
    }
    n[0] = a[1] * b[2] - a[2] * b[1];
    n[1] = a[2] * b[0] - a[0] * b[2];
    n[2] = a[0] * b[1] - a[1] * b[0];
    double dp = 0;
    for (int r = 0; r < 3; ++r) dp += (c[r] - X[3 * f[0] + r]) * n[r];
    if (dp > 0.0) std::swap(f[0], f[1]);
    out.push_back(f);
  }
  return out;
}

// Face force of ComputeDragForces for n = (p2-p1) x (p3-p1) and the relative
// water velocity w: (drag + lift) / 3, each part scaled down to 1000 N.
// Written for T = double and T = complex<double>: the complex instance gives
// the Jacobian by the complex step (df/dx_k = Im f(x + i eps e_k) / eps), with
// every branch decided on real parts.
template <typename T> double re(const T& a) { return std::real(a); }
template <typename T>
void drag_face_t(const T* n, const T* w, T* F) {
  const double cl_x[5] = {-90, -10, -5, 15, 90}, cl_y[5] = {-0.5, -1.26, 0.0, 1.8, 0.5};
  for (int c = 0; c < 3; ++c) F[c] = T(0.0);
  T vv = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
  if (std::sqrt(re(vv)) < 1e-6) return;
  T vl = std::sqrt(vv);
  T vh[3] = {w[0] / vl, w[1] / vl, w[2] / vl};
  T nn = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
  T d[3] = {-(n[0] / nn), -(n[1] / nn), -(n[2] / nn)};
  T area = 0.5 * nn;
  T dd = d[0] * vh[0] + d[1] * vh[1] + d[2] * vh[2];
  T fd[3] = {T(0.0), T(0.0), T(0.0)}, fl[3] = {T(0.0), T(0.0), T(0.0)};
  if (re(dd) > 0.0) {
    T x = re(dd) > 1.0 ? T(1.0) : dd;
    T aoa = 0.5 * M_PI - std::acos(x);
    T cd = (1.0 - 0.0) * (0.5 * (-std::cos(2.0 * aoa) + 1.0)) + 0.0;
    for (int c = 0; c < 3; ++c) fd[c] = 0.5 * 1000.0 * (1.0 * area) * cd * vv * vh[c];
    T u[3] = {vh[1] * d[2] - vh[2] * d[1], vh[2] * d[0] - vh[0] * d[2], vh[0] * d[1] - vh[1] * d[0]};
    T uu = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    if (re(uu) > 0.0) {
      T ul = std::sqrt(uu);
      for (int c = 0; c < 3; ++c) u[c] = u[c] / ul;
    }
    T l[3] = {u[1] * vh[2] - u[2] * vh[1], u[2] * vh[0] - u[0] * vh[2], u[0] * vh[1] - u[1] * vh[0]};
    T theta = aoa * 180.0 / M_PI;
    T cl = T(0.0);
    double th = re(theta);
    if (std::fabs(th) <= 85.0) {
      for (int k = 0; k < 4; ++k) {
        bool in = k < 2 ? (cl_x[k] <= th && th < cl_x[k + 1]) : (cl_x[k] <= th && th <= cl_x[k + 1]);
        if (in) {
          T t = (theta - cl_x[k]) / (cl_x[k + 1] - cl_x[k]);
          cl = (1.0 - t) * cl_y[k] + t * cl_y[k + 1];
          break;
        }
      }
    }
    for (int c = 0; c < 3; ++c) fl[c] = 0.5 * 1000.0 * (1.0 * area) * cl * vv * l[c];
  }
  for (T* f : {fd, fl}) {
    T q = f[0] * f[0] + f[1] * f[1] + f[2] * f[2];
    if (std::sqrt(re(q)) > 1.0e3) {
      T ql = std::sqrt(q);
      for (int c = 0; c < 3; ++c) f[c] = 1.0e3 * (f[c] / ql);
    }
  }
  for (int c = 0; c < 3; ++c) F[c] = fd[c] / 3.0 + fl[c] / 3.0;
}

// J[c][k] = dF_c / d(n, w)_k, k = 0..5.
void drag_face_jac(const double* n, const double* w, double J[3][6]) {
  const double eps = 1e-30;
  for (int k = 0; k < 6; ++k) {
    std::complex<double> nc[3], wc[3], F[3];
    for (int c = 0; c < 3; ++c) {
      nc[c] = std::complex<double>(n[c], k == c ? eps : 0.0);
      wc[c] = std::complex<double>(w[c], k == 3 + c ? eps : 0.0);
    }
    drag_face_t(nc, wc, F);
    for (int c = 0; c < 3; ++c) J[c][k] = std::imag(F[c]) / eps;
  }
}

// Per-vertex drag forces f (3N) of the state (x, v).
void drag_forces(const Scene& s, const double* x, const double* v, double* f) {
  std::fill(f, f + 3 * (size_t)s.N, 0.0);
  for (const auto& t : s.contours) {
    double n[3], w[3], a[3], b[3], F[3];
    for (int r = 0; r < 3; ++r) {
      a[r] = x[3 * t[1] + r] - x[3 * t[0] + r];
      b[r] = x[3 * t[2] + r] - x[3 * t[0] + r];
      w[r] = 0.0 - (v[3 * t[0] + r] + v[3 * t[1] + r] + v[3 * t[2] + r]) / 3.0;
    }
    n[0] = a[1] * b[2] - a[2] * b[1];
    n[1] = a[2] * b[0] - a[0] * b[2];
    n[2] = a[0] * b[1] - a[1] * b[0];
    drag_face_t(n, w, F);
    for (int k = 0; k < 3; ++k)
      for (int r = 0; r < 3; ++r) f[3 * t[k] + r] += F[r];
  }
}

// Reverse of v* += h M^-1 f(x, v): given gvs = dL/dv*, accumulate into gx, gv.
void drag_backward(const Scene& s, const double* x, const double* v, const double* gvs, double* gx, double* gv) {
  for (const auto& t : s.contours) {
    double n[3], w[3], a[3], b[3], gF[3] = {0, 0, 0}, J[3][6];
    for (int r = 0; r < 3; ++r) {
      a[r] = x[3 * t[1] + r] - x[3 * t[0] + r];
      b[r] = x[3 * t[2] + r] - x[3 * t[0] + r];
      w[r] = 0.0 - (v[3 * t[0] + r] + v[3 * t[1] + r] + v[3 * t[2] + r]) / 3.0;
      for (int k = 0; k < 3; ++k) gF[r] += s.h / s.mass[t[k]] * gvs[3 * t[k] + r];
    }
    n[0] = a[1] * b[2] - a[2] * b[1];
    n[1] = a[2] * b[0] - a[0] * b[2];
    n[2] = a[0] * b[1] - a[1] * b[0];
    drag_face_jac(n, w, J);
    double gn[3], gw[3];
    for (int k = 0; k < 3; ++k) {
      gn[k] = gF[0] * J[0][k] + gF[1] * J[1][k] + gF[2] * J[2][k];
      gw[k] = gF[0] * J[0][3 + k] + gF[1] * J[1][3 + k] + gF[2] * J[2][3 + k];
    }
    // n = a x b with a = p2 - p1, b = p3 - p1
    double ga[3] = {b[1] * gn[2] - b[2] * gn[1], b[2] * gn[0] - b[0] * gn[2], b[0] * gn[1] - b[1] * gn[0]};
    double gb[3] = {gn[1] * a[2] - gn[2] * a[1], gn[2] * a[0] - gn[0] * a[2], gn[0] * a[1] - gn[1] * a[0]};
    for (int r = 0; r < 3; ++r) {
      gx[3 * t[0] + r] -= ga[r] + gb[r];
      gx[3 * t[1] + r] += ga[r];
      gx[3 * t[2] + r] += gb[r];
      for (int k = 0; k < 3; ++k) gv[3 * t[k] + r] -= gw[r] / 3.0;
    }
  }
}

// One time step. x,v in/out (3N). act: activations per group (or null).
void step(Scene& s, double* x, double* v, const double* act, StepTape* tape, int* it_pd, int* it_c, int* outer) {
  int N = s.N, n3 = 3 * N;
  // Environment::Step gravity kick (Environment.h:236-245) + TimeStepping (World.h:692-699)
  std::vector<double> vstar(v, v + n3);
  if (s.underwater) {  // Environment.h:246-250, World.h:692
    std::vector<double> f(n3);
    drag_forces(s, x, v, f.data());
    for (int i = 0; i < N; ++i)
      for (int r = 0; r < 3; ++r) vstar[3 * i + r] += s.h * ((1.0 / s.mass[i]) * f[3 * i + r]);
  } else if (s.gravity_on) {
    for (int i = 0; i < N; ++i)
      for (int r = 0; r < 3; ++r) vstar[3 * i + r] += s.g[r] * s.h;
  }
  std::vector<double> xn(x, x + n3), qn(n3);
  for (int k = 0; k < n3; ++k) qn[k] = xn[k] + s.h * vstar[k];
  // constant part c = M v*/h - L_full xn ; per iteration rhs = c + J d(y)
  std::vector<double> cvec(n3, 0.0);
  for (int i = 0; i < N; ++i)
    for (int r = 0; r < 3; ++r) cvec[3 * i + r] = s.mass[i] * vstar[3 * i + r] / s.h;
  add_Lx(s, xn.data(), -1.0, cvec.data());
  if (tape) {
    tape->xn = xn;
    tape->vstar = vstar;
    tape->vn.assign(v, v + n3);
    tape->ys.clear();
    tape->citers.clear();
    tape->final_latest = true;
    tape->ys.push_back(qn);
    if (act) tape->act.assign(act, act + s.n_groups);
    else tape->act.assign(s.n_groups, 0.0);
  }
  std::vector<double> y = qn, yold(n3, 0.0), rhs(n3);
  int iters = 0;
  for (int i = 0; i < s.max_iter; ++i) {  // World.h:885-916
    double e2 = 0;
    for (int k = 0; k < n3; ++k) e2 += (y[k] - yold[k]) * (y[k] - yold[k]);
    if (std::sqrt(e2) / n3 < kEps) break;
    yold = y;
    rhs = cvec;
    add_Jd(s, y.data(), act, rhs.data());
    s.sky.solve3(rhs.data());
    for (int k = 0; k < n3; ++k) y[k] = xn[k] + rhs[k];
    ++iters;
    if (tape) tape->ys.push_back(y);
  }
  if (it_pd) *it_pd = iters;
  std::vector<double> ori_x = y;
  int cit = 0, outer_n = 0;
  const bool selfon = s.self_collision && !s.sc_face.empty();
  if (s.contact && !s.contact_vtx.empty() &&
      (!s.planes.empty() || !s.spheres.empty() || !s.tris.empty() || selfon)) {
    // Contact phase, World.h:921-1023
    std::vector<Impact> impacts;
    std::vector<SelfImpact> simps;
    std::vector<SelfGroup> groups;
    // self contact: the reactions of the previous handling feed the next one
    // (double buffers of World.h:1446-1600, cleared at the step start)
    std::vector<double> self_cur(n3, 0.0), self_next(n3, 0.0), con_next(n3, 0.0);
    std::vector<double> xit = y, oold(n3, 0.0), prev(n3, 0.0), rc(n3);
    bool last_begin = true;
    if (tape && selfon) tape->self_on = true;
    while (true) {
      std::vector<Impact> fresh;
      // position read by this detection: the last taped output, or the input
      // of the last taped iteration when the pass ran out of iterations
      const int det_k = tape ? (int)tape->citers.size() - (last_begin ? 0 : 1) : 0;
      const int nplanes = (int)s.planes.size();
      for (int ci = 0; ci < (int)s.contact_vtx.size(); ++ci) {
        int vtx = s.contact_vtx[ci];
        double nrm[3];
        int ob = 0;
        for (const Plane& pl : s.planes) {
          if (detect_plane(pl, &xn[3 * vtx], &ori_x[3 * vtx], nrm)) fresh.push_back({vtx, ob, {nrm[0], nrm[1], nrm[2]}});
          ++ob;
        }
        for (const Sphere& sp : s.spheres) {
          double tt = 1.0;
          bool rt = false;
          if (detect_sphere(sp, &xn[3 * vtx], &ori_x[3 * vtx], nrm, &tt, &rt)) {
            Impact im{vtx, ob, {nrm[0], nrm[1], nrm[2]}};
            im.sph = ob - nplanes;
            im.det = det_k;
            im.t = tt;
            im.root = rt;
            fresh.push_back(im);
          }
          ++ob;
        }
        for (const auto& tr : s.tris) {
          if (detect_tri(tr.data(), &xn[3 * vtx], &ori_x[3 * vtx], nrm)) fresh.push_back({vtx, ob, {nrm[0], nrm[1], nrm[2]}});
          ++ob;
        }
      }
      for (auto& f : fresh) {  // add-or-replace by (vertex, face) key (collisionBrutal.h:312-327)
        bool found = false;
        for (auto& im : impacts)
          if (im.vtx == f.vtx && im.obs == f.obs) {
            im = f;
            found = true;
            break;
          }
        if (!found) impacts.push_back(f);
      }
      size_t ngroups_before = groups.size();
      if (selfon) {
        // vertex vs surface face of the body mesh (collision.cpp:57-64, 127-138)
        std::vector<SelfImpact> sfresh;
        int nsc = (int)s.sc_vtx.size();
        for (int a = 0; a < nsc; ++a) {
          int va = s.sc_vtx[a];
          for (const auto& f : s.sc_face) {
            if (f[0] == a || f[1] == a || f[2] == a) continue;
            const double* y0[3] = {&xn[3 * s.sc_vtx[f[0]]], &xn[3 * s.sc_vtx[f[1]]], &xn[3 * s.sc_vtx[f[2]]]};
            const double* y1[3] = {&ori_x[3 * s.sc_vtx[f[0]]], &ori_x[3 * s.sc_vtx[f[1]]], &ori_x[3 * s.sc_vtx[f[2]]]};
            SelfImpact im;
            if (vf_moving(&xn[3 * va], &ori_x[3 * va], y0, y1, im)) {
              im.a = a;
              for (int k = 0; k < 3; ++k) im.f[k] = f[k];
              if (s.sn_mode) {
                const std::array<int, 6> key = {s.cur_step, outer_n, a, f[0], f[1], f[2]};
                if (s.sn_mode == 1) {
                  s.sn_tab[key] = {im.n[0], im.n[1], im.n[2]};
                } else {
                  auto it = s.sn_tab.find(key);
                  if (it != s.sn_tab.end())
                    for (int k = 0; k < 3; ++k) im.n[k] = it->second[k];
                }
              }
              sfresh.push_back(im);
            }
          }
        }
        for (auto& f : sfresh) {  // add-or-replace by the four node ids (collisionBrutal.h:291-306)
          bool found = false;
          for (auto& im : simps)
            if (im.a == f.a && im.f[0] == f.f[0] && im.f[1] == f.f[1] && im.f[2] == f.f[2]) {
              im = f;
              found = true;
              break;
            }
          if (!found) simps.push_back(f);
        }
        groups = self_groups(simps);
      }
      // BaseUpdateCollisions: done when no obstacle impact is new and the
      // number of self-contact groups did not change (collisionBrutal.h:404)
      if (fresh.empty() && ngroups_before == groups.size()) break;
      if (outer_n == s.outer_cap) break;
      ++outer_n;
      if (tape) tape->contact_active = true;
      std::vector<double> Bs(impacts.size() * 9);
      for (size_t k = 0; k < impacts.size(); ++k) contact_basis(impacts[k].n, &Bs[9 * k]);
      for (int i = 0; i < s.max_iter1; ++i) {
        ori_x = xit;
        last_begin = true;
        double err = 0;
        for (int k = 0; k < n3; ++k) err = std::max(err, std::fabs(ori_x[k] - oold[k]));
        if (err < kEps) break;
        // rhs_speed = M v* + h (J d - L xn) - h C (ori_x - xn)
        rc = cvec;
        add_Jd(s, ori_x.data(), act, rc.data());
        std::vector<double> dy(n3);
        for (int k = 0; k < n3; ++k) dy[k] = ori_x[k] - xn[k];
        add_Cy(s, dy.data(), -1.0, rc.data());
        for (int k = 0; k < n3; ++k) rc[k] *= s.h;
        if (i == 0) prev = rc;
        std::vector<double> orig = rc;
        // the obstacle projection's input: rhs plus the previous handling's
        // self-contact forces
        std::vector<double> finall = orig;
        if (selfon)
          for (int k = 0; k < n3; ++k) finall[k] += self_cur[k];
        std::vector<ContactIter::Group> gtape;
        std::vector<double> tape_F;
        const bool have_self = selfon && !groups.empty();
        if (!impacts.empty() || have_self) {  // CollisionHandling, World.h:1317-1600
          for (size_t k = 0; k < impacts.size(); ++k) {
            int vtx = impacts[k].vtx;
            double fin[3], out[3];
            for (int r = 0; r < 3; ++r) fin[r] = orig[3 * vtx + r] + (selfon ? self_cur[3 * vtx + r] : 0.0);
            const double* B = &Bs[9 * k];
            if (B[0] * fin[0] + B[3] * fin[1] + B[6] * fin[2] > 0.0) continue;  // take-off
            project_contact(B, s.mu, fin, out);
            for (int r = 0; r < 3; ++r) {
              if (!have_self) rc[3 * vtx + r] = out[r];
              if (selfon) con_next[3 * vtx + r] = out[r];
            }
          }
          if (have_self) {
            for (const SelfGroup& g : groups) {
              double B[9];
              contact_basis(g.n, B);
              auto toloc = [&](const double* w, double* l) {
                for (int c = 0; c < 3; ++c) l[c] = B[0 * 3 + c] * w[0] + B[1 * 3 + c] * w[1] + B[2 * 3 + c] * w[2];
              };
              auto toworld = [&](const double* l, double* w) {
                for (int r = 0; r < 3; ++r) w[r] = B[r * 3 + 0] * l[0] + B[r * 3 + 1] * l[1] + B[r * 3 + 2] * l[2];
              };
              int fv[3] = {s.sc_vtx[g.f[0]], s.sc_vtx[g.f[1]], s.sc_vtx[g.f[2]]};
              double m0 = s.Kdiag[fv[0]] + s.Kdiag[fv[1]] + s.Kdiag[fv[2]];
              double f0w[3] = {0, 0, 0}, f0[3], r0[3] = {0, 0, 0};
              for (int q = 0; q < 3; ++q)
                for (int r = 0; r < 3; ++r) f0w[r] += orig[3 * fv[q] + r] + con_next[3 * fv[q] + r];
              toloc(f0w, f0);
              std::vector<int> vl;
              for (int k : g.verts) {
                int vk = s.sc_vtx[k];
                double fw[3], fl[3];
                for (int r = 0; r < 3; ++r) fw[r] = orig[3 * vk + r] + con_next[3 * vk + r];
                toloc(fw, fl);
                if (fl[0] / s.Kdiag[vk] - f0[0] / m0 <= 0) vl.push_back(vk);
              }
              int nv = (int)vl.size();
              std::vector<double> ms(nv);
              std::vector<std::array<double, 3>> fs(nv), rs(nv);
              double sm = 0, smv = 0;
              for (int k = 0; k < nv; ++k) {
                ms[k] = s.Kdiag[vl[k]];
                sm += ms[k];
                double fw[3];
                for (int r = 0; r < 3; ++r) fw[r] = orig[3 * vl[k] + r] + con_next[3 * vl[k] + r];
                toloc(fw, fs[k].data());
                smv += fs[k][0];
              }
              for (int k = 0; k < nv; ++k) {
                double fn = (f0[0] + smv) * ms[k] / (m0 + sm) - fs[k][0];
                rs[k] = {fn, 0.0, 0.0};
                r0[0] -= fn;
              }
              std::vector<char> stick(nv, 1);
              while (true) {
                solve_group_friction(f0, m0, r0, fs, ms, rs, stick, s.self_mu);
                bool changed = false;
                for (int k = 0; k < nv; ++k)
                  if (stick[k] && rs[k][1] * rs[k][1] + rs[k][2] * rs[k][2] > s.self_mu * rs[k][0] * rs[k][0]) {
                    changed = true;
                    stick[k] = 0;
                  }
                if (!changed) break;
              }
              if (tape) {
                ContactIter::Group gt;
                for (int q = 0; q < 3; ++q) gt.fv[q] = fv[q];
                std::copy(B, B + 9, gt.B);
                gt.vl = vl;
                gt.stick = stick;
                gtape.push_back(std::move(gt));
              }
              double r0w[3], bf0[3], v0[3];
              toworld(r0, r0w);
              toworld(f0, bf0);
              for (int r = 0; r < 3; ++r) v0[r] = (bf0[r] + r0w[r]) / m0;
              for (int q = 0; q < 3; ++q)
                for (int r = 0; r < 3; ++r)
                  self_next[3 * fv[q] + r] += v0[r] * s.Kdiag[fv[q]] - (orig[3 * fv[q] + r] + con_next[3 * fv[q] + r]);
              for (int k = 0; k < nv; ++k) {
                double w[3];
                toworld(rs[k].data(), w);
                for (int r = 0; r < 3; ++r) self_next[3 * vl[k] + r] += w[r];
              }
            }
            if (tape) {
              tape_F.resize(n3);
              for (int k = 0; k < n3; ++k) tape_F[k] = orig[k] + con_next[k];
            }
            for (int k = 0; k < n3; ++k) rc[k] += con_next[k] + self_next[k];
          }
          if (selfon) {
            self_cur.swap(self_next);
            std::fill(self_next.begin(), self_next.end(), 0.0);
            std::fill(con_next.begin(), con_next.end(), 0.0);
          }
        }
        for (int k = 0; k < n3; ++k) prev[k] = rc[k] * s.alpha + prev[k] * (1.0 - s.alpha);
        for (int vtx = 0; vtx < N; ++vtx)
          for (int r = 0; r < 3; ++r) xit[3 * vtx + r] = xn[3 * vtx + r] + (prev[3 * vtx + r] / s.h) / s.Kdiag[vtx];
        oold = ori_x;
        ++cit;
        last_begin = false;
        if (tape) {
          ContactIter ci;
          ci.ori = ori_x;
          ci.xit = xit;
          ci.first = i == 0;
          if (have_self) {
            ci.grp = true;
            ci.groups.swap(gtape);
            ci.F.swap(tape_F);
          }
          std::map<int, size_t> win;  // last impact of each vertex wins
          for (size_t k = 0; k < impacts.size(); ++k) win[impacts[k].vtx] = k;
          for (auto& kv : win) {
            ci.cv.push_back(kv.first);
            ci.B.insert(ci.B.end(), &Bs[9 * kv.second], &Bs[9 * kv.second] + 9);
            ci.f.insert(ci.f.end(), &finall[3 * kv.first], &finall[3 * kv.first] + 3);
            const Impact& w = impacts[kv.second];
            ci.sph.push_back(w.sph);
            ci.det.push_back(w.det);
            ci.st.push_back(w.t);
            ci.sroot.push_back(w.root ? 1 : 0);
          }
          tape->citers.push_back(std::move(ci));
        }
      }
    }
    if (tape) tape->final_latest = last_begin;
  }
  if (it_c) *it_c = cit;
  if (outer) *outer = outer_n;
  // UpdatePositionsAndVelocities + damping (World.h:757-766, 708)
  for (int k = 0; k < n3; ++k) {
    v[k] = (ori_x[k] - xn[k]) * (1.0 / s.h);
    x[k] = ori_x[k];
    v[k] *= s.damping;
  }
}


// VJP of the cone projection at a contact vertex (frozen basis and branch):
// given qbar (adjoint of the projected rhs) return the adjoint of the rhs.
void project_contact_vjp(const double* B, double mu, const double* f_in, const double* qbar, double* out) {
  double fl[3], rb[3];
  for (int c = 0; c < 3; ++c) {
    fl[c] = B[0 * 3 + c] * f_in[0] + B[1 * 3 + c] * f_in[1] + B[2 * 3 + c] * f_in[2];
    rb[c] = B[0 * 3 + c] * qbar[0] + B[1 * 3 + c] * qbar[1] + B[2 * 3 + c] * qbar[2];
  }
  if (fl[0] > 0.0) {
    for (int r = 0; r < 3; ++r) out[r] = qbar[r];
    return;
  }
  double ft = std::sqrt(fl[1] * fl[1] + fl[2] * fl[2]);
  double fb[3];
  if (!(ft > mu * (-fl[0]))) {  // stick: r = -f
    fb[0] = -rb[0];
    fb[1] = -rb[1];
    fb[2] = -rb[2];
  } else {  // slide: r_n = -f_n, r_t = mu f_n f_t / |f_t|
    double u1 = fl[1] / ft, u2 = fl[2] / ft;
    double ur = u1 * rb[1] + u2 * rb[2];
    fb[0] = -rb[0] + mu * ur;
    double sc = mu * fl[0] / ft;
    fb[1] = sc * (rb[1] - ur * u1);
    fb[2] = sc * (rb[2] - ur * u2);
  }
  for (int r = 0; r < 3; ++r) out[r] = B[r * 3 + 0] * fb[0] + B[r * 3 + 1] * fb[1] + B[r * 3 + 2] * fb[2];
}

// Adjoint of the cone projection with respect to its unit normal n (branch
// frozen). The take-off reply f and the stick reply -f do not depend on n; the
// sliding reply is out = -phi n + mu phi e with phi = n.f, e = f_t/|f_t|,
// f_t = f - phi n, whose differential in n gives
// nbar = f (-ob.n + mu ob.e - g.n) - phi (ob + g), g = mu phi (ob - e (e.ob))/|f_t|.
void project_contact_nvjp(const double* n, double mu, const double* f, const double* ob, double* nbar) {
  nbar[0] = nbar[1] = nbar[2] = 0.0;
  const double phi = n[0] * f[0] + n[1] * f[1] + n[2] * f[2];
  if (phi > 0.0) return;
  double ft[3];
  for (int r = 0; r < 3; ++r) ft[r] = f[r] - phi * n[r];
  const double tau = std::sqrt(ft[0] * ft[0] + ft[1] * ft[1] + ft[2] * ft[2]);
  if (!(tau > mu * (-phi))) return;
  double e[3] = {ft[0] / tau, ft[1] / tau, ft[2] / tau};
  const double obn = ob[0] * n[0] + ob[1] * n[1] + ob[2] * n[2];
  const double obe = ob[0] * e[0] + ob[1] * e[1] + ob[2] * e[2];
  const double sc = mu * phi / tau;
  const double coef = -obn + mu * obe - sc * obn;
  for (int r = 0; r < 3; ++r) nbar[r] = f[r] * coef - phi * (ob[r] + sc * (ob[r] - e[r] * obe));
}

// Adjoint of a sphere contact normal n = sg p/|p|, p = x0 + t (x1 - x0) - c, to
// the two positions: t = 1 moves with neither, a crossing root t of
// |p(t)| = r moves as dt = -p.q / p.d (d = x1 - x0, q = (1-t) dx0 + t dx1), so
// the adjoint of q is w = pbar - p (pbar.d)/(p.d).
void sphere_normal_vjp(const double* c, const double* x0, const double* x1, double t, bool root, const double* n,
                       const double* nbar, double* x0bar, double* x1bar) {
  double p[3], d[3];
  for (int r = 0; r < 3; ++r) {
    d[r] = x1[r] - x0[r];
    p[r] = x0[r] + t * d[r] - c[r];
  }
  const double len = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
  const double sg = (n[0] * p[0] + n[1] * p[1] + n[2] * p[2]) > 0.0 ? 1.0 : -1.0;
  const double nn = n[0] * nbar[0] + n[1] * nbar[1] + n[2] * nbar[2];
  double pb[3];
  for (int r = 0; r < 3; ++r) pb[r] = sg * (nbar[r] - n[r] * nn) / len;
  if (!root) {
    for (int r = 0; r < 3; ++r) x1bar[r] += pb[r];
    return;
  }
  const double pd = p[0] * d[0] + p[1] * d[1] + p[2] * d[2];
  const double bd = pb[0] * d[0] + pb[1] * d[1] + pb[2] * d[2];
  for (int r = 0; r < 3; ++r) {
    const double w = pb[r] - p[r] * bd / pd;
    x0bar[r] += (1.0 - t) * w;
    x1bar[r] += t * w;
  }
}

// Dual numbers (value and one directional derivative): the self-contact
// group's Jacobian by forward differentiation, one input direction at a time.
struct Dual {
  double v = 0.0, d = 0.0;
  Dual() = default;
  Dual(double a, double b = 0.0) : v(a), d(b) {}
};
inline Dual operator+(Dual a, Dual b) { return Dual(a.v + b.v, a.d + b.d); }
inline Dual operator-(Dual a, Dual b) { return Dual(a.v - b.v, a.d - b.d); }
inline Dual operator*(Dual a, Dual b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
inline Dual operator/(Dual a, Dual b) { return Dual(a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)); }
inline Dual sqrt(Dual a) {
  const double r = std::sqrt(a.v);
  return Dual(r, a.d / (2.0 * r));
}
inline double val(double a) { return a; }
inline double val(Dual a) { return a.v; }

// One self-contact group of a handling (World.h:1451-1530) as executed: the
// members that approached (in order) and the final stick flags of
// do_solve_friction (World.h:1246-1315) are given. F: forces (rhs + obstacle
// reactions) of the three face corners, then of the members, 3 each; m: their
// masses (diagonal of K_s). out: the changes the group adds to the
// self-contact forces, same layout. B: the frozen basis [n t b] (row-major).
template <typename T>
void group_apply(const double* B, int nv, const T* F, const T* m, const char* stick, double mu, T* out) {
  using std::sqrt;
  auto toloc = [&](const T* w, T* l) {
    for (int c = 0; c < 3; ++c) l[c] = T(B[0 * 3 + c]) * w[0] + T(B[1 * 3 + c]) * w[1] + T(B[2 * 3 + c]) * w[2];
  };
  auto toworld = [&](const T* l, T* w) {
    for (int r = 0; r < 3; ++r) w[r] = T(B[r * 3 + 0]) * l[0] + T(B[r * 3 + 1]) * l[1] + T(B[r * 3 + 2]) * l[2];
  };
  T f0w[3] = {T(0.0), T(0.0), T(0.0)}, f0[3];
  for (int q = 0; q < 3; ++q)
    for (int r = 0; r < 3; ++r) f0w[r] = f0w[r] + F[3 * q + r];
  toloc(f0w, f0);
  const T m0 = m[0] + m[1] + m[2];
  std::vector<std::array<T, 3>> fs(nv), rs(nv);
  std::vector<T> ms(nv);
  T sm(0.0), smv(0.0);
  for (int k = 0; k < nv; ++k) {
    ms[k] = m[3 + k];
    toloc(&F[3 * (3 + k)], fs[k].data());
    sm = sm + ms[k];
    smv = smv + fs[k][0];
  }
  T r0[3] = {T(0.0), T(0.0), T(0.0)};
  for (int k = 0; k < nv; ++k) {
    T fn = (f0[0] + smv) * ms[k] / (m0 + sm) - fs[k][0];
    rs[k] = {fn, T(0.0), T(0.0)};
    r0[0] = r0[0] - fn;
  }
  T pst[3] = {f0[0], f0[1], f0[2]}, mst = m0, psl[3] = {T(0.0), T(0.0), T(0.0)}, msl(0.0), fnl(0.0);
  for (int k = 0; k < nv; ++k) {
    if (stick[k]) {
      for (int r = 0; r < 3; ++r) pst[r] = pst[r] + fs[k][r];
      mst = mst + ms[k];
    } else {
      for (int r = 0; r < 3; ++r) psl[r] = psl[r] + fs[k][r];
      msl = msl + ms[k];
      fnl = fnl + rs[k][0];
    }
  }
  T v[3] = {pst[0] / mst, pst[1] / mst, pst[2] / mst};
  r0[1] = v[1] * m0 - f0[1];
  r0[2] = v[2] * m0 - f0[2];
  for (int k = 0; k < nv; ++k)
    if (stick[k]) {
      rs[k][1] = v[1] * ms[k] - fs[k][1];
      rs[k][2] = v[2] * ms[k] - fs[k][2];
    }
  if (val(msl) > 0) {
    T rel[3];
    for (int r = 0; r < 3; ++r) rel[r] = psl[r] / msl - pst[r] / mst;
    T tl = sqrt(rel[1] * rel[1] + rel[2] * rel[2]);
    for (int r = 0; r < 3; ++r) rel[r] = rel[r] / tl;
    T mult = T(mu) * fnl / (msl + mst);
    r0[1] = r0[1] + rel[1] * mult * m0;
    r0[2] = r0[2] + rel[2] * mult * m0;
    for (int k = 0; k < nv; ++k)
      if (stick[k]) {
        rs[k][1] = rs[k][1] + rel[1] * mult * ms[k];
        rs[k][2] = rs[k][2] + rel[2] * mult * ms[k];
      }
    T av[3] = {(f0[0] + r0[0]) / m0, (f0[1] + r0[1]) / m0, (f0[2] + r0[2]) / m0};
    for (int k = 0; k < nv; ++k)
      if (!stick[k]) {
        T rv[3];
        for (int r = 0; r < 3; ++r) rv[r] = fs[k][r] / ms[k] - av[r];
        T t2 = sqrt(rv[1] * rv[1] + rv[2] * rv[2]);
        for (int r = 0; r < 3; ++r) rv[r] = rv[r] / t2;
        T mm = T(0.0) - T(mu) * rs[k][0] * ms[k] / (ms[k] + mst);
        rs[k][1] = mm * rv[1];
        rs[k][2] = mm * rv[2];
      }
  }
  T r0w[3], bf0[3];
  toworld(r0, r0w);
  toworld(f0, bf0);
  for (int q = 0; q < 3; ++q)
    for (int r = 0; r < 3; ++r) out[3 * q + r] = (bf0[r] + r0w[r]) / m0 * m[q] - F[3 * q + r];
  for (int k = 0; k < nv; ++k) toworld(rs[k].data(), &out[3 * (3 + k)]);
}

// Vector-Jacobian product of group_apply: Fb, mb += J^T ob.
void group_vjp(const double* B, int nv, const double* F, const double* m, const char* stick, double mu,
               const double* ob, double* Fb, double* mb) {
  const int nf = 3 * (3 + nv), nm = 3 + nv;
  std::vector<Dual> Fd(nf), md(nm), od(nf);
  for (int i = 0; i < nf + nm; ++i) {
    for (int k = 0; k < nf; ++k) Fd[k] = Dual(F[k], k == i ? 1.0 : 0.0);
    for (int k = 0; k < nm; ++k) md[k] = Dual(m[k], nf + k == i ? 1.0 : 0.0);
    group_apply<Dual>(B, nv, Fd.data(), md.data(), stick, mu, od.data());
    double g = 0.0;
    for (int o = 0; o < nf; ++o) g += ob[o] * od[o].d;
    if (i < nf)
      Fb[i] += g;
    else
      mb[i - nf] += g;
  }
}

// Reverse sweep of the taped contact iterations. In: ybar = adjoint of the
// step's final positions. Out: ybar = adjoint of the PD result y; xbar, zsum
// and the parameter gradients accumulate.
void contact_backward(Scene& s, const StepTape& tp, std::vector<double>& ybar, std::vector<double>& xbar,
                      std::vector<double>& zsum, double* gE, double* gnu, double* gact) {
  int N = s.N, n3 = 3 * N;
  const double h = s.h, al = s.alpha;
  const double* xn = tp.xn.data();
  const double* act = tp.act.empty() ? nullptr : tp.act.data();
  int J = (int)tp.citers.size();
  std::vector<double> gfin = ybar, XB(n3, 0.0), PB(n3, 0.0), rcb(n3), z(n3), coefD(N), ob(n3);
  // adjoints of the contact-phase positions P_k read by sphere detections
  std::vector<std::vector<double>> pend(J + 1);
  // with self contact: adjoint of the self-contact forces entering the
  // iteration being reversed (the previous handling's, carried over)
  std::vector<double> SCB(n3, 0.0);
  if (tp.final_latest) XB = gfin;
  for (int j = J - 1; j >= 0; --j) {
    const ContactIter& it = tp.citers[j];
    if (!pend[j + 1].empty())
      for (int k = 0; k < n3; ++k) XB[k] += pend[j + 1][k];
    for (int k = 0; k < n3; ++k) xbar[k] += XB[k];
    for (int v = 0; v < N; ++v) {
      double cd = 0;
      for (int c = 0; c < 3; ++c) {
        PB[3 * v + c] += XB[3 * v + c] / (s.Kdiag[v] * h);
        cd -= XB[3 * v + c] * (it.xit[3 * v + c] - xn[3 * v + c]) / s.Kdiag[v];
      }
      coefD[v] = cd;
    }
    for (int k = 0; k < n3; ++k) rcb[k] = al * PB[k];
    if (tp.self_on) {
      // rhs = forces + obstacle reactions + self-contact forces while a group
      // acts (World.h:1585-1588), else the obstacle reaction replaces the
      // forces; the obstacle projection reads forces + the carried self-contact
      // forces; group membership, stick flags and bases frozen as executed
      std::vector<double> Q = rcb, cnb, scb_new(n3, 0.0);
      if (it.grp) {
        cnb = Q;
        std::vector<double> Fb(n3, 0.0);
        for (const ContactIter::Group& g : it.groups) {
          const int nv = (int)g.vl.size();
          std::vector<int> vs(g.fv, g.fv + 3);
          vs.insert(vs.end(), g.vl.begin(), g.vl.end());
          std::vector<double> Fl(3 * vs.size()), ml(vs.size()), ob(3 * vs.size()), Fbl(3 * vs.size(), 0.0),
              mbl(vs.size(), 0.0);
          for (size_t a = 0; a < vs.size(); ++a) {
            ml[a] = s.Kdiag[vs[a]];
            for (int c = 0; c < 3; ++c) {
              Fl[3 * a + c] = it.F[3 * vs[a] + c];
              ob[3 * a + c] = Q[3 * vs[a] + c] + SCB[3 * vs[a] + c];
            }
          }
          group_vjp(g.B, nv, Fl.data(), ml.data(), g.stick.data(), s.self_mu, ob.data(), Fbl.data(), mbl.data());
          for (size_t a = 0; a < vs.size(); ++a) {
            coefD[vs[a]] += mbl[a];
            for (int c = 0; c < 3; ++c) Fb[3 * vs[a] + c] += Fbl[3 * a + c];
          }
        }
        for (int k = 0; k < n3; ++k) {
          cnb[k] += Fb[k];
          rcb[k] += Fb[k];
        }
      }
      for (size_t q = 0; q < it.cv.size(); ++q) {
        const int v = it.cv[q];
        const double* B = &it.B[9 * q];
        const double* fin = &it.f[3 * q];
        if (B[0] * fin[0] + B[3] * fin[1] + B[6] * fin[2] > 0.0) continue;  // take-off: no reaction
        const double* qb = it.grp ? &cnb[3 * v] : &Q[3 * v];
        double o[3];
        project_contact_vjp(B, s.mu, fin, qb, o);
        for (int c = 0; c < 3; ++c) {
          rcb[3 * v + c] = it.grp ? rcb[3 * v + c] + o[c] : o[c];
          scb_new[3 * v + c] = o[c];
        }
        if (it.sph[q] >= 0) {
          double n[3] = {B[0], B[3], B[6]}, nb[3];
          project_contact_nvjp(n, s.mu, fin, qb, nb);
          if (nb[0] != 0.0 || nb[1] != 0.0 || nb[2] != 0.0) {
            const int k = it.det[q];
            const std::vector<double>& Pk = k < J ? tp.citers[k].ori : tp.citers[J - 1].xit;
            if (pend[k].empty()) pend[k].assign(n3, 0.0);
            sphere_normal_vjp(s.spheres[it.sph[q]].c, &xn[3 * v], &Pk[3 * v], it.st[q], it.sroot[q] != 0, n, nb,
                              &xbar[3 * v], &pend[k][3 * v]);
          }
        }
      }
      SCB.swap(scb_new);
    }
    for (size_t q = 0; q < (tp.self_on ? 0 : it.cv.size()); ++q) {
      int v = it.cv[q];
      double qb[3] = {al * PB[3 * v], al * PB[3 * v + 1], al * PB[3 * v + 2]}, o[3];
      project_contact_vjp(&it.B[9 * q], s.mu, &it.f[3 * q], qb, o);
      for (int c = 0; c < 3; ++c) rcb[3 * v + c] = o[c];
      if (it.sph[q] >= 0) {  // the normal moves with the positions it was detected on
        const double* B = &it.B[9 * q];
        double n[3] = {B[0], B[3], B[6]}, nb[3];
        project_contact_nvjp(n, s.mu, &it.f[3 * q], qb, nb);
        if (nb[0] != 0.0 || nb[1] != 0.0 || nb[2] != 0.0) {
          const int k = it.det[q];
          const std::vector<double>& Pk = k < J ? tp.citers[k].ori : tp.citers[J - 1].xit;
          if (pend[k].empty()) pend[k].assign(n3, 0.0);
          sphere_normal_vjp(s.spheres[it.sph[q]].c, &xn[3 * v], &Pk[3 * v], it.st[q], it.sroot[q] != 0, n, nb,
                            &xbar[3 * v], &pend[k][3 * v]);
        }
      }
    }
    if (it.first) {
      for (int k = 0; k < n3; ++k) {
        rcb[k] += (1.0 - al) * PB[k];
        PB[k] = 0.0;
      }
    } else {
      for (int k = 0; k < n3; ++k) PB[k] *= (1.0 - al);
    }
    for (int k = 0; k < n3; ++k) {
      z[k] = h * rcb[k];
      zsum[k] += z[k];
    }
    add_Cy(s, z.data(), 1.0, xbar.data());
    std::fill(ob.begin(), ob.end(), 0.0);
    add_Cy(s, z.data(), -1.0, ob.data());
    const double* ori = it.ori.data();
    for (size_t t = 0; t < s.tets.size(); ++t) {
      const Tet& te = s.tets[t];
      double F[9], Fz[9], md[9], mdv[9], Fbar[9];
      tet_F(te, ori, F);
      tet_F(te, z.data(), Fz);
      corotate_project(F, md, mdv);
      double extra = 0;
      for (int a = 0; a < 4; ++a) {
        int v = te.v[a];
        double g2 = te.G[a][0] * te.G[a][0] + te.G[a][1] * te.G[a][1] + te.G[a][2] * te.G[a][2];
        double zd = 0;
        for (int c = 0; c < 3; ++c) zd += z[3 * v + c] * (ori[3 * v + c] - xn[3 * v + c]);
        extra += g2 * (zd + coefD[v]);
      }
      double pr = extra, pv = extra;
      for (int i = 0; i < 9; ++i) {
        pr += Fz[i] * (md[i] - F[i]);
        pv += Fz[i] * (mdv[i] - F[i]);
      }
      double opn = 1.0 + s.nu_fwd_t[t], vol = te.vol;
      if (gE) gE[t] += pr * vol / opn + pv * vol * s.nu_ctor[t] / opn;
      if (gnu) *gnu += -(pr * s.wr[t] + pv * s.wv[t]) / opn;
      corotate_vjp(F, s.wr[t], s.wv[t], Fz, Fbar);
      for (int k = 0; k < 4; ++k)
        for (int r = 0; r < 3; ++r)
          ob[3 * te.v[k] + r] += Fbar[r * 3 + 0] * te.G[k][0] + Fbar[r * 3 + 1] * te.G[k][1] + Fbar[r * 3 + 2] * te.G[k][2];
    }
    for (const Muscle& m : s.muscles) {
      double a = act ? act[m.group] : 0.0;
      double Ff[3] = {0, 0, 0}, Fz[3] = {0, 0, 0};
      for (int k = 0; k < 4; ++k)
        for (int r = 0; r < 3; ++r) {
          Ff[r] += ori[3 * m.v[k] + r] * m.g[k];
          Fz[r] += z[3 * m.v[k] + r] * m.g[k];
        }
      if (gact) gact[m.group] += -m.w * (Fz[0] * Ff[0] + Fz[1] * Ff[1] + Fz[2] * Ff[2]);
      double sc = m.w * (1.0 - a);
      for (int k = 0; k < 4; ++k)
        for (int r = 0; r < 3; ++r) ob[3 * m.v[k] + r] += sc * Fz[r] * m.g[k];
    }
    XB = ob;
    if (j == J - 1 && !tp.final_latest)
      for (int k = 0; k < n3; ++k) XB[k] += gfin[k];
  }
  if (!pend[0].empty())
    for (int k = 0; k < n3; ++k) XB[k] += pend[0][k];
  ybar = XB;
}

// Reverse sweep of one step (contact iterations first, then the PD loop). Inputs: gx, gv = dL/dx_{n+1}, dL/dv_{n+1}.
// Outputs (accumulated): gx_n, gv_n, gE (per tet), gnu, gact (per group).
void step_backward(Scene& s, const StepTape& tp, const double* gx, const double* gv, double* gxn, double* gvn,
                   double* gE, double* gnu, double* gact) {
  int N = s.N, n3 = 3 * N;
  double gam = s.damping, h = s.h;
  int K = (int)tp.ys.size() - 1;
  std::vector<double> ybar(n3), xbar(n3, 0.0), zsum(n3, 0.0), z(n3);
  for (int k = 0; k < n3; ++k) {
    ybar[k] = gx[k] + gam / h * gv[k];
    xbar[k] = -gam / h * gv[k];
  }
  if (!tp.citers.empty()) contact_backward(s, tp, ybar, xbar, zsum, gE, gnu, gact);
  const double* act = tp.act.empty() ? nullptr : tp.act.data();
  for (int it = K - 1; it >= 0; --it) {
    const double* yk = tp.ys[it].data();
    const double* yk1 = tp.ys[it + 1].data();
    z = ybar;
    s.sky.solve3(z.data());
    for (int k = 0; k < n3; ++k) {
      xbar[k] += ybar[k];
      zsum[k] += z[k];
    }
    std::fill(ybar.begin(), ybar.end(), 0.0);
    for (size_t t = 0; t < s.tets.size(); ++t) {
      const Tet& te = s.tets[t];
      double F[9], Fz[9], F1[9], md[9], mdv[9], Fbar[9];
      tet_F(te, yk, F);
      tet_F(te, z.data(), Fz);
      tet_F(te, yk1, F1);
      corotate_project(F, md, mdv);
      // parameter gradient: (A z).(d - A y_{k+1}) per weight
      double gr = 0, gvv = 0;
      for (int i = 0; i < 9; ++i) {
        gr += Fz[i] * (md[i] - F1[i]);
        gvv += Fz[i] * (mdv[i] - F1[i]);
      }
      double opn = 1.0 + s.nu_fwd_t[t];
      double vol = te.vol;
      // wr = E vol/(1+nu), wv = E vol nu_c/(1+nu)
      if (gE) gE[t] += gr * vol / opn + gvv * vol * s.nu_ctor[t] / opn;
      if (gnu) *gnu += -(gr * s.wr[t] + gvv * s.wv[t]) / opn;
      corotate_vjp(F, s.wr[t], s.wv[t], Fz, Fbar);
      for (int k = 0; k < 4; ++k)
        for (int r = 0; r < 3; ++r)
          ybar[3 * te.v[k] + r] += Fbar[r * 3 + 0] * te.G[k][0] + Fbar[r * 3 + 1] * te.G[k][1] + Fbar[r * 3 + 2] * te.G[k][2];
    }
    for (const Muscle& m : s.muscles) {
      double a = act ? act[m.group] : 0.0;
      double Ff[3] = {0, 0, 0}, Fz[3] = {0, 0, 0};
      for (int k = 0; k < 4; ++k)
        for (int r = 0; r < 3; ++r) {
          Ff[r] += yk[3 * m.v[k] + r] * m.g[k];
          Fz[r] += z[3 * m.v[k] + r] * m.g[k];
        }
      if (gact) gact[m.group] += -m.w * (Fz[0] * Ff[0] + Fz[1] * Ff[1] + Fz[2] * Ff[2]);
      double sc = m.w * (1.0 - a);
      for (int k = 0; k < 4; ++k)
        for (int r = 0; r < 3; ++r) ybar[3 * m.v[k] + r] += sc * Fz[r] * m.g[k];
    }
  }
  // y_0 = xn + h v*  ;  c = M v*/h - L_full xn
  std::vector<double> vsb(n3);
  for (int k = 0; k < n3; ++k) {
    xbar[k] += ybar[k];
    vsb[k] = h * ybar[k];
  }
  for (int i = 0; i < N; ++i)
    for (int r = 0; r < 3; ++r) vsb[3 * i + r] += s.mass[i] / h * zsum[3 * i + r];
  add_Lx(s, zsum.data(), -1.0, xbar.data());
  // c depends on the weights through -L xn; that term is already folded into
  // the per-iteration (A z).(d - A y_{k+1}) expression above.
  for (int k = 0; k < n3; ++k) {
    gxn[k] += xbar[k];
    gvn[k] += vsb[k];
  }
  if (s.underwater) drag_backward(s, tp.xn.data(), tp.vn.data(), vsb.data(), gxn, gvn);
}

std::vector<int> surface_vertices(int N, const std::vector<Tet>& tets) {
  std::map<std::array<int, 3>, int> cnt;
  for (const Tet& t : tets) {
    int sv[4] = {t.v[0], t.v[1], t.v[2], t.v[3]};
    std::sort(sv, sv + 4);
    const int fc[4][3] = {{0, 1, 2}, {0, 1, 3}, {0, 2, 3}, {1, 2, 3}};
    for (auto& f : fc) cnt[{sv[f[0]], sv[f[1]], sv[f[2]]}]++;
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

// ---- Muscle::Initialize per-segment tet selection (Muscle.h:109-215)
inline double signed_vol(const double* p0, const double* p1, const double* p2, const double* p3) {
  double a[3], b[3], c[3];
  for (int r = 0; r < 3; ++r) {
    a[r] = p1[r] - p0[r];
    b[r] = p2[r] - p0[r];
    c[r] = p3[r] - p0[r];
  }
  double cx = a[1] * b[2] - a[2] * b[1], cy = a[2] * b[0] - a[0] * b[2], cz = a[0] * b[1] - a[1] * b[0];
  return (cx * c[0] + cy * c[1] + cz * c[2]) / 6.0;
}
bool line_tri(const double* p0, const double* p1, const double* p2, const double* s, const double* e) {
  if (signed_vol(s, p0, p1, p2) * signed_vol(e, p0, p1, p2) < 0.0) {
    double a = signed_vol(s, e, p0, p1), b = signed_vol(s, e, p1, p2), c = signed_vol(s, e, p2, p0);
    if (a > 0 && b > 0 && c > 0) return true;
    if (a < 0 && b < 0 && c < 0) return true;
  }
  return false;
}
double pos_det(const double* q, const double* p1, const double* p2, const double* p3) {
  // GetDm (Muscle.h:273-290): determinant made non-negative by swapping columns
  double Dm[9];
  for (int r = 0; r < 3; ++r) {
    Dm[r * 3 + 0] = p1[r] - q[r];
    Dm[r * 3 + 1] = p2[r] - q[r];
    Dm[r * 3 + 2] = p3[r] - q[r];
  }
  double d = det3(Dm);
  return d < 0 ? -d : d;
}

}  // namespace

extern "C" {

const char* orc_last_error() { return g_err.c_str(); }

void* orc_scene_new(int N, const double* xrest, int T, const int* tets, int mass_same, double total_mass) {
  Scene* s = new Scene();
  s->N = N;
  s->xrest.assign(xrest, xrest + 3 * (size_t)N);
  s->tets.resize(T);
  std::vector<double> m(N, 0.0);
  double totvol = 0;
  for (int t = 0; t < T; ++t) {
    build_tet(xrest, tets + 4 * t, s->tets[t]);
    double vol = s->tets[t].vol;
    totvol += vol;
    for (int k = 0; k < 4; ++k) m[tets[4 * t + k]] += 0.25 * vol;  // Dfobj.h:506-512
  }
  for (int i = 0; i < N; ++i) m[i] = mass_same ? total_mass / totvol * m[i] : total_mass / N;  // :515-525
  s->mass = m;
  s->E.assign(T, 2e5);
  s->nu_ctor.assign(T, 0.4);
  s->nu_fwd_t.assign(T, 0.4);
  s->contact_vtx = surface_vertices(N, s->tets);
  return s;
}

// Several bodies in one world (one Dfobj per mesh, Environment.h:109-210):
// each body's "Mass" is lumped on its own tets (Dfobj.h:506-525).
int orc_set_bodies(void* p, int nb, const int* body_of_tet, const double* body_mass, int mass_same) {
  Scene* s = (Scene*)p;
  int N = s->N;
  size_t T = s->tets.size();
  std::vector<double> vol_b(nb, 0.0), m(N, 0.0);
  std::vector<int> vbody(N, -1), nv_b(nb, 0);
  for (size_t t = 0; t < T; ++t) {
    int b = body_of_tet[t];
    if (b < 0 || b >= nb) {
      g_err = "body index out of range";
      return 1;
    }
    vol_b[b] += s->tets[t].vol;
    for (int k = 0; k < 4; ++k) vbody[s->tets[t].v[k]] = b;
  }
  for (int i = 0; i < N; ++i)
    if (vbody[i] >= 0) nv_b[vbody[i]]++;
  for (size_t t = 0; t < T; ++t) {
    int b = body_of_tet[t];
    for (int k = 0; k < 4; ++k) m[s->tets[t].v[k]] += 0.25 * s->tets[t].vol * body_mass[b] / vol_b[b];
  }
  if (!mass_same)
    for (int i = 0; i < N; ++i) m[i] = vbody[i] >= 0 ? body_mass[vbody[i]] / nv_b[vbody[i]] : 0.0;
  s->mass = m;
  s->factored = false;
  return 0;
}

// Per-tet nu entering mMu (bodies built with their own Poisson ratio).
void orc_set_nu_fwd_tet(void* p, const double* nu_t) {
  Scene* s = (Scene*)p;
  s->nu_fwd_t.assign(nu_t, nu_t + s->tets.size());
  s->factored = false;
}

void orc_scene_free(void* p) { delete (Scene*)p; }

int orc_num_tets(void* p) { return (int)((Scene*)p)->tets.size(); }

void orc_get_masses(void* p, double* m) {
  Scene* s = (Scene*)p;
  std::copy(s->mass.begin(), s->mass.end(), m);
}

void orc_get_tets(void* p, int* v, double* vol, double* G) {
  Scene* s = (Scene*)p;
  for (size_t t = 0; t < s->tets.size(); ++t) {
    for (int k = 0; k < 4; ++k) v[4 * t + k] = s->tets[t].v[k];
    vol[t] = s->tets[t].vol;
    for (int k = 0; k < 4; ++k)
      for (int j = 0; j < 3; ++j) G[12 * t + 3 * k + j] = s->tets[t].G[k][j];
  }
}

int orc_num_contact_vertices(void* p) { return (int)((Scene*)p)->contact_vtx.size(); }
void orc_get_contact_vertices(void* p, int* out) {
  Scene* s = (Scene*)p;
  std::copy(s->contact_vtx.begin(), s->contact_vtx.end(), out);
}

void orc_set_material(void* p, const double* E_tet, double nu_fwd, const double* nu_ctor_tet) {
  Scene* s = (Scene*)p;
  size_t T = s->tets.size();
  s->E.assign(E_tet, E_tet + T);
  s->nu_fwd_t.assign(T, nu_fwd);
  s->nu_ctor.assign(nu_ctor_tet, nu_ctor_tet + T);
  s->factored = false;
}

void orc_set_attachments(void* p, int n, const int* idx, double w) {
  Scene* s = (Scene*)p;
  s->attach_idx.assign(idx, idx + n);
  s->attach_w.assign(n, w);
  s->factored = false;
}

// Explicit muscle constraints: indices (4 per constraint, orientation is fixed
// here as Muscle.h:143-154 does), fibre direction, weight, group.
void orc_set_muscles(void* p, int n, const int* idx4, const double* fiber, const double* weight, double k_muscle,
                     const int* group, int n_groups) {
  Scene* s = (Scene*)p;
  s->muscles.clear();
  s->n_groups = n_groups;
  for (int c = 0; c < n; ++c) {
    Tet te;
    build_tet(s->xrest.data(), idx4 + 4 * c, te);
    Muscle m;
    for (int k = 0; k < 4; ++k) m.v[k] = te.v[k];
    // v = invDm f ; A = [-(a+b+c), a, b, c] (LinearMuscleConstraint.h:156-166)
    const double* f = fiber + 3 * c;
    double vv[3];
    // invDm rows are G[1..3]: invDm(m, j) = G[m+1][j]
    for (int mm = 0; mm < 3; ++mm) vv[mm] = te.G[mm + 1][0] * f[0] + te.G[mm + 1][1] * f[1] + te.G[mm + 1][2] * f[2];
    m.g[0] = -(vv[0] + vv[1] + vv[2]);
    m.g[1] = vv[0];
    m.g[2] = vv[1];
    m.g[3] = vv[2];
    m.w = te.vol * k_muscle * weight[c];  // mStiffness *= mWeight (:109); J = vol*k*A^T (:168)
    m.group = group[c];
    s->muscles.push_back(m);
  }
  s->factored = false;
}

// Per-segment tet selection of Muscle::Initialize for one segment [start,end].
// Writes up to cap constraints (tet index, weight); returns the count.
int orc_muscle_segment_select(void* p, const int* tets, int T, const double* start, const double* end, int cap,
                              int* out_tet, double* out_w) {
  Scene* s = (Scene*)p;
  const double* X = s->xrest.data();
  int n = 0;
  double dir[3] = {end[0] - start[0], end[1] - start[1], end[2] - start[2]};
  double len = std::sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
  for (int r = 0; r < 3; ++r) dir[r] /= len;
  auto push = [&](int t, double w) {
    if (n < cap) {
      out_tet[n] = t;
      out_w[n] = w;
    }
    ++n;
  };
  for (int j = 0; j < T; ++j) {
    int i0 = tets[4 * j], i1 = tets[4 * j + 1], i2 = tets[4 * j + 2], i3 = tets[4 * j + 3];
    double Dm[9];
    for (int r = 0; r < 3; ++r) {
      Dm[r * 3 + 0] = X[3 * i1 + r] - X[3 * i0 + r];
      Dm[r * 3 + 1] = X[3 * i2 + r] - X[3 * i0 + r];
      Dm[r * 3 + 2] = X[3 * i3 + r] - X[3 * i0 + r];
    }
    if (det3(Dm) < 0) std::swap(i1, i2);
    const double *p0 = &X[3 * i0], *p1 = &X[3 * i1], *p2 = &X[3 * i2], *p3 = &X[3 * i3];
    double center[3], v[3];
    for (int r = 0; r < 3; ++r) {
      center[r] = (p0[r] + p1[r] + p2[r] + p3[r]) / 4.0;
      v[r] = center[r] - start[r];
    }
    double t = v[0] * dir[0] + v[1] * dir[1] + v[2] * dir[2];
    double pp[3];
    for (int r = 0; r < 3; ++r) pp[r] = start[r] + t * dir[r];
    auto dist3 = [](const double* a, const double* b) {
      return std::sqrt((a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]) + (a[2] - b[2]) * (a[2] - b[2]));
    };
    double dist;
    if (std::max(dist3(pp, start), dist3(pp, end)) / dist3(end, start) < 1.0)
      dist = dist3(pp, center);
    else
      dist = 1e5;
    dist = std::min(dist, std::min(dist3(center, start), dist3(center, end)));
    double weight = std::exp(-std::fabs(dist) / 0.005);
    bool inside = line_tri(p0, p1, p2, start, end) || line_tri(p0, p1, p3, start, end) ||
                  line_tri(p0, p2, p3, start, end) || line_tri(p1, p2, p3, start, end);
    if (inside)
      push(j, weight);
    else if (weight > 0.2)
      push(j, weight);
    double D = det3(Dm) < 0 ? -det3(Dm) : det3(Dm);
    double S = (pos_det(start, p1, p2, p3) + pos_det(start, p2, p3, p0) + pos_det(start, p3, p0, p1) +
                pos_det(start, p0, p1, p2)) / D;
    if (std::fabs(S - 1.0) < 1e-4) push(j, weight);
    double Ee = (pos_det(end, p1, p2, p3) + pos_det(end, p2, p3, p0) + pos_det(end, p3, p0, p1) +
                 pos_det(end, p0, p1, p2)) / D;
    if (std::fabs(Ee - 1.0) < 1e-4) push(j, weight);
  }
  return n;
}

void orc_set_tris(void* p, int n, const double* tris) {
  Scene* s = (Scene*)p;
  s->tris.resize(n);
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < 9; ++k) s->tris[i][k] = tris[9 * i + k];
}
int orc_detect_tri(const double* tri, const double* x0, const double* x1, double* n) {
  return detect_tri(tri, x0, x1, n) ? 1 : 0;
}
// Self contact: builds the surface mesh in the reference's contact numbering
// (Dfobj.h:497-566, World.h:540-558).
void orc_set_self_collision(void* p, int on, double mu) {
  Scene* s = (Scene*)p;
  s->self_collision = on;
  s->self_mu = mu;
  s->sc_vtx.clear();
  s->sc_face.clear();
  if (!on) return;
  std::vector<std::array<int, 3>> tri;
  for (const Tet& t : s->tets) {
    int sv[4] = {t.v[0], t.v[1], t.v[2], t.v[3]};
    std::sort(sv, sv + 4);
    tri.push_back({sv[0], sv[1], sv[2]});
    tri.push_back({sv[0], sv[1], sv[3]});
    tri.push_back({sv[0], sv[2], sv[3]});
    tri.push_back({sv[1], sv[2], sv[3]});
  }
  std::map<std::array<int, 3>, int> cnt;
  for (auto& f : tri) cnt[f]++;
  std::map<int, int> cidx;
  for (auto& f : tri) {
    if (cnt[f] != 1) continue;
    std::array<int, 3> cf;
    for (int k = 0; k < 3; ++k) {
      auto it = cidx.find(f[k]);
      if (it == cidx.end()) {
        it = cidx.emplace(f[k], (int)s->sc_vtx.size()).first;
        s->sc_vtx.push_back(f[k]);
      }
      cf[k] = it->second;
    }
    s->sc_face.push_back(cf);
  }
}
int orc_cubic_roots01(double a, double b, double c, double d, double* x) { return cubic_roots01(a, b, c, d, x); }
int orc_vf_moving(const double* p0, const double* p1, const double* y0, const double* y1, double* t, double* d,
                  double* n) {
  const double* a0[3] = {y0, y0 + 3, y0 + 6};
  const double* a1[3] = {y1, y1 + 3, y1 + 6};
  SelfImpact im;
  if (!vf_moving(p0, p1, a0, a1, im)) return 0;
  *t = im.t;
  *d = im.d;
  for (int r = 0; r < 3; ++r) n[r] = im.n[r];
  return 1;
}
// One self-contact group's response (World.h:1451-1530): face forces f3 (3x3)
// and masses m3, members' forces fv (nv x 3) and masses mv, normal n; outputs
// the face corners' and the members' self-contact forces (world frame).
void orc_self_group(const double* n, const double* f3, const double* m3, int nv, const double* fv, const double* mv,
                    double mu, double* out3, double* outv) {
  double B[9];
  contact_basis(n, B);
  auto toloc = [&](const double* w, double* l) {
    for (int c = 0; c < 3; ++c) l[c] = B[0 * 3 + c] * w[0] + B[1 * 3 + c] * w[1] + B[2 * 3 + c] * w[2];
  };
  auto toworld = [&](const double* l, double* w) {
    for (int r = 0; r < 3; ++r) w[r] = B[r * 3 + 0] * l[0] + B[r * 3 + 1] * l[1] + B[r * 3 + 2] * l[2];
  };
  double m0 = m3[0] + m3[1] + m3[2], f0w[3] = {0, 0, 0}, f0[3], r0[3] = {0, 0, 0};
  for (int q = 0; q < 3; ++q)
    for (int r = 0; r < 3; ++r) f0w[r] += f3[3 * q + r];
  toloc(f0w, f0);
  std::vector<int> vl;
  for (int k = 0; k < nv; ++k) {
    double fl[3];
    toloc(fv + 3 * k, fl);
    if (fl[0] / mv[k] - f0[0] / m0 <= 0) vl.push_back(k);
  }
  int m = (int)vl.size();
  std::vector<double> ms(m);
  std::vector<std::array<double, 3>> fs(m), rs(m);
  double sm = 0, smv = 0;
  for (int k = 0; k < m; ++k) {
    ms[k] = mv[vl[k]];
    sm += ms[k];
    toloc(fv + 3 * vl[k], fs[k].data());
    smv += fs[k][0];
  }
  for (int k = 0; k < m; ++k) {
    double fn = (f0[0] + smv) * ms[k] / (m0 + sm) - fs[k][0];
    rs[k] = {fn, 0.0, 0.0};
    r0[0] -= fn;
  }
  std::vector<char> stick(m, 1);
  while (true) {
    solve_group_friction(f0, m0, r0, fs, ms, rs, stick, mu);
    bool changed = false;
    for (int k = 0; k < m; ++k)
      if (stick[k] && rs[k][1] * rs[k][1] + rs[k][2] * rs[k][2] > mu * rs[k][0] * rs[k][0]) {
        changed = true;
        stick[k] = 0;
      }
    if (!changed) break;
  }
  double r0w[3], bf0[3], v0[3];
  toworld(r0, r0w);
  toworld(f0, bf0);
  for (int r = 0; r < 3; ++r) v0[r] = (bf0[r] + r0w[r]) / m0;
  for (int q = 0; q < 3; ++q)
    for (int r = 0; r < 3; ++r) out3[3 * q + r] = v0[r] * m3[q] - f3[3 * q + r];
  std::fill(outv, outv + 3 * nv, 0.0);
  for (int k = 0; k < m; ++k) toworld(rs[k].data(), outv + 3 * vl[k]);
}
void orc_set_underwater(void* p, int on) {
  Scene* s = (Scene*)p;
  s->underwater = on;
  s->contours = on ? build_contours(*s) : std::vector<std::array<int, 3>>();
}
int orc_num_drag_faces(void* p) { return (int)((Scene*)p)->contours.size(); }
void orc_get_drag_faces(void* p, int* out) {
  Scene* s = (Scene*)p;
  for (size_t i = 0; i < s->contours.size(); ++i)
    for (int k = 0; k < 3; ++k) out[3 * i + k] = s->contours[i][k];
}
void orc_drag_face(const double* n, const double* w, double* F, double* J) {
  drag_face_t(n, w, F);
  if (J) {
    double Jm[3][6];
    drag_face_jac(n, w, Jm);
    for (int c = 0; c < 3; ++c)
      for (int k = 0; k < 6; ++k) J[6 * c + k] = Jm[c][k];
  }
}
void orc_drag_forces(void* p, const double* x, const double* v, double* f) { drag_forces(*(Scene*)p, x, v, f); }

void orc_set_sim(void* p, double h, int max_iter, double damping, const double* g, int gravity_on, int contact,
                 int max_iter1, double alpha, double mu) {
  Scene* s = (Scene*)p;
  s->h = h;
  s->max_iter = max_iter;
  s->damping = damping;
  for (int r = 0; r < 3; ++r) s->g[r] = g[r];
  s->gravity_on = gravity_on;
  s->contact = contact;
  s->max_iter1 = max_iter1;
  s->alpha = alpha;
  s->mu = mu;
  s->factored = false;
}

void orc_set_obstacles(void* p, int n_planes, const double* planes /*6 each*/, int n_spheres,
                       const double* spheres /*4 each*/) {
  Scene* s = (Scene*)p;
  s->planes.clear();
  s->spheres.clear();
  for (int k = 0; k < n_planes; ++k) {
    Plane pl;
    double nn = 0;
    for (int r = 0; r < 3; ++r) {
      pl.p[r] = planes[6 * k + r];
      pl.n[r] = planes[6 * k + 3 + r];
      nn += pl.n[r] * pl.n[r];
    }
    nn = std::sqrt(nn);
    for (int r = 0; r < 3; ++r) pl.n[r] /= nn;
    s->planes.push_back(pl);
  }
  for (int k = 0; k < n_spheres; ++k) {
    Sphere sp;
    for (int r = 0; r < 3; ++r) sp.c[r] = spheres[4 * k + r];
    sp.r = spheres[4 * k + 3];
    s->spheres.push_back(sp);
  }
}

int orc_factor(void* p) { return factorize(*(Scene*)p) ? 0 : 1; }
double orc_reg_used(void* p) { return ((Scene*)p)->reg_used; }

// Dense scalar K_s (N x N, row-major) for tests.
void orc_dense_K(void* p, double* out) {
  Scene* s = (Scene*)p;
  if (!s->factored) factorize(*s);
  std::fill(out, out + (size_t)s->N * s->N, 0.0);
  for (int i = 0; i < s->N; ++i)
    for (auto& kv : s->K[i]) out[(size_t)i * s->N + kv.first] = kv.second;
}

// Brute-force 3N x 3N H2ML = M/h^2 + L from the reference's per-constraint
// 9x12 / 3x12 A_i stencils (CorotateFEMConstraint.h:221-350,
// LinearMuscleConstraint.h:150-247, AttachmentConstraint.h:77-84), dense.
void orc_dense_H2ML_bruteforce(void* p, double* out) {
  Scene* s = (Scene*)p;
  update_weights(*s);
  int n3 = 3 * s->N;
  std::fill(out, out + (size_t)n3 * n3, 0.0);
  for (int i = 0; i < s->N; ++i)
    for (int c = 0; c < 3; ++c) out[(size_t)(3 * i + c) * n3 + 3 * i + c] += s->mass[i] / (s->h * s->h);
  for (size_t t = 0; t < s->tets.size(); ++t) {
    const Tet& te = s->tets[t];
    // invDm(m,j) = G[m+1][j]
    double d[3][3];
    for (int m = 0; m < 3; ++m)
      for (int j = 0; j < 3; ++j) d[m][j] = te.G[m + 1][j];
    double Ai[9][12];
    std::memset(Ai, 0, sizeof(Ai));
    for (int j = 0; j < 3; ++j)      // F column j
      for (int c = 0; c < 3; ++c) {  // component
        int row = 3 * j + c;
        Ai[row][0 + c] = -d[0][j] - d[1][j] - d[2][j];
        Ai[row][3 + c] = d[0][j];
        Ai[row][6 + c] = d[1][j];
        Ai[row][9 + c] = d[2][j];
      }
    for (int pass = 0; pass < 2; ++pass) {
      double w = pass == 0 ? s->wr[t] : s->wv[t];
      for (int a = 0; a < 12; ++a)
        for (int b = 0; b < 12; ++b) {
          double acc = 0;
          for (int r = 0; r < 9; ++r) acc += Ai[r][a] * Ai[r][b];
          int ga = 3 * te.v[a / 3] + a % 3, gb = 3 * te.v[b / 3] + b % 3;
          out[(size_t)ga * n3 + gb] += w * acc;
        }
    }
  }
  for (const Muscle& m : s->muscles) {
    double Ai[3][12];
    std::memset(Ai, 0, sizeof(Ai));
    for (int c = 0; c < 3; ++c)
      for (int k = 0; k < 4; ++k) Ai[c][3 * k + c] = m.g[k];
    for (int a = 0; a < 12; ++a)
      for (int b = 0; b < 12; ++b) {
        double acc = 0;
        for (int r = 0; r < 3; ++r) acc += Ai[r][a] * Ai[r][b];
        int ga = 3 * m.v[a / 3] + a % 3, gb = 3 * m.v[b / 3] + b % 3;
        out[(size_t)ga * n3 + gb] += m.w * acc;
      }
  }
  for (size_t k = 0; k < s->attach_idx.size(); ++k)
    for (int c = 0; c < 3; ++c) out[(size_t)(3 * s->attach_idx[k] + c) * n3 + 3 * s->attach_idx[k] + c] += s->attach_w[k];
}

// Solve K_s X = B for a 3-component AoS vector with the oracle's factor.
int orc_solve3(void* p, double* x) {
  Scene* s = (Scene*)p;
  if (!s->factored && !factorize(*s)) return 1;
  s->sky.solve3(x);
  return 0;
}

// J*d - L_full*x at positions x (the per-iteration force term) for tests.
void orc_force_terms(void* p, const double* x, const double* act, double* Jd, double* Lx) {
  Scene* s = (Scene*)p;
  update_weights(*s);
  std::fill(Jd, Jd + 3 * s->N, 0.0);
  std::fill(Lx, Lx + 3 * s->N, 0.0);
  add_Jd(*s, x, act, Jd);
  add_Lx(*s, x, 1.0, Lx);
}

// Rollout of `steps` steps. act: steps x n_groups (or null). iters: 3 ints per
// step (pd iterations, contact inner iterations, contact outer passes), may be null.
int orc_rollout(void* p, int steps, const double* x0, const double* v0, const double* act, double* xT, double* vT,
                int* iters, int record) {
  Scene* s = (Scene*)p;
  if (!s->factored && !factorize(*s)) return 1;
  int n3 = 3 * s->N;
  std::vector<double> x(x0, x0 + n3), v(v0, v0 + n3);
  if (record) s->tape.assign(steps, StepTape());
  else s->tape.clear();
  if (s->sn_mode == 1) s->sn_tab.clear();
  for (int n = 0; n < steps; ++n) {
    int a = 0, b = 0, c = 0;
    s->cur_step = n;
    step(*s, x.data(), v.data(), act ? act + (size_t)n * s->n_groups : nullptr, record ? &s->tape[n] : nullptr, &a,
         &b, &c);
    if (iters) {
      iters[3 * n] = a;
      iters[3 * n + 1] = b;
      iters[3 * n + 2] = c;
    }
  }
  std::copy(x.begin(), x.end(), xT);
  std::copy(v.begin(), v.end(), vT);
  return 0;
}

// Self-contact normals: 0 computed, 1 recorded by the next orc_rollout, 2
// replayed from the last recording (single rollouts only).
int orc_self_normals(void* p, int mode) {
  Scene* s = (Scene*)p;
  if (mode < 0 || mode > 2) {
    g_err = "mode must be 0, 1 or 2";
    return 1;
  }
  s->sn_mode = mode;
  return 0;
}

// Several independent rollouts (shared factor), OpenMP over rollouts when
// built with it. x0,v0,xT,vT: R x 3N; act: R x steps x n_groups (or null).
int orc_rollout_batch(void* p, int R, int steps, const double* x0, const double* v0, const double* act, double* xT,
                      double* vT, int nthreads) {
  Scene* s = (Scene*)p;
  if (s->sn_mode) {
    g_err = "self-contact normal recording/replay is for single rollouts";
    return 1;
  }
  if (!s->factored && !factorize(*s)) return 1;
  int n3 = 3 * s->N;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
  for (int r = 0; r < R; ++r) {
    std::vector<double> x(x0 + (size_t)r * n3, x0 + (size_t)(r + 1) * n3), v(v0 + (size_t)r * n3, v0 + (size_t)(r + 1) * n3);
    for (int n = 0; n < steps; ++n)
      step(*s, x.data(), v.data(), act ? act + ((size_t)r * steps + n) * s->n_groups : nullptr, nullptr, nullptr,
           nullptr, nullptr);
    std::copy(x.begin(), x.end(), xT + (size_t)r * n3);
    std::copy(v.begin(), v.end(), vT + (size_t)r * n3);
  }
  return 0;
}

// Forward + reverse sweep of R independent rollouts (shared factor), OpenMP
// over rollouts: the multi-core CPU baseline. Each rollout records its own
// tape, then seeds the reverse sweep with gxT/gvT (R x 3N) and writes
// gx0 (R x 3N); iters (R x steps, optional) receives the executed PD
// iterations. Per-tet E and activation gradients are accumulated per rollout
// and summed into gE (T) after the loop, in rollout order.
int orc_fwd_adj_batch(void* p, int R, int steps, const double* x0, const double* v0, const double* act,
                      const double* gxT, const double* gvT, double* gx0, double* gE, int* iters, int nthreads) {
  Scene* s = (Scene*)p;
  if (s->sn_mode) {
    g_err = "self-contact normal recording/replay is for single rollouts";
    return 1;
  }
  if (!s->factored && !factorize(*s)) return 1;
  const int n3 = 3 * s->N;
  const size_t T = s->tets.size();
  std::vector<std::vector<double>> gEr(R, std::vector<double>(gE ? T : 0, 0.0));
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
  for (int r = 0; r < R; ++r) {
    std::vector<double> x(x0 + (size_t)r * n3, x0 + (size_t)(r + 1) * n3), v(v0 + (size_t)r * n3, v0 + (size_t)(r + 1) * n3);
    std::vector<StepTape> tape(steps);
    for (int n = 0; n < steps; ++n) {
      int a = 0, b = 0, c = 0;
      step(*s, x.data(), v.data(), act ? act + ((size_t)r * steps + n) * s->n_groups : nullptr, &tape[n], &a, &b, &c);
      if (iters) iters[(size_t)r * steps + n] = a;
    }
    std::vector<double> gx(gxT + (size_t)r * n3, gxT + (size_t)(r + 1) * n3), gv(gvT + (size_t)r * n3, gvT + (size_t)(r + 1) * n3);
    std::vector<double> gact((size_t)std::max(s->n_groups, 1), 0.0);
    double gnu = 0.0;
    for (int n = steps - 1; n >= 0; --n) {
      std::vector<double> gxn(n3, 0.0), gvn(n3, 0.0);
      step_backward(*s, tape[n], gx.data(), gv.data(), gxn.data(), gvn.data(), gE ? gEr[r].data() : nullptr, &gnu,
                    gact.data());
      gx.swap(gxn);
      gv.swap(gvn);
    }
    std::copy(gx.begin(), gx.end(), gx0 + (size_t)r * n3);
  }
  if (gE) {
    std::fill(gE, gE + T, 0.0);
    for (int r = 0; r < R; ++r)
      for (size_t t = 0; t < T; ++t) gE[t] += gEr[r][t];
  }
  return 0;
}

int orc_max_threads() {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

// Gradient of <gx, x_T> + <gv, v_T> w.r.t. x0, v0, per-tet E, nu (the
// forward value entering mMu) and activations (steps x n_groups), over the tape
// of the last recorded rollout; contact iterations are reversed with their
// branches (take-off / stick / slide) frozen and sphere normals differentiated
// through the positions they were detected on.
int orc_backward(void* p, const double* gxT, const double* gvT, double* gx0, double* gv0, double* gE, double* gnu,
                 double* gact) {
  Scene* s = (Scene*)p;
  int n3 = 3 * s->N;
  int steps = (int)s->tape.size();
  std::vector<double> gx(gxT, gxT + n3), gv(gvT, gvT + n3);
  size_t T = s->tets.size();
  if (gE) std::fill(gE, gE + T, 0.0);
  if (gnu) *gnu = 0.0;
  if (gact) std::fill(gact, gact + (size_t)steps * s->n_groups, 0.0);
  int rc = 0;
  for (int n = steps - 1; n >= 0; --n) {
    std::vector<double> gxn(n3, 0.0), gvn(n3, 0.0);
    step_backward(*s, s->tape[n], gx.data(), gv.data(), gxn.data(), gvn.data(), gE, gnu,
                  gact ? gact + (size_t)n * s->n_groups : nullptr);
    gx.swap(gxn);
    gv.swap(gvn);
  }
  std::copy(gx.begin(), gx.end(), gx0);
  std::copy(gv.begin(), gv.end(), gv0);
  return rc;
}

// Unit-level access to the local projection (for closed-form / numpy pins).
void orc_corotate_project(const double* F, double* md, double* mdv) { corotate_project(F, md, mdv); }
void orc_corotate_vjp(const double* F, double wr, double wv, const double* Z, double* Fbar) {
  corotate_vjp(F, wr, wv, Z, Fbar);
}
void orc_svd3(const double* F, double* U, double* S, double* V) { svd3(F, U, S, V); }
void orc_volume_projection(const double* S, double* D) { volume_projection(S, D); }
void orc_volume_projection_jac(const double* S, double* D, double* J, double* rdd) {
  double Jm[3][3];
  volume_projection_jac(S, D, Jm, rdd);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) J[i * 3 + j] = Jm[i][j];
}
int orc_detect_plane(const double* p, const double* n, const double* x0, const double* x1, double* nout) {
  Plane pl;
  double nn = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
  for (int r = 0; r < 3; ++r) {
    pl.p[r] = p[r];
    pl.n[r] = n[r] / nn;
  }
  return detect_plane(pl, x0, x1, nout) ? 1 : 0;
}
void orc_project_contact(const double* n, double mu, const double* f, double* out) {
  double B[9];
  contact_basis(n, B);
  project_contact(B, mu, f, out);
}

}  // extern "C"
