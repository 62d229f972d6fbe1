// Device math shared by the libpdm kernels: 3x3 SVD (Jacobi eigen-analysis of
// F^T F followed by Gram-Schmidt QR of F V), the corotated projection of
// CorotateFEMConstraint::EvaluateDVector and its vector-Jacobian product.
#pragma once
#include <hip/hip_runtime.h>

namespace pdm {

template <typename T>
struct Num;
// FP64: reciprocal and inverse square root from the hardware estimates
// (v_rcp_f64, v_rsq_f64) refined by two Newton steps, to about 1 ulp (not
// always correctly rounded; the plain path is held to 1e-9); IEEE square root.
template <>
struct Num<double> {
  static __device__ __forceinline__ double rcp(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
  }
  // y <- y + (y / 2) (1 - x y^2), twice
  static __device__ __forceinline__ double rsqrt(double x) {
    double y = __builtin_amdgcn_rsq(x);
    double e = fma(-x * y, y, 1.0);
    y = fma(0.5 * y, e, y);
    e = fma(-x * y, y, 1.0);
    return fma(0.5 * y, e, y);
  }
  // one Newton step: relative error about the square of the estimate's,
  // for quantities that only steer a rotation (its orthogonality comes from
  // the fully refined rsqrt)
  static __device__ __forceinline__ double rcp1(double x) {
    double r = __builtin_amdgcn_rcp(x);
    return fma(r, fma(-x, r, 1.0), r);
  }
  static __device__ __forceinline__ double rsqrt1(double x) {
    double y = __builtin_amdgcn_rsq(x);
    return fma(0.5 * y, fma(-x * y, y, 1.0), y);
  }
  static __device__ __forceinline__ double sqrt_(double x) { return sqrt(x); }
  static __device__ __forceinline__ double div(double a, double b) { return a * rcp(b); }
  static constexpr int kSweeps = 6;
  static constexpr double kTiny = 1e-300;
  static constexpr double kRel = 1e-15;
};
// FP32 (mixed path): the hardware v_rcp/v_sqrt/v_rsq approximations (about
// 1 ulp) instead of the multi-instruction IEEE sequences; the FP32 local step
// is held to 1e-5, far above the extra rounding.
template <>
struct Num<float> {
  static __device__ __forceinline__ float rsqrt(float x) { return __builtin_amdgcn_rsqf(x); }
  static __device__ __forceinline__ float sqrt_(float x) { return __builtin_amdgcn_sqrtf(x); }
  static __device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
  static __device__ __forceinline__ float div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
  static constexpr int kSweeps = 4;
  static constexpr float kTiny = 1e-30f;
  static constexpr float kRel = 1e-6f;
};

// Jacobi rotation (c, s) that zeroes a_pq of the symmetric 2x2 block
// [[a_pp, a_pq], [a_pq, a_qq]]: t = tan(theta) = sign(tau) / (|tau| + sqrt(1 + tau^2)),
// tau = (a_qq - a_pp) / (2 a_pq), multiplied through by |2 a_pq|:
// t = sign(tau) |e| / (|d| + sqrt(d^2 + e^2)), d = a_qq - a_pp, e = 2 a_pq,
// which needs one reciprocal and no branch; a_pq = 0 gives t = 0 (the tiny
// addend keeps d = e = 0 finite). FP64: t to about 1e-14 relative (one
// Newton step on the hardware estimates): the angle is then off by as much,
// which leaves an off-diagonal residue of that relative size for the next
// rotations to remove; c = 1/sqrt(1 + t^2) fully refined and s = t c keep
// the rotation orthogonal to rounding.
template <typename T>
__device__ __forceinline__ bool jacobi_cs(T app, T aqq, T apq, T* t_out, T* c_out, T* s_out) {
  if constexpr (sizeof(T) == 8) {
    T d = aqq - app, e = T(2) * apq;
    T ae = fabs(e);
    T num = d * e >= T(0) ? ae : -ae;
    T h2 = d * d + e * e;
    T hyp = h2 * Num<T>::rsqrt1(h2 + T(1e-300));
    T t = num * Num<T>::rcp1(fabs(d) + hyp + T(1e-300));
    T c = Num<T>::rsqrt(T(1) + t * t);
    *t_out = t;
    *c_out = c;
    *s_out = t * c;
    return true;
  } else {
    T d = aqq - app, e = T(2) * apq;
    T ae = fabsf(e);
    T num = d * e >= T(0) ? ae : -ae;
    T t = num * Num<T>::rcp(fabsf(d) + Num<T>::sqrt_(d * d + e * e) + T(1e-30f));
    T c = Num<T>::rsqrt(T(1) + t * t);
    *t_out = t;
    *c_out = c;
    *s_out = t * c;
    return true;
  }
}

// F = U diag(S) V^T, S >= 0 descending, U, V orthogonal (row-major 3x3).
template <typename T>
__device__ __forceinline__ void svd3(const T* F, T* U, T* S, T* V) {
  T a00 = F[0] * F[0] + F[3] * F[3] + F[6] * F[6];
  T a01 = F[0] * F[1] + F[3] * F[4] + F[6] * F[7];
  T a02 = F[0] * F[2] + F[3] * F[5] + F[6] * F[8];
  T a11 = F[1] * F[1] + F[4] * F[4] + F[7] * F[7];
  T a12 = F[1] * F[2] + F[4] * F[5] + F[7] * F[8];
  T a22 = F[2] * F[2] + F[5] * F[5] + F[8] * F[8];
  T v[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  // one cyclic Jacobi sweep over (0,1), (0,2), (1,2)
  auto sweep = [&]() {
    // (p,q) = (0,1)
    {
      T apq = a01;
      T t, c, s;
      if (jacobi_cs<T>(a00, a11, apq, &t, &c, &s)) {
        a00 -= t * apq;
        a11 += t * apq;
        a01 = 0;
        T r0 = a02, r1 = a12;
        a02 = c * r0 - s * r1;
        a12 = s * r0 + c * r1;
#pragma unroll
        for (int r = 0; r < 3; ++r) {
          T vp = v[r * 3 + 0], vq = v[r * 3 + 1];
          v[r * 3 + 0] = c * vp - s * vq;
          v[r * 3 + 1] = s * vp + c * vq;
        }
      }
    }
    // (0,2)
    {
      T apq = a02;
      T t, c, s;
      if (jacobi_cs<T>(a00, a22, apq, &t, &c, &s)) {
        a00 -= t * apq;
        a22 += t * apq;
        a02 = 0;
        T r0 = a01, r1 = a12;
        a01 = c * r0 - s * r1;
        a12 = s * r0 + c * r1;
#pragma unroll
        for (int r = 0; r < 3; ++r) {
          T vp = v[r * 3 + 0], vq = v[r * 3 + 2];
          v[r * 3 + 0] = c * vp - s * vq;
          v[r * 3 + 2] = s * vp + c * vq;
        }
      }
    }
    // (1,2)
    {
      T apq = a12;
      T t, c, s;
      if (jacobi_cs<T>(a11, a22, apq, &t, &c, &s)) {
        a11 -= t * apq;
        a22 += t * apq;
        a12 = 0;
        T r0 = a01, r1 = a02;
        a01 = c * r0 - s * r1;
        a02 = s * r0 + c * r1;
#pragma unroll
        for (int r = 0; r < 3; ++r) {
          T vp = v[r * 3 + 1], vq = v[r * 3 + 2];
          v[r * 3 + 1] = c * vp - s * vq;
          v[r * 3 + 2] = s * vp + c * vq;
        }
      }
    }
    };
  if constexpr (sizeof(T) == 8) {
    // three sweeps, then up to three more while the off-diagonal part of
    // F^T F is above 1e-15 of its trace (cyclic Jacobi converges
    // quadratically: one sweep from there reaches rounding level)
#pragma unroll
    for (int sw = 0; sw < 3; ++sw) sweep();
    for (int sw = 3; sw < Num<T>::kSweeps; ++sw) {
      T off = fabs(a01) + fabs(a02) + fabs(a12);
      if (!(off > T(1e-15) * (a00 + a11 + a22))) break;
      sweep();
    }
  } else {
    // FP32: three sweeps, a fourth only while the off-diagonal part of F^T F
    // is above 1e-7 of its trace (rare for the strains met here; after two
    // sweeps 8-23% of the tets of a muscle box on the floor are still above,
    // so a wave of 64 rollouts almost always needs the third)
#pragma unroll
    for (int sw = 0; sw < 3; ++sw) sweep();
    T off = fabs(a01) + fabs(a02) + fabs(a12);
    if (off > T(1e-7) * (a00 + a11 + a22)) sweep();
  }
  // Columns of V in the descending order of F^T F's eigenvalues (the
  // converged diagonal, a_ii = |F v_i|^2 up to rounding), so that the columns
  // of B = F V come out sorted by norm: only V's nine entries are swapped.
  auto swapv = [&](int i, int j) {
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      T t = v[r * 3 + i];
      v[r * 3 + i] = v[r * 3 + j];
      v[r * 3 + j] = t;
    }
  };
  if (a00 < a11) {
    swapv(0, 1);
    T t = a00; a00 = a11; a11 = t;
  }
  if (a00 < a22) {
    swapv(0, 2);
    T t = a00; a00 = a22; a22 = t;
  }
  if (a11 < a22) swapv(1, 2);
  T b[9];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) b[r * 3 + c] = F[r * 3 + 0] * v[0 * 3 + c] + F[r * 3 + 1] * v[1 * 3 + c] + F[r * 3 + 2] * v[2 * 3 + c];
  const T n0 = b[0] * b[0] + b[3] * b[3] + b[6] * b[6];
  // Gram-Schmidt QR of b
  T s0 = Num<T>::sqrt_(n0);
  T u0[3], u1[3], u2[3];
  if (s0 > Num<T>::kTiny) {
    T i0 = Num<T>::rcp(s0);
    u0[0] = b[0] * i0; u0[1] = b[3] * i0; u0[2] = b[6] * i0;
  } else {
    u0[0] = 1; u0[1] = 0; u0[2] = 0;
  }
  T d01 = u0[0] * b[1] + u0[1] * b[4] + u0[2] * b[7];
  T w1[3] = {b[1] - d01 * u0[0], b[4] - d01 * u0[1], b[7] - d01 * u0[2]};
  T s1 = Num<T>::sqrt_(w1[0] * w1[0] + w1[1] * w1[1] + w1[2] * w1[2]);
  if (s1 > Num<T>::kRel * s0 + Num<T>::kTiny) {
    T i1 = Num<T>::rcp(s1);
    u1[0] = w1[0] * i1; u1[1] = w1[1] * i1; u1[2] = w1[2] * i1;
  } else {
    // any unit vector orthogonal to u0
    T ax = fabs(u0[0]), ay = fabs(u0[1]), az = fabs(u0[2]);
    T e[3] = {0, 0, 0};
    if (ax <= ay && ax <= az) e[0] = 1; else if (ay <= az) e[1] = 1; else e[2] = 1;
    T d = u0[0] * e[0] + u0[1] * e[1] + u0[2] * e[2];
    T w[3] = {e[0] - d * u0[0], e[1] - d * u0[1], e[2] - d * u0[2]};
    T iw = Num<T>::rsqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
    u1[0] = w[0] * iw; u1[1] = w[1] * iw; u1[2] = w[2] * iw;
    s1 = 0;
  }
  u2[0] = u0[1] * u1[2] - u0[2] * u1[1];
  u2[1] = u0[2] * u1[0] - u0[0] * u1[2];
  u2[2] = u0[0] * u1[1] - u0[1] * u1[0];
  T s2 = u2[0] * b[2] + u2[1] * b[5] + u2[2] * b[8];
  if (s2 < 0) {
    s2 = -s2;
    u2[0] = -u2[0]; u2[1] = -u2[1]; u2[2] = -u2[2];
  }
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    U[r * 3 + 0] = u0[r];
    U[r * 3 + 1] = u1[r];
    U[r * 3 + 2] = u2[r];
  }
  S[0] = s0; S[1] = s1; S[2] = s2;
#pragma unroll
  for (int i = 0; i < 9; ++i) V[i] = v[i];
}

template <typename T>
__device__ __forceinline__ T det3(const T* a) {
  return a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) + a[2] * (a[3] * a[7] - a[4] * a[6]);
}

// 5-step linearised volume projection (CorotateFEMConstraint.h:362-374).
template <typename T>
__device__ __forceinline__ void volproj(const T* S, T* D) {
  D[0] = D[1] = D[2] = 0;
#pragma unroll
  for (int it = 0; it < 5; ++it) {
    T a0 = S[0] + D[0], a1 = S[1] + D[1], a2 = S[2] + D[2];
    T cd = a0 * a1 * a2 - T(1);
    T g0 = a1 * a2, g1 = a0 * a2, g2 = a0 * a1;
    T coef = Num<T>::div(g0 * D[0] + g1 * D[1] + g2 * D[2] - cd, g0 * g0 + g1 * g1 + g2 * g2);
    D[0] = coef * g0;
    D[1] = coef * g1;
    D[2] = coef * g2;
  }
}

// Volume projection with the two things its VJP needs: gS = J^T hd, where
// J = d(S + D)/dS, by a reverse sweep over the five stored iterates (one
// adjoint vector instead of nine forward tangents), and the divided
// differences rdd = (phi_i - phi_j)/(S_i - S_j) of phi = S + D for the pairs
// (0,1), (0,2), (1,2), carried through the iterations as ratios (with a the
// current iterate and k the third index, g_i - g_j = -a_k (a_i - a_j), so a
// ratio r becomes 1 - coef a_k r; r = 1 at D = 0): no cancellation where
// S_i ~ S_j, and the value at S_i = S_j is the derivative limit.
template <typename T>
__device__ __forceinline__ void volproj_vjp(const T* S, const T* hd, T* D, T* rdd, T* gS) {
  T dk[5][3];
  T r0 = 1, r1 = 1, r2 = 1;
  T d0 = 0, d1 = 0, d2 = 0;
#pragma unroll
  for (int it = 0; it < 5; ++it) {
    dk[it][0] = d0;
    dk[it][1] = d1;
    dk[it][2] = d2;
    T a0 = S[0] + d0, a1 = S[1] + d1, a2 = S[2] + d2;
    T cd = a0 * a1 * a2 - T(1);
    T g0 = a1 * a2, g1 = a0 * a2, g2 = a0 * a1;
    T coef = (g0 * d0 + g1 * d1 + g2 * d2 - cd) * Num<T>::rcp(g0 * g0 + g1 * g1 + g2 * g2);
    r0 = T(1) - coef * a2 * r0;
    r1 = T(1) - coef * a1 * r1;
    r2 = T(1) - coef * a0 * r2;
    d0 = coef * g0;
    d1 = coef * g1;
    d2 = coef * g2;
  }
  D[0] = d0;
  D[1] = d1;
  D[2] = d2;
  rdd[0] = r0;
  rdd[1] = r1;
  rdd[2] = r2;
  // reverse: phi = S + d_5; bar d_5 = hd, bar S = hd + sum over iterations
  T bd0 = hd[0], bd1 = hd[1], bd2 = hd[2];
  T bs0 = hd[0], bs1 = hd[1], bs2 = hd[2];
#pragma unroll
  for (int it = 4; it >= 0; --it) {
    T e0 = dk[it][0], e1 = dk[it][1], e2 = dk[it][2];
    T a0 = S[0] + e0, a1 = S[1] + e1, a2 = S[2] + e2;
    T cd = a0 * a1 * a2 - T(1);
    T g0 = a1 * a2, g1 = a0 * a2, g2 = a0 * a1;
    T igg = Num<T>::rcp(g0 * g0 + g1 * g1 + g2 * g2);
    T num = g0 * e0 + g1 * e1 + g2 * e2 - cd;
    T coef = num * igg;
    // d_new = coef g, coef = num / gg, num = g.e - cd, gg = |g|^2
    T bcoef = bd0 * g0 + bd1 * g1 + bd2 * g2;
    T bnum = bcoef * igg;
    T bgg = -bcoef * coef * igg;
    T bg0 = bd0 * coef + T(2) * g0 * bgg + bnum * e0;
    T bg1 = bd1 * coef + T(2) * g1 * bgg + bnum * e1;
    T bg2 = bd2 * coef + T(2) * g2 * bgg + bnum * e2;
    // num -> e (old iterate) directly and cd = a0 a1 a2 - 1 (bar cd = -bnum)
    T ba0 = -bnum * g0 + bg1 * a2 + bg2 * a1;
    T ba1 = -bnum * g1 + bg0 * a2 + bg2 * a0;
    T ba2 = -bnum * g2 + bg0 * a1 + bg1 * a0;
    bs0 += ba0;
    bs1 += ba1;
    bs2 += ba2;
    bd0 = ba0 + bnum * g0;
    bd1 = ba1 + bnum * g1;
    bd2 = ba2 + bnum * g2;
  }
  gS[0] = bs0;
  gS[1] = bs1;
  gS[2] = bs2;
}

// P = wr * md + wv * mdv, where md = U V^T (third column negated when
// det F < 0) and mdv = U diag(S + D) V^T.
template <typename T>
__device__ __forceinline__ void corotate_P(const T* F, T wr, T wv, T* P) {
  T U[9], S[3], V[9], D[3];
  svd3<T>(F, U, S, V);
  volproj<T>(S, D);
  bool flip = det3<T>(F) < 0;
  T e0 = S[0] + D[0], e1 = S[1] + D[1], e2 = S[2] + D[2];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      T rot = U[r * 3 + 0] * V[c * 3 + 0] + U[r * 3 + 1] * V[c * 3 + 1] + U[r * 3 + 2] * V[c * 3 + 2];
      T vol = U[r * 3 + 0] * e0 * V[c * 3 + 0] + U[r * 3 + 1] * e1 * V[c * 3 + 1] + U[r * 3 + 2] * e2 * V[c * 3 + 2];
      if (flip && c == 2) rot = -rot;
      P[r * 3 + c] = wr * rot + wv * vol;
    }
}

// Adjoint of the weighted projection: Fbar = (dP/dF)^T [Z]; also returns the
// projections md, mdv at F (needed for the parameter gradient), or, when md is
// null, only their inner products with Z in zmd, zmdv.
template <typename T>
__device__ __forceinline__ void corotate_vjp(const T* F, T wr, T wv, const T* Z, T* Fbar, T* md, T* mdv,
                                             T* zmd = nullptr, T* zmdv = nullptr) {
  T U[9], S[3], V[9], D[3], rdd[3];
  svd3<T>(F, U, S, V);
  bool flip = det3<T>(F) < 0;
  // H = U^T Z V and Hr = U^T (Z Df) V
  T ZU[9], H[9], Hr[9];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int c = 0; c < 3; ++c) ZU[i * 3 + c] = U[0 * 3 + i] * Z[0 * 3 + c] + U[1 * 3 + i] * Z[1 * 3 + c] + U[2 * 3 + i] * Z[2 * 3 + c];
  // Hr = H - 2 (U^T Z)[:, 2] V[2, :] when the third column is flipped
  const T fl = flip ? T(2) : T(0);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      T h = ZU[i * 3 + 0] * V[0 * 3 + j] + ZU[i * 3 + 1] * V[1 * 3 + j] + ZU[i * 3 + 2] * V[2 * 3 + j];
      H[i * 3 + j] = h;
      Hr[i * 3 + j] = h - fl * ZU[i * 3 + 2] * V[2 * 3 + j];
    }
  // volume projection, its divided differences and the diagonal of the
  // volume term's adjoint: O_ii = wv (J^T diag H)_i
  T O[9];
  {
    T hd[3] = {H[0], H[4], H[8]}, gS[3];
    volproj_vjp<T>(S, hd, D, rdd, gS);
#pragma unroll
    for (int i = 0; i < 3; ++i) O[i * 3 + i] = wv * gS[i];
  }
  T phi[3] = {S[0] + D[0], S[1] + D[1], S[2] + D[2]};
  if (md) {
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        T rot = U[r * 3 + 0] * V[c * 3 + 0] + U[r * 3 + 1] * V[c * 3 + 1] + U[r * 3 + 2] * V[c * 3 + 2];
        md[r * 3 + c] = (flip && c == 2) ? -rot : rot;
        mdv[r * 3 + c] = U[r * 3 + 0] * phi[0] * V[c * 3 + 0] + U[r * 3 + 1] * phi[1] * V[c * 3 + 1] +
                         U[r * 3 + 2] * phi[2] * V[c * 3 + 2];
      }
  } else {
    // Z : (U Df V^T) = tr(U^T Z Df V) = tr Hr and Z : (U diag(phi) V^T) = sum phi_k H_kk
    *zmd = Hr[0] + Hr[4] + Hr[8];
    *zmdv = phi[0] * H[0] + phi[1] * H[4] + phi[2] * H[8];
  }
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      if (i == j) continue;
      T ssum = S[i] + S[j];
      T isum = ssum > 0 ? Num<T>::rcp(ssum) : T(0);
      T br = T(2) * isum;
      T av = rdd[i + j - 1];  // pair (0,1) -> 0, (0,2) -> 1, (1,2) -> 2
      T bv = (phi[i] + phi[j]) * isum;
      T rot = T(0.5) * br * (Hr[i * 3 + j] - Hr[j * 3 + i]);
      T vol = T(0.5) * (av + bv) * H[i * 3 + j] + T(0.5) * (av - bv) * H[j * 3 + i];
      O[i * 3 + j] = wr * rot + wv * vol;
    }
  // Fbar = U O V^T
  T UO[9];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int j = 0; j < 3; ++j) UO[r * 3 + j] = U[r * 3 + 0] * O[0 * 3 + j] + U[r * 3 + 1] * O[1 * 3 + j] + U[r * 3 + 2] * O[2 * 3 + j];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) Fbar[r * 3 + c] = UO[r * 3 + 0] * V[c * 3 + 0] + UO[r * 3 + 1] * V[c * 3 + 1] + UO[r * 3 + 2] * V[c * 3 + 2];
}

// ------------------------------------------------------------ surface drag
// Underwater drag and lift on one surface face (Dfobj::ComputeDragForces,
// Dfobj.h:846-994). Inputs: n = (p2 - p1) x (p3 - p1) of the outward-oriented
// face and w = v_water - mean face velocity (v_water = 0). Output: the force
// added to each of the face's three vertices, drag/3 + lift/3. With a = n/|n|
// and wh = w/|w|: x = -a.wh, aoa = pi/2 - acos(x); faces with x <= 0 (moving
// inwards) get no force; drag = 500 area Cd(aoa) |w|^2 wh with Cd = sin^2(aoa);
// lift = 500 area Cl(aoa) |w|^2 l with l = normalize(wh x (-a)) x wh and Cl
// the reference's piecewise-linear table in degrees, zero beyond 85 degrees;
// each of drag and lift is scaled down to 1000 N if longer. Nothing under
// |w| < 1e-6.
//
// S is double (the force) or Dual1 (one forward-mode tangent: a Jacobian
// column). The branches are taken on values, as the reference's taped
// comparisons are. Cd(aoa(x)) is identically x^2, and its derivative is taken
// as 2x (the composite through acos is singular at x = 1, its product is not).
struct Dual1 {
  double v, d;
  __host__ __device__ Dual1(double v_ = 0.0, double d_ = 0.0) : v(v_), d(d_) {}
};
__host__ __device__ inline Dual1 operator+(Dual1 a, Dual1 b) { return Dual1(a.v + b.v, a.d + b.d); }
__host__ __device__ inline Dual1 operator-(Dual1 a, Dual1 b) { return Dual1(a.v - b.v, a.d - b.d); }
__host__ __device__ inline Dual1 operator-(Dual1 a) { return Dual1(-a.v, -a.d); }
__host__ __device__ inline Dual1 operator*(Dual1 a, Dual1 b) { return Dual1(a.v * b.v, a.d * b.v + a.v * b.d); }
__host__ __device__ inline Dual1 operator/(Dual1 a, Dual1 b) {
  double q = a.v / b.v;
  return Dual1(q, (a.d - q * b.d) / b.v);
}
__host__ __device__ inline double dval(double a) { return a; }
__host__ __device__ inline double dval(Dual1 a) { return a.v; }
__host__ __device__ inline double droot(double a) { return sqrt(a); }
__host__ __device__ inline Dual1 droot(Dual1 a) {
  double r = sqrt(a.v);
  return Dual1(r, a.d * 0.5 / r);
}

constexpr double kDragMaxForce = 1.0e3, kDragCutoff = 85.0, kPi = 3.14159265358979323846;

__host__ __device__ inline double drag_aoa(double x) {
  return 0.5 * kPi - acos(x > 1.0 ? 1.0 : (x < -1.0 ? -1.0 : x));
}
__host__ __device__ inline double drag_cd(double x) {
  double aoa = drag_aoa(x);
  return 1.0 * (0.5 * (-cos(2.0 * aoa) + 1.0)) + 0.0;
}
__host__ __device__ inline Dual1 drag_cd(Dual1 x) { return Dual1(drag_cd(x.v), 2.0 * x.v * x.d); }
// Cl at theta degrees and its slope d Cl / d theta.
__host__ __device__ inline double drag_cl_table(double theta, double* slope) {
  const double X[5] = {-90.0, -10.0, -5.0, 15.0, 90.0}, Y[5] = {-0.5, -1.26, 0.0, 1.8, 0.5};
  *slope = 0.0;
  if (fabs(theta) > kDragCutoff) return 0.0;
  int k = -1;
  if (X[0] <= theta && theta < X[1]) k = 0;
  else if (X[1] <= theta && theta < X[2]) k = 1;
  else if (X[2] <= theta && theta <= X[3]) k = 2;
  else if (X[3] <= theta && theta <= X[4]) k = 3;
  if (k < 0) return 0.0;
  double t = (theta - X[k]) / (X[k + 1] - X[k]);
  *slope = (Y[k + 1] - Y[k]) / (X[k + 1] - X[k]);
  return (1.0 - t) * Y[k] + t * Y[k + 1];
}
__host__ __device__ inline double drag_cl(double x) {
  double sl;
  return drag_cl_table(drag_aoa(x) * 180.0 / kPi, &sl);
}
__host__ __device__ inline Dual1 drag_cl(Dual1 x) {
  double sl;
  double c = drag_cl_table(drag_aoa(x.v) * 180.0 / kPi, &sl);
  if (sl == 0.0) return Dual1(c, 0.0);
  return Dual1(c, sl * (180.0 / kPi) / sqrt(1.0 - x.v * x.v) * x.d);
}

template <typename S>
__host__ __device__ inline void drag_clamp(S* f) {
  S q = f[0] * f[0] + f[1] * f[1] + f[2] * f[2];
  if (sqrt(dval(q)) > kDragMaxForce) {
    S l = droot(q);
    for (int c = 0; c < 3; ++c) f[c] = S(kDragMaxForce) * (f[c] / l);
  }
}

template <typename S>
__host__ __device__ inline void drag_force(const S* n, const S* w, S* F) {
  F[0] = F[1] = F[2] = S(0.0);
  S w2 = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
  if (sqrt(dval(w2)) < 1e-6) return;
  S wl = droot(w2);
  S wh[3] = {w[0] / wl, w[1] / wl, w[2] / wl};
  S nl = droot(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
  S d[3] = {-(n[0] / nl), -(n[1] / nl), -(n[2] / nl)};
  S area = S(0.5) * nl;
  S x = d[0] * wh[0] + d[1] * wh[1] + d[2] * wh[2];
  S drag[3] = {S(0.0), S(0.0), S(0.0)}, lift[3] = {S(0.0), S(0.0), S(0.0)};
  if (dval(x) > 0.0) {
    S k = S(500.0) * area * drag_cd(x) * w2;
    for (int c = 0; c < 3; ++c) drag[c] = k * wh[c];
    S u[3] = {wh[1] * d[2] - wh[2] * d[1], wh[2] * d[0] - wh[0] * d[2], wh[0] * d[1] - wh[1] * d[0]};
    S u2 = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    if (dval(u2) > 0.0) {
      S ul = droot(u2);
      for (int c = 0; c < 3; ++c) u[c] = u[c] / ul;
    }
    S l[3] = {u[1] * wh[2] - u[2] * wh[1], u[2] * wh[0] - u[0] * wh[2], u[0] * wh[1] - u[1] * wh[0]};
    S kl = S(500.0) * area * drag_cl(x) * w2;
    for (int c = 0; c < 3; ++c) lift[c] = kl * l[c];
  }
  drag_clamp(drag);
  drag_clamp(lift);
  for (int c = 0; c < 3; ++c) F[c] = drag[c] / S(3.0) + lift[c] / S(3.0);
}

// Transposed Jacobian of drag_force applied to gF: gn = (dF/dn)^T gF and
// gw = (dF/dw)^T gF, one forward-mode pass per input.
__host__ __device__ inline void drag_force_vjp(const double* n, const double* w, const double* gF, double* gn,
                                               double* gw) {
  for (int s = 0; s < 6; ++s) {
    Dual1 nd[3], wd[3], F[3];
    for (int c = 0; c < 3; ++c) {
      nd[c] = Dual1(n[c], s == c ? 1.0 : 0.0);
      wd[c] = Dual1(w[c], s == 3 + c ? 1.0 : 0.0);
    }
    drag_force(nd, wd, F);
    double g = gF[0] * F[0].d + gF[1] * F[1].d + gF[2] * F[2].d;
    if (s < 3) gn[s] = g;
    else gw[s - 3] = g;
  }
}

}  // namespace pdm
