// pdm_self_adj.h — reverse of one self-contact group of a handling
// (World.h:1451-1530 with do_solve_friction, World.h:1246-1315), for the
// device adjoint (PDM_HD = __host__ __device__ there; plain C++ otherwise).
//
// Frozen as executed: the basis [e0 t b] (from the leader's impact normal),
// which members approached (in member order), and the final stick flags; the
// last do_solve_friction call recomputes every output from these, so the
// group is a smooth function of the forces F (rhs + obstacle reactions, world
// frame) and masses m (diagonal of K_s) of its three face corners and its
// approaching members. Outputs: the face corners' changes v0 m_q - F_q and
// the members' reactions E rs_k.
#pragma once
#include <cmath>

#ifndef PDM_HD
#define PDM_HD
#endif

namespace pdmself {

constexpr int kMaxMem = 16;

// E[c][r]: component r of basis column c. Ff, mf: face corners; Fm, mm, stick:
// the nv approaching members. GFb, GVb: adjoints of the outputs. Ffb, Fmb, mfb,
// mmb: adjoints of the inputs (accumulated).
PDM_HD inline void group_reverse(const double (*E)[3], int nv, const double* mf, const double* mm,
                                 const double (*Ff)[3], const double (*Fm)[3], const bool* stick, double mu,
                                 const double (*GFb)[3], const double (*GVb)[3], double (*Ffb)[3],
                                 double (*Fmb)[3], double* mfb, double* mmb) {
  auto toloc = [&](const double* w, double* l) {
    for (int c = 0; c < 3; ++c) l[c] = E[c][0] * w[0] + E[c][1] * w[1] + E[c][2] * w[2];
  };
  auto toworld = [&](const double* l, double* w) {
    for (int r = 0; r < 3; ++r) w[r] = E[0][r] * l[0] + E[1][r] * l[1] + E[2][r] * l[2];
  };
  // ---- forward, keeping what the reverse needs
  double f0w[3] = {0.0, 0.0, 0.0}, f0[3];
  for (int q = 0; q < 3; ++q)
    for (int c = 0; c < 3; ++c) f0w[c] += Ff[q][c];
  toloc(f0w, f0);
  const double m0 = mf[0] + mf[1] + mf[2];
  double fs[kMaxMem][3], rs[kMaxMem][3], sm = 0.0, smv = 0.0;
  for (int k = 0; k < nv; ++k) {
    toloc(Fm[k], fs[k]);
    sm += mm[k];
    smv += fs[k][0];
  }
  const double A = f0[0] + smv, D = m0 + sm;
  double r0[3] = {0.0, 0.0, 0.0};
  for (int k = 0; k < nv; ++k) {
    const double fn = A * mm[k] / D - fs[k][0];
    rs[k][0] = fn;
    rs[k][1] = 0.0;
    rs[k][2] = 0.0;
    r0[0] -= fn;
  }
  double pst[3] = {f0[0], f0[1], f0[2]}, mst = m0, psl[3] = {0.0, 0.0, 0.0}, msl = 0.0, fnl = 0.0;
  for (int k = 0; k < nv; ++k) {
    if (stick[k]) {
      for (int c = 0; c < 3; ++c) pst[c] += fs[k][c];
      mst += mm[k];
    } else {
      for (int c = 0; c < 3; ++c) psl[c] += fs[k][c];
      msl += mm[k];
      fnl += rs[k][0];
    }
  }
  const double v[3] = {pst[0] / mst, pst[1] / mst, pst[2] / mst};
  r0[1] = v[1] * m0 - f0[1];
  r0[2] = v[2] * m0 - f0[2];
  for (int k = 0; k < nv; ++k)
    if (stick[k]) {
      rs[k][1] = v[1] * mm[k] - fs[k][1];
      rs[k][2] = v[2] * mm[k] - fs[k][2];
    }
  const bool sl = msl > 0.0;
  double tl = 1.0, rho[3] = {0.0, 0.0, 0.0}, mult = 0.0, rhov[kMaxMem][3], t2[kMaxMem], mmk[kMaxMem];
  if (sl) {
    double rel[3];
    for (int c = 0; c < 3; ++c) rel[c] = psl[c] / msl - pst[c] / mst;
    tl = sqrt(rel[1] * rel[1] + rel[2] * rel[2]);
    for (int c = 0; c < 3; ++c) rho[c] = rel[c] / tl;
    mult = mu * fnl / (msl + mst);
    r0[1] += rho[1] * mult * m0;
    r0[2] += rho[2] * mult * m0;
    for (int k = 0; k < nv; ++k)
      if (stick[k]) {
        rs[k][1] += rho[1] * mult * mm[k];
        rs[k][2] += rho[2] * mult * mm[k];
      }
    double av[3];
    for (int c = 0; c < 3; ++c) av[c] = (f0[c] + r0[c]) / m0;
    for (int k = 0; k < nv; ++k)
      if (!stick[k]) {
        double rv[3];
        for (int c = 0; c < 3; ++c) rv[c] = fs[k][c] / mm[k] - av[c];
        t2[k] = sqrt(rv[1] * rv[1] + rv[2] * rv[2]);
        for (int c = 0; c < 3; ++c) rhov[k][c] = rv[c] / t2[k];
        mmk[k] = -mu * rs[k][0] * mm[k] / (mm[k] + mst);
        rs[k][1] = mmk[k] * rhov[k][1];
        rs[k][2] = mmk[k] * rhov[k][2];
      }
  }
  double r0w[3], bf0[3], v0[3];
  toworld(r0, r0w);
  toworld(f0, bf0);
  for (int c = 0; c < 3; ++c) v0[c] = (bf0[c] + r0w[c]) / m0;

  // ---- reverse
  double f0b[3] = {0.0, 0.0, 0.0}, r0b[3] = {0.0, 0.0, 0.0}, m0b = 0.0, v0b[3] = {0.0, 0.0, 0.0};
  double fsb[kMaxMem][3], rsb[kMaxMem][3], mml[kMaxMem];
  for (int q = 0; q < 3; ++q)
    for (int c = 0; c < 3; ++c) {
      v0b[c] += GFb[q][c] * mf[q];
      mfb[q] += GFb[q][c] * v0[c];
      Ffb[q][c] -= GFb[q][c];
    }
  for (int k = 0; k < nv; ++k) {
    toloc(GVb[k], rsb[k]);
    fsb[k][0] = fsb[k][1] = fsb[k][2] = 0.0;
    mml[k] = 0.0;
  }
  {  // v0 = (E f0 + E r0) / m0
    double b[3], lb[3];
    for (int c = 0; c < 3; ++c) {
      b[c] = v0b[c] / m0;
      m0b -= v0b[c] * v0[c] / m0;
    }
    toloc(b, lb);
    for (int c = 0; c < 3; ++c) {
      f0b[c] += lb[c];
      r0b[c] += lb[c];
    }
  }
  double pstb[3] = {0.0, 0.0, 0.0}, pslb[3] = {0.0, 0.0, 0.0}, mstb = 0.0, mslb = 0.0, fnlb = 0.0;
  if (sl) {
    double avb[3] = {0.0, 0.0, 0.0}, rhob[3] = {0.0, 0.0, 0.0}, multb = 0.0;
    for (int k = 0; k < nv; ++k)
      if (!stick[k]) {
        const double mkb = rsb[k][1] * rhov[k][1] + rsb[k][2] * rhov[k][2];
        const double hb1 = mmk[k] * rsb[k][1], hb2 = mmk[k] * rsb[k][2];
        rsb[k][1] = rsb[k][2] = 0.0;
        const double den = mm[k] + mst;
        rsb[k][0] += mkb * (-mu * mm[k] / den);
        mml[k] += mkb * (-mu * rs[k][0]) * (mst / (den * den));
        mstb += mkb * (mu * rs[k][0] * mm[k] / (den * den));
        const double dot = rhov[k][1] * hb1 + rhov[k][2] * hb2;
        const double rb1 = (hb1 - rhov[k][1] * dot) / t2[k], rb2 = (hb2 - rhov[k][2] * dot) / t2[k];
        fsb[k][1] += rb1 / mm[k];
        fsb[k][2] += rb2 / mm[k];
        mml[k] -= (rb1 * fs[k][1] + rb2 * fs[k][2]) / (mm[k] * mm[k]);
        avb[1] -= rb1;
        avb[2] -= rb2;
      }
    for (int c = 1; c < 3; ++c) {  // av = (f0 + r0) / m0
      f0b[c] += avb[c] / m0;
      r0b[c] += avb[c] / m0;
      m0b -= avb[c] * (f0[c] + r0[c]) / (m0 * m0);
    }
    for (int k = 0; k < nv; ++k)
      if (stick[k])
        for (int c = 1; c < 3; ++c) {  // rs_k += rho mult m_k
          rhob[c] += rsb[k][c] * mult * mm[k];
          multb += rsb[k][c] * rho[c] * mm[k];
          mml[k] += rsb[k][c] * rho[c] * mult;
        }
    for (int c = 1; c < 3; ++c) {  // r0 += rho mult m0
      rhob[c] += r0b[c] * mult * m0;
      multb += r0b[c] * rho[c] * m0;
      m0b += r0b[c] * rho[c] * mult;
    }
    const double dd = msl + mst;  // mult = mu fnl / (msl + mst)
    fnlb += multb * mu / dd;
    mslb -= multb * mult / dd;
    mstb -= multb * mult / dd;
    const double dot = rho[1] * rhob[1] + rho[2] * rhob[2];  // rho = rel / |rel_12|
    for (int c = 1; c < 3; ++c) {
      const double relb = (rhob[c] - rho[c] * dot) / tl;  // rel = psl/msl - pst/mst
      pslb[c] += relb / msl;
      mslb -= relb * psl[c] / (msl * msl);
      pstb[c] -= relb / mst;
      mstb += relb * pst[c] / (mst * mst);
    }
  }
  double vb[3] = {0.0, 0.0, 0.0};
  for (int k = 0; k < nv; ++k)
    if (stick[k])
      for (int c = 1; c < 3; ++c) {  // rs_k = v m_k - fs_k
        vb[c] += rsb[k][c] * mm[k];
        mml[k] += rsb[k][c] * v[c];
        fsb[k][c] -= rsb[k][c];
      }
  for (int c = 1; c < 3; ++c) {  // r0 = v m0 - f0
    vb[c] += r0b[c] * m0;
    m0b += r0b[c] * v[c];
    f0b[c] -= r0b[c];
  }
  for (int c = 1; c < 3; ++c) {  // v = pst / mst
    pstb[c] += vb[c] / mst;
    mstb -= vb[c] * pst[c] / (mst * mst);
  }
  for (int k = 0; k < nv; ++k) {
    if (stick[k]) {
      for (int c = 0; c < 3; ++c) fsb[k][c] += pstb[c];
      mml[k] += mstb;
    } else {
      for (int c = 0; c < 3; ++c) fsb[k][c] += pslb[c];
      mml[k] += mslb;
      rsb[k][0] += fnlb;
    }
  }
  for (int c = 0; c < 3; ++c) f0b[c] += pstb[c];
  m0b += mstb;
  double Ab = 0.0, Db = 0.0;
  for (int k = 0; k < nv; ++k) {  // fn_k = A m_k / D - fs_k0; rs_k0 = fn_k; r0_0 = -sum fn
    const double fnb = rsb[k][0] - r0b[0];
    Ab += fnb * mm[k] / D;
    mml[k] += fnb * A / D;
    Db -= fnb * A * mm[k] / (D * D);
    fsb[k][0] -= fnb;
  }
  f0b[0] += Ab;
  m0b += Db;
  for (int k = 0; k < nv; ++k) {
    fsb[k][0] += Ab;  // smv
    mml[k] += Db;     // sm
    double w[3];
    toworld(fsb[k], w);
    for (int c = 0; c < 3; ++c) Fmb[k][c] += w[c];
    mmb[k] += mml[k];
  }
  double w[3];
  toworld(f0b, w);
  for (int q = 0; q < 3; ++q) {
    for (int c = 0; c < 3; ++c) Ffb[q][c] += w[c];
    mfb[q] += m0b;
  }
}

}  // namespace pdmself
