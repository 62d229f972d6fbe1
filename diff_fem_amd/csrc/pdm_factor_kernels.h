// pdm_factor_kernels.h — numeric re-factorisation of K_s = M/h^2 + L on the
// device (the reference's World::PreComputation / FactorizeLDLT,
// sim/fem/World.h:1089-1139, 1212-1232, re-done on every forward_all,
// python/pydifem.cc:207-221), over the host's symbolic analysis: supernodes
// in postorder, processed by elimination level (leaves first).
//
// Per re-factorisation:
//   k_fac_assemble_L  values of the scalar matrix L from the per-tet weights
//                     (the host assembly's sums, in its order, rounded alike)
//   k_fac_kdiag       diag(K_s) for the contact Jacobi update
//   k_fac_scatter/diag  fronts start from A's lower entries + m_i/h^2 + reg
// and per level:
//   k_fac_extend      extend-add of the children's update blocks
//   k_fac_panel       partial Cholesky of the supernode's w columns, then
//                     inv(L_ss) into W's top rows
//   k_fac_update      W's lower rows -L_below inv(L_ss) and the update block
//                     U = U_front - L_below L_below^T, as 64x64 output tiles
//
// Layouts (offsets from the symbolic analysis): L of supernode s is m x w
// column-major at lptr[s]; W is stored column-major (Wc) and row-major (WR) at
// the same offsets; the update block of s is (m-w) x (m-w) row-major (lower
// part used) at upos[s]. Every sum runs over its terms in the host
// factorisation's order; the factor kernels contract to FMA.
#pragma once

namespace pdmfac {

constexpr int kTile = 64;     // output tile edge of k_fac_update and row tile of k_fac_extend
constexpr int kChunkK = 16;   // inner-dimension chunk staged in LDS

// One thread per stored entry q of L (CSR): the sum of its contributions in
// the host assembly's order (tets ascending, then muscles, attachments and the
// structural 0 of the diagonal). A tet term is (wr_t + wv_t) * gg with the
// static gradient product gg; other terms are static values (ctet = -1).
__global__ void k_fac_assemble_L(const long long* __restrict__ cptr, const int* __restrict__ ctet,
                                 const double* __restrict__ cval, const double* __restrict__ wr,
                                 const double* __restrict__ wv, long long nnz, double* __restrict__ Lval) {
#pragma clang fp contract(off)
  const long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (q >= nnz) return;
  double v = 0.0;
  for (long long c = cptr[q]; c < cptr[q + 1]; ++c) {
    const int t = ctet[c];
    if (t >= 0)
      v += (wr[t] + wv[t]) * cval[c];
    else
      v += cval[c];
  }
  Lval[q] = v;
}

// Kdiag_i = m_i / h^2 + L_ii.
__global__ void k_fac_kdiag(const int* __restrict__ Lptr, const int* __restrict__ Lcol,
                            const double* __restrict__ Lval, const double* __restrict__ mass, double h, int N,
                            double* __restrict__ Kdiag) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  double kd = mass[i] / (h * h);
  for (int q = Lptr[i]; q < Lptr[i + 1]; ++q)
    if (Lcol[q] == i) kd += Lval[q];
  Kdiag[i] = kd;
}

// Front entries from A (amap: position in Lf of each CSR entry on or below
// the diagonal, -1 above); Lf is zero beforehand.
__global__ void k_fac_scatter(const double* __restrict__ Lval, const long long* __restrict__ amap, long long nnz,
                              double* __restrict__ Lf) {
  const long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (q >= nnz) return;
  const long long a = amap[q];
  if (a >= 0) Lf[a] = 0.0 + Lval[q];
}

// Diagonal shift m_i/h^2 + reg (the regularisation ladder's reg, 0 first).
__global__ void k_fac_diag(const long long* __restrict__ dpos, const double* __restrict__ mass, double h, double reg,
                           int N, double* __restrict__ Lf) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  Lf[dpos[i]] += mass[i] / (h * h) + reg;
}

__device__ inline int lower_bound_i(const int* a, int n, int x) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (a[mid] < x)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

// Extend-add, children in order. Block = (supernode s, parent row tile
// [64 t, 64 t + 64)): it adds the child entries (a, b), b <= a, whose parent row
// rel[a] falls in the tile, so no two blocks write one entry; the waves take
// child rows in turn, the lanes the columns b. Entries of the w pivot columns
// go to L, the others to the update block of s.
__global__ __launch_bounds__(256) void k_fac_extend(const int2* __restrict__ items, const int* __restrict__ first,
                                                    const long long* __restrict__ rowptr,
                                                    const long long* __restrict__ lptr,
                                                    const long long* __restrict__ upos,
                                                    const int* __restrict__ child_ptr, const int* __restrict__ child,
                                                    const int* __restrict__ rel, const long long* __restrict__ relptr,
                                                    double* __restrict__ Lf, double* __restrict__ Uf) {
  const int2 it = items[blockIdx.x];
  const int s = it.x;
  const int w = first[s + 1] - first[s], m = (int)(rowptr[s + 1] - rowptr[s]), mb = m - w;
  const int r0 = it.y * kTile, r1 = min(r0 + kTile, m);
  double* L = Lf + lptr[s];
  double* U = Uf + upos[s];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int q = child_ptr[s]; q < child_ptr[s + 1]; ++q) {
    const int c = child[q];
    const int wc = first[c + 1] - first[c], mbc = (int)(rowptr[c + 1] - rowptr[c]) - wc;
    const int* rc = rel + relptr[c];
    const double* Uc = Uf + upos[c];
    const int a0 = lower_bound_i(rc, mbc, r0), a1 = lower_bound_i(rc, mbc, r1);
    for (int a = a0 + wave; a < a1; a += 4) {
      const int pa = rc[a];
      const double* ua = Uc + (long long)a * mbc;
      for (int b = lane; b <= a; b += 64) {
        const int pb = rc[b];
        const double v = ua[b];
        if (pb < w)
          L[(long long)pb * m + pa] += v;
        else
          U[(long long)(pa - w) * mb + (pb - w)] += v;
      }
    }
    __syncthreads();
  }
}

// Partial Cholesky of supernode s's w columns (m x w, column-major, in place),
// right-looking, so every entry receives its updates in column order: column j
// takes its pivot sqrt(d) and is scaled by 1/sqrt(d), then updates the later
// columns' rows at and below their diagonal. Blocked by nb <= 16 columns: a
// block column (rows j0..m-1) is factored in LDS (dynamic, nb * m doubles,
// rows contiguous), written back, and updates the trailing columns from LDS. A
// pivot that is not positive sets *fail (the host then retries with the next
// regularisation). Then inv(L_ss) by forward substitution, thread k owning
// column k, by blocks of nb rows whose L rows are staged in LDS, into WR's top
// w rows (row-major, zero above the diagonal) and Wc's.
__global__ __launch_bounds__(256) void k_fac_panel(const int* __restrict__ sns, const int* __restrict__ first,
                                                   const long long* __restrict__ rowptr,
                                                   const long long* __restrict__ lptr, double* __restrict__ Lf,
                                                   double* __restrict__ Wc, double* __restrict__ WR,
                                                   int* __restrict__ fail, int nb) {
  extern __shared__ double P[];
  __shared__ double sh_inv;
  const int s = sns[blockIdx.x];
  const int w = first[s + 1] - first[s], m = (int)(rowptr[s + 1] - rowptr[s]);
  double* L = Lf + lptr[s];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  for (int j0 = 0; j0 < w; j0 += nb) {
    const int bw = min(nb, w - j0), mr = m - j0;
    for (int e = tid; e < bw * mr; e += 256) {
      const int kk = e / mr, rr = e - kk * mr;
      P[kk * mr + rr] = L[(long long)(j0 + kk) * m + j0 + rr];
    }
    __syncthreads();
    for (int jj = 0; jj < bw; ++jj) {
      double* cj = P + jj * mr;
      if (tid == 0) {
        double d = cj[jj];
        if (!(d > 0.0)) {
          *fail = 1;
          d = 1.0;
        }
        d = sqrt(d);
        cj[jj] = d;
        sh_inv = 1.0 / d;
      }
      __syncthreads();
      const double inv = sh_inv;
      for (int i = jj + 1 + tid; i < mr; i += 256) cj[i] *= inv;
      __syncthreads();
      for (int jp = jj + 1 + wave; jp < bw; jp += 4) {
        const double f = cj[jp];
        double* cp = P + jp * mr;
        for (int i = jp + lane; i < mr; i += 64) cp[i] -= f * cj[i];
      }
      __syncthreads();
    }
    for (int e = tid; e < bw * mr; e += 256) {
      const int kk = e / mr, rr = e - kk * mr;
      L[(long long)(j0 + kk) * m + j0 + rr] = P[kk * mr + rr];
    }
    for (int jp = j0 + bw + wave; jp < w; jp += 4) {
      double f[16];
#pragma unroll
      for (int kk = 0; kk < 16; ++kk) f[kk] = kk < bw ? P[kk * mr + (jp - j0)] : 0.0;
      double* cp = L + (long long)jp * m;
      for (int i = jp + lane; i < m; i += 64) {
        double acc = cp[i];
#pragma unroll
        for (int kk = 0; kk < 16; ++kk)
          if (kk < bw) acc -= f[kk] * P[kk * mr + (i - j0)];
        cp[i] = acc;
      }
    }
    __syncthreads();
  }
  double* D = WR + lptr[s];
  for (int r0 = 0; r0 < w; r0 += nb) {
    const int bh = min(nb, w - r0), ncol = r0 + bh;
    for (int e = tid; e < bh * ncol; e += 256) {  // Q[rr][q] = L[r0 + rr][q]
      const int q = e / bh, rr = e - q * bh;
      P[rr * ncol + q] = L[(long long)q * m + r0 + rr];
    }
    __syncthreads();
    for (int k = tid; k < w; k += 256) {
      double acc[16], dblk[16];
#pragma unroll
      for (int rr = 0; rr < 16; ++rr) {
        acc[rr] = (r0 + rr == k) ? 1.0 : 0.0;
        dblk[rr] = 0.0;
      }
#pragma unroll 4
      for (int q = k; q < r0; ++q) {
        const double dq = D[(long long)q * w + k];
#pragma unroll
        for (int rr = 0; rr < 16; ++rr)
          if (rr < bh) acc[rr] -= P[rr * ncol + q] * dq;
      }
#pragma unroll
      for (int rr = 0; rr < 16; ++rr) {
        if (rr >= bh) continue;
        const int r = r0 + rr;
        if (r < k) {
          D[(long long)r * w + k] = 0.0;
          continue;
        }
#pragma unroll
        for (int qq = 0; qq < 16; ++qq)
          if (qq < rr && r0 + qq >= k) acc[rr] -= P[rr * ncol + r0 + qq] * dblk[qq];
        dblk[rr] = acc[rr] / P[rr * ncol + r];
        D[(long long)r * w + k] = dblk[rr];
      }
    }
    __syncthreads();
  }
  double* Wt = Wc + lptr[s];
  for (int e = tid; e < w * w; e += 256) {
    const int k = e / w, r = e - k * w;
    Wt[(long long)k * m + r] = D[(long long)r * w + k];
  }
}

// 64x64 output tiles, 16x16 threads with 4x4 outputs each (rows ty + 16 i,
// columns tx + 16 j), the inner dimension staged 16 at a time through LDS.
// kind 0: W's lower rows, W[w + r][k] = -sum_{jj >= k} L[w + r][jj] inv(L_ss)[jj][k];
// kind 1: the update block, U[a][b] -= sum_k L[w + a][k] L[w + b][k], b <= a.
__global__ __launch_bounds__(256) void k_fac_update(const int4* __restrict__ items, const int* __restrict__ first,
                                                    const long long* __restrict__ rowptr,
                                                    const long long* __restrict__ lptr,
                                                    const long long* __restrict__ upos,
                                                    const double* __restrict__ Lf, double* __restrict__ Uf,
                                                    double* __restrict__ Wc, double* __restrict__ WR) {
  const int4 it = items[blockIdx.x];
  const int s = it.x, kind = it.y;
  const int w = first[s + 1] - first[s], m = (int)(rowptr[s + 1] - rowptr[s]), mb = m - w;
  const double* L21 = Lf + lptr[s] + w;  // L21[r][k] at L21[k * m + r]
  const double* D = WR + lptr[s];        // inv(L_ss), row-major w x w
  double* U = Uf + upos[s];
  const int row0 = it.z * kTile, col0 = it.w * kTile;
  const int ncol = kind == 0 ? w : mb;
  const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
  __shared__ double As[kChunkK][kTile], Bs[kChunkK][kTile];
  double acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int a = row0 + ty + 16 * i, b = col0 + tx + 16 * j;
      acc[i][j] = (kind == 1 && a < mb && b <= a) ? U[(long long)a * mb + b] : 0.0;
    }
  const int kbeg = kind == 0 ? col0 : 0;
  for (int k0 = kbeg; k0 < w; k0 += kChunkK) {
#pragma unroll
    for (int e = tid; e < kChunkK * kTile; e += 256) {
      const int kk = e / kTile, rr = e % kTile;
      const int k = k0 + kk, r = row0 + rr, c = col0 + rr;
      As[kk][rr] = (k < w && r < mb) ? L21[(long long)k * m + r] : 0.0;
      double bv = 0.0;
      if (k < w && c < ncol) bv = kind == 0 ? D[(long long)k * w + c] : L21[(long long)k * m + c];
      Bs[kk][rr] = bv;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < kChunkK; ++kk) {
      double av[4], bv[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) av[i] = As[kk][ty + 16 * i];
#pragma unroll
      for (int j = 0; j < 4; ++j) bv[j] = Bs[kk][tx + 16 * j];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] -= av[i] * bv[j];
    }
    __syncthreads();
  }
  const long long lp = lptr[s];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int a = row0 + ty + 16 * i, b = col0 + tx + 16 * j;
      if (a >= mb || b >= ncol) continue;
      if (kind == 0) {
        Wc[lp + (long long)b * m + w + a] = acc[i][j];
        WR[lp + (long long)(w + a) * w + b] = acc[i][j];
      } else if (b <= a) {
        U[(long long)a * mb + b] = acc[i][j];
      }
    }
}

}  // namespace pdmfac
