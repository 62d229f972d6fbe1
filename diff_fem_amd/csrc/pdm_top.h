// Dense separator blocks at the top of the elimination trees (included by
// pdm_engine.hip after the supernodal sweep kernels; uses v4d, mfma16,
// mfma_chain, FactorView and xcd_block from there).
//
// In each elimination tree the supernodes of height >= a per-tree cut form an
// upward-closed set that holds the root: the top separators of the nested
// dissection. Let t be their n rows (engine order, so the factor restricted to
// them, L_t, is lower triangular) and f_t the front the supernodal forward
// sweep would assemble on them: the right-hand side of those rows plus the
// update blocks of the top supernodes' children below the cut. Both sweeps
// over the top rows then reduce to one dense product
//   x_t = (L_t L_t^T)^{-1} f_t = Z f_t,   Z = Linv^T Linv,  Linv = inv(L_t),
// and the rows below the cut see x_t in the backward sweep exactly as they
// would see the ancestors' results of the level sweeps. Z is formed once per
// factorisation on the device: Linv by applying the top supernodes' own
// elimination steps [y_s; T_below] = W_s T_s (postorder) to the identity,
// then the Gram product. The block of tree r is np_r x np_r (n_r padded to a
// multiple of 32, zero padding), row-major.

// Identity of one block of np columns, zero in the rows past np (rows x np).
__global__ void k_top_eye(double* __restrict__ M, int np, int rows) {
  const long long e = blockIdx.x * 256LL + threadIdx.x;
  if (e >= (long long)rows * np) return;
  M[e] = (e / np == e % np) ? 1.0 : 0.0;
}

// Elimination wave, part 1: the own rows of each item's supernode (item =
// (supernode s, tree block r, first local row of s, front-row offset)) copied
// out of the block, T_s = M_r[p_s : p_s + w_s, :], before part 2 overwrites
// them. Grid: (element chunks, items of the wave).
__global__ void k_top_copy(const int4* __restrict__ items, const long long* __restrict__ toff,
                           const int* __restrict__ first, const int* __restrict__ np_of,
                           const long long* __restrict__ zoff, const double* __restrict__ M, double* __restrict__ T) {
  const int4 it = items[blockIdx.y];
  const int w = first[it.x + 1] - first[it.x];
  const int np = np_of[it.y];
  const long long e = blockIdx.x * 256LL + threadIdx.x;
  if (e >= (long long)w * np) return;
  const long long j = e / np, c = e % np;
  T[toff[blockIdx.y] + e] = M[zoff[it.y] + (it.z + j) * np + c];
}

// Elimination wave, part 2: for every front row i of s and column c of the
// block, out = sum_k W_s[i][k] T_s[k][c] (k < w_s); own rows (i < w_s) take
// it, the rows below add it: M_r <- E_s M_r. One wave per 16 x 16 tile. Items
// of one wave belong to different trees, so no element is written twice.
__global__ __launch_bounds__(64) void k_top_elim(FactorView f, const int4* __restrict__ items,
                                                 const long long* __restrict__ toff, const int* __restrict__ frow,
                                                 const int* __restrict__ np_of, const long long* __restrict__ zoff,
                                                 const double* __restrict__ T, double* __restrict__ M) {
  const int4 it = items[blockIdx.y];
  const int s = it.x;
  const int w = f.first[s + 1] - f.first[s];
  const int m = (int)(f.rowptr[s + 1] - f.rowptr[s]);
  const int np = np_of[it.y];
  const int ntc = np >> 4, nrt = (m + 15) >> 4;
  const int tile = blockIdx.x;
  if (tile >= nrt * ntc) return;
  const int rt = (tile / ntc) * 16, ct = (tile % ntc) * 16;
  const int lane = threadIdx.x & 63, lr = lane & 15, lk = lane >> 4;
  const double* Ws = f.W + f.lptr[s];
  const double* Ts = T + toff[blockIdx.y] + ct + lr;
  const int row = rt + lr;
  v4d acc = {0.0, 0.0, 0.0, 0.0};
  acc = mfma_chain(
      acc, 0, w, lk, [&](int k) { return row < m ? Ws[(long long)k * m + row] : 0.0; },
      [&](int k) { return Ts[(long long)k * np]; });
  double* Mr = M + zoff[it.y] + ct + lr;
  const int* fr = frow + it.w;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int r = rt + 4 * i + lk;
    if (r < m) {
      double* p = Mr + (long long)fr[r] * np;
      if (r < w)
        *p = acc[i];
      else
        *p += acc[i];
    }
  }
}

// Z = Linv^T Linv for one block: Z[i][j] = sum_{k >= max(i, j)} Linv[k][i] Linv[k][j]
// (Linv is lower triangular). One wave per 16 x 16 tile.
__global__ __launch_bounds__(64) void k_top_gram(const double* __restrict__ Mi, double* __restrict__ Z, int np) {
  const int ntc = np >> 4;
  const int tile = blockIdx.x;
  if (tile >= ntc * ntc) return;
  const int i0 = (tile / ntc) * 16, j0 = (tile % ntc) * 16;
  const int lane = threadIdx.x & 63, lr = lane & 15, lk = lane >> 4;
  v4d acc = {0.0, 0.0, 0.0, 0.0};
  acc = mfma_chain(
      acc, max(i0, j0), np, lk, [&](int k) { return Mi[(long long)k * np + i0 + lr]; },
      [&](int k) { return Mi[(long long)k * np + j0 + lr]; });
#pragma unroll
  for (int i = 0; i < 4; ++i) Z[(long long)(i0 + 4 * i + lk) * np + j0 + lr] = acc[i];
}

// Solve, part 1: the top fronts f_t (all blocks, padded rows zero), columns
// [cbase, cbase + cw): F[a][c] = b[row(a)][c] + the children's update rows
// that add into row a, in a fixed order (tslot: maxp entries per row, -1 none).
__global__ __launch_bounds__(256) void k_top_gather(const int* __restrict__ trow, const int* __restrict__ tslot,
                                                    int maxp, int nrows, const double* __restrict__ B,
                                                    const double* __restrict__ U, double* __restrict__ F, int C,
                                                    int cbase, int cw, const int* __restrict__ any) {
  if (any && *any == 0) return;
  const long long idx = (long long)xcd_block() * 256 + threadIdx.x;
  if (idx >= (long long)nrows * cw) return;
  const int a = (int)(idx / cw), c = cbase + (int)(idx % cw);
  const int r = trow[a];
  double v = 0.0;
  if (r >= 0) {
    v = B[(long long)r * C + c];
    for (int k = 0; k < maxp; ++k) {
      const int q = tslot[(long long)a * maxp + k];
      if (q >= 0) v += U[(long long)q * C + c];
    }
  }
  F[(long long)a * C + c] = v;
}

// Solve, part 2: x_t = Z f_t, written to the top rows of B. Block = (tree
// block r, 32 output rows, 32 columns), 4 waves; wave q sums the q-th quarter
// of k (np_r / 4 each) for the whole 32 x 32 output as 2 x 2 MFMA tiles with
// operands from global memory (Z read as Z[k][a], its symmetric twin, so both
// operands are 16 consecutive doubles per k). Operands come in batches of
// kTopBatch k-steps, double-buffered: the loads of the next batch are in
// flight while the products of the current one run. The quarters are added
// through LDS in a fixed order, so the result does not depend on the launch
// (columns, chains).
constexpr int kTopBatch = 2;  // k-steps (4 k each) of operands per load batch
struct TopOps {
  double a[2][kTopBatch], b[2][kTopBatch];
};
// Loads of one full batch: byte offsets zo / fo of the batch's first k-step
// from uniform bases, advancing by zs / fs per k-step (32-bit offsets from a
// scalar base). Columns past the chain's end are read too (the buffers are
// padded so every such read stays inside them); they only feed output columns
// that are not written.
__device__ __forceinline__ double ldo(const char* __restrict__ base, unsigned off) {
  return *(const double*)(base + off);
}
__device__ __forceinline__ void top_load(TopOps& o, const char* __restrict__ zb, const char* __restrict__ fb,
                                         unsigned zo, unsigned fo, unsigned zs, unsigned fs) {
#pragma unroll
  for (int u = 0; u < kTopBatch; ++u) {
    o.a[0][u] = ldo(zb, zo + u * zs);
    o.a[1][u] = ldo(zb, zo + u * zs + 128);
    o.b[0][u] = ldo(fb, fo + u * fs);
    o.b[1][u] = ldo(fb, fo + u * fs + 128);
  }
}
__device__ __forceinline__ void top_mul(const TopOps& o, v4d (&acc)[2][2]) {
#pragma unroll
  for (int u = 0; u < kTopBatch; ++u) {
    acc[0][0] = mfma16(o.a[0][u], o.b[0][u], acc[0][0]);
    acc[0][1] = mfma16(o.a[0][u], o.b[1][u], acc[0][1]);
    acc[1][0] = mfma16(o.a[1][u], o.b[0][u], acc[1][0]);
    acc[1][1] = mfma16(o.a[1][u], o.b[1][u], acc[1][1]);
  }
}

// One wave's part of a dense product tile: acc[i][j] += sum over its nk
// k-steps (4 k each) of A(k, rows 16 i + lr) B(k, columns 16 j + lr), with A
// and B read at byte offsets zo0 / fo0 (its first k-step, lanes at their k
// row lk and column lr) advancing zs / fs per k-step. Batch j + 1's loads are
// issued before batch j's products; scheduling barriers keep the compiler
// from sinking loads down to their uses (which would serialise load latency
// and products again). The last even step reloads batch nb - 1 in place of
// one past the end.
__device__ __forceinline__ void dense_tile(v4d (&acc)[2][2], const char* __restrict__ zb, unsigned zo0, unsigned zs,
                                           const char* __restrict__ fb, unsigned fo0, unsigned fs, int nk) {
  const int nb = nk / kTopBatch;  // full batches, then single k-steps
  const unsigned zbs = kTopBatch * zs, fbs = kTopBatch * fs;
  if (nb > 0) {
    TopOps o0, o1;
    top_load(o0, zb, fb, zo0, fo0, zs, fs);
    int j = 0;
    for (; j + 2 <= nb; j += 2) {
      const unsigned j2 = (unsigned)min(j + 2, nb - 1);
      top_load(o1, zb, fb, zo0 + (unsigned)(j + 1) * zbs, fo0 + (unsigned)(j + 1) * fbs, zs, fs);
      __builtin_amdgcn_sched_barrier(0);
      top_mul(o0, acc);
      __builtin_amdgcn_sched_barrier(0);
      top_load(o0, zb, fb, zo0 + j2 * zbs, fo0 + j2 * fbs, zs, fs);
      __builtin_amdgcn_sched_barrier(0);
      top_mul(o1, acc);
      __builtin_amdgcn_sched_barrier(0);
    }
    if (j < nb) top_mul(o0, acc);  // odd count: the last batch is in o0
  }
  for (int t = nb * kTopBatch; t < nk; ++t) {
    const unsigned zo = zo0 + (unsigned)t * zs, fo = fo0 + (unsigned)t * fs;
    const double a0v = ldo(zb, zo), a1v = ldo(zb, zo + 128);
    const double b0v = ldo(fb, fo), b1v = ldo(fb, fo + 128);
    acc[0][0] = mfma16(a0v, b0v, acc[0][0]);
    acc[0][1] = mfma16(a0v, b1v, acc[0][1]);
    acc[1][0] = mfma16(a1v, b0v, acc[1][0]);
    acc[1][1] = mfma16(a1v, b1v, acc[1][1]);
  }
}

// The four waves' partial 32 x 32 tiles added in a fixed order through LDS;
// out(rr, cc, v) is called for the 1,024 outputs, four per thread.
template <typename Out>
__device__ __forceinline__ void dense_reduce(const v4d (&acc)[2][2], double (*red)[16 * 64], Out out) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int q = 0; q < 4; ++q) red[wave][((i * 2 + j) * 4 + q) * 64 + lane] = acc[i][j][q];
  __syncthreads();
  // output element e = 32 rr + cc: tile (rr / 16, cc / 16), accumulator
  // element (rr % 16) / 4 of lane 16 ((rr % 16) % 4) + cc % 16
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int e = tid + 256 * u;
    const int rr = e >> 5, cc = e & 31;
    const int rl = rr & 15;
    const int idx = (((rr >> 4) * 2 + (cc >> 4)) * 4 + (rl >> 2)) * 64 + (rl & 3) * 16 + (cc & 15);
    out(rr, cc, ((red[0][idx] + red[1][idx]) + red[2][idx]) + red[3][idx]);
  }
}

// Solve, part 2: x_t = Z f_t, written to the top rows of B. Block = (tree
// block r, 32 output rows, 32 columns), 4 waves; wave q sums the q-th quarter
// of k (np_r / 4 each) for the whole 32 x 32 output as 2 x 2 MFMA tiles with
// operands from global memory (Z read as Z[k][a], its symmetric twin, so both
// operands are 16 consecutive doubles per k). The quarters are added in a
// fixed order, so the result does not depend on the launch (columns, chains).
__global__ __launch_bounds__(256) void k_top_apply(const double* __restrict__ Z, const long long* __restrict__ zoff,
                                                   const int* __restrict__ a0_of, const int* __restrict__ np_of,
                                                   const int2* __restrict__ items, int ncolt,
                                                   const double* __restrict__ F, const int* __restrict__ trow,
                                                   double* __restrict__ B, int C, int cbase, int cw,
                                                   const int* __restrict__ any) {
  if (any && *any == 0) return;
  __shared__ double red[4][16 * 64];
  const int lb = xcd_block();
  const int2 it = items[lb / ncolt];
  const int r = it.x, row0 = 32 * it.y;
  const int c0 = cbase + 32 * (lb % ncolt);
  const int np = np_of[r], a0 = a0_of[r];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, lr = lane & 15, lk = lane >> 4;
  const int kq = np >> 2;  // np is a multiple of 32: quarters are multiples of 8
  const int kb = wave * kq;
  const int cend = cbase + cw;
  // per k-step 4 k: Z advances 4 rows of np, F 4 rows of C
  v4d acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = v4d{0.0, 0.0, 0.0, 0.0};
  dense_tile(acc, (const char*)(Z + zoff[r]), 8u * ((unsigned)(kb + lk) * (unsigned)np + (unsigned)(row0 + lr)),
             32u * (unsigned)np, (const char*)(F + (long long)a0 * C),
             8u * ((unsigned)(kb + lk) * (unsigned)C + (unsigned)(c0 + lr)), 32u * (unsigned)C, kq >> 2);
  dense_reduce(acc, red, [&](int rr, int cc, double v) {
    const int col = c0 + cc;
    const int er = trow[a0 + row0 + rr];
    if (er >= 0 && col < cend) B[(long long)er * C + col] = v;
  });
}

// ----------------------------------------------------------------------------
// Dense leaf subtrees. A leaf subtree (a subtree below the top separators
// whose supernodes own at most a capped number of rows, nsub, in the
// contiguous engine rows [j0, j0 + nsub)) touches only its own rows and the
// nab rows of its root's update block ("above" rows). With L_s its factor
// block, Linv = inv(L_s) and G = -L_{above,s} Linv (what its elimination does
// to the above rows), both sweeps over the subtree reduce to
//   forward:  z = K_s^{-1} b_s (K_s = L_s L_s^T), u = G b_s   (u: the root's
//             update block, for the parent's front);
//   backward: x_s = z + G^T x_above.
// Formed at factorisation like the top blocks: M = [I; 0] ((np + nabp) x np,
// the above rows after the padded own rows) through the subtree's elimination
// steps gives [Linv; G]; the Gram product gives Zs = Linv^T Linv; the forward
// operator is stored k-major as P[k][a] = [Zs | G^T] (np x (np + nabp)), the
// backward one is G itself (rows np .. np + nabp of M, k-major already).
struct LeafView {
  const double* P;           // forward operators (k-major)
  const double* M;           // [Linv; G] per leaf (the backward reads its G rows)
  const long long* poff;     // per leaf: offset of P
  const long long* moff;     // per leaf: offset of M
  const int* np_of;          // padded own rows (multiple of 32)
  const int* nabp_of;        // padded above rows (multiple of 32)
  const int* nsub_of;        // own rows
  const int* nab_of;         // above rows
  const int* j0_of;          // first own engine row
  const long long* uoff;     // per leaf: first update-block row of its root
  const long long* xaoff;    // per leaf: first row of its above rows in XA
};

// P[k][a] of one leaf: a < np: Zs[k][a]; a >= np: G[a - np][k] (= M[np + a - np][k]).
__global__ void k_leaf_pack(const double* __restrict__ Zs, const double* __restrict__ M, double* __restrict__ P,
                            int np, int nabp) {
  const int w = np + nabp;
  const long long e = blockIdx.x * 256LL + threadIdx.x;
  if (e >= (long long)np * w) return;
  const int k = (int)(e / w), a = (int)(e % w);
  P[e] = a < np ? Zs[(long long)k * np + a] : M[(long long)(np + a - np) * np + k];
}

// The above rows of every leaf, compacted: XA[xaoff + i][c] = B[rows[i]][c]
// (rows -1: padding, zero), columns [cbase, cbase + cw).
__global__ __launch_bounds__(256) void k_leaf_gather(const int* __restrict__ xrow, int nrows,
                                                     const double* __restrict__ B, double* __restrict__ XA, int C,
                                                     int cbase, int cw, const int* __restrict__ any) {
  if (any && *any == 0) return;
  const long long idx = (long long)xcd_block() * 256 + threadIdx.x;
  if (idx >= (long long)nrows * cw) return;
  const int a = (int)(idx / cw), c = cbase + (int)(idx % cw);
  const int r = xrow[a];
  XA[(long long)a * C + c] = r >= 0 ? B[(long long)r * C + c] : 0.0;
}

// Leaf products. FWD: out = P^T b (K = np; b rows j0 + k: rows past the own
// ones meet zero rows of P, and the right-hand-side buffers carry 32 zero rows
// past N for the last leaf); rows a < nsub -> Yb (z), np <= a < np + nab -> U
// at the root's update block (u). !FWD: out = z + G^T xa (K = nabp; xa from
// XA, zero padded), rows a < nsub -> B (x). Block = (leaf, 32 output rows,
// 32 columns), four waves each a quarter of k, quarters added in a fixed
// order (dense_tile, dense_reduce).
template <bool FWD>
__global__ __launch_bounds__(256) void k_leaf_apply(LeafView L, const int2* __restrict__ items, int ncolt,
                                                    const double* __restrict__ Bsrc, const double* __restrict__ XA,
                                                    double* __restrict__ Yb, double* __restrict__ U,
                                                    double* __restrict__ B, int C, int cbase, int cw,
                                                    const int* __restrict__ any) {
  if (any && *any == 0) return;
  __shared__ double red[4][16 * 64];
  const int lb = xcd_block();
  const int2 it = items[lb / ncolt];
  const int l = it.x, row0 = 32 * it.y;
  const int c0 = cbase + 32 * (lb % ncolt);
  const int np = L.np_of[l], nabp = L.nabp_of[l], nsub = L.nsub_of[l], j0 = L.j0_of[l];
  const int K = FWD ? np : nabp;
  const int lda = FWD ? np + nabp : np;
  const double* A = FWD ? L.P + L.poff[l] : L.M + L.moff[l] + (long long)np * np;
  const double* Bop = FWD ? Bsrc + (long long)j0 * C : XA + L.xaoff[l] * C;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, lr = lane & 15, lk = lane >> 4;
  const int kq = K >> 2;  // K is a multiple of 32
  const int kb = wave * kq;
  const int cend = cbase + cw;
  v4d acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = v4d{0.0, 0.0, 0.0, 0.0};
  dense_tile(acc, (const char*)A, 8u * ((unsigned)(kb + lk) * (unsigned)lda + (unsigned)(row0 + lr)),
             32u * (unsigned)lda, (const char*)Bop, 8u * ((unsigned)(kb + lk) * (unsigned)C + (unsigned)(c0 + lr)),
             32u * (unsigned)C, kq >> 2);
  const int nab = L.nab_of[l];
  dense_reduce(acc, red, [&](int rr, int cc, double v) {
    const int col = c0 + cc;
    if (col >= cend) return;
    const int a = row0 + rr;
    if (FWD) {
      if (a < nsub)
        Yb[(long long)(j0 + a) * C + col] = v;
      else if (a >= np && a - np < nab)
        U[(L.uoff[l] + (a - np)) * C + col] = v;
    } else if (a < nsub) {
      B[(long long)(j0 + a) * C + col] = Yb[(long long)(j0 + a) * C + col] + v;
    }
  });
}
