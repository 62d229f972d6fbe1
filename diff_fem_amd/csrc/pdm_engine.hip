// libpdm engine: device data, HIP kernels of the PD step / adjoint, and the
// C ABI declared in include/pdm.h.
//
// Layout in HBM. Vertices are in the engine's nested-dissection order. Every
// per-vertex state array is row-major [N][C] with C = 3R columns
// (column 3r+c = component c of rollout r), so rollouts are the contiguous
// dimension: a wavefront of the local step reads 64 rollouts of one tet with
// unit stride, and the supernodal solves see R*3 right-hand sides per row.
// Tet data is SoA ([k][T]); tet->vertex forces go through per-(element,corner)
// slots grouped by vertex ([slot][C]), summed per vertex in a fixed order
// (deterministic segmented reduction, no atomics).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/pdm.h"
#include "pdm_internal.h"
#include "pdm_kernels.h"
#include "pdm_factor_kernels.h"
#define PDM_HD __host__ __device__
#include "pdm_self_adj.h"

using namespace pdm;

namespace {

thread_local std::string g_err;

#define HIPX(expr)                                                                       \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess) {                                                              \
      g_err = std::string(#expr " failed: ") + hipGetErrorString(e_);                    \
      return 1;                                                                          \
    }                                                                                    \
  } while (0)

constexpr double kEps = 5e-7;  // World.h:214
constexpr int kThreads = 256;
constexpr int kMaxSplit = 4;  // column chains of the split solve
constexpr int kWslItem = 6144;  // doubles of W one grouped level-kernel block stages in LDS (48 KB)
// largest top-separator block (rows) per elimination tree solved as one dense
// product; 0 turns the dense blocks off
constexpr int kTopCapDefault = 1536;
// largest leaf subtree (rows) below the top separators solved as dense
// products; 0 turns the dense leaf blocks off
constexpr int kLeafCapDefault = 512;

inline int cdiv(long long a, long long b) { return (int)((a + b - 1) / b); }

// MI355X dispatches workgroups round-robin over its 8 XCDs (each with its
// own L2). Remap the hardware block index so that every XCD processes one
// contiguous range of logical blocks: neighbouring tets (sharing vertices) and
// the column tiles of one supernode then hit the same L2.
__device__ __forceinline__ int xcd_block() {
  const int nb = gridDim.x, b = blockIdx.x;
  const int q = nb >> 3, rem = nb & 7, k = b & 7;
  return k * q + min(k, rem) + (b >> 3);
}

// The same over the first nb blocks only (those that carry work when the
// work is packed at the front of a larger grid); -1 for the others. The busy
// blocks then still spread over all eight XCDs.
__device__ __forceinline__ int xcd_block_n(int nb) {
  const int b = blockIdx.x;
  if (b >= nb) return -1;
  const int q = nb >> 3, rem = nb & 7, k = b & 7;
  return k * q + min(k, rem) + (b >> 3);
}

// ------------------------------------------------------------------ kernels
__global__ void k_caller_to_engine(const double* __restrict__ src, double* __restrict__ dst,
                                   const int* __restrict__ perm, int N, int R) {
  long long C = 3LL * R;
  long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (idx >= (long long)N * C) return;
  int v = (int)(idx / C);
  int j = (int)(idx % C);
  int r = j / 3, c = j % 3;
  dst[idx] = src[(long long)r * 3 * N + 3LL * perm[v] + c];
}

__global__ void k_engine_to_caller(const double* __restrict__ src, double* __restrict__ dst,
                                   const int* __restrict__ perm, int N, int R) {
  long long C = 3LL * R;
  long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (idx >= (long long)N * C) return;
  int v = (int)(idx / C);
  int j = (int)(idx % C);
  int r = j / 3, c = j % 3;
  dst[(long long)r * 3 * N + 3LL * perm[v] + c] = src[idx];
}

// Y = A X for the scalar CSR A applied to every column.
__global__ void k_spmv(const int* __restrict__ ptr, const int* __restrict__ col, const double* __restrict__ val,
                       const double* __restrict__ X, double* __restrict__ Y, int N, int C) {
  long long idx = (long long)xcd_block() * blockDim.x + threadIdx.x;
  if (idx >= (long long)N * C) return;
  int v = (int)(idx / C);
  int j = (int)(idx % C);
  double s = 0.0;
  for (int q = ptr[v]; q < ptr[v + 1]; ++q) s += val[q] * X[(long long)col[q] * C + j];
  Y[idx] = s;
}

// Start of a step: gravity kick (Environment.h:236-245), qn = x + h v*
// (World.h:692-699) and the iteration-invariant rhs c = M v*/h - L x + W p.
// Underwater (FF non-null) the kick is v* = v + h M^-1 f with f the vertex's
// sum of its faces' drag forces in face order (Environment.h:246-250,
// World.h:692), and there is no gravity.
__global__ void k_step_prep(const double* __restrict__ X, const double* __restrict__ V, const double* __restrict__ Lx,
                            const double* __restrict__ mass, const double* __restrict__ attw,
                            const double* __restrict__ xrest, double* __restrict__ Xn, double* __restrict__ Cc,
                            double* __restrict__ Xit, int N, int C, double h, double g0, double g1, double g2,
                            int gravity_on, const double* __restrict__ FF, const int* __restrict__ vfptr,
                            const int* __restrict__ vfq) {
  long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (idx >= (long long)N * C) return;
  int v = (int)(idx / C);
  int j = (int)(idx - (long long)v * C);
  int c = j % 3;
  double vs;
  if (FF) {
    double f = 0.0;
    for (int q = vfptr[v]; q < vfptr[v + 1]; ++q) f += FF[(long long)(vfq[q] / 3) * C + j];
    vs = V[idx] + h * ((1.0 / mass[v]) * f);
  } else {
    double g = c == 0 ? g0 : (c == 1 ? g1 : g2);
    vs = V[idx] + (gravity_on ? g * h : 0.0);
  }
  double x = X[idx];
  Xn[idx] = x;
  Xit[idx] = x + h * vs;
  Cc[idx] = mass[v] * vs / h - Lx[idx] + attw[v] * xrest[3 * v + c];
}

// Drag force of contour face f for column triple r at the step-start state:
// FF[f][3r..3r+2] = drag/3 + lift/3 (Dfobj.h:874-994).
__global__ void k_drag_face(const double* __restrict__ X, const double* __restrict__ V, const int* __restrict__ cf,
                            double* __restrict__ FF, int nf, int R, int C) {
  long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (idx >= (long long)nf * R) return;
  int f = (int)(idx / R), r = (int)(idx - (long long)f * R);
  const long long o0 = (long long)cf[3 * f] * C + 3 * r, o1 = (long long)cf[3 * f + 1] * C + 3 * r,
                  o2 = (long long)cf[3 * f + 2] * C + 3 * r;
  double a[3], b[3], n[3], w[3], F[3];
  for (int c = 0; c < 3; ++c) {
    a[c] = X[o1 + c] - X[o0 + c];
    b[c] = X[o2 + c] - X[o0 + c];
    w[c] = 0.0 - (V[o0 + c] + V[o1 + c] + V[o2 + c]) / 3.0;
  }
  n[0] = a[1] * b[2] - a[2] * b[1];
  n[1] = a[2] * b[0] - a[0] * b[2];
  n[2] = a[0] * b[1] - a[1] * b[0];
  drag_force(n, w, F);
  for (int c = 0; c < 3; ++c) FF[(long long)f * C + 3 * r + c] = F[c];
}

// Reverse of the drag kick for face f, column triple r: gF = sum over the
// face's vertices of h m^-1 gVs; then, through n = (p2-p1) x (p3-p1) and
// w = -(v1+v2+v3)/3, per-corner position gradients GXf[3f+k] and the shared
// velocity gradient GWf[f] (= -gw/3).
__global__ void k_drag_vjp(const double* __restrict__ X, const double* __restrict__ V, const double* __restrict__ gVs,
                           const double* __restrict__ mass, const int* __restrict__ cf, double* __restrict__ GXf,
                           double* __restrict__ GWf, int nf, int R, int C, double h) {
  long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (idx >= (long long)nf * R) return;
  int f = (int)(idx / R), r = (int)(idx - (long long)f * R);
  int vi[3] = {cf[3 * f], cf[3 * f + 1], cf[3 * f + 2]};
  long long o[3];
  for (int k = 0; k < 3; ++k) o[k] = (long long)vi[k] * C + 3 * r;
  double a[3], b[3], n[3], w[3], gF[3] = {0, 0, 0};
  for (int c = 0; c < 3; ++c) {
    a[c] = X[o[1] + c] - X[o[0] + c];
    b[c] = X[o[2] + c] - X[o[0] + c];
    w[c] = 0.0 - (V[o[0] + c] + V[o[1] + c] + V[o[2] + c]) / 3.0;
    for (int k = 0; k < 3; ++k) gF[c] += h * (1.0 / mass[vi[k]]) * gVs[o[k] + c];
  }
  n[0] = a[1] * b[2] - a[2] * b[1];
  n[1] = a[2] * b[0] - a[0] * b[2];
  n[2] = a[0] * b[1] - a[1] * b[0];
  double gn[3], gw[3];
  drag_force_vjp(n, w, gF, gn, gw);
  // n = a x b: ga = b x gn, gb = gn x a
  double ga[3] = {b[1] * gn[2] - b[2] * gn[1], b[2] * gn[0] - b[0] * gn[2], b[0] * gn[1] - b[1] * gn[0]};
  double gb[3] = {gn[1] * a[2] - gn[2] * a[1], gn[2] * a[0] - gn[0] * a[2], gn[0] * a[1] - gn[1] * a[0]};
  const long long base = (long long)3 * f * C + 3 * r;
  for (int c = 0; c < 3; ++c) {
    GXf[base + c] = -(ga[c] + gb[c]);
    GXf[base + C + c] = ga[c];
    GXf[base + 2 * C + c] = gb[c];
    GWf[(long long)f * C + 3 * r + c] = -gw[c] / 3.0;
  }
}

// gX += sum of the vertex's corner position gradients, gV += sum of its faces'
// velocity gradients (ascending face order).
__global__ void k_drag_gather(const double* __restrict__ GXf, const double* __restrict__ GWf,
                              const int* __restrict__ vfptr, const int* __restrict__ vfq, double* __restrict__ gX,
                              double* __restrict__ gV, int N, int C) {
  long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (idx >= (long long)N * C) return;
  int v = (int)(idx / C);
  int j = (int)(idx - (long long)v * C);
  double sx = 0.0, sv = 0.0;
  for (int q = vfptr[v]; q < vfptr[v + 1]; ++q) {
    sx += GXf[(long long)vfq[q] * C + j];
    sv += GWf[(long long)(vfq[q] / 3) * C + j];
  }
  gX[idx] += sx;
  gV[idx] += sv;
}

// Active-rollout list: alist[0..*acount) = the rollouts r with active[r] != 0,
// ascending. One block of blockDim.x (a multiple of 64, at most 1024) threads.
__global__ __launch_bounds__(1024) void k_compact(const int* __restrict__ active, int R, int* __restrict__ alist,
                                                  int* __restrict__ acount) {
  // per blockDim.x rollouts: a wave ballot gives each listed rollout its
  // rank in its wave, the waves' counts (in wave order) its offset
  __shared__ int wsum[16];
  __shared__ int base;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  if (threadIdx.x == 0) base = 0;
  __syncthreads();
  for (int c0 = 0; c0 < R; c0 += blockDim.x) {
    const int i = c0 + threadIdx.x;
    const bool f = i < R && active[i] != 0;
    const unsigned long long m = __ballot(f);
    if (lane == 0) wsum[wave] = __popcll(m);
    __syncthreads();
    int off = base + __popcll(m & ((1ull << lane) - 1ull));
    for (int w = 0; w < wave; ++w) off += wsum[w];
    if (f) alist[off] = i;
    __syncthreads();
    if (threadIdx.x == 0) {
      int t = 0;
      for (int w = 0; w < nw; ++w) t += wsum[w];
      base += t;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) *acount = base;
}

// The list's launch: as many threads as rollouts (rounded up to a wave), at
// most 1024.
inline void launch_compact(const int* active, int R, int* alist, int* acount, hipStream_t st) {
  const int threads = std::min(1024, std::max(64, (R + 63) / 64 * 64));
  hipLaunchKernelGGL(k_compact, dim3(1), dim3(threads), 0, st, active, R, alist, acount);
}

// Active-rollout indexing of the element kernels. Work items are (tet t,
// listed rollout j < A = *acount). Two variants share the work:
//   wave-uniform (R a multiple of 64, A >= kDenseBelow): item = t P + j with
//     P = A rounded up to 64, so the 64 lanes of a wave share one tet; its
//     index is made wave-uniform and the per-tet data come in by scalar loads
//     (lanes j >= A idle);
//   dense (A < kDenseBelow when R is a multiple of 64, every A otherwise):
//     item = t A + j, lanes span several tets. Per item it costs about three
//     times the wave-uniform variant, so it pays only while most lanes of the
//     wave-uniform one would idle.
// A launch of the variant that does not apply exits after one load. On
// wave-uniform scenes the element kernels walk tet strips (k_local_strip)
// or tets (k_vjp_tet<., true>) with their items numbered per strip or tet.
constexpr int kDenseBelow = 24;

// Wave-uniform item of this thread: (unit u = tet or strip, listed rollout
// j); false for idle lanes and below kDenseBelow listed rollouts. 32-bit: the
// host keeps T R below 2^31.
__device__ __forceinline__ bool uniform_item(const int* __restrict__ acount, int Un, int* u, int* j) {
  const int A = *acount;
  if (A < kDenseBelow) return false;
  const unsigned P = (unsigned)(A + 63) & ~63u;
  const unsigned n = (unsigned)Un * P;
  const int b = xcd_block_n((int)((n + blockDim.x - 1) / blockDim.x));
  if (b < 0) return false;
  const unsigned idx = (unsigned)b * blockDim.x + threadIdx.x;
  if (idx >= n) return false;
  *u = __builtin_amdgcn_readfirstlane((int)(idx / P));
  *j = (int)(idx - (unsigned)*u * P);
  return *j < A;
}

// Dense item of virtual block vb (a grid smaller than the work loops over
// virtual blocks vb = blockIdx.x + m gridDim.x; gridDim.x is a multiple of 8 or
// covers the work, so vb stays on the hardware block's XCD): 1 with (t, j),
// 0 for a thread past the work in a block that has some, -1 past the work or
// when the dense variant does not apply.
__device__ __forceinline__ int dense_item(const int* __restrict__ acount, int Tn, int R, int vb, int* t, int* j) {
  const int A = *acount;
  if ((R % 64 == 0 && A >= kDenseBelow) || A == 0) return -1;
  const unsigned n = (unsigned)Tn * (unsigned)A;
  const int nb = (int)((n + blockDim.x - 1) / blockDim.x);
  if (vb >= nb) return -1;
  const int q = nb >> 3, rem = nb & 7, k = vb & 7;
  const int b = k * q + min(k, rem) + (vb >> 3);
  const unsigned idx = (unsigned)b * blockDim.x + threadIdx.x;
  if (idx >= n) return 0;
  *t = (int)(idx / (unsigned)A);
  *j = (int)(idx - (unsigned)*t * (unsigned)A);
  return 1;
}

// Blocks of a dense element launch: on wave-uniform scenes the dense variant
// runs only below kDenseBelow listed rollouts and its launch is mostly empty,
// so the grid is capped (a multiple of 8) and the blocks loop.
constexpr int kDenseGridCap = 1024;
inline int dense_blocks(long long Tn, int R) {
  const long long nb = R % 64 == 0 ? (Tn * (kDenseBelow - 1) + 255) / 256 : (Tn * R + 255) / 256;
  return R % 64 == 0 ? (int)std::min<long long>(nb, kDenseGridCap) : (int)nb;
}

// The wave-uniform item of this thread with the wave's idle lanes kept
// (*live = false, j clamped to 0): false only for waves without work.
__device__ __forceinline__ bool element_wave(const int* __restrict__ acount, int Tn, int* t, int* j, bool* live) {
  const int A = *acount;
  if (A < kDenseBelow) return false;
  const unsigned P = (unsigned)(A + 63) & ~63u;
  const unsigned n = (unsigned)Tn * P;
  const int b = xcd_block_n((int)((n + blockDim.x - 1) / blockDim.x));
  if (b < 0) return false;
  const unsigned idx = (unsigned)b * blockDim.x + threadIdx.x;
  if (idx >= n) return false;  // n is a multiple of 64: whole waves
  *t = __builtin_amdgcn_readfirstlane((int)(idx / P));
  const int jj = (int)(idx - (unsigned)*t * P);
  *live = jj < A;
  *j = *live ? jj : jj & ~63;  // idle lanes take their wave's first (listed) item
  return true;
}

// Corner forces of tet t for rollout r (the local step's projection P, then
// f_k = P g_k, plus the tet's muscles), in the tet's corner order.
template <typename T>
__device__ __forceinline__ void tet_forces(const double* __restrict__ X, const int* __restrict__ tv,
                                           const double* __restrict__ tG, const double* __restrict__ wr,
                                           const double* __restrict__ wv, int t, int Tn, int r, int C,
                                           const int* __restrict__ mptr, const double* __restrict__ mg,
                                           const double* __restrict__ mw, const int* __restrict__ mgroup, int Mn,
                                           const double* __restrict__ act, int act_stride_r, int act_off, T* f) {
  int v0 = tv[t], v1 = tv[Tn + t], v2 = tv[2 * Tn + t], v3 = tv[3 * Tn + t];
  const double* x0 = X + (long long)v0 * C + 3 * r;
  const double* x1 = X + (long long)v1 * C + 3 * r;
  const double* x2 = X + (long long)v2 * C + 3 * r;
  const double* x3 = X + (long long)v3 * C + 3 * r;
  double b0 = x0[0], b1 = x0[1], b2 = x0[2];
  // Ds = [x1-x0, x2-x0, x3-x0] in FP64, then F = Ds invDm in T
  T ds[9];
  ds[0] = (T)(x1[0] - b0); ds[1] = (T)(x2[0] - b0); ds[2] = (T)(x3[0] - b0);
  ds[3] = (T)(x1[1] - b1); ds[4] = (T)(x2[1] - b1); ds[5] = (T)(x3[1] - b1);
  ds[6] = (T)(x1[2] - b2); ds[7] = (T)(x2[2] - b2); ds[8] = (T)(x3[2] - b2);
  T G[12];
#pragma unroll
  for (int q = 0; q < 12; ++q) G[q] = (T)tG[(long long)q * Tn + t];
  T F[9];
#pragma unroll
  for (int rr = 0; rr < 3; ++rr)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      F[rr * 3 + j] = ds[rr * 3 + 0] * G[3 + j] + ds[rr * 3 + 1] * G[6 + j] + ds[rr * 3 + 2] * G[9 + j];
  T P[9];
  corotate_P<T>(F, (T)wr[t], (T)wv[t], P);
#pragma unroll
  for (int k = 0; k < 4; ++k)
#pragma unroll
    for (int rr = 0; rr < 3; ++rr)
      f[3 * k + rr] = P[rr * 3 + 0] * G[3 * k + 0] + P[rr * 3 + 1] * G[3 * k + 1] + P[rr * 3 + 2] * G[3 * k + 2];
  // Linear muscles on this tet: d = (1 - a) F f (LinearMuscleConstraint.h:265-271),
  // corner force w (1 - a) g_k (sum_j g_j x_j), g in this tet's corner order.
  for (int m = mptr[t], me = mptr[t + 1]; m < me; ++m) {
    double a = act ? act[(long long)r * act_stride_r + act_off + mgroup[m]] : 0.0;
    T sc = (T)(mw[m] * (1.0 - a));
    T g[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) g[k] = (T)mg[(long long)k * Mn + m];
    T ff[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) ff[c] = g[1] * ds[c * 3 + 0] + g[2] * ds[c * 3 + 1] + g[3] * ds[c * 3 + 2];
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
      for (int c = 0; c < 3; ++c) f[3 * k + c] += sc * ff[c] * g[k];
  }
}

// Dense element variant: thread (tet, listed rollout), corner forces into the
// tet's corner slots. LOOP: a capped grid whose blocks loop over the virtual
// blocks (wave-uniform scenes); otherwise one item per thread.
template <typename T, bool LOOP>
__global__ __launch_bounds__(256) void k_local_tet(const double* __restrict__ X, const int* __restrict__ tv,
                                                   const double* __restrict__ tG, const double* __restrict__ wr,
                                                   const double* __restrict__ wv, const int* __restrict__ tslot,
                                                   const int* __restrict__ alist, const int* __restrict__ acount,
                                                   T* __restrict__ corner, int Tn,
                                                   int R, const int* __restrict__ mptr, const double* __restrict__ mg,
                                                   const double* __restrict__ mw, const int* __restrict__ mgroup,
                                                   int Mn, const double* __restrict__ act, int act_stride_r,
                                                   int act_off, const int* __restrict__ step_dev, int ngroups) {
  if (step_dev) act_off = *step_dev * ngroups;
  const int C = 3 * R;
  for (int vb = blockIdx.x;; vb += gridDim.x) {
    int t, j;
    const int st = dense_item(acount, Tn, R, vb, &t, &j);
    if (st < 0 || (!LOOP && st == 0)) return;
    if (st == 0) continue;
    const int r = alist[j];
    T f[12];
    tet_forces<T>(X, tv, tG, wr, wv, t, Tn, r, C, mptr, mg, mw, mgroup, Mn, act, act_stride_r, act_off, f);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      T* o = corner + (long long)tslot[k * Tn + t] * C + 3 * r;
#pragma unroll
      for (int rr = 0; rr < 3; ++rr) o[rr] = f[3 * k + rr];
    }
    if (!LOOP) return;
  }
}

// a += f, never contracted with the product that formed f: the dense variant
// stores f and k_strip_fold adds the stored value, so both must add the same
// rounded f.
template <typename T>
__device__ __forceinline__ void strip_add(T* a, const T* f) {
#pragma clang fp contract(off)
  a[0] = a[0] + f[0];
  a[1] = a[1] + f[1];
  a[2] = a[2] + f[2];
}

// Wave-uniform element variant (kDenseBelow or more listed rollouts, R a
// multiple of 64): wave (strip, 64 listed rollouts) walks the strip's tets and
// adds each corner force into the register slot its code names (uniform, so
// the slot choice is a scalar branch); after each tet the slots of its flush
// mask are written as strip sums and cleared.
template <typename T>
__global__ __launch_bounds__(256) void k_local_strip(const double* __restrict__ X, const int* __restrict__ tv,
                                                     const double* __restrict__ tG, const double* __restrict__ wr,
                                                     const double* __restrict__ wv, const int* __restrict__ stet,
                                                     const int* __restrict__ scode, const int* __restrict__ sfbase,
                                                     const int* __restrict__ ssum_of, const int* __restrict__ alist,
                                                     const int* __restrict__ acount, T* __restrict__ ssum, int Tn,
                                                     int nstrip, int R, const int* __restrict__ mptr,
                                                     const double* __restrict__ mg, const double* __restrict__ mw,
                                                     const int* __restrict__ mgroup, int Mn,
                                                     const double* __restrict__ act, int act_stride_r, int act_off,
                                                     const int* __restrict__ step_dev, int ngroups) {
  int s, j;
  if (!uniform_item(acount, nstrip, &s, &j)) return;
  const int r = alist[j];
  if (step_dev) act_off = *step_dev * ngroups;
  const int C = 3 * R;
  T acc[kStripSlots][3];
#pragma unroll
  for (int q = 0; q < kStripSlots; ++q) acc[q][0] = acc[q][1] = acc[q][2] = (T)0;
  int fl = sfbase[s];
  for (int i = 0; i < kStripLen; ++i) {
    const int t = __builtin_amdgcn_readfirstlane(stet[s * kStripLen + i]);
    if (t < 0) break;
    const int code = __builtin_amdgcn_readfirstlane(scode[s * kStripLen + i]);
    T f[12];
    tet_forces<T>(X, tv, tG, wr, wv, t, Tn, r, C, mptr, mg, mw, mgroup, Mn, act, act_stride_r, act_off, f);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int sl = (code >> (3 * k)) & 7;
      // the empty asm keeps each add behind its (uniform) branch: speculated,
      // the compiler would add every corner into every slot and select
#pragma unroll
      for (int q = 0; q < kStripSlots; ++q)
        if (sl == q) {
          asm volatile("");
          strip_add<T>(acc[q], f + 3 * k);
        }
    }
#pragma unroll
    for (int q = 0; q < kStripSlots; ++q)
      if ((code >> (12 + q)) & 1) {
        const int p = __builtin_amdgcn_readfirstlane(ssum_of[fl++]);
        T* o = ssum + (long long)p * C + 3 * r;
        o[0] = acc[q][0];
        o[1] = acc[q][1];
        o[2] = acc[q][2];
        acc[q][0] = acc[q][1] = acc[q][2] = (T)0;
      }
  }
}

// Strip sums from the dense variant's corner slots, each the sum of its
// corner slots in the strip kernel's order, starting from zero as the strip
// kernel's slots do. Runs exactly when the dense variant does, on a
// wave-uniform scene.
template <typename T>
__global__ __launch_bounds__(256) void k_strip_fold(const T* __restrict__ corner, const int* __restrict__ sfptr,
                                                    const int* __restrict__ sfq, const int* __restrict__ alist,
                                                    const int* __restrict__ acount, T* __restrict__ ssum, int nsum,
                                                    int R) {
  const int A = *acount;
  if (A >= kDenseBelow || A == 0) return;
  const long long n = (long long)nsum * 3 * A;
  const long long C = 3LL * R;
  for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < n;
       idx += (long long)gridDim.x * blockDim.x) {
    const int p = (int)(idx / (3 * A));
    const int jj = (int)(idx - (long long)p * 3 * A);
    const int r3 = jj / 3;
    const int col = 3 * alist[r3] + (jj - 3 * r3);
    T a = (T)0;
    for (int k = sfptr[p], ke = sfptr[p + 1]; k < ke; ++k) a = a + corner[(long long)sfq[k] * C + col];
    ssum[(long long)p * C + col] = a;
  }
}

// Segmented gather over the listed rollouts only: thread (vertex, jj) sums
// column 3 alist[jj/3] + jj%3 over the vertex's slot range (four partial sums,
// fixed order), so that a wave reads contiguous slot entries. MODE 0:
// B = Cc + sum. MODE 1: B = sum. Columns of unlisted rollouts are left as
// they are. UNI: 3R is a multiple of 64 (one vertex per
// wave, its slot range by scalar loads).
template <typename T, int MODE, bool UNI>
__global__ __launch_bounds__(256) void k_gather_list(const double* __restrict__ Cc, const T* __restrict__ corner,
                                                     const int* __restrict__ vptr, const int* __restrict__ alist,
                                                     const int* __restrict__ acount, double* __restrict__ B, int N,
                                                     int R, const int* __restrict__ any = nullptr) {
  if (any && *any == 0) return;
  const long long C = 3LL * R;
  long long idx = (long long)xcd_block() * blockDim.x + threadIdx.x;
  if (idx >= (long long)N * C) return;
  int v = (int)(idx / C);
  const int jj = (int)(idx - (long long)v * C);
  if (UNI) v = __builtin_amdgcn_readfirstlane(v);
  if (jj >= 3 * *acount) return;
  const int r3 = jj / 3;
  const int col = 3 * alist[r3] + (jj - 3 * r3);
  const long long o = (long long)v * C + col;
  const double c0 = MODE == 0 ? Cc[o] : 0.0;
  double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
  int q = vptr[v], qe = vptr[v + 1];
  const T* p = corner + (long long)q * C + col;
  // slot k of the range goes to partial sum k mod 4; eight loads in flight
  for (; q + 8 <= qe; q += 8, p += 8 * C) {
    double x[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) x[u] = (double)p[u * C];
    s0 += x[0];
    s1 += x[1];
    s2 += x[2];
    s3 += x[3];
    s0 += x[4];
    s1 += x[5];
    s2 += x[6];
    s3 += x[7];
  }
  if (q + 4 <= qe) {
    s0 += (double)p[0];
    s1 += (double)p[C];
    s2 += (double)p[2 * C];
    s3 += (double)p[3 * C];
    q += 4;
    p += 4 * C;
  }
  for (; q < qe; ++q, p += C) s0 += (double)p[0];
  const double sum = (s0 + s1) + (s2 + s3);
  B[o] = MODE == 0 ? c0 + sum : sum;
}

// Per supernode s (w columns, m rows, the first w its own) the sweeps use
//   W = [inv(L_ss); -L_below inv(L_ss)]   (m x w, top block lower triangular)
// so that each sweep is one dense product per supernode:
//   forward:  [y_s; U_s] = [0; T_below] + W T_top
//   backward: x_s = W^T [y_s; x_below]
struct FactorView {
  const double* W;      // column-major m x w at lptr[s] (forward sweep reads it down columns)
  const double* WR;     // the same, row-major (backward sweep reads it along rows)
  const int* first;
  const long long* rowptr;
  const int* rows;
  const long long* lptr;
  const int* child_ptr;
  const int* child;
  const long long* relptr;
  const int* rel;
  const long long* uptr;
};

// FP64 MFMA 16x16x4 (gfx950): lane l supplies A[l%16][l/16] and B[l/16][l%16];
// accumulator element i of lane l is D[4i + l/16][l%16].
typedef double v4d __attribute__((ext_vector_type(4)));
__device__ __forceinline__ v4d mfma16(double a, double b, v4d c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// Element loops of the level kernels. Ct is a power of two and the block size
// a multiple of it, so thread tid always handles column c = tid & (Ct-1) and
// rows tid >> lg, tid >> lg + rs, ... (rs = blockDim / Ct): no divisions.
struct TileIdx {
  int c, r0, rs;
};
__device__ __forceinline__ TileIdx tile_idx(int tid, int nt, int lg) {
  return TileIdx{tid & ((1 << lg) - 1), tid >> lg, nt >> lg};
}

// Front load: T[row][c] = src(row, c) for row < nvalid, c < ct; zero for the
// other rows < nrows. rows == nullptr: source row c0 + row; else rows[row].
// Eight loads per thread are issued before they are stored.
__device__ __forceinline__ void load_front(const double* __restrict__ B, const int* __restrict__ rows, int c0,
                                           int nvalid, int nrows, int lg, int ct, int C, int col0,
                                           double* __restrict__ T, TileIdx ti) {
  const int Ct = 1 << lg;
  const bool cok = ti.c < ct;
  const double* Bc = B + col0 + ti.c;
  for (int rb = ti.r0; rb < nrows; rb += 8 * ti.rs) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      int row = rb + u * ti.rs;
      bool ok = cok && row < nvalid;
      long long src = rows ? (ok ? rows[row] : 0) : (long long)(c0 + row);
      v[u] = ok ? Bc[src * C] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      int row = rb + u * ti.rs;
      if (row < nrows) T[row * Ct + ti.c] = v[u];
    }
  }
}

// acc += A(16 x K) * Bm(K x 16) over k in [kbeg, kend) in steps of 4, with
// operands fetched by the callables (lane-level element getters). Operands of
// eight MFMAs are fetched before they are consumed, so that eight loads per
// wave are in flight; the tail (fewer than 32 k) is masked.
template <typename FA, typename FB>
__device__ __forceinline__ v4d mfma_chain(v4d acc, int kbeg, int kend, int lk, FA fa, FB fb) {
  int k = kbeg;
  for (; k + 32 <= kend; k += 32) {
    double a[8], b[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      a[u] = fa(k + 4 * u + lk);
      b[u] = fb(k + 4 * u + lk);
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) acc = mfma16(a[u], b[u], acc);
  }
  if (k < kend) {
    const int nrem = (kend - k + 3) >> 2;
    double a[8], b[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      int kk = k + 4 * u + lk;
      bool ok = u < nrem && kk < kend;
      a[u] = ok ? fa(kk) : 0.0;
      b[u] = ok ? fb(kk) : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (u < nrem) acc = mfma16(a[u], b[u], acc);
  }
  return acc;
}

// Forward substitution (L y = b) of one elimination-tree level. Block =
// (supernode, column tile of Ct = 16*nq right-hand sides), 4 waves. The front
// T (m x Ct) is staged in LDS: own rows from B, children's update blocks added
// in child order. Then one dense product on the matrix cores, split into
// 16x16 output tiles distributed over the waves:
//   y = inv(L_ss) T[0:w]                       (top tiles, into Y)
//   T[w:m] += (-L_below inv(L_ss)) T[0:w]      (update block U_s for the parent)
// Both read only T[0:w] and write disjoint rows, so no barrier separates them.
__device__ __forceinline__ void fwd_supernode(const FactorView& f, int s, int col0, int ct, int lg, int C,
                                              double* __restrict__ B, double* __restrict__ U, double* __restrict__ sh) {
  const int Ct = 1 << lg;
  const int c0 = f.first[s];
  const int w = f.first[s + 1] - c0;
  const long long rp = f.rowptr[s];
  const int m = (int)(f.rowptr[s + 1] - rp);
  const int mpad = (m + 15) & ~15;
  double* Y = sh + mpad * Ct;  // wpad x Ct
  const double* Ws = f.W + f.lptr[s];
  const int tid = threadIdx.x, nt = blockDim.x;
  const int lane = tid & 63, wave = tid >> 6, nwave = nt >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  const int nq = Ct >> 4;
  const TileIdx ti = tile_idx(tid, nt, lg);
  load_front(B, nullptr, c0, w, mpad, lg, ct, C, col0, sh, ti);
  __syncthreads();
  for (int q = f.child_ptr[s]; q < f.child_ptr[s + 1]; ++q) {
    int ch = f.child[q];
    int wc = f.first[ch + 1] - f.first[ch];
    int mbc = (int)(f.rowptr[ch + 1] - f.rowptr[ch]) - wc;
    const int* rel = f.rel + f.relptr[ch];
    const double* Uc = U + f.uptr[ch] * C + col0 + ti.c;
    if (ti.c < ct) {
      for (int ab = ti.r0; ab < mbc; ab += 8 * ti.rs) {
        double v[8];
        int dst[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          int a = ab + u * ti.rs;
          v[u] = a < mbc ? Uc[(long long)a * C] : 0.0;
          dst[u] = a < mbc ? rel[a] : -1;
        }
#pragma unroll
        for (int u = 0; u < 8; ++u)
          if (dst[u] >= 0) sh[dst[u] * Ct + ti.c] += v[u];
      }
    }
    __syncthreads();
  }
  const int wpad = (w + 15) & ~15;
  const int nrt = wpad >> 4;
  const int mb = m - w;
  const int nbt = (mb + 15) >> 4;
  for (int tq = wave; tq < (nrt + nbt) * nq; tq += nwave) {
    const int tr = tq / nq, q = tq % nq;
    const double* pb = sh + q * 16 + lr;
    if (tr < nrt) {
      // y rows (inv(L_ss) is lower triangular: k <= row)
      const int rt = tr * 16;
      const int row = rt + lr;
      const double* pa = Ws + row;
      v4d acc = {0.0, 0.0, 0.0, 0.0};
      acc = mfma_chain(
          acc, 0, min(w, rt + 16), lk, [&](int k) { return row < w ? pa[(long long)k * m] : 0.0; },
          [&](int k) { return pb[k * Ct]; });
#pragma unroll
      for (int i = 0; i < 4; ++i) Y[(rt + 4 * i + lk) * Ct + q * 16 + lr] = acc[i];
    } else {
      // update rows (row tiles start at w)
      const int rt = w + (tr - nrt) * 16;
      const int row = rt + lr;
      const double* pa = Ws + row;
      v4d acc;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        int r = rt + 4 * i + lk;
        acc[i] = r < m ? sh[r * Ct + q * 16 + lr] : 0.0;
      }
      acc = mfma_chain(
          acc, 0, w, lk, [&](int k) { return row < m ? pa[(long long)k * m] : 0.0; },
          [&](int k) { return pb[k * Ct]; });
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        int r = rt + 4 * i + lk;
        if (r < m) sh[r * Ct + q * 16 + lr] = acc[i];
      }
    }
  }
  __syncthreads();
  if (ti.c < ct) {
    double* Bo = B + (long long)c0 * C + col0 + ti.c;
    for (int r = ti.r0; r < w; r += ti.rs) Bo[(long long)r * C] = Y[r * Ct + ti.c];
    if (mb > 0) {
      double* Us = U + f.uptr[s] * C + col0 + ti.c;
      for (int a = ti.r0; a < mb; a += ti.rs) Us[(long long)a * C] = sh[(w + a) * Ct + ti.c];
    }
  }
}

__global__ __launch_bounds__(1024) void k_fwd_level(FactorView f, const int* __restrict__ lev, int ntile, int lg, int C,
                                                   double* __restrict__ B, double* __restrict__ U,
                                                   const int* __restrict__ any, int cbase, int cw) {
  if (any && *any == 0) return;
  extern __shared__ double sh[];
  const int Ct = 1 << lg;
  const int lb = xcd_block();
  const int col0 = cbase + (lb % ntile) * Ct;
  fwd_supernode(f, lev[lb / ntile], col0, min(Ct, cbase + cw - col0), lg, C, B, U, sh);
}

// Forward substitution of whole subtrees (the lower levels of the tree): block
// = (subtree, column tile); the block walks the subtree's supernodes in
// postorder (a contiguous index range, children first), so no kernel boundary
// separates a child from its parent. Update blocks go through global memory
// and are read back by the same block.
__global__ __launch_bounds__(1024) void k_fwd_subtree(FactorView f, const int2* __restrict__ sub, int ntile, int lg,
                                                     int C, double* __restrict__ B, double* __restrict__ U,
                                                     const int* __restrict__ any, int cbase, int cw,
                                                     const char* __restrict__ skip) {
  if (any && *any == 0) return;
  extern __shared__ double sh[];
  const int Ct = 1 << lg;
  const int lb = xcd_block();
  const int2 rg = sub[lb / ntile];
  const int col0 = cbase + (lb % ntile) * Ct;
  const int ct = min(Ct, cbase + cw - col0);
  for (int s = rg.x; s <= rg.y; ++s) {
    if (skip && skip[s]) continue;  // a dense leaf (its root's update block is already in U)
    fwd_supernode(f, s, col0, ct, lg, C, B, U, sh);
    __syncthreads();
  }
}

// Backward substitution (L^T x = y) of one level as one dense product:
//   x_s = inv(L_ss)^T y_s + (-L_below inv(L_ss))^T x_below = W^T T
// with T = [y_s; x_below] (rows below are final). W's top block is lower
// triangular, so output row i needs k >= i only.
__device__ __forceinline__ void bwd_supernode(const FactorView& f, int s, int col0, int ct, int lg, int C,
                                              double* __restrict__ B, double* __restrict__ sh) {
  const int Ct = 1 << lg;
  const int c0 = f.first[s];
  const int w = f.first[s + 1] - c0;
  const long long rp = f.rowptr[s];
  const int m = (int)(f.rowptr[s + 1] - rp);
  const int mpad = (m + 15) & ~15;
  const int wpad = (w + 15) & ~15;
  double* Tt = sh + mpad * Ct;  // wpad x Ct
  const int* rows = f.rows + rp;
  const double* WRs = f.WR + f.lptr[s];
  const int tid = threadIdx.x, nt = blockDim.x;
  const int lane = tid & 63, wave = tid >> 6, nwave = nt >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  const int nq = Ct >> 4;
  const TileIdx ti = tile_idx(tid, nt, lg);
  load_front(B, rows, 0, m, mpad, lg, ct, C, col0, sh, ti);
  __syncthreads();
  const int nrt = wpad >> 4;
  for (int tq = wave; tq < nrt * nq; tq += nwave) {
    const int rt = (tq / nq) * 16, q = tq % nq;
    const int col = rt + lr;  // column of W (output row of x_s)
    const double* pa = WRs + col;
    const double* pb = sh + q * 16 + lr;
    v4d acc = {0.0, 0.0, 0.0, 0.0};
    acc = mfma_chain(
        acc, rt, m, lk, [&](int k) { return col < w ? pa[(long long)k * w] : 0.0; },
        [&](int k) { return pb[k * Ct]; });
#pragma unroll
    for (int i = 0; i < 4; ++i) Tt[(rt + 4 * i + lk) * Ct + q * 16 + lr] = acc[i];
  }
  __syncthreads();
  if (ti.c < ct) {
    double* Bo = B + (long long)c0 * C + col0 + ti.c;
    for (int r = ti.r0; r < w; r += ti.rs) Bo[(long long)r * C] = Tt[r * Ct + ti.c];
  }
}

__global__ __launch_bounds__(1024) void k_bwd_level(FactorView f, const int* __restrict__ lev, int ntile, int lg, int C,
                                                   double* __restrict__ B, const int* __restrict__ any, int cbase,
                                                   int cw) {
  if (any && *any == 0) return;
  extern __shared__ double sh[];
  const int Ct = 1 << lg;
  const int lb = xcd_block();
  const int col0 = cbase + (lb % ntile) * Ct;
  bwd_supernode(f, lev[lb / ntile], col0, min(Ct, cbase + cw - col0), lg, C, B, sh);
}

// Backward substitution of whole subtrees: parents before children (reverse
// postorder), after the levels above them.
__global__ __launch_bounds__(1024) void k_bwd_subtree(FactorView f, const int2* __restrict__ sub, int ntile, int lg,
                                                     int C, double* __restrict__ B, const int* __restrict__ any,
                                                     int cbase, int cw, const char* __restrict__ skip) {
  if (any && *any == 0) return;
  extern __shared__ double sh[];
  const int Ct = 1 << lg;
  const int lb = xcd_block();
  const int2 rg = sub[lb / ntile];
  const int col0 = cbase + (lb % ntile) * Ct;
  const int ct = min(Ct, cbase + cw - col0);
  for (int s = rg.y; s >= rg.x; --s) {
    if (skip && skip[s]) continue;  // a dense leaf (solved after the sweeps)
    bwd_supernode(f, s, col0, ct, lg, C, B, sh);
    __syncthreads();
  }
}

// ----------------------------------------------------------------------------
// Level kernels with one load phase. A block owns (supernode, a contiguous
// range of its 16-row output tiles, one column tile). All it reads from global
// memory it requests at once: the front rows it needs, the children's update
// rows that add into them (pulled through per-row lists, in child order) and
// the slice of W its tiles multiply with; then the MFMA chains run from LDS
// and the results go straight to global memory. The arithmetic (values, their
// order of summation, the MFMA chain over k) is that of fwd_supernode /
// bwd_supernode, so results are bitwise the same; what changes is the number
// of dependent global round trips (two instead of one per child plus one per
// eight MFMA operand batches) and that a large supernode is spread over
// several blocks. The forward writes y into Yb (not in place), so that blocks
// of one supernode never overwrite rows another block still reads; the
// backward reads its own rows' y from Yb and the rows below from B (final
// x of the ancestors) and writes x into B.
struct PullView {
  const long long* pbase;  // per supernode: offset of its m x maxp slot table in ptab
  const int* maxp;         // per supernode: slots per front row (its most contributions to one row)
  const int* ptab;         // per front row, maxp global update-block rows in child order (-1: none)
};
constexpr int kPullFast = 4;  // slots whose loads are issued together; more go one by one

// value of front row r (< m) at column c: own b (rows < w) plus the pulls.
// The slot indices are read with the own values, then the update rows of up
// to kPullFast slots at once (two dependent round trips); the sum runs in
// slot (child) order.
__device__ __forceinline__ void load_rows_pulled(const double* __restrict__ B, const double* __restrict__ U,
                                                 const int* __restrict__ ptab, int maxp, int c0, int w, int r0,
                                                 int nrows, int m, int lg, int ct, int C, int col0,
                                                 double* __restrict__ T, TileIdx ti) {
  const int Ct = 1 << lg;
  const bool cok = ti.c < ct;
  const int np = min(maxp, kPullFast);
  for (int rb = ti.r0; rb < nrows; rb += 4 * ti.rs) {
    double v[4];
    int q[4][kPullFast];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int j = rb + u * ti.rs;
      const int r = r0 + j;
      const bool ok = cok && j < nrows && r < m;
      v[u] = ok && r < w ? B[(long long)(c0 + r) * C + col0 + ti.c] : 0.0;
#pragma unroll
      for (int k = 0; k < kPullFast; ++k) q[u][k] = ok && k < np ? ptab[(long long)r * maxp + k] : -1;
    }
    double a[4][kPullFast];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int k = 0; k < kPullFast; ++k) a[u][k] = q[u][k] >= 0 ? U[(long long)q[u][k] * C + col0 + ti.c] : 0.0;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
#pragma unroll
      for (int k = 0; k < kPullFast; ++k)
        if (q[u][k] >= 0) v[u] += a[u][k];
      const int j = rb + u * ti.rs;
      const int r = r0 + j;
      if (cok && j < nrows && r < m)
        for (int k = kPullFast; k < maxp; ++k) {
          const int qq = ptab[(long long)r * maxp + k];
          if (qq >= 0) v[u] += U[(long long)qq * C + col0 + ti.c];
        }
      if (j < nrows) T[j * Ct + ti.c] = v[u];
    }
  }
}

// Forward tiles [t0, t1) of supernode s: tiles below nrt = wpad / 16 are the
// y rows (inv(L_ss) T_top, written to Yb), the others update rows (written to
// U). LDS: T_top (wpad x Ct), the group's update rows (16 per update tile x
// Ct), then W per tile as [k][16] (k < wpad).
__device__ __forceinline__ void fwd_group(const FactorView& f, const PullView& p, int s, int t0, int t1, int col0,
                                          int ct, int lg, int C, const double* __restrict__ B, double* __restrict__ Yb,
                                          double* __restrict__ U, double* __restrict__ sh) {
  const int Ct = 1 << lg;
  const int c0 = f.first[s];
  const int w = f.first[s + 1] - c0;
  const int m = (int)(f.rowptr[s + 1] - f.rowptr[s]);
  const int wpad = (w + 15) & ~15;
  const int nrt = wpad >> 4;
  const int u0 = max(t0, nrt) - nrt;  // first update tile of the group
  const int nupd = t1 - max(t0, nrt);
  const int ntg = t1 - t0;
  double* Ttop = sh;
  double* Tupd = sh + wpad * Ct;
  double* Wsl = Tupd + (nupd > 0 ? nupd : 0) * 16 * Ct;
  const int tid = threadIdx.x, nt = blockDim.x;
  const int lane = tid & 63, wave = tid >> 6, nwave = nt >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  const int nq = Ct >> 4;
  const TileIdx ti = tile_idx(tid, nt, lg);
  const int* ptab = p.ptab + p.pbase[s];
  const int maxp = p.maxp[s];
  load_rows_pulled(B, U, ptab, maxp, c0, w, 0, wpad, w, lg, ct, C, col0, Ttop, ti);
  if (nupd > 0) load_rows_pulled(B, U, ptab, maxp, c0, w, w + 16 * u0, 16 * nupd, m, lg, ct, C, col0, Tupd, ti);
  {
    // W slice: tile g rows (y tile: rt + r, update tile: w + 16 (t - nrt) + r), k < wpad
    const double* Ws = f.W + f.lptr[s];
    const int tot = ntg * wpad * 16;
    for (int e0 = tid; e0 < tot; e0 += 8 * nt) {
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int e = e0 + u * nt;
        double val = 0.0;
        if (e < tot) {
          const int r = e & 15, k = (e >> 4) % wpad, g = (e >> 4) / wpad;
          const int t = t0 + g;
          const int row = t < nrt ? 16 * t + r : w + 16 * (t - nrt) + r;
          const bool ok = k < w && (t < nrt ? row < w : row < m);
          val = ok ? Ws[(long long)k * m + row] : 0.0;
        }
        v[u] = val;
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int e = e0 + u * nt;
        if (e < tot) Wsl[e] = v[u];
      }
    }
  }
  __syncthreads();
  for (int tq = wave; tq < ntg * nq; tq += nwave) {
    const int g = tq / nq, q = tq % nq;
    const int t = t0 + g;
    const double* pa = Wsl + (size_t)g * wpad * 16 + lr;
    const double* pb = Ttop + q * 16 + lr;
    const int col = q * 16 + lr;
    if (t < nrt) {
      const int rt = 16 * t;
      v4d acc = {0.0, 0.0, 0.0, 0.0};
      acc = mfma_chain(acc, 0, min(w, rt + 16), lk, [&](int k) { return pa[k * 16]; },
                       [&](int k) { return pb[k * Ct]; });
      if (col < ct) {
        double* Yo = Yb + col0 + col;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int r = rt + 4 * i + lk;
          if (r < w) Yo[(long long)(c0 + r) * C] = acc[i];
        }
      }
    } else {
      const int ru = 16 * (t - nrt);  // update row offset (rows w + ru ...)
      const double* pu = Tupd + (size_t)(ru - 16 * u0) * Ct + q * 16 + lr;
      v4d acc;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int a = ru + 4 * i + lk;
        acc[i] = w + a < m ? pu[(4 * i + lk) * Ct] : 0.0;
      }
      acc = mfma_chain(acc, 0, w, lk, [&](int k) { return pa[k * 16]; }, [&](int k) { return pb[k * Ct]; });
      if (col < ct) {
        double* Us = U + f.uptr[s] * C + col0 + col;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int a = ru + 4 * i + lk;
          if (w + a < m) Us[(long long)a * C] = acc[i];
        }
      }
    }
  }
}

// Backward tiles [t0, t1) of supernode s (output rows 16 t ..., < w):
// x = W^T [y_s; x_below]. LDS: the front (mpad x Ct: y_s from Yb, the rows
// below from B), then W per tile as [k - rt][16] for k in [rt, m).
__device__ __forceinline__ void bwd_group(const FactorView& f, int s, int t0, int t1, int col0, int ct, int lg, int C,
                                          double* __restrict__ B, const double* __restrict__ Yb,
                                          double* __restrict__ sh) {
  const int Ct = 1 << lg;
  const int c0 = f.first[s];
  const int w = f.first[s + 1] - c0;
  const long long rp = f.rowptr[s];
  const int m = (int)(f.rowptr[s + 1] - rp);
  const int mpad = (m + 15) & ~15;
  const int* rows = f.rows + rp;
  const double* WRs = f.WR + f.lptr[s];
  double* T = sh;
  double* Wsl = sh + mpad * Ct;
  const int tid = threadIdx.x, nt = blockDim.x;
  const int lane = tid & 63, wave = tid >> 6, nwave = nt >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  const int nq = Ct >> 4;
  const TileIdx ti = tile_idx(tid, nt, lg);
  {
    const bool cok = ti.c < ct;
    for (int rb = ti.r0; rb < mpad; rb += 8 * ti.rs) {
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int r = rb + u * ti.rs;
        const bool ok = cok && r < m;
        const double* src = r < w ? Yb : B;
        v[u] = ok ? src[(long long)rows[r] * C + col0 + ti.c] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int r = rb + u * ti.rs;
        if (r < mpad) T[r * Ct + ti.c] = v[u];
      }
    }
  }
  // W slice offsets: tile g starts after the (m - 16 (t0 + j)) * 16 values of the tiles before it
  const int rt0 = 16 * t0;
  const int tot = [&]() {
    int n = 0;
    for (int t = t0; t < t1; ++t) n += (m - 16 * t) * 16;
    return n;
  }();
  for (int e0 = tid; e0 < tot; e0 += 8 * nt) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int e = e0 + u * nt;
      double val = 0.0;
      if (e < tot) {
        // locate tile g: offsets sum over j < g of (m - rt0 - 16 j) * 16
        int g = 0, base = 0;
        while (base + (m - rt0 - 16 * g) * 16 <= e) {
          base += (m - rt0 - 16 * g) * 16;
          ++g;
        }
        const int rt = rt0 + 16 * g;
        const int kk = (e - base) >> 4, cc = (e - base) & 15;
        const int col = rt + cc;
        val = col < w ? WRs[(long long)(rt + kk) * w + col] : 0.0;
      }
      v[u] = val;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int e = e0 + u * nt;
      if (e < tot) Wsl[e] = v[u];
    }
  }
  __syncthreads();
  const int ntg = t1 - t0;
  for (int tq = wave; tq < ntg * nq; tq += nwave) {
    const int g = tq / nq, q = tq % nq;
    const int rt = rt0 + 16 * g;
    int base = 0;
    for (int j = 0; j < g; ++j) base += (m - rt0 - 16 * j) * 16;
    const double* pa = Wsl + base + lr;  // pa[(k - rt) * 16] = W^T(rt + lr, k) for k >= rt
    const double* pb = T + q * 16 + lr;
    v4d acc = {0.0, 0.0, 0.0, 0.0};
    acc = mfma_chain(acc, rt, m, lk, [&](int k) { return pa[(k - rt) * 16]; }, [&](int k) { return pb[k * Ct]; });
    const int col = q * 16 + lr;
    if (col < ct) {
      double* Bo = B + col0 + col;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int r = rt + 4 * i + lk;
        if (r < w) Bo[(long long)(c0 + r) * C] = acc[i];
      }
    }
  }
}

// items: (supernode, first tile, end tile, 0) per block row; block = item x column tile
__global__ __launch_bounds__(1024) void k_fwd_level_g(FactorView f, PullView p, const int4* __restrict__ items,
                                                     int ntile, int lg, int C, const double* __restrict__ B,
                                                     double* __restrict__ Yb, double* __restrict__ U,
                                                     const int* __restrict__ any, int cbase, int cw) {
  if (any && *any == 0) return;
  extern __shared__ double sh[];
  const int Ct = 1 << lg;
  const int lb = xcd_block();
  const int4 it = items[lb / ntile];
  const int col0 = cbase + (lb % ntile) * Ct;
  fwd_group(f, p, it.x, it.y, it.z, col0, min(Ct, cbase + cw - col0), lg, C, B, Yb, U, sh);
}

__global__ __launch_bounds__(1024) void k_bwd_level_g(FactorView f, const int4* __restrict__ items, int ntile, int lg,
                                                     int C, double* __restrict__ B, const double* __restrict__ Yb,
                                                     const int* __restrict__ any, int cbase, int cw) {
  if (any && *any == 0) return;
  extern __shared__ double sh[];
  const int Ct = 1 << lg;
  const int lb = xcd_block();
  const int4 it = items[lb / ntile];
  const int col0 = cbase + (lb % ntile) * Ct;
  bwd_group(f, it.x, it.y, it.z, col0, min(Ct, cbase + cw - col0), lg, C, B, Yb, sh);
}

#include "pdm_top.h"

__global__ void k_set_int(int* __restrict__ p, int v) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *p = v;
}

// x_new = x_n + B for active rollouts; per-(chunk, column) partial sums of
// (x_new - x_old)^2 for the early-exit test. Block (64 columns, 4 rows).
constexpr int kChunk = 64;
__global__ void k_update(const double* __restrict__ Xn, const double* __restrict__ B, double* __restrict__ Xit,
                         const int* __restrict__ active, double* __restrict__ part, double* __restrict__ tape, int N,
                         int C, int* __restrict__ any, const int* __restrict__ slot, long long sstride) {
  __shared__ double red[4][64];
  if (tape && slot) tape += (long long)(*slot) * sstride;
  // any: "some rollout still iterates", rebuilt by k_err_final right after
  if (any && blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0 && threadIdx.y == 0) *any = 0;
  int col = blockIdx.x * 64 + threadIdx.x;
  int chunk = blockIdx.y;
  double acc = 0.0;
  if (col < C) {
    bool act = active[col / 3] != 0;
    int vend = min(N, (chunk + 1) * kChunk);
    int v = chunk * kChunk + threadIdx.y;
    // four rows per trip with their loads issued together; the squares are
    // added in row order
    for (; v + 12 < vend; v += 16) {
      double old[4], nw[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const long long i = (long long)(v + 4 * u) * C + col;
        old[u] = Xit[i];
        nw[u] = act ? Xn[i] + B[i] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const long long i = (long long)(v + 4 * u) * C + col;
        if (!act) nw[u] = old[u];
        const double d = nw[u] - old[u];
        acc += d * d;
        Xit[i] = nw[u];
        if (tape) tape[i] = nw[u];
      }
    }
    for (; v < vend; v += 4) {
      long long i = (long long)v * C + col;
      double old = Xit[i];
      double nw = act ? Xn[i] + B[i] : old;
      double d = nw - old;
      acc += d * d;
      Xit[i] = nw;
      if (tape) tape[i] = nw;
    }
  }
  red[threadIdx.y][threadIdx.x] = acc;
  __syncthreads();
  if (threadIdx.y == 0 && col < C)
    part[(long long)chunk * C + col] = red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] + red[3][threadIdx.x];
}

// Per rollout: err = ||x_new - x_old||_2 / (3N) (World.h:890-895); a rollout
// whose err falls under 5e-7 stops iterating. iters counts executed iterations.
// Per-rollout reductions of the (chunk, column) partials: one wave per
// rollout, lanes stride over the chunks, then a fixed-order wave reduction.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_down(v, off, 64));
  return v;
}

__global__ __launch_bounds__(256) void k_err_final(const double* __restrict__ part, int nchunk, int R, int N,
                                                   int* __restrict__ active, int* __restrict__ iters,
                                                   double* __restrict__ err, int* __restrict__ any) {
  const int r = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (r >= R || !active[r]) return;
  const long long C = 3LL * R;
  double sacc = 0.0;
  for (int ch = lane; ch < nchunk; ch += 64)
    sacc += part[ch * C + 3 * r] + part[ch * C + 3 * r + 1] + part[ch * C + 3 * r + 2];
  sacc = wave_sum(sacc);
  if (lane != 0) return;
  double e = sqrt(sacc) / (3.0 * N);
  err[r] = e;
  iters[r] += 1;
  if (e < kEps)
    active[r] = 0;
  else if (any)
    atomicOr(any, 1);
}
__global__ __launch_bounds__(256) void k_err_init(const double* __restrict__ part, int nchunk, int R, int N,
                                                  int* __restrict__ active, int* __restrict__ iters,
                                                  int* __restrict__ any) {
  const int r = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (r >= R) return;
  const long long C = 3LL * R;
  double sacc = 0.0;
  for (int ch = lane; ch < nchunk; ch += 64)
    sacc += part[ch * C + 3 * r] + part[ch * C + 3 * r + 1] + part[ch * C + 3 * r + 2];
  sacc = wave_sum(sacc);
  if (lane != 0) return;
  int a = (sqrt(sacc) / (3.0 * N) < kEps) ? 0 : 1;
  active[r] = a;
  iters[r] = 0;
  if (any && a) atomicOr(any, 1);
}

__global__ void k_sqsum_chunks(const double* __restrict__ X, double* __restrict__ part, int N, int C,
                               int* __restrict__ any) {
  __shared__ double red[4][64];
  if (any && blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0 && threadIdx.y == 0) *any = 0;
  int col = blockIdx.x * 64 + threadIdx.x;
  int chunk = blockIdx.y;
  double acc = 0.0;
  if (col < C) {
    int vend = min(N, (chunk + 1) * kChunk);
    for (int v = chunk * kChunk + threadIdx.y; v < vend; v += 4) {
      double x = X[(long long)v * C + col];
      acc += x * x;
    }
  }
  red[threadIdx.y][threadIdx.x] = acc;
  __syncthreads();
  if (threadIdx.y == 0 && col < C)
    part[(long long)chunk * C + col] = red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] + red[3][threadIdx.x];
}

// Constant attachment part of J d: W p.
__global__ void k_attach_rhs(const double* __restrict__ attw, const double* __restrict__ xrest, double* __restrict__ Cc,
                             int N, int C) {
  long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (idx >= (long long)N * C) return;
  int v = (int)(idx / C);
  int c = (int)(idx % C) % 3;
  Cc[idx] = attw[v] * xrest[3 * v + c];
}

// End of step: v = damping (x+ - x)/h ; x = x+ (World.h:757-766, 708).
// Device copy of FP64 arrays, two values per thread (the runtime's blit
// kernel moves these state-sized buffers at a fraction of HBM speed).
// slot (optional): dst is advanced by *slot * sstride doubles (the tape slot of
// a step, read on the device so that a captured graph can serve every step).
__global__ __launch_bounds__(256) void k_copy_f64(double* __restrict__ dst, const double* __restrict__ src,
                                                  long long n, const int* __restrict__ slot, long long sstride) {
  if (slot) dst += (long long)(*slot) * sstride;
  long long i = 2 * ((long long)blockIdx.x * blockDim.x + threadIdx.x);
  if (i + 1 < n) {
    *reinterpret_cast<double2*>(dst + i) = *reinterpret_cast<const double2*>(src + i);
  } else if (i < n) {
    dst[i] = src[i];
  }
}
__global__ __launch_bounds__(256) void k_copy_f64_1(double* __restrict__ dst, const double* __restrict__ src,
                                                    long long n, const int* __restrict__ slot, long long sstride) {
  if (slot) dst += (long long)(*slot) * sstride;
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

// End of a step: v = (x_it - x_n)/h * damping, x = x_it (World.h:757-766,
// 708). Threads r < R also record the step's counters into the tapes.
__global__ void k_step_finish(const double* __restrict__ Xit, const double* __restrict__ Xn, double* __restrict__ X,
                              double* __restrict__ V, long long n, double h, double damping, int R,
                              const int* __restrict__ iters, int* __restrict__ tIters, const int* __restrict__ npass,
                              int* __restrict__ tNpass, const int* __restrict__ citers, int* __restrict__ tCiters) {
  long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (idx < R) {
    if (tIters) tIters[idx] = iters[idx];
    if (tNpass) tNpass[idx] = npass[idx];
    if (tCiters) tCiters[idx] = citers[idx];
  }
  if (idx >= n) return;
  double x1 = Xit[idx];
  V[idx] = (x1 - Xn[idx]) * (1.0 / h) * damping;
  X[idx] = x1;
}

// ------------------------------------------------------------ adjoint kernels
// ybar = gx + (gamma/h) gv ; xbar = -(gamma/h) gv ; zsum = 0
__global__ void k_adj_seed(const double* __restrict__ gX, const double* __restrict__ gV, double* __restrict__ Yb,
                           double* __restrict__ Xb, double* __restrict__ Zs, long long n, double gh) {
  long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (idx >= n) return;
  Yb[idx] = gX[idx] + gh * gV[idx];
  Xb[idx] = -gh * gV[idx];
  Zs[idx] = 0.0;
}

__global__ void k_copy_masked(const double* __restrict__ src, double* __restrict__ dst, const int* __restrict__ active,
                              long long n, int C) {
  long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (idx >= n) return;
  int j = (int)(idx % C);
  dst[idx] = active[j / 3] ? src[idx] : 0.0;
}

// After z = K^-1 ybar: xbar += ybar, zsum += z for active rollouts.
__global__ void k_adj_post_solve(const double* __restrict__ Z, const double* __restrict__ Yb, double* __restrict__ Xb,
                                 double* __restrict__ Zs, const int* __restrict__ active, long long n, int C) {
  long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (idx >= n) return;
  int j = (int)(idx % C);
  if (!active[j / 3]) return;
  Xb[idx] += Yb[idx];
  Zs[idx] += Z[idx];
}

// Local-step VJP of one tet at iterate y_k with z = K^-1 ybar_{k+1}:
// corners get Fbar G_k; the per-(tet,rollout) E-gradient accumulates
// (A z).(d_c(y_k) - A y_{k+1}) dw_c/dE over the two weights.
template <typename T, bool UNI>
__global__ __launch_bounds__(256, (sizeof(T) == 4 && UNI) ? 4 : 1) void k_vjp_tet(const double* __restrict__ Yk, const double* __restrict__ Yk1,
                                                 const double* __restrict__ Z, const int* __restrict__ tv,
                                                 const double* __restrict__ tG, const double* __restrict__ wr,
                                                 const double* __restrict__ wv, const double* __restrict__ dEr,
                                                 const double* __restrict__ dEv, const int* __restrict__ tslot,
                                                 const int* __restrict__ alist, const int* __restrict__ acount,
                                                 T* __restrict__ corner, double* __restrict__ gEpart, int Tn, int R,
                                                 const double* __restrict__ Xn, const double* __restrict__ CD,
                                                 const int* __restrict__ mptr, const double* __restrict__ mg,
                                                 const double* __restrict__ mw, const int* __restrict__ mgroup, int Mn,
                                                 const double* __restrict__ act, int act_stride_r, int act_off,
                                                 double* __restrict__ gMpart, double* __restrict__ gEw) {
  // UNI: the whole wave stays (its idle lanes compute on the wave's first
  // rollout and write nothing) so that the E-gradient terms of its 64
  // rollouts are summed across the wave and added once per (tet, wave) into
  // gEw; otherwise per (tet, rollout) into gEpart
  int t, j;
  bool live = true;
  if (UNI) {
    if (!element_wave(acount, Tn, &t, &j, &live)) return;
  } else if (dense_item(acount, Tn, R, blockIdx.x, &t, &j) <= 0) {
    return;
  }
  const int r = alist[j];
  int C = 3 * R;
  int v[4] = {tv[t], tv[Tn + t], tv[2 * Tn + t], tv[3 * Tn + t]};
  T G[12];
#pragma unroll
  for (int q = 0; q < 12; ++q) G[q] = (T)tG[(long long)q * Tn + t];
  // FP64 quantities of the parameter gradient first (short live ranges):
  // Fz = A z, F1 = A y_{k+1}; the FP32 path keeps only Zt = (T)Fz and the
  // scalars sum Fz.F1 and the contact term across the SVD.
  T F[9], Zt[9];
  double fzf1 = 0.0, extra = 0.0;
  double F1[9], Fz[9];
  {
    T ds[9];
    const double* b = Yk + (long long)v[0] * C + 3 * r;
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int k = 1; k < 4; ++k) ds[c * 3 + (k - 1)] = (T)(Yk[(long long)v[k] * C + 3 * r + c] - b[c]);
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int jj = 0; jj < 3; ++jj)
        F[c * 3 + jj] = ds[c * 3 + 0] * G[3 + jj] + ds[c * 3 + 1] * G[6 + jj] + ds[c * 3 + 2] * G[9 + jj];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int jj = 0; jj < 3; ++jj) {
        double a = 0, z = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          double g = tG[(long long)(3 * k + jj) * Tn + t];
          a += Yk1[(long long)v[k] * C + 3 * r + c] * g;
          z += Z[(long long)v[k] * C + 3 * r + c] * g;
        }
        F1[c * 3 + jj] = a;
        Fz[c * 3 + jj] = z;
        Zt[c * 3 + jj] = (T)z;
      }
    if (sizeof(T) == 4) {
#pragma unroll
      for (int i = 0; i < 9; ++i) fzf1 += Fz[i] * F1[i];
    }
    if (CD) {
      // contact iteration: C and D of the Jacobi update depend on the weights
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        double g0 = tG[(long long)(3 * k) * Tn + t], g1 = tG[(long long)(3 * k + 1) * Tn + t],
               g2 = tG[(long long)(3 * k + 2) * Tn + t];
        long long bb = (long long)v[k] * C + 3 * r;
        double zd = Z[bb] * (Yk[bb] - Xn[bb]) + Z[bb + 1] * (Yk[bb + 1] - Xn[bb + 1]) + Z[bb + 2] * (Yk[bb + 2] - Xn[bb + 2]);
        extra += (g0 * g0 + g1 * g1 + g2 * g2) * (zd + CD[(long long)v[k] * R + r]);
      }
    }
  }
  T Fb[9];
  double gr, gv;
  if (sizeof(T) == 8) {
    // FP64 path: the exact formula sum Fz (md - F1)
    T md[9], mdv[9];
    corotate_vjp<T>(F, (T)wr[t], (T)wv[t], Zt, Fb, md, mdv);
    gr = extra;
    gv = extra;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      gr += Fz[i] * ((double)md[i] - F1[i]);
      gv += Fz[i] * ((double)mdv[i] - F1[i]);
    }
  } else {
    T zmd, zmdv;
    corotate_vjp<T>(F, (T)wr[t], (T)wv[t], Zt, Fb, nullptr, nullptr, &zmd, &zmdv);
    gr = (double)zmd - fzf1 + extra;
    gv = (double)zmdv - fzf1 + extra;
  }
  const double ge = gr * dEr[t] + gv * dEv[t];
  if (UNI) {
    const double sw = wave_sum(live ? ge : 0.0);  // fixed-order butterfly over the wave
    if ((threadIdx.x & 63) == 0) gEw[(long long)t * (R >> 6) + (j >> 6)] += sw;
  } else {
    gEpart[(long long)t * R + r] += ge;
  }
  T f[12];
#pragma unroll
  for (int k = 0; k < 4; ++k)
#pragma unroll
    for (int c = 0; c < 3; ++c)
      f[3 * k + c] = Fb[c * 3 + 0] * G[3 * k + 0] + Fb[c * 3 + 1] * G[3 * k + 1] + Fb[c * 3 + 2] * G[3 * k + 2];
  // Muscles on this tet: ybar += w (1-a) A^T (A z); dL/da += -w (A z).(F(y_k) f).
  for (int m = mptr[t], me = mptr[t + 1]; m < me; ++m) {
    double a = act ? act[(long long)r * act_stride_r + act_off + mgroup[m]] : 0.0;
    double g[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) g[k] = mg[(long long)k * Mn + m];
    double fz[3] = {0, 0, 0}, ff[3] = {0, 0, 0};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      double b = Yk[(long long)v[0] * C + 3 * r + c];
#pragma unroll
      for (int k = 0; k < 4; ++k) fz[c] += Z[(long long)v[k] * C + 3 * r + c] * g[k];
#pragma unroll
      for (int k = 1; k < 4; ++k) ff[c] += (Yk[(long long)v[k] * C + 3 * r + c] - b) * g[k];
    }
    double wm = mw[m];
    if (live) gMpart[(long long)m * R + r] += -wm * (fz[0] * ff[0] + fz[1] * ff[1] + fz[2] * ff[2]);
    double sc = wm * (1.0 - a);
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
      for (int c = 0; c < 3; ++c) f[3 * k + c] += (T)(sc * fz[c] * g[k]);
  }
  if (!live) return;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    T* o = corner + (long long)tslot[k * Tn + t] * C + 3 * r;
#pragma unroll
    for (int c = 0; c < 3; ++c) o[c] = f[3 * k + c];
  }
}

// gact[r, step, g] = sum over the group's muscles (fixed order, 8 partial sums).
__global__ void k_reduce_groups(const double* __restrict__ gMpart, const int* __restrict__ gptr,
                                const int* __restrict__ gmus, double* __restrict__ gact, int G, int R, int steps,
                                int step) {
  int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= G * R) return;
  int g = idx / R, r = idx % R;
  double p[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  int q = gptr[g], qe = gptr[g + 1];
  for (; q + 8 <= qe; q += 8) {
#pragma unroll
    for (int u = 0; u < 8; ++u) p[u] += gMpart[(long long)gmus[q + u] * R + r];
  }
  for (; q < qe; ++q) p[0] += gMpart[(long long)gmus[q] * R + r];
  gact[((long long)r * steps + step) * G + g] = ((p[0] + p[1]) + (p[2] + p[3])) + ((p[4] + p[5]) + (p[6] + p[7]));
}

// End of a step's reverse sweep:
// gx_n = xbar + ybar_0 - L zsum ; gv_n = h ybar_0 + (M/h) zsum.
__global__ void k_adj_step_end(const double* __restrict__ Xb, const double* __restrict__ Yb,
                               const double* __restrict__ LZ, const double* __restrict__ Zs,
                               const double* __restrict__ mass, double* __restrict__ gX, double* __restrict__ gV,
                               int N, int C, double h) {
  long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (idx >= (long long)N * C) return;
  int v = (int)(idx / C);
  gX[idx] = Xb[idx] + Yb[idx] - LZ[idx];
  gV[idx] = h * Yb[idx] + mass[v] / h * Zs[idx];
}

__global__ void k_active_rev(const int* __restrict__ iters, int k, int R, int* __restrict__ active) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < R) active[r] = k < iters[r] ? 1 : 0;
}

// gE of tet t (caller order): its per-rollout partials, then its per-wave ones.
__global__ void k_reduce_rows(const double* __restrict__ part, const double* __restrict__ wpart,
                              const int* __restrict__ tet_order, double* __restrict__ out, int Tn, int R, int W) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= Tn) return;
  double s = 0.0;
  for (int r = 0; r < R; ++r) s += part[(long long)t * R + r];
  for (int w = 0; w < W; ++w) s += wpart[(long long)t * W + w];
  out[tet_order[t]] = s;
}


// ------------------------------------------------------------ contact phase
// Static-obstacle CCD of a vertex moving x0 -> x1 (arcsim collision_test,
// collision.cpp:211-278, for a static face): an impact when the path crosses
// the surface in [0,1] or ends within the thickness 1e-4; the contact normal
// opposes the motion (collision.cpp:269-273).
constexpr double kThick = 1e-4;
__device__ bool detect_plane_d(const double* pl, const double* x0, const double* x1, double* n) {
  double nn = sqrt(pl[3] * pl[3] + pl[4] * pl[4] + pl[5] * pl[5]);
  double u[3] = {pl[3] / nn, pl[4] / nn, pl[5] / nn};
  double ds = 0, de = 0, vn = 0;
  for (int r = 0; r < 3; ++r) {
    ds += (x0[r] - pl[r]) * u[r];
    de += (x1[r] - pl[r]) * u[r];
    vn += (x1[r] - x0[r]) * u[r];
  }
  bool hit = false;
  if (ds != de) {
    double t = ds / (ds - de);
    hit = t >= 0 && t <= 1;
  }
  if (!hit) hit = fabs(de) < kThick;
  if (!hit) return false;
  double sg = vn > 0 ? -1.0 : 1.0;
  for (int r = 0; r < 3; ++r) n[r] = sg * u[r];
  return true;
}

// t_out: the contact time (a crossing root of |x0 + t (x1 - x0) - c| = r, or 1
// for end-of-step proximity); root_out: whether t is such a root.
__device__ bool detect_sphere_d(const double* sp, const double* x0, const double* x1, double* n, double* t_out,
                                int* root_out) {
  double d[3], e[3];
  for (int r = 0; r < 3; ++r) {
    d[r] = x1[r] - x0[r];
    e[r] = x0[r] - sp[r];
  }
  double a = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
  double b = 2 * (d[0] * e[0] + d[1] * e[1] + d[2] * e[2]);
  double c = e[0] * e[0] + e[1] * e[1] + e[2] * e[2] - sp[3] * sp[3];
  double ts[3];
  int nt = 0;
  if (a > 0) {
    double disc = b * b - 4 * a * c;
    if (disc >= 0) {
      double sq = sqrt(disc);
      double q = -0.5 * (b + (b >= 0 ? sq : -sq));
      double r1 = q / a, r2 = q != 0 ? c / q : r1;
      if (r1 > r2) {
        double t = r1; r1 = r2; r2 = t;
      }
      ts[nt++] = r1;
      ts[nt++] = r2;
    }
  }
  ts[nt++] = 1.0;
  for (int k = 0; k < nt; ++k) {
    double t = ts[k];
    if (t < 0 || t > 1) continue;
    double p[3];
    for (int r = 0; r < 3; ++r) p[r] = x0[r] + t * d[r] - sp[r];
    double len = sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
    if (fabs(len - sp[3]) < kThick && len > 0) {
      double vn = (p[0] * d[0] + p[1] * d[1] + p[2] * d[2]) / len;
      double sg = vn > 0 ? -1.0 : 1.0;
      for (int r = 0; r < 3; ++r) n[r] = sg * p[r] / len;
      *t_out = t;
      *root_out = k < nt - 1;
      return true;
    }
  }
  return false;
}

// ArcSim's kDOP-18 box test in single precision (bvh.cpp:54-96, bvh.hpp:62-150):
// the swept box of the vertex (x0, x1) against the triangle's box dilated by
// d on the three axes and sqrt(2) d on the six diagonal ones.
__device__ inline void kdop_add(const double* p, float* mn, float* mx) {
  float a = (float)p[0], b = (float)p[1], c = (float)p[2];
  float q[9] = {a, b, c, a + b, a + c, b + c, a - b, a - c, b - c};
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    mn[i] = fminf(mn[i], q[i]);
    mx[i] = fmaxf(mx[i], q[i]);
  }
}
__device__ inline bool kdop_overlap(const double* x0, const double* x1, const double* Y, double thick) {
  float an[9], ax[9], bn[9], bx[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    an[i] = bn[i] = 3.0e38f;
    ax[i] = bx[i] = -3.0e38f;
  }
  kdop_add(x1, an, ax);
  kdop_add(x0, an, ax);
  kdop_add(Y, bn, bx);
  kdop_add(Y + 3, bn, bx);
  kdop_add(Y + 6, bn, bx);
  const double d = (double)(float)thick, d2 = sqrt(2.0) * d;
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    float lo = (float)((double)bn[i] - (i < 3 ? d : d2)), hi = (float)((double)bx[i] + (i < 3 ? d : d2));
    if (an[i] > hi || ax[i] < lo) return false;
  }
  return true;
}

__device__ inline double stp3(const double* a, const double* b, const double* c) {
  return a[0] * (b[1] * c[2] - b[2] * c[1]) + a[1] * (b[2] * c[0] - b[0] * c[2]) + a[2] * (b[0] * c[1] - b[1] * c[0]);
}

// Vertex x0 -> x1 against a static obstacle triangle Y (3 corners): ArcSim's
// vertex-face test (collision.cpp:127-138, 211-278) with the face at rest. The
// coplanarity cubic then has a single root, t = -a0/a1 (the quadratic and
// cubic terms vanish); at that t and at t = 1 the vertex hits when it lies
// within the thickness of the face's plane and inside the face (barycentric
// weights >= -1e-6), or at t = 1 within the thickness of a corner or of an
// edge's line. The normal is the face normal turned against the vertex's
// motion, unnormalised as the reference leaves it.
__device__ bool detect_tri_d(const double* Y, const double* x0, const double* x1, double* nout) {
  if (!kdop_overlap(x0, x1, Y, kThick)) return false;
  const double *y0 = Y, *y1 = Y + 3, *y2 = Y + 6;
  double v0[3], X1[3], X2[3], X3[3], V[3], e1[3], e2[3];
  for (int r = 0; r < 3; ++r) {
    v0[r] = x1[r] - x0[r];
    X1[r] = y0[r] - x0[r];
    X2[r] = y1[r] - x0[r];
    X3[r] = y2[r] - x0[r];
    V[r] = 0.0 - v0[r];
    e1[r] = y1[r] - y0[r];
    e2[r] = y2[r] - y0[r];
  }
  double l1 = sqrt(e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2]), l2 = sqrt(e2[0] * e2[0] + e2[1] * e2[1] + e2[2] * e2[2]);
  for (int r = 0; r < 3; ++r) {
    e1[r] /= l1;
    e2[r] /= l2;
  }
  double fn[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
  double f2 = fn[0] * fn[0] + fn[1] * fn[1] + fn[2] * fn[2];
  if (!(f2 >= 1e-6)) return false;  // degenerate face: infinite distance
  double fl = sqrt(f2), nu[3] = {fn[0] / fl, fn[1] / fl, fn[2] / fl};
  double a0 = stp3(X1, X2, X3);
  double a1 = stp3(V, X2, X3) + stp3(X1, V, X3) + stp3(X1, X2, V);
  double ts[2];
  int nt = 0;
  if (a1 != 0.0) ts[nt++] = -a0 / a1;
  ts[nt++] = 1.0;
  for (int k = 0; k < nt; ++k) {
    double t = ts[k];
    if (t < 0 || t > 1) continue;
    double p[3], q0[3], q1[3], q2[3];
    for (int r = 0; r < 3; ++r) {
      p[r] = x0[r] + t * v0[r];
      q0[r] = y0[r] - p[r];
      q1[r] = y1[r] - p[r];
      q2[r] = y2[r] - p[r];
    }
    double h = (p[0] - y0[0]) * nu[0] + (p[1] - y0[1]) * nu[1] + (p[2] - y0[2]) * nu[2];
    double b0 = stp3(q1, q2, nu), b1 = stp3(q2, q0, nu), b2 = stp3(q0, q1, nu);
    double sb = b0 + b1 + b2;
    double w1 = -b0 / sb, w2 = -b1 / sb, w3 = -b2 / sb;
    bool inside = fmin(fmin(-w1, -w2), -w3) >= -1e-6;
    if (t == 1.0) {
      const double* Yk[3] = {y0, y1, y2};
      for (int c = 0; c < 3; ++c) {
        const double *A = Yk[c], *B = Yk[(c + 1) % 3];
        double pa[3] = {p[0] - A[0], p[1] - A[1], p[2] - A[2]};
        double ab[3] = {B[0] - A[0], B[1] - A[1], B[2] - A[2]};
        inside = inside || sqrt(pa[0] * pa[0] + pa[1] * pa[1] + pa[2] * pa[2]) < kThick;
        double dd = pa[0] * ab[0] + pa[1] * ab[1] + pa[2] * ab[2], n2 = ab[0] * ab[0] + ab[1] * ab[1] + ab[2] * ab[2];
        double e[3];
        for (int r = 0; r < 3; ++r) e[r] = pa[r] - dd * ab[r] / n2;
        double el = sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
        inside = inside || el < kThick;
      }
    }
    double wv[3] = {w1 * V[0] + w2 * V[0] + w3 * V[0], w1 * V[1] + w2 * V[1] + w3 * V[1],
                    w1 * V[2] + w2 * V[2] + w3 * V[2]};
    double sg = (nu[0] * wv[0] + nu[1] * wv[1] + nu[2] * wv[2]) > 0 ? -1.0 : 1.0;
    if (fabs(h) < kThick && inside) {
      for (int r = 0; r < 3; ++r) nout[r] = sg * fn[r];
      return true;
    }
  }
  return false;
}

// Detection pass for (contact vertex, rollout). The reference keeps a list of
// impacts keyed by (vertex, obstacle face), appending new ones and refreshing
// known ones in place (collisionBrutal.h:291-327), and a vertex's response
// uses the last of its impacts in that list (World.h:1417-1425). Only that
// winner matters, so each (vertex, rollout) keeps: the winner's stamp
// (pass * nobs + obstacle, the list position order) and normal, and the set of
// obstacles seen this step (up to kSeen; more sets cactive[R], an error). A
// known obstacle only refreshes its normal; a new one wins when its stamp is
// the largest. Obstacles: planes, spheres, then triangles (via their BVH).
constexpr int kSeen = 8;
__global__ void k_contact_detect(const int* __restrict__ cvtx, int ncv, const double* __restrict__ Xn,
                                 const double* __restrict__ Ori, int R, const double* __restrict__ planes, int np,
                                 const double* __restrict__ spheres, int nsph, const double* __restrict__ tris,
                                 int ntri, const double* __restrict__ bvh_box, const int* __restrict__ bvh_a,
                                 const int* __restrict__ bvh_b, const int* __restrict__ bvh_tri,
                                 int* __restrict__ win_st, int* __restrict__ win_o, double* __restrict__ win_n,
                                 int* __restrict__ nseen, int* __restrict__ seen, int pass,
                                 int* __restrict__ cactive, int* __restrict__ fresh, int* __restrict__ win_sph,
                                 double* __restrict__ win_t, double* __restrict__ win_x1) {
  long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (idx >= (long long)ncv * R) return;
  int ci = (int)(idx / R), r = (int)(idx % R);
  if (!cactive[r]) return;
  int C = 3 * R;
  int v = cvtx[ci];
  const double* x0 = Xn + (long long)v * C + 3 * r;
  const double* x1 = Ori + (long long)v * C + 3 * r;
  const int nobs = np + nsph + ntri;
  int wst = win_st[idx], wo = win_o[idx], ns = nseen[idx];
  double wn[3] = {win_n[3 * idx], win_n[3 * idx + 1], win_n[3 * idx + 2]};
  bool any = false, over = false;
  // sphere data of the winner's latest detection (win_sph = 4 (pass * nsph +
  // sphere) + 2 root + 1, or 0 for another obstacle): the reverse sweep
  // differentiates its normal
  int wsph = win_sph ? win_sph[idx] : 0;
  double wt = 0.0;
  auto record = [&](int o, const double* n, int sinfo, double t) {
    any = true;
    for (int k = 0; k < ns; ++k)
      if (seen[(long long)k * ncv * R + idx] == o) {
        if (o == wo) {
          for (int c = 0; c < 3; ++c) wn[c] = n[c];
          wsph = sinfo;
          wt = t;
        }
        return;
      }
    if (ns == kSeen) {
      over = true;
      return;
    }
    seen[(long long)ns * ncv * R + idx] = o;
    ++ns;
    int st = pass * nobs + o;
    if (st > wst) {
      wst = st;
      wo = o;
      for (int c = 0; c < 3; ++c) wn[c] = n[c];
      wsph = sinfo;
      wt = t;
    }
  };
  double n[3];
  for (int o = 0; o < np; ++o)
    if (detect_plane_d(planes + 6 * o, x0, x1, n)) record(o, n, 0, 0.0);
  for (int o = 0; o < nsph; ++o) {
    double t;
    int root;
    if (detect_sphere_d(spheres + 4 * o, x0, x1, n, &t, &root)) record(np + o, n, 4 * (pass * nsph + o) + 2 * root + 1, t);
  }
  if (ntri > 0) {
    // swept box of the vertex, widened by the thickness and a margin for the
    // single-precision box test, against the BVH's double boxes
    double lo[3], hi[3];
    for (int c = 0; c < 3; ++c) {
      double m = 2.0 * kThick + 1e-6 * (1.0 + fabs(x0[c]) + fabs(x1[c]));
      lo[c] = fmin(x0[c], x1[c]) - m;
      hi[c] = fmax(x0[c], x1[c]) + m;
    }
    int stack[64], sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
      int k = stack[--sp];
      const double* b = bvh_box + 6 * (long long)k;
      if (lo[0] > b[3] || lo[1] > b[4] || lo[2] > b[5] || hi[0] < b[0] || hi[1] < b[1] || hi[2] < b[2]) continue;
      int a = bvh_a[k];
      if (a >= 0) {
        if (sp + 2 <= 64) {
          stack[sp++] = bvh_b[k];
          stack[sp++] = a;
        } else {
          over = true;
        }
        continue;
      }
      for (int q = -1 - a, qe = -1 - a + bvh_b[k]; q < qe; ++q) {
        int t = bvh_tri[q];
        if (detect_tri_d(tris + 9 * (long long)t, x0, x1, n)) record(np + nsph + t, n, 0, 0.0);
      }
    }
  }
  win_st[idx] = wst;
  win_o[idx] = wo;
  nseen[idx] = ns;
  win_n[3 * idx] = wn[0];
  win_n[3 * idx + 1] = wn[1];
  win_n[3 * idx + 2] = wn[2];
  if (win_sph && win_sph[idx] != wsph) {  // a new sphere detection won: keep where it read
    win_sph[idx] = wsph;
    win_t[idx] = wt;
    for (int c = 0; c < 3; ++c) win_x1[3 * idx + c] = x1[c];
  }
  if (any) atomicOr(&fresh[r], 1);
  if (over) atomicOr(&cactive[R], 1);
}

__global__ void k_contact_reset(int* __restrict__ win_st, int* __restrict__ win_o, int* __restrict__ nseen, long long n,
                                int* __restrict__ cactive, int* __restrict__ cfirst, int* __restrict__ fresh,
                                int* __restrict__ clast, int* __restrict__ npass, int* __restrict__ citers, int R,
                                int* __restrict__ win_sph) {
  long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (idx < n) {
    win_st[idx] = -1;
    win_o[idx] = -1;
    nseen[idx] = 0;
    if (win_sph) win_sph[idx] = 0;
  }
  if (idx < R) {
    cactive[idx] = 1;
    cfirst[idx] = 1;
    fresh[idx] = 0;
    clast[idx] = 1;
    npass[idx] = 0;
    citers[idx] = 0;
  }
  if (idx == 0) cactive[R] = 0;
}

// Sphere data of the winning normals at the start of a pass (0 where the
// winner is no sphere), and per rollout whether the detection read x_it after
// the last executed iteration (clast = 1) or that iteration's input.
__global__ void k_contact_tape_sph(const int* __restrict__ win_st, const int* __restrict__ win_sph,
                                   const double* __restrict__ win_t, const double* __restrict__ win_x1,
                                   const int* __restrict__ clast, long long ni, int R, int* __restrict__ tSi,
                                   double* __restrict__ tSt, double* __restrict__ tSx1, int* __restrict__ tDl) {
  long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (idx < ni) {
    const int si = win_st[idx] >= 0 ? win_sph[idx] : 0;
    tSi[idx] = si;
    tSt[idx] = si ? win_t[idx] : 0.0;
    for (int c = 0; c < 3; ++c) tSx1[3 * idx + c] = si ? win_x1[3 * idx + c] : 0.0;
  }
  if (idx < R) tDl[idx] = clast[idx];
}

// ------------------------------------------------------------ self contact
// Surface vertices against surface faces of the body mesh ("Self-collision",
// collisionBrutal.h:281-406, 643-700; World.h:1246-1315, 1446-1600). Contact
// indices a number the surface vertices as the reference does.
constexpr int kSelfImp = 8;    // impacts kept per (vertex, rollout) in a step
constexpr int kSelfMem = 16;   // vertices of one self-contact group

// Real roots in [0, 1] of f(t) = a3 t^3 + a2 t^2 + a1 t + a0, ascending: the
// times at which a moving vertex is coplanar with a moving face. [0, 1] is
// cut at the stationary points of f inside it (roots of f' by the
// cancellation-free quadratic formula), so f is monotone on every piece. A
// piece whose end values differ in sign holds one root, found by bisection
// down to an interval of width 2^-40 (< 1e-12) and reported as its midpoint;
// a piece end where f vanishes is a root itself (reported once). An
// identically zero f is coplanar throughout: its earliest root t = 0.
__device__ int cubic_roots01(double a3, double a2, double a1, double a0, double t[3]) {
  auto f = [&](double x) { return ((a3 * x + a2) * x + a1) * x + a0; };
  if (a3 == 0.0 && a2 == 0.0 && a1 == 0.0 && a0 == 0.0) {
    t[0] = 0.0;
    return 1;
  }
  double cut[4];
  int nc = 0;
  cut[nc++] = 0.0;
  {
    // f'(t) = A t^2 + B t + c
    const double A = 3.0 * a3, B = 2.0 * a2, c = a1;
    double s0 = 2.0, s1 = 2.0;  // stationary points (2: none)
    if (A == 0.0) {
      if (B != 0.0) s0 = -c / B;
    } else {
      const double disc = B * B - 4.0 * A * c;
      if (disc > 0.0) {
        const double q = -0.5 * (B + copysign(sqrt(disc), B));
        s0 = q / A;
        if (q != 0.0) s1 = c / q;
        if (s1 < s0) {
          double tmp = s0;
          s0 = s1;
          s1 = tmp;
        }
      }
    }
    if (s0 > 0.0 && s0 < 1.0) cut[nc++] = s0;
    if (s1 > 0.0 && s1 < 1.0 && s1 != s0) cut[nc++] = s1;
  }
  cut[nc++] = 1.0;
  int n = 0;
  double fl = f(cut[0]);
  for (int k = 0; k + 1 < nc; ++k) {
    double lo = cut[k], hi = cut[k + 1];
    const double fh = f(hi);
    const double flo = fl;
    fl = fh;
    double r;
    if (flo == 0.0) {
      r = lo;
    } else if (fh == 0.0) {
      r = hi;
    } else if ((flo < 0.0) != (fh < 0.0)) {
      const bool neg_lo = flo < 0.0;
      for (int it = 0; it < 40; ++it) {
        const double mid = 0.5 * (lo + hi);
        const double fm = f(mid);
        if (fm == 0.0) {
          lo = hi = mid;
          break;
        }
        if ((fm < 0.0) == neg_lo)
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

// Moving vertex (p0 -> p1) against a moving face (y0[k] -> y1[k]) with
// thickness 2e-4 (find_face_impacts uses mult 2): kDOP box filter, the
// coplanarity cubic's roots in [0,1] and t = 1, distance and inside tests, the
// normal from the step-start corners turned against the relative motion.
__device__ bool vf_moving_d(const double* p0, const double* p1, const double* const* y0, const double* const* y1,
                            double* tout, double* dout, double* nout) {
  const double thk = 2.0 * kThick;
  {
    float an[9], ax[9], bn[9], bx[9];
    for (int i = 0; i < 9; ++i) {
      an[i] = bn[i] = 3.0e38f;
      ax[i] = bx[i] = -3.0e38f;
    }
    kdop_add(p1, an, ax);
    kdop_add(p0, an, ax);
    for (int k = 0; k < 3; ++k) {
      kdop_add(y1[k], bn, bx);
      kdop_add(y0[k], bn, bx);
    }
    const double d = (double)(float)thk, d2 = sqrt(2.0) * d;
    for (int i = 0; i < 9; ++i) {
      float lo = (float)((double)bn[i] - (i < 3 ? d : d2)), hi = (float)((double)bx[i] + (i < 3 ? d : d2));
      if (an[i] > hi || ax[i] < lo) return false;
    }
  }
  double v0[3], x[3][3], v[3][3];
  for (int r = 0; r < 3; ++r) {
    v0[r] = p1[r] - p0[r];
    for (int k = 0; k < 3; ++k) {
      x[k][r] = y0[k][r] - p0[r];
      v[k][r] = (y1[k][r] - y0[k][r]) - v0[r];
    }
  }
  double a0 = stp3(x[0], x[1], x[2]);
  double a1 = stp3(v[0], x[1], x[2]) + stp3(x[0], v[1], x[2]) + stp3(x[0], x[1], v[2]);
  double a2 = stp3(x[0], v[1], v[2]) + stp3(v[0], x[1], v[2]) + stp3(v[0], v[1], x[2]);
  double a3 = stp3(v[0], v[1], v[2]);
  double ts[4];
  int nsol = cubic_roots01(a3, a2, a1, a0, ts);
  ts[nsol++] = 1;
  double c[3];
  {
    double s1[3], s2[3];
    for (int r = 0; r < 3; ++r) {
      s1[r] = y0[1][r] - y0[0][r];
      s2[r] = y0[2][r] - y0[0][r];
    }
    double m1 = sqrt(s1[0] * s1[0] + s1[1] * s1[1] + s1[2] * s1[2]), m2 = sqrt(s2[0] * s2[0] + s2[1] * s2[1] + s2[2] * s2[2]);
    for (int r = 0; r < 3; ++r) {
      s1[r] /= m1;
      s2[r] /= m2;
    }
    c[0] = s1[1] * s2[2] - s1[2] * s2[1];
    c[1] = s1[2] * s2[0] - s1[0] * s2[2];
    c[2] = s1[0] * s2[1] - s1[1] * s2[0];
  }
  for (int i = 0; i < nsol; i++) {
    double t = ts[i];
    if (t < 0 || t > 1) continue;
    double P[3], Y[3][3];
    for (int r = 0; r < 3; ++r) {
      P[r] = p0[r] + t * (p1[r] - p0[r]);
      for (int k = 0; k < 3; ++k) Y[k][r] = y0[k][r] + t * (y1[k][r] - y0[k][r]);
    }
    double e1[3], e2[3];
    for (int r = 0; r < 3; ++r) {
      e1[r] = Y[1][r] - Y[0][r];
      e2[r] = Y[2][r] - Y[0][r];
    }
    double l1 = sqrt(e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2]), l2 = sqrt(e2[0] * e2[0] + e2[1] * e2[1] + e2[2] * e2[2]);
    for (int r = 0; r < 3; ++r) {
      e1[r] /= l1;
      e2[r] /= l2;
    }
    double nn[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
    double n2 = nn[0] * nn[0] + nn[1] * nn[1] + nn[2] * nn[2];
    if (!(n2 >= 1e-6)) continue;
    double nl = sqrt(n2);
    for (int r = 0; r < 3; ++r) nn[r] /= nl;
    double h = (P[0] - Y[0][0]) * nn[0] + (P[1] - Y[0][1]) * nn[1] + (P[2] - Y[0][2]) * nn[2];
    double q[3][3];
    for (int k = 0; k < 3; ++k)
      for (int r = 0; r < 3; ++r) q[k][r] = Y[k][r] - P[r];
    double b0 = stp3(q[1], q[2], nn), b1 = stp3(q[2], q[0], nn), b2 = stp3(q[0], q[1], nn);
    double sb = b0 + b1 + b2;
    double w[3] = {-b0 / sb, -b1 / sb, -b2 / sb};
    bool inside = fmin(fmin(-w[0], -w[1]), -w[2]) >= -1e-6;
    if (t == 1) {
      for (int k = 0; k < 3; ++k) {
        double dv[3] = {P[0] - Y[k][0], P[1] - Y[k][1], P[2] - Y[k][2]};
        inside = inside || sqrt(dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2]) < thk;
      }
      for (int k = 0; k < 3; ++k) {
        const double *A = Y[k], *B = Y[(k + 1) % 3];
        double pa[3] = {P[0] - A[0], P[1] - A[1], P[2] - A[2]}, ab[3] = {B[0] - A[0], B[1] - A[1], B[2] - A[2]};
        double dp = pa[0] * ab[0] + pa[1] * ab[1] + pa[2] * ab[2], ll = ab[0] * ab[0] + ab[1] * ab[1] + ab[2] * ab[2];
        double e[3];
        for (int r = 0; r < 3; ++r) e[r] = pa[r] - dp * ab[r] / ll;
        inside = inside || sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]) < thk;
      }
    }
    double rv[3];
    for (int r = 0; r < 3; ++r) rv[r] = w[0] * v[0][r] + w[1] * v[1][r] + w[2] * v[2][r];
    double sg = (nn[0] * rv[0] + nn[1] * rv[1] + nn[2] * rv[2]) > 0 ? -1.0 : 1.0;
    if (fabs(h) < thk && inside) {
      *tout = t;
      *dout = fabs(h);
      for (int r = 0; r < 3; ++r) nout[r] = sg * c[r];
      return true;
    }
  }
  return false;
}

// Per-step reset of the self-contact state: impact lists, the force carried
// between handlings and the group counts.
__global__ void k_self_reset(int* __restrict__ nimp, long long nar, double* __restrict__ self_cur, long long nc,
                             int* __restrict__ ngrp, int* __restrict__ ngrp_prev, int R) {
  long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (idx < nar) nimp[idx] = 0;
  if (idx < nc) self_cur[idx] = 0.0;
  if (idx < R) {
    ngrp[idx] = 0;
    ngrp_prev[idx] = 0;
  }
}

// Detection pass for (contact vertex a, rollout): every surface face not
// containing a; hits are added to a's impact list or replace the entry of the
// same face (collisionBrutal.h:291-306), new ones in face order.
__global__ void k_self_detect(const int* __restrict__ sc_vtx, const int* __restrict__ sc_face, int ncs, int nsf,
                              const double* __restrict__ Xn, const double* __restrict__ Ori, int R,
                              const int* __restrict__ cactive, int* __restrict__ nimp, int* __restrict__ imp_f,
                              double* __restrict__ imp_t, double* __restrict__ imp_d, double* __restrict__ imp_n,
                              int* __restrict__ flag) {
  long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (idx >= (long long)ncs * R) return;
  int a = (int)(idx / R), r = (int)(idx % R);
  if (!cactive[r]) return;
  const long long C = 3LL * R, NA = (long long)ncs * R;
  const double* p0 = Xn + (long long)sc_vtx[a] * C + 3 * r;
  const double* p1 = Ori + (long long)sc_vtx[a] * C + 3 * r;
  double lo[3], hi[3];
  for (int c = 0; c < 3; ++c) {
    double m = 4.0 * kThick + 1e-6 * (1.0 + fabs(p0[c]) + fabs(p1[c]));
    lo[c] = fmin(p0[c], p1[c]) - m;
    hi[c] = fmax(p0[c], p1[c]) + m;
  }
  int n = nimp[idx];
  for (int f = 0; f < nsf; ++f) {
    int f0 = sc_face[3 * f], f1 = sc_face[3 * f + 1], f2 = sc_face[3 * f + 2];
    if (f0 == a || f1 == a || f2 == a) continue;
    const double* y0[3] = {Xn + (long long)sc_vtx[f0] * C + 3 * r, Xn + (long long)sc_vtx[f1] * C + 3 * r,
                           Xn + (long long)sc_vtx[f2] * C + 3 * r};
    const double* y1[3] = {Ori + (long long)sc_vtx[f0] * C + 3 * r, Ori + (long long)sc_vtx[f1] * C + 3 * r,
                           Ori + (long long)sc_vtx[f2] * C + 3 * r};
    bool out = false;
    for (int c = 0; c < 3 && !out; ++c) {
      double fl = fmin(fmin(fmin(y0[0][c], y1[0][c]), fmin(y0[1][c], y1[1][c])), fmin(y0[2][c], y1[2][c]));
      double fh = fmax(fmax(fmax(y0[0][c], y1[0][c]), fmax(y0[1][c], y1[1][c])), fmax(y0[2][c], y1[2][c]));
      out = fl > hi[c] || fh < lo[c];
    }
    if (out) continue;
    double t, d, nv[3];
    if (!vf_moving_d(p0, p1, y0, y1, &t, &d, nv)) continue;
    int k = 0;
    while (k < n && imp_f[(long long)k * NA + idx] != f) ++k;
    if (k == n) {
      if (n == kSelfImp) {
        atomicOr(flag, 1);
        continue;
      }
      imp_f[(long long)k * NA + idx] = f;
      ++n;
    }
    imp_t[(long long)k * NA + idx] = t;
    imp_d[(long long)k * NA + idx] = d;
    for (int c = 0; c < 3; ++c) imp_n[(long long)(3 * k + c) * NA + idx] = nv[c];
  }
  nimp[idx] = n;
}

__global__ void k_self_group_reset(int* __restrict__ gcount, int* __restrict__ glead, long long nfr,
                                   int* __restrict__ ngrp, int* __restrict__ ngrp_prev, const int* __restrict__ cactive,
                                   int R) {
  long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (idx < nfr) {
    gcount[idx] = 0;
    glead[idx] = 0x7fffffff;
  }
  if (idx < R && cactive[idx]) {
    ngrp_prev[idx] = ngrp[idx];
    ngrp[idx] = 0;
  }
}

// Closest impact of each vertex (collisionBrutal.h:648-668: smaller t; with
// t = 1 on the kept one, smaller d), and its registration with the group of
// the face it hit: member slot, group count, lowest member (the leader, whose
// impact normal the group uses).
__global__ void k_self_closest(int ncs, int nsf, int R, const int* __restrict__ cactive, const int* __restrict__ nimp,
                               const int* __restrict__ imp_f, const double* __restrict__ imp_t,
                               const double* __restrict__ imp_d, int* __restrict__ cface,
                               int* __restrict__ gcount, int* __restrict__ glead, int* __restrict__ gmem,
                               int* __restrict__ ngrp, int* __restrict__ flag) {
  long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (idx >= (long long)ncs * R) return;
  int a = (int)(idx / R), r = (int)(idx % R);
  if (!cactive[r]) return;
  const long long NA = (long long)ncs * R;
  int n = nimp[idx], best = -1;
  double bt = 0, bd = 0;
  for (int k = 0; k < n; ++k) {
    double t = imp_t[(long long)k * NA + idx], d = imp_d[(long long)k * NA + idx];
    if (best < 0 || t < bt || (bt == 1 && bd > d)) {
      best = k;
      bt = t;
      bd = d;
    }
  }
  cface[idx] = best;  // slot of the closest impact, -1: none
  if (best < 0) return;
  int f = imp_f[(long long)best * NA + idx];
  long long fr = (long long)f * R + r;
  int slot = atomicAdd(&gcount[fr], 1);
  if (slot < kSelfMem)
    gmem[(long long)slot * nsf * R + fr] = a;
  else
    atomicOr(flag, 1);
  if (slot == 0) atomicAdd(&ngrp[r], 1);
  atomicMin(&glead[fr], a);
}

// One self-contact group (face f, rollout r) of a handling (World.h:1451-1530):
// momentum-exchange among the face and its approaching vertices with the
// stick/slide friction of do_solve_friction (World.h:1246-1315), in the basis
// of the leader's impact normal. Writes the face corners' changes GF and the
// members' reactions GV (zero for members that recede).
__global__ void k_self_group(const int* __restrict__ sc_vtx, const int* __restrict__ sc_face, int ncs, int nsf,
                             int R, const int* __restrict__ iactive, const int* __restrict__ gcount,
                             const int* __restrict__ glead, const int* __restrict__ gmem,
                             const int* __restrict__ cface, const double* __restrict__ imp_n,
                             const double* __restrict__ Fo, const double* __restrict__ CN,
                             const double* __restrict__ Kdiag, double mu, double* __restrict__ GF,
                             double* __restrict__ GV, int* __restrict__ tFlags = nullptr) {
  long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (idx >= (long long)nsf * R) return;
  int f = (int)(idx / R), r = (int)(idx % R);
  if (!iactive[r]) return;
  int cnt = gcount[idx];
  if (cnt == 0) return;
  cnt = min(cnt, kSelfMem);
  const long long C = 3LL * R, NA = (long long)ncs * R;
  int mem[kSelfMem];
  for (int k = 0; k < cnt; ++k) mem[k] = gmem[(long long)k * nsf * R + idx];
  for (int i = 1; i < cnt; ++i)  // ascending contact index
    for (int j = i; j > 0 && mem[j - 1] > mem[j]; --j) {
      int t = mem[j];
      mem[j] = mem[j - 1];
      mem[j - 1] = t;
    }
  int lead = glead[idx];
  long long li = (long long)lead * R + r;
  int ls = cface[li];
  double n[3];
  for (int c = 0; c < 3; ++c) n[c] = imp_n[(long long)(3 * ls + c) * NA + li];
  // computeBasis (collisionBrutal.h:408-426): columns e0, t, e2
  double t[3] = {0.0, n[2], -n[1]};
  double tn = sqrt(t[1] * t[1] + t[2] * t[2]);
  if (tn < 1e-15) {
    t[0] = -n[2]; t[1] = 0.0; t[2] = n[0];
    tn = sqrt(t[0] * t[0] + t[2] * t[2]);
  }
  t[0] /= tn; t[1] /= tn; t[2] /= tn;
  double b[3] = {n[1] * t[2] - n[2] * t[1], n[2] * t[0] - n[0] * t[2], n[0] * t[1] - n[1] * t[0]};
  double bn = sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
  double nn = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
  double E[3][3];  // E[c] = column c of the basis
  for (int r2 = 0; r2 < 3; ++r2) {
    E[0][r2] = n[r2] / nn;
    E[1][r2] = t[r2];
    E[2][r2] = b[r2] / bn;
  }
  auto toloc = [&](const double* w, double* l) {
    for (int c = 0; c < 3; ++c) l[c] = E[c][0] * w[0] + E[c][1] * w[1] + E[c][2] * w[2];
  };
  auto toworld = [&](const double* l, double* w) {
    for (int r2 = 0; r2 < 3; ++r2) w[r2] = E[0][r2] * l[0] + E[1][r2] * l[1] + E[2][r2] * l[2];
  };
  int fv[3] = {sc_vtx[sc_face[3 * f]], sc_vtx[sc_face[3 * f + 1]], sc_vtx[sc_face[3 * f + 2]]};
  double m0 = Kdiag[fv[0]] + Kdiag[fv[1]] + Kdiag[fv[2]];
  double f0w[3] = {0, 0, 0}, f0[3], r0[3] = {0, 0, 0};
  for (int q = 0; q < 3; ++q)
    for (int c = 0; c < 3; ++c) {
      long long o = (long long)fv[q] * C + 3 * r + c;
      f0w[c] += Fo[o] + CN[o];
    }
  toloc(f0w, f0);
  int vl[kSelfMem], nv = 0;
  double ms[kSelfMem], fs[kSelfMem][3], rs[kSelfMem][3];
  bool stick[kSelfMem];
  double sm = 0, smv = 0;
  for (int k = 0; k < cnt; ++k) {
    int vk = sc_vtx[mem[k]];
    double fw[3], fl[3];
    for (int c = 0; c < 3; ++c) {
      long long o = (long long)vk * C + 3 * r + c;
      fw[c] = Fo[o] + CN[o];
    }
    toloc(fw, fl);
    for (int c = 0; c < 3; ++c) GV[(long long)mem[k] * C + 3 * r + c] = 0.0;
    if (fl[0] / Kdiag[vk] - f0[0] / m0 <= 0) {
      vl[nv] = k;
      ms[nv] = Kdiag[vk];
      for (int c = 0; c < 3; ++c) fs[nv][c] = fl[c];
      ++nv;
    }
  }
  for (int k = 0; k < nv; ++k) {
    sm += ms[k];
    smv += fs[k][0];
  }
  for (int k = 0; k < nv; ++k) {
    double fn = (f0[0] + smv) * ms[k] / (m0 + sm) - fs[k][0];
    rs[k][0] = fn;
    rs[k][1] = 0.0;
    rs[k][2] = 0.0;
    r0[0] -= fn;
    stick[k] = true;
  }
  while (true) {
    // do_solve_friction
    double pst[3] = {f0[0], f0[1], f0[2]}, mst = m0, psl[3] = {0, 0, 0}, msl = 0, fnl = 0;
    for (int k = 0; k < nv; ++k) {
      if (stick[k]) {
        for (int c = 0; c < 3; ++c) pst[c] += fs[k][c];
        mst += ms[k];
      } else {
        for (int c = 0; c < 3; ++c) psl[c] += fs[k][c];
        msl += ms[k];
        fnl += rs[k][0];
      }
    }
    double v[3] = {pst[0] / mst, pst[1] / mst, pst[2] / mst};
    r0[1] = v[1] * m0 - f0[1];
    r0[2] = v[2] * m0 - f0[2];
    for (int k = 0; k < nv; ++k)
      if (stick[k]) {
        rs[k][1] = v[1] * ms[k] - fs[k][1];
        rs[k][2] = v[2] * ms[k] - fs[k][2];
      }
    if (msl > 0) {
      double rel[3];
      for (int c = 0; c < 3; ++c) rel[c] = psl[c] / msl - pst[c] / mst;
      double tl = sqrt(rel[1] * rel[1] + rel[2] * rel[2]);
      for (int c = 0; c < 3; ++c) rel[c] /= tl;
      double mult = mu * fnl / (msl + mst);
      r0[1] += rel[1] * mult * m0;
      r0[2] += rel[2] * mult * m0;
      for (int k = 0; k < nv; ++k)
        if (stick[k]) {
          rs[k][1] += rel[1] * mult * ms[k];
          rs[k][2] += rel[2] * mult * ms[k];
        }
      double av[3] = {(f0[0] + r0[0]) / m0, (f0[1] + r0[1]) / m0, (f0[2] + r0[2]) / m0};
      for (int k = 0; k < nv; ++k)
        if (!stick[k]) {
          double rv[3];
          for (int c = 0; c < 3; ++c) rv[c] = fs[k][c] / ms[k] - av[c];
          double t2 = sqrt(rv[1] * rv[1] + rv[2] * rv[2]);
          for (int c = 0; c < 3; ++c) rv[c] /= t2;
          double m = -mu * rs[k][0] * ms[k] / (ms[k] + mst);
          rs[k][1] = m * rv[1];
          rs[k][2] = m * rv[2];
        }
    }
    bool changed = false;
    for (int k = 0; k < nv; ++k)
      if (stick[k] && rs[k][1] * rs[k][1] + rs[k][2] * rs[k][2] > mu * rs[k][0] * rs[k][0]) {
        changed = true;
        stick[k] = false;
      }
    if (!changed) break;
  }
  if (tFlags) {  // the reverse sweep's frozen decisions: members that approached, final stick flags
    int fl = 0;
    for (int k = 0; k < nv; ++k) fl |= (1 << vl[k]) | (stick[k] ? (1 << (16 + vl[k])) : 0);
    tFlags[idx] = fl;
  }
  double r0w[3], bf0[3], v0[3];
  toworld(r0, r0w);
  toworld(f0, bf0);
  for (int c = 0; c < 3; ++c) v0[c] = (bf0[c] + r0w[c]) / m0;
  for (int q = 0; q < 3; ++q)
    for (int c = 0; c < 3; ++c) {
      long long o = (long long)fv[q] * C + 3 * r + c;
      GF[(long long)(3 * f + q) * C + 3 * r + c] = v0[c] * Kdiag[fv[q]] - (Fo[o] + CN[o]);
    }
  for (int k = 0; k < nv; ++k) {
    double w[3];
    toworld(rs[k], w);
    for (int c = 0; c < 3; ++c) GV[(long long)mem[vl[k]] * C + 3 * r + c] = w[c];
  }
}

// End of a handling with self contact (World.h:1585-1600): rhs = forces +
// obstacle reactions + self-contact forces, which are also carried to the next
// handling; then the relaxed Jacobi update of k_contact_jacobi.
__global__ void k_self_finish(const int* __restrict__ scf_ptr, const int* __restrict__ scf_q,
                              const int* __restrict__ sc_of_v, int ncs, int R, const int* __restrict__ gcount,
                              const int* __restrict__ glead, const int* __restrict__ imp_f,
                              const int* __restrict__ cface, const int* __restrict__ ngrp,
                              const double* __restrict__ GF, const double* __restrict__ GV,
                              const double* __restrict__ RC, const double* __restrict__ CN,
                              const double* __restrict__ Fo, double* __restrict__ self_cur, double* __restrict__ Prev,
                              double* __restrict__ Xit, const double* __restrict__ Xn,
                              const double* __restrict__ Kdiag, double alpha, double h, int first_iter,
                              const int* __restrict__ alist, const int* __restrict__ acount, int N) {
  long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)N * R) return;
  const int v = (int)(idx / R), j = (int)(idx % R);
  if (j >= *acount) return;
  const int r = alist[j];
  if (ngrp[r] == 0) return;
  const long long C = 3LL * R, base = (long long)v * C + 3 * r;
  // the vertex's shares from the groups of its incident faces and from its own
  // group, added in group order (ascending leader), as the handling adds them
  double sf[3] = {0, 0, 0};
  int own_lead = 0x7fffffff;
  const int a = sc_of_v[v];
  if (a >= 0) {
    const long long ai = (long long)a * R + r;
    const int slot = cface[ai];
    if (slot >= 0) own_lead = glead[(long long)imp_f[(long long)slot * ncs * R + ai] * R + r];
  }
  for (int last = -1;;) {
    int best = 0x7fffffff, bq = -1;
    for (int q = scf_ptr[v]; q < scf_ptr[v + 1]; ++q) {
      long long fr = (long long)(scf_q[q] / 3) * R + r;
      if (gcount[fr] == 0) continue;
      int L = glead[fr];
      if (L > last && L < best) {
        best = L;
        bq = q;
      }
    }
    if (own_lead > last && own_lead < best) {
      best = own_lead;
      bq = -2;
    }
    if (best == 0x7fffffff) break;
    const double* g = bq == -2 ? GV + (long long)a * C + 3 * r : GF + (long long)scf_q[bq] * C + 3 * r;
    for (int c = 0; c < 3; ++c) sf[c] += g[c];
    last = best;
  }
  double kd = Kdiag[v];
  for (int c = 0; c < 3; ++c) {
    double rc = RC[base + c] + (CN[base + c] + sf[c]);
    self_cur[base + c] = sf[c];
    double pv = first_iter ? Fo[base + c] : Prev[base + c];
    double p = rc * alpha + pv * (1.0 - alpha);
    Prev[base + c] = p;
    Xit[base + c] = Xn[base + c] + (p / h) / kd;
  }
}

// Self-contact groups of a pass for the reverse sweep: counts, leaders,
// members (-1 past the count), the leader's impact normal, each contact
// vertex's own group face (-1: none) and the groups per rollout.
__global__ void k_self_tape_pass(int ncs, int nsf, int R, const int* __restrict__ gcount, const int* __restrict__ glead,
                                 const int* __restrict__ gmem, const int* __restrict__ cface,
                                 const int* __restrict__ imp_f, const double* __restrict__ imp_n,
                                 const int* __restrict__ ngrp, int* __restrict__ tCount, int* __restrict__ tLead,
                                 int* __restrict__ tMem, double* __restrict__ tNrm, int* __restrict__ tOwn,
                                 int* __restrict__ tNgrp) {
  const long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  const long long NA = (long long)ncs * R, NF = (long long)nsf * R;
  if (idx < NF) {
    const int cnt = gcount[idx];
    tCount[idx] = cnt;
    tLead[idx] = glead[idx];
    for (int k = 0; k < kSelfMem; ++k) tMem[(long long)k * NF + idx] = k < cnt ? gmem[(long long)k * NF + idx] : -1;
    double n[3] = {0.0, 0.0, 0.0};
    if (cnt > 0) {
      const long long li = (long long)glead[idx] * R + idx % R;
      const int ls = cface[li];
      if (ls >= 0)
        for (int c = 0; c < 3; ++c) n[c] = imp_n[(long long)(3 * ls + c) * NA + li];
    }
    for (int c = 0; c < 3; ++c) tNrm[3 * idx + c] = n[c];
  }
  if (idx < NA) {
    const int slot = cface[idx];
    tOwn[idx] = slot >= 0 ? imp_f[(long long)slot * NA + idx] : -1;
  }
  if (idx < R) tNgrp[idx] = ngrp[idx];
}

// The forces the groups of an iteration read (rhs + obstacle reactions) at the
// surface vertices, for the reverse sweep.
__global__ void k_self_tape_F(const int* __restrict__ sc_vtx, int ncs, int C, const double* __restrict__ Fo,
                              const double* __restrict__ CN, double* __restrict__ tF) {
  const long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (idx >= (long long)ncs * C) return;
  const int a = (int)(idx / C), col = (int)(idx % C);
  const long long o = (long long)sc_vtx[a] * C + col;
  tF[idx] = Fo[o] + CN[o];
}

// Rollouts with no fresh impact leave the contact phase; the ones that stay
// start an inner Jacobi pass (World.h:938-944).
__global__ void k_contact_pass_begin(int* __restrict__ cactive, int* __restrict__ fresh, int* __restrict__ iactive,
                                     int* __restrict__ npass, int R, const int* __restrict__ ngrp = nullptr,
                                     const int* __restrict__ ngrp_prev = nullptr) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= R) return;
  // with self contact a pass also follows when the number of groups changed
  // (collisionBrutal.h:404)
  int a = cactive[r] && (fresh[r] || (ngrp && ngrp[r] != ngrp_prev[r]));
  cactive[r] = a;
  iactive[r] = a;
  npass[r] += a;
  fresh[r] = 0;
}

// Start of an inner Jacobi iteration: ori_x = x_it, err = max|ori_x - ori_x_old|
// (World.h:945-955). Partial maxima per (chunk, column).
__global__ void k_contact_begin(const double* __restrict__ Xit, double* __restrict__ Ori,
                                const int* __restrict__ iactive, double* __restrict__ part, int N, int C,
                                const double* __restrict__ Xn, double* __restrict__ Dn) {
  __shared__ double red[4][64];
  int col = blockIdx.x * 64 + threadIdx.x;
  int chunk = blockIdx.y;
  double mx = 0.0;
  if (col < C && iactive[col / 3]) {
    int vend = min(N, (chunk + 1) * kChunk);
    int v = chunk * kChunk + threadIdx.y;
    for (; v + 12 < vend; v += 16) {  // four rows per trip, loads issued together
      double x[4], o[4], xn[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const long long i = (long long)(v + 4 * u) * C + col;
        x[u] = Xit[i];
        o[u] = Ori[i];
        xn[u] = Xn[i];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const long long i = (long long)(v + 4 * u) * C + col;
        mx = fmax(mx, fabs(x[u] - o[u]));
        Ori[i] = x[u];
        Dn[i] = x[u] - xn[u];  // ori - x_n, read by the neighbours' Jacobi update
      }
    }
    for (; v < vend; v += 4) {
      long long i = (long long)v * C + col;
      double x = Xit[i];
      mx = fmax(mx, fabs(x - Ori[i]));
      Ori[i] = x;
      Dn[i] = x - Xn[i];
    }
  }
  red[threadIdx.y][threadIdx.x] = mx;
  __syncthreads();
  if (threadIdx.y == 0 && col < C)
    part[(long long)chunk * C + col] =
        fmax(fmax(red[0][threadIdx.x], red[1][threadIdx.x]), fmax(red[2][threadIdx.x], red[3][threadIdx.x]));
}

__global__ __launch_bounds__(256) void k_contact_err(const double* __restrict__ part, int nchunk, int R,
                                                     int* __restrict__ iactive, int* __restrict__ cfirst,
                                                     int* __restrict__ citers, int* __restrict__ clast,
                                                     int* __restrict__ exec) {
  const int r = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (r >= R) return;
  if (!iactive[r]) {
    if (exec && lane == 0) exec[r] = 0;
    return;
  }
  const long long C = 3LL * R;
  double m = 0.0;
  for (int ch = lane; ch < nchunk; ch += 64)
    m = fmax(m, fmax(part[ch * C + 3 * r], fmax(part[ch * C + 3 * r + 1], part[ch * C + 3 * r + 2])));
  m = wave_max(m);
  if (lane != 0) return;
  bool go = cfirst[r] || !(m < kEps);
  cfirst[r] = 0;
  iactive[r] = go ? 1 : 0;
  citers[r] += go ? 1 : 0;
  clast[r] = go ? 0 : 1;  // 1: the result is this begin's ori_x = x_it
  if (exec) exec[r] = go ? 1 : 0;
}

// One Jacobi update of the contact phase per (vertex, rollout), World.h:957-1016:
//   rhs = h (c + J d(ori) - C (ori - x_n))  with C = offdiag(M/h^2 + L),
//   contact vertices: rhs replaced by the cone-projected reaction B r
//   (World.h:1363-1436), prev = alpha rhs + (1-alpha) prev,
//   x_it = x_n + diag(M/h^2 + L)^-1 prev / h.
// UNI: R is a multiple of 64, so a wave's 64 threads share one vertex (its
// matrix row then comes in by scalar loads).
template <bool UNI>
__global__ void k_contact_jacobi(const int* __restrict__ Lptr, const int* __restrict__ Lcol,
                                 const double* __restrict__ Lval, const double* __restrict__ B,
                                 const double* __restrict__ Dn, const double* __restrict__ Xn,
                                 const double* __restrict__ Kdiag, const int* __restrict__ cv_of_v,
                                 const int* __restrict__ win_st, const double* __restrict__ win_n,
                                 double mu, double alpha, double h, int first_iter, const int* __restrict__ alist,
                                 const int* __restrict__ acount, double* __restrict__ Prev, double* __restrict__ Xit,
                                 int N, int R, double* __restrict__ tRC, double* __restrict__ tN,
                                 double* __restrict__ tX, double* __restrict__ self_cur = nullptr,
                                 double* __restrict__ Fo = nullptr,
                                 double* __restrict__ CN = nullptr, double* __restrict__ RC = nullptr,
                                 const int* __restrict__ ngrp = nullptr, const double* __restrict__ Cc = nullptr,
                                 const double* __restrict__ corner = nullptr, const int* __restrict__ vptr = nullptr) {
  // self_cur non-null: self contact is on (World.h:1373-1430 with
  // m_handle_self_collision): the obstacle reaction is computed from the
  // forces plus the previous self-contact forces and kept in CN; it replaces
  // the rhs only while the rollout has no self-contact group; with groups the
  // update is left to k_self_finish (RC = rhs before the self-contact sums).
  long long idx = (long long)xcd_block() * blockDim.x + threadIdx.x;
  if (idx >= (long long)N * R) return;
  int v = (int)(idx / R);
  const int j = (int)(idx % R);
  if (UNI) v = __builtin_amdgcn_readfirstlane(v);
  if (j >= *acount) return;
  const int r = alist[j];
  long long C = 3LL * R;
  long long base = (long long)v * C + 3 * r;
  double rc[3];
  if (corner) {
    // c + J d gathered here from the local step's FP64 per-vertex sums
    // (strip sums or corner slots; the gather's order: entry k of the range
    // to partial sum k mod 4 while four remain, the rest to the first, then
    // (s0 + s1) + (s2 + s3))
    double sp[3][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}};
    int q = vptr[v];
    const int qe = vptr[v + 1];
    const double* p = corner + (long long)q * C + 3 * r;
    for (; q + 4 <= qe; q += 4, p += 4 * C) {
      double x[4][3];
#pragma unroll
      for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int c = 0; c < 3; ++c) x[k][c] = p[k * C + c];
#pragma unroll
      for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int c = 0; c < 3; ++c) sp[c][k] += x[k][c];
    }
    for (; q < qe; ++q, p += C)
#pragma unroll
      for (int c = 0; c < 3; ++c) sp[c][0] += p[c];
#pragma unroll
    for (int c = 0; c < 3; ++c) rc[c] = Cc[base + c] + ((sp[c][0] + sp[c][1]) + (sp[c][2] + sp[c][3]));
  } else {
    rc[0] = B[base];
    rc[1] = B[base + 1];
    rc[2] = B[base + 2];
  }
  {
    // neighbours in row order; the loads of four are issued together
    int q = Lptr[v];
    const int qe = Lptr[v + 1];
    for (; q + 4 <= qe; q += 4) {
      int u[4];
      double w[4], d[4][3];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        u[k] = Lcol[q + k];
        w[k] = Lval[q + k];
        const double* p = Dn + (long long)u[k] * C + 3 * r;
        d[k][0] = p[0];
        d[k][1] = p[1];
        d[k][2] = p[2];
      }
#pragma unroll
      for (int k = 0; k < 4; ++k)
        if (u[k] != v) {
          rc[0] -= w[k] * d[k][0];
          rc[1] -= w[k] * d[k][1];
          rc[2] -= w[k] * d[k][2];
        }
    }
    for (; q < qe; ++q) {
      int u = Lcol[q];
      if (u == v) continue;
      long long ub = (long long)u * C + 3 * r;
      double w = Lval[q];
      rc[0] -= w * Dn[ub];
      rc[1] -= w * Dn[ub + 1];
      rc[2] -= w * Dn[ub + 2];
    }
  }
  rc[0] *= h;
  rc[1] *= h;
  rc[2] *= h;
  double pv[3];
  if (first_iter) {
    pv[0] = rc[0]; pv[1] = rc[1]; pv[2] = rc[2];
  } else {
    pv[0] = Prev[base]; pv[1] = Prev[base + 1]; pv[2] = Prev[base + 2];
  }
  const bool self = self_cur != nullptr;
  const bool grp = self && ngrp[r] > 0;
  if (self) {
    for (int c = 0; c < 3; ++c) {
      Fo[base + c] = rc[c];
      CN[base + c] = 0.0;
    }
  }
  int ci = cv_of_v[v];
  if (ci >= 0 && tRC) {  // the obstacle projection's input (with self contact: + the carried forces)
    long long tb = (long long)ci * C + 3 * r;
    for (int c = 0; c < 3; ++c) tRC[tb + c] = self ? rc[c] + self_cur[base + c] : rc[c];
    if (tN) {  // once per pass: the winners do not change within a pass
      tN[tb] = 0.0;
      tN[tb + 1] = 0.0;
      tN[tb + 2] = 0.0;
    }
  }
  if (ci >= 0) {
    const long long wi = (long long)ci * R + r;
    if (win_st[wi] >= 0) {
      const double* n = win_n + 3 * wi;
      if (tN) {
        long long tb = (long long)ci * C + 3 * r;
        tN[tb] = n[0];
        tN[tb + 1] = n[1];
        tN[tb + 2] = n[2];
      }
      // computeBasis (collisionBrutal.h:408-426): [n, t, b]
      double t[3] = {0.0, n[2], -n[1]};
      double tn = sqrt(t[1] * t[1] + t[2] * t[2]);
      if (tn < 1e-15) {
        t[0] = -n[2]; t[1] = 0.0; t[2] = n[0];
        tn = sqrt(t[0] * t[0] + t[2] * t[2]);
      }
      t[0] /= tn; t[1] /= tn; t[2] /= tn;
      double b[3] = {n[1] * t[2] - n[2] * t[1], n[2] * t[0] - n[0] * t[2], n[0] * t[1] - n[1] * t[0]};
      double bn = sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
      double nn = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
      double e0[3] = {n[0] / nn, n[1] / nn, n[2] / nn};
      double e2[3] = {b[0] / bn, b[1] / bn, b[2] / bn};
      double fin[3] = {rc[0], rc[1], rc[2]};
      if (self)
        for (int c = 0; c < 3; ++c) fin[c] += self_cur[base + c];
      double f0 = e0[0] * fin[0] + e0[1] * fin[1] + e0[2] * fin[2];
      if (!(f0 > 0.0)) {
        double f1 = t[0] * fin[0] + t[1] * fin[1] + t[2] * fin[2];
        double f2 = e2[0] * fin[0] + e2[1] * fin[1] + e2[2] * fin[2];
        double r0 = -f0, r1 = -f1, r2 = -f2;
        double rt = sqrt(r1 * r1 + r2 * r2);
        if (rt > mu * r0) {
          r1 *= mu * r0 / rt;
          r2 *= mu * r0 / rt;
        }
        for (int c = 0; c < 3; ++c) {
          double o = e0[c] * r0 + t[c] * r1 + e2[c] * r2;
          if (self) CN[base + c] = o;
          if (!grp) rc[c] = o;
        }
      }
    }
  }
  if (grp) {
    for (int c = 0; c < 3; ++c) RC[base + c] = rc[c];
    return;
  }
  if (self)
    for (int c = 0; c < 3; ++c) self_cur[base + c] = 0.0;
  double kd = Kdiag[v];
  for (int c = 0; c < 3; ++c) {
    double p = rc[c] * alpha + pv[c] * (1.0 - alpha);
    Prev[base + c] = p;
    const double x = Xn[base + c] + (p / h) / kd;
    Xit[base + c] = x;
    if (tX) tX[base + c] = x;  // the iteration's tape slot
  }
}


// ------------------------------------------------------- contact-phase adjoint
// Adjoint of the cone projection at a contact vertex with frozen branch and
// normal n (zero normal: no impact, identity).
__device__ void cone_vjp(const double* n, double mu, const double* f_in, const double* qbar, double* out) {
  double nn = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
  if (!(nn > 0.0)) {
    out[0] = qbar[0]; out[1] = qbar[1]; out[2] = qbar[2];
    return;
  }
  double t[3] = {0.0, n[2], -n[1]};
  double tn = sqrt(t[1] * t[1] + t[2] * t[2]);
  if (tn < 1e-15) {
    t[0] = -n[2]; t[1] = 0.0; t[2] = n[0];
    tn = sqrt(t[0] * t[0] + t[2] * t[2]);
  }
  t[0] /= tn; t[1] /= tn; t[2] /= tn;
  double b[3] = {n[1] * t[2] - n[2] * t[1], n[2] * t[0] - n[0] * t[2], n[0] * t[1] - n[1] * t[0]};
  double bn = sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
  double e0[3] = {n[0] / nn, n[1] / nn, n[2] / nn}, e2[3] = {b[0] / bn, b[1] / bn, b[2] / bn};
  double f0 = e0[0] * f_in[0] + e0[1] * f_in[1] + e0[2] * f_in[2];
  if (f0 > 0.0) {
    out[0] = qbar[0]; out[1] = qbar[1]; out[2] = qbar[2];
    return;
  }
  double f1 = t[0] * f_in[0] + t[1] * f_in[1] + t[2] * f_in[2];
  double f2 = e2[0] * f_in[0] + e2[1] * f_in[1] + e2[2] * f_in[2];
  double r0 = e0[0] * qbar[0] + e0[1] * qbar[1] + e0[2] * qbar[2];
  double r1 = t[0] * qbar[0] + t[1] * qbar[1] + t[2] * qbar[2];
  double r2 = e2[0] * qbar[0] + e2[1] * qbar[1] + e2[2] * qbar[2];
  double ft = sqrt(f1 * f1 + f2 * f2);
  double g0, g1, g2;
  if (!(ft > mu * (-f0))) {  // stick
    g0 = -r0; g1 = -r1; g2 = -r2;
  } else {  // slide
    double u1 = f1 / ft, u2 = f2 / ft;
    double ur = u1 * r1 + u2 * r2;
    g0 = -r0 + mu * ur;
    double sc = mu * f0 / ft;
    g1 = sc * (r1 - ur * u1);
    g2 = sc * (r2 - ur * u2);
  }
  for (int c = 0; c < 3; ++c) out[c] = e0[c] * g0 + t[c] * g1 + e2[c] * g2;
}

// Adjoint of the cone projection with respect to its unit normal n (branch
// frozen): take-off (f) and stick (-f) replies do not depend on n; the sliding
// reply out = -phi n + mu phi e (phi = n.f, e = f_t/|f_t|, f_t = f - phi n) gives
// nbar = f (-ob.n + mu ob.e - g.n) - phi (ob + g), g = mu phi (ob - e (e.ob))/|f_t|.
__device__ void cone_nvjp(const double* n, double mu, const double* f, const double* ob, double* nbar) {
  nbar[0] = nbar[1] = nbar[2] = 0.0;
  const double phi = n[0] * f[0] + n[1] * f[1] + n[2] * f[2];
  if (phi > 0.0) return;
  double ft[3];
  for (int r = 0; r < 3; ++r) ft[r] = f[r] - phi * n[r];
  const double tau = sqrt(ft[0] * ft[0] + ft[1] * ft[1] + ft[2] * ft[2]);
  if (!(tau > mu * (-phi))) return;
  const double e[3] = {ft[0] / tau, ft[1] / tau, ft[2] / tau};
  const double obn = ob[0] * n[0] + ob[1] * n[1] + ob[2] * n[2];
  const double obe = ob[0] * e[0] + ob[1] * e[1] + ob[2] * e[2];
  const double sc = mu * phi / tau;
  const double coef = -obn + mu * obe - sc * obn;
  for (int r = 0; r < 3; ++r) nbar[r] = f[r] * coef - phi * (ob[r] + sc * (ob[r] - e[r] * obe));
}

// Adjoint of a sphere normal n = sg p/|p|, p = x0 + t (x1 - x0) - c, to x0 and
// x1: t = 1 is fixed; a crossing root of |p(t)| = r moves as dt = -p.q / p.d
// (d = x1 - x0, q = (1-t) dx0 + t dx1), so q's adjoint is w = pbar - p (pbar.d)/(p.d).
__device__ void sphere_normal_vjp(const double* c, const double* x0, const double* x1, double t, int root,
                                  const double* n, const double* nbar, double* x0bar, double* x1bar) {
  double p[3], d[3];
  for (int r = 0; r < 3; ++r) {
    d[r] = x1[r] - x0[r];
    p[r] = x0[r] + t * d[r] - c[r];
  }
  const double len = sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
  const double sg = (n[0] * p[0] + n[1] * p[1] + n[2] * p[2]) > 0.0 ? 1.0 : -1.0;
  const double nn = n[0] * nbar[0] + n[1] * nbar[1] + n[2] * nbar[2];
  double pb[3];
  for (int r = 0; r < 3; ++r) pb[r] = sg * (nbar[r] - n[r] * nn) / len;
  if (!root) {
    for (int r = 0; r < 3; ++r) x1bar[r] = pb[r];
    return;
  }
  const double pd = p[0] * d[0] + p[1] * d[1] + p[2] * d[2];
  const double bd = pb[0] * d[0] + pb[1] * d[1] + pb[2] * d[2];
  for (int r = 0; r < 3; ++r) {
    const double w = pb[r] - p[r] * bd / pd;
    x0bar[r] = (1.0 - t) * w;
    x1bar[r] = t * w;
  }
}

// Adds the adjoint of the positions a detection pass read (NB, per contact
// vertex and rollout) into the running x_it adjoint of the rollouts whose
// detection read x_it (want = 1) or the last iteration's input (want = 0).
__global__ void k_cadj_nb(double* __restrict__ XB, const double* __restrict__ NB, const int* __restrict__ dl, int want,
                          const int* __restrict__ cvtx, long long ni, int R) {
  long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (idx >= ni) return;
  const int ci = (int)(idx / R), r = (int)(idx % R);
  if (dl[r] != want) return;
  const long long base = (long long)cvtx[ci] * 3 * R + 3 * r;
  for (int c = 0; c < 3; ++c) XB[base + c] += NB[3 * idx + c];
}

__global__ void k_cadj_init(double* __restrict__ XB, const double* __restrict__ Yb, const int* __restrict__ flast,
                            long long n, int C) {
  long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (idx >= n) return;
  XB[idx] = flast[(int)(idx % C) / 3] ? Yb[idx] : 0.0;
}

// Reverse of x_it = x_n + D^-1 prev / h, prev = alpha q + (1-alpha) P and
// q = cone(rc) for one executed iteration, per (vertex, rollout). Produces
// z = h rc_bar, the D-dependence coefficient CD for the E-gradient, and the
// running adjoints of x_n (Xb), prev (PB) and c (Zs).
template <bool UNI>
__global__ void k_cadj_pre(const double* __restrict__ XB, double* __restrict__ PB, double* __restrict__ Xb,
                           const double* __restrict__ Xit1, const double* __restrict__ Xn,
                           const double* __restrict__ Kdiag, const int* __restrict__ cv_of_v,
                           const double* __restrict__ tRC, const double* __restrict__ tN, double mu, double alpha,
                           double h, int first, const int* __restrict__ alist, const int* __restrict__ acount,
                           double* __restrict__ Zb, double* __restrict__ Zs, double* __restrict__ CD, int N, int R,
                           const double* __restrict__ spheres, int nsph, const int* __restrict__ tSi,
                           const double* __restrict__ tSt, const double* __restrict__ tSx1, double* __restrict__ NB,
                           long long nbstride) {
  long long idx = (long long)xcd_block() * blockDim.x + threadIdx.x;
  if (idx >= (long long)N * R) return;
  int v = (int)(idx / R);
  const int j = (int)(idx % R);
  if (UNI) v = __builtin_amdgcn_readfirstlane(v);
  if (j >= *acount) return;  // entries of other rollouts are never read
  const int r = alist[j];
  long long C = 3LL * R;
  long long base = (long long)v * C + 3 * r;
  double kd = Kdiag[v];
  double pb[3], cd = 0.0;
  for (int c = 0; c < 3; ++c) {
    double xb = XB[base + c];
    Xb[base + c] += xb;
    pb[c] = PB[base + c] + xb / (kd * h);
    cd -= xb * (Xit1[base + c] - Xn[base + c]) / kd;
  }
  double q[3] = {alpha * pb[0], alpha * pb[1], alpha * pb[2]}, rcb[3] = {q[0], q[1], q[2]};
  int ci = cv_of_v[v];
  if (ci >= 0) {
    long long tb = (long long)ci * C + 3 * r;
    cone_vjp(tN + tb, mu, tRC + tb, q, rcb);
    const long long wi = (long long)ci * R + r;
    const int si = tSi ? tSi[wi] : 0;
    if (si > 0) {  // sphere normal: it moves with x_n and the position its detection read
      double nb[3];
      cone_nvjp(tN + tb, mu, tRC + tb, q, nb);
      if (nb[0] != 0.0 || nb[1] != 0.0 || nb[2] != 0.0) {
        const int k = (si - 1) >> 2, o = k % nsph, ps = k / nsph;
        double x0b[3] = {0.0, 0.0, 0.0}, x1b[3];
        sphere_normal_vjp(spheres + 4 * o, Xn + base, tSx1 + 3 * wi, tSt[wi], ((si - 1) >> 1) & 1, tN + tb, nb, x0b,
                          x1b);
        for (int c = 0; c < 3; ++c) {
          Xb[base + c] += x0b[c];
          NB[ps * nbstride + tb + c] += x1b[c];
        }
      }
    }
  }
  for (int c = 0; c < 3; ++c) {
    if (first) {
      rcb[c] += (1.0 - alpha) * pb[c];
      PB[base + c] = 0.0;
    } else {
      PB[base + c] = (1.0 - alpha) * pb[c];
    }
    double z = h * rcb[c];
    Zb[base + c] = z;
    Zs[base + c] += z;
  }
  CD[(long long)v * R + r] = cd;
}

// End of one executed contact iteration's reverse, per (vertex, column) of
// the listed rollouts. rc = h (c + J d(ori) - C (ori - x_n)) with C = offdiag
// of the scalar matrix without masses gives s = (C z)_v: Xb += s. The adjoint
// of x_it_{g-1} (= ori) is XB = (-s + the VJP's corner sums at v) + the
// final-position adjoint Gf for rollouts whose result is this iteration's
// input. The neighbour sum has the row order, loads four at a time; the
// corner sums the gather's partial-sum order. UNI: 3R is a multiple of 64 (one
// vertex per wave).
template <typename T, bool UNI>
__global__ __launch_bounds__(256) void k_cadj_gather(const int* __restrict__ Lptr, const int* __restrict__ Lcol,
                                                     const double* __restrict__ Lval, const double* __restrict__ Zb,
                                                     double* __restrict__ Xb, const T* __restrict__ corner,
                                                     const int* __restrict__ vptr, const double* __restrict__ Gf,
                                                     const int* __restrict__ inj, const int* __restrict__ alist,
                                                     const int* __restrict__ acount, double* __restrict__ XB, int N,
                                                     int R) {
  const long long C = 3LL * R;
  long long idx = (long long)xcd_block() * blockDim.x + threadIdx.x;
  if (idx >= (long long)N * C) return;
  int v = (int)(idx / C);
  const int jj = (int)(idx - (long long)v * C);
  if (UNI) v = __builtin_amdgcn_readfirstlane(v);
  if (jj >= 3 * *acount) return;
  const int r = alist[jj / 3];
  const int col = 3 * r + jj % 3;
  const long long o = (long long)v * C + col;
  double cz = 0.0;
  {
    int q = Lptr[v];
    const int qe = Lptr[v + 1];
    for (; q + 4 <= qe; q += 4) {
      int u[4];
      double w[4], z[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        u[k] = Lcol[q + k];
        w[k] = Lval[q + k];
        z[k] = Zb[(long long)u[k] * C + col];
      }
#pragma unroll
      for (int k = 0; k < 4; ++k)
        if (u[k] != v) cz += w[k] * z[k];
    }
    for (; q < qe; ++q) {
      int u = Lcol[q];
      if (u != v) cz += Lval[q] * Zb[(long long)u * C + col];
    }
  }
  Xb[o] += cz;
  double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
  int q = vptr[v], qe = vptr[v + 1];
  const T* p = corner + (long long)q * C + col;
  for (; q + 4 <= qe; q += 4, p += 4 * C) {
    double x[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) x[u] = (double)p[u * C];
    s0 += x[0];
    s1 += x[1];
    s2 += x[2];
    s3 += x[3];
  }
  for (; q < qe; ++q, p += C) s0 += (double)p[0];
  XB[o] = (-cz + ((s0 + s1) + (s2 + s3))) + (inj[r] ? Gf[o] : 0.0);
}

// ---------------------------------------- contact-phase adjoint, self contact
// With self contact on (World.h:1373-1430, 1446-1600) an iteration's rhs is
// forces + obstacle reactions + self-contact forces while the rollout has a
// group, else the obstacle reaction replaces the forces; the obstacle
// projection reads forces + the previous handling's self-contact forces
// (carried, adjoint SCB). Group members, stick flags and bases are frozen as
// executed; masses (diag of K_s) and forces are differentiated.

// First part of one iteration's reverse: QB = alpha * adjoint of prev (the rhs
// adjoint), QF = the first iteration's (1 - alpha) share that goes to the
// forces, the running prev adjoint, x_n's adjoint and the K_s-diagonal
// coefficient of the Jacobi update.
template <bool UNI>
__global__ void k_sadj_pre(const double* __restrict__ XB, double* __restrict__ PB, double* __restrict__ Xb,
                           const double* __restrict__ Xit1, const double* __restrict__ Xn,
                           const double* __restrict__ Kdiag, double alpha, double h, int first,
                           const int* __restrict__ alist, const int* __restrict__ acount, double* __restrict__ QB,
                           double* __restrict__ QF, double* __restrict__ CD, int N, int R) {
  long long idx = (long long)xcd_block() * blockDim.x + threadIdx.x;
  if (idx >= (long long)N * R) return;
  int v = (int)(idx / R);
  const int j = (int)(idx % R);
  if (UNI) v = __builtin_amdgcn_readfirstlane(v);
  if (j >= *acount) return;
  const int r = alist[j];
  const long long base = (long long)v * 3 * R + 3 * r;
  const double kd = Kdiag[v];
  double cd = 0.0;
  for (int c = 0; c < 3; ++c) {
    const double xb = XB[base + c];
    Xb[base + c] += xb;
    const double pb = PB[base + c] + xb / (kd * h);
    cd -= xb * (Xit1[base + c] - Xn[base + c]) / kd;
    QB[base + c] = alpha * pb;
    QF[base + c] = first ? (1.0 - alpha) * pb : 0.0;
    PB[base + c] = first ? 0.0 : (1.0 - alpha) * pb;
  }
  CD[(long long)v * R + r] = cd;
}

// Reverse of one group (face f, rollout r) that acted in this iteration:
// output adjoints QB + SCB at its vertices; input adjoints into per-group
// slots (face corners: SFb/SFm; members: SVb/SVm, zero for members that
// receded), summed per vertex by k_sadj_finish.
__global__ void k_sadj_group(const int* __restrict__ sc_vtx, const int* __restrict__ sc_face, int ncs, int nsf, int R,
                             const int* __restrict__ exec, const int* __restrict__ tNgrp,
                             const int* __restrict__ tCount, const int* __restrict__ tMem,
                             const double* __restrict__ tNrm, const int* __restrict__ tFlags,
                             const double* __restrict__ tF, const double* __restrict__ Kdiag, double mu,
                             const double* __restrict__ QB, const double* __restrict__ SCB, double* __restrict__ SFb,
                             double* __restrict__ SFm, double* __restrict__ SVb, double* __restrict__ SVm) {
  const long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (idx >= (long long)nsf * R) return;
  const int f = (int)(idx / R), r = (int)(idx % R);
  if (!exec[r] || tNgrp[r] == 0) return;
  const int cnt = min(tCount[idx], kSelfMem);
  if (cnt == 0) return;
  const long long C = 3LL * R, NF = (long long)nsf * R;
  int mem[kSelfMem];
  for (int k = 0; k < cnt; ++k) mem[k] = tMem[(long long)k * NF + idx];
  for (int i = 1; i < cnt; ++i)  // ascending contact index, as the forward orders them
    for (int j = i; j > 0 && mem[j - 1] > mem[j]; --j) {
      int t = mem[j];
      mem[j] = mem[j - 1];
      mem[j - 1] = t;
    }
  const double* n = tNrm + 3 * idx;
  double t[3] = {0.0, n[2], -n[1]};
  double tn = sqrt(t[1] * t[1] + t[2] * t[2]);
  if (tn < 1e-15) {
    t[0] = -n[2]; t[1] = 0.0; t[2] = n[0];
    tn = sqrt(t[0] * t[0] + t[2] * t[2]);
  }
  t[0] /= tn; t[1] /= tn; t[2] /= tn;
  double b[3] = {n[1] * t[2] - n[2] * t[1], n[2] * t[0] - n[0] * t[2], n[0] * t[1] - n[1] * t[0]};
  const double bn = sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
  const double nn = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
  double E[3][3];
  for (int c = 0; c < 3; ++c) {
    E[0][c] = n[c] / nn;
    E[1][c] = t[c];
    E[2][c] = b[c] / bn;
  }
  const int fl = tFlags[idx];
  int fa[3], fv[3];
  double Ff[3][3], mf[3], GFb[3][3];
  for (int q = 0; q < 3; ++q) {
    fa[q] = sc_face[3 * f + q];
    fv[q] = sc_vtx[fa[q]];
    mf[q] = Kdiag[fv[q]];
    for (int c = 0; c < 3; ++c) {
      Ff[q][c] = tF[(long long)fa[q] * C + 3 * r + c];
      const long long o = (long long)fv[q] * C + 3 * r + c;
      GFb[q][c] = QB[o] + SCB[o];
    }
  }
  int nv = 0, who[kSelfMem];
  double Fm[kSelfMem][3], mm[kSelfMem], GVb[kSelfMem][3];
  bool stick[kSelfMem];
  for (int k = 0; k < cnt; ++k) {
    if (!((fl >> k) & 1)) continue;
    const int vk = sc_vtx[mem[k]];
    who[nv] = k;
    mm[nv] = Kdiag[vk];
    stick[nv] = (fl >> (16 + k)) & 1;
    for (int c = 0; c < 3; ++c) {
      Fm[nv][c] = tF[(long long)mem[k] * C + 3 * r + c];
      const long long o = (long long)vk * C + 3 * r + c;
      GVb[nv][c] = QB[o] + SCB[o];
    }
    ++nv;
  }
  double Ffb[3][3] = {}, Fmb[kSelfMem][3] = {}, mfb[3] = {0.0, 0.0, 0.0}, mmb[kSelfMem] = {};
  pdmself::group_reverse(E, nv, mf, mm, Ff, Fm, stick, mu, GFb, GVb, Ffb, Fmb, mfb, mmb);
  for (int q = 0; q < 3; ++q) {
    for (int c = 0; c < 3; ++c) SFb[(long long)(3 * f + q) * C + 3 * r + c] = Ffb[q][c];
    SFm[(long long)(3 * f + q) * R + r] = mfb[q];
  }
  for (int k = 0; k < cnt; ++k) {
    for (int c = 0; c < 3; ++c) SVb[(long long)mem[k] * C + 3 * r + c] = 0.0;
    SVm[(long long)mem[k] * R + r] = 0.0;
  }
  for (int kk = 0; kk < nv; ++kk) {
    const int a = mem[who[kk]];
    for (int c = 0; c < 3; ++c) SVb[(long long)a * C + 3 * r + c] = Fmb[kk][c];
    SVm[(long long)a * R + r] = mmb[kk];
  }
}

// Second part of one iteration's reverse, per (vertex, listed rollout): the
// groups' shares summed in the forward's order (ascending leader), the
// obstacle projection's reverse (its output adjoint: the rhs adjoint plus the
// groups' while the rollout has a group), the forces' adjoint z = h * (...)
// for the element VJP, and the carried forces' adjoint SCB.
__global__ void k_sadj_finish(const int* __restrict__ scf_ptr, const int* __restrict__ scf_q,
                              const int* __restrict__ sc_of_v, int R, const int* __restrict__ tNgrp,
                              const int* __restrict__ tCount, const int* __restrict__ tLead,
                              const int* __restrict__ tOwn, const double* __restrict__ SFb,
                              const double* __restrict__ SFm, const double* __restrict__ SVb,
                              const double* __restrict__ SVm, const double* __restrict__ QB,
                              const double* __restrict__ QF, const int* __restrict__ cv_of_v,
                              const double* __restrict__ tRC, const double* __restrict__ tN, double mu, double h,
                              const double* __restrict__ Xn, const double* __restrict__ spheres, int nsph,
                              const int* __restrict__ tSi, const double* __restrict__ tSt,
                              const double* __restrict__ tSx1, double* __restrict__ NB, long long nbstride,
                              const int* __restrict__ alist, const int* __restrict__ acount,
                              double* __restrict__ Xb, double* __restrict__ CD, double* __restrict__ SCB,
                              double* __restrict__ Zb, double* __restrict__ Zs, int N) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)N * R) return;
  const int v = (int)(idx / R), j = (int)(idx % R);
  if (j >= *acount) return;
  const int r = alist[j];
  const long long C = 3LL * R, base = (long long)v * C + 3 * r;
  const bool grp = tNgrp[r] > 0;
  double og[3] = {QB[base], QB[base + 1], QB[base + 2]};
  if (grp) {
    double sf[3] = {0.0, 0.0, 0.0}, sm = 0.0;
    int own_lead = 0x7fffffff;
    const int a = sc_of_v[v];
    if (a >= 0) {
      const int of = tOwn[(long long)a * R + r];
      if (of >= 0) own_lead = tLead[(long long)of * R + r];
    }
    for (int last = -1;;) {
      int best = 0x7fffffff, bq = -1;
      for (int q = scf_ptr[v]; q < scf_ptr[v + 1]; ++q) {
        const long long fr = (long long)(scf_q[q] / 3) * R + r;
        if (tCount[fr] == 0) continue;
        const int L = tLead[fr];
        if (L > last && L < best) {
          best = L;
          bq = q;
        }
      }
      if (own_lead > last && own_lead < best) {
        best = own_lead;
        bq = -2;
      }
      if (best == 0x7fffffff) break;
      if (bq == -2) {
        for (int c = 0; c < 3; ++c) sf[c] += SVb[(long long)a * C + 3 * r + c];
        sm += SVm[(long long)a * R + r];
      } else {
        for (int c = 0; c < 3; ++c) sf[c] += SFb[(long long)scf_q[bq] * C + 3 * r + c];
        sm += SFm[(long long)scf_q[bq] * R + r];
      }
      last = best;
    }
    for (int c = 0; c < 3; ++c) og[c] += sf[c];
    CD[(long long)v * R + r] += sm;
  }
  double scb[3] = {0.0, 0.0, 0.0};
  const int ci = cv_of_v[v];
  if (ci >= 0) {
    const long long tb = (long long)ci * C + 3 * r;
    const double* n = tN + tb;
    const double* fin = tRC + tb;
    const double nn = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    if (nn > 0.0 && !((n[0] * fin[0] + n[1] * fin[1] + n[2] * fin[2]) / nn > 0.0)) {
      const double qb[3] = {og[0], og[1], og[2]};  // grp: rhs + groups' adjoint; else the rhs adjoint
      double o[3];
      cone_vjp(n, mu, fin, qb, o);
      for (int c = 0; c < 3; ++c) {
        og[c] = grp ? og[c] + o[c] : o[c];
        scb[c] = o[c];
      }
      const long long wi = (long long)ci * R + r;
      const int si = tSi ? tSi[wi] : 0;
      if (si > 0) {
        double nb[3];
        cone_nvjp(n, mu, fin, qb, nb);
        if (nb[0] != 0.0 || nb[1] != 0.0 || nb[2] != 0.0) {
          const int k = (si - 1) >> 2, os = k % nsph, ps = k / nsph;
          double x0b[3] = {0.0, 0.0, 0.0}, x1b[3];
          sphere_normal_vjp(spheres + 4 * os, Xn + base, tSx1 + 3 * wi, tSt[wi], ((si - 1) >> 1) & 1, n, nb, x0b,
                            x1b);
          for (int c = 0; c < 3; ++c) {
            Xb[base + c] += x0b[c];
            NB[ps * nbstride + tb + c] += x1b[c];
          }
        }
      }
    }
  }
  for (int c = 0; c < 3; ++c) {
    const double z = h * (og[c] + QF[base + c]);
    Zb[base + c] = z;
    Zs[base + c] += z;
    SCB[base + c] = scb[c];
  }
}

// ------------------------------------------------------------------ engine
}  // namespace

// Working memory of pdm_solve_noalloc: one fixed array per GPU, held until the
// process ends. Host code must not pass `scratch` itself to a kernel (on the
// host it names a stand-in, not device memory); it uses the device address
// that noalloc_scratch() obtains once per device with hipGetSymbolAddress.
static __device__ __align__(256) double scratch[PDM_NOALLOC_SCRATCH_BYTES / sizeof(double)];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device (looked up on the first
// call for that device, with that device current, then kept).
int noalloc_scratch(double** out) {
  int dev = -1;
  HIPX(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) {
    g_err = "device index out of range for the no-allocation scratch";
    return 1;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr) {
      g_err = std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e);
      return 1;
    }
    g_scratch_ptr[dev] = (double*)p;
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

template <typename X>
struct DevBuf {
  X* p = nullptr;
  size_t n = 0;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  int alloc(size_t count) {
    if (count <= n && p) return 0;
    release();
    if (count == 0) return 0;
    hipError_t e = hipMalloc((void**)&p, count * sizeof(X));
    if (e != hipSuccess) {
      g_err = std::string("hipMalloc failed: ") + hipGetErrorString(e);
      p = nullptr;
      return 1;
    }
    n = count;
    return 0;
  }
  // h may be null (an empty host vector padded to one element): zero-filled
  int upload(const X* h, size_t count) {
    if (alloc(count)) return 1;
    if (count) {
      hipError_t e = h ? hipMemcpy(p, h, count * sizeof(X), hipMemcpyHostToDevice) : hipMemset(p, 0, count * sizeof(X));
      if (e != hipSuccess) {
        g_err = std::string("hipMemcpy failed: ") + hipGetErrorString(e);
        return 1;
      }
    }
    return 0;
  }
};

// A right-hand-side buffer of the solve ([N][C] doubles): 32 more rows and 32
// doubles, zeroed, so that the dense leaf products may read past the last
// leaf's own rows (they meet zero operator rows there) and past the last
// column without leaving the buffer or meeting a non-finite value.
int alloc_rhs(DevBuf<double>& b, size_t nc, int C, hipStream_t st) {
  const size_t tail = 32 * (size_t)C + 32;
  if (b.alloc(nc + tail)) return 1;
  HIPX(hipMemsetAsync(b.p + nc, 0, tail * sizeof(double), st));
  return 0;
}

// Grow a device buffer to `count` elements keeping its first `keep` elements.
// The old buffer is freed at once, after the copy out of it: keeping it to
// the end of the call would hold both generations of a whole-rollout contact
// tape at the same time. Growth happens only when a step needs more contact
// iterations than any step before it (in practice the first call of a batch).
template <typename X>
int grow_keep(DevBuf<X>& b, size_t count, size_t keep, hipStream_t st) {
  if (count <= b.n && b.p) return 0;
  DevBuf<X> nb;
  if (nb.alloc(count)) return 1;
  if (keep && b.p) HIPX(hipMemcpyAsync(nb.p, b.p, keep * sizeof(X), hipMemcpyDeviceToDevice, st));
  HIPX(hipStreamSynchronize(st));
  std::swap(b.p, nb.p);
  std::swap(b.n, nb.n);
  return 0;
}

// Tape of the contact phase of one step: every launched Jacobi iteration g
// (all passes) keeps x_it after it, the pre-projection rhs of each contact
// vertex and which rollouts executed it; each detection pass keeps the winning
// normals (a rollout that executes any iteration of a pass executes its first,
// which writes them).
struct ContactTape {
  int G = 0;
  std::vector<char> first;     // iteration opened a detection pass
  DevBuf<double> xit;          // (G+1) x NC, xit[0] = PD result
  int P = 0;                   // detection passes taped
  std::vector<int> pass_of;    // pass of each iteration
  DevBuf<double> rc;           // G x ncv x C
  DevBuf<double> nrm;          // P x ncv x C: the winning normals of each pass
  DevBuf<int> exec;            // G x R
  DevBuf<int> flast;           // R: result is x_it of the last executed iteration
  // with spheres, per pass: the winners' sphere data (code, contact time,
  // position the detection read; ncv x R, 3 ncv x R) and per rollout whether
  // that position was x_it (1) or the last iteration's input (0)
  DevBuf<int> sinfo, dl;
  DevBuf<double> st, sx1;
  // with self contact, per pass: group counts, leaders, members, leader
  // normals, each contact vertex's own group face, groups per rollout; per
  // iteration: the forces the groups read (surface vertices) and each group's
  // frozen decisions (approached / stick bits)
  DevBuf<int> gCount, gLead, gMem, gOwn, gNgrp, gFlags;
  DevBuf<double> gNrm, gF;
};

template <typename X>
std::vector<X> soa(const std::vector<X>& aos, int n, int k) {
  std::vector<X> o((size_t)n * k);
  for (int i = 0; i < n; ++i)
    for (int q = 0; q < k; ++q) o[(size_t)q * n + i] = aos[(size_t)i * k + q];
  return o;
}

}  // namespace

struct pdm_scene_s {
  HostScene S;
  Csr Ls;
  Factor F;
  Corners Cn;
  Strips Sp;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  int R = 1, C = 3, precision = PDM_FP64, flags = PDM_HOST_EARLY_EXIT, ckpt = 20;
  // static device data
  DevBuf<int> perm, tv, tslot, vptr, mv, mptr, mgroup, gptr, gmus, tet_order;
  DevBuf<double> mass, attw, xrest, tG, wr, wv, dEr, dEv, mg, mw;
  DevBuf<int> Lptr, Lcol;
  DevBuf<double> Lval;
  // factor
  DevBuf<int> f_first, f_rows, f_child_ptr, f_child, f_rel, f_lev;
  DevBuf<long long> f_rowptr, f_lptr, f_relptr, f_uptr;
  DevBuf<double> f_W, f_WR;
  // device re-factorisation plan, built once from the symbolic analysis: the
  // contributions of every stored L entry, the front position of every CSR
  // entry and diagonal, update-block offsets, and per level the extend and
  // update work items (fc_ext_ptr / fc_upd_ptr)
  bool fac_plan = false;
  DevBuf<long long> fc_cptr, fc_amap, fc_dpos, fc_upos;
  DevBuf<int> fc_ctet, fc_fail;
  DevBuf<double> fc_cval, fc_L, fc_U;
  DevBuf<int2> fc_ext;
  DevBuf<int4> fc_upd;
  std::vector<int> fc_ext_ptr, fc_upd_ptr, fc_lev_maxm;
  long long fc_usize = 0;
  std::vector<int> lev_ptr, lev_rows;  // lev_rows: max (mpad + wpad) over the level's supernodes
  // grouped level kernels: per sweep (0 fwd, 1 bwd) the work items of every
  // level (supernode, tile range) at item_ptr[pass][l], and per level the
  // largest front rows and W-slice doubles of one item (their LDS need)
  DevBuf<long long> f_pbase;
  DevBuf<int> f_maxp, f_ptab;
  DevBuf<int4> f_items[2];
  std::vector<int> item_ptr[2], item_front[2], item_wsl[2];
  // solve launch shape per level and sweep (autotuned per right-hand-side count)
  struct LevelCfg {
    int Ct = 16, nthr = 256;
  };
  std::vector<LevelCfg> cfg_fwd, cfg_bwd;
  int tuned_C = -1;
  bool autotune = true;
  // dynamic
  DevBuf<double> X, V, Xn, Cc, Xit, B, Lx, U, Ys, part, err;  // Ys: y of the grouped forward sweep
  DevBuf<char> corner;
  // strip sums of the forward local step (wave-uniform scenes, R % 64 == 0)
  DevBuf<char> ssum;
  DevBuf<int> stet, scode, sfbase, ssum_of, ssptr, sfptr, sfq;
  DevBuf<int> active, iters_dev;
  std::vector<int> iters_host;  // R*steps of the last forward
  int last_steps = 0;
  // tape
  DevBuf<double> ckX, ckV;      // checkpoints
  DevBuf<double> tY, tXn, tVn;  // segment tape (tVn: step-start velocities, drag only)
  DevBuf<int> tIters;           // per step of segment: R
  // per forward call: PD iterations, contact passes and Jacobi iterations per
  // (step, rollout); the reduced E gradient of the backward. Kept between
  // calls: a free inside a call would synchronise the device, stalling other
  // scenes' streams.
  DevBuf<int> it_all, np_all, ci_all;
  DevBuf<double> gEd;
  DevBuf<double> act_dev;       // activations of the recorded forward
  bool have_tape = false;
  int tape_steps = 0;
  bool act_null = true;
  // self contact: surface mesh, per (vertex, rollout) impact lists and closest
  // impact, per (face, rollout) groups, and the handling's force buffers
  int ncs = 0, nsf = 0;
  DevBuf<int> d_sc_vtx, d_sc_face, d_sc_of_v, d_scf_ptr, d_scf_q;
  DevBuf<int> s_nimp, s_impf, s_cface, s_gcount, s_glead, s_gmem, s_ngrp, s_ngrp_prev;
  DevBuf<double> s_impt, s_impd, s_impn, s_Fo, s_CN, s_RC, s_cur, s_GF, s_GV;
  // underwater drag: contour faces, incident faces per vertex, face forces and
  // the reverse's per-corner / per-face gradients
  int nfaces = 0;
  DevBuf<int> dcf, dvfptr, dvfq;
  DevBuf<double> FF, GXf, GWf;
  // contact
  int ncv = 0, nobs = 0;
  DevBuf<int> cvtx, cv_of_v, win_st, win_o, nseen, seen, cactive, iactive, cfirst, fresh, npass, citers, clast, cinj;
  int ntri = 0;
  DevBuf<double> d_tris, d_bvh_box;
  DevBuf<int> d_bvh_a, d_bvh_b, d_bvh_tri;
  std::vector<std::unique_ptr<ContactTape>> ctape;  // per step of the segment being reversed
  DevBuf<double> cXB, cPB, cGf, cCD, cNB;
  DevBuf<double> cQB, cQF, cSCB, sSFb, sSFm, sSVb, sSVm;  // self-contact reverse
  DevBuf<int> win_sph;
  DevBuf<double> win_t, win_x1;
  DevBuf<int> cinjG;          // contact reverse: injection flags per (iteration, rollout)
  std::vector<int> cinj_h;
  DevBuf<double> Kdiag, d_planes, d_spheres, Ori, Dn, Prev, win_n;
  std::vector<int> npass_host, citers_host;  // R*steps of the last forward
  // adjoint buffers
  DevBuf<double> gX, gV, Yb, Xb, Zs, Zb, LZ, gEpart, gMpart, gact;
  DevBuf<double> gEw;  // E-gradient sums per (tet, wave of 64 listed rollouts) of the wave-uniform VJP
  DevBuf<int> ractive;
  double t_fwd = 0, t_bwd = 0, t_factor = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  int fwd_taped_seg = -1;  // checkpoint segment whose tapes the forward pass wrote
  DevBuf<int> alist, acount;  // rollouts that still iterate (k_compact)
  // subtree cut: levels below `cut` run as whole subtrees (k_*_subtree)
  int cut = 0;
  LevelCfg cfg_sub;
  std::vector<std::vector<int2>> sub_ranges;  // per cut level: (first, last) supernode of each subtree
  DevBuf<int2> d_sub;
  int d_sub_cut = -1;
  // split solve: column chains on auxiliary streams (created on first use)
  int nsplit = 1;
  hipStream_t aux[kMaxSplit - 1] = {};
  hipEvent_t ev_fork = nullptr, ev_join[kMaxSplit - 1] = {};
  int ensure_aux_streams() {
    if (ev_fork) return 0;
    for (int k = 0; k < kMaxSplit - 1; ++k) {
      if (hipStreamCreateWithFlags(&aux[k], hipStreamNonBlocking) != hipSuccess ||
          hipEventCreateWithFlags(&ev_join[k], hipEventDisableTiming) != hipSuccess) {
        g_err = "cannot create the solve streams";
        return 1;
      }
    }
    if (hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming) != hipSuccess) {
      g_err = "cannot create the solve streams";
      return 1;
    }
    return 0;
  }
  // PDM_USE_GRAPH: captured step block and its device-side controls
  hipGraphExec_t gexec[2] = {nullptr, nullptr};  // untaped / taped step block
  long long graph_launches = 0;                  // step-block replays since the scene was made
  const double* g_act = nullptr;
  int g_act_steps = -1;
  bool capturing = false;
  DevBuf<int> any_dev, step_dev;
  // PDM_TIME_SOLVES: event pairs around each solve, folded into solve_ms
  std::vector<hipEvent_t> sev;
  size_t sev_used = 0;
  long long n_solves = 0;
  double solve_ms = 0.0;

  // dense separator blocks (pdm_top.h): per tree with a top, its padded
  // offset a0, rows n and padded rows np (a multiple of 32), the offset of its
  // np x np block; per padded top row its engine row (-1: padding) and the
  // update rows that add into it; the elimination waves of the factor-time
  // inverse; the product's work items (block, 32-row tile)
  int top_cap = kTopCapDefault;
  bool top_on = false;
  std::vector<char> is_top;
  int top_rows = 0, top_maxp = 0;
  std::vector<int> top_a0, top_n, top_np;
  std::vector<long long> top_zoff;
  long long top_zsize = 0, top_tsize = 0;
  std::vector<int> top_wave_ptr, top_wave_sn, top_wave_blk;
  DevBuf<int> t_row, t_slot, t_a0, t_np, t_frow;
  DevBuf<long long> t_zoff, t_toff;
  DevBuf<int4> t_wave;
  DevBuf<int2> t_items;
  int t_nitems = 0;
  DevBuf<double> tZ, tM, tT, TF;  // Z, Linv, elimination scratch, fronts [top_rows][C]
  // dense leaf subtrees (pdm_top.h): per leaf its root, own rows (j0, nsub,
  // padded np), above rows (nab, padded nabp), offsets of its operators and
  // of its above rows in XA; the elimination waves; the products' work items
  int leaf_cap = kLeafCapDefault;
  bool leaf_on = false;
  std::vector<char> is_leaf;
  std::vector<int> leaf_np, leaf_nabp, leaf_nsub, leaf_nab, leaf_j0;
  std::vector<long long> leaf_moff, leaf_poff, leaf_zoff, leaf_uoff, leaf_xaoff;
  long long leaf_msize = 0, leaf_psize = 0, leaf_zsize = 0, leaf_tsize = 0;
  int leaf_xrows = 0;
  std::vector<int> leaf_wave_ptr, leaf_wave_sn, leaf_wave_blk;
  DevBuf<int> l_np, l_nabp, l_nsub, l_nab, l_j0, l_frow, l_xrow;
  DevBuf<long long> l_moff, l_poff, l_uoff, l_xaoff, l_toff;
  DevBuf<int4> l_wave;
  DevBuf<int2> l_items[2];
  int l_nitems[2] = {0, 0};
  DevBuf<char> d_skip;             // per supernode: inside a dense leaf (the subtree kernels skip it)
  DevBuf<double> lM, lZ, lP, lT, XA;  // [Linv; G], Zs, forward operators, scratch, above rows [xrows][C]
  LeafView lview() const {
    return LeafView{lP.p, lM.p, l_poff.p, l_moff.p, l_np.p, l_nabp.p, l_nsub.p, l_nab.p, l_j0.p, l_uoff.p, l_xaoff.p};
  }

  PullView pview() const { return PullView{f_pbase.p, f_maxp.p, f_ptab.p}; }
  FactorView view() const {
    return FactorView{f_W.p, f_WR.p, f_first.p, f_rowptr.p, f_rows.p, f_lptr.p, f_child_ptr.p, f_child.p,
                      f_relptr.p, f_rel.p, f_uptr.p};
  }
  size_t NC() const { return (size_t)S.N * C; }
  // the forward local step's per-vertex sums: strip sums on wave-uniform
  // scenes, corner slots otherwise (with their vertex ranges)
  const void* fsum() const { return R % 64 == 0 ? (const void*)ssum.p : (const void*)corner.p; }
  const int* fsum_ptr() const { return R % 64 == 0 ? ssptr.p : vptr.p; }
};

namespace {

int upload_static(pdm_scene s) {
  HostScene& S = s->S;
  if (s->perm.upload(S.perm.data(), S.N)) return 1;
  if (s->tet_order.upload(S.tet_order.data(), S.T)) return 1;
  if (s->tv.upload(soa(S.tv, S.T, 4).data(), 4 * (size_t)S.T)) return 1;
  if (s->tG.upload(soa(S.tG, S.T, 12).data(), 12 * (size_t)S.T)) return 1;
  if (s->tslot.upload(soa(s->Cn.tet_slot, S.T, 4).data(), 4 * (size_t)S.T)) return 1;
  if (s->vptr.upload(s->Cn.vptr.data(), S.N + 1)) return 1;
  const Strips& P = s->Sp;
  if (s->stet.upload(P.tet.data(), P.tet.size()) || s->scode.upload(P.code.data(), P.code.size()) ||
      s->sfbase.upload(P.fbase.data(), P.fbase.size()) || s->ssum_of.upload(P.sum_of.data(), P.sum_of.size()) ||
      s->ssptr.upload(P.ssptr.data(), P.ssptr.size()) || s->sfptr.upload(P.fptr.data(), P.fptr.size()) ||
      s->sfq.upload(P.fq.data(), P.fq.size()))
    return 1;
  if (s->mass.upload(S.mass.data(), S.N)) return 1;
  std::vector<double> aw(S.N, 0.0);
  for (size_t k = 0; k < S.att_idx.size(); ++k) aw[S.att_idx[k]] += S.att_w[k];
  if (s->attw.upload(aw.data(), S.N)) return 1;
  if (s->xrest.upload(S.xrest.data(), 3 * (size_t)S.N)) return 1;
  s->nfaces = S.underwater ? (int)(S.cf.size() / 3) : 0;
  if (s->nfaces > 0 && (s->dcf.upload(S.cf.data(), S.cf.size()) || s->dvfptr.upload(S.vfptr.data(), S.N + 1) ||
                        s->dvfq.upload(S.vfq.data(), S.vfq.size())))
    return 1;
  s->ntri = (int)(S.tris.size() / 9);
  s->nobs = (int)(S.planes.size() / 6 + S.spheres.size() / 4) + s->ntri;
  s->ncs = (S.contact && S.self_collision) ? (int)S.sc_vtx.size() : 0;
  s->nsf = s->ncs > 0 ? (int)(S.sc_face.size() / 3) : 0;
  if (s->ncs > 0 && (s->d_sc_vtx.upload(S.sc_vtx.data(), S.sc_vtx.size()) ||
                     s->d_sc_face.upload(S.sc_face.data(), S.sc_face.size()) ||
                     s->d_sc_of_v.upload(S.sc_of_v.data(), S.N) || s->d_scf_ptr.upload(S.scf_ptr.data(), S.N + 1) ||
                     s->d_scf_q.upload(S.scf_q.data(), S.scf_q.size())))
    return 1;
  s->ncv = (S.contact && (s->nobs > 0 || s->ncs > 0)) ? (int)S.contact_vtx.size() : 0;
  if (s->ncv > 0) {
    std::vector<int> cvo(S.N, -1);
    for (int k = 0; k < s->ncv; ++k) cvo[S.contact_vtx[k]] = k;
    if (s->cvtx.upload(S.contact_vtx.data(), s->ncv) || s->cv_of_v.upload(cvo.data(), S.N)) return 1;
    std::vector<double> pl(S.planes), sp(S.spheres);
    pl.push_back(0);
    sp.push_back(0);
    if (s->d_planes.upload(pl.data(), pl.size()) || s->d_spheres.upload(sp.data(), sp.size())) return 1;
    if (s->ntri > 0 &&
        (s->d_tris.upload(S.tris.data(), S.tris.size()) || s->d_bvh_box.upload(S.bvh_box.data(), S.bvh_box.size()) ||
         s->d_bvh_a.upload(S.bvh_a.data(), S.bvh_a.size()) || s->d_bvh_b.upload(S.bvh_b.data(), S.bvh_b.size()) ||
         s->d_bvh_tri.upload(S.bvh_tri.data(), S.bvh_tri.size())))
      return 1;
  }
  {
    std::vector<int> mp(S.mptr);
    if (mp.empty()) mp.assign(S.T + 1, 0);
    if (s->mptr.upload(mp.data(), mp.size())) return 1;
  }
  if (S.M > 0) {
    if (s->mv.upload(soa(S.mv, S.M, 4).data(), 4 * (size_t)S.M)) return 1;
    if (s->mg.upload(soa(S.mg, S.M, 4).data(), 4 * (size_t)S.M)) return 1;
    if (s->mw.upload(S.mw.data(), S.M)) return 1;
    if (s->mgroup.upload(S.mgroup.data(), S.M)) return 1;
    std::vector<int> gp(S.n_groups + 1, 0), gm(S.M);
    for (int m = 0; m < S.M; ++m) gp[S.mgroup[m] + 1]++;
    for (int g = 0; g < S.n_groups; ++g) gp[g + 1] += gp[g];
    std::vector<int> fill(gp.begin(), gp.end() - 1);
    for (int m = 0; m < S.M; ++m) gm[fill[S.mgroup[m]]++] = m;
    if (s->gptr.upload(gp.data(), gp.size())) return 1;
    if (s->gmus.upload(gm.data(), gm.size())) return 1;
  }
  return 0;
}

int upload_material(pdm_scene s) {
  HostScene& S = s->S;
  if (s->wr.upload(S.wr.data(), S.T)) return 1;
  if (s->wv.upload(S.wv.data(), S.T)) return 1;
  std::vector<double> der(S.T), dev(S.T);
  for (int t = 0; t < S.T; ++t) {
    der[t] = S.tvol[t] / (1.0 + S.nu_fwd_t[t]);
    dev[t] = S.tvol[t] * S.nu_ctor[t] / (1.0 + S.nu_fwd_t[t]);
  }
  if (s->dEr.upload(der.data(), S.T)) return 1;
  if (s->dEv.upload(dev.data(), S.T)) return 1;
  std::vector<double> kd(S.N);
  for (int i = 0; i < S.N; ++i) {
    kd[i] = S.mass[i] / (S.h * S.h);
    for (int q = s->Ls.ptr[i]; q < s->Ls.ptr[i + 1]; ++q)
      if (s->Ls.col[q] == i) kd[i] += s->Ls.val[q];
  }
  if (s->Kdiag.upload(kd.data(), S.N)) return 1;
  if (s->Lptr.upload(s->Ls.ptr.data(), s->Ls.ptr.size())) return 1;
  if (s->Lcol.upload(s->Ls.col.data(), s->Ls.col.size())) return 1;
  if (s->Lval.upload(s->Ls.val.data(), s->Ls.val.size())) return 1;
  return 0;
}

// Structure of the factor (and with host_w the W blocks from the host
// factor's L and inv(L_ss); otherwise W is only allocated).
int plan_solve(pdm_scene s);
int build_top_z(pdm_scene s);
int build_leaf_ops(pdm_scene s);

int upload_factor(pdm_scene s, bool host_w) {
  Factor& F = s->F;
  if (s->f_first.upload(F.first.data(), F.first.size())) return 1;
  if (s->f_rows.upload(F.rows.data(), F.rows.size())) return 1;
  if (s->f_child_ptr.upload(F.child_ptr.data(), F.child_ptr.size())) return 1;
  if (s->f_child.upload(F.child.data(), std::max<size_t>(1, F.child.size()))) return 1;
  if (s->f_rel.upload(F.rel.data(), std::max<size_t>(1, F.rel.size()))) return 1;
  if (s->f_lev.upload(F.level_sn.data(), F.level_sn.size())) return 1;
  std::vector<long long> tmp(F.rowptr.begin(), F.rowptr.end());
  if (s->f_rowptr.upload(tmp.data(), tmp.size())) return 1;
  tmp.assign(F.lptr.begin(), F.lptr.end());
  if (s->f_lptr.upload(tmp.data(), tmp.size())) return 1;
  tmp.assign(F.relptr.begin(), F.relptr.end());
  if (s->f_relptr.upload(tmp.data(), tmp.size())) return 1;
  tmp.assign(F.uptr.begin(), F.uptr.end());
  if (s->f_uptr.upload(tmp.data(), tmp.size())) return 1;
  if (!host_w) {
    if (s->f_W.alloc(F.nnz) || s->f_WR.alloc(F.nnz)) return 1;
  } else {
    // W = [inv(L_ss); -L_below inv(L_ss)] per supernode, column-major and
    // row-major (coalesced operand loads in both sweeps)
    int ns = (int)F.first.size() - 1;
    std::vector<double> wc(F.L.size()), wr(F.L.size());
    std::vector<double> row;
    for (int sn = 0; sn < ns; ++sn) {
      int w = F.first[sn + 1] - F.first[sn];
      int m = (int)(F.rowptr[sn + 1] - F.rowptr[sn]);
      const double* D = &F.Dinv[F.dptr[sn]];  // row-major, zero above the diagonal
      const double* Lc = &F.L[F.lptr[sn]];    // column-major m x w
      double* Wc = &wc[F.lptr[sn]];
      double* Wr = &wr[F.lptr[sn]];
      row.assign(w, 0.0);
      for (int r = 0; r < m; ++r) {
        if (r < w) {
          for (int k = 0; k < w; ++k) row[k] = D[(size_t)r * w + k];
        } else {
          // -sum_j L[r][j] inv(L_ss)[j][k], j >= k
          std::fill(row.begin(), row.end(), 0.0);
          for (int jj = 0; jj < w; ++jj) {
            const double l = Lc[(size_t)jj * m + r];
            const double* Dj = D + (size_t)jj * w;
            for (int k = 0; k <= jj; ++k) row[k] -= l * Dj[k];
          }
        }
        for (int k = 0; k < w; ++k) {
          Wc[(size_t)k * m + r] = row[k];
          Wr[(size_t)r * w + k] = row[k];
        }
      }
    }
    if (s->f_W.upload(wc.data(), wc.size()) || s->f_WR.upload(wr.data(), wr.size())) return 1;
  }
  int nlev = (int)F.level_ptr.size() - 1;
  std::vector<int> old_ptr = s->lev_ptr, old_rows = s->lev_rows;
  s->lev_ptr = F.level_ptr;
  s->lev_rows.assign(nlev, 16);
  for (int l = 0; l < nlev; ++l)
    for (int q = F.level_ptr[l]; q < F.level_ptr[l + 1]; ++q) {
      int sn = F.level_sn[q];
      int m = (int)(F.rowptr[sn + 1] - F.rowptr[sn]), w = F.first[sn + 1] - F.first[sn];
      s->lev_rows[l] = std::max(s->lev_rows[l], ((m + 15) & ~15) + ((w + 15) & ~15));
    }
  if (old_ptr != s->lev_ptr || old_rows != s->lev_rows) s->tuned_C = -1;  // level structure changed
  return plan_solve(s);
}

// Top-separator blocks of every elimination tree (pdm_top.h): per tree the
// largest upward-closed set of whole heights whose supernodes own at most
// top_cap rows. Host tables and their device copies; Z itself is formed by
// build_top_z once W is on the device.
int plan_top(pdm_scene s) {
  Factor& F = s->F;
  const int ns = (int)F.first.size() - 1;
  const int nlev = (int)F.level_ptr.size() - 1;
  std::vector<int> height(ns, 0), tree(ns, -1);
  for (int l = 0; l < nlev; ++l)
    for (int q = F.level_ptr[l]; q < F.level_ptr[l + 1]; ++q) height[F.level_sn[q]] = l;
  for (int sn = ns - 1; sn >= 0; --sn) tree[sn] = F.parent[sn] < 0 ? sn : tree[F.parent[sn]];
  // rows per (tree, height)
  std::vector<int> cutof(ns, -1);  // per root: lowest height in its top set (-1: none)
  {
    std::vector<std::vector<long long>> wh(ns);
    for (int sn = 0; sn < ns; ++sn) {
      auto& v = wh[tree[sn]];
      if ((int)v.size() <= height[sn]) v.resize(height[sn] + 1, 0);
      v[height[sn]] += F.first[sn + 1] - F.first[sn];
    }
    for (int rt = 0; rt < ns; ++rt) {
      if (F.parent[rt] >= 0 || s->top_cap <= 0) continue;
      long long acc = 0;
      for (int h = (int)wh[rt].size() - 1; h >= 0; --h) {
        acc += wh[rt][h];
        if (acc > s->top_cap) break;
        cutof[rt] = h;
      }
    }
  }
  s->is_top.assign(ns, 0);
  for (int sn = 0; sn < ns; ++sn) s->is_top[sn] = cutof[tree[sn]] >= 0 && height[sn] >= cutof[tree[sn]];
  s->top_a0.clear();
  s->top_n.clear();
  s->top_np.clear();
  s->top_zoff.clear();
  std::vector<int> blk(ns, -1);  // per root: its block
  std::vector<int> tmap(F.n, -1), plocal(ns, -1);
  int rows = 0;
  long long zsz = 0;
  for (int rt = 0; rt < ns; ++rt) {
    if (cutof[rt] < 0) continue;
    blk[rt] = (int)s->top_a0.size();
    s->top_a0.push_back(rows);
    s->top_zoff.push_back(zsz);
    s->top_n.push_back(0);
    s->top_np.push_back(0);
  }
  // local rows: top supernodes of a tree in ascending (post)order, so the
  // local numbering is ascending engine order and L restricted to it is lower
  // triangular
  for (int sn = 0; sn < ns; ++sn)
    if (s->is_top[sn]) {
      const int b = blk[tree[sn]];
      plocal[sn] = s->top_n[b];
      s->top_n[b] += F.first[sn + 1] - F.first[sn];
    }
  for (size_t b = 0; b < s->top_a0.size(); ++b) {
    s->top_np[b] = (s->top_n[b] + 31) & ~31;
    s->top_a0[b] = rows;
    s->top_zoff[b] = zsz;
    rows += s->top_np[b];
    zsz += (long long)s->top_np[b] * s->top_np[b];
  }
  s->top_on = !s->top_a0.empty();
  s->top_rows = rows;
  s->top_zsize = zsz;
  if (!s->top_on) return 0;
  std::vector<int> trow(rows, -1);
  for (int sn = 0; sn < ns; ++sn)
    if (s->is_top[sn]) {
      const int b = blk[tree[sn]];
      for (int j = F.first[sn]; j < F.first[sn + 1]; ++j) {
        const int a = s->top_a0[b] + plocal[sn] + (j - F.first[sn]);
        tmap[j] = a;
        trow[a] = j;
      }
    }
  // fronts: the update rows of the top supernodes' children below the cut
  std::vector<std::vector<int>> per(rows);
  for (int sn = 0; sn < ns; ++sn) {
    if (!s->is_top[sn]) continue;
    for (int q = F.child_ptr[sn]; q < F.child_ptr[sn + 1]; ++q) {
      const int ch = F.child[q];
      if (s->is_top[ch]) continue;
      const int wc = F.first[ch + 1] - F.first[ch];
      const int mbc = (int)(F.rowptr[ch + 1] - F.rowptr[ch]) - wc;
      for (int a = 0; a < mbc; ++a) {
        const int g = tmap[F.rows[F.rowptr[ch] + wc + a]];
        if (g < 0) {
          g_err = "top separator plan: an update row lies outside the top rows";
          return 1;
        }
        per[g].push_back((int)(F.uptr[ch] + a));
      }
    }
  }
  int mp = 0;
  for (int a = 0; a < rows; ++a) mp = std::max(mp, (int)per[a].size());
  s->top_maxp = mp;
  std::vector<int> slot((size_t)rows * std::max(mp, 1), -1);
  for (int a = 0; a < rows; ++a)
    for (size_t k = 0; k < per[a].size(); ++k) slot[(size_t)a * mp + k] = per[a][k];
  // elimination waves: wave k holds the k-th top supernode of every tree;
  // per item its front rows in local numbering and its scratch offset
  std::vector<std::vector<int>> order(s->top_a0.size());
  for (int sn = 0; sn < ns; ++sn)
    if (s->is_top[sn]) order[blk[tree[sn]]].push_back(sn);
  size_t nw = 0;
  for (auto& o : order) nw = std::max(nw, o.size());
  std::vector<int4> wave;
  std::vector<long long> toff;
  std::vector<int> frow;
  s->top_wave_ptr.assign(1, 0);
  s->top_wave_sn.clear();
  s->top_wave_blk.clear();
  long long tsz = 0;
  for (size_t k = 0; k < nw; ++k) {
    for (size_t b = 0; b < order.size(); ++b) {
      if (k >= order[b].size()) continue;
      const int sn = order[b][k];
      const int m = (int)(F.rowptr[sn + 1] - F.rowptr[sn]);
      wave.push_back(make_int4(sn, (int)b, plocal[sn], (int)frow.size()));
      s->top_wave_sn.push_back(sn);
      s->top_wave_blk.push_back((int)b);
      for (int i = 0; i < m; ++i) {
        const int g = tmap[F.rows[F.rowptr[sn] + i]];
        if (g < 0) {
          g_err = "top separator plan: a front row lies outside the top rows";
          return 1;
        }
        frow.push_back(g - s->top_a0[b]);
      }
      toff.push_back(tsz);
      tsz += (long long)(F.first[sn + 1] - F.first[sn]) * s->top_np[b];
    }
    s->top_wave_ptr.push_back((int)wave.size());
  }
  s->top_tsize = tsz;
  std::vector<int2> items;
  for (size_t b = 0; b < s->top_a0.size(); ++b)
    for (int t = 0; t < s->top_np[b] / 32; ++t) items.push_back(make_int2((int)b, t));
  s->t_nitems = (int)items.size();
  if (s->t_row.upload(trow.data(), trow.size()) || s->t_slot.upload(slot.data(), slot.size()) ||
      s->t_a0.upload(s->top_a0.data(), s->top_a0.size()) || s->t_np.upload(s->top_np.data(), s->top_np.size()) ||
      s->t_zoff.upload(s->top_zoff.data(), s->top_zoff.size()) || s->t_frow.upload(frow.data(), frow.size()) ||
      s->t_toff.upload(toff.data(), toff.size()) || s->t_wave.upload(wave.data(), wave.size()) ||
      s->t_items.upload(items.data(), items.size()))
    return 1;
  if (s->tZ.alloc(zsz) || s->tM.alloc(zsz) || s->tT.alloc(std::max<long long>(1, tsz))) return 1;
  return 0;
}

// Z = (L_t L_t^T)^{-1} of every top block from the device factor (W): the
// identity, the elimination waves in postorder, then the Gram product.
int build_top_z(pdm_scene s) {
  if (!s->top_on) return 0;
  const hipStream_t st = s->stream;
  const int nb = (int)s->top_a0.size();
  for (int b = 0; b < nb; ++b) {
    const int np = s->top_np[b];
    hipLaunchKernelGGL(k_top_eye, dim3(cdiv((long long)np * np, 256)), dim3(256), 0, st, s->tM.p + s->top_zoff[b], np,
                       np);
  }
  const FactorView fv = s->view();
  const Factor& F = s->F;
  for (size_t k = 0; k + 1 < s->top_wave_ptr.size(); ++k) {
    const int i0 = s->top_wave_ptr[k], ni = s->top_wave_ptr[k + 1] - i0;
    // grid: the wave's largest w * np (copy) and 16 x 16 tiles (elimination)
    long long maxe = 0, maxt = 0;
    for (int i = i0; i < i0 + ni; ++i) {
      const int sn = s->top_wave_sn[i], b = s->top_wave_blk[i];
      const int w = F.first[sn + 1] - F.first[sn];
      const int m = (int)(F.rowptr[sn + 1] - F.rowptr[sn]);
      maxe = std::max(maxe, (long long)w * s->top_np[b]);
      maxt = std::max(maxt, (long long)((m + 15) / 16) * (s->top_np[b] / 16));
    }
    hipLaunchKernelGGL(k_top_copy, dim3(cdiv(maxe, 256), ni), dim3(256), 0, st, s->t_wave.p + i0, s->t_toff.p + i0,
                       s->f_first.p, s->t_np.p, s->t_zoff.p, s->tM.p, s->tT.p);
    hipLaunchKernelGGL(k_top_elim, dim3((unsigned)maxt, ni), dim3(64), 0, st, fv, s->t_wave.p + i0, s->t_toff.p + i0,
                       s->t_frow.p, s->t_np.p, s->t_zoff.p, s->tT.p, s->tM.p);
  }
  for (int b = 0; b < nb; ++b) {
    const int np = s->top_np[b];
    hipLaunchKernelGGL(k_top_gram, dim3((np / 16) * (np / 16)), dim3(64), 0, st, s->tM.p + s->top_zoff[b],
                       s->tZ.p + s->top_zoff[b], np);
  }
  HIPX(hipGetLastError());
  return 0;
}

// The dense leaf operators from the device factor (W): M = [I; 0] through
// each leaf's elimination steps (postorder) gives [Linv; G]; the Gram product
// of Linv gives Zs = K_s^{-1}; the forward operator is packed as [Zs | G^T].
int build_leaf_ops(pdm_scene s) {
  if (!s->leaf_on) return 0;
  const hipStream_t st = s->stream;
  const int nl = (int)s->leaf_np.size();
  for (int l = 0; l < nl; ++l) {
    const int np = s->leaf_np[l], rows = np + s->leaf_nabp[l];
    hipLaunchKernelGGL(k_top_eye, dim3(cdiv((long long)rows * np, 256)), dim3(256), 0, st, s->lM.p + s->leaf_moff[l],
                       np, rows);
  }
  const FactorView fv = s->view();
  const Factor& F = s->F;
  for (size_t k = 0; k + 1 < s->leaf_wave_ptr.size(); ++k) {
    const int i0 = s->leaf_wave_ptr[k], ni = s->leaf_wave_ptr[k + 1] - i0;
    long long maxe = 0, maxt = 0;
    for (int i = i0; i < i0 + ni; ++i) {
      const int sn = s->leaf_wave_sn[i], l = s->leaf_wave_blk[i];
      const int w = F.first[sn + 1] - F.first[sn];
      const int m = (int)(F.rowptr[sn + 1] - F.rowptr[sn]);
      maxe = std::max(maxe, (long long)w * s->leaf_np[l]);
      maxt = std::max(maxt, (long long)((m + 15) / 16) * (s->leaf_np[l] / 16));
    }
    if (ni > 65535) {
      g_err = "leaf plan: more than 65535 leaves in one elimination wave";
      return 1;
    }
    hipLaunchKernelGGL(k_top_copy, dim3(cdiv(maxe, 256), ni), dim3(256), 0, st, s->l_wave.p + i0, s->l_toff.p + i0,
                       s->f_first.p, s->l_np.p, s->l_moff.p, s->lM.p, s->lT.p);
    hipLaunchKernelGGL(k_top_elim, dim3((unsigned)maxt, ni), dim3(64), 0, st, fv, s->l_wave.p + i0, s->l_toff.p + i0,
                       s->l_frow.p, s->l_np.p, s->l_moff.p, s->lT.p, s->lM.p);
  }
  for (int l = 0; l < nl; ++l) {
    const int np = s->leaf_np[l], nabp = s->leaf_nabp[l];
    hipLaunchKernelGGL(k_top_gram, dim3((np / 16) * (np / 16)), dim3(64), 0, st, s->lM.p + s->leaf_moff[l],
                       s->lZ.p + s->leaf_zoff[l], np);
    hipLaunchKernelGGL(k_leaf_pack, dim3(cdiv((long long)np * (np + nabp), 256)), dim3(256), 0, st,
                       s->lZ.p + s->leaf_zoff[l], s->lM.p + s->leaf_moff[l], s->lP.p + s->leaf_poff[l], np, nabp);
  }
  HIPX(hipGetLastError());
  return 0;
}

// Dense leaf subtrees (pdm_top.h): the maximal subtrees below the top
// separators whose supernodes own at most leaf_cap rows together; host tables
// and their device copies (the operators are formed by build_top_z).
int plan_leaf(pdm_scene s) {
  Factor& F = s->F;
  const int ns = (int)F.first.size() - 1;
  std::vector<long long> rows(ns, 0);
  std::vector<int> fd(ns);
  for (int sn = 0; sn < ns; ++sn) {
    rows[sn] += F.first[sn + 1] - F.first[sn];
    fd[sn] = sn;
    for (int q = F.child_ptr[sn]; q < F.child_ptr[sn + 1]; ++q) fd[sn] = std::min(fd[sn], fd[F.child[q]]);
    if (F.parent[sn] >= 0) rows[F.parent[sn]] += rows[sn];
  }
  s->is_leaf.assign(ns, 0);
  s->leaf_np.clear(); s->leaf_nabp.clear(); s->leaf_nsub.clear(); s->leaf_nab.clear(); s->leaf_j0.clear();
  s->leaf_moff.clear(); s->leaf_poff.clear(); s->leaf_zoff.clear(); s->leaf_uoff.clear(); s->leaf_xaoff.clear();
  s->leaf_wave_ptr.assign(1, 0);
  s->leaf_wave_sn.clear();
  s->leaf_wave_blk.clear();
  std::vector<int> roots;
  if (s->leaf_cap > 0)
    for (int sn = 0; sn < ns; ++sn) {
      if (s->is_top[sn] || rows[sn] > s->leaf_cap) continue;
      const int par = F.parent[sn];
      if (par >= 0 && !s->is_top[par] && rows[par] <= s->leaf_cap) continue;  // inside a larger leaf
      roots.push_back(sn);
    }
  long long msz = 0, psz = 0, zsz = 0, tsz = 0;
  int xrows = 0;
  std::vector<int> xrow, frow;
  std::vector<int4> wave;
  std::vector<long long> toff;
  std::vector<std::vector<int>> order(roots.size());
  for (size_t l = 0; l < roots.size(); ++l) {
    const int rt = roots[l], a0 = fd[rt];
    const int j0 = F.first[a0], nsub = F.first[rt + 1] - j0;
    const int wr = F.first[rt + 1] - F.first[rt];
    const int nab = (int)(F.rowptr[rt + 1] - F.rowptr[rt]) - wr;
    const int np = (nsub + 31) & ~31, nabp = (nab + 31) & ~31;
    s->leaf_np.push_back(np);
    s->leaf_nabp.push_back(nabp);
    s->leaf_nsub.push_back(nsub);
    s->leaf_nab.push_back(nab);
    s->leaf_j0.push_back(j0);
    s->leaf_moff.push_back(msz);
    s->leaf_poff.push_back(psz);
    s->leaf_zoff.push_back(zsz);
    s->leaf_uoff.push_back(F.uptr[rt]);
    s->leaf_xaoff.push_back(xrows);
    msz += (long long)(np + nabp) * np;
    psz += (long long)np * (np + nabp);
    zsz += (long long)np * np;
    for (int i = 0; i < nabp; ++i) xrow.push_back(i < nab ? F.rows[F.rowptr[rt] + wr + i] : -1);
    xrows += nabp;
    for (int sn = a0; sn <= rt; ++sn) {
      s->is_leaf[sn] = 1;
      order[l].push_back(sn);
    }
  }
  s->leaf_on = !roots.empty();
  s->leaf_msize = msz;
  s->leaf_psize = psz;
  s->leaf_zsize = zsz;
  s->leaf_xrows = xrows;
  if (!s->leaf_on) return 0;
  // elimination waves: wave k holds the k-th supernode (postorder) of every
  // leaf; per item its front rows in the leaf's local numbering (own rows
  // first, then the above rows after the padded own rows)
  size_t nw = 0;
  for (auto& o : order) nw = std::max(nw, o.size());
  for (size_t k = 0; k < nw; ++k) {
    for (size_t l = 0; l < order.size(); ++l) {
      if (k >= order[l].size()) continue;
      const int sn = order[l][k], rt = roots[l];
      const int j0 = s->leaf_j0[l], nsub = s->leaf_nsub[l], np = s->leaf_np[l];
      const int wr = F.first[rt + 1] - F.first[rt];
      const int m = (int)(F.rowptr[sn + 1] - F.rowptr[sn]);
      wave.push_back(make_int4(sn, (int)l, F.first[sn] - j0, (int)frow.size()));
      s->leaf_wave_sn.push_back(sn);
      s->leaf_wave_blk.push_back((int)l);
      for (int i = 0; i < m; ++i) {
        const int row = F.rows[F.rowptr[sn] + i];
        int lr;
        if (row >= j0 && row < j0 + nsub) {
          lr = row - j0;
        } else {
          const int* b = F.rows.data() + F.rowptr[rt] + wr;
          const int* e = b + s->leaf_nab[l];
          const int* p = std::lower_bound(b, e, row);
          if (p == e || *p != row) {
            g_err = "leaf plan: a front row lies outside the leaf and its root's rows";
            return 1;
          }
          lr = np + (int)(p - b);
        }
        frow.push_back(lr);
      }
      toff.push_back(tsz);
      tsz += (long long)(F.first[sn + 1] - F.first[sn]) * np;
    }
    s->leaf_wave_ptr.push_back((int)wave.size());
  }
  s->leaf_tsize = tsz;
  std::vector<int2> items[2];
  for (size_t l = 0; l < roots.size(); ++l) {
    for (int t = 0; t < (s->leaf_np[l] + s->leaf_nabp[l]) / 32; ++t) items[0].push_back(make_int2((int)l, t));
    for (int t = 0; t < s->leaf_np[l] / 32; ++t) items[1].push_back(make_int2((int)l, t));
  }
  s->l_nitems[0] = (int)items[0].size();
  s->l_nitems[1] = (int)items[1].size();
  std::vector<char> skip(s->is_leaf.begin(), s->is_leaf.end());
  if (s->l_np.upload(s->leaf_np.data(), s->leaf_np.size()) || s->l_nabp.upload(s->leaf_nabp.data(), s->leaf_nabp.size()) ||
      s->l_nsub.upload(s->leaf_nsub.data(), s->leaf_nsub.size()) || s->l_nab.upload(s->leaf_nab.data(), s->leaf_nab.size()) ||
      s->l_j0.upload(s->leaf_j0.data(), s->leaf_j0.size()) || s->l_moff.upload(s->leaf_moff.data(), s->leaf_moff.size()) ||
      s->l_poff.upload(s->leaf_poff.data(), s->leaf_poff.size()) || s->l_uoff.upload(s->leaf_uoff.data(), s->leaf_uoff.size()) ||
      s->l_xaoff.upload(s->leaf_xaoff.data(), s->leaf_xaoff.size()) || s->l_frow.upload(frow.data(), frow.size()) ||
      s->l_toff.upload(toff.data(), toff.size()) || s->l_wave.upload(wave.data(), wave.size()) ||
      s->l_xrow.upload(xrow.data(), xrow.size()) || s->l_items[0].upload(items[0].data(), items[0].size()) ||
      s->l_items[1].upload(items[1].data(), items[1].size()) || s->d_skip.upload(skip.data(), skip.size()))
    return 1;
  if (s->lM.alloc(msz) || s->lZ.alloc(zsz) || s->lP.alloc(psz) || s->lT.alloc(std::max<long long>(1, tsz))) return 1;
  return 0;
}

// Pull tables, the grouped level kernels' work items and the subtree ranges
// (all without the top-separator and dense-leaf supernodes when dense blocks
// are on).
int plan_solve(pdm_scene s) {
  Factor& F = s->F;
  const int nlev = (int)F.level_ptr.size() - 1;
  if (plan_top(s) || plan_leaf(s)) return 1;
  {
    // pull tables: for every front row of every supernode, the update rows of
    // its children that add into it, in child order, padded with -1 to the
    // supernode's largest count per row
    int ns = (int)F.first.size() - 1;
    std::vector<long long> pbase(ns + 1, 0);
    std::vector<int> maxp(ns, 0), ptab;
    std::vector<std::vector<int>> per;
    for (int sn = 0; sn < ns; ++sn) {
      int m = (int)(F.rowptr[sn + 1] - F.rowptr[sn]);
      per.assign(m, {});
      for (int q = F.child_ptr[sn]; q < F.child_ptr[sn + 1]; ++q) {
        int ch = F.child[q];
        int wc = F.first[ch + 1] - F.first[ch];
        int mbc = (int)(F.rowptr[ch + 1] - F.rowptr[ch]) - wc;
        for (int a = 0; a < mbc; ++a) per[F.rel[F.relptr[ch] + a]].push_back((int)(F.uptr[ch] + a));
      }
      int mp = 0;
      for (int r = 0; r < m; ++r) mp = std::max(mp, (int)per[r].size());
      maxp[sn] = mp;
      pbase[sn] = (long long)ptab.size();
      for (int r = 0; r < m; ++r)
        for (int k = 0; k < mp; ++k) ptab.push_back(k < (int)per[r].size() ? per[r][k] : -1);
    }
    pbase[ns] = (long long)ptab.size();
    if (s->f_pbase.upload(pbase.data(), pbase.size()) || s->f_maxp.upload(maxp.data(), std::max<size_t>(1, maxp.size())) ||
        s->f_ptab.upload(ptab.data(), std::max<size_t>(1, ptab.size())))
      return 1;
    // work items: consecutive 16-row output tiles of one supernode, as many as
    // keep the W slice within kWslItem doubles (at least one tile)
    for (int pass = 0; pass < 2; ++pass) {
      std::vector<int4> items;
      s->item_ptr[pass].assign(nlev + 1, 0);
      s->item_front[pass].assign(nlev, 16);
      s->item_wsl[pass].assign(nlev, 0);
      for (int l = 0; l < nlev; ++l) {
        for (int q = F.level_ptr[l]; q < F.level_ptr[l + 1]; ++q) {
          int sn = F.level_sn[q];
          if (s->is_top[sn] || s->is_leaf[sn]) continue;
          int m = (int)(F.rowptr[sn + 1] - F.rowptr[sn]), w = F.first[sn + 1] - F.first[sn];
          int wpad = (w + 15) & ~15, mpad = (m + 15) & ~15, nrt = wpad / 16, nbt = (m - w + 15) / 16;
          int ntl = pass == 0 ? nrt + nbt : nrt;
          int t = 0;
          while (t < ntl) {
            int t1 = t, wsl = 0;
            while (t1 < ntl) {
              int per_tile = pass == 0 ? wpad * 16 : (m - 16 * t1) * 16;
              if (t1 > t && wsl + per_tile > kWslItem) break;
              wsl += per_tile;
              ++t1;
            }
            items.push_back(make_int4(sn, t, t1, 0));
            int nupd = pass == 0 ? std::max(0, t1 - std::max(t, nrt)) : 0;
            int front = pass == 0 ? wpad + 16 * nupd : mpad;
            s->item_front[pass][l] = std::max(s->item_front[pass][l], front);
            s->item_wsl[pass][l] = std::max(s->item_wsl[pass][l], wsl);
            t = t1;
          }
        }
        s->item_ptr[pass][l + 1] = (int)items.size();
      }
      if (s->f_items[pass].upload(items.data(), std::max<size_t>(1, items.size()))) return 1;
    }
  }
  {
    // subtrees for every cut level c: roots of height < c whose parent has
    // height >= c (or none); a subtree is the postorder range [first, root]
    int ns = (int)F.first.size() - 1;
    std::vector<int> height(ns, 0), fd(ns);
    for (int l = 0; l < nlev; ++l)
      for (int q = F.level_ptr[l]; q < F.level_ptr[l + 1]; ++q) height[F.level_sn[q]] = l;
    for (int sn = 0; sn < ns; ++sn) {
      fd[sn] = sn;
      for (int q = F.child_ptr[sn]; q < F.child_ptr[sn + 1]; ++q) fd[sn] = std::min(fd[sn], fd[F.child[q]]);
    }
    s->sub_ranges.assign(nlev + 1, {});
    for (int c = 1; c <= nlev; ++c)
      for (int sn = 0; sn < ns; ++sn) {
        int par = F.parent[sn];
        if (s->is_top[sn] || s->is_leaf[sn] || height[sn] >= c) continue;
        if (par < 0 || height[par] >= c || s->is_top[par]) s->sub_ranges[c].push_back(make_int2(fd[sn], sn));
      }
    s->d_sub_cut = -1;
  }
  return 0;
}

int host_factor(pdm_scene s);

// Re-factorisation on the host CPU (multifrontal factor and W precompute in
// host code, then uploaded).
int refactor_host(pdm_scene s) {
  auto t0 = std::chrono::steady_clock::now();
  if (host_factor(s)) return 1;
  if (upload_material(s) || upload_factor(s, true) || build_top_z(s) || build_leaf_ops(s)) return 1;
  s->fac_plan = false;  // the structure upload above replaced the device plan's inputs
  s->t_factor = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return 0;
}

// Device re-factorisation plan (once per scene; the structure never changes).
int build_fac_plan(pdm_scene s) {
#pragma clang fp contract(off)
  HostScene& S = s->S;
  const Csr& A = s->Ls;
  Factor& F = s->F;
  const int N = S.N, ns = F.ns;
  const long long nnzA = (long long)A.col.size();
  auto find_q = [&](int i, int j) -> long long {
    const int* b = A.col.data() + A.ptr[i];
    const int* e = A.col.data() + A.ptr[i + 1];
    const int* f = std::lower_bound(b, e, j);
    return (f != e && *f == j) ? (long long)(f - A.col.data()) : -1;
  };
  // contributions of each stored entry, in the host assembly's order (tets
  // ascending, muscles, attachments, the structural 0 of the diagonal)
  size_t ncon = 16 * (size_t)S.T + 16 * (size_t)S.M + S.att_idx.size() + (size_t)N;
  std::vector<long long> cq;
  std::vector<int> ct;
  std::vector<double> cv;
  cq.reserve(ncon);
  ct.reserve(ncon);
  cv.reserve(ncon);
  for (int t = 0; t < S.T; ++t) {
    const double* G = &S.tG[12 * (size_t)t];
    for (int a = 0; a < 4; ++a)
      for (int b = 0; b < 4; ++b) {
        double gg = G[3 * a] * G[3 * b] + G[3 * a + 1] * G[3 * b + 1] + G[3 * a + 2] * G[3 * b + 2];
        cq.push_back(find_q(S.tv[4 * t + a], S.tv[4 * t + b]));
        ct.push_back(t);
        cv.push_back(gg);
      }
  }
  for (int m = 0; m < S.M; ++m)
    for (int a = 0; a < 4; ++a)
      for (int b = 0; b < 4; ++b) {
        cq.push_back(find_q(S.mv[4 * m + a], S.mv[4 * m + b]));
        ct.push_back(-1);
        cv.push_back(S.mw[m] * S.mg[4 * m + a] * S.mg[4 * m + b]);
      }
  for (size_t k = 0; k < S.att_idx.size(); ++k) {
    cq.push_back(find_q(S.att_idx[k], S.att_idx[k]));
    ct.push_back(-1);
    cv.push_back(S.att_w[k]);
  }
  for (int i = 0; i < N; ++i) {
    cq.push_back(find_q(i, i));
    ct.push_back(-1);
    cv.push_back(0.0);
  }
  std::vector<long long> cptr(nnzA + 1, 0);
  for (long long q : cq) {
    if (q < 0) {
      g_err = "device factor plan: a contribution has no stored entry";
      return 1;
    }
    cptr[q + 1]++;
  }
  for (long long q = 0; q < nnzA; ++q) cptr[q + 1] += cptr[q];
  std::vector<int> ct2(cq.size());
  std::vector<double> cv2(cq.size());
  {
    std::vector<long long> fill(cptr.begin(), cptr.end() - 1);
    for (size_t k = 0; k < cq.size(); ++k) {  // stable: keeps the order above within an entry
      long long d = fill[cq[k]]++;
      ct2[d] = ct[k];
      cv2[d] = cv[k];
    }
  }
  std::vector<long long>().swap(cq);
  std::vector<int>().swap(ct);
  std::vector<double>().swap(cv);
  // the lists re-sum, in k_fac_assemble_L's arithmetic, to the host assembly's
  // values bit for bit (same terms, same order, same roundings)
  for (long long q = 0; q < nnzA; ++q) {
    double v = 0.0;
    for (long long c = cptr[q]; c < cptr[q + 1]; ++c) {
      const int t = ct2[c];
      if (t >= 0)
        v += (S.wr[t] + S.wv[t]) * cv2[c];
      else
        v += cv2[c];
    }
    if (std::memcmp(&v, &A.val[q], sizeof(double)) != 0) {
      g_err = "device factor plan: contribution lists do not reproduce the assembled matrix";
      return 1;
    }
  }
  // front positions of A's lower entries and of the diagonals; update blocks
  std::vector<long long> amap(nnzA, -1), dpos(N), upos(ns + 1, 0);
  for (int sn = 0; sn < ns; ++sn) {
    const int c0 = F.first[sn], w = F.first[sn + 1] - c0;
    const int m = (int)(F.rowptr[sn + 1] - F.rowptr[sn]);
    const int* rows = &F.rows[F.rowptr[sn]];
    for (int j = 0; j < w; ++j) {
      const int col = c0 + j;
      dpos[col] = F.lptr[sn] + (long long)j * m + j;
      for (int q = A.ptr[col]; q < A.ptr[col + 1]; ++q) {
        const int r = A.col[q];
        if (r < col) continue;
        const int k = (int)(std::lower_bound(rows, rows + m, r) - rows);
        if (k >= m || rows[k] != r) {
          g_err = "device factor plan: an entry of A is outside its supernode's rows";
          return 1;
        }
        amap[q] = F.lptr[sn] + (long long)j * m + k;
      }
    }
    upos[sn + 1] = upos[sn] + (long long)(m - w) * (m - w);
  }
  s->fc_usize = upos[ns];
  // per level: extend items (supernode, parent row tile) of supernodes with
  // children; update items (supernode, kind, row tile, column tile)
  const int nlev = (int)F.level_ptr.size() - 1;
  std::vector<int2> ext;
  std::vector<int4> upd;
  s->fc_ext_ptr.assign(nlev + 1, 0);
  s->fc_upd_ptr.assign(nlev + 1, 0);
  s->fc_lev_maxm.assign(nlev, 1);
  const int T = pdmfac::kTile;
  for (int l = 0; l < nlev; ++l) {
    for (int q = F.level_ptr[l]; q < F.level_ptr[l + 1]; ++q) {
      const int sn = F.level_sn[q];
      const int w = F.first[sn + 1] - F.first[sn], m = (int)(F.rowptr[sn + 1] - F.rowptr[sn]), mb = m - w;
      s->fc_lev_maxm[l] = std::max(s->fc_lev_maxm[l], m);
      if (F.child_ptr[sn + 1] > F.child_ptr[sn])
        for (int t = 0; t < cdiv(m, T); ++t) ext.push_back(make_int2(sn, t));
      for (int ti = 0; ti < cdiv(mb, T); ++ti) {
        for (int tj = 0; tj < cdiv(w, T); ++tj) upd.push_back(make_int4(sn, 0, ti, tj));
        for (int tj = 0; tj <= ti; ++tj) upd.push_back(make_int4(sn, 1, ti, tj));
      }
    }
    s->fc_ext_ptr[l + 1] = (int)ext.size();
    s->fc_upd_ptr[l + 1] = (int)upd.size();
  }
  if (s->fc_cptr.upload(cptr.data(), cptr.size()) || s->fc_ctet.upload(ct2.data(), std::max<size_t>(1, ct2.size())) ||
      s->fc_cval.upload(cv2.data(), std::max<size_t>(1, cv2.size())) || s->fc_amap.upload(amap.data(), std::max<size_t>(1, amap.size())) ||
      s->fc_dpos.upload(dpos.data(), dpos.size()) || s->fc_upos.upload(upos.data(), upos.size()) ||
      s->fc_ext.upload(ext.data(), std::max<size_t>(1, ext.size())) ||
      s->fc_upd.upload(upd.data(), std::max<size_t>(1, upd.size())) || s->fc_fail.alloc(1) ||
      s->fc_L.alloc(std::max<long long>(1, F.nnz)) || s->fc_U.alloc(std::max<long long>(1, s->fc_usize)))
    return 1;
  if (s->Lptr.upload(A.ptr.data(), A.ptr.size()) || s->Lcol.upload(A.col.data(), A.col.size()) ||
      s->Lval.alloc(std::max<long long>(1, nnzA)) || s->Kdiag.alloc(N))
    return 1;
  return 0;
}

// Per-tet weights and the E-gradient factors (host, O(T)), uploaded.
int upload_weights(pdm_scene s) {
  HostScene& S = s->S;
  update_weights(S);
  if (s->wr.upload(S.wr.data(), S.T) || s->wv.upload(S.wv.data(), S.T)) return 1;
  std::vector<double> der(S.T), dev(S.T);
  for (int t = 0; t < S.T; ++t) {
    der[t] = S.tvol[t] / (1.0 + S.nu_fwd_t[t]);
    dev[t] = S.tvol[t] * S.nu_ctor[t] / (1.0 + S.nu_fwd_t[t]);
  }
  if (s->dEr.upload(der.data(), S.T) || s->dEv.upload(dev.data(), S.T)) return 1;
  return 0;
}

// Re-factorisation on the device: L's values and diag(K_s) from the weights,
// then the multifrontal factor level by level and W, with the regularisation
// ladder of World::FactorizeLDLT (World.h:1212-1232): reg = 1e-5, 1e-4, ...
// while a pivot is not positive. One host synchronisation per rung.
int refactor_device(pdm_scene s) {
  auto t0 = std::chrono::steady_clock::now();
  HostScene& S = s->S;
  Factor& F = s->F;
  if (!s->fac_plan) {
    if (upload_factor(s, false) || build_fac_plan(s)) return 1;
    s->fac_plan = true;
  }
  if (upload_weights(s)) return 1;
  const int N = S.N;
  const long long nnzA = (long long)s->Ls.col.size();
  const hipStream_t st = s->stream;
  hipLaunchKernelGGL(pdmfac::k_fac_assemble_L, dim3(cdiv(nnzA, 256)), dim3(256), 0, st, s->fc_cptr.p, s->fc_ctet.p,
                     s->fc_cval.p, s->wr.p, s->wv.p, nnzA, s->Lval.p);
  hipLaunchKernelGGL(pdmfac::k_fac_kdiag, dim3(cdiv(N, 256)), dim3(256), 0, st, s->Lptr.p, s->Lcol.p, s->Lval.p,
                     s->mass.p, S.h, N, s->Kdiag.p);
  const int nlev = (int)F.level_ptr.size() - 1;
  // panel block width per level: 16 columns, fewer where 16 x front rows of
  // doubles would not fit the 160 KiB of LDS
  constexpr size_t kPanelLds = 160 * 1024 - 256;
  static bool panel_attr = false;
  if (!panel_attr) {
    HIPX(hipFuncSetAttribute((const void*)pdmfac::k_fac_panel, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)kPanelLds));
    panel_attr = true;
  }
  std::vector<int> nbl(nlev);
  for (int l = 0; l < nlev; ++l) {
    int nb = 16;
    while (nb > 1 && (size_t)nb * s->fc_lev_maxm[l] * sizeof(double) > kPanelLds) nb /= 2;
    if ((size_t)nb * s->fc_lev_maxm[l] * sizeof(double) > kPanelLds) {
      g_err = "device factorisation: a front has more rows than one block column fits in LDS";
      return 1;
    }
    nbl[l] = nb;
  }
  double reg = 0.0, reg_step = 1e-6;
  for (int tries = 0;; ++tries) {
    HIPX(hipMemsetAsync(s->fc_L.p, 0, (size_t)F.nnz * sizeof(double), st));
    if (s->fc_usize > 0) HIPX(hipMemsetAsync(s->fc_U.p, 0, (size_t)s->fc_usize * sizeof(double), st));
    HIPX(hipMemsetAsync(s->fc_fail.p, 0, sizeof(int), st));
    hipLaunchKernelGGL(pdmfac::k_fac_scatter, dim3(cdiv(nnzA, 256)), dim3(256), 0, st, s->Lval.p, s->fc_amap.p,
                       nnzA, s->fc_L.p);
    hipLaunchKernelGGL(pdmfac::k_fac_diag, dim3(cdiv(N, 256)), dim3(256), 0, st, s->fc_dpos.p, s->mass.p, S.h, reg,
                       N, s->fc_L.p);
    for (int l = 0; l < nlev; ++l) {
      const int ne = s->fc_ext_ptr[l + 1] - s->fc_ext_ptr[l];
      if (ne > 0)
        hipLaunchKernelGGL(pdmfac::k_fac_extend, dim3(ne), dim3(256), 0, st, s->fc_ext.p + s->fc_ext_ptr[l],
                           s->f_first.p, s->f_rowptr.p, s->f_lptr.p, s->fc_upos.p, s->f_child_ptr.p, s->f_child.p,
                           s->f_rel.p, s->f_relptr.p, s->fc_L.p, s->fc_U.p);
      const int nsn = F.level_ptr[l + 1] - F.level_ptr[l];
      hipLaunchKernelGGL(pdmfac::k_fac_panel, dim3(nsn), dim3(256), (size_t)nbl[l] * s->fc_lev_maxm[l] * sizeof(double),
                         st, s->f_lev.p + F.level_ptr[l], s->f_first.p, s->f_rowptr.p, s->f_lptr.p, s->fc_L.p, s->f_W.p,
                         s->f_WR.p, s->fc_fail.p, nbl[l]);
      const int nu = s->fc_upd_ptr[l + 1] - s->fc_upd_ptr[l];
      if (nu > 0)
        hipLaunchKernelGGL(pdmfac::k_fac_update, dim3(nu), dim3(256), 0, st, s->fc_upd.p + s->fc_upd_ptr[l],
                           s->f_first.p, s->f_rowptr.p, s->f_lptr.p, s->fc_upos.p, s->fc_L.p, s->fc_U.p, s->f_W.p,
                           s->f_WR.p);
    }
    HIPX(hipGetLastError());
    int fail = 0;
    HIPX(hipMemcpyAsync(&fail, s->fc_fail.p, sizeof(int), hipMemcpyDeviceToHost, st));
    HIPX(hipStreamSynchronize(st));
    if (!fail) break;
    if (tries == 40) {
      g_err = "factorisation failed: non-positive pivot";
      return 1;
    }
    reg_step *= 10;
    reg = reg_step;
  }
  F.reg_used = reg;
  if (build_top_z(s) || build_leaf_ops(s)) return 1;
  HIPX(hipStreamSynchronize(st));
  s->t_factor = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return 0;
}

// Re-factorisation after a material change (on the device).
int refactor(pdm_scene s) { return refactor_device(s); }


// Scene construction shared by pdm_scene_create and the host-only diagnostics.
int host_setup(const pdm_scene_desc* d, pdm_scene s) {
  HostScene& S = s->S;
  std::string err;
  if (!build_scene(S, d->n_vertices, d->x_rest, d->n_tets, d->tets, d->mass_same, d->total_mass, err)) {
    g_err = err;
    return 1;
  }
  S.nu_fwd = d->nu_fwd;
  S.nu_fwd_t.assign(S.T, d->nu_fwd);
  for (int t = 0; t < S.T; ++t) {
    int ct = S.tet_order[t];
    S.E[t] = d->E ? d->E[ct] : 2e5;
    S.nu_ctor[t] = d->nu_ctor ? d->nu_ctor[ct] : 0.4;
    if (d->nu_fwd_tet) S.nu_fwd_t[t] = d->nu_fwd_tet[ct];
  }
  if (d->n_bodies > 0 &&
      !set_body_masses(S, d->tets, d->n_bodies, d->body_of_tet, d->body_mass, d->mass_same, err)) {
    g_err = err;
    return 1;
  }
  for (int k = 0; k < d->n_attach; ++k) {
    int a = d->attach_idx[k];
    if (a < 0 || a >= S.N) {
      g_err = "attachment index out of range";
      return 1;
    }
    S.att_idx.push_back(S.iperm[a]);
    S.att_w.push_back(d->attach_w);
  }
  if (d->n_segments > 0 &&
      !add_muscles(S, d->x_rest, d->tets, d->n_segments, d->seg_start, d->seg_end, d->seg_group, d->n_groups,
                   d->k_muscle, err)) {
    g_err = err;
    return 1;
  }
  S.n_groups = std::max(S.n_groups, d->n_groups);
  S.h = d->h;
  S.max_iter = d->max_iter;
  S.damping = d->damping;
  for (int r = 0; r < 3; ++r) S.g[r] = d->gravity[r];
  S.gravity_on = d->gravity_on;
  S.contact = d->contact;
  S.max_iter1 = d->max_iter1;
  S.alpha = d->alpha;
  S.mu = d->mu;
  S.planes.assign(d->planes, d->planes + 6 * (size_t)d->n_planes);
  S.spheres.assign(d->spheres, d->spheres + 4 * (size_t)d->n_spheres);
  if (d->n_tris > 0) {
    if (!d->tris) {
      g_err = "n_tris > 0 without tris";
      return 1;
    }
    S.tris.assign(d->tris, d->tris + 9 * (size_t)d->n_tris);
    build_tri_bvh(S);
  }
  S.self_collision = d->self_collision;
  S.self_mu = d->self_mu;
  if (S.self_collision) build_self_mesh(S, d->tets, d->n_tets);
  S.underwater = d->underwater;
  if (S.underwater) {
    std::vector<int> cf = contour_faces(d->x_rest, d->tets, d->n_tets);
    S.cf.resize(cf.size());
    for (size_t k = 0; k < cf.size(); ++k) S.cf[k] = S.iperm[cf[k]];
    vertex_faces(S.N, S.cf, S.vfptr, S.vfq);
  }
  if (!(S.h > 0) || S.max_iter < 0) {
    g_err = "invalid stepping parameters";
    return 1;
  }
  update_weights(S);
  assemble_L(S, s->Ls);
  if (!symbolic(s->Ls, s->F, err)) {
    g_err = "symbolic analysis failed: " + err;
    return 1;
  }
  build_corners(S, s->Cn);
  build_strips(S, s->Cn, s->Sp);
  return 0;
}

// Numeric factorisation with the regularisation ladder of
// World::FactorizeLDLT (World.h:1212-1232): reg = 1e-5, 1e-4, ... on failure.
int host_factor(pdm_scene s) {
  HostScene& S = s->S;
  update_weights(S);
  assemble_L(S, s->Ls);
  std::vector<double> dadd(S.N);
  for (int i = 0; i < S.N; ++i) dadd[i] = S.mass[i] / (S.h * S.h);
  std::string err;
  bool ok = numeric(s->Ls, dadd, s->F, err);
  double reg = 1e-6;
  s->F.reg_used = 0;
  int tries = 0;
  while (!ok && tries < 40) {
    reg *= 10;
    for (int i = 0; i < S.N; ++i) dadd[i] = S.mass[i] / (S.h * S.h) + reg;
    ok = numeric(s->Ls, dadd, s->F, err);
    s->F.reg_used = reg;
    ++tries;
  }
  if (!ok) {
    g_err = "factorisation failed: " + err;
    return 1;
  }
  return 0;
}

int autotune_levels(pdm_scene s);

int ensure_batch_buffers(pdm_scene s) {
  size_t nc = s->NC();
  int N = s->S.N;
  if ((long long)s->S.T * ((s->R + 63) & ~63) >= (1LL << 31)) {
    g_err = "tets x rollouts must stay below 2^31 (element kernels index items in 32 bits)";
    return 1;
  }
  if (s->X.alloc(nc) || s->V.alloc(nc) || s->Xn.alloc(nc) || s->Cc.alloc(nc) || s->Xit.alloc(nc) ||
      alloc_rhs(s->B, nc, s->C, s->stream) ||
      s->Lx.alloc(nc))
    return 1;
  if (s->nfaces > 0 && (s->FF.alloc((size_t)s->nfaces * s->C) || s->GXf.alloc((size_t)3 * s->nfaces * s->C) ||
                        s->GWf.alloc((size_t)s->nfaces * s->C)))
    return 1;
  if (s->U.alloc(std::max<long long>(1, s->F.uptr.back()) * (size_t)s->C)) return 1;
  if (s->Ys.alloc(nc)) return 1;
  // fronts of the top blocks; 32 doubles of tail padding keep the product's
  // reads of columns past a chain's end inside the buffer
  if (s->top_on && s->TF.alloc((size_t)s->top_rows * s->C + 32)) return 1;
  if (s->leaf_on && s->XA.alloc((size_t)s->leaf_xrows * s->C + 32)) return 1;
  // forward slots are FP32 only in PDM_MIXED; the FP32 reverse sweep of
  // PDM_MIXED_ADJOINT writes FP32 slots into the same (larger) buffer
  size_t esz = s->precision == PDM_MIXED ? sizeof(float) : sizeof(double);
  if (s->corner.alloc((size_t)s->Cn.nslot * s->C * esz)) return 1;
  if (s->R % 64 == 0 && s->ssum.alloc((size_t)std::max(1, s->Sp.nsum) * s->C * esz)) return 1;
  if (s->part.alloc((size_t)cdiv(N, kChunk) * s->C)) return 1;
  if (s->err.alloc(s->R) || s->active.alloc(s->R) || s->iters_dev.alloc(s->R)) return 1;
  if (s->alist.alloc(s->R) || s->acount.alloc(1)) return 1;
  if (s->ncv > 0) {
    size_t ni = (size_t)s->ncv * s->R;
    if (s->Ori.alloc(nc) || s->Dn.alloc(nc) || s->Prev.alloc(nc) || s->win_st.alloc(ni) || s->win_o.alloc(ni) ||
        s->nseen.alloc(ni) || s->seen.alloc(kSeen * ni) || s->win_n.alloc(3 * ni) || s->cactive.alloc(s->R + 1) ||
        s->iactive.alloc(s->R) || s->cfirst.alloc(s->R) || s->fresh.alloc(s->R) || s->npass.alloc(s->R) ||
        s->citers.alloc(s->R) || s->clast.alloc(s->R) || s->cinj.alloc(s->R))
      return 1;
    if (!s->S.spheres.empty() && (s->win_sph.alloc(ni) || s->win_t.alloc(ni) || s->win_x1.alloc(3 * ni))) return 1;
  }
  if (s->ncs > 0) {
    size_t na = (size_t)s->ncs * s->R, nf = (size_t)s->nsf * s->R;
    if (s->s_nimp.alloc(na) || s->s_impf.alloc(kSelfImp * na) || s->s_impt.alloc(kSelfImp * na) ||
        s->s_impd.alloc(kSelfImp * na) || s->s_impn.alloc(3 * kSelfImp * na) || s->s_cface.alloc(na) ||
        s->s_gcount.alloc(nf) || s->s_glead.alloc(nf) || s->s_gmem.alloc(kSelfMem * nf) || s->s_ngrp.alloc(s->R) ||
        s->s_ngrp_prev.alloc(s->R) || s->s_Fo.alloc(nc) || s->s_CN.alloc(nc) || s->s_RC.alloc(nc) ||
        s->s_cur.alloc(nc) || s->s_GF.alloc(3 * nf * 3) || s->s_GV.alloc(na * 3))
      return 1;
  }
  if (s->tuned_C != s->C && autotune_levels(s)) return 1;
  return 0;
}

// Launch shape of the level kernels: column tile Ct = 16*nq right-hand sides
// (nq in {1,2,4}) and the block size. LDS per block = front (mpad x Ct) +
// scratch (wpad x Ct) doubles, at most 160 KB. The default shape keeps LDS
// within 64 KB where possible, shrinks nq while the level has too few blocks to
// fill the GPU, and sizes the block so that about 16 waves share a CU; the
// autotuner replaces it by the fastest measured shape (shapes only partition
// the columns and the work of a block, so results do not depend on them).
constexpr size_t kLdsSoft = 64 * 1024, kLdsMax = 160 * 1024;
size_t level_lds(const pdm_scene_s* s, int l, int Ct) { return (size_t)s->lev_rows[l] * Ct * 8; }

pdm_scene_s::LevelCfg level_default(const pdm_scene_s* s, int l, int C) {
  int cnt = s->lev_ptr[l + 1] - s->lev_ptr[l];
  int ncol16 = cdiv(C, 16);
  int nq = 4;
  while (nq > 1 && level_lds(s, l, 16 * nq) > kLdsSoft) nq /= 2;
  while (nq > 1 && (long long)cnt * cdiv(ncol16, nq) < 512) nq /= 2;
  pdm_scene_s::LevelCfg c;
  c.Ct = 16 * nq;
  size_t lds = level_lds(s, l, c.Ct);
  int per_cu = (int)std::max<size_t>(1, std::min<size_t>(8, kLdsMax / std::max<size_t>(lds, 1)));
  int waves = std::min(16, std::max(4, 16 / per_cu));
  if ((long long)cnt * cdiv(C, c.Ct) < 256) waves = 16;
  c.nthr = 64 * waves;
  return c;
}
pdm_scene_s::LevelCfg level_default(const pdm_scene_s* s, int l) { return level_default(s, l, s->C); }

// The grouped level kernels (k_fwd_level_g, k_bwd_level_g): LDS = the item's
// front rows x Ct plus its W slice; the default shape follows level_default.
size_t level_lds_g(const pdm_scene_s* s, int pass, int l, int Ct) {
  return ((size_t)s->item_front[pass][l] * Ct + (size_t)s->item_wsl[pass][l]) * 8;
}
pdm_scene_s::LevelCfg level_default_g(const pdm_scene_s* s, int pass, int l) {
  int cnt = s->item_ptr[pass][l + 1] - s->item_ptr[pass][l];
  int ncol16 = cdiv(s->C, 16);
  int nq = 4;
  while (nq > 1 && level_lds_g(s, pass, l, 16 * nq) > kLdsSoft) nq /= 2;
  while (nq > 1 && (long long)cnt * cdiv(ncol16, nq) < 512) nq /= 2;
  pdm_scene_s::LevelCfg c;
  c.Ct = 16 * nq;
  size_t lds = level_lds_g(s, pass, l, c.Ct);
  int per_cu = (int)std::max<size_t>(1, std::min<size_t>(8, kLdsMax / std::max<size_t>(lds, 1)));
  int waves = std::min(16, std::max(4, 16 / per_cu));
  if ((long long)cnt * cdiv(s->C, c.Ct) < 256) waves = 16;
  c.nthr = 64 * waves;
  return c;
}

// One level of one sweep over the columns [cbase, cbase + cw) of C on stream
// st, with the update blocks in Up.
int launch_level_into(pdm_scene s, int pass, int l, const pdm_scene_s::LevelCfg& c, int C, double* Bp, double* Up,
                      const int* any, int cbase, int cw, hipStream_t st) {
  const FactorView fv = s->view();
  int cnt = s->lev_ptr[l + 1] - s->lev_ptr[l];
  size_t lds = level_lds(s, l, c.Ct);
  if (lds > kLdsMax) {
    g_err = "supernode front exceeds the LDS budget";
    return 1;
  }
  int ntile = cdiv(cw, c.Ct);
  int lg = 0;
  while ((1 << lg) < c.Ct) ++lg;
  if ((1 << lg) != c.Ct || c.nthr % c.Ct != 0) {
    g_err = "invalid solve launch shape";
    return 1;
  }
  if (pass == 0)
    hipLaunchKernelGGL(k_fwd_level, dim3(cnt * ntile), dim3(c.nthr), lds, st, fv, s->f_lev.p + s->lev_ptr[l], ntile,
                       lg, C, Bp, Up, any, cbase, cw);
  else
    hipLaunchKernelGGL(k_bwd_level, dim3(cnt * ntile), dim3(c.nthr), lds, st, fv, s->f_lev.p + s->lev_ptr[l], ntile,
                       lg, C, Bp, any, cbase, cw);
  return 0;
}
// One level of one sweep with the grouped kernels (the solve's own path; y
// goes through the scene's Yb).
int launch_level(pdm_scene s, int pass, int l, const pdm_scene_s::LevelCfg& c, double* Bp, const int* any, int cbase,
                 int cw, hipStream_t st) {
  const FactorView fv = s->view();
  int cnt = s->item_ptr[pass][l + 1] - s->item_ptr[pass][l];
  size_t lds = level_lds_g(s, pass, l, c.Ct);
  if (lds > kLdsMax) {
    g_err = "supernode front exceeds the LDS budget";
    return 1;
  }
  int ntile = cdiv(cw, c.Ct);
  int lg = 0;
  while ((1 << lg) < c.Ct) ++lg;
  if ((1 << lg) != c.Ct || c.nthr % c.Ct != 0) {
    g_err = "invalid solve launch shape";
    return 1;
  }
  if (cnt == 0 || ntile == 0) return 0;
  if (pass == 0)
    hipLaunchKernelGGL(k_fwd_level_g, dim3(cnt * ntile), dim3(c.nthr), lds, st, fv, s->pview(),
                       s->f_items[0].p + s->item_ptr[0][l], ntile, lg, s->C, Bp, s->Ys.p, s->U.p, any, cbase, cw);
  else
    hipLaunchKernelGGL(k_bwd_level_g, dim3(cnt * ntile), dim3(c.nthr), lds, st, fv,
                       s->f_items[1].p + s->item_ptr[1][l], ntile, lg, s->C, Bp, s->Ys.p, any, cbase, cw);
  return 0;
}

// Dense top-separator blocks over the columns [cbase, cbase + cw): fronts,
// then x_t = Z f_t into B (between the sweeps' level kernels).
int launch_top(pdm_scene s, double* Bp, const int* any, int cbase, int cw, hipStream_t st) {
  if (cw <= 0) return 0;
  if (s->TF.n < (size_t)s->top_rows * s->C + 32) {
    g_err = "top separator fronts not allocated for this batch";
    return 1;
  }
  hipLaunchKernelGGL(k_top_gather, dim3(cdiv((long long)s->top_rows * cw, 256)), dim3(256), 0, st, s->t_row.p,
                     s->t_slot.p, s->top_maxp, s->top_rows, Bp, s->U.p, s->TF.p, s->C, cbase, cw, any);
  const int ncolt = cdiv(cw, 32);
  hipLaunchKernelGGL(k_top_apply, dim3(s->t_nitems * ncolt), dim3(256), 0, st, s->tZ.p, s->t_zoff.p, s->t_a0.p,
                     s->t_np.p, s->t_items.p, ncolt, s->TF.p, s->t_row.p, Bp, s->C, cbase, cw, any);
  return 0;
}

// Dense leaf subtrees over the columns [cbase, cbase + cw): pass 0 before the
// forward sweep (z into Ys, the roots' update blocks into U), pass 1 after the
// backward sweep (the above rows compacted, then x = z + G^T x_above into B).
int launch_leaf(pdm_scene s, int pass, double* Bp, const int* any, int cbase, int cw, hipStream_t st) {
  if (cw <= 0) return 0;
  const int ncolt = cdiv(cw, 32);
  const LeafView lv = s->lview();
  if (pass == 0) {
    if (s->l_nitems[0] > 0)
      hipLaunchKernelGGL(k_leaf_apply<true>, dim3(s->l_nitems[0] * ncolt), dim3(256), 0, st, lv, s->l_items[0].p,
                         ncolt, Bp, s->XA.p, s->Ys.p, s->U.p, Bp, s->C, cbase, cw, any);
    return 0;
  }
  if (s->XA.n < (size_t)s->leaf_xrows * s->C + 32) {
    g_err = "leaf above rows not allocated for this batch";
    return 1;
  }
  // (leaves that are whole trees have no rows above them)
  if (s->leaf_xrows > 0)
    hipLaunchKernelGGL(k_leaf_gather, dim3(cdiv((long long)s->leaf_xrows * cw, 256)), dim3(256), 0, st, s->l_xrow.p,
                       s->leaf_xrows, Bp, s->XA.p, s->C, cbase, cw, any);
  if (s->l_nitems[1] > 0)
    hipLaunchKernelGGL(k_leaf_apply<false>, dim3(s->l_nitems[1] * ncolt), dim3(256), 0, st, lv, s->l_items[1].p,
                       ncolt, Bp, s->XA.p, s->Ys.p, s->U.p, Bp, s->C, cbase, cw, any);
  return 0;
}

// Subtree kernels of cut level `cut` over the columns [cbase, cbase + cw).
int launch_subtrees(pdm_scene s, int pass, int cut, const pdm_scene_s::LevelCfg& c, double* Bp, const int* any,
                    int cbase, int cw, hipStream_t st) {
  const FactorView fv = s->view();
  int nsub = (int)s->sub_ranges[cut].size();
  size_t rows = 16;
  for (int l = 0; l < cut; ++l) rows = std::max<size_t>(rows, s->lev_rows[l]);
  size_t lds = rows * c.Ct * 8;
  if (lds > kLdsMax) {
    g_err = "supernode front exceeds the LDS budget";
    return 1;
  }
  int lg = 0;
  while ((1 << lg) < c.Ct) ++lg;
  int ntile = cdiv(cw, c.Ct);
  if (nsub == 0) return 0;
  if (pass == 0)
    hipLaunchKernelGGL(k_fwd_subtree, dim3(nsub * ntile), dim3(c.nthr), lds, st, fv, s->d_sub.p, ntile, lg, s->C, Bp,
                       s->U.p, any, cbase, cw, s->leaf_on ? s->d_skip.p : nullptr);
  else
    hipLaunchKernelGGL(k_bwd_subtree, dim3(nsub * ntile), dim3(c.nthr), lds, st, fv, s->d_sub.p, ntile, lg, s->C, Bp,
                       any, cbase, cw, s->leaf_on ? s->d_skip.p : nullptr);
  return 0;
}

// Device copy of the subtree ranges of cut level c (outside any capture).
int set_cut(pdm_scene s, int c) {
  s->cut = c;
  if (c > 0 && s->d_sub_cut != c) {
    const auto& r = s->sub_ranges[c];
    if (s->d_sub.upload(r.data(), std::max<size_t>(1, r.size()))) return 1;
    s->d_sub_cut = c;
  }
  return 0;
}

// Column chunks of the split solve: nsplit contiguous ranges, multiples of 16.
void split_columns(int C, int nsplit, std::vector<int>& base, std::vector<int>& width) {
  int n16 = cdiv(C, 16);
  nsplit = std::max(1, std::min(nsplit, n16));
  base.assign(nsplit, 0);
  width.assign(nsplit, 0);
  int b = 0;
  for (int i = 0; i < nsplit; ++i) {
    int w16 = n16 / nsplit + (i < n16 % nsplit ? 1 : 0);
    base[i] = b;
    width[i] = std::min(C - b, 16 * w16);
    b += 16 * w16;
  }
}

int set_level_attrs() {
  static bool attr_set = false;
  if (!attr_set) {
    HIPX(hipFuncSetAttribute((const void*)k_fwd_level, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kLdsMax));
    HIPX(hipFuncSetAttribute((const void*)k_bwd_level, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kLdsMax));
    HIPX(hipFuncSetAttribute((const void*)k_fwd_subtree, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kLdsMax));
    HIPX(hipFuncSetAttribute((const void*)k_bwd_subtree, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kLdsMax));
    HIPX(hipFuncSetAttribute((const void*)k_fwd_level_g, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kLdsMax));
    HIPX(hipFuncSetAttribute((const void*)k_bwd_level_g, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kLdsMax));
    attr_set = true;
  }
  return 0;
}

int launch_solve_levels(pdm_scene s, double* Bp, const int* any = nullptr);

// Autotuner: for every level and sweep, time each admissible launch shape
// (nq in {1,2,4} x {256,512,1024} threads) with HIP events on the library
// stream, on the engine's own rhs/update buffers (cleared first, before any
// right-hand side is placed there), and keep the fastest. Runs once per
// right-hand-side count.
int autotune_levels(pdm_scene s) {
  int nlev = (int)s->lev_ptr.size() - 1;
  s->cfg_fwd.resize(nlev);
  s->cfg_bwd.resize(nlev);
  for (int l = 0; l < nlev; ++l) {
    s->cfg_fwd[l] = level_default_g(s, 0, l);
    s->cfg_bwd[l] = level_default_g(s, 1, l);
  }
  if (set_level_attrs()) return 1;
  s->cut = 0;
  s->nsplit = 1;
  if (!s->autotune) {
    s->tuned_C = s->C;
    return 0;
  }
  if (set_level_attrs()) return 1;
  HIPX(hipMemsetAsync(s->B.p, 0, s->NC() * 8, s->stream));
  HIPX(hipMemsetAsync(s->U.p, 0, std::max<long long>(1, s->F.uptr.back()) * (size_t)s->C * 8, s->stream));
  HIPX(hipMemsetAsync(s->Ys.p, 0, s->NC() * 8, s->stream));
  hipEvent_t e0, e1;
  HIPX(hipEventCreate(&e0));
  HIPX(hipEventCreate(&e1));
  const int reps = 3;
  for (int pass = 0; pass < 2; ++pass)
    for (int l = 0; l < nlev; ++l) {
      float best = 1e30f;
      pdm_scene_s::LevelCfg bc = pass == 0 ? s->cfg_fwd[l] : s->cfg_bwd[l];
      for (int nq = 1; nq <= 4; nq *= 2)
        for (int nthr = 256; nthr <= 1024; nthr *= 2) {
          pdm_scene_s::LevelCfg c;
          c.Ct = 16 * nq;
          c.nthr = nthr;
          if (level_lds_g(s, pass, l, c.Ct) > kLdsMax) continue;
          if (launch_level(s, pass, l, c, s->B.p, nullptr, 0, s->C, s->stream)) return 1;  // warm-up
          HIPX(hipEventRecord(e0, s->stream));
          for (int r = 0; r < reps; ++r)
            if (launch_level(s, pass, l, c, s->B.p, nullptr, 0, s->C, s->stream)) return 1;
          HIPX(hipEventRecord(e1, s->stream));
          HIPX(hipEventSynchronize(e1));
          float ms = 0;
          HIPX(hipEventElapsedTime(&ms, e0, e1));
          if (ms < best) {
            best = ms;
            bc = c;
          }
        }
      (pass == 0 ? s->cfg_fwd : s->cfg_bwd)[l] = bc;
    }
  // whole-solve choices, timed with the tuned level shapes: the subtree cut
  // (levels below it run as whole subtrees) with its launch shape, then the
  // number of independent column chains on separate streams
  s->tuned_C = s->C;
  s->nsplit = 1;
  auto time_solve_now = [&](float* out) -> int {
    if (launch_solve_levels(s, s->B.p)) return 1;
    HIPX(hipEventRecord(e0, s->stream));
    for (int r = 0; r < reps; ++r)
      if (launch_solve_levels(s, s->B.p)) return 1;
    HIPX(hipEventRecord(e1, s->stream));
    HIPX(hipEventSynchronize(e1));
    HIPX(hipEventElapsedTime(out, e0, e1));
    return 0;
  };
  {
    float bestc = 1e30f;
    int best_cut = 0;
    pdm_scene_s::LevelCfg best_sc;
    const pdm_scene_s::LevelCfg cands[3] = {{16, 256}, {16, 512}, {32, 512}};
    for (int c = 0; c <= std::min(nlev - 1, 6); ++c)
      for (int k = 0; k < (c == 0 ? 1 : 3); ++k) {
        size_t rows = 16;
        for (int l = 0; l < c; ++l) rows = std::max<size_t>(rows, s->lev_rows[l]);
        if (rows * cands[k].Ct * 8 > kLdsMax) continue;
        if (set_cut(s, c)) return 1;
        s->cfg_sub = cands[k];
        float ms;
        if (time_solve_now(&ms)) return 1;
        if (ms < bestc) {
          bestc = ms;
          best_cut = c;
          best_sc = cands[k];
        }
      }
    if (set_cut(s, best_cut)) return 1;
    s->cfg_sub = best_sc;
  }
  float best = 1e30f;
  int best_split = 1;
  for (int ns = 1; ns <= kMaxSplit; ++ns) {
    if (ns > 1 && cdiv(s->C, 16) < ns) break;
    s->nsplit = ns;
    if (launch_solve_levels(s, s->B.p)) return 1;
    HIPX(hipEventRecord(e0, s->stream));
    for (int r = 0; r < reps; ++r)
      if (launch_solve_levels(s, s->B.p)) return 1;
    HIPX(hipEventRecord(e1, s->stream));
    HIPX(hipEventSynchronize(e1));
    float ms = 0;
    HIPX(hipEventElapsedTime(&ms, e0, e1));
    if (ms < best) {
      best = ms;
      best_split = ns;
    }
  }
  s->nsplit = best_split;
  HIPX(hipEventDestroy(e0));
  HIPX(hipEventDestroy(e1));
  HIPX(hipGetLastError());
  return 0;
}

// Sum the recorded solve event pairs into solve_ms (waits for the stream).
int fold_solve_events(pdm_scene s) {
  if (s->sev_used == 0) return 0;
  HIPX(hipStreamSynchronize(s->stream));
  for (size_t k = 0; k + 1 < s->sev_used; k += 2) {
    float ms = 0;
    HIPX(hipEventElapsedTime(&ms, s->sev[k], s->sev[k + 1]));
    s->solve_ms += ms;
  }
  s->sev_used = 0;
  return 0;
}


// any: optional device flag; the level kernels return at once when it is 0
// (no rollout iterates any more), which graph replays rely on.
int launch_solve(pdm_scene s, double* Bp, const int* any = nullptr) {
  if (!(s->flags & PDM_TIME_SOLVES) || s->capturing) return launch_solve_levels(s, Bp, any);
  if (s->sev_used + 2 > s->sev.size()) {
    if (s->sev.size() >= 4096) {
      if (fold_solve_events(s)) return 1;
    } else {
      size_t n0 = s->sev.size();
      s->sev.resize(n0 + 256);
      for (size_t k = n0; k < s->sev.size(); ++k) HIPX(hipEventCreate(&s->sev[k]));
    }
  }
  HIPX(hipEventRecord(s->sev[s->sev_used], s->stream));
  if (launch_solve_levels(s, Bp)) return 1;
  HIPX(hipEventRecord(s->sev[s->sev_used + 1], s->stream));
  s->sev_used += 2;
  s->n_solves += 1;
  return 0;
}

int launch_solve_levels(pdm_scene s, double* Bp, const int* any) {
  if (set_level_attrs()) return 1;
  if (s->tuned_C != s->C) {
    g_err = "solve launch shapes not set up for this batch";
    return 1;
  }
  int nlev = (int)s->lev_ptr.size() - 1;
  std::vector<int> cb, cwid;
  split_columns(s->C, s->nsplit, cb, cwid);
  const int ns = (int)cb.size();
  // chunk 0 runs on the library stream, the others on auxiliary streams that
  // fork from it and join it again (default priority)
  if (ns > 1) {
    if (s->ensure_aux_streams()) return 1;
    HIPX(hipEventRecord(s->ev_fork, s->stream));
    for (int k = 1; k < ns; ++k) HIPX(hipStreamWaitEvent(s->aux[k - 1], s->ev_fork, 0));
  }
  const int cut = s->cut;
  for (int k = 0; k < ns; ++k) {
    hipStream_t st = k == 0 ? s->stream : s->aux[k - 1];
    if (s->leaf_on && launch_leaf(s, 0, Bp, any, cb[k], cwid[k], st)) return 1;
    if (cut > 0 && launch_subtrees(s, 0, cut, s->cfg_sub, Bp, any, cb[k], cwid[k], st)) return 1;
    for (int l = cut; l < nlev; ++l)
      if (launch_level(s, 0, l, s->cfg_fwd[l], Bp, any, cb[k], cwid[k], st)) return 1;
    if (s->top_on && launch_top(s, Bp, any, cb[k], cwid[k], st)) return 1;
    for (int l = nlev - 1; l >= cut; --l)
      if (launch_level(s, 1, l, s->cfg_bwd[l], Bp, any, cb[k], cwid[k], st)) return 1;
    if (cut > 0 && launch_subtrees(s, 1, cut, s->cfg_sub, Bp, any, cb[k], cwid[k], st)) return 1;
    if (s->leaf_on && launch_leaf(s, 1, Bp, any, cb[k], cwid[k], st)) return 1;
  }
  if (ns > 1)
    for (int k = 1; k < ns; ++k) {
      HIPX(hipEventRecord(s->ev_join[k - 1], s->aux[k - 1]));
      HIPX(hipStreamWaitEvent(s->stream, s->ev_join[k - 1], 0));
    }
  HIPX(hipGetLastError());
  return 0;
}

template <typename T>
void launch_local(pdm_scene s, const double* Xsrc, const double* act, int act_steps, int step, const int* mask,
                  const int* step_dev = nullptr, const int* any = nullptr, bool gather = true) {
  HostScene& S = s->S;
  T* corner = (T*)s->corner.p;
  long long nt = (long long)S.T * s->R;
  const int G = S.n_groups;
  // element work only for the rollouts that still iterate
  launch_compact(mask, s->R, s->alist.p, s->acount.p, s->stream);
  // wave-uniform variant for kDenseBelow or more listed rollouts, dense one for fewer
  const bool uni = s->R % 64 == 0;
  if (uni)
    hipLaunchKernelGGL((k_local_strip<T>), dim3(cdiv((long long)s->Sp.nstrip * s->R, kThreads)), dim3(kThreads), 0,
                       s->stream, Xsrc, s->tv.p, s->tG.p, s->wr.p, s->wv.p, s->stet.p, s->scode.p, s->sfbase.p,
                       s->ssum_of.p, s->alist.p, s->acount.p, (T*)s->ssum.p, S.T, s->Sp.nstrip, s->R, s->mptr.p,
                       s->mg.p, s->mw.p, s->mgroup.p, S.M, act, act_steps * G, step * G, step_dev, G);
  hipLaunchKernelGGL((uni ? k_local_tet<T, true> : k_local_tet<T, false>), dim3(dense_blocks(S.T, s->R)), dim3(kThreads), 0, s->stream,
                     Xsrc, s->tv.p, s->tG.p, s->wr.p, s->wv.p, s->tslot.p, s->alist.p, s->acount.p, corner, S.T, s->R,
                     s->mptr.p, s->mg.p, s->mw.p, s->mgroup.p, S.M, act, act_steps * G, step * G, step_dev, G);
  if (uni)
    hipLaunchKernelGGL((k_strip_fold<T>),
                       dim3(std::min(cdiv((long long)s->Sp.nsum * 3 * (kDenseBelow - 1), kThreads), kDenseGridCap)),
                       dim3(kThreads), 0, s->stream, corner, s->sfptr.p, s->sfq.p, s->alist.p, s->acount.p,
                       (T*)s->ssum.p, s->Sp.nsum, s->R);
  if (gather)
    hipLaunchKernelGGL((uni ? k_gather_list<T, 0, true> : k_gather_list<T, 0, false>),
                       dim3(cdiv((long long)S.N * 3 * s->R, kThreads)), dim3(kThreads), 0, s->stream, s->Cc.p,
                       (const T*)s->fsum(), s->fsum_ptr(), s->alist.p, s->acount.p, s->B.p, S.N, s->R, any);
}


// Contact phase of one step (World.h:921-1023) for every rollout that has
// fresh impacts; the host reads back after each detection pass whether any
// rollout continues. Rollouts finish with x = ori_x.
// FP64 device-to-device copy on the scene's stream (two values per thread when
// both pointers are 16-byte aligned).
// dst (+ *slot * sstride doubles when slot is given) <- src, n doubles.
int dcopy(pdm_scene s, double* dst, const double* src, size_t n, const int* slot = nullptr, long long sstride = 0) {
  if (n == 0) return 0;
  bool al = (((uintptr_t)dst | (uintptr_t)src) & 15) == 0 && (sstride & 1) == 0;
  if (al)
    hipLaunchKernelGGL(k_copy_f64, dim3(cdiv((long long)(n + 1) / 2, 256)), dim3(256), 0, s->stream, dst, src,
                       (long long)n, slot, sstride);
  else
    hipLaunchKernelGGL(k_copy_f64_1, dim3(cdiv((long long)n, 256)), dim3(256), 0, s->stream, dst, src, (long long)n,
                       slot, sstride);
  HIPX(hipGetLastError());
  return 0;
}

int contact_phase(pdm_scene s, const double* act, int act_steps, int step, ContactTape* tp) {
  HostScene& S = s->S;
  int N = S.N, C = s->C, R = s->R;
  size_t nc = s->NC();
  int nchunk = cdiv(N, kChunk);
  dim3 ugrid(cdiv(C, 64), nchunk), ublock(64, 4);
  int np = (int)(S.planes.size() / 6), nsph = (int)(S.spheres.size() / 4);
  long long ni = (long long)s->ncv * R;
  if (dcopy(s, s->Ori.p, s->Xit.p, nc)) return 1;
  hipLaunchKernelGGL(k_contact_reset, dim3(cdiv(std::max<long long>(ni, R), kThreads)), dim3(kThreads), 0, s->stream,
                     s->win_st.p, s->win_o.p, s->nseen.p, ni, s->cactive.p, s->cfirst.p, s->fresh.p, s->clast.p,
                     s->npass.p, s->citers.p, R, nsph > 0 ? s->win_sph.p : nullptr);
  const bool self = s->ncs > 0;
  if (self) {
    long long na = (long long)s->ncs * R;
    hipLaunchKernelGGL(k_self_reset, dim3(cdiv(std::max<long long>(std::max<long long>(na, (long long)nc), R), kThreads)),
                       dim3(kThreads), 0, s->stream, s->s_nimp.p, na, s->s_cur.p, (long long)nc, s->s_ngrp.p,
                       s->s_ngrp_prev.p, R);
  }
  std::vector<int> ca(R + 1);
  size_t cvc = (size_t)s->ncv * C;
  if (tp) {
    tp->G = 0;
    tp->P = 0;
    tp->first.clear();
    tp->pass_of.clear();
    if (grow_keep(tp->xit, nc, 0, s->stream)) return 1;
    if (dcopy(s, tp->xit.p, s->Xit.p, nc)) return 1;
  }
  for (int pass = 0; pass <= S.outer_cap; ++pass) {
    hipLaunchKernelGGL(k_contact_detect, dim3(cdiv((long long)s->ncv * R, kThreads)), dim3(kThreads), 0, s->stream,
                       s->cvtx.p, s->ncv, s->Xn.p, s->Ori.p, R, s->d_planes.p, np, s->d_spheres.p, nsph, s->d_tris.p,
                       s->ntri, s->d_bvh_box.p, s->d_bvh_a.p, s->d_bvh_b.p, s->d_bvh_tri.p, s->win_st.p, s->win_o.p,
                       s->win_n.p, s->nseen.p, s->seen.p, pass, s->cactive.p, s->fresh.p,
                       nsph > 0 ? s->win_sph.p : nullptr, s->win_t.p, s->win_x1.p);
    if (self) {
      long long na = (long long)s->ncs * R, nf = (long long)s->nsf * R;
      hipLaunchKernelGGL(k_self_detect, dim3(cdiv(na, kThreads)), dim3(kThreads), 0, s->stream, s->d_sc_vtx.p,
                         s->d_sc_face.p, s->ncs, s->nsf, s->Xn.p, s->Ori.p, R, s->cactive.p, s->s_nimp.p, s->s_impf.p,
                         s->s_impt.p, s->s_impd.p, s->s_impn.p, s->cactive.p + R);
      hipLaunchKernelGGL(k_self_group_reset, dim3(cdiv(std::max<long long>(nf, R), kThreads)), dim3(kThreads), 0,
                         s->stream, s->s_gcount.p, s->s_glead.p, nf, s->s_ngrp.p, s->s_ngrp_prev.p, s->cactive.p, R);
      hipLaunchKernelGGL(k_self_closest, dim3(cdiv(na, kThreads)), dim3(kThreads), 0, s->stream, s->ncs, s->nsf, R,
                         s->cactive.p, s->s_nimp.p, s->s_impf.p, s->s_impt.p, s->s_impd.p, s->s_cface.p,
                         s->s_gcount.p, s->s_glead.p, s->s_gmem.p, s->s_ngrp.p, s->cactive.p + R);
    }
    if (pass == S.outer_cap) break;  // cap reached: keep ori_x
    hipLaunchKernelGGL(k_contact_pass_begin, dim3(cdiv(R, 64)), dim3(64), 0, s->stream, s->cactive.p, s->fresh.p,
                       s->iactive.p, s->npass.p, R, self ? s->s_ngrp.p : nullptr, self ? s->s_ngrp_prev.p : nullptr);
    HIPX(hipMemcpyAsync(ca.data(), s->cactive.p, (R + 1) * sizeof(int), hipMemcpyDeviceToHost, s->stream));
    HIPX(hipStreamSynchronize(s->stream));
    if (ca[R]) {
      g_err = "contact: a vertex met more than 8 distinct obstacle faces, or a self-contact list overflowed, in one step";
      return 1;
    }
    bool any = false;
    for (int r = 0; r < R; ++r) any |= ca[r] != 0;
    if (!any) break;
    if (tp) {
      // x_it before the pass's first iteration, for every rollout; within the
      // pass the executing rollouts only drop out, so the Jacobi kernel's own
      // store of x_it fills the later slots for every rollout that reads them
      if (tp->G > 0 && !self && dcopy(s, tp->xit.p + (size_t)tp->G * nc, s->Xit.p, nc)) return 1;
      size_t g1 = (size_t)tp->G + S.max_iter1;
      if (grow_keep(tp->xit, (g1 + 1) * nc, (tp->G + 1) * nc, s->stream) ||
          grow_keep(tp->rc, g1 * cvc, tp->G * cvc, s->stream) ||
          grow_keep(tp->nrm, (size_t)(tp->P + 1) * cvc, (size_t)tp->P * cvc, s->stream) ||
          grow_keep(tp->exec, g1 * R, (size_t)tp->G * R, s->stream))
        return 1;
      if (self) {
        const size_t P = (size_t)tp->P, NF = (size_t)s->nsf * R, NA = (size_t)s->ncs * R;
        const size_t NFc = (size_t)s->ncs * C;
        if (grow_keep(tp->gCount, (P + 1) * NF, P * NF, s->stream) ||
            grow_keep(tp->gLead, (P + 1) * NF, P * NF, s->stream) ||
            grow_keep(tp->gMem, (P + 1) * kSelfMem * NF, P * kSelfMem * NF, s->stream) ||
            grow_keep(tp->gNrm, (P + 1) * 3 * NF, P * 3 * NF, s->stream) ||
            grow_keep(tp->gOwn, (P + 1) * NA, P * NA, s->stream) || grow_keep(tp->gNgrp, (P + 1) * R, P * R, s->stream) ||
            grow_keep(tp->gF, g1 * NFc, (size_t)tp->G * NFc, s->stream) ||
            grow_keep(tp->gFlags, g1 * NF, (size_t)tp->G * NF, s->stream))
          return 1;
        hipLaunchKernelGGL(k_self_tape_pass, dim3(cdiv((long long)std::max(std::max(NF, NA), (size_t)R), kThreads)),
                           dim3(kThreads), 0, s->stream, s->ncs, s->nsf, R, s->s_gcount.p, s->s_glead.p, s->s_gmem.p,
                           s->s_cface.p, s->s_impf.p, s->s_impn.p, s->s_ngrp.p, tp->gCount.p + P * NF,
                           tp->gLead.p + P * NF, tp->gMem.p + P * kSelfMem * NF, tp->gNrm.p + P * 3 * NF,
                           tp->gOwn.p + P * NA, tp->gNgrp.p + P * R);
      }
      if (nsph > 0) {
        size_t P = (size_t)tp->P;
        if (grow_keep(tp->sinfo, (P + 1) * ni, P * ni, s->stream) || grow_keep(tp->st, (P + 1) * ni, P * ni, s->stream) ||
            grow_keep(tp->sx1, (P + 1) * 3 * ni, P * 3 * ni, s->stream) ||
            grow_keep(tp->dl, (P + 1) * R, P * R, s->stream))
          return 1;
        hipLaunchKernelGGL(k_contact_tape_sph, dim3(cdiv(std::max<long long>(ni, R), kThreads)), dim3(kThreads), 0,
                           s->stream, s->win_st.p, s->win_sph.p, s->win_t.p, s->win_x1.p, s->clast.p, ni, R,
                           tp->sinfo.p + P * ni, tp->st.p + P * ni, tp->sx1.p + P * 3 * ni, tp->dl.p + P * R);
      }
    }
    for (int i = 0; i < S.max_iter1; ++i) {
      hipLaunchKernelGGL(k_contact_begin, ugrid, ublock, 0, s->stream, s->Xit.p, s->Ori.p, s->iactive.p, s->part.p, N,
                         C, s->Xn.p, s->Dn.p);
      hipLaunchKernelGGL(k_contact_err, dim3(cdiv(R, 4)), dim3(256), 0, s->stream, s->part.p, nchunk, R,
                         s->iactive.p, s->cfirst.p, s->citers.p, s->clast.p,
                         tp ? tp->exec.p + (size_t)tp->G * R : nullptr);
      // FP64 corner slots: the Jacobi kernel gathers them itself
      const bool fused = s->precision != PDM_MIXED;
      if (fused)
        launch_local<double>(s, s->Ori.p, act, act_steps, step, s->iactive.p, nullptr, nullptr, false);
      else
        launch_local<float>(s, s->Ori.p, act, act_steps, step, s->iactive.p);
      hipLaunchKernelGGL((R % 64 == 0 ? k_contact_jacobi<true> : k_contact_jacobi<false>), dim3(cdiv((long long)N * R, kThreads)), dim3(kThreads), 0, s->stream,
                         s->Lptr.p, s->Lcol.p, s->Lval.p, s->B.p, s->Dn.p, s->Xn.p, s->Kdiag.p, s->cv_of_v.p,
                         s->win_st.p, s->win_n.p, S.mu, S.alpha, S.h, i == 0 ? 1 : 0, s->alist.p,
                         s->acount.p, s->Prev.p, s->Xit.p, N, R, tp ? tp->rc.p + (size_t)tp->G * cvc : nullptr,
                         tp && i == 0 ? tp->nrm.p + (size_t)tp->P * cvc : nullptr,
                         tp && !self ? tp->xit.p + (size_t)(tp->G + 1) * nc : nullptr, self ? s->s_cur.p : nullptr,
                         self ? s->s_Fo.p : nullptr, self ? s->s_CN.p : nullptr, self ? s->s_RC.p : nullptr,
                         self ? s->s_ngrp.p : nullptr, fused ? s->Cc.p : nullptr,
                         fused ? (const double*)s->fsum() : nullptr, fused ? s->fsum_ptr() : nullptr);
      if (self) {
        if (tp)
          hipLaunchKernelGGL(k_self_tape_F, dim3(cdiv((long long)s->ncs * C, kThreads)), dim3(kThreads), 0, s->stream,
                             s->d_sc_vtx.p, s->ncs, C, s->s_Fo.p, s->s_CN.p, tp->gF.p + (size_t)tp->G * s->ncs * C);
        hipLaunchKernelGGL(k_self_group, dim3(cdiv((long long)s->nsf * R, kThreads)), dim3(kThreads), 0, s->stream,
                           s->d_sc_vtx.p, s->d_sc_face.p, s->ncs, s->nsf, R, s->iactive.p, s->s_gcount.p,
                           s->s_glead.p, s->s_gmem.p, s->s_cface.p, s->s_impn.p, s->s_Fo.p, s->s_CN.p, s->Kdiag.p,
                           S.self_mu, s->s_GF.p, s->s_GV.p, tp ? tp->gFlags.p + (size_t)tp->G * s->nsf * R : nullptr);
        hipLaunchKernelGGL(k_self_finish, dim3(cdiv((long long)N * R, kThreads)), dim3(kThreads), 0, s->stream,
                           s->d_scf_ptr.p, s->d_scf_q.p, s->d_sc_of_v.p, s->ncs, R, s->s_gcount.p, s->s_glead.p,
                           s->s_impf.p, s->s_cface.p, s->s_ngrp.p,
                           s->s_GF.p, s->s_GV.p, s->s_RC.p, s->s_CN.p, s->s_Fo.p, s->s_cur.p, s->Prev.p, s->Xit.p,
                           s->Xn.p, s->Kdiag.p, S.alpha, S.h, i == 0 ? 1 : 0, s->alist.p, s->acount.p, N);
      }
      if (tp) {
        // with self contact k_self_finish also writes x_it: keep the whole copy
        if (self && dcopy(s, tp->xit.p + (size_t)(tp->G + 1) * nc, s->Xit.p, nc)) return 1;
        tp->first.push_back(i == 0 ? 1 : 0);
        tp->pass_of.push_back(tp->P);
        tp->G += 1;
      }
    }
    if (tp) tp->P += 1;
  }
  if (tp) {
    if (grow_keep(tp->flast, R, 0, s->stream)) return 1;
    HIPX(hipMemcpyAsync(tp->flast.p, s->clast.p, R * sizeof(int), hipMemcpyDeviceToDevice, s->stream));
  }
  if (dcopy(s, s->Xit.p, s->Ori.p, nc)) return 1;
  HIPX(hipGetLastError());
  return 0;
}

// Start of a step and its PD iterations (World.h:673-699, 835-919). any and
// step_dev non-null: device-side control for graph capture (no host reads;
// activations indexed by the device step counter; the gather and the solve
// return at once when no rollout iterates any more).
// slot_dev (graph capture of a taped step): tapeY, tapeXn and tapeVn are the
// tapes' slot-0 addresses and each write goes to the slot *slot_dev.
int launch_pd_block(pdm_scene s, const double* act, int act_steps, int step, double* tapeY, double* tapeXn,
                    double* tapeVn, bool host_exit, int* any, const int* step_dev, const int* slot_dev = nullptr) {
  HostScene& S = s->S;
  int N = S.N, C = s->C;
  size_t nc = s->NC();
  const long long K1nc = (long long)(S.max_iter + 1) * (long long)nc;
  int nchunk = cdiv(N, kChunk);
  dim3 ugrid(cdiv(C, 64), nchunk), ublock(64, 4);
  hipLaunchKernelGGL(k_spmv, dim3(cdiv(nc, kThreads)), dim3(kThreads), 0, s->stream, s->Lptr.p, s->Lcol.p, s->Lval.p,
                     s->X.p, s->Lx.p, N, C);
  if (s->nfaces > 0)
    hipLaunchKernelGGL(k_drag_face, dim3(cdiv((long long)s->nfaces * s->R, kThreads)), dim3(kThreads), 0, s->stream,
                       s->X.p, s->V.p, s->dcf.p, s->FF.p, s->nfaces, s->R, C);
  hipLaunchKernelGGL(k_step_prep, dim3(cdiv(nc, kThreads)), dim3(kThreads), 0, s->stream, s->X.p, s->V.p, s->Lx.p,
                     s->mass.p, s->attw.p, s->xrest.p, s->Xn.p, s->Cc.p, s->Xit.p, N, C, S.h, S.g[0], S.g[1], S.g[2],
                     S.gravity_on, s->nfaces > 0 ? s->FF.p : nullptr, s->dvfptr.p, s->dvfq.p);
  if (tapeXn && dcopy(s, tapeXn, s->Xn.p, nc, slot_dev, (long long)nc)) return 1;
  if (tapeVn && dcopy(s, tapeVn, s->V.p, nc, slot_dev, (long long)nc)) return 1;
  if (tapeY && dcopy(s, tapeY, s->Xit.p, nc, slot_dev, K1nc)) return 1;
  hipLaunchKernelGGL(k_sqsum_chunks, ugrid, ublock, 0, s->stream, s->Xit.p, s->part.p, N, C, any);
  hipLaunchKernelGGL(k_err_init, dim3(cdiv(s->R, 4)), dim3(256), 0, s->stream, s->part.p, nchunk, s->R, N,
                     s->active.p, s->iters_dev.p, any);
  std::vector<int> act_h(s->R);
  for (int it = 0; it < S.max_iter; ++it) {
    if (host_exit) {
      HIPX(hipMemcpyAsync(act_h.data(), s->active.p, s->R * sizeof(int), hipMemcpyDeviceToHost, s->stream));
      HIPX(hipStreamSynchronize(s->stream));
      bool some = false;
      for (int r = 0; r < s->R; ++r) some |= act_h[r] != 0;
      if (!some) break;
    }
    if (s->precision == PDM_MIXED)
      launch_local<float>(s, s->Xit.p, act, act_steps, step, s->active.p, step_dev, any);
    else
      launch_local<double>(s, s->Xit.p, act, act_steps, step, s->active.p, step_dev, any);
    if (launch_solve(s, s->B.p, any)) return 1;
    hipLaunchKernelGGL(k_update, ugrid, ublock, 0, s->stream, s->Xn.p, s->B.p, s->Xit.p, s->active.p, s->part.p,
                       tapeY ? tapeY + (size_t)(it + 1) * nc : nullptr, N, C, any, slot_dev, K1nc);
    hipLaunchKernelGGL(k_err_final, dim3(cdiv(s->R, 4)), dim3(256), 0, s->stream, s->part.p, nchunk, s->R, N,
                       s->active.p, s->iters_dev.p, s->err.p, any);
  }
  HIPX(hipGetLastError());
  return 0;
}

void drop_step_graph(pdm_scene s) {
  for (int k = 0; k < 2; ++k) {
    if (s->gexec[k]) (void)hipGraphExecDestroy(s->gexec[k]);
    s->gexec[k] = nullptr;
  }
}

// PDM_USE_GRAPH: the prep + PD block of a step, captured once per forward call
// (buffers and the activation pointer are fixed within it) and replayed every
// step after the device step counter is set. Graph 1 also writes the step's
// tapes, at the slot held by the device slot counter.
int ensure_step_graph(pdm_scene s, const double* act, int act_steps, bool taped) {
  hipGraphExec_t& gx = s->gexec[taped ? 1 : 0];
  if (gx && s->g_act == act && s->g_act_steps == act_steps) return 0;
  if (s->g_act != act || s->g_act_steps != act_steps) drop_step_graph(s);
  if (s->any_dev.alloc(1) || s->step_dev.alloc(2)) return 1;
  HIPX(hipStreamBeginCapture(s->stream, hipStreamCaptureModeRelaxed));
  s->capturing = true;
  int rc = taped ? launch_pd_block(s, act, act_steps, 0, s->tY.p, s->tXn.p, s->nfaces > 0 ? s->tVn.p : nullptr, false,
                                   s->any_dev.p, s->step_dev.p, s->step_dev.p + 1)
                 : launch_pd_block(s, act, act_steps, 0, nullptr, nullptr, nullptr, false, s->any_dev.p, s->step_dev.p);
  s->capturing = false;
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(s->stream, &g);
  if (rc || e != hipSuccess || !g) {
    if (g) (void)hipGraphDestroy(g);
    if (!rc) g_err = std::string("graph capture failed: ") + hipGetErrorString(e);
    return 1;
  }
  e = hipGraphInstantiate(&gx, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) {
    gx = nullptr;
    g_err = std::string("graph instantiation failed: ") + hipGetErrorString(e);
    return 1;
  }
  s->g_act = act;
  s->g_act_steps = act_steps;
  return 0;
}

__global__ void k_set_int2(int* __restrict__ p, int v0, int v1) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    p[0] = v0;
    p[1] = v1;
  }
}

// One time step on device. tapeY: optional (max_iter+1) x NC buffer for the
// iterates, slot `slot` of the segment tapes (tY, tXn, tVn) when they are
// given; tIters: optional R ints receiving the executed iterations.
int run_step(pdm_scene s, const double* act, int act_steps, int step, double* tapeY, double* tapeXn, double* tapeVn,
             int* tIters, int* tNpass = nullptr, int* tCiters = nullptr, ContactTape* ctp = nullptr, int slot = 0) {
  HostScene& S = s->S;
  size_t nc = s->NC();
  bool host_exit = (s->flags & PDM_HOST_EARLY_EXIT) != 0;
  const bool taped = tapeY != nullptr;
  if (s->flags & PDM_USE_GRAPH) {
    if (ensure_step_graph(s, act, act_steps, taped)) return 1;
    hipLaunchKernelGGL(k_set_int2, dim3(1), dim3(64), 0, s->stream, s->step_dev.p, step, slot);
    HIPX(hipGraphLaunch(s->gexec[taped ? 1 : 0], s->stream));
    ++s->graph_launches;
  } else if (launch_pd_block(s, act, act_steps, step, tapeY, tapeXn, tapeVn, host_exit, nullptr, nullptr)) {
    return 1;
  }
  if (s->ncv > 0 && contact_phase(s, act, act_steps, step, ctp)) return 1;
  const bool cs = s->ncv > 0;
  hipLaunchKernelGGL(k_step_finish, dim3(cdiv(nc, kThreads)), dim3(kThreads), 0, s->stream, s->Xit.p, s->Xn.p, s->X.p,
                     s->V.p, (long long)nc, S.h, S.damping, s->R, s->iters_dev.p, tIters, s->npass.p,
                     cs ? tNpass : nullptr, s->citers.p, cs ? tCiters : nullptr);
  HIPX(hipGetLastError());
  return 0;
}

int ensure_tapes(pdm_scene s, int steps);

int forward_impl(pdm_scene s, int steps, const double* x0, const double* v0, const double* act, double* xT, double* vT,
                 int record) {
  if (steps < 0) {
    g_err = "steps must be >= 0";
    return 1;
  }
  if (ensure_batch_buffers(s)) return 1;
  drop_step_graph(s);  // recaptured on the first step of this call
  HostScene& S = s->S;
  size_t nc = s->NC();
  int N = S.N;
  HIPX(hipEventRecord(s->ev0, s->stream));
  hipLaunchKernelGGL(k_caller_to_engine, dim3(cdiv(nc, kThreads)), dim3(kThreads), 0, s->stream, x0, s->X.p, s->perm.p,
                     N, s->R);
  hipLaunchKernelGGL(k_caller_to_engine, dim3(cdiv(nc, kThreads)), dim3(kThreads), 0, s->stream, v0, s->V.p, s->perm.p,
                     N, s->R);
  s->have_tape = false;
  int nck = record ? cdiv(std::max(steps, 1), s->ckpt) : 0;
  if (record) {
    if (s->ckX.alloc((size_t)nck * nc) || s->ckV.alloc((size_t)nck * nc)) return 1;
    size_t na = act ? (size_t)s->R * steps * S.n_groups : 0;
    if (na) {
      if (s->act_dev.alloc(na)) return 1;
      HIPX(hipMemcpyAsync(s->act_dev.p, act, na * 8, hipMemcpyDeviceToDevice, s->stream));
    }
    s->act_null = act == nullptr;
  }
  if (s->iters_host.size() < (size_t)s->R * steps) s->iters_host.resize((size_t)s->R * steps);
  DevBuf<int>& it_all = s->it_all;
  DevBuf<int>& np_all = s->np_all;
  DevBuf<int>& ci_all = s->ci_all;
  if (it_all.alloc((size_t)s->R * std::max(steps, 1))) return 1;
  if (s->ncv > 0 && (np_all.alloc((size_t)s->R * std::max(steps, 1)) || ci_all.alloc((size_t)s->R * std::max(steps, 1))))
    return 1;
  // The last checkpoint segment is recorded as it runs, so that the reverse
  // sweep starts without recomputing it.
  s->fwd_taped_seg = -1;
  const int last_s0 = record && steps > 0 ? ((steps - 1) / s->ckpt) * s->ckpt : steps;
  if (record && steps > 0 && ensure_tapes(s, steps)) return 1;
  const int K1 = S.max_iter + 1;
  for (int n = 0; n < steps; ++n) {
    if (record && n % s->ckpt == 0) {
      size_t k = n / s->ckpt;
      if (dcopy(s, s->ckX.p + k * nc, s->X.p, nc) || dcopy(s, s->ckV.p + k * nc, s->V.p, nc)) return 1;
    }
    int* np = s->ncv ? np_all.p + (size_t)n * s->R : nullptr;
    int* ci = s->ncv ? ci_all.p + (size_t)n * s->R : nullptr;
    if (n >= last_s0) {
      int q = n - last_s0;
      if (run_step(s, act, steps, n, s->tY.p + (size_t)q * K1 * nc, s->tXn.p + (size_t)q * nc,
                   s->nfaces > 0 ? s->tVn.p + (size_t)q * nc : nullptr, s->tIters.p + (size_t)q * s->R, np, ci,
                   s->ncv > 0 ? s->ctape[q].get() : nullptr, q))
        return 1;
      HIPX(hipMemcpyAsync(it_all.p + (size_t)n * s->R, s->tIters.p + (size_t)q * s->R, s->R * sizeof(int),
                          hipMemcpyDeviceToDevice, s->stream));
    } else if (run_step(s, act, steps, n, nullptr, nullptr, nullptr, it_all.p + (size_t)n * s->R, np, ci)) {
      return 1;
    }
  }
  if (record && steps > 0) s->fwd_taped_seg = (steps - 1) / s->ckpt;
  hipLaunchKernelGGL(k_engine_to_caller, dim3(cdiv(nc, kThreads)), dim3(kThreads), 0, s->stream, s->X.p, xT, s->perm.p,
                     N, s->R);
  hipLaunchKernelGGL(k_engine_to_caller, dim3(cdiv(nc, kThreads)), dim3(kThreads), 0, s->stream, s->V.p, vT, s->perm.p,
                     N, s->R);
  HIPX(hipEventRecord(s->ev1, s->stream));
  std::vector<int> ih((size_t)s->R * steps);
  if (steps) HIPX(hipMemcpyAsync(ih.data(), it_all.p, ih.size() * sizeof(int), hipMemcpyDeviceToHost, s->stream));
  std::vector<int> nph(s->ncv && steps ? ih.size() : 0), cih(nph.size());
  if (!nph.empty()) {
    HIPX(hipMemcpyAsync(nph.data(), np_all.p, nph.size() * sizeof(int), hipMemcpyDeviceToHost, s->stream));
    HIPX(hipMemcpyAsync(cih.data(), ci_all.p, cih.size() * sizeof(int), hipMemcpyDeviceToHost, s->stream));
  }
  HIPX(hipStreamSynchronize(s->stream));
  HIPX(hipGetLastError());
  float ms = 0;
  HIPX(hipEventElapsedTime(&ms, s->ev0, s->ev1));
  s->t_fwd = ms;
  // iters_host layout: [r][step]
  s->npass_host.assign((size_t)s->R * steps, 0);
  s->citers_host.assign((size_t)s->R * steps, 0);
  for (int n = 0; n < steps; ++n)
    for (int r = 0; r < s->R; ++r) {
      s->iters_host[(size_t)r * steps + n] = ih[(size_t)n * s->R + r];
      if (!nph.empty()) {
        s->npass_host[(size_t)r * steps + n] = nph[(size_t)n * s->R + r];
        s->citers_host[(size_t)r * steps + n] = cih[(size_t)n * s->R + r];
      }
    }
  s->last_steps = steps;
  if (record) {
    s->have_tape = true;
    s->tape_steps = steps;
  }
  return 0;
}

template <typename T>
int vjp_iteration(pdm_scene s, const int* active, const double* Yk, const double* Yk1, const double* act,
                  int act_steps, int step, const double* CDp = nullptr, const double* Xn = nullptr,
                  bool gather = true) {
  HostScene& S = s->S;
  T* corner = (T*)s->corner.p;
  long long nt = (long long)S.T * s->R;
  // active null: the rollout list is already compacted
  if (active)
    launch_compact(active, s->R, s->alist.p, s->acount.p, s->stream);
  const bool uni = s->R % 64 == 0;
  if (uni)
    hipLaunchKernelGGL((k_vjp_tet<T, true>), dim3(cdiv(nt, kThreads)), dim3(kThreads), 0, s->stream, Yk, Yk1,
                       s->Zb.p, s->tv.p, s->tG.p, s->wr.p, s->wv.p, s->dEr.p, s->dEv.p, s->tslot.p, s->alist.p,
                       s->acount.p, corner, s->gEpart.p, S.T, s->R, CDp ? Xn : nullptr, CDp, s->mptr.p, s->mg.p,
                       s->mw.p, s->mgroup.p, S.M, act, act_steps * S.n_groups, step * S.n_groups,
                       S.M > 0 ? s->gMpart.p : nullptr, s->gEw.p);
  hipLaunchKernelGGL((k_vjp_tet<T, false>), dim3(cdiv((long long)S.T * (uni ? kDenseBelow - 1 : s->R), kThreads)),
                     dim3(kThreads), 0, s->stream,
                     Yk, Yk1, s->Zb.p, s->tv.p, s->tG.p, s->wr.p, s->wv.p, s->dEr.p, s->dEv.p, s->tslot.p, s->alist.p,
                     s->acount.p, corner, s->gEpart.p, S.T, s->R, CDp ? Xn : nullptr, CDp, s->mptr.p, s->mg.p,
                     s->mw.p, s->mgroup.p, S.M, act, act_steps * S.n_groups, step * S.n_groups,
                     S.M > 0 ? s->gMpart.p : nullptr, s->gEw.p);
  if (!gather) return 0;  // the caller gathers the corner slots itself
  hipLaunchKernelGGL((uni ? k_gather_list<T, 1, true> : k_gather_list<T, 1, false>),
                     dim3(cdiv((long long)S.N * 3 * s->R, kThreads)), dim3(kThreads), 0, s->stream, nullptr, corner,
                     s->vptr.p, s->alist.p, s->acount.p, s->Yb.p, S.N, s->R, nullptr);
  return 0;
}


int contact_reverse_iter(pdm_scene s, ContactTape& tp, int g, const double* Xn, const double* act, int act_steps,
                         int step, const int* dinj, bool sph, int nsph);

// Reverse of the taped contact phase of one step (branches frozen; sphere
// normals differentiated, other normals constant). In: Yb = adjoint of the step's final positions. Out: Yb =
// adjoint of the PD result; Xb, Zs, gEpart, gMpart accumulate.
int contact_reverse(pdm_scene s, ContactTape& tp, const double* Xn, const double* act, int act_steps, int step) {
  HostScene& S = s->S;
  int N = S.N, C = s->C, R = s->R, G = tp.G;
  size_t nc = s->NC(), cvc = (size_t)s->ncv * C;
  std::vector<int> ex((size_t)G * R), fl(R);
  HIPX(hipMemcpyAsync(ex.data(), tp.exec.p, ex.size() * sizeof(int), hipMemcpyDeviceToHost, s->stream));
  HIPX(hipMemcpyAsync(fl.data(), tp.flast.p, R * sizeof(int), hipMemcpyDeviceToHost, s->stream));
  HIPX(hipStreamSynchronize(s->stream));
  std::vector<int> glast(R, -1);
  for (int g = 0; g < G; ++g)
    for (int r = 0; r < R; ++r)
      if (ex[(size_t)g * R + r]) glast[r] = g;
  if (dcopy(s, s->cGf.p, s->Yb.p, nc)) return 1;
  hipLaunchKernelGGL(k_cadj_init, dim3(cdiv(nc, kThreads)), dim3(kThreads), 0, s->stream, s->cXB.p, s->Yb.p,
                     tp.flast.p, (long long)nc, C);
  HIPX(hipMemsetAsync(s->cPB.p, 0, nc * 8, s->stream));
  // injection flags per (iteration, rollout); the host copy stays alive until
  // the next call, whose first synchronisation ends this upload
  std::vector<int>& inj = s->cinj_h;
  inj.assign((size_t)G * R, 0);
  for (int g = 0; g < G; ++g)
    for (int r = 0; r < R; ++r) inj[(size_t)g * R + r] = (g == glast[r] && !fl[r]) ? 1 : 0;
  if (s->cinjG.alloc(inj.size())) return 1;
  DevBuf<int>& dinj = s->cinjG;
  HIPX(hipMemcpyAsync(dinj.p, inj.data(), inj.size() * sizeof(int), hipMemcpyHostToDevice, s->stream));
  // sphere normals: adjoints of the positions each detection pass read, added
  // into the x_it adjoint where the reverse sweep reaches those positions
  const int nsph = (int)(S.spheres.size() / 4);
  const bool sph = nsph > 0 && tp.P > 0;
  const long long ni = (long long)s->ncv * R;
  if (sph) {
    if (s->cNB.alloc((size_t)tp.P * cvc)) return 1;
    HIPX(hipMemsetAsync(s->cNB.p, 0, (size_t)tp.P * cvc * 8, s->stream));
  }
  auto add_nb = [&](int pass, int want) {
    hipLaunchKernelGGL(k_cadj_nb, dim3(cdiv(ni, kThreads)), dim3(kThreads), 0, s->stream, s->cXB.p,
                       s->cNB.p + (size_t)pass * cvc, tp.dl.p + (size_t)pass * R, want, s->cvtx.p, ni, R);
  };
  if (s->ncs > 0) HIPX(hipMemsetAsync(s->cSCB.p, 0, nc * sizeof(double), s->stream));
  int pend_input = -1;  // pass whose input-reading rollouts get their NB after the next iteration
  for (int g = G - 1; g >= 0; --g) {
    bool any = false;
    for (int r = 0; r < R; ++r) any |= ex[(size_t)g * R + r] != 0;
    if (any && contact_reverse_iter(s, tp, g, Xn, act, act_steps, step, dinj.p, sph, nsph)) return 1;
    if (sph) {
      if (pend_input >= 0) add_nb(pend_input, 0);
      pend_input = -1;
      if (tp.first[g]) {
        add_nb(tp.pass_of[g], 1);
        pend_input = tp.pass_of[g];
      }
    }
  }
  if (dcopy(s, s->Yb.p, s->cXB.p, nc)) return 1;
  HIPX(hipGetLastError());
  return 0;
}

// One executed contact iteration g of the reverse sweep.
int contact_reverse_iter(pdm_scene s, ContactTape& tp, int g, const double* Xn, const double* act, int act_steps,
                         int step, const int* dinj, bool sph, int nsph) {
  HostScene& S = s->S;
  int N = S.N, C = s->C, R = s->R;
  size_t nc = s->NC(), cvc = (size_t)s->ncv * C;
  const long long ni = (long long)s->ncv * R;
  const size_t pg = (size_t)tp.pass_of[g];
  {
    const double* ori = tp.xit.p + (size_t)g * nc;
    launch_compact(tp.exec.p + (size_t)g * R, R, s->alist.p, s->acount.p, s->stream);
    if (s->ncs > 0) {
      const size_t NF = (size_t)s->nsf * R, NA = (size_t)s->ncs * R;
      hipLaunchKernelGGL((R % 64 == 0 ? k_sadj_pre<true> : k_sadj_pre<false>), dim3(cdiv((long long)N * R, kThreads)),
                         dim3(kThreads), 0, s->stream, s->cXB.p, s->cPB.p, s->Xb.p,
                         tp.xit.p + (size_t)(g + 1) * nc, Xn, s->Kdiag.p, S.alpha, S.h, (int)tp.first[g], s->alist.p,
                         s->acount.p, s->cQB.p, s->cQF.p, s->cCD.p, N, R);
      hipLaunchKernelGGL(k_sadj_group, dim3(cdiv((long long)NF, kThreads)), dim3(kThreads), 0, s->stream,
                         s->d_sc_vtx.p, s->d_sc_face.p, s->ncs, s->nsf, R, tp.exec.p + (size_t)g * R,
                         tp.gNgrp.p + pg * R, tp.gCount.p + pg * NF, tp.gMem.p + pg * kSelfMem * NF,
                         tp.gNrm.p + pg * 3 * NF, tp.gFlags.p + (size_t)g * NF, tp.gF.p + (size_t)g * s->ncs * C,
                         s->Kdiag.p, S.self_mu, s->cQB.p, s->cSCB.p, s->sSFb.p, s->sSFm.p, s->sSVb.p, s->sSVm.p);
      hipLaunchKernelGGL(k_sadj_finish, dim3(cdiv((long long)N * R, kThreads)), dim3(kThreads), 0, s->stream,
                         s->d_scf_ptr.p, s->d_scf_q.p, s->d_sc_of_v.p, R, tp.gNgrp.p + pg * R, tp.gCount.p + pg * NF,
                         tp.gLead.p + pg * NF, tp.gOwn.p + pg * NA, s->sSFb.p, s->sSFm.p, s->sSVb.p, s->sSVm.p,
                         s->cQB.p, s->cQF.p, s->cv_of_v.p, tp.rc.p + (size_t)g * cvc, tp.nrm.p + pg * cvc, S.mu, S.h,
                         Xn, s->d_spheres.p, nsph, sph ? tp.sinfo.p + pg * ni : nullptr,
                         sph ? tp.st.p + pg * ni : nullptr, sph ? tp.sx1.p + pg * 3 * ni : nullptr, s->cNB.p,
                         (long long)cvc, s->alist.p, s->acount.p, s->Xb.p, s->cCD.p, s->cSCB.p, s->Zb.p, s->Zs.p, N);
    } else
    hipLaunchKernelGGL((R % 64 == 0 ? k_cadj_pre<true> : k_cadj_pre<false>), dim3(cdiv((long long)N * R, kThreads)), dim3(kThreads), 0, s->stream, s->cXB.p,
                       s->cPB.p, s->Xb.p, tp.xit.p + (size_t)(g + 1) * nc, Xn, s->Kdiag.p, s->cv_of_v.p,
                       tp.rc.p + (size_t)g * cvc, tp.nrm.p + pg * cvc, S.mu, S.alpha, S.h, (int)tp.first[g],
                       s->alist.p, s->acount.p, s->Zb.p, s->Zs.p, s->cCD.p, N, R, s->d_spheres.p, nsph,
                       sph ? tp.sinfo.p + pg * ni : nullptr, sph ? tp.st.p + pg * ni : nullptr,
                       sph ? tp.sx1.p + pg * 3 * ni : nullptr, s->cNB.p, (long long)cvc);
    // local-step VJP into the corner slots, then the iteration's end (C z,
    // the corner sums, x_it_{g-1}'s adjoint) in one kernel
    const dim3 gg(cdiv(nc, kThreads));
    const bool uni = R % 64 == 0;
    if (s->precision != PDM_FP64) {
      if (vjp_iteration<float>(s, nullptr, ori, ori, act, act_steps, step, s->cCD.p, Xn, false)) return 1;
      hipLaunchKernelGGL((uni ? k_cadj_gather<float, true> : k_cadj_gather<float, false>), gg, dim3(kThreads), 0,
                         s->stream, s->Lptr.p, s->Lcol.p, s->Lval.p, s->Zb.p, s->Xb.p, (const float*)s->corner.p,
                         s->vptr.p, s->cGf.p, dinj + (size_t)g * R, s->alist.p, s->acount.p, s->cXB.p, N, R);
    } else {
      if (vjp_iteration<double>(s, nullptr, ori, ori, act, act_steps, step, s->cCD.p, Xn, false)) return 1;
      hipLaunchKernelGGL((uni ? k_cadj_gather<double, true> : k_cadj_gather<double, false>), gg, dim3(kThreads), 0,
                         s->stream, s->Lptr.p, s->Lcol.p, s->Lval.p, s->Zb.p, s->Xb.p, (const double*)s->corner.p,
                         s->vptr.p, s->cGf.p, dinj + (size_t)g * R, s->alist.p, s->acount.p, s->cXB.p, N, R);
    }
  }
  return 0;
}

// Tape buffers for one checkpoint segment: every iterate of every step, the
// step-start states, iteration counts and the contact-phase tapes.
int ensure_tapes(pdm_scene s, int steps) {
  HostScene& S = s->S;
  size_t nc = s->NC();
  int K1 = S.max_iter + 1;
  int seg = std::min(s->ckpt, std::max(steps, 1));
  if (s->tY.alloc((size_t)seg * K1 * nc) || s->tXn.alloc((size_t)seg * nc) || s->tIters.alloc((size_t)seg * s->R))
    return 1;
  if (s->nfaces > 0 && s->tVn.alloc((size_t)seg * nc)) return 1;
  if (s->ncv > 0)
    while ((int)s->ctape.size() < seg) s->ctape.emplace_back(new ContactTape());
  return 0;
}

int backward_impl(pdm_scene s, const double* gxT, const double* gvT, double* gx0, double* gv0, double* gE, double* gnu,
                  double* gact) {
  if (!s->have_tape) {
    g_err = "backward needs a recorded forward (record=1)";
    return 1;
  }
  HostScene& S = s->S;
  int N = S.N, C = s->C, R = s->R, steps = s->tape_steps;
  size_t nc = s->NC();
  int K1 = S.max_iter + 1;
  int seg = std::min(s->ckpt, std::max(steps, 1));
  if (ensure_tapes(s, steps)) return 1;
  if (s->gX.alloc(nc) || s->gV.alloc(nc) || s->Yb.alloc(nc) || s->Xb.alloc(nc) || s->Zs.alloc(nc) ||
      alloc_rhs(s->Zb, nc, C, s->stream) || s->LZ.alloc(nc) || s->gEpart.alloc((size_t)S.T * R) ||
      s->gEw.alloc((size_t)S.T * std::max(1, R / 64)) ||
      s->ractive.alloc(R))
    return 1;
  if (S.M > 0 && s->gMpart.alloc((size_t)S.M * R)) return 1;
  if (s->ncv > 0) {
    if (s->cXB.alloc(nc) || s->cPB.alloc(nc) || s->cGf.alloc(nc) || s->cCD.alloc((size_t)N * R))
      return 1;
    if (s->ncs > 0 && (s->cQB.alloc(nc) || s->cQF.alloc(nc) || s->cSCB.alloc(nc) ||
                       s->sSFb.alloc((size_t)3 * s->nsf * C) || s->sSFm.alloc((size_t)3 * s->nsf * R) ||
                       s->sSVb.alloc((size_t)s->ncs * C) || s->sSVm.alloc((size_t)s->ncs * R)))
      return 1;
  }
  size_t nact = (size_t)R * steps * std::max(S.n_groups, 1);
  if (s->gact.alloc(nact)) return 1;
  HIPX(hipEventRecord(s->ev0, s->stream));
  HIPX(hipMemsetAsync(s->gEpart.p, 0, (size_t)S.T * R * 8, s->stream));
  HIPX(hipMemsetAsync(s->gEw.p, 0, (size_t)S.T * std::max(1, R / 64) * 8, s->stream));
  HIPX(hipMemsetAsync(s->gact.p, 0, nact * 8, s->stream));
  hipLaunchKernelGGL(k_caller_to_engine, dim3(cdiv(nc, kThreads)), dim3(kThreads), 0, s->stream, gxT, s->gX.p,
                     s->perm.p, N, R);
  hipLaunchKernelGGL(k_caller_to_engine, dim3(cdiv(nc, kThreads)), dim3(kThreads), 0, s->stream, gvT, s->gV.p,
                     s->perm.p, N, R);
  const double* act = s->act_null ? nullptr : s->act_dev.p;
  int nseg = cdiv(std::max(steps, 1), s->ckpt);
  std::vector<int> it_h((size_t)seg * R);
  double gh = S.damping / S.h;
  for (int sg = nseg - 1; sg >= 0 && steps > 0; --sg) {
    int s0 = sg * s->ckpt, s1 = std::min(steps, s0 + s->ckpt);
    // recompute the segment from its checkpoint, recording every iterate
    // (the last segment was recorded by the forward pass itself)
    bool taped = sg == s->fwd_taped_seg;
    s->fwd_taped_seg = -1;
    if (!taped) {
      if (dcopy(s, s->X.p, s->ckX.p + (size_t)sg * nc, nc) || dcopy(s, s->V.p, s->ckV.p + (size_t)sg * nc, nc)) return 1;
    }
    for (int n = s0; n < s1 && !taped; ++n) {
      int q = n - s0;
      if (run_step(s, act, steps, n, s->tY.p + (size_t)q * K1 * nc, s->tXn.p + (size_t)q * nc,
                   s->nfaces > 0 ? s->tVn.p + (size_t)q * nc : nullptr, s->tIters.p + (size_t)q * R, nullptr, nullptr,
                   s->ncv > 0 ? s->ctape[q].get() : nullptr, q))
        return 1;
    }
    HIPX(hipMemcpyAsync(it_h.data(), s->tIters.p, (size_t)(s1 - s0) * R * sizeof(int), hipMemcpyDeviceToHost,
                        s->stream));
    HIPX(hipStreamSynchronize(s->stream));
    for (int n = s1 - 1; n >= s0; --n) {
      int q = n - s0;
      int kmax = 0;
      for (int r = 0; r < R; ++r) kmax = std::max(kmax, it_h[(size_t)q * R + r]);
      hipLaunchKernelGGL(k_adj_seed, dim3(cdiv(nc, kThreads)), dim3(kThreads), 0, s->stream, s->gX.p, s->gV.p, s->Yb.p,
                         s->Xb.p, s->Zs.p, (long long)nc, gh);
      const double* Y = s->tY.p + (size_t)q * K1 * nc;
      if (S.M > 0) HIPX(hipMemsetAsync(s->gMpart.p, 0, (size_t)S.M * R * 8, s->stream));
      if (s->ncv > 0 && s->ctape[q]->G > 0 &&
          contact_reverse(s, *s->ctape[q], s->tXn.p + (size_t)q * nc, act, steps, n))
        return 1;
      for (int k = kmax - 1; k >= 0; --k) {
        hipLaunchKernelGGL(k_active_rev, dim3(cdiv(R, 64)), dim3(64), 0, s->stream, s->tIters.p + (size_t)q * R, k, R,
                           s->ractive.p);
        hipLaunchKernelGGL(k_copy_masked, dim3(cdiv(nc, kThreads)), dim3(kThreads), 0, s->stream, s->Yb.p, s->Zb.p,
                           s->ractive.p, (long long)nc, C);
        if (launch_solve(s, s->Zb.p)) return 1;
        hipLaunchKernelGGL(k_adj_post_solve, dim3(cdiv(nc, kThreads)), dim3(kThreads), 0, s->stream, s->Zb.p,
                           s->Yb.p, s->Xb.p, s->Zs.p, s->ractive.p, (long long)nc, C);
        int rc = s->precision != PDM_FP64
                     ? vjp_iteration<float>(s, s->ractive.p, Y + (size_t)k * nc, Y + (size_t)(k + 1) * nc, act,
                                            steps, n)
                     : vjp_iteration<double>(s, s->ractive.p, Y + (size_t)k * nc, Y + (size_t)(k + 1) * nc, act,
                                             steps, n);
        if (rc) return 1;
      }
      if (S.M > 0)
        hipLaunchKernelGGL(k_reduce_groups, dim3(cdiv(S.n_groups * R, 64)), dim3(64), 0, s->stream, s->gMpart.p,
                           s->gptr.p, s->gmus.p, s->gact.p, S.n_groups, R, steps, n);
      hipLaunchKernelGGL(k_spmv, dim3(cdiv(nc, kThreads)), dim3(kThreads), 0, s->stream, s->Lptr.p, s->Lcol.p,
                         s->Lval.p, s->Zs.p, s->LZ.p, N, C);
      hipLaunchKernelGGL(k_adj_step_end, dim3(cdiv(nc, kThreads)), dim3(kThreads), 0, s->stream, s->Xb.p, s->Yb.p,
                         s->LZ.p, s->Zs.p, s->mass.p, s->gX.p, s->gV.p, N, C, S.h);
      if (s->nfaces > 0) {
        // gV holds dL/dv* here; the drag kick v* = v + h M^-1 f(x, v) adds its reverse
        hipLaunchKernelGGL(k_drag_vjp, dim3(cdiv((long long)s->nfaces * R, kThreads)), dim3(kThreads), 0, s->stream,
                           s->tXn.p + (size_t)q * nc, s->tVn.p + (size_t)q * nc, s->gV.p, s->mass.p, s->dcf.p,
                           s->GXf.p, s->GWf.p, s->nfaces, R, C, S.h);
        hipLaunchKernelGGL(k_drag_gather, dim3(cdiv(nc, kThreads)), dim3(kThreads), 0, s->stream, s->GXf.p, s->GWf.p,
                           s->dvfptr.p, s->dvfq.p, s->gX.p, s->gV.p, N, C);
      }
    }
  }
  if (gx0)
    hipLaunchKernelGGL(k_engine_to_caller, dim3(cdiv(nc, kThreads)), dim3(kThreads), 0, s->stream, s->gX.p, gx0,
                       s->perm.p, N, R);
  if (gv0)
    hipLaunchKernelGGL(k_engine_to_caller, dim3(cdiv(nc, kThreads)), dim3(kThreads), 0, s->stream, s->gV.p, gv0,
                       s->perm.p, N, R);
  DevBuf<double>& gEd = s->gEd;
  if (gEd.alloc(S.T)) return 1;
  hipLaunchKernelGGL(k_reduce_rows, dim3(cdiv(S.T, kThreads)), dim3(kThreads), 0, s->stream, s->gEpart.p, s->gEw.p,
                     s->tet_order.p, gEd.p, S.T, R, std::max(1, R / 64));
  if (gE) HIPX(hipMemcpyAsync(gE, gEd.p, (size_t)S.T * 8, hipMemcpyDeviceToDevice, s->stream));
  if (gact && S.n_groups > 0)
    HIPX(hipMemcpyAsync(gact, s->gact.p, (size_t)R * steps * S.n_groups * 8, hipMemcpyDeviceToDevice, s->stream));
  HIPX(hipEventRecord(s->ev1, s->stream));
  std::vector<double> gEh(S.T);
  HIPX(hipMemcpyAsync(gEh.data(), gEd.p, (size_t)S.T * 8, hipMemcpyDeviceToHost, s->stream));
  HIPX(hipStreamSynchronize(s->stream));
  HIPX(hipGetLastError());
  if (gnu) {
    // mMu = E/(1+nu_t): dL/dnu (one nu shifted for every tet) = -sum_t E_t dL/dE_t / (1+nu_t)
    double val = 0;
    for (int t = 0; t < S.T; ++t) val -= S.E[t] * gEh[S.tet_order[t]] / (1.0 + S.nu_fwd_t[t]);
    // on the scene's stream: a null-stream copy would wait for every other stream
    HIPX(hipMemcpyAsync(gnu, &val, 8, hipMemcpyHostToDevice, s->stream));
    HIPX(hipStreamSynchronize(s->stream));
  }
  float ms = 0;
  HIPX(hipEventElapsedTime(&ms, s->ev0, s->ev1));
  s->t_bwd = ms;
  return 0;
}

}  // namespace

// ===================================================================== C ABI
extern "C" {

const char* pdm_last_error(void) { return g_err.c_str(); }

int pdm_device_count(int* n) {
  int c = 0;
  hipError_t e = hipGetDeviceCount(&c);
  if (e != hipSuccess) c = 0;
  *n = c;
  return 0;
}

int pdm_scene_create(const pdm_scene_desc* d, pdm_scene* out) {
  if (!d || !out) {
    g_err = "null argument";
    return 1;
  }
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    g_err = "no HIP device available";
    return 1;
  }
  std::unique_ptr<pdm_scene_s> s(new pdm_scene_s());
  if (host_setup(d, s.get())) return 1;
  HIPX(hipStreamCreate(&s->stream));
  s->own_stream = true;
  HIPX(hipEventCreate(&s->ev0));
  HIPX(hipEventCreate(&s->ev1));
  if (upload_static(s.get())) return 1;
  if (refactor(s.get())) return 1;
  *out = s.release();
  return 0;
}

int pdm_host_analyse(const pdm_scene_desc* d, pdm_info* info) {
  if (!d || !info) {
    g_err = "null argument";
    return 1;
  }
  std::unique_ptr<pdm_scene_s> s(new pdm_scene_s());
  if (host_setup(d, s.get())) return 1;
  if (host_factor(s.get())) return 1;
  s->R = 1;
  return pdm_scene_info(s.get(), info);
}

int pdm_host_solve(const pdm_scene_desc* d, int R, const double* b, double* x) {
  if (!d || !b || !x || R < 1) {
    g_err = "invalid argument";
    return 1;
  }
  std::unique_ptr<pdm_scene_s> s(new pdm_scene_s());
  if (host_setup(d, s.get())) return 1;
  if (host_factor(s.get())) return 1;
  const Factor& F = s->F;
  int N = s->S.N;
  // supernodal forward / backward substitution on the host, same schedule
  // as the device kernels (update blocks per supernode, children in order)
  for (int r = 0; r < R; ++r) {
    std::vector<double> y(3 * (size_t)N);
    for (int v = 0; v < N; ++v)
      for (int c = 0; c < 3; ++c) y[3 * v + c] = b[(size_t)r * 3 * N + 3 * s->S.perm[v] + c];
    std::vector<std::vector<double>> U(F.ns);
    for (int sn = 0; sn < F.ns; ++sn) {
      int c0 = F.first[sn], w = F.first[sn + 1] - c0;
      int m = (int)(F.rowptr[sn + 1] - F.rowptr[sn]);
      const double* Ls = &F.L[F.lptr[sn]];
      std::vector<double> Tm((size_t)m * 3, 0.0);
      for (int i = 0; i < w; ++i)
        for (int c = 0; c < 3; ++c) Tm[3 * i + c] = y[3 * (c0 + i) + c];
      for (int q = F.child_ptr[sn]; q < F.child_ptr[sn + 1]; ++q) {
        int ch = F.child[q];
        int mbc = (int)(F.rowptr[ch + 1] - F.rowptr[ch]) - (F.first[ch + 1] - F.first[ch]);
        const int* rel = &F.rel[F.relptr[ch]];
        for (int a = 0; a < mbc; ++a)
          for (int c = 0; c < 3; ++c) Tm[3 * rel[a] + c] += U[ch][3 * a + c];
        std::vector<double>().swap(U[ch]);
      }
      // y = inv(L_ss) T[0:w], then T_below -= L_below y
      const double* Di = &F.Dinv[F.dptr[sn]];
      std::vector<double> yb((size_t)w * 3);
      for (int r = 0; r < w; ++r)
        for (int c = 0; c < 3; ++c) {
          double acc = 0.0;
          for (int k = 0; k <= r; ++k) acc += Di[(size_t)r * w + k] * Tm[3 * k + c];
          yb[3 * r + c] = acc;
        }
      for (int i = w; i < m; ++i)
        for (int c = 0; c < 3; ++c) {
          double acc = 0.0;
          for (int k = 0; k < w; ++k) acc += Ls[(size_t)k * m + i] * yb[3 * k + c];
          Tm[3 * i + c] -= acc;
        }
      for (int r = 0; r < w; ++r)
        for (int c = 0; c < 3; ++c) Tm[3 * r + c] = yb[3 * r + c];
      for (int i = 0; i < w; ++i)
        for (int c = 0; c < 3; ++c) y[3 * (c0 + i) + c] = Tm[3 * i + c];
      U[sn].assign(Tm.begin() + 3 * (size_t)w, Tm.end());
    }
    for (int sn = F.ns - 1; sn >= 0; --sn) {
      int c0 = F.first[sn], w = F.first[sn + 1] - c0;
      int m = (int)(F.rowptr[sn + 1] - F.rowptr[sn]);
      const int* rows = &F.rows[F.rowptr[sn]];
      const double* Ls = &F.L[F.lptr[sn]];
      // t = y_s - L_below^T x_below, x_s = inv(L_ss)^T t
      const double* Di = &F.Dinv[F.dptr[sn]];
      std::vector<double> tb((size_t)w * 3);
      for (int j = 0; j < w; ++j)
        for (int c = 0; c < 3; ++c) {
          double acc = y[3 * (c0 + j) + c];
          for (int i = w; i < m; ++i) acc -= Ls[(size_t)j * m + i] * y[3 * rows[i] + c];
          tb[3 * j + c] = acc;
        }
      for (int j = 0; j < w; ++j)
        for (int c = 0; c < 3; ++c) {
          double acc = 0.0;
          for (int k = j; k < w; ++k) acc += Di[(size_t)k * w + j] * tb[3 * k + c];
          y[3 * (c0 + j) + c] = acc;
        }
    }
    for (int v = 0; v < N; ++v)
      for (int c = 0; c < 3; ++c) x[(size_t)r * 3 * N + 3 * s->S.perm[v] + c] = y[3 * v + c];
  }
  return 0;
}

void pdm_scene_destroy(pdm_scene s) {
  if (!s) return;
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  for (hipEvent_t e : s->sev) (void)hipEventDestroy(e);
  for (int k = 0; k < kMaxSplit - 1; ++k) {
    if (s->aux[k]) (void)hipStreamDestroy(s->aux[k]);
    if (s->ev_join[k]) (void)hipEventDestroy(s->ev_join[k]);
  }
  if (s->ev_fork) (void)hipEventDestroy(s->ev_fork);
  drop_step_graph(s);
  if (s->ev0) (void)hipEventDestroy(s->ev0);
  if (s->ev1) (void)hipEventDestroy(s->ev1);
  if (s->own_stream && s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
}

int pdm_scene_info(pdm_scene s, pdm_info* info) {
  if (!s || !info) {
    g_err = "null argument";
    return 1;
  }
  info->n_vertices = s->S.N;
  info->n_tets = s->S.T;
  info->n_muscles = s->S.M;
  info->n_groups = s->S.n_groups;
  info->n_supernodes = s->F.ns;
  info->n_levels = (int)s->F.level_ptr.size() - 1;
  info->max_front = s->F.max_m;
  info->max_width = s->F.max_w;
  info->nnz_factor = (long long)s->F.nnz;
  {
    long long nl = 0;
    for (int sn = 0; sn < s->F.ns; ++sn) {
      long long w = s->F.first[sn + 1] - s->F.first[sn];
      long long m = s->F.rowptr[sn + 1] - s->F.rowptr[sn];
      nl += w * (w + 1) / 2 + (m - w) * w;
    }
    info->nnz_lower = nl;
  }
  info->factor_flops = s->F.flops_factor;
  info->reg_used = s->F.reg_used;
  info->n_contact_vertices = (int)s->S.contact_vtx.size();
  info->n_drag_faces = s->nfaces;
  info->n_obstacle_tris = s->ntri;
  info->n_self_faces = s->nsf;
  info->top_rows = s->top_on ? s->top_rows : 0;
  info->top_blocks = s->top_on ? (int)s->top_a0.size() : 0;
  info->leaf_blocks = s->leaf_on ? (int)s->leaf_np.size() : 0;
  info->batch = s->R;
  info->precision = s->precision;
  return 0;
}

int pdm_set_stream(pdm_scene s, void* stream) {
  if (!s) {
    g_err = "null scene";
    return 1;
  }
  HIPX(hipStreamSynchronize(s->stream));
  if (stream) {
    if (s->own_stream) (void)hipStreamDestroy(s->stream);
    s->stream = (hipStream_t)stream;
    s->own_stream = false;
  } else if (!s->own_stream) {
    HIPX(hipStreamCreate(&s->stream));
    s->own_stream = true;
  }
  return 0;
}

int pdm_set_material(pdm_scene s, const double* E, double nu_fwd) {
  if (!s) {
    g_err = "null scene";
    return 1;
  }
  HIPX(hipStreamSynchronize(s->stream));
  HostScene& S = s->S;
  if (E)
    for (int t = 0; t < S.T; ++t) S.E[t] = E[S.tet_order[t]];
  if (!std::isnan(nu_fwd)) {
    S.nu_fwd = nu_fwd;
    S.nu_fwd_t.assign(S.T, nu_fwd);
  }
  s->have_tape = false;
  return refactor(s);
}

int pdm_refactor(pdm_scene s, int on_host) {
  if (!s) {
    g_err = "null scene";
    return 1;
  }
  HIPX(hipStreamSynchronize(s->stream));
  s->have_tape = false;
  return on_host ? refactor_host(s) : refactor_device(s);
}

int pdm_get_factor_w(pdm_scene s, double* W, long long n) {
  if (!s || !W || n != (long long)s->F.nnz) {
    g_err = "pdm_get_factor_w: n must equal nnz_factor";
    return 1;
  }
  HIPX(hipMemcpyAsync(W, s->f_W.p, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, s->stream));
  HIPX(hipStreamSynchronize(s->stream));
  return 0;
}

int pdm_set_batch(pdm_scene s, int R, int precision, int flags, int checkpoint_every) {
  if (!s || R < 1 || (precision != PDM_FP64 && precision != PDM_MIXED && precision != PDM_MIXED_ADJOINT) ||
      checkpoint_every < 1) {
    g_err = "invalid batch arguments";
    return 1;
  }
  HIPX(hipStreamSynchronize(s->stream));
  s->R = R;
  s->C = 3 * R;
  s->precision = precision;
  s->flags = flags;
  bool at = (flags & PDM_NO_AUTOTUNE) == 0;
  if (at != s->autotune) s->tuned_C = -1;
  s->autotune = at;
  s->ckpt = checkpoint_every;
  s->have_tape = false;
  return 0;
}

int pdm_forward(pdm_scene s, int steps, const double* x0, const double* v0, const double* act, double* xT, double* vT,
                int record) {
  if (!s || !x0 || !v0 || !xT || !vT) {
    g_err = "null argument";
    return 1;
  }
  return forward_impl(s, steps, x0, v0, act, xT, vT, record);
}

int pdm_forward_host(pdm_scene s, int steps, const double* x0, const double* v0, const double* act, double* xT,
                     double* vT, int record) {
  if (!s || !x0 || !v0 || !xT || !vT) {
    g_err = "null argument";
    return 1;
  }
  size_t n = (size_t)s->R * 3 * s->S.N;
  size_t na = act ? (size_t)s->R * steps * s->S.n_groups : 0;
  DevBuf<double> dx, dv, da, dxt, dvt;
  if (dx.upload(x0, n) || dv.upload(v0, n) || dxt.alloc(n) || dvt.alloc(n)) return 1;
  if (na && da.upload(act, na)) return 1;
  if (forward_impl(s, steps, dx.p, dv.p, na ? da.p : nullptr, dxt.p, dvt.p, record)) return 1;
  HIPX(hipMemcpy(xT, dxt.p, n * 8, hipMemcpyDeviceToHost));
  HIPX(hipMemcpy(vT, dvt.p, n * 8, hipMemcpyDeviceToHost));
  return 0;
}

int pdm_backward(pdm_scene s, const double* gxT, const double* gvT, double* gx0, double* gv0, double* gE, double* gnu,
                 double* gact) {
  if (!s || !gxT || !gvT) {
    g_err = "null argument";
    return 1;
  }
  return backward_impl(s, gxT, gvT, gx0, gv0, gE, gnu, gact);
}

int pdm_backward_host(pdm_scene s, const double* gxT, const double* gvT, double* gx0, double* gv0, double* gE,
                      double* gnu, double* gact) {
  if (!s || !gxT || !gvT) {
    g_err = "null argument";
    return 1;
  }
  size_t n = (size_t)s->R * 3 * s->S.N;
  size_t na = (size_t)s->R * s->tape_steps * s->S.n_groups;
  DevBuf<double> dgx, dgv, ox, ov, oE, onu, oa;
  if (dgx.upload(gxT, n) || dgv.upload(gvT, n) || ox.alloc(n) || ov.alloc(n) || oE.alloc(s->S.T) || onu.alloc(1) ||
      oa.alloc(std::max<size_t>(1, na)))
    return 1;
  if (backward_impl(s, dgx.p, dgv.p, ox.p, ov.p, oE.p, onu.p, na ? oa.p : nullptr)) return 1;
  if (gx0) HIPX(hipMemcpy(gx0, ox.p, n * 8, hipMemcpyDeviceToHost));
  if (gv0) HIPX(hipMemcpy(gv0, ov.p, n * 8, hipMemcpyDeviceToHost));
  if (gE) HIPX(hipMemcpy(gE, oE.p, (size_t)s->S.T * 8, hipMemcpyDeviceToHost));
  if (gnu) HIPX(hipMemcpy(gnu, onu.p, 8, hipMemcpyDeviceToHost));
  if (gact && na) HIPX(hipMemcpy(gact, oa.p, na * 8, hipMemcpyDeviceToHost));
  return 0;
}

int pdm_get_iterations(pdm_scene s, int* iters) {
  if (!s || !iters) {
    g_err = "null argument";
    return 1;
  }
  std::copy(s->iters_host.begin(), s->iters_host.begin() + (size_t)s->R * s->last_steps, iters);
  return 0;
}

int pdm_get_contact_stats(pdm_scene s, int* passes, int* jacobi_iters) {
  if (!s || !passes || !jacobi_iters) {
    g_err = "null argument";
    return 1;
  }
  size_t n = (size_t)s->R * s->last_steps;
  for (size_t k = 0; k < n; ++k) {
    passes[k] = k < s->npass_host.size() ? s->npass_host[k] : 0;
    jacobi_iters[k] = k < s->citers_host.size() ? s->citers_host[k] : 0;
  }
  return 0;
}

int pdm_solve_host(pdm_scene s, const double* b, double* x) {
  if (!s || !b || !x) {
    g_err = "null argument";
    return 1;
  }
  if (ensure_batch_buffers(s)) return 1;
  size_t n = s->NC();
  DevBuf<double> db;
  if (db.upload(b, n)) return 1;
  hipLaunchKernelGGL(k_caller_to_engine, dim3(cdiv(n, kThreads)), dim3(kThreads), 0, s->stream, db.p, s->B.p,
                     s->perm.p, s->S.N, s->R);
  if (launch_solve(s, s->B.p)) return 1;
  hipLaunchKernelGGL(k_engine_to_caller, dim3(cdiv(n, kThreads)), dim3(kThreads), 0, s->stream, s->B.p, db.p,
                     s->perm.p, s->S.N, s->R);
  HIPX(hipStreamSynchronize(s->stream));
  HIPX(hipMemcpy(x, db.p, n * 8, hipMemcpyDeviceToHost));
  return 0;
}

int pdm_solve_noalloc(pdm_scene s, int R, const double* b, double* x) {
  if (!s || !b || !x || R < 1) {
    g_err = "invalid argument";
    return 1;
  }
  const int N = s->S.N;
  const unsigned long long C = 3ull * (unsigned long long)R;
  const unsigned long long urows = (unsigned long long)std::max<long long>(1, s->F.uptr.back());
  auto up256 = [](unsigned long long v) { return (v + 255ull) & ~255ull; };
  const unsigned long long bbytes = up256((unsigned long long)N * C * 8ull), ubytes = up256(urows * C * 8ull);
  if (bbytes + ubytes > PDM_NOALLOC_SCRATCH_BYTES) {
    g_err = "pdm_solve_noalloc: " + std::to_string(R) + " rollouts need " + std::to_string(bbytes + ubytes) +
            " bytes of scratch, above PDM_NOALLOC_SCRATCH_BYTES (" + std::to_string(PDM_NOALLOC_SCRATCH_BYTES) + ")";
    return 1;
  }
  double* base = nullptr;
  if (noalloc_scratch(&base)) return 1;
  double* Bp = base;
  double* Up = (double*)((char*)base + bbytes);
  if (set_level_attrs()) return 1;
  const int Ci = (int)C;
  const size_t n = (size_t)N * C;
  hipLaunchKernelGGL(k_caller_to_engine, dim3(cdiv(n, kThreads)), dim3(kThreads), 0, s->stream, b, Bp, s->perm.p, N,
                     R);
  const int nlev = (int)s->lev_ptr.size() - 1;
  for (int l = 0; l < nlev; ++l)
    if (launch_level_into(s, 0, l, level_default(s, l, Ci), Ci, Bp, Up, nullptr, 0, Ci, s->stream)) return 1;
  for (int l = nlev - 1; l >= 0; --l)
    if (launch_level_into(s, 1, l, level_default(s, l, Ci), Ci, Bp, Up, nullptr, 0, Ci, s->stream)) return 1;
  hipLaunchKernelGGL(k_engine_to_caller, dim3(cdiv(n, kThreads)), dim3(kThreads), 0, s->stream, Bp, x, s->perm.p, N,
                     R);
  HIPX(hipGetLastError());
  return 0;
}

int pdm_force_host(pdm_scene s, const double* x, const double* act_step, double* Jd) {
  if (!s || !x || !Jd) {
    g_err = "null argument";
    return 1;
  }
  if (ensure_batch_buffers(s)) return 1;
  size_t n = s->NC();
  DevBuf<double> dx, da;
  if (dx.upload(x, n)) return 1;
  size_t na = act_step ? (size_t)s->R * s->S.n_groups : 0;
  if (na && da.upload(act_step, na)) return 1;
  hipLaunchKernelGGL(k_caller_to_engine, dim3(cdiv(n, kThreads)), dim3(kThreads), 0, s->stream, dx.p, s->Xit.p,
                     s->perm.p, s->S.N, s->R);
  std::vector<int> ones(s->R, 1);
  HIPX(hipMemcpy(s->active.p, ones.data(), s->R * sizeof(int), hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_attach_rhs, dim3(cdiv(n, kThreads)), dim3(kThreads), 0, s->stream, s->attw.p, s->xrest.p,
                     s->Cc.p, s->S.N, s->C);
  if (s->precision == PDM_MIXED)
    launch_local<float>(s, s->Xit.p, na ? da.p : nullptr, 1, 0, s->active.p);
  else
    launch_local<double>(s, s->Xit.p, na ? da.p : nullptr, 1, 0, s->active.p);
  hipLaunchKernelGGL(k_engine_to_caller, dim3(cdiv(n, kThreads)), dim3(kThreads), 0, s->stream, s->B.p, dx.p,
                     s->perm.p, s->S.N, s->R);
  HIPX(hipStreamSynchronize(s->stream));
  HIPX(hipMemcpy(Jd, dx.p, n * 8, hipMemcpyDeviceToHost));
  return 0;
}

int pdm_time_solve(pdm_scene s, int reps, double* ms_per_solve) {
  if (!s || reps < 1 || !ms_per_solve) {
    g_err = "invalid argument";
    return 1;
  }
  if (ensure_batch_buffers(s)) return 1;
  if (launch_solve(s, s->B.p)) return 1;  // warm-up
  HIPX(hipEventRecord(s->ev0, s->stream));
  for (int k = 0; k < reps; ++k)
    if (launch_solve(s, s->B.p)) return 1;
  HIPX(hipEventRecord(s->ev1, s->stream));
  HIPX(hipEventSynchronize(s->ev1));
  float ms = 0;
  HIPX(hipEventElapsedTime(&ms, s->ev0, s->ev1));
  *ms_per_solve = ms / reps;
  return 0;
}

int pdm_get_solve_shapes(pdm_scene s, int cap, int* n_levels, int* ct_fwd, int* thr_fwd, int* ct_bwd, int* thr_bwd,
                         int* n_chains, int* subtree_cut) {
  if (!s || !n_levels || cap < 0) {
    g_err = "invalid argument";
    return 1;
  }
  if (ensure_batch_buffers(s)) return 1;
  int nlev = (int)s->cfg_fwd.size();
  *n_levels = nlev;
  if (n_chains) {
    std::vector<int> cb, cw;
    split_columns(s->C, s->nsplit, cb, cw);
    *n_chains = (int)cb.size();
  }
  if (subtree_cut) *subtree_cut = s->cut;
  for (int l = 0; l < std::min(cap, nlev); ++l) {
    if (ct_fwd) ct_fwd[l] = s->cfg_fwd[l].Ct;
    if (thr_fwd) thr_fwd[l] = s->cfg_fwd[l].nthr;
    if (ct_bwd) ct_bwd[l] = s->cfg_bwd[l].Ct;
    if (thr_bwd) thr_bwd[l] = s->cfg_bwd[l].nthr;
  }
  return 0;
}

int pdm_set_solve_plan(pdm_scene s, int subtree_cut, int n_chains) {
  if (!s) {
    g_err = "null scene";
    return 1;
  }
  if (ensure_batch_buffers(s)) return 1;
  int nlev = (int)s->lev_ptr.size() - 1;
  if (subtree_cut < 0 || subtree_cut >= std::max(1, nlev) || n_chains < 1 || n_chains > kMaxSplit) {
    g_err = "invalid solve plan";
    return 1;
  }
  size_t rows = 16;
  for (int l = 0; l < subtree_cut; ++l) rows = std::max<size_t>(rows, s->lev_rows[l]);
  if (rows * s->cfg_sub.Ct * 8 > kLdsMax) {
    g_err = "subtree fronts exceed the LDS budget";
    return 1;
  }
  HIPX(hipStreamSynchronize(s->stream));
  if (set_cut(s, subtree_cut)) return 1;
  s->nsplit = n_chains;
  return 0;
}

int pdm_set_dense_top(pdm_scene s, int max_rows) {
  if (!s || max_rows < 0) {
    g_err = "pdm_set_dense_top: invalid argument";
    return 1;
  }
  HIPX(hipStreamSynchronize(s->stream));
  s->top_cap = max_rows;
  if (plan_solve(s) || build_top_z(s) || build_leaf_ops(s)) return 1;
  HIPX(hipStreamSynchronize(s->stream));
  s->tuned_C = -1;  // launch shapes and subtree ranges are re-chosen on the next call
  s->have_tape = false;
  return 0;
}

int pdm_set_dense_leaves(pdm_scene s, int max_rows) {
  if (!s || max_rows < 0) {
    g_err = "pdm_set_dense_leaves: invalid argument";
    return 1;
  }
  HIPX(hipStreamSynchronize(s->stream));
  s->leaf_cap = max_rows;
  if (plan_solve(s) || build_top_z(s) || build_leaf_ops(s)) return 1;
  HIPX(hipStreamSynchronize(s->stream));
  s->tuned_C = -1;
  s->have_tape = false;
  return 0;
}

int pdm_get_solve_timing(pdm_scene s, long long* n_solves, double* total_ms) {
  if (!s || !n_solves || !total_ms) {
    g_err = "invalid argument";
    return 1;
  }
  if (fold_solve_events(s)) return 1;
  *n_solves = s->n_solves;
  *total_ms = s->solve_ms;
  s->n_solves = 0;
  s->solve_ms = 0.0;
  return 0;
}

int pdm_get_graph_launches(pdm_scene s, long long* n) {
  if (!s || !n) {
    g_err = "invalid argument";
    return 1;
  }
  *n = s->graph_launches;
  return 0;
}

int pdm_get_timing(pdm_scene s, double* fwd_ms, double* bwd_ms, double* factor_ms) {
  if (!s) {
    g_err = "null scene";
    return 1;
  }
  if (fwd_ms) *fwd_ms = s->t_fwd;
  if (bwd_ms) *bwd_ms = s->t_bwd;
  if (factor_ms) *factor_ms = s->t_factor;
  return 0;
}

}  // extern "C"
