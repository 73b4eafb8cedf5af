// Plane-fused kernels: one work-group holds a whole x-y plane of one box in
// LDS and applies both in-plane axes there.
//
//   k_plane_fwd  plane j3 of W (wavelet x,y at wavelet z index j3):
//                K  = K_x W + K_y W              (kinetic, + ekin partial)
//                T2 = G_y G_x W                  (real x,y at wavelet z)
//   k_plane_bwd  plane j3 of T2 after the z pass (G_z^T V G_z):
//                out = G_x^T G_y^T T2 + K        -> hpsi += compress(out)
//
// Together with the z-line kernels (k_zpot, k_kin<2>) this replaces the
// x/y line passes and the compress pass: the (L1,2N2,2N3) intermediate and
// the separate wavelet output volume never reach memory.  Operators and
// filters are those of the line kernels (growshrink.f90 comb_grow/comb_shrink,
// convolut_ib_optim.f90 ConvolkineticT).  Chunks whose stencil lies inside
// the line read it directly; the forward pass's other chunks read through a
// per-box line map (periodic wrap, or a zero slot for free padding), the
// backward pass's apply the wrap or padding by index arithmetic.
#include "bdft_internal.h"

namespace bdft {

// threads per plane work-group: a large FP64 plane fills most of a CU's
// LDS, so one work-group runs per CU and it takes 12 waves (smaller boxes'
// forward planes, launched in their own group, two per CU); FP32 planes run
// two to four work-groups of 8 waves per CU
template <typename T>
struct PlaneNT {
  static constexpr int v = sizeof(T) == 8 ? 768 : 512;
};

template <typename T>
struct PLine {
  const T* p;  // element 0 of the line
  int es;      // element stride
};

// Work item it = tid + k step (k = 0, 1, ...) of a loop over ch * n + j
// (0 <= j < n): one division per thread, then a carry per step.
struct ItDiv {
  int ch, j, dch, dj;
  __device__ __forceinline__ ItDiv(int tid, int n, int step) {
    ch = tid / n;
    j = tid - ch * n;
    dch = step / n;
    dj = step - dch * n;
  }
  __device__ __forceinline__ void next(int n) {
    ch += dch;
    j += dj;
    if (j >= n) {
      j -= n;
      ++ch;
    }
  }
};

// LDS row stride for rows of n values: odd (n + 1 or n + 2), so that the
// lanes of a wave that each walk their own row hit distinct banks
__host__ __device__ constexpr int odd_stride(int n) { return (n + 1) | 1; }

// A contiguous n1 x n2 plane (x fastest) from memory into LDS rows of
// stride rs > n1: element i = tid + k NTP of the plane, its row carried from
// step to step (no division per element), UR loads in flight per thread.
// TRANS: stored transposed, x = j of row ch at j rs + ch (rs >= n2, odd).
template <typename T, int NTP, bool TRANS = false>
__device__ __forceinline__ void plane_load(T* dst, int rs, const T* __restrict__ src, int n1,
                                           int n2) {
  constexpr int UR = 8;
  const int tot = n1 * n2;
  ItDiv d(threadIdx.x, n1, NTP);
  for (int base = threadIdx.x; base < tot; base += NTP * UR) {
    T v[UR];
    int a[UR];
#pragma unroll
    for (int u = 0; u < UR; ++u) {
      const int i = base + u * NTP;
      v[u] = i < tot ? src[i] : (T)0;
      // row r, column i - r n1 at r rs + i - r n1 (transposed: (i - r n1) rs + r)
      a[u] = TRANS ? d.j * rs + d.ch : i + d.ch * (rs - n1);
      d.next(n1);
    }
#pragma unroll
    for (int u = 0; u < UR; ++u)
      if (base + u * NTP < tot) dst[a[u]] = v[u];
  }
}

// the reverse: LDS rows of stride rs into a contiguous n1 x n2 plane
template <typename T, int NTP>
__device__ __forceinline__ void plane_store(T* __restrict__ dst, const T* src, int rs, int n1,
                                            int n2) {
  const int tot = n1 * n2;
  ItDiv d(threadIdx.x, n1, NTP);
  for (int i = threadIdx.x; i < tot; i += NTP) {
    dst[i] = src[i + d.ch * (rs - n1)];
    d.next(n1);
  }
}

// G on a wavelet line [c(0..N-1); d(0..N-1)]: the taps of coarse input
// k = k0 + kk on the 8 fine outputs of a chunk (k0 = first coarse index)
template <typename T>
__device__ __forceinline__ void grow_tap(int kk, T c, T d, T o[8]) {
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    const int j = r + 16 - 2 * kk;
    if (j >= -14 && j <= 16) {
      o[r] = fma((T)Fil::F2S[j + 14], c, o[r]);
      o[r] = fma((T)Fil::F2W[j + 14], d, o[r]);
    }
  }
}

// wavelet-basis kinetic (unscaled): the taps of coarse input k = i0-14+kk
// (kk < 32) on the 4 coarse outputs i0..i0+3 (c and d)
template <typename T>
__device__ __forceinline__ void kin_tap(int kk, T c, T d, T oc[4], T od[4]) {
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int l = kk - 14 - r;
    if (l >= -14 && l <= 14) {
      const int al = l < 0 ? -l : l;
      oc[r] = fma((T)Fil::KA[al], c, oc[r]);
      oc[r] = fma((T)Fil::KC[-l + 14], d, oc[r]);
      od[r] = fma((T)Fil::KC[l + 14], c, od[r]);
      od[r] = fma((T)Fil::KE[al], d, od[r]);
    }
  }
}

// Line map of one axis of a box's plane in LDS: entry k + MOFF, for coarse
// index k = -MOFF .. N + 17, holds the byte offsets from the line's element
// 0 of c(k) and d(k): k taken modulo N on a periodic line; on a free line a
// k outside 0..N-1 gives, for both, the offset of a slot that holds zero.
// The stencil chunks that reach past a line end read through it, one map
// entry per tap in place of wrap or bounds arithmetic.  Every tap of the
// grow (k0 >= -15, k0 + 18 <= N + 17 for L <= 2N + 29) and kinetic chunks
// (i0 - 14 >= -14, i0 + 17 <= N + 16) lies in that range.
constexpr int MOFF = 15;
__host__ __device__ constexpr int map_len(int N) { return N + MOFF + 18; }

// es: byte stride of the line's elements; zero: element index of the zero slot
__device__ __forceinline__ void build_map(int2* m, int N, int per, int es, int zero, int tid,
                                          int nt) {
  for (int e = tid; e < map_len(N); e += nt) {
    const int k = e - MOFF;
    int c, d;
    if (per) {
      int kw = k % N;
      kw += kw < 0 ? N : 0;
      c = kw;
      d = N + kw;
    } else if ((unsigned)k < (unsigned)N) {
      c = k;
      d = N + k;
    } else {
      c = d = zero;
    }
    m[e] = make_int2(c * es, d * es);
  }
}

// G on a wavelet line: 8 fine outputs from k = k0 .. k0+18, read directly
// where the chunk lies inside the line, else through the line map m
template <typename T>
__device__ __forceinline__ void grow8_any(PLine<T> ln, const int2* m, int N, int k0, T o[8]) {
#pragma unroll
  for (int r = 0; r < 8; ++r) o[r] = (T)0;
  if (k0 >= 0 && k0 + 18 < N) {
#pragma unroll
    for (int kk = 0; kk < 19; ++kk)
      grow_tap(kk, ln.p[(k0 + kk) * ln.es], ln.p[(N + k0 + kk) * ln.es], o);
  } else {
    const int2* mm = m + (k0 + MOFF);
    const char* p = (const char*)ln.p;
#pragma unroll
    for (int kk = 0; kk < 19; ++kk) {
      const int2 e = mm[kk];
      grow_tap(kk, *(const T*)(p + e.x), *(const T*)(p + e.y), o);
    }
  }
}

// wavelet-basis kinetic (unscaled) on a wavelet line: 4 coarse outputs
// i0..i0+3 (c and d) from k = i0-14 .. i0+17, direct or through the map
template <typename T>
__device__ __forceinline__ void kin4_any(PLine<T> ln, const int2* m, int N, int i0, T oc[4],
                                         T od[4]) {
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    oc[r] = (T)0;
    od[r] = (T)0;
  }
  if (i0 - 14 >= 0 && i0 + 17 < N) {
#pragma unroll
    for (int kk = 0; kk < 32; ++kk) {
      const int k = i0 - 14 + kk;
      kin_tap(kk, ln.p[k * ln.es], ln.p[(N + k) * ln.es], oc, od);
    }
  } else {
    const int2* mm = m + (i0 - 14 + MOFF);
    const char* p = (const char*)ln.p;
#pragma unroll
    for (int kk = 0; kk < 32; ++kk) {
      const int2 e = mm[kk];
      kin_tap(kk, *(const T*)(p + e.x), *(const T*)(p + e.y), oc, od);
    }
  }
}

// G^T on a real line of L points (periodic: fine index p, L = 2N; free: fine
// index p-14): the taps of fine input I = 2i0-14+qq (qq < 2 NO + 29) on the
// NO coarse outputs i0..i0+NO-1 (c and d)
template <typename T, int NO>
__device__ __forceinline__ void shrink_tap(int qq, T x, T oc[NO], T od[NO]) {
#pragma unroll
  for (int r = 0; r < NO; ++r) {
    const int j = qq - 2 * r - 14;
    if (j >= -14 && j <= 16) {
      oc[r] = fma((T)Fil::F2S[j + 14], x, oc[r]);
      od[r] = fma((T)Fil::F2W[j + 14], x, od[r]);
    }
  }
}

// Fine-line map of a periodic line for the backward pass: entry e = 2 i0 + qq
// of a shrink chunk holds the byte offset of its input, fine index e - 14
// modulo L.  Every chunk's entries lie below 2 N + 2 NO + 27 <= fmap_len(N).
__host__ __device__ constexpr int fmap_len(int N) { return 2 * N + 40; }
__device__ __forceinline__ void build_fmap(int* m, int N, int L, int es, int tid, int nt) {
  for (int e = tid; e < fmap_len(N); e += nt) {
    int p = (e - 14) % L;
    p += p < 0 ? L : 0;
    m[e] = p * es;
  }
}

// G^T on a fine line stored as an LDS row: NO coarse outputs, read directly
// where the chunk's inputs lie inside the line; else, periodic, through the
// map m, free (only the line's last chunk), with fine points past L as zero
template <typename T, int NO = 4>
__device__ __forceinline__ void shrink4_any(const T* ln, const int* m, int L, int i0, int per,
                                            T oc[NO], T od[NO]) {
#pragma unroll
  for (int r = 0; r < NO; ++r) {
    oc[r] = (T)0;
    od[r] = (T)0;
  }
  const bool in = per ? (2 * i0 - 14 >= 0 && 2 * i0 + 2 * NO + 14 < L)
                      : (2 * i0 + 2 * NO + 28 < L);
  if (in) {
    const T* q = ln + 2 * i0 - (per ? 14 : 0);
#pragma unroll
    for (int qq = 0; qq < 2 * NO + 29; ++qq) shrink_tap<T, NO>(qq, q[qq], oc, od);
  } else if (per) {
    const int* mm = m + 2 * i0;
    const char* p = (const char*)ln;
#pragma unroll
    for (int qq = 0; qq < 2 * NO + 29; ++qq)
      shrink_tap<T, NO>(qq, *(const T*)(p + mm[qq]), oc, od);
  } else {
#pragma unroll
    for (int qq = 0; qq < 2 * NO + 29; ++qq) {
      const int p = 2 * i0 + qq;
      const bool v = p < L;
      const T x = ln[v ? p : 0];
      shrink_tap<T, NO>(qq, v ? x : (T)0, oc, od);
    }
  }
}

template <int NW>
__device__ __forceinline__ double block_sum_p(double v, double* red) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  const int w = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < NW; ++i) s += red[i];
  return s;
}

// first coarse index of the grow chunk starting at fine output p0
__device__ __forceinline__ int grow_k0(int p0, int per) {
  return per ? (p0 >> 1) - 8 : (p0 >> 1) - 15;
}

// LDS layout of the forward and preconditioner plane kernels: the W plane A
// (W2 rows of stride RA = W1 + 1, then a zero row; column W1 of every row is
// zero: the free lines' zero slots), the second plane C (W2 + 1 rows of
// stride RC), then the line maps (x on the rows of A, y on the columns of A,
// y on the columns of C).  Byte offset of the maps:
__host__ __device__ inline size_t plane_maps_off(int N1, int N2, int RC, int elem) {
  const size_t rows = 2 * N2 + 1, RA = 2 * N1 + 1;
  return ((rows * RA + rows * RC) * elem + 7) & ~(size_t)7;
}

// zero slots: column W1 and row W2 of A, row W2 of C (stride RC)
template <typename T, int NTP>
__device__ __forceinline__ void plane_zero_slots(T* A, int RA, int W1, int W2, T* C, int RC) {
  for (int i = threadIdx.x; i < W2; i += NTP) A[i * RA + W1] = (T)0;
  for (int i = threadIdx.x; i < RA; i += NTP) A[W2 * RA + i] = (T)0;
  for (int i = threadIdx.x; i < RC; i += NTP) C[W2 * RC + i] = (T)0;
}

// ---------------------------------------------------------------------------
// forward plane kernel; KIN = false: T2 only (the density's x and y grow,
// daub_to_isf in sumrho.f90:377-469), no kinetic, no K, no ekin.  tt: the
// launch's tiles (all boxes, or one group of them); ft: the table over all
// boxes whose tile numbers index the kinetic partials
// ---------------------------------------------------------------------------
template <typename T, bool KIN = true>
__global__ __launch_bounds__(PlaneNT<T>::v) void k_plane_fwd(Batch B, TileTable tt,
                                                             TileTable ft, double* part) {
  constexpr int NTP = PlaneNT<T>::v;
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ double red[NTP / 64];
  const int tile = tile_of(tt);
  const int lb = find_box(tt, tile);
  const int b = table_box(tt, lb);
  const BoxDesc bd = B.box[b];
  const int j3 = tile - box_first_tile(tt, lb);
  const int N1 = bd.N[0], N2 = bd.N[1], L1 = bd.L[0], L2 = bd.L[1], per = bd.per;
  const int W1 = 2 * N1, W2 = 2 * N2;
  const int RA = W1 + 1;            // row stride of the wavelet plane
  const int RB = odd_stride(L1);    // row stride of the G_x output
  const int RC = max(RA, RB);
  T* A = (T*)smem;                  // W plane: A[j2*RA + j1]
  T* C = A + (W2 + 1) * RA;         // kinetic accumulator, later G_x output
  int2* mx = (int2*)(smem + plane_maps_off(N1, N2, RC, sizeof(T)));
  int2* mya = mx + map_len(N1);
  int2* myc = mya + map_len(N2);
  const T* src = (const T*)B.W + bd.w_off + (long long)W1 * W2 * j3;
  const int tid = threadIdx.x;
  // load the plane (contiguous), UNR loads in flight per thread
  plane_load<T, NTP>(A, RA, src, W1, W2);
  build_map(mx, N1, per, sizeof(T), W1, tid, NTP);
  build_map(mya, N2, per, RA * sizeof(T), W2, tid, NTP);
  build_map(myc, N2, per, RB * sizeof(T), W2, tid, NTP);
  plane_zero_slots<T, NTP>(A, RA, W1, W2, C, RB);
  __syncthreads();
  if constexpr (KIN) {
  // K_x: rows j2, chunks of 4 coarse x positions -> C
  const T kx = (T)bd.kscale[0], ky = (T)bd.kscale[1];
  const int cx = (N1 + 3) >> 2;
  for (ItDiv d(tid, W2, NTP); d.ch < cx; d.next(W2)) {
    const int ch = d.ch, j2 = d.j;
    const int i0 = 4 * ch;
    T oc[4], od[4];
    kin4_any(PLine<T>{A + j2 * RA, 1}, mx, N1, i0, oc, od);
#pragma unroll
    for (int r = 0; r < 4; ++r)
      if (i0 + r < N1) {
        C[j2 * RA + i0 + r] = oc[r] * kx;
        C[j2 * RA + N1 + i0 + r] = od[r] * kx;
      }
  }
  __syncthreads();
  // K_y: columns j1, chunks of 4 coarse y positions, accumulated into C
  double ek = 0.0;
  const int cy = (N2 + 3) >> 2;
  for (ItDiv d(tid, W1, NTP); d.ch < cy; d.next(W1)) {
    const int ch = d.ch, j1 = d.j;
    const int i0 = 4 * ch;
    T oc[4], od[4];
    kin4_any(PLine<T>{A + j1, RA}, mya, N2, i0, oc, od);
#pragma unroll
    for (int r = 0; r < 4; ++r)
      if (i0 + r < N2) {
        const int a = (i0 + r) * RA + j1, d = (N2 + i0 + r) * RA + j1;
        const T vc = C[a] + oc[r] * ky, vd = C[d] + od[r] * ky;
        C[a] = vc;
        C[d] = vd;
        ek += (double)vc * (double)A[a] + (double)vd * (double)A[d];
      }
  }
  const double eks = block_sum_p<NTP / 64>(ek, red);
  if (tid == 0) part[box_first_tile(ft, b) + j3] = eks;
  __syncthreads();
  // K plane -> memory (contiguous)
  {
    T* kd = (T*)B.K + bd.w_off + (long long)W1 * W2 * j3;
    plane_store<T, NTP>(kd, C, RA, W1, W2);
  }
  __syncthreads();
  }
  // G_x: rows j2 -> C as (W2 rows) x (L1 cols), row stride RB
  const int gx = (L1 + 7) >> 3;
  for (ItDiv d(tid, W2, NTP); d.ch < gx; d.next(W2)) {
    const int ch = d.ch, j2 = d.j;
    const int p0 = 8 * ch;
    T o[8];
    grow8_any(PLine<T>{A + j2 * RA, 1}, mx, N1, grow_k0(p0, per), o);
#pragma unroll
    for (int r = 0; r < 8; ++r)
      if (p0 + r < L1) C[j2 * RB + p0 + r] = o[r];
  }
  __syncthreads();
  // G_y: columns p1 of C -> T2 plane (real x,y at wavelet z j3)
  T* dst = (T*)B.T2 + bd.t2_off + (long long)L1 * L2 * j3;
  const int gy = (L2 + 7) >> 3;
  for (ItDiv d(tid, L1, NTP); d.ch < gy; d.next(L1)) {
    const int ch = d.ch, p1 = d.j;
    const int p0 = 8 * ch;
    T o[8];
    grow8_any(PLine<T>{C + p1, RB}, myc, N2, grow_k0(p0, per), o);
#pragma unroll
    for (int r = 0; r < 8; ++r)
      if (p0 + r < L2) dst[p1 + (long long)L1 * (p0 + r)] = o[r];
  }
}

// ---------------------------------------------------------------------------
// Preconditioner operator, x and y part, of one x-y plane of a box:
//   K = c W + Ktilde_x W + Ktilde_y W      (c = cprecr of the box, pc[9 b])
// the work of the x kinetic line pass (k_pc_kinx) and the y one (k_kin<1>)
// in one read of the W plane and one write of K, in the same order of
// operations (c W + K_x first, then + K_y)
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(PlaneNT<T>::v) void k_plane_kin_pc(Batch B, TileTable tt,
                                                                const double* pc) {
  constexpr int NTP = PlaneNT<T>::v;
  extern __shared__ __align__(16) unsigned char smem[];
  const int tile = tile_of(tt);
  const int lb = find_box(tt, tile);
  const int b = table_box(tt, lb);
  const BoxDesc bd = B.box[b];
  const int j3 = tile - box_first_tile(tt, lb);
  const int N1 = bd.N[0], N2 = bd.N[1], per = bd.per;
  const int W1 = 2 * N1, W2 = 2 * N2;
  const int RA = W1 + 1;
  T* A = (T*)smem;                  // W plane: A[j2*RA + j1]
  T* C = A + (W2 + 1) * RA;         // K plane, same layout
  int2* mx = (int2*)(smem + plane_maps_off(N1, N2, RA, sizeof(T)));
  int2* mya = mx + map_len(N1);
  const T* src = (const T*)B.W + bd.w_off + (long long)W1 * W2 * j3;
  const int tid = threadIdx.x;
  plane_load<T, NTP>(A, RA, src, W1, W2);
  build_map(mx, N1, per, sizeof(T), W1, tid, NTP);
  build_map(mya, N2, per, RA * sizeof(T), W2, tid, NTP);
  plane_zero_slots<T, NTP>(A, RA, W1, W2, C, RA);
  __syncthreads();
  const T kx = (T)bd.kscale[0], ky = (T)bd.kscale[1], c = (T)pc[9 * b];
  const int cx = (N1 + 3) >> 2;
  for (ItDiv d(tid, W2, NTP); d.ch < cx; d.next(W2)) {
    const int ch = d.ch, j2 = d.j;
    const int i0 = 4 * ch;
    T oc[4], od[4];
    kin4_any(PLine<T>{A + j2 * RA, 1}, mx, N1, i0, oc, od);
#pragma unroll
    for (int r = 0; r < 4; ++r)
      if (i0 + r < N1) {
        const int a = j2 * RA + i0 + r, d = j2 * RA + N1 + i0 + r;
        C[a] = fma(c, A[a], oc[r] * kx);
        C[d] = fma(c, A[d], od[r] * kx);
      }
  }
  __syncthreads();
  const int cy = (N2 + 3) >> 2;
  for (ItDiv d(tid, W1, NTP); d.ch < cy; d.next(W1)) {
    const int ch = d.ch, j1 = d.j;
    const int i0 = 4 * ch;
    T oc[4], od[4];
    kin4_any(PLine<T>{A + j1, RA}, mya, N2, i0, oc, od);
#pragma unroll
    for (int r = 0; r < 4; ++r)
      if (i0 + r < N2) {
        const int a = (i0 + r) * RA + j1, d = (N2 + i0 + r) * RA + j1;
        C[a] += oc[r] * ky;
        C[d] += od[r] * ky;
      }
  }
  __syncthreads();
  T* kd = (T*)B.K + bd.w_off + (long long)W1 * W2 * j3;
  plane_store<T, NTP>(kd, C, RA, W1, W2);
}

size_t plane_lds_kin(int N1, int N2, int elem) {
  return plane_maps_off(N1, N2, 2 * N1 + 1, elem) + 8 * (size_t)(map_len(N1) + map_len(N2));
}

hipError_t launch_plane_kin_pc(const Batch& B, const TileTable& t, int ntiles, size_t lds,
                               const double* pc, hipStream_t s) {
  if (ntiles > 0) k_plane_kin_pc<double><<<ntiles, PlaneNT<double>::v, lds, s>>>(B, t, pc);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// backward plane kernel with compress-accumulate
// ---------------------------------------------------------------------------
// OCC2: bounded to six waves per SIMD (two 768-thread work-groups per CU),
// for the FP64 boxes whose planes fit half a CU's LDS
template <typename T, bool OCC2 = false>
__global__ __launch_bounds__(PlaneNT<T>::v, OCC2 ? 6 : 1) void k_plane_bwd(Batch B,
                                                                         TileTable tt) {
  constexpr int NTP = PlaneNT<T>::v;
  extern __shared__ __align__(16) unsigned char smem[];
  const int tile = tile_of(tt);
  const int lb = find_box(tt, tile);   // tables over a subset carry a box map
  const BoxDesc bd = B.box[table_box(tt, lb)];
  const int j3 = tile - box_first_tile(tt, lb);
  const int N1 = bd.N[0], N2 = bd.N[1], N3 = bd.N[2], L1 = bd.L[0], L2 = bd.L[1],
            per = bd.per;
  const int W1 = 2 * N1, W2 = 2 * N2;
  // odd row strides: a free box's fine extents are odd already, so its rows
  // take no padding (with the periodic maps, the plane takes no more LDS
  // than with untransposed rows of L1 + 1)
  const int RU = L2 | 1;            // T2 plane, transposed: y lines as rows
  const int RQ = L1 | 1;            // G_y^T output rows (read by rows)
  const int RO = W1 + 1;
  T* U = (T*)smem;                // T2 plane: U[p1*RU + p2], later output O
  T* Bq = U + max(L1 * RU, W2 * RO);  // G_y^T output: Bq[j2*RQ + p1]
  int* fmx = (int*)(Bq + W2 * RQ);     // periodic fine-line maps (x: rows of Bq,
  int* fmy = fmx + fmap_len(N1);       // y: rows of U)
  const int tid = threadIdx.x;
  const T* src = (const T*)B.T2 + bd.t2_off + (long long)L1 * L2 * j3;
  // the compress's first point per thread (indices, K, old hpsi) is loaded
  // before the plane, so that its latency overlaps the plane load and the
  // two passes (one point: more registers held through the passes cost
  // occupancy); later steps take UC points each
  constexpr int UC = NTP >= 768 ? 3 : 2, UP = 1;
  const T* Kp = (const T*)B.K + bd.w_off + (long long)W1 * W2 * j3;
  CompressStep<T, UP> st0;
  compress_load<UP, NTP>(bd, j3, Kp, B.hpsi, tid, st0);
  plane_load<T, NTP, true>(U, RU, src, L1, L2);
  if (per) {
    build_fmap(fmx, N1, L1, sizeof(T), tid, NTP);
    build_fmap(fmy, N2, L2, sizeof(T), tid, NTP);
  }
  __syncthreads();
  // G_y^T: y lines p1 (rows of U) -> Bq rows (c: i2, d: N2+i2), NOY coarse outputs
  // per work item (NOX for G_x^T): two in FP32, so that the last round of
  // items leaves fewer lanes idle between the barriers (C4 FP32 backward
  // pass 3.41 -> 3.22 ms); four in FP64 (two was slower on the periodic
  // 48^3 planes, 2.27 -> 2.47 ms)
  constexpr int NOY = sizeof(T) == 4 ? 2 : 4, NOX = NOY;
  const int cy = (N2 + NOY - 1) / NOY;
  for (ItDiv d(tid, L1, NTP); d.ch < cy; d.next(L1)) {
    const int ch = d.ch, p1 = d.j;
    const int i0 = NOY * ch;
    T oc[NOY], od[NOY];
    shrink4_any<T, NOY>(U + p1 * RU, fmy, L2, i0, per, oc, od);
#pragma unroll
    for (int r = 0; r < NOY; ++r)
      if (i0 + r < N2) {
        Bq[(i0 + r) * RQ + p1] = oc[r];
        Bq[(N2 + i0 + r) * RQ + p1] = od[r];
      }
  }
  __syncthreads();
  // G_x^T: rows j2 of Bq -> output plane O (reuses U), row stride RO
  T* O = U;
  const int cx = (N1 + NOX - 1) / NOX;
  for (ItDiv d(tid, W2, NTP); d.ch < cx; d.next(W2)) {
    const int ch = d.ch, j2 = d.j;
    const int i0 = NOX * ch;
    T oc[NOX], od[NOX];
    shrink4_any<T, NOX>(Bq + j2 * RQ, fmx, L1, i0, per, oc, od);
#pragma unroll
    for (int r = 0; r < NOX; ++r)
      if (i0 + r < N1) {
        O[j2 * RO + i0 + r] = oc[r];
        O[j2 * RO + N1 + i0 + r] = od[r];
      }
  }
  __syncthreads();
  // out = O + K, accumulated into hpsi at the compressed positions of the
  // plane's four components (s3 fixed by j3), compress_and_accumulate_mixed
  // (un-compress.f90:2174-2238)
  compress_store<T, UP, NTP>(bd, j3, O, RO, B.hpsi, tid, st0);
  for (int i0 = tid + NTP * UP; i0 < N1 * N2; i0 += NTP * UC) {
    CompressStep<T, UC> st;
    compress_load<UC, NTP>(bd, j3, Kp, B.hpsi, i0, st);
    compress_store<T, UC, NTP>(bd, j3, O, RO, B.hpsi, i0, st);
  }
}

// LDS bytes of the two plane kernels for the largest box extents
size_t plane_lds_fwd(int N1, int N2, int L1, int elem) {
  const int RA = 2 * N1 + 1, RB = odd_stride(L1);
  return plane_maps_off(N1, N2, RA > RB ? RA : RB, elem) +
         8 * (size_t)(map_len(N1) + 2 * map_len(N2));
}
size_t plane_lds_bwd(int N1, int N2, int L1, int L2, int elem) {
  const size_t RU = L2 | 1, RQ = L1 | 1, RO = 2 * N1 + 1;
  const size_t u = (size_t)L1 * RU > 2 * N2 * RO ? (size_t)L1 * RU : 2 * N2 * RO;
  const bool per = L1 == 2 * N1;
  return (u + 2 * N2 * RQ) * elem + (per ? 4 * (size_t)(fmap_len(N1) + fmap_len(N2)) : 0);
}

template <typename T>
hipError_t plane_setup(int bytes) {
  hipError_t e = hipFuncSetAttribute((const void*)k_plane_fwd<T>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  if (e == hipSuccess)
    e = hipFuncSetAttribute((const void*)k_plane_fwd<T, false>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  if (e == hipSuccess)
    e = hipFuncSetAttribute((const void*)k_plane_kin_pc<T>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  if (e == hipSuccess)
    e = hipFuncSetAttribute((const void*)k_plane_bwd<T>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  if (e == hipSuccess)
    e = hipFuncSetAttribute((const void*)k_plane_bwd<T, true>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  return e;
}

// LDS bytes one work-group may use on the current device; the plane kernels'
// dynamic-LDS limit is raised to it where the runtime asks for that.
size_t plane_attributes() {
  int dev = 0, v = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) !=
          hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  if (plane_setup<double>(v) != hipSuccess || plane_setup<float>(v) != hipSuccess)
    (void)hipGetLastError();  // not required on this runtime; clear the error
  return (size_t)v;
}

template <typename T>
hipError_t launch_plane_fwd(const Batch& B, const TileTable& t, const TileTable& ft,
                            int ntiles, size_t lds, double* part, hipStream_t s) {
  k_plane_fwd<T><<<ntiles, PlaneNT<T>::v, lds, s>>>(B, t, ft, part);
  return hipGetLastError();
}
template <typename T>
hipError_t launch_plane_grow(const Batch& B, const TileTable& t, int ntiles, size_t lds,
                             hipStream_t s) {
  k_plane_fwd<T, false><<<ntiles, PlaneNT<T>::v, lds, s>>>(B, t, t, nullptr);
  return hipGetLastError();
}
template <typename T>
hipError_t launch_plane_bwd(const Batch& B, const TileTable& t, int ntiles, size_t lds,
                            hipStream_t s, bool occ2) {
  if (occ2) k_plane_bwd<T, true><<<ntiles, PlaneNT<T>::v, lds, s>>>(B, t);
  else k_plane_bwd<T><<<ntiles, PlaneNT<T>::v, lds, s>>>(B, t);
  return hipGetLastError();
}
template hipError_t launch_plane_fwd<double>(const Batch&, const TileTable&, const TileTable&,
                                             int, size_t, double*, hipStream_t);
template hipError_t launch_plane_fwd<float>(const Batch&, const TileTable&, const TileTable&,
                                            int, size_t, double*, hipStream_t);
template hipError_t launch_plane_grow<double>(const Batch&, const TileTable&, int, size_t,
                                              hipStream_t);
template hipError_t launch_plane_grow<float>(const Batch&, const TileTable&, int, size_t,
                                             hipStream_t);
template hipError_t launch_plane_bwd<double>(const Batch&, const TileTable&, int, size_t,
                                             hipStream_t, bool);
template hipError_t launch_plane_bwd<float>(const Batch&, const TileTable&, int, size_t,
                                            hipStream_t, bool);

}  // namespace bdft
