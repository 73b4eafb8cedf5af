// Pass kernels of the local-Hamiltonian apply, templated on the working type
// T of the convolutions (double: plain FP64; float: mixed precision with FP64
// input/potential/energies/accumulation).
//
// hpsi += G^T V G psi + K psi, evaluated as separable line passes over
// bundles of 32 lines staged in LDS:
//   G   = (synthesis o magic filter) per axis, the combined filter fil2 of
//         bigdft/src/orbitals/growshrink.f90 (comb_grow_all), periodic wrap
//         or zero padding;
//   G^T = (magic^T o analysis) per axis (comb_shrink,
//         growshrink.f90:310-340);
//   K   = kinetic operator in the wavelet basis, 2x2 block of 29-tap filters
//         per axis (ConvolkineticT, convolut_ib_optim.f90:921-1740).
// For periodic boxes the reference applies the fine-scaling kinetic between
// synthesis and analysis (convolut_kinetic_per_T, convolut_optim_per.f90:
// 1169-1260); since analysis o synthesis = identity per periodic axis, that
// equals the wavelet-basis operator used here.
//
// Pass order for one box (x fastest in every volume):
//   uncompress   psi -> W                      (wavelet volume)
//   xfwd         W  x-lines: T1 = G_x W, K = K_x W        (+ekin)
//   yfwd         T1 y-lines: T2 = G_y T1
//   kin(y), kin(z) W lines: K += K_d W                    (+ekin)
//   zpot         T2 z-lines: T2 = G_z^T V G_z T2 (in place) (+epot)
//   ybwd         T2 y-lines: T1 = G_y^T T2
//   xbwd         T1 x-lines: W = G_x^T T1 + K
//   compress     hpsi += compress(W)
#include "bdft_internal.h"

namespace bdft {

// i mod n for i in a small window around [0,n): conditional steps, no division
__device__ __forceinline__ int wrap(int i, int n) {
  while (i < 0) i += n;
  while (i >= n) i -= n;
  return i;
}

// a / b for 0 <= a < 2^22 with a float reciprocal, corrected to exact

// ---------------------------------------------------------------------------
// LDS staging.  Layout: lds[row * LDW + line].
// ---------------------------------------------------------------------------

// Wavelet line bundle: two sections (c, d) of NS = N + HL + HR rows; row r of
// a section holds coarse index k = r - HL (wrapped or zero outside [0,N)).
// CONTIG: lines are contiguous in memory (x pass), else positions are.
// Loads are issued in batches of UNR per thread before their LDS stores, so
// that many global loads are in flight at once.
// SCALE (positions along lines only): every value is multiplied by mul, a
// factor of this thread's line (a thread always loads the same line l =
// threadIdx.x mod TW there).
constexpr int UNR = 8;

template <bool CONTIG, typename T, typename S, bool SCALE = false>
__device__ __forceinline__ void fill_wav(T* lds, const S* __restrict__ src,
                                         long long lstride, long long pstride,
                                         int nl, int N, int per, T mul = (T)1) {
  const int NS = N + HL + HR;
  const int rows = 2 * NS;
  if (CONTIG) {
    // lines contiguous (pstride 1, lstride = 2N): read the tile as one flat
    // run of TW*2N elements (lines >= nl read as zero), then fill the halo
    // rows from LDS
    const int W = 2 * N;
    const int tot = TW * W;
    const float inv = 1.0f / (float)W;
    for (int base = threadIdx.x; base < tot; base += NT * UNR) {
      T v[UNR];
#pragma unroll
      for (int u = 0; u < UNR; ++u) {
        const int idx = base + u * NT;
        v[u] = (idx < nl * W) ? (T)src[idx] : (T)0;
      }
#pragma unroll
      for (int u = 0; u < UNR; ++u) {
        const int idx = base + u * NT;
        if (idx < tot) {
          const int l = qdiv(idx, W, inv);
          const int pos = idx - l * W;
          const int sec = pos >= N;
          lds[(sec * NS + (pos - sec * N) + HL) * LDW + l] = v[u];
        }
      }
    }
    (void)lstride;
    __syncthreads();
    const int H = HL + HR;
    for (int idx = threadIdx.x; idx < 2 * H * TW; idx += NT) {
      const int hr = idx >> 5, l = idx & (TW - 1);
      const int sec = hr >= H;
      const int r = hr - sec * H;
      const int k = r < HL ? r - HL : N + (r - HL);  // halo coarse index
      T v = (T)0;
      if (per) v = lds[(sec * NS + wrap(k, N) + HL) * LDW + l];
      lds[(sec * NS + k + HL) * LDW + l] = v;
    }
  } else {
    const int tot = rows * TW;
    for (int base = threadIdx.x; base < tot; base += NT * UNR) {
      T v[UNR];
#pragma unroll
      for (int u = 0; u < UNR; ++u) {
        const int idx = base + u * NT;
        const int row = idx >> 5, l = idx & (TW - 1);
        const int sec = row >= NS;
        int k = row - sec * NS - HL;
        bool ok = l < nl && idx < tot;
        if (per) k = wrap(k, N);
        else ok = ok && k >= 0 && k < N;
        v[u] = ok ? (T)src[(long long)l * lstride + (long long)(sec * N + k) * pstride]
                  : (T)0;
        if (SCALE) v[u] *= mul;
      }
#pragma unroll
      for (int u = 0; u < UNR; ++u) {
        const int idx = base + u * NT;
        if (idx < tot) lds[(idx >> 5) * LDW + (idx & (TW - 1))] = v[u];
      }
    }
  }
}

// Real-space line bundle for G^T: rows q = I + 14 with fine index
// I in [-14, 2N+22); periodic wraps mod 2N, free reads position p = q < L.
template <bool CONTIG, typename T>
__device__ __forceinline__ void fill_real(T* lds, const T* __restrict__ src,
                                          long long lstride, long long pstride,
                                          int nl, int N, int L, int per) {
  const int Q = 2 * N + 36;
  if (CONTIG) {
    const int tot = TW * L;
    const float inv = 1.0f / (float)L;
    for (int base = threadIdx.x; base < tot; base += NT * UNR) {
      T v[UNR];
#pragma unroll
      for (int u = 0; u < UNR; ++u) {
        const int idx = base + u * NT;
        v[u] = (idx < nl * L) ? src[idx] : (T)0;
      }
#pragma unroll
      for (int u = 0; u < UNR; ++u) {
        const int idx = base + u * NT;
        if (idx < tot) {
          const int l = qdiv(idx, L, inv);
          const int p = idx - l * L;
          lds[(per ? p + 14 : p) * LDW + l] = v[u];
        }
      }
    }
    (void)lstride;
    __syncthreads();
    // rows not covered above: periodic halos [0,14) and [2N+14,Q), free tail
    const int nh = per ? Q - 2 * N : Q - L;
    for (int idx = threadIdx.x; idx < nh * TW; idx += NT) {
      const int hr = idx >> 5, l = idx & (TW - 1);
      int q;
      T v = (T)0;
      if (per) {
        q = hr < 14 ? hr : 2 * N + hr;
        v = lds[(wrap(q - 14, 2 * N) + 14) * LDW + l];
      } else {
        q = L + hr;
      }
      lds[q * LDW + l] = v;
    }
  } else {
    const int tot = Q * TW;
    for (int base = threadIdx.x; base < tot; base += NT * UNR) {
      T v[UNR];
#pragma unroll
      for (int u = 0; u < UNR; ++u) {
        const int idx = base + u * NT;
        const int q = idx >> 5, l = idx & (TW - 1);
        int pos;
        bool ok = l < nl && idx < tot;
        if (per) pos = wrap(q - 14, 2 * N);
        else {
          pos = q;
          ok = ok && q < L;
        }
        v[u] = ok ? src[(long long)l * lstride + (long long)pos * pstride] : (T)0;
      }
#pragma unroll
      for (int u = 0; u < UNR; ++u) {
        const int idx = base + u * NT;
        if (idx < tot) lds[(idx >> 5) * LDW + (idx & (TW - 1))] = v[u];
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Register-blocked 1D operators on one LDS column.
// ---------------------------------------------------------------------------

// G: outputs at fine positions I0..I0+7 (I0 even) from coarse k = I0/2-8+kk,
// kk = 0..18: out(I) += fil2(I-2k,1) c(k) + fil2(I-2k,2) d(k).
// cc points at the c-section row of k = I0/2-8, dd at the d-section row.
template <typename T>
__device__ __forceinline__ void grow8(const T* __restrict__ cc,
                                      const T* __restrict__ dd, T o[8]) {
#pragma unroll
  for (int r = 0; r < 8; ++r) o[r] = (T)0;
#pragma unroll
  for (int kk = 0; kk < 19; ++kk) {
    const T c = cc[kk * LDW];
    const T d = dd[kk * LDW];
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      const int j = r + 16 - 2 * kk;  // fil2 index, valid in [-14, 16]
      if (j >= -14 && j <= 16) {
        o[r] = fma((T)Fil::F2S[j + 14], c, o[r]);
        o[r] = fma((T)Fil::F2W[j + 14], d, o[r]);
      }
    }
  }
}

// G^T: coarse outputs i0..i0+3 (c and d) from fine rows q = 2*i0 + qq,
// qq = 0..36 (row q holds fine index q - 14): c(i) = sum_j fil2(j,1) x(2i+j).
template <typename T>
__device__ __forceinline__ void shrink4(const T* __restrict__ xq, T oc[4], T od[4]) {
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    oc[r] = (T)0;
    od[r] = (T)0;
  }
#pragma unroll
  for (int qq = 0; qq < 37; ++qq) {
    const T x = xq[qq * LDW];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int j = qq - 2 * r - 14;
      if (j >= -14 && j <= 16) {
        oc[r] = fma((T)Fil::F2S[j + 14], x, oc[r]);
        od[r] = fma((T)Fil::F2W[j + 14], x, od[r]);
      }
    }
  }
}

// Wavelet-basis kinetic (unscaled): coarse outputs i0..i0+3 from rows of
// k = i0-14+kk, kk = 0..31 (cc, dd point at k = i0-14 in each section).
//   c_out(i) += a(l) c(i+l) + c(-l) d(i+l);  d_out(i) += c(l) c(i+l) + e(l) d(i+l)
template <typename T>
__device__ __forceinline__ void kin4(const T* __restrict__ cc, const T* __restrict__ dd,
                                     T oc[4], T od[4]) {
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    oc[r] = (T)0;
    od[r] = (T)0;
  }
#pragma unroll
  for (int kk = 0; kk < 32; ++kk) {
    const T c = cc[kk * LDW];
    const T d = dd[kk * LDW];
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
}

// Block reduction of one double per thread (work-group of NTB threads, red
// holds NTB / 64 entries); result valid in thread 0.
template <int NTB = NT>
__device__ __forceinline__ double block_sum(double v, double* red) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  const int w = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < NTB / 64; ++i) s += red[i];
  return s;
}

// grow-pass row of the first coarse k used by output chunk starting at p0
__device__ __forceinline__ int grow_row0(int p0, int per) {
  // periodic: I0 = p0, k0 = p0/2 - 8, row = k0 + HL;
  // free:     I0 = p0 - 14, k0 = p0/2 - 15, row = k0 + HL
  return per ? (p0 >> 1) + 7 : (p0 >> 1);
}

// ---------------------------------------------------------------------------
// uncompress / compress (keys pre-decoded into dense coarse linear indices)
// ---------------------------------------------------------------------------
constexpr int CPT = 256;  // compressed points per tile

// psig index of coarse point lin = i1 + N1*(i2 + N2*i3), component bits s
__device__ __forceinline__ long long psig_index(const BoxDesc& bd, int lin, int s1,
                                                int s2, int s3) {
  const int N1 = bd.N[0], N2 = bd.N[1], N3 = bd.N[2];
  const int t = qdiv(lin, N1, 1.0f / (float)N1);
  const int i1 = lin - t * N1;
  const int i3 = qdiv(t, N2, 1.0f / (float)N2);
  const int i2 = t - i3 * N2;
  return (long long)(i1 + N1 * s1) +
         2LL * N1 * ((long long)(i2 + N2 * s2) + 2LL * N2 * (i3 + N3 * s3));
}

template <typename T>
__global__ __launch_bounds__(NT) void k_zero(T* __restrict__ p, long long n) {
  for (long long i = blockIdx.x * (long long)NT + threadIdx.x; i < n;
       i += (long long)gridDim.x * NT)
    p[i] = (T)0;
}

// W offset of component t = 1..7 of fine point lin: (s1,s2,s3) = bits of t
// (un-compress.f90:930-936), each bit a fixed step from the (0,0,0) value
__device__ __forceinline__ long long fine_step(const BoxDesc& bd, int t) {
  const long long N1 = bd.N[0], N2 = bd.N[1], N3 = bd.N[2];
  return (t & 1) * N1 + ((t >> 1) & 1) * 2 * N1 * N2 + ((t >> 2) & 1) * 4 * N1 * N2 * N3;
}

// A fine tile's records (7 consecutive coefficients per point, CPT points)
// move between memory and LDS in one coalesced sweep (7 accesses per thread
// in flight) instead of a 56-byte stride per lane; each thread then works
// on its own record in LDS (stride 7: no bank conflicts).  FP64 uncompress
// 1.55 -> 1.19 ms on 2048 uneven boxes, 0.85 -> 0.65 on the periodic 48^3
// batch.
__device__ __forceinline__ void records_in(double* st, const double* __restrict__ src,
                                           int n) {
  double v[7];
#pragma unroll
  for (int u = 0; u < 7; ++u) {
    const int i = threadIdx.x + u * NT;
    v[u] = i < n ? src[i] : 0.0;
  }
#pragma unroll
  for (int u = 0; u < 7; ++u) {
    const int i = threadIdx.x + u * NT;
    if (i < n) st[i] = v[u];
  }
}

// uncompress (psi -> W, converted to T) or compress-accumulate (hpsi += W)
template <typename T, bool COMPRESS>
__global__ __launch_bounds__(NT) void k_uncompress(Batch B, TileTable tt) {
  static_assert(CPT == NT, "one fine record per thread");
  __shared__ double st[7 * CPT];
  const int tile = tile_of(tt);
  const int b = find_box(tt, tile);
  const BoxDesc bd = B.box[b];
  const int lt = tile - box_first_tile(tt, b);
  const int ctiles = (bd.nvctr_c + CPT - 1) / CPT;
  T* W = (T*)B.W + bd.w_off;
  if (lt < ctiles) {
    const int k = lt * CPT + threadIdx.x;
    if (k < bd.nvctr_c) {
      const long long w = psig_index(bd, bd.cmap[k], 0, 0, 0);
      if (COMPRESS) B.hpsi[bd.psi_off + k] += (double)W[w];
      else W[w] = (T)B.psi[bd.psi_off + k];
    }
  } else if constexpr (!COMPRESS && sizeof(T) == 4) {
    // FP32 grids: each thread reads its own record (the staged form below
    // measured slower here, 0.86 -> 0.90 ms on 2048 uneven boxes)
    const int k = (lt - ctiles) * CPT + threadIdx.x;
    if (k < bd.nvctr_f) {
      const long long w0 = psig_index(bd, bd.fmap[k], 0, 0, 0);
      const long long base = bd.psi_off + bd.nvctr_c + 7LL * k;
#pragma unroll
      for (int t = 1; t <= 7; ++t) W[w0 + fine_step(bd, t)] = (T)B.psi[base + t - 1];
    }
  } else {
    const int k0 = (lt - ctiles) * CPT;
    const int nk = min(CPT, bd.nvctr_f - k0);
    double* rec = (COMPRESS ? B.hpsi : (double*)B.psi) + bd.psi_off + bd.nvctr_c + 7LL * k0;
    records_in(st, rec, 7 * nk);
    __syncthreads();
    if ((int)threadIdx.x < nk) {
      const long long w0 = psig_index(bd, bd.fmap[k0 + threadIdx.x], 0, 0, 0);
      double* r = st + 7 * threadIdx.x;
#pragma unroll
      for (int t = 1; t <= 7; ++t) {
        const long long w = w0 + fine_step(bd, t);
        if (COMPRESS) r[t - 1] += (double)W[w];
        else W[w] = (T)r[t - 1];
      }
    }
    if (COMPRESS) {
      __syncthreads();
      for (int i = threadIdx.x; i < 7 * nk; i += NT) rec[i] = st[i];
    }
  }
}

// ---------------------------------------------------------------------------
// x pass, forward: T1 = G_x W, K = K_x W, ekin partial
// ---------------------------------------------------------------------------
// BIG: lines longer than the register-held outputs cover (N1 > 64 or
// L1 > 192): every thread stores its outputs straight to memory, 8 (G) or
// 4 + 4 (kinetic) consecutive values of its line at a time
template <typename T, bool BIG>
__global__ __launch_bounds__(NT) void k_xfwd(Batch B, TileTable tt, double* part) {
  extern __shared__ __align__(16) unsigned char smem[];
  T* lds = (T*)smem;
  __shared__ double red[NT / 64];
  const int tile = tile_of(tt);
  const int b = find_box(tt, tile);
  const BoxDesc bd = B.box[b];
  const int lt = tile - box_first_tile(tt, b);
  const int N1 = bd.N[0], L1 = bd.L[0], per = bd.per;
  const int W1 = 2 * N1;
  const int nlines = 4 * bd.N[1] * bd.N[2];
  const int l0 = lt * TW;
  const int nl = min(TW, nlines - l0);
  const T* src = (const T*)B.W + bd.w_off + (long long)l0 * W1;
  fill_wav<true>(lds, src, W1, 1, nl, N1, per);
  __syncthreads();
  const int tx = threadIdx.x & (TW - 1), ty = threadIdx.x >> 5;
  const int NS = N1 + HL + HR;
  const T* col = lds + tx;
  if (BIG) {
    T* dst = (T*)B.T1 + bd.t1_off + (long long)(l0 + tx) * L1;
    for (int p0 = 8 * ty; p0 < L1; p0 += 8 * TY) {
      T g[8];
      const int r0 = grow_row0(p0, per);
      grow8(col + r0 * LDW, col + (NS + r0) * LDW, g);
      if (tx < nl) {
#pragma unroll
        for (int r = 0; r < 8; ++r)
          if (p0 + r < L1) dst[p0 + r] = g[r];
      }
    }
    T* kdst = (T*)B.K + bd.w_off + (long long)(l0 + tx) * W1;
    const T ks = (T)bd.kscale[0];
    double ek = 0.0;
    for (int i0 = 4 * ty; i0 < N1; i0 += 4 * TY) {
      T kc[4], kd[4];
      kin4(col + (i0 + 1) * LDW, col + (NS + i0 + 1) * LDW, kc, kd);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        kc[r] *= ks;
        kd[r] *= ks;
        if (i0 + r < N1 && tx < nl) {
          ek += (double)kc[r] * (double)col[(i0 + r + HL) * LDW] +
                (double)kd[r] * (double)col[(NS + i0 + r + HL) * LDW];
          kdst[i0 + r] = kc[r];
          kdst[N1 + i0 + r] = kd[r];
        }
      }
    }
    const double eks = block_sum(ek, red);
    if (threadIdx.x == 0) part[tile] = eks;
    return;
  }
  T g[3][8];
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    const int p0 = 8 * (ty + m * TY);
    if (p0 < L1) {
      const int r0 = grow_row0(p0, per);
      grow8(col + r0 * LDW, col + (NS + r0) * LDW, g[m]);
    }
  }
  T kc[2][4], kd[2][4];
  double ek = 0.0;
  const T ks = (T)bd.kscale[0];
#pragma unroll
  for (int m = 0; m < 2; ++m) {
    const int i0 = 4 * (ty + m * TY);
    if (i0 < N1) {
      kin4(col + (i0 + 1) * LDW, col + (NS + i0 + 1) * LDW, kc[m], kd[m]);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        kc[m][r] *= ks;
        kd[m][r] *= ks;
        if (i0 + r < N1 && tx < nl)
          ek += (double)kc[m][r] * (double)col[(i0 + r + HL) * LDW] +
                (double)kd[m][r] * (double)col[(NS + i0 + r + HL) * LDW];
      }
    }
  }
  const double eks = block_sum(ek, red);
  if (threadIdx.x == 0) part[tile] = eks;
  __syncthreads();
  // stage G outputs as [p][line] and store contiguous lines of T1
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    const int p0 = 8 * (ty + m * TY);
#pragma unroll
    for (int r = 0; r < 8; ++r)
      if (p0 + r < L1) lds[(p0 + r) * LDW + tx] = g[m][r];
  }
  __syncthreads();
  T* dst = (T*)B.T1 + bd.t1_off + (long long)l0 * L1;
  {
    const float inv = 1.0f / (float)L1;
    for (int idx = threadIdx.x; idx < nl * L1; idx += NT) {
      const int l = qdiv(idx, L1, inv), p = idx - l * L1;
      dst[idx] = lds[p * LDW + l];
    }
  }
  __syncthreads();
  // stage K outputs and store contiguous lines of K
#pragma unroll
  for (int m = 0; m < 2; ++m) {
    const int i0 = 4 * (ty + m * TY);
#pragma unroll
    for (int r = 0; r < 4; ++r)
      if (i0 + r < N1) {
        lds[(i0 + r) * LDW + tx] = kc[m][r];
        lds[(N1 + i0 + r) * LDW + tx] = kd[m][r];
      }
  }
  __syncthreads();
  T* kdst = (T*)B.K + bd.w_off + (long long)l0 * W1;
  {
    const float inv = 1.0f / (float)W1;
    for (int idx = threadIdx.x; idx < nl * W1; idx += NT) {
      const int l = qdiv(idx, W1, inv), p = idx - l * W1;
      kdst[idx] = lds[p * LDW + l];
    }
  }
}

// ---------------------------------------------------------------------------
// y pass, forward: T2 = G_y T1
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(NT) void k_yfwd(Batch B, TileTable tt) {
  extern __shared__ __align__(16) unsigned char smem[];
  T* lds = (T*)smem;
  const int tile = tile_of(tt);
  const int b = find_box(tt, tile);
  const BoxDesc bd = B.box[b];
  const int lt = tile - box_first_tile(tt, b);
  const int L1 = bd.L[0], N2 = bd.N[1], L2 = bd.L[1], per = bd.per;
  const int W2 = 2 * N2;
  const int nxt = (L1 + TW - 1) / TW;
  const int j3 = lt / nxt;
  const int j1 = (lt - j3 * nxt) * TW;
  const int nl = min(TW, L1 - j1);
  const T* src = (const T*)B.T1 + bd.t1_off + j1 + (long long)L1 * W2 * j3;
  fill_wav<false>(lds, src, 1, L1, nl, N2, per);
  __syncthreads();
  const int tx = threadIdx.x & (TW - 1), ty = threadIdx.x >> 5;
  const int NS = N2 + HL + HR;
  const T* col = lds + tx;
  T* dst = (T*)B.T2 + bd.t2_off + j1 + tx + (long long)L1 * L2 * j3;
  for (int p0 = 8 * ty; p0 < L2; p0 += 8 * TY) {
    T g[8];
    const int r0 = grow_row0(p0, per);
    grow8(col + r0 * LDW, col + (NS + r0) * LDW, g);
    if (tx < nl) {
#pragma unroll
      for (int r = 0; r < 8; ++r)
        if (p0 + r < L2) dst[(long long)L1 * (p0 + r)] = g[r];
    }
  }
}

// ---------------------------------------------------------------------------
// kinetic along y (AX=1) or z (AX=2) on the wavelet volume: K += K_d W
// ---------------------------------------------------------------------------
template <typename T, int AX>
__global__ __launch_bounds__(NT) void k_kin(Batch B, TileTable tt, double* part) {
  extern __shared__ __align__(16) unsigned char smem[];
  T* lds = (T*)smem;
  __shared__ double red[NT / 64];
  const int tile = tile_of(tt);
  const int b = find_box(tt, tile);
  const BoxDesc bd = B.box[b];
  const int lt = tile - box_first_tile(tt, b);
  const int W1 = 2 * bd.N[0], W2 = 2 * bd.N[1];
  const int Nd = bd.N[AX];
  // AX = 1: 32 lines of one z (x = j1..); AX = 2: 32 consecutive z lines
  // x + W1 y of the box (P_KZ)
  const int nxt = (W1 + TW - 1) / TW;
  const int jo = AX == 1 ? lt / nxt : 0;   // z for AX = 1
  const int j1 = AX == 1 ? (lt - jo * nxt) * TW : lt * TW;
  const int nl = min(TW, (AX == 1 ? W1 : W1 * W2) - j1);
  const long long pstride = AX == 1 ? (long long)W1 : (long long)W1 * W2;
  const long long base = bd.w_off + j1 + (AX == 1 ? (long long)W1 * W2 * jo : 0LL);
  fill_wav<false>(lds, (const T*)B.W + base, 1, pstride, nl, Nd, bd.per);
  __syncthreads();
  const int tx = threadIdx.x & (TW - 1), ty = threadIdx.x >> 5;
  const int NS = Nd + HL + HR;
  const T* col = lds + tx;
  const T ks = (T)bd.kscale[AX];
  double ek = 0.0;
  T* kp = (T*)B.K + base + tx;
  for (int i0 = 4 * ty; i0 < Nd; i0 += 4 * TY) {
    // the chunk's 8 K values are read before the stencil and written after
    // it: read one at a time between the writes (the compiler cannot tell
    // the rows apart), each read waited for the previous write
    T k0[4], k1[4];
    if (tx < nl) {
#pragma unroll
      for (int r = 0; r < 4; ++r)
        if (i0 + r < Nd) {
          k0[r] = kp[pstride * (i0 + r)];
          k1[r] = kp[pstride * (Nd + i0 + r)];
        }
    }
    T oc[4], od[4];
    kin4(col + (i0 + 1) * LDW, col + (NS + i0 + 1) * LDW, oc, od);
    if (tx < nl) {
#pragma unroll
      for (int r = 0; r < 4; ++r)
        if (i0 + r < Nd) {
          const T vc = oc[r] * ks, vd = od[r] * ks;
          ek += (double)vc * (double)col[(i0 + r + HL) * LDW] +
                (double)vd * (double)col[(NS + i0 + r + HL) * LDW];
          kp[pstride * (i0 + r)] = k0[r] + vc;
          kp[pstride * (Nd + i0 + r)] = k1[r] + vd;
        }
    }
  }
  const double eks = block_sum(ek, red);
  if (threadIdx.x == 0) part[tile] = eks;
}

// ---------------------------------------------------------------------------
// z pass with the potential: T2 <- G_z^T ( V . G_z T2 ), epot partial
// ---------------------------------------------------------------------------
// work-group size of the z potential pass (k_zpot) per working type: the
// FP32 pass runs fastest as 128 threads (two waves, more chunks per thread),
// the FP64 one as 256 (measured on 2048 uneven free boxes)
template <typename T>
struct ZNT;
template <>
struct ZNT<float> {
  static constexpr int v = 128;
};
template <>
struct ZNT<double> {
  static constexpr int v = NT;
};


// one tile (32 consecutive z lines of T2, P_ZPF) of the z potential pass, by a
// work-group of NTZ threads: TY = NTZ / 32 threads per line, each holding
// its fine chunks of 8 points (at most MZ for lines of up to LMAX points)
template <typename T, int NTZ, int LMAX>
__device__ __forceinline__ void zpot_tile(const Batch& B, const TileTable& tt, double* part,
                                          int tile, T* lds, double* red) {
  const int b = find_box(tt, tile);
  const BoxDesc bd = B.box[b];
  const int lt = tile - box_first_tile(tt, b);
  const int L1 = bd.L[0], L2 = bd.L[1], N3 = bd.N[2], L3 = bd.L[2], per = bd.per;
  constexpr int TY = NTZ / TW, MZ = (LMAX + 8 * TY - 1) / (8 * TY);
  const int tx = threadIdx.x & (TW - 1), ty = threadIdx.x >> 5;
  const int NS = N3 + HL + HR;
  // the tile's lines: x + L1 y = q0 .. q0 + nl - 1 (P_ZPF); this thread's
  // line (x, y) (lanes past the box take the first line)
  const int q0 = lt * TW;
  const int nl = min(TW, L1 * L2 - q0);
  const int ql = q0 + (tx < nl ? tx : 0);
  const int ly = qdiv(ql, L1, 1.0f / (float)L1), lx = ql - ly * L1;
  // byte offsets from the tile's first line (a uniform base): one box's T2
  // holds L1 L2 2 N3 values, below 2^32 bytes for boxes of up to 256 coarse
  // points per axis; lanes past the box (tx >= nl) address line 0
  char* lineb = (char*)((T*)B.T2 + bd.t2_off + q0);
  const unsigned ls = (unsigned)sizeof(T) * (unsigned)(L1 * L2);   // bytes per z row
  const unsigned lo = (unsigned)sizeof(T) * (unsigned)(tx < nl ? tx : 0);
  {
    // this thread's line tx, wavelet rows j = ty, ty + TY, ... of each
    // section (c rows j, d rows N3 + j) into LDS rows HL + j and NS + HL + j;
    // lanes past the box store zeros; the F16X3 forward pass leaves its
    // table scales on the line (Batch::t2_scaled)
    const T mul = (sizeof(T) == 4 && B.t2_scaled)
                      ? (T)t2_line_unscale(B, bd, lx, ly) : (T)1;
    const T keep = tx < nl ? mul : (T)0;
#pragma unroll
    for (int sec = 0; sec < 2; ++sec) {
      const unsigned g0 = lo + ls * (unsigned)(sec * N3 + ty);
      T* l0 = lds + (sec * NS + HL + ty) * LDW + tx;
      for (int k0 = 0; ty + k0 * TY < N3; k0 += UNR) {
        T v[UNR];
#pragma unroll
        for (int u = 0; u < UNR; ++u) {
          const int k = k0 + u;
          v[u] = ty + k * TY < N3 ? *(const T*)(lineb + (g0 + ls * (unsigned)(k * TY))) : (T)0;
        }
#pragma unroll
        for (int u = 0; u < UNR; ++u) {
          const int k = k0 + u;
          if (ty + k * TY < N3) l0[k * TY * LDW] = v[u] * keep;
        }
      }
    }
    // halo rows (HL before and HR after each section): zero for free
    // boxes, the wrapped rows of the line for periodic ones
    if (!per) {
      for (int h = ty; h < HL + HR; h += TY) {
        const int row = h < HL ? h : N3 + h;   // coarse index h - HL or N3 + h - HL
        lds[row * LDW + tx] = (T)0;
        lds[(NS + row) * LDW + tx] = (T)0;
      }
    } else {
      __syncthreads();
      for (int h = ty; h < 2 * (HL + HR); h += TY) {
        const int sec = h >= HL + HR, r = h - sec * (HL + HR);
        const int k = r < HL ? r - HL : N3 + (r - HL);   // coarse index of the row
        lds[(sec * NS + k + HL) * LDW + tx] = lds[(sec * NS + wrap(k, N3) + HL) * LDW + tx];
      }
    }
  }
  __syncthreads();
  const T* col = lds + tx;
  // potential index of this line; outside the potential box V = 0, which
  // reproduces the zeroing of psir there (potential.f90:874-947).  Lanes off
  // the potential box read its first value (a valid address) and discard it,
  // so every potential load is unconditional.
  const int q1 = lx - bd.ishift[0], q2 = ly - bd.ishift[1];
  const bool in12 = tx < nl && q1 >= 0 && q1 < bd.potdim[0] && q2 >= 0 &&
                    q2 < bd.potdim[1];
  // byte offsets from the box's potential base (uniform): the host keeps a
  // potential box under 2^29 points, so they fit in 32 bits
  const char* potb = (const char*)(B.pot + bd.pot_off);
  const unsigned off12 = 8u * (unsigned)(in12 ? q1 + bd.potdim[0] * q2 : 0);
  const unsigned pz = 8u * (unsigned)(bd.potdim[0] * bd.potdim[1]);
  const int pd3 = bd.potdim[2];
  const double cxy = bd.conf_on ? conf_axis2(bd, 0, lx) + conf_axis2(bd, 1, ly) : 0.0;
  // potential box = the box, no confinement, and (periodic) no chunk past
  // the end of the line: no per-point selection (uniform over the box)
  const bool vexact = bd.ishift[0] == 0 && bd.ishift[1] == 0 && bd.ishift[2] == 0 &&
                      bd.potdim[0] == L1 && bd.potdim[1] == L2 && pd3 == L3 &&
                      !bd.conf_on && (!per || (L3 & 7) == 0);
  T g[MZ][8];
  double ep = 0.0;
#pragma unroll
  for (int m = 0; m < MZ; ++m) {
    const int p0 = 8 * (ty + m * TY);
    if (p0 < L3) {
      // the chunk's potential values are loaded first, with no use between
      // the loads, so they are in flight together while the grow runs
      const int q30 = p0 - bd.ishift[2];
      double v[8];
      if (vexact) {
#pragma unroll
        for (int r = 0; r < 8; ++r)
          v[r] = *(const double*)(potb + (off12 + pz * (unsigned)min(q30 + r, pd3 - 1)));
      } else {
#pragma unroll
        for (int r = 0; r < 8; ++r)
          v[r] = *(const double*)(potb + (off12 + pz * (unsigned)min(max(q30 + r, 0), pd3 - 1)));
      }
      const int r0 = grow_row0(p0, per);
      grow8(col + r0 * LDW, col + (NS + r0) * LDW, g[m]);
      if (vexact) {
        // the potential box is the box: every point of a line in the box
        // reads its own value, and the points without one (lines past the
        // box edge, free-BC points past L3) have psir = 0, so the values
        // they read (valid addresses) are multiplied by zero
#pragma unroll
        for (int r = 0; r < 8; ++r) {
          const double pr = (double)g[m][r];
          ep = fma(v[r] * pr, pr, ep);
          g[m][r] = (T)(v[r] * pr);
        }
      } else {
        // points r of the chunk inside the box and the potential box
        const int rlo = max(0, -q30);
        const int rhi = in12 ? min(L3 - p0, pd3 - q30) : 0;
#pragma unroll
        for (int r = 0; r < 8; ++r) {
          const bool ok = r >= rlo && r < rhi;
          double vr = ok ? v[r] : 0.0;
          if (bd.conf_on && ok) vr += conf_term(bd, cxy + conf_axis2(bd, 2, p0 + r));
          const double pr = (double)g[m][r];
          ep = fma(vr * pr, pr, ep);
          g[m][r] = (T)(vr * pr);
        }
      }
    }
  }
  const double eps = block_sum<NTZ>(ep, red);
  if (threadIdx.x == 0) part[tile] = eps;
  __syncthreads();
  // V psir as a real line bundle: row q holds fine index q (free) or q - 14
  // wrapped (periodic), rows 0 .. Q - 1.  Every chunk stores its eight rows:
  // free-BC points past L3 hold V psir = 0 (psir = 0 there, and V = 0 off
  // the exact case), and for periodic boxes the rows past the line are halo
  // rows, written again below.
  const int Q = 2 * N3 + 36;
  const int rofs = per ? 14 : 0;
#pragma unroll
  for (int m = 0; m < MZ; ++m) {
    const int p0 = 8 * (ty + m * TY);
    if (p0 < L3) {
      T* dst = lds + (p0 + rofs) * LDW + tx;
#pragma unroll
      for (int r = 0; r < 8; ++r) dst[r * LDW] = g[m][r];
    }
  }
  if (!per) {
    for (int q = 8 * ((L3 + 7) >> 3) + ty; q < Q; q += TY) lds[q * LDW + tx] = (T)0;
  } else {
    __syncthreads();
    // rows 0..13 and 2 N3 + 14 .. Q - 1 from the wrapped rows of the line
    for (int h = ty; h < 36; h += TY) {
      const int q = h < 14 ? h : 2 * N3 + h;
      lds[q * LDW + tx] = lds[(wrap(q - 14, 2 * N3) + 14) * LDW + tx];
    }
  }
  __syncthreads();
  for (int i0 = 4 * ty; i0 < N3; i0 += 4 * TY) {
    T oc[4], od[4];
    shrink4(col + 2 * i0 * LDW, oc, od);
    if (tx < nl) {
      char* dc = lineb + (lo + ls * (unsigned)i0);
#pragma unroll
      for (int r = 0; r < 4; ++r)
        if (i0 + r < N3) {
          *(T*)(dc + ls * (unsigned)r) = oc[r];
          *(T*)(dc + ls * (unsigned)(N3 + r)) = od[r];
        }
    }
  }
}

template <typename T, int NTZ, int LMAX>
__global__ __launch_bounds__(NTZ) void k_zpot(Batch B, TileTable tt, double* part) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ double red[NTZ / 64];
  zpot_tile<T, NTZ, LMAX>(B, tt, part, tile_of(tt), (T*)smem, red);
}

// The same pass as a persistent launch over a work queue: a grid of
// resident work-groups walks the queue of tiles (the boxes in order of
// decreasing work, each box's tiles together), every XCD one contiguous
// stretch of it, work-group i of an XCD its entries i, i + g, i + 2g, ...
// (g work-groups on that XCD).  Each work-group stops at the end of its
// stretch.  Same work-group size as k_zpot, so the epot partials are
// summed in the same order.
template <typename T, int NTZ>
__global__ __launch_bounds__(NTZ) void k_zpot_q(Batch B, TileTable tt, double* part,
                                                 const int* __restrict__ queue, int n) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ double red[NTZ / 64];
  const int G = gridDim.x, x = blockIdx.x & 7, i = blockIdx.x >> 3;
  const int g = (G >> 3) + (x < (G & 7) ? 1 : 0);
  const int q = n >> 3, r = n & 7;
  const int lo = x * q + min(x, r), hi = lo + q + (x < r ? 1 : 0);
  for (int k = lo + i; k < hi; k += g) {
    zpot_tile<T, NTZ, REG_L>(B, tt, part, queue[k], (T*)smem, red);
    __syncthreads();
  }
}

// the z potential pass over its tiles, in the short-line form when every
// line of the batch has at most ZSHORT_L points
template <typename T>
static void launch_zpot(const Batch& B, const TileTable& t, int ntiles, int zshort,
                        size_t lds, double* part, hipStream_t s) {
  constexpr int n = ZNT<T>::v;
  if (zshort) k_zpot<T, n, ZSHORT_L><<<ntiles, n, lds, s>>>(B, t, part);
  else k_zpot<T, n, REG_L><<<ntiles, n, lds, s>>>(B, t, part);
}

// ---------------------------------------------------------------------------
// The same pass for z lines longer than its register-held outputs cover
// (L3 > 192), in two kernels through the real volume R (L1, L2, L3) of the
// box: k_zgrowpot R = V G_z T2 (+ epot partial), k_zshrink T2 = G_z^T R.
// Same tiles, operators and order of operations as k_zpot.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(NT) void k_zgrowpot(Batch B, TileTable tt, double* part) {
  extern __shared__ __align__(16) unsigned char smem[];
  T* lds = (T*)smem;
  __shared__ double red[NT / 64];
  const int tile = tile_of(tt);
  const int b = find_box(tt, tile);
  const BoxDesc bd = B.box[b];
  const int lt = tile - box_first_tile(tt, b);
  const int L1 = bd.L[0], L2 = bd.L[1], N3 = bd.N[2], L3 = bd.L[2], per = bd.per;
  const int nxt = (L1 + TW - 1) / TW;
  const int j2 = lt / nxt;
  const int j1 = (lt - j2 * nxt) * TW;
  const int nl = min(TW, L1 - j1);
  const long long pstride = (long long)L1 * L2;
  const T* line = (const T*)B.T2 + bd.t2_off + j1 + (long long)L1 * j2;
  fill_wav<false>(lds, line, 1, pstride, nl, N3, per);
  __syncthreads();
  const int tx = threadIdx.x & (TW - 1), ty = threadIdx.x >> 5;
  const int NS = N3 + HL + HR;
  const T* col = lds + tx;
  const int q1 = j1 + tx - bd.ishift[0], q2 = j2 - bd.ishift[1];
  const bool in12 = tx < nl && q1 >= 0 && q1 < bd.potdim[0] && q2 >= 0 &&
                    q2 < bd.potdim[1];
  // lanes off the potential box read its first value (a valid address) and
  // discard it, so every potential load below is unconditional
  const double* pot = B.pot + bd.pot_off + (in12 ? q1 + (long long)bd.potdim[0] * q2 : 0);
  const long long pz = in12 ? (long long)bd.potdim[0] * bd.potdim[1] : 0;
  const double cxy = bd.conf_on ? conf_axis2(bd, 0, j1 + tx) + conf_axis2(bd, 1, j2) : 0.0;
  T* R = (T*)B.T3 + bd.t3_off + j1 + tx + (long long)L1 * j2;
  double ep = 0.0;
  for (int p0 = 8 * ty; p0 < L3; p0 += 8 * TY) {
    // the chunk's potential values first, in flight while the grow runs
    double pv[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      const int q3 = p0 + r - bd.ishift[2];
      pv[r] = pot[pz * min(max(q3, 0), bd.potdim[2] - 1)];
    }
    T g[8];
    const int r0 = grow_row0(p0, per);
    grow8(col + r0 * LDW, col + (NS + r0) * LDW, g);
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      const int q3 = p0 + r - bd.ishift[2];
      const bool ok = in12 && p0 + r < L3 && q3 >= 0 && q3 < bd.potdim[2];
      double v = ok ? pv[r] : 0.0;
      if (bd.conf_on && ok) v += conf_term(bd, cxy + conf_axis2(bd, 2, p0 + r));
      const double pr = (double)g[r];
      ep = fma(v * pr, pr, ep);
      if (tx < nl && p0 + r < L3) R[pstride * (p0 + r)] = (T)(v * pr);
    }
  }
  const double eps = block_sum(ep, red);
  if (threadIdx.x == 0) part[tile] = eps;
}

template <typename T>
__global__ __launch_bounds__(NT) void k_zshrink(Batch B, TileTable tt) {
  extern __shared__ __align__(16) unsigned char smem[];
  T* lds = (T*)smem;
  const int tile = tile_of(tt);
  const int b = find_box(tt, tile);
  const BoxDesc bd = B.box[b];
  const int lt = tile - box_first_tile(tt, b);
  const int L1 = bd.L[0], L2 = bd.L[1], N3 = bd.N[2], L3 = bd.L[2], per = bd.per;
  const int nxt = (L1 + TW - 1) / TW;
  const int j2 = lt / nxt;
  const int j1 = (lt - j2 * nxt) * TW;
  const int nl = min(TW, L1 - j1);
  const long long pstride = (long long)L1 * L2;
  const T* R = (const T*)B.T3 + bd.t3_off + j1 + (long long)L1 * j2;
  fill_real<false>(lds, R, 1, pstride, nl, N3, L3, per);
  __syncthreads();
  const int tx = threadIdx.x & (TW - 1), ty = threadIdx.x >> 5;
  const T* col = lds + tx;
  T* line = (T*)B.T2 + bd.t2_off + j1 + (long long)L1 * j2;
  for (int i0 = 4 * ty; i0 < N3; i0 += 4 * TY) {
    T oc[4], od[4];
    shrink4(col + 2 * i0 * LDW, oc, od);
    if (tx < nl) {
#pragma unroll
      for (int r = 0; r < 4; ++r)
        if (i0 + r < N3) {
          line[tx + pstride * (i0 + r)] = oc[r];
          line[tx + pstride * (N3 + i0 + r)] = od[r];
        }
    }
  }
}

// ---------------------------------------------------------------------------
// y pass, backward: T1 = G_y^T T2
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(NT) void k_ybwd(Batch B, TileTable tt) {
  extern __shared__ __align__(16) unsigned char smem[];
  T* lds = (T*)smem;
  const int tile = tile_of(tt);
  const int b = find_box(tt, tile);
  const BoxDesc bd = B.box[b];
  const int lt = tile - box_first_tile(tt, b);
  const int L1 = bd.L[0], N2 = bd.N[1], L2 = bd.L[1], per = bd.per;
  const int W2 = 2 * N2;
  const int nxt = (L1 + TW - 1) / TW;
  const int j3 = lt / nxt;
  const int j1 = (lt - j3 * nxt) * TW;
  const int nl = min(TW, L1 - j1);
  const T* src = (const T*)B.T2 + bd.t2_off + j1 + (long long)L1 * L2 * j3;
  fill_real<false>(lds, src, 1, L1, nl, N2, L2, per);
  __syncthreads();
  const int tx = threadIdx.x & (TW - 1), ty = threadIdx.x >> 5;
  const T* col = lds + tx;
  T* dst = (T*)B.T1 + bd.t1_off + j1 + tx + (long long)L1 * W2 * j3;
  for (int i0 = 4 * ty; i0 < N2; i0 += 4 * TY) {
    T oc[4], od[4];
    shrink4(col + 2 * i0 * LDW, oc, od);
    if (tx < nl) {
#pragma unroll
      for (int r = 0; r < 4; ++r)
        if (i0 + r < N2) {
          dst[(long long)L1 * (i0 + r)] = oc[r];
          dst[(long long)L1 * (N2 + i0 + r)] = od[r];
        }
    }
  }
}

// ---------------------------------------------------------------------------
// x pass, backward: W = G_x^T T1 + K
// ---------------------------------------------------------------------------
// BIG: N1 > 64 (more coarse outputs than the registers hold): outputs
// plus K go straight to memory
template <typename T, bool BIG>
__global__ __launch_bounds__(NT) void k_xbwd(Batch B, TileTable tt) {
  extern __shared__ __align__(16) unsigned char smem[];
  T* lds = (T*)smem;
  const int tile = tile_of(tt);
  const int b = find_box(tt, tile);
  const BoxDesc bd = B.box[b];
  const int lt = tile - box_first_tile(tt, b);
  const int N1 = bd.N[0], L1 = bd.L[0], per = bd.per;
  const int W1 = 2 * N1;
  const int nlines = 4 * bd.N[1] * bd.N[2];
  const int l0 = lt * TW;
  const int nl = min(TW, nlines - l0);
  const T* src = (const T*)B.T1 + bd.t1_off + (long long)l0 * L1;
  fill_real<true>(lds, src, L1, 1, nl, N1, L1, per);
  __syncthreads();
  const int tx = threadIdx.x & (TW - 1), ty = threadIdx.x >> 5;
  const T* col = lds + tx;
  if (BIG) {
    const long long o = bd.w_off + (long long)(l0 + tx) * W1;
    T* W = (T*)B.W + o;
    const T* K = (const T*)B.K + o;
    for (int i0 = 4 * ty; i0 < N1; i0 += 4 * TY) {
      T oc[4], od[4];
      shrink4(col + 2 * i0 * LDW, oc, od);
      if (tx < nl) {
#pragma unroll
        for (int r = 0; r < 4; ++r)
          if (i0 + r < N1) {
            W[i0 + r] = oc[r] + K[i0 + r];
            W[N1 + i0 + r] = od[r] + K[N1 + i0 + r];
          }
      }
    }
    return;
  }
  T oc[2][4], od[2][4];
#pragma unroll
  for (int m = 0; m < 2; ++m) {
    const int i0 = 4 * (ty + m * TY);
    if (i0 < N1) shrink4(col + 2 * i0 * LDW, oc[m], od[m]);
  }
  __syncthreads();
#pragma unroll
  for (int m = 0; m < 2; ++m) {
    const int i0 = 4 * (ty + m * TY);
#pragma unroll
    for (int r = 0; r < 4; ++r)
      if (i0 + r < N1) {
        lds[(i0 + r) * LDW + tx] = oc[m][r];
        lds[(N1 + i0 + r) * LDW + tx] = od[m][r];
      }
  }
  __syncthreads();
  const long long o = bd.w_off + (long long)l0 * W1;
  T* W = (T*)B.W;
  const T* K = (const T*)B.K;
  const float inv = 1.0f / (float)W1;
  const int tot = nl * W1;
  for (int base = threadIdx.x; base < tot; base += NT * UNR) {
    T kv[UNR];
#pragma unroll
    for (int u = 0; u < UNR; ++u) {
      const int idx = base + u * NT;
      kv[u] = idx < tot ? K[o + idx] : (T)0;
    }
#pragma unroll
    for (int u = 0; u < UNR; ++u) {
      const int idx = base + u * NT;
      if (idx < tot) {
        const int l = qdiv(idx, W1, inv), p = idx - l * W1;
        W[o + idx] = lds[p * LDW + l] + kv[u];
      }
    }
  }
}

// ---------------------------------------------------------------------------
// energies: fixed-order sums of the per-tile partials of each box
// ---------------------------------------------------------------------------
// one work-group per box; fixed summation order (strided per thread, then a
// fixed tree) so the energies are reproducible run to run
__global__ __launch_bounds__(128) void k_energy(EnergyTables et, int nbox, double* ekin,
                                                double* epot) {
  __shared__ double sh[128];
  const int b = blockIdx.x;
  double e[2] = {0.0, 0.0};
  for (int s = 0; s < et.n; ++s) {
    const TileTable& tt = et.t[s];
    const int t0 = box_first_tile(tt, b);
    const int t1 = tt.uniform > 0 ? t0 + tt.uniform : tt.start[b + 1];
    double acc = 0.0;
    for (int t = t0 + (int)threadIdx.x; t < t1; t += 128) acc += et.p[s][t];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 64; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) e[s < et.nk ? 0 : 1] += sh[0];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    ekin[b] = e[0];
    epot[b] = e[1];
  }
}

// ---------------------------------------------------------------------------
// density (local_partial_density, sumrho.f90:377-469; partial_density,
// sumrho.f90:472-560): z lines of T2 = G_y G_x W for the orbitals of one
// group in turn, psir = G_z line, and the group's partial density
// sum hfac * psir^2 per spin kept in registers, written once.
// ---------------------------------------------------------------------------
template <typename T, int NSPIN>
__global__ __launch_bounds__(NT) void k_zden(Batch B, DenArgs D) {
  extern __shared__ __align__(16) unsigned char smem[];
  T* lds = (T*)smem;
  const int g = blockIdx.x / D.nbund;
  const int bund = blockIdx.x - g * D.nbund;
  const BoxDesc bd0 = B.box[0];
  const int L1 = bd0.L[0], L2 = bd0.L[1], N3 = bd0.N[2], L3 = bd0.L[2], per = bd0.per;
  const int j2 = bund / D.nxt;
  const int j1 = (bund - j2 * D.nxt) * TW;
  const int nl = min(TW, L1 - j1);
  const long long pstride = (long long)L1 * L2;
  const int tx = threadIdx.x & (TW - 1), ty = threadIdx.x >> 5;
  const int NS = N3 + HL + HR;
  const T* col = lds + tx;
  const int o0 = g * D.gsize, o1 = min(D.norb, o0 + D.gsize);
  // z lines longer than the 192 outputs the registers hold are done in
  // rounds of 192 outputs, each reading the group's lines again
  for (int pb = 0; pb < L3; pb += 24 * TY) {
    double acc[NSPIN][3][8];
#pragma unroll
    for (int s = 0; s < NSPIN; ++s)
#pragma unroll
      for (int m = 0; m < 3; ++m)
#pragma unroll
        for (int r = 0; r < 8; ++r) acc[s][m][r] = 0.0;
    for (int o = o0; o < o1; ++o) {
      const double hf = D.hfac[o];
      const int sp = NSPIN == 2 ? D.spin[o] : 0;
      const T* line = (const T*)B.T2 + B.box[o].t2_off + j1 + (long long)L1 * j2;
      fill_wav<false>(lds, line, 1, pstride, nl, N3, per);
      __syncthreads();
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        const int p0 = pb + 8 * (ty + m * TY);
        if (p0 < L3) {
          T gv[8];
          grow8(col + grow_row0(p0, per) * LDW, col + (NS + grow_row0(p0, per)) * LDW, gv);
#pragma unroll
          for (int r = 0; r < 8; ++r) {
            const double v = (double)gv[r];
            if (NSPIN == 2 && sp == 1) acc[NSPIN - 1][m][r] += hf * (v * v);
            else acc[0][m][r] += hf * (v * v);
          }
        }
      }
      __syncthreads();
    }
    if (tx < nl) {
#pragma unroll
      for (int s = 0; s < NSPIN; ++s) {
        double* dst = D.part + ((long long)g * D.nspin + s) * D.nisf + j1 + tx +
                      (long long)L1 * j2;
#pragma unroll
        for (int m = 0; m < 3; ++m) {
          const int p0 = pb + 8 * (ty + m * TY);
#pragma unroll
          for (int r = 0; r < 8; ++r)
            if (p0 + r < L3) dst[pstride * (p0 + r)] = acc[s][m][r];
        }
      }
    }
  }
}

// rho += sum of the group partials, groups in order
__global__ __launch_bounds__(NT) void k_rho_reduce(DenArgs D) {
  const long long n = (long long)D.nspin * D.nisf;
  for (long long i = blockIdx.x * (long long)NT + threadIdx.x; i < n;
       i += (long long)gridDim.x * NT) {
    const int s = (int)(i / D.nisf);
    const long long p = i - (long long)s * D.nisf;
    double a = 0.0;
    for (int g = 0; g < D.ngroups; ++g) a += D.part[((long long)g * D.nspin + s) * D.nisf + p];
    D.rho[i] += a;
  }
}

template <typename T>
static hipError_t launch_den(const Batch& B, const PassTables& PT, const DenArgs& D, int maxN,
                             long long wtotal, bool need_zero, hipStream_t s) {
  const size_t lds = lds_bytes(maxN, sizeof(T));
  if (need_zero) {
    const long long nb = (wtotal + NT - 1) / NT;
    k_zero<T><<<(unsigned)(nb < 65536 ? nb : 65536), NT, 0, s>>>((T*)B.W, wtotal);
  }
  if (PT.ntiles[P_UNC] > 0) k_uncompress<T, false><<<PT.ntiles[P_UNC], NT, 0, s>>>(B, PT.t[P_UNC]);
  if (PT.plane) {
    // x and y grow of each plane in one pass through LDS
    if (PT.ntiles[P_PF] > 0) {
      const hipError_t e = launch_plane_grow<T>(B, PT.t[P_PF], PT.ntiles[P_PF], PT.lds_pf, s);
      if (e != hipSuccess) return e;
    }
  } else {
    if (PT.ntiles[P_XF] > 0) {
      if (PT.xbig) k_xfwd<T, true><<<PT.ntiles[P_XF], NT, lds, s>>>(B, PT.t[P_XF], PT.part[0]);
      else k_xfwd<T, false><<<PT.ntiles[P_XF], NT, lds, s>>>(B, PT.t[P_XF], PT.part[0]);
    }
    if (PT.ntiles[P_YF] > 0) k_yfwd<T><<<PT.ntiles[P_YF], NT, lds, s>>>(B, PT.t[P_YF]);
  }
  const int nblk = D.ngroups * D.nbund;
  if (D.nspin == 2) k_zden<T, 2><<<nblk, NT, lds, s>>>(B, D);
  else k_zden<T, 1><<<nblk, NT, lds, s>>>(B, D);
  const long long n = (long long)D.nspin * D.nisf;
  const long long nb = (n + NT - 1) / NT;
  k_rho_reduce<<<(unsigned)(nb < 65536 ? nb : 65536), NT, 0, s>>>(D);
  return hipGetLastError();
}

int zpot_queue_grid(int fp32, int maxN) {
  const size_t lds = lds_bytes(maxN, fp32 ? sizeof(float) : sizeof(double));
  int per_cu = 0, dev = 0, ncu = 0;
  const hipError_t e =
      fp32 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(
                 &per_cu, k_zpot_q<float, ZNT<float>::v>, ZNT<float>::v, lds)
           : hipOccupancyMaxActiveBlocksPerMultiprocessor(
                 &per_cu, k_zpot_q<double, ZNT<double>::v>, ZNT<double>::v, lds);
  if (e != hipSuccess || hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      per_cu <= 0 || ncu <= 0) {
    (void)hipGetLastError();
    return 0;
  }
  return ((per_cu * ncu) >> 3) << 3;
}

hipError_t launch_density(const Batch& B, const PassTables& T, const DenArgs& D, int maxN,
                          long long wtotal, bool need_zero, int fp32, hipStream_t s) {
  return fp32 ? launch_den<float>(B, T, D, maxN, wtotal, need_zero, s)
              : launch_den<double>(B, T, D, maxN, wtotal, need_zero, s);
}

// ---------------------------------------------------------------------------
// host-side launcher of the whole FP64 sequence
// ---------------------------------------------------------------------------
size_t lds_bytes(int maxN, int elem) {
  return (size_t)2 * (maxN + HL + HR) * LDW * elem;
}

// z potential pass over the P_ZP tiles: one kernel, or two through T3 for
// z lines longer than the registers hold
template <typename T>
static void zpot_pass(const Batch& B, const PassTables& PT, size_t lds, hipStream_t s) {
  if (PT.ntiles[P_ZP] <= 0) return;
  if (PT.zbig) {
    k_zgrowpot<T><<<PT.ntiles[P_ZP], NT, lds, s>>>(B, PT.t[P_ZP], PT.part[3]);
    k_zshrink<T><<<PT.ntiles[P_ZP], NT, lds, s>>>(B, PT.t[P_ZP]);
  } else if (PT.zq && PT.zq_grid > 0) {
    k_zpot_q<T, ZNT<T>::v><<<PT.zq_grid, ZNT<T>::v, lds, s>>>(B, PT.t[P_ZPF], PT.part[3],
                                                             PT.zq, PT.ntiles[P_ZPF]);
  } else {
    launch_zpot<T>(B, PT.t[P_ZPF], PT.ntiles[P_ZPF], PT.zshort, lds, PT.part[3], s);
  }
}

hipError_t launch_zpot_f32(const Batch& B, const PassTables& PT, int maxN, hipStream_t s) {
  zpot_pass<float>(B, PT, lds_bytes(maxN, sizeof(float)), s);
  return hipGetLastError();
}

template <typename T>
static hipError_t launch_seq(const Batch& B, const PassTables& PT, int maxN,
                             long long wtotal, bool need_zero, hipStream_t s,
                             hipEvent_t* ev) {
  const size_t lds = lds_bytes(maxN, sizeof(T));
  auto mark = [&](int stage, int end) {
    if (ev) (void)hipEventRecord(ev[2 * stage + end], s);
  };
  mark(S_ZERO, 0);
  if (need_zero) {
    const long long nb = (wtotal + NT - 1) / NT;
    k_zero<T><<<(unsigned)(nb < 65536 ? nb : 65536), NT, 0, s>>>((T*)B.W, wtotal);
  }
  mark(S_ZERO, 1);
#define BDFT_STAGE(S, P, SH, KERNEL, ...)                              \
  mark(S, 0);                                                          \
  if (PT.ntiles[P] > 0) KERNEL<<<PT.ntiles[P], NT, SH, s>>>(__VA_ARGS__); \
  mark(S, 1);
  BDFT_STAGE(S_UNC, P_UNC, 0, (k_uncompress<T, false>), B, PT.t[P_UNC])
  EnergyTables et;
  if (PT.plane) {
    mark(S_PF, 0);
    if (PT.n_pfg > 0) {
      for (int k = 0; k < PT.n_pfg; ++k)
        if (PT.nt_pfg[k] > 0)
          launch_plane_fwd<T>(B, PT.pfg[k], PT.t[P_PF], PT.nt_pfg[k], PT.lds_pfg[k],
                              PT.part[0], s);
    } else if (PT.ntiles[P_PF] > 0) {
      launch_plane_fwd<T>(B, PT.t[P_PF], PT.t[P_PF], PT.ntiles[P_PF], PT.lds_pf, PT.part[0],
                          s);
    }
    mark(S_PF, 1);
    BDFT_STAGE(S_KZ, P_KZ, lds, (k_kin<T, 2>), B, PT.t[P_KZ], PT.part[2])
    mark(S_ZP, 0);
    zpot_pass<T>(B, PT, lds, s);
    mark(S_ZP, 1);
    mark(S_PB, 0);
    if (PT.n_pbg > 0) {
      for (int k = 0; k < PT.n_pbg; ++k)
        if (PT.nt_pbg[k] > 0)
          launch_plane_bwd<T>(B, PT.pbg[k], PT.nt_pbg[k], PT.lds_pbg[k], s,
                              sizeof(T) == 8 && PT.fit_pbg[k] == 2);
    } else if (PT.ntiles[P_PB] > 0) {
      launch_plane_bwd<T>(B, PT.t[P_PB], PT.ntiles[P_PB], PT.lds_pb, s);
    }
    mark(S_PB, 1);
    et.t[0] = PT.t[P_PF];
    et.p[0] = PT.part[0];
    et.t[1] = PT.t[P_KZ];
    et.p[1] = PT.part[2];
    et.t[2] = zpot_table(PT);
    et.p[2] = PT.part[3];
    et.nk = 2;
    et.n = 3;
  } else {
    if (PT.xbig) {
      BDFT_STAGE(S_XF, P_XF, lds, (k_xfwd<T, true>), B, PT.t[P_XF], PT.part[0])
    } else {
      BDFT_STAGE(S_XF, P_XF, lds, (k_xfwd<T, false>), B, PT.t[P_XF], PT.part[0])
    }
    BDFT_STAGE(S_YF, P_YF, lds, k_yfwd<T>, B, PT.t[P_YF])
    BDFT_STAGE(S_KY, P_KY, lds, (k_kin<T, 1>), B, PT.t[P_KY], PT.part[1])
    BDFT_STAGE(S_KZ, P_KZ, lds, (k_kin<T, 2>), B, PT.t[P_KZ], PT.part[2])
    mark(S_ZP, 0);
    zpot_pass<T>(B, PT, lds, s);
    mark(S_ZP, 1);
    BDFT_STAGE(S_YB, P_YB, lds, k_ybwd<T>, B, PT.t[P_YB])
    if (PT.xbig) {
      BDFT_STAGE(S_XB, P_XB, lds, (k_xbwd<T, true>), B, PT.t[P_XB])
    } else {
      BDFT_STAGE(S_XB, P_XB, lds, (k_xbwd<T, false>), B, PT.t[P_XB])
    }
    BDFT_STAGE(S_COMP, P_UNC, 0, (k_uncompress<T, true>), B, PT.t[P_UNC])
    et.t[0] = PT.t[P_XF];
    et.t[1] = PT.t[P_KY];
    et.t[2] = PT.t[P_KZ];
    et.t[3] = zpot_table(PT);
    for (int i = 0; i < 4; ++i) et.p[i] = PT.part[i];
    et.nk = 3;
    et.n = 4;
  }
#undef BDFT_STAGE
  mark(S_ENERGY, 0);
  k_energy<<<B.nbox, 128, 0, s>>>(et, B.nbox, B.ekin, B.epot);
  mark(S_ENERGY, 1);
  if (ev) {  // stages of the other sequence: empty event pairs
    const int unused_plane[] = {S_XF, S_YF, S_KY, S_YB, S_XB, S_COMP};
    const int unused_lines[] = {S_PF, S_PB};
    if (PT.plane)
      for (int st : unused_plane) { mark(st, 0); mark(st, 1); }
    else
      for (int st : unused_lines) { mark(st, 0); mark(st, 1); }
  }
  return hipGetLastError();
}

hipError_t launch_passes(const Batch& B, const PassTables& T, int maxN, long long wtotal,
                         bool need_zero, int fp32, hipStream_t s, hipEvent_t* ev) {
  return fp32 ? launch_seq<float>(B, T, maxN, wtotal, need_zero, s, ev)
              : launch_seq<double>(B, T, maxN, wtotal, need_zero, s, ev);
}

hipError_t launch_transfer(const Batch& B, const TileTable& t, int ntiles,
                           long long wtotal, bool compress, hipStream_t s) {
  if (!compress) {
    const long long nb = (wtotal + NT - 1) / NT;
    k_zero<double><<<(unsigned)(nb < 65536 ? nb : 65536), NT, 0, s>>>((double*)B.W, wtotal);
    k_uncompress<double, false><<<ntiles, NT, 0, s>>>(B, t);
  } else {
    k_uncompress<double, true><<<ntiles, NT, 0, s>>>(B, t);
  }
  return hipGetLastError();
}

void extern_zero_float(float* p, long long n, int nblocks, hipStream_t s) {
  k_zero<float><<<nblocks, NT, 0, s>>>(p, n);
}
// uncompress into the FP32 grid and record each tile's largest |psi| in
// tmax[tile]; k_box_max then reduces the tiles of each box into bmax[box]
__global__ __launch_bounds__(NT) void k_uncompress_max(Batch B, TileTable tt, float* tmax) {
  __shared__ float red[NT / 64];
  const int tile = tile_of(tt);
  const int b = find_box(tt, tile);
  const BoxDesc bd = B.box[b];
  const int lt = tile - box_first_tile(tt, b);
  const int ctiles = (bd.nvctr_c + CPT - 1) / CPT;
  float* W = (float*)B.W + bd.w_off;
  float mx = 0.0f;
  if (lt < ctiles) {
    const int k = lt * CPT + threadIdx.x;
    if (k < bd.nvctr_c) {
      const float v = (float)B.psi[bd.psi_off + k];
      W[psig_index(bd, bd.cmap[k], 0, 0, 0)] = v;
      mx = fabsf(v);
    }
  } else {
    const int k = (lt - ctiles) * CPT + threadIdx.x;
    if (k < bd.nvctr_f) {
      const long long w0 = psig_index(bd, bd.fmap[k], 0, 0, 0);
      const long long base = bd.psi_off + bd.nvctr_c + 7LL * k;
#pragma unroll
      for (int t = 1; t <= 7; ++t) {
        const float v = (float)B.psi[base + t - 1];
        W[w0 + fine_step(bd, t)] = v;
        mx = fmaxf(mx, fabsf(v));
      }
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    float m = red[0];
    for (int w = 1; w < NT / 64; ++w) m = fmaxf(m, red[w]);
    tmax[tile] = m;
  }
}

// one block per box: bmax[box] = max over the box's tiles (as float bits)
__global__ __launch_bounds__(NT) void k_box_max(Batch B, TileTable tt, const float* tmax) {
  __shared__ float red[NT / 64];
  const int b = blockIdx.x;
  const int t0 = box_first_tile(tt, b);
  const int t1 = tt.uniform > 0 ? t0 + tt.uniform : tt.start[b + 1];
  float mx = 0.0f;
  for (int t = t0 + (int)threadIdx.x; t < t1; t += NT) mx = fmaxf(mx, tmax[t]);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    float m = red[0];
    for (int w = 1; w < NT / 64; ++w) m = fmaxf(m, red[w]);
    B.bmax[b] = __float_as_uint(m);
  }
}

void launch_uncompress_float(const Batch& B, const TileTable& t, int ntiles, bool track_max,
                             hipStream_t s) {
  if (track_max) {
    k_uncompress_max<<<ntiles, NT, 0, s>>>(B, t, B.tmax);
    k_box_max<<<B.nbox, NT, 0, s>>>(B, t, B.tmax);
  } else {
    k_uncompress<float, false><<<ntiles, NT, 0, s>>>(B, t);
  }
}
void launch_energy(const EnergyTables& et, const Batch& B, hipStream_t s) {
  k_energy<<<B.nbox, 128, 0, s>>>(et, B.nbox, B.ekin, B.epot);
}


// ---------------------------------------------------------------------------
// Conjugate-gradient preconditioner: ocl_preconditioner_generic_k
// (bigdft/src/OpenCL/Preconditioner.c:245-358), real case, batched over
// orbitals.  apply_hp (:198-243) = S Ana (K + c) Syn S on the compressed
// vector equals S (K~ + c) S with K~ the wavelet-basis kinetic, as
// Ana Syn = I on the box: W = S v scattered, K = c W + K~_x W (k_pc_kinx),
// K += K~_y W, K~_z W (k_kin<double, 1/2>), q = S K gathered.  The CG
// scalars stay on the device (per box ns, alpha, beta), reduced from
// per-tile partials in a fixed order.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double pc_scal(const PcArgs& A, int b, int t) {
  return A.pc[9 * b + 1 + t];
}

// Element range of compressed tile lt of a box: CPT coarse coefficients, or
// the 7 CPT contiguous coefficients of CPT fine points (the vector side of
// the preconditioner is read and written through these ranges, coalesced;
// the 7 components of a fine point go through LDS to the thread that owns
// the point's grid positions)
struct PcRange {
  long long e0;   // first vector element
  int ne;         // elements
  int k0;         // first point
  int fine;
};
__device__ __forceinline__ PcRange pc_range(const BoxDesc& bd, int lt) {
  const int ctiles = (bd.nvctr_c + CPT - 1) / CPT;
  PcRange R;
  R.fine = lt >= ctiles;
  if (!R.fine) {
    R.k0 = lt * CPT;
    R.e0 = bd.psi_off + R.k0;
    R.ne = min(CPT, bd.nvctr_c - R.k0);
  } else {
    R.k0 = (lt - ctiles) * CPT;
    R.e0 = bd.psi_off + bd.nvctr_c + 7LL * R.k0;
    R.ne = 7 * min(CPT, bd.nvctr_f - R.k0);
  }
  return R;
}

// W <- S v at the region's points; update: d <- r + beta d first, v = d
__global__ __launch_bounds__(NT) void k_pc_scatter(Batch B, TileTable tt, PcArgs A,
                                                   const double* src, int update) {
  __shared__ double st[7 * CPT];
  const int tile = tile_of(tt);
  const int b = find_box(tt, tile);
  const BoxDesc bd = B.box[b];
  const PcRange R = pc_range(bd, tile - box_first_tile(tt, b));
  double* W = (double*)B.W + bd.w_off;
  const double beta = update ? A.sc[4 * b + 2] : 0.0;
  // the tile's vector elements, coalesced (d updated in place)
  for (int e = threadIdx.x; e < R.ne; e += NT) {
    const long long i = R.e0 + e;
    double v;
    if (update) {
      v = fma(beta, A.d[i], A.r[i]);
      A.d[i] = v;
    } else {
      v = src[i];
    }
    st[e] = v;
  }
  __syncthreads();
  const int k = R.k0 + threadIdx.x;
  if (!R.fine) {
    if (threadIdx.x < R.ne)
      W[psig_index(bd, bd.cmap[k], 0, 0, 0)] = pc_scal(A, b, 0) * st[threadIdx.x];
  } else if (7 * (int)threadIdx.x < R.ne) {
    const int lin = bd.fmap[k];
#pragma unroll
    for (int t = 1; t <= 7; ++t)
      W[psig_index(bd, lin, t & 1, (t >> 1) & 1, (t >> 2) & 1)] =
          pc_scal(A, b, t) * st[7 * threadIdx.x + t - 1];
  }
}

// K = c W + K~_x W along x lines (the line loading of k_xfwd); BIG: more
// than 64 coarse points per line, outputs stored as they are made
template <bool BIG>
__global__ __launch_bounds__(NT) void k_pc_kinx(Batch B, TileTable tt, PcArgs A) {
  extern __shared__ __align__(16) unsigned char smem[];
  double* lds = (double*)smem;
  const int tile = tile_of(tt);
  const int b = find_box(tt, tile);
  const BoxDesc bd = B.box[b];
  const int lt = tile - box_first_tile(tt, b);
  const int N1 = bd.N[0], per = bd.per;
  const int W1 = 2 * N1;
  const int nlines = 4 * bd.N[1] * bd.N[2];
  const int l0 = lt * TW;
  const int nl = min(TW, nlines - l0);
  const double* src = (const double*)B.W + bd.w_off + (long long)l0 * W1;
  fill_wav<true>(lds, src, W1, 1, nl, N1, per);
  __syncthreads();
  const int tx = threadIdx.x & (TW - 1), ty = threadIdx.x >> 5;
  const int NS = N1 + HL + HR;
  const double* col = lds + tx;
  const double ks = bd.kscale[0], c = A.pc[9 * b];
  if (BIG) {
    double* kl = (double*)B.K + bd.w_off + (long long)(l0 + tx) * W1;
    for (int i0 = 4 * ty; i0 < N1; i0 += 4 * TY) {
      double oc[4], od[4];
      kin4(col + (i0 + 1) * LDW, col + (NS + i0 + 1) * LDW, oc, od);
#pragma unroll
      for (int r = 0; r < 4; ++r)
        if (i0 + r < N1 && tx < nl) {
          kl[i0 + r] = fma(c, col[(i0 + r + HL) * LDW], oc[r] * ks);
          kl[N1 + i0 + r] = fma(c, col[(NS + i0 + r + HL) * LDW], od[r] * ks);
        }
    }
    return;
  }
  double kc[2][4], kd[2][4];
#pragma unroll
  for (int m = 0; m < 2; ++m) {
    const int i0 = 4 * (ty + m * TY);
    if (i0 < N1) {
      kin4(col + (i0 + 1) * LDW, col + (NS + i0 + 1) * LDW, kc[m], kd[m]);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = min(i0 + r, N1 - 1);
        kc[m][r] = fma(c, col[(i + HL) * LDW], kc[m][r] * ks);
        kd[m][r] = fma(c, col[(NS + i + HL) * LDW], kd[m][r] * ks);
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int m = 0; m < 2; ++m) {
    const int i0 = 4 * (ty + m * TY);
#pragma unroll
    for (int r = 0; r < 4; ++r)
      if (i0 + r < N1) {
        lds[(i0 + r) * LDW + tx] = kc[m][r];
        lds[(N1 + i0 + r) * LDW + tx] = kd[m][r];
      }
  }
  __syncthreads();
  double* kdst = (double*)B.K + bd.w_off + (long long)l0 * W1;
  const float inv = 1.0f / (float)W1;
  for (int idx = threadIdx.x; idx < nl * W1; idx += NT) {
    const int l = qdiv(idx, W1, inv), p = idx - l * W1;
    kdst[idx] = lds[p * LDW + l];
  }
}

// v = S K at the region's points.  init: r = b - v, d = r, partial |r|^2;
// else q = v, partial q.d
__global__ __launch_bounds__(NT) void k_pc_gather(Batch B, TileTable tt, PcArgs A, int init) {
  __shared__ double red[NT / 64];
  __shared__ double st[7 * CPT];
  const int tile = tile_of(tt);
  const int b = find_box(tt, tile);
  const BoxDesc bd = B.box[b];
  const PcRange R = pc_range(bd, tile - box_first_tile(tt, b));
  const double* K = (const double*)B.K + bd.w_off;
  const int k = R.k0 + threadIdx.x;
  if (!R.fine) {
    if (threadIdx.x < R.ne)
      st[threadIdx.x] = pc_scal(A, b, 0) * K[psig_index(bd, bd.cmap[k], 0, 0, 0)];
  } else if (7 * (int)threadIdx.x < R.ne) {
    const int lin = bd.fmap[k];
#pragma unroll
    for (int t = 1; t <= 7; ++t)
      st[7 * threadIdx.x + t - 1] =
          pc_scal(A, b, t) * K[psig_index(bd, lin, t & 1, (t >> 1) & 1, (t >> 2) & 1)];
  }
  __syncthreads();
  double acc = 0.0;
  for (int e = threadIdx.x; e < R.ne; e += NT) {
    const long long i = R.e0 + e;
    const double v = st[e];
    if (init) {
      const double r = A.b[i] - v;
      A.r[i] = r;
      A.d[i] = r;
      acc = fma(r, r, acc);
    } else {
      A.q[i] = v;
      acc = fma(v, A.d[i], acc);
    }
  }
  const double sum = block_sum(acc, red);
  if (threadIdx.x == 0) A.part[tile] = sum;
}

// x += alpha d; unless last: r -= alpha q, partial |r|^2 (elementwise over
// the tile's contiguous range)
__global__ __launch_bounds__(NT) void k_pc_update(Batch B, TileTable tt, PcArgs A, int last) {
  __shared__ double red[NT / 64];
  const int tile = tile_of(tt);
  const int b = find_box(tt, tile);
  const BoxDesc bd = B.box[b];
  const PcRange R = pc_range(bd, tile - box_first_tile(tt, b));
  const double alpha = A.sc[4 * b + 1];
  double acc = 0.0;
  for (int e = threadIdx.x; e < R.ne; e += NT) {
    const long long i = R.e0 + e;
    A.x[i] = fma(alpha, A.d[i], A.x[i]);
    if (!last) {
      const double r = fma(-alpha, A.q[i], A.r[i]);
      A.r[i] = r;
      acc = fma(r, r, acc);
    }
  }
  if (!last) {
    const double sum = block_sum(acc, red);
    if (threadIdx.x == 0) A.part[tile] = sum;
  }
}

// x <- S x (scale_psi, OpenCL/Uncompress.c:377-470)
__global__ __launch_bounds__(NT) void k_pc_final(Batch B, TileTable tt, PcArgs A) {
  const int tile = tile_of(tt);
  const int b = find_box(tt, tile);
  const BoxDesc bd = B.box[b];
  const PcRange R = pc_range(bd, tile - box_first_tile(tt, b));
  for (int e = threadIdx.x; e < R.ne; e += NT)
    A.x[R.e0 + e] *= pc_scal(A, b, R.fine ? 1 + e % 7 : 0);
}

// one block per box: s = sum of the box's tile partials (fixed order).
// mode 0: ns = s; 1: alpha = ns / s; 2: beta = s / ns, ns = s.  A zero
// denominator gives 0 (the reference divides: 0/0 for a zero orbital).
__global__ __launch_bounds__(NT) void k_pc_scalars(TileTable tt, PcArgs A, int mode) {
  __shared__ double red[NT / 64];
  const int b = blockIdx.x;
  const int t0 = box_first_tile(tt, b);
  const int t1 = tt.uniform > 0 ? t0 + tt.uniform : tt.start[b + 1];
  double acc = 0.0;
  for (int t = t0 + (int)threadIdx.x; t < t1; t += NT) acc += A.part[t];
  const double s = block_sum(acc, red);
  if (threadIdx.x == 0) {
    double* sc = A.sc + 4 * b;
    if (mode == 0) {
      sc[0] = s;
    } else if (mode == 1) {
      sc[1] = s != 0.0 ? sc[0] / s : 0.0;
    } else {
      sc[2] = sc[0] != 0.0 ? s / sc[0] : 0.0;
      sc[0] = s;
    }
  }
}

hipError_t launch_precond(const Batch& B, const PassTables& PT, int maxN, long long wtotal,
                          bool need_zero, const PcArgs& A, hipStream_t s) {
  const size_t lds = lds_bytes(maxN, sizeof(double));
  const int nu = PT.ntiles[P_UNC];
  const TileTable& tu = PT.t[P_UNC];
  if (need_zero) {
    const long long nb = (wtotal + NT - 1) / NT;
    k_zero<double><<<(unsigned)(nb < 65536 ? nb : 65536), NT, 0, s>>>((double*)B.W, wtotal);
  }
  auto apply = [&](const double* src, int update) {
    if (nu > 0) k_pc_scatter<<<nu, NT, 0, s>>>(B, tu, A, src, update);
    if (PT.pc_plane) {
      // c W + K_x W + K_y W of each plane in one pass through LDS
      (void)launch_plane_kin_pc(B, PT.t[P_PF], PT.ntiles[P_PF], PT.lds_pc, A.pc, s);
    } else {
      if (PT.ntiles[P_XF] > 0) {
        if (PT.xbig) k_pc_kinx<true><<<PT.ntiles[P_XF], NT, lds, s>>>(B, PT.t[P_XF], A);
        else k_pc_kinx<false><<<PT.ntiles[P_XF], NT, lds, s>>>(B, PT.t[P_XF], A);
      }
      if (PT.ntiles[P_KY] > 0)
        k_kin<double, 1><<<PT.ntiles[P_KY], NT, lds, s>>>(B, PT.t[P_KY], PT.part[1]);
    }
    if (PT.ntiles[P_KZ] > 0)
      k_kin<double, 2><<<PT.ntiles[P_KZ], NT, lds, s>>>(B, PT.t[P_KZ], PT.part[2]);
  };
  if (nu == 0) return hipGetLastError();
  apply(A.x, 0);
  k_pc_gather<<<nu, NT, 0, s>>>(B, tu, A, 1);
  k_pc_scalars<<<B.nbox, NT, 0, s>>>(tu, A, 0);
  for (int it = 0; it < A.ncong; ++it) {
    apply(A.d, it > 0);
    k_pc_gather<<<nu, NT, 0, s>>>(B, tu, A, 0);
    k_pc_scalars<<<B.nbox, NT, 0, s>>>(tu, A, 1);
    const int last = it == A.ncong - 1;
    k_pc_update<<<nu, NT, 0, s>>>(B, tu, A, last);
    if (!last) k_pc_scalars<<<B.nbox, NT, 0, s>>>(tu, A, 2);
  }
  k_pc_final<<<nu, NT, 0, s>>>(B, tu, A);
  return hipGetLastError();
}

// the line kernels may take up to lds_bytes(MAXN, 8) of dynamic LDS
template <typename T>
static void line_attr(int bytes) {
  const void* k[] = {(const void*)k_xfwd<T, false>, (const void*)k_xfwd<T, true>,
                     (const void*)k_yfwd<T>, (const void*)k_kin<T, 1>, (const void*)k_kin<T, 2>,
                     (const void*)k_zpot<T, ZNT<T>::v, ZSHORT_L>,
                     (const void*)k_zpot<T, ZNT<T>::v, REG_L>, (const void*)k_zgrowpot<T>,
                     (const void*)k_zshrink<T>, (const void*)k_ybwd<T>,
                     (const void*)k_xbwd<T, false>, (const void*)k_xbwd<T, true>,
                     (const void*)k_zden<T, 1>, (const void*)k_zden<T, 2>};
  for (const void* f : k)
    if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, bytes) != hipSuccess)
      (void)hipGetLastError();
}
void line_attributes(size_t lds) {
  line_attr<double>((int)lds);
  line_attr<float>((int)lds);
  const void* k[] = {(const void*)k_pc_kinx<false>, (const void*)k_pc_kinx<true>};
  for (const void* f : k)
    if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) !=
        hipSuccess)
      (void)hipGetLastError();
}

}  // namespace bdft
