// Split-BF16 matrix-core kernels of the mixed-precision local-Hamiltonian
// apply (BDFT_PREC_BF16X3).
//
// Every 1D operator of the path along an axis -- G (synthesis o magic,
// combined filter fil2, growshrink.f90 comb_grow), G^T (comb_shrink) and the
// wavelet-basis kinetic (ConvolkineticT, unscaled) -- is a fixed matrix T
// (Dout x Din) for a given line length and boundary condition, with the
// periodic wrap or the zero padding built in.  The host builds T once per
// (operator, N, BC), pads it to 32-row blocks and 16-column K steps, splits
// every entry into bf16 hi + lo, and records per 32-row block which K steps
// hold non-zeros (the band).  A pass over a bundle of 32 lines stored in LDS
// as bf16 hi/lo [line][k] then computes, per 32-row output block,
//     D += T_hi X_hi + T_hi X_lo + T_lo X_hi      (FP32 accumulation)
// with v_mfma_f32_32x32x16_bf16.  Lane l of a wave holds D rows
// 8*(r/4) + 4*(l/32) + r%4 (r = 0..15) of line l%32, so results land with the
// line index across lanes, which is both the coalesced direction in memory
// and the k-contiguous layout the next pass along the other axis needs.
//
// Sequence: uncompress (FP32 W) -> kinetic x, y, z (K, ekin) -> x-y plane
// G_x G_y (T2) -> z lines G_z, V, G_z^T (in place, epot) -> x-y plane
// G_y^T G_x^T + K, compress-accumulate into hpsi -> energies.
#include "bdft_internal.h"

namespace bdft {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int NTM = 256;    // threads of the pencil kernels (4 waves)
constexpr int NTMP = 512;   // threads of the plane kernels (8 waves)

__device__ __forceinline__ void split2(float x, __bf16& hi, __bf16& lo) {
  hi = (__bf16)x;
  lo = (__bf16)(x - (float)hi);
}

// D row held by lane in accumulator register r
__device__ __forceinline__ int drow(int r, int lane) {
  return 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
}

// acc += T[m0:m0+32, :] X[:, line bundle], X in LDS as [line*XS + k].
// T = T0 + T1 + T2 and X = X0 + X1 (bf16 pieces); the products kept are
// T0 X0 + T0 X1 + T1 X0 (NTERM 3) and also T2 X0 (NTERM 4), which removes
// the fixed representation error of the filter (2^-16 relative).
template <int NTERM>
__device__ __forceinline__ void mfma_block(const Batch& B, const OpTab& t, int m0,
                                           const __bf16* __restrict__ xh,
                                           const __bf16* __restrict__ xl, int XS,
                                           f32x16& acc) {
  const int lane = threadIdx.x & 63;
  const int kh = 8 * (lane >> 5);
  const long long ro = t.off + (long long)(m0 + (lane & 31)) * t.Kp + kh;
  const __bf16* th = (const __bf16*)B.thi + ro;
  const __bf16* tl = (const __bf16*)B.tlo + ro;
  const __bf16* t2 = (const __bf16*)B.tlo2 + ro;
  const __bf16* bh = xh + (lane & 31) * XS + kh;
  const __bf16* bl = xl + (lane & 31) * XS + kh;
  unsigned mask = B.tmask[t.moff + (m0 >> 5)];
  while (mask) {
    const int s = __builtin_ctz(mask);
    mask &= mask - 1;
    const bf16x8 ah = *(const bf16x8*)(th + 16 * s);
    const bf16x8 al = *(const bf16x8*)(tl + 16 * s);
    const bf16x8 xh8 = *(const bf16x8*)(bh + 16 * s);
    const bf16x8 xl8 = *(const bf16x8*)(bl + 16 * s);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, xh8, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, xl8, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, xh8, acc, 0, 0, 0);
    if (NTERM == 4) {
      const bf16x8 a2 = *(const bf16x8*)(t2 + 16 * s);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a2, xh8, acc, 0, 0, 0);
    }
  }
}

__device__ __forceinline__ void zero_lds(unsigned* p, int nwords) {
  for (int i = threadIdx.x; i < nwords; i += blockDim.x) p[i] = 0u;
}

__device__ __forceinline__ double wave_block_sum(double v, double* red, int nwaves) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  const int w = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < nwaves; ++i) s += red[i];
  return s;
}

__device__ __forceinline__ int rup(int a, int m) { return (a + m - 1) / m * m; }

// ---------------------------------------------------------------------------
// kinetic along AX on the wavelet volume: K (=, AX 0) or (+=) Ktilde_AX W
// ---------------------------------------------------------------------------
template <int AX, int NT4>
__global__ __launch_bounds__(NTM) void k_mkin(Batch B, TileTable tt, double* part) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ double red[NTM / 64];
  const int tile = blockIdx.x;
  const int b = find_box(tt, tile);
  const BoxDesc bd = B.box[b];
  const int lt = tile - box_first_tile(tt, b);
  const int W1 = 2 * bd.N[0], W2 = 2 * bd.N[1], W3 = 2 * bd.N[2];
  const int Wd = AX == 0 ? W1 : (AX == 1 ? W2 : W3);
  const OpTab t = B.tabs[bd.opK[AX]];
  const int XS = t.Kp + 8;
  __bf16* xh = (__bf16*)smem;
  __bf16* xl = xh + 32 * XS;
  float* xf = (float*)(xl + 32 * XS);  // FP32 copy [line][k], row stride Wd+1
  const int XF = Wd + 1;
  // line bundle
  long long base, lstride, pstride;
  int nl;
  if (AX == 0) {
    const int nlines = W2 * W3;
    const int l0 = lt * 32;
    nl = min(32, nlines - l0);
    base = bd.w_off + (long long)l0 * W1;
    lstride = W1;
    pstride = 1;
  } else {
    const int nxt = (W1 + 31) / 32;
    const int jo = lt / nxt;
    const int j1 = (lt - jo * nxt) * 32;
    nl = min(32, W1 - j1);
    base = bd.w_off + j1 + (AX == 1 ? (long long)W1 * W2 * jo : (long long)W1 * jo);
    lstride = 1;
    pstride = AX == 1 ? (long long)W1 : (long long)W1 * W2;
  }
  zero_lds((unsigned*)smem, 32 * XS);  // hi and lo planes (2 bytes each)
  __syncthreads();
  const float* W = (const float*)B.W;
  const int tot = 32 * Wd;
  for (int i0 = threadIdx.x; i0 < tot; i0 += NTM * 8) {
    float v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int i = i0 + u * NTM;
      int l, k;
      if (AX == 0) { l = i / Wd; k = i - l * Wd; } else { k = i >> 5; l = i & 31; }
      v[u] = (i < tot && l < nl) ? W[base + l * lstride + k * pstride] : 0.0f;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int i = i0 + u * NTM;
      if (i < tot) {
        int l, k;
        if (AX == 0) { l = i / Wd; k = i - l * Wd; } else { k = i >> 5; l = i & 31; }
        __bf16 h, lo;
        split2(v[u], h, lo);
        xh[l * XS + k] = h;
        xl[l * XS + k] = lo;
        xf[l * XF + k] = v[u];
      }
    }
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const float ks = (float)bd.kscale[AX];
  float* K = (float*)B.K;
  double ek = 0.0;
  const int line = lane & 31;
  for (int ob = wv; ob < t.Dp / 32; ob += NTM / 64) {
    f32x16 acc = {};
    mfma_block<NT4>(B, t, 32 * ob, xh, xl, XS, acc);
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int o = 32 * ob + drow(r, lane);
      if (o < Wd && line < nl) {
        const float v = acc[r] * ks;
        ek += (double)v * (double)xf[line * XF + o];
        const long long a = base + line * lstride + o * pstride;
        if (AX == 0) K[a] = v;
        else K[a] += v;
      }
    }
  }
  const double eks = wave_block_sum(ek, red, NTM / 64);
  if (threadIdx.x == 0) part[tile] = eks;
}

// ---------------------------------------------------------------------------
// x-y plane: T2 = G_y G_x W (plane j3)
// ---------------------------------------------------------------------------
template <int NT4>
__global__ __launch_bounds__(NTMP) void k_mgrow_plane(Batch B, TileTable tt) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int tile = blockIdx.x;
  const int b = find_box(tt, tile);
  const BoxDesc bd = B.box[b];
  const int j3 = tile - box_first_tile(tt, b);
  const int W1 = 2 * bd.N[0], W2 = 2 * bd.N[1], L1 = bd.L[0], L2 = bd.L[1];
  const OpTab tx = B.tabs[bd.opG[0]], ty = B.tabs[bd.opG[1]];
  const int XS1 = tx.Kp + 8, XS2 = ty.Kp + 8;
  const int n1 = rup(W2, 32), n2 = rup(L1, 32);  // lines of each buffer
  __bf16* rh = (__bf16*)smem;  // W plane rows j2: [j2*XS1 + j1]
  __bf16* rl = rh + n1 * XS1;
  __bf16* ch = rl + n1 * XS1;  // G_x output columns p1: [p1*XS2 + j2]
  __bf16* cl = ch + n2 * XS2;
  zero_lds((unsigned*)smem, n1 * XS1 + n2 * XS2);
  __syncthreads();
  const float* src = (const float*)B.W + bd.w_off + (long long)W1 * W2 * j3;
  const int tot = W1 * W2;
  for (int i0 = threadIdx.x; i0 < tot; i0 += NTMP * 8) {
    float v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int i = i0 + u * NTMP;
      v[u] = i < tot ? src[i] : 0.0f;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int i = i0 + u * NTMP;
      if (i < tot) {
        const int j2 = i / W1, j1 = i - j2 * W1;
        __bf16 h, lo;
        split2(v[u], h, lo);
        rh[j2 * XS1 + j1] = h;
        rl[j2 * XS1 + j1] = lo;
      }
    }
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  // G_x: lines j2, outputs p1
  {
    const int nob = tx.Dp / 32, nlb = n1 / 32;
    for (int it = wv; it < nob * nlb; it += NTMP / 64) {
      const int ob = it % nob, lb = it / nob;
      f32x16 acc = {};
      mfma_block<NT4>(B, tx, 32 * ob, rh + lb * 32 * XS1, rl + lb * 32 * XS1, XS1, acc);
      const int j2 = 32 * lb + (lane & 31);
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int p1 = 32 * ob + drow(r, lane);
        if (p1 < L1 && j2 < W2) {
          __bf16 h, lo;
          split2(acc[r], h, lo);
          ch[p1 * XS2 + j2] = h;
          cl[p1 * XS2 + j2] = lo;
        }
      }
    }
  }
  __syncthreads();
  // G_y: lines p1, outputs p2 -> T2 plane
  float* dst = (float*)B.T2 + bd.t2_off + (long long)L1 * L2 * j3;
  {
    const int nob = ty.Dp / 32, nlb = n2 / 32;
    for (int it = wv; it < nob * nlb; it += NTMP / 64) {
      const int ob = it % nob, lb = it / nob;
      f32x16 acc = {};
      mfma_block<NT4>(B, ty, 32 * ob, ch + lb * 32 * XS2, cl + lb * 32 * XS2, XS2, acc);
      const int p1 = 32 * lb + (lane & 31);
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int p2 = 32 * ob + drow(r, lane);
        if (p1 < L1 && p2 < L2) dst[p1 + (long long)L1 * p2] = acc[r];
      }
    }
  }
}

// ---------------------------------------------------------------------------
// z lines of T2: G_z, V, epot, G_z^T, in place
// ---------------------------------------------------------------------------
template <int NT4>
__global__ __launch_bounds__(NTM) void k_mzpot(Batch B, TileTable tt, double* part) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ double red[NTM / 64];
  const int tile = blockIdx.x;
  const int b = find_box(tt, tile);
  const BoxDesc bd = B.box[b];
  const int lt = tile - box_first_tile(tt, b);
  const int L1 = bd.L[0], L2 = bd.L[1], W3 = 2 * bd.N[2], L3 = bd.L[2];
  const OpTab tg = B.tabs[bd.opG[2]], tsh = B.tabs[bd.opS[2]];
  const int XS = tg.Kp + 8, RS = tsh.Kp + 8;
  const int nxt = (L1 + 31) / 32;
  const int j2 = lt / nxt;
  const int j1 = (lt - j2 * nxt) * 32;
  const int nl = min(32, L1 - j1);
  const long long pstride = (long long)L1 * L2;
  float* line0 = (float*)B.T2 + bd.t2_off + j1 + (long long)L1 * j2;
  __bf16* xh = (__bf16*)smem;
  __bf16* xl = xh + 32 * XS;
  __bf16* rh = xl + 32 * XS;
  __bf16* rl = rh + 32 * RS;
  zero_lds((unsigned*)smem, 32 * XS + 32 * RS);
  __syncthreads();
  const int tot = 32 * W3;
  for (int i0 = threadIdx.x; i0 < tot; i0 += NTM * 8) {
    float v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int i = i0 + u * NTM;
      const int k = i >> 5, l = i & 31;
      v[u] = (i < tot && l < nl) ? line0[l + pstride * k] : 0.0f;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int i = i0 + u * NTM;
      if (i < tot) {
        const int k = i >> 5, l = i & 31;
        __bf16 h, lo;
        split2(v[u], h, lo);
        xh[l * XS + k] = h;
        xl[l * XS + k] = lo;
      }
    }
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int line = lane & 31;
  const int q1 = j1 + line - bd.ishift[0], q2 = j2 - bd.ishift[1];
  const bool in12 = line < nl && q1 >= 0 && q1 < bd.potdim[0] && q2 >= 0 &&
                    q2 < bd.potdim[1];
  const double* pot = B.pot + bd.pot_off + q1 + (long long)bd.potdim[0] * q2;
  const long long pz = (long long)bd.potdim[0] * bd.potdim[1];
  double ep = 0.0;
  for (int ob = wv; ob < tg.Dp / 32; ob += NTM / 64) {
    f32x16 acc = {};
    mfma_block<NT4>(B, tg, 32 * ob, xh, xl, XS, acc);
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int p3 = 32 * ob + drow(r, lane);
      if (p3 < L3) {
        const int q3 = p3 - bd.ishift[2];
        double v = 0.0;
        if (in12 && q3 >= 0 && q3 < bd.potdim[2]) v = pot[pz * q3];
        const double pr = (double)acc[r];
        ep = fma(v * pr, pr, ep);
        __bf16 h, lo;
        split2((float)(v * pr), h, lo);
        rh[line * RS + p3] = h;
        rl[line * RS + p3] = lo;
      }
    }
  }
  const double eps = wave_block_sum(ep, red, NTM / 64);
  if (threadIdx.x == 0) part[tile] = eps;
  __syncthreads();
  for (int ob = wv; ob < tsh.Dp / 32; ob += NTM / 64) {
    f32x16 acc = {};
    mfma_block<NT4>(B, tsh, 32 * ob, rh, rl, RS, acc);
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int o = 32 * ob + drow(r, lane);
      if (o < W3 && line < nl) line0[line + pstride * o] = acc[r];
    }
  }
}

// ---------------------------------------------------------------------------
// x-y plane: out = G_x^T G_y^T T2 + K, compress-accumulate into hpsi
// ---------------------------------------------------------------------------
template <int NT4>
__global__ __launch_bounds__(NTMP) void k_mshrink_plane(Batch B, TileTable tt) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int tile = blockIdx.x;
  const int b = find_box(tt, tile);
  const BoxDesc bd = B.box[b];
  const int j3 = tile - box_first_tile(tt, b);
  const int N1 = bd.N[0], N2 = bd.N[1], N3 = bd.N[2], L1 = bd.L[0], L2 = bd.L[1];
  const int W1 = 2 * N1, W2 = 2 * N2;
  const OpTab sy = B.tabs[bd.opS[1]], sx = B.tabs[bd.opS[0]];
  const int XU = sy.Kp + 8, XB = sx.Kp + 8;
  const int nu = rup(L1, 32), nb = rup(W2, 32);
  const int ureg = max(2 * nu * XU, 2 * W2 * (W1 + 1));  // bf16 units
  __bf16* uh = (__bf16*)smem;  // T2 plane columns p1: [p1*XU + p2]
  __bf16* ul = uh + nu * XU;
  __bf16* bh = uh + ureg;      // G_y^T output rows j2: [j2*XB + p1]
  __bf16* bl = bh + nb * XB;
  const int RO = W1 + 1;
  float* O = (float*)smem;     // output plane [j2*RO + j1], reuses the U planes
  zero_lds((unsigned*)smem, ureg / 2 + nb * XB);
  __syncthreads();
  const float* src = (const float*)B.T2 + bd.t2_off + (long long)L1 * L2 * j3;
  const int tot = L1 * L2;
  for (int i0 = threadIdx.x; i0 < tot; i0 += NTMP * 8) {
    float v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int i = i0 + u * NTMP;
      v[u] = i < tot ? src[i] : 0.0f;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int i = i0 + u * NTMP;
      if (i < tot) {
        const int p2 = i / L1, p1 = i - p2 * L1;
        __bf16 h, lo;
        split2(v[u], h, lo);
        uh[p1 * XU + p2] = h;
        ul[p1 * XU + p2] = lo;
      }
    }
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  {  // G_y^T: lines p1, outputs j2
    const int nob = sy.Dp / 32, nlb = nu / 32;
    for (int it = wv; it < nob * nlb; it += NTMP / 64) {
      const int ob = it % nob, lb = it / nob;
      f32x16 acc = {};
      mfma_block<NT4>(B, sy, 32 * ob, uh + lb * 32 * XU, ul + lb * 32 * XU, XU, acc);
      const int p1 = 32 * lb + (lane & 31);
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int o2 = 32 * ob + drow(r, lane);
        if (o2 < W2 && p1 < L1) {
          __bf16 h, lo;
          split2(acc[r], h, lo);
          bh[o2 * XB + p1] = h;
          bl[o2 * XB + p1] = lo;
        }
      }
    }
  }
  __syncthreads();
  {  // G_x^T: lines j2, outputs j1 -> O
    const int nob = sx.Dp / 32, nlb = nb / 32;
    for (int it = wv; it < nob * nlb; it += NTMP / 64) {
      const int ob = it % nob, lb = it / nob;
      f32x16 acc = {};
      mfma_block<NT4>(B, sx, 32 * ob, bh + lb * 32 * XB, bl + lb * 32 * XB, XB, acc);
      const int j2 = 32 * lb + (lane & 31);
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int o1 = 32 * ob + drow(r, lane);
        if (o1 < W1 && j2 < W2) O[j2 * RO + o1] = acc[r];
      }
    }
  }
  __syncthreads();
  // compress_and_accumulate_mixed (un-compress.f90:2174-2238) of the plane's
  // four components (s3 fixed by j3)
  const int s3 = j3 >= N3;
  const int i3 = j3 - s3 * N3;
  const float* Kp = (const float*)B.K + bd.w_off + (long long)W1 * W2 * j3;
  const int npl = N1 * N2;
  double* hc = B.hpsi + bd.psi_off;
  double* hf = B.hpsi + bd.psi_off + bd.nvctr_c;
  for (int i = threadIdx.x; i < npl; i += NTMP) {
    const int i2 = i / N1, i1 = i - i2 * N1;
    const int lin = i + npl * i3;
    const int ci = bd.cinv ? bd.cinv[lin] : lin;
    const int fi = bd.finv ? bd.finv[lin] : lin;
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const int s1 = s & 1, s2 = s >> 1;
      const int j1 = i1 + N1 * s1, jj2 = i2 + N2 * s2;
      const int t = s1 + 2 * s2 + 4 * s3;
      const double v = (double)O[jj2 * RO + j1] + (double)Kp[j1 + W1 * jj2];
      if (t == 0) {
        if (ci >= 0) hc[ci] += v;
      } else if (fi >= 0) {
        hf[7LL * fi + t - 1] += v;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// LDS sizes and launcher
// ---------------------------------------------------------------------------
static inline int rup_h(int a, int m) { return (a + m - 1) / m * m; }

size_t mfma_lds_kin(int Wd, int Kp) {
  return (size_t)2 * 32 * (Kp + 8) * 2 + (size_t)32 * (Wd + 1) * 4;
}
size_t mfma_lds_grow(int W2, int L1, int Kp1, int Kp2) {
  return (size_t)2 * 2 * (rup_h(W2, 32) * (Kp1 + 8) + rup_h(L1, 32) * (Kp2 + 8));
}
size_t mfma_lds_zpot(int KpG, int KpS) {
  return (size_t)2 * 2 * 32 * ((KpG + 8) + (KpS + 8));
}
size_t mfma_lds_shrink(int L1, int W1, int W2, int KpU, int KpB) {
  const size_t u = (size_t)2 * 2 * rup_h(L1, 32) * (KpU + 8);
  const size_t o = (size_t)4 * W2 * (W1 + 1);
  return (u > o ? u : o) + (size_t)2 * 2 * rup_h(W2, 32) * (KpB + 8);
}

template <int NT4>
static void set_attr(size_t lds) {
  const void* k[] = {(const void*)k_mkin<0, NT4>, (const void*)k_mkin<1, NT4>,
                     (const void*)k_mkin<2, NT4>, (const void*)k_mgrow_plane<NT4>,
                     (const void*)k_mzpot<NT4>, (const void*)k_mshrink_plane<NT4>};
  for (const void* f : k)
    if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) !=
        hipSuccess)
      (void)hipGetLastError();
}

size_t mfma_attributes(size_t lds) {
  set_attr<3>(lds);
  set_attr<4>(lds);
  return lds;
}

template <int NT4>
static hipError_t launch_mfma_t(const Batch& B, const PassTables& PT, const MfmaLds& L,
                                long long wtotal, bool need_zero, hipStream_t s,
                                hipEvent_t* ev) {
  auto mark = [&](int stage, int end) {
    if (ev) (void)hipEventRecord(ev[2 * stage + end], s);
  };
  mark(S_ZERO, 0);
  if (need_zero) {
    const long long nb = (wtotal + 255) / 256;
    extern_zero_float((float*)B.W, wtotal, (int)(nb < 65536 ? nb : 65536), s);
  }
  mark(S_ZERO, 1);
  mark(S_UNC, 0);
  if (PT.ntiles[P_UNC] > 0) launch_uncompress_float(B, PT.t[P_UNC], PT.ntiles[P_UNC], s);
  mark(S_UNC, 1);
  mark(S_XF, 0);
  if (PT.ntiles[P_XF] > 0)
    k_mkin<0, NT4><<<PT.ntiles[P_XF], NTM, L.kin[0], s>>>(B, PT.t[P_XF], PT.part[0]);
  mark(S_XF, 1);
  mark(S_KY, 0);
  if (PT.ntiles[P_KY] > 0)
    k_mkin<1, NT4><<<PT.ntiles[P_KY], NTM, L.kin[1], s>>>(B, PT.t[P_KY], PT.part[1]);
  mark(S_KY, 1);
  mark(S_KZ, 0);
  if (PT.ntiles[P_KZ] > 0)
    k_mkin<2, NT4><<<PT.ntiles[P_KZ], NTM, L.kin[2], s>>>(B, PT.t[P_KZ], PT.part[2]);
  mark(S_KZ, 1);
  mark(S_PF, 0);
  if (PT.ntiles[P_PF] > 0)
    k_mgrow_plane<NT4><<<PT.ntiles[P_PF], NTMP, L.grow, s>>>(B, PT.t[P_PF]);
  mark(S_PF, 1);
  mark(S_ZP, 0);
  if (PT.ntiles[P_ZP] > 0)
    k_mzpot<NT4><<<PT.ntiles[P_ZP], NTM, L.zpot, s>>>(B, PT.t[P_ZP], PT.part[3]);
  mark(S_ZP, 1);
  mark(S_PB, 0);
  if (PT.ntiles[P_PB] > 0)
    k_mshrink_plane<NT4><<<PT.ntiles[P_PB], NTMP, L.shrink, s>>>(B, PT.t[P_PB]);
  mark(S_PB, 1);
  EnergyTables et;
  et.t[0] = PT.t[P_XF];
  et.p[0] = PT.part[0];
  et.t[1] = PT.t[P_KY];
  et.p[1] = PT.part[1];
  et.t[2] = PT.t[P_KZ];
  et.p[2] = PT.part[2];
  et.t[3] = PT.t[P_ZP];
  et.p[3] = PT.part[3];
  et.nk = 3;
  et.n = 4;
  mark(S_ENERGY, 0);
  launch_energy(et, B, s);
  mark(S_ENERGY, 1);
  if (ev) {
    const int unused[] = {S_YF, S_YB, S_XB, S_COMP};
    for (int st : unused) {
      mark(st, 0);
      mark(st, 1);
    }
  }
  return hipGetLastError();
}

hipError_t launch_mfma(const Batch& B, const PassTables& PT, const MfmaLds& L,
                       long long wtotal, bool need_zero, int nterm, hipStream_t s,
                       hipEvent_t* ev) {
  return nterm == 4 ? launch_mfma_t<4>(B, PT, L, wtotal, need_zero, s, ev)
                    : launch_mfma_t<3>(B, PT, L, wtotal, need_zero, s, ev);
}

}  // namespace bdft
