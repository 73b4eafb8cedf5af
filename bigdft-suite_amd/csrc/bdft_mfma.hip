// Matrix-core kernels of the mixed-precision local-Hamiltonian apply
// (BDFT_PREC_BF16X3, BF16X4, F16X3).
//
// Every 1D operator of the path along an axis -- G (synthesis o magic,
// combined filter fil2, growshrink.f90 comb_grow), G^T (comb_shrink) and the
// wavelet-basis kinetic (ConvolkineticT, unscaled) -- is a fixed matrix T
// (Dout x Din) for a given line length and boundary condition, with the
// periodic wrap or the zero padding built in.  The host builds T once per
// (operator, N, BC, piece format), splits every entry into 16-bit pieces and
// stores, per 32-row block, only the 16-column K steps that hold non-zeros,
// in matrix-core operand order.  A pass over bundles of 32 lines held in LDS
// as 16-bit pieces [line][k] computes, per 32-row output block,
//     D += T0 X0 + T0 X1 + T1 X0 (+ T2 X0)      (FP32 accumulation)
// with v_mfma_f32_32x32x16_{bf16,f16}.  Lane l of a wave holds D rows
// 8*(r/4) + 4*(l/32) + r%4 (r = 0..15) of column l%32; with the table as the
// A operand the columns are lines, with the table as the B operand they are
// outputs, which lets each pass choose where its results land (coalesced
// along x in memory, or k-contiguous for the next pass in LDS).
//
// Plane-fused sequence: uncompress (FP32 W) -> k_mfwd (x-y plane: G_x, G_y
// -> T2; x+y kinetic -> K) -> k_mkz (z kinetic, ekin) -> k_mzp (z lines:
// G_z, V, epot, G_z^T) -> k_mbwd (x-y plane: G_x^T, G_y^T, + K, compress
// into hpsi) -> energies.  A per-pass sequence covers boxes whose planes do
// not fit in LDS.
#include "bdft_internal.h"
#include "../../include/bigdft_hip.h"

namespace bdft {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));


constexpr int NTM = 256;    // threads of the pencil kernels (4 waves)
constexpr int NTMP = 512;   // threads of the plane kernels (8 waves)

typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));

// Piece formats of the matrix-core passes (template argument M):
//   3  : bf16 pieces, products T0 X0 + T0 X1 + T1 X0
//   4  : bf16 pieces, and T2 X0 as well (filter kept to its third piece)
//   16 : fp16 pieces of data scaled by a power of two per box (largest
//        |value| below 2^12), products T0 X0 + T0 X1 + T1 X0; pieces of
//        11 bits make each value and filter entry exact to 2^-22 relative
//        (bf16: 2^-17).  fp16 values live in the same 16-bit LDS slots.
template <int M>
__device__ __forceinline__ void split2(float x, __bf16& hi, __bf16& lo) {
  if (M == 16) {
    const _Float16 h = (_Float16)x;
    const _Float16 l = (_Float16)(x - (float)h);
    hi = __builtin_bit_cast(__bf16, h);
    lo = __builtin_bit_cast(__bf16, l);
  } else {
    hi = (__bf16)x;
    lo = (__bf16)(x - (float)hi);
  }
}

template <int M>
__device__ __forceinline__ f32x16 mf(const bf16x8& a, const bf16x8& b, const f32x16& c) {
  if (M == 16)
    return __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(f16x8, a),
                                                  __builtin_bit_cast(f16x8, b), c, 0, 0, 0);
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

// scale of the data split into fp16 pieces: 2^(12 - e) with the box's
// largest |psi| below 2^e (the grids derived from it stay within a few times
// that); 1 for the bf16 formats
template <int M>
__device__ __forceinline__ float box_scale(const Batch& B, int b) {
  if (M != 16) return 1.0f;
  const float mx = __uint_as_float(B.bmax[b]);
  if (!(mx > 0.0f)) return 1.0f;
  int e;
  (void)frexpf(mx, &e);
  e = min(max(e, -100), 100);
  return ldexpf(1.0f, 12 - e);
}

// factor that removes the fp16 kinetic table's scale (KIN16) from a product
template <int M>
__device__ __forceinline__ float kin_unscale() {
  return M == 16 ? (float)(1.0 / KIN16) : 1.0f;
}
// inverse row scale u_r of a G table (fp16 tables scale each row, see
// OpTab::soff; 1 for unscaled tables and padding rows)
__device__ __forceinline__ float grow_u(const Batch& B, const OpTab& t, int r) {
  return (t.soff >= 0 && r < t.rows) ? B.trs[t.soff + r] : 1.0f;
}

// D row held by lane in accumulator register r
__device__ __forceinline__ int drow(int r, int lane) {
  return 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
}

typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

// One 32x32 output tile of an operator T (table rows o0..o0+31, all K steps
// flagged in the row block's mask) applied to a bundle of 32 lines held in
// LDS as bf16 pieces [line * XS + k].
//   TABB false: D = T X      -> D[m = output o][n = line]
//   TABB true : D = X^T T^T  -> D[m = line][n = output o]
// Both read the same fragments (lane l: row l%32, k 8(l/32)..+8 of the
// step); only the operand slots differ, which picks the lane <-> index map
// of the result (lane l, register r: m = drow(r, l), n = l % 32).
// T = T0 + T1 + T2 and X = X0 + X1 (bf16 pieces); the products kept are
// T0 X0 + T0 X1 + T1 X0 (NTERM 3) and also T2 X0 (NTERM 4), which removes
// the fixed representation error of the filter (2^-16 relative).
// one K step: acc[j] += (fragments f) x (bundle j's columns 16s..16s+15)
template <int NTERM, bool TABB, int NBMAX>
__device__ __forceinline__ void mma_step(const bf16x8 (&f)[3], const __bf16* bh,
                                         const __bf16* bl, int LS, int s, int nb,
                                         f32x16* acc) {
#pragma unroll
  for (int j = 0; j < NBMAX; ++j) {
    if (j < nb) {
      const bf16x8 xh8 = *(const bf16x8*)(bh + j * LS + 16 * s);
      const bf16x8 xl8 = *(const bf16x8*)(bl + j * LS + 16 * s);
      if (TABB) {
        acc[j] = mf<NTERM>(xh8, f[0], acc[j]);
        acc[j] = mf<NTERM>(xl8, f[0], acc[j]);
        acc[j] = mf<NTERM>(xh8, f[1], acc[j]);
        if (NTERM == 4) acc[j] = mf<NTERM>(xh8, f[2], acc[j]);
      } else {
        acc[j] = mf<NTERM>(f[0], xh8, acc[j]);
        acc[j] = mf<NTERM>(f[0], xl8, acc[j]);
        acc[j] = mf<NTERM>(f[1], xh8, acc[j]);
        if (NTERM == 4) acc[j] = mf<NTERM>(f[2], xh8, acc[j]);
      }
    }
  }
}

template <int NTERM>
__device__ __forceinline__ void load_frag(const bf16x8* fp, bf16x8 (&f)[3]) {
  f[0] = fp[0];
  f[1] = fp[64];
  if (NTERM == 4) f[2] = fp[128];
}

// Sweep of the K steps of one 32-row block for nb <= NBMAX bundles; the
// fragments of the next step are loaded while the current step computes
// (two register sets used in turn, so no copy waits on a load).
template <int NTERM, bool TABB, int NBMAX>
__device__ __forceinline__ void mma_sweep(const Batch& B, const OpTab& t, int o0,
                                          const __bf16* __restrict__ xh,
                                          const __bf16* __restrict__ xl, int XS, int LS,
                                          int nb, f32x16* acc) {
  const int lane = threadIdx.x & 63;
  const int kh = 8 * (lane >> 5);
  const uint2 mq = B.tmq[t.moff + (o0 >> 5)];
  unsigned mask = mq.x;
  if (!mask) return;
  const bf16x8* fp = (const bf16x8*)B.tfrag + (size_t)mq.y * 192 + lane;
  const __bf16* bh = xh + (lane & 31) * XS + kh;
  const __bf16* bl = xl + (lane & 31) * XS + kh;
  bf16x8 fa[3], fb[3];
  load_frag<NTERM>(fp, fa);
  while (true) {
    int s = __builtin_ctz(mask);
    mask &= mask - 1;
    if (mask) load_frag<NTERM>(fp += 192, fb);
    mma_step<NTERM, TABB, NBMAX>(fa, bh, bl, LS, s, nb, acc);
    if (!mask) break;
    s = __builtin_ctz(mask);
    mask &= mask - 1;
    if (mask) load_frag<NTERM>(fp += 192, fa);
    mma_step<NTERM, TABB, NBMAX>(fb, bh, bl, LS, s, nb, acc);
    if (!mask) break;
  }
}

template <int NTERM, bool TABB>
__device__ __forceinline__ void mma(const Batch& B, const OpTab& t, int o0,
                                    const __bf16* __restrict__ xh,
                                    const __bf16* __restrict__ xl, int XS, f32x16& acc) {
  mma_sweep<NTERM, TABB, 1>(B, t, o0, xh, xl, XS, 0, 1, &acc);
}

// D = T X with X read from a row-major plane: B[k][n] = X[k * XS + n0 + n]
// (the K steps run down the rows), so an operator along the rows needs no
// transposed copy; lane l gathers rows 8(l/32)..+8 of the step in column
// n0 + l%32.  Same fragments, products and orientation as mma<NTERM, false>.
template <int NTERM>
__device__ __forceinline__ void mma_rows(const Batch& B, const OpTab& t, int o0,
                                         const __bf16* __restrict__ xh,
                                         const __bf16* __restrict__ xl, int XS, int n0,
                                         f32x16& acc) {
  const int lane = threadIdx.x & 63;
  const uint2 mq = B.tmq[t.moff + (o0 >> 5)];
  unsigned mask = mq.x;
  const bf16x8* fp = (const bf16x8*)B.tfrag + (size_t)mq.y * 192 + lane;
  const int c0 = 8 * (lane >> 5) * XS + n0 + (lane & 31);
  while (mask) {
    const int s = __builtin_ctz(mask);
    mask &= mask - 1;
    bf16x8 f[3];
    load_frag<NTERM>(fp, f);
    fp += 192;
    const __bf16* ph = xh + c0 + 16 * s * XS;
    const __bf16* pl = xl + c0 + 16 * s * XS;
    bf16x8 h8, l8;
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      h8[u] = ph[u * XS];
      l8[u] = pl[u * XS];
    }
    acc = mf<NTERM>(f[0], h8, acc);
    acc = mf<NTERM>(f[0], l8, acc);
    acc = mf<NTERM>(f[1], h8, acc);
    if (NTERM == 4) acc = mf<NTERM>(f[2], h8, acc);
  }
}

template <int NTERM>
__device__ __forceinline__ void mfma_block(const Batch& B, const OpTab& t, int m0,
                                           const __bf16* __restrict__ xh,
                                           const __bf16* __restrict__ xl, int XS,
                                           f32x16& acc) {
  mma<NTERM, false>(B, t, m0, xh, xl, XS, acc);
}

// Bundles per sweep of the forward plane kernel's G_y phase: 2 (one
// fragment load feeds two bundles) except with the fourth product term,
// where the extra fragment registers would leave room for only one
// work-group per CU.
template <int NTERM>
constexpr int mbf() { return NTERM == 4 ? 1 : 2; }
constexpr int MBB = 2;  // bundles per sweep: backward plane kernel

// The same product for up to NBMAX line bundles at once (bundle j at
// xh + j*LS), so that each table fragment is fetched once per K step for all
// of them: acc[j] += T[o0..o0+31, :] X_j (orientation as in mma).
template <int NTERM, bool TABB, int NBMAX>
__device__ __forceinline__ void mma_multi(const Batch& B, const OpTab& t, int o0,
                                          const __bf16* __restrict__ xh,
                                          const __bf16* __restrict__ xl, int XS, int LS,
                                          int nb, f32x16 (&acc)[NBMAX]) {
  mma_sweep<NTERM, TABB, NBMAX>(B, t, o0, xh, xl, XS, LS, nb, acc);
}

// Store the 16 results of a lane (one column n, rows m in four groups of four
// consecutive m) into a bf16 hi/lo buffer laid out [n * XS + m], 8 bytes per
// group and piece.
template <int M>
__device__ __forceinline__ void store_cols(const f32x16& acc, __bf16* __restrict__ yh,
                                           __bf16* __restrict__ yl, int XS, int n, int m0) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int g = 0; g < 4; ++g) {
    bf16x4 h, l;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      __bf16 a, b;
      split2<M>(acc[4 * g + u], a, b);
      h[u] = a;
      l[u] = b;
    }
    const int o = n * XS + m0 + 8 * g + 4 * (lane >> 5);
    *(bf16x4*)(yh + o) = h;
    *(bf16x4*)(yl + o) = l;
  }
}

// Work split of a phase: nunits (operator, output block) units over nlb line
// bundles, cut into chunks so that about nw waves have work; task q covers
// bundles [lb_begin, lb_end) of unit u.
struct Task {
  int u, lb_begin, lb_end;
};
__device__ __forceinline__ int task_chunks(int nunits, int nlb, int nw) {
  const int c = nw / max(nunits, 1);
  return max(1, min(nlb, c));
}
__device__ __forceinline__ Task task_of(int q, int chunks, int nlb) {
  const int cs = (nlb + chunks - 1) / chunks;
  Task t;
  t.u = q / chunks;
  const int c = q - t.u * chunks;
  t.lb_begin = c * cs;
  t.lb_end = min(nlb, t.lb_begin + cs);
  return t;
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
// kinetic along AX on the wavelet volume: K (=, AX 0) or (+=) Ktilde_AX W.
// The per-tile partial of ekin is sum W * (Ktilde_AX W) or, with TOTAL, the
// sum of W times the complete K after this pass (the kinetic along the other
// axes already accumulated in K), i.e. the whole <psi|T psi> of the tile.
// x lines take the table as the B operand so that lanes run along x, the
// contiguous direction of K; y and z lines run along x across lanes anyway.
// ---------------------------------------------------------------------------
template <int AX, int NT4, bool TOTAL>
__global__ __launch_bounds__(NTM) void k_mkin(Batch B, TileTable tt, double* part) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ double red[NTM / 64];
  const int tile = tile_of(tt);
  const int b = find_box(tt, tile);
  const BoxDesc bd = B.box[b];
  const int lt = tile - box_first_tile(tt, b);
  const float sc = box_scale<NT4>(B, b), isc = 1.0f / sc;
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
  } else if (AX == 1) {
    const int nxt = (W1 + 31) / 32;
    const int jo = lt / nxt;
    const int j1 = (lt - jo * nxt) * 32;
    nl = min(32, W1 - j1);
    base = bd.w_off + j1 + (long long)W1 * W2 * jo;
    lstride = 1;
    pstride = W1;
  } else {
    // z lines x + W1 y of the box, 32 consecutive ones per tile (P_KZ)
    const int l0 = lt * 32;
    nl = min(32, W1 * W2 - l0);
    base = bd.w_off + l0;
    lstride = 1;
    pstride = (long long)W1 * W2;
  }
  // the K steps read columns k < t.Kp: columns Wd..Kp-1 and lines nl..31
  // are zero
  const float* W = (const float*)B.W;
  const int KpW = t.Kp;
  const float ikp = 1.0f / (float)KpW;
  const int tot = 32 * KpW;
  for (int i0 = threadIdx.x; i0 < tot; i0 += NTM * 8) {
    float v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int i = i0 + u * NTM;
      int l, k;
      if (AX == 0) { l = qdiv(i, KpW, ikp); k = i - l * KpW; } else { k = i >> 5; l = i & 31; }
      v[u] = (i < tot && l < nl && k < Wd) ? W[base + l * lstride + k * pstride] : 0.0f;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int i = i0 + u * NTM;
      if (i < tot) {
        int l, k;
        if (AX == 0) { l = qdiv(i, KpW, ikp); k = i - l * KpW; } else { k = i >> 5; l = i & 31; }
        __bf16 h, lo;
        split2<NT4>(v[u] * sc, h, lo);
        xh[l * XS + k] = h;
        xl[l * XS + k] = lo;
        if (k < Wd) xf[l * XF + k] = v[u];
      }
    }
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const float ks = (float)bd.kscale[AX] * kin_unscale<NT4>();
  float* K = (float*)B.K;
  double ek = 0.0;
  for (int ob = wv; ob < t.Dp / 32; ob += NTM / 64) {
    f32x16 acc = {};
    constexpr bool TB = AX == 0;
    mma<NT4, TB>(B, t, 32 * ob, xh, xl, XS, acc);
    float kold[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int line = TB ? drow(r, lane) : (lane & 31);
      const int o = 32 * ob + (TB ? (lane & 31) : drow(r, lane));
      kold[r] = 0.0f;
      if (AX != 0 && o < Wd && line < nl) kold[r] = K[base + line * lstride + o * pstride];
    }
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int line = TB ? drow(r, lane) : (lane & 31);
      const int o = 32 * ob + (TB ? (lane & 31) : drow(r, lane));
      if (o < Wd && line < nl) {
        const float v = acc[r] * (ks * isc);
        const float kn = v + kold[r];
        ek += (double)(TOTAL ? kn : v) * (double)xf[line * XF + o];
        K[base + line * lstride + o * pstride] = kn;
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
  const int tile = tile_of(tt);
  const int b = find_box(tt, tile);
  const BoxDesc bd = B.box[b];
  const int j3 = tile - box_first_tile(tt, b);
  const float sc = box_scale<NT4>(B, b), isc = 1.0f / sc;
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
  const float iw1 = 1.0f / (float)W1;
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
        const int j2 = qdiv(i, W1, iw1), j1 = i - j2 * W1;
        __bf16 h, lo;
        split2<NT4>(v[u] * sc, h, lo);
        rh[j2 * XS1 + j1] = h;
        rl[j2 * XS1 + j1] = lo;
      }
    }
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  // G_x: lines j2, outputs p1
  {
    const int nob = tx.Dp / 32, nlb = n1 / 32;
    for (int it = wv; it < nob * nlb; it += NTMP / 64) {
      const int ob = it % nob, lb = it / nob;
      f32x16 acc = {};
      mfma_block<NT4>(B, tx, 32 * ob, rh + lb * 32 * XS1, rl + lb * 32 * XS1, XS1, acc);
      if (NT4 == 16) acc *= G16_OUT;   // line p1 keeps its row scale s(p1)
      const int j2 = 32 * lb + (lane & 31);
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int p1 = 32 * ob + drow(r, lane);
        if (p1 < L1 && j2 < W2) {
          __bf16 h, lo;
          split2<NT4>(acc[r], h, lo);
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
      // fp16 tables: the line keeps the row scales of p1 (G_x) and p2 (G_y)
      // and G16_OUT; the z potential pass removes them (Batch::t2_scaled)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int p2 = 32 * ob + drow(r, lane);
        if (p1 < L1 && p2 < L2) dst[p1 + (long long)L1 * p2] = acc[r] * isc;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// z lines of T2: G_z, V, epot, G_z^T, in place
// ---------------------------------------------------------------------------
template <int NT4, int BM>
__global__ __launch_bounds__(NTM) void k_mzpot(Batch B, TileTable tt, double* part) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ double red[NTM / 64];
  const int tile = tile_of(tt);
  const int b = find_box(tt, tile);
  const BoxDesc bd = B.box[b];
  const int lt = tile - box_first_tile(tt, b);
  const float sc = box_scale<NT4>(B, b), isc = 1.0f / sc;
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
        split2<NT4>(v[u] * sc, h, lo);
        xh[l * XS + k] = h;
        xl[l * XS + k] = lo;
      }
    }
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int line = lane & 31;
  const int q1 = j1 + line - bd.ishift[0], q2 = j2 - bd.ishift[1];
  const bool in12 = line < nl && q1 >= 0 && q1 < bd.potdim[0] && q2 >= 0 &&
                    q2 < bd.potdim[1];
  const double* pot = B.pot + bd.pot_off + q1 + (long long)bd.potdim[0] * q2;
  const long long pz = (long long)bd.potdim[0] * bd.potdim[1];
  const double cxy = bd.conf_on ? conf_axis2(bd, 0, j1 + line) + conf_axis2(bd, 1, j2) : 0.0;
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
        if (in12 && q3 >= 0 && q3 < bd.potdim[2]) {
          v = pot[pz * q3];
          if (bd.conf_on) v += conf_term(bd, cxy + conf_axis2(bd, 2, p3));
        }
        const double pr = (double)(acc[r] * isc * grow_u(B, tg, p3));
        ep = fma(v * pr, pr, ep);
        __bf16 h, lo;
        split2<BM>((float)(v * pr), h, lo);
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
    mfma_block<BM>(B, tsh, 32 * ob, rh, rl, RS, acc);
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
template <int BM>
__global__ __launch_bounds__(NTMP) void k_mshrink_plane(Batch B, TileTable tt) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int tile = tile_of(tt);
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
  const float il1 = 1.0f / (float)L1;
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
        const int p2 = qdiv(i, L1, il1), p1 = i - p2 * L1;
        __bf16 h, lo;
        split2<BM>(v[u], h, lo);
        uh[p1 * XU + p2] = h;
        ul[p1 * XU + p2] = lo;
      }
    }
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  {  // G_y^T: lines p1, outputs j2
    const int nob = sy.Dp / 32, nlb = nu / 32;
    for (int it = wv; it < nob * nlb; it += NTMP / 64) {
      const int ob = it % nob, lb = it / nob;
      f32x16 acc = {};
      mfma_block<BM>(B, sy, 32 * ob, uh + lb * 32 * XU, ul + lb * 32 * XU, XU, acc);
      const int p1 = 32 * lb + (lane & 31);
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int o2 = 32 * ob + drow(r, lane);
        if (o2 < W2 && p1 < L1) {
          __bf16 h, lo;
          split2<BM>(acc[r], h, lo);
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
      mfma_block<BM>(B, sx, 32 * ob, bh + lb * 32 * XB, bl + lb * 32 * XB, XB, acc);
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
// Plane-fused forward pass, one work-group per x-y plane j3 of a box:
//   K  = (Ktilde_x + Ktilde_y) W      (kinetic of the plane)
//   T2 = G_y G_x W                    (real space in x, y; wavelet z)
// LDS: R = the W plane as lines j2 [j2][j1] (bf16 hi, lo); C = G_x W as lines
// p1 [p1][j2].
// Two phases: the kinetic tiles and the G_x units (both read R only), then
// G_y (reads C).
// Orientation: x kinetic with the table as B (rows j2, lanes j1 -> K stored
// along x); y kinetic with the table as A and R's rows as the K steps (rows
// j2, lanes j1 -> added to K along x); G_x as B (lanes p1 -> columns of C);
// G_y as A (lanes p1 -> T2).
// ---------------------------------------------------------------------------
template <int NT4>
__global__ __launch_bounds__(NTMP, 4) void k_mfwd(Batch B, TileTable tt) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int tile = tile_of(tt);
  const int lb = find_box(tt, tile);
  const int b = table_box(tt, lb);
  const BoxDesc bd = B.box[b];
  const int j3 = tile - box_first_tile(tt, lb);
  const float sc = box_scale<NT4>(B, b), isc = 1.0f / sc;
  const int W1 = 2 * bd.N[0], W2 = 2 * bd.N[1], L1 = bd.L[0], L2 = bd.L[1];
  const OpTab tx = B.tabs[bd.opG[0]], ty = B.tabs[bd.opG[1]];
  const OpTab kx = B.tabs[bd.opK[0]], ky = B.tabs[bd.opK[1]];
  const int KX = rup(W1, 32), nR = rup(W2, 32);
  const int XR = KX + 8, XC = nR + 8;
  // G_x output blocks of p1, produced and consumed in bd.mfp parts (two for
  // large planes: the C buffer then holds half of them, which keeps two
  // work-groups per CU)
  const int nobG = tx.Dp / 32;
  const int np = bd.mfp > 1 ? bd.mfp : 1;
  const int hb = (nobG + np - 1) / np;
  const int nC = 32 * hb;
  __bf16* rh = (__bf16*)smem;  // [j2 * XR + j1], j2 < nR, j1 < KX (zero padded)
  __bf16* rl = rh + nR * XR;
  __bf16* ch = rl + nR * XR;   // [p1 * XC + j2]
  __bf16* cl = ch + nC * XC;
  // load the plane, two x-neighbours per thread (W1 is even)
  {
    const float2* src =
        (const float2*)((const float*)B.W + bd.w_off + (long long)W1 * W2 * j3);
    const int hp = KX / 2;
    const float ihp = 1.0f / (float)hp;
    const int tot = nR * hp;
    for (int i0 = threadIdx.x; i0 < tot; i0 += NTMP * 8) {
      float2 v[8];
      int jj[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int i = i0 + u * NTMP;
        const int j2 = qdiv(i, hp, ihp), j1 = 2 * (i - j2 * hp);
        jj[u] = j2 * XR + j1;
        v[u] = (i < tot && j2 < W2 && j1 < W1) ? src[(j2 * W1 + j1) >> 1]
                                                : make_float2(0.0f, 0.0f);
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int i = i0 + u * NTMP;
        if (i < tot) {
          __bf16 h0, l0, h1, l1;
          split2<NT4>(v[u].x * sc, h0, l0);
          split2<NT4>(v[u].y * sc, h1, l1);
          *(bf16x2*)(rh + jj[u]) = bf16x2{h0, h1};
          *(bf16x2*)(rl + jj[u]) = bf16x2{l0, l1};
        }
      }
    }
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  constexpr int NW = NTMP / 64;
  const int nlb = nR / 32;          // line blocks j2
  const int nobK = KX / 32;         // output blocks j1 of the kinetic
  float* K = (float*)B.K + bd.w_off + (long long)W1 * W2 * j3;
  const float ksx = (float)bd.kscale[0] * isc * kin_unscale<NT4>(),
              ksy = (float)bd.kscale[1] * isc * kin_unscale<NT4>();
  // The kinetic tiles and G_x -> C: both read only R, so they share one
  // phase.  The nK kinetic tiles (two operator sweeps each) come first, then
  // the nobG x nlb (output block, line block) units of G_x (one sweep); the
  // waves take
  // them in snake order (wave w gets w, 2 NW - 1 - w, 2 NW + w, ...), which
  // evens out the sweeps per wave.
  // Kinetic tile (j2 block lb, j1 block ob): x part from R with the table as
  // B operand, y part down R's rows with the table as A operand; both land
  // as rows j2, lanes j1, and K is stored once along x.
  for (int h = 0; h < np; ++h) {
  const int ob0 = h * hb, ob1 = min(nobG, ob0 + hb);
  {
    const int nK = h == 0 ? nlb * nobK : 0, nT = nK + (ob1 - ob0) * nlb;
    for (int r = 0; r * NW < nT; ++r) {
      const int q = r * NW + ((r & 1) ? NW - 1 - wv : wv);
      if (q >= nT) continue;
      if (q < nK) {
        const int lb = q / nobK, ob = q - lb * nobK;
        f32x16 ax = {}, ay = {};
        mma<NT4, true>(B, kx, 32 * ob, rh + 32 * lb * XR, rl + 32 * lb * XR, XR, ax);
        mma_rows<NT4>(B, ky, 32 * lb, rh, rl, XR, 32 * ob, ay);
        const int j1 = 32 * ob + (lane & 31);
        if (j1 < W1) {
          // byte offsets from the plane's K (uniform base): row j2 of
          // register rr is drow(0, lane) + 8 (rr / 4) + rr % 4
          char* kb = (char*)K;
          const unsigned rs = 4u * (unsigned)W1;
          const unsigned o0 = 4u * (unsigned)(j1 + W1 * (32 * lb + drow(0, lane)));
          const bool full = 32 * lb + 32 <= W2;
#pragma unroll
          for (int rr = 0; rr < 16; ++rr) {
            const int dr = 8 * (rr >> 2) + (rr & 3);
            if (full || 32 * lb + drow(rr, lane) < W2)
              *(float*)(kb + (o0 + rs * (unsigned)dr)) = ax[rr] * ksx + ay[rr] * ksy;
          }
        }
      } else {
        const int g = q - nK, obl = g / nlb, lb = g - obl * nlb, ob = ob0 + obl;
        f32x16 acc = {};
        mma<NT4, true>(B, tx, 32 * ob, rh + 32 * lb * XR, rl + 32 * lb * XR, XR, acc);
        if (NT4 == 16) acc *= G16_OUT;   // line p1 keeps its row scale s(p1)
        store_cols<NT4>(acc, ch, cl, XC, 32 * obl + (lane & 31), 32 * lb);
      }
    }
  }
  __syncthreads();
  // G_y: lines p1, outputs p2 -> T2 plane, lanes along p1
  {
    float* dst = (float*)B.T2 + bd.t2_off + (long long)L1 * L2 * j3;
    const int nl1 = ob1 - ob0, nob = ty.Dp / 32;   // line blocks p1 of this part
    const int ch3 = task_chunks(nob, nl1, NW);
    for (int q = wv; q < nob * ch3; q += NW) {
      const Task tk = task_of(q, ch3, nl1);
      const int ob = tk.u;
      constexpr int MBF = mbf<NT4>();
      for (int lb0 = tk.lb_begin; lb0 < tk.lb_end; lb0 += MBF) {
        const int nb = min(MBF, tk.lb_end - lb0);
        f32x16 acc[MBF] = {};
        mma_multi<NT4, false, MBF>(B, ty, 32 * ob, ch + 32 * lb0 * XC, cl + 32 * lb0 * XC, XC,
                              32 * XC, nb, acc);
#pragma unroll
        for (int j = 0; j < MBF; ++j) {
          if (j >= nb) break;
          const int p1 = 32 * (ob0 + lb0 + j) + (lane & 31);
          // fp16 tables: the line keeps the row scales of p1 (G_x) and p2
          // (G_y) and G16_OUT; the z potential pass removes them
          // (Batch::t2_scaled)
          if (p1 < L1) {
            // byte offsets from the plane's T2 (uniform base; a plane holds
            // L1 L2 < 2^30 values): row p2 of register r is
            // drow(0, lane) + 8 (r / 4) + r % 4
            char* db = (char*)dst;
            const unsigned rs = 4u * (unsigned)L1;
            const unsigned o0 = 4u * (unsigned)(p1 + L1 * (32 * ob + drow(0, lane)));
            const bool full = 32 * ob + 32 <= L2;
#pragma unroll
            for (int r = 0; r < 16; ++r) {
              const int dr = 8 * (r >> 2) + (r & 3);
              if (full || 32 * ob + drow(r, lane) < L2)
                *(float*)(db + (o0 + rs * (unsigned)dr)) = acc[j][r] * isc;
            }
          }
        }
      }
    }
  }
  if (h + 1 < np) __syncthreads();   // the next part overwrites C
  }
}

// ---------------------------------------------------------------------------
// Plane-fused backward pass, one work-group per plane j3:
//   hpsi += compress(G_y^T G_x^T T2 + K)        (compress_and_accumulate_*)
// LDS: R = the T2 plane as lines p2 [p2][p1]; C = G_x^T T2 as lines j1
// [j1][p2] (table as B operand); G_y^T with the table as A operand puts
// lanes along j1 for the K reads and the compress.  One wave per output
// block, sweeping all line bundles.
// ---------------------------------------------------------------------------
template <int BM>
__global__ __launch_bounds__(NTMP) void k_mbwd(Batch B, TileTable tt) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int tile = tile_of(tt);
  const int lb = find_box(tt, tile);
  const int b = table_box(tt, lb);
  const BoxDesc bd = B.box[b];
  const int j3 = tile - box_first_tile(tt, lb);
  const int N1 = bd.N[0], N2 = bd.N[1], N3 = bd.N[2], L1 = bd.L[0], L2 = bd.L[1];
  const int W1 = 2 * N1, W2 = 2 * N2;
  const OpTab sx = B.tabs[bd.opS[0]], sy = B.tabs[bd.opS[1]];
  const int KX = rup(L1, 32), nR = rup(L2, 32);
  const int XR = KX + 8, XC = nR + 8;
  const int nC = rup(W1, 32);
  __bf16* rh = (__bf16*)smem;  // [p2 * XR + p1]
  __bf16* rl = rh + nR * XR;
  __bf16* ch = rl + nR * XR;   // [j1 * XC + p2]
  __bf16* cl = ch + nC * XC;
  {
    const float* src = (const float*)B.T2 + bd.t2_off + (long long)L1 * L2 * j3;
    const int tot = nR * KX;
    const float ikx = 1.0f / (float)KX;
    for (int i0 = threadIdx.x; i0 < tot; i0 += NTMP * 8) {
      float v[8];
      int oo[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int i = i0 + u * NTMP;
        const int p2 = qdiv(i, KX, ikx), p1 = i - p2 * KX;
        oo[u] = p2 * XR + p1;
        v[u] = (i < tot && p2 < L2 && p1 < L1) ? src[p1 + L1 * p2] : 0.0f;
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int i = i0 + u * NTMP;
        if (i < tot) {
          __bf16 h, l;
          split2<BM>(v[u], h, l);
          rh[oo[u]] = h;
          rl[oo[u]] = l;
        }
      }
    }
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  constexpr int NW = NTMP / 64;
  {  // G_x^T: lines p2, outputs j1 -> C[j1][p2]
    const int nlb = nR / 32, nob = sx.Dp / 32;
    const int chx = task_chunks(nob, nlb, NW);
    for (int q = wv; q < nob * chx; q += NW) {
      const Task tk = task_of(q, chx, nlb);
      const int ob = tk.u;
      for (int lb0 = tk.lb_begin; lb0 < tk.lb_end; lb0 += MBB) {
        const int nb = min(MBB, tk.lb_end - lb0);
        f32x16 acc[MBB] = {};
        mma_multi<BM, true, MBB>(B, sx, 32 * ob, rh + 32 * lb0 * XR, rl + 32 * lb0 * XR, XR,
                             32 * XR, nb, acc);
#pragma unroll
        for (int j = 0; j < MBB; ++j) {
          if (j >= nb) break;
          store_cols<BM>(acc[j], ch, cl, XC, 32 * ob + (lane & 31), 32 * (lb0 + j));
        }
      }
    }
  }
  __syncthreads();
  // G_y^T: lines j1, outputs j2, into the plane O [j2][j1] (FP32, over the
  // dead R space), lanes along j1
  const int RO = W1 + 1;
  float* O = (float*)smem;
  const int nl1 = nC / 32, nob = sy.Dp / 32;
  const int chy = task_chunks(nob, nl1, NW);
  for (int q = wv; q < nob * chy; q += NW) {
    const Task tk = task_of(q, chy, nl1);
    const int ob = tk.u;
    for (int lb0 = tk.lb_begin; lb0 < tk.lb_end; lb0 += MBB) {
      const int nb = min(MBB, tk.lb_end - lb0);
      f32x16 acc[MBB] = {};
      mma_multi<BM, false, MBB>(B, sy, 32 * ob, ch + 32 * lb0 * XC, cl + 32 * lb0 * XC, XC,
                                32 * XC, nb, acc);
#pragma unroll
      for (int j = 0; j < MBB; ++j) {
        if (j >= nb) break;
        const int j1 = 32 * (lb0 + j) + (lane & 31);
        if (j1 >= W1) continue;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const int j2 = 32 * ob + drow(r, lane);
          if (j2 < W2) O[j2 * RO + j1] = acc[j][r];
        }
      }
    }
  }
  __syncthreads();
  // + K, compress-accumulate (un-compress.f90:2174-2238 mixed, 2105-2170
  // standard: the same scatter in this layout)
  compress_plane<double, 4, NTMP>(bd, j3, O, RO,
                                  (const float*)B.K + bd.w_off + (long long)W1 * W2 * j3,
                                  B.hpsi);
}

// ---------------------------------------------------------------------------
// z lines, ZG bundles of 32 lines per work-group (each table fragment read
// once for all of them).  Kinetic z: K += Ktilde_z W, and the tile's share of
// ekin = sum W * K (K complete after this pass).
// ---------------------------------------------------------------------------
constexpr int ZU = 24;   // loads in flight per thread in the z-line loaders

// Stages ZG bundles of 32 consecutive-x z lines of a float grid into LDS as
// split pieces [j][line][k] for k < Kp (rows nrows <= k < Kp are zero), and
// with KEEPF the float values of rows k < nrows in xf [j][line][k].  src
// points at (x 0, y line 0, z 0) of the box (x stride 1, y stride rowstride,
// z stride pstride); bundle bi = ZG*lt + j holds x lines 32*(bi % nxt)..+31
// of y line bi / nxt, valid for x < nvalid.  Thread t keeps the lines
// LPT*(t % TPR) + (0..LPT-1) and two consecutive rows per unrolled step, so
// bundle and row follow from the step index with no division, and both rows
// go to LDS in one 32-bit store per piece.  Every load of the thread is in
// flight before the first store.  Requires Kp <= 128 (n <= 64).
// FLAT: bundle bi is the 32 consecutive lines 32 bi .. of the plane (a
// bundle may span two rows; nxt and rowstride unused), valid for line index
// < nvalid.
template <int NT4, bool PAIRS, bool KEEPF, int ZG, bool FLAT = false>
__device__ __forceinline__ void stage_zlines(const float* __restrict__ src, int rowstride,
                                             long long pstride, int nxt, int lt, int nb,
                                             int nvalid, int nrows, int Kp, float sc,
                                             __bf16* xh, __bf16* xl, int XS, float* xf,
                                             int XF) {
  constexpr int LPT = PAIRS ? 2 : 1;       // x lines per thread
  constexpr int TPR = 32 / LPT;            // threads per row of a bundle
  constexpr int RPS = 2 * (NTM / TPR);     // rows per step
  constexpr int SPB = 128 / RPS;           // steps per bundle
  constexpr int NS = ZG * SPB;
  const int t = threadIdx.x;
  const int l = LPT * (t % TPR);
  const int kq = 2 * (t / TPR);
  long long base[ZG];
  int x0[ZG];
#pragma unroll
  for (int j = 0; j < ZG; ++j) {
    const int bi = ZG * lt + j;
    if (FLAT) {
      x0[j] = 32 * bi;
      base[j] = x0[j];
    } else {
      const int y = bi / nxt;
      x0[j] = (bi - y * nxt) * 32;
      base[j] = x0[j] + (long long)rowstride * y;
    }
  }
  float v[NS][2][LPT];
#pragma unroll
  for (int u = 0; u < NS; ++u) {
    const int j = u / SPB;
    const int k = kq + RPS * (u % SPB);
    const bool ok = j < nb && x0[j] + l < nvalid && k < nrows;
    const float* a = src + base[j] + l + k * pstride;
    if (PAIRS) {
      float2 p = make_float2(0.0f, 0.0f), q = p;
      if (ok) {
        p = *(const float2*)a;
        q = *(const float2*)(a + pstride);
      }
      v[u][0][0] = p.x;
      v[u][1][0] = q.x;
      v[u][0][LPT - 1] = p.y;
      v[u][1][LPT - 1] = q.y;
    } else {
      v[u][0][0] = ok ? a[0] : 0.0f;
      v[u][1][0] = ok ? a[pstride] : 0.0f;
    }
  }
#pragma unroll
  for (int u = 0; u < NS; ++u) {
    const int j = u / SPB;
    const int k = kq + RPS * (u % SPB);
    if (j < nb && k < Kp) {
#pragma unroll
      for (int i = 0; i < LPT; ++i) {
        const int o = (j * 32 + l + i) * XS + k;   // k and XS even: 4-byte aligned
        bf16x2 hh, ll;
        __bf16 h, lo;
        split2<NT4>(v[u][0][i] * sc, h, lo);
        hh[0] = h;
        ll[0] = lo;
        split2<NT4>(v[u][1][i] * sc, h, lo);
        hh[1] = h;
        ll[1] = lo;
        *(bf16x2*)(xh + o) = hh;
        *(bf16x2*)(xl + o) = ll;
        if (KEEPF && k < nrows) {
          xf[(j * 32 + l + i) * XF + k] = v[u][0][i];
          xf[(j * 32 + l + i) * XF + k + 1] = v[u][1][i];
        }
      }
    }
  }
}

template <int NT4, int ZG>
__global__ __launch_bounds__(NTM) void k_mkz(Batch B, TileTable tt, double* part) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ double red[NTM / 64];
  const int tile = tile_of(tt);
  const int b = find_box(tt, tile);
  const BoxDesc bd = B.box[b];
  const int lt = tile - box_first_tile(tt, b);
  const float sc = box_scale<NT4>(B, b), isc = 1.0f / sc;
  const int W1 = 2 * bd.N[0], W2 = 2 * bd.N[1], W3 = 2 * bd.N[2];
  const OpTab t = B.tabs[bd.opK[2]];
  const int XS = t.Kp + 8, XF = W3 + 1, KpW = t.Kp;
  const int nlin = W1 * W2;                  // z lines x + W1 y of the box
  const int nbund = (nlin + 31) / 32;
  const int nb = min(ZG, nbund - ZG * lt);
  const long long pstride = (long long)W1 * W2;
  __bf16* xh = (__bf16*)smem;          // [j][line][k]
  __bf16* xl = xh + ZG * 32 * XS;
  float* xf = (float*)(xl + ZG * 32 * XS);  // [j][line][k], k < W3
  // bundles of 32 consecutive lines (W1 even: pairs of neighbours per load)
  stage_zlines<NT4, true, true, ZG, true>((const float*)B.W + bd.w_off, W1, pstride, 1, lt, nb,
                                          nlin, W3, KpW, sc, xh, xl, XS, xf, XF);
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const float ks = (float)bd.kscale[2] * isc * kin_unscale<NT4>();
  // byte offsets from the box's K (uniform base; a box holds W1 W2 W3 <
  // 2^30 values): line li, row o at 4 (li + o W1 W2)
  char* kb = (char*)((float*)B.K + bd.w_off);
  const unsigned rs = 4u * (unsigned)(W1 * W2);
  double ek = 0.0;
  const int nobk = t.Dp / 32;
  if (2 * nobk <= NTM / 64 && nb > 1) {
    // at most half as many output blocks as waves: one (block, bundle) pair
    // per step, so every wave works (each reads the table fragments for its
    // own bundle; with three blocks for four waves sharing them is faster)
    for (int it = wv; it < nobk * nb; it += NTM / 64) {
      const int ob = it / nb, j = it - ob * nb;
      const int li = 32 * (ZG * lt + j) + (lane & 31);
      const bool okj = li < nlin;
      const unsigned o0 = 4u * (unsigned)(okj ? li : 0) + rs * (unsigned)(32 * ob + drow(0, lane));
      float kold[16];
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int o = 32 * ob + drow(r, lane);
        const unsigned dr = (unsigned)(8 * (r >> 2) + (r & 3));
        kold[r] = (okj && o < W3) ? *(const float*)(kb + (o0 + rs * dr)) : 0.0f;
      }
      f32x16 acc[1] = {};
      mma_multi<NT4, false, 1>(B, t, 32 * ob, xh + j * 32 * XS, xl + j * 32 * XS, XS,
                               32 * XS, 1, acc);
      if (okj) {
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const int o = 32 * ob + drow(r, lane);
          if (o < W3) {
            const float kn = kold[r] + acc[0][r] * ks;
            ek += (double)kn * (double)xf[(j * 32 + (lane & 31)) * XF + o];
            *(float*)(kb + (o0 + rs * (unsigned)(8 * (r >> 2) + (r & 3)))) = kn;
          }
        }
      }
    }
  } else
  for (int ob = wv; ob < t.Dp / 32; ob += NTM / 64) {
    // the K values this wave updates, read while the matrix cores run
    float kold[ZG][16];
    unsigned base[ZG];
    bool ok[ZG];
#pragma unroll
    for (int j = 0; j < ZG; ++j) {
      const int li = 32 * (ZG * lt + j) + (lane & 31);
      ok[j] = j < nb && li < nlin;
      base[j] = 4u * (unsigned)(ok[j] ? li : 0) + rs * (unsigned)(32 * ob + drow(0, lane));
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int o = 32 * ob + drow(r, lane);
        const unsigned dr = (unsigned)(8 * (r >> 2) + (r & 3));
        kold[j][r] = (ok[j] && o < W3) ? *(const float*)(kb + (base[j] + rs * dr)) : 0.0f;
      }
    }
    f32x16 acc[ZG] = {};
    mma_multi<NT4, false, ZG>(B, t, 32 * ob, xh, xl, XS, 32 * XS, nb, acc);
#pragma unroll
    for (int j = 0; j < ZG; ++j) {
      if (!ok[j]) continue;
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int o = 32 * ob + drow(r, lane);
        if (o < W3) {
          const float kn = kold[j][r] + acc[j][r] * ks;
          ek += (double)kn * (double)xf[(j * 32 + (lane & 31)) * XF + o];
          *(float*)(kb + (base[j] + rs * (unsigned)(8 * (r >> 2) + (r & 3)))) = kn;
        }
      }
    }
  }
  const double eks = wave_block_sum(ek, red, NTM / 64);
  if (threadIdx.x == 0) part[tile] = eks;
}

// z lines of T2, ZG bundles per work-group: G_z, V (epot), G_z^T in place.
template <int NT4, int BM, int ZG>
__global__ __launch_bounds__(NTM, 2) void k_mzp(Batch B, TileTable tt, double* part) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ double red[NTM / 64];
  const int tile = tile_of(tt);
  const int b = find_box(tt, tile);
  const BoxDesc bd = B.box[b];
  const int lt = tile - box_first_tile(tt, b);
  const float sc = box_scale<NT4>(B, b), isc = 1.0f / sc;
  const int L1 = bd.L[0], L2 = bd.L[1], W3 = 2 * bd.N[2], L3 = bd.L[2];
  const OpTab tg = B.tabs[bd.opG[2]], tsh = B.tabs[bd.opS[2]];
  const int XS = tg.Kp + 8, KpG = tg.Kp;
  const int RS = tg.Dp + 8;   // rows p3 < tg.Dp are written; tsh reads p3 < tsh.Kp
  const int nxt = (L1 + 31) / 32;
  const int nbund = nxt * L2;
  const int nb = min(ZG, nbund - ZG * lt);
  const long long pstride = (long long)L1 * L2;
  float* T2 = (float*)B.T2 + bd.t2_off;
  __bf16* xh = (__bf16*)smem;
  __bf16* xl = xh + ZG * 32 * XS;
  __bf16* rh = xl + ZG * 32 * XS;
  __bf16* rl = rh + ZG * 32 * RS;
  // z lines of 32 consecutive x; pairs of x-neighbours when L1 is even
  if ((L1 & 1) == 0)
    stage_zlines<NT4, true, false, ZG>(T2, L1, pstride, nxt, lt, nb, L1, W3, KpG, sc, xh, xl,
                                       XS, nullptr, 0);
  else
    stage_zlines<NT4, false, false, ZG>(T2, L1, pstride, nxt, lt, nb, L1, W3, KpG, sc, xh, xl,
                                        XS, nullptr, 0);
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const long long pz = (long long)bd.potdim[0] * bd.potdim[1];
  double ep = 0.0;
  // G_z, V: rows p3 < tg.Dp (zero beyond L3) -> R[j][line][p3].  The potential
  // of bundle j+1 is read while bundle j is finished (bundle 0: during the
  // matrix-core sweep).
  auto load_pot = [&](int j, int ob, double (&vv)[16]) {
    const int bi = ZG * lt + j;
    const int p2 = bi / nxt;
    const int p1 = (bi - p2 * nxt) * 32 + (lane & 31);
    const int q1 = p1 - bd.ishift[0], q2 = p2 - bd.ishift[1];
    const bool in12 = j < nb && p1 < L1 && q1 >= 0 && q1 < bd.potdim[0] && q2 >= 0 &&
                      q2 < bd.potdim[1];
    const double* pot = B.pot + bd.pot_off + q1 + (long long)bd.potdim[0] * q2;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int p3 = 32 * ob + drow(r, lane);
      const int q3 = p3 - bd.ishift[2];
      vv[r] = (in12 && p3 < L3 && q3 >= 0 && q3 < bd.potdim[2]) ? pot[pz * q3] : 0.0;
    }
    if (bd.conf_on) {
      const double cxy = conf_axis2(bd, 0, p1) + conf_axis2(bd, 1, p2);
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int p3 = 32 * ob + drow(r, lane);
        const int q3 = p3 - bd.ishift[2];
        if (in12 && p3 < L3 && q3 >= 0 && q3 < bd.potdim[2])
          vv[r] += conf_term(bd, cxy + conf_axis2(bd, 2, p3));
      }
    }
  };
  auto finish = [&](int j, int ob, const f32x16& a, const double (&vv)[16]) {
    const int line = lane & 31;
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      bf16x4 h, l;
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int r = 4 * g + u;
        const double pr = (double)(a[r] * isc * grow_u(B, tg, 32 * ob + drow(r, lane)));
        ep = fma(vv[r] * pr, pr, ep);
        __bf16 x, y;
        split2<BM>((float)(vv[r] * pr), x, y);
        h[u] = x;
        l[u] = y;
      }
      const int o = (j * 32 + line) * RS + 32 * ob + 8 * g + 4 * (lane >> 5);
      *(bf16x4*)(rh + o) = h;
      *(bf16x4*)(rl + o) = l;
    }
  };
  const int nob1 = tg.Dp / 32;
  if (nob1 < NTM / 64 && nb > 1) {
    // fewer output blocks than waves: one (block, bundle) pair per step
    for (int it = wv; it < nob1 * nb; it += NTM / 64) {
      const int ob = it / nb, j = it - ob * nb;
      double va[16];
      load_pot(j, ob, va);
      f32x16 acc[1] = {};
      mma_multi<NT4, false, 1>(B, tg, 32 * ob, xh + j * 32 * XS, xl + j * 32 * XS, XS,
                               32 * XS, 1, acc);
      finish(j, ob, acc[0], va);
    }
  } else
  for (int ob = wv; ob < tg.Dp / 32; ob += NTM / 64) {
    double va[16], vb[16];
    load_pot(0, ob, va);
    f32x16 acc[ZG] = {};
    mma_multi<NT4, false, ZG>(B, tg, 32 * ob, xh, xl, XS, 32 * XS, nb, acc);
    static_assert(ZG == 2 || ZG == 3, "the potential pipeline below handles two or three bundles");
    if (nb > 1) load_pot(1, ob, vb);
    finish(0, ob, acc[0], va);
    if (nb > 1) {
      if constexpr (ZG > 2) {
        if (nb > 2) load_pot(2, ob, va);
      }
      finish(1, ob, acc[1], vb);
    }
    if constexpr (ZG > 2) {
      if (nb > 2) finish(2, ob, acc[2], va);
    }
  }
  const double eps = wave_block_sum(ep, red, NTM / 64);
  if (threadIdx.x == 0) part[tile] = eps;
  __syncthreads();
  // G_z^T back to the wavelet z lines, in place
  const int nob2 = tsh.Dp / 32;
  if (nob2 < NTM / 64 && nb > 1) {
    for (int it = wv; it < nob2 * nb; it += NTM / 64) {
      const int ob = it / nb, j = it - ob * nb;
      f32x16 acc[1] = {};
      mma_multi<BM, false, 1>(B, tsh, 32 * ob, rh + j * 32 * RS, rl + j * 32 * RS, RS,
                              32 * RS, 1, acc);
      const int bi = ZG * lt + j;
      const int p2 = bi / nxt;
      const int p1 = (bi - p2 * nxt) * 32 + (lane & 31);
      if (p1 < L1) {
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const int o = 32 * ob + drow(r, lane);
          if (o < W3) T2[p1 + (long long)L1 * p2 + o * pstride] = acc[0][r];
        }
      }
    }
    return;
  }
  for (int ob = wv; ob < tsh.Dp / 32; ob += NTM / 64) {
    f32x16 acc[ZG] = {};
    mma_multi<BM, false, ZG>(B, tsh, 32 * ob, rh, rl, RS, 32 * RS, nb, acc);
#pragma unroll
    for (int j = 0; j < ZG; ++j) {
      if (j >= nb) break;
      const int bi = ZG * lt + j;
      const int p2 = bi / nxt;
      const int p1 = (bi - p2 * nxt) * 32 + (lane & 31);
      if (p1 >= L1) continue;
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int o = 32 * ob + drow(r, lane);
        if (o < W3) T2[p1 + (long long)L1 * p2 + o * pstride] = acc[j][r];
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

size_t mfma_lds_fwd(int W1, int W2, int L1, int parts) {
  const int nR = rup_h(W2, 32), KX = rup_h(W1, 32);
  const int nob = rup_h(L1, 32) / 32;
  const int nC = 32 * ((nob + parts - 1) / parts);
  return (size_t)4 * nR * (KX + 8) + (size_t)4 * nC * (nR + 8);
}
size_t mfma_lds_kz(int W3, int Kp, int zg) {
  return (size_t)zg * ((size_t)2 * 32 * (Kp + 8) * 2 + (size_t)32 * (W3 + 1) * 4);
}
size_t mfma_lds_zp(int KpG, int DpG, int zg) {
  return (size_t)zg * ((size_t)2 * 32 * (KpG + 8) * 2 + (size_t)2 * 32 * (DpG + 8) * 2);
}
size_t mfma_lds_bwd(int W1, int L1, int L2) {
  const int nR = rup_h(L2, 32), KX = rup_h(L1, 32), nC = rup_h(W1, 32);
  return (size_t)4 * nR * (KX + 8) + (size_t)4 * nC * (nR + 8);
}

template <int FM, int BM>
static void set_attr(size_t lds) {
  const void* k[] = {(const void*)k_mkin<0, FM, false>, (const void*)k_mkin<1, FM, false>,
                     (const void*)k_mkin<2, FM, false>, (const void*)k_mkin<2, FM, true>,
                     (const void*)k_mgrow_plane<FM>, (const void*)k_mzpot<FM, BM>,
                     (const void*)k_mshrink_plane<BM>, (const void*)k_mfwd<FM>,
                     (const void*)k_mbwd<BM>, (const void*)k_mkz<FM, 2>,
                     (const void*)k_mkz<FM, 3>, (const void*)k_mzp<FM, BM, 2>,
                     (const void*)k_mzp<FM, BM, 3>};
  for (const void* f : k)
    if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) !=
        hipSuccess)
      (void)hipGetLastError();
}

size_t mfma_attributes(size_t lds) {
  set_attr<3, 3>(lds);
  set_attr<4, 4>(lds);
  set_attr<16, 4>(lds);
  return lds;
}

// FM: format of the forward transforms and the kinetic; BM: format of the
// transposed transforms
template <int FM, int BM>
static hipError_t launch_mfma_t(const Batch& B, const PassTables& PT, const MfmaLds& L,
                                long long wtotal, bool need_zero, hipStream_t s,
                                hipEvent_t* ev, const Fork* fk) {
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
  if (PT.ntiles[P_UNC] > 0)
    launch_uncompress_float(B, PT.t[P_UNC], PT.ntiles[P_UNC], FM == 16, s);
  mark(S_UNC, 1);
  EnergyTables et;
  if (L.fused) {
    mark(S_PF, 0);
    if (PT.ntiles[P_PF] > 0)
      k_mfwd<FM><<<PT.ntiles[P_PF], NTMP, L.fwd, s>>>(B, PT.t[P_PF]);
    if (PT.ntiles_pf2 > 0) k_mfwd<FM><<<PT.ntiles_pf2, NTMP, L.fwd2, s>>>(B, PT.pf2);
    mark(S_PF, 1);
    // z kinetic (W, K) and z potential (T2) touch disjoint grids: with a
    // fork they run side by side (not while stages are being timed)
    const bool fork = fk && !ev;
    hipStream_t sz = s;
    if (fork) {
      hipError_t e = hipEventRecord(fk->a, s);
      if (e == hipSuccess) e = hipStreamWaitEvent(fk->s2, fk->a, 0);
      if (e != hipSuccess) return e;
      sz = fk->s2;
    }
    mark(S_ZP, 0);
    if (L.zvalu || FM == 16) {
      // G_z, V, G_z^T in FP32 on the vector units (the pass where they beat
      // the split-precision matrix-core form); T2 is unscaled FP32 in both
      const hipError_t e = launch_zpot_f32(B, PT, L.maxN, sz);
      if (e != hipSuccess) return e;
    } else if (PT.ntiles[P_ZPG] > 0) {
      if (L.zg == 2)
        k_mzp<FM, BM, 2><<<PT.ntiles[P_ZPG], NTM, L.zp, sz>>>(B, PT.t[P_ZPG], PT.part[3]);
      else
        k_mzp<FM, BM, 3><<<PT.ntiles[P_ZPG], NTM, L.zp, sz>>>(B, PT.t[P_ZPG], PT.part[3]);
    }
    mark(S_ZP, 1);
    mark(S_KZ, 0);
    if (PT.ntiles[P_KZG] > 0) {
      if (L.zg == 2)
        k_mkz<FM, 2><<<PT.ntiles[P_KZG], NTM, L.kz, s>>>(B, PT.t[P_KZG], PT.part[2]);
      else
        k_mkz<FM, 3><<<PT.ntiles[P_KZG], NTM, L.kz, s>>>(B, PT.t[P_KZG], PT.part[2]);
    }
    mark(S_KZ, 1);
    if (fork) {
      hipError_t e = hipEventRecord(fk->b, fk->s2);
      if (e == hipSuccess) e = hipStreamWaitEvent(s, fk->b, 0);
      if (e != hipSuccess) return e;
    }
    mark(S_PB, 0);
    if (L.bvalu) {
      // G_y^T G_x^T + K and compress in FP32 on the vector units (the plane
      // kernel of the FP32 sequence; T2 and K are unscaled FP32 in both)
      hipError_t e = hipSuccess;
      for (int k = 0; k < PT.n_vbg && e == hipSuccess; ++k)
        if (PT.nt_vbg[k] > 0)
          e = launch_plane_bwd<float>(B, PT.vbg[k], PT.nt_vbg[k], L.lds_vbg[k], s);
      if (e != hipSuccess) return e;
    } else {
      if (PT.ntiles[P_PB] > 0)
        k_mbwd<BM><<<PT.ntiles[P_PB], NTMP, L.bwd, s>>>(B, PT.t[P_PB]);
      if (PT.ntiles_pb2 > 0) k_mbwd<BM><<<PT.ntiles_pb2, NTMP, L.bwd2, s>>>(B, PT.pb2);
    }
    mark(S_PB, 1);
    et.t[0] = PT.t[P_KZG];
    et.p[0] = PT.part[2];
    et.t[1] = (L.zvalu || FM == 16) ? zpot_table(PT) : PT.t[P_ZPG];
    et.p[1] = PT.part[3];
    et.nk = 1;
    et.n = 2;
  } else {
    mark(S_XF, 0);
    if (PT.ntiles[P_XF] > 0)
      k_mkin<0, FM, false><<<PT.ntiles[P_XF], NTM, L.kin[0], s>>>(B, PT.t[P_XF],
                                                                   PT.part[0]);
    mark(S_XF, 1);
    mark(S_KY, 0);
    if (PT.ntiles[P_KY] > 0)
      k_mkin<1, FM, false><<<PT.ntiles[P_KY], NTM, L.kin[1], s>>>(B, PT.t[P_KY],
                                                                   PT.part[1]);
    mark(S_KY, 1);
    mark(S_KZ, 0);
    if (PT.ntiles[P_KZ] > 0)
      k_mkin<2, FM, false><<<PT.ntiles[P_KZ], NTM, L.kin[2], s>>>(B, PT.t[P_KZ],
                                                                   PT.part[2]);
    mark(S_KZ, 1);
    mark(S_PF, 0);
    if (PT.ntiles[P_PF] > 0)
      k_mgrow_plane<FM><<<PT.ntiles[P_PF], NTMP, L.grow, s>>>(B, PT.t[P_PF]);
    mark(S_PF, 1);
    mark(S_ZP, 0);
    if (FM == 16) {
      // F16X3: G_z, V, G_z^T in FP32 (the z lines of a free box's x-y halo
      // are far below the box scale the fp16 pieces are cut for)
      const hipError_t e = launch_zpot_f32(B, PT, L.maxN, s);
      if (e != hipSuccess) return e;
    } else if (PT.ntiles[P_ZP] > 0) {
      k_mzpot<FM, BM><<<PT.ntiles[P_ZP], NTM, L.zpot, s>>>(B, PT.t[P_ZP], PT.part[3]);
    }
    mark(S_ZP, 1);
    mark(S_PB, 0);
    if (PT.ntiles[P_PB] > 0)
      k_mshrink_plane<BM><<<PT.ntiles[P_PB], NTMP, L.shrink, s>>>(B, PT.t[P_PB]);
    mark(S_PB, 1);
    et.t[0] = PT.t[P_XF];
    et.p[0] = PT.part[0];
    et.t[1] = PT.t[P_KY];
    et.p[1] = PT.part[1];
    et.t[2] = PT.t[P_KZ];
    et.p[2] = PT.part[2];
    et.t[3] = FM == 16 ? zpot_table(PT) : PT.t[P_ZP];
    et.p[3] = PT.part[3];
    et.nk = 3;
    et.n = 4;
  }
  mark(S_ENERGY, 0);
  launch_energy(et, B, s);
  mark(S_ENERGY, 1);
  if (ev) {
    const int unused[] = {S_YF, S_YB, S_XB, S_COMP};
    for (int st : unused) {
      mark(st, 0);
      mark(st, 1);
    }
    if (L.fused) {
      mark(S_XF, 0);
      mark(S_XF, 1);
      mark(S_KY, 0);
      mark(S_KY, 1);
    }
  }
  return hipGetLastError();
}

hipError_t launch_mfma(const Batch& B, const PassTables& PT, const MfmaLds& L,
                       long long wtotal, bool need_zero, int precision, hipStream_t s,
                       hipEvent_t* ev, const Fork* fk) {
  switch (precision) {
    case BDFT_PREC_BF16X3: return launch_mfma_t<3, 3>(B, PT, L, wtotal, need_zero, s, ev, fk);
    case BDFT_PREC_BF16X4: return launch_mfma_t<4, 4>(B, PT, L, wtotal, need_zero, s, ev, fk);
    default: return launch_mfma_t<16, 4>(B, PT, L, wtotal, need_zero, s, ev, fk);
  }
}

}  // namespace bdft
