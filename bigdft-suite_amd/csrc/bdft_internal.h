// Internal declarations shared by the HIP kernels and the C-ABI layer.
//
// Data model (per orbital "box", i.e. one localisation region instance):
//   N[d]   coarse points along axis d (reference: n_d + 1)
//   L[d]   real-space (ISF) points along d: 2N (periodic) or 2N+29 (free),
//          the reference's n1i rule (bigdft/src/locregs/locregs.f90:871-884)
//   wavelet volume  "psig" layout of the reference (0:n1,2,0:n2,2,0:n3,2),
//          i.e. per axis [c(0..N-1) ; d(0..N-1)], x fastest
//          (bigdft/src/wavelib/un-compress.f90:877-948)
//   real volume     (L1, L2, L3), x fastest; free-BC index p <-> fine -14+p
//                   (bigdft/src/orbitals/growshrink.f90:12-37)
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bdft {

// Line-bundle tiling shared by every pass kernel.
constexpr int TW = 32;          // lines per tile
constexpr int TY = 8;           // chunk-threads per line
constexpr int NT = TW * TY;     // threads per block
constexpr int LDW = TW + 1;     // padded LDS row (doubles)
constexpr int HL = 15;          // halo rows left of a coarse section
constexpr int HR = 18;          // halo rows right of a coarse section
constexpr int MAXN = 256;       // largest coarse extent handled per axis
// longest lines whose outputs the line kernels hold in registers: 64 coarse
// points (x shrink, x kinetic), 192 fine points (x and z grow); longer
// lines take the kernels' BIG forms (outputs stored as they are made) and
// the two-kernel z potential pass
constexpr int REG_N = 64;
constexpr int REG_L = 192;
constexpr int ZSHORT_L = 128;
// largest coarse extent of the matrix-core kernels (z-line loaders of at
// most 128 columns); batches with a larger box run the FP32 vector passes
constexpr int MFMA_MAXN = 64;

// Per-box descriptor (device resident).
struct BoxDesc {
  int N[3];
  int L[3];
  int per;              // 1 periodic, 0 free
  int nvctr_c, nvctr_f;
  int potdim[3];        // potential box extents (n1ip, n2ip, n3ip)
  int ishift[3];        // potential box offset in wavefunction box coordinates
  int pad_;
  double kscale[3];     // -1/2/h_d^2, coarse grid spacing
  long long w_off;      // offset of the wavelet volume (8 N1 N2 N3)
  long long t1_off;     // offset of (L1, 2N2, 2N3)
  long long t2_off;     // offset of (L1, L2, 2N3)
  long long t3_off;     // offset of the real volume (L1, L2, L3): z lines
                        // longer than 192 points only (k_zgrowpot)
  long long pot_off;    // offset into the potential array
  long long psi_off;    // offset of this orbital's compressed coefficients
  const int* cmap;      // coarse points: dense linear index i1+N1*(i2+N2*i3)
  const int* fmap;      // fine points: same encoding
  const int* cinv;      // grid point -> coarse coefficient index or -1
  const int* finv;      // grid point -> fine point index or -1
                        // (null: identity, every point present in key order)
  int opG[3], opS[3], opK[3];  // operator tables per axis (matrix-core path)
  int mfp;              // matrix-core forward plane kernel: parts of the G_x
                        // output held in LDS at a time (1 or 2)
  int pad2_;
  // confining potential prefac * r2^conf_half added to V (conf_on = 0: none);
  // r2 = sum_d (conf_hh_d (p_d + 1 + conf_ioff_d) - conf_r_d)^2 at the fine
  // point p of the box (0-based)
  int conf_on, conf_half;
  int conf_ioff[3];
  double conf_prefac;
  double conf_hh[3];
  double conf_r[3];
};

// a / b for 0 <= a < 2^24 and 0 < b with inv = 1/b: one float product and
// one correction step each way instead of an integer division
__device__ __forceinline__ int qdiv(int a, int b, float inv) {
  int q = (int)((float)a * inv);
  if ((q + 1) * b <= a) ++q;
  if (q * b > a) --q;
  return q;
}

// the confining term at fine point (p1, p2, p3); callers test bd.conf_on
__device__ __forceinline__ double conf_axis2(const BoxDesc& bd, int d, int p) {
  const double x = bd.conf_hh[d] * (double)(p + 1 + bd.conf_ioff[d]) - bd.conf_r[d];
  return x * x;
}
__device__ __forceinline__ double conf_term(const BoxDesc& bd, double r2) {
  double t = 1.0;
  for (int k = 0; k < bd.conf_half; ++k) t *= r2;
  return bd.conf_prefac * t;
}

// A 1D operator as a padded matrix for the matrix-core path: rows = outputs,
// cols = inputs, split in bf16 hi/lo; mask bit s of 32-row block q set when
// columns [16s,16s+16) of those rows hold non-zeros.
struct OpTab {
  int rows, cols;   // Dout, Din
  int Dp, Kp;       // padded to multiples of 32 and 16
  long long off;    // unused (kept for layout)
  int moff;         // offset of the row blocks' {mask, fragment} entries
  int soff;         // fp16 G tables: offset of the rows' inverse scales u_r
                    // in Batch::trs (row r holds s_r = 1/u_r times G's row);
                    // -1: unscaled
};
// fp16 tables: scale of the kinetic table, and of every G output before it
// is split into pieces again (G rows are scaled to sums in (2^12, 2^13])
constexpr double KIN16 = 512.0;
constexpr float G16_OUT = 0x1p-13f;

// Tile bookkeeping for one pass: tiles of box b are [start[b], start[b+1]).
struct TileTable {
  const int* start;     // nbox + 1 cumulative tile counts (device)
  int nbox;
  int uniform;          // >0: every box has exactly this many tiles
  const int* bmap;      // null, or the batch index of each of the nbox boxes
  int xcd;              // 1: tiles dealt in contiguous ranges per XCD (xcd_tile)
  const int* tbox;      // null, or the box (index into start) of every tile
};

// Tile of work-group b in a launch of n work-groups.  Work-groups are
// dispatched to the 8 XCDs in turn (b mod 8), each XCD with its own L2; the
// tiles are dealt so that every XCD gets one contiguous range of them, and
// tiles that share cache lines (neighbouring line bundles of a plane, the
// two planes that write one fine record of hpsi, the operator tables of a
// box) meet in the same L2.  A bijection of [0, n).
__device__ __forceinline__ int xcd_tile(int b, int n) {
  const int q = n >> 3, r = n & 7, x = b & 7, i = b >> 3;
  return x * q + min(x, r) + i;
}

// tile of this work-group in a launch over table tt
__device__ __forceinline__ int tile_of(const TileTable& tt) {
  return tt.xcd ? xcd_tile(blockIdx.x, gridDim.x) : (int)blockIdx.x;
}

__device__ __forceinline__ int find_box(const TileTable& tt, int tile) {
  if (tt.uniform > 0) return tile / tt.uniform;
  if (tt.tbox) return tt.tbox[tile];
  int lo = 0, hi = tt.nbox;  // invariant start[lo] <= tile < start[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (tt.start[mid] <= tile) lo = mid; else hi = mid;
  }
  return lo;
}
// batch index of the table's box lb (tables over a subset of the batch)
__device__ __forceinline__ int table_box(const TileTable& tt, int lb) {
  return tt.bmap ? tt.bmap[lb] : lb;
}
__device__ __forceinline__ int box_first_tile(const TileTable& tt, int b) {
  return tt.uniform > 0 ? b * tt.uniform : tt.start[b];
}

// hpsi += compress(O + K) over the x-y plane j3 of a box: the four
// components of each coarse point (i1, i2) of the plane (s3 = j3 >= N3 fixed)
// added at their compressed positions, compress_and_accumulate_mixed
// (bigdft/src/wavelib/un-compress.f90:2174-2238).  O: the plane's operator
// output [j2 * RO + j1]; Kp: its kinetic term [j1 + W1 * j2]; each sum is
// formed in TS.  The work goes in steps of U points per thread (points
// i0 + u NT): compress_load issues every load of a step that does not need
// O (the compressed indices, K, the old hpsi values), compress_store adds O
// and stores, so the memory latencies of the U points overlap, and a kernel
// may load its first step before it computes O.  No two points share an
// hpsi entry, and the two planes j3, j3 + N3 of a coarse plane write
// disjoint components.
template <typename TK, int U>
struct CompressStep {
  int ci[U], fi[U];   // coarse / fine record of each point, -1: none
  TK k[U][4];         // K of the four components
  double old[U][4];   // hpsi before the step
};

template <int U, int NT, typename TK>
__device__ __forceinline__ void compress_load(const BoxDesc& bd, int j3, const TK* Kp,
                                              const double* hpsi, int i0,
                                              CompressStep<TK, U>& st) {
  const int N1 = bd.N[0], N2 = bd.N[1], N3 = bd.N[2], W1 = 2 * N1;
  const int s3 = j3 >= N3;
  const int i3 = j3 - s3 * N3;
  const int npl = N1 * N2;
  const double* hc = hpsi + bd.psi_off;
  const double* hf = hpsi + bd.psi_off + bd.nvctr_c + 3 * s3;  // components 1-3 or 4-7
  const float in1 = 1.0f / (float)N1;
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int i = i0 + u * NT;
    st.ci[u] = -1;
    st.fi[u] = -1;
#pragma unroll
    for (int s = 0; s < 4; ++s) st.k[u][s] = (TK)0;
    if (i < npl) {
      const int i2 = qdiv(i, N1, in1), i1 = i - i2 * N1;
      const int lin = i + npl * i3;
      st.ci[u] = s3 ? -1 : (bd.cinv ? bd.cinv[lin] : lin);
      st.fi[u] = bd.finv ? bd.finv[lin] : lin;
#pragma unroll
      for (int s = 0; s < 4; ++s)
        st.k[u][s] = Kp[i1 + N1 * (s & 1) + W1 * (i2 + N2 * (s >> 1))];
    }
  }
  // s3 = 0: coarse value s = 0, fine components 1-3 from s = 1..3;
  // s3 = 1: fine components 4-7 from s = 0..3
  const int sf = 1 - s3;
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const double* f = hf + 7LL * max(st.fi[u], 0) - sf;
    st.old[u][0] = st.ci[u] >= 0 ? hc[max(st.ci[u], 0)]
                                 : (s3 && st.fi[u] >= 0 ? f[0] : 0.0);
#pragma unroll
    for (int s = 1; s < 4; ++s) st.old[u][s] = st.fi[u] >= 0 ? f[s] : 0.0;
  }
}

template <typename TS, int U, int NT, typename TO, typename TK>
__device__ __forceinline__ void compress_store(const BoxDesc& bd, int j3, const TO* O, int RO,
                                               double* hpsi, int i0,
                                               const CompressStep<TK, U>& st) {
  const int N1 = bd.N[0], N2 = bd.N[1], N3 = bd.N[2];
  const int s3 = j3 >= N3;
  const int npl = N1 * N2;
  double* hc = hpsi + bd.psi_off;
  double* hf = hpsi + bd.psi_off + bd.nvctr_c + 3 * s3;
  const float in1 = 1.0f / (float)N1;
  const int sf = 1 - s3;
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int i = i0 + u * NT;
    if (i < npl) {
      const int i2 = qdiv(i, N1, in1), i1 = i - i2 * N1;
      double v[4];
#pragma unroll
      for (int s = 0; s < 4; ++s)
        v[s] = (double)((TS)O[(i2 + N2 * (s >> 1)) * RO + i1 + N1 * (s & 1)] +
                        (TS)st.k[u][s]);
      double* f = hf + 7LL * max(st.fi[u], 0) - sf;
      if (st.ci[u] >= 0) hc[st.ci[u]] = st.old[u][0] + v[0];
      if (st.fi[u] >= 0) {
        if (s3) f[0] = st.old[u][0] + v[0];
#pragma unroll
        for (int s = 1; s < 4; ++s) f[s] = st.old[u][s] + v[s];
      }
    }
  }
}

template <typename TS, int U, int NT, typename TO, typename TK>
__device__ __forceinline__ void compress_plane(const BoxDesc& bd, int j3, const TO* O, int RO,
                                               const TK* Kp, double* hpsi) {
  const int npl = bd.N[0] * bd.N[1];
  for (int i0 = threadIdx.x; i0 < npl; i0 += NT * U) {
    CompressStep<TK, U> st;
    compress_load<U, NT>(bd, j3, Kp, hpsi, i0, st);
    compress_store<TS, U, NT>(bd, j3, O, RO, hpsi, i0, st);
  }
}

// ---------------------------------------------------------------------------
// Filter tables (values of the reference; host copies live in bdft_filters.cpp)
// fil2(j,s), j=-14..16: bigdft/src/orbitals/v.inc:13-76
// kinetic a,c,e:        bigdft/src/orbitals/convolut_ib_optim.f90:962-1038
// ---------------------------------------------------------------------------
struct Fil {
  static constexpr double F2S[31] = {
      -9.2250149921474082015e-9, 1.739330656197903949e-7,
      -1.5578172714074496288e-6, 9.7488426741565769756e-6,
      -0.000028480442793096733611, -0.000034162902320983755896,
      0.00054731040109973893777, -0.0034543155764822098308,
      -0.0043299995331208240697, 0.031835414345970668137,
      0.023424614735397134395, -0.13852477567405100085,
      -0.11023008167833446826, 0.41606693676621257536,
      0.78046048025074402849, 0.4259879826969471933,
      -0.014743648317973476936, -0.031312897214485659659,
      0.0437933264161550095, 0.0080208847559691655459,
      -0.01336723942112679464, -0.0017135145673762282241,
      0.0015547744749237938993, 0.00023987828928177507517,
      0.000025057266882581997398, -0.00001414752521104005874,
      2.0755675626515876328e-6, -4.0033729064957944938e-7,
      1.5691554225771085691e-7, -2.4646355857433108359e-8,
      1.5938753881171638248e-9};
  static constexpr double F2W[31] = {
      -5.1545464563622292493e-9, 9.7186407588266778366e-8,
      -8.7796306212502899553e-7, 5.5890429568834202313e-6,
      -0.000017169835895815067013, -0.000011402461428743860613,
      0.00028479503466830096257, -0.0019699454319373247538,
      -0.0019602139595384833547, 0.015281781287179308495,
      0.0072888369266039604623, -0.039920047232423474172,
      -0.052935814515665507352, 0.049247868599583189247,
      0.40346722112554718347, -0.78110468414482695849,
      0.43810509543421050422, 0.077858113707232306275,
      -0.11950673101135450945, -0.022857677317437822508,
      0.028206423027424364667, 0.0037436631993987390831,
      -0.0027213277354099953613, -0.0004645045737819340081,
      -0.000040668222535397420851, 0.000024486891695262527278,
      -3.3186107540444071037e-6, 6.5211830898659553296e-7,
      -2.7666729880873517609e-7, 4.4109215856612868616e-8,
      -2.8525350339693075057e-9};
  static constexpr double KA[15] = {
      -3.5536922899131901941296809374, 2.2191465938911163898794546405,
      -0.6156141465570069496314853949, 0.2371780582153805636239247476,
      -0.0822663999742123340987663521, 0.02207029188482255523789911295638968409,
      -0.409765689342633823899327051188315485e-2,
      0.45167920287502235349480037639758496e-3,
      -0.2398228524507599670405555359023135e-4,
      2.0904234952920365957922889447361e-6,
      -3.7230763047369275848791496973044e-7,
      -1.05857055496741470373494132287e-8,
      -5.813879830282540547959250667e-11,
      2.70800493626319438269856689037647576e-13,
      -6.924474940639200152025730585882e-18};
  static constexpr double KC[29] = {
      -3.869102413147656535541850057188e-18,
      1.5130616560866154733900029272077362e-13,
      -3.2264702314010525539061647271983988409e-11,
      -5.96264938781402337319841002642e-9,
      -2.1656830629214041470164889350342e-7,
      8.7969704055286288323596890609625e-7,
      -0.00001133456724516819987751818232711775,
      0.00021710795484646138591610188464622454,
      -0.0021356291838797986414312219042358542,
      0.00713761218453631422925717625758502986,
      -0.0284696165863973422636410524436931061,
      0.14327329352510759457155821037742893841,
      -0.42498050943780130143385739554118569733,
      0.65703074007121357894896358254040272157,
      -0.42081655293724308770919536332797729898,
      -0.21716117505137104371463587747283267899,
      0.63457035267892488185929915286969303251,
      -0.53298223962800395684936080758073568406,
      0.23370490631751294307619384973520033236,
      -0.05657736973328755112051544344507997075,
      0.0080872029411844780634067667008050127,
      -0.00093423623304808664741804536808932984,
      0.00005075807947289728306309081261461095,
      -4.62561497463184262755416490048242e-6,
      6.3919128513793415587294752371778e-7,
      1.87909235155149902916133888931e-8,
      1.04757345962781829480207861447155543883e-10,
      -4.84665690596158959648731537084025836e-13,
      1.2392629629188986192855777620877e-17};
  static constexpr double KE[15] = {
      -24.875846029392331358907766562, -7.1440597663471719869313377994,
      -0.04251705323669172315864542163525830944,
      -0.26995931336279126953587091167128839196,
      0.08207454169225172612513390763444496516,
      -0.02207327034586634477996701627614752761,
      0.00409765642831595181639002667514310145,
      -0.00045167920287507774929432548999880117,
      0.00002398228524507599670405555359023135,
      -2.0904234952920365957922889447361e-6,
      3.7230763047369275848791496973044e-7,
      1.05857055496741470373494132287e-8,
      5.8138798302825405479592506674648873655e-11,
      -2.70800493626319438269856689037647576e-13,
      6.924474940639200152025730585882e-18};
};

// Launch arguments of one batched local-H application.
struct Batch {
  const BoxDesc* box;   // device, nbox entries
  int nbox;
  const double* psi;    // compressed coefficients (all orbitals)
  double* hpsi;         // accumulated output, same layout as psi
  const double* pot;    // potential(s)
  void* W;              // wavelet volumes (uncompressed psi, later output)
  void* K;              // kinetic accumulator volumes (wavelet layout)
  void* T1;             // (L1,2N2,2N3) volumes
  void* T2;             // (L1,L2,2N3) volumes
  void* T3;             // (L1,L2,L3) volumes (long z lines only)
                        // (element type: double, or float for mixed precision)
  double* part;         // energy partial sums, one slot per (pass, tile)
  double* ekin;         // per-box kinetic energy (device)
  double* epot;         // per-box potential energy (device)
  const OpTab* tabs;    // matrix-core operator tables (device)
  const void* tfrag;    // operator fragments, matrix-core operand order
  const uint2* tmq;     // per 32-row block: {K-step mask, first fragment}
  unsigned* bmax;       // per box: largest |psi| (float bits), fp16 scaling
  float* tmax;          // per uncompress tile: largest |psi|
  const float* trs;     // inverse row scales of the fp16 G tables (OpTab::soff)
  // 1 (F16X3 plans): the forward plane pass leaves in T2 the z line at
  // (p1, p2) times G16_OUT / (u_x(p1) u_y(p2)), u the row scales of the x
  // and y G tables; the z potential pass removes that factor as it loads
  // the line
  int t2_scaled;
};

// the factor a z line (p1, p2) of box bd carries in T2 when t2_scaled
__device__ __forceinline__ float t2_line_unscale(const Batch& B, const BoxDesc& bd, int p1,
                                                 int p2) {
  const OpTab tx = B.tabs[bd.opG[0]], ty = B.tabs[bd.opG[1]];
  const float ux = (tx.soff >= 0 && p1 < tx.rows) ? B.trs[tx.soff + p1] : 1.0f;
  const float uy = (ty.soff >= 0 && p2 < ty.rows) ? B.trs[ty.soff + p2] : 1.0f;
  return ux * uy * (1.0f / G16_OUT);
}

// Passes of one application; each has its own tile table.
// P_KZG, P_ZPG: z-line passes of the matrix-core sequence, zg (2 or 3)
// bundles of 32 lines per tile, chosen per plan (MfmaLds::zg).
// P_ZP: z lines by rows of x (32 lines of one y per tile); P_ZPF: the same
// lines as one run x + L1 y, 32 consecutive lines per tile (a tile may span
// two rows), so that only a box's last tile is partly empty (k_zpot).
enum Pass { P_UNC = 0, P_XF, P_YF, P_KY, P_KZ, P_ZP, P_YB, P_XB, P_PF, P_PB, P_KZG, P_ZPG,
            P_ZPF, P_COUNT };
constexpr int ZGROUP = 3;   // bundles per z-line tile unless LDS asks for 2

struct PassTables {
  TileTable t[P_COUNT];
  int ntiles[P_COUNT];   // total tiles (grid size) of each pass
  double* part[4];       // energy partials: xfwd|plane_fwd, kin_y, kin_z, zpot
  int plane;             // 1: plane-fused sequence, 0: line passes
  size_t lds_pf, lds_pb; // LDS bytes of the plane kernels
  // matrix-core plane kernels: boxes needing more than half a CU's LDS get
  // launches of their own (t[P_PF], t[P_PB] then cover the other boxes)
  TileTable pf2, pb2;
  int ntiles_pf2, ntiles_pb2;
  // FP32 backward plane pass of the matrix-core sequence: the boxes in up
  // to four groups by how many of their work-groups fit a CU's LDS, one
  // launch per group
  TileTable vbg[4];
  int nt_vbg[4];
  int n_vbg;
  // forward plane pass of the FP64 / FP32 plane-fused sequence: the boxes
  // in groups by how many of their work-groups fit a CU (LDS and threads),
  // one launch per group; the kinetic partials keep t[P_PF]'s tile numbers
  TileTable pfg[4];
  int nt_pfg[4];
  size_t lds_pfg[4];
  int n_pfg;
  // backward plane pass of the FP64 / FP32 plane-fused sequence, grouped
  // the same way; fit_pbg: work-groups per CU of each group (an FP64 group
  // of two runs the kernel bounded to six waves per SIMD)
  TileTable pbg[4];
  int nt_pbg[4];
  size_t lds_pbg[4];
  int fit_pbg[4];
  int n_pbg;
  // preconditioner: x and y operator in one plane pass (FP64 planes that fit)
  int pc_plane;
  size_t lds_pc;
  // lines beyond the register-held outputs: x passes in their BIG forms
  // (some N1 > REG_N or L1 > REG_L), z potential pass in two kernels through
  // T3 (some L3 > REG_L)
  int xbig, zbig;
  // z potential pass as a persistent launch (BDFT_OPT_PERSIST): the P_ZP
  // tiles in work-queue order and the grid of resident work-groups (0: one
  // work-group per tile)
  const int* zq;
  int zq_grid;
  // every z line of the batch within ZSHORT_L points: the z potential pass
  // holds fewer fine chunks per thread (fewer registers, same arithmetic)
  int zshort;
};

// Launch stages in order; profiling records one event pair per stage.
enum Stage { S_ZERO = 0, S_UNC, S_XF, S_YF, S_KY, S_KZ, S_ZP, S_YB, S_XB,
             S_COMP, S_ENERGY, S_PF, S_PB, S_COUNT };

struct EnergyTables {
  TileTable t[4];       // partial sums of ekin (first nk entries) then epot
  const double* p[4];
  int nk, n;
};

struct MfmaLds {
  size_t kin[3], grow, zpot, shrink;  // per-pass sequence
  size_t fwd, bwd, kz, zp;            // plane-fused sequence
  int fused;                          // 1: plane-fused sequence
  int zg;                             // bundles per z-line tile (2 or 3)
  size_t fwd2, bwd2;                  // plane kernels of the large-box launches
  int zvalu;                          // 1: z potential pass by the FP32 line kernel
  int maxN;                           // largest coarse extent (line-kernel LDS)
  int bvalu;                          // 1: backward pass by the FP32 plane kernel
  size_t vbwd, vbwd2;                 // its LDS (box group of pb / pb2)
  size_t lds_vbg[4];                  // its LDS per occupancy group (vbg)
};

// density: orbitals sharing one box are split into ngroups groups of gsize;
// tile = group * nbund + z-line bundle; part[(g*nspin + s)*nisf + p] holds the
// group's partial density, reduced in group order into rho
struct DenArgs {
  const double* hfac;   // per orbital, kwgt*occup/volume_element
  const int* spin;      // per orbital, 0 or 1
  double* part;
  double* rho;          // nspin x nisf, accumulated
  long long nisf;
  int norb, nspin, ngroups, gsize, nbund, nxt;
};

// region transfers: one descriptor per local region, tiles of 256 points
// (coarse then fine) starting at tile0
struct XferDesc {
  const int* cmap;      // local compressed -> local grid (0-based linear)
  const int* fmap;
  int nvctr_c, nvctr_f;
  int m1, m2;           // local points per axis 1, 2 (n + 1)
  int ns[3];            // local box origin in the global box
  long long loff, goff; // offsets of the local and global orbitals
  int tile0;
};
struct XferArgs {
  const XferDesc* d;
  int nd;
  const int* gcinv;     // global grid -> compressed (null: identity)
  const int* gfinv;
  int g1, g2, g3;       // global points per axis
  int gnvctr_c;
  const double* src;
  double* dst;
  int to_local;         // 1: local = global(map), 0: global(map) = local
  int* miss;            // local points with no global counterpart
};
hipError_t launch_xfer(const XferArgs& A, int ntiles, hipStream_t s);
long long xfer_tiles(int nvctr_c, int nvctr_f);

// conjugate-gradient preconditioner, apply = S (K + c) S (bdft_pass.hip)
struct PcArgs {
  const double* pc;  // per box: c, scal[0..7] (9 doubles)
  double* x;         // solution, in/out
  const double* b;   // right-hand side
  double* r;         // residual
  double* d;         // search direction
  double* q;         // A d
  double* part;      // per compressed tile partial sums
  double* sc;        // per box: ns, alpha, beta, unused
  int ncong;
};
hipError_t launch_precond(const Batch& B, const PassTables& T, int maxN, long long wtotal,
                          bool need_zero, const PcArgs& A, hipStream_t s);

// convolution bounds of a free-BC region (bdft_bounds.hip)
constexpr int BOUNDS_NARR = 17;
constexpr int BOUNDS_MAXN = 484;  // values stay below the 1000 sentinel
int bounds_layout(const int n[3], const int nfl[3], const int nfu[3], int64_t off[18],
                  int32_t dims[68]);
struct BoundsArgs {
  int n[3], nfl[3], nfu[3];
  const int* cinv;      // grid point -> coarse coefficient (null: all present)
  const int* finv;      // grid point -> fine coefficient (null: all present)
  int32_t* out;         // bounds_layout buffer
};
hipError_t launch_bounds(const BoundsArgs& A, hipStream_t s);

size_t lds_bytes(int maxN, int elem);
// raise the dynamic-LDS limit of the line kernels to lds bytes
void line_attributes(size_t lds);
hipError_t launch_density(const Batch& B, const PassTables& T, const DenArgs& D, int maxN,
                          long long wtotal, bool need_zero, int fp32, hipStream_t s);
void extern_zero_float(float* p, long long n, int nblocks, hipStream_t s);
void launch_uncompress_float(const Batch& B, const TileTable& t, int ntiles, bool track_max,
                             hipStream_t s);
void launch_energy(const EnergyTables& et, const Batch& B, hipStream_t s);
size_t mfma_lds_kin(int Wd, int Kp);
size_t mfma_lds_grow(int W2, int L1, int Kp1, int Kp2);
size_t mfma_lds_zpot(int KpG, int KpS);
size_t mfma_lds_shrink(int L1, int W1, int W2, int KpU, int KpB);
size_t mfma_lds_fwd(int W1, int W2, int L1, int parts);
size_t mfma_lds_kz(int W3, int Kp, int zg);
size_t mfma_lds_zp(int KpG, int DpG, int zg);
size_t mfma_lds_bwd(int W1, int L1, int L2);
size_t mfma_attributes(size_t lds);
// the FP32 z-line potential pass over its tiles (k_zpot<float> on P_ZPF, or
// the two kernels through T3 for lines beyond REG_L), in place on T2, epot
// partials per tile
hipError_t launch_zpot_f32(const Batch& B, const PassTables& PT, int maxN, hipStream_t s);
// tile table of the z potential pass's epot partials (P_ZPF for k_zpot,
// P_ZP for the two-kernel form of long lines)
inline const TileTable& zpot_table(const PassTables& PT) {
  return PT.zbig ? PT.t[P_ZP] : PT.t[P_ZPF];
}
// resident work-groups of the persistent z potential pass on the current
// device (a multiple of the 8 XCDs), for working type FP32 or FP64
int zpot_queue_grid(int fp32, int maxN);
// a second stream and two events: the z kinetic and z potential passes,
// which read and write disjoint grids, run side by side between the forward
// and the backward plane passes (fork after one, join before the other)
struct Fork {
  hipStream_t s2;
  hipEvent_t a, b;
};
hipError_t launch_mfma(const Batch& B, const PassTables& PT, const MfmaLds& L,
                       long long wtotal, bool need_zero, int precision, hipStream_t s,
                       hipEvent_t* ev, const Fork* fk = nullptr);
size_t plane_lds_fwd(int N1, int N2, int L1, int elem);
size_t plane_lds_bwd(int N1, int N2, int L1, int L2, int elem);
// preconditioner: K = c W + K_x W + K_y W per x-y plane (P_PF tiles)
size_t plane_lds_kin(int N1, int N2, int elem);
hipError_t launch_plane_kin_pc(const Batch& B, const TileTable& t, int ntiles, size_t lds,
                               const double* pc, hipStream_t s);
size_t plane_attributes();
template <typename T>
hipError_t launch_plane_fwd(const Batch& B, const TileTable& t, const TileTable& ft,
                            int ntiles, size_t lds, double* part, hipStream_t s);
template <typename T>
hipError_t launch_plane_grow(const Batch& B, const TileTable& t, int ntiles, size_t lds,
                             hipStream_t s);
template <typename T>
hipError_t launch_plane_bwd(const Batch& B, const TileTable& t, int ntiles, size_t lds,
                            hipStream_t s, bool occ2 = false);
hipError_t launch_transfer(const Batch& B, const TileTable& t, int ntiles,
                           long long wtotal, bool compress, hipStream_t s);
// ev: null, or 2*S_COUNT events recorded before/after each stage;
// fp32 selects the mixed-precision (float working type) kernels
hipError_t launch_passes(const Batch& B, const PassTables& T, int maxN,
                         long long wtotal, bool need_zero, int fp32,
                         hipStream_t s, hipEvent_t* ev = nullptr);

}  // namespace bdft
