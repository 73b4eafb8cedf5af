// Convolution bounds of a free-BC region: the 17 integer arrays of
// convolutions_bounds (kinetic, grow, shrink, real space) computed by
// locreg_bounds (bigdft/src/locregs/bounds.f90:555-613) through make_bounds
// (:690-775) and make_all_ib (:777-865).
//
// Each array is ib(2, lo1:hi1, lo2:hi2): the first and last index along one
// axis of the points of a row, 1000/-1000 when the row is empty.  The
// reference builds every array after the first by painting the previous
// array's intervals, widened by the free-BC filter footprint [2i-14, 2i+16],
// into a flat logical grid and reading it back with its axes taken in a
// rotated order.  Here every output element is one thread that scans the
// previous array directly: for its row it looks for the first and the last
// input index whose widened interval covers the row.  Same integers, no
// work grid.  "squares" (2x2 block min/max) and "squares_1d" (pairs along
// the first axis) are one thread per block.  Setup work: O(n^3) integer
// reads per region, run once.
#include "bdft_internal.h"

namespace bdft {

namespace {

constexpr int BT = 256;
constexpr int32_t EMPTY_LO = 1000, EMPTY_HI = -1000;

// ib(2, lo1:lo1+n1-1, lo2:lo2+n2-1), column-major
struct Ib {
  int32_t* p;
  int lo1, n1, lo2, n2;
};

__device__ __forceinline__ int32_t* ib_at(const Ib& v, int i, int j) {
  return v.p + 2 * ((long long)(i - v.lo1) + (long long)v.n1 * (j - v.lo2));
}

__device__ __forceinline__ bool present(const int* inv, long long lin) {
  return inv ? inv[lin] >= 0 : true;
}

// make_bounds for the coarse (blockIdx.y = 0) and fine (1) grids: one thread
// per row along x (ibyz), then along y (ibxz), then along z (ibxy)
__global__ __launch_bounds__(BT) void k_rows(BoundsArgs A, Ib yz_c, Ib xz_c, Ib xy_c,
                                             Ib yz_f, Ib xz_f, Ib xy_f) {
  const bool f = blockIdx.y == 1;
  const int* inv = f ? A.finv : A.cinv;
  const int N1 = A.n[0] + 1, N2 = A.n[1] + 1, N3 = A.n[2] + 1;
  long long t = (long long)blockIdx.x * BT + threadIdx.x;
  const long long r1 = (long long)N2 * N3, r2 = (long long)N1 * N3,
                  r3 = (long long)N1 * N2;
  int lo = EMPTY_LO, hi = EMPTY_HI;
  int32_t* dst;
  if (t < r1) {
    const int i2 = (int)(t % N2), i3 = (int)(t / N2);
    const long long base = (long long)N1 * (i2 + (long long)N2 * i3);
    for (int i = 0; i < N1; ++i)
      if (present(inv, base + i)) { lo = i; break; }
    for (int i = N1 - 1; i >= 0; --i)
      if (present(inv, base + i)) { hi = i; break; }
    dst = ib_at(f ? yz_f : yz_c, i2, i3);
  } else if ((t -= r1) < r2) {
    const int i1 = (int)(t % N1), i3 = (int)(t / N1);
    const long long base = i1 + (long long)N1 * N2 * i3;
    for (int i = 0; i < N2; ++i)
      if (present(inv, base + (long long)N1 * i)) { lo = i; break; }
    for (int i = N2 - 1; i >= 0; --i)
      if (present(inv, base + (long long)N1 * i)) { hi = i; break; }
    dst = ib_at(f ? xz_f : xz_c, i1, i3);
  } else if ((t -= r2) < r3) {
    const int i1 = (int)(t % N1), i2 = (int)(t / N1);
    const long long base = i1 + (long long)N1 * i2, st = (long long)N1 * N2;
    for (int i = 0; i < N3; ++i)
      if (present(inv, base + st * i)) { lo = i; break; }
    for (int i = N3 - 1; i >= 0; --i)
      if (present(inv, base + st * i)) { hi = i; break; }
    dst = ib_at(f ? xy_f : xy_c, i1, i2);
  } else {
    return;
  }
  dst[0] = lo;
  dst[1] = hi;
}

__device__ __forceinline__ bool covers(const int32_t* v, int C) {
  return C >= 2 * v[0] - 14 && C <= 2 * v[1] + 16;
}

// rot (grow chain, ib_to_logrid_rot + ib_from_logrid):
//   out(b, C) = first/last a with C in footprint(in(a, b))
// inv (shrink chain, ib_to_logrid_inv + ib_from_logrid_inv):
//   out(C, a) = first/last b with C in footprint(in(a, b))
__global__ __launch_bounds__(BT) void k_spread(Ib in, Ib out, int inv_mode) {
  const long long t = (long long)blockIdx.x * BT + threadIdx.x;
  if (t >= (long long)out.n1 * out.n2) return;
  const int o1 = out.lo1 + (int)(t % out.n1), o2 = out.lo2 + (int)(t / out.n1);
  int lo = EMPTY_LO, hi = EMPTY_HI;
  if (!inv_mode) {
    const int b = o1, C = o2;
    for (int a = in.lo1; a < in.lo1 + in.n1; ++a)
      if (covers(ib_at(in, a, b), C)) { lo = a; break; }
    for (int a = in.lo1 + in.n1 - 1; a >= in.lo1; --a)
      if (covers(ib_at(in, a, b), C)) { hi = a; break; }
  } else {
    const int C = o1, a = o2;
    for (int b = in.lo2; b < in.lo2 + in.n2; ++b)
      if (covers(ib_at(in, a, b), C)) { lo = b; break; }
    for (int b = in.lo2 + in.n2 - 1; b >= in.lo2; --b)
      if (covers(ib_at(in, a, b), C)) { hi = b; break; }
  }
  int32_t* d = ib_at(out, o1, o2);
  d[0] = lo;
  d[1] = hi;
}

// squares (bounds.f90:1097-1123): 2x2 blocks from the array's first element
__global__ __launch_bounds__(BT) void k_squares(Ib v, int nb1, int nb2) {
  const long long t = (long long)blockIdx.x * BT + threadIdx.x;
  if (t >= (long long)nb1 * nb2) return;
  const int j1 = 2 * (int)(t % nb1), j2 = 2 * (int)(t / nb1);
  int32_t* q[4] = {v.p + 2 * (j1 + (long long)v.n1 * j2),
                   v.p + 2 * (j1 + 1 + (long long)v.n1 * j2),
                   v.p + 2 * (j1 + (long long)v.n1 * (j2 + 1)),
                   v.p + 2 * (j1 + 1 + (long long)v.n1 * (j2 + 1))};
  int mn = q[0][0], mx = q[0][1];
#pragma unroll
  for (int k = 1; k < 4; ++k) {
    mn = min(mn, q[k][0]);
    mx = max(mx, q[k][1]);
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    q[k][0] = mn;
    q[k][1] = mx;
  }
}

// squares_1d (bounds.f90:1069-1091): pairs (first+2p, first+2p+1) of the
// first axis, every column
__global__ __launch_bounds__(BT) void k_squares_1d(Ib v, int first, int npairs) {
  const long long t = (long long)blockIdx.x * BT + threadIdx.x;
  if (t >= (long long)npairs * v.n2) return;
  const int j1 = first + 2 * (int)(t % npairs), j2 = (int)(t / npairs);
  int32_t* a = v.p + 2 * (j1 + (long long)v.n1 * j2);
  int32_t* b = a + 2;
  const int mn = min(a[0], b[0]), mx = max(a[1], b[1]);
  a[0] = b[0] = mn;
  a[1] = b[1] = mx;
}

// dst(i, j) = src(i, j) over dst's index range (ibxy_ff, ibyz_ff)
__global__ __launch_bounds__(BT) void k_copy_box(Ib src, Ib dst) {
  const long long t = (long long)blockIdx.x * BT + threadIdx.x;
  if (t >= (long long)dst.n1 * dst.n2) return;
  const int i = dst.lo1 + (int)(t % dst.n1), j = dst.lo2 + (int)(t / dst.n1);
  const int32_t* s = ib_at(src, i, j);
  int32_t* d = ib_at(dst, i, j);
  d[0] = s[0];
  d[1] = s[1];
}

// ibyyzz_r from ibyyzz_c (bounds.f90:821-833): fine-grid extent of the
// coarse interval in the 0-based psir index, before squares
__global__ __launch_bounds__(BT) void k_real(Ib c, Ib r) {
  const long long t = (long long)blockIdx.x * BT + threadIdx.x;
  if (t >= (long long)r.n1 * r.n2) return;
  const int32_t* s = c.p + 2 * t;
  int32_t* d = r.p + 2 * t;
  if (s[0] != EMPTY_LO) {
    d[0] = 2 * s[0];
    d[1] = 2 * s[1] + 30;
  } else {
    d[0] = EMPTY_LO;
    d[1] = EMPTY_HI;
  }
}

unsigned nblk(long long n) { return (unsigned)((n + BT - 1) / BT); }

bool inside(const Ib& outer, int lo1, int n1, int lo2, int n2) {
  return lo1 >= outer.lo1 && lo1 + n1 <= outer.lo1 + outer.n1 && lo2 >= outer.lo2 &&
         lo2 + n2 <= outer.lo2 + outer.n2;
}

}  // namespace

int bounds_layout(const int n[3], const int nfl[3], const int nfu[3], int64_t off[18],
                  int32_t dims[68]) {
  for (int d = 0; d < 3; ++d)
    if (n[d] < 0 || nfl[d] > nfu[d]) return 1;
  const int n1 = n[0], n2 = n[1], n3 = n[2];
  const int l1 = nfl[0], l2 = nfl[1], l3 = nfl[2], u1 = nfu[0], u2 = nfu[1], u3 = nfu[2];
  // extents of the arrays allocated by locreg_bounds (bounds.f90:572-601),
  // in the order of the kinetic, grow and shrink types, then ibyyzz_r
  const int32_t e[BOUNDS_NARR][4] = {
      {0, n2, 0, n3}, {0, n1, 0, n3}, {0, n1, 0, n2},
      {0, n2, 0, n3}, {0, n1, 0, n3}, {0, n1, 0, n2},
      {0, n3, -14, 2 * n1 + 16},
      {-14, 2 * n1 + 16, -14, 2 * n2 + 16},
      {l2, u2, l3, u3},
      {l3, u3, 2 * l1 - 14, 2 * u1 + 16},
      {2 * l1 - 14, 2 * u1 + 16, 2 * l2 - 14, 2 * u2 + 16},
      {-14, 2 * n3 + 16, 0, n1},
      {-14, 2 * n2 + 16, -14, 2 * n3 + 16},
      {l1, u1, l2, u2},
      {2 * l3 - 14, 2 * u3 + 16, l1, u1},
      {2 * l2 - 14, 2 * u2 + 16, 2 * l3 - 14, 2 * u3 + 16},
      {-14, 2 * n2 + 16, -14, 2 * n3 + 16}};
  int64_t o = 0;
  for (int k = 0; k < BOUNDS_NARR; ++k) {
    off[k] = o;
    for (int j = 0; j < 4; ++j) dims[4 * k + j] = e[k][j];
    o += 2LL * (e[k][1] - e[k][0] + 1) * (e[k][3] - e[k][2] + 1);
  }
  off[BOUNDS_NARR] = o;
  return 0;
}

hipError_t launch_bounds(const BoundsArgs& A, hipStream_t s) {
  int64_t off[BOUNDS_NARR + 1];
  int32_t dims[4 * BOUNDS_NARR];
  if (bounds_layout(A.n, A.nfl, A.nfu, off, dims)) return hipErrorInvalidValue;
  Ib v[BOUNDS_NARR];
  for (int k = 0; k < BOUNDS_NARR; ++k)
    v[k] = Ib{A.out + off[k], dims[4 * k], dims[4 * k + 1] - dims[4 * k] + 1,
              dims[4 * k + 2], dims[4 * k + 3] - dims[4 * k + 2] + 1};
  const int n1 = A.n[0], n2 = A.n[1], n3 = A.n[2];
  // index ranges that the kernels take from one array into another
  const bool ok = inside(v[5], v[13].lo1, v[13].n1, v[13].lo2, v[13].n2) &&
                  inside(v[3], v[8].lo1, v[8].n1, v[8].lo2, v[8].n2) &&
                  v[11].lo2 == v[2].lo1 && v[11].n2 == v[2].n1 &&     // inv: a range
                  v[12].lo2 == v[11].lo1 && v[12].n2 == v[11].n1 &&
                  v[14].lo2 == v[13].lo1 && v[14].n2 == v[13].n1 &&
                  v[15].lo2 == v[14].lo1 && v[15].n2 == v[14].n1 &&
                  v[6].lo1 == v[0].lo2 && v[6].n1 == v[0].n2 &&       // rot: b range
                  v[7].lo1 == v[6].lo2 && v[7].n1 == v[6].n2 &&
                  v[9].lo1 == v[8].lo2 && v[9].n1 == v[8].n2 &&
                  v[10].lo1 == v[9].lo2 && v[10].n1 == v[9].n2 &&
                  v[16].n1 == v[12].n1 && v[16].n2 == v[12].n2;
  if (!ok) return hipErrorInvalidValue;
  auto spread = [&](int in, int out, int inv_mode) {
    hipLaunchKernelGGL(k_spread, dim3(nblk((long long)v[out].n1 * v[out].n2)), dim3(BT), 0,
                       s, v[in], v[out], inv_mode);
  };
  auto squares = [&](int k) {
    const int d1 = v[k].n1 - 1, d2 = v[k].n2 - 1;
    if (d1 == 0 || d2 == 0) return;
    const int nb1 = (d1 - 1) / 2 + 1, nb2 = (d2 - 1) / 2 + 1;
    hipLaunchKernelGGL(k_squares, dim3(nblk((long long)nb1 * nb2)), dim3(BT), 0, s, v[k],
                       nb1, nb2);
  };
  auto copy = [&](int src, int dst) {
    hipLaunchKernelGGL(k_copy_box, dim3(nblk((long long)v[dst].n1 * v[dst].n2)), dim3(BT),
                       0, s, v[src], v[dst]);
  };
  // make_bounds, coarse and fine
  const long long rows = (long long)(n2 + 1) * (n3 + 1) + (long long)(n1 + 1) * (n3 + 1) +
                         (long long)(n1 + 1) * (n2 + 1);
  hipLaunchKernelGGL(k_rows, dim3(nblk(rows), 2), dim3(BT), 0, s, A, v[0], v[1], v[2], v[3],
                     v[4], v[5]);
  // shrink (make_ib_inv, coarse then fine)
  copy(5, 13);
  spread(2, 11, 1);
  spread(11, 12, 1);
  spread(13, 14, 1);
  spread(14, 15, 1);
  // real space
  hipLaunchKernelGGL(k_real, dim3(nblk((long long)v[16].n1 * v[16].n2)), dim3(BT), 0, s,
                     v[12], v[16]);
  squares(16);
  // grow: make_ib_c squares the coarse kinetic ibyz in place before use
  copy(3, 8);
  squares(0);
  spread(0, 6, 0);
  squares(6);
  spread(6, 7, 0);
  squares(7);
  spread(8, 9, 0);
  spread(9, 10, 0);
  // squares_1d over the first axis of ibxxyy_f, index lo1 = 2 nfl1 - 14
  // (even): pairs j = lo1/2 .. (hi1-1)/2 with C (truncating) division
  {
    const int lo = v[10].lo1, hi = v[10].lo1 + v[10].n1 - 1;
    const int jlo = lo / 2, jhi = (hi - 1) / 2;
    const int npairs = jhi - jlo + 1, first = 2 * jlo - lo;
    if (npairs > 0) {
      if (first < 0 || first + 2 * npairs > v[10].n1) return hipErrorInvalidValue;
      hipLaunchKernelGGL(k_squares_1d, dim3(nblk((long long)npairs * v[10].n2)), dim3(BT), 0,
                         s, v[10], first, npairs);
    }
  }
  (void)n1;
  return hipGetLastError();
}

}  // namespace bdft
