// Transfers between localisation regions and the global region, in the
// compressed form: psi_to_locreg2 (locregs/locreg_operations.f90:2143-2318)
// and Lpsi_to_global2 (locreg_operations.f90:1665-1837).  The reference walks
// the local segments (shifted into the global box, shift_locreg_indexes,
// :2321-2380) against the global segments; here every local compressed point
// is located through the region map (local grid point + box offset -> global
// grid point -> global compressed index through the inverse map of the
// global region), and every coefficient is one thread: a coarse point moves
// one value, a fine point's seven values are moved by seven consecutive
// threads, so loads and stores of a wave cover contiguous ranges of both
// orbitals.  Same elements, same values, no floating-point arithmetic.
#include "bdft_internal.h"

namespace bdft {

constexpr int XT = 256;        // threads per tile
constexpr int XU = 8;          // coefficients per thread
constexpr int XC = XT * XU;    // coefficients per tile

__device__ __forceinline__ int xfer_find(const XferArgs& A, int tile) {
  int lo = 0, hi = A.nd - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (A.d[mid].tile0 <= tile) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__global__ __launch_bounds__(XT) void k_xfer(XferArgs A) {
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const XferDesc D = A.d[xfer_find(A, tile)];
  const int lt = tile - D.tile0;
  const int ctiles = (D.nvctr_c + XC - 1) / XC;
  const bool fine = lt >= ctiles;
  const int ne = fine ? 7 * D.nvctr_f : D.nvctr_c;
  const int e0 = (fine ? lt - ctiles : lt) * XC + threadIdx.x;
  // coefficient e of the coarse or the fine part (fine: point e / 7,
  // wavelet t); XU of them per thread, all located before any is moved
  long long li[XU], gx[XU];
  bool ok[XU], hit[XU];
#pragma unroll
  for (int u = 0; u < XU; ++u) {
    const int e = e0 + u * XT;
    ok[u] = e < ne;
    hit[u] = false;
    li[u] = gx[u] = 0;
    if (!ok[u]) continue;
    const int k = fine ? e / 7 : e;
    const int t = fine ? e - 7 * k : 0;
    const int lin = fine ? D.fmap[k] : D.cmap[k];
    const int i1 = lin % D.m1, r = lin / D.m1;
    const int i2 = r % D.m2, i3 = r / D.m2;
    const int g1 = i1 + D.ns[0], g2 = i2 + D.ns[1], g3 = i3 + D.ns[2];
    int gi = -1;
    if (g1 >= 0 && g1 < A.g1 && g2 >= 0 && g2 < A.g2 && g3 >= 0 && g3 < A.g3) {
      const int g = g1 + A.g1 * (g2 + A.g2 * g3);
      const int* inv = fine ? A.gfinv : A.gcinv;
      gi = inv ? inv[g] : g;
    }
    hit[u] = gi >= 0;
    li[u] = fine ? D.loff + D.nvctr_c + e : D.loff + k;
    gx[u] = fine ? D.goff + A.gnvctr_c + 7LL * gi + t : D.goff + gi;
    if (!hit[u] && !A.to_local && t == 0) atomicAdd(A.miss, 1);  // once per point
  }
  double v[XU];
#pragma unroll
  for (int u = 0; u < XU; ++u)
    v[u] = hit[u] ? A.src[A.to_local ? gx[u] : li[u]] : 0.0;
#pragma unroll
  for (int u = 0; u < XU; ++u) {
    if (!ok[u]) continue;
    if (A.to_local) A.dst[li[u]] = v[u];
    else if (hit[u]) A.dst[gx[u]] = v[u];
  }
}

hipError_t launch_xfer(const XferArgs& A, int ntiles, hipStream_t s) {
  if (ntiles > 0) k_xfer<<<ntiles, XT, 0, s>>>(A);
  return hipGetLastError();
}

long long xfer_tiles(int nvctr_c, int nvctr_f) {
  return ((long long)nvctr_c + XC - 1) / XC + (7LL * nvctr_f + XC - 1) / XC;
}

}  // namespace bdft
