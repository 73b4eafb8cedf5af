// Transfers between localisation regions and the global region, in the
// compressed form: psi_to_locreg2 (locregs/locreg_operations.f90:2143-2318)
// and Lpsi_to_global2 (locreg_operations.f90:1665-1837).  The reference walks
// the local segments (shifted into the global box, shift_locreg_indexes,
// :2321-2380) against the global segments; here every local compressed point
// is one thread: local grid point (region map) + box offset -> global grid
// point -> global compressed index (inverse map of the global region), then
// one gather or scatter of 1 (coarse) or 7 (fine) values.  Same elements,
// same values, no floating-point arithmetic.
#include "bdft_internal.h"

namespace bdft {

constexpr int XT = 256;   // points per tile

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
  const int tile = blockIdx.x;
  const XferDesc D = A.d[xfer_find(A, tile)];
  const int lt = tile - D.tile0;
  const int ctiles = (D.nvctr_c + XT - 1) / XT;
  const bool fine = lt >= ctiles;
  const int k = (fine ? lt - ctiles : lt) * XT + threadIdx.x;
  if (k >= (fine ? D.nvctr_f : D.nvctr_c)) return;
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
  if (!fine) {
    if (A.to_local) {
      A.dst[D.loff + k] = gi >= 0 ? A.src[D.goff + gi] : 0.0;
    } else if (gi >= 0) {
      A.dst[D.goff + gi] = A.src[D.loff + k];
    } else {
      atomicAdd(A.miss, 1);
    }
  } else {
    const long long l0 = D.loff + D.nvctr_c + 7LL * k;
    const long long g0 = D.goff + A.gnvctr_c + 7LL * gi;
    if (A.to_local) {
#pragma unroll
      for (int t = 0; t < 7; ++t) A.dst[l0 + t] = gi >= 0 ? A.src[g0 + t] : 0.0;
    } else if (gi >= 0) {
#pragma unroll
      for (int t = 0; t < 7; ++t) A.dst[g0 + t] = A.src[l0 + t];
    } else {
      atomicAdd(A.miss, 1);
    }
  }
}

hipError_t launch_xfer(const XferArgs& A, int ntiles, hipStream_t s) {
  if (ntiles > 0) k_xfer<<<ntiles, XT, 0, s>>>(A);
  return hipGetLastError();
}

int xfer_tiles(int nvctr_c, int nvctr_f) {
  return (nvctr_c + XT - 1) / XT + (nvctr_f + XT - 1) / XT;
}

}  // namespace bdft
