// C-ABI layer: contexts, localisation regions (descriptor validation and
// upload), batch plans and their hipGraph capture.
#include <algorithm>
#include <map>
#include <tuple>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/bigdft_hip.h"
#include "bdft_internal.h"

using namespace bdft;


struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  bool owned = true;   // false: a slice of a fixed device array (Arena)
};

// A plan made without allocations takes its buffers from one device array
// of fixed size and uploads its descriptors with stream-ordered copies; the
// host copies of those descriptors are kept in `keep` until the copies have
// been read.
struct Arena {
  char* base = nullptr;
  size_t size = 0, used = 0;
  hipStream_t stream = nullptr;
  std::vector<std::vector<unsigned char>>* keep = nullptr;
  bool overflow = false;  // a request did not fit
};

struct bdft_ctx {
  int device = 0;
  hipStream_t own = nullptr;
  hipStream_t stream = nullptr;
  std::string err;
  // grow-only workspace used by one-shot calls
  DevBuf ws;
  // cache of the last one-shot plan
  bdft_plan* cached = nullptr;
  std::vector<uintptr_t> cached_key;
  // preconditioner: plan of the last batch geometry and its vectors
  bdft_plan* pc_plan = nullptr;
  std::vector<uintptr_t> pc_key;
  DevBuf pc_buf;
  // density: plan of the last (region, orbital count, precision) and its
  // per-group partial sums
  bdft_plan* den_plan = nullptr;
  std::vector<uintptr_t> den_key;
  DevBuf den_buf;
  // region transfers: descriptors and the miss counter, and a host copy of
  // the descriptors last uploaded (a repeated batch skips the upload)
  DevBuf xfer_buf;
  std::vector<unsigned char> xfer_last;
  // stage profiling (non-graph launches)
  bool profile = false;
  std::vector<hipEvent_t> ev_free;                 // reusable events
  std::vector<std::vector<hipEvent_t>> ev_pending; // one set per launch
  std::vector<char> ev_counts;  // per set: 1 when it starts an application
                                // (the later parts of a chunked plan add 0)
  double prof_ms[S_COUNT] = {0};
  long long prof_n = 0;
  bool force_lines = false;  // disable the plane-fused sequence
  int opt_zg = 0;            // z-line bundles per work-group (0: automatic)
  int opt_zpot = 0;          // matrix-core z potential pass: 0 auto, 1 MFMA, 2 FP32
  int opt_bwd = 0;           // matrix-core backward pass: 0 auto, 1 MFMA, 2 FP32
  int opt_chunk_mb = 0;      // workspace per chunk of orbitals (MiB), 0: one chunk
  int opt_order = 0;         // plane and (un)compress tiles: 0 dispatch order, 1 per XCD
  int opt_persist = 0;       // z potential pass: 0 one work-group per tile, 1 work queue
  size_t plane_lds = 0;      // LDS bytes the plane kernels may use
  // device address of this device's fixed array (bdft_locham_batch_fixed_d),
  // from hipGetSymbolAddress at creation; null with the reason in fixed_err
  char* fixed = nullptr;
  std::string fixed_err;
  // regions and captured plans made on this context and not yet destroyed:
  // destroying the context releases their device memory and detaches them
  // (their own destroy calls then only free the host objects)
  std::vector<bdft_lr*> regions;
  std::vector<bdft_plan*> plans;
};

struct bdft_lr {
  bdft_ctx* ctx = nullptr;
  char geocode = 'P';
  int n[3] = {0, 0, 0}, nfl[3] = {0, 0, 0}, nfu[3] = {0, 0, 0};
  int N[3] = {0, 0, 0}, L[3] = {0, 0, 0};
  double hgrid[3] = {0, 0, 0};
  int nvctr_c = 0, nvctr_f = 0;
  int* d_cmap = nullptr;
  int* d_fmap = nullptr;
  int* d_cinv = nullptr;  // null when the coarse map is the identity
  int* d_finv = nullptr;  // null when the fine map is the identity
  bool dense = false;
};

struct bdft_plan {
  bdft_ctx* ctx = nullptr;
  int nbox = 0;
  int precision = 0;
  int maxN = 0;
  long long wtotal = 0;
  bool need_zero = true;
  bool owns_ws = false;
  DevBuf ws;           // own workspace (persistent plans)
  // workspace not bound yet (a part of a chunked plan): bytes it needs and
  // the offsets of W, K, T1, T2, T3 in it
  size_t ws_need = 0;
  size_t ws_at[5] = {0, 0, 0, 0, 0};
  // chunked plan: parts over consecutive orbital ranges, run one after the
  // other on one workspace (ws of this plan) sized for the largest part
  std::vector<bdft_plan*> parts;
  DevBuf meta;         // box descriptors + tile starts + energy partials
  Batch B{};
  PassTables T{};
  MfmaLds mlds{};
  DevBuf tabbuf;       // matrix-core operator tables
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

static int set_err(bdft_ctx* ctx, int code, const char* fmt, ...) {
  if (ctx) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    ctx->err = buf;
  }
  return code;
}

#define HIPCHK(ctx, call)                                                     \
  do {                                                                        \
    hipError_t e_ = (call);                                                   \
    if (e_ != hipSuccess)                                                     \
      return set_err(ctx, BDFT_ERR_HIP, "%s failed: %s", #call,               \
                     hipGetErrorString(e_));                                  \
  } while (0)

static int dev_alloc(bdft_ctx* ctx, DevBuf& b, size_t bytes) {
  if (b.bytes >= bytes && b.p) return BDFT_OK;
  if (b.p) {
    hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
  }
  if (bytes == 0) return BDFT_OK;
  hipError_t e = hipMalloc(&b.p, bytes);
  if (e != hipSuccess) {
    b.p = nullptr;
    return set_err(ctx, BDFT_ERR_NOMEM, "hipMalloc(%zu) failed: %s", bytes,
                   hipGetErrorString(e));
  }
  b.bytes = bytes;
  return BDFT_OK;
}

static void dev_free(DevBuf& b) {
  if (b.p && b.owned) hipFree(b.p);
  b.p = nullptr;
  b.bytes = 0;
  b.owned = true;
}

// a buffer of the plan: from the arena when there is one, else allocated
static int plan_buf(bdft_ctx* ctx, Arena* a, DevBuf& b, size_t bytes) {
  if (!a) return dev_alloc(ctx, b, bytes);
  const size_t at = (a->used + 255) & ~(size_t)255;
  if (at + bytes > a->size) {
    a->overflow = true;
    return set_err(ctx, BDFT_ERR_UNSUPPORTED,
                   "the batch needs more than the %zu bytes of the fixed device array", a->size);
  }
  b.p = a->base + at;
  b.bytes = bytes;
  b.owned = false;
  a->used = at + bytes;
  return BDFT_OK;
}

// host -> device copy of plan data: synchronous, or stream-ordered from a
// host copy kept alive in the arena's keep list
static hipError_t plan_h2d(Arena* a, void* dst, const void* src, size_t n) {
  if (!n) return hipSuccess;
  if (!a) return hipMemcpy(dst, src, n, hipMemcpyHostToDevice);
  a->keep->emplace_back((const unsigned char*)src, (const unsigned char*)src + n);
  return hipMemcpyAsync(dst, a->keep->back().data(), n, hipMemcpyHostToDevice, a->stream);
}

// ---------------------------------------------------------------------------
// Fixed device array of the allocation-free entry point
// (bdft_locham_batch_fixed_d): for each chunk of orbitals in turn it holds
// the plan's descriptors (box table, tile starts, energy partials), the
// matrix-core operator tables and the working grids W, K, T2.  Every device
// the library runs on holds its own copy until the process ends.  The name
// is a host-side stand-in: host code uses only the device address that
// hipGetSymbolAddress returns (bdft_ctx::fixed, looked up per context with
// its device current).
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[BDFT_FIXED_BYTES];

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
extern "C" int bdft_ctx_create(int device, bdft_ctx** out) {
  if (!out) return BDFT_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev)
    return BDFT_ERR_ARG;
  bdft_ctx* ctx = new bdft_ctx();
  ctx->device = device;
  if (hipSetDevice(device) != hipSuccess ||
      hipStreamCreateWithFlags(&ctx->own, hipStreamNonBlocking) != hipSuccess) {
    delete ctx;
    return BDFT_ERR_HIP;
  }
  ctx->stream = ctx->own;
  ctx->plane_lds = plane_attributes();
  if (ctx->plane_lds) {
    mfma_attributes(ctx->plane_lds);
    line_attributes(ctx->plane_lds);
  }
  const char* env = getenv("BDFT_LINE_PASSES");
  ctx->force_lines = env && env[0] == '1';
  // this device's copy of the fixed array (its device is current here)
  void* fx = nullptr;
  const hipError_t fe = hipGetSymbolAddress(&fx, HIP_SYMBOL(scratch));
  if (fe == hipSuccess && fx) {
    ctx->fixed = (char*)fx;
  } else {
    ctx->fixed_err = std::string("hipGetSymbolAddress: ") + hipGetErrorString(fe);
    (void)hipGetLastError();
  }
  *out = ctx;
  return BDFT_OK;
}

extern "C" int bdft_plan_destroy(bdft_plan* plan);
static void plan_release(bdft_plan* plan);
static void lr_release(bdft_lr* lr);

extern "C" int bdft_ctx_destroy(bdft_ctx* ctx) {
  if (!ctx) return BDFT_ERR_ARG;
  hipSetDevice(ctx->device);
  hipStreamSynchronize(ctx->stream);
  for (bdft_plan* p : ctx->plans) {
    plan_release(p);
    p->ctx = nullptr;
  }
  ctx->plans.clear();
  for (bdft_lr* l : ctx->regions) {
    lr_release(l);
    l->ctx = nullptr;
  }
  ctx->regions.clear();
  for (auto& set : ctx->ev_pending)
    for (auto x : set) hipEventDestroy(x);
  for (auto x : ctx->ev_free) hipEventDestroy(x);
  if (ctx->cached) bdft_plan_destroy(ctx->cached);
  if (ctx->pc_plan) bdft_plan_destroy(ctx->pc_plan);
  dev_free(ctx->pc_buf);
  if (ctx->den_plan) bdft_plan_destroy(ctx->den_plan);
  dev_free(ctx->den_buf);
  dev_free(ctx->xfer_buf);
  dev_free(ctx->ws);
  if (ctx->own) hipStreamDestroy(ctx->own);
  delete ctx;
  return BDFT_OK;
}

// The context's buffers (one-shot plan workspace, density, preconditioner
// and transfer buffers) are shared by every call, so the new stream first
// waits for everything queued on the old one: later calls cannot overwrite
// a buffer that earlier work still reads, and synchronising the current
// stream (plan and region destruction) also covers the old stream's work.
extern "C" int bdft_ctx_set_stream(bdft_ctx* ctx, void* s) {
  if (!ctx) return BDFT_ERR_ARG;
  hipStream_t ns = s ? (hipStream_t)s : ctx->own;
  if (ns == ctx->stream) return BDFT_OK;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  hipEvent_t ev = nullptr;
  HIPCHK(ctx, hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  hipError_t e = hipEventRecord(ev, ctx->stream);
  if (e == hipSuccess) e = hipStreamWaitEvent(ns, ev, 0);
  hipEventDestroy(ev);
  if (e != hipSuccess)
    return set_err(ctx, BDFT_ERR_HIP, "stream switch: %s", hipGetErrorString(e));
  ctx->stream = ns;
  return BDFT_OK;
}

extern "C" const char* bdft_last_error(const bdft_ctx* ctx) {
  return ctx ? ctx->err.c_str() : "null context";
}

extern "C" int bdft_sync(bdft_ctx* ctx) {
  if (!ctx) return BDFT_ERR_ARG;
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return BDFT_OK;
}

// ---------------------------------------------------------------------------
// localisation regions: key validation and decode into dense index maps
// (segment decode of uncompress_per, bigdft/src/wavelib/un-compress.f90:903-913)
// ---------------------------------------------------------------------------
static int build_map(bdft_ctx* ctx, const char* what, const int N[3], int nseg,
                     int nvctr, const int32_t* keyg, const int32_t* keyv,
                     const int* lo, const int* hi, std::vector<int>& map,
                     std::vector<unsigned char>& occ) {
  map.assign((size_t)nvctr, -1);
  if (nvctr == 0) return BDFT_OK;
  if (nseg <= 0 || !keyg || !keyv)
    return set_err(ctx, BDFT_ERR_DESC, "%s: nvctr=%d but nseg=%d", what, nvctr, nseg);
  const long long M = (long long)N[0] * N[1] * N[2];
  for (int s = 0; s < nseg; ++s) {
    const long long j0 = keyg[2 * s], j1 = keyg[2 * s + 1], jj = keyv[s];
    if (j0 < 1 || j1 < j0 || j1 > M)
      return set_err(ctx, BDFT_ERR_DESC, "%s segment %d: keyg (%lld,%lld) outside 1..%lld",
                     what, s + 1, j0, j1, M);
    long long ii = j0 - 1;
    const int i3 = (int)(ii / ((long long)N[0] * N[1]));
    ii -= (long long)i3 * N[0] * N[1];
    const int i2 = (int)(ii / N[0]);
    const int i0 = (int)(ii - (long long)i2 * N[0]);
    const int i1 = i0 + (int)(j1 - j0);
    if (i1 >= N[0])
      return set_err(ctx, BDFT_ERR_DESC, "%s segment %d crosses a row", what, s + 1);
    if (jj < 1 || jj - 1 + (j1 - j0) >= nvctr)
      return set_err(ctx, BDFT_ERR_DESC, "%s segment %d: keyv %lld out of range", what,
                     s + 1, jj);
    if (lo && (i2 < lo[1] || i2 > hi[1] || i3 < lo[2] || i3 > hi[2] || i0 < lo[0] ||
               i1 > hi[0]))
      return set_err(ctx, BDFT_ERR_DESC, "%s segment %d outside the fine box", what, s + 1);
    for (int i = i0; i <= i1; ++i) {
      const long long v = jj - 1 + (i - i0);
      const long long lin = i + (long long)N[0] * (i2 + (long long)N[1] * i3);
      if (map[v] != -1)
        return set_err(ctx, BDFT_ERR_DESC, "%s: coefficient %lld addressed twice", what, v + 1);
      if (occ[lin])
        return set_err(ctx, BDFT_ERR_DESC, "%s: grid point %lld addressed twice", what, lin + 1);
      occ[lin] = 1;
      map[v] = (int)lin;
    }
  }
  for (int v = 0; v < nvctr; ++v)
    if (map[v] < 0)
      return set_err(ctx, BDFT_ERR_DESC, "%s: coefficient %d not covered by keys", what, v + 1);
  return BDFT_OK;
}

extern "C" int bdft_lr_create(bdft_ctx* ctx, char geocode, const int n[3],
                              const int nfl[3], const int nfu[3],
                              const double hgrid[3], int nseg_c, int nvctr_c,
                              const int32_t* keyg_c, const int32_t* keyv_c,
                              int nseg_f, int nvctr_f, const int32_t* keyg_f,
                              const int32_t* keyv_f, bdft_lr** out) {
  if (!ctx || !out || !n || !hgrid) return set_err(ctx, BDFT_ERR_ARG, "null argument");
  *out = nullptr;
  if (geocode != 'P' && geocode != 'F')
    return set_err(ctx, BDFT_ERR_UNSUPPORTED, "geocode '%c' not supported (P or F)", geocode);
  if (nvctr_c < 0 || nvctr_f < 0 || nseg_c < 0 || nseg_f < 0)
    return set_err(ctx, BDFT_ERR_ARG, "negative descriptor size");
  bdft_lr lr;
  lr.ctx = ctx;
  lr.geocode = geocode;
  for (int d = 0; d < 3; ++d) {
    lr.n[d] = n[d];
    lr.N[d] = n[d] + 1;
    if (lr.N[d] < 1 || lr.N[d] > 4096)
      return set_err(ctx, BDFT_ERR_UNSUPPORTED, "box extent n%d=%d outside 0..4095", d + 1,
                     n[d]);
    lr.L[d] = geocode == 'P' ? 2 * lr.N[d] : 2 * lr.N[d] + 29;
    if (!(hgrid[d] > 0.0)) return set_err(ctx, BDFT_ERR_ARG, "hgrid must be positive");
    lr.hgrid[d] = hgrid[d];
    lr.nfl[d] = (geocode == 'F' && nfl) ? nfl[d] : 0;
    lr.nfu[d] = (geocode == 'F' && nfu) ? nfu[d] : n[d];
  }
  if (geocode == 'F' && nvctr_f > 0) {
    if (!nfl || !nfu) return set_err(ctx, BDFT_ERR_ARG, "free BC needs nfl/nfu");
    for (int d = 0; d < 3; ++d)
      if (lr.nfl[d] < 0 || lr.nfu[d] > n[d] || lr.nfl[d] > lr.nfu[d])
        return set_err(ctx, BDFT_ERR_DESC, "fine box [%d,%d] outside 0..%d", lr.nfl[d],
                       lr.nfu[d], n[d]);
  }
  const long long M = (long long)lr.N[0] * lr.N[1] * lr.N[2];
  if (nvctr_c > M || nvctr_f > M)
    return set_err(ctx, BDFT_ERR_DESC, "more coefficients than grid points");
  std::vector<int> cmap, fmap;
  std::vector<unsigned char> occ((size_t)M, 0);
  int rc = build_map(ctx, "coarse", lr.N, nseg_c, nvctr_c, keyg_c, keyv_c, nullptr,
                     nullptr, cmap, occ);
  if (rc) return rc;
  std::fill(occ.begin(), occ.end(), 0);
  const bool fbox = geocode == 'F';
  rc = build_map(ctx, "fine", lr.N, nseg_f, nvctr_f, keyg_f, keyv_f,
                 fbox ? lr.nfl : nullptr, fbox ? lr.nfu : nullptr, fmap, occ);
  if (rc) return rc;
  lr.nvctr_c = nvctr_c;
  lr.nvctr_f = nvctr_f;
  lr.dense = (nvctr_c == M && nvctr_f == M);
  HIPCHK(ctx, hipSetDevice(ctx->device));
  if (nvctr_c > 0) {
    HIPCHK(ctx, hipMalloc(&lr.d_cmap, sizeof(int) * nvctr_c));
    HIPCHK(ctx, hipMemcpy(lr.d_cmap, cmap.data(), sizeof(int) * nvctr_c,
                          hipMemcpyHostToDevice));
  }
  if (nvctr_f > 0) {
    HIPCHK(ctx, hipMalloc(&lr.d_fmap, sizeof(int) * nvctr_f));
    HIPCHK(ctx, hipMemcpy(lr.d_fmap, fmap.data(), sizeof(int) * nvctr_f,
                          hipMemcpyHostToDevice));
  }
  // inverse maps (grid point -> coefficient index, -1 if absent), skipped
  // when every point is present in grid order
  auto inverse = [&](const std::vector<int>& map, int*& dptr) -> int {
    bool ident = (long long)map.size() == M;
    for (size_t v = 0; ident && v < map.size(); ++v) ident = map[v] == (int)v;
    if (ident) return BDFT_OK;
    std::vector<int> inv((size_t)M, -1);
    for (size_t v = 0; v < map.size(); ++v) inv[map[v]] = (int)v;
    HIPCHK(ctx, hipMalloc(&dptr, sizeof(int) * M));
    HIPCHK(ctx, hipMemcpy(dptr, inv.data(), sizeof(int) * M, hipMemcpyHostToDevice));
    return BDFT_OK;
  };
  rc = inverse(cmap, lr.d_cinv);
  if (!rc) rc = inverse(fmap, lr.d_finv);
  if (rc) {
    if (lr.d_cmap) hipFree(lr.d_cmap);
    if (lr.d_fmap) hipFree(lr.d_fmap);
    if (lr.d_cinv) hipFree(lr.d_cinv);
    return rc;
  }
  *out = new bdft_lr(lr);
  ctx->regions.push_back(*out);
  return BDFT_OK;
}

// device memory of a region (its coarse/fine maps and inverse maps)
static void lr_release(bdft_lr* lr) {
  if (lr->d_cmap) hipFree(lr->d_cmap);
  if (lr->d_fmap) hipFree(lr->d_fmap);
  if (lr->d_cinv) hipFree(lr->d_cinv);
  if (lr->d_finv) hipFree(lr->d_finv);
  lr->d_cmap = lr->d_fmap = lr->d_cinv = lr->d_finv = nullptr;
}

extern "C" int bdft_lr_destroy(bdft_lr* lr) {
  if (!lr) return BDFT_ERR_ARG;
  if (lr->ctx) {
    auto& rs = lr->ctx->regions;
    rs.erase(std::remove(rs.begin(), rs.end(), lr), rs.end());
    hipSetDevice(lr->ctx->device);
    hipStreamSynchronize(lr->ctx->stream);
    if (lr->ctx->cached) {  // a cached plan may reference this region's maps
      bdft_plan_destroy(lr->ctx->cached);
      lr->ctx->cached = nullptr;
      lr->ctx->cached_key.clear();
    }
    if (lr->ctx->pc_plan) {
      bdft_plan_destroy(lr->ctx->pc_plan);
      lr->ctx->pc_plan = nullptr;
      lr->ctx->pc_key.clear();
    }
    if (lr->ctx->den_plan) {
      bdft_plan_destroy(lr->ctx->den_plan);
      lr->ctx->den_plan = nullptr;
      lr->ctx->den_key.clear();
    }
    lr_release(lr);
  }
  delete lr;
  return BDFT_OK;
}

extern "C" int bdft_lr_info(const bdft_lr* lr, int nfine[3], int64_t* ncoef) {
  if (!lr) return BDFT_ERR_ARG;
  if (nfine)
    for (int d = 0; d < 3; ++d) nfine[d] = lr->L[d];
  if (ncoef) *ncoef = (int64_t)lr->nvctr_c + 7 * (int64_t)lr->nvctr_f;
  return BDFT_OK;
}

// ---------------------------------------------------------------------------
// operator tables of the matrix-core path
// ---------------------------------------------------------------------------
static uint16_t f2bf(float x) {  // round to nearest even
  uint32_t u;
  memcpy(&u, &x, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
static double bf2d(uint16_t h) {
  const uint32_t u = (uint32_t)h << 16;
  float f;
  memcpy(&f, &u, 4);
  return (double)f;
}
// IEEE binary16 bits of v, round to nearest even (subnormals kept, overflow
// to infinity)
static uint16_t d2h(double v) {
  const uint16_t sign = std::signbit(v) ? 0x8000 : 0;
  const double a = std::fabs(v);
  if (a != a) return 0x7e00;
  if (a == 0.0) return sign;
  if (a >= 65520.0) return sign | 0x7c00;
  int e;
  (void)std::frexp(a, &e);             // a = f 2^e, f in [0.5, 1)
  if (e >= -13) {                      // a >= 2^-14: normal
    const double q = std::ldexp(1.0, e - 11);
    const double r = std::nearbyint(a / q) * q;
    int er;
    const double fr = std::frexp(r, &er);
    const int E = er - 1 + 15;
    const int mant = (int)std::ldexp(2.0 * fr - 1.0, 10);
    return sign | (uint16_t)(E << 10) | (uint16_t)mant;
  }
  const double r = std::nearbyint(std::ldexp(a, 24));  // subnormal quanta of 2^-24
  if (r >= 1024.0) return sign | (uint16_t)(1 << 10);
  return sign | (uint16_t)r;
}
static double h2d(uint16_t h) {
  const int E = (h >> 10) & 31, mant = h & 1023;
  double v;
  if (E == 0) v = std::ldexp((double)mant, -24);
  else if (E == 31) v = mant ? NAN : INFINITY;
  else v = std::ldexp(1.0 + mant / 1024.0, E - 15);
  return (h & 0x8000) ? -v : v;
}

// Dense operator of one axis: op 0 = G (synthesis+magic, fil2 of
// bigdft/src/orbitals/v.inc), 1 = G^T, 2 = wavelet-basis kinetic (unscaled;
// ConvolkineticT filters, convolut_ib_optim.f90:962-1038).  Periodic boxes
// fold every image of the stencil into the matrix; free boxes cut it.
static void build_op(int op, int N, int per, std::vector<double>& M, int& rows, int& cols) {
  const int W = 2 * N, L = per ? 2 * N : 2 * N + 29;
  if (op == 2) {
    rows = cols = W;
    M.assign((size_t)rows * cols, 0.0);
    for (int so = 0; so < 2; ++so)
      for (int i = 0; i < N; ++i)
        for (int si = 0; si < 2; ++si)
          for (int l = -14; l <= 14; ++l) {
            int k = i + l;
            if (per) k = ((k % N) + N) % N;
            else if (k < 0 || k >= N) continue;
            const int al = l < 0 ? -l : l;
            const double f = so == 0 ? (si == 0 ? Fil::KA[al] : Fil::KC[-l + 14])
                                     : (si == 0 ? Fil::KC[l + 14] : Fil::KE[al]);
            M[(size_t)(so * N + i) * cols + si * N + k] += f;
          }
    return;
  }
  // G: rows = fine outputs, cols = [c ; d]
  std::vector<double> G((size_t)L * W, 0.0);
  for (int k = 0; k < N; ++k)
    for (int s = 0; s < 2; ++s)
      for (int j = -14; j <= 16; ++j) {
        int p = 2 * k + j;
        if (per) p = ((p % W) + W) % W;
        else p += 14;
        G[(size_t)p * W + s * N + k] += s == 0 ? Fil::F2S[j + 14] : Fil::F2W[j + 14];
      }
  if (op == 0) {
    rows = L;
    cols = W;
    M.swap(G);
  } else {
    rows = W;
    cols = L;
    M.assign((size_t)W * L, 0.0);
    for (int p = 0; p < L; ++p)
      for (int q = 0; q < W; ++q) M[(size_t)q * L + p] = G[(size_t)p * W + q];
  }
}

// Operator tables for the matrix-core kernels.  Every entry of the dense
// operator (rows = outputs, cols = inputs, periodic wrap or zero padding
// folded in) is split into three bf16 pieces (hi, lo = bf16(v - hi),
// lo2 = bf16(v - hi - lo)).  Storage is in matrix-core operand order: for
// 32-row block q and each K step s (16 columns) holding a non-zero, one
// fragment of 64 lanes x 8 bf16 per piece (lane l: row 32q + l%32, columns
// 16s + 8(l/32) .. +8), pieces of one step adjacent, so a wave reads each
// piece as one contiguous 1 KB block.  Per row block: {bit mask of the K
// steps stored, index of its first fragment}.
//
// fp16 tables (F16X3) keep every entry to 2^-22 only where its low piece
// stays a normal fp16 number (|v| >= 2^-2), and fp16 stops at 2^-24: the
// far tails of fil2 (down to 1.6e-9) would be lost, and they alone make
// psir in the outer halo of a free box, where an orbital far from the
// potential's wells gets all of its epot.  So the fp16 tables are scaled:
//   G (op 0): row r by s_r = 2^(13 - ceil(log2 sum_c |G_rc|)), every row's
//             entries as large as fp16 allows (row sum in (2^12, 2^13]);
//             the kernels multiply a G output by 2^-13 before splitting it
//             for the next pass and remove the rest, u_r = 1/s_r of each
//             row used, in FP32 at the end (exact powers of two);
//   K (op 2): all entries by KIN16 = 2^9 (largest 24.9 * 2^9 < 65504),
//             removed with the kinetic scale.
// rscale holds u_r for the rows of every scaled G table (OpTab::soff).
struct TabSet {
  std::vector<OpTab> tabs;
  std::vector<uint16_t> frag;
  std::vector<uint32_t> mq;
  std::vector<float> rscale;
  std::map<std::tuple<int, int, int, int>, int> index;
  // f16: fp16 pieces (hi, lo) instead of bf16 (hi, lo, lo2)
  int get(int op, int N, int per, bool f16) {
    const auto key = std::make_tuple(op, N, per, (int)f16);
    auto it = index.find(key);
    if (it != index.end()) return it->second;
    std::vector<double> M;
    int rows = 0, cols = 0;
    build_op(op, N, per, M, rows, cols);
    OpTab t{};
    t.rows = rows;
    t.cols = cols;
    t.Dp = (rows + 31) / 32 * 32;
    t.Kp = (cols + 15) / 16 * 16;
    t.off = 0;
    t.soff = -1;
    if (f16 && op == 0) {
      t.soff = (int)rscale.size();
      for (int r = 0; r < rows; ++r) {
        double sum = 0.0;
        for (int c = 0; c < cols; ++c) sum += std::fabs(M[(size_t)r * cols + c]);
        int e = 0;
        if (sum > 0.0) {
          (void)std::frexp(sum, &e);       // sum in [2^(e-1), 2^e)
          if (std::ldexp(1.0, e - 1) == sum) --e;
        }
        const double sr = std::ldexp(1.0, 13 - e);
        for (int c = 0; c < cols; ++c) M[(size_t)r * cols + c] *= sr;
        rscale.push_back((float)std::ldexp(1.0, e - 13));
      }
    } else if (f16 && op == 2) {
      for (double& v : M) v *= KIN16;
    }
    t.moff = (int)(mq.size() / 2);
    for (int q = 0; q < t.Dp / 32; ++q) {
      uint32_t m = 0;
      for (int r = 32 * q; r < std::min(rows, 32 * q + 32); ++r)
        for (int c = 0; c < cols; ++c)
          if (M[(size_t)r * cols + c] != 0.0) m |= 1u << (c / 16);
      const uint32_t first = (uint32_t)(frag.size() / (3 * 512));
      mq.push_back(m);
      mq.push_back(first);
      for (int s = 0; s < t.Kp / 16; ++s) {
        if (!(m & (1u << s))) continue;
        const size_t base = frag.size();
        frag.resize(base + 3 * 512, 0);
        for (int l = 0; l < 64; ++l)
          for (int u = 0; u < 8; ++u) {
            const int r = 32 * q + (l & 31), c = 16 * s + 8 * (l >> 5) + u;
            if (r >= rows || c >= cols) continue;
            const double v = M[(size_t)r * cols + c];
            uint16_t h, l1, l2;
            if (f16) {
              h = d2h(v);
              l1 = d2h(v - h2d(h));
              l2 = 0;
            } else {
              h = f2bf((float)v);
              l1 = f2bf((float)(v - bf2d(h)));
              l2 = f2bf((float)(v - bf2d(h) - bf2d(l1)));
            }
            frag[base + 0 * 512 + 8 * l + u] = h;
            frag[base + 1 * 512 + 8 * l + u] = l1;
            frag[base + 2 * 512 + 8 * l + u] = l2;
          }
      }
    }
    const int id = (int)tabs.size();
    tabs.push_back(t);
    index[key] = id;
    return id;
  }
};

// ---------------------------------------------------------------------------
// plans
// ---------------------------------------------------------------------------
static inline int cdiv(long long a, long long b) { return (int)((a + b - 1) / b); }

static int box_tiles(const BoxDesc& b, int pass, int zg) {
  const int N1 = b.N[0], N2 = b.N[1], N3 = b.N[2], L1 = b.L[0], L2 = b.L[1];
  switch (pass) {
    case P_UNC: return cdiv(b.nvctr_c, 256) + cdiv(b.nvctr_f, 256);
    case P_XF:
    case P_XB: return cdiv(4LL * N2 * N3, TW);
    case P_YF:
    case P_YB: return cdiv(L1, TW) * 2 * N3;
    case P_KY: return cdiv(2 * N1, TW) * 2 * N3;
    case P_KZ: return cdiv(4LL * N1 * N2, TW);
    case P_ZP: return cdiv(L1, TW) * L2;
    case P_PF:
    case P_PB: return 2 * N3;
    case P_KZG: return cdiv((long long)cdiv(4LL * N1 * N2, TW), zg);
    case P_ZPG: return cdiv((long long)cdiv(L1, TW) * L2, zg);
    case P_ZPF: return cdiv((long long)L1 * L2, TW);
  }
  return 0;
}

extern "C" int bdft_plan_destroy(bdft_plan* plan);

// Launch choices that depend on every box of a batch (the largest box's
// needs): working precision, z-line bundling, kernel sequence and units.  A
// chunked plan takes them from the whole batch, so each part runs the
// kernels the one-chunk plan would run on its orbitals.
struct PlanChoice {
  int precision;
  int zg;
  int fused, zvalu, bvalu;  // matrix-core sequence
  int plane;                // vector-unit sequence: plane-fused or line passes
  int xbig, zbig;
};

// point the plan's grids at their places in a workspace of ws_need bytes
static void bind_ws(bdft_plan* plan, char* w) {
  plan->B.W = w + plan->ws_at[0];
  plan->B.K = w + plan->ws_at[1];
  plan->B.T1 = w + plan->ws_at[2];
  plan->B.T2 = w + plan->ws_at[3];
  plan->B.T3 = w + plan->ws_at[4];
}

static int make_plan(bdft_ctx* ctx, int norb, const bdft_lr* const* lrs,
                     const double* pot, const int64_t* pot_off, const int32_t* potdim,
                     const int32_t* ishift, const double* psi, const int64_t* psi_off,
                     double* hpsi, double* ekin, double* epot, int precision,
                     bool own_ws, bdft_plan** out, bool density = false,
                     const bdft_confpot* conf = nullptr, bool grow_plane = false,
                     Arena* arena = nullptr, bool defer_ws = false,
                     const PlanChoice* force = nullptr, PlanChoice* got = nullptr) {
  if (!ctx || !out) return BDFT_ERR_ARG;
  *out = nullptr;
  if (norb <= 0 || !lrs || !psi || (!density && (!pot || !hpsi || !ekin || !epot)))
    return set_err(ctx, BDFT_ERR_ARG, "null pointer or norb <= 0");
  if (precision < BDFT_PREC_FP64 || precision > BDFT_PREC_F16X3 ||
      (density && precision > BDFT_PREC_FP32))
    return set_err(ctx, BDFT_ERR_UNSUPPORTED, "precision %d not available", precision);
  // split precisions on a batch with a box beyond the matrix-core kernels
  // (more than MFMA_MAXN coarse points on an axis) run the FP32 vector
  // passes, which are at least as accurate
  if (force)
    precision = force->precision;
  else if (precision >= BDFT_PREC_BF16X3)
    for (int i = 0; i < norb; ++i)
      for (int d = 0; d < 3; ++d)
        if (lrs[i] && lrs[i]->N[d] > MFMA_MAXN) precision = BDFT_PREC_FP32;
  const size_t esz = precision == BDFT_PREC_FP64 ? sizeof(double) : sizeof(float);
  const bool mfma = precision >= BDFT_PREC_BF16X3;
  const bool f16 = precision == BDFT_PREC_F16X3;
  TabSet tabs;
  std::vector<BoxDesc> box((size_t)norb);
  long long wtot = 0, t1tot = 0, t2tot = 0, t3tot = 0, psi_run = 0;
  int maxN = 0;
  bool xbig = false, zbig = false;
  int maxL3 = 0;
  bool need_zero = false;
  for (int i = 0; i < norb; ++i) {
    const bdft_lr* lr = lrs[i];
    if (!lr || lr->ctx != ctx)
      return set_err(ctx, BDFT_ERR_ARG, "orbital %d: region missing or from another context", i);
    BoxDesc& b = box[i];
    memset(&b, 0, sizeof b);
    for (int d = 0; d < 3; ++d) {
      b.N[d] = lr->N[d];
      b.L[d] = lr->L[d];
      b.kscale[d] = -0.5 / (lr->hgrid[d] * lr->hgrid[d]);
      b.potdim[d] = potdim ? potdim[3 * i + d] : lr->L[d];
      b.ishift[d] = ishift ? ishift[3 * i + d] : 0;
      if (b.potdim[d] <= 0)
        return set_err(ctx, BDFT_ERR_ARG, "orbital %d: potential extent <= 0", i);
      if (d == 2 && (long long)b.potdim[0] * b.potdim[1] * b.potdim[2] >= (1LL << 29))
        return set_err(ctx, BDFT_ERR_UNSUPPORTED,
                       "orbital %d: potential box of 2^29 or more points (4 GiB)", i);
      maxN = std::max(maxN, lr->N[d]);
      if (lr->N[d] > MAXN)
        return set_err(ctx, BDFT_ERR_UNSUPPORTED,
                       "orbital %d: %d coarse points along axis %d, at most %d supported", i,
                       lr->N[d], d + 1, MAXN);
    }
    b.per = lr->geocode == 'P';
    if (conf && conf[i].potorder != 0) {
      const bdft_confpot& c = conf[i];
      if (c.potorder < 0)
        return set_err(ctx, BDFT_ERR_ARG, "orbital %d: negative confinement order", i);
      b.conf_on = 1;
      b.conf_half = c.potorder / 2;
      b.conf_prefac = c.prefac;
      for (int d = 0; d < 3; ++d) {
        b.conf_ioff[d] = c.ioffset[d];
        b.conf_hh[d] = c.hh[d];
        b.conf_r[d] = c.rxyzConf[d];
      }
    }
    b.nvctr_c = lr->nvctr_c;
    b.nvctr_f = lr->nvctr_f;
    b.cmap = lr->d_cmap;
    b.fmap = lr->d_fmap;
    b.cinv = lr->d_cinv;
    b.finv = lr->d_finv;
    b.w_off = wtot;
    b.t1_off = t1tot;
    b.t2_off = t2tot;
    b.t3_off = t3tot;
    b.pot_off = pot_off ? pot_off[i] : 0;
    const long long ncoef = (long long)lr->nvctr_c + 7LL * lr->nvctr_f;
    // no psi_off: orbitals stored one after the other
    b.psi_off = psi_off ? psi_off[i] : psi_run;
    psi_run += ncoef;
    if (mfma) {
      for (int d = 0; d < 3; ++d) {
        b.opG[d] = tabs.get(0, b.N[d], b.per, f16);
        b.opS[d] = tabs.get(1, b.N[d], b.per, false);
        b.opK[d] = tabs.get(2, b.N[d], b.per, f16);
      }
      // forward plane kernel in two parts where that brings its LDS within
      // half a CU's (two work-groups per CU)
      const size_t half = ctx->plane_lds / 2;
      b.mfp = (mfma_lds_fwd(2 * b.N[0], 2 * b.N[1], b.L[0], 1) > half &&
               mfma_lds_fwd(2 * b.N[0], 2 * b.N[1], b.L[0], 2) <= half) ? 2 : 1;
    }
    wtot += 8LL * b.N[0] * b.N[1] * b.N[2];
    t1tot += (long long)b.L[0] * 4 * b.N[1] * b.N[2];
    t2tot += (long long)b.L[0] * b.L[1] * 2 * b.N[2];
    t3tot += (long long)b.L[0] * b.L[1] * b.L[2];
    xbig = xbig || b.N[0] > REG_N || b.L[0] > REG_L;
    zbig = zbig || b.L[2] > REG_L;
    maxL3 = std::max(maxL3, b.L[2]);
    if (!lr->dense) need_zero = true;
  }
  // bundles per z-line tile of the matrix-core sequence: three, unless a
  // z-line kernel would then take more than half of a CU's LDS (one
  // work-group per CU); two halve the waste of a partial last tile too
  int zg = ZGROUP;
  if (mfma) {
    const OpTab* tb = tabs.tabs.data();
    size_t mx = 0;
    for (int i = 0; i < norb; ++i) {
      const BoxDesc& b = box[i];
      mx = std::max(mx, mfma_lds_kz(2 * b.N[2], tb[b.opK[2]].Kp, ZGROUP));
      mx = std::max(mx, mfma_lds_zp(tb[b.opG[2]].Kp, tb[b.opG[2]].Dp, ZGROUP));
    }
    if (ctx->opt_zg == 2 || (ctx->opt_zg == 0 && mx > ctx->plane_lds / 2) ||
        mx > ctx->plane_lds)
      zg = 2;
    if (force) zg = force->zg;
  }
  // tile tables
  std::vector<int> starts;
  PassTables T{};
  int start_off[P_COUNT];
  for (int p = 0; p < P_COUNT; ++p) {
    const int t0 = box_tiles(box[0], p, zg);
    bool uni = true;
    long long tot = 0;
    for (int i = 0; i < norb; ++i) {
      const int t = box_tiles(box[i], p, zg);
      uni = uni && t == t0;
      tot += t;
    }
    if (tot > 0x7fffffffLL) return set_err(ctx, BDFT_ERR_UNSUPPORTED, "batch too large");
    T.ntiles[p] = (int)tot;
    T.t[p].nbox = norb;
    T.t[p].uniform = uni ? t0 : 0;
    start_off[p] = -1;
    if (!uni) {
      start_off[p] = (int)starts.size();
      int acc = 0;
      for (int i = 0; i < norb; ++i) {
        starts.push_back(acc);
        acc += box_tiles(box[i], p, zg);
      }
      starts.push_back(acc);
    }
  }
  // plane-fused sequence when every box qualifies: periodic axes need at
  // least 18 coarse points (single wrap of the in-plane stencils) and both
  // plane kernels must fit in LDS
  MfmaLds ml{};
  ml.zg = zg;
  bool mfma_bwd_ok = true, valu_bwd_ok = false;  // units that fit the backward pass
  if (mfma) {
    for (int i = 0; i < norb; ++i) {
      const BoxDesc& b = box[i];
      const OpTab* tb = tabs.tabs.data();
      for (int d = 0; d < 3; ++d)
        ml.kin[d] = std::max(ml.kin[d], mfma_lds_kin(2 * b.N[d], tb[b.opK[d]].Kp));
      ml.grow = std::max(ml.grow, mfma_lds_grow(2 * b.N[1], b.L[0], tb[b.opG[0]].Kp,
                                                tb[b.opG[1]].Kp));
      ml.zpot = std::max(ml.zpot, mfma_lds_zpot(tb[b.opG[2]].Kp, tb[b.opS[2]].Kp));
      ml.shrink = std::max(ml.shrink, mfma_lds_shrink(b.L[0], 2 * b.N[0], 2 * b.N[1],
                                                      tb[b.opS[1]].Kp, tb[b.opS[0]].Kp));
    }
    for (int i = 0; i < norb; ++i) {
      const BoxDesc& b = box[i];
      const OpTab* tb = tabs.tabs.data();
      ml.fwd = std::max(ml.fwd, mfma_lds_fwd(2 * b.N[0], 2 * b.N[1], b.L[0], b.mfp));
      ml.bwd = std::max(ml.bwd, mfma_lds_bwd(2 * b.N[0], b.L[0], b.L[1]));
      ml.kz = std::max(ml.kz, mfma_lds_kz(2 * b.N[2], tb[b.opK[2]].Kp, zg));
      ml.zp = std::max(ml.zp, mfma_lds_zp(tb[b.opG[2]].Kp, tb[b.opG[2]].Dp, zg));
    }
    // the FP32 plane backward kernel can stand in for the matrix-core one
    // where its planes fit
    valu_bwd_ok = ctx->opt_bwd != 1;
    size_t vb_max = 0;
    for (int i = 0; i < norb; ++i) {
      const BoxDesc& b = box[i];
      vb_max = std::max(vb_max, plane_lds_bwd(b.N[0], b.N[1], b.L[0], b.L[1],
                                              (int)sizeof(float)));
    }
    valu_bwd_ok = valu_bwd_ok && vb_max <= ctx->plane_lds;
    mfma_bwd_ok = ml.bwd <= ctx->plane_lds;
    // plane-fused sequence when the forward plane kernel fits and the
    // backward pass fits on either unit; the per-pass one otherwise (or
    // when line passes are forced)
    ml.fused = !ctx->force_lines && ml.fwd <= ctx->plane_lds &&
               (mfma_bwd_ok || valu_bwd_ok);
    // z potential pass of the fused sequence: the FP32 line kernel where it
    // applies (every extent within the line kernels' range), unless the
    // matrix cores are asked for
    bool line_ok = true;
    for (int i = 0; i < norb; ++i)
      line_ok = line_ok && box[i].L[2] <= REG_L;
    // (F16X3 always: its fp16 pieces are cut for the box scale, and the z
    // lines of a free box's x-y halo lie far below it)
    ml.zvalu = ml.fused && line_ok && (ctx->opt_zpot != 1 || f16);
    if (force) {
      ml.fused = force->fused;
      ml.zvalu = force->zvalu;
    }
    ml.maxN = maxN;
    size_t mx = std::max(ml.kin[2], ml.zpot);
    if (ml.fused)
      mx = std::max(std::max(ml.kz, ml.zvalu ? (size_t)0 : ml.zp),
                    std::max(ml.fwd, mfma_bwd_ok ? ml.bwd : (size_t)0));
    else
      mx = std::max(std::max(mx, std::max(ml.kin[0], ml.kin[1])),
                    std::max(ml.grow, ml.shrink));
    if (mx > ctx->plane_lds) {
      // automatic units: the FP32 vector passes instead; a matrix-core unit
      // asked for by an option is refused
      if (ctx->opt_bwd != 1 && ctx->opt_zpot != 1)
        return make_plan(ctx, norb, lrs, pot, pot_off, potdim, ishift, psi, psi_off, hpsi,
                         ekin, epot, BDFT_PREC_FP32, own_ws, out, density, conf, grow_plane,
                         arena, defer_ws, nullptr, got);
      return set_err(ctx, BDFT_ERR_UNSUPPORTED,
                     "split-bf16 kernels need %zu bytes of LDS per work-group, the device "
                     "offers %zu",
                     mx, ctx->plane_lds);
    }
  }
  // plane-fused matrix-core kernels: the launch reserves the largest box's
  // LDS for every work-group; boxes whose plane kernels fit in half a CU's
  // LDS get launches of their own (two work-groups per CU), the others a
  // second pair of launches
  int g_off[4] = {-1, -1, -1, -1};  // small starts, small map, large starts, large map
  int pf_off[4][2] = {{-1, -1}, {-1, -1}, {-1, -1}, {-1, -1}};  // forward plane groups
  int pb_off[4][2] = {{-1, -1}, {-1, -1}, {-1, -1}, {-1, -1}};  // backward plane groups
  int vb_off[4][2] = {{-1, -1}, {-1, -1}, {-1, -1}, {-1, -1}};  // FP32 backward groups
  if (mfma && ml.fused) {
    const size_t half = ctx->plane_lds / 2;
    std::vector<int> sm, lg;
    size_t fs = 0, bs = 0, fl = 0, bl = 0;
    size_t vs = 0, vl = 0;
    bool all_free = true;
    for (int i = 0; i < norb; ++i) all_free = all_free && !box[i].per;
    // FP32 backward pass: asked for, automatic on free-BC batches, or the
    // only unit whose planes fit
    const bool bv = force ? force->bvalu != 0
                          : valu_bwd_ok && (ctx->opt_bwd == 2 || all_free || !mfma_bwd_ok);
    for (int i = 0; i < norb; ++i) {
      const BoxDesc& b = box[i];
      const size_t f = mfma_lds_fwd(2 * b.N[0], 2 * b.N[1], b.L[0], b.mfp);
      const size_t w = mfma_lds_bwd(2 * b.N[0], b.L[0], b.L[1]);
      const size_t v = plane_lds_bwd(b.N[0], b.N[1], b.L[0], b.L[1], (int)sizeof(float));
      if ((bv ? f : std::max(f, w)) <= half) {
        sm.push_back(i);
        fs = std::max(fs, f);
        bs = std::max(bs, w);
        vs = std::max(vs, v);
      } else {
        lg.push_back(i);
        fl = std::max(fl, f);
        bl = std::max(bl, w);
        vl = std::max(vl, v);
      }
    }
    if (sm.empty() || lg.empty()) vs = vl = std::max(vs, vl);  // one launch
    ml.bvalu = bv;
    ml.vbwd = vs;
    ml.vbwd2 = vl;
    auto add = [&](const std::vector<int>& g, int& so, int& mo) {
      so = (int)starts.size();
      int acc = 0;
      for (int i : g) {
        starts.push_back(acc);
        acc += box_tiles(box[i], P_PF, zg);
      }
      starts.push_back(acc);
      mo = (int)starts.size();
      for (int i : g) starts.push_back(i);
      return acc;
    };
    // FP32 backward plane pass: the boxes grouped by how many of their
    // work-groups fit a CU's LDS (1-4), one launch per group with the
    // group's largest LDS, so the largest box does not set the LDS of every
    // work-group
    T.n_vbg = 0;
    if (bv) {
      std::vector<int> grp[4];
      size_t gl[4] = {0, 0, 0, 0};
      for (int i = 0; i < norb; ++i) {
        const BoxDesc& b = box[i];
        const size_t v = plane_lds_bwd(b.N[0], b.N[1], b.L[0], b.L[1], (int)sizeof(float));
        const size_t fit = ctx->plane_lds / std::max(v, (size_t)1);
        const int c = (int)std::min((size_t)4, std::max((size_t)1, fit));
        grp[c - 1].push_back(i);
        gl[c - 1] = std::max(gl[c - 1], v);
      }
      for (int c = 4; c >= 1; --c)
        if (!grp[c - 1].empty()) {
          const int k = T.n_vbg++;
          T.nt_vbg[k] = add(grp[c - 1], vb_off[k][0], vb_off[k][1]);
          T.vbg[k].nbox = (int)grp[c - 1].size();
          T.vbg[k].uniform = 0;
          ml.lds_vbg[k] = gl[c - 1];
        }
    }
    if (!sm.empty() && !lg.empty()) {
      T.ntiles[P_PF] = T.ntiles[P_PB] = add(sm, g_off[0], g_off[1]);
      T.ntiles_pf2 = T.ntiles_pb2 = add(lg, g_off[2], g_off[3]);
      for (int p : {P_PF, P_PB}) {
        T.t[p].nbox = (int)sm.size();
        T.t[p].uniform = 0;
      }
      T.pf2.nbox = T.pb2.nbox = (int)lg.size();
      ml.fwd = fs;
      ml.bwd = bs;
      ml.fwd2 = fl;
      ml.bwd2 = bl;
    }
  }
  {
    // density plans (grow_plane) use the forward plane kernel without its
    // kinetic part; the preconditioner's plans keep the line passes
    bool plane = !ctx->force_lines && !mfma && (!density || grow_plane);
    size_t lpf = 0, lpb = 0;
    for (int i = 0; i < norb && plane; ++i) {
      const BoxDesc& b = box[i];
      lpf = std::max(lpf, plane_lds_fwd(b.N[0], b.N[1], b.L[0], (int)esz));
      lpb = std::max(lpb, plane_lds_bwd(b.N[0], b.N[1], b.L[0], b.L[1], (int)esz));
    }
    if (std::max(lpf, lpb) > ctx->plane_lds) plane = false;
    T.plane = plane ? 1 : 0;
    if (force) T.plane = force->plane;
    T.lds_pf = lpf;
    T.lds_pb = lpb;
    // forward plane pass: the boxes grouped by how many of their
    // work-groups fit a CU, by LDS and by threads (FP64: 768 per work-group,
    // at most two; FP32: 512, at most four), one launch per group with the
    // group's largest LDS, so that the largest box does not set every
    // work-group's LDS; one group: the single launch over t[P_PF]
    auto group = [&](int pass, int cap, auto lds_of, TileTable* tg, int* nt, size_t* lg,
                     int* fitv, int& n, int off[4][2]) {
      n = 0;
      std::vector<int> grp[4];
      size_t gl[4] = {0, 0, 0, 0};
      for (int i = 0; i < norb; ++i) {
        const size_t v = lds_of(box[i]);
        const size_t fit = ctx->plane_lds / std::max(v, (size_t)1);
        const int c = (int)std::min((size_t)cap, std::max((size_t)1, fit));
        grp[c - 1].push_back(i);
        gl[c - 1] = std::max(gl[c - 1], v);
      }
      int ng = 0;
      for (int c = 0; c < 4; ++c) ng += grp[c].empty() ? 0 : 1;
      if (ng < 2) return;
      for (int c = 4; c >= 1; --c)
        if (!grp[c - 1].empty()) {
          const int k = n++;
          off[k][0] = (int)starts.size();
          int acc = 0;
          for (int i : grp[c - 1]) {
            starts.push_back(acc);
            acc += box_tiles(box[i], pass, zg);
          }
          starts.push_back(acc);
          off[k][1] = (int)starts.size();
          for (int i : grp[c - 1]) starts.push_back(i);
          nt[k] = acc;
          if (fitv) fitv[k] = c;
          tg[k].nbox = (int)grp[c - 1].size();
          tg[k].uniform = 0;
          lg[k] = gl[c - 1];
        }
    };
    T.n_pfg = T.n_pbg = 0;
    if (T.plane && !mfma && !density && ctx->plane_lds > 0) {
      group(P_PF, esz == 8 ? 2 : 4,
            [&](const BoxDesc& b) { return plane_lds_fwd(b.N[0], b.N[1], b.L[0], (int)esz); },
            T.pfg, T.nt_pfg, T.lds_pfg, nullptr, T.n_pfg, pf_off);
      group(P_PB, esz == 8 ? 2 : 4,
            [&](const BoxDesc& b) {
              return plane_lds_bwd(b.N[0], b.N[1], b.L[0], b.L[1], (int)esz);
            },
            T.pbg, T.nt_pbg, T.lds_pbg, T.fit_pbg, T.n_pbg, pb_off);
    }
    // preconditioner plans (FP64): the x and y operator as one plane pass
    // where the W and K planes of the largest box fit in LDS
    size_t lk = 0;
    for (int i = 0; i < norb; ++i)
      lk = std::max(lk, plane_lds_kin(box[i].N[0], box[i].N[1], (int)sizeof(double)));
    T.pc_plane = density && !grow_plane && !ctx->force_lines && !mfma && esz == 8 &&
                 lk <= ctx->plane_lds;
    T.lds_pc = lk;
  }
  // tile order: the z-line and line passes deal their tiles in contiguous
  // ranges per XCD (neighbouring bundles share cache lines); the plane and
  // (un)compress passes as BDFT_OPT_TILE_ORDER asks
  for (int p = 0; p < P_COUNT; ++p) T.t[p].xcd = 1;
  {
    const int po = ctx->opt_order == 1 ? 1 : 0;
    T.t[P_PF].xcd = T.t[P_PB].xcd = T.t[P_UNC].xcd = po;
    T.pf2.xcd = T.pb2.xcd = po;
    for (int k = 0; k < 4; ++k) T.vbg[k].xcd = T.pfg[k].xcd = T.pbg[k].xcd = po;
  }
  if (force) {
    xbig = force->xbig;
    zbig = force->zbig;
  }
  if (got) {
    got->precision = precision;
    got->zg = zg;
    got->fused = ml.fused;
    got->zvalu = ml.zvalu;
    got->bvalu = ml.bvalu;
    got->plane = T.plane;
    got->xbig = xbig;
    got->zbig = zbig;
    return BDFT_OK;  // choices only: nothing allocated
  }
  bdft_plan* plan = new bdft_plan();
  plan->ctx = ctx;
  plan->nbox = norb;
  plan->precision = precision;
  plan->maxN = maxN;
  plan->wtotal = wtot;
  plan->need_zero = need_zero;
  plan->owns_ws = own_ws;
  // metadata: boxes | starts | partials
  const size_t nb_box = sizeof(BoxDesc) * norb;
  // tile -> box tables of the non-uniform passes (one load per work-group
  // instead of a binary search over the box starts), up to 4M tiles each
  int tb_off[P_COUNT];
  int tb_g[2] = {-1, -1};  // the two plane-kernel box groups
  int tb_v[4] = {-1, -1, -1, -1};  // the FP32 backward groups
  int tb_f[4] = {-1, -1, -1, -1};  // the forward plane groups
  int tb_b[4] = {-1, -1, -1, -1};  // the backward plane groups
  {
    constexpr int TB_MAX = 1 << 22;
    auto tiles_of = [&](int so, int nb) {
      const int off = (int)starts.size();
      for (int i = 0; i < nb; ++i)
        for (int t = starts[so + i]; t < starts[so + i + 1]; ++t) starts.push_back(i);
      return off;
    };
    for (int p = 0; p < P_COUNT; ++p)
      tb_off[p] = (start_off[p] >= 0 && T.ntiles[p] <= TB_MAX)
                      ? tiles_of(start_off[p], norb) : -1;
    if (g_off[0] >= 0) {
      if (T.ntiles[P_PF] <= TB_MAX) tb_g[0] = tiles_of(g_off[0], T.t[P_PF].nbox);
      if (T.ntiles_pf2 <= TB_MAX) tb_g[1] = tiles_of(g_off[2], T.pf2.nbox);
    }
    for (int k = 0; k < T.n_vbg; ++k)
      tb_v[k] = T.nt_vbg[k] <= TB_MAX ? tiles_of(vb_off[k][0], T.vbg[k].nbox) : -1;
    for (int k = 0; k < T.n_pfg; ++k)
      tb_f[k] = T.nt_pfg[k] <= TB_MAX ? tiles_of(pf_off[k][0], T.pfg[k].nbox) : -1;
    for (int k = 0; k < T.n_pbg; ++k)
      tb_b[k] = T.nt_pbg[k] <= TB_MAX ? tiles_of(pb_off[k][0], T.pbg[k].nbox) : -1;
  }
  // work queue of the persistent z potential pass: the boxes in order of
  // decreasing work (fine points), each box's tiles together
  int zq_off = -1;
  if (ctx->opt_persist == 1 && T.ntiles[P_ZPF] > 0) {
    std::vector<int> order((size_t)norb);
    for (int i = 0; i < norb; ++i) order[i] = i;
    auto work = [&](int i) { return (long long)box[i].L[0] * box[i].L[1] * box[i].L[2]; };
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b2) { return work(a) > work(b2); });
    zq_off = (int)starts.size();
    for (int i : order) {
      const int cnt = box_tiles(box[i], P_ZPF, zg);
      const int first = start_off[P_ZPF] >= 0 ? starts[start_off[P_ZPF] + i] : i * cnt;
      for (int t = 0; t < cnt; ++t) starts.push_back(first + t);
    }
  }
  const size_t nb_st = sizeof(int) * starts.size();
  int epass[4] = {P_XF, P_KY, P_KZ, P_ZP};
  if (T.plane) epass[0] = P_PF;
  plan->mlds = ml;
  if (mfma) {
    const size_t nt = sizeof(OpTab) * tabs.tabs.size();
    const size_t nf = sizeof(uint16_t) * tabs.frag.size();
    const size_t nm = sizeof(uint32_t) * tabs.mq.size();
    const size_t ns = sizeof(float) * tabs.rscale.size();
    auto al2 = [](size_t x) { return (x + 255) & ~(size_t)255; };
    int rc2 = plan_buf(ctx, arena, plan->tabbuf, al2(nt) + al2(nf) + al2(nm) + al2(ns));
    if (rc2) {
      delete plan;
      return rc2;
    }
    char* q = (char*)plan->tabbuf.p;
    hipError_t e = plan_h2d(arena, q, tabs.tabs.data(), nt);
    if (e == hipSuccess) e = plan_h2d(arena, q + al2(nt), tabs.frag.data(), nf);
    if (e == hipSuccess) e = plan_h2d(arena, q + al2(nt) + al2(nf), tabs.mq.data(), nm);
    if (e == hipSuccess)
      e = plan_h2d(arena, q + al2(nt) + al2(nf) + al2(nm), tabs.rscale.data(), ns);
    plan->B.trs = ns ? (const float*)(q + al2(nt) + al2(nf) + al2(nm)) : nullptr;
    // F16X3: T2 lines carry their fp16 table scales into the FP32 z pass
    plan->B.t2_scaled = f16 ? 1 : 0;
    if (e != hipSuccess) {
      dev_free(plan->tabbuf);
      delete plan;
      return set_err(ctx, BDFT_ERR_HIP, "operator table upload failed");
    }
    plan->B.tabs = (const OpTab*)q;
    plan->B.tfrag = q + al2(nt);
    plan->B.tmq = (const uint2*)(q + al2(nt) + al2(nf));
  }
  size_t nb_part = 0;
  for (int e = 0; e < 4; ++e) nb_part += sizeof(double) * T.ntiles[epass[e]];
  const size_t nb_max = f16 ? sizeof(unsigned) * norb + sizeof(float) * T.ntiles[P_UNC] + 256 : 0;
  auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  int rc = plan_buf(ctx, arena, plan->meta, al(nb_box) + al(nb_st) + al(nb_part) + al(nb_max));
  if (rc) {
    dev_free(plan->tabbuf);
    delete plan;
    return rc;
  }
  char* m = (char*)plan->meta.p;
  BoxDesc* dbox = (BoxDesc*)m;
  int* dst = (int*)(m + al(nb_box));
  double* dpart = (double*)(m + al(nb_box) + al(nb_st));
  plan->B.bmax = f16 ? (unsigned*)(m + al(nb_box) + al(nb_st) + al(nb_part)) : nullptr;
  plan->B.tmax = f16 ? (float*)(m + al(nb_box) + al(nb_st) + al(nb_part) +
                                al(sizeof(unsigned) * norb))
                     : nullptr;
  hipError_t e1 = plan_h2d(arena, dbox, box.data(), nb_box);
  hipError_t e2 = plan_h2d(arena, dst, starts.data(), nb_st);
  if (e1 != hipSuccess || e2 != hipSuccess) {
    dev_free(plan->meta);
    dev_free(plan->tabbuf);
    delete plan;
    return set_err(ctx, BDFT_ERR_HIP, "descriptor upload failed");
  }
  for (int p = 0; p < P_COUNT; ++p) {
    T.t[p].start = start_off[p] >= 0 ? dst + start_off[p] : nullptr;
    T.t[p].tbox = tb_off[p] >= 0 ? dst + tb_off[p] : nullptr;
  }
  for (int k = 0; k < T.n_vbg; ++k) {
    T.vbg[k].start = dst + vb_off[k][0];
    T.vbg[k].bmap = dst + vb_off[k][1];
    T.vbg[k].tbox = tb_v[k] >= 0 ? dst + tb_v[k] : nullptr;
  }
  for (int k = 0; k < T.n_pfg; ++k) {
    T.pfg[k].start = dst + pf_off[k][0];
    T.pfg[k].bmap = dst + pf_off[k][1];
    T.pfg[k].tbox = tb_f[k] >= 0 ? dst + tb_f[k] : nullptr;
  }
  for (int k = 0; k < T.n_pbg; ++k) {
    T.pbg[k].start = dst + pb_off[k][0];
    T.pbg[k].bmap = dst + pb_off[k][1];
    T.pbg[k].tbox = tb_b[k] >= 0 ? dst + tb_b[k] : nullptr;
  }
  T.zq = zq_off >= 0 ? dst + zq_off : nullptr;
  T.zq_grid = 0;
  if (zq_off >= 0 && hipSetDevice(ctx->device) == hipSuccess)
    T.zq_grid = std::min(zpot_queue_grid(esz == sizeof(float) ? 1 : 0, maxN), T.ntiles[P_ZPF]);
  if (g_off[0] >= 0) {
    for (int p : {P_PF, P_PB}) {
      T.t[p].start = dst + g_off[0];
      T.t[p].bmap = dst + g_off[1];
      T.t[p].tbox = tb_g[0] >= 0 ? dst + tb_g[0] : nullptr;
    }
    T.pf2.start = T.pb2.start = dst + g_off[2];
    T.pf2.bmap = T.pb2.bmap = dst + g_off[3];
    T.pf2.tbox = T.pb2.tbox = tb_g[1] >= 0 ? dst + tb_g[1] : nullptr;
  }
  {
    double* q = dpart;
    for (int e = 0; e < 4; ++e) {
      T.part[e] = q;
      q += T.ntiles[epass[e]];
    }
  }
  // workspace: W | K | T1 | T2 | T3
  if (T.plane || mfma) t1tot = 0;
  if (!zbig) t3tot = 0;
  T.xbig = xbig;
  T.zbig = zbig;
  T.zshort = maxL3 <= ZSHORT_L;
  const size_t ws_bytes = esz * (2 * al(wtot) + al(t1tot) + al(t2tot) + al(t3tot));
  plan->ws_need = ws_bytes;
  plan->ws_at[0] = 0;
  plan->ws_at[1] = esz * al(wtot);
  plan->ws_at[2] = esz * 2 * al(wtot);
  plan->ws_at[3] = esz * (2 * al(wtot) + al(t1tot));
  plan->ws_at[4] = esz * (2 * al(wtot) + al(t1tot) + al(t2tot));
  Batch& B = plan->B;
  B.box = dbox;
  B.nbox = norb;
  B.psi = psi;
  B.hpsi = hpsi;
  B.pot = pot;
  B.ekin = ekin;
  B.epot = epot;
  plan->T = T;
  if (!defer_ws) {
    DevBuf& ws = own_ws ? plan->ws : ctx->ws;
    rc = plan_buf(ctx, arena, ws, ws_bytes);
    if (rc) {
      dev_free(plan->meta);
      dev_free(plan->tabbuf);
      delete plan;
      return rc;
    }
    bind_ws(plan, (char*)ws.p);
  }
  *out = plan;
  return BDFT_OK;
}

// Chunked plan: orbitals split into consecutive ranges whose working grids
// (W, K, T2, and T1 for line passes) take at most opt_chunk_mb MiB, each
// range a plan of its own on one shared workspace, run in order.  A chunk's
// grids are written and read again while they are still in the device's
// last-level cache (256 MiB on MI355X) instead of making an HBM round trip
// per pass; psi, hpsi and the potential stream through once.  Offsets and
// defaults are resolved per orbital first, so every part sees explicit
// psi/potential placements.
static int make_plan_chunked(bdft_ctx* ctx, int norb, const bdft_lr* const* lrs,
                             const double* pot, const int64_t* pot_off,
                             const int32_t* potdim, const int32_t* ishift,
                             const double* psi, const int64_t* psi_off, double* hpsi,
                             double* ekin, double* epot, int precision, bool own_ws,
                             bdft_plan** out, const bdft_confpot* conf = nullptr) {
  if (!ctx || !out) return BDFT_ERR_ARG;
  const size_t budget = (size_t)ctx->opt_chunk_mb << 20;
  std::vector<size_t> need;
  size_t total = 0;
  if (budget && norb > 1 && lrs && psi) {
    const size_t esz = precision == BDFT_PREC_FP64 ? 8 : 4;
    need.resize((size_t)norb);
    for (int i = 0; i < norb; ++i) {
      const bdft_lr* lr = lrs[i];
      if (!lr) return set_err(ctx, BDFT_ERR_ARG, "orbital %d: region missing", i);
      const size_t W = 8ull * lr->N[0] * lr->N[1] * lr->N[2];
      const size_t T2 = (size_t)lr->L[0] * lr->L[1] * 2 * lr->N[2];
      const size_t T1 = (size_t)lr->L[0] * 4 * lr->N[1] * lr->N[2];
      const size_t T3 = lr->L[2] > REG_L ? (size_t)lr->L[0] * lr->L[1] * lr->L[2] : 0;
      need[i] = esz * (2 * W + T2 + T1 + T3);
      total += need[i];
    }
  }
  if (!budget || total <= budget)
    return make_plan(ctx, norb, lrs, pot, pot_off, potdim, ishift, psi, psi_off, hpsi, ekin,
                     epot, precision, own_ws, out, false, conf);
  *out = nullptr;
  // the whole batch's launch choices, imposed on every part
  PlanChoice choice{};
  bdft_plan* none = nullptr;
  int rc0 = make_plan(ctx, norb, lrs, pot, pot_off, potdim, ishift, psi, psi_off, hpsi, ekin,
                      epot, precision, own_ws, &none, false, conf, false, nullptr, true,
                      nullptr, &choice);
  if (rc0) return rc0;
  std::vector<int64_t> so((size_t)norb), po((size_t)norb);
  long long run = 0;
  for (int i = 0; i < norb; ++i) {
    so[i] = psi_off ? psi_off[i] : run;
    run += (long long)lrs[i]->nvctr_c + 7LL * lrs[i]->nvctr_f;
    po[i] = pot_off ? pot_off[i] : 0;
  }
  bdft_plan* parent = new bdft_plan();
  parent->ctx = ctx;
  parent->nbox = norb;
  parent->precision = precision;
  parent->owns_ws = true;
  size_t wsmax = 0;
  for (int o0 = 0; o0 < norb;) {
    int o1 = o0 + 1;
    size_t acc = need[o0];
    while (o1 < norb && acc + need[o1] <= budget) acc += need[o1++];
    bdft_plan* part = nullptr;
    const int rc = make_plan(ctx, o1 - o0, lrs + o0, pot, po.data() + o0,
                             potdim ? potdim + 3 * o0 : nullptr,
                             ishift ? ishift + 3 * o0 : nullptr, psi, so.data() + o0, hpsi,
                             ekin + o0, epot + o0, precision, true, &part, false,
                             conf ? conf + o0 : nullptr, false, nullptr, true, &choice);
    if (rc) {
      for (bdft_plan* q : parent->parts) bdft_plan_destroy(q);
      parent->parts.clear();
      delete parent;
      return rc;
    }
    parent->parts.push_back(part);
    wsmax = std::max(wsmax, part->ws_need);
    o0 = o1;
  }
  const int rc = dev_alloc(ctx, parent->ws, wsmax);
  if (rc) {
    bdft_plan_destroy(parent);
    return rc;
  }
  for (bdft_plan* q : parent->parts) bind_ws(q, (char*)parent->ws.p);
  *out = parent;
  return BDFT_OK;
}

// the kernel sequence of one plan (or one part of a chunked plan) on s
static hipError_t launch_plan_seq(const bdft_plan* plan, hipStream_t s, hipEvent_t* ev,
                                  const Fork* fk) {
  if (plan->precision >= BDFT_PREC_BF16X3)
    return launch_mfma(plan->B, plan->T, plan->mlds, plan->wtotal, plan->need_zero,
                       plan->precision, s, ev, fk);
  return launch_passes(plan->B, plan->T, plan->maxN, plan->wtotal, plan->need_zero,
                       plan->precision == BDFT_PREC_FP32, s, ev);
}

static int run_plan(bdft_plan* plan) {
  bdft_ctx* ctx = plan->ctx;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  if (plan->exec) {
    HIPCHK(ctx, hipGraphLaunch(plan->exec, ctx->stream));
    return BDFT_OK;
  }
  const size_t nparts = plan->parts.empty() ? 1 : plan->parts.size();
  for (size_t i = 0; i < nparts; ++i) {
    const bdft_plan* part = plan->parts.empty() ? plan : plan->parts[i];
    hipEvent_t* ev = nullptr;
    if (ctx->profile) {
      std::vector<hipEvent_t> set(2 * S_COUNT);
      for (auto& x : set) {
        if (!ctx->ev_free.empty()) {
          x = ctx->ev_free.back();
          ctx->ev_free.pop_back();
        } else {
          HIPCHK(ctx, hipEventCreate(&x));
        }
      }
      ctx->ev_pending.push_back(std::move(set));
      ctx->ev_counts.push_back(i == 0 ? 1 : 0);
      ev = ctx->ev_pending.back().data();
    }
    const hipError_t e = launch_plan_seq(part, ctx->stream, ev, nullptr);
    if (e != hipSuccess)
      return set_err(ctx, BDFT_ERR_HIP, "kernel launch failed: %s", hipGetErrorString(e));
  }
  return BDFT_OK;
}

extern "C" int bdft_plan_create(bdft_ctx* ctx, int norb, const bdft_lr* const* lrs,
                                const double* pot, const int64_t* pot_off,
                                const int32_t* potdim, const int32_t* ishift,
                                const double* psi, const int64_t* psi_off, double* hpsi,
                                double* ekin, double* epot, int precision,
                                bdft_plan** out) {
  int rc = make_plan_chunked(ctx, norb, lrs, pot, pot_off, potdim, ishift, psi, psi_off, hpsi,
                             ekin, epot, precision, true, out);
  if (rc) return rc;
  bdft_plan* plan = *out;
  // capture the launch sequence once; replays reuse the instantiated graph.
  // A second capture stream (default priority) lets independent passes
  // become parallel branches of the graph.
  hipStream_t cs = nullptr;
  Fork fk{nullptr, nullptr, nullptr};
  hipError_t e = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&fk.s2, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&fk.a, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&fk.b, hipEventDisableTiming);
  if (e == hipSuccess) e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) {
    hipError_t el = hipSuccess;
    const size_t nparts = plan->parts.empty() ? 1 : plan->parts.size();
    for (size_t i = 0; i < nparts && el == hipSuccess; ++i)
      el = launch_plan_seq(plan->parts.empty() ? plan : plan->parts[i], cs, nullptr, &fk);
    e = hipStreamEndCapture(cs, &plan->graph);
    if (el != hipSuccess) e = el;
  }
  if (e == hipSuccess) e = hipGraphInstantiate(&plan->exec, plan->graph, nullptr, nullptr, 0);
  if (fk.a) hipEventDestroy(fk.a);
  if (fk.b) hipEventDestroy(fk.b);
  if (fk.s2) hipStreamDestroy(fk.s2);
  if (cs) hipStreamDestroy(cs);
  if (e != hipSuccess) {
    bdft_plan_destroy(plan);
    *out = nullptr;
    return set_err(ctx, BDFT_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e));
  }
  ctx->plans.push_back(plan);
  return BDFT_OK;
}

extern "C" int bdft_plan_execute(bdft_plan* plan) {
  if (!plan) return BDFT_ERR_ARG;
  if (!plan->ctx) return BDFT_ERR_ARG;  // its context was destroyed
  return run_plan(plan);
}

// device side of a plan: graph, parts, descriptors, tables, workspace
static void plan_release(bdft_plan* plan) {
  if (plan->ctx) {
    hipSetDevice(plan->ctx->device);
    hipStreamSynchronize(plan->ctx->stream);
  }
  if (plan->exec) hipGraphExecDestroy(plan->exec);
  if (plan->graph) hipGraphDestroy(plan->graph);
  plan->exec = nullptr;
  plan->graph = nullptr;
  for (bdft_plan* q : plan->parts) bdft_plan_destroy(q);
  plan->parts.clear();
  dev_free(plan->meta);
  dev_free(plan->tabbuf);
  if (plan->owns_ws) dev_free(plan->ws);
}

extern "C" int bdft_plan_destroy(bdft_plan* plan) {
  if (!plan) return BDFT_ERR_ARG;
  if (plan->ctx) {
    auto& ps = plan->ctx->plans;
    ps.erase(std::remove(ps.begin(), ps.end(), plan), ps.end());
  }
  plan_release(plan);
  delete plan;
  return BDFT_OK;
}

// one-shot batch: reuses the context's last plan when called again with the
// same arguments (the common case of an SCF loop)
static int locham_batch(bdft_ctx* ctx, int norb, const bdft_lr* const* lrs,
                        const double* pot, const int64_t* pot_off, const int32_t* potdim,
                        const int32_t* ishift, const double* psi, const int64_t* psi_off,
                        double* hpsi, double* ekin, double* epot,
                        const bdft_confpot* conf, int precision) {
  if (!ctx) return BDFT_ERR_ARG;
  if (norb <= 0 || !lrs) return set_err(ctx, BDFT_ERR_ARG, "null pointer or norb <= 0");
  std::vector<uintptr_t> key;
  key.reserve(16 + (size_t)norb * 9);
  key.push_back((uintptr_t)norb);
  key.push_back((uintptr_t)precision);
  key.push_back((uintptr_t)pot);
  key.push_back((uintptr_t)psi);
  key.push_back((uintptr_t)hpsi);
  key.push_back((uintptr_t)ekin);
  key.push_back((uintptr_t)epot);
  key.push_back((uintptr_t)(pot_off != nullptr) | ((uintptr_t)(potdim != nullptr) << 1) |
                ((uintptr_t)(ishift != nullptr) << 2) | ((uintptr_t)(psi_off != nullptr) << 3));
  for (int i = 0; i < norb; ++i) {
    key.push_back((uintptr_t)lrs[i]);
    if (pot_off) key.push_back((uintptr_t)pot_off[i]);
    if (psi_off) key.push_back((uintptr_t)psi_off[i]);
    for (int d = 0; d < 3; ++d) {
      if (potdim) key.push_back((uintptr_t)(uint32_t)potdim[3 * i + d]);
      if (ishift) key.push_back((uintptr_t)(uint32_t)ishift[3 * i + d]);
    }
    if (conf) {  // the confinement parameters, bit for bit
      const uintptr_t* w = (const uintptr_t*)&conf[i];
      for (size_t k = 0; k < sizeof(bdft_confpot) / sizeof(uintptr_t); ++k)
        key.push_back(w[k]);
    }
  }
  key.push_back((uintptr_t)(conf != nullptr));
  if (!ctx->cached || key != ctx->cached_key) {
    if (ctx->cached) {
      HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
      bdft_plan_destroy(ctx->cached);
      ctx->cached = nullptr;
      ctx->cached_key.clear();
    }
    bdft_plan* plan = nullptr;
    int rc = make_plan_chunked(ctx, norb, lrs, pot, pot_off, potdim, ishift, psi, psi_off,
                               hpsi, ekin, epot, precision, false, &plan, conf);
    if (rc) return rc;
    ctx->cached = plan;
    ctx->cached_key.swap(key);
  }
  return run_plan(ctx->cached);
}

extern "C" int bdft_locham_batch_d(bdft_ctx* ctx, int norb, const bdft_lr* const* lrs,
                                   const double* pot, const int64_t* pot_off,
                                   const int32_t* potdim, const int32_t* ishift,
                                   const double* psi, const int64_t* psi_off, double* hpsi,
                                   double* ekin, double* epot, int precision) {
  return locham_batch(ctx, norb, lrs, pot, pot_off, potdim, ishift, psi, psi_off, hpsi, ekin,
                      epot, nullptr, precision);
}

extern "C" int bdft_locham_conf_batch_d(bdft_ctx* ctx, int norb, const bdft_lr* const* lrs,
                                        const double* pot, const int64_t* pot_off,
                                        const int32_t* potdim, const int32_t* ishift,
                                        const double* psi, const int64_t* psi_off,
                                        double* hpsi, double* ekin, double* epot,
                                        const bdft_confpot* conf, int precision) {
  return locham_batch(ctx, norb, lrs, pot, pot_off, potdim, ishift, psi, psi_off, hpsi, ekin,
                      epot, conf, precision);
}

// The batch of bdft_locham_batch_d with no device allocation: the orbitals
// run in consecutive chunks, each chunk's plan (descriptors uploaded with
// stream-ordered copies, operator tables, working grids) laid out in the
// device's fixed array, one chunk after the other on the context's stream.
// The host waits for a chunk's kernels before the next chunk's descriptors
// are written over theirs.  Every buffer the kernels need zeroed is zeroed
// by a kernel of the sequence itself.  Returns after the stream has
// finished the batch (the host copies of the descriptors live until then).
extern "C" int bdft_locham_batch_fixed_d(bdft_ctx* ctx, int norb, const bdft_lr* const* lrs,
                                         const double* pot, const int64_t* pot_off,
                                         const int32_t* potdim, const int32_t* ishift,
                                         const double* psi, const int64_t* psi_off,
                                         double* hpsi, double* ekin, double* epot,
                                         int precision) {
  if (!ctx) return BDFT_ERR_ARG;
  if (norb <= 0 || !lrs || !pot || !psi || !hpsi || !ekin || !epot)
    return set_err(ctx, BDFT_ERR_ARG, "null pointer or norb <= 0");
  if (!ctx->fixed)
    return set_err(ctx, BDFT_ERR_HIP, "fixed device array unavailable (%s)",
                   ctx->fixed_err.c_str());
  if (precision < BDFT_PREC_FP64 || precision > BDFT_PREC_F16X3)
    return set_err(ctx, BDFT_ERR_UNSUPPORTED, "precision %d not available", precision);
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const size_t cap = BDFT_FIXED_BYTES;
  // working grids per orbital (W, K, T2, and T1 for line passes, T3 for
  // long z lines), the bound a chunk is first cut to
  const size_t esz = precision == BDFT_PREC_FP64 ? 8 : 4;
  std::vector<size_t> need((size_t)norb);
  std::vector<int64_t> so((size_t)norb), po((size_t)norb);
  long long run = 0;
  for (int i = 0; i < norb; ++i) {
    const bdft_lr* lr = lrs[i];
    if (!lr || lr->ctx != ctx)
      return set_err(ctx, BDFT_ERR_ARG, "orbital %d: region missing or from another context", i);
    const size_t W = 8ull * lr->N[0] * lr->N[1] * lr->N[2];
    const size_t T2 = (size_t)lr->L[0] * lr->L[1] * 2 * lr->N[2];
    const size_t T1 = (size_t)lr->L[0] * 4 * lr->N[1] * lr->N[2];
    const size_t T3 = lr->L[2] > REG_L ? (size_t)lr->L[0] * lr->L[1] * lr->L[2] : 0;
    need[i] = esz * (2 * W + T2 + T1 + T3);
    if (need[i] > cap)
      return set_err(ctx, BDFT_ERR_UNSUPPORTED,
                     "orbital %d needs %zu bytes of working grids, the fixed device array "
                     "holds %zu",
                     i, need[i], cap);
    so[i] = psi_off ? psi_off[i] : run;
    run += (long long)lr->nvctr_c + 7LL * lr->nvctr_f;
    po[i] = pot_off ? pot_off[i] : 0;
  }
  // the whole batch's launch choices, imposed on every chunk (the kernels
  // bdft_locham_batch_d runs)
  PlanChoice choice{};
  bdft_plan* none = nullptr;
  int rc = make_plan(ctx, norb, lrs, pot, po.data(), potdim, ishift, psi, so.data(), hpsi, ekin,
                     epot, precision, true, &none, false, nullptr, false, nullptr, true, nullptr,
                     &choice);
  if (rc) return rc;
  std::vector<std::vector<unsigned char>> keep;
  for (int o0 = 0; o0 < norb && rc == BDFT_OK;) {
    int o1 = o0 + 1;
    size_t acc = need[o0];
    while (o1 < norb && acc + need[o1] <= cap) acc += need[o1++];
    if (o0 > 0) {
      const hipError_t ew = hipStreamSynchronize(ctx->stream);
      if (ew != hipSuccess) {
        rc = set_err(ctx, BDFT_ERR_HIP, "hipStreamSynchronize failed: %s",
                     hipGetErrorString(ew));
        break;
      }
      keep.clear();
    }
    for (;;) {
      Arena a;
      a.base = ctx->fixed;
      a.size = cap;
      a.stream = ctx->stream;
      a.keep = &keep;
      bdft_plan* part = nullptr;
      rc = make_plan(ctx, o1 - o0, lrs + o0, pot, po.data() + o0,
                     potdim ? potdim + 3 * o0 : nullptr, ishift ? ishift + 3 * o0 : nullptr,
                     psi, so.data() + o0, hpsi, ekin + o0, epot + o0, precision, true, &part,
                     false, nullptr, false, &a, false, &choice);
      if (rc == BDFT_ERR_UNSUPPORTED && a.overflow && o1 - o0 > 1) {
        o1 = o0 + (o1 - o0) / 2;  // descriptors and tables did not fit beside the grids
        rc = BDFT_OK;
        continue;
      }
      if (rc) break;
      const hipError_t e = launch_plan_seq(part, ctx->stream, nullptr, nullptr);
      delete part;  // its buffers are slices of the fixed array
      if (e != hipSuccess)
        rc = set_err(ctx, BDFT_ERR_HIP, "kernel launch failed: %s", hipGetErrorString(e));
      break;
    }
    o0 = o1;
  }
  const hipError_t es = hipStreamSynchronize(ctx->stream);
  if (rc) return rc;
  if (es != hipSuccess)
    return set_err(ctx, BDFT_ERR_HIP, "hipStreamSynchronize failed: %s", hipGetErrorString(es));
  return BDFT_OK;
}

// ---------------------------------------------------------------------------
// density: local_partial_density_OCL (interface_ocl.f90:908-1040) /
// local_partial_density (sumrho.f90:377-469)
// ---------------------------------------------------------------------------
extern "C" int bdft_density_d(bdft_ctx* ctx, const bdft_lr* lr, int norb, const double* psi,
                              const double* hfac, const int32_t* spin, int nspin, double* rho,
                              int precision) {
  if (!ctx) return BDFT_ERR_ARG;
  if (!lr || norb <= 0 || !psi || !hfac || !rho)
    return set_err(ctx, BDFT_ERR_ARG, "null pointer or norb <= 0");
  if (nspin != 1 && nspin != 2) return set_err(ctx, BDFT_ERR_ARG, "nspin must be 1 or 2");
  if (precision != BDFT_PREC_FP64 && precision != BDFT_PREC_FP32)
    return set_err(ctx, BDFT_ERR_UNSUPPORTED,
                   "density: precision %d not available (FP64 or FP32)", precision);
  std::vector<int32_t> sp((size_t)norb, 0);
  if (spin)
    for (int i = 0; i < norb; ++i) {
      if (spin[i] < 0 || spin[i] >= nspin)
        return set_err(ctx, BDFT_ERR_ARG, "orbital %d: spin index %d outside [0,%d)", i,
                       spin[i], nspin);
      sp[i] = spin[i];
    }
  HIPCHK(ctx, hipSetDevice(ctx->device));
  // the plan (tile tables and work grids) depends on the region, the
  // orbital count and the precision only: reused while they stay the same,
  // with the orbitals' address set per call
  const std::vector<uintptr_t> key{(uintptr_t)lr, (uintptr_t)norb, (uintptr_t)precision};
  if (!ctx->den_plan || key != ctx->den_key) {
    if (ctx->den_plan) {
      HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
      bdft_plan_destroy(ctx->den_plan);
      ctx->den_plan = nullptr;
      ctx->den_key.clear();
    }
    std::vector<const bdft_lr*> lrs((size_t)norb, lr);
    bdft_plan* np = nullptr;
    int rc0 = make_plan(ctx, norb, lrs.data(), nullptr, nullptr, nullptr, nullptr, psi, nullptr,
                        nullptr, nullptr, nullptr, precision, true, &np, true, nullptr, true);
    if (rc0) return rc0;
    ctx->den_plan = np;
    ctx->den_key = key;
  }
  bdft_plan* plan = ctx->den_plan;
  plan->B.psi = psi;
  DenArgs D{};
  D.nisf = (long long)lr->L[0] * lr->L[1] * lr->L[2];
  D.norb = norb;
  D.nspin = nspin;
  D.nxt = cdiv(lr->L[0], TW);
  D.nbund = D.nxt * lr->L[1];
  const int want = std::max(1, std::min(std::min(norb, 64), cdiv(4096, D.nbund)));
  D.gsize = cdiv(norb, want);
  D.ngroups = cdiv(norb, D.gsize);
  auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  const size_t nb_h = sizeof(double) * norb, nb_s = sizeof(int32_t) * norb;
  const size_t nb_p = sizeof(double) * (size_t)D.ngroups * nspin * (size_t)D.nisf;
  DevBuf& buf = ctx->den_buf;
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));  // the buffer may be reallocated
  int rc = dev_alloc(ctx, buf, al(nb_h) + al(nb_s) + al(nb_p));
  if (rc) return rc;
  char* q = (char*)buf.p;
  hipError_t e = hipMemcpy(q, hfac, nb_h, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(q + al(nb_h), sp.data(), nb_s, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    D.hfac = (const double*)q;
    D.spin = (const int*)(q + al(nb_h));
    D.part = (double*)(q + al(nb_h) + al(nb_s));
    D.rho = rho;
    e = launch_density(plan->B, plan->T, D, plan->maxN, plan->wtotal, plan->need_zero,
                       precision == BDFT_PREC_FP32, ctx->stream);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess)
    return set_err(ctx, BDFT_ERR_HIP, "density: %s", hipGetErrorString(e));
  return BDFT_OK;
}

// host arrays in and out (the Fortran side passes host psi and rho_p, as
// local_partial_density_OCL does)
extern "C" int bdft_density(bdft_ctx* ctx, const bdft_lr* lr, int norb, const double* psi,
                            const double* hfac, const int32_t* spin, int nspin, double* rho,
                            int precision) {
  if (!ctx) return BDFT_ERR_ARG;
  if (!lr || norb <= 0 || !psi || !hfac || !rho || (nspin != 1 && nspin != 2))
    return set_err(ctx, BDFT_ERR_ARG, "null pointer, norb <= 0 or nspin not 1 or 2");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const size_t ncoef = (size_t)lr->nvctr_c + 7 * (size_t)lr->nvctr_f;
  const size_t nrho = (size_t)lr->L[0] * lr->L[1] * lr->L[2] * nspin;
  double *dpsi = nullptr, *drho = nullptr;
  int rc = BDFT_OK;
  hipError_t e = hipMalloc(&dpsi, ncoef * norb * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&drho, nrho * sizeof(double));
  if (e != hipSuccess) rc = set_err(ctx, BDFT_ERR_NOMEM, "device buffers: %s", hipGetErrorString(e));
  hipStream_t s = ctx->stream;
  if (!rc) {
    e = hipMemcpyAsync(dpsi, psi, ncoef * norb * sizeof(double), hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
      e = hipMemcpyAsync(drho, rho, nrho * sizeof(double), hipMemcpyHostToDevice, s);
    if (e != hipSuccess) rc = set_err(ctx, BDFT_ERR_HIP, "H2D copy: %s", hipGetErrorString(e));
  }
  if (!rc) rc = bdft_density_d(ctx, lr, norb, dpsi, hfac, spin, nspin, drho, precision);
  if (!rc) {
    e = hipMemcpyAsync(rho, drho, nrho * sizeof(double), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) rc = set_err(ctx, BDFT_ERR_HIP, "D2H copy: %s", hipGetErrorString(e));
  }
  hipStreamSynchronize(s);
  if (dpsi) hipFree(dpsi);
  if (drho) hipFree(drho);
  return rc;
}

// ---------------------------------------------------------------------------
// region transfers: psi_to_locreg2 / Lpsi_to_global2
// (locregs/locreg_operations.f90:2143-2318, 1665-1837)
// ---------------------------------------------------------------------------
static int xfer(bdft_ctx* ctx, const bdft_lr* glr, int norb, const bdft_lr* const* llrs,
                const int32_t* ns, const double* src, const int64_t* soff, double* dst,
                const int64_t* doff, bool to_local) {
  if (!ctx) return BDFT_ERR_ARG;
  if (!glr || norb <= 0 || !llrs || !ns || !src || !dst)
    return set_err(ctx, BDFT_ERR_ARG, "null pointer or norb <= 0");
  if (glr->ctx != ctx) return set_err(ctx, BDFT_ERR_ARG, "global region from another context");
  if (glr->geocode != 'F')
    return set_err(ctx, BDFT_ERR_UNSUPPORTED,
                   "region transfers are defined between free-BC regions "
                   "(shift_locreg_indexes, locreg_operations.f90:2338-2343)");
  const long long gnc = (long long)glr->nvctr_c + 7LL * glr->nvctr_f;
  std::vector<XferDesc> d((size_t)norb);
  long long lacc = 0;
  long long tiles = 0;
  for (int i = 0; i < norb; ++i) {
    const bdft_lr* L = llrs[i];
    if (!L || L->ctx != ctx)
      return set_err(ctx, BDFT_ERR_ARG, "region %d missing or from another context", i);
    if (L->geocode != 'F')
      return set_err(ctx, BDFT_ERR_UNSUPPORTED, "region %d is not a free-BC region", i);
    if (7LL * L->nvctr_f > 0x7fffffffLL)
      return set_err(ctx, BDFT_ERR_UNSUPPORTED, "region %d: too many fine points", i);
    XferDesc& x = d[i];
    x.cmap = L->d_cmap;
    x.fmap = L->d_fmap;
    x.nvctr_c = L->nvctr_c;
    x.nvctr_f = L->nvctr_f;
    x.m1 = L->n[0] + 1;
    x.m2 = L->n[1] + 1;
    for (int k = 0; k < 3; ++k) x.ns[k] = ns[3 * i + k];
    const long long lnc = (long long)L->nvctr_c + 7LL * L->nvctr_f;
    if (to_local) {
      x.goff = soff ? soff[i] : gnc * i;
      x.loff = doff ? doff[i] : lacc;
    } else {
      x.loff = soff ? soff[i] : lacc;
      x.goff = doff ? doff[i] : gnc * i;
    }
    lacc += lnc;
    x.tile0 = (int)tiles;
    tiles += xfer_tiles(L->nvctr_c, L->nvctr_f);
  }
  if (tiles > 0x7fffffffLL) return set_err(ctx, BDFT_ERR_UNSUPPORTED, "batch too large");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  const size_t nb_d = sizeof(XferDesc) * norb;
  DevBuf& buf = ctx->xfer_buf;
  const unsigned char* dh = (const unsigned char*)d.data();
  const bool same = buf.p && ctx->xfer_last.size() == nb_d &&
                    std::memcmp(ctx->xfer_last.data(), dh, nb_d) == 0;
  hipError_t e = hipSuccess;
  if (!same) {
    // the previous call's kernel may still read the descriptors
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    ctx->xfer_last.clear();
    int rc = dev_alloc(ctx, buf, al(nb_d) + al(sizeof(int)));
    if (rc) return rc;
    e = hipMemcpy(buf.p, dh, nb_d, hipMemcpyHostToDevice);
    if (e == hipSuccess) ctx->xfer_last.assign(dh, dh + nb_d);
  }
  char* q = (char*)buf.p;
  int* miss = (int*)(q + al(nb_d));
  // only scatters into the global orbital count missing points
  if (e == hipSuccess && !to_local) e = hipMemsetAsync(miss, 0, sizeof(int), ctx->stream);
  int nmiss = 0;
  if (e == hipSuccess) {
    XferArgs A{};
    A.d = (const XferDesc*)q;
    A.nd = norb;
    A.gcinv = glr->d_cinv;
    A.gfinv = glr->d_finv;
    A.g1 = glr->n[0] + 1;
    A.g2 = glr->n[1] + 1;
    A.g3 = glr->n[2] + 1;
    A.gnvctr_c = glr->nvctr_c;
    A.src = src;
    A.dst = dst;
    A.to_local = to_local ? 1 : 0;
    A.miss = miss;
    e = launch_xfer(A, (int)tiles, ctx->stream);
  }
  if (e == hipSuccess && !to_local)
    e = hipMemcpyAsync(&nmiss, miss, sizeof(int), hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return set_err(ctx, BDFT_ERR_HIP, "transfer: %s", hipGetErrorString(e));
  if (nmiss > 0)
    return set_err(ctx, BDFT_ERR_DESC,
                   "%d local points have no counterpart in the global region", nmiss);
  return BDFT_OK;
}

extern "C" int bdft_psi_to_locreg_d(bdft_ctx* ctx, const bdft_lr* glr, int norb,
                                    const bdft_lr* const* llrs, const int32_t* ns,
                                    const double* gpsi, const int64_t* goff, double* lpsi,
                                    const int64_t* loff) {
  return xfer(ctx, glr, norb, llrs, ns, gpsi, goff, lpsi, loff, true);
}

extern "C" int bdft_locreg_to_psi_d(bdft_ctx* ctx, const bdft_lr* glr, int norb,
                                    const bdft_lr* const* llrs, const int32_t* ns,
                                    const double* lpsi, const int64_t* loff, double* gpsi,
                                    const int64_t* goff) {
  return xfer(ctx, glr, norb, llrs, ns, lpsi, loff, gpsi, goff, false);
}

// ---------------------------------------------------------------------------
// preconditioner: the GPU part of preconditionall_OCL (interface_ocl.f90:
// 622-880), ocl_preconditioner_generic_k (OpenCL/Preconditioner.c:245-358)
// ---------------------------------------------------------------------------
extern "C" int bdft_precond_cg_d(bdft_ctx* ctx, int norb, const bdft_lr* const* lrs,
                                 const int64_t* psi_off, const double* cprecr, int ncong,
                                 double* x, const double* b) {
  if (!ctx) return BDFT_ERR_ARG;
  if (norb <= 0 || !lrs || !cprecr || !x || !b)
    return set_err(ctx, BDFT_ERR_ARG, "null pointer or norb <= 0");
  if (ncong < 0 || ncong > 10000) return set_err(ctx, BDFT_ERR_ARG, "ncong %d", ncong);
  HIPCHK(ctx, hipSetDevice(ctx->device));
  long long ntot = 0, run = 0;
  std::vector<double> pc((size_t)norb * 9);
  for (int i = 0; i < norb; ++i) {
    const bdft_lr* lr = lrs[i];
    if (!lr) return set_err(ctx, BDFT_ERR_ARG, "orbital %d: null region", i);
    const long long nc = (long long)lr->nvctr_c + 7LL * lr->nvctr_f;
    const long long off = psi_off ? (long long)psi_off[i] : run;  // as make_plan
    run += nc;
    if (off < 0) return set_err(ctx, BDFT_ERR_ARG, "negative offset");
    ntot = std::max(ntot, off + nc);
    const double c = cprecr[i];
    if (!(c >= 0.0)) return set_err(ctx, BDFT_ERR_ARG, "orbital %d: cprecr %g", i, c);
    // scale_psi_d / uncompress_scale_d (OpenCL/Uncompress.c:377-470), h = hgrid/2
    double hh[3];
    for (int d = 0; d < 3; ++d) {
      const double h = 0.5 * lr->hgrid[d];
      hh[d] = 0.125 / (h * h);
    }
    pc[9 * i] = c;
    for (int t = 0; t < 8; ++t) {
      double v = c;
      for (int d = 0; d < 3; ++d) v += ((t >> d) & 1 ? 24.8758460293923314 : 3.55369228991319019) * hh[d];
      pc[9 * i + 1 + t] = 1.0 / std::sqrt(v);
    }
  }
  // the plan (descriptors, tile tables, W and K) depends on the geometry
  // only: reused while the regions and offsets stay the same
  std::vector<uintptr_t> key;
  key.reserve(2 + 2 * (size_t)norb);
  key.push_back((uintptr_t)norb);
  key.push_back((uintptr_t)(psi_off != nullptr));
  for (int i = 0; i < norb; ++i) {
    key.push_back((uintptr_t)lrs[i]);
    if (psi_off) key.push_back((uintptr_t)psi_off[i]);
  }
  if (!ctx->pc_plan || key != ctx->pc_key) {
    if (ctx->pc_plan) {
      HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
      bdft_plan_destroy(ctx->pc_plan);
      ctx->pc_plan = nullptr;
      ctx->pc_key.clear();
    }
    bdft_plan* np = nullptr;
    int rc0 = make_plan(ctx, norb, lrs, nullptr, nullptr, nullptr, nullptr, x, psi_off,
                        nullptr, nullptr, nullptr, BDFT_PREC_FP64, true, &np, true);
    if (rc0) return rc0;
    ctx->pc_plan = np;
    ctx->pc_key.swap(key);
  }
  bdft_plan* plan = ctx->pc_plan;
  auto al = [](size_t v) { return (v + 255) & ~(size_t)255; };
  const size_t nvec = sizeof(double) * (size_t)ntot;
  const size_t npc = sizeof(double) * pc.size(), nsc = sizeof(double) * 4 * (size_t)norb;
  const size_t npart = sizeof(double) * std::max(1, plan->T.ntiles[P_UNC]);
  DevBuf& buf = ctx->pc_buf;
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));  // the buffer may be reallocated
  int rc = dev_alloc(ctx, buf, 3 * al(nvec) + al(npc) + al(nsc) + al(npart));
  if (rc) return rc;
  char* q = (char*)buf.p;
  PcArgs A{};
  A.r = (double*)q;
  A.d = (double*)(q + al(nvec));
  A.q = (double*)(q + 2 * al(nvec));
  A.pc = (const double*)(q + 3 * al(nvec));
  A.sc = (double*)(q + 3 * al(nvec) + al(npc));
  A.part = (double*)(q + 3 * al(nvec) + al(npc) + al(nsc));
  A.x = x;
  A.b = b;
  A.ncong = ncong;
  hipError_t e = hipMemcpyAsync((void*)A.pc, pc.data(), npc, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess)
    e = launch_precond(plan->B, plan->T, plan->maxN, plan->wtotal, plan->need_zero, A,
                       ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return set_err(ctx, BDFT_ERR_HIP, "preconditioner: %s", hipGetErrorString(e));
  return BDFT_OK;
}

// host arrays: norb orbitals of one region (the per-orbital transfers of
// preconditionall_OCL, interface_ocl.f90:769-843, done once for the batch)
extern "C" int bdft_precondition(bdft_ctx* ctx, const bdft_lr* lr, int norb,
                                 const double* cprecr, int ncong, double* x, const double* b) {
  if (!ctx) return BDFT_ERR_ARG;
  if (!lr || norb <= 0 || !x || !b || !cprecr)
    return set_err(ctx, BDFT_ERR_ARG, "null pointer or norb <= 0");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const size_t n = ((size_t)lr->nvctr_c + 7 * (size_t)lr->nvctr_f) * norb;
  DevBuf buf;
  int rc = dev_alloc(ctx, buf, 2 * sizeof(double) * std::max<size_t>(n, 1));
  if (rc) return rc;
  double* dx = (double*)buf.p;
  double* db = dx + std::max<size_t>(n, 1);
  hipError_t e = hipMemcpy(dx, x, sizeof(double) * n, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(db, b, sizeof(double) * n, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    dev_free(buf);
    return set_err(ctx, BDFT_ERR_HIP, "H2D copy: %s", hipGetErrorString(e));
  }
  std::vector<const bdft_lr*> lrs((size_t)norb, lr);
  rc = bdft_precond_cg_d(ctx, norb, lrs.data(), nullptr, cprecr, ncong, dx, db);
  if (!rc) {
    e = hipMemcpy(x, dx, sizeof(double) * n, hipMemcpyDeviceToHost);
    if (e != hipSuccess) rc = set_err(ctx, BDFT_ERR_HIP, "D2H copy: %s", hipGetErrorString(e));
  }
  dev_free(buf);
  return rc;
}

// ---------------------------------------------------------------------------
// convolution bounds (locregs/bounds.f90:555-613, locreg_bounds)
// ---------------------------------------------------------------------------
extern "C" int bdft_bounds_layout(const int n[3], const int nfl[3], const int nfu[3],
                                  int64_t offset[18], int32_t dims[68]) {
  if (!n || !nfl || !nfu || !offset || !dims) return BDFT_ERR_ARG;
  return bounds_layout(n, nfl, nfu, offset, dims) ? BDFT_ERR_ARG : BDFT_OK;
}

static int bounds_check(bdft_ctx* ctx, const bdft_lr* lr, int64_t* total) {
  if (!lr) return set_err(ctx, BDFT_ERR_ARG, "null region");
  if (lr->ctx != ctx) return set_err(ctx, BDFT_ERR_ARG, "region from another context");
  if (lr->geocode != 'F')
    return set_err(ctx, BDFT_ERR_UNSUPPORTED, "convolution bounds are defined for free BC");
  for (int d = 0; d < 3; ++d) {
    if (lr->n[d] > BOUNDS_MAXN)
      return set_err(ctx, BDFT_ERR_UNSUPPORTED,
                     "n%d=%d: bounds need n <= %d (the 1000 sentinel)", d + 1, lr->n[d],
                     BOUNDS_MAXN);
    if (lr->nfl[d] < 0 || lr->nfl[d] > lr->nfu[d] || lr->nfu[d] > lr->n[d])
      return set_err(ctx, BDFT_ERR_DESC, "fine box [%d,%d] outside 0..%d", lr->nfl[d],
                     lr->nfu[d], lr->n[d]);
  }
  int64_t off[BOUNDS_NARR + 1];
  int32_t dims[4 * BOUNDS_NARR];
  bounds_layout(lr->n, lr->nfl, lr->nfu, off, dims);
  *total = off[BOUNDS_NARR];
  return BDFT_OK;
}

extern "C" int bdft_locreg_bounds_d(bdft_ctx* ctx, const bdft_lr* lr, int32_t* bounds) {
  if (!ctx) return BDFT_ERR_ARG;
  if (!bounds) return set_err(ctx, BDFT_ERR_ARG, "null output");
  int64_t total = 0;
  int rc = bounds_check(ctx, lr, &total);
  if (rc) return rc;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  BoundsArgs A{};
  for (int d = 0; d < 3; ++d) {
    A.n[d] = lr->n[d];
    A.nfl[d] = lr->nfl[d];
    A.nfu[d] = lr->nfu[d];
  }
  A.cinv = lr->d_cinv;
  A.finv = lr->d_finv;
  A.out = bounds;
  HIPCHK(ctx, launch_bounds(A, ctx->stream));
  return BDFT_OK;
}

extern "C" int bdft_locreg_bounds(bdft_ctx* ctx, const bdft_lr* lr, int32_t* bounds) {
  if (!ctx) return BDFT_ERR_ARG;
  if (!bounds) return set_err(ctx, BDFT_ERR_ARG, "null output");
  int64_t total = 0;
  int rc = bounds_check(ctx, lr, &total);
  if (rc) return rc;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  DevBuf buf;
  rc = dev_alloc(ctx, buf, sizeof(int32_t) * (size_t)total);
  if (rc) return rc;
  rc = bdft_locreg_bounds_d(ctx, lr, (int32_t*)buf.p);
  if (!rc) {
    hipError_t e = hipMemcpyAsync(bounds, buf.p, sizeof(int32_t) * (size_t)total,
                                  hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) rc = set_err(ctx, BDFT_ERR_HIP, "D2H copy: %s", hipGetErrorString(e));
  }
  hipStreamSynchronize(ctx->stream);
  dev_free(buf);
  return rc;
}

// potential component of each orbital: spin up (spinsgn > 0, or no spinsgn)
// reads pot[0..npot), spin down pot[npot..2 npot), as local_hamiltonian_OCL
// picks rhopot_up / rhopot_down from orbs%spinsgn (interface_ocl.f90:487-491)
static int spin_offsets(bdft_ctx* ctx, int norb, int nspin, const double* spinsgn,
                        long long npot, std::vector<int64_t>& off) {
  if (nspin != 1 && nspin != 2)
    return set_err(ctx, BDFT_ERR_ARG, "nspin must be 1 or 2 (got %d)", nspin);
  off.assign((size_t)norb, 0);
  if (!spinsgn) return BDFT_OK;
  for (int i = 0; i < norb; ++i) {
    if (spinsgn[i] > 0.0) continue;
    if (nspin == 1)
      return set_err(ctx, BDFT_ERR_ARG,
                     "orbital %d is spin down (spinsgn %g) but nspin = 1", i, spinsgn[i]);
    off[i] = npot;
  }
  return BDFT_OK;
}

extern "C" int bdft_locham_d(bdft_ctx* ctx, const bdft_lr* lr, int norb, int nspin,
                             const double* pot, const double* spinsgn, const double* psi,
                             double* hpsi, double* ekin, double* epot, int precision) {
  if (!ctx) return BDFT_ERR_ARG;
  if (!lr || norb <= 0) return set_err(ctx, BDFT_ERR_ARG, "null region or norb <= 0");
  std::vector<int64_t> poff;
  const long long npot = (long long)lr->L[0] * lr->L[1] * lr->L[2];
  int rc = spin_offsets(ctx, norb, nspin, spinsgn, npot, poff);
  if (rc) return rc;
  std::vector<const bdft_lr*> lrs((size_t)norb, lr);
  return bdft_locham_batch_d(ctx, norb, lrs.data(), pot, poff.data(), nullptr, nullptr, psi,
                             nullptr, hpsi, ekin, epot, precision);
}

extern "C" int bdft_local_hamiltonian(bdft_ctx* ctx, const bdft_lr* lr, int norb, int nspin,
                                      const double* pot, const double* spinsgn,
                                      const double* psi, double* hpsi, double* ekin,
                                      double* epot, int precision) {
  if (!ctx) return BDFT_ERR_ARG;
  if (!lr || norb <= 0 || !pot || !psi || !hpsi || !ekin || !epot)
    return set_err(ctx, BDFT_ERR_ARG, "null pointer or norb <= 0");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const size_t ncoef = (size_t)lr->nvctr_c + 7 * (size_t)lr->nvctr_f;
  const size_t npot = (size_t)lr->L[0] * lr->L[1] * lr->L[2];
  std::vector<int64_t> poff;
  int rc = spin_offsets(ctx, norb, nspin, spinsgn, (long long)npot, poff);
  if (rc) return rc;
  double *dpot = nullptr, *dpsi = nullptr, *dh = nullptr, *den = nullptr;
  hipError_t e = hipMalloc(&dpot, npot * nspin * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&dpsi, ncoef * norb * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&dh, ncoef * norb * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&den, 2 * (size_t)norb * sizeof(double));
  if (e != hipSuccess) rc = set_err(ctx, BDFT_ERR_NOMEM, "device buffers: %s", hipGetErrorString(e));
  hipStream_t s = ctx->stream;
  if (!rc) {
    e = hipMemcpyAsync(dpot, pot, npot * nspin * sizeof(double), hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
      e = hipMemcpyAsync(dpsi, psi, ncoef * norb * sizeof(double), hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
      e = hipMemcpyAsync(dh, hpsi, ncoef * norb * sizeof(double), hipMemcpyHostToDevice, s);
    if (e != hipSuccess) rc = set_err(ctx, BDFT_ERR_HIP, "H2D copy: %s", hipGetErrorString(e));
  }
  if (!rc) {
    std::vector<const bdft_lr*> lrs((size_t)norb, lr);
    bdft_plan* plan = nullptr;
    rc = make_plan_chunked(ctx, norb, lrs.data(), dpot, poff.data(), nullptr, nullptr, dpsi,
                           nullptr, dh, den, den + norb, precision, true, &plan);
    if (!rc) {
      rc = run_plan(plan);
      if (!rc) {
        e = hipMemcpyAsync(hpsi, dh, ncoef * norb * sizeof(double), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess)
          e = hipMemcpyAsync(ekin, den, norb * sizeof(double), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess)
          e = hipMemcpyAsync(epot, den + norb, norb * sizeof(double), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) rc = set_err(ctx, BDFT_ERR_HIP, "D2H copy: %s", hipGetErrorString(e));
      }
      bdft_plan_destroy(plan);
    }
  }
  hipStreamSynchronize(s);
  if (dpot) hipFree(dpot);
  if (dpsi) hipFree(dpsi);
  if (dh) hipFree(dh);
  if (den) hipFree(den);
  return rc;
}

// ---------------------------------------------------------------------------
// compressed <-> mixed layout transfers
// ---------------------------------------------------------------------------
static int transfer(bdft_ctx* ctx, const bdft_lr* lr, int norb, const double* psi,
                    double* hpsi, double* psig, bool compress) {
  if (!ctx) return BDFT_ERR_ARG;
  if (!lr || norb <= 0 || !psig || (compress ? !hpsi : !psi))
    return set_err(ctx, BDFT_ERR_ARG, "null pointer or norb <= 0");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  std::vector<BoxDesc> box((size_t)norb);
  const long long vol = 8LL * lr->N[0] * lr->N[1] * lr->N[2];
  const long long ncoef = (long long)lr->nvctr_c + 7LL * lr->nvctr_f;
  for (int i = 0; i < norb; ++i) {
    BoxDesc& b = box[i];
    memset(&b, 0, sizeof b);
    for (int d = 0; d < 3; ++d) {
      b.N[d] = lr->N[d];
      b.L[d] = lr->L[d];
    }
    b.per = lr->geocode == 'P';
    b.nvctr_c = lr->nvctr_c;
    b.nvctr_f = lr->nvctr_f;
    b.cmap = lr->d_cmap;
    b.fmap = lr->d_fmap;
    b.w_off = vol * i;
    b.psi_off = ncoef * i;
  }
  DevBuf meta;
  int rc = dev_alloc(ctx, meta, sizeof(BoxDesc) * norb);
  if (rc) return rc;
  hipError_t e = hipMemcpy(meta.p, box.data(), sizeof(BoxDesc) * norb, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    Batch B{};
    B.box = (const BoxDesc*)meta.p;
    B.nbox = norb;
    B.psi = psi;
    B.hpsi = hpsi;
    B.W = psig;
    TileTable t{nullptr, norb, box_tiles(box[0], P_UNC, ZGROUP)};
    const long long nt = (long long)t.uniform * norb;
    if (t.uniform == 0 || nt > 0x7fffffffLL) {
      dev_free(meta);
      return t.uniform == 0 ? BDFT_OK : set_err(ctx, BDFT_ERR_UNSUPPORTED, "too many points");
    }
    e = launch_transfer(B, t, (int)nt, vol * norb, compress, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  }
  dev_free(meta);
  if (e != hipSuccess) return set_err(ctx, BDFT_ERR_HIP, "transfer: %s", hipGetErrorString(e));
  return BDFT_OK;
}

extern "C" int bdft_uncompress_d(bdft_ctx* ctx, const bdft_lr* lr, int norb, const double* psi,
                                 double* psig) {
  return transfer(ctx, lr, norb, psi, nullptr, psig, false);
}

extern "C" int bdft_compress_acc_d(bdft_ctx* ctx, const bdft_lr* lr, int norb,
                                   const double* psig, double* hpsi) {
  return transfer(ctx, lr, norb, nullptr, hpsi, const_cast<double*>(psig), true);
}

// ---------------------------------------------------------------------------
// stage profiling with HIP events
// ---------------------------------------------------------------------------
static const char* kStageNames[S_COUNT] = {
    "zero",   "uncompress", "x_grow_kin", "y_grow", "y_kin",    "z_kin",
    "z_pot",  "y_shrink",   "x_shrink",   "compress", "energy", "xy_fwd",
    "xy_bwd_compress"};

extern "C" int bdft_profile_enable(bdft_ctx* ctx, int on) {
  if (!ctx) return BDFT_ERR_ARG;
  ctx->profile = on != 0;
  return BDFT_OK;
}

extern "C" int bdft_profile_read(bdft_ctx* ctx, int nstage, double* ms, int64_t* nlaunch,
                                 int reset) {
  if (!ctx || nstage < 0 || (nstage > 0 && !ms)) return BDFT_ERR_ARG;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  for (auto& set : ctx->ev_pending) {
    for (int s = 0; s < S_COUNT; ++s) {
      float t = 0.f;
      HIPCHK(ctx, hipEventElapsedTime(&t, set[2 * s], set[2 * s + 1]));
      ctx->prof_ms[s] += t;
    }
    for (auto x : set) ctx->ev_free.push_back(x);
  }
  for (char c : ctx->ev_counts) ctx->prof_n += c;
  ctx->ev_pending.clear();
  ctx->ev_counts.clear();
  for (int s = 0; s < nstage && s < S_COUNT; ++s) ms[s] = ctx->prof_ms[s];
  if (nlaunch) *nlaunch = ctx->prof_n;
  if (reset) {
    for (int s = 0; s < S_COUNT; ++s) ctx->prof_ms[s] = 0.0;
    ctx->prof_n = 0;
  }
  return BDFT_OK;
}

static void drop_cached_plans(bdft_ctx* ctx) {
  hipSetDevice(ctx->device);
  hipStreamSynchronize(ctx->stream);
  if (ctx->cached) bdft_plan_destroy(ctx->cached);
  if (ctx->pc_plan) bdft_plan_destroy(ctx->pc_plan);
  if (ctx->den_plan) bdft_plan_destroy(ctx->den_plan);
  ctx->cached = ctx->pc_plan = ctx->den_plan = nullptr;
  ctx->cached_key.clear();
  ctx->pc_key.clear();
  ctx->den_key.clear();
}

extern "C" int bdft_ctx_set_option(bdft_ctx* ctx, int option, int value) {
  if (!ctx) return BDFT_ERR_ARG;
  switch (option) {
    case BDFT_OPT_LINE_PASSES:
      if (value != 0 && value != 1)
        return set_err(ctx, BDFT_ERR_ARG, "line passes: 0 or 1 (got %d)", value);
      if (ctx->force_lines != (value == 1)) {
        drop_cached_plans(ctx);
        ctx->force_lines = value == 1;
      }
      return BDFT_OK;
    case BDFT_OPT_ZPOT:
      if (value < 0 || value > 2)
        return set_err(ctx, BDFT_ERR_ARG, "z potential pass: 0, 1 or 2 (got %d)", value);
      if (ctx->opt_zpot != value) {
        drop_cached_plans(ctx);
        ctx->opt_zpot = value;
      }
      return BDFT_OK;
    case BDFT_OPT_BWD:
      if (value < 0 || value > 2)
        return set_err(ctx, BDFT_ERR_ARG, "backward pass: 0, 1 or 2 (got %d)", value);
      if (ctx->opt_bwd != value) {
        drop_cached_plans(ctx);
        ctx->opt_bwd = value;
      }
      return BDFT_OK;
    case BDFT_OPT_CHUNK_MB:
      if (value < 0 || value > (1 << 20))
        return set_err(ctx, BDFT_ERR_ARG, "chunk size: 0 .. 2^20 MiB (got %d)", value);
      if (ctx->opt_chunk_mb != value) {
        drop_cached_plans(ctx);
        ctx->opt_chunk_mb = value;
      }
      return BDFT_OK;
    case BDFT_OPT_TILE_ORDER:
      if (value != 0 && value != 1)
        return set_err(ctx, BDFT_ERR_ARG, "tile order: 0 or 1 (got %d)", value);
      if (ctx->opt_order != value) {
        drop_cached_plans(ctx);
        ctx->opt_order = value;
      }
      return BDFT_OK;
    case BDFT_OPT_PERSIST:
      if (value != 0 && value != 1)
        return set_err(ctx, BDFT_ERR_ARG, "persistent z pass: 0 or 1 (got %d)", value);
      if (ctx->opt_persist != value) {
        drop_cached_plans(ctx);
        ctx->opt_persist = value;
      }
      return BDFT_OK;
    case BDFT_OPT_ZGROUP:
      if (value != 0 && value != 2 && value != 3)
        return set_err(ctx, BDFT_ERR_ARG, "z-line bundles per work-group: 0, 2 or 3 (got %d)",
                       value);
      if (ctx->opt_zg != value) {
        drop_cached_plans(ctx);
        ctx->opt_zg = value;
      }
      return BDFT_OK;
  }
  return set_err(ctx, BDFT_ERR_ARG, "unknown option %d", option);
}

extern "C" int bdft_ctx_get_option(const bdft_ctx* ctx, int option, int* value) {
  if (!ctx || !value) return BDFT_ERR_ARG;
  switch (option) {
    case BDFT_OPT_LINE_PASSES: *value = ctx->force_lines ? 1 : 0; return BDFT_OK;
    case BDFT_OPT_ZGROUP: *value = ctx->opt_zg; return BDFT_OK;
    case BDFT_OPT_ZPOT: *value = ctx->opt_zpot; return BDFT_OK;
    case BDFT_OPT_BWD: *value = ctx->opt_bwd; return BDFT_OK;
    case BDFT_OPT_CHUNK_MB: *value = ctx->opt_chunk_mb; return BDFT_OK;
    case BDFT_OPT_TILE_ORDER: *value = ctx->opt_order; return BDFT_OK;
    case BDFT_OPT_PERSIST: *value = ctx->opt_persist; return BDFT_OK;
  }
  return BDFT_ERR_ARG;
}

extern "C" int bdft_ctx_info(const bdft_ctx* ctx, int64_t* lds_per_block, int* plane_enabled) {
  if (!ctx) return BDFT_ERR_ARG;
  if (lds_per_block) *lds_per_block = (int64_t)ctx->plane_lds;
  if (plane_enabled) *plane_enabled = ctx->force_lines ? 0 : 1;
  return BDFT_OK;
}

extern "C" const char* bdft_stage_name(int stage) {
  return (stage >= 0 && stage < S_COUNT) ? kStageNames[stage] : "";
}
