"""Per-shape choice of the launch configuration of the batched Hamiltonian.

The reference's OpenCL layer launches fixed work-group shapes (e.g.
kinetic_generic, bigdft/src/OpenCL/Kinetic.c:18-28: 32 x min(16,
MAX_WORK_GROUP_SIZE/32) per group).  Here the kernel sequence and the z-line
bundling are options of the context (BDFT_OPT_LINE_PASSES, BDFT_OPT_ZGROUP);
this module times the candidates on the caller's own batch and keeps the
fastest per shape key in a JSON cache:

    key = "<device>|<BC>:<n1>x<n2>x<n3>:<precision>"   every box of the batch alike
          "<device>|<BC>:max<n1>x<n2>x<n3>:mixed<digest>:<precision>"   otherwise
          (the largest extents, which set the LDS sizes the choices depend on,
          and a digest of the multiset of per-box (BC, extents), so batches
          that differ in their boxes never share an entry)
    <device> = device name, LDS bytes per work-group and the option-table
    version (TUNE_VERSION), so a table never answers for another GPU or
    for a library whose options mean something else.

Candidates: the plane-fused and the line/per-pass sequences, and for the
matrix-core precisions in the plane-fused sequence 2 or 3 z-line bundles per
work-group, and the z potential pass and the backward plane pass each on the
matrix cores or on the FP32 vector units; then, with the fastest of those,
the orbitals in chunks of CHUNK_MB MiB of working grids
(BDFT_OPT_CHUNK_MB), then the plane and (un)compress tiles in one contiguous
range per XCD (BDFT_OPT_TILE_ORDER), then the z potential pass as a
persistent launch over a work queue (BDFT_OPT_PERSIST).  Timing runs
replay a captured plan into scratch outputs, so the caller's hpsi, ekin and
epot are not touched.  A cached choice is checked on the batch before it is
applied (a plan is made and dropped); one the batch does not fit
(BDFT_ERR_UNSUPPORTED) is removed from the table and the batch is tuned
again.  The cache file is plain JSON ({key: {"options": {...},
"ms": {candidate: ms}}}); its default location is
$BDFT_TUNE_CACHE or ~/.cache/bigdft_hip/tune.json.
"""
import hashlib
import json
import os

from . import PRECISIONS, BigDFTError

# version of the option table: bumped when an option's meaning changes, or
# when kernel changes move which options win (a cached choice is then re-timed)
TUNE_VERSION = 7

ERR_UNSUPPORTED = 5   # BDFT_ERR_UNSUPPORTED: the option set does not fit this batch

PREC_NAMES = {v: k for k, v in PRECISIONS.items()}
MATRIX_CORE = ("bf16x3", "bf16x4", "f16x3")


def default_path():
    return os.environ.get("BDFT_TUNE_CACHE") or os.path.join(
        os.path.expanduser("~"), ".cache", "bigdft_hip", "tune.json")


def device_tag(ctx):
    """device name, LDS bytes per work-group and TUNE_VERSION of a context
    ("" for stand-ins without a device)"""
    if not hasattr(ctx, "info") or not hasattr(ctx, "device"):
        return ""
    import torch
    lds, _ = ctx.info()
    name = torch.cuda.get_device_name(ctx.device).replace("|", "/")
    return f"{name}/lds{lds}/v{TUNE_VERSION}"


def shape_key(regions, precision, device=""):
    """cache key of a batch (regions: descriptors with geocode and n)"""
    prec = PREC_NAMES.get(precision, str(precision))
    geos = {r.geocode for r in regions}
    geo = geos.pop() if len(geos) == 1 else "mixed"
    boxes = sorted((r.geocode,) + tuple(int(x) for x in r.n) for r in regions)
    shapes = {b[1:] for b in boxes}
    pre = f"{device}|" if device else ""
    if len(shapes) == 1:
        n = shapes.pop()
        return f"{pre}{geo}:{n[0]}x{n[1]}x{n[2]}:{prec}"
    m = [max(s[d] for s in shapes) for d in range(3)]
    dig = hashlib.sha1(repr(boxes).encode()).hexdigest()[:12]
    return f"{pre}{geo}:max{m[0]}x{m[1]}x{m[2]}:mixed{dig}:{prec}"


# workspace per chunk of orbitals (MiB) tried after the kernel options
# (BDFT_OPT_CHUNK_MB; 0, one chunk, is among the candidates)
CHUNK_MB = (64, 128, 192)
# candidates within REFINE (relative) of the fastest are timed REFINE_ROUNDS
# more times before the choice
REFINE = 0.04
REFINE_ROUNDS = 3


def candidates(precision):
    """option sets tried for a precision, the automatic choice first (the
    bundling and the z potential unit only matter in the plane-fused
    matrix-core sequence)"""
    auto = {"line_passes": 0, "zgroup": 0, "zpot": 0, "bwd": 0, "chunk_mb": 0,
            "tile_order": 0, "persist": 0}
    lines = dict(auto, line_passes=1)
    if PREC_NAMES.get(precision) not in MATRIX_CORE:
        return [auto, lines]
    # F16X3 keeps its z potential pass in FP32 whatever zpot says
    zpots = (0,) if PREC_NAMES.get(precision) == "f16x3" else (0, 1)
    fused = [dict(auto, zgroup=z, zpot=p, bwd=w) for p in zpots for z in (0, 2, 3)
             for w in (1, 2)]
    return [auto] + fused + [lines]


def cand_name(opts):
    return ",".join(f"{k}={opts[k]}" for k in sorted(opts))


class TuneCache:
    """the JSON table (read once, written after each new entry)"""

    def __init__(self, path=None):
        self.path = path or default_path()
        self.table = {}
        if os.path.exists(self.path):
            with open(self.path) as f:
                data = json.load(f)
            if not isinstance(data, dict):
                raise ValueError(f"{self.path}: not a tuning table")
            self.table = data

    def get(self, key):
        e = self.table.get(key)
        return None if e is None else dict(e["options"])

    def put(self, key, options, times):
        self.table[key] = {"options": dict(options), "ms": dict(times)}
        self._write()

    def drop(self, key):
        if self.table.pop(key, None) is not None:
            self._write()

    def _write(self):
        d = os.path.dirname(self.path)
        if d:
            os.makedirs(d, exist_ok=True)
        tmp = self.path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(self.table, f, indent=1, sort_keys=True)
        os.replace(tmp, self.path)


def apply(ctx, options):
    for k, v in options.items():
        ctx.set_option(k, v)


def time_plan(ctx, lrs, pot, psi, precision, batch_kw, steps=3, warmup=1):
    """ms per replay of a plan made with the context's current options, on
    scratch outputs"""
    import torch
    norb = len(lrs)
    h = torch.zeros_like(psi)
    ek = torch.zeros(norb, dtype=torch.float64, device=psi.device)
    ep = torch.zeros_like(ek)
    plan = ctx.plan(lrs, pot, psi, h, ek, ep, precision=precision, **batch_kw)
    try:
        for _ in range(warmup):
            plan.execute()
        s = torch.cuda.current_stream()
        t0 = torch.cuda.Event(enable_timing=True)
        t1 = torch.cuda.Event(enable_timing=True)
        t0.record(s)
        for _ in range(steps):
            plan.execute()
        t1.record(s)
        t1.synchronize()
        return t0.elapsed_time(t1) / steps
    finally:
        plan.close()


def check_plan(ctx, lrs, pot, psi, precision, batch_kw):
    """make and drop a plan with the context's current options (raises
    BDFT_ERR_UNSUPPORTED when they do not fit the batch); stand-ins without
    plans pass"""
    if not hasattr(ctx, "plan"):
        return
    import torch
    norb = len(lrs)
    h = torch.zeros_like(psi)
    ek = torch.zeros(norb, dtype=torch.float64, device=psi.device)
    ctx.plan(lrs, pot, psi, h, ek, ek.clone(), precision=precision, **batch_kw).close()


def tune(ctx, lrs, pot, psi, precision, cache=None, timer=time_plan, check=check_plan,
         **batch_kw):
    """set the context's options for this batch: from the cache (after
    checking that the cached choice fits the batch), or by timing every
    candidate (and caching the fastest).  Returns (key, options, times or
    None when the cache answered)."""
    cache = cache if cache is not None else TuneCache()
    key = shape_key([l.region for l in lrs], precision, device_tag(ctx))
    hit = cache.get(key)
    if hit is not None:
        apply(ctx, hit)
        try:
            if check is not None:
                check(ctx, lrs, pot, psi, precision, batch_kw)
            return key, hit, None
        except BigDFTError as e:
            if e.code != ERR_UNSUPPORTED:
                raise
            cache.drop(key)   # stale entry: tune again
    times = {}
    seen = {}
    best = None

    def trial(opts):
        nonlocal best
        apply(ctx, opts)
        try:
            ms = timer(ctx, lrs, pot, psi, precision, batch_kw)
        except BigDFTError as e:
            if e.code != ERR_UNSUPPORTED:
                raise
            times[cand_name(opts)] = None   # does not fit this batch
            return
        times[cand_name(opts)] = ms
        seen[cand_name(opts)] = opts
        if best is None or ms < best[0]:
            best = (ms, opts)
    for opts in candidates(precision):
        trial(opts)
    if best is None:
        raise BigDFTError(ERR_UNSUPPORTED, f"no launch option set fits the batch {key}")
    # then the orbital chunking, with the kernel options chosen above
    base = best[1]
    for mb in CHUNK_MB:
        trial(dict(base, chunk_mb=mb))
    # then the tile order of the plane and (un)compress passes, then the
    # persistent z potential pass
    trial(dict(best[1], tile_order=1))
    trial(dict(best[1], persist=1))
    # one timing carries a few per cent of noise: the candidates within
    # REFINE of the fastest are timed REFINE_ROUNDS more times, in turn, and
    # the mean of all their timings decides
    close = [n for n, ms in times.items() if ms is not None and ms <= (1 + REFINE) * best[0]]
    if len(close) > 1:
        runs = {n: [times[n]] for n in close}
        for _ in range(REFINE_ROUNDS):
            for n in close:
                apply(ctx, seen[n])
                runs[n].append(timer(ctx, lrs, pot, psi, precision, batch_kw))
        for n in close:
            times[n] = sum(runs[n]) / len(runs[n])
        n = min(close, key=lambda c: times[c])
        best = (times[n], seen[n])
    apply(ctx, best[1])
    cache.put(key, best[1], times)
    return key, best[1], times
