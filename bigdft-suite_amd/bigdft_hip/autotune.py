"""Per-shape choice of the launch configuration of the batched Hamiltonian.

The reference's OpenCL layer launches fixed work-group shapes (e.g.
kinetic_generic, bigdft/src/OpenCL/Kinetic.c:18-28: 32 x min(16,
MAX_WORK_GROUP_SIZE/32) per group).  Here the kernel sequence and the z-line
bundling are options of the context (BDFT_OPT_LINE_PASSES, BDFT_OPT_ZGROUP);
this module times the candidates on the caller's own batch and keeps the
fastest per shape key in a JSON cache:

    key = "<BC>:<n1>x<n2>x<n3>:<precision>"      every box of the batch alike
          "<BC>:max<n1>x<n2>x<n3>:mixed:<precision>"   otherwise (the largest
          extents, which set the LDS sizes the choices depend on)

Candidates: the plane-fused and the line/per-pass sequences, and for the
matrix-core precisions in the plane-fused sequence 2 or 3 z-line bundles per
work-group, and the z potential pass and the backward plane pass each on the
matrix cores or on the FP32 vector units.  Timing runs
replay a captured plan into scratch outputs, so the caller's hpsi, ekin and
epot are not touched.  The cache file is plain JSON ({key: {"options": {...},
"ms": {candidate: ms}}}); its default location is
$BDFT_TUNE_CACHE or ~/.cache/bigdft_hip/tune.json.
"""
import json
import os

from . import PRECISIONS, BigDFTError

ERR_UNSUPPORTED = 5   # BDFT_ERR_UNSUPPORTED: the option set does not fit this batch

PREC_NAMES = {v: k for k, v in PRECISIONS.items()}
MATRIX_CORE = ("bf16x3", "bf16x4", "f16x3")


def default_path():
    return os.environ.get("BDFT_TUNE_CACHE") or os.path.join(
        os.path.expanduser("~"), ".cache", "bigdft_hip", "tune.json")


def shape_key(regions, precision):
    """cache key of a batch (regions: descriptors with geocode and n)"""
    prec = PREC_NAMES.get(precision, str(precision))
    geos = {r.geocode for r in regions}
    geo = geos.pop() if len(geos) == 1 else "mixed"
    shapes = {tuple(int(x) for x in r.n) for r in regions}
    if len(shapes) == 1:
        n = shapes.pop()
        return f"{geo}:{n[0]}x{n[1]}x{n[2]}:{prec}"
    m = [max(s[d] for s in shapes) for d in range(3)]
    return f"{geo}:max{m[0]}x{m[1]}x{m[2]}:mixed:{prec}"


def candidates(precision):
    """option sets tried for a precision, the automatic choice first (the
    bundling and the z potential unit only matter in the plane-fused
    matrix-core sequence)"""
    auto = {"line_passes": 0, "zgroup": 0, "zpot": 0, "bwd": 0}
    lines = dict(auto, line_passes=1)
    if PREC_NAMES.get(precision) not in MATRIX_CORE:
        return [auto, lines]
    fused = [dict(auto, zgroup=z, zpot=p, bwd=w) for p in (0, 1) for z in (0, 2, 3)
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


def tune(ctx, lrs, pot, psi, precision, cache=None, timer=time_plan, **batch_kw):
    """set the context's options for this batch: from the cache, or by
    timing every candidate (and caching the fastest).  Returns (key,
    options, times or None when the cache answered)."""
    cache = cache if cache is not None else TuneCache()
    key = shape_key([l.region for l in lrs], precision)
    hit = cache.get(key)
    if hit is not None:
        apply(ctx, hit)
        return key, hit, None
    times = {}
    best = None
    for opts in candidates(precision):
        apply(ctx, opts)
        try:
            ms = timer(ctx, lrs, pot, psi, precision, batch_kw)
        except BigDFTError as e:
            if e.code != ERR_UNSUPPORTED:
                raise
            times[cand_name(opts)] = None   # does not fit this batch
            continue
        times[cand_name(opts)] = ms
        if best is None or ms < best[0]:
            best = (ms, opts)
    if best is None:
        raise BigDFTError(ERR_UNSUPPORTED, f"no launch option set fits the batch {key}")
    apply(ctx, best[1])
    cache.put(key, best[1], times)
    return key, best[1], times
