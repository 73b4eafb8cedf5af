#!/usr/bin/env python3
"""Benchmark of the batched local-Hamiltonian apply on one MI355X.

Headline workload (default, the largest single-GPU one): 2048 free-BC
localisation boxes with N_d ~ U{16..40} coarse points per axis, hgrid 0.4
bohr, coefficients N(0,1) times exp(-r^2/32), each box's potential sampled
from one 8-Gaussian field (1.26e9 grid points per apply), in the split-FP16
matrix-core precision (F16X3) next to FP64, FP32 and BF16X4.  The same JSON
line carries the other workloads (--also): the periodic 48^3 x 256 and
64^3 x 128 batches (8-Gaussian potential, minimum-image distance) and 1024
compressed spheres in a 96^3 global box, each with its timings, errors
against FP64, roofline entry and CPU baseline.

One step = one full application hpsi += H_loc psi (+ ekin, epot per
orbital) over the batch, replayed from a captured hipGraph.  Inputs are
resident in HBM before the timed region.  Before each precision is timed,
the launch options (kernel sequence, z bundling, units of the mixed
sequences' z potential and backward passes) are chosen by timing the
candidates on the same batch (bigdft_hip.autotune; --no-autotune keeps the
automatic choice).  Errors are per orbital against the FP64 GPU run, with
no floor; FP64 is checked against the oracle on a seeded sample, which is
also the timed CPU baseline (oracle on the host cores, bounded by
--cpu-seconds per workload).  Progress lines go to stderr.

Other workloads (--workload): precond (the conjugate-gradient
preconditioner), density (rho += occ psir^2) and transfers (global <->
localisation-region copies), each printing one JSON line of its own.

Prints one JSON line.  value = grid-point*orbital applications per second
(grid point = fine ISF point, n1i*n2i*n3i per orbital, the reference's unit,
bigdft/src/memguess.f90:2269).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "bigdft-suite_amd"))

METRIC = ("grid-pt·orbital/s of local-H apply: mixed vs FP64 at "
          "≤1e-6 rel err")
PEAK_FP64_TFLOPS = 78.6     # MI355X dense FP64 (vector and matrix)
PEAK_FP32_TFLOPS = 157.3    # MI355X dense FP32
PEAK_BF16_TFLOPS = 2500.0   # MI355X dense BF16 matrix
PEAK_HBM_GBS = 8000.0

# Algorithmic work of each timed stage.  Flops follow the reference's count
# per fine grid point, 524 = 192 (synthesis+magic) + 46*3 (kinetic) + 192
# (magic^T+analysis) + 2 (bigdft/src/memguess.f90:2269): one 1-D grow or
# shrink pass is 64, one kinetic pass 46, V*psi 2.
STAGE_FLOPS = {"uncompress": 0, "zero": 0, "xy_fwd": 220, "z_kin": 46, "z_pot": 130,
               "xy_bwd_compress": 128, "x_grow_kin": 110, "y_grow": 64, "y_kin": 46,
               "y_shrink": 64, "x_shrink": 64, "compress": 0}
# per-pass matrix-core sequence: the x kinetic alone, then one plane grow
STAGE_FLOPS_MFMA_PASSES = dict(STAGE_FLOPS, x_grow_kin=46, xy_fwd=128)


def stage_bytes(batch, esz):
    """compulsory HBM bytes per apply of each stage, from the batch's box
    geometry: working grids of esz bytes -- W (wavelet volume, 8 N1 N2 N3
    per box), K (same), T2 (L1 L2 2 N3), T1 (L1 2 N2 2 N3, line passes
    only) -- FP64 psi read once (8 B per coefficient), hpsi read and written
    once (16 B), each potential value read once (8 B)"""
    regs = batch["regions"]
    W = sum(8 * (r.n[0] + 1) * (r.n[1] + 1) * (r.n[2] + 1) for r in regs)
    T2 = sum(r.nfine[0] * r.nfine[1] * 2 * (r.n[2] + 1) for r in regs)
    T1 = sum(r.nfine[0] * 4 * (r.n[1] + 1) * (r.n[2] + 1) for r in regs)
    nc = int(batch["ncoef"].sum())
    V = 8 * min(int(batch["pot"].numel()), int(sum(r.n_isf for r in regs)))
    return {"zero": esz * W, "uncompress": 8 * nc + esz * W,
            "xy_fwd": esz * (2 * W + T2), "z_kin": esz * 3 * W, "z_pot": esz * 2 * T2 + V,
            "xy_bwd_compress": esz * (T2 + W) + 16 * nc,
            "x_grow_kin": esz * (2 * W + T1), "y_grow": esz * (T1 + T2),
            "y_kin": esz * 3 * W, "y_shrink": esz * (T2 + T1),
            "x_shrink": esz * (T1 + 2 * W), "compress": esz * W + 16 * nc}


def path_of(prec, plane_enabled, stages):
    """kernel sequence that ran, from the precision and the stages timed"""
    if prec in ("bf16x3", "bf16x4", "f16x3"):
        return "mfma_passes" if stages.get("x_grow_kin", 0.0) > 0.02 else "mfma"
    return "plane" if plane_enabled else "lines"


def measured_traffic(precision, config, stage):
    """HBM bytes per apply of a stage from the committed rocprofv3 counter
    summaries (the ```json blocks of profiles/*.md: FETCH_SIZE + WRITE_SIZE,
    calibrated on a device copy of known size), or None when none covers
    this precision and workload"""
    pdir = os.path.join(ROOT, "profiles")
    for f in sorted(os.listdir(pdir)) if os.path.isdir(pdir) else []:
        if not f.endswith(".md"):
            continue
        text = open(os.path.join(pdir, f)).read()
        for blob in text.split("```json")[1:]:
            try:
                data = json.loads(blob.split("```", 1)[0])
            except ValueError:
                continue
            if data.get("config") == config and data.get("precision") == precision:
                v = data["bytes_per_launch"].get(stage)
                if v is not None:
                    return v, f"profiles/{f}"
    return None, None


def roofline(path, avg, units, bytes_of, peak_c):
    """roofline entry of the dominant (longest) kernel stage: its
    algorithmic bytes (bytes_of, from the box geometry) and flops (the
    reference's per-point count) over its measured duration"""
    fl_tab = STAGE_FLOPS_MFMA_PASSES if path == "mfma_passes" else STAGE_FLOPS
    dom = max((k for k in avg if k in bytes_of and avg[k] > 0.02), key=lambda k: avg[k])
    fl = fl_tab[dom]
    t_dom = avg[dom] * 1e-3
    ach_c = fl * units / t_dom / 1e12
    ach_b = bytes_of[dom] / t_dom / 1e9
    if ach_c / peak_c >= ach_b / PEAK_HBM_GBS:
        roof = {"bound": "mfma", "achieved": ach_c, "peak": peak_c,
                "unit": "TFLOP/s", "frac": ach_c / peak_c,
                "hbm_GBps": ach_b, "hbm_frac": ach_b / PEAK_HBM_GBS}
    else:
        roof = {"bound": "hbm", "achieved": ach_b, "peak": PEAK_HBM_GBS,
                "unit": "GB/s", "frac": ach_b / PEAK_HBM_GBS,
                "flop_TFLOPs": ach_c, "flop_frac": ach_c / peak_c}
    roof.update({"traffic": None, "kernel": dom, "kernel_ms": avg[dom],
                 "path": path, "flops_per_pt": fl,
                 "algorithmic_bytes": bytes_of[dom],
                 "bytes_per_pt": bytes_of[dom] / units})
    return roof


CONFIGS = {
    "c2": dict(workload="periodic 48^3 coarse (96^3 fine), 256 orbitals, "
               "hgrid 0.35, 8-Gaussian potential", kind="dense", geocode="P", n=47,
               hgrid=0.35, norb=256),
    "c3": dict(workload="periodic 64^3 coarse (128^3 fine), 128 orbitals, "
               "hgrid 0.35, 8-Gaussian potential", kind="dense", geocode="P", n=63,
               hgrid=0.35, norb=128),
    "c4": dict(workload="2048 free boxes, N_d ~ U{16..40} coarse points per axis, "
               "hgrid 0.4, dense keys, envelope exp(-r^2/32), potential per box "
               "from one 8-Gaussian field", kind="uneven", norb=2048),
    "c5": dict(workload="1024 compressed free orbitals in a 96^3 global box, "
               "hgrid 0.4, coarse r<=7, fine r<=3.5 bohr, one global 8-Gaussian "
               "potential read through ishift", kind="spheres", norb=1024),
}


def build_batch(cfg, seed, device):
    """regions, device arrays and offsets of a configuration"""
    from bigdft_hip import workloads
    if cfg["kind"] == "uneven":
        return workloads.uneven_free(cfg["norb"], seed, device=device)
    if cfg["kind"] == "spheres":
        return workloads.sphere_batch(cfg["norb"], seed, device=device)
    return workloads.dense_periodic(cfg["n"], cfg["norb"], seed, hgrid=cfg["hgrid"],
                                    device=device)


# Algorithmic FP64 bytes per coefficient and CG step of the preconditioner
# (dense keys: one working-grid point per coefficient): d = r + beta d and
# W = S d (32), K = c W + K_x W (16), K += K_y W, K_z W (24 + 24), q = S K and
# q.d (24), x += alpha d, r -= alpha q, |r|^2 (48).
PC_BYTES_PER_COEF_STEP = 168


def bench_precond(args, cfg, batch, ctx, lrs, stream):
    """one step = one preconditioner call over the batch: the initial
    residual, ncong CG steps and the final scaling"""
    import numpy as np
    import torch
    norb = len(lrs)
    b = batch["psi"].clone()
    x = 0.1 * batch["psi"]
    cpr = np.full(norb, 0.5)
    ncoef = int(batch["ncoef"].sum())
    for _ in range(args.warmup):
        ctx.precond_cg(lrs, cpr, args.ncong, x, b, psi_off=batch["psi_off"])
    torch.cuda.synchronize()
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    w0 = time.perf_counter()
    t0.record(stream)
    for _ in range(args.steps):
        ctx.precond_cg(lrs, cpr, args.ncong, x, b, psi_off=batch["psi_off"])
    t1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    ms = t0.elapsed_time(t1) / args.steps
    applies = args.ncong + 1
    units = batch["units"]
    gbs = PC_BYTES_PER_COEF_STEP * ncoef * applies / (ms * 1e-3) / 1e9
    roof = {"bound": "hbm", "achieved": gbs, "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": gbs / PEAK_HBM_GBS, "traffic": None,
            "kernel": "whole call (scatter, kinetic x/y/z, gather, update per CG step)",
            "bytes_per_coef_step": PC_BYTES_PER_COEF_STEP}
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        regions = batch["regions"]
        r = regions[0]
        keys = oracle.Keys(r.nseg_c, r.nvctr_c, r.keyg_c, r.keyv_c, r.nseg_f, r.nvctr_f,
                           r.keyg_f, r.keyv_f)
        nc = int(batch["ncoef"][0])
        xb = x[:nc].cpu().numpy()
        bb = b[:nc].cpu().numpy()
        done, ts = 0, 0.0
        while done < norb and (done < 1 or ts < args.cpu_seconds):
            tt = time.perf_counter()
            oracle.precond_cg(r.geocode, r.n, r.hgrid, 0.5, args.ncong, keys, xb, bb)
            ts += time.perf_counter() - tt
            done += 1
        cpu = {"value": r.n_isf * done * applies / ts,
               "unit": "grid-pt*orbital*CG-step/s", "cores": oracle.num_threads(),
               "kind": "port", "sample": f"{done} orbital(s) of the batch, ncong "
               f"{args.ncong}, oracle (C+OpenMP restatement of the OpenCL "
               f"preconditioner), {ts:.1f} s"}
    out = {
        "metric": "grid-pt*orbital*CG-step/s of the preconditioner "
                  "(ocl_preconditioner_generic_k), FP64",
        "value": units * applies / (ms * 1e-3), "unit": "grid-pt*orbital*CG-step/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded N(0,1) right-hand sides, guess 0.1 b)",
        "config": {"workload": "preconditioner on " + cfg["workload"], "orbitals": norb,
                   "ncong": args.ncong, "cprecr": 0.5, "wall_s": wall},
        "roofline": roof, "cpu_baseline": cpu,
    }
    print(json.dumps(out))


def timed_calls(fn, steps, warmup, stream):
    """ms per call of fn over steps calls after warmup, with events on the
    library's stream, and the wall time"""
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    w0 = time.perf_counter()
    t0.record(stream)
    for _ in range(steps):
        fn()
    t1.record(stream)
    torch.cuda.synchronize()
    return t0.elapsed_time(t1) / steps, time.perf_counter() - w0


def oracle_keys(oracle, reg):
    return oracle.Keys(reg.nseg_c, reg.nvctr_c, reg.keyg_c, reg.keyv_c, reg.nseg_f,
                       reg.nvctr_f, reg.keyg_f, reg.keyv_f)


def bench_density(args, cfg, batch, ctx, lrs, stream):
    """one step = rho += hfac psir^2 over every orbital of the batch (one
    region shared by all orbitals: the dense configurations)"""
    import numpy as np
    import torch
    import bigdft_hip
    if len({id(r) for r in batch["regions"]}) != 1:
        raise SystemExit("the density workload needs one shared region (c2 or c3)")
    norb = len(lrs)
    reg = batch["regions"][0]
    prec = bigdft_hip.PREC_FP32 if args.precision == "fp32" else bigdft_hip.PREC_FP64
    rho = torch.zeros((1, reg.n_isf), dtype=torch.float64, device="cuda")
    hf = np.full(norb, 2.0)
    psi = batch["psi"]
    ms, wall = timed_calls(lambda: ctx.density(lrs[0], psi, hf, rho, precision=prec),
                           args.steps, args.warmup, stream)
    units = batch["units"]
    # algorithmic cost per grid point and orbital: psi read once (8 B: one
    # coefficient per fine point on a dense box) and rho read and written
    # once per call (16 B over the orbitals); three grow passes (synthesis +
    # magic filter, 64 flops each) and hfac*psir^2 (2)
    esz = 8 if prec == bigdft_hip.PREC_FP64 else 4
    bpp = 8 + 16 / norb
    fpp = 3 * 64 + 2
    t = ms * 1e-3
    gbs, tfs = bpp * units / t / 1e9, fpp * units / t / 1e12
    peak_c = 78.6 if esz == 8 else 157.3
    if tfs / peak_c >= gbs / PEAK_HBM_GBS:
        roof = {"bound": "mfma", "achieved": tfs, "peak": peak_c,
                "unit": "TFLOP/s", "frac": tfs / peak_c, "hbm_GBps": gbs}
    else:
        roof = {"bound": "hbm", "achieved": gbs, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": gbs / PEAK_HBM_GBS, "flop_TFLOPs": tfs}
    roof.update({"traffic": None, "bytes_per_pt": bpp, "flops_per_pt": fpp,
                 "kernel": "whole call (uncompress, x and y grow, z grow + hfac psir^2, "
                           "group reduction)"})
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        keys = oracle_keys(oracle, reg)
        nc = int(batch["ncoef"][0])
        done, ts = 0, 0.0
        while done < norb and (done < 1 or ts < args.cpu_seconds):
            p = psi[done * nc:(done + 1) * nc].cpu().numpy()[None, :]
            tt = time.perf_counter()
            oracle.density(reg.geocode, reg.n, reg.nfl, reg.nfu, keys, p, [2.0])
            ts += time.perf_counter() - tt
            done += 1
        cpu = {"value": reg.n_isf * done / ts, "unit": "grid-pt*orbital/s",
               "cores": oracle.num_threads(), "kind": "port",
               "sample": f"{done} orbital(s), oracle density, {ts:.1f} s"}
    print(json.dumps({
        "metric": "grid-pt*orbital/s of the density accumulation (local_partial_density)",
        "value": units / (ms * 1e-3), "unit": "grid-pt*orbital/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if esz == 8 else "f32", "data": "synthetic (seeded N(0,1) coefficients)",
        "config": {"workload": "density on " + cfg["workload"], "orbitals": norb, "wall_s": wall},
        "roofline": roof,
        "cpu_baseline": cpu}))


def bench_transfers(args):
    """one step = psi_to_locreg2 of every localisation region of the
    compressed-sphere batch out of its own global orbital, then
    Lpsi_to_global2 back (bit-exact copies, no arithmetic)"""
    import numpy as np
    import torch
    import bigdft_hip
    from bigdft_hip import workloads
    from bigdft_hip.descriptors import dense_region
    norb = 256
    sb = workloads.sphere_batch(norb, args.seed, device="cuda")
    ng = (95, 95, 95)
    glob = dense_region("F", ng, (0.4, 0.4, 0.4))
    ctx = bigdft_hip.Context(0)
    stream = torch.cuda.Stream()
    ctx.set_stream(stream)
    torch.cuda.synchronize()
    torch.cuda.set_stream(stream)   # one stream for torch and the library
    g = ctx.region(glob)
    cache = {}
    lrs = [cache.setdefault(id(r), ctx.region(r)) for r in sb["regions"]]
    ns = np.array(sb["nl"], dtype=np.int32).reshape(-1, 3)
    gnc = glob.ncoef
    gpsi = torch.randn(norb * gnc, dtype=torch.float64, device="cuda")
    goff = np.arange(norb, dtype=np.int64) * gnc
    lpsi = torch.empty_like(sb["psi"])
    loff = sb["psi_off"]
    ms_g, _ = timed_calls(lambda: ctx.psi_to_locreg(g, lrs, ns, gpsi, lpsi, goff=goff, loff=loff),
                          args.steps, args.warmup, stream)
    ms_l, _ = timed_calls(lambda: ctx.locreg_to_psi(g, lrs, ns, lpsi, gpsi, loff=loff, goff=goff),
                          args.steps, args.warmup, stream)
    nloc = int(sb["ncoef"].sum())
    npts = int(sum(r.nvctr_c + r.nvctr_f for r in sb["regions"]))
    # bytes per direction: every local coefficient read once and written
    # once (16 B), plus per point the local map and the global inverse-map
    # entry (8 B; the dense global region needs no inverse map: 4 B)
    byts = 16 * nloc + 4 * npts
    ms = ms_g + ms_l
    gbs = 2 * byts / (ms * 1e-3) / 1e9
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        gk = oracle_keys(oracle, glob)
        done, ts, moved = 0, 0.0, 0
        while done < norb and (done < 1 or ts < args.cpu_seconds):
            r = sb["regions"][done]
            gp = gpsi[done * gnc:(done + 1) * gnc].cpu().numpy()
            lk = oracle_keys(oracle, r)
            tt = time.perf_counter()
            lp = oracle.psi_to_locreg2(r.n, ng, sb["nl"][done], lk, gk, gp)
            oracle.lpsi_to_global2(r.n, ng, sb["nl"][done], lk, gk, lp, gp)
            ts += time.perf_counter() - tt
            moved += 2 * r.ncoef
            done += 1
        cpu = {"value": moved / ts, "unit": "coefficient/s", "cores": 1, "kind": "port",
               "sample": f"{done} orbital(s) there and back, oracle psi_to_locreg2 + "
                         f"lpsi_to_global2, {ts:.1f} s"}
    print(json.dumps({
        "metric": "compressed coefficients/s moved between global and localisation "
                  "regions (psi_to_locreg2 + Lpsi_to_global2)",
        "value": 2 * nloc / (ms * 1e-3), "unit": "coefficient/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "ms_to_local": ms_g, "ms_to_global": ms_l, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded spheres, N(0,1) global orbitals)",
        "config": {"workload": f"{norb} compressed spheres (coarse r 7, fine r 3.5 bohr) of a "
                               "96^3 free global box, one global orbital each",
                   "local_coefficients": nloc},
        "roofline": {"bound": "hbm", "achieved": gbs, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": gbs / PEAK_HBM_GBS, "traffic": None, "kernel": "k_xfer",
                     "bytes_per_call": byts},
        "cpu_baseline": cpu}))
    for l in cache.values():
        l.close()
    g.close()
    ctx.close()


def cpu_model():
    """model name of the host CPU (first /proc/cpuinfo entry)"""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="c4", choices=sorted(CONFIGS),
                    help="workload of the headline value (default: the largest "
                         "single-GPU workload, 2048 uneven free boxes)")
    ap.add_argument("--also", default="c2,c3,c5",
                    help="other workloads carried in the same JSON line (comma "
                         "separated, '' for none)")
    ap.add_argument("--precision", default="f16x3",
                    choices=["fp64", "fp32", "bf16x3", "bf16x4", "f16x3"],
                    help="arithmetic of the headline run (fp32: FP32 working "
                         "grids; bf16x4: split-BF16 matrix-core path)")
    ap.add_argument("--variants", default="fp64,fp32,bf16x4,f16x3",
                    help="precisions timed beside the headline")
    ap.add_argument("--also-variants", default="fp64,fp32,f16x3",
                    help="precisions timed on the other workloads")
    ap.add_argument("--cpu-seconds", type=float, default=10.0,
                    help="time budget of the CPU baseline sample of each workload")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--workload", default="locham",
                    choices=["locham", "precond", "density", "transfers"],
                    help="locham: the local Hamiltonian (the headline); precond: the "
                         "conjugate-gradient preconditioner on the same batch; density: "
                         "rho += occ psir^2 over the batch; transfers: global <-> "
                         "localisation-region copies of compressed orbitals")
    ap.add_argument("--ncong", type=int, default=6,
                    help="CG steps of the preconditioner workload")
    ap.add_argument("--no-autotune", dest="autotune", action="store_false",
                    help="keep the automatic launch options instead of timing the "
                         "candidates on the batch first (bigdft_hip.autotune)")
    ap.add_argument("--options", default="",
                    help="launch options set on the context instead of tuning, "
                         "e.g. 'chunk_mb=128,zgroup=2' (names of bigdft_hip.OPTIONS)")
    ap.add_argument("--host-build", action="store_true",
                    help="generate the synthetic batch on the host and copy it to the "
                         "device (no generator kernels on the GPU; profiler counter runs)")
    ap.add_argument("--copy-mb", type=int, default=0,
                    help="run a plain device copy of this many MiB first (the "
                         "known-size kernel that scales profiler byte counters)")
    return ap.parse_args()


DTYPES = {"fp64": "f64", "fp32": "f32", "bf16x3": "bf16", "bf16x4": "bf16", "f16x3": "f16"}
MATRIX = ("bf16x3", "bf16x4", "f16x3")


def run_config(args, name, names, headline, ctx, stream, tune_cache, progress):
    """time the precisions `names` on workload `name` (FP64 first: the
    reference of the error columns), then check FP64 against the oracle and
    time the oracle on a bounded sample (the CPU baseline)"""
    import numpy as np
    import torch

    import bigdft_hip
    cfg = CONFIGS[name]
    dev = "cuda"
    t_build = time.perf_counter()
    if args.host_build:
        batch = build_batch(cfg, args.seed, "cpu")
        for k in ("psi", "pot"):
            batch[k] = batch[k].to(dev)
    else:
        batch = build_batch(cfg, args.seed, dev)
    regions, psi = batch["regions"], batch["psi"]
    norb, units = len(regions), batch["units"]
    progress(f"{name}: {norb} orbitals, {units} grid points, built in "
             f"{time.perf_counter() - t_build:.1f} s")
    lr_of = {}
    for r in regions:
        if id(r) not in lr_of:
            lr_of[id(r)] = ctx.region(r)
    lrs = [lr_of[id(r)] for r in regions]
    bargs = dict(pot_off=batch["pot_off"], potdim=batch["potdim"],
                 ishift=batch["ishift"], psi_off=batch["psi_off"])
    chosen = {}
    tuned_ms = {}

    def timed(prec):
        """launch options (autotuned on this batch unless --no-autotune),
        graph-replay timing, then per-stage HIP-event timing, then one
        clean application for the error analysis"""
        if tune_cache is not None:
            from bigdft_hip import autotune
            _, opts, ttimes = autotune.tune(ctx, lrs, batch["pot"], psi, prec,
                                            cache=tune_cache, **bargs)
            chosen[prec] = opts
            if ttimes:
                tuned_ms[prec] = {k: (None if v is None else round(v, 4))
                                  for k, v in ttimes.items()}
        h = torch.zeros_like(psi)
        ek = torch.zeros(norb, dtype=torch.float64, device=dev)
        ep = torch.zeros_like(ek)
        plan = ctx.plan(lrs, batch["pot"], psi, h, ek, ep, precision=prec, **bargs)
        for _ in range(args.warmup):
            plan.execute()
        torch.cuda.synchronize()
        t0 = torch.cuda.Event(enable_timing=True)
        t1 = torch.cuda.Event(enable_timing=True)
        w0 = time.perf_counter()
        t0.record(stream)
        for _ in range(args.steps):
            plan.execute()
        t1.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - w0
        ms = t0.elapsed_time(t1) / args.steps
        plan.close()
        ctx.profile(True)
        ctx.profile_read(reset=True)
        for _ in range(max(3, min(args.steps, 10))):
            ctx.locham_batch(lrs, batch["pot"], psi, h, ek, ep, precision=prec, **bargs)
        stages, nl = ctx.profile_read(reset=True)
        ctx.profile(False)
        h.zero_()
        ctx.locham_batch(lrs, batch["pot"], psi, h, ek, ep, precision=prec, **bargs)
        ctx.sync()
        return ms, wall, {k: v / nl for k, v in stages.items()}, h, ek, ep

    _, plane_enabled = ctx.info()
    if "fp64" not in names:
        names = ["fp64"] + names
    ms64, wall64, st64, h64, ek64, ep64 = timed(bigdft_hip.PREC_FP64)
    progress(f"{name}: fp64 {ms64:.3f} ms per apply")
    hmax = h64.abs().max()

    def errors(hm, ekm, epm):
        d = (hm - h64).abs()
        guard = torch.maximum(h64.abs(), 1e-3 * hmax)
        rk = (ekm - ek64).abs() / ek64.abs()
        rp = (epm - ep64).abs() / ep64.abs()
        return {"hpsi_max_abs_over_max": float(d.max() / hmax),
                "hpsi_rms_over_rms": float(torch.sqrt((d * d).mean() /
                                                      (h64 * h64).mean())),
                "hpsi_max_rel_guarded": float((d / guard).max()),
                "ekin_max_rel": float(rk.max()), "epot_max_rel": float(rp.max()),
                "ekin_rms_rel": float(torch.sqrt((rk * rk).mean())),
                "epot_rms_rel": float(torch.sqrt((rp * rp).mean())),
                "orbitals_above_1e-6": int(((rk > 1e-6) | (rp > 1e-6)).sum()),
                "ekin_sum_rel": float(abs(ekm.sum() - ek64.sum()) / abs(ek64.sum())),
                "epot_sum_rel": float(abs(epm.sum() - ep64.sum()) / abs(ep64.sum())),
                "reference": "FP64 GPU run on the same inputs, per orbital, no floor"}

    variants = {}
    for vname in names:
        prec = bigdft_hip.PRECISIONS[vname]
        if vname == "fp64":
            ms, wall, avg, err = ms64, wall64, st64, None
        else:
            ms, wall, avg, hm, ekm, epm = timed(prec)
            err = errors(hm, ekm, epm)
            del hm
            progress(f"{name}: {vname} {ms:.3f} ms per apply, energies within "
                     f"{max(err['ekin_max_rel'], err['epot_max_rel']):.2e}")
        o = chosen.get(prec)
        plane = plane_enabled if o is None else o.get("line_passes", 0) == 0
        vpath = path_of(vname, plane, avg)
        esz = 8 if vname == "fp64" else 4
        vpeak = (PEAK_BF16_TFLOPS if vname in MATRIX else
                 PEAK_FP64_TFLOPS if vname == "fp64" else PEAK_FP32_TFLOPS)
        roof = roofline(vpath, avg, units, stage_bytes(batch, esz), vpeak)
        traffic, src = measured_traffic(vname, name, roof["kernel"])
        if traffic is not None:
            roof["traffic"] = traffic
            roof["traffic_bytes_per_pt"] = traffic / units
            roof["traffic_source"] = src + " (FETCH_SIZE+WRITE_SIZE per apply, copy-calibrated)"
        variants[vname] = {"value": units / (ms * 1e-3), "ms_per_step": ms,
                           "speedup_vs_fp64": ms64 / ms,
                           "launch_options": chosen.get(prec),
                           "tuning_ms": tuned_ms.get(prec), "wall_s": round(wall, 4),
                           "path": vpath,
                           "stages_ms": {k: round(v, 4) for k, v in avg.items() if v > 0.02},
                           "roofline": roof, "error_vs_fp64": err}
    # ---- parity on a sample + CPU baseline (oracle, host cores) ----
    parity, cpu = None, None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        from bigdft_hip import workloads
        keys_of, pot_of = {}, {}
        done, ts, pts, errs, eks, eps = 0, 0.0, 0, [], [], []
        ek_g, ep_g = ek64.cpu().numpy(), ep64.cpu().numpy()
        order = np.random.default_rng(args.seed + 9).permutation(norb)
        while done < norb and (done < 4 or ts < args.cpu_seconds):
            b = int(order[done])
            r = regions[b]
            if id(r) not in keys_of:
                keys_of[id(r)] = oracle.Keys(r.nseg_c, r.nvctr_c, r.keyg_c, r.keyv_c,
                                             r.nseg_f, r.nvctr_f, r.keyg_f, r.keyv_f)
            pkey = (int(batch["pot_off"][b]), None if batch["ishift"] is None else
                    tuple(batch["ishift"][3 * b:3 * b + 3]), r.nfine)
            if pkey not in pot_of:
                pot_of = {pkey: workloads.local_potential(batch, b)}
            o0, nc = int(batch["psi_off"][b]), int(batch["ncoef"][b])
            p = psi[o0:o0 + nc].cpu().numpy()[None, :]
            tt = time.perf_counter()
            ho, eo, po = oracle.locham(r.geocode, r.n, r.nfl, r.nfu, r.hgrid,
                                       keys_of[id(r)], pot_of[pkey], p)
            ts += time.perf_counter() - tt
            pts += r.n_isf
            hg = h64[o0:o0 + nc].cpu().numpy()
            sc = np.maximum(np.abs(ho[0]), 1e-3 * np.abs(ho[0]).max())
            errs.append(float((np.abs(hg - ho[0]) / sc).max()))
            eks.append(abs(ek_g[b] - eo[0]) / abs(eo[0]))
            eps.append(abs(ep_g[b] - po[0]) / abs(po[0]))
            done += 1
        parity = {"orbitals": done, "fp64_hpsi_max_rel": max(errs),
                  "fp64_ekin_max_rel": float(max(eks)),
                  "fp64_epot_max_rel": float(max(eps)), "reference": "oracle"}
        cpu = {"value": pts / ts, "unit": "grid-pt*orbital/s",
               "cores": oracle.num_threads(), "kind": "port", "cpu_model": cpu_model(),
               "sample": f"{done} of the {norb} orbitals (seeded order, {pts} grid "
                         f"points), oracle (C+OpenMP restatement of the reference "
                         f"Fortran), {ts:.1f} s"}
        progress(f"{name}: oracle parity on {done} orbitals, CPU {pts / ts:.3g} pts/s")
    head = variants[headline] if headline in variants else variants["fp64"]
    out = {"workload": cfg["workload"], "orbitals": norb, "grid_pts": units,
           "precision": headline, "value": head["value"], "ms_per_step": head["ms_per_step"],
           "speedup_vs_fp64": head["speedup_vs_fp64"], "error_vs_fp64": head["error_vs_fp64"],
           "roofline": head["roofline"], "launch_options": head["launch_options"],
           "wall_s": head["wall_s"], "variants": variants, "cpu_baseline": cpu,
           "parity": parity}
    for lr in lr_of.values():
        lr.close()
    del batch, h64, psi
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the "
              "batch shards by orbital across GPUs but no multi-GPU run is "
              "part of this build.", file=sys.stderr)
        sys.exit(2)
    if args.workload == "transfers":
        bench_transfers(args)
        return
    import torch

    import bigdft_hip

    def progress(msg):
        print(f"[bench] {msg}", file=sys.stderr, flush=True)

    if args.copy_mb > 0:
        src = torch.ones(args.copy_mb << 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    ctx = bigdft_hip.Context(0)
    stream = torch.cuda.Stream()
    ctx.set_stream(stream)
    torch.cuda.synchronize()
    torch.cuda.set_stream(stream)   # one stream for torch and the library
    if args.workload in ("precond", "density"):
        cfg = CONFIGS[args.config]
        batch = build_batch(cfg, args.seed, "cuda")
        lr_of = {}
        for r in batch["regions"]:
            if id(r) not in lr_of:
                lr_of[id(r)] = ctx.region(r)
        lrs = [lr_of[id(r)] for r in batch["regions"]]
        (bench_precond if args.workload == "precond" else bench_density)(
            args, cfg, batch, ctx, lrs, stream)
        for lr in lr_of.values():
            lr.close()
        ctx.close()
        return

    tune_cache = None
    if args.options:
        for kv in args.options.split(","):
            k, v = kv.split("=")
            ctx.set_option(k.strip(), int(v))
    elif args.autotune:
        import tempfile

        from bigdft_hip import autotune
        tune_cache = autotune.TuneCache(os.path.join(tempfile.mkdtemp(), "tune.json"))
    names = [v for v in args.variants.split(",") if v]
    if args.precision not in names:
        names.append(args.precision)
    main_cfg = run_config(args, args.config, names, args.precision, ctx, stream, tune_cache,
                          progress)
    others = {}
    also_names = [v for v in args.also_variants.split(",") if v]
    if args.precision not in also_names:
        also_names.append(args.precision)
    for name in [c for c in args.also.split(",") if c and c != args.config]:
        others[name] = run_config(args, name, also_names, args.precision, ctx, stream,
                                  tune_cache, progress)
    units = main_cfg["grid_pts"]
    ms_step = main_cfg["ms_per_step"]
    out = {
        "metric": METRIC, "value": main_cfg["value"],
        "unit": "grid-pt*orbital/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": DTYPES[args.precision],
        "data": "synthetic (seeded N(0,1) coefficients, 8-Gaussian potential)",
        "config": {"workload": main_cfg["workload"], "name": args.config,
                   "orbitals": main_cfg["orbitals"], "grid_pts": units,
                   "precision": args.precision, "graph": True,
                   "launch_options": main_cfg["launch_options"],
                   "wall_s": main_cfg["wall_s"]},
        "roofline": main_cfg["roofline"],
        "speedup_vs_fp64": main_cfg["speedup_vs_fp64"],
        "error_vs_fp64": main_cfg["error_vs_fp64"],
        "variants": main_cfg["variants"],
        "apply_equiv_tflops": 524.0 * units / (ms_step * 1e-3) / 1e12,
        "cpu_baseline": main_cfg["cpu_baseline"],
        "parity": main_cfg["parity"],
        "other_workloads": others,
    }
    print(json.dumps(out))
    ctx.close()


if __name__ == "__main__":
    main()
