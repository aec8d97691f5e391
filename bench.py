#!/usr/bin/env python3
"""Benchmark of the MI355X system-matrix generator.

One step = one full pass of the photoelectric (PE) system-matrix kernel over
the full-size geometry: 64 x 64 x 32 voxels at 0.5 mm, 2048 detector bars in
two layers, 400-hole tungsten plate, 3 x 3 x 3 sub-cells per bar: 268,435,456
matrix elements, 7.25e9 sub-cell rays.  The geometry and the output buffer
(1 GiB) are resident on the GPU before the timed region; the output stays on
the GPU.  Prints one JSON line.

Unless --no-extras is given, the line also carries (under "extras") the
all-FP64 time of the same step, the primary-Compton scatter throughputs with
the reference's scatter points (crystal scatter on 32x32x1 voxels x 2048 bars,
collimator scatter on 32x32x16 voxels x 64 bars), with scatter points on a
0.5 mm grid over the collimator tungsten and on a 0.25 mm grid in every bar
(8 detector rows of 2048 bars), the matrix-free forward and back-projection
times with the headline matrix, the forward and adjoint (dL/dmu) times on
the 32x32x1-voxel, 64-bar geometry and on the full-size geometry, and one
PE launch at a 200 x 200 2D field of view with 6,000 bars and 8 x 16 x 8
sub-cells.

    python bench.py [--gpus 1] [--steps 5] [--warmup 1] [--precision fp32|fp64]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "gpu-based-system-matrix-calculation-for-spect-pet_amd")
sys.path.insert(0, PKG)

METRIC = "matrix elements/s (PE, scatter) at ≤1e-5 rel err vs FP64; fwd+grad wall time"
# MI355X published peaks
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0
# measured HBM traffic of one FP32 launch of the full-size PE step (rocprofv3 PMC
# FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0, factors from a plain 1 GiB copy)
TRAFFIC_BYTES = 1.154e9
# executed work of the same launch (rocprofv3 PMC, FP32 path): FLOPs per sub-cell ray
# the lanes execute (64 x SQ_INSTS_VALU_FLOPS_FP32 + _FP64 per ray), VALU wave
# instructions per ray x 64, transcendental instructions per ray x 64 (FP32 + FP64),
# VALUBusy, lane utilisation and mean resident waves per SIMD (of the 4 the
# kernel's registers and LDS allow)
EXEC_SOURCE = "profiles/pe_forward_instruction_mix_fp32_fp64_64x64x32_2048bars.md"
EXEC_FLOPS_PER_RAY = 250.9 + 182.1
VALU_INSTS_PER_RAY = 819.6
TRANS_INSTS_PER_RAY = 23.0 + 4.8
VALU_BUSY = 0.841
# the same instructions by issue cost (bin/valu_issue_probe on this GPU: 4 cycles
# per wave instruction for FP32/FP64 arithmetic, integer and conversion, ~2 for
# compares, selects, min/max and moves, ~8 for a transcendental)
VALU_4CYCLE_INSTS_PER_RAY = 441.7
VALU_2CYCLE_INSTS_PER_RAY = 377.9
VALU_CYCLE_SOURCE = "profiles/valu_issue_rates_gfx950.md"
LANE_UTILISATION = 0.738
WAVES_PER_SIMD = 3.73
MAX_WAVES_PER_SIMD = 4
# unpacked FP32 (= FP64) vector FMA peak, the transcendental peak (a quarter of
# the unpacked FMA rate in operations), and the VALU issue reference rate (256 CUs x
# 4 SIMDs x 2.4 GHz, one wave64 instruction per 4 cycles: the unpacked FMA rate; the
# bitmap kernel's FP32 compare-and-select mix issues ~1.2 per 4 cycles, so this is a
# yardstick, not a ceiling for every instruction mix)
PEAK_FP32_UNPACKED_TFLOPS = 78.6
PEAK_TRANS_OPS = PEAK_FP32_UNPACKED_TFLOPS * 1e12 / 2 / 4
PEAK_VALU_WAVE_INSTS_PER_S = 256 * 4 * 2.4e9 / 4


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(p, sub, target_s: float):
    """FP64 host oracle (OpenMP, all host cores) on a seeded sample of the same
    workload, scaled to elements/s."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle

    o = oracle.Oracle(p)
    threads = oracle.max_threads()
    rng = np.random.default_rng(3)
    n = 4096
    t_used = 0.0
    done = 0
    while True:
        d = rng.integers(0, o.n_det, n)
        v = rng.integers(0, o.n_vox, n)
        t0 = time.perf_counter()
        o.pe_pairs(d, v, sub=sub, nthreads=threads)
        dt = time.perf_counter() - t0
        t_used += dt
        done += n
        if t_used >= target_s:
            break
        n = int(min(max(n * 2, 4096), max(4096, n * (target_s - t_used) / max(dt, 1e-3))))
    return {
        "value": done / t_used,
        "unit": "matrix elements/s",
        "cores": threads,
        "cpu_model": _cpu_model(),
        "kind": "port",
        "sample": f"{done} seeded random (bar, voxel) elements of the same geometry and "
                  f"sub-cells, FP64 oracle, {t_used:.1f} s on {threads} host threads",
    }


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def ops_per_ray(p, sub):
    """Algorithmic FP operations per sub-cell ray of the reference formula,
    counted by the oracle's op counter on a seeded sample of elements."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle

    o = oracle.Oracle(p)
    rng = np.random.default_rng(12)
    d = rng.integers(0, o.n_det, 512)
    v = rng.integers(0, o.n_vox, 512)
    return o.pe_ops_per_ray(d, v, sub=sub)


# executed work of scatter_kernel<false> per (row, voxel, scatter point, cell) term
# (rocprofv3 PMC, FP32): FLOPs the lanes execute, VALU instructions x 64, lane
# utilisation and VALUBusy
SCATTER_EXEC_SOURCE = "profiles/scatter_kernels_4_workloads.md"
SCATTER_EXEC = {
    "crystal_scatter": {"flops_per_term": 89.3, "valu_insts_per_term": 110.9, "lane_utilisation": 0.754,
                        "valu_busy": 0.68},
    "collimator_grid": {"flops_per_term": 275.8, "valu_insts_per_term": 359.7, "lane_utilisation": 0.703,
                        "valu_busy": 0.56},
}


def scatter_roofline(params, p, kind, d, v, kernel_ms, terms, sample, exec_key=None, **kw):
    """Roofline of a scatter launch: algorithmic operations per (row, voxel,
    scatter point, cell) term of the reference's scatter element body, counted
    by the oracle's op counter on a seeded sample, times the launch's terms;
    with exec_key, also the FLOPs the lanes execute (PMC profile) over the
    workload's kernel time (all its kernels, so a lower bound for the scatter
    kernel itself)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    o = oracle.Oracle(p)
    o.set_mu_tables(*params.synthetic_mu_tables(True))
    f, n = o.scatter_ops_per_term(kind, d, v, **kw)
    achieved = f * terms / (kernel_ms * 1e-3) / 1e12
    out = {"bound": "valu", "alg_ops_per_term": round(f, 1), "achieved": round(achieved, 3),
           "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
           "op_count_sample": f"{sample} ({int(n)} terms)"}
    if exec_key:
        e = SCATTER_EXEC[exec_key]
        ex = e["flops_per_term"] * terms / (kernel_ms * 1e-3) / 1e12
        out["executed"] = {"what": "FLOPs the lanes execute (PMC) over the workload's kernel time, "
                                   "against the unpacked FP32/FP64 vector peak",
                           "flops_per_term": e["flops_per_term"], "achieved": round(ex, 2),
                           "peak": PEAK_FP32_UNPACKED_TFLOPS, "unit": "TFLOP/s",
                           "frac": round(ex / PEAK_FP32_UNPACKED_TFLOPS, 4),
                           "valu_insts_per_term": e["valu_insts_per_term"],
                           "lane_utilisation": e["lane_utilisation"], "valu_busy": e["valu_busy"],
                           "source": SCATTER_EXEC_SOURCE}
    return out


def timed(fn, reps):
    """Median kernel time (ms) of reps calls of fn() -> (result, stats)."""
    ks = sorted(fn()[1]["kernel_ms"] for _ in range(reps))
    return ks[len(ks) // 2]


WORKLOADS = ("crystal_scatter", "collimator_scatter", "collimator_grid", "crystal_grid",
             "bitmaps", "adjoint", "large_fov_pe")


def extras(smg, params, pe_ctx, out, precision, sections=None, precs=("fp32", "fp64")):
    """Secondary measurements: FP64 vs FP32 on the headline step, scatter
    throughputs, bitmaps, adjoint cost.  sections: the WORKLOADS to run (all,
    plus the ones on the headline context, when None); precs: the precisions
    of the scatter workloads."""
    import numpy as np

    def want(name):
        return sections is None or name in sections

    ex = {}
    if pe_ctx is not None:
        ex.update(_headline_extras(np, pe_ctx, out, precision))
    if want("crystal_scatter"):
        _crystal_scatter(np, smg, params, ex, precs)
    if want("collimator_scatter"):
        _collimator_scatter(np, smg, params, ex, precs)
    if want("collimator_grid"):
        _collimator_grid(np, smg, params, ex, precs)
    if want("crystal_grid"):
        _crystal_grid(np, smg, params, ex, precs)
    if want("bitmaps"):
        _bitmaps(np, smg, params, ex)
    if want("adjoint"):
        _adjoint(np, smg, params, ex)
    if want("large_fov_pe"):
        _large_fov_pe(smg, params, ex)
    if pe_ctx is not None:
        ms_this = ex["projection"]["matrix_kernel_ms"]
        _adjoint_full_size(np, pe_ctx, ms_this, ex)
    return ex


def _headline_extras(np, pe_ctx, out, precision):
    ex = {}
    other = "fp64" if precision == "fp32" else "fp32"
    pe_ctx.pe_forward(0, other, device_ptr=out.data_ptr())
    ms_other = timed(lambda: pe_ctx.pe_forward(0, other, device_ptr=out.data_ptr()), 1)
    ms_this = timed(lambda: pe_ctx.pe_forward(0, precision, device_ptr=out.data_ptr()), 1)
    ex[f"pe_{other}_kernel_ms"] = round(ms_other, 2)
    ex["pe_fp64_over_fp32_time"] = round((ms_other / ms_this) if other == "fp64" else (ms_this / ms_other), 3)
    log(f"PE {other}: {ms_other:.1f} ms")

    # generation time as the drivers see it: the same step into host memory
    # (kernels + device-to-host copy of the 1 GiB block, overlapped by row chunks)
    host = np.empty((pe_ctx.dims.n_det, pe_ctx.dims.n_vox), np.float32)
    host.fill(0.0)  # first touch outside the timing
    pe_ctx.pe_forward(0, precision, out=host)
    walls = []
    for _ in range(3):
        t0 = time.perf_counter()
        pe_ctx.pe_forward(0, precision, out=host)
        walls.append((time.perf_counter() - t0) * 1e3)
    wall = sorted(walls)[1]
    ex["pe_generation_incl_d2h"] = {
        "ms": round(wall, 2), "elements_per_s": round(host.size / wall * 1e3, 1),
        "bytes_to_host": host.nbytes,
        "what": "wall time of smg_pe_forward with a pageable host output: kernels plus the "
                "device-to-host copy, which overlaps the kernels in 8 row chunks"}
    log(f"PE into host memory: {wall:.1f} ms")

    # matrix-free projections with the same matrix (elements formed in-kernel, never stored)
    rng = np.random.default_rng(4)
    xv = rng.uniform(0, 1, pe_ctx.dims.n_vox).astype(np.float32)
    yv = rng.uniform(0, 1, pe_ctx.dims.n_det).astype(np.float32)
    pe_ctx.pe_project(xv, 0, precision)
    f_ms = timed(lambda: pe_ctx.pe_project(xv, 0, precision), 1)
    b_ms = timed(lambda: pe_ctx.pe_backproject(yv, 0, precision), 1)
    ex["projection"] = {"forward_kernel_ms": round(f_ms, 2), "back_kernel_ms": round(b_ms, 2),
                        "matrix_kernel_ms": round(ms_this, 2),
                        "elements_per_s_forward": round(pe_ctx.dims.n_det * pe_ctx.dims.n_vox / f_ms * 1e3, 1),
                        "workload": "y = A x and x = A^T y with the full-size PE matrix, matrix-free "
                                    "(no 1 GiB matrix in HBM or on disk)"}
    log(f"projection: forward {f_ms:.1f} ms, back {b_ms:.1f} ms (matrix {ms_this:.1f} ms)")
    return ex


def _crystal_scatter(np, smg, params, ex, precs):
    # crystal scatter (ScatterGen_Crystal path): 32x32x1 voxels, 2048 bars, 2047 scatter points
    p4 = params.config(4)
    with smg.open_context(p4) as c:
        c.set_pe_subsamples(*p4.meta["pe_subsamples"])
        pe, _ = c.pe_forward(0, "fp32")
        c.set_scatter_subsamples(*p4.meta["subsamples"])
        c.set_mu_tables(*params.synthetic_mu_tables(True))
        c.scatter_crystal(pe, 0, "fp32")  # builds the bitmap, warms up
        for prec in precs:
            _, st = c.scatter_crystal(pe, 0, prec)
            ms = timed(lambda: c.scatter_crystal(pe, 0, prec), 3 if prec == "fp32" else 1)
            ex[f"crystal_scatter_{prec}"] = {
                "elements_per_s": round(st["elements"] / ms * 1e3, 1),
                "pairs_per_s": round(st["rays"] / ms * 1e3, 1),
                "kernel_ms": round(ms, 3), "fp64_fallbacks": st["fp64_fallbacks"]}
            log(f"crystal scatter {prec}: {ms:.2f} ms")
            if prec == "fp32":
                rng = np.random.default_rng(13)
                d, v = rng.integers(0, p4.n_det, 16), rng.integers(0, p4.n_vox, 16)
                ex["crystal_scatter_fp32"]["roofline"] = scatter_roofline(
                    params, p4, 0, d, v, ms, st["rays"], "16 seeded elements", exec_key="crystal_scatter",
                    bitmap=c.bitmap_crystal())
        ex["crystal_scatter_workload"] = ("32x32x1 voxels x 2048 bars, scatter points at the "
                                          "2047 other bar centres (reference semantics), 1x1x1 cells, "
                                          "64-node mu(E) tables")


def _collimator_scatter(np, smg, params, ex, precs):
    # collimator scatter (combined variant): 32x32x16 voxels, 64 bars, 16 hole centres
    p3 = params.config(3)
    with smg.open_context(p3) as c:
        c.set_scatter_subsamples(*p3.meta["subsamples"])
        c.set_mu_tables(*params.synthetic_mu_tables(True))
        c.scatter_collimator(0, "combined", "fp32")
        for prec in precs:
            _, st = c.scatter_collimator(0, "combined", prec)
            ms = timed(lambda: c.scatter_collimator(0, "combined", prec), 3)
            ex[f"collimator_scatter_{prec}"] = {
                "elements_per_s": round(st["elements"] / ms * 1e3, 1),
                "pairs_per_s": round(st["rays"] / ms * 1e3, 1),
                "kernel_ms": round(ms, 3), "fp64_fallbacks": st["fp64_fallbacks"]}
            log(f"collimator scatter {prec}: {ms:.2f} ms")
            if prec == "fp32":
                rng = np.random.default_rng(14)
                d, v = rng.integers(0, p3.n_det, 64), rng.integers(0, p3.n_vox, 64)
                ex["collimator_scatter_fp32"]["roofline"] = scatter_roofline(
                    params, p3, 1, d, v, ms, st["rays"], "64 seeded elements", bitmap=c.bitmap_collimator())
        ex["collimator_scatter_workload"] = ("32x32x16 voxels x 64 bars, scatter points at the 16 "
                                             "hole centres (reference semantics), 1x1x1 cells, "
                                             "64-node mu(E) tables")


def _collimator_grid(np, smg, params, ex, precs):
    # collimator scatter from a 0.5 mm grid over the tungsten (first- and second-leg
    # collimator attenuation), 32x32x16 voxels, 64 bars
    p3 = params.config(3)
    with smg.open_context(p3) as c:
        c.set_mu_tables(*params.synthetic_mu_tables(True))
        for prec in precs:
            _, st, n = c.scatter_collimator_grid(0.5, 0, prec)
            ex[f"collimator_grid_scatter_{prec}"] = {
                "elements_per_s": round(st["elements"] / st["kernel_ms"] * 1e3, 1),
                "pairs_per_s": round(st["rays"] / st["kernel_ms"] * 1e3, 1),
                "kernel_ms": round(st["kernel_ms"], 1), "fp64_fallbacks": st["fp64_fallbacks"]}
            log(f"collimator grid scatter {prec}: {st['kernel_ms']:.0f} ms, {n} points")
            if prec == "fp32":
                rng = np.random.default_rng(15)
                d, v = rng.integers(0, p3.n_det, 2), rng.integers(0, p3.n_vox, 2)
                ex["collimator_grid_scatter_fp32"]["roofline"] = scatter_roofline(
                    params, p3, 2, d, v, st["kernel_ms"], st["rays"], "2 seeded elements",
                    exec_key="collimator_grid", pitch=0.5)
        ex["collimator_grid_scatter_workload"] = (f"32x32x16 voxels x 64 bars, {n} scatter points on "
                                                  "a 0.5 mm grid over the 60x5x60 mm tungsten, "
                                                  "1x1x1 cells, 64-node mu(E) tables")


def _crystal_grid(np, smg, params, ex, precs):
    # crystal scatter from a 0.25 mm grid in every bar (10x32x10 cells of the 2.5x8x2.5 mm
    # bars), 32x32x1 voxels, 2048 bars, 8 detector rows (every 256th bar)
    p4 = params.config(4)
    rows = np.arange(0, p4.n_det, 256, dtype=np.int32)
    with smg.open_context(p4) as c:
        c.set_mu_tables(*params.synthetic_mu_tables(True))
        for prec in precs:
            _, st = c.scatter_crystal_grid((10, 32, 10), rows, 0, prec)
            ex[f"crystal_grid_scatter_{prec}"] = {
                "elements_per_s": round(st["elements"] / st["kernel_ms"] * 1e3, 1),
                "pairs_per_s": round(st["rays"] / st["kernel_ms"] * 1e3, 1),
                "kernel_ms": round(st["kernel_ms"], 1), "fp64_fallbacks": st["fp64_fallbacks"]}
            log(f"crystal grid scatter {prec}: {st['kernel_ms']:.0f} ms")
            if prec == "fp32":
                rng = np.random.default_rng(16)
                d, v = rows[rng.integers(0, len(rows), 2)], rng.integers(0, p4.n_vox, 2)
                ex["crystal_grid_scatter_fp32"]["roofline"] = scatter_roofline(
                    params, p4, 3, d, v, st["kernel_ms"], st["rays"],
                    "2 seeded elements of the timed rows on a 2x8x2 grid (the per-term count does not "
                    "depend on the grid pitch; the 10x32x10 count takes ~30 s per element)",
                    bitmap=c.bitmap_crystal(), grid=(2, 8, 2))
        ex["crystal_grid_scatter_workload"] = (
            f"32x32x1 voxels x {len(rows)} detector rows of 2048 bars, scatter points at the "
            f"centres of 10x32x10 cells (0.25 mm) of every bar ({p4.n_det * 3200} points), "
            "1x1x1 cells, 64-node mu(E) tables")


def _bitmaps(np, smg, params, ex):
    # geometry bitmaps of the scatter kernels: the full-size detector and 6,000 bars
    for tag, pb in (("2048_bars", params.config(2)), ("6000_bars", params.large_detector())):
        with smg.open_context(pb) as c:
            n = pb.n_det
            c.build_bitmap(False)
            ms_c = min(c.build_bitmap(False) for _ in range(2))
            c.build_bitmap(True)
            ms_l = min(c.build_bitmap(True) for _ in range(2))
            bits_c = n * (n + 1) // 2 * n
            bits_l = c.dims.n_holes_l0 * n * n
        ex[f"bitmaps_{tag}"] = {
            "crystal_ms": round(ms_c, 2), "crystal_bits_per_s": round(bits_c / ms_c * 1e3, 1),
            "crystal_GB": round(bits_c / 8e9, 3),
            "collimator_ms": round(ms_l, 2), "collimator_bits_per_s": round(bits_l / ms_l * 1e3, 1),
            "collimator_GB": round(bits_l / 8e9, 3),
            "workload": f"{n} bars, 400 layer-0 holes; wall time of the call (kernel + sync)"}
        log(f"bitmaps {tag}: crystal {ms_c:.1f} ms, collimator {ms_l:.1f} ms")


def _adjoint(np, smg, params, ex):
    # adjoint: forward + dL/dmu on the 32x32x1-voxel, 64-bar geometry, 4x4x4 sub-cells
    p5 = params.config(5)
    w = np.random.default_rng(2).uniform(0, 1, (p5.n_det, p5.n_vox)).astype(np.float32)
    with smg.open_context(p5) as c:
        c.set_pe_subsamples(*p5.meta["subsamples"])
        c.pe_forward(0, "fp32")
        c.pe_backward(w)
        f_ms = timed(lambda: c.pe_forward(0, "fp32"), 5)
        b_ms = timed(lambda: c.pe_backward(w)[1:], 5)
    ex["adjoint"] = {"forward_kernel_ms": round(f_ms, 3), "backward_kernel_ms": round(b_ms, 3),
                     "fwd_plus_grad_ms": round(f_ms + b_ms, 3),
                     "backward_over_forward": round(b_ms / f_ms, 2),
                     "workload": "32x32x1 voxels x 64 bars, 4x4x4 sub-cells, 1 + 128 mu"}
    log(f"adjoint: forward {f_ms:.3f} ms, backward {b_ms:.3f} ms")


def _large_fov_pe(smg, params, ex):
    # the PE matrix at the size the reference's README quotes (200 x 200 2D FOV, 6,000
    # bars, 8 x 16 x 8 sub-cells): one FP32 launch into a 0.96 GB device buffer
    import torch
    p = params.large_fov_pe()
    with smg.open_context(p) as c:
        c.set_pe_subsamples(*p.meta["subsamples"])
        buf = torch.empty(c.dims.n_det * c.dims.n_vox, dtype=torch.float32, device="cuda")
        _, st = c.pe_forward(0, "fp32", device_ptr=buf.data_ptr())
        del buf
    ms = st["kernel_ms"]
    ex["large_fov_pe"] = {"kernel_ms": round(ms, 1), "elements_per_s": round(st["elements"] / ms * 1e3, 1),
                          "sub_cell_rays_per_s": round(st["rays"] / ms * 1e3, 1),
                          "fp64_fallback_rays": st["fp64_fallbacks"], "overflow_waves": st["overflow_waves"],
                          "workload": "200x200x1 voxels at 0.5 mm x 6000 bars (3 layers of 50x40), "
                                      "400-hole plate, 8x16x8 sub-cells, FP32; one launch"}
    log(f"PE 200x200 FOV, 6000 bars: {ms:.0f} ms")


def _adjoint_full_size(np, pe_ctx, ms_this, ex):
    # the adjoint on the headline geometry (2048 bars, 64x64x32 voxels, 1 + 4096 mu)
    wf = np.random.default_rng(2).uniform(0, 1, (pe_ctx.dims.n_det, pe_ctx.dims.n_vox)).astype(np.float32)
    bf_ms = timed(lambda: pe_ctx.pe_backward(wf)[1:], 1)
    ex["adjoint_full_size"] = {"forward_kernel_ms": round(ms_this, 2), "backward_kernel_ms": round(bf_ms, 2),
                               "backward_over_forward": round(bf_ms / ms_this, 2),
                               "workload": "64x64x32 voxels x 2048 bars, 3x3x3 sub-cells, 1 + 4096 mu"}
    log(f"adjoint, full size: backward {bf_ms:.1f} ms (forward {ms_this:.1f} ms)")


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", choices=("fp32", "fp64"), default="fp32")
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="host time budget of the CPU-baseline sample (0 to skip)")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the FP64, scatter and adjoint measurements")
    ap.add_argument("--workloads", default=None,
                    help="only run these secondary workloads (comma list of "
                         + ", ".join(WORKLOADS) + ") and print their timings as one JSON "
                         "line; with --precision, the scatter workloads run in that "
                         "precision only (profiling runs)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device-to-device copy three times (the "
                         "reference kernel for scaling FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(a multi-GPU run would give each GPU its own detector rows)", file=sys.stderr)
        sys.exit(2)
    names = [w for w in (args.workloads or "").split(",") if w]
    unknown = [w for w in names if w not in WORKLOADS]
    if unknown:
        print(f"bench.py: unknown workloads {unknown}; choose from {WORKLOADS}", file=sys.stderr)
        sys.exit(2)

    import torch

    if args.copy_calibration:
        n = 1 << 28
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "copies": 3}))
        return

    import smg
    from smg import params

    if names:
        print(json.dumps(extras(smg, params, None, None, args.precision, sections=names,
                                precs=(args.precision,))))
        return

    p = params.config(2)
    sub = p.meta["subsamples"]
    ctx = smg.open_context(p, device=0)
    ctx.set_pe_subsamples(*sub)
    n_el = ctx.dims.n_det * ctx.dims.n_vox
    out = torch.empty(n_el, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()

    for i in range(args.warmup):
        _, st = ctx.pe_forward(0, args.precision, device_ptr=out.data_ptr())
        log(f"warmup {i}: kernel {st['kernel_ms']:.1f} ms")
    torch.cuda.synchronize()
    ctx.synchronize()
    kernel_ms = []
    fallbacks = 0
    t0 = time.perf_counter()
    for i in range(args.steps):
        _, st = ctx.pe_forward(0, args.precision, device_ptr=out.data_ptr())
        kernel_ms.append(st["kernel_ms"])
        fallbacks += st["fp64_fallbacks"]
        log(f"step {i}: kernel {st['kernel_ms']:.1f} ms")
    ctx.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    rays = st["rays"]
    ms_per_step = wall / args.steps * 1e3
    value = n_el * args.steps / wall
    avg_kernel_s = sum(kernel_ms) / len(kernel_ms) / 1e3

    f_ray = ops_per_ray(p, sub)
    achieved = f_ray * rays / avg_kernel_s / 1e12
    roofline = {
        "bound": "valu",
        "achieved": round(achieved, 3),
        "peak": PEAK_FP32_TFLOPS,
        "unit": "TFLOP/s",
        "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
        # HBM bytes per launch: FETCH_SIZE + WRITE_SIZE from rocprofv3 --pmc passes of this
        # command, scaled by the same counters on a plain 1 GiB copy (bench.py --copy-calibration)
        "traffic": TRAFFIC_BYTES if args.precision == "fp32" else None,
        "traffic_source": "profiles/pe_forward_kernel_64x64x32_2048bars.md",
        "kernel": "pe_forward_kernel",
        "flops_per_ray": round(f_ray, 1),
        "rays_per_launch": rays,
        "write_GBps": round(n_el * 4 / avg_kernel_s / 1e9, 2),
    }
    if args.precision == "fp32":
        ex_achieved = EXEC_FLOPS_PER_RAY * rays / avg_kernel_s / 1e12
        issue = VALU_INSTS_PER_RAY * rays / 64 / avg_kernel_s
        valu_cycles = (4 * VALU_4CYCLE_INSTS_PER_RAY + 2 * VALU_2CYCLE_INSTS_PER_RAY
                       + 4 * TRANS_INSTS_PER_RAY) * rays / 64
        trans = TRANS_INSTS_PER_RAY * rays / avg_kernel_s
        roofline["executed"] = {
            "what": "FLOPs the lanes execute (PMC), against the unpacked FP32/FP64 vector peak",
            "flops_per_ray": EXEC_FLOPS_PER_RAY,
            "achieved": round(ex_achieved, 2),
            "peak": PEAK_FP32_UNPACKED_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(ex_achieved / PEAK_FP32_UNPACKED_TFLOPS, 4),
            "valu_insts_per_ray": VALU_INSTS_PER_RAY,
            "valu_issue_frac": round(issue / PEAK_VALU_WAVE_INSTS_PER_S, 4),
            "valu_busy": VALU_BUSY,
            "valu_cycle_frac": round(valu_cycles / (avg_kernel_s * 256 * 4 * 2.4e9), 4),
            "valu_cycle_model": "4/2/8 cycles per arithmetic/compare-select/transcendental wave "
                                "instruction, " + VALU_CYCLE_SOURCE,
            "lane_utilisation": LANE_UTILISATION,
            "transcendental_per_ray": TRANS_INSTS_PER_RAY,
            "transcendental_frac": round(trans / PEAK_TRANS_OPS, 4),
            "occupancy_waves_per_simd": WAVES_PER_SIMD,
            "occupancy_limit": f"{MAX_WAVES_PER_SIMD} waves per SIMD (128 VGPRs, 10 KB of LDS per wave)",
            "source": EXEC_SOURCE,
        }
    ex = None if args.no_extras else extras(smg, params, ctx, out, args.precision)
    cpu = cpu_baseline(p, sub, args.cpu_seconds) if args.cpu_seconds > 0 else None
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "matrix elements/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.precision == "fp32" else "f64",
        "data": "synthetic",
        "config": {
            "workload": "PE matrix, full size: 64x64x32 voxels @0.5 mm x 2048 bars "
                        "(2 layers of 32x32, 2.5x8x2.5 mm), 80x80x8 mm tungsten plate with "
                        "400 holes, 3x3x3 sub-cells, 1 rotation",
            "elements": n_el,
            "rays_per_step": rays,
            "precision": args.precision,
            "fp64_fallback_rays_per_step": fallbacks // max(args.steps, 1),
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
        "extras": ex,
    }
    print(json.dumps(line), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
