#!/usr/bin/env python3
"""One-GPU benchmark of the radiation-transfer hot path; prints ONE JSON line.

Default workload (config 3, the diurnal series): an orchard of 25 trees (1,000,000 triangles + 1,600
ground patches), 64 sun positions x 3 bands (LW with emission, NIR, PAR), 200 direct + 200 diffuse rays
per primitive (d = n = 15), scattering depth 4.  One step = one diurnal series solved by the batched
backend (all 64 suns in one set of persistent launches).  `value` = solves/s of the batched backend
(64 solves per step / step time).  The plain one-solve-per-launch backend is timed on a few suns of the
same series and reported beside it, with the speedup.

The default run also serves a config-4 request stream (Poisson arrivals over random canopies of 10k-2M
primitives) through the request server and through the plain loop and reports both under "serving"
(--no-serve skips it).

Other workloads: --workload config2 (200k-triangle periodic canopy, one 2-band solve per step),
--workload castrays (CollisionDetection::castRaysSoA of 1e8 rays over the 5x tiled orchard),
--workload camera (a 1024x1024 x 64-sample camera image of the same 5x orchard),
--workload homogeneous (the reference's published radiation benchmark, timed as it times runBand) and
--workload serve (the serving leg alone).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from helios_amd import native, scenes  # noqa: E402
from helios_amd import workloads as W  # noqa: E402

METRIC = "Mrays/s & solves/s batched vs plain at <=0.5% flux err; p50/p99 solve latency"
HBM_PEAK_GBS = 8000.0
# canonical algorithmic bytes per traced ray: 64 B per BVH node visit, 48 B per primitive test,
# 32 B of ray/result I/O (node/prim counts are measured by the kernels' own counters)
B_NODE, B_PRIM, B_IO = 64, 48, 32


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def device_sync():
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except Exception:
        pass


def pct(a, q):
    return float(np.percentile(np.asarray(a, np.float64), q)) if len(a) else None


# position of the COUNT template argument: k_cast<STK, COUNT, FLAGS>, k_camera / k_direct_src<STK, COUNT, FLAGS>,
# k_direct / k_diffuse<WAVES, STK, COUNT, FLAGS>
COUNT_ARG = {"k_cast": 1, "k_camera": 1, "k_direct_src": 1, "k_direct": 2, "k_diffuse": 2}


def counting_variant(name):
    m = re.search(r"(k_\w+)<([^>]*)>", name)
    if not m or m.group(1) not in COUNT_ARG:
        return False
    return [a.strip() for a in m.group(2).split(",")][COUNT_ARG[m.group(1)]] == "true"


def kernel_flags(name):
    """The traversal-flags template argument of a ray kernel's name (its last integer argument; boolean
    arguments such as the camera's specular switch may follow it), as an int, else None."""
    m = re.search(r"(k_\w+)<([^>]*)>", name)
    if not m:
        return None
    nums = [a.strip() for a in m.group(2).split(",") if a.strip().isdigit()]
    return int(nums[-1]) if nums else None


def pmc_profile(kernel, profile="pmc_traffic_config3_diurnal.json", flags=0):
    """HBM bytes per dispatch of `kernel` from the committed rocprofv3 PMC summary: FETCH_SIZE + WRITE_SIZE of
    its largest timed (non-counting) dispatch with the traversal flags of the timed launch (0: no masks,
    reference predicates; other variants in the same profile, such as the watertight solves of the predicate
    comparison, are different kernels), each scaled by what the counter read for a copy of known size, and
    that kernel's L2 hit rate when the profile has one.  None without a profile of the kernel."""
    path = os.path.join(ROOT, "profiles", profile)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    best = None
    for name, cs in d.get("kernels", {}).items():
        if (re.search(re.escape(kernel) + r"<", name) and not counting_variant(name)
                and (flags is None or kernel_flags(name) == flags) and "FETCH_SIZE" in cs and "WRITE_SIZE" in cs):
            b = cs["FETCH_SIZE"]["max_bytes"] + cs["WRITE_SIZE"]["max_bytes"]
            if best is None or b > best["bytes"]:
                best = {"bytes": b, "fetch": cs["FETCH_SIZE"]["max_bytes"], "write": cs["WRITE_SIZE"]["max_bytes"],
                        "l2_hit_rate": cs.get("l2_hit_rate"), "profiled_kernel": name}
    return best


def pmc_traffic(kernel, profile="pmc_traffic_config3_diurnal.json", flags=0):
    p = pmc_profile(kernel, profile, flags)
    return None if p is None else p["bytes"]


def roofline(kernel, avg_launch_ms, canonical_bytes, profile, flags=0, **extra):
    """Roofline of the dominant kernel.  achieved = the kernel's HBM bytes per launch (rocprofv3 FETCH_SIZE +
    WRITE_SIZE, calibrated, from the committed profile) / its HIP-event launch time, so frac is a fraction of the
    8 TB/s HBM peak.  The canonical per-ray byte model (64 B per BVH node visit, 48 B per primitive test, 32 B of
    I/O, counted by the kernels) is reported under "canonical": most of those bytes come from L2 and the
    Infinity Cache, so its rate is a measure of work, not a bandwidth, and may exceed the peak."""
    prof = pmc_profile(kernel, profile, flags)
    sec = avg_launch_ms / 1e3
    ach = None if prof is None else prof["bytes"] / sec / 1e9
    canon_gbs = canonical_bytes / sec / 1e9
    return {"bound": "hbm", "kernel": kernel, "achieved": None if ach is None else round(ach, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": None if ach is None else round(ach / HBM_PEAK_GBS, 4),
            "traffic": None if prof is None else round(prof["bytes"]), "avg_launch_ms": round(avg_launch_ms, 3),
            "traffic_fetch": None if prof is None else round(prof["fetch"]),
            "traffic_write": None if prof is None else round(prof["write"]),
            "l2_hit_rate": None if prof is None else prof["l2_hit_rate"],
            "canonical": {"bytes_per_launch": round(canonical_bytes), "gbs": round(canon_gbs, 1),
                          "exceeds_hbm_peak": canon_gbs > HBM_PEAK_GBS, **extra}}


# occlusion-grid direct pass: 8 B cell range per lookup ("nodes" counts lookups), 4 B list entry + 48 B
# primitive record per candidate tested ("prims")
B_CELL, B_CAND = 8, 52


def bytes_per_ray(c, grid=False):
    if not c["rays"]:
        return None
    if grid:
        return B_CELL * c["nodes"] / c["rays"] + B_CAND * c["prims"] / c["rays"] + B_IO
    return B_NODE * c["nodes"] / c["rays"] + B_PRIM * c["prims"] / c["rays"] + B_IO


FRACS = (0.1, 0.3, 0.5, 0.7, 0.9)  # starts of the sampled primitive ranges, as fractions of P


def oracle_samples(o, s, base, sun_dir, direct_seed, width):
    """The CPU oracle (the restatement of the reference CPU path) on primitive ranges [f P, f P + width) for the
    fractions FRACS: one direct pass towards sun_dir with direct_seed and one diffuse pass (both faces) with
    the request's primary diffuse seed -- the bench's own seeds, so the results double as a parity sample.
    Returns the per-range outputs and the oracle's time and rays for each pass."""
    rp = W.rays_per_solve(s, base)
    d, n = rp["d"], rp["n"]
    sd = np.ascontiguousarray(sun_dir, np.float32).reshape(3)
    P = s.P
    out = {"ranges": [], "hits": [], "sums": [], "t_dif": 0.0, "r_dif": 0, "t_dir": 0.0, "r_dir": 0}
    for f in FRACS:
        p0 = int(f * P)
        p1 = min(P, p0 + width)
        t = time.perf_counter()
        hits = o.diffuse_hits_range(n, int(base.dseeds[0]), p0, p1)
        out["t_dif"] += time.perf_counter() - t
        out["r_dif"] += int(np.sum(hits != 0xFFFFFFFE))
        t = time.perf_counter()
        sums = o.direct_sums_range(d, int(direct_seed), sd, p0, p1)
        out["t_dir"] += time.perf_counter() - t
        out["r_dir"] += (p1 - p0) * d * d
        out["ranges"].append((p0, p1))
        out["hits"].append(hits)
        out["sums"].append(sums)
    return out


def cpu_baseline_diurnal(o, s, base, sun_dir, direct_seed, seconds):
    """Time the oracle on bounded samples of primitives (about `seconds` of CPU work); extrapolate to plain
    solves/s by ray count.  Returns the baseline and the samples (reused as the parity sample)."""
    rp = W.rays_per_solve(s, base)
    d, n = rp["d"], rp["n"]
    probe = oracle_samples(o, s, base, sun_dir, direct_seed, 200)
    per = (probe["t_dif"] + probe["t_dir"]) / 200.0
    width = max(200, int(seconds / max(per, 1e-9) / len(FRACS)))
    smp = oracle_samples(o, s, base, sun_dir, direct_seed, width)
    dif_rate, dir_rate = smp["r_dif"] / smp["t_dif"], smp["r_dir"] / smp["t_dir"]
    t_solve = rp["direct"] / dir_rate + rp["diffuse"] / dif_rate
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    nprim = sum(p1 - p0 for p0, p1 in smp["ranges"])
    return {"value": 1.0 / t_solve, "unit": "solves/s", "cores": cores, "kind": "port",
            "sample": (f"CPU oracle (FP32 rays, FP64 sums, OpenMP) on {nprim} primitives x (one direct pass d={d} + "
                       f"one diffuse pass n={n}, both faces) of the config-3 orchard: {smp['r_dir'] + smp['r_dif']} rays "
                       f"in {smp['t_dif'] + smp['t_dir']:.1f} s ({dir_rate / 1e6:.2f} Mrays/s direct, "
                       f"{dif_rate / 1e6:.2f} Mrays/s diffuse); plain solve = {rp['total']} rays -> {t_solve:.0f} s"),
            "mrays_per_s": (smp["r_dir"] + smp["r_dif"]) / (smp["t_dif"] + smp["t_dir"]) / 1e6}, smp


def parity_diurnal(g, base, sun_dir, direct_seed, smp):
    """The GPU's diffuse hit words and direct strength sums of the full scene, at the bench's seeds, against the
    oracle's on the sampled primitive ranges: hit words must be equal, sums within 1e-5 (they are fractions of
    the source flux in [0, 1])."""
    rp_d = int(round(math.sqrt(smp["hits"][0].shape[2])))
    hits = g.diffuse_hits(rp_d, int(base.dseeds[0]))
    d = int(math.ceil(math.sqrt(base.direct_rays)))
    sums = g.direct_sums(d, int(direct_seed), [sun_dir])
    mism, rays, worst, nprim = 0, 0, 0.0, 0
    for (p0, p1), ho, so in zip(smp["ranges"], smp["hits"], smp["sums"]):
        hg = hits[p0:p1]
        launched = ho != 0xFFFFFFFE
        mism += int(np.sum(hg != ho))
        rays += int(np.sum(launched))
        worst = max(worst, float(np.max(np.abs(sums[p0:p1].astype(np.float64) - so))))
        nprim += p1 - p0
    return {"sampled_prims": nprim, "sampled_diffuse_rays": rays, "hit_word_mismatches": mism,
            "direct_sampled_rays": nprim * d * d, "direct_sum_max_abs_err": worst,
            "seeds": {"diffuse": int(base.dseeds[0]), "direct": int(direct_seed)}}


def predicates_diurnal(ctx, g, base, dirs, flux, sky, seeds):
    """One plain solve of the series' first sun with the reference's intersection predicates and again with the
    watertight ones (rt_set_predicates): per-primitive |R_wt - R_ref| / max(|R_ref|, 1 W/m^2) over primitives
    and bands (max, 99.9th percentile, counts above 1e-4 and above 0.5%) and each mode's solve time."""
    req = W.single_sun_request(base, dirs[0], flux[0], sky[0], seeds[0])
    out, ms = {}, {}
    try:
        for mode in ("reference", "watertight"):
            ctx.set_predicates(mode)
            native.run_band(g, req)  # warm
            device_sync()
            t0 = time.perf_counter()
            out[mode] = native.run_band(g, req)
            ms[mode] = (time.perf_counter() - t0) * 1e3
    finally:
        ctx.set_predicates("reference")
    a, b = out["watertight"].astype(np.float64), out["reference"].astype(np.float64)
    e = (np.abs(a - b) / np.maximum(np.abs(b), 1.0)).ravel()
    return {"max_rel_diff": float(e.max()), "p99_9_rel_diff": float(np.percentile(e, 99.9)),
            "n_above_1e-4": int(np.sum(e > 1e-4)), "n_above_0.5pct": int(np.sum(e > 5e-3)), "n": int(e.size),
            "plain_solve_ms": {k: round(v, 1) for k, v in ms.items()}}


def energy_diurnal(ctx, g, base, dirs, flux, sky, seeds, out_ptr):
    """FP64 energy bookkeeping of the series under both pass schedules (one untimed series each): the largest
    |residual| / inputs over the channels, and the channel sums."""
    K, B = dirs.shape[0], base.B
    res = {}
    native.set_energy_accounting(ctx, True)
    try:
        for name, sched in (("reference", 0), ("conserving", 1)):
            base.schedule = sched
            native.run_band_batch(g, base, dirs, flux, sky, seeds, out_device_ptr=out_ptr)
            e = native.energy_fp64(ctx, K * B)
            inputs = e["source"] + e["sky"] + e["emitted"]
            rel = np.abs(e["residual"]) / np.maximum(inputs, 1e-30)
            res[name] = {"max_rel_residual": float(rel.max()), "mean_rel_residual": float(rel.mean()),
                         "inputs_W": float(inputs.sum()), "absorbed_W": float(e["absorbed"].sum()),
                         "escaped_W": float(e["escaped"].sum()), "residual_W": float(e["residual"].sum())}
    finally:
        base.schedule = 0
        native.set_energy_accounting(ctx, False)
    return res


def calib_copy(mib):
    """A plain device-to-device copy of known size, so a PMC run can scale FETCH_SIZE / WRITE_SIZE."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    log(f"[bench] calibration copy: {n * 4} bytes read, {n * 4} bytes written")
    del src, dst


def bench_diurnal(args):
    K = args.suns
    if args.calib_copy_mib:
        calib_copy(args.calib_copy_mib)
    log(f"[bench] generating config-3 orchard ...")
    s = scenes.orchard(3)
    ctx = native.Context(0)
    t0 = time.perf_counter()
    g = ctx.scene(s.transforms, s.types, s.twosided)
    build_s = time.perf_counter() - t0
    info = g.info()
    log(f"[bench] scene {info}; build {build_s * 1e3:.1f} ms")
    base = W.diurnal_request(s)
    dirs, flux, sky, seeds = W.diurnal_suns(K)
    rp = W.rays_per_solve(s, base)
    P, B = s.P, base.B
    import ctypes as C
    out_dev = C.c_void_p()
    lib = ctx.lib
    hip = C.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    if hip.hipMalloc(C.byref(out_dev), K * P * B * 4) != 0:
        raise RuntimeError("hipMalloc failed")
    tune = None
    if not args.no_autotune:
        t0 = time.perf_counter()
        tune = ctx.autotune(g, base, nsuns=K)
        tune["seconds"] = round(time.perf_counter() - t0, 2)
        log(f"[bench] autotune {tune}")
    ctx.enable_timing(True)

    def step():
        native.run_band_batch(g, base, dirs, flux, sky, seeds, out_device_ptr=out_dev.value)

    for w in range(args.warmup):
        log(f"[bench] warmup {w}")
        step()
    # counted traversal work (one untimed series)
    ctx.set_counting(True)
    step()
    counts = ctx.counters()
    ctx.set_counting(False)
    log(f"[bench] counters {counts}")
    device_sync()
    lat = []
    kt = {"direct_ms": 0.0, "diffuse_ms": 0.0, "gather_ms": 0.0, "other_ms": 0.0, "direct_launches": 0,
          "diffuse_launches": 0}
    grid, grid_entries = False, 0.0
    t_start = time.perf_counter()
    for k in range(args.steps):
        t = time.perf_counter()
        step()
        lat.append((time.perf_counter() - t) * 1e3)
        tm = ctx.timing()
        for key in kt:
            kt[key] += tm[key]
        grid = bool(tm["direct_grid"])
        grid_entries = tm["grid_entries"] / 100.0
        log(f"[bench] step {k}: {lat[-1]:.1f} ms  {tm}")
    device_sync()
    wall = time.perf_counter() - t_start
    ms_per_step = wall * 1e3 / args.steps
    solves_s = K / (ms_per_step / 1e3)

    # plain backend: one solve per launch sequence (rt_run_band's path: a batch of one sun), a few suns of the
    # same series, output left in HBM like the batched leg's
    plain_lat = []
    idx = np.linspace(0, K - 1, args.plain_suns).round().astype(int)
    native.run_band_batch(g, base, dirs[idx[0]:idx[0] + 1], flux[idx[0]:idx[0] + 1], sky[idx[0]:idx[0] + 1],
                          seeds[idx[0]:idx[0] + 1], out_device_ptr=out_dev.value)
    for k in idx:
        t = time.perf_counter()
        native.run_band_batch(g, base, dirs[k:k + 1], flux[k:k + 1], sky[k:k + 1], seeds[k:k + 1],
                              out_device_ptr=out_dev.value)
        plain_lat.append((time.perf_counter() - t) * 1e3)
        log(f"[bench] plain sun {k}: {plain_lat[-1]:.1f} ms")
    plain_solves_s = 1e3 / float(np.mean(plain_lat))
    # with the results downloaded to host memory, as runBand hands them to the Context (PCIe-inclusive)
    t = time.perf_counter()
    native.run_band_batch(g, base, dirs, flux, sky, seeds)
    series_host_ms = (time.perf_counter() - t) * 1e3
    t = time.perf_counter()
    native.run_band(g, W.single_sun_request(base, dirs[idx[0]], flux[idx[0]], sky[idx[0]], seeds[idx[0]]))
    plain_host_ms = (time.perf_counter() - t) * 1e3

    # roofline of the dominant kernel class (HIP events on the solve stream)
    per_step = {c: kt[f"{c}_ms"] / args.steps for c in ("direct", "diffuse", "gather", "other")}
    dom = max(("direct", "diffuse"), key=lambda c: per_step[c])
    launches = kt[f"{dom}_launches"] / args.steps
    avg_launch_ms = per_step[dom] / max(launches, 1)
    brays = bytes_per_ray(counts[dom], grid=(dom == "direct" and grid))
    rays_per_launch = counts[dom]["rays"] / max(launches, 1)
    traced_rays = counts["direct"]["rays"] + counts["diffuse"]["rays"]
    # the canonical bytes of the whole step against its wall time: above the HBM peak, they cannot all be HBM bytes
    canon_step = sum(bytes_per_ray(counts[c], grid=(c == "direct" and grid)) * counts[c]["rays"]
                     for c in ("direct", "diffuse"))
    # the timed series' kernels: no texture masks, reference predicates (flags 0); direct rays with occlusion
    # grids run k_direct_q
    dom_kernel = "k_direct_q" if (dom == "direct" and grid) else f"k_{dom}"
    rl = roofline(dom_kernel, avg_launch_ms, brays * rays_per_launch, "pmc_traffic_config3_diurnal.json",
                  bytes_per_ray=round(brays, 1), rays_per_launch=int(rays_per_launch),
                  nodes_per_ray=round(counts[dom]["nodes"] / counts[dom]["rays"], 2),
                  prims_per_ray=round(counts[dom]["prims"] / counts[dom]["rays"], 2),
                  step_gbs=round(canon_step / (ms_per_step / 1e3) / 1e9, 1))
    res = {
        "metric": METRIC, "value": round(solves_s, 4), "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 2), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": (f"config3 diurnal series: orchard 25 trees ({P} prims: 1,000,000 triangles + 1,600 "
                                f"ground patches), {K} suns x {B} bands (LW emission, NIR, PAR), d=n=15, depth 4, "
                                f"batched (one series per step)"),
                   "solves_per_step": K, "prims": P, "bands": B, "suns": K},
        "mrays_per_s_traced": round(traced_rays / (ms_per_step / 1e3) / 1e6, 1),
        "mrays_per_s_reference_equivalent": round(rp["total"] * solves_s / 1e6, 1),
        "plain_solves_per_s": round(plain_solves_s, 4),
        "plain_mrays_per_s": round(rp["total"] * plain_solves_s / 1e6, 1),
        "speedup_batched_vs_plain": round(solves_s / plain_solves_s, 2),
        "latency_ms": {"series_p50": pct(lat, 50), "series_p99": pct(lat, 99),
                       "plain_solve_p50": pct(plain_lat, 50), "plain_solve_p99": pct(plain_lat, 99)},
        "with_host_download_ms": {"series": round(series_host_ms, 1), "plain_solve": round(plain_host_ms, 1)},
        "kernel_ms_per_step": {("grid_build" if k == "other" else k): round(v, 2) for k, v in per_step.items()},
        "direct_occlusion": "grid" if grid else "bvh4-packet",
        "grid_entries_per_prim": round(grid_entries, 2) if grid else None,
        "direct_bytes_per_ray": round(bytes_per_ray(counts["direct"], grid=grid), 1),
        "direct_candidates_per_ray": round(counts["direct"]["prims"] / counts["direct"]["rays"], 2),
        "rays_per_solve": rp["total"],
        "scene_build_ms": round(build_s * 1e3, 1),
        "autotune": tune,
        "roofline": rl,
    }
    if not args.no_energy:
        log("[bench] energy bookkeeping (both schedules) ...")
        res["energy_fp64"] = energy_diurnal(ctx, g, base, dirs, flux, sky, seeds, out_dev.value)
    log("[bench] watertight vs reference predicates ...")
    res["predicates"] = predicates_diurnal(ctx, g, base, dirs, flux, sky, seeds)
    hip.hipFree(out_dev)
    from oracle.oracle import OracleScene
    o = OracleScene.from_arrays(s)
    if not args.no_cpu_baseline:
        log("[bench] cpu baseline (oracle) ...")
        res["cpu_baseline"], smp = cpu_baseline_diurnal(o, s, base, dirs[0][0], seeds[0], args.cpu_seconds)
    else:
        smp = oracle_samples(o, s, base, dirs[0][0], seeds[0], 400)
    log("[bench] parity sample ...")
    res["parity"] = parity_diurnal(g, base, dirs[0][0], seeds[0], smp)
    return res


def bench_config2(args):
    if args.calib_copy_mib:
        calib_copy(args.calib_copy_mib)
    s = scenes.config2(2)
    ctx = native.Context(0)
    g = ctx.scene(s.transforms, s.types, s.twosided, periodic=s.periodic)
    req = W.config2_request(s)
    rp = W.rays_per_solve(s, req)
    ctx.enable_timing(True)
    for _ in range(args.warmup):
        native.run_band(g, req)
    ctx.set_counting(True)
    native.run_band(g, req)
    counts = ctx.counters()
    ctx.set_counting(False)
    lat = []
    kt = {"direct_ms": 0.0, "diffuse_ms": 0.0, "gather_ms": 0.0, "diffuse_launches": 0, "direct_launches": 0}
    device_sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        t = time.perf_counter()
        native.run_band(g, req)
        lat.append((time.perf_counter() - t) * 1e3)
        tm = ctx.timing()
        for key in kt:
            kt[key] += tm[key]
    device_sync()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    dom = "diffuse"
    per = kt["diffuse_ms"] / args.steps / max(kt["diffuse_launches"] / args.steps, 1)
    brays = bytes_per_ray(counts[dom])
    rpl = counts[dom]["rays"] / max(kt["diffuse_launches"] / args.steps, 1)
    return {"metric": METRIC, "value": round(1e3 / ms, 4), "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": round(ms, 2), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": "config2: 200k leaf triangles + 400 ground patches, periodic, NIR+PAR, d=n=32, depth 5"},
            "mrays_per_s": round(rp["total"] / (ms / 1e3) / 1e6, 1),
            "latency_ms": {"p50": pct(lat, 50), "p99": pct(lat, 99)},
            "roofline": roofline("k_diffuse", per, brays * rpl, "pmc_traffic_config2.json", bytes_per_ray=round(brays, 1))}


# published runBand wall times of the homogeneous-canopy benchmark (seconds: direct, diffuse)
PUBLISHED_HOMOGENEOUS = {"NVIDIA RTX 6000 Ada": (1.054, 1.910), "NVIDIA L40S": (0.950, 1.821),
                         "NVIDIA A100 80GB PCIe": (1.727, 14.186)}


def bench_homogeneous(args):
    """The reference's published radiation benchmark (benchmarks/radiation_homogeneous_canopy): LAI 1 over a
    20 x 20 m periodic domain of 5 cm disk-masked tile leaves (2 x 2 sub-patches) + a 100 x 100 ground tile,
    1024 direct rays (band "direct", sun at 36 degrees zenith) and 1024 diffuse rays (band "diffuse"),
    scattering depth 0, one band per runBand.  One step = runBand("direct") + runBand("diffuse"), each a
    full rt_run_band call with host inputs and outputs, as the harness times runBand."""
    if args.calib_copy_mib:
        calib_copy(args.calib_copy_mib)
    t0 = time.perf_counter()
    s, masks, mid, uv, uvid = scenes.homogeneous_canopy(1)
    gen_s = time.perf_counter() - t0
    ctx = native.Context(0)
    # the geometry update (BVH build + mask upload) timed on its second run: the first also loads the
    # kernels' code objects
    for _ in range(2):
        t0 = time.perf_counter()
        g = ctx.scene(s.transforms, s.types, s.twosided, periodic=s.periodic)
        g.set_masks(masks, mid, uv, uvid)
        device_sync()
        geom_s = time.perf_counter() - t0
    P = s.P
    zero = np.zeros((P, 1), np.float32)
    th = 0.2 * math.pi
    sun = native.sun_direction(math.degrees(th), 0.0)
    direct = native.BandRequest(P, [native.BandSpec("direct", 0, False, 0.0)], rho=zero, tau=zero, direct_rays=1024,
                                diffuse_rays=1000, source_dirs=[sun], source_flux=[[1.0 / math.cos(th)]],
                                direct_seed=3, diffuse_seeds=[5])
    diffuse = native.BandRequest(P, [native.BandSpec("diffuse", 0, False, 1.0)], rho=zero, tau=zero,
                                 direct_rays=100, diffuse_rays=1024, source_dirs=[sun], source_flux=[[0.0]],
                                 direct_seed=7, diffuse_seeds=[9])
    # default launch shapes (the autotuner picks the same ones for this scene)
    ctx.enable_timing(True)
    for _ in range(args.warmup):
        native.run_band(g, direct)
        native.run_band(g, diffuse)
    ctx.set_counting(True)
    native.run_band(g, diffuse)
    counts = ctx.counters()
    ctx.set_counting(False)
    td, tf, kms = [], [], 0.0
    device_sync()
    t_all = time.perf_counter()
    for _ in range(args.steps):
        t = time.perf_counter()
        Rd = native.run_band(g, direct)
        td.append(time.perf_counter() - t)
        t = time.perf_counter()
        Rf = native.run_band(g, diffuse)
        tf.append(time.perf_counter() - t)
        kms += ctx.timing()["diffuse_ms"]
    device_sync()
    ms = (time.perf_counter() - t_all) * 1e3 / args.steps
    d_s, f_s = float(np.mean(td)), float(np.mean(tf))
    nl = int((s.twosided == 1).sum())
    direct_rays = P * 1024
    diffuse_rays = (2 * nl + (P - nl)) * 1024
    brays = bytes_per_ray(counts["diffuse"])
    ref = PUBLISHED_HOMOGENEOUS["NVIDIA RTX 6000 Ada"]
    return {"metric": "runBand seconds (direct + diffuse), homogeneous canopy benchmark", "value": round(d_s + f_s, 4),
            "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 2),
            "higher_is_better": False, "scaling": "weak", "vs_baseline": round((d_s + f_s) / sum(ref), 4),
            "dtype": "f32", "data": "synthetic",
            "config": {"workload": f"radiation_homogeneous_canopy: {P} prims ({nl} masked leaf sub-patches, 800x800 "
                                   f"analytic disk mask), periodic 20x20 m, 1024 direct + 1024 diffuse rays, depth 0",
                       "direct_s": round(d_s, 4), "diffuse_s": round(f_s, 4), "geometry_update_s": round(geom_s, 3),
                       "scene_generation_s": round(gen_s, 2),
                       "direct_grays_per_s": round(direct_rays / d_s / 1e9, 3),
                       "diffuse_grays_per_s": round(diffuse_rays / f_s / 1e9, 3),
                       "published_s": {k: {"direct": v[0], "diffuse": v[1]} for k, v in PUBLISHED_HOMOGENEOUS.items()},
                       "absorbed_mean_direct": round(float(Rd.mean()), 5), "absorbed_mean_diffuse": round(float(Rf.mean()), 5)},
            "roofline": roofline("k_diffuse", kms / args.steps, brays * counts["diffuse"]["rays"],
                                 "pmc_traffic_homogeneous_canopy.json", flags=1,  # texture-masked leaves
                                 bytes_per_ray=round(brays, 1))}


def bench_castrays(args):
    import ctypes as C
    if args.calib_copy_mib:
        calib_copy(args.calib_copy_mib)
    s = scenes.orchard(3, tiles=5, jitter=0.05)
    ctx = native.Context(0)
    g = ctx.scene(s.transforms, s.types, s.twosided)
    n = args.rays
    T = s.transforms.reshape(-1, 16)
    lo = [float(T[:, 3].min()), float(T[:, 7].min()), 0.0]
    hi = [float(T[:, 3].max()), float(T[:, 7].max()), 3.6]
    O, D = scenes.rays(11, n, lo, hi)
    hip = C.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    bufs = [C.c_void_p() for _ in range(5)]
    sizes = [n * 12, n * 12, n * 4, n * 12, n * 4]
    for b, sz in zip(bufs, sizes):
        assert hip.hipMalloc(C.byref(b), sz) == 0
    hip.hipMemcpy(bufs[0], O.ctypes.data, n * 12, 1)
    hip.hipMemcpy(bufs[1], D.ctypes.data, n * 12, 1)
    ctx.enable_timing(True)
    for _ in range(args.warmup):
        g.cast_rays_device(bufs[0].value, bufs[1].value, n, -1.0, bufs[2].value, bufs[3].value, bufs[4].value)
    ctx.set_counting(True)
    g.cast_rays_device(bufs[0].value, bufs[1].value, n, -1.0, bufs[2].value, bufs[3].value, bufs[4].value)
    counts = ctx.counters()["cast"]
    ctx.set_counting(False)
    device_sync()
    t0 = time.perf_counter()
    kms = 0.0
    for _ in range(args.steps):
        g.cast_rays_device(bufs[0].value, bufs[1].value, n, -1.0, bufs[2].value, bufs[3].value, bufs[4].value)
        kms += ctx.timing()["direct_ms"]
    device_sync()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    brays = bytes_per_ray(counts)
    return {"metric": "castRays Mrays/s", "value": round(n / (ms / 1e3) / 1e6, 1), "unit": "Mrays/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 2), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": f"config5 castRaysSoA: {n} rays (uniform in scene AABB, uniform directions) over "
                                   f"the 5x tiled orchard ({s.P} prims)"},
            "roofline": roofline("k_cast", kms / args.steps, brays * n, "pmc_traffic_config5_castrays.json",
                                 bytes_per_ray=round(brays, 1), nodes_per_ray=round(counts["nodes"] / counts["rays"], 2),
                                 prims_per_ray=round(counts["prims"] / counts["rays"], 2),
                                 result_bytes_per_launch=20 * n),
            **({} if args.no_cpu_baseline else {"cpu_baseline": cpu_baseline_cast(s, O, D, args.cpu_seconds)})}


def cpu_baseline_cast(s, O, D, seconds):
    """The oracle's castRaysSoA restatement (OpenMP, CD predicates) on leading chunks of the same ray
    stream until about `seconds`."""
    from oracle.oracle import OracleScene
    o = OracleScene.from_arrays(s)
    n, chunk, t0 = 0, 1 << 20, time.perf_counter()
    while time.perf_counter() - t0 < seconds and n < O.shape[0]:
        m = min(chunk, O.shape[0] - n)
        o.cast_rays(O[n:n + m], D[n:n + m])
        n += m
    dt = time.perf_counter() - t0
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": round(n / dt / 1e6, 3), "unit": "Mrays/s", "cores": cores, "kind": "port",
            "sample": f"first {n} rays of the same stream, oracle castRaysSoA restatement (OpenMP), {dt:.1f} s"}


def orchard_camera(s, res, spp):
    """A camera south of the orchard (at 30 % of its depth beyond the first row), 8 m up, looking north
    and 25 degrees down with a 60-degree horizontal field of view."""
    T = s.transforms.reshape(-1, 16)
    x0, x1, y0, y1 = float(T[:, 3].min()), float(T[:, 3].max()), float(T[:, 7].min()), float(T[:, 7].max())
    return native.Camera(((x0 + x1) / 2, y0 - 0.3 * (y1 - y0), 8.0), (math.radians(115.0), 0.0), (res, res),
                         viewplane_length=0.5 / math.tan(math.radians(30.0)), antialiasing_samples=spp)


def bench_camera(args):
    """Config-5 camera: a 1024 x 1024 image at 64 samples per pixel of the 5x tiled orchard, 3 bands,
    one collimated sun with its solar disk, a Prague sky."""
    if args.calib_copy_mib:
        calib_copy(args.calib_copy_mib)
    s = scenes.orchard(3, tiles=5, jitter=0.05)
    ctx = native.Context(0)
    be = native.Backend(ctx)
    B = 3
    be.update_geometry(s.transforms, s.types, s.twosided)
    be.build_acceleration_structure()
    be.update_materials(np.full((s.P, B), 0.1, np.float32), np.full((s.P, B), 0.05, np.float32))
    sun = native.sun_direction(40.0, 150.0)
    be.update_sources([sun])
    be.zero_radiation_buffers(B)
    rng = np.random.default_rng(7)
    # radiation_out of the last scattering pass: synthetic values (the render's work does not depend on them)
    be.upload_radiation_out(rng.uniform(0, 300, (s.P, B)).astype(np.float32),
                            rng.uniform(0, 100, (s.P, B)).astype(np.float32))
    be.upload_source_fluxes(np.array([[400.0, 300.0, 500.0]], np.float32))
    be.update_sky_model(np.tile(np.array([2.0, 30.0, 1.5, 0.25], np.float32), (B, 1)),
                        np.array([20.0, 15.0, 30.0], np.float32), sun, np.array([4e3, 3e3, 5e3], np.float32),
                        math.cos(math.radians(0.27)), np.zeros(B, np.float32), np.zeros(B, np.uint32))
    cam = orchard_camera(s, args.cam_res, args.cam_spp)
    nrays = args.cam_res * args.cam_res * args.cam_spp
    ctx.enable_timing(True)
    for _ in range(args.warmup):
        be.launch_camera_rays(cam, seed=1)
    ctx.set_counting(True)
    be.launch_camera_rays(cam, seed=1)
    counts = ctx.counters()["camera"]
    ctx.set_counting(False)
    device_sync()
    t0 = time.perf_counter()
    kms = 0.0
    for k in range(args.steps):
        be.launch_camera_rays(cam, seed=2 + k)
        kms += ctx.timing()["direct_ms"]
    device_sync()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    # pixel labels of the same view
    be.zero_camera_pixel_buffers(cam.resolution)
    t1 = time.perf_counter()
    be.launch_pixel_label_rays(cam)
    lab_ms = (time.perf_counter() - t1) * 1e3
    R = be.get_camera_results(cam.resolution)
    hit = float((R["pixel_labels"] > 0).mean())
    brays = bytes_per_ray(counts)
    res = {"metric": "camera Mrays/s", "value": round(nrays / (ms / 1e3) / 1e6, 1), "unit": "Mrays/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 2), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": f"config5 camera: {args.cam_res}x{args.cam_res} pixels x {args.cam_spp} samples, "
                                  f"{B} bands, over the 5x tiled orchard ({s.P} prims); one step = one image",
                      "pixels_hit": round(hit, 3), "pixel_label_ms": round(lab_ms, 2)},
           "roofline": roofline("k_camera", kms / args.steps, brays * nrays, "pmc_traffic_config5_camera.json",
                                bytes_per_ray=round(brays, 1), nodes_per_ray=round(counts["nodes"] / counts["rays"], 2),
                                prims_per_ray=round(counts["prims"] / counts["rays"], 2))}
    if not args.no_cpu_baseline:
        res["cpu_baseline"] = cpu_baseline_camera(s, cam, args.cpu_seconds)
    return res


def cpu_baseline_camera(s, cam, seconds):
    """The oracle's camera restatement (one thread) on rows of the same image until about `seconds`."""
    from oracle.oracle import OracleScene
    o = OracleScene.from_arrays(s)
    P, B = s.P, 3
    top = np.full((P, B), 100.0, np.float32)
    sky = dict(params=np.tile(np.array([2.0, 30.0, 1.5, 0.25], np.float32), (B, 1)),
               radiance=np.full(B, 20.0, np.float32), sun=(0.0, 0.0, 1.0), solar=np.zeros(B, np.float32), solar_cos=2.0)
    fx, fy = cam.resolution
    rows, t0 = 0, time.perf_counter()
    out = np.zeros((B, fx * fy))
    while time.perf_counter() - t0 < seconds and rows < fy:
        o.camera_pixels(cam, top, top, np.zeros((0, B)), np.zeros((0, 3)), sky, seed=1, tile=(0, rows, fx, 8),
                        pixels=out)
        rows += 8
    dt = time.perf_counter() - t0
    n = fx * rows * cam.antialiasing_samples
    return {"value": round(n / dt / 1e6, 4), "unit": "Mrays/s", "cores": 1, "kind": "port",
            "sample": f"{rows} image rows x {fx} pixels x {cam.antialiasing_samples} samples ({n} rays) of the same "
                      f"view, oracle camera restatement, {dt:.1f} s"}


def bench_serve(args, ctx=None):
    """Config 4: a Poisson stream of single-sun requests over seeded canopies (P log-uniform 10k-2M,
    1-4 bands), served by the request server (resident models, memory-footprint admission, continuous
    batching over `slots` streams) and, on the same stream, by the plain loop (one rt_run_band per request
    in arrival order).  The arrival rate is `serve_load` x the plain loop's capacity."""
    import threading
    ctx = ctx or native.Context(0)
    models = W.serve_models(seed=4, nscenes=args.serve_scenes)
    log(f"[serve] model sizes {[s.P for s, _ in models]}, bands {[r.B for _, r in models]}")
    scs = [ctx.scene(s.transforms, s.types, s.twosided) for s, _ in models]
    reqs = W.serve_requests(models, args.serve_requests, seed=5)
    single = [W.single_sun_request(models[m][1], d, f, q, seed) for (m, d, f, q, seed) in reqs]
    # plain service time per model (one warm solve each) -> arrival rate
    t_model = []
    for (s, req), g in zip(models, scs):
        native.run_band(g, req)
        t = time.perf_counter()
        native.run_band(g, req)
        t_model.append(time.perf_counter() - t)
    mean_service = float(np.mean([t_model[m] for (m, *_r) in reqs]))
    rate = args.serve_load / mean_service
    rng = np.random.default_rng(6)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, len(reqs)))
    arrivals -= arrivals[0]
    log(f"[serve] {len(reqs)} requests, mean plain service {mean_service * 1e3:.1f} ms, rate {rate:.2f}/s")

    # plain loop: one solve at a time, in arrival order
    plain_out, plain_lat = [], []
    t0 = time.perf_counter()
    for k, (m, *_r) in enumerate(reqs):
        now = time.perf_counter() - t0
        if now < arrivals[k]:
            time.sleep(arrivals[k] - now)
        plain_out.append(native.run_band(scs[m], single[k]))
        plain_lat.append((time.perf_counter() - t0 - arrivals[k]) * 1e3)
    plain_span = time.perf_counter() - t0
    log(f"[serve] plain loop {plain_span:.2f} s")

    srv = native.Server(ctx, slots=args.serve_slots, max_batch_suns=32)
    ids = [srv.add_model(g, req) for g, (_, req) in zip(scs, models)]
    tickets = [None] * len(reqs)

    def producer():
        t1 = time.perf_counter()
        for k, (m, d, f, q, seed) in enumerate(reqs):
            now = time.perf_counter() - t1
            if now < arrivals[k]:
                time.sleep(arrivals[k] - now)
            tickets[k] = srv.submit(ids[m], d, f, q, seed)

    t0 = time.perf_counter()
    th = threading.Thread(target=producer)
    th.start()
    th.join()
    worst = 0.0
    for k in range(len(reqs)):
        R, _ = srv.wait(tickets[k])
        a = plain_out[k].astype(np.float64)
        worst = max(worst, float(np.max(np.abs(R - a) / np.maximum(np.abs(a), 1.0))))
    served_span = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    log(f"[serve] server {served_span:.2f} s, stats {st}")
    return {
        "workload": (f"config4: {len(reqs)} Poisson requests at {rate:.2f}/s ({args.serve_load}x plain capacity) over "
                     f"{len(models)} random canopies (P = {sorted(s.P for s, _ in models)}, 1-4 bands each), "
                     f"d=n=10, depth 2, one sun per request"),
        "solves_per_s": round(len(reqs) / served_span, 3),
        "p50_ms": round(st["p50_ms"], 1), "p99_ms": round(st["p99_ms"], 1), "mean_ms": round(st["mean_ms"], 1),
        "plain_solves_per_s": round(len(reqs) / plain_span, 3),
        "plain_p50_ms": round(pct(plain_lat, 50), 1), "plain_p99_ms": round(pct(plain_lat, 99), 1),
        "speedup_vs_plain": round(plain_span / served_span, 2),
        "slots": args.serve_slots, "batches": st["batches"], "max_batch": st["max_batch"],
        "memory_budget_bytes": st["memory_budget"], "peak_workspace_bytes": st["peak_workspace_bytes"],
        "resident_model_bytes": st["resident_model_bytes"],
        "max_rel_err_vs_plain": worst,
    }


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="diurnal",
                    choices=["diurnal", "config2", "castrays", "camera", "homogeneous", "serve"])
    ap.add_argument("--cam-res", type=int, default=1024)
    ap.add_argument("--cam-spp", type=int, default=64)
    ap.add_argument("--suns", type=int, default=64)
    ap.add_argument("--plain-suns", type=int, default=4)
    ap.add_argument("--rays", type=int, default=100_000_000)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-serve", action="store_true", help="skip the config-4 serving leg of the default run")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch shapes")
    ap.add_argument("--no-energy", action="store_true", help="skip the energy-bookkeeping series")
    ap.add_argument("--serve-requests", type=int, default=64)
    ap.add_argument("--serve-scenes", type=int, default=6)
    ap.add_argument("--serve-slots", type=int, default=4)
    ap.add_argument("--serve-load", type=float, default=2.0)
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="run a plain copy of this many MiB first (PMC byte-counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}): the solver shards by primitives or "
              f"suns across GPUs only on paper, and no multi-GPU path exists to measure", file=sys.stderr)
        sys.exit(2)
    if args.workload == "diurnal":
        res = bench_diurnal(args)
        if not args.no_serve:
            log("[bench] config-4 serving leg ...")
            res["serving"] = bench_serve(args)
    elif args.workload == "serve":
        res = bench_serve(args)
        res = {"metric": "served solves/s with p50/p99 latency (config 4)", "value": res["solves_per_s"],
               "unit": "solves/s", "n_gpus": 1, "steps": 1, "warmup": 1, "higher_is_better": True, "scaling": "weak",
               "vs_baseline": None, "dtype": "f32", "data": "synthetic", "config": {"workload": res["workload"]},
               "serving": res}
    elif args.workload == "config2":
        res = bench_config2(args)
    elif args.workload == "camera":
        res = bench_camera(args)
    elif args.workload == "homogeneous":
        res = bench_homogeneous(args)
    else:
        res = bench_castrays(args)
    print(json.dumps(res))


if __name__ == "__main__":
    main()
