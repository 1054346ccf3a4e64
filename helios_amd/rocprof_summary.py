"""Summarise rocprofv3 SQLite results (``*_results.db``) into short text: per-kernel timing and,
for PMC runs, per-dispatch counter values of the ray kernels and of the byte-counter calibration copy.

usage: python -m helios_amd.rocprof_summary KERNEL_TRACE_DB [PMC_DB ...] [--json OUT.json] [--brief]

With --json, also writes HBM bytes per dispatch of each ray kernel: FETCH_SIZE and WRITE_SIZE
(KiB) scaled by what the same counters read for the calibration copy of known size.  A PMC run with
TCC_HIT_sum / TCC_MISS_sum adds each ray kernel's L2 hit rate (its largest dispatch); other counters (SQ_*
and the like) are listed for each ray kernel's largest dispatch.  --brief leaves out the per-dispatch
listing of FETCH_SIZE / WRITE_SIZE.
"""
from __future__ import annotations

import collections
import json
import re
import sqlite3
import sys

RAY_KERNELS = ("k_direct", "k_diffuse", "k_gather", "k_cast", "k_camera", "k_pixel_label")


def short(name: str) -> str:
    return re.sub(r"\(.*", "", name.replace("(anonymous namespace)::", ""))[:90]


def kernel_rows(db):
    cur = sqlite3.connect(db).cursor()
    return list(cur.execute("select name, start, end, grid_x, workgroup_x, vgpr_count, sgpr_count, lds_size "
                            "from kernels order by start"))


def counter_rows(db):
    cur = sqlite3.connect(db).cursor()
    return list(cur.execute("select dispatch_id, kernel_name, counter_name, value from counters_collection "
                            "order by dispatch_id"))


def summarize(kt_db, pmc_dbs, calib_bytes=1 << 30, brief=False):
    out = []
    rows = kernel_rows(kt_db)
    agg = collections.OrderedDict()
    res = {}
    for name, s, e, g, w, v, sg, lds in rows:
        agg.setdefault(short(name), []).append((e - s) / 1e6)
        if any(t in name for t in RAY_KERNELS):
            res[short(name)] = (g, w, v, sg, lds)
    out.append("== kernel trace (durations in ms)")
    out.append(f"{'kernel':90s} {'calls':>6s} {'total_ms':>12s} {'avg_ms':>12s} {'min_ms':>10s} {'max_ms':>10s}")
    for k, v in sorted(agg.items(), key=lambda kv: -sum(kv[1])):
        out.append(f"{k:90s} {len(v):6d} {sum(v):12.3f} {sum(v) / len(v):12.3f} {min(v):10.3f} {max(v):10.3f}")
    out.append("\n== per-dispatch durations of the ray kernels (ms, launch order)")
    for k, v in agg.items():
        if any(t in k for t in RAY_KERNELS):
            out.append(f"{k}: " + ", ".join(f"{x:.3f}" for x in v))
    for k, (g, w, v, sg, lds) in res.items():
        out.append(f"resources {k}: grid {g} workgroup {w} vgpr {v} sgpr {sg} lds {lds}")
    traffic = collections.defaultdict(lambda: collections.defaultdict(list))
    scale = {}
    l2 = collections.defaultdict(dict)  # kernel -> dispatch -> {counter: value}
    other = collections.defaultdict(dict)  # kernel -> dispatch -> {counter: value} (SQ_* and the like)
    for db in pmc_dbs:
        rows = counter_rows(db)
        if not brief:
            out.append(f"\n== {db.split('/')[-1]}: counter value per dispatch (KiB)")
        for d, n, c, val in rows:
            sn = short(n)
            if c.startswith("TCC_HIT") or c.startswith("TCC_MISS"):
                if any(t in sn for t in RAY_KERNELS):
                    l2[sn].setdefault(d, {})[c] = val
                continue
            if c not in ("FETCH_SIZE", "WRITE_SIZE"):
                if any(t in sn for t in RAY_KERNELS):
                    cs = other[sn].setdefault(d, {})
                    cs[c] = cs.get(c, 0.0) + val
                continue
            if "copyBuffer" in sn and val >= calib_bytes / 1024 / 4:
                scale[c] = calib_bytes / 1024.0 / val
                out.append(f"dispatch {d:6d} {sn:60s} {c:12s} {val:16.1f}  <- calibration copy of {calib_bytes} B")
            elif any(t in sn for t in RAY_KERNELS):
                traffic[sn][c].append(val)
                if not brief:
                    out.append(f"dispatch {d:6d} {sn:60s} {c:12s} {val:16.1f}")
    l2_rate = {}
    if l2:
        out.append("\n== L2 (TCC) hit rate of each ray kernel's largest dispatch")
        for k, ds in l2.items():
            d, cs = max(ds.items(), key=lambda kv: sum(kv[1].values()))
            hit = sum(v for c, v in cs.items() if c.startswith("TCC_HIT"))
            miss = sum(v for c, v in cs.items() if c.startswith("TCC_MISS"))
            if hit + miss > 0:
                l2_rate[k] = hit / (hit + miss)
                out.append(f"{k}: dispatch {d} hits {hit:.3e} misses {miss:.3e} hit rate {hit / (hit + miss):.3f}")
    if other:
        out.append("\n== other counters of each ray kernel's largest dispatch (raw values)")
        for k, ds in other.items():
            d, cs = max(ds.items(), key=lambda kv: sum(kv[1].values()))
            out.append(f"{k}: dispatch {d}: " + ", ".join(f"{c} {v:.4g}" for c, v in sorted(cs.items())))
    summary = {}
    if scale:
        out.append("\n== HBM bytes per dispatch (counter KiB x 1024 x calibration factor)")
        out.append("calibration factors: " + ", ".join(f"{c} x{f:.4f}" for c, f in scale.items()))
        for k, cs in traffic.items():
            tot = {}
            for c, vals in cs.items():
                f = scale.get(c, 1.0)
                tot[c] = [v * 1024.0 * f for v in vals]
            out.append(f"{k}: " + "; ".join(f"{c} max {max(v) / 1e9:.2f} GB min {min(v) / 1e9:.2f} GB"
                                            for c, v in tot.items()))
            summary[k] = {c: {"max_bytes": max(v), "min_bytes": min(v)} for c, v in tot.items()}
            if k in l2_rate:
                summary[k]["l2_hit_rate"] = round(l2_rate[k], 4)
    return "\n".join(out), summary, scale


def main(argv):
    js = None
    if "--json" in argv:
        i = argv.index("--json")
        js = argv[i + 1]
        argv = argv[:i] + argv[i + 2:]
    brief = "--brief" in argv
    argv = [a for a in argv if a != "--brief"]
    text, summary, scale = summarize(argv[0], argv[1:], brief=brief)
    print(text)
    if js:
        with open(js, "w") as f:
            json.dump({"calibration_factors": scale, "kernels": summary}, f, indent=1)


if __name__ == "__main__":
    main(sys.argv[1:])
