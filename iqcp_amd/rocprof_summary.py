"""Reduce rocprofv3 CSV output of bench.py runs to the short summaries kept under profiles/.

    python -m iqcp_amd.rocprof_summary stats    --csv KERNEL_STATS.csv --md OUT.md [--command CMD ...]
    python -m iqcp_amd.rocprof_summary counters --pmc DIR [--pmc DIR ...] --stats KERNEL_STATS.csv --md OUT.md
    python -m iqcp_amd.rocprof_summary traffic  --fetch DIR --write DIR --copy-fetch DIR --copy-write DIR
                                                --tiles TILES.json --json OUT.json --md OUT.md
    python -m iqcp_amd.rocprof_summary mfma     --pmc DIR [--pmc DIR ...] --stats KERNEL_STATS.csv
                                                --match SUBSTRING [--match ...] --md OUT.md

`stats` tabulates a `rocprofv3 --kernel-trace --stats` kernel_stats.csv.  `counters` averages the
counters of one or more `rocprofv3 --pmc` runs per dispatch of each J/K kernel and derives wave
states (fractions of SQ_WAVE_CYCLES), VALU issue rate and counted FP64 FLOP/s against the average
durations of a kernel-trace run.  `traffic` turns FETCH_SIZE / WRITE_SIZE runs into HBM bytes per
launch, calibrated by the same counters on `bench.py --copy-probe` (a 1 GiB device copy), and writes
the JSON bench.py reads for its roofline entry.  `mfma` does the same per dispatch for the kernels
whose names contain a --match string (the XC grid products): wave states, MFMA instructions, MFMA busy
fraction and the MFMA FLOP/s those instructions amount to.  DIR is a rocprofv3 output directory
(searched for *counter_collection.csv); each --command is copied into the summary verbatim.
"""
from __future__ import annotations

import argparse
import collections
import csv
import glob
import json
import re
from pathlib import Path

# MI355X: 256 CUs x 4 SIMDs, 2.4 GHz peak engine clock; a wave64 VALU instruction occupies a SIMD for
# 4 cycles; FP64 vector peak 78.6 TFLOP/s (= 1024 SIMDs x 16 lanes x 2 flops x 2.4 GHz)
SIMDS, CLOCK_HZ, CYCLES_PER_VALU, PEAK_F64 = 1024, 2.4e9, 4, 78.6e12
# dense MFMA peaks (FP32 157.3, FP64 78.6 TFLOP/s); a 16x16x4 MFMA instruction is 2048 flops
PEAK_MFMA = {"F32": 157.3e12, "F64": 78.6e12}
MFMA_FLOPS = 2 * 16 * 16 * 4

KERNEL_RE = re.compile(r"(jk_wave_kernel_mp|jk_kernel_mp|jk_wave_kernel|jk_kernel)<(\d), (\d), (\d), (\d)"
                       r"(?:, (double|float))?(?:, (\d))?>")


def short_name(name: str) -> str:
    """Kernel name without its argument list, 'void ' and the qcg:: namespace."""
    return re.sub(r"\(.*", "", name).replace("void ", "").replace("qcg::", "").strip()


def bench_label(name: str) -> str | None:
    """The name bench.py gives a J/K kernel (no LDS-row template argument), None for other kernels."""
    m = KERNEL_RE.search(name)
    if not m:
        return None
    kind, la, lb, lc, ld, typ, _ = m.groups()
    if kind.endswith("_mp"):
        return f"{kind}<{la}, {lb}, {lc}, {ld}>"
    return f"{kind}<{la}, {lb}, {lc}, {ld}, {typ}>"


def counter_csv(d: str) -> str:
    hits = glob.glob(str(Path(d) / "**" / "*counter_collection.csv"), recursive=True)
    if not hits:
        raise SystemExit(f"no *counter_collection.csv under {d}")
    return hits[0]


def load_counters(d: str):
    """{kernel: {counter: value summed per dispatch, averaged over dispatches}}, dispatch counts, and
    the VGPR count and LDS bytes of each kernel."""
    per = collections.defaultdict(float)
    disp = collections.defaultdict(set)
    res = {}
    for r in csv.DictReader(open(counter_csv(d))):
        k = short_name(r["Kernel_Name"])
        per[(k, r["Counter_Name"])] += float(r["Counter_Value"])
        disp[k].add(r["Dispatch_Id"])
        res[k] = (int(float(r.get("VGPR_Count", 0) or 0)), int(float(r.get("LDS_Block_Size", 0) or 0)))
    avg = collections.defaultdict(dict)
    for (k, c), v in per.items():
        avg[k][c] = v / len(disp[k])
    return avg, {k: len(v) for k, v in disp.items()}, res


def load_stats(path: str):
    """{kernel: (calls, average µs, min µs, max µs, percent)} from a kernel_stats.csv."""
    out = {}
    for r in csv.DictReader(open(path)):
        k = short_name(r["Name"])
        out[k] = (int(r["Calls"]), float(r["AverageNs"]) / 1e3, float(r["MinNs"]) / 1e3,
                  float(r["MaxNs"]) / 1e3, float(r["Percentage"]))
    return out


def header(title: str, commands: list[str]) -> str:
    s = f"# {title}\n\n"
    if commands:
        s += "Commands (run from /tmp on the GPU box with TMPDIR=/tmp; <out> is any scratch directory):\n\n```\n"
        s += "\n".join(commands) + "\n```\n\n"
    return s


def cmd_stats(a):
    st = load_stats(a.csv)
    tot_ms = sum(c * avg for c, avg, *_ in st.values()) / 1e3
    calls = sum(v[0] for v in st.values())
    rows = sorted(st.items(), key=lambda kv: -kv[1][0] * kv[1][1])[: a.top]
    s = header(a.title, a.command)
    s += f"Total GPU kernel time {tot_ms:.2f} ms over {calls} dispatches; the {len(rows)} largest kernels:\n\n"
    s += "| kernel | calls | avg µs | min µs | max µs | % of kernel time |\n|---|---:|---:|---:|---:|---:|\n"
    for k, (c, avg, mn, mx, pct) in rows:
        s += f"| `{k}` | {c} | {avg:.1f} | {mn:.1f} | {mx:.1f} | {pct:.2f} |\n"
    Path(a.md).write_text(s + a.note)


def cmd_counters(a):
    cnt, ndisp, res = {}, {}, {}
    for d in a.pmc:
        c, n, r = load_counters(d)
        for k, v in c.items():
            cnt.setdefault(k, {}).update(v)
            ndisp[k] = n[k]
            res[k] = r[k]
    st = load_stats(a.stats) if a.stats else {}
    keys = [k for k in cnt if bench_label(k)]
    keys.sort(key=lambda k: -(st.get(k, (0, 0))[0] * st.get(k, (0, 0))[1]) if st else -cnt[k].get("SQ_WAVES", 0))
    have = set().union(*(cnt[k].keys() for k in keys)) if keys else set()
    states = {"SQ_WAVE_CYCLES", "SQ_WAIT_ANY", "SQ_WAIT_INST_ANY", "SQ_ACTIVE_INST_VALU"} <= have
    f64 = {"SQ_INSTS_VALU_FMA_F64", "SQ_INSTS_VALU_MUL_F64", "SQ_INSTS_VALU_ADD_F64"} <= have
    s = header(a.title, a.command)
    cols = ["kernel", "dispatches", "waves / dispatch", "VGPR_Count", "LDS_Block_Size", "avg µs"]
    if states:
        s += ("Wave states are fractions of SQ_WAVE_CYCLES (wave lifetime) summed over each dispatch: waiting on a "
              "counter (SQ_WAIT_ANY: s_waitcnt on vector memory, LDS or scalar loads), waiting to issue "
              "(SQ_WAIT_INST_ANY), issuing VALU / scalar / LDS instructions (SQ_ACTIVE_INST_*).  VALU issue rate = "
              f"SQ_INSTS_VALU x {CYCLES_PER_VALU} cycles / ({SIMDS} SIMDs x 2.4 GHz x avg duration).\n\n")
        cols += ["wait on counter", "wait to issue", "VALU", "scalar", "LDS", "VALU issue rate"]
    if f64:
        s += ("Counted FP64 FLOP/s = 64 lanes x (2 FMA + MUL + ADD) FP64 VALU instructions per dispatch "
              "(SQ_INSTS_VALU_*_F64; transcendental instructions not counted) / avg duration; FP64 peak "
              "78.6 TFLOP/s.\n\n")
        cols += ["FP64 instr / dispatch", "counted FP64 TFLOP/s", "of FP64 peak"]
    s += ("VGPR_Count and LDS_Block_Size are the fields of rocprofv3's dispatch records as written (LDS_Block_Size "
          "without the dynamic LDS of the launch).\n\n")
    if st:
        s += "Average durations come from the kernel-trace run (last command).\n\n"
    s += "| " + " | ".join(cols) + " |\n|---|" + "---:|" * (len(cols) - 1) + "\n"
    for k in keys:
        c = cnt[k]
        dur = st.get(k, (0, 0.0))[1] * 1e-6
        row = [f"`{k}`", str(ndisp[k]), f"{c.get('SQ_WAVES', 0):.0f}", str(res[k][0]), str(res[k][1]),
               f"{dur * 1e6:.1f}" if dur else "-"]
        if states:
            wc = c.get("SQ_WAVE_CYCLES", 0) or 1.0
            row += [f"{c.get(x, 0) / wc:.2f}" for x in ("SQ_WAIT_ANY", "SQ_WAIT_INST_ANY", "SQ_ACTIVE_INST_VALU",
                                                          "SQ_ACTIVE_INST_SCA", "SQ_ACTIVE_INST_LDS")]
            iss = c.get("SQ_INSTS_VALU", 0) * CYCLES_PER_VALU / (SIMDS * CLOCK_HZ * dur) if dur else 0
            row.append(f"{iss:.2f}" if dur and "SQ_INSTS_VALU" in c else "-")
        if f64:
            n = sum(c.get(x, 0) for x in ("SQ_INSTS_VALU_FMA_F64", "SQ_INSTS_VALU_MUL_F64", "SQ_INSTS_VALU_ADD_F64",
                                          "SQ_INSTS_VALU_TRANS_F64"))
            fl = 64 * (2 * c.get("SQ_INSTS_VALU_FMA_F64", 0) + c.get("SQ_INSTS_VALU_MUL_F64", 0) +
                       c.get("SQ_INSTS_VALU_ADD_F64", 0))
            row += [f"{n / 1e6:.2f} M", f"{fl / dur / 1e12:.2f}" if dur else "-",
                    f"{fl / dur / PEAK_F64:.3f}" if dur else "-"]
        s += "| " + " | ".join(row) + " |\n"
    Path(a.md).write_text(s + a.note)


def cmd_traffic(a):
    def per_kernel(d, counter):
        c, n, _ = load_counters(d)
        return {k: v[counter] for k, v in c.items() if counter in v}, n

    cf, _ = per_kernel(a.copy_fetch, "FETCH_SIZE")
    cw, _ = per_kernel(a.copy_write, "WRITE_SIZE")
    copyk = max((k for k in cf if "copy" in k.lower()), key=lambda k: cf[k])
    gib = float(1 << 30)
    ff, fw = gib / (cf[copyk] * 1024.0), gib / (cw[copyk] * 1024.0)
    F, nF = per_kernel(a.fetch, "FETCH_SIZE")
    W, _ = per_kernel(a.write, "WRITE_SIZE")
    tiles = json.loads(Path(a.tiles).read_text())
    order = [(cb, ck) for cb in range(5, -1, -1) for ck in range(cb, -1, -1)]
    pair_l = [(0, 0), (1, 0), (1, 1), (2, 0), (2, 1), (2, 2)]
    res = {}
    for k in F:
        m = KERNEL_RE.search(k)
        if not m:
            continue
        kind, la, lb, lc, ld, typ, rows = m.groups()
        ci = order.index((pair_l.index((int(la), int(lb))), pair_l.index((int(lc), int(ld)))))
        mixed = kind.endswith("_mp") or typ == "float"
        tile = tiles["mixed" if mixed else "fp64"][ci][1 if typ == "float" else 0]
        fb, wb = F[k] * 1024.0 * ff, W.get(k, 0.0) * 1024.0 * fw
        res[bench_label(k)] = {"bytes_per_launch": round(fb + wb), "fetch_bytes": round(fb), "write_bytes": round(wb),
                               "tile": tile, "dispatches": nF[k], "lds_rows": rows}
    Path(a.json).write_text(json.dumps({"calibration": {"copy_kernel": copyk[:80], "fetch_factor": ff,
                                                        "write_factor": fw}, "kernels": res}, indent=1))
    s = header(a.title, a.command)
    s += (f"Calibration: the 1 GiB device copy of `bench.py --copy-probe` (`{copyk[:60]}`) reads "
          f"FETCH_SIZE = {cf[copyk]:.1f} KiB and WRITE_SIZE = {cw[copyk]:.1f} KiB per dispatch, so FETCH_SIZE is "
          f"scaled by {ff:.3f} and WRITE_SIZE by {fw:.3f}.  Averages over all dispatches of each kernel; tile = "
          "(threads per block, 0 = wave-cooperative kernel; kets per item) from the tiles file.\n\n")
    s += "| kernel | tile | dispatches | HBM read MB | HBM write MB | total MB per launch |\n|---|---|---:|---:|---:|---:|\n"
    for lab, v in sorted(res.items(), key=lambda kv: -kv[1]["bytes_per_launch"]):
        s += (f"| `{lab}` | {tuple(v['tile'])} | {v['dispatches']} | {v['fetch_bytes'] / 1e6:.2f} | "
              f"{v['write_bytes'] / 1e6:.2f} | {v['bytes_per_launch'] / 1e6:.2f} |\n")
    Path(a.md).write_text(s + a.note)


def cmd_mfma(a):
    cnt, ndisp, res = {}, {}, {}
    for d in a.pmc:
        c, n, r = load_counters(d)
        for k, v in c.items():
            cnt.setdefault(k, {}).update(v)
            ndisp[k] = n[k]
            res[k] = r[k]
    st = load_stats(a.stats) if a.stats else {}
    keys = [k for k in cnt if any(m in k for m in a.match)]
    keys.sort(key=lambda k: -(st.get(k, (0, 0))[0] * st.get(k, (0, 0))[1]))
    s = header(a.title, a.command)
    s += ("Per dispatch: wave states as fractions of SQ_WAVE_CYCLES (waiting on a counter, waiting to issue, "
          "issuing VALU or LDS instructions); MFMA instructions (SQ_INSTS_VALU_MFMA_F32 / _F64, "
          f"{MFMA_FLOPS} flops each for the 16x16x4 shapes these kernels use); MFMA busy = "
          f"SQ_VALU_MFMA_BUSY_CYCLES / ({SIMDS} SIMDs x 2.4 GHz x avg duration); MFMA TFLOP/s against the dense "
          "peak of the product type (FP32 157.3, FP64 78.6).  Average durations come from the kernel-trace run.\n\n")
    cols = ["kernel", "dispatches", "VGPR_Count", "avg µs", "wait on counter", "wait to issue", "VALU", "LDS",
            "MFMA instr / dispatch", "MFMA busy", "MFMA TFLOP/s", "of peak"]
    s += "| " + " | ".join(cols) + " |\n|---|" + "---:|" * (len(cols) - 1) + "\n"
    for k in keys:
        c = cnt[k]
        dur = st.get(k, (0, 0.0))[1] * 1e-6
        wc = c.get("SQ_WAVE_CYCLES", 0)
        states = [f"{c.get(x, 0) / wc:.2f}" if wc else "-" for x in
                  ("SQ_WAIT_ANY", "SQ_WAIT_INST_ANY", "SQ_ACTIVE_INST_VALU", "SQ_ACTIVE_INST_LDS")]
        kind = "F64" if c.get("SQ_INSTS_VALU_MFMA_F64", 0) > c.get("SQ_INSTS_VALU_MFMA_F32", 0) else "F32"
        n = c.get(f"SQ_INSTS_VALU_MFMA_{kind}", 0)
        busy = c.get("SQ_VALU_MFMA_BUSY_CYCLES", 0) / (SIMDS * CLOCK_HZ * dur) if dur else 0.0
        fl = n * MFMA_FLOPS / dur if dur else 0.0
        row = [f"`{k}`", str(ndisp[k]), str(res[k][0]), f"{dur * 1e6:.1f}" if dur else "-"] + states + [
            f"{n / 1e6:.2f} M ({kind})", f"{busy:.2f}" if dur else "-", f"{fl / 1e12:.1f}" if dur else "-",
            f"{fl / PEAK_MFMA[kind]:.3f}" if dur else "-"]
        s += "| " + " | ".join(row) + " |\n"
    Path(a.md).write_text(s + a.note)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    sub = ap.add_subparsers(dest="what", required=True)
    for name in ("stats", "counters", "traffic", "mfma"):
        p = sub.add_parser(name)
        p.add_argument("--md", required=True)
        p.add_argument("--title", default="rocprofv3 summary")
        p.add_argument("--command", action="append", default=[], help="command line that produced the data")
        p.add_argument("--note", default="", help="text appended after the table")
        if name == "stats":
            p.add_argument("--csv", required=True)
            p.add_argument("--top", type=int, default=40)
        elif name in ("counters", "mfma"):
            p.add_argument("--pmc", action="append", required=True)
            p.add_argument("--stats", default=None)
            if name == "mfma":
                p.add_argument("--match", action="append", required=True, help="substring of the kernel names")
        else:
            for o in ("--fetch", "--write", "--copy-fetch", "--copy-write", "--tiles", "--json"):
                p.add_argument(o, required=True)
    a = ap.parse_args(argv)
    {"stats": cmd_stats, "counters": cmd_counters, "traffic": cmd_traffic, "mfma": cmd_mfma}[a.what](a)


if __name__ == "__main__":
    main()
