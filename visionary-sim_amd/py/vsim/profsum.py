"""Summarises rocprofv3 CSV output for the solver's kernels.

    python -m vsim.profsum trace <dir>/..._kernel_trace.csv
    python -m vsim.profsum pmc <fetch_counter_collection.csv> <write_counter_collection.csv> \
        <copy_fetch.csv> <copy_write.csv>
    python -m vsim.profsum valu <counter_collection.csv with SQ_INSTS_VALU [VALUBusy ...]>
    python -m vsim.profsum phases <..._kernel_trace.csv> [frames]

`trace` prints per (kernel, launch grid) call counts and average durations, so
the batched launches (large grids) separate from the per-scene ones. `pmc`
prints FETCH_SIZE / WRITE_SIZE per dispatch, scaled to bytes by what the same
counters read for a device-to-device copy of known size. `valu` prints
SQ_INSTS_VALU (wave instructions) per dispatch against the VALU issue peak
of the dispatch's duration: one wave instruction per CU per cycle, 256 CUs
at 2.4 GHz. `phases` splits the batched colour-pair launches of a bench.py
run (no autotuner) by phase and, in the graph-replayed frames, by row class.
"""
import collections
import csv
import sys


def short(name):
    name = name.replace("(anonymous namespace)::", "").replace("void ", "").strip('"')
    return name.split("(")[0].replace("vsim::", "")


def trace(path, top=20, batched_min=1 << 20):
    d = collections.defaultdict(list)
    for r in csv.DictReader(open(path)):
        grid = int(r.get("Grid_Size_X") or r.get("Grid_Size") or 0)
        d[(short(r["Kernel_Name"]), grid)].append(int(r["End_Timestamp"]) - int(r["Start_Timestamp"]))
    rows = sorted(d.items(), key=lambda kv: -sum(kv[1]))[:top]
    out = ["| kernel | grid (threads) | calls | avg us | total ms |", "|---|---:|---:|---:|---:|"]
    for (k, g), v in rows:
        out.append(f"| {k} | {g} | {len(v)} | {sum(v) / len(v) / 1e3:.2f} | {sum(v) / 1e6:.2f} |")
    # The batched launches, one per pass across every scene, have grids far
    # above any single scene's.
    agg = collections.defaultdict(list)
    for (k, g), v in d.items():
        if g >= batched_min:
            agg[k].extend(v)
    out += ["", f"Batched launches (grid >= {batched_min} threads), all grids pooled:", "",
            "| kernel | calls | avg us | min us | max us |", "|---|---:|---:|---:|---:|"]
    for k, v in sorted(agg.items(), key=lambda kv: -sum(kv[1])):
        out.append(f"| {k} | {len(v)} | {sum(v) / len(v) / 1e3:.2f} | {min(v) / 1e3:.2f} | {max(v) / 1e3:.2f} |")
    return "\n".join(out)


def _dispatches(path, counter=None):
    """(counter value, duration in ns) per dispatch, keyed by (kernel, grid).
    The splat pass runs in two output modes with one kernel: dispatches that
    take over 1.5 times the fastest one of their grid (the engine axes'
    Jacobi sweeps) get their own row."""
    rows = collections.defaultdict(list)
    for r in csv.DictReader(open(path)):
        if counter and r["Counter_Name"] != counter:
            continue
        dur = int(r.get("End_Timestamp") or 0) - int(r.get("Start_Timestamp") or 0)
        rows[(short(r["Kernel_Name"]), int(r.get("Grid_Size") or 0))].append((float(r["Counter_Value"]), dur))
    d = collections.defaultdict(list)
    for (k, g), v in rows.items():
        fastest = min(t for _, t in v)
        for value, t in v:
            mode = " (cov+axes)" if k.startswith("k_splats") and fastest > 0 and t > 1.5 * fastest else ""
            d[(k + mode, g)].append((value, t))
    return d


def _per_kernel(path):
    return {key: [value for value, _ in v] for key, v in _dispatches(path).items()}


def pmc(fetch, write, cfetch, cwrite, copy_bytes=1 << 30):
    cf = max((sum(v) / len(v) for (k, _), v in _per_kernel(cfetch).items() if "copyBuffer" in k), default=None)
    cw = max((sum(v) / len(v) for (k, _), v in _per_kernel(cwrite).items() if "copyBuffer" in k), default=None)
    fs = copy_bytes / (cf * 1024) if cf else 1.0
    ws = copy_bytes / (cw * 1024) if cw else 1.0
    F, W = _per_kernel(fetch), _per_kernel(write)
    out = [f"copy calibration: 1 GiB copy reads FETCH_SIZE {cf:.0f} KiB (scale {fs:.3f}), "
           f"WRITE_SIZE {cw:.0f} KiB (scale {ws:.3f})", "",
           "| kernel | grid (threads) | dispatches | FETCH MB (scaled) | WRITE MB (scaled) | total MB |",
           "|---|---:|---:|---:|---:|---:|"]
    for key in sorted(F, key=lambda k: -sum(F[k])):
        if key not in W or "rocclr" in key[0]:
            continue
        f = sum(F[key]) / len(F[key]) * 1024 * fs / 1e6
        w = sum(W[key]) / len(W[key]) * 1024 * ws / 1e6
        out.append(f"| {key[0]} | {key[1]} | {len(F[key])} | {f:.1f} | {w:.1f} | {f + w:.1f} |")
    return "\n".join(out)


VALU_ISSUE_PER_S = 256 * 2.4e9  # wave instructions per second, MI355X


def valu(path, batched_min=1 << 19):
    """Per (kernel, grid): SQ_INSTS_VALU per dispatch against one wave
    instruction per CU per cycle, and, where the run collected them, the
    derived VALUBusy (share of cycles the VALUs issue) and VALUUtilization
    (share of a wave's lanes active per VALU instruction), the quad-cycles in
    which a SIMD issued two VALU instructions, and F32 transcendentals."""
    counters = {}
    for name in ("SQ_INSTS_VALU", "VALUBusy", "VALUUtilization", "SQ_ACTIVE_INST_VALU2", "SQ_INSTS_VALU_TRANS_F32"):
        d = _dispatches(path, name)
        if d:
            counters[name] = d
    d = counters["SQ_INSTS_VALU"]
    extra = [c for c in ("VALUBusy", "VALUUtilization", "SQ_ACTIVE_INST_VALU2", "SQ_INSTS_VALU_TRANS_F32") if c in counters]
    head = {"VALUBusy": "VALUBusy %", "VALUUtilization": "lanes active %",
            "SQ_ACTIVE_INST_VALU2": "dual-issue quad-cycles (M)", "SQ_INSTS_VALU_TRANS_F32": "F32 transcendentals (M)"}
    out = ["| kernel | grid (threads) | dispatches | VALU wave instructions (M) | avg us | VALU instructions / (CU x cycle) |"
           + "".join(f" {head[c]} |" for c in extra),
           "|---|---:|---:|---:|---:|---:|" + "---:|" * len(extra)]
    mean = lambda v: sum(x for x, _ in v) / len(v)
    for key in sorted(d, key=lambda k: -sum(v for v, _ in d[k])):
        if key[1] < batched_min or "rocclr" in key[0]:
            continue
        v = d[key]
        inst = mean(v)
        dur = sum(y for _, y in v) / len(v) * 1e-9
        row = f"| {key[0]} | {key[1]} | {len(v)} | {inst / 1e6:.1f} | {dur * 1e6:.1f} | {inst / (dur * VALU_ISSUE_PER_S):.2f} |"
        for c in extra:
            x = mean(counters[c][key]) if key in counters[c] else float("nan")
            row += f" {x:.1f} |" if c in ("VALUBusy", "VALUUtilization") else f" {x / 1e6:.2f} |"
        out.append(row)
    return "\n".join(out)


def phases(path, frames=4, per_frame=40, kernel="k_elastic_rows", batched_min=1 << 20):
    """The batched launches of `kernel` in start-time order, as bench.py
    makes them with the autotuner off: `frames` timed frames (the first
    captures the frame graph, the rest replay it), one profiled frame (HIP
    events around each launch), then one `vs_world_time_pass` burst, each of
    `per_frame` launches; and the replayed frames' launches by row class
    (launch index mod 4)."""
    rows = []
    for r in csv.DictReader(open(path)):
        grid = int(r.get("Grid_Size_X") or r.get("Grid_Size") or 0)
        if short(r["Kernel_Name"]).startswith(kernel) and grid >= batched_min:
            rows.append((int(r["Start_Timestamp"]), int(r["End_Timestamp"]) - int(r["Start_Timestamp"])))
    rows.sort()
    t = [d for _, d in rows]
    avg = lambda v: sum(v) / len(v) / 1e3 if v else float("nan")
    cut = [0, per_frame, frames * per_frame, (frames + 1) * per_frame, (frames + 2) * per_frame]
    names = ["first frame (graph capture)", "graph-replayed frames", "profiled frame (HIP events around every launch)",
             "`vs_world_time_pass` burst (back to back, no folding)"]
    out = [f"{len(t)} batched `{kernel}` launches by phase (start-time order):", "",
           "| phase | launches | avg us |", "|---|---:|---:|"]
    for name, a, b in zip(names, cut, cut[1:]):
        out.append(f"| {name} | {len(t[a:b])} | {avg(t[a:b]):.2f} |")
    replay = t[cut[1]:cut[2]]
    out += ["", "Graph-replayed frames by row class:", "", "| class (colours) | launches | avg us |", "|---|---:|---:|"]
    for c in range(4):
        v = replay[c::4]
        out.append(f"| {c} ({2 * c}, {2 * c + 1}) | {len(v)} | {avg(v):.2f} |")
    return "\n".join(out)


if __name__ == "__main__":
    if sys.argv[1] == "trace":
        print(trace(sys.argv[2]))
    elif sys.argv[1] == "valu":
        print(valu(sys.argv[2]))
    elif sys.argv[1] == "phases":
        print(phases(sys.argv[2], *(int(a) for a in sys.argv[3:4])))
    else:
        print(pmc(*sys.argv[2:6]))
