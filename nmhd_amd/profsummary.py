"""Summarise rocprofv3 result databases (rocpd SQLite) per kernel.

    python -m nmhd_amd.profsummary TRACE_DB [--pmc DB ...]

TRACE_DB comes from `rocprofv3 --kernel-trace --stats -d DIR -o run -- CMD`;
each --pmc DB from `rocprofv3 --pmc COUNTERS -d DIR -o run -- CMD`.  Prints
one line per kernel: calls, mean duration (us), and the mean per dispatch of
every counter found (FETCH_SIZE / WRITE_SIZE are KiB per dispatch).
"""
import argparse
import collections
import sqlite3


def short(name):
    name = name.replace("(anonymous namespace)::", "")
    name = name.split("(")[0]
    return name.replace("void ", "")


def kernel_times(db):
    con = sqlite3.connect(db)
    out = {}
    for name, calls, total, avg, pct in con.execute(
            "select name, total_calls, total_duration, average, percentage from top_kernels"):
        out[short(name)] = (calls, avg, pct)  # average is in us
    return out


def counters(db):
    con = sqlite3.connect(db)
    acc = collections.defaultdict(lambda: collections.defaultdict(list))
    for name, cname, value in con.execute(
            "select kernel_name, counter_name, value from counters_collection"):
        acc[short(name)][cname].append(value)
    return {k: {c: sum(v) / len(v) for c, v in d.items()} for k, d in acc.items()}


# rocprof kernel names -> per-launch slots of the tiled step (the stage-2
# update is the instantiation with the CFL epilogue: <TX, TY, false, true>)
SLOTS = ((("k_faces<", ""), ("faces_s1", "faces_s2")),
         (("k_update_march<", ", false, true>"), ("update_s2",)),
         (("k_update_march<", ", false, false>"), ("update_s1",)))


def traffic(fetch_db, write_db, fetch_scale, write_scale, valu_db=None):
    """HBM bytes per launch per step slot: FETCH_SIZE * fetch_scale +
    WRITE_SIZE * write_scale (KiB per dispatch -> bytes).  The scales are
    what the counters read for a copy kernel of known size.  With valu_db
    (a run with SQ_INSTS_VALU) also the VALU wave instructions per launch."""
    f = counters(fetch_db)
    w = counters(write_db)
    v = counters(valu_db) if valu_db else {}
    out = {}
    for (pat, suffix), slots in SLOTS:
        for k in f:
            if pat in k and k.endswith(suffix) and k in w:
                fk = f[k]["FETCH_SIZE"]
                wk = w[k]["WRITE_SIZE"]
                rec = {"rocprof_kernel": k, "fetch_kib": fk, "write_kib": wk,
                       "fetch_scale": fetch_scale, "write_scale": write_scale,
                       "bytes_per_launch": 1024.0 * (fk * fetch_scale + wk * write_scale)}
                if "SQ_INSTS_VALU" in v.get(k, {}):
                    rec["valu_insts"] = v[k]["SQ_INSTS_VALU"]
                for sl in slots:
                    out[sl] = rec
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("trace")
    ap.add_argument("--pmc", nargs="+", action="append", default=[])
    ap.add_argument("--traffic-json", help="write HBM bytes per launch (needs --fetch, --write)")
    ap.add_argument("--fetch")
    ap.add_argument("--write")
    ap.add_argument("--fetch-scale", type=float, default=1.0)
    ap.add_argument("--write-scale", type=float, default=1.0)
    ap.add_argument("--valu", help="a --pmc run with SQ_INSTS_VALU (traffic-json mode)")
    ap.add_argument("--global-size", type=int, nargs=3, default=(256, 256, 256))
    ap.add_argument("--path", default="tiled")
    ap.add_argument("--command", default="")
    a = ap.parse_args()
    if a.traffic_json:
        import json
        kern = traffic(a.fetch, a.write, a.fetch_scale, a.write_scale, a.valu)
        # the kernel-trace mean duration beside the counts: a reader (and the
        # bench) can tell whether the counts belong to the kernel being timed
        times = kernel_times(a.trace)
        for r in kern.values():
            if r["rocprof_kernel"] in times:
                r["trace_mean_us"] = times[r["rocprof_kernel"]][1]
        rec = {"global_size": list(a.global_size), "path": a.path, "command": a.command,
               "kernels": kern}
        with open(a.traffic_json, "w") as fh:
            json.dump(rec, fh, indent=1)
        return
    times = kernel_times(a.trace)
    pmc = {}
    for db in [d for group in a.pmc for d in group]:
        for k, d in counters(db).items():
            pmc.setdefault(k, {}).update(d)
    names = sorted(times, key=lambda k: -times[k][0] * times[k][1])
    cols = sorted({c for d in pmc.values() for c in d})
    print("%-34s %6s %10s %6s " % ("kernel", "calls", "mean_us", "pct") +
          " ".join("%14s" % c for c in cols))
    for k in names:
        calls, avg, pct = times[k]
        row = "%-34s %6d %10.1f %6.1f " % (k[-34:], calls, avg, pct)
        row += " ".join("%14.1f" % pmc.get(k, {}).get(c, float("nan")) for c in cols)
        print(row)


if __name__ == "__main__":
    main()
