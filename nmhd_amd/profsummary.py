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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("trace")
    ap.add_argument("--pmc", nargs="+", action="append", default=[])
    a = ap.parse_args()
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
