"""Per-kernel summary of rocprofv3 result databases (rocpd SQLite).

    python verkko-hem-repo_amd/rocpd_summary.py KT_DB [FETCH_DB WRITE_DB [--calib-bytes N]]

KT_DB: a --kernel-trace run; prints calls, average and total duration per
kernel.  FETCH_DB / WRITE_DB: --pmc FETCH_SIZE and --pmc WRITE_SIZE runs;
prints the average KiB per dispatch of each counter and, when the run holds
the calibration copy kernel (k_copy16 of --calib-bytes bytes), the factor
that maps each counter to true bytes (known traffic / counter reading).
Further --pmc databases (SQ_ event counters and the like) print the average
per dispatch of each counter as read.
"""
import argparse
import collections
import re
import sqlite3


def short(n):
    n = n.replace("(anonymous namespace)::", "").replace("void ", "")
    return re.sub(r"\((int|unsigned|glh::RefTables|double|float|const|size_t).*", "", n)[:100]


def kernel_times(db):
    d = collections.defaultdict(list)
    for n, du in sqlite3.connect(db).execute("select name, duration from kernels"):
        d[short(n)].append(du)
    return d


def counters(db):
    """(kernel, counter) -> value per dispatch.  An SQ counter comes as one row
    per shader engine; the rows of a dispatch are summed."""
    per = collections.defaultdict(float)
    for n, cn, di, v in sqlite3.connect(db).execute(
            "select name, counter_name, dispatch_id, counter_value from pmc_events"):
        per[(short(n), cn, di)] += v
    d = collections.defaultdict(list)
    for (n, cn, _), v in per.items():
        d[(n, cn)].append(v)
    return d


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("kt")
    ap.add_argument("pmc", nargs="*")
    ap.add_argument("--calib-bytes", type=float, default=float(1 << 30))
    a = ap.parse_args()
    d = kernel_times(a.kt)
    tot = sum(sum(v) for v in d.values())
    print("kernel trace: calls, average and total duration")
    for n, v in sorted(d.items(), key=lambda kv: -sum(kv[1])):
        print("  %6d calls  avg %10.1f us  total %9.2f ms  %5.1f%%  %s" % (len(v), sum(v) / len(v) / 1e3,
                                                                         sum(v) / 1e6, 100 * sum(v) / tot, n))
    for db in a.pmc:
        c = counters(db)
        factor = {}
        for (n, cn), v in c.items():
            if n.endswith("k_copy16") and cn in ("FETCH_SIZE", "WRITE_SIZE"):
                kib = sum(v) / len(v)
                factor[cn] = a.calib_bytes / (kib * 1024.0)
        for (n, cn), v in sorted(c.items()):
            avg = sum(v) / len(v)
            if cn not in ("FETCH_SIZE", "WRITE_SIZE"):  # event counts: the average per dispatch as read
                print("  %-24s %14.4g per dispatch  n=%3d  %s" % (cn, avg, len(v), n))
                continue
            f = factor.get(cn)
            print("  %-10s %7.1f MiB/dispatch  x%.3f -> %8.3f GB  n=%3d  %s" % (
                cn, avg / 1024, f or 1.0, avg * 1024 * (f or 1.0) / 1e9, len(v), n))


if __name__ == "__main__":
    main()
