"""HBM traffic per launch from two rocprofv3 --pmc runs (FETCH_SIZE, then
WRITE_SIZE, both in KiB per dispatch), each scaled by what the same counter
reads for the plain copy kernel of known size that bench.py launches first
(k_axpby<double> copying 256 MiB: 256 MiB read, 256 MiB written).

    python -m krans_amd.pmc_traffic <fetch.db> <write.db> <kernel-substring> [--largest]
    python -m krans_amd.pmc_traffic --avg <db> <counter> <kernel-substring>...

The second form prints, per kernel name containing one of the substrings,
the dispatch count and the average of one counter per dispatch (any
counter of a --pmc run, e.g. TCP_TCC_READ_REQ_sum, unscaled).

With --largest only the dispatches whose FETCH_SIZE is at least half the
largest one count (one launch shape out of several, e.g. the 96-column Gram
pass among the 32-column ones).

Prints JSON: per-launch algorithmic bytes are not known here; the caller
compares the HBM bytes with its own algorithmic count.
"""
import json
import sqlite3
import sys

COPY_BYTES = 256 * (1 << 20)


def per_dispatch(db, counter, like):
    c = sqlite3.connect(db)
    rows = c.execute("select dispatch_id, kernel_name, sum(value), grid_size from "
                     "counters_collection where counter_name = ? and kernel_name like ? "
                     "group by dispatch_id", (counter, f"%{like}%")).fetchall()
    return rows


def copy_factor(db, counter):
    rows = per_dispatch(db, counter, "k_axpby<double>")
    # the copy probe is the dispatch with the largest grid among k_axpby<double>
    big = max(rows, key=lambda r: (r[3], r[2]))
    kib = big[2]
    return COPY_BYTES / (kib * 1024.0), kib


def main(fetch_db, write_db, like, largest=False):
    ff, fk = copy_factor(fetch_db, "FETCH_SIZE")
    wf, wk = copy_factor(write_db, "WRITE_SIZE")
    f = per_dispatch(fetch_db, "FETCH_SIZE", like)
    w = per_dispatch(write_db, "WRITE_SIZE", like)
    if largest and f:
        top = max(r[2] for r in f)
        keep = {r[0] for r in f if r[2] >= 0.5 * top}
        f = [r for r in f if r[0] in keep]
        # the write run dispatches carry the same ids when the program repeats
        # itself exactly; fall back to the largest writes otherwise
        w2 = [r for r in w if r[0] in keep]
        if w2:
            w = w2
        elif w:
            wt = max(r[2] for r in w)
            w = [r for r in w if r[2] >= 0.5 * wt]
    fetch = sum(r[2] for r in f) * 1024.0 * ff / max(1, len(f))
    write = sum(r[2] for r in w) * 1024.0 * wf / max(1, len(w))
    out = {"kernel_like": like, "largest_only": largest,
           "dispatches_fetch_run": len(f), "dispatches_write_run": len(w),
           "copy_probe_fetch_KiB": fk, "copy_probe_write_KiB": wk,
           "fetch_scale": ff, "write_scale": wf,
           "hbm_read_bytes_per_launch": fetch, "hbm_write_bytes_per_launch": write,
           "hbm_bytes_per_launch": fetch + write}
    print(json.dumps(out, indent=1))


def averages(db, counter, likes):
    c = sqlite3.connect(db)
    out = {}
    for like in likes:
        rows = c.execute("select kernel_name, dispatch_id, sum(value) from counters_collection "
                         "where counter_name = ? and kernel_name like ? "
                         "group by dispatch_id", (counter, f"%{like}%")).fetchall()
        for name, _, v in rows:
            n, tot = out.get(name, (0, 0.0))
            out[name] = (n + 1, tot + v)
    res = {name: {"dispatches": n, f"avg_{counter}": tot / n} for name, (n, tot) in out.items()}
    print(json.dumps(res, indent=1))


if __name__ == "__main__":
    if sys.argv[1] == "--avg":
        averages(sys.argv[2], sys.argv[3], sys.argv[4:])
    else:
        main(sys.argv[1], sys.argv[2], sys.argv[3], "--largest" in sys.argv[4:])
