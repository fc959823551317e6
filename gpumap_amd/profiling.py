"""Kernel summary of a rocprofv3 database (``--kernel-trace`` run, the
``*_results.db`` file): calls, total, average, minimum and maximum duration
per kernel, as a markdown table.

    python -m gpumap_amd.profiling gpurun_out/prof/run_results.db [--top 30]
        [--after NAME:N]

``--after NAME:N`` keeps the kernels that start after the N-th launch (in
time order) of the kernels whose name contains NAME has ended: the timed
part of a run that first warms up.  ``--until NAME:N`` keeps those that
start before the N-th launch of NAME has ended: with ``--after`` a window
such as one session of a run of several.  ``--pmc`` summarises a ``--pmc`` run
instead: the mean and the maximum of each counter per dispatch, per kernel
(the maximum picks out one large dispatch among many small ones of the same
kernel, such as a calibration copy).

rocprofv3's own ``--stats`` CSV holds the same figures; this reads the
database, which is complete even when the profiler ends abnormally after the
program has finished.
"""

from __future__ import annotations

import argparse
import sqlite3


def _end_of(con, spec: str) -> int:
    """End time of the N-th launch (in time order) of the kernels whose name
    contains NAME, for spec NAME:N."""
    name, n = spec.rsplit(":", 1)
    ends = con.execute(
        "SELECT end FROM kernels WHERE instr(name, ?) > 0 ORDER BY start",
        (name,)).fetchall()
    if len(ends) < int(n) or int(n) < 1:
        raise ValueError(f"fewer than {n} launches of {name!r}")
    return ends[int(n) - 1][0]


def kernel_table(db_path: str, top: int = 30, after: str | None = None,
                 until: str | None = None) -> str:
    con = sqlite3.connect(db_path)
    try:
        t0 = _end_of(con, after) if after else 0
        t1 = _end_of(con, until) if until else (1 << 63) - 1
        rows = con.execute(
            "SELECT name, COUNT(*), SUM(end - start), AVG(end - start), "
            "MIN(end - start), MAX(end - start) FROM kernels "
            "WHERE start > ? AND start < ? "
            "GROUP BY name ORDER BY SUM(end - start) DESC", (t0, t1)).fetchall()
    finally:
        con.close()
    total = sum(r[2] for r in rows) or 1
    out = [f"Total kernel time {total / 1e6:.1f} ms.", "",
           "| kernel | calls | total ms | avg us | min us | max us | % |",
           "|---|---|---|---|---|---|---|"]
    for name, calls, tot, avg, mn, mx in rows[:top]:
        short = name.split("(")[0]
        if len(short) > 64:
            short = short[:61] + "..."
        out.append(f"| `{short}` | {calls} | {tot / 1e6:.2f} | {avg / 1e3:.2f} "
                   f"| {mn / 1e3:.2f} | {mx / 1e3:.2f} | "
                   f"{100.0 * tot / total:.1f} |")
    return "\n".join(out)


def pmc_table(db_path: str, top: int = 30) -> str:
    con = sqlite3.connect(db_path)
    try:
        rows = con.execute(
            "SELECT kernel_name, counter_name, COUNT(*), AVG(value), "
            "SUM(value), MAX(value) FROM counters_collection "
            "GROUP BY kernel_name, counter_name").fetchall()
    finally:
        con.close()
    per: dict[str, dict[str, tuple[int, float, float, float]]] = {}
    for k, c, n, avg, tot, mx in rows:
        per.setdefault(k, {})[c] = (n, avg, tot, mx)
    counters = sorted({c for v in per.values() for c in v})
    order = sorted(per, key=lambda k: -max(v[2] for v in per[k].values()))
    heads = [f"{c} mean | {c} max" for c in counters]
    out = ["| kernel | dispatches | " + " | ".join(heads) + " |",
           "|---|---|" + "---|---|" * len(counters)]
    for k in order[:top]:
        short = k.split("(")[0]
        if len(short) > 64:
            short = short[:61] + "..."
        n = max(v[0] for v in per[k].values())
        cells = [f"{per[k][c][1]:.6g} | {per[k][c][3]:.6g}" if c in per[k]
                 else " | " for c in counters]
        out.append(f"| `{short}` | {n} | " + " | ".join(cells) + " |")
    return "\n".join(out)


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("db")
    ap.add_argument("--top", type=int, default=30)
    ap.add_argument("--after", default=None, metavar="NAME:N")
    ap.add_argument("--until", default=None, metavar="NAME:N")
    ap.add_argument("--pmc", action="store_true")
    a = ap.parse_args()
    if a.pmc:
        print(pmc_table(a.db, a.top))
    else:
        print(kernel_table(a.db, a.top, a.after, a.until))


if __name__ == "__main__":
    main()
