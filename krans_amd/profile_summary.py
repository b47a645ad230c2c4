"""Summarise a rocprofv3 kernel-trace database as a Markdown table.

    python -m krans_amd.profile_summary <run_results.db> [--title T] [--cmd C]
"""
import argparse
import sqlite3


def short(name):
    name = name.replace("kh::(anonymous namespace)::", "").replace("void ", "")
    return name.split("(")[0]


def summary(db, title="", cmd="", top=25, detail=()):
    c = sqlite3.connect(db)
    rows = c.execute("select name, count(*), sum(duration), avg(duration) from kernels "
                     "group by name order by sum(duration) desc").fetchall()
    tot = sum(r[2] for r in rows)
    out = [f"# {title}", "", "Command:", "", f"    {cmd}", "",
           f"Total kernel time {tot / 1e9:.3f} s over {sum(r[1] for r in rows)} dispatches.", "",
           "| kernel | calls | total ms | avg us | share |", "|---|---|---|---|---|"]
    for r in rows[:top]:
        out.append(f"| `{short(r[0])}` | {r[1]} | {r[2] / 1e6:.1f} | {r[3] / 1e3:.2f} | "
                   f"{100 * r[2] / tot:.1f}% |")
    if detail:
        out += ["", "By launch shape:", "", "| kernel | grid_x | calls | avg us |",
                "|---|---|---|---|"]
        for nm in detail:
            for r in c.execute("select name, grid_x, count(*), avg(duration) from kernels "
                               "where name like ? group by name, grid_x order by name, grid_x",
                               (f"%{nm}%",)):
                out.append(f"| `{short(r[0])}` | {r[1]} | {r[2]} | {r[3] / 1e3:.2f} |")
    return "\n".join(out) + "\n"


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("db")
    ap.add_argument("--title", default="rocprofv3 kernel summary")
    ap.add_argument("--cmd", default="")
    ap.add_argument("--detail", nargs="*", default=[])
    ap.add_argument("--top", type=int, default=25)
    a = ap.parse_args()
    print(summary(a.db, a.title, a.cmd, top=a.top, detail=a.detail), end="")
