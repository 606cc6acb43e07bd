"""Summarise a rocprofv3 kernel-trace database (rocpd SQLite) as a markdown table.

Usage: python -m kaiju_amd.tools_rocpd_summary <results.db> <command string> > profiles/<name>.md
"""
import sqlite3
import sys


def summary(db_path: str, command: str) -> str:
    db = sqlite3.connect(db_path)
    rows = db.execute("select name, total_calls, total_duration, average, percentage from top_kernels").fetchall()
    out = ["# rocprofv3 kernel trace summary", "", "Command:", "", "```", command, "```", "",
           "| kernel | calls | total (us) | average (us) | % |", "|---|---:|---:|---:|---:|"]
    for name, calls, tot, avg, pct in rows:
        short = name.split("(")[0]
        out.append(f"| `{short}` | {calls} | {tot:.1f} | {avg:.1f} | {pct:.2f} |")
    return "\n".join(out) + "\n"


if __name__ == "__main__":
    print(summary(sys.argv[1], sys.argv[2]), end="")
