"""Short text summaries of rocprofv3 output databases (rocpd SQLite).

    python -m dynamite_amd.rocprof_summary trace <dir> [last_k]
        per-kernel dispatch count and duration statistics (kernel trace), optionally
        over the last `last_k` dispatches of each kernel;
    python -m dynamite_amd.rocprof_summary pmc <dir> [last_k]
        mean of every collected counter per kernel over its last `last_k` dispatches;
    python -m dynamite_amd.rocprof_summary steps <dir> [last_k]
        the device loop's steps (dispatches from one step start - loop_begin_kernel, or
        loop_end_begin_kernel, which ends a step and begins the next - to the next) over the
        last `last_k` complete steps: wall time per step, kernel-busy time per kernel, and
        the idle time between dispatches.
<dir> is the rocprofv3 -d directory (searched recursively for *.db).
"""
import glob
import sqlite3
import statistics as st
import sys


def _db(path):
    found = glob.glob(path + "/**/*.db", recursive=True)
    if not found:
        raise SystemExit(f"no rocpd database under {path}")
    return sqlite3.connect(found[0])


def _short(name):
    s = name.replace("dmfe::(anonymous namespace)::", "").replace("dmfe::", "")
    s = s.split("(")[0]
    return s[5:] if s.startswith("void ") else s


def trace(path, k):
    con = _db(path)
    cols = [r[1] for r in con.execute("pragma table_info(kernels)")]
    nm = "name" if "name" in cols else "kernel_name"
    agg = {}
    for name, a, b in con.execute(f"select {nm}, start, end from kernels order by start"):
        agg.setdefault(name, []).append((b - a) / 1e3)
    print(f"{'kernel':58s} {'calls':>6s} {'total_us':>11s} {'avg_us':>8s} {'min_us':>8s} {'max_us':>8s}")
    for name, ds in sorted(agg.items(), key=lambda kv: -sum(kv[1][-k:])):
        d = ds[-k:]
        print(f"{_short(name)[:58]:58s} {len(d):6d} {sum(d):11.1f} {st.mean(d):8.2f} {min(d):8.2f} {max(d):8.2f}")


def pmc(path, k):
    con = _db(path)
    agg = {}
    for name, c, v in con.execute("select kernel_name, counter_name, value from counters_collection order by dispatch_id"):
        agg.setdefault((name, c), []).append(v)
    for (name, c), vs in sorted(agg.items()):
        last = vs[-k:]
        print(f"{_short(name)[:58]:58s} {c:14s} dispatches={len(vs):6d} last={len(last):5d} mean={st.mean(last):14.1f}")


def steps(path, k):
    con = _db(path)
    cols = [r[1] for r in con.execute("pragma table_info(kernels)")]
    nm = "name" if "name" in cols else "kernel_name"
    rows = [(_short(n), a, b) for n, a, b in con.execute(f"select {nm}, start, end from kernels order by start")]
    rows = [r for r in rows if r[0].startswith("loop_")]
    heads = [i for i, r in enumerate(rows) if r[0] in ("loop_begin_kernel", "loop_end_begin_kernel")]
    # steps in which the instance ran (a launch after the loop stopped exits at once)
    spans = [(a, b) for a, b in zip(heads[:-1], heads[1:])
             if max((r[2] - r[1] for r in rows[a:b] if r[0].startswith(("loop_rhs", "loop_fused"))), default=0) > 4000][-k:]
    if not spans:
        raise SystemExit("no complete device-loop steps in the trace")
    wall, busy, per = [], [], {}
    for a, b in spans:
        wall.append((rows[b][1] - rows[a][1]) / 1e3)
        busy.append(sum(r[2] - r[1] for r in rows[a:b]) / 1e3)
        for r in rows[a:b]:
            per.setdefault(r[0], []).append((r[2] - r[1]) / 1e3)
    n = len(spans)
    print(f"steps={n}  wall/step={st.mean(wall):.2f} us  kernel-busy/step={st.mean(busy):.2f} us  "
          f"idle between dispatches/step={st.mean(wall) - st.mean(busy):.2f} us")
    print(f"{'kernel':58s} {'per_step':>8s} {'avg_us':>8s} {'us/step':>8s}")
    for name, ds in sorted(per.items(), key=lambda kv: -sum(kv[1])):
        print(f"{name[:58]:58s} {len(ds) / n:8.2f} {st.mean(ds):8.2f} {sum(ds) / n:8.2f}")


if __name__ == "__main__":
    mode, path = sys.argv[1], sys.argv[2]
    k = int(sys.argv[3]) if len(sys.argv) > 3 else 10**9
    {"trace": trace, "pmc": pmc, "steps": steps}[mode](path, k)
