"""Summaries of rocprofv3 output for the kernels of libbsam_gpu.so.

Two commands, both writing Markdown (and, for counters, a JSON digest that
bench.py reads for the roofline's measured-traffic figure):

  python3 -m bsam_amd.profsum trace <run_results.db> --cmd "<command>" --title "<title>"
      kernel-trace statistics (rocprofv3 --kernel-trace --stats), per kernel
      and per kernel x launch size.

  python3 -m bsam_amd.profsum pmc --fetch <counter_collection.csv> --write <counter_collection.csv>
                                  --copy-bytes B --json out.json --cmd "..." --title "..."
      HBM traffic per dispatch from two counter runs (FETCH_SIZE, then
      WRITE_SIZE, in KiB per dispatch).  Each counter is scaled by the
      factor that makes it read the known byte count of the plain copy
      kernel k_copy_span (dst[i] = src[i] over B bytes each way) in the same
      run.
"""
import argparse
import collections
import csv
import json
import sqlite3
import sys

COPY_KERNEL = "k_copy_span"


def short_name(n):
    n = n.replace("(anonymous namespace)::", "").replace("void ", "")
    return n.split("(")[0]


def trace(args):
    db = sqlite3.connect(args.db)
    rows = db.execute("select name, grid_x, workgroup_x, duration from kernels").fetchall()
    byname = collections.defaultdict(lambda: [0, 0.0])
    bygrid = collections.defaultdict(lambda: [0, 0.0])
    for n, g, _w, d in rows:
        k = short_name(n)
        byname[k][0] += 1
        byname[k][1] += d
        bygrid[(k, g)][0] += 1
        bygrid[(k, g)][1] += d
    tot = sum(v[1] for v in byname.values())
    out = [f"# {args.title}", "", "Command (from the repository root, one MI355X):", "", "```", args.cmd, "```", ""]
    out += ["## Per kernel (all launches)", "", "| kernel | calls | total ms | avg us | % |", "|---|---:|---:|---:|---:|"]
    for k, v in sorted(byname.items(), key=lambda kv: -kv[1][1]):
        out.append(f"| `{k}` | {v[0]} | {v[1] / 1e6:.3f} | {v[1] / v[0] / 1e3:.2f} | {100 * v[1] / tot:.1f} |")
    out += ["", "## Per kernel and launch size (grid in threads; the finest level is a kernel's largest grid)", "",
            "| kernel | grid | calls | avg us | total ms |", "|---|---:|---:|---:|---:|"]
    for (k, g), v in sorted(bygrid.items(), key=lambda kv: -kv[1][1])[: args.top]:
        out.append(f"| `{k}` | {g} | {v[0]} | {v[1] / v[0] / 1e3:.2f} | {v[1] / 1e6:.3f} |")
    print("\n".join(out))


def _load_counter(path):
    per = collections.defaultdict(list)
    for r in csv.DictReader(open(path)):
        per[(short_name(r["Kernel_Name"]), int(r["Grid_Size"]))].append(float(r["Counter_Value"]))
    return per


def pmc(args):
    fetch = _load_counter(args.fetch)
    write = _load_counter(args.write)
    copies = [k for k in fetch if k[0] == COPY_KERNEL]
    if not copies:
        sys.exit(f"no {COPY_KERNEL} dispatch in {args.fetch}")
    ck = max(copies, key=lambda k: k[1])
    f_kib = sum(fetch[ck]) / len(fetch[ck])
    w_kib = sum(write[ck]) / len(write[ck])
    sf = args.copy_bytes / (f_kib * 1024.0)
    sw = args.copy_bytes / (w_kib * 1024.0)
    kernels = {}
    for k in fetch:
        fv = sum(fetch[k]) / len(fetch[k])
        wv = sum(write.get(k, [0.0])) / max(1, len(write.get(k, [0.0])))
        kernels[f"{k[0]}@{k[1]}"] = {"kernel": k[0], "grid": k[1], "dispatches": len(fetch[k]),
                                     "fetch_kib": fv, "write_kib": wv,
                                     "bytes": fv * 1024.0 * sf + wv * 1024.0 * sw}
    top = sorted(kernels.items(), key=lambda kv: -kv[1]["bytes"] * kv[1]["dispatches"])[: args.top]
    digest = {"copy_kernel": f"{ck[0]}@{ck[1]}", "copy_bytes": args.copy_bytes,
              "fetch_scale": sf, "write_scale": sw, "kernels": dict(top), "cmd": args.cmd}
    if args.json:
        with open(args.json, "w") as fh:
            json.dump(digest, fh, indent=1, sort_keys=True)
    out = [f"# {args.title}", "", "Commands (from the repository root, one MI355X):", "", "```", args.cmd, "```", "",
           f"Scale factors from `{ck[0]}` (grid {ck[1]}), which reads and writes {args.copy_bytes} bytes: "
           f"FETCH_SIZE reads {f_kib:.0f} KiB, factor {sf:.4f}; WRITE_SIZE reads {w_kib:.0f} KiB, factor {sw:.4f}.", "",
           "| kernel | grid | dispatches | FETCH KiB | WRITE KiB | scaled HBM MB per dispatch |",
           "|---|---:|---:|---:|---:|---:|"]
    for key, v in top:
        out.append(f"| `{v['kernel']}` | {v['grid']} | {v['dispatches']} | {v['fetch_kib']:.0f} | "
                   f"{v['write_kib']:.0f} | {v['bytes'] / 1e6:.1f} |")
    print("\n".join(out))


def main(argv=None):
    ap = argparse.ArgumentParser(prog="bsam_amd.profsum")
    sub = ap.add_subparsers(dest="what", required=True)
    t = sub.add_parser("trace")
    t.add_argument("db")
    t.add_argument("--cmd", required=True)
    t.add_argument("--title", required=True)
    t.add_argument("--top", type=int, default=30)
    p = sub.add_parser("pmc")
    p.add_argument("--fetch", required=True)
    p.add_argument("--write", required=True)
    p.add_argument("--copy-bytes", type=float, required=True)
    p.add_argument("--json")
    p.add_argument("--cmd", required=True)
    p.add_argument("--title", required=True)
    p.add_argument("--top", type=int, default=25)
    a = ap.parse_args(argv)
    trace(a) if a.what == "trace" else pmc(a)


if __name__ == "__main__":
    main()
