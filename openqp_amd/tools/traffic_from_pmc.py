"""HBM traffic per launch of the benchmark's kernels from rocprofv3 counter runs.

Usage:
  python openqp_amd/tools/traffic_from_pmc.py FETCH_CSV WRITE_CSV COPY_FETCH_CSV COPY_WRITE_CSV COPY_BYTES OUT_JSON

FETCH_CSV / WRITE_CSV: ``rocprofv3 --pmc FETCH_SIZE`` / ``--pmc WRITE_SIZE``
counter_collection.csv files of ``bench.py --steps 1 --warmup 0 --fp32-cutoff
1e-3 --no-cpu-baseline --no-grad --no-scf --no-fshell`` (per class: dispatch
1 is the FP64 build, the later ones the mixed builds).  The COPY files are the
same counters for a plain device copy of COPY_BYTES bytes: the counters read
KiB in units that this calibrates (bytes = counter * COPY_BYTES / counter of
the copy).  Writes, per kernel name as bench.py prints it, the FETCH and
WRITE bytes of one FP64 and one mixed launch.
"""
import collections
import csv
import json
import re
import sys


def per_dispatch(path):
    d = collections.defaultdict(list)
    for r in csv.DictReader(open(path)):
        d[r["Kernel_Name"]].append((int(r["Dispatch_Id"]), float(r["Counter_Value"])))
    return {k: [v for _, v in sorted(x)] for k, x in d.items()}


def bench_name(kernel):
    # <la, lb, lc, ld, pure[, open-shell pass[, tiled]]>: closed-shell
    # instances only; the tiled form of k_jkg is reported as k_jkt
    m = re.match(r"void oqph::(k_\w+)<(\d), (\d), (\d), (\d), (true|false)((?:, (?:true|false))*)>", kernel)
    if not m:
        return None
    flags = [x.strip() for x in m.group(7).split(",") if x.strip()]
    if flags and flags[0] == "true":
        return None
    kern = "k_jkt" if m.group(1) == "k_jkg" and len(flags) > 1 and flags[1] == "true" else m.group(1)
    return f"{kern}<{m.group(2)},{m.group(3)},{m.group(4)},{m.group(5)}>"


def copy_scale(path, nbytes):
    vals = [v for k, x in per_dispatch(path).items() if "copy" in k.lower() for v in x]
    return nbytes / max(vals)


def main(fetch, write, cfetch, cwrite, cbytes, out):
    sf, sw = copy_scale(cfetch, float(cbytes)), copy_scale(cwrite, float(cbytes))
    f, w = per_dispatch(fetch), per_dispatch(write)
    ker = {}
    for k in f:
        n = bench_name(k)
        if not n or k not in w or "true" not in k:
            continue
        ker[n] = {"fp64": {"fetch_bytes": f[k][0] * sf, "write_bytes": w[k][0] * sw},
                  "mixed": {"fetch_bytes": f[k][-1] * sf, "write_bytes": w[k][-1] * sw}}
    json.dump({"source": "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of bench.py on (H2O)64/cc-pVDZ, scaled by a "
                         f"{int(float(cbytes))}-byte device copy (FETCH x{sf:.4g}, WRITE x{sw:.4g} bytes per unit)",
               "kernels": ker}, open(out, "w"), indent=1)


if __name__ == "__main__":
    main(*sys.argv[1:7])
