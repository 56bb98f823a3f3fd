"""Summaries of rocprofv3 runs of bench.py (kernel times and HBM traffic);
measurement tooling beside bench.py, not part of the library.

    python bench_profile.py stats  <kernel_stats.csv> <out.md> <title> <command>
    python bench_profile.py db     <results.db> <out.md> <title> <command>
    python bench_profile.py traffic <fetch_dir> <write_dir> <n> <k> <out.json> <command...>
    python bench_profile.py counters <pmc_dir> <out.md> <title> <command...>

`stats` condenses a `rocprofv3 --kernel-trace --stats` kernel_stats CSV into a
markdown table; `db` does the same from the rocpd SQLite database that
rocprofv3 writes by default.  `counters` tabulates any `--pmc` run's counters per field kernel (mean per
dispatch), with VALU instructions per wave.  `traffic` reads the counter_collection CSVs of two `--pmc`
runs (FETCH_SIZE, then WRITE_SIZE, in KiB per dispatch) and scales each
counter by what it reads for a copy of known size: the n^3 x 4-byte
device-to-device copies that `load_particles` makes (`__amd_rocclr_copyBuffer`
dispatches), which move exactly n^3 x 4 bytes each way.  The result is the
HBM bytes per launch of each field kernel, which bench.py reports as
`roofline.traffic`.
"""
from __future__ import annotations

import csv
import json
import re
import sqlite3
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent

# library kernel -> the stage name bench.py reports it under
STAGE_OF = {
    "k_ppart": "ppart_dens",
    "k_vvel": "vvel_sph_vw",
    "k_ddens_scatter": "ddens_scatter",
    "k_ddens_finalize": "ddens_finalize",
    "k_div_finalize": "mean_div",
    "k_mark": "mark_flags",
}


def short(name: str) -> str:
    n = name.replace("avism::(anonymous namespace)::", "").replace("avism::", "")
    n = re.sub(r"\(.*", "", n)
    if "rocprim" in n:
        n = "rocprim::" + ("radix_sort" if "radix" in n else "scan" if "scan" in n else "other")
    if n.startswith("void at::") or n.startswith("at::"):
        n = "torch elementwise (input generation)"
    if n.startswith("fft_"):
        n = "rocFFT (input generation)"
    return n.replace("void ", "")


def stats(src: str, dst: str, title: str, cmd: str) -> None:
    rows = list(csv.DictReader(open(src)))
    agg: dict[str, list] = {}
    for r in rows:
        a = agg.setdefault(short(r["Name"]), [0, 0.0])
        a[0] += int(r["Calls"])
        a[1] += float(r["TotalDurationNs"])
    tot = sum(v[1] for v in agg.values())
    out = [f"# {title}", "", "Command (run from /tmp, TMPDIR=/tmp, repository at $R):", "", "```", cmd, "```", "",
           "| kernel | calls | total ms | avg ms | % |", "|---|---:|---:|---:|---:|"]
    for name, (c, t) in sorted(agg.items(), key=lambda kv: -kv[1][1])[:25]:
        out.append(f"| {name} | {c} | {t / 1e6:.3f} | {t / 1e6 / c:.3f} | {100 * t / tot:.1f} |")
    out.append(f"| (all kernels) | | {tot / 1e6:.3f} | | 100 |")
    Path(dst).write_text("\n".join(out) + "\n")


def db_stats(src: str, dst: str, title: str, cmd: str) -> None:
    """Kernel time table from a rocpd database (the `kernels` view: one row per
    dispatch, durations in ns), with VGPRs and LDS bytes per workgroup."""
    con = sqlite3.connect(src)
    rows = con.execute("select name, duration, vgpr_count, accum_vgpr_count, lds_size from kernels").fetchall()
    agg: dict[str, list] = {}
    for name, dur, vg, ag, lds in rows:
        a = agg.setdefault(short(name), [0, 0.0, vg, ag, lds])
        a[0] += 1
        a[1] += float(dur)
    tot = sum(v[1] for v in agg.values())
    out = [f"# {title}", "", "Command (run from the repository root, TMPDIR=/tmp):", "", "```", cmd, "```", "",
           "| kernel | calls | total ms | avg ms | % | VGPR | AGPR | LDS B |", "|---|---:|---:|---:|---:|---:|---:|---:|"]
    for name, (c, t, vg, ag, lds) in sorted(agg.items(), key=lambda kv: -kv[1][1])[:30]:
        out.append(f"| {name} | {c} | {t / 1e6:.3f} | {t / 1e6 / c:.3f} | {100 * t / tot:.1f} | {vg} | {ag} | {lds} |")
    out.append(f"| (all kernels) | | {tot / 1e6:.3f} | | 100 | | | |")
    Path(dst).write_text("\n".join(out) + "\n")


def sources_hash() -> str:
    sys.path.insert(0, str(ROOT))
    import bench
    return bench.sources_hash()


def _counters(d: str, counter: str) -> list[tuple[int, str, float]]:
    f = next(Path(d).glob("**/*counter_collection.csv"))
    return [(int(r["Dispatch_Id"]), r["Kernel_Name"], float(r["Counter_Value"])) for r in csv.DictReader(open(f))
            if r["Counter_Name"] == counter]


def traffic(fetch_dir: str, write_dir: str, n: int, k: int, dst: str, cmd: str) -> dict:
    fetch = _counters(fetch_dir, "FETCH_SIZE")
    write = _counters(write_dir, "WRITE_SIZE")
    copy_kib = n ** 3 * 4 / 1024.0
    # the particle copies: copyBuffer dispatches that write exactly n^3 x 4 bytes;
    # the two runs execute the same program, so the same dispatch ids are the
    # same copies in the fetch run
    ids = {i for i, name, v in write if "copyBuffer" in name and v == copy_kib}
    fv_copy = [v for i, name, v in fetch if i in ids and "copyBuffer" in name]
    if not ids or len(fv_copy) != len(ids):
        raise SystemExit("no matching copies of n^3 x 4 bytes found for the calibration")
    wc = copy_kib
    fc = sum(fv_copy) / len(fv_copy)
    ffac, wfac = copy_kib / fc, copy_kib / wc
    kernels = {}
    for base, stage in STAGE_OF.items():
        fv = [v for _, name, v in fetch if short(name).split("<")[0] == base]
        wv = [v for _, name, v in write if short(name).split("<")[0] == base]
        if not fv or not wv:
            continue
        fk, wk = sum(fv) / len(fv), sum(wv) / len(wv)
        kernels[stage] = {"kernel": base, "n": n, "k": k, "fetch_kib": round(fk, 1), "write_kib": round(wk, 1),
                          "bytes": int(round((fk * ffac + wk * wfac) * 1024))}
    out = {"command": cmd, "sources": sources_hash(),
           "calibration": {"copy_bytes": n ** 3 * 4, "fetch_kib": round(fc, 1), "write_kib": round(wc, 1),
                           "fetch_factor": round(ffac, 4), "write_factor": round(wfac, 4)},
           "kernels": kernels}
    Path(dst).write_text(json.dumps(out, indent=1) + "\n")
    return out


def counters(pmc_dir: str, dst: str, title: str, cmd: str) -> None:
    f = next(Path(pmc_dir).glob("**/*counter_collection.csv"))
    acc: dict[str, dict[str, list[float]]] = {}
    for r in csv.DictReader(open(f)):
        base = short(r["Kernel_Name"]).split("<")[0]
        if base not in STAGE_OF:
            continue
        acc.setdefault(base, {}).setdefault(r["Counter_Name"], []).append(float(r["Counter_Value"]))
    names = sorted({c for d in acc.values() for c in d})
    out = [f"# {title}", "", "Command (run from /tmp, TMPDIR=/tmp, repository at $R):", "", "```", cmd, "```", "",
           "Mean per dispatch.", "", "| kernel | " + " | ".join(names) + " | VALU instr / wave |",
           "|---|" + "---:|" * (len(names) + 1)]
    for base, d in acc.items():
        mean = {c: sum(v) / len(v) for c, v in d.items()}
        per_wave = (mean["SQ_INSTS_VALU"] / mean["SQ_WAVES"]) if mean.get("SQ_WAVES") and "SQ_INSTS_VALU" in mean \
            else float("nan")
        out.append(f"| {base} | " + " | ".join(f"{mean.get(c, float('nan')):.4g}" for c in names) +
                   f" | {per_wave:.0f} |")
    Path(dst).write_text("\n".join(out) + "\n")


def main(argv: list[str]) -> None:
    if argv[0] == "stats":
        stats(argv[1], argv[2], argv[3], argv[4])
    elif argv[0] == "db":
        db_stats(argv[1], argv[2], argv[3], " ".join(argv[4:]))
    elif argv[0] == "counters":
        counters(argv[1], argv[2], argv[3], " ".join(argv[4:]))
    elif argv[0] == "traffic":
        print(json.dumps(traffic(argv[1], argv[2], int(argv[3]), int(argv[4]), argv[5], " ".join(argv[6:]))))
    else:
        raise SystemExit(__doc__)


if __name__ == "__main__":
    main(sys.argv[1:])
