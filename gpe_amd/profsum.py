"""Summarise rocprofv3 output of bench.py into the text tables kept under
profiles/.

    python -m gpe_amd.profsum KT_DIR FETCH_DIR WRITE_DIR --n 512 --out profiles/

KT_DIR holds the --kernel-trace --stats run (kt_kernel_stats.csv), FETCH_DIR
and WRITE_DIR the --pmc FETCH_SIZE and --pmc WRITE_SIZE runs (per-dispatch
counters in KiB).  Both counter runs include one plain copy kernel of known
size (bench.py --copy-probe-mib); each counter is scaled by
(bytes the copy moved) / (what the counter read for it), which is how the
per-dispatch HBM bytes ("traffic") of the stencil kernels are obtained.

    ... --sq SQ_A_DIR SQ_B_DIR

adds the SQ counter table of the same kernels from two --pmc runs (the
counter sets SQ_PASS_A and SQ_PASS_B), with rocprofv3's derived formulas:
VALUBusy = SQ_ACTIVE_INST_VALU / CUs / GRBM_GUI_ACTIVE, LdsUtil =
SQ_LDS_IDX_ACTIVE / CUs / GRBM_GUI_ACTIVE, SALUBusy = SQ_INST_CYCLES_SALU /
CUs / GRBM_GUI_ACTIVE, with GRBM_GUI_ACTIVE per XCD (rocprofv3 reports it
summed over the XCDs).
"""
from __future__ import annotations

import argparse
import csv
import glob
import json
import os
import re

# (pattern in the kernel name, kind) in match order
KINDS = [
    ("OpRk4First", "rk4_stage1"), ("OpRk4Second", "rk4_stage2"), ("OpRk4Third", "rk4_stage3"),
    ("OpRk4Last", "rk4_stage4"), ("OpAxpySelf", "euler_or_rk2_stage1"), ("OpAxpyP0", "rk2_stage2"),
    ("OpCk1", "rkck_stage1"), ("OpCkLast", "rkck_stage6_error"),
    ("OpCkMid<double, 2>", "rkck_stage2"), ("OpCkMid<float, 2>", "rkck_stage2"),
    ("OpCkMid<double, 3>", "rkck_stage3"), ("OpCkMid<float, 3>", "rkck_stage3"),
    ("OpCkMid<double, 4>", "rkck_stage4"), ("OpCkMid<float, 4>", "rkck_stage4"),
    ("OpCkMid<double, 5>", "rkck_stage5"), ("OpCkMid<float, 5>", "rkck_stage5"),
    ("OpMxRk4K<16, 24>", "mx_rk4_stage1"), ("OpMxRk4K<17, 32>", "mx_rk4_stage2_3"), ("OpMxRk4Last", "mx_rk4_stage4"),
    ("OpMxCk1", "mx_rkck_stage1"), ("OpMxCk2", "mx_rkck_stage2"), ("OpMxCkMid<3>", "mx_rkck_stage3"),
    ("OpMxCkMid<4>", "mx_rkck_stage4"), ("OpMxCkMid<5>", "mx_rkck_stage5"), ("OpMxCkLast", "mx_rkck_stage6_error"),
]
# algorithmic field bytes per cell: accesses x element size, or fixed bytes (mixed storage)
ACCESS = {"rk4_stage1": 2, "rk4_stage2": 3, "rk4_stage3": 5, "rk4_stage4": 3,
          "euler_or_rk2_stage1": 2, "rk2_stage2": 3,
          "rkck_stage1": 3, "rkck_stage2": 5, "rkck_stage3": 6, "rkck_stage4": 7, "rkck_stage5": 8,
          "rkck_stage6_error": 7}
MIXED_BYTES = {"mx_rk4_stage1": 24, "mx_rk4_stage2_3": 32, "mx_rk4_stage4": 56,
               "mx_rkck_stage1": 24, "mx_rkck_stage2": 40, "mx_rkck_stage3": 56, "mx_rkck_stage4": 64,
               "mx_rkck_stage5": 72, "mx_rkck_stage6_error": 72}
FUSED_ACCESS = {"rk4_fused12": 2, "rk4_fused34": 3, "rk2_fused": 2, "euler2_fused": 2,  # field accesses per cell
                "rkck_fused12": 4, "rkck_fused34": 7, "rkck_fused56_error": 7}
CK_PAIR = {"1": "rkck_fused12", "2": "rkck_fused34", "3": "rkck_fused56_error"}
MIXED_BYTES.update({"mx_rkck_fused12": 40, "mx_rkck_fused34": 64, "mx_rkck_fused56_error": 72})
# kernel name forms the summaries cover
KERNEL_FORMS = ("k_stencil", "k_rk4_f", "k_rk2_fused", "k_euler2_fused", "k_ck_pair", "k_ckmx_pair")
# kernels whose calibrated traffic bench.py reads back (traffic_<kind>_<prec>_<n>.json)
TRAFFIC_KINDS = {"rk4_fused12", "rk4_fused34", "rk4_stage3", "mx_rk4_stage4"}


def cell_bytes(kind: str, prec: str) -> int:
    if kind in FUSED_ACCESS:
        return FUSED_ACCESS[kind] * (16 if prec == "c128" else 8)
    if prec == "mixed":
        return MIXED_BYTES.get(kind, 0)
    return ACCESS.get(kind, 0) * (16 if prec == "c128" else 8)


def _one(pattern):
    files = sorted(glob.glob(pattern, recursive=True))
    if not files:
        raise SystemExit(f"no file matches {pattern}")
    return files[0]


def _gm(g) -> str:
    return "" if g in (None, "1") else (" imaginary-dt" if g == "2" else " complex-dt")


def kernel_label(name: str) -> str:
    m = re.search(r"k_stencil<(\w+), (\d), (\d+), (\d+), gpe::(\w+)", name)
    ms = re.search(r"k_stencil_sum<(\d), (\d+), (\d+), gpe::(\w+)", name)
    if m or ms:
        if ms:
            prec, order, tx, ty, op = "mixed", ms.group(1), ms.group(2), ms.group(3), ms.group(4)
        else:
            prec = "c128" if m.group(1) == "double" else "c64"
            order, tx, ty, op = m.group(2), m.group(3), m.group(4), m.group(5)
        kind = next((v for k, v in KINDS if name.find(k) >= 0), op)
        if kind.startswith("mx_"):
            prec = "mixed"
        return f"{kind} {prec} order{order} tile{tx}x{ty}"
    # the last template argument of the fused kernels is the form of the step
    # constant G = i pim dt: 1 real dt (real time), 2 imaginary dt, 0 complex
    mf = re.search(r"k_rk4_f(12|34)<(double|float), (\d), (\d+)(?:, (\d))?>", name)
    if mf:
        prec = "c128" if mf.group(2) == "double" else "c64"
        h = int(mf.group(3)) // 2
        trap = " with-trap" if mf.group(4) != "0" else ""
        return f"rk4_fused{mf.group(1)} {prec} order{mf.group(3)} tile{64 + 2 * h}x15{trap}{_gm(mf.group(5))}"
    mp = re.search(r"k_(rk2_fused|euler2_fused)<(double|float), (\d), (\d+)(?:, (\d))?>", name)
    if mp:
        prec = "c128" if mp.group(2) == "double" else "c64"
        h = int(mp.group(3)) // 2
        trap = " with-trap" if mp.group(4) != "0" else ""
        return f"{mp.group(1)} {prec} order{mp.group(3)} tile{64 + 2 * h}x15{trap}{_gm(mp.group(5))}"
    mc = re.search(r"k_ck_pair<(double|float), (\d), (\d+), (\d)(?:, (\d))?>", name)
    if mc:
        prec = "c128" if mc.group(1) == "double" else "c64"
        h = int(mc.group(2)) // 2
        trap = " with-trap" if mc.group(3) != "0" else ""
        return f"{CK_PAIR[mc.group(4)]} {prec} order{mc.group(2)} tile{64 + 2 * h}x15{trap}{_gm(mc.group(5))}"
    mm = re.search(r"k_ckmx_pair<(\d), (\d+), (\d)(?:, (\d))?>", name)
    if mm:
        h = int(mm.group(1)) // 2
        trap = " with-trap" if mm.group(2) != "0" else ""
        return f"mx_{CK_PAIR[mm.group(3)]} mixed order{mm.group(1)} tile{64 + 2 * h}x15{trap}{_gm(mm.group(4))}"
    m = re.search(r"gpe::(\w+)", name)
    return m.group(1) if m else name[:60]


SQ_PASS_A = ("SQ_WAVE_CYCLES SQ_BUSY_CYCLES SQ_ACTIVE_INST_VALU SQ_ACTIVE_INST_LDS SQ_LDS_IDX_ACTIVE "
             "SQ_LDS_BANK_CONFLICT SQ_WAIT_INST_LDS SQ_WAIT_ANY GRBM_GUI_ACTIVE")
SQ_PASS_B = ("SQ_INSTS_VALU SQ_INSTS_LDS SQ_INSTS_SALU SQ_INST_CYCLES_SALU SQ_LDS_DATA_FIFO_FULL "
             "SQ_LDS_ADDR_CONFLICT SQ_WAVES SQ_INSTS_VMEM_RD GRBM_GUI_ACTIVE")


def sq_counters(d):
    """{kernel name: {counter: mean per dispatch}} and (XCDs, CUs) of one --pmc run."""
    per = {}
    for r in csv.DictReader(open(_one(os.path.join(d, "**", "*counter_collection.csv")))):
        k = (r["Kernel_Name"], r["Dispatch_Id"])
        per.setdefault(k, {})[r["Counter_Name"]] = float(r["Counter_Value"])
    out = {}
    for (name, _), cs in per.items():
        acc = out.setdefault(name, {"_n": 0})
        acc["_n"] += 1
        for c, v in cs.items():
            acc[c] = acc.get(c, 0.0) + v
    for acc in out.values():
        n = acc.pop("_n")
        for c in acc:
            acc[c] /= n
    agent = next(r for r in csv.DictReader(open(_one(os.path.join(d, "**", "*agent_info.csv"))))
                 if r["Agent_Type"] == "GPU")
    return out, (int(agent["Num_Xcc"]), int(agent["Cu_Count"]))


def sq_lines(a_dir, b_dir, run_cmd):
    A, (nxcc, ncu) = sq_counters(a_dir)
    B, _ = sq_counters(b_dir)
    lines = ["", "## SQ counters", "",
             "    rocprofv3 --pmc " + SQ_PASS_A + " -d <SQ_A_DIR> -o c --output-format csv -- \\",
             f"        {run_cmd}",
             "    rocprofv3 --pmc " + SQ_PASS_B + " -d <SQ_B_DIR> -o c --output-format csv -- \\",
             f"        {run_cmd}", "",
             f"{nxcc} XCDs, {ncu} CUs.  Busy figures are fractions of the kernel's GPU-active cycles; the wait "
             "figures are fractions of the waves' own cycles; instructions are per wave.", "",
             "| kernel | VALUBusy | LdsUtil | LDS bank-conflict stall | SALUBusy | waves waiting (any) | "
             "waves waiting for LDS issue | VALU / LDS / SALU / VMEM-read instructions per wave |",
             "|---|---|---|---|---|---|---|---|"]
    for name, a in A.items():
        if not any(f in name for f in KERNEL_FORMS) or name not in B:
            continue
        label = kernel_label(name)
        if len(label.split()) < 2:
            continue
        b = B[name]
        ga, gb = a["GRBM_GUI_ACTIVE"] / nxcc, b["GRBM_GUI_ACTIVE"] / nxcc
        w = b["SQ_WAVES"]
        lines.append(
            f"| {label} | {100 * a['SQ_ACTIVE_INST_VALU'] / ncu / ga:.1f}% | {100 * a['SQ_LDS_IDX_ACTIVE'] / ncu / ga:.1f}% | "
            f"{100 * a['SQ_LDS_BANK_CONFLICT'] / ncu / ga:.1f}% | {100 * b['SQ_INST_CYCLES_SALU'] / ncu / gb:.1f}% | "
            f"{100 * a['SQ_WAIT_ANY'] / a['SQ_WAVE_CYCLES']:.1f}% | {100 * a['SQ_WAIT_INST_LDS'] / a['SQ_WAVE_CYCLES']:.1f}% | "
            f"{b['SQ_INSTS_VALU'] / w:.0f} / {b['SQ_INSTS_LDS'] / w:.0f} / {b['SQ_INSTS_SALU'] / w:.0f} / "
            f"{b['SQ_INSTS_VMEM_RD'] / w:.0f} |")
    return lines


def counters(path):
    per = {}
    for r in csv.DictReader(open(path)):
        per.setdefault(r["Kernel_Name"], []).append(float(r["Counter_Value"]) * 1024.0)
    return {k: sum(v) / len(v) for k, v in per.items()}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("kt")
    ap.add_argument("fetch")
    ap.add_argument("write")
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--copy-mib", type=int, default=1024)
    ap.add_argument("--out", default="profiles")
    ap.add_argument("--workload", choices=["rk4", "rkck"], default="rk4",
                    help="rk4: bench.py runs; rkck: gpe_amd.studies attempts runs")
    ap.add_argument("--sq", nargs=2, metavar=("SQ_A_DIR", "SQ_B_DIR"), default=None,
                    help="the two SQ counter runs (SQ_PASS_A, SQ_PASS_B) of the same command")
    a = ap.parse_args(argv)

    stats = list(csv.DictReader(open(_one(os.path.join(a.kt, "**", "*kernel_stats.csv")))))
    fetch = counters(_one(os.path.join(a.fetch, "**", "*counter_collection.csv")))
    write = counters(_one(os.path.join(a.write, "**", "*counter_collection.csv")))
    copy_bytes = a.copy_mib << 20
    kcopy = next(k for k in fetch if "k_copy16" in k)
    f_scale = copy_bytes / fetch[kcopy]
    w_scale = copy_bytes / write[kcopy]
    cells = a.n ** 3
    os.makedirs(a.out, exist_ok=True)
    if a.workload == "rk4":
        title = f"# RK4 step kernels, {a.n}^3 random-phase workload (bench.py)"
        legs = "--cpu-steps 0 --ref-steps 0 --rk45-steps 0 --ttt-T 0 --ens-steps 0"
        run_kt = f"python3 bench.py --steps 20 --warmup 5 {legs}"
        run_pmc = f"python3 bench.py --steps 2 --warmup 1 {legs} --copy-probe-mib {a.copy_mib}"
    else:
        title = f"# RKF45 (Cash-Karp) stage and fused kernels, {a.n}^3 random-phase workload, tolerance 1e-6"
        run_kt = f"python3 -m gpe_amd.studies attempts --n {a.n} --attempts 6 --storages c128/stages,c128,mixed/stages,mixed,c64/stages,c64"
        run_pmc = (f"python3 -m gpe_amd.studies attempts --n {a.n} --attempts 1 --storages c128/stages,c128,mixed/stages,mixed,c64/stages,c64 "
                   f"--copy-probe-mib {a.copy_mib}")
    lines = [title, "",
             "Commands (run from the repository root; TMPDIR=/tmp):", "",
             "    rocprofv3 --kernel-trace --stats -d <KT_DIR> -o kt --output-format csv -- \\",
             f"        {run_kt}",
             "    rocprofv3 --pmc FETCH_SIZE -d <FETCH_DIR> -o fetch --output-format csv -- \\",
             f"        {run_pmc}",
             "    rocprofv3 --pmc WRITE_SIZE -d <WRITE_DIR> -o write --output-format csv -- \\",
             f"        {run_pmc}",
             f"    python3 -m gpe_amd.profsum <KT_DIR> <FETCH_DIR> <WRITE_DIR> --n {a.n} --workload {a.workload}", "",
             f"Counter calibration with a plain copy of {a.copy_mib} MiB (k_copy16): FETCH_SIZE read "
             f"{fetch[kcopy] / 2**20:.1f} MiB -> scale {f_scale:.4f}; WRITE_SIZE read {write[kcopy] / 2**20:.1f} MiB "
             f"-> scale {w_scale:.4f}.", "",
             "| kernel | calls | avg ms | algorithmic GB/launch | algorithmic GB/s | % of 8 TB/s | HBM read GB (cal.) | "
             "HBM write GB (cal.) | traffic / algorithmic |",
             "|---|---|---|---|---|---|---|---|---|"]
    traffic = {}
    for r in stats:
        name = r["Name"]
        if not any(f in name for f in KERNEL_FORMS):
            continue
        label = kernel_label(name)
        if len(label.split()) < 2:  # not a kernel of a known form (e.g. the batched ensemble's)
            continue
        kind, prec = label.split()[:2]
        alg = cell_bytes(kind, prec) * cells
        avg_ms = float(r["AverageNs"]) / 1e6
        fk = next((v for k, v in fetch.items() if k.startswith(name[:120])), None)
        wk = next((v for k, v in write.items() if k.startswith(name[:120])), None)
        rd = fk * f_scale if fk else None
        wr = wk * w_scale if wk else None
        t = (rd + wr) if (rd and wr) else None
        if t and a.workload == "rk4" and kind in TRAFFIC_KINDS:
            traffic[prec] = t
            json.dump({"traffic_bytes_per_launch": t, "read_bytes": rd, "write_bytes": wr,
                       "algorithmic_bytes_per_launch": alg, "fetch_scale": f_scale, "write_scale": w_scale,
                       "kernel": label},
                      open(os.path.join(a.out, f"traffic_{kind}_{prec}_{a.n}.json"), "w"), indent=1)
        lines.append(f"| {label} | {r['Calls']} | {avg_ms:.4f} | {alg / 1e9:.3f} | "
                     f"{alg / (avg_ms * 1e-3) / 1e9:.0f} | {100 * alg / (avg_ms * 1e-3) / 8e12:.1f} | "
                     f"{(rd or 0) / 1e9:.3f} | {(wr or 0) / 1e9:.3f} | {(t / alg) if (t and alg) else float('nan'):.3f} |")
    if a.sq:
        lines += sq_lines(a.sq[0], a.sq[1], run_pmc.replace(f" --copy-probe-mib {a.copy_mib}", ""))
    out = os.path.join(a.out, f"{a.workload}_stencil_{a.n}.md")
    open(out, "w").write("\n".join(lines) + "\n")
    print(open(out).read())
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
