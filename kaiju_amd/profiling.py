"""Profile helpers: a calibration copy for HBM counters and summaries of
rocprofv3 CSV output.

  python -m kaiju_amd.profiling copy [GiB]
      one plain device-to-device copy of GiB (default 1) float64 bytes, the
      yardstick for FETCH_SIZE / WRITE_SIZE (known read and write bytes).
  python -m kaiju_amd.profiling trace <kernel_stats.csv> "<command>"
      markdown table of a --kernel-trace --stats run.
  python -m kaiju_amd.profiling traffic <fetch.csv> <write.csv> <calib_fetch.csv> <calib_write.csv> <GiB> <out.json> "<commands>"
      per-stage HBM bytes per launch from two --pmc runs (FETCH_SIZE, WRITE_SIZE,
      KiB per dispatch), each scaled by what the same counter reads for the
      calibration copy; writes <out.json> (read by bench.py) and prints markdown.
  python -m kaiju_amd.profiling counters <counter_collection.csv> "<command>"
      per-stage mean per launch of every counter of one --pmc run, with the
      VALU-issue time SQ_INSTS_VALU implies when it holds SQ_INSTS_VALU.
  python -m kaiju_amd.profiling mfma <kernel_stats.csv> <N> "<command>"
      the learned-correction network's kernels of a --kernel-trace --stats run
      on an N^3 block: algorithmic FLOPs per launch (real channel counts), the
      FLOP rate of the mean launch and its fraction of the dense BF16 peak.
"""
from __future__ import annotations

import csv
import json
import sys
from collections import defaultdict

# kernel-name prefix -> stage name used by the library's stage timers
STAGE_OF = [
    ("k_clock_begin", "clock"), ("k_predict", "predictor"), ("k_b2f_shell", "bflux_post"), ("k_b2f_rows", "ring"),
    ("k_b2f", "bflux_half"), ("k_ring_", "ring"), ("k_pole_e", "ring"), ("k_fix_e", "ring"),
    ("k_celldt_rows", "celldt"), ("k_bc_wind", "bcs"), ("k_wind_", "bcs"),
    ("k_vf", "mom2face"), ("k_emf_jcol", "emf_j"), ("k_emf<0", "emf_i"), ("k_emf<1", "emf_j"), ("k_emf<2", "emf_k"),
    ("k_flux<0", "flux_i"), ("k_flux<1", "flux_j"), ("k_flux2<1", "flux_j"), ("k_flux2<2", "flux_k"),
    ("k_flux<2", "flux_k"), ("k_ct", "ct_update"), ("k_update", "update"),
    ("k_dtfinal", "dt_final"), ("k_bc", "bcs"), ("k_celldt", "celldt"),
]


def stage_of(name: str) -> str | None:
    n = name.replace("void ", "").strip()
    if n.startswith("k_packed<"):   # a packed launch: the stage of the body b_* it runs
        i = n.find("b_")
        if i < 0:
            return None
        n = "k_" + n[i + 2:]
    for pre, st in STAGE_OF:
        if n.startswith(pre):
            return st
    return None


def copy(gib: float) -> None:
    import torch
    n = int(gib * (1 << 30)) // 8
    x = torch.ones(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    y.copy_(x)
    torch.cuda.synchronize()
    print(f"copied {n * 8} bytes", flush=True)


def trace_summary(stats_csv: str, command: str) -> str:
    rows = list(csv.DictReader(open(stats_csv)))
    out = ["# rocprofv3 kernel trace summary", "", "Command:", "", "```", command, "```", "",
           "| kernel | calls | total (us) | average (us) | % |", "|---|---:|---:|---:|---:|"]
    for r in rows:
        short = r["Name"].replace("void ", "").split("(")[0]
        out.append(f"| `{short}` | {r['Calls']} | {float(r['TotalDurationNs']) / 1e3:.1f} | "
                   f"{float(r['AverageNs']) / 1e3:.1f} | {float(r['Percentage']):.2f} |")
    return "\n".join(out) + "\n"


def _counter(csv_path: str, counter: str):
    """(kernel name, KiB) per dispatch for one counter."""
    per = defaultdict(float)
    names = {}
    for r in csv.DictReader(open(csv_path)):
        if r["Counter_Name"] != counter:
            continue
        d = r["Dispatch_Id"]
        per[d] += float(r["Counter_Value"])
        names[d] = r["Kernel_Name"]
    return [(names[d], per[d]) for d in per]


def traffic(fetch_csv, write_csv, cf_csv, cw_csv, gib, out_json, command) -> str:
    nbytes = int(float(gib) * (1 << 30)) // 8 * 8
    # the calibration copy is the largest dispatch of its run
    cf = max(v for _, v in _counter(cf_csv, "FETCH_SIZE"))
    cw = max(v for _, v in _counter(cw_csv, "WRITE_SIZE"))
    sf, sw = nbytes / cf, nbytes / cw     # bytes per counted KiB
    acc = defaultdict(lambda: [0, 0.0, 0, 0.0])
    for name, v in _counter(fetch_csv, "FETCH_SIZE"):
        st = stage_of(name)
        if st:
            acc[st][0] += 1
            acc[st][1] += v * sf
    for name, v in _counter(write_csv, "WRITE_SIZE"):
        st = stage_of(name)
        if st:
            acc[st][2] += 1
            acc[st][3] += v * sw
    res = {"calibration": {"copy_bytes": nbytes, "fetch_KiB": cf, "write_KiB": cw,
                           "fetch_bytes_per_KiB": sf, "write_bytes_per_KiB": sw},
           "command": command}
    lines = ["# HBM traffic per launch (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE)", "", "Commands:", "",
             "```", command, "```", "",
             f"Calibration: a plain copy of {nbytes} B reads FETCH_SIZE = {cf:.0f} KiB and "
             f"WRITE_SIZE = {cw:.0f} KiB; factors {sf:.1f} and {sw:.1f} bytes per counted KiB.", "",
             "| stage | launches | read MB/launch | write MB/launch | total MB/launch |",
             "|---|---:|---:|---:|---:|"]
    for st in sorted(acc):
        nf, fb, nw, wb = acc[st]
        rd = fb / max(nf, 1)
        wr = wb / max(nw, 1)
        res[st] = {"launches": nf, "read_bytes_per_launch": rd, "write_bytes_per_launch": wr,
                   "bytes_per_launch": rd + wr}
        lines.append(f"| {st} | {nf} | {rd / 1e6:.1f} | {wr / 1e6:.1f} | {(rd + wr) / 1e6:.1f} |")
    with open(out_json, "w") as f:
        json.dump(res, f, indent=1)
    return "\n".join(lines) + "\n"


# MI355X: 256 CUs x 4 SIMDs; a wave64 VALU instruction occupies its SIMD for
# 4 cycles at the 2.4 GHz peak engine clock (FP64 included: 78.6 TFLOP/s of
# FMA = 1024 SIMDs x 64 lanes x 2 flops x 2.4 GHz / 4)
N_SIMD = 1024
CLOCK_HZ = 2.4e9


def counters(csv_path: str, command: str) -> str:
    rows = list(csv.DictReader(open(csv_path)))
    names = sorted({r["Counter_Name"] for r in rows})
    per = defaultdict(lambda: defaultdict(float))   # dispatch -> counter -> value
    kern = {}
    for r in rows:
        d = r["Dispatch_Id"]
        per[d][r["Counter_Name"]] += float(r["Counter_Value"])
        kern[d] = r["Kernel_Name"]
    acc = defaultdict(lambda: [0, defaultdict(float)])
    for d, vals in per.items():
        st = stage_of(kern[d])
        if not st:
            continue
        acc[st][0] += 1
        for c, v in vals.items():
            acc[st][1][c] += v
    valu = "SQ_INSTS_VALU" in names
    head = "| stage | launches | " + " | ".join(names) + (" | VALU issue (ms) |" if valu else " |")
    lines = ["# rocprofv3 counters per launch", "", "Command:", "", "```", command, "```", ""]
    if valu:
        lines += [f"VALU issue time = SQ_INSTS_VALU x 4 cycles / ({N_SIMD} SIMDs x {CLOCK_HZ / 1e9:.1f} GHz): "
                  "the time the step's vector instructions alone would take with every SIMD issuing "
                  "one each 4 cycles.", ""]
    lines += [head, "|---|---:|" + "---:|" * (len(names) + (1 if valu else 0))]
    for st in sorted(acc):
        n, tot = acc[st]
        cells = [f"{tot[c] / n:.4g}" for c in names]
        if valu:
            cells.append(f"{tot['SQ_INSTS_VALU'] / n * 4 / (N_SIMD * CLOCK_HZ) * 1e3:.3f}")
        lines.append(f"| {st} | {n} | " + " | ".join(cells) + " |")
    return "\n".join(lines) + "\n"


# MI355X dense BF16 MFMA peak (no sparsity)
BF16_PEAK_TFLOPS = 2500.0
# per network kernel: multiply-adds per computed cell and per tap (real channels:
# 8 -> 32 -> 32 -> 8); the backward-data launches alternate between the output
# layer's (8 -> 32) and the hidden layer's (32 -> 32) and are averaged
NET_MACS = {"k_conv<32, 0, 4>": 32 * 8, "k_conv<32, 0, 1>": 32 * 32, "k_conv<16, 1, 1>": 8 * 32,
            "k_conv<32, 2, 1>": (8 * 32 + 32 * 32) / 2, "k_wgrad<32, 2>": 32 * 32, "k_wgrad<16, 2>": 8 * 32,
            "k_wgrad<32, 1>": 32 * 8}
NET_ROLE = {"k_conv<32, 0, 4>": "forward, layer 1 (8 -> 32, 4 taps per K)", "k_conv<32, 0, 1>": "forward, layer 2 (32 -> 32)",
            "k_conv<16, 1, 1>": "forward, layer 3 (32 -> 8)", "k_conv<32, 2, 1>": "backward data, layers 3 and 2",
            "k_wgrad<32, 2>": "weight gradient, layer 2", "k_wgrad<16, 2>": "weight gradient, layer 3",
            "k_wgrad<32, 1>": "weight gradient, layer 1"}
# inference layers (k_infer): MACs per stored cell, role, and the layer's stored
# box, the active N^3 block widened by this many cells on every side
NET_INFER = {"k_infer<8, 32, 0, double>": (32 * 8, "inference, features + layer 1 (FP64 state)", 2),
             "k_infer<8, 32, 0, float>": (32 * 8, "inference, features + layer 1 (FP32 state)", 2),
             "k_infer<32, 32, 0, float>": (32 * 32, "inference, layer 2", 1),
             "k_infer<32, 16, 1, float>": (8 * 32, "inference, layer 3", 0)}


def mfma(stats_csv: str, N: int, command: str) -> str:
    """Network kernels' FLOP rates on an N^3 block (ghost-inclusive box N + 8)."""
    Ni = N + 8
    nc = Ni ** 3
    O = Ni * Ni + Ni + 1
    cells = 16 * ((nc - 2 * O) // 16)   # the cells every convolution computes
    lines = ["# Learned-correction network: MFMA kernels", "", "Command:", "", "```", command, "```", "",
             f"Block {N}^3 (box {Ni}^3, {cells} cells computed per training convolution; the inference layers "
             f"store ({N}+4)^3, ({N}+2)^3 and {N}^3 cells).  Algorithmic FLOPs per launch = "
             "2 x cells x 27 taps x (output x input channels, real counts 8 -> 32 -> 32 -> 8); rate = FLOPs / mean "
             f"launch time; fraction of the {BF16_PEAK_TFLOPS:.0f} TFLOP/s dense BF16 MFMA peak.", "",
             "| kernel | role | calls | mean (us) | GFLOP/launch | TFLOP/s | of peak |", "|---|---|---:|---:|---:|---:|---:|"]
    tot_f = tot_t = 0.0
    for r in csv.DictReader(open(stats_csv)):
        short = r["Name"].replace("void ", "").split("(")[0]
        if short in NET_MACS:
            macs, role, nce = NET_MACS[short], NET_ROLE[short], cells
        elif short in NET_INFER:
            macs, role, w = NET_INFER[short]
            nce = (N + 2 * w) ** 3
        else:
            continue
        t = float(r["AverageNs"]) * 1e-9
        f = 2.0 * nce * 27 * macs
        n = int(r["Calls"])
        tot_f += f * n
        tot_t += t * n
        lines.append(f"| `{short}` | {role} | {n} | {t * 1e6:.1f} | {f / 1e9:.1f} | {f / t / 1e12:.1f} | "
                     f"{f / t / 1e12 / BF16_PEAK_TFLOPS * 100:.2f}% |")
    if tot_t > 0:
        lines += ["", f"All network launches: {tot_f / 1e12:.2f} TFLOP in {tot_t * 1e3:.1f} ms = "
                      f"{tot_f / tot_t / 1e12:.1f} TFLOP/s, {tot_f / tot_t / 1e12 / BF16_PEAK_TFLOPS * 100:.2f}% of peak."]
    return "\n".join(lines) + "\n"


if __name__ == "__main__":
    cmd = sys.argv[1] if len(sys.argv) > 1 else ""
    if cmd == "copy":
        copy(float(sys.argv[2]) if len(sys.argv) > 2 else 1.0)
    elif cmd == "trace":
        print(trace_summary(sys.argv[2], sys.argv[3]), end="")
    elif cmd == "traffic":
        print(traffic(*sys.argv[2:9]), end="")
    elif cmd == "counters":
        print(counters(sys.argv[2], sys.argv[3]), end="")
    elif cmd == "mfma":
        print(mfma(sys.argv[2], int(sys.argv[3]), sys.argv[4]), end="")
    else:
        print(__doc__)
        sys.exit(2)
