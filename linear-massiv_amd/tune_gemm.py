#!/usr/bin/env python3
"""Autotuner of the GEMM tile table (csrc/gemm_tiles.inc).

Times every plain-GEMM tile shape (64x64, 128x64, 128x128; the NN kernel or
the general one is chosen inside by operand layout) on the shapes the
library issues -- square products, the LU and Cholesky trailing updates
(k = 256 / 128, full and 256-wide), and the block updates of the triangular
solves with 64 right-hand sides -- in FP64 and FP32, and writes a rule for
every (element size, ceil(log2 m), ceil(log2 n), ceil(log2 k)) bin whose
fastest tile beats the fixed dispatch by at least 3%.  Rebuild the library
after writing the table.

    python linear-massiv_amd/tune_gemm.py --out linear-massiv_amd/csrc/gemm_tiles.inc
"""
import argparse
import math
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

TILES = {1: "64x64", 2: "128x64", 3: "128x128"}


def shapes():
    out = []
    for n in (512, 1024, 2048, 4096, 8192):
        out.append((n, n, n, "N"))
    for m in (1024, 2048, 4096, 8192, 16384):
        out += [(m, m, 256, "N"), (m, 256, 256, "N"), (m, m, 128, "T"), (m, 128, 128, "T"), (m, 64, 1024, "N")]
    return out


def time_gemm(torch, lmx, fn, reps=5):
    for _ in range(2):
        fn()
    torch.cuda.synchronize()
    ts = []
    for _ in range(reps):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        fn()
        b.record()
        torch.cuda.synchronize()
        ts.append(a.elapsed_time(b))
    return sorted(ts)[reps // 2]


def main():
    p = argparse.ArgumentParser()
    p.add_argument("--out", default=os.path.join(HERE, "csrc", "gemm_tiles.inc"))
    p.add_argument("--gain", type=float, default=0.03)
    a = p.parse_args()
    import torch
    import lmx
    lmx.init(0)
    rules = []
    for dtype, bytes_, call in ((torch.float64, 8, lmx.dgemm_dev), (torch.float32, 4, lmx.sgemm_dev)):
        for m, n, k, tb in shapes():
            A = torch.rand(m, k, dtype=dtype, device="cuda")
            B = torch.rand(n, k, dtype=dtype, device="cuda") if tb == "T" else torch.rand(k, n, dtype=dtype,
                                                                                         device="cuda")
            C = torch.zeros(m, n, dtype=dtype, device="cuda")
            t = {}
            for tile in (0, 1, 2, 3):
                lmx.gemm_tile_override(tile)
                t[tile] = time_gemm(torch, lmx, lambda: call(A, B, C, alpha=-1.0, beta=1.0, transb=tb))
            lmx.gemm_tile_override(0)
            best = min((1, 2, 3), key=lambda x: t[x])
            tf = {x: 2.0 * m * n * k / (t[x] * 1e-3) / 1e12 for x in t}
            print(f"fp{bytes_ * 8} {m}x{n}x{k} {tb}: " + ", ".join(
                f"{'fixed' if x == 0 else TILES[x]} {t[x]:.3f} ms ({tf[x]:.1f} TF)" for x in (0, 1, 2, 3))
                  + f" -> {TILES[best]}", flush=True)
            if t[best] < (1.0 - a.gain) * t[0]:
                lm, ln, lk = (math.ceil(math.log2(v)) for v in (m, n, k))
                rules.append((bytes_, lm, ln, lk, best, f"{m}x{n}x{k} {tb}: {t[0]:.3f} -> {t[best]:.3f} ms"))
            del A, B, C
    lines = ["// Written by linear-massiv_amd/tune_gemm.py: GEMM tile rules timed on MI355X.",
             "// {element bytes, m_lo, m_hi, n_lo, n_hi, k_lo, k_hi (ceil(log2) ranges), tile}",
             "// tile: 1 = 64 x 64, 2 = 128 x 64, 3 = 128 x 128.",
             "static const GemmTileRule kGemmTileRules[] = {"]
    for b, lm, ln, lk, tile, why in rules:
        lines.append(f"    {{{b}, {lm}, {lm}, {ln}, {ln}, {lk}, {lk}, {tile}}},  // {why}")
    if not rules:
        lines.append("    {0, 0, -1, 0, -1, 0, -1, 0},  // matches nothing: the fixed rule is fastest everywhere")
    lines.append("};")
    with open(a.out, "w") as f:
        f.write("\n".join(lines) + "\n")
    print(f"wrote {len(rules)} rules to {a.out}", flush=True)


if __name__ == "__main__":
    main()
