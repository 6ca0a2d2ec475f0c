"""Measures the scan tile table for the local GPU and writes tune/<arch>.txt.

For each sample count in --samples, epi_autotune times every VALU tile
shape and the int8 and FP4 MFMA kernels on a synthetic HWE cohort of --snps SNPs and
returns the fastest; with --report every shape's kernel time is printed as
well (one JSON line per sample count).  The table maps a cohort's sample
count to the tile of the nearest measured count (bucket edges at the
geometric midpoints); the library reads it when a context is opened.

    python3 cross-dpc-episdet_amd/tune/make_table.py --arch gfx950 --report
"""
import argparse
import json
import math
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import episdet as epi  # noqa: E402

VALU_SHAPES = [{"variant": 1, "snps_j": j, "waves": w} for j in (1, 2, 4, 8) for w in (4, 8, 16)]
MFMA_SHAPES = [{"variant": 2}, {"variant": 3}]


def shape_times(ctx, n_samples, n_snps, reps=3):
    g, p = epi.generate_hwe(0x7475E5, n_samples, n_snps, n_samples // 2)
    out = []
    with ctx.cohort(g, p) as co:
        ref = None
        for t in VALU_SHAPES + MFMA_SHAPES:
            best = math.inf
            for _ in range(reps):
                r = ctx.scan(co, tile=t)
                best = min(best, r.kernel_ms)
            key = (r.i, r.j, r.k, r.score)
            if ref is None:
                ref = key
            elif key != ref:
                raise RuntimeError(f"tile {t} gave {key}, expected {ref}")
            out.append((dict(t), best))
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--samples", type=int, nargs="+",
                    default=[256, 512, 1024, 2048, 4096, 8192, 16384, 32768])
    ap.add_argument("--snps", type=int, default=1024)
    ap.add_argument("--arch", default="gfx950")
    ap.add_argument("--out", default=None, help="default: tune/<arch>.txt beside this script")
    ap.add_argument("--report", action="store_true", help="also time every shape")
    ap.add_argument("--device", type=int, default=0)
    a = ap.parse_args()
    ns = sorted(set(a.samples))
    rows = []
    with epi.Context(a.device) as ctx:
        ctx.tune_set([])  # measure from the untuned state
        for n in ns:
            tile, ms = ctx.autotune(n, a.snps)
            line = {"samples": n, "snps": a.snps, "best": tile, "kernel_ms": round(ms, 4)}
            if a.report:
                line["shapes"] = [[t, round(x, 4)] for t, x in shape_times(ctx, n, a.snps)]
            print(json.dumps(line), flush=True)
            rows.append((n, tile))
        device = ctx.device_name
    out = a.out or os.path.join(HERE, f"{a.arch}.txt")
    with open(out, "w") as f:
        f.write(f"# scan tile table for {a.arch} ({device}), measured by tune/make_table.py\n")
        f.write(f"# on HWE cohorts of {a.snps} SNPs at sample counts {' '.join(map(str, ns))}\n")
        f.write("# max_samples snps_j waves variant (1 = VALU, 2 = int8 MFMA, 3 = FP4 MFMA)\n")
        for e, (n, tile) in enumerate(rows):
            edge = int(math.sqrt(n * ns[e + 1])) if e + 1 < len(ns) else 4294967295
            f.write(f"{edge} {tile['snps_j']} {tile['waves']} {tile['variant']}\n")
    print(f"wrote {out}", flush=True)


if __name__ == "__main__":
    main()
