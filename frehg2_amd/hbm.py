"""Achievable HBM bandwidth on the current device (frehg_hbm_probe,
frehg_hbm_probe_sweep).

    python -m frehg2_amd.hbm [--n N] [--iterations K] [--sweep]

prints one JSON line: the GB/s of an FP64 copy and of a 6-read/2-write
stream (the PCG direction kernel's mix of compulsory arrays) over eight
arrays of N elements (default: one halo-layout field of the 512x512x100
workload), with the published MI355X HBM peak beside them.  With --sweep it
adds, for every read/write mix (copy, read-only, write-only, 6r2w, 5r1w),
the best rate over access width, cache policy and occupancy and the
configuration that reached it.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json

from . import abi

HBM_PEAK_GBS = 8000.0
FIELD_512x512x100 = 514 * 514 * 100
MIXES = {(1, 1): "copy", (1, 0): "read", (0, 1): "write", (6, 2): "6r2w", (5, 1): "5r1w"}


def measure(n: int = FIELD_512x512x100, iterations: int = 20) -> dict:
    lib = abi.load_library()
    copy, stream = C.c_double(), C.c_double()
    rc = lib.frehg_hbm_probe(int(n), int(iterations), C.byref(copy), C.byref(stream))
    if rc != 0:
        raise abi.FrehgError(rc, lib.frehg_probe_last_error().decode())
    return {"elements": int(n), "iterations": int(iterations), "copy_gbs": copy.value,
            "stream_6r2w_gbs": stream.value, "peak_gbs": HBM_PEAK_GBS,
            "copy_frac_of_peak": copy.value / HBM_PEAK_GBS,
            "stream_6r2w_frac_of_peak": stream.value / HBM_PEAK_GBS}


def sweep(n: int = FIELD_512x512x100, iterations: int = 10) -> dict:
    """Best rate per read/write mix over the probe's configurations, and the
    whole table."""
    lib = abi.load_library()
    cap = 128
    table = (abi.HbmProbeResult * cap)()
    count = C.c_int()
    rc = lib.frehg_hbm_probe_sweep(int(n), int(iterations), table, cap, C.byref(count))
    if rc != 0:
        raise abi.FrehgError(rc, lib.frehg_probe_last_error().decode())
    rows = []
    for r in table[:count.value]:
        rows.append({"mix": MIXES.get((r.reads, r.writes), f"{r.reads}r{r.writes}w"),
                     "bytes_per_lane": r.bytes_per_lane, "nontemporal": bool(r.nontemporal),
                     "blocks_per_cu": r.blocks_per_cu, "mean_ms": r.mean_ms, "gbs": r.gbs})
    best = {}
    for row in rows:
        if row["mix"] not in best or row["gbs"] > best[row["mix"]]["gbs"]:
            best[row["mix"]] = row
    return {"elements": int(n), "iterations": int(iterations), "peak_gbs": HBM_PEAK_GBS,
            "best": best, "table": rows}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=FIELD_512x512x100)
    ap.add_argument("--iterations", type=int, default=20)
    ap.add_argument("--sweep", action="store_true")
    a = ap.parse_args()
    out = measure(a.n, a.iterations)
    if a.sweep:
        out["sweep"] = sweep(a.n, max(1, a.iterations // 2))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
