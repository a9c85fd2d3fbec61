"""Achievable HBM bandwidth on the current device (frehg_hbm_probe).

    python -m frehg2_amd.hbm [--n N] [--iterations K]

prints one JSON line: the GB/s of an FP64 copy and of a 6-read/2-write
stream (the PCG direction kernel's mix of compulsory arrays) over eight
arrays of N elements (default: one halo-layout field of the 512x512x100
workload), with the published MI355X HBM peak beside them.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json

from . import abi

HBM_PEAK_GBS = 8000.0
FIELD_512x512x100 = 514 * 514 * 100


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=FIELD_512x512x100)
    ap.add_argument("--iterations", type=int, default=20)
    a = ap.parse_args()
    print(json.dumps(measure(a.n, a.iterations)), flush=True)


if __name__ == "__main__":
    main()
