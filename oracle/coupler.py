"""CPU restatement of the coupling window (test infrastructure only).

numpy forms of src/coupling/Exchange.cpp (beginWindow :21-52, applySeepage
:54-110, reduceAudit :112-128) and of the surface depth rule
src/swe/WetDry.cpp:24-43, acting in place on the coupling arrays an
oracle.OracleRichards holds.  ``OracleExchange`` plugs into
frehg2_amd.coupler.Coupler in place of the GPU kernels, so the same coupled
sequence runs on the oracle.  Only tests/ use it, as the checker.
"""
from __future__ import annotations

import numpy as np


class OracleExchange:
    def __init__(self, gw, bottom, dx, dy, min_depth):
        self.gw = gw
        self.c = gw._coupling          # the arrays the oracle step reads and mutates
        self.bottom = np.ascontiguousarray(bottom, dtype=np.float64).copy()
        self.cell = float(dx) * float(dy)
        self.min_depth = float(min_depth)
        self.qss = np.zeros_like(self.bottom)

    def refresh_depth(self):
        # WetDry.cpp:37-42: d = eta - bottom, 0 when d <= min_depth, whole plane
        d = self.c["eta"] - self.bottom
        d[d <= self.min_depth] = 0.0
        self.c["dept"][...] = d

    def begin_window(self):
        # Exchange.cpp:40-51: interior avail = dept, gain = 0; bounce cleared
        self.c["avail"][1:-1, 1:-1] = self.c["dept"][1:-1, 1:-1]
        self.c["gain"][1:-1, 1:-1] = 0.0
        if self.c.get("bounce") is not None:
            self.c["bounce"][...] = 0.0

    def apply_seepage(self, dt):
        # Exchange.cpp:74-104, cell by cell over the interior
        eta, dept, gain, acc = self.c["eta"], self.c["dept"], self.c["gain"], self.c["seep_accum"]
        applied = discarded = 0.0
        ny2, nx2 = eta.shape
        for j in range(1, ny2 - 1):
            for i in range(1, nx2 - 1):
                a = acc[j, i]
                self.qss[j, i] = a / dt
                if a < 0.0:
                    before = eta[j, i]
                    eta[j, i] += a
                    if eta[j, i] < self.bottom[j, i]:
                        eta[j, i] = self.bottom[j, i]
                    change = eta[j, i] - before
                    gain[j, i] += change
                    applied += change
                    discarded += change - a
                    acc[j, i] = 0.0
                elif dept[j, i] > self.min_depth or a > self.min_depth:
                    eta[j, i] += a
                    gain[j, i] += a
                    applied += a
                    acc[j, i] = 0.0
        return applied * self.cell, discarded * self.cell

    def reduce_audit(self):
        # Exchange.cpp:118-127; plus the volume of the surface at rest
        s = (slice(1, -1), slice(1, -1))
        gain = float(np.sum(self.c["gain"][s])) * self.cell
        held = float(np.sum(self.c["seep_accum"][s])) * self.cell
        vol = float(np.sum(self.c["eta"][s] - self.bottom[s])) * self.cell
        return gain, held, vol

    def seepage_rate(self):
        return self.qss.copy()
