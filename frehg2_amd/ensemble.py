"""Ensembles of independent groundwater domains advanced as one domain.

Members share the grid shape (nx, ny, nz), the spacing and vertical layering,
the scheme options and the solver settings; soils, initial states, top and
bottom boundary values, boundary schedules and coupling state are their own.
They are stacked along y into one domain: member m takes interior rows
m (ny + 1) ... m (ny + 1) + ny - 1, and one row of fully inactive columns
separates neighbours.  Every face toward an inactive cell is sealed (the
inactive-cell seal of Predictor.cpp:375-385 / Corrector.cpp:264-278), so a
separator is an exact no-flux wall: each member evolves as it would alone,
while one step advances all of them with one launch sequence, one batched
PCG over all member columns and one mass audit.

Conditions, checked on construction: regular meshes (follow_terrain off)
with the same highest bed in every member (the mesh hangs from it), and
no-flux y sides (the sides that face the separators).  The members take one
common dt (the adaptive controller sees all of them) and the linear solve's
tolerance applies to the stacked system.
"""
from __future__ import annotations

import dataclasses
from typing import List

import numpy as np

from . import abi
from .cases import GwCase

_SHARED = ("nx", "ny", "nz", "dx", "dy", "dz", "dz_stretch", "follow_terrain",
           "terrain_layers", "specific_storage", "use_full3d", "face_conductivity",
           "reallocation", "reallocation_surplus", "dt_init", "dt_min", "dt_max", "dq_grow",
           "dq_shrink", "courant_max", "ic_form", "rtol", "atol", "max_iterations",
           "preconditioner", "precision", "initial_guess_previous", "inner_rtol", "dt",
           "adaptive")
_SOIL = ("ksx", "ksy", "ksz", "vg_alpha", "vg_n", "theta_s", "theta_r", "aev")


class Ensemble:
    """Stack compatible member cases; `case` is the stacked GwCase."""

    def __init__(self, members: List[GwCase]):
        if not members:
            raise ValueError("an ensemble needs at least one member")
        m0 = members[0]
        for m in members[1:]:
            for f in _SHARED:
                if getattr(m, f) != getattr(m0, f):
                    raise ValueError(f"ensemble members differ in {f}")
        if m0.follow_terrain:
            raise ValueError("ensembles are stacked on regular meshes only")
        tops = [float(np.max(self._bath(m))) for m in members]
        if max(tops) != min(tops):
            raise ValueError("ensemble members need the same highest bed elevation")
        for m in members:
            for which in ("ym", "yp"):
                if which in m.bc and np.any(m.bc[which][0] != abi.BC_NOFLUX):
                    raise ValueError("ensemble members need no-flux y sides")
        cpl = [m.coupling is not None for m in members]
        if any(cpl) and not all(cpl):
            raise ValueError("either every member is coupled or none is")
        if all(cpl):
            for m in members[1:]:
                for k in ("min_depth", "vent_all"):
                    if m.coupling.get(k) != m0.coupling.get(k):
                        raise ValueError(f"ensemble members differ in coupling {k}")
        self.members = members
        self.count = len(members)
        self.ny = m0.ny
        self.stride = m0.ny + 1
        self.case = self._stack()

    # ---- layout -------------------------------------------------------------
    def rows(self, m: int) -> slice:
        """Interior rows (0-based, stacked domain) of member m."""
        r0 = m * self.stride
        return slice(r0, r0 + self.ny)

    def member_view(self, field: np.ndarray, m: int) -> np.ndarray:
        """Member m's interior (ny, nx, nk) of a stacked halo-layout field."""
        r0 = 1 + m * self.stride
        return field[r0:r0 + self.ny, 1:-1]

    @staticmethod
    def _bath(c: GwCase) -> np.ndarray:
        if c.bath is not None:
            return np.asarray(c.bath, dtype=np.float64)
        return np.full((c.ny, c.nx), c.bath_const)

    # ---- stacking -----------------------------------------------------------
    def _stack(self) -> GwCase:
        ms = self.members
        m0 = ms[0]
        nx, ny, nz, M = m0.nx, m0.ny, m0.nz, self.count
        nys = M * ny + (M - 1)
        depth = sum(m0.dz * m0.dz_stretch ** k for k in range(nz))
        low = min(float(np.min(self._bath(m))) for m in ms)
        bath = np.full((nys, nx), low - depth - 1.0)  # separators: every cell above the bed
        for k, m in enumerate(ms):
            bath[self.rows(k)] = self._bath(m)

        def stack3(arrs):
            out = np.empty((nys, nx, nz))
            for k, a in enumerate(arrs):
                out[self.rows(k)] = a
                if k + 1 < M:
                    out[self.rows(k).stop] = a[-1]  # a valid value in the inactive row
            return out

        soil = {n: stack3([m.soil_array(n) for m in ms]) for n in _SOIL}
        st = dataclasses.replace(m0, ny=nys, bath=bath, **soil)
        if m0.ic_form == "water_table":
            ic = np.zeros((nys, nx))
            for k, m in enumerate(ms):
                v = m.ic_values if m.ic_values is not None else np.full((ny, nx), m.ic_const)
                ic[self.rows(k)] = v
                if k + 1 < M:
                    ic[self.rows(k).stop] = np.asarray(v)[-1]
        else:
            ic = stack3([np.broadcast_to(m.ic_values if m.ic_values is not None
                                         else m.ic_const, (ny, nx, nz)) for m in ms])
        st.ic_values = ic
        st.bc = {}
        for which in ("top", "bot", "xm", "xp"):
            if any(which in m.bc for m in ms):
                codes = np.zeros((nys, nx), dtype=np.int32)
                vals = np.zeros((nys, nx))
                for k, m in enumerate(ms):
                    if which in m.bc:
                        codes[self.rows(k)] = m.bc[which][0]
                        vals[self.rows(k)] = m.bc[which][1]
                st.bc[which] = (codes, vals)
        if any(m.bc_schedule is not None for m in ms):
            st.bc_schedule = self._schedule
        else:
            st.bc_schedule = None
        if m0.coupling is not None:
            st.coupling = self.stack_coupling([m.coupling for m in ms])
        st.name = f"{m0.name}_x{M}"
        return st

    def _schedule(self, step: int) -> dict:
        """Stacked boundary values for one step: {which index: (ny, nx) values},
        each member's scheduled rows over the stacked initial values."""
        names = ("top", "bot", "xm", "xp", "ym", "yp")
        out = {}
        for k, m in enumerate(self.members):
            if m.bc_schedule is None:
                continue
            for which, vals in m.bc_schedule(step).items():
                if which not in out:
                    base = self.case.bc.get(names[which])
                    out[which] = (np.array(base[1]) if base is not None
                                  else np.zeros((self.case.ny, self.case.nx)))
                out[which][self.rows(k)] = np.broadcast_to(vals, (self.ny, self.case.nx))
        return out

    def stack_plane(self, planes: List[np.ndarray], fill: float = 0.0) -> np.ndarray:
        """Stack members' 2D halo-layout planes ((ny+2, nx+2) each) into the
        stacked domain's plane; separator rows take ``fill``, the outer ghost
        rows the first and last members' ghost rows."""
        nx = self.members[0].nx
        nys = self.count * self.ny + (self.count - 1)
        a = np.full((nys + 2, nx + 2), float(fill))
        for k, pl in enumerate(planes):
            r0 = 1 + k * self.stride
            a[r0:r0 + self.ny] = np.asarray(pl)[1:-1]
        a[0] = np.asarray(planes[0])[0]
        a[-1] = np.asarray(planes[-1])[-1]
        return a

    def stack_coupling(self, states: List[dict]) -> dict:
        """Stack members' 2D halo-layout coupling arrays ((ny+2, nx+2) each)."""
        out = {}
        for key in ("eta", "dept", "seep_accum", "avail", "gain", "bounce"):
            if states[0].get(key) is None:
                out[key] = None
                continue
            out[key] = self.stack_plane([s[key] for s in states])
        out["min_depth"] = states[0]["min_depth"]
        out["vent_all"] = states[0]["vent_all"]
        return out

    def member_coupling(self, stacked: np.ndarray, m: int) -> np.ndarray:
        """Member m's (ny+2, nx+2) view of a stacked 2D coupling array
        (ghost rows and columns zero)."""
        nx = self.members[0].nx
        out = np.zeros((self.ny + 2, nx + 2))
        r0 = 1 + m * self.stride
        out[1:-1] = stacked[r0:r0 + self.ny]
        return out
