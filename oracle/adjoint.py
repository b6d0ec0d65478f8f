"""TEST INFRASTRUCTURE ONLY — the adjoint's oracle.

A double-precision NumPy restatement of the wave solver's step (the reference's run(),
cpp/include/acoustic_spaces/fdtd.hpp:786-904: interior, with a speed of sound per cell or
not, special cells with their branch states, the matched layer's split fields, probes read
before each step), driven by a layout whose special
cells' wall terms and branch gains are given as doubles: coef_b = 1 - g, coef_inv =
1 / (1 + g) for each special's g (add_special, fdtd.hpp:3499-3517) and each state's gain
(fdtd.hpp:391-411). Central finite differences of this forward are the ground truth the
GPU adjoint's dJ/dg and dJ/dgain are checked against; with g and gain taken from the
layout's own float coefficients it equals the oracle's C restatement in double
(oracle/step.c) to rounding.
"""

from __future__ import annotations

import numpy as np

STEPS = np.array([
    (-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1),
    (-1, -1, 0), (1, -1, 0), (-1, 1, 0), (1, 1, 0), (-1, 0, -1), (1, 0, -1),
    (-1, 0, 1), (1, 0, 1), (0, -1, -1), (0, 1, -1), (0, -1, 1), (0, 1, 1),
])


class Forward:
    """The grid of ``layout`` stepped in double with each special's g and each state's gain
    given (see the module's notes)."""

    def __init__(self, layout: dict):
        # The matched layer (the reference's update_layer, fdtd.hpp:2242-2266): each layer
        # cell's three parts and six face velocities (its lower and upper face along each
        # axis), stepped from the pressures before the step.
        nx_, ny_, _ = (int(v) for v in layout["shape"])
        self.layer = np.asarray(layout["layer_cells"], np.int64)
        if len(self.layer):
            c = self.layer
            self.layer_at = np.stack([c % nx_, (c // nx_) % ny_, c // (nx_ * ny_)])  # (3, L)
            self.layer_stride = np.array([1, nx_, nx_ * ny_])
            axes = "xyz"
            self.keep_c = np.stack([np.asarray(layout[f"keep_{a}"], np.float64)[self.layer_at[k]] for k, a in enumerate(axes)])
            self.push_c = np.stack([np.asarray(layout[f"push_{a}"], np.float64)[self.layer_at[k]] for k, a in enumerate(axes)])
            self.keep_lo = np.stack([np.asarray(layout[f"keep_face_{a}"], np.float64)[self.layer_at[k]] for k, a in enumerate(axes)])
            self.push_lo = np.stack([np.asarray(layout[f"push_face_{a}"], np.float64)[self.layer_at[k]] for k, a in enumerate(axes)])
            self.keep_hi = np.stack([np.asarray(layout[f"keep_face_{a}"], np.float64)[self.layer_at[k] + 1] for k, a in enumerate(axes)])
            self.push_hi = np.stack([np.asarray(layout[f"push_face_{a}"], np.float64)[self.layer_at[k] + 1] for k, a in enumerate(axes)])
            if len(layout["layer_scale"]):
                self.push_c = self.push_c * np.asarray(layout["layer_scale"], np.float64)[None, :]
        # A speed of sound per cell: the interior's scaled update (fdtd.hpp:240-288),
        # (2 - total s) p + (s axial) sum [+ (s diagonal) edges] - p_prev.
        self.scale = np.asarray(layout["scale"], np.float64) if int(layout["scaled"][0]) else None
        self.total = float(layout["total"][0])
        nx, ny, nz = (int(v) for v in layout["shape"])
        self.n = nx * ny * nz
        offsets = STEPS[:, 0] + STEPS[:, 1] * nx + STEPS[:, 2] * nx * ny
        types = np.asarray(layout["types"])
        self.interior = np.flatnonzero(types == 1)
        self.wide = int(layout["stencil"][0]) == 19
        w = np.asarray(layout["weights"], np.float64)
        self.axial = w[0]
        self.diagonal = w[6] if self.wide else 0.0
        self.centre = float(layout["centre"][0])
        self.nb = self.interior[:, None] + offsets[None, :6]
        self.edges = self.interior[:, None] + offsets[None, 6:] if self.wide else None
        self.special = np.asarray(layout["special"], np.int64)
        K = len(self.special)
        ls = np.asarray(layout["link_start"], np.int64)
        self.link_owner = np.repeat(np.arange(K), np.diff(ls))
        self.link_cell = self.special[self.link_owner] + offsets[np.asarray(layout["link_direction"], np.int64)]
        self.link_weight = np.asarray(layout["link_weight"], np.float64)
        self.coef_a = np.asarray(layout["coef_a"], np.float64)
        self.source = np.asarray(layout["source"], np.float64)
        ss = np.asarray(layout["state_start"], np.int64)
        self.state_owner = np.repeat(np.arange(K), np.diff(ss))
        br = np.asarray(layout["state_branch"], np.int64)
        self.state_branch = br
        self.state_weight = np.asarray(layout["state_weight"], np.float64)
        self.now = np.asarray(layout["branch_now"], np.float64)[br]
        self.before = np.asarray(layout["branch_before"], np.float64)[br]
        ps = np.asarray(layout["probe_start"], np.int64)
        self.probe_owner = np.repeat(np.arange(len(ps) - 1), np.diff(ps))
        self.probe_cells = np.asarray(layout["probe_cells"], np.int64)
        self.probe_weights = np.asarray(layout["probe_weights"], np.float64)
        self.probes = len(ps) - 1

    def layout_coefficients(self, layout: dict):
        """g and gain as the layout's float coefficients hold them (coef_b = 1 - g)."""
        g = 1.0 - np.asarray(layout["coef_b"], np.float64)
        gain = np.asarray(layout["branch_gain"], np.float64)[self.state_branch]
        return g, gain

    def run(self, signal, g, gain, coef_b=None, coef_inv=None) -> np.ndarray:
        """The probes' pressures before each step (probes, steps)."""
        K = len(self.special)
        cb = 1.0 - g if coef_b is None else coef_b
        ci = 1.0 / (1.0 + g) if coef_inv is None else coef_inv
        a = np.zeros(self.n)
        b = np.zeros(self.n)
        wa = np.zeros(len(self.state_weight))
        wb = np.zeros(len(self.state_weight))
        L = len(self.layer)
        part, lo, hi = np.zeros((3, L)), np.zeros((3, L)), np.zeros((3, L))
        out = np.zeros((self.probes, len(signal)))
        for step, s in enumerate(np.asarray(signal, np.float64)):
            out[:, step] = np.bincount(self.probe_owner, self.probe_weights * b[self.probe_cells],
                                       minlength=self.probes)
            nxt = a.copy()
            i = self.interior
            if self.scale is None:
                total = self.centre * b[i] + self.axial * b[self.nb].sum(axis=1)
                if self.wide:
                    total = total + self.diagonal * b[self.edges].sum(axis=1)
            else:
                sc = self.scale[i]
                total = (2.0 - self.total * sc) * b[i] + (sc * self.axial) * b[self.nb].sum(axis=1)
                if self.wide:
                    total = total + (sc * self.diagonal) * b[self.edges].sum(axis=1)
            nxt[i] = total - a[i]
            sp = self.special
            links = np.bincount(self.link_owner, self.link_weight * b[self.link_cell], minlength=K)
            own = self.now * wb - self.before * wa
            flow = np.bincount(self.state_owner, self.state_weight * own, minlength=K)
            prev = a[sp]
            new = ci * (self.coef_a * b[sp] + links - cb * prev - flow + self.source * s)
            nxt[sp] = new
            wa = wa + (new - prev)[self.state_owner] * gain + own
            if L:
                c = self.layer
                for ax in range(3):
                    st = self.layer_stride[ax]
                    lo[ax] = self.keep_lo[ax] * lo[ax] - self.push_lo[ax] * (b[c] - b[c - st])
                    hi[ax] = self.keep_hi[ax] * hi[ax] - self.push_hi[ax] * (b[c + st] - b[c])
                    part[ax] = self.keep_c[ax] * part[ax] - self.push_c[ax] * (hi[ax] - lo[ax])
                nxt[c] = part[0] + part[1] + part[2]
            a, b = b, nxt
            wa, wb = wb, wa
        return out
