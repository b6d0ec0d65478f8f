"""Writes the fixture of the hall with fitted walls from the reference's own solver.

Run in a container holding the reference's sources, after `make -C oracle ref` has built
oracle/_ref/libasref.so from them:

    python tests/golden/make_fitted_hall.py

For each of the hall's two sources (acoustic_spaces_amd.scenes.hall, 10 cm cells) it lays
the grid out with the reference's fitted set-up (set_up_fitted, fdtd.hpp:970-1217) and
runs the reference's run() (fdtd.hpp:786-904) on the scene's whole drive (11,882 steps),
then writes:
- tests/golden/hall_fitted.json: the counts, the SHA-256 of the layout arrays whose bits
  come from the fit's decisions alone (cell kinds, the specials and their links'
  directions, the branch states, the probes' cells), the fit's air volume, and each
  receiver's 63/125/250 Hz band strength G and T30 of the reference's signals (by this
  package's analysis, as bench.py reads its own);
- tests/golden/hall_fitted_signals.npz: the first 1,500 steps of every receiver, float32,
  source 1 then source 2 (arrays "source1", "source2").
The fixtures hold no code: hashes, counts and numbers.
"""

from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
HERE = os.path.dirname(os.path.abspath(__file__))

PREFIX = 1500
TOPOLOGY = ("types", "special", "link_start", "link_direction", "state_start", "state_branch",
            "probe_start", "probe_cells")


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def main() -> None:
    from acoustic_spaces_amd import scenes
    from acoustic_spaces_amd.analysis import octave_results
    from acoustic_spaces_amd.wave import high_pass
    from oracle.reference import ReferenceSolver

    out, prefix = {"config": "hall, fitted walls, 10 cm"}, {}
    for k, scene in enumerate(scenes.hall()):
        ref = ReferenceSolver.from_scene(scene, threads=os.cpu_count() or 1, fitted_walls=True)
        layout = ref.layout()
        y = ref.run(scene.drive)
        rows = [octave_results(high_pass(r, scene.rate), scene.rate, bands=3) for r in y]
        entry = {
            "steps": int(len(scene.drive)),
            "air_cells": int(ref.air_cells),
            "boundary_cells": int(ref.boundary_cells),
            "branch_states": int(ref.branch_states),
            "specials": int(len(layout["special"])),
            "links": int(len(layout["link_direction"])),
            "diagonal_links": int(np.count_nonzero(layout["link_direction"] >= 6)),
            "fit_air_volume": float(ref.cut_rows(scene.mesh.materials)["stats"]["air_volume"]),
            "topology_sha256": {name: sha(layout[name]) for name in TOPOLOGY},
            "bands_hz": rows[0]["bands_hz"],
            "g_db": [r["g_db"] for r in rows],
            "t30_s": [r["t30_s"] for r in rows],
        }
        out[f"source{k + 1}"] = entry
        prefix[f"source{k + 1}"] = y[:, :PREFIX].astype(np.float32)
        print(f"source {k + 1}: {entry['specials']:,} specials, {entry['steps']} steps", flush=True)
    with open(os.path.join(HERE, "hall_fitted.json"), "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")
    np.savez_compressed(os.path.join(HERE, "hall_fitted_signals.npz"), **prefix)


if __name__ == "__main__":
    main()
