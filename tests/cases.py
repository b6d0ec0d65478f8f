"""Small scenes shared by the tests: each a Scene (acoustic_spaces_amd.scenes) sized so the
oracle steps it in seconds, covering what the reference's own tests cover (both stencils,
branch states, a speed per cell, the matched layer, probes next to the source)."""

from __future__ import annotations

import numpy as np

from acoustic_spaces_amd import scenes


def with_branches(scene, per_material=((0.0, 5.0, 0.3), (0.002, 3.0, 0.5))):
    """The scene with a relaxation and a resonant branch on every material."""
    rows, owners = [], []
    for m in range(scene.mesh.materials):
        for r in per_material:
            rows.append(r)
            owners.append(m)
    scene.branches = np.array(rows, float)
    scene.branch_material = np.array(owners, np.int64)
    return scene


def small_box(stencil=7, seed=3, cell=0.1, size=(1.9, 1.4, 1.1), branches=True):
    sc = scenes.shoebox(size=size, cell=cell, seed=seed, length_s=0.1, stencil=stencil)
    return with_branches(sc) if branches else sc


def open_block(layer=4, seed=5, cell=0.1):
    """A rigid block in open air with a matched layer round the grid: the source and the
    probes outside the block."""
    rng = np.random.default_rng(seed)
    block = scenes.box_mesh((0.6, 0.5, 0.4), low=(0.4, 0.4, 0.3), inward=False)
    grid = scenes.grid_round((0.0, 0.0, 0.0), (1.4, 1.3, 1.0), cell, layer=layer)
    rate = scenes.C / (scenes.COURANT * cell)
    steps = 300
    source = np.array([0.15, 0.2, 0.5]) + rng.uniform(-0.02, 0.02, 3)
    probes = np.array([[1.2, 1.1, 0.8], [0.2, 1.1, 0.2], [1.25, 0.2, 0.5]])
    sc = scenes.Scene(
        name="open block", mesh=block, beta=np.full(block.materials, 0.05),
        branches=np.zeros((0, 3)), branch_material=np.zeros(0, np.int64), grid=grid,
        source=source, probes=probes, rate=rate, steps=steps,
        drive=scenes.gaussian_drive(rate, steps, sigma_s=2e-3), layer=layer, layer_damping=0.5,
    )
    return with_branches(sc, per_material=((0.0, 4.0, 0.2),))


def speeds_for(scene, seed=7):
    """A speed of sound per cell in (0.9, 1]: s = r^2 scales each cell's links."""
    rng = np.random.default_rng(seed)
    n = scene.grid.cells
    return (0.9 + 0.1 * rng.random(n)).astype(np.float32)


def probes_near_source(seed=9):
    """Probes within three cells of the source (read trilinearly) and next to walls."""
    sc = small_box(seed=seed)
    s = sc.source
    c = sc.grid.cell
    sc.probes = np.array([s + [1.3 * c, 0.4 * c, -0.7 * c], s + [-2.1 * c, 1.1 * c, 0.3 * c],
                          [0.03, 0.71, 0.55], [1.0, 0.05, 1.07]])
    return sc


def coarse_hall(cell=0.4, length_s=0.4):
    """The hall (tilted ceiling planes, splayed side walls) at 40 cm: 80 x 68 x 35 cells."""
    sc = scenes.hall(cell=cell, length_s=length_s)[0]
    return sc


def listeners13(stencil=7, seed=31):
    """A box whose probes are two listeners' 13 points each (the listener, one cell and
    two cells either side along each axis), one listener a cell from a wall."""
    from acoustic_spaces_amd.wave import probe_points

    sc = small_box(stencil=stencil, seed=seed)
    c = sc.grid.cell
    listeners = [np.array([1.21, 0.73, 0.52]), np.array([0.5 * c + 0.013, 0.61, 0.44])]
    sc.probes = np.vstack([probe_points(p, c) for p in listeners])
    return sc


def baffled_room(stencil=7, seed=41, cell=0.1):
    """A 3 x 2 x 1.6 m room split by two baffles into an S-shaped channel (each baffle a
    solid slab through the floor, the ceiling and one side wall): the air reaches the far
    end only round two bends."""
    rng = np.random.default_rng(seed)
    room = scenes.box_mesh((3.0, 2.0, 1.6))
    a = scenes.box_mesh((0.14, 1.65, 2.0), low=(0.93, -0.2, -0.2), inward=False)
    b = scenes.box_mesh((0.14, 1.65, 2.0), low=(1.93, 0.55, -0.2), inward=False)
    mesh = scenes.merge(room, a, b)
    grid = scenes.grid_round((0.0, 0.0, 0.0), (3.0, 2.0, 1.6), cell)
    courant = scenes.COURANT if stencil == 7 else scenes.ISOTROPIC_COURANT
    rate = scenes.C / (courant * cell)
    steps = 200
    source = np.array([0.41, 0.52, 0.77]) + rng.uniform(-0.02, 0.02, 3)
    probes = np.array([[0.45, 1.71, 0.63], [1.52, 0.33, 0.91], [2.61, 1.42, 1.13]])
    sc = scenes.Scene(
        name="baffled room", mesh=mesh, beta=np.full(mesh.materials, 0.05),
        branches=np.zeros((0, 3)), branch_material=np.zeros(0, np.int64), grid=grid,
        source=source, probes=probes, rate=rate, steps=steps,
        drive=scenes.gaussian_drive(rate, steps), stencil=stencil,
    )
    return with_branches(sc, per_material=((0.0, 4.0, 0.2),))
