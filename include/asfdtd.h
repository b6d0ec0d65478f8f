/*
 * asfdtd.h — C ABI of the MI355X wave-solver library (libasfdtd.so).
 *
 * Two halves:
 *
 *  1. The stepper. It replaces the reference's WebGPU stepper GpuWaveSolver
 *     (src/acoustic_spaces/engine/gpu.py:445-612), which wave_responses picks at
 *     `stepper = GpuWaveSolver(solver) if on_gpu else solver` (engine/wave.py:1178). Like
 *     it, it takes the solver's layout (WaveSolver::layout, cpp/include/acoustic_spaces/
 *     fdtd.hpp:727-782) once and steps it: run(signal) records every probe's pressure
 *     before each step and keeps the state across calls (WaveSolver::run, fdtd.hpp:786-904).
 *
 *  2. The set-up. It replaces the reference's WaveSolver constructor on walls on cell
 *     faces (fdtd.hpp:360-690; bound as _native.WaveSolver, cpp/bindings/
 *     native_module.cpp:1238-1285): from a triangle mesh, materials and a grid it lays out
 *     the grid and hands out the same layout.
 *
 * Plain pointers and sizes only. Functions return 0 on success and a negative code on
 * failure; the message is then in as_fdtd_last_error / as_setup_last_error. A handle is
 * used by one host thread at a time.
 */
#ifndef ASFDTD_H
#define ASFDTD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Error codes. */
#define AS_OK 0
#define AS_EINVAL (-1)   /* an argument or the layout is refused (reference: ValueError) */
#define AS_EDEVICE (-2)  /* a HIP call failed (no device, out of memory, launch error) */
#define AS_ERANGE (-3)   /* a grid or table beyond the library's 32-bit indices */

/*
 * The layout: the arrays WaveSolver::layout hands out (fdtd.hpp:727-782), each a pointer
 * and its length in elements, in the reference's types. keep/push are per axis, the
 * centres' then the faces' coefficients: keep_x, keep_face_x, keep_y, keep_face_y,
 * keep_z, keep_face_z (each of length n_axis + 1, or 0 without a matched layer).
 * Moving air (fdtd.hpp:1680-2120), where the layout has it (flow_mask not NULL): plain,
 * flow_x and flow_y (each cell's shifts a step) with flow_mask the row masks; or the wind
 * round the solids, wind[3] (the wind on each cell's +x, +y, +z faces) and loop[3], with
 * flow_mask 1 where the air moves the cell and 2 at the border of the moving air. The
 * reference steps moving air on its CPU only (engine/wave.py:1159-1178); here the GPU does.
 */
typedef struct as_layout {
    uint64_t nx, ny, nz;
    uint64_t stencil;              /* 7 or 19 */
    const float *weights;          /* [6] or [18]: the update's link weights */
    uint64_t n_weights;
    float centre, total;
    const float *scale;            /* [nx ny nz] each cell's s, or NULL (layout "scaled" 0) */
    const float *layer_scale;      /* [layer_cells] or NULL */
    const uint8_t *types;          /* [nx ny nz] 0 solid, 1 interior, 2 special, 3 layer */
    uint64_t specials;
    const int64_t *special;        /* [specials] cell of each special */
    const uint64_t *link_start;    /* [specials + 1] */
    uint64_t links;
    const uint8_t *link_direction; /* [links] 0..17 */
    const float *link_weight;      /* [links] */
    const float *coef_a, *coef_b, *coef_inv, *source; /* [specials] */
    const uint64_t *state_start;   /* [specials + 1] */
    uint64_t states;
    const int32_t *state_branch;   /* [states] */
    const float *state_weight;     /* [states] */
    uint64_t branches;
    const float *branch_now, *branch_before, *branch_gain; /* [branches] */
    uint64_t layer_cells;
    const int64_t *layer_cell;     /* [layer_cells] */
    const float *keep[6];
    const float *push[6];
    uint64_t probes;
    const uint64_t *probe_start;   /* [probes + 1] */
    uint64_t taps;
    const uint64_t *probe_cells;   /* [taps] */
    const double *probe_weights;   /* [taps] */
    /* Optional, with precision 64 only: coef_b, coef_inv (specials) and branch_gain
     * (branches) in double, for a scheme whose coefficients are not rounded to float (an
     * FP64 reference, finite differences); NULL: the floats above. Given, every special is
     * stepped by its own blocks. */
    const double *coef_b64, *coef_inv64, *branch_gain64;
    /* Moving air, or NULLs (see above): [nx ny nz] each. */
    const float *flow_x, *flow_y;
    const uint8_t *flow_mask;
    const float *wind[3], *loop[3];
} as_layout;

typedef struct as_fdtd as_fdtd;

/* Launch shape of the step kernels (0 = the library's default for the field). */
typedef struct as_tuning {
    int tile_y;        /* rows of a tile (the tile is 64 cells along x) */
    int tile_z;        /* planes a tile marches through */
    int row_align;     /* row pitch in floats: a multiple of this (1, 16, 32, 64) */
    int graph_steps;   /* steps a captured hipGraph replays (0 = 64; negative: no graphs) */
    int split;         /* 1: one launch per kind of cell (specials, interior, layer, probes) */
    int tile_order;    /* 0: interior tiles grouped by die (XCD), 1: in grid order, 2: grouped by
                          die, each die's in columns (z fastest, then x, then y) */
    int no_faces;      /* 1: step every special cell by its own blocks (as the adjoint needs) */
    int face_blocks;   /* the face specials (axial links of the interior's weight): positive, by
                          blocks of their own beside the tiles; negative, each tile's in the
                          tile after its column march; 0, the library's choice (blocks of their
                          own unless the grid is small) */
    int pairs;         /* 1: two steps by two launches where that applies (precision 32, 7 points,
                          one speed of sound): the first steps every interior cell twice where
                          its neighbours are interior, by boxes of 60 x 4 cells marching 8
                          planes (16 here asks for 16); 0: one step a launch */
    int keep_history;  /* the adjoint's record (as_fdtd_record): 1 keeps every step's history of
                          the special cells (steps x specials values), so that the backward
                          pass needs no recompute; -1 keeps only the checkpoints, and the
                          backward pass recomputes each segment from its own; 0, the library's
                          choice: the history where it takes at most 16 GiB and a quarter of
                          the free device memory */
    int special_order; /* the general specials' order in their blocks: 1 by their counts of links
                          and states, 2 by cell (the layout's order), 3 by the interior tile
                          holding the cell; 0, the library's choice (3) */
    int special_kinds; /* 2: the general specials' coefficients, links and states read from tables
                          of their distinct kinds (specials alike in all but their cell and their
                          states' velocities share a row), each special a 16-byte record; 1: each
                          special's own, from its block's slabs; 0, the library's choice (the
                          kinds where a kind serves four specials or more on average) */
} as_tuning;

/* Copies the layout to the device; precision 32 (the reference's float scheme) or 64 (the
 * same scheme in double, coefficients widened). NULL on failure (see as_fdtd_create_error).
 * Replaces GpuWaveSolver.__init__ (engine/gpu.py:445-545).
 * Device memory, e = 4 or 8 bytes and n the padded grid's cells (rows padded to 64 cells):
 * two fields of n e (four with pairs), a class byte per cell, two (four) slots of branch
 * velocities, and with a matched layer 3 e per layer cell plus two slots of 3 n e for its
 * face velocities (one per face, shared by the two cells beside it). */
as_fdtd *as_fdtd_create(const as_layout *layout, int precision, const as_tuning *tuning);
const char *as_fdtd_create_error(void);

/* Advances `steps` steps, adding signal[k] at the source on step k, and writes every
 * probe's pressure before each step to out[probe * steps + k]; host pointers. The state
 * continues across calls. Replaces GpuWaveSolver.run (engine/gpu.py:546-612) and
 * WaveSolver::run (fdtd.hpp:786-904). */
int as_fdtd_run(as_fdtd *h, const float *signal, size_t steps, double *out);

/* The same on device pointers, enqueued on the handle's stream without waiting. */
int as_fdtd_run_device(as_fdtd *h, const float *d_signal, size_t steps, double *d_out);

/* as_fdtd_run without allocating: no hipMalloc, hipFree, graph capture or other allocating
 * call, for callers that cannot allocate (code in a tight serving loop, a harness that
 * forbids device allocation after set-up). It has no counterpart in the reference. The
 * drive and the probes' output go, piece by piece, through one file-scope device array of
 * 1 MiB per device (held until the process ends); the steps are launched one a step from
 * the host. Results equal as_fdtd_run's bit for bit. Refused (AS_EINVAL) when one step's
 * probes do not fit it: more than 131,040 probes. Nothing in the array must start at zero
 * (each piece writes what it reads). Every call on one device shares that device's array,
 * so calls that overlap in time on one device (two host threads, or two handles whose
 * streams run at once) race on it: serialise them. */
int as_fdtd_run_static(as_fdtd *h, const float *signal, size_t steps, double *out);

/* Zeroes the state (WaveSolver::reset, fdtd.hpp:906-915). */
int as_fdtd_reset(as_fdtd *h);

/* In moving air, the pressure squared over the grid's air (its cells not solid) as each step
 * of the last run carried it, averaged over those cells and the run's steps; 0 in still air,
 * before any run and after a reset (WaveSolver::mean_square, fdtd.hpp:711, 895-899). The sum
 * runs in double in an order of this library's own (a block's cells pairwise, then the blocks),
 * so it equals the reference's to rounding. Waits for the handle's stream. */
int as_fdtd_mean_square(as_fdtd *h, double *out);

/* The current pressure, the field the next step's probes read, widened to double
 * (nx ny nz values, x fastest). */
int as_fdtd_field(as_fdtd *h, double *out);

/* The handle's HIP stream (a hipStream_t), for the caller's events. */
void *as_fdtd_stream(as_fdtd *h);

/* Counts (out holds 11): [0] cells stepped as interior, [1] specials, [2] states, [3] layer
 * cells, [4] probes, [5] active tiles, [6] row pitch, [7] steps done, [8] planes a box of
 * two steps marches (0: one step a launch), [9] interior cells the second launch of two
 * steps takes one by one, [10] the general specials' kinds read from tables (0: each
 * special's own slabs; see as_tuning.special_kinds). */
int as_fdtd_info(const as_fdtd *h, uint64_t *out);

const char *as_fdtd_last_error(const as_fdtd *h);
void as_fdtd_destroy(as_fdtd *h);

/*
 * The gradient of a loss J of the probes' signals with respect to each special cell's wall
 * term g (coef_b = 1 - g, coef_inv = 1 / (1 + g), fdtd.hpp:3499-3517) and each branch
 * state's gain (fdtd.hpp:391-411), by the adjoint (time-reversed) scheme. The reference
 * has no adjoint (no counterpart); it serves fitting wall admittance to measured
 * responses. From rest, `steps` steps of `signal` are run, keeping every general special's
 * p(n + 1) - p(n - 1) where that history fits (keep_history) and the state every `every`
 * steps; then, segment by segment from the last, the segment's history is copied in (or
 * recomputed from its checkpoint) and the adjoint stepped back through it. dl_dout holds dJ/dy[probe][step]
 * (probes x steps). Out: the forward probes (probes x steps, or NULL), dJ/dg per layout
 * special (specials) and dJ/dgain per layout state (states). The handle must be made with
 * no_faces set; a speed of sound per cell and a matched layer are allowed (with a matched
 * layer the record keeps its history: keep_history not -1, and the history within its budget). It is left at
 * rest.
 */
int as_fdtd_gradient(as_fdtd *h, const float *signal, size_t steps, size_t every, const double *dl_dout,
                     double *out, double *sens_g, double *sens_gain);

/* as_fdtd_gradient in two halves, for a caller whose forward pass needs the signals
 * before it knows dJ/dy (an autograd function): as_fdtd_record runs the forward run and
 * keeps its checkpoints (and drive) in the handle; as_fdtd_adjoint then takes dJ/dy
 * (probes x the recorded steps), writes dJ/dg and dJ/dgain, frees the record and leaves
 * the handle at rest. No counterpart in the reference (see as_fdtd_gradient). */
int as_fdtd_record(as_fdtd *h, const float *signal, size_t steps, size_t every, double *out);
int as_fdtd_adjoint(as_fdtd *h, const double *dl_dout, double *sens_g, double *sens_gain);

/*
 * The receivers' first-order reading on the device (the reference's _first_order,
 * src/acoustic_spaces/engine/wave.py:757-802), on device pointers: d_probes holds the
 * stepper's output of `listeners` groups of 13 points (13 x listeners rows of `steps`
 * doubles, each group the listener, then one cell either side along x, y, z, then two),
 * d_found one byte per point (read and in sight), d_courant each listener's Courant number.
 * Writes d_out (listeners x 4 x steps doubles): W, then X, Y, Z as the Courant number times
 * the running trapezoidal integral of the cell difference along each axis (fourth-order,
 * second-order or one-sided by the points found; 0 where no point a cell away is), before
 * the trapezoid's gain correction (a spectral step left to the caller). Enqueued on
 * `stream` (a hipStream_t; NULL the default stream). The same values as the host reading,
 * bit for bit.
 */
int as_first_order_device(const double *d_probes, size_t steps, size_t listeners, const uint8_t *d_found,
                          const double *d_courant, double *d_out, void *stream);

/*
 * The set-up on walls on cell faces (the reference's WaveSolver constructor,
 * fdtd.hpp:360-690, bound at cpp/bindings/native_module.cpp:1238-1285): the arguments
 * of the binding, the mesh as vertices (nv x 3) and faces (nf x 3); threads 0 takes
 * OMP_NUM_THREADS where set, else the machine's. NULL on failure (ValueError in the reference; see as_setup_create_error).
 */
typedef struct as_setup as_setup;
as_setup *as_setup_create(const double *vertices, size_t nv, const int64_t *faces, size_t nf,
                          const int64_t *face_material, const double *beta, size_t materials,
                          const double *branches, const int64_t *branch_material, size_t nbranch,
                          const double *origin, double cell, size_t nx, size_t ny, size_t nz,
                          const double *source, const double *probes, size_t nprobes,
                          double edge_admittance, size_t layer, double layer_damping, size_t stencil,
                          const float *speed, size_t probes_per_listener, int threads);
/*
 * as_setup_create with the air's motion and flags. flow_x, flow_y (both or neither, [nx ny
 * nz]): each cell's shifts a step along x and y, the air moving under half a cell a step
 * (the binding's flow_x, flow_y; set_up_flow, fdtd.hpp:1680-1935, solving the wind round
 * the solids where the shifts do not hold as given). AS_SETUP_ON_DEVICE: the passes over
 * the whole grid, the cut links (fdtd.hpp:446-477), the air's flood (498-517), the
 * matched layer (521-542) and the boundary cells (544-625), run on the current GPU
 * (hipMalloc'd temporaries, freed before the call returns); the layout, and every refusal,
 * is the host set-up's bit for bit. Grids of 2^32 cells or more are refused in this mode.
 */
#define AS_SETUP_ON_DEVICE 1u
/*
 * AS_SETUP_FITTED: fitted walls (the reference's set_up_fitted, fdtd.hpp:970-1217, over its
 * CutCells fit, cut_cells.hpp): each cell a face meets is cut exactly, and every air cell
 * that owns cut air holds its own air volume, open areas towards its 18 neighbours and wall
 * areas; the source and probes read the fit's cells. Closed meshes off the grid's planes
 * only; host only (not with AS_SETUP_ON_DEVICE), in still air.
 */
#define AS_SETUP_FITTED 2u
as_setup *as_setup_create_ex(const double *vertices, size_t nv, const int64_t *faces, size_t nf,
                             const int64_t *face_material, const double *beta, size_t materials,
                             const double *branches, const int64_t *branch_material, size_t nbranch,
                             const double *origin, double cell, size_t nx, size_t ny, size_t nz,
                             const double *source, const double *probes, size_t nprobes,
                             double edge_admittance, size_t layer, double layer_damping, size_t stencil,
                             const float *speed, const float *flow_x, const float *flow_y,
                             size_t probes_per_listener, int threads, unsigned flags);
const char *as_setup_create_error(void);
/* Points `out` at the set-up's layout arrays (WaveSolver::layout, fdtd.hpp:727-782); they
 * live as long as the handle. */
int as_setup_layout(const as_setup *h, as_layout *out);
/* [0] air cells, [1] boundary cells, [2] branch states, [3] layer cells, [4] probes,
 * [5] open, [6] open sides (bit k for -x, +x, -y, +y, -z, +z), [7] source on a face. */
int as_setup_info(const as_setup *h, uint64_t *out);
/* The moving air's set-up (WaveSolver::wind_round, wind_fastest, wind_sweeps, fdtd.hpp:705-707;
 * the binding's properties, native_module.cpp:1596-1605): whether the wind was made to flow
 * round the model's solids; the fastest the air moves along an axis, in cells per step
 * (through a face between two cells where round, else the largest given component; 0 in
 * still air); the sweeps its solve took (0 without one). Any pointer may be NULL. */
int as_setup_wind(const as_setup *h, int *round, double *fastest, uint64_t *sweeps);
/* probe_found and probe_in_sight of probe p (fdtd.hpp:713-716). */
int as_setup_probe(const as_setup *h, size_t p, int *found, int *in_sight);
/*
 * Fitted walls only (AS_SETUP_FITTED; AS_EINVAL otherwise): the fit's rows, as the
 * reference's WaveSolver.cut_cells() hands them out (cpp/bindings/native_module.cpp:1614-1632):
 * their count, and pointers (living as long as the handle) to each row's cell (ascending),
 * air volume (cells^3), open area towards the 18 directions (rows x 18, cell faces), wall
 * area per material then the grid's edge (rows x (materials + 1)), sliver thickness per
 * axial face (rows x 6, cells); each special cell's volume after the clamp (in the layout's
 * order); each cell's kind in the fit (0 solid, 1 air, 2 matched layer). stats, where not
 * NULL, gets 18 counts: cubes, regions, ties, farthest, conduit links, conduits scaled, far
 * volume, diagonal area, dropped area, edge area, air volume, least conduit factor, conduit
 * wall moved, faces, planes, leaves, squares, quantum bits. Any pointer may be NULL.
 */
int as_setup_cut_cells(const as_setup *h, uint64_t *rows, const int64_t **cell, const double **air_volume,
                       const double **open_area, const double **wall_area, const double **sliver,
                       const double **special_volume, const uint8_t **kind, double *stats);
void as_setup_destroy(as_setup *h);

#ifdef __cplusplus
}
#endif
#endif /* ASFDTD_H */
