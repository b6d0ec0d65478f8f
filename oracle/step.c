/*
 * TEST INFRASTRUCTURE ONLY — the oracle. Tests, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, as the checker; the product never does.
 *
 * A CPU restatement, in plain C, of the reference wave solver's time step
 * (acoustic_spaces::WaveSolver::run, cpp/include/acoustic_spaces/fdtd.hpp:786-904),
 * driven only by the solver's layout (WaveSolver::layout, fdtd.hpp:727-782), as the
 * reference's own NumPy restatement in tests/test_wave.py:450-513 is. One thread; the
 * operations of each update in the reference's order, compiled -ffp-contract=off
 * (CMakeLists.txt:27-31) and run with denormals flushed to zero (FlushDenormals,
 * fdtd.hpp:314-336), so that in single precision it equals the reference bit for bit.
 *
 * Per step, with a the previous pressure (overwritten by the next) and b the current:
 *   record     out[p][step] = sum_k w_k (double) * b[c_k]                 (fdtd.hpp:3525-3534)
 *   interior   a = centre*b + axial*(6 neighbours, fixed order) - a        (fdtd.hpp:240-288)
 *   specials   CSR links, branch terms, source, coef_inv last              (fdtd.hpp:860-881)
 *   layer      split-field PML from b into a                               (fdtd.hpp:2242-2266)
 *   then a and b swap, and the branch velocities wa and wb with them       (fdtd.hpp:883-885)
 * In moving air both pressures, b then a, are carried along with the air between the record
 * and the interior (fdtd.hpp:838-846): three Runge-Kutta stages (fdtd.hpp:2026-2111) of the
 * plain rows' operator (fdtd.hpp:1945-1974) or the wind round the solids' (fdtd.hpp:1976-2021).
 *
 * With precision 64 the same scheme runs in double: the fields, states and arithmetic
 * are double, the coefficients the layout's floats widened (an FP64 run of the same
 * discrete scheme, the reference of the float run's rounding error).
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#if defined(__SSE__) || defined(__x86_64__)
#include <xmmintrin.h>
#define ORACLE_HAS_MXCSR 1
#endif

/* The layout's arrays (names as WaveSolver::layout has them). */
typedef struct {
    uint64_t nx, ny, nz, stencil;
    const float *weights; /* [stencil == 7 ? 6 : 18] */
    float centre, total;
    const float *scale;       /* [n] or NULL */
    const float *layer_scale; /* [layer_cells] or NULL */
    const uint8_t *types;     /* [n] */
    uint64_t specials;
    const int64_t *special;
    const uint64_t *link_start;
    const uint8_t *link_direction;
    const float *link_weight;
    const float *coef_a, *coef_b, *coef_inv, *source;
    const uint64_t *state_start;
    uint64_t states;
    const int32_t *state_branch;
    const float *state_weight;
    const float *branch_now, *branch_before, *branch_gain;
    uint64_t layer_cells;
    const int64_t *layer_cell;
    const float *keep[6]; /* keep_x, keep_face_x, keep_y, keep_face_y, keep_z, keep_face_z */
    const float *push[6];
    uint64_t probes;
    const uint64_t *probe_start, *probe_cells;
    const double *probe_weights;
    /* Moving air, or NULL: the cells' masks and either the plain shifts or the faces' winds
       and loops along each axis (x, y, z), each [n]. */
    const uint8_t *flow_mask;
    const float *flow_x, *flow_y;
    const float *wind[3], *loop[3];
} oracle_layout;

typedef struct {
    oracle_layout L;
    int precision;
    void *u0, *u1, *w0, *w1, *layer_state;
    void *t1, *t2; /* the carrying's two stages, in moving air */
    ptrdiff_t offset[18];
} oracle_solver;

static const int kSteps[18][3] = {
    {-1, 0, 0},  {1, 0, 0},  {0, -1, 0}, {0, 1, 0},  {0, 0, -1}, {0, 0, 1},
    {-1, -1, 0}, {1, -1, 0}, {-1, 1, 0}, {1, 1, 0},  {-1, 0, -1}, {1, 0, -1},
    {-1, 0, 1},  {1, 0, 1},  {0, -1, -1}, {0, 1, -1}, {0, -1, 1}, {0, 1, 1}};

oracle_solver *oracle_create(const oracle_layout *layout, int precision) {
    oracle_solver *s = (oracle_solver *)calloc(1, sizeof(oracle_solver));
    if (!s) return NULL;
    s->L = *layout;
    s->precision = precision == 64 ? 64 : 32;
    const size_t e = s->precision == 64 ? sizeof(double) : sizeof(float);
    const size_t n = (size_t)(layout->nx * layout->ny * layout->nz);
    s->u0 = calloc(n, e);
    s->u1 = calloc(n, e);
    s->w0 = calloc(layout->states + 1, e);
    s->w1 = calloc(layout->states + 1, e);
    s->layer_state = calloc(9 * layout->layer_cells + 1, e);
    s->t1 = calloc(layout->flow_mask ? n : 1, e);
    s->t2 = calloc(layout->flow_mask ? n : 1, e);
    const ptrdiff_t sy = (ptrdiff_t)layout->nx, sz = (ptrdiff_t)(layout->nx * layout->ny);
    for (int k = 0; k < 18; ++k) s->offset[k] = kSteps[k][0] + kSteps[k][1] * sy + kSteps[k][2] * sz;
    return s;
}

void oracle_destroy(oracle_solver *s) {
    if (!s) return;
    free(s->u0);
    free(s->u1);
    free(s->w0);
    free(s->w1);
    free(s->layer_state);
    free(s->t1);
    free(s->t2);
    free(s);
}

void oracle_reset(oracle_solver *s) {
    const size_t e = s->precision == 64 ? sizeof(double) : sizeof(float);
    const size_t n = (size_t)(s->L.nx * s->L.ny * s->L.nz);
    memset(s->u0, 0, n * e);
    memset(s->u1, 0, n * e);
    memset(s->w0, 0, (s->L.states + 1) * e);
    memset(s->w1, 0, (s->L.states + 1) * e);
    memset(s->layer_state, 0, (9 * s->L.layer_cells + 1) * e);
    memset(s->t1, 0, (s->L.flow_mask ? n : 1) * e);
    memset(s->t2, 0, (s->L.flow_mask ? n : 1) * e);
}

/* Moving air, one body for both precisions. ROW: a quantity's change along a row (stride st)
 * at cell i over a cell, as air moving by `shift` carries it; m's bits say which of the row's
 * cells one (bits 0, 1: below, above) and two (bits 2, 3) away move with it (fdtd.hpp:1945-1966).
 * PLAIN: -(shift_x ROW_x + shift_y ROW_y) (fdtd.hpp:1969-1974). ROUND: over each axis the two
 * faces' winds times the differences across them, halved, and the three loops through the cell
 * (fdtd.hpp:1976-1991); BORDER: the same where a face leads to air that does not move, whose
 * cell's own value leaves through it against the wind (fdtd.hpp:1993-2021). CARRY: q's three
 * stages, q + d, 3/4 q + 1/4 (t1 + d), q + 2/3 ((t2 + d) - q), solid cells skipped and still ones
 * kept (fdtd.hpp:2026-2111). */
#define ORACLE_FLOW(NAME, REAL)                                                                 \
    static REAL NAME##_row(const REAL *q, size_t i, size_t st, unsigned m, REAL shift) {        \
        if ((m & 15u) == 15u)                                                                    \
            return ((REAL)8 * (q[i + st] - q[i - st]) - (q[i + 2 * st] - q[i - 2 * st])) *       \
                   ((REAL)1 / (REAL)12);                                                         \
        const REAL c = q[i];                                                                     \
        REAL faces = 0, loops = 0;                                                               \
        if (m & 2u) {                                                                            \
            faces += (REAL)0.5 * (q[i + st] - c);                                                \
            if (m & 8u) loops += q[i + st] - q[i + 2 * st];                                      \
        } else if (shift < 0) {                                                                  \
            faces -= c;                                                                          \
        }                                                                                        \
        if (m & 1u) {                                                                            \
            faces += (REAL)0.5 * (c - q[i - st]);                                                \
            if (m & 4u) loops += q[i - 2 * st] - q[i - st];                                      \
        } else if (shift > 0) {                                                                  \
            faces += c;                                                                          \
        }                                                                                        \
        if ((m & 3u) == 3u) loops += q[i + st] - q[i - st];                                      \
        return faces + loops * ((REAL)1 / (REAL)12);                                             \
    }                                                                                            \
    static REAL NAME##_plain(const oracle_layout *L, const REAL *q, size_t i) {                  \
        const REAL sx = (REAL)L->flow_x[i], sy = (REAL)L->flow_y[i];                             \
        const unsigned m = L->flow_mask[i];                                                      \
        return -(sx * NAME##_row(q, i, 1, m & 15u, sx) + sy * NAME##_row(q, i, L->nx, m >> 4, sy)); \
    }                                                                                            \
    static REAL NAME##_round(const oracle_layout *L, const REAL *q, size_t i, int border) {      \
        const size_t stride[3] = {1, L->nx, L->nx * L->ny};                                      \
        const REAL c = q[i];                                                                     \
        REAL sum = 0;                                                                            \
        for (int d = 0; d < 3; ++d) {                                                            \
            const size_t st = stride[d];                                                         \
            const float *w = L->wind[d], *l = L->loop[d];                                        \
            const REAL up = q[i + st], down = q[i - st];                                         \
            const REAL l0 = (REAL)l[i] * (up - q[i + 2 * st]), l1 = (REAL)l[i - st] * (up - down);  \
            const REAL l2 = (REAL)l[i - 2 * st] * (q[i - 2 * st] - down);                        \
            if (!border) {                                                                       \
                sum += (REAL)0.5 * ((REAL)w[i] * (up - c) + (REAL)w[i - st] * (c - down)) + l0 + l1 + l2; \
                continue;                                                                        \
            }                                                                                    \
            if (L->flow_mask[i + st]) {                                                          \
                sum += (REAL)0.5 * (REAL)w[i] * (up - c);                                        \
            } else if (w[i] < 0.0f) {                                                            \
                sum -= (REAL)w[i] * c;                                                           \
            }                                                                                    \
            if (L->flow_mask[i - st]) {                                                          \
                sum += (REAL)0.5 * (REAL)w[i - st] * (c - down);                                 \
            } else if (w[i - st] > 0.0f) {                                                       \
                sum += (REAL)w[i - st] * c;                                                      \
            }                                                                                    \
            sum += l0 + l1 + l2;                                                                 \
        }                                                                                        \
        return -sum;                                                                             \
    }                                                                                            \
    static void NAME##_carry(const oracle_layout *L, REAL *q, REAL *t1, REAL *t2) {              \
        const size_t n = (size_t)(L->nx * L->ny * L->nz);                                        \
        const int round = L->wind[0] != NULL;                                                    \
        for (int stage = 0; stage < 3; ++stage) {                                                \
            const REAL *from = stage == 0 ? q : (stage == 1 ? t1 : t2);                          \
            for (size_t i = 0; i < n; ++i) {                                                     \
                if (L->types[i] == 0) continue;                                                  \
                const int still = round ? L->flow_mask[i] == 0                                   \
                                        : (L->flow_x[i] == 0.0f && L->flow_y[i] == 0.0f);        \
                if (still) {                                                                     \
                    if (stage == 0) t1[i] = q[i];                                                \
                    if (stage == 1) t2[i] = q[i];                                                \
                    continue;                                                                    \
                }                                                                                \
                const REAL d = round ? NAME##_round(L, from, i, L->flow_mask[i] == 2)            \
                                     : NAME##_plain(L, from, i);                                 \
                if (stage == 0) {                                                                \
                    t1[i] = q[i] + d;                                                            \
                } else if (stage == 1) {                                                         \
                    t2[i] = (REAL)0.75 * q[i] + (REAL)0.25 * (t1[i] + d);                        \
                } else {                                                                         \
                    q[i] = q[i] + ((REAL)2 / (REAL)3) * ((t2[i] + d) - q[i]);                    \
                }                                                                                \
            }                                                                                    \
        }                                                                                        \
    }

ORACLE_FLOW(flow_f32, float)
ORACLE_FLOW(flow_f64, double)

/* One body for both precisions: REAL is float or double. */
#define ORACLE_RUN(NAME, FLOW, REAL)                                                            \
    static void NAME(oracle_solver *s, const float *signal, size_t steps, double *out) {        \
        const oracle_layout *L = &s->L;                                                          \
        const size_t nx = L->nx, ny = L->ny, nz = L->nz, n = nx * ny * nz;                       \
        const size_t sy = nx, sz = nx * ny;                                                      \
        REAL *a = (REAL *)s->u0, *b = (REAL *)s->u1;                                             \
        REAL *wa = (REAL *)s->w0, *wb = (REAL *)s->w1;                                           \
        REAL *state = (REAL *)s->layer_state;                                                    \
        const int wide = L->stencil == 19;                                                       \
        const REAL axial = (REAL)L->weights[0];                                                  \
        const REAL diagonal = wide ? (REAL)L->weights[6] : (REAL)0;                              \
        const REAL centre = (REAL)L->centre, total = (REAL)L->total;                             \
        REAL *term = (REAL *)calloc(L->states + 1, sizeof(REAL));                                \
        for (size_t step = 0; step < steps; ++step) {                                            \
            /* record: the pressure before the step, summed in double */                         \
            for (size_t p = 0; p < L->probes; ++p) {                                             \
                double value = 0.0;                                                              \
                for (uint64_t k = L->probe_start[p]; k < L->probe_start[p + 1]; ++k)             \
                    value += L->probe_weights[k] * (double)b[L->probe_cells[k]];                 \
                out[p * steps + step] = value;                                                   \
            }                                                                                    \
            /* moving air: both pressures carried along with it */                               \
            if (L->flow_mask) {                                                                  \
                FLOW##_carry(L, b, (REAL *)s->t1, (REAL *)s->t2);                                \
                FLOW##_carry(L, a, (REAL *)s->t1, (REAL *)s->t2);                                \
            }                                                                                    \
            /* interior */                                                                       \
            for (size_t i = 0; i < n; ++i) {                                                     \
                if (L->types[i] != 1) continue;                                                  \
                REAL sum = b[i - 1] + b[i + 1];                                                  \
                sum = sum + b[i - sy];                                                           \
                sum = sum + b[i + sy];                                                           \
                sum = sum + b[i - sz];                                                           \
                sum = sum + b[i + sz];                                                           \
                REAL edges = 0;                                                                  \
                if (wide) {                                                                      \
                    edges = b[i - 1 - sy] + b[i + 1 - sy];                                       \
                    edges = edges + b[i - 1 + sy];                                               \
                    edges = edges + b[i + 1 + sy];                                               \
                    edges = edges + b[i - 1 - sz];                                               \
                    edges = edges + b[i + 1 - sz];                                               \
                    edges = edges + b[i - 1 + sz];                                               \
                    edges = edges + b[i + 1 + sz];                                               \
                    edges = edges + b[i - sy - sz];                                              \
                    edges = edges + b[i + sy - sz];                                              \
                    edges = edges + b[i - sy + sz];                                              \
                    edges = edges + b[i + sy + sz];                                              \
                }                                                                                \
                REAL next;                                                                       \
                if (L->scale) {                                                                  \
                    const REAL sc = (REAL)L->scale[i];                                           \
                    const REAL c0 = (REAL)2 - total * sc;                                        \
                    const REAL ax = sc * axial;                                                  \
                    next = c0 * b[i] + ax * sum;                                                 \
                    if (wide) {                                                                  \
                        const REAL dg = sc * diagonal;                                           \
                        next = next + dg * edges;                                                \
                    }                                                                            \
                } else {                                                                         \
                    next = centre * b[i] + axial * sum;                                          \
                    if (wide) next = next + diagonal * edges;                                    \
                }                                                                                \
                a[i] = next - a[i];                                                              \
            }                                                                                    \
            /* specials */                                                                       \
            const REAL sig = (REAL)signal[step];                                                 \
            for (size_t k = 0; k < L->specials; ++k) {                                           \
                const size_t i = (size_t)L->special[k];                                          \
                REAL sum = 0;                                                                    \
                for (uint64_t l = L->link_start[k]; l < L->link_start[k + 1]; ++l) {             \
                    const ptrdiff_t j = (ptrdiff_t)i + s->offset[L->link_direction[l]];          \
                    sum += (REAL)L->link_weight[l] * b[j];                                       \
                }                                                                                \
                REAL flow = 0;                                                                   \
                for (uint64_t j = L->state_start[k]; j < L->state_start[k + 1]; ++j) {           \
                    const int32_t br = L->state_branch[j];                                       \
                    const REAL own = (REAL)L->branch_now[br] * wb[j] -                           \
                                     (REAL)L->branch_before[br] * wa[j];                         \
                    term[j] = own;                                                               \
                    flow += (REAL)L->state_weight[j] * own;                                      \
                }                                                                                \
                const REAL previous = a[i];                                                      \
                REAL acc = (REAL)L->coef_a[k] * b[i];                                            \
                acc = acc + sum;                                                                 \
                acc = acc - (REAL)L->coef_b[k] * previous;                                       \
                acc = acc - flow;                                                                \
                acc = acc + (REAL)L->source[k] * sig;                                            \
                const REAL next = acc * (REAL)L->coef_inv[k];                                    \
                a[i] = next;                                                                     \
                for (uint64_t j = L->state_start[k]; j < L->state_start[k + 1]; ++j) {           \
                    const int32_t br = L->state_branch[j];                                       \
                    wa[j] += (next - previous) * (REAL)L->branch_gain[br] + term[j];             \
                }                                                                                \
            }                                                                                    \
            /* matched layer */                                                                  \
            for (size_t c = 0; c < L->layer_cells; ++c) {                                       \
                const size_t i = (size_t)L->layer_cell[c];                                       \
                const size_t at[3] = {i % nx, (i / nx) % ny, i / sz};                            \
                REAL *st = state + 9 * c;                                                        \
                REAL tot = 0;                                                                    \
                for (int axis = 0; axis < 3; ++axis) {                                           \
                    const size_t stride = axis == 0 ? 1 : (axis == 1 ? sy : sz);                 \
                    const size_t x = at[axis];                                                   \
                    const float *fk = L->keep[2 * axis + 1], *fp = L->push[2 * axis + 1];        \
                    REAL below = st[3 + 2 * axis], above = st[4 + 2 * axis];                     \
                    below = (REAL)fk[x] * below - (REAL)fp[x] * (b[i] - b[i - stride]);          \
                    above = (REAL)fk[x + 1] * above - (REAL)fp[x + 1] * (b[i + stride] - b[i]);  \
                    st[3 + 2 * axis] = below;                                                    \
                    st[4 + 2 * axis] = above;                                                    \
                    REAL part = st[axis];                                                        \
                    if (L->layer_scale) {                                                        \
                        const REAL ps = (REAL)L->push[2 * axis][x] * (REAL)L->layer_scale[c];    \
                        part = (REAL)L->keep[2 * axis][x] * part - ps * (above - below);         \
                    } else {                                                                     \
                        part = (REAL)L->keep[2 * axis][x] * part -                               \
                               (REAL)L->push[2 * axis][x] * (above - below);                     \
                    }                                                                            \
                    st[axis] = part;                                                             \
                    tot += part;                                                                 \
                }                                                                                \
                a[i] = tot;                                                                      \
            }                                                                                    \
            REAL *t = a; a = b; b = t;                                                           \
            t = wa; wa = wb; wb = t;                                                             \
        }                                                                                        \
        (void)nz;                                                                                \
        free(term);                                                                              \
    }

ORACLE_RUN(run_f32, flow_f32, float)
ORACLE_RUN(run_f64, flow_f64, double)

void oracle_run(oracle_solver *s, const float *signal, size_t steps, double *out) {
#ifdef ORACLE_HAS_MXCSR
    const unsigned saved = _mm_getcsr();
    _mm_setcsr(saved | 0x8040u); /* flush to zero, denormals are zero */
#endif
    if (s->precision == 64) {
        run_f64(s, signal, steps, out);
    } else {
        run_f32(s, signal, steps, out);
    }
#ifdef ORACLE_HAS_MXCSR
    _mm_setcsr(saved);
#endif
    if (steps % 2 == 1) { /* the state continues across calls (fdtd.hpp:900-903) */
        void *t = s->u0; s->u0 = s->u1; s->u1 = t;
        t = s->w0; s->w0 = s->w1; s->w1 = t;
    }
}

/* The current pressure (the field the next step's probes read), widened to double. */
void oracle_field(const oracle_solver *s, double *out) {
    const size_t n = (size_t)(s->L.nx * s->L.ny * s->L.nz);
    if (s->precision == 64) {
        memcpy(out, s->u1, n * sizeof(double));
    } else {
        const float *f = (const float *)s->u1;
        for (size_t i = 0; i < n; ++i) out[i] = (double)f[i];
    }
}
