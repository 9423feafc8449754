/* oracle.c - CPU restatement of the reference's DiffMC / DiffDMC engine (the oracle).
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg, as the checker and as the CPU baseline. Nothing in the product (diso_amd/) links, loads or
 * calls it; the product fails loudly when its HIP library is missing instead of falling back here.
 *
 * It follows the reference's CUDA engine kernel by kernel on the grid exactly as the reference's
 * C++ layer receives it (already padded by the Python wrapper, diso/__init__.py:52-54):
 *   MC : CuMC::forward / backward          src/cumc.cu:265-743
 *   DMC: CUDualMC::forward / backward      src/cudualmc.cu:605-1138
 * Each per-cell kernel is a (optionally OpenMP-parallel) loop; each CUB ExclusiveSum is a plain
 * exclusive prefix sum. Face indices are kept as int64 here (the reference emits int32 and the
 * Python wrapper widens them with .long()).
 *
 * Build: see oracle/Makefile (gcc -O2 -fopenmp -ffp-contract=off; IEEE single/double via SSE2).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../diso_amd/csrc/diso_tables.h"

enum { ORACLE_MC = 0, ORACLE_DMC = 1 };
enum { ORACLE_F32 = 0, ORACLE_F64 = 1 };
enum { ORACLE_STATUS_UNPADDED = 2 };

/* MC corner order (Bourke), src/cumc.cu:107. */
static const int MC_CORNER[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                    {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};

/* Extractor state, the counterpart of CuMC / CUDualMC (src/cumc.h:70-139,
 * src/cudualmc.h:70-150): backward consumes what the last forward left here. */
typedef struct {
    int algo, dtype, threads, status;
    int64_t dims[3], n_cells;
    int64_t n_used, n_verts_mc, n_verts, n_faces;
    int64_t *first_cell_used; /* [n_cells+1] exclusive scan of used flags          */
    int64_t *used_cell_index; /* [n_used]    used cell -> linear cell id             */
    int64_t *vbase;           /* [n_used+1]  MC vertex (= DMC quad) base per used cell */
    int64_t *fbase;           /* [n_used+1]  MC: 3*triangle base; DMC: patch base      */
    uint8_t *code;            /* [n_used]    MC Bourke code / DMC resolved Morton code */
    int8_t *verts_type;       /* MC: axis of each vertex                              */
    int64_t *mc_vert_to_cell; /* DMC: owner used cell of each crossing edge          */
    uint8_t *mc_vert_type;    /* DMC: axis + 3*exiting                                */
    void *verts;              /* [n_verts,3] scalar                                   */
    int64_t *faces;           /* MC: [n_faces,3]; DMC: [n_faces,4]                    */
} oracle_state;

/* In-place exclusive prefix sum over a[0..n); returns the total (the value left in a[n-1] when
 * a[n-1] was 0 on entry, which is how the reference's ExclusiveSum over n+1 elements reports its
 * count). Two-level when threads > 1. */
static int64_t oracle_exclusive_scan(int64_t *a, int64_t n, int threads) {
    if (n <= 0) return 0;
    if (threads <= 1 || n < 65536) {
        int64_t run = 0;
        for (int64_t i = 0; i < n; ++i) {
            int64_t v = a[i];
            a[i] = run;
            run += v;
        }
        return a[n - 1];
    }
    int64_t *part = (int64_t *)calloc((size_t)threads + 1, sizeof(int64_t));
#pragma omp parallel num_threads(threads)
    {
        int t = 0, nt = 1;
#ifdef _OPENMP
        t = omp_get_thread_num();
        nt = omp_get_num_threads();
#endif
        int64_t chunk = (n + nt - 1) / nt;
        int64_t lo = (int64_t)t * chunk < n ? (int64_t)t * chunk : n;
        int64_t hi = lo + chunk < n ? lo + chunk : n;
        int64_t s = 0;
        for (int64_t i = lo; i < hi; ++i) s += a[i];
        part[t + 1] = s;
#pragma omp barrier
#pragma omp single
        for (int k = 1; k <= nt; ++k) part[k] += part[k - 1];
        int64_t run = part[t];
        for (int64_t i = lo; i < hi; ++i) {
            int64_t v = a[i];
            a[i] = run;
            run += v;
        }
    }
    free(part);
    return a[n - 1];
}

#define S float
#define FN(name) name##_f32
#include "oracle_impl.h"
#undef S
#undef FN
#define S double
#define FN(name) name##_f64
#include "oracle_impl.h"
#undef S
#undef FN

static void state_reset(oracle_state *st) {
    free(st->first_cell_used);
    free(st->used_cell_index);
    free(st->vbase);
    free(st->fbase);
    free(st->code);
    free(st->verts_type);
    free(st->mc_vert_to_cell);
    free(st->mc_vert_type);
    free(st->verts);
    free(st->faces);
    int algo = st->algo, dtype = st->dtype, threads = st->threads;
    memset(st, 0, sizeof(*st));
    st->algo = algo;
    st->dtype = dtype;
    st->threads = threads;
}

/* ---------------------------------------------------------------- C ABI (ctypes) */

void *oracle_create(int algo, int dtype, int threads) {
    if ((algo != ORACLE_MC && algo != ORACLE_DMC) || (dtype != ORACLE_F32 && dtype != ORACLE_F64))
        return NULL;
    oracle_state *st = (oracle_state *)calloc(1, sizeof(oracle_state));
    if (!st) return NULL;
    st->algo = algo;
    st->dtype = dtype;
    st->threads = threads < 1 ? 1 : threads;
    return st;
}

void oracle_destroy(void *h) {
    if (!h) return;
    state_reset((oracle_state *)h);
    free(h);
}

int oracle_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* Forward on a grid of dims (X,Y,Z) exactly as the reference's C++ layer receives it (padded
 * by the caller). iso is converted to the scalar type as pybind converts the Python float.
 * Returns 0 on success, ORACLE_STATUS_UNPADDED when some face referenced a cell outside the used
 * set (the reference reads out of bounds there; it cannot happen on an iso+1-padded grid),
 * negative on allocation failure. */
int oracle_forward(void *h, const void *grid, const void *deform, int64_t X, int64_t Y,
                   int64_t Z, double iso) {
    oracle_state *st = (oracle_state *)h;
    state_reset(st);
    st->dims[0] = X;
    st->dims[1] = Y;
    st->dims[2] = Z;
    st->n_cells = X * Y * Z;
    int rc;
    if (st->dtype == ORACLE_F32) {
        rc = st->algo == ORACLE_MC
                 ? mc_forward_f32(st, (const float *)grid, (const float *)deform, (float)iso)
                 : dmc_forward_f32(st, (const float *)grid, (const float *)deform, (float)iso);
    } else {
        rc = st->algo == ORACLE_MC
                 ? mc_forward_f64(st, (const double *)grid, (const double *)deform, iso)
                 : dmc_forward_f64(st, (const double *)grid, (const double *)deform, iso);
    }
    return rc ? rc : st->status;
}

int64_t oracle_n_verts(void *h) { return ((oracle_state *)h)->n_verts; }
int64_t oracle_n_faces(void *h) { return ((oracle_state *)h)->n_faces; }
int64_t oracle_n_used(void *h) { return ((oracle_state *)h)->n_used; }

/* Copies the last forward's vertices ([V,3] scalar) and faces ([F,3|4] int64). */
void oracle_copy_out(void *h, void *verts, int64_t *faces) {
    oracle_state *st = (oracle_state *)h;
    size_t s = st->dtype == ORACLE_F32 ? sizeof(float) : sizeof(double);
    if (verts && st->n_verts) memcpy(verts, st->verts, s * 3 * (size_t)st->n_verts);
    int k = st->algo == ORACLE_MC ? 3 : 4;
    if (faces && st->n_faces) memcpy(faces, st->faces, sizeof(int64_t) * k * (size_t)st->n_faces);
}

/* Copies the used-cell codes in used-cell order (MC: Bourke code; DMC: resolved Morton code). */
void oracle_copy_codes(void *h, uint8_t *codes, int64_t *used_cell_index) {
    oracle_state *st = (oracle_state *)h;
    if (codes && st->n_used) memcpy(codes, st->code, (size_t)st->n_used);
    if (used_cell_index && st->n_used)
        memcpy(used_cell_index, st->used_cell_index, sizeof(int64_t) * (size_t)st->n_used);
}

/* Backward of the last forward (the same grid/deform/iso must be passed again, as in the
 * reference). Accumulates into adj_grid [X*Y*Z] and adj_deform [X*Y*Z*3] (may be NULL when
 * deform is NULL). reference_compat applies to DMC only. */
int oracle_backward(void *h, const void *grid, const void *deform, double iso,
                    const void *adj_verts, void *adj_grid, void *adj_deform,
                    int reference_compat) {
    oracle_state *st = (oracle_state *)h;
    if (st->dtype == ORACLE_F32) {
        if (st->algo == ORACLE_MC)
            mc_backward_f32(st, (const float *)grid, (const float *)deform, (float)iso,
                            (const float *)adj_verts, (float *)adj_grid, (float *)adj_deform);
        else
            dmc_backward_f32(st, (const float *)grid, (const float *)deform, (float)iso,
                             (const float *)adj_verts, (float *)adj_grid, (float *)adj_deform,
                             reference_compat);
    } else {
        if (st->algo == ORACLE_MC)
            mc_backward_f64(st, (const double *)grid, (const double *)deform, iso,
                            (const double *)adj_verts, (double *)adj_grid, (double *)adj_deform);
        else
            dmc_backward_f64(st, (const double *)grid, (const double *)deform, iso,
                             (const double *)adj_verts, (double *)adj_grid,
                             (double *)adj_deform, reference_compat);
    }
    return 0;
}

/* Expands the packed case tables into the reference's own layouts, so tests can compare them
 * with the reference text and check their invariants:
 *   fmi[257], ids[2460]                         firstMarchingCubesId / marchingCubesIds
 *   fpi[257], fei[359], ei[1536], deo[256*12]   mcFirstPatchIndex / mcFirstEdgeIndex /
 *                                               mcEdgeIndex / dmcEdgeOffset
 *   pc[256]                                     problematicConfigs (255 = not ambiguous)   */
void oracle_expand_tables(int32_t *fmi, int32_t *ids, int32_t *fpi, int32_t *fei, int32_t *ei,
                          int32_t *deo, int32_t *pc) {
    int n = 0;
    for (int c = 0; c < 256; ++c) {
        fmi[c] = n;
        uint64_t w = DISO_MC_CASE[c];
        for (int k = 0; k < 3 * DISO_MC_NTRI(w); ++k) ids[n++] = DISO_MC_EDGE(w, k);
    }
    fmi[256] = n;
    int np = 0, ne = 0;
    for (int c = 0; c < 256; ++c) {
        uint64_t w = DISO_DMC_CASE[c];
        fpi[c] = np;
        for (int k = 0; k < DISO_DMC_NPATCH(w); ++k) {
            fei[np++] = ne;
            for (int e = 0; e < 12; ++e)
                if (DISO_DMC_PATCH_OF(w, e) == k) ei[ne++] = e;
        }
        for (int e = 0; e < 12; ++e) {
            int p = DISO_DMC_PATCH_OF(w, e);
            deo[12 * c + e] = p == 7 ? -1 : p;
        }
        pc[c] = DISO_DMC_DIR(w) == 7 ? 255 : DISO_DMC_DIR(w);
    }
    fpi[256] = np;
    fei[np] = ne;
}
