/* diso_hip.h - C ABI of the MI355X (gfx950) DiffMC / DiffDMC engine (libdiso_hip.so).
 *
 * This ABI replaces the reference's pybind11 classes _C.CUMCFloat / CUMCDouble / CUDMCFloat /
 * CUDMCDouble (src/pybind.cpp:16-395,419-447) and the engine structs behind them (CuMC,
 * src/cumc.h:70-139; CUDualMC, src/cudualmc.h:70-150). It takes plain device pointers, sizes and
 * a HIP stream; it never allocates device memory and never synchronises the device. The caller
 * owns every buffer (two workspaces described below, inputs and outputs).
 *
 * Conventions shared by all entry points
 *  - algo: DISO_MC (Marching Cubes) or DISO_DMC (Dual Marching Cubes).
 *  - dtype: DISO_F32 or DISO_F64; sdf, deform, vertices and gradients all use it.
 *  - sdf is a contiguous [X,Y,Z] grid (z fastest); deform, when not NULL, is contiguous
 *    [X,Y,Z,3]. With DISO_FLAG_VIRTUAL_PAD the grid is the user's unpadded grid and the engine
 *    behaves exactly as if it had been padded by one sample of value (iso + 1) (deform by 0), as
 *    the reference's Python wrapper does (diso/__init__.py:52-54) before calling its engine;
 *    without it the grid is taken as given, like the reference's _C classes.
 *  - iso is the Python float; it is rounded to dtype for comparisons and the pad value is
 *    (dtype)(iso + 1.0), matching pybind's conversion and F.pad's fill.
 *  - Every entry point returns 0 on success or a negative DISO_E* code; diso_last_error() then
 *    holds a message (thread-local). The Python layer raises RuntimeError, the reference's
 *    exception type for its input checks (src/pybind.cpp:5-11,57-60).
 *  - stream is a hipStream_t (NULL = the default stream). All work is stream-ordered.
 */
#ifndef DISO_HIP_H
#define DISO_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { DISO_MC = 0, DISO_DMC = 1 };
enum { DISO_F32 = 0, DISO_F64 = 1 };

enum {
    DISO_FLAG_VIRTUAL_PAD = 1 << 0, /* input is unpadded; pad with iso+1 / 0 on the fly          */
    DISO_FLAG_SHIFT = 1 << 1,       /* emit vertices minus 1 (diso/__init__.py:56,110)           */
    DISO_FLAG_NORMALIZE = 1 << 2,   /* then divide by (dims-1) per axis (diso/__init__.py:57-60) */
    DISO_FLAG_INDEX64 = 1 << 3,     /* faces as int64 (the wrapper's .long()); else int32        */
    DISO_FLAG_REFERENCE_COMPAT = 1 << 4, /* DMC backward reads each cell's first patch gradient
                                            for all its patches, as src/cudualmc.cu:990 does    */
    DISO_FLAG_ACCUMULATE = 1 << 5,  /* backward adds into the gradient buffers instead of
                                       overwriting them                                        */
    DISO_FLAG_RECOMPUTE = 1 << 6    /* backward re-derives the crossing masks from the samples
                                       instead of reading the classification words, so only
                                       the bases prefix of seg (diso_seg_bytes with this flag)
                                       must be kept from the forward                           */
};

enum {
    DISO_OK = 0,
    DISO_E_INVALID = -1,  /* bad argument                                                   */
    DISO_E_HIP = -2,      /* HIP runtime error                                              */
    DISO_E_CAPACITY = -3, /* a workspace or output buffer is too small                      */
    DISO_E_TIMEOUT = -4   /* the single-pass scan did not complete (never expected); the
                             control block stays usable: no reset is needed                */
};

/* Status bits of diso_totals_t.flags. */
enum {
    DISO_TOTAL_HAS_BELOW = 1 << 0, /* some sample < iso  (i.e. not min >= iso)               */
    DISO_TOTAL_HAS_ABOVE = 1 << 1, /* some sample > iso  (i.e. not max <= iso)               */
    DISO_TOTAL_HAS_NAN = 1 << 2,   /* some sample is NaN                                      */
    DISO_TOTAL_SCAN_TIMEOUT = 1 << 3
};

/* Result of the counting pass, read back to size the outputs.
 *  MC : n_verts = vertices (= crossing edges), n_faces = triangles.
 *  DMC: n_verts = vertices (= patches),        n_faces = quads (= crossing edges).
 * The wrapper's empty-grid early-out (diso/__init__.py:49,103) applies when
 * (flags & (HAS_BELOW|HAS_ABOVE)) != (HAS_BELOW|HAS_ABOVE). */
typedef struct {
    int64_t n_verts;
    int64_t n_faces;
    int64_t flags;
    int64_t n_crossing;
} diso_totals_t;

/* Size in bytes of the per-call segment workspace ("seg") for a grid of dims (X,Y,Z): the
 * per-segment vertex/face bases and a 16-bit classification word per grid position, written by
 * diso_count and read by diso_emit / diso_backward. It needs no initialisation. Keep it (with the
 * same inputs) to run diso_backward later; that is the state the reference keeps inside CuMC
 * between forward and backward (src/cumc.h:73-99). With DISO_FLAG_RECOMPUTE in `flags` the size
 * is that of the bases prefix alone, which is all a DISO_FLAG_RECOMPUTE backward reads. */
int64_t diso_seg_bytes(int64_t X, int64_t Y, int64_t Z, int flags);

/* Byte offset in `seg` of the totals a diso_count leaves on the device: four int64, (crossing
 * edges, MC triangles | DMC vertices, DISO_TOTAL_* flags, algo). The reference keeps these counts
 * in host members of CuMC / CUDualMC (src/cumc.h:73-99); here they can stay on the device. */
int64_t diso_totals_offset(int64_t X, int64_t Y, int64_t Z, int flags);

/* Size in bytes of the control workspace ("ctl") that serves a grid of dims (X,Y,Z) or any
 * smaller one. It must be zero-filled once when allocated and is then reused across calls; calls
 * that share a ctl must not overlap in time (it holds the single-pass scan's tile states). */
int64_t diso_ctl_bytes(int64_t X, int64_t Y, int64_t Z, int flags);

/* Pass 1: classify every cell, count vertices and faces per segment and scan the counts in one
 * single-pass launch. Replaces count_used_cells_kernel, index_used_cells_kernel,
 * count_cell_mc_verts_kernel, count_cell_mc_tris_kernel / count_cell_patches_kernel and the three
 * CUB scans of CuMC::forward (src/cumc.cu:651-723) and CUDualMC::forward
 * (src/cudualmc.cu:1058-1121). If totals_host is not NULL the totals are copied back and the
 * stream is synchronised (the one host sync the plain API needs to size its outputs). With
 * totals_host NULL the call only enqueues work: the totals stay in `seg` (diso_totals_offset), and
 * diso_emit / diso_backward take capacities instead of counts, so a whole step can be captured in
 * a HIP graph. */
int diso_count(int algo, int dtype, const void *sdf, int64_t X, int64_t Y, int64_t Z, double iso,
               int flags, void *seg, int64_t seg_bytes, void *ctl, int64_t ctl_bytes,
               diso_totals_t *totals_host, void *stream);

/* diso_count without a caller workspace for the scan's control block and without any
 * allocation: no counterpart in the reference. For callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). The
 * control block is the library's own 128 KiB array per device (131,136 bytes; it serves every
 * grid that diso_count accepts, and a grid that needed more would be rejected with
 * DISO_E_CAPACITY). Every call on a device shares it, so calls that overlap in time on one device
 * race (as calls sharing one `ctl` do); calls on different devices do not. Results equal
 * diso_count's. With totals_host set, the totals are copied through a stack buffer (no pinned
 * staging is allocated) and the stream is synchronised; with totals_host NULL the call only
 * enqueues work. */
int diso_count_noalloc(int algo, int dtype, const void *sdf, int64_t X, int64_t Y, int64_t Z,
                       double iso, int flags, void *seg, int64_t seg_bytes,
                       diso_totals_t *totals_host, void *stream);

/* Size in bytes of the scratch buffer diso_backward needs for a counted grid (n_crossing from
 * diso_totals_t): none for MC; for DMC, the per-edge upstream gradients, n_crossing * 3
 * scalars. */
int64_t diso_scratch_bytes(int algo, int dtype, int64_t n_crossing);

/* Pass 2: write vertices [n_verts,3] and faces ([n_faces,3] MC triangles, [n_faces,4] DMC quads)
 * in the reference's order. Replaces create_cell_mc_verts_kernel + create_cell_mc_tris_kernel
 * (src/cumc.cu:370-410,564-649) and create_dmc_verts_kernel + create_quads_kernel
 * (src/cudualmc.cu:907-955,1027-1056), plus the wrapper's shift / normalise / int64 epilogue
 * (diso/__init__.py:56-61,110-116) when the corresponding flags are set. Writes never exceed
 * vert_capacity / face_capacity (rows); if the counts from diso_count do not fit, nothing is
 * written. DMC only: `vinfo` receives one byte per DMC vertex (edges in its patch | patch
 * rank << 3) that diso_backward needs. */
int diso_emit(int algo, int dtype, const void *sdf, const void *deform, int64_t X, int64_t Y,
              int64_t Z, double iso, int flags, const void *seg, int64_t seg_bytes, void *verts,
              int64_t vert_capacity, void *faces, int64_t face_capacity, void *vinfo, void *stream);

/* Backward: gradients of sum(adj_verts * verts) with respect to sdf [X,Y,Z] and deform
 * [X,Y,Z,3] (adj_deform may be NULL). One thread per grid vertex gathers from its at most six
 * incident crossing edges, so there are no atomics and results are deterministic. Replaces
 * adj_create_cell_mc_verts_kernel (src/cumc.cu:474-512,734-743) and adj_create_dmc_verts_kernel
 * (src/cudualmc.cu:957-1005,1130-1138), plus the chain rule through the wrapper's normalise and
 * the slicing of F.pad's backward. adj_verts is the gradient with respect to the vertices as
 * emitted with the same flags. Gradients are written for every grid vertex (zero where no
 * crossing edge touches it) unless DISO_FLAG_ACCUMULATE is set. `seg` must hold the result of a
 * diso_count on the same inputs. DMC only: `faces` / `vinfo` are the quads (in the index width
 * of `flags`) and vertex info written by diso_emit with the same inputs, n_faces the quad count,
 * and `scratch` (diso_scratch_bytes) is overwritten. n_faces and n_verts may be capacities (rows
 * allocated) rather than the counts: rows past the counts in `seg` are never read. */
int diso_backward(int algo, int dtype, const void *sdf, const void *deform, int64_t X, int64_t Y,
                  int64_t Z, double iso, int flags, const void *seg, int64_t seg_bytes,
                  const void *faces, int64_t n_faces, const void *vinfo, const void *adj_verts,
                  int64_t n_verts, void *adj_sdf, void *adj_deform, void *scratch,
                  int64_t scratch_bytes, void *stream);

/* Size in bytes of the scratch buffer diso_split needs for n_quads quads. */
int64_t diso_split_scratch_bytes(int64_t n_quads);

/* DiffDMC quad -> triangle split (replaces diso/__init__.py:117-147): each quad [a,b,c,d] of
 * `quads` ([n_quads,4], int64 with DISO_FLAG_INDEX64 in `flags`, else int32; indices into `verts`
 * [n_verts,3] of `dtype`) is cut along the diagonal whose worse triangle has the smaller largest
 * corner cosine (F.normalize semantics, eps 1e-12): (a,b,d)+(b,c,d) when its score is strictly
 * smaller, else (a,b,c)+(a,c,d). `tris` ([2*n_quads,3] int64) receives the pairs of all quads
 * taking the first cut, in quad order, then those of all the others. `scratch` is overwritten. */
int diso_split(int dtype, const void *verts, int64_t n_verts, const void *quads, int64_t n_quads, int flags,
               int64_t *tris, int64_t tri_capacity, void *scratch, int64_t scratch_bytes, void *stream);

/* ---- Batches: many grids (uneven dims, MC and DMC, with or without deform) in one launch per
 * pass. The reference has no batched form: it extracts grid by grid (diso/__init__.py:9-147). Each grid is described by a diso_grid_t; the whole batch shares one
 * dtype, one workspace ("seg", diso_batch_seg_bytes) and one control block ("ctl",
 * diso_batch_ctl_bytes: zero-filled once, then reused; calls sharing it must not overlap).
 *   1. diso_batch_count: classify and count every grid in one launch (one single-pass scan over
 *      all grids' segments), then fill each grid's `totals` (one host synchronisation).
 *   2. the caller sizes each grid's outputs from its totals and sets verts / faces / vinfo and
 *      the capacities, then diso_batch_emit writes every grid's vertices and faces.
 *   3. later, with the same workspace: set adj_verts / adj_sdf / adj_deform (grids whose adj_sdf is
 *      NULL are skipped) and call diso_batch_backward.
 * Per grid, results are identical to diso_count / diso_emit / diso_backward with the same flags.
 * Each call copies the descriptors to the device and waits for that copy. */
typedef struct {
    int32_t algo;  /* DISO_MC or DISO_DMC */
    int32_t flags; /* DISO_FLAG_* for this grid */
    int64_t X, Y, Z;
    double iso;
    const void *sdf;    /* [X,Y,Z] */
    const void *deform; /* [X,Y,Z,3] or NULL */
    void *verts;        /* outputs, rows as in diso_emit */
    void *faces;
    void *vinfo;
    int64_t vert_capacity, face_capacity;
    const void *adj_verts; /* backward */
    void *adj_sdf;
    void *adj_deform;
    diso_totals_t totals; /* written by diso_batch_count */
} diso_grid_t;

int64_t diso_batch_seg_bytes(const diso_grid_t *grids, int n);
int64_t diso_batch_ctl_bytes(int n);
int diso_batch_count(int dtype, diso_grid_t *grids, int n, void *seg, int64_t seg_bytes, void *ctl,
                     int64_t ctl_bytes, void *stream);
int diso_batch_emit(int dtype, const diso_grid_t *grids, int n, void *seg, int64_t seg_bytes, void *stream);
/* Scratch of diso_batch_backward: the per-edge gradients of all DMC grids. */
int64_t diso_batch_scratch_bytes(int dtype, const diso_grid_t *grids, int n);
int diso_batch_backward(int dtype, const diso_grid_t *grids, int n, void *seg, int64_t seg_bytes,
                        void *scratch, int64_t scratch_bytes, void *stream);

/* Launch parameters, process-wide (set them before launching work; the autotuner in
 * diso_amd/autotune.py picks them per machine and grid size). Results do not depend on them.
 * No counterpart in the reference, whose block size is a compile-time constant
 * (src/cumc.cu:93, BLOCK_SIZE 512).
 *  DISO_PARAM_RUN: segments per emission / backward block, 1..64 (default 32).
 *  DISO_PARAM_COUNT_BLOCKS_PER_CU: resident blocks per CU of the count pass, 1..4 (default 4).
 *  DISO_PARAM_RECOUNT_SPINS: reads of a predecessor's unpublished aggregate before a count block
 *    counts that range itself, 0..2^20 (default 8192; 0 counts every predecessor not already
 *    published, which tests use to exercise that path).
 * The first two have one value per size bucket. A launch takes the bucket of its segment count
 * (a batched count: of all its grids together; a batched emission or backward: each grid its own
 * run length); buckets 0..4 hold grids of about 32^3, 64^3,
 * 128^3, 256^3 and 512^3 samples (limits 3,200 / 21,000 / 130,000 / 900,000 segments).
 * diso_set_param sets every bucket; diso_get_param returns bucket 3's value. */
enum {
    DISO_PARAM_RUN = 0,
    DISO_PARAM_COUNT_BLOCKS_PER_CU = 1,
    DISO_PARAM_RECOUNT_SPINS = 2
};
#define DISO_PARAM_BUCKETS 5
int diso_set_param(int which, int value);
int diso_get_param(int which);
int diso_set_bucket_param(int which, int bucket, int value);
int diso_get_bucket_param(int which, int bucket);
/* Bucket of a single-grid launch on an X x Y x Z grid (flags as for diso_count). */
int diso_param_bucket(int64_t X, int64_t Y, int64_t Z, int flags);

/* Reads the totals of the last diso_count on `seg` (blocking copy on `stream`). */
int diso_read_totals(const void *seg, int64_t seg_bytes, int64_t X, int64_t Y, int64_t Z,
                     int flags, diso_totals_t *totals_host, void *stream);

/* Thread-local message for the last failing call on this thread ("" if none). */
const char *diso_last_error(void);

/* Library version string. */
const char *diso_version(void);

#ifdef __cplusplus
}
#endif

#endif /* DISO_HIP_H */
