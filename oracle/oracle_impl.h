/* CPU restatement of the reference's per-cell MC / DMC engine, one scalar type per inclusion.
 *
 * TEST INFRASTRUCTURE ONLY. This is the oracle: tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it as the checker / CPU baseline. The product (diso_amd) never links,
 * loads or calls it.
 *
 * Included by oracle.c with S (scalar type) and FN(name) (name mangling) defined. Every function
 * restates the reference kernel it names, in the same operation order, so that fp32 results are
 * reproducible bit-for-bit by any IEEE implementation without contraction (-ffp-contract=off).
 */

typedef struct {
    S x, y, z;
} FN(vtx);

/* clamp(x, a, b) = min(max(a, x), b) with max(a,b) = a > b ? a : b, min(a,b) = a < b ? a : b
 * (src/cumc.cu:80-91). */
static inline S FN(clamp)(S v, S lo, S hi) {
    S m = lo > v ? lo : v;
    return m < hi ? m : hi;
}

static inline int64_t FN(gA)(const oracle_state *st, int64_t x, int64_t y, int64_t z) {
    return z + st->dims[2] * (y + st->dims[1] * x); /* src/cumc.h:101-105 */
}

/* computeMcVert (src/cumc.cu:343-368, identical in src/cudualmc.cu:683-708). */
static FN(vtx) FN(mc_vert)(const oracle_state *st, const S *grid, const S *deform, int64_t x,
                          int64_t y, int64_t z, int dim, S iso) {
    int64_t o[3] = {0, 0, 0};
    o[dim] = 1;
    int64_t v0 = FN(gA)(st, x, y, z);
    int64_t v1 = FN(gA)(st, x + o[0], y + o[1], z + o[2]);
    S d0 = grid[v0], d1 = grid[v1];
    S t = (d1 != d0) ? FN(clamp)((iso - d0) / (d1 - d0), (S)0, (S)1) : (S)0.5;
    S p0x = (S)x, p0y = (S)y, p0z = (S)z;
    S p1x = (S)(x + o[0]), p1y = (S)(y + o[1]), p1z = (S)(z + o[2]);
    if (deform) {
        p0x += deform[3 * v0 + 0];
        p0y += deform[3 * v0 + 1];
        p0z += deform[3 * v0 + 2];
        p1x += deform[3 * v1 + 0];
        p1y += deform[3 * v1 + 1];
        p1z += deform[3 * v1 + 2];
    }
    FN(vtx) r;
    r.x = p0x + (p1x - p0x) * t;
    r.y = p0y + (p1y - p0y) * t;
    r.z = p0z + (p1z - p0z) * t;
    return r;
}

/* adjComputeMcVert (src/cumc.cu:412-453, identical in src/cudualmc.cu:710-751): the adjoint of
 * one edge vertex, scattered into adj_grid / adj_deform. `par` selects atomic accumulation (the
 * reference's atomicAdd) for the multi-threaded baseline; the checker runs it serially, which
 * fixes the accumulation order to the reference's sequential cell order. */
static void FN(adj_mc_vert)(const oracle_state *st, const S *grid, const S *deform, int64_t x,
                            int64_t y, int64_t z, int dim, S iso, S *adj_grid, S *adj_deform,
                            S gx, S gy, S gz, int par) {
    int64_t o[3] = {0, 0, 0};
    o[dim] = 1;
    int64_t v0 = FN(gA)(st, x, y, z);
    int64_t v1 = FN(gA)(st, x + o[0], y + o[1], z + o[2]);
    S d0 = grid[v0], d1 = grid[v1];
    S t = (d1 != d0) ? FN(clamp)((iso - d0) / (d1 - d0), (S)0, (S)1) : (S)0.5;
    S p0x = (S)x, p0y = (S)y, p0z = (S)z;
    S p1x = (S)(x + o[0]), p1y = (S)(y + o[1]), p1z = (S)(z + o[2]);
    if (deform) {
        p0x += deform[3 * v0 + 0];
        p0y += deform[3 * v0 + 1];
        p0z += deform[3 * v0 + 2];
        p1x += deform[3 * v1 + 0];
        p1y += deform[3 * v1 + 1];
        p1z += deform[3 * v1 + 2];
    }
    S omt = (S)1 - t;
    S a0x = omt * gx, a0y = omt * gy, a0z = omt * gz;
    S a1x = t * gx, a1y = t * gy, a1z = t * gz;
    S adj_t = (p1x - p0x) * gx + (p1y - p0y) * gy + (p1z - p0z) * gz;
    S dd = (d1 - d0) * (d1 - d0);
    S ad0 = (iso - d1) / dd * adj_t;
    S ad1 = (d0 - iso) / dd * adj_t;
    if (deform && adj_deform) {
        S *q0 = adj_deform + 3 * v0, *q1 = adj_deform + 3 * v1;
        if (par) {
#pragma omp atomic
            q0[0] += a0x;
#pragma omp atomic
            q0[1] += a0y;
#pragma omp atomic
            q0[2] += a0z;
#pragma omp atomic
            q1[0] += a1x;
#pragma omp atomic
            q1[1] += a1y;
#pragma omp atomic
            q1[2] += a1z;
        } else {
            q0[0] += a0x;
            q0[1] += a0y;
            q0[2] += a0z;
            q1[0] += a1x;
            q1[1] += a1y;
            q1[2] += a1z;
        }
    }
    if (par) {
#pragma omp atomic
        adj_grid[v0] += ad0;
#pragma omp atomic
        adj_grid[v1] += ad1;
    } else {
        adj_grid[v0] += ad0;
        adj_grid[v1] += ad1;
    }
}

/* Cube code. MC: Bourke corner order, no ambiguity resolution (getCellCode / getGoodCellCode,
 * src/cumc.cu:455-538, the neighbour test commented out there). DMC: Morton order
 * (src/cudualmc.cu:796-813). */
static int FN(cell_code)(const oracle_state *st, const S *grid, int64_t x, int64_t y, int64_t z,
                         S iso) {
    int code = 0;
    for (int i = 0; i < 8; ++i) {
        int cx, cy, cz;
        if (st->algo == ORACLE_MC) {
            cx = MC_CORNER[i][0];
            cy = MC_CORNER[i][1];
            cz = MC_CORNER[i][2];
        } else {
            cx = i & 1;
            cy = (i >> 1) & 1;
            cz = (i >> 2) & 1;
        }
        if (grid[FN(gA)(st, x + cx, y + cy, z + cz)] >= iso) code |= 1 << i;
    }
    return code;
}

/* DMC getGoodCellCode (src/cudualmc.cu:815-839): an ambiguous configuration is inverted when
 * its face neighbour in the tabulated direction exists and is ambiguous as well. */
static int FN(good_code)(const oracle_state *st, const S *grid, int64_t x, int64_t y, int64_t z,
                         S iso) {
    int code = FN(cell_code)(st, grid, x, y, z, iso);
    int dir = DISO_DMC_DIR(DISO_DMC_CASE[code]);
    if (dir != 7) {
        int64_t n[3] = {x, y, z};
        int comp = dir >> 1;
        n[comp] += (dir & 1) ? 1 : -1;
        if (n[comp] >= 0 && n[comp] < st->dims[comp] - 1) {
            int ncode = FN(cell_code)(st, grid, n[0], n[1], n[2], iso);
            if (DISO_DMC_DIR(DISO_DMC_CASE[ncode]) != 7) code ^= 0xff;
        }
    }
    return code;
}

static inline int FN(crosses)(S d0, S d1, S iso) {
    return (d0 < iso && d1 >= iso) || (d1 < iso && d0 >= iso); /* src/cumc.cu:334-339 */
}

/* Shared first half of both forwards: count_used_cells_kernel + ExclusiveSum +
 * index_used_cells_kernel + count_cell_mc_verts_kernel + ExclusiveSum
 * (src/cumc.cu:265-341,663-694; src/cudualmc.cu:605-681,1070-1101). */
static int FN(used_cells_and_edges)(oracle_state *st, const S *grid, S iso) {
    int64_t n = st->n_cells;
    st->first_cell_used = (int64_t *)calloc((size_t)n + 1, sizeof(int64_t));
    if (!st->first_cell_used) return -1;
#pragma omp parallel for num_threads(st->threads) schedule(static)
    for (int64_t c = 0; c < n; ++c) {
        int64_t x = c / (st->dims[2] * st->dims[1]);
        int64_t y = (c / st->dims[2]) % st->dims[1];
        int64_t z = c % st->dims[2];
        st->first_cell_used[c] = 0;
        if (x >= st->dims[0] - 1 || y >= st->dims[1] - 1 || z >= st->dims[2] - 1) continue;
        int code = FN(cell_code)(st, grid, x, y, z, iso);
        if (code != 0 && code != 255) st->first_cell_used[c] = 1;
    }
    st->n_used = oracle_exclusive_scan(st->first_cell_used, n + 1, st->threads);
    st->used_cell_index = (int64_t *)malloc(sizeof(int64_t) * (size_t)(st->n_used + 1));
    st->vbase = (int64_t *)calloc((size_t)st->n_used + 1, sizeof(int64_t));
    st->fbase = (int64_t *)calloc((size_t)st->n_used + 1, sizeof(int64_t));
    st->code = (uint8_t *)calloc((size_t)st->n_used + 1, 1);
    if (!st->used_cell_index || !st->vbase || !st->fbase || !st->code) return -1;
#pragma omp parallel for num_threads(st->threads) schedule(static)
    for (int64_t c = 0; c < n; ++c) {
        int64_t u = st->first_cell_used[c];
        if (st->first_cell_used[c + 1] - u) st->used_cell_index[u] = c;
    }
#pragma omp parallel for num_threads(st->threads) schedule(static)
    for (int64_t u = 0; u < st->n_used; ++u) {
        int64_t c = st->used_cell_index[u];
        int64_t x = c / (st->dims[2] * st->dims[1]);
        int64_t y = (c / st->dims[2]) % st->dims[1];
        int64_t z = c % st->dims[2];
        S d0 = grid[c];
        S dx = grid[FN(gA)(st, x + 1, y, z)];
        S dy = grid[FN(gA)(st, x, y + 1, z)];
        S dz = grid[FN(gA)(st, x, y, z + 1)];
        st->vbase[u] = FN(crosses)(d0, dx, iso) + FN(crosses)(d0, dy, iso) + FN(crosses)(d0, dz, iso);
    }
    st->n_verts_mc = oracle_exclusive_scan(st->vbase, st->n_used + 1, st->threads);
    return 0;
}

/* CuMC::forward (src/cumc.cu:651-732). */
static int FN(mc_forward)(oracle_state *st, const S *grid, const S *deform, S iso) {
    if (FN(used_cells_and_edges)(st, grid, iso)) return -1;
    int64_t nv = st->n_verts_mc;
    st->n_verts = nv;
    st->verts = malloc(sizeof(S) * 3 * (size_t)(nv + 1));
    st->verts_type = (int8_t *)malloc((size_t)nv + 1);
    if (!st->verts || !st->verts_type) return -1;
    S *verts = (S *)st->verts;
    /* create_cell_mc_verts_kernel (src/cumc.cu:370-410) */
#pragma omp parallel for num_threads(st->threads) schedule(static)
    for (int64_t u = 0; u < st->n_used; ++u) {
        int64_t c = st->used_cell_index[u];
        int64_t x = c / (st->dims[2] * st->dims[1]);
        int64_t y = (c / st->dims[2]) % st->dims[1];
        int64_t z = c % st->dims[2];
        S d0 = grid[c];
        S ds[3] = {grid[FN(gA)(st, x + 1, y, z)], grid[FN(gA)(st, x, y + 1, z)],
                   grid[FN(gA)(st, x, y, z + 1)]};
        int64_t first = st->vbase[u];
        for (int dim = 0; dim < 3; ++dim) {
            if (FN(crosses)(d0, ds[dim], iso)) {
                FN(vtx) p = FN(mc_vert)(st, grid, deform, x, y, z, dim, iso);
                st->verts_type[first] = (int8_t)dim;
                verts[3 * first + 0] = p.x;
                verts[3 * first + 1] = p.y;
                verts[3 * first + 2] = p.z;
                ++first;
            }
        }
    }
    /* count_cell_mc_tris_kernel (src/cumc.cu:540-562) */
#pragma omp parallel for num_threads(st->threads) schedule(static)
    for (int64_t u = 0; u < st->n_used; ++u) {
        int64_t c = st->used_cell_index[u];
        int64_t x = c / (st->dims[2] * st->dims[1]);
        int64_t y = (c / st->dims[2]) % st->dims[1];
        int64_t z = c % st->dims[2];
        int code = FN(cell_code)(st, grid, x, y, z, iso);
        st->code[u] = (uint8_t)code;
        st->fbase[u] = 3 * DISO_MC_NTRI(DISO_MC_CASE[code]);
    }
    int64_t n_ids = oracle_exclusive_scan(st->fbase, st->n_used + 1, st->threads);
    st->n_faces = n_ids / 3;
    st->faces = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n_ids + 1));
    if (!st->faces) return -1;
    int bad = 0;
    /* create_cell_mc_tris_kernel (src/cumc.cu:564-610): the owner cell of each table edge is
     * found through the used-cell scan, then its vertex by a linear search over verts_type. */
#pragma omp parallel for num_threads(st->threads) schedule(static) reduction(| : bad)
    for (int64_t u = 0; u < st->n_used; ++u) {
        int64_t c = st->used_cell_index[u];
        int64_t x = c / (st->dims[2] * st->dims[1]);
        int64_t y = (c / st->dims[2]) % st->dims[1];
        int64_t z = c % st->dims[2];
        uint64_t w = DISO_MC_CASE[st->code[u]];
        int num = 3 * DISO_MC_NTRI(w);
        int64_t out = st->fbase[u];
        for (int i = 0; i < num; ++i) {
            int eid = DISO_MC_EDGE(w, i);
            int64_t ox = x + DISO_EDGE_OFS[eid][0], oy = y + DISO_EDGE_OFS[eid][1],
                    oz = z + DISO_EDGE_OFS[eid][2];
            int axis = DISO_EDGE_AXIS[eid];
            int64_t eu = st->first_cell_used[FN(gA)(st, ox, oy, oz)];
            int64_t id;
            if (eu >= st->n_used) { /* past the last used cell: the reference reads out of bounds */
                bad = 1;
                id = 0;
            } else {
                int64_t v_num = st->vbase[eu + 1] - st->vbase[eu];
                id = st->vbase[eu];
                for (int64_t k = 0; k < v_num; ++k) {
                    if (st->verts_type[id + k] == axis) {
                        id += k;
                        break;
                    }
                }
            }
            st->faces[out + i] = id;
        }
    }
    st->status = bad ? ORACLE_STATUS_UNPADDED : 0;
    return 0;
}

/* CuMC::backward -> adj_create_cell_mc_verts_kernel (src/cumc.cu:474-512,734-743). */
static void FN(mc_backward)(oracle_state *st, const S *grid, const S *deform, S iso,
                            const S *adj_verts, S *adj_grid, S *adj_deform) {
    int par = st->threads > 1;
#pragma omp parallel for num_threads(st->threads) schedule(static)
    for (int64_t u = 0; u < st->n_used; ++u) {
        int64_t c = st->used_cell_index[u];
        int64_t x = c / (st->dims[2] * st->dims[1]);
        int64_t y = (c / st->dims[2]) % st->dims[1];
        int64_t z = c % st->dims[2];
        S d0 = grid[c];
        S ds[3] = {grid[FN(gA)(st, x + 1, y, z)], grid[FN(gA)(st, x, y + 1, z)],
                   grid[FN(gA)(st, x, y, z + 1)]};
        int64_t first = st->vbase[u];
        for (int dim = 0; dim < 3; ++dim) {
            if (FN(crosses)(d0, ds[dim], iso)) {
                const S *g = adj_verts + 3 * first;
                FN(adj_mc_vert)(st, grid, deform, x, y, z, dim, iso, adj_grid, adj_deform, g[0],
                                g[1], g[2], par);
                ++first;
            }
        }
    }
}

/* CUDualMC::forward (src/cudualmc.cu:1058-1128). */
static int FN(dmc_forward)(oracle_state *st, const S *grid, const S *deform, S iso) {
    if (FN(used_cells_and_edges)(st, grid, iso)) return -1;
    int64_t nq = st->n_verts_mc;
    st->n_faces = nq;
    st->mc_vert_to_cell = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nq + 1));
    st->mc_vert_type = (uint8_t *)malloc((size_t)nq + 1);
    if (!st->mc_vert_to_cell || !st->mc_vert_type) return -1;
    /* index_cell_mc_verts_kernel (src/cudualmc.cu:753-794) */
#pragma omp parallel for num_threads(st->threads) schedule(static)
    for (int64_t u = 0; u < st->n_used; ++u) {
        int64_t c = st->used_cell_index[u];
        int64_t x = c / (st->dims[2] * st->dims[1]);
        int64_t y = (c / st->dims[2]) % st->dims[1];
        int64_t z = c % st->dims[2];
        S d0 = grid[c];
        S ds[3] = {grid[FN(gA)(st, x + 1, y, z)], grid[FN(gA)(st, x, y + 1, z)],
                   grid[FN(gA)(st, x, y, z + 1)]};
        int64_t first = st->vbase[u];
        for (int dim = 0; dim < 3; ++dim) {
            int entering = d0 < iso && ds[dim] >= iso;
            int exiting = ds[dim] < iso && d0 >= iso;
            if (entering || exiting) {
                st->mc_vert_to_cell[first] = u;
                st->mc_vert_type[first] = (uint8_t)((exiting ? 3 : 0) + dim);
                ++first;
            }
        }
    }
    /* count_cell_patches_kernel (src/cudualmc.cu:841-863) */
#pragma omp parallel for num_threads(st->threads) schedule(static)
    for (int64_t u = 0; u < st->n_used; ++u) {
        int64_t c = st->used_cell_index[u];
        int64_t x = c / (st->dims[2] * st->dims[1]);
        int64_t y = (c / st->dims[2]) % st->dims[1];
        int64_t z = c % st->dims[2];
        int code = FN(good_code)(st, grid, x, y, z, iso);
        st->code[u] = (uint8_t)code;
        st->fbase[u] = DISO_DMC_NPATCH(DISO_DMC_CASE[code]);
    }
    int64_t nv = oracle_exclusive_scan(st->fbase, st->n_used + 1, st->threads);
    st->n_verts = nv;
    st->verts = malloc(sizeof(S) * 3 * (size_t)(nv + 1));
    st->faces = (int64_t *)malloc(sizeof(int64_t) * 4 * (size_t)(nq + 1));
    if (!st->verts || !st->faces) return -1;
    S *verts = (S *)st->verts;
    /* create_dmc_verts_kernel (src/cudualmc.cu:907-955): each patch vertex is the mean of its
     * MC edge vertices, accumulated from zero in ascending edge order, scaled by 1/num. */
#pragma omp parallel for num_threads(st->threads) schedule(static)
    for (int64_t u = 0; u < st->n_used; ++u) {
        int64_t c = st->used_cell_index[u];
        int64_t x = c / (st->dims[2] * st->dims[1]);
        int64_t y = (c / st->dims[2]) % st->dims[1];
        int64_t z = c % st->dims[2];
        uint64_t w = DISO_DMC_CASE[st->code[u]];
        int np = DISO_DMC_NPATCH(w);
        int64_t first = st->fbase[u];
        for (int k = 0; k < np; ++k) {
            S px = 0, py = 0, pz = 0, num = 0;
            for (int e = 0; e < 12; ++e) {
                if (DISO_DMC_PATCH_OF(w, e) != k) continue;
                FN(vtx) q = FN(mc_vert)(st, grid, deform, x + DISO_EDGE_OFS[e][0],
                                       y + DISO_EDGE_OFS[e][1], z + DISO_EDGE_OFS[e][2],
                                       DISO_EDGE_AXIS[e], iso);
                px += q.x;
                py += q.y;
                pz += q.z;
                num += (S)1;
            }
            S r = (S)1 / num;
            verts[3 * first + 0] = px * r;
            verts[3 * first + 1] = py * r;
            verts[3 * first + 2] = pz * r;
            ++first;
        }
    }
    /* create_quads_kernel + get_quad_index (src/cudualmc.cu:1019-1056) */
    int bad = 0;
#pragma omp parallel for num_threads(st->threads) schedule(static) reduction(| : bad)
    for (int64_t q = 0; q < nq; ++q) {
        int64_t u = st->mc_vert_to_cell[q];
        int64_t c = st->used_cell_index[u];
        int64_t x = c / (st->dims[2] * st->dims[1]);
        int64_t y = (c / st->dims[2]) % st->dims[1];
        int64_t z = c % st->dims[2];
        int ty = st->mc_vert_type[q];
        for (int i = 0; i < 4; ++i) {
            int64_t nx = x + DISO_QUAD[ty][i][0], ny = y + DISO_QUAD[ty][i][1],
                    nz = z + DISO_QUAD[ty][i][2];
            int eid = DISO_QUAD[ty][i][3];
            int64_t id = 0;
            if (nx < 0 || ny < 0 || nz < 0) {
                bad = 1;
            } else {
                int64_t eu = st->first_cell_used[FN(gA)(st, nx, ny, nz)];
                if (eu >= st->n_used) {
                    bad = 1;
                } else {
                    int off = DISO_DMC_PATCH_OF(DISO_DMC_CASE[st->code[eu]], eid);
                    id = st->fbase[eu] + (off == 7 ? -1 : off);
                }
            }
            st->faces[4 * q + i] = id;
        }
    }
    st->status = bad ? ORACLE_STATUS_UNPADDED : 0;
    return 0;
}

/* CUDualMC::backward -> adj_create_dmc_verts_kernel (src/cudualmc.cu:957-1005,1130-1138).
 * reference_compat = 1 reproduces the reference's read of adj_verts[first] for every patch of a
 * cell (src/cudualmc.cu:990, `first` is never advanced); 0 reads each patch's own vertex, which
 * is the true gradient of the forward. */
static void FN(dmc_backward)(oracle_state *st, const S *grid, const S *deform, S iso,
                             const S *adj_verts, S *adj_grid, S *adj_deform, int reference_compat) {
    int par = st->threads > 1;
#pragma omp parallel for num_threads(st->threads) schedule(static)
    for (int64_t u = 0; u < st->n_used; ++u) {
        int64_t c = st->used_cell_index[u];
        int64_t x = c / (st->dims[2] * st->dims[1]);
        int64_t y = (c / st->dims[2]) % st->dims[1];
        int64_t z = c % st->dims[2];
        uint64_t w = DISO_DMC_CASE[st->code[u]];
        int np = DISO_DMC_NPATCH(w);
        int64_t first = st->fbase[u];
        for (int k = 0; k < np; ++k) {
            S num = 0;
            for (int e = 0; e < 12; ++e)
                if (DISO_DMC_PATCH_OF(w, e) == k) num += (S)1;
            const S *g = adj_verts + 3 * (reference_compat ? first : first + k);
            S r = (S)1 / num;
            S gx = g[0] * r, gy = g[1] * r, gz = g[2] * r;
            for (int e = 0; e < 12; ++e) {
                if (DISO_DMC_PATCH_OF(w, e) != k) continue;
                FN(adj_mc_vert)(st, grid, deform, x + DISO_EDGE_OFS[e][0],
                                y + DISO_EDGE_OFS[e][1], z + DISO_EDGE_OFS[e][2],
                                DISO_EDGE_AXIS[e], iso, adj_grid, adj_deform, gx, gy, gz, par);
            }
        }
    }
}
