// diso_hip.hip - DiffMC / DiffDMC engine for MI355X (gfx950), behind the C ABI of diso_hip.h.
//
// Work decomposition
//   The (virtually padded) grid G = [g0,g1,g2] is cut into "segments": one z-row (px,py) of the
//   grid, split into chunks of 64 consecutive z positions. One wave64 owns one segment, lane l
//   the position pz = 64*chunk + l. Segment order (px, py, chunk) is the grid's linear order
//   (x-major, z fastest), which is the order in which the reference emits its vertices and faces
//   (used cells in increasing linear cell id, src/cumc.cu:299-311,370-410,564-610). So the output
//   position of any cell's first vertex / face is
//       base(segment) + exclusive prefix, within the wave, of the per-cell counts,
//   and only the per-segment bases are stored (16 bytes per 64 cells). The count pass also stores
//   one 16-bit classification word per position (code, owned crossings, inside bit, face count),
//   which the later passes read instead of re-classifying (the recompute-classification backward
//   re-derives the crossings from the samples instead). In-wave prefixes come from bit-plane
//   ballots and v_mbcnt.
//
// Kernels
//   count_kernel          classify + count + single-pass decoupled look-back scan of segment
//                         counts (also over the segments of a batch of grids).
//   emit_mc_kernel        MC vertices and triangles (+ shift / normalise / int64 epilogue).
//   emit_dmc_kernel       DMC vertices (patch means) and quads.
//   dmc_edge_grad_kernel  DMC backward, first step: per crossing edge, the sum of its <= 4 quad
//                         corners' vertex gradients / patch sizes.
//   bwd_mc_kernel         MC gather backward (DMC: on those edge gradients): one lane per grid
//                         vertex sums its <= 6 incident edges, deterministic, no atomics.
//   split_*_kernel        DiffDMC's quad -> triangle split.
//   *_batch_kernel        the same passes over many grids in one launch.
//
// Floating point: forward results use only IEEE +, -, *, /, sqrt and compares, in the
// reference's operation order, built with -ffp-contract=off and correctly rounded division, so
// fp32 vertices equal an fp32 CPU evaluation of the same expressions bit for bit. The fp32
// backward uses one hardware reciprocal per edge (gradients are compared within a tolerance).
#include <hip/hip_runtime.h>

#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <mutex>
#include <vector>

#include "diso_hip.h"
#include "diso_scan_word.h"
#include "diso_tables.h"

namespace {

constexpr int kWave = 64;
constexpr int kWavesPerBlock = 4;
constexpr int kBlock = kWave * kWavesPerBlock;
constexpr uint32_t kSpinLimit = 1u << 20;  // bounded look-back wait (well under a second); never expected
constexpr int kCtlHeader = 64;             // bytes

__constant__ uint64_t kMcCase[256] = {DISO_MC_CASE_VALUES};
__constant__ uint64_t kDmcCase[256] = {DISO_DMC_CASE_VALUES};

// ------------------------------------------------------------------------------------------
// Geometry and sampling

struct Geo {
    int g0, g1, g2;  // conceptual grid (padded dims when padding is virtual)
    int n0, n1, n2;  // stored grid dims
    int off;         // stored sample (i,j,k) sits at conceptual (i+off, j+off, k+off)
    int nchunk;      // ceil(g2 / 64)
    int64_t nseg;    // g0 * g1 * nchunk
};

template <typename S>
struct In {
    const S *sdf;
    const S *deform;  // [n0,n1,n2,3] or nullptr
    S iso;
    S pad;            // value of samples outside the stored grid (iso + 1)
    Geo geo;
};

template <typename S>
struct V3 {
    S x, y, z;
};

// Stores of outputs that no later kernel reads (vertices, faces, dense gradients): non-temporal,
// so these streams do not evict from L2 the inputs that the gathers re-read.
template <typename T>
__device__ __forceinline__ void put(T *p, T v) {
    __builtin_nontemporal_store(v, p);
}

// (Bitwise &, not &&: no short-circuit branch, so loads guarded by it stay branch-free.)
template <typename S>
__device__ __forceinline__ bool in_data(const In<S> &in, int x, int y, int z) {
    return ((unsigned)(x - in.geo.off) < (unsigned)in.geo.n0) & ((unsigned)(y - in.geo.off) < (unsigned)in.geo.n1) &
           ((unsigned)(z - in.geo.off) < (unsigned)in.geo.n2);
}

template <typename S>
__device__ __forceinline__ int64_t data_index(const In<S> &in, int x, int y, int z) {
    return ((int64_t)(x - in.geo.off) * in.geo.n1 + (y - in.geo.off)) * in.geo.n2 + (z - in.geo.off);
}

// sdf sample at conceptual (x,y,z); samples outside the stored grid read as the pad value, so no
// coordinate can address memory outside the input.
// (Branch-free: a sample outside the stored grid loads element 0 and is replaced by the pad value,
// so no lane needs its own execution mask around the load.)
template <typename S>
__device__ __forceinline__ S sample(const In<S> &in, int x, int y, int z) {
    const bool ok = in_data(in, x, y, z);
    const S v = in.sdf[ok ? data_index(in, x, y, z) : 0];
    return ok ? v : in.pad;
}

template <typename S, bool DEF>
__device__ __forceinline__ V3<S> deform_at(const In<S> &in, int x, int y, int z) {
    V3<S> r = {S(0), S(0), S(0)};
    if (DEF) {
        const bool ok = in_data(in, x, y, z);
        const S *d = in.deform + 3 * (ok ? data_index(in, x, y, z) : 0);
        const S dx = d[0], dy = d[1], dz = d[2];
        r.x = ok ? dx : S(0);
        r.y = ok ? dy : S(0);
        r.z = ok ? dz : S(0);
    }
    return r;
}

// A cell is anchored at (x,y,z) when its +1 corner lies inside the grid (src/cumc.cu:278).
template <typename S>
__device__ __forceinline__ bool is_anchor(const In<S> &in, int x, int y, int z) {
    return x >= 0 && y >= 0 && z >= 0 && x < in.geo.g0 - 1 && y < in.geo.g1 - 1 && z < in.geo.g2 - 1;
}

template <typename S>
__device__ __forceinline__ bool crosses(S d0, S d1, S iso) {
    return (d0 < iso && d1 >= iso) || (d1 < iso && d0 >= iso);  // src/cumc.cu:334-339
}

// clamp(x, a, b) = min(max(a, x), b) as in src/cumc.cu:80-91.
template <typename S>
__device__ __forceinline__ S clamp01(S v) {
    S m = S(0) > v ? S(0) : v;
    return m < S(1) ? m : S(1);
}

template <typename S>
__device__ __forceinline__ S interp_t(S d0, S d1, S iso) {
    return (d1 != d0) ? clamp01<S>((iso - d0) / (d1 - d0)) : S(0.5);  // src/cumc.cu:356
}

// The eight corner samples of the cell anchored at (x,y,z): c[dx][dy][dz].
template <typename S>
struct Corners {
    S c[2][2][2];
};

template <typename S>
__device__ __forceinline__ Corners<S> load_corners(const In<S> &in, int x, int y, int z) {
    Corners<S> k;
#pragma unroll
    for (int dx = 0; dx < 2; ++dx)
#pragma unroll
        for (int dy = 0; dy < 2; ++dy)
#pragma unroll
            for (int dz = 0; dz < 2; ++dz) k.c[dx][dy][dz] = sample(in, x + dx, y + dy, z + dz);
    return k;
}

// MC cube code, Bourke corner order (src/cumc.cu:107,455-472).
template <typename S>
__device__ __forceinline__ int mc_code(const Corners<S> &k, S iso) {
    int code = 0;
    code |= (k.c[0][0][0] >= iso) << 0;
    code |= (k.c[1][0][0] >= iso) << 1;
    code |= (k.c[1][1][0] >= iso) << 2;
    code |= (k.c[0][1][0] >= iso) << 3;
    code |= (k.c[0][0][1] >= iso) << 4;
    code |= (k.c[1][0][1] >= iso) << 5;
    code |= (k.c[1][1][1] >= iso) << 6;
    code |= (k.c[0][1][1] >= iso) << 7;
    return code;
}

// DMC cube code, Morton corner order (src/cudualmc.cu:95-104,796-813).
template <typename S>
__device__ __forceinline__ int dmc_code(const Corners<S> &k, S iso) {
    int code = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) code |= (k.c[i & 1][(i >> 1) & 1][(i >> 2) & 1] >= iso) << i;
    return code;
}

// getGoodCellCode (src/cudualmc.cu:815-839): an ambiguous case is inverted when its face
// neighbour in the tabulated direction exists and is ambiguous too. `cases` is the DMC case table.
template <typename S>
__device__ __forceinline__ int dmc_good_code(const In<S> &in, const uint64_t *cases, int code, int x,
                                             int y, int z) {
    const int dir = DISO_DMC_DIR(cases[code]);
    if (dir != 7) {
        // neighbour along axis comp, +1 or -1; selected arithmetically (no indexed arrays, so the
        // grid parameters can stay in registers)
        const int comp = dir >> 1, step = (dir & 1) ? 1 : -1;
        const int on0 = comp == 0, on1 = comp == 1, on2 = comp == 2;
        const int nx = x + on0 * step, ny = y + on1 * step, nz = z + on2 * step;
        const int c = on0 * nx + on1 * ny + on2 * nz;
        const int lim = on0 * in.geo.g0 + on1 * in.geo.g1 + on2 * in.geo.g2;
        if (c >= 0 && c < lim - 1) {
            const int ncode = dmc_code(load_corners(in, nx, ny, nz), in.iso);
            if (DISO_DMC_DIR(cases[ncode]) != 7) code ^= 0xff;
        }
    }
    return code;
}

// dmc_good_code for a cell whose eight corners are all stored samples (`own` = index of corner
// (0,0,0)): the face neighbour shares four corners with the cell, whose bits come from `code`
// (Morton order: bit (dx | dy << 1 | dz << 2)); only its far face, the plane two steps past the
// cell (step +1) or one step before it (step -1), is read, at the cell's index plus strides.
template <typename S>
__device__ __forceinline__ int dmc_good_code_stored(const In<S> &in, const uint64_t *cases, int code, int x, int y,
                                                    int z, int64_t own) {
    const int dir = DISO_DMC_DIR(cases[code]);
    if (dir == 7) return code;
    const Geo &g = in.geo;
    const int comp = dir >> 1, plus = dir & 1;
    const int c = comp == 0 ? x : (comp == 1 ? y : z);
    const int lim = comp == 0 ? g.g0 : (comp == 1 ? g.g1 : g.g2);
    const int nc = plus ? c + 1 : c - 1;  // the neighbour's anchor along comp
    if (nc < 0 || nc >= lim - 1) return code;
    const int64_t sz = 1, sy = g.n2, sx = (int64_t)g.n1 * g.n2;
    // strides of comp and of the face's two other axes a < b
    const int64_t sc = comp == 0 ? sx : (comp == 1 ? sy : sz);
    const int64_t sa = comp == 0 ? sy : sx, sb = comp == 2 ? sy : sz;
    const int a = comp == 0 ? 1 : 0, b = comp == 2 ? 1 : 2;
    // far plane: conceptual c + 2 (step +1) or c - 1 (step -1); outside the stored grid it is pad
    const int fc = plus ? c + 2 : c - 1;
    const int nfc = comp == 0 ? g.n0 : (comp == 1 ? g.n1 : g.n2);
    const bool stored = (unsigned)(fc - g.off) < (unsigned)nfc;
    const int64_t f0 = own + (plus ? 2 * sc : -sc);
    const int wc = 1 << comp;  // bit weight of comp in the corner index
    constexpr int kLow[3] = {0x55, 0x33, 0x0f};  // corners with offset 0 along x, y, z
    const int low = comp == 0 ? kLow[0] : (comp == 1 ? kLow[1] : kLow[2]);
    // shared face: the neighbour's offset-0 corners are the cell's offset-1 ones (step +1), or its
    // offset-1 corners the cell's offset-0 ones (step -1)
    int ncode = plus ? ((code >> wc) & low) : ((code << wc) & ~low & 0xff);
    const int fbit = plus ? wc : 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int ja = j & 1, jb = j >> 1;
        const S v = stored ? in.sdf[f0 + ja * sa + jb * sb] : in.pad;
        ncode |= (int)(v >= in.iso) << (fbit | (ja << a) | (jb << b));
    }
    return DISO_DMC_DIR(cases[ncode]) != 7 ? code ^ 0xff : code;
}

__device__ __forceinline__ int popc3(int m) { return (m & 1) + ((m >> 1) & 1) + ((m >> 2) & 1); }

// ------------------------------------------------------------------------------------------
// Wave primitives

__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

// Set bits of a wave-wide mask below the calling lane (v_mbcnt_lo / v_mbcnt_hi: two VALU
// instructions, where masking with a per-lane 64-bit shift and counting takes six).
__device__ __forceinline__ int lanes_below(uint64_t mask) {
    return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// Inclusive prefix over the wave of a 32-bit value: a scan within each row of 16 lanes (DPP row
// shifts by 1, 2, 4, 8; lanes before the row start read 0), then the last lane of row 0 added to
// rows 1 and 3 and the last lane of row 1 (by then the sum of rows 0 and 1) to rows 2 and 3 (DPP
// row broadcasts). Six adds. Several small counts can be scanned at once, each in a bit field wide
// enough for its wave total. Must be reached by all 64 lanes.
__device__ __forceinline__ uint32_t wave_incl_dpp(uint32_t v) {
    int x = (int)v;
    x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, true);   // row_shr:1
    x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, true);   // row_shr:2
    x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, true);   // row_shr:4
    x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, true);   // row_shr:8
    x += __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false);  // row_bcast:15 to rows 1, 3
    x += __builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false);  // row_bcast:31 to rows 2, 3
    return (uint32_t)x;
}

__device__ __forceinline__ int64_t wave_sum64(int64_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

__device__ __forceinline__ uint32_t wave_or(uint32_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v |= (uint32_t)__shfl_xor((int)v, o, kWave);
    return v;
}

// Block index remap: the dispatcher deals consecutive workgroups round-robin to the 8 XCDs;
// this gives each XCD a contiguous run of logical blocks so that neighbouring rows (which a
// block re-reads as halo) share one XCD's L2.
__device__ __forceinline__ int64_t xcd_block(int64_t bid, int64_t nb) {
    int64_t q = nb / 8, r = nb % 8;
    int64_t xcd = bid % 8, k = bid / 8;
    return xcd * q + (xcd < r ? xcd : r) + k;
}

struct SegBase {
    int64_t v;  // first vertex (MC) / first quad (DMC) of the segment
    int64_t f;  // first triangle (MC) / first DMC vertex (patch) of the segment
};

// Segment (px, py, chunk) is (px * g1 + py) * nchunk + chunk, so the ids step by nchunk from row
// (px, py) to (px, py + 1) and by xstride(g) to (px + 1, py): neighbour rows' segments come from
// additions, not products.
__device__ __forceinline__ int64_t xstride(const Geo &g) { return (int64_t)g.g1 * g.nchunk; }

// Position classification word, written by the count pass for every lane of every segment
// (segment-major: word s * 64 + lane), read by the later passes instead of re-deriving it from
// the samples: bits 0-7 the code of the cell anchored at the position (MC: Bourke corner order;
// DMC: Morton corner order after the face-neighbour rule), bits 8-10 the crossing mask of the
// three edges the position owns (+x, +y, +z), bit 11 whether its own sample is >= iso, bits 12-14
// the cell's face count (MC: triangles, 0..5; DMC: patches, 0..4). Bits 0-10 and 12-14 are 0
// where no cell is anchored.
__device__ __forceinline__ int info_code(uint32_t w) { return (int)(w & 0xffu); }
__device__ __forceinline__ int info_mask(uint32_t w) { return (int)((w >> 8) & 7u); }
__device__ __forceinline__ int info_inside(uint32_t w) { return (int)((w >> 11) & 1u); }
__device__ __forceinline__ int info_count(uint32_t w) { return (int)((w >> 12) & 7u); }

__device__ __forceinline__ const uint16_t *info_of(const SegBase *seg, const Geo &g) {
    return reinterpret_cast<const uint16_t *>(seg + g.nseg + 2);
}

// A grid's segments as the emission and backward passes see them: its bases (seg[s], s from 0;
// seg[nseg] is the base after its last segment), its classification words, and the base values
// of its first segment (v0, f0), which are subtracted to index its own outputs. A grid extracted
// alone has v0 = f0 = 0; in a batch the scan runs over all grids' segments in order.
struct SegCtx {
    const SegBase *seg;
    const uint16_t *info;
    int64_t v0, f0;
};

// ------------------------------------------------------------------------------------------
// Register windows and edge geometry shared by the emission and backward kernels

// Samples of a lane's neighbourhood held in registers: v[i][j][k] is the sample at
// (px + X0 + i, py + Y0 + j, pz + Z0 + k). Only ever indexed with compile-time constants. Boxes
// of it are loaded on demand (load_box), so a kernel can test cheaply for work before loading
// the rest. When the whole window of the whole wave lies inside the stored grid (`interior`,
// wave-uniform: px, py and the chunk are), samples are read through a row pointer without
// per-sample range checks.
template <typename S, int X0, int NX, int Y0, int NY, int Z0, int NZ>
struct Win {
    S v[NX][NY][NZ];
    int px, py, pz;
    bool interior;
    // a window whose interior test the caller has made (a count loop keeps it by row)
    __device__ __forceinline__ Win(int px_, int py_, int pz_, bool interior_)
        : px(px_), py(py_), pz(pz_), interior(interior_) {}
    __device__ __forceinline__ Win(const In<S> &in, int px_, int py_, int pz_, int z0) : px(px_), py(py_), pz(pz_) {
        const Geo &g = in.geo;
        interior = px + X0 - g.off >= 0 && px + X0 + NX - 1 - g.off < g.n0 && py + Y0 - g.off >= 0 &&
                   py + Y0 + NY - 1 - g.off < g.n1 && z0 + Z0 - g.off >= 0 &&
                   z0 + kWave - 1 + Z0 + NZ - 1 - g.off < g.n2;
    }
    template <int BX0, int BNX, int BY0, int BNY, int BZ0, int BNZ>
    __device__ __forceinline__ void load_box(const In<S> &in) {
        static_assert(BX0 >= X0 && BX0 + BNX <= X0 + NX && BY0 >= Y0 && BY0 + BNY <= Y0 + NY && BZ0 >= Z0 &&
                          BZ0 + BNZ <= Z0 + NZ,
                      "box outside the window");
        const Geo &g = in.geo;
        const int64_t sy = g.n2, sx = (int64_t)g.n1 * g.n2;
        if (interior) {
            const S *p = in.sdf + ((int64_t)(px + BX0 - g.off) * g.n1 + (py + BY0 - g.off)) * g.n2 + (pz + BZ0 - g.off);
#pragma unroll
            for (int i = 0; i < BNX; ++i)
#pragma unroll
                for (int j = 0; j < BNY; ++j)
#pragma unroll
                    for (int k = 0; k < BNZ; ++k) v[BX0 - X0 + i][BY0 - Y0 + j][BZ0 - Z0 + k] = p[i * sx + j * sy + k];
        } else if (NX * NY > 4) {  // wide windows: per sample (fewer live scalar registers)
#pragma unroll
            for (int i = 0; i < BNX; ++i)
#pragma unroll
                for (int j = 0; j < BNY; ++j)
#pragma unroll
                    for (int k = 0; k < BNZ; ++k)
                        v[BX0 - X0 + i][BY0 - Y0 + j][BZ0 - Z0 + k] = sample(in, px + BX0 + i, py + BY0 + j, pz + BZ0 + k);
        } else {
            // rows outside the stored grid read as the pad value (a wave-uniform test); inside a
            // row, lanes whose z is outside load at a valid address and select the pad value
            const int x0 = px + BX0 - g.off, y0 = py + BY0 - g.off;
            const int64_t r0 = ((int64_t)x0 * g.n1 + y0) * g.n2;  // offset of row (x0, y0), maybe outside
#pragma unroll
            for (int i = 0; i < BNX; ++i)
#pragma unroll
                for (int j = 0; j < BNY; ++j) {
                    if ((unsigned)(x0 + i) < (unsigned)g.n0 && (unsigned)(y0 + j) < (unsigned)g.n1) {
                        const S *p = in.sdf + (r0 + i * sx + j * sy);
#pragma unroll
                        for (int k = 0; k < BNZ; ++k) {
                            const int z = pz + BZ0 + k - g.off;
                            const bool inz = (unsigned)z < (unsigned)g.n2;
                            const S a = p[inz ? z : 0];
                            v[BX0 - X0 + i][BY0 - Y0 + j][BZ0 - Z0 + k] = inz ? a : in.pad;
                        }
                    } else {
#pragma unroll
                        for (int k = 0; k < BNZ; ++k) v[BX0 - X0 + i][BY0 - Y0 + j][BZ0 - Z0 + k] = in.pad;
                    }
                }
        }
    }
    __device__ __forceinline__ S operator()(int dx, int dy, int dz) const { return v[dx - X0][dy - Y0][dz - Z0]; }
};

// Segment of a wave: its row (px, py) and chunk, from the wave-uniform segment id.
struct SegPos {
    int px, py, chunk;
};
__device__ __forceinline__ SegPos seg_pos(const Geo &g, int64_t s) {
    const uint32_t su = (uint32_t)s, nc = (uint32_t)g.nchunk, g1 = (uint32_t)g.g1;
    const uint32_t row = su / nc;
    return SegPos{(int)(row / g1), (int)(row % g1), (int)(su % nc)};
}

// The position of segment lo + i, read from lane i of a lane-parallel seg_pos (i uniform).
__device__ __forceinline__ SegPos pos_of_lane(const SegPos &lp, int i) {
    return SegPos{__builtin_amdgcn_readlane(lp.px, i), __builtin_amdgcn_readlane(lp.py, i),
                  __builtin_amdgcn_readlane(lp.chunk, i)};
}

// A wave-uniform struct read from memory, kept in scalar registers.
template <typename T>
__device__ __forceinline__ T load_uniform(const T *p) {
    static_assert(sizeof(T) % 4 == 0, "dword-sized");
    int w[sizeof(T) / 4];
    __builtin_memcpy(w, p, sizeof(T));
#pragma unroll
    for (int i = 0; i < (int)(sizeof(T) / 4); ++i) w[i] = __builtin_amdgcn_readfirstlane(w[i]);
    T r;
    __builtin_memcpy(&r, w, sizeof(T));
    return r;
}

// Wave index within the block, as a scalar (wave-uniform) value.
__device__ __forceinline__ int wave_id() { return __builtin_amdgcn_readfirstlane(threadIdx.x / kWave); }

// Edge e of a cell (src/cumc.cu:109-122): starts at cell-relative corner (eox, eoy, eoz) and
// runs along axis eax. Bit masks of the edge table, so unrolled loops fold them to constants.
__device__ constexpr int eox(int e) { return (0x622 >> e) & 1; }
__device__ constexpr int eoy(int e) { return (0x0F0 >> e) & 1; }
__device__ constexpr int eoz(int e) { return (0xC44 >> e) & 1; }
__device__ constexpr int eax(int e) { return e >= 8 ? 1 : ((e & 1) ? 2 : 0); }

// Crossing mask of the three edges a cell owns, from its sample and its +x, +y, +z neighbours.
template <typename S>
__device__ __forceinline__ int own_mask(S d0, S dx, S dy, S dz, S iso) {
    return crosses(d0, dx, iso) | (crosses(d0, dy, iso) << 1) | (crosses(d0, dz, iso) << 2);
}

__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <typename S>
__device__ __forceinline__ V3<S> add3(V3<S> a, V3<S> b) {
    return {a.x + b.x, a.y + b.y, a.z + b.z};
}

// computeMcVert from samples and deforms already in registers: p0 + (p1 - p0) * t with
// p = integer position + deform (src/cumc.cu:343-368).
template <typename S, bool DEF>
__device__ __forceinline__ V3<S> lerp_vertex(int x, int y, int z, int axis, S d0, S d1, V3<S> f0, V3<S> f1,
                                             S iso) {
    const S t = interp_t(d0, d1, iso);
    S p0x = S(x), p0y = S(y), p0z = S(z);
    S p1x = S(x + (axis == 0)), p1y = S(y + (axis == 1)), p1z = S(z + (axis == 2));
    if (DEF) {
        p0x += f0.x;
        p0y += f0.y;
        p0z += f0.z;
        p1x += f1.x;
        p1y += f1.y;
        p1z += f1.z;
    }
    V3<S> r;
    r.x = p0x + (p1x - p0x) * t;
    r.y = p0y + (p1y - p0y) * t;
    r.z = p0z + (p1z - p0z) * t;
    return r;
}

// Samples and deforms of both endpoints of the edge from (x,y,z) along `axis`, as sample() and
// deform_at() read them: the second endpoint's index is the first's plus the axis stride (one
// index product per edge instead of two).
template <typename S, bool DEF>
struct EdgeEnds {
    S d0, d1;
    V3<S> f0, f1;
};
template <typename S, bool DEF>
__device__ __forceinline__ EdgeEnds<S, DEF> edge_ends(const In<S> &in, int x, int y, int z, int axis) {
    const Geo &g = in.geo;
    const bool ok0 = in_data(in, x, y, z), ok1 = in_data(in, x + (axis == 0), y + (axis == 1), z + (axis == 2));
    const int64_t i0 = data_index(in, x, y, z);
    const int64_t i1 = i0 + (axis == 0 ? (int64_t)g.n1 * g.n2 : (axis == 1 ? (int64_t)g.n2 : 1));
    const int64_t j0 = ok0 ? i0 : 0, j1 = ok1 ? i1 : 0;
    EdgeEnds<S, DEF> e;
    const S a = in.sdf[j0], b = in.sdf[j1];
    e.d0 = ok0 ? a : in.pad;
    e.d1 = ok1 ? b : in.pad;
    e.f0 = e.f1 = V3<S>{S(0), S(0), S(0)};
    if (DEF) {
        const S *p0 = in.deform + 3 * j0, *p1 = in.deform + 3 * j1;
        const S u0 = p0[0], u1 = p0[1], u2 = p0[2], w0 = p1[0], w1 = p1[1], w2 = p1[2];
        e.f0 = V3<S>{ok0 ? u0 : S(0), ok0 ? u1 : S(0), ok0 ? u2 : S(0)};
        e.f1 = V3<S>{ok1 ? w0 : S(0), ok1 ? w1 : S(0), ok1 ? w2 : S(0)};
    }
    return e;
}

// MC vertex on the edge from (x,y,z) along `axis`: samples and deforms of both endpoints.
template <typename S, bool DEF>
__device__ __forceinline__ V3<S> edge_vertex(const In<S> &in, int x, int y, int z, int axis) {
    const EdgeEnds<S, DEF> e = edge_ends<S, DEF>(in, x, y, z, axis);
    return lerp_vertex<S, DEF>(x, y, z, axis, e.d0, e.d1, e.f0, e.f1, in.iso);
}

// x / c for a divisor c fixed per call, with r = RN(1/c): q = x*r is within an ulp of x/c and
// one FMA correction with the exact residual makes it the correctly rounded quotient
// (Markstein's theorem), i.e. bit-identical to IEEE x / c, in 3 instructions instead of the
// ~10 of a general correctly rounded division.
template <typename S>
__device__ __forceinline__ S div_by(S x, S c, S r) {
    if (c == S(0)) return x / c;  // uniform; keeps IEEE inf/nan for a length-1 axis
    const S q = x * r;
    return fma(fma(-q, c, x), r, q);
}

// Per-axis divisors (dims - 1) and their reciprocals, for the wrapper's normalisation.
template <typename S>
struct Norm {
    S c[3], r[3];
};

// ------------------------------------------------------------------------------------------
// Batch descriptors (every pass of a batch reads them)

template <typename S>
struct BatchGrid {
    In<S> in;
    Norm<S> nm;
    int algo, flags;
    int64_t seg_first;  // first segment in the batch's segment space
    int64_t nv, nf;     // vertex and face rows (from the count pass)
    S *verts;
    void *faces;
    uint8_t *vinfo;
    int64_t vcap, fcap;
    const S *adj;  // upstream vertex gradients [nv, 3]
    S *adj_sdf, *adj_deform;
    int64_t q_first;  // DMC: first row of its per-edge gradients in the batch scratch
    int run;          // segments per emission / backward block: the run length of the grid's size bucket
};

// j such that rf[j] <= b < rf[j + 1] (rf ascending, rf[0] = 0, rf[m] = total).
__device__ __forceinline__ int find_group(const int64_t *__restrict__ rf, int m, int64_t b) {
    int lo = 0, hi = m;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (rf[mid] <= b)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// ------------------------------------------------------------------------------------------
// Pass 1: classify, count, scan (single pass)
//
// The grid of segments is split into G contiguous ranges, one per block (G ~ 4 blocks per CU).
// A block counts its segments (writing each segment's counts into
// `seg`), publishes its aggregate as one self-contained 64-bit word, reads the aggregates of all
// its predecessors in parallel (a look-back over the whole window at once: no chain of inclusive
// prefixes, so the wait is one memory latency rather than one per predecessor), and finally scans
// its own segment counts into segment bases. Aggregate words are accessed with agent-scope
// atomics only, so they are coherent across the XCDs' L2s without fences. Two word buffers
// alternate between calls; each call clears the one the next call uses (up to the extent that
// was written), and the last block advances the epoch once every block has published (and so has
// read the epoch), so no initialisation launch is needed. Words are tagged with their call's
// epoch (diso_scan_word.h: tag, sample flags, face / patch count, vertex / quad count), so a word
// that a block of an earlier call wrote late never counts for a later call, and no reset is
// needed after a timeout. A predecessor that does not publish in time (it may not be running:
// other launches can fill the slots of the XCD it was dispatched to) is counted by the waiting
// block itself; only the last block waits for all, and every block it waits for ends in bounded
// time once dispatched.

struct CtlHeader {
    uint32_t epoch;     // calls completed on this block (diso_scan_word.h)
    uint32_t dirty[2];  // words written in each buffer since it was last cleared
    uint32_t pad[13];
};

constexpr int kCountWaves = 8;
constexpr int kCountBlock = kCountWaves * kWave;
constexpr int64_t kMaxCountBlocks = 8192;
constexpr int kCountBlocksPerCU = 4;

__device__ __forceinline__ uint64_t ld_relaxed(const uint64_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_relaxed(uint64_t *p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Per-position counts for pass 1: nv = crossing edges owned (MC vertices / DMC quads);
// nf = MC triangles / DMC patches; fl = sample flags of the position's own sample; the position's
// classification word.
template <typename S>
using CornerWin = Win<S, 0, 2, 0, 2, 0, 2>;

template <typename S>
__device__ __forceinline__ void count_position(const In<S> &in, int ALGO, const uint64_t *cases,
                                               const CornerWin<S> &w, int *nv, int *nf, uint32_t *fl,
                                               uint32_t *info) {
    const int px = w.px, py = w.py, pz = w.pz;
    const S v = w(0, 0, 0);
    *fl = in_data(in, px, py, pz) ? ((v < in.iso ? DISO_TOTAL_HAS_BELOW : 0u) |
                                     (v > in.iso ? DISO_TOTAL_HAS_ABOVE : 0u) | (v != v ? DISO_TOTAL_HAS_NAN : 0u))
                                  : 0u;
    *nv = 0;
    *nf = 0;
    *info = (uint32_t)(v >= in.iso) << 11;
    if (!is_anchor(in, px, py, pz)) return;
    const int m = own_mask(v, w(1, 0, 0), w(0, 1, 0), w(0, 0, 1), in.iso);
    *nv = popc3(m);
    int code;
    if (ALGO == DISO_MC) {
        Corners<S> k;
#pragma unroll
        for (int i = 0; i < 8; ++i) k.c[i & 1][(i >> 1) & 1][(i >> 2) & 1] = w(i & 1, (i >> 1) & 1, (i >> 2) & 1);
        code = mc_code(k, in.iso);
        *nf = DISO_MC_NTRI(cases[code]);
    } else {
        code = 0;
#pragma unroll
        for (int i = 0; i < 8; ++i) code |= (w(i & 1, (i >> 1) & 1, (i >> 2) & 1) >= in.iso) << i;
        if (code != 0 && code != 255)
            code = w.interior ? dmc_good_code_stored(in, cases, code, px, py, pz, data_index(in, px, py, pz))
                              : dmc_good_code(in, cases, code, px, py, pz);
        *nf = DISO_DMC_NPATCH(cases[code]);
    }
    *info |= (uint32_t)code | ((uint32_t)m << 8) | ((uint32_t)*nf << 12);
}

// Wave sums of two per-lane counts in [0, 7]: one DPP scan of both as 10-bit fields.
__device__ __forceinline__ void wave_totals2(int a, int b, int *ta, int *tb) {
    const uint32_t t = (uint32_t)__builtin_amdgcn_readlane((int)wave_incl_dpp((uint32_t)a | (uint32_t)b << 10), kWave - 1);
    *ta = (int)(t & 0x3ff);
    *tb = (int)((t >> 10) & 0x3ff);
}

__device__ __forceinline__ int64_t wave_incl64(int64_t v) {
    const int lane = lane_id();
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        int64_t u = __shfl_up(v, o, kWave);
        if (lane >= o) v += u;
    }
    return v;
}

// What a count launch classifies: one grid (`in`, algorithm ALGO), or (BATCH) the concatenated
// segments of several grids, each with its own descriptor; then per-grid sample flags are
// or-ed into gflags[grid] (zero on entry) and the scan runs over all segments in order.

template <typename S>
struct CountSrc {
    In<S> in;
    uint32_t *gflags;
    int ngrids;
    uint32_t recount_spins;  // reads of an unpublished predecessor word before counting its range
    int64_t nseg;            // segments of the launch
};

// Counts segments [wlo, whi) (one wave, in order; its position advances without divisions) and
// adds the wave's totals to (wv, wf, wfl). WRITE: also record each segment's counts (lane 0),
// every lane's classification word and (BATCH) the grids' sample flags; without WRITE the range
// is only summed (a block recounting a predecessor's range, see the look-back).
// Whether count_range_tails applies to a grid: rows of at least two chunks whose last chunk holds
// at most 32 positions (the z extent a few positions past a multiple of 64, as 256 + 2 pad
// samples is).
__device__ __forceinline__ bool short_tails(const Geo &g) {
    return g.nchunk >= 2 && g.g2 - kWave * (g.nchunk - 1) <= kWave / 2;
}

// A count loop's position in one grid: the segment's row and chunk, and the offset of the stored
// row (px - off, py - off) in samples, kept by additions as the loop advances segment by segment
// (instead of 64-bit products per segment; meaningful only where the row is stored).
// Whether a segment's corner window is all stored samples is kept the same way: rows (px..px+1,
// py..py+1) stored, updated when the row changes, and the chunk within [clo, chi], the chunks
// whose z extent z0 .. z0 + 64 is stored.
struct CountPos {
    SegPos sp;
    int64_t rowoff;
    bool rowin;
    int clo, chi;
    __device__ __forceinline__ bool interior() const { return rowin && sp.chunk >= clo && sp.chunk <= chi; }
};
__device__ __forceinline__ bool rows_stored(const Geo &g, int px, int py) {
    return px - g.off >= 0 && px + 1 - g.off < g.n0 && py - g.off >= 0 && py + 1 - g.off < g.n1;
}
__device__ __forceinline__ CountPos count_pos(const Geo &g, int64_t s) {
    const SegPos sp = seg_pos(g, s);
    // chunk c is inside when 64 c - off >= 0 and 64 c + 64 - off < n2
    const int clo = (g.off + kWave - 1) / kWave, hi = g.n2 + g.off - kWave - 1;
    return CountPos{sp, ((int64_t)(sp.px - g.off) * g.n1 + (sp.py - g.off)) * g.n2, rows_stored(g, sp.px, sp.py), clo,
                    hi >= 0 ? hi / kWave : -1};
}
__device__ __forceinline__ void count_advance(const Geo &g, CountPos &c) {
    if (++c.sp.chunk == g.nchunk) {
        c.sp.chunk = 0;
        c.rowoff += g.n2;
        if (++c.sp.py == g.g1) {
            c.sp.py = 0;
            ++c.sp.px;
            c.rowoff += (int64_t)(g.n1 - g.g1) * g.n2;
        }
        c.rowin = rows_stored(g, c.sp.px, c.sp.py);
    }
}

// The corner window of a wave's segment at `c` (pz = chunk * 64 + lane): a window of stored
// samples from the row offset, else load_box's general path.
template <typename S>
__device__ __forceinline__ void load_count_window(const In<S> &in, CornerWin<S> &w, const CountPos &c) {
    if (!w.interior) {  // wave-uniform
        w.template load_box<0, 2, 0, 2, 0, 2>(in);
        return;
    }
    const Geo &g = in.geo;
    const int64_t sy = g.n2, sx = (int64_t)g.n1 * g.n2;
    const S *p = in.sdf + c.rowoff + (w.pz - g.off);
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int k = 0; k < 2; ++k) w.v[i][j][k] = p[i * sx + j * sy + k];
}

// Counts the segments [wlo, whi) of one grid whose segment 0 is segment `first` of the launch
// (a short_tails grid): its full segments one per wave iteration, then its last-chunk segments
// ("tails") several per iteration, P lanes each (P = tail positions rounded up to a power of two),
// so that a wave iteration is not spent on a few positions. A tail's remaining words (positions
// past the grid) all hold the pad sample's inside bit. Same counts, words and flags as one segment
// per iteration; returns the grid's sample flags of the range.
template <typename S, bool WRITE>
__device__ __forceinline__ uint32_t count_range_tails(const In<S> &in, int algo, int64_t first, SegBase *seg,
                                                      uint16_t *info, const uint64_t *cases, int64_t wlo, int64_t whi,
                                                      int lane, int64_t &wv, int64_t &wf) {
    const Geo &g = in.geo;
    const int nc = g.nchunk, last = nc - 1, tailr = g.g2 - kWave * last;
    uint32_t gfl = 0;
    CountPos cp = count_pos(g, wlo - first);
    for (int64_t s = wlo; s < whi; ++s) {  // wave-uniform
        const SegPos sp = cp.sp;
        if (sp.chunk != last) {
            int nv, nf;
            uint32_t fl, iw;
            CornerWin<S> w(sp.px, sp.py, sp.chunk * kWave + lane, cp.interior());
            load_count_window(in, w, cp);
            count_position<S>(in, algo, cases, w, &nv, &nf, &fl, &iw);
            int tv, tf;
            wave_totals2(nv, nf, &tv, &tf);
            if constexpr (WRITE) {
                info[s * kWave + lane] = (uint16_t)iw;
                if (lane == 0) {
                    seg[s].v = tv;
                    seg[s].f = tf;
                }
            }
            wv += tv;
            wf += tf;
            gfl |= fl;
        }
        count_advance(g, cp);
    }
    // tails of the range: segments t0, t0 + nc, ... below whi (t0 - first = last mod nc)
    const int64_t t0 = wlo + ((last - (wlo - first) % nc) + nc) % nc;
    const int64_t nt = t0 < whi ? (whi - 1 - t0) / nc + 1 : 0;
    int P = 1;
    while (P < tailr) P <<= 1;
    const int per = kWave / P, t = lane / P, i = lane % P;
    const uint16_t beyond = (uint16_t)((uint32_t)(in.pad >= in.iso) << 11);  // word of a position past the grid
    for (int64_t b = 0; b < nt; b += per) {  // wave-uniform
        const bool valid = b + t < nt && i < tailr;
        const int64_t st = t0 + (b + t < nt ? b + t : 0) * nc;
        const SegPos tp = seg_pos(g, st - first);  // per lane
        int nv, nf;
        uint32_t fl, iw;
        CornerWin<S> w(in, tp.px, tp.py, last * kWave + i, last * kWave);
        w.template load_box<0, 2, 0, 2, 0, 2>(in);
        count_position<S>(in, algo, cases, w, &nv, &nf, &fl, &iw);
        if (!valid) nv = nf = 0, fl = 0;
        int sv = nv, sf = nf;  // per-tail sums over its P lanes
        for (int o = 1; o < P; o <<= 1) {
            sv += __shfl_xor(sv, o, kWave);
            sf += __shfl_xor(sf, o, kWave);
        }
        if constexpr (WRITE) {
            if (valid) info[st * kWave + i] = (uint16_t)iw;
            if (valid && i == 0) {
                seg[st].v = sv;
                seg[st].f = sf;
            }
            const int ng = nt - b < per ? (int)(nt - b) : per, cw = kWave - tailr;
            for (int k = lane; k < ng * cw; k += kWave)
                info[(t0 + (b + k / cw) * nc) * kWave + tailr + k % cw] = beyond;
        }
        int tv, tf;
        wave_totals2(nv, nf, &tv, &tf);
        wv += tv;
        wf += tf;
        gfl |= fl;
    }
    return gfl;
}

template <typename S, int ALGO, bool BATCH, bool WRITE>
__device__ __forceinline__ void count_range(const CountSrc<S> &src, const BatchGrid<S> *__restrict__ grids,
                                            SegBase *seg, uint16_t *info, const uint64_t (*s_case)[256],
                                            int64_t wlo, int64_t whi, int lane, int64_t &wv, int64_t &wf,
                                            uint32_t &wfl) {
    if constexpr (!BATCH) {
        if (short_tails(src.in.geo)) {
            wfl |= count_range_tails<S, WRITE>(src.in, ALGO, 0, seg, info, s_case[0], wlo, whi, lane, wv, wf);
            return;
        }
    }
    // BATCH: the range may span several grids; one pass of the outer loop per grid, with that
    // grid's parameters held in scalar registers for the inner loop.
    int gi = 0;
    if constexpr (BATCH)
        while (gi + 1 < src.ngrids && wlo >= grids[gi + 1].seg_first) ++gi;
    for (int64_t s = wlo; s < whi;) {  // wave-uniform
        In<S> in;
        if constexpr (BATCH)
            in = load_uniform(&grids[gi].in);
        else
            in = src.in;
        const int algo = BATCH ? grids[gi].algo : ALGO;
        const int64_t first = BATCH ? grids[gi].seg_first : 0;
        const int64_t end = (BATCH && gi + 1 < src.ngrids && grids[gi + 1].seg_first < whi) ? grids[gi + 1].seg_first
                                                                                          : whi;
        const uint64_t *cases = s_case[(BATCH && algo != DISO_MC) ? 1 : 0];
        uint32_t gfl = 0;  // the grid's sample flags
        if (BATCH && short_tails(in.geo)) {
            gfl = count_range_tails<S, WRITE>(in, algo, first, seg, info, cases, s, end, lane, wv, wf);
            s = end;
        }
        CountPos cp = count_pos(in.geo, s - first);
        for (; s < end; ++s) {
            int nv, nf;
            uint32_t fl, iw;
            const SegPos sp = cp.sp;
            CornerWin<S> w(sp.px, sp.py, sp.chunk * kWave + lane, cp.interior());
            load_count_window(in, w, cp);
            count_position<S>(in, algo, cases, w, &nv, &nf, &fl, &iw);
            count_advance(in.geo, cp);  // next segment of the grid
            int tv, tf;
            wave_totals2(nv, nf, &tv, &tf);
            if constexpr (WRITE) {
                info[s * kWave + lane] = (uint16_t)iw;
                if (lane == 0) {
                    seg[s].v = tv;
                    seg[s].f = tf;
                }
            }
            wv += tv;
            wf += tf;
            gfl |= fl;
        }
        wfl |= gfl;
        if constexpr (BATCH) {
            if constexpr (WRITE) {
                gfl = wave_or(gfl);
                if (lane == 0 && gfl) atomicOr(src.gflags + gi, gfl);
            }
            ++gi;
        }
    }
}

template <typename S, int ALGO, bool BATCH>
__global__ __launch_bounds__(kCountBlock, (BATCH ? 3 : kCountBlocksPerCU) * kCountWaves / 4) void count_kernel(
    CountSrc<S> src, const BatchGrid<S> *__restrict__ grids, SegBase *seg, uint16_t *info, uint8_t *ctl,
    int64_t nblocks) {
    __shared__ uint64_t s_case[BATCH ? 2 : 1][256];
    __shared__ int64_t s_v[kCountWaves], s_f[kCountWaves];
    __shared__ uint32_t s_fl[kCountWaves];
    __shared__ uint32_t s_epoch, s_dirty_other, s_next;
    __shared__ uint64_t s_word;

    CtlHeader *hdr = reinterpret_cast<CtlHeader *>(ctl);
    uint64_t *words = reinterpret_cast<uint64_t *>(ctl + kCtlHeader);
    const int tid = threadIdx.x, wave = wave_id(), lane = tid % kWave;
    const int64_t nseg = src.nseg;
    const int64_t b = blockIdx.x;
    const int64_t lo = b * nseg / nblocks, hi = (b + 1) * nseg / nblocks;

    for (int i = tid; i < 256; i += kCountBlock) {
        if (BATCH) {
            s_case[0][i] = kMcCase[i];
            s_case[BATCH ? 1 : 0][i] = kDmcCase[i];
        } else {
            s_case[0][i] = ALGO == DISO_MC ? kMcCase[i] : kDmcCase[i];
        }
    }
    if (tid == 0) {
        const uint32_t ep = __hip_atomic_load(&hdr->epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        s_epoch = ep;
        s_dirty_other = __hip_atomic_load(&hdr->dirty[diso_scan::buffer_of(ep) ^ 1u], __ATOMIC_RELAXED,
                                          __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    const uint32_t epoch = (uint32_t)__builtin_amdgcn_readfirstlane((int)s_epoch);
    const uint32_t parity = diso_scan::buffer_of(epoch), tag = diso_scan::tag_of(epoch);
    uint64_t *cur = words + (int64_t)parity * kMaxCountBlocks;

    // 1. Count: wave w takes the w-th eighth of the block's segments, in order; lane 0 records
    // each segment's counts, every lane its classification word.
    int64_t wv = 0, wf = 0;
    uint32_t wfl = 0;
    count_range<S, ALGO, BATCH, true>(src, grids, seg, info, s_case, lo + (hi - lo) * wave / kCountWaves,
                                      lo + (hi - lo) * (wave + 1) / kCountWaves, lane, wv, wf, wfl);
    wfl = wave_or(wfl);
    if (lane == 0) {
        s_v[wave] = wv;
        s_f[wave] = wf;
        s_fl[wave] = wfl;
    }
    __syncthreads();
    int64_t av = 0, af = 0;
    uint32_t afl = 0;
#pragma unroll
    for (int w = 0; w < kCountWaves; ++w) {
        av += s_v[w];
        af += s_f[w];
        afl |= s_fl[w];
    }
    if (tid == 0) st_relaxed(cur + b, diso_scan::pack(av, af, afl, tag));

    // 2. Exclusive prefix of this block = sum of all predecessors' aggregates. A predecessor that
    // has not published within `recount_spins` reads is not waited for any longer: this block
    // counts that range itself (read-only; the same sums), so no block waits on one that may not
    // be running (several launches sharing the device fill the XCDs' slots). The last block alone
    // waits for every predecessor's word: a block publishes after reading the epoch, so once all
    // have published the last block may advance it; every other block ends in bounded time, so the
    // blocks it waits for are all dispatched.
    const bool last = b == nblocks - 1;
    const uint32_t spin_limit = last ? kSpinLimit / (uint32_t)(kMaxCountBlocks / kCountBlock) : src.recount_spins;
    int64_t pv = 0, pf = 0;
    uint32_t pfl = 0;
    uint32_t miss = 0;  // bit k: predecessor tid + k * kCountBlock did not publish in time
    for (int64_t j = tid, k = 0; j < b; j += kCountBlock, ++k) {
        uint64_t w;
        uint32_t spins = 0;
        do {
            w = ld_relaxed(cur + j);
        } while (!diso_scan::published(w, tag) && ++spins < spin_limit);
        if (!diso_scan::published(w, tag)) {
            miss |= 1u << k;
        } else {
            pv += diso_scan::count_v(w);
            pf += diso_scan::count_f(w);
            pfl |= diso_scan::flags_of(w);
        }
    }
    pv = wave_sum64(pv);
    pf = wave_sum64(pf);
    pfl = wave_or(pfl);
    __syncthreads();  // s_v / s_f reuse
    if (lane == 0) {
        s_v[wave] = pv;
        s_f[wave] = pf;
        s_fl[wave] = pfl;
    }
    __syncthreads();
    int64_t ev = 0, ef = 0;
    uint32_t efl = 0;
#pragma unroll
    for (int w = 0; w < kCountWaves; ++w) {
        ev += s_v[w];
        ef += s_f[w];
        efl |= s_fl[w];
    }
    // the missing predecessors, lowest first, one at a time by the whole block
    for (;;) {
        if (tid == 0) s_next = 0xffffffffu;
        __syncthreads();
        if (miss) atomicMin(&s_next, (uint32_t)(tid + __builtin_ctz(miss) * kCountBlock));  // LDS
        __syncthreads();
        const uint32_t j = (uint32_t)__builtin_amdgcn_readfirstlane((int)s_next);
        if (j == 0xffffffffu) break;  // block-uniform
        if (last) efl |= DISO_TOTAL_SCAN_TIMEOUT;  // never expected: the epoch may advance early
        if (tid == (int)(j % kCountBlock)) {  // one read for the whole block: every wave takes the same branch
            miss &= ~(1u << (j / kCountBlock));
            s_word = ld_relaxed(cur + j);
        }
        __syncthreads();
        const uint64_t w = s_word;
        int64_t rv = 0, rf = 0;
        uint32_t rfl = 0;
        if (__builtin_amdgcn_readfirstlane((int)diso_scan::published(w, tag))) {  // published meanwhile
            rv = diso_scan::count_v(w);
            rf = diso_scan::count_f(w);
            rfl = diso_scan::flags_of(w);
            if (wave != 0) rv = rf = 0, rfl = 0;  // count it once
        } else {
            const int64_t jl = (int64_t)j * nseg / nblocks, jh = ((int64_t)j + 1) * nseg / nblocks;
            count_range<S, ALGO, BATCH, false>(src, grids, seg, info, s_case, jl + (jh - jl) * wave / kCountWaves,
                                               jl + (jh - jl) * (wave + 1) / kCountWaves, lane, rv, rf, rfl);
            rfl = wave_or(rfl);
        }
        __syncthreads();  // s_v / s_f reuse
        if (lane == 0) {  // rv, rf are wave-uniform sums
            s_v[wave] = rv;
            s_f[wave] = rf;
            s_fl[wave] = rfl;
        }
        __syncthreads();
#pragma unroll
        for (int w2 = 0; w2 < kCountWaves; ++w2) {
            ev += s_v[w2];
            ef += s_f[w2];
            efl |= s_fl[w2];
        }
    }
    if (last && tid == 0) {
        SegBase *tot = seg + nseg;
        tot[0].v = ev + av;
        tot[0].f = ef + af;
        tot[1].v = (int64_t)(efl | afl);
        tot[1].f = ALGO;
        __hip_atomic_store(&hdr->dirty[parity], (uint32_t)nblocks, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&hdr->dirty[parity ^ 1u], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&hdr->epoch, epoch + 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }

    // 3. Clear this block's share of the other word buffer for the next call.
    {
        const int64_t n = (int64_t)__builtin_amdgcn_readfirstlane((int)s_dirty_other);
        const int64_t per = (n + nblocks - 1) / nblocks;
        const int64_t c0 = b * per, c1 = c0 + per < n ? c0 + per : n;
        uint64_t *other = words + (int64_t)(parity ^ 1u) * kMaxCountBlocks;
        for (int64_t i = c0 + tid; i < c1; i += kCountBlock) st_relaxed(other + i, 0ull);
    }

    // 4. Segment bases: block-wide exclusive scan of this block's segment counts.
    __syncthreads();  // s_v / s_f reuse; phase-1 counts complete
    for (int64_t base = lo; base < hi; base += kCountBlock) {
        const int64_t s = base + tid;
        int64_t cv = 0, cf = 0;
        if (s < hi) {
            cv = (int64_t)__hip_atomic_load(reinterpret_cast<uint64_t *>(&seg[s].v), __ATOMIC_RELAXED,
                                            __HIP_MEMORY_SCOPE_AGENT);
            cf = (int64_t)__hip_atomic_load(reinterpret_cast<uint64_t *>(&seg[s].f), __ATOMIC_RELAXED,
                                            __HIP_MEMORY_SCOPE_AGENT);
        }
        const int64_t iv = wave_incl64(cv), jf = wave_incl64(cf);
        if (lane == kWave - 1) {
            s_v[wave] = iv;
            s_f[wave] = jf;
        }
        __syncthreads();
        int64_t ov = ev, of = ef, tv = 0, tf = 0;
#pragma unroll
        for (int w = 0; w < kCountWaves; ++w) {
            if (w < wave) {
                ov += s_v[w];
                of += s_f[w];
            }
            tv += s_v[w];
            tf += s_f[w];
        }
        if (s < hi) {
            seg[s].v = ov + iv - cv;
            seg[s].f = of + jf - cf;
        }
        ev += tv;
        ef += tf;
        __syncthreads();
    }
}

// ------------------------------------------------------------------------------------------
// Pass 2: emission
//
// Work compaction: each lane publishes its work items (MC vertices and triangle corners; DMC
// patch vertices and quads) into a dense per-wave list in LDS at its in-wave prefix, and the wave
// then processes the list 64 items at a time. Every item is computed exactly once (a lane-per-
// cell loop would make the whole wave execute the union of all lanes' branches), and item k is
// the segment's k-th output row, so consecutive lanes write consecutive rows (coalesced stores).

// Emission and backward blocks each take a run of `run` (<= 64) consecutive segments (so per-block
// setup, such as the case tables in LDS, is paid once per run) and share its segments with work
// among their waves (emission: from a shared counter, run_segments; backward: round-robin); the
// rest cost one lane-parallel load of the run's bases. The run length is a launch parameter
// (diso_set_param).
constexpr int kMaxRun = kWave;

// Bit i: segment lo + i has faces (MC triangles / DMC patches), i.e. seg[s + 1].f != seg[s].f
// (the entry after the last segment is the totals entry).
__device__ __forceinline__ uint64_t run_active(const SegBase *__restrict__ seg, int64_t lo, int64_t hi, int lane) {
    const int64_t s = lo + lane;
    return __ballot(s < hi && seg[s + 1].f != seg[s].f);
}

// The segments of a run with work (bits of `act`, segment lo + bit), taken by the block's waves
// from a shared counter as each becomes free (their costs differ, so a fixed round-robin deal
// leaves waves idle while the block, and its LDS, stays resident). A wave loads its next
// segment (load(bit)) before it processes the current one (work(loaded)). *s_take must be zero
// and visible to every wave of the block on entry.
template <typename Load, typename Work>
__device__ __forceinline__ void run_segments(uint64_t act, uint32_t *s_take, int lane, Load load, Work work) {
    const int nact = __popcll(act);
    auto grab = [&]() {
        uint32_t k = 0;
        if (lane == 0) k = atomicAdd(s_take, 1u);  // LDS
        return __builtin_amdgcn_readfirstlane((int)k);
    };
    auto kth = [&](int k) {  // bit of the k-th active segment
        const bool hit = ((act >> lane) & 1) && lanes_below(act) == k;
        return (int)__builtin_ctzll(__ballot(hit));
    };
    int k = grab();
    if (k >= nact) return;  // wave-uniform
    auto nxt = load(kth(k));
    for (;;) {
        const auto cur = nxt;
        k = grab();
        const bool more = k < nact;  // wave-uniform
        if (more) nxt = load(kth(k));
        work(cur);
        if (!more) break;
    }
}

struct OutCfg {
    int flags;
    int64_t vcap, fcap;  // capacities in rows
};

template <typename S>
__device__ __forceinline__ V3<S> epilogue(V3<S> p, int flags, const Norm<S> &nm) {
    if (flags & DISO_FLAG_SHIFT) {  // verts - 1 (diso/__init__.py:56)
        p.x = p.x - S(1);
        p.y = p.y - S(1);
        p.z = p.z - S(1);
    }
    if (flags & DISO_FLAG_NORMALIZE) {  // verts / (dims - 1) (diso/__init__.py:57-60)
        p.x = div_by(p.x, nm.c[0], nm.r[0]);
        p.y = div_by(p.y, nm.c[1], nm.r[1]);
        p.z = div_by(p.z, nm.c[2], nm.r[2]);
    }
    return p;
}

// Triangle row f = (a, b, c), as three int64 or int32 indices (the int32 row in one 12-byte store).
typedef int i32x3u __attribute__((ext_vector_type(3), aligned(4)));
__device__ __forceinline__ void store_tri(void *__restrict__ faces, int64_t f, int64_t a, int64_t b, int64_t c,
                                          int flags) {
    if (flags & DISO_FLAG_INDEX64) {
        int64_t *p = reinterpret_cast<int64_t *>(faces) + 3 * f;
        put(p + 0, a);
        put(p + 1, b);
        put(p + 2, c);
    } else {
        const i32x3u v = {(int)a, (int)b, (int)c};
        __builtin_nontemporal_store(v, reinterpret_cast<i32x3u *>(reinterpret_cast<int32_t *>(faces) + 3 * f));
    }
}

// Quad q = base + rel[0..3], as four int64 or int32 indices; `vec`: `faces` is 16-byte aligned,
// so the quad goes out in 16-byte stores.
typedef long long i64x2 __attribute__((ext_vector_type(2)));
typedef int i32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ void store_quad(void *__restrict__ faces, int64_t q, int64_t base, const int32_t *rel,
                                           int flags, bool vec) {
    if (flags & DISO_FLAG_INDEX64) {
        int64_t *f = reinterpret_cast<int64_t *>(faces) + 4 * q;
        if (vec) {
            put(reinterpret_cast<i64x2 *>(f), i64x2{base + rel[0], base + rel[1]});
            put(reinterpret_cast<i64x2 *>(f) + 1, i64x2{base + rel[2], base + rel[3]});
        } else {
#pragma unroll
            for (int j = 0; j < 4; ++j) put(f + j, (int64_t)(base + rel[j]));
        }
    } else {
        int32_t *f = reinterpret_cast<int32_t *>(faces) + 4 * q;
        if (vec) {
            put(reinterpret_cast<i32x4 *>(f), i32x4{(int32_t)(base + rel[0]), (int32_t)(base + rel[1]),
                                                  (int32_t)(base + rel[2]), (int32_t)(base + rel[3])});
        } else {
#pragma unroll
            for (int j = 0; j < 4; ++j) put(f + j, (int32_t)(base + rel[j]));
        }
    }
}

__device__ __forceinline__ int64_t load_index(const void *__restrict__ faces, int64_t i, int flags) {
    return (flags & DISO_FLAG_INDEX64) ? reinterpret_cast<const int64_t *>(faces)[i]
                                       : (int64_t) reinterpret_cast<const int32_t *>(faces)[i];
}

template <typename S>
__device__ __forceinline__ void store_vertex(S *__restrict__ verts, int64_t i, int64_t cap, V3<S> p) {
    if (i < cap) {
        put(verts + 3 * i + 0, p.x);
        put(verts + 3 * i + 1, p.y);
        put(verts + 3 * i + 2, p.z);
    }
}

// The vertex ids of a cell's 12 edges, relative to its segment's first MC vertex. A corner on
// table edge e lies on an edge owned by the cell at (cell + edge offset); its id is that owner's
// first vertex plus the rank of the edge's axis among the owner's crossing axes (what
// create_cell_mc_tris_kernel finds by linear search, src/cumc.cu:586-608). Owners are the cell,
// its +x, +y, +x+y neighbours (rows R00, R10, R01, R11 of the chunk; rNN = first vertex, nNN =
// count, mNN = crossing mask) and their +z neighbours, whose first vertex is the current one's
// first vertex plus its count (cNNn = x-crossing flag of the +z neighbour).
__device__ __forceinline__ void edge_ids(int32_t *c, int32_t r00, int32_t r10, int32_t r01, int32_t r11, int m00,
                                         int m10, int m01, int m11, int c00n, int c10n) {
    const int n00 = popc3(m00), n10 = popc3(m10), n01 = popc3(m01), n11 = popc3(m11);
    c[0] = r00;
    c[1] = r10 + n10 - 1;
    c[2] = r00 + n00;
    c[3] = r00 + n00 - 1;
    c[4] = r01;
    c[5] = r11 + n11 - 1;
    c[6] = r01 + n01;
    c[7] = r01 + n01 - 1;
    c[8] = r00 + (m00 & 1);
    c[9] = r10 + (m10 & 1);
    c[10] = r10 + n10 + c10n;
    c[11] = r00 + n00 + c00n;
}

// MC emission. Vertex item = (lane, axis) of an owned crossing edge, in the reference's order
// (cell, then axis x < y < z; src/cumc.cu:370-410); triangle item = (lane, triangle of its cell),
// whose three corners' table edges come from the cell's case word.
// Cell codes and crossing masks come from the count pass's classification words.
struct McEmitLds {
    int32_t cand[kWave][12];  // vertex id of each table edge, relative to the segment's first vertex
    uint64_t cw[kWave];       // case word of each lane's cell
    uint16_t tri[kWave * 5 + 1];  // triangle items: lane << 3 | triangle of the lane's cell
    uint8_t vert[kWave * 3 + 1];  // vertex items: lane * 4 + axis
    // (the last entry of each list takes the writes of the lanes with nothing to list there)
};

// What an MC segment reads from memory before any work (loaded one segment ahead): the
// classification words of rows 00, 10, 01, 11 at z and of rows 00, 10 at z + 1, and the bases.
struct McSegIn {
    int64_t s;
    SegPos sp;
    uint32_t w00, w10, w01, w11, n00, n10;
    int64_t vabs, fabs;
    int32_t d10, d01, d11;  // first vertices of rows +x, +y, +x+y relative to the segment's
};

template <typename S>
__device__ __forceinline__ McSegIn mc_load(const In<S> &in, const SegCtx &sc, int64_t s, SegPos sp, int lane) {
    const Geo &g = in.geo;
    const SegBase *__restrict__ seg = sc.seg;
    const uint16_t *info = sc.info;
    McSegIn d;
    d.s = s;
    d.sp = sp;
    // This row has cells, so rows +x, +y, +x+y exist; +z words are read within the row only.
    const int pz = sp.chunk * kWave + lane;
    const int64_t s10 = s + xstride(g), s01 = s + g.nchunk, s11 = s10 + g.nchunk;
    d.vabs = seg[s].v;
    d.fabs = seg[s].f;
    d.d10 = (int32_t)(seg[s10].v - d.vabs);
    d.d01 = (int32_t)(seg[s01].v - d.vabs);
    d.d11 = (int32_t)(seg[s11].v - d.vabs);
    d.w00 = info[s * kWave + lane];
    d.w10 = info[s10 * kWave + lane];
    d.w01 = info[s01 * kWave + lane];
    d.w11 = info[s11 * kWave + lane];
    const bool zn = pz + 1 < g.g2;
    d.n00 = zn ? info[s * kWave + lane + 1] : 0u;
    d.n10 = zn ? info[s10 * kWave + lane + 1] : 0u;
    return d;
}

// One segment that has triangles.
template <typename S, bool DEF>
__device__ __forceinline__ void emit_mc_segment(const In<S> &in, const SegCtx &sc, const McSegIn &d,
                                                S *__restrict__ verts, void *__restrict__ faces, const OutCfg &out,
                                                const Norm<S> &nm, const uint64_t *s_case, McEmitLds &L,
                                                int lane) {
    wave_lds_sync();  // the previous segment's LDS reads are done
    const int64_t tbase = d.fabs - sc.f0;
    const int px = d.sp.px, py = d.sp.py;
    const int z0 = d.sp.chunk * kWave;
    const int64_t vabs = d.vabs, vbase = vabs - sc.v0;
    const uint32_t w00 = d.w00;
    const int m00 = info_mask(w00), m10 = info_mask(d.w10), m01 = info_mask(d.w01), m11 = info_mask(d.w11);
    const int c00n = info_mask(d.n00) & 1, c10n = info_mask(d.n10) & 1;
    const uint64_t cw = s_case[info_code(w00)];  // Bourke corner order (src/cumc.cu:107)
    const int ntri = DISO_MC_NTRI(cw);

    // exclusive prefixes over the lanes of the rows' vertex counts and of the triangle counts, in
    // two packed scans of 10-bit fields (wave totals <= 320)
    const uint32_t cA = (uint32_t)popc3(m00) | (uint32_t)popc3(m10) << 10 | (uint32_t)popc3(m01) << 20;
    const uint32_t cB = (uint32_t)ntri | (uint32_t)popc3(m11) << 10;
    const uint32_t iA = wave_incl_dpp(cA), iB = wave_incl_dpp(cB);
    const uint32_t eA = iA - cA, eB = iB - cB;
    const uint32_t tA = (uint32_t)__builtin_amdgcn_readlane((int)iA, kWave - 1),
                   tB = (uint32_t)__builtin_amdgcn_readlane((int)iB, kWave - 1);
    const int p00 = (int)(eA & 0x3ff), tv = (int)(tA & 0x3ff);
    const int pt = (int)(eB & 0x3ff), tt = (int)(tB & 0x3ff);
    const int32_t r10 = d.d10 + (int32_t)((eA >> 10) & 0x3ff);
    const int32_t r01 = d.d01 + (int32_t)((eA >> 20) & 0x3ff);
    const int32_t r11 = d.d11 + (int32_t)((eB >> 10) & 0x3ff);

    // list writes without per-lane branches: what is not listed goes to the spare last entry
    {
        int r = p00;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const int on = (m00 >> a) & 1;
            L.vert[on ? r : 3 * kWave] = (uint8_t)(lane * 4 + a);
            r += on;
        }
    }
    edge_ids(L.cand[lane], p00, r10, r01, r11, m00, m10, m01, m11, c00n, c10n);
    L.cw[lane] = cw;
#pragma unroll
    for (int t = 0; t < 5; ++t) L.tri[t < ntri ? pt + t : 5 * kWave] = (uint16_t)((lane << 3) | t);
    wave_lds_sync();

    for (int k = lane; k < tv; k += kWave) {
        const int it = L.vert[k], src = it >> 2, a = it & 3;
        const V3<S> p = edge_vertex<S, DEF>(in, px, py, z0 + src, a);
        store_vertex(verts, vbase + k, out.vcap, epilogue(p, out.flags, nm));
    }
    // triangle k of the segment per lane: its cell's case word once, the three corners' table
    // edges and vertex ids (independent LDS reads), then the three indices as one row
    for (int k = lane; k < tt; k += kWave) {
        const int64_t f = tbase + k;
        if (f >= out.fcap) continue;
        const int it = L.tri[k], src = it >> 3, t = it & 7;
        const uint64_t cw = L.cw[src];
        const int32_t *cand = L.cand[src];
        store_tri(faces, f, vbase + cand[DISO_MC_EDGE(cw, 3 * t)], vbase + cand[DISO_MC_EDGE(cw, 3 * t + 1)],
                  vbase + cand[DISO_MC_EDGE(cw, 3 * t + 2)], out.flags);
    }
}

template <typename S, bool DEF>
__global__ __launch_bounds__(kBlock, 7) void emit_mc_kernel(In<S> in, const SegBase *__restrict__ seg,
                                                         S *__restrict__ verts, void *__restrict__ faces,
                                                         OutCfg out, Norm<S> nm, int run, int64_t nblocks) {
    __shared__ uint64_t s_case[256];
    __shared__ McEmitLds s_lds[kWavesPerBlock];
    __shared__ uint32_t s_take;
    const int tid = threadIdx.x, wave = wave_id(), lane = tid % kWave;
    for (int i = tid; i < 256; i += kBlock) s_case[i] = kMcCase[i];
    if (tid == 0) s_take = 0;
    __syncthreads();
    const Geo &g = in.geo;
    const SegCtx sc{seg, info_of(seg, g), 0, 0};
    const SegBase *tot = seg + g.nseg;
    if (tot[0].v > out.vcap || tot[0].f > out.fcap) return;
    const int64_t lo = xcd_block(blockIdx.x, nblocks) * run, hi = lo + run < g.nseg ? lo + run : g.nseg;
    // segments of the run with triangles (hence owned crossings)
    const uint64_t act = run_active(seg, lo, hi, lane);
    const SegPos lp = seg_pos(g, lo + lane);
    run_segments(
        act, &s_take, lane, [&](int i) { return mc_load<S>(in, sc, lo + i, pos_of_lane(lp, i), lane); },
        [&](const McSegIn &d) { emit_mc_segment<S, DEF>(in, sc, d, verts, faces, out, nm, s_case, s_lds[wave], lane); });
}

// DMC emission. Patch-vertex item = (lane, patch): the mean of the MC vertices on the patch's
// edges, summed from zero in ascending edge order and scaled by 1/num (src/cudualmc.cu:907-955);
// each MC vertex is evaluated with the same expression as the MC pass, so it has the same bits.
// Quad item = (lane, axis, leaving) of an owned crossing edge (src/cudualmc.cu:1027-1056): its
// corners are the vertices of the patches holding the edge in the four cells around it, which
// lie in rows R00, Rm0 (-x), R0m (-y), Rmm (-x-y) at this z and z-1. Cell codes come from the
// count pass's classification words, so no sample is read to classify.

// Quad stencil (src/cudualmc.cu:504-513) as cell index ci and edge id: entry k of type ty.
// Types 3..5 (edges leaving the inside) list the cells of types 0..2 in the order 0,3,2,1.
__device__ constexpr int stencil_ci(int ty, int k) {
    constexpr int q[3][4] = {{0, 2, 3, 1}, {0, 1, 5, 4}, {0, 4, 6, 2}};
    return ty < 3 ? q[ty][k] : q[ty - 3][(4 - k) & 3];
}
__device__ constexpr int stencil_eid(int ty, int k) {
    constexpr int q[3][4] = {{0, 4, 6, 2}, {8, 11, 10, 9}, {3, 1, 5, 7}};
    return ty < 3 ? q[ty][k] : q[ty - 3][(4 - k) & 3];
}

// The edge and quad tables as the kernels evaluate them (eox..eax, stencil_ci / stencil_eid) agree
// with the tables' entries (diso_tables.h: edge offsets and axes, quad cell offsets and edge ids).
__device__ constexpr bool tables_agree() {
    const int ofs[12][3] = {DISO_EDGE_OFS_VALUES};
    const int axis[12] = {DISO_EDGE_AXIS_VALUES};
    const int quad[6][4][4] = {DISO_QUAD_VALUES};
    for (int e = 0; e < 12; ++e)
        if (eox(e) != ofs[e][0] || eoy(e) != ofs[e][1] || eoz(e) != ofs[e][2] || eax(e) != axis[e]) return false;
    for (int ty = 0; ty < 6; ++ty)
        for (int k = 0; k < 4; ++k) {
            const int *c = quad[ty][k];
            if (stencil_ci(ty, k) != -4 * c[0] - 2 * c[1] - c[2] || stencil_eid(ty, k) != c[3]) return false;
        }
    return true;
}
static_assert(tables_agree(), "edge / quad tables");

// Edge kinds of a segment's cells, by owner row (dx, dy) and axis: 0: row 00 x, 1: row 00 y,
// 2: row 00 z, 3: row 10 y, 4: row 10 z, 5: row 01 x, 6: row 01 z, 7: row 11 z (the owner of an
// edge is the position it starts from; row dxy is (px + dx, py + dy)). The x and y edges of the
// last cell's top face start at z0 + 64 (the "halo", kinds 8..11 for 0, 1, 3, 5 there).
// Cell edge e (table numbering) as (kind | top << 3) in 5 bits: top = the edge starts at z + 1.
__device__ constexpr int edge_kind(int e) {
    return eax(e) == 0 ? (eoy(e) ? 5 : 0) : eax(e) == 1 ? (eox(e) ? 3 : 1) : (eox(e) + 2 * eoy(e) == 0 ? 2 : eox(e) + 2 * eoy(e) == 1 ? 4 : eox(e) + 2 * eoy(e) == 2 ? 6 : 7);
}
__device__ constexpr uint64_t edge_kind_table() {
    uint64_t t = 0;
    for (int e = 0; e < 12; ++e) t |= (uint64_t)(edge_kind(e) | (eoz(e) << 3)) << (5 * e);
    return t;
}
// A cell's top edges (dz = 1) are all of kinds 0, 1, 3 or 5.
__device__ constexpr bool top_kinds_ok() {
    for (int e = 0; e < 12; ++e)
        if (eoz(e) && !((1 << edge_kind(e)) & 0x2b)) return false;
    return true;
}
static_assert(top_kinds_ok(), "top edges outside the halo slot entries");
constexpr int kEdgeSlots = 255;  // MC vertices kept per segment; beyond, patches recompute theirs

template <typename S>
struct DmcEmitLds {
    uint8_t code[8][kWave];    // codes of the cells (-i,-j,-k) around the lane, ci = 4i + 2j + k
    int16_t pre[8][kWave];     // their first DMC vertex, relative to the first of row (-i,-j)
    int32_t rowb[4];           // first DMC vertex of row (-i,-j), relative to the segment's first
    uint8_t patch[kWave * 4 + 1];  // patch items: lane * 4 + patch
    uint16_t quad[kWave * 3 + 1];  // quad items: lane * 8 + axis * 2 + leaving
    // (the last entry of each list takes the writes of the lanes with nothing to list there)
    // slot of the lane's edge of each kind (0xff: does not cross); entry kWave of kinds 0, 1, 3
    // and 5 holds the halo position z0 + 64's (the only kinds a cell's top edges have)
    uint8_t slot[8][kWave + 1];
    uint16_t eitem[kEdgeSlots + 1];  // edge items: lane << 4 | kind; the last takes non-crossing edges' writes
    S ev[3][kEdgeSlots + 1];         // their MC vertices; the last entry holds zeros
};

// Edge masks of the patches of every DMC case, 16 bits per patch (patch k in bits 16k..16k+11),
// derived at compile time from the case table.
struct PatchEdgeTable {
    uint64_t m[256];
};
constexpr PatchEdgeTable make_patch_edges() {
    const uint64_t cases[256] = {DISO_DMC_CASE_VALUES};
    PatchEdgeTable t{};
    for (int c = 0; c < 256; ++c) {
        uint64_t w = 0;
        for (int e = 0; e < 12; ++e) {
            const int k = DISO_DMC_PATCH_OF(cases[c], e);
            if (k < 4) w |= 1ull << (16 * k + e);
        }
        t.m[c] = w;
    }
    return t;
}
__constant__ PatchEdgeTable kDmcPatchEdges = make_patch_edges();

// Quad stencil of each edge type as one word: cell index ci of corner j in bits 3j..3j+2, edge id
// in bits 12+4j..15+4j.
struct StencilTable {
    uint32_t w[6];
};
constexpr StencilTable make_stencil() {
    StencilTable t{};
    for (int ty = 0; ty < 6; ++ty)
        for (int j = 0; j < 4; ++j)
            t.w[ty] |= ((uint32_t)stencil_ci(ty, j) << (3 * j)) | ((uint32_t)stencil_eid(ty, j) << (12 + 4 * j));
    return t;
}
__constant__ StencilTable kStencil = make_stencil();

// 1/n for the possible patch sizes (n = 1..7), correctly rounded at compile time.
template <typename S>
struct RcpTable {
    S r[8];
};
template <typename S>
constexpr RcpTable<S> make_rcp() {
    RcpTable<S> t{};
    for (int n = 1; n < 8; ++n) t.r[n] = S(1) / S(n);
    return t;
}
__constant__ RcpTable<float> kRcpF = make_rcp<float>();
__constant__ RcpTable<double> kRcpD = make_rcp<double>();
template <typename S>
__device__ __forceinline__ S rcp_small(int n);
template <>
__device__ __forceinline__ float rcp_small<float>(int n) { return kRcpF.r[n]; }
template <>
__device__ __forceinline__ double rcp_small<double>(int n) { return kRcpD.r[n]; }

// What a DMC segment reads from memory before any work: its classification words (own row,
// rows -x, -y, -x-y at z and z - 1, rows +x, +y, +x+y at z, and the halo position z0 + 64) and the
// bases. Loaded one segment ahead, so the loads of the next segment overlap this one's work.
struct DmcSegIn {
    int64_t s;
    SegPos sp;
    uint32_t w0, w10, w01, w11, h00, h10, h01;
    uint32_t cur[4], prev[4];  // codes of cells (-i,-j,0) and (-i,-j,-1), index 2i + j
    int64_t vabs, pabs;
    int32_t rowd[4];  // first DMC vertex of row (-i, -j), relative to the segment's
};

template <typename S>
__device__ __forceinline__ DmcSegIn dmc_load(const In<S> &in, const SegCtx &sc, int64_t s, SegPos sp, int lane) {
    const Geo &g = in.geo;
    const SegBase *__restrict__ seg = sc.seg;
    const uint16_t *info = sc.info;
    DmcSegIn d;
    d.s = s;
    d.sp = sp;
    const int px = sp.px, py = sp.py, chunk = sp.chunk, pz = chunk * kWave + lane;
    d.vabs = seg[s].v;
    d.pabs = seg[s].f;
    const int64_t s10 = s + xstride(g), s01 = s + g.nchunk, s11 = s10 + g.nchunk;
    d.w0 = info[s * kWave + lane];
    d.w10 = info[s10 * kWave + lane];
    d.w01 = info[s01 * kWave + lane];
    d.w11 = info[s11 * kWave + lane];
    d.h00 = d.h10 = d.h01 = 0;
    if (lane == kWave - 1 && pz + 1 < g.g2) {  // position z0 + 64, in the row's next segment
        d.h00 = info[s * kWave + lane + 1];
        d.h10 = info[s10 * kWave + lane + 1];
        d.h01 = info[s01 * kWave + lane + 1];
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const bool row = px - i >= 0 && py - j >= 0;  // wave-uniform
            const int64_t r = row ? s - i * xstride(g) - j * g.nchunk : s;
            d.rowd[2 * i + j] = row ? (int32_t)(seg[r].f - d.pabs) : 0;
            d.cur[2 * i + j] = (i || j) ? (row ? info[r * kWave + lane] : 0u) : d.w0;
            d.prev[2 * i + j] = (row && !(i && j) && pz >= 1) ? info[r * kWave + lane - 1] : 0u;
        }
    return d;
}

// One segment that has patches (hence cells, so rows +x, +y exist).
template <typename S, bool DEF>
__device__ __forceinline__ void emit_dmc_segment(const In<S> &in, const SegCtx &sc, const DmcSegIn &d,
                                                 S *__restrict__ verts, void *__restrict__ faces,
                                                 uint8_t *__restrict__ vinfo, const OutCfg &out, const Norm<S> &nm,
                                                 DmcEmitLds<S> &L, int lane) {
    wave_lds_sync();  // the previous segment's LDS reads are done
    const int64_t pabs = d.pabs, pbase = pabs - sc.f0;
    const int px = d.sp.px, py = d.sp.py;
    const int z0 = d.sp.chunk * kWave;
    const int64_t qbase = d.vabs - sc.v0;
    const uint32_t w0 = d.w0;
    // Crossing masks of the rows the cells' edges start from: 10, 01, 11 (this z) and the halo.
    const int m00 = info_mask(w0), m10 = info_mask(d.w10), m01 = info_mask(d.w01), m11 = info_mask(d.w11);
    const int h00 = info_mask(d.h00), h10 = info_mask(d.h10), h01 = info_mask(d.h01);
    // Exclusive prefixes over the lanes, in two packed scans of 10-bit fields (wave totals <= 448):
    // the patch counts of rows (0,0) (the segment's own; also the first vertices of row (0,0)'s
    // cells below), (0,-1), (-1,0) and (-1,-1) at this z, and the owned crossing edges (quads).
    const int np = info_count(w0);
    const uint32_t cA = (uint32_t)np | (uint32_t)info_count(d.cur[1]) << 10 | (uint32_t)info_count(d.cur[2]) << 20;
    const uint32_t cB = (uint32_t)info_count(d.cur[3]) | (uint32_t)popc3(m00) << 10;
    const uint32_t iA = wave_incl_dpp(cA), iB = wave_incl_dpp(cB);
    const uint32_t eA = iA - cA, eB = iB - cB;
    const uint32_t tA = (uint32_t)__builtin_amdgcn_readlane((int)iA, kWave - 1),
                   tB = (uint32_t)__builtin_amdgcn_readlane((int)iB, kWave - 1);
    const int pp = (int)(eA & 0x3ff), tp = (int)(tA & 0x3ff);
    const int pq = (int)((eB >> 10) & 0x3ff), tq = (int)((tB >> 10) & 0x3ff);
    const int rowpre[4] = {pp, (int)((eA >> 10) & 0x3ff), (int)((eA >> 20) & 0x3ff), (int)(eB & 0x3ff)};
    // Codes and first vertices of the cells (-i, -j, 0) and (-i, -j, -1).
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const uint32_t cur = d.cur[2 * i + j], prev = d.prev[2 * i + j];
            const int b = rowpre[2 * i + j];
            if (lane == 0) L.rowb[2 * i + j] = d.rowd[2 * i + j];
            L.code[4 * i + 2 * j][lane] = (uint8_t)info_code(cur);
            L.pre[4 * i + 2 * j][lane] = (int16_t)b;
            L.code[4 * i + 2 * j + 1][lane] = (uint8_t)info_code(prev);
            L.pre[4 * i + 2 * j + 1][lane] = (int16_t)(b - info_count(prev));
        }
    // list writes without per-lane branches: what is not listed goes to the spare last entry
#pragma unroll
    for (int k = 0; k < 4; ++k) L.patch[k < np ? pp + k : 4 * kWave] = (uint8_t)(lane * 4 + k);
    {
        // an owned crossing edge leaves the inside when the position's sample is >= iso
        // (type a + 3, src/cudualmc.cu:782-788)
        const int leaving = info_inside(w0);
        int q = pq;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const int on = (m00 >> a) & 1;
            L.quad[on ? q : 3 * kWave] = (uint16_t)(lane * 8 + a * 2 + leaving);
            q += on;
        }
    }
    // Edge slots: kinds 0..7 at every lane, then the halo kinds 8..11 (lane 63 only).
    const int kbits = (m00 & 7) | (((m10 >> 1) & 3) << 3) | ((m01 & 1) << 5) | (((m01 >> 2) & 1) << 6) |
                      (((m11 >> 2) & 1) << 7);
    const int hbits = (h00 & 3) | (((h10 >> 1) & 1) << 2) | ((h01 & 1) << 3);
    // Slots are numbered kind by kind, lanes in order within a kind: the lane's exclusive prefix
    // per kind from two DPP scans of the kind bits spread into 8-bit fields (x * 0x204081 moves bit
    // i of a 4-bit x to bit 8i), the kinds' wave totals from the scans' last lane; the halo kinds
    // (lane 63 only) follow all of them.
    const uint32_t c0 = ((uint32_t)(kbits & 15) * 0x204081u) & 0x01010101u;
    const uint32_t c1 = ((uint32_t)(kbits >> 4) * 0x204081u) & 0x01010101u;
    const uint32_t i0 = wave_incl_dpp(c0), i1 = wave_incl_dpp(c1);
    const uint32_t e0 = i0 - c0, e1 = i1 - c1;
    const uint32_t t0 = (uint32_t)__builtin_amdgcn_readlane((int)i0, kWave - 1),
                   t1 = (uint32_t)__builtin_amdgcn_readlane((int)i1, kWave - 1);
    const int hb = __builtin_amdgcn_readlane(hbits, kWave - 1);
    int nslot = 0, myslot[12];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int sh = 8 * (k & 3);
        const int ex = (int)(((k < 4 ? e0 : e1) >> sh) & 0xff);
        myslot[k] = ((kbits >> k) & 1) ? nslot + ex : -1;
        L.slot[k][lane] = (uint8_t)myslot[k];
        nslot += (int)(((k < 4 ? t0 : t1) >> sh) & 0xff);
    }
#pragma unroll
    for (int h = 0; h < 4; ++h) {
        // hb is wave-uniform, so every lane writes the same halo entry (kinds 0, 1, 3, 5)
        const int hs = ((hb >> h) & 1) ? nslot + __popc(hb & ((1 << h) - 1)) : -1;
        myslot[8 + h] = ((hbits >> h) & 1) ? hs : -1;
        L.slot[h < 2 ? h : 2 * h - 1][kWave] = (uint8_t)hs;
    }
    nslot += __popc(hb);
    const bool slots = nslot <= kEdgeSlots;  // wave-uniform
    if (slots) {
        // slots are < kEdgeSlots here, so the edges that do not cross (-1 & 0xff) all write the
        // spare last item, and no lane needs its own branch
#pragma unroll
        for (int k = 0; k < 12; ++k) L.eitem[myslot[k] & 0xff] = (uint16_t)((lane << 4) | k);
    }
    wave_lds_sync();

    if (slots) {
        // MC vertex of every crossing edge of the cells, once
        for (int i = lane; i < nslot; i += kWave) {
            const int it = L.eitem[i], l = it >> 4, k0 = it & 15;
            const int k = k0 < 8 ? k0 : (k0 == 8 ? 0 : (k0 == 9 ? 1 : (k0 == 10 ? 3 : 5)));
            // kinds 3, 4, 7 start at dx = 1; 5, 6, 7 at dy = 1; 2, 4, 6, 7 run along z, 1 and 3 along y
            const int dx = (0x98 >> k) & 1, dy = (0xE0 >> k) & 1, a = (0xD4 >> k) & 1 ? 2 : (0x0A >> k) & 1;
            const V3<S> p = edge_vertex<S, DEF>(in, px + dx, py + dy, z0 + (k0 < 8 ? l : kWave), a);
            L.ev[0][i] = p.x;
            L.ev[1][i] = p.y;
            L.ev[2][i] = p.z;
        }
        wave_lds_sync();
    }

    constexpr uint64_t kKinds = edge_kind_table();
    for (int k = lane; k < tp; k += kWave) {
        const int it = L.patch[k], src = it >> 2, pk = it & 3;
        unsigned em = (unsigned)(kDmcPatchEdges.m[L.code[0][src]] >> (16 * pk)) & 0xfffu;
        const int num = __popc(em);
        S sx = S(0), sy = S(0), sz = S(0);
        if (slots) {
            // the slots of all twelve edges first (independent reads), then all twelve vertices
            // added in ascending edge order, edges outside the patch (and, on NaN samples only,
            // patch edges that do not cross) reading the zero entry: the running sum starts at
            // +0 and so is never -0, and adding +0 leaves it unchanged
            int sl[12];
#pragma unroll
            for (int e = 0; e < 12; ++e) {
                const int v = (int)(kKinds >> (5 * e)) & 31, kd = v & 7, top = v >> 3;
                // a top edge belongs to the next position (lane src + 1, or the halo entry after lane 63)
                const int sv = L.slot[kd][src + top];
                sl[e] = ((em >> e) & 1) ? sv : kEdgeSlots;
            }
#pragma unroll
            for (int e = 0; e < 12; ++e) {
                sx += L.ev[0][sl[e]];
                sy += L.ev[1][sl[e]];
                sz += L.ev[2][sl[e]];
            }
        } else {
            const int z = z0 + src;
            while (em) {  // ascending edge order
                const int e = __ffs(em) - 1;
                em &= em - 1;
                const V3<S> p = edge_vertex<S, DEF>(in, px + eox(e), py + eoy(e), z + eoz(e), eax(e));
                sx += p.x;
                sy += p.y;
                sz += p.z;
            }
        }
        const S r = rcp_small<S>(num);
        store_vertex(verts, pbase + k, out.vcap, epilogue(V3<S>{sx * r, sy * r, sz * r}, out.flags, nm));
        if (pbase + k < out.vcap) put(vinfo + pbase + k, (uint8_t)(num | (pk << 3)));
    }
    const bool vec = ((uintptr_t)faces & 15) == 0;  // a quad's four indices in 16-byte stores
    for (int k = lane; k < tq; k += kWave) {
        const int it = L.quad[k], src = it >> 3, ty = ((it >> 1) & 3) + 3 * (it & 1);
        const int64_t q = qbase + k;
        if (q >= out.fcap) continue;
        const uint32_t stw = kStencil.w[ty];
        int32_t rel[4];  // corner ids relative to the segment's first DMC vertex
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int ci = (stw >> (3 * j)) & 7, eid = (stw >> (12 + 4 * j)) & 15;
            const int p = DISO_DMC_PATCH_OF(kDmcCase[L.code[ci][src]], eid);
            // dmcEdgeOffset is -1 for a non-crossing edge (never the case on a padded grid)
            rel[j] = L.rowb[ci >> 1] + L.pre[ci][src] + (p == 7 ? -1 : p);
        }
        store_quad(faces, q, pbase, rel, out.flags, vec);
    }
}

template <typename S, bool DEF>
__global__ __launch_bounds__(kBlock, 6) void emit_dmc_kernel(In<S> in, const SegBase *__restrict__ seg,
                                                          S *__restrict__ verts, void *__restrict__ faces,
                                                          uint8_t *__restrict__ vinfo, OutCfg out, Norm<S> nm,
                                                          int run, int64_t nblocks) {
    __shared__ DmcEmitLds<S> s_lds[kWavesPerBlock];
    __shared__ uint32_t s_take;
    const int tid = threadIdx.x, wave = wave_id(), lane = tid % kWave;
    if (tid == 0) s_take = 0;
    if (lane < 3) s_lds[wave].ev[lane][kEdgeSlots] = S(0);  // the zero entry of the patch sums
    __syncthreads();
    const Geo &g = in.geo;
    const SegCtx sc{seg, info_of(seg, g), 0, 0};
    const SegBase *tot = seg + g.nseg;
    if (tot[0].v > out.fcap || tot[0].f > out.vcap) return;
    const int64_t lo = xcd_block(blockIdx.x, nblocks) * run, hi = lo + run < g.nseg ? lo + run : g.nseg;
    // segments of the run with patches (hence owned crossings)
    const uint64_t act = run_active(seg, lo, hi, lane);
    const SegPos lp = seg_pos(g, lo + lane);
    run_segments(
        act, &s_take, lane, [&](int i) { return dmc_load<S>(in, sc, lo + i, pos_of_lane(lp, i), lane); },
        [&](const DmcSegIn &d) { emit_dmc_segment<S, DEF>(in, sc, d, verts, faces, vinfo, out, nm, s_lds[wave], lane); });
}

// ------------------------------------------------------------------------------------------
// Backward (gather, deterministic)
//
// MC: one lane per grid vertex; its incident crossing edges (at most six: the three it owns and
// the ones owned by its -x, -y, -z neighbours) become items of a dense per-wave list. Each item
// computes the adjoint of its edge vertex (adjComputeMcVert, src/cumc.cu:412-453) for the lane's
// endpoint; the owning lane adds its items up in a fixed order (slot 0..5), so results are
// deterministic and no atomics are used. Item slots: 0/1/2 = edge along x/y/z owned by the
// vertex (vertex is its start), 3/4/5 = edge along x/y/z owned by the -x/-y/-z neighbour
// (vertex is its end).
//
// DMC: dmc_edge_grad_kernel first turns the DMC vertex gradients into one upstream gradient per
// crossing edge (MC vertex order = quad order): the sum over the quad's four corners of (gradient
// of that patch vertex) / (edges in its patch), the adjoint of create_dmc_verts_kernel's mean
// (src/cudualmc.cu:907-955); with DISO_FLAG_REFERENCE_COMPAT each corner reads its cell's first
// patch instead (src/cudualmc.cu:990). The MC gather then runs on those edge gradients.

template <typename S>
struct AdjIn {
    const S *__restrict__ adj;  // [n, 3]
    int64_t n;
    int flags;
    Norm<S> nm;  // upstream scaling by 1/(dims-1) when the forward normalised
};

template <typename S>
__device__ __forceinline__ V3<S> load_adj(const AdjIn<S> &a, int64_t i) {
    V3<S> r = {S(0), S(0), S(0)};
    if (i >= 0 && i < a.n) {
        r.x = a.adj[3 * i + 0];
        r.y = a.adj[3 * i + 1];
        r.z = a.adj[3 * i + 2];
        if (a.flags & DISO_FLAG_NORMALIZE) {  // grad of verts / (dims - 1): g / (dims - 1)
            r.x = div_by(r.x, a.nm.c[0], a.nm.r[0]);
            r.y = div_by(r.y, a.nm.c[1], a.nm.r[1]);
            r.z = div_by(r.z, a.nm.c[2], a.nm.r[2]);
        }
    }
    return r;
}

template <typename S>
struct Grad {
    S g, dx, dy, dz;
};

// d t / d d0 and d t / d d1 factors of t = (iso - d0) / (d1 - d0), and t itself.
template <typename S>
__device__ __forceinline__ void interp_adjoint(S d0, S d1, S iso, S *t, S *k0, S *k1) {
    *t = interp_t(d0, d1, iso);
    const S dd = (d1 - d0) * (d1 - d0);
    *k0 = (iso - d1) / dd;
    *k1 = (d0 - iso) / dd;
}
// fp32: gradients are compared within a tolerance, so one hardware reciprocal replaces the three
// correctly rounded divisions (a few ulp).
template <>
__device__ __forceinline__ void interp_adjoint<float>(float d0, float d1, float iso, float *t, float *k0, float *k1) {
    const float r = __builtin_amdgcn_rcpf(d1 - d0);
    *t = d1 != d0 ? clamp01<float>((iso - d0) * r) : 0.5f;
    const float r2 = r * r;
    *k0 = (iso - d1) * r2;
    *k1 = (d0 - iso) * r2;
}

// Adjoint of one edge vertex (v0 -> v0 + e_axis, upstream gradient gv) at endpoint `end`
// (0 = v0, 1 = v1), with the endpoints' samples d0, d1 and deforms f0, f1 (src/cumc.cu:412-453).
template <typename S, bool DEF>
__device__ __forceinline__ Grad<S> edge_adjoint(int x, int y, int z, int axis, int end, S d0, S d1, V3<S> f0,
                                                V3<S> f1, S iso, V3<S> gv) {
    S t, k0, k1;
    interp_adjoint(d0, d1, iso, &t, &k0, &k1);
    S p0x = S(x), p0y = S(y), p0z = S(z);
    S p1x = S(x + (axis == 0)), p1y = S(y + (axis == 1)), p1z = S(z + (axis == 2));
    if (DEF) {
        p0x += f0.x;
        p0y += f0.y;
        p0z += f0.z;
        p1x += f1.x;
        p1y += f1.y;
        p1z += f1.z;
    }
    const S adj_t = (p1x - p0x) * gv.x + (p1y - p0y) * gv.y + (p1z - p0z) * gv.z;
    const S wgt = end ? t : S(1) - t;
    Grad<S> r;
    r.g = (end ? k1 : k0) * adj_t;
    r.dx = wgt * gv.x;
    r.dy = wgt * gv.y;
    r.dz = wgt * gv.z;
    return r;
}

template <typename S>
__device__ __forceinline__ void store_grad(const In<S> &in, S *__restrict__ adj_sdf, S *__restrict__ adj_deform, int x,
                                           int y, int z, Grad<S> acc, int flags) {
    if (!in_data(in, x, y, z)) return;
    const int64_t i = data_index(in, x, y, z);
    if (flags & DISO_FLAG_ACCUMULATE) {
        adj_sdf[i] += acc.g;
        if (adj_deform) {
            adj_deform[3 * i + 0] += acc.dx;
            adj_deform[3 * i + 1] += acc.dy;
            adj_deform[3 * i + 2] += acc.dz;
        }
    } else {
        put(adj_sdf + i, acc.g);
        if (adj_deform) {
            put(adj_deform + 3 * i + 0, acc.dx);
            put(adj_deform + 3 * i + 1, acc.dy);
            put(adj_deform + 3 * i + 2, acc.dz);
        }
    }
}

// Per-wave LDS of the gather: the vertex id of each incident edge relative to the segment's first
// vertex; the item list (lane * 8 + slot) and one round of item results.
template <typename S>
struct BwdLds {
    int32_t id[6][kWave];
    uint16_t item[kWave * 6];
    Grad<S> res[2 * kWave];
};

// Whether segment t (-1: none) owns a crossing edge.
__device__ __forceinline__ bool seg_owns_crossing(const SegBase *__restrict__ seg, int64_t t) {
    return t >= 0 && seg[t + 1].v != seg[t].v;
}

// The edges incident to a segment's vertices are owned by the segment itself, by the segments of
// rows -x and -y, and (the z edge into lane 0) by the previous chunk: whether any of them owns a
// crossing edge.
__device__ __forceinline__ bool seg_touched(const SegBase *__restrict__ seg, const Geo &g, int64_t s, SegPos sp) {
    return seg_owns_crossing(seg, s) || (sp.px > 0 && seg_owns_crossing(seg, s - xstride(g))) ||
           (sp.py > 0 && seg_owns_crossing(seg, s - g.nchunk)) ||
           (sp.chunk > 0 && seg_owns_crossing(seg, s - 1));
}

// What a gather segment reads before any work: the vertex bases of the segment and of rows -x, -y
// (the segment's own base where the row does not exist) and, unless the masks are recomputed from
// the samples, the classification words of the vertex and of its -x, -y, -z neighbours. Loaded one
// segment ahead, so these loads overlap the previous segment's work.
struct BwdSegIn {
    int64_t s;
    SegPos sp;
    int64_t v;
    int32_t dm0, d0m;  // first vertices of rows -x, -y relative to the segment's
    uint32_t w00, wm0, w0m, wzm;
};

template <bool RECOMP>
__device__ __forceinline__ BwdSegIn bwd_load(const Geo &g, const SegCtx &sc, int64_t s, SegPos sp, int lane) {
    const SegBase *__restrict__ seg = sc.seg;
    BwdSegIn d;
    d.s = s;
    d.sp = sp;
    const int64_t sm0 = sp.px > 0 ? s - xstride(g) : -1;
    const int64_t s0m = sp.py > 0 ? s - g.nchunk : -1;
    d.v = seg[s].v;
    d.dm0 = sm0 >= 0 ? (int32_t)(seg[sm0].v - d.v) : 0;
    d.d0m = s0m >= 0 ? (int32_t)(seg[s0m].v - d.v) : 0;
    d.w00 = d.wm0 = d.w0m = d.wzm = 0;
    if constexpr (!RECOMP) {
        const uint16_t *info = sc.info;
        d.w00 = info[s * kWave + lane];
        if (sm0 >= 0) d.wm0 = info[sm0 * kWave + lane];
        if (s0m >= 0) d.w0m = info[s0m * kWave + lane];
        if (sp.chunk * kWave + lane > 0) d.wzm = info[s * kWave + lane - 1];
    }
    return d;
}

// Gather of one touched segment.
template <typename S, bool DEF, bool RECOMP>
__device__ __forceinline__ void bwd_segment(const In<S> &in, const SegCtx &sc, const BwdSegIn &sg,
                                            const AdjIn<S> &adj, S *__restrict__ adj_sdf,
                                            S *__restrict__ adj_deform, BwdLds<S> &L, int lane) {
    wave_lds_sync();  // the previous segment's LDS reads are done
    const int px = sg.sp.px, py = sg.sp.py, chunk = sg.sp.chunk;
    const int z0 = chunk * kWave, pz = z0 + lane;
    const S iso = in.iso;
    Grad<S> acc = {S(0), S(0), S(0), S(0)};
    const int64_t vabs = sg.v, vbase = vabs - sc.v0;
    // Crossing masks of the edges owned by the vertex and by its -x, -y, -z neighbours: from the
    // classification words, or (RECOMP) from the samples of the vertex's neighbourhood.
    int m00, mm0, m0m, mzm;
    if constexpr (RECOMP) {
        Win<S, -1, 3, -1, 3, -1, 3> w(in, px, py, pz, z0);
        w.template load_box<0, 1, 0, 1, -1, 3>(in);
        w.template load_box<-1, 1, 0, 1, 0, 2>(in);
        w.template load_box<-1, 1, 1, 1, 0, 1>(in);
        w.template load_box<1, 1, -1, 2, 0, 1>(in);
        w.template load_box<0, 1, -1, 1, 0, 2>(in);
        w.template load_box<0, 1, 1, 1, 0, 1>(in);
        const S c = w(0, 0, 0);
        m00 = is_anchor(in, px, py, pz) ? own_mask(c, w(1, 0, 0), w(0, 1, 0), w(0, 0, 1), iso) : 0;
        mm0 = is_anchor(in, px - 1, py, pz) ? own_mask(w(-1, 0, 0), c, w(-1, 1, 0), w(-1, 0, 1), iso) : 0;
        m0m = is_anchor(in, px, py - 1, pz) ? own_mask(w(0, -1, 0), w(1, -1, 0), c, w(0, -1, 1), iso) : 0;
        mzm = is_anchor(in, px, py, pz - 1) ? crosses(w(0, 0, -1), c, iso) << 2 : 0;
    } else {
        m00 = info_mask(sg.w00);
        mm0 = info_mask(sg.wm0);
        m0m = info_mask(sg.w0m);
        mzm = info_mask(sg.wzm);
    }
    // slots 0-2: edges the vertex owns; 3-5: the x edge of -x, the y edge of -y, the z edge of -z
    const int inc_m = (mm0 & 1) | (m0m & 2) | (mzm & 4);
    const int mask = in_data(in, px, py, pz) ? (m00 | (inc_m << 3)) : 0;
    // exclusive prefixes over the lanes of the rows' vertex counts (10-bit fields, totals <= 192)
    // and of the items, in two DPP scans
    const uint32_t cA = (uint32_t)popc3(m00) | (uint32_t)popc3(mm0) << 10 | (uint32_t)popc3(m0m) << 20;
    const uint32_t cB = (uint32_t)__popc(mask);
    const uint32_t iA = wave_incl_dpp(cA), iB = wave_incl_dpp(cB);
    const uint32_t eA = iA - cA;
    const int32_t r00 = (int32_t)(eA & 0x3ff);
    const int32_t rm0 = sg.dm0 + (int32_t)((eA >> 10) & 0x3ff);
    const int32_t r0m = sg.d0m + (int32_t)((eA >> 20) & 0x3ff);
    const int first = (int)(iB - cB), total = __builtin_amdgcn_readlane((int)iB, kWave - 1);

    L.id[0][lane] = r00;
    L.id[1][lane] = r00 + (m00 & 1);
    L.id[2][lane] = r00 + (m00 & 1) + ((m00 >> 1) & 1);
    L.id[3][lane] = rm0;
    L.id[4][lane] = r0m + (m0m & 1);
    L.id[5][lane] = r00 - 1;
    {
        // (branch-free writes to a spare entry measured 2 us slower here)
        int r = first;
#pragma unroll
        for (int sl = 0; sl < 6; ++sl)
            if ((mask >> sl) & 1) L.item[r++] = (uint16_t)(lane * 8 + sl);
    }
    wave_lds_sync();

    const int last = first + __popc(mask);
    // Adjoint of item k (an edge and the endpoint that is this segment's vertex).
    auto item = [&](int k) {
        const int it = L.item[k], src = it >> 3, slot = it & 7;
        const int axis = slot % 3, end = slot >= 3;
        // the edge runs from (x, y, z) along `axis`; the vertex is its start (slots 0-2) or its end
        const int x = px - (slot == 3), y = py - (slot == 4), z = z0 + src - (slot == 5);
        const EdgeEnds<S, DEF> e = edge_ends<S, DEF>(in, x, y, z, axis);
        return edge_adjoint<S, DEF>(x, y, z, axis, end, e.d0, e.d1, e.f0, e.f1, iso,
                                    load_adj(adj, vbase + L.id[slot][src]));
    };
    // Rounds of up to 128 items: with more than 64 left, each lane takes two (the second clamped
    // to the last item, its result unused), so both gathers are in flight together.
    for (int b = 0; b < total; b += 2 * kWave) {
        if (b + kWave < total) {  // wave-uniform
            const int k1 = b + kWave + lane < total ? b + kWave + lane : total - 1;
            const Grad<S> r0 = item(b + lane), r1 = item(k1);
            L.res[lane] = r0;
            L.res[kWave + lane] = r1;
        } else if (b + lane < total) {
            L.res[lane] = item(b + lane);
        }
        wave_lds_sync();
        const int lo = first > b ? first : b, hi = last < b + 2 * kWave ? last : b + 2 * kWave;
        for (int j = lo; j < hi; ++j) {  // this lane's items of the round, in slot order
            const Grad<S> r = L.res[j - b];
            acc.g += r.g;
            acc.dx += r.dx;
            acc.dy += r.dy;
            acc.dz += r.dz;
        }
        wave_lds_sync();
    }
    store_grad(in, adj_sdf, adj_deform, px, py, pz, acc, adj.flags);
}

// The touched segments of a block (bits of `act`, segment lo + bit) dealt round-robin over its
// waves, each wave's next segment loaded while it gathers the current one. (Taking them from a
// shared counter, as the emission kernels do, measured slower here: the gathers' costs vary
// less.)
template <typename S, bool DEF, bool RECOMP>
__device__ __forceinline__ void bwd_run(const In<S> &in, const SegCtx &sc, int64_t lo, const SegPos &lp, uint64_t act,
                                        const AdjIn<S> &adj, S *__restrict__ adj_sdf, S *__restrict__ adj_deform,
                                        BwdLds<S> &L, int wave, int lane) {
    uint64_t mine = 0;
    for (int k = 0; act; ++k, act &= act - 1)
        if ((k & (kWavesPerBlock - 1)) == wave) mine |= act & (0ull - act);
    if (!mine) return;
    int i = __builtin_ctzll(mine);
    BwdSegIn nxt = bwd_load<RECOMP>(in.geo, sc, lo + i, pos_of_lane(lp, i), lane);
    while (mine) {
        const BwdSegIn cur = nxt;
        mine &= mine - 1;
        if (mine) {
            i = __builtin_ctzll(mine);
            nxt = bwd_load<RECOMP>(in.geo, sc, lo + i, pos_of_lane(lp, i), lane);
        }
        bwd_segment<S, DEF, RECOMP>(in, sc, cur, adj, adj_sdf, adj_deform, L, lane);
    }
}

template <typename S, bool DEF, bool RECOMP>
__global__ __launch_bounds__(kBlock) void bwd_mc_kernel(In<S> in, const SegBase *__restrict__ seg, AdjIn<S> adj,
                                                        S *__restrict__ adj_sdf, S *__restrict__ adj_deform,
                                                        int run, int64_t nblocks) {
    __shared__ BwdLds<S> s_lds[kWavesPerBlock];
    const int tid = threadIdx.x, wave = wave_id(), lane = tid % kWave;
    const Geo &g = in.geo;
    const SegCtx sc{seg, info_of(seg, g), 0, 0};
    const int64_t lo = xcd_block(blockIdx.x, nblocks) * run, hi = lo + run < g.nseg ? lo + run : g.nseg;
    const int64_t t = lo + lane;
    const SegPos lp = seg_pos(g, t);
    const bool touched = t < hi && seg_touched(seg, g, t, lp);
    uint64_t act = __ballot(touched);
    // untouched segments get zero gradients (unless accumulating), dealt round-robin as well
    uint64_t idle = (adj.flags & DISO_FLAG_ACCUMULATE) ? 0ull : __ballot(t < hi && !touched);
    for (int k = 0; idle; ++k, idle &= idle - 1)
        if ((k & (kWavesPerBlock - 1)) == wave) {
            const SegPos sp = pos_of_lane(lp, __builtin_ctzll(idle));
            store_grad(in, adj_sdf, adj_deform, sp.px, sp.py, sp.chunk * kWave + lane, Grad<S>{S(0), S(0), S(0), S(0)},
                       adj.flags);
        }
    bwd_run<S, DEF, RECOMP>(in, sc, lo, lp, act, adj, adj_sdf, adj_deform, s_lds[wave], wave, lane);
}

// Per-edge upstream gradient of a DMC forward: ge[q] = sum over the quad's corners v of
// (dL/dv) / (edges in v's patch), corners in the quad's order.
template <typename S>
__device__ __forceinline__ void edge_grad_one(const void *__restrict__ quads, int64_t q,
                                              const uint8_t *__restrict__ vinfo, const AdjIn<S> &adj, const S *s_rcp,
                                              S *__restrict__ ge) {
    const bool compat = (adj.flags & DISO_FLAG_REFERENCE_COMPAT) != 0;
    V3<S> sum = {S(0), S(0), S(0)};
    // the quad's four indices: one 16-byte non-temporal load when they are 32-bit and aligned (a
    // stream read once, kept out of the caches the following gather re-reads samples from)
    int64_t qv[4];
    if (!(adj.flags & DISO_FLAG_INDEX64) && ((uintptr_t)quads & 15) == 0) {  // uniform
        const i32x4 r = __builtin_nontemporal_load(reinterpret_cast<const i32x4 *>(quads) + q);
        qv[0] = r.x;
        qv[1] = r.y;
        qv[2] = r.z;
        qv[3] = r.w;
    } else {
#pragma unroll
        for (int k = 0; k < 4; ++k) qv[k] = load_index(quads, 4 * q + k, adj.flags);
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int64_t v = qv[k];
        if (v < 0 || v >= adj.n) continue;
        const int info = vinfo[v];
        const V3<S> a = load_adj(adj, compat ? v - (info >> 3) : v);
        const S r = s_rcp[info & 7];
        sum.x += a.x * r;
        sum.y += a.y * r;
        sum.z += a.z * r;
    }
    ge[3 * q + 0] = sum.x;
    ge[3 * q + 1] = sum.y;
    ge[3 * q + 2] = sum.z;
}

template <typename S>
__global__ __launch_bounds__(kBlock) void dmc_edge_grad_kernel(const void *__restrict__ quads, int64_t nq,
                                                               const SegBase *__restrict__ tot,
                                                               const uint8_t *__restrict__ vinfo, AdjIn<S> adj,
                                                               S *__restrict__ ge) {
    __shared__ S s_rcp[8];
    if (threadIdx.x < 8) s_rcp[threadIdx.x] = threadIdx.x ? S(1) / S(threadIdx.x) : S(0);
    __syncthreads();
    // nq may be a capacity: the quads are the first tot->v rows (the count pass's total)
    const int64_t n = tot->v < nq ? tot->v : nq;
    // each XCD takes a contiguous eighth of the quads, so the vertex rows a quad shares with the
    // quads of the previous plane are still in that XCD's L2
    const int64_t q = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    if (q < n) edge_grad_one(quads, q, vinfo, adj, s_rcp, ge);
}

// ------------------------------------------------------------------------------------------
// Batched extraction: many grids (uneven dims, MC and DMC, with or without deform) per launch.
//
// The count pass runs once over the concatenation of all grids' segments (a grid's segments are
// consecutive, grids in batch order), so one single-pass scan yields every grid's bases; a grid's
// own output rows are its bases minus those of its first segment. Emission and backward launch
// one block per run of segments of one grid: a launch covers the runs of every grid of one
// kind (algorithm x deform x recompute), listed in `gl` with the prefix of their run counts in
// `rf`; a block finds its grid by binary search over `rf`.


template <typename S>
__device__ __forceinline__ SegCtx batch_ctx(const BatchGrid<S> &G, const SegBase *seg, const uint16_t *info) {
    return SegCtx{seg + G.seg_first, info + G.seg_first * kWave, seg[G.seg_first].v, seg[G.seg_first].f};
}

template <typename S>
__global__ void batch_totals_kernel(const BatchGrid<S> *__restrict__ grids, int n, int64_t nseg,
                                    const SegBase *__restrict__ seg, uint32_t *__restrict__ gflags,
                                    diso_totals_t *__restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const SegBase a = seg[grids[i].seg_first], e = seg[i + 1 < n ? grids[i + 1].seg_first : nseg];
    const int64_t cross = e.v - a.v, faces = e.f - a.f;
    const bool dmc = grids[i].algo == DISO_DMC;
    out[i].n_verts = dmc ? faces : cross;
    out[i].n_faces = dmc ? cross : faces;
    out[i].flags = (int64_t)(gflags[i] | ((uint32_t)seg[nseg + 1].v & DISO_TOTAL_SCAN_TIMEOUT));
    out[i].n_crossing = cross;
    gflags[i] = 0u;  // ready for the next batch
}

template <typename S, bool DEF>
__global__ __launch_bounds__(kBlock, 7) void emit_mc_batch_kernel(const BatchGrid<S> *__restrict__ grids,
                                                               const int32_t *__restrict__ gl,
                                                               const int64_t *__restrict__ rf, int m,
                                                               const SegBase *__restrict__ seg,
                                                               const uint16_t *__restrict__ info, int64_t nblocks) {
    __shared__ uint64_t s_case[256];
    __shared__ McEmitLds s_lds[kWavesPerBlock];
    __shared__ uint32_t s_take;
    const int tid = threadIdx.x, wave = wave_id(), lane = tid % kWave;
    for (int i = tid; i < 256; i += kBlock) s_case[i] = kMcCase[i];
    if (tid == 0) s_take = 0;
    __syncthreads();
    const int64_t b = xcd_block(blockIdx.x, nblocks);
    const int j = find_group(rf, m, b);
    const BatchGrid<S> &G = grids[gl[j]];
    if (G.nv > G.vcap || G.nf > G.fcap) return;
    const In<S> in = G.in;
    const SegCtx sc = batch_ctx(G, seg, info);
    const OutCfg out{G.flags, G.vcap, G.fcap};
    const int64_t lo = (b - rf[j]) * G.run, hi = lo + G.run < in.geo.nseg ? lo + G.run : in.geo.nseg;
    const uint64_t act = run_active(sc.seg, lo, hi, lane);
    const SegPos lp = seg_pos(in.geo, lo + lane);
    run_segments(
        act, &s_take, lane, [&](int i) { return mc_load<S>(in, sc, lo + i, pos_of_lane(lp, i), lane); },
        [&](const McSegIn &d) {
            emit_mc_segment<S, DEF>(in, sc, d, G.verts, G.faces, out, G.nm, s_case, s_lds[wave], lane);
        });
}

template <typename S, bool DEF>
__global__ __launch_bounds__(kBlock, 6) void emit_dmc_batch_kernel(const BatchGrid<S> *__restrict__ grids,
                                                                const int32_t *__restrict__ gl,
                                                                const int64_t *__restrict__ rf, int m,
                                                                const SegBase *__restrict__ seg,
                                                                const uint16_t *__restrict__ info, int64_t nblocks) {
    __shared__ DmcEmitLds<S> s_lds[kWavesPerBlock];
    __shared__ uint32_t s_take;
    const int tid = threadIdx.x, wave = wave_id(), lane = tid % kWave;
    if (tid == 0) s_take = 0;
    if (lane < 3) s_lds[wave].ev[lane][kEdgeSlots] = S(0);  // the zero entry of the patch sums
    __syncthreads();
    const int64_t b = xcd_block(blockIdx.x, nblocks);
    const int j = find_group(rf, m, b);
    const BatchGrid<S> &G = grids[gl[j]];
    if (G.nv > G.vcap || G.nf > G.fcap) return;
    const In<S> in = G.in;
    const SegCtx sc = batch_ctx(G, seg, info);
    const OutCfg out{G.flags, G.vcap, G.fcap};
    const int64_t lo = (b - rf[j]) * G.run, hi = lo + G.run < in.geo.nseg ? lo + G.run : in.geo.nseg;
    const uint64_t act = run_active(sc.seg, lo, hi, lane);
    const SegPos lp = seg_pos(in.geo, lo + lane);
    run_segments(
        act, &s_take, lane, [&](int i) { return dmc_load<S>(in, sc, lo + i, pos_of_lane(lp, i), lane); },
        [&](const DmcSegIn &d) {
            emit_dmc_segment<S, DEF>(in, sc, d, (S *)G.verts, G.faces, G.vinfo, out, G.nm, s_lds[wave], lane);
        });
}

// Per-edge upstream gradients of every DMC grid of the batch: blocks of kBlock quads of one grid.
template <typename S>
__global__ __launch_bounds__(kBlock) void dmc_edge_grad_batch_kernel(const BatchGrid<S> *__restrict__ grids,
                                                                     const int32_t *__restrict__ gl,
                                                                     const int64_t *__restrict__ rf, int m,
                                                                     S *__restrict__ ge) {
    __shared__ S s_rcp[8];
    if (threadIdx.x < 8) s_rcp[threadIdx.x] = threadIdx.x ? S(1) / S(threadIdx.x) : S(0);
    __syncthreads();
    const int64_t b = xcd_block(blockIdx.x, gridDim.x);  // as dmc_edge_grad_kernel
    const int j = find_group(rf, m, b);
    const BatchGrid<S> &G = grids[gl[j]];
    const int64_t q = (b - rf[j]) * kBlock + threadIdx.x;
    if (q >= G.nf) return;
    const AdjIn<S> a{G.adj, G.nv, G.flags, G.nm};
    edge_grad_one(G.faces, q, G.vinfo, a, s_rcp, ge + 3 * G.q_first);
}

template <typename S, bool DEF, bool RECOMP>
__global__ __launch_bounds__(kBlock) void bwd_batch_kernel(const BatchGrid<S> *__restrict__ grids,
                                                           const int32_t *__restrict__ gl,
                                                           const int64_t *__restrict__ rf, int m,
                                                           const SegBase *__restrict__ seg,
                                                           const uint16_t *__restrict__ info,
                                                           const S *__restrict__ ge, int64_t nblocks) {
    __shared__ BwdLds<S> s_lds[kWavesPerBlock];
    const int tid = threadIdx.x, wave = wave_id(), lane = tid % kWave;
    const int64_t b = xcd_block(blockIdx.x, nblocks);
    const int j = find_group(rf, m, b);
    const BatchGrid<S> &G = grids[gl[j]];
    const In<S> in = G.in;
    const Geo &g = in.geo;
    const SegCtx sc = batch_ctx(G, seg, info);
    // DMC grids gather the per-edge gradients (already divided by the normalisation)
    const AdjIn<S> adj = G.algo == DISO_DMC ? AdjIn<S>{ge + 3 * G.q_first, G.nf, G.flags & ~DISO_FLAG_NORMALIZE, G.nm}
                                            : AdjIn<S>{G.adj, G.nv, G.flags, G.nm};
    const int64_t lo = (b - rf[j]) * G.run, hi = lo + G.run < g.nseg ? lo + G.run : g.nseg;
    const int64_t t = lo + lane;
    const SegPos lp = seg_pos(g, t);
    const bool touched = t < hi && seg_touched(sc.seg, g, t, lp);
    uint64_t act = __ballot(touched);
    uint64_t idle = (adj.flags & DISO_FLAG_ACCUMULATE) ? 0ull : __ballot(t < hi && !touched);
    for (int k = 0; idle; ++k, idle &= idle - 1)
        if ((k & (kWavesPerBlock - 1)) == wave) {
            const SegPos sp = pos_of_lane(lp, __builtin_ctzll(idle));
            store_grad(in, G.adj_sdf, G.adj_deform, sp.px, sp.py, sp.chunk * kWave + lane,
                       Grad<S>{S(0), S(0), S(0), S(0)}, adj.flags);
        }
    bwd_run<S, DEF, RECOMP>(in, sc, lo, lp, act, adj, G.adj_sdf, G.adj_deform, s_lds[wave], wave, lane);
}

// ------------------------------------------------------------------------------------------
// DiffDMC quad -> triangle split (diso/__init__.py:117-147)
//
// Each quad is cut along the diagonal whose worse triangle has the smaller largest corner cosine:
// cut (0,1,3)+(1,2,3) ("first") when score1 < score2, else (0,1,2)+(0,2,3); the output holds all
// first quads' pairs in quad order, then all the others'. split_score_kernel scores the quads of a
// contiguous range per block and counts the first ones; split_emit_kernel gives each block the
// prefix and the total of those counts (it reads all of them: the grid is ~4 blocks per CU, so a
// few KB), ranks its quads with ballots and writes the pairs in place. The cosine arithmetic is
// F.normalize's (x / max(|x|, 1e-12), |x| = sqrt((x*x + y*y) + z*z)) and a left-to-right dot.

constexpr int kSplitBlock = 256;
constexpr int kSplitWaves = kSplitBlock / kWave;
constexpr int64_t kMaxSplitBlocks = 4096;

template <typename S>
__device__ __forceinline__ V3<S> unit(V3<S> v) {
    const S n = sqrt(v.x * v.x + v.y * v.y + v.z * v.z);
    const S d = n < S(1e-12) ? S(1e-12) : n;  // clamp_min(eps); keeps a NaN norm
    return {v.x / d, v.y / d, v.z / d};
}

template <typename S>
__device__ __forceinline__ S nan_max(S a, S b) {  // torch.max / np.maximum: NaN wins
    return (a != a || a > b) ? a : b;
}

// The unit vectors of a quad's six vertex pairs, e[i][j] = unit(p_j - p_i), the others by
// negation: unit(-v) = -unit(v) exactly (same norm, sign-symmetric division), and a dot product
// of negated vectors is the negated dot product exactly, so each corner cosine below has the bits
// of normalising its two edge vectors itself (24 normalisations per quad become 6).
template <typename S>
struct QuadUnits {
    V3<S> e[4][4];
    __device__ __forceinline__ QuadUnits(const V3<S> *p) {
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = i + 1; j < 4; ++j) {
                e[i][j] = unit(V3<S>{p[j].x - p[i].x, p[j].y - p[i].y, p[j].z - p[i].z});
                e[j][i] = V3<S>{-e[i][j].x, -e[i][j].y, -e[i][j].z};
            }
    }
    // cosine at corner o between the edges to a and to b (u first, then w, left-to-right dot)
    __device__ __forceinline__ S cosine(int o, int a, int b) const {
        const V3<S> &u = e[o][a], &w = e[o][b];
        return u.x * w.x + u.y * w.y + u.z * w.z;
    }
    // largest corner cosine of triangle (a, b, c), corners in the order a, b, c
    // (diso/__init__.py:127-136)
    __device__ __forceinline__ S worst(int a, int b, int c) const {
        return nan_max(nan_max(cosine(a, b, c), cosine(b, c, a)), cosine(c, a, b));
    }
};

template <typename S>
__device__ __forceinline__ V3<S> load_vert(const S *__restrict__ v, int64_t i) {
    return {v[3 * i + 0], v[3 * i + 1], v[3 * i + 2]};
}

template <typename S>
__global__ __launch_bounds__(kSplitBlock) void split_score_kernel(const S *__restrict__ verts, int64_t nv,
                                                                  const void *__restrict__ quads, int64_t nq,
                                                                  int flags, uint8_t *__restrict__ first,
                                                                  uint32_t *__restrict__ bcount, int64_t nblocks) {
    __shared__ uint32_t s_cnt[kSplitWaves];
    const int tid = threadIdx.x, lane = tid % kWave, wave = wave_id();
    const int64_t b = blockIdx.x, lo = b * nq / nblocks, hi = (b + 1) * nq / nblocks;
    uint32_t cnt = 0;
    for (int64_t q = lo + tid; q < hi; q += kSplitBlock) {
        int64_t id[4];
        bool ok = true;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            id[k] = load_index(quads, 4 * q + k, flags);
            ok &= id[k] >= 0 && id[k] < nv;
        }
        uint8_t f = 0;
        if (ok) {
            const V3<S> p[4] = {load_vert(verts, id[0]), load_vert(verts, id[1]), load_vert(verts, id[2]),
                                load_vert(verts, id[3])};
            const QuadUnits<S> qu(p);
            const S s1 = nan_max(qu.worst(0, 1, 3), qu.worst(1, 2, 3));
            const S s2 = nan_max(qu.worst(0, 1, 2), qu.worst(0, 2, 3));
            f = s1 < s2;
        }
        first[q] = f;
        cnt += f;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) cnt += (uint32_t)__shfl_xor((int)cnt, o, kWave);
    if (lane == 0) s_cnt[wave] = cnt;
    __syncthreads();
    if (tid == 0) {
        uint32_t t = 0;
        for (int w = 0; w < kSplitWaves; ++w) t += s_cnt[w];
        bcount[b] = t;
    }
}

__global__ __launch_bounds__(kSplitBlock) void split_emit_kernel(const void *__restrict__ quads, int64_t nq,
                                                                 int flags, const uint8_t *__restrict__ first,
                                                                 const uint32_t *__restrict__ bcount,
                                                                 int64_t nblocks, int64_t *__restrict__ tris) {
    __shared__ int64_t s_pre[kSplitWaves], s_tot[kSplitWaves];
    __shared__ uint32_t s_w[kSplitWaves];
    const int tid = threadIdx.x, lane = tid % kWave, wave = wave_id();
    const int64_t b = blockIdx.x, lo = b * nq / nblocks, hi = (b + 1) * nq / nblocks;
    // First quads in the blocks before this one, and in all blocks.
    int64_t pre = 0, tot = 0;
    for (int64_t j = tid; j < nblocks; j += kSplitBlock) {
        const int64_t c = bcount[j];
        tot += c;
        pre += j < b ? c : 0;
    }
    pre = wave_sum64(pre);
    tot = wave_sum64(tot);
    if (lane == 0) {
        s_pre[wave] = pre;
        s_tot[wave] = tot;
    }
    __syncthreads();
    int64_t run = 0, n1 = 0;
#pragma unroll
    for (int w = 0; w < kSplitWaves; ++w) {
        run += s_pre[w];
        n1 += s_tot[w];
    }
    for (int64_t base = lo; base < hi; base += kSplitBlock) {
        const int64_t q = base + tid;
        const bool f = q < hi && first[q];
        const uint64_t bal = __ballot(f);
        __syncthreads();  // s_w reuse
        if (lane == 0) s_w[wave] = (uint32_t)__popcll(bal);
        __syncthreads();
        int64_t r1 = run + lanes_below(bal);
        uint32_t blk = 0;
#pragma unroll
        for (int w = 0; w < kSplitWaves; ++w) {
            r1 += w < wave ? s_w[w] : 0;
            blk += s_w[w];
        }
        if (q < hi) {
            const int64_t a = load_index(quads, 4 * q + 0, flags), c1 = load_index(quads, 4 * q + 1, flags),
                          c2 = load_index(quads, 4 * q + 2, flags), c3 = load_index(quads, 4 * q + 3, flags);
            // rows 2r, 2r+1 of the first half, or of the second half at rank q - r1 among the others
            int64_t *t = tris + 6 * (f ? r1 : n1 + (q - r1));
            const int64_t v[6] = {a, c1, f ? c3 : c2, f ? c1 : a, c2, c3};
#pragma unroll
            for (int j = 0; j < 6; ++j) put(t + j, v[j]);
        }
        run += blk;
    }
}

// ------------------------------------------------------------------------------------------
// Host side

thread_local char g_err[512] = "";

int fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
int fail(int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

int check_hip(hipError_t e, const char *what) {
    if (e == hipSuccess) return DISO_OK;
    return fail(DISO_E_HIP, "%s: %s", what, hipGetErrorString(e));
}

// Launch parameters (diso_set_param, diso_set_bucket_param): emission / backward run length and
// count blocks per CU, each per size bucket, and the recount spins. A launch takes the bucket of
// its segment count (a batched launch: the segments of all its grids); the bucket limits lie
// between the segment counts of cubic grids of 32, 64, 128, 256 and 512 samples per side with
// the virtual pad (1,156 / 8,712 / 50,700 / 332,820 / 2,377,764 segments).
constexpr int kBuckets = DISO_PARAM_BUCKETS;
constexpr int64_t kBucketLimit[kBuckets - 1] = {3200, 21000, 130000, 900000};
struct Params {
    int run[kBuckets] = {32, 32, 32, 32, 32};
    int count_per_cu[kBuckets] = {kCountBlocksPerCU, kCountBlocksPerCU, kCountBlocksPerCU, kCountBlocksPerCU,
                                  kCountBlocksPerCU};
    int recount_spins = 8192;
};
Params g_params;

int bucket_of(int64_t nseg) {
    int b = 0;
    while (b < kBuckets - 1 && nseg > kBucketLimit[b]) ++b;
    return b;
}
int run_for(int64_t nseg) { return g_params.run[bucket_of(nseg)]; }

// Validates a per-bucket parameter value; returns its table (nullptr after fail()).
int *bucket_table(int which, int value) {
    switch (which) {
        case DISO_PARAM_RUN:
            if (value < 1 || value > kMaxRun) {
                fail(DISO_E_INVALID, "run length must be in [1, %d]", kMaxRun);
                return nullptr;
            }
            return g_params.run;
        case DISO_PARAM_COUNT_BLOCKS_PER_CU:
            if (value < 1 || value > kCountBlocksPerCU) {
                fail(DISO_E_INVALID, "count blocks per CU must be in [1, %d]", kCountBlocksPerCU);
                return nullptr;
            }
            return g_params.count_per_cu;
        default:
            fail(DISO_E_INVALID, "parameter %d has no size buckets", which);
            return nullptr;
    }
}

bool make_geo(int64_t X, int64_t Y, int64_t Z, int flags, Geo *g) {
    if (X < 1 || Y < 1 || Z < 1) return false;
    const int64_t lim = 1 << 20;
    if (X > lim || Y > lim || Z > lim) return false;
    const int pad = (flags & DISO_FLAG_VIRTUAL_PAD) ? 2 : 0;
    g->n0 = (int)X;
    g->n1 = (int)Y;
    g->n2 = (int)Z;
    g->g0 = (int)X + pad;
    g->g1 = (int)Y + pad;
    g->g2 = (int)Z + pad;
    g->off = pad / 2;
    g->nchunk = (g->g2 + kWave - 1) / kWave;
    g->nseg = (int64_t)g->g0 * g->g1 * g->nchunk;
    // Per-block aggregates must fit the 29/30-bit fields of the scan words; segment ids 32 bits.
    return (int64_t)g->g0 * g->g1 * g->g2 < (1ll << 36) && g->nseg < (1ll << 32);
}

// Segment bases (+ two totals entries), then one 16-bit classification word per lane.
int64_t bases_bytes_for(const Geo &g) { return (g.nseg + 2) * (int64_t)sizeof(SegBase); }
int64_t seg_bytes_for(const Geo &g) { return bases_bytes_for(g) + g.nseg * kWave * 2; }
int64_t ctl_bytes_for(const Geo &) { return kCtlHeader + 2 * kMaxCountBlocks * (int64_t)sizeof(uint64_t); }

// Number of count blocks: about kCountBlocksPerCU per CU of the current device, so that every
// block is resident at once, and never more than the segments or the word buffer allow.
int cu_count() {
    static int cus[64] = {0};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
    if (cus[dev] == 0) {
        int n = 0;
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) n = 256;
        cus[dev] = n;
    }
    return cus[dev];
}

// Count blocks of one CU that can be resident at once for count-kernel instance `kernel` (the
// single-pass scan wants every block resident; queried once per instance).
template <auto Kernel>
int count_resident_per_cu() {
    static int n = 0;
    if (n == 0) {
        int r = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&r, Kernel, kCountBlock, 0) != hipSuccess || r < 1) r = 1;
        n = r;
    }
    return n;
}

int64_t count_blocks(int64_t nseg, int resident_per_cu) {
    const int want = g_params.count_per_cu[bucket_of(nseg)];
    const int per_cu = want < resident_per_cu ? want : resident_per_cu;
    int64_t nb = (int64_t)cu_count() * per_cu;
    const int64_t min_nb = (nseg + (1ll << 20) - 1) >> 20;  // <= 2^20 segments per block keeps the
    if (nb < min_nb) nb = min_nb;                          // aggregate fields from overflowing
    if (nb > kMaxCountBlocks) nb = kMaxCountBlocks;
    if (nb > nseg) nb = nseg;
    return nb < 1 ? 1 : nb;
}

template <typename S>
In<S> make_in(const void *sdf, const void *deform, double iso, const Geo &g) {
    In<S> in;
    in.sdf = static_cast<const S *>(sdf);
    in.deform = static_cast<const S *>(deform);
    in.iso = (S)iso;
    in.pad = (S)(iso + 1.0);
    in.geo = g;
    return in;
}

template <typename S>
int launch_count(int algo, const void *sdf, double iso, const Geo &g, void *seg, void *ctl,
                 int64_t ctl_bytes, hipStream_t st) {
    CountSrc<S> src{};
    src.in = make_in<S>(sdf, nullptr, iso, g);
    src.nseg = g.nseg;
    src.recount_spins = (uint32_t)g_params.recount_spins;
    (void)ctl_bytes;
    SegBase *sb = (SegBase *)seg;
    uint16_t *info = reinterpret_cast<uint16_t *>(sb + g.nseg + 2);
    if (algo == DISO_MC) {
        const int64_t nb = count_blocks(g.nseg, count_resident_per_cu<count_kernel<S, DISO_MC, false>>());
        hipLaunchKernelGGL((count_kernel<S, DISO_MC, false>), dim3((unsigned)nb), dim3(kCountBlock), 0, st, src,
                           (const BatchGrid<S> *)nullptr, sb, info, (uint8_t *)ctl, nb);
    } else {
        const int64_t nb = count_blocks(g.nseg, count_resident_per_cu<count_kernel<S, DISO_DMC, false>>());
        hipLaunchKernelGGL((count_kernel<S, DISO_DMC, false>), dim3((unsigned)nb), dim3(kCountBlock), 0, st, src,
                           (const BatchGrid<S> *)nullptr, sb, info, (uint8_t *)ctl, nb);
    }
    return check_hip(hipGetLastError(), "count kernel launch");
}

int64_t row_blocks(const Geo &g) { return (g.nseg + run_for(g.nseg) - 1) / run_for(g.nseg); }

template <typename S>
Norm<S> norm_of(const Geo &g) {
    Norm<S> nm;
    const int n[3] = {g.n0, g.n1, g.n2};
    for (int a = 0; a < 3; ++a) {
        nm.c[a] = (S)n[a] - S(1);
        nm.r[a] = S(1) / nm.c[a];
    }
    return nm;
}

template <typename S>
int launch_emit(int algo, const void *sdf, const void *deform, double iso, int flags, const Geo &g,
                const void *seg, void *verts, int64_t vcap, void *faces, int64_t fcap, void *vinfo, hipStream_t st) {
    In<S> in = make_in<S>(sdf, deform, iso, g);
    const Norm<S> nm = norm_of<S>(g);
    const int64_t nb = row_blocks(g);
    const SegBase *sb = (const SegBase *)seg;
    const OutCfg out{flags, vcap, fcap};
    const dim3 grid((unsigned)nb), block(kBlock);
    const int run = run_for(g.nseg);
    if (algo == DISO_MC) {
        if (deform)
            hipLaunchKernelGGL((emit_mc_kernel<S, true>), grid, block, 0, st, in, sb, (S *)verts, faces, out, nm,
                               run, nb);
        else
            hipLaunchKernelGGL((emit_mc_kernel<S, false>), grid, block, 0, st, in, sb, (S *)verts, faces, out, nm,
                               run, nb);
    } else {
        if (deform)
            hipLaunchKernelGGL((emit_dmc_kernel<S, true>), grid, block, 0, st, in, sb, (S *)verts, faces,
                               (uint8_t *)vinfo, out, nm, run, nb);
        else
            hipLaunchKernelGGL((emit_dmc_kernel<S, false>), grid, block, 0, st, in, sb, (S *)verts, faces,
                               (uint8_t *)vinfo, out, nm, run, nb);
    }
    return check_hip(hipGetLastError(), "emit kernel launch");
}

template <typename S>
int launch_backward(int algo, const void *sdf, const void *deform, double iso, int flags, const Geo &g,
                    const void *seg, const void *faces, int64_t nfaces, const void *vinfo, const void *adj_verts,
                    int64_t nverts, void *adj_sdf, void *adj_deform, void *scratch, hipStream_t st) {
    In<S> in = make_in<S>(sdf, deform, iso, g);
    AdjIn<S> a;
    a.adj = (const S *)adj_verts;
    a.n = nverts;
    a.flags = flags;
    a.nm = norm_of<S>(g);
    const int run = run_for(g.nseg);
    const int64_t nb = row_blocks(g);
    const SegBase *sb = (const SegBase *)seg;
    S *gs = (S *)adj_sdf, *gd = (S *)(deform ? adj_deform : nullptr);
    if (algo == DISO_DMC) {
        // per-edge upstream gradients from the quads, then the MC gather on them
        if (nfaces > 0) {
            const int64_t eb = (nfaces + kBlock - 1) / kBlock;
            hipLaunchKernelGGL((dmc_edge_grad_kernel<S>), dim3((unsigned)eb), dim3(kBlock), 0, st, faces, nfaces,
                               sb + g.nseg, (const uint8_t *)vinfo, a, (S *)scratch);
            int rc = check_hip(hipGetLastError(), "edge gradient kernel launch");
            if (rc) return rc;
        }
        a.adj = (const S *)scratch;
        a.n = nfaces;
        a.flags = flags & ~DISO_FLAG_NORMALIZE;  // applied by the edge gradient kernel
    }
    const dim3 grid((unsigned)nb), block(kBlock);
    const bool rec = (flags & DISO_FLAG_RECOMPUTE) != 0;
    if (deform && rec)
        hipLaunchKernelGGL((bwd_mc_kernel<S, true, true>), grid, block, 0, st, in, sb, a, gs, gd, run, nb);
    else if (deform)
        hipLaunchKernelGGL((bwd_mc_kernel<S, true, false>), grid, block, 0, st, in, sb, a, gs, gd, run, nb);
    else if (rec)
        hipLaunchKernelGGL((bwd_mc_kernel<S, false, true>), grid, block, 0, st, in, sb, a, gs, gd, run, nb);
    else
        hipLaunchKernelGGL((bwd_mc_kernel<S, false, false>), grid, block, 0, st, in, sb, a, gs, gd, run, nb);
    return check_hip(hipGetLastError(), "backward kernel launch");
}

// Split: about 4 blocks per CU, each over a contiguous range of quads.
int64_t split_blocks(int64_t nq) {
    int64_t nb = (int64_t)cu_count() * 4;
    const int64_t need = (nq + kSplitBlock - 1) / kSplitBlock;
    if (nb > need) nb = need;
    if (nb > kMaxSplitBlocks) nb = kMaxSplitBlocks;
    return nb < 1 ? 1 : nb;
}

int64_t split_scratch_bytes(int64_t nq) { return kMaxSplitBlocks * (int64_t)sizeof(uint32_t) + nq; }

template <typename S>
int launch_split(const void *verts, int64_t nv, const void *quads, int64_t nq, int flags, int64_t *tris,
                 void *scratch, hipStream_t st) {
    const int64_t nb = split_blocks(nq);
    uint32_t *bcount = static_cast<uint32_t *>(scratch);
    uint8_t *first = static_cast<uint8_t *>(scratch) + kMaxSplitBlocks * sizeof(uint32_t);
    hipLaunchKernelGGL((split_score_kernel<S>), dim3((unsigned)nb), dim3(kSplitBlock), 0, st, (const S *)verts, nv,
                       quads, nq, flags, first, bcount, nb);
    int rc = check_hip(hipGetLastError(), "split score kernel launch");
    if (rc) return rc;
    hipLaunchKernelGGL(split_emit_kernel, dim3((unsigned)nb), dim3(kSplitBlock), 0, st, quads, nq, flags,
                       (const uint8_t *)first, (const uint32_t *)bcount, nb, tris);
    return check_hip(hipGetLastError(), "split emit kernel launch");
}

int check_common(int algo, int dtype, int64_t X, int64_t Y, int64_t Z, int flags, Geo *g) {
    if (algo != DISO_MC && algo != DISO_DMC) return fail(DISO_E_INVALID, "algo must be DISO_MC or DISO_DMC");
    if (dtype != DISO_F32 && dtype != DISO_F64) return fail(DISO_E_INVALID, "dtype must be DISO_F32 or DISO_F64");
    if (!make_geo(X, Y, Z, flags, g))
        return fail(DISO_E_INVALID, "unsupported grid dims (%lld, %lld, %lld)", (long long)X, (long long)Y,
                    (long long)Z);
    return DISO_OK;
}

// Pinned staging for the totals copy, one per host thread, kept for the life of the process (a
// copy into pinned memory completes without a bounce through a pageable buffer).
SegBase *pinned_totals() {
    thread_local SegBase *p = nullptr;
    if (!p) {
        void *q = nullptr;
        if (hipHostMalloc(&q, 2 * sizeof(SegBase), hipHostMallocDefault) == hipSuccess) p = (SegBase *)q;
    }
    return p;
}

// `pinned`: stage the copy in this thread's pinned buffer (allocated on first use); otherwise in
// a stack buffer (diso_count_noalloc allocates nothing).
int read_totals(const Geo &g, const void *seg, diso_totals_t *t, hipStream_t st, bool pinned = true) {
    SegBase local[2];
    SegBase *tot = pinned ? pinned_totals() : nullptr;
    if (!tot) tot = local;
    int rc = check_hip(hipMemcpyAsync(tot, (const SegBase *)seg + g.nseg, 2 * sizeof(SegBase), hipMemcpyDeviceToHost,
                                      st),
                       "totals copy");
    if (rc) return rc;
    rc = check_hip(hipStreamSynchronize(st), "stream synchronize");
    if (rc) return rc;
    // Segment bases count (crossing edges, MC triangles | DMC patches); the count kernel records
    // which algorithm produced them.
    const bool dmc = tot[1].f == DISO_DMC;
    t->n_verts = dmc ? tot[0].f : tot[0].v;
    t->n_faces = dmc ? tot[0].v : tot[0].f;
    t->flags = tot[1].v;
    t->n_crossing = tot[0].v;
    return DISO_OK;
}

// ---- batches

int64_t align256(int64_t x) { return (x + 255) & ~(int64_t)255; }

constexpr int kBatchGroups = 9;  // emit: MC/DMC x deform; backward: deform x recompute; DMC edges

// Where everything of a batch lives in its workspace: the device descriptors, the launch-group
// tables, the per-grid totals, then the segment bases and classification words of all grids.
struct BatchLayout {
    int n = 0;
    int64_t nseg = 0;
    std::vector<Geo> geo;
    std::vector<int64_t> seg_first;
    int64_t off_desc = 0, off_groups = 0, off_totals = 0, off_seg = 0, off_info = 0, total = 0;
};

int64_t group_bytes(int n) { return align256((int64_t)n * 4) + align256((int64_t)(n + 1) * 8); }

bool batch_layout(const diso_grid_t *grids, int n, BatchLayout *L) {
    if (!grids || n < 1) return false;
    L->n = n;
    L->geo.resize(n);
    L->seg_first.resize(n);
    int64_t nseg = 0;
    for (int i = 0; i < n; ++i) {
        if (!make_geo(grids[i].X, grids[i].Y, grids[i].Z, grids[i].flags, &L->geo[i])) return false;
        L->seg_first[i] = nseg;
        nseg += L->geo[i].nseg;
    }
    if (nseg >= (1ll << 32)) return false;
    L->nseg = nseg;
    L->off_desc = 0;
    L->off_groups = align256((int64_t)n * (int64_t)sizeof(BatchGrid<double>));
    L->off_totals = L->off_groups + kBatchGroups * group_bytes(n);
    L->off_seg = L->off_totals + align256((int64_t)n * (int64_t)sizeof(diso_totals_t));
    L->off_info = L->off_seg + (nseg + 2) * (int64_t)sizeof(SegBase);
    L->total = L->off_info + nseg * kWave * 2;
    return true;
}

int64_t batch_ctl_bytes(int n) { return align256(ctl_bytes_for(Geo{})) + align256((int64_t)n * 4); }

int check_batch(int dtype, const diso_grid_t *grids, int n, BatchLayout *L) {
    if (dtype != DISO_F32 && dtype != DISO_F64) return fail(DISO_E_INVALID, "dtype must be DISO_F32 or DISO_F64");
    if (!grids || n < 1) return fail(DISO_E_INVALID, "empty batch");
    for (int i = 0; i < n; ++i) {
        if (grids[i].algo != DISO_MC && grids[i].algo != DISO_DMC)
            return fail(DISO_E_INVALID, "grid %d: algo must be DISO_MC or DISO_DMC", i);
        if (!grids[i].sdf) return fail(DISO_E_INVALID, "grid %d: null sdf", i);
    }
    if (!batch_layout(grids, n, L)) return fail(DISO_E_INVALID, "unsupported grid dims in the batch");
    return DISO_OK;
}

// Device descriptors of a batch, as far as the host knows them at this point of the pipeline.
template <typename S>
std::vector<BatchGrid<S>> batch_desc(const diso_grid_t *grids, const BatchLayout &L) {
    std::vector<BatchGrid<S>> d(L.n);
    int64_t q = 0;
    for (int i = 0; i < L.n; ++i) {
        const diso_grid_t &G = grids[i];
        BatchGrid<S> &D = d[i];
        memset(&D, 0, sizeof(D));
        D.in = make_in<S>(G.sdf, G.deform, G.iso, L.geo[i]);
        D.nm = norm_of<S>(L.geo[i]);
        D.algo = G.algo;
        D.flags = G.flags;
        D.seg_first = L.seg_first[i];
        D.nv = G.totals.n_verts;
        D.nf = G.totals.n_faces;
        D.verts = (S *)G.verts;
        D.faces = G.faces;
        D.vinfo = (uint8_t *)G.vinfo;
        D.vcap = G.vert_capacity;
        D.fcap = G.face_capacity;
        D.adj = (const S *)G.adj_verts;
        D.adj_sdf = (S *)G.adj_sdf;
        D.adj_deform = (S *)G.adj_deform;
        D.q_first = q;
        D.run = run_for(L.geo[i].nseg);
        if (G.algo == DISO_DMC) q += G.totals.n_faces;
    }
    return d;
}

// One launch group: the grids it covers and the prefix of their block counts.
struct Group {
    std::vector<int32_t> gl;
    std::vector<int64_t> rf{0};
    void add(int i, int64_t blocks) {
        if (blocks <= 0) return;
        gl.push_back(i);
        rf.push_back(rf.back() + blocks);
    }
    int m() const { return (int)gl.size(); }
    int64_t blocks() const { return rf.back(); }
};

// Pinned host buffers of one host thread, kept for the life of the process: `up` stages
// host-to-device copies (reused once the copy recorded in `ev` has run, so the caller never waits
// for the device), `down` receives device-to-host copies.
struct Staging {
    void *up = nullptr, *down = nullptr;
    size_t up_cap = 0, down_cap = 0;
    hipEvent_t ev = nullptr;
    bool pending = false;
};
Staging &staging() {
    thread_local Staging s;
    return s;
}

int grow_pinned(void **p, size_t *cap, size_t n) {
    if (*cap >= n) return DISO_OK;
    if (*p) check_hip(hipHostFree(*p), "pinned free");
    *p = nullptr;
    *cap = 0;
    const size_t want = n < 65536 ? 65536 : n + n / 2;
    int rc = check_hip(hipHostMalloc(p, want, hipHostMallocDefault), "pinned allocation");
    if (rc) return rc;
    *cap = want;
    return DISO_OK;
}

// Copies n host bytes to `dst` on `st` through the pinned staging buffer, without waiting.
int stage_copy(void *dst, const void *src, size_t n, hipStream_t st) {
    Staging &S = staging();
    int rc;
    if (S.pending) {  // the previous copy from the buffer must have run before it is reused
        rc = check_hip(hipEventSynchronize(S.ev), "staging wait");
        if (rc) return rc;
        S.pending = false;
    }
    if (!S.ev && (rc = check_hip(hipEventCreateWithFlags(&S.ev, hipEventDisableTiming), "staging event"))) return rc;
    if ((rc = grow_pinned(&S.up, &S.up_cap, n))) return rc;
    memcpy(S.up, src, n);
    if ((rc = check_hip(hipMemcpyAsync(dst, S.up, n, hipMemcpyHostToDevice, st), "descriptor copy"))) return rc;
    if ((rc = check_hip(hipEventRecord(S.ev, st), "staging event"))) return rc;
    S.pending = true;
    return DISO_OK;
}

// Host staging of descriptors + group tables, copied into the workspace in one transfer.
template <typename S>
int upload_batch(void *ws, const BatchLayout &L, const std::vector<BatchGrid<S>> &d, const Group *groups, int ng,
                 hipStream_t st) {
    std::vector<uint8_t> h(L.off_totals, 0);
    memcpy(h.data() + L.off_desc, d.data(), d.size() * sizeof(BatchGrid<S>));
    for (int k = 0; k < ng; ++k) {
        uint8_t *base = h.data() + L.off_groups + k * group_bytes(L.n);
        memcpy(base, groups[k].gl.data(), groups[k].gl.size() * 4);
        memcpy(base + align256((int64_t)L.n * 4), groups[k].rf.data(), groups[k].rf.size() * 8);
    }
    return stage_copy(ws, h.data(), h.size(), st);
}

template <typename S>
struct BatchPtrs {
    BatchGrid<S> *desc;
    SegBase *seg;
    uint16_t *info;
    diso_totals_t *totals;
    const int32_t *gl(const BatchLayout &L, int k, uint8_t *ws) const {
        return (const int32_t *)(ws + L.off_groups + k * group_bytes(L.n));
    }
    const int64_t *rf(const BatchLayout &L, int k, uint8_t *ws) const {
        return (const int64_t *)(ws + L.off_groups + k * group_bytes(L.n) + align256((int64_t)L.n * 4));
    }
};

template <typename S>
BatchPtrs<S> batch_ptrs(void *ws, const BatchLayout &L) {
    uint8_t *b = (uint8_t *)ws;
    return BatchPtrs<S>{(BatchGrid<S> *)(b + L.off_desc), (SegBase *)(b + L.off_seg), (uint16_t *)(b + L.off_info),
                        (diso_totals_t *)(b + L.off_totals)};
}

template <typename S>
int batch_count(diso_grid_t *grids, const BatchLayout &L, void *ws, void *ctl, hipStream_t st) {
    const std::vector<BatchGrid<S>> d = batch_desc<S>(grids, L);
    int rc = upload_batch<S>(ws, L, d, nullptr, 0, st);
    if (rc) return rc;
    const BatchPtrs<S> P = batch_ptrs<S>(ws, L);
    CountSrc<S> src{};
    src.gflags = (uint32_t *)((uint8_t *)ctl + align256(ctl_bytes_for(Geo{})));
    src.ngrids = L.n;
    src.nseg = L.nseg;
    src.recount_spins = (uint32_t)g_params.recount_spins;
    const int64_t nb = count_blocks(L.nseg, count_resident_per_cu<count_kernel<S, DISO_MC, true>>());
    hipLaunchKernelGGL((count_kernel<S, DISO_MC, true>), dim3((unsigned)nb), dim3(kCountBlock), 0, st, src,
                       (const BatchGrid<S> *)P.desc, P.seg, P.info, (uint8_t *)ctl, nb);
    rc = check_hip(hipGetLastError(), "batch count kernel launch");
    if (rc) return rc;
    hipLaunchKernelGGL((batch_totals_kernel<S>), dim3((unsigned)((L.n + 255) / 256)), dim3(256), 0, st, P.desc, L.n,
                       L.nseg, P.seg, src.gflags, P.totals);
    rc = check_hip(hipGetLastError(), "batch totals kernel launch");
    if (rc) return rc;
    Staging &stg = staging();
    const size_t tb = (size_t)L.n * sizeof(diso_totals_t);
    if ((rc = grow_pinned(&stg.down, &stg.down_cap, tb))) return rc;
    rc = check_hip(hipMemcpyAsync(stg.down, P.totals, tb, hipMemcpyDeviceToHost, st), "batch totals copy");
    if (rc) return rc;
    rc = check_hip(hipStreamSynchronize(st), "stream synchronize");
    if (rc) return rc;
    const diso_totals_t *t = (const diso_totals_t *)stg.down;
    for (int i = 0; i < L.n; ++i) {
        grids[i].totals = t[i];
        if (t[i].flags & DISO_TOTAL_SCAN_TIMEOUT) return fail(DISO_E_TIMEOUT, "single-pass scan did not complete");
    }
    return DISO_OK;
}

int64_t runs_of(const Geo &g, int run) { return (g.nseg + run - 1) / run; }

template <typename S>
int batch_emit(const diso_grid_t *grids, const BatchLayout &L, void *ws, hipStream_t st) {
    Group gr[4];  // MC, MC + deform, DMC, DMC + deform
    for (int i = 0; i < L.n; ++i) {
        const diso_grid_t &G = grids[i];
        if (G.totals.n_verts <= 0) continue;
        if (G.totals.n_verts > G.vert_capacity || G.totals.n_faces > G.face_capacity)
            return fail(DISO_E_CAPACITY, "grid %d: output buffers too small", i);
        if (!G.verts || !G.faces || (G.algo == DISO_DMC && !G.vinfo))
            return fail(DISO_E_INVALID, "grid %d: null output buffer", i);
        gr[2 * (G.algo == DISO_DMC) + (G.deform != nullptr)].add(i, runs_of(L.geo[i], run_for(L.geo[i].nseg)));
    }
    int rc = upload_batch<S>(ws, L, batch_desc<S>(grids, L), gr, 4, st);
    if (rc) return rc;
    const BatchPtrs<S> P = batch_ptrs<S>(ws, L);
    uint8_t *w = (uint8_t *)ws;
    for (int k = 0; k < 4; ++k) {
        const int64_t nb = gr[k].blocks();
        if (!nb) continue;
        const dim3 grid((unsigned)nb), block(kBlock);
        const int32_t *gl = P.gl(L, k, w);
        const int64_t *rf = P.rf(L, k, w);
        const int m = gr[k].m();
        switch (k) {
            case 0: hipLaunchKernelGGL((emit_mc_batch_kernel<S, false>), grid, block, 0, st, P.desc, gl, rf, m, P.seg, P.info, nb); break;
            case 1: hipLaunchKernelGGL((emit_mc_batch_kernel<S, true>), grid, block, 0, st, P.desc, gl, rf, m, P.seg, P.info, nb); break;
            case 2: hipLaunchKernelGGL((emit_dmc_batch_kernel<S, false>), grid, block, 0, st, P.desc, gl, rf, m, P.seg, P.info, nb); break;
            default: hipLaunchKernelGGL((emit_dmc_batch_kernel<S, true>), grid, block, 0, st, P.desc, gl, rf, m, P.seg, P.info, nb); break;
        }
        rc = check_hip(hipGetLastError(), "batch emit kernel launch");
        if (rc) return rc;
    }
    return DISO_OK;
}

int64_t batch_scratch_bytes(int dtype, const diso_grid_t *grids, int n) {
    int64_t q = 0;
    for (int i = 0; i < n; ++i)
        if (grids[i].algo == DISO_DMC && grids[i].totals.n_faces > 0) q += grids[i].totals.n_faces;
    return q * 3 * (dtype == DISO_F32 ? 4 : 8);
}

template <typename S>
int batch_backward(const diso_grid_t *grids, const BatchLayout &L, void *ws, void *scratch, hipStream_t st) {
    Group gr[5];  // 0-3: gather, deform x recompute; 4: DMC per-edge gradients
    for (int i = 0; i < L.n; ++i) {
        const diso_grid_t &G = grids[i];
        if (!G.adj_sdf || G.totals.n_verts <= 0) continue;
        if (G.deform && !G.adj_deform) return fail(DISO_E_INVALID, "grid %d: adj_deform required", i);
        if (!G.adj_verts) return fail(DISO_E_INVALID, "grid %d: null adj_verts", i);
        if (G.algo == DISO_DMC) {
            if (!G.faces || !G.vinfo) return fail(DISO_E_INVALID, "grid %d: DMC backward needs quads and vinfo", i);
            gr[4].add(i, (G.totals.n_faces + kBlock - 1) / kBlock);
        }
        gr[2 * (G.deform != nullptr) + ((G.flags & DISO_FLAG_RECOMPUTE) != 0)].add(i, runs_of(L.geo[i], run_for(L.geo[i].nseg)));
    }
    int rc = upload_batch<S>(ws, L, batch_desc<S>(grids, L), gr, 5, st);
    if (rc) return rc;
    const BatchPtrs<S> P = batch_ptrs<S>(ws, L);
    uint8_t *w = (uint8_t *)ws;
    S *ge = (S *)scratch;
    if (gr[4].blocks()) {
        hipLaunchKernelGGL((dmc_edge_grad_batch_kernel<S>), dim3((unsigned)gr[4].blocks()), dim3(kBlock), 0, st,
                           P.desc, P.gl(L, 4, w), P.rf(L, 4, w), gr[4].m(), ge);
        rc = check_hip(hipGetLastError(), "batch edge gradient kernel launch");
        if (rc) return rc;
    }
    for (int k = 0; k < 4; ++k) {
        const int64_t nb = gr[k].blocks();
        if (!nb) continue;
        const dim3 grid((unsigned)nb), block(kBlock);
        const int32_t *gl = P.gl(L, k, w);
        const int64_t *rf = P.rf(L, k, w);
        const int m = gr[k].m();
        switch (k) {
            case 0: hipLaunchKernelGGL((bwd_batch_kernel<S, false, false>), grid, block, 0, st, P.desc, gl, rf, m, P.seg, P.info, ge, nb); break;
            case 1: hipLaunchKernelGGL((bwd_batch_kernel<S, false, true>), grid, block, 0, st, P.desc, gl, rf, m, P.seg, P.info, ge, nb); break;
            case 2: hipLaunchKernelGGL((bwd_batch_kernel<S, true, false>), grid, block, 0, st, P.desc, gl, rf, m, P.seg, P.info, ge, nb); break;
            default: hipLaunchKernelGGL((bwd_batch_kernel<S, true, true>), grid, block, 0, st, P.desc, gl, rf, m, P.seg, P.info, ge, nb); break;
        }
        rc = check_hip(hipGetLastError(), "batch backward kernel launch");
        if (rc) return rc;
    }
    return DISO_OK;
}

}  // namespace

// ------------------------------------------------------------------------------------------
// Counting without allocation (diso_count_noalloc): the single-pass scan's control block lives
// in this file-scope device array instead of a caller workspace. Each device holds its own copy
// (at its own address, from hipGetSymbolAddress) until the process ends. It is cleared once per
// device by a fill kernel on the first call's stream; after that the scan keeps it consistent
// from call to call (epoch-tagged words, each call clears the buffer the next one uses), so no
// per-call clearing is needed. Every call on a device shares the one array: calls that overlap
// in time on one device race.
constexpr int64_t kNoallocCtlBytes = kCtlHeader + 2 * kMaxCountBlocks * (int64_t)sizeof(uint64_t);
static_assert(kNoallocCtlBytes % 8 == 0, "control block is cleared in 8-byte words");
static __device__ __align__(256) uint8_t scratch[kNoallocCtlBytes];

static __global__ void clear_words_kernel(uint64_t *p, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0ull;
}

namespace {

std::mutex g_scratch_mu;
void *g_scratch_ptr[64];  // device address of `scratch`, per device ordinal

// The current device's copy of `scratch` (looked up, and cleared on `st`, on the device's first
// call). Host code uses only this address, never the array's name.
int noalloc_ctl(hipStream_t st, void **out) {
    int dev = -1;
    int rc = check_hip(hipGetDevice(&dev), "hipGetDevice");
    if (rc) return rc;
    if (dev < 0 || dev >= 64) return fail(DISO_E_INVALID, "device ordinal %d out of range", dev);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        rc = check_hip(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        if (rc) return rc;
        if (!p) return fail(DISO_E_HIP, "hipGetSymbolAddress(scratch) returned a null address");
        hipLaunchKernelGGL(clear_words_kernel, dim3(16), dim3(256), 0, st, (uint64_t *)p, kNoallocCtlBytes / 8);
        rc = check_hip(hipGetLastError(), "control block clear launch");
        if (rc) return rc;
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return DISO_OK;
}

}  // namespace

// ------------------------------------------------------------------------------------------
// C ABI

extern "C" {

const char *diso_last_error(void) { return g_err; }

const char *diso_version(void) { return "diso_hip 0.1.0 gfx950"; }

int64_t diso_seg_bytes(int64_t X, int64_t Y, int64_t Z, int flags) {
    Geo g;
    if (!make_geo(X, Y, Z, flags, &g)) return fail(DISO_E_INVALID, "unsupported grid dims");
    return (flags & DISO_FLAG_RECOMPUTE) ? bases_bytes_for(g) : seg_bytes_for(g);
}

int64_t diso_totals_offset(int64_t X, int64_t Y, int64_t Z, int flags) {
    Geo g;
    if (!make_geo(X, Y, Z, flags, &g)) return fail(DISO_E_INVALID, "unsupported grid dims");
    return g.nseg * (int64_t)sizeof(SegBase);
}

int64_t diso_ctl_bytes(int64_t X, int64_t Y, int64_t Z, int flags) {
    Geo g;
    if (!make_geo(X, Y, Z, flags, &g)) return fail(DISO_E_INVALID, "unsupported grid dims");
    return ctl_bytes_for(g);
}

int diso_count(int algo, int dtype, const void *sdf, int64_t X, int64_t Y, int64_t Z, double iso,
               int flags, void *seg, int64_t seg_bytes, void *ctl, int64_t ctl_bytes,
               diso_totals_t *totals_host, void *stream) {
    g_err[0] = 0;
    Geo g;
    int rc = check_common(algo, dtype, X, Y, Z, flags, &g);
    if (rc) return rc;
    if (!sdf || !seg || !ctl) return fail(DISO_E_INVALID, "null pointer argument");
    if (seg_bytes < seg_bytes_for(g))
        return fail(DISO_E_CAPACITY, "seg workspace too small: %lld < %lld bytes", (long long)seg_bytes,
                    (long long)seg_bytes_for(g));
    if (ctl_bytes < ctl_bytes_for(g))
        return fail(DISO_E_CAPACITY, "ctl workspace too small: %lld < %lld bytes", (long long)ctl_bytes,
                    (long long)ctl_bytes_for(g));
    hipStream_t st = (hipStream_t)stream;
    rc = dtype == DISO_F32 ? launch_count<float>(algo, sdf, iso, g, seg, ctl, ctl_bytes, st)
                           : launch_count<double>(algo, sdf, iso, g, seg, ctl, ctl_bytes, st);
    if (rc) return rc;
    if (totals_host) {
        rc = read_totals(g, seg, totals_host, st);
        if (rc) return rc;
        if (totals_host->flags & DISO_TOTAL_SCAN_TIMEOUT)
            return fail(DISO_E_TIMEOUT, "single-pass scan did not complete");
    }
    return DISO_OK;
}

int diso_count_noalloc(int algo, int dtype, const void *sdf, int64_t X, int64_t Y, int64_t Z, double iso,
                       int flags, void *seg, int64_t seg_bytes, diso_totals_t *totals_host, void *stream) {
    g_err[0] = 0;
    Geo g;
    int rc = check_common(algo, dtype, X, Y, Z, flags, &g);
    if (rc) return rc;
    if (!sdf || !seg) return fail(DISO_E_INVALID, "null pointer argument");
    if (seg_bytes < seg_bytes_for(g))
        return fail(DISO_E_CAPACITY, "seg workspace too small: %lld < %lld bytes", (long long)seg_bytes,
                    (long long)seg_bytes_for(g));
    if (ctl_bytes_for(g) > kNoallocCtlBytes)
        return fail(DISO_E_CAPACITY, "grid needs a %lld-byte control block, more than the %lld bytes of the "
                    "library's own", (long long)ctl_bytes_for(g), (long long)kNoallocCtlBytes);
    hipStream_t st = (hipStream_t)stream;
    void *ctl = nullptr;
    rc = noalloc_ctl(st, &ctl);
    if (rc) return rc;
    rc = dtype == DISO_F32 ? launch_count<float>(algo, sdf, iso, g, seg, ctl, kNoallocCtlBytes, st)
                           : launch_count<double>(algo, sdf, iso, g, seg, ctl, kNoallocCtlBytes, st);
    if (rc) return rc;
    if (totals_host) {
        rc = read_totals(g, seg, totals_host, st, false);
        if (rc) return rc;
        if (totals_host->flags & DISO_TOTAL_SCAN_TIMEOUT)
            return fail(DISO_E_TIMEOUT, "single-pass scan did not complete");
    }
    return DISO_OK;
}

int diso_read_totals(const void *seg, int64_t seg_bytes, int64_t X, int64_t Y, int64_t Z, int flags,
                     diso_totals_t *totals_host, void *stream) {
    g_err[0] = 0;
    Geo g;
    if (!make_geo(X, Y, Z, flags, &g)) return fail(DISO_E_INVALID, "unsupported grid dims");
    if (!seg || !totals_host) return fail(DISO_E_INVALID, "null pointer argument");
    if (seg_bytes < bases_bytes_for(g)) return fail(DISO_E_CAPACITY, "seg workspace too small");
    return read_totals(g, seg, totals_host, (hipStream_t)stream);
}

int64_t diso_scratch_bytes(int algo, int dtype, int64_t n_crossing) {
    if ((algo != DISO_MC && algo != DISO_DMC) || (dtype != DISO_F32 && dtype != DISO_F64) || n_crossing < 0)
        return fail(DISO_E_INVALID, "bad scratch query");
    if (algo == DISO_MC) return 0;
    return n_crossing * 3 * (dtype == DISO_F32 ? 4 : 8);
}

int diso_emit(int algo, int dtype, const void *sdf, const void *deform, int64_t X, int64_t Y, int64_t Z,
              double iso, int flags, const void *seg, int64_t seg_bytes, void *verts, int64_t vert_capacity,
              void *faces, int64_t face_capacity, void *vinfo, void *stream) {
    g_err[0] = 0;
    Geo g;
    int rc = check_common(algo, dtype, X, Y, Z, flags, &g);
    if (rc) return rc;
    if (!sdf || !seg) return fail(DISO_E_INVALID, "null pointer argument");
    if (seg_bytes < seg_bytes_for(g)) return fail(DISO_E_CAPACITY, "seg workspace too small");
    if (vert_capacity < 0 || face_capacity < 0) return fail(DISO_E_INVALID, "negative capacity");
    if ((vert_capacity && !verts) || (face_capacity && !faces)) return fail(DISO_E_INVALID, "null output buffer");
    if (algo == DISO_DMC && vert_capacity && !vinfo) return fail(DISO_E_INVALID, "DMC needs a vinfo buffer");
    hipStream_t st = (hipStream_t)stream;
    return dtype == DISO_F32 ? launch_emit<float>(algo, sdf, deform, iso, flags, g, seg, verts, vert_capacity, faces,
                                                  face_capacity, vinfo, st)
                             : launch_emit<double>(algo, sdf, deform, iso, flags, g, seg, verts, vert_capacity, faces,
                                                   face_capacity, vinfo, st);
}

int diso_backward(int algo, int dtype, const void *sdf, const void *deform, int64_t X, int64_t Y, int64_t Z,
                  double iso, int flags, const void *seg, int64_t seg_bytes, const void *faces, int64_t n_faces,
                  const void *vinfo, const void *adj_verts, int64_t n_verts, void *adj_sdf, void *adj_deform,
                  void *scratch, int64_t scratch_bytes, void *stream) {
    g_err[0] = 0;
    Geo g;
    int rc = check_common(algo, dtype, X, Y, Z, flags, &g);
    if (rc) return rc;
    if (!sdf || !seg || !adj_sdf) return fail(DISO_E_INVALID, "null pointer argument");
    if (deform && !adj_deform) return fail(DISO_E_INVALID, "adj_deform required when deform is given");
    if (n_verts < 0 || (n_verts && !adj_verts)) return fail(DISO_E_INVALID, "bad adj_verts");
    if (seg_bytes < ((flags & DISO_FLAG_RECOMPUTE) ? bases_bytes_for(g) : seg_bytes_for(g)))
        return fail(DISO_E_CAPACITY, "seg workspace too small");
    if (algo == DISO_DMC) {
        if (n_faces < 0 || (n_faces && (!faces || !vinfo))) return fail(DISO_E_INVALID, "DMC backward needs the quads and vinfo");
        if (scratch_bytes < diso_scratch_bytes(algo, dtype, n_faces) || (n_faces && !scratch))
            return fail(DISO_E_CAPACITY, "scratch buffer too small for %lld quads", (long long)n_faces);
    }
    hipStream_t st = (hipStream_t)stream;
    return dtype == DISO_F32 ? launch_backward<float>(algo, sdf, deform, iso, flags, g, seg, faces, n_faces, vinfo,
                                                      adj_verts, n_verts, adj_sdf, adj_deform, scratch, st)
                             : launch_backward<double>(algo, sdf, deform, iso, flags, g, seg, faces, n_faces, vinfo,
                                                       adj_verts, n_verts, adj_sdf, adj_deform, scratch, st);
}

int64_t diso_split_scratch_bytes(int64_t n_quads) {
    if (n_quads < 0) return fail(DISO_E_INVALID, "negative quad count");
    return split_scratch_bytes(n_quads);
}

int diso_split(int dtype, const void *verts, int64_t n_verts, const void *quads, int64_t n_quads, int flags,
               int64_t *tris, int64_t tri_capacity, void *scratch, int64_t scratch_bytes, void *stream) {
    g_err[0] = 0;
    if (dtype != DISO_F32 && dtype != DISO_F64) return fail(DISO_E_INVALID, "dtype must be DISO_F32 or DISO_F64");
    if (n_quads < 0 || n_verts < 0) return fail(DISO_E_INVALID, "negative size");
    if (n_quads == 0) return DISO_OK;
    if (!verts || !quads || !tris || !scratch) return fail(DISO_E_INVALID, "null pointer argument");
    if (tri_capacity < 2 * n_quads)
        return fail(DISO_E_CAPACITY, "triangle buffer too small: %lld < %lld rows", (long long)tri_capacity,
                    (long long)(2 * n_quads));
    if (scratch_bytes < split_scratch_bytes(n_quads)) return fail(DISO_E_CAPACITY, "scratch buffer too small");
    hipStream_t st = (hipStream_t)stream;
    return dtype == DISO_F32 ? launch_split<float>(verts, n_verts, quads, n_quads, flags, tris, scratch, st)
                             : launch_split<double>(verts, n_verts, quads, n_quads, flags, tris, scratch, st);
}

int diso_set_param(int which, int value) {
    g_err[0] = 0;
    switch (which) {
        case DISO_PARAM_RUN:
        case DISO_PARAM_COUNT_BLOCKS_PER_CU: {
            int *t = bucket_table(which, value);
            if (!t) return DISO_E_INVALID;
            for (int b = 0; b < kBuckets; ++b) t[b] = value;
            return DISO_OK;
        }
        case DISO_PARAM_RECOUNT_SPINS:
            if (value < 0 || value > (int)kSpinLimit)
                return fail(DISO_E_INVALID, "recount spins must be in [0, %u]", kSpinLimit);
            g_params.recount_spins = value;
            return DISO_OK;
        default:
            return fail(DISO_E_INVALID, "unknown parameter %d", which);
    }
}

int diso_get_param(int which) {
    switch (which) {
        case DISO_PARAM_RUN: return g_params.run[kBuckets - 2];
        case DISO_PARAM_COUNT_BLOCKS_PER_CU: return g_params.count_per_cu[kBuckets - 2];
        case DISO_PARAM_RECOUNT_SPINS: return g_params.recount_spins;
        default: return fail(DISO_E_INVALID, "unknown parameter %d", which);
    }
}

int diso_set_bucket_param(int which, int bucket, int value) {
    g_err[0] = 0;
    if (bucket < 0 || bucket >= kBuckets) return fail(DISO_E_INVALID, "bucket must be in [0, %d]", kBuckets - 1);
    int *t = bucket_table(which, value);
    if (!t) return DISO_E_INVALID;
    t[bucket] = value;
    return DISO_OK;
}

int diso_get_bucket_param(int which, int bucket) {
    if (bucket < 0 || bucket >= kBuckets) return fail(DISO_E_INVALID, "bucket must be in [0, %d]", kBuckets - 1);
    switch (which) {
        case DISO_PARAM_RUN: return g_params.run[bucket];
        case DISO_PARAM_COUNT_BLOCKS_PER_CU: return g_params.count_per_cu[bucket];
        default: return fail(DISO_E_INVALID, "parameter %d has no size buckets", which);
    }
}

int diso_param_bucket(int64_t X, int64_t Y, int64_t Z, int flags) {
    Geo g;
    if (!make_geo(X, Y, Z, flags, &g)) return fail(DISO_E_INVALID, "bad grid dims");
    return bucket_of(g.nseg);
}

int64_t diso_batch_seg_bytes(const diso_grid_t *grids, int n) {
    BatchLayout L;
    if (!batch_layout(grids, n, &L)) return fail(DISO_E_INVALID, "bad batch");
    return L.total;
}

int64_t diso_batch_ctl_bytes(int n) {
    if (n < 1) return fail(DISO_E_INVALID, "empty batch");
    return batch_ctl_bytes(n);
}

int diso_batch_count(int dtype, diso_grid_t *grids, int n, void *seg, int64_t seg_bytes, void *ctl, int64_t ctl_bytes,
                     void *stream) {
    g_err[0] = 0;
    BatchLayout L;
    int rc = check_batch(dtype, grids, n, &L);
    if (rc) return rc;
    if (!seg || !ctl) return fail(DISO_E_INVALID, "null pointer argument");
    if (seg_bytes < L.total) return fail(DISO_E_CAPACITY, "batch workspace too small");
    if (ctl_bytes < batch_ctl_bytes(n)) return fail(DISO_E_CAPACITY, "batch ctl workspace too small");
    hipStream_t st = (hipStream_t)stream;
    return dtype == DISO_F32 ? batch_count<float>(grids, L, seg, ctl, st) : batch_count<double>(grids, L, seg, ctl, st);
}

int diso_batch_emit(int dtype, const diso_grid_t *grids, int n, void *seg, int64_t seg_bytes, void *stream) {
    g_err[0] = 0;
    BatchLayout L;
    int rc = check_batch(dtype, grids, n, &L);
    if (rc) return rc;
    if (!seg) return fail(DISO_E_INVALID, "null pointer argument");
    if (seg_bytes < L.total) return fail(DISO_E_CAPACITY, "batch workspace too small");
    hipStream_t st = (hipStream_t)stream;
    return dtype == DISO_F32 ? batch_emit<float>(grids, L, seg, st) : batch_emit<double>(grids, L, seg, st);
}

int64_t diso_batch_scratch_bytes(int dtype, const diso_grid_t *grids, int n) {
    if ((dtype != DISO_F32 && dtype != DISO_F64) || !grids || n < 1) return fail(DISO_E_INVALID, "bad batch");
    return batch_scratch_bytes(dtype, grids, n);
}

int diso_batch_backward(int dtype, const diso_grid_t *grids, int n, void *seg, int64_t seg_bytes, void *scratch,
                        int64_t scratch_bytes, void *stream) {
    g_err[0] = 0;
    BatchLayout L;
    int rc = check_batch(dtype, grids, n, &L);
    if (rc) return rc;
    if (!seg) return fail(DISO_E_INVALID, "null pointer argument");
    if (seg_bytes < L.total) return fail(DISO_E_CAPACITY, "batch workspace too small");
    const int64_t need = batch_scratch_bytes(dtype, grids, n);
    if (scratch_bytes < need || (need && !scratch)) return fail(DISO_E_CAPACITY, "batch scratch too small");
    hipStream_t st = (hipStream_t)stream;
    return dtype == DISO_F32 ? batch_backward<float>(grids, L, seg, scratch, st)
                             : batch_backward<double>(grids, L, seg, scratch, st);
}

}  // extern "C"
