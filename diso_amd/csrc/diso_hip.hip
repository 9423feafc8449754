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
//   and only the per-segment bases are stored (16 bytes per 64 cells). Every kernel recomputes
//   per-cell classifications from the sdf (an L1/L2-resident 3x3 row neighbourhood) and takes the
//   in-wave prefix from bit-plane ballots + popcounts.
//
// Kernels
//   count_kernel    classify + count + single-pass decoupled look-back scan of segment counts.
//   emit_mc_kernel  MC vertices and triangles (+ shift / normalise / int64 epilogue).
//   emit_dmc_kernel DMC vertices (patch means) and quads.
//   bwd_mc_kernel   MC gradient: one thread per grid vertex gathers its <= 6 incident edges.
//   bwd_dmc_kernel  DMC gradient: as above, each edge first sums its <= 4 incident patch adjoints.
//
// Floating point: only IEEE +, -, *, / and compares, in the reference's operation order, built
// with -ffp-contract=off and correctly rounded division, so fp32 results equal an fp32 CPU
// evaluation of the same expressions bit for bit.
#include <hip/hip_runtime.h>

#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include "diso_hip.h"
#include "diso_tables.h"

namespace {

constexpr int kWave = 64;
constexpr int kWavesPerBlock = 4;
constexpr int kBlock = kWave * kWavesPerBlock;
constexpr uint32_t kSpinLimit = 1u << 22;  // bounded look-back wait (~seconds); never expected
constexpr int kCtlHeader = 64;             // bytes

__constant__ uint64_t kMcCase[256] = {DISO_MC_CASE_VALUES};
__constant__ uint64_t kDmcCase[256] = {DISO_DMC_CASE_VALUES};
__constant__ int kEdgeOfs[12][3] = {DISO_EDGE_OFS_VALUES};
__constant__ int kEdgeAxis[12] = {DISO_EDGE_AXIS_VALUES};
__constant__ int kQuad[6][4][4] = {DISO_QUAD_VALUES};

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

template <typename S>
__device__ __forceinline__ bool in_data(const In<S> &in, int x, int y, int z) {
    return (unsigned)(x - in.geo.off) < (unsigned)in.geo.n0 &&
           (unsigned)(y - in.geo.off) < (unsigned)in.geo.n1 &&
           (unsigned)(z - in.geo.off) < (unsigned)in.geo.n2;
}

template <typename S>
__device__ __forceinline__ int64_t data_index(const In<S> &in, int x, int y, int z) {
    return ((int64_t)(x - in.geo.off) * in.geo.n1 + (y - in.geo.off)) * in.geo.n2 + (z - in.geo.off);
}

// sdf sample at conceptual (x,y,z); samples outside the stored grid read as the pad value, so no
// coordinate can address memory outside the input.
template <typename S>
__device__ __forceinline__ S sample(const In<S> &in, int x, int y, int z) {
    return in_data(in, x, y, z) ? in.sdf[data_index(in, x, y, z)] : in.pad;
}

template <typename S, bool DEF>
__device__ __forceinline__ V3<S> deform_at(const In<S> &in, int x, int y, int z) {
    V3<S> r = {S(0), S(0), S(0)};
    if (DEF && in_data(in, x, y, z)) {
        const S *d = in.deform + 3 * data_index(in, x, y, z);
        r.x = d[0];
        r.y = d[1];
        r.z = d[2];
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

// computeMcVert (src/cumc.cu:343-368): p0 + (p1 - p0) * t with p = integer position + deform.
template <typename S, bool DEF>
__device__ __forceinline__ V3<S> edge_vertex(const In<S> &in, int x, int y, int z, int axis) {
    int x1 = x + (axis == 0), y1 = y + (axis == 1), z1 = z + (axis == 2);
    S d0 = sample(in, x, y, z), d1 = sample(in, x1, y1, z1);
    S t = interp_t(d0, d1, in.iso);
    S p0x = S(x), p0y = S(y), p0z = S(z);
    S p1x = S(x1), p1y = S(y1), p1z = S(z1);
    if (DEF) {
        V3<S> a = deform_at<S, DEF>(in, x, y, z), b = deform_at<S, DEF>(in, x1, y1, z1);
        p0x += a.x;
        p0y += a.y;
        p0z += a.z;
        p1x += b.x;
        p1y += b.y;
        p1z += b.z;
    }
    V3<S> r;
    r.x = p0x + (p1x - p0x) * t;
    r.y = p0y + (p1y - p0y) * t;
    r.z = p0z + (p1z - p0z) * t;
    return r;
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
    int dir = DISO_DMC_DIR(cases[code]);
    if (dir != 7) {
        int n[3] = {x, y, z};
        int comp = dir >> 1;
        n[comp] += (dir & 1) ? 1 : -1;
        int lim = comp == 0 ? in.geo.g0 : (comp == 1 ? in.geo.g1 : in.geo.g2);
        if (n[comp] >= 0 && n[comp] < lim - 1) {
            int ncode = dmc_code(load_corners(in, n[0], n[1], n[2]), in.iso);
            if (DISO_DMC_DIR(cases[ncode]) != 7) code ^= 0xff;
        }
    }
    return code;
}

// Resolved DMC code of the cell anchored at (x,y,z), or 0 (no patches) when it is not a cell.
template <typename S>
__device__ __forceinline__ int dmc_cell(const In<S> &in, const uint64_t *cases, int x, int y, int z) {
    if (!is_anchor(in, x, y, z)) return 0;
    int code = dmc_code(load_corners(in, x, y, z), in.iso);
    if (code == 0 || code == 255) return code;
    return dmc_good_code(in, cases, code, x, y, z);
}

// Crossing flags (bit a = edge along axis a) of the edges owned by the cell at (x,y,z).
template <typename S>
__device__ __forceinline__ int owned_crossings(const In<S> &in, int x, int y, int z) {
    if (!is_anchor(in, x, y, z)) return 0;
    S d0 = sample(in, x, y, z);
    int m = 0;
    m |= crosses(d0, sample(in, x + 1, y, z), in.iso) << 0;
    m |= crosses(d0, sample(in, x, y + 1, z), in.iso) << 1;
    m |= crosses(d0, sample(in, x, y, z + 1), in.iso) << 2;
    return m;
}

__device__ __forceinline__ int popc3(int m) { return (m & 1) + ((m >> 1) & 1) + ((m >> 2) & 1); }

// ------------------------------------------------------------------------------------------
// Wave primitives

__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

// Exclusive prefix over the wave of a per-lane value in [0, 7] (three ballots); *total gets the
// wave sum. Must be reached by all 64 lanes.
__device__ __forceinline__ int wave_excl_small(int v, int *total) {
    uint64_t lt = (1ull << lane_id()) - 1ull;
    uint64_t b0 = __ballot(v & 1), b1 = __ballot(v & 2), b2 = __ballot(v & 4);
    *total = __popcll(b0) + 2 * __popcll(b1) + 4 * __popcll(b2);
    return __popcll(b0 & lt) + 2 * __popcll(b1 & lt) + 4 * __popcll(b2 & lt);
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

__device__ __forceinline__ int64_t seg_index(const Geo &g, int px, int py, int chunk) {
    return ((int64_t)px * g.g1 + py) * g.nchunk + chunk;
}

// ------------------------------------------------------------------------------------------
// Pass 1: classify, count, scan (single pass)
//
// The grid of segments is split into G contiguous ranges, one per block (G ~ 4 blocks per CU, so
// all blocks are resident). A block counts its segments (writing each segment's counts into
// `seg`), publishes its aggregate as one self-contained 64-bit word, reads the aggregates of all
// its predecessors in parallel (a look-back over the whole window at once: no chain of inclusive
// prefixes, so the wait is one memory latency rather than one per predecessor), and finally scans
// its own segment counts into segment bases. Aggregate words are accessed with agent-scope
// atomics only, so they are coherent across the XCDs' L2s without fences. Two word buffers
// alternate between calls; each call clears the one the next call uses (up to the extent that
// was written), and the last block flips the parity, so no initialisation launch is needed.
//
// Aggregate word: [63:62] status (1 = published), [61:59] sample flags, [58:29] face / patch
// count, [28:0] vertex / quad count.

struct CtlHeader {
    uint32_t parity;
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
__device__ __forceinline__ uint64_t pack_agg(int64_t v, int64_t f, uint32_t fl) {
    return (1ull << 62) | ((uint64_t)fl << 59) | ((uint64_t)f << 29) | (uint64_t)v;
}

// Per-position counts for pass 1: nv = crossing edges owned (MC vertices / DMC quads);
// nf = MC triangles / DMC patches; fl = sample flags of the position's own sample.
template <typename S, int ALGO>
__device__ __forceinline__ void count_position(const In<S> &in, const uint64_t *cases, int px, int py,
                                               int pz, int *nv, int *nf, uint32_t *fl) {
    *nv = 0;
    *nf = 0;
    *fl = 0;
    if (in_data(in, px, py, pz)) {
        S v = in.sdf[data_index(in, px, py, pz)];
        *fl = (v < in.iso ? DISO_TOTAL_HAS_BELOW : 0u) | (v > in.iso ? DISO_TOTAL_HAS_ABOVE : 0u) |
              (v != v ? DISO_TOTAL_HAS_NAN : 0u);
    }
    if (!is_anchor(in, px, py, pz)) return;
    Corners<S> k = load_corners(in, px, py, pz);
    S d0 = k.c[0][0][0];
    *nv = crosses(d0, k.c[1][0][0], in.iso) + crosses(d0, k.c[0][1][0], in.iso) +
          crosses(d0, k.c[0][0][1], in.iso);
    if (ALGO == DISO_MC) {
        *nf = DISO_MC_NTRI(cases[mc_code(k, in.iso)]);
    } else {
        int code = dmc_code(k, in.iso);
        if (code != 0 && code != 255) code = dmc_good_code(in, cases, code, px, py, pz);
        *nf = DISO_DMC_NPATCH(cases[code]);
    }
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

template <typename S, int ALGO>
__global__ __launch_bounds__(kCountBlock) void count_kernel(In<S> in, SegBase *seg, uint8_t *ctl,
                                                            int64_t nblocks) {
    __shared__ uint64_t s_case[256];
    __shared__ int64_t s_v[kCountWaves], s_f[kCountWaves];
    __shared__ uint32_t s_fl[kCountWaves];
    __shared__ uint32_t s_parity, s_dirty_other;

    CtlHeader *hdr = reinterpret_cast<CtlHeader *>(ctl);
    uint64_t *words = reinterpret_cast<uint64_t *>(ctl + kCtlHeader);
    const int tid = threadIdx.x, wave = tid / kWave, lane = tid % kWave;
    const Geo &g = in.geo;
    const int64_t b = blockIdx.x;
    const int64_t lo = b * g.nseg / nblocks, hi = (b + 1) * g.nseg / nblocks;

    for (int i = tid; i < 256; i += kCountBlock) s_case[i] = ALGO == DISO_MC ? kMcCase[i] : kDmcCase[i];
    if (tid == 0) {
        s_parity = __hip_atomic_load(&hdr->parity, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        s_dirty_other =
            __hip_atomic_load(&hdr->dirty[s_parity ^ 1u], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    const uint32_t parity = s_parity;
    uint64_t *cur = words + (int64_t)parity * kMaxCountBlocks;

    // 1. Count: wave w takes segments lo+w, lo+w+8, ...; lane 0 records each segment's counts.
    int64_t wv = 0, wf = 0;
    uint32_t wfl = 0;
    for (int64_t s = lo + wave; s < hi; s += kCountWaves) {  // wave-uniform
        const int chunk = (int)(s % g.nchunk);
        const int64_t row = s / g.nchunk;
        int nv, nf;
        uint32_t fl;
        count_position<S, ALGO>(in, s_case, (int)(row / g.g1), (int)(row % g.g1), chunk * kWave + lane, &nv,
                                &nf, &fl);
        int tv, tf;
        (void)wave_excl_small(nv, &tv);
        (void)wave_excl_small(nf, &tf);
        if (lane == 0) {
            seg[s].v = tv;
            seg[s].f = tf;
        }
        wv += tv;
        wf += tf;
        wfl |= fl;
    }
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
    if (tid == 0) st_relaxed(cur + b, pack_agg(av, af, afl));

    // 2. Exclusive prefix of this block = sum of all predecessors' aggregates.
    int64_t pv = 0, pf = 0;
    uint32_t pfl = 0;
    bool timeout = false;
    for (int64_t j = tid; j < b; j += kCountBlock) {
        uint64_t w;
        uint32_t spins = 0;
        do {
            w = ld_relaxed(cur + j);
        } while ((w >> 62) == 0 && ++spins < kSpinLimit);
        timeout |= (w >> 62) == 0;
        pv += (int64_t)(w & ((1ull << 29) - 1));
        pf += (int64_t)((w >> 29) & ((1ull << 30) - 1));
        pfl |= (uint32_t)((w >> 59) & 7);
    }
    pv = wave_sum64(pv);
    pf = wave_sum64(pf);
    pfl = wave_or(pfl);
    const bool wave_timeout = __ballot(timeout) != 0;
    __syncthreads();  // s_v / s_f reuse
    if (lane == 0) {
        s_v[wave] = pv;
        s_f[wave] = pf;
        s_fl[wave] = pfl | (wave_timeout ? DISO_TOTAL_SCAN_TIMEOUT : 0u);
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
    if (b == nblocks - 1 && tid == 0) {
        SegBase *tot = seg + g.nseg;
        tot[0].v = ev + av;
        tot[0].f = ef + af;
        tot[1].v = (int64_t)(efl | afl);
        tot[1].f = ALGO;
        // Every block has read the parity (it published a word, which this block waited for).
        __hip_atomic_store(&hdr->dirty[parity], (uint32_t)nblocks, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&hdr->dirty[parity ^ 1u], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&hdr->parity, parity ^ 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }

    // 3. Clear this block's share of the other word buffer for the next call.
    {
        const int64_t n = s_dirty_other;
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

struct OutCfg {
    int flags;
    int64_t vcap, fcap;  // capacities in rows
};

template <typename S>
__device__ __forceinline__ void store_vertex(S *verts, int64_t i, V3<S> p, int flags, S nx, S ny,
                                             S nz) {
    if (flags & DISO_FLAG_SHIFT) {  // verts - 1 (diso/__init__.py:56)
        p.x = p.x - S(1);
        p.y = p.y - S(1);
        p.z = p.z - S(1);
    }
    if (flags & DISO_FLAG_NORMALIZE) {  // verts / (dims - 1) (diso/__init__.py:57-60)
        p.x = p.x / nx;
        p.y = p.y / ny;
        p.z = p.z / nz;
    }
    verts[3 * i + 0] = p.x;
    verts[3 * i + 1] = p.y;
    verts[3 * i + 2] = p.z;
}

__device__ __forceinline__ void store_index(void *faces, int64_t i, int64_t v, int flags) {
    if (flags & DISO_FLAG_INDEX64)
        reinterpret_cast<int64_t *>(faces)[i] = v;
    else
        reinterpret_cast<int32_t *>(faces)[i] = (int32_t)v;
}

// MC: vertices of the owned crossing edges in axis order, then triangles. A triangle corner on
// table edge e lies on an edge owned by the cell at (cell + DISO_EDGE_OFS[e]); its vertex id is
// that owner's first vertex plus the rank of the edge's axis among the owner's crossing axes
// (what create_cell_mc_tris_kernel finds by linear search, src/cumc.cu:586-608). Owners are
// the cell itself, its +x, +y, +x+y neighbours (rows R00, R10, R01, R11 of this chunk) and their
// +z neighbours, whose first vertex is the current one's first vertex plus its count.
template <typename S, bool DEF>
__global__ __launch_bounds__(kBlock) void emit_mc_kernel(In<S> in, const SegBase *seg, S *verts,
                                                         void *faces, OutCfg out, S nx, S ny, S nz,
                                                         int64_t nblocks) {
    __shared__ uint64_t s_case[256];
    const int tid = threadIdx.x, wave = tid / kWave, lane = tid % kWave;
    for (int i = tid; i < 256; i += kBlock) s_case[i] = kMcCase[i];
    __syncthreads();
    const Geo &g = in.geo;
    const int64_t s = xcd_block(blockIdx.x, nblocks) * kWavesPerBlock + wave;
    if (s >= g.nseg) return;
    const int chunk = (int)(s % g.nchunk);
    const int64_t row = s / g.nchunk;
    const int px = (int)(row / g.g1), py = (int)(row % g.g1);
    if (px >= g.g0 - 1 || py >= g.g1 - 1) return;  // no cells on this row (wave-uniform)
    const int pz = chunk * kWave + lane;
    const SegBase *tot = seg + g.nseg;
    const bool fits = tot[0].v <= out.vcap && tot[0].f <= out.fcap;

    const int m00 = owned_crossings(in, px, py, pz);
    const int m10 = owned_crossings(in, px + 1, py, pz);
    const int m01 = owned_crossings(in, px, py + 1, pz);
    const int m11 = owned_crossings(in, px + 1, py + 1, pz);
    int t;
    const int64_t vb00 = seg[s].v + wave_excl_small(popc3(m00), &t);
    const int64_t vb10 = seg[seg_index(g, px + 1, py, chunk)].v + wave_excl_small(popc3(m10), &t);
    const int64_t vb01 = seg[seg_index(g, px, py + 1, chunk)].v + wave_excl_small(popc3(m01), &t);
    const int64_t vb11 = seg[seg_index(g, px + 1, py + 1, chunk)].v + wave_excl_small(popc3(m11), &t);
    const int n00 = popc3(m00), n10 = popc3(m10), n01 = popc3(m01), n11 = popc3(m11);

    int ntri = 0;
    uint64_t w = 0;
    if (is_anchor(in, px, py, pz)) {
        w = s_case[mc_code(load_corners(in, px, py, pz), in.iso)];
        ntri = DISO_MC_NTRI(w);
    }
    const int64_t tb = seg[s].f + wave_excl_small(ntri, &t);
    if (!fits) return;

    // vertices
    {
        int64_t v = vb00;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            if ((m00 >> a) & 1) {
                if (v < out.vcap) store_vertex(verts, v, edge_vertex<S, DEF>(in, px, py, pz, a), out.flags, nx, ny, nz);
                ++v;
            }
        }
    }
    // triangles
    if (ntri) {
        const int c00n = owned_crossings(in, px, py, pz + 1) & 1;
        const int c10n = owned_crossings(in, px + 1, py, pz + 1) & 1;
        for (int i = 0; i < 3 * ntri; ++i) {
            int64_t id;
            switch (DISO_MC_EDGE(w, i)) {
                case 0: id = vb00; break;
                case 1: id = vb10 + n10 - 1; break;
                case 2: id = vb00 + n00; break;
                case 3: id = vb00 + n00 - 1; break;
                case 4: id = vb01; break;
                case 5: id = vb11 + n11 - 1; break;
                case 6: id = vb01 + n01; break;
                case 7: id = vb01 + n01 - 1; break;
                case 8: id = vb00 + (m00 & 1); break;
                case 9: id = vb10 + (m10 & 1); break;
                case 10: id = vb10 + n10 + c10n; break;
                default: id = vb00 + n00 + c00n; break;
            }
            int64_t o = 3 * tb + i;
            if (o < 3 * out.fcap) store_index(faces, o, id, out.flags);
        }
    }
}

// DMC: patch vertices of the cell (mean of the patch's edge vertices, src/cudualmc.cu:907-955)
// and one quad per owned crossing edge (src/cudualmc.cu:1027-1056). A quad corner is the vertex
// of the patch that holds the edge in one of the four cells around it; those cells are the
// cell itself and its -x, -y, -z neighbours (rows R00, Rm0, R0m, Rmm at this z and z-1).
struct DmcCell {
    int code;      // resolved code (0 when not a cell)
    int64_t base;  // first DMC vertex of the cell
};

template <typename S>
__device__ __forceinline__ int64_t dmc_row_base(const SegBase *seg, const Geo &g, int px, int py,
                                                int chunk) {
    return (px >= 0 && py >= 0) ? seg[seg_index(g, px, py, chunk)].f : 0;
}

// Resolved codes and first-vertex ids of row (px,py) at this lane's z and at z-1.
template <typename S>
__device__ __forceinline__ void dmc_row(const In<S> &in, const uint64_t *cases, const SegBase *seg,
                                        int px, int py, int chunk, int pz, DmcCell *cur,
                                        DmcCell *prev) {
    const int code = dmc_cell(in, cases, px, py, pz);
    const int np = DISO_DMC_NPATCH(cases[code]);
    int t;
    cur->code = code;
    cur->base = dmc_row_base<S>(seg, in.geo, px, py, chunk) + wave_excl_small(np, &t);
    int pcode = __shfl_up(code, 1, kWave);
    if (lane_id() == 0) pcode = dmc_cell(in, cases, px, py, pz - 1);
    prev->code = pcode;
    prev->base = cur->base - DISO_DMC_NPATCH(cases[pcode]);
}

__device__ __forceinline__ int64_t dmc_vertex_id(const uint64_t *cases, const DmcCell &c, int eid) {
    int p = DISO_DMC_PATCH_OF(cases[c.code], eid);
    return c.base + (p == 7 ? -1 : p);  // dmcEdgeOffset -1 for a non-crossing edge
}

template <typename S, bool DEF>
__global__ __launch_bounds__(kBlock) void emit_dmc_kernel(In<S> in, const SegBase *seg, S *verts,
                                                          void *faces, OutCfg out, S nx, S ny, S nz,
                                                          int64_t nblocks) {
    __shared__ uint64_t s_case[256];
    const int tid = threadIdx.x, wave = tid / kWave, lane = tid % kWave;
    for (int i = tid; i < 256; i += kBlock) s_case[i] = kDmcCase[i];
    __syncthreads();
    const Geo &g = in.geo;
    const int64_t s = xcd_block(blockIdx.x, nblocks) * kWavesPerBlock + wave;
    if (s >= g.nseg) return;
    const int chunk = (int)(s % g.nchunk);
    const int64_t row = s / g.nchunk;
    const int px = (int)(row / g.g1), py = (int)(row % g.g1);
    if (px >= g.g0 - 1 || py >= g.g1 - 1) return;
    const int pz = chunk * kWave + lane;
    const SegBase *tot = seg + g.nseg;
    const bool fits = tot[0].v <= out.fcap && tot[0].f <= out.vcap;

    DmcCell c00, c00m, c0m, c0mm, cm0, cm0m, cmm, cmmm;
    dmc_row(in, s_case, seg, px, py, chunk, pz, &c00, &c00m);
    dmc_row(in, s_case, seg, px, py - 1, chunk, pz, &c0m, &c0mm);
    dmc_row(in, s_case, seg, px - 1, py, chunk, pz, &cm0, &cm0m);
    dmc_row(in, s_case, seg, px - 1, py - 1, chunk, pz, &cmm, &cmmm);
    (void)cmmm;
    const int m00 = owned_crossings(in, px, py, pz);
    int t;
    const int64_t qb = seg[s].v + wave_excl_small(popc3(m00), &t);
    if (!fits) return;

    // patch vertices
    const uint64_t w = s_case[c00.code];
    const int np = DISO_DMC_NPATCH(w);
    if (np) {
        V3<S> ev[12];
#pragma unroll
        for (int e = 0; e < 12; ++e) {
            if (DISO_DMC_PATCH_OF(w, e) != 7)
                ev[e] = edge_vertex<S, DEF>(in, px + kEdgeOfs[e][0], py + kEdgeOfs[e][1],
                                           pz + kEdgeOfs[e][2], kEdgeAxis[e]);
        }
        for (int k = 0; k < np; ++k) {
            S sx = S(0), sy = S(0), sz = S(0), num = S(0);
#pragma unroll
            for (int e = 0; e < 12; ++e) {
                if (DISO_DMC_PATCH_OF(w, e) == k) {
                    sx += ev[e].x;
                    sy += ev[e].y;
                    sz += ev[e].z;
                    num += S(1);
                }
            }
            S r = S(1) / num;
            V3<S> p = {sx * r, sy * r, sz * r};
            int64_t v = c00.base + k;
            if (v < out.vcap) store_vertex(verts, v, p, out.flags, nx, ny, nz);
        }
    }
    // quads
    if (m00) {
        const S d0 = sample(in, px, py, pz);
        int64_t q = qb;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            if (!((m00 >> a) & 1)) continue;
            const S d1 = sample(in, px + (a == 0), py + (a == 1), pz + (a == 2));
            const bool exiting = d1 < in.iso && d0 >= in.iso;  // type a + 3 (src/cudualmc.cu:782-788)
            int64_t id[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                int64_t v[2];
#pragma unroll
                for (int o = 0; o < 2; ++o) {
                    const int *q4 = kQuad[a + 3 * o][k];
                    const int sel = (-q4[0]) * 4 + (-q4[1]) * 2 + (-q4[2]);
                    const DmcCell &c = sel == 0 ? c00 : sel == 1 ? c00m : sel == 2 ? c0m : sel == 3 ? c0mm
                                     : sel == 4 ? cm0 : sel == 5 ? cm0m : cmm;
                    v[o] = dmc_vertex_id(s_case, c, q4[3]);
                }
                id[k] = exiting ? v[1] : v[0];
            }
            if (q < out.fcap) {
#pragma unroll
                for (int k = 0; k < 4; ++k) store_index(faces, 4 * q + k, id[k], out.flags);
            }
            ++q;
        }
    }
}

// ------------------------------------------------------------------------------------------
// Backward (gather, deterministic)

template <typename S>
struct Grad {
    S g, dx, dy, dz;
};

// Contribution of one crossing edge (v0 -> v0 + e_axis, upstream gradient gv) to the endpoint
// `end` (0 = v0, 1 = v1): adjComputeMcVert (src/cumc.cu:412-453).
template <typename S, bool DEF>
__device__ __forceinline__ void edge_adjoint(const In<S> &in, int x, int y, int z, int axis, int end,
                                             V3<S> gv, Grad<S> *acc) {
    const int x1 = x + (axis == 0), y1 = y + (axis == 1), z1 = z + (axis == 2);
    const S d0 = sample(in, x, y, z), d1 = sample(in, x1, y1, z1);
    const S t = interp_t(d0, d1, in.iso);
    S p0x = S(x), p0y = S(y), p0z = S(z);
    S p1x = S(x1), p1y = S(y1), p1z = S(z1);
    if (DEF) {
        V3<S> a = deform_at<S, DEF>(in, x, y, z), b = deform_at<S, DEF>(in, x1, y1, z1);
        p0x += a.x;
        p0y += a.y;
        p0z += a.z;
        p1x += b.x;
        p1y += b.y;
        p1z += b.z;
    }
    const S adj_t = (p1x - p0x) * gv.x + (p1y - p0y) * gv.y + (p1z - p0z) * gv.z;
    const S dd = (d1 - d0) * (d1 - d0);
    if (end == 0) {
        acc->g += (in.iso - d1) / dd * adj_t;
        const S w = S(1) - t;
        acc->dx += w * gv.x;
        acc->dy += w * gv.y;
        acc->dz += w * gv.z;
    } else {
        acc->g += (d0 - in.iso) / dd * adj_t;
        acc->dx += t * gv.x;
        acc->dy += t * gv.y;
        acc->dz += t * gv.z;
    }
}

template <typename S>
struct AdjIn {
    const S *adj;  // [n_verts, 3]
    int64_t n;
    int flags;
    S nx, ny, nz;  // upstream scaling by 1/(dims-1) when the forward normalised
};

template <typename S>
__device__ __forceinline__ V3<S> load_adj(const AdjIn<S> &a, int64_t i) {
    V3<S> r = {S(0), S(0), S(0)};
    if (i >= 0 && i < a.n) {
        r.x = a.adj[3 * i + 0];
        r.y = a.adj[3 * i + 1];
        r.z = a.adj[3 * i + 2];
        if (a.flags & DISO_FLAG_NORMALIZE) {  // grad of verts / (dims - 1)
            r.x = r.x / a.nx;
            r.y = r.y / a.ny;
            r.z = r.z / a.nz;
        }
    }
    return r;
}

template <typename S>
__device__ __forceinline__ void store_grad(const In<S> &in, S *adj_sdf, S *adj_deform, int x, int y,
                                           int z, Grad<S> acc, int flags) {
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
        adj_sdf[i] = acc.g;
        if (adj_deform) {
            adj_deform[3 * i + 0] = acc.dx;
            adj_deform[3 * i + 1] = acc.dy;
            adj_deform[3 * i + 2] = acc.dz;
        }
    }
}

// MC: the grid vertex v gathers the edges v->v+e_a (owned by v, ids base(v)+rank) and
// v-e_a->v (owned by v-e_a: the -x owner's x edge is its first vertex, the -y owner's y edge is
// its first vertex plus its x flag, the -z owner's z edge is the vertex just before base(v)).
template <typename S, bool DEF>
__global__ __launch_bounds__(kBlock) void bwd_mc_kernel(In<S> in, const SegBase *seg, AdjIn<S> adj,
                                                        S *adj_sdf, S *adj_deform, int64_t nblocks) {
    const int tid = threadIdx.x, wave = tid / kWave, lane = tid % kWave;
    const Geo &g = in.geo;
    const int64_t s = xcd_block(blockIdx.x, nblocks) * kWavesPerBlock + wave;
    if (s >= g.nseg) return;
    const int chunk = (int)(s % g.nchunk);
    const int64_t row = s / g.nchunk;
    const int px = (int)(row / g.g1), py = (int)(row % g.g1);
    const int pz = chunk * kWave + lane;

    const int m00 = owned_crossings(in, px, py, pz);
    const int mm0 = owned_crossings(in, px - 1, py, pz);
    const int m0m = owned_crossings(in, px, py - 1, pz);
    const int m00m = owned_crossings(in, px, py, pz - 1);
    int t;
    const int64_t vb00 = seg[s].v + wave_excl_small(popc3(m00), &t);
    const int64_t vbm0 = (px > 0 ? seg[seg_index(g, px - 1, py, chunk)].v : 0) + wave_excl_small(popc3(mm0), &t);
    const int64_t vb0m = (py > 0 ? seg[seg_index(g, px, py - 1, chunk)].v : 0) + wave_excl_small(popc3(m0m), &t);
    if (!in_data(in, px, py, pz)) return;

    Grad<S> acc = {S(0), S(0), S(0), S(0)};
    int r = 0;
    if (m00 & 1) edge_adjoint<S, DEF>(in, px, py, pz, 0, 0, load_adj(adj, vb00 + r++), &acc);
    if (m00 & 2) edge_adjoint<S, DEF>(in, px, py, pz, 1, 0, load_adj(adj, vb00 + r++), &acc);
    if (m00 & 4) edge_adjoint<S, DEF>(in, px, py, pz, 2, 0, load_adj(adj, vb00 + r++), &acc);
    if (mm0 & 1) edge_adjoint<S, DEF>(in, px - 1, py, pz, 0, 1, load_adj(adj, vbm0), &acc);
    if (m0m & 2) edge_adjoint<S, DEF>(in, px, py - 1, pz, 1, 1, load_adj(adj, vb0m + (m0m & 1)), &acc);
    if (m00m & 4) edge_adjoint<S, DEF>(in, px, py, pz - 1, 2, 1, load_adj(adj, vb00 - 1), &acc);
    store_grad(in, adj_sdf, adj_deform, px, py, pz, acc, adj.flags);
}

// DMC: as MC, but an edge's upstream gradient is the sum over the (up to four) cells around it
// of (gradient of the patch vertex holding the edge) / (edges in that patch), i.e. the adjoint
// of create_dmc_verts_kernel's mean. With DISO_FLAG_REFERENCE_COMPAT each cell contributes its
// first patch's gradient instead (src/cudualmc.cu:990).
template <typename S>
__device__ __forceinline__ S recip_count(int n) {
    switch (n) {  // S(1)/S(n), correctly rounded
        case 1: return S(1) / S(1);
        case 2: return S(1) / S(2);
        case 3: return S(1) / S(3);
        case 4: return S(1) / S(4);
        case 5: return S(1) / S(5);
        case 6: return S(1) / S(6);
        default: return S(1) / S(7);
    }
}

__device__ __forceinline__ int patch_size(uint64_t w, int p) {
    int n = 0;
#pragma unroll
    for (int e = 0; e < 12; ++e) n += DISO_DMC_PATCH_OF(w, e) == p;
    return n;
}

template <typename S>
__device__ __forceinline__ V3<S> dmc_cell_adjoint(const uint64_t *cases, const AdjIn<S> &adj,
                                                  const DmcCell &c, int eid, bool compat) {
    const uint64_t w = cases[c.code];
    const int p = DISO_DMC_PATCH_OF(w, eid);
    V3<S> r = {S(0), S(0), S(0)};
    if (p == 7) return r;
    V3<S> a = load_adj(adj, compat ? c.base : c.base + p);
    const S k = recip_count<S>(patch_size(w, p));
    r.x = a.x * k;
    r.y = a.y * k;
    r.z = a.z * k;
    return r;
}

template <typename S>
__device__ __forceinline__ V3<S> add4(V3<S> a, V3<S> b, V3<S> c, V3<S> d) {
    V3<S> r;
    r.x = ((a.x + b.x) + c.x) + d.x;
    r.y = ((a.y + b.y) + c.y) + d.y;
    r.z = ((a.z + b.z) + c.z) + d.z;
    return r;
}

template <typename S, bool DEF>
__global__ __launch_bounds__(kBlock) void bwd_dmc_kernel(In<S> in, const SegBase *seg, AdjIn<S> adj,
                                                         S *adj_sdf, S *adj_deform, int64_t nblocks) {
    __shared__ uint64_t s_case[256];
    const int tid = threadIdx.x, wave = tid / kWave, lane = tid % kWave;
    for (int i = tid; i < 256; i += kBlock) s_case[i] = kDmcCase[i];
    __syncthreads();
    const Geo &g = in.geo;
    const int64_t s = xcd_block(blockIdx.x, nblocks) * kWavesPerBlock + wave;
    if (s >= g.nseg) return;
    const int chunk = (int)(s % g.nchunk);
    const int64_t row = s / g.nchunk;
    const int px = (int)(row / g.g1), py = (int)(row % g.g1);
    const int pz = chunk * kWave + lane;
    const bool compat = (adj.flags & DISO_FLAG_REFERENCE_COMPAT) != 0;

    DmcCell c00, c00m, c0m, c0mm, cm0, cm0m, cmm, cmmm;
    dmc_row(in, s_case, seg, px, py, chunk, pz, &c00, &c00m);
    dmc_row(in, s_case, seg, px, py - 1, chunk, pz, &c0m, &c0mm);
    dmc_row(in, s_case, seg, px - 1, py, chunk, pz, &cm0, &cm0m);
    dmc_row(in, s_case, seg, px - 1, py - 1, chunk, pz, &cmm, &cmmm);
    if (!in_data(in, px, py, pz)) return;

    const int m00 = owned_crossings(in, px, py, pz);
    const int mm0 = owned_crossings(in, px - 1, py, pz);
    const int m0m = owned_crossings(in, px, py - 1, pz);
    const int m00m = owned_crossings(in, px, py, pz - 1);
    Grad<S> acc = {S(0), S(0), S(0), S(0)};
#define DMC_EDGE_GRAD(C0, E0, C1, E1, C2, E2, C3, E3)                                               \
    add4(dmc_cell_adjoint(s_case, adj, C0, E0, compat), dmc_cell_adjoint(s_case, adj, C1, E1, compat), \
         dmc_cell_adjoint(s_case, adj, C2, E2, compat), dmc_cell_adjoint(s_case, adj, C3, E3, compat))
    // Cells around each incident edge, in the order of the reference's quad stencil for the
    // edge's axis (src/cudualmc.cu:504-507).
    if (m00 & 1) edge_adjoint<S, DEF>(in, px, py, pz, 0, 0, DMC_EDGE_GRAD(c00, 0, c0m, 4, c0mm, 6, c00m, 2), &acc);
    if (m00 & 2) edge_adjoint<S, DEF>(in, px, py, pz, 1, 0, DMC_EDGE_GRAD(c00, 8, c00m, 11, cm0m, 10, cm0, 9), &acc);
    if (m00 & 4) edge_adjoint<S, DEF>(in, px, py, pz, 2, 0, DMC_EDGE_GRAD(c00, 3, cm0, 1, cmm, 5, c0m, 7), &acc);
    if (mm0 & 1) edge_adjoint<S, DEF>(in, px - 1, py, pz, 0, 1, DMC_EDGE_GRAD(cm0, 0, cmm, 4, cmmm, 6, cm0m, 2), &acc);
    if (m0m & 2) edge_adjoint<S, DEF>(in, px, py - 1, pz, 1, 1, DMC_EDGE_GRAD(c0m, 8, c0mm, 11, cmmm, 10, cmm, 9), &acc);
    if (m00m & 4) edge_adjoint<S, DEF>(in, px, py, pz - 1, 2, 1, DMC_EDGE_GRAD(c00m, 3, cm0m, 1, cmmm, 5, c0mm, 7), &acc);
#undef DMC_EDGE_GRAD
    store_grad(in, adj_sdf, adj_deform, px, py, pz, acc, adj.flags);
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
    // Per-block aggregates must fit the 29/30-bit fields of the scan words.
    return (int64_t)g->g0 * g->g1 * g->g2 < (1ll << 36);
}

int64_t seg_bytes_for(const Geo &g) { return (g.nseg + 2) * (int64_t)sizeof(SegBase); }
int64_t ctl_bytes_for(const Geo &) { return kCtlHeader + 2 * kMaxCountBlocks * (int64_t)sizeof(uint64_t); }

// Number of count blocks: about kCountBlocksPerCU per CU of the current device, so that every
// block is resident at once, and never more than the segments or the word buffer allow.
int64_t count_blocks(const Geo &g) {
    static int cus[64] = {0};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
    if (cus[dev] == 0) {
        int n = 0;
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) n = 256;
        cus[dev] = n;
    }
    int64_t nb = (int64_t)cus[dev] * kCountBlocksPerCU;
    const int64_t min_nb = (g.nseg + (1ll << 20) - 1) >> 20;  // <= 2^20 segments per block keeps the
    if (nb < min_nb) nb = min_nb;                            // aggregate fields from overflowing
    if (nb > kMaxCountBlocks) nb = kMaxCountBlocks;
    if (nb > g.nseg) nb = g.nseg;
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
    In<S> in = make_in<S>(sdf, nullptr, iso, g);
    const int64_t nb = count_blocks(g);
    (void)ctl_bytes;
    if (algo == DISO_MC)
        hipLaunchKernelGGL((count_kernel<S, DISO_MC>), dim3((unsigned)nb), dim3(kCountBlock), 0, st, in,
                           (SegBase *)seg, (uint8_t *)ctl, nb);
    else
        hipLaunchKernelGGL((count_kernel<S, DISO_DMC>), dim3((unsigned)nb), dim3(kCountBlock), 0, st, in,
                           (SegBase *)seg, (uint8_t *)ctl, nb);
    return check_hip(hipGetLastError(), "count kernel launch");
}

int64_t row_blocks(const Geo &g) { return (g.nseg + kWavesPerBlock - 1) / kWavesPerBlock; }

template <typename S>
void norm_of(const Geo &g, S *nx, S *ny, S *nz) {
    *nx = (S)g.n0 - S(1);
    *ny = (S)g.n1 - S(1);
    *nz = (S)g.n2 - S(1);
}

template <typename S>
int launch_emit(int algo, const void *sdf, const void *deform, double iso, int flags, const Geo &g,
                const void *seg, void *verts, int64_t vcap, void *faces, int64_t fcap,
                hipStream_t st) {
    In<S> in = make_in<S>(sdf, deform, iso, g);
    OutCfg out{flags, vcap, fcap};
    S nx, ny, nz;
    norm_of(g, &nx, &ny, &nz);
    const int64_t nb = row_blocks(g);
    const SegBase *sb = (const SegBase *)seg;
    if (algo == DISO_MC) {
        if (deform)
            hipLaunchKernelGGL((emit_mc_kernel<S, true>), dim3((unsigned)nb), dim3(kBlock), 0, st, in, sb,
                               (S *)verts, faces, out, nx, ny, nz, nb);
        else
            hipLaunchKernelGGL((emit_mc_kernel<S, false>), dim3((unsigned)nb), dim3(kBlock), 0, st, in, sb,
                               (S *)verts, faces, out, nx, ny, nz, nb);
    } else {
        if (deform)
            hipLaunchKernelGGL((emit_dmc_kernel<S, true>), dim3((unsigned)nb), dim3(kBlock), 0, st, in, sb,
                               (S *)verts, faces, out, nx, ny, nz, nb);
        else
            hipLaunchKernelGGL((emit_dmc_kernel<S, false>), dim3((unsigned)nb), dim3(kBlock), 0, st, in, sb,
                               (S *)verts, faces, out, nx, ny, nz, nb);
    }
    return check_hip(hipGetLastError(), "emit kernel launch");
}

template <typename S>
int launch_backward(int algo, const void *sdf, const void *deform, double iso, int flags,
                    const Geo &g, const void *seg, const void *adj_verts, int64_t nverts,
                    void *adj_sdf, void *adj_deform, hipStream_t st) {
    In<S> in = make_in<S>(sdf, deform, iso, g);
    AdjIn<S> a;
    a.adj = (const S *)adj_verts;
    a.n = nverts;
    a.flags = flags;
    norm_of(g, &a.nx, &a.ny, &a.nz);
    const int64_t nb = row_blocks(g);
    const SegBase *sb = (const SegBase *)seg;
    S *gs = (S *)adj_sdf, *gd = (S *)(deform ? adj_deform : nullptr);
    if (algo == DISO_MC) {
        if (deform)
            hipLaunchKernelGGL((bwd_mc_kernel<S, true>), dim3((unsigned)nb), dim3(kBlock), 0, st, in, sb, a, gs, gd, nb);
        else
            hipLaunchKernelGGL((bwd_mc_kernel<S, false>), dim3((unsigned)nb), dim3(kBlock), 0, st, in, sb, a, gs, gd, nb);
    } else {
        if (deform)
            hipLaunchKernelGGL((bwd_dmc_kernel<S, true>), dim3((unsigned)nb), dim3(kBlock), 0, st, in, sb, a, gs, gd, nb);
        else
            hipLaunchKernelGGL((bwd_dmc_kernel<S, false>), dim3((unsigned)nb), dim3(kBlock), 0, st, in, sb, a, gs, gd, nb);
    }
    return check_hip(hipGetLastError(), "backward kernel launch");
}

int check_common(int algo, int dtype, int64_t X, int64_t Y, int64_t Z, int flags, Geo *g) {
    if (algo != DISO_MC && algo != DISO_DMC) return fail(DISO_E_INVALID, "algo must be DISO_MC or DISO_DMC");
    if (dtype != DISO_F32 && dtype != DISO_F64) return fail(DISO_E_INVALID, "dtype must be DISO_F32 or DISO_F64");
    if (!make_geo(X, Y, Z, flags, g))
        return fail(DISO_E_INVALID, "unsupported grid dims (%lld, %lld, %lld)", (long long)X, (long long)Y,
                    (long long)Z);
    return DISO_OK;
}

int read_totals(const Geo &g, const void *seg, diso_totals_t *t, hipStream_t st) {
    SegBase tot[2];
    int rc = check_hip(hipMemcpyAsync(tot, (const SegBase *)seg + g.nseg, sizeof(tot), hipMemcpyDeviceToHost, st),
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

}  // namespace

// ------------------------------------------------------------------------------------------
// C ABI

extern "C" {

const char *diso_last_error(void) { return g_err; }

const char *diso_version(void) { return "diso_hip 0.1.0 gfx950"; }

int64_t diso_seg_bytes(int64_t X, int64_t Y, int64_t Z, int flags) {
    Geo g;
    if (!make_geo(X, Y, Z, flags, &g)) return fail(DISO_E_INVALID, "unsupported grid dims");
    return seg_bytes_for(g);
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

int diso_read_totals(const void *seg, int64_t seg_bytes, int64_t X, int64_t Y, int64_t Z, int flags,
                     diso_totals_t *totals_host, void *stream) {
    g_err[0] = 0;
    Geo g;
    if (!make_geo(X, Y, Z, flags, &g)) return fail(DISO_E_INVALID, "unsupported grid dims");
    if (!seg || !totals_host) return fail(DISO_E_INVALID, "null pointer argument");
    if (seg_bytes < seg_bytes_for(g)) return fail(DISO_E_CAPACITY, "seg workspace too small");
    return read_totals(g, seg, totals_host, (hipStream_t)stream);
}

int diso_emit(int algo, int dtype, const void *sdf, const void *deform, int64_t X, int64_t Y,
              int64_t Z, double iso, int flags, const void *seg, int64_t seg_bytes, void *verts,
              int64_t vert_capacity, void *faces, int64_t face_capacity, void *stream) {
    g_err[0] = 0;
    Geo g;
    int rc = check_common(algo, dtype, X, Y, Z, flags, &g);
    if (rc) return rc;
    if (!sdf || !seg) return fail(DISO_E_INVALID, "null pointer argument");
    if (seg_bytes < seg_bytes_for(g)) return fail(DISO_E_CAPACITY, "seg workspace too small");
    if (vert_capacity < 0 || face_capacity < 0) return fail(DISO_E_INVALID, "negative capacity");
    if ((vert_capacity && !verts) || (face_capacity && !faces)) return fail(DISO_E_INVALID, "null output buffer");
    hipStream_t st = (hipStream_t)stream;
    return dtype == DISO_F32
               ? launch_emit<float>(algo, sdf, deform, iso, flags, g, seg, verts, vert_capacity, faces,
                                    face_capacity, st)
               : launch_emit<double>(algo, sdf, deform, iso, flags, g, seg, verts, vert_capacity, faces,
                                     face_capacity, st);
}

int diso_backward(int algo, int dtype, const void *sdf, const void *deform, int64_t X, int64_t Y,
                  int64_t Z, double iso, int flags, const void *seg, int64_t seg_bytes,
                  const void *adj_verts, int64_t n_verts, void *adj_sdf, void *adj_deform,
                  void *stream) {
    g_err[0] = 0;
    Geo g;
    int rc = check_common(algo, dtype, X, Y, Z, flags, &g);
    if (rc) return rc;
    if (!sdf || !seg || !adj_sdf) return fail(DISO_E_INVALID, "null pointer argument");
    if (deform && !adj_deform) return fail(DISO_E_INVALID, "adj_deform required when deform is given");
    if (n_verts < 0 || (n_verts && !adj_verts)) return fail(DISO_E_INVALID, "bad adj_verts");
    if (seg_bytes < seg_bytes_for(g)) return fail(DISO_E_CAPACITY, "seg workspace too small");
    hipStream_t st = (hipStream_t)stream;
    return dtype == DISO_F32 ? launch_backward<float>(algo, sdf, deform, iso, flags, g, seg, adj_verts,
                                                      n_verts, adj_sdf, adj_deform, st)
                             : launch_backward<double>(algo, sdf, deform, iso, flags, g, seg, adj_verts,
                                                       n_verts, adj_sdf, adj_deform, st);
}

}  // extern "C"
