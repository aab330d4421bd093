// kernels.hpp -- gfx950 device code of the implicit FEM step.
//
// Layout in HBM (one system):
//   CRS  ptr[n+1], col[nnz] int32; A[nnz] f64 (A32[nnz] f32 in IR mode)
//   vectors f64[n] (f32[n] for the IR inner solve), Jacobi diagonal d[n]
//   elements: conn[E*npe] int32, colour-ordered element ids, slot[E*npe*npe]
//   int32 (CRS position of local entry (i,j)), node coordinates f64 SoA.
//
// n-sized kernels split the rows into G contiguous chunks (G a multiple of
// 8); block b works on chunk (b % 8) * (G / 8) + b / 8, so the chunks that
// one XCD receives (blocks are dealt to the 8 XCDs round-robin) are adjacent
// and the x entries a 27-point row gathers stay in that XCD's L2.
//
// Reductions are deterministic and atomic-free: every block writes one
// partial per dot product; the kernel that consumes a dot reduces all G
// partials in a fixed order (each block redundantly), so results depend only
// on G and the data, never on scheduling.
#pragma once

#include <float.h>

#include <type_traits>

#include "ogsx_internal.h"

namespace ogsx {

// Device pointers read from descriptors in memory (not kernel arguments) are
// generic to the compiler, which then emits flat loads/stores.  Kernels
// re-type them as global-address-space pointers (gl) so they get
// global_load/global_store; the hot loops measured ~40% slower with flat.
#define OGSX_GL __attribute__((address_space(1)))

template <typename T>
__device__ __forceinline__ OGSX_GL T* gl(T* p)
{
    return (OGSX_GL T*)p;
}

// ------------------------------------------------------------------------
// generic helpers
// ------------------------------------------------------------------------
// Block b of G (G a multiple of 8) works on chunk (b % 8) * (G / 8) + b / 8
// of n items.  In a group launch, b and G are the block's index and count
// within its own system.
__device__ __forceinline__ void chunk_of(int n, int G, int b, int& beg, int& end)
{
    const int c = (b % kNumXcd) * (G / kNumXcd) + b / kNumXcd;
    beg = (int)((int64_t)n * c / G);
    end = (int)((int64_t)n * (c + 1) / G);
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// Block-reduces NV values and stores them as partials[k*G + b].
template <int NV>
__device__ __forceinline__ void store_partials(double (&v)[NV], OGSX_GL double* partials, int G,
                                               int b)
{
    __shared__ double sm[NV][kBlock / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NV; k++) {
        double w = wave_sum(v[k]);
        if (lane == 0) sm[k][wave] = w;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; k++) {
            double s = sm[k][0];
#pragma unroll
            for (int w = 1; w < kBlock / 64; w++) s += sm[k][w];
            partials[(size_t)k * G + b] = s;
        }
    }
}

// Sums G partials per value in a fixed order; every thread gets the result.
// G is the grid size of the producer, which equals this kernel's grid, and
// at most kMaxGrid (grid_for's cap).  Thread t adds partials t, t + kBlock,
// t + 2 kBlock, ... in that order, all loaded before the first add, so a
// block waits for one round trip to L2 instead of one per partial (the
// partials were written by blocks on every XCD and miss the local L2).
constexpr int kMaxGrid = 2048;
constexpr int kPartialBatch = kMaxGrid / kBlock;
template <int NV>
struct PartialLoads {
    double t[NV][kPartialBatch];
};

// The first kPartialBatch partials per value of thread t, loaded (not yet
// added).  A kernel that also streams rows issues these loads first and its
// rows' loads after them, so that the reduction can wait for the partials
// (from L2) while the rows (from HBM) are still in flight: a wait on the
// memory counter covers every load issued before the one awaited.
template <int NV>
__device__ __forceinline__ void load_partials(const OGSX_GL double* partials, int G,
                                              PartialLoads<NV>& P)
{
#pragma unroll
    for (int k = 0; k < NV; k++)
#pragma unroll
        for (int j = 0; j < kPartialBatch; j++) {
            // clamped index: every load is issued unpredicated (G >= 1)
            const int i = threadIdx.x + j * kBlock;
            const double x = partials[(size_t)k * G + (i < G ? i : G - 1)];
            P.t[k][j] = i < G ? x : 0.0;
        }
}

// Completes reduce_partials from loads made by load_partials.
template <int NV>
__device__ __forceinline__ void finish_partials(const OGSX_GL double* partials, int G,
                                                const PartialLoads<NV>& P, double (&out)[NV])
{
    __shared__ double sm[NV][kBlock / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const auto& t = P.t;
    (void)partials;
    double v[NV];
#pragma unroll
    for (int k = 0; k < NV; k++) {
        // s starts at +0.0, so adding +0.0 for a missing partial changes no bit
        double s = 0.0;
#pragma unroll
        for (int j = 0; j < kPartialBatch; j++) s += t[k][j];
        v[k] = wave_sum(s);
    }
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < NV; k++) sm[k][wave] = v[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NV; k++) {
        double s = sm[k][0];
#pragma unroll
        for (int w = 1; w < kBlock / 64; w++) s += sm[k][w];
        out[k] = s;
    }
    __syncthreads();
}

template <int NV>
__device__ __forceinline__ void reduce_partials(const OGSX_GL double* partials, int G,
                                                double (&out)[NV])
{
    PartialLoads<NV> P;
    load_partials<NV>(partials, G, P);
    finish_partials<NV>(partials, G, P, out);
}

// Sum over one CRS row by a group of LPR lanes; every lane of the group gets
// the row sum.  Lanes of one group share `row`, so they enter and leave the
// row loop together and the width-LPR shuffles stay inside active lanes.
template <int LPR, typename VT, typename XT, typename ACC>
__device__ __forceinline__ ACC row_dot(const OGSX_GL int* __restrict__ ptr,
                                       const OGSX_GL int* __restrict__ col,
                                       const OGSX_GL VT* __restrict__ val,
                                       const OGSX_GL XT* __restrict__ x, int row, int lane)
{
    const int beg = ptr[row], end = ptr[row + 1];
    ACC s = 0;
    for (int k = beg + lane; k < end; k += LPR) s += (ACC)val[k] * (ACC)x[col[k]];
#pragma unroll
    for (int off = LPR / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, LPR);
    return s;
}

// Loops the block's chunk of rows with LPR lanes per row and calls
// epi(row, Ax_row) on the group's lane 0.
template <int LPR, typename VT, typename XT, typename ACC, typename Epi>
__device__ __forceinline__ void spmv_chunk(int n, const OGSX_GL int* __restrict__ ptr,
                                           const OGSX_GL int* __restrict__ col,
                                           const OGSX_GL VT* __restrict__ val,
                                           const OGSX_GL XT* __restrict__ x, int G, int b, Epi&& epi)
{
    constexpr int RPB = kBlock / LPR;
    int beg, end;
    chunk_of(n, G, b, beg, end);
    const int lane = threadIdx.x % LPR;
    for (int r = beg + threadIdx.x / LPR; r < end; r += RPB) {
        ACC s = row_dot<LPR, VT, XT, ACC>(ptr, col, val, x, r, lane);
        if (lane == 0) epi(r, s);
    }
}

template <typename Body>
__device__ __forceinline__ void vec_chunk(int n, int G, int b, Body&& body)
{
    int beg, end;
    chunk_of(n, G, b, beg, end);
    for (int i = beg + threadIdx.x; i < end; i += kBlock) body(i);
}

// The rows of chunk b, kVecU per thread at a time: load(k, u) reads row k
// into slot u for all kVecU slots (rows past the chunk's end read its last
// row and are not used), then apply(k, u) runs for the slots inside the
// chunk, in ascending row order.  A thread thus waits for memory once per
// batch instead of once (or, where stores may alias the next row's loads,
// several times) per row; per-thread order, hence every sum, is that of
// vec_chunk.  mid() runs on every thread between the first batch's loads
// and their use -- the reduction of the previous kernel's partials, whose
// loads the caller issued first, so it overlaps the rows in flight; when it
// returns false the rows are left alone.
constexpr int kVecU = 4;

template <typename Load, typename Mid, typename Apply>
__device__ __forceinline__ bool vec_chunk_batched_mid(int n, int G, int b, Load&& load, Mid&& mid,
                                                      Apply&& apply)
{
    int beg, end;
    chunk_of(n, G, b, beg, end);
    const int last = end - 1;
    int i = beg + threadIdx.x;
    if (i < end) {
#pragma unroll
        for (int u = 0; u < kVecU; u++) load(min(i + u * kBlock, last), u);
    }
    if (!mid()) return false;
    for (bool loaded = true; i < end; i += kVecU * kBlock, loaded = false) {
        if (!loaded) {
#pragma unroll
            for (int u = 0; u < kVecU; u++) load(min(i + u * kBlock, last), u);
        }
#pragma unroll
        for (int u = 0; u < kVecU; u++)
            if (i + u * kBlock < end) apply(i + u * kBlock, u);
    }
    return true;
}

// Matrix as the SpMV kernels see it: CRS (ptr, col, val) plus the row tiles
// of the tiled format (tile_row: first row of every tile; tile_cb: the
// smallest column of the tile, col16 = col - tile_cb when it fits 16 bits).
template <typename VT>
struct MatView {
    int n;
    int n_tiles;
    const int* ptr;
    const int* col;
    const uint16_t* col16;
    const VT* val;
    const int* tile_row;
    const int* tile_cb;
    // SELL-64: slice s holds rows [64 s, 64 s + 64) column-major: entry j of
    // lane l at sell_ptr[s] + 64 j + l, j < sell_w[s]; padding has value 0
    // and column offset 0.  Columns are sell_cb[s] + sell_c16 (16-bit) or
    // sell_c32.
    int n_slices;
    const int* sell_ptr;
    const int* sell_w;
    const int* sell_cb;
    const uint16_t* sell_c16;
    const int* sell_c32;
    const VT* sval;
    // stencil dictionary: row r's columns are r + pat_delta[sell_pid[r] *
    // pat_w + j] (rows padded to the slice, padding rows use the last row)
    const uint8_t* sell_pid;
    const int* pat_delta;
    int pat_w;
    int pat_cnt;
    // slice triples: slice s holds rows [64 s, 64 s + 64); its stencil is
    // tri_cnt[q] triples (q = tri_id[s]) of column offsets o, o+1, o+2 with
    // o = tri_off[q * kTriMax + k]; entry 3 k + t of lane l (column
    // 64 s + l + o + t) at tri_ptr[s] + 64 (3 k + t) + l, value 0 where the
    // row has no such column.  Values live in sval (the layout of the
    // system's current format).
    const int* tri_ptr;
    const uint16_t* tri_id;
    const int* tri_cnt;
    const int* tri_off;
    // symmetric diagonal storage (FP32 inner solves of symmetric systems on
    // banded patterns): dia_val[0 .. n) the diagonal, dia_val[(k + 1) n + i]
    // the entry (i, i + dia_off[k]) for the dia_k positive offsets (0 where
    // the pattern has no such entry); entry (i, i - o) is the stored (i - o, i)
    const int* dia_off;
    const VT* dia_val;
    int dia_k;
};

// MatView with global-address-space pointers, as the device loops use it.
template <typename VT>
struct GMat {
    int n;
    int n_tiles;
    const OGSX_GL int* ptr;
    const OGSX_GL int* col;
    const OGSX_GL uint16_t* col16;
    const OGSX_GL VT* val;
    const OGSX_GL int* tile_row;
    const OGSX_GL int* tile_cb;
    int n_slices;
    const OGSX_GL int* sell_ptr;
    const OGSX_GL int* sell_w;
    const OGSX_GL int* sell_cb;
    const OGSX_GL uint16_t* sell_c16;
    const OGSX_GL int* sell_c32;
    const OGSX_GL VT* sval;
    const OGSX_GL uint8_t* sell_pid;
    const OGSX_GL int* pat_delta;
    int pat_w;
    int pat_cnt;
    const OGSX_GL int* tri_ptr;
    const OGSX_GL uint16_t* tri_id;
    const OGSX_GL int* tri_cnt;
    const OGSX_GL int* tri_off;
    const OGSX_GL int* dia_off;
    const OGSX_GL VT* dia_val;
    int dia_k;
};

template <typename VT>
__device__ __forceinline__ GMat<VT> gmat(const MatView<VT>& m)
{
    return GMat<VT>{m.n,          m.n_tiles,       gl(m.ptr),      gl(m.col),
                    gl(m.col16),  gl(m.val),       gl(m.tile_row), gl(m.tile_cb),
                    m.n_slices,   gl(m.sell_ptr),  gl(m.sell_w),   gl(m.sell_cb),
                    gl(m.sell_c16), gl(m.sell_c32), gl(m.sval),    gl(m.sell_pid),
                    gl(m.pat_delta), m.pat_w,       m.pat_cnt,     gl(m.tri_ptr),
                    gl(m.tri_id),   gl(m.tri_cnt),  gl(m.tri_off), gl(m.dia_off),
                    gl(m.dia_val),  m.dia_k};
}

constexpr int kTileU = 8;                 // nonzeros per thread per tile
constexpr int kTileNnz = kTileU * kBlock; // nonzeros per tile (2048)
constexpr int kFmtTile32 = 1;             // tiled, 32-bit columns
constexpr int kFmtTile16 = 2;             // tiled, 16-bit column offsets
constexpr int kFmtSell32 = 3;             // SELL-64, 32-bit columns
constexpr int kFmtSell16 = 5;             // SELL-64, 16-bit column offsets
constexpr int kFmtSell16U8 = 6;           // ... 8 entries in flight per lane
constexpr int kFmtSell16NT = 7;           // ... and non-temporal matrix loads
constexpr int kFmtSell16U16 = 10;         // SELL-64, 16-bit, 16 in flight, NT
constexpr int kFmtSell16X2 = 11;          // SELL-64, 16-bit, two slices per wave, NT
constexpr int kFmtSell16U7 = 12;          // SELL-64, 16-bit, 7 entries in flight, NT
constexpr int kFmtSellPat8 = 13;          // SELL-64, stencil dictionary, 8 in flight, NT
constexpr int kFmtSellPat7 = 14;          // SELL-64, stencil dictionary, 7 in flight, NT
constexpr int kFmtSellPat8C = 17;         // ... 8 in flight, cached loads
constexpr int kFmtSellPat7C = 18;         // ... 7 in flight, cached loads
constexpr int kFmtSellPat16 = 21;         // ... a row's entries 16 at a time, the last
                                          // batch predicated (one round trip for
                                          // rows of up to 16 entries), NT
constexpr int kFmtSellPat16C = 22;        // ... the same with cached loads
constexpr int kSellC = 64;                // rows per slice = wavefront size
constexpr int kPatLds = 4096;             // dictionary entries staged in LDS
constexpr int kFmtTri5 = 19;              // SELL-64 slice triples, 5 in flight, NT
constexpr int kFmtTri4C = 20;             // ... 4 in flight, cached loads
constexpr int kTriMax = 16;               // triples per slice stencil
constexpr int kFmtSymDia = 25;            // symmetric diagonal storage, FP32 inner
                                          // solves only (chosen by structure, not timed)
constexpr int kDiaMax = 16;               // positive offsets it takes at most
constexpr int kFmtCrs1 = 9;               // plain CRS, one thread per row (needs only
                                          // ptr, col, val)

// Tiled CRS SpMV: a block streams the contiguous nonzeros of one tile of
// rows (<= kTileNnz nonzeros, <= kBlock rows) with kTileU independent
// coalesced loads per thread, gathers x, parks the products in LDS and then
// sums each row from LDS with one thread per row.  Many independent loads
// per thread keep enough bytes in flight to stream at HBM rate, which a
// lane-group-per-row loop (one dependent load chain per row) does not.
// Tiles are split into G contiguous ranges dealt XCD-aware like rows.  Rows
// are summed left to right over rounded products, as the SELL kernel does.
// A tile whose single row exceeds kTileNnz is summed by the whole block.
template <int C16, typename VT, typename XT, typename ACC, typename Epi>
__device__ __forceinline__ void spmv_tiles(const GMat<VT>& m, const OGSX_GL XT* __restrict__ x,
                                           int G, int b, Epi&& epi)
{
#pragma clang fp contract(off)
    __shared__ ACC prod[kTileNnz];
    int tb, te;
    chunk_of(m.n_tiles, G, b, tb, te);
    const int tid = threadIdx.x;
    for (int t = tb; t < te; t++) {
        const int r0 = m.tile_row[t], r1 = m.tile_row[t + 1];
        const int k0 = m.ptr[r0], k1 = m.ptr[r1];
        const int cb = C16 ? m.tile_cb[t] : 0;
        if (k1 - k0 > kTileNnz) {
            // one long row (r1 == r0 + 1): block-wide reduction
            ACC s = 0;
            for (int k = k0 + tid; k < k1; k += kBlock) {
                const int c = C16 ? cb + (int)m.col16[k] : m.col[k];
                const ACC pr = (ACC)m.val[k] * (ACC)x[c];
                s = s + pr;
            }
            double w = wave_sum((double)s);
            __shared__ double red[kBlock / 64];
            if ((tid & 63) == 0) red[tid >> 6] = w;
            __syncthreads();
            if (tid == 0) {
                double tot = red[0];
                for (int q = 1; q < kBlock / 64; q++) tot += red[q];
                epi(r0, (ACC)tot);
            }
            __syncthreads();
            continue;
        }
        VT v[kTileU];
        int c[kTileU];
#pragma unroll
        for (int j = 0; j < kTileU; j++) {
            const int k = k0 + tid + j * kBlock;
            if (k < k1) {
                v[j] = m.val[k];
                c[j] = C16 ? cb + (int)m.col16[k] : m.col[k];
            }
        }
        XT xv[kTileU];
#pragma unroll
        for (int j = 0; j < kTileU; j++) {
            const int k = k0 + tid + j * kBlock;
            if (k < k1) xv[j] = x[c[j]];
        }
#pragma unroll
        for (int j = 0; j < kTileU; j++) {
            const int k = k0 + tid + j * kBlock;
            if (k < k1) prod[tid + j * kBlock] = (ACC)v[j] * (ACC)xv[j];
        }
        __syncthreads();
        const int r = r0 + tid;
        if (r < r1) {
            const int q0 = m.ptr[r] - k0, q1 = m.ptr[r + 1] - k0;
            ACC s = 0;
            for (int q = q0; q < q1; q++) s += prod[q];
            epi(r, s);
        }
        __syncthreads();
    }
}

// SELL-64 SpMV: one wavefront per slice, one lane per row.  Each row is
// summed left to right from 0 over rounded products (no FMA contraction):
// the order and rounding of CSparseMatrix::multiVec (matrix_class.cpp:1527-
// 1538), so the result is bit-identical to the reference's SpMV and to the
// tiled kernel, whichever of them the autotuner picks.  Loads of the
// values and column offsets are fully coalesced (64 consecutive entries per
// instruction), rows need no cross-lane reduction, and for meshes numbered
// along a grid line the 64 lanes' x gathers of one entry j touch 64
// consecutive x values.  Blocks take contiguous slice ranges, XCD-aware.
// NT selects non-temporal loads (the matrix streams through once per SpMV;
// keeping it out of the caches leaves them to x and the Krylov vectors).  A
// compile-time choice: a run-time select between the two loads of the same
// address gets merged into one plain load, losing the hint.
template <bool NT, typename P>
__device__ __forceinline__ auto ld(P p)
{
    if constexpr (NT)
        return __builtin_nontemporal_load(p);
    else
        return *p;
}

template <bool NT, typename P, typename V>
__device__ __forceinline__ void store_hint(P p, V v)
{
    if constexpr (NT)
        __builtin_nontemporal_store(v, p);
    else
        *p = v;
}

template <int C16, int U, bool NT, typename VT, typename XT, typename ACC, typename Epi>
__device__ __forceinline__ void spmv_sell(const GMat<VT>& m, const OGSX_GL XT* __restrict__ x,
                                          int G, int b, Epi&& epi)
{
#pragma clang fp contract(off)
    int sb, se;
    chunk_of(m.n_slices, G, b, sb, se);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int sl = sb + wave; sl < se; sl += kBlock / 64) {
        const int base = m.sell_ptr[sl] + lane;
        const int w = m.sell_w[sl];
        const int cb = C16 ? m.sell_cb[sl] : 0;
        ACC s = 0;
        int j = 0;
        for (; j + U <= w; j += U) {
            int c[U];
            VT v[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int k = base + (j + u) * kSellC;
                if constexpr (C16)
                    c[u] = cb + (int)ld<NT>(m.sell_c16 + k);
                else
                    c[u] = ld<NT>(m.sell_c32 + k);
                v[u] = ld<NT>(m.sval + k);
            }
            ACC xv[U];
#pragma unroll
            for (int u = 0; u < U; u++) xv[u] = (ACC)x[c[u]];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const ACC prod = (ACC)v[u] * xv[u];
                s = s + prod;
            }
        }
        for (; j < w; j++) {
            const int k = base + j * kSellC;
            const int c0 = C16 ? cb + (int)m.sell_c16[k] : m.sell_c32[k];
            const ACC prod = (ACC)m.sval[k] * (ACC)x[c0];
            s = s + prod;
        }
        const int row = sl * kSellC + lane;
        if (row < m.n) epi(row, s);
    }
}

// SELL-64 with a stencil dictionary instead of column indices: rows whose
// columns sit at the same offsets from the row (the interior rows of a
// structured mesh, and each kind of boundary row) share one dictionary
// entry, so a row costs one byte (its pattern id) instead of 2-4 bytes per
// stored entry.  The block stages the dictionary in LDS; lane l of a slice
// reads entry j's column as row + delta[pid][j].  Values and summation are
// those of the other SELL formats (same value array, CRS order, rounded
// products), so results are bit-identical.  Padding entries have value 0
// and delta 0; lanes past the last row take the last row's column.
//
// PRED: the row's entries are taken U at a time down to the last, partial
// batch, whose missing entries read the row's last entry again and are not
// added (the slice width w is the same for the wave, so the test is
// uniform).  With U = 16 a row of up to 16 entries (2-D quads 9, tets 15)
// costs one round trip for its values and one for its x gathers, where the
// other forms add the entries past the last full batch one at a time.
template <int U, bool NT, bool PRED, typename VT, typename XT, typename ACC, typename Epi>
__device__ __forceinline__ void spmv_sell_pat(const GMat<VT>& m, const OGSX_GL XT* __restrict__ x,
                                              int G, int b, Epi&& epi)
{
#pragma clang fp contract(off)
    __shared__ int pd[kPatLds];
    const int tot = m.pat_cnt * m.pat_w;
    for (int i = threadIdx.x; i < tot; i += blockDim.x) pd[i] = m.pat_delta[i];
    __syncthreads();
    int sb, se;
    chunk_of(m.n_slices, G, b, sb, se);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int sl = sb + wave; sl < se; sl += kBlock / 64) {
        const int row = sl * kSellC + lane;
        const int rowc = row < m.n ? row : m.n - 1;
        const int base = m.sell_ptr[sl] + lane;
        const int w = m.sell_w[sl];
        const int* pr = pd + (int)m.sell_pid[row] * m.pat_w;
        ACC s = 0;
        if constexpr (PRED) {
            for (int j = 0; j < w; j += U) {
                int c[U];
                VT v[U];
#pragma unroll
                for (int u = 0; u < U; u++) {
                    const int jj = min(j + u, w - 1);
                    v[u] = ld<NT>(m.sval + base + jj * kSellC);
                    c[u] = rowc + pr[jj];
                }
                ACC xv[U];
#pragma unroll
                for (int u = 0; u < U; u++) xv[u] = (ACC)x[c[u]];
#pragma unroll
                for (int u = 0; u < U; u++) {
                    if (j + u < w) {
                        const ACC prod = (ACC)v[u] * xv[u];
                        s = s + prod;
                    }
                }
            }
            if (row < m.n) epi(row, s);
            continue;
        }
        int j = 0;
        for (; j + U <= w; j += U) {
            int c[U];
            VT v[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                v[u] = ld<NT>(m.sval + base + (j + u) * kSellC);
                c[u] = rowc + pr[j + u];
            }
            ACC xv[U];
#pragma unroll
            for (int u = 0; u < U; u++) xv[u] = (ACC)x[c[u]];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const ACC prod = (ACC)v[u] * xv[u];
                s = s + prod;
            }
        }
        for (; j < w; j++) {
            const ACC prod = (ACC)m.sval[base + j * kSellC] * (ACC)x[rowc + pr[j]];
            s = s + prod;
        }
        if (row < m.n) epi(row, s);
    }
}

// Lane l of the wave takes lane l + 1's value; lane 63 takes `edge` (DPP
// wave_shl:1, lanes whose source is past the wave keep the old value).
__device__ __forceinline__ float wave_next(float v, float edge)
{
    return __int_as_float(
        __builtin_amdgcn_update_dpp(__float_as_int(edge), __float_as_int(v), 0x130, 0xF, 0xF, false));
}
__device__ __forceinline__ double wave_next(double v, double edge)
{
    const long long vb = __double_as_longlong(v), eb = __double_as_longlong(edge);
    const int lo = __builtin_amdgcn_update_dpp((int)eb, (int)vb, 0x130, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp((int)(eb >> 32), (int)(vb >> 32), 0x130, 0xF, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// SELL-64 with one stencil per slice, in triples of consecutive column
// offsets: for a mesh numbered along grid lines, the 64 rows of a slice and
// their neighbours o, o+1, o+2 read x[r0 + l + o ...]: one coalesced load of
// x[r0 + l + o] per triple, the +1 and +2 neighbours by shifting it one lane
// across the wave (the wave's last lanes take x[r0 + 64 + o], x[r0 + 65 + o]
// from scalar loads).  A third of the gathers of the per-row formats, and
// none of their address arithmetic.  The slice stencil is the union of its
// rows' offsets, so a row may hold zero entries for columns it lacks; a zero
// product adds nothing (x finite), the others are summed left to right over
// rounded products in column order: bit-identical to the other formats.
// Gathers of slices near the first or last row are clamped to [0, n).
template <int UT, bool NT, typename VT, typename XT, typename ACC, typename Epi>
__device__ __forceinline__ void spmv_tri(const GMat<VT>& m, const OGSX_GL XT* __restrict__ x,
                                         int G, int b, Epi&& epi)
{
#pragma clang fp contract(off)
    int sb, se;
    chunk_of(m.n_slices, G, b, sb, se);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nm1 = m.n - 1;
    for (int sl0 = sb + wave; sl0 < se; sl0 += kBlock / 64) {
        const int sl = __builtin_amdgcn_readfirstlane(sl0);  // wave-uniform
        const int r0 = sl * kSellC;
        const int q = m.tri_id[sl];
        const int K = m.tri_cnt[q];
        const OGSX_GL int* of = m.tri_off + q * kTriMax;
        const OGSX_GL VT* vb = m.sval + m.tri_ptr[sl];
        const bool inside = r0 + of[0] >= 0 && r0 + of[K - 1] + kSellC + 1 <= nm1;
        ACC s = 0;
        int k = 0;
        for (; k + UT <= K; k += UT) {
            VT v[3 * UT];
            XT xb[UT], e1[UT], e2[UT];
#pragma unroll
            for (int u = 0; u < UT; u++) {
                const int o = of[k + u];
#pragma unroll
                for (int t = 0; t < 3; t++) v[3 * u + t] = ld<NT>(vb + (3 * (k + u) + t) * kSellC + lane);
                if (inside) {
                    const OGSX_GL XT* xo = x + (r0 + o);
                    xb[u] = xo[lane];
                    e1[u] = xo[kSellC];
                    e2[u] = xo[kSellC + 1];
                } else {
                    xb[u] = x[min(max(r0 + lane + o, 0), nm1)];
                    e1[u] = x[min(max(r0 + kSellC + o, 0), nm1)];
                    e2[u] = x[min(max(r0 + kSellC + 1 + o, 0), nm1)];
                }
            }
#pragma unroll
            for (int u = 0; u < UT; u++) {
                const XT x1 = wave_next(xb[u], e1[u]);
                const XT x2 = wave_next(x1, e2[u]);
                ACC pr = (ACC)v[3 * u] * (ACC)xb[u];
                s = s + pr;
                pr = (ACC)v[3 * u + 1] * (ACC)x1;
                s = s + pr;
                pr = (ACC)v[3 * u + 2] * (ACC)x2;
                s = s + pr;
            }
        }
        for (; k < K; k++) {
            const int o = of[k];
            const XT x0 = x[min(max(r0 + lane + o, 0), nm1)];
            const XT x1 = wave_next(x0, x[min(max(r0 + kSellC + o, 0), nm1)]);
            const XT x2 = wave_next(x1, x[min(max(r0 + kSellC + 1 + o, 0), nm1)]);
            ACC pr = (ACC)vb[(3 * k) * kSellC + lane] * (ACC)x0;
            s = s + pr;
            pr = (ACC)vb[(3 * k + 1) * kSellC + lane] * (ACC)x1;
            s = s + pr;
            pr = (ACC)vb[(3 * k + 2) * kSellC + lane] * (ACC)x2;
            s = s + pr;
        }
        const int row = r0 + lane;
        if (row < m.n) epi(row, s);
    }
}

// Two slices per wavefront in one loop: twice the independent loads in
// flight; each row is still summed in CRS order.
template <int U, typename VT, typename XT, typename ACC, typename Epi>
__device__ __forceinline__ void spmv_sell_x2(const GMat<VT>& m, const OGSX_GL XT* __restrict__ x,
                                             int G, int b, Epi&& epi)
{
#pragma clang fp contract(off)
    int sb, se;
    chunk_of(m.n_slices, G, b, sb, se);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    constexpr int W = kBlock / 64;
    for (int sl = sb + wave; sl < se; sl += 2 * W) {
        const int sl2 = sl + W;
        const bool two = sl2 < se;
        const int base0 = m.sell_ptr[sl] + lane, w0 = m.sell_w[sl], cb0 = m.sell_cb[sl];
        const int base1 = two ? m.sell_ptr[sl2] + lane : 0, w1 = two ? m.sell_w[sl2] : 0;
        const int cb1 = two ? m.sell_cb[sl2] : 0;
        ACC s0 = 0, s1 = 0;
        const int wmax = w0 > w1 ? w0 : w1;
        for (int j = 0; j < wmax; j += U) {
            int c0[U], c1[U];
            VT v0[U], v1[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int k0 = base0 + (j + u) * kSellC, k1 = base1 + (j + u) * kSellC;
                if (j + u < w0) {
                    c0[u] = cb0 + (int)__builtin_nontemporal_load(m.sell_c16 + k0);
                    v0[u] = __builtin_nontemporal_load(m.sval + k0);
                }
                if (j + u < w1) {
                    c1[u] = cb1 + (int)__builtin_nontemporal_load(m.sell_c16 + k1);
                    v1[u] = __builtin_nontemporal_load(m.sval + k1);
                }
            }
            ACC x0[U], x1[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                if (j + u < w0) x0[u] = (ACC)x[c0[u]];
                if (j + u < w1) x1[u] = (ACC)x[c1[u]];
            }
#pragma unroll
            for (int u = 0; u < U; u++) {
                if (j + u < w0) {
                    const ACC p0 = (ACC)v0[u] * x0[u];
                    s0 = s0 + p0;
                }
                if (j + u < w1) {
                    const ACC p1 = (ACC)v1[u] * x1[u];
                    s1 = s1 + p1;
                }
            }
        }
        const int row0 = sl * kSellC + lane;
        if (row0 < m.n) epi(row0, s0);
        if (two) {
            const int row1 = sl2 * kSellC + lane;
            if (row1 < m.n) epi(row1, s1);
        }
    }
}

// Plain CRS with one thread per row: each row summed left to right over
// rounded products (CSparseMatrix::multiVec's order,
// FEM/matrix_class.cpp:1527-1538).  Needs no layout beyond ptr, col and
// val, so it serves matrices the library did not build (the zero-allocation
// solve); the tuned formats stream faster.  Rows map to blocks and threads
// as in the SELL-64 formats -- block b takes the rows of its chunk of 64-row
// slices, thread t the rows R0 + t + 256 j -- so the fused dot products of
// the solve kernels (accumulated per thread in that order, then reduced in
// a fixed tree) round exactly as with every SELL-64 format.
template <typename VT, typename XT, typename ACC, typename Epi>
__device__ __forceinline__ void spmv_crs1(const GMat<VT>& m, const OGSX_GL XT* __restrict__ x,
                                          int G, int b, Epi&& epi)
{
#pragma clang fp contract(off)
    int sb, se;
    chunk_of((m.n + kSellC - 1) / kSellC, G, b, sb, se);
    const int beg = sb * kSellC, end = min(se * kSellC, m.n);
    for (int r = beg + (int)threadIdx.x; r < end; r += kBlock) {
        const int k1 = m.ptr[r + 1];
        ACC s = 0;
        for (int k = m.ptr[r]; k < k1; k++) s += (ACC)m.val[k] * (ACC)x[m.col[k]];
        epi(r, s);
    }
}

// Format dispatch: FMT = 4..64 lanes per row (vector CRS), or a tiled,
// SELL-64 or plain CRS format.
// Symmetric diagonal storage: y_i = d_i x_i + sum_k (u_k[i] x[i + o_k] +
// u_k[i - o_k] x[i - o_k]), one thread per row, offsets in ascending order
// and for each the upper, then the lower entry.  Half the matrix values of
// the full formats cross HBM; the lower entries are the upper ones of rows
// o_k back, read again from cache (the blocks of one XCD take consecutive
// chunks).  Not the reference's summation order: this form serves the FP32
// matrix inside the FP64 refinement, whose FP64 residual stays exact; the
// FP64 matrix uses spmv_symdia_crs_k below.
#ifndef OGSX_DIA_ROWS
#define OGSX_DIA_ROWS 2
#endif
// KC > 0: the offset count known at compile time (the loop over them fully
// unrolled); KC = 0: m.dia_k at run time.
template <int KC, typename VT, typename XT, typename ACC, typename Epi>
__device__ __forceinline__ void spmv_symdia_k(const GMat<VT>& m, const OGSX_GL XT* __restrict__ x,
                                              int G, int b, Epi&& epi)
{
    int beg, end;
    chunk_of(m.n, G, b, beg, end);
    const int n = m.n, K = KC > 0 ? KC : m.dia_k;
    const OGSX_GL VT* dv = m.dia_val;
    // R rows per thread at a time (i, i + kBlock, ...; those past the chunk
    // clamped to its last row and not stored), sums before any epilogue, so
    // the loads of the R rows are in flight together (128^3 FP32: R = 1, 2,
    // 4 gave 36.1, 35.5, 43.8 us before the offset loop was unrolled; unrolled
    // R = 1, 2: 31.5, 31.4 us)
    constexpr int R = OGSX_DIA_ROWS;
    for (int i = beg + threadIdx.x; i < end; i += R * kBlock) {
        int r[R];
#pragma unroll
        for (int q = 0; q < R; q++) r[q] = min(i + q * kBlock, end - 1);
        ACC s[R];
#pragma unroll
        for (int q = 0; q < R; q++) s[q] = (ACC)dv[r[q]] * (ACC)x[r[q]];
        constexpr int kUnroll = KC > 0 ? KC : 4;
#pragma unroll kUnroll
        for (int k = 0; k < K; k++) {
            const int o = m.dia_off[k];
            const OGSX_GL VT* u = dv + (int64_t)(k + 1) * n;
#pragma unroll
            for (int q = 0; q < R; q++) {
                const int up = min(r[q] + o, n - 1), lo = r[q] - o;
                const ACC a1 = (ACC)u[r[q]], x1 = (ACC)x[up];  // u[i] = 0 where i + o >= n
                const ACC a2 = lo >= 0 ? (ACC)u[max(lo, 0)] : (ACC)0, x2 = (ACC)x[max(lo, 0)];
                s[q] = s[q] + a1 * x1;
                s[q] = s[q] + a2 * x2;
            }
        }
        epi(r[0], s[0]);
#pragma unroll
        for (int q = 1; q < R; q++)
            if (i + q * kBlock < end) epi(r[q], s[q]);
    }
}

// The same storage summed in the CRS row's column order: the lower entries
// from the largest offset down, the diagonal, the upper entries from the
// smallest offset up, no fused multiply-adds.  A stored zero (a column the
// row does not have) adds +-0 to the running sum, which leaves it unchanged
// unless it is itself zero, so the result equals the CRS kernels' bit for
// bit (up to the sign of an all-zero row): the FP64 matrix of a symmetric
// system reads half its values.
template <int KC, typename VT, typename XT, typename ACC, typename Epi>
__device__ __forceinline__ void spmv_symdia_crs_k(const GMat<VT>& m, const OGSX_GL XT* __restrict__ x,
                                                  int G, int b, Epi&& epi)
{
#pragma clang fp contract(off)
    // rows to blocks and threads as in the CRS and SELL-64 formats (block b
    // the rows of its chunk of 64-row slices, thread t the rows R0 + t +
    // 256 j in that order), so the solve kernels' fused dot products round
    // as with those formats
    int sb, se;
    chunk_of((m.n + kSellC - 1) / kSellC, G, b, sb, se);
    const int beg = sb * kSellC, end = min(se * kSellC, m.n);
    const int n = m.n, K = KC > 0 ? KC : m.dia_k;
    const OGSX_GL VT* dv = m.dia_val;
    // one row per thread at a time: two (the FP32 form's) took 229 VGPRs
    // for 13 offsets, 2 waves per SIMD, 94 us on 128^3 against 72 us here
    // (72 VGPRs, 7 waves)
    constexpr int R = 1;
    constexpr int kUnroll = KC > 0 ? KC : 4;
    for (int i = beg + threadIdx.x; i < end; i += R * kBlock) {
        int r[R];
#pragma unroll
        for (int q = 0; q < R; q++) r[q] = min(i + q * kBlock, end - 1);
        ACC s[R];
#pragma unroll
        for (int q = 0; q < R; q++) s[q] = 0;
#pragma unroll kUnroll
        for (int k = K - 1; k >= 0; k--) {
            const int o = m.dia_off[k];
            const OGSX_GL VT* u = dv + (int64_t)(k + 1) * n;
#pragma unroll
            for (int q = 0; q < R; q++) {
                const int lo = max(r[q] - o, 0);
                const ACC a = r[q] - o >= 0 ? (ACC)u[lo] : (ACC)0;
                const ACC prod = a * (ACC)x[lo];
                s[q] = s[q] + prod;
            }
        }
#pragma unroll
        for (int q = 0; q < R; q++) {
            const ACC prod = (ACC)dv[r[q]] * (ACC)x[r[q]];
            s[q] = s[q] + prod;
        }
#pragma unroll kUnroll
        for (int k = 0; k < K; k++) {
            const int o = m.dia_off[k];
            const OGSX_GL VT* u = dv + (int64_t)(k + 1) * n;
#pragma unroll
            for (int q = 0; q < R; q++) {
                const int up = min(r[q] + o, n - 1);  // u[i] = 0 where i + o >= n
                const ACC prod = (ACC)u[r[q]] * (ACC)x[up];
                s[q] = s[q] + prod;
            }
        }
        epi(r[0], s[0]);
#pragma unroll
        for (int q = 1; q < R; q++)
            if (i + q * kBlock < end) epi(r[q], s[q]);
    }
}

// The offset counts of the structured meshes (4: 9-point quads, 7: 15-point
// tets, 13: 27-point hexes) get fully unrolled instances.  FP64 values are
// summed in CRS order (spmv_symdia_crs_k), FP32 ones per offset pair.
template <typename VT, typename XT, typename ACC, typename Epi>
__device__ __forceinline__ void spmv_symdia(const GMat<VT>& m, const OGSX_GL XT* __restrict__ x,
                                            int G, int b, Epi&& epi)
{
    if constexpr (sizeof(VT) == 8) {
        switch (m.dia_k) {
            case 4: spmv_symdia_crs_k<4, VT, XT, ACC>(m, x, G, b, epi); break;
            case 7: spmv_symdia_crs_k<7, VT, XT, ACC>(m, x, G, b, epi); break;
            case 13: spmv_symdia_crs_k<13, VT, XT, ACC>(m, x, G, b, epi); break;
            default: spmv_symdia_crs_k<0, VT, XT, ACC>(m, x, G, b, epi); break;
        }
        return;
    }
    switch (m.dia_k) {
        case 4: spmv_symdia_k<4, VT, XT, ACC>(m, x, G, b, epi); break;
        case 7: spmv_symdia_k<7, VT, XT, ACC>(m, x, G, b, epi); break;
        case 13: spmv_symdia_k<13, VT, XT, ACC>(m, x, G, b, epi); break;
        default: spmv_symdia_k<0, VT, XT, ACC>(m, x, G, b, epi); break;
    }
}

template <int FMT, typename VT, typename XT, typename ACC, typename Epi>
__device__ __forceinline__ void spmv_rows(const GMat<VT>& m, const OGSX_GL XT* __restrict__ x,
                                          int G, int b, Epi&& epi)
{
    if constexpr (FMT == kFmtTile32)
        spmv_tiles<0, VT, XT, ACC>(m, x, G, b, epi);
    else if constexpr (FMT == kFmtTile16)
        spmv_tiles<1, VT, XT, ACC>(m, x, G, b, epi);
    else if constexpr (FMT == kFmtSell32)
        spmv_sell<0, 4, false, VT, XT, ACC>(m, x, G, b, epi);
    else if constexpr (FMT == kFmtSell16)
        spmv_sell<1, 4, false, VT, XT, ACC>(m, x, G, b, epi);
    else if constexpr (FMT == kFmtSell16U8)
        spmv_sell<1, 8, false, VT, XT, ACC>(m, x, G, b, epi);
    else if constexpr (FMT == kFmtSell16NT)
        spmv_sell<1, 8, true, VT, XT, ACC>(m, x, G, b, epi);
    else if constexpr (FMT == kFmtSell16U16)
        spmv_sell<1, 16, true, VT, XT, ACC>(m, x, G, b, epi);
    else if constexpr (FMT == kFmtSell16U7)
        spmv_sell<1, 7, true, VT, XT, ACC>(m, x, G, b, epi);
    else if constexpr (FMT == kFmtSellPat8)
        spmv_sell_pat<8, true, false, VT, XT, ACC>(m, x, G, b, epi);
    else if constexpr (FMT == kFmtSellPat7)
        spmv_sell_pat<7, true, false, VT, XT, ACC>(m, x, G, b, epi);
    else if constexpr (FMT == kFmtSellPat8C)
        spmv_sell_pat<8, false, false, VT, XT, ACC>(m, x, G, b, epi);
    else if constexpr (FMT == kFmtSellPat7C)
        spmv_sell_pat<7, false, false, VT, XT, ACC>(m, x, G, b, epi);
    else if constexpr (FMT == kFmtSellPat16)
        spmv_sell_pat<16, true, true, VT, XT, ACC>(m, x, G, b, epi);
    else if constexpr (FMT == kFmtSellPat16C)
        spmv_sell_pat<16, false, true, VT, XT, ACC>(m, x, G, b, epi);
    else if constexpr (FMT == kFmtSell16X2)
        spmv_sell_x2<4, VT, XT, ACC>(m, x, G, b, epi);
    else if constexpr (FMT == kFmtTri5)
        spmv_tri<5, true, VT, XT, ACC>(m, x, G, b, epi);
    else if constexpr (FMT == kFmtTri4C)
        spmv_tri<4, false, VT, XT, ACC>(m, x, G, b, epi);
    else if constexpr (FMT == kFmtCrs1)
        spmv_crs1<VT, XT, ACC>(m, x, G, b, epi);
    else if constexpr (FMT == kFmtSymDia)
        spmv_symdia<VT, XT, ACC>(m, x, G, b, epi);
    else
        spmv_chunk<FMT, VT, XT, ACC>(m.n, m.ptr, m.col, m.val, x, G, b, epi);
}

// ------------------------------------------------------------------------
// element assembly (one thread per element of one colour)
// ------------------------------------------------------------------------
struct AsmArgs {
    const int32_t* conn;
    const int32_t* color_elems;
    const int32_t* slot;
    const double* x;
    const double* y;
    const double* z;
    const double* mass_coef;
    const double* cond;
    const double* vel;      // [E][3][ngp] or null
    const double* tabN;     // [ngp][npe]
    const double* tabdN;    // [ngp][edim][npe]
    const double* tabW;     // [ngp][3]
    const double* u_old;
    double* A;
    double* b;
    double* eb;         // two-pass assembly, per (element, node) pair at its
                        // incidence position q = pos[e npe + i]: the
                        // element's row i at eb[q npe], its right-hand side
                        // at rb[q s] and, for s = 2, its gravity term (or 0)
                        // at rb[2 q + 1]; s = rb_stride<ET>(gravity)
    double* rb;
    const int32_t* pos;
    int cond_stride;
    int ngp;
    int steady;
    int pad;
    double tuf, mu, geo_area, adv_coef, rho_g, dt, theta;
};

template <int ET> struct ETraits;
template <> struct ETraits<OGSX_LINE2> { static constexpr int NPE = 2, D = 1, NW = 1; };
template <> struct ETraits<OGSX_QUAD4> { static constexpr int NPE = 4, D = 2, NW = 2; };
template <> struct ETraits<OGSX_HEX8> { static constexpr int NPE = 8, D = 3, NW = 3; };
template <> struct ETraits<OGSX_TET4> { static constexpr int NPE = 4, D = 3, NW = 1; };

// Doubles per (element, node) record of the two-pass right-hand sides: the
// right-hand side alone, or with the gravity term beside it (one 16-byte
// store).  Hexes keep the pair without gravity too: on the 128^3 mesh the
// 8-byte records made the assembly 5 % slower (2.15 against 2.04 ms), while
// on tets and quads they save 3 % (0.78 against 0.81 ms on 96^3 x 6 tets).
template <int ET>
__host__ __device__ constexpr int rb_stride(bool gravity)
{
    return gravity || ET == OGSX_HEX8 ? 2 : 1;
}

// Jacobian determinant and global shape-function gradients at one Gauss
// point: CElement::computeJacobian (FEM/fem_ele.cpp:585-729) and
// ComputeGradShapefct (:1128-1160).  dNg are the local gradients [D][NPE];
// dshp receives the global ones [D][NPE].
template <int ET>
__device__ __forceinline__ double elem_gradients(const double* X, const double* Y,
                                                 const double* Z, const double* dNg,
                                                 double* dshp)
{
    constexpr int NPE = ETraits<ET>::NPE, D = ETraits<ET>::D;
    double J[D * D];
#pragma unroll
    for (int i = 0; i < D * D; i++) J[i] = 0.0;
    double inv[D * D], det;
    if constexpr (D == 1) {
        const double dx = X[1] - X[0];
        J[0] = 0.5 * dx;
        inv[0] = 2.0 / dx;
        det = J[0];
    } else if constexpr (D == 2) {
#pragma unroll
        for (int i = 0; i < NPE; i++) {
            J[0] += X[i] * dNg[i];
            J[1] += Y[i] * dNg[i];
            J[2] += X[i] * dNg[NPE + i];
            J[3] += Y[i] * dNg[NPE + i];
        }
        det = J[0] * J[3] - J[1] * J[2];
        inv[0] = J[3];
        inv[1] = -J[1];
        inv[2] = -J[2];
        inv[3] = J[0];
#pragma unroll
        for (int i = 0; i < 4; i++) inv[i] /= det;
    } else {
#pragma unroll
        for (int i = 0; i < NPE; i++) {
            const double d0 = dNg[i], d1 = dNg[NPE + i], d2 = dNg[2 * NPE + i];
            J[0] += X[i] * d0;
            J[1] += Y[i] * d0;
            J[2] += Z[i] * d0;
            J[3] += X[i] * d1;
            J[4] += Y[i] * d1;
            J[5] += Z[i] * d1;
            J[6] += X[i] * d2;
            J[7] += Y[i] * d2;
            J[8] += Z[i] * d2;
        }
        det = J[0] * (J[4] * J[8] - J[7] * J[5]) + J[6] * (J[1] * J[5] - J[4] * J[2]) +
              J[3] * (J[2] * J[7] - J[8] * J[1]);
        inv[0] = J[4] * J[8] - J[7] * J[5];
        inv[1] = J[2] * J[7] - J[1] * J[8];
        inv[2] = J[1] * J[5] - J[2] * J[4];
        inv[3] = J[5] * J[6] - J[8] * J[3];
        inv[4] = J[0] * J[8] - J[6] * J[2];
        inv[5] = J[2] * J[3] - J[5] * J[0];
        inv[6] = J[3] * J[7] - J[6] * J[4];
        inv[7] = J[1] * J[6] - J[7] * J[0];
        inv[8] = J[0] * J[4] - J[3] * J[1];
#pragma unroll
        for (int i = 0; i < 9; i++) inv[i] /= det;
    }
    // global gradients dshp[k][i] = sum_m inv[k][m] dN_local[m][i]
#pragma unroll
    for (int i = 0; i < NPE; i++) {
#pragma unroll
        for (int k = 0; k < D; k++) {
            double s = 0.0;
#pragma unroll
            for (int m = 0; m < D; m++) s += inv[k * D + m] * dNg[m * NPE + i];
            dshp[k * NPE + i] = s;
        }
    }
    return det;
}

// Local matrices follow CElement::computeJacobian / ComputeGradShapefct
// (FEM/fem_ele.cpp:585-729, 1128-1160), CalcMass (FEM/fem_ele_std.cpp:3753),
// CalcLaplace (:4948), CalcAdvection (:5280), Assemble_Gravity (:5859);
// the time discretisation follows AssembleParabolicEquation (:7702, LIQUID)
// and AssembleMixedHyperbolicParabolicEquation (:8504, HEAT).  ISO selects
// an isotropic (diagonal) tensor: the zero off-diagonal products of the
// reference's k,l loop add exact zeros and are skipped, and K is formed on
// its lower triangle and mirrored.
template <int ET, int PROC, bool ISO, bool STORE>
__device__ __forceinline__ void assemble_elem(const AsmArgs& a, int e)
{
    constexpr int NPE = ETraits<ET>::NPE, D = ETraits<ET>::D, NW = ETraits<ET>::NW;
    int nd[NPE];
    double X[NPE], Y[NPE], Z[NPE];
#pragma unroll
    for (int i = 0; i < NPE; i++) {
        nd[i] = a.conn[(int64_t)e * NPE + i];
        X[i] = a.x[nd[i]];
        Y[i] = (D >= 2) ? a.y[nd[i]] : 0.0;
        Z[i] = (D >= 3) ? a.z[nd[i]] : 0.0;
    }
    // CalCoefLaplace: LIQUID tuf*K/mu (fem_ele_std.cpp:2298); HEAT tensor
    double mat[D * D];
#pragma unroll
    for (int i = 0; i < D * D; i++) mat[i] = 0.0;
    if (ISO) {
#pragma unroll
        for (int i = 0; i < D; i++) mat[i * D + i] = a.cond[e];
    } else {
#pragma unroll
        for (int i = 0; i < D * D; i++) mat[i] = a.cond[(int64_t)e * a.cond_stride + i];
    }
    if (PROC == OGSX_LIQUID_FLOW) {
#pragma unroll
        for (int i = 0; i < D * D; i++) mat[i] = a.tuf * mat[i] / a.mu;
    }
    double mass_fac0 = a.mass_coef[e];
    if (PROC == OGSX_HEAT_TRANSPORT) mass_fac0 /= a.tuf;
    const bool do_mass = !a.steady;
    const bool do_adv = (PROC == OGSX_HEAT_TRANSPORT) && a.vel != nullptr;
    const bool do_grav = (PROC == OGSX_LIQUID_FLOW) && a.rho_g != 0.0;

    // Everything else the element needs from memory is requested here, right
    // after the connectivity: the old node values, the incidence positions of
    // the two-pass store and, for tets (whose rule always has 5 points,
    // SamplePointTet5), the Gauss-point velocities.  A thread then waits for
    // memory twice (connectivity, then the rest) instead of once per Gauss
    // point and once more after the loop.  Hexes keep the late loads: their
    // 256 registers leave no room to hold them across the loop; so do tets
    // with a full tensor, whose unrolled loop would need more registers than
    // 2 waves per SIMD leave.
    constexpr bool kEarly = ET != OGSX_HEX8 && STORE;
    constexpr int kFixedGp = ET == OGSX_TET4 && ISO ? 5 : 0;
    double u0[NPE];
    int32_t qpos[NPE];
    if constexpr (kEarly) {
#pragma unroll
        for (int i = 0; i < NPE; i++) {
            u0[i] = a.u_old[nd[i]];
            if (STORE) qpos[i] = a.pos[(int64_t)e * NPE + i];
        }
    }
    double velt[kFixedGp > 0 ? 3 * kFixedGp : 1];
    if constexpr (kFixedGp > 0) {
        if (do_adv) {
#pragma unroll
            for (int k = 0; k < 3 * kFixedGp; k++)
                velt[k] = a.adv_coef * a.vel[(int64_t)e * 3 * kFixedGp + k];
        }
    }

    double M[NPE * NPE], L[NPE * NPE], Adv[NPE * NPE], G[NPE];
#pragma unroll
    for (int i = 0; i < NPE * NPE; i++) {
        M[i] = 0.0;
        L[i] = 0.0;
        Adv[i] = 0.0;
    }
#pragma unroll
    for (int i = 0; i < NPE; i++) G[i] = 0.0;

    // Linear tets and lines have the same local gradients at every Gauss
    // point, so the Jacobian, its determinant and the global gradients are
    // the same at every point: computed once, with the values the per-point
    // computation gives.
    constexpr bool kConstGrad = ET == OGSX_TET4 || ET == OGSX_LINE2;
    double dshp0[D * NPE], det0 = 0.0;
    if constexpr (kConstGrad) det0 = elem_gradients<ET>(X, Y, Z, a.tabdN, dshp0);
    auto gauss_point = [&](const int gp) {
        const double* Ng = a.tabN + gp * NPE;
        double dshp[D * NPE];
        double det;
        if constexpr (kConstGrad) {
            det = det0;
#pragma unroll
            for (int i = 0; i < D * NPE; i++) dshp[i] = dshp0[i];
        } else {
            det = elem_gradients<ET>(X, Y, Z, a.tabdN + gp * D * NPE, dshp);
        }
        double fkt = fabs(det);
        if (D != 3) fkt *= a.geo_area;
        fkt = fkt * a.tabW[gp * 3 + 0];
        if (NW >= 2) fkt = fkt * a.tabW[gp * 3 + 1];
        if (NW >= 3) fkt = fkt * a.tabW[gp * 3 + 2];

        if (do_mass) {
            const double mf = mass_fac0 * fkt;
#pragma unroll
            for (int i = 0; i < NPE; i++)
#pragma unroll
                for (int j = 0; j <= i; j++) M[i * NPE + j] += mf * Ng[i] * Ng[j];
        }
        if (ISO) {
#pragma unroll
            for (int i = 0; i < NPE; i++)
#pragma unroll
                for (int j = 0; j <= i; j++)
#pragma unroll
                    for (int k = 0; k < D; k++)
                        L[i * NPE + j] += fkt * dshp[k * NPE + i] * mat[k * D + k] * dshp[k * NPE + j];
        } else {
#pragma unroll
            for (int i = 0; i < NPE; i++)
#pragma unroll
                for (int j = 0; j < NPE; j++)
#pragma unroll
                    for (int k = 0; k < D; k++)
#pragma unroll
                        for (int l = 0; l < D; l++)
                            L[i * NPE + j] += fkt * dshp[k * NPE + i] * mat[D * k + l] * dshp[l * NPE + j];
        }
        if (do_adv) {
            double vel[D];
#pragma unroll
            for (int k = 0; k < D; k++) {
                if constexpr (kFixedGp > 0)
                    vel[k] = velt[k * kFixedGp + gp];
                else
                    vel[k] = a.adv_coef * a.vel[((int64_t)e * 3 + k) * a.ngp + gp];
            }
#pragma unroll
            for (int i = 0; i < NPE; i++)
#pragma unroll
                for (int j = 0; j < NPE; j++)
#pragma unroll
                    for (int k = 0; k < D; k++)
                        Adv[i * NPE + j] += fkt * Ng[i] * vel[k] * dshp[k * NPE + j];
        }
        if (do_grav) {
            const double f = fkt * a.rho_g;
#pragma unroll
            for (int i = 0; i < NPE; i++)
#pragma unroll
                for (int k = 0; k < D; k++) G[i] -= f * dshp[k * NPE + i] * mat[D * k + D - 1];
        }
    };
    if constexpr (kFixedGp > 0) {
#pragma unroll
        for (int gp = 0; gp < kFixedGp; gp++) gauss_point(gp);
    } else {
        for (int gp = 0; gp < a.ngp; gp++) gauss_point(gp);
    }
    // mirror the lower triangles (CalcMass :3890-3896)
#pragma unroll
    for (int i = 0; i < NPE; i++)
#pragma unroll
        for (int j = 0; j < i; j++) {
            M[j * NPE + i] = M[i * NPE + j];
            if (ISO) L[j * NPE + i] = L[i * NPE + j];
        }

    const double dt_inverse = 1.0 / a.dt;
    if constexpr (!kEarly) {
#pragma unroll
        for (int i = 0; i < NPE; i++) u0[i] = a.u_old[nd[i]];
    }
    double lhs[NPE * NPE], rhs[NPE];
    if (PROC == OGSX_LIQUID_FLOW) {
        double relax0 = a.theta;
        if (relax0 < DBL_MIN) relax0 = 1.0;
        const double relax1 = 1.0 - relax0;
#pragma unroll
        for (int i = 0; i < NPE; i++) {
            double val = 0.0;
#pragma unroll
            for (int j = 0; j < NPE; j++) {
                const double m = M[i * NPE + j] * dt_inverse;
                lhs[i * NPE + j] = m + L[i * NPE + j] * relax0;
                val += (m - L[i * NPE + j] * relax1) * u0[j];
            }
            rhs[i] = val;
        }
    } else {
        const double th = a.theta, fl = -(1.0 - th);
#pragma unroll
        for (int i = 0; i < NPE; i++) {
            double val = 0.0;
#pragma unroll
            for (int j = 0; j < NPE; j++) {
                const double m = M[i * NPE + j] * dt_inverse;
                lhs[i * NPE + j] = (m + L[i * NPE + j] * th) + Adv[i * NPE + j] * th;
                val += ((m + L[i * NPE + j] * fl) + Adv[i * NPE + j] * fl) * u0[j];
            }
            rhs[i] = val;
        }
    }
    if (STORE) {
        // two-pass assembly: each element row goes, with its right-hand
        // side (and gravity term), to its node's incidence list (node-major),
        // where k_assemble_rows reads it in sequence
#pragma unroll
        for (int i = 0; i < NPE; i++) {
            const int64_t q = kEarly ? qpos[i] : a.pos[(int64_t)e * NPE + i];
            double* rec = a.eb + q * NPE;
#pragma unroll
            for (int j = 0; j < NPE; j++) rec[j] = lhs[i * NPE + j];
            if (rb_stride<ET>(do_grav) == 2) {  // one 16-byte store
                a.rb[2 * q] = rhs[i];
                a.rb[2 * q + 1] = do_grav ? G[i] : 0.0;
            } else {
                a.rb[q] = rhs[i];
            }
        }
        return;
    }
    // atomic-free scatter: elements of one colour own disjoint rows
    const int32_t* sl = a.slot + (int64_t)e * NPE * NPE;
#pragma unroll
    for (int i = 0; i < NPE; i++) {
        a.b[nd[i]] += rhs[i];
        if (do_grav) a.b[nd[i]] += G[i];
#pragma unroll
        for (int j = 0; j < NPE; j++) a.A[sl[i * NPE + j]] += lhs[i * NPE + j];
    }
}

template <int ET, int PROC, bool ISO, bool STORE>
__global__ __launch_bounds__(64) void k_assemble(AsmArgs a, int off, int cnt)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= cnt) return;
    const int e = STORE ? off + t : a.color_elems[off + t];
    assemble_elem<ET, PROC, ISO, STORE>(a, e);
}

// Second pass of the two-pass assembly: G lanes per row r, lane c owning
// entry c of the row.  The rows of the element matrices that belong to node
// r sit in node r's incidence order (ascending element order, the order of
// the reference's element loop, FEM/rf_pcs.cpp:5741-5762); the lanes walk
// them together, record by record, and lane c adds the one value of each
// record whose column position (loc) is c.  Every entry is therefore summed
// from A's current value in the same order as A += contribution element by
// element, and no two lanes touch one entry (no colouring, no atomics).  The
// row's lanes load its records together, one record per lane, through LDS
// (one contiguous stretch per row), so a wave holds a few cache lines at a
// time, each line is fetched once, and a row costs one memory round trip per
// G records.  Lane 0 adds the element right-hand sides, then the gravity
// terms, to b[r] as the scatter does.
constexpr int kAsmRowMax = 32;      // longest row the two-pass assembly takes
constexpr int kAsmRowThreads = 256; // threads per block of the row pass

// lanes per row for rows of at most row_max entries
__host__ __device__ constexpr int asm_row_lanes(int row_max) { return row_max <= 16 ? 16 : 32; }

// Column positions of one record (NPE bytes at q NPE) as one word; words
// past the row's last record read as 0xff bytes, which match no lane.
template <int NPE>
using LocWord = typename std::conditional<NPE == 2, uint16_t,
                typename std::conditional<NPE == 4, uint32_t, uint64_t>::type>::type;

template <int NPE>
__device__ __forceinline__ int loc_match(uint64_t w, int c)
{
    int jm = -1;
#pragma unroll
    for (int j = 0; j < NPE; j++) jm = (int)((w >> (8 * j)) & 0xff) == c ? j : jm;
    return jm;
}

// LDS written by some lanes of a wavefront and read by others of the same
// wavefront: the wavefront's LDS accesses complete in order once waited for.
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The row's records are taken G at a time: lane t loads record t of the
// batch (a contiguous stretch of the row's records across the lanes) and its
// column positions into LDS, then lane c walks the batch's records in order.
// The lanes of a row sit in one wavefront; every lane of a wavefront runs
// every batch any of its rows needs (no early exit), so the wavefront-level
// syncs are uniform.  lds holds kAsmRowThreads records, ldw as many words.
template <int ET, bool GRAV, int G>
__device__ __forceinline__ void assemble_row(const AsmArgs& a, const int32_t* inc_ptr,
                                             const uint8_t* loc, const int32_t* ptr, int n,
                                             int64_t r, double* lds, uint64_t* ldw)
{
    constexpr int NPE = ETraits<ET>::NPE;
    constexpr int RS = NPE + (GRAV ? 2 : 1);
    using W = LocWord<NPE>;
    static_assert(64 % G == 0, "a row's lanes must sit in one wavefront");
    const int tid = threadIdx.x;
    const int c = tid % G, g0 = tid - c;
    const bool live = r < n;
    const OGSX_GL double* eb = gl(a.eb);
    const OGSX_GL double* rb = gl(a.rb);
    const OGSX_GL W* lw = (const OGSX_GL W*)gl(loc);
    OGSX_GL double* A = gl(a.A);
    int p0 = 0, len = 0;
    int64_t q0 = 0, q1 = 0;
    double acc = 0.0, br = 0.0;
    if (live) {
        p0 = ptr[r];
        len = ptr[r + 1] - p0;
        q0 = inc_ptr[r];
        q1 = inc_ptr[r + 1];
        if (c < len) acc = A[p0 + c];
        if (c == 0) br = gl(a.b)[r];
    }
    for (int64_t base = 0; __any(live && q0 + base < q1); base += G) {
        const int64_t q = q0 + base + c;
        if (live && q < q1) {
            double v[RS];
#pragma unroll
            for (int j = 0; j < NPE; j++) v[j] = eb[q * NPE + j];
            v[NPE] = rb[rb_stride<ET>(GRAV) * q];
            if (GRAV) v[NPE + 1] = rb[2 * q + 1];
            const uint64_t w = lw[q];
#pragma unroll
            for (int j = 0; j < RS; j++) lds[tid * RS + j] = v[j];
            ldw[tid] = w;
        }
        wave_lds_sync();
        const int64_t rem = q1 - q0 - base;
        const int cnt = live && rem > 0 ? (int)(rem < G ? rem : G) : 0;
        for (int k = 0; k < cnt; k++) {
            const double* rec = lds + (g0 + k) * RS;
            const int jm = c < len ? loc_match<NPE>(ldw[g0 + k], c) : -1;
            if (jm >= 0) acc += rec[jm];
            if (c == 0) {
                br += rec[NPE];
                if (GRAV) br += rec[NPE + 1];
            }
        }
        wave_lds_sync();
    }
    if (live) {
        if (c < len) A[p0 + c] = acc;
        if (c == 0) gl(a.b)[r] = br;
    }
}

template <int ET, bool GRAV, int G>
__global__ __launch_bounds__(kAsmRowThreads) void k_assemble_rows(AsmArgs a, const int32_t* inc_ptr,
                                                                  const uint8_t* loc,
                                                                  const int32_t* ptr, int n)
{
    __shared__ double lds[kAsmRowThreads * (ETraits<ET>::NPE + (GRAV ? 2 : 1))];
    __shared__ uint64_t ldw[kAsmRowThreads];
    const int64_t r = ((int64_t)blockIdx.x * kAsmRowThreads + threadIdx.x) / G;
    assemble_row<ET, GRAV, G>(a, inc_ptr, loc, ptr, n, r, lds, ldw);
}

// Two-pass assembly tables, per incidence position q (node-major, inc_ei[q] =
// e npe + i): pos[e npe + i] = q, and loc[q npe + j] = the position of entry
// (i, j) of element e in its CRS row.
__global__ __launch_bounds__(256) void k_row_positions(int64_t count, int npe, const int32_t* inc_ei,
                                                        const int32_t* conn, const int32_t* slot,
                                                        const int32_t* ptr, int32_t* pos,
                                                        uint8_t* loc)
{
    for (int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x; q < count;
         q += (int64_t)gridDim.x * 256) {
        const int64_t ei = inc_ei[q];
        const int r = conn[ei];
        pos[ei] = (int32_t)q;
        for (int j = 0; j < npe; j++) loc[q * npe + j] = (uint8_t)(slot[ei * npe + j] - ptr[r]);
    }
}

// ------------------------------------------------------------------------
// Gauss-point Darcy velocity of a LIQUID_FLOW system (one thread per element)
// ------------------------------------------------------------------------
struct VelArgs {
    const int32_t* conn;
    const double* x;
    const double* y;
    const double* z;
    const double* cond;
    const double* tabdN;  // [ngp][edim][npe]
    const double* u;      // nodal pressure (head) at the new time level
    double* vel;          // out: [E][3][ngp]
    int n_elem;
    int cond_stride;
    int ngp;
    int pad;
    double tuf, mu, rho_g;
};

// CFiniteElementStd::Cal_Velocity (FEM/fem_ele_std.cpp:6554-6868) for a
// single-phase flow: at each Gauss point g = grad N . u (:6655-6661), plus
// rho*g on the last component when gravity is on (:6748), then
// v_i = -sum_j mat_ij g_j / tuf with mat = tuf*K/mu from CalCoefLaplace
// (:2298-2301), components >= dim left zero.  ISO skips the zero
// off-diagonal products, which subtract exact zeros.  Linear tetrahedra
// have the same local gradients at all Gauss points, so their velocity is
// computed once and repeated (the same values the per-point loop gives).
//
// STAGED: the block's velocities ([elements of the block][3][ngp], one
// contiguous range of the output) are gathered in LDS and written with
// coalesced stores; per-thread stores of 3*ngp consecutive doubles leave
// every wave store touching 64 partial cache lines.  Needs
// blockDim.x * 3 * ngp doubles of dynamic LDS.
template <int ET, bool ISO, bool STAGED>
__global__ __launch_bounds__(256) void k_gauss_velocity(VelArgs a)
{
    constexpr int NPE = ETraits<ET>::NPE, D = ETraits<ET>::D;
    extern __shared__ double vstage[];
    const int e0 = blockIdx.x * blockDim.x;
    const int e = e0 + threadIdx.x;
    const int ngp = a.ngp, per = 3 * ngp;
    if (e < a.n_elem) {
        double X[NPE], Y[NPE], Z[NPE], nv[NPE];
#pragma unroll
        for (int i = 0; i < NPE; i++) {
            const int nd = a.conn[(int64_t)e * NPE + i];
            X[i] = a.x[nd];
            Y[i] = (D >= 2) ? a.y[nd] : 0.0;
            Z[i] = (D >= 3) ? a.z[nd] : 0.0;
            nv[i] = a.u[nd];
        }
        double mat[D * D];
#pragma unroll
        for (int i = 0; i < D * D; i++) mat[i] = 0.0;
        if (ISO) {
#pragma unroll
            for (int i = 0; i < D; i++) mat[i * D + i] = a.tuf * a.cond[e] / a.mu;
        } else {
#pragma unroll
            for (int i = 0; i < D * D; i++)
                mat[i] = a.tuf * a.cond[(int64_t)e * a.cond_stride + i] / a.mu;
        }
        double* out = STAGED ? vstage + threadIdx.x * per : a.vel + (int64_t)e * per;
        const int n_eval = (ET == OGSX_TET4) ? 1 : ngp;
        double v[3] = {0.0, 0.0, 0.0};
        for (int gp = 0; gp < ngp; gp++) {
            if (gp < n_eval) {
                double dshp[D * NPE];
                elem_gradients<ET>(X, Y, Z, a.tabdN + gp * D * NPE, dshp);
                double g[D];
#pragma unroll
                for (int i = 0; i < D; i++) {
                    double s = 0.0;
#pragma unroll
                    for (int j = 0; j < NPE; j++) s += nv[j] * dshp[i * NPE + j];
                    g[i] = s;
                }
                if (a.rho_g != 0.0) g[D - 1] += a.rho_g;
#pragma unroll
                for (int i = 0; i < D; i++) {
                    double vi = 0.0;
#pragma unroll
                    for (int j = 0; j < D; j++) {
                        if (ISO && j != i) continue;
                        vi -= mat[D * i + j] * g[j] / a.tuf;
                    }
                    v[i] = vi;
                }
            }
#pragma unroll
            for (int i = 0; i < 3; i++) out[i * ngp + gp] = v[i];
        }
    }
    if constexpr (STAGED) {
        __syncthreads();
        const int cnt = min((int)blockDim.x, a.n_elem - e0);
        const int total = cnt * per;
        double* dst = a.vel + (int64_t)e0 * per;
        for (int i = threadIdx.x; i < total; i += blockDim.x) dst[i] = vstage[i];
    }
}

// ------------------------------------------------------------------------
// sparse matrix-vector product (one system; autotuner, bench, ogsx_spmv)
// ------------------------------------------------------------------------
template <int FMT, typename VT, typename XT>
__global__ __launch_bounds__(kBlock) void k_spmv(MatView<VT> m, const XT* __restrict__ x,
                                                 XT* __restrict__ y)
{
    using ACC = XT;
    OGSX_GL XT* __restrict__ yg = gl(y);
    spmv_rows<FMT, VT, XT, ACC>(gmat(m), gl(x), gridDim.x, blockIdx.x,
                                [&](int r, ACC s) { yg[r] = (XT)s; });
}

// ------------------------------------------------------------------------
// Krylov kernels.  Every kernel reads its system from a device descriptor
// (SysDev); a group launch covers many systems, each with its own G blocks,
// partial sums and state, so a system's arithmetic is the same alone and in
// any mix of other systems.
// ------------------------------------------------------------------------
template <typename T>
struct KVec {
    T* x;
    T* b;
    T* d;   // Jacobi diagonal (entries with |a_ii| < DBL_MIN are 1)
    T* r;
    T* p;
    T* s;
    T* z;
    T* r0;
    T* v;
    T* t;
    T* ph;
    T* sh;
};

// KVec with global-address-space pointers.
template <typename T>
struct GVec {
    OGSX_GL T* x;
    OGSX_GL T* b;
    OGSX_GL T* d;
    OGSX_GL T* r;
    OGSX_GL T* p;
    OGSX_GL T* s;
    OGSX_GL T* z;
    OGSX_GL T* r0;
    OGSX_GL T* v;
    OGSX_GL T* t;
    OGSX_GL T* ph;
    OGSX_GL T* sh;
};

template <typename T>
__device__ __forceinline__ GVec<T> gvec(const KVec<T>& k)
{
    return GVec<T>{gl(k.x), gl(k.b), gl(k.d),  gl(k.r), gl(k.p),  gl(k.s),
                   gl(k.z), gl(k.r0), gl(k.v), gl(k.t), gl(k.ph), gl(k.sh)};
}

// Per-solve parameters of one system.  The recurrence fields are written by
// the host (FP64 solves) or by k_ir_decide (the FP32 recurrence of IR); the
// ir_* fields by the host when an IR solve starts.
struct KrylovParams {
    double tol;         // stopping test ||r||/||b|| < tol
    double bnorm_prev;  // CheckNormRHS reference (0 switches the test off)
    double upd_thresh;  // reliable updates: stop the recurrence when ||r|| < it
    double scale;       // IR: r32 = scale * r64
    double unscale;     // IR: x64 += unscale * x32
    int max_iter;
    int jacobi;         // 1 Jacobi, 0 no preconditioner
    double ir_tol;        // FP64 stopping test of the refinement
    double ir_bnorm_prev; // CheckNormRHS reference of the refinement
    double ir_delta;      // restart: FP32 solve tolerance; reliable: update drop
    int ir_max_iter;
    int ir_max_outer;     // restart: refinement steps
    int ir_reliable;      // 1 reliable-update CG, 0 restarted refinement
    int ir_scaled;        // 1: the FP32 recurrence runs on S A S with
                          // S = diag(1/sqrt|a_ii|) (single-reduction CG)
};

// Refinement state of an IR solve, advanced by k_ir_decide.
struct IRState {
    double bnorm;   // FP64 ||b||
    double error;   // FP64 ||b - A x|| / ||b|| at the last residual step
    double maxr;    // reliable updates: scaled ||r|| at the last update
    int started;    // an FP32 recurrence has run
    int status;     // -1 running; else finished: 1 converged, 0 max_iter,
                    // 2 skipped by CheckNormRHS, 3 breakdown
    int action;     // of the last decision: 0 none, 1 new recurrence, 2 resume
    int total;      // Krylov iterations so far
    int updates;    // residual steps after the first
    int pad;
};

struct SysDev {
    MatView<double> m64;
    MatView<float> m32;
    KVec<double> v64;
    KVec<float> v32;
    double* A;           // CRS values (FP64)
    float* A32;          // CRS values (FP32, IR)
    double* sval64;      // SELL values
    float* sval32;
    const int* diagpos;
    const int* sell_src;
    const int* tri_src;  // CRS position of every slice-triple entry (-1 = zero)
    const int* dia_src;  // ... of every symmetric-diagonal entry (-1 = zero)
    double* r64;         // IR residual
    double* part;        // partial sums [16][G]
    double* red;         // reduced scalars
    KrylovState* st;
    KrylovParams* prm;
    IRState* irs;        // IR systems
    int n, G, nnz, npad, npad3;
};

template <typename VT>
__device__ __forceinline__ const MatView<VT>& mv_of(const SysDev& S);
template <>
__device__ __forceinline__ const MatView<double>& mv_of<double>(const SysDev& S) { return S.m64; }
template <>
__device__ __forceinline__ const MatView<float>& mv_of<float>(const SysDev& S) { return S.m32; }
template <typename T>
__device__ __forceinline__ const KVec<T>& kv_of(const SysDev& S);
template <>
__device__ __forceinline__ const KVec<double>& kv_of<double>(const SysDev& S) { return S.v64; }
template <>
__device__ __forceinline__ const KVec<float>& kv_of<float>(const SysDev& S) { return S.v32; }

// Launch geometry: blk2sys = nullptr for one system (grid = its G blocks);
// otherwise blk2sys maps a launched block to its system and blk0[i] is
// system i's first block.  Kernels with one block per system use blk2sys as
// the block -> system list.  (Launching fewer blocks that each work through
// several of a system's G row chunks was measured on a 256-system ensemble:
// 4-7% slower, and the loop costs ~10 VGPRs in the SpMV kernels.)
struct Grp {
    const SysDev* sys;
    const int* blk2sys;
    const int* blk0;
};

struct Ctx {
    const SysDev* S;
    int G, b;  // the block is chunk b of the system's G
};

// The system and chunk of virtual block vb (a launched block, or one of the
// blocks a persistent launch works through).
__device__ __forceinline__ Ctx ctx_of(const Grp& g, int vb)
{
    if (!g.blk2sys) return Ctx{g.sys, g.sys->G, vb};
    const int si = g.blk2sys[vb];
    const SysDev* S = g.sys + si;
    return Ctx{S, S->G, vb - g.blk0[si]};
}

__device__ __forceinline__ Ctx ctx_of(const Grp& g) { return ctx_of(g, (int)blockIdx.x); }

__device__ __forceinline__ const SysDev* sys_of_block(const Grp& g)
{
    return g.sys + (g.blk2sys ? g.blk2sys[blockIdx.x] : (int)blockIdx.x);
}

// Vector body over the rows of the block's chunk.
template <typename Body>
__device__ __forceinline__ void vec_vblocks(const Ctx& c, int n, Body&& body)
{
    vec_chunk(n, c.G, c.b, body);
}

// Jacobi diagonal: Precond_Jacobi's a_ii with |a_ii| < DBL_MIN -> 1
// (FEM/matrix_class.cpp:1786-1821); no preconditioner -> 1.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_diag(Grp g)
{
    const Ctx c = ctx_of(g);
    const SysDev& S = *c.S;
    OGSX_GL T* d = gl(kv_of<T>(S).d);
    const OGSX_GL int* dp = gl(S.diagpos);
    const OGSX_GL double* A = gl(S.A);
    const int jac = gl(S.prm)->jacobi;
    vec_vblocks(c, S.n, [&](int i) {
        double v = (jac && dp[i] >= 0) ? A[dp[i]] : 1.0;
        if (fabs(v) < DBL_MIN) v = 1.0;
        d[i] = (T)v;
    });
}

// CRS values -> SELL-64 values (padding 0) of one precision, in the SELL
// layout (tri = 0) or the slice-triple layout (tri = 1).
// Value copies into a slice layout: store(p, A[src[p]]) (0 where src[p] < 0)
// for the block's chunk of the layout's positions, 8 positions per thread at
// a time so that a thread's index loads, then its gathers of A, are each one
// memory round trip.
template <typename Store>
__device__ __forceinline__ void copy_to_layout(const Ctx& c, int npos, const OGSX_GL int* src,
                                               const OGSX_GL double* A, Store&& store)
{
    constexpr int U = 8;
    int beg, end;
    chunk_of(npos, c.G, c.b, beg, end);
    for (int i = beg + threadIdx.x; i < end; i += U * kBlock) {
        int k[U];
#pragma unroll
        for (int u = 0; u < U; u++) k[u] = i + u * kBlock < end ? src[i + u * kBlock] : -1;
        double v[U];
#pragma unroll
        for (int u = 0; u < U; u++) v[u] = k[u] >= 0 ? A[k[u]] : 0.0;
#pragma unroll
        for (int u = 0; u < U; u++)
            if (i + u * kBlock < end) store(i + u * kBlock, v[u]);
    }
}

template <typename VT>
__global__ __launch_bounds__(kBlock) void k_to_sell_g(Grp g, int tri)
{
    const Ctx c = ctx_of(g);
    const SysDev& S = *c.S;
    OGSX_GL VT* sv;
    if constexpr (sizeof(VT) == 8)
        sv = gl(S.sval64);
    else
        sv = gl(S.sval32);
    const OGSX_GL int* src = gl(tri ? S.tri_src : S.sell_src);
    const OGSX_GL double* A = gl(S.A);
    copy_to_layout(c, tri ? S.npad3 : S.npad, src, A, [&](int p, double v) { sv[p] = (VT)v; });
}

// Both value copies of an IR system in one pass when they share a layout:
// A read once, the FP64 and the FP32 value written.
__global__ __launch_bounds__(kBlock) void k_to_sell_both_g(Grp g, int tri)
{
    const Ctx c = ctx_of(g);
    const SysDev& S = *c.S;
    OGSX_GL double* sv64 = gl(S.sval64);
    OGSX_GL float* sv32 = gl(S.sval32);
    const OGSX_GL int* src = gl(tri ? S.tri_src : S.sell_src);
    const OGSX_GL double* A = gl(S.A);
    copy_to_layout(c, tri ? S.npad3 : S.npad, src, A, [&](int p, double v) {
        sv64[p] = v;
        sv32[p] = (float)v;
    });
}

// Symmetric diagonal copy of A (see spmv_symdia; T = float: the FP32
// matrix, T = double: the FP64 one) through its position map: one thread
// per row, its K + 1 entries 8 at a time (map reads and value writes
// coalesced across the threads; the gathered values of one row share its
// few cache lines).  Taken position by position, as the slice-layout copies
// are, every diagonal pass fetched A's lines anew: 561 us against 331 us on
// 128^3 (FP32).  ALSO32 (with T = double): the FP32 copy written in the
// same pass, from the same gathered values.
template <typename T, bool ALSO32 = false>
__global__ __launch_bounds__(kBlock) void k_to_dia_g(Grp g)
{
    const Ctx c = ctx_of(g);
    const SysDev& S = *c.S;
    const MatView<T>& M = mv_of<T>(S);
    OGSX_GL T* dv = (OGSX_GL T*)gl(M.dia_val);
    OGSX_GL float* dv32 = ALSO32 ? (OGSX_GL float*)gl(S.m32.dia_val) : nullptr;
    const OGSX_GL int* src = gl(S.dia_src);
    const OGSX_GL double* A = gl(S.A);
    const int n = S.n, K1 = M.dia_k + 1;
    constexpr int U = 8;
    vec_vblocks(c, n, [&](int i) {
        for (int k0 = 0; k0 < K1; k0 += U) {
            int p[U];
#pragma unroll
            for (int u = 0; u < U; u++) p[u] = k0 + u < K1 ? src[(int64_t)(k0 + u) * n + i] : -1;
            double v[U];
#pragma unroll
            for (int u = 0; u < U; u++) v[u] = p[u] >= 0 ? A[p[u]] : 0.0;
#pragma unroll
            for (int u = 0; u < U; u++)
                if (k0 + u < K1) {
                    dv[(int64_t)(k0 + u) * n + i] = (T)v[u];
                    if constexpr (ALSO32) dv32[(int64_t)(k0 + u) * n + i] = (float)v[u];
                }
        }
    });
}

__global__ __launch_bounds__(kBlock) void k_cast_f32_g(Grp g)
{
    const Ctx c = ctx_of(g);
    const SysDev& S = *c.S;
    OGSX_GL float* A32 = gl(S.A32);
    const OGSX_GL double* A = gl(S.A);
    vec_vblocks(c, S.nnz, [&](int k) { A32[k] = (float)A[k]; });
}

// IR event kernels run only for the systems whose last k_ir_decide chose
// `action` (1 new recurrence, 2 resume); FP64 solves pass want = 0.
__device__ __forceinline__ bool skip_action(const SysDev& S, int want)
{
    return want != 0 && gl(S.irs)->action != want;
}

// ---- Jacobi-preconditioned CG (Linear_EQS::CG, FEM/equation_class.cpp:1129)

// s = A x; r = b - s; z = r/d; p = z; partials (b,b), (r,r), (r,z)
template <int FMT, typename VT, typename T>
__global__ __launch_bounds__(kBlock) void k_cg_init(Grp g, int want)
{
    const Ctx c = ctx_of(g);
    const SysDev& S = *c.S;
    if (skip_action(S, want)) return;
    const GMat<VT> m = gmat(mv_of<VT>(S));
    const GVec<T> v = gvec(kv_of<T>(S));
    {
        const int b = c.b;
        double acc[3] = {0.0, 0.0, 0.0};
        spmv_rows<FMT, VT, T, T>(m, v.x, c.G, b, [&](int i, T ax) {
            const T ri = v.b[i] - ax;
            const T zi = ri / v.d[i];
            v.r[i] = ri;
            v.p[i] = zi;
            acc[0] += (double)v.b[i] * (double)v.b[i];
            acc[1] += (double)ri * (double)ri;
            acc[2] += (double)ri * (double)zi;
        });
        store_partials<3>(acc, gl(S.part), c.G, b);
    }
}

// One block per system: CheckNormRHS (:1104), initial residual test, rr.
__global__ void k_cg_init_final(Grp g, int want)
{
    const SysDev& S = *sys_of_block(g);
    if (skip_action(S, want)) return;
    double s[3];
    reduce_partials<3>(gl(S.part), S.G, s);
    if (threadIdx.x != 0) return;
    OGSX_GL KrylovState* st = gl(S.st);
    const KrylovParams P = *S.prm;  // one block per system: a generic load is fine
    const double bnorm_new = sqrt(s[0]);
    st->iter = 0;
    st->sr_it = 0;
    st->max_iter = P.max_iter;
    st->done = 0;
    st->pend = 0;
    if (P.bnorm_prev > 0.0 && bnorm_new / P.bnorm_prev < P.tol) {
        st->error = bnorm_new / P.bnorm_prev;
        st->bnorm = bnorm_new;
        st->done = 4;
        return;
    }
    st->bnorm = bnorm_new;
    if (bnorm_new < DBL_MIN) {
        st->error = 0.0;
        st->done = 4;
        return;
    }
    st->error = sqrt(s[1]) / bnorm_new;
    if (st->error < P.tol) {
        st->done = 1;
        return;
    }
    st->rr = s[2];
    st->rr_old = s[2];
}

// K1: s = A p; partial (p,s)
template <int FMT, typename VT, typename T>
__device__ __forceinline__ void cg_spmv_body(const Grp& g, int vb)
{
    const Ctx c = ctx_of(g, vb);
    const SysDev& S = *c.S;
    if (gl(S.st)->done) return;
    const GMat<VT> m = gmat(mv_of<VT>(S));
    const GVec<T> v = gvec(kv_of<T>(S));
    {
        const int b = c.b;
        double acc[1] = {0.0};
        spmv_rows<FMT, VT, T, T>(m, v.p, c.G, b, [&](int i, T ax) {
            v.s[i] = ax;
            acc[0] += (double)v.p[i] * (double)ax;
        });
        store_partials<1>(acc, gl(S.part), c.G, b);
    }
}

template <int FMT, typename VT, typename T>
__global__ __launch_bounds__(kBlock) void k_cg_spmv(Grp g)
{
    cg_spmv_body<FMT, VT, T>(g, (int)blockIdx.x);
}

// K2: alpha = rr/(p,s); x += alpha p; r -= alpha s; partials (r,r), (r/d, r).
// VNT: the vector streams bypass the caches (non-temporal loads and stores),
// chosen by the host when the matrix of the launch cannot stay in the
// Infinity Cache anyway: the vectors then do not evict the part of it that
// can (128^3 FP64: the solve's SpMV 100.6 -> 87.5 us, the job 549 -> 519
// ms).  Where matrix and vectors fit, cached streams are faster.
//
// SMALL: a member of the launch has chunks of fewer than U kBlock rows
// (systems of up to 262,144 rows), whose threads would otherwise take their
// 1-3 rows one memory round trip at a time.  Such a chunk's rows are then
// loaded as one batch, before the partials are reduced, with indices past
// the chunk clamped to its last row (loaded, not used).  Rows are updated
// and summed in the same order either way, so both forms give the same
// bits; the host picks SMALL only when some running member needs it (the
// 256-instance ensemble: 5-7 % faster; on 128^3 hexes the SMALL form ran
// 0.6-0.9 % slower, though it takes the same path there).
template <typename T, bool VNT, bool SMALL = false>
__device__ __forceinline__ void cg_update_body(const Grp& g, int vb)
{
    const Ctx c = ctx_of(g, vb);
    const SysDev& S = *c.S;
    OGSX_GL KrylovState* st = gl(S.st);
    if (st->done) return;
    const GVec<T> v = gvec(kv_of<T>(S));
    OGSX_GL T* __restrict__ x = v.x;
    OGSX_GL T* __restrict__ r = v.r;
    const OGSX_GL T* __restrict__ p = v.p;
    const OGSX_GL T* __restrict__ sv = v.s;
    const OGSX_GL T* __restrict__ d = v.d;
    const int b = c.b;
    int beg, end;
    chunk_of(S.n, c.G, b, beg, end);
    constexpr int U = 4;
    int i = beg + threadIdx.x;
    // the first batch of the stream does not depend on alpha: its loads are
    // issued before the partials are reduced, so their latency overlaps
    // partials first, then the first batch of rows (see load_partials)
    PartialLoads<1> pl;
    load_partials<1>(gl(S.part), c.G, pl);
    T xi[U], pi[U], ri[U], si[U], di[U];
    const bool first = i + (U - 1) * kBlock < end;
    // a chunk of one full batch plus fewer than kBlock rows (1024-1280
    // rows: 2048 chunks of a 2.1-2.6 M-row system) loads its extra row per
    // thread with the batch, so the chunk costs one memory round trip
    const bool tail = !SMALL && first && end - beg <= (U + 1) * kBlock && i + U * kBlock < end;
    T xt = 0, pt = 0, rt = 0, st1 = 0, dt = 1;
    if (first) {
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int k = i + u * kBlock;
            xi[u] = ld<VNT>(x + k);
            pi[u] = ld<VNT>(p + k);
            ri[u] = ld<VNT>(r + k);
            si[u] = ld<VNT>(sv + k);
            di[u] = ld<VNT>(d + k);
        }
        if (tail) {
            const int k = i + U * kBlock;
            xt = ld<VNT>(x + k);
            pt = ld<VNT>(p + k);
            rt = ld<VNT>(r + k);
            st1 = ld<VNT>(sv + k);
            dt = ld<VNT>(d + k);
        }
    } else if (SMALL && i < end) {
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int k = min(i + u * kBlock, end - 1);
            xi[u] = ld<VNT>(x + k);
            pi[u] = ld<VNT>(p + k);
            ri[u] = ld<VNT>(r + k);
            si[u] = ld<VNT>(sv + k);
            di[u] = ld<VNT>(d + k);
        }
    }
    double ps[1];
    finish_partials<1>(gl(S.part), c.G, pl, ps);
    const double rr = st->rr;
    const T alpha = (T)(rr / ps[0]);
    {
        double acc[2] = {0.0, 0.0};
        if (SMALL && !first && i < end) {  // the chunk's one, partial batch
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int k = i + u * kBlock;
                if (k < end) {
                    store_hint<VNT>(x + k, (T)(xi[u] + alpha * pi[u]));
                    const T rn = ri[u] - alpha * si[u];
                    store_hint<VNT>(r + k, rn);
                    const T zi = rn / di[u];
                    acc[0] += (double)rn * (double)rn;
                    acc[1] += (double)zi * (double)rn;
                }
            }
            i = end;
        }
        for (bool loaded = first; i + (U - 1) * kBlock < end; i += U * kBlock, loaded = false) {
            if (!loaded) {
#pragma unroll
                for (int u = 0; u < U; u++) {
                    const int k = i + u * kBlock;
                    xi[u] = ld<VNT>(x + k);
                    pi[u] = ld<VNT>(p + k);
                    ri[u] = ld<VNT>(r + k);
                    si[u] = ld<VNT>(sv + k);
                    di[u] = ld<VNT>(d + k);
                }
            }
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int k = i + u * kBlock;
                store_hint<VNT>(x + k, (T)(xi[u] + alpha * pi[u]));
                const T rn = ri[u] - alpha * si[u];
                store_hint<VNT>(r + k, rn);
                const T zi = rn / di[u];
                acc[0] += (double)rn * (double)rn;
                acc[1] += (double)zi * (double)rn;
            }
        }
        if (tail) {  // the batch loop ran once: i is the thread's extra row
            store_hint<VNT>(x + i, (T)(xt + alpha * pt));
            const T rn = rt - alpha * st1;
            store_hint<VNT>(r + i, rn);
            const T zi = rn / dt;
            acc[0] += (double)rn * (double)rn;
            acc[1] += (double)zi * (double)rn;
            i = end;
        }
        for (; i < end; i += kBlock) {
            x[i] += alpha * p[i];
            const T rn = r[i] - alpha * sv[i];
            r[i] = rn;
            const T zi = rn / d[i];
            acc[0] += (double)rn * (double)rn;
            acc[1] += (double)zi * (double)rn;
        }
        store_partials<2>(acc, gl(S.part) + 4 * (size_t)c.G, c.G, b);
    }
    if (c.b == 0 && threadIdx.x == 0) {
        st->rr_old = rr;
        st->alpha = rr / ps[0];
    }
}

template <typename T, bool VNT, bool SMALL>
__global__ __launch_bounds__(kBlock) void k_cg_update(Grp g)
{
    cg_update_body<T, VNT, SMALL>(g, (int)blockIdx.x);
}

// K3: error test; beta = (z,r)/rr_old; p = r/d + beta p.  With reliable
// updates (upd_thresh > 0) the recurrence stops with done = 5 once ||r|| <
// upd_thresh, leaving p for the resumed iteration.  Every block takes the
// same branch (same reduced values), so a stop leaves p untouched everywhere.
template <typename T, bool VNT, bool SMALL = false>
__device__ __forceinline__ void cg_direction_body(const Grp& g, int vb)
{
    const Ctx c = ctx_of(g, vb);
    const SysDev& S = *c.S;
    OGSX_GL KrylovState* st = gl(S.st);
    if (st->done) return;
    const GVec<T> v = gvec(kv_of<T>(S));
    OGSX_GL T* __restrict__ p = v.p;
    const OGSX_GL T* __restrict__ r = v.r;
    const OGSX_GL T* __restrict__ d = v.d;
    int beg, end;
    chunk_of(S.n, c.G, c.b, beg, end);
    constexpr int U = 4;
    int i = beg + threadIdx.x;
    // first batch loaded before the reduction (independent of beta)
    PartialLoads<2> pl;  // partials first, then the first batch of rows
    load_partials<2>(gl(S.part) + 4 * (size_t)c.G, c.G, pl);
    T pi[U], ri[U], di[U];
    const bool first = i + (U - 1) * kBlock < end;
    // the extra row of a chunk of 1024-1280 rows, as in the update
    const bool tail = !SMALL && first && end - beg <= (U + 1) * kBlock && i + U * kBlock < end;
    T pt = 0, rt = 0, dt = 1;
    if (first) {
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int k = i + u * kBlock;
            pi[u] = ld<VNT>(p + k);
            ri[u] = ld<VNT>(r + k);
            di[u] = ld<VNT>(d + k);
        }
        if (tail) {
            const int k = i + U * kBlock;
            pt = ld<VNT>(p + k);
            rt = ld<VNT>(r + k);
            dt = ld<VNT>(d + k);
        }
    } else if (SMALL && i < end) {  // a chunk's one, partial batch (as in the update)
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int k = min(i + u * kBlock, end - 1);
            pi[u] = ld<VNT>(p + k);
            ri[u] = ld<VNT>(r + k);
            di[u] = ld<VNT>(d + k);
        }
    }
    double s[2];
    finish_partials<2>(gl(S.part) + 4 * (size_t)c.G, c.G, pl, s);
    const double rnorm = sqrt(s[0]);
    const double error = rnorm / st->bnorm;
    const double tol = gl(S.prm)->tol;
    if (error < tol || rnorm < gl(S.prm)->upd_thresh) {
        if (c.b == 0 && threadIdx.x == 0) {
            st->error = error;
            st->iter = st->iter + 1;
            st->done = error < tol ? 1 : 5;
        }
        return;
    }
    const T beta = (T)(s[1] / st->rr_old);
    {
        if (SMALL && !first && i < end) {
#pragma unroll
            for (int u = 0; u < U; u++)
                if (i + u * kBlock < end) p[i + u * kBlock] = ri[u] / di[u] + beta * pi[u];
            i = end;
        }
        for (bool loaded = first; i + (U - 1) * kBlock < end; i += U * kBlock, loaded = false) {
            if (!loaded) {
#pragma unroll
                for (int u = 0; u < U; u++) {
                    const int k = i + u * kBlock;
                    pi[u] = ld<VNT>(p + k);
                    ri[u] = ld<VNT>(r + k);
                    di[u] = ld<VNT>(d + k);
                }
            }
#pragma unroll
            for (int u = 0; u < U; u++) p[i + u * kBlock] = ri[u] / di[u] + beta * pi[u];
        }
        if (tail) {  // the batch loop ran once: i is the thread's extra row
            p[i] = rt / dt + beta * pt;
            i = end;
        }
        for (; i < end; i += kBlock) p[i] = r[i] / d[i] + beta * p[i];
    }
    if (c.b == 0 && threadIdx.x == 0) {
        const int it = st->iter + 1;
        st->error = error;
        st->rr = s[1];
        st->iter = it;
        if (it >= st->max_iter) st->done = 2;
    }
}

template <typename T, bool VNT, bool SMALL>
__global__ __launch_bounds__(kBlock) void k_cg_direction(Grp g)
{
    cg_direction_body<T, VNT, SMALL>(g, (int)blockIdx.x);
}

// Reliable update: partials (r,r) and (r/d, r) of the replaced residual...
template <typename T>
__global__ __launch_bounds__(kBlock) void k_cg_rescan(Grp g)
{
    const Ctx c = ctx_of(g);
    const SysDev& S = *c.S;
    if (skip_action(S, 2)) return;
    const GVec<T> v = gvec(kv_of<T>(S));
    {
        const int b = c.b;
        double acc[2] = {0.0, 0.0};
        vec_chunk(S.n, c.G, b, [&](int i) {
            const T ri = v.r[i];
            const T zi = ri / v.d[i];
            acc[0] += (double)ri * (double)ri;
            acc[1] += (double)zi * (double)ri;
        });
        store_partials<2>(acc, gl(S.part) + 4 * (size_t)c.G, c.G, b);
    }
}

// ... then the direction step the update interrupted: beta = (z,r)/rr_old,
// p = z + beta p, and the recurrence resumes.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_cg_resume(Grp g)
{
    const Ctx c = ctx_of(g);
    const SysDev& S = *c.S;
    if (skip_action(S, 2)) return;
    double s[2];
    reduce_partials<2>(gl(S.part) + 4 * (size_t)c.G, c.G, s);
    OGSX_GL KrylovState* st = gl(S.st);
    const T beta = (T)(s[1] / st->rr_old);
    const GVec<T> v = gvec(kv_of<T>(S));
    vec_vblocks(c, S.n, [&](int i) { v.p[i] = v.r[i] / v.d[i] + beta * v.p[i]; });
    if (c.b == 0 && threadIdx.x == 0) {
        st->rr = s[1];
        st->done = 0;
    }
}

// ---- Right-preconditioned Jacobi BiCGStab (Linear_EQS::BiCGStab :1287)
// Partial layout: [0,2G) (r,r),(r0,r); [4G,5G) (r0,v); [8G,9G) (s,s);
// [12G,14G) (t,t),(t,s).

// s = A x; r0 = b - s; r = r0; v = p = 0; partials (r,r), (r0,r), (b,b)
template <int FMT, typename VT, typename T>
__global__ __launch_bounds__(kBlock) void k_bi_init(Grp g, int want)
{
    const Ctx c = ctx_of(g);
    const SysDev& S = *c.S;
    if (skip_action(S, want)) return;
    const GMat<VT> m = gmat(mv_of<VT>(S));
    const GVec<T> v = gvec(kv_of<T>(S));
    {
        const int b = c.b;
        double acc[3] = {0.0, 0.0, 0.0};
        spmv_rows<FMT, VT, T, T>(m, v.x, c.G, b, [&](int i, T ax) {
            const T ri = v.b[i] - ax;
            v.r0[i] = ri;
            v.r[i] = ri;
            v.v[i] = 0;
            v.p[i] = 0;
            acc[0] += (double)ri * (double)ri;
            acc[1] += (double)ri * (double)ri;
            acc[2] += (double)v.b[i] * (double)v.b[i];
        });
        store_partials<3>(acc, gl(S.part), c.G, b);
    }
}

__global__ void k_bi_init_final(Grp g, int want)
{
    const SysDev& S = *sys_of_block(g);
    if (skip_action(S, want)) return;
    double s[3];
    reduce_partials<3>(gl(S.part), S.G, s);
    if (threadIdx.x != 0) return;
    OGSX_GL KrylovState* st = gl(S.st);
    const KrylovParams P = *S.prm;  // one block per system: a generic load is fine
    const double bnorm_new = sqrt(s[2]);
    st->iter = 0;
    st->max_iter = P.max_iter;
    st->done = 0;
    st->pend = 0;
    st->rho0 = 1.0;
    st->alpha = 1.0;
    st->omega = 1.0;
    if (P.bnorm_prev > 0.0 && bnorm_new / P.bnorm_prev < P.tol) {
        st->error = bnorm_new / P.bnorm_prev;
        st->bnorm = bnorm_new;
        st->done = 4;
        return;
    }
    st->bnorm = bnorm_new;
    if (bnorm_new < DBL_MIN) {
        st->error = 0.0;
        st->done = 4;
        return;
    }
    st->error = sqrt(s[0]) / bnorm_new;
    if (st->error < P.tol) st->done = 1;
}

// B1: test of the previous iterate's r; rho1 = (r0,r); p update; p_h = p/d.
// A convergence on s in B4 (pend) is turned into done here, one kernel
// later, so that no block of B4 can see it before doing its own update.
template <typename T>
__device__ __forceinline__ void bi_dir_body(const Grp& g, int vb)
{
    const Ctx c = ctx_of(g, vb);
    const SysDev& S = *c.S;
    OGSX_GL KrylovState* st = gl(S.st);
    if (st->done) return;
    if (st->pend) {
        if (c.b == 0 && threadIdx.x == 0) st->done = 1;
        return;
    }
    // the partials' loads first, then the first batch of rows (p and v are
    // read at the first iteration too, unused), then the reduction and tests
    PartialLoads<2> pl;
    load_partials<2>(gl(S.part), c.G, pl);
    const int it = st->iter;
    const GVec<T> v = gvec(kv_of<T>(S));
    T rr[kVecU], pp[kVecU], vv[kVecU], dd[kVecU];
    double rho1 = 0.0;
    T beta = 0, omega = 0;
    const bool go = vec_chunk_batched_mid(
        S.n, c.G, c.b,
        [&](int k, int u) {
            rr[u] = v.r[k];
            pp[u] = v.p[k];
            vv[u] = v.v[k];
            dd[u] = v.d[k];
        },
        [&]() -> bool {
            double s[2];
            finish_partials<2>(gl(S.part), c.G, pl, s);
            const double tol = gl(S.prm)->tol;
            if (it > 0) {
                const double err = sqrt(s[0]) / st->bnorm;
                if (err < tol || fabs(st->omega) < DBL_MIN) {
                    if (c.b == 0 && threadIdx.x == 0) {
                        st->error = err;
                        st->done = (err < tol) ? 1 : 3;
                    }
                    return false;
                }
                if (it >= st->max_iter) {
                    if (c.b == 0 && threadIdx.x == 0) {
                        st->error = err;
                        st->done = 2;
                    }
                    return false;
                }
            }
            rho1 = s[1];
            if (fabs(rho1) < DBL_MIN) {
                if (c.b == 0 && threadIdx.x == 0) st->done = 6;
                return false;
            }
            if (it > 0) {
                beta = (T)((rho1 / st->rho0) * (st->alpha / st->omega));
                omega = (T)st->omega;
            }
            return true;
        },
        [&](int k, int u) {
            const T pi = it == 0 ? rr[u] : rr[u] + beta * (pp[u] - omega * vv[u]);
            v.p[k] = pi;
            v.ph[k] = pi / dd[u];
        });
    if (go && c.b == 0 && threadIdx.x == 0) st->rho1 = rho1;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_bi_dir(Grp g)
{
    bi_dir_body<T>(g, (int)blockIdx.x);
}

// B2: v = A p_h; partial (r0,v); counts the iteration
template <int FMT, typename VT, typename T>
__device__ __forceinline__ void bi_spmv_v_body(const Grp& g, int vb)
{
    const Ctx c = ctx_of(g, vb);
    const SysDev& S = *c.S;
    OGSX_GL KrylovState* st = gl(S.st);
    if (st->done || st->pend) return;
    const GMat<VT> m = gmat(mv_of<VT>(S));
    const GVec<T> v = gvec(kv_of<T>(S));
    {
        const int b = c.b;
        double acc[1] = {0.0};
        spmv_rows<FMT, VT, T, T>(m, v.ph, c.G, b, [&](int i, T ax) {
            v.v[i] = ax;
            acc[0] += (double)v.r0[i] * (double)ax;
        });
        store_partials<1>(acc, gl(S.part) + 4 * (size_t)c.G, c.G, b);
    }
    if (c.b == 0 && threadIdx.x == 0) st->iter = st->iter + 1;
}

template <int FMT, typename VT, typename T>
__global__ __launch_bounds__(kBlock) void k_bi_spmv_v(Grp g)
{
    bi_spmv_v_body<FMT, VT, T>(g, (int)blockIdx.x);
}

// B3: alpha = rho1/(r0,v); s = r - alpha v; s_h = s/d; partial (s,s)
template <typename T>
__device__ __forceinline__ void bi_s_body(const Grp& g, int vb)
{
    const Ctx c = ctx_of(g, vb);
    const SysDev& S = *c.S;
    OGSX_GL KrylovState* st = gl(S.st);
    if (st->done || st->pend) return;
    PartialLoads<1> pl;  // partials first, then the first batch of rows
    load_partials<1>(gl(S.part) + 4 * (size_t)c.G, c.G, pl);
    const double rho1 = st->rho1;
    double alpha_d = 0.0;
    T alpha = 0;
    const GVec<T> v = gvec(kv_of<T>(S));
    {
        const int b = c.b;
        double acc[1] = {0.0};
        T rr[kVecU], vv[kVecU], dd[kVecU];
        vec_chunk_batched_mid(
            S.n, c.G, b,
            [&](int k, int u) {
                rr[u] = v.r[k];
                vv[u] = v.v[k];
                dd[u] = v.d[k];
            },
            [&]() -> bool {
                double sv[1];
                finish_partials<1>(gl(S.part) + 4 * (size_t)c.G, c.G, pl, sv);
                alpha_d = rho1 / sv[0];
                alpha = (T)alpha_d;
                return true;
            },
            [&](int k, int u) {
                const T si = rr[u] - alpha * vv[u];
                v.s[k] = si;
                v.sh[k] = si / dd[u];
                acc[0] += (double)si * (double)si;
            });
        store_partials<1>(acc, gl(S.part) + 8 * (size_t)c.G, c.G, b);
    }
    if (c.b == 0 && threadIdx.x == 0) {
        st->alpha = alpha_d;
        st->rho0 = rho1;
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_bi_s(Grp g)
{
    bi_s_body<T>(g, (int)blockIdx.x);
}

// B4: test on s (then x += alpha p_h and stop: pend, done in B1); else
// t = A s_h; partials (t,t), (t,s)
template <int FMT, typename VT, typename T>
__device__ __forceinline__ void bi_spmv_t_body(const Grp& g, int vb)
{
    const Ctx c = ctx_of(g, vb);
    const SysDev& S = *c.S;
    OGSX_GL KrylovState* st = gl(S.st);
    if (st->done || st->pend) return;
    double ss[1];
    reduce_partials<1>(gl(S.part) + 8 * (size_t)c.G, c.G, ss);
    const double err = sqrt(ss[0]) / st->bnorm;
    const GVec<T> v = gvec(kv_of<T>(S));
    if (err < gl(S.prm)->tol) {
        const T alpha = (T)st->alpha;
        vec_vblocks(c, S.n, [&](int i) { v.x[i] += alpha * v.ph[i]; });
        if (c.b == 0 && threadIdx.x == 0) {
            st->error = err;
            st->pend = 1;
        }
        return;
    }
    const GMat<VT> m = gmat(mv_of<VT>(S));
    {
        const int b = c.b;
        double acc[2] = {0.0, 0.0};
        spmv_rows<FMT, VT, T, T>(m, v.sh, c.G, b, [&](int i, T ax) {
            v.t[i] = ax;
            acc[0] += (double)ax * (double)ax;
            acc[1] += (double)ax * (double)v.s[i];
        });
        store_partials<2>(acc, gl(S.part) + 12 * (size_t)c.G, c.G, b);
    }
}

template <int FMT, typename VT, typename T>
__global__ __launch_bounds__(kBlock) void k_bi_spmv_t(Grp g)
{
    bi_spmv_t_body<FMT, VT, T>(g, (int)blockIdx.x);
}

// B5: omega = (t,s)/(t,t); x += alpha p_h + omega s_h; r = s - omega t;
// partials (r,r), (r0,r)
template <typename T>
__device__ __forceinline__ void bi_update_body(const Grp& g, int vb)
{
    const Ctx c = ctx_of(g, vb);
    const SysDev& S = *c.S;
    OGSX_GL KrylovState* st = gl(S.st);
    if (st->done || st->pend) return;
    PartialLoads<2> pl;  // partials first, then the first batch of rows
    load_partials<2>(gl(S.part) + 12 * (size_t)c.G, c.G, pl);
    double omega_d = 0.0;
    T omega = 0;
    const T alpha = (T)st->alpha;
    const GVec<T> v = gvec(kv_of<T>(S));
    {
        const int b = c.b;
        double acc[2] = {0.0, 0.0};
        T xx[kVecU], ph[kVecU], sh[kVecU], ss[kVecU], tt[kVecU], r0[kVecU];
        vec_chunk_batched_mid(
            S.n, c.G, b,
            [&](int k, int u) {
                xx[u] = v.x[k];
                ph[u] = v.ph[k];
                sh[u] = v.sh[k];
                ss[u] = v.s[k];
                tt[u] = v.t[k];
                r0[u] = v.r0[k];
            },
            [&]() -> bool {
                double sv[2];
                finish_partials<2>(gl(S.part) + 12 * (size_t)c.G, c.G, pl, sv);
                omega_d = (sv[0] > DBL_MIN) ? sv[1] / sv[0] : 1.0;
                omega = (T)omega_d;
                return true;
            },
            [&](int k, int u) {
                v.x[k] = xx[u] + (alpha * ph[u] + omega * sh[u]);
                const T ri = ss[u] - omega * tt[u];
                v.r[k] = ri;
                acc[0] += (double)ri * (double)ri;
                acc[1] += (double)r0[u] * (double)ri;
            });
        store_partials<2>(acc, gl(S.part), c.G, b);
    }
    if (c.b == 0 && threadIdx.x == 0) st->omega = omega_d;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_bi_update(Grp g)
{
    bi_update_body<T>(g, (int)blockIdx.x);
}

// ---- Persistent launches: a batch of Krylov iterations in one launch
//
// For systems too small to fill the GPU the cost of an iteration is the
// dispatch of its dependent launches, not their work.  A persistent launch
// keeps a grid of co-resident blocks (cooperative launch) that works
// through the same virtual blocks -- the same per-system chunks, partial-sum
// layout and G -- that separate launches would use, so every result is
// bit-identical, with a grid barrier between two phases instead of a kernel
// boundary.  Measured on MI355X (200 x 200 quads, FP64 CG): 57 us per
// iteration against 13 us with three separate launches; the barrier's
// agent-scope release/acquire across the 8 XCDs costs 12-15 us, a dependent
// kernel dispatch about 4 us, so the path is off unless asked for.

// Barrier state of one persistent launch (device memory, zeroed before it).
struct GridSync {
    unsigned count;  // blocks arrived at the current barrier
    unsigned gen;    // barriers completed
    int aborted;     // some block gave up waiting (the host asked it to)
    int pad;
};

// How often a waiting block looks at the host's abort flag (polls).
constexpr unsigned kSyncAbortPolls = 4096;

// Waits until `target` blocks arrived (target = gridDim.x).  The wait is
// bounded by the host: every kSyncAbortPolls polls a waiting block reads
// the host's abort flag and, when it is set, leaves (returns false); a
// launch whose blocks leave ends normally and the host reports the
// failure.  Agent-scope release/acquire fences make the phase's writes of
// every block (on every XCD) visible to the next phase.
__device__ bool grid_sync(GridSync* gs, const volatile int* host_abort, unsigned target)
{
    __shared__ int ok;
    __syncthreads();
    if (threadIdx.x == 0) {
        ok = 1;
        const unsigned g0 = __hip_atomic_load(&gs->gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __threadfence();  // the block's writes (all its waves passed the barrier above)
        const unsigned arrived =
            __hip_atomic_fetch_add(&gs->count, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (arrived + 1 == target) {
            __hip_atomic_store(&gs->count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_fetch_add(&gs->gen, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            unsigned polls = 0;
            while (__hip_atomic_load(&gs->gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == g0) {
                __builtin_amdgcn_s_sleep(2);
                if (++polls % kSyncAbortPolls == 0 &&
                    (*host_abort ||
                     __hip_atomic_load(&gs->aborted, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
                    __hip_atomic_store(&gs->aborted, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    ok = 0;
                    break;
                }
            }
        }
        __threadfence();
    }
    __syncthreads();
    return ok;
}

// Phases of the recurrences a persistent launch runs.
constexpr int kPersistCg = 0;    // K1 SpMV, K2 update, K3 direction
constexpr int kPersistCgSr = 1;  // single-reduction CG: K1 SpMV, K2 update
constexpr int kPersistBi = 2;    // BiCGStab: B1 .. B5

// Recurrence state `done` of a single-system launch, read after a barrier.
__device__ __forceinline__ int stopped(const Grp& g)
{
    return __hip_atomic_load(&g.sys->st->done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// `iters` iterations of the recurrence over the virtual blocks of gm (SpMV
// phases, nbm blocks) and gv (vector phases, nbv blocks).  A single system
// (gm.blk2sys == nullptr) leaves the loop once its recurrence stopped.  The
// flag is read right after the barrier that follows the only phase writing
// it (CG: K3; single-reduction CG: K2; BiCGStab: B1), so no block can be
// writing it while another reads it: every block leaves at the same point.
template <int FMT, typename VT, typename T>
__device__ __forceinline__ void cg1_spmv_body(const Grp& g, int resume, int vb);
template <typename T>
__device__ __forceinline__ void cg1_update_body(const Grp& g, int vb);

template <int FMT, typename VT, typename T, int KM>
__global__ __launch_bounds__(kBlock) void k_krylov_persistent(Grp gm, int nbm, Grp gv, int nbv,
                                                              int iters, GridSync* gs,
                                                              const int* host_abort)
{
    const unsigned all = gridDim.x;
    const volatile int* ha = host_abort;
    const bool single = gm.blk2sys == nullptr;
#define OGSX_PHASE(NB, CALL)                                                  \
    for (int vb = (int)blockIdx.x; vb < (NB); vb += (int)gridDim.x) {        \
        CALL;                                                                 \
        __syncthreads();                                                      \
    }                                                                         \
    if (!grid_sync(gs, ha, all)) return;
    for (int it = 0; it < iters; it++) {
        if constexpr (KM == kPersistCg) {
            OGSX_PHASE(nbm, (cg_spmv_body<FMT, VT, T>(gm, vb)))
            OGSX_PHASE(nbv, (cg_update_body<T, false>(gv, vb)))
            OGSX_PHASE(nbv, (cg_direction_body<T, false>(gv, vb)))
            if (single && stopped(gm)) break;
        } else if constexpr (KM == kPersistCgSr) {
            OGSX_PHASE(nbm, (cg1_spmv_body<FMT, VT, T>(gm, 0, vb)))
            OGSX_PHASE(nbv, (cg1_update_body<T>(gv, vb)))
            if (single && stopped(gm)) break;
        } else {
            OGSX_PHASE(nbv, (bi_dir_body<T>(gv, vb)))
            if (single && stopped(gm)) break;
            OGSX_PHASE(nbm, (bi_spmv_v_body<FMT, VT, T>(gm, vb)))
            OGSX_PHASE(nbv, (bi_s_body<T>(gv, vb)))
            OGSX_PHASE(nbm, (bi_spmv_t_body<FMT, VT, T>(gm, vb)))
            OGSX_PHASE(nbv, (bi_update_body<T>(gv, vb)))
        }
    }
#undef OGSX_PHASE
}

// Diagnostic of the barrier's give-up path: every block waits for one more
// block than the grid holds, a barrier that cannot complete; with the host's
// abort flag already set, each block leaves after kSyncAbortPolls polls and
// the launch ends normally.
__global__ void k_sync_giveup_probe(GridSync* gs, const int* host_abort)
{
    const volatile int* ha = host_abort;
    (void)grid_sync(gs, ha, gridDim.x + 1);
}

// ---- FP32-inside-FP64 iterative refinement helpers

// r64 = b - A x in FP64; partials (b,b), (r,r) and (S r, S r) -- the norm
// of the residual the scaled FP32 recurrence starts from (= (r,r) unscaled)
template <int FMT>
__global__ __launch_bounds__(kBlock) void k_ir_residual(Grp g)
{
    const Ctx c = ctx_of(g);
    const SysDev& S = *c.S;
    const GMat<double> m = gmat(S.m64);
    const GVec<double> v = gvec(S.v64);
    OGSX_GL double* r = gl(S.r64);
    const bool scaled = gl(S.prm)->ir_scaled;
    const OGSX_GL float* sc = gl(S.v32.d);
    {
        const int b = c.b;
        double acc[3] = {0.0, 0.0, 0.0};
        spmv_rows<FMT, double, double, double>(m, v.x, c.G, b, [&](int i, double ax) {
            const double ri = v.b[i] - ax;
            r[i] = ri;
            acc[0] += v.b[i] * v.b[i];
            acc[1] += ri * ri;
            const double si = scaled ? (double)sc[i] * ri : ri;
            acc[2] += si * si;
        });
        store_partials<3>(acc, gl(S.part), c.G, b);
    }
}

// x64 += unscale * x32: folds the FP32 correction of the recurrence that
// just stopped into the FP64 iterate.
__global__ __launch_bounds__(kBlock) void k_ir_fold(Grp g)
{
    const Ctx c = ctx_of(g);
    const SysDev& S = *c.S;
    if (!gl(S.irs)->started) return;
    const double u = gl(S.prm)->unscale;
    OGSX_GL double* x = gl(S.v64.x);
    const OGSX_GL float* x32 = gl(S.v32.x);
    if (gl(S.prm)->ir_scaled) {
        // the scaled recurrence solved (S A S) y = S r: the correction is S y
        const OGSX_GL float* sc = gl(S.v32.d);
        vec_vblocks(c, S.n, [&](int i) { x[i] += u * ((double)sc[i] * (double)x32[i]); });
    } else {
        vec_vblocks(c, S.n, [&](int i) { x[i] += u * (double)x32[i]; });
    }
}

// One block per listed system: after k_ir_residual, the refinement step.
// First call of a solve: ||b||, CheckNormRHS (equation_class.cpp:1104).
// Then the FP64 stopping test, the iteration / refinement-step limits, and
// the choice of the next FP32 recurrence: a new one on r / ||r|| (restarted
// refinement; reliable updates' first) or the resumption of the running one
// from the true residual (reliable updates, Clark et al., Comput. Phys.
// Commun. 181 (2010) 1517).
__global__ void k_ir_decide(Grp g)
{
    const SysDev& S = *sys_of_block(g);
    double s[3];
    reduce_partials<3>(gl(S.part), S.G, s);
    if (threadIdx.x != 0) return;
    OGSX_GL IRState& R = *gl(S.irs);
    OGSX_GL KrylovParams& P = *gl(S.prm);
    OGSX_GL KrylovState& st = *gl(S.st);
    R.action = 0;
    // a finished solve leaves its recurrence stopped, so batched launches skip it
    auto finish = [&](int status) {
        R.status = status;
        if (st.done == 0) st.done = 4;
    };
    const double rn = sqrt(s[1]);
    const double rs = sqrt(s[2]);  // ||S r||, = rn without scaling
    int inner_status = 0, inner_iters = 0;
    if (!R.started) {
        R.bnorm = sqrt(s[0]);
        R.total = 0;
        R.updates = 0;
        if (P.ir_bnorm_prev > 0.0 && R.bnorm / P.ir_bnorm_prev < P.ir_tol) {
            R.error = R.bnorm / P.ir_bnorm_prev;
            finish(2);
            return;
        }
        if (R.bnorm < DBL_MIN) {
            R.error = 0.0;
            finish(2);
            return;
        }
    } else {
        // st.done == 0 with pend set: BiCGStab converged on s; the
        // single-reduction CG counts its last update in sr_it
        const int its = max(st.iter, st.sr_it);
        inner_iters = its;
        inner_status = st.done == 2 ? 2 : (st.done == 3 || st.done == 6) ? 3 : 1;
        R.total = P.ir_reliable ? its : R.total + max(its, 1);
        R.updates++;
    }
    R.error = rn / R.bnorm;
    if (R.error < P.ir_tol) {
        finish(1);
        return;
    }
    if (R.total >= P.ir_max_iter || (!P.ir_reliable && R.updates >= P.ir_max_outer) ||
        inner_status == 2) {
        finish(0);
        return;
    }
    if (inner_status == 3 && inner_iters == 0) {
        finish(3);
        return;
    }
    if (P.ir_reliable && R.started) {
        // the recurrence resumes from r32 = scale S r64 (||r32|| = rs scale);
        // its stopping target is the FP64 one, mapped by ||S r|| / ||r||
        R.maxr = rs * P.scale;
        P.upd_thresh = P.ir_delta * R.maxr;
        P.tol = P.ir_tol * R.bnorm * P.scale * (rs / rn);
        R.action = 2;
    } else {
        P.scale = 1.0 / rs;
        P.unscale = rs;
        P.bnorm_prev = 0.0;
        if (P.ir_reliable) {
            P.tol = P.ir_tol * R.bnorm / rn;  // the FP64 target in units of ||r32|| = 1
            R.maxr = 1.0;
            P.upd_thresh = P.ir_delta;
        } else {
            P.tol = P.ir_delta;
            P.upd_thresh = 0.0;
        }
        P.max_iter = max(1, P.ir_max_iter - R.total);
        R.action = 1;
    }
    R.started = 1;
}

// The decided recurrence's start vectors: new recurrence b32 = scale r64,
// resumption r32 = scale r64; x32 = 0 either way.
__global__ __launch_bounds__(kBlock) void k_ir_prepare(Grp g)
{
    const Ctx c = ctx_of(g);
    const SysDev& S = *c.S;
    const int act = gl(S.irs)->action;
    if (act == 0) return;
    const double scale = gl(S.prm)->scale;
    OGSX_GL float* dst = gl(act == 1 ? S.v32.b : S.v32.r);
    OGSX_GL float* x32 = gl(S.v32.x);
    const OGSX_GL double* r64 = gl(S.r64);
    if (gl(S.prm)->ir_scaled) {
        const OGSX_GL float* sc = gl(S.v32.d);
        vec_vblocks(c, S.n, [&](int i) {
            dst[i] = (float)(r64[i] * scale * (double)sc[i]);
            x32[i] = 0.0f;
        });
    } else {
        vec_vblocks(c, S.n, [&](int i) {
            dst[i] = (float)(r64[i] * scale);
            x32[i] = 0.0f;
        });
    }
}

// ---- FP32 inner CG with one reduction point per iteration
// (OGSX_IR_RELIABLE_SR): the Chronopoulos-Gear recurrence (J. Comput. Appl.
// Math. 25 (1989) 153-168) on the symmetrically Jacobi-scaled FP32 matrix
// S A S, S = diag(1/sqrt|a_ii|) -- in exact arithmetic Jacobi-PCG, with no
// preconditioner left in the loop.  Per iteration two kernels:
//   K1  w = (SAS) r; partials (r,w), (r,r)
//   K2  gamma = (r,r), delta = (r,w); stopping test; beta = gamma/gamma_old,
//       alpha = gamma/(delta - beta gamma/alpha_old) (beta = 0, alpha =
//       gamma/delta at the first iteration); p = r + beta p, s = w + beta s,
//       x += alpha p, r -= alpha s
// w lives in the z vector.  Only the FP32 recurrence inside FP64 iterative
// refinement uses it; the FP64 solver keeps the reference's recurrence.

// S = diag(1/sqrt|a_ii|) into the FP32 diagonal vector (1 where |a_ii| <
// DBL_MIN, and everywhere without a preconditioner).
__global__ __launch_bounds__(kBlock) void k_diag_scale(Grp g)
{
    const Ctx c = ctx_of(g);
    const SysDev& S = *c.S;
    OGSX_GL float* d = gl(S.v32.d);
    const OGSX_GL int* dp = gl(S.diagpos);
    const OGSX_GL double* A = gl(S.A);
    const int jac = gl(S.prm)->jacobi;
    vec_vblocks(c, S.n, [&](int i) {
        double v = (jac && dp[i] >= 0) ? fabs(A[dp[i]]) : 1.0;
        if (v < DBL_MIN) v = 1.0;
        d[i] = (float)(1.0 / sqrt(v));
    });
}

// A32 = S A S (CRS order), from the FP64 values and the FP32 scale vector.
__global__ __launch_bounds__(kBlock) void k_scale_a32(Grp g)
{
    const Ctx c = ctx_of(g);
    const SysDev& S = *c.S;
    const GMat<double> m = gmat(S.m64);
    OGSX_GL float* A32 = gl(S.A32);
    const OGSX_GL float* sc = gl(S.v32.d);
    vec_vblocks(c, S.n, [&](int i) {
        const double si = (double)sc[i];
        const int k1 = m.ptr[i + 1];
        for (int k = m.ptr[i]; k < k1; k++)
            A32[k] = (float)(m.val[k] * si * (double)sc[m.col[k]]);
    });
}

// FP32 SELL-64 / slice-triple values of S A S straight from the FP64 CRS
// values (tri = 1: the triple layout): one wave per 64-row slice, so the
// stored entries are written and their source positions read coalesced,
// and lane l's row is 64 s + l.
__global__ __launch_bounds__(kBlock) void k_scale_sell32(Grp g, int tri)
{
    const Ctx c = ctx_of(g);
    const SysDev& S = *c.S;
    const GMat<double> m = gmat(S.m64);
    const GMat<float> m32 = gmat(S.m32);
    OGSX_GL float* sv = gl(S.sval32);
    const OGSX_GL int* src = gl(tri ? S.tri_src : S.sell_src);
    const OGSX_GL float* sc = gl(S.v32.d);
    int sb, se;
    chunk_of(m32.n_slices, c.G, c.b, sb, se);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int sl = sb + wave; sl < se; sl += kBlock / 64) {
        int p0, cnt;
        if (tri) {
            p0 = m32.tri_ptr[sl];
            cnt = 3 * m32.tri_cnt[m32.tri_id[sl]];
        } else {
            p0 = m32.sell_ptr[sl];
            cnt = m32.sell_w[sl];
        }
        const int row = sl * kSellC + lane;
        const double si = row < m.n ? (double)sc[row] : 0.0;
        for (int e = 0; e < cnt; e++) {
            const int p = p0 + e * kSellC + lane;
            const int k = src[p];
            sv[p] = k >= 0 ? (float)(m.val[k] * si * (double)sc[m.col[k]]) : 0.0f;
        }
    }
}

// New recurrence: r = b - (SAS) x; partials (b,b), (r,r), (r,r) for
// k_cg_init_final.
template <int FMT, typename VT, typename T>
__global__ __launch_bounds__(kBlock) void k_cg1_init(Grp g, int want)
{
    const Ctx c = ctx_of(g);
    const SysDev& S = *c.S;
    if (skip_action(S, want)) return;
    const GMat<VT> m = gmat(mv_of<VT>(S));
    const GVec<T> v = gvec(kv_of<T>(S));
    double acc[3] = {0.0, 0.0, 0.0};
    spmv_rows<FMT, VT, T, T>(m, v.x, c.G, c.b, [&](int i, T ax) {
        const T ri = v.b[i] - ax;
        v.r[i] = ri;
        acc[0] += (double)v.b[i] * (double)v.b[i];
        acc[1] += (double)ri * (double)ri;
        acc[2] += (double)ri * (double)ri;
    });
    store_partials<3>(acc, gl(S.part), c.G, c.b);
}

// K1: w = (SAS) r; partials (r,w), (r,r).  Block 0 takes over the scalars
// the last update left (sr_*).  resume = 1: the first SpMV after a reliable
// update, run for the systems whose refinement step chose to resume.
template <int FMT, typename VT, typename T>
__device__ __forceinline__ void cg1_spmv_body(const Grp& g, int resume, int vb)
{
    const Ctx c = ctx_of(g, vb);
    const SysDev& S = *c.S;
    OGSX_GL KrylovState* st = gl(S.st);
    if (resume ? skip_action(S, 2) : st->done != 0) return;
    const GMat<VT> m = gmat(mv_of<VT>(S));
    const GVec<T> v = gvec(kv_of<T>(S));
    double acc[2] = {0.0, 0.0};
    spmv_rows<FMT, VT, T, T>(m, v.r, c.G, c.b, [&](int i, T ax) {
        v.z[i] = ax;
        const T ri = v.r[i];
        acc[0] += (double)ri * (double)ax;
        acc[1] += (double)ri * (double)ri;
    });
    store_partials<2>(acc, gl(S.part), c.G, c.b);
    if (c.b == 0 && threadIdx.x == 0) {
        st->rr_old = st->sr_gam;
        st->alpha = st->sr_alp;
        st->iter = st->sr_it;
    }
}

template <int FMT, typename VT, typename T>
__global__ __launch_bounds__(kBlock) void k_cg1_spmv(Grp g, int resume)
{
    cg1_spmv_body<FMT, VT, T>(g, resume, (int)blockIdx.x);
}

// K2: stopping test on the current r, then the update.  With reliable
// updates (upd_thresh > 0) the recurrence stops with done = 5 once ||r|| <
// upd_thresh, before updating, leaving p and s for the resumption.
template <typename T>
__device__ __forceinline__ void cg1_update_body(const Grp& g, int vb)
{
    const Ctx c = ctx_of(g, vb);
    const SysDev& S = *c.S;
    OGSX_GL KrylovState* st = gl(S.st);
    if (st->done) return;
    const GVec<T> v = gvec(kv_of<T>(S));
    OGSX_GL T* __restrict__ x = v.x;
    OGSX_GL T* __restrict__ r = v.r;
    OGSX_GL T* __restrict__ p = v.p;
    OGSX_GL T* __restrict__ sv = v.s;
    const OGSX_GL T* __restrict__ w = v.z;
    int beg, end;
    chunk_of(S.n, c.G, c.b, beg, end);
    constexpr int U = 4;
    int i = beg + threadIdx.x;
    // the first batch of the stream does not depend on the scalars: its
    // loads are issued before the partials are reduced
    T xi[U], pi[U], ri[U], si[U], wi[U];
    const bool first = i + (U - 1) * kBlock < end;
    if (first) {
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int k = i + u * kBlock;
            xi[u] = x[k];
            pi[u] = p[k];
            ri[u] = r[k];
            si[u] = sv[k];
            wi[u] = w[k];
        }
    }
    double red[2];
    reduce_partials<2>(gl(S.part), c.G, red);
    const double delta = red[0], gamma = red[1];
    const double rnorm = sqrt(gamma);
    const double error = rnorm / st->bnorm;
    const double tol = gl(S.prm)->tol;
    if (error < tol || rnorm < gl(S.prm)->upd_thresh) {
        if (c.b == 0 && threadIdx.x == 0) {
            st->error = error;
            st->done = error < tol ? 1 : 5;
        }
        return;
    }
    // the limit is tested here, where every block reads the same count, so
    // either all blocks update or none does
    const int it = st->iter;
    if (it >= st->max_iter) {
        if (c.b == 0 && threadIdx.x == 0) {
            st->error = error;
            st->done = 2;
        }
        return;
    }
    const bool start = it == 0;
    double alpha_d, beta_d = 0.0;
    if (start) {
        alpha_d = gamma / delta;
    } else {
        beta_d = gamma / st->rr_old;
        alpha_d = gamma / (delta - beta_d * gamma / st->alpha);
    }
    const T alpha = (T)alpha_d, beta = (T)beta_d;
    for (bool loaded = first; i + (U - 1) * kBlock < end; i += U * kBlock, loaded = false) {
        if (!loaded) {
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int k = i + u * kBlock;
                xi[u] = x[k];
                pi[u] = p[k];
                ri[u] = r[k];
                si[u] = sv[k];
                wi[u] = w[k];
            }
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int k = i + u * kBlock;
            const T pn = start ? ri[u] : ri[u] + beta * pi[u];
            const T sn = start ? wi[u] : wi[u] + beta * si[u];
            p[k] = pn;
            sv[k] = sn;
            x[k] = xi[u] + alpha * pn;
            r[k] = ri[u] - alpha * sn;
        }
    }
    for (; i < end; i += kBlock) {
        const T pn = start ? r[i] : r[i] + beta * p[i];
        const T sn = start ? w[i] : w[i] + beta * sv[i];
        p[i] = pn;
        sv[i] = sn;
        x[i] += alpha * pn;
        r[i] -= alpha * sn;
    }
    if (c.b == 0 && threadIdx.x == 0) {
        st->sr_gam = gamma;
        st->sr_alp = alpha_d;
        st->sr_it = it + 1;
        st->error = error;
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_cg1_update(Grp g)
{
    cg1_update_body<T>(g, (int)blockIdx.x);
}

// Resumption after a reliable update (r replaced by the scaled true
// residual, x = 0, p and s kept), after k_cg1_spmv(resume): the direction
// step with beta = (r,r)/gamma_old, then the partial (p, s) -- the
// recurrence's (p, A p) identity assumes the residuals it produced itself,
// so the first step after a replacement takes alpha from the explicit dot.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_cg1_resume(Grp g)
{
    const Ctx c = ctx_of(g);
    const SysDev& S = *c.S;
    if (skip_action(S, 2)) return;
    OGSX_GL KrylovState* st = gl(S.st);
    double red[2];
    reduce_partials<2>(gl(S.part), c.G, red);
    const double gamma = red[1];
    const T beta = (T)(gamma / st->rr_old);
    const GVec<T> v = gvec(kv_of<T>(S));
    double acc[1] = {0.0};
    vec_chunk(S.n, c.G, c.b, [&](int i) {
        const T pn = v.r[i] + beta * v.p[i];
        const T sn = v.z[i] + beta * v.s[i];
        v.p[i] = pn;
        v.s[i] = sn;
        acc[0] += (double)pn * (double)sn;
    });
    store_partials<1>(acc, gl(S.part) + 4 * (size_t)c.G, c.G, c.b);
    if (c.b == 0 && threadIdx.x == 0) st->rr = gamma;
}

// ... and its update: alpha = gamma/(p,s), x += alpha p, r -= alpha s; the
// recurrence runs again.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_cg1_resume_step(Grp g)
{
    const Ctx c = ctx_of(g);
    const SysDev& S = *c.S;
    if (skip_action(S, 2)) return;
    OGSX_GL KrylovState* st = gl(S.st);
    double red[1];
    reduce_partials<1>(gl(S.part) + 4 * (size_t)c.G, c.G, red);
    const double gamma = st->rr;
    const double alpha_d = gamma / red[0];
    const T alpha = (T)alpha_d;
    const GVec<T> v = gvec(kv_of<T>(S));
    vec_vblocks(c, S.n, [&](int i) {
        v.x[i] += alpha * v.p[i];
        v.r[i] -= alpha * v.s[i];
    });
    if (c.b == 0 && threadIdx.x == 0) {
        const int it = st->iter + 1;
        st->sr_gam = gamma;
        st->sr_alp = alpha_d;
        st->sr_it = it;
        st->done = it >= st->max_iter ? 2 : 0;
    }
}

// Element -> CRS slot map, built on the device from the uploaded pattern:
// slot[(e npe + i) npe + j] = position of column conn[e][j] in row conn[e][i]
// (lower_bound in the sorted row; the pattern holds every pair of nodes that
// share an element, so the column is found).
__global__ __launch_bounds__(kBlock) void k_slot_map(int64_t ne_npe, int npe, const int* conn,
                                                     const int* ptr, const int* col, int* slot)
{
    for (int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x; t < ne_npe;
         t += (int64_t)gridDim.x * kBlock) {
        const int64_t e = t / npe;
        const int r = conn[t];
        const int lo0 = ptr[r], hi0 = ptr[r + 1];
        for (int j = 0; j < npe; j++) {
            const int c = conn[e * npe + j];
            int lo = lo0, hi = hi0;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (col[mid] < c)
                    lo = mid + 1;
                else
                    hi = mid;
            }
            slot[t * npe + j] = min(lo, hi0 - 1);
        }
    }
}

// Sets the recurrence state of one system (tuning and timing runs of the
// solve's SpMV kernel need it "running"; afterwards it goes back to stopped).
__global__ void k_set_done(KrylovState* st, int done)
{
    if (threadIdx.x == 0) gl(st)->done = done;
}

// ------------------------------------------------------------------------
// small utility kernels
// ------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kBlock) void k_fill(int64_t n, T* p, T v)
{
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kBlock)
        p[i] = v;
}

// Plain streaming copy (bandwidth calibration): 16-byte loads and stores.
__global__ __launch_bounds__(kBlock) void k_stream_copy(int64_t n4, const double2* __restrict__ a,
                                                        double2* __restrict__ b)
{
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n4;
         i += (int64_t)gridDim.x * kBlock)
        b[i] = a[i];
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_copy(int64_t n, const T* a, T* b)
{
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kBlock)
        b[i] = a[i];
}

// Point source terms: one thread per distinct node, values in list order.
__device__ __forceinline__ void add_rhs_node(int k, const int* node, const int* vptr,
                                             const double* val, double* b)
{
    double acc = b[node[k]];
    for (int j = vptr[k]; j < vptr[k + 1]; j++) acc += val[j];
    b[node[k]] = acc;
}

__global__ __launch_bounds__(kBlock) void k_add_rhs(int nu, const int* node, const int* vptr,
                                                    const double* val, double* b)
{
    const int k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= nu) return;
    add_rhs_node(k, node, vptr, val, b);
}

// Dirichlet flags: flag[i] = 1 and g[i] for every listed node
__device__ __forceinline__ void bc_mark_one(int k, const int* idx, const double* val, int8_t* flag,
                                            double* g, int8_t on)
{
    flag[idx[k]] = on;
    if (on) g[idx[k]] = val[k];
}

__global__ __launch_bounds__(kBlock) void k_bc_mark(int m, const int* idx, const double* val,
                                                    int8_t* flag, double* g, int8_t on)
{
    const int k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= m) return;
    bc_mark_one(k, idx, val, flag, g, on);
}

// Row-only elimination (CSparseMatrix::Diagonize CRS branch,
// matrix_class.cpp:1663-1690,1772): one thread per Dirichlet node.
__device__ __forceinline__ void bc_row_one(int k, const int* idx, const double* val, const int* ptr,
                                           const int* col, const int* diagpos, double* A, double* b)
{
    const int id = idx[k];
    const int dp = diagpos[id];
    const double vdiag = dp >= 0 ? A[dp] : 0.0;
    for (int j = ptr[id]; j < ptr[id + 1]; j++)
        if (col[j] != id) A[j] = 0.0;
    b[id] = vdiag * val[k];
}

__global__ __launch_bounds__(kBlock) void k_bc_row(int m, const int* idx, const double* val,
                                                   const int* ptr, const int* col,
                                                   const int* diagpos, double* A, double* b)
{
    const int k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= m) return;
    bc_row_one(k, idx, val, ptr, col, diagpos, A, b);
}

// Symmetric elimination: Dirichlet rows as above; every other row lifts
// b_i -= a_ik g_k and zeroes a_ik for Dirichlet columns k (column order).
// 32 lanes per row: the row's columns are read in one coalesced pass, and
// the lifts are applied one Dirichlet column at a time in column order
// (ballot over the lanes that hold one, then each lane's product in turn).
__device__ __forceinline__ void bc_sym_rows(int n, const int* rows, const int* ptr, const int* col,
                                            const int* diagpos, const int8_t* flag, const double* g,
                                            double* A, double* b, int first, int step)
{
    const int lane = threadIdx.x & 31;
    const int half = (threadIdx.x & 63) >> 5;
    for (int t = first; t < n; t += step) {
        const int i = rows ? rows[t] : t;
        const int p0 = ptr[i], p1 = ptr[i + 1];
        if (flag[i]) {
            const int dp = diagpos[i];
            const double vdiag = dp >= 0 ? A[dp] : 0.0;  // the diagonal is not zeroed
            for (int j = p0 + lane; j < p1; j += 32)
                if (col[j] != i) A[j] = 0.0;
            if (lane == 0) b[i] = vdiag * g[i];
        } else {
            double bi = b[i];
            bool touched = false;
            for (int base = p0; base < p1; base += 32) {
                const int j = base + lane;
                const int c = j < p1 ? col[j] : 0;
                const bool f = j < p1 && flag[c];
                const double av = f ? A[j] : 0.0;
                const double gv = f ? g[c] : 0.0;
                unsigned m = (unsigned)(__ballot(f) >> (32 * half));
                touched |= m != 0;
                while (m) {
                    const int l = __ffs(m) - 1;
                    m &= m - 1;
                    const double al = __shfl(av, l, 32);
                    const double gl_ = __shfl(gv, l, 32);
                    bi -= al * gl_;
                }
                if (f) A[j] = 0.0;
            }
            if (lane == 0 && touched) b[i] = bi;
        }
    }
}

// Symmetric elimination over the m rows of `rows` (rows the elimination can
// change; the others are left as they are).
__global__ __launch_bounds__(kBlock) void k_bc_sym(int m, const int* rows, const int* ptr,
                                                   const int* col, const int* diagpos,
                                                   const int8_t* flag, const double* g, double* A,
                                                   double* b)
{
    bc_sym_rows(m, rows, ptr, col, diagpos, flag, g, A, b,
                blockIdx.x * (kBlock / 32) + (int)(threadIdx.x >> 5), gridDim.x * (kBlock / 32));
}

// ------------------------------------------------------------------------
// Batched time-step preparation over the members of a group: one launch per
// phase (u_old <- x and A = b = 0; element pass; row pass; sources;
// Dirichlet) over every listed member, each member's blocks doing exactly
// what the per-system kernels do for it, so its A and b are bit-identical.
// ------------------------------------------------------------------------
struct StepArgs {
    AsmArgs a;
    const int32_t* inc_ptr;  // two-pass incidence and row positions
    const uint8_t* loc;
    const int32_t* ptr;      // CRS
    const int32_t* col;
    const int32_t* diagpos;
    const int32_t* st_node;  // stored source list (grouped by node)
    const int32_t* st_vptr;
    const double* st_val;
    const int32_t* bc_idx;   // stored Dirichlet list
    const double* bc_val;
    const int32_t* bc_rows;  // rows the symmetric elimination can change
    int8_t* bcflag;
    double* bcg;
    double* uold;
    const double* x;
    int n, nnz, ne, st_nu, bc_m, bc_rows_m, shift;
};

// Launch geometry of one phase: block b works for member blk2mem[b] as its
// block b - blk0[m] of blk0[m + 1] - blk0[m].
struct StepMap {
    const StepArgs* tab;
    const int* blk2mem;
    const int* blk0;
};

struct StepCtx {
    const StepArgs* S;
    int lb, nb;
};

__device__ __forceinline__ StepCtx step_ctx(const StepMap& m)
{
    const int mi = m.blk2mem[blockIdx.x];
    const int b0 = m.blk0[mi];
    return StepCtx{m.tab + mi, (int)blockIdx.x - b0, m.blk0[mi + 1] - b0};
}

constexpr int kStepFill = kBlock * 4;  // entries per block of the fill phase

// u_old <- x (with shift), A = 0, b = 0.
__global__ __launch_bounds__(kBlock) void k_step_init(StepMap m)
{
    const StepCtx c = step_ctx(m);
    const StepArgs& S = *c.S;
    OGSX_GL double* A = gl(S.a.A);
    OGSX_GL double* b = gl(S.a.b);
    OGSX_GL double* uo = gl(S.uold);
    const OGSX_GL double* x = gl(S.x);
    const int shift = S.shift;
    for (int i = c.lb * kStepFill + (int)threadIdx.x; i < (c.lb + 1) * kStepFill; i += kBlock) {
        if (i < S.nnz) A[i] = 0.0;
        if (i < S.n) {
            b[i] = 0.0;
            if (shift) uo[i] = x[i];
        }
    }
}

template <int ET, int PROC, bool ISO>
__global__ __launch_bounds__(64) void k_step_assemble(StepMap m)
{
    const StepCtx c = step_ctx(m);
    const int t = c.lb * 64 + (int)threadIdx.x;
    if (t >= c.S->ne) return;
    const AsmArgs a = c.S->a;
    assemble_elem<ET, PROC, ISO, true>(a, t);
}

template <int ET, bool GRAV, int G>
__global__ __launch_bounds__(kAsmRowThreads) void k_step_rows(StepMap m)
{
    __shared__ double lds[kAsmRowThreads * (ETraits<ET>::NPE + (GRAV ? 2 : 1))];
    __shared__ uint64_t ldw[kAsmRowThreads];
    const StepCtx c = step_ctx(m);
    const StepArgs& S = *c.S;
    const int64_t r = (c.lb * (int64_t)kAsmRowThreads + threadIdx.x) / G;
    const AsmArgs a = S.a;
    assemble_row<ET, GRAV, G>(a, S.inc_ptr, S.loc, S.ptr, S.n, r, lds, ldw);
}

// b[node] += the node's source values in list order (as k_add_rhs)
__global__ __launch_bounds__(kBlock) void k_step_sources(StepMap m)
{
    const StepCtx c = step_ctx(m);
    const StepArgs& S = *c.S;
    const int k = c.lb * kBlock + (int)threadIdx.x;
    if (k >= S.st_nu) return;
    add_rhs_node(k, S.st_node, S.st_vptr, S.st_val, S.a.b);
}

// row-only Dirichlet elimination (as k_bc_row)
__global__ __launch_bounds__(kBlock) void k_step_bc_row(StepMap m)
{
    const StepCtx c = step_ctx(m);
    const StepArgs& S = *c.S;
    const int k = c.lb * kBlock + (int)threadIdx.x;
    if (k >= S.bc_m) return;
    bc_row_one(k, S.bc_idx, S.bc_val, S.ptr, S.col, S.diagpos, S.a.A, S.a.b);
}

// Dirichlet flags on / off (as k_bc_mark)
__global__ __launch_bounds__(kBlock) void k_step_bc_mark(StepMap m, int8_t on)
{
    const StepCtx c = step_ctx(m);
    const StepArgs& S = *c.S;
    const int k = c.lb * kBlock + (int)threadIdx.x;
    if (k >= S.bc_m) return;
    bc_mark_one(k, S.bc_idx, S.bc_val, S.bcflag, S.bcg, on);
}

// symmetric elimination (as k_bc_sym) over the member's listed rows with its blocks
__global__ __launch_bounds__(kBlock) void k_step_bc_sym(StepMap m)
{
    const StepCtx c = step_ctx(m);
    const StepArgs& S = *c.S;
    bc_sym_rows(S.bc_rows_m, S.bc_rows, S.ptr, S.col, S.diagpos, S.bcflag, S.bcg, S.a.A, S.a.b,
                c.lb * (kBlock / 32) + (int)(threadIdx.x >> 5), c.nb * (kBlock / 32));
}

}  // namespace ogsx
