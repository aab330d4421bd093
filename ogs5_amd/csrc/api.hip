// api.hip -- host runtime behind include/ogsx.h: device allocation, element
// upload, assembly launch per colour, boundary conditions, and the Krylov
// drivers (FP64 and FP32-inside-FP64 iterative refinement).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <climits>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <map>
#include <thread>
#include <vector>

#include "kernels.hpp"
#include "ogsx_internal.h"

using namespace ogsx;

namespace {
thread_local std::string g_last_error;

int fail(int code, const std::string& msg)
{
    g_last_error = msg;
    return code;
}

#define HIP_TRY(expr)                                                                  \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess)                                                          \
            return fail(OGSX_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Blocks of the n-sized kernels of one system (each block a contiguous chunk
// of rows).  G sets the partial-sum layout, so it is a function of n alone,
// never tuned at run time: one row per thread up to OGSX_FILL_BLOCKS blocks,
// then OGSX_ROWS_PER_BLOCK rows per block, at most kMaxGrid (2048) blocks
// (the partial sums' reduction loads all of them in one batch).
// Few rows per block keep mid-size systems (latency-bound) spread over the
// CUs; many rows per block cut block dispatches and the per-block reduction
// of partial sums for large systems and for groups of many systems.
#ifndef OGSX_ROWS_PER_BLOCK
#define OGSX_ROWS_PER_BLOCK 1024
#endif
#ifndef OGSX_FILL_BLOCKS
#define OGSX_FILL_BLOCKS 256
#endif

int grid_for(int64_t n)
{
    int64_t g = (n + OGSX_ROWS_PER_BLOCK - 1) / OGSX_ROWS_PER_BLOCK;
    g = std::max<int64_t>(g, std::min<int64_t>((n + kBlock - 1) / kBlock, OGSX_FILL_BLOCKS));
    g = (g + kNumXcd - 1) / kNumXcd * kNumXcd;
    return (int)std::max<int64_t>(kNumXcd, std::min<int64_t>(g, kMaxGrid));
}

// Whether some of the G row chunks of an n-row system hold fewer than the
// 4 kBlock rows of one full batch of the CG vector kernels (their SMALL form).
bool chunk_rows_small(int64_t n, int G)
{
    return n / G < 4 * kBlock;
}

int blocks_for(int64_t n, int per)
{
    int64_t b = (n + per - 1) / per;
    return (int)std::max<int64_t>(1, std::min<int64_t>(b, 65535));
}
}  // namespace

// Device-ready arrays of a pattern (32-bit CRS, row tiles, SELL-64 layout),
// built on the host together with the pattern, so that creating a system is
// allocation and upload only.
struct PatternLayout {
    std::vector<int> ptr32, col32, dg32;
    std::vector<int> trow, tcb;
    std::vector<uint16_t> tc16;
    bool tile16 = false;
    std::vector<int> sptr, sw, scb, sc32, ssrc;
    std::vector<uint16_t> sc16;
    bool sell16 = false;
    int64_t npad = 0;
    // stencil dictionary (formats 13, 14): pattern id per row, padded to
    // whole slices, and the column offsets of every pattern
    std::vector<uint8_t> spid;
    std::vector<int> pdelta;
    int pat_w = 0, pat_cnt = 0;
    bool pat_ok = false;
    // slice triples (formats 19, 20): per slice its stencil id and first
    // entry, per stencil its triples, and the CRS position of every entry
    std::vector<int> tptr, tcnt, toff, tsrc;
    std::vector<uint16_t> tid;
    int64_t npad3 = 0;
    bool tri_ok = false;
    // symmetric diagonal storage (format 25): the positive column offsets of
    // the pattern, ascending, when there are at most kDiaMax of them
    std::vector<int> doff;
    std::vector<int> dsrc;  // CRS position of every stored entry, -1 for a zero
    bool dia_ok = false;
};

struct ogsx_pattern {
    PatternHost h;
    PatternLayout L;
    // systems created on this pattern and not yet destroyed: they keep
    // pointers into h and L, so the pattern outlives them
    mutable std::atomic<int> users{0};
};

template <typename T>
struct DevVecs {
    T *x = nullptr, *b = nullptr, *d = nullptr, *r = nullptr, *p = nullptr, *s = nullptr,
      *z = nullptr, *r0 = nullptr, *v = nullptr, *t = nullptr, *ph = nullptr, *sh = nullptr;
    KVec<T> kv() const { return KVec<T>{x, b, d, r, p, s, z, r0, v, t, ph, sh}; }
};

struct ProfAcc {
    double ms = 0.0;
    int64_t launches = 0;
};

namespace {
struct MemberSet;

// Host-side record of one system's solve (one member of a solve set).
struct SolveCtx {
    int active = 0;  // solve in progress
    int method = OGSX_CG;
    int kmethod = OGSX_CG;  // kernel set (kernel_method)
    ogsx_solver_config cfg{};
    bool ir = false;  // FP32 recurrence inside FP64 refinement
    int total = 0, updates = 0;
    int converged = 0, ret = 0;
    double error = 0.0, bnorm = 0.0;
    // convergence rate of the running recurrence segment, for batch sizing
    int seg_key = -1, seg_it0 = 0;
    double seg_e0 = 0.0;
};
}  // namespace

struct ogsx_system {
    const PatternHost* pat = nullptr;
    const ogsx_pattern* owner = nullptr;  // the pattern pat and lay point into
    int precision = OGSX_FP64;
    int n = 0;
    int64_t nnz = 0;
    int G = 8;
    int fmt64 = 0;  // SpMV format of the FP64 matrix (0 = autotune on first use)
    int fmt32 = 0;  // SpMV format of the FP32 matrix
    // row tiles of the tiled SpMV
    int n_tiles = 0;
    bool tile16 = false;  // every tile's columns fit 16-bit offsets
    int *tile_row = nullptr, *tile_cb = nullptr;
    uint16_t* col16 = nullptr;
    bool extra_layouts = false;  // tiles and 32-bit SELL columns on the device
    const PatternLayout* lay = nullptr;
    std::vector<std::pair<int, double>> tune_log;  // (format, ms) of the last tuning
    size_t tune_n64 = 0;      // leading FP64 entries of tune_log (then FP32)
    bool fmt_forced = false;  // formats set by ogsx_set_spmv_variant
    // SELL-64 copy of the matrix (values refreshed from A when A changed)
    int n_slices = 0;
    int64_t npad = 0;
    bool sell16 = false;
    int *sell_ptr = nullptr, *sell_w = nullptr, *sell_cb = nullptr, *sell_c32 = nullptr,
        *sell_src = nullptr;
    uint16_t* sell_c16 = nullptr;
    bool pat_ok = false;  // stencil dictionary built (formats 13, 14, 17, 18)
    uint8_t* sell_pid = nullptr;
    int* pat_delta = nullptr;
    int pat_w = 0, pat_cnt = 0;
    bool tri_ok = false;  // slice triples built (formats 19, 20)
    int64_t npad3 = 0;
    int *tri_ptr = nullptr, *tri_cnt = nullptr, *tri_off = nullptr, *tri_src = nullptr;
    uint16_t* tri_id = nullptr;
    bool dia_ok = false;  // symmetric diagonal FP32 copy available (format 25)
    int* dia_off = nullptr;
    int* dia_src = nullptr;  // CRS position of every stored entry (-1: zero)
    float* dia_val32 = nullptr;   // FP32 copy (FP32-IR systems)
    double* dia_val64 = nullptr;  // FP64 copy (solves, FP64 residuals)
    int dia_k = 0;
    uint64_t dia32_version = 0, dia64_version = 0;
    int fmt64_tuned = 0;  // the tuner's FP64 format while format 25 stands in
    int fmt32_tuned = 0;  // the tuned FP32 format while format 25 stands in for it
    // values in SELL-64 (layout 0) or slice-triple (layout 1) order, as the
    // precision's format reads them
    double* sval64 = nullptr;
    float* sval32 = nullptr;
    uint64_t a_version = 1, sval64_version = 0, sval32_version = 0, a32_version = 0;
    int sval64_layout = 0, sval32_layout = 0;
    // FP32 copies (A32, sval32) hold A (0) or S A S with S from the Jacobi
    // diagonal (1) or S = I (2)
    int a32_scaled = 0, sval32_scaled = 0;
    hipStream_t stream = nullptr;
    // pattern on device
    int *ptr = nullptr, *col = nullptr, *diagpos = nullptr;
    int *conn = nullptr, *color_elems = nullptr, *slot = nullptr;
    // two-pass assembly: node -> element incidence and the element buffers,
    // allocated on the first assembly that uses them
    int *inc_ptr = nullptr, *inc_ele = nullptr, *asm_pos = nullptr;
    uint8_t* asm_loc = nullptr;
    double* asm_eb = nullptr;  // element rows [E*npe][npe], then right-hand sides
                               // [E*npe][rb_stride] (the second of a pair is
                               // the gravity term or 0)
    // values
    double* A = nullptr;
    float* A32 = nullptr;
    double* uold = nullptr;
    double* ic = nullptr;  // OGSX_VEC_IC: a stored node-value vector
    DevVecs<double> v64;
    DevVecs<float> v32;
    double* r64 = nullptr;  // IR residual
    double* part = nullptr; // partial sums [4][G] x 3 sets
    double* red = nullptr;  // reduced scalars
    KrylovState* st = nullptr;
    // device descriptor of this system and its per-solve parameters
    SysDev devh{};
    SysDev* dev = nullptr;
    KrylovParams* prm = nullptr;
    KrylovParams* prm_host = nullptr;  // pinned
    IRState* irs = nullptr;            // IR systems
    IRState* irs_host = nullptr;       // pinned
    SolveCtx sc;
    MemberSet* solo = nullptr;         // set of one for ogsx_solve
    ogsx_group* group = nullptr;       // group the system belongs to, if any
    hipEvent_t ready = nullptr;        // marks the system's stream for another stream
    // BC / ST
    int8_t* bcflag = nullptr;
    double* bcg = nullptr;
    std::vector<int64_t> last_bc_idx, last_st_idx;
    std::vector<double> last_bc_val, last_st_val;
    int *bc_idx_d = nullptr, *st_node_d = nullptr, *st_vptr_d = nullptr;
    double *bc_val_d = nullptr, *st_val_d = nullptr;
    int bc_cap = 0, bc_val_cap = 0, st_cap = 0, st_nu = 0;
    int bc_m = 0;
    // rows the symmetric elimination can change: the Dirichlet rows and the
    // rows with a Dirichlet column (the pattern is symmetric), ascending
    int* bc_rows_d = nullptr;
    int bc_rows_cap = 0, bc_rows_m = 0;
    // elements
    bool elements_set = false;
    bool use_vel = false;
    size_t vel_cap = 0;
    int process = 0, dim = 0, ngp = 0, cond_stride = 1;
    double tuf = 1, mu = 1, geo_area = 1, adv_coef = 0, rho_g = 0;
    bool tensor_sym = true;   // every element tensor symmetric
    bool a_symmetric = true;  // A is symmetric by construction (symmetric
                              // tensors, no advection, no row-only Dirichlet)
    double *ex = nullptr, *ey = nullptr, *ez = nullptr, *mass_coef = nullptr, *cond = nullptr,
           *vel = nullptr, *tabN = nullptr, *tabdN = nullptr, *tabW = nullptr;
    double* gvel = nullptr;  // Gauss-point velocity computed for the host only
    size_t gvel_cap = 0;
    // persistent Linear_EQS::bNorm (equation_class.cpp:135)
    double bnorm = 1.0;
    // profiling
    bool prof = false;
    std::vector<hipEvent_t> ev_pool;
    std::vector<std::pair<int, int>> ev_pending;  // (kind, pool index of start)
    size_t ev_next = 0;
    ProfAcc prof_acc[4];
    std::vector<void*> allocs;
    int64_t bytes = 0;  // device bytes allocated
};

namespace {

// Systems allocate from the device's stream-ordered memory pool when the
// device has one (freed memory stays in the pool: creating and destroying
// systems in a serving loop then costs no device-wide synchronisation and
// no trip to the driver); plain hipMalloc otherwise.
bool use_pool()
{
    static std::once_flag once;
    static bool pool = false;
    std::call_once(once, [] {
        int dev = 0, sup = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&sup, hipDeviceAttributeMemoryPoolsSupported, dev) != hipSuccess ||
            !sup)
            return;
        hipMemPool_t mp;
        if (hipDeviceGetDefaultMemPool(&mp, dev) != hipSuccess) return;
        uint64_t keep = UINT64_MAX;
        pool = hipMemPoolSetAttribute(mp, hipMemPoolAttrReleaseThreshold, &keep) == hipSuccess;
    });
    return pool;
}

template <typename T>
int dalloc(ogsx_system* s, T** p, size_t count)
{
    if (count == 0) count = 1;
    void* q = nullptr;
    hipError_t e = use_pool() ? hipMallocAsync(&q, count * sizeof(T), s->stream)
                              : hipMalloc(&q, count * sizeof(T));
    if (e != hipSuccess)
        return fail(OGSX_ENOMEM, std::string("hipMalloc failed: ") + hipGetErrorString(e));
    s->allocs.push_back(q);
    s->bytes += (int64_t)(count * sizeof(T));
    *p = (T*)q;
    return OGSX_OK;
}

#define OK_TRY(...)                     \
    do {                                \
        int rc_ = (__VA_ARGS__);        \
        if (rc_ != OGSX_OK) return rc_; \
    } while (0)

inline bool is_pat(int f)
{
    return f == kFmtSellPat8 || f == kFmtSellPat7 || f == kFmtSellPat8C || f == kFmtSellPat7C ||
           f == kFmtSellPat16 || f == kFmtSellPat16C;
}

inline bool is_tri(int f)
{
    return f == kFmtTri5 || f == kFmtTri4C;
}

bool valid_fmt(const ogsx_system* s, int f)
{
    if (f == kFmtTile32 || f == kFmtSell32) return true;
    if (f == kFmtTile16) return s->tile16;
    if (f == kFmtSell16 || f == kFmtSell16U8 || f == kFmtSell16NT || f == kFmtSell16U16 ||
        f == kFmtSell16X2 || f == kFmtSell16U7)
        return s->sell16;
    if (is_pat(f)) return s->pat_ok;
    if (is_tri(f)) return s->tri_ok;
    return f == kFmtCrs1 || f == 4 || f == 8 || f == 16 || f == 32 || f == 64;
}

// Formats that read the row tiles or the 32-bit SELL columns.
inline bool needs_extra_layouts(int f)
{
    return f == kFmtTile32 || f == kFmtTile16 || f == kFmtSell32;
}

template <typename VT>
MatView<VT> mat_view(const ogsx_system* s, const VT* val)
{
    const VT* sv = nullptr;
    if constexpr (sizeof(VT) == 8)
        sv = s->sval64;
    else
        sv = s->sval32;
    const VT* dval = nullptr;
    if constexpr (sizeof(VT) == 4)
        dval = s->dia_val32;
    else
        dval = s->dia_val64;
    return MatView<VT>{s->n,        s->n_tiles,  s->ptr,      s->col,     s->col16,
                       val,         s->tile_row, s->tile_cb,  s->n_slices, s->sell_ptr,
                       s->sell_w,   s->sell_cb,  s->sell_c16, s->sell_c32, sv,
                       s->sell_pid, s->pat_delta, s->pat_w,   s->pat_cnt,  s->tri_ptr,
                       s->tri_id,   s->tri_cnt,  s->tri_off,  dval ? s->dia_off : nullptr,
                       dval,        dval ? s->dia_k : 0};
}

inline bool is_sell(int f)
{
    return f == kFmtSell32 || f == kFmtSell16 || f == kFmtSell16U8 || f == kFmtSell16NT ||
           f == kFmtSell16U16 || f == kFmtSell16X2 || f == kFmtSell16U7 || is_pat(f) || is_tri(f);
}

// SELL-64 layout of the pattern: slice widths, offsets, columns (32-bit and,
// when every slice's columns span < 65536, 16-bit offsets from the slice's
// smallest column) and the CRS position of every SELL entry (-1 = padding).
void build_sell(const PatternHost& h, std::vector<int>& sptr, std::vector<int>& sw,
                std::vector<int>& scb, std::vector<int>& c32, std::vector<uint16_t>& c16,
                std::vector<int>& src, bool& fits16)
{
    const int64_t ns = (h.n + kSellC - 1) / kSellC;
    sptr.assign(ns + 1, 0);
    sw.assign(ns, 0);
    scb.assign(ns, 0);
    fits16 = true;
    for (int64_t sl = 0; sl < ns; sl++) {
        int w = 0;
        int64_t lo = INT64_MAX, hi = -1;
        for (int64_t r = sl * kSellC; r < std::min<int64_t>(h.n, (sl + 1) * kSellC); r++) {
            w = std::max<int>(w, (int)(h.ptr[r + 1] - h.ptr[r]));
            for (int64_t k = h.ptr[r]; k < h.ptr[r + 1]; k++) {
                lo = std::min(lo, h.col[k]);
                hi = std::max(hi, h.col[k]);
            }
        }
        if (hi < 0) lo = hi = 0;
        if (hi - lo > 65535) fits16 = false;
        sw[sl] = w;
        scb[sl] = (int)lo;
        sptr[sl + 1] = sptr[sl] + w * kSellC;
    }
    const int64_t npad = sptr[ns];
    c32.assign(npad, 0);
    src.assign(npad, -1);
    for (int64_t sl = 0; sl < ns; sl++)
        for (int l = 0; l < kSellC; l++) {
            const int64_t r = sl * kSellC + l;
            for (int j = 0; j < sw[sl]; j++) {
                const int64_t p = sptr[sl] + (int64_t)j * kSellC + l;
                if (r < h.n && h.ptr[r] + j < h.ptr[r + 1]) {
                    c32[p] = (int)h.col[h.ptr[r] + j];
                    src[p] = (int)(h.ptr[r] + j);
                } else {
                    c32[p] = scb[sl];  // padding: any valid column, value 0
                }
            }
        }
    c16.assign(fits16 ? npad : 0, 0);
    if (fits16)
        for (int64_t sl = 0; sl < ns; sl++)
            for (int64_t p = sptr[sl]; p < sptr[sl + 1]; p++) c16[p] = (uint16_t)(c32[p] - scb[sl]);
}

// Stencil dictionary of the pattern: rows grouped by the offsets of their
// columns from the row index (in CRS order).  Applicable when at most 255
// distinct offset lists occur and they fit the LDS stage (kPatLds entries);
// one more all-zero entry serves the padding rows of the last slice.
void build_pattern_dict(const PatternHost& h, const std::vector<int>& sw, PatternLayout& L)
{
    L.pat_ok = false;
    L.spid.clear();
    L.pdelta.clear();
    const int64_t ns = (int64_t)sw.size();
    int wmax = 0;
    for (int w : sw) wmax = std::max(wmax, w);
    if (h.n == 0 || wmax == 0) return;
    std::map<std::vector<int>, int> ids;
    std::vector<std::vector<int>> pats;
    std::vector<uint8_t> pid(ns * kSellC, 0);
    std::vector<int> key, prev;
    int prev_id = -1;
    for (int64_t r = 0; r < h.n; r++) {
        key.clear();
        for (int64_t k = h.ptr[r]; k < h.ptr[r + 1]; k++) key.push_back((int)(h.col[k] - r));
        int id;
        if (prev_id >= 0 && key == prev) {
            id = prev_id;
        } else {
            auto it = ids.find(key);
            if (it == ids.end()) {
                if (pats.size() >= 255) return;
                id = (int)pats.size();
                ids.emplace(key, id);
                pats.push_back(key);
            } else {
                id = it->second;
            }
            prev = key;
            prev_id = id;
        }
        pid[r] = (uint8_t)id;
    }
    const int pad_id = (int)pats.size();  // all-zero offsets: padding rows
    pats.emplace_back();
    if ((int64_t)pats.size() * wmax > kPatLds) return;
    for (int64_t r = h.n; r < ns * kSellC; r++) pid[r] = (uint8_t)pad_id;
    L.pdelta.assign(pats.size() * wmax, 0);
    for (size_t p = 0; p < pats.size(); p++)
        for (size_t j = 0; j < pats[p].size(); j++) L.pdelta[p * wmax + j] = pats[p][j];
    L.spid.swap(pid);
    L.pat_w = wmax;
    L.pat_cnt = (int)pats.size();
    L.pat_ok = true;
}

// Brings the SELL value copies up to date with A (after assembly / BC), in
// the layout of the precision's format (SELL-64 or slice triples).
// Refreshes the FP64 symmetric diagonal copy from A when the FP64 format is
// the diagonal one and A changed since the last copy.
int sync_dia64(ogsx_system* s, hipStream_t st)
{
    if (s->fmt64 != kFmtSymDia || s->dia64_version == s->a_version) return OGSX_OK;
    if (!s->dia_val64) return fail(OGSX_ESTATE, "no FP64 diagonal copy on this system");
    k_to_dia_g<double><<<s->G, kBlock, 0, st>>>(Grp{s->dev, nullptr, nullptr});
    HIP_TRY(hipGetLastError());
    s->dia64_version = s->a_version;
    return OGSX_OK;
}

int sync_sell(ogsx_system* s, bool want64, bool want32)
{
    const Grp g1{s->dev, nullptr, nullptr};
    const int lay64 = is_tri(s->fmt64) ? 1 : 0, lay32 = is_tri(s->fmt32) ? 1 : 0;
    const bool stale64 = want64 && (s->sval64_version != s->a_version || s->sval64_layout != lay64);
    const bool stale32 = want32 && s->sval32 &&
                         (s->sval32_version != s->a_version || s->sval32_layout != lay32 ||
                          s->sval32_scaled != 0);
    if (stale64 && stale32 && lay64 == lay32) {
        k_to_sell_both_g<<<s->G, kBlock, 0, s->stream>>>(g1, lay64);
        s->sval64_version = s->sval32_version = s->a_version;
        s->sval64_layout = s->sval32_layout = lay64;
        s->sval32_scaled = 0;
    }
    if (want64 && (s->sval64_version != s->a_version || s->sval64_layout != lay64)) {
        k_to_sell_g<double><<<s->G, kBlock, 0, s->stream>>>(g1, lay64);
        s->sval64_version = s->a_version;
        s->sval64_layout = lay64;
    }
    if (want32 && s->sval32 &&
        (s->sval32_version != s->a_version || s->sval32_layout != lay32 || s->sval32_scaled != 0)) {
        k_to_sell_g<float><<<s->G, kBlock, 0, s->stream>>>(g1, lay32);
        s->sval32_version = s->a_version;
        s->sval32_layout = lay32;
        s->sval32_scaled = 0;
    }
    HIP_TRY(hipGetLastError());
    return OGSX_OK;
}

// Row tiles: consecutive rows with at most kTileNnz nonzeros and kBlock rows
// (a row longer than kTileNnz gets a tile of its own); per tile the smallest
// column, and whether all column offsets fit 16 bits.
void build_tiles(const PatternHost& h, std::vector<int>& trow, std::vector<int>& tcb,
                 std::vector<uint16_t>& c16, bool& fits16)
{
    trow.clear();
    tcb.clear();
    int64_t r = 0;
    fits16 = true;
    while (r < h.n) {
        int64_t r1 = r, nz = 0;
        while (r1 < h.n && r1 - r < kBlock) {
            const int64_t len = h.ptr[r1 + 1] - h.ptr[r1];
            if (nz + len > kTileNnz && r1 > r) break;
            nz += len;
            r1++;
            if (nz > kTileNnz) break;
        }
        int64_t lo = INT64_MAX, hi = -1;
        for (int64_t k = h.ptr[r]; k < h.ptr[r1]; k++) {
            lo = std::min(lo, h.col[k]);
            hi = std::max(hi, h.col[k]);
        }
        if (hi < 0) lo = hi = 0;
        if (hi - lo > 65535) fits16 = false;
        trow.push_back((int)r);
        tcb.push_back((int)lo);
        r = r1;
    }
    trow.push_back((int)h.n);
    c16.assign(h.nnz, 0);
    if (fits16)
        for (size_t t = 0; t + 1 < trow.size(); t++)
            for (int64_t k = h.ptr[trow[t]]; k < h.ptr[trow[t + 1]]; k++)
                c16[k] = (uint16_t)(h.col[k] - tcb[t]);
}

// Slice triples: each slice's stencil is the sorted union of its rows'
// column offsets, cut greedily into triples o, o+1, o+2 (an offset the
// union lacks becomes a zero entry).  Built when every slice needs at most
// kTriMax triples, there are at most 65535 distinct slice stencils, and the
// stored entries are at most twice the nonzeros (grid-numbered meshes: 1.01
// on the 128^3 hexes; small meshes pay more at their boundaries); otherwise
// the formats are not offered.
void build_triples(const PatternHost& h, PatternLayout& L)
{
    const int64_t ns = (h.n + kSellC - 1) / kSellC;
    if (h.n == 0) return;
    std::map<std::vector<int>, int> ids;
    std::vector<std::vector<int>> stencils;
    std::vector<uint16_t> tid(ns);
    std::vector<int> tptr(ns + 1, 0);
    std::vector<int> u, starts, key, prev;
    int64_t total = 0;
    for (int64_t sl = 0; sl < ns; sl++) {
        u.clear();
        prev.clear();
        const int64_t r1 = std::min<int64_t>(h.n, (sl + 1) * kSellC);
        for (int64_t r = sl * kSellC; r < r1; r++) {
            key.clear();
            for (int64_t k = h.ptr[r]; k < h.ptr[r + 1]; k++) key.push_back((int)(h.col[k] - r));
            if (key == prev) continue;  // neighbouring rows mostly share their offsets
            u.insert(u.end(), key.begin(), key.end());
            prev.swap(key);
        }
        std::sort(u.begin(), u.end());
        u.erase(std::unique(u.begin(), u.end()), u.end());
        if (u.empty()) return;
        starts.clear();
        for (size_t i = 0; i < u.size();) {
            const int o = u[i];
            starts.push_back(o);
            while (i < u.size() && u[i] <= o + 2) i++;
        }
        if ((int)starts.size() > kTriMax) return;
        auto it = ids.find(starts);
        int id;
        if (it == ids.end()) {
            if (stencils.size() >= 65535) return;
            id = (int)stencils.size();
            ids.emplace(starts, id);
            stencils.push_back(starts);
        } else {
            id = it->second;
        }
        tid[sl] = (uint16_t)id;
        total += 3 * kSellC * (int64_t)starts.size();
        if (total > 2 * h.nnz + 3 * kSellC || total >= (int64_t(1) << 31)) return;
        tptr[sl + 1] = (int)total;
    }
    L.tsrc.assign(total, -1);
    for (int64_t sl = 0; sl < ns; sl++) {
        const std::vector<int>& st = stencils[tid[sl]];
        const int64_t r1 = std::min<int64_t>(h.n, (sl + 1) * kSellC);
        for (int64_t r = sl * kSellC; r < r1; r++) {
            const int l = (int)(r - sl * kSellC);
            // the row's columns and the stencil's columns both ascend: merge
            int64_t q = h.ptr[r];
            const int64_t q1 = h.ptr[r + 1];
            for (size_t k = 0; k < st.size(); k++)
                for (int t = 0; t < 3; t++) {
                    const int64_t c = r + st[k] + t;
                    while (q < q1 && h.col[q] < c) q++;
                    if (q < q1 && h.col[q] == c)
                        L.tsrc[tptr[sl] + kSellC * (3 * (int64_t)k + t) + l] = (int)q;
                }
        }
    }
    L.tcnt.assign(stencils.size(), 0);
    L.toff.assign(stencils.size() * kTriMax, 0);
    for (size_t q = 0; q < stencils.size(); q++) {
        L.tcnt[q] = (int)stencils[q].size();
        for (size_t k = 0; k < stencils[q].size(); k++) L.toff[q * kTriMax + k] = stencils[q][k];
    }
    L.tid.swap(tid);
    L.tptr.swap(tptr);
    L.npad3 = total;
    L.tri_ok = true;
}

// Positive column offsets of the pattern (format 25), when every row's
// entries right of the diagonal sit at one of at most kDiaMax offsets.
void build_dia(const PatternHost& h, PatternLayout& L)
{
    std::vector<int> offs, row, prev;
    for (int64_t i = 0; i < h.n; i++) {
        row.clear();
        for (int64_t k = h.ptr[i]; k < h.ptr[i + 1]; k++)
            if (h.col[k] > i) row.push_back((int)(h.col[k] - i));
        if (row == prev) continue;  // most rows repeat their neighbour's offsets
        prev = row;
        for (int o : row)
            if (std::find(offs.begin(), offs.end(), o) == offs.end()) {
                if ((int)offs.size() == kDiaMax) return;  // too many: not built
                offs.push_back(o);
            }
    }
    std::sort(offs.begin(), offs.end());
    if (offs.empty()) return;
    // position map: entry (k + 1) n + i holds (i, i + offs[k]), entry i the
    // diagonal (rows and offsets both ascending: one merge per row)
    const int K = (int)offs.size();
    const int64_t n = h.n;
    L.dsrc.assign((size_t)(K + 1) * (size_t)n, -1);
    for (int64_t i = 0; i < n; i++) {
        int k = 0;
        for (int64_t p = h.ptr[i]; p < h.ptr[i + 1]; p++) {
            const int64_t o = h.col[p] - i;
            if (o == 0) L.dsrc[i] = (int)p;
            if (o <= 0) continue;
            while (k < K && offs[k] < o) k++;
            if (k < K && offs[k] == o) L.dsrc[(size_t)(k + 1) * n + i] = (int)p;
        }
    }
    L.doff = offs;
    L.dia_ok = true;
}

void build_layout(const PatternHost& h, PatternLayout& L)
{
    L.ptr32.resize(h.n + 1);
    L.col32.resize(h.nnz);
    L.dg32.resize(h.n);
    for (int64_t i = 0; i <= h.n; i++) L.ptr32[i] = (int)h.ptr[i];
    for (int64_t k = 0; k < h.nnz; k++) L.col32[k] = (int)h.col[k];
    for (int64_t i = 0; i < h.n; i++) L.dg32[i] = (int)h.diag[i];
    build_tiles(h, L.trow, L.tcb, L.tc16, L.tile16);
    if (!L.tile16) L.tc16.clear();
    build_sell(h, L.sptr, L.sw, L.scb, L.sc32, L.sc16, L.ssrc, L.sell16);
    L.npad = L.sptr.back();
    build_pattern_dict(h, L.sw, L);
    build_triples(h, L);
    build_dia(h, L);
}

// Device bytes a system on pattern p allocates: ogsx_system_create plus
// ogsx_system_set_elements without Gauss-point velocities (counted as the
// allocations ask for them, one element for an empty array).
int64_t system_bytes(const ogsx_pattern* p, int precision)
{
    const PatternHost& h = p->h;
    const PatternLayout& L = p->L;
    const int64_t n = h.n, nnz = h.nnz, E = h.n_elem, G = grid_for(n);
    auto b = [](int64_t count, int64_t size) { return std::max<int64_t>(count, 1) * size; };
    const int ed = edim_of(h.etype);
    int64_t t = b(n + 1, 4) + b(nnz, 4) + b(n, 4) + b((int64_t)h.conn.size(), 4) +
                b((int64_t)h.color_elems.size(), 4) + b((int64_t)h.conn.size() * h.npe, 4) + b(nnz, 8) +
                2 * b(n, 8) + 12 * b(n, 8) + b(16 * G, 8) + b(8, 8) + b(1, sizeof(KrylovState)) +
                b(n, 1) + b(n, 8) + b(1, sizeof(SysDev)) + b(1, sizeof(KrylovParams));
    if (precision == OGSX_FP32_IR) t += b(nnz, 4) + 12 * b(n, 4) + b(n, 8) + b(1, sizeof(IRState));
    if (!L.sell16)  // tiles and 32-bit SELL columns, built at creation only then
        t += b((int64_t)L.trow.size(), 4) + b((int64_t)L.tcb.size(), 4) +
             (L.tile16 ? b((int64_t)L.tc16.size(), 2) : 0) + b((int64_t)L.sc32.size(), 4);
    t += b((int64_t)L.sptr.size(), 4) + b((int64_t)L.sw.size(), 4) + b((int64_t)L.scb.size(), 4) +
         b((int64_t)L.ssrc.size(), 4) + b(L.npad, 8) +
         (precision == OGSX_FP32_IR ? b(L.npad, 4) : 0) + (L.sell16 ? b((int64_t)L.sc16.size(), 2) : 0);
    if (L.pat_ok) t += b((int64_t)L.spid.size(), 1) + b((int64_t)L.pdelta.size(), 4);
    if (L.tri_ok) {  // tables, entry map, value arrays sized for the longer layout
        const int64_t extra = std::max<int64_t>(0, L.npad3 - L.npad);
        t += b((int64_t)L.tptr.size(), 4) + b((int64_t)L.tid.size(), 2) + b((int64_t)L.tcnt.size(), 4) +
             b((int64_t)L.toff.size(), 4) + b(L.npad3, 4) + extra * 8 +
             (precision == OGSX_FP32_IR ? extra * 4 : 0);
    }
    if (L.dia_ok)  // offsets, position map, FP64 values (and FP32 ones for IR)
        t += b((int64_t)L.doff.size(), 4) + b(((int64_t)L.doff.size() + 1) * n, 4) +
             b(((int64_t)L.doff.size() + 1) * n, 8) +
             (precision == OGSX_FP32_IR ? b(((int64_t)L.doff.size() + 1) * n, 4) : 0);
    // element data
    t += 3 * b(n, 8) + b(E, 8) + b(E * ed * ed, 8) + b(kMaxGaussPts * 8, 8) +
         b(kMaxGaussPts * 24, 8) + b(kMaxGaussPts * 3, 8);
    // two-pass assembly: incidence, row positions, element matrices and
    // right-hand sides
    if (!h.inc_ptr.empty() && h.row_max <= kAsmRowMax) {
        const int64_t np = h.npe;
        t += b(n + 1, 4) + 2 * b(E * np, 4) + b(E * np * np, 1) + b(E * np * (np + 2), 8);
    }
    return t;
}

// ---- profiling helpers ----------------------------------------------------
int prof_begin(ogsx_system* s, int kind)
{
    if (!s->prof) return -1;
    if (s->ev_next + 2 > s->ev_pool.size()) {
        // pool exhausted: drain pending pairs
        HIP_TRY(hipStreamSynchronize(s->stream));
        for (auto& pr : s->ev_pending) {
            float ms = 0.f;
            HIP_TRY(hipEventElapsedTime(&ms, s->ev_pool[pr.second], s->ev_pool[pr.second + 1]));
            s->prof_acc[pr.first].ms += ms;
            s->prof_acc[pr.first].launches++;
        }
        s->ev_pending.clear();
        s->ev_next = 0;
    }
    int k = (int)s->ev_next;
    s->ev_next += 2;
    (void)hipEventRecord(s->ev_pool[k], s->stream);
    s->ev_pending.push_back({kind, k});
    return k;
}
void prof_end(ogsx_system* s, int k)
{
    if (k >= 0) (void)hipEventRecord(s->ev_pool[k + 1], s->stream);
}
int prof_flush(ogsx_system* s)
{
    if (!s->prof || s->ev_pending.empty()) return OGSX_OK;
    HIP_TRY(hipStreamSynchronize(s->stream));
    for (auto& pr : s->ev_pending) {
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, s->ev_pool[pr.second], s->ev_pool[pr.second + 1]));
        s->prof_acc[pr.first].ms += ms;
        s->prof_acc[pr.first].launches++;
    }
    s->ev_pending.clear();
    s->ev_next = 0;
    return OGSX_OK;
}

// ---- SpMV-family launch dispatch on lanes-per-row -------------------------
#define FMT_DISPATCH(fmt, CALL)                                 \
    switch (fmt) {                                              \
        case kFmtTile32: { constexpr int L = kFmtTile32; CALL; } break; \
        case kFmtTile16: { constexpr int L = kFmtTile16; CALL; } break; \
        case kFmtSell32: { constexpr int L = kFmtSell32; CALL; } break; \
        case kFmtSell16: { constexpr int L = kFmtSell16; CALL; } break; \
        case kFmtSell16U8: { constexpr int L = kFmtSell16U8; CALL; } break; \
        case kFmtSell16NT: { constexpr int L = kFmtSell16NT; CALL; } break; \
        case kFmtSell16U16: { constexpr int L = kFmtSell16U16; CALL; } break; \
        case kFmtSell16X2: { constexpr int L = kFmtSell16X2; CALL; } break; \
        case kFmtSell16U7: { constexpr int L = kFmtSell16U7; CALL; } break; \
        case kFmtSellPat8: { constexpr int L = kFmtSellPat8; CALL; } break; \
        case kFmtSellPat7: { constexpr int L = kFmtSellPat7; CALL; } break; \
        case kFmtSellPat8C: { constexpr int L = kFmtSellPat8C; CALL; } break; \
        case kFmtSellPat7C: { constexpr int L = kFmtSellPat7C; CALL; } break; \
        case kFmtSellPat16: { constexpr int L = kFmtSellPat16; CALL; } break; \
        case kFmtSellPat16C: { constexpr int L = kFmtSellPat16C; CALL; } break; \
        case kFmtTri5: { constexpr int L = kFmtTri5; CALL; } break;           \
        case kFmtTri4C: { constexpr int L = kFmtTri4C; CALL; } break;         \
        case kFmtCrs1: { constexpr int L = kFmtCrs1; CALL; } break;           \
        case kFmtSymDia: { constexpr int L = kFmtSymDia; CALL; } break;       \
        case 4: { constexpr int L = 4; CALL; } break;           \
        case 8: { constexpr int L = 8; CALL; } break;           \
        case 16: { constexpr int L = 16; CALL; } break;         \
        case 32: { constexpr int L = 32; CALL; } break;         \
        default: { constexpr int L = 64; CALL; } break;         \
    }

template <typename VT>
int fmt_of(const ogsx_system* s)
{
    return sizeof(VT) == 8 ? s->fmt64 : s->fmt32;
}

// The solve's SpMV kernel (k_cg_spmv: s = A p and the partial (p, s)) on
// this system alone; the caller has set the state to running.
template <typename VT, typename T>
void launch_solve_spmv(ogsx_system* s, int fmt)
{
    FMT_DISPATCH(fmt, (k_cg_spmv<L, VT, T><<<s->G, kBlock, 0, s->stream>>>(
                          Grp{s->dev, nullptr, nullptr})));
}

// Autotuner: times the solve's SpMV kernel in every slice-based format
// (SELL-64 variants, stencil dictionary, slice triples) on this matrix and
// keeps the fastest.  All candidates sum each row in CRS order and hand it
// to the same thread of the same block, so the row sums and the dot
// products fused into the solve kernels are bit-identical between them:
// the choice (which depends on timing) never changes a result.  The Krylov
// vectors are scratch at this point; p is set to 1.
template <typename VT, typename T>
int tune_one(ogsx_system* s, int* fmt_out)
{
    // the slice-based formats only: they hand every row to the same thread
    // of the same block, so the dot products fused into the solve kernels
    // round identically with each of them (the tiled formats, selectable
    // by hand, accumulate rows in another order)
    const int cands[] = {kFmtSell16,    kFmtSell16U8, kFmtSell16NT, kFmtSell16U16,
                         kFmtSell16X2,  kFmtSell16U7, kFmtSellPat8, kFmtSellPat7,
                         kFmtSellPat8C, kFmtSellPat7C, kFmtSellPat16, kFmtSellPat16C,
                         kFmtTri5,      kFmtTri4C,    kFmtSell32};
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    T* p = sizeof(T) == 8 ? (T*)s->v64.p : (T*)s->v32.p;
    k_fill<T><<<blocks_for(s->n, kBlock * 4), kBlock, 0, s->stream>>>(s->n, p, (T)1);
    k_set_done<<<1, 64, 0, s->stream>>>(s->st, 0);
    int best = kFmtSell32;
    float best_ms = 1e30f;
    const int reps = 4;
    for (int f : cands) {
        if (!valid_fmt(s, f) || (needs_extra_layouts(f) && !s->extra_layouts)) continue;
        launch_solve_spmv<VT, T>(s, f);
        (void)hipEventRecord(e0, s->stream);
        for (int r = 0; r < reps; r++) launch_solve_spmv<VT, T>(s, f);
        (void)hipEventRecord(e1, s->stream);
        HIP_TRY(hipEventSynchronize(e1));
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
        s->tune_log.push_back({f, ms / reps});
        if (ms < best_ms) {
            best_ms = ms;
            best = f;
        }
    }
    k_set_done<<<1, 64, 0, s->stream>>>(s->st, 4);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    HIP_TRY(hipGetLastError());
    *fmt_out = best;
    return OGSX_OK;
}

int autotune(ogsx_system* s)
{
    if (s->fmt64 == 0) {
        s->tune_log.clear();
        OK_TRY((tune_one<double, double>(s, &s->fmt64)));
        s->tune_n64 = s->tune_log.size();
    }
    if (s->precision == OGSX_FP32_IR && s->fmt32 == 0) OK_TRY((tune_one<float, float>(s, &s->fmt32)));
    return OGSX_OK;
}

double spmv_bytes(const ogsx_system* s, int fmt, int sval, int sx)
{
    const double n = s->n, nnz = (double)s->nnz;
    const int sidx = (fmt == kFmtTile16 || fmt == kFmtSell16 || fmt == kFmtSell16U8 ||
                      fmt == kFmtSell16NT || fmt == kFmtSell16U16 || fmt == kFmtSell16X2 ||
                      fmt == kFmtSell16U7)
                         ? 2
                         : 4;
    if (fmt == kFmtSymDia)  // diagonal and the upper diagonals (zeros included)
        return (double)(s->dia_k + 1) * n * sval + 2 * n * sx;
    if (is_tri(fmt))  // values incl. zero entries, slice table (first entry, stencil id)
        return (double)s->npad3 * sval + 6.0 * s->n_slices + 2 * n * sx;
    if (is_pat(fmt))  // values, slice table, one id per row
        return (double)s->npad * sval + 8.0 * s->n_slices + (double)s->n_slices * kSellC +
               2 * n * sx;
    if (is_sell(fmt))  // stored entries incl. padding, slice table (ptr, width, base)
        return (double)s->npad * (sval + sidx) + 12.0 * s->n_slices + 2 * n * sx;
    const double tiles = (fmt == kFmtTile16 || fmt == kFmtTile32) ? 8.0 * s->n_tiles : 0.0;
    return nnz * (sval + sidx) + (n + 1) * 4 + 2 * n * sx + tiles;
}

// Row tiles (formats 1, 2) and 32-bit SELL columns (format 3): allocated and
// uploaded on first need (the caller synchronises the stream).
int alloc_extra_layouts(ogsx_system* s)
{
    if (s->extra_layouts) return OGSX_OK;
    const PatternLayout& Ly = *s->lay;
    OK_TRY(dalloc(s, &s->tile_row, Ly.trow.size()));
    OK_TRY(dalloc(s, &s->tile_cb, Ly.tcb.size()));
    if (s->tile16) OK_TRY(dalloc(s, &s->col16, Ly.tc16.size()));
    OK_TRY(dalloc(s, &s->sell_c32, Ly.sc32.size()));
    auto up = [&](void* d, const void* src, size_t bytes) {
        return hipMemcpyAsync(d, src, bytes, hipMemcpyHostToDevice, s->stream);
    };
    if (up(s->tile_row, Ly.trow.data(), 4 * Ly.trow.size()) ||
        up(s->tile_cb, Ly.tcb.data(), 4 * Ly.tcb.size()) ||
        (s->tile16 && up(s->col16, Ly.tc16.data(), 2 * Ly.tc16.size())) ||
        up(s->sell_c32, Ly.sc32.data(), 4 * Ly.sc32.size()))
        return fail(OGSX_EHIP, "upload of the tiled / 32-bit SELL layouts failed");
    s->extra_layouts = true;
    return OGSX_OK;
}

template <typename T>
int alloc_vecs(ogsx_system* s, DevVecs<T>& v)
{
    T** f[] = {&v.x, &v.b, &v.d, &v.r, &v.p, &v.s, &v.z, &v.r0, &v.v, &v.t, &v.ph, &v.sh};
    for (T** q : f) OK_TRY(dalloc(s, q, (size_t)s->n));
    return OGSX_OK;
}

}  // namespace

#include "solver.inc"

namespace {
int alloc_two_pass(ogsx_system* s);
bool use_two_pass(const ogsx_system* s);
}  // namespace

// =========================================================================
// C ABI
// =========================================================================
extern "C" {

const char* ogsx_last_error(void) { return g_last_error.c_str(); }

const char* ogsx_version(void) { return "ogsx 0.1 gfx950"; }

int ogsx_device_count(int* count)
{
    if (!count) return fail(OGSX_EINVAL, "null count");
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
    *count = c;
    return OGSX_OK;
}

int ogsx_pattern_create(int64_t n_nodes, int64_t n_elem, int etype, const int64_t* conn,
                        ogsx_pattern** out)
{
    if (!out) return fail(OGSX_EINVAL, "null output handle");
    *out = nullptr;
    auto* p = new (std::nothrow) ogsx_pattern();
    if (!p) return fail(OGSX_ENOMEM, "host allocation failed");
    std::string msg;
    int rc = build_pattern(n_nodes, n_elem, etype, conn, p->h, msg);
    if (rc != OGSX_OK) {
        delete p;
        return fail(rc, msg);
    }
    build_layout(p->h, p->L);
    *out = p;
    return OGSX_OK;
}

int ogsx_pattern_info(const ogsx_pattern* p, int64_t* n_rows, int64_t* nnz, int* n_colors)
{
    if (!p) return fail(OGSX_EINVAL, "null pattern");
    if (n_rows) *n_rows = p->h.n;
    if (nnz) *nnz = p->h.nnz;
    if (n_colors) *n_colors = p->h.n_colors;
    return OGSX_OK;
}

int ogsx_pattern_export(const ogsx_pattern* p, int64_t* row_ptr, int64_t* col, int64_t* diag)
{
    if (!p) return fail(OGSX_EINVAL, "null pattern");
    if (row_ptr) std::memcpy(row_ptr, p->h.ptr.data(), sizeof(int64_t) * (p->h.n + 1));
    if (col) std::memcpy(col, p->h.col.data(), sizeof(int64_t) * p->h.nnz);
    if (diag) std::memcpy(diag, p->h.diag.data(), sizeof(int64_t) * p->h.n);
    return OGSX_OK;
}

int ogsx_pattern_colors(const ogsx_pattern* p, int32_t* color)
{
    if (!p || !color) return fail(OGSX_EINVAL, "null argument");
    std::memcpy(color, p->h.color.data(), sizeof(int32_t) * p->h.n_elem);
    return OGSX_OK;
}

int ogsx_pattern_footprint(const ogsx_pattern* p, int precision, int64_t* bytes)
{
    if (!p || !bytes) return fail(OGSX_EINVAL, "null argument");
    if (precision != OGSX_FP64 && precision != OGSX_FP32_IR) return fail(OGSX_EINVAL, "unknown precision");
    *bytes = system_bytes(p, precision);
    return OGSX_OK;
}

int ogsx_system_bytes(const ogsx_system* s, int64_t* bytes)
{
    if (!s || !bytes) return fail(OGSX_EINVAL, "null argument");
    *bytes = s->bytes;
    return OGSX_OK;
}

int ogsx_pattern_destroy(ogsx_pattern* p)
{
    if (!p) return OGSX_OK;
    const int users = p->users.load();
    if (users > 0)
        return fail(OGSX_ESTATE, "pattern still used by " + std::to_string(users) +
                                     " system(s): destroy them first");
    delete p;
    return OGSX_OK;
}

int ogsx_system_destroy(ogsx_system* s)
{
    if (!s) return OGSX_OK;
    if (s->group) return fail(OGSX_ESTATE, "system belongs to a group: destroy the group first");
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    delete s->solo;
    if (s->ready) (void)hipEventDestroy(s->ready);
    if (use_pool() && s->stream) {
        for (void* q : s->allocs) (void)hipFreeAsync(q, s->stream);
        (void)hipStreamSynchronize(s->stream);  // back in the pool before the stream goes
    } else {
        for (void* q : s->allocs) (void)hipFree(q);
    }
    if (s->prm_host) (void)hipHostFree(s->prm_host);
    if (s->irs_host) (void)hipHostFree(s->irs_host);
    for (auto e : s->ev_pool) (void)hipEventDestroy(e);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    if (s->owner) s->owner->users--;
    delete s;
    return OGSX_OK;
}

int ogsx_system_create(const ogsx_pattern* p, int precision, ogsx_system** out)
{
    if (!p || !out) return fail(OGSX_EINVAL, "null argument");
    if (precision != OGSX_FP64 && precision != OGSX_FP32_IR)
        return fail(OGSX_EINVAL, "unknown precision");
    *out = nullptr;
    if (p->h.n == 0) return fail(OGSX_EINVAL, "empty pattern (no nodes)");
    // a node in no element has an empty CRS row and no diagonal, which the
    // Jacobi preconditioner and Dirichlet elimination read
    for (int64_t i = 0; i < p->h.n; i++)
        if (p->h.diag[i] < 0)
            return fail(OGSX_EINVAL, "node " + std::to_string(i) +
                                         " belongs to no element (empty matrix row)");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return fail(OGSX_EHIP, "no HIP device available");
    auto* s = new (std::nothrow) ogsx_system();
    if (!s) return fail(OGSX_ENOMEM, "host allocation failed");
    s->pat = &p->h;
    s->owner = p;
    p->users++;
    s->precision = precision;
    s->n = (int)p->h.n;
    s->nnz = p->h.nnz;
    s->G = grid_for(s->n);
    int rc = OGSX_OK;
    auto cleanup = [&](int code) {
        ogsx_system_destroy(s);
        return code;
    };
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess)
        return cleanup(fail(OGSX_EHIP, "hipStreamCreate failed"));
    const PatternHost& h = p->h;
    const PatternLayout& Ly = p->L;
    if ((rc = dalloc(s, &s->ptr, h.n + 1)) || (rc = dalloc(s, &s->col, h.nnz)) ||
        (rc = dalloc(s, &s->diagpos, h.n)) || (rc = dalloc(s, &s->conn, h.conn.size())) ||
        (rc = dalloc(s, &s->color_elems, h.color_elems.size())) ||
        (rc = dalloc(s, &s->slot, h.conn.size() * h.npe)) || (rc = dalloc(s, &s->A, h.nnz)) ||
        (rc = dalloc(s, &s->uold, h.n)) || (rc = dalloc(s, &s->ic, h.n)) ||
        (rc = alloc_vecs(s, s->v64)) ||
        (rc = dalloc(s, &s->part, (size_t)16 * s->G)) || (rc = dalloc(s, &s->red, 8)) ||
        (rc = dalloc(s, &s->st, 1)) || (rc = dalloc(s, &s->bcflag, h.n)) ||
        (rc = dalloc(s, &s->bcg, h.n)))
        return cleanup(rc);
    if (precision == OGSX_FP32_IR) {
        if ((rc = dalloc(s, &s->A32, h.nnz)) || (rc = alloc_vecs(s, s->v32)) ||
            (rc = dalloc(s, &s->r64, h.n)))
            return cleanup(rc);
    }
    if ((rc = dalloc(s, &s->dev, 1)) || (rc = dalloc(s, &s->prm, 1))) return cleanup(rc);
    if (hipHostMalloc((void**)&s->prm_host, sizeof(KrylovParams)) != hipSuccess)
        return cleanup(fail(OGSX_ENOMEM, "hipHostMalloc failed"));
    if (precision == OGSX_FP32_IR) {
        if ((rc = dalloc(s, &s->irs, 1))) return cleanup(rc);
        if (hipHostMalloc((void**)&s->irs_host, sizeof(IRState)) != hipSuccess)
            return cleanup(fail(OGSX_ENOMEM, "hipHostMalloc failed"));
    }
    std::memset(s->prm_host, 0, sizeof(KrylovParams));
    auto up = [&](void* d, const void* src, size_t bytes) {
        return hipMemcpyAsync(d, src, bytes, hipMemcpyHostToDevice, s->stream);
    };
    if (up(s->ptr, Ly.ptr32.data(), 4 * (h.n + 1)) || up(s->col, Ly.col32.data(), 4 * h.nnz) ||
        up(s->diagpos, Ly.dg32.data(), 4 * h.n) ||
        up(s->conn, h.conn.data(), 4 * h.conn.size()) ||
        up(s->color_elems, h.color_elems.data(), 4 * h.color_elems.size()))
        return cleanup(fail(OGSX_EHIP, "upload of the pattern failed"));
    if (!h.conn.empty())
        k_slot_map<<<blocks_for((int64_t)h.conn.size(), kBlock), kBlock, 0, s->stream>>>(
            (int64_t)h.conn.size(), h.npe, s->conn, s->ptr, s->col, s->slot);
    const int nb = blocks_for(h.nnz, kBlock * 4);
    k_fill<double><<<nb, kBlock, 0, s->stream>>>(h.nnz, s->A, 0.0);
    double* zv[] = {s->v64.x, s->v64.b, s->uold, s->ic};
    for (double* q : zv) k_fill<double><<<blocks_for(h.n, kBlock * 4), kBlock, 0, s->stream>>>(h.n, q, 0.0);
    k_fill<int8_t><<<blocks_for(h.n, kBlock * 4), kBlock, 0, s->stream>>>(h.n, s->bcflag, (int8_t)0);
    if (hipStreamSynchronize(s->stream) != hipSuccess || hipGetLastError() != hipSuccess)
        return cleanup(fail(OGSX_EHIP, "initialisation kernels failed"));
    s->lay = &p->L;
    s->tile16 = Ly.tile16;
    s->n_tiles = (int)Ly.trow.size() - 1;
    s->sell16 = Ly.sell16;
    s->n_slices = (int)Ly.sw.size();
    s->npad = Ly.npad;
    s->npad3 = Ly.tri_ok ? Ly.npad3 : 0;
    {
        const int64_t nval = std::max(s->npad, s->npad3);
        if ((rc = dalloc(s, &s->sell_ptr, Ly.sptr.size())) || (rc = dalloc(s, &s->sell_w, Ly.sw.size())) ||
            (rc = dalloc(s, &s->sell_cb, Ly.scb.size())) ||
            (rc = dalloc(s, &s->sell_src, Ly.ssrc.size())) || (rc = dalloc(s, &s->sval64, nval)))
            return cleanup(rc);
        if (precision == OGSX_FP32_IR && (rc = dalloc(s, &s->sval32, nval))) return cleanup(rc);
        if (s->sell16 && (rc = dalloc(s, &s->sell_c16, Ly.sc16.size()))) return cleanup(rc);
        if (up(s->sell_ptr, Ly.sptr.data(), 4 * Ly.sptr.size()) || up(s->sell_w, Ly.sw.data(), 4 * Ly.sw.size()) ||
            up(s->sell_cb, Ly.scb.data(), 4 * Ly.scb.size()) ||
            up(s->sell_src, Ly.ssrc.data(), 4 * Ly.ssrc.size()) ||
            (s->sell16 && up(s->sell_c16, Ly.sc16.data(), 2 * Ly.sc16.size())))
            return cleanup(fail(OGSX_EHIP, "upload of the SELL layout failed"));
        if (Ly.pat_ok) {
            if ((rc = dalloc(s, &s->sell_pid, Ly.spid.size())) ||
                (rc = dalloc(s, &s->pat_delta, Ly.pdelta.size())))
                return cleanup(rc);
            if (up(s->sell_pid, Ly.spid.data(), Ly.spid.size()) ||
                up(s->pat_delta, Ly.pdelta.data(), 4 * Ly.pdelta.size()))
                return cleanup(fail(OGSX_EHIP, "upload of the stencil dictionary failed"));
            s->pat_ok = true;
            s->pat_w = Ly.pat_w;
            s->pat_cnt = Ly.pat_cnt;
        }
        if (Ly.tri_ok) {
            if ((rc = dalloc(s, &s->tri_ptr, Ly.tptr.size())) || (rc = dalloc(s, &s->tri_id, Ly.tid.size())) ||
                (rc = dalloc(s, &s->tri_cnt, Ly.tcnt.size())) || (rc = dalloc(s, &s->tri_off, Ly.toff.size())) ||
                (rc = dalloc(s, &s->tri_src, Ly.tsrc.size())))
                return cleanup(rc);
            if (up(s->tri_ptr, Ly.tptr.data(), 4 * Ly.tptr.size()) ||
                up(s->tri_id, Ly.tid.data(), 2 * Ly.tid.size()) ||
                up(s->tri_cnt, Ly.tcnt.data(), 4 * Ly.tcnt.size()) ||
                up(s->tri_off, Ly.toff.data(), 4 * Ly.toff.size()) ||
                up(s->tri_src, Ly.tsrc.data(), 4 * Ly.tsrc.size()))
                return cleanup(fail(OGSX_EHIP, "upload of the slice triples failed"));
            s->tri_ok = true;
        }
        if (Ly.dia_ok) {
            s->dia_k = (int)Ly.doff.size();
            const size_t nd = ((size_t)s->dia_k + 1) * (size_t)h.n;
            if ((rc = dalloc(s, &s->dia_off, Ly.doff.size())) ||
                (rc = dalloc(s, &s->dia_src, Ly.dsrc.size())) ||
                (rc = dalloc(s, &s->dia_val64, nd)) ||
                (precision == OGSX_FP32_IR && (rc = dalloc(s, &s->dia_val32, nd))))
                return cleanup(rc);
            if (up(s->dia_off, Ly.doff.data(), 4 * Ly.doff.size()) ||
                up(s->dia_src, Ly.dsrc.data(), 4 * Ly.dsrc.size()))
                return cleanup(fail(OGSX_EHIP, "upload of the diagonal layout failed"));
            s->dia_ok = true;
        }
    }
    // the 32-bit-column SELL and tiled layouts are tuner candidates only when
    // 16-bit slice offsets do not fit; otherwise they are built on request
    if (!s->sell16 && (rc = alloc_extra_layouts(s))) return cleanup(rc);
    if (hipStreamSynchronize(s->stream) != hipSuccess || hipGetLastError() != hipSuccess)
        return cleanup(fail(OGSX_EHIP, "upload of the SpMV layouts failed"));
    {
        SysDev& d = s->devh;
        d.m64 = mat_view(s, (const double*)s->A);
        d.m32 = mat_view(s, (const float*)s->A32);
        d.v64 = s->v64.kv();
        d.v32 = s->v32.kv();
        d.A = s->A;
        d.A32 = s->A32;
        d.sval64 = s->sval64;
        d.sval32 = s->sval32;
        d.diagpos = s->diagpos;
        d.sell_src = s->sell_src;
        d.tri_src = s->tri_src;
        d.dia_src = s->dia_src;
        d.r64 = s->r64;
        d.part = s->part;
        d.red = s->red;
        d.st = s->st;
        d.prm = s->prm;
        d.irs = s->irs;
        d.n = s->n;
        d.G = s->G;
        d.nnz = (int)s->nnz;
        d.npad = (int)s->npad;
        d.npad3 = (int)s->npad3;
        // no solve yet: the state reads "stopped" so group launches skip it
        KrylovState st0;
        std::memset(&st0, 0, sizeof(st0));
        st0.done = 4;
        if (up(s->dev, &d, sizeof(SysDev)) || up(s->st, &st0, sizeof(st0)) ||
            up(s->prm, s->prm_host, sizeof(KrylovParams)) || hipStreamSynchronize(s->stream))
            return cleanup(fail(OGSX_EHIP, "upload of the system descriptor failed"));
    }
    *out = s;
    return OGSX_OK;
}

int ogsx_system_set_elements(ogsx_system* s, const ogsx_elements* e)
{
    if (!s || !e) return fail(OGSX_EINVAL, "null argument");
    if (s->sc.active) return fail(OGSX_ESTATE, "system has a solve in progress");
    const PatternHost& h = *s->pat;
    const int ed = edim_of(h.etype);
    if (e->process != OGSX_LIQUID_FLOW && e->process != OGSX_GROUNDWATER_FLOW &&
        e->process != OGSX_HEAT_TRANSPORT)
        return fail(OGSX_EINVAL, "unknown process");
    const bool head_form = e->process == OGSX_GROUNDWATER_FLOW;
    if (e->dim != ed) return fail(OGSX_EINVAL, "problem dimension must equal the element dimension");
    if (e->cond_stride != 1 && e->cond_stride != ed * ed)
        return fail(OGSX_EINVAL, "cond_stride must be 1 or dim*dim");
    if (!e->x || (ed >= 2 && !e->y) || (ed >= 3 && !e->z) || !e->mass_coef || !e->cond)
        return fail(OGSX_EINVAL, "missing element data");
    if ((!head_form && e->mu == 0.0) || e->tuf == 0.0)
        return fail(OGSX_EINVAL, "tuf and mu must be nonzero");
    GaussTable t;
    if (build_gauss_table(h.etype, e->n_gauss, t) != OGSX_OK)
        return fail(OGSX_EINVAL, "unsupported n_gauss");
    if (!s->elements_set) {
        OK_TRY(dalloc(s, &s->ex, h.n));
        OK_TRY(dalloc(s, &s->ey, h.n));
        OK_TRY(dalloc(s, &s->ez, h.n));
        OK_TRY(dalloc(s, &s->mass_coef, h.n_elem));
        OK_TRY(dalloc(s, &s->cond, (size_t)h.n_elem * ed * ed));
        OK_TRY(dalloc(s, &s->tabN, kMaxGaussPts * 8));
        OK_TRY(dalloc(s, &s->tabdN, kMaxGaussPts * 24));
        OK_TRY(dalloc(s, &s->tabW, kMaxGaussPts * 3));
    }
    s->process = e->process;
    s->dim = e->dim;
    s->ngp = t.ngp;
    s->cond_stride = e->cond_stride;
    s->tuf = e->tuf;
    // the HEAD form is the flow operator with mat = tuf*K/1 (= K*tuf*1 of
    // the reference, exactly) and no gravity
    s->mu = head_form ? 1.0 : e->mu;
    s->geo_area = e->geo_area;
    s->adv_coef = e->adv_coef;
    s->rho_g = head_form ? 0.0 : e->rho_g;
    auto up = [&](void* d, const void* src, size_t bytes) {
        return hipMemcpyAsync(d, src, bytes, hipMemcpyHostToDevice, s->stream);
    };
    HIP_TRY(up(s->ex, e->x, 8 * h.n));
    if (ed >= 2) HIP_TRY(up(s->ey, e->y, 8 * h.n));
    if (ed >= 3) HIP_TRY(up(s->ez, e->z, 8 * h.n));
    HIP_TRY(up(s->mass_coef, e->mass_coef, 8 * h.n_elem));
    HIP_TRY(up(s->cond, e->cond, 8 * (size_t)h.n_elem * e->cond_stride));
    HIP_TRY(up(s->tabN, t.N.data(), 8 * t.N.size()));
    HIP_TRY(up(s->tabdN, t.dN.data(), 8 * t.dN.size()));
    HIP_TRY(up(s->tabW, t.w.data(), 8 * t.w.size()));
    // a symmetric conductivity/permeability tensor (and no advection) makes
    // the assembled operator symmetric
    s->tensor_sym = true;
    if (e->cond_stride == ed * ed && ed > 1)
        for (int64_t el = 0; el < h.n_elem && s->tensor_sym; el++) {
            const double* T = e->cond + (size_t)el * ed * ed;
            for (int i = 0; i < ed; i++)
                for (int j = i + 1; j < ed; j++)
                    if (T[i * ed + j] != T[j * ed + i]) s->tensor_sym = false;
        }
    s->use_vel = (e->vel != nullptr) && e->process == OGSX_HEAT_TRANSPORT;
    if (s->use_vel) {
        const size_t need = (size_t)h.n_elem * 3 * t.ngp;
        if (need > s->vel_cap) {
            OK_TRY(dalloc(s, &s->vel, need));
            s->vel_cap = need;
        }
        HIP_TRY(up(s->vel, e->vel, 8 * need));
    }
    if (use_two_pass(s)) OK_TRY(alloc_two_pass(s));
    s->elements_set = true;
    HIP_TRY(hipStreamSynchronize(s->stream));
    return OGSX_OK;
}

}  // extern "C"

namespace {

template <int ET, int PROC, bool ISO>
void launch_assemble_colours(ogsx_system* s, const AsmArgs& a)
{
    const PatternHost& h = *s->pat;
    if (a.eb) {  // two passes: every element at once, then row by row
        if (h.n_elem > 0)
            k_assemble<ET, PROC, ISO, true><<<(int)((h.n_elem + 63) / 64), 64, 0, s->stream>>>(
                a, 0, (int)h.n_elem);
        const bool grav = PROC == OGSX_LIQUID_FLOW && a.rho_g != 0.0;
        const int lanes = asm_row_lanes(h.row_max);
        const int blocks = (int)(((int64_t)s->n * lanes + kAsmRowThreads - 1) / kAsmRowThreads);
        if (blocks == 0) return;
#define ROWS_(GR, L)                                                                         \
    k_assemble_rows<ET, GR, L><<<blocks, kAsmRowThreads, 0, s->stream>>>(a, s->inc_ptr, s->asm_loc, \
                                                                        s->ptr, s->n)
        if (lanes == 16) {
            if (grav) ROWS_(true, 16); else ROWS_(false, 16);
        } else {
            if (grav) ROWS_(true, 32); else ROWS_(false, 32);
        }
#undef ROWS_
        return;
    }
    for (int c = 0; c < h.n_colors; c++) {
        const int off = h.color_ptr[c], cnt = h.color_ptr[c + 1] - off;
        if (cnt == 0) continue;
        k_assemble<ET, PROC, ISO, false><<<(cnt + 63) / 64, 64, 0, s->stream>>>(a, off, cnt);
    }
}

// The two-pass assembly's arrays (incidence lists; element matrices and
// right-hand sides), allocated and uploaded once.
int alloc_two_pass(ogsx_system* s)
{
    if (s->asm_eb) return OGSX_OK;
    const PatternHost& h = *s->pat;
    const int64_t npe = h.npe;
    const int64_t cnt = h.n_elem * npe * npe;
    OK_TRY(dalloc(s, &s->inc_ptr, h.inc_ptr.size()));
    OK_TRY(dalloc(s, &s->inc_ele, h.inc_ele.size()));
    OK_TRY(dalloc(s, &s->asm_pos, h.inc_ele.size()));
    OK_TRY(dalloc(s, &s->asm_loc, (size_t)cnt));
    OK_TRY(dalloc(s, &s->asm_eb, (size_t)(h.n_elem * npe * (npe + 2))));
    HIP_TRY(hipMemcpyAsync(s->inc_ptr, h.inc_ptr.data(), 4 * h.inc_ptr.size(), hipMemcpyHostToDevice,
                           s->stream));
    HIP_TRY(hipMemcpyAsync(s->inc_ele, h.inc_ele.data(), 4 * h.inc_ele.size(), hipMemcpyHostToDevice,
                           s->stream));
    const int64_t ni = (int64_t)h.inc_ele.size();
    if (ni > 0)
        k_row_positions<<<blocks_for(ni, 256 * 4), 256, 0, s->stream>>>(
            ni, (int)npe, s->inc_ele, s->conn, s->slot, s->ptr, s->asm_pos, s->asm_loc);
    HIP_TRY(hipGetLastError());
    return OGSX_OK;
}

// Two-pass assembly where every CRS row fits the row pass's LDS accumulator;
// the coloured scatter otherwise (and for patterns with 2^31 or more
// element-node pairs, which have no incidence list).
bool use_two_pass(const ogsx_system* s)
{
    return !s->pat->inc_ptr.empty() && s->pat->row_max <= kAsmRowMax;
}

template <int ET>
void launch_assemble_et(ogsx_system* s, const AsmArgs& a)
{
    const bool iso = s->cond_stride == 1;
    if (s->process != OGSX_HEAT_TRANSPORT) {  // LIQUID_FLOW, GROUNDWATER_FLOW
        if (iso)
            launch_assemble_colours<ET, OGSX_LIQUID_FLOW, true>(s, a);
        else
            launch_assemble_colours<ET, OGSX_LIQUID_FLOW, false>(s, a);
    } else {
        if (iso)
            launch_assemble_colours<ET, OGSX_HEAT_TRANSPORT, true>(s, a);
        else
            launch_assemble_colours<ET, OGSX_HEAT_TRANSPORT, false>(s, a);
    }
}

// Elements per block of the staged velocity kernel: a multiple of 64 up to
// 256 whose 3*ngp velocities fit 32 KiB of LDS; 0 when even 64 elements do
// not fit 64 KiB (hexahedra with 4^3 points), which then store directly.
int velocity_block(int ngp)
{
    const int per_wave = 64 * 3 * ngp;  // doubles for 64 elements
    if (per_wave * 8 > 65536) return 0;
    return 64 * std::max(1, std::min(4, 4096 / per_wave));
}

template <int ET, bool ISO>
void launch_velocity_iso(ogsx_system* s, const VelArgs& a)
{
    const int epb = velocity_block(a.ngp);
    if (epb > 0) {
        const size_t lds = (size_t)epb * 3 * a.ngp * sizeof(double);
        k_gauss_velocity<ET, ISO, true><<<(a.n_elem + epb - 1) / epb, epb, lds, s->stream>>>(a);
    } else {
        k_gauss_velocity<ET, ISO, false><<<(a.n_elem + 255) / 256, 256, 0, s->stream>>>(a);
    }
}

template <int ET>
void launch_velocity_et(ogsx_system* s, const VelArgs& a)
{
    if (s->cond_stride == 1)
        launch_velocity_iso<ET, true>(s, a);
    else
        launch_velocity_iso<ET, false>(s, a);
}

// Algorithmic HBM bytes of one k_gauss_velocity launch: every node's
// coordinates and pressure once, per element its connectivity and tensor,
// and the 3 x ngp velocities written.
double velocity_bytes(const ogsx_system* s)
{
    const PatternHost& h = *s->pat;
    const int npe = npe_of(h.etype), ed = edim_of(h.etype);
    return (double)h.n * 8.0 * (ed + 1) +
           (double)h.n_elem * (4.0 * npe + 8.0 * s->cond_stride + 24.0 * s->ngp);
}

// Makes `waiter`'s stream wait for the work queued so far on `s`'s stream.
int order_after(ogsx_system* s, ogsx_system* waiter)
{
    if (!s->ready) HIP_TRY(hipEventCreateWithFlags(&s->ready, hipEventDisableTiming));
    HIP_TRY(hipEventRecord(s->ready, s->stream));
    HIP_TRY(hipStreamWaitEvent(waiter->stream, s->ready, 0));
    return OGSX_OK;
}

// Copies host lists to device buffers that grow on demand.
template <typename T>
int upload_list(ogsx_system* s, T** dptr, int* cap, const T* src, int count)
{
    if (count > *cap) {
        int newcap = std::max(count, 2 * *cap);
        OK_TRY(dalloc(s, dptr, (size_t)newcap));
        *cap = newcap;
    }
    if (count > 0)
        HIP_TRY(hipMemcpyAsync(*dptr, src, sizeof(T) * count, hipMemcpyHostToDevice, s->stream));
    return OGSX_OK;
}

int check_vec(ogsx_system* s, int which, double** p)
{
    switch (which) {
        case OGSX_VEC_X: *p = s->v64.x; return OGSX_OK;
        case OGSX_VEC_B: *p = s->v64.b; return OGSX_OK;
        case OGSX_VEC_UOLD: *p = s->uold; return OGSX_OK;
        case OGSX_VEC_IC: *p = s->ic; return OGSX_OK;
    }
    return fail(OGSX_EINVAL, "unknown vector id");
}

}  // namespace

extern "C" {

int ogsx_system_initialize(ogsx_system* s)
{
    if (!s) return fail(OGSX_EINVAL, "null system");
    if (s->sc.active) return fail(OGSX_ESTATE, "system has a solve in progress");
    k_fill<double><<<blocks_for(s->nnz, kBlock * 4), kBlock, 0, s->stream>>>(s->nnz, s->A, 0.0);
    k_fill<double><<<blocks_for(s->n, kBlock * 4), kBlock, 0, s->stream>>>(s->n, s->v64.b, 0.0);
    s->a_version++;
    HIP_TRY(hipGetLastError());
    return OGSX_OK;
}

}  // extern "C"

namespace {
// The element pass computes gravity terms (LIQUID_FLOW with rho g != 0;
// GROUNDWATER_FLOW's head form has rho_g = 0).
bool asm_args_gravity(const ogsx_system* s)
{
    return s->process == OGSX_LIQUID_FLOW && s->rho_g != 0.0;
}

// Arguments of the assembly kernels for system s (dt, theta, steady).
AsmArgs asm_args(const ogsx_system* s, double dt, double theta, int steady)
{
    AsmArgs a;
    std::memset(&a, 0, sizeof(a));
    a.conn = s->conn;
    a.color_elems = s->color_elems;
    a.slot = s->slot;
    a.x = s->ex;
    a.y = s->ey;
    a.z = s->ez;
    a.mass_coef = s->mass_coef;
    a.cond = s->cond;
    a.vel = s->use_vel ? s->vel : nullptr;
    a.tabN = s->tabN;
    a.tabdN = s->tabdN;
    a.tabW = s->tabW;
    a.u_old = s->uold;
    a.A = s->A;
    a.b = s->v64.b;
    a.cond_stride = s->cond_stride;
    a.ngp = s->ngp;
    a.steady = steady;
    a.tuf = s->tuf;
    a.mu = s->mu;
    a.geo_area = s->geo_area;
    a.adv_coef = s->adv_coef;
    a.rho_g = s->rho_g;
    a.dt = dt;
    a.theta = theta;
    if (s->asm_eb) {  // two-pass assembly (buffers from ogsx_system_set_elements)
        a.eb = s->asm_eb;
        a.rb = s->asm_eb + s->pat->n_elem * s->pat->npe * s->pat->npe;
        a.pos = s->asm_pos;
    }
    return a;
}
}  // namespace

extern "C" {

int ogsx_assemble(ogsx_system* s, double dt, double theta, int steady)
{
    if (!s) return fail(OGSX_EINVAL, "null system");
    if (s->sc.active) return fail(OGSX_ESTATE, "system has a solve in progress");
    if (!s->elements_set) return fail(OGSX_ESTATE, "ogsx_system_set_elements not called");
    if (!(dt > 0.0)) return fail(OGSX_EINVAL, "dt must be positive");
    s->a_symmetric = s->tensor_sym && !(s->process == OGSX_HEAT_TRANSPORT && s->use_vel);
    const AsmArgs a = asm_args(s, dt, theta, steady);
    int pk = prof_begin(s, OGSX_PROF_ASSEMBLY);
    switch (s->pat->etype) {
        case OGSX_LINE2: launch_assemble_et<OGSX_LINE2>(s, a); break;
        case OGSX_QUAD4: launch_assemble_et<OGSX_QUAD4>(s, a); break;
        case OGSX_HEX8: launch_assemble_et<OGSX_HEX8>(s, a); break;
        case OGSX_TET4: launch_assemble_et<OGSX_TET4>(s, a); break;
        default: return fail(OGSX_EINVAL, "unsupported element type");
    }
    prof_end(s, pk);
    s->a_version++;
    HIP_TRY(hipGetLastError());
    return OGSX_OK;
}

int ogsx_gauss_velocity(ogsx_system* flow, ogsx_system* heat, double* host_out)
{
    if (!flow) return fail(OGSX_EINVAL, "null flow system");
    if (!heat && !host_out) return fail(OGSX_EINVAL, "heat and host_out are both NULL");
    if (heat == flow) return fail(OGSX_EINVAL, "heat and flow are the same system");
    if (flow->sc.active || (heat && heat->sc.active))
        return fail(OGSX_ESTATE, "system has a solve in progress");
    if (!flow->elements_set || flow->process == OGSX_HEAT_TRANSPORT)
        return fail(OGSX_EINVAL, "flow must be a LIQUID_FLOW or GROUNDWATER_FLOW system with "
                                 "its elements set");
    const PatternHost& h = *flow->pat;
    const size_t need = (size_t)h.n_elem * 3 * flow->ngp;
    double* dst = nullptr;
    if (heat) {
        if (!heat->elements_set || heat->process != OGSX_HEAT_TRANSPORT)
            return fail(OGSX_EINVAL, "heat must be a HEAT_TRANSPORT system with its elements set");
        if (heat->pat->n_elem != h.n_elem || heat->pat->etype != h.etype || heat->ngp != flow->ngp)
            return fail(OGSX_EINVAL, "heat is not on the flow mesh with the same n_gauss");
        if (need > heat->vel_cap) {
            OK_TRY(dalloc(heat, &heat->vel, need));
            heat->vel_cap = need;
        }
        dst = heat->vel;
        // heat's queued work (an assembly reading the old velocity, the
        // allocation) comes first
        OK_TRY(order_after(heat, flow));
    } else {
        if (need > flow->gvel_cap) {
            OK_TRY(dalloc(flow, &flow->gvel, need));
            flow->gvel_cap = need;
        }
        dst = flow->gvel;
    }
    if (h.n_elem > 0) {
        VelArgs a;
        std::memset(&a, 0, sizeof(a));
        a.conn = flow->conn;
        a.x = flow->ex;
        a.y = flow->ey;
        a.z = flow->ez;
        a.cond = flow->cond;
        a.tabdN = flow->tabdN;
        a.u = flow->v64.x;
        a.vel = dst;
        a.n_elem = (int)h.n_elem;
        a.cond_stride = flow->cond_stride;
        a.ngp = flow->ngp;
        a.tuf = flow->tuf;
        a.mu = flow->mu;
        a.rho_g = flow->rho_g;
        const int pk = prof_begin(flow, OGSX_PROF_VELOCITY);
        switch (h.etype) {
            case OGSX_LINE2: launch_velocity_et<OGSX_LINE2>(flow, a); break;
            case OGSX_QUAD4: launch_velocity_et<OGSX_QUAD4>(flow, a); break;
            case OGSX_HEX8: launch_velocity_et<OGSX_HEX8>(flow, a); break;
            case OGSX_TET4: launch_velocity_et<OGSX_TET4>(flow, a); break;
            default: return fail(OGSX_EINVAL, "unsupported element type");
        }
        prof_end(flow, pk);
        HIP_TRY(hipGetLastError());
    }
    if (heat) {
        OK_TRY(order_after(flow, heat));
        heat->use_vel = true;
    }
    if (host_out) {
        if (need > 0)
            HIP_TRY(hipMemcpyAsync(host_out, dst, 8 * need, hipMemcpyDeviceToHost, flow->stream));
        HIP_TRY(hipStreamSynchronize(flow->stream));
    }
    return OGSX_OK;
}

int ogsx_add_rhs(ogsx_system* s, int64_t n, const int64_t* idx, const double* val)
{
    if (!s) return fail(OGSX_EINVAL, "null system");
    if (s->sc.active) return fail(OGSX_ESTATE, "system has a solve in progress");
    if (n < 0 || (n > 0 && (!idx || !val))) return fail(OGSX_EINVAL, "bad source-term list");
    if (n == 0) return OGSX_OK;
    for (int64_t k = 0; k < n; k++)
        if (idx[k] < 0 || idx[k] >= s->n) return fail(OGSX_EINVAL, "source-term node out of range");
    const bool same = (int64_t)s->last_st_idx.size() == n &&
                      std::equal(idx, idx + n, s->last_st_idx.begin()) &&
                      std::memcmp(val, s->last_st_val.data(), 8 * n) == 0;
    if (!same) {
        // group by node, keeping list order inside a node
        std::vector<int64_t> ord(n);
        for (int64_t k = 0; k < n; k++) ord[k] = k;
        std::stable_sort(ord.begin(), ord.end(), [&](int64_t a, int64_t b) { return idx[a] < idx[b]; });
        std::vector<int> node, vptr(1, 0);
        std::vector<double> v;
        for (int64_t k = 0; k < n; k++) {
            const int64_t o = ord[k];
            if (node.empty() || node.back() != idx[o]) {
                if (!node.empty()) vptr.push_back((int)v.size());
                node.push_back((int)idx[o]);
            }
            v.push_back(val[o]);
        }
        vptr.push_back((int)v.size());
        int want = (int)std::max<size_t>(v.size(), vptr.size());
        if (want > s->st_cap) {
            OK_TRY(dalloc(s, &s->st_node_d, want));
            OK_TRY(dalloc(s, &s->st_vptr_d, want + 1));
            OK_TRY(dalloc(s, &s->st_val_d, want));
            s->st_cap = want;
        }
        HIP_TRY(hipMemcpyAsync(s->st_node_d, node.data(), 4 * node.size(), hipMemcpyHostToDevice, s->stream));
        HIP_TRY(hipMemcpyAsync(s->st_vptr_d, vptr.data(), 4 * vptr.size(), hipMemcpyHostToDevice, s->stream));
        HIP_TRY(hipMemcpyAsync(s->st_val_d, v.data(), 8 * v.size(), hipMemcpyHostToDevice, s->stream));
        HIP_TRY(hipStreamSynchronize(s->stream));
        s->st_nu = (int)node.size();
        s->last_st_idx.assign(idx, idx + n);
        s->last_st_val.assign(val, val + n);
    }
    k_add_rhs<<<blocks_for(s->st_nu, kBlock), kBlock, 0, s->stream>>>(s->st_nu, s->st_node_d,
                                                                      s->st_vptr_d, s->st_val_d,
                                                                      s->v64.b);
    HIP_TRY(hipGetLastError());
    return OGSX_OK;
}

int ogsx_set_dirichlet(ogsx_system* s, int64_t n, const int64_t* idx, const double* val, int mode)
{
    if (!s) return fail(OGSX_EINVAL, "null system");
    if (s->sc.active) return fail(OGSX_ESTATE, "system has a solve in progress");
    if (mode != OGSX_DIRICHLET_ROW && mode != OGSX_DIRICHLET_SYMMETRIC)
        return fail(OGSX_EINVAL, "unknown Dirichlet mode");
    if (n < 0 || (n > 0 && (!idx || !val))) return fail(OGSX_EINVAL, "bad Dirichlet list");
    if (n == 0) return OGSX_OK;
    for (int64_t k = 0; k < n; k++)
        if (idx[k] < 0 || idx[k] >= s->n) return fail(OGSX_EINVAL, "Dirichlet node out of range");
    const bool same = (int64_t)s->last_bc_idx.size() == n &&
                      std::equal(idx, idx + n, s->last_bc_idx.begin()) &&
                      std::memcmp(val, s->last_bc_val.data(), 8 * n) == 0;
    if (!same) {
        // SetKnownX_i applied in list order: a node listed twice keeps the last value
        std::vector<int64_t> last(s->n, -1);
        for (int64_t k = 0; k < n; k++) last[idx[k]] = k;
        std::vector<int> di;
        std::vector<double> dv;
        for (int64_t k = 0; k < n; k++)
            if (last[idx[k]] == k) {
                di.push_back((int)idx[k]);
                dv.push_back(val[k]);
            }
        std::vector<char> touched(s->n, 0);
        for (int d : di)
            for (int64_t k = s->pat->ptr[d]; k < s->pat->ptr[d + 1]; k++) touched[s->pat->col[k]] = 1;
        for (int d : di) touched[d] = 1;
        std::vector<int> rows;
        for (int64_t i = 0; i < s->n; i++)
            if (touched[i]) rows.push_back((int)i);
        OK_TRY(upload_list(s, &s->bc_idx_d, &s->bc_cap, di.data(), (int)di.size()));
        OK_TRY(upload_list(s, &s->bc_val_d, &s->bc_val_cap, dv.data(), (int)dv.size()));
        OK_TRY(upload_list(s, &s->bc_rows_d, &s->bc_rows_cap, rows.data(), (int)rows.size()));
        HIP_TRY(hipStreamSynchronize(s->stream));
        s->bc_m = (int)di.size();
        s->bc_rows_m = (int)rows.size();
        s->last_bc_idx.assign(idx, idx + n);
        s->last_bc_val.assign(val, val + n);
    }
    const int m = s->bc_m;
    if (mode == OGSX_DIRICHLET_ROW) s->a_symmetric = false;
    if (mode == OGSX_DIRICHLET_ROW) {
        k_bc_row<<<blocks_for(m, kBlock), kBlock, 0, s->stream>>>(m, s->bc_idx_d, s->bc_val_d, s->ptr,
                                                                  s->col, s->diagpos, s->A, s->v64.b);
    } else {
        k_bc_mark<<<blocks_for(m, kBlock), kBlock, 0, s->stream>>>(m, s->bc_idx_d, s->bc_val_d,
                                                                   s->bcflag, s->bcg, (int8_t)1);
        // only the rows the elimination can change (2.1 M -> 0.1 M rows with
        // two Dirichlet faces of the 128^3 hex mesh)
        k_bc_sym<<<blocks_for(s->bc_rows_m, kBlock / 32), kBlock, 0, s->stream>>>(
            s->bc_rows_m, s->bc_rows_d, s->ptr, s->col, s->diagpos, s->bcflag, s->bcg, s->A,
            s->v64.b);
        k_bc_mark<<<blocks_for(m, kBlock), kBlock, 0, s->stream>>>(m, s->bc_idx_d, s->bc_val_d,
                                                                   s->bcflag, s->bcg, (int8_t)0);
    }
    s->a_version++;
    HIP_TRY(hipGetLastError());
    return OGSX_OK;
}

int ogsx_solve(ogsx_system* s, const ogsx_solver_config* cfg, ogsx_solve_info* info)
{
    if (!s || !cfg) return fail(OGSX_EINVAL, "null argument");
    OK_TRY(check_config(cfg));
    if (s->group) return fail(OGSX_ESTATE, "system belongs to a group: solve it through the group");
    if (!s->solo) OK_TRY(make_set({s}, s->stream, &s->solo));
    MemberSet* m = s->solo;
    OK_TRY(prepare_member(s, cfg, s->stream));
    OK_TRY(set_start(m, {0}));
    OK_TRY(set_advance(m, INT_MAX));
    HIP_TRY(hipGetLastError());
    if (info) info_of(s, info);
    return OGSX_OK;
}

}  // extern "C"

struct ogsx_group {
    int precision = OGSX_FP64;
    hipStream_t stream = nullptr;
    MemberSet* set = nullptr;
    // batched step preparation: the members' StepArgs (pinned staging and
    // device copy), the event after its upload, and the event after the
    // preparation that the members' streams wait on
    StepArgs* tab_host = nullptr;
    StepArgs* tab_dev = nullptr;
    size_t tab_cap = 0;
    hipEvent_t tab_copied = nullptr;
    hipEvent_t prep_done = nullptr;
    ~ogsx_group()
    {
        if (tab_copied) (void)hipEventSynchronize(tab_copied);
        if (tab_host) (void)hipHostFree(tab_host);
        if (tab_dev) (void)hipFree(tab_dev);
        if (tab_copied) (void)hipEventDestroy(tab_copied);
        if (prep_done) (void)hipEventDestroy(prep_done);
    }
};

namespace {
int member_index(const ogsx_group* g, const ogsx_system* s)
{
    const auto& mem = g->set->mem;
    for (size_t i = 0; i < mem.size(); i++)
        if (mem[i] == s) return (int)i;
    return -1;
}

// A starting system may take a format slower (its own tuner timings) than
// its fastest by this fraction, or by this many ms per launch — about what
// one more launch per Krylov phase costs the group — whichever is larger,
// to share a format with the others.
constexpr double kGroupFmtSlack = 0.10;
constexpr double kGroupFmtSlackMs = 0.004;

// Formats of systems that start together in a group: every distinct format
// among the running systems costs one more SpMV launch per Krylov phase
// (measured on 256 2-D instances: one common format runs 11-14 % faster
// than each instance's own fastest).  A starting system takes a format the
// running systems already use, or the one the most starters can share,
// among the formats its tuning timed within the slack of its fastest.
// Formats set by ogsx_set_spmv_variant stay.  Results do not depend on the
// format (all candidates sum rows in CRS order).
void share_formats(MemberSet* m, const std::vector<int>& which, bool ir)
{
    for (int slot = 0; slot < (ir ? 2 : 1); slot++) {
        auto fmt = [&](ogsx_system* s) -> int& { return slot == 0 ? s->fmt64 : s->fmt32; };
        std::vector<int> inuse;
        for (ogsx_system* s : m->mem)
            if (s->sc.active) inuse.push_back(fmt(s));
        struct Starter {
            ogsx_system* s;
            std::vector<std::pair<double, int>> ok;  // (ms, format), fastest first
        };
        std::vector<Starter> todo;
        for (int i : which) {
            ogsx_system* s = m->mem[i];
            const size_t b = slot == 0 ? 0 : s->tune_n64;
            const size_t e = slot == 0 ? s->tune_n64 : s->tune_log.size();
            if (s->fmt_forced || b >= e) {
                inuse.push_back(fmt(s));
                continue;
            }
            double best = 1e300;
            for (size_t k = b; k < e; k++) best = std::min(best, s->tune_log[k].second);
            Starter st{s, {}};
            for (size_t k = b; k < e; k++)
                if (s->tune_log[k].second - best <= std::max(kGroupFmtSlack * best, kGroupFmtSlackMs))
                    st.ok.push_back({s->tune_log[k].second, s->tune_log[k].first});
            std::sort(st.ok.begin(), st.ok.end());
            todo.push_back(st);
        }
        auto in_use = [&](int f) { return std::find(inuse.begin(), inuse.end(), f) != inuse.end(); };
        std::vector<char> done(todo.size(), 0);
        size_t left = todo.size();
        while (left > 0) {
            // a format already in use, if acceptable
            for (size_t t = 0; t < todo.size(); t++) {
                if (done[t]) continue;
                for (const auto& c : todo[t].ok)
                    if (in_use(c.second)) {
                        fmt(todo[t].s) = c.second;
                        done[t] = 1;
                        left--;
                        break;
                    }
            }
            if (left == 0) break;
            // else the format acceptable to the most remaining starters; ties
            // go to the format more of them hold already (no re-layout of
            // their values, and a system starting alone keeps its format
            // from step to step), then to the smaller summed time
            struct Vote {
                int count = 0, current = 0;
                double ms = 0.0;
            };
            std::map<int, Vote> votes;
            for (size_t t = 0; t < todo.size(); t++)
                if (!done[t])
                    for (const auto& c : todo[t].ok) {
                        Vote& v = votes[c.second];
                        v.count++;
                        v.current += fmt(todo[t].s) == c.second;
                        v.ms += c.first;
                    }
            int pick = 0;
            Vote best;
            best.count = -1;
            for (const auto& kv : votes) {
                const Vote& v = kv.second;
                if (v.count > best.count ||
                    (v.count == best.count &&
                     (v.current > best.current || (v.current == best.current && v.ms < best.ms)))) {
                    pick = kv.first;
                    best = v;
                }
            }
            if (best.count <= 0) break;
            inuse.push_back(pick);
        }
    }
}
}  // namespace

extern "C" {

int ogsx_group_create(int precision, int use_graphs, ogsx_group** out)
{
    if (!out) return fail(OGSX_EINVAL, "null output handle");
    *out = nullptr;
    if (precision != OGSX_FP64 && precision != OGSX_FP32_IR) return fail(OGSX_EINVAL, "unknown precision");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(OGSX_EHIP, "no HIP device available");
    auto* g = new (std::nothrow) ogsx_group();
    if (!g) return fail(OGSX_ENOMEM, "host allocation failed");
    g->precision = precision;
    if (hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking) != hipSuccess) {
        delete g;
        return fail(OGSX_EHIP, "hipStreamCreate failed");
    }
    const int rc = make_set({}, g->stream, &g->set);
    if (rc != OGSX_OK) {
        (void)hipStreamDestroy(g->stream);
        delete g;
        return rc;
    }
    g->set->use_graphs = use_graphs != 0;
    *out = g;
    return OGSX_OK;
}

int ogsx_group_destroy(ogsx_group* g)
{
    if (!g) return OGSX_OK;
    (void)hipStreamSynchronize(g->stream);
    for (ogsx_system* s : g->set->mem) {
        s->group = nullptr;
        s->sc.active = 0;
    }
    delete g->set;
    (void)hipStreamDestroy(g->stream);
    delete g;
    return OGSX_OK;
}

int ogsx_group_add(ogsx_group* g, ogsx_system* s)
{
    if (!g || !s) return fail(OGSX_EINVAL, "null argument");
    if (s->group) return fail(OGSX_ESTATE, "system already belongs to a group");
    if (s->precision != g->precision) return fail(OGSX_EINVAL, "system precision differs from the group's");
    std::vector<ogsx_system*> mem = g->set->mem;
    mem.push_back(s);
    OK_TRY(set_members(g->set, mem));
    s->group = g;
    s->sc.active = 0;
    return OGSX_OK;
}

int ogsx_group_remove(ogsx_group* g, ogsx_system* s)
{
    if (!g || !s) return fail(OGSX_EINVAL, "null argument");
    const int i = member_index(g, s);
    if (i < 0) return fail(OGSX_EINVAL, "system is not a member of this group");
    if (s->sc.active) return fail(OGSX_ESTATE, "system has a solve in progress");
    HIP_TRY(hipStreamSynchronize(g->stream));
    std::vector<ogsx_system*> mem = g->set->mem;
    mem.erase(mem.begin() + i);
    OK_TRY(set_members(g->set, mem));
    s->group = nullptr;
    return OGSX_OK;
}

int ogsx_group_start(ogsx_group* g, ogsx_system* const* systems, int count,
                     const ogsx_solver_config* cfg)
{
    if (!g || !cfg || count < 0 || (count > 0 && !systems)) return fail(OGSX_EINVAL, "bad argument");
    OK_TRY(check_config(cfg));
    MemberSet* m = g->set;
    int kmethod = -1;
    for (ogsx_system* s : m->mem)
        if (s->sc.active) kmethod = s->sc.kmethod;
    std::vector<int> which;
    for (int k = 0; k < count; k++) {
        const int i = member_index(g, systems[k]);
        if (i < 0) return fail(OGSX_EINVAL, "system is not a member of this group");
        if (systems[k]->sc.active) return fail(OGSX_ESTATE, "system has a solve in progress");
        if (std::find(which.begin(), which.end(), i) != which.end())
            return fail(OGSX_EINVAL, "system listed twice");
        const int km = kernel_method(cfg, systems[k]);
        if (kmethod >= 0 && km != kmethod)
            return fail(OGSX_EINVAL, "solves running together in a group must share the method "
                                     "(and the form of the FP32 recurrence)");
        kmethod = km;
        which.push_back(i);
    }
    for (int i : which) OK_TRY(autotune(m->mem[i]));
    share_formats(m, which, g->precision == OGSX_FP32_IR);
    for (int i : which) {
        ogsx_system* s = m->mem[i];
        OK_TRY(prepare_member(s, cfg, s->stream));
        if (!s->ready) HIP_TRY(hipEventCreateWithFlags(&s->ready, hipEventDisableTiming));
        HIP_TRY(hipEventRecord(s->ready, s->stream));
        HIP_TRY(hipStreamWaitEvent(g->stream, s->ready, 0));
    }
    return set_start(m, which);
}

}  // extern "C"

namespace {
// Block maps of one phase of the batched step preparation: per listed
// member (compact index) its block count; returns the total block count.
int upload_step_map(MemberSet* m, const std::vector<int64_t>& blocks, const int** blk2mem,
                    const int** blk0, int* total)
{
    std::vector<int> host;
    int64_t t = 0;
    for (int64_t b : blocks) t += b;
    if (t >= INT_MAX) return fail(OGSX_EINVAL, "batched step: too many blocks");
    host.reserve((size_t)t + blocks.size() + 1);
    std::vector<int> b0(blocks.size() + 1, 0);
    for (size_t i = 0; i < blocks.size(); i++) {
        b0[i] = (int)host.size();
        for (int64_t k = 0; k < blocks[i]; k++) host.push_back((int)i);
    }
    b0[blocks.size()] = (int)host.size();
    const size_t nb = host.size();
    host.insert(host.end(), b0.begin(), b0.end());
    OK_TRY(arena_reserve(m, m->tmp_arena, host.size()));
    int* dev = nullptr;
    OK_TRY(upload_ints(m, m->tmp_arena, host.data(), host.size(), &dev));
    *blk2mem = dev;
    *blk0 = dev + nb;
    *total = (int)nb;
    return OGSX_OK;
}

int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

template <int ET, int PROC, bool ISO>
void launch_step_assemble(hipStream_t st, const StepMap& sm, int blocks)
{
    k_step_assemble<ET, PROC, ISO><<<blocks, 64, 0, st>>>(sm);
}

template <int ET>
void launch_step_rows(hipStream_t st, const StepMap& sm, int blocks, bool grav, int lanes)
{
    if (lanes == 16) {
        if (grav)
            k_step_rows<ET, true, 16><<<blocks, kAsmRowThreads, 0, st>>>(sm);
        else
            k_step_rows<ET, false, 16><<<blocks, kAsmRowThreads, 0, st>>>(sm);
    } else {
        if (grav)
            k_step_rows<ET, true, 32><<<blocks, kAsmRowThreads, 0, st>>>(sm);
        else
            k_step_rows<ET, false, 32><<<blocks, kAsmRowThreads, 0, st>>>(sm);
    }
}

template <int ET>
void launch_step_assemble_et(hipStream_t st, const StepMap& sm, int blocks, bool heat, bool iso)
{
    if (heat) {
        if (iso)
            launch_step_assemble<ET, OGSX_HEAT_TRANSPORT, true>(st, sm, blocks);
        else
            launch_step_assemble<ET, OGSX_HEAT_TRANSPORT, false>(st, sm, blocks);
    } else {
        if (iso)
            launch_step_assemble<ET, OGSX_LIQUID_FLOW, true>(st, sm, blocks);
        else
            launch_step_assemble<ET, OGSX_LIQUID_FLOW, false>(st, sm, blocks);
    }
}
}  // namespace

extern "C" {

int ogsx_group_step(ogsx_group* g, ogsx_system* const* systems, int count,
                    const ogsx_step_params* params, const ogsx_solver_config* cfg)
{
    if (!g || !cfg || count < 0 || (count > 0 && (!systems || !params)))
        return fail(OGSX_EINVAL, "bad argument");
    if (count == 0) return OGSX_OK;
    OK_TRY(check_config(cfg));
    MemberSet* m = g->set;
    for (int k = 0; k < count; k++) {
        ogsx_system* s = systems[k];
        const ogsx_step_params& p = params[k];
        if (member_index(g, s) < 0) return fail(OGSX_EINVAL, "system is not a member of this group");
        if (s->sc.active) return fail(OGSX_ESTATE, "system has a solve in progress");
        if (!s->elements_set) return fail(OGSX_ESTATE, "ogsx_system_set_elements not called");
        if (!(p.dt > 0.0)) return fail(OGSX_EINVAL, "dt must be positive");
        if (p.dirichlet != -1 && p.dirichlet != OGSX_DIRICHLET_ROW &&
            p.dirichlet != OGSX_DIRICHLET_SYMMETRIC)
            return fail(OGSX_EINVAL, "unknown Dirichlet mode");
        if (!s->asm_eb || !use_two_pass(s))
            return fail(OGSX_EINVAL, "batched steps need the two-pass assembly (rows of at most 32 "
                                     "entries)");
        for (int j = 0; j < k; j++)
            if (systems[j] == s) return fail(OGSX_EINVAL, "system listed twice");
    }
    // the members' queued work (uploads, the last copy-back) comes first
    for (int k = 0; k < count; k++) {
        ogsx_system* s = systems[k];
        if (!s->ready) HIP_TRY(hipEventCreateWithFlags(&s->ready, hipEventDisableTiming));
        HIP_TRY(hipEventRecord(s->ready, s->stream));
        HIP_TRY(hipStreamWaitEvent(g->stream, s->ready, 0));
    }
    // the members' arguments, through pinned staging (the previous call's
    // copy must have left it)
    if ((size_t)count > g->tab_cap) {
        if (g->tab_copied) HIP_TRY(hipEventSynchronize(g->tab_copied));
        if (g->tab_host) (void)hipHostFree(g->tab_host);
        if (g->tab_dev) (void)hipFree(g->tab_dev);
        g->tab_host = nullptr;
        g->tab_dev = nullptr;
        g->tab_cap = 0;
        const size_t cap = std::max<size_t>(count, 64);
        if (hipHostMalloc((void**)&g->tab_host, cap * sizeof(StepArgs)) != hipSuccess ||
            hipMalloc((void**)&g->tab_dev, cap * sizeof(StepArgs)) != hipSuccess)
            return fail(OGSX_ENOMEM, "allocation of the step table failed");
        g->tab_cap = cap;
    }
    if (!g->tab_copied) HIP_TRY(hipEventCreateWithFlags(&g->tab_copied, hipEventDisableTiming));
    if (!g->prep_done) HIP_TRY(hipEventCreateWithFlags(&g->prep_done, hipEventDisableTiming));
    HIP_TRY(hipEventSynchronize(g->tab_copied));
    for (int k = 0; k < count; k++) {
        ogsx_system* s = systems[k];
        const ogsx_step_params& p = params[k];
        StepArgs& A = g->tab_host[k];
        std::memset(&A, 0, sizeof(A));
        A.a = asm_args(s, p.dt, p.theta, p.steady);
        A.inc_ptr = s->inc_ptr;
        A.loc = s->asm_loc;
        A.ptr = s->ptr;
        A.col = s->col;
        A.diagpos = s->diagpos;
        A.st_node = s->st_node_d;
        A.st_vptr = s->st_vptr_d;
        A.st_val = s->st_val_d;
        A.bc_idx = s->bc_idx_d;
        A.bc_val = s->bc_val_d;
        A.bc_rows = s->bc_rows_d;
        A.bcflag = s->bcflag;
        A.bcg = s->bcg;
        A.uold = s->uold;
        A.x = s->v64.x;
        A.n = s->n;
        A.nnz = (int)s->nnz;
        A.ne = (int)s->pat->n_elem;
        A.st_nu = p.sources ? s->st_nu : 0;
        A.bc_m = p.dirichlet != -1 ? s->bc_m : 0;
        A.bc_rows_m = p.dirichlet != -1 ? s->bc_rows_m : 0;
        A.shift = p.shift != 0;
    }
    HIP_TRY(hipMemcpyAsync(g->tab_dev, g->tab_host, sizeof(StepArgs) * count,
                           hipMemcpyHostToDevice, g->stream));
    HIP_TRY(hipEventRecord(g->tab_copied, g->stream));
    hipStream_t st = g->stream;
    auto phase = [&](auto blocks_of, auto launch) -> int {
        std::vector<int64_t> nb(count);
        int64_t any = 0;
        for (int k = 0; k < count; k++) any += (nb[k] = blocks_of(k));
        if (any == 0) return OGSX_OK;
        StepMap sm{g->tab_dev, nullptr, nullptr};
        int total = 0;
        OK_TRY(upload_step_map(m, nb, &sm.blk2mem, &sm.blk0, &total));
        launch(sm, total);
        return OGSX_OK;
    };
    // u_old <- x, A = 0, b = 0 (ogsx_copy_vector, Linear_EQS::Initialize)
    OK_TRY(phase([&](int k) { return ceil_div(std::max<int64_t>(systems[k]->nnz, systems[k]->n), kStepFill); },
                 [&](const StepMap& sm, int total) { k_step_init<<<total, kBlock, 0, st>>>(sm); }));
    // the element pass, one launch per (element type, process, tensor) class
    for (int et : {OGSX_LINE2, OGSX_QUAD4, OGSX_HEX8, OGSX_TET4})
        for (int heat = 0; heat < 2; heat++)
            for (int iso = 0; iso < 2; iso++) {
                auto in = [&](int k) {
                    const ogsx_system* s = systems[k];
                    return s->pat->etype == et && (s->process == OGSX_HEAT_TRANSPORT) == (heat != 0) &&
                           (s->cond_stride == 1) == (iso != 0);
                };
                OK_TRY(phase([&](int k) { return in(k) ? ceil_div(systems[k]->pat->n_elem, 64) : 0; },
                             [&](const StepMap& sm, int total) {
                                 switch (et) {
                                     case OGSX_LINE2: launch_step_assemble_et<OGSX_LINE2>(st, sm, total, heat, iso); break;
                                     case OGSX_QUAD4: launch_step_assemble_et<OGSX_QUAD4>(st, sm, total, heat, iso); break;
                                     case OGSX_HEX8: launch_step_assemble_et<OGSX_HEX8>(st, sm, total, heat, iso); break;
                                     default: launch_step_assemble_et<OGSX_TET4>(st, sm, total, heat, iso); break;
                                 }
                             }));
            }
    // the row pass, per (element type, gravity, lanes per row)
    for (int et : {OGSX_LINE2, OGSX_QUAD4, OGSX_HEX8, OGSX_TET4})
        for (int grav = 0; grav < 2; grav++)
            for (int lanes : {16, 32}) {
                auto in = [&](int k) {
                    const ogsx_system* s = systems[k];
                    const bool gr = asm_args_gravity(s);
                    return s->pat->etype == et && gr == (grav != 0) &&
                           asm_row_lanes(s->pat->row_max) == lanes;
                };
                OK_TRY(phase(
                    [&](int k) { return in(k) ? ceil_div((int64_t)systems[k]->n * lanes, kAsmRowThreads) : 0; },
                    [&](const StepMap& sm, int total) {
                        switch (et) {
                            case OGSX_LINE2: launch_step_rows<OGSX_LINE2>(st, sm, total, grav, lanes); break;
                            case OGSX_QUAD4: launch_step_rows<OGSX_QUAD4>(st, sm, total, grav, lanes); break;
                            case OGSX_HEX8: launch_step_rows<OGSX_HEX8>(st, sm, total, grav, lanes); break;
                            default: launch_step_rows<OGSX_TET4>(st, sm, total, grav, lanes); break;
                        }
                    }));
            }
    // IncorporateSourceTerms with the stored lists
    OK_TRY(phase([&](int k) { return params[k].sources ? ceil_div(systems[k]->st_nu, kBlock) : 0; },
                 [&](const StepMap& sm, int total) { k_step_sources<<<total, kBlock, 0, st>>>(sm); }));
    // Dirichlet with the stored lists: row-only members, then symmetric ones
    auto bc_blocks = [&](int mode) {
        return [&, mode](int k) { return params[k].dirichlet == mode ? ceil_div(systems[k]->bc_m, kBlock) : 0; };
    };
    OK_TRY(phase(bc_blocks(OGSX_DIRICHLET_ROW),
                 [&](const StepMap& sm, int total) { k_step_bc_row<<<total, kBlock, 0, st>>>(sm); }));
    OK_TRY(phase(bc_blocks(OGSX_DIRICHLET_SYMMETRIC), [&](const StepMap& sm, int total) {
        k_step_bc_mark<<<total, kBlock, 0, st>>>(sm, (int8_t)1);
    }));
    OK_TRY(phase(
        [&](int k) {
            return params[k].dirichlet == OGSX_DIRICHLET_SYMMETRIC && systems[k]->bc_m > 0
                       ? std::min<int64_t>(ceil_div(systems[k]->bc_rows_m, kBlock / 32), 2048)
                       : 0;
        },
        [&](const StepMap& sm, int total) { k_step_bc_sym<<<total, kBlock, 0, st>>>(sm); }));
    OK_TRY(phase(bc_blocks(OGSX_DIRICHLET_SYMMETRIC), [&](const StepMap& sm, int total) {
        k_step_bc_mark<<<total, kBlock, 0, st>>>(sm, (int8_t)0);
    }));
    HIP_TRY(hipGetLastError());
    // the members' own work (format tuning, value copies, the solve start)
    // follows the preparation
    HIP_TRY(hipEventRecord(g->prep_done, g->stream));
    for (int k = 0; k < count; k++) {
        ogsx_system* s = systems[k];
        HIP_TRY(hipStreamWaitEvent(s->stream, g->prep_done, 0));
        s->a_version++;
        s->a_symmetric = s->tensor_sym && !(s->process == OGSX_HEAT_TRANSPORT && s->use_vel) &&
                         !(params[k].dirichlet == OGSX_DIRICHLET_ROW && s->bc_m > 0);
    }
    return ogsx_group_start(g, systems, count, cfg);
}

int ogsx_group_advance(ogsx_group* g, int max_iterations, int* n_active)
{
    if (!g || max_iterations < 0) return fail(OGSX_EINVAL, "bad argument");
    OK_TRY(set_advance(g->set, max_iterations, n_active));
    HIP_TRY(hipGetLastError());
    return OGSX_OK;
}

int ogsx_solve_status(ogsx_system* s, int* active, ogsx_solve_info* info)
{
    if (!s) return fail(OGSX_EINVAL, "null system");
    if (active) *active = s->sc.active;
    if (info) info_of(s, info);
    return OGSX_OK;
}

int ogsx_group_solve(ogsx_group* g, const ogsx_solver_config* cfg, ogsx_solve_info* infos)
{
    if (!g || !cfg) return fail(OGSX_EINVAL, "null argument");
    const std::vector<ogsx_system*> mem = g->set->mem;
    OK_TRY(ogsx_group_start(g, mem.data(), (int)mem.size(), cfg));
    int active = 1;
    while (active) OK_TRY(set_advance(g->set, INT_MAX, &active));
    if (infos)
        for (size_t i = 0; i < mem.size(); i++) info_of(mem[i], &infos[i]);
    return OGSX_OK;
}

int ogsx_set_vector(ogsx_system* s, int which, const double* host)
{
    if (!s || !host) return fail(OGSX_EINVAL, "null argument");
    if (s->sc.active) return fail(OGSX_ESTATE, "system has a solve in progress");
    double* d = nullptr;
    OK_TRY(check_vec(s, which, &d));
    HIP_TRY(hipMemcpyAsync(d, host, 8 * (size_t)s->n, hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return OGSX_OK;
}

int ogsx_get_vector(ogsx_system* s, int which, double* host)
{
    if (!s || !host) return fail(OGSX_EINVAL, "null argument");
    double* d = nullptr;
    OK_TRY(check_vec(s, which, &d));
    HIP_TRY(hipMemcpyAsync(host, d, 8 * (size_t)s->n, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return OGSX_OK;
}

int ogsx_copy_vector(ogsx_system* s, int dst, int src)
{
    if (!s) return fail(OGSX_EINVAL, "null system");
    if (s->sc.active) return fail(OGSX_ESTATE, "system has a solve in progress");
    double *d = nullptr, *q = nullptr;
    OK_TRY(check_vec(s, dst, &d));
    OK_TRY(check_vec(s, src, &q));
    if (d != q) k_copy<double><<<blocks_for(s->n, kBlock * 4), kBlock, 0, s->stream>>>(s->n, q, d);
    HIP_TRY(hipGetLastError());
    return OGSX_OK;
}

int ogsx_get_matrix(ogsx_system* s, double* values)
{
    if (!s || !values) return fail(OGSX_EINVAL, "null argument");
    HIP_TRY(hipMemcpyAsync(values, s->A, 8 * (size_t)s->nnz, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return OGSX_OK;
}

int ogsx_spmv(ogsx_system* s, const double* x, double* y)
{
    if (!s || !x || !y) return fail(OGSX_EINVAL, "null argument");
    if (s->sc.active) return fail(OGSX_ESTATE, "system has a solve in progress");
    // the Krylov work vectors r and s serve as scratch
    HIP_TRY(hipMemcpyAsync(s->v64.r, x, 8 * (size_t)s->n, hipMemcpyHostToDevice, s->stream));
    int pk = prof_begin(s, OGSX_PROF_SPMV64);
    OK_TRY(autotune(s));
    OK_TRY(sync_sell(s, is_sell(s->fmt64), false));
    OK_TRY(sync_dia64(s, s->stream));
    FMT_DISPATCH(s->fmt64, (k_spmv<L, double, double><<<s->G, kBlock, 0, s->stream>>>(
                               mat_view(s, (const double*)s->A), s->v64.r, s->v64.s)));
    prof_end(s, pk);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(y, s->v64.s, 8 * (size_t)s->n, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    OK_TRY(prof_flush(s));
    return OGSX_OK;
}

int ogsx_write_bin(ogsx_system* s, const char* path)
{
    if (!s || !path) return fail(OGSX_EINVAL, "null argument");
    const PatternHost& h = *s->pat;
    std::vector<double> a(h.nnz), b(h.n);
    OK_TRY(ogsx_get_matrix(s, a.data()));
    OK_TRY(ogsx_get_vector(s, OGSX_VEC_B, b.data()));
    FILE* f = std::fopen(path, "wb");
    if (!f) return fail(OGSX_EINVAL, std::string("cannot open ") + path);
    int64_t n = h.n;
    bool ok = std::fwrite(&n, 8, 1, f) == 1 &&
              std::fwrite(h.ptr.data(), 8, h.n + 1, f) == (size_t)(h.n + 1) &&
              std::fwrite(h.col.data(), 8, h.nnz, f) == (size_t)h.nnz &&
              std::fwrite(a.data(), 8, h.nnz, f) == (size_t)h.nnz &&
              std::fwrite(b.data(), 8, h.n, f) == (size_t)h.n;
    std::fclose(f);
    return ok ? OGSX_OK : fail(OGSX_EINVAL, "short write");
}

int ogsx_profile_enable(ogsx_system* s, int on)
{
    if (!s) return fail(OGSX_EINVAL, "null system");
    if (on && s->ev_pool.empty()) {
        s->ev_pool.resize(1024);
        for (auto& e : s->ev_pool) HIP_TRY(hipEventCreate(&e));
    }
    s->prof = on != 0;
    for (auto& a : s->prof_acc) a = ProfAcc();
    return OGSX_OK;
}

int ogsx_profile_read(ogsx_system* s, int kind, double* total_ms, int64_t* launches,
                      double* bytes_per_launch)
{
    if (!s || kind < 0 || kind > 3) return fail(OGSX_EINVAL, "bad argument");
    OK_TRY(prof_flush(s));
    if (total_ms) *total_ms = s->prof_acc[kind].ms;
    if (launches) *launches = s->prof_acc[kind].launches;
    if (bytes_per_launch) {
        if (kind == OGSX_PROF_SPMV64)
            *bytes_per_launch = spmv_bytes(s, s->fmt64, 8, 8);
        else if (kind == OGSX_PROF_SPMV32)
            *bytes_per_launch = spmv_bytes(s, s->fmt32, 4, 4);
        else if (kind == OGSX_PROF_VELOCITY)
            *bytes_per_launch = velocity_bytes(s);
        else
            *bytes_per_launch = 0.0;
    }
    return OGSX_OK;
}

int ogsx_set_spmv_variant(ogsx_system* s, int lanes_per_row)
{
    if (!s) return fail(OGSX_EINVAL, "null system");
    if (s->sc.active) return fail(OGSX_ESTATE, "system has a solve in progress");
    if (lanes_per_row == 0) {
        s->fmt64 = s->fmt32 = 0;  // autotune on next use
        s->fmt_forced = false;
        return OGSX_OK;
    }
    if (!valid_fmt(s, lanes_per_row))
        return fail(OGSX_EINVAL, "SpMV variant must be 1/2 (tiled CRS, 32/16-bit columns), "
                                 "3/5/6/7/10/11/12 (SELL-64), 13/14/17/18 (SELL-64 with a stencil "
                                 "dictionary), 19/20 (SELL-64 slice triples), 9 (plain CRS, one "
                                 "thread per row), or 4, 8, 16, 32, 64 lanes per row (16-bit, "
                                 "dictionary and triple forms only when the pattern allows)");
    if (needs_extra_layouts(lanes_per_row) && !s->extra_layouts) {
        OK_TRY(alloc_extra_layouts(s));
        // the matrix views of the device descriptor now hold the new arrays
        s->devh.m64 = mat_view(s, (const double*)s->A);
        s->devh.m32 = mat_view(s, (const float*)s->A32);
        HIP_TRY(hipMemcpyAsync(s->dev, &s->devh, sizeof(SysDev), hipMemcpyHostToDevice, s->stream));
        HIP_TRY(hipStreamSynchronize(s->stream));
        delete s->solo;
        s->solo = nullptr;
        if (s->group) OK_TRY(set_members(s->group->set, s->group->set->mem));
    }
    s->fmt64 = s->fmt32 = lanes_per_row;
    s->fmt_forced = true;
    return OGSX_OK;
}

int ogsx_get_spmv_variant(ogsx_system* s, int* lanes_per_row)
{
    if (!s || !lanes_per_row) return fail(OGSX_EINVAL, "null argument");
    OK_TRY(autotune(s));
    *lanes_per_row = s->fmt64;
    return OGSX_OK;
}

int ogsx_get_spmv_formats(ogsx_system* s, int* fmt64, int* fmt32)
{
    if (!s || !fmt64 || !fmt32) return fail(OGSX_EINVAL, "null argument");
    OK_TRY(autotune(s));
    *fmt64 = s->fmt64;
    *fmt32 = s->precision == OGSX_FP32_IR ? s->fmt32 : 0;
    return OGSX_OK;
}

int ogsx_spmv_tuning(ogsx_system* s, int max_entries, int* formats, double* ms, int* count)
{
    if (!s || !count) return fail(OGSX_EINVAL, "null argument");
    OK_TRY(autotune(s));
    const int m = (int)s->tune_log.size();
    for (int i = 0; i < m && i < max_entries; i++) {
        if (formats) formats[i] = s->tune_log[i].first;
        if (ms) ms[i] = s->tune_log[i].second;
    }
    *count = m;
    return OGSX_OK;
}

int ogsx_bench_spmv(ogsx_system* s, int fp32, int reps, double* avg_ms, double* bytes_per_launch)
{
    if (!s || reps < 1 || !avg_ms) return fail(OGSX_EINVAL, "bad argument");
    if (s->sc.active) return fail(OGSX_ESTATE, "system has a solve in progress");
    if (fp32 && s->precision != OGSX_FP32_IR) return fail(OGSX_EINVAL, "system has no FP32 matrix");
    OK_TRY(autotune(s));
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    // the solve's own SpMV kernel (k_cg_spmv), on p = 1
    if (!fp32) {
        OK_TRY(sync_sell(s, is_sell(s->fmt64), false));
        OK_TRY(sync_dia64(s, s->stream));
        k_fill<double><<<blocks_for(s->n, kBlock * 4), kBlock, 0, s->stream>>>(s->n, s->v64.p, 1.0);
        k_set_done<<<1, 64, 0, s->stream>>>(s->st, 0);
        (void)hipEventRecord(e0, s->stream);
        for (int r = 0; r < reps; r++) launch_solve_spmv<double, double>(s, s->fmt64);
        (void)hipEventRecord(e1, s->stream);
    } else {
        if (!is_sell(s->fmt32)) {
            k_cast_f32_g<<<s->G, kBlock, 0, s->stream>>>(Grp{s->dev, nullptr, nullptr});
            s->a32_scaled = 0;
        }
        OK_TRY(sync_sell(s, false, is_sell(s->fmt32)));
        k_fill<float><<<blocks_for(s->n, kBlock * 4), kBlock, 0, s->stream>>>(s->n, s->v32.p, 1.0f);
        k_set_done<<<1, 64, 0, s->stream>>>(s->st, 0);
        (void)hipEventRecord(e0, s->stream);
        for (int r = 0; r < reps; r++) launch_solve_spmv<float, float>(s, s->fmt32);
        (void)hipEventRecord(e1, s->stream);
    }
    k_set_done<<<1, 64, 0, s->stream>>>(s->st, 4);
    HIP_TRY(hipEventSynchronize(e1));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    HIP_TRY(hipGetLastError());
    *avg_ms = ms / reps;
    if (bytes_per_launch)
        *bytes_per_launch = fp32 ? spmv_bytes(s, s->fmt32, 4, 4) : spmv_bytes(s, s->fmt64, 8, 8);
    return OGSX_OK;
}

int ogsx_stream_copy(int64_t bytes, int reps, double* gbs)
{
    if (bytes < 16 || reps < 1 || !gbs) return fail(OGSX_EINVAL, "bad argument");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return fail(OGSX_EHIP, "no HIP device available");
    const int64_t n4 = bytes / 16;
    double2 *a = nullptr, *b = nullptr;
    HIP_TRY(hipMalloc(&a, n4 * 16));
    if (hipMalloc(&b, n4 * 16) != hipSuccess) {
        (void)hipFree(a);
        return fail(OGSX_ENOMEM, "hipMalloc failed");
    }
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    const int nb = 2048 * 8;
    (void)hipMemset(a, 0, n4 * 16);
    k_stream_copy<<<nb, kBlock>>>(n4, a, b);
    (void)hipEventRecord(e0, nullptr);
    for (int r = 0; r < reps; r++) k_stream_copy<<<nb, kBlock>>>(n4, a, b);
    (void)hipEventRecord(e1, nullptr);
    (void)hipEventSynchronize(e1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    *gbs = 2.0 * n4 * 16 * reps / (ms * 1e-3) / 1e9;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    HIP_TRY(hipGetLastError());
    return OGSX_OK;
}

int ogsx_set_persistent(int mode)
{
    if (mode < 0 || mode > 2) return fail(OGSX_EINVAL, "persistent mode must be 0, 1 or 2");
    g_persist_mode.store(mode);
    return OGSX_OK;
}

int ogsx_get_persistent(int* mode, int* supported)
{
    if (mode) *mode = g_persist_mode.load();
    if (supported) *supported = coop_supported() ? 1 : 0;
    return OGSX_OK;
}

int ogsx_persistent_giveup_probe(int* aborted)
{
    if (!aborted) return fail(OGSX_EINVAL, "null argument");
    if (!coop_supported()) return fail(OGSX_EHIP, "cooperative launches are not supported");
    GridSync* gs = nullptr;
    int *ah = nullptr, *ad = nullptr;
    hipStream_t st = nullptr;
    int rc = OGSX_OK;
    if (hipMalloc((void**)&gs, sizeof(GridSync)) != hipSuccess ||
        hipHostMalloc((void**)&ah, sizeof(int), hipHostMallocMapped) != hipSuccess ||
        hipHostGetDevicePointer((void**)&ad, ah, 0) != hipSuccess ||
        hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
        rc = fail(OGSX_EHIP, "set-up of the probe failed");
    } else {
        *ah = 1;  // the host has already asked the waits to give up
        k_zero_sync<<<1, 64, 0, st>>>(gs);
        void* args[] = {&gs, &ad};
        hipError_t e = hipLaunchCooperativeKernel((const void*)k_sync_giveup_probe, dim3(2),
                                                  dim3(kBlock), args, 0, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        GridSync h{};
        if (e == hipSuccess) e = hipMemcpy(&h, gs, sizeof(h), hipMemcpyDeviceToHost);
        if (e != hipSuccess)
            rc = fail(OGSX_EHIP, std::string("probe: ") + hipGetErrorString(e));
        else
            *aborted = h.aborted;
    }
    if (st) (void)hipStreamDestroy(st);
    if (ah) (void)hipHostFree(ah);
    if (gs) (void)hipFree(gs);
    return rc;
}

int ogsx_synchronize(ogsx_system* s)
{
    if (!s) return fail(OGSX_EINVAL, "null system");
    HIP_TRY(hipStreamSynchronize(s->stream));
    return OGSX_OK;
}

}  // extern "C"

// ogsx_solve_crs_noalloc: the zero-allocation FP64 solve over a file-scope
// scratch array
#include "noalloc.inc"
