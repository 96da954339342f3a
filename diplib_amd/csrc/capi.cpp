// C ABI of the Gaussian FIR path: argument checks, tap generation, the pass
// planner (the MI355X counterpart of dip::Framework::Separable) and launches.
//
// Reference behaviour restated here:
//   dip::Gauss / GaussFIR           derivative.cpp:62-86, gauss.cpp:238-286
//   HalfGaussianSize / taps         gauss.cpp:38-146, numeric.h:69-76
//   SeparableConvolution            convolution.cpp:319-421 (filter internalisation :67-162,
//                                   border :355-368, process mask :370-387)
//   Framework::Separable            framework_separable.cpp:35-431 (pass order by input
//                                   stride :146-163, buffer-type intermediates :186-206,
//                                   clamp_cast on the last write :384-418)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <cstdio>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "dipamd.h"
#include "dipamd_internal.hpp"
#include "gauss_taps.hpp"
#include "sep_common.hpp"
#include "sep_config.hpp"
#include "sep_kernels.hpp"
#include "fused_rc.hpp"
#include "fused3d.hpp"

namespace dipamd {

namespace {
thread_local std::string g_last_error;

// Kernel trace: HIP events recorded on the launch stream around every kernel the
// planner issues, for per-kernel timing inside a caller's timed region.
struct TraceRec {
    char name[48];
    int64_t bytes, ops;
    hipEvent_t a, b;
};
std::atomic<bool> g_trace_on{false};
std::mutex g_trace_mu;
std::vector<TraceRec> g_trace;
std::vector<hipEvent_t> g_event_pool;

hipEvent_t pooled_event() {
    if (!g_event_pool.empty()) {
        hipEvent_t e = g_event_pool.back();
        g_event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
}
}  // namespace

TraceScope::TraceScope(const char *name, int64_t bytes, hipStream_t stream, int64_t ops)
    : stream_(stream), index_(-1) {
    if (!g_trace_on.load(std::memory_order_relaxed)) return;
    std::lock_guard<std::mutex> lk(g_trace_mu);
    TraceRec r;
    std::snprintf(r.name, sizeof(r.name), "%s", name);
    r.bytes = bytes;
    r.ops = ops;
    r.a = pooled_event();
    r.b = pooled_event();
    (void)hipEventRecord(r.a, stream);
    g_trace.push_back(r);
    index_ = static_cast<int>(g_trace.size()) - 1;
}

TraceScope::~TraceScope() {
    if (index_ < 0) return;
    std::lock_guard<std::mutex> lk(g_trace_mu);
    if (index_ < static_cast<int>(g_trace.size())) (void)hipEventRecord(g_trace[static_cast<size_t>(index_)].b, stream_);
}

int set_error(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

const char *last_error() { return g_last_error.c_str(); }

size_t dt_size(int dt) {
    switch (dt) {
        case DIPAMD_DT_UINT8:
        case DIPAMD_DT_SINT8:
        case DIPAMD_DT_BIN: return 1;
        case DIPAMD_DT_UINT16:
        case DIPAMD_DT_SINT16: return 2;
        case DIPAMD_DT_UINT32:
        case DIPAMD_DT_SINT32:
        case DIPAMD_DT_SFLOAT: return 4;
        case DIPAMD_DT_DFLOAT: return 8;
        case DIPAMD_DT_SCOMPLEX: return 8;
        case DIPAMD_DT_DCOMPLEX: return 16;
        default: return 0;
    }
}

// DataType::SuggestFlex (datatype.cpp:109-124), real types.
int suggest_flex(int dt) {
    return (dt == DIPAMD_DT_UINT32 || dt == DIPAMD_DT_SINT32 || dt == DIPAMD_DT_DFLOAT) ? DIPAMD_DT_DFLOAT
                                                                                      : DIPAMD_DT_SFLOAT;
}

// ---------------------------------------------------------------------------
// Taps
// ---------------------------------------------------------------------------
int gauss_radius(double sigma, int order, double truncation, bool is_double) {
    if (truncation <= 0) truncation = 3;
    const double eps = is_double ? std::numeric_limits<double>::epsilon()
                                 : static_cast<double>(std::numeric_limits<float>::epsilon());
    const double max_trunc = std::sqrt(-2.0 * std::log(eps));
    truncation = std::min(truncation, max_trunc);
    truncation += 0.5 * static_cast<double>(order);
    const double v = std::ceil(truncation * sigma);
    if (!(v > 0)) return 0;
    if (v > 2147483647.0) return 2147483647;
    return static_cast<int>(v);
}

// Half Gaussian of n = r+1 taps, x = 0 LAST (gauss.cpp:53-146); see gauss_taps.hpp.
std::vector<double> half_gaussian_n(double sigma, int order, int n) { return half_gaussian_t<double>(sigma, order, n); }

// MakeHalfGaussian (gauss.cpp:151-165).
std::vector<double> half_gaussian(double sigma, int order, double truncation, bool is_double) {
    if (sigma == 0.0) return {1.0};
    int n = 1 + gauss_radius(sigma, order, truncation, is_double);
    if (order > 2 && n < 2) n = 2;
    return half_gaussian_n(sigma, order, n);
}

// ---------------------------------------------------------------------------
// Filter internalisation (convolution.cpp:67-162).
// ---------------------------------------------------------------------------
int intern_filter(const double *taps, int n, int sym, int origin_in, InternFilter &F) {
    if (sym == DIPAMD_SYM_CONJ) sym = DIPAMD_SYM_EVEN;
    if (sym == DIPAMD_SYM_D_CONJ) sym = DIPAMD_SYM_D_EVEN;
    if (sym < 0 || sym > DIPAMD_SYM_D_CONJ) return set_error(DIPAMD_E_PARAMETER, "Symmetry string not recognized");
    F.sym = sym;
    F.data.assign(taps, taps + n);
    std::reverse(F.data.begin(), F.data.end());
    F.size = n;
    F.origin = 0;
    F.border = 0;
    F.shift = 0;
    if (n == 0) return DIPAMD_OK;
    if (sym == DIPAMD_SYM_EVEN || sym == DIPAMD_SYM_ODD) F.size = 2 * n - 1;
    if (sym == DIPAMD_SYM_D_EVEN || sym == DIPAMD_SYM_D_ODD) F.size = 2 * n;
    int origin = origin_in < 0 ? F.size / 2 : origin_in;
    if (origin >= F.size) return set_error(DIPAMD_E_PARAMETER, "Origin outside of filter");
    F.origin = F.size - origin - 1;
    F.border = std::max(F.origin, F.size - F.origin - 1);
    F.shift = (sym == DIPAMD_SYM_GENERAL) ? F.origin : F.origin - (n - 1);
    F.zpad = 0;
    if ((sym == DIPAMD_SYM_D_EVEN || sym == DIPAMD_SYM_D_ODD) && F.origin >= n) {
        // The d-* line filter (convolution.cpp:256-283) reads x[i - origin - 1] for its
        // last tap, one sample past the border of :355-368 when the origin lies left of
        // the centre.  The reference's line buffer has no sample there (uninitialised
        // memory, 0 in every observed run); here it is one more border sample, read as 0.
        F.border = F.origin + 1;
        F.zpad = 1;
    }
    return DIPAMD_OK;
}

bool meaningless(const InternFilter &F, bool use_double) {
    if (F.size == 0) return true;
    if (F.size == 1) return use_double ? (F.data[0] == 1.0) : (static_cast<float>(F.data[0]) == 1.0f);
    return false;
}

// Algorithmic VALU lane-ops per output sample of one filter (the trace's ops field):
// symmetric forms, n stored taps: a pair add, a multiply and an add per pair plus the
// centre multiply (fast: add + fma); general form, n taps: multiply + add each (fast: fma).
int64_t filter_ops(const InternFilter &F, bool fast) {
    const int64_t n = static_cast<int64_t>(F.data.size());
    if (F.sym == SYM_NONE || n == 0) return 0;
    if (F.sym == DIPAMD_SYM_GENERAL) return fast ? n : 2 * n - 1;
    if (F.sym == DIPAMD_SYM_EVEN || F.sym == DIPAMD_SYM_ODD) return fast ? 2 * (n - 1) + 1 : 3 * (n - 1) + 1;
    return fast ? 2 * n : 3 * n - 1;  // D_EVEN / D_ODD: n pairs, no centre
}

// ---------------------------------------------------------------------------
// Device scratch: a bump arena per (device, stream).
//
// hipMallocAsync / hipFreeAsync hold the calling thread until the device has drained
// the work queued before them (measured: a malloc behind 144 ms of queued work returned
// after 144 ms), which would make every call that needs an intermediate wait for the
// device.  Calls on one stream run in stream order, so the scratch of one call can be
// handed to the next call on the same stream at once: each stream gets blocks of
// device memory, allocations bump through the last block, and when a call has freed
// everything the arena starts over (blocks allocated while it grew are merged into one
// at the next allocation).  Memory stays with the arena until the process ends (at
// most kMaxArenas arenas; beyond that an idle one is released).  Under graph capture
// the stream-ordered calls are used instead, so a captured graph owns its memory.
// ---------------------------------------------------------------------------
namespace {
std::once_flag g_pool_once[64];

struct ScratchArena {
    std::vector<std::pair<char *, size_t>> blocks;  // allocations bump through the last
    size_t used = 0;                                // bytes used in the last block
    int outstanding = 0;                            // allocations not yet freed
    uint64_t last_use = 0;
};
std::mutex g_arena_mu;
std::map<std::pair<int, hipStream_t>, ScratchArena> g_arenas;
uint64_t g_arena_clock = 0;
constexpr size_t kArenaAlign = 256;
constexpr size_t kArenaMinBlock = size_t(1) << 20;
constexpr size_t kMaxArenas = 16;

bool capturing(hipStream_t stream) {
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    return hipStreamIsCapturing(stream, &st) == hipSuccess && st != hipStreamCaptureStatusNone;
}
}  // namespace

hipError_t scratch_alloc(void **p, size_t bytes, hipStream_t stream) {
    int dev = 0;
    (void)hipGetDevice(&dev);
    if (dev >= 0 && dev < 64) {
        std::call_once(g_pool_once[dev], [dev]() {
            hipMemPool_t pool;
            if (hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess) {
                uint64_t thr = UINT64_MAX;
                (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &thr);
            }
        });
    }
    if (capturing(stream)) return hipMallocAsync(p, bytes ? bytes : 256, stream);
    const size_t need = ((bytes ? bytes : 1) + kArenaAlign - 1) / kArenaAlign * kArenaAlign;
    std::lock_guard<std::mutex> lk(g_arena_mu);
    const auto key = std::make_pair(dev, stream);
    auto it = g_arenas.find(key);
    if (it == g_arenas.end()) {
        if (g_arenas.size() >= kMaxArenas) {  // release the least recently used idle arena
            auto victim = g_arenas.end();
            for (auto jt = g_arenas.begin(); jt != g_arenas.end(); ++jt)
                if (jt->second.outstanding == 0 && (victim == g_arenas.end() || jt->second.last_use < victim->second.last_use))
                    victim = jt;
            if (victim != g_arenas.end()) {
                // its stream may be gone: wait for the device, then nothing uses the blocks
                (void)hipSetDevice(victim->first.first);
                (void)hipDeviceSynchronize();
                for (auto &b : victim->second.blocks) (void)hipFreeAsync(b.first, nullptr);
                (void)hipSetDevice(dev);
                g_arenas.erase(victim);
            }
        }
        it = g_arenas.emplace(key, ScratchArena{}).first;
    }
    ScratchArena &A = it->second;
    A.last_use = ++g_arena_clock;
    if (A.outstanding == 0 && A.blocks.size() > 1) {  // merge the blocks of the last growth
        size_t total = 0;
        for (auto &b : A.blocks) {
            total += b.second;
            (void)hipFreeAsync(b.first, stream);
        }
        A.blocks.clear();
        A.used = 0;
        void *q = nullptr;
        const hipError_t e = hipMallocAsync(&q, total, stream);
        if (e != hipSuccess) return e;
        A.blocks.emplace_back(static_cast<char *>(q), total);
    }
    if (A.blocks.empty() || A.used + need > A.blocks.back().second) {
        void *q = nullptr;
        const hipError_t e = hipMallocAsync(&q, std::max(need, kArenaMinBlock), stream);
        if (e != hipSuccess) return e;
        A.blocks.emplace_back(static_cast<char *>(q), std::max(need, kArenaMinBlock));
        A.used = 0;
    }
    *p = A.blocks.back().first + A.used;
    A.used += need;
    ++A.outstanding;
    return hipSuccess;
}

void scratch_free(void *p, hipStream_t stream) {
    if (!p) return;
    if (capturing(stream)) {
        (void)hipFreeAsync(p, stream);
        return;
    }
    int dev = 0;
    (void)hipGetDevice(&dev);
    std::lock_guard<std::mutex> lk(g_arena_mu);
    auto it = g_arenas.find(std::make_pair(dev, stream));
    if (it == g_arenas.end()) return;
    ScratchArena &A = it->second;
    if (A.outstanding > 0 && --A.outstanding == 0) A.used = 0;
}

// ---------------------------------------------------------------------------
// Views and geometry helpers
// ---------------------------------------------------------------------------
void byte_range(const View &v, int ndim, uintptr_t *lo, uintptr_t *hi) {
    const int64_t es = static_cast<int64_t>(dt_size(v.dt));
    int64_t a = 0, b = 0;
    for (int d = 0; d < ndim; ++d) {
        const int64_t ext = (v.sizes[d] - 1) * v.strides[d];
        if (ext < 0) a += ext; else b += ext;
    }
    const uintptr_t base = reinterpret_cast<uintptr_t>(v.p);
    *lo = base + static_cast<uintptr_t>(a * es);
    *hi = base + static_cast<uintptr_t>(b * es + es);
}

bool views_overlap(const View &x, const View &y, int ndim) {
    uintptr_t a0, a1, b0, b1;
    byte_range(x, ndim, &a0, &a1);
    byte_range(y, ndim, &b0, &b1);
    return a0 < b1 && b0 < a1;
}

View contiguous_view(void *p, int dt, int ndim, const int64_t *sizes) {
    View v;
    v.p = p;
    v.dt = dt;
    int64_t s = 1;
    for (int d = 0; d < ndim; ++d) {
        v.sizes[d] = sizes[d];
        v.strides[d] = s;
        s *= sizes[d];
    }
    return v;
}

void fill_outer(Outer &o, const View &src, const View &dst, int ndim, int skip0, int skip1) {
    // remaining dims, ascending |input stride| for locality
    int dims[DIPAMD_MAX_DIMS];
    int n = 0;
    for (int d = 0; d < ndim; ++d)
        if (d != skip0 && d != skip1 && src.sizes[d] > 1) dims[n++] = d;
    std::sort(dims, dims + n, [&](int a, int b) { return std::llabs(src.strides[a]) < std::llabs(src.strides[b]); });
    o.n = n;
    o.pad_ = 0;
    for (int i = 0; i < kMaxOuter; ++i) {
        o.size[i] = 1;
        o.sin[i] = 0;
        o.sout[i] = 0;
    }
    for (int i = 0; i < n; ++i) {
        o.size[i] = src.sizes[dims[i]];
        o.sin[i] = src.strides[dims[i]];
        o.sout[i] = dst.strides[dims[i]];
    }
}

static int64_t npx_of(const View &v, int ndim) {
    int64_t p = 1;
    for (int d = 0; d < ndim; ++d) p *= v.sizes[d];
    return p;
}

static int64_t product_except(const View &v, int ndim, int skip0, int skip1) {
    int64_t p = 1;
    for (int d = 0; d < ndim; ++d)
        if (d != skip0 && d != skip1) p *= v.sizes[d];
    return p;
}

template <typename TB>
static TB const_value(int bc) {
    switch (bc) {
        case DIPAMD_BC_ADD_MAX_VALUE: return std::numeric_limits<TB>::max();
        case DIPAMD_BC_ADD_MIN_VALUE: return std::numeric_limits<TB>::lowest();
        default: return TB(0);
    }
}

bool bc_supported(int bc) {
    return bc == DIPAMD_BC_SYMMETRIC_MIRROR || bc == DIPAMD_BC_ASYMMETRIC_MIRROR || bc == DIPAMD_BC_PERIODIC ||
           bc == DIPAMD_BC_ASYMMETRIC_PERIODIC || bc == DIPAMD_BC_ADD_ZEROS || bc == DIPAMD_BC_ADD_MAX_VALUE ||
           bc == DIPAMD_BC_ADD_MIN_VALUE || bc == DIPAMD_BC_ZERO_ORDER_EXTRAPOLATE;
}

// Conditions whose extension is computed from the edge values (k_expand.hip).
bool bc_value_based(int bc) {
    return bc == DIPAMD_BC_ANTISYMMETRIC_REFLECT || bc == DIPAMD_BC_FIRST_ORDER_EXTRAPOLATE ||
           bc == DIPAMD_BC_SECOND_ORDER_EXTRAPOLATE || bc == DIPAMD_BC_THIRD_ORDER_EXTRAPOLATE;
}

// ---------------------------------------------------------------------------
// Single-axis and fused passes
// ---------------------------------------------------------------------------
// Bytes of a pass (input + output) that the device's caches hold whole (the 256 MB
// last-level cache, with room to spare).
constexpr int64_t kCachedBytes = int64_t(64) << 20;

static int64_t pick_chunk(int64_t len, int64_t base, int granule, int border, int64_t target_blocks,
                          bool cached = false) {
    int64_t want = std::max<int64_t>(1, (target_blocks + base - 1) / std::max<int64_t>(base, 1));
    int64_t chunk = (len + want - 1) / want;
    chunk = ((chunk + granule - 1) / granule) * granule;
    // the halo (2*border rows) is re-read by each chunk: keep it under ~1/8 of the chunk,
    // unless that leaves fewer than 4 workgroups per CU (small images: then up to 1/2;
    // images that the caches hold whole, `cached`: no limit, the re-reads hit the caches
    // and the waves are what the pass lacks)
    // (a filter without border, e.g. a single tap, needs no halo: minc is one granule)
    int64_t minc = std::max<int64_t>(granule, ((16 * static_cast<int64_t>(border) + granule - 1) / granule) * granule);
    if (base * ((len + minc - 1) / minc) < 1024)
        minc = cached ? granule
                      : std::max<int64_t>(granule, ((4 * static_cast<int64_t>(border) + granule - 1) / granule) * granule);
    chunk = std::max<int64_t>(chunk, std::max<int64_t>(granule, minc));
    int64_t full = ((len + granule - 1) / granule) * granule;
    return std::min(chunk, full);
}

template <typename TB>
static int run_fused(const View &src, const View &dst, int ndim, int dx, int dy, const InternFilter &Fx,
                     const InternFilter &Fy, int bcx, int bcy, bool fast, int cfg, hipStream_t stream);

template <typename TB>
static int run_single(const View &src, const View &dst, int ndim, int d, const InternFilter &F, int bc, bool fast,
                      hipStream_t stream) {
    if (F.border > kLineCap - 1 || F.data.size() > static_cast<size_t>(kLineCap))
        return set_error(DIPAMD_E_UNSUPPORTED, "filter longer than the kernels' tap capacity");
    PassArgs<TB> a;
    std::memset(&a, 0, sizeof(a));
    a.in = src.p;
    a.out = dst.p;
    a.n = src.sizes[d];
    a.sin = src.strides[d];
    a.sout = dst.strides[d];
    a.border = F.border;
    a.ntaps = static_cast<int32_t>(F.data.size());
    a.shift = F.shift;
    a.bc = bc;
    a.zlo = F.zpad ? -(F.border - F.zpad) : INT32_MIN;
    a.cval = const_value<TB>(bc);
    for (size_t k = 0; k < F.data.size(); ++k) a.taps[k] = static_cast<TB>(F.data[k]);
    const InType ti = (src.dt == DIPAMD_DT_UINT8) ? IN_U8 : IN_BUF;
    const bool is_double = sizeof(TB) == 8;
    // line kernel when the filter axis has the smallest |stride| of the source
    int lane = -1;
    for (int k = 0; k < ndim; ++k) {
        if (k == d || src.sizes[k] <= 1) continue;
        if (lane < 0 || std::llabs(src.strides[k]) < std::llabs(src.strides[lane])) lane = k;
    }
    hipError_t err;
    if (lane < 0 || std::llabs(src.strides[d]) <= std::llabs(src.strides[lane])) {
        fill_outer(a.outer, src, dst, ndim, d, -1);
        a.nlines = product_except(src, ndim, d, -1);
        if (line_lds_bytes(F.border, is_double) > 65536)
            return set_error(DIPAMD_E_UNSUPPORTED, "filter border too large for the line kernel");
        TraceScope ts("line", npx_of(src, ndim) * static_cast<int64_t>(dt_size(src.dt) + sizeof(TB)), stream,
                      npx_of(src, ndim) * filter_ops(F, fast));
        err = launch_line(ti, is_double, F.sym, fast, &a, a.n, a.nlines, F.border, stream);
    } else {
        // strided axis: the LDS-ring sweep (the fused kernel with a copy x stage) when
        // the lane axis is contiguous and the window fits; else the global sweep
        const bool ring_ok = (F.sym == DIPAMD_SYM_EVEN || F.sym == DIPAMD_SYM_ODD) && F.shift == 0 &&
                             F.border <= kFusedCap - 1 && src.strides[lane] == 1 &&
                             src.sizes[lane] < (int64_t(1) << 31) && src.sizes[d] < (int64_t(1) << 31);
        // 0 column tiles, 1 LDS ring, 2 global sweep; by default the column tiles for
        // long filters (VALU-bound: the LDS tile keeps the window loads cheap) and the
        // global sweep for short ones (HBM-bound: no staging phase, full occupancy)
        int kind = plan_ysweep(is_double, fast, F.border);
        const bool lane_ok = src.strides[lane] == 1 && src.sizes[d] < (int64_t(1) << 31);
        const int rows = colsweep_rows(F.border, is_double);
        if (kind == 0 && lane_ok && rows > 0) {
            a.e = src.sizes[lane];
            a.ein = src.strides[lane];
            a.eout = dst.strides[lane];
            fill_outer(a.outer, src, dst, ndim, d, lane);
            const int64_t nouter = product_except(src, ndim, d, lane);
            a.chunk = rows;
            const int64_t nchunks = (a.n + rows - 1) / rows;
            TraceScope ts("colsweep", npx_of(src, ndim) * static_cast<int64_t>(dt_size(src.dt) + sizeof(TB)), stream,
                          npx_of(src, ndim) * filter_ops(F, fast));
            err = launch_colsweep(ti, is_double, F.sym, fast, &a, a.e, nchunks, nouter, stream);
            if (err != hipSuccess)
                return set_error(DIPAMD_E_HIP, std::string("kernel launch: ") + hipGetErrorString(err));
            return DIPAMD_OK;
        }
        const int cfg = (ring_ok && kind != 2) ? choose_ysweep_config(is_double, F.border) : -1;
        if (cfg >= 0) {
            InternFilter ident;
            ident.data = {1.0};
            ident.size = 1;
            ident.sym = SYM_NONE;
            return run_fused<TB>(src, dst, ndim, lane, d, ident, F, DIPAMD_BC_SYMMETRIC_MIRROR, bc, fast, cfg, stream);
        }
        a.e = src.sizes[lane];
        a.ein = src.strides[lane];
        a.eout = dst.strides[lane];
        fill_outer(a.outer, src, dst, ndim, d, lane);
        const int64_t nouter = product_except(src, ndim, d, lane);
        const int bs = a.e >= 256 ? 256 : 64;
        const int64_t ntiles = (a.e + bs - 1) / bs;
        const bool zsw = zsweep_enabled() && (F.sym == DIPAMD_SYM_EVEN || F.sym == DIPAMD_SYM_ODD) && F.shift == 0 &&
                         F.border >= 1 && F.border <= kZSweepMaxR &&
                         static_cast<int>(F.data.size()) == F.border + 1;
        // the register-window sweep gains from long chunks (no halo re-reads, long
        // streams), so it aims for fewer workgroups
        const bool cached = npx_of(src, ndim) * static_cast<int64_t>(dt_size(src.dt) + sizeof(TB)) <= kCachedBytes;
        a.chunk = pick_chunk(a.n, ntiles * nouter, 8, F.border, sweep_target_blocks(zsw ? 1024 : 2048), cached);
        const int64_t nchunks = (a.n + a.chunk - 1) / a.chunk;
        TraceScope ts(zsw ? "zsweep" : "sweep",
                      npx_of(src, ndim) * static_cast<int64_t>(dt_size(src.dt) + sizeof(TB)), stream,
                      npx_of(src, ndim) * filter_ops(F, fast));
        err = zsw ? launch_zsweep(ti, is_double, F.border, F.sym, fast, &a, a.e, nchunks, nouter, stream)
                  : launch_sweep(ti, is_double, F.sym, fast, &a, a.e, nchunks, nouter, stream);
    }
    if (err != hipSuccess) return set_error(DIPAMD_E_HIP, std::string("kernel launch: ") + hipGetErrorString(err));
    return DIPAMD_OK;
}

// A pass whose boundary condition is computed from the edge values: the lines are
// written, extended by `border` samples each side, into a buffer-type scratch
// (k_expand), filtered there (the interior outputs never read past the written
// extension, so the condition given to the filter kernel is immaterial), and the
// interior is copied to `dst`.  The reference extends its line buffer the same way
// (CopyBuffer then ExpandBuffer, framework_separable.cpp:356-380).
static int run_cast_out(const View &src, const View &dst, int ndim, hipStream_t stream);

template <typename TB>
static int run_extended_pass(const View &src, const View &dst, int ndim, int d, const InternFilter &F, int bc,
                             bool fast, hipStream_t stream) {
    const int B = F.border;
    if (B == 0) return run_single<TB>(src, dst, ndim, d, F, DIPAMD_BC_ZERO_ORDER_EXTRAPOLATE, fast, stream);
    const int buf_dt = sizeof(TB) == 8 ? DIPAMD_DT_DFLOAT : DIPAMD_DT_SFLOAT;
    int64_t psz[DIPAMD_MAX_DIMS];
    int64_t total = 1;
    for (int k = 0; k < ndim; ++k) {
        psz[k] = src.sizes[k] + (k == d ? 2 * static_cast<int64_t>(B) : 0);
        total *= psz[k];
    }
    void *p1 = nullptr, *p2 = nullptr;
    hipError_t e = scratch_alloc(&p1, static_cast<size_t>(total) * sizeof(TB), stream);
    if (e == hipSuccess) e = scratch_alloc(&p2, static_cast<size_t>(total) * sizeof(TB), stream);
    if (e != hipSuccess) {
        if (p1) scratch_free(p1, stream);
        return set_error(DIPAMD_E_HIP, std::string("scratch allocation: ") + hipGetErrorString(e));
    }
    const View t1 = contiguous_view(p1, buf_dt, ndim, psz);
    const View t2 = contiguous_view(p2, buf_dt, ndim, psz);
    CastGeo g;
    std::memset(&g, 0, sizeof(g));
    g.ndim = ndim;
    g.total = total;
    for (int k = 0; k < ndim; ++k) {
        g.size[k] = psz[k];
        g.sin[k] = src.strides[k];
        g.sout[k] = t1.strides[k];
    }
    int st = DIPAMD_OK;
    {
        TraceScope ts("expand", total * static_cast<int64_t>(sizeof(TB)) +
                                    npx_of(src, ndim) * static_cast<int64_t>(dt_size(src.dt)), stream);
        e = launch_expand(src.dt == DIPAMD_DT_UINT8 ? IN_U8 : IN_BUF, sizeof(TB) == 8, src.p, t1.p, &g, d,
                          static_cast<int>(src.sizes[d]), B - F.zpad, F.zpad, bc, stream);
        if (e != hipSuccess) st = set_error(DIPAMD_E_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
    }
    if (st == DIPAMD_OK) st = run_single<TB>(t1, t2, ndim, d, F, DIPAMD_BC_ZERO_ORDER_EXTRAPOLATE, fast, stream);
    if (st == DIPAMD_OK) {
        View inner = t2;
        inner.p = static_cast<TB *>(t2.p) + static_cast<int64_t>(B) * t2.strides[d];
        for (int k = 0; k < ndim; ++k) inner.sizes[k] = src.sizes[k];
        st = run_cast_out(inner, dst, ndim, stream);
    }
    scratch_free(p2, stream);
    scratch_free(p1, stream);
    return st;
}

template <typename TB>
static int run_fused(const View &src, const View &dst, int ndim, int dx, int dy, const InternFilter &Fx,
                     const InternFilter &Fy, int bcx, int bcy, bool fast, int cfg, hipStream_t stream) {
    FusedArgs<TB> a;
    std::memset(&a, 0, sizeof(a));
    a.in = src.p;
    a.out = dst.p;
    a.nx = src.sizes[dx];
    a.ny = src.sizes[dy];
    a.sxin = src.strides[dx];
    a.syin = src.strides[dy];
    a.sxout = dst.strides[dx];
    a.syout = dst.strides[dy];
    a.rx = Fx.border;
    a.ry = Fy.border;
    a.bcx = bcx;
    a.bcy = bcy;
    a.cvalx = const_value<TB>(bcx);
    a.cvaly = const_value<TB>(bcy);
    for (size_t k = 0; k < Fx.data.size(); ++k) a.tx[k] = static_cast<TB>(Fx.data[k]);
    for (size_t k = 0; k < Fy.data.size(); ++k) a.ty[k] = static_cast<TB>(Fy.data[k]);
    fill_outer(a.outer, src, dst, ndim, dx, dy);
    const int64_t nouter = product_except(src, ndim, dx, dy);
    const int txc = kFusedConfigs[cfg].txc;
    const int64_t ntiles = (a.nx + txc - 1) / txc;
    // about 1024 workgroups (four resident per CU) for float sources: longer chunks than
    // at 2048 halve the x-filtered halo rows per output row (32 x 1024^2 sfloat sigma 3:
    // 91.8 -> 84.9 us); uint8 sources stay at 2048 (8192^2 sigma 4: 200 us against 204);
    // DIPAMD_FUSED_BLOCKS overrides both for experiments
    static const int fused_target = [] {
        const char *v = std::getenv("DIPAMD_FUSED_BLOCKS");
        return v ? std::atoi(v) : 0;
    }();
    const int64_t target = fused_target > 0 ? fused_target : (src.dt == DIPAMD_DT_UINT8 ? 2048 : 1024);
    a.chunk = pick_chunk(a.ny, ntiles * nouter, kFusedStep, Fy.border, target);
    const int64_t nchunks = (a.ny + a.chunk - 1) / a.chunk;
    const InType ti = (src.dt == DIPAMD_DT_UINT8) ? IN_U8 : IN_BUF;
    TraceScope ts(Fx.sym == SYM_NONE ? "ringsweep" : "fused2d",
                  npx_of(src, ndim) * static_cast<int64_t>(dt_size(src.dt) + sizeof(TB)), stream,
                  npx_of(src, ndim) * (filter_ops(Fx, fast) + filter_ops(Fy, fast)));
    // the radius-class kernel (x window in registers) where the plane's rows allow
    // vector loads; the runtime-radius kernel otherwise
    const int rxc = Fx.sym == SYM_NONE ? 0 : std::max(4, rc_class(Fx.border));
    hipError_t err = hipErrorNotSupported;
    if (fused_rc_enabled() && src.strides[dx] == 1 && rxc <= kFusedRcMaxClass &&
        fused_rc_lds_bytes(cfg, rxc, Fy.border, sizeof(TB) == 8) <= lds_limit_bytes()) {
        const int64_t es = static_cast<int64_t>(dt_size(src.dt));
        const int64_t vb = src.dt == DIPAMD_DT_UINT8 ? 4 : 16;  // bytes per vector load
        bool aligned = reinterpret_cast<uintptr_t>(src.p) % static_cast<uintptr_t>(vb) == 0;
        for (int d = 0; d < ndim; ++d)
            if (d != dx && src.sizes[d] > 1 && (src.strides[d] * es) % vb != 0) aligned = false;
        if (aligned)
            err = launch_fused_rc(ti, sizeof(TB) == 8, cfg, Fx.sym, Fy.sym, fast, &a, a.nx, nchunks, nouter, rxc,
                                  Fy.border, stream);
    }
    if (err == hipErrorNotSupported)
        err = launch_fused(ti, sizeof(TB) == 8, cfg, Fx.sym, Fy.sym, fast, &a, a.nx, nchunks, nouter, Fx.border,
                           Fy.border, stream);
    if (err != hipSuccess) return set_error(DIPAMD_E_HIP, std::string("kernel launch: ") + hipGetErrorString(err));
    return DIPAMD_OK;
}

// The fused x+y+z pass (fused3d.hpp) over dims d0 (contiguous), d1, d2 of a float
// source into dst; -100 when its conditions do not hold (the caller then plans the
// passes separately).
bool fused3d_launch_exact_x4(int, int, int, const Fused3Args &, int64_t, size_t, hipStream_t, int *);
bool fused3d_launch_exact_x8(int, int, int, const Fused3Args &, int64_t, size_t, hipStream_t, int *);
bool fused3d_launch_exact_x12(int, int, int, const Fused3Args &, int64_t, size_t, hipStream_t, int *);
bool fused3d_launch_fast_x4(int, int, int, const Fused3Args &, int64_t, size_t, hipStream_t, int *);
bool fused3d_launch_fast_x8(int, int, int, const Fused3Args &, int64_t, size_t, hipStream_t, int *);
bool fused3d_launch_fast_x12(int, int, int, const Fused3Args &, int64_t, size_t, hipStream_t, int *);

static bool fused3d_enabled() {
    static bool on = [] {
        const char *v = std::getenv("DIPAMD_FUSE3D");
        return !(v && v[0] == '0');
    }();
    return on;
}

// Whether the fused 3-D pass takes dims d0, d1, d2 of `src` (float, d0 contiguous).
static bool fused3d_applies(const View &src, int ndim, int d0, int d1, int d2, const InternFilter &Fx,
                            const InternFilter &Fy, const InternFilter &Fz, const int *bc, bool fast) {
    if (!fused3d_enabled() || src.dt != DIPAMD_DT_SFLOAT || src.strides[d0] != 1) return false;
    for (const InternFilter *f : {&Fx, &Fy, &Fz})
        if ((f->sym != DIPAMD_SYM_EVEN && f->sym != DIPAMD_SYM_ODD) || f->shift != 0 ||
            static_cast<int>(f->data.size()) != f->border + 1)
            return false;
    if (Fz.sym != DIPAMD_SYM_EVEN) return false;
    if (Fx.border < 1 || Fx.border > 12 || Fy.border < 1 || Fy.border > kF3MaxRy || Fz.border < 1 || Fz.border > 4)
        return false;
    // index-based conditions only: every sample outside is a sample inside, maybe negated
    for (int d : {d0, d1, d2})
        if (bc[d] != DIPAMD_BC_SYMMETRIC_MIRROR && bc[d] != DIPAMD_BC_ASYMMETRIC_MIRROR &&
            bc[d] != DIPAMD_BC_PERIODIC && bc[d] != DIPAMD_BC_ASYMMETRIC_PERIODIC &&
            bc[d] != DIPAMD_BC_ZERO_ORDER_EXTRAPOLATE)
            return false;
    // The fused pass moves about 8 bytes of HBM per sample against 16 for the separate
    // passes, but is VALU-bound: past about 56 operations per sample (the x stage also
    // filters the 2 ry halo rows of each 48-row tile) the separate passes are faster.
    const double ops = filter_ops(Fx, fast) * (kF3TY + 2.0 * Fy.border) / kF3TY + filter_ops(Fy, fast) +
                       filter_ops(Fz, fast);
    if (ops > 56.0) return false;
    // a tile of 64 x 48 mostly inside and enough planes to stream
    if (src.sizes[d0] < 64 || src.sizes[d1] < 32 || src.sizes[d2] < 8) return false;
    // in-plane offsets of the staged samples are kept as int, two bits of flags below them
    if (src.strides[d1] <= 0 || src.strides[d1] * src.sizes[d1] >= (int64_t(1) << 29)) return false;
    for (int d : {d0, d1, d2})
        if (src.sizes[d] >= (int64_t(1) << 30)) return false;
    return f3_lds_bytes(std::max(4, rc_class(Fx.border)), kF3MaxPlanes) <= lds_limit_bytes();
}

static int run_fused3d(const View &src, const View &dst, int ndim, int d0, int d1, int d2, const InternFilter &Fx,
                       const InternFilter &Fy, const InternFilter &Fz, const int *bc, bool fast, hipStream_t stream) {
    Fused3Args a;
    std::memset(&a, 0, sizeof(a));
    a.in = src.p;
    a.out = static_cast<float *>(dst.p);
    a.nx = static_cast<int>(src.sizes[d0]);
    a.ny = static_cast<int>(src.sizes[d1]);
    a.nz = static_cast<int>(src.sizes[d2]);
    a.syin = src.strides[d1];
    a.szin = src.strides[d2];
    a.sxout = dst.strides[d0];
    a.syout = dst.strides[d1];
    a.szout = dst.strides[d2];
    a.rx = Fx.border;
    a.ry = Fy.border;
    a.bcx = bc[d0];
    a.bcy = bc[d1];
    a.bcz = bc[d2];
    for (size_t k = 0; k < Fx.data.size(); ++k) a.tx[k] = static_cast<float>(Fx.data[k]);
    for (size_t k = 0; k < Fy.data.size(); ++k) a.ty[k] = static_cast<float>(Fy.data[k]);
    for (size_t k = 0; k < Fz.data.size(); ++k) a.tz[k] = static_cast<float>(Fz.data[k]);
    // dims beyond the three
    int dims[DIPAMD_MAX_DIMS];
    int n = 0;
    int64_t nouter = 1;
    for (int d = 0; d < ndim; ++d)
        if (d != d0 && d != d1 && d != d2 && src.sizes[d] > 1) dims[n++] = d;
    a.outer.n = n;
    for (int i = 0; i < kMaxOuter; ++i) {
        a.outer.size[i] = 1;
        a.outer.sin[i] = 0;
        a.outer.sout[i] = 0;
    }
    for (int i = 0; i < n; ++i) {
        a.outer.size[i] = src.sizes[dims[i]];
        a.outer.sin[i] = src.strides[dims[i]];
        a.outer.sout[i] = dst.strides[dims[i]];
        nouter *= src.sizes[dims[i]];
    }
    bool aligned = reinterpret_cast<uintptr_t>(src.p) % 16 == 0;
    for (int d = 0; d < ndim; ++d)
        if (d != d0 && src.sizes[d] > 1 && (src.strides[d] * 4) % 16 != 0) aligned = false;
    a.vec_ok = aligned ? 1 : 0;
    a.ntx = (a.nx + kF3TX - 1) / kF3TX;
    a.nty = (a.ny + kF3TY - 1) / kF3TY;
    const int64_t tiles = static_cast<int64_t>(a.ntx) * a.nty * nouter;
    const int rxc = std::max(4, rc_class(Fx.border));
    const int rz = Fz.border;
    auto launch = [&](int64_t nblocks, size_t lds, int *occ) {
        switch (rxc) {
            case 4: return fast ? fused3d_launch_fast_x4(Fx.sym, Fy.sym, rz, a, nblocks, lds, stream, occ)
                                : fused3d_launch_exact_x4(Fx.sym, Fy.sym, rz, a, nblocks, lds, stream, occ);
            case 8: return fast ? fused3d_launch_fast_x8(Fx.sym, Fy.sym, rz, a, nblocks, lds, stream, occ)
                                : fused3d_launch_exact_x8(Fx.sym, Fy.sym, rz, a, nblocks, lds, stream, occ);
            case 12: return fast ? fused3d_launch_fast_x12(Fx.sym, Fy.sym, rz, a, nblocks, lds, stream, occ)
                                 : fused3d_launch_exact_x12(Fx.sym, Fy.sym, rz, a, nblocks, lds, stream, occ);
        }
        return false;
    };
    // Resident workgroups on the device (per instance: x class, symmetries, z radius,
    // arithmetic), at the largest plane map
    static std::atomic<int> occ_cache[2][3][4][4];
    std::atomic<int> &oc = occ_cache[fast ? 1 : 0][rxc / 4 - 1][(Fx.sym == DIPAMD_SYM_ODD) * 2 + (Fy.sym == DIPAMD_SYM_ODD)][rz - 1];
    int per_cu = oc.load(std::memory_order_relaxed);
    if (per_cu <= 0) {
        if (!launch(0, f3_lds_bytes(rxc, kF3MaxPlanes), &per_cu)) return -100;
        oc.store(per_cu, std::memory_order_relaxed);
    }
    int dev = 0, ncu = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    const int64_t slots = static_cast<int64_t>(per_cu) * std::max(ncu, 1);
    // Plane chunks: a chunk streams 2 rz planes beyond its outputs (and its set-up costs
    // about two more), and a launch of a little more than a whole number of resident
    // rounds leaves most of the device idle in its last round.  The count of chunks
    // minimising (planes per workgroup) x (rounds) is taken: whole rounds up to two,
    // then fractional ones (the scheduler balances many rounds) plus half a round for
    // the tail.
    static const int force = [] {
        const char *v = std::getenv("DIPAMD_FUSE3D_ZCHUNK");
        return v ? std::atoi(v) : 0;
    }();
    const int zmax = kF3MaxPlanes - 2 * rz;
    int64_t zchunk = 0;
    if (force > 0) {
        zchunk = force;
    } else {
        double best = 1e300;
        const int64_t nmax = std::max<int64_t>(1, a.nz / 8);
        for (int64_t nn = 1; nn <= nmax && nn <= 4096; ++nn) {
            const int64_t ch = (a.nz + nn - 1) / nn;
            if (ch > zmax) continue;
            const int64_t n2 = (a.nz + ch - 1) / ch;
            if (n2 != nn) continue;  // the same chunking as a smaller count
            const double wg = static_cast<double>(tiles * n2);
            const double rounds = wg <= 2.0 * slots ? std::ceil(wg / slots) : wg / slots + 0.5;
            const double cost = static_cast<double>(ch + 2 * rz + 2) * rounds;
            if (cost < best * 0.999) {
                best = cost;
                zchunk = ch;
            }
        }
        if (zchunk == 0) zchunk = std::min<int64_t>(a.nz, zmax);
    }
    zchunk = std::min<int64_t>(std::min<int64_t>(zchunk, zmax), a.nz);
    a.zchunk = static_cast<int>(zchunk);
    a.nzchunks = static_cast<int>((a.nz + zchunk - 1) / zchunk);
    const int64_t nblocks = tiles * a.nzchunks;
    if (nblocks >= (int64_t(1) << 31)) return -100;
    const size_t lds = f3_lds_bytes(rxc, a.zchunk + 2 * rz);
    TraceScope ts("fused3d", npx_of(src, ndim) * 8, stream,
                  npx_of(src, ndim) * (filter_ops(Fx, fast) + filter_ops(Fy, fast) + filter_ops(Fz, fast)));
    const bool ok = launch(nblocks, lds, nullptr);
    if (!ok) return -100;
    const hipError_t err = hipGetLastError();
    if (err != hipSuccess) return set_error(DIPAMD_E_HIP, std::string("kernel launch: ") + hipGetErrorString(err));
    return DIPAMD_OK;
}

// Geometry, boundary folds and the wave split of the register-streaming kernels
// (stream2d.hpp, k_stream2d_deriv.hip) for a uint8 source plane stack; -100 when their
// conditions on the source do not hold.  Taps, radii and outputs are left to the caller.
static int stream_setup(StreamArgs &a, const View &src, const View &dst, int ndim, int dx, int dy, int bcx, int bcy) {
    if (src.dt != DIPAMD_DT_UINT8) return -100;
    if (src.strides[dx] != 1 || dst.strides[dx] != 1) return -100;
    if (src.sizes[dx] >= (int64_t(1) << 30) || src.sizes[dy] >= (int64_t(1) << 30)) return -100;
    if (src.sizes[dx] < 16 || src.sizes[dy] < 16) return -100;  // its boundary folds reflect once
    if (src.sizes[dx] & 3) return -100;                          // whole 4-sample words per row
    if (!bc_supported(bcx) || !bc_supported(bcy)) return -100;
    // word loads need 4-byte aligned rows in every plane
    if ((reinterpret_cast<uintptr_t>(src.p) & 3) || (src.strides[dy] & 3)) return -100;
    for (int d = 0; d < ndim; ++d)
        if (d != dx && d != dy && src.sizes[d] > 1 && (src.strides[d] & 3)) return -100;
    std::memset(&a, 0, sizeof(a));
    a.in = static_cast<const uint8_t *>(src.p);
    a.nx = static_cast<int>(src.sizes[dx]);
    a.ny = static_cast<int>(src.sizes[dy]);
    a.syin = src.strides[dy];
    a.syout = dst.strides[dy];
    a.bcx = bcx;
    a.bcy = bcy;
    a.cvalx = const_value<float>(bcx);
    a.cvaly = const_value<float>(bcy);
    fill_outer(a.outer, src, dst, ndim, dx, dy);
    // one-reflection row fold coefficients (fold_once in stream2d.hpp, as arithmetic)
    const int ny = static_cast<int>(src.sizes[dy]);
    switch (bcy) {
        case DIPAMD_BC_SYMMETRIC_MIRROR:
        case DIPAMD_BC_ASYMMETRIC_MIRROR:
            a.ylo_m = -1, a.ylo_c = 0, a.yhi_m = -1, a.yhi_c = 2 * (ny - 1);
            a.yflag = bcy == DIPAMD_BC_ASYMMETRIC_MIRROR ? 1 : 0;
            break;
        case DIPAMD_BC_PERIODIC:
        case DIPAMD_BC_ASYMMETRIC_PERIODIC:
            a.ylo_m = 1, a.ylo_c = ny, a.yhi_m = 1, a.yhi_c = -ny;
            a.yflag = bcy == DIPAMD_BC_ASYMMETRIC_PERIODIC ? 1 : 0;
            break;
        case DIPAMD_BC_ZERO_ORDER_EXTRAPOLATE:
            a.ylo_m = 0, a.ylo_c = 0, a.yhi_m = 0, a.yhi_c = ny - 1, a.yflag = 0;
            break;
        default:  // constant conditions: any valid row, value replaced
            a.ylo_m = 0, a.ylo_c = 0, a.yhi_m = 0, a.yhi_c = 0, a.yflag = 2;
            break;
    }
    const int64_t nouter = product_except(src, ndim, dx, dy);
    a.nstrips = static_cast<int>((a.nx + 255) / 256);
    // enough waves to fill the chip (~24 per CU), chunks of at least 32 rows
    const int64_t target = stream2d_target_waves();
    const int64_t base = static_cast<int64_t>(a.nstrips) * nouter;
    int64_t ch = (static_cast<int64_t>(a.ny) * base + target - 1) / target;
    ch = std::max<int64_t>(32, std::min<int64_t>(ch, a.ny));
    a.chunk = static_cast<int>(ch);
    a.nchunks = static_cast<int>((a.ny + ch - 1) / ch);
    a.nwaves = base * a.nchunks;
    a.nplanes = nouter;
    return DIPAMD_OK;
}

// Output rows of a streaming kernel: 16-byte stores of 4 samples.
static bool stream_out_aligned(const View &dst, int ndim, int dx, int dy) {
    if (dst.strides[dx] != 1) return false;
    bool aligned = (reinterpret_cast<uintptr_t>(dst.p) & 15) == 0 && (dst.strides[dy] & 3) == 0;
    for (int d = 0; d < ndim; ++d)
        if (d != dx && d != dy && dst.sizes[d] > 1 && (dst.strides[d] & 3)) aligned = false;
    return aligned;
}

// The register-streaming fused kernel for uint8 planes and radii <= 4 (stream2d.hpp).
// Returns -100 when its conditions do not hold (the caller then uses the LDS kernel).
static int run_stream2d(const View &src, const View &dst, int ndim, int dx, int dy, const InternFilter &Fx,
                        const InternFilter &Fy, int bcx, int bcy, bool fast, hipStream_t stream) {
    if (!stream2d_enabled()) return -100;
    auto sym_ok = [](const InternFilter &F) {
        return (F.sym == DIPAMD_SYM_EVEN || F.sym == DIPAMD_SYM_ODD) && F.shift == 0 && F.border >= 1 &&
               F.border <= kStreamRMaxHost && static_cast<int>(F.data.size()) == F.border + 1;
    };
    if (!sym_ok(Fx) || !sym_ok(Fy)) return -100;
    if (!stream_out_aligned(dst, ndim, dx, dy)) return -100;  // 16-byte stores of 4 outputs
    StreamArgs a;
    if (stream_setup(a, src, dst, ndim, dx, dy, bcx, bcy) != DIPAMD_OK) return -100;
    a.out = static_cast<float *>(dst.p);
    a.aligned_out = 1;
    a.rx = Fx.border;
    a.ry = Fy.border;
    for (size_t k = 0; k < Fx.data.size(); ++k) a.tx[k] = static_cast<float>(Fx.data[k]);
    for (size_t k = 0; k < Fy.data.size(); ++k) a.ty[k] = static_cast<float>(Fy.data[k]);
    TraceScope ts("stream2d", npx_of(src, ndim) * static_cast<int64_t>(1 + sizeof(float)), stream,
                  npx_of(src, ndim) * (filter_ops(Fx, fast) + filter_ops(Fy, fast)));
    hipError_t err = launch_stream2d(Fx.sym, Fy.sym, fast, &a, stream);
    if (err != hipSuccess) return set_error(DIPAMD_E_HIP, std::string("kernel launch: ") + hipGetErrorString(err));
    return DIPAMD_OK;
}

int run_stream2d_deriv(int mode, const View &src, const View *outs, int nouts, const InternFilter *const *fx,
                       const InternFilter *const *fy, int bcx, int bcy, bool fast, hipStream_t stream) {
    if (!stream2d_enabled() || nouts < 1 || nouts > 3) return -100;
    const int ndim = 2, dx = 0, dy = 1;
    for (int o = 0; o < 3; ++o) {
        const InternFilter *f[2] = {fx[o], fy[o]};
        for (const InternFilter *F : f)
            if (F && !((F->sym == (o == 1 ? DIPAMD_SYM_ODD : DIPAMD_SYM_EVEN)) && F->shift == 0 && F->border >= 1 &&
                       F->border <= kStreamRMaxHost && static_cast<int>(F->data.size()) == F->border + 1))
                return -100;
    }
    for (int k = 0; k < nouts; ++k) {
        if (outs[k].dt != DIPAMD_DT_SFLOAT || !stream_out_aligned(outs[k], ndim, dx, dy)) return -100;
        if (views_overlap(outs[k], src, ndim)) return -100;
        for (int j = 0; j < k; ++j)
            if (views_overlap(outs[k], outs[j], ndim)) return -100;
    }
    DerivStreamArgs a;
    std::memset(&a, 0, sizeof(a));
    if (stream_setup(a.s, src, outs[0], ndim, dx, dy, bcx, bcy) != DIPAMD_OK) return -100;
    for (int k = 0; k < nouts; ++k) {
        a.out[k] = static_cast<float *>(outs[k].p);
        a.syout[k] = outs[k].strides[dy];
    }
    int64_t ops = 0;
    for (int o = 0; o < 3; ++o) {
        // an order the function does not use: radius 1, zero taps (never read)
        a.rx[o] = fx[o] ? fx[o]->border : 1;
        a.ry[o] = fy[o] ? fy[o]->border : 1;
        if (fx[o]) {
            for (size_t k = 0; k < fx[o]->data.size(); ++k) a.tx[o][k] = static_cast<float>(fx[o]->data[k]);
            ops += filter_ops(*fx[o], fast);
        }
        if (fy[o]) {
            for (size_t k = 0; k < fy[o]->data.size(); ++k) a.ty[o][k] = static_cast<float>(fy[o]->data[k]);
            ops += filter_ops(*fy[o], fast);
        }
    }
    TraceScope ts("stream2d_deriv", npx_of(src, ndim) * static_cast<int64_t>(1 + nouts * sizeof(float)), stream,
                  npx_of(src, ndim) * ops);
    const hipError_t err = launch_stream2d_deriv(mode, fast, &a, stream);
    if (err != hipSuccess) return set_error(DIPAMD_E_HIP, std::string("kernel launch: ") + hipGetErrorString(err));
    return DIPAMD_OK;
}

static int run_cast_in(const View &src, const View &dst, int ndim, bool to_double, hipStream_t stream) {
    CastGeo g;
    std::memset(&g, 0, sizeof(g));
    g.ndim = ndim;
    g.total = 1;
    for (int d = 0; d < ndim; ++d) {
        g.size[d] = src.sizes[d];
        g.sin[d] = src.strides[d];
        g.sout[d] = dst.strides[d];
        g.total *= src.sizes[d];
    }
    TraceScope ts("cast_in", g.total * static_cast<int64_t>(dt_size(src.dt) + dt_size(dst.dt)), stream);
    hipError_t err = launch_cast_in(src.dt, to_double, src.p, dst.p, &g, g.total, stream);
    if (err != hipSuccess) return set_error(DIPAMD_E_HIP, std::string("cast launch: ") + hipGetErrorString(err));
    return DIPAMD_OK;
}

static int run_cast_out(const View &src, const View &dst, int ndim, hipStream_t stream) {
    CastGeo g;
    std::memset(&g, 0, sizeof(g));
    g.ndim = ndim;
    g.total = 1;
    for (int d = 0; d < ndim; ++d) {
        g.size[d] = src.sizes[d];
        g.sin[d] = src.strides[d];
        g.sout[d] = dst.strides[d];
        g.total *= src.sizes[d];
    }
    TraceScope ts("cast_out", g.total * static_cast<int64_t>(dt_size(src.dt) + dt_size(dst.dt)), stream);
    hipError_t err = launch_cast_out(src.dt == DIPAMD_DT_DFLOAT, dst.dt, src.p, dst.p, &g, g.total, stream);
    if (err != hipSuccess) return set_error(DIPAMD_E_HIP, std::string("cast launch: ") + hipGetErrorString(err));
    return DIPAMD_OK;
}

// Framework::Separable over device views.  `filters[d]` is the internal filter of
// dim d; `process[d]` the final process mask (already cleared for size-1 dims
// and meaningless filters).
int separable_device(const View &in, const View &out, int ndim, const InternFilter *const *filters,
                     const int *bc, const bool *process, uint32_t flags, hipStream_t stream, const View *keep,
                     bool *kept) {
    if (kept) *kept = false;
    const int buf_dt = suggest_flex(in.dt);
    const bool is_double = buf_dt == DIPAMD_DT_DFLOAT;
    const bool fast = (flags & DIPAMD_FAST) != 0;
    // pass order: processed dims sorted by (signed) input stride, stable (sortIndices)
    std::vector<int> order;
    for (int d = 0; d < ndim; ++d)
        if (process[d] && in.sizes[d] > 1) order.push_back(d);
    for (size_t i = 1; i < order.size(); ++i) {
        int e = order[i];
        size_t j = i;
        while (j > 0 && in.strides[order[j - 1]] > in.strides[e]) {
            order[j] = order[j - 1];
            --j;
        }
        order[j] = e;
    }
    for (int d : order) {
        if (bc[d] == DIPAMD_BC_ALREADY_EXPANDED)  // ExpandBuffer's default case (copy_buffer.cpp:745)
            return set_error(DIPAMD_E_UNSUPPORTED, "Functionality has not (yet) been implemented");
        if (!bc_supported(bc[d]) && !bc_value_based(bc[d]))
            return set_error(DIPAMD_E_UNSUPPORTED, "boundary condition not supported");
    }
    std::vector<void *> scratch;
    auto alloc = [&](int dt, View *v) -> int {
        int64_t n = 1;
        for (int d = 0; d < ndim; ++d) n *= in.sizes[d];
        void *p = nullptr;
        hipError_t e = scratch_alloc(&p, static_cast<size_t>(n) * dt_size(dt), stream);
        if (e != hipSuccess) return set_error(DIPAMD_E_HIP, std::string("scratch allocation: ") + hipGetErrorString(e));
        scratch.push_back(p);
        *v = contiguous_view(p, dt, ndim, in.sizes);
        return DIPAMD_OK;
    };
    auto release = [&]() {
        for (void *p : scratch) scratch_free(p, stream);
    };
    int st = DIPAMD_OK;
    if (order.empty()) {
        // output.Copy( input ) with conversion
        if (views_overlap(in, out, ndim) && in.p == out.p && in.dt == out.dt) return DIPAMD_OK;
        View t;
        if ((st = alloc(buf_dt, &t)) != DIPAMD_OK) return st;
        if ((st = run_cast_in(in, t, ndim, is_double, stream)) == DIPAMD_OK) st = run_cast_out(t, out, ndim, stream);
        release();
        return st;
    }

    // Build the pass list: optionally fuse the first two dims.
    struct PassPlan {
        bool fused;
        int d0, d1;
        int plan;  // fused passes: plan_fusion's choice; -3: the fused x+y+z pass (d2)
        int d2;
    };
    // the first pass reads uint8 directly; other types are converted to the buffer type
    const int first_src_dt = (in.dt == DIPAMD_DT_UINT8 && !is_double) ? DIPAMD_DT_UINT8 : buf_dt;
    std::vector<PassPlan> passes;
    size_t i0 = 0;
    if (!(flags & DIPAMD_NO_FUSION) && !is_double && order.size() >= 3 && in.dt == DIPAMD_DT_SFLOAT &&
        fused3d_applies(in, ndim, order[0], order[1], order[2], *filters[order[0]], *filters[order[1]],
                        *filters[order[2]], bc, fast)) {
        passes.push_back({true, order[0], order[1], -3, order[2]});
        i0 = 3;
    }
    if (i0 == 0 && !(flags & DIPAMD_NO_FUSION) && order.size() >= 2) {
        const InternFilter &Fx = *filters[order[0]];
        const InternFilter &Fy = *filters[order[1]];
        const bool symok = (Fx.sym == DIPAMD_SYM_EVEN || Fx.sym == DIPAMD_SYM_ODD) &&
                           (Fy.sym == DIPAMD_SYM_EVEN || Fy.sym == DIPAMD_SYM_ODD) && Fx.shift == 0 && Fy.shift == 0;
        // the fused kernel keeps column offsets (x * x-stride) and row indices in 32 bits
        const bool fits = Fx.border <= kFusedCap - 1 && Fy.border <= kFusedCap - 1 &&
                          in.sizes[order[0]] * std::llabs(in.strides[order[0]]) < (int64_t(1) << 31) &&
                          in.sizes[order[1]] < (int64_t(1) << 31);
        const bool index_bcs = !bc_value_based(bc[order[0]]) && !bc_value_based(bc[order[1]]);
        // taps beyond the border would be dropped: a symmetric filter's ntaps is border + 1
        const bool whole = static_cast<int>(Fx.data.size()) == Fx.border + 1 &&
                           static_cast<int>(Fy.data.size()) == Fy.border + 1;
        if (symok && fits && index_bcs && whole) {
            const int plan = plan_fusion(first_src_dt, is_double, fast, Fx.border, Fy.border, in.sizes[order[0]],
                                         in.sizes[order[1]], flags);
            if (plan >= 0) {
                passes.push_back({true, order[0], order[1], plan, -1});
                i0 = 2;
            }
        }
    }
    for (size_t i = i0; i < order.size(); ++i) passes.push_back({false, order[i], -1, -1, -1});

    // Source of the first pass: the input, converted first if its type has no kernel.
    View src = in;
    if (!(in.dt == DIPAMD_DT_UINT8 && !is_double) && in.dt != buf_dt) {
        View t;
        if ((st = alloc(buf_dt, &t)) != DIPAMD_OK) return st;
        if ((st = run_cast_in(in, t, ndim, is_double, stream)) != DIPAMD_OK) {
            release();
            return st;
        }
        src = t;
    }
    // Destination of the last pass: the output if it holds the buffer type, else a scratch.
    const bool out_is_buf = out.dt == buf_dt;
    View final_dst = out;
    if (!out_is_buf) {
        if ((st = alloc(buf_dt, &final_dst)) != DIPAMD_OK) return st;
    }
    View s1, s2;
    bool have_s1 = false, have_s2 = false;
    const size_t m = passes.size();
    for (size_t k = 0; k < m && st == DIPAMD_OK; ++k) {
        // choose a destination that does not overlap the pass's source
        View dst;
        const bool last = k + 1 == m;
        const size_t from_end = m - 1 - k;
        bool want_final = (from_end % 2) == 0;
        if (want_final && !views_overlap(final_dst, src, ndim)) {
            dst = final_dst;
        } else {
            if (!have_s1) {
                if ((st = alloc(buf_dt, &s1)) != DIPAMD_OK) break;
                have_s1 = true;
            }
            if (!views_overlap(s1, src, ndim)) {
                dst = s1;
            } else {
                if (!have_s2) {
                    if ((st = alloc(buf_dt, &s2)) != DIPAMD_OK) break;
                    have_s2 = true;
                }
                dst = s2;
            }
        }
        const PassPlan &pp = passes[k];
        if (k == 0 && keep && kept && pp.fused && pp.plan != -3 && pp.d0 == 0 && pp.d1 == 1 && m >= 2 &&
            keep->dt == buf_dt && !views_overlap(*keep, src, ndim)) {
            dst = *keep;  // the intermediate after dims 0 and 1, for the caller
            *kept = true;
        }
        if (pp.fused && pp.plan == -3) {
            st = run_fused3d(src, dst, ndim, pp.d0, pp.d1, pp.d2, *filters[pp.d0], *filters[pp.d1], *filters[pp.d2], bc,
                             fast, stream);
            if (st == -100) st = set_error(DIPAMD_E_HIP, "fused 3-D pass: no kernel instance");
        } else if (pp.fused) {
            st = -100;
            if (pp.plan == 3 && !is_double)
                st = run_stream2d(src, dst, ndim, pp.d0, pp.d1, *filters[pp.d0], *filters[pp.d1], bc[pp.d0],
                                  bc[pp.d1], fast, stream);
            if (st == -100) {
                const int cfg = pp.plan != 3 ? pp.plan
                                             : plan_fused_config(src.dt == DIPAMD_DT_UINT8 ? DIPAMD_DT_UINT8 : buf_dt,
                                                                 is_double, fast, filters[pp.d0]->border,
                                                                 filters[pp.d1]->border, src.sizes[pp.d0],
                                                                 src.sizes[pp.d1]);
                st = is_double ? run_fused<double>(src, dst, ndim, pp.d0, pp.d1, *filters[pp.d0], *filters[pp.d1],
                                                   bc[pp.d0], bc[pp.d1], fast, cfg, stream)
                               : run_fused<float>(src, dst, ndim, pp.d0, pp.d1, *filters[pp.d0], *filters[pp.d1],
                                                  bc[pp.d0], bc[pp.d1], fast, cfg, stream);
            }
        } else if (bc_value_based(bc[pp.d0])) {
            st = is_double
                     ? run_extended_pass<double>(src, dst, ndim, pp.d0, *filters[pp.d0], bc[pp.d0], fast, stream)
                     : run_extended_pass<float>(src, dst, ndim, pp.d0, *filters[pp.d0], bc[pp.d0], fast, stream);
        } else {
            st = is_double ? run_single<double>(src, dst, ndim, pp.d0, *filters[pp.d0], bc[pp.d0], fast, stream)
                           : run_single<float>(src, dst, ndim, pp.d0, *filters[pp.d0], bc[pp.d0], fast, stream);
        }
        if (last && st == DIPAMD_OK && dst.p != final_dst.p) {
            // the last pass could not write the final view directly (it overlapped the source)
            st = run_cast_out(dst, final_dst, ndim, stream);
        }
        src = dst;
    }
    if (st == DIPAMD_OK && !out_is_buf) st = run_cast_out(final_dst, out, ndim, stream);
    release();
    return st;
}

// One filter pass along dim d from `src` (buffer type or uint8) into `dst` (buffer type).
int run_pass(const View &src, const View &dst, int ndim, int d, const InternFilter &F, int bc, bool is_double,
             bool fast, hipStream_t stream) {
    return is_double ? run_single<double>(src, dst, ndim, d, F, bc, fast, stream)
                     : run_single<float>(src, dst, ndim, d, F, bc, fast, stream);
}

template <typename TB>
static int run_pass_dot_t(const View &src, const View &u, int ndim, int d, const InternFilter &F, int bc, bool fast,
                          double *out, hipStream_t stream) {
    // a strided axis (the register sweep's lane axis is the one of smallest stride)
    int lane = -1;
    for (int k = 0; k < ndim; ++k) {
        if (k == d || src.sizes[k] <= 1) continue;
        if (lane < 0 || std::llabs(src.strides[k]) < std::llabs(src.strides[lane])) lane = k;
    }
    if (lane < 0 || std::llabs(src.strides[d]) <= std::llabs(src.strides[lane])) return -100;
    if ((F.sym != DIPAMD_SYM_EVEN && F.sym != DIPAMD_SYM_ODD) || F.shift != 0 || F.border < 1 ||
        F.border > kZSweepMaxR || static_cast<int>(F.data.size()) != F.border + 1 || bc_value_based(bc) ||
        src.sizes[d] >= (int64_t(1) << 31))
        return -100;
    PassArgs<TB> a;
    std::memset(&a, 0, sizeof(a));
    a.in = src.p;
    a.out = nullptr;
    a.n = src.sizes[d];
    a.sin = src.strides[d];
    a.sout = u.strides[d];
    a.border = F.border;
    a.ntaps = static_cast<int32_t>(F.data.size());
    a.bc = bc;
    a.zlo = F.zpad ? -(F.border - F.zpad) : INT32_MIN;
    a.cval = const_value<TB>(bc);
    for (size_t k = 0; k < F.data.size(); ++k) a.taps[k] = static_cast<TB>(F.data[k]);
    a.e = src.sizes[lane];
    a.ein = src.strides[lane];
    a.eout = u.strides[lane];
    fill_outer(a.outer, src, u, ndim, d, lane);
    const int64_t nouter = product_except(src, ndim, d, lane);
    const int bs = a.e >= 256 ? 256 : 64;
    const int64_t ntiles = (a.e + bs - 1) / bs;
    const bool cached = npx_of(src, ndim) * static_cast<int64_t>(2 * sizeof(TB)) <= kCachedBytes;
    a.chunk = pick_chunk(a.n, ntiles * nouter, 8, F.border, sweep_target_blocks(1024), cached);
    const int64_t nchunks = (a.n + a.chunk - 1) / a.chunk;
    const int64_t nblk = ntiles * nchunks * nouter;
    if (nblk >= (int64_t(1) << 31)) return -100;
    double *partials = nullptr;
    hipError_t e = scratch_alloc(reinterpret_cast<void **>(&partials), static_cast<size_t>(nblk) * sizeof(double), stream);
    if (e != hipSuccess) return set_error(DIPAMD_E_HIP, std::string("scratch allocation: ") + hipGetErrorString(e));
    {
        TraceScope ts("zsweep_dot", npx_of(src, ndim) * static_cast<int64_t>(2 * sizeof(TB)), stream,
                      npx_of(src, ndim) * (filter_ops(F, fast) + 2));
        e = launch_zsweep_dot(sizeof(TB) == 8, F.border, F.sym, fast, &a, a.e, nchunks, nouter, u.p, partials, out,
                              stream);
    }
    scratch_free(partials, stream);
    if (e != hipSuccess) return set_error(DIPAMD_E_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
    return DIPAMD_OK;
}

int run_pass_dot(const View &src, const View &u, int ndim, int d, const InternFilter &F, int bc, bool is_double,
                 bool fast, double *out, hipStream_t stream) {
    if (src.dt != (is_double ? DIPAMD_DT_DFLOAT : DIPAMD_DT_SFLOAT) || u.dt != src.dt) return -100;
    return is_double ? run_pass_dot_t<double>(src, u, ndim, d, F, bc, fast, out, stream)
                     : run_pass_dot_t<float>(src, u, ndim, d, F, bc, fast, out, stream);
}

// Element-wise conversion passes (exposed for the backward form).
int cast_to_buffer(const View &src, const View &dst, int ndim, bool to_double, hipStream_t stream) {
    return run_cast_in(src, dst, ndim, to_double, stream);
}
int cast_from_buffer(const View &src, const View &dst, int ndim, hipStream_t stream) {
    return run_cast_out(src, dst, ndim, stream);
}

// ---------------------------------------------------------------------------
// Argument checks shared by the entry points
// ---------------------------------------------------------------------------
int check_image(const dipamd_image *im, const char *what) {
    if (!im) return set_error(DIPAMD_E_PARAMETER, std::string(what) + ": null image");
    if (!im->origin) return set_error(DIPAMD_E_NOT_FORGED, "Image is not forged");
    if (im->ndim < 1 || im->ndim > DIPAMD_MAX_DIMS)
        return set_error(DIPAMD_E_PARAMETER, "Dimensionality not supported");
    if (dt_size(im->dtype) == 0) return set_error(DIPAMD_E_PARAMETER, "Data type not supported");
    for (int d = 0; d < im->ndim; ++d)
        if (im->sizes[d] < 1) return set_error(DIPAMD_E_PARAMETER, "Sizes must be positive");
    return DIPAMD_OK;
}

int require_real(const dipamd_image *im) {
    if (im && (im->dtype == DIPAMD_DT_SCOMPLEX || im->dtype == DIPAMD_DT_DCOMPLEX))
        return set_error(DIPAMD_E_UNSUPPORTED, "complex images are taken by the separable convolution and the FIR "
                                               "Gaussian only");
    return DIPAMD_OK;
}

View to_view(const dipamd_image *im) {
    View v;
    v.p = im->origin;
    v.dt = im->dtype;
    for (int d = 0; d < DIPAMD_MAX_DIMS; ++d) {
        v.sizes[d] = d < im->ndim ? im->sizes[d] : 1;
        v.strides[d] = d < im->ndim ? im->strides[d] : 0;
    }
    return v;
}

int broadcast_bc(const int32_t *bc, int nbc, int ndim, int *out) {
    if (nbc != 0 && nbc != 1 && nbc != ndim) return set_error(DIPAMD_E_ARRAY_LENGTH, "Array parameter has the wrong number of elements");
    for (int d = 0; d < ndim; ++d) {
        const int b = nbc == 0 ? DIPAMD_BC_SYMMETRIC_MIRROR : (nbc == 1 ? bc[0] : bc[d]);
        if (b < 0 || b > 12) return set_error(DIPAMD_E_PARAMETER, "Boundary condition not recognized");
        out[d] = b;
    }
    return DIPAMD_OK;
}

}  // namespace dipamd

using namespace dipamd;

extern "C" {

const char *dipamd_last_error(void) { return last_error(); }

int dipamd_trace_enable(int32_t on) {
    std::lock_guard<std::mutex> lk(g_trace_mu);
    for (TraceRec &r : g_trace) {
        g_event_pool.push_back(r.a);
        g_event_pool.push_back(r.b);
    }
    g_trace.clear();
    g_trace_on.store(on != 0);
    return DIPAMD_OK;
}

int dipamd_trace_read(dipamd_trace_record *records, int32_t capacity) {
    std::lock_guard<std::mutex> lk(g_trace_mu);
    int n = 0;
    for (TraceRec &r : g_trace) {
        if (n >= capacity) break;
        (void)hipEventSynchronize(r.b);
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, r.a, r.b) != hipSuccess) ms = -1.f;
        std::snprintf(records[n].name, sizeof(records[n].name), "%s", r.name);
        records[n].ms = ms;
        records[n].bytes = r.bytes;
        records[n].ops = r.ops;
        ++n;
    }
    return n;
}
const char *dipamd_version(void) { return "dipamd 0.1 (DIPlib 3.6.1 Gauss FIR path, gfx950)"; }

int dipamd_gauss_radius(double sigma, int32_t order, double truncation, int32_t is_double) {
    return gauss_radius(sigma, order, truncation, is_double != 0);
}

int dipamd_make_half_gaussian(double sigma, int32_t order, double truncation, int32_t is_double, double *taps,
                              int32_t capacity) {
    if (order < 0 || order > 3) return set_error(DIPAMD_E_ORDER, "Gaussian FIR filter not implemented for order > 3");
    std::vector<double> f = half_gaussian(sigma, order, truncation, is_double != 0);
    if (static_cast<int>(f.size()) > capacity) return set_error(DIPAMD_E_PARAMETER, "capacity too small");
    std::memcpy(taps, f.data(), f.size() * sizeof(double));
    return static_cast<int>(f.size());
}

static int separable_convolution_impl(const dipamd_image *in, const dipamd_image *out, const dipamd_filter *filters,
                                      int32_t nfilters, const int32_t *bc, int32_t nbc, const uint8_t *process,
                                      uint32_t flags, void *stream, const View *keep, bool *kept) {
    int st;
    if ((st = check_image(in, "in")) != DIPAMD_OK) return st;
    if ((st = check_image(out, "out")) != DIPAMD_OK) return st;
    const int nd = in->ndim;
    if (out->ndim != nd) return set_error(DIPAMD_E_SIZES, "Sizes don't match");
    for (int d = 0; d < nd; ++d)
        if (out->sizes[d] != in->sizes[d]) return set_error(DIPAMD_E_SIZES, "Sizes don't match");
    if (nfilters != 1 && nfilters != nd)
        return set_error(DIPAMD_E_ARRAY_LENGTH, "Array parameter has the wrong number of elements");
    int bcs[DIPAMD_MAX_DIMS];
    if ((st = broadcast_bc(bc, nbc, nd, bcs)) != DIPAMD_OK) return st;
    // a complex filter or a complex image: the complex buffer path (convolution.cpp:331-346)
    bool any_complex = false;
    for (int i = 0; i < nfilters; ++i) any_complex = any_complex || filters[i].is_complex != 0;
    if (any_complex || in->dtype == DIPAMD_DT_SCOMPLEX || in->dtype == DIPAMD_DT_DCOMPLEX)
        return separable_complex(to_view(in), to_view(out), nd, filters, nfilters, bcs, process, any_complex,
                                 static_cast<hipStream_t>(stream));
    if (out->dtype == DIPAMD_DT_SCOMPLEX || out->dtype == DIPAMD_DT_DCOMPLEX)
        return set_error(DIPAMD_E_UNSUPPORTED, "a real result into a complex output image");
    const bool use_double = suggest_flex(in->dtype) == DIPAMD_DT_DFLOAT;
    std::vector<InternFilter> F(static_cast<size_t>(nfilters));
    for (int i = 0; i < nfilters; ++i) {
        if (filters[i].ntaps < 0 || (filters[i].ntaps > 0 && !filters[i].taps))
            return set_error(DIPAMD_E_PARAMETER, "bad filter");
        if ((st = intern_filter(filters[i].taps, filters[i].ntaps, filters[i].symmetry, filters[i].origin, F[i])) !=
            DIPAMD_OK)
            return st;
    }
    const InternFilter *per_dim[DIPAMD_MAX_DIMS];
    bool proc[DIPAMD_MAX_DIMS];
    for (int d = 0; d < nd; ++d) {
        per_dim[d] = nfilters == 1 ? &F[0] : &F[static_cast<size_t>(d)];
        proc[d] = process ? process[d] != 0 : true;
        if (nfilters == 1) {
            if (meaningless(F[0], use_double)) proc[d] = false;
        } else if (in->sizes[d] <= 1 || meaningless(F[static_cast<size_t>(d)], use_double)) {
            proc[d] = false;
        }
        if (in->sizes[d] == 1) proc[d] = false;
    }
    return separable_device(to_view(in), to_view(out), nd, per_dim, bcs, proc, flags,
                            static_cast<hipStream_t>(stream), keep, kept);
}

int dipamd_separable_convolution(const dipamd_image *in, const dipamd_image *out, const dipamd_filter *filters,
                                 int32_t nfilters, const int32_t *bc, int32_t nbc, const uint8_t *process,
                                 uint32_t flags, void *stream) {
    return separable_convolution_impl(in, out, filters, nfilters, bc, nbc, process, flags, stream, nullptr, nullptr);
}

static int gauss_fir_impl(const dipamd_image *in, const dipamd_image *out, const double *sigmas, int32_t nsigmas,
                          const int32_t *orders, int32_t norders, const int32_t *bc, int32_t nbc, double truncation,
                          uint32_t flags, void *stream, const View *keep, bool *kept) {
    int st;
    if ((st = check_image(in, "in")) != DIPAMD_OK) return st;
    const int nd = in->ndim;
    if (nsigmas != 1 && nsigmas != nd)
        return set_error(DIPAMD_E_ARRAY_LENGTH, "Array parameter has the wrong number of elements");
    if (norders != 0 && norders != 1 && norders != nd)
        return set_error(DIPAMD_E_ARRAY_LENGTH, "Array parameter has the wrong number of elements");
    // the taps' precision (gauss.cpp:254): double for dfloat and dcomplex images
    const bool compute_double = in->dtype == DIPAMD_DT_DFLOAT || in->dtype == DIPAMD_DT_DCOMPLEX;
    std::vector<dipamd_filter> fl(static_cast<size_t>(nd));
    std::vector<std::vector<double>> taps(static_cast<size_t>(nd));
    uint8_t proc[DIPAMD_MAX_DIMS];
    for (int d = 0; d < nd; ++d) {
        const double s = nsigmas == 1 ? sigmas[0] : sigmas[d];
        const int o = norders == 0 ? 0 : (norders == 1 ? orders[0] : orders[d]);
        fl[d] = dipamd_filter{nullptr, 0, -1, DIPAMD_SYM_EVEN, 0};
        proc[d] = 0;
        if (s > 0.0 && in->sizes[d] > 1) {
            if (o < 0 || o > 3) return set_error(DIPAMD_E_ORDER, "Gaussian FIR filter not implemented for order > 3");
            const int r = gauss_radius(s, o, truncation, compute_double);
            if (r > DIPAMD_MAX_RADIUS) return set_error(DIPAMD_E_UNSUPPORTED, "Gaussian radius above DIPAMD_MAX_RADIUS");
            taps[d] = half_gaussian(s, o, truncation, compute_double);
            fl[d].taps = taps[d].data();
            fl[d].ntaps = static_cast<int32_t>(taps[d].size());
            fl[d].symmetry = (o & 1) ? DIPAMD_SYM_ODD : DIPAMD_SYM_EVEN;
            proc[d] = 1;
        }
    }
    return separable_convolution_impl(in, out, fl.data(), nd, bc, nbc, proc, flags, stream, keep, kept);
}

int dipamd_gauss_fir(const dipamd_image *in, const dipamd_image *out, const double *sigmas, int32_t nsigmas,
                     const int32_t *orders, int32_t norders, const int32_t *bc, int32_t nbc, double truncation,
                     uint32_t flags, void *stream) {
    return gauss_fir_impl(in, out, sigmas, nsigmas, orders, norders, bc, nbc, truncation, flags, stream, nullptr,
                          nullptr);
}

int dipamd_gauss_fir_keep(const dipamd_image *in, const dipamd_image *out, const double *sigmas, int32_t nsigmas,
                          const int32_t *orders, int32_t norders, const int32_t *bc, int32_t nbc, double truncation,
                          uint32_t flags, const dipamd_image *keep, int32_t *kept, void *stream) {
    int st;
    if (kept) *kept = 0;
    if ((st = check_image(keep, "keep")) != DIPAMD_OK) return st;
    if ((st = check_image(in, "in")) != DIPAMD_OK) return st;
    if (keep->ndim != in->ndim) return set_error(DIPAMD_E_SIZES, "Sizes don't match");
    for (int d = 0; d < in->ndim; ++d)
        if (keep->sizes[d] != in->sizes[d]) return set_error(DIPAMD_E_SIZES, "Sizes don't match");
    if (keep->dtype != suggest_flex(in->dtype))
        return set_error(DIPAMD_E_PARAMETER, "keep must have the buffer type of the input");
    const View kv = to_view(keep);
    if (views_overlap(kv, to_view(in), in->ndim) || (out && out->origin && views_overlap(kv, to_view(out), in->ndim)))
        return set_error(DIPAMD_E_PARAMETER, "keep must not overlap in or out");
    bool k = false;
    st = gauss_fir_impl(in, out, sigmas, nsigmas, orders, norders, bc, nbc, truncation, flags, stream, &kv, &k);
    if (kept) *kept = k ? 1 : 0;
    return st;
}

}  // extern "C"
