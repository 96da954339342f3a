// Batched form: many 2-D images of uneven sizes and per-image sigmas through one
// persistent kernel fed by a work queue (k_batched), optionally replayed from a
// hipGraph.  Each image's result equals dip::Gauss(image, {sx, sy}, {0}, "FIR",
// bc, truncation) bit for bit in the exact mode: the tiles run the same fused code
// as the plain drop-in, and the per-image taps come from the same generator.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "batch_kernels.hpp"
#include "dipamd.h"
#include "dipamd_internal.hpp"
#include "sep_common.hpp"
#include "sep_kernels.hpp"

using namespace dipamd;

namespace {
// Row-chunk rule of the work items: max(minimum rows, rows per unit of y radius).
// Long chunks keep the ring warm-up (2*ry rows x-filtered and discarded per item)
// small; 256 rows or 16 per unit of radius measured best on the 512-image uneven
// batch (2.4 -> 2.2 ms against 64 / 8).  DIPAMD_BATCH_CHUNK="<min>,<per radius>"
// overrides it for experiments.
std::pair<int, int> batch_chunk_rule() {
    static std::pair<int, int> rule = [] {
        std::pair<int, int> r{256, 16};
        if (const char *v = std::getenv("DIPAMD_BATCH_CHUNK")) {
            int a = 0, b = 0;
            if (std::sscanf(v, "%d,%d", &a, &b) == 2 && a >= 16 && b >= 1) r = {a, b};
        }
        return r;
    }();
    return rule;
}
}  // namespace

// Images are grouped by radius class: the x radius rounded up to a multiple of 4 (the
// radius-class tile, up to 24) and the ring depth of the y radius (A = ceil(2 ry / 16)).
// Each group is one persistent launch whose LDS and register use fit its own radii, with
// its own tile config.  Radii beyond the classes run the runtime-radius tile (rxc 0).
struct GroupKey {
    int rxc;  // 4..24, or 0: the runtime-radius tile
    int ya;   // ring depth class of the y radius
    int wc;   // floor(log2) of the group's mean image width (a tuning key, not a grouping key)
    bool operator<(const GroupKey &o) const { return rxc != o.rxc ? rxc < o.rxc : ya < o.ya; }
    bool operator==(const GroupKey &o) const { return rxc == o.rxc && ya == o.ya; }
};

// One persistent launch of a plan: a tile config, its dynamic LDS bytes, a slice of the
// work-item table and the grid.
struct BatchGroup {
    GroupKey key;
    int cfg;
    size_t lds;
    int64_t item0, nitems;
    int grid;
};

struct dipamd_batch_plan {
    int n = 0;
    int in_dtype = 0, out_dtype = 0;
    bool is_double = false;
    bool fast = false;
    int cfg = 1;
    int64_t nitems = 0;
    int bc[2] = {0, 0};
    double truncation = 3;
    uint32_t flags = 0;
    void *d_images = nullptr;
    void *d_items = nullptr;
    void *d_taps = nullptr;
    int32_t *d_counter = nullptr;
    std::vector<dipamd_batch_image> host_images;
    std::vector<BatchGroup> groups;
    std::vector<std::vector<uint8_t>> args;  // BatchArgs<TB> bytes, one per group
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t cap_stream = nullptr;
    // concurrent groups: side streams (default priority) and the fork / join events
    std::vector<hipStream_t> side;
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> join;
};

namespace {

// Counters of the groups sit 64 bytes apart (one per cache line).
constexpr int kCounterStride = 16;
constexpr int kMaxGroups = 16;

int ring_class(int ry) { return (2 * ry + kFusedStep - 1) / kFusedStep; }
// Largest y radius of a ring class.
int ring_class_max_ry(int ya) { return (ya * kFusedStep) / 2; }

bool use_rc_tile(int rx, int ry) { return fused_rc_enabled() && rc_class_of(rx) <= kFusedRcMaxClass && ry <= 24; }

// Cache key of a group's tile config.
std::string group_tune_key(int in_dtype, bool fast, const int bc[2], const GroupKey &k) {
    return "batch in" + std::to_string(in_dtype) + (fast ? " fast" : " exact") + " bc" + std::to_string(bc[0]) + "," +
           std::to_string(bc[1]) + " rxc" + std::to_string(k.rxc) + " ya" + std::to_string(k.ya) + " w" +
           std::to_string(k.wc);
}

// The host side of a plan: per-image descriptors, taps, and per group its work items
// (longest-processing-time order: the longest items first) and launch shape.  No device memory is touched.
// `cfg_of` picks each group's tile config (falling back to one whose LDS fits).
template <typename TB>
struct BatchTables {
    std::vector<BatchImage<TB>> ims;
    std::vector<BatchItem> items;
    std::vector<TB> taps;
    std::vector<BatchGroup> groups;
};

template <typename TB, typename CfgOf>
int plan_tables(const dipamd_batch_image *images, int n, int in_dtype, double truncation, bool fast, CfgOf cfg_of,
                BatchTables<TB> &T) {
    const bool compute_double = in_dtype == DIPAMD_DT_DFLOAT;
    const bool is_double = sizeof(TB) == 8;
    T.ims.assign(static_cast<size_t>(n), BatchImage<TB>{});
    T.items.clear();
    T.taps.clear();
    T.groups.clear();
    const int64_t vb = in_dtype == DIPAMD_DT_UINT8 ? 4 : 16;  // bytes per vector load
    const int64_t es = static_cast<int64_t>(dt_size(in_dtype));
    std::map<GroupKey, std::vector<int>> members;
    for (int i = 0; i < n; ++i) {
        const dipamd_batch_image &b = images[i];
        if (!b.in || !b.out) return set_error(DIPAMD_E_NOT_FORGED, "Image is not forged");
        if (b.width < 1 || b.height < 1) return set_error(DIPAMD_E_PARAMETER, "Sizes must be positive");
        if (b.order_x != 0 || b.order_y != 0)
            return set_error(DIPAMD_E_UNSUPPORTED, "the batched form takes derivative order 0");
        auto axis = [&](double s, int size, int *r) -> std::vector<double> {
            // a dim of size 1 or sigma <= 0 is not filtered (GaussFIR): identity tap
            if (!(s > 0.0) || size <= 1) {
                *r = 0;
                return {1.0};
            }
            std::vector<double> h = half_gaussian(s, 0, truncation, compute_double);
            *r = static_cast<int>(h.size()) - 1;
            std::reverse(h.begin(), h.end());  // correlation order: centre first
            return h;
        };
        int rx = 0, ry = 0;
        std::vector<double> hx = axis(b.sigma_x, b.width, &rx);
        std::vector<double> hy = axis(b.sigma_y, b.height, &ry);
        if (rx > kFusedCap - 1 || ry > kFusedCap - 1)
            return set_error(DIPAMD_E_UNSUPPORTED, "batched form: radius above 64 (sigma too large)");
        BatchImage<TB> &im = T.ims[static_cast<size_t>(i)];
        im.in = b.in;
        im.out = static_cast<TB *>(b.out);
        im.nx = b.width;
        im.ny = b.height;
        im.syin = b.in_stride_y;
        im.syout = b.out_stride_y;
        im.rx = rx;
        im.ry = ry;
        im.tap_offset = static_cast<int32_t>(T.taps.size());
        im.vec_ok = (reinterpret_cast<uintptr_t>(b.in) % static_cast<uintptr_t>(vb) == 0 &&
                     (b.in_stride_y * es) % vb == 0)
                        ? 1
                        : 0;
        for (double v : hx) T.taps.push_back(static_cast<TB>(v));
        for (double v : hy) T.taps.push_back(static_cast<TB>(v));
        const GroupKey key = use_rc_tile(rx, ry) ? GroupKey{rc_class_of(rx), ring_class(ry), 0} : GroupKey{0, 0, 0};
        members[key].push_back(i);
    }
    if (members.size() > static_cast<size_t>(kMaxGroups))
        return set_error(DIPAMD_E_UNSUPPORTED, "batched form: too many launch groups");
    int dev = 0, ncu = 256;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    const size_t limit = lds_limit_bytes();
    const InType ti = in_dtype == DIPAMD_DT_UINT8 ? IN_U8 : IN_BUF;
    // groups with the largest radii first: their items are the longest
    for (auto it = members.rbegin(); it != members.rend(); ++it) {
        GroupKey key = it->first;
        std::vector<int> &mem = it->second;
        int rmax_x = 0, rmax_y = 0;
        double wsum = 0;
        for (int i : mem) {
            rmax_x = std::max(rmax_x, T.ims[static_cast<size_t>(i)].rx);
            rmax_y = std::max(rmax_y, T.ims[static_cast<size_t>(i)].ry);
            wsum += static_cast<double>(T.ims[static_cast<size_t>(i)].nx);
        }
        key.wc = std::max(0, static_cast<int>(std::floor(std::log2(std::max(1.0, wsum / mem.size())))));
        // the LDS of a group is sized for its ring class, so one config serves the class
        const int ry_lds = key.rxc > 0 ? ring_class_max_ry(key.ya) : rmax_y;
        auto lds_of = [&](int c) {
            return key.rxc > 0 ? fused_rc_lds_bytes(c, key.rxc, ry_lds, is_double)
                               : fused_lds_bytes(c, rmax_x, rmax_y, is_double);
        };
        auto usable = [&](int c) { return c >= 0 && c < kNumFusedConfigs && !(is_double && c == 2) && lds_of(c) <= limit; };
        int cfg = cfg_of(key);
        if (!usable(cfg)) {
            cfg = -1;
            for (int c = kNumFusedConfigs - 1; c >= 0; --c)
                if (usable(c)) {
                    cfg = c;
                    break;
                }
            if (cfg < 0) return set_error(DIPAMD_E_UNSUPPORTED, "batched form: tile does not fit in LDS");
        }
        const int txc = kFusedConfigs[cfg].txc;
        std::stable_sort(mem.begin(), mem.end(), [&](int a, int b) {
            const auto &A = T.ims[static_cast<size_t>(a)];
            const auto &B = T.ims[static_cast<size_t>(b)];
            return (A.rx + A.ry + 2) > (B.rx + B.ry + 2);
        });
        BatchGroup g;
        g.key = key;
        g.cfg = cfg;
        g.lds = lds_of(cfg);
        g.item0 = static_cast<int64_t>(T.items.size());
        for (int i : mem) {
            const auto &im = T.ims[static_cast<size_t>(i)];
            const int ntiles = static_cast<int>((im.nx + txc - 1) / txc);
            // row chunks: long enough that the y halo stays small, short enough to balance
            int ch = std::max(batch_chunk_rule().first, ((batch_chunk_rule().second * im.ry + 15) / 16) * 16);
            ch = std::min<int>(ch, static_cast<int>(((im.ny + 15) / 16) * 16));
            for (int64_t z = 0; z < im.ny; z += ch)
                for (int t = 0; t < ntiles; ++t)
                    T.items.push_back(
                        {i, t, static_cast<int32_t>(z), static_cast<int32_t>(std::min<int64_t>(z + ch, im.ny))});
        }
        g.nitems = static_cast<int64_t>(T.items.size()) - g.item0;
        // longest items first (the persistent workgroups take them in table order, so the
        // short ones fill the gaps at the end): cost ~ (x-filtered rows + output rows) x
        // operations per sample
        std::stable_sort(T.items.begin() + g.item0, T.items.end(), [&](const BatchItem &A, const BatchItem &B) {
            auto cost = [&](const BatchItem &it) {
                const auto &im = T.ims[static_cast<size_t>(it.image)];
                return static_cast<int64_t>(2 * (it.z1 - it.z0) + 2 * im.ry) * (1 + 3 * std::max(im.rx, im.ry));
            };
            return cost(A) > cost(B);
        });
        if (T.items.size() > static_cast<size_t>(INT32_MAX))
            return set_error(DIPAMD_E_UNSUPPORTED, "batched form: too many work items");
        int occ = 1;
        if (key.rxc > 0) {
            if (launch_batched_rc(ti, is_double, cfg, fast, key.rxc, nullptr, 0, g.lds, nullptr, &occ) != hipSuccess)
                occ = 1;
        } else {
            occ = batched_occupancy(ti, is_double, cfg, fast, g.lds);
        }
        g.grid = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(static_cast<int64_t>(ncu) * occ, g.nitems)));
        T.groups.push_back(g);
    }
    return DIPAMD_OK;
}

// Kernel arguments of group `gi` over device tables at the given addresses.
template <typename TB>
std::vector<uint8_t> group_args(const BatchTables<TB> &T, size_t gi, const void *d_images, const void *d_items,
                                const void *d_taps, int32_t *d_counter, const int bc[2]) {
    const BatchGroup &g = T.groups[gi];
    BatchArgs<TB> a;
    std::memset(&a, 0, sizeof(a));
    a.images = static_cast<const BatchImage<TB> *>(d_images);
    a.items = static_cast<const BatchItem *>(d_items) + g.item0;
    a.taps = static_cast<const TB *>(d_taps);
    a.counter = d_counter + gi * kCounterStride;
    a.nitems = static_cast<int32_t>(g.nitems);
    a.bcx = bc[0];
    a.bcy = bc[1];
    auto cv = [](int b) -> TB {
        if (b == DIPAMD_BC_ADD_MAX_VALUE) return std::numeric_limits<TB>::max();
        if (b == DIPAMD_BC_ADD_MIN_VALUE) return std::numeric_limits<TB>::lowest();
        return TB(0);
    };
    a.cvalx = cv(bc[0]);
    a.cvaly = cv(bc[1]);
    std::vector<uint8_t> bytes(sizeof(a));
    std::memcpy(bytes.data(), &a, sizeof(a));
    return bytes;
}

template <typename TB, typename CfgOf>
int build_plan(dipamd_batch_plan *P, const dipamd_batch_image *images, int n, CfgOf cfg_of) {
    BatchTables<TB> T;
    int st = plan_tables<TB>(images, n, P->in_dtype, P->truncation, P->fast, cfg_of, T);
    if (st != DIPAMD_OK) return st;
    P->cfg = T.groups[0].cfg;
    P->nitems = static_cast<int64_t>(T.items.size());
    P->groups = T.groups;
    // device tables
    hipError_t e;
    if ((e = hipMalloc(&P->d_images, T.ims.size() * sizeof(BatchImage<TB>))) != hipSuccess ||
        (e = hipMalloc(&P->d_items, std::max<size_t>(1, T.items.size()) * sizeof(BatchItem))) != hipSuccess ||
        (e = hipMalloc(&P->d_taps, T.taps.size() * sizeof(TB))) != hipSuccess ||
        (e = hipMalloc(reinterpret_cast<void **>(&P->d_counter), kMaxGroups * kCounterStride * sizeof(int32_t))) !=
            hipSuccess)
        return set_error(DIPAMD_E_HIP, std::string("batch plan allocation: ") + hipGetErrorString(e));
    if ((e = hipMemcpy(P->d_images, T.ims.data(), T.ims.size() * sizeof(BatchImage<TB>), hipMemcpyHostToDevice)) !=
            hipSuccess ||
        (e = hipMemcpy(P->d_items, T.items.data(), T.items.size() * sizeof(BatchItem), hipMemcpyHostToDevice)) !=
            hipSuccess ||
        (e = hipMemcpy(P->d_taps, T.taps.data(), T.taps.size() * sizeof(TB), hipMemcpyHostToDevice)) != hipSuccess)
        return set_error(DIPAMD_E_HIP, std::string("batch plan upload: ") + hipGetErrorString(e));
    P->args.clear();
    for (size_t gi = 0; gi < T.groups.size(); ++gi)
        P->args.push_back(group_args<TB>(T, gi, P->d_images, P->d_items, P->d_taps, P->d_counter, P->bc));
    return DIPAMD_OK;
}

void free_plan_buffers(dipamd_batch_plan *P) {
    if (P->exec) (void)hipGraphExecDestroy(P->exec);
    if (P->graph) (void)hipGraphDestroy(P->graph);
    if (P->cap_stream) (void)hipStreamDestroy(P->cap_stream);
    P->exec = nullptr;
    P->graph = nullptr;
    P->cap_stream = nullptr;
    for (hipStream_t sk : P->side) (void)hipStreamDestroy(sk);
    for (hipEvent_t ek : P->join) (void)hipEventDestroy(ek);
    if (P->fork) (void)hipEventDestroy(P->fork);
    P->side.clear();
    P->join.clear();
    P->fork = nullptr;
    if (P->d_images) (void)hipFree(P->d_images);
    if (P->d_items) (void)hipFree(P->d_items);
    if (P->d_taps) (void)hipFree(P->d_taps);
    if (P->d_counter) (void)hipFree(P->d_counter);
    P->d_images = P->d_items = P->d_taps = nullptr;
    P->d_counter = nullptr;
}

template <typename CfgOf>
int make_plan(dipamd_batch_plan *P, CfgOf cfg_of) {
    return P->is_double ? build_plan<double>(P, P->host_images.data(), P->n, cfg_of)
                        : build_plan<float>(P, P->host_images.data(), P->n, cfg_of);
}

// One group's launch (its counter cleared first).
hipError_t launch_group(const BatchGroup &g, const std::vector<uint8_t> &args, int32_t *counter, int in_dtype,
                        bool is_double, bool fast, hipStream_t stream) {
    const InType ti = in_dtype == DIPAMD_DT_UINT8 ? IN_U8 : IN_BUF;
    if (g.key.rxc > 0) return launch_batched_rc(ti, is_double, g.cfg, fast, g.key.rxc, args.data(), g.grid, g.lds, stream, nullptr);
    return launch_batched(ti, is_double, g.cfg, fast, args.data(), g.grid, g.lds, stream);
}

// The groups' launches on `stream`: counters cleared by a kernel (no memset node
// under graph capture), then one persistent launch per group.
int launch_groups(const std::vector<BatchGroup> &groups, const std::vector<std::vector<uint8_t>> &args,
                  int32_t *d_counter, int in_dtype, bool is_double, bool fast, hipStream_t stream) {
    hipError_t e = launch_reset_counter(d_counter, static_cast<int>(groups.size()) * kCounterStride, stream);
    for (size_t gi = 0; e == hipSuccess && gi < groups.size(); ++gi)
        e = launch_group(groups[gi], args[gi], d_counter, in_dtype, is_double, fast, stream);
    if (e != hipSuccess) return set_error(DIPAMD_E_HIP, std::string("batched launch: ") + hipGetErrorString(e));
    return DIPAMD_OK;
}

// Side streams for concurrent groups: DIPAMD_BATCH_STREAMS (default 3, besides the
// caller's; 0 runs the groups one after another).
int batch_side_streams() {
    static int n = [] {
        const char *v = std::getenv("DIPAMD_BATCH_STREAMS");
        const int k = v ? std::atoi(v) : 3;
        return std::max(0, std::min(k, 8));
    }();
    return n;
}

// The plan's groups on `stream`.  With several groups and side streams, the groups run
// concurrently: the counters are cleared on `stream`, every side stream waits for that,
// group g goes to stream (g mod (1 + sides)) -- 0 being the caller's --, and `stream`
// waits for every side stream's last group.  A persistent group that runs out of items
// frees its CUs for the others, so one group's tail overlaps the next group's work.
int run_plan(dipamd_batch_plan *P, hipStream_t stream) {
    const int ns = batch_side_streams();
    if (P->groups.size() < 2 || ns == 0)
        return launch_groups(P->groups, P->args, P->d_counter, P->in_dtype, P->is_double, P->fast, stream);
    hipError_t e = hipSuccess;
    const int nside = std::min<int>(ns, static_cast<int>(P->groups.size()) - 1);
    if (P->side.empty()) {
        for (int k = 0; k < nside && e == hipSuccess; ++k) {
            hipStream_t sk = nullptr;
            e = hipStreamCreateWithFlags(&sk, hipStreamNonBlocking);
            if (e == hipSuccess) P->side.push_back(sk);
        }
        if (e == hipSuccess) e = hipEventCreateWithFlags(&P->fork, hipEventDisableTiming);
        for (int k = 0; k < nside && e == hipSuccess; ++k) {
            hipEvent_t ek = nullptr;
            e = hipEventCreateWithFlags(&ek, hipEventDisableTiming);
            if (e == hipSuccess) P->join.push_back(ek);
        }
        if (e != hipSuccess) return set_error(DIPAMD_E_HIP, std::string("batch streams: ") + hipGetErrorString(e));
    }
    e = launch_reset_counter(P->d_counter, static_cast<int>(P->groups.size()) * kCounterStride, stream);
    if (e == hipSuccess) e = hipEventRecord(P->fork, stream);
    for (size_t k = 0; e == hipSuccess && k < P->side.size(); ++k) e = hipStreamWaitEvent(P->side[k], P->fork, 0);
    const size_t nq = P->side.size() + 1;
    for (size_t gi = 0; e == hipSuccess && gi < P->groups.size(); ++gi) {
        const size_t q = gi % nq;
        e = launch_group(P->groups[gi], P->args[gi], P->d_counter, P->in_dtype, P->is_double, P->fast,
                         q == 0 ? stream : P->side[q - 1]);
    }
    for (size_t k = 0; e == hipSuccess && k < P->side.size(); ++k) {
        e = hipEventRecord(P->join[k], P->side[k]);
        if (e == hipSuccess) e = hipStreamWaitEvent(stream, P->join[k], 0);
    }
    if (e != hipSuccess) return set_error(DIPAMD_E_HIP, std::string("batched launch: ") + hipGetErrorString(e));
    return DIPAMD_OK;
}

// ---------------------------------------------------------------------------
// The entry point that allocates nothing: the plan's tables and counters go to a
// file-scope device array (one copy per GPU), whose device address is looked up
// once per device.
// ---------------------------------------------------------------------------
constexpr size_t kScratchBytes = size_t(8) << 20;  // DIPAMD_BATCH_NOALLOC_SCRATCH_BYTES
static_assert(kScratchBytes == DIPAMD_BATCH_NOALLOC_SCRATCH_BYTES, "scratch size and header disagree");
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::once_flag g_scratch_once[kMaxDevices];
void *g_scratch_ptr[kMaxDevices];
hipError_t g_scratch_err[kMaxDevices];

int scratch_address(void **ptr) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return set_error(DIPAMD_E_HIP, std::string("hipGetDevice: ") + hipGetErrorString(e));
    if (dev < 0 || dev >= kMaxDevices) return set_error(DIPAMD_E_UNSUPPORTED, "device index out of range");
    std::call_once(g_scratch_once[dev], [dev]() {
        void *p = nullptr;
        g_scratch_err[dev] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        g_scratch_ptr[dev] = g_scratch_err[dev] == hipSuccess ? p : nullptr;
    });
    if (g_scratch_err[dev] != hipSuccess || !g_scratch_ptr[dev])
        return set_error(DIPAMD_E_HIP, std::string("scratch symbol address: ") + hipGetErrorString(g_scratch_err[dev]));
    *ptr = g_scratch_ptr[dev];
    return DIPAMD_OK;
}

size_t align256(size_t v) { return (v + 255) & ~size_t(255); }

template <typename TB, typename CfgOf>
int run_noalloc(const dipamd_batch_image *images, int n, int in_dtype, const int bc[2], double truncation, bool fast,
                CfgOf cfg_of, hipStream_t stream) {
    BatchTables<TB> T;
    int st = plan_tables<TB>(images, n, in_dtype, truncation, fast, cfg_of, T);
    if (st != DIPAMD_OK) return st;
    if (T.groups.size() > static_cast<size_t>(kMaxGroups))
        return set_error(DIPAMD_E_UNSUPPORTED, "batched form: too many launch groups");
    // layout: counters, image descriptors, work items, taps (each 256-byte aligned)
    const size_t off_counter = 0;
    const size_t off_images = align256(kMaxGroups * kCounterStride * sizeof(int32_t));
    const size_t off_items = align256(off_images + T.ims.size() * sizeof(BatchImage<TB>));
    const size_t off_taps = align256(off_items + T.items.size() * sizeof(BatchItem));
    const size_t total = off_taps + T.taps.size() * sizeof(TB);
    if (total > kScratchBytes)
        return set_error(DIPAMD_E_UNSUPPORTED,
                         "batched form without allocation: the batch needs " + std::to_string(total) +
                             " bytes of tables, above the fixed scratch of " + std::to_string(kScratchBytes) +
                             " bytes (use dipamd_batch_create)");
    void *base = nullptr;
    if ((st = scratch_address(&base)) != DIPAMD_OK) return st;
    unsigned char *b = static_cast<unsigned char *>(base);
    // host tables are pageable memory: these copies complete before the calls return
    hipError_t e;
    if ((e = hipMemcpyAsync(b + off_images, T.ims.data(), T.ims.size() * sizeof(BatchImage<TB>),
                            hipMemcpyHostToDevice, stream)) != hipSuccess ||
        (e = hipMemcpyAsync(b + off_items, T.items.data(), T.items.size() * sizeof(BatchItem), hipMemcpyHostToDevice,
                            stream)) != hipSuccess ||
        (e = hipMemcpyAsync(b + off_taps, T.taps.data(), T.taps.size() * sizeof(TB), hipMemcpyHostToDevice, stream)) !=
            hipSuccess)
        return set_error(DIPAMD_E_HIP, std::string("batch table upload: ") + hipGetErrorString(e));
    int32_t *counters = reinterpret_cast<int32_t *>(b + off_counter);
    std::vector<std::vector<uint8_t>> args;
    for (size_t gi = 0; gi < T.groups.size(); ++gi)
        args.push_back(group_args<TB>(T, gi, b + off_images, b + off_items, b + off_taps, counters, bc));
    return launch_groups(T.groups, args, counters, in_dtype, sizeof(TB) == 8, fast, stream);
}

// Arguments shared by dipamd_batch_create and dipamd_batch_gauss_noalloc.
int check_batch(const dipamd_batch_image *images, int32_t n, int32_t in_dtype, int32_t out_dtype, const int32_t *bc,
                int32_t nbc, int *bcs) {
    if (n < 1 || !images) return set_error(DIPAMD_E_PARAMETER, "empty batch");
    if (in_dtype != DIPAMD_DT_UINT8 && in_dtype != DIPAMD_DT_SFLOAT && in_dtype != DIPAMD_DT_DFLOAT)
        return set_error(DIPAMD_E_UNSUPPORTED, "batched form: input must be uint8, sfloat or dfloat");
    if (out_dtype != suggest_flex(in_dtype))
        return set_error(DIPAMD_E_UNSUPPORTED, "batched form: output type must be SuggestFlex(input)");
    int st = broadcast_bc(bc, nbc, 2, bcs);
    if (st != DIPAMD_OK) return st;
    for (int k = 0; k < 2; ++k)
        if (!bc_supported(bcs[k])) return set_error(DIPAMD_E_UNSUPPORTED, "boundary condition not supported");
    // The tiles of one image run concurrently and read halo rows that other tiles
    // write: an image's output must not share memory with its input (the per-image
    // loop would handle it through a scratch copy; the persistent kernel cannot).
    for (int i = 0; i < n; ++i) {
        const dipamd_batch_image &b = images[i];
        if (!b.in || !b.out || b.width < 1 || b.height < 1) continue;  // reported by the planner
        View vi, vo;
        vi.p = const_cast<void *>(b.in);
        vi.dt = in_dtype;
        vo.p = b.out;
        vo.dt = out_dtype;
        vi.sizes[0] = vo.sizes[0] = b.width;
        vi.sizes[1] = vo.sizes[1] = b.height;
        vi.strides[0] = vo.strides[0] = 1;
        vi.strides[1] = b.in_stride_y;
        vo.strides[1] = b.out_stride_y;
        if (views_overlap(vi, vo, 2))
            return set_error(DIPAMD_E_PARAMETER,
                             "batched form: image " + std::to_string(i) + " has an output that overlaps its input");
    }
    return DIPAMD_OK;
}

}  // namespace

extern "C" {

int dipamd_batch_num_configs(void) { return kNumFusedConfigs; }

int dipamd_batch_create(dipamd_batch_plan **plan, const dipamd_batch_image *images, int32_t n, int32_t in_dtype,
                        int32_t out_dtype, const int32_t *bc, int32_t nbc, double truncation, uint32_t flags,
                        int32_t tile_config, void *stream) {
    if (!plan) return set_error(DIPAMD_E_PARAMETER, "null plan pointer");
    *plan = nullptr;
    int bcs[2];
    int st = check_batch(images, n, in_dtype, out_dtype, bc, nbc, bcs);
    if (st != DIPAMD_OK) return st;
    if (tile_config >= kNumFusedConfigs) return set_error(DIPAMD_E_PARAMETER, "tile config out of range");
    auto *P = new dipamd_batch_plan();
    P->n = n;
    P->in_dtype = in_dtype;
    P->out_dtype = out_dtype;
    P->is_double = out_dtype == DIPAMD_DT_DFLOAT;
    P->fast = (flags & DIPAMD_FAST) != 0;
    P->flags = flags;
    P->bc[0] = bcs[0];
    P->bc[1] = bcs[1];
    P->truncation = truncation;
    P->host_images.assign(images, images + n);
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (tile_config >= 0) {
        const int c = (P->is_double && tile_config == 2) ? 1 : tile_config;
        st = make_plan(P, [c](const GroupKey &) { return c; });
    } else {
        // each group's tile config: the persisted measurement where there is one, else
        // measured now on this batch (every config, each group timed alone on the
        // caller's stream) and persisted
        std::map<GroupKey, int> chosen;
        std::vector<GroupKey> keys;
        if ((st = make_plan(P, [](const GroupKey &) { return 1; })) == DIPAMD_OK)
            for (const BatchGroup &g : P->groups) keys.push_back(g.key);
        free_plan_buffers(P);
        std::vector<GroupKey> todo;
        for (const GroupKey &k : keys) {
            int v = -1;
            if (tune_get(group_tune_key(in_dtype, P->fast, bcs, k), &v) && v >= 0 && v < kNumFusedConfigs)
                chosen[k] = v;
            else
                todo.push_back(k);
        }
        if (st == DIPAMD_OK && !todo.empty()) {
            std::map<GroupKey, float> best;
            hipEvent_t e0, e1;
            (void)hipEventCreate(&e0);
            (void)hipEventCreate(&e1);
            for (int c = 0; c < kNumFusedConfigs; ++c) {
                if (P->is_double && c == 2) continue;
                if (make_plan(P, [c](const GroupKey &) { return c; }) != DIPAMD_OK) {
                    free_plan_buffers(P);
                    continue;
                }
                for (size_t gi = 0; gi < P->groups.size(); ++gi) {
                    const BatchGroup &g = P->groups[gi];
                    if (g.cfg != c || std::find(todo.begin(), todo.end(), g.key) == todo.end()) continue;
                    int32_t *ctr = P->d_counter + gi * kCounterStride;
                    float tot = 0, ms = 0;
                    bool ok = true;
                    for (int rep = 0; ok && rep < 3; ++rep) {  // the first run warms up
                        (void)hipEventRecord(e0, s);
                        ok = launch_reset_counter(ctr, 1, s) == hipSuccess &&
                             launch_group(g, P->args[gi], ctr, P->in_dtype, P->is_double, P->fast, s) == hipSuccess;
                        (void)hipEventRecord(e1, s);
                        (void)hipEventSynchronize(e1);
                        (void)hipEventElapsedTime(&ms, e0, e1);
                        if (rep > 0) tot += ms;
                    }
                    auto it = best.find(g.key);
                    if (ok && (it == best.end() || tot < it->second)) {
                        best[g.key] = tot;
                        chosen[g.key] = c;
                    }
                }
                free_plan_buffers(P);
            }
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            for (const GroupKey &k : todo)
                if (chosen.count(k)) tune_put(group_tune_key(in_dtype, P->fast, bcs, k), chosen[k]);
        }
        st = make_plan(P, [&chosen](const GroupKey &k) {
            auto it = chosen.find(k);
            return it != chosen.end() ? it->second : 1;
        });
    }
    if (st != DIPAMD_OK) {
        free_plan_buffers(P);
        delete P;
        return st;
    }
    *plan = P;
    return DIPAMD_OK;
}

int dipamd_batch_run(dipamd_batch_plan *plan, void *stream) {
    if (!plan) return set_error(DIPAMD_E_PARAMETER, "null plan");
    return run_plan(plan, static_cast<hipStream_t>(stream));
}

int dipamd_batch_run_graph(dipamd_batch_plan *plan, void *stream) {
    if (!plan) return set_error(DIPAMD_E_PARAMETER, "null plan");
    hipError_t e;
    if (!plan->exec) {
        if ((e = hipStreamCreateWithFlags(&plan->cap_stream, hipStreamNonBlocking)) != hipSuccess)
            return set_error(DIPAMD_E_HIP, std::string("capture stream: ") + hipGetErrorString(e));
        if ((e = hipStreamBeginCapture(plan->cap_stream, hipStreamCaptureModeThreadLocal)) != hipSuccess)
            return set_error(DIPAMD_E_HIP, std::string("begin capture: ") + hipGetErrorString(e));
        const int st = run_plan(plan, plan->cap_stream);
        e = hipStreamEndCapture(plan->cap_stream, &plan->graph);
        if (st != DIPAMD_OK) return st;
        if (e != hipSuccess) return set_error(DIPAMD_E_HIP, std::string("end capture: ") + hipGetErrorString(e));
        if ((e = hipGraphInstantiate(&plan->exec, plan->graph, nullptr, nullptr, 0)) != hipSuccess)
            return set_error(DIPAMD_E_HIP, std::string("graph instantiate: ") + hipGetErrorString(e));
    }
    if ((e = hipGraphLaunch(plan->exec, static_cast<hipStream_t>(stream))) != hipSuccess)
        return set_error(DIPAMD_E_HIP, std::string("graph launch: ") + hipGetErrorString(e));
    return DIPAMD_OK;
}

int dipamd_batch_run_per_image(dipamd_batch_plan *plan, void *stream) {
    if (!plan) return set_error(DIPAMD_E_PARAMETER, "null plan");
    for (const dipamd_batch_image &b : plan->host_images) {
        dipamd_image in{}, out{};
        in.origin = const_cast<void *>(b.in);
        in.dtype = plan->in_dtype;
        in.ndim = 2;
        in.sizes[0] = b.width;
        in.sizes[1] = b.height;
        in.strides[0] = 1;
        in.strides[1] = b.in_stride_y;
        out.origin = b.out;
        out.dtype = plan->out_dtype;
        out.ndim = 2;
        out.sizes[0] = b.width;
        out.sizes[1] = b.height;
        out.strides[0] = 1;
        out.strides[1] = b.out_stride_y;
        const double sig[2] = {b.sigma_x, b.sigma_y};
        const int32_t ord[2] = {b.order_x, b.order_y};
        const int st = dipamd_gauss_fir(&in, &out, sig, 2, ord, 2, plan->bc, 2, plan->truncation, plan->flags, stream);
        if (st != DIPAMD_OK) return st;
    }
    return DIPAMD_OK;
}

int dipamd_batch_gauss_noalloc(const dipamd_batch_image *images, int32_t n, int32_t in_dtype, int32_t out_dtype,
                               const int32_t *bc, int32_t nbc, double truncation, uint32_t flags,
                               int32_t tile_config, void *stream) {
    int bcs[2];
    int st = check_batch(images, n, in_dtype, out_dtype, bc, nbc, bcs);
    if (st != DIPAMD_OK) return st;
    if (tile_config >= kNumFusedConfigs) return set_error(DIPAMD_E_PARAMETER, "tile config out of range");
    const bool is_double = out_dtype == DIPAMD_DT_DFLOAT;
    const bool fast = (flags & DIPAMD_FAST) != 0;
    // each group's tile config: the caller's, or the persisted measurement, or the default
    auto cfg_of = [&](const GroupKey &k) {
        if (tile_config >= 0) return (is_double && tile_config == 2) ? 1 : static_cast<int>(tile_config);
        int v = -1;
        if (tune_get(group_tune_key(in_dtype, fast, bcs, k), &v) && v >= 0 && v < kNumFusedConfigs) return v;
        return 1;
    };
    hipStream_t s = static_cast<hipStream_t>(stream);
    return is_double ? run_noalloc<double>(images, n, in_dtype, bcs, truncation, fast, cfg_of, s)
                     : run_noalloc<float>(images, n, in_dtype, bcs, truncation, fast, cfg_of, s);
}

int dipamd_batch_destroy(dipamd_batch_plan *plan) {
    if (!plan) return DIPAMD_OK;
    free_plan_buffers(plan);
    delete plan;
    return DIPAMD_OK;
}

int dipamd_batch_info(const dipamd_batch_plan *plan, int64_t *ntiles, int32_t *config) {
    if (!plan) return set_error(DIPAMD_E_PARAMETER, "null plan");
    if (ntiles) *ntiles = plan->nitems;
    if (config) *config = plan->cfg;
    return DIPAMD_OK;
}

}  // extern "C"
