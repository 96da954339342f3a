// Batched form: many 2-D images of uneven sizes and per-image sigmas through one
// persistent kernel fed by a work queue (k_batched), optionally replayed from a
// hipGraph.  Each image's result equals dip::Gauss(image, {sx, sy}, {0}, "FIR",
// bc, truncation) bit for bit in the exact mode: the tiles run the same fused code
// as the plain drop-in, and the per-image taps come from the same generator.
#include <hip/hip_runtime.h>

#include <algorithm>
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

// One persistent launch of a plan: a tile config, its dynamic LDS bytes, a slice of the
// work-item table and the grid.
struct BatchGroup {
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
};

namespace {

std::mutex g_tune_mutex;
std::map<std::tuple<int, int, int, int>, int> g_batch_tuned;  // (in dtype, fast, rmax, bc) -> cfg

// Counters of the groups sit 64 bytes apart (one per cache line).
constexpr int kCounterStride = 16;
constexpr int kMaxGroups = 16;

// The host side of a plan: per-image descriptors, work items (longest-processing-time
// order), taps, and the launch groups.  No device memory is touched.
template <typename TB>
struct BatchTables {
    std::vector<BatchImage<TB>> ims;
    std::vector<BatchItem> items;
    std::vector<TB> taps;
    std::vector<BatchGroup> groups;
};

template <typename TB>
int plan_tables(const dipamd_batch_image *images, int n, int in_dtype, double truncation, bool fast, int cfg,
                BatchTables<TB> &T) {
    const bool compute_double = in_dtype == DIPAMD_DT_DFLOAT;
    T.ims.assign(static_cast<size_t>(n), BatchImage<TB>{});
    T.items.clear();
    T.taps.clear();
    T.groups.clear();
    int rmax_x = 0, rmax_y = 0;
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
        im.pad_ = 0;
        for (double v : hx) T.taps.push_back(static_cast<TB>(v));
        for (double v : hy) T.taps.push_back(static_cast<TB>(v));
        rmax_x = std::max(rmax_x, rx);
        rmax_y = std::max(rmax_y, ry);
    }
    // a tile config whose LDS (staging + ring for the largest radii) fits; the
    // requested one when it does
    const size_t limit = lds_limit_bytes();
    if (fused_lds_bytes(cfg, rmax_x, rmax_y, sizeof(TB) == 8) > limit) {
        int fit = -1;
        for (int c = kNumFusedConfigs - 1; c >= 0; --c)
            if (!(sizeof(TB) == 8 && c == 2) && fused_lds_bytes(c, rmax_x, rmax_y, sizeof(TB) == 8) <= limit) {
                fit = c;
                break;
            }
        if (fit < 0) return set_error(DIPAMD_E_UNSUPPORTED, "batched form: tile does not fit in LDS");
        cfg = fit;
    }
    // work items, costliest images first (longest-processing-time order)
    const int txc = kFusedConfigs[cfg].txc;
    std::vector<int> order(static_cast<size_t>(n));
    for (int i = 0; i < n; ++i) order[static_cast<size_t>(i)] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        const auto &A = T.ims[static_cast<size_t>(a)];
        const auto &B = T.ims[static_cast<size_t>(b)];
        return (A.rx + A.ry + 2) > (B.rx + B.ry + 2);
    });
    for (int i : order) {
        const auto &im = T.ims[static_cast<size_t>(i)];
        const int ntiles = static_cast<int>((im.nx + txc - 1) / txc);
        // row chunks: long enough that the y halo stays small, short enough to balance
        int ch = std::max(batch_chunk_rule().first, ((batch_chunk_rule().second * im.ry + 15) / 16) * 16);
        ch = std::min<int>(ch, static_cast<int>(((im.ny + 15) / 16) * 16));
        for (int64_t z = 0; z < im.ny; z += ch)
            for (int t = 0; t < ntiles; ++t)
                T.items.push_back({i, t, static_cast<int32_t>(z), static_cast<int32_t>(std::min<int64_t>(z + ch, im.ny))});
    }
    if (T.items.size() > static_cast<size_t>(INT32_MAX))
        return set_error(DIPAMD_E_UNSUPPORTED, "batched form: too many work items");
    BatchGroup g;
    g.cfg = cfg;
    g.lds = fused_lds_bytes(cfg, rmax_x, rmax_y, sizeof(TB) == 8);
    g.item0 = 0;
    g.nitems = static_cast<int64_t>(T.items.size());
    int dev = 0, ncu = 256;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    const InType ti = in_dtype == DIPAMD_DT_UINT8 ? IN_U8 : IN_BUF;
    const int occ = batched_occupancy(ti, sizeof(TB) == 8, g.cfg, fast, g.lds);
    g.grid = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(static_cast<int64_t>(ncu) * occ, g.nitems)));
    T.groups.push_back(g);
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

template <typename TB>
int build_plan(dipamd_batch_plan *P, const dipamd_batch_image *images, int n) {
    BatchTables<TB> T;
    int st = plan_tables<TB>(images, n, P->in_dtype, P->truncation, P->fast, P->cfg, T);
    if (st != DIPAMD_OK) return st;
    if (T.groups.size() > static_cast<size_t>(kMaxGroups))
        return set_error(DIPAMD_E_UNSUPPORTED, "batched form: too many launch groups");
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
    if (P->d_images) (void)hipFree(P->d_images);
    if (P->d_items) (void)hipFree(P->d_items);
    if (P->d_taps) (void)hipFree(P->d_taps);
    if (P->d_counter) (void)hipFree(P->d_counter);
    P->d_images = P->d_items = P->d_taps = nullptr;
    P->d_counter = nullptr;
}

int make_plan(dipamd_batch_plan *P, int cfg) {
    P->cfg = cfg;
    return P->is_double ? build_plan<double>(P, P->host_images.data(), P->n)
                        : build_plan<float>(P, P->host_images.data(), P->n);
}

// The groups' launches on `stream`: counters cleared by a kernel (no memset node
// under graph capture), then one persistent launch per group.
int launch_groups(const std::vector<BatchGroup> &groups, const std::vector<std::vector<uint8_t>> &args,
                  int32_t *d_counter, int in_dtype, bool is_double, bool fast, hipStream_t stream) {
    hipError_t e = launch_reset_counter(d_counter, static_cast<int>(groups.size()) * kCounterStride, stream);
    const InType ti = in_dtype == DIPAMD_DT_UINT8 ? IN_U8 : IN_BUF;
    for (size_t gi = 0; e == hipSuccess && gi < groups.size(); ++gi)
        e = launch_batched(ti, is_double, groups[gi].cfg, fast, args[gi].data(), groups[gi].grid, groups[gi].lds,
                           stream);
    if (e != hipSuccess) return set_error(DIPAMD_E_HIP, std::string("batched launch: ") + hipGetErrorString(e));
    return DIPAMD_OK;
}

int run_plan(dipamd_batch_plan *P, hipStream_t stream) {
    return launch_groups(P->groups, P->args, P->d_counter, P->in_dtype, P->is_double, P->fast, stream);
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

template <typename TB>
int run_noalloc(const dipamd_batch_image *images, int n, int in_dtype, const int bc[2], double truncation, bool fast,
                int cfg, hipStream_t stream) {
    BatchTables<TB> T;
    int st = plan_tables<TB>(images, n, in_dtype, truncation, fast, cfg, T);
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
    int cfg = tile_config;
    if (P->is_double && cfg == 2) cfg = 1;
    if (cfg < 0) {
        // autotune over the tile configs on this batch (timed on the caller's stream),
        // memoised per (input type, fast, largest radius, boundary condition)
        int rmax = 0;
        for (int i = 0; i < n; ++i) {
            const double s2 = std::max(images[i].sigma_x, images[i].sigma_y);
            rmax = std::max(rmax, gauss_radius(s2, 0, truncation, in_dtype == DIPAMD_DT_DFLOAT));
        }
        const auto key = std::make_tuple(in_dtype, P->fast ? 1 : 0, rmax, bcs[0] * 16 + bcs[1]);
        {
            std::lock_guard<std::mutex> lk(g_tune_mutex);
            auto it = g_batch_tuned.find(key);
            if (it != g_batch_tuned.end()) cfg = it->second;
        }
        if (cfg < 0) {
            float best = 1e30f;
            int best_cfg = 0;
            hipEvent_t e0, e1;
            (void)hipEventCreate(&e0);
            (void)hipEventCreate(&e1);
            for (int c = 0; c < kNumFusedConfigs; ++c) {
                if (P->is_double && c == 2) continue;
                if ((st = make_plan(P, c)) != DIPAMD_OK) {
                    free_plan_buffers(P);
                    continue;
                }
                float ms = 0, tot = 0;
                bool ok = run_plan(P, s) == DIPAMD_OK;  // warm-up
                for (int rep = 0; ok && rep < 2; ++rep) {
                    (void)hipEventRecord(e0, s);
                    ok = run_plan(P, s) == DIPAMD_OK;
                    (void)hipEventRecord(e1, s);
                    (void)hipEventSynchronize(e1);
                    (void)hipEventElapsedTime(&ms, e0, e1);
                    tot += ms;
                }
                free_plan_buffers(P);
                if (ok && tot < best) {
                    best = tot;
                    best_cfg = c;
                }
            }
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            cfg = best_cfg;
            std::lock_guard<std::mutex> lk(g_tune_mutex);
            g_batch_tuned[key] = cfg;
        }
    }
    if ((st = make_plan(P, cfg)) != DIPAMD_OK) {
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
    int cfg = tile_config;
    if (cfg < 0) {
        cfg = 1;
        int rmax = 0;
        for (int i = 0; i < n; ++i)
            rmax = std::max(rmax, gauss_radius(std::max(images[i].sigma_x, images[i].sigma_y), 0, truncation,
                                               in_dtype == DIPAMD_DT_DFLOAT));
        const auto key = std::make_tuple(static_cast<int>(in_dtype), fast ? 1 : 0, rmax, bcs[0] * 16 + bcs[1]);
        std::lock_guard<std::mutex> lk(g_tune_mutex);
        auto it = g_batch_tuned.find(key);
        if (it != g_batch_tuned.end()) cfg = it->second;
    }
    if (is_double && cfg == 2) cfg = 1;
    hipStream_t s = static_cast<hipStream_t>(stream);
    return is_double ? run_noalloc<double>(images, n, in_dtype, bcs, truncation, fast, cfg, s)
                     : run_noalloc<float>(images, n, in_dtype, bcs, truncation, fast, cfg, s);
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
