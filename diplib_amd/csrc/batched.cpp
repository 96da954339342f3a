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

struct dipamd_batch_plan {
    int n = 0;
    int in_dtype = 0, out_dtype = 0;
    bool is_double = false;
    bool fast = false;
    int cfg = 1;
    int grid = 0;
    size_t lds = 0;
    int64_t nitems = 0;
    int bc[2] = {0, 0};
    double truncation = 3;
    uint32_t flags = 0;
    void *d_images = nullptr;
    void *d_items = nullptr;
    void *d_taps = nullptr;
    int32_t *d_counter = nullptr;
    std::vector<dipamd_batch_image> host_images;
    std::vector<uint8_t> args;  // BatchArgs<TB> bytes
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t cap_stream = nullptr;
};

namespace {

std::mutex g_tune_mutex;
std::map<std::tuple<int, int, int, int>, int> g_batch_tuned;  // (in dtype, fast, rmax, bc) -> cfg

template <typename TB>
int build_plan(dipamd_batch_plan *P, const dipamd_batch_image *images, int n) {
    const bool compute_double = P->in_dtype == DIPAMD_DT_DFLOAT;
    std::vector<BatchImage<TB>> ims(static_cast<size_t>(n));
    std::vector<TB> taps;
    int rmax_x = 0, rmax_y = 0;
    std::vector<std::vector<double>> cache;
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
            std::vector<double> h = half_gaussian(s, 0, P->truncation, compute_double);
            *r = static_cast<int>(h.size()) - 1;
            std::reverse(h.begin(), h.end());  // correlation order: centre first
            return h;
        };
        int rx = 0, ry = 0;
        std::vector<double> hx = axis(b.sigma_x, b.width, &rx);
        std::vector<double> hy = axis(b.sigma_y, b.height, &ry);
        if (rx > kFusedCap - 1 || ry > kFusedCap - 1)
            return set_error(DIPAMD_E_UNSUPPORTED, "batched form: radius above 64 (sigma too large)");
        BatchImage<TB> &im = ims[static_cast<size_t>(i)];
        im.in = b.in;
        im.out = static_cast<TB *>(b.out);
        im.nx = b.width;
        im.ny = b.height;
        im.syin = b.in_stride_y;
        im.syout = b.out_stride_y;
        im.rx = rx;
        im.ry = ry;
        im.tap_offset = static_cast<int32_t>(taps.size());
        im.pad_ = 0;
        for (double v : hx) taps.push_back(static_cast<TB>(v));
        for (double v : hy) taps.push_back(static_cast<TB>(v));
        rmax_x = std::max(rmax_x, rx);
        rmax_y = std::max(rmax_y, ry);
    }
    // a tile config whose LDS (staging + ring for the largest radii) fits; the
    // requested one when it does
    const size_t limit = lds_limit_bytes();
    if (fused_lds_bytes(P->cfg, rmax_x, rmax_y, sizeof(TB) == 8) > limit) {
        int fit = -1;
        for (int c = kNumFusedConfigs - 1; c >= 0; --c)
            if (!(sizeof(TB) == 8 && c == 2) && fused_lds_bytes(c, rmax_x, rmax_y, sizeof(TB) == 8) <= limit) {
                fit = c;
                break;
            }
        if (fit < 0) return set_error(DIPAMD_E_UNSUPPORTED, "batched form: tile does not fit in LDS");
        P->cfg = fit;
    }
    // work items, costliest images first (longest-processing-time order)
    const int txc = kFusedConfigs[P->cfg].txc;
    std::vector<int> order(static_cast<size_t>(n));
    for (int i = 0; i < n; ++i) order[static_cast<size_t>(i)] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        const auto &A = ims[static_cast<size_t>(a)];
        const auto &B = ims[static_cast<size_t>(b)];
        return (A.rx + A.ry + 2) > (B.rx + B.ry + 2);
    });
    std::vector<BatchItem> items;
    for (int i : order) {
        const auto &im = ims[static_cast<size_t>(i)];
        const int ntiles = static_cast<int>((im.nx + txc - 1) / txc);
        // row chunks: long enough that the y halo stays small, short enough to balance
        int ch = std::max(batch_chunk_rule().first, ((batch_chunk_rule().second * im.ry + 15) / 16) * 16);
        ch = std::min<int>(ch, static_cast<int>(((im.ny + 15) / 16) * 16));
        for (int64_t z = 0; z < im.ny; z += ch)
            for (int t = 0; t < ntiles; ++t)
                items.push_back({i, t, static_cast<int32_t>(z), static_cast<int32_t>(std::min<int64_t>(z + ch, im.ny))});
    }
    P->nitems = static_cast<int64_t>(items.size());
    P->lds = fused_lds_bytes(P->cfg, rmax_x, rmax_y, sizeof(TB) == 8);
    // device tables
    hipError_t e;
    if ((e = hipMalloc(&P->d_images, ims.size() * sizeof(BatchImage<TB>))) != hipSuccess ||
        (e = hipMalloc(&P->d_items, std::max<size_t>(1, items.size()) * sizeof(BatchItem))) != hipSuccess ||
        (e = hipMalloc(&P->d_taps, taps.size() * sizeof(TB))) != hipSuccess ||
        (e = hipMalloc(reinterpret_cast<void **>(&P->d_counter), 256)) != hipSuccess)
        return set_error(DIPAMD_E_HIP, std::string("batch plan allocation: ") + hipGetErrorString(e));
    if ((e = hipMemcpy(P->d_images, ims.data(), ims.size() * sizeof(BatchImage<TB>), hipMemcpyHostToDevice)) !=
            hipSuccess ||
        (e = hipMemcpy(P->d_items, items.data(), items.size() * sizeof(BatchItem), hipMemcpyHostToDevice)) !=
            hipSuccess ||
        (e = hipMemcpy(P->d_taps, taps.data(), taps.size() * sizeof(TB), hipMemcpyHostToDevice)) != hipSuccess)
        return set_error(DIPAMD_E_HIP, std::string("batch plan upload: ") + hipGetErrorString(e));
    BatchArgs<TB> a;
    std::memset(&a, 0, sizeof(a));
    a.images = static_cast<const BatchImage<TB> *>(P->d_images);
    a.items = static_cast<const BatchItem *>(P->d_items);
    a.taps = static_cast<const TB *>(P->d_taps);
    a.counter = P->d_counter;
    a.nitems = static_cast<int32_t>(items.size());
    a.bcx = P->bc[0];
    a.bcy = P->bc[1];
    auto cv = [](int bc) -> TB {
        if (bc == DIPAMD_BC_ADD_MAX_VALUE) return std::numeric_limits<TB>::max();
        if (bc == DIPAMD_BC_ADD_MIN_VALUE) return std::numeric_limits<TB>::lowest();
        return TB(0);
    };
    a.cvalx = cv(P->bc[0]);
    a.cvaly = cv(P->bc[1]);
    P->args.resize(sizeof(a));
    std::memcpy(P->args.data(), &a, sizeof(a));
    int dev = 0, ncu = 256;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    const InType ti = P->in_dtype == DIPAMD_DT_UINT8 ? IN_U8 : IN_BUF;
    const int occ = batched_occupancy(ti, P->is_double, P->cfg, P->fast, P->lds);
    P->grid = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(static_cast<int64_t>(ncu) * occ, P->nitems)));
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

int run_plan(dipamd_batch_plan *P, hipStream_t stream) {
    hipError_t e = launch_reset_counter(P->d_counter, stream);
    if (e == hipSuccess) {
        const InType ti = P->in_dtype == DIPAMD_DT_UINT8 ? IN_U8 : IN_BUF;
        e = launch_batched(ti, P->is_double, P->cfg, P->fast, P->args.data(), P->grid, P->lds, stream);
    }
    if (e != hipSuccess) return set_error(DIPAMD_E_HIP, std::string("batched launch: ") + hipGetErrorString(e));
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
    if (n < 1 || !images) return set_error(DIPAMD_E_PARAMETER, "empty batch");
    if (in_dtype != DIPAMD_DT_UINT8 && in_dtype != DIPAMD_DT_SFLOAT && in_dtype != DIPAMD_DT_DFLOAT)
        return set_error(DIPAMD_E_UNSUPPORTED, "batched form: input must be uint8, sfloat or dfloat");
    if (out_dtype != suggest_flex(in_dtype))
        return set_error(DIPAMD_E_UNSUPPORTED, "batched form: output type must be SuggestFlex(input)");
    int bcs[2];
    int st = broadcast_bc(bc, nbc, 2, bcs);
    if (st != DIPAMD_OK) return st;
    for (int b : bcs)
        if (!bc_supported(b)) return set_error(DIPAMD_E_UNSUPPORTED, "boundary condition not supported");
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
