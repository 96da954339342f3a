// Schedule choices of the plain path and the persistent tuning cache.
//
// Decisions (per device, keyed by the buffer type, exact/fast arithmetic, a radius class
// and for fusion a plane-size class):
//   fuse    the first two passes of a 2-D (or stacked) image: -1 one pass per axis,
//           0..2 the fused kernel with that tile config, 3 the register-streaming
//           kernel for uint8 planes (radius <= 4), falling back to "fusecfg";
//   fusecfg the best fused tile config when the streaming kernel cannot run;
//   ysweep  a single pass along a strided axis: 0 column tiles, 1 LDS ring, 2 the
//           register-window / global sweep.
// They are measured by dipamd_autotune (every candidate timed on synthetic planes of the
// class's size) and kept in the cache: DIPAMD_TUNE_FILE, else tune_cache.txt beside the
// library, lines "<device> <key>\t<value>", later lines winning.  tune_defaults.txt
// beside the library holds the same measurements made on an MI355X and is read first,
// so an untuned install starts from measured choices; a decision with no entry at all
// takes the fallback noted at each decision.  Environment overrides for experiments:
// DIPAMD_FUSED_CONFIG=<cfg or -1>, DIPAMD_YSWEEP=col|ring|global, DIPAMD_FUSE_MAX_RADIUS.
#include <dlfcn.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "dipamd.h"
#include "dipamd_internal.hpp"
#include "sep_config.hpp"
#include "sep_kernels.hpp"

namespace dipamd {

namespace {

int g_forced = -2;  // DIPAMD_FUSED_CONFIG
std::once_flag g_env_once;

// Choices forced by the tuner while it times a candidate (this thread only).
thread_local int t_force_fuse = -100;
thread_local int t_force_ysweep = -100;

// The persistent key -> value cache of measured choices.
std::once_flag g_cache_once;
std::mutex g_cache_mu;
std::map<std::string, int> g_cache;
std::string g_cache_path;

std::string library_dir() {
    Dl_info info;
    if (dladdr(reinterpret_cast<void *>(&library_dir), &info) && info.dli_fname) {
        std::string p(info.dli_fname);
        size_t s = p.find_last_of('/');
        if (s != std::string::npos) return p.substr(0, s);
    }
    return ".";
}

void read_env() {
    if (const char *f = std::getenv("DIPAMD_FUSED_CONFIG")) g_forced = std::atoi(f);
}

std::string cache_path() {
    if (const char *f = std::getenv("DIPAMD_TUNE_FILE")) return f;
    return library_dir() + "/tune_cache.txt";
}

// Lines "<key>\t<value>"; a later line overrides an earlier one for the same key.
void read_table(const std::string &path) {
    FILE *fp = std::fopen(path.c_str(), "r");
    if (!fp) return;
    char line[512];
    while (std::fgets(line, sizeof(line), fp)) {
        if (line[0] == '#') continue;
        char *tab = std::strchr(line, '\t');
        if (!tab) continue;
        *tab = 0;
        g_cache[line] = std::atoi(tab + 1);
    }
    std::fclose(fp);
}

void load_cache() {
    g_cache_path = cache_path();
    read_table(library_dir() + "/tune_defaults.txt");
    read_table(g_cache_path);
}

bool fits(int cfg, bool is_double, int rx, int ry) {
    if (cfg < 0 || cfg >= kNumFusedConfigs) return false;
    if (is_double && cfg == 2) return false;
    return fused_lds_bytes(cfg, rx, ry, is_double) <= lds_limit_bytes();
}

// Radius classes the tuner measures; a radius maps to the smallest class >= it.
constexpr int kRadiusClasses[] = {1, 2, 3, 4, 5, 6, 8, 10, 12, 14, 16, 20, 24, 32, 48, 64};
int radius_class(int r) {
    for (int c : kRadiusClasses)
        if (r <= c) return c;
    return 0;  // beyond the tuned range
}

// Plane-size class: planes up to 2048^2 pixels, and larger ones.
int plane_class(int64_t px) { return px <= int64_t(2048) * 2048 ? 0 : 1; }

const char *src_name(int src_dt) {
    return src_dt == DIPAMD_DT_UINT8 ? "u8" : (src_dt == DIPAMD_DT_DFLOAT ? "f64" : "f32");
}

std::string fuse_key(const char *what, int src_dt, bool fast, int rclass, int pc) {
    return std::string(what) + " src" + src_name(src_dt) + (fast ? " fast" : " exact") + " r" + std::to_string(rclass) +
           " px" + std::to_string(pc);
}

std::string ysweep_key(bool is_double, bool fast, int rclass) {
    return std::string("ysweep ") + (is_double ? "f64" : "f32") + (fast ? " fast" : " exact") + " r" +
           std::to_string(rclass);
}

}  // namespace

std::string tune_device_tag() {
    static std::string tag = [] {
        int dev = 0, ncu = 0;
        (void)hipGetDevice(&dev);
        (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
        hipDeviceProp_t prop;
        std::string arch = "gfx";
        if (hipGetDeviceProperties(&prop, dev) == hipSuccess) arch = prop.gcnArchName;
        const size_t colon = arch.find(':');
        if (colon != std::string::npos) arch = arch.substr(0, colon);
        return arch + "-" + std::to_string(ncu) + "cu";
    }();
    return tag;
}

bool tune_get(const std::string &key, int *value) {
    std::call_once(g_cache_once, load_cache);
    std::lock_guard<std::mutex> lk(g_cache_mu);
    auto it = g_cache.find(tune_device_tag() + " " + key);
    if (it == g_cache.end()) return false;
    *value = it->second;
    return true;
}

void tune_put(const std::string &key, int value) {
    std::call_once(g_cache_once, load_cache);
    const std::string full = tune_device_tag() + " " + key;
    std::lock_guard<std::mutex> lk(g_cache_mu);
    g_cache[full] = value;
    if (const char *v = std::getenv("DIPAMD_TUNE_WRITE"))
        if (v[0] == '0') return;
    FILE *fp = std::fopen(g_cache_path.c_str(), "a");
    if (!fp) return;  // read-only location: the choice stays in memory for this process
    std::fprintf(fp, "%s\t%d\n", full.c_str(), value);
    std::fclose(fp);
}

// Largest radius fused by default when nothing is measured (and the override).
int fuse_max_radius(int64_t px) {
    static int r = [] {
        const char *v = std::getenv("DIPAMD_FUSE_MAX_RADIUS");
        return v ? std::atoi(v) : -1;
    }();
    if (r >= 0) return r;
    return px <= int64_t(2048) * 2048 ? 24 : 16;
}

bool fused_rc_enabled() {
    static bool on = [] {
        const char *v = std::getenv("DIPAMD_FUSED_RC");
        return !(v && v[0] == '0');
    }();
    return on;
}

bool stream2d_enabled() {
    static bool on = [] {
        const char *v = std::getenv("DIPAMD_STREAM2D");
        return !(v && v[0] == '0');
    }();
    return on;
}

int sweep_target_blocks(int dflt) {
    static int n = [] {
        const char *v = std::getenv("DIPAMD_SWEEP_BLOCKS");
        return v ? std::atoi(v) : 0;
    }();
    return n > 0 ? n : dflt;
}

bool zsweep_enabled() {
    static bool on = [] {
        const char *v = std::getenv("DIPAMD_ZSWEEP");
        return !(v && v[0] == '0');
    }();
    return on;
}

int stream2d_target_waves() {
    static int n = [] {
        const char *v = std::getenv("DIPAMD_STREAM_WAVES");
        const int k = v ? std::atoi(v) : 0;
        return k > 0 ? k : 6144;
    }();
    return n;
}

size_t lds_limit_bytes() {
    static size_t limit = [] {
        int dev = 0, v = 0;
        (void)hipGetDevice(&dev);
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || v <= 0)
            v = 65536;
        return static_cast<size_t>(v < 65536 ? v : 65536);
    }();
    return limit;
}

int plan_fusion(int src_dt, bool is_double, bool fast, int rx, int ry, int64_t nx, int64_t ny, uint32_t flags) {
    std::call_once(g_env_once, read_env);
    const int r = std::max(rx, ry);
    auto fitted = [&](int cfg) -> int {
        if (fits(cfg, is_double, rx, ry)) return cfg;
        for (int c = 0; c < kNumFusedConfigs; ++c)
            if (fits(c, is_double, rx, ry)) return c;
        return -1;
    };
    if (t_force_fuse != -100) return t_force_fuse == 3 ? 3 : (t_force_fuse < 0 ? -1 : fitted(t_force_fuse));
    if (g_forced != -2) return g_forced < 0 ? -1 : (fits(g_forced, is_double, rx, ry) ? g_forced : -1);
    const int rc = radius_class(r);
    int v = -2;
    if (rc > 0 && tune_get(fuse_key("fuse", src_dt, fast, rc, plane_class(nx * ny)), &v)) {
        if (v == 3) return 3;
        if (v < 0) return (flags & DIPAMD_FORCE_FUSION) ? fitted(is_double ? 0 : 1) : -1;
        return fitted(v);
    }
    // nothing measured: fuse short filters (the fused pass saves the intermediate round
    // trip); long ones run as a line pass plus a strided sweep, which recompute nothing
    if (r > fuse_max_radius(nx * ny) && !(flags & DIPAMD_FORCE_FUSION)) return -1;
    if (src_dt == DIPAMD_DT_UINT8 && !is_double && r <= kStreamRMaxHost) return 3;
    return fitted((!is_double && r <= 16 && nx >= 256) ? 1 : 0);
}

int plan_fused_config(int src_dt, bool is_double, bool fast, int rx, int ry, int64_t nx, int64_t ny) {
    std::call_once(g_env_once, read_env);
    const int r = std::max(rx, ry);
    const int rc = radius_class(r);
    int v = -1;
    if (t_force_fuse >= 0 && t_force_fuse < kNumFusedConfigs) v = t_force_fuse;
    else if (!(rc > 0 && tune_get(fuse_key("fusecfg", src_dt, fast, rc, plane_class(nx * ny)), &v)))
        v = (!is_double && r <= 16 && nx >= 256) ? 1 : 0;
    if (fits(v, is_double, rx, ry)) return v;
    for (int c = 0; c < kNumFusedConfigs; ++c)
        if (fits(c, is_double, rx, ry)) return c;
    return -1;
}

int plan_ysweep(bool is_double, bool fast, int r) {
    if (t_force_ysweep != -100) return t_force_ysweep;
    static int env = [] {
        const char *v = std::getenv("DIPAMD_YSWEEP");
        if (!v) return -1;
        std::string s(v);
        return s == "col" ? 0 : (s == "ring" ? 1 : (s == "global" ? 2 : -1));
    }();
    if (env >= 0) return env;
    const int rc = radius_class(r);
    int v = -1;
    if (rc > 0 && tune_get(ysweep_key(is_double, fast, rc), &v) && v >= 0 && v <= 2) return v;
    // nothing measured: column tiles for long filters (VALU-bound: the LDS tile keeps the
    // window loads cheap), the global sweep for short ones (HBM-bound: no staging)
    return r >= kColsweepMinBorder ? 0 : 2;
}

int choose_ysweep_config(bool is_double, int ry) {
    std::call_once(g_env_once, read_env);
    if (g_forced == -1) return -1;
    const int pref = is_double ? 0 : 1;
    if (fits(pref, is_double, 0, ry)) return pref;
    return fits(0, is_double, 0, ry) ? 0 : -1;
}

}  // namespace dipamd

using namespace dipamd;

namespace {

// Times `run` (median of `reps` after one warm-up) on `stream`; -1 on failure.
template <typename F>
float time_ms(F run, hipStream_t stream, int reps) {
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess) return -1.f;
    if (hipEventCreate(&b) != hipSuccess) {
        (void)hipEventDestroy(a);
        return -1.f;
    }
    std::vector<float> t;
    bool ok = run() == DIPAMD_OK;
    for (int i = 0; ok && i < reps; ++i) {
        (void)hipEventRecord(a, stream);
        ok = run() == DIPAMD_OK;
        (void)hipEventRecord(b, stream);
        (void)hipEventSynchronize(b);
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, a, b);
        t.push_back(ms);
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (!ok || t.empty()) return -1.f;
    std::sort(t.begin(), t.end());
    return t[t.size() / 2];
}

struct ForceScope {
    ForceScope(int fuse, int ysweep) {
        t_force_fuse = fuse;
        t_force_ysweep = ysweep;
    }
    ~ForceScope() {
        t_force_fuse = -100;
        t_force_ysweep = -100;
    }
};

}  // namespace

extern "C" int dipamd_autotune(uint32_t flags, void *stream_) {
    hipStream_t stream = static_cast<hipStream_t>(stream_);
    const bool fast = (flags & DIPAMD_FAST) != 0;
    const bool quick = (flags & DIPAMD_TUNE_QUICK) != 0;
    const int reps = quick ? 2 : 5;
    // synthetic planes (uniform content: the kernels' speed does not depend on values)
    const int64_t side[2] = {quick ? 1024 : 2048, quick ? 2048 : 8192};
    struct Src {
        int dt;
        size_t es;
    };
    const Src srcs[3] = {{DIPAMD_DT_UINT8, 1}, {DIPAMD_DT_SFLOAT, 4}, {DIPAMD_DT_DFLOAT, 8}};
    const size_t max_bytes = static_cast<size_t>(side[1]) * side[1] * 16;
    void *in = nullptr, *out = nullptr;
    hipError_t e = hipMalloc(&in, max_bytes);
    if (e == hipSuccess) e = hipMalloc(&out, max_bytes);
    if (e != hipSuccess) {
        if (in) (void)hipFree(in);
        return set_error(DIPAMD_E_HIP, std::string("tuning buffers: ") + hipGetErrorString(e));
    }
    (void)hipMemsetAsync(in, 0x3c, max_bytes, stream);
    int st = DIPAMD_OK;
    for (const Src &s : srcs) {
        const bool dbl = s.dt == DIPAMD_DT_DFLOAT;
        const int buf_dt = dbl ? DIPAMD_DT_DFLOAT : DIPAMD_DT_SFLOAT;
        for (int pc = 0; pc < 2 && st == DIPAMD_OK; ++pc) {
            const int64_t n = (dbl && pc == 1) ? side[1] / 2 : side[pc];  // f64: a quarter of the pixels
            dipamd_image a{}, b{};
            a.origin = in;
            a.dtype = s.dt;
            a.ndim = 2;
            a.sizes[0] = a.sizes[1] = n;
            a.strides[0] = 1;
            a.strides[1] = n;
            b = a;
            b.origin = out;
            b.dtype = buf_dt;
            for (int rc : kRadiusClasses) {
                if (rc > 48) break;
                // a Gaussian whose radius is the class (sigma = r / 3 at truncation 3)
                const double sigma = rc / 3.0;
                const double sig[2] = {sigma, sigma};
                const int32_t ord[2] = {0, 0};
                auto run = [&]() {
                    return dipamd_gauss_fir(&a, &b, sig, 2, ord, 2, nullptr, 0, 3.0, flags & DIPAMD_FAST, stream);
                };
                float best = 1e30f, best_fused = 1e30f;
                int choice = -1, fused_choice = -1;
                for (int cand = -1; cand <= 3; ++cand) {
                    if (cand == 3 && !(s.dt == DIPAMD_DT_UINT8 && rc <= kStreamRMaxHost)) continue;
                    if (cand >= 0 && cand < 3 && !fits(cand, dbl, rc, rc)) continue;
                    float ms;
                    {
                        ForceScope f(cand, -100);
                        ms = time_ms(run, stream, reps);
                    }
                    if (ms < 0) continue;
                    if (ms < best) {
                        best = ms;
                        choice = cand;
                    }
                    if (cand >= 0 && cand < 3 && ms < best_fused) {
                        best_fused = ms;
                        fused_choice = cand;
                    }
                }
                if (best < 1e30f) tune_put(fuse_key("fuse", s.dt, fast, rc, pc), choice);
                if (best_fused < 1e30f) tune_put(fuse_key("fusecfg", s.dt, fast, rc, pc), fused_choice);
                // the strided pass alone (the y axis of a buffer-type plane)
                if (pc == 1 && s.dt != DIPAMD_DT_UINT8) {
                    const double sy[2] = {0.0, sigma};
                    auto runy = [&]() {
                        return dipamd_gauss_fir(&a, &b, sy, 2, ord, 2, nullptr, 0, 3.0, flags & DIPAMD_FAST, stream);
                    };
                    float by = 1e30f;
                    int ky = -1;
                    for (int kind = 0; kind <= 2; ++kind) {
                        float ms;
                        {
                            ForceScope f(-100, kind);
                            ms = time_ms(runy, stream, reps);
                        }
                        if (ms >= 0 && ms < by) {
                            by = ms;
                            ky = kind;
                        }
                    }
                    if (ky >= 0) tune_put(ysweep_key(dbl, fast, rc), ky);
                }
            }
        }
    }
    (void)hipStreamSynchronize(stream);
    (void)hipFree(in);
    (void)hipFree(out);
    return st;
}
