// Schedule choice for the fused 2-D kernel, keyed by (dtype, radius bucket).
//
// A table of measured choices is read from the file named by DIPAMD_TUNE_FILE (or
// diplib_amd/tune_cache.txt next to the library) when present; lines are
//     <is_double> <radius_lo> <radius_hi> <config or -1>
// with -1 meaning "run one pass per axis".  Without a table the built-in defaults
// below apply.  DIPAMD_FUSED_CONFIG=<n> forces a config (or -1) for every call.
#include <dlfcn.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "dipamd_internal.hpp"
#include "sep_kernels.hpp"

namespace dipamd {

namespace {

struct TuneRow {
    int is_double, rlo, rhi, cfg;
};

std::once_flag g_tune_once;
std::vector<TuneRow> g_tune;
int g_forced = -2;

// The persistent key -> value cache of measured choices.
std::once_flag g_cache_once;
std::mutex g_cache_mu;
std::map<std::string, int> g_cache;
std::string g_cache_path;

std::string library_dir() {
    Dl_info info;
    if (dladdr(reinterpret_cast<void *>(&choose_fused_config), &info) && info.dli_fname) {
        std::string p(info.dli_fname);
        size_t s = p.find_last_of('/');
        if (s != std::string::npos) return p.substr(0, s);
    }
    return ".";
}

void load_table() {
    if (const char *f = std::getenv("DIPAMD_FUSED_CONFIG")) g_forced = std::atoi(f);
    std::string path;
    if (const char *f = std::getenv("DIPAMD_TUNE_FILE")) {
        path = f;
    } else {
        path = library_dir() + "/tune_cache.txt";
    }
    FILE *fp = std::fopen(path.c_str(), "r");
    if (!fp) return;
    char line[256];
    while (std::fgets(line, sizeof(line), fp)) {
        if (line[0] == '#') continue;
        TuneRow r;
        if (std::sscanf(line, "%d %d %d %d", &r.is_double, &r.rlo, &r.rhi, &r.cfg) == 4) g_tune.push_back(r);
    }
    std::fclose(fp);
}

std::string cache_path() {
    if (const char *f = std::getenv("DIPAMD_TUNE_FILE")) return f;
    return library_dir() + "/tune_cache.txt";
}

// Lines "<key>\t<value>"; a later line overrides an earlier one for the same key.
void load_cache() {
    g_cache_path = cache_path();
    FILE *fp = std::fopen(g_cache_path.c_str(), "r");
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

bool fits(int cfg, bool is_double, int rx, int ry) {
    if (cfg < 0 || cfg >= kNumFusedConfigs) return false;
    if (is_double && cfg == 2) return false;
    return fused_lds_bytes(cfg, rx, ry, is_double) <= lds_limit_bytes();
}

}  // namespace

// Largest radius the fused 2-D kernel takes for a plane of px pixels.  On large planes
// beyond radius 16 the two single passes (K1 line + the strided sweeps) are as fast or
// faster (8192^2 u8, sigma 6: 0.306 vs 0.317 ms, sigma 8 order 0: 0.352 vs 0.381 ms;
// sigma 4, radius 12-16: fused 0.20-0.21 vs 0.27-0.30); on planes up to 2048^2 the
// single launch without an intermediate image wins up to radius 24 (512 images of
// 64..2048 squared, sigma 0.5..8, one call each: 9.4 ms with 24, 15.5 ms with 16).
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

int ysweep_kind() {
    static int kind = [] {
        const char *v = std::getenv("DIPAMD_YSWEEP");
        if (!v) return -1;
        std::string s(v);
        return s == "col" ? 0 : (s == "ring" ? 1 : (s == "global" ? 2 : -1));
    }();
    return kind;
}

int choose_ysweep_config(bool is_double, int ry) {
    std::call_once(g_tune_once, load_table);
    if (g_forced == -1) return -1;
    const int pref = is_double ? 0 : 1;
    if (fits(pref, is_double, 0, ry)) return pref;
    return fits(0, is_double, 0, ry) ? 0 : -1;
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

int choose_fused_config(bool is_double, int rx, int ry, int64_t nx, int64_t ny, uint32_t flags) {
    std::call_once(g_tune_once, load_table);
    const int r = rx > ry ? rx : ry;
    if (g_forced != -2) return fits(g_forced, is_double, rx, ry) ? g_forced : -1;
    int cfg = -2;
    for (const TuneRow &t : g_tune) {
        if ((t.is_double != 0) == is_double && r >= t.rlo && r <= t.rhi) {
            cfg = t.cfg;
            break;
        }
    }
    if (cfg == -2) {
        // defaults: fuse short filters (HBM-bound: the fused pass saves the intermediate
        // round trip); long filters are VALU-bound and run as a line pass plus a
        // column-tile sweep, which recompute nothing.  The 128-column tile while its
        // ring fits comfortably, else 64 columns.
        if (r > fuse_max_radius(nx * ny) && !(flags & DIPAMD_FORCE_FUSION)) return -1;
        cfg = (!is_double && r <= 16 && nx >= 256) ? 1 : 0;
    }
    if (cfg < 0) {
        if (!(flags & DIPAMD_FORCE_FUSION)) return -1;
        cfg = is_double ? 0 : 1;
    }
    if (!fits(cfg, is_double, rx, ry)) {
        for (int c = 0; c < kNumFusedConfigs; ++c)
            if (fits(c, is_double, rx, ry)) return c;
        return -1;
    }
    return cfg;
}

}  // namespace dipamd
