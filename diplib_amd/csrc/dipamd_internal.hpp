// Internal host-side declarations shared by the C ABI translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "dipamd.h"

namespace dipamd {

// A strided device view; dims beyond ndim have size 1.
struct View {
    void *p;
    int dt;
    int64_t sizes[DIPAMD_MAX_DIMS];
    int64_t strides[DIPAMD_MAX_DIMS];
};

// InternOneDimensionalFilter (convolution.cpp:67-162), real part: `data` holds the
// taps reversed (correlation order), `border` the boundary extension, `shift` where
// the tap loop starts relative to the output sample.
struct InternFilter {
    std::vector<double> data;
    int size = 0;
    int origin = 0;
    int border = 0;
    int shift = 0;
    int sym = DIPAMD_SYM_GENERAL;
    // 1 when the filter reads one sample past the reference's border on the left
    // (d-even / d-odd with an origin left of the centre): `border` then includes that
    // sample, which reads as 0.
    int zpad = 0;
};

// Records HIP events around a kernel launch while dipamd_trace_enable(1) is on.
class TraceScope {
  public:
    TraceScope(const char *name, int64_t bytes, hipStream_t stream, int64_t ops = 0);
    ~TraceScope();
    TraceScope(const TraceScope &) = delete;
    TraceScope &operator=(const TraceScope &) = delete;

  private:
    hipStream_t stream_;
    int index_;
};

struct Outer;
// Geometry of the dims other than skip0 and skip1 (size > 1), ascending |input stride|.
void fill_outer(Outer &o, const View &src, const View &dst, int ndim, int skip0, int skip1);

int set_error(int code, const std::string &msg);
const char *last_error();
size_t dt_size(int dt);
int suggest_flex(int dt);
int gauss_radius(double sigma, int order, double truncation, bool is_double);
std::vector<double> half_gaussian_n(double sigma, int order, int n);
std::vector<double> half_gaussian(double sigma, int order, double truncation, bool is_double);
int intern_filter(const double *taps, int n, int sym, int origin_in, InternFilter &F);
bool meaningless(const InternFilter &F, bool use_double);
bool bc_supported(int bc);
bool bc_value_based(int bc);
int check_image(const dipamd_image *im, const char *what);
View to_view(const dipamd_image *im);
View contiguous_view(void *p, int dt, int ndim, const int64_t *sizes);
bool views_overlap(const View &x, const View &y, int ndim);
int broadcast_bc(const int32_t *bc, int nbc, int ndim, int *out);
hipError_t scratch_alloc(void **p, size_t bytes, hipStream_t stream);
void scratch_free(void *p, hipStream_t stream);

// Persistent tuning cache (autotune.cpp): measured choices keyed by a string naming the
// decision and its parameters, per device ("gfx950-256cu").  The file is DIPAMD_TUNE_FILE
// or tune_cache.txt beside the library, one "<device> <key>\t<value>" per line, a later
// line overriding an earlier one; tune_put appends (DIPAMD_TUNE_WRITE=0: memory only).
std::string tune_device_tag();
bool tune_get(const std::string &key, int *value);
void tune_put(const std::string &key, int value);

// Dynamic LDS a workgroup may use on this device (64 KiB unless the device reports
// less); kernels are sized against it.
size_t lds_limit_bytes();

int run_pass(const View &src, const View &dst, int ndim, int d, const InternFilter &F, int bc, bool is_double,
             bool fast, hipStream_t stream);
// *out = <u, pass(src)> in FP64 (the pass along dim d, no output written), for a
// strided d and a symmetric filter of radius 1..12 with an index-based or constant
// condition; -100 otherwise (run_pass and a dot product then).
int run_pass_dot(const View &src, const View &u, int ndim, int d, const InternFilter &F, int bc, bool is_double,
                 bool fast, double *out, hipStream_t stream);
int cast_to_buffer(const View &src, const View &dst, int ndim, bool to_double, hipStream_t stream);
int cast_from_buffer(const View &src, const View &dst, int ndim, hipStream_t stream);

// `keep` (optional, buffer type, the input's sizes): receives the intermediate image after
// dims 0 and 1 when the first pass is the fused 2-D pass over them and a further pass
// follows; *kept tells whether it was written.
int separable_device(const View &in, const View &out, int ndim, const InternFilter *const *filters, const int *bc,
                     const bool *process, uint32_t flags, hipStream_t stream, const View *keep = nullptr,
                     bool *kept = nullptr);
// The separable convolution over a complex buffer type (complex.cpp): a complex filter
// in the array (any_complex_filter) or a complex input.
int separable_complex(const View &in, const View &out, int ndim, const dipamd_filter *filters, int nfilters,
                      const int *bc, const uint8_t *process, bool any_complex_filter, hipStream_t stream);
// The derivative family's one-launch form on a 2-D uint8 image (k_stream2d_deriv.hip):
// mode kDeriv*, fx[o] / fy[o] the interned x / y Gaussian derivative of order o (null
// where the function does not use it), outs the sfloat outputs.  -100 when its
// conditions do not hold (the caller then makes the separate calls).
int run_stream2d_deriv(int mode, const View &src, const View *outs, int nouts, const InternFilter *const *fx,
                       const InternFilter *const *fy, int bcx, int bcy, bool fast, hipStream_t stream);
// A real-valued entry point rejects complex images with DIPAMD_E_UNSUPPORTED.
int require_real(const dipamd_image *im);

// Schedule choices (autotune.cpp), measured by dipamd_autotune where available.
// plan_fusion: whether and how the first two passes fuse (-1 one pass per axis, 0..2
// the fused kernel's tile config, 3 the uint8 register-streaming kernel first);
// src_dt is the first pass's source type (uint8 or the buffer type).
int plan_fusion(int src_dt, bool is_double, bool fast, int rx, int ry, int64_t nx, int64_t ny, uint32_t flags);
// The fused tile config where the streaming kernel does not run (-1: none fits).
int plan_fused_config(int src_dt, bool is_double, bool fast, int rx, int ry, int64_t nx, int64_t ny);
// Kernel of a single pass along a strided axis: 0 column tiles, 1 LDS ring, 2 sweep.
int plan_ysweep(bool is_double, bool fast, int r);
// Whether the register-streaming fused kernel may be used (DIPAMD_STREAM2D=0 disables).
bool stream2d_enabled();
// DIPAMD_ZSWEEP=0 turns the register-window sweep off (experiments).
bool zsweep_enabled();
// DIPAMD_FUSED_RC=0 turns the radius-class fused kernel off (experiments).
bool fused_rc_enabled();
// Workgroups a strided sweep aims for when it splits the filter axis into chunks:
// `dflt`, or DIPAMD_SWEEP_BLOCKS when set (experiments).
int sweep_target_blocks(int dflt);
int stream2d_target_waves();  // waves per launch of the register-streaming kernel
// Tile config of the LDS-ring sweep along a strided axis (y-only), or -1.
int choose_ysweep_config(bool is_double, int ry);

}  // namespace dipamd
