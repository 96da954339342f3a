/*
 * dipamd.h -- C ABI of the MI355X-native Gaussian FIR path (DIPlib-compatible).
 *
 * Every entry point works on DEVICE-resident images (HIP device pointers), is
 * stream-ordered (asynchronous on `stream`, a hipStream_t or NULL for the null
 * stream), returns 0 on success or a negative DIPAMD_E_* code, and leaves a
 * thread-local message for dipamd_last_error().  Plain pointers and sizes only.
 *
 * Axis convention is DIPlib's: dim 0 is the fastest-varying axis (x), strides
 * are in samples, sigmas / orders / boundary conditions are per dim in that order
 * (or a single value broadcast to every dim, as dip::ArrayUseParameter does).
 */
#ifndef DIPAMD_H
#define DIPAMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DIPAMD_MAX_DIMS 8
/* Largest 1-D filter half size (radius) the kernels take; sigma <= ~84 at the
 * default truncation of 3. */
#define DIPAMD_MAX_RADIUS 255

/* Sample types (dip::DataType, datatype.h).  The complex types hold (re, im) pairs of
 * float / double; strides of a complex image count complex samples.  Complex images
 * and complex filters are taken by dipamd_separable_convolution and dipamd_gauss_fir
 * (a Gaussian filters the real and imaginary parts, as the reference's does); the
 * other entry points take real images. */
enum {
    DIPAMD_DT_UINT8 = 0,
    DIPAMD_DT_UINT16 = 1,
    DIPAMD_DT_UINT32 = 2,
    DIPAMD_DT_SINT8 = 3,
    DIPAMD_DT_SINT16 = 4,
    DIPAMD_DT_SINT32 = 5,
    DIPAMD_DT_SFLOAT = 6,
    DIPAMD_DT_DFLOAT = 7,
    DIPAMD_DT_BIN = 8,
    DIPAMD_DT_SCOMPLEX = 9,
    DIPAMD_DT_DCOMPLEX = 10
};

/* Boundary conditions (dip::BoundaryCondition, boundary.h:65-80, same order). */
enum {
    DIPAMD_BC_SYMMETRIC_MIRROR = 0,   /* "mirror", "default", "" */
    DIPAMD_BC_ASYMMETRIC_MIRROR = 1,  /* "asym mirror" */
    DIPAMD_BC_ANTISYMMETRIC_REFLECT = 2, /* "antisym reflect" */
    DIPAMD_BC_PERIODIC = 3,           /* "periodic" */
    DIPAMD_BC_ASYMMETRIC_PERIODIC = 4,/* "asym periodic" */
    DIPAMD_BC_ADD_ZEROS = 5,          /* "add zeros" */
    DIPAMD_BC_ADD_MAX_VALUE = 6,      /* "add max" */
    DIPAMD_BC_ADD_MIN_VALUE = 7,      /* "add min" */
    DIPAMD_BC_ZERO_ORDER_EXTRAPOLATE = 8,   /* "zero order" */
    DIPAMD_BC_FIRST_ORDER_EXTRAPOLATE = 9,  /* "first order" */
    DIPAMD_BC_SECOND_ORDER_EXTRAPOLATE = 10, /* "second order" */
    DIPAMD_BC_THIRD_ORDER_EXTRAPOLATE = 11, /* "third order" */
    DIPAMD_BC_ALREADY_EXPANDED = 12         /* "already expanded": an error in the separable
                                               framework, as in the reference (ExpandBuffer,
                                               copy_buffer.cpp:745, has no case for it) */
};
/* Codes 2 and 9-11 are computed from the edge values: the plain (non-batched) calls take
   them (one extra pass over the image per axis that uses one); the batched and backward
   forms return DIPAMD_E_UNSUPPORTED for them. */

/* Filter symmetry (convolution.cpp:40-48 FilterSymmetry). */
enum {
    DIPAMD_SYM_GENERAL = 0,
    DIPAMD_SYM_EVEN = 1,
    DIPAMD_SYM_ODD = 2,
    DIPAMD_SYM_CONJ = 3,   /* without a complex filter in the array: same as EVEN */
    DIPAMD_SYM_D_EVEN = 4,
    DIPAMD_SYM_D_ODD = 5,
    DIPAMD_SYM_D_CONJ = 6  /* without a complex filter in the array: same as D_EVEN */
};

/* Flags. */
enum {
    DIPAMD_EXACT = 0,          /* default: the reference's arithmetic, bit-identical output */
    DIPAMD_FAST = 1,           /* fused multiply-add in the tap loop (not bit-identical) */
    DIPAMD_NO_FUSION = 2,      /* one pass per axis (no fused 2-D kernel) */
    DIPAMD_FORCE_FUSION = 4,   /* fused 2-D kernel wherever it applies (ignores the tuner) */
    /* Derivative family and structure tensor: the Gaussian derivatives by dip::GaussIIR
     * (discrete time fit, default filter orders: what method "IIR", or "best" with a sigma
     * above 10, runs -- derivative.cpp:35-58, 111-114); DIPAMD_TENSOR_IIR: the structure
     * tensor's smoothing too (its method resolves with the tensor sigmas). */
    DIPAMD_GAUSS_IIR = 8,
    DIPAMD_TENSOR_IIR = 16,
    /* The same through dip::GaussFT ("FT", or "best" with a sigma below 0.8; an empty
     * boundary condition array means "mirror" there, as dip::Gauss passes it). */
    DIPAMD_GAUSS_FT = 32,
    DIPAMD_TENSOR_FT = 64,
    /* dipamd_gauss_fir_backward: grad_sigma is ONE device double, the sum over the dims
     * of dL/dsigma_d -- the derivative for one sigma shared by every filtered dim. */
    DIPAMD_GRAD_SIGMA_TOTAL = 128,
    /* dipamd_gauss_ft: the input is a spectrum ("frequency" inRepresentation) / the
     * result is one ("frequency" outRepresentation); see dipamd_gauss_ft. */
    DIPAMD_FT_IN_FREQUENCY = 256,
    DIPAMD_FT_OUT_FREQUENCY = 512
};

/* Error codes. */
enum {
    DIPAMD_OK = 0,
    DIPAMD_E_PARAMETER = -1,     /* dip::ParameterError */
    DIPAMD_E_NOT_FORGED = -2,    /* E::IMAGE_NOT_FORGED */
    DIPAMD_E_ARRAY_LENGTH = -3,  /* E::ARRAY_PARAMETER_WRONG_LENGTH */
    DIPAMD_E_SIZES = -4,         /* output sizes differ from the input's */
    DIPAMD_E_UNSUPPORTED = -5,   /* boundary condition / type / radius not supported */
    DIPAMD_E_HIP = -6,           /* a HIP runtime call failed */
    DIPAMD_E_ORDER = -7          /* "Gaussian FIR filter not implemented for order > 3" */
};

/* A strided image view (the dip::Image subset the path uses: image.h:458-491). */
typedef struct dipamd_image {
    void *origin;                       /* device pointer to the first pixel */
    int32_t dtype;                      /* DIPAMD_DT_* */
    int32_t ndim;                       /* 1..DIPAMD_MAX_DIMS */
    int64_t sizes[DIPAMD_MAX_DIMS];     /* dim 0 = x */
    int64_t strides[DIPAMD_MAX_DIMS];   /* in samples, may be negative */
} dipamd_image;

/* A 1-D filter (dip::OneDimensionalFilter, linear.h:84-90). `taps` is HOST memory,
 * in the user's (convolution) order; origin -1 = centre.  With is_complex != 0 the
 * filter is complex: `taps` holds ntaps (re, im) pairs (2 * ntaps doubles), as
 * OneDimensionalFilter::filter holds them with isComplex. */
typedef struct dipamd_filter {
    const double *taps;
    int32_t ntaps;
    int32_t origin;
    int32_t symmetry;                   /* DIPAMD_SYM_* */
    int32_t is_complex;
} dipamd_filter;

/* Replaces dip::GaussFIR (include/diplib/linear.h:300-318; src/linear/gauss.cpp:238-286)
 * and, with method "FIR", dip::Gauss (linear.h:452-460; derivative.cpp:62-86).
 * out must have the sizes of in; its dtype is the caller's choice (dip::Gauss picks
 * SuggestFlex(in): sfloat, or dfloat for 32-bit integer and dfloat input); a
 * different dtype is written with clamp_cast as for a protected dip::Image.
 * out may alias in (same view).  nsigmas, norders, nbc: 1 or in->ndim (nbc 0 = mirror,
 * norders 0 = all zero). */
int dipamd_gauss_fir(const dipamd_image *in, const dipamd_image *out, const double *sigmas, int32_t nsigmas,
                     const int32_t *orders, int32_t norders, const int32_t *bc, int32_t nbc, double truncation,
                     uint32_t flags, void *stream);

/* dip::GaussIIR (src/linear/gaussiir.cpp:741-806; include/diplib/linear.h:407-416):
 * the recursive Gaussian and its derivatives up to order 4, in double, bit-identical to
 * the reference.  sigmas / orders / filter_orders broadcast from 0 (defaults: sigma 1,
 * order 0, filter order 3 + order capped at 5), 1 or ndim elements; dims with sigma <= 0
 * or size 1 are not filtered.  design_method is one of DIPAMD_IIR_*.  `out` has the
 * sizes of `in` (normally dtype SuggestFlex(in)); it may be `in` itself. */
enum {
    DIPAMD_IIR_FORWARD_BACKWARD = 1, /* "forward backward" (Young and van Vliet 1995) */
    DIPAMD_IIR_DISCRETE_TIME_FIT = 2 /* "discrete time fit" (van Vliet et al. 1998), the default */
};
int dipamd_gauss_iir(const dipamd_image *in, const dipamd_image *out, const double *sigmas, int32_t nsigmas,
                     const int32_t *orders, int32_t norders, const int32_t *bc, int32_t nbc,
                     const int32_t *filter_orders, int32_t nfilter_orders, int32_t design_method,
                     double truncation, uint32_t flags, void *stream);

/* dip::GaussFT (src/linear/gauss.cpp:397-487; include/diplib/linear.h:355-364): the
 * Gaussian and its derivatives of any order as a product in the frequency domain
 * (rocFFT; single precision for 8-bit and sfloat images, double for dfloat, as the
 * reference's DFT).  Spatial domain in and out (real images): nbc == 0 transforms the
 * image as it is (periodic); otherwise the image is extended by the boundary condition
 * (the ones that copy samples) to a 2-3-5-smooth size holding its Gaussian border, and
 * cropped back.  dip::Gauss with method "FT" (and "best" for sigma < 0.8 or order > 3)
 * passes "mirror".  At most three dims may be filtered.
 *
 * With DIPAMD_FT_OUT_FREQUENCY the result is the spectrum: the centred DFT of the input
 * (origin in the middle, unscaled: dip::FourierTransform's default) times the
 * separable response; with DIPAMD_FT_IN_FREQUENCY the input is such a (complex)
 * spectrum, multiplied by the response and, unless the output is a spectrum too,
 * transformed back (centred inverse DFT, scaled by 1/N) into a complex image.  These
 * forms take no boundary condition (nbc == 0: the reference extends only spatial
 * inputs, to a larger size), any real or complex input, at most three dims, and a
 * complex `out` (DIPAMD_DT_SCOMPLEX / DCOMPLEX, SuggestComplex(in) normally).  Where
 * the reference transforms a real image whose real-to-complex dim has an odd length,
 * its spectrum deviates from the DFT (that path of its DFT is not exact); this
 * implementation gives the DFT there, as the reference does for complex input.
 * Results match the reference to FFT rounding. */
int dipamd_gauss_ft(const dipamd_image *in, const dipamd_image *out, const double *sigmas, int32_t nsigmas,
                    const int32_t *orders, int32_t norders, const int32_t *bc, int32_t nbc, double truncation,
                    uint32_t flags, void *stream);

/* The derivative family on the Gaussian FIR path (src/linear/derivative.cpp:120-521;
 * include/diplib/linear.h:666, 697, 862, 904): dip::Gradient, dip::GradientMagnitude,
 * dip::Hessian, dip::Laplace and the Dgg family with method "FIR" (or "best" resolving to FIR), each Gaussian derivative
 * computed as dipamd_gauss_fir computes it and combined in the reference's order and
 * arithmetic type.  The differentiated dims are those with process != 0 (NULL: all),
 * size > 1 and sigma > 0; sigmas apply to every dim (a dim that is not differentiated
 * is still smoothed).  `outs` holds dipamd_derivative_outputs(function, n) images
 * of the sizes of `in`, n = the number of differentiated dims: Gradient the partial
 * derivatives in dim order; Hessian the symmetric matrix as dip stores it, the n
 * diagonal elements first, then the upper triangle column by column; GradientMagnitude,
 * Laplace and the Dgg family one image, of type SuggestFlex(in). */
enum {
    DIPAMD_FN_GRADIENT = 0,
    DIPAMD_FN_GRADIENT_MAGNITUDE = 1,
    DIPAMD_FN_HESSIAN = 2,
    DIPAMD_FN_LAPLACE = 3,
    DIPAMD_FN_DGG = 4,               /* dip::Dgg (derivative.cpp:486-496) */
    DIPAMD_FN_LAPLACE_PLUS_DGG = 5,  /* dip::LaplacePlusDgg (:498-508) */
    DIPAMD_FN_LAPLACE_MINUS_DGG = 6, /* dip::LaplaceMinusDgg (:510-520) */
    DIPAMD_FN_STRUCTURE_TENSOR = 7   /* only through dipamd_structure_tensor */
};
int dipamd_derivative_outputs(int32_t function, int32_t ndiff);
int dipamd_derivative(int32_t function, const dipamd_image *in, const dipamd_image *outs, int32_t nouts,
                      const double *sigmas, int32_t nsigmas, const int32_t *bc, int32_t nbc, const uint8_t *process,
                      int32_t nprocess, double truncation, uint32_t flags, void *stream);

/* Replaces dip::StructureTensor without a mask (src/analysis/structure_tensor.cpp:32-57;
 * include/diplib/analysis.h:454): the gradient (gradient_sigmas, default 1), its outer
 * product g g^T as dip stores a symmetric matrix (diagonal first, then the upper
 * triangle column by column), each element smoothed by dip::Gauss with tensor_sigmas
 * (default 5), the FIR method.  outs: n(n+1)/2 images of type SuggestFlex(in). */
int dipamd_structure_tensor(const dipamd_image *in, const dipamd_image *outs, int32_t nouts,
                            const double *gradient_sigmas, int32_t ngs, const double *tensor_sigmas, int32_t nts,
                            const int32_t *bc, int32_t nbc, double truncation, uint32_t flags, void *stream);

/* The masked forms.  Replaces dip::NormalizedConvolution (src/linear/derivative.cpp:522-542;
 * include/diplib/linear.h) -- SafeDivide( Gauss( in x mask ), Gauss( mask ) ) -- and
 * dip::NormalizedDifferentialConvolution (derivative.cpp:544-600) along `dimension`:
 * SafeDivide( Derivative( in x mask ) - NC x Derivative( mask ), Gauss( mask ) ) with
 * NC the normalized convolution; a binary mask (DIPAMD_DT_BIN) zeroes the pixels outside
 * it, a real one multiplies; each image operation rounded in SuggestFlex(in), the
 * result of that type, bit-identical to the reference.  flags select the Gaussian method
 * (DIPAMD_GAUSS_IIR / DIPAMD_GAUSS_FT; FIR otherwise).  `out` may be `in`. */
int dipamd_normalized_convolution(const dipamd_image *in, const dipamd_image *mask, const dipamd_image *out,
                                  const double *sigmas, int32_t nsigmas, const int32_t *bc, int32_t nbc,
                                  double truncation, uint32_t flags, void *stream);
int dipamd_normalized_differential_convolution(const dipamd_image *in, const dipamd_image *mask,
                                               const dipamd_image *out, int32_t dimension, const double *sigmas,
                                               int32_t nsigmas, const int32_t *bc, int32_t nbc, double truncation,
                                               uint32_t flags, void *stream);
/* Replaces dip::StructureTensor with a mask (structure_tensor.cpp:46-56): the normalized
 * derivative of every dim as the gradient, then as dipamd_structure_tensor; outs:
 * ndim(ndim+1)/2 images of type SuggestFlex(in). */
int dipamd_structure_tensor_masked(const dipamd_image *in, const dipamd_image *mask, const dipamd_image *outs,
                                   int32_t nouts, const double *gradient_sigmas, int32_t ngs,
                                   const double *tensor_sigmas, int32_t nts, const int32_t *bc, int32_t nbc,
                                   double truncation, uint32_t flags, void *stream);

/* Replaces dip::SeparableConvolution (include/diplib/linear.h:114-130;
 * src/linear/convolution.cpp:319-421).  nfilters: 1 or ndim.  process: NULL (all) or
 * ndim bytes.  With a complex filter in the array the computation runs in
 * SuggestComplex(in) (scomplex, or dcomplex for 32-bit integer and double input), with
 * a complex input in SuggestFlex(in) (its own complex type); the output must then be
 * complex (a real output would receive the modulus in the reference; not taken here). */
int dipamd_separable_convolution(const dipamd_image *in, const dipamd_image *out, const dipamd_filter *filters,
                                 int32_t nfilters, const int32_t *bc, int32_t nbc, const uint8_t *process,
                                 uint32_t flags, void *stream);

/* The line buffers of the separable framework for user line filters -- the plug-in
 * operator API of dip::Framework::Separable (include/diplib/framework.h:589-669,
 * 760-770; framework_separable.cpp:146-331), whose device form is the template
 * dip::Framework::Separable in include/diplib_amd_framework.hpp.
 *
 * A line buffer holds every line of an image along `dim` interleaved: sample i of
 * line l at element i * nlines + l (lines numbered over the other dims, dim 0
 * fastest; nlines from dipamd_separable_lines_count), so one thread per line reads
 * and writes whole cache lines across a wave.  buffer_dt is DIPAMD_DT_SFLOAT or
 * DIPAMD_DT_DFLOAT.
 *
 * dipamd_separable_lines_gather replaces the framework's input copy (CopyBuffer with the
 * cast to the buffer type, copy_buffer.cpp:284-350) and boundary extension
 * (ExpandBuffer, copy_buffer.cpp:751-813, every condition): `buffer` receives
 * nlines * (sizes[dim] + 2 border) samples, sample i of a line being its extended
 * sample i - border.  dipamd_separable_lines_scatter replaces the output copy
 * (CopyBuffer with clamp_cast to the image's type): `buffer` holds nlines * sizes[dim]
 * filtered samples, written to `out` along `dim`. */
int64_t dipamd_separable_lines_count(const dipamd_image *im, int32_t dim);
int dipamd_separable_lines_gather(const dipamd_image *in, int32_t dim, int32_t border, int32_t bc, int32_t buffer_dt,
                                  void *buffer, void *stream);
int dipamd_separable_lines_scatter(const void *buffer, int32_t buffer_dt, const dipamd_image *out, int32_t dim,
                                   void *stream);

/* Replaces dip::MakeHalfGaussian (include/diplib/generation.h:638-653;
 * src/linear/gauss.cpp:151-165): writes r+1 host doubles, x = 0 LAST, returns the
 * count (or a negative error code). */
int dipamd_make_half_gaussian(double sigma, int32_t order, double truncation, int32_t is_double, double *taps,
                              int32_t capacity);

/* ---- batched form (no reference counterpart: a per-image loop of dip::Gauss) ----
 * Many 2-D images of uneven sizes and per-image sigmas in ONE persistent kernel that
 * pulls tiles from a work queue.  A plan fixes the device pointers, sizes, sigmas and
 * tile shape; running it is graph-capturable (no allocation, no host sync). */
typedef struct dipamd_batch_image {
    const void *in;          /* device pointer, dtype in_dtype */
    void *out;               /* device pointer, dtype out_dtype (sfloat or dfloat) */
    int32_t width, height;   /* dip sizes (x, y) */
    int64_t in_stride_y;     /* samples between rows (x stride is 1) */
    int64_t out_stride_y;
    double sigma_x, sigma_y;
    int32_t order_x, order_y;
} dipamd_batch_image;

typedef struct dipamd_batch_plan dipamd_batch_plan;

/* tile_config: -1 = autotuned choice, else an index < dipamd_batch_num_configs().
 * An image's output must not overlap its own input (DIPAMD_E_PARAMETER otherwise):
 * the tiles of one image run concurrently and read halo rows other tiles write. */
int dipamd_batch_create(dipamd_batch_plan **plan, const dipamd_batch_image *images, int32_t n, int32_t in_dtype,
                        int32_t out_dtype, const int32_t *bc, int32_t nbc, double truncation, uint32_t flags,
                        int32_t tile_config, void *stream);
int dipamd_batch_run(dipamd_batch_plan *plan, void *stream);
/* Captures dipamd_batch_run into a hipGraph owned by the plan (once), then launches it. */
int dipamd_batch_run_graph(dipamd_batch_plan *plan, void *stream);
/* The same images through a per-image loop of dipamd_gauss_fir (the plain drop-in):
 * the baseline the batched kernel is measured against, and its reference result. */
int dipamd_batch_run_per_image(dipamd_batch_plan *plan, void *stream);
int dipamd_batch_destroy(dipamd_batch_plan *plan);

/* No counterpart in the reference.  For callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc):
 * the batched form in one call, with no plan object and no allocation -- the work
 * queue, its counters, the per-image descriptors and the taps are written into one
 * fixed device array of the library (DIPAMD_BATCH_NOALLOC_SCRATCH_BYTES per GPU, held
 * until the process ends), the counters cleared by a kernel, then the persistent
 * kernel runs.  Same arguments and results as dipamd_batch_create + dipamd_batch_run
 * (tile_config -1 takes the tuned choice of an earlier plan with the same key, else
 * the default tile).  A batch whose tables exceed the array is rejected with
 * DIPAMD_E_UNSUPPORTED (roughly 100 000 work items, or 60 000 images).  Every call on
 * a GPU shares that GPU's array: calls that overlap on one GPU (different streams, or
 * host threads) race.  The descriptor upload reads host memory before the call returns. */
#define DIPAMD_BATCH_NOALLOC_SCRATCH_BYTES (8u << 20)
int dipamd_batch_gauss_noalloc(const dipamd_batch_image *images, int32_t n, int32_t in_dtype, int32_t out_dtype,
                               const int32_t *bc, int32_t nbc, double truncation, uint32_t flags,
                               int32_t tile_config, void *stream);
int dipamd_batch_num_configs(void);
/* Number of work-queue tiles and the config index of a plan. */
int dipamd_batch_info(const dipamd_batch_plan *plan, int64_t *ntiles, int32_t *config);

/* ---- backward (no reference counterpart: the adjoint of dip::Gauss FIR) ----
 * grad_in = J^T grad_out, the exact adjoint of the boundary-extended separable
 * convolution (flipped taps, halo contributions folded back by the boundary
 * condition); grad_sigma[d] (DEVICE doubles, ndim of them, overwritten) =
 * <grad_out, d out / d sigma_d>, differentiating the normalised discrete taps at a
 * fixed radius.  x and grad_out are sfloat or dfloat, grad_in has the dtype of
 * grad_out; either output may be NULL (origin NULL) to skip it.  With
 * DIPAMD_GRAD_SIGMA_TOTAL in flags, grad_sigma is one double: the sum over the dims. */
int dipamd_gauss_fir_backward(const dipamd_image *x, const dipamd_image *grad_out, const dipamd_image *grad_in,
                              double *grad_sigma, const double *sigmas, int32_t nsigmas, const int32_t *orders,
                              int32_t norders, const int32_t *bc, int32_t nbc, double truncation, uint32_t flags,
                              void *stream);

/* The forward with its intermediate kept for the backward (no reference counterpart; for
 * autograd callers of images with three or more filtered dims): dipamd_gauss_fir, and when
 * its first pass is the fused one over dims 0 and 1 and a further dim follows, that pass
 * writes its result (x filtered along dims 0 and 1, the buffer type SuggestFlex(in), the
 * sizes of in) into `keep` and *kept = 1; else *kept = 0 and keep is untouched.  keep
 * must not overlap in or out. */
int dipamd_gauss_fir_keep(const dipamd_image *in, const dipamd_image *out, const double *sigmas, int32_t nsigmas,
                          const int32_t *orders, int32_t norders, const int32_t *bc, int32_t nbc, double truncation,
                          uint32_t flags, const dipamd_image *keep, int32_t *kept, void *stream);

/* dipamd_gauss_fir_backward taking that intermediate (NULL: computed again), which the
 * sigma derivative of the dims after 0 and 1 needs (no reference counterpart). */
int dipamd_gauss_fir_backward_kept(const dipamd_image *x, const dipamd_image *grad_out, const dipamd_image *grad_in,
                                   double *grad_sigma, const double *sigmas, int32_t nsigmas,
                                   const int32_t *orders, int32_t norders, const int32_t *bc, int32_t nbc,
                                   double truncation, uint32_t flags, const dipamd_image *kept, void *stream);

/* ---- schedule tuning (no reference counterpart) ----
 * Measures, on synthetic planes of this device, every candidate schedule of the plain
 * path per (source type uint8 / sfloat / dfloat, radius class, plane size class): one
 * pass per axis, the fused kernel with each tile config, the uint8 streaming kernel,
 * and the kernel of a single strided pass; the fastest of each goes to the persistent
 * tuning cache (DIPAMD_TUNE_FILE, else tune_cache.txt beside the library), which the
 * planner consults on every call.  flags: DIPAMD_FAST tunes the fast arithmetic;
 * DIPAMD_TUNE_QUICK uses smaller planes and fewer repetitions.  Allocates its own
 * planes and synchronises `stream`; results do not depend on the choices. */
#define DIPAMD_TUNE_QUICK 256u
int dipamd_autotune(uint32_t flags, void *stream);

/* Radius of the tap array used for (sigma, order) as GaussFIR computes it (gauss.cpp:38-51). */
int dipamd_gauss_radius(double sigma, int32_t order, double truncation, int32_t is_double);

/* Kernel trace (no reference counterpart; the reference times with dip::testing::Timer,
 * testing.h:221-265): while enabled, every kernel the library launches is bracketed by
 * HIP events on its launch stream.  dipamd_trace_enable(1) clears the record list. */
typedef struct dipamd_trace_record {
    char name[48];     /* "fused2d", "line", "sweep", "cast_in", "cast_out", ... */
    double ms;         /* event-timed duration of the launch */
    int64_t bytes;     /* algorithmic bytes of the launch: samples read + written once */
    int64_t ops;       /* algorithmic VALU lane-ops of the launch: per output sample and filtered
                          axis, 3r+1 for a symmetric filter of radius r in exact mode (add, mul,
                          add per tap pair and the centre multiply), 2r+1 in fast mode */
} dipamd_trace_record;
int dipamd_trace_enable(int32_t on);
/* Waits for the recorded events and copies up to `capacity` records; returns the count. */
int dipamd_trace_read(dipamd_trace_record *records, int32_t capacity);

/* Thread-local message of the last failed call on this thread. */
const char *dipamd_last_error(void);
/* Library version string. */
const char *dipamd_version(void);

#ifdef __cplusplus
}
#endif
#endif /* DIPAMD_H */
