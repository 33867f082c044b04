/*
 * enblend_amd -- C ABI of the MI355X (gfx950) implementation of the
 * enblend/enfuse multiresolution path.
 *
 * Every entry point takes plain device pointers, sizes and pitches, returns
 * an int status (ENBL_OK or a negative ENBL_E_* code, never throws) and is
 * ordered on the context's HIP stream.  Images are interleaved like the
 * reference's vigra::RGBValue images; pyramids are planar per channel.  All
 * pitches are in elements of the buffer's own type.
 *
 * The reference exposes this path as C++ templates, not a C ABI; each
 * declaration names the template it replaces (jackmitch/enblend-enfuse,
 * VERSION 4.3-af6cddaed31b, paths relative to the repository root).
 *
 * Type codes follow the reference's numeric traits (src/numerictraits.h:168-189):
 *   8-bit images   -> ENBL_I16 pyramid, 7 fraction bits; mask pyramid ENBL_I16, 7
 *   16-bit images  -> ENBL_I32 pyramid, 7 fraction bits; mask pyramid ENBL_I32, 15
 *   float images   -> ENBL_F64 pyramid (log transform); mask pyramid ENBL_I32, 15
 */
#ifndef ENBLEND_AMD_H
#define ENBLEND_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define ENBL_OK 0
#define ENBL_E_PRECONDITION (-1) /* vigra_precondition violated in the reference */
#define ENBL_E_SIZE (-2)         /* inconsistent level sizes */
#define ENBL_E_ARG (-3)          /* null pointer, bad pitch or alignment */
#define ENBL_E_UNSUPPORTED (-4)  /* type / channel combination not provided */
#define ENBL_E_OOM (-5)          /* device allocation failed (std::bad_alloc) */
#define ENBL_E_HIP (-6)          /* HIP runtime error */
#define ENBL_E_IO (-7)           /* file missing, unreadable or malformed */

/* pixel type codes */
#define ENBL_I16 0
#define ENBL_I32 1
#define ENBL_F64 2
#define ENBL_U8 10
#define ENBL_U16 11
#define ENBL_F32 12
#define ENBL_U32 13   /* only as the component type of enbl_checkpoint's output */

/* grayscale projectors of enfuse's exposure and contrast weights */
#define ENBL_GRAY_AVERAGE 0
#define ENBL_GRAY_LSTAR 1
#define ENBL_GRAY_PLSTAR 2
#define ENBL_GRAY_LIGHTNESS 3
#define ENBL_GRAY_VALUE 4
#define ENBL_GRAY_ANTI_VALUE 5
#define ENBL_GRAY_LUMINANCE 6
#define ENBL_GRAY_MIXER 7

/* dither modes of the pyramid-to-image conversion */
#define ENBL_DITHER_NONE 0 /* round half up inside the dither band */
#define ENBL_DITHER_DET 1  /* counter-based draw keyed by (seed, x, y, channel) */

typedef struct enbl_context enbl_context;
typedef struct enbl_pyramid enbl_pyramid;

/* ---- context --------------------------------------------------------- */

/* Create a context on HIP device `device` with its own default-priority
 * stream. */
int enbl_context_create(int device, enbl_context** out);
int enbl_context_destroy(enbl_context* ctx);
/* Order all later work on an external HIP stream (NULL restores the
 * context's own stream). */
int enbl_context_set_stream(enbl_context* ctx, void* hip_stream);
int enbl_context_get_stream(const enbl_context* ctx, void** hip_stream);
int enbl_synchronize(enbl_context* ctx);
/* Kernel selection of reduce/expand for every context of the process:
 * AUTO picks the one-thread-per-output kernels for small levels and the
 * streaming kernels for large ones; DIRECT and STREAMING force one family
 * (used by the parity tests and the tile autotuner).  Not thread-safe. */
#define ENBL_KERNELS_AUTO 0
#define ENBL_KERNELS_DIRECT 1
#define ENBL_KERNELS_STREAMING 2
int enbl_set_kernel_policy(int policy);
/* Strip shape of the fused Laplacian and collapse passes for every context of
 * the process: rows per strip such that about target_waves waves cover a
 * level, at most max_rows (defaults 8192, 32).  A tuning knob with no
 * counterpart in the reference; results do not depend on it.  Not
 * thread-safe. */
int enbl_set_strip_shape(int target_waves, int max_rows);
/* Message of the last failing call on this thread. */
const char* enbl_last_error(void);
/* Kernel timing: when enabled every launch is bracketed by HIP events on the
 * context's stream; the report synchronises and writes one line per kernel
 * name: "name total_ms launches credited_bytes own_bytes".  credited_bytes
 * are the algorithmic bytes of the reference's unfused operations the kernel
 * performs (operands at the reference's storage types, each read once and
 * each result written once); own_bytes are the fused kernel's own minimum
 * traffic (each of its inputs read once, each of its outputs written once),
 * equal to credited_bytes for a kernel that replaces one operation. */
int enbl_profile_enable(enbl_context* ctx, int on);
int enbl_profile_report(enbl_context* ctx, char* buf, size_t buflen);

/* Device memory of the context's device, for hosts that do not link HIP
 * themselves (e.g. the C++ drop-in headers).  enbl_copy is ordered on the
 * context's stream and returns after the copy completed. */
int enbl_device_alloc(enbl_context* ctx, size_t bytes, void** out);
int enbl_device_free(enbl_context* ctx, void* p);
int enbl_copy(enbl_context* ctx, void* dst, const void* src, size_t bytes);

/* ---- single-level operations ---------------------------------------- */

/* Replaces reduce<SKIPSMImagePixelType, SKIPSMAlphaPixelType>(wraparound,
 * src, mask, dest, destMask) -- src/pyramid.h:166-573 -- when src_alpha is
 * non-NULL, and reduce<SKIPSMImagePixelType>(wraparound, src, dest) --
 * src/pyramid.h:601-848 -- when it is NULL.  `channels` planes of `ptype`
 * (plane strides splane/dplane); dst is ceil(sw/2) x ceil(sh/2).
 * ENBL_E_PRECONDITION when sw < 2 or sh < 2 ("src image too small in reduce"). */
int enbl_reduce(enbl_context* ctx, int ptype, int channels, int wraparound,
                const void* src, int sw, int sh, long spitch, long splane,
                const uint8_t* src_alpha, long sapitch,
                void* dst, int dw, int dh, long dpitch, long dplane,
                uint8_t* dst_alpha, long dapitch);

/* Replaces expand<SKIPSMImagePixelType>(add, wraparound, src, dest) --
 * src/pyramid.h:1102-1419 with the combine functors of :1425-1457:
 * add != 0: dest = fromPromote(dest + expand(src)) (collapse),
 * add == 0: dest = dest - expand(src) (Laplacian).  dw in {2sw-1, 2sw}. */
int enbl_expand(enbl_context* ctx, int ptype, int channels, int add, int wraparound,
                const void* src, int sw, int sh, long spitch, long splane,
                void* dst, int dw, int dh, long dpitch, long dplane);

/* ---- pyramids --------------------------------------------------------- */

/* Device pyramid of `levels` levels; level l is ceil(w/2^l) x ceil(h/2^l)
 * (src/pyramid.h:1509-1510), `channels` planes per level.  Replaces the
 * std::vector<PyramidImageType*> the reference allocates. */
int enbl_pyramid_create(enbl_context* ctx, int ptype, int channels, int w, int h, int levels,
                        enbl_pyramid** out);
int enbl_pyramid_destroy(enbl_pyramid* p);
int enbl_pyramid_levels(const enbl_pyramid* p);
int enbl_pyramid_level(const enbl_pyramid* p, int level, void** data, int* w, int* h,
                       long* pitch, long* plane);

/* Copy level `level` between the pyramid and a packed buffer of
 * channels x h x w elements (host or device memory).  to_pyramid != 0
 * writes the buffer into the level.  Synchronous with respect to the host.
 * Replaces direct element access to the reference's BasicImage levels. */
int enbl_pyramid_copy_level(enbl_pyramid* p, int level, void* buf, int to_pyramid);

/* Level count of roiBounds (src/bounds.h:121-133): halve the short side of
 * the ROI until it is <= 8, starting from minimum_levels, capped at 29;
 * exact_levels > 0 caps it, exact_levels < 0 subtracts (--levels). */
int enbl_auto_levels(int w, int h, int minimum_levels, int exact_levels);

/* Replaces copyToPyramidImage<SrcImageType, PyramidImageType, IntBits,
 * FracBits>(src, dest) -- src/fixmath.h:2018-2067, identity colour space --
 * into level 0 of `out`: integer images v << frac_bits (ConvertScalarTo-
 * PyramidFunctor, src/fixmath.h:116-206), float images into integer
 * pyramids fromRealPromote(v * 2^frac_bits), float images into double
 * pyramids 1 + log1p(v) for v >= 0, else 1 / (1 - v).  `img` is an
 * interleaved device image with the pyramid's channel count. */
int enbl_copy_to_pyramid(enbl_context* ctx, int itype, int frac_bits, const void* img, long ipitch,
                         enbl_pyramid* out);

/* Replaces gaussianPyramid<...>(numLevels, wraparound, src, alpha) --
 * src/pyramid.h:1467-1559 (alpha != NULL) and :1562-1636 (alpha == NULL) --
 * with copyToPyramidImage (src/fixmath.h:2018-2067) for level 0.  `img` is
 * an interleaved device image of `itype` (ENBL_U8, ENBL_U16, ENBL_F32) with
 * the pyramid's channel count; frac_bits is PyramidFractionBits. */
int enbl_gaussian_pyramid(enbl_context* ctx, int wraparound, int itype, int frac_bits,
                          const void* img, long ipitch, const uint8_t* alpha, long apitch,
                          enbl_pyramid* out);

/* No counterpart in the reference: enbl_gaussian_pyramid for callers that
 * cannot allocate device memory (code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc).  Same arguments, results and
 * errors; the per-level alpha lives in a 16 MiB device array that the
 * library holds on every GPU it runs on until the process ends, so the call
 * allocates nothing.  ENBL_E_SIZE when the level-1 alpha ping-pong,
 * 2 * round_up(level-1 width, 256) * level-1 height bytes, exceeds it
 * (e.g. above 8192 x 4096 images).  All calls on one GPU share that array:
 * calls that overlap on one GPU race.  No profile events are recorded. */
int enbl_gaussian_pyramid_noalloc(enbl_context* ctx, int wraparound, int itype, int frac_bits,
                                  const void* img, long ipitch, const uint8_t* alpha, long apitch,
                                  enbl_pyramid* out);

/* Replaces laplacianPyramid<...>(exportName, numLevels, wraparound, src,
 * alpha) -- src/pyramid.h:1646-1747. */
int enbl_laplacian_pyramid(enbl_context* ctx, int wraparound, int itype, int frac_bits,
                           const void* img, long ipitch, const uint8_t* alpha, long apitch,
                           enbl_pyramid* out);

/* The levels above 0 of a pyramid whose level 0 is already written (for
 * instance by enbl_rgb_to_jch): the reduces of gaussianPyramid
 * (src/pyramid.h:1467-1559; alpha may be NULL) and, with laplacian != 0,
 * the expand-subtracts of laplacianPyramid (src/pyramid.h:1646-1747). */
int enbl_pyramid_build_levels(enbl_context* ctx, enbl_pyramid* p, int wrap, const uint8_t* alpha, long apitch,
                              int laplacian);

/* Replaces collapsePyramid<SKIPSMImagePixelType>(wraparound, p) --
 * src/pyramid.h:1757-1784.  In place; level 0 holds the result. */
int enbl_collapse_pyramid(enbl_context* ctx, int wraparound, enbl_pyramid* p);

/* Replaces blend(maskGP, whiteLP, blackLP, maskPyramidWhiteValue) --
 * src/blend.h:81-121 with CartesianBlendFunctor (:41-76).  Writes black_lp. */
int enbl_blend(enbl_context* ctx, const enbl_pyramid* mask_gp, const enbl_pyramid* white_lp,
               enbl_pyramid* black_lp, double white_value);

/* Replaces copyFromPyramidImageIf<...>(src, mask, dest) --
 * src/fixmath.h:2157-2294 with SpecializedConvertPyramidToScalarFunctor
 * (:209-358) for the identity colour space: writes level `level` of `p` into
 * the interleaved device image `img` of `itype` where mask != 0.  Integer
 * pyramids dither inside the band [1/4, 3/4) of a unit per `dither_mode`. */
int enbl_copy_from_pyramid_if(enbl_context* ctx, const enbl_pyramid* p, int level, int frac_bits,
                              const uint8_t* mask, long mpitch, void* img, int itype, long ipitch,
                              int dither_mode, uint64_t seed);

/* ---- enfuse weights and fusion -------------------------------------- */

/* built-in exposure weight functions (src/exposure_weight.h:44-120) */
#define ENBL_EXPOSURE_GAUSSIAN 0
#define ENBL_EXPOSURE_LORENTZIAN 1
#define ENBL_EXPOSURE_HALFSINE 2
#define ENBL_EXPOSURE_FULLSINE 3
#define ENBL_EXPOSURE_BISQUARE 4

/* enfuse weighting options (src/enfuse.cc:118-146; defaults exposure 1.0,
 * optimum 0.5, width 0.2, gaussian, contrast window 5, entropy window 3). */
typedef struct enbl_enfuse_params {
    double w_exposure, w_saturation, w_contrast, w_entropy;
    double exposure_optimum, exposure_width;
    int exposure_function;        /* ENBL_EXPOSURE_* */
    int contrast_window;          /* odd, >= 3 */
    int entropy_window;           /* odd, >= 3, window area <= 121 */
    int hard_mask;                /* 0: soft masks; 1: --hard-mask, the first image of largest
                                     weight takes a pixel whole (src/enfuse.h:1402-1443) */
    /* Host table of ExposureWeight::weight(code / max) for every gray code
     * of an integer image (256 or 65536 doubles), e.g. from a user plugin
     * (src/exposure_weight_base.h:29-88).  NULL: the built-in function.
     * With w_exposure > 0 the table is checked as check_weight_function
     * checks the plugin (src/exposure_weight.cc:237-264): a negative entry,
     * an entry above 1 or all entries zero give ENBL_E_PRECONDITION (enfuse
     * exits there, src/enfuse.cc:1920-1945).  A table for an ENBL_F32
     * image gives ENBL_E_UNSUPPORTED (the reference evaluates the plugin per
     * float pixel, src/enfuse.h:637-640). */
    const double* exposure_table;
    /* Contrast by Laplacian-of-Gaussian edges (--contrast-edge-scale,
     * src/enfuse.h:1020-1093) when edge_scale > 0: the edge scale, the
     * local-contrast-enhancement scale and factor (0: none; lce_scale in
     * pixels, lce_factor as a fraction), and the minimum curvature as the
     * reference instantiates it for the image's scalar type
     * (static_cast<double>(MinCurvature.instantiate<ScalarType>())).  Zero
     * edge_scale: the local standard deviation. */
    double edge_scale, lce_scale, lce_factor, min_curvature;
    /* --exposure-cutoff (CutoffExposureFunctor, src/enfuse.h:654-720, with
     * the default projectors: minimum channel >= lower, maximum channel <=
     * upper, else weight 0) and --entropy-cutoff (the image clamped before
     * the local entropy, src/enfuse.h:1145-1188): 1 when the reference finds
     * the cutoff effective, with the cutoffs instantiated for the image's
     * scalar type. */
    int exposure_cutoff, entropy_cutoff;
    double exposure_lower_cutoff, exposure_upper_cutoff;
    double entropy_lower_cutoff, entropy_upper_cutoff;
    /* --gray-projector (MultiGrayscaleAccessor, src/mga.h:60-250) of the
     * exposure and contrast weights: ENBL_GRAY_*; the channel mixer's
     * weights as given (divided by their sum; negative or zero sum is
     * ENBL_E_ARG). */
    int gray_projector;
    double mixer_red, mixer_green, mixer_blue;
} enbl_enfuse_params;

/* Replaces enfuseMask<ImageType, AlphaType, MaskType>(src, mask, result) --
 * src/enfuse.h:962-1212 -- for an interleaved RGB device image: the float
 * weight image (exposure, contrast, saturation, entropy, accumulated in that
 * order where alpha != 0, zero elsewhere). */
int enbl_enfuse_mask(enbl_context* ctx, int itype, int w, int h, const void* img, long ipitch,
                     const uint8_t* alpha, long apitch, const enbl_enfuse_params* prm,
                     float* mask, long mpitch);

/* Replaces the fusion of enfuseMain<ImagePixelType> -- src/enfuse.h:1217-1609,
 * soft masks, identity colour space -- for K interleaved RGB device images of
 * one size: weights, union alpha, normalisation, Laplacian pyramids, mask
 * Gaussian pyramids over the union alpha, multiply, accumulate, collapse and
 * the dithered masked copy.  imgs/alphas/pitches are host arrays of K
 * entries.  `out` receives the fused image where the union alpha is set
 * (other pixels untouched), `out_alpha` the union alpha. */
int enbl_enfuse(enbl_context* ctx, int itype, int w, int h, int K,
                const void* const* imgs, const long* ipitches,
                const uint8_t* const* alphas, const long* apitches,
                const enbl_enfuse_params* prm, int levels, int wraparound,
                void* out, long opitch, uint8_t* out_alpha, long oapitch,
                int dither_mode, uint64_t seed);

/* ---- batched enfuse ------------------------------------------------------ */

/* One exposure stack of a batched enfuse call: the arguments of one
 * enbl_enfuse call (host arrays of K device pointers / pitches). */
typedef struct enbl_stack {
    int w, h;                       /* image size */
    int K;                          /* images in the stack, >= 1 */
    int levels;                     /* pyramid levels, 1..29 */
    const void* const* imgs;        /* K interleaved RGB device images of itype */
    const long* ipitches;           /* K row pitches (elements) */
    const uint8_t* const* alphas;   /* K 0/255 device alpha planes */
    const long* apitches;           /* K alpha row pitches (bytes) */
    void* out; long opitch;         /* fused image, written where the union alpha is set */
    uint8_t* out_alpha; long oapitch;  /* union alpha */
    uint64_t seed;                  /* dither seed */
} enbl_stack;

/* Batched form of enbl_enfuse -- no counterpart in the reference, which runs
 * enfuseMain (src/enfuse.h:1217-1609) once per stack: `nstacks` stacks of
 * any sizes, depths and level counts fused with the same parameters, with
 * results identical to one enbl_enfuse call per stack.  8/16-bit and float
 * images without wraparound take the batched pyramid schedule (one
 * persistent launch per pyramid level over all stacks, replayed from a HIP
 * graph when the same batch comes again); their weights are batched too for
 * integer images with exposure and saturation weights, and made stack by
 * stack otherwise.  A blend colour space or wraparound runs enbl_enfuse per
 * stack.  The context keeps the intermediate pyramids and the captured graph
 * until it is destroyed. */
int enbl_enfuse_batch(enbl_context* ctx, int itype, int nstacks, const enbl_stack* stacks,
                      const enbl_enfuse_params* prm, int wraparound, int dither_mode);

/* Launch shape of the batched passes: rows of a 64-column tile and resident
 * blocks per compute unit of the persistent grid, per pass kind (0 weights,
 * 1 reduce, 2 merge + collapse).  Zeroes select the built-in defaults. */
typedef struct enbl_batch_tuning {
    int rows[3];
    int blocks_per_cu[3];
} enbl_batch_tuning;

int enbl_batch_get_tuning(const enbl_context* ctx, enbl_batch_tuning* t);
int enbl_batch_set_tuning(enbl_context* ctx, const enbl_batch_tuning* t);

/* Autotuner: times every candidate launch shape of each pass kind on this
 * batch (HIP events, `repeats` runs each, best-of), installs the fastest in
 * the context and returns it.  The batch's outputs are written as by
 * enbl_enfuse_batch. */
int enbl_batch_autotune(enbl_context* ctx, int itype, int nstacks, const enbl_stack* stacks,
                        const enbl_enfuse_params* prm, int dither_mode, int repeats,
                        enbl_batch_tuning* best);

/* ---- differentiable enfuse (relaxed variant; not reference parity) ------- */

/* The reference's arithmetic cannot be differentiated: its mask pyramid is
 * fixed point and its float exposure weight is constant (y = gray / FLT_MAX,
 * src/enfuse.h:637).  The relaxed variant keeps enfuseMain's dataflow
 * (src/enfuse.h:1217-1609) in double throughout: y = gray, smooth saturation
 * (max-min)/(max+min+1e-6), contrast sqrt(var + 1e-12) on the n x n window,
 * weight w_e E + w_c sd + w_s s + 1e-12, normalisation, Gaussian pyramids of
 * the log-transformed images and of the normalised weights, merge of the
 * Laplacian levels, collapse, inverse transform.  Parameters: */
typedef struct enbl_diff_params {
    double w_exposure, w_saturation, w_contrast;
    double exposure_optimum, exposure_width;   /* Gaussian exposure weight */
} enbl_diff_params;

/* Forward: K interleaved RGB double device images (K*h*w*3) -> out (h*w*3). */
int enbl_diff_enfuse_forward(enbl_context* ctx, int w, int h, int K, const double* imgs,
                             const enbl_diff_params* prm, int levels, int contrast_window, double* out);

/* Adjoint: for the output gradient gout (h*w*3) writes d/d imgs (K*h*w*3) and
 * d/d {w_exposure, w_saturation, w_contrast, exposure_optimum,
 * exposure_width} (5 doubles, device).  Recomputes the forward pyramids; the
 * Laplacian levels are evaluated on the fly, never stored. */
int enbl_diff_enfuse_backward(enbl_context* ctx, int w, int h, int K, const double* imgs,
                              const enbl_diff_params* prm, int levels, int contrast_window,
                              const double* gout, double* gimgs, double* gparams);

/* Transposed pyramid operators on one double plane (accumulate into the
 * output): adjoint of expand (src/pyramid.h:873-1419, no wraparound) maps a
 * fine gradient to the coarse level, adjoint of reduce with an all-set alpha
 * (src/pyramid.h:166-573) maps a coarse gradient to the fine level. */
int enbl_expand_adjoint(enbl_context* ctx, const double* g, long gpitch, int fw, int fh,
                        double* gc, long cpitch, int cw, int ch);
int enbl_reduce_adjoint(enbl_context* ctx, const double* g, long gpitch, int dw, int dh,
                        double* gf, long fpitch, int sw, int sh);

/* ---- seam mask ------------------------------------------------------------ */

/* Replaces vigra::omp::distanceTransform (src/openmp_vigra.h:442-470, the
 * OpenMP build's Felzenszwalb-Huttenlocher transform, src/openmp_vigra.h:
 * 244-438): dst = distance of every pixel to the nearest non-zero pixel of
 * src, in float with the reference's operation order; norm 1 = Manhattan,
 * 2 = Euclidean (square root of the squared distance).  Pixels of an image
 * without any non-zero pixel get the transform of FLT_MAX.  Device buffers,
 * pitches in elements. */
int enbl_distance_transform(enbl_context* ctx, int w, int h, const uint8_t* src, long spitch, int norm,
                            float* dst, long dpitch);

/* The quotient of the envelope transform's intersection,
 * (sum_q - sum_v) / (2 * (q - v)) (src/openmp_vigra.h:322, :328: a float
 * divided by an int converted to float), as the row pass of
 * enbl_distance_transform computes it on the device: out[i] = num[i] /
 * (float)den[i] for i < count, which must equal the IEEE quotient bit for
 * bit.  Exposed so that equality can be checked on the device; device
 * buffers, stream-ordered. */
int enbl_nft_quotients(enbl_context* ctx, const float* num, const int* den, float* out, long count);

/* Replaces nearestFeatureTransform with open boundaries (src/nearest.h:
 * 424-552), the primary seam generator of enblend's fine mask without
 * optimisation (src/mask.h:1101-1104, :1130-1135, :1176-1179): the
 * distance transforms of white-minus-black and black-minus-white alpha
 * (saturating), mask = 255 where the pixel is nearer to the white-only than
 * to the black-only pixels, else 0.  The mask is written in any case;
 * ENBL_E_PRECONDITION when neither image has 2 * 2 * (w + h) exclusive
 * pixels (the reference's "excessive image overlap" exit). */
int enbl_nft_mask(enbl_context* ctx, int w, int h, const uint8_t* white_alpha, long wapitch,
                  const uint8_t* black_alpha, long bapitch, int norm, uint8_t* mask, long mpitch);

/* nearestFeatureTransform with strip boundaries (src/nearest.h:474-520):
 * boundary 1 = HorizontalStrip (enblend --wrap=horizontal), 2 =
 * VerticalStrip, 3 = DoubleStrip.  Each difference image is tiled twice
 * along the wrapped axes (quadruple_image), transformed with open
 * boundaries and cut back (quater_image: position p + n for p < n / 2 on a
 * doubled axis of length n); the overlap tally is that of the original
 * images.  Same results and errors as enbl_nft_mask otherwise. */
int enbl_nft_mask_periodic(enbl_context* ctx, int w, int h, const uint8_t* white_alpha, long wapitch,
                           const uint8_t* black_alpha, long bapitch, int norm, int boundary, uint8_t* mask,
                           long mpitch);

/* ---- blend colour spaces ------------------------------------------------- */

/* --blend-colorspace (BlendColorspace, src/enfuse.cc:110, :1628-1650,
 * src/enblend.cc:124): the space the pyramids of RGB images are built,
 * blended and collapsed in.  enfuse and enblend choose CIELUV by themselves
 * for integer images that carry an ICC profile (src/enfuse.cc:2380-2383). */
#define ENBL_COLORSPACE_IDENTITY 0
#define ENBL_COLORSPACE_CIELAB 1
#define ENBL_COLORSPACE_CIELUV 2
#define ENBL_COLORSPACE_CIECAM 3   /* CIECAM02 JCh */

/* Sets the context's blend colour space for enbl_enblend_pair,
 * enbl_enblend_step, enbl_enblend_layers, enbl_enfuse and
 * enbl_enfuse_batch (default identity).  With CIELAB or CIELUV those convert
 * 8- and 16-bit RGB images as the reference's
 * ConvertVectorTo{Lab,Luv}PyramidFunctor and Convert{Lab,Luv}PyramidTo-
 * VectorFunctor do (src/fixmath.h:429-1330; sRGB input profile, the
 * reference's assumption for images without an ICC profile); with CIECAM as
 * ConvertVectorToJCHPyramidFunctor / ConvertJCHPyramidToVectorFunctor
 * (src/fixmath.h:1336-1990: CIECAM02 under the viewing conditions of
 * src/enblend.cc:2235-2244, JCh in the pyramid as fixed point); enfuse's
 * weights stay computed on the RGB images.  Float RGB images with a
 * non-identity space give ENBL_E_UNSUPPORTED; one-channel images are not
 * converted (the reference's scalar path ignores the space). */
int enbl_set_blend_colorspace(enbl_context* ctx, int space);

/* The forward conversion alone: interleaved 8- or 16-bit RGB `img` into the
 * scaled blend-space image `out` of the same type (L' = round(s L / 100),
 * a' = round(s (200 + a) / 400), s = 2^(IntegerBits - 2) = 128 for 8-bit,
 * 32768 for 16-bit images; PyramidScale, src/fixmath.h:430-548), the values
 * the pyramid's level 0 holds before the fraction shift. */
int enbl_rgb_to_blend_space(enbl_context* ctx, int space, int itype, int w, int h, const void* img, long ipitch,
                            void* out, long opitch);

/* The inverse conversion alone: a scaled blend-space image (as the integer
 * output conversion of the collapsed pyramid leaves it) into RGB `img`,
 * fromRealPromote(max * rgb), where mask != 0; other pixels are kept.
 * Out-of-gamut colours are clamped (the reference polishes them with a
 * numeric minimiser, src/fixmath.h:835-1045). */
int enbl_blend_space_to_rgb_if(enbl_context* ctx, int space, int itype, int w, int h, const void* src,
                               long spitch, const uint8_t* mask, long mpitch, void* img, long ipitch);

/* CIECAM02 (ConvertVectorToJCHPyramidFunctor, src/fixmath.h:1406-1464):
 * interleaved 8- or 16-bit RGB `img` into the JCh pyramid level 0 `out`,
 * three planes `oplane` elements apart of Int16 (8-bit) or Int32 (16-bit)
 * fixed point with 7 fraction bits: (J s, C sin h s, C cos h s) with
 * s = 2^(IntegerBits - 8) = 2 for 8-bit, 512 for 16-bit images. */
int enbl_rgb_to_jch(enbl_context* ctx, int itype, int w, int h, const void* img, long ipitch, void* out, long opitch,
                    long oplane);

/* ConvertJCHPyramidToVectorFunctor (src/fixmath.h:1854-1977): a JCh level
 * 0 (planar, as enbl_rgb_to_jch writes it) into RGB `img` where mask != 0;
 * J <= 0 gives black; out-of-gamut colours are clamped (the reference
 * searches a close in-gamut colour with GSL minimisers). */
int enbl_jch_to_rgb_if(enbl_context* ctx, int itype, int w, int h, const void* lv, long lpitch, long lplane,
                       const uint8_t* mask, long mpitch, void* img, long ipitch);

/* ---- pipelines ---------------------------------------------------------- */

/* One enblend blend step on the ROI canvas, replacing src/enblend.h:378-510
 * of enblendMain: mask Gaussian pyramid (no alpha) of the 0/255 seam mask,
 * Laplacian pyramids of the white and black images under their own alpha,
 * black alpha |= white alpha, blend, collapse and the masked copy into the
 * black image.  black/black_alpha are updated in place.  Device buffers;
 * intermediate pyramids are cached in the context and reused by later calls
 * of the same geometry. */
int enbl_enblend_pair(enbl_context* ctx, int itype, int channels, int w, int h, int levels,
                      int wraparound,
                      const void* white, long wpitch, const uint8_t* white_alpha, long wapitch,
                      void* black, long bpitch, uint8_t* black_alpha, long bapitch,
                      const uint8_t* mask, long mpitch,
                      int dither_mode, uint64_t seed);

/* ---- driver around the blend: import, assembly, overlap, ROI, export ---- */

/* vigra::Rect2D: [x0, x1) x [y0, y1), empty when x1 <= x0 or y1 <= y0. */
typedef struct enbl_rect { int x0, y0, x1, y1; } enbl_rect;

/* One decoded input image as vigra::importImageAlpha reads it: interleaved
 * bands of the file's pixel type (ENBL_U8, ENBL_U16 or ENBL_F32), `channels`
 * image bands followed by the alpha band when bands == channels + 1.  Its
 * upper-left corner sits at (x, y) of the union canvas. */
typedef struct enbl_layer {
    const void* data;   /* device pointer */
    long pitch;         /* elements per row */
    int w, h, x, y;
    int ftype;
    int bands;
} enbl_layer;

/* Replaces assemble() with import() (src/assemble.h:231-444): the first
 * layer with taken[i] == 0 goes onto a zeroed canvas of union_w x union_h
 * (all of its rectangle); unless one_at_a_time, every later untaken layer
 * whose thresholded alpha meets none of the canvas alpha is copied where its
 * alpha is set, in list order.  The alpha band becomes 255 where
 * T(max/2) <= a <= T(max) (T = the image component type, max = the file
 * type's range), else 0; no alpha band means 255.  Samples of a narrower
 * integer file are mapped to the image's range as vigra::linearRangeMapping
 * does (8-bit into 16-bit or float, 16-bit into float).  taken[] is updated;
 * *bb receives the bounding box of the canvas alpha.  Returns the number of
 * layers placed (0: every layer was taken, the canvas is untouched) or a
 * negative status.  Synchronises the context's stream. */
int enbl_assemble(enbl_context* ctx, int itype, int channels, int union_w, int union_h, int nlayers,
                  const enbl_layer* layers, int one_at_a_time, int* taken,
                  void* image, long ipitch, uint8_t* alpha, long apitch, enbl_rect* bb);

/* Replaces inspectOverlap (src/bounds.h:41-80) over w x h alpha windows:
 * 0 NoOverlap, 1 PartialOverlap, 2 CompleteOverlap in *overlap (enum Overlap,
 * src/common.h:193).  Synchronises. */
int enbl_inspect_overlap(enbl_context* ctx, int w, int h, const uint8_t* alpha1, long p1,
                         const uint8_t* alpha2, long p2, int* overlap);

/* The transition-line box of maskBounds (src/mask.h:520-560), relative to the
 * mask; *found = 0 when the mask holds one value only.  Synchronises. */
int enbl_mask_bounds(enbl_context* ctx, int w, int h, const uint8_t* mask, long mpitch, enbl_rect* bb,
                     int* found);

/* One iteration of enblendMain's blending loop (src/enblend.h:127-545) for
 * the fine nearest-feature-transform seam mask without optimisation
 * (--fine-mask --no-optimize): union and intersection boxes, inspectOverlap
 * (a redundant white image leaves black as it is: *outcome 1; images that do
 * not overlap, with exact_levels == 0, are combined by copying: *outcome 2),
 * the NFT mask over the union box (norm 1 or 2), maskBounds, roiBounds with
 * minimum_levels and exact_levels (--levels), then the blend step of
 * enbl_enblend_pair on the ROI (*outcome 0).  black/black_alpha and *black_bb
 * are updated as the reference updates blackPair and blackBB.  Canvases are
 * union_w x union_h; *black_bb and *white_bb are the bounding boxes of the
 * two alphas, as enbl_assemble returns them.  With wraparound and a union box
 * as wide as the canvas the seam mask wraps horizontally
 * (enbl_nft_mask_periodic, boundary 1).  Synchronises. */
int enbl_enblend_step(enbl_context* ctx, int itype, int channels, int union_w, int union_h, int wraparound,
                      int exact_levels, int minimum_levels, int norm,
                      void* black, long bpitch, uint8_t* black_alpha, long bapitch, enbl_rect* black_bb,
                      const void* white, long wpitch, const uint8_t* white_alpha, long wapitch,
                      const enbl_rect* white_bb, int dither_mode, uint64_t seed, int* outcome);

/* enblendMain's whole loop (src/enblend.h:56-594) over n decoded layers on a
 * union_w x union_h canvas, with the fine NFT seam mask and no optimiser:
 * the first enbl_assemble becomes black, every later one white, each white
 * is folded into black by enbl_enblend_step with dither seed seed + m for
 * the m-th step.  pre_assemble is --pre-assemble (assemble non-overlapping
 * layers; off: one layer at a time).  The white canvas lives in the
 * context's scratch.  counts (optional) receives the number of steps that
 * blended, that found white redundant and that copied without blending. */
int enbl_enblend_layers(enbl_context* ctx, int itype, int channels, int union_w, int union_h, int nlayers,
                        const enbl_layer* layers, int pre_assemble, int wraparound, int exact_levels,
                        int minimum_levels, int norm, int dither_mode, uint64_t seed,
                        void* black, long bpitch, uint8_t* black_alpha, long bapitch, enbl_rect* black_bb,
                        int* counts);

/* Replaces checkpoint()'s conversion before the export call (src/assemble.h:
 * 130-224): for an output file type otype (ENBL_U8, ENBL_U16, ENBL_F32), the
 * image as the reference hands it to vigra::exportImageAlpha -- unchanged when
 * the ranges agree, narrowed by linearRangeMapping in the image's own type
 * when the image range covers the output range, or a float image rescaled
 * into UInt32 -- interleaved with the alpha band (alpha != 0 -> the
 * component type's alpha maximum).  out holds channels + 1 bands of the type
 * written to *out_ctype (the image type, or ENBL_U32).  An 8-bit image written
 * as 16-bit is ENBL_E_UNSUPPORTED (the reference's conversion overflows). */
int enbl_checkpoint(enbl_context* ctx, int itype, int channels, int w, int h, const void* image, long ipitch,
                    const uint8_t* alpha, long apitch, int otype, void* out, long opitch, int* out_ctype);

/* ---- TIFF layer files (host side of import() and checkpoint()) -------- */

/* What vigra::ImageImportInfo reports for one directory (layer) of a TIFF
 * file as enblend's import() reads it (src/assemble.h:231-289): size, bands
 * (extra samples, i.e. alpha, included and last), the component type, the
 * canvas position (XPosition / YPosition times the resolution, rounded) and
 * the canvas size of the PIXAR_IMAGEFULLWIDTH/LENGTH tags (0 when absent). */
typedef struct enbl_tiff_layer_info {
    int w, h;
    int bands;          /* samples per pixel */
    int extra;          /* extra samples (alpha) among them */
    int ftype;          /* ENBL_U8, ENBL_U16, ENBL_U32 or ENBL_F32 */
    int x, y;           /* position on the canvas */
    int canvas_w, canvas_h;
    int compression;    /* TIFF Compression tag */
    int pages;          /* directories in the file */
} enbl_tiff_layer_info;

/* One layer to write: interleaved host samples, alpha bands last. */
typedef struct enbl_tiff_page {
    int w, h;
    int ftype;          /* ENBL_U8, ENBL_U16, ENBL_U32 or ENBL_F32 */
    int bands, extra;
    const void* data;   /* host pointer */
    long pitch;         /* elements per row */
    int x, y;           /* canvas position (>= 0) */
    int canvas_w, canvas_h;  /* 0: no canvas tags */
} enbl_tiff_page;

/* TIFF compression codes accepted by enbl_tiff_write */
#define ENBL_TIFF_NONE 1
#define ENBL_TIFF_LZW 5
#define ENBL_TIFF_DEFLATE 8
#define ENBL_TIFF_PACKBITS 32773

/* Header of directory `page` (vigra::ImageImportInfo, src/assemble.h:231-289).
 * Reads uncompressed, PackBits, LZW and Deflate strips or tiles, chunky or
 * planar, either byte order, predictors 1-3. */
int enbl_tiff_info(const char* path, int page, enbl_tiff_layer_info* info);

/* The samples of directory `page`, interleaved as vigra::importImageAlpha
 * hands them over (colour bands, then alpha), into the host buffer `buf`
 * (pitch >= w * bands elements).  enbl_assemble takes them from there once
 * they are on the device. */
int enbl_tiff_read(const char* path, int page, void* buf, long pitch);

/* Writes `npages` layers as one little-endian TIFF (one directory each):
 * what vigra::exportImageAlpha does with checkpoint()'s buffer
 * (src/assemble.h:130-224).  compression is one of ENBL_TIFF_*; LZW and
 * Deflate use horizontal differencing on integer samples.  Alpha bands are
 * written as unassociated alpha. */
int enbl_tiff_write(const char* path, int npages, const enbl_tiff_page* pages, int compression);

#ifdef __cplusplus
}
#endif

#endif /* ENBLEND_AMD_H */
