/* C ABI of the MI355X perturbation path (libmxhip.so).
 *
 * The reference engine has no C ABI for this path (its only C-ABI header,
 * src/engine/dll_interface.h:7-11, is declared but never defined); its boundary is
 * the C++ backend plugin IComputeBackend (src/engine/compute_backend.h:102-115) over
 * the tiled frame API Mandel::Compute (src/engine/mandel_perturbation.h:72-75).
 * These entry points are what that plugin binds.  Plain pointers and sizes only.
 *
 * Ownership: host buffers are the caller's; device buffers belong to the library.
 * One context per host thread; a context runs one call at a time.
 * Every int-returning call returns MX_OK or a negative MX_E_* code, and the message
 * for the last failure on a context is available from mx_last_error().
 */
#ifndef MX_H
#define MX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    MX_OK = 0,
    MX_E_INVALID = -1,     /* request fails validation (compute_backend.cpp:263-272) */
    MX_E_NODEVICE = -2,    /* no HIP device */
    MX_E_HIP = -3,         /* a HIP runtime call failed */
    MX_E_CANCELLED = -4,   /* mx_cancel() was called (compute_backend.cpp:116) */
    MX_E_BUSY = -5,        /* the context is already computing (compute_backend.cpp:48) */
    MX_E_NOMEM = -6,
    MX_E_UNSUPPORTED = -7  /* a request outside this path (coloring, Julia, ...) */
};

/* Kernel forms. */
enum {
    MX_FORM_FE64 = 0,  /* plain: IEEE FP64 + FP64-mantissa floatexp, engine op order */
    MX_FORM_FE32 = 1,  /* mixed: FP32-mantissa floatexp delta loop, FP64 decisions */
    MX_FORM_DOUBLE = 2, /* plain double perturbation (scale 1e6..1e280) */
    MX_FORM_SC64 = 3    /* the mixed form's power-of-two-scaled algorithm with an FP64 delta */
};

/* ColoringMethod bits accepted here (mandel_perturbation.h:30-49). */
enum { MX_SUPER_SAMPLING = 1, MX_EXTERIOR_DIST_EST = 2 };

#define MX_EMPTYPIXEL (-10.0f) /* mandel_perturbation.h:23 */

typedef struct mx_ctx mx_ctx;
typedef struct mx_ref mx_ref;
typedef struct mx_plan mx_plan;
typedef struct mx_batch mx_batch;
typedef struct mx_graph mx_graph;
typedef struct mx_frame mx_frame;

/* ---- context ------------------------------------------------------------ */

/* Library ABI version. */
int mx_version(void);
/* Number of visible HIP devices (0 when there is none). */
int mx_device_count(void);
/* Open device `device`; replaces createComputeBackend("hip") device set-up
 * (the D3D11 precedent: compute_backend_d3d11.cpp:213-279). */
int mx_ctx_create(int device, mx_ctx** out);
void mx_ctx_destroy(mx_ctx* ctx);
const char* mx_last_error(const mx_ctx* ctx);
/* Compute units of the context's device, for sizing mx_tile.grid_blocks (the D3D11
 * backend reads its adapter's properties the same way, compute_backend_d3d11.cpp:213-279). */
int mx_ctx_cu_count(const mx_ctx* ctx);
/* IComputeBackend::cancel / resetCancellation (compute_backend.cpp:119-128).
 * cancel() may be called from any thread; running kernels stop taking samples. */
void mx_cancel(mx_ctx* ctx);
void mx_reset_cancellation(mx_ctx* ctx);

/* ---- reference upload (once per reference orbit) -------------------------- */

/* Host arrays describing one reference: the orbit shadows filled by createRef /
 * calCoefficient (mandel_perturbation.cpp:4113-4249) and the BLA table of buildBLA
 * (2613-2709).  zf[2k], zf[2k+1] = X_{k+1}.  BLA level p has lev_size[p] entries
 * starting at lev_off[p], each skipping 2^p iterations. */
typedef struct {
    int32_t len;             /* entries = mx_ref_it + 1 (a periodic reference: the period) */
    const double* zf;        /* [2*len] */
    const double* zfe_m;     /* [2*len] FloatExp mantissas (re, im) */
    const int32_t* zfe_e;    /* [2*len] FloatExp exponents (re, im) */
    int32_t nlev;            /* 0: no BLA table */
    const int32_t* lev_off;  /* [nlev] */
    const int32_t* lev_size; /* [nlev] */
    const double* bla_r2;    /* [total] validity radius^2 */
    const double* bla_coef;  /* [4*total] ar, ai, br, bi */
    double rmax2;            /* largest radius^2 */
    /* plain-double path only (NULL / 0 otherwise): the glitch metric's orbit
     * magnitudes _z_m3 (mandel_perturbation.cpp:4249) and the series approximation
     * of calCoefficient (4273-4303): dz0 = sum A_x (dc/delta)^(x+1), start at sa_it+1 */
    const double* z_m3;      /* [len] |X|^2 / 1e6 */
    int32_t sa_it;
    int32_t sa_order;
    const double* sa_coef;   /* [60] A_0..A_29 as (re, im) */
    double sa_delta_re, sa_delta_im;
    /* plain-double path with the exterior distance estimate (NULL otherwise): the
     * reference derivative D (re, im interleaved, [2*len]), the BLA derivative couplings
     * (cr, ci, er, ei per entry, [4*total]; mandel_perturbation.h:203-206) and the
     * derivative-delta series B_0..B_29 ([60]) */
    const double* df;
    const double* bla_der;
    const double* sa_coef_b;
    /* > 0: a periodic-nucleus reference (createPeriodicRef, mandel_perturbation.cpp:
     * 1108-1172): the floatexp kernels read X_m at (m-1) mod period, never rebase at
     * its end, and land BLA skips inside one period; len must equal it */
    int32_t period;
} mx_ref_desc;

int mx_ref_upload(mx_ctx* ctx, const mx_ref_desc* desc, mx_ref** out);
void mx_ref_release(mx_ref* ref);

/* ---- one tile of samples against one reference --------------------------- */

/* The per-pixel inputs of stepParallel's floatexp path (mandel_perturbation.cpp:
 * 2150-2168): dc is separable, Dcr = dcol[column], Dci = drow[row]. */
typedef struct {
    int32_t width, height;   /* sample grid; outputs are width*height */
    int32_t stride;          /* 1: all samples; k: only rows/cols multiple of k (probe) */
    int32_t max_iter;
    double esc2;             /* escape radius squared (Mandel::escapeRadiusSquared) */
    int32_t form;            /* MX_FORM_* */
    int32_t ede;             /* exterior distance estimate output */
    int32_t interior;        /* periodicity interior detection (_use_interior) */
    int32_t int_rep;         /* exact state-repeat interior detection */
    int32_t bla;             /* apply BLA skips */
    int32_t bla_minlevel;    /* only skips of 2^minlevel or more */
    double interior_eps2;
    int32_t interior_confirm;
    const double* dcol_m;    /* [width] */
    const int32_t* dcol_e;
    const double* drow_m;    /* [height] */
    const int32_t* drow_e;
    double dx_m;             /* pixel spacing (EDE normalisation) */
    int32_t dx_e;
    int32_t skip_sample;     /* sample index left untouched (the reference pixel), -1: none */
    /* plain-double path (form MX_FORM_DOUBLE): dc = (dcol_d[column], drow_d[row]),
     * truncated from dx*(column - ref column) as stepParallel does
     * (mandel_perturbation.cpp:2300-2312); the FloatExp tables above may be NULL */
    const double* dcol_d;
    const double* drow_d;
    /* optional list of sample indices (row-major order) to compute instead of the
     * grid: a later reference pass over the samples the previous pass left unresolved */
    const int32_t* list;
    int32_t list_len;
    /* launch tuning (0: defaults); results do not depend on either.  chunk: lean
     * steps per trip between the full step's checks (default 32); grid_blocks:
     * workgroups of 256 lanes in the persistent grid (default 8 per CU, at most one
     * per 64 samples). */
    int32_t chunk;
    int32_t grid_blocks;
    /* floatexp forms: explicit samples instead of the grid (adaptive supersampling's
     * sub-samples, mandel_perturbation.cpp:1886-1981): n_samples entries whose dc is
     * (sample_dc_m[2q] 2^sample_dc_e[2q], sample_dc_m[2q+1] 2^sample_dc_e[2q+1]); outputs
     * are indexed by entry, width/height/stride and the dc tables are not used */
    int32_t n_samples;
    const double* sample_dc_m;
    const int32_t* sample_dc_e;
    /* launch tuning (0: default 32), results do not depend on it: a wave ends a run of
     * lean steps as soon as fewer than quorum/64 of the lanes that started it are still
     * stepping (the others would idle until the run ends) */
    int32_t quorum;
    /* plain-double form only, results do not depend on it: 0 stages the orbit (all of
     * it, or the prefix that fits) in LDS unless the tile has both EDE and interior
     * detection, 1 always, -1 never */
    int32_t lds_stage;
} mx_tile;

/* A plan holds the device copy of a tile's inputs and outputs so that launches
 * can be repeated (and timed) with everything resident in HBM. */
int mx_plan_create(mx_ctx* ctx, const mx_ref* ref, const mx_tile* tile, mx_plan** out);
int mx_plan_launch(mx_plan* plan);                      /* asynchronous */
int mx_plan_wait(mx_plan* plan, double* kernel_ms);     /* synchronise; kernel time of the last launch */
/* Copy results to host buffers (any may be NULL).  n_esc: escape iteration, -1 =
 * reached max_iter, -2 = interior detected.  value: the double the engine casts to
 * float.  out: the engine's float buffer value (-2 interior).  stats4: executed
 * steps, BLA applies, BLA-skipped iterations, rebases. */
int mx_plan_read(mx_plan* plan, int32_t* n_esc, double* value, float* out, uint64_t* stats4);
void mx_plan_destroy(mx_plan* plan);
/* Plain-double path: the samples the last launch left unresolved (glitch, or the end of
 * the reference reached; mandel_perturbation.cpp:2556-2577) with their glitch metric
 * glitch_p = |z|^2 / _z_m3[iteration], in no particular order.  *count is the number of
 * such samples; at most cap are copied. */
int mx_plan_unresolved(mx_plan* plan, int32_t* idx, double* glitch, int32_t cap, int32_t* count);
/* ---- batches: many tiles in one persistent launch ------------------------- */

/* The engine renders a large or multi-view job as a loop of Compute calls, one per
 * strip or tile (src/gui/gui_export.cpp:119-140 issues one ComputeRequest per strip;
 * compute_backend.cpp:263-290 serves one request per call).  A batch launches the
 * tiles of several plans, each with its own reference, size and max_iter, as one
 * persistent grid: waves walk the plans in the given order (largest work first keeps
 * the tail short) and move on as each plan's queues drain.  The plans must be of one
 * floatexp form (FE64, FE32 or SC64) with the same ede flag, each plan
 * at most once; they stay
 * owned by the caller, must outlive the batch, and their results are read with
 * mx_plan_read.  A batch launch equals launching its plans one by one, sample for
 * sample. */
int mx_batch_create(mx_ctx* ctx, mx_plan* const* plans, int32_t n, mx_batch** out);
int mx_batch_launch(mx_batch* batch);                   /* asynchronous */
int mx_batch_wait(mx_batch* batch, double* kernel_ms);  /* synchronise; time of the last launch */
void mx_batch_destroy(mx_batch* batch);

/* ---- captured sequences: a zoom animation's frames as one HIP graph --------- */

/* The engine renders an animation as one Compute call per frame (the export loop of
 * src/gui/gui_export.cpp:119-140 per image).  A graph records, once, every plan's
 * reset and launch in the given order on the context's stream; each mx_graph_launch
 * replays the whole sequence with no per-frame host launch work.  Plans: floatexp
 * forms, owned by the caller, outliving the graph; results are read per plan with
 * mx_plan_read after mx_graph_wait.  A replay equals launching the plans one by one. */
int mx_graph_create(mx_ctx* ctx, mx_plan* const* plans, int32_t n, mx_graph** out);
int mx_graph_launch(mx_graph* graph);                    /* asynchronous */
int mx_graph_wait(mx_graph* graph, double* ms);          /* synchronise; time of the last replay */
void mx_graph_destroy(mx_graph* graph);

/* create + launch + wait + read + destroy */
int mx_tile_render(mx_ctx* ctx, const mx_ref* ref, const mx_tile* tile, int32_t* n_esc, double* value, float* out, uint64_t* stats4);

/* mx_tile_render without allocating: no hipMalloc, hipFree or other allocating call.
 * No counterpart in the reference; for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid it).  The
 * tile's dc tables, its outputs and its per-launch state (work queues, counters, the
 * job record) are staged in a fixed scratch area in device memory, one per device,
 * reserved when the library is loaded and held until the process ends; the reference
 * must already be uploaded.  Limits: a grid tile (no sample list, no explicit samples)
 * of a floatexp form with at most MX_NOALLOC_MAX_SAMPLES samples and at most
 * MX_NOALLOC_MAX_TABLE columns plus rows; larger tiles are refused with MX_E_INVALID.
 * Every call on a device shares that device's scratch area (which keeps the last
 * call's contents): calls that overlap on one device race, and a context runs one
 * call at a time (MX_E_BUSY otherwise). */
#define MX_NOALLOC_MAX_SAMPLES (1 << 20)
#define MX_NOALLOC_MAX_TABLE 16384
int mx_tile_render_noalloc(mx_ctx* ctx, const mx_ref* ref, const mx_tile* tile, int32_t* n_esc, double* value, float* out,
                           uint64_t* stats4);

/* ---- whole frame: Mandel::Compute (mandel_perturbation.cpp:1388-1788) ----- */

typedef struct {
    const char* center_re;   /* decimal strings (ComputeRequest::centerRe/centerIm/scale) */
    const char* center_im;
    const char* scale;
    int32_t width, height;   /* ComputeRequest::width/height */
    int32_t full_height;     /* strip export: 0 = the whole image (mandel_perturbation.cpp:1399-1424) */
    int32_t row_base;
    int32_t sub;             /* output cells per pixel side, odd; the iteration (and
                                values, n_esc) buffers hold (height sub) x (width sub)
                                cells, pixel (i, j) at its centre (getIndex,
                                mandel_perturbation.cpp:4334-4340).  With
                                MX_SUPER_SAMPLING the other cells of pixels that differ
                                from a neighbour are computed (adaptive supersampling,
                                floatexp path); cells not computed are left as they are */
    int32_t max_iterations;
    int32_t coloring_method; /* 0 or MX_EXTERIOR_DIST_EST */
    float escape_radius;     /* 0 => 1e8f (MANDEL_BAILOUT) */
    int32_t precision_bits;  /* 0 => digits(scale)*log2(10)+64 (render.cpp:170) */
    int32_t form;            /* floatexp path (scale > 1e280): MX_FORM_FE64, FE32 or SC64;
                                the plain-double path (1e6 < scale <= 1e280) ignores it */
    int32_t use_bla;         /* -1 default (on) */
    int32_t bla_minskip;     /* 0 => 8 (MANDEL_BLA_MINSKIP) */
    int32_t interior;        /* -1 default (on, gated by the probe) */
    float* iterations;       /* [width*height] engine float buffer, caller-owned */
    double* values;          /* optional [width*height] */
    int32_t* n_esc;          /* optional [width*height] */
    int32_t options;         /* MX_OPT_* bits; 0 = the engine's behaviour */
} mx_request;

/* Request options outside the engine's behaviour, for frame sequences that share one
 * reference (a zoom animation around a fixed centre):
 *   MX_OPT_VIEW_CENTER_REF  reference at the exact view centre c0 + d*(n-1)/2 (the
 *                           engine's fallback reference, mandel_perturbation.cpp:
 *                           1326-1373) instead of the centre pixel, so that frames of
 *                           one centre share it;
 *   MX_OPT_FLOATEXP         the floatexp path at any scale above 1e6 (the engine takes
 *                           the plain-double path up to 1e280);
 *   MX_OPT_NO_PERIODIC      no search for a nucleus in view: always the full reference
 *                           from the centre pixel (the engine's MANDEL_PERIODIC=0,
 *                           mandel_perturbation.cpp:1240-1250);
 *   MX_OPT_PERIODIC         the search at every scale, not only above 1e280 (the
 *                           engine's MANDEL_PERIODIC=1, mandel_perturbation.cpp:
 *                           1240-1245); on the floatexp path only (with MX_OPT_FLOATEXP
 *                           below 1e280): the periodic reference on the plain-double
 *                           path (its mod-p delta loop, 2345-2349, 2315) is refused with
 *                           MX_E_UNSUPPORTED. */
enum { MX_OPT_VIEW_CENTER_REF = 1, MX_OPT_FLOATEXP = 2, MX_OPT_NO_PERIODIC = 4, MX_OPT_PERIODIC = 8 };

typedef struct {
    int32_t path;            /* 1 = plain-double path, 2 = floatexp rescaled path */
    int32_t ref_it;          /* reference length - 1 */
    int32_t ref_escaped;
    int32_t ref_virtual;     /* exact view-centre reference */
    double ref_x, ref_y;     /* reference position in pixels */
    int32_t interior_used;   /* interior detection after the probe gate */
    int32_t cutoff_sensitive;/* view-centre reference escaping within 16 of mxit, or bounded
                                but escaping by mxit + 64: the fallback below ran */
    int32_t precision_bits;
    int32_t bla_levels;
    double ref_seconds;      /* host GMP orbit */
    double bla_seconds;      /* host BLA build */
    double upload_seconds;
    double probe_ms;         /* probe kernel */
    double kernel_ms;        /* main pass kernel */
    double total_seconds;
    uint64_t steps, bla_applies, bla_skipped, rebases;
    uint64_t nominal_iterations; /* sum of n_esc, max_iterations for interior */
    int32_t passes;          /* reference passes (plain-double path re-references on glitches) */
    int32_t sa_it;           /* first reference's series-approximation skip */
    int32_t escalated;       /* scale <= 1e280 but run on the floatexp path because the
                                reference passes too close to zero (mandel_perturbation.cpp:1336-1355) */
    int32_t ss_pixels;       /* pixels refined by adaptive supersampling */
    double ss_ms;            /* their sub-sample pass (kernel) */
    int32_t period;          /* > 0: the reference was the period-`period` nucleus in view */
    int32_t subpixel;        /* a nucleus was found but spans < 4 pixels: BLA was turned off */
    int32_t gmp_fallback;    /* cutoff-sensitive frame: cells whose value was negative or above
                                mxit - 16, recomputed in arbitrary precision on the host
                                (mandel_perturbation.cpp:2173-2183) */
} mx_frame_info;

int mx_compute(mx_ctx* ctx, const mx_request* req, mx_frame_info* info);

/* Measurement aid, no counterpart in the reference: one plain device-to-device copy
 * kernel over `bytes` (two fresh device buffers), timed.  Its hardware byte counters
 * (rocprofv3 FETCH_SIZE / WRITE_SIZE) read bytes / 1024 KiB each for an exact copy:
 * the factor that turns the same counters of the path's kernels into bytes. */
int mx_copy_probe(mx_ctx* ctx, size_t bytes, double* kernel_ms);

/* ---- colouring and downsampling (tools/render.cpp:252-302) ------------------ */

/* The engine's palette (render.cpp:21-30: colorMapInit through mono_cubic_interpolate,
 * engine/interpolate.cpp:7-65): out[3*2048], channels r, g, b of 2048 floats in 0..255.
 * Host only. */
int mx_palette(float* out);

/* render.cpp's colour pass on the device: the (height*ss) x (width*ss) float buffer
 * `iterations` (the engine's layout: smooth value or distance estimate, negative for
 * interior / empty) mapped through the palette (getColor, render.cpp:69-85: density
 * 60 is the engine's default, MANDEL_DENS), averaged over each ss x ss block in linear
 * light (exponent 2.2) and written as width*height RGB bytes.  coloring_method selects
 * the distance-estimate mapping (MX_EXTERIOR_DIST_EST). */
int mx_colorize(mx_ctx* ctx, const float* iterations, int32_t width, int32_t height, int32_t ss, int32_t coloring_method,
                float density, uint8_t* rgb);

/* The same from a launched plan's device output (a full grid tile, stride 1, no sample
 * list, sides multiples of ss): the tile's W/ss x H/ss pixels, without copying its
 * float samples to the host.  No counterpart in the engine, which colours on the host. */
int mx_plan_colorize(mx_plan* plan, int32_t ss, int32_t coloring_method, float density, uint8_t* rgb);

/* ---- host-only frame preparation (no GPU needed) --------------------------- */

/* Everything Mandel::Compute prepares on the host before the delta loop: grid,
 * reference orbit (GMP), BLA table and dc tables.  View pointers stay valid until
 * mx_frame_free. */
typedef struct {
    int32_t len;
    const double* zf;
    const double* zfe_m;
    const int32_t* zfe_e;
    int32_t nlev;
    const int32_t* lev_off;
    const int32_t* lev_size;
    int64_t bla_total;
    const double* bla_r2;
    const double* bla_coef;
    double rmax2;
    double dcmax;
    const double* dcol_m;
    const int32_t* dcol_e;
    const double* drow_m;
    const int32_t* drow_e;
    double dx_m;
    int32_t dx_e;
    double dy_m;
    int32_t dy_e;
    double esc2;
    int32_t bla_minlevel;
    int32_t ref_it, ref_escaped, ref_virtual;
    double ref_x, ref_y;
    int32_t ref_pixel_n;     /* n_esc of the reference pixel (non-virtual references) */
    double ref_pixel_value;  /* its value (getEscapeTime / accuratePointCompute) */
    int32_t cutoff_sensitive;
    int32_t precision_bits;
    double ref_seconds, bla_seconds;
    /* plain-double path (path == 1) */
    int32_t path;
    const double* z_m3;      /* [len] */
    int32_t sa_it, sa_order;
    const double* sa_coef;   /* [60] */
    double sa_delta_re, sa_delta_im;
    const double* dcol_d;    /* [width] */
    const double* drow_d;    /* [height] */
    const double* dfr;       /* [len] reference derivative (EDE) */
    const double* dfi;
    const double* bla_der;   /* [4*bla_total] (EDE) */
    const double* sa_coef_b; /* [60] (EDE) */
    int32_t escalated;       /* see mx_frame_info */
    int32_t period;          /* > 0: periodic-nucleus reference, len == period */
    int32_t subpixel;        /* a nucleus in view smaller than 4 pixels: full reference, BLA off */
} mx_frame_view;

int mx_frame_prepare(const mx_request* req, mx_frame** out, mx_frame_view* view);
void mx_frame_free(mx_frame* frame);

#ifdef __cplusplus
}
#endif

#endif
