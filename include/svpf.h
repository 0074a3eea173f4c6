/* libsvpf — MI355X (gfx950) implementation of the open-svpflow motion path:
 * svpflow1.Super pyramid build, svpflow1.Analyse block search and
 * svpflow2.SmoothFps motion-compensated render.
 *
 * Plain C ABI: no exceptions, int status codes (0 = ok, negative = error, text
 * in svpf_last_error()), caller-owned buffers with explicit byte strides, and
 * opaque context handles.  Every entry point is safe to call from several host
 * threads as long as each thread uses its own context (a context owns one HIP
 * stream); calls on one context are serialised by the caller.
 *
 * Pointers named d_* are device pointers (allocated with svpf_device_alloc or any
 * HIP allocator on the context's device); h_* are host pointers.
 *
 * Each entry point cites the reference function it replaces (paths relative to
 * the open-svpflow repository).
 */
#ifndef SVPF_H
#define SVPF_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SVPF_OK 0
#define SVPF_ERR_ARG (-1)    /* invalid argument / unsupported shape */
#define SVPF_ERR_HIP (-2)    /* HIP runtime error (message names the call) */
#define SVPF_ERR_OPT (-3)    /* option string rejected (message = reference error text) */
#define SVPF_ERR_NOMEM (-4)  /* allocation failed */
#define SVPF_ERR_STATE (-5)  /* call out of order / resource exhausted */

typedef struct svpf_ctx svpf_ctx;

/* One YUV420P8 frame: three planes with byte strides. */
typedef struct svpf_frame {
    uint8_t* y;
    uint8_t* u;
    uint8_t* v;
    int32_t stride_y;
    int32_t stride_u;
    int32_t stride_v;
} svpf_frame;

/* ------------------------------------------------------------------ runtime */
const char* svpf_last_error(void);
int svpf_version(void); /* svpGetVersion value, crates/svpflow1/src/abi.rs:3-8 */
int svpf_create(int device, svpf_ctx** out);
void svpf_destroy(svpf_ctx* ctx);
int svpf_synchronize(svpf_ctx* ctx);
void* svpf_stream(svpf_ctx* ctx); /* the context's hipStream_t */
int svpf_device_alloc(svpf_ctx* ctx, size_t bytes, void** d_out);
int svpf_device_free(svpf_ctx* ctx, void* d_ptr);
int svpf_copy_h2d(svpf_ctx* ctx, void* d_dst, const void* h_src, size_t bytes);
int svpf_copy_d2h(svpf_ctx* ctx, void* h_dst, const void* d_src, size_t bytes);
/* Device-to-device copy on the context's stream (asynchronous). */
int svpf_copy_d2d(svpf_ctx* ctx, void* d_dst, const void* d_src, size_t bytes);
int svpf_memset_d(svpf_ctx* ctx, void* d_dst, int value, size_t bytes);
/* Event timing on the context's stream: start, then stop returns elapsed ms. */
int svpf_timer_start(svpf_ctx* ctx);
int svpf_timer_stop(svpf_ctx* ctx, float* ms);
/* Per-kernel timing for benchmarks (no counterpart in the reference): when enabled,
 * HIP events are recorded on the context's stream around each launch of "super"
 * (level 0 + pyramid), "search" (one Analyse level's block search), "analyse_aux"
 * (reserved), "smooth_prep" (vector prep, plans and masks) and "render".  Enabling
 * (or disabling) synchronizes and clears the records; read synchronizes and returns
 * the summed milliseconds and the number of timed launches of one kernel. */
int svpf_profile_enable(svpf_ctx* ctx, int on);
int svpf_profile_read(svpf_ctx* ctx, const char* kernel, double* total_ms, int32_t* launches);
/* Analyse sweep statistics of one pyramid level since profiling was enabled (no
 * counterpart in the reference; gathered on the device while enabled, without a host
 * round trip; this call waits for the context's stream): out4 = {level runs,
 * sweeps launched, block evaluations skipped because nothing the block reads had
 * changed, block evaluations run (sweep 0 included)}. */
int svpf_analyse_stats(svpf_ctx* ctx, int32_t level, int64_t* out4);

/* ------------------------------------------------------------------ Super */
/* Replaces SuperOpts (crates/svpflow1/src/super_opts.rs:3-95). */
typedef struct svpf_super_params {
    int32_t width, height; /* source luma size */
    int32_t pel;           /* 1, 2 or 4 */
    int32_t gpu;           /* 0..2 */
    int32_t scale_down;    /* 4 = build the pyramid */
    int32_t scale_up;      /* 0 bilinear, 2 bicubic (forced 0 when gpu != 0) */
    int32_t full;          /* bool */
    int32_t rc;            /* bool */
    int32_t levels;        /* auto_levels(width, height) */
} svpf_super_params;

/* SuperOpts::from_opt (super_opts.rs:17-49) over the reference's option-string
 * syntax (crates/svpflow-core/src/params.rs:1-11).  On SVPF_ERR_OPT `err` holds
 * the reference's message.  *sdata = SuperOpts::pack_data (super_opts.rs:51-63). */
int svpf_super_params_from_opt(const char* opt, size_t opt_len, int32_t width, int32_t height,
                               svpf_super_params* out, int64_t* sdata, char* err, size_t err_len);
/* super_plane_height (super_opts.rs:89-95): rows of the luma super plane; chroma has half. */
int32_t svpf_super_height(const svpf_super_params* p);

/* build_plane x3 for one frame (super_build.rs:8-93 as called by get_frame_super,
 * super_filter.rs:143-268).  d_src/d_dst hold device plane pointers.  Every byte of
 * the super frame (stride x rows) is written, so the destination needs no clearing. */
int svpf_super_build(svpf_ctx* ctx, const svpf_super_params* p, const svpf_frame* d_src,
                     const svpf_frame* d_dst);
/* The same for n frames of one clip in one launch per pyramid stage (batched form). */
int svpf_super_build_batch(svpf_ctx* ctx, const svpf_super_params* p, int32_t n, const svpf_frame* d_src,
                           const svpf_frame* d_dst);

/* ------------------------------------------------------------------ Analyse */
/* Replaces AnalyseOpts (crates/svpflow1/src/analyse_opts.rs:4-240). */
typedef struct svpf_analyse_params {
    int32_t vectors;            /* 1 bwd, 2 fwd, 3 both */
    int32_t block_w, block_h;
    int32_t overlap_mode, overlap_x, overlap_y;
    int32_t delta;
    int32_t pel, selector, levels, super_levels;
    int32_t width, height;
    int32_t lambda, lsad;
    double plevel;
    int32_t pnew, pglobal, pzero, pnbour, prev;
    int32_t search_type, coarse_search_type, distance, coarse_distance;
    int32_t coarse_trymany, coarse_bad_sad, coarse_bad_range;
    int32_t fine_satd, coarse_satd;
    int32_t has_refine, refine_thsad, refine_search_type, refine_distance, refine_satd;
    int32_t sort;
} svpf_analyse_params;

/* AnalyseOpts::from_opt (analyse_opts.rs:45-219, sdata unpacked as :243-257), with
 * width/height overridden by the source clip when > 0 (analyse_filter.rs:92-97).
 * vdata16 = pack_vdata_header (analyse_search.rs:2813-2841); *payload_len =
 * the vector-clip width (analyse_filter.rs:122-129). */
int svpf_analyse_params_from_opt(const char* opt, size_t opt_len, int64_t sdata, int32_t width, int32_t height,
                                 svpf_analyse_params* out, int32_t* vdata16, int32_t* payload_len, char* err,
                                 size_t err_len);

/* analyse_pair + pack_vector_frame for one frame pair (analyse_search.rs:68-92,
 * 2740-2811, as called by get_frame_analyse, analyse_filter.rs:229-334):
 * d_cur = super frame n, d_ref = super frame n+delta (device), d_payload = device
 * buffer of payload_len bytes (4-byte aligned) receiving the vector frame. */
int svpf_analyse_pair(svpf_ctx* ctx, const svpf_analyse_params* p, const svpf_frame* d_cur, const svpf_frame* d_ref,
                      uint8_t* d_payload);

/* Batched form: n pairs of one clip searched together (one wavefront per pair and
 * direction, one launch per pyramid level and stage).  Arrays hold n entries. */
int svpf_analyse_batch(svpf_ctx* ctx, const svpf_analyse_params* p, int32_t n, const svpf_frame* d_cur,
                       const svpf_frame* d_ref, uint8_t* const* d_payload);

/* Multi-clip batch: pair i uses params[i] (clips of different sizes and pyramid
 * depths share every launch; levels are aligned at level 0). */
int svpf_analyse_multi(svpf_ctx* ctx, int32_t n, const svpf_analyse_params* const* params, const svpf_frame* d_cur,
                       const svpf_frame* d_ref, uint8_t* const* d_payload);

/* Runtime knob, no reference counterpart: the number of parallel sweeps a pyramid
 * level may take before the context falls back to the sequential walk of the
 * reference's visit order (one wavefront per frame pair and direction).  Both give
 * the same vectors; 0 restores the default of 4096, the maximum. */
int svpf_analyse_set_max_sweeps(svpf_ctx* ctx, int32_t max_sweeps);

/* Runtime tuning of the Analyse level search (no reference counterpart; every setting
 * gives the same vectors).  Zero fields keep the defaults.
 *   chase        visit-order successors a later sweep evaluates right after a block
 *                changes (default 0: none; measured slower on large batches, whose
 *                early sweeps then wait on their longest chase);
 *   slot_scale   factor on the finest level's sweep-0 slot count (default 1.0);
 *   coarse_slots sweep-0 slots of the coarser levels (default 256: one segment per chain
 *                from 256 chains on; their visit order is sorted by parent score, so a
 *                block's earlier neighbours lie anywhere before it);
 *   sweep_grid   workgroups of a later sweep (default 2048);
 *   fine_run     consecutive visit positions a finest-level sweep-0 item searches in
 *                order (1, 2, 4, 8, 16 or 32; default 32);
 *   coarse_run   the same at the coarser levels (4, 8, 16 or 32; default 16);
 *   graphs       1 (default): the later sweeps of a level are replayed from a captured
 *                hipGraph of six sweeps (cached per launch configuration); -1: launched
 *                one by one.  The chase and svpf_analyse_pair_noalloc never use graphs;
 *   fast_list_min at a level whose sweep 0 ran the straight-line tile-plan kernel, a later
 *                sweep whose list holds at least this many blocks runs that kernel too,
 *                with a deferred pass for the blocks it passes on (default 16384; -1:
 *                never). */
typedef struct svpf_analyse_tuning {
    int32_t chase;
    double slot_scale;
    int32_t coarse_slots;
    int32_t sweep_grid;
    int32_t fine_run;
    int32_t coarse_run;
    int32_t graphs;
    int32_t fast_list_min;
} svpf_analyse_tuning;
int svpf_analyse_set_tuning(svpf_ctx* ctx, const svpf_analyse_tuning* t);

/* No counterpart in the reference.  svpf_analyse_pair for callers that cannot
 * allocate device memory (serving loops, harnesses that forbid hipMalloc): the same
 * arguments and payload, with the search's workspace, sort storage and descriptors
 * carved from a fixed 128 MiB device array per GPU instead of the context's
 * grow-only buffers.  A pair that needs more (beyond about one 2560x1440 pair with
 * 8x8 blocks and overlap 4) fails with SVPF_ERR_ARG and says how much it needs.
 * Every call on one GPU shares that array: calls that overlap on a GPU (two threads
 * or two contexts) race and must be serialised by the caller.  Like the other entry
 * points it synchronises the context's stream while it runs. */
int svpf_analyse_pair_noalloc(svpf_ctx* ctx, const svpf_analyse_params* p, const svpf_frame* d_cur,
                              const svpf_frame* d_ref, uint8_t* d_payload);

/* ------------------------------------------------------------------ SmoothFps */
typedef struct svpf_smooth svpf_smooth;

/* Creates a render session for one clip: the SmoothFps option string, the source
 * frame rate and length, the Super clip's `sdata` and the vector clip's 16-word
 * `vdata` header.  Replaces svpflow2 create_filter + Options::from_value
 * (crates/svpflow2/src/filter.rs:60-130, options.rs:150-260): the same option
 * keys and defaults, and on rejection SVPF_ERR_OPT with the reference's error text
 * in `err` ("SVSmoothFps: incorrect 'algo' value", ...).  *out_frames receives
 * the output clip length (frame_math.rs:128-135).  Renders YUV420P8 with the
 * reference's integer CPU renderer semantics (renderer.rs).  With ctx and out both
 * NULL it only validates the options and sets *out_frames (no device work). */
int svpf_smooth_create(svpf_ctx* ctx, const char* opt, size_t opt_len, int32_t width, int32_t height, int64_t fps_num,
                       int64_t fps_den, int32_t nframes, int64_t sdata, const int32_t* vdata16, svpf_smooth** out,
                       int32_t* out_frames, char* err, size_t err_len);
void svpf_smooth_destroy(svpf_smooth* s);  /* before svpf_destroy of its context */
/* Binds source frames first..first+count-1: the source planes, their Super frames
 * (d_super may be NULL when the vectors are pel 1) and the device vector payloads
 * written by svpf_analyse_* (the frame's getFrame inputs, core.rs:962-1051).  All
 * source frames share strides, as do all Super frames.  Binding marks the frame's
 * decoded vectors stale. */
int svpf_smooth_set_sources(svpf_smooth* s, int32_t first, int32_t count, const svpf_frame* d_src,
                            const svpf_frame* d_super, uint8_t* const* d_vectors);
/* Renders output frames k0..k0+count-1 into d_out[0..count-1] in one batch of
 * launches on the context's stream (get_frame -> try_render -> render_into_output,
 * core.rs:752-1770), each of svpf_smooth_output_size (with light.* padding, the light
 * border of output frame k is drawn around it: light.rs, core.rs:1866-1897).  If
 * h_decision is not NULL it receives, per frame, the pair
 * (kind, algo): kind 0 = copy of source n, 1 = copy of n+1, 2 = interpolated with
 * `algo`; the call then synchronizes. */
int svpf_smooth_render(svpf_smooth* s, int32_t k0, int32_t count, const svpf_frame* d_out, int32_t* h_decision);

/* Routes the session's rendered frames through the reference's float GPU renderer
 * (the OpenCL kernel of crates/svpflow2/src/gpu.rs:1039-1168, which the reference uses
 * on a GPU, core.rs:1429, 1630-1653) instead of the bit-exact integer CPU renderer
 * (renderer.rs), when `on` != 0: B-spline cubic vector sampling, bilinear taps, luma
 * blended in linear light.  Its results depend on floating-point rounding, so they
 * agree with the integer renderer only within a tolerance.  Sessions with block
 * vectors keep the integer renderer, as the reference does; so do sessions with
 * scene blending (the reference falls back only for their zero-origin frames). */
int svpf_smooth_set_float_render(svpf_smooth* s, int32_t on);
int32_t svpf_smooth_output_frames(const svpf_smooth* s);
/* output_info (crates/svpflow2/src/core.rs:603-613): the output frame size, the source
 * size plus twice the light.* padding (Options::padding, options.rs:256-277).  Output
 * frames passed to svpf_smooth_render have this size; the rendered frame sits at
 * (pad_x, pad_y) (chroma at half) and the light border (light.rs) fills the rest. */
int svpf_smooth_output_size(const svpf_smooth* s, int32_t* width, int32_t* height);
/* Output frame rate of the session (FilterState::output_info, core.rs:603-613). */
int svpf_smooth_output_rate(const svpf_smooth* s, int64_t* fps_num, int64_t* fps_den);
/* What output frame k reads, for a host that fetches frames on demand (the
 * getFrame requests of FilterState::request_frame, core.rs:150-201): its source
 * frame (source_frame, core.rs:2054-2077) and the range first..last of source
 * frames whose planes, Super frames and vector payloads the render of k can read;
 * raw_phase (Timing::raw_phase_256) and ratio (frame_num / frame_den) feed the
 * output's duration and PTS (copy_interpolated_timing, frame.rs:111-138). */
typedef struct svpf_smooth_frame {
    int32_t source, first, last;
    double raw_phase, ratio;
} svpf_smooth_frame;
int svpf_smooth_frame_info(const svpf_smooth* s, int32_t k, svpf_smooth_frame* out);
/* Unbinds source frames first..first+count-1 and frees the device state the session
 * derived from them (decoded vectors, motion planes, masks), for callers that stream
 * a long clip through a window of frames (no counterpart in the reference, whose
 * frame cache is bounded by VapourSynth).  Synchronizes if any state was held.  A
 * frame is bound again with svpf_smooth_set_sources before a render that needs it. */
int svpf_smooth_release(svpf_smooth* s, int32_t first, int32_t count);

#ifdef __cplusplus
}
#endif
#endif /* SVPF_H */
