/*
 * hipvr.h -- C ABI of libhipvr.so, the MI355X (gfx950) direct volume
 * renderer behind the Python VolumeRenderer.
 *
 * The reference renders with an OpenGL fragment shader driven through
 * ModernGL; this ABI replaces that device-resource layer (the reference's
 * ModernGLManager, pyvr/moderngl_renderer/manager.py:8-256) and the shader
 * (pyvr/shaders/volume.frag.glsl).  Paths below are relative to the reference
 * repository.
 *
 * Conventions
 *   - Every function returns HVR_OK (0) or a negative HVR_E* code; the
 *     message of the last failure on the calling thread is hvr_last_error().
 *   - Host arrays are copied at upload (as the reference's tobytes() copies);
 *     the caller keeps ownership of every host pointer.
 *   - A context (hvr_ctx) is used by one host thread at a time, like a GL
 *     context.  Different contexts may be used from different threads.
 *   - Volumes belong to the context that created them and must be destroyed
 *     before it.
 */
#ifndef HIPVR_H
#define HIPVR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HVR_ABI_VERSION 1

/* return codes */
#define HVR_OK 0
#define HVR_EINVAL -1   /* bad argument (null pointer, size, format)        */
#define HVR_EDEVICE -2  /* HIP runtime error                                 */
#define HVR_ENOMEM -3   /* device allocation failed                          */
#define HVR_ESTATE -4   /* call not valid in the current state               */

/* storage format of the scalar field in a volume */
#define HVR_SCALAR_NONE 0
#define HVR_SCALAR_F32 1   /* float32                                       */
#define HVR_SCALAR_F16 2   /* IEEE half, round to nearest even              */
#define HVR_SCALAR_U8 3    /* UNORM8: round(255 * clamp(v, 0, 1))           */

/* storage format of the normal field in a volume */
#define HVR_NORMAL_NONE 0
#define HVR_NORMAL_F32X3 1 /* 3 x float32                                   */
#define HVR_NORMAL_SN8X3 2 /* SNORM8: round(127 * clamp(n, -1, 1)) per axis */
#define HVR_NORMAL_F16X3 3 /* 3 x IEEE half, round to nearest even; scalar   */
                           /* NONE or F16 only (8-byte texels)               */

/* which field of a volume a sampler reads (see hvr_bind) */
#define HVR_ROLE_SCALAR 0  /* the scalar field                               */
#define HVR_ROLE_NORMAL 1  /* the normal field                               */

typedef struct hvr_ctx hvr_ctx;
typedef struct hvr_volume hvr_volume;
typedef struct hvr_batch hvr_batch;

/* Uniform block of one frame.  Replaces the shader uniforms of
 * volume.frag.glsl:3-19 set by renderer.py:99-105, 136-141, 163-172,
 * 303-316.  view/proj are the float32 matrices exactly as the reference
 * uploads them (numpy row-major bytes of Camera.get_view_matrix() and
 * Camera.get_projection_matrix(aspect), camera.py:292-369).  The library
 * inverts them in double precision. */
typedef struct {
    float view[16];
    float proj[16];
    float cam_pos[3];
    float bmin[3], bmax[3];        /* volume_min/max_bounds                  */
    float step;                    /* step_size                              */
    float ref_step;                /* reference_step_size                    */
    int32_t max_steps;
    float ambient, diffuse;        /* ambient_light, diffuse_light           */
    float light_pos[3], light_target[3];
    float et_threshold;            /* early termination alpha; the shader's  */
                                   /* constant is 0.99 (volume.frag.glsl:87) */
    int32_t camera_set;            /* 0: matrices never set; the shader then */
                                   /* sees zero matrices and every pixel is 0 */
    uint8_t clear_rgba[4];         /* framebuffer clear colour, 8-bit        */
                                   /* (clear_framebuffer, manager.py:212-215) */
    int32_t blend;                 /* 1: SRC_ALPHA, ONE_MINUS_SRC_ALPHA over */
                                   /* the clear colour (setup_blending,      */
                                   /* manager.py:217-220); 0: overwrite      */
} hvr_params;

/* ---- library / device ---- */

/* Version of this ABI (HVR_ABI_VERSION). */
int hvr_abi_version(void);

/* Debugging aid (no counterpart in the reference): the entry point named
 * `function` (e.g. "hvr_batch_run_async") fails with `code` (a negative
 * HVR_E* value) on its `after`-th call from now (0: the next call), once,
 * before doing any device work, with the message "injected fault in
 * <function>".  NULL disarms.  For testing how callers handle failures
 * without a failing device.  Process-wide; one armed entry point at a time.
 * Entry points with an injection point: hvr_create, hvr_resize,
 * hvr_volume_create, hvr_volume_upload_scalar, hvr_volume_upload_normals,
 * hvr_volume_upload_scalar_normals, hvr_volume_generate_blobs, hvr_set_lut,
 * hvr_set_params, hvr_bind, hvr_render_async, hvr_render_async_to,
 * hvr_read, hvr_sample_counts, hvr_batch_create, hvr_batch_add,
 * hvr_batch_run_async, hvr_batch_wait, hvr_batch_query. */
int hvr_debug_inject_fault(const char *function, int code, int after);

/* Message of the last failed call on this thread ("" if none). */
const char *hvr_last_error(void);

/* Number of visible HIP devices. */
int hvr_device_count(int *count);

/* Free and total device memory in bytes (no counterpart in the reference;
 * for memory-budgeted serving). */
int hvr_device_memory(int device, unsigned long long *free_bytes, unsigned long long *total_bytes);

/* ---- context: replaces ModernGLManager.__init__ (manager.py:14-41:
 * standalone context + RGBA8 framebuffer of width x height) ----
 * The march reads volumes below 4 GiB through buffer loads (32-bit
 * offsets); with the environment variable HVR_GATHER=global set when the
 * context is created it uses 64-bit global loads, as for larger volumes
 * (same results; for testing that path). */
int hvr_create(int device, int width, int height, hvr_ctx **out);
int hvr_destroy(hvr_ctx *ctx);
/* Change the framebuffer size (no counterpart in the reference, whose
 * renderer keeps the size it was created with: a new size there means a new
 * ModernGLManager, manager.py:14-41).  Volumes, LUT and bindings stay; the
 * parameters' matrices are the caller's (aspect = width / height); buffers
 * grow when the new size needs more pixels and are kept otherwise; only
 * growing waits for the context's stream (frames in flight keep their own
 * sizes).  Until the next frame at the new size, the framebuffer (and
 * hvr_read) holds the earlier frames' pixels at their own size. */
int hvr_resize(hvr_ctx *ctx, int width, int height);

/* ---- volumes: replace create_volume_texture (manager.py:77-108) and
 * create_normal_texture (manager.py:110-135) ----
 * A volume is one 3-D texture of GL size (s0, s1, s2) = the numpy shape of
 * the host array, holding a scalar field, a normal field, or both
 * interleaved in one texel.  Formats are fixed at creation. */
int hvr_volume_create(hvr_ctx *ctx, int s0, int s1, int s2, int scalar_fmt, int normal_fmt,
                      hvr_volume **out);
/* Upload the scalar field: `data` is s0*s1*s2 float32 values in numpy C
 * order (the bytes the reference passes to texture3d). */
int hvr_volume_upload_scalar(hvr_volume *vol, const float *data);
/* Upload the normal field: s0*s1*s2*3 float32 values in numpy C order. */
int hvr_volume_upload_normals(hvr_volume *vol, const float *data);
/* Device-side construction (no counterpart in the reference, whose normals
 * come from compute_normal_volume on the host, pyvr/datasets/synthetic.py:
 * 109-122, and are uploaded by create_normal_texture, manager.py:110-135).
 *
 * hvr_volume_upload_scalar_normals: upload the float32 scalar field (numpy
 * C order) once and compute the normals on the device, bit for bit as
 * compute_normal_volume computes them from that float32 array (np.gradient,
 * norm summed left to right, + 1e-8, division); every axis needs >= 2
 * voxels.  The volume must hold both fields.
 *
 * hvr_volume_generate_blobs: fill the scalar field (and the normals, when
 * the volume holds them) with a seeded Gaussian-blob field generated on the
 * device: axes linspace(-1, 1, s); per blob amp * exp(-(axis - centre)^2 /
 * (2 sigma^2)) per axis, multiplied in double; plus uniform noise in
 * [-noise, noise] from a counter-based hash of (seed, voxel index); clipped
 * to [0, 1].  centers: 3 * n_blobs values, n_blobs <= 32.  For large
 * serving volumes that would take minutes to build on the host. */
int hvr_volume_upload_scalar_normals(hvr_volume *vol, const float *data);
int hvr_volume_generate_blobs(hvr_volume *vol, int n_blobs, const double *centers, const double *sigmas,
                              const double *amps, double noise, unsigned long long seed);

/* Device bytes held by the volume's texels. */
int hvr_volume_bytes(const hvr_volume *vol, uint64_t *bytes);
/* The stored fields read back from the device, in numpy C order of the
 * volume's shape (no counterpart in the reference, whose textures are never
 * read back): `scalar` s0*s1*s2 float32 as the march reads them (FP16
 * widened, UNORM8 times 1/255), `normals` s0*s1*s2*3 float32 as stored
 * (SNORM8 as its integer in [-127, 127], FP16 widened); either may be NULL.
 * For checking device-built volumes against a CPU renderer.  Waits. */
int hvr_volume_download(hvr_volume *vol, float *scalar, float *normals);
int hvr_volume_destroy(hvr_volume *vol);

/* ---- transfer function: replaces create_rgba_transfer_function_texture
 * (manager.py:137-186) ----  `rgba` is n x 4 float32 (RGB from the colour
 * LUT, A from the opacity LUT).  n == 0 unbinds the LUT (an unbound sampler
 * reads (0, 0, 0, 1)).  The copy is queued on the context's stream behind
 * the frames already enqueued (through pinned staging), so a LUT change
 * does not wait for them; a LUT larger than any before it does. */
int hvr_set_lut(hvr_ctx *ctx, const float *rgba, int n);

/* ---- uniforms: replace set_uniform_* (manager.py:188-210) ---- */
int hvr_set_params(hvr_ctx *ctx, const hvr_params *params);

/* ---- sampler bindings: replace set_uniform_int("volume_texture" /
 * "normal_volume", unit) (renderer.py:133, 146) ----
 * The scalar sampler (volume.frag.glsl:93, `.r`) reads field `scalar_role`
 * of `scalar_vol`: the scalar, or the x component of the normal field.  The
 * normal sampler (:103, `.rgb`) reads field `normal_role` of `normal_vol`:
 * the normal, or (scalar, 0, 0) -- what a 3-channel GL sampler returns from
 * a 1-channel texture.  A NULL volume is an unbound sampler: the scalar
 * reads 0 and the normal (0, 0, 0) (GL incomplete texture).  The volume
 * must hold the requested field. */
int hvr_bind(hvr_ctx *ctx, hvr_volume *scalar_vol, int scalar_role, hvr_volume *normal_vol,
             int normal_role);

/* ---- render: replaces clear_framebuffer + setup_blending + render_quad +
 * read_pixels (manager.py:212-230; renderer.py:209-219) ----
 * Writes width*height*4 bytes, RGBA8, bottom row first, into `out` and
 * returns when they are there. */
int hvr_render(hvr_ctx *ctx, uint8_t *out);

/* Asynchronous form: enqueue the frame on the context's stream, leaving the
 * image in the context's device framebuffer (or, with HVR_READBACK_DIRECT,
 * its pinned host buffer).  hvr_read copies it out (and waits). */
int hvr_render_async(hvr_ctx *ctx);
int hvr_read(hvr_ctx *ctx, uint8_t *out);
int hvr_synchronize(hvr_ctx *ctx);

/* The read step without a host copy (replaces fbo.read of read_pixels,
 * manager.py:228-230): copies the framebuffer into the context's pinned host
 * buffer (HVR_READBACK_DIRECT: the frame is already there), waits, and sets
 * *rgba to it (width*height*4 bytes, bottom row
 * first; valid until the next read or render on this context), so that a
 * caller building its own object (a Python bytes) copies the image once. */
int hvr_read_pinned(hvr_ctx *ctx, const uint8_t **rgba);

/* Device pointer of the context's RGBA8 framebuffer (width*height*4 bytes;
 * frames of HVR_READBACK_DIRECT are not written there). */
int hvr_framebuffer(hvr_ctx *ctx, void **device_ptr);

/* Where frames of hvr_render / hvr_render_async go (no counterpart in the
 * reference, whose read_pixels copies the framebuffer, manager.py:228-230).
 * HVR_READBACK_COPY: the device framebuffer, copied out by hvr_read (in
 * chunks through pinned memory) or hvr_read_pinned.  HVR_READBACK_DIRECT
 * (the default): the march writes each pixel straight into the context's
 * pinned host buffer over PCIe as its ray finishes, so the image is on the
 * host when the kernel ends; hvr_read then waits and copies it out once,
 * hvr_read_pinned only waits.  Same bytes either way. */
#define HVR_READBACK_COPY 0
#define HVR_READBACK_DIRECT 1
int hvr_set_readback(hvr_ctx *ctx, int mode);

/* Samples composited per pixel for the current state (a separate counting
 * pass; not part of hvr_render).  `counts` is width*height uint32, bottom
 * row first. */
int hvr_sample_counts(hvr_ctx *ctx, uint32_t *counts);
/* The samples among those that load texels (no counterpart in the
 * reference; for the texel-traffic account of a frame): hvr_sample_counts
 * less the samples stepped over in transparent bricks when the frame's
 * march skips them (hvr_set_empty_skip, hvr_skip_info). */
int hvr_shaded_sample_counts(hvr_ctx *ctx, uint32_t *counts);
/* The bricks the current frame's samples load texels from (no counterpart
 * in the reference; the compulsory texel traffic of a frame is `touched` x
 * `brick_bytes`): a counting pass that marks them, of `bricks` in the bound
 * scalar volume.  Needs the fast path's conditions (8^3 or 16^3 bricks, a
 * LUT of at most 2048 entries); waits. */
int hvr_touched_bricks(hvr_ctx *ctx, unsigned long long *touched, unsigned long long *bricks,
                       unsigned long long *brick_bytes);

/* Milliseconds of the last frame's march kernel and ray-setup kernel,
 * measured with HIP events on the context's stream. */
int hvr_last_kernel_ms(hvr_ctx *ctx, float *march_ms, float *setup_ms);

/* Render into a caller-provided device buffer of width*height*4 bytes
 * instead of the context's framebuffer (no counterpart in the reference:
 * lets a caller keep many frames resident).  NULL means the framebuffer. */
int hvr_render_async_to(hvr_ctx *ctx, void *device_rgba8);

/* Device memory owned by the caller (for hvr_render_async_to). */
int hvr_device_alloc(hvr_ctx *ctx, uint64_t bytes, void **device_ptr);
int hvr_device_free(hvr_ctx *ctx, void *device_ptr);
/* Copy `bytes` from device memory to host memory (waits). */
int hvr_copy_to_host(hvr_ctx *ctx, void *host, const void *device_ptr, uint64_t bytes);
/* Device-to-device copy on the context's stream (waits); a plain copy of
 * known size, used to calibrate hardware traffic counters. */
int hvr_device_copy(hvr_ctx *ctx, void *dst, const void *src, uint64_t bytes);

/* Per-launch kernel timing over a region: after hvr_timing_begin every
 * frame records HIP events around its ray-setup and march kernels on the
 * context's stream (up to `max_frames` frames); hvr_timing_end waits and
 * returns the summed kernel milliseconds and the number of frames timed. */
int hvr_timing_begin(hvr_ctx *ctx, int max_frames);
int hvr_timing_end(hvr_ctx *ctx, double *march_ms, double *setup_ms, int *frames);

/* Transparent-brick skipping (no counterpart in the reference).  Samples in
 * bricks where every LUT entry that a density of the brick can weight has
 * alpha 0 (and a finite colour) contribute nothing; the march can count
 * them and step over them without loading texels.  The image and the sample
 * counts are identical in every mode.  HVR_SKIP_AUTO (the default) steps
 * over bricks when at least a quarter of the bound volume's bricks are
 * transparent for the current LUT.  The first frame after a LUT or volume
 * change computes the brick flags and waits for their count. */
#define HVR_SKIP_OFF 0
#define HVR_SKIP_ON 1
#define HVR_SKIP_AUTO 2
int hvr_set_empty_skip(hvr_ctx *ctx, int mode);

/* Transparent bricks of the volume the last frame's flags were computed for
 * (0 and 0 when none were), and whether the last frame's march stepped over
 * them. */
int hvr_skip_info(hvr_ctx *ctx, unsigned long long *clear_bricks, unsigned long long *bricks, int *used);

/* Cost-ordered dispatch of single frames (no counterpart in the reference).
 * A frame ends with its longest waves; in row order, long 16x16 ray blocks
 * that start after the first round of resident blocks set its end.  With
 * ordering, ray setup files every block under its rays' largest iteration
 * bound and the march starts the longest first.  The image and the sample
 * counts are identical in every mode.  HVR_ORDER_AUTO (the default) orders
 * frames of more blocks than the device keeps resident at once (8 per
 * compute unit) and at least 8 block rows. */
#define HVR_ORDER_OFF 0
#define HVR_ORDER_ON 1
#define HVR_ORDER_AUTO 2
int hvr_set_dispatch_order(hvr_ctx *ctx, int mode);
/* Whether the last frame was cost-ordered. */
int hvr_dispatch_order_info(hvr_ctx *ctx, int *used);

/* Tuning knobs (no counterpart in the reference): log2 of the brick edge
 * used by volumes created afterwards: 2, 3 or 4 (4^3, 8^3 or 16^3 bricks,
 * each stored with a one-texel apron; default 3). */
int hvr_set_brick_log2(hvr_ctx *ctx, int log2_edge);

/* ---- batches (no counterpart in the reference: PyVR renders one frame per
 * ModernGLVolumeRenderer.render() call, renderer.py:209-219).  A batch
 * holds many render requests, each with its own volume, image size,
 * camera/light/step parameters and LUT, and renders them all with one
 * persistent launch per volume layout, captured once in a hipGraph and
 * replayed.  Each request's image is the one hvr_render gives for the same
 * bindings, LUT and parameters (byte for byte, sample counts included).
 *
 * A request reads its scalar field from an uploaded volume of the batch's
 * context with brick edge 8 or 16 (volumes of 4 GiB and more are read with
 * 64-bit loads, below with buffer loads), and its normals from the
 * same volume (HVR_ROLE_NORMAL), its scalar (HVR_ROLE_SCALAR: (rho, 0, 0)),
 * or none (n = NULL); other bindings and LUTs of more than 2048 entries
 * need the single-request path.  Volumes must outlive the requests that
 * reference them (until hvr_batch_clear or hvr_batch_destroy).  A batch
 * owns its stream and buffers: several batches of one context may be in
 * flight at once; one batch is used by one host thread at a time. */
#define HVR_TILE_8X8 0   /* ray tile of a wavefront: 8 x 8 pixels (block 16 x 16) */
#define HVR_TILE_16X4 1  /* 16 x 4 pixels (block 32 x 8) */
int hvr_batch_create(hvr_ctx *ctx, hvr_batch **out);
int hvr_batch_destroy(hvr_batch *batch);
int hvr_batch_clear(hvr_batch *batch);
int hvr_batch_set_tile_shape(hvr_batch *batch, int shape);
/* Kernel variant of the batch's non-counting runs (counting runs use the
 * default): waves per block and whether transparent bricks are stepped over
 * (skip) and normals interpolated only for samples of non-zero alpha (gate).
 * Supported: (4, 1, 1) -- the default --, (2, 1, 1), (8, 1, 1), (4, 1, 0),
 * (4, 0, 1).  Every variant gives the same image; the autotuner times them. */
int hvr_batch_set_kernel(hvr_batch *batch, int waves_per_block, int skip, int gate);
/* Where the batch's march writes its images (no counterpart in the
 * reference): HVR_READBACK_DIRECT (the default) straight into the pinned
 * host pool through its device address, with no copy after the launch;
 * HVR_READBACK_COPY into a device pool copied to the pinned pool after the
 * launch.  Same bytes either way; applies from the next run. */
int hvr_batch_set_readback(hvr_batch *batch, int mode);
/* Add a request; lut_rgba holds lut_n RGBA float entries (lut_n = 0: no
 * LUT bound, reads (0, 0, 0, 1)); *index receives the request's index. */
int hvr_batch_add(hvr_batch *batch, hvr_volume *scalar_vol, int scalar_role, hvr_volume *normal_vol,
                  int normal_role, const float *lut_rgba, int lut_n, const hvr_params *params, int width,
                  int height, int *index);
/* Enqueue the whole batch on the batch's stream (count != 0: also record
 * per-pixel sample counts); the images end in pinned host memory (copied
 * there, or written there, per hvr_batch_set_readback).  hvr_batch_run
 * waits for them. */
/* Grow the batch's pools now to what a batch of up to `max_requests`
 * requests, `max_pixels` pixels in all and volumes of up to `max_bricks`
 * bricks needs (outputs, brick flags, staged block with LUTs of up to 2048
 * entries), so that runs within those limits never allocate (growing a
 * pool frees the old one, which waits for the device).  Waits for the
 * batch's stream. */
int hvr_batch_reserve(hvr_batch *batch, int max_requests, unsigned long long max_pixels,
                      unsigned long long max_bricks);
int hvr_batch_run_async(hvr_batch *batch, int count);
int hvr_batch_wait(hvr_batch *batch);
/* Non-blocking: *done = 1 when the last run's images (and counts) are in
 * host memory (also when the batch has not been run), else 0. */
int hvr_batch_query(hvr_batch *batch, int *done);
int hvr_batch_run(hvr_batch *batch, int count);
/* After the run completed: request `index`'s width*height*4 RGBA8 bytes
 * (bottom row first) / its per-pixel sample counts, valid until the next
 * run of the batch. */
int hvr_batch_output(hvr_batch *batch, int index, const uint8_t **rgba);
int hvr_batch_counts(hvr_batch *batch, int index, const uint32_t **counts);
/* Device time of the last run's graph (ms), without the copy back. */
int hvr_batch_last_ms(hvr_batch *batch, float *ms);

/* ---- frames without allocation ----
 * Size of the static device array (one per GPU, held until the process
 * ends) that hvr_render_async_to_noalloc uses: 4 KiB of request block and
 * one transparency flag per brick of the bound volume. */
#define HVR_NOALLOC_MAX_BRICKS (1u << 22)
#define HVR_NOALLOC_SCRATCH_BYTES (4096u + HVR_NOALLOC_MAX_BRICKS)

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory while rendering (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): the frame hvr_render_async_to renders for
 * the context's current bindings, LUT and parameters -- the same bytes --
 * into `device_rgba8` (width*height*4 bytes, bottom row first), enqueued on
 * the context's stream without waiting.  Everything the frame needs beyond
 * its inputs and output (request block, tile counter, one transparency flag
 * per brick) lives in a static device array of HVR_NOALLOC_SCRATCH_BYTES that
 * each GPU holds until the process ends; the call allocates nothing and
 * calls no allocating routine.  Limits (HVR_EINVAL / HVR_ESTATE with a
 * message otherwise): the scalar sampler reads an uploaded scalar field of a
 * volume with brick edge 8 or 16, the normals come from that volume or none,
 * the LUT has at most 2048 entries, the volume at most HVR_NOALLOC_MAX_BRICKS
 * bricks (4,194,304: e.g. 1024^3 with 8^3 bricks).  Every call on a GPU
 * shares that GPU's array, which keeps the last call's contents: calls that
 * overlap on one GPU (from any context or stream) race; calls on one stream
 * are ordered.  hvr_last_kernel_ms reports the frame's staging + flags and
 * march times. */
int hvr_render_async_to_noalloc(hvr_ctx *ctx, void *device_rgba8);

#ifdef __cplusplus
}
#endif

#endif /* HIPVR_H */
