/*
 * autoscoper_hip.h -- C ABI of the MI355X (gfx950) backend for Autoscoper's
 * 2D-3D registration cost path: DRR ray march -> image filters -> DRR mask
 * -> normalised cross-correlation.
 *
 * Plain pointers and sizes only. Every entry point returns an int status
 * (ASC_OK == 0, negative on error; asc_last_error() gives the message for
 * the calling thread). Nothing here calls exit(). Device pointers are HIP
 * device addresses; work is issued on the context's stream (the context's
 * own stream unless asc_context_set_stream() installed another one).
 *
 * Each declaration names the reference interface it replaces (paths relative
 * to the reference checkout, libautoscoper/src/...). The reference binds its
 * backend at compile time (gpu/opencl or gpu/cuda); the C++ facade in
 * autoscoper_xromm.hpp maps those class signatures onto this ABI.
 */
#ifndef AUTOSCOPER_HIP_H
#define AUTOSCOPER_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ASC_API __attribute__((visibility("default")))

/* ---- status codes -------------------------------------------------------- */
#define ASC_OK 0
#define ASC_ERR_INVALID (-1)      /* bad argument / handle */
#define ASC_ERR_HIP (-2)          /* HIP runtime error */
#define ASC_ERR_NOMEM (-3)        /* allocation failed / over budget */
#define ASC_ERR_EMPTY_VOLUME (-4) /* VolumeDescription.cpp:196-199 (exit(0) there) */
#define ASC_ERR_TOO_LARGE (-5)    /* View.cpp:448-455, Tracker.cpp:625-629 (throw there) */
#define ASC_ERR_UNSUPPORTED (-6)  /* e.g. bit depth, RadRenderer.cpp:93-96 */
#define ASC_ERR_NO_DEVICE (-7)
#define ASC_ERR_IO (-8)           /* a file is missing or malformed (the reference throws runtime_error) */

/* ---- handles and records ------------------------------------------------- */
typedef struct asc_context_s* asc_context;
typedef struct asc_volume_s* asc_volume;
typedef struct asc_image_s* asc_image;
typedef struct asc_scene_s* asc_scene;
typedef struct asc_server_s* asc_server;

/* Volume storage on the device. U16 is exact (the reference uploads
 * CL_UNORM_INT16, VolumeDescription.cpp:287-303); F16/BF16 store the same
 * normalised value rounded; F32 stores it exactly. */
enum { ASC_STORAGE_U16 = 0, ASC_STORAGE_F16 = 1, ASC_STORAGE_BF16 = 2, ASC_STORAGE_F32 = 3 };

/* Filter types, numbered as the reference's Filter enum (Filter.hpp:60-66). */
enum {
  ASC_FILTER_CONTRAST = 0, /* p = {alpha, beta, size}         ContrastFilter.cpp:57-86 */
  ASC_FILTER_SOBEL = 1,    /* p = {scale, blend}              SobelFilter.cpp:56-83 */
  ASC_FILTER_MEDIAN = 2,   /* declared but never implemented by the reference: rejected */
  ASC_FILTER_GAUSSIAN = 3, /* p = {radius}                    GaussianFilter.cpp:75-150 */
  ASC_FILTER_SHARPEN = 4   /* p = {radius, contrast, threshold} SharpenFilter.cpp:80-173 */
};

typedef struct {
  int type;
  int enabled; /* Filter::enabled(), Filter.hpp:86-87; a disabled filter is a copy */
  float p[4];
} asc_filter;

/* xromm::Camera state (Camera.hpp): coord_frame (rotation columns = camera
 * x, y, z axes in world space; translation = camera centre), image_plane
 * (4 world-space corners), viewport (film-plane centre and extent at z=-2),
 * size (pixels). */
typedef struct {
  double rotation[9];
  double translation[3];
  double image_plane[12];
  double viewport[4];
  double size[2];
} asc_camera;

/* gpu::VolumeDescription accessors (VolumeDescription.hpp:62-87). */
typedef struct {
  int dims[3];     /* cropped extent */
  int crop_min[3]; /* first non-zero voxel per axis after flipping */
  int crop_max[3];
  float inv_scale[3];
  float inv_trans[3];
  double trans_center[3];
  float min_value;
  float max_value;
  int storage;
  int reserved;
} asc_volume_desc;

/* ---- context, memory (OpenCL.cpp:668-782, 1027-1095) ---------------------- */
ASC_API const char* asc_last_error(void);
ASC_API int asc_version(void);
ASC_API int asc_device_count(int* count);
ASC_API int asc_context_create(int device, asc_context* out);
ASC_API int asc_context_destroy(asc_context ctx);
/* Install a caller-owned hipStream_t (NULL restores the context's own). */
ASC_API int asc_context_set_stream(asc_context ctx, void* hip_stream);
ASC_API void* asc_context_stream(asc_context ctx);
/* The HIP device index the context was created on. */
ASC_API int asc_context_device(asc_context ctx, int* device);
/* Blocking mode: each per-operator call (asc_drr_render ... asc_composite)
 * waits for its launch to finish before it returns, as the reference's
 * OpenCL backend does after every kernel (Kernel::launch, clFinish,
 * OpenCL.cpp:978). Off by default; scenes are not affected. */
ASC_API int asc_context_set_blocking(asc_context ctx, int enable);
ASC_API int asc_synchronize(asc_context ctx);

ASC_API int asc_malloc(asc_context ctx, size_t bytes, void** dptr);
ASC_API int asc_free(asc_context ctx, void* dptr);
/* Buffer::read (host -> device, blocking), OpenCL.cpp:1043-1049. */
ASC_API int asc_memcpy_h2d(asc_context ctx, void* dst, const void* src, size_t bytes);
/* Buffer::write (device -> host, blocking), OpenCL.cpp:1051-1057. */
ASC_API int asc_memcpy_d2h(asc_context ctx, void* dst, const void* src, size_t bytes);
/* Buffer::copy, OpenCL.cpp:1059-1067 (stream-ordered). */
ASC_API int asc_memcpy_d2d(asc_context ctx, void* dst, const void* src, size_t bytes);
/* Buffer::fill(float), OpenCL.cpp:1077-1095; gpu::fill in the CUDA backend. */
ASC_API int asc_fill_f32(asc_context ctx, float* dst, float value, size_t count);
/* Buffer::fill(char), OpenCL.cpp:1069-1075. */
ASC_API int asc_fill_u8(asc_context ctx, void* dst, unsigned char value, size_t bytes);

/* ---- a1: volume (VolumeDescription.cpp:152-305) --------------------------- */
/* data: width*height*depth voxels, x fastest, bps 8 or 16 (unsigned).
 * flip: Volume::flipX/Y/Z; scale: mm per voxel. Flips, crops to the
 * bounding box of non-zero voxels and uploads in `storage` form. */
ASC_API int asc_volume_create(asc_context ctx, const void* data, int width, int height, int depth, int bps,
                              const int flip[3], const double scale[3], int storage, asc_volume* out,
                              asc_volume_desc* desc);
ASC_API int asc_volume_describe(asc_volume vol, asc_volume_desc* desc);
/* Device layout used by the ray march: LINEAR (one texel per element),
 * QUAD (each element carries its 2x2 corner block in x/y: 4x memory, 2 loads
 * per sample), OCTET (each element carries its 2x2x2 cell: 8x memory, 1 load
 * per sample, x-fastest rows) or BRICK (three copies of the OCTET cells, one
 * per ray axis, each grouped in bricks of 1 x 2 x 4 cells across that axis,
 * so that the cells neighbouring rays sample share cache lines; a launch
 * reads the copy of its rays' main axis: 24x memory). Values and results are
 * identical; only the access pattern changes. asc_volume_create picks the
 * densest that fits (BRICK, else OCTET, QUAD, LINEAR): OCTET and QUAD need
 * slice strides below 8 MiB for the march's 24-bit address products, BRICK
 * needs W + H + D + 3 <= 4096 for its offset tables, and every array must
 * stay below 2 GiB (ASC_ERR_TOO_LARGE otherwise). */
enum { ASC_LAYOUT_LINEAR = 0, ASC_LAYOUT_QUAD = 1, ASC_LAYOUT_OCTET = 2, ASC_LAYOUT_BRICK = 3 };
ASC_API int asc_volume_set_layout(asc_volume vol, int layout);
ASC_API int asc_volume_destroy(asc_volume vol);
/* Device bytes held by the volume (for memory-budget admission). */
ASC_API int asc_volume_bytes(asc_volume vol, size_t* bytes);

/* ---- 2-D images: radiograph (RadRenderer::set_rad, RadRenderer.cpp:82-100)
 * and background (BackgroundRenderer::set_back, BackgroundRenderer.cpp:76-103).
 * bps 8 / 16: UNORM; bps 32: float32 as stored. */
ASC_API int asc_image_create(asc_context ctx, const void* data, int width, int height, int bps, asc_image* out);
ASC_API int asc_image_update(asc_image img, const void* data);
ASC_API int asc_image_destroy(asc_image img);

/* ---- a2/a3: geometry (host, double precision, reference operation order) - */
/* Frames are 12 doubles {rotation[9], translation[3]} as CoordFrame. */
ASC_API int asc_frame_from_xyzypr(const double xyzypr[6], double frame[12]);      /* CoordFrame.cpp:100-108 */
ASC_API int asc_frame_from_xyz_axis_angle(const double p[6], double frame[12]);   /* CoordFrame.cpp:163-175 */
ASC_API int asc_frame_to_xyzypr(const double frame[12], double xyzypr[6]);        /* CoordFrame.cpp:177-189 */
ASC_API int asc_frame_mul(const double a[12], const double b[12], double out[12]); /* CoordFrame.cpp:478-527 */
ASC_API int asc_frame_inverse(const double a[12], double out[12]);                 /* CoordFrame.cpp:335-353 */
/* a.linear_extrap(b): Tracker::optimize's guess 2 (CoordFrame.cpp:419-476). */
ASC_API int asc_frame_linear_extrap(const double a[12], const double b[12], double out[12]);
/* Camera::loadMayaCam2 arithmetic (Camera.cpp:229-362) from the parsed K, R
 * (file rows, row-major 3x3) and t, plus image size. */
ASC_API int asc_camera_from_mayacam2(const double K[9], const double R[9], const double t[3], const double size[2],
                                     asc_camera* out);
/* RayCaster::setInvModelView, RayCaster.cpp:92-142. */
ASC_API int asc_inv_model_view(const double imv16[16], const float inv_scale[3], const float inv_trans[3],
                               float out16[16]);
/* Tracker::calculate_viewport (Tracker.cpp:923-1032) for a model-view
 * frame (12 doubles): film-plane bounding box {x, y, w, h} of the volume's
 * box, clipped to the radiograph. Returns 1 in view, 0 not in view. */
ASC_API int asc_calculate_viewport(const double modelview[12], const asc_camera* cam, const asc_volume_desc* desc,
                                   double vp_out[4]);
/* Tracker.cpp:604-629 for one camera: model-view, inverse model-view (a2),
 * calculate_viewport (a3, Tracker.cpp:923-1032) and render size.
 * Returns 1 in view, 0 not in view (cost 9999), <0 error. */
ASC_API int asc_view_geometry(const asc_camera* cam, const double xyzypr[6], const asc_volume_desc* desc,
                              int render_w, int render_h, float imv_out[12], double vp_out[4],
                              unsigned size_out[2]);

/* ---- a5: DRR (RayCaster::render, RayCaster.cpp:158-193; RayCaster.cl) ----- */
ASC_API int asc_drr_render(asc_context ctx, asc_volume vol, const float imv[12], const float viewport[4],
                           float step, float intensity, float cutoff, float* out, unsigned width, unsigned height);
/* ---- a6: radiograph resample (RadRenderer::render, RadRenderer.cpp:117-145) */
ASC_API int asc_rad_render(asc_context ctx, asc_image img, const float image_plane[4], const float viewport[4],
                           float* out, unsigned width, unsigned height);
/* ---- a12: background mask (BackgroundRenderer::render, BackgroundRenderer.cpp:105-129) */
ASC_API int asc_background_render(asc_context ctx, asc_image img, const float image_plane[4],
                                  const float viewport[4], float threshold, float* out, unsigned width,
                                  unsigned height);

/* ---- a8-a11: filters (Filter::apply) --------------------------------------- */
ASC_API int asc_sobel(asc_context ctx, const float* in, float* out, int width, int height, float scale,
                      float blend);                                   /* SobelFilter.cpp:66-83 */
ASC_API int asc_contrast(asc_context ctx, const float* in, float* out, int width, int height, float alpha,
                         float beta, int size);                       /* ContrastFilter.cpp:66-86 */
/* GaussianFilter::set_radius weights (GaussianFilter.cpp:75-127); *size = N
 * (N*N weights written when capacity allows; N == 1 means "copy"). */
ASC_API int asc_gaussian_weights(float radius, float* out, int capacity, int* size);
ASC_API int asc_gaussian(asc_context ctx, const float* in, float* out, int width, int height,
                         float radius);                               /* GaussianFilter.cpp:129-150 */
ASC_API int asc_sharpen(asc_context ctx, const float* in, float* out, int width, int height, float radius,
                        float contrast, float threshold);             /* SharpenFilter.cpp:149-173 */
ASC_API int asc_filter_apply(asc_context ctx, const asc_filter* f, const float* in, float* out, int width,
                             int height);
/* View::filter (View.cpp:475-555): applies the chain with `scratch` as the
 * ping-pong buffer so that the last stage lands in `out`. */
ASC_API int asc_filter_chain(asc_context ctx, const asc_filter* filters, int count, const float* in, float* out,
                             float* scratch, int width, int height);

/* ---- a13: DRR mask and multiplies ----------------------------------------- */
ASC_API int asc_drr_background(asc_context ctx, const float* src, float* dest, unsigned width,
                               unsigned height);                      /* DRRBackground.cl */
ASC_API int asc_multiply(asc_context ctx, const float* a, const float* b, float* dest, unsigned width,
                         unsigned height);                            /* Mult.cl */
/* Tracker.cpp:747-752 in one pass: m = bg * (drr != 0); rad *= m; drr *= m.
 * bg may be NULL (all ones). mask receives m. */
ASC_API int asc_mask_multiply(asc_context ctx, float* drr, float* rad, const float* bg, float* mask,
                              unsigned width, unsigned height);

/* ---- a14: NCC (gpu::ncc, Ncc.cpp:180-220) ----------------------------------- */
/* Intended semantics of the reference (exact sums): returns 1e5 when the
 * denominator is < 1e-5. moments (optional, 9 doubles): {sum mask, sum f,
 * sum g, count(mask>0.5), S_f, S_g, S_ff, S_gg, S_fg} with the S_ over
 * mask > 0.5. Blocking (returns a host value, like the reference).
 * asc_context_set_ncc_form selects how asc_ncc computes:
 *   ASC_NCC_MOMENTS (default): compensated moments, FP64 combine;
 *   ASC_NCC_TREE_CUDA, ASC_NCC_TREE_OPENCL: the reference's own float
 *     computation, launch by launch: six tree sums (NccSum.cl / the CUDA
 *     sum kernel; 256-thread blocks with contiguous indexing, or 128-thread
 *     blocks with the OpenCL kernel's index t + b * blocks as written), each
 *     ended by a blocking read, float means, the products kernel (Ncc.cl),
 *     den = sqrt(float * float), 1e5 below 1e-5. The values are the
 *     reference backend's; moments then holds {sum mask, sum f, sum g,
 *     sum products, sum f dev^2, sum g dev^2} as tree-summed floats
 *     (products NaN when the reference skips that sum) and NaN in 6..8. */
enum { ASC_NCC_MOMENTS = 0, ASC_NCC_TREE_CUDA = 1, ASC_NCC_TREE_OPENCL = 2 };
ASC_API int asc_context_set_ncc_form(asc_context ctx, int form);
ASC_API int asc_ncc(asc_context ctx, const float* f, const float* g, const float* mask, unsigned n, double* ncc,
                    double moments[9]);
/* No counterpart in the reference. asc_ncc (ASC_NCC_MOMENTS form, whatever
 * asc_context_set_ncc_form says) for callers that cannot allocate: code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc.
 * It calls no allocation routine and does not wait: the two moment kernels
 * run on the context's stream (blocking contexts wait, as every per-operator
 * call does), and out10 (caller-owned device memory, 10 doubles) receives the
 * nine moments in asc_ncc's order, then the NCC, the same bits as asc_ncc.
 * The block partials live in a static device array of the library, one per
 * GPU, held until the process ends (316 KB): n is limited to
 * ASC_NCC_NOALLOC_MAX_N (3000 x 3000, the reference View's buffer size);
 * larger n returns ASC_ERR_TOO_LARGE. All calls on one GPU share that array:
 * calls that overlap on one GPU (two streams, two threads) race, so keep
 * them on one stream. */
#define ASC_NCC_NOALLOC_MAX_N 9000000u
ASC_API int asc_ncc_noalloc(asc_context ctx, const float* f, const float* g, const float* mask, unsigned n,
                            double* out10);
/* ncc_init / ncc_deinit (Ncc.cpp:151-178): reserve scratch for n elements. */
ASC_API int asc_ncc_init(asc_context ctx, unsigned max_n);
ASC_API int asc_ncc_deinit(asc_context ctx);

/* ---- a19/a20: display helpers ---------------------------------------------- */
ASC_API int asc_merge(asc_context ctx, const float* a, const float* b, float* dest, unsigned width,
                      unsigned height);                               /* Merger.cl */
ASC_API int asc_composite(asc_context ctx, const float* src1, const float* src2, const float* src3,
                          const float* src4, float* dest, unsigned width, unsigned height); /* Compositor.cl */

/* ---- a15/a16 batched: scene + fused evaluation ----------------------------
 * A scene holds what Tracker::load sets up (Tracker.cpp:244-335): volumes
 * with their DRR parameters and pivots, cameras with radiograph, optional
 * background and DRR/rad filter chains, and the render resolution. */
ASC_API int asc_scene_create(asc_context ctx, int render_width, int render_height, asc_scene* out);
ASC_API int asc_scene_destroy(asc_scene scene);
/* pivot: 12-double frame (Trial::getVolumeMatrix); NULL = identity. */
ASC_API int asc_scene_add_volume(asc_scene scene, asc_volume vol, float step, float intensity, float cutoff,
                                 const double pivot[12], int* index);
ASC_API int asc_scene_set_drr_params(asc_scene scene, int volume, float step, float intensity, float cutoff);
/* Trial::getVolumeMatrix(volume) as a 12-double frame. */
ASC_API int asc_scene_volume_pivot(asc_scene scene, int volume, double pivot[12]);
ASC_API int asc_scene_add_camera(asc_scene scene, const asc_camera* cam, asc_image rad, int* index);
ASC_API int asc_scene_set_rad(asc_scene scene, int camera, asc_image rad);
ASC_API int asc_scene_set_background(asc_scene scene, int camera, asc_image bg, float threshold);
/* which: 0 = DRR chain, 1 = radiograph chain (View::drrFilters/radFilters). */
ASC_API int asc_scene_set_filters(asc_scene scene, int camera, int which, const asc_filter* filters, int count);
/* Evaluate n absolute poses (Tracker::trackFrame, Tracker.cpp:583-773).
 * volumes[n] picks the volume per pose. per_camera (n*ncam, optional) gets
 * 1 - NCC or 9999 per camera; costs (n, optional) the minimizationFunc
 * total (Tracker.cpp:814-826). Fused: all poses x cameras in one set of
 * launches. */
ASC_API int asc_scene_eval_poses(asc_scene scene, const int* volumes, const double* xyzypr, int n, double* costs,
                                 double* per_camera);
/* Evaluate n manipulator offsets relative to stored float poses
 * (Tracker::minimizationFunc, Tracker.cpp:775-827): stores n*7 floats
 * {x, y, z, qw, qx, qy, qz}, manips n*6 doubles. */
ASC_API int asc_scene_eval_manips(asc_scene scene, const int* volumes, const float* stores, const double* manips,
                                  int n, double* costs, double* per_camera);

/* Several jobs in one scene (a serving loop): volumes and cameras are
 * reusable slots; a camera set is one job's views with its render size
 * (Tracker's render_width/height); a frame is one radiograph per camera of
 * a set (Video frame, Tracker.cpp:380-385). These have no counterpart in the
 * reference, which holds one trial per process; they let requests of
 * different jobs share one fused launch. */
ASC_API int asc_scene_remove_volume(asc_scene scene, int volume);
ASC_API int asc_scene_remove_camera(asc_scene scene, int camera);
ASC_API int asc_scene_add_camera_set(asc_scene scene, const int* cameras, int count, int render_width,
                                     int render_height, int* set);
ASC_API int asc_scene_remove_camera_set(asc_scene scene, int set);
ASC_API int asc_scene_add_frame(asc_scene scene, int set, const asc_image* rads, int* frame);
ASC_API int asc_scene_remove_frame(asc_scene scene, int frame);
/* n requests, each a pose of volumes[i] seen by camera set sets[i] at frame
 * frames[i] (frames may be NULL or -1: each camera's own radiograph), all in
 * one fused launch. per_camera (optional) is n x 8: camera k of request i at
 * i*8 + k. costs[i] = sum over the set's cameras of 1 - NCC (9999 rules of
 * Tracker.cpp:614-618, 814-826). */
ASC_API int asc_scene_eval_requests(asc_scene scene, int n, const int* volumes, const int* sets, const int* frames,
                                    const double* xyzypr, double* costs, double* per_camera);
/* The same batch in two halves, so that a caller can prepare and queue the
 * next batch while one runs (no counterpart in the reference; the
 * registration server's double-buffered batcher): submit queues the fused
 * launches on the scene's stream behind any batch already submitted and
 * returns a ticket at once; finish waits for that batch and writes its costs
 * and per-camera values as eval_requests does. At most two batches may be
 * submitted and not finished; they finish in any order, and the last one
 * finished is the "last batch" of asc_scene_last_*, which refuse while a
 * batch is submitted. Submit refuses while graph replay is on. */
ASC_API int asc_scene_submit_requests(asc_scene scene, int n, const int* volumes, const int* sets,
                                      const int* frames, const double* xyzypr, int* ticket);
ASC_API int asc_scene_finish_requests(asc_scene scene, int ticket, double* costs, double* per_camera);
/* Blocks until the submitted batch's launches are done (no counterpart in
 * the reference). It touches no other scene state, so a thread may wait here
 * while another thread submits (the scene's other calls are not
 * thread-safe). */
ASC_API int asc_scene_wait_requests(asc_scene scene, int ticket);
/* Device bytes the scene holds for batch scratch (DRR images, moment
 * partials, results, descriptors); they grow with the largest batch and are
 * kept for reuse. No counterpart in the reference (memory accounting of the
 * registration server). */
ASC_API int asc_scene_scratch_bytes(asc_scene scene, size_t* bytes);
/* Record the fused launches of an n-pose batch in a hipGraph and replay it
 * for later batches of the same size (1 = on, 0 = off). */
ASC_API int asc_scene_use_graph(asc_scene scene, int enable);
/* Diagnostics of the last batch: info[n*ncam*8] = {in_view, w, h,
 * mask_count, tiles, 0, 0, 0}, samples = total ray samples. */
ASC_API int asc_scene_last_info(asc_scene scene, int* info, int capacity, int64_t* samples);
/* Keep the filtered, masked DRR/radiograph and mask of every item of the
 * following batches for asc_scene_last_images (costs extra stores). */
ASC_API int asc_scene_set_debug(asc_scene scene, int enable);
/* Launch-shape knobs of the fused path (asc_scene_autotune sets them):
 * ASC_TUNE_UNROLL: ray samples per loop iteration (1, 2 or 4; default 4);
 * ASC_TUNE_GRID: persistent grid size in blocks, multiple of 8 (0 = CUs x
 * occupancy);
 * ASC_TUNE_WAVE_PATCH: width of the ray patch one wave marches inside a
 * 16 x 16 tile: 8 (8 x 8, default), 16 (16 x 4) or 4 (4 x 16).
 * None of them changes a result.
 * ASC_TUNE_GATHER_PROBE (no counterpart in the reference; a measurement
 * tool): 1 replaces the ray march by its memory access stream alone (the
 * same rays, sample counts and 16-byte cell loads, with as little
 * arithmetic as the addresses need), so that the kernel's duration is the
 * gather bandwidth the memory hierarchy sustains for this access pattern;
 * results are meaningless while it is set (16-bit OCTET/BRICK volumes). */
enum { ASC_TUNE_UNROLL = 0, ASC_TUNE_GRID = 1, ASC_TUNE_WAVE_PATCH = 2, ASC_TUNE_GATHER_PROBE = 3 };
ASC_API int asc_scene_set_tuning(asc_scene scene, int key, int value);

/* Autotuner: times every combination of unroll {2, 4}, wave patch {8, 16,
 * 4} and grid {1, 2} x (CUs x occupancy) on a representative batch (n poses
 * of `volumes`, all cameras), each `reps` times with HIP events around the
 * ray-march kernel, keeps the fastest by median and sets it on the scene.
 * result = {unroll, wave patch, grid, best ms, default ms} (the default is
 * unroll 4, patch 8, grid 1x). */
typedef struct {
  int unroll, wave_patch, grid;
  double best_ms, default_ms;
  int candidates;
} asc_tune_result;
ASC_API int asc_scene_autotune(asc_scene scene, int n, const int* volumes, const double* poses, int reps,
                               asc_tune_result* result);
/* Time the fused kernels of eager (non-graph) batches with HIP events on
 * the scene's stream; ms = {ray march, filters+moments, finalize} of the
 * last batch. */
ASC_API int asc_scene_set_timing(asc_scene scene, int enable);
ASC_API int asc_scene_last_timing(asc_scene scene, float ms[3]);
/* moments[n*ncam*10]: the nine NCC moments (asc_ncc order) + NCC per item. */
ASC_API int asc_scene_last_moments(asc_scene scene, double* moments, int capacity);
/* Copy the last batch's filtered, masked DRR and radiograph and the mask of
 * (pose i, camera c) to host buffers of w*h floats (any may be NULL). With
 * submitted batches, only when the batch finished last was also submitted
 * last (the device holds the images of the last one submitted). */
ASC_API int asc_scene_last_images(asc_scene scene, int pose, int camera, float* drr, float* rad, float* mask);
/* Plain form: one pose evaluated camera by camera with one launch per
 * operator and a blocking NCC readback per camera, as trackFrame does. */
ASC_API int asc_scene_track_frame_plain(asc_scene scene, int volume, const double xyzypr[6], double* per_camera,
                                        double* cost);
/* Display composite of one scene camera over its whole image plane
 * (View::render, View.cpp:355-394, with the viewport of Tracker::getFullDRR,
 * Tracker.cpp:1056-1064): every volume's DRR at its pose (poses: n_poses x 6
 * doubles, xyzypr, one per scene volume slot) merged, the DRR filters, the radiograph and
 * its filters, the background mask, then the Compositor's 3 floats per
 * pixel into device memory `rgb` (capacity in floats). The size is
 * (unsigned)(viewport_w * render_w / viewport_w) by the same for h, returned
 * in width/height; rgb == NULL only reports it. */
ASC_API int asc_scene_render_view(asc_scene scene, int camera, const double* poses, int n_poses, float* rgb,
                                  size_t rgb_capacity, unsigned* width, unsigned* height);

/* ---- pose store helpers (KeyCurve<float> / Quatf, KeyCurve.hpp:330-356,
 * Quaternion.hpp:80-94,220-243) -------------------------------------------- */
ASC_API int asc_pose_store_set(float store[7], const double xyzypr[6]);
ASC_API int asc_pose_store_get(const float store[7], double xyzypr[6]);
/* minimizationFunc pose composition (Tracker.cpp:777-796). */
ASC_API int asc_compose_pose(const float store[7], const double pivot[12], const double manip[6],
                             double xyzypr[6]);

/* ---- a17/a18: optimizers (Tracker::optimize, Tracker.cpp:337-580) --------- */
typedef struct {
  int method;          /* 0 PSO + downhill simplex, 1 simplex only */
  unsigned max_iter;   /* PSO MAX_EPOCHS */
  unsigned max_stall;  /* PSO MAX_STALL */
  double min_limit;    /* PSO START_RANGE_MIN */
  double max_limit;    /* PSO START_RANGE_MAX */
  double offsets[6];   /* simplex vertex offsets (Trial::offsets) */
  unsigned seed;       /* srand seed (Autoscoper_RANDOM_SEED) */
  int particles;       /* NUM_OF_PARTICLES (100 in the reference) */
  int synchronous;     /* 0: reference asynchronous PSO, one eval per call;
                          1: synchronous PSO, all particles of an epoch in one fused batch */
  int volume;          /* volume index (Trial::current_volume) */
  int speculative;     /* 0: one simplex evaluation per call, in the reference's order;
                          1: each simplex iteration evaluates its reflection, expansion and
                          both contraction candidates together, and a shrink all its new
                          vertices together: the same trajectory in fewer, larger batches */
} asc_optimize_options;

typedef struct {
  int pso_epochs;
  int simplex_iterations;
  double pso_best_cost;
  double final_cost;
  double initial_cost;
  int64_t evaluations;
} asc_optimize_stats;

/* One frame of Tracker::optimize with guess 0: store (7 floats) is updated
 * in place with the optimised pose. */
ASC_API int asc_optimize_frame(asc_scene scene, float store[7], const asc_optimize_options* opt,
                               asc_optimize_stats* stats);

/* Cost of n points x[n][ndim] into costs[n]; returns ASC_OK or an error
 * that stops the optimiser. The reference binds its cost through the
 * global callbacks PSO_FUNC / FUNC (PSO.hpp:23, DownhillSimplex.hpp:27,
 * Tracker.cpp:88-96); n is 1 except for a synchronous PSO epoch. */
typedef int (*asc_cost_fn)(const double* x, int n, double* costs, void* user);

/* Replaces pso(start_range_min, start_range_max, MAX_EPOCHS, MAX_STALL)
 * (PSO.hpp:58, PSO.cpp:198-295) with srand(seed) (PSO.cpp:131-149) on a
 * private generator. particles = NUM_OF_PARTICLES (100 in the reference);
 * synchronous = 0 is the reference's order. best = gBest.Position,
 * best_cost = gBest.NCC, epochs = the epoch count the reference prints. */
ASC_API int asc_pso(asc_cost_fn fn, void* user, unsigned seed, float min_limit, float max_limit,
                    unsigned max_epochs, unsigned max_stall, int particles, int synchronous, float best[6],
                    float* best_cost, int* epochs);

/* Replaces AMOEBA(P, Y, NDIM, FTOL, ITER) (DownhillSimplex.hpp:28,
 * DownhillSimplex.cpp:39-172): P is (ndim+1) x ndim row-major and Y the
 * costs of its rows, both 0-indexed here (1-indexed MAT in the reference);
 * ndim <= 20 as the reference's MAT. */
ASC_API int asc_simplex(asc_cost_fn fn, void* user, double* P, double* Y, int ndim, double ftol,
                        int* iterations);

/* Tracker::optimize for one frame (Tracker.cpp:337-580, guess handled by
 * the caller) on any cost of manipulator offsets: fn(x, n, costs, user)
 * evaluates n offsets relative to the pose in `store`, which is updated in
 * place after the PSO and at the end, as asc_optimize_frame does it on a
 * scene. pivot = Trial::getVolumeMatrix. opt->volume is not used. */
ASC_API int asc_optimize_frame_fn(asc_cost_fn fn, void* user, float store[7], const double pivot[12],
                                  const asc_optimize_options* opt, asc_optimize_stats* stats);

/* ---- serving: continuous batching of tracking jobs ------------------------ */
/* No counterpart in the reference (one trial per GUI process, one blocking
 * frame at a time). A job is a trial: bones (16-bit volumes), a camera set
 * with render size and filter chains, a stack of radiograph frames, and
 * Tracker::optimize options; each bone runs the frame loop with the given
 * guess (1: previous frame, 2: linear extrapolation of the two previous,
 * Tracker.cpp:380-427; frame 0 starts at initial_poses). The server admits
 * jobs first-come first-served while they fit the device-memory budget and
 * the number of active jobs, and evaluates the pending cost requests of all
 * running bones in one fused launch per tick. Results do not depend on what
 * else runs. */
typedef struct {
  int n_volumes;
  const void* const* volumes; /* n_volumes pointers to W*H*D u16 voxels, x fastest */
  const int* volume_dims;     /* n_volumes x {W, H, D} */
  const double* voxel_mm;     /* n_volumes x 3 */
  const double* initial_poses; /* n_volumes x 6: xyzypr of frame 0 */
  int storage;                /* ASC_STORAGE_* */
  float drr_step, drr_intensity, drr_cutoff;
  int n_cameras;              /* 1 .. 8 */
  const asc_camera* cameras;
  int render_width, render_height;
  const asc_filter* drr_filters;
  int n_drr_filters;
  const asc_filter* rad_filters;
  int n_rad_filters;
  int n_frames;
  const void* const* frames;  /* n_frames * n_cameras pointers, frame-major */
  const int* frame_dims;      /* n_cameras x {width, height} */
  int frame_bps;              /* 8, 16 or 32 */
  asc_optimize_options options; /* options.volume is ignored */
  int guess;
} asc_job_desc;

typedef struct {
  double queue_ms;  /* arrival -> admission */
  double total_ms;  /* arrival -> last frame done */
  int64_t evaluations;
  int frames;
} asc_job_stats;

typedef struct {
  int64_t jobs_done, frames_done, batches, requests, evaluations;
  double busy_ms;     /* time inside fused launches */
  double elapsed_ms;  /* since the server started */
  double latency_p50_ms, latency_p99_ms, latency_max_ms; /* per-frame latency of finished jobs */
  int active_jobs, waiting_jobs;
  size_t device_bytes_in_use;
} asc_serve_stats;

ASC_API int asc_server_create(asc_context ctx, size_t hbm_budget, int max_active_jobs, asc_server* out);
/* Launches the batcher keeps in flight (no counterpart in the reference,
 * which evaluates one pose at a time): 2 (default) queues the next fused
 * launch while one runs. With nothing in flight it starts a launch once half
 * of the running bones and streams have a request pending, if those hold at
 * least 32 poses; the rest start the next launch while that one runs. The
 * requesters so settle into two cohorts whose launches alternate, and the
 * GPU does not wait for one cohort's host work (optimiser steps, batch
 * preparation). 1 launches once every running bone has a request pending
 * and waits for it before the next. Results are the same either way. */
ASC_API int asc_server_set_pipeline(asc_server server, int depth);
/* Copies the job's host data; returns at once. */
ASC_API int asc_server_submit(asc_server server, const asc_job_desc* job, int* job_id);
/* Blocks until the job is done: poses n_frames x n_volumes x 6, per-frame
 * latency n_frames (ms). A job can be waited for once. */
ASC_API int asc_server_wait(asc_server server, int job_id, double* poses, double* frame_latency_ms,
                            asc_job_stats* stats);
ASC_API int asc_server_stats(asc_server server, asc_serve_stats* out);
/* Front ends that keep their own trial state (the socket session,
 * asc_remote_create_shared; reference: the socket's getNCC and optimizeFrame
 * messages, Socket.cpp:282-302,345-372, which run one client's requests
 * alone there) share the server's scene and its fused launches. Under
 * asc_server_lock_scene the batcher starts no launch (those it started may
 * still run; they read no entry of a front end whose own requests are
 * done): the front end adds or
 * removes its volumes, cameras, camera sets and frames with asc_scene_*
 * calls, then unlocks (same thread). asc_server_eval_requests evaluates n
 * requests as asc_scene_eval_requests does, in the batcher's next launch
 * together with every other pending request, and blocks until done; a
 * thread holding the scene lock may not call it. */
ASC_API int asc_server_lock_scene(asc_server server, asc_scene* scene);
ASC_API int asc_server_unlock_scene(asc_server server);
ASC_API int asc_server_eval_requests(asc_server server, int n, const int* volumes, const int* sets,
                                     const int* frames, const double* xyzypr, double* costs, double* per_camera);
/* A front end about to post a stream of requests, one after another (an
 * optimisation: Tracker::optimize's cost calls, Tracker.cpp:476-570),
 * announces it; until the matching asc_server_stream_end the batcher, before
 * it launches, waits for a request from every announced stream as it does
 * for the running jobs' bones (at most 0.5 ms for stragglers), so that
 * concurrent streams share launches. Requests of streams and any other
 * external requests count alike. */
ASC_API int asc_server_stream_begin(asc_server server);
ASC_API int asc_server_stream_end(asc_server server);
/* Drops queued jobs, lets running ones finish, then frees everything. */
ASC_API int asc_server_destroy(asc_server server);

/* ---- on-disk inputs and outputs (host only, no device) --------------------- */

/* TIFF images as libautoscoper reads them through libtiff (TiffImage.cpp,
 * Volume.cpp:51-140, Video.cpp:66-239): strips, uncompressed, LZW or
 * PackBits, predictor 1 or 2, 8/16/32 bits per sample. A volume file holds
 * one slice per page (all pages alike, one sample per pixel, unsigned). The
 * pixels come out pages x height x width, x fastest, in host byte order. */
ASC_API int asc_tiff_info(const char* path, int* width, int* height, int* pages, int* bits_per_sample);
ASC_API int asc_tiff_read(const char* path, void* dst, size_t dst_bytes);

/* xromm::Camera(filename) (Camera.cpp:111-140): a ".json" file is a VTKCam
 * 1.0 camera, a file whose first line is "image size" a MayaCam 2.0 camera,
 * any other a MayaCam 1.0 camera. */
ASC_API int asc_read_camera(const char* path, asc_camera* out);
/* Camera::loadMayaCam1 after parsing (Camera.cpp:141-227): the file's five
 * lines of three values, row after row. */
ASC_API int asc_camera_from_mayacam1(const double values[15], asc_camera* out);
/* Camera::loadVTKCamera after parsing (Camera.cpp:364-422). */
ASC_API int asc_camera_from_vtkcam(const double position[3], const double focal_point[3], const double view_up[3],
                                   double view_angle_deg, const double size[2], asc_camera* out);

/* A filter settings file (.vie) as the GUI applies it to a camera
 * (FilterTreeWidget::loadFilterSettings, FilterTreeWidget.cpp:290-329):
 * every value passes the GUI's spin box (rounded to 2 decimals, then bounded
 * to its range), the Sobel keys are swapped as the GUI reads them (`Scale`
 * sets blend, `Blend` sets scale; FilterTreeWidgetItem.cpp:158-165,298-299),
 * and the DRR renderer's values map to step = exp(7 SD - 5) and intensity =
 * exp(15 RI - 5) (ModelViewTreeWidgetItem.cpp:321-338). The cutoff is a
 * fraction of the volume's value range: asc_drr_cutoff maps it. */
#define ASC_MAX_FILTERS 16
typedef struct {
  int has_drr_renderer, has_drr_filters, has_rad_filters; /* sections present in the file */
  int n_drr_filters, n_rad_filters;
  asc_filter drr_filters[ASC_MAX_FILTERS];
  asc_filter rad_filters[ASC_MAX_FILTERS];
  double sample_distance, ray_intensity, cutoff_fraction; /* spin box values (defaults 0.62, 0.49, 0) */
  float drr_step, drr_intensity;                          /* RayCaster::setSampleDistance, setRayIntensity */
} asc_filter_settings;
ASC_API int asc_read_filter_settings(const char* path, asc_filter_settings* out);
/* RayCaster::setCutoff(fraction * (max - min) + min) over the volume's
 * normalised value range (ModelViewTreeWidgetItem.cpp:336-338). */
ASC_API int asc_drr_cutoff(double cutoff_fraction, const asc_volume_desc* desc, float* cutoff);

/* A trial configuration (.cfg v1.0 to 1.2; xromm::Trial(filename),
 * Trial.cpp:74-257): cameras are read at open, volumes and frames on
 * demand. Paths of version 1.1 and later are relative to the file. Frames
 * of a camera are the sorted *.tif* files of its directory; a frame index
 * past a camera's last frame reads its last frame (Video::set_frame). */
typedef struct asc_trial_s* asc_trial;
typedef struct {
  int version[2];
  int n_cameras, n_volumes;
  int n_frames;                     /* the longest camera's frame count */
  int render_width, render_height;  /* RenderResolution, default 512 x 512 */
  double offsets[6];                /* OptimizationOffsets, default 0.1 */
} asc_trial_info;
ASC_API int asc_trial_open(const char* cfg_path, asc_trial* out);
ASC_API int asc_trial_get_info(asc_trial trial, asc_trial_info* out);
ASC_API int asc_trial_camera(asc_trial trial, int camera, asc_camera* out);
/* dims = {W, H, D} of the stack as stored (before VolumeFlip, which
 * asc_volume_create applies); voxel_mm = VoxelSize; flip = VolumeFlip. */
ASC_API int asc_trial_volume_info(asc_trial trial, int volume, int dims[3], int* bits_per_sample, double voxel_mm[3],
                                  int flip[3]);
ASC_API int asc_trial_volume_read(asc_trial trial, int volume, void* dst, size_t dst_bytes);
ASC_API int asc_trial_frame_info(asc_trial trial, int camera, int frame, int dims[2], int* bits_per_sample);
ASC_API int asc_trial_frame_read(asc_trial trial, int camera, int frame, void* dst, size_t dst_bytes);
ASC_API int asc_trial_frame_path(asc_trial trial, int camera, int frame, char* dst, size_t dst_bytes);
/* Video::create_background_image (Video.cpp:147-218): per pixel, the
 * largest normalised value over all of the camera's frames (count = W*H). */
ASC_API int asc_trial_background(asc_trial trial, int camera, float* dst, size_t count);
ASC_API int asc_trial_close(asc_trial trial);

/* Tracking results (.tra; AutoscoperMainWindow::save_tracking_results and
 * load_tracking_results, AutoscoperMainWindow.cpp:736-830,983-1092): one line
 * per frame, per volume 6 values x y z yaw pitch roll or the 16 of the 4x4
 * matrix (column order, or row order with ASC_TRA_ROWS), printed "%.16f";
 * a pose with any NaN component (an untracked frame) is written as NaNs.
 * poses = n_frames x n_volumes x 6 (mm, degrees). Reading fills the poses
 * of the lines present and leaves the others NaN. */
enum { ASC_TRA_MATRIX = 1, ASC_TRA_ROWS = 2, ASC_TRA_COMMAS = 4, ASC_TRA_CM = 8, ASC_TRA_RADIANS = 16 };
ASC_API int asc_write_tra(const char* path, const double* poses, int n_frames, int n_volumes, int flags);
ASC_API int asc_read_tra(const char* path, int n_frames, int n_volumes, int flags, double* poses, int* frames_read);

/* Display output, headless (the GUI shows the composite with
 * glDrawPixels(GL_RGB, GL_FLOAT), GLView.cpp:473-527).
 * asc_write_png: 8-bit greyscale (channels 1) or RGB (3) rows, top first.
 * asc_write_view_png: an asc_scene_render_view composite copied to the
 *   host (3 floats per pixel, row 0 at the bottom as GL draws it), clamped
 *   to [0, 1] and rounded to 8 bits, written top row first.
 * asc_write_full_drr_pgm: save_full_drr (Tracker.cpp:155-204), the image of
 *   Tracker::getFullDRR: min-max stretched, inverted, ASCII PGM. */
ASC_API int asc_write_png(const char* path, const uint8_t* pixels, int width, int height, int channels);
ASC_API int asc_write_view_png(const char* path, const float* rgb, int width, int height);
ASC_API int asc_write_full_drr_pgm(const char* path, const float* image, int width, int height);

/* ---- remote control: the reference's socket protocol, version 2 ----------- */
/* The server side of autoscoper/src/net/Socket.cpp:53-444 (the protocol
 * PyAutoscoper speaks, scripts/python/PyAutoscoper/connect.py): a session
 * holding a trial, its cameras, volumes, frames, filter settings,
 * background threshold and per-volume pose curves (KeyCurve.hpp: cubic
 * Hermite x, y, z with smooth tangents, slerped orientation), answering
 * the 18 message types on 127.0.0.1. Costs and optimisation (messages 8
 * and 11) run on this library's fused scene, on `ctx` (NULL: a session
 * without a device, whose device requests fail). Only the NCC cost is
 * built (message 11 with cf_model 1, SAD, fails). */
typedef struct asc_remote_s* asc_remote;
ASC_API int asc_remote_create(asc_context ctx, asc_remote* out);
/* A session whose costs and optimisations (messages 8 and 11,
 * Socket.cpp:282-302,345-372) run as requests of the registration server
 * `server` (asc_server_eval_requests): its volumes, cameras and camera set
 * live in the server's scene, so the requests of several sessions and of
 * the server's own jobs share fused launches; a request's cost does not
 * depend on what shares its launch, so the answers are the same bits as a
 * session of its own. `ctx` (same device) holds the session's uploads and
 * its full-view renders (messages 10 and 12). */
ASC_API int asc_remote_create_shared(asc_context ctx, asc_server server, asc_remote* out);
/* Serves on 127.0.0.1:port from a thread of its own (port 0: any free
 * port; the reference's client uses 30007). */
ASC_API int asc_remote_listen(asc_remote remote, int port);
ASC_API int asc_remote_port(asc_remote remote, int* port);
/* Answers one message without a socket: reply_length gets the answer's
 * size (the first `capacity` bytes are copied); returns 1 for message 13
 * (close the connection). */
ASC_API int asc_remote_handle(asc_remote remote, const void* message, size_t length, void* reply, size_t capacity,
                              size_t* reply_length);
/* Directory for message 12 (saveFullDRR: Tracker::getFullDRR of the
 * current volume, one image_camNN.pgm per camera, NN counting up over the
 * session; default "."). The reference writes to a fixed path. */
ASC_API int asc_remote_set_output_dir(asc_remote remote, const char* dir);
ASC_API int asc_remote_destroy(asc_remote remote);

#ifdef __cplusplus
}
#endif
#endif /* AUTOSCOPER_HIP_H */
