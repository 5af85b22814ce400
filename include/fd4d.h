/*
 * fd4d.h — C ABI of the MI355X soft-shadow visibility library (libfd4d.so).
 *
 * Drop-in for the FDZB_Soft_Shadow call surface of galaxy-wolf/4DZB
 * (namespace FD, 4DZB/FDZB_Soft_Shadow/api.h:10-56).  Each entry point below
 * names the reference function it replaces.  Differences that follow from
 * the boundary, not from the algorithm:
 *   - OpenGL handles become device pointers (the caller owns every buffer);
 *   - the reference's process globals become an explicit context;
 *   - errors are returned as status codes instead of terminating the process;
 *   - outputs are dense (y * width + x); the reference writes the shadow value
 *     at y << ceil(log2 width) | x, which only differs for non-power-of-two
 *     widths;
 *   - besides the reference's per-receiver shadow value (occluded fraction,
 *     1 - visibility) the library also writes the per-receiver light-sample
 *     bitmask (bit set = sample occluded).
 *
 * Mask layout: receiver i owns fd_mask_words(res) consecutive uint32 words:
 * res rows of light samples, row r = r steps along the light's x_delt
 * (rightDir * -2 * rightRadius / res), bit c = c steps along y_delt.  Grids
 * up to 32 samples wide use one word per row; wider grids (up to 64) use two
 * (low word first).  Back-facing receivers have every bit set and shadow 1.
 *
 * Threading: one context per host thread; calls are stream-ordered on the
 * stream passed to the launch functions.
 */
#ifndef FD4D_H
#define FD4D_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fd_ctx fd_ctx;

/* Status codes (return values). */
#define FD_OK 0
#define FD_ERR_ARG (-1)               /* invalid argument */
#define FD_ERR_STATE (-2)             /* required set-up call missing */
#define FD_ERR_HIP (-3)               /* HIP runtime error */
#define FD_ERR_CAPACITY (-4)          /* pair list larger than the pair arena */
#define FD_ERR_UNSUPPORTED_LIGHT (-5) /* light grid not square or outside 1..64 */
#define FD_ERR_NO_DEVICE (-6)         /* no GPU / kernels not loadable */
#define FD_ERR_SCRATCH (-7)           /* workspace larger than FD_NOALLOC_SCRATCH_BYTES */

/* Frame exit codes (fd_stats.exit_code), in the order of FD::FDLaunch's
 * early returns (4DZB/FDZB_Soft_Shadow/api.cpp:495-559). */
#define FD_EXIT_OK 0              /* occlusion evaluated */
#define FD_EXIT_SAMPLE_MAXZ 1     /* no receiver in front of the light */
#define FD_EXIT_TRI_MAXZ 2        /* no vertex in front of the light */
#define FD_EXIT_NO_VALID_SAMPLE 3 /* every receiver back-facing or empty */
#define FD_EXIT_EMPTY_AABB 4
#define FD_EXIT_EMPTY_LP 5        /* receiver and occluder projections disjoint */
#define FD_EXIT_NO_VALID_TRI 6
#define FD_EXIT_NO_PAIRS 7
#define FD_EXIT_CAPACITY 9

/* AreaLightDes, 4DZB/FDZB_Soft_Shadow/AreaLightDesc.h:3-13. */
typedef struct fd_area_light {
    float right_radius; /* half width */
    float top_radius;   /* half height */
    float position[3];
    float view_dir[3];
    float up_dir[3];
    int res_w;          /* light samples per row; res_w == res_h, 1..64 */
    int res_h;
} fd_area_light;

/* Per-frame record (the reference only returns the elapsed ms). */
typedef struct fd_stats {
    int exit_code;           /* FD_EXIT_* */
    int valid_receivers;     /* receivers bound to a light-plane bin */
    int front_receivers;     /* non-empty, front-facing receivers */
    int valid_triangles;     /* triangles reaching the light plane */
    unsigned long long pairs;/* (bin, triangle) pairs */
    int key_out_of_range;    /* bound receivers whose bin key left the grid */
    int work_items;          /* occlusion work items (64 receivers each) */
    float elapsed_ms;        /* device time of the frame (fd_launch only) */
    float light_plane_begin[2];
    float light_plane_end[2];
    float far_plane;         /* f = max(receiver z, vertex z) + 0.1 */
} fd_stats;

/* FD::FDinit, FD/api.cpp:48-144.  Creates a context on the current device. */
int fd_init(fd_ctx** out);
/* No counterpart (the reference keeps process globals): frees the context. */
int fd_destroy(fd_ctx* ctx);

/* FD::FDsetBackfaceCull, FD/api.cpp:146-150 (default on, FourD.cu:61). */
int fd_set_backface_cull(fd_ctx* ctx, int enable);

/* FD::FDaddModel, FD/api.cpp:152-160.  Vertex positions are the first three
 * floats of each vertex; vertex_size_bytes is the stride (multiple of 4,
 * >= 12).  Indices are int3 per triangle.  Device pointers, caller-owned.
 * As in the reference only the first model added is used
 * (FD/api.cpp:430,441-445); fd_clear_models drops them all. */
int fd_add_model(fd_ctx* ctx, const float* d_vertices, size_t vertex_size_bytes, int num_vertices,
                 const int* d_indices, int num_triangles);
int fd_clear_models(fd_ctx* ctx);

/* FD::FDsetSampleGLTex, FD/api.cpp:162-166.  Receiver G-buffer: float4
 * positions and normals, width*height each, row-major; pos.x NaN = empty. */
int fd_set_sample_buffers(fd_ctx* ctx, const float* d_position4, const float* d_normal4);

/* FD::FDsetOutputGLBuffer, FD/api.cpp:168-171.  d_shadow: width*height
 * floats.  d_masks: width*height*fd_mask_words(res) words, or NULL. */
int fd_set_output_buffers(fd_ctx* ctx, float* d_shadow, uint32_t* d_masks);

/* FD::FDregisterGLres / FD::FDunregisterGLres, FD/api.cpp:173-192.  There is
 * no interop to register: these validate the bound buffers / drop them. */
int fd_register_res(fd_ctx* ctx);
int fd_unregister_res(fd_ctx* ctx);

/* FD::FDsetBinNum, declared at FD/api.h:36 (body commented out,
 * FD/api.cpp:193-202): bins per light-plane side, a power of two in 8..64
 * (default 16 = FD_MAX_LP_GRID_SIZE, constants.h:13,45). */
int fd_set_bin_num(fd_ctx* ctx, size_t bins_per_side);

/* No counterpart (a layout knob of this implementation): receivers are
 * walked in tiles of tile_width x (64 / tile_width) pixels, tile_width in
 * {8, 16, 32, 64} (default 8).  Changes speed only, never a result. */
int fd_set_tile_width(fd_ctx* ctx, int tile_width);

/* Occlusion kernel of fd_launch / fd_launch_async.  FD_OCC_PER_SAMPLE is
 * shadowCal_per_sample_kernel (FDc/Step5_2_shadowCalPerSample.inl:2-259),
 * the kernel FD::FDLaunch runs and the bit-exact contract of this library.
 * FD_OCC_BIN_QUEUE is the reference's unused bin-queue kernel
 * shadowCal_CTA_kernel (FDc/Step5_2_shadowCal.inl:408-1100, with
 * prepareBin, FDc/Step5_1_prepareBin.inl:75-139): other arithmetic (a
 * fourth, triangle plane; other boundary tests) and mask rows along y_delt
 * (bit j of row i is the sample upRight + j x_delt + i y_delt).  The batched
 * and serving forms always use FD_OCC_PER_SAMPLE. */
#define FD_OCC_PER_SAMPLE 0
#define FD_OCC_BIN_QUEUE 1
int fd_set_occlusion_variant(fd_ctx* ctx, int variant);

/* No counterpart (a schedule knob of this implementation): the order in
 * which the per-sample occlusion walks each bin's triangle list.  Masks and
 * shadow values do not depend on it (each receiver's mask is the OR of its
 * triangles' masks; rows already full are skipped, which changes nothing);
 * the order only decides how soon receivers fill up and stop.
 *   FD_TRI_ORDER_LIST: the reference's order (ascending triangle id, the
 *     stable sort of FDc/Step4_RasterizeTriangleToBin.inl:223);
 *   FD_TRI_ORDER_AREA: by the area of the triangle projected on the light
 *     plane in 16 powers of two, largest first, ascending id within a
 *     class;
 *   FD_TRI_ORDER_MESH: the area order, except that triangles whose
 *     projection covers 1/512 of the light plane or more come last (in
 *     list order): on meshes those are mostly triangles close to the light,
 *     which a receiver seldom needs, while putting them first scatters the
 *     triangles of each object;
 *   FD_TRI_ORDER_AUTO (default): decided per frame on the device -- when
 *     consecutive triangles mostly overlap on the light plane (a connected
 *     mesh) the mesh order, or the list order if the frame has more than
 *     7168 x 256 pairs (dense scenes: lists longer than 7168 entries in
 *     16 x 16 bins); else the area order (a triangle soup).
 * The batched form takes it per instance (fd_instance.tri_order, else the
 * context's setting), the serving form per class (fd_tune_entry.tri_order,
 * else automatic); fd_debug_read("pair_tri") always returns the reference's
 * order. */
#define FD_TRI_ORDER_AUTO 0
#define FD_TRI_ORDER_LIST 1
#define FD_TRI_ORDER_AREA 2
#define FD_TRI_ORDER_MESH 3
int fd_set_triangle_order(fd_ctx* ctx, int order);

/* No counterpart (a schedule knob of this implementation): waves per work
 * item of the per-sample occlusion (64 receivers of one bin, whose triangle
 * list the waves split): 2, 4 or 8; 0 (the default) selects 8 for light
 * grids above 32 x 32 and for frames of up to 8192 work items (about
 * 710 x 710 receivers at 16 x 16 bins), 4 otherwise.  Results do not depend on it.  The
 * batched form takes it per instance (fd_instance.occ_waves), the serving
 * form per class (fd_tune_entry.occ_waves). */
int fd_set_occlusion_waves(fd_ctx* ctx, int waves);

/* No counterpart (a schedule knob of this implementation): stretches per
 * bin triangle list.  The per-sample occlusion gives each 64 receivers of a
 * bin one work item that walks the bin's whole list; with S stretches it
 * gives them S work items, one per stretch of the list, whose masks merge by
 * atomic OR into the (zeroed) mask output before a pass counts them.  OR
 * does not depend on the order, so the bits are the same for every S.
 * 0 (the default): a frame of up to 2048 work items with fewer than 6144
 * waves in all (work items x waves per item) splits into about 49152
 * waves (at most 16 stretches); larger ones do not (nor does a batch whose
 * instances total more than 2048 work items); 1..64: that many
 * stretches.  The
 * batched form takes it per instance (fd_instance.list_split), the serving
 * form per class (fd_tune_entry.list_split). */
#define FD_LIST_SPLIT_MAX 64
int fd_set_list_split(fd_ctx* ctx, int segs);

/* FD::FDsetView, FD/api.cpp:205-237. */
int fd_set_view(fd_ctx* ctx, int width, int height);

/* FD::FDsetScenesEsp, FD/api.cpp:238-241 (default 1e-6, FourD.cu:62). */
int fd_set_scenes_esp(fd_ctx* ctx, float eps);

/* FD::FDsetSceneParams, FD/api.cpp:243-294: capacity for triangles/vertices. */
int fd_set_scene_params(fd_ctx* ctx, size_t max_triangles, size_t max_vertices);

/* FD::FDsetAreaLight, FD/api.cpp:303-400.  Unlike the reference, launches
 * stay correct without calling this again before every frame.  A grid that is
 * not square or outside 1..64 returns FD_ERR_UNSUPPORTED_LIGHT and still
 * replaces the previous light: launches return FD_ERR_UNSUPPORTED_LIGHT until
 * a supported light is set. */
int fd_set_area_light(fd_ctx* ctx, const fd_area_light* light);

/* FD::FDLaunch, FD/api.cpp:402-582.  Runs one frame on `hip_stream` (NULL =
 * default stream), waits for it and fills *stats (may be NULL).  Returns
 * FD_OK for every exit code except FD_EXIT_CAPACITY (FD_ERR_CAPACITY). */
int fd_launch(fd_ctx* ctx, void* hip_stream, fd_stats* stats);

/* No counterpart: enqueue one frame without any host synchronisation (the
 * sequence is capturable into a HIP graph); fd_get_stats reads the result. */
int fd_launch_async(fd_ctx* ctx, void* hip_stream);
int fd_get_stats(fd_ctx* ctx, void* hip_stream, fd_stats* stats);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops or harnesses that
 * forbid hipMalloc).  fd_launch_async with the frame workspace in a static
 * device array of FD_NOALLOC_SCRATCH_BYTES (128 MiB) that each GPU holds until
 * the process ends: calls no hipMalloc / hipFree / allocating routine.  The
 * pair arena takes what the rest of the workspace leaves, up to the context's
 * pair capacity (more pairs: the frame exits with FD_EXIT_CAPACITY).  A frame
 * whose workspace does not fit returns FD_ERR_SCRATCH and launches nothing
 * (1M receivers with 200k triangles and caller-provided masks fit; masks left
 * NULL are placed in the scratch as well).  Needs fd_init on the current
 * device; profiling (fd_set_profiling) does not apply.  Every call on one GPU
 * shares that GPU's array: fd_get_stats / fd_debug_read read the last call's
 * state only until the next call on that GPU, and calls that overlap on one
 * GPU (other streams, threads or contexts) race. */
#define FD_NOALLOC_SCRATCH_BYTES (128u << 20)
int fd_launch_async_noalloc(fd_ctx* ctx, void* hip_stream);

/* No counterpart (the reference's pair buffer is fixed at
 * FD_MAX_BT_PIAR_Num * 1.4 entries, FD/api.cpp:132-138): pair arena size. */
int fd_set_pair_capacity(fd_ctx* ctx, size_t max_pairs);

/* FD::moveModel, FD/api.cpp:17-20 / cudaModeTransform, FDc/moveModel.inl:16-51:
 * in-place 4x4 (column-major) transform of model 0's vertex positions. */
int fd_move_model(fd_ctx* ctx, const float m[16], void* hip_stream);

/* ---- Batched form (no counterpart: the reference renders one view with one
 * light per FDLaunch).  Many frames over the context's scene -- each with its
 * own receivers, light and outputs -- run as ONE launch per stage (instance
 * table in device memory), optionally captured into a HIP graph.  Results
 * are bit-identical to running each instance alone through fd_launch. */
typedef struct fd_instance {
    const float* d_position4;   /* receiver G-buffer (as fd_set_sample_buffers) */
    const float* d_normal4;
    int width, height;
    fd_area_light light;
    float* d_shadow;            /* width*height floats */
    uint32_t* d_masks;          /* width*height*fd_mask_words(res) words, or NULL */
    int bins_per_side;          /* 0 = the context's setting */
    int tile_width;             /* 0 = the context's setting (fd_set_tile_width) */
    unsigned long long max_pairs; /* pair arena of this instance; 0 = 16M */
    int occ_waves;              /* 0 = the context's setting (fd_set_occlusion_waves) */
    int tri_order;              /* 0 = the context's setting (fd_set_triangle_order), else
                                   FD_TRI_ORDER_LIST, _AREA or _MESH */
    int list_split;             /* 0 = the context's setting (fd_set_list_split) */
    int pad;
} fd_instance;
typedef struct fd_batch fd_batch;
int fd_batch_create(fd_ctx* ctx, const fd_instance* instances, int n, fd_batch** out);
/* Build a HIP graph of the batch on `hip_stream` (a created stream); later
 * fd_batch_launch calls replay it. */
int fd_batch_capture(fd_batch* batch, void* hip_stream);
/* Enqueue (or replay) the batch; no host synchronisation. */
int fd_batch_launch(fd_batch* batch, void* hip_stream);
int fd_batch_get_stats(fd_batch* batch, int i, fd_stats* stats);
int fd_batch_destroy(fd_batch* batch);

/* ---- Serving form (no counterpart).  Resident scenes; requests of uneven
 * size arrive at any time; a worker thread admits them in arrival order as
 * batch slots (streams) and HBM (the workspace pool) free up, runs each
 * admitted group as one batched launch, and overlaps batches across streams.
 * Results are bit-identical to fd_launch of the request alone.  Device
 * buffers named in a request must stay valid until its completion is
 * polled. */
typedef struct fd_server fd_server;
/* One tuned layout: requests of this scene (-1 = any), receiver size and
 * light grid run with these bins and tile width (see fd_autotune). */
typedef struct fd_tune_entry {
    int scene;
    int width, height, light_res;
    int bins_per_side;
    int tile_width;
    float ms;                 /* measured frame time; < 0: not bit-exact */
    int occ_waves;            /* waves per occlusion work item: 2, 4 or 8 (0 = default) */
    int tri_order;            /* walk order: FD_TRI_ORDER_LIST / _AREA / _MESH (0 = automatic) */
    int list_split;           /* stretches per bin list (0 = automatic; fd_set_list_split) */
} fd_tune_entry;
typedef struct fd_server_config {
    size_t hbm_budget_bytes;  /* workspace pool (>= 64 MiB) */
    int streams;              /* batches in flight (default 4) */
    int max_batch;            /* requests per batch (default 16) */
    int bins_per_side;        /* default 16 */
    int backface_cull;
    float eps;
    int tile_width;           /* default 8 */
    unsigned long long max_pairs; /* per request; 0 = 16M */
    const fd_tune_entry* tune;  /* optional per-class layouts (copied at creation) */
    int ntune;
    int pad;
} fd_server_config;
typedef struct fd_request {
    int scene;                /* id from fd_server_add_scene */
    int width, height;
    int pad;
    const float* d_position4;
    const float* d_normal4;
    fd_area_light light;
    float* d_shadow;
    uint32_t* d_masks;        /* or NULL */
    unsigned long long user_tag;
} fd_request;
typedef struct fd_completion {
    long long id;
    unsigned long long user_tag;
    int status;               /* FD_OK or FD_ERR_CAPACITY */
    int exit_code;            /* FD_EXIT_* */
    int batch_size;           /* requests in the batch that ran it */
    int pad;
    unsigned long long pairs;
    double t_submit_ms, t_start_ms, t_done_ms;  /* fd_server_now_ms clock */
} fd_completion;
int fd_server_create(const fd_server_config* cfg, fd_server** out);
/* returns the scene id (>= 0) */
int fd_server_add_scene(fd_server* srv, const float* d_vertices, size_t vertex_size_bytes, int num_vertices,
                        const int* d_indices, int num_triangles);
/* returns the request id (>= 0); FD_ERR_CAPACITY if it can never fit */
long long fd_server_submit(fd_server* srv, const fd_request* req);
/* FD::moveModel (FDZB_Soft_Shadow/api.cpp moveModel, FDc/moveModel.inl:6-30)
 * for a resident scene, ordered with the requests: requests submitted before
 * the move see the old vertices, requests submitted after it the transformed
 * ones (m: column-major 4x4, applied in place to the scene's vertex buffer).
 * The move completes like a request (batch_size 0); returns its id. */
long long fd_server_move_scene(fd_server* srv, int scene, const float m[16], unsigned long long user_tag);
/* non-blocking; copies up to max completions, returns how many */
int fd_server_poll(fd_server* srv, fd_completion* out, int max);
/* waits until every submitted request completed (FD_ERR_STATE on timeout) */
int fd_server_drain(fd_server* srv, int timeout_ms);
double fd_server_now_ms(fd_server* srv);
int fd_server_destroy(fd_server* srv);

/* No counterpart: the autotuner.  Runs `inst` over the context's scene once
 * per candidate layout (cands[i].bins_per_side, cands[i].tile_width,
 * cands[i].occ_waves, cands[i].tri_order, cands[i].list_split), timing
 * `reps` launches on the device, and compares every output bit with the
 * default layout (16x16 bins, 8-wide tiles).  cands[i].ms receives the median
 * frame time, -1 when the candidate's shadow values or masks differ (the bin
 * grid is observable at exact light-plane-edge ties), or -2 when its pairs do
 * not fit the instance's pair arena (max_pairs).  *best receives the fastest
 * bit-exact candidate.  The instance's own output buffers are not
 * written.  Synchronises the device. */
int fd_autotune(fd_ctx* ctx, const fd_instance* inst, fd_tune_entry* cands, int n, int reps, int* best);

/* No counterpart (the reference times the whole FDLaunch with cudaEvents,
 * FD/api.cpp:405-409,574-581): per-stage device times of the next frames
 * (light plane, receiver binning, triangle pairing, tables, occlusion) and,
 * with count_ops, FD_PROFILE_COUNTERS occlusion work counters (a slower
 * counting variant of the occlusion kernel runs):
 *   [0] candidate (receiver, triangle) tests   [1] cull passes
 *   [2] rows visited by the kernel             [3] work items (64 receivers)
 *   [4] wave triangle iterations               [5] wave row iterations
 *   [6] rows the reference visits (row-range mask minus full rows)
 *   [7] pairs whose planes are set up (cull passes not on the triangle). */
#define FD_PROFILE_COUNTERS 8
int fd_set_profiling(fd_ctx* ctx, int stage_events, int count_ops);
int fd_get_profile(fd_ctx* ctx, float* stage_ms, int nstages, unsigned long long* ops);

/* No counterpart (the reference receives GL objects): device memory and
 * streams for callers without their own HIP allocator.  Copies are
 * synchronous except fd_memcpy_dtod. */
int fd_device_count(int* n);
int fd_dev_alloc(size_t bytes, void** out);
int fd_dev_free(void* p);
int fd_memcpy_htod(void* dst, const void* src, size_t bytes);
int fd_memcpy_dtoh(void* dst, const void* src, size_t bytes);
int fd_memcpy_dtod(void* dst, const void* src, size_t bytes, void* hip_stream);
int fd_stream_create(void** out);
int fd_stream_destroy(void* hip_stream);
int fd_stream_sync(void* hip_stream);
int fd_device_sync(void);

/* ---- Application stages on either side of the visibility path --------- */

/* One object's material, as the reference's G-buffer pass packs it
 * (Resources/shader/gatherData.shf:25-44): ka, kd, ks (rgb), shininess. */
typedef struct fd_material {
    float ka[3], kd[3], ks[3], shininess;
} fd_material;

/* A ground plane z = 0 (|x|, |y| <= ground_half) with nbox axis-aligned boxes
 * standing on it, seen by a pinhole camera (eye, at, up; vertical field of
 * view in degrees).  materials: nbox + 1 entries, [0] the ground's, [1 + b]
 * box b's (may be NULL when g2 / g3 are not requested). */
typedef struct fd_gbuffer_scene {
    int width, height;
    int nbox, pad;
    const double* boxes;          /* host: nbox x {min x, y, z, max x, y, z} */
    const fd_material* materials; /* host */
    double eye[3], at[3], up[3];
    double fovy_deg, ground_half;
} fd_gbuffer_scene;

/* Replaces Render::pass0 + gatherData.shf (4DZB/Render/Render.cpp:390-431,
 * 4DZB/Resources/shader/gatherData.shf:22-46): ray-casts the scene into the
 * four RGBA32F G-buffer targets (float4 x width*height, row-major y*W + x):
 * g0 = (position, ks.r), g1 = (normal, ks.g), g2 = (ka, ks.b),
 * g3 = (kd, shininess).  Background texels are NaN in every channel (the
 * reference clears to NaN, Render.cpp:423-426).  g0 / g1 are the visibility
 * path's receiver buffers (fd_set_sample_buffers); g2 / g3 may be NULL
 * (then, without materials, g0.w = g1.w = 1).  Double-precision ray cast,
 * stream-ordered on hip_stream (no host synchronisation). */
int fd_gbuffer_boxes(const fd_gbuffer_scene* scene, float* d_g0, float* d_g1, float* d_g2, float* d_g3,
                     void* hip_stream);

/* Light terms of the shading pass (phoneShading.shf:12-22). */
typedef struct fd_shade_params {
    int width, height;
    int nsamples, pad;      /* light sample positions (the reference binds 1, Render.cpp:550) */
    const float* samples;   /* host: nsamples x 3 world positions, in accumulation order */
    float camera[3];
    float La[3], Ld[3], Ls[3];
} fd_shade_params;

/* Replaces Render::pass2 + phoneShading.shf (4DZB/Render/Render.cpp:510-565,
 * 4DZB/Resources/shader/phoneShading.shf:23-61): per texel,
 * intensity = 1 - shadow (the visibility output); a texel whose intensity is
 * NaN keeps the cleared colour (0, 0, 0, 0); otherwise the Phong terms of
 * each light sample, ambient + (diffuse + specular) * intensity, summed and
 * divided by the sample count.  Float arithmetic with IEEE division and
 * square root (normalize(v) = v / sqrt(dot(v, v)), pow = powf).  Writes
 * float4 x width*height; stream-ordered. */
int fd_shade(const fd_shade_params* params, const float* d_shadow, const float* d_g0, const float* d_g1,
             const float* d_g2, const float* d_g3, float* d_rgba, void* hip_stream);

/* Words per receiver in the mask output for a res x res light grid. */
int fd_mask_words(int light_res);

/* Stage outputs of the last frame, for stage-by-stage verification:
 * "rect" (8 floats per receiver), "key", "sorted_id", "bin_start", "bin_end",
 * "bin_maxz", "bin_minrange", "tri_aabb", "tri_valid", "tri_data",
 * "tri_count", "tri_range", "pair_tri", "bin_tri_start", "bin_tri_end",
 * "light" (light block floats), "state" (raw state).  "sorted_id" lists each
 * bin's receivers by ascending id (the reference's stable order; the kernels
 * themselves walk a bin in 8x8-pixel tile order).  No reference counterpart:
 * "tri_class" (uint8 area class per triangle) and "pair_tri_walk" (the bins'
 * lists in the order the occlusion walked them, see fd_set_triangle_order).
 * Synchronises the device.
 * *bytes receives the size; copies min(size, max_bytes). */
int fd_debug_read(fd_ctx* ctx, const char* name, void* host_dst, size_t max_bytes, size_t* bytes);

/* Verification only: evaluates the light-row mask primitive of the
 * occlusion kernel (Step5_2_shadowCalPerSample.inl:14-33, clipped to a
 * res-sample row) on n host-given (D, B) pairs, one 64-bit mask each.
 * mode 0: the kernel's general form; mode 1: its fast form wherever that
 * form's precondition (|B| <= 2^100) holds, the general form elsewhere. */
int fd_selftest_rowmask(const float* D, const float* B, int n, int res, int mode, unsigned long long* out);

/* Verification only: the occlusion kernel's row window (rows outside it are
 * skipped because no light sample of them can be set; the reference visits
 * every row of its row-range mask, Step5_2_shadowCalPerSample.inl:191-222).
 * For n edge-plane triples (coef: n x {A0, B0, C0, A1, B1, C1, A2, B2, C2}),
 * plane k at row r, column c being A_k r + B_k c + C_k, over a res x res grid,
 * out receives 4 row masks per triple: the rows whose three-edge mask is
 * non-empty by the general form, the same by the fast form, the rows of the
 * window (every row where the kernel would not apply it) and the row-range
 * mask (the rows the reference visits). */
int fd_selftest_rowwindow(const float* coef, int n, int res, unsigned long long* out);

const char* fd_status_string(int code);

#ifdef __cplusplus
}
#endif

#endif /* FD4D_H */
