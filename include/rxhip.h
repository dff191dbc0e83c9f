/*
 * rxhip.h -- C ABI of librxhip.so, the MI355X (gfx950) path-tracing hot path
 * of RenderX: KD-tree ray traversal with triangle intersection and the
 * PathTracer bounce loop, as hand-written HIP kernels.
 *
 * Every entry point returns RX_OK (0) or a negative rx_status; none throws or
 * exits.  A context is bound to one device and one HIP stream and is not
 * thread-safe (it mirrors the reference's single static GL state,
 * include/IntersectionGPU.h:11-45).  All host pointers are read or written
 * synchronously unless the entry point name ends in _dev, in which case the
 * pointers are device pointers and the work is queued on the context stream.
 *
 * Reference citations are file:line in the RenderX tree.
 */
#ifndef RXHIP_H
#define RXHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum rx_status {
  RX_OK = 0,
  RX_ERR_INVALID = -1,   /* bad argument or size */
  RX_ERR_HIP = -2,       /* a HIP runtime call failed */
  RX_ERR_NO_MEMORY = -3, /* device or host allocation failed */
  RX_ERR_STATE = -4,     /* call order (e.g. render before load_scene) */
  RX_ERR_NO_DEVICE = -5, /* no usable gfx950 device */
  RX_ERR_OVERFLOW = -6   /* a device queue or traversal stack overflowed */
};

/* object types, one per SceneObject subclass on the path */
enum rx_object_type {
  RX_OBJ_DIFFUSE = 0,    /* SceneDiffuseObject (include/SceneDiffuseObject.h:14-19) */
  RX_OBJ_GLOSSY = 1,     /* SceneGlossyObject (include/SceneGlossyObject.h:12-17) */
  RX_OBJ_PHONG = 2,      /* ScenePhongObject = glossy + diffuse lobe (include/ScenePhongObject.h:15-22) */
  RX_OBJ_REFRACTIVE = 3, /* SceneRefractiveObject (include/SceneRefractiveObject.h:12-20) */
  RX_OBJ_EMISSIVE = 4,   /* SceneEmissiveObject (include/SceneEmissiveObject.h) */
  RX_OBJ_VPM = 5         /* SceneVPMObject, homogeneous medium (include/SceneVPMObject.h:8-104) */
};

/* One scene object: a contiguous range of the global triangle list plus its
 * material (the setters of include/GlossyMaterial.h:16-25,
 * include/SceneVPMObject.h:84-104, include/RefractiveMaterial.h).  64 bytes. */
typedef struct rx_object_desc {
  int32_t type;          /* rx_object_type */
  uint32_t tri_begin;    /* first global triangle of the object */
  uint32_t tri_count;
  uint32_t pad0;
  float color[3];        /* diffuse rho | glossy/phong ks | refractive surfColor | emissive Le | vpm sigma_s */
  float exponent;        /* glossy/phong Phong exponent n */
  float color2[3];       /* phong kd | refractive decayColor | vpm sigma_a */
  float ior;             /* refractive refrCoeff | vpm IOR ("inrate") */
  float g;               /* vpm Henyey-Greenstein g */
  float pad1[3];
} rx_object_desc;

/* Pinhole camera (include/Camera.h:10-31): integer focal length in pixels. */
typedef struct rx_camera_desc {
  float position[3];
  float focus[3];
  float up[3];
  uint32_t sight_dist;
  int32_t width;
  int32_t height;
} rx_camera_desc;

enum rx_render_mode {
  RX_MODE_PT = 0,      /* PathTracer: BSDF sampling only, emitter-terminated paths (src/PathTracer.cpp:17-89) */
  RX_MODE_NEE = 1,     /* next-event estimation to area lights + MIS, unbiased w.r.t. RX_MODE_PT */
  RX_MODE_GUIDED = 2,  /* RX_MODE_NEE + learned vMF-mixture proposal combined by one-sample MIS */
  RX_MODE_BPT = 3      /* BidirectionalPathTracer (src/BidirectionalPathTracer.cpp:6-398): per pixel and
                          sample one eye and one light subpath, every connection weighted by the reference's
                          balance heuristic; light-tracing contributions land on the pixel
                          Camera::transToPixel names.  Media (RX_OBJ_VPM) enter through the Ray methods
                          the reference's connection code calls.  max_depth <= 30,
                          at most 2^22 pixels */
};

typedef struct rx_render_params {
  uint32_t spp;
  uint32_t max_depth;      /* MCRenderer::maxDepth: bounces whose outgoing ray is traced */
  uint32_t rr_start_depth; /* 0 = Russian roulette at every bounce (the reference) */
  uint32_t mode;           /* rx_render_mode */
  uint64_t seed;
  uint32_t sample_offset;  /* first sample index: resumes a render with the same counters */
  uint32_t flags;          /* rx_render_flags */
} rx_render_params;

enum rx_render_flags {
  RX_RENDER_TIMING = 1,     /* time each kernel with HIP events (ms_extend/ms_shade/ms_shadow) */
  RX_RENDER_COUNT_WORK = 2, /* count box and triangle tests of the extend kernels */
  RX_RENDER_BPT_REGISTERED = 4 /* RX_MODE_BPT: light-tracing contributions land on the pixel whose camera rays
                                  cover them (floor of the raster position) instead of transToPixel's
                                  (int)(x - 0.5), which shifts them by half a pixel and doubles the first
                                  row's and column's footprint; the image then has RX_MODE_PT's expectation */
};

/* Launch shape of rx_render; no counterpart in the reference.  Every setting
 * gives bit-identical images (the film adds each pixel's samples in sample
 * order and the random numbers are keyed by pixel and sample); only the time
 * changes.  Zero fields take the defaults. */
typedef struct rx_render_tuning {
  uint32_t path_slots;    /* path states resident at once (default 2^26): a pass renders
                             min(spp, path_slots / pixels) samples of every pixel (>= 1) */
  uint32_t blocks_per_cu; /* resident 128-thread blocks per CU of the render kernels
                             (default and maximum: the traversal kernel's occupancy limit) */
  uint32_t use_graph;     /* rx_batch_render: 1 records one pass into a hipGraph and replays it for every pass;
                             2 runs the whole batch (every pass and bounce) as one persistent cooperative launch
                             with grid-wide barriers between the stages.  rx_render: ignored */
  uint32_t stack_depth;   /* LDS traversal stack entries per lane of the extend kernel: 0 (default 10), 6, 8, 12, 16
                             with block_size 0/128; deeper entries spill to a per-thread global stack */
  uint32_t block_size;    /* threads per block of the extend kernel: 0 (default 128), 64 or 256 (with the default
                             stack_depth); the grid keeps the default shape's thread count */
  uint32_t camera_rays;   /* camera rays: 1 one ray per lane like the later bounces, 2 packets of 64 (a wave walks
                             the tree once for neighbouring samples), 0 the default: packets for rx_render (a
                             wave holds the samples of two pixels), lanes for batches (a wave holds 64
                             neighbouring pixels of a scene, too far apart to share a walk).  The hits are
                             identical (the persistent batch launch traces them per lane) */
  uint32_t tail_bounce;   /* rx_render in RX_MODE_PT on a scene without media: the first bounce whose paths
                             one launch carries to their end (each lane traces, shades and traces again)
                             instead of one extend and one shade launch per bounce; 0 the default (3),
                             a bounce above max_depth none */
} rx_render_tuning;

typedef struct rx_render_stats {
  uint64_t rays;           /* all traversal queries: camera, bounce, shadow, containment */
  uint64_t extend_rays;    /* nearest-hit queries of the extend kernels */
  uint64_t shadow_rays;    /* next-event shadow queries (one per boundary crossing) */
  uint64_t segments;       /* shaded path segments */
  uint64_t samples;        /* camera paths */
  uint64_t box_tests;      /* extend traversal box tests (RX_RENDER_COUNT_WORK) */
  uint64_t tri_tests;      /* extend traversal triangle tests (RX_RENDER_COUNT_WORK) */
  double ms_total;         /* wall time of the render on the device stream */
  double ms_extend;        /* summed extend kernel time (RX_RENDER_TIMING) */
  double ms_shade;
  double ms_shadow;
  uint64_t extend_launches;
  uint64_t guide_rows;       /* guided mode: network queries evaluated (one per guided vertex) */
  uint64_t guide_train_rows; /* guided mode: training rows over all Adam steps (batch_size per step) */
  /* per-bounce breakdown, summed over the render's passes; bounce b >= RX_STATS_BOUNCES-1 adds to the
     last entry.  RX_RENDER_TIMING: time of the extend launch that starts at bounce b (k_tail's whole
     launch, traversal and shading of every later bounce, at its first bounce).  RX_RENDER_COUNT_WORK:
     the counted extend launch's nearest-hit queries, box tests and triangle tests at bounce b. */
  double ms_extend_bounce[16];
  uint64_t extend_rays_bounce[16];
  uint64_t box_tests_bounce[16];
  uint64_t tri_tests_bounce[16];
} rx_render_stats;
#define RX_STATS_BOUNCES 16

typedef struct rx_ctx rx_ctx;
typedef struct rx_tree rx_tree;

/* ------------------------------------------------------------------ host tree */
enum rx_tree_flags {
  RX_TREE_ABS_DOUBLE = 0, /* split criterion |ratio-0.5| in double (libc++/MSVC binding of abs) */
  RX_TREE_ABS_INT = 1     /* int abs(int) binding under libstdc++: x axis only (src/KDTree.cpp:191) */
};

/* KDTree::build(BEST) -> splitBest (src/KDTree.cpp:230-246,129-228).
 * verts: 3*n_verts floats (world positions, KDTree::vertexPositionList);
 * tris: 3*n_tris global vertex indices (KDTree::triangleList). */
int rx_tree_build(uint32_t n_verts, const float* verts, uint32_t n_tris, const uint32_t* tris,
                  uint32_t flags, rx_tree** out);
void rx_tree_free(rx_tree* tree);
int rx_tree_info(const rx_tree* tree, uint32_t* n_nodes, uint32_t* n_leaf_tris, uint32_t* max_depth);
/* KDTree::serializeForGPU (src/KDTree.cpp:346-383): pre-order nodes,
 * nodes4[i] = (parent, right child, leafStart, leafEnd) as floats, min4/max4
 * node boxes, v1/v2/v3 leaf triangle vertices with v1.w = triangle id.
 * Sizes: nodes4/min4/max4 4*n_nodes, v1/v2/v3 4*n_leaf_tris floats. */
int rx_tree_serialize(const rx_tree* tree, float* nodes4, float* min4, float* max4, float* v1, float* v2,
                      float* v3);

/* ------------------------------------------------------------------ context */
/* IntersectionGPU::init (src/IntersectionGPU.cpp:16-44) */
int rx_create(int device, rx_ctx** out);
void rx_destroy(rx_ctx* ctx);
const char* rx_last_error(const rx_ctx* ctx);
int rx_synchronize(rx_ctx* ctx);
/* Test hook, no counterpart in the reference: the nth device allocation the
 * library makes from now on (0 = the next one) fails as out of memory, once;
 * negative switches it off.  Process-wide.  It shows that a load interrupted
 * by a failed allocation leaves the context without a scene or tree
 * (RX_ERR_STATE from rx_render / rx_intersect) instead of half-built. */
int rx_debug_fail_device_alloc(int32_t nth);

/* IntersectionGPU::loadKDTreeToShader (src/IntersectionGPU.cpp:243-255):
 * takes the arrays KDTree::serializeForGPU produces, unchanged.  Replaces
 * any loaded tree and clears legality data (load it again afterwards). */
int rx_load_tree(rx_ctx* ctx, uint32_t n_nodes, const float* nodes4, const float* min4, const float* max4,
                 uint32_t n_leaf_tris, const float* v1, const float* v2, const float* v3);

/* Per-triangle object id and face normal (SimpleShape::getWorldNormal flat
 * branch, src/SimpleShape.cpp:393-421), indexed by triangle id; enables the
 * NoSelfIntersectionCondition legality test (src/NoSelfIntersectionCondition.cpp:12-27). */
int rx_load_legality(rx_ctx* ctx, uint32_t n_tris, const int32_t* tri_obj, const float* tri_normal);

enum rx_intersect_flags {
  RX_ISECT_ORDERED = 0,     /* nearer child first, conservative distance culling */
  RX_ISECT_EXHAUSTIVE = 1,  /* every child whose box is hit: the reference recursion's visits */
  RX_ISECT_LEGALITY = 2     /* apply NoSelfIntersectionCondition with (inside, contact) per ray */
};

/* IntersectionGPU::query (src/IntersectionGPU.cpp:216-241) with the result of
 * KDTree::intersect(ray, triangleID, condition) (src/KDTree.cpp:335-339):
 * org/dir 3n floats; inside/contact n object ids (-1 = none), used with
 * RX_ISECT_LEGALITY, else may be NULL; skip_tid (n triangle ids, -1 = none,
 * may be NULL) is IntersectionGPU::Ray::last_tid (include/IntersectionGPU.h:14-25):
 * the named triangle is passed over, as the shader does (Shader/Intersection.frag:63-64)
 * for the re-test rays of Scene::fillIntersectObject / testVisibility
 * (src/Scene.cpp:438,463,487).  Outputs: dist (-1 = miss), tid (-1 = miss),
 * optional bary (2n, FP32 of an FP64 solve: p = (1-u-v)v0+u v1+v v2),
 * optional counts (2n: box tests, triangle tests). */
int rx_intersect(rx_ctx* ctx, uint32_t n, const float* org, const float* dir, const int32_t* inside,
                 const int32_t* contact, const int32_t* skip_tid, uint32_t flags, float* dist, int32_t* tid,
                 float* bary, uint32_t* counts);
/* Same with device pointers, queued on the context stream.  A traversal
 * stack overflow of a queued call is reported by the next rx_synchronize
 * (RX_ERR_OVERFLOW). */
int rx_intersect_dev(rx_ctx* ctx, uint32_t n, const float* org, const float* dir, const int32_t* inside,
                     const int32_t* contact, const int32_t* skip_tid, uint32_t flags, float* dist, int32_t* tid,
                     float* bary, uint32_t* counts);
/* No counterpart in the reference: rx_intersect_dev for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments and results; the
 * traversal-stack spill and the work counters live in a fixed 32 MiB device
 * array of the library (one per GPU, held until the process ends) instead of
 * buffers the context allocates, and they are cleared by a kernel, not a
 * memset.  Trees whose depth needs more spill than that array holds for one
 * 128-thread block are rejected with RX_ERR_INVALID.  Calls that overlap on
 * one GPU (any context, any stream) share the array and race: order them on
 * one stream or synchronize between them.  A stack overflow is reported by
 * the next rx_synchronize; rx_last_kernel_ms does not time these calls. */
int rx_intersect_dev_noalloc(rx_ctx* ctx, uint32_t n, const float* org, const float* dir, const int32_t* inside,
                             const int32_t* contact, const int32_t* skip_tid, uint32_t flags, float* dist,
                             int32_t* tid, float* bary, uint32_t* counts);
/* Duration in ms of the last rx_intersect(_dev) traversal kernel (HIP events
 * on the context stream; valid after rx_synchronize). */
int rx_last_kernel_ms(rx_ctx* ctx, float* ms);
/* No counterpart in the reference: a plain streaming device-to-device copy
 * (each byte read once and written once; 16-byte aligned pointers and size),
 * the known-traffic kernel that calibrates hardware byte counters against
 * the traversal kernels.  Synchronous. */
int rx_copy_kernel(rx_ctx* ctx, void* dst, const void* src, uint64_t bytes);

/* ------------------------------------------------------------------ scene + render */
/* Scene::buildKDTree + Scene::buildObjKDTrees (src/Scene.cpp:248-320) with
 * the objects' materials; builds the global tree and one tree per medium,
 * computes legality data and the emitter sampler (src/Scene.cpp:19-33,193-204). */
int rx_load_scene(rx_ctx* ctx, uint32_t n_verts, const float* verts, uint32_t n_tris, const uint32_t* tris,
                  uint32_t n_objects, const rx_object_desc* objects, uint32_t tree_flags);

/* MCRenderer::renderPixels / PathTracer::renderPixels (include/MCRenderer.h:145-148,
 * src/PathTracer.cpp:6-89): renders spp samples per pixel of the camera's
 * image.  Outputs are row-major y*W+x RGB like vector<vec3f>: mean (3WH
 * floats, may be NULL), and FP64 per-pixel sum and sum of squares (3WH
 * doubles each, may be NULL) for variance estimates.  The film is
 * deterministic: identical outputs for identical inputs. */
int rx_render(rx_ctx* ctx, const rx_camera_desc* camera, const rx_render_params* params, float* rgb_mean,
              double* sum, double* sumsq, rx_render_stats* stats);

/* Sets / reads the context's rx_render_tuning (validated; zero fields = defaults). */
int rx_set_render_tuning(rx_ctx* ctx, const rx_render_tuning* tuning);
int rx_get_render_tuning(rx_ctx* ctx, rx_render_tuning* tuning);
/* Times rx_render with each candidate launch shape on `params` (spp capped
 * at params->spp) and keeps the fastest in the context: path slots x blocks
 * per CU, then the extend kernel's stack depth and block size, then camera
 * rays per lane instead of in packets.  `best` (may be NULL) receives it and
 * `best_ms` its time.  No counterpart in the reference. */
int rx_autotune_render(rx_ctx* ctx, const rx_camera_desc* camera, const rx_render_params* params,
                       rx_render_tuning* best, float* best_ms);

/* ------------------------------------------------------------------ multi-scene batch */
/* No counterpart in the reference, which renders one scene per run
 * (src/Renderer.cpp:50): many scenes resident together, rendered as one
 * wavefront job whose passes mix the scenes' paths (a pass holds at most one
 * sample of any pixel; with rx_render_tuning.use_graph one pass is recorded
 * into a hipGraph and replayed).  Each scene's film is bit-identical to
 * rx_load_scene + rx_render of that scene with the same camera and params.
 * Scenes with media (RX_OBJ_VPM) are rejected; all scenes share one mode
 * (RX_MODE_PT or RX_MODE_NEE). */
typedef struct rx_batch rx_batch;
int rx_batch_create(rx_ctx* ctx, rx_batch** out);
void rx_batch_destroy(rx_batch* batch);
/* Builds the scene's tree (as rx_load_scene) and queues it with its camera
 * and render params; *index (may be NULL) receives its position. */
int rx_batch_add_scene(rx_batch* batch, uint32_t n_verts, const float* verts, uint32_t n_tris, const uint32_t* tris,
                       uint32_t n_objects, const rx_object_desc* objects, uint32_t tree_flags,
                       const rx_camera_desc* camera, const rx_render_params* params, uint32_t* index);
/* Renders every queued scene; films stay on the device. */
int rx_batch_render(rx_batch* batch, rx_render_stats* stats);
/* Times the first `passes` passes of rx_batch_render (0: all) under each
 * candidate launch shape and keeps the fastest (time per camera path) as the
 * context's rx_render_tuning: path slots (the wavefront size, which is also
 * every queue's capacity) x blocks per CU with graph replay, then the extend
 * kernel's stack depth and block size, then camera rays per lane, then
 * stream launches and the one persistent launch.  `best` and `best_ms` (ms per camera path) may be NULL.
 * No counterpart in the reference. */
int rx_batch_autotune(rx_batch* batch, uint32_t passes, rx_render_tuning* best, float* best_ms);
/* One scene's film as rx_render returns it (any output may be NULL). */
int rx_batch_get_image(rx_batch* batch, uint32_t index, float* rgb_mean, double* sum, double* sumsq);

/* ------------------------------------------------------------------ film I/O and comparison (host only) */
/* Layout: RGB float32 row-major y*W+x, y = 0 the top row (rx_render's output). */
/* MCRenderer::saveImagePFM (src/MCRenderer.cpp:556-574): "PF\n<W> <H>\n-1.000000\n"
 * header, rows from the bottom of the image up, R G B float32 little endian. */
int rx_save_pfm(const char* path, int32_t width, int32_t height, const float* rgb);
/* The time-stamped copy of MCRenderer::showCurrentResult (src/MCRenderer.cpp:465-490):
 * writes save_path with its extension replaced by "_time_<time>_iter_<iter>.pfm". */
int rx_save_pfm_snapshot(const char* save_path, uint32_t time, uint32_t iter, int32_t width, int32_t height,
                         const float* rgb);
/* readImagePFM (include/visualizeUtils.h:23-48): rgb may be NULL to query the
 * size; clamp != 0 clamps channels to [0, 4] like the reference.  The header
 * is parsed per the PFM format (scale sign = byte order), unlike the
 * reference's %d parse of the scale. */
int rx_read_pfm(const char* path, int32_t* width, int32_t* height, float* rgb, int32_t clamp);
/* CalculateRMSE (cmp/compare.cpp:85-124): sqrt(mean over masked pixels of
 * |ref - img|^2); mask (W*H bytes, non-zero = counted) may be NULL. */
int rx_image_rmse(int32_t width, int32_t height, const float* img, const float* ref, const uint8_t* mask,
                  double* rmse);
/* No counterpart in the reference: relMSE = mean over pixels and channels of
 * (img - ref)^2 / (ref^2 + eps), the path-guiding convention (eps = 1e-2). */
int rx_image_relmse(int32_t width, int32_t height, const float* img, const float* ref, float eps, double* relmse);

/* ------------------------------------------------------------------ scene files (host only) */
/* ConfigManager::load (src/ConfigManager.cpp:166-405) for this path: the
 * Config XML, the Scene XML it names (GroupObject / Object / Material /
 * Camera nodes, findNode's Name matching, paths relative to the config's
 * folder), OBJ and SHP shapes (SimpleShape::loadShape, src/SimpleShape.cpp:44-313:
 * quads split in two, GroupObjects split by "g" groups and unitized), object
 * transforms applied like getWorldVertexPosition (include/SimpleShape.h:44).
 * Materials: Diffuse, Glossy, Phong, Refractive, Emissive, VPM; renderer
 * PathTracer/PT (RX_MODE_PT) or PTtest (RX_MODE_NEE).  Textures, other
 * materials and other renderers are rejected with a message in err. */
typedef struct rx_scene_file rx_scene_file;
int rx_scene_load(const char* config_path, rx_scene_file** out, char* err, uint32_t err_cap);
void rx_scene_file_free(rx_scene_file* file);
int rx_scene_file_sizes(const rx_scene_file* file, uint32_t* n_verts, uint32_t* n_tris, uint32_t* n_objects);
/* Copies the arrays rx_load_scene takes (3*n_verts floats, 3*n_tris indices,
 * n_objects descriptors), the camera, the render params (spp from
 * <Renderer><spp>, default 1; maxDepth default 20), savePath and the <time>
 * budget in seconds.  Any output may be NULL. */
int rx_scene_file_get(const rx_scene_file* file, float* verts, uint32_t* tris, rx_object_desc* objects,
                      rx_camera_desc* camera, rx_render_params* params, char* save_path, uint32_t save_cap,
                      double* runtime_s);

/* ------------------------------------------------------------------ path guiding */
typedef struct rx_guide_params {
  uint32_t train_passes;   /* online warm-up passes: each renders enough samples per pixel for
                              its store of radiance samples (at least one), the leading samples
                              of every pixel, and trains on them */
  uint32_t batch_size;     /* radiance samples per Adam step (>= 256) */
  uint32_t steps_per_pass; /* Adam steps after each training pass, over a deterministic
                              permutation of that pass's store; steps beyond store / batch_size
                              go over it again (epochs) */
  float learning_rate;     /* Adam */
  float beta1, beta2, eps;
  uint32_t seed;
  uint32_t flags;          /* rx_guide_flags */
  uint32_t samples_per_pass; /* radiance samples a training pass keeps (the store: an even subset
                              of its records); 0 = batch_size * steps_per_pass */
} rx_guide_params;

enum rx_guide_flags {
  RX_GUIDE_LR_DECAY = 1,   /* the Adam learning rate falls geometrically from learning_rate to a tenth of it
                              over the online warm-up's train_passes * steps_per_pass steps */
  RX_GUIDE_KEEP_SAMPLES = 2 /* each warm-up pass adds its samples_per_pass samples to those of the earlier
                              passes, and its Adam steps draw from all of them (pass k of the warm-up
                              trains on k * samples_per_pass samples) */
};

/* The guiding network has no counterpart in the reference.  It maps a
 * frequency encoding of (position normalised to the scene box: 12 sin/cos
 * octaves; view direction: 4 octaves) = 96 inputs through 3 ReLU layers of 64
 * to 48 outputs (41 used): per lobe j of 8 a weight logit, an unnormalised
 * mean direction and a log concentration, then one selection logit.
 * Parameters are FP32 masters (RX_GUIDE_PARAMS floats: W1 96x64, W2 64x64,
 * W3 64x64, W4 64x48 row-major in x out, then b1, b2, b3, b4); the kernels
 * compute in BF16 on MFMA with FP32 accumulation. */
#define RX_GUIDE_PARAMS 17648

/* Configures and (re)initialises the guiding network used by RX_MODE_GUIDED. */
int rx_guide_init(rx_ctx* ctx, const rx_guide_params* params);
/* Box the position encoding normalises to (rx_load_scene sets the scene box). */
int rx_guide_set_bounds(rx_ctx* ctx, const float* lo3, const float* hi3);
int rx_guide_set_weights(rx_ctx* ctx, const float* params);
int rx_guide_get_weights(rx_ctx* ctx, float* params);
/* Network outputs (48 floats per query) for n positions/directions (3n floats each). */
int rx_guide_eval(rx_ctx* ctx, uint32_t n, const float* pos, const float* dir, float* out48);
/* The same queries decoded as RX_MODE_GUIDED's renderer reads them, 52 floats
 * per query: 8 x (unit mean direction, concentration kappa), 8 x w_j kappa_j /
 * (2 pi (1 - e^{-2 kappa_j})), 8 cumulative lobe weights (the last is 1), the
 * selection probability alpha, 3 floats of padding. */
int rx_guide_eval_mixture(rx_ctx* ctx, uint32_t n, const float* pos, const float* dir, float* out52);
/* One Adam step on n radiance samples: sampled direction omega (3n), BSDF
 * density at omega, weight F/p_sample; loss = mean of -weight log(alpha q +
 * (1-alpha) p_bsdf).  Optional outputs: mean loss, mean gradient (RX_GUIDE_PARAMS). */
int rx_guide_train_batch(rx_ctx* ctx, uint32_t n, const float* pos, const float* dir, const float* omega,
                         const float* pdf_bsdf, const float* weight, float* loss, float* grad);

#ifdef __cplusplus
}
#endif

#endif /* RXHIP_H */
