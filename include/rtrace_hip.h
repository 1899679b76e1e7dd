/*
 * rtrace_hip.h — C ABI of the MI355X (gfx950) per-pixel trace path.
 *
 * Drop-in boundary for rtrace's `Renderer::render(&self, &Scene) -> Result<RgbImage, _>`
 * (src/renderer.rs:120-382) and `render_to_file` (src/renderer.rs:914-923). Plain C types
 * only; no exception crosses this boundary; every call returns RT_OK (0) or a negative
 * RT_E_* code and leaves a message in rt_last_error().
 *
 * Threading: a context is used by one host thread at a time; calls block until results
 * are in caller memory (the reference's render call also blocks). Output is independent
 * of launch shape, as the reference's is independent of thread count
 * (src/renderer.rs:1119-1171).
 */
#ifndef RTRACE_HIP_H
#define RTRACE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (mirroring the reference's error cases) ---- */
enum {
    RT_OK = 0,
    RT_E_INVALID_ARG = -1, /* e.g. "Samples must be greater than 0" (src/renderer.rs:122-124) */
    RT_E_SCENE = -2,       /* scene JSON / serde errors (src/scene.rs:462-480) */
    RT_E_IO = -3,          /* file and STL errors (src/mesh.rs:751-1003) */
    RT_E_CAMERA = -4,      /* bad FOV / camera kind / hex colour (src/camera.rs:62,115-117; src/scene.rs:15-30) */
    RT_E_DEVICE = -5,      /* no usable gfx950 device or a HIP runtime failure */
    RT_E_OOM = -6,         /* device allocation failed */
    RT_E_INTERNAL = -99
};

/* Renderer.anti_aliasing_mode (src/renderer.rs:17-24) */
enum { RT_AA_NO_JITTER = 0, RT_AA_QUINCUNX = 1, RT_AA_STOCHASTIC = 2 };

/* Object kinds (src/scene.rs:314-344) */
enum { RT_OBJ_SPHERE = 0, RT_OBJ_PLANE = 1, RT_OBJ_CUBE = 2, RT_OBJ_MESH = 3 };

typedef struct rt_ctx rt_ctx;
typedef struct rt_scene rt_scene;

/* ---- flattened scene (what a host that already parsed the Scene hands over) ---- */

/* Camera config (src/scene.rs:34-46); the basis is built per render from the aspect
 * ratio as Camera::from_config does (src/camera.rs:23-152). */
typedef struct {
    int32_t kind; /* 0 = "ortho", 1 = "perspective" */
    double position[3], target[3], up[3];
    double width, height;
    int32_t has_fov;
    double fov;
    int32_t has_grid; /* grid_pitch, grid_color and grid_thickness all present */
    double grid_pitch, grid_thickness, grid_color[3];
} rt_camera_desc;

/* Material (src/scene.rs:67-105) with colours already parsed. */
typedef struct {
    double color[3];
    double ambient, diffuse, specular, shininess;
    int32_t has_reflectivity;
    double reflectivity;
    int32_t texture; /* 0 none, 1 grid, 2 checkerboard */
    /* grid: the line colour after the reference's "#{:02X}" round trip (src/lighting.rs:28-36) */
    double grid_line_color[3];
    double line_width, cell_size;
    int32_t material_b; /* checkerboard: index of the alternate material in the table */
} rt_material_desc;

/* One World object in scene order (object index == material index, src/renderer.rs:141),
 * transforms already baked as World assembly does (src/renderer.rs:137-310). */
typedef struct {
    int32_t kind;
    int32_t material;
    double center[3];     /* sphere, cube */
    double radius;        /* sphere */
    double point[3];      /* plane */
    double normal[3];     /* plane, unit */
    double half_size[3];  /* cube */
    double world_to_local[16]; /* cube: M^-1, column-major (identity when untransformed) */
    double local_to_world[16]; /* cube: M, column-major */
    int32_t mesh;         /* mesh: index into rt_scene_desc.meshes */
} rt_object_desc;

/* Light (src/scene.rs:348-353); is_area when `diameter` is Some.
 * Extension fields (no reference counterpart; the reference ignores the JSON keys):
 *   shape   RT_LIGHT_DISK (reference) or RT_LIGHT_SQUARE (JSON "shape": "square");
 *   size    square side length (JSON "size"; falls back to "diameter");
 *   samples shadow samples per hit (JSON "samples"; default 16, the reference's
 *           hard-coded SAMPLES at src/lighting.rs:158). Disks keep the reference's
 *           rejection sampling for any count; squares use a jittered grid. */
enum { RT_LIGHT_DISK = 0, RT_LIGHT_SQUARE = 1 };
typedef struct {
    double position[3];
    double color[3];
    double intensity;
    int32_t is_area;
    double diameter;
    int32_t shape;
    int32_t samples;
    double size;
} rt_light_desc;

/* Triangle mesh, vertices in f64 after the object's transform (src/renderer.rs:270-296).
 * The k-d tree is built on the host from these with the reference algorithm
 * (src/mesh.rs:75-197) and flattened to device arrays. */
typedef struct {
    int32_t n_tris;
    const double* vertices; /* n_tris * 9 */
} rt_mesh_desc;

typedef struct {
    rt_camera_desc camera;
    int32_t n_objects;
    const rt_object_desc* objects;
    int32_t n_materials;
    const rt_material_desc* materials;
    int32_t n_lights;
    const rt_light_desc* lights;
    int32_t n_meshes;
    const rt_mesh_desc* meshes;
    double ambient_color[3];
    double ambient_intensity;
    int32_t has_fog;
    double fog_color[3], fog_density, fog_start, fog_end;
    double background[3];
} rt_scene_desc;

/* Mirrors the Renderer fields (src/renderer.rs:37-47) plus backend knobs. */
typedef struct {
    int32_t width, height;
    int32_t max_depth;   /* default 10 */
    int32_t use_kdtree;  /* default 1; 0 = brute-force mesh intersection */
    int32_t samples;     /* default 1; 0 is rejected like the reference */
    int32_t aa_mode;     /* RT_AA_* (default quincunx) */
    int32_t has_seed;    /* Renderer.seed = Some(seed) */
    uint64_t seed;
    /* Renderer.outline_config (src/outline.rs:5-18) */
    int32_t outline_enabled;
    int32_t outline_use_8_neighbors;
    double outline_depth_weight, outline_normal_weight, outline_threshold;
    double outline_color[3];
    double outline_thickness;
    /* backend knobs */
    int32_t count_work;  /* 1 = also collect traversal counters into rt_render_stats */
    int32_t profile;     /* 1 = time each pipeline stage with hipEvents (rt_render_stats.ms_stage) */
    int32_t light_samples; /* > 0: shadow samples per hit for every area light (extension) */
} rt_render_opts;

/* Pipeline stages timed when rt_render_opts.profile is set. */
enum {
    RT_STAGE_RAYGEN = 0,   /* counters init + camera rays */
    RT_STAGE_CLOSEST = 1,  /* closest-hit traversal, all levels */
    RT_STAGE_SHADE = 2,    /* exact hit records, light samples, reflection queue */
    RT_STAGE_SHADOW = 3,   /* any-hit shadow traversal, all levels */
    RT_STAGE_FINISH = 4,   /* Phong + fog per hit */
    RT_STAGE_POST = 5,     /* path combine, AA resolve, outline, quantise */
    RT_STAGE_COUNT = 6
};

/* Work counters of the two traversal kernels (count_work). */
enum { RT_KERNEL_CLOSEST = 0, RT_KERNEL_SHADOW = 1 };

typedef struct {
    uint64_t nodes;    /* k-d nodes fetched (32 B box and links, + the 8 B split of an inner node) */
    uint64_t tris;     /* FP32 triangle tests (48 B record each) */
    uint64_t retests;  /* FP64 exact triangle re-tests (72 B of f64 vertices each) */
    uint64_t spheres;  /* sphere tests */
    uint64_t planes;   /* plane tests */
    uint64_t cubes;    /* cube tests */
    uint64_t skipped;  /* shadow queries not traced: no sample of their light faces the hit point, so
                          their answers cannot change the image */
    uint64_t bvh_nodes; /* mesh BVH nodes fetched (128 B: four slot boxes and links) */
} rt_work;

typedef struct {
    uint64_t rays_primary;  /* camera rays traced (quincunx: centres + unique corners) */
    uint64_t rays_reflect;  /* reflection rays traced */
    uint64_t rays_shadow;   /* shadow queries (World::hit calls of the shading), traced or skipped */
    rt_work work[2];        /* per traversal kernel (RT_KERNEL_*), when count_work */
    double ms_render;       /* device time of the render, hipEvents */
    double ms_upload;       /* host->device copies of the last scene upload */
    double ms_stage[RT_STAGE_COUNT]; /* per-stage device time, when profile */
    int32_t launches_stage[RT_STAGE_COUNT];
    int32_t levels;         /* trace levels launched (max_depth) */
} rt_render_stats;

/* ---- context ---- */
int rt_ctx_create(int device, rt_ctx** out);
void rt_ctx_destroy(rt_ctx* ctx);
/* Message of the last failed call on ctx (thread-local when ctx is NULL). */
const char* rt_last_error(rt_ctx* ctx);

/* ---- scenes ---- */
/* Scene::from_json_file (src/scene.rs:462-470) + World assembly; meshes resolve
 * relative to the scene file. */
int rt_scene_load_json_file(rt_ctx* ctx, const char* path, rt_scene** out);
/* Scene::from_json_str (src/scene.rs:473-480); meshes resolve relative to base_dir
 * (NULL = current directory, as the reference). */
int rt_scene_load_json_str(rt_ctx* ctx, const char* json, const char* base_dir, rt_scene** out);
/* Upload a scene a host already flattened (the Rust `Renderer` would build this from
 * its `World`). The library copies what it needs; desc may be freed afterwards. */
int rt_scene_upload(rt_ctx* ctx, const rt_scene_desc* desc, rt_scene** out);
void rt_scene_free(rt_scene* scene);
/* Camera aspect width/height and the scene_settings.outline block, for CLI glue
 * (cli/src/main.rs:79-124). outline_* are filled only when enabled. */
int rt_scene_info(rt_scene* scene, double* cam_width, double* cam_height, int32_t* outline_enabled,
                  rt_render_opts* outline_out);

/* scene.objects.len() and scene.lights.len() of the loaded Scene, as the reference CLI
 * prints them after loading (cli/src/main.rs:63-67). */
int rt_scene_counts(rt_scene* scene, int32_t* n_objects, int32_t* n_lights);

/* Defaults of Renderer::new(width, height) (src/renderer.rs:50-62). */
void rt_render_opts_default(rt_render_opts* opts, int32_t width, int32_t height);

/* ---- render ---- */
/* Renderer::render: writes width*height*3 u8 (row-major RGB) to rgb_out (host memory). */
int rt_render(rt_ctx* ctx, rt_scene* scene, const rt_render_opts* opts, uint8_t* rgb_out,
              rt_render_stats* stats);
/* Same, into a device buffer (no host copy); the call still blocks until done. */
int rt_render_device(rt_ctx* ctx, rt_scene* scene, const rt_render_opts* opts, void* d_rgb_out,
                     rt_render_stats* stats);
/* ---- auxiliary outputs (extension; no reference counterpart) ----
 * Per-pixel first-hit features for the denoiser and tile-masked re-rendering. Features are
 * taken from the pixel's centre ray (quincunx) or its sample 0; misses leave zeros:
 *   0-2 radiance RGB (the pixel's colour before quantisation), 3-5 albedo RGB (material
 *   colour after textures), 6-8 normal XYZ (facing the ray), 9 depth (distance to the
 *   camera position), 10 mean visibility of the area-light samples (of all shadow rays
 *   when no light is an area light; the samples of a light none of whose samples faces the
 *   hit point count as not visible, as they are not traced), 11 its binomial variance
 *   mean*(1-mean)/n, 12 fog factor. */
enum { RT_FEAT_COUNT = 13 };
typedef struct {
    float* features;           /* device, W*H*RT_FEAT_COUNT f32, or NULL */
    float* radiance;           /* device, W*H*3 f32 linear colour before quantisation, or NULL */
    const uint8_t* tile_mask;  /* device, ceil(W/T)*ceil(H/T) bytes, row-major; only pixels of tiles
                                  != 0 are traced and written (others keep the output's contents);
                                  NULL = every pixel. Not combinable with the outline pass. */
    int32_t mask_tile;         /* T, tile edge in pixels (> 0 with a mask) */
} rt_render_aux;
/* ---- rendering without allocation (extension) ----
 * No counterpart in the reference. For callers that cannot allocate device memory: code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc.
 * rt_render_device with the frame's buffers (ray queues, hit and shade records, light
 * points, per-level colours) in a fixed device array of RT_NOALLOC_SCRATCH_BYTES that every
 * GPU the library runs on holds until the process ends; the call allocates nothing (no
 * hipMalloc / hipFree, no graph instantiation: it launches eagerly, and opts->profile is
 * ignored). Frames whose buffers need more than the array, or max_depth above 64, are
 * refused with RT_E_INVALID_ARG and a message giving the size. Every call on one GPU shares
 * that GPU's array: calls that overlap on one GPU (two contexts, two threads) race.
 * rt_last_primary_hits after such a call reads the shared array. The image equals
 * rt_render_device's. */
#define RT_NOALLOC_SCRATCH_BYTES (128u << 20)
int rt_render_device_noalloc(rt_ctx* ctx, rt_scene* scene, const rt_render_opts* opts, void* d_rgb_out,
                             rt_render_stats* stats);

/* rt_render_device plus auxiliary outputs. */
int rt_render_device_aux(rt_ctx* ctx, rt_scene* scene, const rt_render_opts* opts, void* d_rgb_out,
                         const rt_render_aux* aux, rt_render_stats* stats);

/* Primary-ray hit records of the last render on ctx, in canonical order
 * (quincunx: [W*H centres][(W+1)*(H+1) corners]; else pixel-major, sample-minor):
 * object index (-1 miss), triangle index (-1 non-mesh), t. n = capacity. */
int rt_last_primary_hits(rt_ctx* ctx, int32_t* obj, int32_t* tri, double* t, size_t n);

/* ---- on-disk formats ---- */
/* Mesh::from_stl_file (src/mesh.rs:751-1003) on the host: ASCII or binary detection, the
 * reference's parse rules and error messages. Writes the triangle count to *n_tris and,
 * with vertices != NULL, up to capacity_tris triangles as 9 f64 each. Errors through
 * rt_last_error(NULL). */
int rt_mesh_load_stl(const char* path, double* vertices, int64_t capacity_tris, int64_t* n_tris);

/* ---- k-d tree inspection ---- */
/* No counterpart in the reference: the flattened k-d tree of KdTree::new (src/mesh.rs:75-197)
 * for a mesh of n_tris triangles (9 f64 each), built on the host (on_device = 0) or on the
 * device (on_device = 1, what rt_scene_load_* uses for meshes of 50,000 triangles or more),
 * for checking one build against the other (ctx may be NULL for a host build, which needs no
 * device; errors then go to rt_last_error(NULL)). Pass capacities of 0 to get the counts; arrays
 * are copied out up to their capacities: nodes and records are 48 B each, bounds 6 f64 per
 * node. */
typedef struct {
    void* nodes;
    int64_t nodes_cap;
    void* recs;
    int64_t recs_cap;
    double* bounds;
    int64_t bounds_cap;   /* nodes */
    int64_t n_nodes;      /* out */
    int64_t n_recs;       /* out */
    int64_t root_kd;      /* out */
    int64_t root_bf;      /* out */
    int64_t leaves;       /* out */
    int64_t max_leaf;     /* out */
} rt_kd_arrays;
int rt_kd_build(rt_ctx* ctx, const double* vertices, int64_t n_tris, int32_t on_device, rt_kd_arrays* out);

/* ---- automatic cameras ---- */
/* AutoCamera::generate_cameras (src/auto_camera.rs:9-130): left, front, top (ortho) and a
 * 50-degree perspective view of the scene's finite bounds (Scene::compute_finite_bounds,
 * src/scene.rs:538-660; planes excluded) with a 15% margin, written to out[0..3] in that
 * order. RT_E_SCENE "Scene has no finite objects to compute bounds" when nothing is finite. */
int rt_scene_auto_cameras(rt_scene* scene, rt_camera_desc out[4]);
/* The same from a Scene JSON file, on the host only (no device needed; errors through
 * rt_last_error(NULL)), as the reference's auto_camera CLI (cli/src/auto_camera_cli.rs). */
int rt_auto_cameras_file(const char* scene_path, rt_camera_desc out[4]);

/* ---- batched rendering (extension; no reference counterpart) ----
 * Many frames (scenes x camera views x resolutions) rendered by one sequence of persistent
 * kernels, captured once into a hipGraph and replayed: each stage kernel takes chunks of
 * work from all jobs through one atomic counter, so jobs of uneven cost balance on the
 * device. Each job's image is the one rt_render_device gives for the same scene, camera
 * and options. */
typedef struct {
    rt_scene* scene;
    const rt_camera_desc* camera; /* NULL = the scene's camera */
    rt_render_opts opts;          /* size, anti-aliasing, samples, depth (>= 1), seed, light_samples;
                                     no outline pass, no count_work / profile */
    void* d_rgb_out;              /* device, width*height*3 u8 */
} rt_batch_job;

typedef struct {
    int32_t tile;           /* raygen tile edge in pixels: 4, 8 or 16 */
    int32_t chunk;          /* items a workgroup (one wavefront) takes at a time, a multiple of 64 */
    int32_t blocks_per_cu;  /* persistent workgroups per compute unit */
    int32_t use_graph;      /* 1 = capture the launch sequence into a hipGraph and replay it */
    double max_gib;         /* refuse batches whose buffers need more device memory (GiB) */
    double group_gib;       /* jobs run in consecutive groups whose frame buffers fit this size (GiB),
                               all groups in the same launch, sharing one allocation; 0 = one group */
    int32_t concurrent;     /* groups in flight at once (1-8), each on its own stream branch with
                               its own frame buffers: a group's stage tails overlap another's work;
                               the device holds this many groups' buffers */
} rt_batch_opts;

typedef struct {
    double ms;              /* device time of the last run (hipEvents around the graph launch) */
    double ms_build;        /* host time of create: layout, allocation, upload, capture */
    int32_t n_jobs;
    int32_t graph_nodes;    /* kernels in the captured graph (0 without a graph) */
    int32_t groups;         /* job groups (group_gib), `concurrent` of them in flight at once */
    int64_t bytes;          /* device memory of the frame buffers (concurrent x the largest group's) */
    uint64_t pixels;
    uint64_t rays_primary, rays_reflect, rays_shadow;
} rt_batch_stats;

typedef struct rt_batch rt_batch;

void rt_batch_opts_default(rt_batch_opts* bo);
/* Lays out and allocates every job's buffers and captures the launch graph. The batch keeps
 * raw device pointers of each job's scene and d_rgb_out: both must outlive the rt_batch
 * (free the batch before rt_scene_free or before releasing an output buffer). */
int rt_batch_create(rt_ctx* ctx, const rt_batch_job* jobs, int32_t n_jobs, const rt_batch_opts* bo,
                    rt_batch** out);
/* Renders every job into its d_rgb_out; blocks until done. */
int rt_batch_run(rt_batch* b, rt_batch_stats* stats);
void rt_batch_destroy(rt_batch* b);
/* Chooses tile, chunk and blocks_per_cu for these jobs by timing candidate settings
 * (coordinate descent from *bo, which also supplies use_graph and max_gib); writes the
 * fastest to *bo and its time to *best_ms. */
int rt_batch_autotune(rt_ctx* ctx, const rt_batch_job* jobs, int32_t n_jobs, rt_batch_opts* bo, double* best_ms);

/* ---- ray queries (the `World::hit` seam, src/ray.rs:699-711) ---- */
/* Closest hit of n rays (origins/dirs: n*3 f64; dirs are normalized as Ray::new does). */
int rt_intersect(rt_ctx* ctx, rt_scene* scene, const double* origins, const double* dirs, size_t n,
                 double t_min, double t_max, int use_kdtree, int32_t* obj_out, int32_t* tri_out,
                 double* t_out);

#ifdef __cplusplus
}
#endif
#endif /* RTRACE_HIP_H */
