/*
 * rt_oracle.h — CPU f64 restatement of rtrace's per-pixel trace path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the parity oracle for the HIP
 * path: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it. The product (rtrace_amd, librtrace_hip.so) never links it.
 *
 * Every function restates the reference (SeaRyanC/rtrace, Rust) with the same
 * operation order in IEEE f64 and no FMA contraction; each one cites the file
 * and lines it follows. See rt_oracle.cpp for the citations.
 */
#ifndef RT_ORACLE_H
#define RT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Mirrors the public fields of `Renderer` (src/renderer.rs:37-47). */
typedef struct {
    int32_t width;
    int32_t height;
    int32_t max_depth;       /* Renderer.max_depth (default 10) */
    int32_t use_kdtree;      /* Renderer.use_kdtree (default 1) */
    int32_t threads;         /* 0 = hardware_concurrency() */
    int32_t samples;         /* Renderer.samples (default 1) */
    int32_t aa_mode;         /* 0 = NoJitter, 1 = Quincunx, 2 = Stochastic */
    int32_t has_seed;        /* Renderer.seed is Some(seed) when 1 */
    uint64_t seed;
    /* Renderer.outline_config (src/outline.rs:5-18); enabled when outline_enabled != 0 */
    int32_t outline_enabled;
    int32_t outline_use_8_neighbors;
    double outline_depth_weight;
    double outline_normal_weight;
    double outline_threshold;
    double outline_color[3];
    double outline_thickness;
    /* extension: > 0 overrides every area light's shadow samples per hit */
    int32_t light_samples;
} oracle_opts;

typedef struct {
    uint64_t rays_primary;    /* World::hit calls made for camera rays */
    uint64_t rays_reflect;    /* World::hit calls for reflection rays */
    uint64_t rays_shadow;     /* World::hit calls for shadow rays */
    uint64_t kd_inner_visits; /* reference-order k-d inner nodes visited */
    uint64_t kd_leaf_visits;  /* leaves whose bounds test ran */
    uint64_t tri_tests;       /* Moller-Trumbore calls */
    uint64_t sphere_tests;
    double render_seconds;    /* wall clock of the render (scene parse and k-d build excluded) */
    int32_t threads_used;
} oracle_stats;

/* Render `scene_path` (a Scene JSON file, meshes resolved relative to it).
 * rgb_out: width*height*3 bytes. hit_obj/hit_tri/hit_t (nullable): one entry per
 * primary ray in canonical order — quincunx: [W*H centres][(W+1)*(H+1) corners];
 * otherwise [pixel-major, sample-minor] W*H*samples. Returns 0 or a negative code
 * with a message in err. */
int oracle_render(const char* scene_path, const oracle_opts* opts, uint8_t* rgb_out,
                  int32_t* hit_obj, int32_t* hit_tri, double* hit_t,
                  oracle_stats* stats, char* err, int errlen);

/* Render plus the extension's per-pixel outputs (the device's rt_render_aux):
 * features W*H*13 f32 (radiance RGB, albedo RGB, normal XYZ, depth, area-light visibility
 * mean and binomial variance, fog factor; first hit of the centre ray or sample 0; zeros
 * on a miss) and radiance W*H*3 f32 (colour before quantisation). Either may be NULL. */
int oracle_render_aux(const char* scene_path, const oracle_opts* opts, uint8_t* rgb_out, float* features,
                      float* radiance, oracle_stats* stats, char* err, int errlen);

/* AutoCamera::generate_cameras (src/auto_camera.rs:9-130) of a scene file: 4 cameras
 * (left, front, top, perspective) x 14 doubles: kind (0 ortho, 1 perspective), position,
 * target, up, width, height, fov (0 for ortho), 0. */
int oracle_auto_cameras(const char* scene_path, double* out, char* err, int errlen);
/* Mesh::from_stl_file (src/mesh.rs:751-1003): triangle count, and with out != NULL the
 * f64 vertices (9 per triangle, up to capacity_tris). */
int oracle_load_stl(const char* path, double* out, int64_t capacity_tris, int64_t* n_tris, char* err, int errlen);

/* Unit-level hooks used by the tests (each restates one reference function). */
void oracle_chacha_block(const uint32_t in[16], int double_rounds, uint32_t out[16]);
void oracle_seed_key(uint64_t seed, uint32_t key_out[8]);
/* First n u64 draws of StdRng::seed_from_u64(seed) (next_u64). */
void oracle_stdrng_u64(uint64_t seed, int n, uint64_t* out);
/* n draws of rng.gen::<f64>() */
void oracle_stdrng_f64(uint64_t seed, int n, double* out);
/* rand 0.8's StdRng construction sequence: StdRng::from_seed(seed) -> next_u64 -> from_rng(that rng)
 * (fill_bytes of a 32-byte seed) -> next_u64; writes the two draws. */
void oracle_stdrng_construction(const uint8_t seed[32], uint64_t out[2]);
/* n draws from rand_pcg's Pcg32::new(state, stream) through the oracle's f64 recipes:
 * kind 0 = gen::<f64>() (Standard), kind 1 = gen_range(low..high) (UniformFloat::sample_single). */
void oracle_pcg32_f64(uint64_t state, uint64_t stream, int kind, double low, double high, int n, double* out);
/* n calls of sample_disk_point(rng, radius); writes 2n doubles */
void oracle_sample_disk(uint64_t seed, double radius, int n, double* out_xy);
/* n shadow-sample points of an area light for a hit at p with path seed `path_seed`
 * (calculate_diffuse_light_contribution seeding, src/lighting.rs:161-166): the reference's
 * disk (square = 0, extent = diameter) or the square extension (square = 1, extent = side). */
void oracle_area_samples(uint64_t path_seed, const double p[3], const double light_pos[3], int square,
                         double extent, int n, double* out);
/* k-th output (k >= 1) of SplitMix64 seeded with `seed`; jittered-grid columns for n. */
uint64_t oracle_splitmix64_nth(uint64_t seed, uint64_t k);
int oracle_jitter_cols(int n);
int oracle_parse_transforms(const char* joined, double m_out[16]); /* '\n'-separated */
int oracle_hex_to_color(const char* hex, double out[3]);
int oracle_kd_stats(const char* stl_path, int32_t* n_tris, int32_t* n_leaves, int32_t* max_leaf,
                    int32_t* n_nodes, char* err, int errlen);
/* KdTree::new (src/mesh.rs:75-197) of n_tris triangles (9 f64 each), nodes in depth-first
 * pre-order (left child = node + 1): leaf flag, axis (-1 leaf), split, right child (-1 leaf),
 * the leaf's triangle list as [first, first + count) of tri_ids (leaves in node order), and
 * the node's f64 bounds (min xyz, max xyz). Arrays are filled up to their capacities;
 * n_nodes / n_ids receive the full sizes. */
int oracle_kd_tree(const double* verts, int64_t n_tris, int32_t* leaf, int32_t* axis, double* split, int32_t* right,
                   int32_t* first, int32_t* count, double* bounds, int64_t node_cap, int32_t* tri_ids,
                   int64_t id_cap, int64_t* n_nodes, int64_t* n_ids);
/* Closest hit of one ray against a mesh: k-d (use_kdtree=1) or brute force.
 * Returns triangle index or -1; writes t. */
int oracle_mesh_hit(const char* stl_path, int use_kdtree, const double origin[3], const double dir[3],
                    double t_min, double t_max, double* t_out);
/* Outline on raw buffers: depth (NaN = None), normal xyz per pixel (NaN = None). */
void oracle_outline(int32_t w, int32_t h, const double* depth, const double* normal,
                    const oracle_opts* opts, double* color_inout);

#ifdef __cplusplus
}
#endif
#endif
