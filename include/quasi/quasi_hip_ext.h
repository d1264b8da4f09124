/* quasi_hip_ext.h -- optional extension exports of libquasi_hip.so.
 *
 * None of these exist in the reference.  A host discovers them with dlsym
 * exactly as the reference loader discovers its optional v2/v3 symbols
 * (loader.hpp:114-128), so the ABI version stays 3 and an unmodified host that
 * never looks them up still works.
 *
 * Conventions (all functions):
 *   - every struct starts with uint32_t struct_size = sizeof(struct); structs only
 *     grow at the end, and a caller built against an older header passes its
 *     smaller size (Q_render_params and Q_render_stats: at least the size they
 *     had without the bvh_builder fields);
 *   - int32_t return: Q_EXT_OK (0) or a negative Q_EXT_ERR_* code, with a message
 *     sent through Q_plugin_context.log;
 *   - inputs are caller-owned and copied before the call returns;
 *   - no C++ exception and no HIP error crosses the ABI.
 */
#ifndef QUASI_HIP_EXT_H
#define QUASI_HIP_EXT_H

#include <stddef.h>
#include <stdint.h>

#include "plugin_interface.h"

#ifdef __cplusplus
extern "C" {
#endif

#define Q_EXT_OK               0
#define Q_EXT_ERR_ARGUMENT    -1
#define Q_EXT_ERR_STATE       -2
#define Q_EXT_ERR_DEVICE      -3
#define Q_EXT_ERR_OUT_OF_MEMORY -4
#define Q_EXT_ERR_UNSUPPORTED -5

/* Random stream seeding.
 *   Q_RNG_REFERENCE: state = pcg_hash(x + y*1920 + sample*1920*1080) (plugin.mm:295)
 *   Q_RNG_HASHED:    state = pcg_hash((sample*W*H + y*W + x) ^ pcg_hash(seed))
 * The stream itself is always state = pcg_hash(state), u = float(state)/2^32
 * (plugin.mm:98-107). */
#define Q_RNG_REFERENCE 0u
#define Q_RNG_HASHED    1u

/* Light transport.
 *   Q_INTEGRATOR_REFERENCE: the reference bounce loop (plugin.mm:223-276): emission
 *     collected when a BSDF-sampled ray hits an emitter, no next-event estimation.
 *   Q_INTEGRATOR_NEE: emission seen directly only by camera rays; at every
 *     non-emissive hit before the last segment one light triangle is sampled
 *     (probability proportional to area x summed radiance) and a shadow ray is
 *     traced.  Same expectation as the reference loop at the same max_depth. */
#define Q_INTEGRATOR_REFERENCE 0u
#define Q_INTEGRATOR_NEE       1u

/* Execution engine for the sample passes. */
#define Q_ENGINE_MEGAKERNEL 0u  /* one thread per (pixel, sample) path, whole path in one kernel */
#define Q_ENGINE_WAVEFRONT  1u  /* raygen / extend / shade / shadow kernels over compacted queues */
#define Q_ENGINE_PERSISTENT 2u  /* single launch per pass: one persistent grid runs every path from its camera
                                   ray to its end (closest hits, shading, shadow rays), lanes refilled from a
                                   shared path queue; then the per-pixel accumulation.  Same images as
                                   Q_ENGINE_WAVEFRONT bit for bit; block_size 128 or 256 */

/* Q_render_params.flags */
#define Q_FLAG_RECORD_IDS    1u  /* record primary triangle ids of sample id_sample */
#define Q_FLAG_TIME_KERNELS  2u  /* bracket every kernel launch with hipEvents */

/* Indexed triangle mesh with per-triangle material.  Triangles whose largest
 * emission channel exceeds 0.1 are emitters (plugin.mm:243, :255). */
typedef struct Q_scene_desc {
    uint32_t struct_size;
    uint32_t num_vertices;
    uint32_t num_triangles;
    uint32_t reserved;
    const float* vertices;    /* xyz, 3 * num_vertices */
    const uint32_t* indices;  /* 3 * num_triangles */
    const float* albedo;      /* rgb per triangle, 3 * num_triangles */
    const float* emission;    /* rgb radiance per triangle, 3 * num_triangles */
} Q_scene_desc;

/* Defaults after Q_plugin_create mirror the reference: spp_per_render 1,
 * max_depth 5 (plugin.mm:22), rr_start 3 (plugin.mm:264, "bounce > 2"),
 * seed 0, Q_RNG_REFERENCE, Q_INTEGRATOR_REFERENCE, Q_ENGINE_MEGAKERNEL. */
typedef struct Q_render_params {
    uint32_t struct_size;
    uint32_t spp_per_render;   /* samples per pixel added by one Q_plugin_render call */
    uint32_t max_depth;        /* closest-hit traces per path, 1..16 */
    uint32_t rr_start;         /* Russian roulette when bounce index >= rr_start */
    uint32_t seed;
    uint32_t rng_mode;         /* Q_RNG_* */
    uint32_t integrator;       /* Q_INTEGRATOR_* */
    uint32_t engine;           /* Q_ENGINE_* */
    uint32_t bvh_width;        /* wavefront traversal: 2 or 4 (0 = 4); the megakernel walks BVH2 */
    uint32_t samples_per_pass; /* wavefront: samples per pixel in flight at once (0 = auto: about
                                  2^27 paths per pass; for a batch, the largest power of two that
                                  divides the samples requested and stays within that) */
    uint32_t block_size;       /* wavefront traversal threads per block: 64, 128, 256 (0 = 128 for
                                  bvh_width 4, 256 for bvh_width 2) */
    uint32_t flags;            /* Q_FLAG_* */
    uint32_t id_sample;        /* sample index whose primary ids are recorded */
    uint32_t queue_chunk;      /* wavefront: 1 = each idle lane takes its next ray from the queue
                                  (persistent waves, refilled when 32 lanes are idle); 64 or 256 =
                                  each persistent wave takes that many rays at a time;
                                  Q_QUEUE_LDS_TOP = as 1, with the top BVH4 levels staged in LDS
                                  (single scenes, width 4: four levels at block 256, three at 128;
                                  batches fetch as 1);
                                  Q_QUEUE_QUANT = as Q_QUEUE_LDS_TOP over 64-byte BVH4 nodes whose
                                  child boxes are quantised to 8 bits per plane against the node's
                                  box (four loads per node step instead of seven; batches too,
                                  without the LDS staging; the same images);
                                  0 = Q_QUEUE_QUANT where a kernel has it (width 4, block 128 or
                                  256), else Q_QUEUE_LDS_TOP, else 1;
                                  Q_QUEUE_FLAT = no persistence: one thread per queue slot, the
                                  plain wavefront tracer (block 128 only) */
    uint32_t bvh_builder;      /* Q_BVH_BUILD_*: where Q_plugin_set_scene builds the BVH */
    uint32_t reserved;
} Q_render_params;

/* Q_render_params.bvh_builder (the same binned-SAH algorithm either way; any
 * valid tree renders the same image, because the closest hit does not
 * depend on the tree) */
/* Q_render_params.queue_chunk: the plain (non-persistent) wavefront traversal */
#define Q_QUEUE_FLAT 0xFFFFFFFFu
#define Q_QUEUE_LDS_TOP 2u
#define Q_QUEUE_QUANT 3u

#define Q_BVH_BUILD_AUTO 0u   /* on the GPU from 2^14 triangles, else on the host */
#define Q_BVH_BUILD_HOST 1u   /* host threads */
#define Q_BVH_BUILD_GPU  2u   /* GPU kernels (falls back to the host when the GPU build fails) */

#define Q_STATS_MAX_KERNELS 16

typedef struct Q_kernel_stat {
    char name[48];
    uint64_t launches;
    double total_ms;           /* summed hipEvent time (Q_FLAG_TIME_KERNELS only) */
} Q_kernel_stat;

typedef struct Q_render_stats {
    uint32_t struct_size;
    uint32_t width;
    uint32_t height;
    uint32_t sample_count;     /* samples per pixel in the accumulation */
    uint32_t num_triangles;
    uint32_t num_lights;
    uint32_t bvh_nodes;
    uint32_t bvh_depth;
    uint64_t rays_camera;      /* BVH queries since the last Q_plugin_reset_stats */
    uint64_t rays_extension;
    uint64_t rays_shadow;
    double bvh_build_ms;       /* host wall time of the last BVH build + upload */
    double render_ms;          /* hipEvent time of all render calls since the reset */
    uint32_t num_kernels;
    uint32_t reserved;
    Q_kernel_stat kernels[Q_STATS_MAX_KERNELS];
    uint32_t bvh_builder;      /* Q_BVH_BUILD_HOST or Q_BVH_BUILD_GPU: where the current BVH was built */
    uint32_t reserved2;
    double bvh_sah_cost;       /* SAH cost of the current BVH2 (traversal 1, triangle 1, root area 1) */
    double bvh_device_ms;      /* GPU build: hipEvent time of the build kernels */
    uint32_t bvh_references;   /* triangle records in the BVH leaves: num_triangles plus the extra
                                  references of triangles split before the build (a large
                                  triangle is split into up to four with smaller boxes) */
    uint32_t reserved3;
} Q_render_stats;

/* Replaces the hard-coded scene of plugin.mm:584 (cornell_box.hpp:95-134) with a
 * caller mesh; builds the BVH and resets the accumulation. */
int32_t Q_plugin_set_scene(Q_plugin_handle* handle, const Q_scene_desc* scene);

/* Reads / replaces the render parameters; set_params resets the accumulation. */
int32_t Q_plugin_get_params(Q_plugin_handle* handle, Q_render_params* out);
int32_t Q_plugin_set_params(Q_plugin_handle* handle, const Q_render_params* params);

/* Primary-hit triangle index (input order; 0xFFFFFFFF for a miss) of sample
 * id_sample for every pixel, row-major top-to-bottom.  Synchronises. */
int32_t Q_plugin_readback_ids(Q_plugin_handle* handle, uint32_t* out, uint64_t count);

/* Replaces the display pass of plugin.mm:334-350 (tonemap_frag, drawn to the
 * drawable): the beauty mean, Reinhard c / (c + 1), gamma 1/2.2, as 8-bit RGBA
 * (unorm: clamped, times 255, rounded to nearest even; alpha 255), row-major
 * top-to-bottom, 4 * width * height bytes into out.  Synchronises.
 * Q_EXT_ERR_ARGUMENT when out holds fewer bytes, Q_EXT_ERR_STATE before the
 * first sample. */
int32_t Q_plugin_readback_display(Q_plugin_handle* handle, uint8_t* out, uint64_t bytes);

/* Synchronises the plugin's streams, then fills *out. */
int32_t Q_plugin_get_stats(Q_plugin_handle* handle, Q_render_stats* out);
int32_t Q_plugin_reset_stats(Q_plugin_handle* handle);

/* Waits until every queued render has finished. */
int32_t Q_plugin_synchronize(Q_plugin_handle* handle);

/* ---- Batched scenes --------------------------------------------------------
 * Many scenes of uneven size (and resolution) traced together: every sample
 * pass covers all scenes with shared, compacted ray queues, and the pass is
 * captured once in a hipGraph and replayed per pass.  The render parameters
 * (depth, integrator, seeding mode, traversal shape, samples_per_pass: samples
 * of every scene per replayed pass) are those of Q_plugin_set_params;
 * each scene brings its own camera, resolution and seed.  With the same
 * parameters, scene i of a batch renders exactly what a single-scene render
 * of it with seed = scenes[i].seed and the same samples_per_pass gives. */
typedef struct Q_batch_scene {
    uint32_t width;
    uint32_t height;
    uint32_t seed;             /* seed of this scene's Q_RNG_HASHED streams */
    uint32_t reserved;
    Q_camera camera;
    Q_scene_desc mesh;
} Q_batch_scene;

typedef struct Q_batch_desc {
    uint32_t struct_size;
    uint32_t num_scenes;
    const Q_batch_scene* scenes;
} Q_batch_desc;

/* Builds every scene's BVH (host threads), uploads all scenes and clears the
 * batch accumulation.  Replaces any previous batch. */
int32_t Q_plugin_set_batch(Q_plugin_handle* handle, const Q_batch_desc* batch);

/* Adds spp samples to every scene of the batch, asynchronously on `stream`
 * (a hipStream_t; NULL: the plugin's stream).  spp must be a multiple of an
 * explicit samples_per_pass. */
int32_t Q_plugin_render_batch(Q_plugin_handle* handle, uint32_t spp, void* stream);

/* The mean AOVs of one scene (malloc'd, free with Q_plugin_readback_aov_free). */
int32_t Q_plugin_readback_batch(Q_plugin_handle* handle, uint32_t scene, Q_readback_aov_result* out);

/* Primary ids of one scene's sample id_sample (see Q_plugin_readback_ids). */
int32_t Q_plugin_readback_batch_ids(Q_plugin_handle* handle, uint32_t scene, uint32_t* out, uint64_t count);

/* ---- Autotuner --------------------------------------------------------------
 * Times every compiled traversal shape (BVH width x block size x queue chunk)
 * on probe_spp samples of the current single scene (Q_AUTOTUNE_SINGLE, with
 * the last camera given to Q_plugin_render) or of the current batch
 * (Q_AUTOTUNE_BATCH), then, with the fastest shape, the pass sizes (samples
 * per pass, i.e. the ray-queue lengths: powers of two up to about 2^27 paths
 * and probe_spp, the six largest; a batch's divide probe_spp) when the probe
 * leaves more than one; keeps the fastest in the render parameters (bvh_width,
 * block_size, queue_chunk, samples_per_pass) and clears the accumulation.  The probes' rays, launches and kernel times are not
 * counted: Q_plugin_get_stats reads the same before and after.  Timing the
 * single target needs a frame rendered first (Q_EXT_ERR_STATE otherwise; a
 * choice served from the cache does not).  Results are cached, one JSON
 * object per line, in $QUASI_HIP_TUNE_CACHE (default
 * ~/.cache/quasi_hip/autotune.jsonl), keyed by GPU architecture, target,
 * workload signature and a fingerprint of the scene content. */
#define Q_AUTOTUNE_SINGLE 0u
#define Q_AUTOTUNE_BATCH  1u
#define Q_AUTOTUNE_USE_CACHE 1u   /* flags: take a cached choice without timing */
#define Q_AUTOTUNE_NO_STORE  2u   /* flags: do not write the cache */
#define Q_AUTOTUNE_MAX_ENTRIES 32

typedef struct Q_autotune_entry {
    uint32_t bvh_width;
    uint32_t block_size;
    uint32_t queue_chunk;
    uint32_t samples_per_pass; /* pass size of the probe (samples per pixel in flight: the queue length) */
    double ms;                 /* hipEvent time of the probe */
} Q_autotune_entry;

typedef struct Q_autotune_result {
    uint32_t struct_size;
    uint32_t num_entries;
    uint32_t best;             /* index into entries */
    uint32_t from_cache;
    Q_autotune_entry entries[Q_AUTOTUNE_MAX_ENTRIES];
} Q_autotune_result;

int32_t Q_plugin_autotune(Q_plugin_handle* handle, uint32_t target, uint32_t probe_spp, uint32_t flags,
                          Q_autotune_result* out);

/* ---------------------------------------------------------------- adjoint
 * Gradients of a scalar image loss L with respect to the current scene's
 * per-triangle albedo and vertex positions (no reference counterpart; the
 * reference renders a fixed scene and has no derivatives).
 *
 * The image is I = (1/spp) sum_s C_s over samples sample_begin ..
 * sample_begin + spp - 1 of the current camera (the last rendered frame's),
 * resolution, seed (or desc->seed with Q_ADJOINT_SEED), depth, Russian
 * roulette start and integrator.  Each path is re-traced from its RNG seed
 * and differentiated with its discrete decisions held fixed: the triangles
 * hit, facing sides, light choices, shadow-ray visibility, Russian-roulette
 * outcomes and probabilities.  Emitter geometry (light sampling) is held
 * fixed; the gradients of vertices are those of the non-emitting triangles
 * that use them.
 *
 * mode Q_ADJOINT_IMAGE_GRADIENT: image = dL/dI, RGBA32F top-to-bottom, the
 *   alpha channel ignored (the pixels with dL/dI = 0 are skipped).
 * mode Q_ADJOINT_L2_TARGET: image = target T (RGBA32F); L = sum over pixels
 *   and RGB of (M - T)^2 with M the current accumulated mean image, so
 *   dL/dI = 2 (M - T); L is returned in result->loss.
 * Gradients are written in FP64, in input order: grad_albedo 3 per
 * triangle, grad_vertices 3 per vertex (either may be NULL).
 *
 * Albedo parameter map: with albedo_param, triangle t's albedo is parameter
 * albedo_param[t] (< num_albedo_params; 0xFFFFFFFF: no parameter), shared by
 * every triangle mapped to it, and grad_albedo holds 3 * num_albedo_params
 * gradients: each the sum of its triangles' gradients, reduced on the device
 * in FP64.
 *
 * Passes and memory: each pass traces and records at most max_pass_paths
 * paths (0: Q_ADJOINT_PASS_PATHS): samples of the whole image, or, when one
 * sample of the image holds more pixels, one sample of a band of full rows
 * (per-tile path recomputation: every band is traced, recorded and
 * replayed on its own).  The working memory beyond the scene is then
 * max_pass_paths * (9 * max_depth + 188) bytes of path records and path state,
 * 4 * 8 * 3 * (num_triangles + num_vertices) bytes of FP32 gradient
 * replicas, 8 * 3 * (num_triangles + num_vertices + num_albedo_params) bytes
 * of FP64 gradients, 3 * 16 * width * height bytes of images and 16 bytes per
 * triangle of vertex indices and map (result->working_bytes reports it for
 * the call; buffers kept from earlier, larger calls are not released).
 */
#define Q_ADJOINT_IMAGE_GRADIENT 0u
#define Q_ADJOINT_L2_TARGET      1u
#define Q_ADJOINT_SEED           1u   /* flags: use desc->seed instead of the render seed */
#define Q_ADJOINT_PASS_PATHS     (1u << 24)
#define Q_ADJOINT_NO_PARAM       0xFFFFFFFFu

typedef struct Q_adjoint_desc {
    uint32_t struct_size;
    uint32_t mode;
    uint32_t spp;              /* samples replayed per pixel (>= 1) */
    uint32_t sample_begin;     /* first sample index */
    uint32_t seed;             /* with Q_ADJOINT_SEED */
    uint32_t flags;
    const float* image;        /* width * height * 4 floats (host memory) */
    double* grad_albedo;       /* 3 * num_triangles (3 * num_albedo_params with a map), or NULL */
    double* grad_vertices;     /* 3 * num_vertices, or NULL */
    /* fields below are 0 for callers whose struct_size ends before them */
    uint32_t max_pass_paths;   /* paths per recorded pass (0: Q_ADJOINT_PASS_PATHS) */
    uint32_t num_albedo_params;
    const uint32_t* albedo_param;   /* num_triangles entries (host memory), or NULL: per-triangle albedo */
} Q_adjoint_desc;

typedef struct Q_adjoint_result {
    uint32_t struct_size;
    uint32_t reserved;
    uint64_t paths;            /* paths replayed */
    double loss;               /* Q_ADJOINT_L2_TARGET: L; otherwise 0 */
    double kernel_ms;          /* hipEvent time of the adjoint and reduction kernels */
    /* fields below are written only when struct_size covers them */
    uint64_t working_bytes;    /* device memory the call needs beyond the scene and framebuffer */
    uint32_t passes;           /* recorded passes (samples x row bands) */
    uint32_t bands;            /* row bands per sample block (1: whole image) */
} Q_adjoint_result;

int32_t Q_plugin_adjoint(Q_plugin_handle* handle, const Q_adjoint_desc* desc, Q_adjoint_result* result);

/* ---------------------------------------------------------------- no-alloc render
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc): the wavefront render of Q_plugin_render with every piece of
 * working memory (path state, ray queues, queue counters, traversal-stack
 * overflow) in one fixed device array of Q_NOALLOC_SCRATCH_BYTES per device,
 * set up by Q_plugin_create.  The call allocates nothing.
 *
 * Adds samples sample_begin .. sample_begin + spp - 1 of the current scene,
 * seen from *camera at width x height (render parameters: those of
 * Q_plugin_set_params; the engine is always the wavefront one), to the
 * caller's device buffer `sums`: 4 planes of width * height RGBA32F sums,
 * beauty, albedo, normal, depth (the Q_aov_type order), row-major
 * top-to-bottom, each sample adding 1 to alpha -- what Q_plugin_render
 * accumulates, before Q_plugin_readback divides by the sample count.
 * Asynchronous on `stream` (a hipStream_t; NULL: the plugin's stream).
 * Primary ids are not recorded and the render statistics are not changed.
 *
 * A pass holds samples_per_pass samples (0: as many as fit) of every pixel;
 * Q_EXT_ERR_ARGUMENT when one pass of the image does not fit the array
 * (2.1M to 2.8M paths, depending on the traversal-stack overflow the
 * scene's BVH needs: 1920 x 1080 fits at 1 sample per pass).  With the same
 * samples_per_pass the sums equal Q_plugin_render's bit for bit.  Every call
 * on one device shares that device's array: calls that overlap on one device
 * (two streams, two plugin instances) race. */
#define Q_NOALLOC_SCRATCH_BYTES (512ull << 20)

int32_t Q_plugin_render_noalloc(Q_plugin_handle* handle, const Q_camera* camera, uint32_t width, uint32_t height,
                                uint32_t sample_begin, uint32_t spp, float* sums, void* stream);

#ifdef __cplusplus
}  /* extern "C" */
#endif

#endif /* QUASI_HIP_EXT_H */
