/*
 * cg_device.h — the C-ABI between the crystal::graphics host layer and the
 * MI355X (gfx950) HIP backend.  It replaces the reference's WebGPU backend
 * (src/wgpu/) and its WGSL compute pass; plain pointers and sizes only.
 *
 * Conventions (mirroring the reference's std::expected<..., Error> returns,
 * src/error_stack.h:12-28):
 *   - every call returns CG_OK (0) or a CG_ERR_* code;
 *   - on failure the message is kept per env (cg_env_last_error) and per
 *     thread (cg_last_error) until the next failing call;
 *   - device work is enqueued on the env's HIP stream and the call returns
 *     without waiting (as View's submit does, src/wgpu/view.cpp:81-87), except
 *     where a function says it synchronises;
 *   - one env is not re-entrant: calls on one env come from one thread.
 * Entry points that read or write "device" buffers take pointers obtained
 * from cg_device_alloc (HBM); "host" buffers are ordinary memory.
 */
#ifndef CG_DEVICE_H_
#define CG_DEVICE_H_

#include <stddef.h>
#include <stdint.h>

#include "cg_scene.h"

#ifdef __cplusplus
extern "C" {
#endif

enum {
  CG_OK = 0,
  CG_ERR_INVALID = 1,      /* bad argument or unsupported input */
  CG_ERR_HIP = 2,          /* HIP runtime error */
  CG_ERR_OUT_OF_MEMORY = 3,
  CG_ERR_UNSUPPORTED = 4,  /* scene feature the kernels do not implement */
  CG_ERR_STATE = 5,        /* call out of order (e.g. view before upload) */
  CG_ERR_FULL = 6          /* no room now (batch slots or tables); retry later */
};

typedef struct cg_env cg_env;
typedef struct cg_scene cg_scene;

/* Traversal counters accumulated by cg_view / cg_trace_* while counting is
 * enabled (and always by a batch): closest-hit and any-hit queries, 128-B
 * 8-wide BVH nodes visited, 48-B triangle records tested, 80-B instance
 * records entered. */
typedef struct cg_stats {
  uint64_t rays;
  uint64_t shadow_rays;
  uint64_t nodes;
  uint64_t tris;
  uint64_t insts;
} cg_stats;

/* Create the backend on HIP device `device` for `conf` (render size, spp,
 * depth, lobe caps), allocating the history and target buffers.  Replaces
 * wgpu::CreateEnv (src/wgpu/environment.cpp:19-108); trace_depth is clamped
 * to [1, 64] as there (:20). */
int cg_env_create(const cg_conf* conf, int device, cg_env** out);
/* Release everything the env owns, after waiting for all its streams (view
 * batches included); wavefront batches still alive are destroyed with it.
 * Replaces wgpu::Env::Term (src/wgpu/environment.cpp:126-138). */
int cg_env_destroy(cg_env* env);
/* Last error message of this env / of this thread ("" if none). */
const char* cg_env_last_error(const cg_env* env);
const char* cg_last_error(void);
/* Wait for all enqueued work; replaces wgpu::Env::Poll
 * (src/wgpu/environment.cpp:110-116).  Reports CG_ERR_STATE if a traversal
 * since the last check needed a deeper stack than the scene's bound (its
 * walk dropped entries; the kernels still finish).  cg_read_target,
 * cg_view_batch_wait and cg_batch_poll report it too; reporting clears it. */
int cg_env_sync(cg_env* env);
/* The env's hipStream_t, for callers that time or order work on it. */
void* cg_env_stream(cg_env* env);
/* Enable (1) or disable (0) traversal counting and, for cg_view, per-pixel
 * path records (hash of the pixel's discrete decisions -- hit ids, shadow
 * outcomes with the emitter each light sample picked, sky texels -- and its
 * query count). */
int cg_env_set_counting(cg_env* env, int enable);
int cg_env_set_path_records(cg_env* env, int enable);
int cg_stats_read(cg_env* env, cg_stats* out);  /* synchronises */
/* Acceleration structure built by later cg_scene_upload calls:
 *   CG_ACCEL_REFERENCE: the SceneData's own BLAS/TLAS trees, node for node;
 *   CG_ACCEL_REBUILT (default): every distinct BLAS rebuilt with binned SAH
 *     (leaves of up to 4 triangles), identity-transform instances of at most 16
 *     triangles merged into one world-space BLAS, and a SAH TLAS;
 *   CG_ACCEL_DEVICE: the same instances and TLAS, every BLAS built on the GPU
 *     as a linear BVH over Morton codes (no counterpart in the reference; for
 *     callers that upload large or changing meshes).
 * All give the same closest hits (up to triangles tied at the same t). */
enum { CG_ACCEL_REFERENCE = 0, CG_ACCEL_REBUILT = 1, CG_ACCEL_DEVICE = 2 };
int cg_env_set_accel(cg_env* env, int mode);
/* Random-number streams of cg_view (and of cg_view_job through
 * cg_job.rng_mode):
 *   CG_RNG_REFERENCE (default): one stream per pixel consumed in the
 *     reference's order (compute.wesl:44-56, rng.wesl:1-21), so a pixel's
 *     samples run one after another on one lane;
 *   CG_RNG_SAMPLE (no counterpart in the reference): one stream per sample,
 *     seeded from the pixel's base seed b (the pixel index on frame 0, the
 *     carried seed after) as step(step(b) ^ (i + 1) * 0x9E3779B9) for sample
 *     i, where step is the reference's PCG step; the carried seed is
 *     step(step(b)).  Samples are traced on separate lanes and summed in
 *     sample order, so the image does not depend on the schedule.  The same
 *     estimator with independent streams: images converge to the
 *     reference-order ones. */
enum { CG_RNG_REFERENCE = 0, CG_RNG_SAMPLE = 1 };
int cg_env_set_rng_mode(cg_env* env, int mode);
/* Work order of cg_view and cg_view_job (no counterpart in the reference, a
 * scheduling switch for callers that time or compare orders).  On (default):
 * a render of the same size as the env's last render in the same RNG mode
 * takes its work items (tiles, or a tile's samples in chunks) longest first
 * by that render's durations; off: every render takes them in image order.
 * Either way the results are identical. */
int cg_env_set_schedule(cg_env* env, int enable);
int cg_stats_reset(cg_env* env);

/* Upload a flattened scene and build the device layout (quantised 8-wide BVH
 * nodes, triangle records, instance records, material and emitter records,
 * alias table, environment and material texture arrays).  Replaces
 * WriteScene + WriteMaterialTextures + UpdateComputeBindGroup2
 * (src/wgpu/resource.cpp:338-414,471-546, src/wgpu/compute.cpp:332-372).
 * Texture arrays larger than 16384 x 16384 per layer, or of 65535 layers or
 * more, give CG_ERR_UNSUPPORTED. */
int cg_scene_upload(cg_env* env, const cg_scene_data* sd, cg_scene** out);
int cg_scene_release(cg_env* env, cg_scene* scene);
/* Refit (no counterpart in the reference): a scene uploaded with
 * CG_ACCEL_DEVICE takes new scene data with the same triangles and BLASes
 * (vertex positions, instance transforms, materials, emitters may differ):
 * every device-built BLAS keeps its tree and gets new boxes, the TLAS is
 * rebuilt, the other records are replaced.  Waits for the env's stream;
 * other launches on the scene must have finished.  CG_ERR_INVALID if the
 * scene was uploaded otherwise or the structure differs. */
int cg_scene_refit(cg_env* env, cg_scene* scene, const cg_scene_data* sd);
/* Host-only check of the device layout cg_scene_upload would build (no GPU
 * needed): every quantised BVH node's decoded child boxes contain the exact
 * ones.  Reports the node count and the traversal stack bound. */
int cg_layout_check(const cg_scene_data* sd, int accel, uint64_t* nodes, uint32_t* stack_depth);
/* Device bytes held by a scene, and the traversal stack depth it needs. */
int cg_scene_footprint(const cg_scene* scene, uint64_t* bytes, uint32_t* stack_depth);

/* One progressive frame of the compute pass (shader/compute.wesl:32-62):
 * camera ray per pixel, spp spectral samples of the depth-first transport
 * tree in the reference's per-pixel RNG order, running-mean CIE XYZ history,
 * sRGB rgba8 target.  iter_count == 0 starts a new accumulation.  Replaces
 * wgpu::Env::View (src/wgpu/view.cpp:16-88). */
int cg_view(cg_env* env, cg_scene* scene, const cg_camera* camera, uint32_t iter_count);
/* Headless read-back of the target (new; the reference only presents):
 * history (W*H cg_history_pixel) and/or rgba8 (W*H*4).  Synchronises. */
int cg_read_target(cg_env* env, cg_history_pixel* history, uint8_t* rgba8);
/* Per-pixel path records of the last cg_view (W*H each).  Synchronises. */
int cg_read_path_records(cg_env* env, uint32_t* path_hash, uint32_t* ray_count);
/* Overwrite the history buffer (W*H entries) from host memory. */
int cg_write_history(cg_env* env, const cg_history_pixel* history);

/* Batch closest hit, one query per ray: the reference's rayHit
 * (shader/compute/structural/bvh.wesl:52-130) as a standalone entry point.
 * rays / hits (and uv: 2 floats per ray, nullable) are device buffers. */
int cg_trace_closest(cg_env* env, cg_scene* scene, const cg_ray* rays, uint32_t n, cg_hit* hits, float* uv);
/* cg_trace_closest without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * results equal cg_trace_closest's.  Its work-queue counters and overflow
 * traversal stack live in a fixed 32 MiB scratch array per device, resolved
 * when the env is created; the counters are cleared by a kernel at each call.
 * Scenes whose traversal stack would need more than that scratch for one
 * block per CU give CG_ERR_INVALID.  All calls on one GPU share its scratch:
 * calls that overlap on one GPU (two envs or streams of one device) race. */
int cg_trace_closest_noalloc(cg_env* env, cg_scene* scene, const cg_ray* rays, uint32_t n, cg_hit* hits, float* uv);
/* Batch occlusion: occluded[i] = 1 iff a hit exists with t < tmax[i]
 * (trace.wesl:104-111 asks exactly this of its shadow rays). */
int cg_trace_occluded(cg_env* env, cg_scene* scene, const cg_ray* rays, const float* tmax, uint32_t n,
                      uint32_t* occluded);

/* The emitter next-event estimation picks for given uniforms (u: 2n floats,
 * pairs (u1, u2); out: n indices; device buffers): the reference's
 * sampleEmissivePrim (shader/compute/shading/pdf/emission.wesl:181-191) on
 * the scene's uploaded alias table, with the same device code the renders
 * use.  bin = u32(u1 * N) reaches N at u1 = 1.0; reads and the returned
 * emitter are clamped to the tables as WebGPU robustness clamps them. */
int cg_alias_pick(cg_env* env, cg_scene* scene, const float* u, uint32_t n, uint32_t* out);

/* Measurement probe (no reference counterpart): copy `bytes` (a multiple of
 * 16) from src to dst, both device buffers, with a plain coalesced 16-byte
 * vector copy on the env's stream.  Its rate is the achievable HBM
 * bandwidth, and its counters the scale for hardware byte counters. */
int cg_stream_copy(cg_env* env, void* dst, const void* src, size_t bytes);
/* Measurement probe (no reference counterpart): read `bytes` (a multiple of
 * 16) of the device buffer src once, four 16-byte loads in flight per lane,
 * writing one XOR word per workgroup to the device buffer sink (at least
 * 4096 words).  Its rate is the achievable read-only HBM bandwidth, the
 * denominator for read-dominated kernels such as traversal. */
int cg_stream_read(cg_env* env, const void* src, size_t bytes, uint32_t* sink);

/* HBM buffers for the batch entry points. */
int cg_device_alloc(cg_env* env, size_t bytes, void** dptr);
int cg_device_free(cg_env* env, void* dptr);
int cg_memcpy_h2d(cg_env* env, void* dst, const void* src, size_t bytes);  /* synchronises */
int cg_memcpy_d2h(cg_env* env, void* dst, const void* src, size_t bytes);  /* synchronises */

/* ---------------------------------------------------------------------------
 * Batched rendering (a form the reference does not have: it renders one
 * View at a time, src/wgpu/view.cpp:16-88).  A job is that View's work -- one
 * progressive frame of one scene and camera -- at its own size, spp, depth
 * and lobe caps, with its own history.  A batch keeps a pool of pixel slots
 * whose path state lives in HBM and advances every admitted job together in
 * rounds (trace the queued rays, advance each path one event, queue its next
 * ray), `rounds_per_launch` rounds per launch, captured once as a hipGraph.
 * Jobs are admitted between launches (continuous batching); each pixel's
 * result equals cg_view's for the same scene, camera, conf and history, bit
 * for bit (the same per-pixel operations in the same order).
 * Scenes must stay uploaded while a job on them is active. */
typedef struct cg_batch cg_batch;
typedef struct cg_batch_conf {
  uint32_t slot_capacity;      /* pixels resident at once (rounded up to 64) */
  uint32_t max_trace_depth;    /* deepest job accepted (1-64) */
  uint32_t rounds_per_launch;  /* even, >= 2 */
  uint32_t use_graph;          /* 1: one launch = one replay of a captured hipGraph */
  uint32_t max_jobs;           /* job table entries (1-4096) */
  uint32_t max_stack_depth;    /* deepest traversal stack of an accepted scene */
  uint32_t block_size;         /* workgroup size of the round kernels: 64, 128 or 256 (0: 256) */
  uint32_t short_stack;        /* LDS traversal-stack entries per lane: 8, 16 or 24 (0: 16) */
} cg_batch_conf;
typedef struct cg_job {
  cg_scene* scene;
  cg_camera camera;
  cg_conf conf;                /* render size, sample_count, trace_depth, lobe caps */
  uint32_t iter_count;         /* 0 starts a new accumulation */
  uint32_t rng_mode;           /* CG_RNG_REFERENCE or CG_RNG_SAMPLE (cg_view_job, cg_view_batch_launch;
                                  the wavefront engine takes CG_RNG_REFERENCE only) */
  cg_history_pixel* history;   /* device, W*H: read when iter_count > 0, written */
  uint8_t* rgba8;              /* device, W*H*4, nullable */
  uint32_t* path_hash;         /* device, W*H, nullable (path records) */
  uint32_t* ray_count;         /* device, W*H, nullable */
} cg_job;
/* One job with the plain one-lane-per-pixel kernel of cg_view, on the job's
 * own conf and buffers (cg_view renders the env's conf into the env's
 * buffers).  The one-job-at-a-time form a batch is measured against. */
int cg_view_job(cg_env* env, const cg_job* job);
/* Many jobs in one persistent launch of the plain per-pixel kernel: the work
 * items are the 8x8 tiles of all n jobs (the most samples x trace depth per
 * pixel first), so the tiles of other jobs fill the GPU while one job's heavy
 * tiles finish.  The counter reset and the kernel are captured as one
 * hipGraph and launched as such.  Each job's
 * result equals cg_view_job's, bit for bit.  The launch runs on a stream of
 * its own (taken from a pool; default priority), so successive launches
 * overlap; query / wait on the returned handle, then release it.  The job
 * table is copied at launch; buffers must stay valid until it finishes. */
typedef struct cg_vbatch cg_vbatch;
int cg_view_batch_launch(cg_env* env, const cg_job* jobs, uint32_t n, cg_vbatch** out);
int cg_view_batch_query(cg_vbatch* batch, int* done);
int cg_view_batch_wait(cg_vbatch* batch);
int cg_view_batch_release(cg_vbatch* batch);  /* waits, then returns the stream to the pool */
int cg_batch_create(cg_env* env, const cg_batch_conf* conf, cg_batch** out);
int cg_batch_destroy(cg_batch* batch);
/* Device bytes held, and the largest run of free slots (what fits now). */
int cg_batch_footprint(const cg_batch* batch, uint64_t* bytes, uint32_t* slots_free);
/* Admit a job: CG_ERR_FULL when its (ceil(W/8)*ceil(H/8)*64) slots or a
 * table entry are not free now (job entries and the 64 scene entries free up
 * as jobs finish; a scene entry is reused once no active job reads it). */
int cg_batch_submit(cg_batch* batch, const cg_job* job, uint32_t* job_id);
/* Enqueue `launches` launches of rounds (asynchronous). */
int cg_batch_launch(cg_batch* batch, uint32_t launches);
/* Jobs finished since the last poll (ids into `done`, up to cap), and the
 * number still active.  Synchronises. */
int cg_batch_poll(cg_batch* batch, uint32_t* done, uint32_t cap, uint32_t* n_done, uint32_t* n_active);
/* Launch and poll until no job is active.  Synchronises. */
int cg_batch_drain(cg_batch* batch);
/* Traversal counters of the batch's rounds so far (optionally reset). */
int cg_batch_stats(cg_batch* batch, cg_stats* out, int reset);

#ifdef __cplusplus
}
#endif

#endif /* CG_DEVICE_H_ */
