// cg_device.hip — HIP backend of the spectral path tracer for MI355X
// (gfx950): scene upload into the device layout, the per-pixel compute pass
// and the batch traversal entry points behind the C-ABI of cg_device.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../../include/cg_device.h"
#include "cg_core.h"
#include "cg_layout.h"
#include "cg_render.h"
#include "cg_upload.h"

using namespace cg;

namespace {

thread_local std::string t_last_error;

// ---------------------------------------------------------------------------
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
};

}  // namespace

struct cg_scene {
  DevBuf bvh, nodes, tris, insts, vidx, vattr, mats, emit, alias, emit_order, env;
  DevScene view{};
  uint32_t stack_depth = 0;
  uint64_t bytes = 0;
};

struct cg_env {
  cg_conf conf{};
  int device = 0;
  hipStream_t stream = nullptr;
  std::string last_error;
  DevBuf history, rgba, path_hash, ray_count, stats;
  DevBuf wl_cdf, wl_leaf, queue, ovf;
  int cu_count = 0;
  float wl_avg = 0.0f;
  bool counting = false;
  bool path_records = false;
  int accel = CG_ACCEL_REBUILT;
  std::vector<cg_scene*> scenes;
};

namespace {

int fail(cg_env* env, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  t_last_error = buf;
  if (env) env->last_error = buf;
  return code;
}
#define CG_HIP(env, call)                                                                      \
  do {                                                                                         \
    hipError_t e_ = (call);                                                                    \
    if (e_ != hipSuccess)                                                                      \
      return fail((env), e_ == hipErrorOutOfMemory ? CG_ERR_OUT_OF_MEMORY : CG_ERR_HIP,        \
                  "%s: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, __LINE__);         \
  } while (0)

int dev_alloc(cg_env* env, DevBuf& b, size_t bytes) {
  b.bytes = bytes;
  b.p = nullptr;
  if (bytes == 0) return CG_OK;
  CG_HIP(env, hipMalloc(&b.p, bytes));
  return CG_OK;
}
void dev_free(DevBuf& b) {
  if (b.p) (void)hipFree(b.p);
  b.p = nullptr;
  b.bytes = 0;
}
int dev_upload(cg_env* env, DevBuf& b, const void* src, size_t bytes) {
  int rc = dev_alloc(env, b, bytes);
  if (rc != CG_OK || bytes == 0) return rc;
  CG_HIP(env, hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, env->stream));
  return CG_OK;
}

// Work distribution.  Blocks are issued round-robin over the 8 XCDs
// (block b runs on XCD b % 8).  Work items (64 rays, or an 8x8 pixel tile:
// one wave's worth) are split into 8 contiguous ranges, one per XCD, so an
// XCD's L2 sees neighbouring rays or pixels.  Each wave takes items from its
// own XCD's range through an atomic counter and, once that range is
// exhausted, steals from the next XCD's.  The grid is the resident size, so
// waves loop until every range is empty; waves never wait for each other.
constexpr uint32_t kXcds = 8u;
constexpr uint32_t kNoItem = 0xFFFFFFFFu;
constexpr uint32_t kCounterStride = 32u;  // one 128-byte line per counter
struct WorkQueue {
  uint32_t* counters;  // kXcds counters, kCounterStride apart, zero at launch
  uint32_t total;      // items
};
__device__ __forceinline__ uint32_t grab_item(const WorkQueue& q, uint32_t& probe) {
  uint32_t item = kNoItem;
  if ((threadIdx.x & 63u) == 0u) {
    const uint32_t per = (q.total + kXcds - 1u) / kXcds;
    const uint32_t home = blockIdx.x % kXcds;
    while (probe < kXcds) {
      const uint32_t x = (home + probe) % kXcds;
      const uint32_t begin = x * per;
      const uint32_t end = begin + per < q.total ? begin + per : q.total;
      if (begin < end) {
        const uint32_t c = atomicAdd(q.counters + x * kCounterStride, 1u);
        if (begin + c < end) {
          item = begin + c;
          break;
        }
      }
      ++probe;
    }
  }
  return __shfl(item, 0);
}

__global__ void k_zero_u32(uint32_t* p, uint32_t n) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

__device__ __forceinline__ uint64_t wave_sum(uint64_t v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
__device__ __forceinline__ void flush_stats(unsigned long long* stats, uint64_t rays, uint64_t shadows,
                                            uint64_t nodes, uint64_t tris, uint64_t insts) {
  rays = wave_sum(rays);
  shadows = wave_sum(shadows);
  nodes = wave_sum(nodes);
  tris = wave_sum(tris);
  insts = wave_sum(insts);
  if ((threadIdx.x & 63u) == 0u) {
    atomicAdd(stats + 0, static_cast<unsigned long long>(rays));
    atomicAdd(stats + 1, static_cast<unsigned long long>(shadows));
    atomicAdd(stats + 2, static_cast<unsigned long long>(nodes));
    atomicAdd(stats + 3, static_cast<unsigned long long>(tris));
    atomicAdd(stats + 4, static_cast<unsigned long long>(insts));
  }
}

// ---------------------------------------------------------------------------
// Batch traversal: one ray per lane, the short traversal stack in LDS
// interleaved by lane (entry k of lane i at k * 256 + i: conflict-free), the
// overflow in device memory interleaved over the whole grid.
constexpr int kBlock = 256;

struct StackArgs {
  uint2* ovf;
  int ovf_cap;
};
__device__ __forceinline__ TStack lane_stack(uint32_t* s_ref, float* s_t, const StackArgs& sa) {
  const uint32_t g = blockIdx.x * kBlock + threadIdx.x;
  return TStack{s_ref + threadIdx.x, s_t + threadIdx.x, kBlock, sa.ovf ? sa.ovf + g : nullptr, gridDim.x * kBlock,
                sa.ovf ? sa.ovf_cap : 0};
}

template <bool kAnyHit>
__global__ __launch_bounds__(kBlock) void k_trace(DevScene S, const cg_ray* __restrict__ rays,
                                                  const float* __restrict__ tmax, uint32_t n,
                                                  cg_hit* __restrict__ hits, float* __restrict__ uv,
                                                  uint32_t* __restrict__ occluded,
                                                  unsigned long long* __restrict__ stats, WorkQueue q, StackArgs sa) {
  __shared__ uint32_t s_ref[kShortStack * kBlock];
  __shared__ float s_t[kShortStack * kBlock];
  const TStack stk = lane_stack(s_ref, s_t, sa);
  TravStats st{0u, 0u, 0u};
  uint64_t nq = 0;
  uint32_t probe = 0;
  for (;;) {
    const uint32_t item = grab_item(q, probe);
    if (item == kNoItem) break;
    const uint32_t i = item * 64u + (threadIdx.x & 63u);
    if (i >= n) continue;
    ++nq;
    const float4* rp = reinterpret_cast<const float4*>(rays + i);
    float4 a = rp[0], b = rp[1];
    f3 o = mk3(a.x, a.y, a.z), d = mk3(b.x, b.y, b.z);
    if (kAnyHit) {
      HitRec h = traverse<true>(S, o, d, fmn(tmax[i], kMax), stk, st);
      occluded[i] = h.inst != CG_MISS_ID ? 1u : 0u;
    } else {
      HitRec h = traverse<false>(S, o, d, kMax, stk, st);
      cg_hit out;
      if (h.inst != CG_MISS_ID) {
        out.t = h.t;
        out.inst = h.inst;
        out.tri = S.tris[h.slot].tri_id;
      } else {
        out.t = CG_MISS_T;
        out.inst = CG_MISS_ID;
        out.tri = CG_MISS_ID;
      }
      out._pad = 0u;
      hits[i] = out;
      if (uv) {
        uv[2 * i] = h.inst != CG_MISS_ID ? h.u : 0.0f;
        uv[2 * i + 1] = h.inst != CG_MISS_ID ? h.v : 0.0f;
      }
    }
  }
  if (stats) flush_stats(stats, kAnyHit ? 0u : nq, kAnyHit ? nq : 0u, st.nodes, st.tris, st.insts);
}

// ---------------------------------------------------------------------------
// The compute pass (compute.wesl:32-62, render.wesl:7-18, trace.wesl:22-113)
// with one lane per pixel, so each pixel consumes its RNG stream in exactly
// the reference's depth-first order.  The transport tree's open frames live
// in a private array (MF frames = trace_depth - 1); the BVH stack in LDS
// (overflow in device memory).
template <int MF>
__global__ __launch_bounds__(kBlock, 4) void k_view(ViewArgs A, WorkQueue q, StackArgs sa) {
  __shared__ uint32_t s_ref[kShortStack * kBlock];
  __shared__ float s_t[kShortStack * kBlock];
  const TStack stk = lane_stack(s_ref, s_t, sa);
  const uint32_t tiles_x = (A.W + 7u) / 8u;
  TravStats st{0u, 0u, 0u};
  uint64_t nrays = 0, nshadow = 0;
  uint32_t probe = 0;
  const uint32_t lane = threadIdx.x & 63u;
  for (;;) {
    const uint32_t tile = grab_item(q, probe);
    if (tile == kNoItem) break;
    const uint32_t px = (tile % tiles_x) * 8u + (lane & 7u);
    const uint32_t py = (tile / tiles_x) * 8u + (lane >> 3);
    if (px < A.W && py < A.H) render_pixel<MF>(A, px, py, stk, st, nrays, nshadow);
  }
  if (A.stats) flush_stats(A.stats, nrays, nshadow, st.nodes, st.tris, st.insts);
}

// Plain streaming copy: 16 B per lane per iteration, grid-stride.
__global__ __launch_bounds__(kBlock) void k_stream_copy(float4* __restrict__ dst, const float4* __restrict__ src,
                                                        size_t n) {
  size_t stride = static_cast<size_t>(gridDim.x) * kBlock;
  for (size_t i = static_cast<size_t>(blockIdx.x) * kBlock + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

template <typename T>
int upload_vec(cg_env* env, DevBuf& b, const std::vector<T>& v, uint64_t& total) {
  total += v.size() * sizeof(T);
  return dev_upload(env, b, v.data(), v.size() * sizeof(T));
}

void release_scene(cg_scene* s) {
  dev_free(s->bvh);
  dev_free(s->nodes);
  dev_free(s->tris);
  dev_free(s->insts);
  dev_free(s->vidx);
  dev_free(s->vattr);
  dev_free(s->mats);
  dev_free(s->emit);
  dev_free(s->alias);
  dev_free(s->emit_order);
  dev_free(s->env);
}

// Resident grid of a kernel: blocks per CU at full occupancy x CUs, capped by
// the number of work items.
template <typename K>
uint32_t resident_grid(K kernel, const cg_env* env, uint32_t items) {
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kBlock, 0) != hipSuccess || per_cu < 1) per_cu = 1;
  const uint64_t g = static_cast<uint64_t>(per_cu) * static_cast<uint64_t>(env->cu_count > 0 ? env->cu_count : 1);
  return static_cast<uint32_t>(std::max<uint64_t>(1, std::min<uint64_t>(g, items)));
}
WorkQueue reset_queue(cg_env* env, uint32_t items) {
  uint32_t* c = static_cast<uint32_t*>(env->queue.p);
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(kXcds * kCounterStride), 0, env->stream, c, kXcds * kCounterStride);
  return WorkQueue{c, items};
}
// Overflow stack area for a launch of `grid` blocks on a scene whose
// traversal stack is at most `depth` entries deep.
int stack_args(cg_env* env, uint32_t grid, uint32_t depth, StackArgs& sa) {
  sa.ovf = nullptr;
  sa.ovf_cap = 0;
  if (depth <= static_cast<uint32_t>(kShortStack)) return CG_OK;
  const int cap = static_cast<int>(depth) - kShortStack;
  const size_t bytes = static_cast<size_t>(grid) * kBlock * static_cast<size_t>(cap) * sizeof(uint2);
  if (env->ovf.bytes < bytes) {
    if (env->ovf.p) {
      CG_HIP(env, hipStreamSynchronize(env->stream));
      dev_free(env->ovf);
    }
    int rc = dev_alloc(env, env->ovf, bytes);
    if (rc != CG_OK) return rc;
  }
  sa.ovf = static_cast<uint2*>(env->ovf.p);
  sa.ovf_cap = cap;
  return CG_OK;
}
template <int MF>
int launch_view_mf(cg_env* env, const ViewArgs& a, uint32_t tiles, uint32_t depth) {
  const uint32_t grid = resident_grid(k_view<MF>, env, (tiles + 3u) / 4u);
  StackArgs sa;
  int rc = stack_args(env, grid, depth, sa);
  if (rc != CG_OK) return rc;
  WorkQueue q = reset_queue(env, tiles);
  hipLaunchKernelGGL((k_view<MF>), dim3(grid), dim3(kBlock), 0, env->stream, a, q, sa);
  return CG_OK;
}
int launch_view(cg_env* env, const ViewArgs& a, int frames, uint32_t tiles, uint32_t depth) {
  if (frames <= 4) return launch_view_mf<4>(env, a, tiles, depth);
  if (frames <= 8) return launch_view_mf<8>(env, a, tiles, depth);
  if (frames <= 16) return launch_view_mf<16>(env, a, tiles, depth);
  return launch_view_mf<64>(env, a, tiles, depth);
}
template <bool kAnyHit>
int launch_trace(cg_env* env, const cg_scene* scene, const cg_ray* rays, const float* tmax, uint32_t n, cg_hit* hits,
                 float* uv, uint32_t* occ, unsigned long long* st) {
  const uint32_t items = (n + 63u) / 64u;
  const uint32_t grid = resident_grid(k_trace<kAnyHit>, env, (items + 3u) / 4u);
  StackArgs sa;
  int rc = stack_args(env, grid, scene->stack_depth, sa);
  if (rc != CG_OK) return rc;
  WorkQueue q = reset_queue(env, items);
  hipLaunchKernelGGL((k_trace<kAnyHit>), dim3(grid), dim3(kBlock), 0, env->stream, scene->view, rays, tmax, n, hits, uv,
                     occ, st, q, sa);
  return CG_OK;
}

}  // namespace

// ===========================================================================
// C-ABI
extern "C" {

const char* cg_last_error(void) { return t_last_error.c_str(); }
const char* cg_env_last_error(const cg_env* env) { return env ? env->last_error.c_str() : ""; }

int cg_env_create(const cg_conf* conf, int device, cg_env** out) {
  if (!conf || !out) return fail(nullptr, CG_ERR_INVALID, "cg_env_create: null argument");
  *out = nullptr;
  if (conf->render_width == 0 || conf->render_height == 0)
    return fail(nullptr, CG_ERR_INVALID, "render size must be positive");
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count == 0)
    return fail(nullptr, CG_ERR_HIP, "no HIP device available (%s)", hipGetErrorString(e));
  if (device < 0 || device >= count) return fail(nullptr, CG_ERR_INVALID, "device %d out of range", device);
  cg_env* env = new cg_env();
  env->conf = *conf;
  env->conf.trace_depth = std::min(std::max(conf->trace_depth, 1u), 64u);  // environment.cpp:20
  env->device = device;
  int rc = CG_OK;
  auto bail = [&](int code) {
    t_last_error = env->last_error.empty() ? t_last_error : env->last_error;
    cg_env_destroy(env);
    return code;
  };
  if (hipSetDevice(device) != hipSuccess) return bail(fail(env, CG_ERR_HIP, "hipSetDevice(%d) failed", device));
  if (hipStreamCreateWithFlags(&env->stream, hipStreamNonBlocking) != hipSuccess)
    return bail(fail(env, CG_ERR_HIP, "hipStreamCreate failed"));
  const size_t px = static_cast<size_t>(conf->render_width) * conf->render_height;
  if ((rc = dev_alloc(env, env->history, px * sizeof(cg_history_pixel))) != CG_OK) return bail(rc);
  if ((rc = dev_alloc(env, env->rgba, px * 4)) != CG_OK) return bail(rc);
  if ((rc = dev_alloc(env, env->stats, 8 * sizeof(unsigned long long))) != CG_OK) return bail(rc);
  if ((rc = dev_alloc(env, env->queue, kXcds * kCounterStride * sizeof(uint32_t))) != CG_OK) return bail(rc);
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) return bail(fail(env, CG_ERR_HIP, "hipGetDeviceProperties failed"));
  env->cu_count = prop.multiProcessorCount;
  if (hipMemsetAsync(env->history.p, 0, env->history.bytes, env->stream) != hipSuccess ||
      hipMemsetAsync(env->stats.p, 0, env->stats.bytes, env->stream) != hipSuccess)
    return bail(fail(env, CG_ERR_HIP, "hipMemsetAsync failed"));
  WlTable wl;
  build_wl_table(wl);
  env->wl_avg = wl.avg_integral;
  if ((rc = dev_upload(env, env->wl_cdf, wl.cdf.data(), wl.cdf.size() * sizeof(float))) != CG_OK) return bail(rc);
  if ((rc = dev_upload(env, env->wl_leaf, wl.leaf.data(), wl.leaf.size() * sizeof(WlLeaf))) != CG_OK)
    return bail(rc);
  if (hipStreamSynchronize(env->stream) != hipSuccess) return bail(fail(env, CG_ERR_HIP, "stream sync failed"));
  *out = env;
  return CG_OK;
}

int cg_env_destroy(cg_env* env) {
  if (!env) return fail(nullptr, CG_ERR_INVALID, "cg_env_destroy: null env");
  if (env->stream) (void)hipStreamSynchronize(env->stream);
  for (cg_scene* s : env->scenes) {
    release_scene(s);
    delete s;
  }
  env->scenes.clear();
  dev_free(env->history);
  dev_free(env->rgba);
  dev_free(env->path_hash);
  dev_free(env->ray_count);
  dev_free(env->stats);
  dev_free(env->wl_cdf);
  dev_free(env->wl_leaf);
  dev_free(env->queue);
  dev_free(env->ovf);
  if (env->stream) (void)hipStreamDestroy(env->stream);
  delete env;
  return CG_OK;
}

int cg_env_sync(cg_env* env) {
  if (!env) return fail(nullptr, CG_ERR_INVALID, "null env");
  CG_HIP(env, hipStreamSynchronize(env->stream));
  return CG_OK;
}

void* cg_env_stream(cg_env* env) { return env ? static_cast<void*>(env->stream) : nullptr; }

int cg_env_set_counting(cg_env* env, int enable) {
  if (!env) return fail(nullptr, CG_ERR_INVALID, "null env");
  env->counting = enable != 0;
  return CG_OK;
}

int cg_env_set_path_records(cg_env* env, int enable) {
  if (!env) return fail(nullptr, CG_ERR_INVALID, "null env");
  env->path_records = enable != 0;
  if (env->path_records && !env->path_hash.p) {
    const size_t px = static_cast<size_t>(env->conf.render_width) * env->conf.render_height;
    int rc = dev_alloc(env, env->path_hash, px * 4);
    if (rc != CG_OK) return rc;
    rc = dev_alloc(env, env->ray_count, px * 4);
    if (rc != CG_OK) return rc;
  }
  return CG_OK;
}

int cg_env_set_accel(cg_env* env, int mode) {
  if (!env) return fail(nullptr, CG_ERR_INVALID, "null env");
  if (mode != CG_ACCEL_REFERENCE && mode != CG_ACCEL_REBUILT) return fail(env, CG_ERR_INVALID, "unknown accel mode %d", mode);
  env->accel = mode;
  return CG_OK;
}

int cg_stats_read(cg_env* env, cg_stats* out) {
  if (!env || !out) return fail(env, CG_ERR_INVALID, "null argument");
  unsigned long long v[8];
  CG_HIP(env, hipMemcpyAsync(v, env->stats.p, sizeof(v), hipMemcpyDeviceToHost, env->stream));
  CG_HIP(env, hipStreamSynchronize(env->stream));
  out->rays = v[0];
  out->shadow_rays = v[1];
  out->nodes = v[2];
  out->tris = v[3];
  out->insts = v[4];
  return CG_OK;
}

int cg_stats_reset(cg_env* env) {
  if (!env) return fail(nullptr, CG_ERR_INVALID, "null env");
  CG_HIP(env, hipMemsetAsync(env->stats.p, 0, env->stats.bytes, env->stream));
  return CG_OK;
}

int cg_scene_upload(cg_env* env, const cg_scene_data* sd, cg_scene** out) {
  if (!env || !sd || !out) return fail(env, CG_ERR_INVALID, "cg_scene_upload: null argument");
  *out = nullptr;
  Layout L;
  std::string err;
  int rc = build_layout(sd, L, err, env->accel);
  if (rc != CG_OK) return fail(env, rc, "%s", err.c_str());
  if (L.stack_depth > 4096u)
    return fail(env, CG_ERR_UNSUPPORTED, "BVH needs a %u-entry traversal stack (max 4096)", L.stack_depth);
  cg_scene* s = new cg_scene();
  uint64_t total = 0;
  auto up = [&](int r) {
    if (r != CG_OK && rc == CG_OK) rc = r;
  };
  up(upload_vec(env, s->bvh, L.bvh, total));
  up(upload_vec(env, s->nodes, L.nodes, total));
  up(upload_vec(env, s->tris, L.tris, total));
  up(upload_vec(env, s->insts, L.insts, total));
  up(upload_vec(env, s->vidx, L.vidx, total));
  up(upload_vec(env, s->vattr, L.vattr, total));
  up(upload_vec(env, s->mats, L.mats, total));
  up(upload_vec(env, s->emit, L.emit, total));
  up(upload_vec(env, s->alias, L.alias, total));
  up(upload_vec(env, s->emit_order, L.emit_order, total));
  size_t env_bytes = static_cast<size_t>(sd->env_width) * sd->env_height * 4 * sizeof(float);
  total += env_bytes;
  up(dev_upload(env, s->env, sd->env_rgba32f, env_bytes));
  if (rc == CG_OK && hipStreamSynchronize(env->stream) != hipSuccess)
    rc = fail(env, CG_ERR_HIP, "scene upload failed");
  if (rc != CG_OK) {
    release_scene(s);
    delete s;
    return rc;
  }
  DevScene& v = s->view;
  v.bvh = static_cast<const Node4*>(s->bvh.p);
  v.nodes = static_cast<const PairNode*>(s->nodes.p);
  v.tris = static_cast<const TriRec*>(s->tris.p);
  v.insts = static_cast<const InstRec*>(s->insts.p);
  v.tri_vidx = static_cast<const uint4*>(s->vidx.p);
  v.vattr = static_cast<const VAttr*>(s->vattr.p);
  v.mats = static_cast<const MatRec*>(s->mats.p);
  v.mat_count = static_cast<uint32_t>(L.mats.size());
  v.emit = static_cast<const EmitRec*>(s->emit.p);
  v.emit_count = static_cast<uint32_t>(L.emit.size());
  v.alias = static_cast<const float2*>(s->alias.p);
  v.alias_count = static_cast<uint32_t>(L.alias.size());
  v.emit_order = static_cast<const uint32_t*>(s->emit_order.p);
  v.emit_root = L.emit_root;
  v.env = static_cast<const float4*>(s->env.p);
  v.env_w = sd->env_width;
  v.env_h = sd->env_height;
  v.root = L.root;
  s->stack_depth = L.stack_depth;
  s->bytes = total;
  env->scenes.push_back(s);
  *out = s;
  return CG_OK;
}

int cg_scene_release(cg_env* env, cg_scene* scene) {
  if (!env || !scene) return fail(env, CG_ERR_INVALID, "cg_scene_release: null argument");
  auto it = std::find(env->scenes.begin(), env->scenes.end(), scene);
  if (it == env->scenes.end()) return fail(env, CG_ERR_INVALID, "scene does not belong to this env");
  CG_HIP(env, hipStreamSynchronize(env->stream));
  release_scene(scene);
  delete scene;
  env->scenes.erase(it);
  return CG_OK;
}

int cg_scene_footprint(const cg_scene* scene, uint64_t* bytes, uint32_t* stack_depth) {
  if (!scene) return fail(nullptr, CG_ERR_INVALID, "null scene");
  if (bytes) *bytes = scene->bytes;
  if (stack_depth) *stack_depth = scene->stack_depth;
  return CG_OK;
}

int cg_view(cg_env* env, cg_scene* scene, const cg_camera* camera, uint32_t iter_count) {
  if (!env || !scene || !camera) return fail(env, CG_ERR_INVALID, "cg_view: null argument");
  ViewArgs a{};
  a.S = scene->view;
  a.P.D = static_cast<int>(env->conf.trace_depth);
  a.P.T = env->conf.max_transport_sample_count;
  a.P.E = env->conf.max_emission_sample_count;
  a.P.DIF = env->conf.max_diffuse_sample_count;
  a.P.ROUGH = env->conf.max_rough_sample_count;
  a.W = env->conf.render_width;
  a.H = env->conf.render_height;
  a.spp = env->conf.sample_count;
  a.iter = iter_count;
  for (int k = 0; k < 3; ++k) {
    a.cam_pos[k] = camera->position[k];
    a.cam_dir[k] = camera->direction[k];
  }
  a.viewport[0] = camera->viewport[0];
  a.viewport[1] = camera->viewport[1];
  a.wl_cdf = static_cast<const float*>(env->wl_cdf.p);
  a.wl_leaf = static_cast<const WlLeaf*>(env->wl_leaf.p);
  a.wl_avg = env->wl_avg;
  a.history = static_cast<cg_history_pixel*>(env->history.p);
  a.rgba = static_cast<uint32_t*>(env->rgba.p);
  a.path_hash = env->path_records ? static_cast<uint32_t*>(env->path_hash.p) : nullptr;
  a.ray_count = env->path_records ? static_cast<uint32_t*>(env->ray_count.p) : nullptr;
  a.stats = env->counting ? static_cast<unsigned long long*>(env->stats.p) : nullptr;
  const uint32_t tiles = ((a.W + 7u) / 8u) * ((a.H + 7u) / 8u);
  const int frames = std::max(1, a.P.D - 1);
  int rc = launch_view(env, a, frames, tiles, scene->stack_depth);
  if (rc != CG_OK) return rc;
  CG_HIP(env, hipGetLastError());
  return CG_OK;
}

int cg_read_target(cg_env* env, cg_history_pixel* history, uint8_t* rgba8) {
  if (!env) return fail(nullptr, CG_ERR_INVALID, "null env");
  if (history) CG_HIP(env, hipMemcpyAsync(history, env->history.p, env->history.bytes, hipMemcpyDeviceToHost, env->stream));
  if (rgba8) CG_HIP(env, hipMemcpyAsync(rgba8, env->rgba.p, env->rgba.bytes, hipMemcpyDeviceToHost, env->stream));
  CG_HIP(env, hipStreamSynchronize(env->stream));
  return CG_OK;
}

int cg_read_path_records(cg_env* env, uint32_t* path_hash, uint32_t* ray_count) {
  if (!env) return fail(nullptr, CG_ERR_INVALID, "null env");
  if (!env->path_hash.p) return fail(env, CG_ERR_STATE, "path records are not enabled");
  if (path_hash)
    CG_HIP(env, hipMemcpyAsync(path_hash, env->path_hash.p, env->path_hash.bytes, hipMemcpyDeviceToHost, env->stream));
  if (ray_count)
    CG_HIP(env, hipMemcpyAsync(ray_count, env->ray_count.p, env->ray_count.bytes, hipMemcpyDeviceToHost, env->stream));
  CG_HIP(env, hipStreamSynchronize(env->stream));
  return CG_OK;
}

int cg_write_history(cg_env* env, const cg_history_pixel* history) {
  if (!env || !history) return fail(env, CG_ERR_INVALID, "null argument");
  CG_HIP(env, hipMemcpyAsync(env->history.p, history, env->history.bytes, hipMemcpyHostToDevice, env->stream));
  CG_HIP(env, hipStreamSynchronize(env->stream));
  return CG_OK;
}

int cg_trace_closest(cg_env* env, cg_scene* scene, const cg_ray* rays, uint32_t n, cg_hit* hits, float* uv) {
  if (!env || !scene || (!rays && n) || (!hits && n)) return fail(env, CG_ERR_INVALID, "cg_trace_closest: null argument");
  if (n == 0) return CG_OK;
  unsigned long long* st = env->counting ? static_cast<unsigned long long*>(env->stats.p) : nullptr;
  int rc = launch_trace<false>(env, scene, rays, nullptr, n, hits, uv, nullptr, st);
  if (rc != CG_OK) return rc;
  CG_HIP(env, hipGetLastError());
  return CG_OK;
}

int cg_trace_occluded(cg_env* env, cg_scene* scene, const cg_ray* rays, const float* tmax, uint32_t n,
                      uint32_t* occluded) {
  if (!env || !scene || (n && (!rays || !tmax || !occluded)))
    return fail(env, CG_ERR_INVALID, "cg_trace_occluded: null argument");
  if (n == 0) return CG_OK;
  unsigned long long* st = env->counting ? static_cast<unsigned long long*>(env->stats.p) : nullptr;
  int rc = launch_trace<true>(env, scene, rays, tmax, n, nullptr, nullptr, occluded, st);
  if (rc != CG_OK) return rc;
  CG_HIP(env, hipGetLastError());
  return CG_OK;
}

int cg_stream_copy(cg_env* env, void* dst, const void* src, size_t bytes) {
  if (!env || !dst || !src) return fail(env, CG_ERR_INVALID, "cg_stream_copy: null argument");
  if (bytes % 16 != 0) return fail(env, CG_ERR_INVALID, "cg_stream_copy: size must be a multiple of 16");
  const size_t n = bytes / 16;
  if (n == 0) return CG_OK;
  const size_t blocks = std::min<size_t>((n + kBlock - 1) / kBlock, 256u * 64u);
  hipLaunchKernelGGL(k_stream_copy, dim3(static_cast<uint32_t>(blocks)), dim3(kBlock), 0, env->stream,
                     static_cast<float4*>(dst), static_cast<const float4*>(src), n);
  CG_HIP(env, hipGetLastError());
  return CG_OK;
}

int cg_device_alloc(cg_env* env, size_t bytes, void** dptr) {
  if (!env || !dptr) return fail(env, CG_ERR_INVALID, "cg_device_alloc: null argument");
  *dptr = nullptr;
  if (bytes == 0) return fail(env, CG_ERR_INVALID, "cg_device_alloc: zero bytes");
  CG_HIP(env, hipMalloc(dptr, bytes));
  return CG_OK;
}

int cg_device_free(cg_env* env, void* dptr) {
  if (!env) return fail(nullptr, CG_ERR_INVALID, "null env");
  if (!dptr) return CG_OK;
  CG_HIP(env, hipStreamSynchronize(env->stream));
  CG_HIP(env, hipFree(dptr));
  return CG_OK;
}

int cg_memcpy_h2d(cg_env* env, void* dst, const void* src, size_t bytes) {
  if (!env || (bytes && (!dst || !src))) return fail(env, CG_ERR_INVALID, "cg_memcpy_h2d: null argument");
  if (!bytes) return CG_OK;
  CG_HIP(env, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, env->stream));
  CG_HIP(env, hipStreamSynchronize(env->stream));
  return CG_OK;
}

int cg_memcpy_d2h(cg_env* env, void* dst, const void* src, size_t bytes) {
  if (!env || (bytes && (!dst || !src))) return fail(env, CG_ERR_INVALID, "cg_memcpy_d2h: null argument");
  if (!bytes) return CG_OK;
  CG_HIP(env, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, env->stream));
  CG_HIP(env, hipStreamSynchronize(env->stream));
  return CG_OK;
}

}  // extern "C"
