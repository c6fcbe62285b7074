// cg_device.hip — HIP backend of the spectral path tracer for MI355X
// (gfx950): scene upload into the device layout, the per-pixel compute pass
// and the batch traversal entry points behind the C-ABI of cg_device.h.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../../include/cg_device.h"
#include "cg_core.h"
#include "cg_layout.h"
#include "cg_render.h"
#include "cg_upload.h"
#include "cg_wavefront.h"

using namespace cg;

// Fixed device scratch of cg_trace_closest_noalloc (32 MiB per device).
constexpr size_t kScratchBytes = size_t{32} << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

thread_local std::string t_last_error;

// ---------------------------------------------------------------------------
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
};

}  // namespace

// A BLAS built on the device, kept for refits: its triangles' sorted order,
// the binary tree, and each wide node's binary children.
struct BlasState {
  DevBuf ids, left, right, parent, kids;
  uint32_t n = 0, slot_base = 0, node_base = 0, nodes = 0, stack = 0;
};

struct cg_scene {
  DevBuf bvh, nodes, tris, insts, vidx, vattr, mats, emit, alias, emit_order, env, tex;
  DevScene view{};
  uint32_t stack_depth = 0;
  uint64_t bytes = 0;
  bool device_built = false;      // uploaded with CG_ACCEL_DEVICE: BLAS nodes first, then the TLAS
  std::vector<BlasState> blas;
  uint32_t blas_nodes = 0;        // BLAS nodes before the TLAS's
  uint64_t bvh_capacity = 0;      // nodes the bvh buffer holds
};

struct cg_vbatch {
  cg_env* env = nullptr;
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  DevBuf jobs, first, queue, ovf;
  DevBuf jobs_s, first_s, pfirst_s, samples;  // the launch's sample-stream jobs
  std::vector<ViewArgs> host_jobs;  // sources of the in-flight copies: kept until reuse
  std::vector<uint32_t> host_first;
  std::vector<ViewArgs> host_jobs_s;
  std::vector<uint32_t> host_first_s, host_pfirst_s;
  hipGraphExec_t exec = nullptr;  // the last launch's graph (counter reset + kernel)
  bool in_flight = false;
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
  std::vector<cg_vbatch*> vbatch_pool;  // finished view batches, for reuse
  std::vector<cg_vbatch*> vbatch_all;
  void* scratch = nullptr;  // this device's copy of `scratch` (hipGetSymbolAddress)
  DevBuf errflag;           // device error flag: a traversal stack overflowed
  DevBuf samples;           // sample-stream mode: per-sample results of the last render
  // Work schedules of the per-pixel renders ([0] reference order, [1]
  // sample-stream mode): item durations, the next order, the sort's buffers.
  DevBuf sched[2];
  uint32_t sched_items[2] = {0u, 0u};  // work items the buffers are sized for
  size_t sched_tmp[2] = {0u, 0u};      // the sort's temporary bytes
  bool sched_ready[2] = {false, false};  // the order holds a complete render's durations
  bool sched_on = true;                  // cg_env_set_schedule
  uint32_t rng_mode = CG_RNG_REFERENCE;
  int test_stack_cap = -1;  // CG_TEST_STACK_CAP: caps overflow entries (tests force an overflow)
  std::vector<cg_batch*> batches;  // live wavefront batches, destroyed with the env
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
#ifndef CG_VB_LPT
#define CG_VB_LPT 1
#endif
constexpr uint32_t kNoItem = 0xFFFFFFFFu;
constexpr uint32_t kCounterStride = 32u;  // one 128-byte line per counter
struct WorkQueue {
  uint32_t* counters;  // kXcds counters, kCounterStride apart, zero at launch
  uint32_t total;      // items
  uint32_t interleave; // 0: XCD x owns a contiguous range; 1: XCD x owns items x, x + 8, x + 16, ...
};
__device__ __forceinline__ uint32_t grab_item(const WorkQueue& q, uint32_t& probe) {
  uint32_t item = kNoItem;
  if ((threadIdx.x & 63u) == 0u && q.interleave) {
    const uint32_t home = blockIdx.x % kXcds;
    while (probe < kXcds) {
      const uint32_t x = (home + probe) % kXcds;
      uint32_t* ctr = q.counters + x * kCounterStride;
      if (x + uint64_t{kXcds} * __hip_atomic_load(ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < q.total) {
        const uint64_t i = x + uint64_t{kXcds} * atomicAdd(ctr, 1u);
        if (i < q.total) {
          item = static_cast<uint32_t>(i);
          break;
        }
      }
      ++probe;
    }
  } else if ((threadIdx.x & 63u) == 0u) {
    const uint32_t per = (q.total + kXcds - 1u) / kXcds;
    const uint32_t home = blockIdx.x % kXcds;
    while (probe < kXcds) {
      const uint32_t x = (home + probe) % kXcds;
      const uint32_t begin = x * per;
      const uint32_t end = begin + per < q.total ? begin + per : q.total;
      uint32_t* ctr = q.counters + x * kCounterStride;
      // the counter only grows: a plain coherent load rules out an exhausted
      // range without a read-modify-write that would queue on its address
      if (begin < end && begin + __hip_atomic_load(ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < end) {
        const uint32_t c = atomicAdd(ctr, 1u);
        if (begin + c < end) {
          item = begin + c;
          break;
        }
      }
      ++probe;
    }
  }
  // Lane 0 is active at every call (it takes the item), so it is the first
  // active lane: the item becomes a wave-uniform (scalar) value.
  return __builtin_amdgcn_readfirstlane(item);
}

__global__ void k_zero_u32(uint32_t* p, uint32_t n) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

__device__ __forceinline__ uint64_t wave_sum(uint64_t v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
// Traversal counters: kStatCopies copies of 8 words, one chosen per wave, so
// the adds of thousands of waves do not queue on one address; readers sum.
constexpr uint32_t kStatCopies = 64u;
constexpr size_t kStatBytes = kStatCopies * 8 * sizeof(unsigned long long);
__device__ __forceinline__ void flush_stats(unsigned long long* stats, uint64_t rays, uint64_t shadows,
                                            uint64_t nodes, uint64_t tris, uint64_t insts) {
  rays = wave_sum(rays);
  shadows = wave_sum(shadows);
  nodes = wave_sum(nodes);
  tris = wave_sum(tris);
  insts = wave_sum(insts);
  if ((rays | shadows | nodes | tris | insts) == 0u) return;
  if ((threadIdx.x & 63u) == 0u) {
    unsigned long long* c = stats + 8u * ((blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) % kStatCopies);
    if (rays) atomicAdd(c + 0, static_cast<unsigned long long>(rays));
    if (shadows) atomicAdd(c + 1, static_cast<unsigned long long>(shadows));
    if (nodes) atomicAdd(c + 2, static_cast<unsigned long long>(nodes));
    if (tris) atomicAdd(c + 3, static_cast<unsigned long long>(tris));
    if (insts) atomicAdd(c + 4, static_cast<unsigned long long>(insts));
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
  uint32_t* err;  // the env's device error flag (stack overflow)
};
// The stack of this lane: `short_n` entries in the block's LDS arrays, the
// rest in the overflow area interleaved over the whole grid.
template <int BLOCK>
__device__ __forceinline__ TStack lane_stack(uint2* s_stk, const StackArgs& sa, int short_n) {
  const uint32_t g = blockIdx.x * BLOCK + threadIdx.x;
  return TStack{s_stk + threadIdx.x, BLOCK, short_n, sa.ovf ? sa.ovf + g : nullptr,
                gridDim.x * BLOCK, sa.ovf ? sa.ovf_cap : 0, sa.err};
}

// Persistent traversal of queued rays with lane refill: a wave takes rays
// 64 at a time from the work queue and hands them to its idle lanes; each
// lane advances its own ray one visit per step, and once a quarter of the
// wave is idle the idle lanes take the next rays.  Lanes stay busy instead of
// waiting for the slowest ray of the wave.  fetch(i, scene, o, d, tmax, any)
// sets up ray i; done(i, ray) stores its result.
// LDS stack entries per lane in the one-job per-pixel kernel.  One frame is
// bound by its longest pixel chains (latency): more registers (3 waves per
// SIMD) and a deeper LDS stack (24 x 8 B x 256 lanes, three blocks per CU)
// shorten them.  The many-job kernel is bound by throughput and keeps 4 waves
// per SIMD and 16 entries.
constexpr int kViewShortStack = 24;
// LDS stack entries per lane of the sample-stream kernels (k_view_s,
// k_view_batch_s): their chains are short and their throughput needs waves.
// They run 5 waves per SIMD (96 VGPRs, some spills: 5 waves beat 4 by 3 % on
// one frame and 7 % on the 64-job mix; 3 waves lose 18 %); 12 entries (24 KB
// per block, five blocks per CU) beat 16, 10 and 8 on the 64-job mix and the
// serving mix.
#ifndef CG_SAMPLE_SHORT
#define CG_SAMPLE_SHORT 12
#endif
constexpr int kSampleShortStack = CG_SAMPLE_SHORT;
constexpr int kRefillLanes = 16;
constexpr int kStepsPerRefill = 4;
template <int kMode, class Fetch, class Done>
__device__ __forceinline__ void trace_refill(const WorkQueue& q, uint32_t n, const TStack& stk, TravStats& st,
                                             Fetch fetch, Done done) {
  uint32_t probe = 0;
  uint32_t cbeg = 0, cend = 0;  // rays of the current chunk not yet handed out (wave-uniform)
  bool more = true;             // the queue may hold more chunks (wave-uniform)
  bool has = false;
  uint32_t my = 0;
  TravScene ts{};
  TravRay r;
  for (;;) {
    if (more && __popcll(__ballot(!has)) >= kRefillLanes) {
      for (;;) {
        const unsigned long long idle = __ballot(!has);
        if (idle == 0ull) break;
        if (cbeg == cend) {
          const uint32_t item = grab_item(q, probe);
          if (item == kNoItem) {
            more = false;
            break;
          }
          cbeg = item * 64u;
          cend = cbeg + 64u < n ? cbeg + 64u : n;
          continue;
        }
        const uint32_t rank = __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(idle >> 32),
                                                        __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(idle), 0u));
        const uint32_t avail = cend - cbeg;
        if (!has && rank < avail) {
          my = cbeg + rank;
          f3 o, d;
          float tmax;
          bool any;
          fetch(my, ts, o, d, tmax, any);
          if (trav_begin(ts, r, o, d, tmax, any))
            has = true;
          else
            done(my, r);
        }
        const uint32_t taken = static_cast<uint32_t>(__popcll(idle));
        cbeg += avail < taken ? avail : taken;
      }
    }
    if (!more && __ballot(has) == 0ull) break;
    for (int k = 0; k < kStepsPerRefill; ++k) {
      if (has && trav_step<kMode>(ts, r, stk, st)) {
        done(my, r);
        has = false;
      }
    }
  }
}

template <bool kAnyHit>
__global__ __launch_bounds__(kBlock, 5) void k_trace(DevScene S, const cg_ray* __restrict__ rays,
                                                  const float* __restrict__ tmax, uint32_t n,
                                                  cg_hit* __restrict__ hits, float* __restrict__ uv,
                                                  uint32_t* __restrict__ occluded,
                                                  unsigned long long* __restrict__ stats, WorkQueue q, StackArgs sa) {
  __shared__ uint2 s_stk[kShortStack * kBlock];
  const TStack stk = lane_stack<kBlock>(s_stk, sa, kShortStack);
  const TravScene scene = trav_scene(S);
  TravStats st{0u, 0u, 0u};
  uint64_t nq = 0;
  auto fetch = [&](uint32_t i, TravScene& ts, f3& o, f3& d, float& tm, bool& any) {
    const float4* rp = reinterpret_cast<const float4*>(rays + i);
    const float4 a = rp[0], b = rp[1];
    o = mk3(a.x, a.y, a.z);
    d = mk3(b.x, b.y, b.z);
    tm = kAnyHit ? fmn(tmax[i], kMax) : kMax;
    any = kAnyHit;
    ts = scene;
  };
  auto done = [&](uint32_t i, const TravRay& r) {
    ++nq;
    const HitRec& h = r.best;
    if (kAnyHit) {
      occluded[i] = h.inst != CG_MISS_ID ? 1u : 0u;
      return;
    }
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
  };
  trace_refill<kAnyHit ? 1 : 0>(q, n, stk, st, fetch, done);
  if (stats) flush_stats(stats, kAnyHit ? 0u : nq, kAnyHit ? nq : 0u, st.nodes, st.tris, st.insts);
}

// An item's duration (wall-clock ticks) as a sort key: four steps per octave,
// at most 127, so one 8-bit radix pass orders the items and items of one step
// keep their image order.
__device__ __forceinline__ uint32_t duration_key(uint64_t ticks) {
  const uint32_t t = ticks > 0xFFFFFFFFull ? 0xFFFFFFFFu : static_cast<uint32_t>(ticks);
  if (t < 4u) return t;
  const uint32_t e = 31u - static_cast<uint32_t>(__clz(t));
  return (e << 2) | ((t >> (e - 2u)) & 3u);
}
// The position in the longest-first order that interleaved queue slot s
// takes: XCD x's slots x, x + 8, x + 16, ... walk runs of kDealRun
// consecutive positions, so each XCD takes neighbouring items of one duration
// step (mostly one tile's samples) together and its L2 sees them together.
constexpr uint32_t kDealRun = 16u;
__device__ __forceinline__ uint32_t dealt(uint32_t s, uint32_t total) {
  constexpr uint32_t g = kXcds * kDealRun;
  if (s >= total / g * g) return s;  // the last partial group in queue order
  return s / g * g + (s % kXcds) * kDealRun + (s % g) / kXcds;
}
// ---------------------------------------------------------------------------
// The compute pass (compute.wesl:32-62, render.wesl:7-18, trace.wesl:22-113)
// with one lane per pixel, so each pixel consumes its RNG stream in exactly
// the reference's depth-first order.  The transport tree's open frames live
// in a private array (MF frames = trace_depth - 1); the BVH stack in LDS
// (overflow in device memory).
template <int MF>
__global__ __launch_bounds__(kBlock, 3) void k_view(ViewArgs A, WorkQueue q, StackArgs sa,
                                                    const uint32_t* __restrict__ order,
                                                    uint32_t* __restrict__ duration) {
  __shared__ uint2 s_stk[kViewShortStack * kBlock];
  smits_lds_init();
  const TStack stk = lane_stack<kBlock>(s_stk, sa, kViewShortStack);
  const uint32_t tiles_x = (A.W + 7u) / 8u;
  TravStats st{0u, 0u, 0u};
  uint32_t nrays = 0, nshadow = 0;  // per lane; summed in 64 bits at the flush
  uint32_t probe = 0;
  const uint32_t lane = threadIdx.x & 63u;
  for (;;) {
    const uint32_t slot = grab_item(q, probe);
    if (slot == kNoItem) break;
    const uint32_t tile = order ? order[dealt(slot, q.total)] : slot;
    const uint64_t t0 = wall_clock64();
    const uint32_t px = (tile % tiles_x) * 8u + (lane & 7u);
    const uint32_t py = (tile / tiles_x) * 8u + (lane >> 3);
    if (px < A.W && py < A.H) render_pixel<MF>(A, px, py, stk, st, nrays, nshadow);
    if (duration && lane == 0u) duration[tile] = duration_key(static_cast<uint64_t>(wall_clock64()) - t0);
  }
  if (A.stats) flush_stats(A.stats, nrays, nshadow, st.nodes, st.tris, st.insts);
}

// The lane task of a sample-stream work item: item c of a tile covers lane
// tasks t = 64c .. 64c + 63 of the tile's 64 x nchunks (pixel, chunk)
// pairs, taken pixel-major, so a wave holds few pixels and many of their
// samples (at 16 spp, a 2x2 block x 16 samples): one pixel's samples share
// the camera ray and the first surface, and their paths are alike in length.
// Pixels are in Morton order within the tile.  The image does not depend on
// the mapping (every sample draws from its own stream).
struct SampleTask {
  uint32_t px, py, s0, s1;
};
__device__ __forceinline__ SampleTask sample_task(uint32_t tile, uint32_t c, uint32_t lane, uint32_t tiles_x,
                                                  uint32_t nchunks, uint32_t chunk, uint32_t spp) {
  const uint32_t t = c * 64u + lane;
  const uint32_t p = t / nchunks, q = t - p * nchunks;
  const uint32_t mx = (p & 1u) | ((p >> 1) & 2u) | ((p >> 2) & 4u);
  const uint32_t my = ((p >> 1) & 1u) | ((p >> 2) & 2u) | ((p >> 3) & 4u);
  const uint32_t s0 = q * chunk;
  return SampleTask{(tile % tiles_x) * 8u + mx, (tile / tiles_x) * 8u + my, s0, s0 + chunk < spp ? s0 + chunk : spp};
}

// Sample-stream mode: a tile's 64 pixels x nchunks chunks of `chunk` samples
// are nchunks work items of 64 lane tasks (sample_task), so a pixel's samples
// run on separate lanes and no lane carries a long chain of dependent
// queries; the per-sample results go to B and k_view_finish sums them.
template <int MF>
__global__ __launch_bounds__(kBlock, 5) void k_view_s(ViewArgs A, SampleBufs B, uint32_t chunk, uint32_t nchunks,
                                                      WorkQueue q, StackArgs sa, const uint32_t* __restrict__ order,
                                                      uint32_t* __restrict__ duration) {
  __shared__ uint2 s_stk[kSampleShortStack * kBlock];
  smits_lds_init();
  const TStack stk = lane_stack<kBlock>(s_stk, sa, kSampleShortStack);
  const uint32_t tiles_x = (A.W + 7u) / 8u;
  TravStats st{0u, 0u, 0u};
  uint32_t nrays = 0, nshadow = 0;
  uint32_t probe = 0;
  const uint32_t lane = threadIdx.x & 63u;
  for (;;) {
    const uint32_t slot = grab_item(q, probe);
    if (slot == kNoItem) break;
    const uint32_t item = order ? order[dealt(slot, q.total)] : slot;
    const uint64_t t0 = wall_clock64();
    const uint32_t tile = item / nchunks, c = item - tile * nchunks;
    const SampleTask k = sample_task(tile, c, lane, tiles_x, nchunks, chunk, A.spp);
    if (k.px < A.W && k.py < A.H) render_samples<MF>(A, B, k.px, k.py, k.s0, k.s1, stk, st, nrays, nshadow);
    if (duration && lane == 0u) duration[item] = duration_key(static_cast<uint64_t>(wall_clock64()) - t0);
  }
  if (A.stats) flush_stats(A.stats, nrays, nshadow, st.nodes, st.tris, st.insts);
}
__global__ __launch_bounds__(kBlock) void k_view_finish(ViewArgs A, SampleBufs B) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i < A.W * A.H) finish_pixel(A, B, i);
}

// Many render jobs in one persistent launch of the same per-pixel kernel:
// the work items are the 8x8 tiles of every job, job j owning items
// [first[j], first[j+1]).  A wave finding a job's heavy tiles no longer leaves
// the rest of the GPU idle: other jobs' tiles fill it.
template <int MF>
__global__ __launch_bounds__(kBlock, 4) void k_view_batch(const ViewArgs* __restrict__ jobs,
                                                          const uint32_t* __restrict__ first, uint32_t njobs,
                                                          WorkQueue q, StackArgs sa,
                                                          unsigned long long* __restrict__ stats) {
  __shared__ uint2 s_stk[kShortStack * kBlock];
  smits_lds_init();
  const TStack stk = lane_stack<kBlock>(s_stk, sa, kShortStack);
  TravStats st{0u, 0u, 0u};
  uint32_t nrays = 0, nshadow = 0;  // per lane; summed in 64 bits at the flush
  uint32_t probe = 0;
  const uint32_t lane = threadIdx.x & 63u;
  for (;;) {
    const uint32_t item = grab_item(q, probe);
    if (item == kNoItem) break;
    uint32_t lo = 0, hi = njobs;  // the job with first[lo] <= item < first[lo + 1]
    while (hi - lo > 1u) {
      const uint32_t mid = (lo + hi) >> 1;
      if (first[mid] <= item)
        lo = mid;
      else
        hi = mid;
    }
    const ViewArgs& A = jobs[lo];
    const uint32_t tile = item - first[lo];
    const uint32_t tiles_x = (A.W + 7u) / 8u;
    const uint32_t px = (tile % tiles_x) * 8u + (lane & 7u);
    const uint32_t py = (tile / tiles_x) * 8u + (lane >> 3);
    if (px < A.W && py < A.H) render_pixel<MF>(A, px, py, stk, st, nrays, nshadow);
  }
  if (stats) flush_stats(stats, nrays, nshadow, st.nodes, st.tris, st.insts);
}

// The job of work item `item`: first[lo] <= item < first[lo + 1].
__device__ __forceinline__ uint32_t job_of_item(const uint32_t* __restrict__ first, uint32_t njobs, uint32_t item) {
  uint32_t lo = 0, hi = njobs;
  while (hi - lo > 1u) {
    const uint32_t mid = (lo + hi) >> 1;
    if (first[mid] <= item)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}
// Sample-stream jobs of a many-job launch: the items are every job's (tile,
// sample chunk) pairs; k_view_batch_finish then sums each pixel's samples.
template <int MF>
__global__ __launch_bounds__(kBlock, 5) void k_view_batch_s(const ViewArgs* __restrict__ jobs,
                                                            const uint32_t* __restrict__ first, uint32_t njobs,
                                                            WorkQueue q, StackArgs sa,
                                                            unsigned long long* __restrict__ stats) {
  __shared__ uint2 s_stk[kSampleShortStack * kBlock];
  smits_lds_init();
  const TStack stk = lane_stack<kBlock>(s_stk, sa, kSampleShortStack);
  TravStats st{0u, 0u, 0u};
  uint32_t nrays = 0, nshadow = 0;
  uint32_t probe = 0;
  const uint32_t lane = threadIdx.x & 63u;
  for (;;) {
    const uint32_t item = grab_item(q, probe);
    if (item == kNoItem) break;
    const uint32_t j = job_of_item(first, njobs, item);
    const ViewArgs& A = jobs[j];
    const uint32_t local = item - first[j];
    const uint32_t tile = local / A.nchunks, c = local - tile * A.nchunks;
    const SampleTask k = sample_task(tile, c, lane, (A.W + 7u) / 8u, A.nchunks, A.chunk, A.spp);
    if (k.px < A.W && k.py < A.H)
      render_samples<MF>(A, SampleBufs{A.sxyz, A.srec}, k.px, k.py, k.s0, k.s1, stk, st, nrays, nshadow);
  }
  if (stats) flush_stats(stats, nrays, nshadow, st.nodes, st.tris, st.insts);
}
__global__ __launch_bounds__(kBlock) void k_view_batch_finish(const ViewArgs* __restrict__ jobs,
                                                              const uint32_t* __restrict__ pfirst, uint32_t njobs,
                                                              uint32_t npix) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= npix) return;
  const uint32_t j = job_of_item(pfirst, njobs, i);
  const ViewArgs& A = jobs[j];
  finish_pixel(A, SampleBufs{A.sxyz, A.srec}, i - pfirst[j]);
}

// Linear-BVH build of one BLAS (cg_lbvh.h): one thread per triangle, per
// internal node, or per node of the current collapse level.
__global__ __launch_bounds__(kBlock) void k_lbvh_morton(LbvhArrays A) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i < A.n) lbvh_morton(A, i);
}
__global__ __launch_bounds__(kBlock) void k_lbvh_nodes(LbvhArrays A) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i + 1u < A.n) lbvh_node(A, i);
}
__global__ __launch_bounds__(kBlock) void k_lbvh_bounds(LbvhArrays A) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i < A.n) lbvh_bounds(A, i);
}
__global__ __launch_bounds__(kBlock) void k_lbvh_collapse(LbvhArrays A, uint32_t count) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i < count) lbvh_collapse(A, A.front_in[i]);
}
__global__ __launch_bounds__(kBlock) void k_lbvh_refit(LbvhArrays A, uint32_t nodes) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i < nodes) lbvh_refit_wide(A, i);
}
__global__ __launch_bounds__(kBlock) void k_lbvh_gather(const TriRec* __restrict__ tin, const uint4* __restrict__ vin,
                                                        TriRec* __restrict__ tout, uint4* __restrict__ vout,
                                                        const uint32_t* __restrict__ ids, uint32_t n) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) {
    tout[i] = tin[ids[i]];
    vout[i] = vin[ids[i]];
  }
}
__global__ __launch_bounds__(kBlock) void k_fill_u32(uint32_t* p, uint32_t v, uint32_t n) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) p[i] = v;
}

// Plain streaming copy: 16 B per lane per iteration, grid-stride.
__global__ __launch_bounds__(kBlock) void k_stream_copy(float4* __restrict__ dst, const float4* __restrict__ src,
                                                        size_t n) {
  size_t stride = static_cast<size_t>(gridDim.x) * kBlock;
  for (size_t i = static_cast<size_t>(blockIdx.x) * kBlock + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

// Alias-table picks for given uniform pairs (the check of the emitter
// indices next-event estimation draws).
__global__ __launch_bounds__(kBlock) void k_alias_pick(DevScene S, const float2* __restrict__ u, uint32_t n,
                                                       uint32_t* __restrict__ out) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) {
    const float2 v = u[i];
    out[i] = alias_pick(S, v.x, v.y);
  }
}

// Read-only probe: every 16-byte vector of src is loaded once (four loads in
// flight per lane per step), folded into a per-lane XOR and one word per
// block is written, so the kernel's traffic is the read stream alone.
__global__ __launch_bounds__(kBlock) void k_stream_read(const uint4* __restrict__ src, size_t n,
                                                        uint32_t* __restrict__ sink) {
  const size_t stride = static_cast<size_t>(gridDim.x) * kBlock;
  size_t i = static_cast<size_t>(blockIdx.x) * kBlock + threadIdx.x;
  uint32_t acc = 0u;
  for (; i + 3 * stride < n; i += 4 * stride) {
    const uint4 a = src[i], b = src[i + stride], c = src[i + 2 * stride], d = src[i + 3 * stride];
    acc ^= (a.x ^ a.y ^ a.z ^ a.w) ^ (b.x ^ b.y ^ b.z ^ b.w) ^ (c.x ^ c.y ^ c.z ^ c.w) ^ (d.x ^ d.y ^ d.z ^ d.w);
  }
  for (; i < n; i += stride) {
    const uint4 a = src[i];
    acc ^= a.x ^ a.y ^ a.z ^ a.w;
  }
  for (int off = 32; off > 0; off >>= 1) acc ^= __shfl_xor(acc, off, 64);
  __shared__ uint32_t s_acc[kBlock / 64];
  if ((threadIdx.x & 63u) == 0u) s_acc[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t v = 0u;
    for (int w = 0; w < kBlock / 64; ++w) v ^= s_acc[w];
    sink[blockIdx.x] = v;
  }
}

// ---------------------------------------------------------------------------
// Wavefront rounds (cg_wavefront.h).  Queues hold slot ids; a wave appends
// the slots of its lanes that continue with one atomic (ballot, popcount,
// mbcnt prefix), so 64 neighbouring slots stay neighbours in the next queue.
__device__ __forceinline__ void wave_push(uint32_t* queue, uint32_t* len, bool push, uint32_t value) {
  const unsigned long long m = __ballot(push);
  if (m == 0ull) return;
  uint32_t base = 0;
  if ((threadIdx.x & 63u) == 0u) base = atomicAdd(len, static_cast<uint32_t>(__popcll(m)));
  base = __shfl(base, 0);
  const uint32_t below =
      __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(m >> 32), __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(m), 0u));
  if (push) queue[base + below] = value;
}

// Admission: initialise the slots of job j (tile-major) and queue them on
// queue 0.  One lane per slot; lanes past the picture only vote.
__global__ __launch_bounds__(kBlock) void k_wf_admit(WfTables T, uint32_t j, uint32_t nslots) {
  const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
  const WfJob& J = T.jobs[j];
  const uint32_t slot = J.slot_base + t;
  bool push = false;
  if (t < nslots) {
    uint32_t px, py;
    slot_pixel(J, t, px, py);
    if (px < J.W && py < J.H) {
      T.slot_job[slot] = j;
      const uint32_t hidx = px + py * J.W;
      WfState s;
      s.hash = 2166136261u;
      s.acc = mk3(0.0f, 0.0f, 0.0f);
      s.sample = 0u;
      s.queries = 0u;
      s.rng = J.iter > 0u ? J.history[hidx].rand_seed : hidx;
      rng_step(s.rng);
      if (J.spp > 0u) {
        wf_start_sample(T, J, px, py, s);
        wf_store(T.path[slot], s);
        push = true;
      } else {
        wf_finish(J, px, py, s.acc, s.rng, s.hash, 0u);
        atomicSub(T.job_left + j, 1u);
      }
    }
  }
  wave_push(T.queue[0], T.counters + kWfQueueLen, push, slot);
}

// Trace the ray in flight of every slot on queue q.  BLOCK threads per
// workgroup, SHORT LDS stack entries per lane (autotuned).
template <int BLOCK, int SHORT>
__global__ __launch_bounds__(BLOCK, 5) void k_wf_trace(WfTables T, uint32_t q, StackArgs sa) {
  __shared__ uint2 s_stk[SHORT * BLOCK];
  const TStack stk = lane_stack<BLOCK>(s_stk, sa, SHORT);
  const uint32_t n = T.counters[kWfQueueLen + 32u * q];
  const WorkQueue wq{T.counters + kWfTraceWork, (n + 63u) / 64u};
  const uint32_t* queue = T.queue[q];
  TravStats st{0u, 0u, 0u};
  uint64_t nray = 0, nshadow = 0;
  uint32_t slot = 0;
  auto fetch = [&](uint32_t i, TravScene& ts, f3& o, f3& d, float& tm, bool& any) {
    slot = queue[i];
    ts = trav_scene(T.scenes[T.jobs[T.slot_job[slot]].scene]);
    const float4* rp = reinterpret_cast<const float4*>(T.path + slot);
    const float4 a = rp[0], b = rp[1];
    o = mk3(a.x, a.y, a.z);
    d = mk3(b.x, b.y, b.z);
    tm = a.w;
    any = (__float_as_uint(b.w) & 1u) != 0u;
  };
  auto done = [&](uint32_t, const TravRay& r) {
    if (r.any)
      ++nshadow;
    else
      ++nray;
    const HitRec& h = r.best;
    float4* hp = reinterpret_cast<float4*>(T.hits + slot);
    hp[0] = make_float4(h.t, h.u, h.v, __uint_as_float(h.inst));
    hp[1] = make_float4(__uint_as_float(h.slot), 0.0f, 0.0f, 0.0f);
  };
  trace_refill<2>(wq, n, stk, st, fetch, done);
  if (T.stats) flush_stats(T.stats, nray, nshadow, st.nodes, st.tris, st.insts);
}

// Advance every slot on queue q by one event; continuing slots go to q ^ 1.
// 3 waves/SIMD: the shading code fits its registers without spilling (4
// waves spill, and the 64-job mix runs 5 % slower).
template <int BLOCK, int SHORT>
__global__ __launch_bounds__(BLOCK, 3) void k_wf_shade(WfTables T, uint32_t q, StackArgs sa) {
  __shared__ uint2 s_stk[SHORT * BLOCK];
  smits_lds_init();
  const TStack stk = lane_stack<BLOCK>(s_stk, sa, SHORT);
  const uint32_t n = T.counters[kWfQueueLen + 32u * q];
  const WorkQueue wq{T.counters + kWfShadeWork, (n + 63u) / 64u};
  const uint32_t* queue = T.queue[q];
  uint32_t probe = 0;
  for (;;) {
    const uint32_t item = grab_item(wq, probe);
    if (item == kNoItem) break;
    const uint32_t i = item * 64u + (threadIdx.x & 63u);
    bool push = false, finished = false;
    uint32_t slot = 0u, j = 0u;
    if (i < n) {
      slot = queue[i];
      j = T.slot_job[slot];
      const WfJob& J = T.jobs[j];
      const DevScene& S = T.scenes[J.scene];
      uint32_t px, py;
      slot_pixel(J, slot - J.slot_base, px, py);
      WfState s = wf_load(T.path[slot]);
      const HitW hw = T.hits[slot];
      push = wf_step(T, J, S, slot, px, py, s, hw, stk);
      if (push)
        wf_store(T.path[slot], s);
      else
        finished = true;
    }
    // finished slots leave their jobs: one subtraction per job among the
    // wave's lanes (neighbouring slots almost always share a job)
    for (unsigned long long m = __ballot(finished); m != 0ull;) {
      const uint32_t lead = static_cast<uint32_t>(__ffsll(static_cast<long long>(m)) - 1);
      const uint32_t lj = __shfl(j, static_cast<int>(lead));
      const unsigned long long same = __ballot(finished && j == lj);
      if ((threadIdx.x & 63u) == lead) atomicSub(T.job_left + lj, static_cast<uint32_t>(__popcll(same)));
      m &= ~same;
    }
    wave_push(T.queue[q ^ 1u], T.counters + kWfQueueLen + 32u * (q ^ 1u), push, slot);
  }
}

// End of a round: queue q is consumed; clear its length and the work
// counters of the next round's kernels (a kernel, not a memset, so the round
// can sit in a captured graph).
__global__ void k_wf_flip(WfTables T, uint32_t q) {
  const uint32_t t = threadIdx.x;
  if (t == 0u) T.counters[kWfQueueLen + 32u * q] = 0u;
  if (t < kXcds) {
    T.counters[kWfTraceWork + kCounterStride * t] = 0u;
    T.counters[kWfShadeWork + kCounterStride * t] = 0u;
  }
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
  dev_free(s->tex);
  for (BlasState& b : s->blas)
    for (DevBuf* d : {&b.ids, &b.left, &b.right, &b.parent, &b.kids}) dev_free(*d);
  s->blas.clear();
}

// Every record but the BVH nodes and triangles (emitter BVH, instances,
// vertex attributes, materials, emitters, alias table, environment, textures).
int upload_records(cg_env* env, cg_scene* s, const cg_scene_data* sd, const Layout& L, uint64_t& total) {
  int rc = CG_OK;
  auto up = [&](int r) {
    if (r != CG_OK && rc == CG_OK) rc = r;
  };
  up(upload_vec(env, s->nodes, L.nodes, total));
  up(upload_vec(env, s->insts, L.insts, total));
  up(upload_vec(env, s->vattr, L.vattr, total));
  up(upload_vec(env, s->mats, L.mats, total));
  up(upload_vec(env, s->emit, L.emit, total));
  up(upload_vec(env, s->alias, L.alias, total));
  up(upload_vec(env, s->emit_order, L.emit_order, total));
  const size_t env_bytes = static_cast<size_t>(sd->env_width) * sd->env_height * 4 * sizeof(float);
  total += env_bytes;
  up(dev_upload(env, s->env, sd->env_rgba32f, env_bytes));
  const size_t tex_bytes = static_cast<size_t>(sd->tex_width) * sd->tex_height * 4 * sd->tex_layers;
  if (tex_bytes) {
    total += tex_bytes;
    up(dev_upload(env, s->tex, sd->tex_rgba8, tex_bytes));
  }
  return rc;
}
// The kernels' view of an uploaded scene.
void set_view(cg_scene* s, const cg_scene_data* sd, const Layout& L) {
  DevScene& v = s->view;
  v.bvh = static_cast<const Node8Q*>(s->bvh.p);
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
  const bool tex = s->tex.p != nullptr;
  v.tex = static_cast<const uint32_t*>(s->tex.p);
  v.tex_w = tex ? sd->tex_width : 0u;
  v.tex_h = tex ? sd->tex_height : 0u;
  v.root = L.root;
  s->stack_depth = L.stack_depth;
}

// The host-built TLAS nodes (refs among themselves from 0) moved to start at
// `base`: their interior refs and the root shift; instance and leaf refs stay.
void place_tlas(Layout& L, uint32_t base) {
  auto shift = [base](uint32_t& r) {
    if (r != CG_REF_NONE && (r & (CG_REF_LEAF | CG_REF_INST)) == 0u) r += base;
  };
  for (Node8Q& q : L.bvhq)
    for (int k = 0; k < CG_WIDE; ++k) shift(q.ref[k]);
  shift(L.root);
}

// Builds one BLAS with the linear-BVH kernels: sorts its triangles, writes
// their records into tris / vidx at the job's slots (from the source-order
// copies), and returns its wide nodes (global refs from node_base) in `nodes`.
int lbvh_build_device(cg_env* env, BlasJob& j, uint32_t node_base, const TriRec* tris_src, const uint4* vidx_src,
                      TriRec* tris, uint4* vidx, DevBuf& nodes, BlasState& keep) {
  const uint32_t n = static_cast<uint32_t>(j.boxes.size());
  DevBuf prims, keys, keys2, ids, ids2, left, right, first, count, parent, box, flag, fa, fb, ctr, tmp, kids, price;
  auto release = [&] {
    for (DevBuf* b : {&prims, &keys, &keys2, &ids, &ids2, &left, &right, &first, &count, &parent, &box, &flag, &fa,
                      &fb, &ctr, &tmp, &kids, &price})
      dev_free(*b);
  };
  auto grid = [](uint32_t items) { return dim3((items + kBlock - 1u) / kBlock); };
  int rc = CG_OK;
  auto alloc = [&](DevBuf& b, size_t bytes) {
    if (rc == CG_OK) rc = dev_alloc(env, b, bytes);
  };
  alloc(prims, n * sizeof(PrimBox));
  for (DevBuf* b : {&keys, &keys2, &ids, &ids2}) alloc(*b, n * sizeof(uint32_t));
  for (DevBuf* b : {&left, &right, &first, &count, &flag}) alloc(*b, (n - 1u) * sizeof(uint32_t));
  alloc(parent, (2u * n - 1u) * sizeof(uint32_t));
  alloc(box, (n - 1u) * sizeof(PrimBox));
  alloc(price, (n - 1u) * sizeof(Price));
  alloc(fa, (n - 1u) * sizeof(uint2));
  alloc(fb, (n - 1u) * sizeof(uint2));
  alloc(ctr, 4 * sizeof(uint32_t));
  alloc(nodes, (n - 1u) * sizeof(Node8Q));
  alloc(kids, (n - 1u) * CG_WIDE * sizeof(uint32_t));
  size_t tmp_bytes = 0;
  if (rc == CG_OK &&
      hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, static_cast<uint32_t*>(keys.p), static_cast<uint32_t*>(keys2.p),
                                         static_cast<uint32_t*>(ids.p), static_cast<uint32_t*>(ids2.p),
                                         static_cast<int>(n), 0, 30, env->stream) != hipSuccess)
    rc = fail(env, CG_ERR_HIP, "device BVH build: sort sizing failed");
  alloc(tmp, tmp_bytes);
  if (rc != CG_OK) {
    release();
    dev_free(nodes);
    return rc;
  }
  LbvhArrays A{};
  A.n = n;
  A.prims = static_cast<const PrimBox*>(prims.p);
  A.pad = j.pad;
  lbvh_grid(j, A.cmin, A.cscale);
  A.keys = static_cast<uint32_t*>(keys.p);
  A.ids = static_cast<uint32_t*>(ids.p);
  A.left = static_cast<uint32_t*>(left.p);
  A.right = static_cast<uint32_t*>(right.p);
  A.first = static_cast<uint32_t*>(first.p);
  A.count = static_cast<uint32_t*>(count.p);
  A.parent = static_cast<uint32_t*>(parent.p);
  A.box = static_cast<PrimBox*>(box.p);
  A.flag = static_cast<uint32_t*>(flag.p);
  A.price = static_cast<Price*>(price.p);
  uint32_t* c = static_cast<uint32_t*>(ctr.p);  // [0] next frontier length, [1] wide nodes, [2] most children
  A.front_out_n = c;
  A.wide_n = c + 1;
  A.max_children = c + 2;
  A.nodes = static_cast<Node8Q*>(nodes.p);
  A.exact = nullptr;
  A.kids = static_cast<uint32_t*>(kids.p);
  A.node_base = node_base;
  A.slot_base = j.slot_base;
  auto fin = [&](int code) {
    release();
    if (code != CG_OK) dev_free(nodes);
    return code;
  };
  if (hipMemcpyAsync(prims.p, j.boxes.data(), n * sizeof(PrimBox), hipMemcpyHostToDevice, env->stream) != hipSuccess)
    return fin(fail(env, CG_ERR_HIP, "device BVH build: box upload failed"));
  hipLaunchKernelGGL(k_lbvh_morton, grid(n), dim3(kBlock), 0, env->stream, A);
  if (hipcub::DeviceRadixSort::SortPairs(tmp.p, tmp_bytes, A.keys, static_cast<uint32_t*>(keys2.p), A.ids,
                                         static_cast<uint32_t*>(ids2.p), static_cast<int>(n), 0, 30,
                                         env->stream) != hipSuccess)
    return fin(fail(env, CG_ERR_HIP, "device BVH build: sort failed"));
  A.keys = static_cast<uint32_t*>(keys2.p);
  A.ids = static_cast<uint32_t*>(ids2.p);
  hipLaunchKernelGGL(k_fill_u32, grid(2u * n - 1u), dim3(kBlock), 0, env->stream, A.parent, kLbvhNone, 2u * n - 1u);
  hipLaunchKernelGGL(k_fill_u32, grid(n - 1u), dim3(kBlock), 0, env->stream, A.flag, 0u, n - 1u);
  hipLaunchKernelGGL(k_lbvh_nodes, grid(n - 1u), dim3(kBlock), 0, env->stream, A);
  hipLaunchKernelGGL(k_lbvh_bounds, grid(n), dim3(kBlock), 0, env->stream, A);
  const uint2 root = make_uint2(0u, 0u);
  const uint32_t init[4] = {0u, 1u, 0u, 0u};
  if (hipMemcpyAsync(fa.p, &root, sizeof root, hipMemcpyHostToDevice, env->stream) != hipSuccess ||
      hipMemcpyAsync(c, init, sizeof init, hipMemcpyHostToDevice, env->stream) != hipSuccess)
    return fin(fail(env, CG_ERR_HIP, "device BVH build: counter upload failed"));
  uint2* cur = static_cast<uint2*>(fa.p);
  uint2* nxt = static_cast<uint2*>(fb.p);
  uint32_t width = 1, stack = 0, got[4] = {0u, 0u, 0u, 0u};
  for (int level = 0; width > 0u; ++level) {
    if (level >= 64) return fin(fail(env, CG_ERR_UNSUPPORTED, "device BVH build: tree deeper than 64 levels"));
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(kBlock), 0, env->stream, c, 0u, 1u);      // next frontier length
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(kBlock), 0, env->stream, c + 2, 0u, 1u);  // most children
    A.front_in = cur;
    A.front_out = nxt;
    hipLaunchKernelGGL(k_lbvh_collapse, grid(width), dim3(kBlock), 0, env->stream, A, width);
    if (hipMemcpyAsync(got, c, sizeof got, hipMemcpyDeviceToHost, env->stream) != hipSuccess ||
        hipStreamSynchronize(env->stream) != hipSuccess)
      return fin(fail(env, CG_ERR_HIP, "device BVH build: collapse level %d failed", level));
    stack += got[2] > 0u ? got[2] - 1u : 0u;
    width = got[0];
    std::swap(cur, nxt);
  }
  hipLaunchKernelGGL(k_lbvh_gather, grid(n), dim3(kBlock), 0, env->stream, tris_src + j.slot_base,
                     vidx_src + j.slot_base, tris + j.slot_base, vidx + j.slot_base, A.ids, n);
  if (hipStreamSynchronize(env->stream) != hipSuccess || hipGetLastError() != hipSuccess)
    return fin(fail(env, CG_ERR_HIP, "device BVH build: kernels failed"));
  j.root = node_base;
  j.nodes = got[1];
  j.stack = stack;
  keep.ids = ids2;
  keep.left = left;
  keep.right = right;
  keep.parent = parent;
  keep.kids = kids;
  for (DevBuf* b : {&ids2, &left, &right, &parent, &kids}) *b = DevBuf{};
  keep.n = n;
  keep.slot_base = j.slot_base;
  keep.node_base = node_base;
  keep.nodes = j.nodes;
  keep.stack = stack;
  return fin(CG_OK);
}

// Resident grid of a kernel: blocks per CU at full occupancy x CUs, capped by
// the number of work items.
template <typename K>
uint32_t resident_grid(K kernel, const cg_env* env, uint32_t items, int block = kBlock) {
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, 0) != hipSuccess || per_cu < 1) per_cu = 1;
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
int stack_args(cg_env* env, uint32_t grid, uint32_t depth, StackArgs& sa, int short_n = kShortStack) {
  sa.ovf = nullptr;
  sa.ovf_cap = 0;
  sa.err = static_cast<uint32_t*>(env->errflag.p);
  if (depth <= static_cast<uint32_t>(short_n)) return CG_OK;
  int cap = static_cast<int>(depth) - short_n;
  if (env->test_stack_cap >= 0) cap = std::min(cap, env->test_stack_cap);
  if (cap == 0) return CG_OK;
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
// Work schedule of the per-pixel renders (one per kernel: k = 0 reference
// order, k = 1 sample-stream mode).  A render of the same size as the
// kernel's last one takes its work items longest first by that render's
// durations (progressive frames of one view repeat the same work), so its
// long items start early and the frame does not end waiting on one of them;
// the first render, or one of another size, takes them in image order.  The
// order decides which wave runs an item, never its result (each pixel or
// sample draws from its own stream).  Buffers: durations, the order, the
// sorted keys, the identity, sort storage.
struct Schedule {
  uint32_t* duration = nullptr;
  const uint32_t* order = nullptr;  // null: image order
  uint32_t* next = nullptr;         // the order the sort writes for the next render
  uint32_t* keys = nullptr;
  uint32_t* ids = nullptr;
  void* tmp = nullptr;
};
__global__ void k_iota_u32(uint32_t* p, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = i;
}
int schedule(cg_env* env, int k, uint32_t items, Schedule& sc) {
  if (!env->sched_on) return CG_OK;  // image order, nothing recorded
  size_t tmp = 0;
  if (hipcub::DeviceRadixSort::SortPairsDescending(nullptr, tmp, static_cast<const uint32_t*>(nullptr),
                                                   static_cast<uint32_t*>(nullptr),
                                                   static_cast<const uint32_t*>(nullptr),
                                                   static_cast<uint32_t*>(nullptr), static_cast<int>(items), 0, 8,
                                                   env->stream) != hipSuccess)
    return fail(env, CG_ERR_HIP, "render schedule: sort sizing failed");
  const size_t head = (4u * static_cast<size_t>(items) * sizeof(uint32_t) + 255u) / 256u * 256u;
  if (head + tmp > env->sched[k].bytes) {  // the buffers only grow: renders of many sizes reuse them
    if (env->sched[k].p) {
      CG_HIP(env, hipStreamSynchronize(env->stream));
      dev_free(env->sched[k]);
    }
    env->sched_items[k] = 0;
    int rc = dev_alloc(env, env->sched[k], head + tmp);
    if (rc != CG_OK) return rc;
  }
  uint32_t* base = static_cast<uint32_t*>(env->sched[k].p);
  const size_t n = items;
  if (env->sched_items[k] != items) {  // another size: image order this time, and a new identity
    env->sched_ready[k] = false;
    env->sched_items[k] = items;
    hipLaunchKernelGGL(k_iota_u32, dim3((items + kBlock - 1u) / kBlock), dim3(kBlock), 0, env->stream, base + 3u * n,
                       items);
  }
  env->sched_tmp[k] = tmp;
  sc.duration = base;
  sc.next = base + n;
  sc.order = env->sched_ready[k] ? sc.next : nullptr;
  sc.keys = base + 2u * n;
  sc.ids = base + 3u * n;
  sc.tmp = static_cast<char*>(env->sched[k].p) + head;
  return CG_OK;
}
// After the render (same stream): the next order, longest first; items of
// equal key keep their image order (the sort is stable).
int reorder(cg_env* env, int k, uint32_t items, const Schedule& sc) {
  if (!sc.duration) return CG_OK;
  size_t tmp = env->sched_tmp[k];
  if (hipcub::DeviceRadixSort::SortPairsDescending(sc.tmp, tmp, sc.duration, sc.keys, sc.ids, sc.next,
                                                   static_cast<int>(items), 0, 8,
                                                   env->stream) != hipSuccess)
    return fail(env, CG_ERR_HIP, "render schedule: sort failed");
  env->sched_ready[k] = true;
  return CG_OK;
}
template <int MF>
int launch_view_mf(cg_env* env, const ViewArgs& a, uint32_t tiles, uint32_t depth) {
  const uint32_t grid = resident_grid(k_view<MF>, env, (tiles + 3u) / 4u);
  StackArgs sa;
  int rc = stack_args(env, grid, depth, sa, kViewShortStack);
  if (rc != CG_OK) return rc;
  Schedule sc;
  rc = schedule(env, 0, tiles, sc);
  if (rc != CG_OK) return rc;
  WorkQueue q = reset_queue(env, tiles);
  q.interleave = sc.order ? 1u : 0u;  // the longest tiles dealt to the XCDs in runs
  hipLaunchKernelGGL((k_view<MF>), dim3(grid), dim3(kBlock), 0, env->stream, a, q, sa, sc.order, sc.duration);
  return reorder(env, 0, tiles, sc);
}
int launch_view_ref(cg_env* env, const ViewArgs& a, int frames, uint32_t tiles, uint32_t depth) {
  if (frames <= 4) return launch_view_mf<4>(env, a, tiles, depth);
  if (frames <= 8) return launch_view_mf<8>(env, a, tiles, depth);
  if (frames <= 16) return launch_view_mf<16>(env, a, tiles, depth);
  return launch_view_mf<64>(env, a, tiles, depth);
}
// Samples per lane of the sample-stream mode: one up to 64 samples per pixel
// (a 64-spp wave is one pixel's samples), then enough that a tile has at
// most 64 items (the image does not depend on it: samples are summed in
// sample order).
uint32_t sample_chunk(uint32_t spp) { return spp <= 64u ? 1u : (spp + 63u) / 64u; }
template <int MF>
int launch_view_s_mf(cg_env* env, const ViewArgs& a, const SampleBufs& b, uint32_t tiles, uint32_t depth) {
  const uint32_t chunk = sample_chunk(a.spp), nchunks = (a.spp + chunk - 1u) / chunk;
  const uint64_t items = static_cast<uint64_t>(tiles) * nchunks;
  if (items >= kNoItem) return fail(env, CG_ERR_INVALID, "render too large for the sample-stream mode");
  const uint32_t grid = resident_grid(k_view_s<MF>, env, static_cast<uint32_t>((items + 3u) / 4u));
  StackArgs sa;
  int rc = stack_args(env, grid, depth, sa, kSampleShortStack);
  if (rc != CG_OK) return rc;
  Schedule sc;
  rc = schedule(env, 1, static_cast<uint32_t>(items), sc);
  if (rc != CG_OK) return rc;
  WorkQueue q = reset_queue(env, static_cast<uint32_t>(items));
  q.interleave = sc.order ? 1u : 0u;  // the longest items dealt round-robin to the XCDs
  hipLaunchKernelGGL((k_view_s<MF>), dim3(grid), dim3(kBlock), 0, env->stream, a, b, chunk, nchunks, q, sa, sc.order,
                     sc.duration);
  const uint32_t npix = a.W * a.H;
  hipLaunchKernelGGL(k_view_finish, dim3((npix + kBlock - 1u) / kBlock), dim3(kBlock), 0, env->stream, a, b);
  return reorder(env, 1, static_cast<uint32_t>(items), sc);
}
// The sample-stream buffers of a render: W*H*spp XYZ values (and path records).
int sample_bufs(cg_env* env, const ViewArgs& a, SampleBufs& b) {
  const size_t n = static_cast<size_t>(a.W) * a.H * a.spp;
  const size_t bytes = n * sizeof(float4) + (a.path_hash ? n * sizeof(uint2) : 0);
  if (env->samples.bytes < bytes) {
    if (env->samples.p) {
      CG_HIP(env, hipStreamSynchronize(env->stream));
      dev_free(env->samples);
    }
    int rc = dev_alloc(env, env->samples, bytes);
    if (rc != CG_OK) return rc;
  }
  b.xyz = static_cast<float4*>(env->samples.p);
  b.records = a.path_hash ? reinterpret_cast<uint2*>(static_cast<char*>(env->samples.p) + n * sizeof(float4)) : nullptr;
  return CG_OK;
}
int launch_view(cg_env* env, const ViewArgs& a, int frames, uint32_t tiles, uint32_t depth, uint32_t rng_mode) {
  if (rng_mode == CG_RNG_REFERENCE || a.spp == 0u) return launch_view_ref(env, a, frames, tiles, depth);
  SampleBufs b{};
  int rc = sample_bufs(env, a, b);
  if (rc != CG_OK) return rc;
  if (frames <= 4) return launch_view_s_mf<4>(env, a, b, tiles, depth);
  if (frames <= 8) return launch_view_s_mf<8>(env, a, b, tiles, depth);
  if (frames <= 16) return launch_view_s_mf<16>(env, a, b, tiles, depth);
  return launch_view_s_mf<64>(env, a, b, tiles, depth);
}
// ---------------------------------------------------------------------------
// Fixed scratch of cg_trace_closest_noalloc: the work-queue counters in its
// first kScratchQueueBytes, then the overflow traversal stack.  Every device
// the library runs on holds one copy for the life of the process; host code
// uses only the per-device address that hipGetSymbolAddress returns.
constexpr uint32_t kStreamReadBlocks = 256u * 16u;  // grid of the read-only probe (16 blocks per CU)
constexpr size_t kScratchQueueBytes = 4096;
static_assert(kXcds * kCounterStride * sizeof(uint32_t) <= kScratchQueueBytes, "queue counters fit the scratch head");
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Resolve this device's copy of `scratch` (the device must be current).
int scratch_resolve(cg_env* env) {
  if (env->device < 0 || env->device >= kMaxDevices)
    return fail(env, CG_ERR_UNSUPPORTED, "device %d: the fixed scratch table holds %d devices", env->device,
                kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[env->device]) {
    void* p = nullptr;
    CG_HIP(env, hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(env, CG_ERR_HIP, "hipGetSymbolAddress(scratch) returned a null address");
    g_scratch_ptr[env->device] = p;
  }
  env->scratch = g_scratch_ptr[env->device];
  return CG_OK;
}

// Wait for everything in flight on the env: its stream and every view
// batch's stream.
void wait_all(cg_env* env) {
  if (env->stream) (void)hipStreamSynchronize(env->stream);
  for (cg_vbatch* v : env->vbatch_all)
    if (v->stream) (void)hipStreamSynchronize(v->stream);
}

// Report (and clear) the device error flag after the work that may set it
// has finished.
int check_device_error(cg_env* env) {
  uint32_t flag = 0;
  CG_HIP(env, hipMemcpyAsync(&flag, env->errflag.p, sizeof(flag), hipMemcpyDeviceToHost, env->stream));
  CG_HIP(env, hipStreamSynchronize(env->stream));
  if (!flag) return CG_OK;
  CG_HIP(env, hipMemsetAsync(env->errflag.p, 0, sizeof(flag), env->stream));
  CG_HIP(env, hipStreamSynchronize(env->stream));
  return fail(env, CG_ERR_STATE,
              "traversal stack overflow: a walk needed more entries than the scene's stack bound; results since the "
              "last check are incomplete");
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

// The closest-hit launch of launch_trace with its counters and overflow
// stack in the fixed scratch: the grid shrinks until its overflow stacks fit
// (the kernel is persistent, so fewer blocks only take more rays each), but
// not below one block per CU.
int launch_trace_noalloc(cg_env* env, const cg_scene* scene, const cg_ray* rays, uint32_t n, cg_hit* hits, float* uv,
                         unsigned long long* st) {
  unsigned char* base = static_cast<unsigned char*>(env->scratch);
  if (!base) return fail(env, CG_ERR_STATE, "cg_trace_closest_noalloc: scratch not resolved for device %d", env->device);
  const uint32_t items = (n + 63u) / 64u;
  uint32_t grid = resident_grid(k_trace<false>, env, (items + 3u) / 4u);
  StackArgs sa{nullptr, 0, static_cast<uint32_t*>(env->errflag.p)};
  if (scene->stack_depth > static_cast<uint32_t>(kShortStack)) {
    const size_t cap = scene->stack_depth - kShortStack;
    const size_t per_block = static_cast<size_t>(kBlock) * cap * sizeof(uint2);
    const size_t fit = (kScratchBytes - kScratchQueueBytes) / per_block;
    const uint32_t floor_blocks = std::min<uint32_t>(grid, static_cast<uint32_t>(std::max(env->cu_count, 1)));
    if (fit < floor_blocks)
      return fail(env, CG_ERR_INVALID,
                  "cg_trace_closest_noalloc: a traversal stack of %u entries needs %zu bytes of scratch for %u "
                  "blocks, beyond the fixed %zu",
                  scene->stack_depth, kScratchQueueBytes + per_block * floor_blocks, floor_blocks, kScratchBytes);
    grid = std::min<uint32_t>(grid, static_cast<uint32_t>(fit));
    sa.ovf = reinterpret_cast<uint2*>(base + kScratchQueueBytes);
    sa.ovf_cap = static_cast<int>(cap);
  }
  uint32_t* ctr = reinterpret_cast<uint32_t*>(base);
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(kXcds * kCounterStride), 0, env->stream, ctr, kXcds * kCounterStride);
  WorkQueue q{ctr, items};
  hipLaunchKernelGGL((k_trace<false>), dim3(grid), dim3(kBlock), 0, env->stream, scene->view, rays, nullptr, n, hits,
                     uv, nullptr, st, q, sa);
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
  if ((rc = dev_alloc(env, env->stats, kStatBytes)) != CG_OK) return bail(rc);
  if ((rc = dev_alloc(env, env->queue, kXcds * kCounterStride * sizeof(uint32_t))) != CG_OK) return bail(rc);
  if ((rc = dev_alloc(env, env->errflag, 256)) != CG_OK) return bail(rc);
  if (const char* cap = std::getenv("CG_TEST_STACK_CAP")) env->test_stack_cap = std::max(0, std::atoi(cap));
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) return bail(fail(env, CG_ERR_HIP, "hipGetDeviceProperties failed"));
  env->cu_count = prop.multiProcessorCount;
  if ((rc = scratch_resolve(env)) != CG_OK) return bail(rc);
  if (hipMemsetAsync(env->history.p, 0, env->history.bytes, env->stream) != hipSuccess ||
      hipMemsetAsync(env->stats.p, 0, env->stats.bytes, env->stream) != hipSuccess ||
      hipMemsetAsync(env->errflag.p, 0, env->errflag.bytes, env->stream) != hipSuccess)
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
  wait_all(env);  // no launch may still read what is freed below
  while (!env->batches.empty()) cg_batch_destroy(env->batches.back());
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
  dev_free(env->errflag);
  dev_free(env->samples);
  dev_free(env->sched[0]);
  dev_free(env->sched[1]);
  for (cg_vbatch* v : env->vbatch_all) {
    if (v->stream) (void)hipStreamSynchronize(v->stream);
    for (DevBuf* d : {&v->jobs, &v->first, &v->queue, &v->ovf, &v->jobs_s, &v->first_s, &v->pfirst_s, &v->samples})
      dev_free(*d);
    if (v->exec) (void)hipGraphExecDestroy(v->exec);
    if (v->done) (void)hipEventDestroy(v->done);
    if (v->stream) (void)hipStreamDestroy(v->stream);
    delete v;
  }
  env->vbatch_all.clear();
  env->vbatch_pool.clear();
  if (env->stream) (void)hipStreamDestroy(env->stream);
  delete env;
  return CG_OK;
}

int cg_env_sync(cg_env* env) {
  if (!env) return fail(nullptr, CG_ERR_INVALID, "null env");
  CG_HIP(env, hipStreamSynchronize(env->stream));
  return check_device_error(env);
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

int cg_env_set_rng_mode(cg_env* env, int mode) {
  if (!env) return fail(nullptr, CG_ERR_INVALID, "null env");
  if (mode != CG_RNG_REFERENCE && mode != CG_RNG_SAMPLE) return fail(env, CG_ERR_INVALID, "unknown rng mode %d", mode);
  env->rng_mode = static_cast<uint32_t>(mode);
  return CG_OK;
}

int cg_env_set_schedule(cg_env* env, int enable) {
  if (!env) return fail(nullptr, CG_ERR_INVALID, "null env");
  env->sched_on = enable != 0;
  env->sched_ready[0] = env->sched_ready[1] = false;  // durations are recorded again from the next render
  return CG_OK;
}

int cg_env_set_accel(cg_env* env, int mode) {
  if (!env) return fail(nullptr, CG_ERR_INVALID, "null env");
  if (mode != CG_ACCEL_REFERENCE && mode != CG_ACCEL_REBUILT && mode != CG_ACCEL_DEVICE)
    return fail(env, CG_ERR_INVALID, "unknown accel mode %d", mode);
  env->accel = mode;
  return CG_OK;
}

namespace {
void sum_stats(const unsigned long long* v, cg_stats* out) {
  *out = cg_stats{0, 0, 0, 0, 0};
  for (uint32_t c = 0; c < kStatCopies; ++c) {
    const unsigned long long* w = v + 8u * c;
    out->rays += w[0];
    out->shadow_rays += w[1];
    out->nodes += w[2];
    out->tris += w[3];
    out->insts += w[4];
  }
}
}  // namespace

int cg_stats_read(cg_env* env, cg_stats* out) {
  if (!env || !out) return fail(env, CG_ERR_INVALID, "null argument");
  std::vector<unsigned long long> v(kStatCopies * 8);
  CG_HIP(env, hipMemcpyAsync(v.data(), env->stats.p, kStatBytes, hipMemcpyDeviceToHost, env->stream));
  CG_HIP(env, hipStreamSynchronize(env->stream));
  sum_stats(v.data(), out);
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
  std::vector<BlasJob> jobs;
  const bool on_device = env->accel == CG_ACCEL_DEVICE;
  int rc;
  if (on_device) {  // TLAS and small records on the host, every BLAS below on the device
    rc = build_common(sd, L, err);
    if (rc == CG_OK) rc = build_sah_accel(sd, L, err, &jobs);
    if (rc == CG_OK) {
      collapse_scene_bvh(L);
      quantize_scene_bvh(L);
      build_emitter_bvh(L);
    }
  } else {
    rc = build_layout(sd, L, err, env->accel);
  }
  if (rc != CG_OK) return fail(env, rc, "%s", err.c_str());
  cg_scene* s = new cg_scene();
  uint64_t total = 0;
  auto up = [&](int r) {
    if (r != CG_OK && rc == CG_OK) rc = r;
  };
  if (on_device) {
    // source-order triangle records, one build per BLAS into node ranges from
    // 0, then the host-built TLAS nodes after them (so a refit that changes the
    // TLAS leaves the BLAS refs in place)
    DevBuf tsrc, vsrc;
    up(dev_upload(env, tsrc, L.tris.data(), L.tris.size() * sizeof(TriRec)));
    up(dev_upload(env, vsrc, L.vidx.data(), L.vidx.size() * sizeof(uint4)));
    up(dev_alloc(env, s->tris, L.tris.size() * sizeof(TriRec)));
    up(dev_alloc(env, s->vidx, L.vidx.size() * sizeof(uint4)));
    if (rc == CG_OK && L.tris.size() &&
        (hipMemcpyAsync(s->tris.p, tsrc.p, tsrc.bytes, hipMemcpyDeviceToDevice, env->stream) != hipSuccess ||
         hipMemcpyAsync(s->vidx.p, vsrc.p, vsrc.bytes, hipMemcpyDeviceToDevice, env->stream) != hipSuccess))
      rc = fail(env, CG_ERR_HIP, "scene upload failed");
    std::vector<DevBuf> built(jobs.size());
    s->blas.resize(jobs.size());
    uint64_t node_count = 0;
    for (size_t k = 0; k < jobs.size() && rc == CG_OK; ++k) {
      if (node_count >= CG_REF_NONE) {
        rc = fail(env, CG_ERR_UNSUPPORTED, "too many BVH nodes");
        break;
      }
      rc = lbvh_build_device(env, jobs[k], static_cast<uint32_t>(node_count), static_cast<const TriRec*>(tsrc.p),
                             static_cast<const uint4*>(vsrc.p), static_cast<TriRec*>(s->tris.p),
                             static_cast<uint4*>(s->vidx.p), built[k], s->blas[k]);
      node_count += jobs[k].nodes;
    }
    const uint64_t tlas_base = node_count;
    node_count += L.bvhq.size();
    if (rc == CG_OK && node_count >= CG_REF_NONE) rc = fail(env, CG_ERR_UNSUPPORTED, "too many BVH nodes");
    if (rc == CG_OK) place_tlas(L, static_cast<uint32_t>(tlas_base));
    up(dev_alloc(env, s->bvh, node_count * sizeof(Node8Q)));
    if (rc == CG_OK && !L.bvhq.empty() &&
        hipMemcpyAsync(static_cast<Node8Q*>(s->bvh.p) + tlas_base, L.bvhq.data(), L.bvhq.size() * sizeof(Node8Q),
                       hipMemcpyHostToDevice, env->stream) != hipSuccess)
      rc = fail(env, CG_ERR_HIP, "scene upload failed");
    for (size_t k = 0; k < jobs.size() && rc == CG_OK; ++k)
      if (hipMemcpyAsync(static_cast<Node8Q*>(s->bvh.p) + jobs[k].root, built[k].p, jobs[k].nodes * sizeof(Node8Q),
                         hipMemcpyDeviceToDevice, env->stream) != hipSuccess)
        rc = fail(env, CG_ERR_HIP, "scene upload failed");
    if (hipStreamSynchronize(env->stream) != hipSuccess && rc == CG_OK) rc = fail(env, CG_ERR_HIP, "scene upload failed");
    for (DevBuf& b : built) dev_free(b);
    dev_free(tsrc);
    dev_free(vsrc);
    total += node_count * sizeof(Node8Q) + L.tris.size() * sizeof(TriRec) + L.vidx.size() * sizeof(uint4);
    if (rc == CG_OK) lbvh_finish(L, jobs);
    s->device_built = true;
    s->blas_nodes = static_cast<uint32_t>(tlas_base);
    s->bvh_capacity = node_count;
  } else {
    up(upload_vec(env, s->bvh, L.bvhq, total));
    up(upload_vec(env, s->tris, L.tris, total));
    up(upload_vec(env, s->vidx, L.vidx, total));
  }
  if (rc == CG_OK && L.stack_depth > 4096u)
    rc = fail(env, CG_ERR_UNSUPPORTED, "BVH needs a %u-entry traversal stack (max 4096)", L.stack_depth);
  up(upload_records(env, s, sd, L, total));
  if (rc == CG_OK && hipStreamSynchronize(env->stream) != hipSuccess)
    rc = fail(env, CG_ERR_HIP, "scene upload failed");
  if (rc != CG_OK) {
    release_scene(s);
    delete s;
    return rc;
  }
  set_view(s, sd, L);
  s->bytes = total;
  env->scenes.push_back(s);
  *out = s;
  return CG_OK;
}

int cg_scene_refit(cg_env* env, cg_scene* s, const cg_scene_data* sd) {
  if (!env || !s || !sd) return fail(env, CG_ERR_INVALID, "cg_scene_refit: null argument");
  if (std::find(env->scenes.begin(), env->scenes.end(), s) == env->scenes.end())
    return fail(env, CG_ERR_INVALID, "cg_scene_refit: scene does not belong to this env");
  if (!s->device_built) return fail(env, CG_ERR_INVALID, "cg_scene_refit: the scene was not uploaded with CG_ACCEL_DEVICE");
  Layout L;
  std::string err;
  std::vector<BlasJob> jobs;
  int rc = build_common(sd, L, err);
  if (rc == CG_OK) rc = build_sah_accel(sd, L, err, &jobs);
  if (rc != CG_OK) return fail(env, rc, "%s", err.c_str());
  collapse_scene_bvh(L);
  quantize_scene_bvh(L);
  build_emitter_bvh(L);
  bool same = jobs.size() == s->blas.size() && L.tris.size() * sizeof(TriRec) == s->tris.bytes;
  for (size_t k = 0; same && k < jobs.size(); ++k)
    same = jobs[k].boxes.size() == s->blas[k].n && jobs[k].slot_base == s->blas[k].slot_base;
  if (!same)
    return fail(env, CG_ERR_INVALID, "cg_scene_refit: the scene data has other triangles or BLASes than the scene");
  CG_HIP(env, hipStreamSynchronize(env->stream));
  const uint64_t need = uint64_t{s->blas_nodes} + L.bvhq.size();
  if (need >= CG_REF_NONE) return fail(env, CG_ERR_UNSUPPORTED, "too many BVH nodes");
  if (need > s->bvh_capacity) {  // a larger TLAS: move the BLAS nodes to a larger buffer
    DevBuf nb;
    rc = dev_alloc(env, nb, need * sizeof(Node8Q));
    if (rc != CG_OK) return rc;
    if (s->blas_nodes &&
        hipMemcpyAsync(nb.p, s->bvh.p, uint64_t{s->blas_nodes} * sizeof(Node8Q), hipMemcpyDeviceToDevice, env->stream) !=
            hipSuccess) {
      dev_free(nb);
      return fail(env, CG_ERR_HIP, "cg_scene_refit: copy failed");
    }
    CG_HIP(env, hipStreamSynchronize(env->stream));
    dev_free(s->bvh);
    s->bvh = nb;
    s->bvh_capacity = need;
  }
  auto up = [&](int r) {
    if (r != CG_OK && rc == CG_OK) rc = r;
  };
  DevBuf tsrc, vsrc;
  up(dev_upload(env, tsrc, L.tris.data(), L.tris.size() * sizeof(TriRec)));
  up(dev_upload(env, vsrc, L.vidx.data(), L.vidx.size() * sizeof(uint4)));
  if (rc == CG_OK && L.tris.size() &&
      (hipMemcpyAsync(s->tris.p, tsrc.p, tsrc.bytes, hipMemcpyDeviceToDevice, env->stream) != hipSuccess ||
       hipMemcpyAsync(s->vidx.p, vsrc.p, vsrc.bytes, hipMemcpyDeviceToDevice, env->stream) != hipSuccess))
    rc = fail(env, CG_ERR_HIP, "cg_scene_refit: copy failed");
  Node8Q* bvh = static_cast<Node8Q*>(s->bvh.p);
  auto grid = [](uint32_t items) { return dim3((items + kBlock - 1u) / kBlock); };
  for (size_t k = 0; k < jobs.size() && rc == CG_OK; ++k) {
    BlasState& b = s->blas[k];
    BlasJob& j = jobs[k];
    DevBuf prims, box, flag;
    up(dev_upload(env, prims, j.boxes.data(), j.boxes.size() * sizeof(PrimBox)));
    up(dev_alloc(env, box, (b.n - 1u) * sizeof(PrimBox)));
    up(dev_alloc(env, flag, (b.n - 1u) * sizeof(uint32_t)));
    if (rc == CG_OK) {
      LbvhArrays A{};
      A.n = b.n;
      A.prims = static_cast<const PrimBox*>(prims.p);
      A.pad = j.pad;
      A.ids = static_cast<uint32_t*>(b.ids.p);
      A.left = static_cast<uint32_t*>(b.left.p);
      A.right = static_cast<uint32_t*>(b.right.p);
      A.parent = static_cast<uint32_t*>(b.parent.p);
      A.box = static_cast<PrimBox*>(box.p);
      A.flag = static_cast<uint32_t*>(flag.p);
      A.kids = static_cast<uint32_t*>(b.kids.p);
      A.nodes = bvh + b.node_base;
      A.node_base = b.node_base;
      A.slot_base = b.slot_base;
      hipLaunchKernelGGL(k_fill_u32, grid(b.n - 1u), dim3(kBlock), 0, env->stream, A.flag, 0u, b.n - 1u);
      hipLaunchKernelGGL(k_lbvh_bounds, grid(b.n), dim3(kBlock), 0, env->stream, A);
      hipLaunchKernelGGL(k_lbvh_refit, grid(b.nodes), dim3(kBlock), 0, env->stream, A, b.nodes);
      hipLaunchKernelGGL(k_lbvh_gather, grid(b.n), dim3(kBlock), 0, env->stream,
                         static_cast<const TriRec*>(tsrc.p) + b.slot_base, static_cast<const uint4*>(vsrc.p) + b.slot_base,
                         static_cast<TriRec*>(s->tris.p) + b.slot_base, static_cast<uint4*>(s->vidx.p) + b.slot_base,
                         A.ids, b.n);
      if (hipStreamSynchronize(env->stream) != hipSuccess || hipGetLastError() != hipSuccess)
        rc = fail(env, CG_ERR_HIP, "cg_scene_refit: kernels failed");
    }
    dev_free(prims);
    dev_free(box);
    dev_free(flag);
    j.root = b.node_base;
    j.nodes = b.nodes;
    j.stack = b.stack;
  }
  dev_free(tsrc);
  dev_free(vsrc);
  if (rc != CG_OK) return rc;
  place_tlas(L, s->blas_nodes);
  if (!L.bvhq.empty() && hipMemcpyAsync(bvh + s->blas_nodes, L.bvhq.data(), L.bvhq.size() * sizeof(Node8Q),
                                        hipMemcpyHostToDevice, env->stream) != hipSuccess)
    return fail(env, CG_ERR_HIP, "cg_scene_refit: TLAS upload failed");
  lbvh_finish(L, jobs);
  if (L.stack_depth > 4096u)
    return fail(env, CG_ERR_UNSUPPORTED, "BVH needs a %u-entry traversal stack (max 4096)", L.stack_depth);
  for (DevBuf* d : {&s->nodes, &s->insts, &s->vattr, &s->mats, &s->emit, &s->alias, &s->emit_order, &s->env, &s->tex})
    dev_free(*d);
  uint64_t total = s->bvh_capacity * sizeof(Node8Q) + L.tris.size() * sizeof(TriRec) + L.vidx.size() * sizeof(uint4);
  rc = upload_records(env, s, sd, L, total);
  if (rc == CG_OK && hipStreamSynchronize(env->stream) != hipSuccess) rc = fail(env, CG_ERR_HIP, "cg_scene_refit failed");
  if (rc != CG_OK) return rc;
  set_view(s, sd, L);
  s->bytes = total;
  return CG_OK;
}

int cg_layout_check(const cg_scene_data* sd, int accel, uint64_t* nodes, uint32_t* stack_depth) {
  if (!sd) return fail(nullptr, CG_ERR_INVALID, "cg_layout_check: null scene data");
  Layout L;
  std::string err;
  int rc = build_layout(sd, L, err, accel);
  if (rc != CG_OK) return fail(nullptr, rc, "%s", err.c_str());
  if (L.bvh.size() != L.bvhq.size()) return fail(nullptr, CG_ERR_INVALID, "node arrays differ in length");
  for (size_t i = 0; i < L.bvh.size(); ++i)
    if (!quantized_contains(L.bvh[i], L.bvhq[i]))
      return fail(nullptr, CG_ERR_INVALID, "quantised node %zu does not contain its children's boxes", i);
  // every triangle slot in exactly one leaf (of a node, or an instance's root)
  std::vector<uint8_t> seen(L.tris.size(), 0);
  auto leaf = [&](uint32_t r) {
    if (r == CG_REF_NONE || (r & CG_REF_LEAF) == 0u) return true;
    const uint32_t slot = r & CG_REF_SLOT_MASK, cnt = ((r >> CG_REF_COUNT_SHIFT) & 3u) + 1u;
    for (uint32_t k = 0; k < cnt; ++k)
      if (slot + k >= seen.size() || seen[slot + k]++) return false;
    return true;
  };
  for (const Node8& n : L.bvh)
    for (int i = 0; i < CG_WIDE; ++i)
      if (!leaf(n.ref[i])) return fail(nullptr, CG_ERR_INVALID, "a triangle slot is in two leaves or out of range");
  for (const InstRec& in : L.insts)
    if (!leaf(in.root)) return fail(nullptr, CG_ERR_INVALID, "a triangle slot is in two leaves or out of range");
  for (size_t k = 0; k < seen.size(); ++k)
    if (!seen[k]) return fail(nullptr, CG_ERR_INVALID, "triangle slot %zu is in no leaf", k);
  if (nodes) *nodes = L.bvh.size();
  if (stack_depth) *stack_depth = L.stack_depth;
  return CG_OK;
}

int cg_scene_release(cg_env* env, cg_scene* scene) {
  if (!env || !scene) return fail(env, CG_ERR_INVALID, "cg_scene_release: null argument");
  auto it = std::find(env->scenes.begin(), env->scenes.end(), scene);
  if (it == env->scenes.end()) return fail(env, CG_ERR_INVALID, "scene does not belong to this env");
  wait_all(env);  // view batches on their own streams may read the scene too
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
  int rc = launch_view(env, a, frames, tiles, scene->stack_depth, env->rng_mode);
  if (rc != CG_OK) return rc;
  CG_HIP(env, hipGetLastError());
  return CG_OK;
}

int cg_view_job(cg_env* env, const cg_job* job) {
  if (!env || !job || !job->scene || !job->history) return fail(env, CG_ERR_INVALID, "cg_view_job: null argument");
  if (std::find(env->scenes.begin(), env->scenes.end(), job->scene) == env->scenes.end())
    return fail(env, CG_ERR_INVALID, "job scene does not belong to this env");
  const cg_conf& c = job->conf;
  if (c.render_width == 0 || c.render_height == 0) return fail(env, CG_ERR_INVALID, "job render size must be positive");
  ViewArgs a{};
  a.S = job->scene->view;
  a.P.D = static_cast<int>(std::min(std::max(c.trace_depth, 1u), 64u));
  a.P.T = c.max_transport_sample_count;
  a.P.E = c.max_emission_sample_count;
  a.P.DIF = c.max_diffuse_sample_count;
  a.P.ROUGH = c.max_rough_sample_count;
  a.W = c.render_width;
  a.H = c.render_height;
  a.spp = c.sample_count;
  a.iter = job->iter_count;
  for (int k = 0; k < 3; ++k) {
    a.cam_pos[k] = job->camera.position[k];
    a.cam_dir[k] = job->camera.direction[k];
  }
  a.viewport[0] = job->camera.viewport[0];
  a.viewport[1] = job->camera.viewport[1];
  a.wl_cdf = static_cast<const float*>(env->wl_cdf.p);
  a.wl_leaf = static_cast<const WlLeaf*>(env->wl_leaf.p);
  a.wl_avg = env->wl_avg;
  a.history = job->history;
  a.rgba = reinterpret_cast<uint32_t*>(job->rgba8);
  a.path_hash = job->path_hash;
  a.ray_count = job->path_hash ? job->ray_count : nullptr;
  a.stats = env->counting ? static_cast<unsigned long long*>(env->stats.p) : nullptr;
  if (a.path_hash && !a.ray_count) return fail(env, CG_ERR_INVALID, "cg_view_job: path_hash needs ray_count");
  const uint32_t tiles = ((a.W + 7u) / 8u) * ((a.H + 7u) / 8u);
  if (job->rng_mode != CG_RNG_REFERENCE && job->rng_mode != CG_RNG_SAMPLE)
    return fail(env, CG_ERR_INVALID, "cg_view_job: unknown rng_mode %u", job->rng_mode);
  int rc = launch_view(env, a, std::max(1, a.P.D - 1), tiles, job->scene->stack_depth, job->rng_mode);
  if (rc != CG_OK) return rc;
  CG_HIP(env, hipGetLastError());
  return CG_OK;
}

}  // extern "C"

namespace {
int grow(cg_env* env, DevBuf& b, size_t bytes) {
  if (b.bytes >= bytes && b.p) return CG_OK;
  dev_free(b);
  return dev_alloc(env, b, bytes);
}
template <int MF>
void launch_view_batch_mf(cg_vbatch* v, uint32_t grid, uint32_t njobs, const WorkQueue& q, const StackArgs& sa,
                          unsigned long long* stats) {
  hipLaunchKernelGGL((k_view_batch<MF>), dim3(grid), dim3(kBlock), 0, v->stream,
                     static_cast<const ViewArgs*>(v->jobs.p), static_cast<const uint32_t*>(v->first.p), njobs, q, sa,
                     stats);
}
template <int MF>
uint32_t view_batch_grid(cg_env* env, uint32_t items) {
  return resident_grid(k_view_batch<MF>, env, (items + 3u) / 4u);
}
template <int MF>
void launch_view_batch_s_mf(cg_vbatch* v, uint32_t grid, uint32_t njobs, const WorkQueue& q, const StackArgs& sa,
                            unsigned long long* stats) {
  hipLaunchKernelGGL((k_view_batch_s<MF>), dim3(grid), dim3(kBlock), 0, v->stream,
                     static_cast<const ViewArgs*>(v->jobs_s.p), static_cast<const uint32_t*>(v->first_s.p), njobs, q,
                     sa, stats);
}
template <int MF>
uint32_t view_batch_s_grid(cg_env* env, uint32_t items) {
  return resident_grid(k_view_batch_s<MF>, env, (items + 3u) / 4u);
}
}  // namespace

extern "C" {

int cg_view_batch_launch(cg_env* env, const cg_job* jobs, uint32_t n, cg_vbatch** out) {
  if (!env || !out || (n && !jobs)) return fail(env, CG_ERR_INVALID, "cg_view_batch_launch: null argument");
  *out = nullptr;
  if (n == 0) return fail(env, CG_ERR_INVALID, "cg_view_batch_launch: no jobs");
  // Jobs in order of their expected cost (samples x trace depth per pixel),
  // dealt round-robin to the XCDs: every XCD starts on its most expensive
  // tiles and the launch ends on cheap ones instead of a few long tiles.
  std::vector<uint32_t> ord(n);
  for (uint32_t i = 0; i < n; ++i) ord[i] = i;
#if CG_VB_LPT
  auto cost = [&](uint32_t i) {
    return static_cast<uint64_t>(jobs[i].conf.sample_count) * std::min(std::max(jobs[i].conf.trace_depth, 1u), 64u);
  };
  std::stable_sort(ord.begin(), ord.end(), [&](uint32_t a, uint32_t b) { return cost(a) > cost(b); });
#endif
  std::vector<ViewArgs> va, vs;                      // reference-order jobs, sample-stream jobs
  std::vector<uint32_t> first{0u}, first_s{0u}, pfirst_s{0u};
  std::vector<size_t> sample_off;                     // each sample-stream job's first sample slot
  size_t samples = 0;
  bool records_s = false;
  uint32_t depth = 1, frames = 1, frames_s = 1;
  for (uint32_t k = 0; k < n; ++k) {
    const uint32_t i = ord[k];
    const cg_job* job = jobs + i;
    if (!job->scene || !job->history) return fail(env, CG_ERR_INVALID, "job %u: null scene or history", i);
    if (std::find(env->scenes.begin(), env->scenes.end(), job->scene) == env->scenes.end())
      return fail(env, CG_ERR_INVALID, "job %u: scene does not belong to this env", i);
    const cg_conf& c = job->conf;
    if (c.render_width == 0 || c.render_height == 0) return fail(env, CG_ERR_INVALID, "job %u: empty render size", i);
    if (job->path_hash && !job->ray_count) return fail(env, CG_ERR_INVALID, "job %u: path_hash needs ray_count", i);
    if (job->rng_mode != CG_RNG_REFERENCE && job->rng_mode != CG_RNG_SAMPLE)
      return fail(env, CG_ERR_INVALID, "job %u: unknown rng_mode %u", i, job->rng_mode);
    ViewArgs a{};
    a.S = job->scene->view;
    a.P.D = static_cast<int>(std::min(std::max(c.trace_depth, 1u), 64u));
    a.P.T = c.max_transport_sample_count;
    a.P.E = c.max_emission_sample_count;
    a.P.DIF = c.max_diffuse_sample_count;
    a.P.ROUGH = c.max_rough_sample_count;
    a.W = c.render_width;
    a.H = c.render_height;
    a.spp = c.sample_count;
    a.iter = job->iter_count;
    for (int d = 0; d < 3; ++d) {
      a.cam_pos[d] = job->camera.position[d];
      a.cam_dir[d] = job->camera.direction[d];
    }
    a.viewport[0] = job->camera.viewport[0];
    a.viewport[1] = job->camera.viewport[1];
    a.wl_cdf = static_cast<const float*>(env->wl_cdf.p);
    a.wl_leaf = static_cast<const WlLeaf*>(env->wl_leaf.p);
    a.wl_avg = env->wl_avg;
    a.history = job->history;
    a.rgba = reinterpret_cast<uint32_t*>(job->rgba8);
    a.path_hash = job->path_hash;
    a.ray_count = job->path_hash ? job->ray_count : nullptr;
    a.stats = nullptr;
    const uint64_t tiles = static_cast<uint64_t>((c.render_width + 7u) / 8u) * ((c.render_height + 7u) / 8u);
    depth = std::max(depth, job->scene->stack_depth);
    const uint32_t fr = static_cast<uint32_t>(std::max(1, a.P.D - 1));
    if (job->rng_mode == CG_RNG_SAMPLE && a.spp > 0u) {
      a.chunk = sample_chunk(a.spp);
      a.nchunks = (a.spp + a.chunk - 1u) / a.chunk;
      const uint64_t items = tiles * a.nchunks;
      const uint64_t px = static_cast<uint64_t>(a.W) * a.H;
      if (first_s.back() + items > 0x7FFFFF00ull || pfirst_s.back() + px > 0x7FFFFF00ull)
        return fail(env, CG_ERR_INVALID, "batch too large");
      first_s.push_back(first_s.back() + static_cast<uint32_t>(items));
      pfirst_s.push_back(pfirst_s.back() + static_cast<uint32_t>(px));
      sample_off.push_back(samples);
      samples += px * a.spp;
      records_s = records_s || a.path_hash;
      frames_s = std::max(frames_s, fr);
      vs.push_back(a);
    } else {
      if (first.back() + tiles > 0x7FFFFF00ull) return fail(env, CG_ERR_INVALID, "batch too large");
      first.push_back(first.back() + static_cast<uint32_t>(tiles));
      frames = std::max(frames, fr);
      va.push_back(a);
    }
  }
  cg_vbatch* v = nullptr;
  if (!env->vbatch_pool.empty()) {
    v = env->vbatch_pool.back();
    env->vbatch_pool.pop_back();
  } else {
    v = new cg_vbatch();
    v->env = env;
    env->vbatch_all.push_back(v);
    if (hipStreamCreateWithFlags(&v->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&v->done, hipEventDisableTiming) != hipSuccess) {
      env->vbatch_pool.push_back(v);
      return fail(env, CG_ERR_HIP, "cg_view_batch_launch: stream or event creation failed");
    }
  }
  auto give_back = [&](int rc) {
    env->vbatch_pool.push_back(v);
    return rc;
  };
  const uint32_t nref = static_cast<uint32_t>(va.size()), ns = static_cast<uint32_t>(vs.size());
  const uint32_t items = first.back(), items_s = first_s.back(), npix_s = pfirst_s.back();
  const uint32_t grid = nref == 0 ? 0u
                        : frames <= 4  ? view_batch_grid<4>(env, items)
                        : frames <= 8  ? view_batch_grid<8>(env, items)
                        : frames <= 16 ? view_batch_grid<16>(env, items)
                                       : view_batch_grid<64>(env, items);
  const uint32_t grid_s = ns == 0 ? 0u
                          : frames_s <= 4  ? view_batch_s_grid<4>(env, items_s)
                          : frames_s <= 8  ? view_batch_s_grid<8>(env, items_s)
                          : frames_s <= 16 ? view_batch_s_grid<16>(env, items_s)
                                           : view_batch_s_grid<64>(env, items_s);
  int rc = CG_OK;
  if ((rc = grow(env, v->queue, kXcds * kCounterStride * sizeof(uint32_t))) != CG_OK) return give_back(rc);
  if (nref) {
    if ((rc = grow(env, v->jobs, nref * sizeof(ViewArgs))) != CG_OK) return give_back(rc);
    if ((rc = grow(env, v->first, (nref + 1) * sizeof(uint32_t))) != CG_OK) return give_back(rc);
  }
  if (ns) {
    if ((rc = grow(env, v->jobs_s, ns * sizeof(ViewArgs))) != CG_OK) return give_back(rc);
    if ((rc = grow(env, v->first_s, (ns + 1) * sizeof(uint32_t))) != CG_OK) return give_back(rc);
    if ((rc = grow(env, v->pfirst_s, (ns + 1) * sizeof(uint32_t))) != CG_OK) return give_back(rc);
    const size_t bytes = samples * (sizeof(float4) + (records_s ? sizeof(uint2) : 0));
    if ((rc = grow(env, v->samples, bytes)) != CG_OK) return give_back(rc);
    float4* xyz = static_cast<float4*>(v->samples.p);
    uint2* rec = records_s ? reinterpret_cast<uint2*>(xyz + samples) : nullptr;
    for (uint32_t k = 0; k < ns; ++k) {
      vs[k].sxyz = xyz + sample_off[k];
      vs[k].srec = vs[k].path_hash ? rec + sample_off[k] : nullptr;
    }
  }
  // overflow stacks: one area, sized for the larger need of the two kernels
  StackArgs sa{nullptr, 0, static_cast<uint32_t*>(env->errflag.p)}, sa_s = sa;
  const int cap = nref ? std::max(0, static_cast<int>(depth) - kShortStack) : 0;
  const int cap_s = ns ? std::max(0, static_cast<int>(depth) - kSampleShortStack) : 0;
  const size_t ovf_bytes = std::max(static_cast<size_t>(grid) * kBlock * cap, static_cast<size_t>(grid_s) * kBlock * cap_s) *
                           sizeof(uint2);
  if (ovf_bytes) {
    if ((rc = grow(env, v->ovf, ovf_bytes)) != CG_OK) return give_back(rc);
    sa.ovf = sa_s.ovf = static_cast<uint2*>(v->ovf.p);
    sa.ovf_cap = cap;
    sa_s.ovf_cap = cap_s;
    if (!cap) sa.ovf = nullptr;
    if (!cap_s) sa_s.ovf = nullptr;
  }
  v->host_jobs = std::move(va);
  v->host_first = std::move(first);
  v->host_jobs_s = std::move(vs);
  v->host_first_s = std::move(first_s);
  v->host_pfirst_s = std::move(pfirst_s);
  auto h2d = [&](DevBuf& d, const void* src, size_t bytes) {
    return bytes == 0 || hipMemcpyAsync(d.p, src, bytes, hipMemcpyHostToDevice, v->stream) == hipSuccess;
  };
  if (!h2d(v->jobs, v->host_jobs.data(), nref * sizeof(ViewArgs)) ||
      !h2d(v->first, v->host_first.data(), nref ? (nref + 1) * sizeof(uint32_t) : 0) ||
      !h2d(v->jobs_s, v->host_jobs_s.data(), ns * sizeof(ViewArgs)) ||
      !h2d(v->first_s, v->host_first_s.data(), ns ? (ns + 1) * sizeof(uint32_t) : 0) ||
      !h2d(v->pfirst_s, v->host_pfirst_s.data(), ns ? (ns + 1) * sizeof(uint32_t) : 0))
    return give_back(fail(env, CG_ERR_HIP, "cg_view_batch_launch: job table copy failed"));
  // The counter resets (kernels, not memsets) and the persistent kernels --
  // reference-order jobs, then sample-stream jobs and their per-pixel sums --
  // are captured as one hipGraph and launched as such.  The vbatch came from
  // the pool, so its previous graph has finished.
  if (v->exec) {
    (void)hipGraphExecDestroy(v->exec);
    v->exec = nullptr;
  }
  (void)hipGetLastError();
  if (hipStreamBeginCapture(v->stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
    return give_back(fail(env, CG_ERR_HIP, "cg_view_batch_launch: graph capture failed to begin"));
  uint32_t* ctr = static_cast<uint32_t*>(v->queue.p);
  unsigned long long* stats = env->counting ? static_cast<unsigned long long*>(env->stats.p) : nullptr;
  if (nref) {
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(kXcds * kCounterStride), 0, v->stream, ctr, kXcds * kCounterStride);
    const WorkQueue q{ctr, items, CG_VB_LPT ? 1u : 0u};
    if (frames <= 4)
      launch_view_batch_mf<4>(v, grid, nref, q, sa, stats);
    else if (frames <= 8)
      launch_view_batch_mf<8>(v, grid, nref, q, sa, stats);
    else if (frames <= 16)
      launch_view_batch_mf<16>(v, grid, nref, q, sa, stats);
    else
      launch_view_batch_mf<64>(v, grid, nref, q, sa, stats);
  }
  if (ns) {
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(kXcds * kCounterStride), 0, v->stream, ctr, kXcds * kCounterStride);
    const WorkQueue q{ctr, items_s, CG_VB_LPT ? 1u : 0u};
    if (frames_s <= 4)
      launch_view_batch_s_mf<4>(v, grid_s, ns, q, sa_s, stats);
    else if (frames_s <= 8)
      launch_view_batch_s_mf<8>(v, grid_s, ns, q, sa_s, stats);
    else if (frames_s <= 16)
      launch_view_batch_s_mf<16>(v, grid_s, ns, q, sa_s, stats);
    else
      launch_view_batch_s_mf<64>(v, grid_s, ns, q, sa_s, stats);
    hipLaunchKernelGGL(k_view_batch_finish, dim3((npix_s + kBlock - 1u) / kBlock), dim3(kBlock), 0, v->stream,
                       static_cast<const ViewArgs*>(v->jobs_s.p), static_cast<const uint32_t*>(v->pfirst_s.p), ns,
                       npix_s);
  }
  const hipError_t le = hipGetLastError();
  hipGraph_t graph = nullptr;
  const hipError_t ce = hipStreamEndCapture(v->stream, &graph);
  const bool made = le == hipSuccess && ce == hipSuccess && graph &&
                    hipGraphInstantiate(&v->exec, graph, nullptr, nullptr, 0) == hipSuccess;
  if (graph) (void)hipGraphDestroy(graph);
  if (!made) {
    v->exec = nullptr;
    return give_back(fail(env, CG_ERR_HIP, "cg_view_batch_launch: graph capture failed"));
  }
  if (hipGraphLaunch(v->exec, v->stream) != hipSuccess || hipEventRecord(v->done, v->stream) != hipSuccess)
    return give_back(fail(env, CG_ERR_HIP, "cg_view_batch_launch: launch failed"));
  v->in_flight = true;
  *out = v;
  return CG_OK;
}

int cg_view_batch_query(cg_vbatch* v, int* done) {
  if (!v || !done) return fail(nullptr, CG_ERR_INVALID, "cg_view_batch_query: null argument");
  const hipError_t e = hipEventQuery(v->done);
  if (e == hipSuccess) {
    *done = 1;
    return CG_OK;
  }
  if (e == hipErrorNotReady) {
    *done = 0;
    return CG_OK;
  }
  return fail(v->env, CG_ERR_HIP, "cg_view_batch_query: %s", hipGetErrorString(e));
}

int cg_view_batch_wait(cg_vbatch* v) {
  if (!v) return fail(nullptr, CG_ERR_INVALID, "cg_view_batch_wait: null batch");
  CG_HIP(v->env, hipEventSynchronize(v->done));
  return check_device_error(v->env);
}

int cg_view_batch_release(cg_vbatch* v) {
  if (!v) return fail(nullptr, CG_ERR_INVALID, "cg_view_batch_release: null batch");
  CG_HIP(v->env, hipEventSynchronize(v->done));
  v->in_flight = false;
  v->env->vbatch_pool.push_back(v);
  return CG_OK;
}

int cg_read_target(cg_env* env, cg_history_pixel* history, uint8_t* rgba8) {
  if (!env) return fail(nullptr, CG_ERR_INVALID, "null env");
  if (history) CG_HIP(env, hipMemcpyAsync(history, env->history.p, env->history.bytes, hipMemcpyDeviceToHost, env->stream));
  if (rgba8) CG_HIP(env, hipMemcpyAsync(rgba8, env->rgba.p, env->rgba.bytes, hipMemcpyDeviceToHost, env->stream));
  CG_HIP(env, hipStreamSynchronize(env->stream));
  return check_device_error(env);
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

int cg_trace_closest_noalloc(cg_env* env, cg_scene* scene, const cg_ray* rays, uint32_t n, cg_hit* hits, float* uv) {
  if (!env || !scene || (!rays && n) || (!hits && n))
    return fail(env, CG_ERR_INVALID, "cg_trace_closest_noalloc: null argument");
  if (n == 0) return CG_OK;
  unsigned long long* st = env->counting ? static_cast<unsigned long long*>(env->stats.p) : nullptr;
  int rc = launch_trace_noalloc(env, scene, rays, n, hits, uv, st);
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

int cg_alias_pick(cg_env* env, cg_scene* scene, const float* u, uint32_t n, uint32_t* out) {
  if (!env || !scene || (n && (!u || !out))) return fail(env, CG_ERR_INVALID, "cg_alias_pick: null argument");
  if (scene->view.alias_count == 0 || scene->view.emit_count == 0)
    return fail(env, CG_ERR_STATE, "cg_alias_pick: the scene has no emitters");
  if (n == 0) return CG_OK;
  hipLaunchKernelGGL(k_alias_pick, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, env->stream, scene->view,
                     reinterpret_cast<const float2*>(u), n, out);
  CG_HIP(env, hipGetLastError());
  return CG_OK;
}

int cg_stream_read(cg_env* env, const void* src, size_t bytes, uint32_t* sink) {
  if (!env || !src || !sink) return fail(env, CG_ERR_INVALID, "cg_stream_read: null argument");
  if (bytes % 16 != 0) return fail(env, CG_ERR_INVALID, "cg_stream_read: size must be a multiple of 16");
  const size_t n = bytes / 16;
  if (n == 0) return CG_OK;
  const uint32_t blocks = static_cast<uint32_t>(std::min<size_t>((n + kBlock - 1) / kBlock, kStreamReadBlocks));
  hipLaunchKernelGGL(k_stream_read, dim3(blocks), dim3(kBlock), 0, env->stream, static_cast<const uint4*>(src), n,
                     sink);
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

// ===========================================================================
// Batched rendering: the wavefront engine.
struct cg_batch {
  cg_env* env = nullptr;
  cg_batch_conf conf{};
  uint32_t frames_per_slot = 0;
  DevBuf path, hits, frames, slot_job, q0, q1, counters, job_left, jobs, scenes, stats, ovf;
  WfTables T{};
  StackArgs sa{};
  uint32_t grid_trace = 0, grid_shade = 0;
  void (*round)(cg_batch*, uint32_t) = nullptr;  // enqueue one round for the chosen kernel shape
  struct Job {
    bool active = false;
    uint32_t slot_base = 0, nslots = 0, scene = 0;
  };
  std::vector<Job> job;
  std::vector<WfJob> job_dev;        // host images of the device job table
  std::vector<uint32_t> left_init;   // host images of job_left writes
  std::vector<uint32_t> left_read;
  uint32_t qlen_read[64];
  std::vector<const cg_scene*> scene_ptr;
  std::vector<DevScene> scene_dev;
  std::vector<uint32_t> scene_jobs;  // active jobs per scene-table entry (0: the entry may be reused)
  std::vector<std::pair<uint32_t, uint32_t>> free_list;  // (first slot, count), sorted
  hipGraphExec_t exec = nullptr;
  uint64_t bytes = 0;
  uint32_t active = 0;
};

namespace {

constexpr uint32_t kBatchMaxScenes = 64u;

void batch_release(cg_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  b->exec = nullptr;
  for (DevBuf* d : {&b->path, &b->hits, &b->frames, &b->slot_job, &b->q0, &b->q1, &b->counters, &b->job_left,
                    &b->jobs, &b->scenes, &b->stats, &b->ovf})
    dev_free(*d);
}

bool slots_alloc(cg_batch* b, uint32_t n, uint32_t& first) {
  for (size_t i = 0; i < b->free_list.size(); ++i) {
    auto& r = b->free_list[i];
    if (r.second < n) continue;
    first = r.first;
    r.first += n;
    r.second -= n;
    if (r.second == 0) b->free_list.erase(b->free_list.begin() + static_cast<long>(i));
    return true;
  }
  return false;
}
void slots_free(cg_batch* b, uint32_t first, uint32_t n) {
  auto& fl = b->free_list;
  auto it = std::lower_bound(fl.begin(), fl.end(), std::make_pair(first, 0u));
  it = fl.insert(it, std::make_pair(first, n));
  size_t i = static_cast<size_t>(it - fl.begin());
  if (i + 1 < fl.size() && fl[i].first + fl[i].second == fl[i + 1].first) {
    fl[i].second += fl[i + 1].second;
    fl.erase(fl.begin() + static_cast<long>(i) + 1);
  }
  if (i > 0 && fl[i - 1].first + fl[i - 1].second == fl[i].first) {
    fl[i - 1].second += fl[i].second;
    fl.erase(fl.begin() + static_cast<long>(i));
  }
}

template <int BLOCK, int SHORT>
void round_t(cg_batch* b, uint32_t q) {
  hipStream_t st = b->env->stream;
  hipLaunchKernelGGL((k_wf_trace<BLOCK, SHORT>), dim3(b->grid_trace), dim3(BLOCK), 0, st, b->T, q, b->sa);
  hipLaunchKernelGGL((k_wf_shade<BLOCK, SHORT>), dim3(b->grid_shade), dim3(BLOCK), 0, st, b->T, q, b->sa);
  hipLaunchKernelGGL(k_wf_flip, dim3(1), dim3(64), 0, st, b->T, q);
}
template <int BLOCK, int SHORT>
void shape_t(cg_batch* b, cg_env* env) {
  b->round = round_t<BLOCK, SHORT>;
  b->grid_trace = resident_grid(k_wf_trace<BLOCK, SHORT>, env, 1u << 30, BLOCK);
  b->grid_shade = resident_grid(k_wf_shade<BLOCK, SHORT>, env, 1u << 30, BLOCK);
}
// Kernel shapes the autotuner chooses from: workgroup size x LDS stack entries.
bool batch_shape(cg_batch* b, cg_env* env, uint32_t block, uint32_t short_n) {
  switch (block * 100u + short_n) {
    case 6408: shape_t<64, 8>(b, env); return true;
    case 6416: shape_t<64, 16>(b, env); return true;
    case 6424: shape_t<64, 24>(b, env); return true;
    case 12808: shape_t<128, 8>(b, env); return true;
    case 12816: shape_t<128, 16>(b, env); return true;
    case 12824: shape_t<128, 24>(b, env); return true;
    case 25608: shape_t<256, 8>(b, env); return true;
    case 25616: shape_t<256, 16>(b, env); return true;
    case 25624: shape_t<256, 24>(b, env); return true;
    default: return false;
  }
}
void enqueue_round(cg_batch* b, uint32_t q) { b->round(b, q); }

int batch_build_graph(cg_batch* b) {
  cg_env* env = b->env;
  hipGraph_t g = nullptr;
  CG_HIP(env, hipStreamBeginCapture(env->stream, hipStreamCaptureModeThreadLocal));
  for (uint32_t r = 0; r < b->conf.rounds_per_launch; ++r) enqueue_round(b, r & 1u);
  hipError_t e = hipStreamEndCapture(env->stream, &g);
  if (e != hipSuccess) return fail(env, CG_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e));
  e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) return fail(env, CG_ERR_HIP, "graph instantiation failed: %s", hipGetErrorString(e));
  return CG_OK;
}

}  // namespace

extern "C" {

int cg_batch_create(cg_env* env, const cg_batch_conf* conf, cg_batch** out) {
  if (!env || !conf || !out) return fail(env, CG_ERR_INVALID, "cg_batch_create: null argument");
  *out = nullptr;
  cg_batch_conf c = *conf;
  if (c.slot_capacity == 0 || c.slot_capacity > (1u << 28))
    return fail(env, CG_ERR_INVALID, "slot_capacity must be in [1, 2^28]");
  c.slot_capacity = (c.slot_capacity + 63u) & ~63u;
  if (c.max_trace_depth < 1 || c.max_trace_depth > 64) return fail(env, CG_ERR_INVALID, "max_trace_depth must be in [1, 64]");
  if (c.rounds_per_launch < 2 || (c.rounds_per_launch & 1u))
    return fail(env, CG_ERR_INVALID, "rounds_per_launch must be even and >= 2");
  if (c.max_jobs == 0 || c.max_jobs > 4096) return fail(env, CG_ERR_INVALID, "max_jobs must be in [1, 4096]");
  if (c.max_stack_depth == 0 || c.max_stack_depth > 4096)
    return fail(env, CG_ERR_INVALID, "max_stack_depth must be in [1, 4096]");
  if (c.block_size == 0) c.block_size = kBlock;
  if (c.short_stack == 0) c.short_stack = kShortStack;
  cg_batch* b = new cg_batch();
  b->env = env;
  b->conf = c;
  b->frames_per_slot = std::max(1u, c.max_trace_depth - 1u);
  const size_t S = c.slot_capacity;
  int rc = CG_OK;
  auto bail = [&](int code) {
    batch_release(b);
    delete b;
    return code;
  };
  auto alloc = [&](DevBuf& d, size_t bytes) {
    if (rc == CG_OK) rc = dev_alloc(env, d, bytes);
    b->bytes += bytes;
  };
  alloc(b->path, S * sizeof(PathRec));
  alloc(b->hits, S * sizeof(HitW));
  alloc(b->frames, S * b->frames_per_slot * sizeof(Frame5));
  alloc(b->slot_job, S * sizeof(uint32_t));
  alloc(b->q0, S * sizeof(uint32_t));
  alloc(b->q1, S * sizeof(uint32_t));
  alloc(b->counters, kWfCounterWords * sizeof(uint32_t));
  alloc(b->job_left, c.max_jobs * sizeof(uint32_t));
  alloc(b->jobs, c.max_jobs * sizeof(WfJob));
  alloc(b->scenes, kBatchMaxScenes * sizeof(DevScene));
  alloc(b->stats, kStatBytes);
  if (rc != CG_OK) return bail(rc);
  if (!batch_shape(b, env, c.block_size, c.short_stack))
    return bail(fail(env, CG_ERR_INVALID, "unsupported kernel shape: block %u, short stack %u", c.block_size,
                     c.short_stack));
  // The traversal-stack overflow area is the batch's own (its address is
  // baked into the graph), sized for the larger of the two resident grids.
  b->sa.ovf = nullptr;
  b->sa.ovf_cap = 0;
  b->sa.err = static_cast<uint32_t*>(env->errflag.p);
  if (c.max_stack_depth > c.short_stack) {
    b->sa.ovf_cap = static_cast<int>(c.max_stack_depth - c.short_stack);
    alloc(b->ovf, static_cast<size_t>(std::max(b->grid_trace, b->grid_shade)) * c.block_size *
                      static_cast<size_t>(b->sa.ovf_cap) * sizeof(uint2));
    if (rc != CG_OK) return bail(rc);
    b->sa.ovf = static_cast<uint2*>(b->ovf.p);
  }
  if (hipMemsetAsync(b->counters.p, 0, b->counters.bytes, env->stream) != hipSuccess ||
      hipMemsetAsync(b->job_left.p, 0, b->job_left.bytes, env->stream) != hipSuccess ||
      hipMemsetAsync(b->stats.p, 0, b->stats.bytes, env->stream) != hipSuccess)
    return bail(fail(env, CG_ERR_HIP, "hipMemsetAsync failed"));
  WfTables& T = b->T;
  T.path = static_cast<PathRec*>(b->path.p);
  T.hits = static_cast<HitW*>(b->hits.p);
  T.frames = static_cast<Frame5*>(b->frames.p);
  T.slot_job = static_cast<uint32_t*>(b->slot_job.p);
  T.queue[0] = static_cast<uint32_t*>(b->q0.p);
  T.queue[1] = static_cast<uint32_t*>(b->q1.p);
  T.counters = static_cast<uint32_t*>(b->counters.p);
  T.job_left = static_cast<uint32_t*>(b->job_left.p);
  T.jobs = static_cast<const WfJob*>(b->jobs.p);
  T.scenes = static_cast<const DevScene*>(b->scenes.p);
  T.wl_cdf = static_cast<const float*>(env->wl_cdf.p);
  T.wl_leaf = static_cast<const WlLeaf*>(env->wl_leaf.p);
  T.wl_avg = env->wl_avg;
  T.slot_cap = c.slot_capacity;
  T.stats = static_cast<unsigned long long*>(b->stats.p);
  b->job.resize(c.max_jobs);
  b->job_dev.resize(c.max_jobs);
  b->left_init.assign(c.max_jobs, 0u);
  b->left_read.assign(c.max_jobs, 0u);
  b->scene_dev.resize(kBatchMaxScenes);
  b->free_list.push_back({0u, c.slot_capacity});
  if (c.use_graph && (rc = batch_build_graph(b)) != CG_OK) return bail(rc);
  if (hipStreamSynchronize(env->stream) != hipSuccess) return bail(fail(env, CG_ERR_HIP, "stream sync failed"));
  env->batches.push_back(b);
  *out = b;
  return CG_OK;
}

int cg_batch_destroy(cg_batch* b) {
  if (!b) return fail(nullptr, CG_ERR_INVALID, "null batch");
  cg_env* env = b->env;
  (void)hipStreamSynchronize(env->stream);
  batch_release(b);
  env->batches.erase(std::remove(env->batches.begin(), env->batches.end(), b), env->batches.end());
  delete b;
  return CG_OK;
}

int cg_batch_footprint(const cg_batch* b, uint64_t* bytes, uint32_t* slots_free) {
  if (!b) return fail(nullptr, CG_ERR_INVALID, "null batch");
  if (bytes) *bytes = b->bytes;
  if (slots_free) {
    uint32_t f = 0;
    for (const auto& r : b->free_list) f = std::max(f, r.second);
    *slots_free = f;
  }
  return CG_OK;
}

int cg_batch_submit(cg_batch* b, const cg_job* job, uint32_t* job_id) {
  if (!b || !job || !job->scene || !job->history) return fail(b ? b->env : nullptr, CG_ERR_INVALID, "cg_batch_submit: null argument");
  cg_env* env = b->env;
  const cg_conf& c = job->conf;
  if (c.render_width == 0 || c.render_height == 0) return fail(env, CG_ERR_INVALID, "job render size must be positive");
  if (job->rng_mode != CG_RNG_REFERENCE)
    return fail(env, CG_ERR_UNSUPPORTED, "cg_batch_submit: the wavefront engine renders in reference order only");
  const uint32_t depth = std::min(std::max(c.trace_depth, 1u), 64u);
  if (depth > b->conf.max_trace_depth)
    return fail(env, CG_ERR_INVALID, "job trace_depth %u exceeds the batch's %u", depth, b->conf.max_trace_depth);
  if (std::find(env->scenes.begin(), env->scenes.end(), job->scene) == env->scenes.end())
    return fail(env, CG_ERR_INVALID, "job scene does not belong to this env");
  if (job->scene->stack_depth > b->conf.max_stack_depth)
    return fail(env, CG_ERR_INVALID, "scene needs a %u-entry stack; the batch supports %u", job->scene->stack_depth,
                b->conf.max_stack_depth);
  const uint32_t tiles_x = (c.render_width + 7u) / 8u, tiles_y = (c.render_height + 7u) / 8u;
  const uint64_t nslots = static_cast<uint64_t>(tiles_x) * tiles_y * 64u;
  if (nslots > b->conf.slot_capacity)
    return fail(env, CG_ERR_INVALID, "job needs %llu slots; the batch has %u", static_cast<unsigned long long>(nslots),
                b->conf.slot_capacity);
  uint32_t j = 0;
  while (j < b->job.size() && b->job[j].active) ++j;
  if (j == b->job.size()) return fail(env, CG_ERR_FULL, "job table full");
  // Scene-table entry: the scene's own while a job of it is active (the same
  // address and device view), else any entry no active job uses.  Entries
  // free up as jobs finish, so a full table is "retry later".
  uint32_t sc = 0;
  while (sc < b->scene_ptr.size() &&
         !(b->scene_ptr[sc] == job->scene && b->scene_jobs[sc] > 0 &&
           std::memcmp(&b->scene_dev[sc], &job->scene->view, sizeof(DevScene)) == 0))
    ++sc;
  if (sc == b->scene_ptr.size()) {
    sc = 0;
    while (sc < b->scene_ptr.size() && b->scene_jobs[sc] > 0) ++sc;
    if (sc == b->scene_ptr.size()) {
      if (sc == kBatchMaxScenes) return fail(env, CG_ERR_FULL, "scene table full: %u scenes have active jobs", sc);
      b->scene_ptr.push_back(nullptr);
      b->scene_jobs.push_back(0u);
    }
  }
  uint32_t first = 0;
  if (!slots_alloc(b, static_cast<uint32_t>(nslots), first)) return fail(env, CG_ERR_FULL, "no room for %llu slots now",
                                                                         static_cast<unsigned long long>(nslots));
  if (b->scene_jobs[sc] == 0u) {
    b->scene_ptr[sc] = job->scene;
    b->scene_dev[sc] = job->scene->view;
    if (hipMemcpyAsync(static_cast<DevScene*>(b->scenes.p) + sc, &b->scene_dev[sc], sizeof(DevScene),
                       hipMemcpyHostToDevice, env->stream) != hipSuccess) {
      slots_free(b, first, static_cast<uint32_t>(nslots));
      return fail(env, CG_ERR_HIP, "cg_batch_submit: scene table copy failed");
    }
  }
  WfJob& J = b->job_dev[j];
  J = WfJob{};
  J.scene = sc;
  J.W = c.render_width;
  J.H = c.render_height;
  J.spp = c.sample_count;
  J.iter = job->iter_count;
  J.tiles_x = tiles_x;
  J.slot_base = first;
  J.P.D = static_cast<int>(depth);
  J.P.T = c.max_transport_sample_count;
  J.P.E = c.max_emission_sample_count;
  J.P.DIF = c.max_diffuse_sample_count;
  J.P.ROUGH = c.max_rough_sample_count;
  for (int k = 0; k < 3; ++k) {
    J.cam_pos[k] = job->camera.position[k];
    J.cam_dir[k] = job->camera.direction[k];
  }
  J.viewport[0] = job->camera.viewport[0];
  J.viewport[1] = job->camera.viewport[1];
  J.history = job->history;
  J.rgba = reinterpret_cast<uint32_t*>(job->rgba8);
  J.path_hash = job->path_hash;
  J.ray_count = job->ray_count;
  b->left_init[j] = c.render_width * c.render_height;
  CG_HIP(env, hipMemcpyAsync(static_cast<WfJob*>(b->jobs.p) + j, &J, sizeof(WfJob), hipMemcpyHostToDevice, env->stream));
  CG_HIP(env, hipMemcpyAsync(static_cast<uint32_t*>(b->job_left.p) + j, &b->left_init[j], sizeof(uint32_t),
                             hipMemcpyHostToDevice, env->stream));
  hipLaunchKernelGGL(k_wf_admit, dim3(static_cast<uint32_t>((nslots + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                     env->stream, b->T, j, static_cast<uint32_t>(nslots));
  CG_HIP(env, hipGetLastError());
  b->job[j].active = true;
  b->job[j].slot_base = first;
  b->job[j].nslots = static_cast<uint32_t>(nslots);
  b->job[j].scene = sc;
  ++b->scene_jobs[sc];
  ++b->active;
  if (job_id) *job_id = j;
  return CG_OK;
}

int cg_batch_launch(cg_batch* b, uint32_t launches) {
  if (!b) return fail(nullptr, CG_ERR_INVALID, "null batch");
  cg_env* env = b->env;
  for (uint32_t l = 0; l < launches; ++l) {
    if (b->exec) {
      CG_HIP(env, hipGraphLaunch(b->exec, env->stream));
    } else {
      for (uint32_t r = 0; r < b->conf.rounds_per_launch; ++r) enqueue_round(b, r & 1u);
      CG_HIP(env, hipGetLastError());
    }
  }
  return CG_OK;
}

int cg_batch_poll(cg_batch* b, uint32_t* done, uint32_t cap, uint32_t* n_done, uint32_t* n_active) {
  if (!b) return fail(nullptr, CG_ERR_INVALID, "null batch");
  cg_env* env = b->env;
  CG_HIP(env, hipMemcpyAsync(b->left_read.data(), b->job_left.p, b->left_read.size() * sizeof(uint32_t),
                             hipMemcpyDeviceToHost, env->stream));
  CG_HIP(env, hipMemcpyAsync(b->qlen_read, b->counters.p, sizeof(b->qlen_read), hipMemcpyDeviceToHost, env->stream));
  CG_HIP(env, hipStreamSynchronize(env->stream));
  uint32_t nd = 0;
  for (uint32_t j = 0; j < b->job.size(); ++j) {
    if (!b->job[j].active || b->left_read[j] != 0u) continue;
    if (done && nd < cap) {
      done[nd++] = j;
    } else if (done) {
      break;  // report the rest next time
    }
    b->job[j].active = false;
    slots_free(b, b->job[j].slot_base, b->job[j].nslots);
    --b->scene_jobs[b->job[j].scene];
    --b->active;
  }
  if (n_done) *n_done = nd;
  if (n_active) *n_active = b->active;
  if (b->active > 0 && b->qlen_read[kWfQueueLen] == 0u && b->qlen_read[kWfQueueLen + 32u] == 0u)
    return fail(env, CG_ERR_STATE, "batch: %u jobs active but no rays queued", b->active);
  return check_device_error(env);
}

int cg_batch_drain(cg_batch* b) {
  if (!b) return fail(nullptr, CG_ERR_INVALID, "null batch");
  uint32_t n_active = b->active;
  while (n_active > 0) {
    int rc = cg_batch_launch(b, 1);
    if (rc != CG_OK) return rc;
    rc = cg_batch_poll(b, nullptr, 0, nullptr, &n_active);
    if (rc != CG_OK) return rc;
  }
  return CG_OK;
}

int cg_batch_stats(cg_batch* b, cg_stats* out, int reset) {
  if (!b || !out) return fail(b ? b->env : nullptr, CG_ERR_INVALID, "null argument");
  std::vector<unsigned long long> v(kStatCopies * 8);
  CG_HIP(b->env, hipMemcpyAsync(v.data(), b->stats.p, kStatBytes, hipMemcpyDeviceToHost, b->env->stream));
  if (reset) CG_HIP(b->env, hipMemsetAsync(b->stats.p, 0, b->stats.bytes, b->env->stream));
  CG_HIP(b->env, hipStreamSynchronize(b->env->stream));
  sum_stats(v.data(), out);
  return CG_OK;
}

}  // extern "C"
