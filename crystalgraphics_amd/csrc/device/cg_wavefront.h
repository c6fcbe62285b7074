// cg_wavefront.h — the compute pass (compute.wesl:32-62, render.wesl:7-18,
// trace.wesl:22-113) as a wavefront: every pixel of every admitted render job
// owns a slot whose path state lives in HBM, and each round traces one ray per
// queued slot (one kernel) and then advances each slot's walk by one event
// (another kernel), which queues the slot's next ray.  The events of one
// pixel happen in the order of the reference's depth-first walk, so a pixel
// draws its RNG stream exactly as there and as the one-lane-per-pixel kernel
// (cg_render.h) does: same samples, same hits, same XYZ.
#ifndef CG_WAVEFRONT_H_
#define CG_WAVEFRONT_H_

#include "cg_core.h"

namespace cg {

// Path state of one slot, 64 B.  The first 32 B are the ray in flight (what
// the trace kernel reads); flags: bit 0 = shadow ray (any hit below tmax),
// bits 8-15 = depth of the open frame.
struct __attribute__((aligned(16))) PathRec {
  float ro[3];
  float tmax;
  float rd[3];
  uint32_t flags;
  uint32_t rng;
  uint32_t hash;    // running hash of the hit-id sequence (path record)
  uint32_t sample;  // samples finished
  uint32_t leaf;    // wavelength table leaf of the sample in progress
  float acc[3];     // sum of finished samples' XYZ contributions
  uint32_t queries; // closest-hit + shadow queries so far
};
static_assert(sizeof(PathRec) == 64, "PathRec");

// Hit of the ray in flight, 32 B.
struct __attribute__((aligned(16))) HitW {
  float t, u, v;
  uint32_t inst, slot;
  uint32_t _pad[3];
};
static_assert(sizeof(HitW) == 32, "HitW");

// One render job, as the kernels see it.
struct WfJob {
  uint32_t scene;      // index into the scene table
  uint32_t W, H, spp, iter;
  uint32_t tiles_x;    // 8x8 pixel tiles per row; slots are tile-major
  uint32_t slot_base;  // first slot; the job owns tiles * 64 slots
  uint32_t _pad;
  TraceParams P;
  float cam_pos[3], cam_dir[3], viewport[2];
  cg_history_pixel* history;
  uint32_t* rgba;
  uint32_t* path_hash;  // nullable
  uint32_t* ray_count;  // nullable
};

// Slot `t` of a job (tile-major: tile t / 64, lane t % 64) -> pixel.
CG_HD void slot_pixel(const WfJob& J, uint32_t t, uint32_t& px, uint32_t& py) {
  const uint32_t tile = t >> 6, lane = t & 63u;
  px = (tile % J.tiles_x) * 8u + (lane & 7u);
  py = (tile / J.tiles_x) * 8u + (lane >> 3);
}

// Everything the wavefront kernels address (fixed for the life of a batch,
// so a captured graph of rounds stays valid).
struct WfTables {
  PathRec* path;
  HitW* hits;
  Frame5* frames;        // frame d of slot s at frames[d * slot_cap + s]
  uint32_t* slot_job;
  uint32_t* queue[2];
  uint32_t* counters;    // see kWf* offsets
  uint32_t* job_left;    // pixels of job j not yet finished
  const WfJob* jobs;
  const DevScene* scenes;
  const float* wl_cdf;
  const WlLeaf* wl_leaf;
  float wl_avg;
  uint32_t slot_cap;
  unsigned long long* stats;  // nullable: traversal counters
};
// Counter words (each on its own 128-byte line): queue lengths, and the
// per-XCD work counters of the trace and shade kernels.
constexpr uint32_t kWfQueueLen = 0u;         // + 32 * q
constexpr uint32_t kWfTraceWork = 64u;       // 8 counters, 32 apart
constexpr uint32_t kWfShadeWork = 64u + 256u;
constexpr uint32_t kWfCounterWords = 64u + 512u;

// sampleWavelength (physics/wavelength.wesl:20-45): walk the bisection heap.
CG_HD uint32_t wf_sample_leaf(const float* wl_cdf, uint32_t& rng) {
  float target = clampf(rng_float(rng), 0.0f, 0.999999f);
  uint32_t node = 1u;
#pragma unroll
  for (int k = 0; k < CG_WL_LEVELS; ++k) node = 2u * node + (wl_cdf[node] < target ? 1u : 0u);
  return node - CG_WL_LEAVES;
}

CG_HD void wf_camera(const WfJob& J, uint32_t px, uint32_t py, f3& o, f3& d) {
  camera_ray(J.cam_pos, J.cam_dir, J.viewport, J.W, J.H, px, py, o, d);
}

// Final write of a pixel (compute.wesl:41-61): running mean with the
// history, carried RNG state, sRGB target, path record.
CG_HD void wf_finish(const WfJob& J, uint32_t px, uint32_t py, f3 acc, uint32_t rng, uint32_t hash,
                     uint32_t queries) {
  const uint32_t hidx = px + py * J.W;
  const float w = 1.0f / static_cast<float>(J.iter + 1u);
  f3 lum = mk3(0.0f, 0.0f, 0.0f);
  if (J.iter > 0u) {
    cg_history_pixel hp = J.history[hidx];
    lum = mk3(hp.xyz[0], hp.xyz[1], hp.xyz[2]) * (1.0f - w);
  }
  lum = lum + acc * w;
  cg_history_pixel out;
  out.xyz[0] = lum.x;
  out.xyz[1] = lum.y;
  out.xyz[2] = lum.z;
  rng_step(rng);
  out.rand_seed = rng;
  J.history[hidx] = out;
  if (J.rgba) J.rgba[hidx] = xyz_to_rgba8(lum);
  if (J.path_hash) J.path_hash[hidx] = hash;
  if (J.ray_count) J.ray_count[hidx] = queries;
}

// The state of a slot between events, in registers.
struct WfState {
  f3 ro, rd;
  float tmax;
  bool shadow;
  int depth;
  uint32_t rng, hash, sample, leaf, queries;
  f3 acc;
};
CG_HD WfState wf_load(const PathRec& p) {
  WfState s;
  s.ro = mk3(p.ro[0], p.ro[1], p.ro[2]);
  s.rd = mk3(p.rd[0], p.rd[1], p.rd[2]);
  s.tmax = p.tmax;
  s.shadow = (p.flags & 1u) != 0u;
  s.depth = static_cast<int>((p.flags >> 8) & 0xFFu);
  s.rng = p.rng;
  s.hash = p.hash;
  s.sample = p.sample;
  s.leaf = p.leaf;
  s.acc = mk3(p.acc[0], p.acc[1], p.acc[2]);
  s.queries = p.queries;
  return s;
}
CG_HD void wf_store(PathRec& p, const WfState& s) {
  p.ro[0] = s.ro.x;
  p.ro[1] = s.ro.y;
  p.ro[2] = s.ro.z;
  p.tmax = s.tmax;
  p.rd[0] = s.rd.x;
  p.rd[1] = s.rd.y;
  p.rd[2] = s.rd.z;
  p.flags = (s.shadow ? 1u : 0u) | (static_cast<uint32_t>(s.depth) << 8);
  p.rng = s.rng;
  p.hash = s.hash;
  p.sample = s.sample;
  p.leaf = s.leaf;
  p.acc[0] = s.acc.x;
  p.acc[1] = s.acc.y;
  p.acc[2] = s.acc.z;
  p.queries = s.queries;
}

// Start the next sample: wavelength, camera ray.
CG_HD void wf_start_sample(const WfTables& T, const WfJob& J, uint32_t px, uint32_t py, WfState& s) {
  s.leaf = wf_sample_leaf(T.wl_cdf, s.rng);
  s.depth = 0;
  wf_camera(J, px, py, s.ro, s.rd);
  s.tmax = kMax;
  s.shadow = false;
}

// Advance a slot by one event: consume the hit of the ray in flight and set up
// the next ray.  Returns false when the pixel has finished (outputs written).
CG_HD bool wf_step(const WfTables& T, const WfJob& J, const DevScene& S, uint32_t slot, uint32_t px, uint32_t py,
                   WfState& s, const HitW& hw, const TStack& stk) {
  const WlLeaf L = T.wl_leaf[s.leaf];
  const SmitsBin sb = smits_bin(L.lambda);
  Frame5* F = T.frames + slot;
  const size_t fs = T.slot_cap;  // stride between a slot's frames
  float ret = 0.0f;
  bool settled = false;  // `ret` is the radiance of the node at `depth`
  ++s.queries;
  if (s.shadow) {
    // solveEmission's visibility test (trace.wesl:91-113)
    Frame5 f = F[s.depth * fs];
    const bool occ = hw.inst != CG_MISS_ID;
    s.hash = hmix(hmix(s.hash, 0xA5000000u | (occ ? 1u : 0u)), f.ecount - 1u);  // outcome, emitter
    if (occ) f.erad = 0.0f;
    s.shadow = false;
    if (f.tcount != 0u) {
      f.iter = 1u;
      s.rd = f.tdir[0];
      s.ro = spawn_origin(f.pos, f.gn, s.rd);
      s.tmax = kMax;
      F[s.depth * fs] = f;
      ++s.depth;
    } else {
      float rad = 0.0f;
      rad += f.erad;
      ret = f.emission + rad;
      settled = true;
    }
  } else {
    s.hash = hmix(hmix(s.hash, hw.inst), hw.inst != CG_MISS_ID ? S.tris[hw.slot].tri_id : CG_MISS_ID);
    if (hw.inst == CG_MISS_ID) {
      uint32_t texel = 0u;
      ret = env_radiance(S, s.rd, sb, texel);
      s.hash = hmix(s.hash, 0x5E000000u ^ texel);
      settled = true;
    } else if (s.depth >= J.P.D - 1) {
      ret = 0.0f;
      settled = true;
    } else {
      HitRec h{hw.t, hw.u, hw.v, hw.inst, hw.slot};
      Surface sf = surface_at(S, s.ro, s.rd, h);
      Frame5 f;
      forward_pass(S, s.rd, L.lambda, sb, sf, s.depth, J.P, s.rng, f, stk);
      if (f.ecount != 0u) {
        f3 lp = sf.pos + f.edir * f.edist;
        s.ro = spawn_origin(sf.pos, sf.gn, f.edir);
        s.rd = f.edir;
        s.tmax = fmn(length(s.ro - lp) - 1e-3f, kMax);
        s.shadow = true;
        F[s.depth * fs] = f;
      } else if (f.tcount != 0u) {
        f.iter = 1u;
        s.rd = f.tdir[0];
        s.ro = spawn_origin(f.pos, f.gn, s.rd);
        s.tmax = kMax;
        F[s.depth * fs] = f;
        ++s.depth;
      } else {
        ret = f.emission;
        settled = true;
      }
    }
  }
  if (!settled) return true;
  // return `ret` up the tree (collect.wesl:4-17, shading.wesl:48-53)
  while (s.depth > 0) {
    Frame5 p = F[(s.depth - 1) * fs];
    p.acc += p.tcoeff[p.iter - 1u] * ret;
    if (p.iter < p.tcount) {
      s.rd = p.tdir[p.iter];
      s.ro = spawn_origin(p.pos, p.gn, s.rd);
      s.tmax = kMax;
      p.iter++;
      F[(s.depth - 1) * fs] = p;
      return true;
    }
    float rad = p.acc;
    if (p.ecount != 0u) rad += p.erad;
    ret = p.emission + rad;
    --s.depth;
  }
  // illuminate (luminance.wesl:13-16), averaged over spp (render.wesl:13-16)
  const float spp_f = static_cast<float>(J.spp);
  s.acc.x += L.x * ret / T.wl_avg / L.prob / spp_f;
  s.acc.y += L.y * ret / T.wl_avg / L.prob / spp_f;
  s.acc.z += L.z * ret / T.wl_avg / L.prob / spp_f;
  if (++s.sample < J.spp) {
    wf_start_sample(T, J, px, py, s);
    return true;
  }
  wf_finish(J, px, py, s.acc, s.rng, s.hash, s.queries);
  return false;
}

}  // namespace cg

#endif  // CG_WAVEFRONT_H_
