// cg_render.h — the compute pass for one pixel (compute.wesl:32-62,
// render.wesl:7-18, trace.wesl:22-113), shared by the HIP kernel.
//
// The pixel consumes its RNG stream in exactly the reference's order: spp
// samples, each a depth-first walk of the transport tree.  The open frames of
// the walk live in F (one per level above the deepest, trace_depth - 1), the
// BVH traversal stack behind `stk`.
#ifndef CG_RENDER_H_
#define CG_RENDER_H_

#include "cg_core.h"

namespace cg {

struct ViewArgs {
  DevScene S;
  TraceParams P;
  uint32_t W, H, spp, iter;
  float cam_pos[3], cam_dir[3], viewport[2];
  const float* wl_cdf;
  const WlLeaf* wl_leaf;
  float wl_avg;
  cg_history_pixel* history;
  uint32_t* rgba;
  uint32_t* path_hash;
  uint32_t* ray_count;
  unsigned long long* stats;
  // sample-stream mode in a many-job launch: the job's per-sample buffers and
  // its split of the samples into work items
  float4* sxyz;
  uint2* srec;
  uint32_t chunk, nchunks;
};


// One spectral sample of a pixel (render.wesl:7-18 for one iteration): the
// wavelength, then the depth-first walk of the transport tree, drawing from
// `rng` and folding the path's discrete decisions into `hash`; returns the
// sample's XYZ contribution, illuminate(lambda, L) / p(lambda) / spp.
template <int MF>
CG_HD f3 trace_sample(const ViewArgs& A, uint32_t px, uint32_t py, uint32_t& rng, uint32_t& hash,
                                      Frame5* F, const TStack& stk, TravStats& st, uint32_t& nrays,
                                      uint32_t& nshadow) {
  const DevScene& S = A.S;
  // sampleWavelength: walk the bisection heap.
  float target = clampf(rng_float(rng), 0.0f, 0.999999f);
  uint32_t node = 1u;
#pragma unroll
  for (int k = 0; k < CG_WL_LEVELS; ++k) node = 2u * node + (A.wl_cdf[node] < target ? 1u : 0u);
  const uint32_t leaf = node - CG_WL_LEAVES;
  const float lambda = A.wl_leaf[leaf].lambda;
  const SmitsBin sb = smits_bin(lambda);
  // trace(): depth-first over the transport tree.
  int depth = 0;
  f3 ro, rd;
  camera_ray(A.cam_pos, A.cam_dir, A.viewport, A.W, A.H, px, py, ro, rd);
  float ret = 0.0f;
  for (;;) {
    HitRec h = traverse<false>(S, ro, rd, kMax, stk, st);
    ++nrays;
    if (A.path_hash)  // the triangle id is read for the path record only
      hash = hmix(hmix(hash, h.inst), h.inst != CG_MISS_ID ? S.tris[h.slot].tri_id : CG_MISS_ID);
    bool descended = false;
    if (h.inst == CG_MISS_ID) {
      uint32_t texel = 0u;
      ret = env_radiance(S, rd, sb, texel);
      hash = hmix(hash, 0x5E000000u ^ texel);
    } else if (depth >= A.P.D - 1) {
      ret = 0.0f;
    } else {
      Surface sf = surface_at(S, ro, rd, h);
      Frame5& f = F[depth];
      forward_pass(S, rd, lambda, sb, sf, depth, A.P, rng, f, stk);
      if (f.ecount != 0u) {  // solveEmission (trace.wesl:91-113)
        f3 lp = sf.pos + f.edir * f.edist;
        f3 so = spawn_origin(sf.pos, sf.gn, f.edir);
        float expected = length(so - lp);
        HitRec sh = traverse<true>(S, so, f.edir, fmn(expected - 1e-3f, kMax), stk, st);
        ++nshadow;
        bool occ = sh.inst != CG_MISS_ID;
        hash = hmix(hmix(hash, 0xA5000000u | (occ ? 1u : 0u)), f.ecount - 1u);  // outcome, emitter
        if (occ) f.erad = 0.0f;
      }
      if (f.tcount != 0u) {
        f.iter = 1u;
        rd = f.tdir[0];
        ro = spawn_origin(f.pos, f.gn, rd);
        ++depth;
        descended = true;
      } else {
        float rad = 0.0f;
        if (f.ecount != 0u) rad += f.erad;
        ret = f.emission + rad;
      }
    }
    if (descended) continue;
    // return `ret` up the tree (collect.wesl:4-17, shading.wesl:48-53)
    bool next_sibling = false;
    while (depth > 0) {
      Frame5& p = F[depth - 1];
      p.acc += p.tcoeff[p.iter - 1u] * ret;
      if (p.iter < p.tcount) {
        rd = p.tdir[p.iter];
        ro = spawn_origin(p.pos, p.gn, rd);
        p.iter++;
        next_sibling = true;
        break;
      }
      float rad = p.acc;
      if (p.ecount != 0u) rad += p.erad;
      ret = p.emission + rad;
      --depth;
    }
    if (!next_sibling) break;
  }
  // illuminate (luminance.wesl:13-16), averaged over spp (render.wesl:13-16)
  const WlLeaf L = A.wl_leaf[leaf];
  const float spp_f = static_cast<float>(A.spp);
  return mk3(L.x * ret / A.wl_avg / L.prob / spp_f, L.y * ret / A.wl_avg / L.prob / spp_f,
             L.z * ret / A.wl_avg / L.prob / spp_f);
}

// The running mean and the pixel's outputs (compute.wesl:41-61).
CG_HD void store_pixel(const ViewArgs& A, uint32_t hidx, f3 acc, uint32_t seed_out, uint32_t hash,
                                       uint32_t rays) {
  const float w = 1.0f / static_cast<float>(A.iter + 1u);
  f3 lum = mk3(0.0f, 0.0f, 0.0f);
  if (A.iter > 0u) {
    const cg_history_pixel hp = A.history[hidx];
    lum = mk3(hp.xyz[0], hp.xyz[1], hp.xyz[2]) * (1.0f - w);
  }
  lum = lum + acc * w;
  cg_history_pixel out;
  out.xyz[0] = lum.x;
  out.xyz[1] = lum.y;
  out.xyz[2] = lum.z;
  out.rand_seed = seed_out;
  A.history[hidx] = out;
  if (A.rgba) A.rgba[hidx] = xyz_to_rgba8(lum);
  if (A.path_hash) {
    A.path_hash[hidx] = hash;
    A.ray_count[hidx] = rays;
  }
}

// Reference order: the pixel's spp samples from its one stream, in the
// reference's order.  Only what the sample loop needs stays live across it:
// the camera ray is recomputed per sample, the wavelength leaf reread for the
// sample's XYZ, and the running mean read at the end (the pixel is this
// lane's alone).
template <int MF>
CG_HD void render_pixel(const ViewArgs& A, uint32_t px, uint32_t py, const TStack& stk, TravStats& st,
                        uint32_t& nrays, uint32_t& nshadow) {
  const uint32_t hidx = px + py * A.W;
  const uint32_t rays_before = nrays + nshadow;  // the lane's counters span several pixels
  uint32_t rng = A.iter > 0u ? A.history[hidx].rand_seed : hidx;
  rng_step(rng);
  uint32_t hash = 2166136261u;
  Frame5 F[MF];
  f3 acc = mk3(0.0f, 0.0f, 0.0f);
  for (uint32_t s = 0; s < A.spp; ++s) {
    const f3 v = trace_sample<MF>(A, px, py, rng, hash, F, stk, st, nrays, nshadow);
    acc.x += v.x;
    acc.y += v.y;
    acc.z += v.z;
  }
  rng_step(rng);
  store_pixel(A, hidx, acc, rng, hash, nrays + nshadow - rays_before);
}

// Sample-stream mode: per-sample results, sample-major (sample s of pixel p
// at s * W * H + p), summed in sample order by finish_pixel.
struct SampleBufs {
  float4* xyz;      // the sample's XYZ contribution (w unused)
  uint2* records;   // path records: the sample's hash and query count (nullable)
};

// Samples [s0, s1) of one pixel, each from its own stream.
template <int MF>
CG_HD void render_samples(const ViewArgs& A, const SampleBufs& B, uint32_t px, uint32_t py, uint32_t s0, uint32_t s1,
                          const TStack& stk, TravStats& st, uint32_t& nrays, uint32_t& nshadow) {
  const uint32_t hidx = px + py * A.W;
  const uint32_t base = A.iter > 0u ? A.history[hidx].rand_seed : hidx;
  const size_t npix = static_cast<size_t>(A.W) * A.H;
  Frame5 F[MF];
  for (uint32_t s = s0; s < s1; ++s) {
    uint32_t rng = sample_seed(base, s);
    uint32_t hash = 2166136261u;
    const uint32_t before = nrays + nshadow;
    const f3 v = trace_sample<MF>(A, px, py, rng, hash, F, stk, st, nrays, nshadow);
    const size_t k = static_cast<size_t>(s) * npix + hidx;
    B.xyz[k] = make_float4(v.x, v.y, v.z, 0.0f);
    if (B.records) B.records[k] = make_uint2(hash, nrays + nshadow - before);
  }
}

// Sample-stream mode: the pixel's samples summed in sample order, its path
// hash the fold of the samples' hashes, the carried seed sample_carry(base).
CG_HD void finish_pixel(const ViewArgs& A, const SampleBufs& B, uint32_t hidx) {
  const uint32_t base = A.iter > 0u ? A.history[hidx].rand_seed : hidx;
  const size_t npix = static_cast<size_t>(A.W) * A.H;
  f3 acc = mk3(0.0f, 0.0f, 0.0f);
  uint32_t hash = 2166136261u, rays = 0u;
  for (uint32_t s = 0; s < A.spp; ++s) {
    const size_t k = static_cast<size_t>(s) * npix + hidx;
    const float4 v = B.xyz[k];
    acc.x += v.x;
    acc.y += v.y;
    acc.z += v.z;
    if (B.records) {
      const uint2 r = B.records[k];
      hash = hmix(hash, r.x);
      rays += r.y;
    }
  }
  store_pixel(A, hidx, acc, sample_carry(base), hash, rays);
}

}  // namespace cg

#endif  // CG_RENDER_H_
