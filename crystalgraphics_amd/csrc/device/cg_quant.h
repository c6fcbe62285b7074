// cg_quant.h — quantisation of 8-wide BVH nodes (Node8 -> Node8Q), shared by
// the host build and the device build.
//
// The decoder (trav_step) computes a bound as origin + float(q) * step with
// step = 2^(e - 127): one exact product and one rounded sum.  The encoder
// picks per axis the smallest step whose 255th grid line reaches the node's
// upper bound, then per child the largest q whose decoded value is <= the
// child's lower bound and the smallest q whose decoded value is >= its upper
// bound, evaluating the decoder's own expression, so every decoded box
// contains the exact one whatever the initial estimates were.
#ifndef CG_QUANT_H_
#define CG_QUANT_H_

#include <math.h>
#include <string.h>

#include "cg_layout.h"

#if defined(__HIPCC__) || defined(__HIP__)
#include <hip/hip_runtime.h>
#define CG_QHD __host__ __device__ __forceinline__
#else
#define CG_QHD inline
#endif

namespace cg {

CG_QHD float q_step(uint32_t biased) {
  const uint32_t bits = biased << 23;
  float f;
  memcpy(&f, &bits, sizeof f);
  return f;
}
CG_QHD float q_decode(float origin, uint32_t q, float step) { return origin + static_cast<float>(q) * step; }

CG_QHD Node8Q quantize_node(const Node8& n) {
  Node8Q out;
  memset(&out, 0, sizeof out);
  int used[CG_WIDE];
  int nu = 0;
  for (int i = 0; i < CG_WIDE; ++i) {
    out.ref[i] = n.ref[i];
    if (n.ref[i] != CG_REF_NONE) used[nu++] = i;
  }
  uint32_t exps = 0;
  for (int a = 0; a < 3; ++a) {
    float blo = __builtin_inff(), bhi = -__builtin_inff();
    for (int k = 0; k < nu; ++k) {
      blo = fminf(blo, n.lo[a][used[k]]);
      bhi = fmaxf(bhi, n.hi[a][used[k]]);
    }
    if (nu == 0) blo = bhi = 0.0f;
    out.origin[a] = blo;
    const double extent = static_cast<double>(bhi) - static_cast<double>(blo);
    int be = 1;
    if (extent > 0.0) {
      const int e = static_cast<int>(floor(log2(extent / 255.0))) + 127 - 1;
      be = e > 1 ? e : 1;
    }
    while (be < 254 && q_decode(blo, 255u, q_step(static_cast<uint32_t>(be))) < bhi) ++be;
    const float step = q_step(static_cast<uint32_t>(be));
    exps |= static_cast<uint32_t>(be) << (8 * a);
    for (int k = 0; k < nu; ++k) {
      const int i = used[k];
      const double fl = floor((static_cast<double>(n.lo[a][i]) - blo) / step);
      uint32_t ql = static_cast<uint32_t>(fmin(255.0, fmax(0.0, fl)));
      while (ql > 0u && q_decode(blo, ql, step) > n.lo[a][i]) --ql;
      const double fh = ceil((static_cast<double>(n.hi[a][i]) - blo) / step);
      uint32_t qh = static_cast<uint32_t>(fmin(255.0, fmax(0.0, fh)));
      while (qh < 255u && q_decode(blo, qh, step) < n.hi[a][i]) ++qh;
      out.qlo[a][i] = static_cast<uint8_t>(ql);
      out.qhi[a][i] = static_cast<uint8_t>(qh);
    }
  }
  out.exps = exps;
  return out;
}

// Every decoded child box contains the exact one.
CG_QHD bool quantized_contains(const Node8& n, const Node8Q& q) {
  for (int i = 0; i < CG_WIDE; ++i) {
    if (n.ref[i] != q.ref[i]) return false;
    if (n.ref[i] == CG_REF_NONE) continue;
    for (int a = 0; a < 3; ++a) {
      const float step = q_step((q.exps >> (8 * a)) & 0xFFu);
      if (!(q_decode(q.origin[a], q.qlo[a][i], step) <= n.lo[a][i])) return false;
      if (!(q_decode(q.origin[a], q.qhi[a][i], step) >= n.hi[a][i])) return false;
    }
  }
  return true;
}

// The exact child boxes a quantised node decodes to (for checks).
CG_QHD void dequantize_box(const Node8Q& q, int i, float* lo, float* hi) {
  for (int a = 0; a < 3; ++a) {
    const float step = q_step((q.exps >> (8 * a)) & 0xFFu);
    lo[a] = q_decode(q.origin[a], q.qlo[a][i], step);
    hi[a] = q_decode(q.origin[a], q.qhi[a][i], step);
  }
}

}  // namespace cg

#endif  // CG_QUANT_H_
